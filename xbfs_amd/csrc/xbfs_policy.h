// Per-level direction / queue-mode decision, usable from host code and from
// device code (the persistent kernel evaluates it on the GPU).
//
// The reference heuristic restated from BFS_ControlFlow,
// bfs_adaptiveFQ_wb_async.cuh:1047-1078 (decision) and 1078-1167 (branches):
//   count[L] = sum of degrees of level-L vertices (1042-1046)
//   TD_adaptive trigger, once (1047-1064):
//     L==0 && count>=INF                                   -> TD_adaptive=1
//     L>0 && count[L-1]>=INF && count>=10*INF && count/count[L-1]>=8 -> =L
//     (L>0 && count>=100*INF) || (count[L-1]<INF && count>=10*INF &&
//                                 count/count[L-1]>=1000)  -> =L
//   ratio = (float)count/(float)E                               (1067)
//   ratio>=alpha && tag[L-1]>=0  -> hub_before=L                (1069)
//   ratio<alpha  && tag[L-1]==-1 -> hub_after=L                 (1071)
//   ratio<alpha: hub_before==0 && L<TD_adaptive -> scan-free (tag 0)
//                else hub_after>0 && L>hub_after -> last phase (tag 0)
//                else single-scan (tag 1)
//   else bottom-up (tag -1)
// Level 0 reads tag[-1] / count[-1] in the reference; both reads are inert
// (they can only assign hub_* = 0 or are short-circuited) and are modelled as
// tag[-1] = -2, count[-1] = 0.
#ifndef XBFS_POLICY_H
#define XBFS_POLICY_H
#include <stdint.h>

#if defined(__HIPCC__)
#define XBFS_HD __host__ __device__
#else
#define XBFS_HD
#endif

#define XBFS_INFINITE 10000 /* bfs_adaptiveFQ_wb_async.cuh:29 */

struct XbfsHeurState {
    int32_t td_adaptive;
    int32_t is_change;
    int32_t hub_before;
    int32_t hub_after;
    int64_t count_prev;
    int32_t tag_prev;
    int32_t pad;
};

XBFS_HD inline void xbfs_heur_init(XbfsHeurState *s)
{
    s->td_adaptive = XBFS_INFINITE;
    s->is_change = 0;
    s->hub_before = 0;
    s->hub_after = 0;
    s->count_prev = 0;
    s->tag_prev = -2;
    s->pad = 0;
}

// Returns the reference tag (0, 1, -1); *sub receives 0 scan-free,
// 2 last phase, 1 single-scan, -1 bottom-up.  Advances the state.
XBFS_HD inline int32_t xbfs_heur_decide(XbfsHeurState *s, int32_t L, int64_t c,
                                        int64_t E, float alpha, int32_t *sub)
{
    const int64_t INF = XBFS_INFINITE;
    const int64_t cp = (L > 0) ? s->count_prev : 0;
    const int32_t tp = (L > 0) ? s->tag_prev : -2;
    if (!s->is_change) {
        if (L == 0 && c >= INF) {
            s->is_change = 1;
            s->td_adaptive = L + 1;
        } else if (L > 0 && cp >= INF && c >= INF * 10 && c / cp >= 8) {
            s->is_change = 1;
            s->td_adaptive = L;
        } else if ((L > 0 && c >= INF * 100) ||
                   (cp < INF && c >= INF * 10 && cp != 0 && c / cp >= 1000)) {
            // cp != 0: a non-empty level L needs an edge out of level L-1, and
            // at L == 0 the first branch already took c >= INF.
            s->is_change = 1;
            s->td_adaptive = L;
        }
    }
    const float ratio = (float)c / (float)E;
    if (ratio >= alpha && tp >= 0) s->hub_before = L;
    if (ratio < alpha && tp == -1) s->hub_after = L;
    int32_t t, sb;
    if (ratio < alpha) {
        t = 0;
        if (s->hub_before == 0 && L < s->td_adaptive) {
            sb = 0;
        } else if (s->hub_after > 0 && L > s->hub_after) {
            sb = 2;
        } else {
            t = 1;
            sb = 1;
        }
    } else {
        t = -1;
        sb = -1;
    }
    s->count_prev = c;
    s->tag_prev = t;
    if (sub) *sub = sb;
    return t;
}

// tag -> execution mode (XBFS_MODE_*: 0 TD scan-free, 1 TD single-scan, 2 BU)
XBFS_HD inline int32_t xbfs_tag_to_mode(int32_t tag)
{
    return tag == 0 ? 0 : (tag == 1 ? 1 : 2);
}
XBFS_HD inline int32_t xbfs_mode_to_tag(int32_t mode)
{
    return mode == 0 ? 0 : (mode == 1 ? 1 : -1);
}

// ---- learned cost model ---------------------------------------------------
// FP32 MLP 8 -> 32 -> 32 -> 3 (ReLU), parameters packed as
// W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32], b3[3] (1443 floats).
// Output j is the predicted log2 time (us) of mode j (XBFS_MODE_* order:
// 0 top-down scan-free, 1 top-down single-scan, 2 bottom-up).
#define XB_MLP_NI 8
#define XB_MLP_NH 32
#define XB_MLP_NO 3
// Domain of the trained model.  It was trained on skewed RMAT graphs (the
// training family's largest degree is 363-1366 times the mean degree of the
// non-isolated vertices) and on levels below 16 (the level feature L/16 stays
// below 1).  On low-skew graphs (Erdos-Renyi about 2.4, grids about 1) the
// modes' costs follow other rules, and deeper levels lie outside the features
// it has seen: there the heuristic decides, and the level records say so.
#define XBFS_MLP_MIN_SKEW 16.0
#define XBFS_MLP_MAX_LEVEL 15
#define XB_MLP_NPARAM (XB_MLP_NH * XB_MLP_NI + XB_MLP_NH + XB_MLP_NH * XB_MLP_NH + XB_MLP_NH + \
                       XB_MLP_NO * XB_MLP_NH + XB_MLP_NO)

XBFS_HD inline float xb_log2p1(long long v)
{
    // log2(1+v) without libm: the exponent from the top set bit of 1+v (the
    // power-of-two scaling to [1,2) is exact), the mantissa part by a
    // degree-5 least-squares fit of log2(1+f) on [0,1] that is exact at
    // f = 0 and f = 1 (max error 2.2e-5; features only)
    double x = (double)v + 1.0;
    int e = 0;
    if (v >= 1) {
        e = 63 - __builtin_clzll((unsigned long long)v + 1ull);
        x *= 1.0 / (double)(1ull << e);
    }
    const double f = x - 1.0;
    const double l =
        f * (1.4417403 + f * (-0.70777018 + f * (0.41234422 + f * (-0.19031903 + f * 0.04400469))));
    return (float)(e + l);
}

// Features of level L: log2 sizes of the frontier (vertices, edges), the
// log2 ratios of the frontier to the unvisited part (edges, vertices), the
// level index, the previous direction and whether an exact queue is available
// (previous level was top-down scan-free), and a constant.  The two ratios
// carry the direction decision independently of the graph's size (the
// unvisited counts alone grow with the graph, beyond the training graphs').
XBFS_HD inline void xbfs_mlp_features(int32_t L, long long n_f, long long m_f, long long n_u,
                                      long long m_u, int32_t prev_mode, float *x)
{
    const float lnf = xb_log2p1(n_f), lmf = xb_log2p1(m_f);
    x[0] = lnf * (1.0f / 32.0f);
    x[1] = lmf * (1.0f / 32.0f);
    x[2] = (lmf - xb_log2p1(m_u)) * (1.0f / 32.0f);
    x[3] = (lnf - xb_log2p1(n_u)) * (1.0f / 32.0f);
    x[4] = (float)(L < 64 ? L : 64) * (1.0f / 16.0f);
    x[5] = prev_mode == 2 ? 1.0f : (prev_mode < 0 ? -1.0f : 0.0f);
    x[6] = prev_mode == 0 ? 1.0f : 0.0f;
    x[7] = 1.0f;
}

XBFS_HD inline void xbfs_mlp_forward1(const float *p, const float *x, float *y)
{
    const float *W1 = p, *b1 = W1 + XB_MLP_NH * XB_MLP_NI;
    const float *W2 = b1 + XB_MLP_NH, *b2 = W2 + XB_MLP_NH * XB_MLP_NH;
    const float *W3 = b2 + XB_MLP_NH, *b3 = W3 + XB_MLP_NO * XB_MLP_NH;
    float h1[XB_MLP_NH], h2[XB_MLP_NH];
    for (int o = 0; o < XB_MLP_NH; o++) {
        float s = b1[o];
        for (int i = 0; i < XB_MLP_NI; i++) s += W1[o * XB_MLP_NI + i] * x[i];
        h1[o] = s > 0.0f ? s : 0.0f;
    }
    for (int o = 0; o < XB_MLP_NH; o++) {
        float s = b2[o];
        for (int i = 0; i < XB_MLP_NH; i++) s += W2[o * XB_MLP_NH + i] * h1[i];
        h2[o] = s > 0.0f ? s : 0.0f;
    }
    for (int o = 0; o < XB_MLP_NO; o++) {
        float s = b3[o];
        for (int i = 0; i < XB_MLP_NH; i++) s += W3[o * XB_MLP_NH + i] * h2[i];
        y[o] = s;
    }
}

// Mode for level L under `policy` (XBFS_POLICY_* values 0..5).  The
// reference heuristic state always advances, so *htag reports the reference's
// tag for this level whatever the policy; *used is the policy that decided.
// pred: the learned model's three predicted costs for this level (NULL when
// no model is loaded); LEARNED falls back to HEURISTIC without a model or on
// a non-finite prediction.
XBFS_HD inline int32_t xbfs_choose_mode(int32_t policy, const int8_t *sched, int32_t sched_len,
                                        const float *pred, XbfsHeurState *hs, int32_t L,
                                        long long m_f, long long E, float alpha,
                                        int32_t *htag, int32_t *used)
{
    int32_t sub = 0;
    const int32_t t = xbfs_heur_decide(hs, L, m_f, E, alpha, &sub);
    *htag = t;
    *used = policy;
    switch (policy) {
    case 2: return 0;
    case 3: return 1;
    case 4: return 2;
    case 5: return sched[L < sched_len ? L : sched_len - 1];
    case 1:
        if (pred && L <= XBFS_MLP_MAX_LEVEL) {
            bool finite = true;
            for (int j = 0; j < XB_MLP_NO; j++)
                if (!(pred[j] > -1e30f && pred[j] < 1e30f)) finite = false;
            if (finite) {
                int32_t best = 0;
                for (int j = 1; j < XB_MLP_NO; j++)
                    if (pred[j] < pred[best]) best = j;
                return best;
            }
        }
        *used = 0;
        return xbfs_tag_to_mode(t);
    default: *used = 0; return xbfs_tag_to_mode(t);
    }
}

#endif
