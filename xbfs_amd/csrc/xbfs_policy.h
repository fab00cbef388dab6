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

#endif
