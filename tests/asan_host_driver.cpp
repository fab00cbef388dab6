// Host-side checks of libxbfs's graph code (graph_host.cpp) built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_asan.py:
// the generators, the binary loader and writer, the text edge-list reader,
// the CSR validation, the concatenation and the degree statistics, on valid
// inputs and on the malformed inputs the reference's own loader asserts on
// (graph.hpp:32,37,61-62).  Every call must return the expected code without
// a sanitizer report.  Argument: a scratch directory.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "xbfs.h"
#include "xbfs_host_internal.h"

static int failures = 0;

#define EXPECT(cond)                                                        \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
            failures++;                                                     \
        }                                                                   \
    } while (0)

static void write_bytes(const std::string &path, const void *p, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::printf("FAIL cannot write %s\n", path.c_str());
        std::exit(2);
    }
    if (n) std::fwrite(p, 1, n, f);
    std::fclose(f);
}

static bool same_graph(const xbfs_host_graph &a, const xbfs_host_graph &b)
{
    return a.n == b.n && a.m == b.m &&
           std::memcmp(a.beg_pos, b.beg_pos, sizeof(int64_t) * (size_t)(a.n + 1)) == 0 &&
           (a.m == 0 || std::memcmp(a.adj, b.adj, sizeof(int32_t) * (size_t)a.m) == 0);
}

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    const std::string dir = argv[1];

    // generators, write -> load round trip, degree statistics
    xbfs_host_graph r = {}, back = {};
    EXPECT(xbfs_gen_rmat(10, 8, 0.57, 0.19, 0.19, 7, 1, &r) == XBFS_OK);
    EXPECT(xbfs_check_csr(r.n, r.m, r.beg_pos, r.adj) == XBFS_OK);
    const std::string pre = dir + "/rmat10";
    EXPECT(xbfs_graph_write(pre.c_str(), &r) == XBFS_OK);
    EXPECT(xbfs_graph_load((pre + "_beg_pos.bin").c_str(), (pre + "_csr.bin").c_str(),
                           (pre + "_weight.bin").c_str(), &back) == XBFS_OK);
    EXPECT(same_graph(r, back));
    EXPECT(back.weight == nullptr);
    xbfs_degree_info di;
    EXPECT(xbfs_degree_stats(&r, &di) == XBFS_OK);
    EXPECT(di.small + di.medium + di.large == r.n);
    xbfs_graph_free(&back);

    xbfs_host_graph e = {}, gr = {};
    EXPECT(xbfs_gen_er(500, 2000, 3, &e) == XBFS_OK);
    EXPECT(xbfs_check_csr(e.n, e.m, e.beg_pos, e.adj) == XBFS_OK);
    EXPECT(xbfs_gen_grid(7, 9, &gr) == XBFS_OK);
    EXPECT(gr.n == 63 && gr.m == 2 * (6 * 9 + 7 * 8));

    // concatenation: offsets and the packed graph
    xbfs_host_graph parts[3] = {r, e, gr}, cat = {};
    int64_t off[4];
    EXPECT(xbfs_graph_concat(parts, 3, &cat, off) == XBFS_OK);
    EXPECT(off[0] == 0 && off[1] == r.n && off[2] == r.n + e.n && off[3] == cat.n);
    EXPECT(cat.m == r.m + e.m + gr.m);
    EXPECT(xbfs_check_csr(cat.n, cat.m, cat.beg_pos, cat.adj) == XBFS_OK);
    EXPECT(cat.adj[cat.beg_pos[off[2]]] == gr.adj[0] + (int32_t)off[2]);
    xbfs_graph_free(&cat);
    EXPECT(xbfs_graph_concat(parts, 0, &cat, off) != XBFS_OK);

    // loader: missing, truncated and malformed files
    xbfs_host_graph bad = {};
    EXPECT(xbfs_graph_load((dir + "/none_beg.bin").c_str(), (pre + "_csr.bin").c_str(), nullptr,
                           &bad) == XBFS_ERR_IO);
    // csr shorter than beg_pos[n]
    write_bytes(dir + "/short_csr.bin", r.adj, sizeof(int32_t) * (size_t)(r.m - 1));
    EXPECT(xbfs_graph_load((pre + "_beg_pos.bin").c_str(), (dir + "/short_csr.bin").c_str(), nullptr,
                           &bad) != XBFS_OK);
    // beg_pos file of 1 entry (n = 0) and of a size that is not a multiple of 8
    const int64_t one = 0;
    write_bytes(dir + "/one_beg.bin", &one, sizeof one);
    const int32_t none = 0;
    write_bytes(dir + "/empty_csr.bin", &none, 0);
    int rc = xbfs_graph_load((dir + "/one_beg.bin").c_str(), (dir + "/empty_csr.bin").c_str(),
                             nullptr, &bad);
    EXPECT(rc == XBFS_OK || rc == XBFS_ERR_FORMAT);
    if (rc == XBFS_OK) xbfs_graph_free(&bad);
    write_bytes(dir + "/odd_beg.bin", r.beg_pos, 8 * 3 + 5);
    EXPECT(xbfs_graph_load((dir + "/odd_beg.bin").c_str(), (pre + "_csr.bin").c_str(), nullptr,
                           &bad) != XBFS_OK);
    // non-monotone offsets, beg_pos[0] != 0, adjacency id out of range
    std::vector<int64_t> beg(r.beg_pos, r.beg_pos + r.n + 1);
    std::vector<int32_t> adj(r.adj, r.adj + r.m);
    std::swap(beg[5], beg[6]);
    if (beg[5] != beg[6]) EXPECT(xbfs_check_csr(r.n, r.m, beg.data(), adj.data()) == XBFS_ERR_FORMAT);
    std::swap(beg[5], beg[6]);
    beg[0] = 1;
    EXPECT(xbfs_check_csr(r.n, r.m, beg.data(), adj.data()) == XBFS_ERR_FORMAT);
    beg[0] = 0;
    adj[3] = (int32_t)r.n;
    EXPECT(xbfs_check_csr(r.n, r.m, beg.data(), adj.data()) == XBFS_ERR_FORMAT);
    adj[3] = -1;
    EXPECT(xbfs_check_csr(r.n, r.m, beg.data(), adj.data()) == XBFS_ERR_FORMAT);
    write_bytes(dir + "/neg_csr.bin", adj.data(), sizeof(int32_t) * adj.size());
    EXPECT(xbfs_graph_load((pre + "_beg_pos.bin").c_str(), (dir + "/neg_csr.bin").c_str(), nullptr,
                           &bad) == XBFS_ERR_FORMAT);

    // text edge lists: comments, blank lines, weights, self-loops, duplicates
    const char *txt = "# comment\n% other comment\n\n0 1\n1 2 7\n2 2\n1 0\n3 1 0.5\n  4   0\n";
    write_bytes(dir + "/edges.txt", txt, std::strlen(txt));
    xbfs_host_graph t = {};
    EXPECT(xbfs_graph_from_edge_text((dir + "/edges.txt").c_str(), 1, &t) == XBFS_OK);
    EXPECT(t.n == 5);
    EXPECT(xbfs_check_csr(t.n, t.m, t.beg_pos, t.adj) == XBFS_OK);
    EXPECT(t.m == 8); // {0-1, 1-2, 1-3, 0-4} both ways
    xbfs_graph_free(&t);
    const char *junk = "0 1\nabc def\n";
    write_bytes(dir + "/junk.txt", junk, std::strlen(junk));
    rc = xbfs_graph_from_edge_text((dir + "/junk.txt").c_str(), 1, &t);
    if (rc == XBFS_OK) xbfs_graph_free(&t);
    EXPECT(rc == XBFS_OK || rc == XBFS_ERR_FORMAT);
    EXPECT(xbfs_graph_from_edge_text((dir + "/missing.txt").c_str(), 1, &t) == XBFS_ERR_IO);

    // reference-semantics tuples: header lines, duplicates, self-loops kept,
    // a lone trailing id, no final newline; then written with weights and loaded
    const char *tup = "h1\nh2\n0 1\n1\t2\n0 1\n2 2\r\n3 0\n7";
    write_bytes(dir + "/tuples.txt", tup, std::strlen(tup));
    EXPECT(xbfs_graph_from_tuples((dir + "/tuples.txt").c_str(), 1, 2, &t) == XBFS_OK);
    EXPECT(t.n == 8 && t.m == 10 && t.weight != nullptr);
    EXPECT(xbfs_check_csr(t.n, t.m, t.beg_pos, t.adj) == XBFS_OK);
    EXPECT(xbfs_graph_write((dir + "/tw").c_str(), &t) == XBFS_OK);
    xbfs_graph_free(&t);
    EXPECT(xbfs_graph_load((dir + "/tw_beg_pos.bin").c_str(), (dir + "/tw_csr.bin").c_str(),
                           (dir + "/tw_weight.bin").c_str(), &t) == XBFS_OK);
    EXPECT(t.m == 10 && t.weight != nullptr);
    xbfs_graph_free(&t);
    EXPECT(xbfs_graph_from_tuples((dir + "/tuples.txt").c_str(), 0, 9, &t) == XBFS_ERR_FORMAT);
    EXPECT(xbfs_graph_from_tuples((dir + "/junk.txt").c_str(), 0, 0, &t) == XBFS_ERR_FORMAT);
    EXPECT(xbfs_graph_from_tuples((dir + "/missing.txt").c_str(), 0, 0, &t) == XBFS_ERR_IO);
    const char *empty = "only a header\n";
    write_bytes(dir + "/empty.txt", empty, std::strlen(empty));
    EXPECT(xbfs_graph_from_tuples((dir + "/empty.txt").c_str(), 0, 1, &t) == XBFS_ERR_FORMAT);

    xbfs_graph_free(&r);
    xbfs_graph_free(&e);
    xbfs_graph_free(&gr);
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
