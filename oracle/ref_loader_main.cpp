// Driver of our own around the reference's graph loader, compiled from the
// reference tree by oracle/Makefile (target "ref") into oracle/_ref/.  It
// loads a graph with graph<int,long,int,int,long,char> (graph.hpp:4-104) and
// calls workload_gap1 (workload_gap.cuh:27-38) on the degree array, then prints
//   n m max_degree printed_vertex fnv1a64(beg_pos) fnv1a64(csr)
// so tests can compare our loader and degree statistics with the reference's.
#include <cstdint>
#include <cstdio>
#include <sstream>
#include <iostream>
#include "graph.h"
#include "workload_gap.cuh"

static uint64_t fnv(const void *p, size_t bytes)
{
    const unsigned char *c = (const unsigned char *)p;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < bytes; i++) { h ^= c[i]; h *= 1099511628211ull; }
    return h;
}

int main(int argc, char **argv)
{
    if (argc != 4) { fprintf(stderr, "usage: ref_loader beg csr weight\n"); return 1; }
    std::streambuf *saved = std::cout.rdbuf();
    std::ostringstream sink;
    std::cout.rdbuf(sink.rdbuf()); // the reference prints progress lines
    graph<int, long, int, int, long, char> *g =
        new graph<int, long, int, int, long, char>(argv[1], argv[2], argv[3]);
    long *deg = new long[g->vert_count];
    for (long i = 0; i < g->vert_count; i++) deg[i] = g->beg_pos[i + 1] - g->beg_pos[i];
    long mx = workload_gap1(deg, (int)g->vert_count);
    std::string printed = sink.str();
    std::cout.rdbuf(saved);
    long vid = -1;
    size_t at = printed.find("vertexID = ");
    if (at != std::string::npos) vid = atol(printed.c_str() + at + 11);
    printf("%ld %ld %ld %ld %llu %llu\n", (long)g->vert_count, (long)g->edge_count, mx, vid,
           (unsigned long long)fnv(g->beg_pos, sizeof(long) * (g->vert_count + 1)),
           (unsigned long long)fnv(g->csr, sizeof(int) * g->edge_count));
    return 0;
}
