// text_to_bin: the reference's CSR converter command line
// (CSRGenerator/tuple_text_to_bin.cpp:59-72):
//   ./text_to_bin tuple_file reverse(1 reverse, 0 not) lines_to_skip thread_count
// writes <tuple_file>_beg_pos.bin (int64[n+1]), <tuple_file>_csr.bin and
// <tuple_file>_weight.bin (int32[m] each: the layout ./bfs and graph.hpp load)
// through xbfs_graph_from_tuples / xbfs_graph_write.  thread_count is
// accepted for compatibility; the conversion is one pass over memory.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "xbfs.h"

int main(int argc, char **argv)
{
    std::cout << "Input: ./exe tuple_file(text) "
              << "reverse_the_edge(1 reverse, 0 not reverse) lines_to_skip thread_count\n";
    if (argc < 5) {
        printf("Wrong input\n");
        return 1;
    }
    const int reverse = atol(argv[2]) == 1;
    const long skip = atol(argv[3]);
    const auto t0 = std::chrono::steady_clock::now();
    xbfs_host_graph g;
    int rc = xbfs_graph_from_tuples(argv[1], reverse, skip, &g);
    if (rc != XBFS_OK) {
        std::cout << argv[1] << ": " << xbfs_strerror(rc) << "\n";
        return 1;
    }
    int64_t vmin = g.n - 1;
    for (int64_t j = 0; j < g.m; j++)
        if (g.adj[j] < vmin) vmin = g.adj[j];
    for (int64_t v = 0; v < g.n; v++)
        if (g.beg_pos[v + 1] > g.beg_pos[v]) {
            if (v < vmin) vmin = v;
            break;
        }
    std::cout << "edge count: " << g.m << std::endl;
    std::cout << "max vertex id: " << g.n - 1 << std::endl;
    std::cout << "min vertex id: " << vmin << std::endl;
    std::cout << "vert count: " << g.n << std::endl;
    rc = xbfs_graph_write(argv[1], &g);
    xbfs_graph_free(&g);
    if (rc != XBFS_OK) {
        std::cout << "writing " << argv[1] << "_*.bin: " << xbfs_strerror(rc) << "\n";
        return 1;
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf("%f seconds elapsed\n\n", s);
    return 0;
}
