// Internal declarations shared by the host C++ and HIP translation units of
// libxbfs.  Not part of the public C ABI.
#ifndef XBFS_HOST_INTERNAL_H
#define XBFS_HOST_INTERNAL_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
// Checks beg[0]==0, beg[n]==m, beg non-decreasing, 0 <= adj[j] < n.
int xbfs_check_csr(int64_t n, int64_t m, const int64_t *beg, const int32_t *adj);
#ifdef __cplusplus
}
#endif
#endif
