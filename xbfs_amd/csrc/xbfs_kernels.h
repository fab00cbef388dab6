// Kernel argument blocks and launch declarations shared by the HIP
// translation units of libxbfs.
#ifndef XBFS_KERNELS_H
#define XBFS_KERNELS_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_device.h"

#define XB_TILE_WORDS 64                     /* 64-bit words per bottom-up tile */
#define XB_TILE_VERTS (XB_TILE_WORDS * 64)   /* 4096 vertices per tile          */
#define XB_SCAN_CHUNK 4096                   /* elements per block, generic scan */

struct XbBuCountArgs {
    const int32_t *level;
    const uint64_t *nz;
    int64_t n, nwords;
    int32_t L;
    uint64_t *fb, *nb, *ub;
    int32_t *tile_cnt;
};

#endif
