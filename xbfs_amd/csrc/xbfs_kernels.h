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

struct XbTdArgs {
    int32_t *level;
    int32_t *parent;
    const int64_t *beg;
    const int32_t *adj;
    const int32_t *qc0, *qc1, *qc2; // current frontier bins
    const int32_t *cnt_c;           // their lengths (3)
    int32_t *qn0, *qn1, *qn2;       // next frontier bins (scan-free only)
    int32_t *cnt_n;                 // their lengths (3)
    unsigned long long *slot_n1, *slot_m1; // ring slot of level L+1
    int32_t L, t1, t2;
};

struct XbGenArgs {
    const int32_t *level;
    const int64_t *beg;
    int64_t n;
    int32_t L, t1, t2;
    int32_t *q0, *q1, *q2;
    int32_t *cnt;
    unsigned long long *slot_n, *slot_m;
};

struct XbFilterArgs {
    const int32_t *list;
    const int32_t *list_cnt;
    const int32_t *level;
    const int64_t *beg;
    int32_t L, t1, t2, carry;
    int32_t *qc0, *qc1, *qc2;
    int32_t *cnt_c;
    int32_t *qn0, *qn1, *qn2;
    int32_t *cnt_n;
};

struct XbBuCountArgs {
    const int32_t *level;
    const uint64_t *nz;
    int64_t n, nwords;
    int32_t L;
    uint64_t *fb, *nb, *ub;
    int32_t *tile_cnt;
};

struct XbBuArgs {
    int32_t *level;
    int32_t *parent;
    const int64_t *beg;
    const int32_t *adj;
    const int32_t *list;
    const int32_t *list_cnt;
    const uint64_t *fb;
    uint64_t *nb;
    unsigned long long *slot_n1, *slot_m1, *slot_n2, *slot_m2;
    int32_t L;
};

#endif
