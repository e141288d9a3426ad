// Device-side constant tables and shared launch helpers of the SO(4)/SNAP
// kernels (host declarations; definitions in so4_kernels.hip).
#pragma once
#include "pxff_common.h"

namespace pxff {

// Per-(device, 2J, normalisation) constant tables in device memory, built
// once from So4HostTables and kept until process exit.
struct So4Dev {
    int twol, nu, nhalf, nz, nb, ncg, nblk;
    double *cg;
    int *blk_level_start, *blk_zoff, *blk_tb;
    double *blk_tf;
    int *b_zoff, *b_level;
    int *b_src;       // [nb*3] blocks targeting b (-1 = none)
    double *b_srcf;   // [nb*3] their factors
    int *zh;          // [nz] half-U element paired with each Z element
    double *hscale;   // [nhalf] 2 hw(h) u_norm(l(h)): Y scale per half element
    double *rootpq, *lnorm;
    int *hoff, *h_level, *h_diag;
    double *h_weight;
    int *idxu_block;
    int nzgroups;
    int *zgroups;
    int nbrow;        // B row partials per atom
    int *b_brow0;     // [nb] first B row partial of b
    int zlevel_max;   // largest level of Z blocks: sum of its blocks' half-row entries
    int nytask;       // Y tasks: (level, run of half elements) per wavefront
    int *ytask;       // [nytask * YTASK_STRIDE]
};

// A Y task is {l, loc0, count <= YTASK_ROWS, first block, end block, h0}: the
// half elements h0 .. h0 + count - 1 (locations loc0.. of level l) summed
// over the blocks [first, end) of level l.
constexpr int YTASK_ROWS = 4;
constexpr int YTASK_STRIDE = 6;

int get_so4_tables(int twol, int norm, const So4Dev **out);
int so4_check_params(const pxff_so4_params *p);
int num_cus();
int tune_value(int knob);

__host__ __device__ constexpr int nhalf_of(int twoj) {
    int s = 0;
    for (int l = 0; l <= twoj; l++) s += (l / 2 + 1) * (l + 1);
    return s;
}
// Half-row entries of the largest level of Z blocks: sum over the blocks
// (l1, l2, l) of level l (l1 >= l2, |l1 - l2| <= l <= l1 + l2, l1 + l2 - l
// even) of (l/2 + 1)(l + 1), maximised over l.
__host__ __device__ constexpr int zlevel_max_of(int twoj) {
    int best = 0;
    for (int l = 0; l <= twoj; l++) {
        int nbl = 0;
        for (int l1 = 0; l1 <= twoj; l1++)
            for (int l2 = 0; l2 <= l1; l2++)
                if (l >= l1 - l2 && l <= l1 + l2 && ((l1 + l2 - l) & 1) == 0) nbl++;
        const int n = nbl * (l / 2 + 1) * (l + 1);
        if (n > best) best = n;
    }
    return best;
}
__host__ __device__ constexpr int nb_of(int twoj) {
    int s = 0;
    for (int l1 = 0; l1 <= twoj; l1++)
        for (int l2 = 0; l2 <= l1; l2++)
            for (int l = l1 - l2; l <= (twoj < l1 + l2 ? twoj : l1 + l2); l += 2)
                if (l >= l1) s++;
    return s;
}
__host__ __device__ constexpr int hoff_of(int l) {
    int s = 0;
    for (int k = 0; k < l; k++) s += (k / 2 + 1) * (k + 1);
    return s;
}

#define PXFF_DISPATCH_TWOJ(twol, ...)                                             \
    switch (twol) {                                                               \
    case 0: { constexpr int TJ = 0; __VA_ARGS__; } break;                         \
    case 2: { constexpr int TJ = 2; __VA_ARGS__; } break;                         \
    case 4: { constexpr int TJ = 4; __VA_ARGS__; } break;                         \
    case 6: { constexpr int TJ = 6; __VA_ARGS__; } break;                         \
    case 8: { constexpr int TJ = 8; __VA_ARGS__; } break;                         \
    case 10: { constexpr int TJ = 10; __VA_ARGS__; } break;                       \
    default: ::pxff::set_error("unsupported 2J"); return PXFF_ENOTIMPL;           \
    }

}  // namespace pxff
