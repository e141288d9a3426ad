// Segmented sums over contiguous row ranges (per-structure energies, forces
// statistics and virials of a batch of structures).  Replaces the
// per-structure accumulation of the reference's network
// (pyxtal_ff/models/neuralnetwork.py:603-677, energies and stresses summed
// over the atoms of each structure).  Rows of one segment are reduced by one
// workgroup in a fixed tree order: deterministic, no atomics.
#include <hip/hip_runtime.h>

#include "pxff_common.h"

namespace pxff {

constexpr int SEG_THREADS = PXFF_SUM_THREADS;
constexpr int SEG_COLS = 16;  // columns reduced per pass

__global__ __launch_bounds__(SEG_THREADS) void segment_sum_kernel(int64_t nseg, int ncol,
                                                                  const int64_t *__restrict__ seg,
                                                                  const double *__restrict__ in,
                                                                  double *__restrict__ out) {
    __shared__ double red[SEG_THREADS][SEG_COLS + 1];
    const int tid = threadIdx.x;
    if (ncol == 1) {
        // one column (energies): block_sum_1col per segment (fixed order)
        for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
            const double t = block_sum_1col(in, seg[s], seg[s + 1]);
            if (tid == 0) out[s] = t;
        }
        return;
    }
    for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const int64_t r0 = seg[s], r1 = seg[s + 1];
        for (int c0 = 0; c0 < ncol; c0 += SEG_COLS) {
            const int nc = ncol - c0 < SEG_COLS ? ncol - c0 : SEG_COLS;
            double acc[SEG_COLS];
#pragma unroll
            for (int c = 0; c < SEG_COLS; c++) acc[c] = 0.0;
            for (int64_t r = r0 + tid; r < r1; r += SEG_THREADS) {
                const double *row = in + (size_t)r * ncol + c0;
#pragma unroll
                for (int c = 0; c < SEG_COLS; c++)
                    if (c < nc) acc[c] += row[c];
            }
#pragma unroll
            for (int c = 0; c < SEG_COLS; c++) red[tid][c] = acc[c];
            __syncthreads();
            for (int stride = SEG_THREADS / 2; stride > 0; stride >>= 1) {
                if (tid < stride) {
#pragma unroll
                    for (int c = 0; c < SEG_COLS; c++) red[tid][c] += red[tid + stride][c];
                }
                __syncthreads();
            }
            if (tid < nc) out[(size_t)s * ncol + c0 + tid] = red[0][tid];
            __syncthreads();
        }
    }
}

}  // namespace pxff

using namespace pxff;

extern "C" int pxff_segment_sum(int64_t nseg, int ncol, const int64_t *d_seg_start, const double *d_in,
                                double *d_out, void *stream) {
    PXFF_REQUIRE(nseg >= 0 && ncol >= 0, "negative size");
    if (nseg == 0 || ncol == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int grid = (int)(nseg < 65536 ? nseg : 65536);
    segment_sum_kernel<<<grid, SEG_THREADS, 0, s>>>(nseg, ncol, d_seg_start, d_in, d_out);
    PXFF_LAUNCH_CHECK();
    return 0;
}
