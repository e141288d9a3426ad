// Shared host/device helpers of the pxff library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/pxff.h"

namespace pxff {

void set_error(const std::string &msg);

// A failed call is reported once: the runtime's last-error slot is cleared
// so that a later launch check does not report it again.
#define PXFF_HIP_CHECK(expr)                                                                  \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            (void)hipGetLastError();                                                          \
            ::pxff::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return PXFF_EHIP;                                                                 \
        }                                                                                     \
    } while (0)

#define PXFF_LAUNCH_CHECK()                                                                   \
    do {                                                                                      \
        hipError_t _e = hipGetLastError();                                                    \
        if (_e != hipSuccess) {                                                               \
            ::pxff::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));       \
            return PXFF_EHIP;                                                                 \
        }                                                                                     \
    } while (0)

#define PXFF_REQUIRE(cond, msg)                                                               \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            ::pxff::set_error(msg);                                                           \
            return PXFF_EINVAL;                                                               \
        }                                                                                     \
    } while (0)

// ---------------------------------------------------------------- complex
struct cd {
    double re, im;
};

__host__ __device__ inline cd cmake(double r, double i) { return cd{r, i}; }
__host__ __device__ inline cd cadd(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cd csub(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cd cmul(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline cd cscale(cd a, double s) { return cd{a.re * s, a.im * s}; }
// a * b + c (complex multiply-add; four fused multiply-adds)
__host__ __device__ inline cd cmadd(cd a, cd b, cd c) {
    return cd{fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
}
__host__ __device__ inline cd cconj(cd a) { return cd{a.re, -a.im}; }
// Re(conj(a) * b)
__host__ __device__ inline double cdotre(cd a, cd b) { return a.re * b.re + a.im * b.im; }

// ---------------------------------------------------------------- cutoffs
// SNAP.py:336-440 (compute_sfac / compute_dsfac, SNAP.py:512-588): 0 for r > rc.
__device__ inline void cutoff_fdf(int id, double r, double rc, double &f, double &df) {
    if (r > rc) {
        f = 0.0;
        df = 0.0;
        return;
    }
    double x = r / rc;
    switch (id) {
    case 1: {
        double s, c;
        sincos(M_PI * r / rc, &s, &c);
        f = 0.5 * (c + 1.0);
        df = -0.5 * M_PI / rc * s;
        break;
    }
    case 2: {
        double t = tanh(1.0 - x);
        double t2 = t * t;
        f = t2 * t;
        df = -(3.0 / rc) * t2 * (1.0 - t2);
        break;
    }
    case 3:
        f = x * x * (2 * x - 3) + 1;
        df = (6.0 / (rc * rc)) * r * (x - 1.0);
        break;
    case 4:
        f = x * x * x * (x * (15 - 6 * x) - 10) + 1;
        df = (-30.0 / rc) * (x * x * (x - 1) * (x - 1));
        break;
    case 5:
        f = x * x * x * x * (x * (x * (20 * x - 70) + 84) - 35) + 1;
        df = (140.0 / rc) * (x * x * x * (x - 1) * (x - 1) * (x - 1));
        break;
    case 6: {
        f = x * x * x * x * x * (x * (x * (x * (315 - 70 * x) - 540) + 420) - 126) + 1;
        double y = x * (x - 1);
        df = (-630.0 / rc) * y * y * y * y;
        break;
    }
    case 7: {
        double e = exp(1.0 - 1.0 / (1.0 - x * x));
        double d = 1 + x * x;
        f = e;
        df = 2 * x * e / (d * d);
        break;
    }
    default:
        f = 0.0;
        df = 0.0;
    }
}

// Cutoff value at r = 0 (addself_uarraytot, SNAP.py:590-627).
inline double cutoff_f0_host(int id) {
    switch (id) {
    case 1: return 1.0;
    case 2: { double t = std::tanh(1.0); return t * t * t; }
    case 3: case 4: case 5: case 6: return 1.0;
    case 7: return 1.0;
    default: return 0.0;
    }
}

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Sum of in[r0, r1) by one workgroup of PXFF_SUM_THREADS threads in a fixed
// order: strided sums with four loads in flight, a butterfly per
// wavefront, then the four wavefront sums in order by thread 0.  Thread 0
// returns the total (the others 0.0).  Every thread of the workgroup must
// call it; it ends with a barrier, so the workgroup may call it again.
constexpr int PXFF_SUM_THREADS = 256;
__device__ inline double block_sum_1col(const double *__restrict__ in, int64_t r0, int64_t r1) {
    __shared__ double wred[PXFF_SUM_THREADS / 64];
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int64_t rb = r0 + tid; rb < r1; rb += 4 * PXFF_SUM_THREADS) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t r = rb + (int64_t)u * PXFF_SUM_THREADS;
            v[u] = r < r1 ? in[r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc += v[u];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((tid & 63) == 0) wred[tid >> 6] = acc;
    __syncthreads();
    double t = 0.0;
    if (tid == 0)
        for (int w = 0; w < PXFF_SUM_THREADS / 64; w++) t += wred[w];
    __syncthreads();
    return t;
}

}  // namespace pxff
