// Per-element network energies and their gradients with respect to the
// descriptors, for energy + force evaluation (no parameter gradients).
// Mirrors the reference's per-element feed-forward networks
// (pyxtal_ff/models/neuralnetwork.py:80-200: nn.Linear layers with one
// activation each, descriptors scaled by the element's drange,
// neuralnetwork.py:938-1029) evaluated forward, then backward to dE/dx.
// Three kernels, by what fits in LDS: 16-atom tiles on the FP64 matrix
// cores (mlp_mfma_kernel, scratch and one element's weights within 160 KB);
// else one wavefront per atom with lane n holding neuron n of a layer
// (widths <= 64), the weights of every element in LDS (mlp_lds_kernel) or
// streamed from L2 (mlp_energy_grad_kernel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>

#include "pxff_common.h"
#include "so4_dev.h"

namespace pxff {

constexpr int MLP_MAXW = 64;     // hidden / output width limit (one lane per neuron)
constexpr int MLP_MAXL = 8;      // layers per element
constexpr int MLP_MAXIN = 1024;  // descriptor length limit
constexpr int MLP_LSTRIDE = 6;   // ints per layer in the layout
constexpr int MLP_STRIDE = 1 + MLP_LSTRIDE * MLP_MAXL;

__device__ inline void act_fdf(int id, double z, double &a, double &da) {
    switch (id) {
    case 1: {  // Tanh
        const double t = tanh(z);
        a = t;
        da = 1.0 - t * t;
        break;
    }
    case 2: {  // Sigmoid
        const double s = 1.0 / (1.0 + exp(-z));
        a = s;
        da = s * (1.0 - s);
        break;
    }
    case 3:  // ReLU
        a = z > 0.0 ? z : 0.0;
        da = z > 0.0 ? 1.0 : 0.0;
        break;
    case 4: {  // Softplus (beta 1, threshold 20, as torch.nn.Softplus)
        if (z > 20.0) {
            a = z;
            da = 1.0;
        } else {
            a = log1p(exp(z));
            da = 1.0 / (1.0 + exp(-z));
        }
        break;
    }
    case 5:  // ELU (alpha 1)
        if (z > 0.0) {
            a = z;
            da = 1.0;
        } else {
            a = expm1(z);
            da = a + 1.0;
        }
        break;
    default:  // Linear
        a = z;
        da = 1.0;
    }
}

// layout per element e (ints, at d_layout[e * MLP_STRIDE]):
//   nlayers, then per layer {nin, nout, act, w_off, b_off, wt_off}: W
//   row-major [nout][nin] as torch.nn.Linear stores it (read by the
//   backward: lanes over inputs, consecutive addresses) and its transpose
//   [nin][nout] (read by the forward: lanes over outputs), offsets into
//   the parameter array.

// Workgroup-wide barrier (one wavefront per workgroup).
struct WgSync {
    __device__ void operator()() const { __syncthreads(); }
};
// The wavefront's own LDS writes are visible to its later reads (LDS
// operations of one wavefront execute in order; the fences keep the
// compiler from moving them).
struct WaveSync {
    __device__ void operator()() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// One atom on one wavefront: lane n holds neuron n of a layer; h0 [nin],
// dact / hout [layers][hw] (hw = the widest layer) are the wavefront's LDS scratch.
template <typename Sync>
__device__ inline void mlp_atom(int64_t i, int lane, int nin, int nelem, const int64_t *__restrict__ species,
                                const int *__restrict__ layout, const double *params, const double *__restrict__ dmin,
                                const double *__restrict__ scale, const double *__restrict__ x,
                                double *__restrict__ e_out, double *__restrict__ beta, double *h0, double *dact,
                                double *hout, int hw, Sync sync) {
    const int64_t el64 = species[i];
    if (el64 < 0 || el64 >= nelem) {  // no network for this atom: NaN results
        if (lane == 0) e_out[i] = __builtin_nan("");
        for (int j = lane; j < nin; j += 64) beta[(size_t)i * nin + j] = __builtin_nan("");
        return;  // whole wavefront
    }
    const int el = (int)el64;
    const int *lay = layout + el * MLP_STRIDE;
    const int nl = lay[0];
    const double *xi = x + (size_t)i * nin;
    const double *dm = dmin + (size_t)el * nin, *sc = scale + (size_t)el * nin;
    for (int j = lane; j < nin; j += 64) h0[j] = (xi[j] - dm[j]) * sc[j];
    sync();
    // forward
    const double *hin = h0;
    for (int k = 0; k < nl; k++) {
        const int *L = lay + 1 + MLP_LSTRIDE * k;
        const int ni = L[0], no = L[1], act = L[2];
        const double *WT = params + L[5], *b = params + L[4];
        if (lane < no) {
            // four partial sums: four independent FMA chains
            double z0 = b[lane], z1 = 0.0, z2 = 0.0, z3 = 0.0;
            int j = 0;
            for (; j + 4 <= ni; j += 4) {
                z0 += WT[(size_t)j * no + lane] * hin[j];
                z1 += WT[(size_t)(j + 1) * no + lane] * hin[j + 1];
                z2 += WT[(size_t)(j + 2) * no + lane] * hin[j + 2];
                z3 += WT[(size_t)(j + 3) * no + lane] * hin[j + 3];
            }
            for (; j < ni; j++) z0 += WT[(size_t)j * no + lane] * hin[j];
            const double z = (z0 + z1) + (z2 + z3);
            double a, da;
            act_fdf(act, z, a, da);
            hout[k * hw + lane] = a;
            dact[k * hw + lane] = da;
        }
        sync();
        hin = hout + k * hw;
    }
    if (lane == 0) e_out[i] = hin[0];
    // backward: dact[k] becomes the pre-activation gradient of layer k; for
    // the output layer (one neuron, dE/dE = 1) it already is act'
    for (int k = nl - 1; k >= 1; k--) {
        const int *L = lay + 1 + MLP_LSTRIDE * k;
        const int ni = L[0], no = L[1];
        const double *W = params + L[3];
        // dE/dh_{k-1}[j] = sum_n W[n][j] g_pre[n]; then times act'_{k-1}
        if (lane < ni) {
            const double *gk = dact + k * hw;
            double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
            int n = 0;
            for (; n + 4 <= no; n += 4) {
                g0 += W[(size_t)n * ni + lane] * gk[n];
                g1 += W[(size_t)(n + 1) * ni + lane] * gk[n + 1];
                g2 += W[(size_t)(n + 2) * ni + lane] * gk[n + 2];
                g3 += W[(size_t)(n + 3) * ni + lane] * gk[n + 3];
            }
            for (; n < no; n++) g0 += W[(size_t)n * ni + lane] * gk[n];
            dact[(k - 1) * hw + lane] *= (g0 + g1) + (g2 + g3);
        }
        sync();
    }
    {
        const int *L = lay + 1;
        const int ni = L[0], no = L[1];
        const double *W = params + L[3];
        for (int j = lane; j < ni; j += 64) {
            double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
            int n = 0;
            for (; n + 4 <= no; n += 4) {
                g0 += W[(size_t)n * ni + j] * dact[n];
                g1 += W[(size_t)(n + 1) * ni + j] * dact[n + 1];
                g2 += W[(size_t)(n + 2) * ni + j] * dact[n + 2];
                g3 += W[(size_t)(n + 3) * ni + j] * dact[n + 3];
            }
            for (; n < no; n++) g0 += W[(size_t)n * ni + j] * dact[n];
            beta[(size_t)i * nin + j] = ((g0 + g1) + (g2 + g3)) * sc[j];
        }
    }
    sync();  // the next atom reuses the scratch
}

// One wavefront (= one workgroup) per atom, weights read from global memory
// (L2).
__global__ __launch_bounds__(64) void mlp_energy_grad_kernel(int64_t natoms, int nin, int nelem,
                                                              const int64_t *__restrict__ species,
                                                              const int *__restrict__ layout,
                                                              const double *__restrict__ params,
                                                              const double *__restrict__ dmin,
                                                              const double *__restrict__ scale,
                                                              const double *__restrict__ x, double *__restrict__ e_out,
                                                              double *__restrict__ beta, int nlm, int hw) {
    extern __shared__ double sm[];
    double *h0 = sm;
    double *dact = h0 + nin;
    double *hout = dact + nlm * hw;
    mlp_atom(blockIdx.x, threadIdx.x, nin, nelem, species, layout, params, dmin, scale, x, e_out, beta, h0, dact, hout,
             hw, WgSync());
}

// Persistent workgroups of MLP_LW wavefronts: the parameters of every
// element network are copied into LDS once per workgroup, then each
// wavefront runs atoms (grid stride) with its own scratch, reading the
// weights from LDS instead of L2.
constexpr int MLP_LW = 4;
__global__ __launch_bounds__(64 * MLP_LW) void mlp_lds_kernel(int64_t natoms, int nin, int nelem,
                                                               const int64_t *__restrict__ species,
                                                               const int *__restrict__ layout,
                                                               const double *__restrict__ params, int nparams,
                                                               const double *__restrict__ dmin,
                                                               const double *__restrict__ scale,
                                                               const double *__restrict__ x,
                                                               double *__restrict__ e_out, double *__restrict__ beta,
                                                               int nlm, int hw) {
    extern __shared__ double sm[];
    double *P = sm;
    // eight loads in flight per thread before their LDS stores
    for (int k0 = threadIdx.x; k0 < nparams; k0 += 8 * blockDim.x) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + u * blockDim.x;
            v[u] = k < nparams ? params[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (k0 + u * (int)blockDim.x < nparams) P[k0 + u * blockDim.x] = v[u];
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *h0 = sm + nparams + (size_t)wv * (nin + 2 * nlm * hw);
    double *dact = h0 + nin;
    double *hout = dact + nlm * hw;
    for (int64_t i = (int64_t)blockIdx.x * MLP_LW + wv; i < natoms; i += (int64_t)gridDim.x * MLP_LW)
        mlp_atom(i, lane, nin, nelem, species, layout, P, dmin, scale, x, e_out, beta, h0, dact, hout, hw,
                 WaveSync());
}

// ---------------------------------------------------------- matrix cores
// 16 atoms per workgroup of four wavefronts; every layer is a GEMM on the
// FP64 matrix cores (v_mfma_f64_16x16x4_f64: A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4),
// one 16-wide column tile per wavefront, activations and their derivatives
// through LDS.  Forward, per 16-neuron tile: Z^T = W_k H_{k-1}^T (A = W_k,
// B = H^T), H_k = act(Z + b_k); backward: G_{k-1} = (G_k W_k) * act'_{k-1}
// (A = G_k, B = W_k), dE/dx = (G_0 W_0) * scale.
// The element's weights are staged in LDS first, zero-padded, and every
// activation row is zero past its width, so the GEMM steps need no bounds
// tests (the K loops run over multiples of 4).  LDS rows are 2 doubles past
// a multiple of 32: the operand reads whose lanes walk down 16 rows fall in
// distinct banks.  The 16 atoms may belong to different elements: the tile
// runs the network of every element present and keeps each atom's own
// results.
typedef double mlp_d4 __attribute__((ext_vector_type(4)));
constexpr int MLP_TILE = 16;

__host__ __device__ inline int mlp_r16(int n) { return (n + 15) & ~15; }
__host__ __device__ inline int mlp_r4(int n) { return (n + 3) & ~3; }
__host__ __device__ inline int mlp_row(int n) { return ((n + 31) & ~31) + 2; }  // LDS row stride

// acc = sum_k A[lane & 15][k] B[k][col], k < K (a multiple of 4):
// a = &A[lane & 15][lane >> 4] (row-major A, k steps of 4 elements),
// b = &B[lane >> 4][col], bstep = 4 x the row stride of B.  Four
// accumulators take the k steps in turn (four independent MFMA chains).
__device__ inline mlp_d4 mlp_gemm_tile(int K, const double *a, const double *b, int bstep) {
    mlp_d4 c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0, c2 = c0, c3 = c0;
    const int S = K >> 2;
    int t = 0;
    for (; t + 4 <= S; t += 4) {
        const double a0 = a[4 * t], a1 = a[4 * t + 4], a2 = a[4 * t + 8], a3 = a[4 * t + 12];
        const double b0 = b[t * bstep], b1 = b[(t + 1) * bstep], b2 = b[(t + 2) * bstep], b3 = b[(t + 3) * bstep];
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a3, b3, c3, 0, 0, 0);
    }
    for (; t < S; t++) c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[4 * t], b[t * bstep], c0, 0, 0, 0);
    return (c0 + c1) + (c2 + c3);
}

__global__ __launch_bounds__(256) void mlp_mfma_kernel(int64_t natoms, int nin, int nelem,
                                                        const int64_t *__restrict__ species,
                                                        const int *__restrict__ layout,
                                                        const double *__restrict__ params,
                                                        const double *__restrict__ dmin,
                                                        const double *__restrict__ scale,
                                                        const double *__restrict__ x, double *__restrict__ e_out,
                                                        double *__restrict__ beta, int nlm, int hw) {
    extern __shared__ double sm[];
    __shared__ int s_sp[MLP_TILE];  // element of each atom; -1 no network, -2 past the end
    __shared__ int s_wo[MLP_MAXL];  // offset of layer k's padded W, then its b, in wl
    const int P0 = mlp_row(nin), HP = mlp_row(hw);
    double *h0 = sm;                                 // [16][P0] scaled inputs, zero-padded
    double *hact = h0 + MLP_TILE * P0;               // [nlm][16][HP] activations
    double *hder = hact + (size_t)nlm * MLP_TILE * HP;  // [nlm][16][HP] act', then gradients
    double *wl = hder + (size_t)nlm * MLP_TILE * HP;    // W_k [r16(nout)][row(nin)], b_k [r16(nout)]
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int ri = lane & 15, kq = lane >> 4;
    // tiles [t0, t1) of this workgroup, in order: an element's weights stay
    // staged from one tile to the next (atoms come sorted by element)
    const int64_t ntiles = (natoms + MLP_TILE - 1) / MLP_TILE;
    const int64_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
    int staged = -1;  // element whose weights are in wl
    for (int64_t tl = t0; tl < t1; tl++) {
        const int64_t a0 = tl * MLP_TILE;
        __syncthreads();  // the previous tile is done with s_sp and the scratch
        if (tid < MLP_TILE) {
            const int64_t a = a0 + tid;
            int e = -2;
            if (a < natoms) {
                const int64_t sp = species[a];
                e = (sp >= 0 && sp < nelem) ? (int)sp : -1;
            }
            s_sp[tid] = e;
        }
        __syncthreads();
        for (int el = 0; el < nelem; el++) {
            bool any = false;
            for (int t = 0; t < MLP_TILE; t++) any = any || s_sp[t] == el;
            if (!any) continue;  // uniform over the workgroup
            const int *lay = layout + el * MLP_STRIDE;
            const int nl = lay[0];
            const double *dm = dmin + (size_t)el * nin, *sc = scale + (size_t)el * nin;
            // staging: eight loads in flight per thread before their LDS stores
            for (int q0 = tid; q0 < MLP_TILE * P0; q0 += 256 * 8) {
                double v[8];
    #pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int q = q0 + 256 * u;
                    const int t = q / P0, j = q - t * P0;
                    const int64_t a = a0 + t;
                    v[u] = (q < MLP_TILE * P0 && a < natoms && j < nin) ? (x[(size_t)a * nin + j] - dm[j]) * sc[j]
                                                                         : 0.0;
                }
    #pragma unroll
                for (int u = 0; u < 8; u++)
                    if (q0 + 256 * u < MLP_TILE * P0) h0[q0 + 256 * u] = v[u];
            }
            if (el != staged) {  // uniform over the workgroup
                staged = el;
                int wo = 0;
                for (int k = 0; k < nl; k++) {
                    const int *L = lay + 1 + MLP_LSTRIDE * k;
                    const int ni = L[0], no = L[1], R = mlp_r16(no), C = mlp_row(ni);
                    const double *W = params + L[3], *bias = params + L[4];
                    for (int q0 = tid; q0 < R * C + R; q0 += 256 * 8) {
                        double v[8];
    #pragma unroll
                        for (int u = 0; u < 8; u++) {
                            const int q = q0 + 256 * u;
                            v[u] = 0.0;
                            if (q < R * C) {
                                const int r = q / C, c = q - r * C;
                                if (r < no && c < ni) v[u] = W[r * ni + c];
                            } else if (q - R * C < no) {
                                v[u] = bias[q - R * C];
                            }
                        }
    #pragma unroll
                        for (int u = 0; u < 8; u++)
                            if (q0 + 256 * u < R * C + R) wl[wo + q0 + 256 * u] = v[u];
                    }
                    if (tid == 0) s_wo[k] = wo;
                    wo += R * C + R;
                }
            }
            __syncthreads();
            // forward
            const double *hin = h0;
            int ldi = P0;
            for (int k = 0; k < nl; k++) {
                const int *L = lay + 1 + MLP_LSTRIDE * k;
                const int ni = L[0], no = L[1], act = L[2], C = mlp_row(ni);
                const double *W = wl + s_wo[k], *bias = W + mlp_r16(no) * C;
                double *ho = hact + (size_t)k * MLP_TILE * HP, *hd = hder + (size_t)k * MLP_TILE * HP;
                for (int ct = wv; ct * 16 < no; ct += 4) {
                    // D[j][i] = sum_k W[j][k] H[i][k]: A rows = neurons of the
                    // tile, B columns = atoms
                    const mlp_d4 acc = mlp_gemm_tile(mlp_r4(ni), W + (ct * 16 + ri) * C + kq, hin + ri * ldi + kq, 4);
    #pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int j = ct * 16 + kq + 4 * r;  // neuron; the lane's atom is ri
                        double av = 0.0, da = 0.0;          // zero past the layer width
                        if (j < no) act_fdf(act, acc[r] + bias[j], av, da);
                        ho[ri * HP + j] = av;
                        hd[ri * HP + j] = da;
                    }
                }
                __syncthreads();
                hin = ho;
                ldi = HP;
            }
            if (tid < MLP_TILE && s_sp[tid] == el) e_out[a0 + tid] = hin[tid * HP];
            // backward; hder of the output layer (one neuron) already is dE/dz
            for (int k = nl - 1; k >= 1; k--) {
                const int *L = lay + 1 + MLP_LSTRIDE * k;
                const int ni = L[0], no = L[1], C = mlp_row(ni);
                const double *W = wl + s_wo[k];
                const double *g = hder + (size_t)k * MLP_TILE * HP;
                double *gp = hder + (size_t)(k - 1) * MLP_TILE * HP;
                for (int ct = wv; ct * 16 < ni; ct += 4) {
                    const int j = ct * 16 + ri;
                    // B[n][j] = W[n][j]
                    const mlp_d4 acc = mlp_gemm_tile(mlp_r4(no), g + ri * HP + kq, W + kq * C + j, 4 * C);
    #pragma unroll
                    for (int r = 0; r < 4; r++) gp[(kq + 4 * r) * HP + j] *= acc[r];
                }
                __syncthreads();
            }
            {
                const int *L = lay + 1;
                const int no = L[1], C = mlp_row(nin);
                const double *W = wl + s_wo[0];
                for (int ct = wv; ct * 16 < nin; ct += 4) {
                    const int j = ct * 16 + ri;
                    const mlp_d4 acc = mlp_gemm_tile(mlp_r4(no), hder + ri * HP + kq, W + kq * C + j, 4 * C);
                    if (j < nin) {
                        const double sj = sc[j];
    #pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int row = kq + 4 * r;
                            if (s_sp[row] == el) beta[(size_t)(a0 + row) * nin + j] = acc[r] * sj;
                        }
                    }
                }
            }
            __syncthreads();  // the next element reuses the scratch
        }
        // atoms without a network: NaN results
        for (int q = tid; q < MLP_TILE * nin; q += 256) {
            const int t = q / nin;
            if (s_sp[t] == -1) beta[(size_t)(a0 + t) * nin + (q - t * nin)] = __builtin_nan("");
        }
        if (tid < MLP_TILE && s_sp[tid] == -1) e_out[a0 + tid] = __builtin_nan("");
    }
}

}  // namespace pxff

using namespace pxff;

extern "C" int pxff_mlp_energy_grad(int64_t natoms, int nin, int nelem, const int64_t *d_species,
                                    const int *h_layout, const int *d_layout, const double *d_params,
                                    const double *d_dmin, const double *d_scale, const double *d_x, double *d_e,
                                    double *d_beta, void *stream) {
    PXFF_REQUIRE(nin > 0 && nin <= MLP_MAXIN, "descriptor length must be in 1..1024");
    PXFF_REQUIRE(nelem > 0 && h_layout != nullptr, "need at least one element network");
    for (int e = 0; e < nelem; e++) {
        const int *lay = h_layout + e * MLP_STRIDE;
        const int nl = lay[0];
        PXFF_REQUIRE(nl >= 1 && nl <= MLP_MAXL, "1..8 layers per element network");
        int prev = nin;
        for (int k = 0; k < nl; k++) {
            const int *L = lay + 1 + MLP_LSTRIDE * k;
            PXFF_REQUIRE(L[0] == prev, "layer input width does not match the previous layer");
            PXFF_REQUIRE(L[1] >= 1 && L[1] <= MLP_MAXW, "layer widths must be in 1..64");
            PXFF_REQUIRE(L[2] >= 0 && L[2] <= 5, "unknown activation id");
            prev = L[1];
        }
        PXFF_REQUIRE(prev == 1, "the last layer must have one output (the atomic energy)");
    }
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    // parameter array length: the furthest end of any W, b or W^T
    int64_t nparams = 0;
    for (int e = 0; e < nelem; e++) {
        const int *lay = h_layout + e * MLP_STRIDE;
        for (int k = 0; k < lay[0]; k++) {
            const int *L = lay + 1 + MLP_LSTRIDE * k;
            const int64_t nw = (int64_t)L[0] * L[1];
            nparams = std::max({nparams, L[3] + nw, (int64_t)L[4] + L[1], L[5] + nw});
        }
    }
    // scratch per wavefront: the scaled input, then activations and their
    // derivatives of every layer at a stride of the widest layer
    int nlm = 1, hw = 1;
    for (int e = 0; e < nelem; e++) {
        const int *lay = h_layout + e * MLP_STRIDE;
        nlm = std::max(nlm, lay[0]);
        for (int k = 0; k < lay[0]; k++) hw = std::max(hw, lay[1 + MLP_LSTRIDE * k + 1]);
    }
    const size_t scratch = sizeof(double) * ((size_t)nin + 2 * (size_t)nlm * hw);
    // matrix cores: 16 atoms per workgroup, their scratch and one element's
    // zero-padded weights and biases within the 160 KB of a CU
    size_t wmax = 0;
    for (int e = 0; e < nelem; e++) {
        const int *lay = h_layout + e * MLP_STRIDE;
        size_t w = 0;
        for (int k = 0; k < lay[0]; k++) {
            const int *L = lay + 1 + MLP_LSTRIDE * k;
            w += (size_t)mlp_r16(L[1]) * mlp_row(L[0]) + mlp_r16(L[1]);
        }
        wmax = std::max(wmax, w);
    }
    const size_t ldsm =
        sizeof(double) * (MLP_TILE * ((size_t)mlp_row(nin) + 2 * (size_t)nlm * mlp_row(hw)) + wmax);
    // above 64 KB (long descriptors, e.g. SO3 6/6: 147 inputs) one workgroup
    // per CU; the kernel's static LDS (a few hundred bytes) comes on top
    if (ldsm <= 160 * 1024 - 1024) {
        // dynamic LDS limit set so far; raised under a lock so that
        // concurrent callers never lower it below what another relies on
        static std::mutex mfma_attr_mu;
        static int mfma_dyn_max = 65536;
        {
            std::lock_guard<std::mutex> lock(mfma_attr_mu);
            if ((int)ldsm > mfma_dyn_max) {
                PXFF_HIP_CHECK(hipFuncSetAttribute((const void *)mlp_mfma_kernel,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsm));
                mfma_dyn_max = (int)ldsm;
            }
        }
        // persistent: as many workgroups as the LDS keeps resident (at most
        // 8 per CU), each over a contiguous range of 16-atom tiles
        const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(8, (int64_t)(160 * 1024) / (int64_t)(ldsm + 1024)));
        const int64_t nwg = std::min<int64_t>(div_up(natoms, MLP_TILE), (int64_t)num_cus() * per_cu);
        mlp_mfma_kernel<<<(unsigned)nwg, 256, ldsm, s>>>(
            natoms, nin, nelem, d_species, d_layout, d_params, d_dmin, d_scale, d_x, d_e, d_beta, nlm, hw);
        PXFF_LAUNCH_CHECK();
        return 0;
    }
    const size_t ldsp = sizeof(double) * (size_t)nparams + MLP_LW * scratch;
    if (ldsp <= 160 * 1024) {  // the weights fit in LDS next to the scratch
        static std::atomic<int> attr_set{0};  // the limit is one value: setting it twice is harmless
        if (ldsp > 65536 && !attr_set.load(std::memory_order_relaxed)) {
            PXFF_HIP_CHECK(hipFuncSetAttribute((const void *)mlp_lds_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr_set.store(1, std::memory_order_relaxed);
        }
        // as many resident workgroups as the LDS holds (at most 4 per CU)
        const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(4, (int64_t)(160 * 1024) / (int64_t)ldsp));
        const int64_t nwg = std::min<int64_t>((natoms + MLP_LW - 1) / MLP_LW, (int64_t)num_cus() * per_cu);
        mlp_lds_kernel<<<(unsigned)nwg, 64 * MLP_LW, ldsp, s>>>(natoms, nin, nelem, d_species, d_layout, d_params,
                                                               (int)nparams, d_dmin, d_scale, d_x, d_e, d_beta, nlm,
                                                               hw);
        PXFF_LAUNCH_CHECK();
        return 0;
    }
    const size_t lds = scratch;
    if (lds > 65536)
        PXFF_HIP_CHECK(hipFuncSetAttribute((const void *)mlp_energy_grad_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    PXFF_REQUIRE(natoms <= 0x7fffffff, "too many atoms for one launch");
    mlp_energy_grad_kernel<<<(unsigned)natoms, 64, lds, s>>>(natoms, nin, nelem, d_species, d_layout, d_params, d_dmin,
                                                               d_scale, d_x, d_e, d_beta, nlm, hw);
    PXFF_LAUNCH_CHECK();
    return 0;
}
