// Z of the SO(4)/SNAP bispectrum (compute_zi, SNAP.py:1050-1098) for gfx950,
// one atom per lane.  Kept in its own translation unit: the kernel is
// instantiated for every (l1, l2) pair of every supported 2J.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "pxff_common.h"
#include "so4_dev.h"
#include "so4_tables.h"

namespace pxff {

// Z-group chunk tables per (device, 2J, normalisation, nchunks)
static std::mutex g_chunk_mutex;
static std::map<std::tuple<int, int, int>, int *> g_chunks;
static int get_zchunks(int twol, int norm, int nch, const int **out) {
    int dev = 0;
    PXFF_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_chunk_mutex);
    auto key = std::make_tuple(dev, twol * 2 + (norm ? 1 : 0), nch);
    auto it = g_chunks.find(key);
    if (it != g_chunks.end()) {
        *out = it->second;
        return 0;
    }
    So4HostTables t;
    if (!build_so4_tables(twol, norm, t)) {
        set_error("unsupported 2J");
        return PXFF_ENOTIMPL;
    }
    std::vector<int> c = chunk_zgroups(t, nch);
    int *d = nullptr;
    PXFF_HIP_CHECK(hipMalloc((void **)&d, c.size() * sizeof(int)));
    PXFF_HIP_CHECK(hipMemcpy(d, c.data(), c.size() * sizeof(int), hipMemcpyHostToDevice));
    g_chunks[key] = d;
    *out = d;
    return 0;
}


// ------------------------------------------------- Z, atom per lane
// compute_zi (SNAP.py:1050-1098) with one atom per lane: a wavefront owns a
// tile of 64 atoms and a chunk of Z row groups, so the loop structure and
// the Clebsch-Gordan coefficients are the same for all lanes (uniform
// branches, coefficients as scalar operands).  Row (l1, l2, l, mb) is
//   Z[ma] = sum_{mb1+mb2=K} C[mb1][mb2] sum_{ma1+ma2=ma+s} C[ma1][ma2] U1[mb1][ma1] U2[mb2][ma2],
// s = (l1 + l2 - l)/2, K = mb + s.  The rows of one (l1, l2) with the same
// K (one per l) share every (mb1, mb2) pair and hence every product
// U1[mb1][ma1] U2[mb2][ma2]: per pair a group loads the U^{l2} row mb2
// into registers and streams the U^{l1} row mb1 one element at a time
// (coalesced loads from the tile-major U_tot, rows beyond l/2 by inversion
// symmetry), forms each product once and adds it, with that l's
// coefficients, into the register accumulators of every l whose ma range
// holds ma1 + ma2.  Accumulators of l without a row in the group get a zero
// coefficient (uniform predication instead of branches).  Z is stored
// tile-major, [tile][nz][64 lanes], so every store is one coalesced 1 KB
// row.  Workgroup ids are mapped so that the chunks of one tile run on the
// same XCD (workgroups are dispatched to the 8 XCDs round-robin), sharing
// its L2.
// Compile-time loop: f(std::integral_constant<int, 0>) ... f(<N-1>).
template <typename F, int... I>
__device__ inline void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ inline void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Sign flip by XOR on the high word (m is 0 or 0x80000000).
__device__ __forceinline__ double xsign(double x, uint32_t m) {
    return __hiloint2double(__double2hiint(x) ^ (int)m, __double2loint(x));
}

// Element access to row mb of U^LV in the tile-major half table: rows
// beyond LV/2 come from the inversion symmetry
//   U[LV-mb][LV-ma] = (-1)^(ma+mb) conj(U[mb][ma]),
// so a row is a uniform start, a uniform direction (+-1 column) and a sign
// pattern applied by XOR: element ma gets sign masks (sr, si) = (m0, m0^SB)
// for even ma and (m0^SB, m0) for odd ma, m0 = 0 for stored rows.
struct URow {
    const cd *p;  // element 0 of the row in lane 0 (wave-uniform)
    int step;     // +-64 (elements)
    uint32_t se, so;
    int lane;
};
template <int LV>
__device__ __forceinline__ URow urow(const cd *__restrict__ ut, int mb, int hoffl, int lane) {
    URow r;
    r.lane = lane;
    if (2 * mb > LV) {
        r.p = ut + (hoffl + (LV - mb) * (LV + 1) + LV) * 64;
        r.step = -64;
        r.se = (uint32_t)(mb & 1) << 31;
        r.so = r.se ^ 0x80000000u;
    } else {
        r.p = ut + (hoffl + mb * (LV + 1)) * 64;
        r.step = 64;
        r.se = 0;
        r.so = 0;
    }
    return r;
}
template <int MA>
__device__ __forceinline__ cd urow_get(const URow &r) {
    // uniform row element address, lane as the only per-lane offset: the
    // address arithmetic stays scalar
    const cd *e = r.p + MA * r.step;
    const cd v = e[r.lane];
    const uint32_t sr = (MA & 1) ? r.so : r.se;
    const uint32_t si = (MA & 1) ? r.se : r.so;
    // stored rows: se = so = 0 and the XORs are no-ops; mirrored rows:
    // re gets (-1)^(ma+mb), im gets -(-1)^(ma+mb)
    return cmake(xsign(v.re, sr), xsign(v.im, si));
}

// The l of (L1, L2): LMIN + 2 j, j < nl_of; accumulator offset of l_j.
__host__ __device__ constexpr int zg_nl(int TWOJ, int L1, int L2) {
    return ((TWOJ < L1 + L2 ? TWOJ : L1 + L2) - (L1 - L2)) / 2 + 1;
}
__host__ __device__ constexpr int zg_accoff(int L1, int L2, int j) {
    int o = 0;
    for (int q = 0; q < j; q++) o += L1 - L2 + 2 * q + 1;
    return o;
}
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

template <int TWOJ, int L1, int L2>
__device__ __forceinline__ void zgroup(const cd *__restrict__ ut, int lane, const double *__restrict__ cg,
                                       const int *__restrict__ g, cd *__restrict__ zo, double *__restrict__ bpo) {
    constexpr int NL = zg_nl(TWOJ, L1, L2);
    constexpr int NACC = zg_accoff(L1, L2, NL);
    constexpr int L2P = L2 + 1;
    const int K = __builtin_amdgcn_readfirstlane(g[2]);
    const int mb1lo = __builtin_amdgcn_readfirstlane(g[3]);
    const int nbb = __builtin_amdgcn_readfirstlane(g[4]);
    const int h1 = __builtin_amdgcn_readfirstlane(g[5]);
    const int h2 = __builtin_amdgcn_readfirstlane(g[6]);
    const int mask = __builtin_amdgcn_readfirstlane(g[7]);
    int cgblk[NL];
#pragma unroll
    for (int j = 0; j < NL; j++) cgblk[j] = __builtin_amdgcn_readfirstlane(g[ZG_CGBLK + j]);
    cd acc[NACC];
#pragma unroll
    for (int t = 0; t < NACC; t++) acc[t] = cmake(0.0, 0.0);
    for (int ib = 0; ib < nbb; ib++) {
        const int mb1 = mb1lo + ib, mb2 = K - mb1;
        const URow r1 = urow<L1>(ut, mb1, h1, lane);
        const URow r2 = urow<L2>(ut, mb2, h2, lane);
        cd u2[L2 + 1];
        static_for<L2 + 1>([&](auto C) { u2[decltype(C)::value] = urow_get<decltype(C)::value>(r2); });
        double cgb[NL];
#pragma unroll
        for (int j = 0; j < NL; j++) {
            cgb[j] = ((mask >> j) & 1) ? cg[cgblk[j] + mb1 * L2P + mb2] : 0.0;
            // in a VGPR: the product with a scalar coefficient per term then
            // takes one scalar operand (a VALU instruction reads at most
            // one), no copy of the coefficient into a VGPR per term
            asm volatile("" : "+v"(cgb[j]));
        }
        // keep the CG reads of the product loop inside the pair loop:
        // hoisted, they would not fit in registers beside acc and U
        asm volatile("" ::: "memory");
        static_for<L1 + 1>([&](auto A) {
            constexpr int a = decltype(A)::value;
            // one ma1 at a time: its coefficient loads are not hoisted over
            // the whole pair (they would not fit in the scalar registers)
            asm volatile("" ::: "memory");
            const cd u1a = urow_get<a>(r1);
            static_for<L2 + 1>([&](auto C) {
                constexpr int c = decltype(C)::value;
                constexpr int t = a + c;
                constexpr bool used = (t >= (L1 + L2 - (NL - 1) * 2 - (L1 - L2)) / 2) &&
                                      (t <= (L1 + L2 + (L1 - L2) + 2 * (NL - 1)) / 2);
                if constexpr (used) {
                    const cd pr = cmul(u1a, u2[c]);
                    static_for<NL>([&](auto J) {
                        constexpr int j = decltype(J)::value;
                        constexpr int l = L1 - L2 + 2 * j;
                        constexpr int s = (L1 + L2 - l) / 2;
                        if constexpr (t >= s && t <= s + l) {
                            constexpr int o = zg_accoff(L1, L2, j) + t - s;
                            const double w = cgb[j] * cg[cgblk[j] + a * L2P + c];
                            acc[o].re += w * pr.re;
                            acc[o].im += w * pr.im;
                        }
                    });
                }
            });
        });
    }
    // every lane stores (padding lanes hold zeros: their U is zero)
    static_for<NL>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr int l = L1 - L2 + 2 * j;
        if ((mask >> j) & 1) {
            const int zrow0 = __builtin_amdgcn_readfirstlane(g[ZG_ZROW + j]);
#pragma unroll
            for (int m = 0; m <= l; m++) zo[(zrow0 + m) * 64] = acc[zg_accoff(L1, L2, j) + m];
        }
    });
    // B row partials (compute_bi, SNAP.py:1101-1141): the rows of a B
    // component (l1, l2, l), l >= l1, contribute
    //   sum_ma w Re(conj(U^l[mb][ma]) Z[mb][ma]),
    // w = 1 on rows mb < l/2 and, on the middle row of even l, 1 for
    // ma < mb, 1/2 at ma = mb, 0 beyond.  Rows of one component lie in
    // several groups, so each row's sum is stored and summed afterwards.
    if (bpo != nullptr) {
        static_for<NL>([&](auto J) {
            constexpr int j = decltype(J)::value;
            constexpr int l = L1 - L2 + 2 * j;
            if constexpr (l >= L1) {
                const int brow = __builtin_amdgcn_readfirstlane(g[ZG_BROW + j]);
                if (brow >= 0) {
                    const int urow = __builtin_amdgcn_readfirstlane(g[ZG_UROW + j]);
                    const int mb = K - (L1 + L2 - l) / 2;
                    const bool mid = 2 * mb == l;
                    double sf = 0.0, sh = 0.0;
#pragma unroll
                    for (int m = 0; m <= l; m++) {
                        const cd u = ut[(urow + m) * 64 + lane];
                        const double d = cdotre(u, acc[zg_accoff(L1, L2, j) + m]);
                        if (!mid || m < mb)
                            sf += d;
                        else if (m == mb)
                            sh = d;
                    }
                    bpo[brow * 64] = sf + 0.5 * sh;
                }
            }
        });
    }
}

// Uniform dispatch to zgroup<TWOJ, l1, l2>.
template <int TWOJ, int L1, int L2>
__device__ __forceinline__ void zgroup_dispatch(int l1, int l2, const cd *ut, int lane, const double *cg,
                                                const int *g, cd *zo, double *bpo) {
    if constexpr (L1 <= TWOJ) {
        if (l1 == L1 && l2 == L2) {
            zgroup<TWOJ, L1, L2>(ut, lane, cg, g, zo, bpo);
            return;
        }
        if constexpr (L2 < L1)
            zgroup_dispatch<TWOJ, L1, L2 + 1>(l1, l2, ut, lane, cg, g, zo, bpo);
        else
            zgroup_dispatch<TWOJ, L1 + 1, 0>(l1, l2, ut, lane, cg, g, zo, bpo);
    }
}

template <int TWOJ>
__global__ __launch_bounds__(256) void so4_z_tile_kernel(int64_t natoms, int ntiles, int nch, const int *__restrict__ zchunk,
                                                          const int *__restrict__ zgroups, const double *__restrict__ cg_g,
                                                          int ncg, const double *__restrict__ utile,
                                                          double *__restrict__ zout, int nz,
                                                          double *__restrict__ bpart, int nbrow) {
    constexpr int NH = nhalf_of(TWOJ);
    // CG coefficients by wave-uniform addresses: scalar loads into SGPR
    // operands (an LDS copy would cost 512 B of LDS bandwidth per coefficient
    // read, all lanes receiving the same value)
    const double *__restrict__ cg = cg_g;
    (void)ncg;
    // workgroup = 4 wavefronts on 4 consecutive chunks of one tile
    const int b = blockIdx.x;
    const int nq = (nch + 3) / 4;
    const int xcd = b & 7, k = b >> 3;
    const int tile = xcd + 8 * (k / nq);
    const int ch = 4 * (k % nq) + (threadIdx.x >> 6);
    if (tile >= ntiles || ch >= nch) return;
    const int lane = threadIdx.x & 63;
    const cd *ut = reinterpret_cast<const cd *>(utile) + (size_t)tile * NH * 64;
    cd *zo = reinterpret_cast<cd *>(zout) + (size_t)tile * nz * 64 + lane;
    double *bpo = bpart ? bpart + (size_t)tile * nbrow * 64 + lane : nullptr;
    const int r0 = __builtin_amdgcn_readfirstlane(zchunk[ch]);
    const int r1 = __builtin_amdgcn_readfirstlane(zchunk[ch + 1]);
    for (int r = r0; r < r1; r++) {
        const int *g = zgroups + ZG_STRIDE * r;
        const int l1 = __builtin_amdgcn_readfirstlane(g[0]);
        const int l2 = __builtin_amdgcn_readfirstlane(g[1]);
        zgroup_dispatch<TWOJ, 0, 0>(l1, l2, ut, lane, cg, g, zo, bpo);
    }
}

// B[i][b] = 2 sum_{mb <= l_b/2} P[b_brow0[b] + mb] from the row partials P
// of the Z kernel ([tile][nbrow][64], one atom per lane): one thread per
// (atom, b), the four wavefronts of a workgroup on four components of one
// tile, so each thread's few loads are independent and coalesced.
__global__ __launch_bounds__(256) void so4_brow_sum_kernel(int64_t natoms, int nb, int nbrow,
                                                            const int *__restrict__ b_brow0,
                                                            const int *__restrict__ b_level,
                                                            const double *__restrict__ bpart, double *__restrict__ bout) {
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int64_t tile = blockIdx.x;
    const int64_t i = tile * 64 + lane;
    if (i >= natoms || b >= nb) return;
    const double *pt = bpart + (size_t)tile * nbrow * 64 + lane;
    const int r0 = b_brow0[b];
    const int nr = b_level[b] / 2 + 1;
    double s = 0.0;
    for (int r = 0; r < nr; r++) s += pt[(size_t)(r0 + r) * 64];
    bout[(size_t)i * nb + b] = 2.0 * s;
}

}  // namespace pxff

// =================================================================== C ABI
using namespace pxff;

extern "C" int pxff_so4_zb(const pxff_so4_params *p, int64_t natoms, const double *d_utot, double *d_z, double *d_b,
                           void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int ntiles = (int)((natoms + 63) / 64);
    const int nch = std::min(tune_value(PXFF_TUNE_Z_CHUNKS), T->nzgroups);
    const int *zch = nullptr;
    if ((rc = get_zchunks(p->twol, p->normalize_u, nch, &zch))) return rc;
    const int grid = 8 * ((ntiles + 7) / 8) * ((nch + 3) / 4);
    // B row partials after Z in the same buffer
    double *bpart = d_b ? d_z + 2 * (size_t)ntiles * 64 * T->nz : nullptr;
    PXFF_DISPATCH_TWOJ(p->twol, (so4_z_tile_kernel<TJ><<<grid, 256, 0, s>>>(natoms, ntiles, nch, zch, T->zgroups,
                                                                           T->cg, T->ncg, d_utot, d_z, T->nz, bpart,
                                                                           T->nbrow)));
    PXFF_LAUNCH_CHECK();
    if (d_b) {
        const dim3 gb((unsigned)ntiles, (unsigned)div_up(T->nb, 4));
        so4_brow_sum_kernel<<<gb, 256, 0, s>>>(natoms, T->nb, T->nbrow, T->b_brow0, T->b_level, bpart, d_b);
        PXFF_LAUNCH_CHECK();
    }
    return 0;
}

extern "C" int pxff_so4_zlen(int twol, int64_t natoms, int64_t *len) {
    PXFF_REQUIRE(len != nullptr, "null output");
    PXFF_REQUIRE(natoms >= 0, "natoms must be >= 0");
    So4HostTables t;
    if (twol < 0 || (twol & 1) || !build_so4_tables(twol, 1, t)) {
        set_error("unsupported 2J=" + std::to_string(twol) + " (even 2J in 0..10 supported)");
        return PXFF_ENOTIMPL;
    }
    *len = (natoms + 63) / 64 * 64 * (2 * (int64_t)t.nz + t.nbrow);
    return 0;
}

