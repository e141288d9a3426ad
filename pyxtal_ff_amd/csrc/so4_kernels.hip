// SO(4)/SNAP bispectrum kernels for gfx950 (FP64, wave64).
//
// Device layout (HBM), per atom i:
//   utot[i][nhalf]  complex  U_tot half rows (mb <= l/2), self term, u_norm
//   z[i][nz]        complex  Z in the reference's idxz order
//   b[i][nb]        real     B in the reference's idxb order
//   y[i][nhalf]     complex  adjoint Y (dE/dB-weighted Z, half weights,
//                            factor 2 and u_norm folded in)
// per image e (centre i, neighbour j, translation T): r_e[3], dE/dr_e[3].
//
// The Wigner-U recursion (compute_uarray_recursive / compute_duarray_recursive,
// SNAP.py:870-1017) is run "one lane per row": a group of G = 2J/2+1 lanes
// serves one neighbour image, lane mb carrying row mb of every level in
// registers.  Row mb of level l only needs row mb of level l-1, so the lanes
// never exchange data; lane mb starts from row 0 (which it recurses up to
// level mb) and turns it into row mb of level mb with the inversion symmetry
// u[l][l-mb][l-ma] = (-1)^(ma+mb) conj(u[l][mb][ma]) (SNAP.py:908-925).
// Every lane executes the same fully unrolled instruction stream
// (sum_l l recursion steps), so groups never diverge.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "pxff_common.h"
#include "so4_tables.h"

namespace pxff {

// ------------------------------------------------------------ device tables
struct So4Dev {
    int twol, nu, nhalf, nz, nb, ncg, nblk;
    double *cg;
    int *zsched;
    int *blk_level_start, *blk_zoff, *blk_tb;
    double *blk_tf;
    int *b_zoff, *b_level;
    double *rootpq, *lnorm;
    int *hoff, *h_level, *h_diag;
    double *h_weight;
    int *idxu_block;
    int nzrows, nzchunks;
    int *zrows, *zchunk;
};

namespace {
std::mutex g_tab_mutex;
std::map<std::pair<int, int>, So4Dev> g_tabs;  // (device, 2J*2+norm)

template <typename T>
int upload(const std::vector<T> &v, T **dst) {
    size_t n = v.size() ? v.size() : 1;
    PXFF_HIP_CHECK(hipMalloc((void **)dst, n * sizeof(T)));
    if (!v.empty()) PXFF_HIP_CHECK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}
}  // namespace

int get_so4_tables(int twol, int norm, const So4Dev **out) {
    int dev = 0;
    PXFF_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tab_mutex);
    auto key = std::make_pair(dev, twol * 2 + (norm ? 1 : 0));
    auto it = g_tabs.find(key);
    if (it != g_tabs.end()) {
        *out = &it->second;
        return 0;
    }
    So4HostTables t;
    if (!build_so4_tables(twol, norm, t)) {
        set_error("unsupported 2J=" + std::to_string(twol) + " (even 2J in 0..10 supported)");
        return PXFF_ENOTIMPL;
    }
    So4Dev d;
    d.twol = twol;
    d.nu = t.nu;
    d.nhalf = t.nhalf;
    d.nz = t.nz;
    d.nb = t.nb;
    d.ncg = t.ncg;
    d.nblk = t.nblk;
    int rc = 0;
    if ((rc = upload(t.cglist, &d.cg))) return rc;
    if ((rc = upload(t.zsched, &d.zsched))) return rc;
    if ((rc = upload(t.blk_level_start, &d.blk_level_start))) return rc;
    if ((rc = upload(t.blk_zoff, &d.blk_zoff))) return rc;
    if ((rc = upload(t.blk_tb, &d.blk_tb))) return rc;
    if ((rc = upload(t.blk_tf, &d.blk_tf))) return rc;
    if ((rc = upload(t.b_zoff, &d.b_zoff))) return rc;
    if ((rc = upload(t.b_level, &d.b_level))) return rc;
    if ((rc = upload(t.rootpq, &d.rootpq))) return rc;
    if ((rc = upload(t.lnorm, &d.lnorm))) return rc;
    if ((rc = upload(t.hoff, &d.hoff))) return rc;
    if ((rc = upload(t.h_level, &d.h_level))) return rc;
    if ((rc = upload(t.h_diag, &d.h_diag))) return rc;
    if ((rc = upload(t.h_weight, &d.h_weight))) return rc;
    if ((rc = upload(t.idxu_block, &d.idxu_block))) return rc;
    d.nzrows = t.nzrows;
    d.nzchunks = 16;
    if ((rc = upload(t.zrows, &d.zrows))) return rc;
    if ((rc = upload(chunk_zrows(t, d.nzchunks), &d.zchunk))) return rc;
    g_tabs[key] = d;
    *out = &g_tabs[key];
    return 0;
}

// ------------------------------------------------------------ tuning knobs
// Process-wide launch parameters set by the autotuner (pxff_tune_set).
static int g_tune[PXFF_TUNE_COUNT] = {16, 16, 4};
static const int g_tune_lo[PXFF_TUNE_COUNT] = {1, 1, 1};
static const int g_tune_hi[PXFF_TUNE_COUNT] = {256, 256, 4};

// Z-row chunk tables per (device, 2J, normalisation, nchunks)
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
    std::vector<int> c = chunk_zrows(t, nch);
    int *d = nullptr;
    PXFF_HIP_CHECK(hipMalloc((void **)&d, c.size() * sizeof(int)));
    PXFF_HIP_CHECK(hipMemcpy(d, c.data(), c.size() * sizeof(int), hipMemcpyHostToDevice));
    g_chunks[key] = d;
    *out = d;
    return 0;
}

// ------------------------------------------------------- compile-time sizes
__host__ __device__ constexpr int nhalf_of(int twoj) {
    int s = 0;
    for (int l = 0; l <= twoj; l++) s += (l / 2 + 1) * (l + 1);
    return s;
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

// Cayley-Klein parameters of one image (compute_uarray_recursive,
// SNAP.py:909-914) and their gradients (compute_duarray_recursive,
// SNAP.py:965-976).
struct CK {
    cd ca, cb;          // conj(a), conj(b)
    cd cda[3], cdb[3];  // conj(da/dr_k), conj(db/dr_k)
};

__device__ inline void cayley_klein(double x, double y, double z, double r, double psi, bool grad, CK &ck) {
    double sinpsi, cospsi;
    sincos(psi, &sinpsi, &cospsi);
    // a = cos psi - i sin psi z/r ; b = sin psi/r (y - i x)
    ck.ca = cmake(cospsi, sinpsi / r * z);
    ck.cb = cmake(sinpsi / r * y, sinpsi / r * x);
    if (grad) {
        double g[3] = {x / r, y / r, z / r};
        double rr = r * r;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            // da = -sinpsi*psi/r*g - i*z*cospsi/r^2*psi*g + i*z*sinpsi/r^2*g  (+ -i sinpsi/r on z)
            double dre = -sinpsi * psi / r * g[k];
            double dim = -z * cospsi / rr * psi * g[k] + z * sinpsi / rr * g[k];
            if (k == 2) dim += -sinpsi / r;
            ck.cda[k] = cmake(dre, -dim);
            // db = cospsi/r^2*psi*(y - i x)*g - sinpsi/r^2*(y - i x)*g  (+ sinpsi/r on y, -i sinpsi/r on x)
            double f = cospsi / rr * psi * g[k] - sinpsi / rr * g[k];
            double bre = f * y, bim = -f * x;
            if (k == 1) bre += sinpsi / r;
            if (k == 0) bim += -sinpsi / r;
            ck.cdb[k] = cmake(bre, -bim);
        }
    }
}

// One level of the row recursion, in place, descending ma.
//   new[m] = rootpq[l-m][q] ca old[m] - rootpq[m][q] cb old[m-1],  q = l - mb
// (SNAP.py:921-931) and, with GRAD, its derivative (SNAP.py:982-991).
template <int TWOJ, bool GRAD>
__device__ inline void row_step(int l, int q, const double *__restrict__ rpq, const CK &ck, cd *u, cd (*du)[TWOJ + 1]) {
#pragma unroll
    for (int m = TWOJ; m >= 0; --m) {
        if (m > l) continue;
        cd un = cmake(0.0, 0.0);
        cd dn[3] = {cmake(0, 0), cmake(0, 0), cmake(0, 0)};
        if (m < l) {
            double c1 = rpq[(l - m) * (TWOJ + 1) + q];
            cd c1a = cscale(ck.ca, c1);
            un = cmul(c1a, u[m]);
            if (GRAD) {
#pragma unroll
                for (int k = 0; k < 3; k++)
                    dn[k] = cscale(cadd(cmul(ck.cda[k], u[m]), cmul(ck.ca, du[k][m])), c1);
            }
        }
        if (m >= 1) {
            double c2 = rpq[m * (TWOJ + 1) + q];
            cd c2b = cscale(ck.cb, -c2);
            un = cadd(un, cmul(c2b, u[m - 1]));
            if (GRAD) {
#pragma unroll
                for (int k = 0; k < 3; k++)
                    dn[k] = cadd(dn[k], cscale(cadd(cmul(ck.cdb[k], u[m - 1]), cmul(ck.cb, du[k][m - 1])), -c2));
            }
        }
        u[m] = un;
        if (GRAD) {
#pragma unroll
            for (int k = 0; k < 3; k++) du[k][m] = dn[k];
        }
    }
}

// Row 0 -> row l of the same level: new[m] = (-1)^(l-m) conj(old[l-m]).
template <int TWOJ, bool GRAD>
__device__ inline void row_flip(int l, cd *u, cd (*du)[TWOJ + 1]) {
    cd tu[TWOJ + 1];
    cd td[3][TWOJ + 1];
#pragma unroll
    for (int m = 0; m <= TWOJ; m++) {
        if (m > l) continue;
        double s = ((l - m) & 1) ? -1.0 : 1.0;
        cd o = u[l - m];
        tu[m] = cmake(s * o.re, -s * o.im);
        if (GRAD) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                cd od = du[k][l - m];
                td[k][m] = cmake(s * od.re, -s * od.im);
            }
        }
    }
#pragma unroll
    for (int m = 0; m <= TWOJ; m++) {
        if (m > l) continue;
        u[m] = tu[m];
        if (GRAD) {
#pragma unroll
            for (int k = 0; k < 3; k++) du[k][m] = td[k][m];
        }
    }
}

// Sweep the levels of one image.  After every level l all lanes call
// f(l, u_row, du_row, half) with half = (2 mb <= l), i.e. whether the lane's
// row is one of the half rows the bispectrum uses.  Lanes whose row is not a
// half row still run the recursion, so control flow stays uniform.
template <int TWOJ, bool GRAD, typename F>
__device__ inline void sweep_rows(int mb, const double *__restrict__ rpq, const CK &ck, F &&f) {
    cd u[TWOJ + 1];
    cd du[3][TWOJ + 1];
#pragma unroll
    for (int m = 0; m <= TWOJ; m++) {
        u[m] = cmake(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 3; k++) du[k][m] = cmake(0.0, 0.0);
    }
    u[0] = cmake(1.0, 0.0);
    f(0, u, du, mb == 0);
#pragma unroll
    for (int l = 1; l <= TWOJ; l++) {
        const int mbe = (l <= mb) ? 0 : mb;
        row_step<TWOJ, GRAD>(l, l - mbe, rpq, ck, u, du);
        if (l == mb) row_flip<TWOJ, GRAD>(l, u, du);
        f(l, u, du, 2 * mb <= l);
    }
}

// ---------------------------------------------------------------- U_tot
// One wavefront per centre (64 threads); NG groups of G lanes each take one
// image per round; per-group partial sums live in LDS and are reduced in a
// fixed order at the end (deterministic).  With TANGENT the kernel instead
// accumulates dU_tot . dr_e (forward-mode derivative of U_tot).
template <int TWOJ, bool TANGENT, bool TILED>
__global__ __launch_bounds__(64) void so4_accum_kernel(int64_t natoms, const int64_t *__restrict__ img_start,
                                                        const int64_t *__restrict__ img_j,
                                                        const double *__restrict__ img_r,
                                                        const double *__restrict__ atom_w,
                                                        const double *__restrict__ site_w,
                                                        const double *__restrict__ dr,
                                                        const int64_t *__restrict__ work_order, const double *rootpq_g,
                                                        const double *lnorm_g, double rcut, double rfac0, int cutoff_id,
                                                        double fcut0, double *__restrict__ out) {
    constexpr int G = TWOJ / 2 + 1;
    constexpr int NG = 64 / G;
    constexpr int NH = nhalf_of(TWOJ);
    __shared__ cd part[NG][NH];
    __shared__ double rpq[(TWOJ + 1) * (TWOJ + 1)];
    __shared__ double lnorm[TWOJ + 1];
    const int lane = threadIdx.x;
    const int g = lane / G, mb = lane % G;
    for (int k = lane; k < (TWOJ + 1) * (TWOJ + 1); k += 64) rpq[k] = rootpq_g[k];
    for (int k = lane; k <= TWOJ; k += 64) lnorm[k] = lnorm_g[k];

    for (int64_t c = blockIdx.x; c < natoms; c += gridDim.x) {
        const int64_t i = work_order ? work_order[c] : c;
        __syncthreads();
        for (int k = lane; k < NG * NH; k += 64) (&part[0][0])[k] = cmake(0.0, 0.0);
        __syncthreads();
        const int64_t s = img_start[i], e = img_start[i + 1];
        for (int64_t base = s; base < e; base += NG) {
            const int64_t img = base + g;
            bool valid = (g < NG) && (img < e);
            double x = 0, y = 0, z = 0, r = 1.0, w = 0.0;
            if (valid) {
                x = img_r[3 * img];
                y = img_r[3 * img + 1];
                z = img_r[3 * img + 2];
                r = sqrt(x * x + y * y + z * z);
                w = atom_w[img_j[img]];
                if (r < 1e-8) valid = false;  // SNAP.py:1487 (padding rows)
            }
            if (!valid) {
                x = 0;
                y = 0;
                z = 1.0;
                r = 1.0;
            }
            double fc, dfc;
            cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
            CK ck;
            cayley_klein(x, y, z, r, rfac0 * M_PI * r / rcut, TANGENT, ck);
            if (!TANGENT) {
                const double sc = fc * w;
                cd *prow = part[g < NG ? g : 0];
                sweep_rows<TWOJ, false>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
                    if (!valid || !half) return;
                    const int h0 = hoff_of(l) + mb * (l + 1);
#pragma unroll
                    for (int m = 0; m <= TWOJ; m++) {
                        if (m > l) continue;
                        cd v = prow[h0 + m];
                        prow[h0 + m] = cmake(v.re + sc * u[m].re, v.im + sc * u[m].im);
                    }
                });
            } else {
                double d0 = 0, d1 = 0, d2 = 0;
                if (valid) {
                    d0 = dr[3 * img];
                    d1 = dr[3 * img + 1];
                    d2 = dr[3 * img + 2];
                }
                // D . dr = fc (dU . dr) + dfc (rhat . dr) U
                const double gu = dfc * (x * d0 + y * d1 + z * d2) / r;
                cd *prow = part[g < NG ? g : 0];
                sweep_rows<TWOJ, true>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
                    if (!valid || !half) return;
                    const int h0 = hoff_of(l) + mb * (l + 1);
#pragma unroll
                    for (int m = 0; m <= TWOJ; m++) {
                        if (m > l) continue;
                        double re = fc * (du[0][m].re * d0 + du[1][m].re * d1 + du[2][m].re * d2) + gu * u[m].re;
                        double im = fc * (du[0][m].im * d0 + du[1][m].im * d1 + du[2][m].im * d2) + gu * u[m].im;
                        cd v = prow[h0 + m];
                        prow[h0 + m] = cmake(v.re + w * re, v.im + w * im);
                    }
                });
            }
        }
        __syncthreads();
        for (int h = lane; h < NH; h += 64) {
            cd acc = cmake(0.0, 0.0);
            for (int q = 0; q < NG; q++) acc = cadd(acc, part[q][h]);
            // level / row / column of h
            int l = 0;
            while (l < TWOJ && h >= hoff_of(l + 1)) l++;
            const int loc = h - hoff_of(l);
            const int hm = loc / (l + 1), hma = loc % (l + 1);
            if (!TANGENT && hm == hma) acc.re += fcut0 * site_w[i];  // addself_uarraytot
            acc = cscale(acc, lnorm[l]);                               // ulisttot *= u_norm
            if (TILED)  // [atom / 64][h][atom % 64], the layout of the Z kernel's loads
                reinterpret_cast<cd *>(out)[((size_t)(i >> 6) * NH + h) * 64 + (i & 63)] = acc;
            else
                reinterpret_cast<cd *>(out)[(size_t)i * NH + h] = acc;
        }
    }
}

// ------------------------------------------------------------------- Z, B
// One workgroup (256 threads) per atom, persistent over atoms.  The CG table
// and the full U_tot (rebuilt from the half rows by inversion symmetry) live
// in LDS; one thread per Z element runs compute_zi's double loop
// (SNAP.py:1050-1098), elements scheduled in decreasing-cost order.  B then
// follows compute_bi (SNAP.py:1101-1141).
template <int TWOJ>
__global__ __launch_bounds__(256) void so4_zb_kernel(int64_t natoms, const double *__restrict__ utot,
                                                       const double *__restrict__ cg_g, int ncg,
                                                       const int *__restrict__ zsched, int nz,
                                                       const int *__restrict__ b_zoff, const int *__restrict__ b_level,
                                                       int nb, double *__restrict__ zout, double *__restrict__ bout) {
    constexpr int NH = nhalf_of(TWOJ);
    constexpr int NU = (TWOJ + 1) * (TWOJ + 2) * (2 * TWOJ + 3) / 6;
    extern __shared__ double smem[];
    double *cg = smem;
    cd *U = reinterpret_cast<cd *>(smem + ((ncg + 1) & ~1));
    const int tid = threadIdx.x;
    for (int k = tid; k < ncg; k += 256) cg[k] = cg_g[k];
    for (int64_t i = blockIdx.x; i < natoms; i += gridDim.x) {
        __syncthreads();
        const cd *uh = reinterpret_cast<const cd *>(utot) + (size_t)i * NH;
        for (int k = tid; k < NU; k += 256) {
            int l = 0, base = 0;
            while (l < TWOJ && k >= base + (l + 1) * (l + 1)) {
                base += (l + 1) * (l + 1);
                l++;
            }
            int loc = k - base;
            int mb = loc / (l + 1), ma = loc % (l + 1);
            cd v;
            if (2 * mb <= l) {
                v = uh[hoff_of(l) + mb * (l + 1) + ma];
            } else {
                cd o = uh[hoff_of(l) + (l - mb) * (l + 1) + (l - ma)];
                double s = ((ma + mb) & 1) ? -1.0 : 1.0;
                v = cmake(s * o.re, -s * o.im);
            }
            U[k] = v;
        }
        __syncthreads();
        cd *zo = reinterpret_cast<cd *>(zout) + (size_t)i * nz;
        for (int k = tid; k < nz; k += 256) {
            const int *zs = zsched + 6 * k;
            const int jjz = zs[0];
            int u1 = zs[1], u2 = zs[2];
            const int cga0 = zs[3];
            int cgb = zs[4];
            const int pk = zs[5];
            const int na = pk & 15, nbb = (pk >> 4) & 15, l1p = (pk >> 8) & 15, l2p = (pk >> 12) & 15;
            const int l2 = pk >> 16;
            cd zsum = cmake(0.0, 0.0);
            for (int ib = 0; ib < nbb; ib++) {
                cd sa = cmake(0.0, 0.0);
                int cga = cga0;
                for (int ia = 0; ia < na; ia++) {
                    cd p = cmul(U[u1 + ia], U[u2 - ia]);
                    double c = cg[cga];
                    sa.re += c * p.re;
                    sa.im += c * p.im;
                    cga += l2;
                }
                double cb = cg[cgb];
                zsum.re += cb * sa.re;
                zsum.im += cb * sa.im;
                u1 += l1p;
                u2 -= l2p;
                cgb += l2;
            }
            zo[jjz] = zsum;
        }
        if (bout) {
            __syncthreads();
            for (int b = tid; b < nb; b += 256) {
                const int l = b_level[b];
                const cd *zb = zo + b_zoff[b];
                const int ub = l * (l + 1) * (2 * l + 1) / 6;  // idxu_block[l]
                double s = 0.0;
                int llz = 0;
                for (int mb = 0; 2 * mb < l; mb++)
                    for (int ma = 0; ma <= l; ma++, llz++) s += cdotre(U[ub + llz], zb[llz]);
                if (l % 2 == 0) {
                    const int mb = l / 2;
                    for (int ma = 0; ma < mb; ma++, llz++) s += cdotre(U[ub + llz], zb[llz]);
                    s += 0.5 * cdotre(U[ub + llz], zb[llz]);
                }
                bout[(size_t)i * nb + b] = 2.0 * s;
            }
        }
    }
}

// ------------------------------------------------- Z, atom per lane
// compute_zi (SNAP.py:1050-1098) with one atom per lane: a wavefront owns a
// tile of 64 atoms and a chunk of Z rows (block (l1, l2, l), row mb), so the
// loop structure and the Clebsch-Gordan coefficients are the same for all
// lanes (scalar loads, uniform branches).  For a row, every (mb1, mb2) pair
// of the ib loop brings U^{l1} row mb1 and U^{l2} row mb2 into registers
// (coalesced loads from the tile-major U_tot, rows beyond l/2 by inversion
// symmetry), and all (ma1, ma2) products are accumulated by ma1 + ma2 =
// ma + (l1 + l2 - l)/2 into registers with static indices:
//   Z[ma][mb] = sum_{mb1+mb2=mb+s} C[mb1][mb2] sum_{ma1+ma2=ma+s} C[ma1][ma2] U1[mb1][ma1] U2[mb2][ma2].
// Workgroup ids are mapped so that the chunks of one tile run on the same XCD
// (workgroups are dispatched to the 8 XCDs round-robin), sharing its L2.
// Compile-time loop: f(std::integral_constant<int, 0>) ... f(<N-1>).
template <typename F, int... I>
__device__ inline void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ inline void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int TWOJ>
__device__ inline void load_urow(const cd *__restrict__ ut, int lv, int mb, int hoffl, int lane, cd *u) {
    const bool flip = 2 * mb > lv;
    const int mrow = flip ? lv - mb : mb;
    const int base = hoffl + mrow * (lv + 1);
#pragma unroll
    for (int ma = 0; ma <= TWOJ; ma++) {
        if (ma <= lv) {
            const int col = flip ? lv - ma : ma;
            cd v = ut[(base + col) * 64 + lane];
            if (flip) {
                const double sg = ((ma + mb) & 1) ? -1.0 : 1.0;
                v = cmake(sg * v.re, -sg * v.im);
            }
            u[ma] = v;
        } else {
            u[ma] = cmake(0.0, 0.0);
        }
    }
}

// One Z row (block (L1, L2, l), row mb) for the lane's atom, with the (l1,
// l2) of the block known at compile time: the loads of U rows are exact,
// and the pairs (a, c) of one t = a + c are branch-free (only the t range
// [sh, sh+l], which depends on l, is tested -- one uniform branch per t).
template <int TWOJ, int L1, int L2>
__device__ __forceinline__ void zrow(const cd *__restrict__ ut, int lane, const double *__restrict__ cg, int cgblk, int l,
                                  int sh, int mb, int mb1lo, int nbb, int h1, int h2, cd *__restrict__ zo, bool valid,
                                  int zrow0) {
    constexpr int NT = L1 + L2 + 1;
    constexpr int L2P = L2 + 1;
    cd acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = cmake(0.0, 0.0);
    const double *cgr = cg + cgblk;
    for (int ib = 0; ib < nbb; ib++) {
        const int mb1 = mb1lo + ib, mb2 = mb + sh - mb1;
        cd u1[L1 + 1], u2[L2 + 1];
        load_urow<L1>(ut, L1, mb1, h1, lane, u1);
        load_urow<L2>(ut, L2, mb2, h2, lane, u2);
        const double cgb = cgr[mb1 * L2P + mb2];
        // keep the CG reads of the pair loop inside the row loop: hoisted,
        // (L1+1)(L2+1) doubles would not fit in registers beside acc and U
        asm volatile("" ::: "memory");
        static_for<NT>([&](auto T) {
            constexpr int t = decltype(T)::value;
            if (t >= sh && t <= sh + l) {
                static_for<L1 + 1>([&](auto A) {
                    constexpr int a = decltype(A)::value;
                    constexpr int c = t - a;
                    if constexpr (c >= 0 && c <= L2) {
                        const double w = cgb * cgr[a * L2P + c];
                        const cd pr = cmul(u1[a], u2[c]);
                        acc[t].re += w * pr.re;
                        acc[t].im += w * pr.im;
                    }
                });
            }
        });
    }
    if (valid) {
#pragma unroll
        for (int t = 0; t < NT; t++)
            if (t >= sh && t <= sh + l) zo[zrow0 + t - sh] = acc[t];
    }
}

// Uniform dispatch to zrow<TWOJ, l1, l2>.
template <int TWOJ, int L1, int L2>
__device__ inline void zrow_dispatch(int l1, int l2, const cd *ut, int lane, const double *cg, int cgblk, int l, int sh,
                                     int mb, int mb1lo, int nbb, int h1, int h2, cd *zo, bool valid, int zrow0) {
    if constexpr (L1 <= TWOJ) {
        if (l1 == L1 && l2 == L2) {
            zrow<TWOJ, L1, L2>(ut, lane, cg, cgblk, l, sh, mb, mb1lo, nbb, h1, h2, zo, valid, zrow0);
            return;
        }
        if constexpr (L2 < L1)
            zrow_dispatch<TWOJ, L1, L2 + 1>(l1, l2, ut, lane, cg, cgblk, l, sh, mb, mb1lo, nbb, h1, h2, zo, valid, zrow0);
        else
            zrow_dispatch<TWOJ, L1 + 1, 0>(l1, l2, ut, lane, cg, cgblk, l, sh, mb, mb1lo, nbb, h1, h2, zo, valid, zrow0);
    }
}

template <int TWOJ>
__global__ __launch_bounds__(256) void so4_z_tile_kernel(int64_t natoms, int ntiles, int nch, const int *__restrict__ zchunk,
                                                          const int *__restrict__ zrows, const double *__restrict__ cg_g,
                                                          int ncg, const double *__restrict__ utile,
                                                          double *__restrict__ zout, int nz) {
    constexpr int NH = nhalf_of(TWOJ);
    extern __shared__ double cg[];  // the whole CG table; uniform-address reads broadcast
    for (int k = threadIdx.x; k < ncg; k += blockDim.x) cg[k] = cg_g[k];
    __syncthreads();
    // workgroup = 4 wavefronts on 4 consecutive chunks of one tile
    const int b = blockIdx.x;
    const int nq = (nch + 3) / 4;
    const int xcd = b & 7, k = b >> 3;
    const int tile = xcd + 8 * (k / nq);
    const int ch = 4 * (k % nq) + (threadIdx.x >> 6);
    if (tile >= ntiles || ch >= nch) return;
    const int lane = threadIdx.x & 63;
    const int64_t atom = (int64_t)tile * 64 + lane;
    const bool valid = atom < natoms;
    const cd *ut = reinterpret_cast<const cd *>(utile) + (size_t)tile * NH * 64;
    cd *zo = reinterpret_cast<cd *>(zout) + (size_t)atom * nz;
    const int r0 = __builtin_amdgcn_readfirstlane(zchunk[ch]);
    const int r1 = __builtin_amdgcn_readfirstlane(zchunk[ch + 1]);
    for (int r = r0; r < r1; r++) {
        const int *row = zrows + 12 * r;
        const int zrow0 = __builtin_amdgcn_readfirstlane(row[0]);
        const int l1 = __builtin_amdgcn_readfirstlane(row[1]);
        const int l2 = __builtin_amdgcn_readfirstlane(row[2]);
        const int l = __builtin_amdgcn_readfirstlane(row[3]);
        const int sh = __builtin_amdgcn_readfirstlane(row[4]);
        const int mb = __builtin_amdgcn_readfirstlane(row[5]);
        const int mb1lo = __builtin_amdgcn_readfirstlane(row[6]);
        const int nbb = __builtin_amdgcn_readfirstlane(row[7]);
        const int cgblk = __builtin_amdgcn_readfirstlane(row[8]);
        const int h1 = __builtin_amdgcn_readfirstlane(row[9]);
        const int h2 = __builtin_amdgcn_readfirstlane(row[10]);
        zrow_dispatch<TWOJ, 0, 0>(l1, l2, ut, lane, cg, cgblk, l, sh, mb, mb1lo, nbb, h1, h2, zo, valid, zrow0);
    }
}

// B (compute_bi, SNAP.py:1101-1141) from the tile-major U_tot and Z: one
// thread per (atom, b).
template <int TWOJ>
__global__ void so4_b_kernel(int64_t natoms, const double *__restrict__ utile, const double *__restrict__ z, int nz,
                             int nb, const int *__restrict__ b_zoff, const int *__restrict__ b_level,
                             double *__restrict__ bout) {
    constexpr int NH = nhalf_of(TWOJ);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= natoms * nb) return;
    const int64_t i = t / nb;
    const int b = (int)(t - i * nb);
    const cd *ut = reinterpret_cast<const cd *>(utile) + (size_t)(i >> 6) * NH * 64 + (i & 63);
    const cd *zb = reinterpret_cast<const cd *>(z) + (size_t)i * nz + b_zoff[b];
    const int l = b_level[b];
    const int h0 = hoff_of(l);
    double s = 0.0;
    int llz = 0;
    for (int mb = 0; 2 * mb < l; mb++)
        for (int ma = 0; ma <= l; ma++, llz++) s += cdotre(ut[(h0 + llz) * 64], zb[llz]);
    if (l % 2 == 0) {
        const int mb = l / 2;
        for (int ma = 0; ma < mb; ma++, llz++) s += cdotre(ut[(h0 + llz) * 64], zb[llz]);
        s += 0.5 * cdotre(ut[(h0 + llz) * 64], zb[llz]);
    }
    bout[t] = 2.0 * s;
}

// ---------------------------------------------------------------------- Y
// Y[h] = 2 hw(h) norm_l sum_{blocks at level l} c_blk Z[blk][loc(h)],
// c_blk = sum_targets f beta[b].  One thread per (atom, half element).
__global__ void so4_yi_kernel(int64_t natoms, int nhalf, int nz, int nb, const double *__restrict__ z,
                              const double *__restrict__ beta, const int *__restrict__ hoff,
                              const int *__restrict__ h_level, const double *__restrict__ h_weight,
                              const double *__restrict__ lnorm, const int *__restrict__ blk_level_start,
                              const int *__restrict__ blk_zoff, const int *__restrict__ blk_tb,
                              const double *__restrict__ blk_tf, double *__restrict__ y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= natoms * nhalf) return;
    const int64_t i = t / nhalf;
    const int h = (int)(t - i * nhalf);
    const int l = h_level[h];
    const int loc = h - hoff[l];
    const cd *zi = reinterpret_cast<const cd *>(z) + (size_t)i * nz;
    const double *bi = beta + (size_t)i * nb;
    cd acc = cmake(0.0, 0.0);
    const double hw = h_weight[h];
    if (hw != 0.0) {
        for (int k = blk_level_start[l]; k < blk_level_start[l + 1]; k++) {
            double c = 0.0;
#pragma unroll
            for (int s = 0; s < 3; s++) {
                int b = blk_tb[3 * k + s];
                if (b >= 0) c += blk_tf[3 * k + s] * bi[b];
            }
            cd v = zi[blk_zoff[k] + loc];
            acc.re += c * v.re;
            acc.im += c * v.im;
        }
    }
    const double f = 2.0 * hw * lnorm[l];
    reinterpret_cast<cd *>(y)[(size_t)i * nhalf + h] = cmake(f * acc.re, f * acc.im);
}

// -------------------------------------------------------- adjoint dE/dr
// Flat over images: each wavefront serves NG images (any centres).  Per
// image: dE/dr_k = w Re sum_h conj(fc dU_k[h] + dfc rhat_k U[h]) Y_i[h].
template <int TWOJ>
__global__ __launch_bounds__(256) void so4_deidrj_kernel(int64_t nimg, const int64_t *__restrict__ img_atom,
                                                           const int64_t *__restrict__ img_j,
                                                           const double *__restrict__ img_r,
                                                           const double *__restrict__ atom_w,
                                                           const double *__restrict__ y, const double *rootpq_g,
                                                           double rcut, double rfac0, int cutoff_id,
                                                           double *__restrict__ dedr) {
    constexpr int G = TWOJ / 2 + 1;
    constexpr int NG = 64 / G;
    constexpr int NH = nhalf_of(TWOJ);
    __shared__ double rpq[(TWOJ + 1) * (TWOJ + 1)];
    __shared__ double red[4][64][4];
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63;
    const int wpb = blockDim.x >> 6;
    const int g = lane / G, mb = lane % G;
    for (int k = tid; k < (TWOJ + 1) * (TWOJ + 1); k += blockDim.x) rpq[k] = rootpq_g[k];
    __syncthreads();
    const int64_t img = ((int64_t)blockIdx.x * wpb + wv) * NG + g;
    bool valid = (g < NG) && (img < nimg);
    double x = 0, yy = 0, zz = 1.0, r = 1.0, w = 0.0;
    int64_t i = 0;
    if (valid) {
        x = img_r[3 * img];
        yy = img_r[3 * img + 1];
        zz = img_r[3 * img + 2];
        r = sqrt(x * x + yy * yy + zz * zz);
        w = atom_w[img_j[img]];
        i = img_atom[img];
        if (r < 1e-8) {
            valid = false;
            x = 0;
            yy = 0;
            zz = 1.0;
            r = 1.0;
        }
    }
    double fc, dfc;
    cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
    CK ck;
    cayley_klein(x, yy, zz, r, rfac0 * M_PI * r / rcut, true, ck);
    const cd *yi = reinterpret_cast<const cd *>(y) + (size_t)i * NH;
    double au = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
    sweep_rows<TWOJ, true>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
        if (!valid || !half) return;
        const int h0 = hoff_of(l) + mb * (l + 1);
#pragma unroll
        for (int m = 0; m <= TWOJ; m++) {
            if (m > l) continue;
            const cd yv = yi[h0 + m];
            au += cdotre(u[m], yv);
            a0 += cdotre(du[0][m], yv);
            a1 += cdotre(du[1][m], yv);
            a2 += cdotre(du[2][m], yv);
        }
    });
    red[wv][lane][0] = au;
    red[wv][lane][1] = a0;
    red[wv][lane][2] = a1;
    red[wv][lane][3] = a2;
    __syncthreads();
    if (valid && mb == 0) {
        double s[4] = {0, 0, 0, 0};
        for (int q = 0; q < G; q++)
            for (int k = 0; k < 4; k++) s[k] += red[wv][lane + q][k];
        const double rh[3] = {x / r, yy / r, zz / r};
        for (int k = 0; k < 3; k++) dedr[3 * img + k] = w * (fc * s[1 + k] + dfc * rh[k] * s[0]);
    }
}

// --------------------------------------------------- plain dB/dr per image
// One wavefront per NG images.  Level by level the groups put their D rows
// (D_k = fc dU_k + dfc rhat_k U, i.e. dU_tot/dr without w and u_norm) into
// LDS; then all 64 lanes form the block inner products
// p[blk][k] = Re <D_k^l, 2 hw norm_l Z^blk> of that level and add w f p to
// every (b, factor) target of the block -- the three terms of
// compute_dbidrj (SNAP.py:1143-1257).  Within one level a (b, k) output
// receives at most one block, so the accumulation needs no atomics.
template <int TWOJ>
__global__ __launch_bounds__(64) void so4_dbdr_kernel(
    int64_t nimg, const int64_t *__restrict__ img_atom, const int64_t *__restrict__ img_j,
    const double *__restrict__ img_r, const double *__restrict__ atom_w, const double *__restrict__ z, int nz, int nb,
    const double *rootpq_g, const double *__restrict__ lnorm, const int *__restrict__ blk_level_start,
    const int *__restrict__ blk_zoff, const int *__restrict__ blk_tb, const double *__restrict__ blk_tf, double rcut,
    double rfac0, int cutoff_id, double *__restrict__ dbdr) {
    constexpr int G = TWOJ / 2 + 1;
    constexpr int NG = 64 / G;
    constexpr int NL = (TWOJ / 2 + 1) * (TWOJ + 1);
    constexpr int NBMAX = nb_of(TWOJ);
    __shared__ double rpq[(TWOJ + 1) * (TWOJ + 1)];
    __shared__ cd Dl[NG][3][NL];
    __shared__ double acc[NG][NBMAX * 3];
    __shared__ int64_t s_atom[NG];
    __shared__ double s_w[NG];
    const int lane = threadIdx.x;
    const int g = lane / G, mb = lane % G;
    for (int k = lane; k < (TWOJ + 1) * (TWOJ + 1); k += 64) rpq[k] = rootpq_g[k];
    for (int k = lane; k < NG * NBMAX * 3; k += 64) (&acc[0][0])[k] = 0.0;
    const int64_t img = (int64_t)blockIdx.x * NG + g;
    bool valid = (g < NG) && (img < nimg);
    double x = 0, yy = 0, zz = 1.0, r = 1.0, w = 0.0;
    if (valid) {
        x = img_r[3 * img];
        yy = img_r[3 * img + 1];
        zz = img_r[3 * img + 2];
        r = sqrt(x * x + yy * yy + zz * zz);
        w = atom_w[img_j[img]];
        if (r < 1e-8) {
            valid = false;
            x = 0;
            yy = 0;
            zz = 1.0;
            r = 1.0;
        }
    }
    if (g < NG && mb == 0) {
        s_atom[g] = valid ? img_atom[img] : -1;
        s_w[g] = w;
    }
    double fc, dfc;
    cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
    const double rh[3] = {x / r, yy / r, zz / r};
    CK ck;
    cayley_klein(x, yy, zz, r, rfac0 * M_PI * r / rcut, true, ck);
    __syncthreads();
    sweep_rows<TWOJ, true>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
        if (g < NG && half) {
#pragma unroll
            for (int m = 0; m <= TWOJ; m++) {
                if (m > l) continue;
#pragma unroll
                for (int k = 0; k < 3; k++)
                    Dl[g][k][mb * (l + 1) + m] = cmake(fc * du[k][m].re + dfc * rh[k] * u[m].re,
                                                       fc * du[k][m].im + dfc * rh[k] * u[m].im);
            }
        }
        __syncthreads();
        const int b0 = blk_level_start[l], nbl = blk_level_start[l + 1] - b0;
        const int nloc = (l / 2 + 1) * (l + 1);
        const double nrm2 = 2.0 * lnorm[l];
        for (int it = lane; it < NG * nbl * 3; it += 64) {
            const int gg = it / (nbl * 3);
            const int rem = it - gg * nbl * 3;
            const int bk = rem / 3, k = rem - bk * 3;
            const int64_t a = s_atom[gg];
            if (a < 0) continue;
            const int blk = b0 + bk;
            const cd *zb = reinterpret_cast<const cd *>(z) + (size_t)a * nz + blk_zoff[blk];
            double p = 0.0;
            int loc = 0;
            for (int hm = 0; 2 * hm < l; hm++)
                for (int ma = 0; ma <= l; ma++, loc++) p += cdotre(Dl[gg][k][loc], zb[loc]);
            if ((l & 1) == 0) {
                const int hm = l / 2;
                for (int ma = 0; ma < hm; ma++, loc++) p += cdotre(Dl[gg][k][loc], zb[loc]);
                p += 0.5 * cdotre(Dl[gg][k][loc], zb[loc]);
            }
            (void)nloc;
            p *= nrm2 * s_w[gg];
#pragma unroll
            for (int s = 0; s < 3; s++) {
                const int b = blk_tb[3 * blk + s];
                if (b >= 0) acc[gg][3 * b + k] += blk_tf[3 * blk + s] * p;
            }
        }
        __syncthreads();
    });
    for (int gg = 0; gg < NG; gg++) {
        const int64_t im = (int64_t)blockIdx.x * NG + gg;
        if (im >= nimg) break;
        for (int k = lane; k < nb * 3; k += 64) dbdr[(size_t)im * nb * 3 + k] = acc[gg][k];
    }
}

// -------------------------------------------------------------- row scatter
// Reference layout rows (SNAP.py:1458-1468).  One workgroup per centre;
// threads over the (b, k) / (b, a, c) components, looping over the centre's
// rows in order -- no atomics, rows are owned by their centre.
__global__ __launch_bounds__(256) void so4_rows_kernel(int64_t natoms, int nb, const int64_t *__restrict__ row_start,
                                                        const int64_t *__restrict__ row_img,
                                                        const int64_t *__restrict__ seq,
                                                        const int64_t *__restrict__ nsite,
                                                        const double *__restrict__ pos,
                                                        const double *__restrict__ img_r,
                                                        const double *__restrict__ dbdr,
                                                        const double *__restrict__ inv_vol, double *__restrict__ dxdr,
                                                        double *__restrict__ rdxdr) {
    const int64_t i = blockIdx.x;
    if (i >= natoms) return;
    const int64_t r0 = row_start[i], r1 = row_start[i + 1], ns = nsite[i];
    const int nbk = nb * 3;
    for (int t = threadIdx.x; t < nbk; t += blockDim.x) {
        double self = 0.0;
        for (int64_t rw = r0; rw < r1; rw++) {
            if (rw == ns) continue;
            double v = 0.0;
            for (int64_t e = row_img[rw]; e < row_img[rw + 1]; e++) v += dbdr[(size_t)e * nbk + t];
            dxdr[(size_t)rw * nbk + t] = v;
            self -= v;
        }
        dxdr[(size_t)ns * nbk + t] = self;
    }
    if (!rdxdr) return;
    const double iv = inv_vol[i];
    const double Ri[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const int nb9 = nb * 9;
    for (int t = threadIdx.x; t < nb9; t += blockDim.x) {
        const int b = t / 9, ac = t - b * 9, a = ac / 3, c = ac - a * 3;
        double site = 0.0;
        for (int64_t rw = r0; rw < r1; rw++) {
            double v = 0.0;
            for (int64_t e = row_img[rw]; e < row_img[rw + 1]; e++) {
                const double db = dbdr[(size_t)e * nbk + 3 * b + c];
                const double Rj = img_r[3 * e + a] + Ri[a];
                site += Ri[a] * db;
                v -= Rj * db;
            }
            if (rw == ns) {
                site += v;
            } else {
                rdxdr[(size_t)rw * nb9 + t] = -v * iv;
            }
        }
        rdxdr[(size_t)ns * nb9 + t] = -site * iv;
    }
}

// ----------------------------------------------------- force contractions
// Plain path: dE/dR_m = sum over rows (i, m) of sum_b beta[i,b] dxdr[row,b,:]
// (neuralnetwork.py:484-497), gathered over centre m's rows through the
// reverse-row index.  One wavefront per atom.
__global__ __launch_bounds__(64) void dxdr_contract_kernel(int64_t natoms, int nb, const int64_t *__restrict__ row_start,
                                                            const int64_t *__restrict__ seq,
                                                            const int64_t *__restrict__ row_rev,
                                                            const double *__restrict__ beta,
                                                            const double *__restrict__ dxdr, double *__restrict__ grad) {
    const int64_t m = blockIdx.x;
    if (m >= natoms) return;
    const int lane = threadIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int64_t rw = row_start[m]; rw < row_start[m + 1]; rw++) {
        const int64_t i = seq[2 * rw + 1];
        const int64_t rr = row_rev[rw];
        const double *d = dxdr + (size_t)rr * nb * 3;
        for (int b = lane; b < nb; b += 64) {
            const double bt = beta[(size_t)i * nb + b];
            a0 += bt * d[3 * b];
            a1 += bt * d[3 * b + 1];
            a2 += bt * d[3 * b + 2];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off);
        a1 += __shfl_down(a1, off);
        a2 += __shfl_down(a2, off);
    }
    if (lane == 0) {
        grad[3 * m] = a0;
        grad[3 * m + 1] = a1;
        grad[3 * m + 2] = a2;
    }
}

// Adjoint path: dE/dR_i = sum_{e of i} (dE/dr_rev(e) - dE/dr_e); virial of
// centre i = sum_e r_e (x) dE/dr_e.  One thread per atom.
__global__ void so4_gather_kernel(int64_t natoms, const int64_t *__restrict__ img_start,
                                  const int64_t *__restrict__ img_rev, const double *__restrict__ img_r,
                                  const double *__restrict__ dedr, double *__restrict__ grad,
                                  double *__restrict__ virial) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms) return;
    double g[3] = {0, 0, 0};
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t e = img_start[i]; e < img_start[i + 1]; e++) {
        const int64_t q = img_rev[e];
#pragma unroll
        for (int k = 0; k < 3; k++) g[k] += dedr[3 * q + k] - dedr[3 * e + k];
        if (virial) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int c = 0; c < 3; c++) v[3 * a + c] += img_r[3 * e + a] * dedr[3 * e + c];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) grad[3 * i + k] = g[k];
    if (virial)
#pragma unroll
        for (int k = 0; k < 9; k++) virial[9 * i + k] = v[k];
}

// Per-atom rdxdr reduction (utilities/__init__.py:226-235).
__global__ void rdxdr_reduce_kernel(int64_t natoms, int nb, const int64_t *__restrict__ row_start,
                                    const double *__restrict__ rdxdr, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= natoms * nb) return;
    const int64_t i = t / nb;
    const int b = (int)(t - i * nb);
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t rw = row_start[i]; rw < row_start[i + 1]; rw++)
        for (int k = 0; k < 9; k++) s[k] += rdxdr[((size_t)rw * nb + b) * 9 + k];
    const int vo[6] = {0, 4, 8, 1, 2, 5};
    for (int k = 0; k < 6; k++) out[((size_t)i * nb + b) * 6 + k] = s[vo[k]];
}

// ----------------------------------------------------------- tangent of B
// dB_b = Re sum_targets f <dU^level, 2 hw Z^blk> with dU = dU_tot (u_norm
// already applied by the accumulation kernel).  One thread per (atom, b).
__global__ void so4_tangent_b_kernel(int64_t natoms, int nhalf, int nz, int nb, const double *__restrict__ dutot,
                                     const double *__restrict__ z, const int *__restrict__ hoff,
                                     const int *__restrict__ blk_level_start, const int *__restrict__ blk_zoff,
                                     const int *__restrict__ blk_tb, const double *__restrict__ blk_tf, int ldim,
                                     double *__restrict__ db) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= natoms * nb) return;
    const int64_t i = t / nb;
    const int b = (int)(t - i * nb);
    const cd *du = reinterpret_cast<const cd *>(dutot) + (size_t)i * nhalf;
    const cd *zi = reinterpret_cast<const cd *>(z) + (size_t)i * nz;
    double acc = 0.0;
    for (int l = 0; l < ldim; l++) {
        for (int k = blk_level_start[l]; k < blk_level_start[l + 1]; k++) {
            double f = 0.0;
            for (int s = 0; s < 3; s++)
                if (blk_tb[3 * k + s] == b) f += blk_tf[3 * k + s];
            if (f == 0.0) continue;
            const cd *zb = zi + blk_zoff[k];
            const cd *ub = du + hoff[l];
            double p = 0.0;
            int loc = 0;
            for (int hm = 0; 2 * hm < l; hm++)
                for (int ma = 0; ma <= l; ma++, loc++) p += cdotre(ub[loc], zb[loc]);
            if ((l & 1) == 0) {
                const int hm = l / 2;
                for (int ma = 0; ma < hm; ma++, loc++) p += cdotre(ub[loc], zb[loc]);
                p += 0.5 * cdotre(ub[loc], zb[loc]);
            }
            acc += 2.0 * f * p;
        }
    }
    db[t] = acc;
}

}  // namespace pxff

// =================================================================== C ABI
using namespace pxff;

#define PXFF_DISPATCH_TWOJ(twol, ...)                                             \
    switch (twol) {                                                               \
    case 0: { constexpr int TJ = 0; __VA_ARGS__; } break;                         \
    case 2: { constexpr int TJ = 2; __VA_ARGS__; } break;                         \
    case 4: { constexpr int TJ = 4; __VA_ARGS__; } break;                         \
    case 6: { constexpr int TJ = 6; __VA_ARGS__; } break;                         \
    case 8: { constexpr int TJ = 8; __VA_ARGS__; } break;                         \
    case 10: { constexpr int TJ = 10; __VA_ARGS__; } break;                       \
    default: set_error("unsupported 2J"); return PXFF_ENOTIMPL;                   \
    }

static int check_params(const pxff_so4_params *p) {
    PXFF_REQUIRE(p != nullptr, "null params");
    if (p->twol < 0 || p->twol > 10 || (p->twol & 1)) {
        set_error("unsupported 2J=" + std::to_string(p->twol) + " (even 2J in 0..10 supported)");
        return PXFF_ENOTIMPL;
    }
    PXFF_REQUIRE(p->rcut > 0, "rcut must be > 0");
    PXFF_REQUIRE(p->cutoff_id >= 1 && p->cutoff_id <= 7, "cutoff id must be 1..7");
    return 0;
}

static int g_num_cus = 0;
static int num_cus() {
    if (g_num_cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
        g_num_cus = prop.multiProcessorCount;
    }
    return g_num_cus;
}

extern "C" int pxff_so4_sizes(int twol, int *nu, int *nhalf, int *nz, int *nb) {
    So4HostTables t;
    if (!build_so4_tables(twol, 0, t)) {
        set_error("unsupported 2J=" + std::to_string(twol));
        return PXFF_ENOTIMPL;
    }
    if (nu) *nu = t.nu;
    if (nhalf) *nhalf = t.nhalf;
    if (nz) *nz = t.nz;
    if (nb) *nb = t.nb;
    return 0;
}

extern "C" int pxff_so4_ui(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                           const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                           const double *d_site_w, const int64_t *d_work_order, double *d_utot, void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int grid = (int)std::min<int64_t>(natoms, (int64_t)num_cus() * g_tune[PXFF_TUNE_ACCUM_WAVES_PER_CU]);
    const double fcut0 = cutoff_f0_host(p->cutoff_id);
    PXFF_DISPATCH_TWOJ(p->twol, (so4_accum_kernel<TJ, false, true><<<grid, 64, 0, s>>>(
                                    natoms, d_img_start, d_img_j, d_img_r, d_atom_w, d_site_w, nullptr, d_work_order,
                                    T->rootpq, T->lnorm, p->rcut, p->rfac0, p->cutoff_id, fcut0, d_utot)));
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_zb(const pxff_so4_params *p, int64_t natoms, const double *d_utot, double *d_z, double *d_b,
                           void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int ntiles = (int)((natoms + 63) / 64);
    const int nch = std::min(g_tune[PXFF_TUNE_Z_CHUNKS], T->nzrows);
    const int *zch = nullptr;
    if ((rc = get_zchunks(p->twol, p->normalize_u, nch, &zch))) return rc;
    const int grid = 8 * ((ntiles + 7) / 8) * ((nch + 3) / 4);
    const size_t lds = sizeof(double) * (size_t)T->ncg;
    PXFF_DISPATCH_TWOJ(p->twol, {
        auto kern = so4_z_tile_kernel<TJ>;
        if (lds > 65536)
            PXFF_HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        kern<<<grid, 256, lds, s>>>(natoms, ntiles, nch, zch, T->zrows, T->cg, T->ncg, d_utot, d_z, T->nz);
    });
    PXFF_LAUNCH_CHECK();
    if (d_b) return pxff_so4_b(p, natoms, d_utot, d_z, d_b, stream);
    return 0;
}

extern "C" int pxff_so4_b(const pxff_so4_params *p, int64_t natoms, const double *d_utot, const double *d_z,
                          double *d_b, void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    PXFF_REQUIRE(d_b != nullptr, "null B output");
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    PXFF_DISPATCH_TWOJ(p->twol, (so4_b_kernel<TJ><<<div_up(natoms * T->nb, 256), 256, 0, s>>>(
                                    natoms, d_utot, d_z, T->nz, T->nb, T->b_zoff, T->b_level, d_b)));
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_yi(const pxff_so4_params *p, int64_t natoms, const double *d_z, const double *d_beta,
                           double *d_y, void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = natoms * T->nhalf;
    so4_yi_kernel<<<div_up(n, 256), 256, 0, s>>>(natoms, T->nhalf, T->nz, T->nb, d_z, d_beta, T->hoff, T->h_level,
                                                 T->h_weight, T->lnorm, T->blk_level_start, T->blk_zoff, T->blk_tb,
                                                 T->blk_tf, d_y);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_deidrj(const pxff_so4_params *p, int64_t nimg, const int64_t *d_img_atom,
                               const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                               const double *d_y, double *d_dedr, void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if (nimg <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int ng = 64 / (p->twol / 2 + 1);
    const int wpb = g_tune[PXFF_TUNE_IMAGE_WAVES_PER_BLOCK];
    const int grid = div_up(nimg, wpb * ng);
    PXFF_DISPATCH_TWOJ(p->twol, (so4_deidrj_kernel<TJ><<<grid, 64 * wpb, 0, s>>>(nimg, d_img_atom, d_img_j, d_img_r, d_atom_w,
                                                                           d_y, T->rootpq, p->rcut, p->rfac0,
                                                                           p->cutoff_id, d_dedr)));
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_gather(int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_rev,
                               const double *d_img_r, const double *d_dedr, double *d_grad, double *d_virial,
                               void *stream) {
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    so4_gather_kernel<<<div_up(natoms, 256), 256, 0, s>>>(natoms, d_img_start, d_img_rev, d_img_r, d_dedr, d_grad,
                                                          d_virial);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_tangent(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                                const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                                const double *d_dr, const double *d_z, double *d_dutot, double *d_db, void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int grid = (int)std::min<int64_t>(natoms, (int64_t)num_cus() * g_tune[PXFF_TUNE_ACCUM_WAVES_PER_CU]);
    PXFF_DISPATCH_TWOJ(p->twol, (so4_accum_kernel<TJ, true, false><<<grid, 64, 0, s>>>(
                                    natoms, d_img_start, d_img_j, d_img_r, d_atom_w, nullptr, d_dr, nullptr,
                                    T->rootpq, T->lnorm, p->rcut, p->rfac0, p->cutoff_id, 0.0, d_dutot)));
    PXFF_LAUNCH_CHECK();
    so4_tangent_b_kernel<<<div_up(natoms * T->nb, 256), 256, 0, s>>>(natoms, T->nhalf, T->nz, T->nb, d_dutot, d_z,
                                                                     T->hoff, T->blk_level_start, T->blk_zoff,
                                                                     T->blk_tb, T->blk_tf, p->twol + 1, d_db);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_dbdr(const pxff_so4_params *p, int64_t nimg, const int64_t *d_img_atom,
                             const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w, const double *d_z,
                             double *d_dbdr, void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if (nimg <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int ng = 64 / (p->twol / 2 + 1);
    const int grid = div_up(nimg, ng);
    PXFF_DISPATCH_TWOJ(p->twol, (so4_dbdr_kernel<TJ><<<grid, 64, 0, s>>>(
                                    nimg, d_img_atom, d_img_j, d_img_r, d_atom_w, d_z, T->nz, T->nb, T->rootpq,
                                    T->lnorm, T->blk_level_start, T->blk_zoff, T->blk_tb, T->blk_tf, p->rcut,
                                    p->rfac0, p->cutoff_id, d_dbdr)));
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_rows(int64_t natoms, int nb, const int64_t *d_row_start, const int64_t *d_row_img,
                             const int64_t *d_seq, const int64_t *d_nsite, const double *d_pos, const double *d_img_r,
                             const double *d_dbdr, const double *d_inv_vol, double *d_dxdr, double *d_rdxdr,
                             void *stream) {
    if (natoms <= 0) return 0;
    PXFF_REQUIRE(nb > 0, "nb must be > 0");
    hipStream_t s = (hipStream_t)stream;
    so4_rows_kernel<<<(int)natoms, 256, 0, s>>>(natoms, nb, d_row_start, d_row_img, d_seq, d_nsite, d_pos, d_img_r,
                                                 d_dbdr, d_inv_vol, d_dxdr, d_rdxdr);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_dxdr_contract(int64_t natoms, int nb, const int64_t *d_row_start, const int64_t *d_seq,
                                  const int64_t *d_row_rev, const double *d_beta, const double *d_dxdr, double *d_grad,
                                  void *stream) {
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    dxdr_contract_kernel<<<(int)natoms, 64, 0, s>>>(natoms, nb, d_row_start, d_seq, d_row_rev, d_beta, d_dxdr, d_grad);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_rdxdr_reduce(int64_t natoms, int nb, const int64_t *d_row_start, const double *d_rdxdr,
                                 double *d_out, void *stream) {
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    rdxdr_reduce_kernel<<<div_up(natoms * nb, 256), 256, 0, s>>>(natoms, nb, d_row_start, d_rdxdr, d_out);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_tune_set(int knob, int value) {
    PXFF_REQUIRE(knob >= 0 && knob < PXFF_TUNE_COUNT, "unknown tuning knob");
    PXFF_REQUIRE(value >= g_tune_lo[knob] && value <= g_tune_hi[knob], "tuning value out of range");
    if (knob == PXFF_TUNE_IMAGE_WAVES_PER_BLOCK) PXFF_REQUIRE(value == 1 || value == 2 || value == 4, "waves per block must be 1, 2 or 4");
    g_tune[knob] = value;
    return 0;
}

extern "C" int pxff_tune_get(int knob) {
    if (knob < 0 || knob >= PXFF_TUNE_COUNT) return -1;
    return g_tune[knob];
}
