// SO(4)/SNAP bispectrum kernels for gfx950 (FP64, wave64).
//
// Device layout (HBM), per atom i (tile t = i/64, lane = i%64):
//   utot[t][nhalf][lane] complex  U_tot half rows (mb <= l/2), self term,
//                                 u_norm (tile-major; the Z, B and Y kernels
//                                 run one atom per lane)
//   z[t][nz][lane]  complex  Z in the reference's idxz order, tile-major
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

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <atomic>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "pxff_common.h"
#include "so4_dev.h"
#include "so4_tables.h"

namespace pxff {

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

int find_so4_tables(int twol, int norm, const So4Dev **out) {
    int dev = 0;
    PXFF_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tab_mutex);
    auto it = g_tabs.find(std::make_pair(dev, twol * 2 + (norm ? 1 : 0)));
    if (it == g_tabs.end()) {
        set_error("the 2J=" + std::to_string(twol) + " tables are not built on this device yet: call any allocating "
                  "SO4 entry point (e.g. pxff_so4_tangent) with the same 2J and normalisation once first");
        return PXFF_EINVAL;
    }
    *out = &it->second;
    return 0;
}

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
    d.nbrow = t.nbrow;
    if ((rc = upload(t.b_brow0, &d.b_brow0))) return rc;
    d.zlevel_max = 0;
    for (int l = 0; l + 1 < (int)t.blk_level_start.size(); l++)
        d.zlevel_max = std::max(d.zlevel_max,
                                (t.blk_level_start[l + 1] - t.blk_level_start[l]) * (l / 2 + 1) * (l + 1));
    d.nzgroups = t.nzgroups;
    if ((rc = upload(t.zgroups, &d.zgroups))) return rc;
    // inverse of the block targets: the (block, factor) pairs of each b
    std::vector<int> bsrc(3 * (size_t)t.nb, -1);
    std::vector<double> bsrcf(3 * (size_t)t.nb, 0.0);
    for (int k = 0; k < t.nblk; k++)
        for (int q = 0; q < 3; q++) {
            const int b = t.blk_tb[3 * k + q];
            if (b < 0) continue;
            int slot = 0;
            while (slot < 3 && bsrc[3 * b + slot] >= 0) slot++;
            if (slot == 3) {
                set_error("more than three blocks target one bispectrum component");
                return PXFF_EINVAL;
            }
            bsrc[3 * b + slot] = k;
            bsrcf[3 * b + slot] = t.blk_tf[3 * k + q];
        }
    if ((rc = upload(bsrc, &d.b_src))) return rc;
    if ((rc = upload(bsrcf, &d.b_srcf))) return rc;
    // half-U element of each Z element: Z block (l1, l2, l) element loc pairs
    // with U^l element loc (same (mb, ma), mb <= l/2)
    std::vector<int> zh((size_t)t.nz, 0);
    for (int k = 0; k < t.nblk; k++) {
        int l = 0;
        while (l + 1 < t.ldim && k >= t.blk_level_start[l + 1]) l++;
        const int nloc = (l / 2 + 1) * (l + 1);
        for (int q = 0; q < nloc; q++) {
            zh[t.blk_zoff[k] + q] = t.hoff[l] + q;
        }
    }
    if ((rc = upload(zh, &d.zh))) return rc;
    std::vector<double> hscale((size_t)t.nhalf, 0.0);
    for (int h = 0; h < t.nhalf; h++) hscale[h] = 2.0 * t.h_weight[h] * t.lnorm[t.h_level[h]];
    if ((rc = upload(hscale, &d.hscale))) return rc;
    // Y tasks, most blocks first (longest wavefronts start first)
    std::vector<std::array<int, YTASK_STRIDE>> yt;
    for (int l = 0; l < t.ldim; l++) {
        const int nloc = (l / 2 + 1) * (l + 1);
        for (int loc0 = 0; loc0 < nloc; loc0 += YTASK_ROWS)
            yt.push_back({l, loc0, std::min(YTASK_ROWS, nloc - loc0), t.blk_level_start[l],
                          t.blk_level_start[l + 1], t.hoff[l] + loc0});
    }
    std::stable_sort(yt.begin(), yt.end(), [](const std::array<int, YTASK_STRIDE> &a,
                                              const std::array<int, YTASK_STRIDE> &b) {
        return (a[4] - a[3]) * a[2] > (b[4] - b[3]) * b[2];
    });
    std::vector<int> ytf;
    for (auto &q : yt) ytf.insert(ytf.end(), q.begin(), q.end());
    d.nytask = (int)yt.size();
    if ((rc = upload(ytf, &d.ytask))) return rc;
    g_tabs[key] = d;
    *out = &g_tabs[key];
    return 0;
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
        // the coefficient is folded into one factor of each product (c1 ca,
        // c1 u), so every derivative term is a complex multiply-add
        if (m < l) {
            double c1 = rpq[(l - m) * (TWOJ + 1) + q];
            cd c1a = cscale(ck.ca, c1);
            un = cmul(c1a, u[m]);
            if (GRAD) {
                const cd c1u = cscale(u[m], c1);
#pragma unroll
                for (int k = 0; k < 3; k++) dn[k] = cmadd(c1a, du[k][m], cmul(ck.cda[k], c1u));
            }
        }
        if (m >= 1) {
            double c2 = rpq[m * (TWOJ + 1) + q];
            cd c2b = cscale(ck.cb, -c2);
            un = cmadd(c2b, u[m - 1], un);
            if (GRAD) {
                const cd c2u = cscale(u[m - 1], -c2);
#pragma unroll
                for (int k = 0; k < 3; k++) dn[k] = cmadd(c2b, du[k][m - 1], cmadd(ck.cdb[k], c2u, dn[k]));
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
        // without derivative rows: keep the coefficient loads of a level
        // inside it (hoisted loads of later levels would hold registers
        // across the whole sweep)
        if (!GRAD) asm volatile("" ::: "memory");
        const int mbe = (l <= mb) ? 0 : mb;
        row_step<TWOJ, GRAD>(l, l - mbe, rpq, ck, u, du);
        if (l == mb) row_flip<TWOJ, GRAD>(l, u, du);
        f(l, u, du, 2 * mb <= l);
    }
}

// The same sweep with the recursion coefficients taken out.  The
// coefficients are sqrt(p / q) (init_rootpqarray, SNAP.py:861-868):
// c1 = sqrt(l - m) / sqrt(q), c2 = sqrt(m) / sqrt(q) with q = l - mb (q = l
// on row 0), so with
//   u^l_mb[m] = sqrt(m! (l - m)! / (mb! (l - mb)!)) W^l[m]
// the row recursion becomes W^l[m] = ca W^{l-1}[m] - cb W^{l-1}[m-1]
// (binomial, no coefficients: W^l on row 0 is C(l, m) ca^(l-m) (-cb)^m),
// and the inversion flip keeps its form (the factor is symmetric under
// m <-> l - m).  f(l, w, half) receives W; the factor (so4_wscale) is
// image-independent, so sums over images take it once.  |W| <= C(l, l/2)
// (252 at 2J = 10).
template <int TWOJ, typename F>
__device__ inline void sweep_rows_w(int mb, const CK &ck, F &&f) {
    cd w[TWOJ + 1];
#pragma unroll
    for (int m = 0; m <= TWOJ; m++) w[m] = cmake(0.0, 0.0);
    w[0] = cmake(1.0, 0.0);
    f(0, w, mb == 0);
    const cd mcb = cmake(-ck.cb.re, -ck.cb.im);
#pragma unroll
    for (int l = 1; l <= TWOJ; l++) {
#pragma unroll
        for (int m = TWOJ; m >= 0; --m) {
            if (m > l) continue;
            if (m == l)
                w[m] = cmul(mcb, w[m - 1]);
            else if (m >= 1)
                w[m] = cmadd(mcb, w[m - 1], cmul(ck.ca, w[m]));
            else
                w[m] = cmul(ck.ca, w[m]);
        }
        if (l == mb) {
            cd t[TWOJ + 1];
#pragma unroll
            for (int m = 0; m <= TWOJ; m++) {
                if (m > l) continue;
                const cd o = w[l - m];
                t[m] = ((l - m) & 1) ? cmake(-o.re, o.im) : cmake(o.re, -o.im);
            }
#pragma unroll
            for (int m = 0; m <= TWOJ; m++)
                if (m <= l) w[m] = t[m];
        }
        f(l, w, 2 * mb <= l);
    }
}

// sqrt(m! (l - m)! / (mb! (l - mb)!)), the factor of sweep_rows_w, from
// sf[k] = sqrt(k!) and isf[k] = 1 / sqrt(k!).
__device__ inline double so4_wscale(int l, int m, int mb, const double *sf, const double *isf) {
    return (sf[m] * sf[l - m]) * (isf[mb] * isf[l - mb]);
}

// ------------------------------------------------------ su(2) generators
// U^l is a representation of SU(2): U^l(g h) = U^l(g) U^l(h) for the
// Cayley-Klein matrix g = [[conj a, -conj b], [b, a]] of an image.  Hence
// dU^l/dr_k = U^l M^l(xi_k) with xi_k = g^-1 dg/dr_k = [[i p, q], [-conj q,
// -i p]] in su(2), and M^l(xi) is tridiagonal in the column index:
//   M[m][m] = i p (l - 2m),  M[m][m+1] = c_m q,  M[m+1][m] = -c_m conj q,
//   c_m = sqrt((m + 1)(l - m)).
// Right multiplication keeps the row, so on the half rows the adjoint
// contraction of compute_duarray_recursive / compute_dbidrj (SNAP.py:
// 958-1017, 1143-1257) with Y becomes
//   Re sum conj(dU_k) Y = p_k A1 + Re q_k A2 + Im q_k A3,
//   A1 = sum (l-2m) Im(conj U_m Y_m),
//   A2 = sum c_m Re(conj U_m Y_m+1 - conj U_m+1 Y_m),
//   A3 = sum c_m Im(conj U_m Y_m+1 + conj U_m+1 Y_m),
// so the sweep needs U alone (no dU recursion, no derivative registers).
// p_k, q_k from a, b and their gradients: i p = a conj(da) + conj(b) db,
// q = conj(b) da - a conj(db).
__device__ inline void su2_generators(const CK &ck, double p[3], cd q[3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const cd ca = ck.ca, cb = ck.cb, cda = ck.cda[k], cdb = ck.cdb[k];  // conjugates of a, b, da, db
        p[k] = (ca.re * cda.im - ca.im * cda.re) + (cb.im * cdb.re - cb.re * cdb.im);
        q[k] = cmake((cb.re * cda.re + cb.im * cda.im) - (ca.re * cdb.re + ca.im * cdb.im),
                     (cb.im * cda.re - cb.re * cda.im) - (ca.re * cdb.im - ca.im * cdb.re));
    }
}

// Accumulate A_u = sum Re(conj U_m Y_m) and A1..A3 (above) of one half row
// of level l (u: the row, y: Y of the same row).
template <int TWOJ>
__device__ inline void generator_dots(int l, const cd *u, const cd *y, double &au, double &a1, double &a2,
                                      double &a3) {
#pragma unroll
    for (int m = 0; m <= TWOJ; m++) {
        if (m > l) continue;
        au = fma(u[m].re, y[m].re, fma(u[m].im, y[m].im, au));
        if (l != 2 * m) a1 = fma((double)(l - 2 * m), u[m].re * y[m].im - u[m].im * y[m].re, a1);
        if (m < l) {
            const double c = sqrt((double)((m + 1) * (l - m)));  // folded per (l, m)
            const double t1 = fma(u[m].re, y[m + 1].re, u[m].im * y[m + 1].im) -
                              fma(u[m + 1].re, y[m].re, u[m + 1].im * y[m].im);
            const double t2 = fma(u[m].re, y[m + 1].im, -u[m].im * y[m + 1].re) +
                              fma(u[m + 1].re, y[m].im, -u[m + 1].im * y[m].re);
            a2 = fma(c, t1, a2);
            a3 = fma(c, t2, a3);
        }
    }
}

// ---------------------------------------------------------------- U_tot
// One wavefront per centre (64 threads); NG groups of G lanes each take one
// image per round; per-group partial sums live in LDS and are reduced in a
// fixed order at the end (deterministic).  With TANGENT the kernel instead
// accumulates dU_tot . dr_e (forward-mode derivative of U_tot).
// The U_tot sums hold no derivative rows: registers are kept within three
// wavefronts per SIMD (the tangent variant needs more than that).
template <int TWOJ, bool TANGENT, bool TILED, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void so4_accum_kernel(int64_t natoms, const int64_t *__restrict__ img_start,
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
    // MODE 0: each group keeps its own partial array, summed in group order
    // at the end.  MODE 1: all groups add into one LDS array (ds_add_f64;
    // one wavefront per centre issues every add, so the order is fixed).
    // MODE 2 / 3: MODE groups share each partial array (groups g, g + NP,
    // g + 2 NP, ...); they add in MODE masked passes of one wavefront, in
    // program order: 1/MODE of the LDS, so more resident wavefronts.
    constexpr bool LDSADD = MODE == 1;
    constexpr int NPASS = MODE >= 2 ? MODE : 1;
    constexpr int NP = MODE == 1 ? 1 : (NG + NPASS - 1) / NPASS;
    __shared__ cd part[NP][NH];
    __shared__ double rpq[(TWOJ + 1) * (TWOJ + 1)];
    __shared__ double lnorm[TWOJ + 1];
    // staged per image (64 at a time): ca, cb, fc w (0: no image); with
    // TANGENT ca, cb, w fc p(eta), w fc q(eta), w dfc (rhat . dr), valid
    __shared__ double stg[TANGENT ? 9 : 5][64];
    const int lane = threadIdx.x;
    const int g = lane / G, mb = lane % G;
    for (int k = lane; k < (TWOJ + 1) * (TWOJ + 1); k += 64) rpq[k] = rootpq_g[k];
    for (int k = lane; k <= TWOJ; k += 64) lnorm[k] = lnorm_g[k];
    if (TILED && blockIdx.x == 0 && (natoms & 63)) {
        // padding lanes of the last tile (read by the Z kernel) are zero
        cd *last = reinterpret_cast<cd *>(out) + (size_t)(natoms >> 6) * NH * 64;
        for (int k = lane; k < NH * 64; k += 64)
            if ((k & 63) >= (natoms & 63)) last[k] = cmake(0.0, 0.0);
    }

    for (int64_t c = blockIdx.x; c < natoms; c += gridDim.x) {
        const int64_t i = work_order ? work_order[c] : c;
        __syncthreads();
        for (int k = lane; k < NP * NH; k += 64) (&part[0][0])[k] = cmake(0.0, 0.0);
        __syncthreads();
        const int64_t s = img_start[i], e = img_start[i + 1];
        if constexpr (!TANGENT) {
            // up to 64 images at a time: their Cayley-Klein parameters and
            // weights, one lane each, staged in LDS for the row sweeps (the G
            // lanes of an image would otherwise each compute them)
            for (int64_t b64 = s; b64 < e; b64 += 64) {
                const int64_t e64 = b64 + 64 < e ? b64 + 64 : e;
                {
                    const int64_t img = b64 + lane;
                    bool ok = img < e64;
                    double x = 0, y = 0, z = 1.0, r = 1.0, w = 0.0;
                    if (ok) {
                        x = img_r[3 * img];
                        y = img_r[3 * img + 1];
                        z = img_r[3 * img + 2];
                        r = sqrt(x * x + y * y + z * z);
                        w = atom_w[img_j[img]];
                        if (r < 1e-8) ok = false;  // SNAP.py:1487 (padding rows)
                    }
                    if (!ok) {
                        x = 0;
                        y = 0;
                        z = 1.0;
                        r = 1.0;
                    }
                    double fc, dfc;
                    cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
                    CK ck;
                    cayley_klein(x, y, z, r, rfac0 * M_PI * r / rcut, false, ck);
                    stg[0][lane] = ck.ca.re;
                    stg[1][lane] = ck.ca.im;
                    stg[2][lane] = ck.cb.re;
                    stg[3][lane] = ck.cb.im;
                    stg[4][lane] = ok ? fc * w : 0.0;  // a zero weight adds nothing either way
                }
                __syncthreads();
                for (int64_t base = b64; base < e64; base += NG) {
                    const int64_t img = base + g;
                    const int si = (g < NG && img < e64) ? (int)(img - b64) : 0;
                    const bool valid = (g < NG) && (img < e64) && stg[4][si] != 0.0;
                    CK ck;
                    ck.ca = cmake(stg[0][si], stg[1][si]);
                    ck.cb = cmake(stg[2][si], stg[3][si]);
                    const double sc = stg[4][si];
                    const int pidx = g < NG ? g % NP : 0;  // index the __shared__ array itself: LDS atomics
                    sweep_rows<TWOJ, false>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
                        const bool act = valid && half;
                        const int h0 = hoff_of(l) + mb * (l + 1);
#pragma unroll
                        for (int pass = 0; pass < NPASS; pass++) {
                            // MODE >= 2: the groups sharing an array add in
                            // separate passes, ordered by workgroup barriers
                            if (act && (NPASS == 1 || g / NP == pass)) {
#pragma unroll
                                for (int m = 0; m <= TWOJ; m++) {
                                    if (m > l) continue;
                                    if (LDSADD) {
                                        atomicAdd(&part[pidx][h0 + m].re, sc * u[m].re);
                                        atomicAdd(&part[pidx][h0 + m].im, sc * u[m].im);
                                    } else {
                                        cd v = part[pidx][h0 + m];
                                        part[pidx][h0 + m] = cmake(v.re + sc * u[m].re, v.im + sc * u[m].im);
                                    }
                                }
                            }
                            if (NPASS > 1) __syncthreads();
                        }
                    });
                }
                __syncthreads();  // the next 64 images reuse the stage
            }
        } else {
            // forward-mode: dU . dr_e = U M(eta_e) with eta_e = sum_k dr_k xi_k
            // (the su(2) generators of the adjoint kernel), so a row needs U
            // alone: (U M(eta))_m = i p (l-2m) U_m + c_m-1 q U_m-1 - c_m conj(q) U_m+1
            for (int64_t b64 = s; b64 < e; b64 += 64) {
                const int64_t e64 = b64 + 64 < e ? b64 + 64 : e;
                {
                    const int64_t img = b64 + lane;
                    bool ok = img < e64;
                    double x = 0, y = 0, z = 0, r = 1.0, w = 0.0;
                    if (ok) {
                        x = img_r[3 * img];
                        y = img_r[3 * img + 1];
                        z = img_r[3 * img + 2];
                        r = sqrt(x * x + y * y + z * z);
                        w = atom_w[img_j[img]];
                        if (r < 1e-8) ok = false;  // SNAP.py:1487 (padding rows)
                    }
                    if (!ok) {
                        x = 0;
                        y = 0;
                        z = 1.0;
                        r = 1.0;
                    }
                    double fc, dfc;
                    cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
                    CK ck;
                    cayley_klein(x, y, z, r, rfac0 * M_PI * r / rcut, true, ck);
                    double p[3];
                    cd q[3];
                    su2_generators(ck, p, q);
                    double d0 = 0, d1 = 0, d2 = 0;
                    if (ok) {
                        d0 = dr[3 * img];
                        d1 = dr[3 * img + 1];
                        d2 = dr[3 * img + 2];
                    }
                    const double wf = ok ? w * fc : 0.0;
                    stg[0][lane] = ck.ca.re;
                    stg[1][lane] = ck.ca.im;
                    stg[2][lane] = ck.cb.re;
                    stg[3][lane] = ck.cb.im;
                    stg[4][lane] = wf * (p[0] * d0 + p[1] * d1 + p[2] * d2);
                    stg[5][lane] = wf * (q[0].re * d0 + q[1].re * d1 + q[2].re * d2);
                    stg[6][lane] = wf * (q[0].im * d0 + q[1].im * d1 + q[2].im * d2);
                    // D . dr = fc (dU . dr) + dfc (rhat . dr) U
                    stg[7][lane] = ok ? w * dfc * (x * d0 + y * d1 + z * d2) / r : 0.0;
                    stg[8][lane] = ok ? 1.0 : 0.0;
                }
                __syncthreads();
                for (int64_t base = b64; base < e64; base += NG) {
                    const int64_t img = base + g;
                    const int si = (g < NG && img < e64) ? (int)(img - b64) : 0;
                    const bool valid = (g < NG) && (img < e64) && stg[8][si] != 0.0;
                    CK ck;
                    ck.ca = cmake(stg[0][si], stg[1][si]);
                    ck.cb = cmake(stg[2][si], stg[3][si]);
                    const double P = stg[4][si], Qr = stg[5][si], Qi = stg[6][si], Gu = stg[7][si];
                    const int pidx = g < NG ? g % NP : 0;  // index the __shared__ array itself: LDS atomics
                    sweep_rows<TWOJ, false>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
                        const bool act = valid && half;
                        const int h0 = hoff_of(l) + mb * (l + 1);
#pragma unroll
                        for (int pass = 0; pass < NPASS; pass++) {
                            if (act && (NPASS == 1 || g / NP == pass)) {
#pragma unroll
                                for (int m = 0; m <= TWOJ; m++) {
                                    if (m > l) continue;
                                    const double pm = P * (double)(l - 2 * m);
                                    double re = fma(-pm, u[m].im, Gu * u[m].re);
                                    double im = fma(pm, u[m].re, Gu * u[m].im);
                                    if (m >= 1) {
                                        const double c = sqrt((double)(m * (l - m + 1)));  // c_(m-1)
                                        re = fma(c, Qr * u[m - 1].re - Qi * u[m - 1].im, re);
                                        im = fma(c, Qr * u[m - 1].im + Qi * u[m - 1].re, im);
                                    }
                                    if (m < l) {
                                        const double c = sqrt((double)((m + 1) * (l - m)));  // c_m
                                        re = fma(-c, Qr * u[m + 1].re + Qi * u[m + 1].im, re);
                                        im = fma(-c, Qr * u[m + 1].im - Qi * u[m + 1].re, im);
                                    }
                                    if (LDSADD) {
                                        atomicAdd(&part[pidx][h0 + m].re, re);
                                        atomicAdd(&part[pidx][h0 + m].im, im);
                                    } else {
                                        cd v = part[pidx][h0 + m];
                                        part[pidx][h0 + m] = cmake(v.re + re, v.im + im);
                                    }
                                }
                            }
                            if (NPASS > 1) __syncthreads();
                        }
                    });
                }
                __syncthreads();  // the next 64 images reuse the stage
            }
        }
        __syncthreads();
        for (int h = lane; h < NH; h += 64) {
            cd acc = cmake(0.0, 0.0);
            for (int q = 0; q < NP; q++) acc = cadd(acc, part[q][h]);
            // level / row / column of h (compile-time level offsets and
            // divisors: compares and selects instead of loops and divisions)
            int l = 0, hm = 0, hma = 0;
#pragma unroll
            for (int L = 0; L <= TWOJ; L++)
                if (h >= hoff_of(L) && h < hoff_of(L + 1)) {
                    l = L;
                    hm = (h - hoff_of(L)) / (L + 1);
                    hma = (h - hoff_of(L)) % (L + 1);
                }
            if (!TANGENT && hm == hma) acc.re += fcut0 * site_w[i];  // addself_uarraytot
            acc = cscale(acc, lnorm[l]);                               // ulisttot *= u_norm
            if (TILED)  // [atom / 64][h][atom % 64], the layout of the Z kernel's loads
                reinterpret_cast<cd *>(out)[((size_t)(i >> 6) * NH + h) * 64 + (i & 63)] = acc;
            else
                reinterpret_cast<cd *>(out)[(size_t)i * NH + h] = acc;
        }
    }
}

// The per-image values the U_tot sweeps read (compute_uarray_recursive's
// a, b, SNAP.py:1515-1555, with the cutoff and weight of add_uarraytot):
// conj(a), conj(b) and w fc (0 for coincident atoms, SNAP.py:1487), and for
// the tangent w fc p(eta), w fc q(eta), w dfc (rhat . dr) and a valid flag
// in place of w fc.
template <bool TANGENT>
__device__ inline void so4_stage_image(bool ok, int64_t img, const int64_t *__restrict__ img_j,
                                       const double *__restrict__ img_r, const double *__restrict__ atom_w,
                                       const double *__restrict__ dr, double rcut, double rfac0, int cutoff_id,
                                       double *v) {
    double x = 0, y = 0, z = 1.0, r = 1.0, w = 0.0;
    if (ok) {
        x = img_r[3 * img];
        y = img_r[3 * img + 1];
        z = img_r[3 * img + 2];
        r = sqrt(x * x + y * y + z * z);
        w = atom_w[img_j[img]];
        if (r < 1e-8) ok = false;  // SNAP.py:1487 (padding rows)
    }
    if (!ok) {
        x = 0;
        y = 0;
        z = 1.0;
        r = 1.0;
    }
    double fc, dfc;
    cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
    CK ck;
    cayley_klein(x, y, z, r, rfac0 * M_PI * r / rcut, TANGENT, ck);
    v[0] = ck.ca.re;
    v[1] = ck.ca.im;
    v[2] = ck.cb.re;
    v[3] = ck.cb.im;
    if constexpr (!TANGENT) {
        v[4] = ok ? fc * w : 0.0;  // a zero weight adds nothing either way
    } else {
        double p[3];
        cd q[3];
        su2_generators(ck, p, q);
        double d0 = 0, d1 = 0, d2 = 0;
        if (ok) {
            d0 = dr[3 * img];
            d1 = dr[3 * img + 1];
            d2 = dr[3 * img + 2];
        }
        const double wf = ok ? w * fc : 0.0;
        v[4] = wf * (p[0] * d0 + p[1] * d1 + p[2] * d2);
        v[5] = wf * (q[0].re * d0 + q[1].re * d1 + q[2].re * d2);
        v[6] = wf * (q[0].im * d0 + q[1].im * d1 + q[2].im * d2);
        // D . dr = fc (dU . dr) + dfc (rhat . dr) U
        v[7] = ok ? w * dfc * (x * d0 + y * d1 + z * d2) / r : 0.0;
        v[8] = ok ? 1.0 : 0.0;
    }
}

// Record stride (doubles) of the staged values: 5 (9 with the tangent) used,
// padded to 16-byte multiples.
template <bool TANGENT>
constexpr int so4_stage_stride() { return TANGENT ? 10 : 6; }

// The staged values of every image e < min(img_start[natoms], cap), one
// thread per image (grid stride): the U_tot sweep kernel then reads one
// record per image instead of evaluating the cutoff and the Cayley-Klein
// parameters between its sweeps.
template <bool TANGENT>
__global__ __launch_bounds__(256) void so4_stage_kernel(int64_t natoms, int64_t cap,
                                                        const int64_t *__restrict__ img_start,
                                                        const int64_t *__restrict__ img_j,
                                                        const double *__restrict__ img_r,
                                                        const double *__restrict__ atom_w,
                                                        const double *__restrict__ dr, double rcut, double rfac0,
                                                        int cutoff_id, double *__restrict__ stage) {
    constexpr int NS = so4_stage_stride<TANGENT>();
    const int64_t n = img_start[natoms] < cap ? img_start[natoms] : cap;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double v[TANGENT ? 9 : 5];
        so4_stage_image<TANGENT>(true, e, img_j, img_r, atom_w, dr, rcut, rfac0, cutoff_id, v);
        double *o = stage + e * NS;
#pragma unroll
        for (int k = 0; k < (TANGENT ? 9 : 5); k++) o[k] = v[k];
    }
}

// U_tot (and its tangent) with the partial sums in registers.  One
// wavefront per centre, G = 2J/2+1 lanes per image as in so4_accum_kernel,
// but lane (g, mb) adds row mb of every level of all its images into
// registers (one slot per (l, m), 45 complex at 2J=8), so a centre's images
// cost no LDS traffic; once per centre the NG groups' sums are reduced in
// LDS in a fixed order (groups g < NP store, groups g + NP add, then the NP
// arrays are summed in order: deterministic).  2J <= 8 (at 2J=10 the
// register partials would exceed the register file; so4_accum_kernel).
template <int TWOJ, bool TANGENT, bool STAGED>
__device__ __forceinline__ void so4_utot_body(
    int64_t natoms, const int64_t *__restrict__ img_start, const int64_t *__restrict__ img_j,
    const double *__restrict__ img_r, const double *__restrict__ atom_w, const double *__restrict__ site_w,
    const double *__restrict__ dr, const int64_t *__restrict__ work_order, const double *rootpq_g,
    const double *lnorm_g, double rcut, double rfac0, int cutoff_id, double fcut0, double *__restrict__ out,
    const double *__restrict__ stage) {
    constexpr int G = TWOJ / 2 + 1;
    constexpr int NG = 64 / G;
    constexpr int NH = nhalf_of(TWOJ);
    constexpr int NSLOT = (TWOJ + 1) * (TWOJ + 2) / 2;  // (l, m <= l) of every level
    constexpr int NP = (NG + 1) / 2;                    // LDS partial arrays
    __shared__ cd part[NP][NH];
    __shared__ double sf[TWOJ + 1], isf[TWOJ + 1];  // sqrt(k!), 1 / sqrt(k!)
    __shared__ double lnorm[TWOJ + 1];
    // staged per image: ca, cb, fc w (0: no image); with TANGENT ca, cb,
    // w fc p(eta), w fc q(eta), w dfc (rhat . dr), valid
    __shared__ double stg[TANGENT ? 9 : 5][64];
    const int lane = threadIdx.x;
    const int g = lane / G, mb = lane % G;
    (void)rootpq_g;  // the W recursion has no coefficients
    if (lane <= TWOJ) {
        double f = 1.0;
        for (int q = 2; q <= lane; q++) f *= q;
        sf[lane] = sqrt(f);
        isf[lane] = 1.0 / sqrt(f);
    }
    for (int k = lane; k <= TWOJ; k += 64) lnorm[k] = lnorm_g[k];
    if (blockIdx.x == 0 && (natoms & 63)) {
        // padding lanes of the last tile (read by the Z kernel) are zero
        cd *last = reinterpret_cast<cd *>(out) + (size_t)(natoms >> 6) * NH * 64;
        for (int k = lane; k < NH * 64; k += 64)
            if ((k & 63) >= (natoms & 63)) last[k] = cmake(0.0, 0.0);
    }
    __syncthreads();
    for (int64_t c = blockIdx.x; c < natoms; c += gridDim.x) {
        // without a work order, the eight centres of a 64-atom tile whose
        // elements share a 128-byte line of the tile-major output run on one
        // XCD (workgroups go to the XCDs round-robin, and the grid is a
        // multiple of 8 or a single round): swap the 3-bit fields c % 8 and
        // (c / 8) % 8, a bijection on every whole tile (a partial last tile
        // keeps the identity)
        const int64_t cx = (c & ~(int64_t)63) | ((c & 7) << 3) | ((c >> 3) & 7);
        const int64_t i = work_order ? work_order[c] : ((c | 63) < natoms ? cx : c);
        const int64_t s = img_start[i], e = img_start[i + 1];
        const double wself = TANGENT ? 0.0 : site_w[i];  // loaded ahead of the sweeps
        if (s == e) {  // no image: the partial arrays start (and stay) at zero
            for (int k = lane; k < NP * NH; k += 64) part[k / NH][k % NH] = cmake(0.0, 0.0);
            __syncthreads();
        }
        for (int64_t b64 = s; b64 < e; b64 += 64) {
            const int64_t e64 = b64 + 64 < e ? b64 + 64 : e;
            if constexpr (!STAGED) {
                const int64_t img = b64 + lane;
                double v[TANGENT ? 9 : 5];
                so4_stage_image<TANGENT>(img < e64, img, img_j, img_r, atom_w, dr, rcut, rfac0, cutoff_id, v);
#pragma unroll
                for (int q = 0; q < (TANGENT ? 9 : 5); q++) stg[q][lane] = v[q];
            }
            if constexpr (!STAGED) __syncthreads();
            // the chunk's sums: live only over its sweeps (not across the
            // staging above), flushed to the partial arrays below
            cd acc[NSLOT];
#pragma unroll
            for (int q = 0; q < NSLOT; q++) acc[q] = cmake(0.0, 0.0);
            constexpr int NV = TANGENT ? 9 : 5;
            for (int64_t base = b64; base < e64; base += NG) {
                const int64_t img = base + g;
                const int si = (g < NG && img < e64) ? (int)(img - b64) : 0;
                double sv[NV];
                if constexpr (STAGED) {
                    // the group's record (loaded where it is used: a copy
                    // loaded one pass ahead costs registers, and its spills
                    // cost more than the wait: 39 against 36 us)
                    const bool ok = g < NG && img < e64;
                    const double *rec = stage + (ok ? img : b64) * so4_stage_stride<TANGENT>();
#pragma unroll
                    for (int q = 0; q < NV; q++) sv[q] = ok ? rec[q] : 0.0;
                } else {
#pragma unroll
                    for (int q = 0; q < NV; q++) sv[q] = stg[q][si];
                }
                const bool valid = (g < NG) && (img < e64) && sv[TANGENT ? 8 : 4] != 0.0;
                CK ck;
                ck.ca = cmake(sv[0], sv[1]);
                ck.cb = cmake(sv[2], sv[3]);
                if constexpr (!TANGENT) {
                    const double sc = sv[4];
                    sweep_rows_w<TWOJ>(mb, ck, [&](int l, const cd *u, bool half) {
                        if (!(valid && half)) return;
                        const int q0 = l * (l + 1) / 2;
#pragma unroll
                        for (int m = 0; m <= TWOJ; m++) {
                            if (m > l) continue;
                            acc[q0 + m].re = fma(sc, u[m].re, acc[q0 + m].re);
                            acc[q0 + m].im = fma(sc, u[m].im, acc[q0 + m].im);
                        }
                    });
                } else {
                    // dU . dr_e = U M(eta_e): (U M)_m = i p (l-2m) U_m + c_m-1 q U_m-1 - c_m conj(q) U_m+1,
                    // c_m = sqrt((m+1)(l-m)); on W (U_m = s_m W_m, s_m-1 / s_m = sqrt((l-m+1)/m))
                    // the coefficients become c_m-1 s_m-1 / s_m = l-m+1 and c_m s_m+1 / s_m = m+1
                    const double P = sv[4], Qr = sv[5], Qi = sv[6], Gu = sv[7];
                    sweep_rows_w<TWOJ>(mb, ck, [&](int l, const cd *u, bool half) {
                        if (!(valid && half)) return;
                        const int q0 = l * (l + 1) / 2;
#pragma unroll
                        for (int m = 0; m <= TWOJ; m++) {
                            if (m > l) continue;
                            const double pm = P * (double)(l - 2 * m);
                            double re = fma(-pm, u[m].im, Gu * u[m].re);
                            double im = fma(pm, u[m].re, Gu * u[m].im);
                            if (m >= 1) {
                                const double cc = (double)(l - m + 1);
                                re = fma(cc, Qr * u[m - 1].re - Qi * u[m - 1].im, re);
                                im = fma(cc, Qr * u[m - 1].im + Qi * u[m - 1].re, im);
                            }
                            if (m < l) {
                                const double cc = (double)(m + 1);
                                re = fma(-cc, Qr * u[m + 1].re + Qi * u[m + 1].im, re);
                                im = fma(-cc, Qr * u[m + 1].im - Qi * u[m + 1].re, im);
                            }
                            acc[q0 + m].re += re;
                            acc[q0 + m].im += im;
                        }
                    });
                }
            }
            __syncthreads();  // the next 64 images reuse the stage
            // reduce the NG groups' register sums: g < NP store (add after
            // the centre's first chunk), g >= NP add
#pragma unroll
            for (int pass = 0; pass < 2; pass++) {
                const bool mine = g < NG && (pass == 0 ? g < NP : g >= NP);
                const bool store = pass == 0 && b64 == s;
                const int pidx = g < NG ? (pass == 0 ? g : g - NP) : 0;
                if (mine) {
#pragma unroll
                    for (int l = 0; l <= TWOJ; l++) {
                        if (2 * mb > l) continue;
                        const int h0 = hoff_of(l) + mb * (l + 1), q0 = l * (l + 1) / 2;
#pragma unroll
                        for (int m = 0; m <= l; m++) {
                            const cd v = cscale(acc[q0 + m], so4_wscale(l, m, mb, sf, isf));
                            if (store)
                                part[pidx][h0 + m] = v;
                            else
                                part[pidx][h0 + m] = cadd(part[pidx][h0 + m], v);
                        }
                    }
                }
                __syncthreads();  // (odd NG: array NP - 1 has no second group)
            }
        }
        for (int h = lane; h < NH; h += 64) {
            cd a = cmake(0.0, 0.0);
            for (int q = 0; q < NP; q++) a = cadd(a, part[q][h]);
            int l = 0, hm = 0, hma = 0;
#pragma unroll
            for (int L = 0; L <= TWOJ; L++)
                if (h >= hoff_of(L) && h < hoff_of(L + 1)) {
                    l = L;
                    hm = (h - hoff_of(L)) / (L + 1);
                    hma = (h - hoff_of(L)) % (L + 1);
                }
            if (!TANGENT && hm == hma) a.re += fcut0 * wself;  // addself_uarraytot
            a = cscale(a, lnorm[l]);                                 // ulisttot *= u_norm
            reinterpret_cast<cd *>(out)[((size_t)(i >> 6) * NH + h) * 64 + (i & 63)] = a;
        }
        __syncthreads();  // the next centre rewrites part
    }
}

template <int TWOJ, bool TANGENT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1))) void so4_utot_kernel(
    int64_t natoms, const int64_t *__restrict__ img_start, const int64_t *__restrict__ img_j,
    const double *__restrict__ img_r, const double *__restrict__ atom_w, const double *__restrict__ site_w,
    const double *__restrict__ dr, const int64_t *__restrict__ work_order, const double *rootpq_g,
    const double *lnorm_g, double rcut, double rfac0, int cutoff_id, double fcut0, double *__restrict__ out) {
    so4_utot_body<TWOJ, TANGENT, false>(natoms, img_start, img_j, img_r, atom_w, site_w, dr, work_order, rootpq_g,
                                        lnorm_g, rcut, rfac0, cutoff_id, fcut0, out, nullptr);
}

// U_tot from the staged per-image records (so4_stage_kernel): without the
// cutoff and Cayley-Klein evaluation between its sweeps the kernel fits two
// wavefronts per SIMD.
template <int TWOJ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void so4_utot_staged_kernel(
    int64_t natoms, const int64_t *__restrict__ img_start, const double *__restrict__ site_w,
    const int64_t *__restrict__ work_order, const double *lnorm_g, double fcut0, const double *__restrict__ stage,
    double *__restrict__ out) {
    so4_utot_body<TWOJ, false, true>(natoms, img_start, nullptr, nullptr, nullptr, site_w, nullptr, work_order,
                                     nullptr, lnorm_g, 0.0, 0.0, 1, fcut0, out, stage);
}

// B (compute_bi, SNAP.py:1101-1141) from the tile-major U_tot and Z: one
// atom per lane (64-atom tiles, so the U_tot loads are coalesced), the
// components b spread over the wavefronts of a tile (grid.y); loop bounds
// are wave-uniform.
template <int TWOJ>
__global__ __launch_bounds__(256) void so4_b_kernel(int64_t natoms, const double *__restrict__ utile,
                                                     const double *__restrict__ z, int nz, int nb,
                                                     const int *__restrict__ b_zoff, const int *__restrict__ b_level,
                                                     double *__restrict__ bout) {
    constexpr int NH = nhalf_of(TWOJ);
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.y * 4 + (threadIdx.x >> 6), nw = gridDim.y * 4;
    const int64_t tile = blockIdx.x;
    const int64_t i = tile * 64 + lane;
    if (i >= natoms) return;
    const cd *ut = reinterpret_cast<const cd *>(utile) + (size_t)tile * NH * 64 + lane;
    const cd *zi = reinterpret_cast<const cd *>(z) + (size_t)tile * nz * 64 + lane;  // Z tile-major
    for (int b = w; b < nb; b += nw) {
        const int l = __builtin_amdgcn_readfirstlane(b_level[b]);
        const int z0 = __builtin_amdgcn_readfirstlane(b_zoff[b]);
        const int h0 = hoff_of(l);
        // rows mb < l/2 in full, then (l even) the middle row up to the
        // diagonal, whose element counts half
        const int nfull = ((l + 1) / 2) * (l + 1);
        double s = 0.0;
#pragma unroll 8
        for (int q = 0; q < nfull; q++) s += cdotre(ut[(size_t)(h0 + q) * 64], zi[(size_t)(z0 + q) * 64]);
        if ((l & 1) == 0) {
            const int mb = l / 2;
#pragma unroll 4
            for (int q = nfull; q < nfull + mb; q++)
                s += cdotre(ut[(size_t)(h0 + q) * 64], zi[(size_t)(z0 + q) * 64]);
            s += 0.5 * cdotre(ut[(size_t)(h0 + nfull + mb) * 64], zi[(size_t)(z0 + nfull + mb) * 64]);
        }
        bout[(size_t)i * nb + b] = 2.0 * s;
    }
}

// ---------------------------------------------------------------------- Y
// Y[h] = 2 hw(h) norm_l sum_{blocks k at level l} c_k Z[k][loc(h)],
// c_k = sum_targets f beta[b].  One atom per lane over a 64-atom tile, like
// the Z kernel whose tile-major output it reads: a wavefront owns one task
// (a run of at most YTASK_ROWS half elements of one level), loops over the
// level's blocks in order (fixed summation order, no atomics) and adds
// c_k Z[k][loc] into register accumulators; every Z load is a coalesced
// 1 KB row and every Z element is read by exactly one wavefront.  The four
// wavefronts of a workgroup take four tasks of the same tile and share its
// dE/dB rows, staged in LDS.
__global__ __launch_bounds__(256) void so4_yi_kernel(int64_t natoms, int ntasks, int nz, int nb, int nhalf,
                                                      const int *__restrict__ ytask, const double *__restrict__ z,
                                                      const double *__restrict__ beta,
                                                      const int *__restrict__ blk_zoff, const int *__restrict__ blk_tb,
                                                      const double *__restrict__ blk_tf,
                                                      const double *__restrict__ hscale, double *__restrict__ y) {
    extern __shared__ double sb[];  // [64][nb] dE/dB of the tile (zero beyond natoms)
    const int tg = (ntasks + 3) / 4;
    const int64_t tile = blockIdx.x / tg;
    const int task = (int)(blockIdx.x % tg) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t a0 = tile * 64;
    const int64_t nv = (natoms - a0 < 64 ? natoms - a0 : 64) * nb;
    for (int64_t k = threadIdx.x; k < 64 * (int64_t)nb; k += blockDim.x) sb[k] = k < nv ? beta[a0 * nb + k] : 0.0;
    __syncthreads();
    if (task >= ntasks) return;
    const int *T = ytask + YTASK_STRIDE * task;
    const int loc0 = __builtin_amdgcn_readfirstlane(T[1]);
    const int cnt = __builtin_amdgcn_readfirstlane(T[2]);
    const int k0 = __builtin_amdgcn_readfirstlane(T[3]);
    const int k1 = __builtin_amdgcn_readfirstlane(T[4]);
    const int h0 = __builtin_amdgcn_readfirstlane(T[5]);
    const double *bl = sb + lane * nb;
    const cd *zt = reinterpret_cast<const cd *>(z) + (size_t)tile * nz * 64 + lane;
    cd acc[YTASK_ROWS];
#pragma unroll
    for (int q = 0; q < YTASK_ROWS; q++) acc[q] = cmake(0.0, 0.0);
    // YB blocks per step, so that their Z loads are in flight together
    constexpr int YB = 4;
    for (int kb = k0; kb < k1; kb += YB) {
        double c[YB];
        cd v[YB][YTASK_ROWS];
#pragma unroll
        for (int u = 0; u < YB; u++) {
            const int k = kb + u;
            c[u] = 0.0;
            if (k < k1) {
                const cd *zk = zt + (size_t)(__builtin_amdgcn_readfirstlane(blk_zoff[k]) + loc0) * 64;
#pragma unroll
                for (int q = 0; q < YTASK_ROWS; q++)
                    if (q < cnt) v[u][q] = zk[q * 64];
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    const int b = __builtin_amdgcn_readfirstlane(blk_tb[3 * k + t]);
                    if (b >= 0) c[u] += blk_tf[3 * k + t] * bl[b];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < YB; u++)
            if (kb + u < k1) {
#pragma unroll
                for (int q = 0; q < YTASK_ROWS; q++)
                    if (q < cnt) {
                        acc[q].re += c[u] * v[u][q].re;
                        acc[q].im += c[u] * v[u][q].im;
                    }
            }
    }
    if (a0 + lane < natoms) {
        cd *yo = reinterpret_cast<cd *>(y) + (size_t)(a0 + lane) * nhalf + h0;
#pragma unroll
        for (int q = 0; q < YTASK_ROWS; q++)
            if (q < cnt) yo[q] = cscale(acc[q], hscale[h0 + q]);
    }
}

// -------------------------------------------------------- adjoint dE/dr
constexpr int DEIDRJ_NST = 16;  // staged doubles per image of the dE/dr kernel
// Flat over images: a wavefront serves NB = nr x NG images (any centres).
// It first stages, one lane per image, the Cayley-Klein parameters and the
// per-image factors of the generator contraction in LDS (the G lanes of an
// image would otherwise each compute them), then sweeps them in nr rounds
// of NG images (nr is chosen by the launch so that the grid holds enough
// wavefronts to hide latency):
//   dE/dr_k = w fc (p_k A1 + Re q_k A2 + Im q_k A3) + w dfc rhat_k A_u.
template <int TWOJ>
__global__ __launch_bounds__(256) void so4_deidrj_kernel(int64_t nimg, const int64_t *__restrict__ img_atom,
                                                           const int64_t *__restrict__ img_j,
                                                           const double *__restrict__ img_r,
                                                           const double *__restrict__ atom_w,
                                                           const double *__restrict__ y, const double *rootpq_g,
                                                           double rcut, double rfac0, int cutoff_id,
                                                           double *__restrict__ dedr,
                                                           const int64_t *__restrict__ d_count, int nr) {
    constexpr int G = TWOJ / 2 + 1;
    constexpr int NG = 64 / G;
    const int NB = NG * nr;  // images per wavefront (nr rounds, nr <= 64 / NG)
    constexpr int NST = DEIDRJ_NST;
    constexpr int NH = nhalf_of(TWOJ);
    __shared__ double rpq[(TWOJ + 1) * (TWOJ + 1)];
    // per wavefront (dynamic, sized by the launch's wavefronts per block):
    // NST x 64 staged doubles, 64 x 4 partial sums, 64 centre indices
    extern __shared__ double dsm[];
    // with d_count the grid covers an image capacity and the images in use
    // are counted on the device
    if (d_count) {
        const int64_t c = *d_count;
        nimg = c < nimg ? c : nimg;
    }
    const int tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63;
    const int wpb = blockDim.x >> 6;
    const int g = lane / G, mb = lane % G;
    for (int k = tid; k < (TWOJ + 1) * (TWOJ + 1); k += blockDim.x) rpq[k] = rootpq_g[k];
    __syncthreads();
    const int64_t base = ((int64_t)blockIdx.x * wpb + wv) * NB;
    // a wavefront past the images in use (a launch over an image capacity)
    // has nothing to do; the later synchronisation is wavefront-local
    if (base >= nimg) return;
    double(*S)[64] = reinterpret_cast<double(*)[64]>(dsm + (size_t)wv * (NST + 5) * 64);
    double(*red)[4] = reinterpret_cast<double(*)[4]>(dsm + (size_t)wv * (NST + 5) * 64 + NST * 64);
    int64_t *s_at = reinterpret_cast<int64_t *>(dsm + (size_t)wv * (NST + 5) * 64 + (NST + 4) * 64);
    {
        const int64_t img = base + lane;
        bool ok = lane < NB && img < nimg;
        double x = 0, yy = 0, zz = 1.0, r = 1.0, w = 0.0;
        int64_t i = -1;
        if (ok) {
            x = img_r[3 * img];
            yy = img_r[3 * img + 1];
            zz = img_r[3 * img + 2];
            r = sqrt(x * x + yy * yy + zz * zz);
            w = atom_w[img_j[img]];
            i = img_atom[img];
            if (r < 1e-8) ok = false;  // SNAP.py:1487 (padding rows)
        }
        if (!ok) {
            x = 0;
            yy = 0;
            zz = 1.0;
            r = 1.0;
        }
        double fc, dfc;
        cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
        CK ck;
        cayley_klein(x, yy, zz, r, rfac0 * M_PI * r / rcut, true, ck);
        double p[3];
        cd q[3];
        su2_generators(ck, p, q);
        S[0][lane] = ck.ca.re;
        S[1][lane] = ck.ca.im;
        S[2][lane] = ck.cb.re;
        S[3][lane] = ck.cb.im;
        const double wf = w * fc, wd = w * dfc / r;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            S[4 + k][lane] = wf * p[k];
            S[7 + k][lane] = wf * q[k].re;
            S[10 + k][lane] = wf * q[k].im;
        }
        S[13][lane] = wd * x;
        S[14][lane] = wd * yy;
        S[15][lane] = wd * zz;
        s_at[lane] = ok ? i : -1;
    }
    // the rounds read lanes of the same wavefront only
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int rd = 0; rd < nr; rd++) {
        if (base + rd * NG >= nimg) break;  // uniform
        const int slot = rd * NG + (g < NG ? g : 0);
        const int64_t i = s_at[slot];
        const bool valid = g < NG && i >= 0;
        CK ck;
        ck.ca = cmake(S[0][slot], S[1][slot]);
        ck.cb = cmake(S[2][slot], S[3][slot]);
        const cd *yi = reinterpret_cast<const cd *>(y) + (size_t)(valid ? i : 0) * NH;
        double au = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        sweep_rows<TWOJ, false>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
            if (!valid || !half) return;
            const cd *yr = yi + hoff_of(l) + mb * (l + 1);
            cd yv[TWOJ + 1];
#pragma unroll
            for (int m = 0; m <= TWOJ; m++)
                if (m <= l) yv[m] = yr[m];
            generator_dots<TWOJ>(l, u, yv, au, a1, a2, a3);
        });
        red[lane][0] = au;
        red[lane][1] = a1;
        red[lane][2] = a2;
        red[lane][3] = a3;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (valid && mb == 0) {
            double sm[4] = {0, 0, 0, 0};
            for (int qq = 0; qq < G; qq++)
                for (int k = 0; k < 4; k++) sm[k] += red[lane + qq][k];
            const int64_t img = base + slot;
            for (int k = 0; k < 3; k++)
                dedr[3 * img + k] = S[4 + k][slot] * sm[1] + S[7 + k][slot] * sm[2] + S[10 + k][slot] * sm[3] +
                                    S[13 + k][slot] * sm[0];
        } else if (g < NG && mb == 0 && base + slot < nimg) {
            // an image of coincident atoms (r < 1e-8) contributes nothing, as
            // the reference skips it (SNAP.py:1487)
            const int64_t img = base + slot;
            for (int k = 0; k < 3; k++) dedr[3 * img + k] = 0.0;
        }
        // the next round rewrites red
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
            // atom-major Z: each lane walks one contiguous block row
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

// --------------------------------------- plain dB/dr, one centre per workgroup
// The same dB/dr per image as so4_dbdr_kernel, organised around the centre:
// a workgroup (one wavefront) takes one centre and runs its images NG at a
// time.  With the su(2) generators of the image (see su2_generators), the
// derivative of a block product needs U alone:
//   Re <dU_tot/dr_k, Z_blk> = fc (p_k A1 + Re q_k A2 + Im q_k A3) + dfc rhat_k Au,
// A the generator sums of the image's U rows against the block's Z rows
// (generator_dots; the middle row of an even level weighted 1 / 0.5 / 0
// for ma < , = , > l/2).  Level by level the groups put their U half rows
// into LDS and the wavefront stages the centre's Z blocks of the level
// there (weights applied), read from the tile-major Z; then each lane sums
// one (image, block) pair over all the block's rows and adds
// 2 norm_l w f (.) to the block's (b, factor) targets (three terms of
// compute_dbidrj, SNAP.py:1143-1257) directly in the output rows: within a
// level each (image, b, k) receives at most one block, and the barriers
// between levels order the additions of different levels.  LDS holds only
// the level's U rows, its Z blocks and per-image factors; NG is chosen so
// that it stays within 20 KB where possible (two workgroups per SIMD).
// Workgroups are mapped to centres so that the eight centres whose Z
// entries share a 128-byte line run on one XCD (the Z reads of the
// tile-major layout then hit that XCD's L2).
template <int TWOJ>
__host__ __device__ constexpr int dbdr_centre_ng() {
    constexpr int G = TWOJ / 2 + 1;
    constexpr int NL = (TWOJ / 2 + 1) * (TWOJ + 1);
    constexpr int fixed = zlevel_max_of(TWOJ) * 16 + (TWOJ + 1) * (TWOJ + 1) * 8;
    constexpr int per = NL * 16 + 12 * 8;
    constexpr int fit = (20480 - fixed) / per;
    return fit >= 64 / G ? 64 / G : (fit >= 4 ? fit : 64 / G);
}

template <int TWOJ>
__global__ __launch_bounds__(64) void so4_dbdr_centre_kernel(
    int64_t natoms, const int64_t *__restrict__ img_start, const int64_t *__restrict__ img_j,
    const double *__restrict__ img_r, const double *__restrict__ atom_w, const double *__restrict__ zt, int nz,
    int nb, const double *rootpq_g, const double *__restrict__ lnorm, const int *__restrict__ blk_level_start,
    const int *__restrict__ blk_zoff, const int *__restrict__ blk_tb, const double *__restrict__ blk_tf, double rcut,
    double rfac0, int cutoff_id, int z_atom_major, double *__restrict__ dbdr) {
    constexpr int G = TWOJ / 2 + 1;
    constexpr int NG = dbdr_centre_ng<TWOJ>();
    constexpr int NL = (TWOJ / 2 + 1) * (TWOJ + 1);
    constexpr int ZLM = zlevel_max_of(TWOJ);
    // per image: w fc p_k, w fc Re q_k, w fc Im q_k, w dfc rhat_k (k = 0..2)
    constexpr int NSC = 12;
    __shared__ double rpq[(TWOJ + 1) * (TWOJ + 1)];
    __shared__ cd Ul[NG][NL];
    __shared__ cd Zl[ZLM];
    __shared__ double sc[NG][NSC];
    // centre of this workgroup: swap the 3-bit fields (b / 8) % 8 and b % 8
    // of the block index (the XCD is b % 8)
    const int64_t bid = blockIdx.x;
    const int64_t c = (bid & ~(int64_t)63) | ((bid & 7) << 3) | ((bid >> 3) & 7);
    if (c >= natoms) return;
    const int lane = threadIdx.x;
    const int g = lane / G, mb = lane % G;
    for (int k = lane; k < (TWOJ + 1) * (TWOJ + 1); k += 64) rpq[k] = rootpq_g[k];
    const int64_t s0 = img_start[c], s1 = img_start[c + 1];
    // the centre's Z entries: stride 64 in the tile-major layout, 1 in the
    // atom-major copy
    const cd *zc = reinterpret_cast<const cd *>(zt) + (z_atom_major ? (size_t)c * nz
                                                                      : (size_t)(c >> 6) * nz * 64 + (c & 63));
    const int zs = z_atom_major ? 1 : 64;
    for (int64_t e0 = s0; e0 < s1; e0 += NG) {
        const int ne = (int)(s1 - e0 < NG ? s1 - e0 : NG);
        // the chunk's output rows start at zero
        for (int k = lane; k < ne * nb * 3; k += 64) dbdr[(size_t)e0 * nb * 3 + k] = 0.0;
        const int64_t e = e0 + g;
        bool valid = g < ne;
        double x = 0, yy = 0, zz = 1.0, r = 1.0, w = 0.0;
        if (valid) {
            x = img_r[3 * e];
            yy = img_r[3 * e + 1];
            zz = img_r[3 * e + 2];
            r = sqrt(x * x + yy * yy + zz * zz);
            w = atom_w[img_j[e]];
            if (r < 1e-8) {  // coincident atoms contribute nothing (SNAP.py:1487)
                valid = false;
                x = 0;
                yy = 0;
                zz = 1.0;
                r = 1.0;
                w = 0.0;
            }
        }
        double fc, dfc;
        cutoff_fdf(cutoff_id, r, rcut, fc, dfc);
        CK ck;
        cayley_klein(x, yy, zz, r, rfac0 * M_PI * r / rcut, true, ck);
        if (g < NG && mb == 0) {
            double p[3];
            cd q[3];
            su2_generators(ck, p, q);
            double *S = sc[g];
            const double wf = w * fc, wd = w * dfc / r;
            S[0] = wf * p[0];
            S[1] = wf * p[1];
            S[2] = wf * p[2];
            S[3] = wf * q[0].re;
            S[4] = wf * q[1].re;
            S[5] = wf * q[2].re;
            S[6] = wf * q[0].im;
            S[7] = wf * q[1].im;
            S[8] = wf * q[2].im;
            S[9] = wd * x;
            S[10] = wd * yy;
            S[11] = wd * zz;
        }
        __syncthreads();
        sweep_rows<TWOJ, false>(mb, rpq, ck, [&](int l, const cd *u, const cd(*du)[TWOJ + 1], bool half) {
            (void)du;
            if (g < NG && half) {
#pragma unroll
                for (int m = 0; m <= TWOJ; m++)
                    if (m <= l) Ul[g][mb * (l + 1) + m] = u[m];
            }
            const int b0 = blk_level_start[l], nbl = blk_level_start[l + 1] - b0;
            const int nloc = (l / 2 + 1) * (l + 1);
            for (int it = lane; it < nbl * nloc; it += 64) {
                const int bk = it / nloc, loc = it - bk * nloc;
                const int row = loc / (l + 1), ma = loc - row * (l + 1);
                const double wt = 2 * row < l ? 1.0 : (ma < row ? 1.0 : (ma == row ? 0.5 : 0.0));
                const cd zv = zc[(size_t)(blk_zoff[b0 + bk] + loc) * zs];
                Zl[it] = cmake(wt * zv.re, wt * zv.im);
            }
            // also orders the previous level's output additions before this
            // level's
            __syncthreads();
            const double nrm2 = 2.0 * lnorm[l];
            const int nrow = l / 2 + 1;
            for (int it = lane; it < ne * nbl; it += 64) {
                const int bk = it / ne, gg = it - bk * ne;
                const int blk = b0 + bk;
                double *o = dbdr + (size_t)(e0 + gg) * nb * 3;
                int tb[3];
                double tf[3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    tb[q] = blk_tb[3 * blk + q];
                    tf[q] = blk_tf[3 * blk + q] * nrm2;
                }
                // the targets' current values, loaded ahead of the sums so
                // that their latency overlaps them (a block's targets are
                // distinct b, duplicates being merged in the table, so none
                // of these loads can see one of the stores below)
                double old[3][3];
#pragma unroll
                for (int q = 0; q < 3; q++)
#pragma unroll
                    for (int k = 0; k < 3; k++) old[q][k] = tb[q] >= 0 ? o[3 * tb[q] + k] : 0.0;
                double au = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
                for (int row = 0; row < nrow; row++)
                    generator_dots<TWOJ>(l, &Ul[gg][row * (l + 1)], &Zl[bk * nloc + row * (l + 1)], au, a1, a2, a3);
                const double *S = sc[gg];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const double pk = S[k] * a1 + S[3 + k] * a2 + S[6 + k] * a3 + S[9 + k] * au;
#pragma unroll
                    for (int q = 0; q < 3; q++)
                        if (tb[q] >= 0) o[3 * tb[q] + k] = fma(tf[q], pk, old[q][k]);
                }
            }
            __syncthreads();
        });
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

// Transpose of the plain force contraction: for a gradient g [N, 3] with
// respect to dE/dR, grad_beta[i, b] = sum over rows (i, m) of
// dxdr[row, b, :] . g[m, :] -- the backward of dxdr_contract_kernel (the
// reference's loss differentiates its dxdr contraction this way,
// neuralnetwork.py:484-497).  One wavefront per centre over its own rows
// (no reverse index), lanes over b, four rows in flight.
__global__ __launch_bounds__(64) void dxdr_contract_t_kernel(int64_t natoms, int nb,
                                                              const int64_t *__restrict__ row_start,
                                                              const int64_t *__restrict__ seq,
                                                              const double *__restrict__ g,
                                                              const double *__restrict__ dxdr,
                                                              double *__restrict__ gbeta) {
    const int64_t i = blockIdx.x;
    if (i >= natoms) return;
    const int lane = threadIdx.x;
    const int64_t r0 = row_start[i], r1 = row_start[i + 1];
    for (int b = lane; b < nb; b += 64) {
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        int64_t rw = r0;
        for (; rw + 4 <= r1; rw += 4) {
            const double *d = dxdr + (size_t)rw * nb * 3 + 3 * b;
            const int64_t m0 = seq[2 * rw + 1], m1 = seq[2 * rw + 3], m2 = seq[2 * rw + 5], m3 = seq[2 * rw + 7];
            const size_t st = (size_t)nb * 3;
            a0 += d[0] * g[3 * m0] + d[1] * g[3 * m0 + 1] + d[2] * g[3 * m0 + 2];
            a1 += d[st] * g[3 * m1] + d[st + 1] * g[3 * m1 + 1] + d[st + 2] * g[3 * m1 + 2];
            a2 += d[2 * st] * g[3 * m2] + d[2 * st + 1] * g[3 * m2 + 1] + d[2 * st + 2] * g[3 * m2 + 2];
            a3 += d[3 * st] * g[3 * m3] + d[3 * st + 1] * g[3 * m3 + 1] + d[3 * st + 2] * g[3 * m3 + 2];
        }
        for (; rw < r1; rw++) {
            const double *d = dxdr + (size_t)rw * nb * 3 + 3 * b;
            const int64_t m = seq[2 * rw + 1];
            a0 += d[0] * g[3 * m] + d[1] * g[3 * m + 1] + d[2] * g[3 * m + 2];
        }
        gbeta[(size_t)i * nb + b] = (a0 + a1) + (a2 + a3);
    }
}

// Adjoint path: dE/dR_i = sum_{e of i} (dE/dr_rev(e) - dE/dr_e); virial of
// centre i = sum_e r_e (x) dE/dr_e.  One wavefront per atom (4 per
// workgroup), lanes over the centre's images, then a butterfly reduction
// across lanes (fixed order: deterministic, no atomics).
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// NEG: the forces -dE/dR_i, summed as (dedr_e - dedr_rev(e)) -- the exact
// negation of the dE/dR sum, term by term in the same order.  With e_out the
// grid has one more workgroup, the last, which stores the sum of
// e_in[0, n_e) (the total energy) to e_out[0] by block_sum_1col: the same
// value as the one-column segment sum, with no launch of its own.
template <bool NEG>
__global__ __launch_bounds__(256) void so4_gather_kernel(int64_t natoms, const int64_t *__restrict__ img_start,
                                                          const int64_t *__restrict__ img_rev,
                                                          const double *__restrict__ img_r,
                                                          const double *__restrict__ dedr, double *__restrict__ grad,
                                                          double *__restrict__ virial,
                                                          const double *__restrict__ e_in = nullptr,
                                                          int64_t n_e = 0, double *__restrict__ e_out = nullptr) {
    static_assert(PXFF_SUM_THREADS == 256, "the energy workgroup is one gather workgroup");
    if (e_out != nullptr && blockIdx.x == gridDim.x - 1) {
        const double t = block_sum_1col(e_in, 0, n_e);
        if (threadIdx.x == 0) e_out[0] = t;
        return;
    }
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= natoms) return;  // whole wavefront
    double g[3] = {0, 0, 0};
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t e = img_start[i] + lane; e < img_start[i + 1]; e += 64) {
        const int64_t q = img_rev[e];
        const double d[3] = {dedr[3 * e], dedr[3 * e + 1], dedr[3 * e + 2]};
#pragma unroll
        for (int k = 0; k < 3; k++) g[k] += NEG ? d[k] - dedr[3 * q + k] : dedr[3 * q + k] - d[k];
        if (virial) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const double ra = img_r[3 * e + a];
#pragma unroll
                for (int c = 0; c < 3; c++) v[3 * a + c] += ra * d[c];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) g[k] = wave_sum(g[k]);
    if (virial) {
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = wave_sum(v[k]);
    }
    if (lane < 3) grad[3 * i + lane] = lane == 0 ? g[0] : (lane == 1 ? g[1] : g[2]);
    if (virial && lane == 0) {
#pragma unroll
        for (int k = 0; k < 9; k++) virial[9 * i + k] = v[k];
    }
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
// dB_b = sum over the (block k, factor f) pairs targeting b of f p_k with
// p_k = 2 Re sum_loc hw(loc) conj(dU^l[loc]) Z_k[loc] and dU = dU_tot (u_norm
// already applied by the accumulation kernel).  One workgroup of 16
// wavefronts per 64-atom tile, one atom per lane: dU and Z are both
// tile-major, so every load is a coalesced 1 KB row; the wavefronts split
// the blocks (p_k into LDS; four partial sums per block keep four loads and
// FMA chains in flight), then the components b.  The sixteen wavefronts of
// a tile keep enough of Z in flight that the kernel runs at the rate Z
// streams from memory.
constexpr int TANB_THREADS = 1024;
__global__ __launch_bounds__(TANB_THREADS) void so4_tangent_b_kernel(int64_t natoms, int nhalf, int nz, int nb, int nblk,
                                                             int ldim, const double *__restrict__ dutile,
                                                             const double *__restrict__ z,
                                                             const double *__restrict__ h_weight,
                                                             const int *__restrict__ hoff,
                                                             const int *__restrict__ blk_level_start,
                                                             const int *__restrict__ blk_zoff,
                                                             const int *__restrict__ b_src,
                                                             const double *__restrict__ b_srcf,
                                                             double *__restrict__ db) {
    extern __shared__ double pk[];  // [nblk][64]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t tile = blockIdx.x;
    const cd *dut = reinterpret_cast<const cd *>(dutile) + (size_t)tile * nhalf * 64 + lane;
    const cd *zt = reinterpret_cast<const cd *>(z) + (size_t)tile * nz * 64 + lane;
    const int nw = blockDim.x >> 6;
    for (int k = wv; k < nblk; k += nw) {
        int l = 0;
        while (l + 1 < ldim && k >= blk_level_start[l + 1]) l++;
        const int nloc = (l / 2 + 1) * (l + 1);
        const int z0 = blk_zoff[k], h0 = hoff[l];
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
        int t = 0;
        for (; t + 4 <= nloc; t += 4) {
            p0 += h_weight[h0 + t] * cdotre(dut[(size_t)(h0 + t) * 64], zt[(size_t)(z0 + t) * 64]);
            p1 += h_weight[h0 + t + 1] * cdotre(dut[(size_t)(h0 + t + 1) * 64], zt[(size_t)(z0 + t + 1) * 64]);
            p2 += h_weight[h0 + t + 2] * cdotre(dut[(size_t)(h0 + t + 2) * 64], zt[(size_t)(z0 + t + 2) * 64]);
            p3 += h_weight[h0 + t + 3] * cdotre(dut[(size_t)(h0 + t + 3) * 64], zt[(size_t)(z0 + t + 3) * 64]);
        }
        for (; t < nloc; t++)
            p0 += h_weight[h0 + t] * cdotre(dut[(size_t)(h0 + t) * 64], zt[(size_t)(z0 + t) * 64]);
        pk[k * 64 + lane] = 2.0 * ((p0 + p1) + (p2 + p3));
    }
    __syncthreads();
    const int64_t i = tile * 64 + lane;
    if (i >= natoms) return;
    for (int b = wv; b < nb; b += nw) {
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int k = b_src[3 * b + t];
            if (k >= 0) acc += b_srcf[3 * b + t] * pk[k * 64 + lane];
        }
        db[(size_t)i * nb + b] = acc;
    }
}

// Atom-major copy of the tile-major Z: za[i][e] = Z[i/64][e][i%64].  A
// workgroup moves 32 rows of one tile through LDS: coalesced 1 KB row loads,
// 512 B contiguous stores per atom.
__global__ __launch_bounds__(256) void so4_z_atom_major_kernel(int64_t natoms, int nz, const double *__restrict__ zt,
                                                                double *__restrict__ za) {
    __shared__ cd buf[32][65];
    const int64_t tile = blockIdx.x;
    const int e0 = blockIdx.y * 32;
    const int lane = threadIdx.x & 63;
    const cd *src = reinterpret_cast<const cd *>(zt) + (size_t)tile * nz * 64;
    for (int r = threadIdx.x >> 6; r < 32; r += 4)
        if (e0 + r < nz) buf[r][lane] = src[(size_t)(e0 + r) * 64 + lane];
    __syncthreads();
    cd *dst = reinterpret_cast<cd *>(za);
    for (int idx = threadIdx.x; idx < 64 * 32; idx += blockDim.x) {
        const int a = idx >> 5, r = idx & 31;
        const int64_t i = tile * 64 + a;
        if (i < natoms && e0 + r < nz) dst[(size_t)i * nz + e0 + r] = buf[r][a];
    }
}

}  // namespace pxff

// =================================================================== C ABI
using namespace pxff;

namespace pxff {
// Process-wide launch parameters set by the autotuner (pxff_tune_set).
// Atomic: a call reads each knob once when it launches, so calls from
// several host threads and streams never race with a concurrent set (a set
// affects the launches that start after it).
static std::atomic<int> g_tune[PXFF_TUNE_COUNT] = {16, 16, 4, 4};
static const int g_tune_lo[PXFF_TUNE_COUNT] = {1, 1, 1, 0};
static const int g_tune_hi[PXFF_TUNE_COUNT] = {256, 256, 4, 4};
int tune_value(int knob) { return g_tune[knob].load(std::memory_order_relaxed); }

int so4_check_params(const pxff_so4_params *p) {
    PXFF_REQUIRE(p != nullptr, "null params");
    if (p->twol < 0 || p->twol > 10 || (p->twol & 1)) {
        set_error("unsupported 2J=" + std::to_string(p->twol) + " (even 2J in 0..10 supported)");
        return PXFF_ENOTIMPL;
    }
    PXFF_REQUIRE(p->rcut > 0, "rcut must be > 0");
    PXFF_REQUIRE(p->cutoff_id >= 1 && p->cutoff_id <= 7, "cutoff id must be 1..7");
    return 0;
}

// queried once; concurrent first calls store the same value
static std::atomic<int> g_num_cus{0};
int num_cus() {
    int n = g_num_cus.load(std::memory_order_relaxed);
    if (n == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
        n = prop.multiProcessorCount;
        g_num_cus.store(n, std::memory_order_relaxed);
    }
    return n;
}
}  // namespace pxff

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
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int grid = (int)std::min<int64_t>(natoms, (int64_t)num_cus() * tune_value(PXFF_TUNE_ACCUM_WAVES_PER_CU));
    const double fcut0 = cutoff_f0_host(p->cutoff_id);
#define PXFF_ACCUM_UI(MODE)                                                                              \
    PXFF_DISPATCH_TWOJ(p->twol, (so4_accum_kernel<TJ, false, true, MODE><<<grid, 64, 0, s>>>(               \
                                    natoms, d_img_start, d_img_j, d_img_r, d_atom_w, d_site_w, nullptr,       \
                                    d_work_order, T->rootpq, T->lnorm, p->rcut, p->rfac0, p->cutoff_id, fcut0, \
                                    d_utot)))
    const int mode = tune_value(PXFF_TUNE_ACCUM_LDSADD);
    if (mode == 4 && p->twol <= 8) {
        PXFF_DISPATCH_TWOJ(p->twol, if constexpr (TJ <= 8) {
            so4_utot_kernel<TJ, false><<<grid, 64, 0, s>>>(natoms, d_img_start, d_img_j, d_img_r, d_atom_w, d_site_w,
                                                          nullptr, d_work_order, T->rootpq, T->lnorm, p->rcut,
                                                          p->rfac0, p->cutoff_id, fcut0, d_utot);
        });
    } else {
        switch (mode) {
        case 1: PXFF_ACCUM_UI(1); break;
        case 2: PXFF_ACCUM_UI(2); break;
        case 3: PXFF_ACCUM_UI(3); break;
        default: PXFF_ACCUM_UI(0); break;
        }
    }
#undef PXFF_ACCUM_UI
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_ui_staged(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                                  const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                                  const double *d_site_w, const int64_t *d_work_order, int64_t cap, double *d_stage,
                                  double *d_utot, void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    PXFF_REQUIRE(cap >= 0 && (cap == 0 || d_stage != nullptr), "stage workspace missing");
    if (natoms <= 0) return 0;
    // the register-sum kernel is the one with a staged form (2J <= 8);
    // the other summation modes stage in LDS themselves
    if (tune_value(PXFF_TUNE_ACCUM_LDSADD) != 4 || p->twol > 8)
        return pxff_so4_ui(p, natoms, d_img_start, d_img_j, d_img_r, d_atom_w, d_site_w, d_work_order, d_utot,
                           stream);
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int grid = (int)std::min<int64_t>(natoms, (int64_t)num_cus() * tune_value(PXFF_TUNE_ACCUM_WAVES_PER_CU));
    const double fcut0 = cutoff_f0_host(p->cutoff_id);
    const int sgrid = (int)std::max<int64_t>(1, std::min<int64_t>(div_up(cap, 256), (int64_t)num_cus() * 8));
    so4_stage_kernel<false><<<sgrid, 256, 0, s>>>(natoms, cap, d_img_start, d_img_j, d_img_r, d_atom_w, nullptr,
                                                  p->rcut, p->rfac0, p->cutoff_id, d_stage);
    PXFF_LAUNCH_CHECK();
    PXFF_DISPATCH_TWOJ(p->twol, if constexpr (TJ <= 8) {
        so4_utot_staged_kernel<TJ><<<grid, 64, 0, s>>>(natoms, d_img_start, d_site_w, d_work_order, T->lnorm, fcut0,
                                                      d_stage, d_utot);
    });
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_b(const pxff_so4_params *p, int64_t natoms, const double *d_utot, const double *d_z,
                          double *d_b, void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    PXFF_REQUIRE(d_b != nullptr, "null B output");
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)div_up(natoms, 64), (unsigned)div_up(T->nb, 16));
    PXFF_DISPATCH_TWOJ(p->twol, (so4_b_kernel<TJ><<<grid, 256, 0, s>>>(natoms, d_utot, d_z, T->nz, T->nb, T->b_zoff,
                                                                       T->b_level, d_b)));
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_yi(const pxff_so4_params *p, int64_t natoms, const double *d_z, const double *d_beta,
                           double *d_y, void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t lds = sizeof(double) * 64 * (size_t)T->nb;
    const int64_t ntiles = (natoms + 63) / 64;
    const int64_t grid = ntiles * ((T->nytask + 3) / 4);
    PXFF_REQUIRE(grid <= 0x7fffffff, "too many atoms for one launch");
    if (lds > 65536)
        PXFF_HIP_CHECK(hipFuncSetAttribute((const void *)so4_yi_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds));
    so4_yi_kernel<<<(unsigned)grid, 256, lds, s>>>(natoms, T->nytask, T->nz, T->nb, T->nhalf, T->ytask, d_z, d_beta,
                                                   T->blk_zoff, T->blk_tb, T->blk_tf, T->hscale, d_y);
    PXFF_LAUNCH_CHECK();
    return 0;
}

static int so4_deidrj_launch(const pxff_so4_params *p, int64_t nimg, const int64_t *d_count,
                             const int64_t *d_img_atom, const int64_t *d_img_j, const double *d_img_r,
                             const double *d_atom_w, const double *d_y, double *d_dedr, void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (nimg <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int ng = 64 / (p->twol / 2 + 1);
    // rounds per wavefront: at most 64 / ng (one staged lane per image), and
    // few enough that the grid has about 16 wavefronts per CU
    const int nr = (int)std::max<int64_t>(1, std::min<int64_t>(64 / ng, nimg / ((int64_t)num_cus() * 16 * ng)));
    const int nb = ng * nr;  // images per wavefront (staged batch)
    const int wpb = tune_value(PXFF_TUNE_IMAGE_WAVES_PER_BLOCK);
    const int grid = div_up(div_up(nimg, nb), wpb);
    const size_t lds = sizeof(double) * (size_t)wpb * (DEIDRJ_NST + 5) * 64;  // per-wavefront stage
    PXFF_DISPATCH_TWOJ(p->twol, (so4_deidrj_kernel<TJ><<<grid, 64 * wpb, lds, s>>>(nimg, d_img_atom, d_img_j, d_img_r, d_atom_w,
                                                                           d_y, T->rootpq, p->rcut, p->rfac0,
                                                                           p->cutoff_id, d_dedr, d_count, nr)));
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_deidrj(const pxff_so4_params *p, int64_t nimg, const int64_t *d_img_atom,
                               const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                               const double *d_y, double *d_dedr, void *stream) {
    return so4_deidrj_launch(p, nimg, nullptr, d_img_atom, d_img_j, d_img_r, d_atom_w, d_y, d_dedr, stream);
}

extern "C" int pxff_so4_deidrj_counted(const pxff_so4_params *p, int64_t cap, const int64_t *d_count,
                                       const int64_t *d_img_atom, const int64_t *d_img_j, const double *d_img_r,
                                       const double *d_atom_w, const double *d_y, double *d_dedr, void *stream) {
    PXFF_REQUIRE(d_count != nullptr, "null image count");
    return so4_deidrj_launch(p, cap, d_count, d_img_atom, d_img_j, d_img_r, d_atom_w, d_y, d_dedr, stream);
}

extern "C" int pxff_so4_gather(int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_rev,
                               const double *d_img_r, const double *d_dedr, double *d_grad, double *d_virial,
                               void *stream) {
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    so4_gather_kernel<false><<<div_up(natoms, 4), 256, 0, s>>>(natoms, d_img_start, d_img_rev, d_img_r, d_dedr,
                                                               d_grad, d_virial);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_forces(int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_rev,
                               const double *d_img_r, const double *d_dedr, double *d_forces, double *d_virial,
                               void *stream) {
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    so4_gather_kernel<true><<<div_up(natoms, 4), 256, 0, s>>>(natoms, d_img_start, d_img_rev, d_img_r, d_dedr,
                                                              d_forces, d_virial);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_forces_energy(int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_rev,
                                      const double *d_img_r, const double *d_dedr, double *d_forces, double *d_virial,
                                      const double *d_e_atom, int64_t n_e, double *d_energy, void *stream) {
    PXFF_REQUIRE(d_energy != nullptr, "null energy output");
    PXFF_REQUIRE(n_e >= 0 && (n_e == 0 || d_e_atom != nullptr), "bad atom energies");
    PXFF_REQUIRE(natoms >= 0, "natoms must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    // workgroups 0 .. div_up(natoms, 4) - 1 gather forces; the last one sums the energies
    so4_gather_kernel<true><<<div_up(natoms, 4) + 1, 256, 0, s>>>(natoms, d_img_start, d_img_rev, d_img_r, d_dedr,
                                                                  d_forces, d_virial, d_e_atom, n_e, d_energy);
    PXFF_LAUNCH_CHECK();
    return 0;
}

static int so4_tangent_launch(const pxff_so4_params *p, const So4Dev *T, int64_t natoms, const int64_t *d_img_start,
                              const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                              const double *d_dr, const double *d_z, double *d_dutot, double *d_db, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const int grid = (int)std::min<int64_t>(natoms, (int64_t)num_cus() * tune_value(PXFF_TUNE_ACCUM_WAVES_PER_CU));
#define PXFF_ACCUM_TAN(MODE)                                                                      \
    PXFF_DISPATCH_TWOJ(p->twol, (so4_accum_kernel<TJ, true, true, MODE><<<grid, 64, 0, s>>>(         \
                                    natoms, d_img_start, d_img_j, d_img_r, d_atom_w, nullptr, d_dr, nullptr, \
                                    T->rootpq, T->lnorm, p->rcut, p->rfac0, p->cutoff_id, 0.0, d_dutot)))
    // the tangent sums in registers whenever they fit (2J <= 8): 0.53 ms
    // against 0.75 ms for the LDS modes on the 33.7k-atom training batch,
    // whatever mode the forward's tuning chose
    const int mode = p->twol <= 8 ? 4 : tune_value(PXFF_TUNE_ACCUM_LDSADD);
    if (mode == 4 && p->twol <= 8) {
        PXFF_DISPATCH_TWOJ(p->twol, if constexpr (TJ <= 8) {
            so4_utot_kernel<TJ, true><<<grid, 64, 0, s>>>(natoms, d_img_start, d_img_j, d_img_r, d_atom_w, nullptr,
                                                         d_dr, nullptr, T->rootpq, T->lnorm, p->rcut, p->rfac0,
                                                         p->cutoff_id, 0.0, d_dutot);
        });
    } else {
        switch (mode) {
        case 1: PXFF_ACCUM_TAN(1); break;
        case 2: PXFF_ACCUM_TAN(2); break;
        case 3: PXFF_ACCUM_TAN(3); break;
        default: PXFF_ACCUM_TAN(0); break;
        }
    }
#undef PXFF_ACCUM_TAN
    PXFF_LAUNCH_CHECK();
    const size_t lds = sizeof(double) * 64 * (size_t)T->nblk;
    if (lds > 65536)
        PXFF_HIP_CHECK(hipFuncSetAttribute((const void *)so4_tangent_b_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    so4_tangent_b_kernel<<<(unsigned)div_up(natoms, 64), TANB_THREADS, lds, s>>>(
        natoms, T->nhalf, T->nz, T->nb, T->nblk, p->twol + 1, d_dutot, d_z, T->h_weight, T->hoff,
        T->blk_level_start, T->blk_zoff, T->b_src, T->b_srcf, d_db);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_tangent(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                                const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                                const double *d_dr, const double *d_z, double *d_dutot, double *d_db, void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    return so4_tangent_launch(p, T, natoms, d_img_start, d_img_j, d_img_r, d_atom_w, d_dr, d_z, d_dutot, d_db, stream);
}

// ------------------------------------------------- no-allocation tangent
// The dU_tot . dr workspace of pxff_so4_tangent in a file-scope device array
// (one copy per GPU, held until the process ends), for callers that cannot
// allocate device memory (HIP graph capture, tight loops).  Host code must
// not use the array's name as a device address: its address on the current
// device comes from hipGetSymbolAddress, resolved once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];
static std::mutex g_scratch_mutex;
static std::map<int, double *> g_scratch_ptr;  // device -> address of scratch there

static int scratch_on_device(double **out) {
    int dev = 0;
    PXFF_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        double *ptr = nullptr;
        PXFF_HIP_CHECK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
        PXFF_REQUIRE(ptr != nullptr, "no device address for the no-allocation scratch array");
        it = g_scratch_ptr.emplace(dev, ptr).first;
    }
    *out = it->second;
    return 0;
}

extern "C" int pxff_so4_tangent_noalloc(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                                        const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                                        const double *d_dr, const double *d_z, double *d_db, void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = find_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    const size_t need = (size_t)((natoms + 63) / 64) * 64 * (size_t)T->nhalf * 2 * sizeof(double);
    if (need > NOALLOC_SCRATCH_BYTES) {
        set_error("pxff_so4_tangent_noalloc: " + std::to_string(natoms) + " atoms at 2J=" + std::to_string(p->twol) +
                  " need " + std::to_string(need) + " bytes of workspace, more than the " +
                  std::to_string(NOALLOC_SCRATCH_BYTES) + " of the static array (use pxff_so4_tangent)");
        return PXFF_EINVAL;
    }
    double *du = nullptr;
    if ((rc = scratch_on_device(&du))) return rc;
    return so4_tangent_launch(p, T, natoms, d_img_start, d_img_j, d_img_r, d_atom_w, d_dr, d_z, du, d_db, stream);
}

extern "C" int pxff_so4_z_atom_major(const pxff_so4_params *p, int64_t natoms, const double *d_z, double *d_za,
                                     void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    const dim3 grid((unsigned)div_up(natoms, 64), (unsigned)div_up(T->nz, 32));
    so4_z_atom_major_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(natoms, T->nz, d_z, d_za);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so4_dbdr(const pxff_so4_params *p, int64_t nimg, const int64_t *d_img_atom,
                             const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w, const double *d_z,
                             double *d_dbdr, void *stream) {
    int rc = so4_check_params(p);
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

extern "C" int pxff_so4_dbdr_centres(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                                     const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w,
                                     const double *d_z, int z_atom_major, double *d_dbdr, void *stream) {
    int rc = so4_check_params(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const So4Dev *T = nullptr;
    if ((rc = get_so4_tables(p->twol, p->normalize_u, &T))) return rc;
    PXFF_REQUIRE(T->zlevel_max <= zlevel_max_of(p->twol) && T->nb <= nb_of(p->twol),
                 "Z level staging exceeds its LDS bound");
    PXFF_REQUIRE(natoms <= ((int64_t)1 << 31) - 64, "too many atoms for one launch");
    hipStream_t s = (hipStream_t)stream;
    // one workgroup per centre, rounded up to whole 64-centre tiles (the
    // centre map permutes within a tile; centres >= natoms exit)
    const unsigned grid = (unsigned)((natoms + 63) / 64 * 64);
    PXFF_DISPATCH_TWOJ(p->twol, (so4_dbdr_centre_kernel<TJ><<<grid, 64, 0, s>>>(
                                    natoms, d_img_start, d_img_j, d_img_r, d_atom_w, d_z, T->nz, T->nb, T->rootpq,
                                    T->lnorm, T->blk_level_start, T->blk_zoff, T->blk_tb, T->blk_tf, p->rcut,
                                    p->rfac0, p->cutoff_id, z_atom_major ? 1 : 0, d_dbdr)));
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

extern "C" int pxff_dxdr_contract_t(int64_t natoms, int nb, const int64_t *d_row_start, const int64_t *d_seq,
                                    const double *d_g, const double *d_dxdr, double *d_gbeta, void *stream) {
    if (natoms <= 0) return 0;
    PXFF_REQUIRE(nb > 0, "nb must be > 0");
    PXFF_REQUIRE(natoms <= 0x7fffffff, "too many atoms for one launch");
    dxdr_contract_t_kernel<<<(unsigned)natoms, 64, 0, (hipStream_t)stream>>>(natoms, nb, d_row_start, d_seq, d_g,
                                                                            d_dxdr, d_gbeta);
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
    g_tune[knob].store(value, std::memory_order_relaxed);
    return 0;
}

extern "C" int pxff_tune_get(int knob) {
    if (knob < 0 || knob >= PXFF_TUNE_COUNT) return -1;
    return tune_value(knob);
}
