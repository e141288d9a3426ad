// SO(3) power spectrum kernels for gfx950 (FP64, wave64).
//
// Reference: pyxtal_ff/descriptors/SO3.py:187-291 (calculate) and
// SO3.py:466-694 (W, g, Gauss-Chebyshev quadrature, compute_cs/compute_dcs).
// Per neighbour image e with separation r (|r| = R) and weight w_e (Z_j,
// sign-flipped when weight_on and Z_j != Z_i), for l <= lmax, n < nmax:
//   I_nl(R)  = sum_q G[n][q] i_l(2 alpha R r_q)
//   dI_nl(R) = sum_q G[n][q] 2 alpha r_q i_l'(2 alpha R r_q)
//   c_nlm    = norm_l w_e 4 pi exp(-alpha R^2) fc(R) I_nl(R) Y_lm(rhat)
// with G[n][q] = g_n(r_q) (pi/NQ)(rc/2) r_q^2 exp(-alpha r_q^2) sqrt(1-x_q^2)
// precomputed on the host, Y_lm with the Condon-Shortley phase (scipy's
// sph_harm, not conjugated) and grad Y_lm from the covariant components of
// SO3.py:649-674.
//
// Work mapping: a group of L = lmax+1 lanes serves one image, lane l owns
// angular order l: it evaluates i_l and i_l' at the NQ quadrature nodes
// (positive power series below a threshold, upward recurrence above it,
// both stable there), contracts them with G into NMAX registers, and reads
// the group's Y table from LDS.  Per-centre sums (ctot, its tangent) use
// per-group LDS partials reduced in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <set>
#include <vector>

#include "pxff_common.h"

namespace pxff {

constexpr int SO3_LMAXC = 32;  // SO3.py:95-101 allows lmax <= 32
constexpr int SO3_KMAX = 400;  // series terms

// Reciprocals of the series: g_so3_rk[n][k] = 1 / (k (2n + 2k + 1)) and
// g_so3_rodd[k] = 1 / (2k + 1), filled once per device (so3_series_init),
// so the series below multiplies instead of dividing.
__device__ double g_so3_rk[(SO3_LMAXC + 2) * SO3_KMAX];
__device__ double g_so3_rodd[SO3_LMAXC + 3];

// Coefficients of the normalised associated Legendre recurrence (same
// expressions as evaluated in place before; filled once per device with
// the series tables): for k >= 2, m <= k - 2
//   g_so3_lega[k][m] = sqrt((4k^2 - 1) / (k^2 - m^2)),
//   g_so3_legb[k][m] = sqrt(((k-1)^2 - m^2) / (4 (k-1)^2 - 1)),
// and per m: g_so3_pmm[m] = -sqrt((2m + 1) / (2m)) (m >= 1),
// g_so3_p1[m] = sqrt(2m + 3).
constexpr int SO3_LEGD = SO3_LMAXC + 2;
__device__ double g_so3_lega[SO3_LEGD * SO3_LEGD];
__device__ double g_so3_legb[SO3_LEGD * SO3_LEGD];
__device__ double g_so3_pmm[SO3_LEGD];
__device__ double g_so3_p1[SO3_LEGD];

// i_l(x) and i_{l-1}(x) (i_1 when l == 0: needed for i_0' = i_1).
__device__ inline void sph_bessel_i_pair(int l, double x, double &il, double &ilm1) {
    const double thr = 16.0 + 0.5 * l * (l + 1);
    if (x < thr) {
        // i_n(x) = x^n / (2n+1)!! sum_k (x^2/2)^k / (k! (2n+3)(2n+5)...(2n+2k+1))
        // both series in one loop (two independent product chains), four
        // terms per convergence test; stopping when both last terms are
        // below 1e-17 of their sums
        const int n2 = (l == 0) ? 1 : l - 1;
        const double h = 0.5 * x * x;
        double pre1 = 1.0, pre2 = 1.0;
        for (int k = 1; k <= l; k++) pre1 *= x * g_so3_rodd[k];
        for (int k = 1; k <= n2; k++) pre2 *= x * g_so3_rodd[k];
        const double *rk1 = g_so3_rk + l * SO3_KMAX, *rk2 = g_so3_rk + n2 * SO3_KMAX;
        double t1 = 1.0, s1 = 1.0, t2 = 1.0, s2 = 1.0;
        for (int k = 1; k + 3 < SO3_KMAX; k += 4) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                t1 *= h * rk1[k + u];
                s1 += t1;
                t2 *= h * rk2[k + u];
                s2 += t2;
            }
            if (t1 < 1e-17 * s1 && t2 < 1e-17 * s2) break;
        }
        il = pre1 * s1;
        ilm1 = pre2 * s2;
        return;
    }
    const double e = exp(x), ei = 1.0 / e;
    const double sh = 0.5 * (e - ei), ch = 0.5 * (e + ei);
    double i0 = sh / x;
    double i1 = (x * ch - sh) / (x * x);
    if (l == 0) {
        il = i0;
        ilm1 = i1;
        return;
    }
    double a = i0, b = i1;
    for (int k = 1; k < l; k++) {
        double c = a - (2 * k + 1) / x * b;
        a = b;
        b = c;
    }
    il = b;
    ilm1 = a;
}

__device__ inline void so3_cutoff(int id, double r, double rc, double &f, double &df) {
    if (id == 8) {  // 'unity': the reference returns ones for the value and the derivative (SO3.py:459-464)
        f = 1.0;
        df = 1.0;
        return;
    }
    double x = r / rc;
    switch (id) {
    case 1: {
        double s, c;
        sincos(M_PI * r / rc, &s, &c);
        f = 0.5 * (c + 1.0);
        df = -0.5 * M_PI / rc * s;
        return;
    }
    case 2: {
        double t = tanh(1.0 - x), t2 = t * t;
        f = t2 * t;
        df = -(3.0 / rc) * t2 * (1.0 - t2);
        return;
    }
    case 3:
        f = x * x * (2 * x - 3) + 1;
        df = (6.0 / (rc * rc)) * r * (x - 1.0);
        return;
    case 4:
        f = x * x * x * (x * (15 - 6 * x) - 10) + 1;
        df = (-30.0 / rc) * (x * x * (x - 1) * (x - 1));
        return;
    case 5:
        f = x * x * x * x * (x * (x * (20 * x - 70) + 84) - 35) + 1;
        df = (140.0 / rc) * (x * x * x * (x - 1) * (x - 1) * (x - 1));
        return;
    case 6: {
        f = x * x * x * x * x * (x * (x * (x * (315 - 70 * x) - 540) + 420) - 126) + 1;
        double y = x * (x - 1);
        df = (-630.0 / rc) * y * y * y * y;
        return;
    }
    case 7: {
        double e = exp(1.0 - 1.0 / (1.0 - x * x)), d = 1 + x * x;
        f = e;
        df = 2 * x * e / (d * d);
        return;
    }
    default:
        f = 0.0;
        df = 0.0;
    }
}

// Y_l^m for l <= L (one value of l per call), m = -l..l, written to
// Y[l*(2*LP+1) + LP + m] (LP = row stride half-width).  Normalised associated
// Legendre recurrence with the Condon-Shortley phase.
__device__ inline void ylm_row(int l, double t, double s, cd eiphi, cd *Y, int LP) {
    const double inv4pi = 0.25 / M_PI;
    cd em = cmake(1.0, 0.0);  // e^{i m phi}
    double pmm = sqrt(inv4pi);  // Pbar_m^m, m = 0
    for (int m = 0; m <= l; m++) {
        if (m > 0) {
            pmm = g_so3_pmm[m] * s * pmm;
            em = cmul(em, eiphi);
        }
        double p;
        if (l == m) {
            p = pmm;
        } else {
            double pm2 = pmm;
            double pm1 = g_so3_p1[m] * t * pmm;
            for (int k = m + 2; k <= l; k++) {
                const double a = g_so3_lega[k * SO3_LEGD + m];
                const double b = g_so3_legb[k * SO3_LEGD + m];
                double pk = a * (t * pm1 - b * pm2);
                pm2 = pm1;
                pm1 = pk;
            }
            p = pm1;
        }
        cd y = cscale(em, p);
        Y[l * (2 * LP + 1) + LP + m] = y;
        if (m > 0) {
            // Y_l^{-m} = (-1)^m conj(Y_l^m)
            double sg = (m & 1) ? -1.0 : 1.0;
            Y[l * (2 * LP + 1) + LP - m] = cmake(sg * y.re, -sg * y.im);
        }
    }
}

struct So3Args {
    int nmax, lmax, nq, cutoff_id, weight_on;
    double rc, alpha;
    const double *G;      // [nmax][nq]
    const double *rq;     // [nq]
    const double *lnorm;  // [lmax+1]
};

// Radial part of lane (image, l): I[n], dI[n] for n < NMAX.
template <int NMAX>
__device__ inline void radial(const So3Args &a, int l, double R, const double *__restrict__ Gs,
                              const double *__restrict__ rqs, double *I,
                              double *dI, bool want_d) {
#pragma unroll
    for (int n = 0; n < NMAX; n++) {
        I[n] = 0.0;
        dI[n] = 0.0;
    }
    const double ta = 2.0 * a.alpha;
    for (int q = 0; q < a.nq; q++) {
        const double rq = rqs[q];
        const double x = ta * R * rq;
        double il, ilm1;
        sph_bessel_i_pair(l, x, il, ilm1);
#pragma unroll
        for (int n = 0; n < NMAX; n++) I[n] += Gs[n * a.nq + q] * il;
        if (want_d) {
            // scipy spherical_in(derivative=True): i_0' = i_1; i_l' = i_{l-1} - (l+1)/x i_l
            const double dil = (l == 0) ? ilm1 : ilm1 - (l + 1) / x * il;
            const double f = ta * rq * dil;
#pragma unroll
            for (int n = 0; n < NMAX; n++) dI[n] += Gs[n * a.nq + q] * f;
        }
    }
}

// Covariant gradient of Y_l^m (SO3.py:649-674), from the group's Y table
// (rows 0..lmax+1, stride 2*LP+1).
__device__ inline void grad_ylm(int l, int m, double R, const cd *Y, int LP, cd g[3]) {
    if (l == 0) {
        g[0] = g[1] = g[2] = cmake(0, 0);
        return;
    }
    auto y = [&](int ll, int mm) -> cd { return Y[ll * (2 * LP + 1) + LP + mm]; };
    const double dl = l, dm = m;
    cd x0 = cscale(y(l + 1, m), -sqrt(((dl + 1) * (dl + 1) - dm * dm) / (2 * dl + 1) / (2 * dl + 3)) * dl / R);
    if (abs(m) <= l - 1)
        x0 = cadd(x0, cscale(y(l - 1, m), sqrt((dl * dl - dm * dm) / (2 * dl - 1) / (2 * dl + 1)) * (dl + 1) / R));
    cd xp = cscale(y(l + 1, m + 1), -sqrt((dl + dm + 1) * (dl + dm + 2) / 2 / (2 * dl + 1) / (2 * dl + 3)) * dl / R);
    if (abs(m + 1) <= l - 1)
        xp = csub(xp, cscale(y(l - 1, m + 1), sqrt((dl - dm - 1) * (dl - dm) / 2 / (2 * dl - 1) / (2 * dl + 1)) * (dl + 1) / R));
    cd xm = cscale(y(l + 1, m - 1), -sqrt((dl - dm + 1) * (dl - dm + 2) / 2 / (2 * dl + 1) / (2 * dl + 3)) * dl / R);
    if (abs(m - 1) <= l - 1)
        xm = csub(xm, cscale(y(l - 1, m - 1), sqrt((dl + dm - 1) * (dl + dm) / 2 / (2 * dl - 1) / (2 * dl + 1)) * (dl + 1) / R));
    const double r2 = 1.0 / sqrt(2.0);
    g[0] = cscale(csub(xm, xp), r2);
    cd s = cadd(xm, xp);
    g[1] = cmake(-s.im * r2, s.re * r2);  // i/sqrt2 * (xm + xp)
    g[2] = x0;
}

enum { SO3_CTOT = 0, SO3_TANGENT = 1, SO3_ADJOINT = 2, SO3_PLAIN = 3 };

// One image's per-lane context.
struct So3Img {
    bool valid;
    int64_t e, i;
    double x, y, z, R, w, fc, dfc, ex, rh[3];
};

__device__ inline void load_image(const So3Args &a, bool valid, int64_t e, int64_t centre, const int64_t *img_j,
                                  const double *img_r, const int64_t *atom_z, So3Img &m) {
    m.valid = valid;
    m.e = e;
    m.i = centre;
    m.x = 0;
    m.y = 0;
    m.z = 1;
    m.R = 1;
    m.w = 0;
    if (m.valid) {
        m.x = img_r[3 * e];
        m.y = img_r[3 * e + 1];
        m.z = img_r[3 * e + 2];
        m.R = sqrt(m.x * m.x + m.y * m.y + m.z * m.z);
        const int64_t zj = atom_z[img_j[e]], zi = atom_z[m.i];
        m.w = (a.weight_on && zj != zi) ? -(double)zj : (double)zj;
    }
    so3_cutoff(a.cutoff_id, m.R, a.rc, m.fc, m.dfc);
    m.ex = 4.0 * M_PI * exp(-a.alpha * m.R * m.R);
    m.rh[0] = m.x / m.R;
    m.rh[1] = m.y / m.R;
    m.rh[2] = m.z / m.R;
}

// Fill the Y table of one image: rows l' = 0..lmax+1 spread over the L lanes
// of the group (lane l also takes row l + L when it exists).
__device__ inline void fill_ylm(const So3Img &m, int l, int L, int lmax, cd *Y, int LP) {
    const double t = m.z / m.R;
    const double rho = sqrt(m.x * m.x + m.y * m.y);
    const double s = rho / m.R;
    const double ph = atan2(m.y, m.x);
    double sp, cp;
    sincos(ph, &sp, &cp);
    const cd eiphi = cmake(cp, sp);
    for (int ll = l; ll <= lmax + 1; ll += L) ylm_row(ll, t, s, eiphi, Y, LP);
}

// Per-image factors of one lane (image, l): with
//   P_m  = w norm_l ex fc Y_lm
//   U_km = w norm_l (r_k ex (dfc - 2 alpha R fc) Y_lm + ex fc dY_lm/dr_k)
//   V_km = w norm_l ex fc r_k Y_lm
// one has c_nlm = I_n P_m and dc_nlm/dr_k = I_n U_km + dI_n V_km
// (SO3.py:676-694: gradient of exp, of the radial integral, of Y_lm, of fc).
__device__ inline void image_factors(const So3Args &a, const So3Img &m, int l, int mm, const cd *Y, int LP, cd &P,
                                     cd U[3], cd V[3]) {
    const double wn = m.w * a.lnorm[l];
    const cd y = Y[l * (2 * LP + 1) + LP + mm];
    cd gy[3];
    grad_ylm(l, mm, m.R, Y, LP, gy);
    P = cscale(y, wn * m.ex * m.fc);
    const double cu = wn * m.ex * (m.dfc - 2.0 * a.alpha * m.R * m.fc);
    const double cv = wn * m.ex * m.fc;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        U[k] = cadd(cscale(y, cu * m.rh[k]), cscale(gy[k], cv));
        V[k] = cscale(y, cv * m.rh[k]);
    }
}

// ------------------------------------------------------------ radial
// rad[e][l][0..NMAX) = I_nl(R_e), rad[e][l][NMAX..2NMAX) = dI_nl/dR: one
// thread per (image, l), no shared memory (the quadrature tables are read
// through wave-uniform addresses), so the quadrature -- the bulk of the
// expansion's work -- runs at full occupancy.
// Grid-stride over (image, l); the image count is read from the list
// (img_start[natoms]), so the launch needs no host-side count.
template <int NMAX>
__global__ __launch_bounds__(256) void so3_radial_kernel(So3Args a, int64_t natoms,
                                                         const int64_t *__restrict__ img_start,
                                                         const double *__restrict__ img_r, double *__restrict__ rad) {
    const int L = a.lmax + 1;
    const int64_t total = img_start[natoms] * L;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t / L;
        const int l = (int)(t - e * L);
        const double x = img_r[3 * e], y = img_r[3 * e + 1], z = img_r[3 * e + 2];
        const double R = sqrt(x * x + y * y + z * z);
        double I[NMAX], dI[NMAX];
        radial<NMAX>(a, l, R, a.G, a.rq, I, dI, true);
        double *ro = rad + (size_t)t * 2 * NMAX;
#pragma unroll
        for (int n = 0; n < NMAX; n++) {
            ro[n] = I[n];
            ro[NMAX + n] = dI[n];
        }
    }
}

// i_0(x) .. i_{LC-1}(x) at once (LC <= 8 orders), for the radial kernel
// below: the positive series for x < 0.1 (h = x^2/2 < 0.005: eight terms
// leave < 1e-22 of the sum), Miller's downward recurrence below XUP
// (f_{N+1} = 0, f_N = 1, f_{k-1} = f_{k+1} + (2k+1)/x f_k from
// N = LC + 12 + x, normalised by sum_k (2k+1) i_k(x) = e^x, the x = 1 case
// of the generating function e^{x cos t} = sum_k (2k+1) i_k(x) P_k(cos t)),
// and the closed forms of i_0, i_1 with the upward recurrence from XUP on,
// where it is stable (XUP = 10 for orders <= 6, 20 above).  Checked against
// scipy.special.spherical_in over x in [1e-3, 100]: <= 1e-14 relative.
// ix = 1/x, which the caller computes once and reuses.
template <int LC>
__device__ inline void sph_bessel_i_all(double x, double ix, double *f) {
    constexpr double XUP = LC <= 7 ? 10.0 : 20.0;
    if (x < 0.1) {
        const double h = 0.5 * x * x;
        double pre = 1.0;
#pragma unroll
        for (int l = 0; l < LC; l++) {
            if (l > 0) pre *= x * g_so3_rodd[l];
            const double *rk = g_so3_rk + l * SO3_KMAX;
            double t = 1.0, sm = 1.0;
#pragma unroll
            for (int k = 1; k <= 8; k++) {
                t *= h * rk[k];
                sm += t;
            }
            f[l] = pre * sm;
        }
        return;
    }
    if (x < XUP) {
        const int N = LC + 12 + (int)x;
        double fkp1 = 0.0, fk = 1.0, sum = 0.0;
        int k = N;
        // two steps per trip (the same operations in the same order as one
        // step per trip, without the register rotation)
        for (; k - 1 >= LC; k -= 2) {
            sum += (2 * k + 1) * fk;
            const double fkm1 = fkp1 + (2 * k + 1) * ix * fk;
            sum += (2 * k - 1) * fkm1;
            const double fkm2 = fk + (2 * k - 1) * ix * fkm1;
            fkp1 = fkm1;
            fk = fkm2;
        }
        if (k >= LC) {
            sum += (2 * k + 1) * fk;
            const double fkm1 = fkp1 + (2 * k + 1) * ix * fk;
            fkp1 = fk;
            fk = fkm1;
        }
#pragma unroll
        for (int k = LC - 1; k >= 1; k--) {
            f[k] = fk;
            sum += (2 * k + 1) * fk;
            const double fkm1 = fkp1 + (2 * k + 1) * ix * fk;
            fkp1 = fk;
            fk = fkm1;
        }
        f[0] = fk;
        sum += fk;
        const double sc = exp(x) / sum;
#pragma unroll
        for (int k = 0; k < LC; k++) f[k] *= sc;
        return;
    }
    const double e = exp(x), ei = 1.0 / e;
    const double sh = 0.5 * (e - ei), ch = 0.5 * (e + ei);
    f[0] = sh * ix;
    if (LC > 1) f[1] = (x * ch - sh) * ix * ix;
#pragma unroll
    for (int k = 1; k + 1 < LC; k++) f[k + 1] = f[k - 1] - (2 * k + 1) * ix * f[k];
}

// Radial integrals of every image, one image per lane, all orders l <= lmax
// (lmax + 1 <= LC) at once: per quadrature node one evaluation of
// i_0..i_{LC-1} feeds the I / dI sums of every (l, n), accumulated in node
// order as in radial().  Same output as so3_radial_kernel.  With img_rev
// (the separations being the list's own, so an image and its reverse have
// bitwise the same |r|), only images with rev(e) >= e are evaluated and
// their integrals are written to both rows: each wavefront queues the
// images it owns that qualify in LDS and runs them 64 at a time.
// I_nl(R), dI_nl/dR for l < LC, n < NMAX by the quadrature (one evaluation
// of i_0..i_{LC-1} per node).
template <int NMAX, int LC>
__device__ inline void so3_radial_quadrature(const So3Args &a, double R, double (&I)[LC][NMAX],
                                             double (&dI)[LC][NMAX]) {
    const int L = a.lmax + 1;
    const double ta = 2.0 * a.alpha;
#pragma unroll
    for (int l = 0; l < LC; l++)
#pragma unroll
        for (int n = 0; n < NMAX; n++) {
            I[l][n] = 0.0;
            dI[l][n] = 0.0;
        }
    for (int q = 0; q < a.nq; q++) {
        const double rq = a.rq[q];
        // the node's quadrature weights, loaded once for every order
        double g[NMAX];
#pragma unroll
        for (int n = 0; n < NMAX; n++) g[n] = a.G[n * a.nq + q];
        const double x = ta * R * rq;
        const double ix = 1.0 / x;
        double f[LC];
        sph_bessel_i_all<LC>(x, ix, f);
        const double tq = ta * rq;
#pragma unroll
        for (int l = 0; l < LC; l++) {
            if (l >= L) continue;  // uniform
            // i_0' = i_1; i_l' = i_{l-1} - (l+1)/x i_l
            const double dil = (l == 0) ? f[1] : f[l - 1] - (l + 1) * ix * f[l];
            const double fd = tq * dil;
#pragma unroll
            for (int n = 0; n < NMAX; n++) {
                I[l][n] += g[n] * f[l];
                dI[l][n] += g[n] * fd;
            }
        }
    }
}

template <int NMAX, int LC>
__device__ inline void so3_radial_image(const So3Args &a, int64_t e, int64_t re, const double *__restrict__ img_r,
                                        double *__restrict__ rad) {
    const int L = a.lmax + 1;
    const double x0 = img_r[3 * e], y0 = img_r[3 * e + 1], z0 = img_r[3 * e + 2];
    const double R = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
    double I[LC][NMAX], dI[LC][NMAX];
    so3_radial_quadrature<NMAX, LC>(a, R, I, dI);
#pragma unroll
    for (int w = 0; w < 2; w++) {
        if (w == 1 && re == e) break;
        const int64_t row = w == 0 ? e : re;
#pragma unroll
        for (int l = 0; l < LC; l++) {
            if (l >= L) continue;
            double *ro = rad + ((size_t)row * L + l) * 2 * NMAX;
#pragma unroll
            for (int n = 0; n < NMAX; n++) {
                ro[n] = I[l][n];
                ro[NMAX + n] = dI[l][n];
            }
        }
    }
}

template <int NMAX, int LC>
__global__ __launch_bounds__(256) void so3_radial_all_kernel(So3Args a, int64_t natoms,
                                                             const int64_t *__restrict__ img_start,
                                                             const double *__restrict__ img_r,
                                                             const int64_t *__restrict__ img_rev,
                                                             double *__restrict__ rad) {
    __shared__ int64_t s_q[4][128];  // per wavefront: queued images (at most 63 + 64)
    const int64_t nimg = img_start[natoms];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t wid = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
    int64_t *q = s_q[wv];
    int pend = 0;  // uniform over the wavefront
    for (int64_t c0 = wid * 64; c0 < nimg; c0 += nw * 64) {
        const int64_t e = c0 + lane;
        bool take = false;
        if (e < nimg) take = img_rev ? img_rev[e] >= e : true;
        const uint64_t mask = __ballot(take);
        if (take) q[pend + (int)__popcll(mask & ((1ull << lane) - 1ull))] = e;
        pend += (int)__popcll(mask);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (pend >= 64) {
            const int64_t my = q[lane];
            const int rest = pend - 64;
            const int64_t mv = lane < rest ? q[64 + lane] : 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < rest) q[lane] = mv;
            pend = rest;
            so3_radial_image<NMAX, LC>(a, my, img_rev ? img_rev[my] : my, img_r, rad);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (lane < pend) {
        const int64_t my = q[lane];
        so3_radial_image<NMAX, LC>(a, my, img_rev ? img_rev[my] : my, img_r, rad);
    }
}

// Scheduling of the radial kernel by |r|: a counting sort of the images
// that the pair store evaluates (rev(e) >= e) into SO3_RBINS bins of |r| over
// (0, rc], so that the 64 lanes of a wavefront hold images of nearly equal
// |r|.  At each quadrature node their Bessel arguments x = 2 alpha |r| r_q
// then fall in the same branch of sph_bessel_i_all (series, Miller,
// closed form), which otherwise diverge and cost the wavefront all three.
// Only the order in which images are processed changes: every image's
// integrals are the same computation as in so3_radial_all_kernel.
// The sort is the two-level one of a radix-sort pass, with no global
// atomics: SO3_RBLK workgroups each histogram a contiguous chunk of the
// images in LDS and store it bin-major, one workgroup scans the
// SO3_RBINS x SO3_RBLK counts into offsets, and each workgroup places its
// chunk from its own offsets (LDS atomics: the order inside a bin of a
// chunk is arbitrary, which the values do not see).
// Workspace (int64): offsets [SO3_RBINS][SO3_RBLK], the number of sorted
// images, the order [image capacity].
constexpr int SO3_RBINS = 256, SO3_RBLK = 64, SO3_RBT = 1024;  // bins, chunks, threads per chunk
constexpr int64_t SO3_RWORK = (int64_t)SO3_RBINS * SO3_RBLK + 1;
__device__ inline int so3_rbin(const double *__restrict__ img_r, int64_t e, double rc) {
    const double x = img_r[3 * e], y = img_r[3 * e + 1], z = img_r[3 * e + 2];
    const double R = sqrt(x * x + y * y + z * z);
    int b = (int)(R / rc * SO3_RBINS);
    return b < 0 ? 0 : (b >= SO3_RBINS ? SO3_RBINS - 1 : b);
}

// the chunk [lo, hi) of workgroup blockIdx.x
__device__ inline void so3_rchunk(int64_t nimg, int64_t &lo, int64_t &hi) {
    const int64_t ch = (nimg + SO3_RBLK - 1) / SO3_RBLK;
    lo = (int64_t)blockIdx.x * ch;
    hi = lo + ch < nimg ? lo + ch : nimg;
}

__global__ __launch_bounds__(SO3_RBT) void so3_rbin_hist_kernel(int64_t natoms, const int64_t *__restrict__ img_start,
                                                            const double *__restrict__ img_r,
                                                            const int64_t *__restrict__ img_rev, double rc,
                                                            int64_t *__restrict__ work) {
    __shared__ unsigned int h[SO3_RBINS];
    for (int k = threadIdx.x; k < SO3_RBINS; k += blockDim.x) h[k] = 0;
    __syncthreads();
    int64_t lo, hi;
    so3_rchunk(img_start[natoms], lo, hi);
    for (int64_t e = lo + threadIdx.x; e < hi; e += blockDim.x)
        if (img_rev[e] >= e) atomicAdd(&h[so3_rbin(img_r, e, rc)], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < SO3_RBINS; k += blockDim.x) work[(int64_t)k * SO3_RBLK + blockIdx.x] = h[k];
}

// exclusive scan of the SO3_RBINS * SO3_RBLK counts in place (bin-major:
// bin b's images from workgroup g follow those of bins < b and of
// workgroups < g in bin b).  One workgroup: thread t sums the 16
// consecutive counts it owns (one 128-byte line), one workgroup scan of the
// 1024 thread sums, then thread t rewrites its counts as offsets.
constexpr int SO3_RSCAN_T = 1024;
__global__ __launch_bounds__(SO3_RSCAN_T) void so3_rbin_scan_kernel(int64_t *__restrict__ work) {
    constexpr int TOT = SO3_RBINS * SO3_RBLK;
    constexpr int PER = TOT / SO3_RSCAN_T;
    static_assert(TOT % SO3_RSCAN_T == 0, "scan layout");
    __shared__ int64_t sc[SO3_RSCAN_T / 64];
    const int t = threadIdx.x;
    int64_t *w = work + (int64_t)t * PER;
    int64_t loc = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) loc += w[k];
    // inclusive scan: within each wavefront by shuffles, then over the 16
    // wavefront totals
    const int lane = t & 63, wv = t >> 6;
    int64_t inc = loc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) sc[wv] = inc;
    __syncthreads();
    int64_t before = 0;
    for (int k = 0; k < wv; k++) before += sc[k];
    const int64_t total = before + (wv == SO3_RSCAN_T / 64 - 1 ? inc : 0);
    int64_t run = before + inc - loc;
    // second pass over the (cached) counts
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int64_t v = w[k];
        w[k] = run;
        run += v;
    }
    if (t == SO3_RSCAN_T - 1) work[SO3_RWORK - 1] = total;  // sorted images
}

__global__ __launch_bounds__(SO3_RBT) void so3_rbin_fill_kernel(int64_t natoms, const int64_t *__restrict__ img_start,
                                                            const double *__restrict__ img_r,
                                                            const int64_t *__restrict__ img_rev, double rc,
                                                            int64_t *__restrict__ work) {
    __shared__ unsigned long long cur[SO3_RBINS];
    for (int k = threadIdx.x; k < SO3_RBINS; k += blockDim.x)
        cur[k] = (unsigned long long)work[(int64_t)k * SO3_RBLK + blockIdx.x];
    __syncthreads();
    int64_t *order = work + SO3_RWORK;
    int64_t lo, hi;
    so3_rchunk(img_start[natoms], lo, hi);
    for (int64_t e = lo + threadIdx.x; e < hi; e += blockDim.x)
        if (img_rev[e] >= e) order[atomicAdd(&cur[so3_rbin(img_r, e, rc)], 1ull)] = e;
}

// ------------------------------------------------ tabulated radial integrals
// For a fixed parameter set I_nl(R) and dI_nl/dR are smooth functions of
// |r| alone on [0, rc].  They are tabulated once, piecewise on NK equal
// intervals, as polynomials of degree SO3_TDEG in the local variable
// t = 2 (R/h - k) - 1 in [-1, 1]: the quadrature above is evaluated at the
// SO3_TDEG + 1 Chebyshev nodes of every interval (so3_rtab_nodes_kernel),
// and a fixed linear map (the Chebyshev interpolant at those nodes, written
// in powers of t; so3_rtab_weights) turns the values into coefficients
// (so3_rtab_coef_kernel).  With NK chosen from the steepness 2 alpha rc^2
// (>= 256 intervals; 2.5 per unit of it, rounded up to a power of two) the
// interpolant matches the quadrature to about 1e-14 of each function's
// largest value (checked against the quadrature with SciPy's spherical_in).
// Table layout: [NK][(lmax+1) * 2 * nmax][SO3_TDEG + 1], the functions in
// the radial store's order (l, {I, dI}, n).
constexpr int SO3_TDEG = 11;
constexpr int SO3_TNP = SO3_TDEG + 1;

struct So3TabW {
    double w[SO3_TNP][SO3_TNP];  // coefficient p of t^p = sum_j w[p][j] f(t_j)
};

// t_j = cos(pi (j + 1/2) / NP) and the map from node values to monomial
// coefficients of the interpolant (Chebyshev coefficients c_i = (2 - d_i0) /
// NP sum_j T_i(t_j) f_j, then T_i written in powers of t).
static So3TabW so3_rtab_weights() {
    So3TabW W{};
    double tn[SO3_TNP], T[SO3_TNP][SO3_TNP], M[SO3_TNP][SO3_TNP];  // T[i][j] = T_i(t_j); M[i][p]: t^p in T_i
    for (int j = 0; j < SO3_TNP; j++) tn[j] = std::cos(M_PI * (j + 0.5) / SO3_TNP);
    for (int j = 0; j < SO3_TNP; j++) {
        T[0][j] = 1.0;
        T[1][j] = tn[j];
        for (int i = 2; i < SO3_TNP; i++) T[i][j] = 2.0 * tn[j] * T[i - 1][j] - T[i - 2][j];
    }
    for (int i = 0; i < SO3_TNP; i++)
        for (int q = 0; q < SO3_TNP; q++) M[i][q] = 0.0;
    M[0][0] = 1.0;
    M[1][1] = 1.0;
    for (int i = 2; i < SO3_TNP; i++)
        for (int q = 0; q < SO3_TNP; q++) M[i][q] = (q > 0 ? 2.0 * M[i - 1][q - 1] : 0.0) - M[i - 2][q];
    for (int q = 0; q < SO3_TNP; q++)
        for (int j = 0; j < SO3_TNP; j++) {
            double acc = 0.0;
            for (int i = 0; i < SO3_TNP; i++) acc += M[i][q] * (i == 0 ? 1.0 : 2.0) / SO3_TNP * T[i][j];
            W.w[q][j] = acc;
        }
    return W;
}

static int so3_rtab_nk(const pxff_so3_params *p) {
    const double steep = 2.0 * p->alpha * p->rcut * p->rcut;
    int nk = 256;
    while (nk < 4096 && nk < 2.5 * steep) nk *= 2;
    return nk;
}

// node values: vals[(k * NP + j) * nf + f], one thread per (interval, node)
template <int NMAX, int LC>
__global__ __launch_bounds__(256) void so3_rtab_nodes_kernel(So3Args a, int nk, double h, double *__restrict__ vals) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * SO3_TNP) return;
    const int k = t / SO3_TNP, j = t - k * SO3_TNP;
    const double tj = cos(M_PI * (j + 0.5) / SO3_TNP);
    const double R = (k + 0.5 + 0.5 * tj) * h;
    double I[LC][NMAX], dI[LC][NMAX];
    so3_radial_quadrature<NMAX, LC>(a, R, I, dI);
    const int L = a.lmax + 1, nf = L * 2 * NMAX;
    double *o = vals + (size_t)t * nf;
#pragma unroll
    for (int l = 0; l < LC; l++) {
        if (l >= L) continue;
#pragma unroll
        for (int n = 0; n < NMAX; n++) {
            o[(l * 2) * NMAX + n] = I[l][n];
            o[(l * 2 + 1) * NMAX + n] = dI[l][n];
        }
    }
}

// coefficients: tab[(k * nf + f) * NP + p], one thread per (interval, function)
__global__ __launch_bounds__(256) void so3_rtab_coef_kernel(So3TabW W, int nk, int nf, const double *__restrict__ vals,
                                                            double *__restrict__ tab) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * nf) return;
    const int k = t / nf, f = t - k * nf;
    double v[SO3_TNP];
#pragma unroll
    for (int j = 0; j < SO3_TNP; j++) v[j] = vals[((size_t)k * SO3_TNP + j) * nf + f];
#pragma unroll
    for (int q = 0; q < SO3_TNP; q++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < SO3_TNP; j++) acc = fma(W.w[q][j], v[j], acc);
        tab[(size_t)t * SO3_TNP + q] = acc;
    }
}

// A wave-uniform copy of lane r's 64-bit value (r uniform).
__device__ inline int64_t so3_readlane64(int64_t x, int r) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, r);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), r);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// The radial store from the table, image pairs in |r| order (as
// so3_radial_sorted_kernel), one wavefront per 64 pairs.  Lane j of the
// wavefront owns the functions 2j and 2j + 1 of the row and the wavefront
// walks its pairs one by one: each pair's row then goes out as one 16-byte
// store per lane, contiguous over the row, to both images of the pair (the
// rows are nf <= 84 doubles; a lane per pair writing its own row scatters
// every store instruction over 64 rows, and those stores bounded the
// kernel).  The coefficients of the lane's two functions stay in registers
// while consecutive pairs share an interval, the common case in |r| order.
template <int NMAX, int LC>
__global__ __launch_bounds__(64) void so3_radial_tab_kernel(So3Args a, const int64_t *__restrict__ work,
                                                            const double *__restrict__ img_r,
                                                            const int64_t *__restrict__ img_rev,
                                                            const double *__restrict__ tab, int nk, double h,
                                                            double *__restrict__ rad) {
    const int64_t n = work[SO3_RWORK - 1];
    const int64_t *order = work + SO3_RWORK;
    const int64_t base = (int64_t)blockIdx.x * 64;
    // the grid covers the image capacity: wavefronts past the sorted pairs
    // leave at once
    if (base >= n) return;
    const int lane = threadIdx.x;
    const int cnt = n - base < 64 ? (int)(n - base) : 64;
    int64_t e = 0, re = 0, tb = 0;
    int k = 0;
    if (lane < cnt) {
        e = order[base + lane];
        re = img_rev[e];
        const double x0 = img_r[3 * e], y0 = img_r[3 * e + 1], z0 = img_r[3 * e + 2];
        const double u = sqrt(x0 * x0 + y0 * y0 + z0 * z0) / h;
        k = (int)u;
        k = k < 0 ? 0 : (k >= nk ? nk - 1 : k);
        tb = __double_as_longlong(2.0 * (u - k) - 1.0);
    }
    const int nf = (a.lmax + 1) * 2 * NMAX;
    const bool act = 2 * lane < nf;
    const int f0 = act ? 2 * lane : 0;
    double c0[SO3_TNP], c1[SO3_TNP];
    int kc = -1;
    for (int r = 0; r < cnt; r++) {
        const int kr = __builtin_amdgcn_readlane(k, r);
        if (kr != kc) {
            kc = kr;
            const double2 *src = reinterpret_cast<const double2 *>(tab + ((size_t)kr * nf + f0) * SO3_TNP);
#pragma unroll
            for (int q = 0; q < SO3_TNP / 2; q++) {
                const double2 v = src[q], w = src[SO3_TNP / 2 + q];
                c0[2 * q] = v.x;
                c0[2 * q + 1] = v.y;
                c1[2 * q] = w.x;
                c1[2 * q + 1] = w.y;
            }
        }
        const double t = __longlong_as_double(so3_readlane64(tb, r));
        double v0 = c0[SO3_TDEG], v1 = c1[SO3_TDEG];
#pragma unroll
        for (int q = SO3_TDEG - 1; q >= 0; q--) {
            v0 = fma(v0, t, c0[q]);
            v1 = fma(v1, t, c1[q]);
        }
        const int64_t er = so3_readlane64(e, r), rer = so3_readlane64(re, r);
        if (act) {
            const double2 v = make_double2(v0, v1);
            *reinterpret_cast<double2 *>(rad + (size_t)er * nf + f0) = v;
            if (rer != er) *reinterpret_cast<double2 *>(rad + (size_t)rer * nf + f0) = v;
        }
    }
}

// The quadrature for the image pairs in |r| order, one lane per pair and
// one wavefront per 64 pairs.  The rows are written one l at a time through
// LDS, 2 nmax doubles of every pair's row at once, so that a store
// instruction covers about 11 row pieces of 16 nmax bytes instead of 64
// scattered 8-byte pieces.
template <int NMAX, int LC>
__global__ __launch_bounds__(64) void so3_radial_sorted_kernel(So3Args a, const int64_t *__restrict__ work,
                                                               const double *__restrict__ img_r,
                                                               const int64_t *__restrict__ img_rev,
                                                               double *__restrict__ rad) {
    __shared__ double buf[64][2 * NMAX + 1];
    __shared__ int64_t s_e[64], s_re[64];
    const int64_t n = work[SO3_RWORK - 1];
    const int64_t *order = work + SO3_RWORK;
    const int64_t base = (int64_t)blockIdx.x * 64;
    if (base >= n) return;
    const int lane = threadIdx.x;
    const int cnt = n - base < 64 ? (int)(n - base) : 64;
    int64_t e = 0, re = 0;
    double R = 1.0;
    if (lane < cnt) {
        e = order[base + lane];
        re = img_rev[e];
        const double x0 = img_r[3 * e], y0 = img_r[3 * e + 1], z0 = img_r[3 * e + 2];
        R = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
    }
    s_e[lane] = e;
    s_re[lane] = re;
    const int L = a.lmax + 1, nf = L * 2 * NMAX;
    double I[LC][NMAX], dI[LC][NMAX];
    so3_radial_quadrature<NMAX, LC>(a, R, I, dI);
#pragma unroll
    for (int l = 0; l < LC; l++) {
        if (l >= L) break;
#pragma unroll
        for (int nn = 0; nn < NMAX; nn++) {
            buf[lane][nn] = I[l][nn];
            buf[lane][NMAX + nn] = dI[l][nn];
        }
        __syncthreads();
        for (int q = lane; q < cnt * NMAX; q += 64) {
            const int r = q / NMAX, c = q - r * NMAX;
            const double2 v = make_double2(buf[r][2 * c], buf[r][2 * c + 1]);
            const int64_t er = s_e[r], rer = s_re[r];
            *reinterpret_cast<double2 *>(rad + (size_t)er * nf + 2 * l * NMAX + 2 * c) = v;
            if (rer != er) *reinterpret_cast<double2 *>(rad + (size_t)rer * nf + 2 * l * NMAX + 2 * c) = v;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ accumulate
// ctot[i][n][l][m] = sum_e c_e  (MODE 0)  or its tangent
// sum_e sum_k dr_e,k dc_e/dr_k (MODE 1).  One workgroup (64 threads) per
// centre, persistent.  Round: NG images; lanes (image, l) compute the radial
// sums and the per-(l, m) factors into LDS; then every thread adds the
// contributions of the round's images to the ctot entries it owns (LDS
// accumulator, fixed order -> deterministic).
template <int NMAX, int MODE>
__global__ __launch_bounds__(64) void so3_accum_kernel(So3Args a, int64_t natoms, const int64_t *__restrict__ img_start,
                                                        const int64_t *__restrict__ img_j,
                                                        const double *__restrict__ img_r,
                                                        const int64_t *__restrict__ atom_z,
                                                        const double *__restrict__ dr, double *__restrict__ out,
                                                        const double *__restrict__ rad_in) {
    constexpr bool TAN = MODE == SO3_TANGENT;
    extern __shared__ double smem[];
    const int L = a.lmax + 1;
    const int LP = a.lmax + 1;
    const int NG = 64 / L;
    const int M2 = 2 * a.lmax + 1;
    const int ysz = (a.lmax + 2) * (2 * LP + 1);
    const int csz = NMAX * L * M2;
    // LDS: the radial quadrature tables stay in global memory (their
    // addresses are wave-uniform); the tangent-only arrays exist only in
    // the tangent instantiation
    double *lnorm = smem;                                    // [L]
    double *Ir = lnorm + L;                                  // [NG][NMAX][L]
    double *dIr = Ir + NG * NMAX * L;                        // [NG][NMAX][L] (tangent)
    cd *Yt = reinterpret_cast<cd *>(dIr + (TAN ? NG * NMAX * L : 0) + ((L + NG * NMAX * L * (TAN ? 2 : 1)) & 1));
    cd *X = Yt + NG * ysz;                                   // [NG][L][M2]
    cd *Z2 = X + NG * L * M2;                                // [NG][L][M2] (tangent)
    cd *acc = Z2 + (TAN ? NG * L * M2 : 0);                  // [csz]
    // the entries with |m| <= l, nval = NMAX L^2 of the csz: kv[2i] = entry,
    // kv[2i+1] = (l M2 + m + lmax) | (n L + l) << 16 (X and radial offsets)
    int *kv = reinterpret_cast<int *>(acc + csz);
    const int nval = NMAX * L * L;
    const int lane = threadIdx.x;
    const int g = lane / L, l = lane % L;
    for (int k = lane; k < L; k += 64) lnorm[k] = a.lnorm[k];
    for (int i = lane; i < nval; i += 64) {
        const int n = i / (L * L), r = i - n * L * L;
        int ll = 0;
        while ((ll + 1) * (ll + 1) <= r) ll++;
        const int mi = r - ll * ll - ll + a.lmax;
        kv[2 * i] = (n * L + ll) * M2 + mi;
        kv[2 * i + 1] = (ll * M2 + mi) | ((n * L + ll) << 16);
    }
    So3Args al = a;
    al.lnorm = lnorm;
    for (int64_t c = blockIdx.x; c < natoms; c += gridDim.x) {
        __syncthreads();
        for (int k = lane; k < csz; k += 64) acc[k] = cmake(0.0, 0.0);
        const int64_t s0 = img_start[c], s1 = img_start[c + 1];
        for (int64_t base = s0; base < s1; base += NG) {
            const int64_t e = base + g;
            const bool valid = (g < NG) && (e < s1);
            So3Img m;
            load_image(al, valid, e, c, img_j, img_r, atom_z, m);
            double I[NMAX], dI[NMAX];
            if (valid) {
                if (rad_in) {  // from so3_radial_kernel
                    const double *ri = rad_in + ((size_t)e * L + l) * 2 * NMAX;
#pragma unroll
                    for (int n = 0; n < NMAX; n++) {
                        I[n] = ri[n];
                        dI[n] = ri[NMAX + n];
                    }
                } else {
                    radial<NMAX>(al, l, m.R, a.G, a.rq, I, dI, TAN);
                }
                fill_ylm(m, l, L, a.lmax, Yt + g * ysz, LP);
#pragma unroll
                for (int n = 0; n < NMAX; n++) {
                    Ir[(g * NMAX + n) * L + l] = I[n];
                    if (TAN) dIr[(g * NMAX + n) * L + l] = dI[n];
                }
            }
            __syncthreads();
            if (valid) {
                double d[3] = {0, 0, 0};
                if (TAN) {
                    d[0] = dr[3 * e];
                    d[1] = dr[3 * e + 1];
                    d[2] = dr[3 * e + 2];
                }
                const cd *Y = Yt + g * ysz;
                for (int mm = -l; mm <= l; mm++) {
                    cd P, U[3], V[3];
                    image_factors(al, m, l, mm, Y, LP, P, U, V);
                    const int o = (g * L + l) * M2 + a.lmax + mm;
                    if (!TAN) {
                        X[o] = P;
                    } else {
                        X[o] = cadd(cadd(cscale(U[0], d[0]), cscale(U[1], d[1])), cscale(U[2], d[2]));
                        Z2[o] = cadd(cadd(cscale(V[0], d[0]), cscale(V[1], d[1])), cscale(V[2], d[2]));
                    }
                }
            }
            __syncthreads();
            const int nvalid = (int)((s1 - base) < NG ? (s1 - base) : NG);
            for (int i = lane; i < nval; i += 64) {
                const int k = kv[2 * i], w = kv[2 * i + 1];
                const int oX = w & 0xffff, oI = w >> 16;
                cd s = acc[k];
                for (int gg = 0; gg < nvalid; gg++) {
                    const int o = gg * L * M2 + oX;
                    const double iv = Ir[gg * NMAX * L + oI];
                    s = cadd(s, cscale(X[o], iv));
                    if (TAN) s = cadd(s, cscale(Z2[o], dIr[gg * NMAX * L + oI]));
                }
                acc[k] = s;
            }
        }
        __syncthreads();
        cd *oc = reinterpret_cast<cd *>(out) + (size_t)c * csz;
        for (int k = lane; k < csz; k += 64) oc[k] = acc[k];
    }
}

// Y_l^|m| without tables, for l <= 7: Pbar_l^|m|(cos theta) e^{i|m|phi} =
// Q(t) z^|m| with t = cos theta, z = (x + i y) / R = sin theta e^{i phi} and
// Q = Pbar_l^|m| / sin^|m| theta a polynomial of degree l - |m| in t.  Its
// coefficients q[j] (of t^j) come from the normalised associated Legendre
// recurrence of ylm_row (the same coefficient tables) run on coefficient
// vectors: P_|m| = sqrt(1/4pi) prod_k pmm_k s^|m|, P_|m|+1 = p1 t P_|m|,
// P_k = lega_k (t P_k-1 - legb_k P_k-2).  With on false q is P_|m| / s^|m|
// alone (the lanes without a row).
constexpr int SO3_QK = 8;
__device__ __forceinline__ void so3_legendre_poly(int l, int am, bool on, double (&q)[SO3_QK]) {
    double q2[SO3_QK];
    double c0 = sqrt(0.25 / M_PI);
#pragma unroll
    for (int k = 1; k < SO3_QK; k++)
        if (k <= am) c0 *= g_so3_pmm[k];
#pragma unroll
    for (int j = 0; j < SO3_QK; j++) q[j] = j == 0 ? c0 : 0.0;
    if (on && l > am) {
        const double p1 = g_so3_p1[am];
#pragma unroll
        for (int j = SO3_QK - 1; j >= 0; j--) {
            q2[j] = q[j];
            q[j] = j > 0 ? p1 * q[j - 1] : 0.0;
        }
#pragma unroll
        for (int k = 2; k < SO3_QK; k++)
            if (k >= am + 2 && k <= l) {
                const double la = g_so3_lega[k * SO3_LEGD + am], lb = g_so3_legb[k * SO3_LEGD + am];
#pragma unroll
                for (int j = SO3_QK - 1; j >= 0; j--) {
                    const double nq = la * ((j > 0 ? q[j - 1] : 0.0) - lb * q2[j]);
                    q2[j] = q[j];
                    q[j] = nq;
                }
            }
    }
}
// Q(t) by Horner's rule
__device__ __forceinline__ double so3_legendre_eval(const double (&q)[SO3_QK], double t) {
    double p = q[SO3_QK - 1];
#pragma unroll
    for (int j = SO3_QK - 2; j >= 0; j--) p = fma(p, t, q[j]);
    return p;
}
// z^am, 0 <= am <= 7: two squarings and two products
__device__ __forceinline__ cd so3_zpow(cd z, int am) {
    const cd z2 = cmul(z, z), z4 = cmul(z2, z2);
    cd w = (am & 1) ? z : cmake(1.0, 0.0);
    if (am & 2) w = cmul(w, z2);
    if (am & 4) w = cmul(w, z4);
    return w;
}

// ctot[i][n][l][m] = sum_e c_e,nlm = sum_e w norm_l ex fc I_nl Y_lm (SO3.py:
// 205-226) with the radial integrals I_nl read from the radial store.  One
// wavefront per centre (persistent), one lane per (l, m), (lmax + 1)^2 <= 64:
// lane (l, m) keeps its NMAX sums in registers over all of the centre's
// images, so no reduction is needed.  The images run in chunks of
// SO3_LM_CH: the per-image scalars are computed once per chunk (one lane
// each) and staged in LDS, where every lane reads the same address
// (broadcast), and the chunk's I_nl rows are copied from the radial store to
// LDS beside them, all loads of a batch in flight together (one memory
// latency per chunk instead of one per image).  Y_lm of the lane's (l, |m|)
// as Q(t) z^|m| (so3_legendre_poly, built once per lane), so per image a
// lane evaluates Q by Horner's rule and z^|m| by two squarings and two
// products; Y_l^-m = (-1)^m conj(Y_l^m).  Entries with |m| > l are written
// as zero.
constexpr int SO3_LM_NST = 4;  // staged per image: t, Re z, Im z, wf (w ex fc)
// images per chunk: a centre of up to 40 neighbours in one chunk, and with
// NMAX <= 6 at most 18 KB of LDS (eight workgroups per CU)
constexpr int SO3_LM_CH = 40;
constexpr int SO3_LM_RB = 12;  // I_nl loads in flight per lane
template <int NMAX>
__global__ __launch_bounds__(64) void so3_ctot_lm_kernel(So3Args a, int64_t natoms, const int64_t *__restrict__ img_start,
                                                          const int64_t *__restrict__ img_j,
                                                          const double *__restrict__ img_r,
                                                          const int64_t *__restrict__ atom_z,
                                                          const double *__restrict__ rad, double *__restrict__ out) {
    constexpr int CH = SO3_LM_CH;
    __shared__ double stg[SO3_LM_NST][CH];
    __shared__ double rI[CH * 8 * NMAX];  // [image][l][n], L <= 8
    const int L = a.lmax + 1;
    const int M2 = 2 * a.lmax + 1;
    const int lane = threadIdx.x;
    // the lane's (l, m): lane = l^2 + l + m
    int l = 0;
    while ((l + 1) * (l + 1) <= lane) l++;
    const int m = lane - l * l - l;
    const bool act = lane < L * L;
    const int am = m < 0 ? -m : m;
    // Q = Pbar_l^|m| / s^|m| of the lane (lmax <= 7)
    double q[SO3_QK];
    so3_legendre_poly(l, am, act, q);
    const double lnorm = act ? a.lnorm[l] : 0.0;
    const double sgn = (m < 0 && (am & 1)) ? -1.0 : 1.0;
    for (int64_t c = blockIdx.x; c < natoms; c += gridDim.x) {
        double accr[NMAX], acci[NMAX];
#pragma unroll
        for (int n = 0; n < NMAX; n++) {
            accr[n] = 0.0;
            acci[n] = 0.0;
        }
        const int64_t s0 = img_start[c], s1 = img_start[c + 1];
        for (int64_t b64 = s0; b64 < s1; b64 += CH) {
            const int nb = (int)((s1 - b64) < CH ? (s1 - b64) : CH);
            __syncthreads();  // the previous chunk's stage has been read
            {
                const int64_t e = b64 + lane;
                const bool ok = lane < nb;
                So3Img mi;
                load_image(a, ok, e, c, img_j, img_r, atom_z, mi);
                if (lane < CH) {
                    stg[0][lane] = mi.z / mi.R;  // t = cos theta
                    stg[1][lane] = mi.x / mi.R;  // z = sin theta e^{i phi}
                    stg[2][lane] = mi.y / mi.R;
                    stg[3][lane] = ok ? mi.w * mi.ex * mi.fc : 0.0;
                }
            }
            {
                // rI[(b L + l) NMAX + n] = I_nl of image b64 + b: the first
                // NMAX of each 2 NMAX-row of the radial store
                const double *rb = rad + (size_t)b64 * L * 2 * NMAX;
                const int nri = nb * L * NMAX;
                for (int k0 = 0; k0 < nri; k0 += 64 * SO3_LM_RB) {
                    double t[SO3_LM_RB];
#pragma unroll
                    for (int u = 0; u < SO3_LM_RB; u++) {
                        const int k = k0 + u * 64 + lane;
                        const int row = k / NMAX;
                        t[u] = k < nri ? rb[(size_t)row * 2 * NMAX + (k - row * NMAX)] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < SO3_LM_RB; u++) {
                        const int k = k0 + u * 64 + lane;
                        if (k < nri) rI[k] = t[u];
                    }
                }
            }
            __syncthreads();
            if (act) {
                for (int b = 0; b < nb; b++) {
                    const double ct = stg[0][b];
                    const cd z = cmake(stg[1][b], stg[2][b]);
                    const double wf = stg[3][b] * lnorm;
                    const double pl = so3_legendre_eval(q, ct);
                    const cd em = so3_zpow(z, am);
                    // Y_l^m w ex fc norm_l, then times I_nl
                    const double yr = pl * em.re * wf * sgn, yi = (m < 0 ? -1.0 : 1.0) * pl * em.im * wf * sgn;
                    const double *ri = rI + (b * L + l) * NMAX;
#pragma unroll
                    for (int n = 0; n < NMAX; n++) {
                        const double in = ri[n];
                        accr[n] = fma(in, yr, accr[n]);
                        acci[n] = fma(in, yi, acci[n]);
                    }
                }
            }
        }
        cd *oc = reinterpret_cast<cd *>(out) + (size_t)c * NMAX * L * M2;
        if (act) {
#pragma unroll
            for (int n = 0; n < NMAX; n++) oc[(n * L + l) * M2 + a.lmax + m] = cmake(accr[n], acci[n]);
        }
        // the entries with |m| > l are zero
        for (int k = lane; k < NMAX * L * M2; k += 64) {
            const int ll = (k / M2) % L, mm = k % M2 - a.lmax;
            if (mm > ll || mm < -ll) oc[k] = cmake(0.0, 0.0);
        }
    }
}

// A double moved across lanes by one DPP control (two 32-bit moves);
// lanes of rows outside ROWMASK read 0.
template <int CTRL, int ROWMASK = 0xf>
__device__ inline double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront, complete in lane 63: pairs and
// quads (quad_perm), half rows and rows (row_half_mirror, row_mirror), then
// row 0 into row 1 and row 2 into row 3 (row_bcast:15), rows 0-1 into rows
// 2-3 (row_bcast:31).  Vector moves only, no LDS round trips.
__device__ inline void wave_sum3_lane63(double &a, double &b, double &c) {
    a += dpp_f64<0xb1>(a); b += dpp_f64<0xb1>(b); c += dpp_f64<0xb1>(c);
    a += dpp_f64<0x4e>(a); b += dpp_f64<0x4e>(b); c += dpp_f64<0x4e>(c);
    a += dpp_f64<0x141>(a); b += dpp_f64<0x141>(b); c += dpp_f64<0x141>(c);
    a += dpp_f64<0x140>(a); b += dpp_f64<0x140>(b); c += dpp_f64<0x140>(c);
    a += dpp_f64<0x142, 0xa>(a); b += dpp_f64<0x142, 0xa>(b); c += dpp_f64<0x142, 0xa>(c);
    a += dpp_f64<0x143, 0xc>(a); b += dpp_f64<0x143, 0xc>(b); c += dpp_f64<0x143, 0xc>(c);
}

// Adjoint dE/dr per image, one lane per (l, m) with l <= lmax + 1
// ((lmax + 2)^2 <= 64): flat over images, a wavefront takes 64 consecutive
// images (their scalars staged in LDS, one lane each) and runs them one at
// a time.  Per image every lane forms its Y_lm (rows <= lmax + 1) as
// Q(cos theta) z^|m| (so3_legendre_poly, built once per lane) into a
// 64-entry LDS table, from which the lanes with l <= lmax read the
// Y_{l+-1, m'} of the covariant gradient (SO3.py:649-674; its coefficients
// depend on (l, m) only and are held in registers).  With
// alpha = sum_n I_nl A_nlm and beta = sum_n dI_nl A_nlm (A of the image's
// centre, reloaded when the centre changes),
//   dE/dr_k = sum_lm  cu rhat_k Re(conj(alpha) Y) + cv Re(conj(alpha) dY/dr_k)
//                   + cv rhat_k Re(conj(beta) Y),
// cu = w norm_l ex (dfc - 2 alpha R fc), cv = w norm_l ex fc: the same sums
// as so3_image_kernel<SO3_ADJOINT> in another order, reduced over the
// lanes with DPP moves (wave_sum3_lane63).
// staged: t, Re z, Im z, (unused), R, w ex fc, w ex (dfc - 2 a R fc), rhat
constexpr int SO3_ADJ_NST = 10;
template <int NMAX>
__global__ __launch_bounds__(64) void so3_deidrj_lm_kernel(So3Args a, int64_t nimg, const int64_t *__restrict__ img_atom,
                                                            const int64_t *__restrict__ img_j,
                                                            const double *__restrict__ img_r,
                                                            const int64_t *__restrict__ atom_z,
                                                            const double *__restrict__ A,
                                                            const double *__restrict__ rad, double *__restrict__ dedr,
                                                            const int64_t *__restrict__ d_count) {
    __shared__ double stg[SO3_ADJ_NST][64];
    __shared__ int64_t s_at[64];
    __shared__ cd ysh[64];
    if (d_count) {
        const int64_t c = *d_count;
        nimg = c < nimg ? c : nimg;
    }
    const int L = a.lmax + 1;
    const int M2 = 2 * a.lmax + 1;
    const int lane = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * 64;
    if (base >= nimg) return;  // a launch over an image capacity
    const int nb = (int)((nimg - base) < 64 ? (nimg - base) : 64);
    int l = 0;
    while ((l + 1) * (l + 1) <= lane) l++;
    const int m = lane - l * l - l;
    const bool ylane = lane < (L + 1) * (L + 1);  // rows 0..lmax+1
    const bool act = lane < L * L;                // rows 0..lmax
    const int am = m < 0 ? -m : m;
    double q[SO3_QK];  // Y rows l <= lmax + 1 <= 7
    so3_legendre_poly(l, am, ylane, q);
    const double msg = (m < 0 && (am & 1)) ? -1.0 : 1.0, imsg = m < 0 ? -msg : msg;
    // covariant gradient coefficients of the lane's (l, m) and the lanes of
    // the Y values they multiply (-1: term absent)
    const double dl = l, dm = m;
    double gc[6] = {0, 0, 0, 0, 0, 0};
    int gi[6] = {-1, -1, -1, -1, -1, -1};
    auto lidx = [&](int ll, int mm) { return ll * ll + ll + mm; };
    if (act && l > 0) {
        gc[0] = -sqrt(((dl + 1) * (dl + 1) - dm * dm) / (2 * dl + 1) / (2 * dl + 3)) * dl;  // x0: Y(l+1, m)
        gi[0] = lidx(l + 1, m);
        if (am <= l - 1) {
            gc[1] = sqrt((dl * dl - dm * dm) / (2 * dl - 1) / (2 * dl + 1)) * (dl + 1);      // x0: Y(l-1, m)
            gi[1] = lidx(l - 1, m);
        }
        gc[2] = -sqrt((dl + dm + 1) * (dl + dm + 2) / 2 / (2 * dl + 1) / (2 * dl + 3)) * dl;  // xp: Y(l+1, m+1)
        gi[2] = lidx(l + 1, m + 1);
        if (abs(m + 1) <= l - 1) {
            gc[3] = -sqrt((dl - dm - 1) * (dl - dm) / 2 / (2 * dl - 1) / (2 * dl + 1)) * (dl + 1);  // xp: Y(l-1, m+1)
            gi[3] = lidx(l - 1, m + 1);
        }
        gc[4] = -sqrt((dl - dm + 1) * (dl - dm + 2) / 2 / (2 * dl + 1) / (2 * dl + 3)) * dl;  // xm: Y(l+1, m-1)
        gi[4] = lidx(l + 1, m - 1);
        if (abs(m - 1) <= l - 1) {
            gc[5] = -sqrt((dl + dm - 1) * (dl + dm) / 2 / (2 * dl - 1) / (2 * dl + 1)) * (dl + 1);  // xm: Y(l-1, m-1)
            gi[5] = lidx(l - 1, m - 1);
        }
    }
    const double lnorm = act ? a.lnorm[l] : 0.0;
    const double r2 = 1.0 / sqrt(2.0);
    {
        const int64_t e = base + lane;
        const bool ok = lane < nb;
        So3Img mi;
        load_image(a, ok, e, ok ? img_atom[e] : 0, img_j, img_r, atom_z, mi);
        stg[0][lane] = mi.z / mi.R;  // t = cos theta
        stg[1][lane] = mi.x / mi.R;  // z = sin theta e^{i phi}
        stg[2][lane] = mi.y / mi.R;
        stg[3][lane] = 0.0;
        stg[4][lane] = mi.R;
        stg[5][lane] = ok ? mi.w * mi.ex * mi.fc : 0.0;
        stg[6][lane] = ok ? mi.w * mi.ex * (mi.dfc - 2.0 * a.alpha * mi.R * mi.fc) : 0.0;
        stg[7][lane] = mi.rh[0];
        stg[8][lane] = mi.rh[1];
        stg[9][lane] = mi.rh[2];
        s_at[lane] = ok ? mi.i : -1;
    }
    __syncthreads();
    int64_t cur = -1;
    double Ar[NMAX], Ai[NMAX];
#pragma unroll
    for (int n = 0; n < NMAX; n++) {
        Ar[n] = 0.0;
        Ai[n] = 0.0;
    }
    for (int b = 0; b < nb; b++) {
        const int64_t i = s_at[b];
        if (i != cur) {  // uniform: the image's centre changed
            cur = i;
            if (act) {
                const cd *ai = reinterpret_cast<const cd *>(A) + (size_t)i * NMAX * L * M2;
#pragma unroll
                for (int n = 0; n < NMAX; n++) {
                    const cd v = ai[(n * L + l) * M2 + a.lmax + m];
                    Ar[n] = v.re;
                    Ai[n] = v.im;
                }
            }
        }
        // Y_lm of every lane (rows 0..lmax+1)
        const double pl = ylane ? so3_legendre_eval(q, stg[0][b]) : 0.0;
        const cd em = so3_zpow(cmake(stg[1][b], stg[2][b]), am);
        const cd y = cmake(msg * pl * em.re, imsg * pl * em.im);
        ysh[lane] = y;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double c0 = 0.0, c1 = 0.0, c2 = 0.0;
        if (act) {
            const double R = stg[4][b], iR = 1.0 / R;
            const double cv = stg[5][b] * lnorm, cu = stg[6][b] * lnorm;
            const double rh0 = stg[7][b], rh1 = stg[8][b], rh2 = stg[9][b];
            cd yv[6];
#pragma unroll
            for (int q = 0; q < 6; q++) yv[q] = gi[q] >= 0 ? ysh[gi[q]] : cmake(0.0, 0.0);
            const cd x0 = cmake((gc[0] * yv[0].re + gc[1] * yv[1].re) * iR, (gc[0] * yv[0].im + gc[1] * yv[1].im) * iR);
            const cd xp = cmake((gc[2] * yv[2].re + gc[3] * yv[3].re) * iR, (gc[2] * yv[2].im + gc[3] * yv[3].im) * iR);
            const cd xm = cmake((gc[4] * yv[4].re + gc[5] * yv[5].re) * iR, (gc[4] * yv[4].im + gc[5] * yv[5].im) * iR);
            // grad Y: (xm - xp)/sqrt2, i (xm + xp)/sqrt2, x0
            const cd g0 = cmake((xm.re - xp.re) * r2, (xm.im - xp.im) * r2);
            const cd g1 = cmake(-(xm.im + xp.im) * r2, (xm.re + xp.re) * r2);
            const double *ri = rad + ((size_t)(base + b) * L + l) * 2 * NMAX;
            double ar = 0.0, ai2 = 0.0, br = 0.0, bi = 0.0;
#pragma unroll
            for (int n = 0; n < NMAX; n++) {
                const double In = ri[n], dIn = ri[NMAX + n];
                ar = fma(In, Ar[n], ar);
                ai2 = fma(In, Ai[n], ai2);
                br = fma(dIn, Ar[n], br);
                bi = fma(dIn, Ai[n], bi);
            }
            // Re(conj(alpha) z) = ar z.re + ai z.im
            const double ray = ar * y.re + ai2 * y.im;
            const double rby = br * y.re + bi * y.im;
            const double s = cu * ray + cv * rby;
            c0 = fma(s, rh0, cv * (ar * g0.re + ai2 * g0.im));
            c1 = fma(s, rh1, cv * (ar * g1.re + ai2 * g1.im));
            c2 = fma(s, rh2, cv * (ar * x0.re + ai2 * x0.im));
        }
        wave_sum3_lane63(c0, c1, c2);
        if (lane == 63) {
            const int64_t e = base + b;
            dedr[3 * e] = c0;
            dedr[3 * e + 1] = c1;
            dedr[3 * e + 2] = c2;
        }
        // the next image rewrites the Y table
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ------------------------------------------------ per-image (flat) kernels
// MODE SO3_ADJOINT: dE/dr_e,k = Re sum_{n,l,m} dc_e,nlm,k conj(A_i,nlm)
// MODE SO3_PLAIN:   dP_e[tril(a,b)][l][k] = Q_k[a][b] + Q_k[b][a],
//                   Q_k[n][n'] = Re sum_m dc_e,nlm,k conj(ctot_i,n'lm)
// (SO3.py:243-247).  One wavefront per NG images; lanes (image, l).
template <int NMAX, int MODE>
__global__ __launch_bounds__(64) void so3_image_kernel(So3Args a, int64_t nimg, const int64_t *__restrict__ img_atom,
                                                        const int64_t *__restrict__ img_j,
                                                        const double *__restrict__ img_r,
                                                        const int64_t *__restrict__ atom_z,
                                                        const double *__restrict__ cin, double *__restrict__ out,
                                                        const double *__restrict__ rad_in,
                                                        const int64_t *__restrict__ d_count) {
    extern __shared__ double smem[];
    // with d_count the grid covers an image capacity and the images in use
    // are counted on the device
    if (d_count) {
        const int64_t c = *d_count;
        nimg = c < nimg ? c : nimg;
    }
    const int L = a.lmax + 1;
    const int LP = a.lmax + 1;
    const int NG = 64 / L;
    const int M2 = 2 * a.lmax + 1;
    const int ysz = (a.lmax + 2) * (2 * LP + 1);
    const int csz = NMAX * L * M2;
    double *lnorm = smem;
    double *red = lnorm + L;  // [64][3]
    cd *Yt = reinterpret_cast<cd *>(red + 64 * 3 + (L & 1));
    const int lane = threadIdx.x;
    const int g = lane / L, l = lane % L;
    for (int k = lane; k < L; k += 64) lnorm[k] = a.lnorm[k];
    So3Args al = a;
    al.lnorm = lnorm;
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * NG + g;
    const bool valid = (g < NG) && (e < nimg);
    So3Img m;
    load_image(al, valid, e, valid ? img_atom[e] : 0, img_j, img_r, atom_z, m);
    double I[NMAX], dI[NMAX];
    if (valid) {
        if (rad_in) {  // radial integrals stored by the forward expansion
            const double *ri = rad_in + ((size_t)e * L + l) * 2 * NMAX;
#pragma unroll
            for (int n = 0; n < NMAX; n++) {
                I[n] = ri[n];
                dI[n] = ri[NMAX + n];
            }
        } else {
            radial<NMAX>(al, l, m.R, a.G, a.rq, I, dI, true);
        }
        fill_ylm(m, l, L, a.lmax, Yt + g * ysz, LP);
    }
    __syncthreads();
    double r3[3] = {0, 0, 0};
    if (valid) {
        const cd *Y = Yt + g * ysz;
        const cd *ci = reinterpret_cast<const cd *>(cin) + (size_t)m.i * csz;
        // alpha_k[n'] = Re sum_m U_km conj(C_n'lm), beta_k[n'] = Re sum_m V_km conj(C_n'lm)
        double al_[3][NMAX], be_[3][NMAX];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int n = 0; n < NMAX; n++) {
                al_[k][n] = 0.0;
                be_[k][n] = 0.0;
            }
        for (int mm = -l; mm <= l; mm++) {
            cd P, U[3], V[3];
            image_factors(al, m, l, mm, Y, LP, P, U, V);
#pragma unroll
            for (int n = 0; n < NMAX; n++) {
                const cd cv = ci[(n * L + l) * M2 + a.lmax + mm];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    al_[k][n] += cdotre(cv, U[k]);
                    be_[k][n] += cdotre(cv, V[k]);
                }
            }
        }
        if (MODE == SO3_ADJOINT) {
            // cin = A: dE/dr_k = sum_n I_n alpha_k[n] + dI_n beta_k[n]
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
#pragma unroll
                for (int n = 0; n < NMAX; n++) s += I[n] * al_[k][n] + dI[n] * be_[k][n];
                r3[k] = s;
            }
        } else {
            // cin = ctot: write dP of this l for every tril pair
            const int ntri = NMAX * (NMAX + 1) / 2;
            double *o = out + (size_t)e * ntri * L * 3;
            int t = 0;
#pragma unroll
            for (int aa = 0; aa < NMAX; aa++)
#pragma unroll
                for (int bb = 0; bb <= aa; bb++, t++)
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        o[(t * L + l) * 3 + k] = I[aa] * al_[k][bb] + dI[aa] * be_[k][bb] + I[bb] * al_[k][aa] +
                                                 dI[bb] * be_[k][aa];
        }
    }
    if (MODE == SO3_ADJOINT) {
        red[3 * lane] = r3[0];
        red[3 * lane + 1] = r3[1];
        red[3 * lane + 2] = r3[2];
        __syncthreads();
        if (valid && l == 0) {
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
                for (int q = 0; q < L; q++) s += red[3 * (lane + q) + k];
                out[3 * e + k] = s;
            }
        }
    }
}

// P[i][tril(a,b)*L + l] = Re sum_m C[a][l][m] conj(C[b][l][m])   (SO3.py:228,259)
// with mode tangent: Re sum_m dC[a] conj(C[b]) + C[a] conj(dC[b]).
__global__ void so3_power_kernel(int64_t natoms, int nmax, int lmax, const double *__restrict__ ct,
                                 const double *__restrict__ dct, double *__restrict__ x) {
    const int L = lmax + 1, M2 = 2 * lmax + 1;
    const int ntri = nmax * (nmax + 1) / 2;
    const int nc = ntri * L;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= natoms * nc) return;
    const int64_t i = t / nc;
    const int q = (int)(t - i * nc);
    const int tri = q / L, l = q - tri * L;
    int aa = 0;
    while ((aa + 1) * (aa + 2) / 2 <= tri) aa++;
    const int bb = tri - aa * (aa + 1) / 2;
    const int csz = nmax * L * M2;
    const cd *c = reinterpret_cast<const cd *>(ct) + (size_t)i * csz;
    const cd *ca = c + (aa * L + l) * M2 + lmax, *cb = c + (bb * L + l) * M2 + lmax;
    double s = 0.0;
    if (!dct) {
        for (int mm = -l; mm <= l; mm++) s += cdotre(cb[mm], ca[mm]);
    } else {
        const cd *d = reinterpret_cast<const cd *>(dct) + (size_t)i * csz;
        const cd *da = d + (aa * L + l) * M2 + lmax, *db = d + (bb * L + l) * M2 + lmax;
        for (int mm = -l; mm <= l; mm++) s += cdotre(cb[mm], da[mm]) + cdotre(db[mm], ca[mm]);
    }
    x[t] = s;
}

// A[i][n][l][m] = sum_n' S[n][n'][l] C[i][n'][l][m], S from dE/dP with the
// diagonal doubled (P is quadratic in C).  One thread per (centre, l, m)
// forms all NMAX outputs from the column C[i][.][l][m] read once into
// registers: every C entry is read by one thread (with a thread per output,
// a centre's entries straddle workgroups that run on different XCDs, and
// each XCD fetched them from HBM through its own L2).  Loads and stores run
// along (l, m), consecutive across the threads.
template <int NMAX>
__global__ __launch_bounds__(256) void so3_adj_a_kernel(int64_t natoms, int lmax, const double *__restrict__ ct,
                                                         const double *__restrict__ beta, double *__restrict__ A) {
    constexpr int NTRI = NMAX * (NMAX + 1) / 2;
    const int L = lmax + 1, M2 = 2 * lmax + 1, LM = L * M2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= natoms * LM) return;
    const int64_t i = t / LM;
    const int k = (int)(t - i * LM);
    const int l = k / M2, mi = k - l * M2;
    const bool in = mi - lmax >= -l && mi - lmax <= l;
    const cd *c = reinterpret_cast<const cd *>(ct) + (size_t)i * NMAX * LM + k;
    cd *o = reinterpret_cast<cd *>(A) + (size_t)i * NMAX * LM + k;
    cd cv[NMAX];
#pragma unroll
    for (int np = 0; np < NMAX; np++) cv[np] = in ? c[(size_t)np * LM] : cmake(0.0, 0.0);
    const double *b = beta + (size_t)i * NTRI * L + l;
#pragma unroll
    for (int n = 0; n < NMAX; n++) {
        cd s = cmake(0, 0);
#pragma unroll
        for (int np = 0; np < NMAX; np++) {
            const int hi = n > np ? n : np, lo = n > np ? np : n;
            double f = b[(hi * (hi + 1) / 2 + lo) * L];
            if (np == n) f *= 2.0;
            s = cadd(s, cscale(cv[np], f));
        }
        o[(size_t)n * LM] = s;
    }
}

}  // namespace pxff

// =================================================================== C ABI
using namespace pxff;

namespace {
int so3_series_init();

// The parameter checks alone (host only: the size queries use them).
static int so3_check_params(const pxff_so3_params *p) {
    PXFF_REQUIRE(p != nullptr, "null params");
    if (p->nmax < 1 || p->nmax > 11 || p->lmax < 0 || p->lmax > SO3_LMAXC) {
        set_error("SO3: nmax must be in 1..11 and lmax in 0..32");
        return PXFF_EINVAL;
    }
    PXFF_REQUIRE(p->rcut > 0 && p->alpha > 0 && p->nquad > 0, "SO3: rcut, alpha, nquad must be > 0");
    PXFF_REQUIRE(p->cutoff_id >= 1 && p->cutoff_id <= 8, "SO3: cutoff id must be 1..8");
    const int L = p->lmax + 1;
    if (p->nmax * L * (2 * p->lmax + 1) > 4096) {
        set_error("SO3: nmax*(lmax+1)*(2lmax+1) > 4096 is not supported on the device path");
        return PXFF_ENOTIMPL;
    }
    return 0;
}

int so3_check(const pxff_so3_params *p) {
    const int rc = so3_check_params(p);
    return rc ? rc : so3_series_init();
}

So3Args make_args(const pxff_so3_params *p, const double *G, const double *rq, const double *lnorm) {
    So3Args a;
    a.nmax = p->nmax;
    a.lmax = p->lmax;
    a.nq = p->nquad;
    a.cutoff_id = p->cutoff_id;
    a.weight_on = p->weight_on;
    a.rc = p->rcut;
    a.alpha = p->alpha;
    a.G = G;
    a.rq = rq;
    a.lnorm = lnorm;
    return a;
}

size_t accum_lds(const So3Args &a, int nmax, bool tangent) {
    const int L = a.lmax + 1, NG = 64 / L, M2 = 2 * a.lmax + 1, LP = L;
    const int ysz = (a.lmax + 2) * (2 * LP + 1);
    size_t d = (size_t)L + (tangent ? 2 : 1) * (size_t)NG * nmax * L + 1;
    d += 2 * ((size_t)NG * ysz + (tangent ? 2 : 1) * (size_t)NG * L * M2 + (size_t)nmax * L * M2);
    d += (size_t)nmax * L * L;  // the valid-entry table: two ints per entry
    return d * sizeof(double);
}

size_t image_lds(const So3Args &a, int nmax) {
    const int L = a.lmax + 1, NG = 64 / L, LP = L;
    const int ysz = (a.lmax + 2) * (2 * LP + 1);
    (void)nmax;
    size_t d = (size_t)L + 64 * 3 + 1 + 2 * (size_t)NG * ysz;
    return d * sizeof(double);
}

// Fill the series reciprocal tables of the current device (once).
std::mutex g_series_mutex;
std::set<int> g_series_done;
int so3_series_init() {
    int dev = 0;
    PXFF_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_series_mutex);
    if (g_series_done.count(dev)) return 0;
    std::vector<double> rk((SO3_LMAXC + 2) * SO3_KMAX, 0.0), rodd(SO3_LMAXC + 3, 0.0);
    for (int n = 0; n < SO3_LMAXC + 2; n++)
        for (int k = 1; k < SO3_KMAX; k++) rk[n * SO3_KMAX + k] = 1.0 / (k * (2.0 * n + 2 * k + 1));
    for (int k = 0; k < SO3_LMAXC + 3; k++) rodd[k] = 1.0 / (2 * k + 1);
    PXFF_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_so3_rk), rk.data(), rk.size() * sizeof(double)));
    PXFF_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_so3_rodd), rodd.data(), rodd.size() * sizeof(double)));
    std::vector<double> la(SO3_LEGD * SO3_LEGD, 0.0), lb(SO3_LEGD * SO3_LEGD, 0.0), pm(SO3_LEGD, 0.0),
        p1(SO3_LEGD, 0.0);
    for (int k = 2; k < SO3_LEGD; k++)
        for (int m = 0; m + 2 <= k; m++) {
            la[k * SO3_LEGD + m] = std::sqrt((4.0 * k * k - 1.0) / ((double)k * k - (double)m * m));
            lb[k * SO3_LEGD + m] =
                std::sqrt(((k - 1.0) * (k - 1.0) - (double)m * m) / (4.0 * (k - 1.0) * (k - 1.0) - 1.0));
        }
    for (int m = 0; m < SO3_LEGD; m++) {
        if (m > 0) pm[m] = -std::sqrt((2.0 * m + 1.0) / (2.0 * m));
        p1[m] = std::sqrt(2.0 * m + 3.0);
    }
    PXFF_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_so3_lega), la.data(), la.size() * sizeof(double)));
    PXFF_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_so3_legb), lb.data(), lb.size() * sizeof(double)));
    PXFF_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_so3_pmm), pm.data(), pm.size() * sizeof(double)));
    PXFF_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_so3_p1), p1.data(), p1.size() * sizeof(double)));
    g_series_done.insert(dev);
    return 0;
}

int g_cus = 0;
int cus() {
    if (!g_cus) {
        int dev = 0;
        hipDeviceProp_t pr;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&pr, dev) != hipSuccess) return 256;
        g_cus = pr.multiProcessorCount;
    }
    return g_cus;
}
}  // namespace

#define SO3_DISPATCH_NMAX(nm, ...)                          \
    switch (nm) {                                           \
    case 1: { constexpr int NM = 1; __VA_ARGS__; } break;   \
    case 2: { constexpr int NM = 2; __VA_ARGS__; } break;   \
    case 3: { constexpr int NM = 3; __VA_ARGS__; } break;   \
    case 4: { constexpr int NM = 4; __VA_ARGS__; } break;   \
    case 5: { constexpr int NM = 5; __VA_ARGS__; } break;   \
    case 6: { constexpr int NM = 6; __VA_ARGS__; } break;   \
    case 7: { constexpr int NM = 7; __VA_ARGS__; } break;   \
    case 8: { constexpr int NM = 8; __VA_ARGS__; } break;   \
    case 9: { constexpr int NM = 9; __VA_ARGS__; } break;   \
    case 10: { constexpr int NM = 10; __VA_ARGS__; } break; \
    case 11: { constexpr int NM = 11; __VA_ARGS__; } break; \
    default: set_error("SO3: nmax out of range"); return PXFF_EINVAL; \
    }

template <typename K>
static int launch_set_lds(K kern, size_t lds) {
    if (lds > 65536)
        PXFF_HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return 0;
}

static int so3_ctot_impl(const pxff_so3_params *p, int64_t natoms, const int64_t *d_img_start,
                         const int64_t *d_img_j, const double *d_img_r, const int64_t *d_img_rev,
                         const int64_t *d_atom_z, const double *d_G, const double *d_rq, const double *d_lnorm,
                         const double *d_dr, double *d_out, double *d_radial, void *stream, int64_t *d_work = nullptr,
                         int64_t cap = 0, const double *d_table = nullptr) {
    int rc = so3_check(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    PXFF_REQUIRE(!(d_dr && d_radial), "SO3: the radial integrals are stored by the plain expansion only");
    So3Args a = make_args(p, d_G, d_rq, d_lnorm);
    hipStream_t s = (hipStream_t)stream;
    const size_t lds = accum_lds(a, p->nmax, d_dr != nullptr);
    PXFF_REQUIRE(lds <= 160 * 1024, "SO3: LDS need exceeds 160 KB");
    const int grid = (int)std::min<int64_t>(natoms, (int64_t)cus() * 8);
    if (d_dr) {
        SO3_DISPATCH_NMAX(p->nmax, {
            auto k = so3_accum_kernel<NM, SO3_TANGENT>;
            if ((rc = launch_set_lds(k, lds))) return rc;
            k<<<grid, 64, lds, s>>>(a, natoms, d_img_start, d_img_j, d_img_r, d_atom_z, d_dr, d_out, nullptr);
        });
    } else {
        if (d_radial) {
            if (p->nmax <= 6 && p->lmax <= 6) {
                // every order of an image from one Bessel evaluation per node
                // (i_0..i_6: l <= 6, and i_1 for l = 0)
                // with pairs, one resident round of wavefronts (2 per SIMD at
                // 256 registers), each queueing several chunks of 64 images so
                // that its batches of 64 qualifying images are full
                if (d_img_rev && d_work && cap > 0) {
                    // pair images sorted into |r| bins, then one image per lane
                    so3_rbin_hist_kernel<<<SO3_RBLK, SO3_RBT, 0, s>>>(natoms, d_img_start, d_img_r, d_img_rev, p->rcut,
                                                                  d_work);
                    so3_rbin_scan_kernel<<<1, SO3_RSCAN_T, 0, s>>>(d_work);
                    so3_rbin_fill_kernel<<<SO3_RBLK, SO3_RBT, 0, s>>>(natoms, d_img_start, d_img_r, d_img_rev, p->rcut,
                                                                  d_work);
                    if (d_table) {
                        // from the tabulated integrals (pxff_so3_radial_table)
                        const int nk = so3_rtab_nk(p);
                        const double h = p->rcut / nk;
                        SO3_DISPATCH_NMAX(p->nmax, if constexpr (NM <= 6) {
                            so3_radial_tab_kernel<NM, 7><<<div_up(cap, 64), 64, 0, s>>>(
                                a, d_work, d_img_r, d_img_rev, d_table, nk, h, d_radial);
                        });
                    } else {
                        SO3_DISPATCH_NMAX(p->nmax, if constexpr (NM <= 6) {
                            so3_radial_sorted_kernel<NM, 7><<<div_up(cap, 64), 64, 0, s>>>(a, d_work, d_img_r,
                                                                                          d_img_rev, d_radial);
                        });
                    }
                } else {
                    const int grid = d_img_rev ? cus() * 2 : cus() * 16;
                    SO3_DISPATCH_NMAX(p->nmax, if constexpr (NM <= 6) {
                        so3_radial_all_kernel<NM, 7><<<grid, 256, 0, s>>>(a, natoms, d_img_start, d_img_r,
                                                                         d_img_rev, d_radial);
                    });
                }
            } else {
                SO3_DISPATCH_NMAX(p->nmax, (so3_radial_kernel<NM><<<cus() * 16, 256, 0, s>>>(
                                               a, natoms, d_img_start, d_img_r, d_radial)));
            }
            PXFF_LAUNCH_CHECK();
        }
        if (d_radial && p->lmax <= 7) {
            // one lane per (l, m), the n-sums in registers
            SO3_DISPATCH_NMAX(p->nmax, (so3_ctot_lm_kernel<NM><<<grid, 64, 0, s>>>(
                                           a, natoms, d_img_start, d_img_j, d_img_r, d_atom_z, d_radial, d_out)));
        } else {
            SO3_DISPATCH_NMAX(p->nmax, {
                auto k = so3_accum_kernel<NM, SO3_CTOT>;
                if ((rc = launch_set_lds(k, lds))) return rc;
                k<<<grid, 64, lds, s>>>(a, natoms, d_img_start, d_img_j, d_img_r, d_atom_z, nullptr, d_out, d_radial);
            });
        }
    }
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so3_ctot(const pxff_so3_params *p, int64_t natoms, const int64_t *d_img_start,
                             const int64_t *d_img_j, const double *d_img_r, const int64_t *d_atom_z,
                             const double *d_G, const double *d_rq, const double *d_lnorm, const double *d_dr,
                             double *d_out, double *d_radial, void *stream) {
    return so3_ctot_impl(p, natoms, d_img_start, d_img_j, d_img_r, nullptr, d_atom_z, d_G, d_rq, d_lnorm, d_dr,
                         d_out, d_radial, stream);
}

extern "C" int pxff_so3_ctot_pairs(const pxff_so3_params *p, int64_t natoms, const int64_t *d_img_start,
                                   const int64_t *d_img_j, const double *d_img_r, const int64_t *d_img_rev,
                                   const int64_t *d_atom_z, const double *d_G, const double *d_rq,
                                   const double *d_lnorm, double *d_out, double *d_radial, int64_t *d_work,
                                   int64_t cap, const double *d_table, void *stream) {
    PXFF_REQUIRE(d_img_rev != nullptr && d_radial != nullptr, "SO3: pairs need the reverse images and the radial store");
    PXFF_REQUIRE(cap >= 0 && (d_work == nullptr || cap > 0), "SO3: the |r| ordering needs the image capacity");
    PXFF_REQUIRE(d_table == nullptr || (d_work != nullptr && p != nullptr && p->nmax <= 6 && p->lmax <= 6),
                 "SO3: the radial table needs the |r| ordering workspace and nmax, lmax <= 6");
    return so3_ctot_impl(p, natoms, d_img_start, d_img_j, d_img_r, d_img_rev, d_atom_z, d_G, d_rq, d_lnorm, nullptr,
                         d_out, d_radial, stream, d_work, cap, d_table);
}

extern "C" int pxff_so3_radial_table_size(const pxff_so3_params *p, int *nk, int64_t *len) {
    int rc = so3_check_params(p);
    if (rc) return rc;
    PXFF_REQUIRE(nk != nullptr && len != nullptr, "null output");
    PXFF_REQUIRE(p->nmax <= 6 && p->lmax <= 6, "SO3: the radial table is for nmax, lmax <= 6");
    *nk = so3_rtab_nk(p);
    *len = (int64_t)(*nk) * (p->lmax + 1) * 2 * p->nmax * SO3_TNP;
    return 0;
}

extern "C" int pxff_so3_radial_table(const pxff_so3_params *p, const double *d_G, const double *d_rq,
                                     const double *d_lnorm, double *d_scratch, double *d_table, void *stream) {
    int rc = so3_check(p);
    if (rc) return rc;
    PXFF_REQUIRE(p->nmax <= 6 && p->lmax <= 6, "SO3: the radial table is for nmax, lmax <= 6");
    PXFF_REQUIRE(d_scratch != nullptr && d_table != nullptr, "null table or scratch");
    So3Args a = make_args(p, d_G, d_rq, d_lnorm);
    hipStream_t s = (hipStream_t)stream;
    const int nk = so3_rtab_nk(p);
    const int nf = (p->lmax + 1) * 2 * p->nmax;
    const double h = p->rcut / nk;
    SO3_DISPATCH_NMAX(p->nmax, if constexpr (NM <= 6) {
        so3_rtab_nodes_kernel<NM, 7><<<div_up(nk * SO3_TNP, 256), 256, 0, s>>>(a, nk, h, d_scratch);
    });
    PXFF_LAUNCH_CHECK();
    static const So3TabW W = so3_rtab_weights();
    so3_rtab_coef_kernel<<<div_up(nk * nf, 256), 256, 0, s>>>(W, nk, nf, d_scratch, d_table);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so3_power(const pxff_so3_params *p, int64_t natoms, const double *d_ctot, const double *d_dctot,
                              double *d_x, void *stream) {
    int rc = so3_check(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    const int64_t nc = (int64_t)p->nmax * (p->nmax + 1) / 2 * (p->lmax + 1);
    so3_power_kernel<<<div_up(natoms * nc, 256), 256, 0, (hipStream_t)stream>>>(natoms, p->nmax, p->lmax, d_ctot,
                                                                                d_dctot, d_x);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so3_adjoint_a(const pxff_so3_params *p, int64_t natoms, const double *d_ctot,
                                  const double *d_beta, double *d_a, void *stream) {
    int rc = so3_check(p);
    if (rc) return rc;
    if (natoms <= 0) return 0;
    // one thread per (centre, l, m)
    const int64_t nlm = (int64_t)(p->lmax + 1) * (2 * p->lmax + 1);
    hipStream_t s = (hipStream_t)stream;
    SO3_DISPATCH_NMAX(p->nmax, (so3_adj_a_kernel<NM><<<div_up(natoms * nlm, 256), 256, 0, s>>>(natoms, p->lmax, d_ctot,
                                                                                            d_beta, d_a)));
    PXFF_LAUNCH_CHECK();
    return 0;
}

static int so3_image_launch(const pxff_so3_params *p, int mode, int64_t nimg, const int64_t *d_img_atom,
                            const int64_t *d_img_j, const double *d_img_r, const int64_t *d_atom_z, const double *d_G,
                            const double *d_rq, const double *d_lnorm, const double *d_cin, double *d_out,
                            const double *d_radial, void *stream, const int64_t *d_count = nullptr) {
    int rc = so3_check(p);
    if (rc) return rc;
    if (nimg <= 0) return 0;
    So3Args a = make_args(p, d_G, d_rq, d_lnorm);
    hipStream_t s = (hipStream_t)stream;
    const size_t lds = image_lds(a, p->nmax);
    PXFF_REQUIRE(lds <= 160 * 1024, "SO3: LDS need exceeds 160 KB");
    const int NG = 64 / (p->lmax + 1);
    const int grid = div_up(nimg, NG);
    if (mode == SO3_ADJOINT && d_radial && p->lmax <= 6) {
        // one lane per (l, m), 64 images per wavefront
        SO3_DISPATCH_NMAX(p->nmax, (so3_deidrj_lm_kernel<NM><<<div_up(nimg, 64), 64, 0, s>>>(
                                       a, nimg, d_img_atom, d_img_j, d_img_r, d_atom_z, d_cin, d_radial, d_out,
                                       d_count)));
    } else if (mode == SO3_ADJOINT) {
        SO3_DISPATCH_NMAX(p->nmax, {
            auto k = so3_image_kernel<NM, SO3_ADJOINT>;
            if ((rc = launch_set_lds(k, lds))) return rc;
            k<<<grid, 64, lds, s>>>(a, nimg, d_img_atom, d_img_j, d_img_r, d_atom_z, d_cin, d_out, d_radial, d_count);
        });
    } else {
        SO3_DISPATCH_NMAX(p->nmax, {
            auto k = so3_image_kernel<NM, SO3_PLAIN>;
            if ((rc = launch_set_lds(k, lds))) return rc;
            k<<<grid, 64, lds, s>>>(a, nimg, d_img_atom, d_img_j, d_img_r, d_atom_z, d_cin, d_out, d_radial, d_count);
        });
    }
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_so3_deidrj(const pxff_so3_params *p, int64_t nimg, const int64_t *d_img_atom,
                               const int64_t *d_img_j, const double *d_img_r, const int64_t *d_atom_z,
                               const double *d_G, const double *d_rq, const double *d_lnorm, const double *d_a,
                               const double *d_radial, double *d_dedr, void *stream) {
    return so3_image_launch(p, SO3_ADJOINT, nimg, d_img_atom, d_img_j, d_img_r, d_atom_z, d_G, d_rq, d_lnorm, d_a,
                            d_dedr, d_radial, stream);
}

extern "C" int pxff_so3_deidrj_counted(const pxff_so3_params *p, int64_t cap, const int64_t *d_count,
                                       const int64_t *d_img_atom, const int64_t *d_img_j, const double *d_img_r,
                                       const int64_t *d_atom_z, const double *d_G, const double *d_rq,
                                       const double *d_lnorm, const double *d_a, const double *d_radial,
                                       double *d_dedr, void *stream) {
    PXFF_REQUIRE(d_count != nullptr, "null image count");
    return so3_image_launch(p, SO3_ADJOINT, cap, d_img_atom, d_img_j, d_img_r, d_atom_z, d_G, d_rq, d_lnorm, d_a,
                            d_dedr, d_radial, stream, d_count);
}

extern "C" int pxff_so3_dpdr(const pxff_so3_params *p, int64_t nimg, const int64_t *d_img_atom,
                             const int64_t *d_img_j, const double *d_img_r, const int64_t *d_atom_z,
                             const double *d_G, const double *d_rq, const double *d_lnorm, const double *d_ctot,
                             const double *d_radial, double *d_dpdr, void *stream) {
    return so3_image_launch(p, SO3_PLAIN, nimg, d_img_atom, d_img_j, d_img_r, d_atom_z, d_G, d_rq, d_lnorm, d_ctot,
                            d_dpdr, d_radial, stream);
}
