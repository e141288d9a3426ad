// Neighbour list for the descriptor hot path on gfx950.
//
// Replaces the ASE NeighborList call and the host regrouping of the
// reference (pyxtal_ff/descriptors/SNAP.py:240-317, SO3.py:317-376): every
// image (i, j, T) with |pos_j + T.cell - pos_i| < rc except (i, T=0), for a
// batch of structures at once.
//
// Algorithm: each structure's atoms are wrapped into the cell along its
// periodic directions (fractional coordinate f = s - floor(s)) and binned on
// an nb0 x nb1 x nb2 grid (nb_d = floor(h_d / rc), h_d = plane spacing).  A
// centre scans the (2S+1)^3 stencil of bins around its own, S_d =
// ceil(rc nb_d / h_d); a stencil offset that leaves the grid wraps into the
// next periodic image, so small cells (nb = 1, S = 1 or 2) are handled by the
// same code -- the stencil then enumerates lattice translations directly.
// Candidate separations are computed from the raw positions in a canonical
// orientation with no FMA contraction (see canon_sep), so the image set is
// bit-identical to the CPU oracle's brute-force enumeration and exactly
// symmetric (r_ji = -r_ij), which the force gather relies on.
#include <hip/hip_runtime.h>

#include "pxff_common.h"

namespace pxff {

__device__ inline int floordiv_i(int a, int b) {
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

// Separation of image (i, j, T) computed as
//   r = (pos_j + ((T0 c0 + T1 c1) + T2 c2)) - pos_i       if (i,j,T) is "forward"
//   r = -((pos_i + ((-T0 c0 - T1 c1) - T2 c2)) - pos_j)   otherwise,
// forward meaning j > i, or j == i and T > 0 lexicographically.  All
// operations are individually rounded (no contraction).  pi, pj: the two
// positions, cl: the cell (rows = lattice vectors).
__device__ inline void canon_sep_v(const double *pi, const double *pj, const double *cl, int64_t i, int64_t j, int T0,
                                   int T1, int T2, double r[3]) {
    bool fwd;
    if (j != i)
        fwd = j > i;
    else
        fwd = (T0 > 0) || (T0 == 0 && T1 > 0) || (T0 == 0 && T1 == 0 && T2 > 0);
    const int s = fwd ? 1 : -1;
    const double t0 = (double)(s * T0), t1 = (double)(s * T1), t2 = (double)(s * T2);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double t = __dmul_rn(t0, cl[k]);
        t = __dadd_rn(t, __dmul_rn(t1, cl[3 + k]));
        t = __dadd_rn(t, __dmul_rn(t2, cl[6 + k]));
        const double pa = fwd ? pj[k] : pi[k], pb = fwd ? pi[k] : pj[k];
        double p = __dadd_rn(pa, t);
        double v = __dsub_rn(p, pb);
        r[k] = fwd ? v : -v;
    }
}

__device__ inline void canon_sep(const double *__restrict__ pos, const double *__restrict__ cell, int64_t i, int64_t j,
                                 int T0, int T1, int T2, double r[3]) {
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double pj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    canon_sep_v(pi, pj, cell, i, j, T0, T1, T2, r);
}

__device__ inline double dist2(const double r[3]) {
    return __dadd_rn(__dadd_rn(__dmul_rn(r[0], r[0]), __dmul_rn(r[1], r[1])), __dmul_rn(r[2], r[2]));
}

// ---------------------------------------------------------------- binning
// Bin of atom a (global bin index) and its lattice wrap, stored to
// atom_bin[a] / wrap[3a..3a+2].
__device__ inline int nl_bin_of(int64_t a, const double *__restrict__ pos, const int64_t *__restrict__ atom_struct,
                                const double *__restrict__ inv_cells, const int32_t *__restrict__ pbc,
                                const int32_t *__restrict__ geom, int32_t *__restrict__ atom_bin,
                                int32_t *__restrict__ wrap) {
    const int64_t s = atom_struct[a];
    const double *inv = inv_cells + 9 * s;
    const int32_t *gm = geom + 8 * s;
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double f = pos[3 * a] * inv[d] + pos[3 * a + 1] * inv[3 + d] + pos[3 * a + 2] * inv[6 + d];
        int k = 0, bd = 0;
        if (pbc[3 * s + d]) {
            const double fl = floor(f);
            k = (int)fl;
            double fr = f - fl;
            bd = (int)(fr * gm[d]);
            if (bd >= gm[d]) bd = gm[d] - 1;
            if (bd < 0) bd = 0;
        }
        wrap[3 * a + d] = k;
        b[d] = bd;
    }
    const int gb = gm[6] + (b[0] * gm[1] + b[1]) * gm[2] + b[2];
    atom_bin[a] = gb;
    return gb;
}

__global__ void nl_bin_kernel(int64_t natoms, const double *__restrict__ pos, const int64_t *__restrict__ atom_struct,
                              const double *__restrict__ inv_cells, const int32_t *__restrict__ pbc,
                              const int32_t *__restrict__ geom, int32_t *__restrict__ atom_bin,
                              int32_t *__restrict__ wrap, int32_t *__restrict__ bin_count) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    atomicAdd(&bin_count[nl_bin_of(a, pos, atom_struct, inv_cells, pbc, geom, atom_bin, wrap)], 1);
}

// Bin and lattice wrap of every atom (no counting): the parallel half of
// the small-list binning.
__global__ void nl_bin_place_kernel(int64_t natoms, const double *__restrict__ pos,
                                    const int64_t *__restrict__ atom_struct, const double *__restrict__ inv_cells,
                                    const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom,
                                    int32_t *__restrict__ atom_bin, int32_t *__restrict__ wrap) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    nl_bin_of(a, pos, atom_struct, inv_cells, pbc, geom, atom_bin, wrap);
}

// The counting half of a small list's binning in one workgroup of 1024
// threads, with the bin counters in LDS: bin sizes from the atoms' bins,
// their exclusive scan (bin_start, nbins + 1 entries) and the atoms of
// every bin -- two launches instead of five.  The order of the atoms
// inside a bin is not fixed (LDS atomics), as with the multi-kernel path;
// the image lists do not depend on it (their fill ranks images by (j, T)).
constexpr int NL_SMALL_BINS = 8192;
constexpr int NL_SMALL_ATOMS = 1 << 16;
__global__ __launch_bounds__(1024) void nl_bin_small_kernel(int64_t natoms, int nbins,
                                                             const int32_t *__restrict__ atom_bin,
                                                             int32_t *__restrict__ bin_start,
                                                             int32_t *__restrict__ bin_atoms) {
    __shared__ int cnt[NL_SMALL_BINS];
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int k = tid; k < nbins; k += 1024) cnt[k] = 0;
    __syncthreads();
    // four atoms per thread per step, their loads in flight together
    for (int64_t a0 = tid; a0 < natoms; a0 += 4 * 1024) {
        int gb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) gb[u] = a0 + u * 1024 < natoms ? atom_bin[a0 + u * 1024] : -1;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (gb[u] >= 0) atomicAdd(&cnt[gb[u]], 1);
    }
    __syncthreads();
    // exclusive scan: a contiguous chunk of bins per thread
    const int chunk = (nbins + 1023) / 1024;
    const int b0 = tid * chunk, b1 = b0 + chunk < nbins ? b0 + chunk : nbins;
    int sum = 0;
    for (int k = b0; k < b1; k++) sum += cnt[k];
    int x = sum;  // inclusive scan of the chunk sums within the wavefront
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int run = x - sum, total = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int t = wsum[q];
        if (q < w) run += t;
        total += t;
    }
    for (int k = b0; k < b1; k++) {
        const int c = cnt[k];
        bin_start[k] = run;
        cnt[k] = run;  // from here on the insertion cursor of bin k
        run += c;
    }
    if (tid == 0) bin_start[nbins] = total;
    __syncthreads();
    for (int64_t a0 = tid; a0 < natoms; a0 += 4 * 1024) {
        int gb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) gb[u] = a0 + u * 1024 < natoms ? atom_bin[a0 + u * 1024] : -1;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (gb[u] >= 0) bin_atoms[atomicAdd(&cnt[gb[u]], 1)] = (int32_t)(a0 + u * 1024);
    }
}

__global__ void nl_bin_fill_kernel(int64_t natoms, const int32_t *__restrict__ atom_bin,
                                   const int32_t *__restrict__ bin_start, int32_t *__restrict__ cursor,
                                   int32_t *__restrict__ bin_atoms) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms) return;
    const int gb = atom_bin[a];
    const int slot = atomicAdd(&cursor[gb], 1);
    bin_atoms[bin_start[gb] + slot] = (int32_t)a;
}

// Enumerate the candidate images of centre i; calls f(j, T0, T1, T2, r) for
// every image inside the cutoff.
template <typename F>
__device__ inline void nl_enumerate(int64_t i, const double *__restrict__ pos, const int64_t *__restrict__ atom_struct,
                                    const double *__restrict__ cells, const int32_t *__restrict__ pbc,
                                    const int32_t *__restrict__ geom, double rc2, const int32_t *__restrict__ bin_start,
                                    const int32_t *__restrict__ atom_bin, const int32_t *__restrict__ wrap,
                                    const int32_t *__restrict__ bin_atoms, F &&f) {
    const int64_t s = atom_struct[i];
    const double *cell = cells + 9 * s;
    const int32_t *gm = geom + 8 * s;
    const int nb0 = gm[0], nb1 = gm[1], nb2 = gm[2];
    const int S0 = gm[3], S1 = gm[4], S2 = gm[5];
    const int off = gm[6];
    const int lb = atom_bin[i] - off;
    const int bi0 = lb / (nb1 * nb2), bi1 = (lb / nb2) % nb1, bi2 = lb % nb2;
    const int ki0 = wrap[3 * i], ki1 = wrap[3 * i + 1], ki2 = wrap[3 * i + 2];
    const bool p0 = pbc[3 * s], p1 = pbc[3 * s + 1], p2 = pbc[3 * s + 2];
    for (int d0 = -S0; d0 <= S0; d0++) {
        int tb0 = bi0 + d0, t0 = 0;
        if (p0) {
            t0 = floordiv_i(tb0, nb0);
            tb0 -= t0 * nb0;
        } else if (tb0 < 0 || tb0 >= nb0)
            continue;
        for (int d1 = -S1; d1 <= S1; d1++) {
            int tb1 = bi1 + d1, t1 = 0;
            if (p1) {
                t1 = floordiv_i(tb1, nb1);
                tb1 -= t1 * nb1;
            } else if (tb1 < 0 || tb1 >= nb1)
                continue;
            for (int d2 = -S2; d2 <= S2; d2++) {
                int tb2 = bi2 + d2, t2 = 0;
                if (p2) {
                    t2 = floordiv_i(tb2, nb2);
                    tb2 -= t2 * nb2;
                } else if (tb2 < 0 || tb2 >= nb2)
                    continue;
                const int gb = off + (tb0 * nb1 + tb1) * nb2 + tb2;
                for (int q = bin_start[gb]; q < bin_start[gb + 1]; q++) {
                    const int64_t j = bin_atoms[q];
                    const int T0 = t0 - wrap[3 * j] + ki0;
                    const int T1 = t1 - wrap[3 * j + 1] + ki1;
                    const int T2 = t2 - wrap[3 * j + 2] + ki2;
                    if (j == i && T0 == 0 && T1 == 0 && T2 == 0) continue;
                    double r[3];
                    canon_sep(pos, cell, i, j, T0, T1, T2, r);
                    if (dist2(r) < rc2) f(j, T0, T1, T2, r);
                }
            }
        }
    }
}

__device__ inline bool key_less(int64_t ja, const int32_t *Ta, int64_t jb, const int32_t *Tb) {
    if (ja != jb) return ja < jb;
    if (Ta[0] != Tb[0]) return Ta[0] < Tb[0];
    if (Ta[1] != Tb[1]) return Ta[1] < Tb[1];
    return Ta[2] < Tb[2];
}

// Images of centre i by one thread, sorted in place in the output arrays
// (insertion sort by (j, T)): the path of centres with more images than
// the wavefront fill's LDS list holds.
__device__ __forceinline__ void nl_fill_serial(int64_t i, const double *__restrict__ pos,
                                      const int64_t *__restrict__ atom_struct, const double *__restrict__ cells,
                                      const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom, double rc2,
                                      const int32_t *__restrict__ bin_start, const int32_t *__restrict__ atom_bin,
                                      const int32_t *__restrict__ wrap, const int32_t *__restrict__ bin_atoms,
                                      const int64_t *__restrict__ img_start, int64_t *__restrict__ img_atom,
                                      int64_t *__restrict__ img_j, int32_t *__restrict__ img_T,
                                      double *__restrict__ img_r, int64_t *__restrict__ row_count) {
    const int64_t s0 = img_start[i], send = img_start[i + 1];
    int64_t c = s0;
    nl_enumerate(i, pos, atom_struct, cells, pbc, geom, rc2, bin_start, atom_bin, wrap, bin_atoms,
                 [&](int64_t j, int T0, int T1, int T2, const double *r) {
                     if (c >= send) return;  // range clamped to the image capacity
                     img_j[c] = j;
                     img_T[3 * c] = T0;
                     img_T[3 * c + 1] = T1;
                     img_T[3 * c + 2] = T2;
                     img_r[3 * c] = r[0];
                     img_r[3 * c + 1] = r[1];
                     img_r[3 * c + 2] = r[2];
                     c++;
                 });
    const int64_t s1 = c;
    // insertion sort of the centre's images by (j, T0, T1, T2)
    for (int64_t a = s0 + 1; a < s1; a++) {
        const int64_t ja = img_j[a];
        const int32_t Ta[3] = {img_T[3 * a], img_T[3 * a + 1], img_T[3 * a + 2]};
        const double ra[3] = {img_r[3 * a], img_r[3 * a + 1], img_r[3 * a + 2]};
        int64_t b = a - 1;
        while (b >= s0 && key_less(ja, Ta, img_j[b], img_T + 3 * b)) {
            img_j[b + 1] = img_j[b];
            for (int k = 0; k < 3; k++) {
                img_T[3 * (b + 1) + k] = img_T[3 * b + k];
                img_r[3 * (b + 1) + k] = img_r[3 * b + k];
            }
            b--;
        }
        img_j[b + 1] = ja;
        for (int k = 0; k < 3; k++) {
            img_T[3 * (b + 1) + k] = Ta[k];
            img_r[3 * (b + 1) + k] = ra[k];
        }
    }
    int64_t rows = 0;
    bool has_self = false;
    for (int64_t a = s0; a < s1; a++) {
        img_atom[a] = i;
        if (a == s0 || img_j[a] != img_j[a - 1]) rows++;
        if (img_j[a] == i) has_self = true;
    }
    row_count[i] = rows + (has_self ? 0 : 1);
}

// seq rows: for centre i, ascending unique j with (i, i) inserted
// (SNAP.py:279-296).
__global__ void nl_rows_kernel(int64_t natoms, int64_t nimg, const int64_t *__restrict__ img_start,
                               const int64_t *__restrict__ img_j, const int64_t *__restrict__ row_start,
                               int64_t *__restrict__ seq, int64_t *__restrict__ row_img, int64_t *__restrict__ img_row,
                               int64_t *__restrict__ nsite) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natoms) return;
    int64_t rw = row_start[i];
    const int64_t s0 = img_start[i], s1 = img_start[i + 1];
    bool self_done = false;
    int64_t a = s0;
    while (a < s1 || !self_done) {
        int64_t j = (a < s1) ? img_j[a] : INT64_MAX;
        if (!self_done && i <= j) {
            // (i, i) row: with the self images if j == i, else empty
            seq[2 * rw] = i;
            seq[2 * rw + 1] = i;
            nsite[i] = rw;
            row_img[rw] = a;
            if (j == i) {
                while (a < s1 && img_j[a] == i) img_row[a++] = rw;
            }
            self_done = true;
            rw++;
            continue;
        }
        seq[2 * rw] = i;
        seq[2 * rw + 1] = j;
        row_img[rw] = a;
        while (a < s1 && img_j[a] == j) img_row[a++] = rw;
        rw++;
    }
    if (i == natoms - 1) row_img[rw] = nimg;
}

// reverse image (j, i, -T) and reverse row (j, i) by binary search in centre
// j's sorted images / rows.
__global__ void nl_rev_kernel(int64_t natoms, int64_t nimg, const int64_t *__restrict__ img_start,
                              const int64_t *__restrict__ img_atom, const int64_t *__restrict__ img_j,
                              const int32_t *__restrict__ img_T, int64_t *__restrict__ img_rev,
                              int32_t *__restrict__ flag) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nimg || e >= img_start[natoms]) return;
    const int64_t i = img_atom[e], j = img_j[e];
    const int32_t T[3] = {-img_T[3 * e], -img_T[3 * e + 1], -img_T[3 * e + 2]};
    int64_t lo = img_start[j], hi = img_start[j + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (key_less(img_j[mid], img_T + 3 * mid, i, T))
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < img_start[j + 1] && img_j[lo] == i && img_T[3 * lo] == T[0] && img_T[3 * lo + 1] == T[1] &&
        img_T[3 * lo + 2] == T[2]) {
        img_rev[e] = lo;
    } else {
        img_rev[e] = e;  // cannot happen for a symmetric list; flagged
        if (flag) atomicOr(flag, 1);
    }
}

// Reverse images without searching on T.  Within centre i's list the images
// with neighbour j form a run sorted by T; in j's list the images with
// neighbour i form a run of the same length sorted by T' = -T, and negation
// reverses the lexicographic order, so the reverse of the k-th image of the
// run (i -> j) of length n is the (n-1-k)-th of the run (j -> i).  One
// wavefront per centre, lanes over its images 64 at a time: k and n come
// from ballots over the lanes' neighbours, the start of the run (j -> i)
// from a lower-bound search of i in j's neighbour indices alone (lanes of one
// run search the same addresses).  A run that crosses a 64-image chunk
// (centres with more than 64 images) is searched with T as in
// nl_rev_kernel.  Assumes the symmetric list every build produces
// (nl_rev_kernel with its flag checks it); an index outside the images in
// use (a list cut at its capacity, which is rebuilt before use) is replaced
// by the image itself.
__global__ __launch_bounds__(64) void nl_rev_runs_kernel(int64_t natoms, int64_t nimg,
                                                         const int64_t *__restrict__ img_start,
                                                         const int64_t *__restrict__ img_j,
                                                         const int32_t *__restrict__ img_T,
                                                         int64_t *__restrict__ img_rev) {
    const int64_t i = blockIdx.x;
    if (i >= natoms) return;
    const int lane = threadIdx.x;
    const int64_t ntot = img_start[natoms] < nimg ? img_start[natoms] : nimg;
    const int64_t s = img_start[i];
    const int64_t e = img_start[i + 1] < ntot ? img_start[i + 1] : ntot;
    for (int64_t base = s; base < e; base += 64) {
        const int64_t idx = base + lane;
        const bool in = idx < e;
        const int64_t j = in ? img_j[idx] : -1;
        const int64_t jp = __shfl_up(j, 1, 64), jn = __shfl_down(j, 1, 64);
        const bool first = in && (lane == 0 || jp != j);
        const bool last = in && (lane == 63 || jn != j);
        const uint64_t fm = __ballot(first), lm = __ballot(last);
        if (!in) continue;
        // run start: the highest first lane <= lane; run end: the lowest
        // last lane >= lane (both exist for a lane in use)
        const int st = 63 - __clzll(fm & (~0ull >> (63 - lane)));
        const int en = __ffsll(lm & (~0ull << lane)) - 1;
        const bool cross = (st == 0 && base > s && img_j[base - 1] == j) ||
                           (en == 63 && base + 64 < e && img_j[base + 64] == j);
        const int64_t jlo = img_start[j], jhi = img_start[j + 1] < ntot ? img_start[j + 1] : ntot;
        int64_t r;
        if (!cross) {
            int64_t lo = jlo, hi = jhi;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (img_j[mid] < i)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            r = lo + (en - lane);
        } else {
            const int32_t T[3] = {-img_T[3 * idx], -img_T[3 * idx + 1], -img_T[3 * idx + 2]};
            int64_t lo = jlo, hi = jhi;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (key_less(img_j[mid], img_T + 3 * mid, i, T))
                    lo = mid + 1;
                else
                    hi = mid;
            }
            r = lo;
        }
        img_rev[idx] = (r >= 0 && r < ntot) ? r : idx;
    }
}

__global__ void nl_row_rev_kernel(int64_t nrows, const int64_t *__restrict__ seq, const int64_t *__restrict__ row_start,
                                  int64_t *__restrict__ row_rev, int32_t *__restrict__ flag) {
    const int64_t rw = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (rw >= nrows) return;
    const int64_t i = seq[2 * rw], j = seq[2 * rw + 1];
    int64_t lo = row_start[j], hi = row_start[j + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (seq[2 * mid + 1] < i)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < row_start[j + 1] && seq[2 * lo + 1] == i) {
        row_rev[rw] = lo;
    } else {
        row_rev[rw] = rw;
        atomicOr(flag, 2);
    }
}

// ------------------------------------------------ wave-per-centre build
// One workgroup = one wavefront = one centre.  The stencil bins are taken 64
// at a time: each lane owns one stencil bin, an exclusive wave prefix sum of
// the bin sizes flattens the candidates (bin, slot), and every lane tests
// NL_U candidates per step (the owner of candidate c is the last bin whose
// prefix is <= c).  Hits are compacted with a ballot into an LDS list; the
// fill pass ranks the list by (j, T) (rank = number of smaller keys) and
// writes it in order.  Centres with more than NL_KMAX images are filled by
// one lane of their wavefront (nl_fill_serial).
constexpr int NL_KMAX = 128;  // 7 KB of LDS per centre: many resident wavefronts

struct NlHit {
    int64_t j;
    int32_t T[4];
    double r[3];
};

constexpr int NL_U = 4;  // candidate rounds per step, their loads in flight together

template <bool FILL>
__device__ __forceinline__ int nl_wave_enumerate(int64_t i, int lane, const double *__restrict__ pos,
                                        const int64_t *__restrict__ atom_struct, const double *__restrict__ cells,
                                        const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom, double rc2,
                                        const int32_t *__restrict__ bin_start, const int32_t *__restrict__ atom_bin,
                                        const int32_t *__restrict__ wrap, const int32_t *__restrict__ bin_atoms,
                                        int *s_pref, int *s_info, NlHit *hits) {
    const int64_t s = atom_struct[i];
    const double *cell = cells + 9 * s;
    const int32_t *gm = geom + 8 * s;
    const int nb0 = gm[0], nb1 = gm[1], nb2 = gm[2];
    const int S0 = gm[3], S1 = gm[4], S2 = gm[5];
    const int off = gm[6];
    const int lb = atom_bin[i] - off;
    const int bi0 = lb / (nb1 * nb2), bi1 = (lb / nb2) % nb1, bi2 = lb % nb2;
    const int ki0 = wrap[3 * i], ki1 = wrap[3 * i + 1], ki2 = wrap[3 * i + 2];
    const bool p0 = pbc[3 * s], p1 = pbc[3 * s + 1], p2 = pbc[3 * s + 2];
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    double cl[9];
#pragma unroll
    for (int k = 0; k < 9; k++) cl[k] = cell[k];
    const int w0 = 2 * S0 + 1, w1 = 2 * S1 + 1, w2 = 2 * S2 + 1;
    const int nsten = w0 * w1 * w2;
    int count = 0;
    for (int sb = 0; sb < nsten; sb += 64) {
        const int st = sb + lane;
        int q0 = 0, t0 = 0, t1 = 0, t2 = 0, sz = 0;
        if (st < nsten) {
            const int d0 = st / (w1 * w2) - S0, d1 = (st / w2) % w1 - S1, d2 = st % w2 - S2;
            int tb0 = bi0 + d0, tb1 = bi1 + d1, tb2 = bi2 + d2;
            bool ok = true;
            if (p0) { t0 = floordiv_i(tb0, nb0); tb0 -= t0 * nb0; } else { ok = ok && tb0 >= 0 && tb0 < nb0; }
            if (p1) { t1 = floordiv_i(tb1, nb1); tb1 -= t1 * nb1; } else { ok = ok && tb1 >= 0 && tb1 < nb1; }
            if (p2) { t2 = floordiv_i(tb2, nb2); tb2 -= t2 * nb2; } else { ok = ok && tb2 >= 0 && tb2 < nb2; }
            if (ok) {
                const int gb = off + (tb0 * nb1 + tb1) * nb2 + tb2;
                q0 = bin_start[gb];
                sz = bin_start[gb + 1] - q0;
            }
        }
        int incl = sz;
        for (int o = 1; o < 64; o <<= 1) {
            int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const int total = __shfl(incl, 63);
        __syncthreads();  // previous chunk's readers are done with s_pref / s_info
        s_pref[lane] = incl - sz;
        s_info[4 * lane] = q0;
        s_info[4 * lane + 1] = t0;
        s_info[4 * lane + 2] = t1;
        s_info[4 * lane + 3] = t2;
        __syncthreads();
        for (int cb = 0; cb < total; cb += 64 * NL_U) {
            // candidate c belongs to the last stencil bin whose prefix is
            // <= c (s_pref is non-decreasing, s_pref[0] = 0); a fixed
            // six-step search, the NL_U candidates of a lane interleaved
            int own[NL_U], q[NL_U];
            bool in[NL_U];
#pragma unroll
            for (int u = 0; u < NL_U; u++) {
                const int c = cb + 64 * u + lane;
                in[u] = c < total;
                int lo = 0;
#pragma unroll
                for (int w = 32; w > 0; w >>= 1)
                    if (s_pref[lo + w] <= c) lo += w;
                own[u] = lo;
                q[u] = s_info[4 * lo] + (c - s_pref[lo]);
            }
            int64_t j[NL_U];
#pragma unroll
            for (int u = 0; u < NL_U; u++) j[u] = in[u] ? (int64_t)bin_atoms[q[u]] : i;
            int wj[NL_U][3];
            double pj[NL_U][3];
#pragma unroll
            for (int u = 0; u < NL_U; u++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    wj[u][k] = wrap[3 * j[u] + k];
                    pj[u][k] = pos[3 * j[u] + k];
                }
#pragma unroll
            for (int u = 0; u < NL_U; u++) {
                const int T0 = s_info[4 * own[u] + 1] - wj[u][0] + ki0;
                const int T1 = s_info[4 * own[u] + 2] - wj[u][1] + ki1;
                const int T2 = s_info[4 * own[u] + 3] - wj[u][2] + ki2;
                double r[3] = {0, 0, 0};
                bool hit = false;
                if (in[u] && !(j[u] == i && T0 == 0 && T1 == 0 && T2 == 0)) {
                    canon_sep_v(pi, pj[u], cl, i, j[u], T0, T1, T2, r);
                    hit = dist2(r) < rc2;
                }
                const uint64_t mask = __ballot(hit);
                if (FILL && hit) {
                    const int rank = count + (int)__popcll(mask & ((1ull << lane) - 1ull));
                    if (rank < NL_KMAX) {
                        NlHit &h = hits[rank];
                        h.j = j[u];
                        h.T[0] = T0;
                        h.T[1] = T1;
                        h.T[2] = T2;
                        h.r[0] = r[0];
                        h.r[1] = r[1];
                        h.r[2] = r[2];
                    }
                }
                count += (int)__popcll(mask);
            }
        }
    }
    return count;
}

__global__ __launch_bounds__(64) void nl_count_wave_kernel(
    int64_t natoms, const double *__restrict__ pos, const int64_t *__restrict__ atom_struct,
    const double *__restrict__ cells, const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom, double rc2,
    const int32_t *__restrict__ bin_start, const int32_t *__restrict__ atom_bin, const int32_t *__restrict__ wrap,
    const int32_t *__restrict__ bin_atoms, int64_t *__restrict__ img_count) {
    __shared__ int s_pref[64];
    __shared__ int s_info[256];
    const int lane = threadIdx.x;
    const int64_t i = blockIdx.x;
    if (i >= natoms) return;  // uniform for the whole workgroup
    const int c = nl_wave_enumerate<false>(i, lane, pos, atom_struct, cells, pbc, geom, rc2, bin_start, atom_bin, wrap,
                                           bin_atoms, s_pref, s_info, nullptr);
    if (lane == 0) img_count[i] = c;
}

__device__ inline bool hit_less(const NlHit &a, const NlHit &b) {
    if (a.j != b.j) return a.j < b.j;
    if (a.T[0] != b.T[0]) return a.T[0] < b.T[0];
    if (a.T[1] != b.T[1]) return a.T[1] < b.T[1];
    return a.T[2] < b.T[2];
}

// Ranks the centre's first k LDS hits by (j, T) and hands each hit of rank
// < kcap to put(rank, hit); returns (in every lane) the centre's seq-row
// count: the distinct neighbours among them, plus one for the (i, i) row
// when i is not its own neighbour.  Starts with the wavefront's LDS keys
// (every lane calls it).
template <typename Put>
__device__ __forceinline__ int64_t nl_rank_hits(int64_t i, int k, int kcap, int lane, const NlHit *hits, uint64_t *keys,
                                       Put &&put) {
    // sort keys (j, T0, T1, T2) packed in 64 bits, T_d + 128 in 8 bits each,
    // when every |T_d| < 128 (the lexicographic order of the packed key is
    // that of hit_less); otherwise the comparison of the fields
    bool wide = false;
    for (int e = lane; e < k; e += 64) {
        const NlHit &h = hits[e];
        wide = wide || h.T[0] < -128 || h.T[0] > 127 || h.T[1] < -128 || h.T[1] > 127 || h.T[2] < -128 ||
               h.T[2] > 127;
        keys[e] = ((uint64_t)h.j << 24) | ((uint64_t)(h.T[0] + 128) << 16) | ((uint64_t)(h.T[1] + 128) << 8) |
                  (uint64_t)(h.T[2] + 128);
    }
    const bool packed = __ballot(wide) == 0;
    __syncthreads();
    int firsts = 0;
    bool self = false;
    for (int e = lane; e < k; e += 64) {
        const NlHit h = hits[e];
        int rank = 0;
        bool first = true;
        if (packed) {
            const uint64_t mk = keys[e];
#pragma unroll 4
            for (int q = 0; q < k; q++) {
                const uint64_t o = keys[q];
                const bool lt = o < mk;
                rank += lt ? 1 : 0;
                if (lt && (o >> 24) == (mk >> 24)) first = false;
            }
        } else {
            for (int q = 0; q < k; q++) {
                const NlHit &o = hits[q];
                if (hit_less(o, h)) {
                    rank++;
                    if (o.j == h.j) first = false;
                }
            }
        }
        if (rank >= kcap) continue;  // (the fill's range was clamped to the image capacity)
        put(rank, h);
        if (first) firsts++;
        if (h.j == i) self = true;
    }
    for (int o = 32; o > 0; o >>= 1) firsts += __shfl_xor(firsts, o);
    const bool any_self = __ballot(self) != 0;
    return firsts + (any_self ? 0 : 1);
}

// The fill of one centre by its wavefront: enumerate into LDS, rank, write
// the images at img_start[i] in (j, T) order and the seq-row count.
__device__ __forceinline__ void nl_fill_centre(int64_t i, int lane, const double *__restrict__ pos,
                                      const int64_t *__restrict__ atom_struct, const double *__restrict__ cells,
                                      const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom, double rc2,
                                      const int32_t *__restrict__ bin_start, const int32_t *__restrict__ atom_bin,
                                      const int32_t *__restrict__ wrap, const int32_t *__restrict__ bin_atoms,
                                      const int64_t *__restrict__ img_start, int64_t *__restrict__ img_atom,
                                      int64_t *__restrict__ img_j, int32_t *__restrict__ img_T,
                                      double *__restrict__ img_r, int64_t *__restrict__ row_count, int *s_pref,
                                      int *s_info, NlHit *hits, uint64_t *keys) {
    const int64_t s0 = img_start[i];
    const int k = (int)(img_start[i + 1] - s0);
    if (k > NL_KMAX) {
        // more images than the LDS list holds: one lane fills the centre
        if (lane == 0)
            nl_fill_serial(i, pos, atom_struct, cells, pbc, geom, rc2, bin_start, atom_bin, wrap, bin_atoms, img_start,
                           img_atom, img_j, img_T, img_r, row_count);
        return;
    }
    const int c = nl_wave_enumerate<true>(i, lane, pos, atom_struct, cells, pbc, geom, rc2, bin_start, atom_bin, wrap,
                                          bin_atoms, s_pref, s_info, hits);
    (void)c;
    __syncthreads();
    const int64_t nrow = nl_rank_hits(i, k, k, lane, hits, keys, [&](int rank, const NlHit &h) {
        const int64_t d = s0 + rank;
        img_atom[d] = i;
        img_j[d] = h.j;
        img_T[3 * d] = h.T[0];
        img_T[3 * d + 1] = h.T[1];
        img_T[3 * d + 2] = h.T[2];
        img_r[3 * d] = h.r[0];
        img_r[3 * d + 1] = h.r[1];
        img_r[3 * d + 2] = h.r[2];
    });
    if (lane == 0) row_count[i] = nrow;
}

__global__ __launch_bounds__(64) void nl_fill_wave_kernel(
    int64_t natoms, const double *__restrict__ pos, const int64_t *__restrict__ atom_struct,
    const double *__restrict__ cells, const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom, double rc2,
    const int32_t *__restrict__ bin_start, const int32_t *__restrict__ atom_bin, const int32_t *__restrict__ wrap,
    const int32_t *__restrict__ bin_atoms, const int64_t *__restrict__ img_start, int64_t *__restrict__ img_atom,
    int64_t *__restrict__ img_j, int32_t *__restrict__ img_T, double *__restrict__ img_r,
    int64_t *__restrict__ row_count) {
    __shared__ int s_pref[64];
    __shared__ int s_info[256];
    __shared__ NlHit hits[NL_KMAX];
    __shared__ uint64_t keys[NL_KMAX];
    const int64_t i = blockIdx.x;
    if (i >= natoms) return;
    nl_fill_centre(i, threadIdx.x, pos, atom_struct, cells, pbc, geom, rc2, bin_start, atom_bin, wrap, bin_atoms,
                   img_start, img_atom, img_j, img_T, img_r, row_count, s_pref, s_info, hits, keys);
}

// The count with the images kept: every centre's wavefront enumerates and
// counts as nl_count_wave_kernel does, and a centre with at most `slot`
// images also ranks them and writes them, in (j, T) order, into its slot
// (slot_j / slot_T / slot_r at [i * slot, i * slot + count)) together with
// its seq-row count; nl_fill_slots_kernel then copies the slots into place
// after the scan, so those centres are enumerated once per build.
__global__ __launch_bounds__(64) void nl_count_slots_kernel(
    int64_t natoms, const double *__restrict__ pos, const int64_t *__restrict__ atom_struct,
    const double *__restrict__ cells, const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom, double rc2,
    const int32_t *__restrict__ bin_start, const int32_t *__restrict__ atom_bin, const int32_t *__restrict__ wrap,
    const int32_t *__restrict__ bin_atoms, int64_t *__restrict__ img_count, int slot, int64_t *__restrict__ slot_j,
    int32_t *__restrict__ slot_T, double *__restrict__ slot_r, int64_t *__restrict__ row_count) {
    __shared__ int s_pref[64];
    __shared__ int s_info[256];
    __shared__ NlHit hits[NL_KMAX];
    __shared__ uint64_t keys[NL_KMAX];
    const int lane = threadIdx.x;
    const int64_t i = blockIdx.x;
    if (i >= natoms) return;  // uniform for the whole workgroup
    const int c = nl_wave_enumerate<true>(i, lane, pos, atom_struct, cells, pbc, geom, rc2, bin_start, atom_bin, wrap,
                                          bin_atoms, s_pref, s_info, hits);
    if (lane == 0) img_count[i] = c;
    if (c > slot) return;  // uniform (the count is the wavefront's); the fill enumerates this centre
    __syncthreads();
    const int64_t nrow = nl_rank_hits(i, c, c, lane, hits, keys, [&](int rank, const NlHit &h) {
        const int64_t d = i * (int64_t)slot + rank;
        slot_j[d] = h.j;
        slot_T[3 * d] = h.T[0];
        slot_T[3 * d + 1] = h.T[1];
        slot_T[3 * d + 2] = h.T[2];
        slot_r[3 * d] = h.r[0];
        slot_r[3 * d + 1] = h.r[1];
        slot_r[3 * d + 2] = h.r[2];
    });
    if (lane == 0) row_count[i] = nrow;
}

// The fill after nl_count_slots_kernel: a centre whose images are in its
// slot copies them into place (its seq-row count is written already); the
// others are filled as by nl_fill_wave_kernel.
__global__ __launch_bounds__(64) void nl_fill_slots_kernel(
    int64_t natoms, const double *__restrict__ pos, const int64_t *__restrict__ atom_struct,
    const double *__restrict__ cells, const int32_t *__restrict__ pbc, const int32_t *__restrict__ geom, double rc2,
    const int32_t *__restrict__ bin_start, const int32_t *__restrict__ atom_bin, const int32_t *__restrict__ wrap,
    const int32_t *__restrict__ bin_atoms, const int64_t *__restrict__ img_start, const int64_t *__restrict__ img_count,
    int slot, const int64_t *__restrict__ slot_j, const int32_t *__restrict__ slot_T,
    const double *__restrict__ slot_r, int64_t *__restrict__ img_atom, int64_t *__restrict__ img_j,
    int32_t *__restrict__ img_T, double *__restrict__ img_r, int64_t *__restrict__ row_count) {
    __shared__ int s_pref[64];
    __shared__ int s_info[256];
    __shared__ NlHit hits[NL_KMAX];
    __shared__ uint64_t keys[NL_KMAX];
    const int lane = threadIdx.x;
    const int64_t i = blockIdx.x;
    if (i >= natoms) return;
    if (img_count[i] <= slot) {
        const int64_t s0 = img_start[i];
        const int k = (int)(img_start[i + 1] - s0);  // <= the count (clamped to the image capacity)
        for (int e = lane; e < k; e += 64) {
            const int64_t d = s0 + e, q = i * (int64_t)slot + e;
            img_atom[d] = i;
            img_j[d] = slot_j[q];
            img_T[3 * d] = slot_T[3 * q];
            img_T[3 * d + 1] = slot_T[3 * q + 1];
            img_T[3 * d + 2] = slot_T[3 * q + 2];
            img_r[3 * d] = slot_r[3 * q];
            img_r[3 * d + 1] = slot_r[3 * q + 1];
            img_r[3 * d + 2] = slot_r[3 * q + 2];
        }
        return;
    }
    nl_fill_centre(i, lane, pos, atom_struct, cells, pbc, geom, rc2, bin_start, atom_bin, wrap, bin_atoms, img_start,
                   img_atom, img_j, img_T, img_r, row_count, s_pref, s_info, hits, keys);
}

// Zeroes n ints.  Used instead of a memset so that a build can be captured
// into a HIP graph (memset nodes have given wrong results on replays after
// the first).
__global__ void nl_zero_i32_kernel(int32_t *__restrict__ p, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

// ------------------------------------------------------------ device scan
// Exclusive scan by one workgroup of 1024 threads (16 wavefronts) over
// tiles of 4096 entries: each thread sums four consecutive entries, the
// wavefronts scan those sums by shuffles, and one LDS exchange of the 16
// wavefront totals per tile gives every thread its offset (two barriers
// per tile).  out has n+1 entries; out[n] = total.  With cap >= 0 every
// output is clamped to cap (so ranges read through it stay inside arrays
// of cap entries) and the unclamped total goes to total_out[0].
constexpr int SCAN_ITEMS = 4;
template <typename T>
__global__ __launch_bounds__(1024) void scan_kernel(int64_t n, const T *__restrict__ in, T *__restrict__ out,
                                                    T cap = -1, T *__restrict__ total_out = nullptr) {
    __shared__ T wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    T carry = 0;
    for (int64_t base = 0; base < n; base += 1024 * SCAN_ITEMS) {
        T v[SCAN_ITEMS];
        T own = 0;
#pragma unroll
        for (int u = 0; u < SCAN_ITEMS; u++) {
            const int64_t k = base + (int64_t)SCAN_ITEMS * tid + u;
            v[u] = k < n ? in[k] : 0;
            own += v[u];
        }
        T x = own;  // inclusive scan of the thread sums within the wavefront
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const T y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        T before = carry, tile = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const T t = wsum[q];
            if (q < w) before += t;
            tile += t;
        }
        T run = before + x - own;
#pragma unroll
        for (int u = 0; u < SCAN_ITEMS; u++) {
            const int64_t k = base + (int64_t)SCAN_ITEMS * tid + u;
            if (k < n) out[k] = (cap >= 0 && run > cap) ? cap : run;
            run += v[u];
        }
        carry += tile;
        __syncthreads();  // wsum is rewritten by the next tile
    }
    if (tid == 0) {
        out[n] = (cap >= 0 && carry > cap) ? cap : carry;
        if (total_out) total_out[0] = carry;
    }
}

}  // namespace pxff

using namespace pxff;

extern "C" int pxff_scan_i64(const int64_t *d_in, int64_t n, int64_t *d_out, void *stream) {
    PXFF_REQUIRE(n >= 0, "n must be >= 0");
    scan_kernel<int64_t><<<1, 1024, 0, (hipStream_t)stream>>>(n, d_in, d_out);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_scan_i64_capped(const int64_t *d_in, int64_t n, int64_t cap, int64_t *d_out, int64_t *d_total,
                                    void *stream) {
    PXFF_REQUIRE(n >= 0 && cap >= 0, "n and cap must be >= 0");
    PXFF_REQUIRE(d_total != nullptr, "null total");
    scan_kernel<int64_t><<<1, 1024, 0, (hipStream_t)stream>>>(n, d_in, d_out, cap, d_total);
    PXFF_LAUNCH_CHECK();
    return 0;
}

// Binning of pxff_nlist_count: the atoms' bins and wraps, the bin starts
// and the bin members.
static int nl_bin_launch(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct, const double *d_inv_cells,
                         const int32_t *d_pbc, const int32_t *d_geom, int64_t nbins_total, int32_t *d_bin_count,
                         int32_t *d_bin_start, int32_t *d_atom_bin, int32_t *d_wrap, int32_t *d_bin_atoms,
                         hipStream_t s) {
    if (natoms <= NL_SMALL_ATOMS && nbins_total <= NL_SMALL_BINS) {
        nl_bin_place_kernel<<<div_up(natoms, 256), 256, 0, s>>>(natoms, d_pos, d_atom_struct, d_inv_cells, d_pbc,
                                                                d_geom, d_atom_bin, d_wrap);
        PXFF_LAUNCH_CHECK();
        nl_bin_small_kernel<<<1, 1024, 0, s>>>(natoms, (int)nbins_total, d_atom_bin, d_bin_start, d_bin_atoms);
        PXFF_LAUNCH_CHECK();
        return 0;
    }
    nl_zero_i32_kernel<<<div_up(nbins_total, 256), 256, 0, s>>>(d_bin_count, nbins_total);
    PXFF_LAUNCH_CHECK();
    nl_bin_kernel<<<div_up(natoms, 256), 256, 0, s>>>(natoms, d_pos, d_atom_struct, d_inv_cells, d_pbc, d_geom,
                                                      d_atom_bin, d_wrap, d_bin_count);
    PXFF_LAUNCH_CHECK();
    scan_kernel<int32_t><<<1, 1024, 0, s>>>(nbins_total, d_bin_count, d_bin_start);
    PXFF_LAUNCH_CHECK();
    nl_zero_i32_kernel<<<div_up(nbins_total, 256), 256, 0, s>>>(d_bin_count, nbins_total);
    PXFF_LAUNCH_CHECK();
    nl_bin_fill_kernel<<<div_up(natoms, 256), 256, 0, s>>>(natoms, d_atom_bin, d_bin_start, d_bin_count, d_bin_atoms);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_nlist_count(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                                const int64_t *d_atom_start, const double *d_cells, const double *d_inv_cells,
                                const int32_t *d_pbc, const int32_t *d_geom, int64_t nbins_total, double rc,
                                int32_t *d_bin_count, int32_t *d_bin_start, int32_t *d_atom_bin, int32_t *d_wrap,
                                int32_t *d_bin_atoms, int64_t *d_img_count, void *stream) {
    (void)d_atom_start;
    PXFF_REQUIRE(rc > 0, "rc must be > 0");
    PXFF_REQUIRE(nbins_total > 0, "nbins_total must be > 0");
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    int rc_ = nl_bin_launch(d_pos, natoms, d_atom_struct, d_inv_cells, d_pbc, d_geom, nbins_total, d_bin_count,
                            d_bin_start, d_atom_bin, d_wrap, d_bin_atoms, s);
    if (rc_) return rc_;
    nl_count_wave_kernel<<<(unsigned)natoms, 64, 0, s>>>(natoms, d_pos, d_atom_struct, d_cells, d_pbc, d_geom, rc * rc,
                                                           d_bin_start, d_atom_bin, d_wrap, d_bin_atoms, d_img_count);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_nlist_count_slots(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                                      const int64_t *d_atom_start, const double *d_cells, const double *d_inv_cells,
                                      const int32_t *d_pbc, const int32_t *d_geom, int64_t nbins_total, double rc,
                                      int32_t *d_bin_count, int32_t *d_bin_start, int32_t *d_atom_bin,
                                      int32_t *d_wrap, int32_t *d_bin_atoms, int64_t *d_img_count, int slot,
                                      int64_t *d_slot_j, int32_t *d_slot_T, double *d_slot_r, int64_t *d_row_count,
                                      void *stream) {
    (void)d_atom_start;
    PXFF_REQUIRE(rc > 0, "rc must be > 0");
    PXFF_REQUIRE(nbins_total > 0, "nbins_total must be > 0");
    PXFF_REQUIRE(slot >= 1 && slot <= NL_KMAX, "slot must be in 1..128");
    PXFF_REQUIRE(d_slot_j && d_slot_T && d_slot_r && d_row_count, "slot arrays missing");
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    int rc_ = nl_bin_launch(d_pos, natoms, d_atom_struct, d_inv_cells, d_pbc, d_geom, nbins_total, d_bin_count,
                            d_bin_start, d_atom_bin, d_wrap, d_bin_atoms, s);
    if (rc_) return rc_;
    nl_count_slots_kernel<<<(unsigned)natoms, 64, 0, s>>>(natoms, d_pos, d_atom_struct, d_cells, d_pbc, d_geom,
                                                            rc * rc, d_bin_start, d_atom_bin, d_wrap, d_bin_atoms,
                                                            d_img_count, slot, d_slot_j, d_slot_T, d_slot_r,
                                                            d_row_count);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_nlist_fill(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                               const int64_t *d_atom_start, const double *d_cells, const int32_t *d_pbc,
                               const int32_t *d_geom, double rc, const int32_t *d_bin_start, const int32_t *d_atom_bin,
                               const int32_t *d_wrap, const int32_t *d_bin_atoms, const int64_t *d_img_start,
                               int64_t *d_img_atom, int64_t *d_img_j, int32_t *d_img_T, double *d_img_r,
                               int64_t *d_row_count, void *stream) {
    (void)d_atom_start;
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    nl_fill_wave_kernel<<<(unsigned)natoms, 64, 0, s>>>(natoms, d_pos, d_atom_struct, d_cells, d_pbc, d_geom, rc * rc,
                                                          d_bin_start, d_atom_bin, d_wrap, d_bin_atoms, d_img_start,
                                                          d_img_atom, d_img_j, d_img_T, d_img_r, d_row_count);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_nlist_fill_slots(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                                     const int64_t *d_atom_start, const double *d_cells, const int32_t *d_pbc,
                                     const int32_t *d_geom, double rc, const int32_t *d_bin_start,
                                     const int32_t *d_atom_bin, const int32_t *d_wrap, const int32_t *d_bin_atoms,
                                     const int64_t *d_img_start, const int64_t *d_img_count, int slot,
                                     const int64_t *d_slot_j, const int32_t *d_slot_T, const double *d_slot_r,
                                     int64_t *d_img_atom, int64_t *d_img_j, int32_t *d_img_T, double *d_img_r,
                                     int64_t *d_row_count, void *stream) {
    (void)d_atom_start;
    PXFF_REQUIRE(slot >= 1 && slot <= NL_KMAX, "slot must be in 1..128");
    PXFF_REQUIRE(d_slot_j && d_slot_T && d_slot_r && d_img_count, "slot arrays missing");
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    nl_fill_slots_kernel<<<(unsigned)natoms, 64, 0, s>>>(natoms, d_pos, d_atom_struct, d_cells, d_pbc, d_geom,
                                                           rc * rc, d_bin_start, d_atom_bin, d_wrap, d_bin_atoms,
                                                           d_img_start, d_img_count, slot, d_slot_j, d_slot_T,
                                                           d_slot_r, d_img_atom, d_img_j, d_img_T, d_img_r,
                                                           d_row_count);
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_nlist_rev(int64_t natoms, int64_t nimg, const int64_t *d_img_start, const int64_t *d_img_atom,
                              const int64_t *d_img_j, const int32_t *d_img_T, int64_t *d_img_rev, int32_t *d_flag,
                              void *stream) {
    if (natoms <= 0 || nimg <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (d_flag) {
        nl_zero_i32_kernel<<<1, 64, 0, s>>>(d_flag, 1);
        PXFF_LAUNCH_CHECK();
    }
    if (d_flag) {
        // the checking search (verify)
        nl_rev_kernel<<<div_up(nimg, 256), 256, 0, s>>>(natoms, nimg, d_img_start, d_img_atom, d_img_j, d_img_T,
                                                        d_img_rev, d_flag);
    } else {
        nl_rev_runs_kernel<<<(unsigned)natoms, 64, 0, s>>>(natoms, nimg, d_img_start, d_img_j, d_img_T, d_img_rev);
    }
    PXFF_LAUNCH_CHECK();
    return 0;
}

extern "C" int pxff_nlist_rows(int64_t natoms, int64_t nimg, int64_t nrows, const int64_t *d_img_start,
                               const int64_t *d_img_j, const int64_t *d_row_start, int64_t *d_seq, int64_t *d_row_img,
                               int64_t *d_img_row, int64_t *d_nsite, int64_t *d_row_rev, int32_t *d_flag,
                               void *stream) {
    if (natoms <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    nl_zero_i32_kernel<<<1, 64, 0, s>>>(d_flag, 1);
    PXFF_LAUNCH_CHECK();
    nl_rows_kernel<<<div_up(natoms, 64), 64, 0, s>>>(natoms, nimg, d_img_start, d_img_j, d_row_start, d_seq, d_row_img,
                                                     d_img_row, d_nsite);
    PXFF_LAUNCH_CHECK();
    nl_row_rev_kernel<<<div_up(nrows, 256), 256, 0, s>>>(nrows, d_seq, d_row_start, d_row_rev, d_flag);
    PXFF_LAUNCH_CHECK();
    return 0;
}
