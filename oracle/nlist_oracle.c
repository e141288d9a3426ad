/*
 * CPU oracle for the neighbour list of the descriptor hot path (TEST
 * INFRASTRUCTURE ONLY; used by tests/, smoke() and bench.py's cpu_baseline).
 *
 * It restates what the reference obtains from ASE's NeighborList with
 * cutoffs [rc/2]*N, self_interaction=False, bothways=True, skin=0
 * (pyxtal_ff/descriptors/SNAP.py:250-252, SO4.py:221-223, SO3.py:326-332):
 * every (j, lattice translation T) with |pos_j + T.cell - pos_i| < rc,
 * except (i, T = 0).  It enumerates images by brute force over all atom
 * pairs and every translation that can reach the cutoff -- an algorithm
 * independent of the device cell list it checks.
 *
 * Conventions shared with the device build (so that image sets compare
 * bit-exactly):
 *   - the separation of image (i, j, T) is computed in a canonical
 *     orientation: for (j > i) or (j == i and T > 0 lexicographically)
 *     r = (pos_j + (T0*c0 + T1*c1 + T2*c2)) - pos_i; otherwise r is the
 *     negation of the reverse image's separation.  This makes the list
 *     exactly symmetric (r_ji = -r_ij) and the cut test identical for
 *     both directions;
 *   - images are ordered by (i, j, T0, T1, T2).
 * This file must be compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static void inv3(const double *c, double *inv) {
    double a = c[0], b = c[1], cc = c[2], d = c[3], e = c[4], f = c[5], g = c[6], h = c[7], k = c[8];
    double det = a * (e * k - f * h) - b * (d * k - f * g) + cc * (d * h - e * g);
    inv[0] = (e * k - f * h) / det;
    inv[1] = -(b * k - cc * h) / det;
    inv[2] = (b * f - cc * e) / det;
    inv[3] = -(d * k - f * g) / det;
    inv[4] = (a * k - cc * g) / det;
    inv[5] = -(a * f - cc * d) / det;
    inv[6] = (d * h - e * g) / det;
    inv[7] = -(a * h - b * g) / det;
    inv[8] = (a * e - b * d) / det;
}

/* separation in the canonical orientation (see header) */
static void canon_sep(const double *pos, const double *cell, int64_t i, int64_t j, const int *T, double *r) {
    int fwd;
    if (j != i)
        fwd = j > i;
    else
        fwd = (T[0] > 0) || (T[0] == 0 && T[1] > 0) || (T[0] == 0 && T[1] == 0 && T[2] > 0);
    if (fwd) {
        for (int k = 0; k < 3; k++) {
            double t = (double)T[0] * cell[k];
            t = t + (double)T[1] * cell[3 + k];
            t = t + (double)T[2] * cell[6 + k];
            double p = pos[3 * j + k] + t;
            r[k] = p - pos[3 * i + k];
        }
    } else {
        for (int k = 0; k < 3; k++) {
            double t = (double)(-T[0]) * cell[k];
            t = t + (double)(-T[1]) * cell[3 + k];
            t = t + (double)(-T[2]) * cell[6 + k];
            double p = pos[3 * i + k] + t;
            r[k] = -(p - pos[3 * j + k]);
        }
    }
}

/*
 * Two calls: with out arrays NULL it returns the number of images; with
 * arrays of that length it fills them.  Outputs per image: centre i, neighbour
 * j, translation T[3], separation r[3].  Ordered by (i, j, T).
 */
static void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void unit3(double *a) {
    double s = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int k = 0; k < 3; k++) a[k] /= s;
}

/*
 * Lattice for the fractional coordinates: the cell with its non-periodic
 * rows replaced by unit vectors normal to the periodic rows (by cross
 * products).  Only periodic directions use fractional coordinates, and a
 * cluster may carry a zero cell along its non-periodic directions, as ASE's
 * NeighborList allows (SNAP.py:250-252).  Separations use the cell itself.
 */
static void frac_lattice(const double *cell, const int *pbc, double *lat) {
    int per[3], non[3], np_ = 0, nn = 0;
    memcpy(lat, cell, 9 * sizeof(double));
    for (int d = 0; d < 3; d++) {
        if (pbc[d])
            per[np_++] = d;
        else
            non[nn++] = d;
    }
    if (np_ == 3) return;
    if (np_ == 2) {
        cross3(cell + 3 * per[0], cell + 3 * per[1], lat + 3 * non[0]);
        unit3(lat + 3 * non[0]);
    } else if (np_ == 1) {
        const double *p = cell + 3 * per[0];
        /* the axis least aligned with p gives a well-conditioned first normal */
        int k = 0;
        for (int d = 1; d < 3; d++)
            if (fabs(p[d]) < fabs(p[k])) k = d;
        double e[3] = {0, 0, 0};
        e[k] = 1.0;
        cross3(p, e, lat + 3 * non[0]);
        unit3(lat + 3 * non[0]);
        cross3(p, lat + 3 * non[0], lat + 3 * non[1]);
        unit3(lat + 3 * non[1]);
    } else {
        memset(lat, 0, 9 * sizeof(double));
        lat[0] = lat[4] = lat[8] = 1.0;
    }
}

int64_t nlist_oracle(const double *pos, int64_t n, const double *cell, const int *pbc, double rc,
                     int64_t *out_i, int64_t *out_j, int *out_T, double *out_r) {
    double inv[9], lat[9];
    frac_lattice(cell, pbc, lat);
    inv3(lat, inv);
    /* plane spacings h_d = 1/|row d of inverse-transpose| */
    double h[3];
    for (int d = 0; d < 3; d++) {
        double s = inv[d] * inv[d] + inv[3 + d] * inv[3 + d] + inv[6 + d] * inv[6 + d];
        h[d] = 1.0 / sqrt(s);
    }
    double *frac = malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    for (int64_t a = 0; a < n; a++)
        for (int d = 0; d < 3; d++)
            frac[3 * a + d] = pos[3 * a] * inv[d] + pos[3 * a + 1] * inv[3 + d] + pos[3 * a + 2] * inv[6 + d];
    const double rc2 = rc * rc;
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < n; j++) {
            int lo[3], hi[3];
            for (int d = 0; d < 3; d++) {
                if (!pbc[d]) {
                    lo[d] = hi[d] = 0;
                    continue;
                }
                /* |ds + T| * h_d <= |r| is necessary for |r| < rc; the
                 * 1e-9 slack covers rounding of the fractional coordinates */
                double ds = frac[3 * j + d] - frac[3 * i + d];
                double span = rc / h[d] + 1e-9;
                lo[d] = (int)ceil(-ds - span);
                hi[d] = (int)floor(-ds + span);
            }
            int T[3];
            for (T[0] = lo[0]; T[0] <= hi[0]; T[0]++)
                for (T[1] = lo[1]; T[1] <= hi[1]; T[1]++)
                    for (T[2] = lo[2]; T[2] <= hi[2]; T[2]++) {
                        if (j == i && T[0] == 0 && T[1] == 0 && T[2] == 0) continue;
                        double r[3];
                        canon_sep(pos, cell, i, j, T, r);
                        double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (!(d2 < rc2)) continue;
                        if (out_i) {
                            out_i[cnt] = i;
                            out_j[cnt] = j;
                            out_T[3 * cnt] = T[0];
                            out_T[3 * cnt + 1] = T[1];
                            out_T[3 * cnt + 2] = T[2];
                            out_r[3 * cnt] = r[0];
                            out_r[3 * cnt + 1] = r[1];
                            out_r[3 * cnt + 2] = r[2];
                        }
                        cnt++;
                    }
        }
    }
    free(frac);
    return cnt;
}
