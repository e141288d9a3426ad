// Host construction of the SO(4)/SNAP index tables (see so4_tables.h).
#include "so4_tables.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace pxff {

namespace {

// The reference's factorial table (angular_momentum.py:9-182): exact up to
// 16!, then 170!, 171!, ... at indices 17.. -- kept as is so that the
// Clebsch-Gordan values equal the reference's for every 2J.
const double kFac[] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800, 39916800, 479001600,
                       6227020800.0, 87178291200.0, 1307674368000.0, 20922789888000.0,
                       7.59070505394721e+286, 1.22969421873945e+289, 2.0044015765453e+291,
                       3.28721858553429e+293, 5.42391066613159e+295, 9.00369170577843e+297,
                       1.503616514865e+300};

double fac(long n) {
    if (n < 0 || n >= (long)(sizeof(kFac) / sizeof(kFac[0]))) return NAN;
    return kFac[n];
}

long fdiv(long a, long b) {
    long q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
    return q;
}

// deltacg, angular_momentum.py:184-190
double deltacg(long l1, long l2, long l) {
    double s = fac(fdiv(l1 + l2 + l, 2) + 1);
    return std::sqrt(fac(fdiv(l1 + l2 - l, 2)) * fac(fdiv(l1 - l2 + l, 2)) * fac(fdiv(-l1 + l2 + l, 2)) / s);
}

// init_clebsch_gordan, SNAP.py:442-510
void clebsch_gordan(int twol, std::vector<double> &cg) {
    int ldim = twol + 1;
    cg.clear();
    for (int l1 = 0; l1 < ldim; l1++)
        for (int l2 = 0; l2 <= l1; l2++)
            for (int l = l1 - l2; l <= std::min(twol, l1 + l2); l += 2)
                for (int m1 = 0; m1 <= l1; m1++) {
                    int aa2 = 2 * m1 - l1;
                    for (int m2 = 0; m2 <= l2; m2++) {
                        int bb2 = 2 * m2 - l2;
                        int m2x = aa2 + bb2 + l;
                        if (m2x < 0 || m2x > 2 * l) {
                            cg.push_back(0.0);
                            continue;
                        }
                        long m = m2x / 2;
                        double sum = 0.0;
                        long zlo = std::max(0L, std::max(fdiv(-(l - l2 + aa2), 2), fdiv(-(l - l1 - bb2), 2)));
                        long zhi = std::min(fdiv(l1 + l2 - l, 2), std::min(fdiv(l1 - aa2, 2), fdiv(l2 + bb2, 2)));
                        for (long z = zlo; z <= zhi; z++) {
                            double ifac = (z % 2 == 0) ? 1.0 : -1.0;
                            sum += ifac / (fac(z) * fac(fdiv(l1 + l2 - l, 2) - z) * fac(fdiv(l1 - aa2, 2) - z) *
                                           fac(fdiv(l2 + bb2, 2) - z) * fac(fdiv(l - l2 + aa2, 2) + z) *
                                           fac(fdiv(l - l1 - bb2, 2) + z));
                        }
                        long cc2 = 2 * m - l;
                        double dcg = deltacg(l1, l2, l);
                        double sf = std::sqrt(fac(fdiv(l1 + aa2, 2)) * fac(fdiv(l1 - aa2, 2)) * fac(fdiv(l2 + bb2, 2)) *
                                              fac(fdiv(l2 - bb2, 2)) * fac(fdiv(l + cc2, 2)) * fac(fdiv(l - cc2, 2)) *
                                              (double)(l + 1));
                        cg.push_back(sum * dcg * sf);
                    }
                }
}

}  // namespace

bool build_so4_tables(int twol, int normalize_u, So4HostTables &t) {
    if (twol < 0 || twol > 10 || (twol % 2) != 0) return false;
    t = So4HostTables();
    const int ldim = twol + 1;
    t.twol = twol;
    t.ldim = ldim;
    auto I3 = [ldim](int a, int b, int c) { return (a * ldim + b) * ldim + c; };

    t.idxcg_block.assign(ldim * ldim * ldim, 0);
    t.idxz_block.assign(ldim * ldim * ldim, 0);
    int cnt = 0;
    for (int l1 = 0; l1 < ldim; l1++)
        for (int l2 = 0; l2 <= l1; l2++)
            for (int l = l1 - l2; l <= std::min(twol, l1 + l2); l += 2) {
                t.idxcg_block[I3(l1, l2, l)] = cnt;
                cnt += (l1 + 1) * (l2 + 1);
            }
    t.ncg = cnt;
    clebsch_gordan(twol, t.cglist);

    t.idxu_block.resize(ldim);
    t.hoff.resize(ldim + 1);
    cnt = 0;
    int hc = 0;
    for (int l = 0; l < ldim; l++) {
        t.idxu_block[l] = cnt;
        cnt += (l + 1) * (l + 1);
        t.hoff[l] = hc;
        hc += (l / 2 + 1) * (l + 1);
    }
    t.hoff[ldim] = hc;
    t.nu = cnt;
    t.nhalf = hc;
    t.lnorm.resize(ldim);
    for (int l = 0; l < ldim; l++) t.lnorm[l] = normalize_u ? 4.0 * M_PI / std::sqrt((double)(l + 1)) : 1.0;
    for (int l = 0; l < ldim; l++)
        for (int mb = 0; 2 * mb <= l; mb++)
            for (int ma = 0; ma <= l; ma++) {
                t.h_level.push_back(l);
                double w = 1.0;
                if (2 * mb == l) w = (ma < mb) ? 1.0 : (ma == mb ? 0.5 : 0.0);
                t.h_weight.push_back(w);
                t.h_diag.push_back(ma == mb ? 1 : 0);
            }

    // idxb, SNAP.py:1343-1356
    for (int l1 = 0; l1 < ldim; l1++)
        for (int l2 = 0; l2 <= l1; l2++)
            for (int l = l1 - l2; l <= std::min(twol, l1 + l2); l += 2)
                if (l >= l1) {
                    t.idxb.push_back(l1);
                    t.idxb.push_back(l2);
                    t.idxb.push_back(l);
                }
    t.nb = (int)t.idxb.size() / 3;
    t.b_brow0.assign(t.nb, 0);
    t.nbrow = 0;
    for (int b = 0; b < t.nb; b++) {
        t.b_brow0[b] = t.nbrow;
        t.nbrow += t.idxb[3 * b + 2] / 2 + 1;
    }

    // idxz, SNAP.py:1358-1416, plus the block list
    struct Blk {
        int l1, l2, l, zoff;
    };
    std::vector<Blk> blks;
    std::vector<int> zraw;  // per jjz: u1, u2, cga, cgb, na, nb, l1, l2
    cnt = 0;
    for (int l1 = 0; l1 < ldim; l1++)
        for (int l2 = 0; l2 <= l1; l2++)
            for (int l = l1 - l2; l <= std::min(twol, l1 + l2); l += 2) {
                t.idxz_block[I3(l1, l2, l)] = cnt;
                blks.push_back({l1, l2, l, cnt});
                for (int mb = 0; 2 * mb <= l; mb++)
                    for (int ma = 0; ma <= l; ma++) {
                        int ma1min = std::max(0, (2 * ma - l - l2 + l1) / 2);
                        int ma2max = (2 * ma - l - (2 * ma1min - l1) + l2) / 2;
                        int na = std::min(l1, (2 * ma - l + l2 + l1) / 2) - ma1min + 1;
                        int mb1min = std::max(0, (2 * mb - l - l2 + l1) / 2);
                        int mb2max = (2 * mb - l - (2 * mb1min - l1) + l2) / 2;
                        int nbb = std::min(l1, (2 * mb - l + l2 + l1) / 2) - mb1min + 1;
                        int cgb0 = t.idxcg_block[I3(l1, l2, l)];
                        zraw.push_back(t.idxu_block[l1] + (l1 + 1) * mb1min + ma1min);
                        zraw.push_back(t.idxu_block[l2] + (l2 + 1) * mb2max + ma2max);
                        zraw.push_back(cgb0 + ma1min * (l2 + 1) + ma2max);
                        zraw.push_back(cgb0 + mb1min * (l2 + 1) + mb2max);
                        zraw.push_back(na);
                        zraw.push_back(nbb);
                        zraw.push_back(l1);
                        zraw.push_back(l2);
                        cnt++;
                    }
            }
    t.nz = cnt;
    // schedule Z elements by decreasing cost na*nb so that the 64 lanes of a
    // wavefront run loops of similar length
    std::vector<int> order(t.nz);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return zraw[8 * a + 4] * zraw[8 * a + 5] > zraw[8 * b + 4] * zraw[8 * b + 5];
    });
    t.zsched.resize(6 * (size_t)t.nz);
    for (int k = 0; k < t.nz; k++) {
        int z = order[k];
        const int *r = &zraw[8 * (size_t)z];
        t.zsched[6 * k + 0] = z;
        t.zsched[6 * k + 1] = r[0];
        t.zsched[6 * k + 2] = r[1];
        t.zsched[6 * k + 3] = r[2];
        t.zsched[6 * k + 4] = r[3];
        // na | nb<<4 | (l1+1)<<8 | (l2+1)<<12 | l2<<16
        t.zsched[6 * k + 5] = r[4] | (r[5] << 4) | ((r[6] + 1) << 8) | ((r[7] + 1) << 12) | (r[7] << 16);
    }

    // b -> block of term 1
    t.b_zoff.resize(t.nb);
    t.b_level.resize(t.nb);
    for (int b = 0; b < t.nb; b++) {
        int l1 = t.idxb[3 * b], l2 = t.idxb[3 * b + 1], l = t.idxb[3 * b + 2];
        t.b_zoff[b] = t.idxz_block[I3(l1, l2, l)];
        t.b_level[b] = l;
    }

    // Targets: the three terms of compute_dbidrj for b = (l1, l2, l) use the
    // blocks (l1,l2,l) with factor 1, (l,l2,l1) with (l+1)/(l1+1) and
    // (l,l1,l2) with (l+1)/(l2+1).
    std::vector<std::vector<std::pair<int, double>>> tg(blks.size());
    auto find_blk = [&](int a, int b, int c) {
        for (size_t k = 0; k < blks.size(); k++)
            if (blks[k].l1 == a && blks[k].l2 == b && blks[k].l == c) return (int)k;
        return -1;
    };
    for (int b = 0; b < t.nb; b++) {
        int l1 = t.idxb[3 * b], l2 = t.idxb[3 * b + 1], l = t.idxb[3 * b + 2];
        int k1 = find_blk(l1, l2, l), k2 = find_blk(l, l2, l1), k3 = find_blk(l, l1, l2);
        if (k1 < 0 || k2 < 0 || k3 < 0) return false;
        tg[k1].push_back({b, 1.0});
        tg[k2].push_back({b, (l + 1) / (l1 + 1.0)});
        tg[k3].push_back({b, (l + 1) / (l2 + 1.0)});
    }
    // merge duplicate (block, b) targets and sort blocks by level
    std::vector<int> border(blks.size());
    std::iota(border.begin(), border.end(), 0);
    std::stable_sort(border.begin(), border.end(), [&](int a, int b) { return blks[a].l < blks[b].l; });
    t.nblk = (int)blks.size();
    t.blk_level_start.assign(ldim + 1, 0);
    for (int k = 0; k < t.nblk; k++) {
        int bi = border[k];
        std::vector<std::pair<int, double>> m;
        for (auto &p : tg[bi]) {
            bool found = false;
            for (auto &q : m)
                if (q.first == p.first) {
                    q.second += p.second;
                    found = true;
                }
            if (!found) m.push_back(p);
        }
        if (m.size() > 3) return false;
        t.blk_zoff.push_back(blks[bi].zoff);
        for (int s = 0; s < 3; s++) {
            t.blk_tb.push_back(s < (int)m.size() ? m[s].first : -1);
            t.blk_tf.push_back(s < (int)m.size() ? m[s].second : 0.0);
        }
        t.blk_level_start[blks[bi].l + 1]++;
    }
    for (int l = 0; l < ldim; l++) t.blk_level_start[l + 1] += t.blk_level_start[l];

    // Z row groups for the atom-per-lane kernel.  Row (l1, l2, l, mb) sums
    // over the pairs (mb1, mb2) with mb1 + mb2 = K = mb + (l1 + l2 - l)/2;
    // all rows of one (l1, l2) with the same K share those pairs (and so
    // the products U^{l1}[mb1][ma1] U^{l2}[mb2][ma2]), one row per l at
    // most.  A group is (l1, l2, K) with the rows of its active l.
    for (int l1 = 0; l1 < ldim; l1++)
        for (int l2 = 0; l2 <= l1; l2++) {
            const int lmin = l1 - l2, lmax = std::min(twol, l1 + l2);
            for (int K = 0; K <= l1 + l2; K++) {
                std::vector<int> g(ZG_STRIDE, -1);
                int mask = 0, cost = 0;
                const int lo = std::max(0, K - l2), hi = std::min(l1, K);
                if (hi < lo) continue;
                for (int l = lmin, j = 0; l <= lmax; l += 2, j++) {
                    const int s2 = (l1 + l2 - l) / 2, mb = K - s2;
                    g[ZG_CGBLK + j] = t.idxcg_block[I3(l1, l2, l)];
                    if (mb < 0 || 2 * mb > l) continue;
                    mask |= 1 << j;
                    g[ZG_ZROW + j] = t.idxz_block[I3(l1, l2, l)] + mb * (l + 1);
                    g[ZG_UROW + j] = t.hoff[l] + mb * (l + 1);
                    for (int b = 0; b < t.nb; b++)
                        if (t.idxb[3 * b] == l1 && t.idxb[3 * b + 1] == l2 && t.idxb[3 * b + 2] == l)
                            g[ZG_BROW + j] = t.b_brow0[b] + mb;
                    cost += (l + 1) * (hi - lo + 1) * (l1 + 1);
                }
                if (!mask) continue;
                g[0] = l1;
                g[1] = l2;
                g[2] = K;
                g[3] = lo;
                g[4] = hi - lo + 1;
                g[5] = t.hoff[l1];
                g[6] = t.hoff[l2];
                g[7] = mask;
                g[8] = cost + (hi - lo + 1) * (l1 + l2 + 2) * 4;
                t.zgroups.insert(t.zgroups.end(), g.begin(), g.end());
                t.nzgroups++;
            }
        }

    // init_rootpqarray, SNAP.py:861-868
    t.rootpq.assign(ldim * ldim, 0.0);
    for (int p = 1; p < ldim; p++)
        for (int q = 1; q < ldim; q++) t.rootpq[p * ldim + q] = std::sqrt((double)p / (double)q);
    return true;
}

std::vector<int> chunk_zgroups(const So4HostTables &t, int nchunks) {
    long total = 0;
    for (int r = 0; r < t.nzgroups; r++) total += t.zgroups[ZG_STRIDE * r + 8];
    std::vector<int> start(1, 0);
    long acc = 0;
    for (int r = 0; r < t.nzgroups; r++) {
        acc += t.zgroups[ZG_STRIDE * r + 8];
        if ((int)start.size() < nchunks && acc * nchunks >= total * (long)start.size() && r + 1 < t.nzgroups)
            start.push_back(r + 1);
    }
    while ((int)start.size() < nchunks) start.push_back(t.nzgroups);
    start.push_back(t.nzgroups);
    return start;
}

}  // namespace pxff
