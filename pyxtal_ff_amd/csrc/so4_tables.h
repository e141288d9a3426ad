// Constant index tables of the SO(4)/SNAP bispectrum for one 2J, built on
// the host and uploaded once per (device, 2J).
//
// The tables follow the reference's index construction
// (pyxtal_ff/descriptors/SNAP.py:1296-1416): idxcg_block / cglist
// (Varshalovich 8.2.1(3), SNAP.py:442-510, with the reference's factorial
// table, angular_momentum.py:9-182), idxu_block, u_norm, idxb (l1 >= l2,
// l >= l1) and idxz (l1, l2, l, ma1min, ma2max, na, mb1min, mb2max, nb).
// On top of them the device path adds: the half-U layout (rows mb <= l/2),
// a cost-sorted Z schedule, and per Z block the (b, factor) targets that
// the three terms of compute_dbidrj (SNAP.py:1143-1257) attach to it.
#pragma once
#include <vector>

namespace pxff {

struct So4HostTables {
    int twol = 0, ldim = 0;
    int nu = 0, nhalf = 0, nz = 0, nb = 0, ncg = 0, nblk = 0;
    std::vector<double> cglist;          // [ncg]
    std::vector<int> idxcg_block;        // [ldim^3]
    std::vector<int> idxu_block;         // [ldim]
    std::vector<int> hoff;               // [ldim+1] half-U level offsets
    std::vector<int> idxz_block;         // [ldim^3]
    std::vector<int> idxb;               // [nb*3]
    // Z schedule: per Z element (in cost order) {jjz, u1, u2, cga, cgb, pack}
    std::vector<int> zsched;             // [nz*6]
    // blocks sorted by level: zoff, level, targets
    std::vector<int> blk_level_start;    // [ldim+1]
    std::vector<int> blk_zoff;           // [nblk]
    std::vector<int> blk_tb;             // [nblk*3] target b (-1 = none)
    std::vector<double> blk_tf;          // [nblk*3] target factors
    std::vector<int> b_zoff;             // [nb] Z offset of block (l1,l2,l) of b
    std::vector<int> b_level;            // [nb] l of b
    std::vector<double> rootpq;          // [ldim*ldim] sqrt(p/q)
    std::vector<double> lnorm;           // [ldim]
    std::vector<int> h_level;            // [nhalf] level of each half element
    std::vector<double> h_weight;        // [nhalf] half weight (1, 0.5 or 0)
    std::vector<int> h_diag;             // [nhalf] 1 if mb == ma
    // B row partials: row (b, mb), mb = 0..l_b/2, is b_brow0[b] + mb
    std::vector<int> b_brow0;            // [nb]
    int nbrow = 0;
    // Z row groups (l1, l2, K) for the atom-per-lane kernel, ZG_STRIDE ints
    // each: {l1, l2, K, mb1lo, nbb, hoff[l1], hoff[l2], active-l mask, cost,
    // Z offset of the row of each l (ZG_ZROW + j, -1 if inactive), CG block
    // of each l (ZG_CGBLK + j), B row of each l (ZG_BROW + j, -1 if the row
    // is inactive or (l1, l2, l) is not a B component), half-U offset of
    // row (l, mb) (ZG_UROW + j)}, l = l1 - l2 + 2 j
    std::vector<int> zgroups;            // [nzgroups*ZG_STRIDE]
    int nzgroups = 0;
};

constexpr int ZG_MAXL = 6;  // at most six l per (l1, l2) for 2J <= 10
constexpr int ZG_ZROW = 9;
constexpr int ZG_CGBLK = ZG_ZROW + ZG_MAXL;
constexpr int ZG_BROW = ZG_CGBLK + ZG_MAXL;
constexpr int ZG_UROW = ZG_BROW + ZG_MAXL;
constexpr int ZG_STRIDE = ZG_UROW + ZG_MAXL + 3;

// Split the Z row groups into nchunks contiguous chunks of about equal cost.
std::vector<int> chunk_zgroups(const So4HostTables &t, int nchunks);

// Build the tables.  Returns false for unsupported 2J.
bool build_so4_tables(int twol, int normalize_u, So4HostTables &t);

}  // namespace pxff
