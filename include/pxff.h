/*
 * pxff.h -- C ABI of the MI355X (gfx950) descriptor hot path of PyXtal_FF.
 *
 * Every pointer argument named d_* is a caller-owned DEVICE pointer; every
 * function enqueues work on the given HIP stream (`stream` is a hipStream_t
 * passed as void*) and returns 0 on success or a negative status, with a
 * message available from pxff_last_error() (thread-local).  No function
 * allocates device memory except the per-2J constant tables, which are
 * built once per (device, 2J) under a mutex and kept until process exit.
 * Arrays are row-major; floating point is FP64; complex arrays are
 * interleaved (re, im) doubles.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   pyxtal_ff/descriptors/SNAP.py:240-317   build_neighbor_list (ASE NeighborList + regrouping)
 *   pyxtal_ff/descriptors/SNAP.py:1289-1762 get_bispectrum_components (numba)
 *   pyxtal_ff/descriptors/SO4.py:1256-1730  get_bispectrum_components (numba, Z weights)
 *   pyxtal_ff/descriptors/SO3.py:187-291,504-694 SO3.calculate / compute_cs / compute_dcs
 *   pyxtal_ff/utilities/__init__.py:226-235 per-atom rdxdr reduction
 *   pyxtal_ff/models/neuralnetwork.py:484-497,657-677 force/stress contraction over dxdr/rdxdr
 */
#ifndef PXFF_H
#define PXFF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXFF_OK 0
#define PXFF_EINVAL -1   /* bad argument */
#define PXFF_EHIP -2     /* HIP runtime error */
#define PXFF_ENOTIMPL -3 /* unsupported parameter (e.g. odd 2J) */

/* Launch-parameter knobs (process-wide), set by the autotuner:
 *   PXFF_TUNE_Z_CHUNKS              Z rows split into this many wavefront chunks per 64-atom tile (1..256)
 *   PXFF_TUNE_ACCUM_WAVES_PER_CU    persistent grid of the per-centre kernels = CUs x this (1..256)
 *   PXFF_TUNE_IMAGE_WAVES_PER_BLOCK wavefronts per workgroup of the per-image adjoint kernel (1, 2, 4)
 *   PXFF_TUNE_ACCUM_LDSADD          per-centre U_tot sums: 0 = per-group LDS partials, 1 = one LDS array
 *                                   with LDS atomic adds (issued by one wavefront in a fixed order),
 *                                   2 (3) = two (three) groups per partial array, added in two
 *                                   (three) ordered passes, 4 (default) = each lane's sums over
 *                                   all of a centre's images in registers, reduced across the
 *                                   lane groups once per centre (2J <= 8; 2J = 10 uses 0) */
#define PXFF_TUNE_Z_CHUNKS 0
#define PXFF_TUNE_ACCUM_WAVES_PER_CU 1
#define PXFF_TUNE_IMAGE_WAVES_PER_BLOCK 2
#define PXFF_TUNE_ACCUM_LDSADD 3
#define PXFF_TUNE_COUNT 4
int pxff_tune_set(int knob, int value);
int pxff_tune_get(int knob); /* -1 for an unknown knob */

/* Bytes of the caller-owned buffers that are neither inputs nor outputs of
 * their entry point (the ABI allocates nothing else on the device apart
 * from the immutable per-parameter tables it builds once, under a mutex):
 *   PXFF_WS_NLIST        the int32 workspace of pxff_nlist_count / _fill
 *                        (bin_count, bin_start, atom_bin, wrap, bin_atoms),
 *                        n0 = natoms, n1 = nbins_total; params unused
 *   PXFF_WS_SO4_Z        d_z of pxff_so4_zb (Z and the B row partials),
 *                        params = pxff_so4_params, n0 = natoms
 *   PXFF_WS_SO4_TANGENT  d_dutot of pxff_so4_tangent, params =
 *                        pxff_so4_params, n0 = natoms
 *   PXFF_WS_SO3_PAIRS    d_work of pxff_so3_ctot_pairs, n1 = image capacity
 *   PXFF_WS_SO3_TABLE    d_table of pxff_so3_radial_table (its d_scratch is
 *                        the same size), params = pxff_so3_params
 *   PXFF_WS_SO4_UI       d_stage of pxff_so4_ui_staged, n1 = image capacity
 *   PXFF_WS_NLIST_SLOTS  the slot arrays of pxff_nlist_count_slots / _fill_slots
 *                        (44 bytes per slot entry), n0 = natoms, n1 = slot
 * The launch knobs below are the only other state: process-wide atomics
 * read once per launch, so calls on several threads and streams are
 * reentrant. */
#define PXFF_WS_NLIST 0
#define PXFF_WS_SO4_Z 1
#define PXFF_WS_SO4_TANGENT 2
#define PXFF_WS_SO3_PAIRS 3
#define PXFF_WS_SO3_TABLE 4
#define PXFF_WS_SO4_UI 5
#define PXFF_WS_NLIST_SLOTS 6
int pxff_workspace_size(int kind, const void *params, int64_t n0, int64_t n1, int64_t *bytes);

/* Copy the calling thread's last error message into buf (NUL-terminated). */
int pxff_last_error(char *buf, size_t len);

/* Library version string. */
const char *pxff_version(void);

/* ------------------------------------------------------------------ *
 * Neighbour list (replaces SNAP.py:240-317 / SO3.py:317-376 and the   *
 * ASE NeighborList call at SNAP.py:250-252).                          *
 *                                                                     *
 * A batch of nstruct structures; atoms of structure s are             *
 * [atom_start[s], atom_start[s+1]).  Per structure: cell [9] (rows =  *
 * lattice vectors), pbc [3], and a host-computed bin geometry         *
 * geom [12] int32 = {nb0, nb1, nb2, S0, S1, S2, bin_offset, 0...}.     *
 * Images (i, j, T) with |pos_j + T.cell - pos_i| < rc, excluding       *
 * (i, T=0), are produced in the order (i, j, T0, T1, T2).             *
 * ------------------------------------------------------------------ */

/* Phase 1: bin atoms and count images per centre.
 *   d_bin_count [nbins_total] and d_bin_start [nbins_total+1] are int32
 *   workspace; d_atom_bin [natoms] int32, d_wrap [natoms*3] int32 and
 *   d_bin_atoms [natoms] int32 are workspace; d_img_count [natoms] int64 out. */
int pxff_nlist_count(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                     const int64_t *d_atom_start, const double *d_cells, const double *d_inv_cells,
                     const int32_t *d_pbc, const int32_t *d_geom, int64_t nbins_total, double rc,
                     int32_t *d_bin_count, int32_t *d_bin_start, int32_t *d_atom_bin, int32_t *d_wrap,
                     int32_t *d_bin_atoms, int64_t *d_img_count, void *stream);

/* Phase 1 with the images kept: pxff_nlist_count, and every centre with at
 * most `slot` images (1..128) also ranks them and writes them in (j, T)
 * order into its slot -- d_slot_j [natoms*slot] int64, d_slot_T
 * [natoms*slot*3] int32, d_slot_r [natoms*slot*3] double, at [i*slot,
 * i*slot + count) -- with its seq-row count in d_row_count [natoms].
 * pxff_nlist_fill_slots then copies the slots into place and fills the other
 * centres as pxff_nlist_fill does, so most centres are enumerated once per
 * build instead of twice.  Workspace: pxff_workspace_size
 * PXFF_WS_NLIST_SLOTS (n0 = natoms, n1 = slot). */
int pxff_nlist_count_slots(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                           const int64_t *d_atom_start, const double *d_cells, const double *d_inv_cells,
                           const int32_t *d_pbc, const int32_t *d_geom, int64_t nbins_total, double rc,
                           int32_t *d_bin_count, int32_t *d_bin_start, int32_t *d_atom_bin, int32_t *d_wrap,
                           int32_t *d_bin_atoms, int64_t *d_img_count, int slot, int64_t *d_slot_j,
                           int32_t *d_slot_T, double *d_slot_r, int64_t *d_row_count, void *stream);

/* Phase 2: fill images (sorted per centre by (j, T)) and count seq rows.
 *   d_img_start [natoms+1] int64 = exclusive scan of d_img_count.
 *   Outputs: d_img_atom [nimg] int64 (centre), d_img_j [nimg] int64,
 *   d_img_T [nimg*3] int32, d_img_r [nimg*3] double, d_row_count [natoms]
 *   int64 (unique j plus the (i,i) row). */
int pxff_nlist_fill(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                    const int64_t *d_atom_start, const double *d_cells, const int32_t *d_pbc,
                    const int32_t *d_geom, double rc, const int32_t *d_bin_start, const int32_t *d_atom_bin,
                    const int32_t *d_wrap, const int32_t *d_bin_atoms, const int64_t *d_img_start,
                    int64_t *d_img_atom, int64_t *d_img_j, int32_t *d_img_T, double *d_img_r,
                    int64_t *d_row_count, void *stream);

/* Phase 2 after pxff_nlist_count_slots (same slot and slot arrays, and its
 * d_img_count): centres whose images are in their slot are copied into
 * place (their seq-row counts are written already), the others filled as
 * by pxff_nlist_fill.  Same output as pxff_nlist_count + pxff_nlist_fill. */
int pxff_nlist_fill_slots(const double *d_pos, int64_t natoms, const int64_t *d_atom_struct,
                          const int64_t *d_atom_start, const double *d_cells, const int32_t *d_pbc,
                          const int32_t *d_geom, double rc, const int32_t *d_bin_start, const int32_t *d_atom_bin,
                          const int32_t *d_wrap, const int32_t *d_bin_atoms, const int64_t *d_img_start,
                          const int64_t *d_img_count, int slot, const int64_t *d_slot_j, const int32_t *d_slot_T,
                          const double *d_slot_r, int64_t *d_img_atom, int64_t *d_img_j, int32_t *d_img_T,
                          double *d_img_r, int64_t *d_row_count, void *stream);

/* Phase 3: reverse images.  d_img_rev [nimg] int64 = index of image
 * (j, i, -T) for image (i, j, T) -- the list is exactly symmetric, which
 * the deterministic force gather relies on.  nimg may be an upper bound
 * (the image capacity): images at or beyond d_img_start[natoms] are
 * skipped.  With d_flag, each reverse is searched by (j, T) and a missing
 * one (cannot happen for this construction) sets d_flag[0] (int32, zeroed
 * by the call).  With d_flag NULL nothing is checked and the symmetry is
 * used instead: the run of images (i -> j) and the run (j -> i) have the
 * same length and opposite order in T, so only the start of the run
 * (j -> i) is searched, by neighbour index (one launch fewer; for a
 * symmetric list the same indices as the checked call). */
int pxff_nlist_rev(int64_t natoms, int64_t nimg, const int64_t *d_img_start, const int64_t *d_img_atom,
                   const int64_t *d_img_j, const int32_t *d_img_T, int64_t *d_img_rev, int32_t *d_flag,
                   void *stream);

/* Phase 4 (only for the reference's row layout): seq rows.  d_row_start
 * [natoms+1] int64 = exclusive scan of d_row_count, nrows =
 * d_row_start[natoms].  Outputs: d_seq [nrows*2] int64 (the reference's
 * seq), d_row_img [nrows+1] int64 (images of row r are [row_img[r],
 * row_img[r+1])), d_img_row [nimg] int64, d_nsite [natoms] int64 (row of
 * (i,i)), d_row_rev [nrows] int64 (row (j,i) of row (i,j)); a missing
 * reverse row sets d_flag[0]. */
int pxff_nlist_rows(int64_t natoms, int64_t nimg, int64_t nrows, const int64_t *d_img_start,
                    const int64_t *d_img_j, const int64_t *d_row_start, int64_t *d_seq, int64_t *d_row_img,
                    int64_t *d_img_row, int64_t *d_nsite, int64_t *d_row_rev, int32_t *d_flag, void *stream);

/* Exclusive scan: d_out [n+1], d_out[n] = total. */
int pxff_scan_i64(const int64_t *d_in, int64_t n, int64_t *d_out, void *stream);

/* Exclusive scan clamped to cap (every d_out[k] <= cap) with the unclamped
 * total in d_total[0]: image offsets for arrays of cap entries, sized before
 * the total is known on the host (no synchronisation between the count and
 * fill phases).  A total above cap means the list was cut; the caller
 * re-runs the scan and fill phases with a larger capacity. */
int pxff_scan_i64_capped(const int64_t *d_in, int64_t n, int64_t cap, int64_t *d_out, int64_t *d_total,
                         void *stream);

/* ------------------------------------------------------------------ *
 * SO(4) / SNAP bispectrum (replaces get_bispectrum_components,         *
 * SNAP.py:1289-1762; SO4.py:1256-1730 is the same with rfac0 = 1 and   *
 * Z weights).                                                          *
 * Layout: per atom a "half" U/Y array of nhalf complex (levels l, rows *
 * mb <= l/2, columns ma), Z [nz] complex in the reference's idxz order, *
 * B [nb] real in the reference's idxb order.  U_tot and Z are stored    *
 * tile-major (64-atom tiles, atom index innermost).                     *
 * ------------------------------------------------------------------ */
typedef struct {
    int twol;      /* 2J = round(2*lmax); even, 0..10 */
    double rcut;   /* cutoff radius */
    double rfac0;  /* psi = rfac0*pi*r/rcut (1.0 for SO4.py) */
    int cutoff_id; /* 1 cosine, 2 tanh, 3..6 poly1..4, 7 exp (SNAP.py:141-170) */
    int normalize_u; /* u_norm = 4pi/sqrt(l+1) (SNAP.py:1333-1341) */
} pxff_so4_params;

/* Table sizes for 2J: nu (full U), nhalf (half U), nz, nb. */
int pxff_so4_sizes(int twol, int *nu, int *nhalf, int *nz, int *nb);

/* U_tot (half rows, self term and u_norm included) per centre, stored
 * tile-major for the Z kernel: element h of atom i at
 * [(i/64)*nhalf + h]*64 + i%64 (complex); d_utot must hold ceil(natoms/64)*64
 * atoms (the padding lanes of the last tile are set to zero).
 *   d_img_start [natoms+1], d_img_j [nimg], d_img_r [nimg*3]; d_atom_w
 *   [natoms] = neighbour weight of each atom (pair weight of image = w[j]);
 *   d_site_w [natoms] = self weight.  d_work_order [natoms] int64 or NULL:
 *   centres in the order they are scheduled (neighbour-count sorted). */
int pxff_so4_ui(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_j,
                const double *d_img_r, const double *d_atom_w, const double *d_site_w,
                const int64_t *d_work_order, double *d_utot, void *stream);

/* pxff_so4_ui with the per-image values of the sweeps (conj(a), conj(b),
 * w fc) computed first by a kernel of their own into d_stage (6 doubles per
 * image, pxff_workspace_size PXFF_WS_SO4_UI, n1 = cap), for images
 * e < min(d_img_start[natoms], cap): the sweep kernel then reads one record
 * per image instead of evaluating cutoff and Cayley-Klein parameters
 * between its sweeps (same construction as pxff_so4_ui, SNAP.py:1480-1500;
 * equal to it to rounding).  cap must be at least the image count.  With
 * a summation mode other than the register sums (pxff_tune_set
 * PXFF_TUNE_ACCUM_LDSADD != 4) or 2J > 8 it is pxff_so4_ui. */
int pxff_so4_ui_staged(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_j,
                       const double *d_img_r, const double *d_atom_w, const double *d_site_w,
                       const int64_t *d_work_order, int64_t cap, double *d_stage, double *d_utot, void *stream);

/* Z (and B) from the tile-major U_tot (compute_zi + compute_bi,
 * SNAP.py:1050-1141).  d_z out, pxff_so4_zlen(2J, natoms) doubles: Z
 * tile-major like U_tot (element e of atom i at [(i/64)*nz + e]*64 + i%64,
 * complex; padding lanes of the last tile are zero), followed by the B row
 * partials the Z kernel writes when d_b != NULL; d_b [natoms*nb] out (or
 * NULL). */
int pxff_so4_zb(const pxff_so4_params *p, int64_t natoms, const double *d_utot, double *d_z, double *d_b,
                void *stream);

/* Length in doubles of the d_z buffer of pxff_so4_zb for natoms atoms. */
int pxff_so4_zlen(int twol, int64_t natoms, int64_t *len);

/* B alone (compute_bi, SNAP.py:1101-1141) from the tile-major U_tot and Z
 * (the layouts pxff_so4_zb writes): d_b [natoms*nb] out.  Same values as
 * pxff_so4_zb with d_b != NULL (which sums the rows in another order). */
int pxff_so4_b(const pxff_so4_params *p, int64_t natoms, const double *d_utot, const double *d_z, double *d_b,
               void *stream);

/* Adjoint, step 1: Y [natoms*nhalf*2] = sum over blocks of the dE/dB-weighted
 * Z (half weights, factor 2 and u_norm folded in).  d_beta [natoms*nb]. */
int pxff_so4_yi(const pxff_so4_params *p, int64_t natoms, const double *d_z, const double *d_beta,
                double *d_y, void *stream);

/* Adjoint, step 2: dE/dr per image [nimg*3] from Y (one dU sweep; dB/dR is
 * never materialised).  d_img_atom [nimg] = centre of each image. */
int pxff_so4_deidrj(const pxff_so4_params *p, int64_t nimg, const int64_t *d_img_atom, const int64_t *d_img_j,
                    const double *d_img_r, const double *d_atom_w, const double *d_y, double *d_dedr,
                    void *stream);

/* pxff_so4_deidrj over an image capacity: images e < min(cap, *d_count)
 * are processed, the count being read on the device (d_count = the capped
 * image offsets' last entry, d_img_start + natoms), so the launch needs no
 * host copy of the image count (HIP graph capture of a whole E+F step).
 * d_dedr [cap*3]; entries at and beyond the count are not written. */
int pxff_so4_deidrj_counted(const pxff_so4_params *p, int64_t cap, const int64_t *d_count,
                            const int64_t *d_img_atom, const int64_t *d_img_j, const double *d_img_r,
                            const double *d_atom_w, const double *d_y, double *d_dedr, void *stream);

/* Forces and virial from dE/dr per image (deterministic gather via the
 * reverse-image index).  d_grad [natoms*3] = dE/dR_i (= -F_i);
 * d_virial [natoms*9] (or NULL) = per-centre sum_e r_e (x) dE/dr_e. */
int pxff_so4_gather(int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_rev, const double *d_img_r,
                    const double *d_dedr, double *d_grad, double *d_virial, void *stream);

/* pxff_so4_gather with the sign of the forces: d_forces [natoms*3] = F_i =
 * -dE/dR_i (bit-identical to negating d_grad); d_virial as above.  The
 * reference's forces are -dE/dR (neuralnetwork.py:491). */
int pxff_so4_forces(int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_rev, const double *d_img_r,
                    const double *d_dedr, double *d_forces, double *d_virial, void *stream);

/* pxff_so4_forces plus the total energy in the same launch: d_energy[0] =
 * sum of d_e_atom[0, n_e) (the structure's energy, the sum of the atomic
 * network energies, neuralnetwork.py:603-677), bit-identical to
 * pxff_segment_sum over the one segment [0, n_e); d_forces / d_virial as
 * pxff_so4_forces. */
int pxff_so4_forces_energy(int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_rev,
                           const double *d_img_r, const double *d_dedr, double *d_forces, double *d_virial,
                           const double *d_e_atom, int64_t n_e, double *d_energy, void *stream);

/* Forward-mode (tangent) of B along per-image displacements dr_e
 * [nimg*3]: dB [natoms*nb] = sum_e dB/dr_e . dr_e.  Used by the double
 * backward of the adjoint (force-matched training).  d_z is the tile-major Z
 * of pxff_so4_zb; d_dutot is workspace [ceil(natoms/64)*64*nhalf*2]. */
int pxff_so4_tangent(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_j,
                     const double *d_img_r, const double *d_atom_w, const double *d_dr, const double *d_z,
                     double *d_dutot, double *d_db, void *stream);

/* pxff_so4_tangent without the d_dutot workspace argument, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving or
 * training loops, harnesses that forbid hipMalloc).  The workspace is a
 * static device array of 32 MiB per GPU (held until the process ends), so
 * ceil(natoms/64)*64*nhalf*16 bytes must fit in it (about 13,500 atoms at
 * 2J=8; larger inputs return PXFF_EINVAL with a message).  The 2J tables must
 * already exist on the device (built by an earlier call of an allocating SO4
 * entry point with the same 2J and normalisation, e.g. pxff_so4_tangent);
 * otherwise PXFF_EINVAL.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams) race.  Same results as
 * pxff_so4_tangent. */
int pxff_so4_tangent_noalloc(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                             const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w, const double *d_dr,
                             const double *d_z, double *d_db, void *stream);

/* Atom-major copy of the tile-major Z of pxff_so4_zb: d_za [natoms*nz*2],
 * element e of atom i at i*nz + e (complex). */
int pxff_so4_z_atom_major(const pxff_so4_params *p, int64_t natoms, const double *d_z, double *d_za, void *stream);

/* Plain (materialised) path, step 1: dB/dr per image [nimg*nb*3] from the
 * atom-major Z (pxff_so4_z_atom_major). */
int pxff_so4_dbdr(const pxff_so4_params *p, int64_t nimg, const int64_t *d_img_atom, const int64_t *d_img_j,
                  const double *d_img_r, const double *d_atom_w, const double *d_z, double *d_dbdr, void *stream);

/* Plain dB/dr per image [nimg*nb*3]: the same values as pxff_so4_dbdr, one
 * workgroup per centre over its images img_start[c] .. img_start[c+1]-1 (the
 * image count is read on the device).  d_z: the tile-major Z of
 * pxff_so4_zb (z_atom_major = 0) or its atom-major copy from
 * pxff_so4_z_atom_major (z_atom_major = 1; contiguous block rows, far fewer
 * cache lines per centre).  Replaces the dblist part of compute_dbidrj /
 * get_bispectrum_components (SNAP.py:1143-1257, 1428-1548). */
int pxff_so4_dbdr_centres(const pxff_so4_params *p, int64_t natoms, const int64_t *d_img_start,
                          const int64_t *d_img_j, const double *d_img_r, const double *d_atom_w, const double *d_z,
                          int z_atom_major, double *d_dbdr, void *stream);

/* Plain path, step 2: reference layout rows.  dxdr [nrows*nb*3] (rows (i,j)
 * = sum over images, row (i,i) = -sum of the others), and, if d_rdxdr is not
 * NULL, per-row -bstress/V [nrows*nb*9] with bstress[nsite] += Ri (x) dB,
 * bstress[row] -= Rj (x) dB (SNAP.py:1458-1468).  d_inv_vol [natoms] = 1/V
 * of each atom's structure. */
int pxff_so4_rows(int64_t natoms, int nb, const int64_t *d_row_start, const int64_t *d_row_img,
                  const int64_t *d_seq, const int64_t *d_nsite, const double *d_pos, const double *d_img_r,
                  const double *d_dbdr, const double *d_inv_vol, double *d_dxdr, double *d_rdxdr, void *stream);

/* Plain path, force contraction (neuralnetwork.py:484-497, 657-677):
 * dE/dR_m = sum_{rows (i,m)} sum_b beta[i,b] dxdr[row,b,:], gathered per m
 * through the reverse-row index.  d_grad [natoms*3]. */
int pxff_dxdr_contract(int64_t natoms, int nb, const int64_t *d_row_start, const int64_t *d_seq,
                       const int64_t *d_row_rev, const double *d_beta, const double *d_dxdr, double *d_grad,
                       void *stream);

/* Transpose of pxff_dxdr_contract, for training on a materialised dxdr:
 * d_gbeta [natoms*nb], gbeta[i,b] = sum over centre i's rows (i,m) of
 * dxdr[row,b,:] . g[m,:] for d_g [natoms*3] -- the derivative of the force
 * contraction with respect to beta = dE/dx that the reference's loss
 * backpropagates (neuralnetwork.py:484-497, 499-545). */
int pxff_dxdr_contract_t(int64_t natoms, int nb, const int64_t *d_row_start, const int64_t *d_seq, const double *d_g,
                         const double *d_dxdr, double *d_gbeta, void *stream);

/* Per-atom rdxdr reduction (utilities/__init__.py:226-235): out [natoms*nb*6]
 * Voigt (xx, yy, zz, xy, xz, yz) of the sum of the centre's rows. */
int pxff_rdxdr_reduce(int64_t natoms, int nb, const int64_t *d_row_start, const double *d_rdxdr, double *d_out,
                      void *stream);

/* Segmented sum over contiguous rows: out [nseg*ncol], out[s][c] = sum of
 * in[r][c] for r in [d_seg_start[s], d_seg_start[s+1]) (d_seg_start [nseg+1]
 * int64).  Per-structure energies / virials of a batch (the reference sums
 * atoms per structure in neuralnetwork.py:603-677); deterministic order,
 * no atomics. */
int pxff_segment_sum(int64_t nseg, int ncol, const int64_t *d_seg_start, const double *d_in, double *d_out,
                     void *stream);

/* Per-element network energies and dE/dx (energy + force evaluation; no
 * parameter gradients).  Replaces the forward of the reference's per-element
 * networks and the descriptor gradient they feed into the force contraction
 * (neuralnetwork.py:80-200 architecture, :484-497 / :657-677 use,
 * :938-1029 drange scaling).  Atom i (element d_species[i] in [0, nelem))
 * uses x' = (x_i - dmin_e) * scale_e (d_dmin, d_scale [nelem*nin]) and the
 * layers of its element; layout per element (ints, stride 1 + 6*8):
 * nlayers (1..8), then per layer {nin, nout, act, w_off, b_off, wt_off}: W
 * [nout][nin] row-major at d_params + w_off, bias at d_params + b_off, the
 * transpose W^T [nin][nout] at d_params + wt_off, act 0 Linear, 1 Tanh,
 * 2 Sigmoid, 3 ReLU, 4 Softplus, 5 ELU.  Widths <= 64,
 * last width 1, nin <= 1024.  h_layout (host copy of d_layout) is
 * validated before launch.  Outputs: d_e [natoms] atomic energies, d_beta
 * [natoms*nin] = dE_i/dx_i.  An out-of-range species gives NaN. */
int pxff_mlp_energy_grad(int64_t natoms, int nin, int nelem, const int64_t *d_species, const int *h_layout,
                         const int *d_layout, const double *d_params, const double *d_dmin, const double *d_scale,
                         const double *d_x, double *d_e, double *d_beta, void *stream);

/* ------------------------------------------------------------------ *
 * SO(3) power spectrum (replaces SO3.calculate / compute_cs /          *
 * compute_dcs, SO3.py:187-291, 504-694).                               *
 * Host-precomputed: d_G [nmax*nquad] = g_n(r_q) (pi/NQ)(rc/2) r_q^2     *
 * exp(-alpha r_q^2) sqrt(1 - x_q^2) (SO3.py:466-502, 589-617), d_rq     *
 * [nquad] = rc (x_q + 1)/2, d_lnorm [lmax+1] = sqrt(2 sqrt2 pi /        *
 * sqrt(2l+1)) (SO3.py:209-210).  d_atom_z [natoms] int64 = atomic        *
 * numbers (neighbour weight Z_j, sign-flipped with weight_on when        *
 * Z_j != Z_i, SO3.py:349-354).                                          *
 * ctot layout per atom: [nmax][lmax+1][2 lmax+1] complex (m + lmax).    *
 * x layout: [ntri(n>=n') * (lmax+1) + l] (SO3.py:259).                  *
 * ------------------------------------------------------------------ */
typedef struct {
    int nmax;      /* 1..11 */
    int lmax;      /* 0..32 (device path: nmax*(lmax+1)*(2lmax+1) <= 4096) */
    double rcut;
    double alpha;
    int cutoff_id; /* 1 cosine, 2 tanh, 3..6 poly1..4, 7 exp, 8 unity */
    int nquad;     /* 10*(nmax+lmax+1) (SO3.py:495-502) */
    int weight_on;
} pxff_so3_params;

/* Per-centre expansion ctot [natoms*csz*2] (sum over the centre's images);
 * with d_dr [nimg*3] non-NULL, its tangent sum_e dc_e/dr_e . dr_e.
 * d_radial (plain expansion only; may be NULL): the radial integrals of
 * every image, [nimg][lmax+1][2][nmax] (I_nl, dI_nl/dR), for the adjoint
 * and plain per-image kernels to reuse. */
int pxff_so3_ctot(const pxff_so3_params *p, int64_t natoms, const int64_t *d_img_start, const int64_t *d_img_j,
                  const double *d_img_r, const int64_t *d_atom_z, const double *d_G, const double *d_rq,
                  const double *d_lnorm, const double *d_dr, double *d_out, double *d_radial, void *stream);

/* pxff_so3_ctot with the radial store (d_radial required, no tangent),
 * where the separations d_img_r are the neighbour list's own, so that image
 * e and its reverse d_img_rev[e] (pxff_nlist_rev) have bitwise the same |r|:
 * the radial integrals are evaluated once per image pair and written to
 * both rows (same values as pxff_so3_ctot).  With d_work (int64 workspace
 * of 256*64 + 1 + cap entries, cap = the image capacity of the arrays) the
 * pairs are first ordered by |r| (a counting sort into 256 bins, over 64
 * chunks of the images) so that the images a wavefront evaluates together
 * take the same branches; the values do not depend on it.  d_work = NULL:
 * the unsorted order.  d_table (non-NULL, with d_work): the integrals from
 * the table of pxff_so3_radial_table instead of the quadrature. */
int pxff_so3_ctot_pairs(const pxff_so3_params *p, int64_t natoms, const int64_t *d_img_start,
                        const int64_t *d_img_j, const double *d_img_r, const int64_t *d_img_rev,
                        const int64_t *d_atom_z, const double *d_G, const double *d_rq, const double *d_lnorm,
                        double *d_out, double *d_radial, int64_t *d_work, int64_t cap, const double *d_table,
                        void *stream);

/* Tabulated radial integrals for pxff_so3_ctot_pairs (d_table; nmax, lmax
 * <= 6): I_nl(R) and dI_nl/dR on nk equal intervals of [0, rc] as local
 * polynomials of degree 11, built on the device from the same quadrature at
 * each interval's Chebyshev nodes (to about 1e-14 of each function's
 * largest value; nk from the steepness 2 alpha rc^2, at least 256).  The
 * size call gives nk and the table length in doubles; the build needs a
 * scratch of the same length.  A table serves every call with the same
 * parameters (it depends on nmax, lmax, rcut, alpha, not on the cutoff
 * function or weights, which the store does not hold). */
int pxff_so3_radial_table_size(const pxff_so3_params *p, int *nk, int64_t *len);
int pxff_so3_radial_table(const pxff_so3_params *p, const double *d_G, const double *d_rq, const double *d_lnorm,
                          double *d_scratch, double *d_table, void *stream);

/* Power spectrum x [natoms*ncoef]; with d_dctot non-NULL the tangent of x. */
int pxff_so3_power(const pxff_so3_params *p, int64_t natoms, const double *d_ctot, const double *d_dctot,
                   double *d_x, void *stream);

/* Adjoint: A [natoms*csz*2] = S(dE/dx) . ctot. */
int pxff_so3_adjoint_a(const pxff_so3_params *p, int64_t natoms, const double *d_ctot, const double *d_beta,
                       double *d_a, void *stream);

/* Adjoint: dE/dr per image [nimg*3] (angular part and expansion recomputed
 * per image; radial integrals from d_radial when not NULL, else recomputed). */
int pxff_so3_deidrj(const pxff_so3_params *p, int64_t nimg, const int64_t *d_img_atom, const int64_t *d_img_j,
                    const double *d_img_r, const int64_t *d_atom_z, const double *d_G, const double *d_rq,
                    const double *d_lnorm, const double *d_a, const double *d_radial, double *d_dedr,
                    void *stream);

/* pxff_so3_deidrj over an image capacity: images e < min(cap, *d_count),
 * the count read on the device (as pxff_so4_deidrj_counted), for a HIP
 * graph capture of a whole SO3 E+F step.  Replaces the same reference
 * contraction as pxff_so3_deidrj (SO3.py:572-694 with neuralnetwork.py:
 * 484-497); d_dedr [cap*3], entries at and beyond the count not written. */
int pxff_so3_deidrj_counted(const pxff_so3_params *p, int64_t cap, const int64_t *d_count,
                            const int64_t *d_img_atom, const int64_t *d_img_j, const double *d_img_r,
                            const int64_t *d_atom_z, const double *d_G, const double *d_rq, const double *d_lnorm,
                            const double *d_a, const double *d_radial, double *d_dedr, void *stream);

/* Plain path: dP/dr per image [nimg*ncoef*3] (then pxff_so4_rows builds the
 * reference's dxdr / rdxdr rows, whose row rules SO3.py:243-265 shares). */
int pxff_so3_dpdr(const pxff_so3_params *p, int64_t nimg, const int64_t *d_img_atom, const int64_t *d_img_j,
                  const double *d_img_r, const int64_t *d_atom_z, const double *d_G, const double *d_rq,
                  const double *d_lnorm, const double *d_ctot, const double *d_radial, double *d_dpdr,
                  void *stream);

#ifdef __cplusplus
}
#endif
#endif
