/*
 * og_celeste.h -- the reference's own GPU call points, one for one, on the
 * MI355X engine.
 *
 * celeste's GPU build calls a flat extern "C" ABI defined in
 * src/celeste/cuda/cuda_setup.cu and cuda_common.cu and declared privately at
 * the top of its two callers (src/celeste/core/SubBox.cpp:5-78,
 * MiniCell.cpp:7-35).  Every function below replaces the cuda_* function of
 * the same suffix with the same argument list and meaning; a reference build
 * switches by renaming the prefix in those two prototype blocks and linking
 * libog_engine.so.  What stays the same: host MiniCell builds the grid
 * (atomids, column heads, float grid-order coordinates) and owns the pinned
 * host buffers; the device enumerates cell pairs, evaluates LJ + zero-dipole
 * (alpha = 0, the form of cuda_setup.cu:513-520) and hands back the FP64
 * gradient ("work") in grid order plus energies {vdw, ele}.
 *
 * What differs, all invisible to the caller's data:
 *  - the int return values are og_status codes (0 = OK, as the reference's
 *    "return 0") instead of exit(EXIT_FAILURE) on error (cuda_common.h:7-22);
 *    the message is available through og_last_error (og_engine.h);
 *  - forces accumulate in 64-bit fixed point, so the gradient is
 *    deterministic (the reference adds doubles with CAS atomics in
 *    nondeterministic order, cuda_setup.cu:8-16,642-674);
 *  - the cell-pair list is built in the reference CPU order
 *    (MiniCell::enumerate_cell_pairs, MiniCell.cpp:938-1149) with its
 *    inclusion test, not the GPU variant's atomic slot order;
 *  - a capacity overflow is an error status (OG_E_CAPACITY), not a device printf.
 *
 * Like the reference, this layer keeps one system per process in file-scope
 * state and is not re-entrant: call it from one host thread.  The handle ABI
 * in og_engine.h is the re-entrant, fully device-resident path.
 * Types: real_pw = float, real_fc = double (CelesteObject.h:11-24);
 * the reference's T*& parameters are T** here, and otherwise every prototype
 * is the one its caller declares (tests/abi_check.cpp checks them).
 */
#ifndef OG_CELESTE_H
#define OG_CELESTE_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#include "og_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cuda_hostalloc_atom_info cuda_setup.cu:268-280: pinned grid-order host
 * buffers crd [max_n_atom_array*3] float, atomids [max_n_atom_array],
 * work [max_n_atom_array*3] double, energy [2] double.  Must run first: it
 * fixes max_n_atom_array for og_alloc_atom_info. */
int og_hostalloc_atom_info(float **h_crd, int **h_atomids, double **h_work, double **h_energy,
                                 int max_n_atom_array);
/* cuda_hostfree_atom_info cuda_setup.cu:316-324, declared by its caller as
 * (real_pw*, int*, real_fc*&, real_fc*&) (MiniCell.cpp:19): work and energy
 * by reference, freed and set to null */
int og_hostfree_atom_info(float *h_crd, int *h_atomids, double **h_work, double **h_energy);
/* cuda_hostalloc_cell_info cuda_setup.cu:282-287: idx_xy_head_cell [n_columns] */
int og_hostalloc_cell_info(int **h_idx_xy_head_cell, int n_columns);
/* cuda_hostfree_cell_info cuda_setup.cu:325-328 */
int og_hostfree_cell_info(int *h_idx_xy_head_cell);
/* cuda_hostalloc_atom_type_charge cuda_setup.cu:257-264: atom order, n_atoms_system */
int og_hostalloc_atom_type_charge(int **h_atom_type, float **h_charge, int n_atoms_system);
/* cuda_hostfree_atom_type_charge cuda_setup.cu:310-314 */
int og_hostfree_atom_type_charge(int *h_atom_type, float *h_charge);

/* cuda_alloc_atom_info cuda_setup.cu:42-77 (SubBox.cpp:1688-1690 passes
 * n_columns + 1 as the last argument) */
int og_alloc_atom_info(int max_n_atoms_exbox, int max_n_cells, int max_n_cell_pairs, int n_columns);
/* cuda_free_atom_info cuda_setup.cu:79-102 */
int og_free_atom_info(void);
/* cuda_alloc_set_lj_params cuda_setup.cu:205-244 as its caller declares it
 * (SubBox.cpp:27-33, seven arguments; the definition reads five): float LJ
 * tables [n_lj_types^2] and the atom-order exclusion table
 * h_nb15off [max_n_atoms_exbox * max_n_nb15off] (-1 padded, MmSystem.cpp:478-525);
 * max_n_atoms and max_n_atom_array are accepted and not needed. */
int og_alloc_set_lj_params(float *h_lj_6term, float *h_lj_12term, int n_lj_types, int *h_nb15off,
                           const int max_n_nb15off, const int max_n_atoms, const int max_n_atom_array);
/* cuda_free_lj_params cuda_setup.cu:246-255 */
int og_free_lj_params(void);
/* cuda_set_pbc cuda_setup.cu:138-146: L[3], lower_bound[3] (float) */
int og_set_pbc(float *l, float *lb);
/* cuda_set_constant cuda_setup.cu:193-203 */
int og_set_constant(float cutoff, float cutoff_pairlist, int n_atomtypes);
/* cuda_zerodipole_constant cuda_setup.cu:148-153 */
int og_zerodipole_constant(float zcore, float bcoeff, float fcoeff);
/* cuda_memcpy_htod_atom_info cuda_setup.cu:115-119: atom-order charges and
 * (0-based) LJ types of the n_atoms_system atoms */
int og_memcpy_htod_atom_info(float **h_charge_orig, int **h_atomtype_orig);
/* cuda_set_cell_constant cuda_setup.cu:157-189 (per grid rebuild) */
int og_set_cell_constant(int n_cells, int n_atoms_exbox, int n_atom_array, const int *n_cells_xyz,
                               int n_columns, const float *l_cell_xyz, const int *n_neighbor_xyz);
/* cuda_memcpy_htod_atomids cuda_setup.cu:104-110: atomids [n_atom_array]
 * (grid slot -> atom, -1 dummy) and idx_xy_head_cell [n_columns + 1]
 * (-1 for an empty column, MiniCell.cpp:807-810) */
int og_memcpy_htod_atomids(int **h_atomids, int **h_idx_xy_head_cell);
/* cuda_set_atominfo cuda_setup.cu:404-425: types, charges and the reverse
 * map into grid order */
int og_set_atominfo(void);
/* cuda_memcpy_htod_crd cuda_setup.cu:124-136: float grid-order xyz [n_atom_array*3] */
int og_memcpy_htod_crd(float **h_crd);
/* cuda_set_crd cuda_setup.cu:427-434: pack xyz + charge into the pair kernel's stream */
int og_set_crd(void);
/* cuda_enumerate_cell_pairs cuda_setup.cu:1095-1216: cell pairs and their
 * 8x8 exclusion masks from the current grid (idx_atom_cell_xy: column of
 * each grid slot, MiniCell.cpp:351; n_neighbor_col is checked against the
 * cell constants) */
int og_enumerate_cell_pairs(int **h_atomids, int n_cells, int n_neighbor_col, const int *idx_atom_cell_xy);
/* cuda_pairwise_ljzd cuda_setup.cu:777-792: asynchronous; the flag is
 * ignored as in the reference */
int og_pairwise_ljzd(const bool flg_mod_15mask);
/* cuda_pair_sync cuda_setup.cu:797-801 */
int og_pair_sync(void);
/* cuda_memcpy_dtoh_work cuda_setup.cu:803-823: work [n_atom_array*3] grid
 * order (dE/dx, kcal/mol/A) and energy [2] = {vdw, ele} */
int og_memcpy_dtoh_work(double **h_work, double **h_energy, int n_atoms, int n_atom_array);

/* The second pair form, hydrophobicity-scale LJ + Debye-Hueckel (the
 * reference's F_HPSCUDA build, SubBox.cpp:1702-1710,1730-1733):
 * cuda_hps_constant cuda_setup.cu:1254-1258 (--hydrophobicity-scale-epsiron) */
int og_hps_constant(float epsiron);
/* cuda_alloc_set_hps_params cuda_setup.cu:1230-1244: float [n_lj_types^2]
 * switch distance and lambda per type pair (after og_alloc_set_lj_params) */
int og_alloc_set_hps_params(float *h_hps_cutoff, float *h_hps_lambda, int n_lj_types);
/* cuda_free_hps_params cuda_setup.cu:1245-1252 */
int og_free_hps_params(void);
/* cuda_debye_huckel_constant cuda_setup.cu:1259-1270 */
int og_debye_huckel_constant(float dielect, float temperature, float ionic_strength);
/* cuda_pairwise_hps_dh cuda_setup.cu:1478-1493: asynchronous; results through
 * og_memcpy_dtoh_work as og_pairwise_ljzd (the flag is ignored as in the reference) */
int og_pairwise_hps_dh(const bool flg_mod_15mask);

/* Inspection for parity tests (no reference counterpart): the cell-pair list
 * of the last og_enumerate_cell_pairs as CellPair records
 * {cell_id1, cell_id2, image, pair_mask[2]} (MiniCell.h:17-26) and
 * idx_head_cell_pairs [n_cells + 1]; n_cell_pairs may be queried with
 * cell_pairs5 == NULL. */
int og_get_celeste_cell_pairs(int32_t *cell_pairs5, int *idx_head_cell_pairs, int *n_cell_pairs);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): og_pairwise_ljzd followed by og_memcpy_dtoh_work, with
 * the same arithmetic and bit-identical results, on buffers the caller owns
 * on the device.  Asynchronous on `stream` (a hipStream_t; NULL = the null
 * stream): four kernel launches, no allocation, memset or copy.
 *   d_xyzq      [n_cells*8*4] grid-order float x, y, z, charge (charge 0 in
 *               empty slots)
 *   d_atom_type [n_cells*8] LJ type of each grid slot (0 in empty slots)
 *   d_lj_pair   [n_lj_types^2 * 2] {6-term, 12-term} per type pair, row-major
 *   d_idx_head_cell_pairs [n_cells+1], d_cell_pair_c2 [n_cell_pairs]
 *               (cell_id2 | image << 24), d_cell_pair_mask [n_cell_pairs]
 *               (pair_mask[0] | pair_mask[1] << 32): the CellPair list of
 *               og_enumerate_cell_pairs (og_get_celeste_cell_pairs) split by field
 *   l [3]       box lengths (host); cutoff, zcore/bcoeff/fcoeff as
 *               og_set_constant and og_zerodipole_constant (alpha = 0)
 *   d_work      [n_cells*8*3] out: FP64 gradient in grid order
 *   d_energy    [2] out: vdw, ele
 *   d_flags     [1] out (may be NULL): non-zero when a gradient was
 *               non-finite or out of the fixed-point range (og_memcpy_dtoh_work
 *               returns OG_E_NAN then)
 * Its working memory is a fixed per-GPU device array held until the process
 * ends; grids above OG_NOALLOC_MAX_ATOM_ARRAY slots (n_cells*8) are rejected
 * with OG_E_CAPACITY.  Calls that overlap on one GPU share that array and
 * race: serialize them (one stream, or events). */
#define OG_NOALLOC_MAX_ATOM_ARRAY 262144
#define OG_NOALLOC_MAX_DEVICES 64
int og_pairwise_ljzd_noalloc(const float *d_xyzq, const int *d_atom_type, const float *d_lj_pair, int n_lj_types,
                             const int *d_idx_head_cell_pairs, const uint32_t *d_cell_pair_c2,
                             const uint64_t *d_cell_pair_mask, int n_cells, const float *l, float cutoff,
                             float zcore, float bcoeff, float fcoeff, double *d_work, double *d_energy, int *d_flags,
                             void *stream);

#ifdef __cplusplus
}
#endif
#endif
