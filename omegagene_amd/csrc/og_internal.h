// og_internal.h -- device-side data layout of the MI355X engine.
//
// One og_dev_sys describes one simulated system (a walker).  Every kernel is
// batched: it receives the device array of descriptors of a batch plus a
// block map (int2: system index, block index within that system), so a
// single system and 64 uneven walkers run through the same kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OG_NAC 8                 // atoms per cell (MiniCell.h:10 N_ATOM_CELL)
#define OG_ENUM_STAGE 320        // cell pairs per cell kept from the counting pass (more: enumerated again)
#define OG_MAX_NB15 32           // exclusion partners per atom (CelesteObject.h:26)
#define OG_SORT_CAP 2048         // max atoms per column handled by the in-LDS sort
#define OG_EMPTY_BIN 0x7fffffff  // uniform-z bin with no cell
#define OG_FX_SCALE 4294967296.0 // fixed-point force scale, 2^32 units per kcal/mol/A
#define OG_FX_INV (1.0 / 4294967296.0)
// j-side (x, y) gradient of the pair kernel: two signed 32-bit fixed-point
// fields at 2^19 units per kcal/mol/A packed linearly into one 64-bit word
// (word = x + y * 2^32), so one 64-bit integer atomic adds both and the sums
// stay exact and order-independent.  Each field holds +-4096 kcal/mol/A; a
// pair-kernel row beyond +-OG_FXJ_ROW_LIMIT flags the evaluation (an overlap).
#define OG_FXJ_SCALE 524288.0
#define OG_FXJ_INV (1.0 / 524288.0)
#define OG_FXJ_ROW_LIMIT 2048.0f
// Words of the fixed-point gradient per system (cap = cap_atom_array, a
// multiple of 32): x, y, z of the i side and of excluded pairs at 2^32 units
// (cap each); the j-side (x, y) words (cap); the j-side z of each slot pair
// (2k, 2k + 1) as two 32-bit fields at 2^19 units in one word (cap / 2), so
// neighbouring rows of one cell2 add both with one atomic; and the value of
// that word each slot last took (cap).  The pair word is never cleared: a slot
// takes the change since its last read, so the two slots of a pair may read it
// in any order (a slot reads it once per evaluation).
#define OG_FX_ZJ(cap) (4 * (size_t)(cap))
#define OG_FX_ZLAST(cap) (OG_FX_ZJ(cap) + (size_t)(cap) / 2)
#define OG_FX_WORDS(cap) (OG_FX_ZLAST(cap) + (size_t)(cap))
#define OG_NO_STEP_END 0x7fffffffffffffffLL  // og_dev_counters::step_end of a system with no segment limit
#define OG_TH_SCALING 1  // og_params::thermostat value of the scaling thermostat (OG_THERMOSTAT_SCALING)
// og_dev_sys::mc_mode values (the public OG_EXT_VMCMD / OG_EXT_VAUS) and coefficients per state (OG_MC_MAX_ORDER + 1)
enum { OG_MC_VMCMD = 2, OG_MC_VAUS = 3 };
#define OG_MC_NPOLY 9

// energy terms kept per system and per step
enum {
    OG_E_VDW = 0,    // LJ pair energy
    OG_E_ELE,        // ZD pair energy (no excess, no self)
    OG_E_EXCESS,     // ZD excluded-pair ("excess") energy
    OG_E_SELF,       // ZD self energy (constant)
    OG_E_BOND,       // harmonic bonds
    OG_E_KIN,        // kinetic energy
    OG_E_VIRIAL,     // scalar pair virial  sum r_ij . f_ij
    OG_E_EXTEND,     // VcMD bias energy
    OG_E_NPAIRS,     // in-cutoff unmasked atom pairs evaluated (exact count, stored as double)
    OG_E_LAMBDA,     // VcMD reaction coordinate
    OG_E_VSTATE,     // VcMD virtual state during the step
    OG_E_ANGLE,      // harmonic angles
    OG_E_TORSION,    // torsions
    OG_E_IMPRO,      // improper torsions
    OG_E_VDW14,      // scaled 1-4 LJ
    OG_E_ELE14,      // scaled 1-4 Coulomb
    OG_E_NTERMS
};

// flags in og_dev_counters::flags
enum {
    OG_F_CP_OVERFLOW   = 1,
    OG_F_COL_OVERFLOW  = 2,
    OG_F_CELL_OVERFLOW = 4,
    OG_F_NAN           = 8,
    OG_F_SHAKE         = 16,
    OG_F_OUTSIDE_BOX   = 32,
};
// flags that stop the system (every kernel skips it) until the host has reported and cleared them
#define OG_F_FATAL (OG_F_CP_OVERFLOW | OG_F_COL_OVERFLOW | OG_F_CELL_OVERFLOW | OG_F_NAN)

struct og_dev_counters {
    int       n_cells;
    int       n_cell_pairs;
    int       flags;
    int       shake_fail;
    long long step;      // steps completed since system creation
    int       vs;        // current VcMD virtual state
    int       force_energy;  // 1: evaluate energies on the next force call whatever the step
    long long step_end;  // the system stands still once step reaches this (segment end)
    long long noise_step;  // the step whose Langevin update takes its normals from og_dev_sys::noise (-1: none)
    int       rebuild_now; // 1: rebuild the grid before the next step whatever its phase (cleared by the step)
    double    mc_scale;    // V-McMD: the gradient factor of the current step
};

struct og_dev_sys {
    // ---- sizes ----
    int n_atoms;
    int n_columns, nx, ny, nnx, nny, n_uni_z;
    int cap_atom_array, cap_cells, cap_cell_pairs;
    int n_types;
    int n_excess, n_bonds, n_trios, n_pairs_c, n_settle;
    int n_angles, n_torsions, n_impros, n_nb14;  // bonded terms beyond bonds (og_bonded.hip)
    int n_bterms;          // n_angles + n_torsions + n_impros + n_nb14: one thread each
    int n_quads;           // SHAKE four-atom groups
    int n_free;            // atoms outside SETTLE waters (fused Langevin step)
    int n_water_blocks;    // fused Langevin step: blocks of OG_WATERS_PER_BLOCK waters; 0 = not fusable
    int pair_split;        // pair kernel: at most this many of the last cells run as two half-cell waves each
    int n_com_groups;
    int integrator;        // 0 leapfrog-presto, 1 langevin
    int constraint;        // 0 none, 1 shake, 2 shake+settle
    int com_cancel;
    int gap_bug;
    int excess_bug;
    int rebuild_intvl;
    int max_loops;
    int log_stride;        // energies stored every log_stride steps
    int log_cap;           // entries in the energy log
    // ---- geometry, reference float semantics (PBC.h:21-27, MiniCell.h:42-54) ----
    float L[3], lb[3], ub[3];
    float L_cell[2];
    float L_z_uni;
    float cutoff_pair2;    // float(nsgrid_cutoff)^2
    float cutoff;          // pair cutoff (float)
    float cutoff2;
    // ---- zero-dipole constants (ZeroMultipoleSum.cpp:108-190) ----
    float zcore, bcoeff, fcoeff, piewald, alpha;
    // higher zero-multipole orders (zero-quadrupole/-octupole/-hexadecapole, ZeroMultipoleSum.cpp:108-190,
    // 218-442): zm_order 0 = dipole (the constants above), 1..3 = quadrupole..hexadecapole.  Pair energy
    // cc (1/r [erfc] - zm_core - sum_k zm_c[k] r^(2k+2)), gradient coefficient
    // -cc (1/r^3 [erfc] [+ piewald exp(-a^2 r^2) / r^2] + sum_k zm_pg[k] r^(2k)); zm_cc[k] = (2k+2) zm_c[k]
    // as the reference's float members (its excluded-pair terms use these; zm_pg differs from zm_cc only
    // where the reference's octupole alpha > 0 pair gradient takes the hexadecapole coefficients)
    int   zm_order;
    float zm_core, zm_c[4], zm_cc[4], zm_pg[4];
    // hydrophobicity-scale LJ + Debye-Hueckel (cal_pair_hps_dh, cuda_setup.cu:1273-1332)
    float hps_eps;         // --hydrophobicity-scale-epsiron
    float dh_diel_inv;     // 1 / --debye-huckel-dielectric
    float dh_kappa;        // inverse Debye length (1/A)
    int   dh_grad_fix;     // 1: Debye-Hueckel gradient coefficient -E (1/r + kappa) / r instead of the reference's -E (1/r + kappa)
    // ---- integration (FP64) ----
    double dt, tol, temperature, gamma;
    // scaling thermostat (leapfrog, no constraints): 1/tau after the clamp by the step, heating ramp,
    // temperature coefficient 2 * 4184 / (R d_free) * KINETIC_COEFF (Thermostat.cpp:13-17)
    int    thermostat, heating_steps;
    double th_tau_inv, th_temp_init, th_temp_coeff;
    double *th_part;       // [n_atom_blocks] kinetic partials of the thermostat
    double *noise;         // [n_atoms*3] injected N(0,1) draws (og_set_langevin_noise), or null
    double settle_dOH, settle_dHH;
    double e_self;
    uint32_t seed_lo, seed_hi;
    // ---- VcMD 1-D bias on the distance between two group mass centres ----
    int    vc_n_states, vc_interval, vc_na, vc_nb;
    double vc_sigma, vc_coef, vc_constk;
    double vc_minv_a, vc_minv_b;       // float 1/group mass, as MmSystem.cpp:572-582
    int    *vc_atoms;                  // [vc_na + vc_nb], group a then group b
    double *vc_crdg;                   // [3*(vc_na+vc_nb)] unwrapped group coordinates
    double *vc_lo, *vc_hi, *vc_qinv;   // [n_states] ranges and 1/q_cano
    double *vc_hist;                   // [n_states] visit counts (q_raw)
    // ---- V-McMD / V-AUS (mc_mode OG_EXT_VMCMD / OG_EXT_VAUS; V-AUS takes its groups from vc_*) ----
    int     mc_mode, mc_n_states, mc_interval;
    double  mc_constk;                 // k_B T in kcal/mol
    int    *mc_order;                  // [n_states]
    double *mc_poly;                   // [n_states * OG_MC_NPOLY]
    double *mc_lo, *mc_hi, *mc_pdown, *mc_pup;  // [n_states]
    // ---- atom-order arrays ----
    double *crd, *crd_prev, *crd_prev2, *vel, *vel_next, *work;
    float  *mass, *mass_inv, *charge;
    int    *atom_type;
    int    *nb15off;        // [n_atoms*32], -1 terminated
    int    *comp;           // exclusion-connectivity component of each atom
    float  *lj6, *lj12;     // [n_types*n_types]
    float2 *lj_pair;        // [n_types*n_types] {6 A6, 12 A12} interleaved for the pair kernel (the gradient's
                            // coefficients; lj_unscaled: {A6, A12} as the reference's table)
    float4 *lj_hps;         // [n_types*n_types] {6 A6, 12 A12, HPS switch distance, HPS lambda} (pair form HPS + DH)
    int     lj_unscaled;    // 1: lj_pair holds {A6, A12} (a caller's table), scaled by the pair kernel
    int    *excess;         // [n_excess*2]
    int    *bond_ids;       // [n_bonds*2]
    double *bond_k, *bond_r0;
    int    *exc_off, *bond_off;  // per-atom CSR offsets of the excess pairs / bonds: [n_atoms+1]
    int4   *exc_at, *bond_at;    // entries {a, b, pair index, 0}: [2*n_excess], [2*n_bonds]
    // bonded terms: each term writes the gradient share of each of its atoms
    // (its "contributions") to bcontrib; an atom sums its shares through the
    // per-atom list bc_off / bc_at in ascending slot order.  Slots: angle t
    // 3t..3t+2, then torsions (4 each), impropers (4 each), 1-4 pairs (2 each).
    int     *ang_ids;       // [n_angles*3]
    double2 *ang_par;       // {k, theta0}
    int     *tor_ids;       // [(n_torsions+n_impros)*4], torsions then impropers
    double4 *tor_par;       // {V, overlaps, symmetry, phase}
    int     *nb14_ids;      // [n_nb14*2]
    int     *nb14_tp;       // index of the entry's type pair in lj6 / lj12
    double2 *nb14_coef;     // {LJ scale, Coulomb scale}
    double  *bcontrib;      // [n_contrib*3]
    int     *bc_off;        // [n_atoms+1]
    int     *bc_at;         // [n_contrib] contribution slots of each atom
    int    *quad_ids;       // [n_quads*4]
    double *quad_d2;        // [n_quads*6]
    int    *trio_ids;       // [n_trios*3]
    double *trio_d2;        // [n_trios*3]
    int    *pairc_ids;      // [n_pairs_c*2]
    double *pairc_d2;
    int    *settle_ids;     // [n_settle*3]
    int    *free_ids;       // [n_free] atoms outside SETTLE waters
    int    *com_off, *com_atoms;   // COM-cancel groups
    float  *com_minv;
    // ---- grid ----
    int      *col_of_atom;  // [n_atoms]
    uint32_t *zkey;         // [n_atoms] order-preserving float-z bits
    int      *n_atoms_col;  // [n_columns]
    int      *col_fill;     // [n_columns]
    int      *head_cell;    // [n_columns+1], -1 = empty column
    int      *head_atom;    // [n_columns+1], -1 = empty column
    int      *n_cells_z;    // [n_columns]
    int      *tmp_ids;      // [cap_atom_array] atoms scattered by column
    int      *atomids;      // [cap_atom_array] grid slot -> atom, -1 dummy
    int      *rev;          // [n_atoms] atom -> grid slot
    double   *gcrd;         // [cap_atom_array*3] grid-order coordinates (float-exact at rebuild)
    float4   *gxyzq;        // [cap_atom_array] float x, y, z, charge
    int      *gtype;        // [cap_atom_array]
    int      *gcomp;        // [cap_atom_array] exclusion component (-1 dummy)
    int      *cell_col;     // [cap_cells]
    int      *cell_z;       // [cap_cells]
    float    *cell_zmin, *cell_zmax;
    int      *cell_u0, *cell_u1;
    int      *u2c_min, *u2c_max;   // [(n_columns+1)*n_uni_z]
    // ---- cell pairs ----
    int      *cp_count;     // [cap_cells]
    int      *cp_stage;     // [cap_cells * OG_ENUM_STAGE] pairs of each cell from the counting pass
    int      *cp_head;      // [cap_cells+1]
    long long *cp_bsum;     // [cap_cells / 1024 + 1] block totals of the multi-block offset scan
    int      *cp_c2;        // [cap_cell_pairs] cell2 | image << 24
    uint64_t *cp_mask;      // [cap_cell_pairs] bit a2*8+a1 set = skip
    // ---- forces / energies ----
    long long *fx;          // [OG_FX_WORDS(cap_atom_array)] fixed-point gradient, slot g at og_fx_slot(g)
                            // (2^32 units) | packed j-side (x, y) (og_fxj_pack)
    long long *e_acc;       // [OG_EACC_SLOTS][OG_EACC_STRIDE] energy accumulators, 2^32 fixed point (the pair
                            // count as an integer): vdw, ele, virial, pairs, excess, bond; a cell / block adds to
                            // slot (index % OG_EACC_SLOTS), so no cache line takes every add; zeroed by the reader
    double    *e_atom;      // [n_ke_parts] kinetic partials (one per block of the finishing kernel)
    double    *com_mom;     // [n_com_groups*3]
    double    *energy;      // [OG_E_NTERMS] last step
    double    *elog;        // [log_cap*OG_E_NTERMS]
    int        n_atom_blocks;
    int        n_ke_parts;   // max(n_atom_blocks, water blocks + 1 of the fused step)
    og_dev_counters *cnt;
};

// ---- block maps ----
enum og_map_kind {
    OG_MAP_ATOM = 0,   // 256 threads per block over atoms
    OG_MAP_SLOT,       // 256 threads per block over grid slots (capacity)
    OG_MAP_CELLW,      // 4 waves (256 thr) per block, one wave per cell (capacity)
    OG_MAP_COL,        // one block per column
    OG_MAP_SYS,        // one block per system
    OG_MAP_MISC,       // 256 threads per block over max(n_excess, n_bonds, n_pairs_c)
    OG_MAP_TRIO,       // 128 threads per block over trios / settle waters
    OG_MAP_CP,         // 4 waves per block, grid-stride over cell pairs (fixed block count per system)
    OG_MAP_BIN,        // 256 threads per block over uniform-z bins
    OG_MAP_NKINDS
};

#define OG_THREADS 256
#define OG_TRIO_THREADS 128
#define OG_QUAD_THREADS 64
// Pair kernel work units: in a system of 2 * pair_split to 8 * pair_split
// cells the last pair_split cells are evaluated as two waves each (their
// cell-pair records halved), so a launch of a few rounds of waves ends on
// short waves instead of a tail of whole cells.  Smaller systems (a batch of
// walkers, where the launch holds many systems) and larger ones (many rounds,
// a short tail) keep one wave per cell.
#define OG_PAIR_SPLIT 4096
__host__ __device__ __forceinline__ int og_pair_split_cells(int split, int n_cells) {
    return n_cells >= 2 * split && n_cells < 8 * split ? split : 0;
}
#define OG_WATER_THREADS 256  // fused Langevin step: threads per block
#define OG_WATERS_PER_WAVE 21  // fused Langevin step: three lanes per water
#define OG_WATERS_PER_BLOCK (OG_WATERS_PER_WAVE * (OG_WATER_THREADS / 64))
#define OG_FREE_MAX 4096      // fused Langevin step: most atoms outside SETTLE waters (one block takes them)
#define OG_CP_BLOCKS_PER_SYS_MIN 64

// ---- small device helpers ----
// A system whose segment is complete, or whose list overflowed or whose
// forces went non-finite, skips every kernel (each kernel reads the counters
// before its first write; step only changes in the last kernel of a step, so
// all kernels of one step agree); a fault never spreads to the other systems
// of a batch and never integrates on without forces.
__device__ __forceinline__ bool og_frozen_c(const og_dev_counters &c) {
    return c.step >= c.step_end || (c.flags & OG_F_FATAL) != 0;
}
__device__ __forceinline__ bool og_frozen(const og_dev_sys &s) { return og_frozen_c(*s.cnt); }
// The grid-rebuild kernels run for a system when its own step is a multiple
// of its rebuild interval (DynamicsMode.cpp:480-486) or a rebuild is forced;
// systems at other phases in the same launch skip them.
__device__ __forceinline__ bool og_rebuild_skip(const og_dev_sys &s) {
    return og_frozen(s) || (s.rebuild_intvl > 0 && (s.cnt->step % s.rebuild_intvl) != 0 && !s.cnt->rebuild_now);
}
// Energies (pair, virial) are evaluated on logged steps and on request; the
// forces and the pair count on every step.
__device__ __forceinline__ bool og_energy_step_c(const og_dev_sys &s, const og_dev_counters &c) {
    // V-McMD needs the potential energy of every step (its lambda)
    return c.force_energy || (c.step % s.log_stride) == 0 || s.mc_mode == OG_MC_VMCMD;
}
__device__ __forceinline__ bool og_energy_step(const og_dev_sys &s) { return og_energy_step_c(s, *s.cnt); }

// PBC::diff_crd_minim_image (PBC.cpp:54-80) with PBC's float box lengths:
// two truncating passes d - L (int)(d / (L/2)).  When d lies within rounding
// of +-L (two atoms close along an axis, one just inside each face of the
// box), d / (L/2) rounds to +-2 exactly and the two passes swap -L and +L,
// leaving |d| ~ L.  A third, nearest-image step, taken only when |d| > 3L/4
// after the two passes (never for any other input), returns the short
// vector there.  Double differences (integration, bonds) and float ones
// (excluded pairs).
__device__ __forceinline__ double og_mimg(double v, float L) {
#pragma clang fp contract(off)
    const float Lh  = L * 0.5f;
    const float Lhi = 1.0f / Lh;
    v               = v - (double)(L * (float)(int)(v * (double)Lhi));
    v               = v - (double)(L * (float)(int)(v * (double)Lhi));
    if (fabs(v) > 0.75 * (double)L) v -= (double)L * nearbyint(v / (double)L);
    return v;
}
__device__ __forceinline__ float og_mimg_f(float v, float L) {
#pragma clang fp contract(off)
    const float Lh  = L * 0.5f;
    const float Lhi = 1.0f / Lh;
    v               = v - L * (float)(int)(v * Lhi);
    v               = v - L * (float)(int)(v * Lhi);
    if (fabsf(v) > 0.75f * L) v -= L * rintf(v / L);
    return v;
}

__device__ __forceinline__ int og_lane() { return __lane_id(); }

__device__ __forceinline__ float og_dpp_sum8(float x) {
    // sum over the 8 lanes sharing lane>>3 (quad swaps, then half-row mirror)
    x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, true));
    x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x4E, 0xF, 0xF, true));
    x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x141, 0xF, 0xF, true));
    return x;
}

__device__ __forceinline__ long long og_to_fixed(double v) { return __double2ll_rn(v * OG_FX_SCALE); }

// Index of grid slot g in each component of the fixed-point gradient: cell
// index bits 0 and 1 swapped, so cells c and c + 2 -- the same-parity cells
// one cell1 takes from a neighbour column -- share a 128-byte line and one
// memory-side atomic request covers both.  A bijection on every group of
// four cells (cap_atom_array is a multiple of 32 slots).
__device__ __forceinline__ int og_fx_slot(int g) {
    const int c = g >> 3;
    return ((((c & ~3) | ((c & 1) << 1) | ((c >> 1) & 1))) << 3) | (g & 7);
}

// Round-to-nearest fixed point without a conversion sequence: adding
// 1.5*2^52 puts round(d) in the low mantissa bits (valid for |d| < 2^51,
// i.e. |gradient| < 2^19 kcal/mol/A at the 2^32 scale).
#define OG_FX_MAGIC 6755399441055744.0
#define OG_FX_MAGIC_BITS 0x4338000000000000LL
#define OG_FX_LIMIT 262144.0
__device__ __forceinline__ long long og_fx_fast(double v) {
    const double t = v * OG_FX_SCALE + OG_FX_MAGIC;
    return __double_as_longlong(t) - OG_FX_MAGIC_BITS;
}
// (x, y) rounded to 2^-19 and packed as x + y * 2^32 (|x|, |y| < 4096)
__device__ __forceinline__ unsigned long long og_fxj_pack(double x, double y) {
    const long long ix = __double_as_longlong(x * OG_FXJ_SCALE + OG_FX_MAGIC) - OG_FX_MAGIC_BITS;
    const long long iy = __double_as_longlong(y * OG_FXJ_SCALE + OG_FX_MAGIC) - OG_FX_MAGIC_BITS;
    return (unsigned long long)ix + ((unsigned long long)iy << 32);
}
// The same word from float inputs (|x|, |y| below 4096): x * 2^19 is exact in
// float and rintf rounds half to even like the FP64 form, so the result is
// bit-identical to og_fxj_pack((double)x, (double)y), in FP32 instructions.
__device__ __forceinline__ unsigned long long og_fxj_pack_f(float x, float y) {
    const long long ix = (long long)(int)__builtin_rintf(x * 524288.0f);
    const long long iy = (long long)(int)__builtin_rintf(y * 524288.0f);
    return (unsigned long long)ix + ((unsigned long long)iy << 32);
}
// the two fields of a packed sum (each within +-2^31 units)
__device__ __forceinline__ void og_fxj_unpack(long long w, double &x, double &y) {
    const long long ix = (long long)(int)(unsigned)(unsigned long long)w;
    const long long iy = (long long)((unsigned long long)w - (unsigned long long)ix) >> 32;
    x = (double)ix * OG_FXJ_INV;
    y = (double)iy * OG_FXJ_INV;
}
// the pair gradient of grid slot g (index f = og_fx_slot(g)) in kcal/mol/A
__device__ __forceinline__ void og_fx_read(const long long *fx, int cap, int f, double *w) {
    double jx, jy, zlo, zhi;
    og_fxj_unpack(fx[3 * cap + f], jx, jy);
    // the slot pair's z word changed by both slots' sums since this slot's last read
    og_fxj_unpack(fx[OG_FX_ZJ(cap) + (f >> 1)] - fx[OG_FX_ZLAST(cap) + f], zlo, zhi);
    w[0] = (double)fx[f] * OG_FX_INV + jx;
    w[1] = (double)fx[cap + f] * OG_FX_INV + jy;
    w[2] = (double)fx[2 * cap + f] * OG_FX_INV + ((f & 1) ? zhi : zlo);
}
// after og_fx_read: clear the slot's own words and note the pair word it has taken
__device__ __forceinline__ void og_fx_consume(long long *fx, int cap, int f) {
    fx[OG_FX_ZLAST(cap) + f] = fx[OG_FX_ZJ(cap) + (f >> 1)];
    fx[3 * cap + f]          = 0;
    for (int k = 0; k < 3; k++) fx[k * cap + f] = 0;
}

#define OG_GLOBAL(T) __attribute__((address_space(1))) T

#define OG_RED_THREADS 1024  // per-system reduction block
enum { OG_EACC_VDW, OG_EACC_ELE, OG_EACC_VIR, OG_EACC_PAIRS, OG_EACC_EXCESS, OG_EACC_BOND, OG_EACC_ANGLE,
       OG_EACC_TORSION, OG_EACC_IMPRO, OG_EACC_VDW14, OG_EACC_ELE14, OG_EACC };
#define OG_EACC_SLOTS 256   // accumulator slots per system
#define OG_EACC_STRIDE 16   // long longs per slot (two 64-byte lines: pair terms, then the others)

// Pair, excluded-pair and bond energies of one system into s.energy from
// the fixed-point accumulator slots (integer adds: the same bits whatever the
// order of the contributions), which are cleared for the next evaluation.
// Block-wide, NT threads (a multiple of 64, at most 1024); sh: 16 *
// OG_EACC_STRIDE long longs of shared memory.  The slots are read as one
// flat array, thread t taking words t, t + NT, ... (coalesced, all of a
// thread's loads in flight together), so each thread sums one term
// (t % OG_EACC_STRIDE) over a share of the slots.
#define OG_EACC_WORDS (OG_EACC_SLOTS * OG_EACC_STRIDE)
template <int NT>
__device__ __forceinline__ long long og_eacc_load(const og_dev_sys &s) {
    static_assert(NT % 64 == 0 && NT <= 1024 && OG_EACC_WORDS % NT == 0, "block shape");
    long long w[OG_EACC_WORDS / NT];
#pragma unroll
    for (int k = 0; k < OG_EACC_WORDS / NT; k++) w[k] = s.e_acc[threadIdx.x + k * NT];
    long long v = 0;
#pragma unroll
    for (int k = 0; k < OG_EACC_WORDS / NT; k++) v += w[k];
    return (threadIdx.x & (OG_EACC_STRIDE - 1)) < OG_EACC ? v : 0;  // the padding words of a slot count zero
}
template <int NT>
__device__ __forceinline__ void og_eacc_clear(const og_dev_sys &s) {
#pragma unroll
    for (int k = 0; k < OG_EACC_WORDS / NT; k++) s.e_acc[threadIdx.x + k * NT] = 0;
}
// the block's loaded term sums into s.energy (and, when e is given, into
// thread 0's copy e of the energy row)
template <int NT>
__device__ __forceinline__ void og_eacc_reduce(const og_dev_sys &s, long long v, long long *sh, double *e) {
    static_assert(OG_EACC_STRIDE == 16 && OG_EACC <= OG_EACC_STRIDE, "one term per lane of a 16-lane row");
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) < OG_EACC_STRIDE) sh[w * OG_EACC_STRIDE + (threadIdx.x & 63)] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t[OG_EACC];
        for (int q = 0; q < OG_EACC; q++) t[q] = 0;
        for (int k = 0; k < NT / 64; k++)
            for (int q = 0; q < OG_EACC; q++) t[q] += sh[k * OG_EACC_STRIDE + q];
        double r[OG_E_NTERMS];
        r[OG_E_VDW]     = (double)t[OG_EACC_VDW] * OG_FX_INV;
        r[OG_E_ELE]     = (double)t[OG_EACC_ELE] * OG_FX_INV;
        r[OG_E_VIRIAL]  = (double)t[OG_EACC_VIR] * OG_FX_INV;
        r[OG_E_NPAIRS]  = (double)t[OG_EACC_PAIRS];
        r[OG_E_EXCESS]  = (double)t[OG_EACC_EXCESS] * OG_FX_INV;
        r[OG_E_BOND]    = (double)t[OG_EACC_BOND] * OG_FX_INV;
        r[OG_E_ANGLE]   = (double)t[OG_EACC_ANGLE] * OG_FX_INV;
        r[OG_E_TORSION] = (double)t[OG_EACC_TORSION] * OG_FX_INV;
        r[OG_E_IMPRO]   = (double)t[OG_EACC_IMPRO] * OG_FX_INV;
        r[OG_E_VDW14]   = (double)t[OG_EACC_VDW14] * OG_FX_INV;
        r[OG_E_ELE14]   = (double)t[OG_EACC_ELE14] * OG_FX_INV;
        r[OG_E_SELF]    = s.e_self;
        const int terms[] = {OG_E_VDW, OG_E_ELE, OG_E_VIRIAL, OG_E_NPAIRS, OG_E_EXCESS, OG_E_BOND, OG_E_ANGLE,
                             OG_E_TORSION, OG_E_IMPRO, OG_E_VDW14, OG_E_ELE14, OG_E_SELF};
        for (int k : terms) {
            s.energy[k] = r[k];
            if (e) e[k] = r[k];
        }
    }
    __syncthreads();
}
template <int NT>
__device__ __forceinline__ void og_take_pair_energies(const og_dev_sys &s, long long *sh) {
    const long long v = og_eacc_load<NT>(s);
    og_eacc_clear<NT>(s);
    og_eacc_reduce<NT>(s, v, sh, nullptr);
}



__device__ __forceinline__ void og_atomic_add_fx(long long *p, long long v) {
    if (v != 0) atomicAdd((unsigned long long *)p, (unsigned long long)v);
}
__device__ __forceinline__ void og_atomic_add_fx_g(OG_GLOBAL(long long) *p, long long v) {
    __hip_atomic_fetch_add((OG_GLOBAL(unsigned long long) *)p, (unsigned long long)v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// float -> orderable uint32 (ascending float order == ascending uint order); -0 -> +0
__device__ __forceinline__ uint32_t og_float_key(float f) {
    if (f == 0.0f) f = 0.0f;
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float og_key_float(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

// kernel launch declarations (defined in og_grid.hip, og_pair.hip, og_integrate.hip)
struct og_launch {
    const og_dev_sys *sys;   // device array
    const int2       *map;   // device block map
    int               n_blocks;
    hipStream_t       stream;
};

// grid build
void og_launch_grid_clear(const og_launch &cols, const og_launch &bins);
void og_launch_grid_bin(const og_launch &atoms);
void og_launch_grid_scan(const og_launch &sys);
void og_launch_grid_scatter(const og_launch &atoms);
void og_launch_grid_sort(const og_launch &cols);
void og_launch_cells_from_grid(const og_launch &cols);
void og_launch_enum_count(const og_launch &cellw);
void og_launch_enum_scan(const og_launch &sys);
void og_launch_enum_scan_blocks(const og_launch &blocks);  // blocks of 1024 cells (cap_cells + 1)
void og_launch_enum_write(const og_launch &cellw);
void og_launch_masks(const og_launch &cp);
// forces
// pair forms of og_launch_pair
enum { OG_FORM_ZD = 0, OG_FORM_ZD_ALPHA = 1, OG_FORM_HPS_DH = 2, OG_FORM_ZM = 3, OG_FORM_ZM_ALPHA = 4 };
void og_launch_pair(const og_launch &cells, int form, int waves);  // waves per block: 1, 2, 4 or 8
void og_launch_forces_to_work(const og_launch &atoms);
void og_launch_energy_reduce(const og_launch &sys);
// integration
void og_launch_leapfrog_vel(const og_launch &atoms);
void og_launch_com_sum(const og_launch &sys);
void og_launch_com_apply(const og_launch &atoms);
void og_launch_thermostat(const og_launch &atoms);  // scaling thermostat: kinetic partials, then the rescale
void og_launch_positions(const og_launch &atoms);
void og_launch_langevin_positions(const og_launch &atoms);
void og_launch_constraints(const og_launch &trio, const og_launch &misc);
// bonded terms beyond bonds and four-atom SHAKE (og_bonded.hip)
void og_launch_bonded(const og_launch &terms);  // OG_THREADS threads per block over n_bterms
void og_launch_shake_quad(const og_launch &quads);  // OG_QUAD_THREADS threads per block over n_quads
void og_launch_langevin_step(const og_launch &a);
void og_launch_finish(const og_launch &atoms);
void og_launch_step_close(const og_launch &sys);  // energies of the step, log entry, step counter
void og_launch_vcmd(const og_launch &sys);  // VcMD and V-AUS group bias
void og_launch_vmcmd(const og_launch &sys, const og_launch &atoms);  // V-McMD: lambda, state, gradient factor
void og_launch_refresh_grid(const og_launch &slots);
