/*
 * oracle.h -- CPU restatement of the celeste (omegagene) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libog_engine.so) never links, loads or calls it.
 *
 * Every function states the reference file:line it restates (paths relative
 * to the reference source tree src/celeste/).  Numeric semantics follow the
 * reference host build (g++ -std=c++14 -O3, x86-64 SSE, no FP contraction):
 * `real` = double, `real_pw` = float, double literals promote.
 *
 * Parity pins (what checks each function) are the oracle test-suite
 * (tests/test_oracle.py): reference fixtures, reference builds, closed forms
 * and brute force; functions with no pin are marked "parity unpinned" here.
 */
#ifndef OG_ORACLE_H
#define OG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_N_ATOM_CELL 8
#define ORC_MAX_N_NB15OFF 32

/* Grid geometry (core/MiniCell.cpp:221-290, :1279-1303, :1327-1362; PBC.cpp:18-41). */
typedef struct {
    float L[3];          /* PBC::L (float) */
    float lb[3];         /* PBC::lower_bound (float) */
    float cutoff_pair;   /* --nsgrid-cutoff (float) */
    float cutoff_pair_half;
    float cutoff_pair2;
    int   n_cells_xyz[3];
    float L_cell_xyz[3];
    int   n_neighbors_xyz[3];
    int   n_columns;
    int   n_uni_z;
    float L_z_uni;
    int   max_n_cells;       /* capacity heuristic (MiniCell.cpp:264-270) */
    int   max_n_cell_pairs;  /* capacity heuristic (MiniCell.cpp:272-280) */
    int   max_n_atoms_column;
} orc_grid_t;

/* PBC from the .cls box block: val[12] = {Lx,0,0, 0,Ly,0, 0,0,Lz, cx,cy,cz} (PBC.cpp:18-41). */
void orc_pbc_from_box(const double val[12], float L[3], float lb[3], float ub[3]);

/* set_grid_parameters + set_box_info + set_max_n_atoms_region + set_grid_xy. */
int orc_grid_setup(orc_grid_t *g, int n_atoms, const double box[12], double nsgrid_cutoff);

/* Result of one grid rebuild (MiniCell::set_atoms_into_grid_xy and callees). */
typedef struct {
    int     n_cells;
    int     n_atom_array;      /* n_cells * 8 */
    int    *atomids;           /* [cap_atom_array] grid slot -> atom id, -1 = dummy */
    double *crd;               /* [cap_atom_array*3] grid-order coordinates (float-exact) */
    int    *cell_crd;          /* [cap_cells*3] (column x, column y, z index) */
    int    *idx_xy_head_cell;  /* [n_columns+1]; -1 for an empty column */
    int    *idx_xy_head_atom;  /* [n_columns+1]; -1 for an empty column */
    int    *n_atoms_col;       /* [n_columns] */
    int    *n_cells_z;         /* [n_columns] */
    int    *column_of_atom;    /* [n_atoms] column id of each atom (atom order) */
    int     cap_atom_array;
    int     cap_cells;
} orc_cells_t;

/* Allocate / free the buffers of an orc_cells_t for n_atoms atoms. */
int  orc_cells_alloc(orc_cells_t *c, const orc_grid_t *g, int n_atoms);
void orc_cells_free(orc_cells_t *c);

/* Full rebuild: binning, column scan, cycle-chasing reorder, dummy padding,
 * in-column z quicksort (MiniCell.cpp:292-402, 404-501, 503-613, 618-691, 785-869).
 * crd: atom-order double coordinates, assumed inside [lb, lb+L). */
int orc_cells_build(orc_cells_t *c, const orc_grid_t *g, int n_atoms, const double *crd);

/* Cell-pair record as the reference lays it out (MiniCell.h:17-26). */
typedef struct {
    int32_t cell_id1;
    int32_t cell_id2;
    int32_t image;
    int32_t pair_mask[2];
} orc_cellpair_t;

/* Enumerate cell pairs in the reference CPU order and set their exclusion
 * bitmasks (MiniCell.cpp:896-936 set_uniform_grid, :938-1149 enumerate_cell_pairs,
 * :1150-1160 check_valid_pair, :1161-1196 add_cell_pair, :1198-1237 set_cell_pair_bitmask).
 * nb15off: [n_atoms*32] exclusion partner lists, -1 terminated (MmSystem.cpp:478-525).
 * gap_bug != 0 reproduces the reference truncation of the uniform-z scan at
 * the first empty bin (MiniCell.cpp:1063-1077); 0 scans every bin in range.
 * Returns the number of pairs, or -1 if cap is exceeded. */
int orc_enumerate_cell_pairs(const orc_grid_t *g, const orc_cells_t *c, const int *nb15off,
                             int gap_bug, orc_cellpair_t *out, int cap, int *idx_head /* [n_cells+1] */);

/* Zero-multipole constants (ZeroMultipoleSum.cpp:108-190; float members).
 * orc_zd_setup fills the dipole members (mode ORC_ZM_DIPOLE); orc_zm_setup
 * every order's and selects one (--electrostatic zero-dipole / -quadrupole /
 * -octupole / -hexadecapole, Config.cpp:114-128). */
enum { ORC_ZM_DIPOLE = 0, ORC_ZM_QUADRUPOLE = 1, ORC_ZM_OCTUPOLE = 2, ORC_ZM_HEXADECAPOLE = 3 };
typedef struct {
    float cutoff, alpha;
    float fcoeff, bcoeff, zcore, piewald;
    int   mode;         /* ORC_ZM_* */
    int   zo_grad_bug;  /* 1: the octupole alpha > 0 pair gradient takes the hexadecapole coefficients
                           (ZeroMultipoleSum.cpp:404-406, the reference's form) */
    float zmcore;       /* the selected order's core constant (cal_self_energy :72-80) */
    float zq2, zq4, zq22, zq44;
    float zo2, zo4, zo6, zo22, zo44, zo66;
    float zh2, zh4, zh6, zh8, zh22, zh44, zh66, zh88;
} orc_zd_t;
void orc_zd_setup(orc_zd_t *zd, double cutoff, double alpha);
void orc_zm_setup(orc_zd_t *zd, double cutoff, double alpha, int mode, int zo_grad_bug);

/* Self energy (ZeroMultipoleSum.cpp:43-106). Returns the sum of per-atom terms
 * (d_self_mon, stored in a float member) and writes the per-atom FP64 terms. */
double orc_self_energy(const orc_zd_t *zd, int n_atoms, const float *charge,
                       int n_excess, const int *excess_pairs /* [n_excess*2] */, double *e_self);

/* Excluded-pair list i<j from nb15off (MmSystem.cpp:700-736). Returns count. */
int orc_excess_pairs(int n_atoms, const int *nb15off, int *out_pairs /* cap*2 */, int cap);

/* Reference-faithful FP32 pair term (ForceField.cpp:333-418 +
 * ZeroMultipoleSum.cpp:309-346).  d12 = crd1 - crd2 already imaged.
 * Returns 1 if r < cutoff (pair counted), 0 otherwise. */
int orc_pair_fp32(const orc_zd_t *zd, const float d12[3], float a6, float a12, float q1, float q2,
                  float *e_vdw, float *e_ele, float work[3]);

/* The other pair forms of ForceField::calc_pairwise (ForceField.cpp:333-418):
 * hydrophobicity-scale LJ (--nonbond hydrophobicity-scale-lj, Config.cpp:333-342)
 * and Debye-Hueckel electrostatics (--electrostatic debye-huckel,
 * ZeroMultipoleSum.cpp:443-455, constants :24).  hps_cutoff / hps_lambda are the
 * [n_types^2] float tables of the .cls "nbpair hps" section (Read.cpp:322-334). */
typedef struct {
    int          hps;          /* 1: HPS switch on the LJ term */
    int          dh;           /* 1: Debye-Hueckel electrostatics (else zero-dipole) */
    int          dh_grad_fix;  /* 0: the reference's grad_coeff = -E (1/r + 1/lambda_D), :453; 1: that / r */
    float        hps_eps;      /* --hydrophobicity-scale-epsiron (passed as real_pw, SubBox.cpp:1048) */
    float        dielectric_inv, debye_inv;  /* real_pw members of ZeroMultipoleSum */
    const float *hps_cutoff, *hps_lambda;
} orc_ff_t;
/* dielectric_inv = 1/eps_r and the inverse Debye length in 1/A from the
 * double cfg values, as the ZeroMultipoleSum constructor (ZeroMultipoleSum.cpp:21-24) */
void orc_dh_setup(orc_ff_t *ff, double dielectric, double ionic_strength, double temperature);
/* one pair with the form of ff (ff == NULL: orc_pair_fp32); t = type pair index */
int orc_pair_ff_fp32(const orc_zd_t *zd, const orc_ff_t *ff, const float d12[3], int t, float a6, float a12,
                     float q1, float q2, float *e_vdw, float *e_ele, float work[3]);
int64_t orc_pairs_celllist_ff(const orc_zd_t *zd, const orc_ff_t *ff, const orc_grid_t *g, const orc_cells_t *c,
                              const orc_cellpair_t *cps, int n_cp, const int *atom_type, int n_types,
                              const float *lj6, const float *lj12, const float *charge,
                              double *work_grid, double energy[2]);
int64_t orc_pairs_allpairs_ff(const orc_zd_t *zd, const orc_ff_t *ff, const float L[3], int n_atoms,
                              const double *crd, const int *atom_type, int n_types, const float *lj6,
                              const float *lj12, const float *charge, const int *nb15off, double *work,
                              double energy[2]);

/* Pair loop over a cell-pair list (SubBox.cpp:956-1125): FP32 pair terms,
 * FP64 accumulation in list order.  work_grid[n_atom_array*3] (grid order,
 * zeroed here), energy[2] = {vdw, ele}.  Returns in-cutoff pair count. */
int64_t orc_pairs_celllist(const orc_zd_t *zd, const orc_grid_t *g, const orc_cells_t *c,
                           const orc_cellpair_t *cps, int n_cp, const int *atom_type, int n_types,
                           const float *lj6, const float *lj12, const float *charge,
                           double *work_grid, double energy[2]);

/* All-pairs loop with float minimum image (SubBox.cpp:1127-1220, the
 * omegagene_wons build).  work[n_atoms*3] atom order (zeroed here). */
int64_t orc_pairs_allpairs_fp32(const orc_zd_t *zd, const float L[3], int n_atoms, const double *crd,
                                const int *atom_type, int n_types, const float *lj6, const float *lj12,
                                const float *charge, const int *nb15off, double *work, double energy[2]);

/* FP64 reference of the same physics (not a restatement of reference
 * rounding): double coordinates, exact minimum image, erfc in double.
 * Used to state force RMS relative error.  Uses the cell list for O(N).
 * energy: vdw, ele, virial (sum over pairs of g r^2 = r_ij . dE/dr_i; the
 * reference has no virial). */
int64_t orc_pairs_celllist_fp64(const orc_zd_t *zd, const orc_grid_t *g, const orc_cells_t *c,
                                const orc_cellpair_t *cps, int n_cp, const double *crd_atom,
                                const int *atom_type, int n_types, const float *lj6, const float *lj12,
                                const float *charge, double *work_atom, double energy[3]);

/* ZD excess over excluded pairs (SubBox.cpp:1347-1367, ForceField.cpp:420-437,
 * ZeroMultipoleSum.cpp:191-217).  bug_compat != 0 keeps the alpha>0 energy
 * term -erf(ar)*r of ZeroMultipoleSum.cpp:214; 0 uses -erf(ar)/r.
 * Adds into work[n_atoms*3]; returns the energy. */
double orc_excess(const orc_zd_t *zd, const float L[3], int n_excess, const int *pairs, const double *crd,
                  const float *charge, int bug_compat, double *work);

/* Harmonic bonds E = k (r - r0)^2 (ForceField.cpp:108-136, SubBox.cpp:1229-1249). */
double orc_bonds(const float L[3], int n_bonds, const int *pairs, const double *k, const double *r0,
                 const double *crd, double *work);

/* ---- integration (all FP64) ---- */

/* v' = v + dt*(-FORCE_VEL)*w*minv  (SubBox.cpp:1445-1454). */
void orc_update_velocities(int n, double dt, const double *vel, const double *work, const float *mass_inv,
                           double *vel_next);
/* x' = x + v*dt (SubBox.cpp:1567-1580). */
void orc_update_coordinates(int n, double dt, const double *crd_prev, const double *vel, double *crd);
/* SHAKE pair (ConstraintShake.cpp:18-49). dist2 = squared distances. */
void orc_shake_pair(const float L[3], int n, const int *ids, const double *dist2, double *crd,
                    const double *crd_prev, const float *mass_inv);
/* SHAKE trio (ConstraintShake.cpp:51-149). Returns number of non-converged trios. */
int orc_shake_trio(const float L[3], int n, const int *ids, const double *dist2 /* n*3 */, double *crd,
                   const double *crd_prev, const float *mass_inv, double tol, int max_loops);
/* Analytic SETTLE (Miyamoto & Kollman 1992) for rigid 3-site waters; new code
 * with no reference counterpart (the reference SETTLE is a stub,
 * ConstraintShake.cpp:284-286).  ids = (O,H1,H2); dOH, dHH distances. */
int orc_settle(const float L[3], int n, const int *ids, double dOH, double dHH, double *crd,
               const double *crd_prev, const float *mass); /* returns waters with no rigid solution (left unconstrained) */
/* v' = (x' - x)/dt (SubBox.cpp:1462-1481). */
void orc_velocity_from_crd(int n, double dt_inv, const double *crd, const double *crd_prev, double *vel_next);
/* Wrap once into [lb, lb+L) (SubBox.cpp:1482-1502). */
void orc_revise_pbc(int n, const float L[3], const float lb[3], double *crd);
/* Remove group COM velocity (COMMotion.cpp:30-55). */
void orc_cancel_com(int n_groups, const int *group_off /* n_groups+1 */, const int *group_atoms,
                    const float *mass_inv_group, const float *mass, double *vel);
/* KE = KINETIC_COEFF * sum m ((v+v')/2)^2 (DynamicsMode.cpp:358-372, SubBox.cpp:1504-1516). */
double orc_kinetic(int n, const float *mass, const double *vel, const double *vel_next);
/* Position Langevin (SubBox.cpp:2077-2114) with caller-supplied N(0,1) noise
 * zeta[n*3] (the reference draws it from mt19937 + std::normal_distribution). */
void orc_langevin_positions(int n, double dt_half, double gamma, double temperature, const float L[3],
                            const double *crd_prev, const double *crd_prev2, const double *work,
                            const float *mass, const float *mass_inv, const double *zeta, double *crd);
/* The reference's N(0,1) stream (oracle_rng.cpp): std::mt19937 seeded with
 * seed, one fresh std::normal_distribution per draw (Random.cpp:10-43);
 * skips `skip` draws, writes the next n. */
void orc_ref_normals(int seed, long long skip, long long n, double *out);
/* ThermostatScaling::apply_thermostat (Thermostat.cpp:57-84) after the
 * leapfrog velocity update: vel = v(t - dt/2), vel_next overwritten.
 * tau_inv = 1/tau with tau = --berendsen-tau clamped by the step as
 * SubBox::init_thermostat (SubBox.cpp:1674-1676); temperature_coeff of
 * ThermostatObject::set_temperature_coeff (Thermostat.cpp:13-17).  Returns
 * the scale factor. */
double orc_thermostat_scaling(int n, double dt, double tau_inv, double temperature, double temperature_coeff,
                              const double *work, const double *vel, const float *mass, const float *mass_inv,
                              double *vel_next);
/* Central-difference velocity (SubBox.cpp:2054-2066). */
void orc_langevin_velocity(int n, double dt, const float L[3], const double *crd, const double *crd_prev2,
                           double *vel);

/* VcMD 1-D bias (Extend.cpp:668-692, :753-797, :1035-1102) and the
 * transition candidates of vs_step/set_vs_next (:808-880).  Adds the bias to
 * work[n*3]; returns the bias energy. */
double orc_vcmd_bias(const float L[3], int na, const int *atoms_a, int nb, const int *atoms_b, double *crdg,
                     const double *crd, const float *mass, float minv_a, float minv_b, int n_states,
                     const double *lo, const double *hi, int vs, double sigma, double coef, double temperature,
                     double *work, double *lambda_out, int *in_range, int cand[2]);

/* V-McMD / V-AUS (Extend.cpp:126-540): the virtual-state move with the two
 * uniform draws given, the V-McMD gradient factor, and the V-AUS bias (adds
 * into work, returns lambda; crdg as for orc_vcmd_bias). */
int    orc_vs_transition(int n_states, const double *lo, const double *hi, const double *p_down, const double *p_up,
                         int vs, double lambda, double u_up, double u_down);
double orc_vmcmd_factor(const double *poly, int order, double lo, double hi, double temperature, double lambda);
double orc_vaus_bias(const float L[3], int na, const int *atoms_a, int nb, const int *atoms_b, double *crdg,
                     const double *crd, const float *mass, float minv_a, float minv_b, const double *poly,
                     int order, double lo, double hi, double sigma, double recov_coef, double temperature,
                     double *work);

#ifdef __cplusplus
}
#endif
/* ---- bonded terms beyond bonds and four-atom SHAKE (oracle_bonded.c) ----
 * Angles (SubBox.cpp:1250-1270, ForceField.cpp:138-177): ids [3n], E = k (theta - theta0)^2. */
double orc_angles(const float L[3], int n, const int *ids, const double *k_eps, const double *theta0,
                  const double *crd, double *work);
/* Torsions and impropers (SubBox.cpp:1272-1321, ForceField.cpp:179-299): ids [4n]. */
double orc_torsions(const float L[3], int n, const int *ids, const double *ene, const int *overlaps,
                    const int *symmetry, const double *phase, const double *crd, double *work);
/* 1-4 pairs (SubBox.cpp:1323-1346, ForceField.cpp:301-331): ids [2n], the entry's
 * own types [2n], scaling of the LJ and Coulomb terms. */
void orc_nb14(const float L[3], int n, const int *ids, const int *types, const double *coeff_vdw,
              const double *coeff_ele, const float *lj6, const float *lj12, int n_types, const float *charge,
              const double *crd, double *work, double *e_vdw, double *e_ele);
/* SHAKE on four-atom groups (ConstraintShake.cpp:151-276): ids [4n], squared
 * distances [6n] of pairs 01 12 20 30 23 13.  Returns the groups not converged. */
int orc_shake_quad(const float L[3], int n, const int *ids, const double *dist2, double *crd, const double *crd_prev,
                   const float *mass_inv, double tol, int max_loops);

#endif
