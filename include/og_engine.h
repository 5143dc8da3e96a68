/*
 * og_engine.h -- C ABI of the MI355X (gfx950) force/integration engine that
 * replaces the CUDA backend of celeste/omegagene.
 *
 * The reference exposes a flat extern "C" ABI from src/celeste/cuda/cuda_setup.cu
 * and cuda_common.cu, declared privately at the top of its callers
 * (src/celeste/core/SubBox.cpp:5-78, MiniCell.cpp:7-35), with device state in
 * file-scope globals (cuda_setup.h:77-133).  This header keeps the call points
 * and their data semantics (grid-order gradient "work" = dE/dx, energies
 * {vdw, ele}) but:
 *   - state lives behind an og_system handle, so many systems can coexist;
 *   - the grid, the cell-pair list, the pair kernel, the reductions and the
 *     integrator all run on the device; no per-step PCIe copies;
 *   - every call returns an og_status instead of calling exit();
 *   - batched calls run many independent systems (walkers) per launch.
 *
 * Threading: one handle is used from one host thread at a time; distinct
 * handles (and batches) are independent.  Ownership: the library owns device
 * memory; callers own every host array and pass it with its count.
 * Units: Angstrom, fs, kcal/mol, amu, e (reference units, CelesteObject.cpp:27-40).
 */
#ifndef OG_ENGINE_H
#define OG_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference aborts instead: cuda_common.h:7-22) ---- */
typedef enum {
    OG_OK = 0,
    OG_E_ARG = 1,       /* invalid argument */
    OG_E_CAPACITY = 2,  /* a device list (cells, cell pairs, column sort) overflowed */
    OG_E_NAN = 3,       /* non-finite coordinate or force */
    OG_E_NOCONV = 4,    /* SHAKE did not converge within --const-max-loops */
    OG_E_DEVICE = 5,    /* HIP runtime error */
    OG_E_STATE = 6      /* call out of order */
} og_status;

typedef struct og_system og_system;
typedef struct og_batch og_batch;

/* integrator / constraint / electrostatics selectors (Config.cpp:71-131) */
enum { OG_INTEGRATOR_LEAPFROG_PRESTO = 0, OG_INTEGRATOR_LANGEVIN = 1 };
enum { OG_CONSTRAINT_NONE = 0, OG_CONSTRAINT_SHAKE = 1, OG_CONSTRAINT_SHAKE_SETTLE = 2 };

/* energy terms returned by og_get_energies (array of OG_NTERMS doubles) */
enum {
    OG_TERM_VDW = 0,    /* LJ pair energy                     (MmSystem::pote_vdw) */
    OG_TERM_ELE,        /* ZD pair energy, no excess/self      */
    OG_TERM_EXCESS,     /* ZD excluded-pair energy             (SubBox.cpp:1347-1367) */
    OG_TERM_SELF,       /* ZD self energy                      (MmSystem::energy_self_sum) */
    OG_TERM_BOND,       /* harmonic bonds                      (MmSystem::pote_bond) */
    OG_TERM_KINETIC,    /* kinetic energy                      (MmSystem::kinetic_e) */
    OG_TERM_VIRIAL,     /* sum over pairs of g*r^2 = r_ij . dE/dr_i (new; the reference has none) */
    OG_TERM_EXTEND,     /* VcMD bias energy                    (MmSystem::pote_extend) */
    OG_TERM_NPAIRS,     /* in-cutoff unmasked atom pairs of the last evaluation (an exact integer) */
    OG_TERM_LAMBDA,     /* VcMD reaction coordinate: distance between the two group mass centres (A) */
    OG_TERM_VSTATE,     /* VcMD virtual state in force during the step (0-based) */
    OG_TERM_ANGLE,      /* harmonic angles                     (MmSystem::pote_angle) */
    OG_TERM_TORSION,    /* torsions                            (MmSystem::pote_torsion) */
    OG_TERM_IMPRO,      /* improper torsions                   (MmSystem::pote_impro) */
    OG_TERM_VDW14,      /* scaled 1-4 LJ                       (MmSystem::pote_14vdw) */
    OG_TERM_ELE14,      /* scaled 1-4 Coulomb                  (MmSystem::pote_14ele) */
    OG_NTERMS
};

/* One system as read from a .cls launch set (Read.cpp:50-794) plus the
 * derived tables MmSystem builds (nb15off MmSystem.cpp:478-525, excess pairs
 * MmSystem.cpp:700-736).  All pointers are host memory owned by the caller and
 * only read during og_system_create. */
typedef struct {
    int           n_atoms;
    const double *box;          /* [12] {Lx,0,0, 0,Ly,0, 0,0,Lz, cx,cy,cz} */
    const double *crd;          /* [n*3] */
    const double *vel;          /* [n*3] velocities as read (vel_just), A/fs */
    const double *charge;       /* [n] */
    const double *mass;         /* [n] */
    const int    *atom_type;    /* [n] 0-based */
    int           n_lj_types;
    const double *lj6;          /* [n_types*n_types] A6 = 4 eps sigma^6 */
    const double *lj12;         /* [n_types*n_types] A12 = 4 eps sigma^12 */
    const int    *nb15off;      /* [n*32] partner lists, -1 terminated */
    int           n_excess;     /* excluded pairs i<j, reference order */
    const int    *excess;       /* [n_excess*2] */
    int           n_bonds;
    const int    *bond_ids;     /* [n_bonds*2] */
    const double *bond_k;       /* E = k (r - r0)^2 */
    const double *bond_r0;
    int           n_shake_pairs;
    const int    *shake_pair_ids;   /* [n*2] */
    const double *shake_pair_d2;    /* squared distances */
    int           n_shake_trios;
    const int    *shake_trio_ids;   /* [n*3] */
    const double *shake_trio_d2;    /* [n*3] squared d12, d23, d31 */
    int           n_settle;
    const int    *settle_ids;       /* [n*3] O, H1, H2 */
    double        settle_dOH, settle_dHH;
    int           n_com_groups;     /* COM-cancel groups (disjoint) */
    const int    *com_group_off;    /* [n_com_groups+1] */
    const int    *com_group_atoms;
    int           d_free;           /* degrees of freedom for the temperature */
    /* VcMD 1-D virtual-state bias on the distance between the mass centres of
     * two atom groups (ExtendedVcMD, Extend.cpp:655-1102; input block
     * Read.cpp:705-779).  vc_n_states = 0 disables it. */
    int           vc_n_states;
    const double *vc_min;           /* [n_states] lower edge of each state's lambda range */
    const double *vc_max;           /* [n_states] upper edge */
    const double *vc_q_cano;        /* [n_states] canonical weights; 0 = the default (Extend.cpp:548-566) */
    int           vc_init_state;    /* 0-based */
    int           vc_interval;      /* steps between transition trials */
    int           vc_n_a, vc_n_b;   /* group sizes */
    const int    *vc_atoms_a;       /* [vc_n_a] */
    const int    *vc_atoms_b;       /* [vc_n_b] */
    /* the .cls "nbpair hps" section (Read.cpp:322-334): per type pair the HPS
     * switch distance and hydrophobicity scale lambda; read when nonbond is
     * OG_NONBOND_HPS (NULL otherwise) */
    const double *hps_cutoff;       /* [n_types*n_types] */
    const double *hps_lambda;       /* [n_types*n_types] */
    /* bonded terms beyond bonds, in FP64 (MmSystem's angle / torsion / impro /
     * nb14 tables, Read.cpp:336-459, MmSystem.cpp:349-407): */
    int           n_angles;         /* E = k (theta - theta0)^2, theta at the middle atom */
    const int    *angle_ids;        /* [n_angles*3] */
    const double *angle_k;
    const double *angle_theta0;     /* radians */
    int           n_torsions;       /* E = V (1 + cos(n phi - phase)) / overlaps */
    const int    *torsion_ids;      /* [n_torsions*4] */
    const double *torsion_ene;      /* V */
    const int    *torsion_overlaps;
    const int    *torsion_symmetry; /* n, 0..10 */
    const double *torsion_phase;    /* radians */
    int           n_impros;         /* improper torsions, same form */
    const int    *impro_ids;
    const double *impro_ene;
    const int    *impro_overlaps;
    const int    *impro_symmetry;
    const double *impro_phase;
    int           n_nb14;           /* scaled 1-4 LJ + Coulomb pairs */
    const int    *nb14_ids;         /* [n_nb14*2] */
    const int    *nb14_types;       /* [n_nb14*2] 0-based LJ types of the entry (the table row used) */
    const double *nb14_coeff_vdw;   /* LJ scale of the pair */
    const double *nb14_coeff_ele;   /* Coulomb scale of the pair */
    /* SHAKE on four-atom groups (ConstraintShake::shake_quad, ConstraintShake.cpp:151-276) */
    int           n_shake_quads;
    const int    *shake_quad_ids;   /* [n*4] */
    const double *shake_quad_d2;    /* [n*6] squared distances of the pairs 01 12 20 30 23 13 */
    /* V-McMD / V-AUS virtual states (ExtendedVMcMD / ExtendedVAUS, Extend.cpp:126-540; input block
     * Read.cpp:521-561); mc_mode 0 disables them.  V-McMD: lambda = the potential energy, the whole
     * gradient multiplied by k_B T d ln P / d lambda.  V-AUS: lambda = the distance between the mass
     * centres of the groups vc_atoms_a and vc_atoms_b (vc_n_states 0), a force along it on those groups,
     * with og_params::enhance_sigma / enhance_recov_coef for the recovery outside a state's range. */
    int           mc_mode;          /* 0, OG_EXT_VMCMD or OG_EXT_VAUS */
    int           mc_n_states;
    int           mc_interval;      /* steps between transition trials */
    double        mc_temperature;   /* K, for k_B T (ExtendedVMcMD::set_temperature) */
    const int    *mc_order;         /* [n_states] order of each state's polynomial, 0..OG_MC_MAX_ORDER */
    const double *mc_poly;          /* [n_states * (OG_MC_MAX_ORDER + 1)] d ln P / d lambda, constant term first */
    const double *mc_lo, *mc_hi;    /* [n_states] lambda range of each state (closed) */
    const double *mc_prob_down;     /* [n_states] probability of a move down, when lambda allows it */
    const double *mc_prob_up;       /* [n_states] probability of a move up */
    int           mc_init_state;    /* 0-based */
    /* unwrapped coordinates of the bias groups at the start, group a then group b ([3*(vc_n_a+vc_n_b)]; the
     * group-coordinate section, Read.cpp:646-703); NULL: the atoms' own coordinates.  VcMD and V-AUS. */
    const double *vc_crd_init;
} og_system_desc;

enum { OG_EXT_VMCMD = 2, OG_EXT_VAUS = 3 };
#define OG_MC_MAX_ORDER 8

/* Run parameters: the .cfg keys of the path (Config.cpp:37-367). */
typedef struct {
    double   cutoff;               /* --cutoff */
    double   nsgrid_cutoff;        /* --nsgrid-cutoff */
    int      nsgrid_update_intvl;  /* --nsgrid-update-intvl */
    double   ele_alpha;            /* --ele-alpha */
    double   time_step;            /* --time-step (fs) */
    int      integrator;           /* OG_INTEGRATOR_* */
    int      constraint;           /* OG_CONSTRAINT_* */
    double   const_tolerance;      /* --const-tolerance */
    int      const_max_loops;      /* --const-max-loops */
    double   temperature;          /* --temperature (K) */
    double   langevin_gamma;       /* --langevin-gamma already in fs^-1 (Config.cpp:345-348) */
    uint64_t random_seed;          /* --random-seed */
    int      com_cancel;           /* --com-motion cancel */
    int      gap_bug;              /* 1: reproduce the uniform-z scan stop at the first empty bin (MiniCell.cpp:1063-1077) */
    int      excess_bug;           /* 1: the reference's excluded-pair forms: the zero-dipole alpha>0 energy
                                      -erf(ar)*r (ZeroMultipoleSum.cpp:214), the zero-octupole r^5 for r^6
                                      (:255,:271), the zero-hexadecapole alpha 0 gradient (:288) and alpha>0
                                      energy (:304) */
    int      log_stride;           /* energies kept every log_stride steps */
    int      log_cap;              /* entries kept in the device energy log */
    double   capacity_factor;      /* cell-pair list capacity multiplier (1.0 default) */
    double   enhance_sigma;        /* --enhance-sigma (VcMD, A) */
    double   enhance_recov_coef;   /* --enhance-recovery-coef (VcMD) */
    int      split_step;           /* 1: run a Langevin step as separate kernels even where the fused
                                      one-kernel step applies (same trajectory; for comparison) */
    int      thermostat;           /* --thermostat: OG_THERMOSTAT_NONE or OG_THERMOSTAT_SCALING (leapfrog-presto,
                                      no constraints: ThermostatScaling::apply_thermostat, Thermostat.cpp:57-84) */
    double   berendsen_tau;        /* --berendsen-tau already in fs (Config.cpp:162-164); the step replaces it
                                      when smaller (SubBox.cpp:1674-1676) */
    double   temperature_init;     /* --temperature-init (K); < 0: --temperature (Config.cpp:364) */
    int      heating_steps;        /* --heating-steps: linear ramp temperature_init -> temperature
                                      (SubBox::update_thermostat, SubBox.cpp:1745-1757) */
    /* the second pair form (the reference GPU kernel kernel_pairwise_hps_dh, cuda_setup.cu:1273-1477):
     * hydrophobicity-scale LJ with Debye-Hueckel electrostatics */
    int      nonbond;              /* --nonbond: OG_NONBOND_LJ or OG_NONBOND_HPS (hydrophobicity-scale-lj) */
    double   hps_epsilon;          /* --hydrophobicity-scale-epsiron (kcal/mol) */
    int      electrostatic;        /* --electrostatic: OG_ELE_ZERO_DIPOLE, _QUADRUPOLE, _OCTUPOLE, _HEXADECAPOLE or
                                      OG_ELE_DEBYE_HUCKEL */
    double   dh_dielectric;        /* --debye-huckel-dielectric */
    double   dh_ionic_strength;    /* --debye-huckel-ionic-strength */
    double   dh_temperature;       /* --debye-huckel-temperature (K) */
    int      dh_grad_fix;          /* 1: Debye-Hueckel gradient -E (1/r + kappa) r_ij / r; 0 (default): the
                                      reference's -E (1/r + kappa) r_ij (ZeroMultipoleSum.cpp:453,
                                      cuda_setup.cu:1327), which is r times the derivative */
    double   cell_edge_scale;      /* MiniCell column width relative to the reference's (1 = the reference
                                      geometry, required with gap_bug 1); og_autotune_cell_edge picks one */
    int      zo_grad_bug;          /* 1 (default): the zero-octupole alpha > 0 pair gradient takes the
                                      zero-hexadecapole coefficients, as ZeroMultipoleSum.cpp:404-406;
                                      0: its own.  (excess_bug likewise keeps the reference's excluded-pair
                                      forms of every order, ZeroMultipoleSum.cpp:203-307.) */
} og_params;

enum { OG_THERMOSTAT_NONE = 0, OG_THERMOSTAT_SCALING = 1 };
enum { OG_NONBOND_LJ = 0, OG_NONBOND_HPS = 1 };
enum { OG_ELE_ZERO_DIPOLE = 0, OG_ELE_DEBYE_HUCKEL = 1, OG_ELE_ZERO_QUADRUPOLE = 2, OG_ELE_ZERO_OCTUPOLE = 3,
       OG_ELE_ZERO_HEXADECAPOLE = 4 };

void og_params_default(og_params *p);

/* ---- device / lifetime (cuda_set_device cuda_common.cu:3-6;
 *      cuda_alloc_atom_info cuda_setup.cu:42-77; cuda_alloc_set_lj_params :205-244;
 *      cuda_set_pbc :138-146; cuda_set_constant :193-203; cuda_zerodipole_constant :148-153;
 *      cuda_memcpy_htod_atom_info :115-119; cuda_free_atom_info :79-102) ---- */
og_status og_set_device(int device_id);
/* Wait for all work on the current device (cuda_thread_sync cuda_setup.cu:793-796). */
og_status og_thread_sync(void);
og_status og_system_create(const og_system_desc *desc, const og_params *params, og_system **out);
og_status og_system_destroy(og_system *sys);

/* Grid rebuild on the device: atom->column binning, column sort, cells, cell
 * pairs and exclusion masks.  Replaces SubBox::nsgrid_update's host MiniCell
 * work plus cuda_set_cell_constant :157-189, cuda_memcpy_htod_atomids :104-110,
 * cuda_set_atominfo :404-425 and cuda_enumerate_cell_pairs :1095-1216. */
og_status og_nsgrid_update(og_system *sys);

/* Pair + excluded-pair + bond forces and energies for the current
 * coordinates.  Replaces cuda_pairwise_ljzd :777-792 + cuda_pair_sync :797-801 +
 * cuda_memcpy_dtoh_work :803-823 (+ the host bonded/excess terms,
 * SubBox.cpp:923-929).  work_atom: [n*3] dE/dx in atom order (may be NULL);
 * energies: [OG_NTERMS] (may be NULL). */
og_status og_calc_energy_force(og_system *sys, int rebuild, double *work_atom, double *energies);

/* n_steps MD steps on the device (DynamicsModePresto/Langevin::calc_in_each_step,
 * DynamicsMode.cpp:468-560 / :798-899), grid rebuilt every nsgrid_update_intvl
 * steps; launched as a captured hipGraph. */
og_status og_md_run(og_system *sys, int n_steps);

/* The N(0,1) draws zeta [n_atoms*3] (atom-major x, y, z) that the next MD
 * step's position Langevin update (SubBox.cpp:2077-2114) uses in place of the
 * engine's counter-based generator; zeta = NULL cancels.  No reference
 * counterpart: it lets a caller drive the device step with the reference's
 * own noise stream (std::mt19937 + std::normal_distribution, Random.cpp:10-43)
 * and compare trajectories bit for bit.  Applies to og_md_run of this system. */
og_status og_set_langevin_noise(og_system *sys, const double *zeta);
/* gradient ("work", dE/dx) of the last force evaluation or MD step, [n*3] atom order */
og_status og_get_work(og_system *sys, double *work_atom);

/* State exchange (host arrays [n*3], atom order). */
og_status og_get_state(og_system *sys, double *crd, double *vel_next, double *vel_just);
og_status og_set_state(og_system *sys, const double *crd, const double *vel_next);
/* positions before the last step's update [n*3]: what the reference writes to
 * trajectories (SubBox::copy_crd_prev in DynamicsMode::sub_output,
 * DynamicsMode.cpp:272-300) and to a Langevin restart (:215-218) */
og_status og_get_crd_prev(og_system *sys, double *crd_prev);
og_status og_get_energies(og_system *sys, double *energies /* [OG_NTERMS] */);
/* energy log of the last run: rows of OG_NTERMS doubles, one per log_stride steps */
og_status og_get_energy_log(og_system *sys, double *out, int max_rows, int *n_rows);
og_status og_get_step(og_system *sys, long long *step);
/* outcome of the system's last run, batch run or evaluation (an og_status):
 * in a batch every member reports its own, and a faulted member stands
 * still while the others go on */
og_status og_system_status(og_system *sys, int *status);
/* VcMD state: current virtual state and the visit counts per state
 * (q_raw, Extend.cpp:828-832); counts may be NULL */
og_status og_get_vcmd(og_system *sys, int *vstate, double *counts, int n_states);

/* Grid / cell-pair inspection in the reference layouts (parity tests):
 * atomids [cap >= n_atom_array] grid slot -> atom (-1 dummy);
 * idx_xy_head_cell [n_columns+1]; cell pairs as CellPair
 * {cell_id1, cell_id2, image, pair_mask[2]} (MiniCell.h:17-26) and
 * idx_head_cell_pairs [n_cells+1]. */
og_status og_get_grid_info(og_system *sys, int *n_cells, int *n_columns, int *n_cell_pairs, int *n_atom_array);
og_status og_get_grid(og_system *sys, int *atomids, double *grid_crd, int *idx_xy_head_cell);
/* the MiniCell geometry of the grid (MiniCell::set_grid_xy, MiniCell.cpp:221-290),
 * in the arguments og_set_cell_constant and og_alloc_atom_info take:
 * n_cells_xyz {nx, ny, 0}, L_cell_xyz {Lx/nx, Ly/ny, 0}, n_neighbors_xyz
 * {nnx, nny, 0} and the capacities max_n_cells / max_n_cell_pairs */
og_status og_get_grid_geometry(og_system *sys, int *n_cells_xyz, float *L_cell_xyz, int *n_neighbors_xyz,
                               int *max_n_cells, int *max_n_cell_pairs);
og_status og_get_cell_pairs(og_system *sys, int32_t *cell_pairs5, int *idx_head_cell_pairs);
/* fixed-point gradient accumulators would be zero after a force call; this
 * returns the last gradient in grid order (reference cuda_memcpy_dtoh_work layout) */
og_status og_get_work_grid(og_system *sys, double *work_grid);

/* ---- batched walkers (the reference runs them as separate processes:
 *      samples/cg_q8_vcmd/1/c2_exe.bash:11-19) ---- */
og_status og_batch_create(og_batch **out);
og_status og_batch_add(og_batch *b, og_system *sys);
og_status og_batch_remove(og_batch *b, og_system *sys);
og_status og_batch_run(og_batch *b, int n_steps);
og_status og_batch_destroy(og_batch *b);
og_status og_batch_size(og_batch *b, int *n);
/* steps per captured hipGraph of a batch (0 = automatic) */
og_status og_batch_set_graph_steps(og_batch *b, int steps);
/* End the system's current segment n_steps from now: past that step every
 * kernel skips the system, so walkers with different segment lengths share
 * one batch.  n_steps <= 0 removes the limit. */
og_status og_set_segment(og_system *sys, long long n_steps);

/* ---- serving: continuous batching of simulation segments (no reference
 *      counterpart: the reference starts one process per walker and segment,
 *      samples/mcmd_ala3/cal03_mcmd_inp/job_mpi.bash:16-32) ----
 * Requests are admitted FIFO against an HBM budget; every og_server_pump
 * advances all admitted segments by chunk_steps in one batch and retires
 * the finished ones; walkers stay resident on the device between segments.
 * Times are seconds on the server's clock: times[3] = {submit, admit, done}. */
/* ---- on-disk outputs (og_io.cpp) ----
 * WriteRestart::write_restart (WriteTrr.cpp:258-313): title record (80 bytes,
 * zero padded), n_atoms twice, step, time, E_total = E_kin + E_pot, E_kin,
 * E_pot, then crd and vel [n*3] as doubles, each in int-marked records. */
og_status og_write_restart(const char *path, const char *title, int n_atoms, int n_steps, double time,
                           double e_potential, double e_kinetic, const double *crd, const double *vel);
/* WriteTrrPresto::write_trr (WriteTrr.cpp:179-254): one frame (header record
 * of step and ten floats closed by a 0 marker as the reference writes it, then
 * float xyz of every atom or of the n_group atoms in group); append = 0
 * starts the file. */
og_status og_write_trr_presto_frame(const char *path, int append, int n_atoms, int step, float time,
                                    float cpu_time, float total_e, float kinetic_e, float temperature,
                                    float potential_e, float vdw_e, const double *crd, int n_group,
                                    const int *group);
/* WriteTrrGromacs::write_trr (WriteTrr.cpp:33-172): GMX_trn_file frame in
 * double precision, box and crd / vel / force (NULL = left out) in nm. */
og_status og_write_trr_gromacs_frame(const char *path, int append, int n_atoms, int step, double time, double lx,
                                     double ly, double lz, const double *crd, const double *vel, const double *force,
                                     int n_group, const int *group);

typedef struct og_server og_server;
og_status og_server_create(uint64_t hbm_budget_bytes, int chunk_steps, og_server **out);
og_status og_server_destroy(og_server *srv);
/* a new walker (the description is copied) and its first segment */
og_status og_server_submit_new(og_server *srv, const og_system_desc *desc, const og_params *params, int n_steps,
                               long long *request_id, long long *walker_id);
/* the next segment of a resident walker, continuing from its device state */
og_status og_server_submit_next(og_server *srv, long long walker_id, int n_steps, long long *request_id);
og_status og_server_pump(og_server *srv, int *n_active, int *n_queued, int *n_done);
/* final coordinates and velocities [n*3] (atom order) and energy rows
 * [n_steps][OG_NTERMS] of a finished request; any pointer may be NULL */
og_status og_server_result(og_server *srv, long long request_id, double *crd, double *vel, double *energy_rows,
                           int max_rows, int *n_rows, double *times);
og_status og_server_forget(og_server *srv, long long request_id);
og_status og_server_release_walker(og_server *srv, long long walker_id);
og_status og_server_stats(og_server *srv, double *t_now, long long *steps_done, uint64_t *bytes_used, int *n_walkers);

/* diagnostics */
const char *og_status_string(og_status s);
og_status   og_last_error(char *buf, int len);
/* device time of the pair kernel over the last og_md_run / og_batch_run (ms) and its launch count;
 * recorded with HIP events around each pair launch on the run's stream when timing is on */
og_status   og_set_pair_timing(og_system *sys, int on);
og_status   og_pair_kernel_time(og_system *sys, double *ms, int *n_launches);
/* Autotuner (no counterpart in the reference): times the pair kernel over its launch shapes (1, 2, 4 or 8 waves
 * per block; for a batch also the block order: 0 system order, 1 round-robin across systems, 2 each system's
 * blocks on one XCD), `trials` launches each, and keeps the fastest for later runs.  Forces are fixed-point
 * integer sums, so the shape never changes a result. */
og_status   og_autotune(og_system *sys, int trials, int *waves, double *pair_ms);
og_status   og_batch_autotune(og_batch *b, int trials, int *waves, int *order, double *pair_ms);
og_status   og_set_pair_shape(og_system *sys, int waves);
/* The pair kernel's launch shape of a system: waves per block and block order (0 system order, 3 the
 * block list in eight contiguous parts, one per XCD); og_autotune may choose either order. */
og_status   og_get_pair_shape(og_system *sys, int *waves, int *order);
og_status   og_set_pair_shape_order(og_system *sys, int waves, int order);
/* Cell-edge autotuner (no counterpart in the reference): for a system on the
 * full z-bin scan (gap_bug 0) times the pair kernel on grids whose column
 * width is 0.8, 0.9, 1, 1.12 and 1.25 times the reference's (temporary
 * systems, each at its best launch shape) and returns the fastest scale for
 * og_params.cell_edge_scale; the pair set within the cutoff is the same on
 * every grid.  With gap_bug 1 it returns 1. */
og_status   og_autotune_cell_edge(const og_system_desc *desc, const og_params *params, int trials,
                                  double *best_scale, double *pair_ms);
og_status   og_batch_set_pair_shape(og_batch *b, int waves, int order);
/* 1 when an MD step of this system runs as the fused one-kernel Langevin step
 * (Langevin, SETTLE waters bound to nothing outside the water, at most 4096 other atoms) */
og_status   og_step_fused(og_system *sys, int *fused);
/* device memory a created system holds (bytes) */
og_status   og_system_device_bytes(og_system *sys, uint64_t *bytes);
/* bytes of device memory a system with this description would hold (an upper estimate, used for admission) */
og_status   og_system_footprint(const og_system_desc *desc, const og_params *params, uint64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
