/*
 * lima_hip.h -- C ABI of the MI355X-native LIMA force / minimization path.
 *
 * The library (lima_amd/lib/liblima_hip.so) replaces LIMA_ENGINE's per-step
 * force path and its energy-minimization entry point.  Every entry point below
 * names the reference interface it stands in for (paths are relative to the
 * LIMA source tree).
 *
 * Conventions
 *   - Plain C, no exceptions cross the ABI.  Every function returns an
 *     lima_status; lima_last_error() returns a thread-local message for the
 *     last failing call on the calling thread.
 *   - The caller owns every host buffer it passes; lima_create deep-copies the
 *     box.  The engine owns all device memory and frees it in lima_destroy.
 *   - An engine handle is not thread-safe.  One HIP stream set per handle.
 *   - Units are LIMA's (code/LIMA_BASE/include/Constants.h:43-66,
 *     LimaTypes.cuh:297-334): length nm, positions as int32 Coord in 1e-8 nm
 *     relative to an integer-nm origo, velocity nm/ns, mass kg/mol, charge
 *     kC/mol, energy J/mol, force J/mol/nm, time steps given in fs.
 */
#ifndef LIMA_HIP_H
#define LIMA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIMA_MAX_COMPOUND_PARTICLES 32 /* Constants.h:91 MAX_COMPOUND_PARTICLES */
#define LIMA_LUT_SLOTS 5               /* LimaTypes.cuh:455 max_bonded_compounds */
#define LIMA_NANO_TO_LIMA 100000000    /* LimaTypes.cuh:299 Coord::nanoToLima_i */

typedef enum lima_status {
    LIMA_OK = 0,
    LIMA_ERR_HIP = 1,          /* a HIP runtime call failed */
    LIMA_ERR_BAD_INPUT = 2,    /* descriptor / parameter validation failed */
    LIMA_ERR_LIST_OVERFLOW = 3,/* a neighbour list or cell column overflowed */
    LIMA_ERR_NONFINITE = 4,    /* NaN / inf force or energy detected */
    LIMA_ERR_NO_DEVICE = 5,    /* no gfx950 device visible */
    LIMA_ERR_UNSUPPORTED = 6   /* feature outside this implementation's scope */
} lima_status;

/* Boundary condition select, LIMA BoundaryConditionSelect (BoundaryConditionPublic.h).
 * Only periodic boxes are implemented. */
typedef enum lima_bc { LIMA_BC_NONE = 0, LIMA_BC_PBC = 1 } lima_bc;

/*
 * Box descriptor: the flat-array form of LIMA's Box (Simulation.cuh:137-157).
 *
 * Compounds keep LIMA's layout (Bodies.cuh:193-279): up to 32 particles per
 * compound, an integer-nm origo (NodeIndex) and int32 Coord positions relative
 * to it (CompoundCoords).  Per-particle arrays are [n_compounds][32].
 * Tinymols are LIMA's single-site LJ solvents (BoxGrid.cuh:12-60): each has
 * the NodeIndex of its solvent block and a Coord relative to that block.
 *
 * Exclusions are LIMA's BondedParticlesLUT set (LimaTypes.cuh:415-490,
 * CompoundBuilder.cpp:17-93): bonded_lut[c][k][i] bit j is set when particle i
 * of compound c and particle j of compound c+k-2 share any bonded term (or are
 * the same particle).  Compounds c and c+k-2 with any bit set are "bonded
 * compounds": their non-excluded pairs interact without cutoff
 * (LennardJonesInteractions.cuh:121-161).
 *
 * Bonded terms are LIMA's BondGroup contents (Bodies.cuh:281-307) flattened;
 * particle ids are compound*32 + slot.  Parameters are in LIMA's stored units
 * (Bodies.cpp:9-47): dihedral k_phi is the GROMACS k converted by
 * DihedralBond::Parameters::CreateFromCharmm, i.e. kJ*1000/2.
 */
typedef struct lima_box_desc {
    double box_nm[3];                 /* orthorhombic edges [nm] */
    int32_t n_compounds;
    int32_t n_tinymols;

    const int32_t*  compound_n_particles;  /* [n_compounds] */
    const int32_t*  compound_origo;        /* [n_compounds][3] NodeIndex [nm] */
    const int32_t*  compound_key_particle; /* [n_compounds] centerparticle_index */
    const int32_t*  particle_coord;        /* [n_compounds][32][3] Coord rel. origo */
    const float*    particle_velocity;     /* [n_compounds][32][3] nm/ns, NULL = 0 */
    const float*    particle_charge;       /* [n_compounds][32] kC/mol */
    const float*    particle_sigma;        /* [n_compounds][32] nm */
    const float*    particle_epsilon;      /* [n_compounds][32] J/mol */
    const float*    particle_mass;         /* [n_compounds][32] kg/mol */
    const uint32_t* bonded_lut;            /* [n_compounds][5][32] */

    const int32_t*  tinymol_block;         /* [n_tinymols][3] NodeIndex */
    const int32_t*  tinymol_coord;         /* [n_tinymols][3] Coord rel. block */
    const float*    tinymol_velocity;      /* [n_tinymols][3], NULL = 0 */
    const float*    tinymol_sigma;         /* [n_tinymols] nm */
    const float*    tinymol_epsilon;       /* [n_tinymols] J/mol */
    const float*    tinymol_mass;          /* [n_tinymols] kg/mol */

    int32_t n_bonds;     const int32_t* bond_ids;     const float* bond_params;     /* [n][2] ids, [n][2] b0 nm, kb J/mol/nm^2 */
    int32_t n_angles;    const int32_t* angle_ids;    const float* angle_params;    /* [n][3] ids, [n][4] theta0 rad, ktheta J/mol/rad^2, ub0 nm, kub J/mol/nm^2 */
    int32_t n_dihedrals; const int32_t* dihedral_ids; const float* dihedral_params; /* [n][4] ids, [n][3] phi0 rad, kphi J/mol, n */
    int32_t n_impropers; const int32_t* improper_ids; const float* improper_params; /* [n][4] ids, [n][2] psi0 rad, kpsi J/mol/rad^2 */
    int32_t n_pairs;     const int32_t* pair_ids;     const float* pair_params;     /* [n][2] ids, [n][2] sigma nm, eps J/mol */
} lima_box_desc;

/* LIMA SimParams (Simulation.cuh:17-46) plus the knobs of this implementation. */
typedef struct lima_simparams {
    int64_t n_steps;
    float   dt_fs;                    /* SimParams::dt, given in fs */
    int32_t em_variant;               /* SimParams::em_variant */
    float   em_force_tolerance;       /* kJ/mol/nm */
    int32_t bc_select;                /* lima_bc, PBC only */
    int32_t enable_electrostatics;
    float   cutoff_nm;
    int32_t data_logging_interval;
    int32_t apply_thermostat;
    int64_t steps_per_temperature_measurement;
    /* additions */
    int32_t pme_grid[3];              /* 0 = derived from pme_spacing_nm */
    float   pme_spacing_nm;           /* target PME spacing, default 0.12 */
    float   nlist_skin_nm;            /* Verlet buffer, default 0.30 */
    int32_t nlist_interval;           /* most steps per list, default 40.  Lists start at 10 steps (or
                                         this, if less) and grow by a step after 20 rebuilds in a row whose
                                         lists stayed complete.  An interval after which the two largest
                                         displacements since the build sum to more than the skin is undone
                                         and run again with lists a sixth shorter */
    int32_t energy_interval;          /* steps between FP64 energy reductions (0 = never) */
    int32_t launch_mode;              /* 0 = each MD step replayed as a HIP graph (captured again only
                                         when the list or the PME bins are reallocated), intra /
                                         bonded terms and PME on a second stream; or-ed flags
                                         LIMA_LAUNCH_EAGER (launch the kernels one by one; the
                                         integrator then also packs the next step's slots) and
                                         LIMA_LAUNCH_ONE_STREAM (PME in line) */
    int32_t log_capacity;             /* frames the device data log holds (0 = no logging); with
                                         data_logging_interval > 0 every step % interval == 0 is
                                         evaluated with energies and logged (lima_read_log) */
    int32_t nlist_capacity;           /* pair-list entries per i-cluster to start with (0 = 2048); a
                                         build that needs more grows it to 1.5 x the largest count */
} lima_simparams;

#define LIMA_LAUNCH_EAGER 1
#define LIMA_LAUNCH_ONE_STREAM 2

/* FP64 energy totals [J/mol]. */
typedef struct lima_energy {
    double lj;            /* LJ, cutoff pairs + intra/bonded-compound pairs */
    double coulomb_sr;    /* erfc-Coulomb, same pair set */
    double bonded;        /* bonds + angles(+UB) + dihedrals + impropers + 1-4 pairs */
    double pme_recip;     /* 1/2 sum_i q_i phi_i */
    double pme_self_lima; /* LIMA's per-charged-particle self term, PME.cuh:417-420,610-630 */
    double pme_self_ewald;/* standard -kappa/sqrt(pi) k sum q^2 */
    double potential;     /* lj + coulomb_sr + bonded + pme_recip + pme_self_ewald */
    double potential_lima;/* sum of LIMA's per-particle potE (uses pme_self_lima) */
    double kinetic;
    double total;         /* potential + kinetic */
    double virial[6];     /* potential virial W = -dU/d(strain) = sum over pairs of r_ij (x) F_ij, plus the
                             bonded terms' sum_a (x_a - x_ref) (x) F_a and the PME reciprocal-space sum;
                             components xx, yy, zz, xy, xz, yz [J/mol].  Pressure tensor
                             P = (sum_i m_i v_i (x) v_i + W) / V.  No counterpart in LIMA. */
    int64_t step;         /* MD step whose evaluation produced these totals (lima_compute_forces: the
                             current step; lima_minimize: the current step after minimising); -1
                             before any energy evaluation */
} lima_energy;

/* LIMA RunStatus mirror (Engine.cuh:63-75). */
typedef struct lima_runstatus {
    int64_t current_step;
    float   current_temperature;
    float   greatest_force;       /* kJ/mol/nm, max |F| at last check */
    int32_t simulation_finished;
    int32_t critical_error_occured;
    int32_t n_list_rebuilds;
    int64_t list_entries;         /* pair-list entries (8 pair slots each) of the last rebuild */
    int64_t list_violations;      /* kept list intervals whose largest plus second-largest
                                     displacement since the build exceeded the skin (a pair could
                                     have come inside the cutoff unlisted).  Such an interval is
                                     undone and run again with shorter lists, so this stays 0 in MD;
                                     see list_stale_rollbacks */
    float   max_displacement_nm;  /* max displacement over the last completed list interval */
    int32_t capacity_rollbacks;   /* times issued steps were undone and run again after a capacity
                                     overflow (see lima_step) */
    int32_t list_capacity;        /* pair-list entry slots per i-cluster */
    int32_t list_max_entries;     /* slots the fullest i-cluster of the current list uses */
    int32_t list_interval;        /* steps per list now in use (at most nlist_interval) */
    int32_t list_stale_rollbacks; /* list intervals undone and run again because the list could have
                                     missed a pair */
} lima_runstatus;

/* State read-back.  Arrays the caller provides (NULL = skip) in LIMA's
 * [n_compounds][32] / [n_tinymols] layouts. */
typedef struct lima_state_out {
    int32_t* compound_origo;      /* [n_compounds][3] */
    int32_t* particle_coord;      /* [n_compounds][32][3] */
    float*   particle_velocity;   /* [n_compounds][32][3] */
    float*   particle_force;      /* [n_compounds][32][3] J/mol/nm, forces of the last evaluation */
    float*   particle_pot_energy; /* [n_compounds][32] J/mol, LIMA per-particle potE of the last energy evaluation */
    int32_t* tinymol_block;       /* [n_tinymols][3] */
    int32_t* tinymol_coord;       /* [n_tinymols][3] */
    float*   tinymol_velocity;    /* [n_tinymols][3] */
    float*   tinymol_force;       /* [n_tinymols][3] */
    float*   tinymol_pot_energy;  /* [n_tinymols] */
} lima_state_out;

/* Steepest-descent result (Programs::EnergyMinimize, Programs.cpp:155-199). */
typedef struct lima_em_result {
    int64_t iterations;
    int64_t accepted;
    double  final_energy;         /* FP64 potential, J/mol */
    double  final_fmax;           /* kJ/mol/nm */
    double  wall_seconds;
    int32_t converged;
} lima_em_result;

typedef struct lima_engine lima_engine;

/* ---- engine lifetime: Engine(std::unique_ptr<Simulation>, BoundaryConditionSelect, ...)
 *      code/LIMA_ENGINE/include/Engine.cuh:78-99, Engine.cu:19-81 ---- */
lima_status lima_create(const lima_box_desc* box, const lima_simparams* params, lima_engine** out);
/* Engine::~Engine / terminateSimulation, Engine.cu:192-199 */
lima_status lima_destroy(lima_engine* engine);

/* ---- Engine::step(), Engine.cu:148-159: nsteps velocity-Verlet steps ----
 * Asynchronous, as LIMA's kernel launches are: the call issues the steps and returns; the device
 * runs at most two neighbour-list intervals behind the host.  Conditions the device meets (a pair
 * list, cell column or PME tile bin over capacity, a non-finite force) are found at a later list
 * rebuild or at the next lima_synchronize: a capacity overflow is undone by restoring the state of
 * the last list rebuild, growing the capacity and running those steps again, so callers never see
 * its results; a non-finite force puts the engine in its critical state (LIMA_ERR_NONFINITE, as
 * RunStatus::critical_error_occured).  Every call that returns data (lima_get_state,
 * lima_get_energy, lima_get_status, lima_read_log, ...) synchronises first.  Energies are evaluated
 * on every energy_interval-th step and on logged steps only. */
lima_status lima_step(lima_engine* engine, int64_t nsteps);

/* Waits for every issued step and reports what the device met (see lima_step).  No counterpart
 * in LIMA, whose step() waits for the device itself (Engine.cu:297-300). */
lima_status lima_synchronize(lima_engine* engine);

/* One force (+ FP64 energy) evaluation at the current positions, no integration.
 * The force half of _deviceMaster (Engine.cu:297-352). */
lima_status lima_compute_forces(lima_engine* engine, lima_energy* energy_out);

/* FP64 energy totals of the last energy evaluation: lima_compute_forces, lima_minimize, or an
 * energy step of lima_step (every energy_interval-th step, every logged step); .step names it. */
lima_status lima_get_energy(lima_engine* engine, lima_energy* energy_out);

/* ---- Programs::EnergyMinimize, Programs.cpp:155-199 (EM entry, SimParams::em_variant):
 *      steepest descent with FP64 energy acceptance ---- */
lima_status lima_minimize(lima_engine* engine, double fmax_tol_kj, int64_t max_iter,
                          double initial_step_nm, lima_em_result* result);

/* ---- takeBackSim / CopyDataToHost, EngineHostside.h:10-14, SimulationDevice.cpp:74-96 ---- */
lima_status lima_get_state(lima_engine* engine, lima_state_out* out);
/* ---- Engine::runstatus, Engine.cuh:95 ---- */
lima_status lima_get_status(lima_engine* engine, lima_runstatus* out);

/* ---- NeighborLists::updateNlists, Neighborlists.cuh:294-317: force a rebuild and
 *      read back the cluster list.  Counts: clusters, list entries. ---- */
lima_status lima_get_neighbor_counts(lima_engine* engine, int64_t* n_clusters, int64_t* n_entries);
/* cluster_atoms[n_clusters*8]: global particle index or -1; entry_offsets[n_clusters+1];
 * entries[n_entries]: j particle index | (shift code << 24); entry_masks[n_entries]: bit i set = pair
 * (cluster slot i, j) is computed in the cutoff kernel. */
lima_status lima_get_neighbor_list(lima_engine* engine, int32_t* cluster_atoms, int64_t* entry_offsets,
                                   uint32_t* entries, uint8_t* entry_masks);

/* ---- Data logging: LIMA's DatabuffersDeviceController (SimulationDevice.cuh:45-75), filled by
 * LogCompoundData / LogSolventData (EngineUtils.cuh:143-173) at every step % data_logging_interval
 * == 0 and offloaded by Engine::offloadLoggingData (Engine.cu:203-240).  Frames use LIMA's buffer
 * layout of total_particles_upperbound = n_compounds * 32 + n_tinymols entries (compound c slot s
 * at c * 32 + s, tinymol t at n_compounds * 32 + t; unused compound slots are zero):
 *   position    [nm] after the step's integration, compound particles as origo + Coord with the
 *               origo at the key particle's wrapped nm node, tinymols wrapped into the box
 *               (LIMAPOSITIONSYSTEM::GetAbsolutePositionNM after applyBC);
 *   pot_energy  the particle's LIMA potE of the step (as lima_state_out.particle_pot_energy);
 *   speed       |v| after the step (thermostat scale applied);
 *   force       the step's force [J/mol/nm].
 * The device holds the last log_capacity frames (a ring); the host copies frames on request, so
 * logging adds no host synchronisation to the step loop.  Frame k is the k-th logged step since
 * lima_create. ---- */
typedef struct lima_log_info {
    int64_t frames_logged;        /* frames written since lima_create */
    int64_t first_held;           /* oldest frame still held (frames_logged - held) */
    int64_t particles_per_frame;  /* n_compounds * 32 + n_tinymols */
    int32_t interval;             /* data_logging_interval */
    int32_t capacity;             /* log_capacity */
} lima_log_info;
lima_status lima_get_log_info(lima_engine* engine, lima_log_info* out);
/* Arrays the caller provides for n_frames frames (NULL = skip), P = particles_per_frame. */
typedef struct lima_log_out {
    int64_t*     step;            /* [n_frames] */
    lima_energy* energy;          /* [n_frames] FP64 totals of the logged step */
    float*       position;        /* [n_frames][P][3] */
    float*       pot_energy;      /* [n_frames][P] */
    float*       speed;           /* [n_frames][P] */
    float*       force;           /* [n_frames][P][3] */
} lima_log_out;
/* Frames first_frame .. first_frame + n_frames - 1; LIMA_ERR_BAD_INPUT when any of them is no
 * longer (or not yet) held. */
lima_status lima_read_log(lima_engine* engine, int64_t first_frame, int64_t n_frames, lima_log_out* out);

/* Thread-local message of the last error on this thread. */
const char* lima_last_error(void);
/* Per-phase GPU times [ms] accumulated since create (HIP events). */
typedef struct lima_timings {
    double pair_ms, intra_ms, bonded_ms, pme_ms, integrate_ms, nlist_ms, total_ms;
    int64_t steps_timed;
    /* host side of lima_step, accumulated over every call: time spent issuing steps (list
     * rebuilds included, the final wait for the device excluded), and the steps issued */
    double host_issue_ms;
    int64_t host_steps;
} lima_timings;
lima_status lima_get_timings(lima_engine* engine, lima_timings* out);
/* Enable per-phase HIP-event timing (no host synchronisation per step; the events are
 * summed when lima_get_timings is called). */
lima_status lima_set_timing(lima_engine* engine, int32_t enabled);
/* Particle pairs inside the cutoff counted by the pair kernel at the last energy evaluation
 * (the algorithmic work unit of the pair kernel), after the issued steps have finished; -1 for a
 * null engine or an error (lima_last_error). */
int64_t lima_get_pair_count(lima_engine* engine);

/* ---- Batched minimisation: Programs::EnergyMinimize (Programs.cpp:155-199) applied to many
 * independent small systems at once.  LIMA minimises one Simulation per Environment; here one
 * handle holds n systems, each minimised by steepest descent with FP64 acceptance exactly as
 * lima_minimize does, all of them in one persistent launch per workgroup shape (one workgroup
 * per system, no host round trip per iteration).  Limits per system: 1..8192 atoms, box edges
 * in (cutoff, 10) nm, PME grid edges in [4, 32]. ---- */
typedef struct lima_batch lima_batch;

/* Workgroup shape per system-size bucket: systems with n atoms use the first bucket with
 * n <= bucket_max_atoms.  threads: 256, 512 or 1024; pair_unroll: unroll factor of the pair loop
 * (1, 2 or 4).  us_per_iter: cost of one SD iteration measured by the autotuner. */
typedef struct lima_batch_tuning {
    int32_t n_buckets;
    int32_t bucket_max_atoms[8];
    int32_t threads[8];
    int32_t pair_unroll[8];
    double  us_per_iter[8];
} lima_batch_tuning;

/* boxes[n_systems]; params apply to every system (cutoff, electrostatics, PME spacing/grid). */
lima_status lima_batch_create(const lima_box_desc* boxes, int32_t n_systems, const lima_simparams* params,
                              lima_batch** out);
lima_status lima_batch_destroy(lima_batch* batch);
/* Times every candidate shape for trial_iterations SD iterations on up to 8 systems per size
 * bucket, keeps the fastest per bucket, and restores the starting positions. */
lima_status lima_batch_autotune(lima_batch* batch, int32_t trial_iterations, lima_batch_tuning* out);
lima_status lima_batch_set_tuning(lima_batch* batch, const lima_batch_tuning* tuning);
lima_status lima_batch_get_tuning(lima_batch* batch, lima_batch_tuning* out);
/* Minimises every system from its current positions; results[n_systems] per system
 * (wall_seconds: the system's own time inside the launch); *wall_seconds: the whole batch. */
lima_status lima_batch_minimize(lima_batch* batch, double fmax_tol_kj, int64_t max_iter, double initial_step_nm,
                                lima_em_result* results, double* wall_seconds);
/* State of one system in LIMA's layout (velocities are zero; forces of the accepted state). */
lima_status lima_batch_get_state(lima_batch* batch, int32_t system, lima_state_out* out);
/* Restores the positions every system had at lima_batch_create. */
lima_status lima_batch_reset(lima_batch* batch);
/* ---- Learned step proposals.  A per-atom MLP (32 features: F / Fmax, 1 + log10(|F| / Fmax) / 6, and 28 counts of
 * neighbours within 1.0 nm in 14 distance bins per charge sign; two SiLU layers of 64; output a
 * step scale s in (0, 4)) proposes x' = x + h s_i F_i / max_j (s_j |F_j|): a per-atom preconditioned
 * step whose largest displacement is h, as for the plain step (a constant s gives the plain step).
 * The proposal is kept only if the FP64 energy decreases; otherwise the plain steepest-descent step
 * is taken.  No counterpart in LIMA; the minimisation is still Programs::EnergyMinimize's
 * (Programs.cpp:155-199). ---- */
typedef struct lima_mlp lima_mlp;
int32_t lima_mlp_n_params(void);
lima_status lima_mlp_create(uint64_t seed, lima_mlp** out);
lima_status lima_mlp_destroy(lima_mlp* model);
/* FP32 parameters: W1 [32][64], b1 [64], W2 [64][64], b2 [64], w3 [64], b3 */
lima_status lima_mlp_get_params(lima_mlp* model, float* params);
lima_status lima_mlp_set_params(lima_mlp* model, const float* params);
/* s for n feature rows [n][32] (BF16 MFMA inference, FP32 accumulation) */
lima_status lima_mlp_forward(lima_mlp* model, const float* features, int64_t n, float* scale_out);
/* Allocation-free variant of lima_mlp_forward.  No counterpart in LIMA: it is for callers that
 * cannot allocate device memory (code under stream capture, tight serving loops).  The features and
 * scales are staged in a fixed device array of LIMA_MLP_NOALLOC_ROWS rows (8.65 MB per GPU, held until
 * the process ends) instead of buffers allocated per call; n > LIMA_MLP_NOALLOC_ROWS returns
 * LIMA_ERR_BAD_INPUT.  The array is shared by every call on a GPU and keeps the last call's contents;
 * the call holds a process-wide lock until its results are on the host, so calls from several host
 * threads run one after another (calls that overlapped on the array would race).  Results equal
 * lima_mlp_forward's bit for bit. */
#define LIMA_MLP_NOALLOC_ROWS 65536
lima_status lima_mlp_forward_noalloc(lima_mlp* model, const float* features, int64_t n, float* scale_out);
/* Adam on 1/2 (s - t)^2, BF16 MFMA forward and backward with FP32 accumulation and master weights;
 * batch a multiple of 16.  loss_per_epoch[epochs] (may be NULL), *seconds (may be NULL). */
lima_status lima_mlp_train(lima_mlp* model, const float* features, const float* targets, int64_t n, int32_t epochs,
                           int32_t batch, float learning_rate, uint64_t seed, float* loss_per_epoch, double* seconds);
/* Gradient of sum_samples 1/2 (s - t)^2 w.r.t. the FP32 parameters (the training kernel's own
 * product), grad_out[lima_mlp_n_params()]; *loss = the summed loss. */
lima_status lima_mlp_gradient(lima_mlp* model, const float* features, const float* targets, int64_t n, float* grad_out,
                              double* loss);
/* Batched minimisation with learned proposals; the step-scale features' neighbour histograms are
 * refreshed every feature_interval iterations; grow_on_learned != 0 grows h by 1.2 after an accepted
 * proposal as after an accepted plain step (0 keeps h, so h tracks the plain step alone).
 * Systems of up to 4096 atoms. */
lima_status lima_batch_minimize_learned(lima_batch* batch, lima_mlp* model, double fmax_tol_kj, int64_t max_iter,
                                        double initial_step_nm, int32_t feature_interval, int32_t grow_on_learned,
                                        lima_em_result* results, double* wall_seconds);
/* Plain batched minimisation that records, at every accepted state, the features, the plain step
 * and the Coord of every atom_stride-th atom (up to max_states states per system). */
lima_status lima_batch_record(lima_batch* batch, double fmax_tol_kj, int64_t max_iter, double initial_step_nm,
                              int32_t atom_stride, int32_t max_states, int64_t* n_states);
/* Training samples from the recording: features [n][32] and the target step scale.  rho, the secant
 * estimate of the atom's best step along its displacement D to the state `lookahead` accepted states
 * later in units of its plain step (|D| f / (f - f') / |d|, f and f' the force components along D
 * before and after), mapped log-linearly from [0.1, 1000] onto [0.05, 3.95].  With features == NULL
 * only *n_samples is set. */
lima_status lima_batch_training_set(lima_batch* batch, int32_t lookahead, float* features, float* targets, int64_t capacity,
                                    int64_t* n_samples);
/* Feature rows [n_atoms][32] of one system as last computed (slots 4..31: neighbour histogram,
 * count / 8; slots 0..3 are filled from the forces at inference time and read here as 0). */
lima_status lima_batch_get_features(lima_batch* batch, int32_t system, float* out);
/* Per system [n][3]: force evaluations, iterations whose learned proposal was accepted, recorded states
 * of the last call. */
lima_status lima_batch_get_counters(lima_batch* batch, int64_t* out);

/* Time each system spent in the last batched minimisation, by phase, as seen by its workgroup's
 * first thread: seconds[n_systems][7] = cutoff tile pairs, per-atom sums + intra pairs + bonded
 * terms, PME, reductions, trial steps, feature histograms, scale inference (the last two in the
 * learned mode only). */
lima_status lima_batch_get_phase_times(lima_batch* batch, double* seconds);

#ifdef __cplusplus
}
#endif
#endif /* LIMA_HIP_H */
