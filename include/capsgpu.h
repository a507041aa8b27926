/* capsgpu — the MI355X force-and-step path of CAPS behind a plain C ABI.
 *
 * The library replaces, for one GPU, the two C++ interfaces CAPS's protocols call for dynamics and energies:
 *   caps::Evaluator (core/include/caps/field.hpp:327-352: construct from a ForceField + EnergyOptions, then
 *     compute(x, cell, f) -> EnergyTerms, energies(bool), list_builds())                      -> capsgpu_create / _compute
 *   caps::run_dynamics (core/include/caps/dynamics.hpp:180; loop core/src/dynamics.cpp:606-774) -> capsgpu_run_dynamics
 * next to the stable C surface of capi (capi/include/caps_c.h:149-193: caps_md, caps_last_error). Conventions follow
 * capi: every call returns 0 on success or a negative error class, and capsgpu_last_error() holds the message of the
 * calling thread's last failure (capi/src/caps_c.cpp:353, 373-383).
 *
 * Arrays are host memory owned by the caller; the state (positions, velocities, lists, meshes) lives in device memory
 * owned by the context. A context is single-caller, as an Evaluator is.
 * Units: Å, fs, kcal/mol, e, g/mol; positions and forces 3N doubles, atom-major (x0 y0 z0 x1 ...).
 * A cell is double[12]: origin, a, b, c (Cartesian edge vectors), periodic along all three axes.
 */
#ifndef CAPSGPU_H
#define CAPSGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAPSGPU_OK 0
#define CAPSGPU_EINVAL (-1)     /* a bad argument (std::invalid_argument in the reference) */
#define CAPSGPU_EFIELD (-2)     /* a force-field error (caps::FieldError): FENE bond past R0, unstable run */
#define CAPSGPU_EDEVICE (-3)    /* a HIP runtime failure or no GPU */
#define CAPSGPU_ECANCEL (-4)    /* the progress callback asked to stop (caps::DynamicsCancelled) */

typedef struct capsgpu_ctx capsgpu_ctx;

/* The flattened force field: what caps::ForceField (core/include/caps/field.hpp:154-249) holds for the plain pair,
 * bonded and FENE terms this path evaluates. */
typedef struct capsgpu_ff {
  int32_t n_atoms;
  int32_t n_types;
  const int32_t* type;          /* [n] type index (ForceField::type_index) */
  const double* charge;         /* [n] charges, e (already divided by √εr) */
  const double* mass;           /* [n] g/mol */
  const double* eps;            /* [nt·nt] mixed ε of each type pair (caps::mixed_pair, field.cpp:2593-2604) */
  const double* sigma;          /* [nt·nt] mixed σ */
  int32_t lj96;                 /* 0: 12-6 4ε[(σ/r)¹² − (σ/r)⁶]; 1: 9-6 ε[2(σ/r)⁹ − 3(σ/r)⁶] */
  const int32_t* excl_offset;   /* [n+1] CSR of the excluded partners, sorted (ForceField::excluded) */
  const int32_t* excl_index;
  const int32_t* p14_offset;    /* [n+1] or NULL: the excluded partners that are 1-4 pairs (Coulomb keeps coul14) */
  const int32_t* p14_index;
  double coul14;
  int32_t n_bonds;              /* harmonic k (r − r0)²: [2·nb] atoms, [2·nb] (k, r0) */
  const int32_t* bond_atoms;
  const double* bond_param;
  int32_t n_angles;             /* harmonic k_θ (θ − θ0)², j the vertex: [3·na] atoms, [2·na] (k_θ, θ0 in radians) */
  const int32_t* angle_atoms;
  const double* angle_param;
  int32_t n_fene;               /* FENE −½KR0² ln(1 − (r/R0)²) + WCA(ε, σ): [2·nf] atoms, [4·nf] (K, R0, ε, σ) */
  const int32_t* fene_atoms;
  const double* fene_param;
} capsgpu_ff;

/* caps::EnergyOptions (core/include/caps/field.hpp:274-308), the fields this path reads */
typedef struct capsgpu_energy_opts {
  double cutoff;                /* Å */
  double skin;                  /* Å */
  int32_t coulomb;              /* 0 off, 1 on */
  int32_t electrostatics;       /* 0 damped shifted force, 1 PME */
  double dsf_alpha;             /* Å⁻¹ */
  double ewald_rtol;
  double pme_spacing;           /* Å */
  int32_t pme_order;            /* 3 … 8 */
  int32_t tail;                 /* LJ tail correction (else LJ shifted to zero at the cut-off) */
  int32_t parts;                /* bit 1 bonded, bit 2 non-bonded (pairs, k-space, self, tail), 3 both (field.hpp:301-303);
                                   bit 4 (no reference counterpart): PME's reciprocal part with the Ewald self and
                                   background terms, without the pair list (the slow force of learned r-RESPA) */
  int32_t no_reciprocal;        /* 1: PME's reciprocal part, self and background terms left out (the real-space part
                                   alone, for checks); 0 the evaluator as the reference's */
} capsgpu_energy_opts;

/* caps::EnergyTerms (core/include/caps/field.hpp:310-320) as 15 doubles */
enum {
  CAPSGPU_T_BOND = 0, CAPSGPU_T_ANGLE, CAPSGPU_T_DIHEDRAL, CAPSGPU_T_IMPROPER, CAPSGPU_T_VDW, CAPSGPU_T_COULOMB,
  CAPSGPU_T_VIRIAL, CAPSGPU_T_WXX, CAPSGPU_T_WYY, CAPSGPU_T_WZZ, CAPSGPU_T_WXY, CAPSGPU_T_WXZ, CAPSGPU_T_WYZ,
  CAPSGPU_T_TAIL_ENERGY, CAPSGPU_T_TAIL_VIRIAL, CAPSGPU_NTERMS
};

/* The messages of the last failed call on this thread (caps_last_error, capi/include/caps_c.h:63). */
const char* capsgpu_last_error(void);

/* Evaluator(ff, options) (field.hpp:329): uploads the force field; the arrays may be freed afterwards. */
int32_t capsgpu_create(const capsgpu_ff* ff, const capsgpu_energy_opts* opts, capsgpu_ctx** out);
int32_t capsgpu_destroy(capsgpu_ctx* ctx);

/* Evaluator::compute(x, cell, f) (field.hpp:337, field.cpp:1249-2539): energy terms and forces at x. The pair list is
 * kept between calls and rebuilt when an atom has moved more than half the skin (field.cpp:800-838), pruned at
 * r_c + 0.8 Å when one has moved 0.4 Å (field.cpp:1306-1312). want_energies = 0 skips the pair energies and the virial
 * (Evaluator::energies(false), field.hpp:341). f: 3N out (overwritten). terms: CAPSGPU_NTERMS out. */
int32_t capsgpu_compute(capsgpu_ctx* ctx, const double* x, const double cell[12], double* f, double* terms,
                        int32_t want_energies);

/* capsgpu_compute without any device allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc). The per-evaluation scratch (PME mesh,
 * spectrum and influence function, spline and tile arrays, bonded member slots, partial sums) comes from one 96 MiB
 * block of device memory per GPU, held from the first call on that GPU until the process ends; the context's own
 * state (positions, forces, force field), allocated by capsgpu_create, is used as capsgpu_compute uses it. Takes
 * contexts with the bonded terms and PME's reciprocal part (parts 1, 4 or 5) and rejects, with CAPSGPU_EINVAL and a
 * message and before any work, a system whose scratch would exceed the block. Calls on one GPU share
 * the block: this library serialises its own calls, and calls that overlap them from elsewhere on the same GPU race.
 * Results are bit-identical to capsgpu_compute's. */
int32_t capsgpu_compute_noalloc(capsgpu_ctx* ctx, const double* x, const double cell[12], double* f, double* terms,
                                int32_t want_energies);

/* Evaluator::list_builds() (field.hpp:338) */
int32_t capsgpu_list_builds(capsgpu_ctx* ctx, int32_t* builds);

/* The device state for dynamics: positions (3N), velocities (3N, Å/fs, NULL: zero), cell. get_state reads them back
 * (a checkpoint, core/src/dynamics.cpp:759-763); any pointer may be NULL. */
int32_t capsgpu_set_state(capsgpu_ctx* ctx, const double* x, const double* v, const double cell[12]);
int32_t capsgpu_get_state(capsgpu_ctx* ctx, double* x, double* v, double* cell);
/* the forces of the last evaluation (3N) */
int32_t capsgpu_get_forces(capsgpu_ctx* ctx, double* f);

/* ---- energy minimisation -------------------------------------------------------------------------------------------
 * caps::minimise (core/src/relax.cpp:65-282; options core/include/caps/relax.hpp:28-32, stage :86-91): steepest descent,
 * Polak–Ribière CG (with its parabola step), L-BFGS (m = 10) or FIRE on the context's evaluation, every vector on the
 * device. Stops when the largest atomic force is below ftol, or the relative energy change below etol, or after
 * max_iterations, or when the line search fails twice. */
typedef struct capsgpu_relax_opts {
  int32_t method;               /* 0 steepest descent, 1 conjugate gradients, 2 L-BFGS, 3 FIRE (caps::Minimiser's order) */
  double ftol;                  /* kcal/mol/Å */
  double etol;
  int32_t max_iterations;
} capsgpu_relax_opts;

typedef struct capsgpu_relax_stage {
  int32_t iterations, evaluations;
  int32_t stopped_by;           /* 0 iterations, 1 force, 2 energy, 3 line search */
  double energy, fmax;          /* kcal/mol, kcal/mol/Å at the end */
} capsgpu_relax_stage;

/* x: 3N in, the minimised positions out (also the context's state); fixed: per atom 1 = held in place (NULL: none) */
int32_t capsgpu_minimise(capsgpu_ctx* ctx, double* x, const double cell[12], const capsgpu_relax_opts* opts, const uint8_t* fixed,
                         capsgpu_relax_stage* out);

/* caps::DynamicsOptions (core/include/caps/dynamics.hpp:60-154), the fields this path reads */
typedef int32_t (*capsgpu_checkpoint_fn)(const double* x, const double* v, const double cell[12], int64_t step, void* user);

typedef struct capsgpu_dyn_opts {
  double dt;                    /* fs */
  int64_t steps;
  double temperature;           /* K, thermostat target and initial velocities */
  int32_t thermostat;           /* 0 none (NVE), 1 Bussi, 2 Langevin (BAOAB) */
  double tau_t;                 /* fs: Bussi relaxation time, Langevin 1/γ */
  uint64_t seed;
  int32_t thermo_every;         /* steps between thermo rows */
  int32_t new_velocities;       /* 1: Maxwell–Boltzmann at temperature (dynamics.cpp:236-256), else the set state */
  int32_t respa;                /* r-RESPA inner steps for the bonded forces (1: off), Bussi or NVE only */
  int64_t step_offset;          /* added to reported steps; Langevin's noise is drawn for step + step_offset, so a run
                                   resumed from a checkpoint draws what the uninterrupted run would have */
  /* learned r-RESPA (no reference counterpart): 0 off; k > 1 a full PME evaluation every k steps, the MLP's predicted
   * increment of the reciprocal-space force in between, falling back to plain PME when its force RMSE exceeds
   * learned_tol (relative to the full force) at a full step */
  int32_t learned_k;
  double learned_tol;
  int32_t learned_train;        /* learned r-RESPA: PME frames (plain steps) the network trains on before it predicts;
                                   < 0: 2000. Afterwards it is frozen: later full steps judge it without training it.
                                   The network persists in the context across runs. */
  /* pressure coupling (DynamicsOptions::barostat ... barostat_every, dynamics.hpp:80-87; dynamics.cpp:701-723):
   * isotropic, every barostat_every steps after the step: 0 none, 1 Berendsen, 2 stochastic cell rescaling
   * (Bernetti & Bussi 2020), the cell and positions scaled about the cell origin, at most 1 % in volume per update */
  int32_t barostat;
  double pressure;              /* atm */
  double tau_p;                 /* fs */
  double compressibility;       /* atm⁻¹ */
  int32_t barostat_every;       /* steps between cell updates (< 1: 10) */
  double temperature_end;       /* K; > 0 ramps the thermostat target linearly from temperature to this over the run
                                   (DynamicsOptions::temperature_end, dynamics.hpp:80); <= 0 no ramp (a zeroed struct) */
  /* checkpoints (DynamicsOptions::checkpoint_every / checkpoint, dynamics.hpp:143-146; dynamics.cpp:758-763): every
   * checkpoint_every steps (0: none) before the last, after the step, the full state — positions (3N, Å, as held),
   * velocities (3N, Å/fs), cell and step (with step_offset) — when every coordinate and velocity is finite. The callback
   * returns 0 to go on; anything else ends the run with CAPSGPU_EINVAL. */
  int64_t checkpoint_every;
  capsgpu_checkpoint_fn checkpoint;
  void* checkpoint_user;
} capsgpu_dyn_opts;

/* caps::ThermoRow (core/include/caps/dynamics.hpp:28-43), the fields this path fills */
typedef struct capsgpu_thermo {
  int64_t step;
  double time_ps, temperature, kinetic, potential, total, pressure, conserved;
  double terms[CAPSGPU_NTERMS];
  int32_t learned;              /* learned r-RESPA: 1 this row's forces came from the MLP path */
  double learned_rmse;          /* learned r-RESPA: the last judged relative force RMSE (0 before the first) */
  double learned_rmse_stale;    /* learned r-RESPA: the same for the stale slow force held since the last full step */
  double volume, density;       /* Å³, g/cm³ */
  double p[6];                  /* pressure tensor, atm (kinetic + virial): xx yy zz xy xz yz */
} capsgpu_thermo;

/* Progress callback with each thermo row: return 0 to cancel (caps::DynamicsOptions::progress, dynamics.hpp:~141) */
typedef int32_t (*capsgpu_progress_fn)(const capsgpu_thermo* row, void* user);

typedef struct capsgpu_dyn_report {
  int64_t steps;
  double seconds, ns_per_day;
  int32_t list_builds, prunes;
  int32_t learned_steps, full_steps;
  int32_t fallbacks;            /* learned r-RESPA: judgments that missed learned_tol (each sends the run back to plain PME) */
  int32_t judgments;            /* learned r-RESPA: full steps that judged a k-step prediction against PME */
} capsgpu_dyn_report;

/* run_dynamics (dynamics.cpp:71-815) from the set state: velocity Verlet + Bussi, Langevin BAOAB, NVE, r-RESPA; the
 * state stays on the device and can be read back with capsgpu_get_state. rows: up to cap thermo rows written (NULL:
 * none), *nrows their count. */
int32_t capsgpu_run_dynamics(capsgpu_ctx* ctx, const capsgpu_dyn_opts* opts, capsgpu_progress_fn progress, void* user,
                             capsgpu_thermo* rows, int32_t cap, int32_t* nrows, capsgpu_dyn_report* report);

/* ---- kernel tile shapes (no counterpart in the reference: its kernels have fixed loop shapes) --------------------------
 * The mesh transforms run a tile of lines per workgroup: line pairs along a (tiles[0]) and lines along b and c
 * (tiles[1]), powers of two; tiles[0] = 0 (the default) runs the passes along a and b together, a whole c-plane per
 * workgroup in LDS, when the grid's planes fit (K2 even, the device's LDS), else 8 line pairs. The force gather runs a
 * workgroup per 4x4x4 tile (shape 0) or per run of 64 atoms of the tile order, one atom per lane (shape 1, the default,
 * where K1 is a multiple of 4; per tile otherwise). capsgpu_tune picks both for this context by timing each candidate on its own mesh (after a
 * reciprocal-space evaluation run here; needs PME and a set state) and keeps the fastest; capsgpu_set_tiles and
 * capsgpu_set_gather set them, capsgpu_get_gather reads the gather's. Every shape gives bit-identical forces. tiles,
 * in_use: the shape now in use (may be NULL). */
int32_t capsgpu_tune(capsgpu_ctx* ctx, int32_t tiles[2]);
int32_t capsgpu_set_tiles(capsgpu_ctx* ctx, int32_t lines_x, int32_t lines_yz, int32_t tiles[2]);
int32_t capsgpu_set_gather(capsgpu_ctx* ctx, int32_t shape, int32_t* in_use);
int32_t capsgpu_get_gather(capsgpu_ctx* ctx, int32_t* shape);

/* ---- replicas stepped together ------------------------------------------------------------------------------------
 * The reference runs the replicas of a cooling scan one after another, each through run_dynamics
 * (core/src/mechanics.cpp:1381-1470 via caps::run_cooling, core/include/caps/mechanics.hpp:323-353); here many contexts
 * (each its own force field, cell, PME grid and temperature) take Langevin BAOAB steps together. A chunk of steps of
 * every replica is captured once into HIP graphs and replayed; the launch shape (how many streams the replicas are
 * spread over, each replaying its own graph, and how many steps one graph holds) is picked by timing the candidates on
 * the batch itself, from a snapshot of the state that is restored afterwards; in the tuned mode each replica's kernel
 * tile shapes are first picked as capsgpu_tune does. Each replica ends bit-identical to
 * capsgpu_run_dynamics of that context alone with thermostat 2, seed + r and the same dt, tau_t and temperature.
 * The replicas' energy options must leave the pair path out (parts 1, 4 or 5): its list rebuilds need the host. */
typedef struct capsgpu_batch capsgpu_batch;

typedef struct capsgpu_batch_opts {
  double dt;                    /* fs */
  int64_t steps;
  double tau_t;                 /* fs: Langevin 1/γ */
  uint64_t seed;                /* replica r draws its noise (and new velocities) from seed + r */
  int32_t new_velocities;       /* 1: Maxwell–Boltzmann at each replica's temperature, else the set state */
  int32_t mode;                 /* 0: the tuned launch shape; 1: graphs with the shape below; 2: direct launches;
                                   3: grouped launches — each kernel of the step launched once for every replica (they
                                   need PME's reciprocal part with one spline order), graph_steps steps per graph */
  int32_t branches;             /* mode 1: streams the replicas are spread over (longest first), each with its graph */
  int32_t graph_steps;          /* modes 1 and 3: steps one graph holds */
} capsgpu_batch_opts;

typedef struct capsgpu_batch_report {
  int64_t steps;
  double seconds;               /* the steps alone: set-up, first forces and tuning left out */
  double atom_steps_per_s;      /* Σ atoms × steps / seconds */
  double ns_per_day;            /* simulated time of each replica per day */
  int32_t mode, branches, graph_steps;   /* the launch shape used */
  double tune_seconds;          /* the timing of the candidates (0 when the shape was given or already tuned) */
} capsgpu_batch_report;

/* a batch over n contexts (held, not owned: they must outlive the batch; one context at most once) */
int32_t capsgpu_batch_create(capsgpu_ctx* const* replicas, int32_t n, capsgpu_batch** out);
/* steps every replica from its set state at temperatures[r] (K); temperature_out[r] (NULL: not written): the kinetic
 * temperature at the end. The states stay on the device (capsgpu_get_state per context). */
int32_t capsgpu_batch_run(capsgpu_batch* batch, const capsgpu_batch_opts* opts, const double* temperatures, double* temperature_out,
                          capsgpu_batch_report* report);
int32_t capsgpu_batch_destroy(capsgpu_batch* batch);

/* ---- the glass-transition scan -------------------------------------------------------------------------------------
 * caps::run_cooling (core/src/mechanics.cpp:1381-1487, core/include/caps/mechanics.hpp:320-353): an unsampled hold at
 * t_start (equilibrate_ps; < 0: one hold; Bussi, seed + 1000), then a hold of ps_per_step at each temperature of the
 * ladder t_start, t_start ± t_step, … to t_end (Bussi τ_t, the barostat given, seed + k), the density and potential
 * energy averaged over each hold's rows from average_from on (the density's error from five blocks), and Tg fitted to the
 * specific volume (property 0) or the potential energy per atom (property 1): the hinge of a continuous two-line fit
 * (fit 0, mechanics.cpp:1218-1308) or the crossing of lines through a glassy and a rubbery range (fit 1, :1310-1379;
 * glassy_max / rubbery_min 0: the lowest and highest third), errors by a bootstrap of residuals. */
typedef struct capsgpu_cooling_opts {
  double t_start, t_end, t_step;   /* K */
  double ps_per_step;              /* hold at each temperature, ps */
  double equilibrate_ps;           /* < 0: one hold */
  double average_from;             /* fraction of each hold left out of the averages */
  double dt, tau_t, tau_p, pressure, compressibility;
  int32_t barostat;                /* as capsgpu_dyn_opts (2: stochastic cell rescaling, the reference's default) */
  int32_t barostat_every;
  uint64_t seed;
  int32_t new_velocities;
  int32_t property, fit;
  double glassy_max, rubbery_min;
} capsgpu_cooling_opts;

typedef struct capsgpu_cooling_point {
  double temperature, density, density_err, specific_volume, potential;   /* K, g/cm³, g/cm³, cm³/g, kcal/mol */
} capsgpu_cooling_point;

/* caps::BilinearFit (mechanics.hpp:305-313); note: 0 none, 1 too few temperatures, 2 the slopes differ by under 10 %,
 * 3 the hinge at the edge of the range, 4 the ranges overlap, 5 a range with under two temperatures, 6 parallel lines,
 * 7 a range of two temperatures (no error estimate), 8 the lines cross outside the gap between the ranges (no Tg) */
typedef struct capsgpu_tg_fit {
  int32_t ok, note;
  double tg, tg_err;               /* K */
  double slope_low, slope_high, value_at_tg, alpha_low, alpha_high, rss;
} capsgpu_tg_fit;

/* progress with each thermo row: index the hold (−1 the equilibration) of total; return 0 to cancel */
typedef int32_t (*capsgpu_cooling_progress_fn)(const capsgpu_thermo* row, int32_t index, int32_t total, void* user);

/* the scan of one context from its set state (left at the end of the scan); points: up to cap, *npoints the ladder's
 * length; fit may be NULL */
int32_t capsgpu_run_cooling(capsgpu_ctx* ctx, const capsgpu_cooling_opts* opts, capsgpu_cooling_progress_fn progress, void* user,
                            capsgpu_cooling_point* points, int32_t cap, int32_t* npoints, capsgpu_tg_fit* fit);
/* Many independent replicas scanned together (no reference counterpart: the reference scans one system at a time): the
 * holds run through the replica batch (Langevin τ_t at constant volume, so barostat 0 and property 1), every replica
 * through the same ladder, its potential energy sampled at `samples` evenly spaced points of the hold's averaged part.
 * points: [replicas][cap]; fits: one per replica (may be NULL). */
int32_t capsgpu_batch_cooling(capsgpu_batch* batch, const capsgpu_cooling_opts* opts, int32_t samples, capsgpu_cooling_point* points,
                              int32_t cap, int32_t* npoints, capsgpu_tg_fit* fits);
/* the Tg fit of given points (fit 0 or 1 as above) */
int32_t capsgpu_tg_fit_points(const double* T, const double* y, int32_t n, int32_t fit, double glassy_max, double rubbery_min,
                              capsgpu_tg_fit* out);

/* ---- equilibration protocols ---------------------------------------------------------------------------------------
 * caps::equilibrate (core/src/equilibrate.cpp:188-262; core/include/caps/equilibrate.hpp:15-37, 65-106) without the
 * convergence blocks: the stages run one after another from the context's state (no new velocities), stage k with seed
 * + 1000003 k, NVT and NPT stages under the thermostat given (none: Bussi), NPT ones under the barostat given (none:
 * stochastic cell rescaling), a thermo row every thermo_ps (0: about 200 per stage, between 10 steps and 0.5 ps); each
 * stage summarised by the means over the second half of its rows. Push-off stages (capped pair forces) are not built. */
typedef struct capsgpu_eq_stage {
  int32_t ensemble;             /* 0 NVT, 1 NPT, 2 NVE */
  double ps;                    /* duration */
  double t_start, t_end;        /* K; t_end > 0 ramps the thermostat target over the stage */
  double pressure;              /* atm (NPT) */
} capsgpu_eq_stage;

typedef struct capsgpu_eq_opts {
  double dt, tau_t, tau_p, compressibility;
  int32_t thermostat, barostat, barostat_every;   /* as capsgpu_dyn_opts */
  uint64_t seed;
  double thermo_ps;
} capsgpu_eq_opts;

typedef struct capsgpu_eq_summary {
  double ps, temperature, pressure, density, potential, density_end;
} capsgpu_eq_summary;

typedef struct capsgpu_protocol_params {      /* caps::ProtocolParams (equilibrate.hpp:24-37) */
  double t_final, t_max, p_final, p_max, time_scale;
  int32_t cycles;
  double t_low, t_high, ramp_ps, hold_ps;
} capsgpu_protocol_params;

/* progress with each thermo row of stage index (of total); return 0 to cancel */
typedef int32_t (*capsgpu_eq_progress_fn)(const capsgpu_thermo* row, int32_t index, int32_t total, void* user);

int32_t capsgpu_equilibrate(capsgpu_ctx* ctx, const capsgpu_eq_stage* stages, int32_t nstages, const capsgpu_eq_opts* opts,
                            capsgpu_eq_progress_fn progress, void* user, capsgpu_eq_summary* out);
/* the stages of a named protocol (equilibrate.cpp:43-75): "larsen21" (the 21-step compression and decompression of
 * Larsen et al. 2011) or "annealing" (NPT heat / hold / cool / hold cycles); *n the count, up to cap written to out */
int32_t capsgpu_protocol(const char* name, const capsgpu_protocol_params* params, capsgpu_eq_stage* out, int32_t cap, int32_t* n);

/* ---- inspection, for the parity checks (no counterpart in the reference's interfaces) ----------------------------- */
/* sizes of the current cluster list: atoms, clusters, slots, outer entries, image codes, columns nu, nv, pruned full
 * entries */
int32_t capsgpu_list_sizes(capsgpu_ctx* ctx, int64_t out[8]);
/* the half list of the last build, as nbnxm::List holds it (core/src/nbnxm.hpp:29-48); any pointer may be NULL */
int32_t capsgpu_list_get(capsgpu_ctx* ctx, uint32_t* perm, uint32_t* iperm, int64_t* first, uint32_t* cj, uint16_t* code,
                         uint16_t* mask, uint16_t* ex, uint16_t* ex14, double* org, double* off, int32_t* ioff,
                         int32_t* qint, double* qvec);
/* the slot coordinates of the last evaluation (List::positions) and the pruned masks per outer entry (List::prune) */
int32_t capsgpu_list_pruned(capsgpu_ctx* ctx, float* X, float* Y, float* Z, uint16_t* pmask, uint16_t* pex,
                            uint16_t* pex14);
/* per-stage device times of the last capsgpu_compute, ms: list build, prune, pair kernel, PME, bonded, rest */
int32_t capsgpu_stage_times(capsgpu_ctx* ctx, double out[6]);
/* stage profiling of every later evaluation, dynamics included (events on the context's stream): on = 1 starts it and
 * zeroes the sums, 0 stops it. read: device time summed per stage (ms, the stages of capsgpu_stage_times) and the number
 * of evaluations summed, since profiling started */
int32_t capsgpu_profile(capsgpu_ctx* ctx, int32_t on);
int32_t capsgpu_profile_read(capsgpu_ctx* ctx, double ms[6], int64_t counts[6]);
/* per-kernel launch times since profiling started, as text: one "name summed_ms launches" line per kernel, truncated
 * to cap − 1 characters and NUL-terminated */
int32_t capsgpu_profile_kernels(capsgpu_ctx* ctx, char* buf, int32_t cap);
/* time only the launches of one kernel (by name, e.g. "k_spread"; NULL or "" for every kernel) and no stages, so a
 * profiled run keeps two events per launch of that kernel and nothing else (no counterpart in the reference) */
int32_t capsgpu_profile_only(capsgpu_ctx* ctx, const char* kernel);

/* ---- learned r-RESPA inspection, for the parity checks (no counterpart in the reference) --------------------------- */
/* The network (96-128-128-3, SiLU; BF16 MFMA, FP32 accumulation) at the given FP32 parameters, laid out w1 [128][96],
 * b1 [128], w2 [128][128], b2 [128], w3 [3][128], b3 [3] (29,315 values): y [m][3] for x [m][96]. Replaces the
 * context's network. */
int32_t capsgpu_learned_forward(capsgpu_ctx* ctx, const float* params, const float* x, int32_t m, float* y);
/* the gradient of L = ½ Σ (y − t)² / m with respect to the parameters (same layout) and L; m a multiple of 32 */
int32_t capsgpu_learned_gradient(capsgpu_ctx* ctx, const float* params, const float* x, const float* t, int32_t m, float* grad,
                                 double* loss);
/* the raw descriptors [n][96] of the set positions: 24 hat-function radial bins (0 … 8 Å) × {Σ q_j h, Σ q_j h d̂_x,
 * Σ q_j h d̂_y, Σ q_j h d̂_z} with the cosine cut-off ½(1 + cos πr/8) */
int32_t capsgpu_learned_descriptors(capsgpu_ctx* ctx, float* raw);

#ifdef __cplusplus
}
#endif
#endif
