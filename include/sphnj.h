/*
 * sphnj.h -- C ABI of libsphnj, the MI355X (gfx950) implementation of SPHNJ's per-step
 * particle pipeline: cell hashing + key sort + cell ranges, wall (dummy) particles, the
 * Riemann fluid pair loop with the gas-liquid interface terms, particle shifting, and the
 * total-Lagrangian solid pass with sub-stepping.
 *
 * Plain C: no C++ types, no exceptions and no torch types cross this boundary.  Every
 * function returns 0 on success or a negative SPHNJ_E* code; sphnj_last_error() gives the
 * message of the last failure on a context (sphnj_global_error() for calls without one).
 *
 * Reference interfaces replaced (file:line into niji127/SPHNJ):
 *   - the operator API is the static class DeviceCalculation (include/calculation.cuh:9-23),
 *     entered through DeviceFunction::GPUCalulation (src/calculation.cu:14-18) from main
 *     (src/main.cu:14-35); its global state lives in Sim::parameter / Sim::data_pointer
 *     (include/sim.cuh:24-25) and __constant__ dev_par / dev_data (src/memory_copy.cu:5-6).
 *     Here that state is an opaque context, so several simulations (and batches of cases)
 *     can live in one process.
 *
 * Conventions
 *   - Ownership: host arrays passed in are copied during the call; the library owns all
 *     device memory of a context until sphnj_destroy.
 *   - Threading: one context per host thread; calls on one context are serialised by the
 *     caller.  Each context owns one HIP stream and its captured step graphs.
 *   - Particle identity: downloads are in the caller's input order (the library carries the
 *     original index of every particle through every sort).
 *   - Synchronisation: sphnj_step returns after enqueueing; sphnj_download, sphnj_status,
 *     sphnj_sync and the stage-inspection calls synchronise the context's stream.
 */
#ifndef SPHNJ_H
#define SPHNJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHNJ_ABI_VERSION 2

/* ---- error codes -------------------------------------------------------------------- */
enum {
  SPHNJ_OK = 0,
  SPHNJ_E_ARG = -1,        /* bad argument (null pointer, size out of range, ...)        */
  SPHNJ_E_PARAM = -2,      /* parameter validation failed (reference: "FAILED TO READ")  */
  SPHNJ_E_IO = -3,         /* file could not be opened / parsed                          */
  SPHNJ_E_COUNT = -4,      /* .dat header count != number of particle lines              */
  SPHNJ_E_HIP = -5,        /* a HIP runtime call failed                                  */
  SPHNJ_E_NODEVICE = -6,   /* no GPU visible                                             */
  SPHNJ_E_STATE = -7,      /* call out of order (e.g. step before init)                  */
  SPHNJ_E_NOMEM = -8,      /* allocation failed                                          */
  SPHNJ_E_UNSUPPORTED = -9 /* feature combination not supported                          */
};

/* ---- precision policies ------------------------------------------------------------- */
enum {
  SPHNJ_PREC_P64 = 0, /* plain FP64 everywhere                                                  */
  SPHNJ_PREC_M32 = 1  /* FP32 pair sums / rates / stress on cell-relative FP32 coordinates;
                         FP64 absolute position, velocity, density and time integration      */
};

/* ---- enum integers used in .dat files and downloads (include/types.h:1-60) ---------- */
enum { SPHNJ_LIQUID = 0, SPHNJ_GAS = 1 };                                      /* PhaseType          */
enum { SPHNJ_SLIP = 0, SPHNJ_NO_SLIP = 1, SPHNJ_MOVING = 2 };                  /* BoundaryType (wall)*/
enum { SPHNJ_SPLASH = 0, SPHNJ_FREE_SURFACE = 1, SPHNJ_VICINITY = 2, SPHNJ_INNER = 3 }; /* FluidBoundaryType */
enum { SPHNJ_NO_PARTICLE = 0, SPHNJ_HAS_PARTICLE = 1, SPHNJ_HAS_NEAR_PARTICLE = 2 };   /* NearParticle      */
enum { SPHNJ_LINEAR_ELASTIC = 0, SPHNJ_AUXETIC = 1,
       SPHNJ_NEO_HOOKEAN = 2 /* extension: no counterpart in the reference */ };        /* MaterialType      */
enum { SPHNJ_UNCONSTRAINED = 0, SPHNJ_FIXED = 1 };                             /* SolidBoundaryType  */

/* ---- status word bits (sphnj_status) ------------------------------------------------ */
enum {
  SPHNJ_ST_OUT_OF_GRID = 1u << 0,   /* a particle hashed outside the cell box; clamped to the
                                       edge cell (the reference indexes out of range here)   */
  SPHNJ_ST_CELL_OVER_64 = 1u << 1,  /* some cell holds > 64 particles: the reference's fixed
                                       CELL_SIZE list would have overflowed (include/global.h:9) */
  SPHNJ_ST_NAN = 1u << 2,           /* a NaN appeared in a fluid or solid state field         */
  SPHNJ_ST_REFLIST_LEFT = 1u << 3,  /* a solid reference-configuration neighbour left the 3x3
                                       current cells: the reference (src/solid.cu:184-199)
                                       would have dropped that pair                           */
  SPHNJ_ST_SYNC_TIMEOUT = 1u << 4   /* a grid barrier of the one-launch solid sub-steps gave up
                                       waiting (the device did not hold every block at once,
                                       e.g. under heavy concurrent use): that step's solid
                                       state is invalid; sphnj_set_persistent_solid(ctx, 0)
                                       selects the 2 S launch form                           */
};

/* ---- parameters: POD mirror of Parameter (include/parameter.h:17-49) ------------------
 * Field names are the INI keys (defaults in sphnj_params_default, matching
 * src/{physics,properties,domain,sim_time,kernel}.cpp readParameter).  Values are held in
 * double precision.  Booleans are ints (0/1). */
typedef struct sphnj_params {
  /* [physics]  src/physics.cpp:9-25 */
  double gravity;              /* 9.81; g = (0, -gravity)                                  */
  int32_t is_steady;           /* 0                                                        */
  /* [fluid]  src/properties.cpp:35-51 */
  double particle_shift_coef;  /* 0.4    */
  int32_t is_viscid;           /* 0      */
  double min_pressure;         /* -1.013e5 */
  double liquid_sound_speed, liquid_density, liquid_gamma, liquid_viscosity; /* 1500 1000 7 1.01e-3 */
  double gas_sound_speed, gas_density, gas_gamma, gas_viscosity;             /* 340 1.29 1.4 17.9e-6 */
  /* [solid]  src/properties.cpp:118-161 */
  double solid_density, youngs_modulus, poisson_ratio;                       /* 2700 67.5e9 0.34 */
  double arti_vis_alpha, arti_vis_beta;                                      /* 2.5 2.5 */
  double solid_fixed_speed_x, solid_fixed_speed_y;                           /* 0 0 */
  int32_t is_2D_strain, is_plastic;                                          /* 0 0 */
  double auxetic_density, auxetic_modulus, auxetic_poisson_ratio;            /* 636 0 -0.2 */
  double yield_strain, yield_stress, harden_strain, harden_coef;             /* 0 2e7 0.5 100 */
  /* [domain]  src/domain.cpp:8-32 (required, non-degenerate) */
  double cell_xmin, cell_ymin, cell_xmax, cell_ymax;
  double domain_xmin, domain_ymin, domain_xmax, domain_ymax;
  /* [time]  src/sim_time.cpp:8-25 */
  double dt;                   /* required */
  int32_t istart, iend;        /* istart != iend required */
  int32_t solid_sub_step, result_interval, sort_every_step, shift_per_step; /* 1 1 1 1 */
  /* [kernel]  src/kernel.cpp:9-23 */
  double smoothing_length, particle_diameter;  /* required */
  double impact_length_hsml_ratio;             /* 3 */
  int32_t thread_num;                          /* 32; read for compatibility, block sizes are
                                                  chosen by the library's autotuner */
  /* extensions (no counterpart in the reference; defaults reproduce it exactly) */
  double background_pressure;  /* 0: added to both phases' Tait EOS                          */
  double kernel_2d_coef;       /* 0 = the reference constant COEF_3D_TO_2D = 2.120729        */
  double adaptive_cfl;         /* 0 = every step takes dt (the reference).  C > 0: each step
                                  takes dt_n = min(dt, C h / (c_max + U), S C h / c_s) with U
                                  the largest fluid speed after the previous step's update (the
                                  input speeds for the first step), c_max the larger phase sound
                                  speed, the last term only with solids                        */
} sphnj_params;

/* ---- host-side particle input (caller-owned, copied during sphnj_upload) -------------- */
typedef struct sphnj_fluid_in {   /* fluid_particle.dat rows: x y u v rho p phase */
  int32_t n;
  const double *x, *y, *u, *v, *rho, *p;
  const int32_t *phase;
} sphnj_fluid_in;

typedef struct sphnj_solid_in {   /* solid_particle.dat rows: x y u v X Y material [bc] */
  int32_t n;
  const double *x, *y, *u, *v, *X, *Y;
  const int32_t *material, *bc;   /* bc may be NULL (= UNCONSTRAINED, as a 7-column row) */
} sphnj_solid_in;

typedef struct sphnj_wall_in {    /* virt_particle.dat rows: x y bc */
  int32_t n;
  const double *x, *y;
  const int32_t *bc;
} sphnj_wall_in;

/* ---- host-side output buffers (any pointer may be NULL = skip), input order ---------- */
typedef struct sphnj_fluid_out {
  double *x, *y, *u, *v, *rho, *p;
  int32_t *phase, *boundary_type, *near;   /* FluidBoundaryType / NearParticle */
  double *color;                            /* colour sum C_i of the last pair pass       */
  double *du_dx;                            /* velocity-gradient L_i, 4 per particle, row-major */
} sphnj_fluid_out;

typedef struct sphnj_solid_out {
  double *x, *y, *u, *v, *X, *Y;
  double *pressure, *normal;                /* normal: 2 per particle                     */
  double *deformation, *correction, *stress; /* F, L, P: 4 per particle, row-major       */
  double *max_strain;                       /* 4 per particle                             */
  double *du_dt, *couple_dudt;              /* 2 per particle                             */
} sphnj_solid_out;

typedef struct sphnj_wall_out {
  double *x, *y, *u, *v, *p, *rho, *normal; /* normal: 2 per particle */
  int32_t *phase, *bc;
} sphnj_wall_out;

typedef struct sphnj_ctx sphnj_ctx;

/* ---- library info --------------------------------------------------------------------- */
int sphnj_abi_version(void);
const char *sphnj_global_error(void);

/* ---- parameters: INI surface ------------------------------------------------------------
 * sphnj_params_default: every key at its reference default (Appendix of readParameter).
 * sphnj_params_from_ini: Config semantics of include/config.h + src/config.cpp:59-129 ('#'
 *   comments, first '=' splits, repeated key overwrites, bool "false/F/no/n/0/none" = false,
 *   value continuation lines); keys not present keep their defaults.
 * sphnj_params_validate: the checks of Parameter::readFieldParameter (src/parameter.cpp:36-52):
 *   smoothing_length / particle_diameter non-zero, dt non-zero, istart != iend, cell and
 *   domain boxes non-degenerate, not both auxetic_modulus and yield_strain zero. */
int sphnj_params_default(sphnj_params *out);
int sphnj_params_from_ini(const char *path, sphnj_params *out);
int sphnj_params_validate(const sphnj_params *p);
/* Derived grid of a parameter set (Domain::initiate(Kernel), src/domain.cpp:37-48). */
int sphnj_params_grid(const sphnj_params *p, int32_t *nx, int32_t *ny, double *cell_dx, double *cell_dy);

/* ---- .dat particle files (Sim::read*Data src/sim.cu:62-161, host_data.cuh:24-115) ------
 * A host "particle set" owns arrays read from or written to data/{fluid,solid,virt}_particle.dat. */
typedef struct sphnj_particles sphnj_particles;
int sphnj_particles_read(const char *data_dir, sphnj_particles **out);
int sphnj_particles_write(const sphnj_particles *ps, const char *data_dir);
int sphnj_particles_views(const sphnj_particles *ps, sphnj_fluid_in *f, sphnj_solid_in *s,
                          sphnj_wall_in *w, int32_t *n_liquid);
void sphnj_particles_destroy(sphnj_particles *ps);

/* ---- case generators (the cases/<name>/particle_generate.cpp surface) -------------------
 * name: "dam_break" (liquid column, optional air phase), "gate_fsi" (dam break against an
 *       elastic gate), "wedge_entry" (elastic wedge into water), "ensemble_member" (one
 *       seeded member of a dam-break ensemble).  Options are "key=value" pairs separated by
 *       ';' (e.g. "dx=0.02;air=0"); the keys of each case are listed at the top of
 *       sphnj_amd/csrc/host_cases.cpp.
 * sphnj_case_write writes <dir>/config/input.ini and the three .dat files under <dir>/data. */
int sphnj_case_create(const char *name, const char *options, sphnj_params *params,
                      sphnj_particles **out);
int sphnj_case_write(const sphnj_params *params, const sphnj_particles *ps, const char *dir);

/* ---- context life cycle -----------------------------------------------------------------
 * sphnj_create: one case (the drop-in).  sphnj_batch_create: ncase independent cases that
 * share one set of launches per phase; each case keeps its own parameters, dt, cell grid,
 * shift_per_step and solid_sub_step.  Device memory is allocated by sphnj_init, after every
 * case is uploaded. */
int sphnj_create(const sphnj_params *params, int precision, sphnj_ctx **out);
int sphnj_batch_create(const sphnj_params *per_case, int32_t ncase, int precision, sphnj_ctx **out);
int sphnj_upload(sphnj_ctx *ctx, int32_t case_index, const sphnj_fluid_in *f,
                 const sphnj_solid_in *s, const sphnj_wall_in *w);
/* Before-loop work of src/calculation.cu:28 plus the library's set-up: wall sort and wall
 * normals (virtNormalCalculation), solid reference-configuration neighbour lists. */
int sphnj_init(sphnj_ctx *ctx);
/* nsteps iterations of the loop body src/calculation.cu:34-41 (output excluded). */
int sphnj_step(sphnj_ctx *ctx, int32_t nsteps);
int sphnj_sync(sphnj_ctx *ctx);
int sphnj_download(sphnj_ctx *ctx, int32_t case_index, const sphnj_fluid_out *f,
                   const sphnj_solid_out *s, const sphnj_wall_out *w);
int sphnj_status(sphnj_ctx *ctx, uint32_t *flags);
int sphnj_counts(const sphnj_ctx *ctx, int32_t case_index, int32_t *n_fluid, int32_t *n_solid,
                 int32_t *n_wall, int32_t *n_cells);
const char *sphnj_last_error(const sphnj_ctx *ctx);
void sphnj_destroy(sphnj_ctx *ctx);

/* ---- execution options ---------------------------------------------------------------- */
/* Capture the step into hipGraphs (default on) and replay them. */
int sphnj_set_graph(sphnj_ctx *ctx, int enable);
/* Run a fluid step's S solid sub-steps (solidCalculation, src/solid.cu:473-492) as one launch
 * whose blocks meet at grid barriers between the deformation and force phases, or as 2 S
 * launches (default).  Both give the same results bit for bit.  On MI355X a grid barrier must
 * write back and invalidate the L2 of every XCD, which costs more than a launch boundary. */
int sphnj_set_persistent_solid(sphnj_ctx *ctx, int enable);
/* Shift passes of a step (fluidCorrection's loop, src/fluid.cu:442-459; no counterpart of its
 * own in the reference): mode 1 (default) records, per particle, the runs of candidates within
 * rc (1 + 0.02) -- in the fluid stage's pair kernel when the pair and the shift use the same
 * block size (their tiles are then the same), else in the step's first shift pass -- and the
 * shift passes walk only those runs while every path since the recording stays within 0.02 rc
 * (else the full windows); mode 0 searches the full pruned windows in every pass; mode 2 has the
 * first shift pass record the runs and never walks them.  All three give the same results bit
 * for bit.  SPHNJ_E_ARG for another mode. */
int sphnj_set_shift_runs(sphnj_ctx *ctx, int mode);
/* Autotuner: times the fluid pair kernel (family 0) and a step's shift passes (family 1) at
 * block sizes 64, 128 and 256 -- a block owns that many sorted particles of a cell column, so
 * the block size also sets the cells each LDS tile spans -- on the current state with HIP
 * events, and keeps the fastest.  The particle state is unchanged afterwards; the per-step
 * transients (colour, near flag, velocity gradient, position views) are recomputed from it. */
int sphnj_autotune(sphnj_ctx *ctx);
/* Milliseconds per launch (per pass for the shift) measured by the last sphnj_autotune for
 * block sizes 64, 128, 256. */
int sphnj_autotune_report(const sphnj_ctx *ctx, int32_t family, double ms_per_launch[3]);
/* Block size in use for a kernel family (0 = pair, 1 = shift, 2 = streaming). */
int sphnj_get_block(const sphnj_ctx *ctx, int32_t family, int32_t *block);
int sphnj_set_block(sphnj_ctx *ctx, int32_t family, int32_t block);

/* ---- stage entry points (one phase of the step, for parity tests and the host mirror) --
 * SPHNJ_STAGE_CELLS  = particleCellUpdate + sortParticle + fillCell (src/calculation.cu:34-36)
 * SPHNJ_STAGE_COUPLE = coupleCalculation  (src/solid.cu:494-507)
 * SPHNJ_STAGE_WALL   = virtUpdate         (src/virt.cu:151-160)
 * SPHNJ_STAGE_FLUID  = fluidCalculation   (src/fluid.cu:289-302)
 * SPHNJ_STAGE_SOLID  = solidCalculation   (src/solid.cu:473-492)
 * SPHNJ_STAGE_SHIFT  = fluidCorrection    (src/fluid.cu:442-459)                            */
enum {
  SPHNJ_STAGE_CELLS = 1, SPHNJ_STAGE_COUPLE = 2, SPHNJ_STAGE_WALL = 3,
  SPHNJ_STAGE_FLUID = 4, SPHNJ_STAGE_SOLID = 5, SPHNJ_STAGE_SHIFT = 6
};
int sphnj_stage(sphnj_ctx *ctx, int32_t stage);

/* Species selector for the inspection calls. */
enum { SPHNJ_FLUID = 0, SPHNJ_SOLID = 1, SPHNJ_WALL = 2 };
/* Result of the last CELLS stage for one species of one case: sorted keys and the original
 * index of each sorted slot (n entries each), and a start table local to the case
 * (start[b]..start[b+1] is bin b's slot range).  Solids and walls are binned by cell: keys
 * are global cell keys (cy + cx ny + the case's cell offset), n_cells + 1 starts.  Fluids are
 * binned by half-cell sub-cell (sx = 2 cx + hx, sy = 2 cy + hy, hx and hy the halves of the
 * cell): keys are sy + sx 2 ny + 4 x the case's cell offset, 4 n_cells + 1 starts; a cell's
 * fluids are its sub-cells (2cx, 2cy), (2cx, 2cy+1), (2cx+1, 2cy), (2cx+1, 2cy+1). */
int sphnj_stage_sorted(sphnj_ctx *ctx, int32_t case_index, int32_t species, int32_t *keys,
                       int32_t *orig, int32_t *cell_start);
/* In-range neighbours of every fluid particle of one case, in the library's iteration
 * order (cells 0..8 of src/sim.cu:175-193, then fluid, solid, wall, then slot order),
 * as CSR over original indices: counts[n_fluid], then up to cap entries of
 * (species << 28 | original index).  Returns the total number of entries (or a negative
 * code; entries beyond cap are not written). */
int sphnj_stage_neighbors(sphnj_ctx *ctx, int32_t case_index, int32_t *counts, int32_t *list,
                          int64_t cap, int64_t *total);
/* Solid reference-configuration neighbour lists (built by sphnj_init) as CSR over original
 * indices: offsets[n_solid + 1], list[offsets[n_solid]]. */
int sphnj_solid_reflist(sphnj_ctx *ctx, int32_t case_index, int32_t *offsets, int32_t *list,
                        int64_t cap);

/* ---- FP64 global reductions (deterministic order) --------------------------------------
 * out[0] fluid kinetic energy sum(0.5 m |u|^2) with m = rho0[phase] d^2,
 * out[1] fluid mass sum(rho d^2), out[2] max |u| over fluid, out[3] solid kinetic energy,
 * out[4] solid strain energy (src/output.cu:426-455 form), out[5] number of fluid particles. */
int sphnj_energy(sphnj_ctx *ctx, int32_t case_index, double out[6]);

/* Simulated time of one case (the sum of the step sizes taken since sphnj_init) and the step
 * size of its last step (SimTime::current_time, src/sim_time.cpp:27-55; with adaptive_cfl the
 * size varies from step to step).  Synchronises the context. */
int sphnj_time(sphnj_ctx *ctx, int32_t case_index, double *time, double *last_dt);

/* ---- snapshots: binary restart files of one case ----------------------------------------
 * sphnj_snapshot_write: the case's state in input order -- what sphnj_upload takes plus what a
 *   step carries into the next (fluid boundary class, solid max_strain, surface pressure and
 *   normal) -- with its parameters and the caller's step count, to path (SoA, doubles and
 *   int32, a 64-bit hash at the end).  Synchronises the context.
 * sphnj_snapshot_info: the parameters, step count and (fluid, solid, wall) counts of a snapshot.
 * sphnj_snapshot_load: in place of sphnj_upload for case_index (before sphnj_init), from a
 *   snapshot whose parameters equal the context's for that case (SPHNJ_E_PARAM otherwise;
 *   SPHNJ_E_IO for a damaged file).  Stepping the restored context gives the same results,
 *   bit for bit, as stepping the context the snapshot was written from.
 * The reference has no restart files (its output is VTK only, src/output.cu:7-281). */
int sphnj_snapshot_write(sphnj_ctx *ctx, int32_t case_index, int64_t step, const char *path);
int sphnj_snapshot_info(const char *path, sphnj_params *params, int64_t *step, int32_t counts[3]);
int sphnj_snapshot_load(sphnj_ctx *ctx, int32_t case_index, const char *path);

/* ---- probes: the reference's global outputs (src/output.cu:283-582), on the device ------
 * Computed from the current state of one case by kernels enqueued on the context stream
 * after the steps already enqueued, outside the step graph; the call synchronises.  The
 * reference's probe points are hard-coded to its case geometries; here they are arguments. */
enum {
  SPHNJ_PROBE_FORCE_X = 0,        /* outputTotalForce (:283-305): sum couple_dudt rho_ref d^2, x  */
  SPHNJ_PROBE_FORCE_Y = 1,        /*   ... y                                                      */
  SPHNJ_PROBE_FRONT = 2,          /* outputEdge (:339-369): largest liquid x (at least 0) + d/2  */
  SPHNJ_PROBE_HEIGHT = 3,         /* outputEdge (:339-369): largest y of a liquid particle within
                                     d/2 of x = height_x, + d/2 (0 when there is none)          */
  SPHNJ_PROBE_VM_ELASTIC = 4,     /* outputVonMisesStress (:503-544): largest von Mises stress of
                                     F^-1 P over the elastic (non-auxetic) solids               */
  SPHNJ_PROBE_VM_AUXETIC = 5,     /*   ... over the auxetic solids                              */
  SPHNJ_PROBE_VM_MAX = 6,         /*   ... the larger of the two                                */
  SPHNJ_PROBE_MEAN_PRESSURE = 7,  /* outputMeanPressure (:546-560): mean solid surface pressure */
  SPHNJ_PROBE_COUNT = 8
};
int sphnj_probe_global(sphnj_ctx *ctx, int32_t case_index, double height_x, double out[SPHNJ_PROBE_COUNT]);
/* Point probe of outputTankBeamEnd (:457-478), outputEndPosition (:399-424),
 * outputMiddlePosition (:371-397) and outputWedgeDisp (:307-337): the current position of the
 * solid particle whose reference position lies within d/2 of (X, Y) (the lowest input index
 * among several); *index receives that index.  SPHNJ_E_ARG (index -1) when there is none. */
int sphnj_probe_solid(sphnj_ctx *ctx, int32_t case_index, double X, double Y, double pos[2], int32_t *index);

/* sphnj_energy without allocating.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture setup, tight serving loops,
 * harnesses that forbid allocation).  Same reductions and results as sphnj_energy, with its
 * constants and partial sums in a fixed per-GPU static array of 1,288 doubles (10 KB, held by
 * each GPU the library runs on until the process ends) instead of memory allocated per call;
 * the array's size does not depend on the case, and a reduction that needed more would be
 * refused with SPHNJ_E_UNSUPPORTED.  Every context on one GPU shares that array, so calls that
 * overlap on one GPU race: serialise them.  Synchronises the context's stream. */
int sphnj_energy_noalloc(sphnj_ctx *ctx, int32_t case_index, double out[6]);

/* ---- timing (HIP events on the context stream) ------------------------------------------ */
/* Elapsed time of the last sphnj_step call. */
int sphnj_last_step_ms(sphnj_ctx *ctx, double *ms);
/* While enabled, steps run without graphs and every stage is bracketed by events; resets
 * the accumulated records. */
int sphnj_set_stage_timing(sphnj_ctx *ctx, int enable);
/* Accumulated time of one stage since timing was enabled, and the number of kernel launches
 * of that stage's main kernel (K shift passes, 2 per solid sub-step, else 1 per step). */
int sphnj_stage_time(sphnj_ctx *ctx, int32_t stage, int32_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif /* SPHNJ_H */
