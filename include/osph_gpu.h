/*
 * osph_gpu.h -- C ABI of the MI355X (gfx950) implementation of OpenSPH's per-timestep hot path.
 *
 * The library keeps the particle state resident in HBM between calls and exposes it at the
 * granularity of the reference's own interfaces:
 *
 *   osph_set_quantity / osph_get_quantity  <- Storage value/dt/d2t buffers
 *                                             (core/quantities/Storage.h:242-575, Quantity.h:58-160)
 *   osph_build_neighbors / osph_get_pairs  <- ISymmetricFinder::build + findLowerRank as used by
 *                                             SymmetricSolver<3>::loop
 *                                             (core/objects/finders/NeighborFinder.h:87-133,
 *                                              core/sph/solvers/SymmetricSolver.cpp:109-141)
 *   osph_integrate                         <- ISolver::integrate, SymmetricSolver<3>::integrate
 *                                             (core/timestepping/ISolver.h:38,
 *                                              core/sph/solvers/SymmetricSolver.cpp:58-95), and
 *                                             AsymmetricSolver::integrate / loop, whose per-particle
 *                                             sums over all neighbours are the kernels' own form
 *                                             (core/sph/solvers/AsymmetricSolver.cpp:70-96,153-202)
 *   osph_step_pc                           <- ITimeStepping::step of PredictorCorrector
 *                                             (core/timestepping/TimeStepping.cpp:34-75,286-346)
 *   osph_timestep                          <- MultiCriterion::compute
 *                                             (core/timestepping/TimeStepCriterion.cpp:389-419)
 *
 * Conventions: every function returns 0 on success and a nonzero code on failure; the message is
 * available from osph_last_error().  No exceptions cross the ABI.  A context owns one HIP stream and
 * all device memory; host buffers are only borrowed for the duration of a call.  A context is not
 * re-entrant and must be used from one thread at a time (as ISolver, core/timestepping/ISolver.h:28-31).
 */
#ifndef OSPH_GPU_H
#define OSPH_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct osph_ctx osph_ctx;

enum {
    OSPH_OK = 0,
    OSPH_ERR_ARG = 1,
    OSPH_ERR_HIP = 2,
    OSPH_ERR_STATE = 3,
    OSPH_ERR_UNSUPPORTED = 4
};

/* Per-material parameters (BodySettings of one body, core/system/Settings.cpp:793-943). */
typedef struct {
    /* Tillotson EoS (TillotsonEos, core/physics/Eos.cpp:184-238) */
    double rho0, A, B, a, b, alpha, beta, u0, uiv, ucv;
    int yielding;          /* 0 = NONE (EosMaterial), 1 = VON_MISES, 2 = DRUCKER_PRAGER (SolidMaterial) */
    int damage;            /* 0 = NONE, 1 = SCALAR_GRADY_KIPP */
    double shear_modulus, elasticity_limit, melt_energy, young_modulus;
    /* IMaterial::range (clamping) and IMaterial::minimal (derivative criterion) */
    double rho_lo, rho_hi, u_lo, u_hi, d_lo, d_hi;
    double rho_min, u_min, s_min, d_min;
    /* DruckerPragerRheology (core/physics/Rheology.cpp:119-181): COHESION, INTERNAL_FRICTION, DRY_FRICTION */
    double cohesion, internal_friction, dry_friction;
} osph_material;

/* Run settings read by the hot path (RunSettings, core/system/Settings.cpp:486-715). */
typedef struct {
    int solid_stress;        /* ForceEnum::SOLID_STRESS: SolidStressForce in the equation set */
    int adaptive_h;          /* SmoothingLengthEnum::CONTINUITY_EQUATION: AdaptiveSmoothingLength */
    int sum_only_undamaged;  /* SPH_SUM_ONLY_UNDAMAGED */
    int has_reduce;          /* the storage holds STRESS_REDUCING (some material has a rheology) */
    int has_damage;          /* the storage holds DAMAGE */
    double av_alpha, av_beta;
    double h_lo, h_hi;       /* SPH_SMOOTHING_LENGTH_RANGE */
    int criteria;            /* bit0 COURANT, bit1 DERIVATIVES, bit2 ACCELERATION, bit3 DIVERGENCE */
    double courant, deriv_factor, div_factor, max_dt;
    int precision;           /* OSPH_PRECISION_FP64 or OSPH_PRECISION_MIXED */
} osph_config;

enum { OSPH_PRECISION_FP64 = 0, OSPH_PRECISION_MIXED = 1 };

/* Quantity ids (a subset of QuantityId, core/quantities/QuantityIds.h) */
enum {
    OSPH_Q_POSITION = 0,          /* Vector: x, y, z, h; orders 0, 1, 2 */
    OSPH_Q_MASS,                  /* double */
    OSPH_Q_DENSITY,               /* double; orders 0, 1 */
    OSPH_Q_ENERGY,                /* double; orders 0, 1 */
    OSPH_Q_DEVIATORIC_STRESS,     /* TracelessTensor; orders 0, 1 */
    OSPH_Q_DAMAGE,                /* double; orders 0, 1 */
    OSPH_Q_PRESSURE,              /* double */
    OSPH_Q_SOUND_SPEED,           /* double */
    OSPH_Q_STRESS_REDUCING,       /* double */
    OSPH_Q_VELOCITY_DIVERGENCE,   /* double */
    OSPH_Q_VELOCITY_GRADIENT,     /* SymmetricTensor */
    OSPH_Q_NEIGHBOR_CNT,          /* uint32 */
    OSPH_Q_FLAG,                  /* uint32 (body index) */
    OSPH_Q_MATERIAL_ID,           /* uint32 (index into the material table) */
    OSPH_Q_EPS_MIN,               /* double */
    OSPH_Q_M_ZERO,                /* double */
    OSPH_Q_EXPLICIT_GROWTH,       /* double */
    OSPH_Q_N_FLAWS,               /* uint32 */
    OSPH_Q_COUNT
};

/* Host element layouts are those of the reference types:
 *   Vector          4 doubles x, y, z, h                  natural stride 32 B (Vector.h:21-26)
 *   TracelessTensor xx, yy, xy, xz at 0..24, yz at 32     natural stride 64 B (TracelessTensor.h:36-45)
 *   SymmetricTensor xx, yy, zz at 0..16, xy, xz, yz at 32..48, natural stride 64 B (SymmetricTensor.h:13-14)
 *   double 8 B, uint32 4 B.
 * A stride of 0 selects the natural stride. */

/* ---- lifetime ------------------------------------------------------------------------------------- */
int osph_create(osph_ctx** ctx, const osph_config* cfg, const osph_material* materials, int n_materials);
void osph_destroy(osph_ctx* ctx);
const char* osph_last_error(const osph_ctx* ctx);
int osph_device_info(char* name, size_t cap, int* compute_units);

/* ---- state (Storage) -------------------------------------------------------------------------------- */
int osph_resize(osph_ctx* ctx, size_t n);
int osph_set_quantity(osph_ctx* ctx, int quantity, int order, const void* host, size_t stride_bytes);
int osph_get_quantity(osph_ctx* ctx, int quantity, int order, void* host, size_t stride_bytes);
/* Storage::zeroHighestDerivatives (core/quantities/Storage.cpp:594-602) */
int osph_zero_derivatives(osph_ctx* ctx);

/* ---- neighbour search (ISymmetricFinder) ----------------------------------------------------------- */
/* Builds the cell grid and the neighbour lists for the current positions. */
int osph_build_neighbors(osph_ctx* ctx);
/* Unordered pairs {i<j} with |r_i - r_j|^2 < (2 hbar)^2 in storage indices, sorted by (i, j).
 * Writes min(count, cap) pairs and the total into *count. */
int osph_get_pairs(osph_ctx* ctx, uint32_t* pi, uint32_t* pj, size_t cap, size_t* count);

/* ---- solver / integrator ---------------------------------------------------------------------------- */
typedef struct {
    uint32_t neighbors_min, neighbors_max;
    double neighbors_mean;   /* StatisticsId::NEIGHBOR_COUNT (SymmetricSolver.cpp:172-179) */
} osph_stats;

int osph_integrate(osph_ctx* ctx, osph_stats* stats);
int osph_step_pc(osph_ctx* ctx, double dt, double* next_dt, int* criterion);
int osph_timestep(osph_ctx* ctx, double* dt, int* criterion);
/* One osph_step_pc of each of `count` independent contexts (an ensemble of runs, each with its own
 * time step dt[k]), interleaved phase by phase so that the contexts' streams run concurrently and the
 * host waits overlap.  Each context's result is bit-identical to osph_step_pc on it alone.  On failure
 * the first failing context's code is returned and its osph_last_error() holds the message.  No
 * reference counterpart: the reference runs ensembles as separate jobs. */
int osph_step_pc_batch(osph_ctx** ctxs, int count, const double* dt, double* next_dt, int* criterion);

/* Replaces ScalarGradyKippModel::setFlaws (core/physics/Damage.cpp:17-126, accumulate mode) for the
 * particles [first, first + count) of one body, on the resident state: flaws p = 1, 2, ... each hit a
 * uniformly drawn particle of the body until every particle holds one (the draw that flaws the last
 * particle ends the sequence); a particle's EPS_MIN / M_ZERO / N_FLAWS follow from its first and last
 * flaw and their count, eps(p) = (p / (k_weibull volume))^(1/m_weibull), and EXPLICIT_GROWTH =
 * growth_speed / (2 h).  volume = the body's sum of m / rho; growth_speed = the Rayleigh fraction of the
 * longitudinal sound speed.  The draws come from a counter-based generator (splitmix64 of seed + p times
 * the golden-ratio increment), not from the reference's BenzAsphaugRng: the same seed gives the same flaws
 * on every device, whatever the particle order in HBM.  The state must hold damage quantities. */
int osph_set_flaws(osph_ctx* ctx, size_t first, size_t count, double volume, double k_weibull, double m_weibull,
                   double growth_speed, uint64_t seed);

/* ---- initial conditions (InitialConditions + HexagonalPacking + CollisionGeometrySetupJob) ------------- */
/* One monolithic spherical body: a MonolithicBody job with a SphericalDomain and HexagonalPacking
 * (BodySettingsId::PARTICLE_COUNT, SMOOTHING_LENGTH_ETA), its material and flag, and the velocity that every
 * particle of the body starts with. */
typedef struct {
    double particle_count;  /* requested count; the lattice inside the sphere decides the actual count */
    double radius;
    double center[3];
    double eta;
    double velocity[3];
    int material;           /* index into the context's material table */
    uint32_t flag;          /* QuantityId::FLAG of the body */
} osph_sphere_body;

/* CollisionGeometrySettingsId::IMPACT_ANGLE (degrees) and IMPACTOR_OFFSET (in target smoothing lengths). */
typedef struct {
    double impact_angle_deg;
    double impactor_offset_h;
} osph_collision;

/* Where a body landed in the storage, and its volume sum(m / rho) (what osph_set_flaws takes). */
typedef struct {
    size_t first, count;
    double volume;
    double h;               /* smoothing length of its particles */
    double mass;            /* mass of each of its particles */
} osph_body_info;

/* Builds the bodies on the device, one after another in the storage, and resizes the context to them:
 *   HexagonalPacking::generate with CENTER in a SphericalDomain (core/sph/initial/Distribution.cpp:126-201,
 *     lattice order of Box::iterateWithIndices, core/objects/geometry/Box.h:182-195);
 *   InitialConditions::addMonolithicBody's quantities (core/sph/initial/Initial.cpp:293-333): h = eta x the
 *     lattice h, masses proportional to h^3 summing to rho0 x the sphere volume, density rho0, zero energy,
 *     unit stress reduction, the body's flag, material and velocity;
 *   with collision set and exactly two bodies (target, then impactor), CollisionGeometrySetupJob::evaluate
 *     (core/run/jobs/ParticleJobs.cpp:532-577): bounding spheres about the centres of mass, the target's
 *     centre moved to the origin, the impactor's to (d cos phi + offset h_target, d sin phi, 0) with d the
 *     sum of the bounding radii (the impactor velocity is the body's own velocity field);
 *   then the density floor of ContinuityEquation::create (core/sph/equations/EquationTerm.cpp:318-336):
 *     each material's lower density bound becomes max(min_i m_i W(0) / h_i^3, its bound).
 * Every position is bit-identical to the sequential construction (the centres of mass are left-to-right
 * sums).  info (optional) receives one osph_body_info per body.  Flaws are drawn afterwards with
 * osph_set_flaws. */
int osph_setup_spheres(osph_ctx* ctx, const osph_sphere_body* bodies, int count, const osph_collision* collision,
                       osph_body_info* info);

/* Replaces GravitySolver<SymmetricSolver<3>> (core/sph/solvers/GravitySolver.cpp:64-91): with a self-gravity
 * context attached (osph_gravity.h), every derivative evaluation of ctx (osph_integrate, osph_step_pc and
 * the batch) first evaluates the self-gravity of the current positions, smoothing lengths and masses
 * with it, and the pair sums then add to those accelerations.  The particles are copied device to device
 * each evaluation (the gravity context's own particles are overwritten); set the gravity context up as
 * Factory::getGravity would (core/system/Factory.cpp:361-413; BruteForceGravity takes Constants::gravity,
 * BarnesHut GRAVITY_CONSTANT).  g = NULL detaches.  The gravity context must outlive the attachment;
 * the adjoint is not available while one is attached.  A gravity context serves one SPH context at a
 * time (its planes hold that context's particles and accelerations between the evaluation and the pair
 * sums): attaching it to a second one fails with OSPH_ERR_ARG until the first detaches or is destroyed. */
struct osph_gravity;
int osph_attach_gravity(osph_ctx* ctx, struct osph_gravity* g);
int osph_sync(osph_ctx* ctx);
/* Integrals of the resident state (TotalEnergy, TotalKineticEnergy, TotalInternalEnergy, TotalMomentum,
 * TotalAngularMomentum about the origin, CenterOfMass, TotalMass; core/physics/Integrals.cpp:9-180):
 * out = { E, E_kin, E_int, p x y z, L x y z, com x y z, M }, FP64 with a fixed reduction order. */
enum { OSPH_INTEGRALS = 13 };
int osph_integrals(osph_ctx* ctx, double out[OSPH_INTEGRALS]);
/* osph_integrals for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, hosts that forbid hipMalloc).  No reference counterpart.  The same kernels and results as
 * osph_integrals (bit for bit), with the per-block partial sums in a static device array of 16384 doubles
 * (128 KiB) that each GPU holds until the process ends, instead of a context buffer allocated on first use.
 * Calls no allocator.  Rejects a state whose partial sums would not fit (none does: the partials are
 * capped at 1025 x 12).  Every context on one GPU shares that GPU's array: calls that overlap on one GPU
 * (from several threads or contexts) race. */
int osph_integrals_noalloc(osph_ctx* ctx, double out[OSPH_INTEGRALS]);

/* ---- reverse mode (no reference counterpart: OpenSPH has no derivatives of its runs) ---------------- */
/* Gradients of a function of the final state of a PredictorCorrector run with the gas terms
 * (PressureForce + StandardAV + ContinuityEquation, Tillotson, constant h; the PC of
 * core/timestepping/TimeStepping.cpp:286-346) with respect to its initial state and the Tillotson
 * parameters a, b, A.  osph_adjoint_forward runs `steps` PC steps with the given time steps, keeping a
 * checkpoint of r, v, a, rho, drho, u, du every `checkpoint_every` steps; the resident state ends at the
 * final state.  osph_adjoint_backward replays each segment from its checkpoint and runs the reverse
 * steps: seeds are dJ/d(final r, v) (n x 3, caller order) and dJ/d(final rho, u) (n); each may be null
 * (zero).  Outputs (each may be null): dJ/d(initial r, v) (n x 3), dJ/d(initial rho, u) (n) and
 * grad_eos[3 * material + q] = dJ/d(a, b, A).  The time steps are held fixed and the neighbour sets are
 * those of the replay (identical to the forward's).  *clamped counts particle-steps whose predicted
 * density or energy sat on its range bound, where the clamped map's derivative is not taken (expect 0).
 * Afterwards the resident state is the initial state again.  While a forward run is recorded the state is
 * not reordered; osph_adjoint_clear drops the checkpoints and restores the reorder setting. */
int osph_adjoint_forward(osph_ctx* ctx, int steps, const double* dt, int checkpoint_every);
int osph_adjoint_backward(osph_ctx* ctx, const double* seed_r, const double* seed_v, const double* seed_rho,
                          const double* seed_u, double* grad_r, double* grad_v, double* grad_rho, double* grad_u,
                          double* grad_eos, unsigned* clamped);
int osph_adjoint_clear(osph_ctx* ctx);
/* Launch-shape tuning of the neighbour-list and pair kernels (no reference counterpart: the reference
 * has no device kernels).  Keys: "pairs_block", "lists_block" (64, 128 or 256 threads) and
 * "pairs_lds", "lists_lds" (0..65536 bytes of dynamic LDS per block, which caps resident blocks per CU),
 * "pairs_lanes" (1, 2 or 4 lanes share one particle's neighbour list in the pair kernels, default 1),
 * "cell_divisor" (1 to 4: neighbour-grid cells of 2 h_max / divisor, default 3), "lists_screen" (1: FP32-screened
 * neighbour test, 0: FP64 only, -1: screened in mixed precision only), "reorder_period" (the resident state
 * is moved into cell order every this many neighbour searches; 0 never; default 40), "fuse_finalize" (1: the
 * PC step's finalize, corrector and criteria run in the pair kernel, the default, 0: as a separate pass,
 * -1: fused for FP64 only). */
int osph_tune(osph_ctx* ctx, const char* key, long long value);
/* Launch-shape autotuner (no reference counterpart): coordinate descent over osph_tune keys.  keys[k] takes
 * the n_values[k] candidates values[sum(n_values[<k]) ...], the first of them the starting value (keys = NULL:
 * pairs_lanes 1/2/4, cell_divisor 3/2/4, lists_block 256/128, pairs_block 256/128, fuse_finalize 1/0/-1).
 * Each candidate is timed over `steps` osph_step_pc calls with time step dt (after one untimed step) on the
 * resident state, which is restored on the device after every candidate: the tuning leaves the run where it
 * was.  One key at a time, in order, the fastest value is kept.  The chosen values are applied and written
 * to chosen[n_keys] (5 for the default space); *best_ms is their time per step; trial_ms[0 ..] (up to
 * trial_cap) the time of every candidate, *n_trials their number.  Every candidate gives the same
 * neighbour sets; the lanes per particle change the order of the pair sums. */
int osph_autotune(osph_ctx* ctx, double dt, int steps, const char* const* keys, const int* n_values,
                  const long long* values, int n_keys, long long* chosen, double* best_ms, double* trial_ms,
                  int trial_cap, int* n_trials);

/* ---- kernel table and profiling --------------------------------------------------------------------- */
/* The 40001-entry M4 look-up tables the device uses (LutKernel<3>, core/sph/kernel/Kernel.h:83-101). */
int osph_kernel_lut(double* values, double* grads);
/* STREAM-style device copy of `bytes` bytes repeated `iters` times; *gbs = read+write GB/s.  Used to
 * re-measure the HBM roof and to calibrate the FETCH_SIZE / WRITE_SIZE counters (no reference
 * counterpart: measurement support). */
int osph_stream_copy(size_t bytes, int iters, double* gbs);
/* enable 1: every kernel launch is bracketed by HIP events on the context's stream and its duration
 * accumulated per kernel name; 2: only the pair-sum launches (the events of every launch add gaps of a
 * few microseconds between kernels); 0: off. */
int osph_profile_enable(osph_ctx* ctx, int enable);
/* Copies up to cap (name, total ms, launches) records; returns the count in *count. */
int osph_profile_read(osph_ctx* ctx, char* names, size_t name_cap, double* ms, uint64_t* launches, int cap,
                      int* count);

#ifdef __cplusplus
}
#endif
#endif
