/*
 * osph_gravity.h -- C ABI of the MI355X (gfx950) self-gravity of OpenSPH.
 *
 * Replaces, at the granularity of the reference's own interface IGravity
 * (core/gravity/IGravity.h:16-64):
 *
 *   osph_gravity_create                     <- Factory::getGravity (core/system/Factory.cpp:361-413):
 *                                              BarnesHut or BruteForceGravity with a GravityLutKernel
 *   osph_gravity_set_precision              (none: FP32 evaluation of the interaction lists)
 *   osph_gravity_tune                       (none: FP32-screened walk)
 *   osph_gravity_set_particles              <- the POSITION / MASS buffers that IGravity::build reads
 *   osph_gravity_build                      <- IGravity::build (BarnesHut.cpp:50-84, BruteForceGravity.h:44-47)
 *   osph_gravity_eval_self                  <- IGravity::evalSelfGravity (BarnesHut.cpp:86-99,212-312;
 *                                              BruteForceGravity.h:49-57)
 *   osph_gravity_eval_points                <- IGravity::evalAcceleration (BarnesHut.cpp:136-178,
 *                                              BruteForceGravity.h:91-101)
 *   osph_gravity_energy                     <- BruteForceGravity::evalEnergy (BruteForceGravity.h:103-115)
 *   osph_gravity_moments                    <- BarnesHut::getMoments (BarnesHut.cpp:489-492)
 *   osph_gravity_step_leapfrog              <- LeapFrog::stepParticles (core/timestepping/TimeStepping.cpp:352-391)
 *                                              with HardSphereSolver::integrate (core/gravity/NBodySolver.cpp:97-117)
 *   osph_gravity_integrals                  <- TotalKineticEnergy, TotalMomentum, TotalAngularMomentum,
 *                                              CenterOfMass (core/physics/Integrals.cpp:24-180)
 *
 * The Barnes-Hut tree is a binary radix tree over 63-bit Morton codes of the positions instead of the
 * reference's sliding-midpoint k-d tree; node moments (centre of mass, traceless quadrupole and
 * octupole), opening radii, the opening test against each leaf's box and the evaluation of the
 * interaction lists follow the reference.  Conventions and status codes as in osph_gpu.h: OSPH_OK on
 * success, the message of a failure from osph_gravity_last_error(), one stream per context, one thread
 * at a time.
 */
#ifndef OSPH_GRAVITY_H
#define OSPH_GRAVITY_H

#include <stddef.h>
#include <stdint.h>

#include "osph_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct osph_gravity osph_gravity;

enum { OSPH_GRAVITY_BRUTE_FORCE = 0, OSPH_GRAVITY_BARNES_HUT = 1 };  /* GravityEnum */
enum { OSPH_GRAVITY_KERNEL_POINT = 0, OSPH_GRAVITY_KERNEL_M4 = 1 };  /* GravityKernelEnum POINT_PARTICLES,
                                                                        SPH_KERNEL with CUBIC_SPLINE */

typedef struct {
    int solver;      /* OSPH_GRAVITY_* (RunSettingsId::GRAVITY_SOLVER) */
    int kernel;      /* OSPH_GRAVITY_KERNEL_* (GRAVITY_KERNEL) */
    double theta;    /* opening angle (GRAVITY_OPENING_ANGLE) */
    int order;       /* 0 monopole, 2 quadrupole, 3 octupole (GRAVITY_MULTIPOLE_ORDER) */
    int leaf_size;   /* 1..64 (FINDER_LEAF_SIZE) */
    double G;        /* GRAVITY_CONSTANT */
} osph_gravity_config;

typedef struct {
    uint64_t nodes_approx;  /* StatisticsId::GRAVITY_NODES_APPROX: nodes evaluated by their moments */
    uint64_t nodes_exact;   /* GRAVITY_NODES_EXACT: leaves evaluated pair by pair */
    uint64_t node_count;    /* GRAVITY_NODE_COUNT */
    uint64_t expansions;    /* particle-node multipole evaluations (no reference counterpart: work count) */
    uint64_t pairs;         /* particle-particle kernel evaluations (likewise) */
} osph_gravity_stats;

int osph_gravity_create(osph_gravity** g, const osph_gravity_config* cfg);
void osph_gravity_destroy(osph_gravity* g);
const char* osph_gravity_last_error(const osph_gravity* g);

/* Arithmetic of the Barnes-Hut evaluation (no reference counterpart: the reference evaluates in FP64).
 * OSPH_PRECISION_FP64 (the default) follows the reference's FP64 sums.  OSPH_PRECISION_MIXED evaluates
 * the interaction lists in FP32 on coordinates relative to each target leaf, with lengths and G m
 * scaled by powers of two (the root box extent and the total G m), and adds each batch of up to 64
 * terms to FP64 accumulators; the tree, its moments, the walk and the integration stay FP64.  The
 * brute-force solver, evaluations at points and the direct sums are FP64 either way. */
int osph_gravity_set_precision(osph_gravity* g, int precision);

/* Launch tuning (no reference counterpart).  Key "walk_screen": 1 takes the Barnes-Hut walk's opening
 * tests in FP32 on 32-byte node records, with the FP64 test wherever the FP32 margin is within its error
 * bound, so the interaction lists are those of the FP64 walk, entry for entry; 0 walks in FP64 only; -1
 * (the default) screens in mixed precision only.  Key "reorder_period": the LeapFrog steps
 * (osph_gravity_step_leapfrog) move the resident state into the tree's Morton order on their first
 * tree build after osph_gravity_set_particles and then every that many builds (default 10; 0 never).
 * Every call still takes and returns particles in the caller's order; sums over all particles
 * (energy, integrals, the direct sum's sources) then run in the resident order. */
int osph_gravity_tune(osph_gravity* g, const char* key, long long value);

/* Particles in the reference's element layouts: r and v are Vectors (x, y, z, h; stride bytes, 0 = 32),
 * m doubles.  v may be NULL (zero velocities).  Accelerations start at zero. */
int osph_gravity_set_particles(osph_gravity* g, size_t n, const double* r, const double* v, size_t vec_stride,
                               const double* m);
/* what: 0 positions, 1 velocities, 2 accelerations (Vectors, h lane of accelerations = 0) */
int osph_gravity_get(osph_gravity* g, int what, double* out, size_t vec_stride);

int osph_gravity_build(osph_gravity* g);
/* dv += acceleration of every particle (dv: n Vectors, read and written; NULL = the context's own
 * accelerations, left on the device). */
int osph_gravity_eval_self(osph_gravity* g, double* dv, size_t vec_stride, osph_gravity_stats* stats);
/* Acceleration at np arbitrary points p (Vectors, h lane unused) into a (Vectors). */
int osph_gravity_eval_points(osph_gravity* g, const double* p, size_t np, double* a, size_t vec_stride);
/* Direct sum for the particles t[0..nt) only (the brute-force check of a sample of targets); a: nt Vectors. */
int osph_gravity_direct(osph_gravity* g, const uint32_t* t, size_t nt, double* a, size_t vec_stride);
/* Potential energy by direct summation, O(n^2). */
int osph_gravity_energy(osph_gravity* g, double* energy);
/* Root moments divided by G: { M, com x y z, q2 xx xy xz yy yz, q3 xxx xxy xxz xyy xyz yyy yyz }. */
int osph_gravity_moments(osph_gravity* g, double out[16]);

/* `steps` LeapFrog steps of dt with the state resident: drift, rebuild, gravity, kick, drift. */
int osph_gravity_step_leapfrog(osph_gravity* g, double dt, int steps);
/* { kinetic energy, momentum x y z, angular momentum x y z, centre of mass x y z, total mass } */
int osph_gravity_integrals(osph_gravity* g, double out[11]);
int osph_gravity_sync(osph_gravity* g);
/* Per-kernel durations as for osph_profile_enable / osph_profile_read. */
int osph_gravity_profile_enable(osph_gravity* g, int enable);
int osph_gravity_profile_read(osph_gravity* g, char* names, size_t name_cap, double* ms, uint64_t* launches,
                              int cap, int* count);

#ifdef __cplusplus
}
#endif
#endif
