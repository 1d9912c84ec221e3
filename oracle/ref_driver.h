/*
 * TEST INFRASTRUCTURE -- not part of the product.
 *
 * extern "C" driver over the real OpenSPH core (compiled from its own sources into
 * oracle/_ref/libosph_ref.so by oracle/Makefile).  It builds a Storage from the oracle's plane layout
 * (osph_oracle.h), runs the reference SymmetricSolver<3> (or AsymmetricSolver) / PredictorCorrector /
 * MultiCriterion on it,
 * and reads the state back.  Used to pin the oracle and as the CPU baseline ("reference" kind).
 *
 * Particles must be grouped by body: each maximal run of equal FLAG values becomes one body with its
 * own material instance (as InitialConditions::addMonolithicBody creates them).
 */
#ifndef OSPH_REF_DRIVER_H
#define OSPH_REF_DRIVER_H

#include "osph_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* nthreads == 0: SequentialScheduler (bit-stable, used for parity); > 0: ThreadPool of that size. */
void* ref_create(const orc_config* cfg, const double* planes, const uint32_t* iplanes, size_t n, double dt0,
                 int nthreads);
/* The same with the solver chosen: asymmetric == 0 SymmetricSolver<3>, 1 AsymmetricSolver (each particle
 * sums over all its neighbours, AsymmetricSolver.cpp:153-202), both under GravitySolver with cfg->gravity. */
void* ref_create_solver(const orc_config* cfg, const double* planes, const uint32_t* iplanes, size_t n, double dt0,
                        int nthreads, int asymmetric);
void ref_destroy(void* h);

/* Storage::zeroHighestDerivatives + SymmetricSolver<3>::integrate on the current state. */
void ref_integrate(void* h);

/* One ITimeStepping::step of PredictorCorrector; returns the next time step, *crit = CriterionId. */
double ref_step(void* h, int* crit);

double ref_get_dt(void* h);

/* Copies the reference Storage back into the plane layout. */
void ref_read(void* h, double* planes, uint32_t* iplanes, size_t n);

/* The run's Storage, RunSettings and IScheduler (as Sph::SharedPtr<Sph::Storage>*, Sph::RunSettings*,
 * Sph::IScheduler*), for drivers that step the same reference-built storage with other solvers
 * (oracle/adapter_driver.cpp). */
void* ref_storage(void* h);
void* ref_settings(void* h);
void* ref_scheduler(void* h);

/* KdTree<KdNode> (leaf 25) findLowerRank(i, 2 h_i) filtered by the SymmetricSolver cutoff; pairs are
 * returned as (min, max) sorted.  Returns the count (writes at most cap). */
size_t ref_pairs(void* h, uint32_t* pi, uint32_t* pj, size_t cap);

/* The reference's self-gravity on n particles (r: n Vectors x, y, z, h; m: masses).  solver 0 =
 * BruteForceGravity, 1 = BarnesHut(theta, order, kernel, leaf); kernel 0 = point particles, 1 =
 * GravityKernel<CubicSpline<3>>.  a (n x 3) receives evalSelfGravity from zero accelerations; when
 * moments != NULL (Barnes-Hut) it receives getMoments() as { M, q2 (xx xy xz yy yz), q3 (xxx xxy xxz xyy
 * xyz yyy yyz) }.  Sequential scheduler. */
void ref_gravity(int solver, int kernel, double theta, int order, int leaf, double G, size_t n, const double* r,
                 const double* m, double* a, double* moments);

/* IGravity::evalAcceleration of the reference's solver at np points p (n x 3 out). */
void ref_gravity_points(int solver, int kernel, double theta, int order, int leaf, double G, size_t n,
                        const double* r, const double* m, const double* p, size_t np, double* a);

#ifdef __cplusplus
}
#endif
#endif
