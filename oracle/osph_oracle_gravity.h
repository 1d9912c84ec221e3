/*
 * TEST INFRASTRUCTURE -- not part of the product.
 *
 * Plain-C, strict-IEEE restatement of OpenSPH's self-gravity (BruteForceGravity, the softened
 * GravityLutKernel of the M4 kernel, the multipole moments and their evaluation of BarnesHut, and the
 * LeapFrog stepper with the gravity-only HardSphereSolver).  It is the checker of the HIP gravity path:
 * only tests/ load it.  Barnes-Hut results depend on the tree, so the checks of the GPU tree code are
 * the direct sum below (at theta -> 0 exactly, otherwise within the reference's own tolerances), the
 * moments of the whole set computed directly, and the multipole evaluation for a distant cluster.
 *
 * Build: gcc -std=c11 -O2 -ffp-contract=off (see Makefile).  Positions are SoA planes x, y, z, h.
 */
#ifndef OSPH_ORACLE_GRAVITY_H
#define OSPH_ORACLE_GRAVITY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_GRAVITY_LUT_ENTRIES 40000

/* LutKernel<3> built from GravityKernel<CubicSpline<3>> (core/sph/kernel/GravityKernel.h:77-114,
 * core/sph/kernel/Kernel.h:83-101): 40001 values and gradients at q^2 = i / 10000. */
void orc_gravity_lut(double* values, double* grads);

/* GravityLutKernel::grad / value (core/sph/kernel/GravityKernel.h:45-70).  grads / values == NULL is
 * the default-constructed kernel of point particles (radius 0). */
void orc_gravity_grad(const double* grads, double rx, double ry, double rz, double h, double out[3]);
double orc_gravity_value(const double* values, double rx, double ry, double rz, double h);

/* BruteForceGravity::evalSelfGravity (core/gravity/BruteForceGravity.h:49-57,112-131) for the targets
 * t[0..nt): a[k] += G sum_{j != t[k]} m_j grad(r_j - r_t, (h_j + h_t) / 2), with the symmetrised kernel.
 * t == NULL means every particle (nt = n).  a is 3 planes of nt values. */
void orc_gravity_direct(size_t n, const double* x, const double* y, const double* z, const double* h,
                        const double* m, const double* grads, double G, const uint32_t* t, size_t nt, double* a);

/* BruteForceGravity::evalAcceleration (BruteForceGravity.h:91-101): unsymmetrised kernel with the
 * source's h, at arbitrary points p (3 planes of np values). */
void orc_gravity_points(size_t n, const double* x, const double* y, const double* z, const double* h,
                        const double* m, const double* grads, double G, const double* p, size_t np, double* a);

/* BruteForceGravity::evalEnergy (BruteForceGravity.h:103-115): 0.5 G sum_{i != j} m_i m_j value. */
double orc_gravity_energy(size_t n, const double* x, const double* y, const double* z, const double* h,
                          const double* m, const double* values, double G);

/* Moments of the whole set about its centre of mass: computeMultipole<0, 2, 3> and
 * computeReducedMultipole (core/gravity/Moments.h:84-130,162-174).  out = { M, com_x, com_y, com_z,
 * q2 (xx, xy, xz, yy, yz), q3 (xxx, xxy, xxz, xyy, xyz, yyy, yyz) } (16 values). */
void orc_gravity_moments(size_t n, const double* x, const double* y, const double* z, const double* m,
                         double out[16]);

/* evaluateGravity (core/gravity/Moments.h:318-360) of moments mom = { M, q2[5], q3[7] } at
 * dr = r - com, up to order 0, 2 or 3 (MONOPOLE, QUADRUPOLE, OCTUPOLE). */
void orc_multipole_accel(const double mom[13], double drx, double dry, double drz, int order, double out[3]);

/* LeapFrog::stepParticles (core/timestepping/TimeStepping.cpp:352-391) with the gravity of
 * HardSphereSolver::integrate (core/gravity/NBodySolver.cpp:97-117) evaluated by brute force; `steps`
 * steps of dt.  r and v are 4 planes each (lane 3 = h and its derivative), a 3 planes (the last
 * accelerations). */
void orc_leapfrog(size_t n, double* r, double* v, double* a, const double* m, const double* grads, double G,
                  double dt, int steps);

#ifdef __cplusplus
}
#endif
#endif
