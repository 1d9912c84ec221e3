/*
 * TEST INFRASTRUCTURE -- not part of the product.
 *
 * Plain-C, strict-IEEE restatement of OpenSPH's per-timestep hot path (SymmetricSolver<3> with its
 * EquationTerm set, Tillotson / von Mises / scalar Grady-Kipp material physics, PredictorCorrector and
 * the time-step criteria).  It is the checker for the HIP path: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  Every function cites the reference file:line it follows
 * (paths relative to the OpenSPH source root).
 *
 * Build: gcc -std=c11 -O2 -ffp-contract=off (no -ffast-math, no -march=native), see Makefile.
 *
 * State layout: plane-major SoA.  `planes` holds ORC_NP double planes of n values each
 * (planes[k*n + i]), `iplanes` holds ORC_NI uint32 planes.  Particle order is the caller's
 * ("storage order"); nothing is reordered.
 */
#ifndef OSPH_ORACLE_H
#define OSPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* double planes */
enum {
    ORC_X, ORC_Y, ORC_Z, ORC_H,             /* POSITION value (lane 3 = smoothing length) */
    ORC_VX, ORC_VY, ORC_VZ, ORC_VH,         /* POSITION dt */
    ORC_AX, ORC_AY, ORC_AZ,                 /* POSITION d2t (lane 3 is always 0 after finalize) */
    ORC_RHO, ORC_DRHO, ORC_U, ORC_DU,
    ORC_SXX, ORC_SYY, ORC_SXY, ORC_SXZ, ORC_SYZ,           /* DEVIATORIC_STRESS (traceless) */
    ORC_DSXX, ORC_DSYY, ORC_DSXY, ORC_DSXZ, ORC_DSYZ,
    ORC_D, ORC_DD,                          /* DAMAGE and its dt */
    ORC_M, ORC_P, ORC_CS, ORC_RED, ORC_DIVV,
    ORC_GXX, ORC_GYY, ORC_GZZ, ORC_GXY, ORC_GXZ, ORC_GYZ, /* VELOCITY_GRADIENT (symmetric) */
    ORC_EPSMIN, ORC_MZERO, ORC_GROWTH,      /* Grady-Kipp flaw data */
    /* highest derivatives saved by the predictor (PredictorCorrector::predictions) */
    ORC_PAX, ORC_PAY, ORC_PAZ, ORC_PDRHO, ORC_PDU,
    ORC_PDSXX, ORC_PDSYY, ORC_PDSXY, ORC_PDSXZ, ORC_PDSYZ, ORC_PDD,
    ORC_NP
};

/* uint32 planes */
enum { ORC_FLAG, ORC_MAT, ORC_NCNT, ORC_NFLAWS, ORC_NI };

#define ORC_LUT_ENTRIES 40000

typedef struct {
    /* Tillotson EoS (core/physics/Eos.cpp:184-196) */
    double rho0, A, B, a, b, alpha, beta, u0, uiv, ucv;
    int yielding;          /* 0 = NONE (EosMaterial), 1 = VON_MISES, 2 = DRUCKER_PRAGER (SolidMaterial) */
    int damage;            /* 0 = NONE, 1 = SCALAR_GRADY_KIPP */
    double shear_modulus;  /* mu, Hooke's law */
    double elasticity_limit, melt_energy, young_modulus;
    /* IMaterial ranges (clamping) and minimals (derivative criterion) */
    double rho_lo, rho_hi, u_lo, u_hi, d_lo, d_hi;
    double rho_min, u_min, s_min, d_min;
    /* DruckerPragerRheology (core/physics/Rheology.cpp:119-181): COHESION, INTERNAL_FRICTION, DRY_FRICTION */
    double cohesion, internal_friction, dry_friction;
} orc_material;

typedef struct {
    int solid_stress;        /* ForceEnum::SOLID_STRESS: SolidStressForce in the term set */
    int adaptive_h;          /* AdaptiveSmoothingLength (1) or ConstSmoothingLength (0) */
    int sum_only_undamaged;  /* RunSettingsId::SPH_SUM_ONLY_UNDAMAGED */
    int has_reduce;          /* storage holds STRESS_REDUCING (some material has a rheology) */
    int has_damage;          /* storage holds DAMAGE */
    double av_alpha, av_beta;
    double h_lo, h_hi;       /* SPH_SMOOTHING_LENGTH_RANGE */
    int criteria;            /* bit0 COURANT, bit1 DERIVATIVES, bit2 ACCELERATION, bit3 DIVERGENCE */
    double courant, deriv_factor, div_factor, max_dt;
    int n_mat;
    const orc_material* mat;
    /* GravitySolver<SymmetricSolver> (core/sph/solvers/GravitySolver.cpp:64-91): 1 adds the self-gravity of
     * BruteForceGravity with the GravityLutKernel (GRAVITY_KERNEL = SPH_KERNEL) to the accelerations
     * before the pair sums; 0 none.  Factory::getGravity (core/system/Factory.cpp:383-385) builds
     * BruteForceGravity with its default constant, Constants::gravity, whatever GRAVITY_CONSTANT holds. */
    int gravity;
} orc_config;

/* M4 look-up table (LutKernel<3> built from CubicSpline<3>): values and grads, 40001 entries each. */
void orc_lut(double* values, double* grads);
double orc_lut_value(const double* values, double q2);
double orc_lut_grad(const double* grads, double q2);

/* Tillotson: (rho, u) -> (p, cs) */
void orc_tillotson(const orc_material* m, double rho, double u, double* p, double* cs);

/* Eigenvalues of a symmetric tensor given as {xx, yy, zz, xy, xz, yz} */
void orc_eigenvalues(const double t[6], double out[3]);

/* Unordered neighbour pairs {i<j} with |r_i - r_j|^2 < (2*0.5*(h_i+h_j))^2, sorted by (i, j).
 * Returns the number of pairs; writes at most cap of them. */
size_t orc_find_pairs(size_t n, const double* x, const double* y, const double* z, const double* h,
                      uint32_t* pi, uint32_t* pj, size_t cap);

/* SymmetricSolver<3>::integrate: material initialize, equation initialize, the pair loop,
 * finalize (reverse term order), material finalize.  Highest derivatives must be zero on entry
 * (as ISolver::integrate requires). */
void orc_integrate(const orc_config* cfg, double* planes, uint32_t* iplanes, size_t n);

/* Zero all highest derivatives (Storage::zeroHighestDerivatives). */
void orc_zero_derivatives(double* planes, size_t n);

/* One PredictorCorrector step with time step dt; returns the criterion's next time step and
 * writes the criterion id (CriterionId enum order of the reference) into *crit. */
double orc_step_pc(const orc_config* cfg, double* planes, uint32_t* iplanes, size_t n, double dt, int* crit);

/* MultiCriterion::compute over the current state. */
double orc_timestep(const orc_config* cfg, const double* planes, const uint32_t* iplanes, size_t n, int* crit);

#ifdef __cplusplus
}
#endif
#endif
