/*
 * TEST INFRASTRUCTURE -- not part of the product.  CPU restatement of OpenSPH's per-timestep hot path,
 * used only as the checker (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).  See osph_oracle.h.
 *
 * Expression order follows the reference term by term (compiled with -ffp-contract=off), so that the
 * only differences to the reference are the summation order of pair contributions.
 */
#include "osph_oracle.h"
#include "osph_oracle_gravity.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_ 3.14159265358979323846264338327950288
#define PL(k) (P + (size_t)(k) * n)
#define IPL(k) (I + (size_t)(k) * n)

/* core/math/MathUtils.h:26-34: EPS and LARGE are float literals widened to double */
static const double EPS_ = (double)1.e-12f;
static const double LARGE_ = (double)1.e20f;

static inline double mind(double a, double b) { return (a < b) ? a : b; } /* core/math/MathBasic.h:15-17 */
static inline double maxd(double a, double b) { return (a > b) ? a : b; } /* core/math/MathBasic.h:20-22 */
static inline double pow3(double v) { return v * v * v; }               /* core/math/MathUtils.h:132-134 */

/* ---- kernel ---------------------------------------------------------------------------------------- */

/* CubicSpline<3>::valueImpl / gradImpl (core/sph/kernel/Kernel.h:148-188), D = 3 normalisation 1/pi */
static double m4_value(double qSqr) {
    const double norm = 1.0 / PI_;
    const double q = sqrt(qSqr);
    if (q < 1.0) {
        return norm * (0.25 * pow3(2.0 - q) - pow3(1.0 - q));
    }
    if (q < 2.0) {
        return norm * (0.25 * pow3(2.0 - q));
    }
    return 0.0;
}

static double m4_grad(double qSqr) {
    const double norm = 1.0 / PI_;
    const double q = sqrt(qSqr);
    if (q == 0.0) {
        return -3.0 * norm;
    }
    if (q < 1.0) {
        return (1.0 / q) * norm * (-0.75 * ((2.0 - q) * (2.0 - q)) + 3.0 * ((1.0 - q) * (1.0 - q)));
    }
    if (q < 2.0) {
        return (1.0 / q) * norm * (-0.75 * ((2.0 - q) * (2.0 - q)));
    }
    return 0.0;
}

/* LutKernel<3>(CubicSpline<3>) constructor (core/sph/kernel/Kernel.h:83-101): radius 2,
 * qSqrToIdx = 40000 / 4, entries at q^2 = i / qSqrToIdx for i in [0, 40000]. */
void orc_lut(double* values, double* grads) {
    const double rad = 2.0;
    const double radInvSqr = 1.0 / (rad * rad);
    const double qSqrToIdx = (double)ORC_LUT_ENTRIES * radInvSqr;
    for (int i = 0; i < ORC_LUT_ENTRIES + 1; ++i) {
        const double qSqr = (double)i / qSqrToIdx;
        values[i] = m4_value(qSqr);
        grads[i] = m4_grad(qSqr);
    }
}

static const double LUT_Q2_TO_IDX = (double)ORC_LUT_ENTRIES * 0.25;

/* LutKernel::valueImpl / gradImpl (core/sph/kernel/Kernel.h:111-144): linear interpolation in q^2 */
double orc_lut_value(const double* values, double q2) {
    if (q2 >= 4.0) return 0.0;
    const double fi = LUT_Q2_TO_IDX * q2;
    const uint32_t i1 = (uint32_t)fi;
    const double ratio = fi - (double)i1;
    return values[i1] * (1.0 - ratio) + values[i1 + 1] * ratio;
}

double orc_lut_grad(const double* grads, double q2) {
    if (q2 >= 4.0) return 0.0;
    const double fi = LUT_Q2_TO_IDX * q2;
    const uint32_t i1 = (uint32_t)fi;
    const double ratio = fi - (double)i1;
    return grads[i1] * (1.0 - ratio) + grads[i1 + 1] * ratio;
}

/* ---- material physics ------------------------------------------------------------------------------ */

/* TillotsonEos::evaluate (core/physics/Eos.cpp:198-238) */
void orc_tillotson(const orc_material* m, double rho, double u, double* p_out, double* cs_out) {
    const double eta = rho / m->rho0;
    const double mu = eta - 1.0;
    const double denom = u / (m->u0 * eta * eta) + 1.0;
    /* compressed phase */
    const double pc = (m->a + m->b / denom) * rho * u + m->A * mu + m->B * mu * mu;
    double dpdu = m->a * rho + m->b * rho / (denom * denom);
    double dpdrho = m->a * u + m->b * u * (3.0 * denom - 2.0) / (denom * denom) + m->A / m->rho0 +
                    2.0 * m->B * mu / m->rho0;
    const double csc = dpdrho + dpdu * pc / (rho * rho);
    /* expanded phase */
    const double rhoExp = m->rho0 / rho - 1.0;
    const double betaExp = exp(-mind(m->beta * rhoExp, 70.0));
    const double alphaExp = exp(-mind(m->alpha * (rhoExp * rhoExp), 70.0));
    const double pe = m->a * rho * u + (m->b * rho * u / denom + m->A * mu * betaExp) * alphaExp;
    dpdu = m->a * rho + alphaExp * m->b * rho / (denom * denom);
    dpdrho = m->a * u + alphaExp * (m->b * u * (3.0 * denom - 2.0) / (denom * denom)) +
             alphaExp * (m->b * u * rho / denom) * m->rho0 * (2.0 * m->alpha * rhoExp) / (rho * rho) +
             alphaExp * m->A * betaExp *
                 (1.0 / m->rho0 + m->rho0 * mu / (rho * rho) * (2.0 * m->alpha * rhoExp + m->beta));
    double cse = dpdrho + dpdu * pe / (rho * rho);
    cse = maxd(cse, 0.0);
    double p = pc, cs = csc;
    if (rho <= m->rho0 && u > m->ucv) {
        p = pe;
        cs = cse;
    } else if (rho <= m->rho0 && u > m->uiv && u <= m->ucv) {
        p = ((u - m->uiv) * pe + (m->ucv - u) * pc) / (m->ucv - m->uiv);
        cs = ((u - m->uiv) * cse + (m->ucv - u) * csc) / (m->ucv - m->uiv);
    }
    cs = maxd(cs, 0.25 * m->A / m->rho0);
    *p_out = p;
    *cs_out = sqrt(cs);
}

/* sqrtInv / sqrtApprox (core/math/MathUtils.h:44-62): single-precision bit-trick estimate with one
 * Newton step, used by norm(Vector) (core/objects/geometry/Vector.h:681-685). */
static double sqrt_inv_approx(double f) {
    int32_t i;
    float x2 = (float)f * 0.5f;
    float y = (float)f;
    memcpy(&i, &y, sizeof(float));
    i = 0x5f3759df - (i >> 1);
    memcpy(&y, &i, sizeof(float));
    return (double)(y * (1.5f - (x2 * y * y)));
}

static double sqrt_approx(double f) {
    if (f == 0.0) return 0.0;
    return 1.0 / sqrt_inv_approx(f);
}

/* findEigenvalues (core/objects/geometry/SymmetricTensor.h:377-398); norm(SymmetricTensor) is
 * norm(max(diag, off)) (:290-294). t = {xx, yy, zz, xy, xz, yz}. */
void orc_eigenvalues(const double t[6], double out[3]) {
    const double d0 = t[0], d1 = t[1], d2 = t[2], o0 = t[3], o1 = t[4], o2 = t[5];
    const double mx0 = maxd(d0, o0), mx1 = maxd(d1, o1), mx2 = maxd(d2, o2);
    const double n = sqrt_approx(mx0 * mx0 + mx1 * mx1 + mx2 * mx2);
    if (n < 1.e-12) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double I1 = d0 * 1.0 + d1 * 1.0 + d2 * 1.0;
    const double I2 = (o0 * o0 + o1 * o1 + o2 * o2) - (d1 * d2 + d2 * d0 + d0 * d1);
    const double I3 = d0 * d1 * d2 + 2.0 * o0 * o1 * o2 - ((o0 * o0) * d2 + (o1 * o1) * d1 + (o2 * o2) * d0);
    const double p = -I1 / n;
    const double q = -I2 / (n * n);
    const double r = -I3 / pow3(n);
    const double a = q - p * p / 3.0;
    const double b = (2.0 * pow3(p) - 9.0 * p * q + 27.0 * r) / 27.0;
    const double aCub = pow3(a) / 27.0;
    if (0.25 * b * b + aCub >= 0.0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double t1 = 2.0 * sqrt(-a / 3.0);
    const double phi = acos(-0.5 * b / sqrt(-aCub));
    const double v0 = phi / 3.0, v1 = (phi + 2.0 * PI_) / 3.0, v2 = (phi + 4.0 * PI_) / 3.0;
    const double p3 = p / 3.0;
    out[0] = (t1 * cos(v0) - p3) * n;
    out[1] = (t1 * cos(v1) - p3) * n;
    out[2] = (t1 * cos(v2) - p3) * n;
}

/* DruckerPragerRheology::initialize (core/physics/Rheology.cpp:119-181) for particle i, without acoustic
 * fluidization (USE_ACOUSTIC_FLUDIZATION defaults to false): the pressure-dependent yield strength
 * Y_i = Y_0 + mu_i p / (1 + mu_i max(p, 0) / (Y_M - Y_0)), the dry-friction strength mu_d p, their blend
 * by damage, the melting reduction and the deviatoric stress reduced to Y / sqrt(J2). */
static void drucker_prager(const orc_config* cfg, const orc_material* m, double* P, size_t n, size_t i) {
    const double d = cfg->has_damage ? pow3(PL(ORC_D)[i]) : 0.0;
    if (PL(ORC_P)[i] < 0.0) {
        PL(ORC_P)[i] = (1.0 - d) * PL(ORC_P)[i];
    }
    const double p = PL(ORC_P)[i];
    const double Y_0 = m->cohesion, Y_M = m->elasticity_limit, mu_i = m->internal_friction, mu_d = m->dry_friction;
    const double Y_i = Y_0 + mu_i * p / (1.0 + mu_i * maxd(p, 0.0) / (Y_M - Y_0));
    const double Y_d = mu_d * p;
    double Y;
    if (Y_d > Y_i) {
        Y = Y_i;
    } else {
        Y = (1.0 - d) * Y_i + d * Y_d;
    }
    const double u = PL(ORC_U)[i];
    Y = (u < 1.e-5 * m->melt_energy) ? Y : Y * maxd(1.0 - u / m->melt_energy, 0.0);
    double* s[5] = { PL(ORC_SXX), PL(ORC_SYY), PL(ORC_SXY), PL(ORC_SXZ), PL(ORC_SYZ) };
    if (Y < EPS_) {
        PL(ORC_RED)[i] = 0.0;
        for (int k = 0; k < 5; ++k) s[k][i] = 0.0;
        return;
    }
    const double xx = s[0][i], yy = s[1][i], zz = -s[0][i] - s[1][i];
    const double dd = (xx * xx + yy * yy + zz * zz) +
                      2.0 * (s[2][i] * s[2][i] + s[3][i] * s[3][i] + s[4][i] * s[4][i]);
    const double J2 = 0.5 * dd + EPS_;
    const double red = mind(Y / sqrt(J2), 1.0);
    PL(ORC_RED)[i] = red;
    for (int k = 0; k < 5; ++k) s[k][i] = s[k][i] * red;
}

/* EosMaterial::initialize (core/sph/Materials.cpp:47-58) followed by VonMisesRheology::initialize
 * (core/physics/Rheology.cpp:36-83) or DruckerPragerRheology::initialize (core/physics/Rheology.cpp:119-181,
 * without acoustic fluidization) for particle i. */
static void material_initialize(const orc_config* cfg, double* P, size_t n, const uint32_t* I, size_t i) {
    const orc_material* m = &cfg->mat[IPL(ORC_MAT)[i]];
    double p, cs;
    orc_tillotson(m, PL(ORC_RHO)[i], PL(ORC_U)[i], &p, &cs);
    PL(ORC_P)[i] = p;
    PL(ORC_CS)[i] = cs;
    if (m->yielding == 2) {
        drucker_prager(cfg, m, P, n, i);
        return;
    }
    if (m->yielding != 1) return;

    const double d = cfg->has_damage ? pow3(PL(ORC_D)[i]) : 0.0;
    if (PL(ORC_P)[i] < 0.0) {
        PL(ORC_P)[i] = (1.0 - d) * PL(ORC_P)[i];
    }
    const double unorm = PL(ORC_U)[i] / m->melt_energy;
    double Y = unorm < 1.e-5 ? m->elasticity_limit : m->elasticity_limit * maxd(1.0 - unorm, 0.0);
    Y = (1.0 - d) * Y;
    double* s[5] = { PL(ORC_SXX), PL(ORC_SYY), PL(ORC_SXY), PL(ORC_SXZ), PL(ORC_SYZ) };
    if (Y < EPS_) {
        PL(ORC_RED)[i] = 0.0;
        for (int k = 0; k < 5; ++k) s[k][i] = 0.0;
        return;
    }
    const double xx = s[0][i], yy = s[1][i], zz = -s[0][i] - s[1][i];
    /* ddot(TracelessTensor, TracelessTensor) (core/objects/geometry/TracelessTensor.h:385-387) */
    const double dd = (xx * xx + yy * yy + zz * zz) +
                      2.0 * (s[2][i] * s[2][i] + s[3][i] * s[3][i] + s[4][i] * s[4][i]);
    const double J2 = 0.5 * dd + 1.e-15;
    const double red = mind(Y / sqrt(3.0 * J2), 1.0);
    PL(ORC_RED)[i] = red;
    for (int k = 0; k < 5; ++k) s[k][i] = s[k][i] * red;
}

/* ScalarGradyKippModel::integrate (core/physics/Damage.cpp:128-170) for particle i. */
static void damage_integrate(const orc_config* cfg, double* P, size_t n, const uint32_t* I, size_t i) {
    const orc_material* m = &cfg->mat[IPL(ORC_MAT)[i]];
    if (m->damage != 1) return;
    const double D = PL(ORC_D)[i];
    if (D >= m->d_hi) {
        PL(ORC_DD)[i] = LARGE_;
        return;
    }
    const double pr = PL(ORC_P)[i];
    const double sxx = PL(ORC_SXX)[i], syy = PL(ORC_SYY)[i];
    const double t[6] = { sxx - pr * 1.0, syy - pr * 1.0, (-sxx - syy) - pr * 1.0,
                          PL(ORC_SXY)[i] - pr * 0.0, PL(ORC_SXZ)[i] - pr * 0.0, PL(ORC_SYZ)[i] - pr * 0.0 };
    double sig[3];
    orc_eigenvalues(t, sig);
    const double sigMax = maxd(maxd(sig[0], sig[1]), sig[2]);
    const double young_red = maxd((1.0 - pow3(D)) * m->young_modulus, 1.e-20);
    const double strain = sigMax / young_red;
    const double ratio = strain / PL(ORC_EPSMIN)[i];
    if (ratio <= 1.0) return;
    PL(ORC_DD)[i] = PL(ORC_GROWTH)[i] * cbrt(mind(pow(ratio, PL(ORC_MZERO)[i]), (double)IPL(ORC_NFLAWS)[i]));
}

/* ---- neighbour pairs ------------------------------------------------------------------------------- */

typedef struct {
    double lo[3], cell;
    int dim[3];
    uint32_t* start; /* dim product + 1 */
    uint32_t* idx;   /* particles sorted by cell, stable */
} grid_t;

static size_t cell_of(const grid_t* g, double x, double y, double z, int c[3]) {
    const double v[3] = { x, y, z };
    for (int k = 0; k < 3; ++k) {
        int ci = (int)floor((v[k] - g->lo[k]) / g->cell);
        if (ci < 0) ci = 0;
        if (ci >= g->dim[k]) ci = g->dim[k] - 1;
        c[k] = ci;
    }
    return ((size_t)c[2] * g->dim[1] + c[1]) * g->dim[0] + c[0];
}

/* Uniform grid with cells no smaller than the largest search radius 2*h_max, so that every pair with
 * |r_ij| < 2*hbar <= 2*h_max lies in adjacent cells. */
static void grid_build(grid_t* g, size_t n, const double* x, const double* y, const double* z, const double* h) {
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    double hmax = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double v[3] = { x[i], y[i], z[i] };
        for (int k = 0; k < 3; ++k) {
            lo[k] = v[k] < lo[k] ? v[k] : lo[k];
            hi[k] = v[k] > hi[k] ? v[k] : hi[k];
        }
        hmax = h[i] > hmax ? h[i] : hmax;
    }
    double cell = 2.0 * hmax * (1.0 + 1.e-9);
    double ext = 0.0;
    for (int k = 0; k < 3; ++k) ext = (hi[k] - lo[k]) > ext ? (hi[k] - lo[k]) : ext;
    if (cell < ext / 256.0) cell = ext / 256.0; /* keep the grid bounded; larger cells stay exact */
    if (!(cell > 0.0)) cell = 1.0;
    g->cell = cell;
    size_t ncell = 1;
    for (int k = 0; k < 3; ++k) {
        g->lo[k] = lo[k];
        g->dim[k] = (int)floor((hi[k] - lo[k]) / cell) + 1;
        ncell *= (size_t)g->dim[k];
    }
    g->start = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
    g->idx = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t* key = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    int c[3];
    for (size_t i = 0; i < n; ++i) {
        key[i] = (uint32_t)cell_of(g, x[i], y[i], z[i], c);
        g->start[key[i] + 1]++;
    }
    for (size_t k = 0; k < ncell; ++k) g->start[k + 1] += g->start[k];
    uint32_t* fill = (uint32_t*)malloc((ncell + 1) * sizeof(uint32_t));
    memcpy(fill, g->start, (ncell + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) g->idx[fill[key[i]]++] = (uint32_t)i;
    free(fill);
    free(key);
}

static void grid_free(grid_t* g) {
    free(g->start);
    free(g->idx);
}

/* The pair test of SymmetricSolver<3>::loop (core/sph/solvers/SymmetricSolver.cpp:126-131): keep the
 * pair iff getSqrLength(r_i - r_j) < sqr(2 * hbar), hbar = 0.5 * (h_i + h_j); getSqrLength sums the three
 * squares left to right (core/objects/geometry/Vector.h:571-574). */
static inline int is_pair(double dx, double dy, double dz, double hi, double hj) {
    const double hbar = 0.5 * (hi + hj);
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double rad = 2.0 * hbar;
    return d2 < rad * rad;
}

typedef void (*pair_fn)(void* ctx, uint32_t i, uint32_t j);

/* Visits every unordered pair once, as (i, j) with i < j, i ascending. */
static void for_each_pair(size_t n, const double* x, const double* y, const double* z, const double* h,
                          pair_fn fn, void* ctx) {
    if (n == 0) return;
    grid_t g;
    grid_build(&g, n, x, y, z, h);
    int c[3];
    for (size_t i = 0; i < n; ++i) {
        cell_of(&g, x[i], y[i], z[i], c);
        for (int dz = -1; dz <= 1; ++dz) {
            const int cz = c[2] + dz;
            if (cz < 0 || cz >= g.dim[2]) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int cy = c[1] + dy;
                if (cy < 0 || cy >= g.dim[1]) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int cx = c[0] + dx;
                    if (cx < 0 || cx >= g.dim[0]) continue;
                    const size_t cell = ((size_t)cz * g.dim[1] + cy) * g.dim[0] + cx;
                    for (uint32_t k = g.start[cell]; k < g.start[cell + 1]; ++k) {
                        const uint32_t j = g.idx[k];
                        if (j <= i) continue;
                        if (is_pair(x[i] - x[j], y[i] - y[j], z[i] - z[j], h[i], h[j])) fn(ctx, (uint32_t)i, j);
                    }
                }
            }
        }
    }
    grid_free(&g);
}

typedef struct {
    uint32_t *pi, *pj;
    size_t cnt, cap;
} pair_list;

static void collect_pair(void* ctx, uint32_t i, uint32_t j) {
    pair_list* l = (pair_list*)ctx;
    if (l->cnt < l->cap) {
        l->pi[l->cnt] = i;
        l->pj[l->cnt] = j;
    }
    l->cnt++;
}

static int cmp_pair(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

size_t orc_find_pairs(size_t n, const double* x, const double* y, const double* z, const double* h,
                      uint32_t* pi, uint32_t* pj, size_t cap) {
    pair_list l = { pi, pj, 0, cap };
    for_each_pair(n, x, y, z, h, collect_pair, &l);
    const size_t m = l.cnt < cap ? l.cnt : cap;
    uint64_t* keys = (uint64_t*)malloc((m ? m : 1) * sizeof(uint64_t));
    for (size_t k = 0; k < m; ++k) keys[k] = ((uint64_t)pi[k] << 32) | pj[k];
    qsort(keys, m, sizeof(uint64_t), cmp_pair);
    for (size_t k = 0; k < m; ++k) {
        pi[k] = (uint32_t)(keys[k] >> 32);
        pj[k] = (uint32_t)(keys[k] & 0xffffffffu);
    }
    free(keys);
    return l.cnt;
}

/* ---- the pair loop --------------------------------------------------------------------------------- */

typedef struct {
    const orc_config* cfg;
    double* P;
    uint32_t* I;
    size_t n;
    double lut_grads[ORC_LUT_ENTRIES + 1];
} loop_ctx;

/* One symmetric pair evaluation: the kernel gradient of SymmetrizeSmoothingLengths<LutKernel<3>>
 * (core/sph/kernel/Kernel.h:32-36,640-643) followed by every derivative's evalSymmetric, in the order
 * the term set requires them (core/sph/solvers/StandardSets.cpp:14-95, SymmetricSolver.cpp:35). */
static void eval_pair(void* vctx, uint32_t i, uint32_t j) {
    loop_ctx* c = (loop_ctx*)vctx;
    const orc_config* cfg = c->cfg;
    double* P = c->P;
    uint32_t* I = c->I;
    const size_t n = c->n;

    const double rx = PL(ORC_X)[i] - PL(ORC_X)[j];
    const double ry = PL(ORC_Y)[i] - PL(ORC_Y)[j];
    const double rz = PL(ORC_Z)[i] - PL(ORC_Z)[j];
    const double hi = PL(ORC_H)[i], hj = PL(ORC_H)[j];
    const double hbar = 0.5 * (hi + hj);
    /* Kernel::grad: r * pow<5>(1/h) * gradImpl(|r|^2 / h^2) */
    const double hInv = 1.0 / hbar;
    const double d2 = rx * rx + ry * ry + rz * rz;
    const double g = orc_lut_grad(c->lut_grads, d2 * (hInv * hInv));
    const double h5 = (hInv * hInv) * (hInv * hInv) * hInv;
    const double gx = rx * h5 * g, gy = ry * h5 * g, gz = rz * h5 * g;

    const double mi = PL(ORC_M)[i], mj = PL(ORC_M)[j];
    const double rhoi = PL(ORC_RHO)[i], rhoj = PL(ORC_RHO)[j];
    const double vxi = PL(ORC_VX)[i], vyi = PL(ORC_VY)[i], vzi = PL(ORC_VZ)[i];
    const double vxj = PL(ORC_VX)[j], vyj = PL(ORC_VY)[j], vzj = PL(ORC_VZ)[j];
    double *ax = PL(ORC_AX), *ay = PL(ORC_AY), *az = PL(ORC_AZ);

    const int undamaged = !(cfg->sum_only_undamaged && cfg->has_reduce) ||
                          !(IPL(ORC_FLAG)[i] != IPL(ORC_FLAG)[j] || PL(ORC_RED)[i] == 0.0 || PL(ORC_RED)[j] == 0.0);

    /* VelocityDivergence<CenterDensityDiscr> (core/sph/equations/DerivativeHelpers.h:355-371,292-304) */
    {
        const double dvdot = (vxj - vxi) * gx + (vyj - vyi) * gy + (vzj - vzi) * gz;
        PL(ORC_DIVV)[i] += mj / rhoi * dvdot;
        PL(ORC_DIVV)[j] += mi / rhoj * dvdot;
    }
    /* PressureGradient<StandardForceDiscr> (core/sph/equations/EquationTerm.cpp:12-24,61-66) through
     * AccelerationTemplate::evalSymmetric (DerivativeHelpers.h:229-242) */
    {
        const double pi_ = PL(ORC_P)[i], pj_ = PL(ORC_P)[j];
        const double coef = pi_ / (rhoi * rhoi) + pj_ / (rhoj * rhoj);
        const double fx = -(coef * gx), fy = -(coef * gy), fz = -(coef * gz);
        ax[i] += mj * fx; ay[i] += mj * fy; az[i] += mj * fz;
        ax[j] -= mi * fx; ay[j] -= mi * fy; az[j] -= mi * fz;
    }
    if (cfg->solid_stress && undamaged) {
        /* VelocityGradient: symmetricOuter(v_j - v_i, grad) (SymmetricTensor.h:362-368) */
        const double wx = vxj - vxi, wy = vyj - vyi, wz = vzj - vzi;
        const double T[6] = { wx * gx, wy * gy, wz * gz, 0.5 * (wx * gy + wy * gx), 0.5 * (wx * gz + wz * gx),
                              0.5 * (wy * gz + wz * gy) };
        const double ci = mj / rhoi, cj = mi / rhoj;
        for (int k = 0; k < 6; ++k) {
            PL(ORC_GXX + k)[i] += ci * T[k];
            PL(ORC_GXX + k)[j] += cj * T[k];
        }
        /* StressDivergence<StandardForceDiscr> (EquationTerm.cpp:113-140): (S_i/rho_i^2 + S_j/rho_j^2) * grad,
         * traceless tensor times vector (TracelessTensor.h:151-155) */
        const double ri2 = rhoi * rhoi, rj2 = rhoj * rhoj;
        const double xx = PL(ORC_SXX)[i] / ri2 + PL(ORC_SXX)[j] / rj2;
        const double yy = PL(ORC_SYY)[i] / ri2 + PL(ORC_SYY)[j] / rj2;
        const double xy = PL(ORC_SXY)[i] / ri2 + PL(ORC_SXY)[j] / rj2;
        const double xz = PL(ORC_SXZ)[i] / ri2 + PL(ORC_SXZ)[j] / rj2;
        const double yz = PL(ORC_SYZ)[i] / ri2 + PL(ORC_SYZ)[j] / rj2;
        const double fx = xx * gx + xy * gy + xz * gz;
        const double fy = xy * gx + yy * gy + yz * gz;
        const double fz = xz * gx + yz * gy + (-xx - yy) * gz;
        ax[i] += mj * fx; ay[i] += mj * fy; az[i] += mj * fz;
        ax[j] -= mi * fx; ay[j] -= mi * fy; az[j] -= mi * fz;
    }
    /* StandardAV::Derivative (core/sph/equations/av/Standard.h:62-83) */
    {
        const double dvdr = (vxi - vxj) * rx + (vyi - vyj) * ry + (vzi - vzj) * rz;
        if (dvdr < 0.0) {
            const double rhobar = 0.5 * (rhoi + rhoj);
            const double csbar = 0.5 * (PL(ORC_CS)[i] + PL(ORC_CS)[j]);
            const double mu = hbar * dvdr / (d2 + 1.e-2 * (hbar * hbar));
            const double av = 1.0 / rhobar * (-cfg->av_alpha * csbar * mu + cfg->av_beta * (mu * mu));
            const double fx = -(av * gx), fy = -(av * gy), fz = -(av * gz);
            const double heating = 0.5 * av * ((vxi - vxj) * gx + (vyi - vyj) * gy + (vzi - vzj) * gz);
            ax[i] += mj * fx; ay[i] += mj * fy; az[i] += mj * fz;
            ax[j] -= mi * fx; ay[j] -= mi * fy; az[j] -= mi * fz;
            PL(ORC_DU)[i] += mj * heating;
            PL(ORC_DU)[j] += mi * heating;
        }
    }
    /* NeighborCountTerm (core/sph/equations/HelperTerms.h:38-46) */
    IPL(ORC_NCNT)[i]++;
    IPL(ORC_NCNT)[j]++;
}

void orc_zero_derivatives(double* P, size_t n) {
    static const int planes[] = { ORC_AX, ORC_AY, ORC_AZ, ORC_DRHO, ORC_DU, ORC_DSXX, ORC_DSYY,
                                  ORC_DSXY, ORC_DSXZ, ORC_DSYZ, ORC_DD };
    for (size_t k = 0; k < sizeof(planes) / sizeof(planes[0]); ++k) memset(PL(planes[k]), 0, n * sizeof(double));
}

/* SymmetricSolver<3>::integrate (core/sph/solvers/SymmetricSolver.cpp:58-95) */
void orc_integrate(const orc_config* cfg, double* P, uint32_t* I, size_t n) {
    /* material->initialize */
    for (size_t i = 0; i < n; ++i) material_initialize(cfg, P, n, I, i);
    /* equations.initialize: AdaptiveSmoothingLength clamps h (EquationTerm.cpp:356-364) */
    if (cfg->adaptive_h) {
        for (size_t i = 0; i < n; ++i) PL(ORC_H)[i] = maxd(cfg->h_lo, mind(PL(ORC_H)[i], cfg->h_hi));
    }
    /* beforeLoop: the accumulated buffers start at zero and replace the storage buffers in afterLoop
     * (core/sph/equations/Accumulated.cpp:38-51,87-100) */
    static const int acc[] = { ORC_AX, ORC_AY, ORC_AZ, ORC_DU, ORC_DIVV, ORC_GXX, ORC_GYY, ORC_GZZ,
                               ORC_GXY, ORC_GXZ, ORC_GYZ };
    for (size_t k = 0; k < sizeof(acc) / sizeof(acc[0]); ++k) memset(PL(acc[k]), 0, n * sizeof(double));
    memset(IPL(ORC_NCNT), 0, n * sizeof(uint32_t));
    /* GravitySolver::loop (core/sph/solvers/GravitySolver.cpp:64-91): the self-gravity goes into the
     * accumulated accelerations first (BruteForceGravity::evalSelfGravity, BruteForceGravity.h:49-57),
     * then the SPH pair sums add to them; the planes AX, AY, AZ are consecutive */
    if (cfg->gravity) {
        double* gvalues = (double*)malloc((ORC_GRAVITY_LUT_ENTRIES + 1) * sizeof(double));
        double* ggrads = (double*)malloc((ORC_GRAVITY_LUT_ENTRIES + 1) * sizeof(double));
        orc_gravity_lut(gvalues, ggrads);
        const double G = 6.67408e-11; /* Constants::gravity (core/physics/Constants.h:32) */
        orc_gravity_direct(n, PL(ORC_X), PL(ORC_Y), PL(ORC_Z), PL(ORC_H), PL(ORC_M), ggrads, G, NULL, n, PL(ORC_AX));
        free(gvalues);
        free(ggrads);
    }

    loop_ctx* c = (loop_ctx*)malloc(sizeof(loop_ctx));
    double* values = (double*)malloc((ORC_LUT_ENTRIES + 1) * sizeof(double));
    orc_lut(values, c->lut_grads);
    free(values);
    c->cfg = cfg;
    c->P = P;
    c->I = I;
    c->n = n;
    for_each_pair(n, PL(ORC_X), PL(ORC_Y), PL(ORC_Z), PL(ORC_H), eval_pair, c);
    free(c);

    /* equations.finalize in reverse term order (EquationTerm.h:293-297) */
    for (size_t i = 0; i < n; ++i) {
        const orc_material* m = &cfg->mat[IPL(ORC_MAT)[i]];
        const double rho = PL(ORC_RHO)[i], divv = PL(ORC_DIVV)[i];
        /* AdaptiveSmoothingLength::finalize (EquationTerm.cpp:366-387) / ConstSmoothingLength::finalize
         * (:427-434); dv[H] is always zeroed, so no acceleration plane is kept for h */
        if (cfg->adaptive_h && PL(ORC_H)[i] > 2.0 * cfg->h_lo) {
            PL(ORC_VH)[i] = PL(ORC_H)[i] / 3.0 * divv;
        } else {
            PL(ORC_VH)[i] = 0.0;
        }
        /* ContinuityEquation::finalize, STANDARD mode (EquationTerm.cpp:289-299) */
        PL(ORC_DRHO)[i] += -rho * divv;
        /* SolidStressForce::finalize (EquationTerm.cpp:177-203), skipped for YieldingEnum::NONE */
        if (cfg->solid_stress && m->yielding != 0) {
            const double sxx = PL(ORC_SXX)[i], syy = PL(ORC_SYY)[i], szz = -sxx - syy;
            const double gxx = PL(ORC_GXX)[i], gyy = PL(ORC_GYY)[i], gzz = PL(ORC_GZZ)[i];
            const double gxy = PL(ORC_GXY)[i], gxz = PL(ORC_GXZ)[i], gyz = PL(ORC_GYZ)[i];
            const double dd = (sxx * gxx + syy * gyy + szz * gzz) +
                              2.0 * (PL(ORC_SXY)[i] * gxy + PL(ORC_SXZ)[i] * gxz + PL(ORC_SYZ)[i] * gyz);
            PL(ORC_DU)[i] += 1.0 / rho * dd;
            const double tr3 = (gxx * 1.0 + gyy * 1.0 + gzz * 1.0) / 3.0;
            const double twomu = 2.0 * m->shear_modulus;
            PL(ORC_DSXX)[i] += twomu * (gxx - tr3);
            PL(ORC_DSYY)[i] += twomu * (gyy - tr3);
            PL(ORC_DSXY)[i] += twomu * gxy;
            PL(ORC_DSXZ)[i] += twomu * gxz;
            PL(ORC_DSYZ)[i] += twomu * gyz;
        }
        /* PressureForce::finalize (EquationTerm.cpp:90-99) */
        PL(ORC_DU)[i] -= PL(ORC_P)[i] / rho * divv;
    }
    /* material->finalize: ScalarGradyKippModel::integrate */
    if (cfg->has_damage) {
        for (size_t i = 0; i < n; ++i) damage_integrate(cfg, P, n, I, i);
    }
}

/* ---- time integration ------------------------------------------------------------------------------ */

/* clampWithDerivative<Float> (core/objects/wrappers/Interval.h:158-162) */
static inline void clamp_with_derivative(double* v, double* dv, double lo, double hi) {
    const int zero = (*v >= hi && *dv > 0.0) || (*v <= lo && *dv < 0.0);
    *v = maxd(lo, mind(*v, hi));
    if (zero) *dv = 0.0;
}

typedef struct {
    int value, deriv, saved;
    int bounded; /* 0 none, 1 rho, 2 u, 3 damage */
} first_order_q;

static int first_order_list(const orc_config* cfg, first_order_q* q) {
    int k = 0;
    q[k++] = (first_order_q){ ORC_RHO, ORC_DRHO, ORC_PDRHO, 1 };
    q[k++] = (first_order_q){ ORC_U, ORC_DU, ORC_PDU, 2 };
    if (cfg->solid_stress) {
        for (int c = 0; c < 5; ++c) q[k++] = (first_order_q){ ORC_SXX + c, ORC_DSXX + c, ORC_PDSXX + c, 0 };
    }
    if (cfg->has_damage) q[k++] = (first_order_q){ ORC_D, ORC_DD, ORC_PDD, 3 };
    return k;
}

static void range_of(const orc_material* m, int which, double* lo, double* hi) {
    switch (which) {
    case 1: *lo = m->rho_lo; *hi = m->rho_hi; break;
    case 2: *lo = m->u_lo; *hi = m->u_hi; break;
    default: *lo = m->d_lo; *hi = m->d_hi; break;
    }
}

/* PredictorCorrector::stepParticles (core/timestepping/TimeStepping.cpp:286-346) followed by the
 * criterion (ITimeStepping::step, :53-73). */
double orc_step_pc(const orc_config* cfg, double* P, uint32_t* I, size_t n, double dt, int* crit) {
    const double dt2 = 0.5 * (dt * dt);
    first_order_q q[8];
    const int nq = first_order_list(cfg, q);
    /* makePredictions: second order (POSITION, unbounded), then first order with clamping */
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            double* r = PL(ORC_X + k);
            double* v = PL(ORC_VX + k);
            const double a = PL(ORC_AX + k)[i];
            r[i] += v[i] * dt + a * dt2;
            v[i] += a * dt;
        }
        PL(ORC_H)[i] += PL(ORC_VH)[i] * dt + 0.0 * dt2;
        const orc_material* m = &cfg->mat[IPL(ORC_MAT)[i]];
        for (int c = 0; c < nq; ++c) {
            double* x = PL(q[c].value) + i;
            double* dx = PL(q[c].deriv) + i;
            *x += *dx * dt;
            if (q[c].bounded) {
                double lo, hi;
                range_of(m, q[c].bounded, &lo, &hi);
                clamp_with_derivative(x, dx, lo, hi);
            }
        }
    }
    /* swap the highest derivatives into the predictions storage and zero them */
    for (int k = 0; k < 3; ++k) memcpy(PL(ORC_PAX + k), PL(ORC_AX + k), n * sizeof(double));
    for (int c = 0; c < nq; ++c) memcpy(PL(q[c].saved), PL(q[c].deriv), n * sizeof(double));
    orc_zero_derivatives(P, n);

    orc_integrate(cfg, P, I, n);

    /* makeCorrections (TimeStepping.cpp:302-322): pdv = new derivative, cdv = saved prediction */
    const double ca = 1.0 / 3.0, cb = 0.5;
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            const double pdv = PL(ORC_AX + k)[i], cdv = PL(ORC_PAX + k)[i];
            PL(ORC_X + k)[i] -= ca * (cdv - pdv) * dt2;
            PL(ORC_VX + k)[i] -= cb * (cdv - pdv) * dt;
        }
        const orc_material* m = &cfg->mat[IPL(ORC_MAT)[i]];
        for (int c = 0; c < nq; ++c) {
            double* x = PL(q[c].value) + i;
            double* pdx = PL(q[c].deriv) + i;
            const double cdx = PL(q[c].saved)[i];
            *x -= 0.5 * (cdx - *pdx) * dt;
            if (q[c].bounded) {
                double lo, hi;
                range_of(m, q[c].bounded, &lo, &hi);
                clamp_with_derivative(x, pdx, lo, hi);
            }
        }
    }
    return orc_timestep(cfg, P, I, n, crit);
}

/* MultiCriterion::compute (core/timestepping/TimeStepCriterion.cpp:389-419) over Courant (:315-356),
 * DerivativeCriterion with the minimal-step mean (:123-211), AccelerationCriterion (:221-257) and
 * DivergenceCriterion (:267-305).  CriterionId values: 1 MAXIMAL_VALUE, 2 DERIVATIVE, 3 CFL_CONDITION,
 * 4 ACCELERATION, 5 DIVERGENCE. */
double orc_timestep(const orc_config* cfg, const double* P, const uint32_t* I, size_t n, int* crit) {
    double best = INFINITY;
    int bestId = 0;
    const double maxStep = cfg->max_dt;
    if (cfg->criteria & 1) {
        double mn = INFINITY;
        for (size_t i = 0; i < n; ++i) {
            const double cs = PL(ORC_CS)[i];
            if (cs > 0.0) mn = mind(mn, cfg->courant * PL(ORC_H)[i] / cs);
        }
        const double v = mn > maxStep ? maxStep : mn;
        const int id = mn > maxStep ? 1 : 3;
        if (v < best) { best = v; bestId = id; }
    }
    if (cfg->criteria & 2) {
        first_order_q q[8];
        const int nq = first_order_list(cfg, q);
        double total = INFINITY;
        int id = 0;
        /* stress components are visited as a SymmetricTensor: xx, yy, zz = -xx-yy, xy, xz, yz */
        for (int c = 0; c < nq; ++c) {
            double mn = INFINITY;
            const int isS = (q[c].value >= ORC_SXX && q[c].value <= ORC_SYZ);
            if (isS && q[c].value != ORC_SXX) continue; /* handled together with SXX */
            for (size_t i = 0; i < n; ++i) {
                const orc_material* m = &cfg->mat[IPL(ORC_MAT)[i]];
                double vs[6], dvs[6], minValue;
                int nc;
                if (isS) {
                    const double xx = PL(ORC_SXX)[i], yy = PL(ORC_SYY)[i];
                    const double dxx = PL(ORC_DSXX)[i], dyy = PL(ORC_DSYY)[i];
                    vs[0] = fabs(xx); vs[1] = fabs(yy); vs[2] = fabs(-xx - yy);
                    vs[3] = fabs(PL(ORC_SXY)[i]); vs[4] = fabs(PL(ORC_SXZ)[i]); vs[5] = fabs(PL(ORC_SYZ)[i]);
                    dvs[0] = fabs(dxx); dvs[1] = fabs(dyy); dvs[2] = fabs(-dxx - dyy);
                    dvs[3] = fabs(PL(ORC_DSXY)[i]); dvs[4] = fabs(PL(ORC_DSXZ)[i]); dvs[5] = fabs(PL(ORC_DSYZ)[i]);
                    nc = 6;
                    minValue = m->s_min;
                } else {
                    vs[0] = fabs(PL(q[c].value)[i]);
                    dvs[0] = fabs(PL(q[c].deriv)[i]);
                    nc = 1;
                    minValue = q[c].bounded == 1 ? m->rho_min : (q[c].bounded == 2 ? m->u_min : m->d_min);
                }
                for (int k = 0; k < nc; ++k) {
                    if (fabs(vs[k]) < 2.0 * minValue) continue;
                    const double value = cfg->deriv_factor * (vs[k] + minValue) / (dvs[k] + EPS_);
                    mn = mind(mn, value);
                }
            }
            if (mn < total) { total = mn; id = 2; }
        }
        if (total > maxStep) { total = maxStep; id = 1; }
        if (total < best) { best = total; bestId = id; }
    }
    if (cfg->criteria & 4) {
        double mn = INFINITY;
        for (size_t i = 0; i < n; ++i) {
            const double ax = PL(ORC_AX)[i], ay = PL(ORC_AY)[i], az = PL(ORC_AZ)[i];
            const double dvNorm = ax * ax + ay * ay + az * az;
            if (dvNorm > EPS_) {
                const double h = PL(ORC_H)[i];
                mn = mind(mn, cfg->deriv_factor * sqrt(sqrt((h * h) / dvNorm)));
            }
        }
        const double v = mn > maxStep ? maxStep : mn;
        const int id = mn > maxStep ? 1 : 4;
        if (v < best) { best = v; bestId = id; }
    }
    if (cfg->criteria & 8) {
        double mn = INFINITY;
        for (size_t i = 0; i < n; ++i) {
            const double dv = fabs(PL(ORC_DIVV)[i]);
            if (dv > EPS_) mn = mind(mn, cfg->div_factor / dv);
        }
        const double v = mn > maxStep ? maxStep : mn;
        const int id = mn > maxStep ? 1 : 5;
        if (v < best) { best = v; bestId = id; }
    }
    if (crit) *crit = bestId;
    return best;
}
