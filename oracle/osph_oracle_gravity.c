/*
 * TEST INFRASTRUCTURE -- not part of the product.  Strict-IEEE C restatement of OpenSPH's self-gravity;
 * see osph_oracle_gravity.h.  Reference paths are relative to the OpenSPH source root.
 */
#include "osph_oracle_gravity.h"
#include "osph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double EPS_ = (double)1.e-12f; /* core/math/MathUtils.h:26-30 */

/* GravityKernel<CubicSpline<3>>::valueImpl (core/sph/kernel/GravityKernel.h:82-92); pow<N> as
 * core/math/MathUtils.h:127-142 (products, left to right). */
static double gk_value(double qSqr) {
    const double q = sqrt(qSqr);
    if (q < 1.0) {
        const double q4 = (q * q) * (q * q), q5 = ((q * q) * (q * q)) * q;
        return 2.0 / 3.0 * qSqr - 3.0 / 10.0 * q4 + 1.0 / 10.0 * q5 - 7.0 / 5.0;
    }
    const double q3 = q * q * q, q5 = ((q * q) * (q * q)) * q;
    return 4.0 / 3.0 * qSqr - q3 + 3.0 / 10.0 * (qSqr * qSqr) - 1.0 / 30.0 * q5 - 8.0 / 5.0 + 1.0 / (15.0 * q);
}

/* GravityKernel<CubicSpline<3>>::gradImpl (GravityKernel.h:94-107) */
static double gk_grad(double qSqr) {
    const double q = sqrt(qSqr);
    if (q == 0.0) {
        return 4.0 / 3.0;
    }
    if (q < 1.0) {
        return 1.0 / q * (4.0 / 3.0 * q - 6.0 / 5.0 * (q * q * q) + 1.0 / 2.0 * (qSqr * qSqr));
    }
    return 1.0 / q *
           (8.0 / 3.0 * q - 3.0 * qSqr + 6.0 / 5.0 * (q * q * q) - 1.0 / 6.0 * (qSqr * qSqr) - 1.0 / (15.0 * qSqr));
}

/* LutKernel<3>(GravityKernel<CubicSpline<3>>) (core/sph/kernel/Kernel.h:83-101): radius 2, the same
 * table layout as the M4 kernel, so orc_lut_value / orc_lut_grad interpolate it. */
void orc_gravity_lut(double* values, double* grads) {
    const double rad = 2.0;
    const double qSqrToIdx = (double)ORC_GRAVITY_LUT_ENTRIES * (1.0 / (rad * rad));
    for (int i = 0; i < ORC_GRAVITY_LUT_ENTRIES + 1; ++i) {
        const double qSqr = (double)i / qSqrToIdx;
        values[i] = gk_value(qSqr);
        grads[i] = gk_grad(qSqr);
    }
}

static inline double dot3(double x, double y, double z) { return x * x + y * y + z * z; } /* Vector.h:566-569 */

/* GravityLutKernel::grad (GravityKernel.h:60-70).  Without a table the kernel radius is 0 and the
 * inverse-square branch is taken for every separation. */
void orc_gravity_grad(const double* grads, double rx, double ry, double rz, double h, double out[3]) {
    if (grads) {
        const double hInv = 1.0 / h;
        const double sx = rx * hInv, sy = ry * hInv, sz = rz * hInv;
        const double qSqr = dot3(sx, sy, sz);
        if (!(qSqr + EPS_ >= 4.0)) {
            const double g = orc_lut_grad(grads, qSqr);
            const double h3 = hInv * hInv * hInv;
            out[0] = (h3 * rx) * g;
            out[1] = (h3 * ry) * g;
            out[2] = (h3 * rz) * g;
            return;
        }
    }
    const double l = sqrt(dot3(rx, ry, rz));
    const double l3 = l * l * l;
    out[0] = rx / l3;
    out[1] = ry / l3;
    out[2] = rz / l3;
}

/* GravityLutKernel::value (GravityKernel.h:45-58) */
double orc_gravity_value(const double* values, double rx, double ry, double rz, double h) {
    if (values) {
        const double hInv = 1.0 / h;
        const double sx = rx * hInv, sy = ry * hInv, sz = rz * hInv;
        const double qSqr = dot3(sx, sy, sz);
        if (!(qSqr + EPS_ >= 4.0)) return hInv * orc_lut_value(values, qSqr);
    }
    return -1.0 / sqrt(dot3(rx, ry, rz));
}

/* BruteForceGravity::evalImpl with SymmetrizeSmoothingLengths (BruteForceGravity.h:112-131,
 * core/sph/kernel/Kernel.h:625-648): the two loops skip the target itself; G multiplies the sum. */
void orc_gravity_direct(size_t n, const double* x, const double* y, const double* z, const double* h,
                        const double* m, const double* grads, double G, const uint32_t* t, size_t nt, double* a) {
    if (!t) nt = n;
    for (size_t k = 0; k < nt; ++k) {
        const size_t i = t ? t[k] : k;
        double s[3] = { 0.0, 0.0, 0.0 };
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double g[3];
            orc_gravity_grad(grads, x[j] - x[i], y[j] - y[i], z[j] - z[i], 0.5 * (h[j] + h[i]), g);
            for (int c = 0; c < 3; ++c) s[c] += m[j] * g[c];
        }
        for (int c = 0; c < 3; ++c) a[c * nt + k] += G * s[c];
    }
}

/* BruteForceGravity::evalAcceleration (BruteForceGravity.h:91-101) */
void orc_gravity_points(size_t n, const double* x, const double* y, const double* z, const double* h,
                        const double* m, const double* grads, double G, const double* p, size_t np, double* a) {
    for (size_t k = 0; k < np; ++k) {
        double s[3] = { 0.0, 0.0, 0.0 };
        for (size_t j = 0; j < n; ++j) {
            double g[3];
            orc_gravity_grad(grads, x[j] - p[k], y[j] - p[np + k], z[j] - p[2 * np + k], h[j], g);
            for (int c = 0; c < 3; ++c) s[c] += m[j] * g[c];
        }
        for (int c = 0; c < 3; ++c) a[c * np + k] = G * s[c];
    }
}

/* BruteForceGravity::evalEnergy (BruteForceGravity.h:103-115), one thread */
double orc_gravity_energy(size_t n, const double* x, const double* y, const double* z, const double* h,
                          const double* m, const double* values, double G) {
    double e = 0.0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j) {
            if (i == j) continue;
            e += m[i] * m[j] * orc_gravity_value(values, x[i] - x[j], y[i] - y[j], z[i] - z[j], 0.5 * (h[i] + h[j]));
        }
    }
    return 0.5 * G * e;
}

/* computeMultipole<N> about the centre of mass (Moments.h:162-174) and computeReducedMultipole
 * (Moments.h:84-130): reducedFactor<2, 0..1> = 1, -1/3; reducedFactor<3, 0..1> = 1, -1/5. */
void orc_gravity_moments(size_t n, const double* x, const double* y, const double* z, const double* m,
                         double out[16]) {
    double M = 0.0, c[3] = { 0.0, 0.0, 0.0 };
    for (size_t i = 0; i < n; ++i) { /* CenterOfMass::evaluate (core/physics/Integrals.cpp:155-180) */
        M += m[i];
        c[0] += m[i] * x[i];
        c[1] += m[i] * y[i];
        c[2] += m[i] * z[i];
    }
    for (int k = 0; k < 3; ++k) c[k] /= M;
    double M2[3][3] = { { 0 } }, M3[3][3][3] = { { { 0 } } };
    for (size_t i = 0; i < n; ++i) {
        const double d[3] = { x[i] - c[0], y[i] - c[1], z[i] - c[2] };
        for (int a = 0; a < 3; ++a) {
            for (int b = 0; b < 3; ++b) {
                M2[a][b] += d[a] * d[b] * m[i];
                for (int e = 0; e < 3; ++e) M3[a][b][e] += d[a] * d[b] * d[e] * m[i];
            }
        }
    }
    const double T2 = M2[0][0] + M2[1][1] + M2[2][2];
    double q2[3][3], q3[3][3][3], T3[3];
    for (int k = 0; k < 3; ++k) T3[k] = M3[0][0][k] + M3[1][1][k] + M3[2][2][k];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) {
            q2[a][b] = M2[a][b] * 1.0 + (a == b ? T2 : 0.0) * (-1.0 / 3.0);
            for (int e = 0; e < 3; ++e) {
                const double perm = (a == b ? T3[e] : 0.0) + (a == e ? T3[b] : 0.0) + (b == e ? T3[a] : 0.0);
                q3[a][b][e] = M3[a][b][e] * 1.0 + perm * (-1.0 / 5.0);
            }
        }
    }
    out[0] = M;
    out[1] = c[0];
    out[2] = c[1];
    out[3] = c[2];
    out[4] = q2[0][0];
    out[5] = q2[0][1];
    out[6] = q2[0][2];
    out[7] = q2[1][1];
    out[8] = q2[1][2];
    out[9] = q3[0][0][0];
    out[10] = q3[0][0][1];
    out[11] = q3[0][0][2];
    out[12] = q3[0][1][1];
    out[13] = q3[0][1][2];
    out[14] = q3[1][1][1];
    out[15] = q3[1][1][2];
}

/* Full tensors from the independent components of the traceless q2 (xx, xy, xz, yy, yz) and q3
 * (xxx, xxy, xxz, xyy, xyz, yyy, yyz); the remaining components follow from the vanishing traces. */
static void expand_q2(const double* s, double q[3][3]) {
    q[0][0] = s[0];
    q[0][1] = q[1][0] = s[1];
    q[0][2] = q[2][0] = s[2];
    q[1][1] = s[3];
    q[1][2] = q[2][1] = s[4];
    q[2][2] = -s[0] - s[3];
}

static void expand_q3(const double* s, double q[3][3][3]) {
    const double xxx = s[0], xxy = s[1], xxz = s[2], xyy = s[3], xyz = s[4], yyy = s[5], yyz = s[6];
    const double xzz = -xxx - xyy, yzz = -xxy - yyy, zzz = -xxz - yyz;
    const double v[3][3][3] = { { { xxx, xxy, xxz }, { xxy, xyy, xyz }, { xxz, xyz, xzz } },
                                { { xxy, xyy, xyz }, { xyy, yyy, yyz }, { xyz, yyz, yzz } },
                                { { xxz, xyz, xzz }, { xyz, yyz, yzz }, { xzz, yzz, zzz } } };
    memcpy(q, v, sizeof(v));
}

/* evaluateGravity and computeMultipoleAcceleration (Moments.h:20-27,318-360): d = -dr = com - r,
 * gamma from computeGreenGamma, order-M term gamma[M+1] d Q0 + gamma[M] Q1 with
 * Q0 = q_M : d^M / M!, Q1 = q_M : d^(M-1) / (M-1)!; the dipole vanishes about the centre of mass. */
void orc_multipole_accel(const double mom[13], double drx, double dry, double drz, int order, double out[3]) {
    const double invDistSqr = 1.0 / dot3(drx, dry, drz);
    double gamma[5];
    gamma[0] = -sqrt(invDistSqr);
    for (int i = 1; i < 5; ++i) gamma[i] = -(2.0 * i - 1.0) * invDistSqr * gamma[i - 1];
    const double d[3] = { -drx, -dry, -drz };
    double a[3] = { 0.0, 0.0, 0.0 };
    if (order >= 3) {
        double q[3][3][3], Q1[3] = { 0.0, 0.0, 0.0 }, Q0 = 0.0;
        expand_q3(mom + 6, q);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                for (int k = 0; k < 3; ++k) Q1[i] += q[i][j][k] * d[j] * d[k];
            }
        }
        for (int i = 0; i < 3; ++i) Q0 += Q1[i] * d[i];
        Q0 *= 1.0 / 6.0;
        for (int i = 0; i < 3; ++i) a[i] += gamma[4] * d[i] * Q0 + gamma[3] * (0.5 * Q1[i]);
    }
    if (order >= 2) {
        double q[3][3], Q1[3] = { 0.0, 0.0, 0.0 }, Q0 = 0.0;
        expand_q2(mom + 1, q);
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) Q1[i] += q[i][j] * d[j];
        }
        for (int i = 0; i < 3; ++i) Q0 += Q1[i] * d[i];
        Q0 *= 0.5;
        for (int i = 0; i < 3; ++i) a[i] += gamma[3] * d[i] * Q0 + gamma[2] * Q1[i];
    }
    for (int i = 0; i < 3; ++i) a[i] += gamma[1] * d[i] * mom[0];
    for (int i = 0; i < 3; ++i) out[i] = a[i];
}

/* LeapFrog::stepParticles (TimeStepping.cpp:352-391) with HardSphereSolver::integrate
 * (NBodySolver.cpp:97-117): drift (all four lanes), zero the derivatives, gravity, zero the h lanes of
 * v and dv, kick, drift.  Collisions are not part of this restatement. */
void orc_leapfrog(size_t n, double* r, double* v, double* a, const double* m, const double* grads, double G,
                  double dt, int steps) {
    double* x = r;
    double* y = r + n;
    double* z = r + 2 * n;
    double* h = r + 3 * n;
    for (int s = 0; s < steps; ++s) {
        for (int c = 0; c < 4; ++c) {
            for (size_t i = 0; i < n; ++i) r[c * n + i] += v[c * n + i] * 0.5 * dt;
        }
        memset(a, 0, 3 * n * sizeof(double));
        orc_gravity_direct(n, x, y, z, h, m, grads, G, NULL, n, a);
        for (size_t i = 0; i < n; ++i) v[3 * n + i] = 0.0;
        for (int c = 0; c < 3; ++c) {
            for (size_t i = 0; i < n; ++i) v[c * n + i] += a[c * n + i] * dt;
        }
        for (int c = 0; c < 4; ++c) {
            for (size_t i = 0; i < n; ++i) r[c * n + i] += v[c * n + i] * 0.5 * dt;
        }
    }
}
