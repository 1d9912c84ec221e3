// Reverse-mode derivatives of the gas-dynamics step on gfx950: the vector-Jacobian product of the pair
// sums of PressureForce + StandardAV + ContinuityEquation (core/sph/equations/EquationTerm.cpp:41-110,
// 263-336; av/Standard.h:36-84), of the Tillotson equation of state (core/physics/Eos.cpp:198-238) and of
// the PredictorCorrector step (core/timestepping/TimeStepping.cpp:286-346) with a given time step.  No
// reference counterpart: OpenSPH has no derivatives of its runs.
//
// One reverse step, for the forward step X_n -> X_{n+1} (X = r, v, a, rho, drho, u, du; a, drho, du the
// derivatives of the last evaluation), given the adjoints L of X_{n+1}:
//   correction  r' = r* + (a' - a) dt^2 / 6, v' = v* + (a' - a) dt / 2, x' = x* + (x.' - x.) dt / 2
//   evaluation  (a', rho.', u.') = F(r*, v*, rho*, u*)   (the pair sums and the finalize)
//   prediction  r* = r + v dt + a dt^2 / 2, v* = v + a dt, x* = x + x. dt
// k_adj_seed turns L into the seeds of F's outputs, k_adj_pairs gathers F's vector-Jacobian product over
// each particle's neighbour list (both the particle's own pair terms and the terms its neighbours receive
// from it: the gather form, no atomics), k_adj_finish chains the equation of state (forward-mode duals),
// the prediction and the correction into the adjoints of X_n and the per-block partials of the gradient
// with respect to the Tillotson parameters a, b and A.
#include "osph_device.h"
#include "osph_internal.h"

namespace osph {
namespace {

// ---- forward-mode dual numbers (value and N directional derivatives) -------------------------------

template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(double x) {
    Dual<N> r;
    r.v = x;
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> dvar(double x, int k) {
    Dual<N> r = dconst<N>(x);
    r.d[k] = 1.0;
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double inv = 1.0 / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
    return r;
}

template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) { return a + dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) { return dconst<N>(a) + b; }
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) { return a - dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) { return dconst<N>(a) - b; }
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) { return a * dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) { return dconst<N>(a) * b; }
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) { return a / dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) { return dconst<N>(a) / b; }

template <int N>
__device__ __forceinline__ Dual<N> dexp(const Dual<N>& a) {
    Dual<N> r;
    r.v = exp(a.v);
    for (int k = 0; k < N; ++k) r.d[k] = r.v * a.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = sqrt(a.v);
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
    return r;
}

// min / max select one operand by value (the derivative of the selected branch)
template <int N> __device__ __forceinline__ Dual<N> dminv(const Dual<N>& a, double b) { return a.v < b ? a : dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> dmaxv(const Dual<N>& a, double b) { return a.v > b ? a : dconst<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> dmaxv(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v ? a : b; }

// TillotsonEos::evaluate (core/physics/Eos.cpp:198-238) over dual numbers, the same expressions in the
// same order as the forward kernel's tillotson(); a, b and A are the differentiated parameters.
constexpr int EOS_DIRS = 5;  // rho, u, a, b, A
using D5 = Dual<EOS_DIRS>;

__device__ void tillotson_dual(const osph_material& m, const D5& rho, const D5& u, const D5& a, const D5& b,
                               const D5& A, D5& p_out, D5& cs_out) {
#pragma clang fp contract(off)
    const D5 eta = rho / m.rho0;
    const D5 mu = eta - 1.0;
    const D5 denom = u / (m.u0 * eta * eta) + 1.0;
    const D5 pc = (a + b / denom) * rho * u + A * mu + m.B * mu * mu;
    D5 dpdu = a * rho + b * rho / (denom * denom);
    D5 dpdrho = a * u + b * u * (3.0 * denom - 2.0) / (denom * denom) + A / m.rho0 + 2.0 * m.B * mu / m.rho0;
    const D5 csc = dpdrho + dpdu * pc / (rho * rho);
    const D5 rhoExp = m.rho0 / rho - 1.0;
    const D5 betaExp = dexp(-dminv(m.beta * rhoExp, 70.0));
    const D5 alphaExp = dexp(-dminv(m.alpha * (rhoExp * rhoExp), 70.0));
    const D5 pe = a * rho * u + (b * rho * u / denom + A * mu * betaExp) * alphaExp;
    dpdu = a * rho + alphaExp * b * rho / (denom * denom);
    dpdrho = a * u + alphaExp * (b * u * (3.0 * denom - 2.0) / (denom * denom)) +
             alphaExp * (b * u * rho / denom) * m.rho0 * (2.0 * m.alpha * rhoExp) / (rho * rho) +
             alphaExp * A * betaExp * (1.0 / m.rho0 + m.rho0 * mu / (rho * rho) * (2.0 * m.alpha * rhoExp + m.beta));
    D5 cse = dpdrho + dpdu * pe / (rho * rho);
    cse = dmaxv(cse, 0.0);
    D5 p = pc, cs = csc;
    if (rho.v <= m.rho0 && u.v > m.ucv) {
        p = pe;
        cs = cse;
    } else if (rho.v <= m.rho0 && u.v > m.uiv && u.v <= m.ucv) {
        p = ((u - m.uiv) * pe + (m.ucv - u) * pc) / (m.ucv - m.uiv);
        cs = ((u - m.uiv) * cse + (m.ucv - u) * csc) / (m.ucv - m.uiv);
    }
    cs = dmaxv(cs, 0.25 * A / m.rho0);
    p_out = p;
    cs_out = dsqrt(cs);
}

// LutKernel<3>::gradImpl (core/sph/kernel/Kernel.h:129-144) and its slope in q^2 (piecewise constant).
__device__ __forceinline__ void lut_grad_slope(const double* __restrict__ grads, double q2, double& g, double& dg) {
    if (!(q2 < 4.0)) {
        g = 0.0;
        dg = 0.0;
        return;
    }
    const double fi = double(LUT_ENTRIES / 4) * q2;
    const uint32_t i1 = uint32_t(fi);
    const double ratio = fi - double(i1);
    g = grads[i1] * (1.0 - ratio) + grads[i1 + 1] * ratio;
    dg = (grads[i1 + 1] - grads[i1]) * double(LUT_ENTRIES / 4);
}

// ---- the pair terms' vector-Jacobian product -------------------------------------------------------

// What the adjoint pair kernel gathers for a neighbour (sorted order): position and h, velocity and
// mass, density, pressure, sound speed, and the seeds of its outputs (adjoints of its acceleration,
// velocity divergence and AV heating sums).
struct alignas(16) AdjRec {
    double x, y, z, h;
    double vx, vy, vz, m;
    double rho, p, cs, ld;
    double lax, lay, laz, lh;
};
static_assert(sizeof(AdjRec) == 128, "AdjRec must be 128 bytes");

struct PairGrad {
    double gr[3];  // d/d(r_o - r_q)
    double gw[3];  // d/d(v_q - v_o)
    double rho_o, rho_q, p_o, p_q, c_o, c_q;
};

// Gradient of L = la . da + ld ddivv + lh dheat, the pair (owner o, other q) contributions to o's
// acceleration, velocity divergence and AV heating (the forward k_pairs_fp64 terms without stress):
//   da = m_q (-(p_o/rho_o^2 + p_q/rho_q^2) - Pi) G,  ddivv = m_q/rho_o w.G,  dheat = -m_q Pi w.G / 2,
//   G = grad W(r) = r h^-5 g(|r|^2/h^2),  w = v_q - v_o,
//   Pi = (-alpha cbar mu + beta mu^2)/rhobar when (v_o - v_q).r < 0, mu = h (v_o - v_q).r/(|r|^2 + eps h^2).
// T is the arithmetic: double for the FP64 adjoint; float for the mixed one, whose separation r and
// velocity difference w are taken in FP64 and rounded once (a close pair's separation is a small
// difference of large coordinates), and whose results the caller adds to FP64 sums.
template <typename T>
__device__ void pair_adjoint(const AdjRec& o, const AdjRec& q, const double* __restrict__ lut, double alpha_d,
                             double beta_d, PairGrad& out) {
#pragma clang fp contract(off)
    const T alpha = T(alpha_d), beta = T(beta_d);
    const T r[3] = { T(o.x - q.x), T(o.y - q.y), T(o.z - q.z) };
    const T w[3] = { T(q.vx - o.vx), T(q.vy - o.vy), T(q.vz - o.vz) };
    const T oh = T(o.h), qh = T(q.h), orho = T(o.rho), qrho = T(q.rho), op = T(o.p), qp = T(q.p);
    const T hbar = T(0.5) * (oh + qh);
    const T hInv = T(1) / hbar;
    const T hInv2 = hInv * hInv;
    const T s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double gd, dgd;
    lut_grad_slope(lut, double(s * hInv2), gd, dgd);
    const T g = T(gd), dg = T(dgd);
    const T h5 = (hInv2 * hInv2) * hInv;
    const T K = h5 * g, Kp = h5 * dg * hInv2;  // G = K r, dK/ds = Kp
    const T G[3] = { K * r[0], K * r[1], K * r[2] };
    const T wg = w[0] * G[0] + w[1] * G[1] + w[2] * G[2];
    const T dvdr = -(w[0] * r[0] + w[1] * r[1] + w[2] * r[2]);
    const T Po = op / (orho * orho), Pq = qp / (qrho * qrho);
    const T mq = T(q.m);
    T Pi = T(0), mu = T(0), den = T(1), rhobar = T(1), csbar = T(0);
    const bool conv = dvdr < T(0);
    if (conv) {
        rhobar = T(0.5) * (orho + qrho);
        csbar = T(0.5) * (T(o.cs) + T(q.cs));
        den = s + T(1.e-2) * (hbar * hbar);
        mu = hbar * dvdr / den;
        Pi = (-alpha * csbar * mu + beta * (mu * mu)) / rhobar;
    }
    const T la[3] = { T(o.lax), T(o.lay), T(o.laz) };
    const T old = T(o.ld), olh = T(o.lh);
    const T A = la[0] * G[0] + la[1] * G[1] + la[2] * G[2];
    const T C = -(Po + Pq) - Pi;
    const T gPi = -mq * A - T(0.5) * mq * olh * wg;
    const T gP = -mq * A;
    const T cw = mq * old / orho - T(0.5) * mq * olh * Pi;  // coefficient of w in dL/dG and of G in dL/dw
    T gG[3], gw[3];
    for (int k = 0; k < 3; ++k) {
        gG[k] = mq * C * la[k] + cw * w[k];
        gw[k] = cw * G[k];
    }
    T gmu = T(0), gc = T(0), grb = T(0), gdvdr = T(0), gs = T(0);
    if (conv) {
        gmu = gPi * (-alpha * csbar + T(2) * beta * mu) / rhobar;
        gc = gPi * (-alpha * mu / rhobar);
        grb = gPi * (-Pi / rhobar);
        gdvdr = gmu * hbar / den;
        gs = gmu * (-mu / den);
    }
    const T gGr = gG[0] * r[0] + gG[1] * r[1] + gG[2] * r[2];
    for (int k = 0; k < 3; ++k) {
        out.gr[k] = double(K * gG[k] + T(2) * Kp * gGr * r[k] + T(2) * gs * r[k] - gdvdr * w[k]);
        out.gw[k] = double(gw[k] - gdvdr * r[k]);
    }
    const T ro3 = orho * orho * orho, rq3 = qrho * qrho * qrho;
    out.p_o = double(gP / (orho * orho));
    out.p_q = double(gP / (qrho * qrho));
    out.rho_o = double(-(mq * old / (orho * orho)) * wg - T(2) * op * gP / ro3 + T(0.5) * grb);
    out.rho_q = double(-T(2) * qp * gP / rq3 + T(0.5) * grb);
    out.c_o = double(T(0.5) * gc);
    out.c_q = double(T(0.5) * gc);
}

// The adjoint planes of one run (slot order): lambda of X (13), the gathered gradients (9).
struct AdjPlanes {
    double* l[13];  // r x y z, v x y z, a x y z, rho, drho, u, du
    double* g[9];   // r* x y z, v* x y z, rho*, p, cs
};

enum { LRX = 0, LVX = 3, LAX = 6, LRHO = 9, LDRHO = 10, LU = 11, LDU = 12 };
enum { GRX = 0, GVX = 3, GRHO = 6, GP = 7, GC = 8 };

// Seeds of F's outputs from the adjoints of X_{n+1}, and the adjoint record of sorted particle k.  The
// state holds the predicted X* with F evaluated at it (p, cs, divv).
__global__ void __launch_bounds__(BLOCK) k_adj_seed(DevState s, AdjPlanes ap, const uint32_t* __restrict__ order,
                                                    double dt, AdjRec* __restrict__ rec, size_t n) {
#pragma clang fp contract(off)
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = order[k];
    const double dt2 = 0.5 * (dt * dt);
    const double rho = s.d[RHO][i], p = s.d[P][i], divv = s.d[DIVV][i];
    double lA[3];
    for (int c = 0; c < 3; ++c) lA[c] = ap.l[LAX + c][i] + dt2 / 3.0 * ap.l[LRX + c][i] + 0.5 * dt * ap.l[LVX + c][i];
    const double lD = ap.l[LDRHO][i] + 0.5 * dt * ap.l[LRHO][i];
    const double lU = ap.l[LDU][i] + 0.5 * dt * ap.l[LU][i];
    // rho.' = -rho divv, u.' = heat - p/rho divv (ContinuityEquation and PressureForce finalize)
    AdjRec r;
    r.x = s.d[X][i];
    r.y = s.d[Y][i];
    r.z = s.d[Z][i];
    r.h = s.d[H][i];
    r.vx = s.d[VX][i];
    r.vy = s.d[VY][i];
    r.vz = s.d[VZ][i];
    r.m = s.d[M][i];
    r.rho = rho;
    r.p = p;
    r.cs = s.d[CS][i];
    r.ld = -rho * lD - p / rho * lU;
    r.lax = lA[0];
    r.lay = lA[1];
    r.laz = lA[2];
    r.lh = lU;
    rec[k] = r;
    ap.g[GRHO][i] = -divv * lD + p / (rho * rho) * divv * lU;
    ap.g[GP][i] = -(divv / rho) * lU;
}

// F's vector-Jacobian product gathered per particle over its neighbour list: its own pair terms (owner
// = itself) and the terms each neighbour receives from it (owner = the neighbour).
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_adj_pairs(AdjPlanes ap, const AdjRec* __restrict__ rec,
                                                     const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ list,
                                                     const uint32_t* __restrict__ count,
                                                     const double* __restrict__ lut, double alpha, double beta,
                                                     uint32_t kmax, size_t n) {
#pragma clang fp contract(off)
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const AdjRec me = rec[k];
    double gr[3] = { 0.0, 0.0, 0.0 }, gv[3] = { 0.0, 0.0, 0.0 };
    double grho = 0.0, gp = 0.0, gc = 0.0;
    const uint32_t cnt = count[k];
    const uint32_t listed = cnt < kmax ? cnt : kmax;
    const uint4* list4 = reinterpret_cast<const uint4*>(list);
    for (uint32_t c = 0; c < listed; c += 4) {
        const uint4 qd = list4[list_quad(k, c >> 2, kmax >> 2)];
        const uint32_t ent[4] = { qd.x, qd.y, qd.z, qd.w };
        for (uint32_t t = 0; t < 4 && c + t < listed; ++t) {
            const AdjRec o = rec[ent[t]];
            PairGrad a, b;
            pair_adjoint<T>(me, o, lut, alpha, beta, a);  // this particle's own terms
            pair_adjoint<T>(o, me, lut, alpha, beta, b);  // the neighbour's terms from this particle
            for (int d = 0; d < 3; ++d) {
                gr[d] += a.gr[d] - b.gr[d];  // r = r_me - r_o in a, r_o - r_me in b
                gv[d] += -a.gw[d] + b.gw[d];  // w = v_o - v_me in a, v_me - v_o in b
            }
            grho += a.rho_o + b.rho_q;
            gp += a.p_o + b.p_q;
            gc += a.c_o + b.c_q;
        }
    }
    const uint32_t i = order[k];
    for (int d = 0; d < 3; ++d) {
        ap.g[GRX + d][i] = gr[d];
        ap.g[GVX + d][i] = gv[d];
    }
    ap.g[GRHO][i] += grho;
    ap.g[GP][i] += gp;
    ap.g[GC][i] = gc;
}

// The equation of state, prediction and correction adjoints, per slot: the adjoints of X_{n+1} in ap.l
// become those of X_n.  Per-block partials of d/d(a, b, A) per material go to part[(3 mat + q) *
// nblocks + block]; *clamped counts particles whose predicted density or energy sits on a range bound.
__global__ void __launch_bounds__(BLOCK) k_adj_finish(DevState s, AdjPlanes ap, const osph_material* __restrict__ mats,
                                                      int nmat, double dt, double* __restrict__ part,
                                                      unsigned* __restrict__ clamped, size_t n) {
#pragma clang fp contract(off)
    __shared__ double red[ADJ_MAX_MATERIALS * 3][BLOCK / 64];
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double th[ADJ_MAX_MATERIALS * 3];
    for (int q = 0; q < ADJ_MAX_MATERIALS * 3; ++q) th[q] = 0.0;
    if (i < n) {
        const double dt2 = 0.5 * (dt * dt);
        const uint32_t mi = s.u[MAT][i];
        const osph_material m = mats[mi];
        const double rho = s.d[RHO][i], u = s.d[U][i];
        if (rho <= m.rho_lo || rho >= m.rho_hi || u <= m.u_lo || u >= m.u_hi) atomicAdd(clamped, 1u);
        D5 p, c;
        tillotson_dual(m, dvar<EOS_DIRS>(rho, 0), dvar<EOS_DIRS>(u, 1), dvar<EOS_DIRS>(m.a, 2),
                       dvar<EOS_DIRS>(m.b, 3), dvar<EOS_DIRS>(m.A, 4), p, c);
        const double gp = ap.g[GP][i], gc = ap.g[GC][i];
        const double grho = ap.g[GRHO][i] + gp * p.d[0] + gc * c.d[0];
        const double gu = gp * p.d[1] + gc * c.d[1];
        if (int(mi) < nmat) {
            for (int q = 0; q < 3; ++q) th[3 * mi + q] = gp * p.d[2 + q] + gc * c.d[2 + q];
        }
        for (int d = 0; d < 3; ++d) {
            const double lr = ap.l[LRX + d][i], lv = ap.l[LVX + d][i];
            const double lrs = lr + ap.g[GRX + d][i];  // adjoint of r*
            const double lvs = lv + ap.g[GVX + d][i];  // adjoint of v*
            const double la_direct = -(dt2 / 3.0) * lr - 0.5 * dt * lv;
            ap.l[LRX + d][i] = lrs;
            ap.l[LVX + d][i] = dt * lrs + lvs;
            ap.l[LAX + d][i] = la_direct + dt2 * lrs + dt * lvs;
        }
        const double lrho = ap.l[LRHO][i], lu = ap.l[LU][i];
        const double lrhos = lrho + grho, lus = lu + gu;
        ap.l[LRHO][i] = lrhos;
        ap.l[LDRHO][i] = -0.5 * dt * lrho + dt * lrhos;
        ap.l[LU][i] = lus;
        ap.l[LDU][i] = -0.5 * dt * lu + dt * lus;
    }
    // per-block partials of the parameter gradient (fixed order: lanes, then waves)
    for (int q = 0; q < ADJ_MAX_MATERIALS * 3; ++q) {
        double x = th[q];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[q][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < ADJ_MAX_MATERIALS * 3) {
        double x = 0.0;
        for (int w = 0; w < int(blockDim.x / 64); ++w) x += red[threadIdx.x][w];
        part[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = x;
    }
}

// Adds the reduced parameter gradient of this step to acc (one block per value, fixed order).
__global__ void __launch_bounds__(BLOCK) k_adj_theta(const double* __restrict__ part, uint32_t nblocks,
                                                     double* __restrict__ acc) {
    __shared__ double red[BLOCK / 64];
    const int q = blockIdx.x;
    double x = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x) x += part[size_t(q) * nblocks + b];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x / 64); ++w) x += red[w];
        acc[q] += x;
    }
}

// Copies of the 13 planes of X (r, v, a, rho, drho, u, du) between the state and a checkpoint.
constexpr int XPLANES[13] = { X, Y, Z, VX, VY, VZ, AX, AY, AZ, RHO, DRHO, U, DU };

struct XPtrs {
    double* p[13];
};

__global__ void __launch_bounds__(BLOCK) k_copy_x(XPtrs dst, XPtrs src, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 13; ++c) dst.p[c][i] = src.p[c][i];
}

// Host arrays (caller order, nc components per element) to and from adjoint planes (slot order).
__global__ void __launch_bounds__(BLOCK) k_scatter_planes(const double* __restrict__ src, int nc, XPtrs dst,
                                                          const uint32_t* __restrict__ slot_of, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = slot_of[i];
    for (int c = 0; c < nc; ++c) dst.p[c][k] = src ? src[i * nc + c] : 0.0;
}

__global__ void __launch_bounds__(BLOCK) k_gather_planes(double* __restrict__ dst, int nc, XPtrs src,
                                                         const uint32_t* __restrict__ slot_of, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = slot_of[i];
    for (int c = 0; c < nc; ++c) dst[i * nc + c] = src.p[c][k];
}

}  // namespace

// ---- launchers --------------------------------------------------------------------------------------

size_t adj_rec_bytes() { return sizeof(AdjRec); }

void launch_adj_step(const Launch& l, DevState s, double* const* lplanes, double* const* gplanes,
                     const uint32_t* order, const uint32_t* list, const uint32_t* count, uint32_t kmax,
                     const double* lut, double alpha, double beta, const osph_material* mats, int nmat, double dt,
                     void* rec, double* part, double* theta, unsigned* clamped, size_t n, bool fp32) {
    if (!n) return;
    AdjPlanes ap;
    for (int c = 0; c < 13; ++c) ap.l[c] = lplanes[c];
    for (int c = 0; c < 9; ++c) ap.g[c] = gplanes[c];
    AdjRec* r = static_cast<AdjRec*>(rec);
    k_adj_seed<<<grid_for(n), BLOCK, 0, l.stream>>>(s, ap, order, dt, r, n);
    if (fp32) {
        k_adj_pairs<float><<<grid_for(n), BLOCK, 0, l.stream>>>(ap, r, order, list, count, lut, alpha, beta, kmax, n);
    } else {
        k_adj_pairs<double><<<grid_for(n), BLOCK, 0, l.stream>>>(ap, r, order, list, count, lut, alpha, beta, kmax, n);
    }
    const unsigned g = grid_for(n);
    k_adj_finish<<<g, BLOCK, 0, l.stream>>>(s, ap, mats, nmat, dt, part, clamped, n);
    k_adj_theta<<<ADJ_MAX_MATERIALS * 3, BLOCK, 0, l.stream>>>(part, g, theta);
}

size_t adj_partials(size_t n) { return size_t(grid_for(n)) * ADJ_MAX_MATERIALS * 3; }

static XPtrs xptrs_state(DevState s) {
    XPtrs x;
    for (int c = 0; c < 13; ++c) x.p[c] = s.d[XPLANES[c]];
    return x;
}

static XPtrs xptrs_block(double* base, size_t cap) {
    XPtrs x;
    for (int c = 0; c < 13; ++c) x.p[c] = base + size_t(c) * cap;
    return x;
}

void launch_save_x(const Launch& l, DevState s, double* block, size_t cap, size_t n) {
    if (!n) return;
    k_copy_x<<<grid_for(n), BLOCK, 0, l.stream>>>(xptrs_block(block, cap), xptrs_state(s), n);
}

void launch_restore_x(const Launch& l, DevState s, const double* block, size_t cap, size_t n) {
    if (!n) return;
    k_copy_x<<<grid_for(n), BLOCK, 0, l.stream>>>(xptrs_state(s), xptrs_block(const_cast<double*>(block), cap), n);
}

void launch_scatter_planes(const Launch& l, const double* src, int nc, double* const* planes, const uint32_t* slot_of,
                           size_t n) {
    if (!n) return;
    XPtrs x{};
    for (int c = 0; c < nc; ++c) x.p[c] = planes[c];
    k_scatter_planes<<<grid_for(n), BLOCK, 0, l.stream>>>(src, nc, x, slot_of, n);
}

void launch_gather_planes(const Launch& l, double* dst, int nc, double* const* planes, const uint32_t* slot_of,
                          size_t n) {
    if (!n) return;
    XPtrs x{};
    for (int c = 0; c < nc; ++c) x.p[c] = planes[c];
    k_gather_planes<<<grid_for(n), BLOCK, 0, l.stream>>>(dst, nc, x, slot_of, n);
}

}  // namespace osph
