// HIP kernels of the OpenSPH per-timestep hot path for gfx950 (MI355X).
//
// Layout: particle state is SoA in HBM (one plane per component, storage order).  Every evaluation
// sorts the particles by the Morton code of a uniform cell grid (cell edge >= 2 h_max), packs the
// per-neighbour data into 128-byte records in that order, builds neighbour lists with the exact FP64
// pair test of SymmetricSolver<3>::loop, and evaluates all pair sums in gather form: each particle
// sums over its own neighbours, so every sum is written by one thread, in a fixed order, without atomics.
#include "osph_device.h"
#include "osph_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace osph {

namespace {


// ---- material physics (per particle; compiled without FMA contraction so that branch decisions
// ---- follow the reference's rounding) -------------------------------------------------------------

// TillotsonEos::evaluate (core/physics/Eos.cpp:198-238)
__device__ void tillotson(const osph_material& m, double rho, double u, double& p_out, double& cs_out) {
#pragma clang fp contract(off)
    const double eta = rho / m.rho0;
    const double mu = eta - 1.0;
    const double denom = u / (m.u0 * eta * eta) + 1.0;
    const double pc = (m.a + m.b / denom) * rho * u + m.A * mu + m.B * mu * mu;
    double dpdu = m.a * rho + m.b * rho / (denom * denom);
    double dpdrho = m.a * u + m.b * u * (3.0 * denom - 2.0) / (denom * denom) + m.A / m.rho0 + 2.0 * m.B * mu / m.rho0;
    const double csc = dpdrho + dpdu * pc / (rho * rho);
    const double rhoExp = m.rho0 / rho - 1.0;
    const double betaExp = exp(-dmin(m.beta * rhoExp, 70.0));
    const double alphaExp = exp(-dmin(m.alpha * (rhoExp * rhoExp), 70.0));
    const double pe = m.a * rho * u + (m.b * rho * u / denom + m.A * mu * betaExp) * alphaExp;
    dpdu = m.a * rho + alphaExp * m.b * rho / (denom * denom);
    dpdrho = m.a * u + alphaExp * (m.b * u * (3.0 * denom - 2.0) / (denom * denom)) +
             alphaExp * (m.b * u * rho / denom) * m.rho0 * (2.0 * m.alpha * rhoExp) / (rho * rho) +
             alphaExp * m.A * betaExp * (1.0 / m.rho0 + m.rho0 * mu / (rho * rho) * (2.0 * m.alpha * rhoExp + m.beta));
    double cse = dpdrho + dpdu * pe / (rho * rho);
    cse = dmax(cse, 0.0);
    double p = pc, cs = csc;
    if (rho <= m.rho0 && u > m.ucv) {
        p = pe;
        cs = cse;
    } else if (rho <= m.rho0 && u > m.uiv && u <= m.ucv) {
        p = ((u - m.uiv) * pe + (m.ucv - u) * pc) / (m.ucv - m.uiv);
        cs = ((u - m.uiv) * cse + (m.ucv - u) * csc) / (m.ucv - m.uiv);
    }
    cs = dmax(cs, 0.25 * m.A / m.rho0);
    p_out = p;
    cs_out = sqrt(cs);
}

// sqrtInv / sqrtApprox (core/math/MathUtils.h:44-62): the single-precision bit-trick estimate with one
// Newton step that norm(Vector) uses (core/objects/geometry/Vector.h:681-685).
__device__ double sqrt_approx(double f) {
#pragma clang fp contract(off)
    if (f == 0.0) return 0.0;
    const float x2 = float(f) * 0.5f;
    float y = float(f);
    int i = __float_as_int(y);
    i = 0x5f3759df - (i >> 1);
    y = __int_as_float(i);
    const float r = y * (1.5f - (x2 * y * y));
    return 1.0 / double(r);
}

// findEigenvalues (core/objects/geometry/SymmetricTensor.h:377-398); t = {xx, yy, zz, xy, xz, yz}
__device__ void eigenvalues(const double t[6], double out[3]) {
#pragma clang fp contract(off)
    const double d0 = t[0], d1 = t[1], d2 = t[2], o0 = t[3], o1 = t[4], o2 = t[5];
    const double mx0 = dmax(d0, o0), mx1 = dmax(d1, o1), mx2 = dmax(d2, o2);
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
    const double r = -I3 / (n * n * n);
    const double a = q - p * p / 3.0;
    const double b = (2.0 * (p * p * p) - 9.0 * p * q + 27.0 * r) / 27.0;
    const double aCub = (a * a * a) / 27.0;
    if (0.25 * b * b + aCub >= 0.0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double t1 = 2.0 * sqrt(-a / 3.0);
    const double phi = acos(-0.5 * b / sqrt(-aCub));
    const double p3 = p / 3.0;
    out[0] = (t1 * cos(phi / 3.0) - p3) * n;
    out[1] = (t1 * cos((phi + 2.0 * PI_D) / 3.0) - p3) * n;
    out[2] = (t1 * cos((phi + 4.0 * PI_D) / 3.0) - p3) * n;
}

// EosMaterial::initialize (core/sph/Materials.cpp:47-58), VonMisesRheology::initialize
// (core/physics/Rheology.cpp:36-83) or DruckerPragerRheology::initialize (core/physics/Rheology.cpp:119-181,
// without acoustic fluidization) and the h clamp of AdaptiveSmoothingLength::initialize
// (core/sph/equations/EquationTerm.cpp:356-364).
__device__ void material_init_at(DevState& s, const osph_material& m, const Params& prm, size_t i) {
#pragma clang fp contract(off)
    double p, cs;
    tillotson(m, s.d[RHO][i], s.d[U][i], p, cs);
    s.d[CS][i] = cs;
    if (m.yielding == 1) {
        const double dmg = prm.has_damage ? s.d[D][i] : 0.0;
        const double d = prm.has_damage ? dmg * dmg * dmg : 0.0;
        if (p < 0.0) p = (1.0 - d) * p;
        const double u = s.d[U][i];
        const double unorm = u / m.melt_energy;
        double Y = unorm < 1.e-5 ? m.elasticity_limit : m.elasticity_limit * dmax(1.0 - unorm, 0.0);
        Y = (1.0 - d) * Y;
        if (Y < EPS_D) {
            s.d[RED][i] = 0.0;
            s.d[SXX][i] = 0.0;
            s.d[SYY][i] = 0.0;
            s.d[SXY][i] = 0.0;
            s.d[SXZ][i] = 0.0;
            s.d[SYZ][i] = 0.0;
        } else {
            const double xx = s.d[SXX][i], yy = s.d[SYY][i], zz = -xx - yy;
            const double xy = s.d[SXY][i], xz = s.d[SXZ][i], yz = s.d[SYZ][i];
            const double dd = (xx * xx + yy * yy + zz * zz) + 2.0 * (xy * xy + xz * xz + yz * yz);
            const double J2 = 0.5 * dd + 1.e-15;
            const double red = dmin(Y / sqrt(3.0 * J2), 1.0);
            s.d[RED][i] = red;
            s.d[SXX][i] = xx * red;
            s.d[SYY][i] = yy * red;
            s.d[SXY][i] = xy * red;
            s.d[SXZ][i] = xz * red;
            s.d[SYZ][i] = yz * red;
        }
    } else if (m.yielding == 2) {
        const double d = prm.has_damage ? s.d[D][i] * s.d[D][i] * s.d[D][i] : 0.0;
        if (p < 0.0) p = (1.0 - d) * p;
        const double Y0 = m.cohesion, YM = m.elasticity_limit, mui = m.internal_friction, mud = m.dry_friction;
        const double Yi = Y0 + mui * p / (1.0 + mui * dmax(p, 0.0) / (YM - Y0));
        const double Yd = mud * p;
        double Y = Yd > Yi ? Yi : (1.0 - d) * Yi + d * Yd;
        const double u = s.d[U][i];
        Y = u < 1.e-5 * m.melt_energy ? Y : Y * dmax(1.0 - u / m.melt_energy, 0.0);
        if (Y < EPS_D) {
            s.d[RED][i] = 0.0;
            s.d[SXX][i] = 0.0;
            s.d[SYY][i] = 0.0;
            s.d[SXY][i] = 0.0;
            s.d[SXZ][i] = 0.0;
            s.d[SYZ][i] = 0.0;
        } else {
            const double xx = s.d[SXX][i], yy = s.d[SYY][i], zz = -xx - yy;
            const double xy = s.d[SXY][i], xz = s.d[SXZ][i], yz = s.d[SYZ][i];
            const double dd = (xx * xx + yy * yy + zz * zz) + 2.0 * (xy * xy + xz * xz + yz * yz);
            const double J2 = 0.5 * dd + EPS_D;
            const double red = dmin(Y / sqrt(J2), 1.0);
            s.d[RED][i] = red;
            s.d[SXX][i] = xx * red;
            s.d[SYY][i] = yy * red;
            s.d[SXY][i] = xy * red;
            s.d[SXZ][i] = xz * red;
            s.d[SYZ][i] = yz * red;
        }
    }
    s.d[P][i] = p;
    if (prm.adaptive_h) {
        s.d[H][i] = dmax(prm.h_lo, dmin(s.d[H][i], prm.h_hi));
    }
}

// ---- grid ------------------------------------------------------------------------------------------

// Block reduction of NV order-encoded values; is_min[k] selects min or max for value k.
template <int NV>
__device__ void block_reduce_ord(unsigned long long (&v)[NV], const bool (&is_min)[NV], unsigned long long* acc) {
    __shared__ unsigned long long red[BLOCK / 64][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const unsigned long long r = is_min[k] ? wave_min(v[k]) : wave_max(v[k]);
        if (lane == 0) red[wave][k] = r;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NV; k += blockDim.x) {
        unsigned long long r = red[0][k];
        for (int w = 1; w < BLOCK / 64; ++w) {
            const unsigned long long x = red[w][k];
            r = is_min[k] ? (x < r ? x : r) : (x > r ? x : r);
        }
        if (is_min[k]) {
            atomicMin(&acc[k], r);
        } else {
            atomicMax(&acc[k], r);
        }
    }
}

__device__ __forceinline__ void init_small_at(unsigned long long* a, int t) {
    bool is_min;
    if (t >= SMALL_CLASSES && t < SMALL_CLASSES + CLASS_STRIDE * MAX_CLASSES) {
        const int k = (t - SMALL_CLASSES) % CLASS_STRIDE;
        is_min = k < 3 || k == 7;
    } else if (t >= SMALL_CRITERIA && t < SMALL_CRITERIA + 4) {
        is_min = true;
    } else if (t >= SMALL_VEL && t < SMALL_VEL + 6) {
        is_min = t < SMALL_VEL + 3;
    } else {
        is_min = t < SMALL_BOUNDS + 3 || t == SMALL_BOUNDS + 7;
    }
    a[t] = is_min ? ~0ull : 0ull;
}

__global__ void __launch_bounds__(SMALL_WORDS) k_init_small(unsigned long long* a) { init_small_at(a, threadIdx.x); }

// acc[0..2] = min x, y, z; acc[3..5] = max x, y, z; acc[6] = max h; acc[7] = min h (order-encoded).
__global__ void __launch_bounds__(BLOCK) k_bounds(DevState s, unsigned long long* acc, size_t n) {
    unsigned long long v[8] = { ~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull, 0ull, ~0ull };
    const bool is_min[8] = { true, true, true, false, false, false, false, true };
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            const unsigned long long b = ord_bits(s.d[X + k][i]);
            v[k] = b < v[k] ? b : v[k];
            v[3 + k] = b > v[3 + k] ? b : v[3 + k];
        }
        const unsigned long long hb = ord_bits(s.d[H][i]);
        v[6] = hb > v[6] ? hb : v[6];
        v[7] = hb < v[7] ? hb : v[7];
    }
    block_reduce_ord<8>(v, is_min, acc);
}

__device__ __forceinline__ int h_class(double h, double hmin, int nclass) {
    if (nclass <= 1) return 0;
    const int c = int(floor(log2(h / hmin)));
    return c < 0 ? 0 : (c >= nclass ? nclass - 1 : c);
}

__device__ __forceinline__ int cell_coord(const GridClass& g, double x, int k) {
    return int(floor((x - g.lo[k]) * g.inv_cell));
}

__device__ __forceinline__ void cell_of(const GridClass& g, double x, double y, double z, int c[3]) {
    const double v[3] = { x, y, z };
    for (int k = 0; k < 3; ++k) {
        int ci = cell_coord(g, v[k], k);
        ci = ci < 0 ? 0 : (ci >= g.dim[k] ? g.dim[k] - 1 : ci);
        c[k] = ci;
    }
}

__device__ __forceinline__ uint32_t cell_key_of(const DevState& s, const Grid& g, size_t i) {
    const int cls = h_class(s.d[H][i], g.hmin, g.nclass);
    int c[3];
    cell_of(g.c[cls], s.d[X][i], s.d[Y][i], s.d[Z][i], c);
    return g.c[cls].base + row_key(uint32_t(c[0]), uint32_t(c[1]), uint32_t(c[2]), g.c[cls].bits);
}

__global__ void __launch_bounds__(BLOCK) k_cell_keys(DevState s, Grid g, uint32_t* keys, uint32_t* idx,
                                                     uint32_t* range, size_t clear, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t t = i; t < clear; t += size_t(gridDim.x) * blockDim.x) range[t] = 0u;
    if (i >= n) return;
    keys[i] = cell_key_of(s, g, i);
    idx[i] = uint32_t(i);
}

__global__ void __launch_bounds__(BLOCK) k_fill_u32(uint32_t* p, size_t count, uint32_t value) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x) {
        p[i] = value;
    }
}

// Cell index of a sort key (the key itself: keys are dense cell indices, monotonic in sorted order).
__device__ __forceinline__ uint32_t cell_index(const Grid&, uint32_t key) { return key; }

// First sorted position in [lo, hi) whose cell index is >= c (hi when none is).
__device__ __forceinline__ uint32_t first_at_least(const uint32_t* __restrict__ keys, const Grid& g, uint32_t c,
                                                   uint32_t lo, uint32_t hi) {
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (cell_index(g, keys[mid]) < c) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo;
}

// Row q of the grid (rows of all non-empty classes numbered in turn): its class, y and z.
__device__ __forceinline__ bool row_of(const Grid& g, size_t q, int& cls, uint32_t& cy, uint32_t& cz) {
    for (int c = 0; c < g.nclass; ++c) {
        const GridClass& G = g.c[c];
        if (G.empty) continue;
        const size_t nr = size_t(G.dim[1]) * size_t(G.dim[2]);
        if (q < nr) {
            cls = c;
            cy = uint32_t(q % size_t(G.dim[1]));
            cz = uint32_t(q / size_t(G.dim[1]));
            return true;
        }
        q -= nr;
    }
    return false;
}

__device__ __forceinline__ void row_range_at(const uint32_t* __restrict__ keys, const Grid& g,
                                             uint32_t* __restrict__ range, size_t q, uint32_t n) {
    int cls;
    uint32_t cy, cz;
    if (!row_of(g, q, cls, cy, cz)) return;
    const GridClass& G = g.c[cls];
    const uint32_t row = G.base + row_key(0u, cy, cz, G.bits);
    range[2 * q] = first_at_least(keys, g, row, 0u, n);
    range[2 * q + 1] = first_at_least(keys, g, row + (1u << G.bits), 0u, n);
}

// Row number (as row_of numbers them) of a cell key, from its class and the (y, z) of its Morton row key;
// *cls = the class.
__device__ __forceinline__ uint32_t row_index_of(const Grid& g, uint32_t key, int* cls) {
    uint32_t off = 0, q = 0;
    for (int c = 0; c < g.nclass; ++c) {
        const GridClass& G = g.c[c];
        if (G.empty) continue;
        if (key >= G.base) {
            const uint32_t m = (key - G.base) >> G.bits;
            q = off + compact2(m >> 1) * uint32_t(G.dim[1]) + compact2(m);
            *cls = c;
        }
        off += uint32_t(G.dim[1]) * uint32_t(G.dim[2]);
    }
    return q;
}

// The cell table of the rows that hold particles, in one pass over the sorted keys (rows are contiguous
// runs of them, cells contiguous runs inside a row).  The thread of sorted position k writes
// first[cell] = k for the cells after the previous particle's cell (from x = 0 when k starts its row)
// up to its own, and, when k ends its row, first[cell] = k + 1 for the cells after its own up to the
// row's spare slot x = dim_x; it also records the row's range of sorted positions.  Rows without
// particles keep the empty range [0, 0) that k_cell_keys left, and k_empty_rows fills their cells.
__global__ void __launch_bounds__(BLOCK) k_row_cells(const uint32_t* __restrict__ keys, Grid g,
                                                     uint32_t* __restrict__ range, uint32_t* __restrict__ first,
                                                     uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int cls = 0, cp = 0, cn = 0;
    const uint32_t key = keys[k];
    const uint32_t q = row_index_of(g, key, &cls);
    const GridClass& G = g.c[cls];
    // the cell's x in its row (the class's keys start at G.base, which need not be a multiple of 2^bits)
    const uint32_t xmask = (1u << G.bits) - 1u;
    const uint32_t x = (key - G.base) & xmask;
    const uint32_t rowbase = key - x;
    uint32_t xlo = 0;
    if (k > 0) {
        const uint32_t kp = keys[k - 1];
        if (row_index_of(g, kp, &cp) == q) xlo = ((kp - G.base) & xmask) + 1u;
    }
    if (xlo == 0) range[2 * size_t(q)] = k;
    for (uint32_t c = xlo; c <= x; ++c) first[rowbase + c] = k;
    const bool last = k + 1 == n || row_index_of(g, keys[k + 1], &cn) != q;
    if (last) {
        range[2 * size_t(q) + 1] = k + 1;
        for (uint32_t c = x + 1; c <= uint32_t(G.dim[0]); ++c) first[rowbase + c] = k + 1;
    }
}

// The cells of the rows without particles (empty range): equal entries, so that every x-run of them
// reads as an empty range of sorted positions.
__global__ void __launch_bounds__(BLOCK) k_empty_rows(Grid g, const uint32_t* __restrict__ range,
                                                      uint32_t* __restrict__ first, size_t rows, int xslots) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t q = t / size_t(xslots);
    const uint32_t x = uint32_t(t % size_t(xslots));
    if (q >= rows) return;
    if (range[2 * q] != range[2 * q + 1]) return;
    int cls;
    uint32_t cy, cz;
    if (!row_of(g, q, cls, cy, cz)) return;
    const GridClass& G = g.c[cls];
    if (x > uint32_t(G.dim[0])) return;
    first[G.base + row_key(x, cy, cz, G.bits)] = 0u;
}

// One thread per cell x in [0, dim_x] of a row: the binary search runs inside the row's own range.
__device__ __forceinline__ void cell_first_at(const uint32_t* __restrict__ keys, const Grid& g,
                                              const uint32_t* __restrict__ range, uint32_t* __restrict__ first,
                                              size_t t, size_t rows, int xslots) {
    const size_t q = t / size_t(xslots);
    const uint32_t x = uint32_t(t % size_t(xslots));
    if (q >= rows) return;
    int cls;
    uint32_t cy, cz;
    if (!row_of(g, q, cls, cy, cz)) return;
    const GridClass& G = g.c[cls];
    if (x > uint32_t(G.dim[0])) return;
    const uint32_t cell = G.base + row_key(x, cy, cz, G.bits);
    first[cell] = first_at_least(keys, g, cell, range[2 * q], range[2 * q + 1]);
}

// Pack the neighbour-side data of particle order[k] into rec[k].
// Origin-relative FP32 position of the screened neighbour test (one rounding of the FP64 offset).
__device__ __forceinline__ float4 screen_pos(const Screen& sc, double x, double y, double z, double h) {
    return make_float4(float(x - sc.o[0]), float(y - sc.o[1]), float(z - sc.o[2]), float(h));
}

__device__ __forceinline__ void pack_at(const DevState& s, const Params& prm, const uint32_t* __restrict__ order,
                                        PairRec* __restrict__ rec, double4* __restrict__ pos,
                                        float4* __restrict__ posf, const Screen& sc, uint32_t* __restrict__ group,
                                        size_t k) {
#pragma clang fp contract(off)
    const uint32_t i = order[k];
    PairRec r;
    r.x = s.d[X][i];
    r.y = s.d[Y][i];
    r.z = s.d[Z][i];
    r.h = s.d[H][i];
    r.vx = s.d[VX][i];
    r.vy = s.d[VY][i];
    r.vz = s.d[VZ][i];
    r.m = s.d[M][i];
    const double rho = s.d[RHO][i];
    const double rho2 = rho * rho;
    r.rho = rho;
    r.cs = s.d[CS][i];
    r.pr = s.d[P][i] / rho2;  // StandardForceDiscr: p / sqr(rho) (EquationTerm.cpp:20-23)
    if (prm.solid_stress) {
        r.sxx = s.d[SXX][i] / rho2;
        r.syy = s.d[SYY][i] / rho2;
        r.sxy = s.d[SXY][i] / rho2;
        r.sxz = s.d[SXZ][i] / rho2;
        r.syz = s.d[SYZ][i] / rho2;
    } else {
        r.sxx = r.syy = r.sxy = r.sxz = r.syz = 0.0;
    }
    rec[k] = r;
    pos[k] = make_double4(r.x, r.y, r.z, r.h);
    if (sc.on) posf[k] = screen_pos(sc, r.x, r.y, r.z, r.h);
    // DerivativeTemplate / AccelerationTemplate sum filter (DerivativeHelpers.h:132-139,262-269):
    // a pair enters the SUM_ONLY_UNDAMAGED sums iff both flags are equal and both reductions nonzero.
    uint32_t gid = 0;
    if (prm.sum_only_undamaged && prm.has_reduce) {
        gid = (s.d[RED][i] == 0.0) ? NO_GROUP : s.u[FLAG][i];
    }
    group[k] = gid;
}

__global__ void __launch_bounds__(BLOCK) k_pack(DevState s, Params prm, const uint32_t* __restrict__ order,
                                                PairRec* __restrict__ rec, double4* __restrict__ pos,
                                                float4* __restrict__ posf, Screen sc, uint32_t* __restrict__ group,
                                                size_t n) {
    const size_t k = size_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (k < n) pack_at(s, prm, order, rec, pos, posf, sc, group, k);
}

// The pair test of SymmetricSolver<3>::loop (core/sph/solvers/SymmetricSolver.cpp:126-131) evaluated
// bit for bit: getSqrLength sums the squares left to right without contraction and
// sqr(2 * 0.5 * (h_i + h_j)) == (h_i + h_j)^2 exactly.
__device__ __forceinline__ bool exact_pair(double xi, double yi, double zi, double hi, double xj, double yj, double zj,
                                           double hj) {
    const double dx = rn_sub(xi, xj), dy = rn_sub(yi, yj), dz = rn_sub(zi, zj);
    const double d2 = rn_add(rn_add(rn_mul(dx, dx), rn_mul(dy, dy)), rn_mul(dz, dz));
    const double hs = rn_add(hi, hj);
    return d2 < rn_mul(hs, hs);
}

// Candidates loaded per batch in the neighbour-list kernel: 4 FP32 positions (16 B each) in the screened
// test, 2 FP64 positions (32 B) in the plain one (measured at the 1M impact: 0.305 / 0.313 ms screened at
// 4 / 2, 0.400 / 0.363 ms plain).
template <bool SCREEN>
constexpr uint32_t LIST_BATCH = SCREEN ? 4u : 2u;

// One thread per particle (sorted position k).  For every smoothing-length class whose bounding box
// lies within h_i + h_max(class), the cells of that class within that radius are scanned with the exact
// pair test.  Neighbour positions are buffered four at a time in registers and stored as one 16-byte
// quad at list_quad(k, j, kmax / 4).
//
// With SCREEN, candidates are first tested in FP32 on origin-relative positions (16 bytes per candidate
// instead of 32).  The FP32 squared distance and radius carry a relative error below sc.band (bounded on
// the host from the domain extent and h_min), so d2 < R2 (1 - band) and d2 > R2 (1 + band) decide the
// exact test; only candidates inside the band are re-tested in FP64, and the pair set is unchanged.
template <bool SCREEN>
__device__ __forceinline__ void neighbor_list_at(const double4* __restrict__ pos, const float4* __restrict__ posf,
                                                 const Screen& sc, const Grid& g, const uint32_t* __restrict__ first,
                                                 uint32_t* __restrict__ list, uint32_t* __restrict__ count,
                                                 uint32_t kmax, uint32_t* overflow, size_t k) {
    const double4 me = pos[k];
    const float4 mf = SCREEN ? posf[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double p[3] = { me.x, me.y, me.z };
    uint4* list4 = reinterpret_cast<uint4*>(list);
    // the last four entries, oldest first (shifted in: no register indexed by the count)
    uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    uint32_t cnt = 0;
    auto append = [&](uint32_t k2) {
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = k2;
        ++cnt;
        if ((cnt & 3u) == 0 && cnt <= kmax) {
            list4[list_quad(k, (cnt >> 2) - 1, kmax >> 2)] = make_uint4(b0, b1, b2, b3);
        }
    };
    for (int cls = 0; cls < g.nclass; ++cls) {
        const GridClass& G = g.c[cls];
        if (G.empty) continue;
        const double R = me.w + G.hmax;  // h_i + h_j <= R for every member j
        double d2b = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double t = dmax(G.lo[a] - p[a], p[a] - G.hi[a]);
            if (t > 0.0) d2b += t * t;
        }
        if (d2b > R * R * (1.0 + 1e-6)) continue;
        // Rows of cells (cy, cz) whose yz-square lies farther than R hold no neighbour; in the others a
        // neighbour lies within the chord half-width w = sqrt(R^2 - d_yz^2) of the particle along x, and
        // the cells of that x-run are one contiguous range of sorted particles.  The planes cz within R,
        // the rows cy of a plane within R (half-width sqrt(R^2 - d_z^2) in y) and the x-runs are computed
        // directly, so every step of the walk yields a row.  The geometry is evaluated in FP32 in cell
        // units with R widened by 1e-3 cells: the FP32 errors (square roots of about 1 ulp included; below
        // 1e-4 cells for the at most a few thousand cells per axis) then only ever add candidates, and the
        // exact test decides.  The ranges of the next row are loaded before the current row's candidates
        // are tested, so the two loads overlap.
        const float pcx = float((p[0] - G.lo[0]) * G.inv_cell);
        const float pcy = float((p[1] - G.lo[1]) * G.inv_cell);
        const float pcz = float((p[2] - G.lo[2]) * G.inv_cell);
        const float Rc = float(R * G.inv_cell) * (1.0f + 1e-6f) + 1e-3f;
        const float R2c = Rc * Rc;
        const int zlast = min(G.dim[2] - 1, int(floorf(pcz + Rc)));
        int cz = max(0, int(floorf(pcz - Rc))) - 1;  // the plane before the first
        int cy = 0, ylast = -1;                       // the row and the last row of the plane within R
        float tz2 = 0.0f;
        uint32_t zkey = 0, ykey = 0;
        // the next row within R: its first and past-the-end sorted positions (an empty range where the
        // x-run misses the grid)
        auto next_row = [&](uint32_t& b, uint32_t& e) -> bool {
            if (++cy > ylast) {
                for (;;) {
                    if (++cz > zlast) return false;
                    const float fz = float(cz);
                    const float tz = fmaxf(fmaxf(fz - pcz, pcz - (fz + 1.0f)), 0.0f);
                    tz2 = tz * tz;
                    if (tz2 > R2c) continue;
                    const float wy = __builtin_amdgcn_sqrtf(R2c - tz2);
                    cy = max(0, int(floorf(pcy - wy)));
                    ylast = min(G.dim[1] - 1, int(floorf(pcy + wy)));
                    if (cy > ylast) continue;
                    zkey = spread2(uint32_t(cz)) << 1;
                    ykey = spread2(uint32_t(cy));
                    break;
                }
            } else {
                // spread2(cy + 1) from spread2(cy): adding one to the even-bit interleaved value with the odd
                // bits set carries through them
                ykey = ((ykey | 0xaaaaaaaau) + 1u) & 0x55555555u;
            }
            const float fy = float(cy);
            const float ty = fmaxf(fmaxf(fy - pcy, pcy - (fy + 1.0f)), 0.0f);
            const float w = __builtin_amdgcn_sqrtf(fmaxf(R2c - (tz2 + ty * ty), 0.0f));
            const int x0 = max(0, int(floorf(pcx - w)));
            const int x1 = min(G.dim[0] - 1, int(floorf(pcx + w)));
            if (x0 > x1) {
                b = e = 0u;
                return true;
            }
            const uint32_t row = G.base + ((ykey | zkey) << G.bits);
            b = first[row + uint32_t(x0)];
            e = first[row + uint32_t(x1) + 1u];
            return true;
        };
        uint32_t nb = 0, ne = 0;
        bool more = next_row(nb, ne);
        while (more) {
            const uint32_t b = nb, e = ne;
            more = next_row(nb, ne);
            // candidates in batches of LIST_BATCH<SCREEN>: all loads of a batch are issued before the first
            // test; a batch may read up to LIST_BATCH - 1 entries past e, which are never tested (e <= n and
            // the arrays hold POS_SLACK >= LIST_BATCH entries past n)
            for (uint32_t k0 = b; k0 < e; k0 += LIST_BATCH<SCREEN>) {
                if (SCREEN) {
                    float4 o[LIST_BATCH<true>];
#pragma unroll
                    for (uint32_t u = 0; u < LIST_BATCH<true>; ++u) o[u] = posf[k0 + u];  // POS_SLACK past n
#pragma unroll
                    for (uint32_t u = 0; u < LIST_BATCH<true>; ++u) {
                        const uint32_t k2 = k0 + u;
                        if (k2 >= e) break;
                        const float dx = mf.x - o[u].x, dy = mf.y - o[u].y, dz = mf.z - o[u].z;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        const float hs = mf.w + o[u].w;
                        const float rr = hs * hs;
                        bool hit = d2 < rr * sc.lo;
                        if (!hit && !(d2 > rr * sc.hi)) {
                            const double4 od = pos[k2];
                            hit = exact_pair(me.x, me.y, me.z, me.w, od.x, od.y, od.z, od.w);
                        }
                        if (hit && k2 != k) append(k2);
                    }
                } else {
                    double4 o[LIST_BATCH<false>];
#pragma unroll
                    for (uint32_t u = 0; u < LIST_BATCH<false>; ++u) o[u] = pos[k0 + u];  // POS_SLACK past n
#pragma unroll
                    for (uint32_t u = 0; u < LIST_BATCH<false>; ++u) {
                        const uint32_t k2 = k0 + u;
                        if (k2 >= e) break;
                        const bool hit = exact_pair(me.x, me.y, me.z, me.w, o[u].x, o[u].y, o[u].z, o[u].w);
                        if (hit && k2 != k) append(k2);
                    }
                }
            }
        }
    }
    if ((cnt & 3u) != 0 && cnt < kmax) {
        // the r = cnt % 4 entries of the last quad are b(4-r) .. b3; the rest of the quad is not read
        const uint32_t r = cnt & 3u;
        const uint4 q = r == 1 ? make_uint4(b3, 0u, 0u, 0u) : (r == 2 ? make_uint4(b2, b3, 0u, 0u) : make_uint4(b1, b2, b3, 0u));
        list4[list_quad(k, cnt >> 2, kmax >> 2)] = q;
    }
    count[k] = cnt;
    if (cnt > kmax) atomicOr(overflow, 1u);
}

template <bool SCREEN>
__global__ void __launch_bounds__(BLOCK) k_neighbor_lists(const double4* __restrict__ pos,
                                                          const float4* __restrict__ posf, Screen sc, Grid g,
                                                          const uint32_t* __restrict__ first,
                                                          uint32_t* __restrict__ list, uint32_t* __restrict__ count,
                                                          uint32_t kmax, uint32_t* overflow, size_t n) {
    const size_t k = size_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (k < n) neighbor_list_at<SCREEN>(pos, posf, sc, g, first, list, count, kmax, overflow, k);
}

// LutKernel<3>::gradImpl (core/sph/kernel/Kernel.h:129-144): linear interpolation in q^2.
__device__ __forceinline__ double lut_grad(const double* __restrict__ grads, double q2) {
    if (!(q2 < 4.0)) return 0.0;  // outside the support (also catches NaN, which must not become an index)
    const double fi = double(LUT_ENTRIES / 4) * q2;
    const uint32_t i1 = uint32_t(fi);
    const double ratio = fi - double(i1);
    return grads[i1] * (1.0 - ratio) + grads[i1 + 1] * ratio;
}

// Sum of v over the LPP adjacent lanes that share one particle (LPP = 1, 2 or 4), by DPP quad
// permutations (lane ^ 1, then lane ^ 2).  Every lane of the group gets the same, order-fixed result:
// (v0 + v1) + (v2 + v3) on lanes 0 and 1, (v2 + v3) + (v0 + v1) on lanes 2 and 3, equal since FP
// addition commutes.  All lanes of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

template <int LPP>
__device__ __forceinline__ double group_sum(double v) {
    static_assert(LPP == 1 || LPP == 2 || LPP == 4, "lanes per particle must be 1, 2 or 4");
    if constexpr (LPP >= 2) v += dpp_move<0xb1>(v);  // quad_perm [1, 0, 3, 2]
    if constexpr (LPP >= 4) v += dpp_move<0x4e>(v);  // quad_perm [2, 3, 0, 1]
    return v;
}

// Particle handled by this lane and the lane's rank among the LPP lanes of that particle.  Lanes past
// the end keep running (with an empty list) so that the group sums see every lane active.
template <int LPP>
struct PairLane {
    size_t k;
    uint32_t sub;
    bool valid;
    __device__ __forceinline__ PairLane(size_t t, size_t n) {
        k = t / LPP;
        sub = uint32_t(t % LPP);
        valid = k < n;
        if (!valid) k = n - 1;
    }
};

// The pair sums of particle i that the finalize, the corrector and the criteria use: acceleration, du,
// velocity divergence and gradient (xx, yy, zz, xy, xz, yz).  The fused pair kernels hand them over in
// registers; the separate passes read them from the state (sums_of).
struct PairSums {
    double a[3], du, divv, g[6];
};
__device__ __forceinline__ PairSums sums_of(const DevState& s, bool solid, size_t i) {
    PairSums ps;
    for (int k = 0; k < 3; ++k) ps.a[k] = s.d[AX + k][i];
    ps.du = s.d[DU][i];
    ps.divv = s.d[DIVV][i];
    for (int k = 0; k < 6; ++k) ps.g[k] = solid ? s.d[GXX + k][i] : 0.0;
    return ps;
}
__device__ void finalize_at(DevState& s, const osph_material& m, const Params& prm, size_t i, const PairSums& ps);
__device__ void correct_at(DevState& s, const osph_material& m, const Params& prm, double dt, size_t i,
                           const double a[3]);
__device__ void criteria_values(const DevState& s, const osph_material& m, const Params& prm, size_t i,
                                const double a[3], double divv, double out[4]);
__device__ void reduce_criteria(double v[4], unsigned long long* mins);


// How the pair kernels gather neighbour records (for_each_neighbor): 0 directly, 1 through the LDS stage,
// 2 through the LDS stage with the next step's loads issued before the current step's compute, 3 directly
// with the next neighbour's record loaded before the current one's terms.  For the 64-B mixed record at
// the 1M impact, in one A/B on one box: 0.432 ms (0), 0.417 ms (3); the 80-B record it replaced ran at
// 0.415 ms directly.
constexpr int PAIRS64_STAGE = 1;
constexpr int PAIRSMX_STAGE = 3;

// Per-wavefront LDS staging of neighbour records: 64 records of CH 16-byte chunks each, at an odd stride
// of chunks (CH | 1) so that the lanes' 16-byte reads of their own record spread over all banks.
template <typename Rec>
struct StageOf {
    static constexpr uint32_t CH = sizeof(Rec) / 16;
    static constexpr uint32_t STRIDE = CH | 1u;
    static constexpr uint32_t WORDS = 64 * STRIDE;  // uint4 per wavefront
    static_assert(sizeof(Rec) % 16 == 0, "records are whole 16-byte chunks");
};

// The neighbour list of one particle, split among its LPP lanes by quads (4 entries, 16 B): lane `sub`
// takes quads sub, sub + LPP, ..., in order, the next quad loaded one quad ahead.  f(record, entry) runs
// for each listed entry of the lane.
//
// STAGED (1, 2): the records are fetched cooperatively.  In each step the wavefront needs one record per lane;
// pass i of the CH load passes has lane L fetch chunk c of the record of lane r, (r, c) = divmod(64 i +
// L, CH), into the wavefront's LDS stage, and each lane then reads its own record back.  Consecutive
// lanes fetch consecutive chunks of one record, so a load instruction touches ~64 / CH records' cache
// lines instead of one line per lane: CH times fewer line accesses in the texture path, for a longer
// dependency chain (shuffle, load, LDS write, LDS read).  It pays for the 128-B FP64 record (1.16 ->
// 1.06 ms at the impact) and not for the (then 80-B) mixed one (0.45 -> 0.51 ms), which gathers directly.
// With STAGED == 2 the next step's loads are issued before the current step's compute.
// Every lane of the wavefront must call this (lanes without entries load their own record as a
// stand-in, and run no f).
template <int LPP, int STAGED, typename Rec, typename F>
__device__ __forceinline__ void for_each_neighbor(const Rec* __restrict__ rec, const uint32_t* __restrict__ list,
                                                  size_t k, uint32_t quads, uint32_t listed, uint32_t sub,
                                                  uint4* __restrict__ stage, F&& f) {
    using S = StageOf<Rec>;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nq = (listed + 3u) >> 2;
    const uint32_t myq = sub < nq ? (nq - sub + LPP - 1) / LPP : 0u;  // quads of this lane
    uint32_t total = 0;  // entries of this lane
    if (myq > 0) {
        const uint32_t tail = listed - 4u * (sub + (myq - 1) * LPP);
        total = 4u * (myq - 1) + (tail < 4u ? tail : 4u);
    }
    const uint4* list4 = reinterpret_cast<const uint4*>(list);
    auto row = [&](uint32_t j) { return list4[list_quad(k, j, quads)]; };
    if constexpr (STAGED == 0 || STAGED == 3) {
        if (myq == 0) return;
        uint4 quad = row(sub);
        uint4 nquad = myq > 1 ? row(sub + LPP) : quad;
        auto entry_of = [&](uint32_t t) -> uint32_t {
            const uint32_t sel = t & 3u;
            if (sel == 0 && t > 0) {
                quad = nquad;
                const uint32_t j = (t >> 2) + 1;
                nquad = j < myq ? row(sub + j * LPP) : quad;
            }
            return sel == 0 ? quad.x : (sel == 1 ? quad.y : (sel == 2 ? quad.z : quad.w));
        };
        if constexpr (STAGED == 0) {
            for (uint32_t t = 0; t < total; ++t) {
                const uint32_t e = entry_of(t);
                const Rec o = rec[e];
                f(o, e);
            }
        } else {
            // the next neighbour's record is loaded before the current one's terms are evaluated
            uint32_t e = entry_of(0);
            Rec o = rec[e];
            for (uint32_t t = 0; t < total; ++t) {
                const uint32_t en = t + 1 < total ? entry_of(t + 1) : e;
                const Rec on = rec[en];
                f(o, e);
                o = on;
                e = en;
            }
        }
        return;
    }
    // (in the staged loops the step index is uniform across the wavefront, so the entry is taken from
    // rotating registers rather than selected by index, which the compiler would place in scratch)
    const uint4 first = myq > 0 ? row(sub) : make_uint4(0, 0, 0, 0);
    uint32_t q0 = first.x, q1 = first.y, q2 = first.z, q3 = first.w;  // entries of the current quad, next first
    uint4 nquad = myq > 1 ? row(sub + LPP) : first;
    // the entry of step t for this lane (its own index as a stand-in when it has none); steps in order
    auto entry = [&](uint32_t t) -> uint32_t {
        const bool act = t < total;
        if (act && (t & 3u) == 0 && t > 0) {
            q0 = nquad.x;
            q1 = nquad.y;
            q2 = nquad.z;
            q3 = nquad.w;
            const uint32_t j = (t >> 2) + 1;
            if (j < myq) nquad = row(sub + j * LPP);
        }
        const uint32_t e = act ? q0 : uint32_t(k);
        q0 = q1;
        q1 = q2;
        q2 = q3;
        return e;
    };
    const uint4* src = reinterpret_cast<const uint4*>(rec);
    // pass i: this lane's chunk of the records of every lane's entry e
    uint4 v[S::CH];
    auto fetch = [&](uint32_t e) {
#pragma unroll
        for (uint32_t i = 0; i < S::CH; ++i) {
            const uint32_t q = i * 64u + lane;
            const uint32_t r = q / S::CH, c = q - r * S::CH;
            const uint32_t er = uint32_t(__shfl(int(e), int(r), 64));
            v[i] = src[size_t(er) * S::CH + c];
        }
    };
    auto put = [&]() {
#pragma unroll
        for (uint32_t i = 0; i < S::CH; ++i) {
            const uint32_t q = i * 64u + lane;
            const uint32_t r = q / S::CH, c = q - r * S::CH;
            stage[r * S::STRIDE + c] = v[i];
        }
    };
    uint32_t e = entry(0);
    if constexpr (STAGED == 2) fetch(e);
    for (uint32_t t = 0; __any(t < total); ++t) {
        const bool act = t < total;
        const uint32_t ecur = e;
        wave_sync();  // the previous step's reads of the stage come first
        if constexpr (STAGED == 2) {
            put();
            // the next step's loads leave before this step's compute (pipelined)
            if (__any(t + 1 < total)) {
                e = entry(t + 1);
                fetch(e);
            }
        } else {
            fetch(ecur);
            put();
            if (__any(t + 1 < total)) e = entry(t + 1);
        }
        wave_sync();
        union {
            uint4 v[S::CH];
            Rec r;
        } o;
#pragma unroll
        for (uint32_t c = 0; c < S::CH; ++c) o.v[c] = stage[lane * S::STRIDE + c];
        if (act) f(o.r, ecur);
    }
}

// FP64 pair sums in gather form.  For particle i (sorted position k) and each neighbour j, the kernel
// gradient of SymmetrizeSmoothingLengths<LutKernel<3>> (Kernel.h:32-36,640-643) at hbar, then the
// contributions every derivative of the term set makes to particle i:
//   VelocityDivergence<CenterDensityDiscr>   DerivativeHelpers.h:355-371
//   PressureGradient<StandardForceDiscr>     EquationTerm.cpp:61-66 via AccelerationTemplate :229-242
//   VelocityGradient (SUM_ONLY_UNDAMAGED)    DerivativeHelpers.h:382-388, SymmetricTensor.h:362-368
//   StressDivergence (SUM_ONLY_UNDAMAGED)    EquationTerm.cpp:134-139, TracelessTensor.h:151-155
//   StandardAV                               av/Standard.h:62-83
//   NeighborCountTerm                        HelperTerms.h:38-46
// The accumulated buffers replace the storage buffers (Accumulated::store, Accumulated.cpp:87-100).
// LPP lanes share one particle: each sums its share of the list, then the partial sums are added across
// the group (fewer particles in flight per XCD, so their neighbour records stay in its L2).
template <int LPP, bool FUSED>
__device__ __forceinline__ void pairs_fp64_body(size_t t, DevState s, const Params& prm,
                                                const PairRec* __restrict__ rec, const uint32_t* __restrict__ group,
                                                const uint32_t* __restrict__ order, const uint32_t* __restrict__ list,
                                                const uint32_t* __restrict__ count, const double* __restrict__ lut,
                                                uint32_t kmax, size_t n, const FuseArgs& fa) {
    const PairLane<LPP> ln(t, n);
    const size_t k = ln.k;
    const PairRec me = rec[k];
    const uint32_t gi = group[k];
    const double rhoInv = 1.0 / me.rho;
    double ax = 0.0, ay = 0.0, az = 0.0, du = 0.0, divv = 0.0;
    double gxx = 0.0, gyy = 0.0, gzz = 0.0, gxy = 0.0, gxz = 0.0, gyz = 0.0;
    const uint32_t cnt = ln.valid ? count[k] : 0u;
    const uint32_t listed = cnt < kmax ? cnt : kmax;
    const bool solid = prm.solid_stress != 0;
    __shared__ uint4 stage[PAIRS64_STAGE ? (BLOCK / 64) * StageOf<PairRec>::WORDS : 1];
    uint4* my_stage = PAIRS64_STAGE ? stage + (threadIdx.x >> 6) * StageOf<PairRec>::WORDS : stage;
    for_each_neighbor<LPP, PAIRS64_STAGE>(rec, list, k, kmax >> 2, listed, ln.sub, my_stage, [&](const PairRec& o, uint32_t k2) {
        const double rx = me.x - o.x, ry = me.y - o.y, rz = me.z - o.z;
        const double hbar = 0.5 * (me.h + o.h);
        const double hInv = 1.0 / hbar;
        const double hInv2 = hInv * hInv;
        const double d2 = rx * rx + ry * ry + rz * rz;
        const double gk = lut_grad(lut, d2 * hInv2);
        const double h5g = (hInv2 * hInv2) * hInv * gk;
        const double gx = rx * h5g, gy = ry * h5g, gz = rz * h5g;
        const double wx = o.vx - me.vx, wy = o.vy - me.vy, wz = o.vz - me.vz;  // v_j - v_i
        const double mj = o.m;
        const double mjr = mj * rhoInv;
        // velocity divergence
        divv += mjr * (wx * gx + wy * gy + wz * gz);
        // pressure gradient
        {
            const double coef = me.pr + o.pr;
            ax -= mj * (coef * gx);
            ay -= mj * (coef * gy);
            az -= mj * (coef * gz);
        }
        if (solid && gi == group[k2] && gi != NO_GROUP) {
            gxx += mjr * (wx * gx);
            gyy += mjr * (wy * gy);
            gzz += mjr * (wz * gz);
            gxy += mjr * (0.5 * (wx * gy + wy * gx));
            gxz += mjr * (0.5 * (wx * gz + wz * gx));
            gyz += mjr * (0.5 * (wy * gz + wz * gy));
            const double xx = me.sxx + o.sxx, yy = me.syy + o.syy, xy = me.sxy + o.sxy;
            const double xz = me.sxz + o.sxz, yz = me.syz + o.syz;
            ax += mj * (xx * gx + xy * gy + xz * gz);
            ay += mj * (xy * gx + yy * gy + yz * gz);
            az += mj * (xz * gx + yz * gy + (-xx - yy) * gz);
        }
        // artificial viscosity, only in convergent flow
        const double dvdr = -(wx * rx + wy * ry + wz * rz);  // (v_i - v_j) . (r_i - r_j)
        if (dvdr < 0.0) {
            const double rhobar = 0.5 * (me.rho + o.rho);
            const double csbar = 0.5 * (me.cs + o.cs);
            const double mu = hbar * dvdr / (d2 + 1.e-2 * (hbar * hbar));
            const double av = (-prm.av_alpha * csbar * mu + prm.av_beta * (mu * mu)) / rhobar;
            ax -= mj * (av * gx);
            ay -= mj * (av * gy);
            az -= mj * (av * gz);
            du += mj * (0.5 * av * -(wx * gx + wy * gy + wz * gz));
        }
    });
    ax = group_sum<LPP>(ax);
    ay = group_sum<LPP>(ay);
    az = group_sum<LPP>(az);
    du = group_sum<LPP>(du);
    divv = group_sum<LPP>(divv);
    if (solid) {
        gxx = group_sum<LPP>(gxx);
        gyy = group_sum<LPP>(gyy);
        gzz = group_sum<LPP>(gzz);
        gxy = group_sum<LPP>(gxy);
        gxz = group_sum<LPP>(gxz);
        gyz = group_sum<LPP>(gyz);
    }
    double cv[4] = { INFINITY, INFINITY, INFINITY, INFINITY };
    if (ln.valid && ln.sub == 0) {
        const uint32_t i = order[k];
        if (prm.add_accel) {
            ax += s.d[AX][i];
            ay += s.d[AY][i];
            az += s.d[AZ][i];
        }
        s.d[AX][i] = ax;
        s.d[AY][i] = ay;
        s.d[AZ][i] = az;
        s.d[DU][i] = du;
        s.d[DIVV][i] = divv;
        if (solid) {
            s.d[GXX][i] = gxx;
            s.d[GYY][i] = gyy;
            s.d[GZZ][i] = gzz;
            s.d[GXY][i] = gxy;
            s.d[GXZ][i] = gxz;
            s.d[GYZ][i] = gyz;
        }
        s.u[NCNT][i] = cnt;
        if (FUSED && !*fa.overflow) {
            const osph_material m = fa.mats[s.u[MAT][i]];
            const PairSums ps{ { ax, ay, az }, du, divv, { gxx, gyy, gzz, gxy, gxz, gyz } };
            finalize_at(s, m, prm, i, ps);
            correct_at(s, m, prm, fa.dt, i, ps.a);
            criteria_values(s, m, prm, i, ps.a, ps.divv, cv);
        }
    }
    if (FUSED && !*fa.overflow) reduce_criteria(cv, fa.mins);
}

template <int LPP, bool FUSED>
__global__ void __launch_bounds__(BLOCK) k_pairs_fp64(DevState s, Params prm, const PairRec* __restrict__ rec,
                                                      const uint32_t* __restrict__ group,
                                                      const uint32_t* __restrict__ order,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count,
                                                      const double* __restrict__ lut, uint32_t kmax, size_t n,
                                                      FuseArgs fa) {
    pairs_fp64_body<LPP, FUSED>(size_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x, s, prm, rec,
                                group, order, list, count, lut, kmax, n, fa);
}

// ---- mixed precision ---------------------------------------------------------------------------------

// Coordinate v on the fixed-point axis from lo with step 1 / inv (a power of two), rounded to nearest.
__device__ __forceinline__ uint32_t fixed_coord(double v, double lo, double inv) {
    const double t = floor((v - lo) * inv + 0.5);
    return uint32_t(t < 0.0 ? 0.0 : (t > 4294967295.0 ? 4294967295.0 : t));
}

// Packs the mixed-precision neighbour record of particle order[k] (and the exact FP64 position used by
// the neighbour-list build): position and velocity as 32-bit fixed-point coordinates, cs rounded to 20
// bits with the group code in the low 4.
__device__ __forceinline__ void pack_mixed_at(const DevState& s, const Params& prm, const Fixed& F,
                                              const uint32_t* __restrict__ order, PairRecF* __restrict__ rec,
                                              double4* __restrict__ pos, float4* __restrict__ posf, const Screen& sc,
                                              uint32_t* __restrict__ group, size_t k) {
    const uint32_t i = order[k];
    const double x[3] = { s.d[X][i], s.d[Y][i], s.d[Z][i] };
    const double h = s.d[H][i];
    PairRecF r;
    r.x = fixed_coord(x[0], F.lo[0], F.inv_step);
    r.y = fixed_coord(x[1], F.lo[1], F.inv_step);
    r.z = fixed_coord(x[2], F.lo[2], F.inv_step);
    r.vx = fixed_coord(s.d[VX][i], F.vlo[0], F.inv_vstep);
    r.vy = fixed_coord(s.d[VY][i], F.vlo[1], F.inv_vstep);
    r.vz = fixed_coord(s.d[VZ][i], F.vlo[2], F.inv_vstep);
    r.h = float(h);
    r.m = float(s.d[M][i]);
    const double rho = s.d[RHO][i];
    const double rho2 = rho * rho;
    r.rho = float(rho);
    r.pr = float(s.d[P][i] / rho2);
    if (prm.solid_stress) {
        r.sxx = float(s.d[SXX][i] / rho2);
        r.syy = float(s.d[SYY][i] / rho2);
        r.sxy = float(s.d[SXY][i] / rho2);
        r.sxz = float(s.d[SXZ][i] / rho2);
        r.syz = float(s.d[SYZ][i] / rho2);
    } else {
        r.sxx = r.syy = r.sxy = r.sxz = r.syz = 0.f;
    }
    uint32_t gid = 0;
    if (prm.sum_only_undamaged && prm.has_reduce) {
        gid = (s.d[RED][i] == 0.0) ? NO_GROUP : s.u[FLAG][i];
    }
    // cs >= 0: rounded to a multiple of 16 units in the last place (ties up), then the code
    const uint32_t cb = __float_as_uint(float(s.d[CS][i]));
    r.cs = __uint_as_float(((cb + 8u) & ~15u) | group_code(gid));
    rec[k] = r;
    pos[k] = make_double4(x[0], x[1], x[2], h);
    if (sc.on) posf[k] = screen_pos(sc, x[0], x[1], x[2], h);
    group[k] = gid;
}

__global__ void __launch_bounds__(BLOCK) k_pack_mixed(DevState s, Params prm, Fixed F,
                                                      const uint32_t* __restrict__ order, PairRecF* __restrict__ rec,
                                                      double4* __restrict__ pos, float4* __restrict__ posf,
                                                      Screen sc, uint32_t* __restrict__ group,
                                                      size_t n) {
    const size_t k = size_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (k < n) pack_mixed_at(s, prm, F, order, rec, pos, posf, sc, group, k);
}

// M4 gradient factor grad W / q of CubicSpline<3> (core/sph/kernel/Kernel.h:172-187) in closed form,
// FP32: for q < 1 the bracket divided by q simplifies to (2.25 q - 3) / pi (finite at q = 0, where the
// reference returns -3 / pi); for 1 <= q < 2 it is -0.75 (2 - q)^2 / (pi q).  The square root and the
// reciprocal are the hardware instructions (about 1 ulp).
__device__ __forceinline__ float m4_grad_f(float q2) {
    constexpr float INV_PI = 0.318309886183790671538f;
    if (!(q2 < 4.0f)) return 0.0f;
    const float q = __builtin_amdgcn_sqrtf(q2);
    if (q < 1.0f) return INV_PI * (2.25f * q - 3.0f);
    const float t = 2.0f - q;
    return (-0.75f * INV_PI) * (t * t) * __builtin_amdgcn_rcpf(q);
}

// FP64 accumulator of FP32 terms: each term is widened exactly and summed in FP64 (on gfx950 an FP64 add
// issues at the rate of an unpacked FP32 add, so this costs two instructions per term).
struct Acc {
    double sum = 0.0;
    __device__ __forceinline__ void add(float v) { sum += double(v); }
};

// Mixed-precision pair sums: the same terms and gather form as k_pairs_fp64, evaluated in FP32 on the
// records' fixed-point positions and velocities and accumulated in FP64; results are written in FP64.  A
// difference of fixed-point coordinates is an exact int32 (each frame spans less than 2^31 steps), rounded
// once by the conversion to FP32 and scaled by the power-of-two step exactly, so it keeps FP32 precision
// whatever the magnitudes of the absolute values.  Divisions are hardware reciprocals
// (about 1 ulp) times the numerator.  Groups (SUM_ONLY_UNDAMAGED) compare by the records' 4-bit codes,
// and through the group array for groups beyond the codes' range.
template <int LPP, bool FUSED>
__device__ __forceinline__ void pairs_mixed_body(size_t t, DevState s, const Params& prm, const Fixed& F,
                                                 const PairRecF* __restrict__ rec, const uint32_t* __restrict__ group,
                                                 const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                                                 uint32_t kmax, size_t n, const FuseArgs& fa) {
    const PairLane<LPP> ln(t, n);
    const size_t k = ln.k;
    const PairRecF me = rec[k];
    const uint32_t gi = group[k];
    const uint32_t gci = __float_as_uint(me.cs) & 15u;
    const float csi = __uint_as_float(__float_as_uint(me.cs) & ~15u);
    const float rhoInv = 1.0f / me.rho;
    const float alpha = float(prm.av_alpha), beta = float(prm.av_beta);
    Acc ax, ay, az, du, divv, gxx, gyy, gzz, gxy, gxz, gyz;
    const uint32_t cnt = ln.valid ? count[k] : 0u;
    const uint32_t listed = cnt < kmax ? cnt : kmax;
    const bool solid = prm.solid_stress != 0;
    const bool grouped = solid && gci != GROUP_CODE_NONE;  // stress terms with the particle's own group
    // the LDS stage exists whenever the gather mode uses it (one uint4 otherwise)
    __shared__ uint4 stage[PAIRSMX_STAGE ? (BLOCK / 64) * StageOf<PairRecF>::WORDS : 1];
    uint4* my_stage = PAIRSMX_STAGE ? stage + (threadIdx.x >> 6) * StageOf<PairRecF>::WORDS : stage;
    for_each_neighbor<LPP, PAIRSMX_STAGE>(rec, list, k, kmax >> 2, listed, ln.sub, my_stage, [&](const PairRecF& o, uint32_t e) {
        const float rx = float(int32_t(me.x - o.x)) * F.stepf;
        const float ry = float(int32_t(me.y - o.y)) * F.stepf;
        const float rz = float(int32_t(me.z - o.z)) * F.stepf;
        const float hbar = 0.5f * (me.h + o.h);
        const float hInv = __builtin_amdgcn_rcpf(hbar);
        const float hInv2 = hInv * hInv;
        const float d2 = rx * rx + ry * ry + rz * rz;
        const float gk = m4_grad_f(d2 * hInv2);
        const float h5g = (hInv2 * hInv2) * hInv * gk;
        const float gx = rx * h5g, gy = ry * h5g, gz = rz * h5g;
        const float wx = float(int32_t(o.vx - me.vx)) * F.vstepf;  // v_j - v_i
        const float wy = float(int32_t(o.vy - me.vy)) * F.vstepf;
        const float wz = float(int32_t(o.vz - me.vz)) * F.vstepf;
        const float mj = o.m;
        const float mjr = mj * rhoInv;
        const float wg = wx * gx + wy * gy + wz * gz;
        divv.add(mjr * wg);
        const float coef = me.pr + o.pr;
        float fx = -(coef * gx), fy = -(coef * gy), fz = -(coef * gz);
        bool same = grouped && (__float_as_uint(o.cs) & 15u) == gci;
        if (same && gci == GROUP_CODE_WIDE) same = group[e] == gi;
        if (same) {
            gxx.add(mjr * (wx * gx));
            gyy.add(mjr * (wy * gy));
            gzz.add(mjr * (wz * gz));
            gxy.add(mjr * (0.5f * (wx * gy + wy * gx)));
            gxz.add(mjr * (0.5f * (wx * gz + wz * gx)));
            gyz.add(mjr * (0.5f * (wy * gz + wz * gy)));
            const float xx = me.sxx + o.sxx, yy = me.syy + o.syy, xy = me.sxy + o.sxy;
            const float xz = me.sxz + o.sxz, yz = me.syz + o.syz;
            fx += xx * gx + xy * gy + xz * gz;
            fy += xy * gx + yy * gy + yz * gz;
            fz += xz * gx + yz * gy + (-xx - yy) * gz;
        }
        const float dvdr = -(wx * rx + wy * ry + wz * rz);
        if (dvdr < 0.0f) {
            const float rhobar = 0.5f * (me.rho + o.rho);
            const float csbar = 0.5f * (csi + __uint_as_float(__float_as_uint(o.cs) & ~15u));
            const float mu = hbar * dvdr * __builtin_amdgcn_rcpf(d2 + 1.e-2f * (hbar * hbar));
            const float av = (-alpha * csbar * mu + beta * (mu * mu)) * __builtin_amdgcn_rcpf(rhobar);
            fx -= av * gx;
            fy -= av * gy;
            fz -= av * gz;
            du.add(mj * (0.5f * av * -wg));
        }
        ax.add(mj * fx);
        ay.add(mj * fy);
        az.add(mj * fz);
    });
    const double sax = group_sum<LPP>(ax.sum), say = group_sum<LPP>(ay.sum), saz = group_sum<LPP>(az.sum);
    const double sdu = group_sum<LPP>(du.sum), sdivv = group_sum<LPP>(divv.sum);
    double g[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
    if (solid) {
        g[0] = group_sum<LPP>(gxx.sum);
        g[1] = group_sum<LPP>(gyy.sum);
        g[2] = group_sum<LPP>(gzz.sum);
        g[3] = group_sum<LPP>(gxy.sum);
        g[4] = group_sum<LPP>(gxz.sum);
        g[5] = group_sum<LPP>(gyz.sum);
    }
    double cv[4] = { INFINITY, INFINITY, INFINITY, INFINITY };
    if (ln.valid && ln.sub == 0) {
        const uint32_t i = order[k];
        const double fax = prm.add_accel ? sax + s.d[AX][i] : sax;
        const double fay = prm.add_accel ? say + s.d[AY][i] : say;
        const double faz = prm.add_accel ? saz + s.d[AZ][i] : saz;
        s.d[AX][i] = fax;
        s.d[AY][i] = fay;
        s.d[AZ][i] = faz;
        s.d[DU][i] = sdu;
        s.d[DIVV][i] = sdivv;
        if (solid) {
            s.d[GXX][i] = g[0];
            s.d[GYY][i] = g[1];
            s.d[GZZ][i] = g[2];
            s.d[GXY][i] = g[3];
            s.d[GXZ][i] = g[4];
            s.d[GYZ][i] = g[5];
        }
        s.u[NCNT][i] = cnt;
        if (FUSED && !*fa.overflow) {
            const osph_material m = fa.mats[s.u[MAT][i]];
            const PairSums ps{ { fax, fay, faz }, sdu, sdivv, { g[0], g[1], g[2], g[3], g[4], g[5] } };
            finalize_at(s, m, prm, i, ps);
            correct_at(s, m, prm, fa.dt, i, ps.a);
            criteria_values(s, m, prm, i, ps.a, ps.divv, cv);
        }
    }
    if (FUSED && !*fa.overflow) reduce_criteria(cv, fa.mins);
}

template <int LPP, bool FUSED>
__global__ void __launch_bounds__(BLOCK) k_pairs_mixed(DevState s, Params prm, Fixed F,
                                                       const PairRecF* __restrict__ rec,
                                                       const uint32_t* __restrict__ group,
                                                       const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ count, uint32_t kmax, size_t n,
                                                       FuseArgs fa) {
    pairs_mixed_body<LPP, FUSED>(size_t(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x, s, prm, F, rec,
                                 group, order, list, count, kmax, n, fa);
}

// ScalarGradyKippModel::integrate (core/physics/Damage.cpp:128-170) for particle i.
__device__ void damage_rate(DevState& s, const osph_material& m, size_t i) {
#pragma clang fp contract(off)
    const double D0 = s.d[D][i];
    if (D0 >= m.d_hi) {
        s.d[DD][i] = LARGE_D;
        return;
    }
    const double pr = s.d[P][i];
    const double sxx = s.d[SXX][i], syy = s.d[SYY][i];
    const double t[6] = { sxx - pr * 1.0, syy - pr * 1.0, (-sxx - syy) - pr * 1.0,
                          s.d[SXY][i] - pr * 0.0, s.d[SXZ][i] - pr * 0.0, s.d[SYZ][i] - pr * 0.0 };
    double sig[3];
    eigenvalues(t, sig);
    const double sigMax = dmax(dmax(sig[0], sig[1]), sig[2]);
    const double young_red = dmax((1.0 - D0 * D0 * D0) * m.young_modulus, 1.e-20);
    const double strain = sigMax / young_red;
    const double ratio = strain / s.d[EPSMIN][i];
    if (ratio <= 1.0) {
        s.d[DD][i] = 0.0;
        return;
    }
    s.d[DD][i] = s.d[GROWTH][i] * cbrt(dmin(pow(ratio, s.d[MZERO][i]), double(s.u[NFLAWS][i])));
}

// EquationHolder::finalize in reverse term order (EquationTerm.h:293-297): AdaptiveSmoothingLength
// (EquationTerm.cpp:366-387) or ConstSmoothingLength (:427-434), ContinuityEquation (:289-299),
// SolidStressForce (:177-203), PressureForce (:90-99); then the material finalize (damage).  Every
// highest derivative the step integrates is assigned here or in the pair sums (the reference zeroes them
// before the evaluation and adds the terms' contributions, which is the same value).
__device__ void finalize_at(DevState& s, const osph_material& m, const Params& prm, size_t i, const PairSums& ps) {
#pragma clang fp contract(off)
    const double rho = s.d[RHO][i], divv = ps.divv;
    const double h = s.d[H][i];
    s.d[VH][i] = (prm.adaptive_h && h > 2.0 * prm.h_lo) ? h / 3.0 * divv : 0.0;
    s.d[DRHO][i] = -rho * divv;
    double du = ps.du;
    if (prm.solid_stress && m.yielding != 0) {
        const double sxx = s.d[SXX][i], syy = s.d[SYY][i], szz = -sxx - syy;
        const double gxx = ps.g[0], gyy = ps.g[1], gzz = ps.g[2];
        const double gxy = ps.g[3], gxz = ps.g[4], gyz = ps.g[5];
        const double dd = (sxx * gxx + syy * gyy + szz * gzz) +
                          2.0 * (s.d[SXY][i] * gxy + s.d[SXZ][i] * gxz + s.d[SYZ][i] * gyz);
        du += 1.0 / rho * dd;
        const double tr3 = (gxx * 1.0 + gyy * 1.0 + gzz * 1.0) / 3.0;
        const double twomu = 2.0 * m.shear_modulus;
        s.d[DSXX][i] = twomu * (gxx - tr3);
        s.d[DSYY][i] = twomu * (gyy - tr3);
        s.d[DSXY][i] = twomu * gxy;
        s.d[DSXZ][i] = twomu * gxz;
        s.d[DSYZ][i] = twomu * gyz;
    } else if (prm.solid_stress) {
        for (int k = 0; k < 5; ++k) s.d[DSXX + k][i] = 0.0;
    }
    du -= s.d[P][i] / rho * divv;
    s.d[DU][i] = du;
    if (prm.has_damage) {
        if (m.damage == 1) {
            damage_rate(s, m, i);
        } else {
            s.d[DD][i] = 0.0;
        }
    }
}

__global__ void __launch_bounds__(BLOCK) k_finalize(DevState s, const osph_material* __restrict__ mats, Params prm,
                                                    size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    finalize_at(s, mats[s.u[MAT][i]], prm, i, sums_of(s, prm.solid_stress != 0, i));
}

// Storage::zeroHighestDerivatives (core/quantities/Storage.cpp:594-602)
__global__ void __launch_bounds__(BLOCK) k_zero_derivatives(DevState s, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int planes[] = { AX, AY, AZ, DRHO, DU, DSXX, DSYY, DSXY, DSXZ, DSYZ, DD };
#pragma unroll
    for (int k = 0; k < 11; ++k) s.d[planes[k]][i] = 0.0;
}

// clampWithDerivative<Float> (core/objects/wrappers/Interval.h:158-162)
__device__ __forceinline__ void clamp_deriv(double& v, double& dv, double lo, double hi) {
    const bool zero = (v >= hi && dv > 0.0) || (v <= lo && dv < 0.0);
    v = dmax(lo, dmin(v, hi));
    if (zero) dv = 0.0;
}

// PredictorCorrector::makePredictions (core/timestepping/TimeStepping.cpp:286-300).  The swap of the
// highest derivatives into the predictions storage and Storage::zeroHighestDerivatives (:331-334) are
// done by the host before this pass, by exchanging the plane pointers of each highest derivative and its
// prediction (AX <-> PAX, DRHO <-> PDRHO, ...): the derivative of the last evaluation is read here from
// the prediction planes, and the derivative planes are left for the next evaluation, which assigns every
// one of them (the pair sums, the finalize).  A derivative zeroed by a clamp is written back to its
// prediction plane, as the reference clamps the derivative it saves.
__device__ void predict_at(DevState& s, const osph_material& m, const Params& prm, double dt, size_t i) {
#pragma clang fp contract(off)
    const double dt2 = 0.5 * (dt * dt);
    for (int k = 0; k < 3; ++k) {
        const double a = s.d[PAX + k][i];
        const double v = s.d[VX + k][i];
        s.d[X + k][i] += v * dt + a * dt2;
        s.d[VX + k][i] = v + a * dt;
    }
    s.d[H][i] += s.d[VH][i] * dt + 0.0 * dt2;
    auto first_order = [&](int q, int pq, double lo, double hi) {
        double x = s.d[q][i];
        const double dx0 = s.d[pq][i];
        double dx = dx0;
        x += dx * dt;
        clamp_deriv(x, dx, lo, hi);
        s.d[q][i] = x;
        if (dx != dx0) s.d[pq][i] = dx;
    };
    first_order(RHO, PDRHO, m.rho_lo, m.rho_hi);
    first_order(U, PDU, m.u_lo, m.u_hi);
    if (prm.solid_stress) {
        for (int k = 0; k < 5; ++k) s.d[SXX + k][i] += s.d[PDSXX + k][i] * dt;
    }
    if (prm.has_damage) first_order(D, PDD, m.d_lo, m.d_hi);
}

__device__ __forceinline__ void atomic_min_pos(unsigned long long* addr, double v) {
    // v >= 0 (or +inf): the IEEE bit patterns of non-negative doubles are ordered like the values
    atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}

// Per-particle values of the four criteria of MultiCriterion (TimeStepCriterion.cpp:315-356 Courant,
// :123-211 DerivativeCriterion with the minimal-step mean, :221-257 Acceleration, :267-305 Divergence),
// reduced to the minimum of each criterion in mins[0..3].
__device__ void criteria_values(const DevState& s, const osph_material& m, const Params& prm, size_t i,
                                const double a[3], double divv, double out[4]) {
#pragma clang fp contract(off)
    out[0] = out[1] = out[2] = out[3] = INFINITY;
    const double h = s.d[H][i];
    if (prm.criteria & 1) {
        const double cs = s.d[CS][i];
        if (cs > 0.0) out[0] = prm.courant * h / cs;
    }
    if (prm.criteria & 2) {
        double mn = INFINITY;
        auto add = [&](double v, double dv, double minValue) {
            const double av = fabs(v), adv = fabs(dv);
            if (av < 2.0 * minValue) return;
            mn = dmin(mn, prm.deriv_factor * (av + minValue) / (adv + EPS_D));
        };
        add(s.d[RHO][i], s.d[DRHO][i], m.rho_min);
        add(s.d[U][i], s.d[DU][i], m.u_min);
        if (prm.solid_stress) {
            const double xx = s.d[SXX][i], yy = s.d[SYY][i];
            const double dxx = s.d[DSXX][i], dyy = s.d[DSYY][i];
            add(xx, dxx, m.s_min);
            add(yy, dyy, m.s_min);
            add(-xx - yy, -dxx - dyy, m.s_min);
            add(s.d[SXY][i], s.d[DSXY][i], m.s_min);
            add(s.d[SXZ][i], s.d[DSXZ][i], m.s_min);
            add(s.d[SYZ][i], s.d[DSYZ][i], m.s_min);
        }
        if (prm.has_damage) add(s.d[D][i], s.d[DD][i], m.d_min);
        out[1] = mn;
    }
    if (prm.criteria & 4) {
        const double ax = a[0], ay = a[1], az = a[2];
        const double dvNorm = ax * ax + ay * ay + az * az;
        if (dvNorm > EPS_D) out[2] = prm.deriv_factor * sqrt(sqrt((h * h) / dvNorm));
    }
    if (prm.criteria & 8) {
        const double dv = fabs(divv);
        if (dv > EPS_D) out[3] = prm.div_factor / dv;
    }
}

__device__ void reduce_criteria(double v[4], unsigned long long* mins) {
    __shared__ double red[BLOCK / 64][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < 4; ++c) {
        double x = v[c];
        for (int o = 32; o > 0; o >>= 1) x = dmin(x, __shfl_xor(x, o, 64));
        if (lane == 0) red[wave][c] = x;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        // the pair kernels run with blocks of 64, 128 or 256 threads: only the block's own waves hold values
        const int waves = int(blockDim.x >> 6);
        double x = red[0][threadIdx.x];
        for (int w = 1; w < waves; ++w) x = dmin(x, red[w][threadIdx.x]);
        // the atomic only when it would lower the current minimum (a device-coherent read): most blocks
        // then skip it, where thousands of same-address atomics would serialise
        const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
        if (x < INFINITY && bits < __hip_atomic_load(&mins[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            atomic_min_pos(&mins[threadIdx.x], x);
        }
    }
}

// PredictorCorrector::makeCorrections (TimeStepping.cpp:302-322): pdv is the new derivative, cdv the
// saved prediction; the clamp zeroes the new derivative at the bounds.
__device__ void correct_at(DevState& s, const osph_material& m, const Params& prm, double dt, size_t i,
                           const double a[3]) {
#pragma clang fp contract(off)
    {
        const double dt2 = 0.5 * (dt * dt);
        const double ca = 1.0 / 3.0, cb = 0.5;
        for (int k = 0; k < 3; ++k) {
            const double pdv = a[k], cdv = s.d[PAX + k][i];
            s.d[X + k][i] -= ca * (cdv - pdv) * dt2;
            s.d[VX + k][i] -= cb * (cdv - pdv) * dt;
        }
        {
            double x = s.d[RHO][i], px = s.d[DRHO][i];
            x -= 0.5 * (s.d[PDRHO][i] - px) * dt;
            clamp_deriv(x, px, m.rho_lo, m.rho_hi);
            s.d[RHO][i] = x;
            s.d[DRHO][i] = px;
        }
        {
            double x = s.d[U][i], px = s.d[DU][i];
            x -= 0.5 * (s.d[PDU][i] - px) * dt;
            clamp_deriv(x, px, m.u_lo, m.u_hi);
            s.d[U][i] = x;
            s.d[DU][i] = px;
        }
        if (prm.solid_stress) {
            for (int k = 0; k < 5; ++k) {
                s.d[SXX + k][i] -= 0.5 * (s.d[PDSXX + k][i] - s.d[DSXX + k][i]) * dt;
            }
        }
        if (prm.has_damage) {
            double x = s.d[D][i], px = s.d[DD][i];
            x -= 0.5 * (s.d[PDD][i] - px) * dt;
            clamp_deriv(x, px, m.d_lo, m.d_hi);
            s.d[D][i] = x;
            s.d[DD][i] = px;
        }
    }
}

// The end of a predictor-corrector step for particle i: EquationHolder::finalize and the material
// finalize, the corrections, then the criteria of the corrected state, reduced into mins[0..3].  When
// the neighbour lists overflowed (*overflow != 0) the kernel changes nothing: the step is repeated with
// larger lists from the same predicted state.
__device__ __forceinline__ void finalize_correct_body(size_t i, DevState s, const osph_material* __restrict__ mats,
                                                      const Params& prm, double dt, unsigned long long* mins,
                                                      const uint32_t* overflow, size_t n) {
    if (*overflow) return;  // uniform over the grid
    double v[4] = { INFINITY, INFINITY, INFINITY, INFINITY };
    if (i < n) {
        const osph_material m = mats[s.u[MAT][i]];
        const PairSums ps = sums_of(s, prm.solid_stress != 0, i);
        finalize_at(s, m, prm, i, ps);
        correct_at(s, m, prm, dt, i, ps.a);
        criteria_values(s, m, prm, i, ps.a, ps.divv, v);
    }
    reduce_criteria(v, mins);
}

__global__ void __launch_bounds__(BLOCK) k_finalize_correct(DevState s, const osph_material* __restrict__ mats,
                                                            Params prm, double dt, unsigned long long* mins,
                                                            const uint32_t* overflow, size_t n) {
    finalize_correct_body(size_t(blockIdx.x) * blockDim.x + threadIdx.x, s, mats, prm, dt, mins, overflow, n);
}

// Per-block partial grid bounds.  Values v = CLASS_VALUES * c + k for class c (k: 0..2 min x, y, z;
// 3..5 max x, y, z; 6 max h; 7 min h; 8 count, min and max as order-preserving bit patterns), then
// VEL_VALUES velocity bounds (min x, y, z, max x, y, z).  part[v * nblocks + block]; k_bounds_final
// reduces them.  Two levels instead of per-block atomics: same-address atomics from thousands of blocks
// serialise (56 us for 1024 blocks at the impact).
constexpr int CLASS_VALUES = 9;
constexpr int VEL_VALUES = 6;
constexpr int BOUND_VALUES = CLASS_VALUES * MAX_CLASSES + VEL_VALUES;

__device__ __forceinline__ bool bound_is_min(int v) {
    if (v >= CLASS_VALUES * MAX_CLASSES) return v - CLASS_VALUES * MAX_CLASSES < 3;
    const int k = v % CLASS_VALUES;
    return k < 3 || k == 7;
}

__device__ __forceinline__ unsigned long long bound_identity(int v) {
    if (v < CLASS_VALUES * MAX_CLASSES && v % CLASS_VALUES == 8) return 0ull;
    return bound_is_min(v) ? ~0ull : 0ull;
}

__device__ __forceinline__ unsigned long long bound_combine(int v, unsigned long long a, unsigned long long b) {
    if (v < CLASS_VALUES * MAX_CLASSES && v % CLASS_VALUES == 8) return a + b;
    return bound_is_min(v) ? (b < a ? b : a) : (b > a ? b : a);
}

// Block-level partial bounds of one particle per thread (cls < 0: no particle).  Every thread of the
// block must call it.
template <bool VEL>
__device__ void block_bounds(int cls, double x, double y, double z, double h, const double* v,
                             unsigned long long* __restrict__ part, uint32_t nblocks, uint32_t block) {
    __shared__ unsigned long long red[BLOCK / 64][BOUND_VALUES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long b[4] = { ord_bits(x), ord_bits(y), ord_bits(z), ord_bits(h) };
    for (int c = 0; c < MAX_CLASSES; ++c) {
        const bool in = cls == c;
        unsigned long long* r = red[wave] + CLASS_VALUES * c;
        if (__any(in)) {
            for (int k = 0; k < 3; ++k) {
                const unsigned long long lo = wave_min(in ? b[k] : ~0ull);
                const unsigned long long hi = wave_max(in ? b[k] : 0ull);
                if (lane == 0) {
                    r[k] = lo;
                    r[3 + k] = hi;
                }
            }
            const unsigned long long hmax = wave_max(in ? b[3] : 0ull), hmin = wave_min(in ? b[3] : ~0ull);
            const unsigned long long cnt = __popcll(__ballot(in));
            if (lane == 0) {
                r[6] = hmax;
                r[7] = hmin;
                r[8] = cnt;
            }
        } else if (lane == 0) {
            for (int k = 0; k < CLASS_VALUES; ++k) r[k] = bound_identity(CLASS_VALUES * c + k);
        }
    }
    if (VEL) {
        for (int k = 0; k < 3; ++k) {
            const bool ok = cls >= 0;
            const unsigned long long bv = ord_bits(v[k]);
            const unsigned long long lo = wave_min(ok ? bv : ~0ull), hi = wave_max(ok ? bv : 0ull);
            if (lane == 0) {
                red[wave][CLASS_VALUES * MAX_CLASSES + k] = lo;
                red[wave][CLASS_VALUES * MAX_CLASSES + 3 + k] = hi;
            }
        }
    }
    __syncthreads();
    const int nv = VEL ? BOUND_VALUES : CLASS_VALUES * MAX_CLASSES;
    if (int(threadIdx.x) < nv) {
        const int vi = threadIdx.x;
        unsigned long long acc = red[0][vi];
        for (int w = 1; w < int(blockDim.x / 64); ++w) acc = bound_combine(vi, acc, red[w][vi]);
        part[size_t(vi) * nblocks + block] = acc;
    }
}

// One pass before the neighbour search, MODE 2: predictions (predict_at) then the material initialize;
// MODE 1: the material initialize only; MODE 0: neither.  Every mode ends with the per-block partial
// grid bounds of the resulting positions (class of h relative to href) and, with VEL, velocities.
template <int MODE, bool VEL>
__device__ __forceinline__ void advance_body(uint32_t bid, uint32_t nb, DevState s,
                                             const osph_material* __restrict__ mats, const Params& prm, double dt,
                                             double href, unsigned long long* __restrict__ part, size_t n) {
    const size_t i = size_t(bid) * BLOCK + threadIdx.x;
    int cls = -1;
    double x = 0.0, y = 0.0, z = 0.0, h = 0.0, v[3] = { 0.0, 0.0, 0.0 };
    if (i < n) {
        if (MODE > 0) {
            const osph_material m = mats[s.u[MAT][i]];
            if (MODE > 1) predict_at(s, m, prm, dt, i);
            material_init_at(s, m, prm, i);
        }
        x = s.d[X][i];
        y = s.d[Y][i];
        z = s.d[Z][i];
        h = s.d[H][i];
        cls = h_class(h, href, MAX_CLASSES);
        if (VEL) {
            v[0] = s.d[VX][i];
            v[1] = s.d[VY][i];
            v[2] = s.d[VZ][i];
        }
    }
    block_bounds<VEL>(cls, x, y, z, h, v, part, nb, bid);
}

template <int MODE, bool VEL>
__global__ void __launch_bounds__(BLOCK) k_advance(DevState s, const osph_material* __restrict__ mats, Params prm,
                                                   double dt, double href, unsigned long long* __restrict__ part,
                                                   size_t n) {
    advance_body<MODE, VEL>(blockIdx.x, gridDim.x, s, mats, prm, dt, href, part, n);
}

// Second level of the bounds: block v reduces value v over the nblocks partials into acc (the class
// values, at CLASS_STRIDE * c + k) or vel.
__device__ __forceinline__ void bounds_final_at(int vi, const unsigned long long* __restrict__ part,
                                                uint32_t nblocks, unsigned long long* __restrict__ acc,
                                                unsigned long long* __restrict__ vel) {
    __shared__ unsigned long long red[BLOCK / 64];
    unsigned long long r = bound_identity(vi);
    for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x) r = bound_combine(vi, r, part[size_t(vi) * nblocks + b]);
    for (int o = 32; o > 0; o >>= 1) r = bound_combine(vi, r, __shfl_xor(r, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x / 64); ++w) r = bound_combine(vi, r, red[w]);
        if (vi < CLASS_VALUES * MAX_CLASSES) {
            acc[CLASS_STRIDE * (vi / CLASS_VALUES) + vi % CLASS_VALUES] = r;
        } else {
            vel[vi - CLASS_VALUES * MAX_CLASSES] = r;
        }
    }
}

// Block v reduces bound value v; block 0 also starts the step's other accumulators (global bounds,
// criterion minima, list overflow flag: words 0 .. SMALL_CLASSES - 1, which no earlier pass of the step
// uses), in place of a launch of its own.
__global__ void __launch_bounds__(BLOCK) k_bounds_final(const unsigned long long* __restrict__ part, uint32_t nblocks,
                                                        unsigned long long* __restrict__ small) {
    if (blockIdx.x == 0 && threadIdx.x < SMALL_CLASSES) init_small_at(small, threadIdx.x);
    bounds_final_at(blockIdx.x, part, nblocks, small + SMALL_CLASSES, small + SMALL_VEL);
}

__global__ void __launch_bounds__(BLOCK) k_criteria(DevState s, const osph_material* __restrict__ mats, Params prm,
                                                    unsigned long long* mins, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double v[4] = { INFINITY, INFINITY, INFINITY, INFINITY };
    if (i < n) {
        const osph_material m = mats[s.u[MAT][i]];
        const double a[3] = { s.d[AX][i], s.d[AY][i], s.d[AZ][i] };
        criteria_values(s, m, prm, i, a, s.d[DIVV][i], v);
    }
    reduce_criteria(v, mins);
}

struct StageDesc {
    int ncomp;
    int elem;
    int off[8];
    void* p[8];
    const uint32_t* slot;  // storage index -> device slot
};

__global__ void __launch_bounds__(BLOCK) k_stage_to_planes(const unsigned char* __restrict__ st, size_t stride,
                                                           StageDesc d, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned char* e = st + i * stride;
    const size_t k = d.slot[i];
    for (int c = 0; c < d.ncomp; ++c) {
        if (d.elem == 8) {
            static_cast<double*>(d.p[c])[k] = *reinterpret_cast<const double*>(e + d.off[c]);
        } else {
            static_cast<uint32_t*>(d.p[c])[k] = *reinterpret_cast<const uint32_t*>(e + d.off[c]);
        }
    }
}

__global__ void __launch_bounds__(BLOCK) k_planes_to_stage(unsigned char* __restrict__ st, size_t stride, StageDesc d,
                                                           size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned char* e = st + i * stride;
    const size_t k = d.slot[i];
    for (int c = 0; c < d.ncomp; ++c) {
        if (d.elem == 8) {
            *reinterpret_cast<double*>(e + d.off[c]) = static_cast<const double*>(d.p[c])[k];
        } else {
            *reinterpret_cast<uint32_t*>(e + d.off[c]) = static_cast<const uint32_t*>(d.p[c])[k];
        }
    }
}

// Moves the state into cell order: slot k of dst receives slot order[k] of src, and the slot <-> storage
// maps follow (perm_out[k] = storage index of slot k, slot_of[storage index] = k).
__device__ __forceinline__ void permute_at(const DevState& src, const DevState& dst, const uint32_t* __restrict__ order,
                                           const uint32_t* __restrict__ perm_in, uint32_t* __restrict__ perm_out,
                                           uint32_t* __restrict__ slot_of, size_t k) {
    const uint32_t i = order[k];
#pragma unroll 4
    for (int p = 0; p < NDPLANES; ++p) dst.d[p][k] = src.d[p][i];
#pragma unroll
    for (int p = 0; p < NIPLANES; ++p) dst.u[p][k] = src.u[p][i];
    const uint32_t g = perm_in[i];
    perm_out[k] = g;
    slot_of[g] = uint32_t(k);
}

__global__ void __launch_bounds__(BLOCK) k_permute_state(DevState src, DevState dst, const uint32_t* __restrict__ order,
                                                         const uint32_t* __restrict__ perm_in,
                                                         uint32_t* __restrict__ perm_out,
                                                         uint32_t* __restrict__ slot_of, size_t n) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) permute_at(src, dst, order, perm_in, perm_out, slot_of, k);
}

__global__ void __launch_bounds__(BLOCK) k_iota_u32(uint32_t* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = uint32_t(i);
}

// STREAM-style copy: b[i] = a[i], 16 bytes per lane per iteration.
__global__ void __launch_bounds__(BLOCK) k_stream_copy(const double2* __restrict__ a, double2* __restrict__ b,
                                                       size_t n2) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += size_t(gridDim.x) * blockDim.x) {
        b[i] = a[i];
    }
}

// Integrals of the state (core/physics/Integrals.cpp:24-180): per block, in a fixed reduction order, then
// one block over the partials.  v[0] kinetic energy, v[1] internal energy, v[2..4] momentum, v[5..7]
// angular momentum about the origin, v[8..10] sum of m r, v[11] total mass.
constexpr int NINT = 12;

__global__ void __launch_bounds__(BLOCK) k_integrals(DevState s, size_t n, double* __restrict__ partial) {
    double v[NINT] = { 0.0 };
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const double m = s.d[M][i];
        const double x = s.d[X][i], y = s.d[Y][i], z = s.d[Z][i];
        const double vx = s.d[VX][i], vy = s.d[VY][i], vz = s.d[VZ][i];
        v[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[1] += m * s.d[U][i];
        v[2] += m * vx;
        v[3] += m * vy;
        v[4] += m * vz;
        v[5] += m * (y * vz - z * vy);
        v[6] += m * (z * vx - x * vz);
        v[7] += m * (x * vy - y * vx);
        v[8] += m * x;
        v[9] += m * y;
        v[10] += m * z;
        v[11] += m;
    }
    __shared__ double red[BLOCK / 64][NINT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < NINT; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[wave][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < NINT) {
        double x = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) x += red[w][threadIdx.x];
        partial[size_t(blockIdx.x) * NINT + threadIdx.x] = x;
    }
}

__global__ void __launch_bounds__(64) k_integrals_final(const double* __restrict__ partial, int nblocks,
                                                        double* __restrict__ out) {
    if (threadIdx.x >= NINT) return;
    double x = 0.0;
    for (int b = 0; b < nblocks; ++b) x += partial[size_t(b) * NINT + threadIdx.x];
    out[threadIdx.x] = x;
}

}  // namespace

size_t integrals_scratch_doubles(size_t n) {
    unsigned g = grid_for(n);
    g = g > 1024 ? 1024 : g;
    return size_t(g) * NINT + NINT;
}

void launch_integrals(const Launch& l, DevState s, double* scratch, size_t n) {
    unsigned g = grid_for(n);
    g = g > 1024 ? 1024 : g;
    k_integrals<<<g, BLOCK, 0, l.stream>>>(s, n, scratch);
    k_integrals_final<<<1, 64, 0, l.stream>>>(scratch, int(g), scratch + size_t(g) * NINT);
}

void launch_stream_copy(hipStream_t stream, const void* a, void* b, size_t bytes) {
    const size_t n2 = bytes / sizeof(double2);
    if (!n2) return;
    unsigned g = grid_for(n2);
    g = g > 8192 ? 8192 : g;
    k_stream_copy<<<g, BLOCK, 0, stream>>>(static_cast<const double2*>(a), static_cast<double2*>(b), n2);
}

// ---- launchers ----------------------------------------------------------------------------------------

void launch_init_small(const Launch& l, unsigned long long* small) {
    k_init_small<<<1, SMALL_WORDS, 0, l.stream>>>(small);
}

void launch_bounds_reduce(const Launch& l, DevState s, unsigned long long* acc, size_t n) {
    if (!n) return;
    unsigned g = grid_for(n);
    g = g > 1024 ? 1024 : g;
    k_bounds<<<g, BLOCK, 0, l.stream>>>(s, acc, n);
}

void launch_cell_keys(const Launch& l, DevState s, Grid g, uint32_t* keys, uint32_t* idx, uint32_t* range,
                      size_t clear, size_t n) {
    if (!n) return;
    k_cell_keys<<<grid_for(n), BLOCK, 0, l.stream>>>(s, g, keys, idx, range, clear, n);
}

void launch_clear_u32(const Launch& l, uint32_t* p, size_t count, uint32_t value) {
    if (!count) return;
    unsigned g = grid_for(count);
    g = g > 4096 ? 4096 : g;
    k_fill_u32<<<g, BLOCK, 0, l.stream>>>(p, count, value);
}

void launch_cell_first(const Launch& l, const uint32_t* keys, Grid g, uint32_t* row_range, uint32_t* first,
                       size_t rows, size_t n) {
    // without particles nothing reads the tables (and the row ranges were not cleared)
    if (!rows || !n) return;
    int xslots = 1;
    for (int c = 0; c < g.nclass; ++c) {
        if (!g.c[c].empty) xslots = xslots > g.c[c].dim[0] + 1 ? xslots : g.c[c].dim[0] + 1;
    }
    k_row_cells<<<grid_for(n), BLOCK, 0, l.stream>>>(keys, g, row_range, first, uint32_t(n));
    k_empty_rows<<<grid_for(rows * size_t(xslots)), BLOCK, 0, l.stream>>>(g, row_range, first, rows, xslots);
}

size_t bounds_partials(size_t n) { return size_t(grid_for(n)) * BOUND_VALUES; }
size_t bounds_values() { return BOUND_VALUES; }

void launch_advance(const Launch& l, DevState s, const osph_material* mats, Params prm, int mode, double dt,
                    double href, unsigned long long* small, bool vel, unsigned long long* part, size_t n) {
    if (!n) return;
    const unsigned g = grid_for(n);
    const bool v = vel;
    if (mode == 2) {
        if (v) k_advance<2, true><<<g, BLOCK, 0, l.stream>>>(s, mats, prm, dt, href, part, n);
        else k_advance<2, false><<<g, BLOCK, 0, l.stream>>>(s, mats, prm, dt, href, part, n);
    } else if (mode == 1) {
        if (v) k_advance<1, true><<<g, BLOCK, 0, l.stream>>>(s, mats, prm, dt, href, part, n);
        else k_advance<1, false><<<g, BLOCK, 0, l.stream>>>(s, mats, prm, dt, href, part, n);
    } else {
        if (v) k_advance<0, true><<<g, BLOCK, 0, l.stream>>>(s, mats, prm, dt, href, part, n);
        else k_advance<0, false><<<g, BLOCK, 0, l.stream>>>(s, mats, prm, dt, href, part, n);
    }
    k_bounds_final<<<v ? BOUND_VALUES : CLASS_VALUES * MAX_CLASSES, BLOCK, 0, l.stream>>>(part, g, small);
}

void launch_pack(const Launch& l, DevState s, Params prm, const uint32_t* order, PairRec* rec, double4* pos,
                 float4* posf, Screen sc, uint32_t* group, size_t n) {
    if (!n) return;
    k_pack<<<grid_for(n), BLOCK, 0, l.stream>>>(s, prm, order, rec, pos, posf, sc, group, n);
}

void launch_neighbor_lists(const Launch& l, const double4* pos, const float4* posf, Screen sc, Grid g,
                           const uint32_t* first, uint32_t* list, uint32_t* count, uint32_t kmax, uint32_t* overflow,
                           size_t n) {
    if (!n) return;
    const int b = l.tune.lists_block;
    if (sc.on) {
        k_neighbor_lists<true><<<grid_for(n, b), b, size_t(l.tune.lists_lds), l.stream>>>(
            pos, posf, sc, g, first, list, count, kmax, overflow, n);
    } else {
        k_neighbor_lists<false><<<grid_for(n, b), b, size_t(l.tune.lists_lds), l.stream>>>(
            pos, posf, sc, g, first, list, count, kmax, overflow, n);
    }
}

void launch_pairs_fp64(const Launch& l, DevState s, Params prm, const PairRec* rec, const uint32_t* group,
                       const uint32_t* order, const uint32_t* list, const uint32_t* count, const double* lut_grads,
                       uint32_t kmax, size_t n, const FuseArgs* fuse) {
    if (!n) return;
    const int b = l.tune.pairs_block, lpp = l.tune.pairs_lanes;
    const unsigned grid = grid_for(n * size_t(lpp), b);
    const size_t lds = size_t(l.tune.pairs_lds);
    const FuseArgs fa = fuse ? *fuse : FuseArgs{};
#define OSPH_PAIRS64(P, F) \
    k_pairs_fp64<P, F><<<grid, b, lds, l.stream>>>(s, prm, rec, group, order, list, count, lut_grads, kmax, n, fa)
    if (lpp == 4) {
        if (fuse) OSPH_PAIRS64(4, true); else OSPH_PAIRS64(4, false);
    } else if (lpp == 2) {
        if (fuse) OSPH_PAIRS64(2, true); else OSPH_PAIRS64(2, false);
    } else {
        if (fuse) OSPH_PAIRS64(1, true); else OSPH_PAIRS64(1, false);
    }
#undef OSPH_PAIRS64
}

void launch_pack_mixed(const Launch& l, DevState s, Params prm, Fixed fx, const uint32_t* order, PairRecF* rec,
                       double4* pos, float4* posf, Screen sc, uint32_t* group, size_t n) {
    if (!n) return;
    k_pack_mixed<<<grid_for(n), BLOCK, 0, l.stream>>>(s, prm, fx, order, rec, pos, posf, sc, group, n);
}

void launch_pairs_mixed(const Launch& l, DevState s, Params prm, Fixed fx, const PairRecF* rec, const uint32_t* group,
                        const uint32_t* order, const uint32_t* list, const uint32_t* count, uint32_t kmax, size_t n,
                        const FuseArgs* fuse) {
    if (!n) return;
    const int b = l.tune.pairs_block, lpp = l.tune.pairs_lanes;
    const unsigned grid = grid_for(n * size_t(lpp), b);
    const size_t lds = size_t(l.tune.pairs_lds);
    const FuseArgs fa = fuse ? *fuse : FuseArgs{};
#define OSPH_PAIRSMX(P, F) \
    k_pairs_mixed<P, F><<<grid, b, lds, l.stream>>>(s, prm, fx, rec, group, order, list, count, kmax, n, fa)
    if (lpp == 4) {
        if (fuse) OSPH_PAIRSMX(4, true); else OSPH_PAIRSMX(4, false);
    } else if (lpp == 2) {
        if (fuse) OSPH_PAIRSMX(2, true); else OSPH_PAIRSMX(2, false);
    } else {
        if (fuse) OSPH_PAIRSMX(1, true); else OSPH_PAIRSMX(1, false);
    }
#undef OSPH_PAIRSMX
}

void launch_finalize(const Launch& l, DevState s, const osph_material* mats, Params prm, size_t n) {
    if (!n) return;
    k_finalize<<<grid_for(n), BLOCK, 0, l.stream>>>(s, mats, prm, n);
}

void launch_zero_derivatives(const Launch& l, DevState s, size_t n) {
    if (!n) return;
    k_zero_derivatives<<<grid_for(n), BLOCK, 0, l.stream>>>(s, n);
}

void launch_finalize_correct(const Launch& l, DevState s, const osph_material* mats, Params prm, double dt,
                             unsigned long long* mins, const uint32_t* overflow, size_t n) {
    if (!n) return;
    k_finalize_correct<<<grid_for(n), BLOCK, 0, l.stream>>>(s, mats, prm, dt, mins, overflow, n);
}

void launch_criteria(const Launch& l, DevState s, const osph_material* mats, Params prm, unsigned long long* mins,
                     size_t n) {
    if (!n) return;
    k_criteria<<<grid_for(n), BLOCK, 0, l.stream>>>(s, mats, prm, mins, n);
}

static StageDesc make_desc(int ncomp, const int* off, int elem, void* const* p) {
    StageDesc d{};
    d.ncomp = ncomp;
    d.elem = elem;
    for (int c = 0; c < ncomp && c < 8; ++c) {
        d.off[c] = off[c];
        d.p[c] = p[c];
    }
    return d;
}

void launch_stage_to_planes(const Launch& l, const unsigned char* staging, size_t stride, int ncomp,
                            const int* offsets_bytes, int elem_bytes, void* const* dst, const uint32_t* slot_of,
                            size_t n) {
    if (!n) return;
    StageDesc d = make_desc(ncomp, offsets_bytes, elem_bytes, dst);
    d.slot = slot_of;
    k_stage_to_planes<<<grid_for(n), BLOCK, 0, l.stream>>>(staging, stride, d, n);
}

void launch_planes_to_stage(const Launch& l, unsigned char* staging, size_t stride, int ncomp,
                            const int* offsets_bytes, int elem_bytes, void* const* src, const uint32_t* slot_of,
                            size_t n) {
    if (!n) return;
    StageDesc d = make_desc(ncomp, offsets_bytes, elem_bytes, src);
    d.slot = slot_of;
    k_planes_to_stage<<<grid_for(n), BLOCK, 0, l.stream>>>(staging, stride, d, n);
}

void launch_permute_state(const Launch& l, DevState src, DevState dst, const uint32_t* order, const uint32_t* perm_in,
                          uint32_t* perm_out, uint32_t* slot_of, size_t n) {
    if (!n) return;
    k_permute_state<<<grid_for(n), BLOCK, 0, l.stream>>>(src, dst, order, perm_in, perm_out, slot_of, n);
}

void launch_iota(const Launch& l, uint32_t* p, size_t n) {
    if (!n) return;
    k_iota_u32<<<grid_for(n), BLOCK, 0, l.stream>>>(p, n);
}

// ---- batched stages: one launch for every member of an ensemble ------------------------------------------
// Block b runs the single-run kernel's block bid = b - first[m] of the member m = member[b], out of the
// member's nb = first[m + 1] - first[m] blocks, on that member's buffers and parameters.

namespace {

__device__ __forceinline__ const BatchMember& batch_block(const BatchMember* __restrict__ M, const BatchBlocks& t,
                                                         uint32_t& m, uint32_t& bid, uint32_t& nb) {
    m = t.member[blockIdx.x];
    const uint32_t f = t.first[m];
    bid = blockIdx.x - f;
    nb = t.first[m + 1] - f;
    return M[m];
}

__device__ __forceinline__ uint32_t* overflow_of(const BatchMember& a) {
    return reinterpret_cast<uint32_t*>(a.small + SMALL_OVERFLOW);
}

__global__ void __launch_bounds__(SMALL_WORDS) kb_init_small(const BatchMember* __restrict__ M) {
    init_small_at(M[blockIdx.x].small, threadIdx.x);
}

template <bool VEL>
__global__ void __launch_bounds__(BLOCK) kb_advance(const BatchMember* __restrict__ M, BatchBlocks pb) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    advance_body<2, VEL>(bid, nb, a.s, a.mats, a.prm, a.dt, a.href, a.part, a.n);
}

__global__ void __launch_bounds__(BLOCK) kb_bounds_final(const BatchMember* __restrict__ M, BatchBlocks pb, int nv) {
    const uint32_t m = blockIdx.x / uint32_t(nv);
    const int vi = int(blockIdx.x % uint32_t(nv));
    const BatchMember& a = M[m];
    bounds_final_at(vi, a.part, pb.first[m + 1] - pb.first[m], a.small + SMALL_CLASSES, a.small + SMALL_VEL);
}

// The sort key of a batch holds the member above the member's own cell key, so that one sort groups the
// particles by member and, within a member, orders them as the member's own sort does (stable, by slot).
__global__ void __launch_bounds__(BLOCK) kb_cell_keys(const BatchMember* __restrict__ M, BatchBlocks pb, int key_bits,
                                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    const size_t i = size_t(bid) * BLOCK + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t at = a.key_off + uint32_t(i);
    keys[at] = (uint64_t(m) << key_bits) | uint64_t(cell_key_of(a.s, a.grid, i));
    vals[at] = at;
}

// The member's part of the sorted arrays: its cell keys and its slot order.
__global__ void __launch_bounds__(BLOCK) kb_split(const BatchMember* __restrict__ M, BatchBlocks pb, int key_bits,
                                                  const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    const size_t j = size_t(bid) * BLOCK + threadIdx.x;
    if (j >= a.n) return;
    const uint32_t at = a.key_off + uint32_t(j);
    a.keys_sorted[j] = uint32_t(keys[at] & ((uint64_t(1) << key_bits) - 1));
    a.order[j] = vals[at] - a.key_off;
}

__global__ void __launch_bounds__(BLOCK) kb_row_range(const BatchMember* __restrict__ M, BatchBlocks pb) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    for (size_t q = size_t(bid) * BLOCK + threadIdx.x; q < a.rows; q += size_t(nb) * BLOCK) {
        row_range_at(a.keys_sorted, a.grid, a.row_range, q, a.n);
    }
}

__global__ void __launch_bounds__(BLOCK) kb_cell_first(const BatchMember* __restrict__ M, BatchBlocks pb) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    const size_t total = size_t(a.rows) * size_t(a.xslots);
    for (size_t t = size_t(bid) * BLOCK + threadIdx.x; t < total; t += size_t(nb) * BLOCK) {
        cell_first_at(a.keys_sorted, a.grid, a.row_range, a.cell_first, t, a.rows, a.xslots);
    }
}

// Members whose step reorders: the state into cell order (s -> s2), and the order becomes the identity
// (each thread reads its own order entry before it overwrites it).
__global__ void __launch_bounds__(BLOCK) kb_permute(const BatchMember* __restrict__ M, BatchBlocks pb) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    if (!a.reorder) return;
    const size_t k = size_t(bid) * BLOCK + threadIdx.x;
    if (k >= a.n) return;
    permute_at(a.s, a.s2, a.order, a.perm, a.perm2, a.slot_of, k);
    a.order[k] = uint32_t(k);
}

template <bool MIXED>
__global__ void __launch_bounds__(BLOCK) kb_pack(const BatchMember* __restrict__ M, BatchBlocks pb) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    const size_t k = size_t(xcd_block(bid, nb)) * BLOCK + threadIdx.x;
    if (k >= a.n) return;
    if constexpr (MIXED) {
        pack_mixed_at(a.s2, a.prm, a.fx, a.order, static_cast<PairRecF*>(a.rec), a.pos, a.posf, a.screen, a.group, k);
    } else {
        pack_at(a.s2, a.prm, a.order, static_cast<PairRec*>(a.rec), a.pos, a.posf, a.screen, a.group, k);
    }
}

template <bool SCREEN>
__global__ void __launch_bounds__(BLOCK) kb_lists(const BatchMember* __restrict__ M, BatchBlocks lb) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, lb, m, bid, nb);
    if ((a.screen.on != 0) != SCREEN) return;
    const size_t k = size_t(xcd_block(bid, nb)) * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    neighbor_list_at<SCREEN>(a.pos, a.posf, a.screen, a.grid, a.cell_first, a.list, a.count, a.kmax, overflow_of(a), k);
}

template <bool MIXED, int LPP, bool FUSED>
__global__ void __launch_bounds__(BLOCK) kb_pairs(const BatchMember* __restrict__ M, BatchBlocks qb,
                                                  const double* __restrict__ lut) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, qb, m, bid, nb);
    const size_t t = size_t(xcd_block(bid, nb)) * blockDim.x + threadIdx.x;
    const FuseArgs fa{ a.mats, a.dt, a.small + SMALL_CRITERIA, overflow_of(a) };
    if constexpr (MIXED) {
        pairs_mixed_body<LPP, FUSED>(t, a.s2, a.prm, a.fx, static_cast<const PairRecF*>(a.rec), a.group, a.order,
                                     a.list, a.count, a.kmax, a.n, fa);
    } else {
        pairs_fp64_body<LPP, FUSED>(t, a.s2, a.prm, static_cast<const PairRec*>(a.rec), a.group, a.order, a.list,
                                    a.count, lut, a.kmax, a.n, fa);
    }
}

__global__ void __launch_bounds__(BLOCK) kb_finalize_correct(const BatchMember* __restrict__ M, BatchBlocks pb) {
    uint32_t m, bid, nb;
    const BatchMember& a = batch_block(M, pb, m, bid, nb);
    finalize_correct_body(size_t(bid) * BLOCK + threadIdx.x, a.s2, a.mats, a.prm, a.dt, a.small + SMALL_CRITERIA,
                          overflow_of(a), a.n);
}

}  // namespace

void batch_advance(const Launch& l, const BatchMember* M, int count, BatchBlocks pb, bool vel) {
    if (count <= 0) return;
    kb_init_small<<<count, SMALL_WORDS, 0, l.stream>>>(M);
    if (vel) {
        kb_advance<true><<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
    } else {
        kb_advance<false><<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
    }
    const int nv = vel ? BOUND_VALUES : CLASS_VALUES * MAX_CLASSES;
    kb_bounds_final<<<unsigned(count) * unsigned(nv), BLOCK, 0, l.stream>>>(M, pb, nv);
}

void batch_cell_keys(const Launch& l, const BatchMember* M, BatchBlocks pb, int key_bits, uint64_t* keys,
                     uint32_t* vals) {
    kb_cell_keys<<<pb.total, BLOCK, 0, l.stream>>>(M, pb, key_bits, keys, vals);
}

void batch_split(const Launch& l, const BatchMember* M, BatchBlocks pb, int key_bits, const uint64_t* keys,
                 const uint32_t* vals) {
    kb_split<<<pb.total, BLOCK, 0, l.stream>>>(M, pb, key_bits, keys, vals);
}

void batch_cells(const Launch& l, const BatchMember* M, BatchBlocks pb) {
    kb_row_range<<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
    kb_cell_first<<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
}

void batch_permute(const Launch& l, const BatchMember* M, BatchBlocks pb) {
    kb_permute<<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
}

void batch_pack(const Launch& l, const BatchMember* M, BatchBlocks pb, bool mixed) {
    if (mixed) {
        kb_pack<true><<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
    } else {
        kb_pack<false><<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
    }
}

void batch_lists(const Launch& l, const BatchMember* M, BatchBlocks lb, bool screened, bool plain) {
    const int b = l.tune.lists_block;
    const size_t lds = size_t(l.tune.lists_lds);
    if (screened) kb_lists<true><<<lb.total, b, lds, l.stream>>>(M, lb);
    if (plain) kb_lists<false><<<lb.total, b, lds, l.stream>>>(M, lb);
}

void batch_pairs(const Launch& l, const BatchMember* M, BatchBlocks qb, bool mixed, bool fuse, const double* lut) {
    const int b = l.tune.pairs_block, lpp = l.tune.pairs_lanes;
    const size_t lds = size_t(l.tune.pairs_lds);
#define OSPH_BPAIRS(MX, P, F) kb_pairs<MX, P, F><<<qb.total, b, lds, l.stream>>>(M, qb, lut)
    if (mixed) {
        if (lpp == 4) {
            if (fuse) OSPH_BPAIRS(true, 4, true); else OSPH_BPAIRS(true, 4, false);
        } else if (lpp == 2) {
            if (fuse) OSPH_BPAIRS(true, 2, true); else OSPH_BPAIRS(true, 2, false);
        } else {
            if (fuse) OSPH_BPAIRS(true, 1, true); else OSPH_BPAIRS(true, 1, false);
        }
    } else {
        if (lpp == 4) {
            if (fuse) OSPH_BPAIRS(false, 4, true); else OSPH_BPAIRS(false, 4, false);
        } else if (lpp == 2) {
            if (fuse) OSPH_BPAIRS(false, 2, true); else OSPH_BPAIRS(false, 2, false);
        } else {
            if (fuse) OSPH_BPAIRS(false, 1, true); else OSPH_BPAIRS(false, 1, false);
        }
    }
#undef OSPH_BPAIRS
}

void batch_finalize_correct(const Launch& l, const BatchMember* M, BatchBlocks pb) {
    kb_finalize_correct<<<pb.total, BLOCK, 0, l.stream>>>(M, pb);
}

// 64-bit keys (member above cell) of the batched sort, rocPRIM's default configuration.
size_t sort_pairs64_temp_bytes(size_t n, int key_bits) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                    (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0u, unsigned(key_bits));
    return bytes;
}

hipError_t sort_pairs64(const Launch& l, void* temp, size_t temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                        const uint32_t* vals_in, uint32_t* vals_out, size_t n, int key_bits) {
    if (!n) return hipSuccess;
    return rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0u,
                                     unsigned(key_bits), l.stream);
}

// Radix sort of the cell keys: rocPRIM's onesweep with launch shapes measured at the 1M impact (24-bit
// keys): 1024 threads x 4 items per block 84 us per sort, against 119 us for its default of 1024 x 16
// (65 blocks per pass on 256 CUs); 512 x 8: 95 us, 256 x 8: 124 us.  12-bit digits exceed the LDS.
constexpr unsigned SORT_BLOCK = 1024, SORT_ITEMS = 4, SORT_BITS = 8;
using SortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<SORT_BLOCK, SORT_ITEMS>,
                                        rocprim::kernel_config<SORT_BLOCK, SORT_ITEMS>, SORT_BITS,
                                        rocprim::block_radix_rank_algorithm::match>>;

size_t sort_pairs_temp_bytes(size_t n, int key_bits) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs<SortConfig>(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0u,
                                                unsigned(key_bits));
    return bytes;
}

hipError_t sort_pairs(const Launch& l, void* temp, size_t temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                      const uint32_t* vals_in, uint32_t* vals_out, size_t n, int key_bits) {
    if (!n) return hipSuccess;
    return rocprim::radix_sort_pairs<SortConfig>(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0u,
                                                 unsigned(key_bits), l.stream);
}

}  // namespace osph

// ---- Grady-Kipp flaws on the device ------------------------------------------------------------------

namespace osph {
namespace {

// Uniform particle index of flaw p (1-based) of a body of nb particles: splitmix64 of seed + p times the
// golden-ratio increment, mapped to [0, nb) by the high word of a 64 x 64-bit product.
__device__ __forceinline__ uint32_t flaw_target(uint64_t seed, uint64_t p, uint32_t nb) {
    uint64_t x = seed + p * 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return uint32_t(__umul64hi(x, uint64_t(nb)));
}

// Flaws p0 .. p0 + count - 1: the first flaw of each hit particle.
__global__ void __launch_bounds__(BLOCK) k_flaws_first(uint64_t seed, uint64_t p0, uint64_t count, uint32_t nb,
                                                       uint32_t* __restrict__ first) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint64_t p = p0 + t;
    atomicMin(&first[flaw_target(seed, p, nb)], uint32_t(p));
}

// Particles without a flaw and the largest first flaw (acc[0], acc[1]); one atomic per wavefront.
__global__ void __launch_bounds__(BLOCK) k_flaws_cover(const uint32_t* __restrict__ first, uint32_t nb,
                                                       uint32_t* __restrict__ acc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t f = i < nb ? first[i] : 0u;
    const bool open = i < nb && f == 0xffffffffu;
    uint32_t mx = open ? 0u : f;
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, uint32_t(__shfl_xor(int(mx), o, 64)));
    const unsigned long long ballot = __ballot(open);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t c = uint32_t(__popcll(ballot));
        if (c) atomicAdd(&acc[0], c);
        atomicMax(&acc[1], mx);
    }
}

// Flaws 1 .. K: count and last flaw of each particle.
__global__ void __launch_bounds__(BLOCK) k_flaws_count(uint64_t seed, uint32_t K, uint32_t nb,
                                                       uint32_t* __restrict__ cnt, uint32_t* __restrict__ last) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= K) return;
    const uint64_t p = t + 1;
    const uint32_t i = flaw_target(seed, p, nb);
    atomicAdd(&cnt[i], 1u);
    atomicMax(&last[i], uint32_t(p));
}

// ScalarGradyKippModel::setFlaws (core/physics/Damage.cpp:17-126), accumulate mode, per particle of the
// body: eps_min / eps_max of its first / last flaw, eps(p) = denom p^(1/m); m_zero = ln n / ln(eps_max /
// eps_min) for n > 1 flaws, else 1; growth = cg / (2 h).
__global__ void __launch_bounds__(BLOCK) k_flaws_finish(DevState s, const uint32_t* __restrict__ slot_of,
                                                        size_t body_first, uint32_t nb,
                                                        const uint32_t* __restrict__ first,
                                                        const uint32_t* __restrict__ last,
                                                        const uint32_t* __restrict__ cnt, double denom,
                                                        double inv_m, double cg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const uint32_t slot = slot_of[body_first + i];
    const double emin = denom * pow(double(first[i]), inv_m);
    const double emax = denom * pow(double(last[i]), inv_m);
    const uint32_t n = cnt[i];
    s.d[EPSMIN][slot] = emin;
    s.d[MZERO][slot] = n > 1 ? log(double(n)) / log(emax / emin) : 1.0;
    s.d[GROWTH][slot] = cg / (2.0 * s.d[H][slot]);
    s.u[NFLAWS][slot] = n;
}

}  // namespace

int flaws_first_pass(const Launch& l, uint64_t seed, uint64_t p0, uint64_t count, uint32_t nb, uint32_t* first) {
    if (!count) return 0;
    const uint64_t blocks = (count + BLOCK - 1) / BLOCK;
    k_flaws_first<<<unsigned(blocks), BLOCK, 0, l.stream>>>(seed, p0, count, nb, first);
    return 0;
}

void flaws_cover(const Launch& l, const uint32_t* first, uint32_t nb, uint32_t* acc) {
    k_fill_u32<<<1, 64, 0, l.stream>>>(acc, 2, 0u);
    k_flaws_cover<<<grid_for(nb), BLOCK, 0, l.stream>>>(first, nb, acc);
}

void flaws_count_pass(const Launch& l, uint64_t seed, uint32_t K, uint32_t nb, uint32_t* cnt, uint32_t* last) {
    k_flaws_count<<<unsigned((uint64_t(K) + BLOCK - 1) / BLOCK), BLOCK, 0, l.stream>>>(seed, K, nb, cnt, last);
}

void flaws_finish(const Launch& l, DevState s, const uint32_t* slot_of, size_t body_first, uint32_t nb,
                  const uint32_t* first, const uint32_t* last, const uint32_t* cnt, double denom, double inv_m,
                  double cg) {
    k_flaws_finish<<<grid_for(nb), BLOCK, 0, l.stream>>>(s, slot_of, body_first, nb, first, last, cnt, denom, inv_m,
                                                          cg);
}

}  // namespace osph
