// Initial conditions on the device: hexagonal close packing of spherical bodies and their placement for
// a two-body collision (osph_setup_spheres in osph_capi.hip).
//
// HexagonalPacking::generate (core/sph/initial/Distribution.cpp:126-201) walks the lattice of
// Box::iterateWithIndices (core/objects/geometry/Box.h:182-195) with z outermost and x innermost, shifts
// every other row and layer, and keeps the points inside the domain in that order.  Here the lattice
// point of linear index (k ny + j) nx + i is evaluated by one thread; a block keeps its inside points in
// index order at the offset given by the inside counts of the blocks before it, so the kept points come
// out in the reference's order.  The axis values themselves are accumulated on the host (v += step, as
// the reference loop does), and every operation on a coordinate is an explicitly rounded IEEE operation,
// so each position is bit-identical to the sequential construction.
//
// The centre of mass that HexagonalPacking::generate (CENTER) and getBoundingSphere
// (core/run/jobs/ParticleJobs.cpp:499-510) use is a left-to-right sum over the particles; k_serial_sums
// evaluates such sums in that order, one thread per sum.
#include "osph_device.h"
#include "osph_internal.h"

namespace osph {

namespace {

__device__ __forceinline__ void hcp_point(const HcpLattice& P, size_t idx, double& x, double& y, double& z) {
    const size_t i = idx % P.nx, t = idx / P.nx;
    const size_t j = t % P.ny, k = t / P.ny;
    x = P.xs[i];
    y = P.ys[j];
    z = P.zs[k];
    const bool even_k = (k & 1u) == 0;
    const bool shift_x = even_k ? (j & 1u) == 1 : (j & 1u) == 0;
    if (shift_x) x = rn_add(x, P.delta_x);
    if (!even_k) y = rn_add(y, P.delta_y);
}

// SphericalDomain::contains (core/objects/geometry/Domain.cpp): |v - centre|^2 <= radius^2, the squares
// summed left to right.
__device__ __forceinline__ bool hcp_inside(const HcpLattice& P, double x, double y, double z) {
    const double dx = rn_sub(x, P.cx), dy = rn_sub(y, P.cy), dz = rn_sub(z, P.cz);
    const double d2 = rn_add(rn_add(rn_mul(dx, dx), rn_mul(dy, dy)), rn_mul(dz, dz));
    return d2 <= P.r2;
}

__global__ void __launch_bounds__(BLOCK) k_hcp_count(HcpLattice P, uint32_t* __restrict__ block_count) {
    const size_t idx = size_t(blockIdx.x) * BLOCK + threadIdx.x;
    bool in = false;
    if (idx < P.points) {
        double x, y, z;
        hcp_point(P, idx, x, y, z);
        in = hcp_inside(P, x, y, z);
    }
    const int c = __syncthreads_count(in);
    if (threadIdx.x == 0) block_count[blockIdx.x] = uint32_t(c);
}

// Writes the inside points of block b at out_first + block_first[b] + (their rank inside the block).
__global__ void __launch_bounds__(BLOCK) k_hcp_fill(HcpLattice P, const uint32_t* __restrict__ block_first,
                                                    double* __restrict__ px, double* __restrict__ py,
                                                    double* __restrict__ pz, double* __restrict__ ph, double h) {
    __shared__ uint32_t wave_in[BLOCK / 64];
    const size_t idx = size_t(blockIdx.x) * BLOCK + threadIdx.x;
    double x = 0.0, y = 0.0, z = 0.0;
    bool in = false;
    if (idx < P.points) {
        hcp_point(P, idx, x, y, z);
        in = hcp_inside(P, x, y, z);
    }
    const unsigned long long ball = __ballot(in);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) wave_in[wave] = uint32_t(__popcll(ball));
    __syncthreads();
    uint32_t rank = uint32_t(__popcll(ball & ((1ull << lane) - 1ull)));
    for (uint32_t w = 0; w < wave; ++w) rank += wave_in[w];
    if (in) {
        const size_t o = size_t(block_first[blockIdx.x]) + rank;
        px[o] = x;
        py[o] = y;
        pz[o] = z;
        ph[o] = h;
    }
}

// One thread per job: out = a[0] + a[1] + ... + a[n-1] (w[i] * a[i] when w is set), added left to right
// as the reference's loops and numpy's cumsum add.
__global__ void __launch_bounds__(64) k_serial_sums(const SerialSum* __restrict__ jobs, int count,
                                                    double* __restrict__ out) {
    const int t = int(blockIdx.x) * 64 + int(threadIdx.x);
    if (t >= count) return;
    const SerialSum J = jobs[t];
    double s = 0.0;
    size_t i = 0;
    // loads are independent of the chain of additions, so eight are issued ahead of it
    for (; i + 8 <= J.n; i += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = J.w ? rn_mul(J.w[i + u], J.a[i + u]) : J.a[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) s = rn_add(s, v[u]);
    }
    for (; i < J.n; ++i) s = rn_add(s, J.w ? rn_mul(J.w[i], J.a[i]) : J.a[i]);
    out[t] = s;
}

// Largest squared distance from c of the points (x, y, z)[0, n), as an order-preserving integer.
__global__ void __launch_bounds__(BLOCK) k_max_dist2(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, size_t n, double cx, double cy,
                                                     double cz, unsigned long long* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x;
    unsigned long long v = 0;
    if (i < n) {
        const double dx = rn_sub(x[i], cx), dy = rn_sub(y[i], cy), dz = rn_sub(z[i], cz);
        v = ord_bits(rn_add(rn_add(rn_mul(dx, dx), rn_mul(dy, dy)), rn_mul(dz, dz)));
    }
    v = wave_max(v);
    if ((threadIdx.x & 63u) == 0 && v) atomicMax(out, v);
}

// Adds (tx, ty, tz) to the positions of [first, first + n) and fills the other quantities of a body:
// h = h_eta, mass, velocity, flag, material, density rho0, zero energy, unit stress reduction.
__global__ void __launch_bounds__(BLOCK) k_body_finish(DevState s, size_t first, size_t n, BodyFill f) {
    const size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x;
    if (k >= n) return;
    const size_t i = first + k;
    if (f.move) {
        s.d[X][i] = rn_add(s.d[X][i], f.t[0]);
        s.d[Y][i] = rn_add(s.d[Y][i], f.t[1]);
        s.d[Z][i] = rn_add(s.d[Z][i], f.t[2]);
    }
    if (f.fill) {
        s.d[H][i] = f.h;
        s.d[M][i] = f.m;
        s.d[VX][i] = f.v[0];
        s.d[VY][i] = f.v[1];
        s.d[VZ][i] = f.v[2];
        s.d[RHO][i] = f.rho;
        s.d[U][i] = 0.0;
        s.d[RED][i] = 1.0;
        s.u[FLAG][i] = f.flag;
        s.u[MAT][i] = f.mat;
    }
}

}  // namespace

unsigned hcp_blocks(const HcpLattice& P) { return unsigned((P.points + BLOCK - 1) / BLOCK); }

void launch_hcp_count(const Launch& l, const HcpLattice& P, uint32_t* block_count) {
    k_hcp_count<<<hcp_blocks(P), BLOCK, 0, l.stream>>>(P, block_count);
}

void launch_hcp_fill(const Launch& l, const HcpLattice& P, const uint32_t* block_first, DevState s, double h) {
    k_hcp_fill<<<hcp_blocks(P), BLOCK, 0, l.stream>>>(P, block_first, s.d[X], s.d[Y], s.d[Z], s.d[H], h);
}

void launch_serial_sums(const Launch& l, const SerialSum* jobs, int count, double* out) {
    if (count <= 0) return;
    k_serial_sums<<<unsigned((count + 63) / 64), 64, 0, l.stream>>>(jobs, count, out);
}

void launch_max_dist2(const Launch& l, const double* x, const double* y, const double* z, size_t n, const double c[3],
                      unsigned long long* out) {
    if (n == 0) return;
    k_max_dist2<<<grid_for(n), BLOCK, 0, l.stream>>>(x, y, z, n, c[0], c[1], c[2], out);
}

void launch_body_finish(const Launch& l, DevState s, size_t first, size_t n, const BodyFill& f) {
    if (n == 0) return;
    k_body_finish<<<grid_for(n), BLOCK, 0, l.stream>>>(s, first, n, f);
}

}  // namespace osph
