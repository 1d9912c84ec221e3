// Self-gravity of OpenSPH on gfx950 (MI355X): BarnesHut (core/gravity/BarnesHut.cpp) on a binary radix
// tree, BruteForceGravity (core/gravity/BruteForceGravity.h), and resident LeapFrog steps with the
// gravity-only HardSphereSolver (core/timestepping/TimeStepping.cpp:352-391, core/gravity/NBodySolver.cpp:97-117).
//
// Build (IGravity::build): the particles are sorted by the 63-bit Morton code of their position; the
// binary radix tree of the sorted codes (internal node i, 0 <= i < n-1, with one end of its range at
// sorted position i; particle k is node n-1+k) is built level-free in one pass, then the node moments
// are accumulated bottom-up.  Each node carries its box, centre of mass, opening radius
// r_open = 2/sqrt(3) |r_max| / theta (BarnesHut.cpp:411-414), and the traceless quadrupole and octupole
// about its centre of mass, combined from its two children with the parallel-axis theorem
// (BarnesHut.cpp:433-486, Moments.h:176-280).  A node of at most leaf_size particles is a leaf: it is
// never opened, and its particles are summed exactly when it is not approximated.
//
// Evaluation (IGravity::evalSelfGravity): every leaf (a node of <= leaf_size particles whose parent is
// larger) is a target.  Its interaction lists follow the reference's check-list walk (BarnesHut.cpp:212-312):
// a node whose opening ball does not overlap the target's box goes to the node list; a leaf whose ball
// overlaps goes to the particle list; other nodes are opened.  Nodes containing the target are always
// opened and the target's own pairs are summed separately, as the reference's walk never holds a node
// together with its ancestors.  The walk runs one lane per target over skip links (no stack); the sums
// run one wavefront per target, its lanes split between the target's particles and the list entries.
#include "osph_device.h"
#include "osph_gpu.h"
#include "osph_gravity.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace osph {
namespace {

constexpr uint32_t END = 0xffffffffu;
constexpr int LUT_ENTRIES = 40000;  // LutKernel<3>::NEntries (core/sph/kernel/Kernel.h:53)
// Bound on the depth of the radix tree: 63 bits of Morton code plus 31 bits of position tie-break.
constexpr int MAX_DEPTH = 96;

// Plane indices of the resident state (storage order).
enum GPlane : int { GX, GY, GZ, GH, GVX, GVY, GVZ, GVH, GAX, GAY, GAZ, GM, NGPLANES };

// Traceless moments about the node's centre of mass; q2 = { xx, xy, xz, yy, yz } and
// q3 = { xxx, xxy, xxz, xyy, xyz, yyy, yyz } (the other components follow from the vanishing traces).
struct NodeMom {
    double M;
    double q2[5];
    double q3[7];
    double pad[3];
};
static_assert(sizeof(NodeMom) == 128, "NodeMom is 128 bytes");

// What the walk reads of an internal node, in one 64-byte record.
struct WalkNode {
    double4 geo;  // centre of mass, r_open
    uint32_t first, count, left, skip;
    uint32_t pad[4];
};
static_assert(sizeof(WalkNode) == 64, "WalkNode is 64 bytes");

// The screened walk's record of an internal node, in FP32: the centre of mass as an offset from the root
// box's centre and r_open, both divided by the power of two 2^e at or above the root box's extent (so the
// offsets lie within [-1/2, 1/2]); count's top bit marks r_open == 0.
struct WalkNodeF {
    float4 geo;
    uint32_t first, count, left, skip;
};
static_assert(sizeof(WalkNodeF) == 32, "WalkNodeF is 32 bytes");
constexpr uint32_t WALK_ZERO_R = 0x80000000u;

struct Tree {
    uint32_t n;        // particles; internal nodes 0..n-2, particle k is node n-1+k
    uint32_t leaf;     // leaf size
    uint32_t* left;    // [n-1]
    uint32_t* right;   // [n-1]
    uint32_t* parent;  // [2n-1]
    uint32_t* skip;    // [2n-1]: the node after this subtree in depth-first order
    uint32_t* first;   // [n-1]
    uint32_t* count;   // [n-1]
    double4* geo;      // [n-1]: centre of mass, r_open
    double4* lo;       // [n-1]: box
    double4* hi;
    NodeMom* mom;      // [n-1]
    WalkNode* wn;      // [n-1]
    WalkNodeF* wnf;    // [n-1], for the screened walk
    double ctr[3];     // the root box's centre and 2^-e, the scale of wnf
    double invL;
    uint32_t* flag;    // [n-1]: node depths while the moments are built
    const double4* pos;  // sorted x, y, z, h
    const double* gm;    // sorted G m
    double thetaInv;
    uint32_t* error;     // bit 0: a climb exceeded MAX_DEPTH; bit 1: a walk exceeded 2n visits
};

struct Planes {
    double* d[NGPLANES];
};

__device__ __forceinline__ bool is_particle(const Tree& t, uint32_t id) { return id >= t.n - 1; }

// ---- build ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(BLOCK) k_bounds3(Planes s, unsigned long long* acc, size_t n) {
    unsigned long long v[6] = { ~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull };
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            const unsigned long long b = ord_bits(s.d[GX + k][i]);
            v[k] = b < v[k] ? b : v[k];
            v[3 + k] = b > v[3 + k] ? b : v[3 + k];
        }
    }
    __shared__ unsigned long long red[BLOCK / 64][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 6; ++k) {
        const unsigned long long r = k < 3 ? wave_min(v[k]) : wave_max(v[k]);
        if (lane == 0) red[wave][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        unsigned long long r = red[0][k];
        for (int w = 1; w < BLOCK / 64; ++w) r = k < 3 ? (red[w][k] < r ? red[w][k] : r) : (red[w][k] > r ? red[w][k] : r);
        if (k < 3) {
            atomicMin(&acc[k], r);
        } else {
            atomicMax(&acc[k], r);
        }
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void __launch_bounds__(BLOCK) k_morton(Planes s, double lo0, double lo1, double lo2, double scale,
                                                  uint64_t* keys, uint32_t* idx, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lo[3] = { lo0, lo1, lo2 };
    uint64_t key = 0;
    for (int k = 0; k < 3; ++k) {
        double q = floor((s.d[GX + k][i] - lo[k]) * scale);
        q = q < 0.0 ? 0.0 : (q > 2097151.0 ? 2097151.0 : q);
        key |= spread21(uint64_t(q)) << k;
    }
    keys[i] = key;
    idx[i] = uint32_t(i);
}

// The state moved into the slots of the Morton order: slot k of dst takes every plane of slot order[k]
// of src, and the particle's index in the caller's order (perm_in of that slot, or the slot itself while
// the state is in the caller's order) goes with it, into perm_out and its inverse slot_of.
__global__ void __launch_bounds__(BLOCK) k_permute_planes(Planes src, Planes dst, const uint32_t* __restrict__ order,
                                                          const uint32_t* __restrict__ perm_in,
                                                          uint32_t* __restrict__ perm_out,
                                                          uint32_t* __restrict__ slot_of, size_t n) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = order[k];
    for (int p = 0; p < NGPLANES; ++p) dst.d[p][k] = src.d[p][i];
    const uint32_t u = perm_in ? perm_in[i] : i;
    perm_out[k] = u;
    slot_of[u] = uint32_t(k);
}

__global__ void __launch_bounds__(BLOCK) k_sequence(uint32_t* __restrict__ p, size_t n) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = uint32_t(k);
}

// Particle indices in the caller's order to the slots that hold them.
__global__ void __launch_bounds__(BLOCK) k_to_slots(uint32_t* __restrict__ idx, const uint32_t* __restrict__ slot_of,
                                                    size_t n) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) idx[k] = slot_of[idx[k]];
}

__global__ void __launch_bounds__(BLOCK) k_gather(Planes s, const uint32_t* __restrict__ order, double G,
                                                  double4* __restrict__ pos, double* __restrict__ gm, size_t n) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = order[k];
    pos[k] = make_double4(s.d[GX][i], s.d[GY][i], s.d[GZ][i], s.d[GH][i]);
    gm[k] = G * s.d[GM][i];  // BarnesHut::build premultiplies the masses (BarnesHut.cpp:58-60)
}

// Common-prefix length of sorted codes i and j; equal codes are told apart by their positions.
__device__ __forceinline__ int delta(const uint64_t* __restrict__ keys, long long i, long long j, long long n) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz(uint32_t(i) ^ uint32_t(j));
    return __clzll(a ^ b);
}

// Internal node i of the binary radix tree of the sorted codes: its range, split and children.
__global__ void __launch_bounds__(BLOCK) k_radix_tree(Tree t, const uint64_t* __restrict__ keys) {
    const long long n = t.n;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, i, i + 1, n) - delta(keys, i, i - 1, n)) >= 0 ? 1 : -1;
    const int dmin_ = delta(keys, i, i - d, n);
    long long lmax = 2;
    while (delta(keys, i, i + lmax * d, n) > dmin_) lmax <<= 1;
    long long l = 0;
    for (long long s = lmax >> 1; s >= 1; s >>= 1) {
        if (delta(keys, i, i + (l + s) * d, n) > dmin_) l += s;
    }
    const long long j = i + l * d;
    const int dnode = delta(keys, i, j, n);
    long long split = 0;
    for (long long div = 2;; div <<= 1) {
        const long long s = (l + div - 1) / div;
        if (delta(keys, i, i + (split + s) * d, n) > dnode) split += s;
        if (s == 1) break;
    }
    const long long gamma = i + split * d + (d < 0 ? -1 : 0);
    const long long lo = i < j ? i : j, hi = i < j ? j : i;
    const uint32_t lc = uint32_t(lo == gamma ? (n - 1 + gamma) : gamma);
    const uint32_t rc = uint32_t(hi == gamma + 1 ? (n - 1 + gamma + 1) : gamma + 1);
    t.left[i] = lc;
    t.right[i] = rc;
    t.first[i] = uint32_t(lo);
    t.count[i] = uint32_t(hi - lo + 1);
    t.wn[i].first = uint32_t(lo);
    t.wn[i].count = uint32_t(hi - lo + 1);
    t.wn[i].left = lc;
    t.parent[lc] = uint32_t(i);
    t.parent[rc] = uint32_t(i);
    if (i == 0) t.parent[0] = END;
}

// Depth-first successor of each node: the right sibling of the nearest ancestor-or-self that is a left
// child, or END.  Only for the nodes a walk can visit: the root and the children of nodes of more than
// leaf_size particles (the nodes inside a leaf are never visited).
__global__ void __launch_bounds__(BLOCK) k_skip(Tree t) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = 2 * size_t(t.n) - 1;
    if (id >= total) return;
    if (id != 0 && t.count[t.parent[id]] <= t.leaf) return;
    uint32_t c = uint32_t(id), next = END;
    for (int level = 0; c != 0u && level < MAX_DEPTH; ++level) {
        const uint32_t p = t.parent[c];
        if (t.left[p] == c) {
            next = t.right[p];
            break;
        }
        c = p;
    }
    t.skip[id] = next;
    if (id + 1 < t.n) t.wn[id].skip = next;
}

struct Moments {
    double M, c[3], lo[3], hi[3], q2[5], q3[7];
};

__device__ void load_node(const Tree& t, uint32_t id, Moments& m) {
    if (is_particle(t, id)) {
        const double4 p = t.pos[id - (t.n - 1)];
        m.M = t.gm[id - (t.n - 1)];
        m.c[0] = m.lo[0] = m.hi[0] = p.x;
        m.c[1] = m.lo[1] = m.hi[1] = p.y;
        m.c[2] = m.lo[2] = m.hi[2] = p.z;
        for (int k = 0; k < 5; ++k) m.q2[k] = 0.0;
        for (int k = 0; k < 7; ++k) m.q3[k] = 0.0;
        return;
    }
    const double4 g = t.geo[id], lo = t.lo[id], hi = t.hi[id];
    const NodeMom& q = t.mom[id];
    m.M = q.M;
    m.c[0] = g.x;
    m.c[1] = g.y;
    m.c[2] = g.z;
    m.lo[0] = lo.x;
    m.lo[1] = lo.y;
    m.lo[2] = lo.z;
    m.hi[0] = hi.x;
    m.hi[1] = hi.y;
    m.hi[2] = hi.z;
    for (int k = 0; k < 5; ++k) m.q2[k] = q.q2[k];
    for (int k = 0; k < 7; ++k) m.q3[k] = q.q3[k];
}

// Adds the moments of child m (shifted by d = com_child - com_parent) to q2 and q3: the parallel-axis
// theorem for traceless moments (Moments.h:176-186, 262-280).  With t = q2 d:
//   q2 += q2c + M (d d - |d|^2/3 I)
//   q3 += q3c + M (d d d - |d|^2/5 sym(I d)) + sym(q2c d) - 2/5 sym(I t)
// where sym sums the three distinct index placements.
__device__ void add_shifted(const Moments& m, const double d[3], double q2[5], double q3[7]) {
    const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double M = m.M;
    q2[0] += m.q2[0] + M * (d[0] * d[0] - dd / 3.0);
    q2[1] += m.q2[1] + M * (d[0] * d[1]);
    q2[2] += m.q2[2] + M * (d[0] * d[2]);
    q2[3] += m.q2[3] + M * (d[1] * d[1] - dd / 3.0);
    q2[4] += m.q2[4] + M * (d[1] * d[2]);
    // full child quadrupole
    const double cxx = m.q2[0], cxy = m.q2[1], cxz = m.q2[2], cyy = m.q2[3], cyz = m.q2[4], czz = -cxx - cyy;
    const double Q[3][3] = { { cxx, cxy, cxz }, { cxy, cyy, cyz }, { cxz, cyz, czz } };
    double tv[3];
    for (int a = 0; a < 3; ++a) tv[a] = Q[a][0] * d[0] + Q[a][1] * d[1] + Q[a][2] * d[2];
    // independent components (i, j, k) of q3
    const int I3[7][3] = { { 0, 0, 0 }, { 0, 0, 1 }, { 0, 0, 2 }, { 0, 1, 1 }, { 0, 1, 2 }, { 1, 1, 1 }, { 1, 1, 2 } };
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const int i = I3[c][0], j = I3[c][1], k = I3[c][2];
        const double dij = i == j ? 1.0 : 0.0, dik = i == k ? 1.0 : 0.0, djk = j == k ? 1.0 : 0.0;
        const double red = d[i] * d[j] * d[k] - dd / 5.0 * (dij * d[k] + dik * d[j] + djk * d[i]);
        const double perm = Q[i][j] * d[k] + Q[i][k] * d[j] + Q[j][k] * d[i];
        const double permt = dij * tv[k] + dik * tv[j] + djk * tv[i];
        q3[c] += m.q3[c] + M * red + perm - 0.4 * permt;
    }
}

// Node `node` from its two children: box, centre of mass and opening radius (BarnesHut::buildInner,
// BarnesHut.cpp:433-468), moments by the parallel-axis theorem.
__device__ void combine(const Tree& t, uint32_t node) {
    Moments a, b;
    load_node(t, t.left[node], a);
    load_node(t, t.right[node], b);
    const double M = a.M + b.M;
    double c[3], lo[3], hi[3];
    for (int q = 0; q < 3; ++q) {
        lo[q] = dmin(a.lo[q], b.lo[q]);
        hi[q] = dmax(a.hi[q], b.hi[q]);
    }
    NodeMom out{};
    double r_open = 0.0;
    if (M == 0.0) {
        c[0] = c[1] = c[2] = 0.0;
    } else {
        for (int q = 0; q < 3; ++q) c[q] = (a.M * a.c[q] + b.M * b.c[q]) / M;
        double rm[3];
        for (int q = 0; q < 3; ++q) rm[q] = dmax(c[q] - lo[q], hi[q] - c[q]);
        r_open = 2.0 / sqrt(3.0) * t.thetaInv * sqrt(rm[0] * rm[0] + rm[1] * rm[1] + rm[2] * rm[2]);
        out.M = M;
        double da[3], db[3];
        for (int q = 0; q < 3; ++q) {
            da[q] = a.c[q] - c[q];
            db[q] = b.c[q] - c[q];
        }
        add_shifted(a, da, out.q2, out.q3);
        add_shifted(b, db, out.q2, out.q3);
    }
    t.geo[node] = make_double4(c[0], c[1], c[2], r_open);
    t.wn[node].geo = make_double4(c[0], c[1], c[2], r_open);
    t.lo[node] = make_double4(lo[0], lo[1], lo[2], 0.0);
    t.hi[node] = make_double4(hi[0], hi[1], hi[2], 0.0);
    t.mom[node] = out;
}

// Moments of a leaf (an internal node of at most leaf_size particles whose parent is larger, or the
// root when it is that small) from its particles, as BarnesHut::buildLeaf computes them
// (BarnesHut.cpp:401-437; computeMultipole, Moments.h:167-180): mass, box, centre of mass, r_open, and
// the traceless moments about the centre summed particle by particle.  The centre is accumulated
// relative to the first particle, so that a leaf of coincident particles has its centre on them exactly
// (and r_open 0).
// The internal nodes inside a leaf are never read by the walks or the evaluations and get no moments.
// LEAF_LANES lanes per entry of the leaf list (k_mark_leaves, compacted in Morton order), each summing
// every LEAF_LANES-th particle of the leaf, the partial sums then added across the group in a fixed
// order; single-particle leaves need nothing.
constexpr uint32_t LEAF_LANES = 8;
__device__ __forceinline__ double leaf_sum(double v) {
    for (int o = LEAF_LANES / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, LEAF_LANES);
    return v;
}
__device__ __forceinline__ double leaf_min(double v) {
    for (int o = LEAF_LANES / 2; o >= 1; o >>= 1) v = dmin(v, __shfl_xor(v, o, LEAF_LANES));
    return v;
}
__device__ __forceinline__ double leaf_max(double v) {
    for (int o = LEAF_LANES / 2; o >= 1; o >>= 1) v = dmax(v, __shfl_xor(v, o, LEAF_LANES));
    return v;
}
__global__ void __launch_bounds__(BLOCK) k_leaf_moments(Tree t, const uint32_t* __restrict__ leaves,
                                                        const uint32_t* __restrict__ nleaves) {
    const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t q = gid / LEAF_LANES;
    const uint32_t sub = uint32_t(gid % LEAF_LANES);
    // uniform over the group of LEAF_LANES lanes, so every shuffle below reads lanes of the same leaf
    if (q >= *nleaves) return;
    const uint32_t i = leaves[q];
    if (is_particle(t, i)) return;
    const uint32_t cnt = t.count[i];
    const uint32_t f = t.first[i];
    const double4 p0 = t.pos[f];
    double lo[3] = { p0.x, p0.y, p0.z }, hi[3] = { p0.x, p0.y, p0.z };
    double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (uint32_t j = f + sub; j < f + cnt; j += LEAF_LANES) {
        const double4 p = t.pos[j];
        const double m = t.gm[j];
        lo[0] = dmin(lo[0], p.x);
        lo[1] = dmin(lo[1], p.y);
        lo[2] = dmin(lo[2], p.z);
        hi[0] = dmax(hi[0], p.x);
        hi[1] = dmax(hi[1], p.y);
        hi[2] = dmax(hi[2], p.z);
        M += m;
        sx += m * (p.x - p0.x);
        sy += m * (p.y - p0.y);
        sz += m * (p.z - p0.z);
    }
    for (int k = 0; k < 3; ++k) {
        lo[k] = leaf_min(lo[k]);
        hi[k] = leaf_max(hi[k]);
    }
    M = leaf_sum(M);
    sx = leaf_sum(sx);
    sy = leaf_sum(sy);
    sz = leaf_sum(sz);
    NodeMom out{};
    double c[3] = { 0.0, 0.0, 0.0 }, r_open = 0.0;
    if (M != 0.0) {
        c[0] = p0.x + sx / M;
        c[1] = p0.y + sy / M;
        c[2] = p0.z + sz / M;
        double rm[3];
        for (int k = 0; k < 3; ++k) rm[k] = dmax(c[k] - lo[k], hi[k] - c[k]);
        r_open = 2.0 / sqrt(3.0) * t.thetaInv * sqrt(rm[0] * rm[0] + rm[1] * rm[1] + rm[2] * rm[2]);
        out.M = M;
        Moments pm{};
        for (uint32_t j = f + sub; j < f + cnt; j += LEAF_LANES) {
            const double4 p = t.pos[j];
            pm.M = t.gm[j];
            const double d[3] = { p.x - c[0], p.y - c[1], p.z - c[2] };
            add_shifted(pm, d, out.q2, out.q3);  // a point mass: only the shift terms
        }
        for (int k = 0; k < 5; ++k) out.q2[k] = leaf_sum(out.q2[k]);
        for (int k = 0; k < 7; ++k) out.q3[k] = leaf_sum(out.q3[k]);
    }
    if (sub != 0) return;
    t.geo[i] = make_double4(c[0], c[1], c[2], r_open);
    t.wn[i].geo = make_double4(c[0], c[1], c[2], r_open);
    t.lo[i] = make_double4(lo[0], lo[1], lo[2], 0.0);
    t.hi[i] = make_double4(hi[0], hi[1], hi[2], 0.0);
    t.mom[i] = out;
}

// Depth of every internal node above the leaves (more than leaf_size particles; the number of its
// ancestors, all of them such nodes too) and the node count per depth; the other nodes get NO_DEPTH.
// Their moments are then built one depth per launch, deepest first, from the leaves' moments: the
// children of a node are complete when the launch of its depth starts, with no fences between
// work-groups.  A bounded grid strides over the nodes, so that few blocks add their per-depth counts
// to the ~100 global counters.
constexpr uint32_t NO_DEPTH = 0xffffffffu;
constexpr unsigned LEVEL_BLOCKS = 1024;
__global__ void __launch_bounds__(BLOCK) k_depth(Tree t, uint32_t* __restrict__ depth, uint32_t* __restrict__ per_level) {
    __shared__ uint32_t hist[MAX_DEPTH + 1];
    for (int d = threadIdx.x; d <= MAX_DEPTH; d += blockDim.x) hist[d] = 0u;
    __syncthreads();
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i + 1 < t.n; i += stride) {
        if (t.count[i] <= t.leaf) {
            depth[i] = NO_DEPTH;
            continue;
        }
        uint32_t c = uint32_t(i), d = 0;
        while (c != 0u && d < uint32_t(MAX_DEPTH)) {
            c = t.parent[c];
            ++d;
        }
        if (c != 0u) {
            atomicOr(t.error, 1u);
            d = 0;
        }
        depth[i] = d;
        atomicAdd(&hist[d], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d <= MAX_DEPTH; d += blockDim.x) {
        if (hist[d]) atomicAdd(&per_level[d], hist[d]);
    }
}

// Scatter of the internal nodes above the leaves into per-depth buckets.  Each block of a bounded grid
// counts its strided share of the nodes per depth in LDS, reserves one range per depth, then places
// its nodes from LDS cursors (the order inside a bucket is free: each node's moments depend only on
// its children).
__global__ void __launch_bounds__(BLOCK) k_bucket(uint32_t n_internal, const uint32_t* __restrict__ depth,
                                                  uint32_t* __restrict__ cursor, uint32_t* __restrict__ nodes) {
    __shared__ uint32_t cnt[MAX_DEPTH + 1];
    __shared__ uint32_t base[MAX_DEPTH + 1];
    for (int d = threadIdx.x; d <= MAX_DEPTH; d += blockDim.x) cnt[d] = 0u;
    __syncthreads();
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t i0 = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t i = i0; i < n_internal; i += stride) {
        const uint32_t d = depth[i];
        if (d != NO_DEPTH) atomicAdd(&cnt[d], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k <= MAX_DEPTH; k += blockDim.x) {
        base[k] = cnt[k] ? atomicAdd(&cursor[k], cnt[k]) : 0u;
        cnt[k] = 0u;
    }
    __syncthreads();
    for (size_t i = i0; i < n_internal; i += stride) {
        const uint32_t d = depth[i];
        if (d != NO_DEPTH) nodes[base[d] + atomicAdd(&cnt[d], 1u)] = uint32_t(i);
    }
}

__global__ void __launch_bounds__(BLOCK) k_moments_level(Tree t, const uint32_t* __restrict__ nodes, uint32_t count) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= count) return;
    combine(t, nodes[k]);
}

// Leaves (targets): nodes of at most leaf_size particles whose parent is larger, and particles whose
// parent is larger.  leaf_at[first] = node id.
__global__ void __launch_bounds__(BLOCK) k_mark_leaves(Tree t, uint32_t* leaf_at) {
    const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = 2 * size_t(t.n) - 1;
    if (id >= total) return;
    const uint32_t p = id == 0 ? END : t.parent[id];
    const uint32_t pc = p == END ? 0xffffffffu : t.count[p];
    if (is_particle(t, uint32_t(id))) {
        if (pc > t.leaf) leaf_at[id - (t.n - 1)] = uint32_t(id);
    } else if (t.count[id] <= t.leaf && pc > t.leaf) {
        leaf_at[t.first[id]] = uint32_t(id);
    }
}

struct NotEnd {
    __host__ __device__ bool operator()(uint32_t v) const { return v != END; }
};

// Sphere::overlaps (core/objects/geometry/Sphere.h:73-79): the ball (c, r) and the box share a point.
__device__ __forceinline__ bool ball_overlaps_box(double4 g, const double lo[3], const double hi[3]) {
#pragma clang fp contract(off)
    const double c[3] = { g.x, g.y, g.z };
    double l2 = 0.0, r2 = 0.0;
    double lv[3], rv[3];
    for (int k = 0; k < 3; ++k) {
        lv[k] = dmax(lo[k] - c[k], 0.0);
        rv[k] = dmax(c[k] - hi[k], 0.0);
    }
    l2 = lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2];
    r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
    return g.w * g.w - l2 - r2 > 0.0;
}

__device__ __forceinline__ void node_range(const Tree& t, uint32_t id, uint32_t& f, uint32_t& c) {
    if (is_particle(t, id)) {
        f = id - (t.n - 1);
        c = 1;
    } else {
        f = t.first[id];
        c = t.count[id];
    }
}

// One lane's interaction list, written 8 entries (32 B) at a time.  Entries wait in registers until a
// group of 8 is complete, so every store writes whole 32-byte sectors: single 4-byte appends to a row
// per lane left partly written lines in L2, which the node-record reads evicted (measured 7x the list
// bytes written to HBM at 8M particles).  Entries from the capacity on are counted, not stored (the host
// grows the lists and walks again); capacities are multiples of 8.
struct ListBuf8 {
    uint32_t* row;
    uint32_t cap, n = 0;
    uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0, b5 = 0, b6 = 0, b7 = 0;
    __device__ __forceinline__ ListBuf8(uint32_t* r, uint32_t c) : row(r), cap(c) {}
    __device__ __forceinline__ void push(uint32_t id) {
        // shift in: after 8 pushes b0..b7 hold the group in order
        b0 = b1;
        b1 = b2;
        b2 = b3;
        b3 = b4;
        b4 = b5;
        b5 = b6;
        b6 = b7;
        b7 = id;
        ++n;
        if ((n & 7u) == 0 && n <= cap) {
            uint4* dst = reinterpret_cast<uint4*>(row + (n - 8));
            dst[0] = make_uint4(b0, b1, b2, b3);
            dst[1] = make_uint4(b4, b5, b6, b7);
        }
    }
    // the last n % 8 entries (b8-r .. b7)
    __device__ __forceinline__ void flush() {
        const uint32_t r = n & 7u;
        if (r == 0 || n > cap) return;
        const uint32_t base = n - r;
        const uint32_t v[8] = { b0, b1, b2, b3, b4, b5, b6, b7 };
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            if (u < r) row[base + u] = v[8 - r + u];
        }
    }
};

// Interaction lists of one target per lane (BarnesHut::evalNode, BarnesHut.cpp:212-312, applied to the
// target leaf's box).  Lists of target q live at nlist[q * cap_n ...] and xlist[q * cap_x ...]; counts
// beyond the capacity are reported so that the host can grow the lists and walk again.
__global__ void __launch_bounds__(BLOCK) k_walk(Tree t, const uint32_t* __restrict__ targets, uint32_t ntargets,
                                                uint32_t* __restrict__ nlist, uint32_t cap_n,
                                                uint32_t* __restrict__ xlist, uint32_t cap_x,
                                                uint32_t* __restrict__ ncount, uint32_t* __restrict__ xcount,
                                                uint32_t* __restrict__ work, uint32_t* need) {
    // each XCD walks one contiguous range of the (Morton-ordered) targets, so the node records that
    // neighbouring targets share stay in its L2
    const uint32_t q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= ntargets) return;
    const uint32_t T = targets[q];
    uint32_t tf, tc;
    node_range(t, T, tf, tc);
    double lo[3], hi[3];
    if (is_particle(t, T)) {
        const double4 p = t.pos[tf];
        lo[0] = hi[0] = p.x;
        lo[1] = hi[1] = p.y;
        lo[2] = hi[2] = p.z;
    } else {
        const double4 a = t.lo[T], b = t.hi[T];
        lo[0] = a.x;
        lo[1] = a.y;
        lo[2] = a.z;
        hi[0] = b.x;
        hi[1] = b.y;
        hi[2] = b.z;
    }
    uint32_t xsum = 0;
    ListBuf8 nl(nlist + size_t(q) * cap_n, cap_n), xl(xlist + size_t(q) * cap_x, cap_x);
    uint32_t id = 0;  // the root: internal node 0, or particle 0 when n == 1
    // a depth-first walk visits each of the 2n - 1 nodes at most once
    for (size_t visits = 0; id != END; ++visits) {
        if (visits > 2 * size_t(t.n)) {
            atomicOr(t.error, 2u);
            break;
        }
        if (id == T) {
            id = t.skip[id];
            continue;
        }
        if (is_particle(t, id)) {
            // a single-particle leaf has r_open = 0: always summed pair by pair (BarnesHut.cpp:237-243)
            xl.push(id);
            ++xsum;
            id = t.skip[id];
            continue;
        }
        const WalkNode w = t.wn[id];
        if (w.first <= tf && tf + tc <= w.first + w.count) {  // an ancestor of the target
            id = w.left;
            continue;
        }
        const bool is_leaf = w.count <= t.leaf;
        const double4 g = w.geo;
        if (is_leaf && g.w == 0.0) {
            xl.push(id);
            xsum += w.count;
            id = w.skip;
        } else if (g.w == 0.0) {
            // more than leaf_size coincident particles: no opening ball; their expansion is singular
            // for a target at the same point, so the node is opened down to its leaves
            id = w.left;
        } else if (!ball_overlaps_box(g, lo, hi)) {
            nl.push(id);
            id = w.skip;
        } else if (is_leaf) {
            xl.push(id);
            xsum += w.count;
            id = w.skip;
        } else {
            id = w.left;
        }
    }
    nl.flush();
    xl.flush();
    const uint32_t nn = nl.n, nx = xl.n;
    ncount[q] = nn;
    xcount[q] = nx;
    work[2 * size_t(q)] = tc * nn;                        // particle-node expansions
    work[2 * size_t(q) + 1] = tc * xsum + tc * (tc - 1);  // particle pairs
    if (nn > cap_n) atomicMax(&need[0], nn);
    if (nx > cap_x) atomicMax(&need[1], nx);
}

// The target's box: its leaf's box, or the particle itself.
__device__ __forceinline__ void target_box(const Tree& t, uint32_t T, uint32_t tf, double lo[3], double hi[3]) {
    if (is_particle(t, T)) {
        const double4 p = t.pos[tf];
        lo[0] = hi[0] = p.x;
        lo[1] = hi[1] = p.y;
        lo[2] = hi[2] = p.z;
    } else {
        const double4 a = t.lo[T], b = t.hi[T];
        lo[0] = a.x;
        lo[1] = a.y;
        lo[2] = a.z;
        hi[0] = b.x;
        hi[1] = b.y;
        hi[2] = b.z;
    }
}

__global__ void __launch_bounds__(BLOCK) k_walk_records(Tree t) {
    const uint32_t i = uint32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i + 1 >= t.n) return;
    const WalkNode w = t.wn[i];
    WalkNodeF f;
    f.geo = make_float4(float((w.geo.x - t.ctr[0]) * t.invL), float((w.geo.y - t.ctr[1]) * t.invL),
                        float((w.geo.z - t.ctr[2]) * t.invL), float(w.geo.w * t.invL));
    f.first = w.first;
    f.count = w.count | (w.geo.w == 0.0 ? WALK_ZERO_R : 0u);
    f.left = w.left;
    f.skip = w.skip;
    t.wnf[i] = f;
}

// k_walk with the opening test screened in FP32: the same interaction lists, in the same order, from
// 32-byte node records instead of 64-byte ones.  In units of the root box (offsets within [-1/2, 1/2]),
// the rounding of the box, the centre and r_open to FP32 moves S = |box - centre|^2 by at most
// 7 delta sqrt(S) (delta = 2^-24) and the FP32 arithmetic adds a few ulps of Q = r_open^2 + S; where
// |r_open^2 - S| is within 16 delta sqrt(Q) + 2e-6 Q, the FP64 test of k_walk decides.
__global__ void __launch_bounds__(BLOCK) k_walk_screened(Tree t, const uint32_t* __restrict__ targets,
                                                         uint32_t ntargets, uint32_t* __restrict__ nlist,
                                                         uint32_t cap_n, uint32_t* __restrict__ xlist, uint32_t cap_x,
                                                         uint32_t* __restrict__ ncount, uint32_t* __restrict__ xcount,
                                                         uint32_t* __restrict__ work, uint32_t* need) {
    // each XCD walks one contiguous range of the (Morton-ordered) targets, so the node records that
    // neighbouring targets share stay in its L2
    const uint32_t q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= ntargets) return;
    const uint32_t T = targets[q];
    uint32_t tf, tc;
    node_range(t, T, tf, tc);
    float lof[3], hif[3];
    {
        double lo[3], hi[3];
        target_box(t, T, tf, lo, hi);
        for (int k = 0; k < 3; ++k) {
            lof[k] = float((lo[k] - t.ctr[k]) * t.invL);
            hif[k] = float((hi[k] - t.ctr[k]) * t.invL);
        }
    }
    uint32_t xsum = 0;
    ListBuf8 nl(nlist + size_t(q) * cap_n, cap_n), xl(xlist + size_t(q) * cap_x, cap_x);
    uint32_t id = 0;
    for (size_t visits = 0; id != END; ++visits) {
        if (visits > 2 * size_t(t.n)) {
            atomicOr(t.error, 2u);
            break;
        }
        if (id == T) {
            id = t.skip[id];
            continue;
        }
        if (is_particle(t, id)) {
            xl.push(id);
            ++xsum;
            id = t.skip[id];
            continue;
        }
        const WalkNodeF w = t.wnf[id];
        const uint32_t cnt = w.count & ~WALK_ZERO_R;
        if (w.first <= tf && tf + tc <= w.first + cnt) {  // an ancestor of the target
            id = w.left;
            continue;
        }
        const bool is_leaf = cnt <= t.leaf;
        const bool zero_r = (w.count & WALK_ZERO_R) != 0u;
        if (is_leaf && zero_r) {
            xl.push(id);
            xsum += cnt;
            id = w.skip;
            continue;
        }
        if (zero_r) {
            id = w.left;
            continue;
        }
        float S = 0.f;
        {
            const float c[3] = { w.geo.x, w.geo.y, w.geo.z };
            for (int k = 0; k < 3; ++k) {
                const float l = fmaxf(lof[k] - c[k], 0.f), r = fmaxf(c[k] - hif[k], 0.f);
                S += l * l + r * r;
            }
        }
        const float R2 = w.geo.w * w.geo.w;
        const float m = R2 - S, Q = R2 + S;
        bool overlaps;
        if (fabsf(m) > 1e-6f * __builtin_amdgcn_sqrtf(Q) + 2e-6f * Q) {
            overlaps = m > 0.f;
        } else {
            double lo[3], hi[3];
            target_box(t, T, tf, lo, hi);
            overlaps = ball_overlaps_box(t.geo[id], lo, hi);
        }
        if (!overlaps) {
            nl.push(id);
            id = w.skip;
        } else if (is_leaf) {
            xl.push(id);
            xsum += cnt;
            id = w.skip;
        } else {
            id = w.left;
        }
    }
    nl.flush();
    xl.flush();
    const uint32_t nn = nl.n, nx = xl.n;
    ncount[q] = nn;
    xcount[q] = nx;
    work[2 * size_t(q)] = tc * nn;
    work[2 * size_t(q) + 1] = tc * xsum + tc * (tc - 1);
    if (nn > cap_n) atomicMax(&need[0], nn);
    if (nx > cap_x) atomicMax(&need[1], nx);
}

// ---- evaluation ---------------------------------------------------------------------------------------

// GravityLutKernel::grad (core/sph/kernel/GravityKernel.h:60-70) times the source mass, added to a.
// grads == nullptr is the point-particle kernel (radius 0).  REF evaluates it as the reference does
// (the direct sum is then bit-identical to it); otherwise the tree sums use the cheaper form below.
template <bool REF>
__device__ __forceinline__ void add_pair(const double* __restrict__ grads, double rx, double ry, double rz, double h,
                                         double m, double a[3]) {
#pragma clang fp contract(off)
    if (!REF) {
        // the same branches, decided on |r|^2 against (4 - EPS) h^2 without dividing by h, and the
        // Newtonian branch with one reciprocal square root
        const double d2 = rx * rx + ry * ry + rz * rz;
        if (grads && d2 < (4.0 - EPS_D) * (h * h)) {
            const double hInv = 1.0 / h;
            const double qSqr = d2 * (hInv * hInv);
            const double fi = double(LUT_ENTRIES / 4) * qSqr;
            const uint32_t i1 = uint32_t(fi);
            const double ratio = fi - double(i1);
            const double gr = grads[i1] * (1.0 - ratio) + grads[i1 + 1] * ratio;
            const double f = m * (hInv * hInv * hInv) * gr;
            a[0] += rx * f;
            a[1] += ry * f;
            a[2] += rz * f;
            return;
        }
        const double il = rsqrt(d2);
        const double f = m * (il * il * il);
        a[0] += rx * f;
        a[1] += ry * f;
        a[2] += rz * f;
        return;
    }
    if (grads) {
        const double hInv = 1.0 / h;
        const double sx = rx * hInv, sy = ry * hInv, sz = rz * hInv;
        const double qSqr = sx * sx + sy * sy + sz * sz;
        if (!(qSqr + EPS_D >= 4.0)) {
            // LutKernel<3>::gradImpl (core/sph/kernel/Kernel.h:129-144)
            const double fi = double(LUT_ENTRIES / 4) * qSqr;
            const uint32_t i1 = uint32_t(fi);
            const double ratio = fi - double(i1);
            const double gr = grads[i1] * (1.0 - ratio) + grads[i1 + 1] * ratio;
            const double h3 = hInv * hInv * hInv;
            a[0] += m * ((h3 * rx) * gr);
            a[1] += m * ((h3 * ry) * gr);
            a[2] += m * ((h3 * rz) * gr);
            return;
        }
    }
    const double l = sqrt(rx * rx + ry * ry + rz * rz);
    const double l3 = l * l * l;
    a[0] += m * (rx / l3);
    a[1] += m * (ry / l3);
    a[2] += m * (rz / l3);
}

// evaluateGravity (core/gravity/Moments.h:318-360) of node moments at dr = r - com, added to a.
__device__ __forceinline__ void add_multipole(const NodeMom& q, double drx, double dry, double drz, int order,
                                              double a[3]) {
    const double il = rsqrt(drx * drx + dry * dry + drz * drz);
    const double inv = il * il;
    const double g0 = -il;
    const double g1 = -1.0 * inv * g0;
    const double g2 = -3.0 * inv * g1;
    const double g3 = -5.0 * inv * g2;
    const double g4 = -7.0 * inv * g3;
    const double d0 = -drx, d1 = -dry, d2 = -drz;
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (order >= 3) {
        const double xxx = q.q3[0], xxy = q.q3[1], xxz = q.q3[2], xyy = q.q3[3], xyz = q.q3[4], yyy = q.q3[5],
                     yyz = q.q3[6];
        const double xzz = -xxx - xyy, yzz = -xxy - yyy, zzz = -xxz - yyz;
        // Q1_i = q_ijk d_j d_k
        const double Q1x = xxx * d0 * d0 + xyy * d1 * d1 + xzz * d2 * d2 + 2.0 * (xxy * d0 * d1 + xxz * d0 * d2 + xyz * d1 * d2);
        const double Q1y = xxy * d0 * d0 + yyy * d1 * d1 + yzz * d2 * d2 + 2.0 * (xyy * d0 * d1 + xyz * d0 * d2 + yyz * d1 * d2);
        const double Q1z = xxz * d0 * d0 + yyz * d1 * d1 + zzz * d2 * d2 + 2.0 * (xyz * d0 * d1 + xzz * d0 * d2 + yzz * d1 * d2);
        const double Q0 = (Q1x * d0 + Q1y * d1 + Q1z * d2) * (1.0 / 6.0);
        ax += g4 * d0 * Q0 + g3 * (0.5 * Q1x);
        ay += g4 * d1 * Q0 + g3 * (0.5 * Q1y);
        az += g4 * d2 * Q0 + g3 * (0.5 * Q1z);
    }
    if (order >= 2) {
        const double xx = q.q2[0], xy = q.q2[1], xz = q.q2[2], yy = q.q2[3], yz = q.q2[4], zz = -xx - yy;
        const double Q1x = xx * d0 + xy * d1 + xz * d2;
        const double Q1y = xy * d0 + yy * d1 + yz * d2;
        const double Q1z = xz * d0 + yz * d1 + zz * d2;
        const double Q0 = 0.5 * (Q1x * d0 + Q1y * d1 + Q1z * d2);
        ax += g3 * d0 * Q0 + g2 * Q1x;
        ay += g3 * d1 * Q0 + g2 * Q1y;
        az += g3 * d2 * Q0 + g2 * Q1z;
    }
    ax += g1 * d0 * q.M;
    ay += g1 * d1 * q.M;
    az += g1 * d2 * q.M;
    a[0] += ax;
    a[1] += ay;
    a[2] += az;
}

// Per-wavefront staging area of k_eval: 64 node expansions (SoA) or 64 source particles.
// Moments of the orders below ORDER take no LDS (one unused row keeps the arrays legal): the stage's size
// bounds the wavefronts per CU.
template <int ORDER>
struct EvalStage {
    double c[3][64];
    double M[64];
    double q2[ORDER >= 2 ? 5 : 1][64];
    double q3[ORDER >= 3 ? 7 : 1][64];
    double4 src[64];
    double sm[64];
    uint32_t sidx[64];
};

// Lanes of a target leaf's wavefront: lane = sp * tc + p, for particle p < tc of the leaf and split
// sp < S = 64 / tc.  Lanes from S * tc on take no entries.  (Splits of a power-of-two width would leave
// up to half the lanes idle for leaves of 17 particles; a 25-particle leaf limit makes most leaves
// 9-25 particles.)
struct LaneSplit {
    uint32_t S, p, sp;
    bool used;
};

__device__ __forceinline__ LaneSplit lane_split(uint32_t lane, uint32_t tc) {
    LaneSplit l;
    l.S = __builtin_amdgcn_readfirstlane(64u / tc);
    l.sp = lane / tc;
    l.p = lane - l.sp * tc;
    l.used = l.sp < l.S;
    return l;
}

// Sums the splits' partial accelerations into the lanes of split 0: split sp adds split sp + d for
// d = 1, 2, 4, ... when sp is a multiple of 2d.
__device__ __forceinline__ void sum_splits(const LaneSplit& l, uint32_t lane, uint32_t tc, double a[3]) {
    for (uint32_t d = 1; d < l.S; d <<= 1) {
        const int src = int(min(lane + d * tc, 63u));
        const bool take = l.used && (l.sp & (2 * d - 1)) == 0 && l.sp + d < l.S;
        for (int c = 0; c < 3; ++c) {
            const double v = __shfl(a[c], src, 64);
            if (take) a[c] += v;
        }
    }
}

// One wavefront per target leaf, its lanes split as lane_split says.  The node list and the source
// particles (every listed leaf followed by the target leaf itself) are staged 64 at a time in the
// wavefront's LDS with coalesced loads; split sp then takes every S-th staged entry, and the splits are
// summed across lanes at the end.  Accelerations are added to the state planes.
template <int ORDER>
__global__ void __launch_bounds__(BLOCK) k_eval(Tree t, Planes s, const uint32_t* __restrict__ order,
                                                const uint32_t* __restrict__ targets, uint32_t ntargets,
                                                const uint32_t* __restrict__ nlist, uint32_t cap_n,
                                                const uint32_t* __restrict__ xlist, uint32_t cap_x,
                                                const uint32_t* __restrict__ ncount,
                                                const uint32_t* __restrict__ xcount, const double* __restrict__ grads) {
    __shared__ EvalStage<ORDER> stage[BLOCK / 64];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    EvalStage<ORDER>& S_ = stage[wave];
    const uint32_t q = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / 64)) + wave;
    if (q >= ntargets) return;  // uniform over the wavefront
    const uint32_t T = targets[q];
    uint32_t tf, tc;
    node_range(t, T, tf, tc);
    const LaneSplit ls = lane_split(lane, tc);
    const uint32_t S = ls.S, sp = ls.sp;
    const bool active = ls.used;
    const uint32_t j0 = active ? sp : 64u;  // unused lanes take no entries
    const uint32_t k = tf + ls.p;
    const double4 me = t.pos[k];
    double a[3] = { 0.0, 0.0, 0.0 };
    const uint32_t nn = min(ncount[q], cap_n), nx = min(xcount[q], cap_x);
    const uint32_t* nl = nlist + size_t(q) * cap_n;
    const uint32_t* xl = xlist + size_t(q) * cap_x;

    // node list
    for (uint32_t base = 0; base < nn; base += 64) {
        const uint32_t cnt = min(64u, nn - base);
        if (lane < cnt) {
            const uint32_t id = nl[base + lane];
            const double4 g = t.geo[id];
            const NodeMom& mm = t.mom[id];
            S_.c[0][lane] = g.x;
            S_.c[1][lane] = g.y;
            S_.c[2][lane] = g.z;
            S_.M[lane] = mm.M;
            if constexpr (ORDER >= 2) {
                for (int c = 0; c < 5; ++c) S_.q2[c][lane] = mm.q2[c];
            }
            if constexpr (ORDER >= 3) {
                for (int c = 0; c < 7; ++c) S_.q3[c][lane] = mm.q3[c];
            }
        }
        wave_sync();
        for (uint32_t j = j0; j < cnt; j += S) {
            NodeMom mm;
            mm.M = S_.M[j];
            if constexpr (ORDER >= 2) {
                for (int c = 0; c < 5; ++c) mm.q2[c] = S_.q2[c][j];
            }
            if constexpr (ORDER >= 3) {
                for (int c = 0; c < 7; ++c) mm.q3[c] = S_.q3[c][j];
            }
            add_multipole(mm, me.x - S_.c[0][j], me.y - S_.c[1][j], me.z - S_.c[2][j], ORDER, a);
        }
        wave_sync();
    }

    // source particles: entries e < nx are the listed leaves, entry nx is the target leaf
    const uint32_t nent = nx + 1;
    for (uint32_t eb = 0; eb < nent; eb += 64) {
        // entry eb + lane: its range and the exclusive prefix of the counts
        uint32_t f = 0, c = 0;
        const uint32_t e = eb + lane;
        if (e < nx) {
            node_range(t, xl[e], f, c);
        } else if (e == nx) {
            f = tf;
            c = tc;
        }
        uint32_t incl = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, 64);
            if (lane >= uint32_t(o)) incl += v;
        }
        const uint32_t excl = incl - c;
        const uint32_t total = __shfl(incl, 63, 64);
        for (uint32_t s0 = 0; s0 < total; s0 += 64) {
            const uint32_t slot = s0 + lane;
            const uint32_t cnt = min(64u, total - s0);
            // the entry holding this slot: the last lane whose exclusive prefix is <= slot (all lanes take
            // part in the cross-lane reads; only lanes with a slot load)
            int lo = 0;
            for (int step = 32; step > 0; step >>= 1) {
                const uint32_t ex = __shfl(excl, (lo + step) & 63, 64);
                if (lo + step < 64 && ex <= slot) lo += step;
            }
            const uint32_t fe = __shfl(f, lo, 64), xe = __shfl(excl, lo, 64);
            if (lane < cnt) {
                const uint32_t src = fe + (slot - xe);
                S_.src[lane] = t.pos[src];
                S_.sm[lane] = t.gm[src];
                S_.sidx[lane] = src;
            }
            wave_sync();
            for (uint32_t j = j0; j < cnt; j += S) {
                if (S_.sidx[j] == k) continue;
                const double4 o = S_.src[j];
                add_pair<false>(grads, o.x - me.x, o.y - me.y, o.z - me.z, 0.5 * (o.w + me.w), S_.sm[j], a);
            }
            wave_sync();
        }
    }
    sum_splits(ls, lane, tc, a);
    if (sp == 0) {
        const uint32_t i = order[k];
        s.d[GAX][i] += a[0];
        s.d[GAY][i] += a[1];
        s.d[GAZ][i] += a[2];
    }
}

// ---- mixed-precision evaluation ---------------------------------------------------------------------

// Largest value over the wavefront (every lane gets it).
__device__ __forceinline__ float wave_max_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Powers of two that bring lengths and G m into FP32's comfortable range: lengths are divided by
// 2^e with the root box extent in [2^(e-1), 2^e), G m by 2^f with the total G m in [2^(f-1), 2^f).
// Both divisions are exact in FP64; the results are scaled back by 2^(f - 2e).
struct EvalScale {
    double invL, invM, back;
};

// A wave-uniform double kept in scalar registers.
__device__ __forceinline__ double uniform_d(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(b)), hi = __builtin_amdgcn_readfirstlane(uint32_t(b >> 32));
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
}

__device__ __forceinline__ EvalScale eval_scale(const Tree& t) {
    EvalScale sc{ 1.0, 1.0, 1.0 };
    if (t.n < 2) return sc;
    const double4 lo = t.lo[0], hi = t.hi[0];
    const double ext = fmax(fmax(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
    int e = 0, f = 0;
    if (ext > 0.0) (void)frexp(ext, &e);
    const double M = t.mom[0].M;
    if (M > 0.0) (void)frexp(M, &f);
    sc.invL = uniform_d(ldexp(1.0, -e));
    sc.invM = uniform_d(ldexp(1.0, -f));
    sc.back = uniform_d(ldexp(1.0, f - 2 * e));
    return sc;
}

// The multipole expansion of add_multipole in FP32, written so that no intermediate leaves FP32's range:
// with d the vector from the target to the node's centre of mass, il = 1/|d|, u = d il and w = il^2,
//   monopole   il^3 M d                                  = w M u
//   quadrupole il^4 (7.5 u (u.q2.u) - 3 q2.u)            = w^2 t2
//   octupole   il^5 (7.5 (q3:uu) - 17.5 u (q3:uuu))      = w^2 il t3
// and an approximated node has size < theta |d|, so w |q2| < theta^2 M and w il |q3| < theta^3 M.
template <int ORDER>
__device__ __forceinline__ void add_multipole_f(float M, const float* q2, const float* q3, float dx, float dy, float dz,
                                                float a[3]) {
    const float il = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
    const float w = il * il;
    const float u0 = dx * il, u1 = dy * il, u2 = dz * il;
    float hx = 0.f, hy = 0.f, hz = 0.f;
    if constexpr (ORDER >= 2) {
        const float xx = q2[0], xy = q2[1], xz = q2[2], yy = q2[3], yz = q2[4], zz = -xx - yy;
        const float Q1x = xx * u0 + xy * u1 + xz * u2;
        const float Q1y = xy * u0 + yy * u1 + yz * u2;
        const float Q1z = xz * u0 + yz * u1 + zz * u2;
        const float Q0 = 7.5f * (Q1x * u0 + Q1y * u1 + Q1z * u2);
        hx = Q0 * u0 - 3.f * Q1x;
        hy = Q0 * u1 - 3.f * Q1y;
        hz = Q0 * u2 - 3.f * Q1z;
    }
    if constexpr (ORDER >= 3) {
        const float xxx = q3[0], xxy = q3[1], xxz = q3[2], xyy = q3[3], xyz = q3[4], yyy = q3[5], yyz = q3[6];
        const float xzz = -xxx - xyy, yzz = -xxy - yyy, zzz = -xxz - yyz;
        const float Q1x = xxx * u0 * u0 + xyy * u1 * u1 + xzz * u2 * u2 + 2.f * (xxy * u0 * u1 + xxz * u0 * u2 + xyz * u1 * u2);
        const float Q1y = xxy * u0 * u0 + yyy * u1 * u1 + yzz * u2 * u2 + 2.f * (xyy * u0 * u1 + xyz * u0 * u2 + yyz * u1 * u2);
        const float Q1z = xxz * u0 * u0 + yyz * u1 * u1 + zzz * u2 * u2 + 2.f * (xyz * u0 * u1 + xzz * u0 * u2 + yzz * u1 * u2);
        const float Q0 = 17.5f * (Q1x * u0 + Q1y * u1 + Q1z * u2);
        hx += il * (7.5f * Q1x - Q0 * u0);
        hy += il * (7.5f * Q1y - Q0 * u1);
        hz += il * (7.5f * Q1z - Q0 * u2);
    }
    a[0] += w * (M * u0 + w * hx);
    a[1] += w * (M * u1 + w * hy);
    a[2] += w * (M * u2 + w * hz);
}

// add_pair<false> in FP32 on scaled coordinates (r = source - target with d2 = |r|^2, h and m scaled).
// The LUT index is clamped: an FP32 q^2 may round up to 4 where the FP64 test admitted it.
__device__ __forceinline__ void add_pair_f(const double* __restrict__ grads, float rx, float ry, float rz, float d2,
                                           float h, float m, float a[3]) {
    if (grads && d2 < 4.f * (h * h)) {
        const float hInv = __builtin_amdgcn_rcpf(h);
        const float qSqr = d2 * (hInv * hInv);
        const float fi = float(LUT_ENTRIES / 4) * qSqr;
        const uint32_t i1 = min(uint32_t(fi), uint32_t(LUT_ENTRIES - 1));
        const float ratio = fi - float(i1);
        const float gr = float(grads[i1]) * (1.f - ratio) + float(grads[i1 + 1]) * ratio;
        const float f = m * (hInv * hInv * hInv) * gr;
        a[0] += rx * f;
        a[1] += ry * f;
        a[2] += rz * f;
        return;
    }
    const float il = __builtin_amdgcn_rsqf(d2);
    const float f = m * (il * il * il);
    a[0] += rx * f;
    a[1] += ry * f;
    a[2] += rz * f;
}

// Per-wavefront staging area of k_eval_mixed: the FP32 half of EvalStage.
template <int ORDER>
struct EvalStageF {
    float c[3][64];
    float M[64];
    float q2[ORDER >= 2 ? 5 : 1][64];
    float q3[ORDER >= 3 ? 7 : 1][64];
    float4 src[64];
    float sm[64];
    uint32_t sidx[64];
};

// k_eval with FP32 arithmetic on offsets from the first particle of the target leaf (R), lengths and
// G m scaled by eval_scale.  A pair separation taken as the difference of two FP32 offsets is off by at
// most 2^-23 O, with O the largest offset among the batch's sources and the wavefront's targets: pairs
// closer than O / 8 (where that is more than 1e-6 of the separation) take theirs in FP64 from the
// source's FP64 position instead.  The lanes, lists, splits and the order of the staged entries are
// those of k_eval.  Each batch of up to 64 staged entries is summed in FP32 and added to FP64
// accumulators; the accelerations are scaled back and added to the state planes in FP64.
template <int ORDER>
__global__ void __launch_bounds__(BLOCK) k_eval_mixed(Tree t, Planes s, const uint32_t* __restrict__ order,
                                                      const uint32_t* __restrict__ targets, uint32_t ntargets,
                                                      const uint32_t* __restrict__ nlist, uint32_t cap_n,
                                                      const uint32_t* __restrict__ xlist, uint32_t cap_x,
                                                      const uint32_t* __restrict__ ncount,
                                                      const uint32_t* __restrict__ xcount,
                                                      const double* __restrict__ grads) {
    __shared__ EvalStageF<ORDER> stage[BLOCK / 64];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    EvalStageF<ORDER>& S_ = stage[wave];
    const uint32_t q = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / 64)) + wave;
    if (q >= ntargets) return;  // uniform over the wavefront
    const EvalScale sc = eval_scale(t);
    const uint32_t T = targets[q];
    uint32_t tf, tc;
    node_range(t, T, tf, tc);
    const LaneSplit ls = lane_split(lane, tc);
    const uint32_t S = ls.S, sp = ls.sp;
    const bool active = ls.used;
    const uint32_t j0 = active ? sp : 64u;  // unused lanes take no entries
    const uint32_t k = tf + ls.p;
    const double4 R = t.pos[tf];
    float mx, my, mz, mh;
    {
        const double4 me = t.pos[k];
        mx = float((me.x - R.x) * sc.invL);
        my = float((me.y - R.y) * sc.invL);
        mz = float((me.z - R.z) * sc.invL);
        mh = float(me.w * sc.invL);
    }
    const float me2 = wave_max_f(active ? mx * mx + my * my + mz * mz : 0.f);
    double a[3] = { 0.0, 0.0, 0.0 };
    const uint32_t nn = min(ncount[q], cap_n), nx = min(xcount[q], cap_x);
    const uint32_t* nl = nlist + size_t(q) * cap_n;
    const uint32_t* xl = xlist + size_t(q) * cap_x;
    const double s2 = sc.invM * sc.invL * sc.invL, s3 = s2 * sc.invL;

    // node list
    for (uint32_t base = 0; base < nn; base += 64) {
        const uint32_t cnt = min(64u, nn - base);
        if (lane < cnt) {
            const uint32_t id = nl[base + lane];
            const double4 g = t.geo[id];
            const NodeMom& mm = t.mom[id];
            S_.c[0][lane] = float((g.x - R.x) * sc.invL);
            S_.c[1][lane] = float((g.y - R.y) * sc.invL);
            S_.c[2][lane] = float((g.z - R.z) * sc.invL);
            S_.M[lane] = float(mm.M * sc.invM);
            if constexpr (ORDER >= 2) {
                for (int c = 0; c < 5; ++c) S_.q2[c][lane] = float(mm.q2[c] * s2);
            }
            if constexpr (ORDER >= 3) {
                for (int c = 0; c < 7; ++c) S_.q3[c][lane] = float(mm.q3[c] * s3);
            }
        }
        wave_sync();
        float b[3] = { 0.f, 0.f, 0.f };
        for (uint32_t j = j0; j < cnt; j += S) {
            float q2[5], q3[7];
            if constexpr (ORDER >= 2) {
                for (int c = 0; c < 5; ++c) q2[c] = S_.q2[c][j];
            }
            if constexpr (ORDER >= 3) {
                for (int c = 0; c < 7; ++c) q3[c] = S_.q3[c][j];
            }
            add_multipole_f<ORDER>(S_.M[j], q2, q3, S_.c[0][j] - mx, S_.c[1][j] - my, S_.c[2][j] - mz, b);
        }
        for (int c = 0; c < 3; ++c) a[c] += double(b[c]);
        wave_sync();
    }

    // source particles: entries e < nx are the listed leaves, entry nx is the target leaf
    const uint32_t nent = nx + 1;
    for (uint32_t eb = 0; eb < nent; eb += 64) {
        uint32_t f = 0, c = 0;
        const uint32_t e = eb + lane;
        if (e < nx) {
            node_range(t, xl[e], f, c);
        } else if (e == nx) {
            f = tf;
            c = tc;
        }
        uint32_t incl = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, 64);
            if (lane >= uint32_t(o)) incl += v;
        }
        const uint32_t excl = incl - c;
        const uint32_t total = __shfl(incl, 63, 64);
        for (uint32_t s0 = 0; s0 < total; s0 += 64) {
            const uint32_t slot = s0 + lane;
            const uint32_t cnt = min(64u, total - s0);
            int lo = 0;
            for (int step = 32; step > 0; step >>= 1) {
                const uint32_t ex = __shfl(excl, (lo + step) & 63, 64);
                if (lo + step < 64 && ex <= slot) lo += step;
            }
            const uint32_t fe = __shfl(f, lo, 64), xe = __shfl(excl, lo, 64);
            if (lane < cnt) {
                const uint32_t src = fe + (slot - xe);
                const double4 o = t.pos[src];
                S_.src[lane] = make_float4(float((o.x - R.x) * sc.invL), float((o.y - R.y) * sc.invL),
                                           float((o.z - R.z) * sc.invL), float(o.w * sc.invL));
                S_.sm[lane] = float(t.gm[src] * sc.invM);
                S_.sidx[lane] = src;
            }
            wave_sync();
            float o2 = 0.f;
            if (lane < cnt) {
                const float4 o = S_.src[lane];
                o2 = o.x * o.x + o.y * o.y + o.z * o.z;
            }
            const float close = fmaxf(wave_max_f(o2), me2) * (1.f / 64.f);  // (O / 8)^2
            float b[3] = { 0.f, 0.f, 0.f };
            // the target itself lies within the close radius (or at zero separation when the radius is zero),
            // so its entry is recognised there, where the source index is read anyway
            for (uint32_t j = j0; j < cnt; j += S) {
                const float4 o = S_.src[j];
                float rx = o.x - mx, ry = o.y - my, rz = o.z - mz;
                float d2 = rx * rx + ry * ry + rz * rz;
                bool self = false;
                if (d2 < close || d2 == 0.f) {
                    const uint32_t src = S_.sidx[j];
                    self = src == k;
                    const double4 od = t.pos[src], me = t.pos[k];
                    rx = float((od.x - me.x) * sc.invL);
                    ry = float((od.y - me.y) * sc.invL);
                    rz = float((od.z - me.z) * sc.invL);
                    d2 = rx * rx + ry * ry + rz * rz;
                }
                if (!self) add_pair_f(grads, rx, ry, rz, d2, 0.5f * (o.w + mh), S_.sm[j], b);
            }
            for (int c = 0; c < 3; ++c) a[c] += double(b[c]);
            wave_sync();
        }
    }
    sum_splits(ls, lane, tc, a);
    if (sp == 0) {
        const uint32_t i = order[k];
        s.d[GAX][i] += a[0] * sc.back;
        s.d[GAY][i] += a[1] * sc.back;
        s.d[GAZ][i] += a[2] * sc.back;
    }
}

// BarnesHut::evalImpl (BarnesHut.cpp:136-178) at arbitrary points: a node is approximated when the point
// lies outside its box and size^2 / (distance^2 + EPS) < theta^2; leaves are summed with the unsymmetrised
// kernel of the source particle.
__global__ void __launch_bounds__(BLOCK) k_points(Tree t, const double4* __restrict__ pts, double4* __restrict__ acc,
                                                  size_t np, const double* __restrict__ grads, int order_m) {
    const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const double4 r0 = pts[q];
    double a[3] = { 0.0, 0.0, 0.0 };
    const double th2 = 1.0 / (t.thetaInv * t.thetaInv);
    uint32_t id = 0;
    for (size_t visits = 0; id != END; ++visits) {
        if (visits > 2 * size_t(t.n)) {
            atomicOr(t.error, 2u);
            break;
        }
        uint32_t f, c;
        node_range(t, id, f, c);
        if (!is_particle(t, id)) {
            const double4 lo = t.lo[id], hi = t.hi[id], g = t.geo[id];
            const bool inside = !(r0.x < lo.x || r0.x > hi.x || r0.y < lo.y || r0.y > hi.y || r0.z < lo.z || r0.z > hi.z);
            const double sx = hi.x - lo.x, sy = hi.y - lo.y, sz = hi.z - lo.z;
            const double cx = 0.5 * (lo.x + hi.x) - r0.x, cy = 0.5 * (lo.y + hi.y) - r0.y, cz = 0.5 * (lo.z + hi.z) - r0.z;
            const double size2 = sx * sx + sy * sy + sz * sz;
            const double dist2 = cx * cx + cy * cy + cz * cz;
            if (!inside && size2 > 0.0 && size2 / (dist2 + EPS_D) < th2) {
                add_multipole(t.mom[id], r0.x - g.x, r0.y - g.y, r0.z - g.z, order_m, a);
                id = t.skip[id];
                continue;
            }
            if (c > t.leaf) {
                id = t.left[id];
                continue;
            }
        }
        // BarnesHut::evalExact (BarnesHut.cpp:376-387)
        for (uint32_t j = f; j < f + c; ++j) {
            const double4 o = t.pos[j];
            add_pair<false>(grads, o.x - r0.x, o.y - r0.y, o.z - r0.z, o.w, t.gm[j], a);
        }
        id = t.skip[id];
    }
    acc[q] = make_double4(a[0], a[1], a[2], 0.0);
}

// BruteForceGravity::evalImpl (BruteForceGravity.h:112-131) for targets tg[0..nt): sources in storage
// order, staged through LDS; the per-target sum runs in source order like the reference's two loops.
__global__ void __launch_bounds__(BLOCK) k_direct(Planes s, size_t n, const uint32_t* __restrict__ tg, size_t nt,
                                                  const double* __restrict__ grads, double G, double4* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double4 sp[BLOCK];
    __shared__ double sm[BLOCK];
    const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool active = q < nt;
    const uint32_t i = active ? tg[q] : 0u;
    const double xi = s.d[GX][i], yi = s.d[GY][i], zi = s.d[GZ][i], hi = s.d[GH][i];
    double a[3] = { 0.0, 0.0, 0.0 };
    for (size_t base = 0; base < n; base += BLOCK) {
        const size_t j = base + threadIdx.x;
        if (j < n) {
            sp[threadIdx.x] = make_double4(s.d[GX][j], s.d[GY][j], s.d[GZ][j], s.d[GH][j]);
            sm[threadIdx.x] = s.d[GM][j];
        }
        __syncthreads();
        const size_t cnt = n - base < size_t(BLOCK) ? n - base : size_t(BLOCK);
        if (active) {
            for (size_t c = 0; c < cnt; ++c) {
                if (base + c == i) continue;
                const double4 o = sp[c];
                add_pair<true>(grads, o.x - xi, o.y - yi, o.z - zi, 0.5 * (o.w + hi), sm[c], a);
            }
        }
        __syncthreads();
    }
    if (active) out[q] = make_double4(G * a[0], G * a[1], G * a[2], 0.0);
}

// BruteForceGravity::evalEnergy (BruteForceGravity.h:103-115): per-particle sums m_i m_j value.
__global__ void __launch_bounds__(BLOCK) k_energy(Planes s, size_t n, const double* __restrict__ values,
                                                  double* __restrict__ e_out) {
#pragma clang fp contract(off)
    __shared__ double4 sp[BLOCK];
    __shared__ double sm[BLOCK];
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool active = i < n;
    const size_t ii = active ? i : 0;
    const double xi = s.d[GX][ii], yi = s.d[GY][ii], zi = s.d[GZ][ii], hi = s.d[GH][ii], mi = s.d[GM][ii];
    double e = 0.0;
    for (size_t base = 0; base < n; base += BLOCK) {
        const size_t j = base + threadIdx.x;
        if (j < n) {
            sp[threadIdx.x] = make_double4(s.d[GX][j], s.d[GY][j], s.d[GZ][j], s.d[GH][j]);
            sm[threadIdx.x] = s.d[GM][j];
        }
        __syncthreads();
        const size_t cnt = n - base < size_t(BLOCK) ? n - base : size_t(BLOCK);
        if (active) {
            for (size_t c = 0; c < cnt; ++c) {
                if (base + c == i) continue;
                const double4 o = sp[c];
                const double rx = xi - o.x, ry = yi - o.y, rz = zi - o.z, h = 0.5 * (hi + o.w);
                double v = 0.0;
                bool outer = true;
                if (values) {
                    // GravityLutKernel::value (GravityKernel.h:45-58)
                    const double hInv = 1.0 / h;
                    const double sx = rx * hInv, sy = ry * hInv, sz = rz * hInv;
                    const double qSqr = sx * sx + sy * sy + sz * sz;
                    if (!(qSqr + EPS_D >= 4.0)) {
                        const double fi = double(LUT_ENTRIES / 4) * qSqr;
                        const uint32_t i1 = uint32_t(fi);
                        const double ratio = fi - double(i1);
                        v = hInv * (values[i1] * (1.0 - ratio) + values[i1 + 1] * ratio);
                        outer = false;
                    }
                }
                if (outer) v = -1.0 / sqrt(rx * rx + ry * ry + rz * rz);
                e += mi * sm[c] * v;
            }
        }
        __syncthreads();
    }
    if (active) e_out[i] = e;
}

// Deterministic reductions: per-block partial sums in a fixed tree order.
template <int NV>
__device__ void block_sum(double (&v)[NV], double* __restrict__ out) {
    __shared__ double red[BLOCK / 64][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[wave][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double x = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) x += red[w][threadIdx.x];
        out[size_t(blockIdx.x) * NV + threadIdx.x] = x;
    }
}

// Integrals (core/physics/Integrals.cpp:24-180): kinetic energy, momentum, angular momentum about the
// origin, mass-weighted position, total mass.
__global__ void __launch_bounds__(BLOCK) k_integrals(Planes s, size_t n, double* __restrict__ partial) {
    double v[11] = { 0.0 };
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const double m = s.d[GM][i];
        const double x = s.d[GX][i], y = s.d[GY][i], z = s.d[GZ][i];
        const double vx = s.d[GVX][i], vy = s.d[GVY][i], vz = s.d[GVZ][i];
        v[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[1] += m * vx;
        v[2] += m * vy;
        v[3] += m * vz;
        v[4] += m * (y * vz - z * vy);
        v[5] += m * (z * vx - x * vz);
        v[6] += m * (x * vy - y * vx);
        v[7] += m * x;
        v[8] += m * y;
        v[9] += m * z;
        v[10] += m;
    }
    block_sum<11>(v, partial);
}

__global__ void __launch_bounds__(BLOCK) k_sum_partials(const double* __restrict__ partial, int nblocks, int nv,
                                                        double* __restrict__ out) {
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        double x = 0.0;
        for (int b = 0; b < nblocks; ++b) x += partial[size_t(b) * nv + k];
        out[k] = x;
    }
}

__global__ void __launch_bounds__(BLOCK) k_sum_vector(const double* __restrict__ v, size_t n, double* __restrict__ partial) {
    double x[1] = { 0.0 };
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) x[0] += v[i];
    block_sum<1>(x, partial);
}

// LeapFrog drift (TimeStepping.cpp:356-360,382-386): r += v dt / 2 in all four lanes.
__global__ void __launch_bounds__(BLOCK) k_drift(Planes s, double dt, size_t n) {
#pragma clang fp contract(off)
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 4; ++k) s.d[GX + k][i] += s.d[GVX + k][i] * 0.5 * dt;
}

__global__ void __launch_bounds__(BLOCK) k_zero_acc(Planes s, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s.d[GAX][i] = 0.0;
    s.d[GAY][i] = 0.0;
    s.d[GAZ][i] = 0.0;
}

// HardSphereSolver zeroes the h lanes of v and dv (NBodySolver.cpp:110-114); then the kick v += dv dt
// (TimeStepping.cpp:372-376) and the second drift.
__global__ void __launch_bounds__(BLOCK) k_kick_drift(Planes s, double dt, size_t n) {
#pragma clang fp contract(off)
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s.d[GVH][i] = 0.0;
    for (int k = 0; k < 3; ++k) s.d[GVX + k][i] += s.d[GAX + k][i] * dt;
    for (int k = 0; k < 4; ++k) s.d[GX + k][i] += s.d[GVX + k][i] * 0.5 * dt;
}

__global__ void __launch_bounds__(BLOCK) k_fill_u32g(uint32_t* p, size_t count, uint32_t value) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x) {
        p[i] = value;
    }
}

// small[0..5] bounds, [6] list sizes needed (2 x u32), [8] traversal error flags.
__global__ void k_init_bounds(unsigned long long* a) {
    const int t = threadIdx.x;
    if (t < 6) a[t] = t < 3 ? ~0ull : 0ull;
    if (t == 6 || t == 8) a[t] = 0ull;
}

// Strided element layouts <-> planes.
// Host rows (the caller's order) into the planes and back; perm (slot -> caller's index) is null while
// the state is in the caller's order.
__global__ void __launch_bounds__(BLOCK) k_unstage(const unsigned char* __restrict__ st, size_t stride, int ncomp,
                                                   Planes s, int plane0, const uint32_t* __restrict__ perm, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* e = reinterpret_cast<const double*>(st + size_t(perm ? perm[i] : i) * stride);
    for (int c = 0; c < ncomp; ++c) s.d[plane0 + c][i] = e[c];
}

__global__ void __launch_bounds__(BLOCK) k_stage(unsigned char* __restrict__ st, size_t stride, int ncomp, Planes s,
                                                 int plane0, const uint32_t* __restrict__ perm, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* e = reinterpret_cast<double*>(st + size_t(perm ? perm[i] : i) * stride);
    for (int c = 0; c < ncomp; ++c) e[c] = s.d[plane0 + c][i];
}

// ---- host-side gravity LUT ------------------------------------------------------------------------------

// GravityKernel<CubicSpline<3>>::gradImpl / valueImpl (core/sph/kernel/GravityKernel.h:82-107) sampled
// by the LutKernel<3> constructor (core/sph/kernel/Kernel.h:83-101).
#pragma STDC FP_CONTRACT OFF
double gk_value(double qSqr) {
    const double q = std::sqrt(qSqr);
    if (q < 1.0) {
        const double q4 = (q * q) * (q * q), q5 = ((q * q) * (q * q)) * q;
        return 2.0 / 3.0 * qSqr - 3.0 / 10.0 * q4 + 1.0 / 10.0 * q5 - 7.0 / 5.0;
    }
    const double q3 = q * q * q, q5 = ((q * q) * (q * q)) * q;
    return 4.0 / 3.0 * qSqr - q3 + 3.0 / 10.0 * (qSqr * qSqr) - 1.0 / 30.0 * q5 - 8.0 / 5.0 + 1.0 / (15.0 * q);
}

double gk_grad(double qSqr) {
    const double q = std::sqrt(qSqr);
    if (q == 0.0) return 4.0 / 3.0;
    if (q < 1.0) return 1.0 / q * (4.0 / 3.0 * q - 6.0 / 5.0 * (q * q * q) + 1.0 / 2.0 * (qSqr * qSqr));
    return 1.0 / q * (8.0 / 3.0 * q - 3.0 * qSqr + 6.0 / 5.0 * (q * q * q) - 1.0 / 6.0 * (qSqr * qSqr) -
                      1.0 / (15.0 * qSqr));
}

}  // namespace
}  // namespace osph

using namespace osph;

struct osph_gravity {
    osph_gravity_config cfg{};
    const void* owner = nullptr;  // the SPH context it is attached to (osph_attach_gravity), if any
    int precision = OSPH_PRECISION_FP64;  // the tree walk's arithmetic (osph_gravity_set_precision)
    hipStream_t stream = nullptr;
    size_t n = 0, cap = 0;
    double* dblock = nullptr;
    Planes s{};
    // LeapFrog steps move the state into the Morton order of the tree every reorder_period builds, so
    // that the tree's gather and the accelerations' writes run over consecutive slots.  perm maps a
    // slot to the particle's index in the caller's order and slot_of back; identity: no move since the
    // particles were set.
    double* dblock2 = nullptr;
    uint32_t *perm = nullptr, *perm2 = nullptr, *slot_of = nullptr;
    bool identity = true;
    int reorder_period = 10, since_reorder = 0;
    // tree
    uint64_t *keys = nullptr, *keys_sorted = nullptr;
    uint32_t *iota = nullptr, *order = nullptr;
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    double4* pos = nullptr;
    double* gm = nullptr;
    uint32_t *left = nullptr, *right = nullptr, *parent = nullptr, *skip = nullptr, *first = nullptr, *count = nullptr,
             *flag = nullptr;
    double4 *geo = nullptr, *lo = nullptr, *hi = nullptr;
    NodeMom* mom = nullptr;
    WalkNode* wn = nullptr;
    WalkNodeF* wnf = nullptr;  // filled on the first screened walk after a build
    bool wnf_ready = false;
    double ctr[3] = { 0.0, 0.0, 0.0 }, invL = 1.0;
    int walk_screen = -1;  // 1: FP32-screened walk, 0: FP64 walk, -1: screened in mixed precision only
    uint32_t *leaf_at = nullptr, *targets = nullptr, *ntargets_dev = nullptr;
    void* select_temp = nullptr;
    size_t select_temp_bytes = 0;
    uint32_t ntargets = 0;
    // interaction lists
    uint32_t *nlist = nullptr, *xlist = nullptr, *ncount = nullptr, *xcount = nullptr, *work = nullptr;
    uint32_t cap_n = 256, cap_x = 64;
    size_t list_targets = 0;
    uint32_t* levels = nullptr;           // node counts per depth, then the bucket cursors
    unsigned long long* small = nullptr;  // [0..5] bounds, [6] need (2 x u32)
    unsigned long long* small_host = nullptr;
    double *lut_grads = nullptr, *lut_values = nullptr;
    double* scratch = nullptr;
    size_t scratch_cap = 0;
    unsigned char* staging = nullptr;
    size_t staging_cap = 0;
    bool built = false;
    // profiling
    bool prof = false;
    struct Ev {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Ev> events;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, uint64_t>> acc;
    std::string err;

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

namespace {

#define GCHK(g, expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return (g)->fail(OSPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int OK = OSPH_OK, ERR_ARG = OSPH_ERR_ARG, ERR_HIP = OSPH_ERR_HIP, ERR_STATE = OSPH_ERR_STATE;

hipEvent_t take_event(osph_gravity* g) {
    if (!g->pool.empty()) {
        hipEvent_t e = g->pool.back();
        g->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

template <typename F>
void timed(osph_gravity* g, const char* name, F&& f) {
    if (!g->prof) {
        f();
        return;
    }
    osph_gravity::Ev ev{ name, take_event(g), take_event(g) };
    if (ev.a) (void)hipEventRecord(ev.a, g->stream);
    f();
    if (ev.b) (void)hipEventRecord(ev.b, g->stream);
    g->events.push_back(ev);
}

void resolve(osph_gravity* g) {
    for (auto& e : g->events) {
        float ms = 0.f;
        if (e.a && e.b && hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            auto& a = g->acc[e.name];
            a.first += ms;
            a.second += 1;
        }
        if (e.a) g->pool.push_back(e.a);
        if (e.b) g->pool.push_back(e.b);
    }
    g->events.clear();
}

int sync(osph_gravity* g) {
    GCHK(g, hipStreamSynchronize(g->stream));
    GCHK(g, hipGetLastError());
    return OK;
}

// Synchronises and fails if a tree climb or walk ran past its bound (a malformed tree).
int check_traversal(osph_gravity* g) {
    GCHK(g, hipMemcpyAsync(g->small_host + 8, g->small + 8, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           g->stream));
    if (int e = sync(g)) return e;
    if (g->small_host[8] != 0) return g->fail(ERR_STATE, "tree traversal exceeded its bound");
    return OK;
}

void free_particles(osph_gravity* g) {
    void* ptrs[] = { g->dblock, g->dblock2, g->perm, g->perm2, g->slot_of, g->keys, g->keys_sorted, g->iota, g->order, g->sort_temp, g->pos, g->gm, g->left,
                     g->right, g->parent, g->skip, g->first, g->count, g->flag, g->geo, g->lo, g->hi, g->mom, g->wn,
                     g->wnf, g->leaf_at, g->targets, g->ntargets_dev, g->select_temp };
    for (void* p : ptrs) (void)hipFree(p);
    g->dblock = g->dblock2 = nullptr;
    g->perm = g->perm2 = g->slot_of = nullptr;
    g->identity = true;
    g->keys = g->keys_sorted = nullptr;
    g->iota = g->order = nullptr;
    g->sort_temp = g->select_temp = nullptr;
    g->pos = nullptr;
    g->gm = nullptr;
    g->left = g->right = g->parent = g->skip = g->first = g->count = g->flag = nullptr;
    g->geo = g->lo = g->hi = nullptr;
    g->mom = nullptr;
    g->wn = nullptr;
    g->wnf = nullptr;
    g->wnf_ready = false;
    g->leaf_at = g->targets = g->ntargets_dev = nullptr;
    g->cap = 0;
}

int ensure_particles(osph_gravity* g, size_t n) {
    if (n <= g->cap) return OK;
    GCHK(g, hipStreamSynchronize(g->stream));
    free_particles(g);
    const size_t c = n, nodes = c > 1 ? c - 1 : 1, ids = 2 * c;
    GCHK(g, hipMalloc(&g->dblock, sizeof(double) * NGPLANES * c));
    GCHK(g, hipMalloc(&g->dblock2, sizeof(double) * NGPLANES * c));
    GCHK(g, hipMalloc(&g->perm, sizeof(uint32_t) * c));
    GCHK(g, hipMalloc(&g->perm2, sizeof(uint32_t) * c));
    GCHK(g, hipMalloc(&g->slot_of, sizeof(uint32_t) * c));
    GCHK(g, hipMalloc(&g->keys, sizeof(uint64_t) * c));
    GCHK(g, hipMalloc(&g->keys_sorted, sizeof(uint64_t) * c));
    GCHK(g, hipMalloc(&g->iota, sizeof(uint32_t) * c));
    GCHK(g, hipMalloc(&g->order, sizeof(uint32_t) * c));
    GCHK(g, hipMalloc(&g->pos, sizeof(double4) * c));
    GCHK(g, hipMalloc(&g->gm, sizeof(double) * c));
    GCHK(g, hipMalloc(&g->left, sizeof(uint32_t) * nodes));
    GCHK(g, hipMalloc(&g->right, sizeof(uint32_t) * nodes));
    GCHK(g, hipMalloc(&g->first, sizeof(uint32_t) * nodes));
    GCHK(g, hipMalloc(&g->count, sizeof(uint32_t) * nodes));
    GCHK(g, hipMalloc(&g->flag, sizeof(uint32_t) * nodes));
    GCHK(g, hipMalloc(&g->geo, sizeof(double4) * nodes));
    GCHK(g, hipMalloc(&g->lo, sizeof(double4) * nodes));
    GCHK(g, hipMalloc(&g->hi, sizeof(double4) * nodes));
    GCHK(g, hipMalloc(&g->mom, sizeof(NodeMom) * nodes));
    GCHK(g, hipMalloc(&g->wn, sizeof(WalkNode) * nodes));
    GCHK(g, hipMalloc(&g->wnf, sizeof(WalkNodeF) * nodes));
    GCHK(g, hipMalloc(&g->parent, sizeof(uint32_t) * ids));
    GCHK(g, hipMalloc(&g->skip, sizeof(uint32_t) * ids));
    GCHK(g, hipMalloc(&g->leaf_at, sizeof(uint32_t) * c));
    GCHK(g, hipMalloc(&g->targets, sizeof(uint32_t) * c));
    GCHK(g, hipMalloc(&g->ntargets_dev, sizeof(uint32_t)));
    size_t b1 = 0, b2 = 0;
    GCHK(g, hipcub::DeviceRadixSort::SortPairs(nullptr, b1, g->keys, g->keys_sorted, g->iota, g->order, int(c), 0, 63,
                                               g->stream));
    GCHK(g, hipcub::DeviceSelect::If(nullptr, b2, g->leaf_at, g->targets, g->ntargets_dev, int(c), NotEnd(), g->stream));
    g->sort_temp_bytes = b1;
    g->select_temp_bytes = b2;
    GCHK(g, hipMalloc(&g->sort_temp, b1 > 0 ? b1 : 1));
    GCHK(g, hipMalloc(&g->select_temp, b2 > 0 ? b2 : 1));
    for (int k = 0; k < NGPLANES; ++k) g->s.d[k] = g->dblock + size_t(k) * c;
    g->cap = c;
    return OK;
}

int ensure_staging(osph_gravity* g, size_t bytes) {
    if (bytes <= g->staging_cap) return OK;
    (void)hipFree(g->staging);
    g->staging = nullptr;
    g->staging_cap = 0;
    GCHK(g, hipMalloc(&g->staging, bytes));
    g->staging_cap = bytes;
    return OK;
}

int ensure_scratch(osph_gravity* g, size_t doubles) {
    if (doubles <= g->scratch_cap) return OK;
    (void)hipFree(g->scratch);
    g->scratch = nullptr;
    g->scratch_cap = 0;
    GCHK(g, hipMalloc(&g->scratch, doubles * sizeof(double)));
    g->scratch_cap = doubles;
    return OK;
}

int ensure_lists(osph_gravity* g, size_t targets) {
    const size_t need = targets;
    if (need <= g->list_targets && g->nlist) return OK;
    (void)hipFree(g->nlist);
    (void)hipFree(g->xlist);
    (void)hipFree(g->ncount);
    (void)hipFree(g->xcount);
    (void)hipFree(g->work);
    g->nlist = g->xlist = g->ncount = g->xcount = g->work = nullptr;
    g->list_targets = 0;
    const size_t t = need > 0 ? need : 1;
    GCHK(g, hipMalloc(&g->nlist, sizeof(uint32_t) * t * g->cap_n));
    GCHK(g, hipMalloc(&g->xlist, sizeof(uint32_t) * t * g->cap_x));
    GCHK(g, hipMalloc(&g->ncount, sizeof(uint32_t) * t));
    GCHK(g, hipMalloc(&g->xcount, sizeof(uint32_t) * t));
    GCHK(g, hipMalloc(&g->work, 2 * sizeof(uint32_t) * t));
    g->list_targets = t;
    return OK;
}

Tree tree_of(osph_gravity* g) {
    Tree t{};
    t.n = uint32_t(g->n);
    t.leaf = uint32_t(g->cfg.leaf_size);
    t.left = g->left;
    t.right = g->right;
    t.parent = g->parent;
    t.skip = g->skip;
    t.first = g->first;
    t.count = g->count;
    t.geo = g->geo;
    t.lo = g->lo;
    t.hi = g->hi;
    t.mom = g->mom;
    t.wn = g->wn;
    t.wnf = g->wnf;
    for (int k = 0; k < 3; ++k) t.ctr[k] = g->ctr[k];
    t.invL = g->invL;
    t.flag = g->flag;
    t.pos = g->pos;
    t.gm = g->gm;
    t.thetaInv = 1.0 / g->cfg.theta;
    t.error = reinterpret_cast<uint32_t*>(g->small + 8);
    return t;
}

double decode(unsigned long long e) { return decode_ord_host(e); }

// Node moments: the leaves' from their particles (k_leaf_moments over the leaf list, which must be
// selected into g->targets first), then the nodes above them, deepest level first (k_depth, k_bucket,
// k_moments_level).  The flag array holds the depths and leaf_at the nodes bucketed by depth; leaf_at
// is free once the leaf list is selected.
int build_moments(osph_gravity* g, const Tree& t) {
    const size_t n = g->n;
    hipStream_t st = g->stream;
    uint32_t* per_level = g->levels;
    uint32_t* cursor = g->levels + MAX_DEPTH + 1;
    timed(g, "depth", [&] {
        k_fill_u32g<<<1, BLOCK, 0, st>>>(per_level, MAX_DEPTH + 1, 0u);
        k_depth<<<std::min<unsigned>(grid_for(n - 1), LEVEL_BLOCKS), BLOCK, 0, st>>>(t, g->flag, per_level);
    });
    std::vector<uint32_t> counts(MAX_DEPTH + 1);
    GCHK(g, hipMemcpyAsync(counts.data(), per_level, counts.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (int e = check_traversal(g)) return e;
    std::vector<uint32_t> off(MAX_DEPTH + 2, 0);
    int deepest = 0;
    for (int d = 0; d <= MAX_DEPTH; ++d) {
        off[d + 1] = off[d] + counts[d];
        if (counts[d]) deepest = d;
    }
    GCHK(g, hipMemcpyAsync(cursor, off.data(), (MAX_DEPTH + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    timed(g, "moments", [&] {
        k_leaf_moments<<<grid_for(n * size_t(LEAF_LANES)), BLOCK, 0, st>>>(t, g->targets, g->ntargets_dev);
        k_bucket<<<std::min<unsigned>(grid_for(n - 1), LEVEL_BLOCKS), BLOCK, 0, st>>>(uint32_t(n - 1), g->flag, cursor,
                                                                                   g->leaf_at);
        for (int d = deepest; d >= 0; --d) {
            if (counts[d]) {
                k_moments_level<<<grid_for(counts[d]), BLOCK, 0, st>>>(t, g->leaf_at + off[d], counts[d]);
            }
        }
    });
    return OK;
}

// With reorder (LeapFrog steps only), the state is moved into the Morton order when due.
int build_tree(osph_gravity* g, bool reorder = false) {
    const size_t n = g->n;
    g->built = false;
    if (n == 0) {
        g->built = true;
        g->ntargets = 0;
        return OK;
    }
    hipStream_t st = g->stream;
    timed(g, "bounds", [&] {
        k_init_bounds<<<1, 64, 0, st>>>(g->small);
        unsigned gb = grid_for(n);
        gb = gb > 1024 ? 1024 : gb;
        k_bounds3<<<gb, BLOCK, 0, st>>>(g->s, g->small, n);
    });
    GCHK(g, hipMemcpyAsync(g->small_host, g->small, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    if (int e = sync(g)) return e;
    double lo[3], ext = 0.0;
    for (int k = 0; k < 3; ++k) {
        lo[k] = decode(g->small_host[k]);
        const double hi = decode(g->small_host[3 + k]);
        if (!std::isfinite(lo[k]) || !std::isfinite(hi)) return g->fail(ERR_STATE, "non-finite positions");
        ext = std::max(ext, hi - lo[k]);
    }
    const double scale = ext > 0.0 ? 2097151.0 / ext : 0.0;
    // the screened walk's frame: the root box's centre, lengths divided by 2^e >= the extent
    for (int k = 0; k < 3; ++k) g->ctr[k] = lo[k] + 0.5 * (decode(g->small_host[3 + k]) - lo[k]);
    int e = 0;
    if (ext > 0.0) (void)std::frexp(ext, &e);
    g->invL = std::ldexp(1.0, -e);
    g->wnf_ready = false;
    Tree t = tree_of(g);
    timed(g, "morton", [&] { k_morton<<<grid_for(n), BLOCK, 0, st>>>(g->s, lo[0], lo[1], lo[2], scale, g->keys, g->iota, n); });
    timed(g, "sort", [&] {
        (void)hipcub::DeviceRadixSort::SortPairs(g->sort_temp, g->sort_temp_bytes, g->keys, g->keys_sorted, g->iota,
                                                 g->order, int(n), 0, 63, st);
    });
    if (reorder && g->reorder_period > 0 && ++g->since_reorder >= g->reorder_period) {
        g->since_reorder = 0;
        Planes dst{};
        for (int k = 0; k < NGPLANES; ++k) dst.d[k] = g->dblock2 + size_t(k) * g->cap;
        timed(g, "reorder", [&] {
            k_permute_planes<<<grid_for(n), BLOCK, 0, st>>>(g->s, dst, g->order, g->identity ? nullptr : g->perm,
                                                            g->perm2, g->slot_of, n);
            k_sequence<<<grid_for(n), BLOCK, 0, st>>>(g->order, n);
        });
        std::swap(g->dblock, g->dblock2);
        std::swap(g->perm, g->perm2);
        g->s = dst;
        g->identity = false;
    }
    timed(g, "gather", [&] { k_gather<<<grid_for(n), BLOCK, 0, st>>>(g->s, g->order, g->cfg.G, g->pos, g->gm, n); });
    if (n > 1) {
        timed(g, "radix_tree", [&] { k_radix_tree<<<grid_for(n - 1), BLOCK, 0, st>>>(t, g->keys_sorted); });
        timed(g, "skip", [&] { k_skip<<<grid_for(2 * n - 1), BLOCK, 0, st>>>(t); });
    } else {
        const uint32_t endv = END;
        GCHK(g, hipMemcpyAsync(g->parent, &endv, sizeof(uint32_t), hipMemcpyHostToDevice, st));
        GCHK(g, hipMemcpyAsync(g->skip, &endv, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    timed(g, "leaves", [&] {
        k_fill_u32g<<<std::min<unsigned>(grid_for(n), 4096), BLOCK, 0, st>>>(g->leaf_at, n, END);
        k_mark_leaves<<<grid_for(2 * n - 1), BLOCK, 0, st>>>(t, g->leaf_at);
        (void)hipcub::DeviceSelect::If(g->select_temp, g->select_temp_bytes, g->leaf_at, g->targets, g->ntargets_dev,
                                       int(n), NotEnd(), st);
    });
    if (n > 1) {
        if (int e = build_moments(g, t)) return e;
    }
    uint32_t nt = 0;
    GCHK(g, hipMemcpyAsync(g->small_host + 7, g->ntargets_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (int e = check_traversal(g)) return e;
    std::memcpy(&nt, g->small_host + 7, sizeof(uint32_t));
    g->ntargets = nt;
    g->built = true;
    return OK;
}

// Barnes-Hut interaction lists and sums, adding to the state's accelerations.
int eval_bh(osph_gravity* g, osph_gravity_stats* stats) {
    const size_t n = g->n;
    if (!n) return OK;
    hipStream_t st = g->stream;
    Tree t = tree_of(g);
    const uint32_t nt = g->ntargets;
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (int e = ensure_lists(g, nt)) return e;
        uint32_t* need = reinterpret_cast<uint32_t*>(g->small + 6);
        const bool screened = g->walk_screen == 1 || (g->walk_screen == -1 && g->precision == OSPH_PRECISION_MIXED);
        if (screened && !g->wnf_ready && n > 1) {
            timed(g, "walk_records", [&] { k_walk_records<<<grid_for(n - 1), BLOCK, 0, st>>>(t); });
            g->wnf_ready = true;
        }
        if (screened && n > 1) {
            timed(g, "walk_screened", [&] {
                k_fill_u32g<<<1, 64, 0, st>>>(need, 2, 0u);
                k_walk_screened<<<grid_for(nt), BLOCK, 0, st>>>(t, g->targets, nt, g->nlist, g->cap_n, g->xlist,
                                                                g->cap_x, g->ncount, g->xcount, g->work, need);
            });
        } else timed(g, "walk", [&] {
            k_fill_u32g<<<1, 64, 0, st>>>(need, 2, 0u);
            k_walk<<<grid_for(nt), BLOCK, 0, st>>>(t, g->targets, nt, g->nlist, g->cap_n, g->xlist, g->cap_x, g->ncount,
                                                   g->xcount, g->work, need);
        });
        GCHK(g, hipMemcpyAsync(g->small_host + 6, need, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        if (int e = check_traversal(g)) return e;
        uint32_t nd[2];
        std::memcpy(nd, g->small_host + 6, sizeof(nd));
        if (nd[0] == 0 && nd[1] == 0) break;
        if (attempt == 3) return g->fail(ERR_STATE, "interaction lists did not fit after growing");
        // grow to the largest list and walk again
        (void)hipFree(g->nlist);
        g->nlist = nullptr;
        if (nd[0]) g->cap_n = (nd[0] + 63u) & ~63u;
        if (nd[1]) g->cap_x = (nd[1] + 15u) & ~15u;
        g->list_targets = 0;
    }
    const int ord = g->cfg.order;
    const double* grads = g->cfg.kernel == OSPH_GRAVITY_KERNEL_M4 ? g->lut_grads : nullptr;
    const unsigned blocks = unsigned((nt + (BLOCK / 64) - 1) / (BLOCK / 64));
    if (g->precision == OSPH_PRECISION_MIXED) {
        timed(g, "eval_mixed", [&] {
            const unsigned nb = blocks > 0 ? blocks : 1;
            if (ord == 3) {
                k_eval_mixed<3><<<nb, BLOCK, 0, st>>>(t, g->s, g->order, g->targets, nt, g->nlist, g->cap_n, g->xlist,
                                                      g->cap_x, g->ncount, g->xcount, grads);
            } else if (ord == 2) {
                k_eval_mixed<2><<<nb, BLOCK, 0, st>>>(t, g->s, g->order, g->targets, nt, g->nlist, g->cap_n, g->xlist,
                                                      g->cap_x, g->ncount, g->xcount, grads);
            } else {
                k_eval_mixed<0><<<nb, BLOCK, 0, st>>>(t, g->s, g->order, g->targets, nt, g->nlist, g->cap_n, g->xlist,
                                                      g->cap_x, g->ncount, g->xcount, grads);
            }
        });
    } else timed(g, "eval", [&] {
        const unsigned nb = blocks > 0 ? blocks : 1;
        if (ord == 3) {
            k_eval<3><<<nb, BLOCK, 0, st>>>(t, g->s, g->order, g->targets, nt, g->nlist, g->cap_n, g->xlist, g->cap_x,
                                            g->ncount, g->xcount, grads);
        } else if (ord == 2) {
            k_eval<2><<<nb, BLOCK, 0, st>>>(t, g->s, g->order, g->targets, nt, g->nlist, g->cap_n, g->xlist, g->cap_x,
                                            g->ncount, g->xcount, grads);
        } else {
            k_eval<0><<<nb, BLOCK, 0, st>>>(t, g->s, g->order, g->targets, nt, g->nlist, g->cap_n, g->xlist, g->cap_x,
                                            g->ncount, g->xcount, grads);
        }
    });
    if (stats) {
        std::vector<uint32_t> a(nt), b(nt), w(2 * size_t(nt));
        if (nt) {
            GCHK(g, hipMemcpyAsync(a.data(), g->ncount, nt * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            GCHK(g, hipMemcpyAsync(b.data(), g->xcount, nt * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            GCHK(g, hipMemcpyAsync(w.data(), g->work, w.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        }
        if (int e = sync(g)) return e;
        *stats = osph_gravity_stats{};
        for (uint32_t k = 0; k < nt; ++k) {
            stats->nodes_approx += a[k];
            stats->nodes_exact += b[k];
            stats->expansions += w[2 * size_t(k)];
            stats->pairs += w[2 * size_t(k) + 1];
        }
        stats->node_count = 2 * uint64_t(n) - 1;
    }
    return OK;
}

// Direct sum for targets (device indices) into out (double4 per target), G applied.
int eval_direct(osph_gravity* g, const uint32_t* tg_dev, size_t nt, double4* out) {
    if (!nt) return OK;
    const double* grads = g->cfg.kernel == OSPH_GRAVITY_KERNEL_M4 ? g->lut_grads : nullptr;
    timed(g, "direct", [&] {
        k_direct<<<grid_for(nt), BLOCK, 0, g->stream>>>(g->s, g->n, tg_dev, nt, grads, g->cfg.G, out);
    });
    return OK;
}

__global__ void k_add_direct(Planes s, const uint32_t* tg, const double4* a, size_t nt) {
    const size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nt) return;
    const uint32_t i = tg[q];
    s.d[GAX][i] += a[q].x;
    s.d[GAY][i] += a[q].y;
    s.d[GAZ][i] += a[q].z;
}

__global__ void k_iota(uint32_t* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = uint32_t(i);
}

// Self-gravity of every particle added to the state's accelerations.
int eval_self(osph_gravity* g, osph_gravity_stats* stats) {
    if (g->cfg.solver == OSPH_GRAVITY_BARNES_HUT) {
        if (!g->built) return g->fail(ERR_STATE, "gravity not built");
        return eval_bh(g, stats);
    }
    const size_t n = g->n;
    if (!n) return OK;
    if (int e = ensure_scratch(g, 4 * n + n / 2 + 1)) return e;
    double4* out = reinterpret_cast<double4*>(g->scratch);
    uint32_t* tg = reinterpret_cast<uint32_t*>(g->scratch + 4 * n);
    k_iota<<<grid_for(n), BLOCK, 0, g->stream>>>(tg, n);
    if (int e = eval_direct(g, tg, n, out)) return e;
    k_add_direct<<<grid_for(n), BLOCK, 0, g->stream>>>(g->s, tg, out, n);
    if (stats) {
        *stats = osph_gravity_stats{};
        stats->pairs = uint64_t(n) * (n - 1);
    }
    return OK;
}

int upload_vectors(osph_gravity* g, const double* host, size_t stride, int plane0, int ncomp, size_t n) {
    if (!n) return OK;
    if (int e = ensure_staging(g, n * stride)) return e;
    GCHK(g, hipMemcpyAsync(g->staging, host, n * stride, hipMemcpyHostToDevice, g->stream));
    k_unstage<<<grid_for(n), BLOCK, 0, g->stream>>>(g->staging, stride, ncomp, g->s, plane0,
                                                    g->identity ? nullptr : g->perm, n);
    return OK;
}

int download_vectors(osph_gravity* g, double* host, size_t stride, int plane0, int ncomp, size_t n) {
    if (!n) return OK;
    if (int e = ensure_staging(g, n * stride)) return e;
    if (stride > size_t(ncomp) * sizeof(double)) {
        // keep the caller's other lanes: start from their current values
        GCHK(g, hipMemcpyAsync(g->staging, host, n * stride, hipMemcpyHostToDevice, g->stream));
    }
    k_stage<<<grid_for(n), BLOCK, 0, g->stream>>>(g->staging, stride, ncomp, g->s, plane0,
                                                  g->identity ? nullptr : g->perm, n);
    GCHK(g, hipMemcpyAsync(host, g->staging, n * stride, hipMemcpyDeviceToHost, g->stream));
    return sync(g);
}

}  // namespace

namespace {

// Positions, smoothing lengths and masses of another context's particles (device planes, n each) into
// this context's planes; accelerations and velocities start at zero.
__global__ void __launch_bounds__(BLOCK) k_load_planes(Planes s, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       const double* __restrict__ h, const double* __restrict__ m,
                                                       size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s.d[GX][i] = x[i];
    s.d[GY][i] = y[i];
    s.d[GZ][i] = z[i];
    s.d[GH][i] = h[i];
    s.d[GM][i] = m[i];
    for (int k = GVX; k <= GVH; ++k) s.d[k][i] = 0.0;
    for (int k = GAX; k <= GAZ; ++k) s.d[k][i] = 0.0;
}

}  // namespace

const void* gravity_owner(const osph_gravity* g) { return g->owner; }
void gravity_set_owner(osph_gravity* g, const void* owner) { g->owner = owner; }

// GravitySolver::loop's gravity half (core/sph/solvers/GravitySolver.cpp:64-80) for particles held in
// another context's device planes: once `ready` has fired, the planes are copied in on this context's
// stream, the tree is built and the self-gravity evaluated; `done` is recorded after the last kernel and
// accel[0..2] then point to the accelerations (n doubles each, in the order of the source planes).
int gravity_self_from_device(osph_gravity* g, size_t n, const double* const src[5], hipEvent_t ready,
                             hipEvent_t done, const double* accel[3]) {
    if (!g || (n && !src)) return ERR_ARG;
    if (n >= (size_t(1) << 31)) return g->fail(ERR_ARG, "particle count must be below 2^31");
    if (int e = ensure_particles(g, n > 0 ? n : 1)) return e;
    g->n = n;
    g->built = false;
    g->identity = true;  // the planes are loaded in the source order and stay in it
    for (int k = 0; k < 3; ++k) accel[k] = g->s.d[GAX + k];
    GCHK(g, hipStreamWaitEvent(g->stream, ready, 0));
    if (n) {
        k_load_planes<<<grid_for(n), BLOCK, 0, g->stream>>>(g->s, src[0], src[1], src[2], src[3], src[4], n);
        GCHK(g, hipPeekAtLastError());
        if (int e = osph_gravity_build(g)) return e;
        if (int e = eval_self(g, nullptr)) return e;
    }
    GCHK(g, hipEventRecord(done, g->stream));
    return OK;
}

extern "C" {

int osph_gravity_create(osph_gravity** out, const osph_gravity_config* cfg) {
    if (!out || !cfg) return ERR_ARG;
    *out = nullptr;
    if (cfg->solver != OSPH_GRAVITY_BRUTE_FORCE && cfg->solver != OSPH_GRAVITY_BARNES_HUT) return ERR_ARG;
    if (cfg->kernel != OSPH_GRAVITY_KERNEL_POINT && cfg->kernel != OSPH_GRAVITY_KERNEL_M4) return ERR_ARG;
    if (cfg->solver == OSPH_GRAVITY_BARNES_HUT) {
        if (!(cfg->theta > 0.0) || !std::isfinite(cfg->theta)) return ERR_ARG;
        if (cfg->order != 0 && cfg->order != 2 && cfg->order != 3) return ERR_ARG;
        if (cfg->leaf_size < 1 || cfg->leaf_size > 64) return ERR_ARG;
    }
    osph_gravity* g = new osph_gravity();
    g->cfg = *cfg;
    auto bail = [&](int code) {
        osph_gravity_destroy(g);
        return code;
    };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) return bail(ERR_HIP);
    if (hipMalloc(&g->small, 16 * sizeof(unsigned long long)) != hipSuccess) return bail(ERR_HIP);
    if (hipMalloc(&g->levels, 2 * (MAX_DEPTH + 1) * sizeof(uint32_t)) != hipSuccess) return bail(ERR_HIP);
    if (hipHostMalloc(&g->small_host, 16 * sizeof(unsigned long long)) != hipSuccess) return bail(ERR_HIP);
    std::vector<double> grads(LUT_ENTRIES + 1), values(LUT_ENTRIES + 1);
    const double qSqrToIdx = double(LUT_ENTRIES) * (1.0 / (2.0 * 2.0));
    for (int i = 0; i <= LUT_ENTRIES; ++i) {
        const double qSqr = double(i) / qSqrToIdx;
        grads[i] = gk_grad(qSqr);
        values[i] = gk_value(qSqr);
    }
    if (hipMalloc(&g->lut_grads, grads.size() * sizeof(double)) != hipSuccess) return bail(ERR_HIP);
    if (hipMalloc(&g->lut_values, values.size() * sizeof(double)) != hipSuccess) return bail(ERR_HIP);
    if (hipMemcpy(g->lut_grads, grads.data(), grads.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lut_values, values.data(), values.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        return bail(ERR_HIP);
    }
    *out = g;
    return OK;
}

void osph_gravity_destroy(osph_gravity* g) {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    resolve(g);
    for (hipEvent_t e : g->pool) (void)hipEventDestroy(e);
    free_particles(g);
    void* ptrs[] = { g->nlist, g->xlist, g->ncount, g->xcount, g->work, g->small, g->levels, g->lut_grads, g->lut_values,
                     g->scratch, g->staging };
    for (void* p : ptrs) (void)hipFree(p);
    if (g->small_host) (void)hipHostFree(g->small_host);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

const char* osph_gravity_last_error(const osph_gravity* g) { return g ? g->err.c_str() : "null context"; }

int osph_gravity_set_precision(osph_gravity* g, int precision) {
    if (!g) return ERR_ARG;
    if (precision != OSPH_PRECISION_FP64 && precision != OSPH_PRECISION_MIXED) {
        return g->fail(ERR_ARG, "precision must be OSPH_PRECISION_FP64 or OSPH_PRECISION_MIXED");
    }
    g->precision = precision;
    return OK;
}

int osph_gravity_tune(osph_gravity* g, const char* key, long long value) {
    if (!g || !key) return ERR_ARG;
    const std::string k(key);
    if (k == "walk_screen" && (value == -1 || value == 0 || value == 1)) {
        g->walk_screen = int(value);
        return OK;
    }
    if (k == "reorder_period" && value >= 0 && value <= 1000000) {
        g->reorder_period = int(value);
        return OK;
    }
    return g->fail(ERR_ARG, "unknown tuning key or value out of range: " + k);
}

int osph_gravity_set_particles(osph_gravity* g, size_t n, const double* r, const double* v, size_t vec_stride,
                               const double* m) {
    if (!g || (n && (!r || !m))) return ERR_ARG;
    if (n >= (size_t(1) << 31)) return g->fail(ERR_ARG, "particle count must be below 2^31");
    const size_t stride = vec_stride ? vec_stride : 4 * sizeof(double);
    if (stride < 4 * sizeof(double) || stride % sizeof(double)) return g->fail(ERR_ARG, "bad vector stride");
    if (int e = ensure_particles(g, n > 0 ? n : 1)) return e;
    g->n = n;
    g->built = false;
    g->identity = true;
    g->since_reorder = g->reorder_period;  // the first LeapFrog build moves the state into Morton order
    if (!n) return sync(g);
    if (int e = upload_vectors(g, r, stride, GX, 4, n)) return e;
    if (v) {
        if (int e = upload_vectors(g, v, stride, GVX, 4, n)) return e;
        if (int e = sync(g)) return e;  // the staging buffer is reused below
    } else {
        GCHK(g, hipMemsetAsync(g->s.d[GVX], 0, 4 * sizeof(double) * g->cap, g->stream));
    }
    GCHK(g, hipMemsetAsync(g->s.d[GAX], 0, 3 * sizeof(double) * g->cap, g->stream));
    GCHK(g, hipMemcpyAsync(g->s.d[GM], m, n * sizeof(double), hipMemcpyHostToDevice, g->stream));
    return sync(g);
}

int osph_gravity_get(osph_gravity* g, int what, double* out, size_t vec_stride) {
    if (!g || !out || what < 0 || what > 2) return ERR_ARG;
    const size_t stride = vec_stride ? vec_stride : 4 * sizeof(double);
    if (stride < 4 * sizeof(double) || stride % sizeof(double)) return g->fail(ERR_ARG, "bad vector stride");
    const int plane0 = what == 0 ? GX : (what == 1 ? GVX : GAX);
    const int ncomp = what == 2 ? 3 : 4;
    if (what == 2) {
        for (size_t i = 0; i < g->n; ++i) out[i * (stride / sizeof(double)) + 3] = 0.0;
    }
    return download_vectors(g, out, stride, plane0, ncomp, g->n);
}

int osph_gravity_build(osph_gravity* g) {
    if (!g) return ERR_ARG;
    if (g->cfg.solver == OSPH_GRAVITY_BRUTE_FORCE) {
        g->built = true;
        return OK;
    }
    return build_tree(g);
}

int osph_gravity_eval_self(osph_gravity* g, double* dv, size_t vec_stride, osph_gravity_stats* stats) {
    if (!g) return ERR_ARG;
    const size_t stride = vec_stride ? vec_stride : 4 * sizeof(double);
    if (dv) {
        if (stride < 4 * sizeof(double) || stride % sizeof(double)) return g->fail(ERR_ARG, "bad vector stride");
        if (int e = upload_vectors(g, dv, stride, GAX, 3, g->n)) return e;
        if (int e = sync(g)) return e;
    }
    if (int e = eval_self(g, stats)) return e;
    if (dv) return download_vectors(g, dv, stride, GAX, 3, g->n);
    return sync(g);
}

int osph_gravity_eval_points(osph_gravity* g, const double* p, size_t np, double* a, size_t vec_stride) {
    if (!g || (np && (!p || !a))) return ERR_ARG;
    const size_t stride = vec_stride ? vec_stride : 4 * sizeof(double);
    if (stride < 4 * sizeof(double) || stride % sizeof(double)) return g->fail(ERR_ARG, "bad vector stride");
    if (!np) return OK;
    if (g->cfg.solver != OSPH_GRAVITY_BARNES_HUT) {
        return g->fail(ERR_STATE, "point evaluation is provided by the Barnes-Hut solver");
    }
    if (!g->built) return g->fail(ERR_STATE, "gravity not built");
    if (int e = ensure_scratch(g, 8 * np)) return e;
    double4* pts = reinterpret_cast<double4*>(g->scratch);
    double4* acc = pts + np;
    std::vector<double4> hp(np), ha(np);
    for (size_t k = 0; k < np; ++k) {
        const double* e = p + k * (stride / sizeof(double));
        hp[k] = make_double4(e[0], e[1], e[2], e[3]);
    }
    GCHK(g, hipMemcpyAsync(pts, hp.data(), np * sizeof(double4), hipMemcpyHostToDevice, g->stream));
    if (g->n == 0) {
        for (size_t k = 0; k < np; ++k) {
            double* e = a + k * (stride / sizeof(double));
            e[0] = e[1] = e[2] = e[3] = 0.0;
        }
        return sync(g);
    }
    const double* grads = g->cfg.kernel == OSPH_GRAVITY_KERNEL_M4 ? g->lut_grads : nullptr;
    Tree t = tree_of(g);
    timed(g, "points", [&] { k_points<<<grid_for(np), BLOCK, 0, g->stream>>>(t, pts, acc, np, grads, g->cfg.order); });
    GCHK(g, hipMemcpyAsync(ha.data(), acc, np * sizeof(double4), hipMemcpyDeviceToHost, g->stream));
    if (int e = check_traversal(g)) return e;
    for (size_t k = 0; k < np; ++k) {
        double* e = a + k * (stride / sizeof(double));
        e[0] = ha[k].x;
        e[1] = ha[k].y;
        e[2] = ha[k].z;
        e[3] = 0.0;
    }
    return OK;
}

int osph_gravity_direct(osph_gravity* g, const uint32_t* t, size_t nt, double* a, size_t vec_stride) {
    if (!g || (nt && (!t || !a))) return ERR_ARG;
    const size_t stride = vec_stride ? vec_stride : 4 * sizeof(double);
    if (stride < 4 * sizeof(double) || stride % sizeof(double)) return g->fail(ERR_ARG, "bad vector stride");
    for (size_t k = 0; k < nt; ++k) {
        if (t[k] >= g->n) return g->fail(ERR_ARG, "target index out of range");
    }
    if (!nt) return OK;
    if (int e = ensure_scratch(g, 4 * nt + nt / 2 + 1)) return e;
    double4* out = reinterpret_cast<double4*>(g->scratch);
    uint32_t* tg = reinterpret_cast<uint32_t*>(g->scratch + 4 * nt);
    GCHK(g, hipMemcpyAsync(tg, t, nt * sizeof(uint32_t), hipMemcpyHostToDevice, g->stream));
    if (!g->identity) k_to_slots<<<grid_for(nt), BLOCK, 0, g->stream>>>(tg, g->slot_of, nt);
    if (int e = eval_direct(g, tg, nt, out)) return e;
    std::vector<double4> ha(nt);
    GCHK(g, hipMemcpyAsync(ha.data(), out, nt * sizeof(double4), hipMemcpyDeviceToHost, g->stream));
    if (int e = sync(g)) return e;
    for (size_t k = 0; k < nt; ++k) {
        double* e = a + k * (stride / sizeof(double));
        e[0] = ha[k].x;
        e[1] = ha[k].y;
        e[2] = ha[k].z;
        e[3] = 0.0;
    }
    return OK;
}

int osph_gravity_energy(osph_gravity* g, double* energy) {
    if (!g || !energy) return ERR_ARG;
    const size_t n = g->n;
    *energy = 0.0;
    if (!n) return OK;
    const unsigned blocks = std::min<unsigned>(grid_for(n), 1024);
    if (int e = ensure_scratch(g, n + blocks + 1)) return e;
    const double* values = g->cfg.kernel == OSPH_GRAVITY_KERNEL_M4 ? g->lut_values : nullptr;
    timed(g, "energy", [&] {
        k_energy<<<grid_for(n), BLOCK, 0, g->stream>>>(g->s, n, values, g->scratch);
        k_sum_vector<<<blocks, BLOCK, 0, g->stream>>>(g->scratch, n, g->scratch + n);
        k_sum_partials<<<1, 64, 0, g->stream>>>(g->scratch + n, int(blocks), 1, g->scratch + n + blocks);
    });
    double sum = 0.0;
    GCHK(g, hipMemcpyAsync(&sum, g->scratch + n + blocks, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    if (int e = sync(g)) return e;
    *energy = 0.5 * g->cfg.G * sum;
    return OK;
}

int osph_gravity_moments(osph_gravity* g, double out[16]) {
    if (!g || !out) return ERR_ARG;
    if (g->cfg.solver != OSPH_GRAVITY_BARNES_HUT || !g->built) return g->fail(ERR_STATE, "no Barnes-Hut tree built");
    std::memset(out, 0, 16 * sizeof(double));
    if (g->n == 0) return OK;
    if (g->n == 1) {
        double4 p;
        double m;
        GCHK(g, hipMemcpy(&p, g->pos, sizeof(p), hipMemcpyDeviceToHost));
        GCHK(g, hipMemcpy(&m, g->gm, sizeof(m), hipMemcpyDeviceToHost));
        out[0] = m / g->cfg.G;
        out[1] = p.x;
        out[2] = p.y;
        out[3] = p.z;
        return OK;
    }
    double4 geo;
    NodeMom mom;
    GCHK(g, hipMemcpy(&geo, g->geo, sizeof(geo), hipMemcpyDeviceToHost));
    GCHK(g, hipMemcpy(&mom, g->mom, sizeof(mom), hipMemcpyDeviceToHost));
    const double inv = 1.0 / g->cfg.G;  // BarnesHut::getMoments divides the premultiplied moments by G
    out[0] = mom.M * inv;
    out[1] = geo.x;
    out[2] = geo.y;
    out[3] = geo.z;
    for (int k = 0; k < 5; ++k) out[4 + k] = mom.q2[k] * inv;
    for (int k = 0; k < 7; ++k) out[9 + k] = mom.q3[k] * inv;
    return OK;
}

int osph_gravity_step_leapfrog(osph_gravity* g, double dt, int steps) {
    if (!g) return ERR_ARG;
    if (!(dt > 0.0) || !std::isfinite(dt) || steps < 0) return g->fail(ERR_ARG, "bad time step or step count");
    const size_t n = g->n;
    if (!n) return OK;
    for (int s = 0; s < steps; ++s) {
        timed(g, "drift", [&] { k_drift<<<grid_for(n), BLOCK, 0, g->stream>>>(g->s, dt, n); });
        k_zero_acc<<<grid_for(n), BLOCK, 0, g->stream>>>(g->s, n);
        if (g->cfg.solver == OSPH_GRAVITY_BARNES_HUT) {
            if (int e = build_tree(g, true)) return e;
        }
        if (int e = eval_self(g, nullptr)) return e;
        timed(g, "kick_drift", [&] { k_kick_drift<<<grid_for(n), BLOCK, 0, g->stream>>>(g->s, dt, n); });
    }
    return sync(g);
}

int osph_gravity_integrals(osph_gravity* g, double out[11]) {
    if (!g || !out) return ERR_ARG;
    std::memset(out, 0, 11 * sizeof(double));
    const size_t n = g->n;
    if (!n) return OK;
    const unsigned blocks = std::min<unsigned>(grid_for(n), 1024);
    if (int e = ensure_scratch(g, size_t(blocks) * 11 + 11)) return e;
    k_integrals<<<blocks, BLOCK, 0, g->stream>>>(g->s, n, g->scratch);
    k_sum_partials<<<1, 64, 0, g->stream>>>(g->scratch, int(blocks), 11, g->scratch + size_t(blocks) * 11);
    GCHK(g, hipMemcpyAsync(out, g->scratch + size_t(blocks) * 11, 11 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    if (int e = sync(g)) return e;
    // CenterOfMass::evaluate (Integrals.cpp:155-180): sum m r / sum m
    for (int k = 7; k < 10; ++k) out[k] /= out[10];
    return OK;
}

int osph_gravity_sync(osph_gravity* g) {
    if (!g) return ERR_ARG;
    return sync(g);
}

int osph_gravity_profile_enable(osph_gravity* g, int enable) {
    if (!g) return ERR_ARG;
    GCHK(g, hipStreamSynchronize(g->stream));
    resolve(g);
    g->prof = enable != 0;
    g->acc.clear();
    return OK;
}

int osph_gravity_profile_read(osph_gravity* g, char* names, size_t name_cap, double* ms, uint64_t* launches, int cap,
                              int* count) {
    if (!g || !count) return ERR_ARG;
    GCHK(g, hipStreamSynchronize(g->stream));
    resolve(g);
    int k = 0;
    for (const auto& kv : g->acc) {
        if (k < cap) {
            if (names && name_cap) {
                std::strncpy(names + size_t(k) * name_cap, kv.first.c_str(), name_cap - 1);
                names[size_t(k) * name_cap + name_cap - 1] = '\0';
            }
            if (ms) ms[k] = kv.second.first;
            if (launches) launches[k] = kv.second.second;
        }
        ++k;
    }
    *count = k;
    return OK;
}

}  // extern "C"
