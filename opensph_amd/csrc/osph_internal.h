// Internal definitions shared by the HIP kernels (osph_kernels.hip) and the C ABI (osph_capi.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "osph_gpu.h"

namespace osph {

// Device particle state: one plane of n values per component, in storage ("slot") order.
enum DPlane : int {
    X, Y, Z, H, VX, VY, VZ, VH, AX, AY, AZ,
    RHO, DRHO, U, DU,
    SXX, SYY, SXY, SXZ, SYZ,
    DSXX, DSYY, DSXY, DSXZ, DSYZ,
    D, DD,
    M, P, CS, RED, DIVV,
    GXX, GYY, GZZ, GXY, GXZ, GYZ,
    EPSMIN, MZERO, GROWTH,
    // highest derivatives saved by the predictor (PredictorCorrector::predictions)
    PAX, PAY, PAZ, PDRHO, PDU, PDSXX, PDSYY, PDSXY, PDSXZ, PDSYZ, PDD,
    NDPLANES
};
enum IPlane : int { FLAG, MAT, NCNT, NFLAWS, NIPLANES };

struct DevState {
    double* d[NDPLANES];
    uint32_t* u[NIPLANES];
};

// Per-particle record the pair kernels read for a neighbour, packed in cell-sorted order.
// 16 doubles = 128 B (one cache line pair): position, velocity, mass, density, sound speed,
// p / rho^2 and S / rho^2 (the per-particle factors of the StandardForceDiscr sums).
struct alignas(16) PairRec {
    double x, y, z, h;
    double vx, vy, vz, m;
    double rho, cs, pr;
    double sxx, syy, sxy, sxz, syz;
};
static_assert(sizeof(PairRec) == 128, "PairRec must be 128 bytes");

// Mixed-precision neighbour record, 64 B (four 16-byte loads per neighbour).  Positions and velocities
// are 32-bit fixed-point coordinates on power-of-two steps from the lower corner of their bounding boxes
// (Fixed): the difference of two coordinates is an exact integer, rounded once when it is converted to
// FP32, so r_i - r_j and v_j - v_i keep FP32 relative precision at any distance from the origin and at any
// bulk speed (bodies move at km/s with m/s internal differences, which absolute FP32 values would round
// away).  The four low mantissa bits of cs carry the particle's group code (group_code): cs is rounded to
// 20 bits (relative 1e-6), and it enters only the mean sound speed of the artificial viscosity.
struct alignas(16) PairRecF {
    uint32_t x, y, z;        // position lo + (x, y, z) * step
    float h;                 // smoothing length
    uint32_t vx, vy, vz;     // velocity vlo + (vx, vy, vz) * vstep
    float m;                 // mass
    float rho, cs, pr, sxx;  // pr = p / rho^2, s.. = S / rho^2; cs with the group code in its low 4 bits
    float syy, sxy, sxz, syz;
};
static_assert(sizeof(PairRecF) == 64, "PairRecF must be 64 bytes");

// FP32 screen of the neighbour test: positions relative to the origin o, and the band [lo, hi] = [1 - b,
// 1 + b] around the squared radius outside which the FP32 comparison equals the exact FP64 one.
struct Screen {
    double o[3];
    float lo, hi;
    int on;
};

// Fixed-point frames of the mixed-precision records: power-of-two steps such that the bounding box of
// the positions (of the velocities) spans at most 2^31 - 2 steps per axis, so that the difference of any
// two coordinates is an exact int32.
struct Fixed {
    double lo[3];
    double step, inv_step;
    float stepf;
    double vlo[3];
    double vstep, inv_vstep;
    float vstepf;
};

// Particles are binned into smoothing-length classes, class = min(nclass - 1, floor(log2(h / h_min))).
// Each class has its own uniform cell grid (cell edge 2 h_max(class) / divisor) keyed by row_key.
constexpr int MAX_CLASSES = 4;

struct GridClass {
    double lo[3], hi[3];  // bounding box of the class members
    double hmax;          // largest h in the class
    double cell, inv_cell;
    int dim[3];
    int bits;             // 2^bits > every dim: x slots per row, and the width of y and z in the row key
    uint32_t base;        // offset of the class's cells in the cell tables and in the sort keys
    int empty;
};

// The sort key of a particle is its cell index: base(class) + row_key(x, y, z), the classes' key ranges
// packed back to back (each as long as its largest row key + 1), so the keys are dense and as short
// as the cell count allows (fewer radix-sort passes).
struct Grid {
    int nclass;
    int key_bits;  // bits of the sort key that can be nonzero
    double hmin;
    GridClass c[MAX_CLASSES];
};

// Run-time options the kernels read.
struct Params {
    int solid_stress, adaptive_h, sum_only_undamaged, has_reduce, has_damage;
    double av_alpha, av_beta, h_lo, h_hi;
    int criteria;
    double courant, deriv_factor, div_factor, max_dt;
    int add_accel;  // the pair sums add to the accelerations already in the planes (attached self-gravity)
};

constexpr int LUT_ENTRIES = 40000;
constexpr uint32_t NO_GROUP = 0xffffffffu;
// Group code of the mixed-precision records (4 bits): group g < 14 -> g + 1, NO_GROUP -> 15, and 0 for
// the larger groups, which the pair kernel compares through the group array.
constexpr uint32_t GROUP_CODE_NONE = 15u;
constexpr uint32_t GROUP_CODE_WIDE = 0u;
__host__ __device__ inline uint32_t group_code(uint32_t g) {
    return g == NO_GROUP ? GROUP_CODE_NONE : (g < 14u ? g + 1u : GROUP_CODE_WIDE);
}

__host__ __device__ inline uint32_t morton_spread(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__host__ __device__ inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return morton_spread(x) | (morton_spread(y) << 1) | (morton_spread(z) << 2);
}

// The inverse of spread2: the even bits of v, packed.
__host__ __device__ inline uint32_t compact2(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}

__host__ __device__ inline uint32_t spread2(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Cell key of a neighbour grid with 2^bits cells per axis: the rows of cells along x are kept whole and
// ordered by the Morton code of (y, z), so that every x-run of cells is one contiguous range of sorted
// particles.
__host__ __device__ inline uint32_t row_key(uint32_t x, uint32_t y, uint32_t z, int bits) {
    return ((spread2(y) | (spread2(z) << 1)) << bits) | x;
}

// Neighbour-list layout: quad j (entries 4j .. 4j + 3) of sorted particle k, with Q = kmax / 4 quads
// per particle, is uint4 element ((k / LIST_TILE) * Q + j) * LIST_TILE + k % LIST_TILE.  Tiles of
// LIST_TILE consecutive particles hold all their quads, quad-major: the list kernel's lanes fill whole
// lines when they store their j-th quads, and the pair kernels still read each particle's quads from
// one small region.  The list holds round_up(n, LIST_TILE) * kmax entries.
constexpr uint32_t LIST_TILE = 8;
// The sorted position arrays (double4, float4) hold this many entries past the last particle, so that the
// list kernel loads whole batches of candidates without clamping each index (entries past a run are
// loaded and never used).
constexpr size_t POS_SLACK = 4;
__host__ __device__ inline size_t list_quad(size_t k, uint32_t j, uint32_t quads) {
    return ((k / LIST_TILE) * quads + j) * LIST_TILE + (k % LIST_TILE);
}

// ---- launchers implemented in osph_kernels.hip -------------------------------------------------------
// Launch shapes of the pair-stage kernels, chosen by the tuner (osph_tune): threads per block and
// dynamic LDS per block (which caps the number of resident blocks per CU, and with it the size of the
// neighbour-record working set competing for each XCD's L2).
struct Tuning {
    int pairs_block = 256;
    int pairs_lds = 0;
    int pairs_lanes = 1;  // lanes per particle in the pair kernels (1, 2 or 4)
    int lists_block = 256;
    int lists_lds = 0;
    int cell_divisor = 3;  // cells of 2 h_max(class) / divisor
    int lists_screen = -1; // FP32 screen in the list kernel: 1 on, 0 off, -1 on for mixed precision only
    int reorder_period = 40;  // move the state into cell order every this many searches (0: never); at the
                              // 1M impact 10 / 20 / 40 / 80 measured 1.738 / 1.724 / 1.723 / 1.736 ms per
                              // FP64 step and 1.243 / 1.223 / 1.216 / 1.222 ms per mixed step (fused)
    int fuse_finalize = 1;    // PC step: finalize, corrector and criteria in the pair kernel: 1 always, 0 never,
                              // -1 for FP64 only (the mixed step measured 1.246 ms unfused, 1.234 ms fused)
    int batch_graph = 1;      // batched steps replay their stages after the sort as a captured HIP graph
};

struct Launch {
    hipStream_t stream;
    Tuning tune{};
};

// Accumulators of one step (64 words of device memory): SMALL_BOUNDS + 0..7 global bounds,
// SMALL_CRITERIA + 0..3 the criterion minima, SMALL_OVERFLOW the list overflow flag (low 32 bits),
// SMALL_CLASSES + CLASS_STRIDE * c + 0..8 the bounds of class c, SMALL_VEL + 0..5 the velocity bounds
// (min x, y, z, max x, y, z; mixed precision only).
constexpr int SMALL_BOUNDS = 0;
constexpr int SMALL_CRITERIA = 8;
constexpr int SMALL_OVERFLOW = 12;
constexpr int SMALL_CLASSES = 16;
constexpr int CLASS_STRIDE = 10;
constexpr int SMALL_VEL = 56;
constexpr int SMALL_WORDS = 64;
static_assert(SMALL_CLASSES + CLASS_STRIDE * MAX_CLASSES <= SMALL_VEL, "accumulator layout");
static_assert(SMALL_VEL + 6 <= SMALL_WORDS, "accumulator layout");

// Clears every accumulator (min slots to all ones, max and count slots to zero) in one launch.
void launch_init_small(const Launch& l, unsigned long long* small);
// acc[0..2] = min x, y, z; acc[3..5] = max x, y, z; acc[6] = max h; acc[7] = min h (order-encoded).
void launch_bounds_reduce(const Launch& l, DevState s, unsigned long long* acc, size_t n);
// mode 2: predictions + material initialize; 1: material initialize; 0: neither.  Then the per-class
// grid bounds (class of h relative to href) into small[SMALL_CLASSES + CLASS_STRIDE * c + 0..8] and, with
// vel, the velocity bounds into small[SMALL_VEL + 0..5]; the step's other accumulators (small[0 ..
// SMALL_CLASSES - 1]) start at their identities.
void launch_advance(const Launch& l, DevState s, const osph_material* mats, Params prm, int mode, double dt,
                    double href, unsigned long long* small, bool vel, unsigned long long* part, size_t n);
// Entries of the per-block partial bounds that launch_advance needs for n particles.
size_t bounds_partials(size_t n);
// Values of the per-block partial bounds (partials per block).
size_t bounds_values();

// ---- adjoint (osph_adjoint.hip) -----------------------------------------------------------------------
// Materials whose Tillotson parameters (a, b, A) get a gradient.
constexpr int ADJ_MAX_MATERIALS = 4;
// Bytes of one adjoint neighbour record (n of them, in sorted order).
size_t adj_rec_bytes();
// Doubles of the per-block partials of the parameter gradient for n particles.
size_t adj_partials(size_t n);
// One reverse PC step: lplanes (13: lambda of r xyz, v xyz, a xyz, rho, drho, u, du) hold the adjoints of
// X_{n+1} and receive those of X_n; the state holds the predicted X* with the derivatives evaluated at it
// (and its neighbour lists); gplanes (9) are scratch; theta[3 mat + q] accumulates d/d(a, b, A).  fp32:
// the pair terms in FP32 (FP64 separations, FP64 sums), the mixed adjoint.
void launch_adj_step(const Launch& l, DevState s, double* const* lplanes, double* const* gplanes,
                     const uint32_t* order, const uint32_t* list, const uint32_t* count, uint32_t kmax,
                     const double* lut, double alpha, double beta, const osph_material* mats, int nmat, double dt,
                     void* rec, double* part, double* theta, unsigned* clamped, size_t n, bool fp32);
// The 13 planes of X (r, v, a, rho, drho, u, du) to and from a block of 13 planes of capacity cap.
void launch_save_x(const Launch& l, DevState s, double* block, size_t cap, size_t n);
void launch_restore_x(const Launch& l, DevState s, const double* block, size_t cap, size_t n);
// Device arrays in the caller's order (nc doubles per element; src null: zeros) to and from planes.
void launch_scatter_planes(const Launch& l, const double* src, int nc, double* const* planes, const uint32_t* slot_of,
                           size_t n);
void launch_gather_planes(const Launch& l, double* dst, int nc, double* const* planes, const uint32_t* slot_of,
                          size_t n);
// Cell keys and the identity permutation; also clears the first `clear` words of `range` (the row
// ranges that launch_cell_first fills).
void launch_cell_keys(const Launch& l, DevState s, Grid g, uint32_t* keys, uint32_t* idx, uint32_t* range,
                      size_t clear, size_t n);
void launch_clear_u32(const Launch& l, uint32_t* p, size_t count, uint32_t value);
// first[cell] = the first sorted position whose cell index is >= cell, for every cell (x, y, z) of a
// row (y < dim_y, z < dim_z) with x in [0, dim_x]; row_range (2 per row) is scratch.
void launch_cell_first(const Launch& l, const uint32_t* keys, Grid g, uint32_t* row_range, uint32_t* first,
                       size_t rows, size_t n);
void launch_pack(const Launch& l, DevState s, Params prm, const uint32_t* order, PairRec* rec, double4* pos,
                 float4* posf, Screen sc, uint32_t* group, size_t n);
// Neighbour lists are row-major: list[k * kmax + c], kmax a multiple of 4.
void launch_neighbor_lists(const Launch& l, const double4* pos, const float4* posf, Screen sc, Grid g,
                           const uint32_t* first, uint32_t* list, uint32_t* count, uint32_t kmax, uint32_t* overflow,
                           size_t n);
// With fuse set, the pair kernels go on, per particle, with what k_finalize_correct does (the finalize
// of every term, the corrector with time step dt and the time-step criteria into mins), without the
// round trip of the sums through memory; skipped, like k_finalize_correct, while *overflow is set.
struct FuseArgs {
    const osph_material* mats;
    double dt;
    unsigned long long* mins;
    const uint32_t* overflow;
};
void launch_pairs_fp64(const Launch& l, DevState s, Params prm, const PairRec* rec, const uint32_t* group,
                       const uint32_t* order, const uint32_t* list, const uint32_t* count, const double* lut_grads,
                       uint32_t kmax, size_t n, const FuseArgs* fuse = nullptr);
void launch_pack_mixed(const Launch& l, DevState s, Params prm, Fixed fx, const uint32_t* order, PairRecF* rec,
                       double4* pos, float4* posf, Screen sc, uint32_t* group, size_t n);
void launch_pairs_mixed(const Launch& l, DevState s, Params prm, Fixed fx, const PairRecF* rec, const uint32_t* group,
                        const uint32_t* order, const uint32_t* list, const uint32_t* count, uint32_t kmax, size_t n,
                        const FuseArgs* fuse = nullptr);
void launch_finalize(const Launch& l, DevState s, const osph_material* mats, Params prm, size_t n);
void launch_zero_derivatives(const Launch& l, DevState s, size_t n);
// Finalize + corrections + criteria of a predictor-corrector step; a no-op while *overflow != 0.
void launch_finalize_correct(const Launch& l, DevState s, const osph_material* mats, Params prm, double dt,
                             unsigned long long* mins, const uint32_t* overflow, size_t n);
void launch_criteria(const Launch& l, DevState s, const osph_material* mats, Params prm,
                     unsigned long long* mins, size_t n);
// Host element i <-> device slot slot_of[i].
void launch_stage_to_planes(const Launch& l, const unsigned char* staging, size_t stride, int ncomp,
                            const int* offsets_bytes, int elem_bytes, void* const* dst, const uint32_t* slot_of,
                            size_t n);
void launch_planes_to_stage(const Launch& l, unsigned char* staging, size_t stride, int ncomp,
                            const int* offsets_bytes, int elem_bytes, void* const* src, const uint32_t* slot_of,
                            size_t n);
// dst slot k <- src slot order[k] for every plane; perm_out[k] = perm_in[order[k]]; slot_of[perm_out[k]] = k.
void launch_permute_state(const Launch& l, DevState src, DevState dst, const uint32_t* order, const uint32_t* perm_in,
                          uint32_t* perm_out, uint32_t* slot_of, size_t n);
void launch_iota(const Launch& l, uint32_t* p, size_t n);
void launch_stream_copy(hipStream_t stream, const void* a, void* b, size_t bytes);
// Integrals of the state into scratch (integrals_scratch_doubles(n) doubles); the 12 results are the
// last 12 doubles.
size_t integrals_scratch_doubles(size_t n);
void launch_integrals(const Launch& l, DevState s, double* scratch, size_t n);
// Temporary storage of the radix sort of n pairs by the low key_bits bits (the algorithm, and with it
// the storage, depends on both).
size_t sort_pairs_temp_bytes(size_t n, int key_bits);
hipError_t sort_pairs(const Launch& l, void* temp, size_t temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                      const uint32_t* vals_in, uint32_t* vals_out, size_t n, int key_bits);

// ---- batched stepping: one launch per stage for every member of an ensemble -----------------------------
// What the kernels of one predictor-corrector step read for one member of a batch (in device memory, one
// entry per member).  Every batched kernel runs the same per-particle code as the single-run kernel,
// with the member's own buffers and parameters, so each member's results are those of stepping it alone.
struct BatchMember {
    DevState s;    // the state before this step's reorder
    DevState s2;   // the state from this step's reorder on (= s when the step does not reorder)
    const osph_material* mats;
    Params prm;
    double dt, href;
    Grid grid;
    Screen screen;
    Fixed fx;
    uint32_t n, kmax;
    uint32_t rows;    // rows of the member's grid (all classes)
    int xslots;       // largest dim_x + 1 over its classes
    int reorder;      // this step moves the state into cell order (s -> s2)
    uint32_t key_off; // first position of the member in the batch's sort arrays
    unsigned long long* small;  // accumulators (SMALL_* layout)
    unsigned long long* part;   // per-block partial bounds
    uint32_t *keys_sorted, *order, *row_range, *cell_first;
    uint32_t *perm, *perm2, *slot_of;
    void* rec;        // PairRec (FP64) or PairRecF (mixed precision)
    double4* pos;
    float4* posf;
    uint32_t* group;
    uint32_t *list, *count;
};

// Block b of a batched launch belongs to member member[b]; the member's blocks are first[m] ..
// first[m + 1] - 1, a multiple of 8 of them, so that each member's blocks are dealt to the 8 XCDs as in
// a launch of its own.
struct BatchBlocks {
    const uint32_t* member;
    const uint32_t* first;
    uint32_t total;
};
// Blocks of `block` threads for n items, rounded up to a multiple of 8.
inline uint32_t batch_blocks_for(size_t n, int block) {
    const size_t g = (n + size_t(block) - 1) / size_t(block);
    return uint32_t(((g > 0 ? g : 1) + 7) / 8 * 8);
}
// Stage launchers over all `count` members (M in device memory).  `pb`: blocks of BLOCK threads per
// particle; `lb`: blocks of the list kernel's size; `qb`: blocks of the pair kernel's size and lanes.
void batch_advance(const Launch& l, const BatchMember* M, int count, BatchBlocks pb, bool vel);
void batch_cell_keys(const Launch& l, const BatchMember* M, BatchBlocks pb, int key_bits, uint64_t* keys,
                     uint32_t* vals);
void batch_split(const Launch& l, const BatchMember* M, BatchBlocks pb, int key_bits, const uint64_t* keys,
                 const uint32_t* vals);
void batch_cells(const Launch& l, const BatchMember* M, BatchBlocks pb);
void batch_permute(const Launch& l, const BatchMember* M, BatchBlocks pb);
void batch_pack(const Launch& l, const BatchMember* M, BatchBlocks pb, bool mixed);
void batch_lists(const Launch& l, const BatchMember* M, BatchBlocks lb, bool screened, bool plain);
void batch_pairs(const Launch& l, const BatchMember* M, BatchBlocks qb, bool mixed, bool fuse, const double* lut);
void batch_finalize_correct(const Launch& l, const BatchMember* M, BatchBlocks pb);
size_t sort_pairs64_temp_bytes(size_t n, int key_bits);
hipError_t sort_pairs64(const Launch& l, void* temp, size_t temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                        const uint32_t* vals_in, uint32_t* vals_out, size_t n, int key_bits);

// Grady-Kipp flaws of one body on the device (osph_set_flaws): flaw p (1-based) hits particle
// flaw_target(seed, p, nb); the first pass records first[i] = min p (first initialised to ~0), cover
// counts the particles still without a flaw and the largest first flaw into acc[0], acc[1]; the count
// pass records the count and the last flaw of p <= K; finish writes the damage planes.
int flaws_first_pass(const Launch& l, uint64_t seed, uint64_t p0, uint64_t count, uint32_t nb, uint32_t* first);
void flaws_cover(const Launch& l, const uint32_t* first, uint32_t nb, uint32_t* acc);
void flaws_count_pass(const Launch& l, uint64_t seed, uint32_t K, uint32_t nb, uint32_t* cnt, uint32_t* last);
void flaws_finish(const Launch& l, DevState s, const uint32_t* slot_of, size_t body_first, uint32_t nb,
                  const uint32_t* first, const uint32_t* last, const uint32_t* cnt, double denom, double inv_m,
                  double cg);

// ---- initial conditions (osph_initial.hip) -------------------------------------------------------------
// The hexagonal lattice of one spherical body: axis values xs[nx], ys[ny], zs[nz] (device), the shifts of
// odd rows and layers, and the sphere (centre, squared radius) whose points are kept.
struct HcpLattice {
    const double *xs, *ys, *zs;
    size_t nx, ny, nz, points;  // points = nx ny nz
    double delta_x, delta_y;
    double cx, cy, cz, r2;
};
// Blocks of the lattice kernels (one thread per lattice point).
unsigned hcp_blocks(const HcpLattice& P);
// One left-to-right sum: a[0..n) (times w[i] when w is set).
struct SerialSum {
    const double* a;
    const double* w;
    size_t n;
};
// What k_body_finish writes: the translation t (move), and (fill) the per-body constants.
struct BodyFill {
    int move, fill;
    double t[3];
    double h, m, rho;
    double v[3];
    uint32_t flag, mat;
};
// block_count[b] = the inside points among lattice indices [256 b, 256 b + 256).
void launch_hcp_count(const Launch& l, const HcpLattice& P, uint32_t* block_count);
// The inside points, in lattice order, to slots block_first[b] + rank of the X, Y, Z planes; H = h.
void launch_hcp_fill(const Launch& l, const HcpLattice& P, const uint32_t* block_first, DevState s, double h);
void launch_serial_sums(const Launch& l, const SerialSum* jobs, int count, double* out);
// atomicMax of the order-encoded squared distances |r_i - c|^2 into *out (which starts at 0).
void launch_max_dist2(const Launch& l, const double* x, const double* y, const double* z, size_t n, const double c[3],
                      unsigned long long* out);
void launch_body_finish(const Launch& l, DevState s, size_t first, size_t n, const BodyFill& f);

}  // namespace osph

// The gravity half of GravitySolver::loop for particles in another context's device planes (defined with
// the self-gravity, osph_gravity.hip): src = x, y, z, h, m (n doubles each); the copy waits for `ready`,
// `done` is recorded on the gravity context's stream after the evaluation, and accel[0..2] then point
// to the accelerations in the order of the source planes.  Returns a status of osph_gravity.h.
struct osph_gravity;
// The SPH context a gravity context is attached to (null: none).  One gravity context serves at most one
// SPH context: its planes hold that context's particles and accelerations between the evaluation and the
// pair sums.
const void* gravity_owner(const osph_gravity* g);
void gravity_set_owner(osph_gravity* g, const void* owner);
int gravity_self_from_device(osph_gravity* g, size_t n, const double* const src[5], hipEvent_t ready,
                             hipEvent_t done, const double* accel[3]);
