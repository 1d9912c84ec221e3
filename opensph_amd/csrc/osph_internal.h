// Internal definitions shared by the HIP kernels (osph_kernels.hip) and the C ABI (osph_capi.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/osph_gpu.h"

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

// Mixed-precision neighbour record: cell-relative FP32 position, FP32 fields, 64 B.
struct alignas(16) PairRecF {
    float x, y, z, h;
    float vx, vy, vz, m;
    float rho, cs, pr, pad;
    float sxx, syy, sxy, sxz;
};

// Particles are binned into smoothing-length classes, class = min(nclass - 1, floor(log2(h / h_min))).
// Each class has its own uniform cell grid (cell edge >= 2 h_max of the class) keyed by the Morton code
// of the cell coordinates; the sort key is (class << 30) | morton.
constexpr int MAX_CLASSES = 4;
constexpr int CLASS_SHIFT = 30;

struct GridClass {
    double lo[3], hi[3];  // bounding box of the class members
    double hmax;          // largest h in the class
    double cell, inv_cell;
    int dim[3];
    int bits;             // bits per axis of the Morton key
    uint32_t base;        // offset of the class's cells in the cell tables
    int empty;
};

struct Grid {
    int nclass;
    double hmin;
    GridClass c[MAX_CLASSES];
};

// Run-time options the kernels read.
struct Params {
    int solid_stress, adaptive_h, sum_only_undamaged, has_reduce, has_damage;
    double av_alpha, av_beta, h_lo, h_hi;
    int criteria;
    double courant, deriv_factor, div_factor, max_dt;
};

constexpr int LUT_ENTRIES = 40000;
constexpr uint32_t NO_GROUP = 0xffffffffu;

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

// ---- launchers implemented in osph_kernels.hip -------------------------------------------------------
struct Launch {
    hipStream_t stream;
};

void launch_material_init(const Launch& l, DevState s, const osph_material* mats, Params prm, size_t n);
void launch_bounds_reduce(const Launch& l, DevState s, unsigned long long* acc, size_t n);
void launch_class_bounds(const Launch& l, DevState s, int nclass, double hmin, unsigned long long* acc, size_t n);
void launch_cell_keys(const Launch& l, DevState s, Grid g, uint32_t* keys, uint32_t* idx, size_t n);
void launch_clear_u32(const Launch& l, uint32_t* p, size_t count, uint32_t value);
void launch_cell_ranges(const Launch& l, const uint32_t* keys, Grid g, uint32_t* cstart, uint32_t* cend, size_t n);
void launch_pack(const Launch& l, DevState s, Params prm, const uint32_t* order, PairRec* rec, uint32_t* group,
                 size_t n);
void launch_neighbor_lists(const Launch& l, const PairRec* rec, Grid g, const uint32_t* cstart, const uint32_t* cend,
                           uint32_t* list, uint32_t* count, uint32_t kmax, uint32_t* overflow, size_t n);
void launch_pairs_fp64(const Launch& l, DevState s, Params prm, const PairRec* rec, const uint32_t* group,
                       const uint32_t* order, const uint32_t* list, const uint32_t* count, const double* lut_grads,
                       uint32_t kmax, size_t n);
void launch_finalize(const Launch& l, DevState s, const osph_material* mats, Params prm, size_t n);
void launch_zero_derivatives(const Launch& l, DevState s, size_t n);
void launch_predict(const Launch& l, DevState s, const osph_material* mats, Params prm, double dt, size_t n);
void launch_correct(const Launch& l, DevState s, const osph_material* mats, Params prm, double dt,
                    unsigned long long* mins, size_t n);
void launch_criteria(const Launch& l, DevState s, const osph_material* mats, Params prm,
                     unsigned long long* mins, size_t n);
void launch_stage_to_planes(const Launch& l, const unsigned char* staging, size_t stride, int ncomp,
                            const int* offsets_bytes, int elem_bytes, void* const* dst, size_t n);
void launch_planes_to_stage(const Launch& l, unsigned char* staging, size_t stride, int ncomp,
                            const int* offsets_bytes, int elem_bytes, void* const* src, size_t n);
void launch_stream_copy(hipStream_t stream, const void* a, void* b, size_t bytes);
size_t sort_pairs_temp_bytes(size_t n);
void sort_pairs(const Launch& l, void* temp, size_t temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                const uint32_t* vals_in, uint32_t* vals_out, size_t n);

}  // namespace osph
