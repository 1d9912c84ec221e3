// Device-side helpers shared by the HIP translation units of libosph_gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace osph {

constexpr int BLOCK = 256;

inline unsigned grid_for(size_t n, int block = BLOCK) {
    size_t g = (n + block - 1) / block;
    return unsigned(g > 0 ? g : 1);
}

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).  Renumber them so that XCD x
// processes one contiguous range of logical blocks: consecutive sorted particles, and therefore their
// shared neighbours, stay in one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
    constexpr uint32_t NXCD = 8;
    const uint32_t q = nblocks / NXCD, r = nblocks % NXCD;
    const uint32_t x = bid % NXCD, local = bid / NXCD;
    return x * q + (x < r ? x : r) + local;
}

// Orders the LDS accesses of one wavefront around this point (no cross-wave synchronisation): a
// wavefront's LDS instructions execute in issue order, so after it every lane sees what every lane of
// the same wavefront wrote before it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Single IEEE operations rounded to nearest, never fused with a neighbouring one.  HIP's __dadd_rn and
// __dmul_rn are plain `+` and `*` in the device headers, so after inlining the compiler contracts a
// product and a sum of them into one FMA (seen as v_fmac_f64 in the gfx950 code): a squared distance
// summed that way differs in the last bit from the reference's separately rounded one.  The pragma keeps
// these operations out of any contraction.
__device__ __forceinline__ double rn_add(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double rn_sub(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double rn_mul(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }  // MathBasic.h:15-17
__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }  // MathBasic.h:20-22

// EPS and LARGE are float literals widened to double (core/math/MathUtils.h:26-34).
__device__ constexpr double EPS_D = double(1.e-12f);
__device__ constexpr double LARGE_D = double(1.e20f);
__device__ constexpr double PI_D = 3.14159265358979323846264338327950288;

// Order-preserving map of doubles onto unsigned 64-bit integers (for atomic min/max).
__device__ __forceinline__ unsigned long long ord_bits(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ unsigned long long wave_min(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Inverse of ord_bits (host side).
inline double decode_ord_host(unsigned long long e) {
    const unsigned long long b = (e >> 63) ? (e & 0x7fffffffffffffffull) : ~e;
    double v;
    __builtin_memcpy(&v, &b, sizeof(v));
    return v;
}

}  // namespace osph
