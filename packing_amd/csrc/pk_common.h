// Common definitions for the packing kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pk {

constexpr int MAXP = 16;        // max convex pieces per polygon
constexpr int MAXV = 4;         // max vertices per convex piece
constexpr int MAXO = 64;        // max outline vertices (boundary term)
constexpr int WAVE = 64;

// Error reporting for the C-ABI: every entry point returns 0 or a negative code and
// leaves a message readable through pk_last_error().
enum Status : int {
    PK_OK = 0,
    PK_EINVAL = -1,
    PK_EHIP = -2,
    PK_ENOMEM = -3,
    PK_ERANGE = -4,
};

void set_error(const std::string& msg);

#define PK_CHECK_HIP(expr)                                                            \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            ::pk::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));      \
            return ::pk::PK_EHIP;                                                     \
        }                                                                             \
    } while (0)

#define PK_REQUIRE(cond, msg)                                                         \
    do {                                                                              \
        if (!(cond)) {                                                                \
            ::pk::set_error(msg);                                                     \
            return ::pk::PK_EINVAL;                                                   \
        }                                                                             \
    } while (0)

// Polygon tables in device memory.  Pieces are CCW, padded to MAXV vertices by
// repeating the last vertex (a zero-length edge, which the kernels treat as neutral).
template <typename T>
struct GeomT {
    int P;                        // number of pieces
    int n_outline;                // outline vertex count
    T radius;                     // bounding radius about the pose centre
    T cutoff_sq;                  // (2 radius)^2
    T px[MAXP * MAXV], py[MAXP * MAXV];   // local piece vertices
    T ccx[MAXP], ccy[MAXP], cr[MAXP];     // local piece bounding circles
    T ox[MAXO], oy[MAXO];                 // outline vertices
};

// Host handle: the same polygon in both precisions.
struct Geometry {
    GeomT<float> hf;
    GeomT<double> hd;
    GeomT<float>* df = nullptr;
    GeomT<double>* dd = nullptr;
    int device = 0;
};

// Deterministic accumulation: float contributions are summed as int64 fixed point
// (2^-40 resolution, +-8.4e6 range), so the total does not depend on the order in
// which lanes, waves or workgroup shapes deliver them.  Doubles use double atomics.
constexpr double FIX_SCALE = 1099511627776.0;   // 2^40
constexpr double FIX_INV = 1.0 / 1099511627776.0;

template <typename T>
struct Acc;

template <>
struct Acc<float> {
    using type = unsigned long long;
    __device__ static inline void add(type* p, float v) {
        long long q = __double2ll_rn((double)v * FIX_SCALE);
        atomicAdd(p, (unsigned long long)q);
    }
    __device__ static inline float get(type v) { return (float)((double)(long long)v * FIX_INV); }
    __device__ static inline type zero() { return 0ull; }
};

template <>
struct Acc<double> {
    using type = double;
    __device__ static inline void add(type* p, double v) { atomicAdd(p, v); }
    __device__ static inline double get(type v) { return v; }
    __device__ static inline type zero() { return 0.0; }
};

__host__ __device__ inline size_t align16(size_t v) { return (v + 15) / 16 * 16; }

__host__ __device__ inline float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
__host__ __device__ inline double fast_rcp(double x) { return 1.0 / x; }

__device__ inline void sincos_t(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ inline void sincos_t(double a, double* s, double* c) { sincos(a, s, c); }

__device__ inline uint64_t lanemask_lt() {
    int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace pk
