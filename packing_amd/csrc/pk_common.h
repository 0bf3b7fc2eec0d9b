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
    T ox[MAXO], oy[MAXO];                 // outline vertices, padded to MAXO with the last
    // Distinct piece vertices: pieces of a decomposition share most of their vertices
    // (and padding repeats one), so the evaluators rotate and store each distinct vertex
    // once per tree.  Vertex v of piece p is distinct vertex (pv[p] >> 8 v) & 0xff.
    int NU;
    uint32_t pv[MAXP];
    T ux[MAXP * MAXV], uy[MAXP * MAXV];
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
// Largest |term| that local() converts: pair terms stay below it because the geometry
// handle keeps every piece vertex within 8 of the tree origin and the table handle
// bounds the table's values and slopes (both checked when the handle is made).
constexpr float FIX_TERM_MAX = 2000.f;

template <typename T>
struct Acc;

template <>
struct Acc<float> {
    using type = unsigned long long;
    // round(v 2^40), ties to even, for |v| < 2^11: v 2^40 is exact in float and in double;
    // adding 1.5 2^52 leaves the nearest integer n in the low mantissa bits, so the bit
    // pattern is 0x4338000000000000 + n (two's complement), and that constant's low word
    // is zero.  Four instructions instead of the nine of a general double -> int64
    // conversion, with the same result in that range.
    __device__ static inline type local(float v) {
        const double d = (double)(v * 0x1p40f) + 0x1.8p52;
        const uint32_t hi = (uint32_t)__double2hiint(d) - 0x43380000u;
        return ((unsigned long long)hi << 32) | (uint32_t)__double2loint(d);
    }
    // the same rounding over the whole int64 range (boundary terms, which grow with the
    // distance of a tree outside its box)
    __device__ static inline type wide(float v) {
        return (unsigned long long)__double2ll_rn((double)v * FIX_SCALE);
    }
    __device__ static inline void add(type* p, float v) { atomicAdd(p, local(v)); }
    __device__ static inline void add_wide(type* p, float v) { atomicAdd(p, wide(v)); }
    __device__ static inline void add_raw(type* p, type v) { atomicAdd(p, v); }
    __device__ static inline float get(type v) { return (float)((double)(long long)v * FIX_INV); }
    __device__ static inline type zero() { return 0ull; }
    __device__ static inline type neg(type v) { return 0ull - v; }   // local(-x) = neg(local(x))
};

template <>
struct Acc<double> {
    using type = double;
    __device__ static inline type local(double v) { return v; }
    __device__ static inline type wide(double v) { return v; }
    __device__ static inline void add(type* p, double v) { atomicAdd(p, v); }
    __device__ static inline void add_wide(type* p, double v) { atomicAdd(p, v); }
    __device__ static inline void add_raw(type* p, type v) { atomicAdd(p, v); }
    __device__ static inline double get(type v) { return v; }
    __device__ static inline type zero() { return 0.0; }
    __device__ static inline type neg(type v) { return -v; }
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

// Orders LDS traffic between the lanes of one wave (no workgroup barrier).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-wide reductions in a fixed order (DPP within rows of 16 lanes, then the four row
// results combined as (r0 + r1) + (r2 + r3)); every lane returns the same value.  The
// full wave must be active.
template <int CTRL>
__device__ inline double dpp_d(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ inline double lane_d(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ inline double wave_sum(double v) {
    v += dpp_d<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);   // row_half_mirror
    v += dpp_d<0x140>(v);   // row_mirror
    return (lane_d(v, 0) + lane_d(v, 16)) + (lane_d(v, 32) + lane_d(v, 48));
}
__device__ inline double wave_max(double v) {
    v = fmax(v, dpp_d<0xB1>(v));
    v = fmax(v, dpp_d<0x4E>(v));
    v = fmax(v, dpp_d<0x141>(v));
    v = fmax(v, dpp_d<0x140>(v));
    return fmax(fmax(lane_d(v, 0), lane_d(v, 16)), fmax(lane_d(v, 32), lane_d(v, 48)));
}

__device__ inline double shfl_xor_d(double v, int mask) {
    const int lo = __shfl_xor(__double2loint(v), mask);
    const int hi = __shfl_xor(__double2hiint(v), mask);
    return __hiloint2double(hi, lo);
}
__device__ inline double shfl_d(double v, int src) {
    const int lo = __shfl(__double2loint(v), src);
    const int hi = __shfl(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ inline float lane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

constexpr int ilog2c(int k) { return k <= 1 ? 0 : 1 + ilog2c(k / 2); }

// Butterfly partner exchange for lane bit `bit` (a compile-time constant after unrolling):
// within a row of 16 lanes by DPP (quad permutes for bits 0 and 1, the half-row and row
// mirrors for bits 2 and 3: the partner differs in that bit and in lower bits only, which
// later stages handle), across rows by a lane permute.  Every value is then reduced by
// one fixed pairing tree.  The full wave must be active.
template <int CTRL>
__device__ inline int dpp_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
__device__ inline int xchg_i(int v, int bit) {
    switch (bit) {
        case 1: return dpp_i<0xB1>(v);    // quad_perm [1,0,3,2]
        case 2: return dpp_i<0x4E>(v);    // quad_perm [2,3,0,1]
        case 4: return dpp_i<0x141>(v);   // row_half_mirror
        case 8: return dpp_i<0x140>(v);   // row_mirror
        default: return __shfl_xor(v, bit);
    }
}
__device__ inline float xchg_f(float v, int bit) { return __int_as_float(xchg_i(__float_as_int(v), bit)); }
__device__ inline double xchg_d(double v, int bit) {
    return __hiloint2double(xchg_i(__double2hiint(v), bit), xchg_i(__double2loint(v), bit));
}

// One stage of the many-value butterfly below for lane bit `bit` on the pair (lo, hi):
// returns a (this lane's) and b (the partner's) such that a lane with the bit clear
// combines lo with its partner's lo and a lane with the bit set hi with its partner's hi.
// Across rows (bits 16 and 32) one v_permlane16/32_swap of the pair delivers both:
// it swaps the odd 16- (32-) lane rows of lo with the even rows of hi, leaving
// a = (lo_l, hi_{l-b}) and b = (lo_{l+b}, hi_l) on the clear and set lanes, the operands
// of the lane-exchange form in swapped order for the set lanes (sum and max are
// symmetric).  Within a row: a bitwise blend (a select between two elements of a register
// array can be rewritten by the compiler into a dynamically indexed load, which moves the
// whole array to scratch memory) and a DPP move.
__device__ inline void stage_i(int lo, int hi, int bit, bool up, int& a, int& b) {
    if (bit == 32) {
        const auto p = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
        a = (int)p[0];
        b = (int)p[1];
    } else if (bit == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
        a = (int)p[0];
        b = (int)p[1];
    } else {
        const int m = up ? ~0 : 0;
        a = (hi & m) | (lo & ~m);
        b = xchg_i((lo & m) | (hi & ~m), bit);
    }
}
__device__ inline void stage_f(float lo, float hi, int bit, bool up, float& a, float& b) {
    int ai, bi;
    stage_i(__float_as_int(lo), __float_as_int(hi), bit, up, ai, bi);
    a = __int_as_float(ai);
    b = __int_as_float(bi);
}
__device__ inline void stage_d(double lo, double hi, int bit, bool up, double& a, double& b) {
    int al, bl, ah, bh;
    stage_i(__double2loint(lo), __double2loint(hi), bit, up, al, bl);
    stage_i(__double2hiint(lo), __double2hiint(hi), bit, up, ah, bh);
    a = __hiloint2double(ah, al);
    b = __hiloint2double(bh, bl);
}


// Sums of K values (K a power of two, 2..64) over the 64 lanes of a wave in one
// butterfly.  At stage s the lanes l and its partner for bit b = 32 >> s (xchg_i: same
// higher bits, opposite bit b) exchange halves of their value lists and add, so after
// log2 K stages lane l holds value l >> (6 - log2 K); the remaining stages add it across
// the other lanes.  Every value is summed by the same pairing tree (partners for bit 5,
// then 4, ..., 0), and each pair sum is symmetric, so a value's total does not depend on
// K or on its position in the list.
// The full wave must be active.  Returns this lane's value.
template <int K>
__device__ inline double wave_sum_many(double (&v)[K]) {
    constexpr int LK = ilog2c(K);
    static_assert((1 << LK) == K && K >= 2 && K <= 64, "K must be a power of two in [2, 64]");
    const int lane = __lane_id();
#pragma unroll
    for (int s = 0; s < LK; ++s) {
        const int bit = 32 >> s;
        const int h = K >> (s + 1);
        const bool up = (lane & bit) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            double a, b;
            stage_d(v[j], v[h + j], bit, up, a, b);
            v[j] = a + b;
        }
    }
    double r = v[0];
#pragma unroll
    for (int s = LK; s < 6; ++s) {
        double a, b;
        stage_d(r, r, 32 >> s, (lane & (32 >> s)) != 0, a, b);
        r = a + b;
    }
    return r;
}
// The same for float partial sums, added in double.
template <int K>
__device__ inline double wave_sum_many(float (&v)[K]) {
    double dv[K];
#pragma unroll
    for (int j = 0; j < K; ++j) dv[j] = (double)v[j];
    return wave_sum_many<K>(dv);
}
// The same butterfly in float (one exchange per value and stage).
template <int K>
__device__ inline float wave_sum_many_f(float (&v)[K]) {
    constexpr int LK = ilog2c(K);
    static_assert((1 << LK) == K && K >= 2 && K <= 64, "K must be a power of two in [2, 64]");
    const int lane = __lane_id();
#pragma unroll
    for (int s = 0; s < LK; ++s) {
        const int bit = 32 >> s;
        const int h = K >> (s + 1);
        const bool up = (lane & bit) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            float a, b;
            stage_f(v[j], v[h + j], bit, up, a, b);
            v[j] = a + b;
        }
    }
    float r = v[0];
#pragma unroll
    for (int s = LK; s < 6; ++s) {
        float a, b;
        stage_f(r, r, 32 >> s, (lane & (32 >> s)) != 0, a, b);
        r = a + b;
    }
    return r;
}
// The same layout with max instead of sum (exact in any order).
template <int K>
__device__ inline double wave_max_many(double (&v)[K]) {
    constexpr int LK = ilog2c(K);
    static_assert((1 << LK) == K && K >= 2 && K <= 64, "K must be a power of two in [2, 64]");
    const int lane = __lane_id();
#pragma unroll
    for (int s = 0; s < LK; ++s) {
        const int bit = 32 >> s;
        const int h = K >> (s + 1);
        const bool up = (lane & bit) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            double a, b;
            stage_d(v[j], v[h + j], bit, up, a, b);
            v[j] = fmax(a, b);
        }
    }
    double r = v[0];
#pragma unroll
    for (int s = LK; s < 6; ++s) {
        double a, b;
        stage_d(r, r, 32 >> s, (lane & (32 >> s)) != 0, a, b);
        r = fmax(a, b);
    }
    return r;
}
// The same in float (max is exact, so the result equals the double version's).
template <int K>
__device__ inline float wave_max_many_f(float (&v)[K]) {
    constexpr int LK = ilog2c(K);
    static_assert((1 << LK) == K && K >= 2 && K <= 64, "K must be a power of two in [2, 64]");
    const int lane = __lane_id();
#pragma unroll
    for (int s = 0; s < LK; ++s) {
        const int bit = 32 >> s;
        const int h = K >> (s + 1);
        const bool up = (lane & bit) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            float a, b;
            stage_f(v[j], v[h + j], bit, up, a, b);
            v[j] = fmaxf(a, b);
        }
    }
    float r = v[0];
#pragma unroll
    for (int s = LK; s < 6; ++s) {
        float a, b;
        stage_f(r, r, 32 >> s, (lane & (32 >> s)) != 0, a, b);
        r = fmaxf(a, b);
    }
    return r;
}
// Lane holding value j of a K-value butterfly result.
template <int K>
__device__ inline int many_src(int j) { return j << (6 - ilog2c(K)); }

__device__ inline uint64_t lanemask_lt() {
    int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace pk
