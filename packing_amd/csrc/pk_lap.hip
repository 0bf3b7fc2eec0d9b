// Batched linear assignment (Hungarian / Kuhn-Munkres with potentials) and the fused
// diversity shortcut.
//
// One wave solves one N x N problem (N <= 256): lane l owns columns l, l+64, l+128,
// l+192 (minimum slack, column potential, used flag in registers); row potentials,
// matches and parent pointers live in LDS, and so does the cost matrix when it fits.
// Each Dijkstra step is the reference's scan (core/lap_batch.py:97-164) done in
// parallel: reduced cost c[row][j] - u[row] - v[j] with strict-< slack update, then the
// minimum slack over unused columns with the lowest column index on ties (exactly what
// the sequential strict-< scan picks), then the same dual update.  Float matrices use
// the reference's float32 operation order; int32 matrices use int64 potentials (exact
// for every int32 input).
#include <cmath>
#include <cstring>
#include <string>

#include "pk_common.h"
#include "packing.h"

namespace pk {

constexpr int LAP_COLS = 4;              // columns per lane
constexpr int LAP_MAXN = WAVE * LAP_COLS;

template <typename C>
struct LapT;
template <>
struct LapT<float> {
    using P = float;   // potentials / slack
    __device__ static float big() { return 1e30f; }
};
template <>
struct LapT<int> {
    using P = long long;
    __device__ static long long big() { return (long long)1 << 62; }
};

// argmin of (value, index) over the wave; lowest index among equal values
template <typename P>
__device__ inline void wave_argmin(P& v, int& idx) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const P ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov < v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
}

// Hungarian solve of the N x N matrix M (LDS or global) by the calling wave, with the LDS
// arrays u, row_match, col_match, par of N entries each; row_match holds the result.
template <typename C>
__device__ inline void lap_solve(const C* __restrict__ M, int N, typename LapT<C>::P* u, int* row_match,
                                 int* col_match, int* par) {
    using P = typename LapT<C>::P;
    const int lane = threadIdx.x;
    for (int q = lane; q < N; q += WAVE) {
        u[q] = P(0);
        row_match[q] = -1;
        col_match[q] = -1;
    }
    P v[LAP_COLS], minv[LAP_COLS];
    unsigned used = 0;
#pragma unroll
    for (int k = 0; k < LAP_COLS; ++k) v[k] = P(0);
    __syncthreads();

    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int k = 0; k < LAP_COLS; ++k) minv[k] = LapT<C>::big();
        used = 0;
        int cur_row = i, cur_col = -1;
        for (;;) {
            const P ur = u[cur_row];
            const C* row = M + (size_t)cur_row * N;
            P best = LapT<C>::big();
            int bcol = 0x7fffffff;
#pragma unroll
            for (int k = 0; k < LAP_COLS; ++k) {
                const int j = lane + k * WAVE;
                if (j < N && !(used & (1u << k))) {
                    const P red = (P)row[j] - ur - v[k];
                    if (red < minv[k]) {
                        minv[k] = red;
                        par[j] = cur_row;
                    }
                    if (minv[k] < best) {
                        best = minv[k];
                        bcol = j;
                    }
                }
            }
            wave_argmin<P>(best, bcol);
            // dual update (each matched row is touched by exactly one used column)
#pragma unroll
            for (int k = 0; k < LAP_COLS; ++k) {
                const int j = lane + k * WAVE;
                if (j < N) {
                    if (used & (1u << k)) {
                        u[col_match[j]] += best;
                        v[k] -= best;
                    } else {
                        minv[k] -= best;
                    }
                }
            }
            __syncthreads();
            if (lane == 0) u[i] += best;
            if ((bcol & (WAVE - 1)) == lane) used |= 1u << (bcol / WAVE);
            cur_col = bcol;
            __syncthreads();
            cur_row = col_match[cur_col];
            if (cur_row < 0) break;
        }
        // flip the alternating path
        if (lane == 0) {
            while (cur_col >= 0) {
                const int pr = par[cur_col];
                const int pc = row_match[pr];
                row_match[pr] = cur_col;
                col_match[cur_col] = pr;
                cur_col = pc;
            }
        }
        __syncthreads();
    }
}

template <typename C, bool STAGE>
__global__ __launch_bounds__(WAVE) void k_hungarian(const C* __restrict__ cost, int B, int N,
                                                    int* __restrict__ row_match_out,
                                                    void* __restrict__ total_out,
                                                    float* __restrict__ total_f32_out) {
    using P = typename LapT<C>::P;
    extern __shared__ __align__(16) char smem[];
    const int b = blockIdx.x;
    if (b >= B) return;
    const int lane = threadIdx.x;
    const C* __restrict__ Cg = cost + (size_t)b * N * N;
    P* u = reinterpret_cast<P*>(smem);                      // [N] row potentials
    int* row_match = reinterpret_cast<int*>(u + N);         // [N]
    int* col_match = row_match + N;                         // [N]
    int* par = col_match + N;                               // [N] parent row of each column
    C* Cs = reinterpret_cast<C*>(smem + align16(sizeof(P) * N + sizeof(int) * 3 * N));
    const C* __restrict__ M = STAGE ? Cs : Cg;
    if (STAGE) {
        const int nn = N * N;
        for (int q = lane; q < nn; q += WAVE) Cs[q] = Cg[q];
    }
    lap_solve<C>(M, N, u, row_match, col_match, par);
    for (int q = lane; q < N; q += WAVE) row_match_out[(size_t)b * N + q] = row_match[q];
    if (lane == 0) {
        // core/lap_batch.py:187-195: sequential float32 sum in row order
        float tf = 0.f;
        long long ti = 0;
        for (int r = 0; r < N; ++r) {
            const int c = row_match[r];
            if (c >= 0 && c < N) {
                tf += (float)M[(size_t)r * N + c];
                ti += (long long)M[(size_t)r * N + c];
            }
        }
        if (total_f32_out) total_f32_out[b] = tf;
        if (total_out) {
            if (sizeof(C) == 4 && sizeof(P) == 8) reinterpret_cast<long long*>(total_out)[b] = ti;
            else reinterpret_cast<float*>(total_out)[b] = tf;
        }
    }
}

// ---------------------------------------------------------------------------------
// Genetic-diversity matrix in one kernel (core/kaggle_support.py:2014-2133): one wave per
// (symmetry transform, population individual, reference individual) builds the N x N
// tree-to-tree distance matrix straight into LDS -- reference trees mirrored / rotated as
// the transform says, distance sqrt(dx^2 + dy^2 + dtheta^2) with both angles wrapped to
// [-pi, pi) -- and solves its assignment there (Hungarian, or the min-cost-row / -column
// sums), writing one total per problem.  The reference materialises the 8 cost tensors
// (8 x N_pop x N x N_ref x N) and then batches them through the LAP.  Entries are
// computed in double from the float32 poses, with numpy's remainder and no fused
// multiply-adds, and rounded to float32 once, the type the assignment runs in.
// ---------------------------------------------------------------------------------
struct DivTransforms {
    double cos_a[8], sin_a[8], rot[8];
    int mirror[8], has_rot[8];
};

__device__ inline double np_remainder(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

__device__ inline void div_ref_tree(const float* R, const DivTransforms& T, int tr, double& x, double& y,
                                    double& th) {
#pragma clang fp contract(off)
    const double PI = 3.141592653589793;
    x = (double)R[0];
    y = (double)R[1];
    th = (double)R[2];
    if (T.mirror[tr]) {
        y = -y;
        th = PI - th;
    }
    if (T.has_rot[tr]) {
        const double nx = x * T.cos_a[tr] - y * T.sin_a[tr];
        const double ny = x * T.sin_a[tr] + y * T.cos_a[tr];
        x = nx;
        y = ny;
        th = th + T.rot[tr];
    }
    th = np_remainder(th + PI, 2.0 * PI) - PI;
}

__device__ inline float div_entry(double px, double py, double pt, double rx, double ry, double rt) {
#pragma clang fp contract(off)
    const double PI = 3.141592653589793;
    const double dx = px - rx, dy = py - ry;
    const double dt = np_remainder((pt - rt) + PI, 2.0 * PI) - PI;
    return (float)sqrt(dx * dx + dy * dy + dt * dt);
}

// algorithm: 0 Hungarian, 1 min cost per row, 2 min cost per column
__global__ __launch_bounds__(WAVE) void k_diversity_lap(const float* __restrict__ pop, const float* __restrict__ ref,
                                                        int n_pop, int n_ref, int N, int n_tr, DivTransforms T,
                                                        int algorithm, float* __restrict__ out) {
    using P = float;
    extern __shared__ __align__(16) char smem[];
    const int b = blockIdx.x;
    if (b >= n_tr * n_pop * n_ref) return;
    const int tr = b / (n_pop * n_ref), ip = (b / n_ref) % n_pop, ir = b % n_ref;
    const int lane = threadIdx.x;
    P* u = reinterpret_cast<P*>(smem);
    int* row_match = reinterpret_cast<int*>(u + N);
    int* col_match = row_match + N;
    int* par = col_match + N;
    float* Cs = reinterpret_cast<float*>(smem + align16(sizeof(P) * N + sizeof(int) * 3 * N));
    const float* Pp = pop + (size_t)ip * N * 3;
    const float* Rr = ref + (size_t)ir * N * 3;
    double rx[LAP_COLS], ry[LAP_COLS], rt[LAP_COLS];
#pragma unroll
    for (int k = 0; k < LAP_COLS; ++k) {
        const int j = lane + k * WAVE;
        rx[k] = ry[k] = rt[k] = 0.0;
        if (j < N) div_ref_tree(Rr + 3 * j, T, tr, rx[k], ry[k], rt[k]);
    }
    for (int i = 0; i < N; ++i) {
        const double px = (double)Pp[3 * i], py = (double)Pp[3 * i + 1], pt = (double)Pp[3 * i + 2];
#pragma unroll
        for (int k = 0; k < LAP_COLS; ++k) {
            const int j = lane + k * WAVE;
            if (j < N) Cs[i * N + j] = div_entry(px, py, pt, rx[k], ry[k], rt[k]);
        }
    }
    __syncthreads();
    if (algorithm == 0) {
        lap_solve<float>(Cs, N, u, row_match, col_match, par);
        if (lane == 0) {
            float tf = 0.f;   // sequential float32 sum in row order (core/lap_batch.py:187-195)
            for (int r = 0; r < N; ++r) tf += Cs[r * N + row_match[r]];
            out[b] = tf;
        }
    } else if (algorithm == 1) {
        for (int i = lane; i < N; i += WAVE) {
            float m = Cs[i * N];
            for (int j = 1; j < N; ++j) m = fminf(m, Cs[i * N + j]);
            u[i] = m;
        }
        __syncthreads();
        if (lane == 0) {
            float tf = 0.f;
            for (int i = 0; i < N; ++i) tf += u[i];
            out[b] = tf;
        }
    } else {
        for (int j = lane; j < N; j += WAVE) {
            float m = Cs[j];
            for (int i = 1; i < N; ++i) m = fminf(m, Cs[i * N + j]);
            u[j] = m;
        }
        __syncthreads();
        if (lane == 0) {
            float tf = 0.f;
            for (int j = 0; j < N; ++j) tf += u[j];
            out[b] = tf;
        }
    }
}

// One wave per (pop, ref) pair: transform the reference (optional mirror y -> -y,
// theta -> pi - theta, then rotation), then sum over population trees of the distance to
// the nearest transformed reference tree in (x, y, wrapped theta)
// (core/lap_batch.py:216-331).
__device__ inline float wrap_mod(float t) {
    const float PI = 3.14159265358979323846f, TWO_PI = 2.0f * PI;
    t = fmodf(t + PI, TWO_PI);
    if (t < 0) t += TWO_PI;
    return t - PI;
}

__global__ __launch_bounds__(WAVE) void k_diversity(const float* __restrict__ pop, const float* __restrict__ ref,
                                                    int n_pop, int n_ref, int n_trees, float cos_a, float sin_a,
                                                    int mirror, float* __restrict__ out) {
    extern __shared__ __align__(16) char smem[];
    float* rx = reinterpret_cast<float*>(smem);
    float* ry = rx + n_trees;
    float* rt = ry + n_trees;
    float* part = rt + n_trees;
    const int b = blockIdx.x;
    if (b >= n_pop * n_ref) return;
    const int ip = b / n_ref, ir = b % n_ref;
    const float* P = pop + (size_t)ip * n_trees * 3;
    const float* R = ref + (size_t)ir * n_trees * 3;
    const float PI = 3.14159265358979323846f;
    const float rot = atan2f(sin_a, cos_a);
    for (int t = threadIdx.x; t < n_trees; t += WAVE) {
        float x = R[3 * t], y = R[3 * t + 1], th = R[3 * t + 2];
        if (mirror) {
            y = -y;
            th = PI - th;
        }
        rx[t] = x * cos_a - y * sin_a;
        ry[t] = x * sin_a + y * cos_a;
        rt[t] = wrap_mod(th + rot);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n_trees; t += WAVE) {
        const float px = P[3 * t], py = P[3 * t + 1], pt = P[3 * t + 2];
        float best = 1e30f;
        for (int j = 0; j < n_trees; ++j) {
            const float dx = px - rx[j], dy = py - ry[j];
            const float dt = wrap_mod(pt - rt[j]);
            const float sq = dx * dx + dy * dy + dt * dt;
            if (sq < best) best = sq;
        }
        part[t] = sqrtf(best);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float tot = 0.f;
        for (int t = 0; t < n_trees; ++t) tot += part[t];
        out[b] = tot;
    }
}

}  // namespace pk

using namespace pk;

extern "C" {

int pk_lap_hungarian(int dtype, const void* cost, int B, int N, int32_t* row_match, void* total,
                     float* total_f32, void* stream) {
    PK_REQUIRE(dtype == PK_F32 || dtype == PK_I32, "pk_lap_hungarian: dtype must be PK_F32 or PK_I32");
    PK_REQUIRE(B >= 0 && N >= 0 && N <= LAP_MAXN, "pk_lap_hungarian: 0 <= N <= 256");
    if (B == 0 || N == 0) return PK_OK;
    PK_REQUIRE(cost && row_match, "pk_lap_hungarian: NULL argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int dev = 0, lim = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    PK_CHECK_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t psz = dtype == PK_I32 ? 8 : 4;
    const size_t head = align16(psz * N + sizeof(int) * 3 * N);
    const size_t full = head + (size_t)4 * N * N;
    const bool stage = full <= (size_t)lim;
    const size_t lds = stage ? full : head;
#define PK_H(C_, S_)                                                                                    \
    {                                                                                                   \
        auto kern = k_hungarian<C_, S_>;                                                                \
        if (lds > 60000)                                                                                \
            PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                       \
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));    \
        hipLaunchKernelGGL(kern, dim3(B), dim3(WAVE), lds, st, (const C_*)cost, B, N, row_match, total, \
                           total_f32);                                                                  \
    }
    if (dtype == PK_F32) {
        if (stage) PK_H(float, true) else PK_H(float, false)
    } else {
        if (stage) PK_H(int, true) else PK_H(int, false)
    }
#undef PK_H
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

int pk_diversity_matrix(const float* pop, const float* ref, int n_pop, int n_ref, int n_trees, int transform,
                        int algorithm, float* out, void* stream) {
    PK_REQUIRE(n_pop >= 0 && n_ref >= 0 && n_trees >= 1 && n_trees <= 200,
               "pk_diversity_matrix: 1 <= n_trees <= 200, sizes >= 0");
    PK_REQUIRE(algorithm >= 0 && algorithm <= 2, "pk_diversity_matrix: algorithm 0, 1 or 2");
    if (n_pop == 0 || n_ref == 0) return PK_OK;
    PK_REQUIRE(pop && ref && out, "pk_diversity_matrix: NULL argument");
    // the transforms of core/kaggle_support.py:2062-2076: rotations by 0, pi/2, pi, 3pi/2,
    // then the same after mirroring; cos and sin as numpy gives them for those doubles
    DivTransforms T;
    const double PI = 3.141592653589793;
    const double rots[4] = {0.0, PI / 2, PI, 3 * PI / 2};
    const int n_tr = transform ? 8 : 1;
    for (int t = 0; t < 8; ++t) {
        const double r = rots[t % 4];
        T.rot[t] = r;
        T.cos_a[t] = std::cos(r);
        T.sin_a[t] = std::sin(r);
        T.has_rot[t] = r != 0.0;
        T.mirror[t] = t >= 4;
    }
    int dev = 0, lim = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    PK_CHECK_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t lds = align16(sizeof(float) * n_trees + sizeof(int) * 3 * n_trees) + sizeof(float) * n_trees * n_trees;
    PK_REQUIRE(lds <= (size_t)lim, "pk_diversity_matrix: matrix does not fit in LDS");
    if (lds > 60000)
        PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_diversity_lap),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long nb = (long long)n_tr * n_pop * n_ref;
    PK_REQUIRE(nb < (1ll << 31), "pk_diversity_matrix: too many problems");
    hipLaunchKernelGGL(k_diversity_lap, dim3((unsigned)nb), dim3(WAVE), lds, reinterpret_cast<hipStream_t>(stream),
                       pop, ref, n_pop, n_ref, n_trees, n_tr, T, algorithm, out);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

int pk_diversity(const float* pop, const float* ref, int n_pop, int n_ref, int n_trees, float cos_a,
                 float sin_a, int mirror, float* out, void* stream) {
    PK_REQUIRE(n_pop >= 0 && n_ref >= 0 && n_trees >= 0, "pk_diversity: negative size");
    if (n_pop == 0 || n_ref == 0) return PK_OK;
    PK_REQUIRE(pop && ref && out, "pk_diversity: NULL argument");
    const size_t lds = sizeof(float) * 4 * (size_t)(n_trees > 0 ? n_trees : 1);
    PK_REQUIRE(lds <= 60000, "pk_diversity: too many trees");
    hipLaunchKernelGGL(k_diversity, dim3(n_pop * n_ref), dim3(WAVE), lds,
                       reinterpret_cast<hipStream_t>(stream), pop, ref, n_pop, n_ref, n_trees, cos_a,
                       sin_a, mirror, out);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

}  // extern "C"
