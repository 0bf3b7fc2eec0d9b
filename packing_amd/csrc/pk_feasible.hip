// FP64 zero-overlap feasibility of final packings.
//
// Two trees overlap when some pair of their convex pieces has intersecting interiors;
// touching is allowed (core/pack_metric.py:130-141, "intersects and not touches").  For
// convex pieces that means: no edge normal of either piece separates the projections
// with max1 <= min2 or max2 <= min1.  Poses, rotations and projections are in double.
#include <cstring>
#include <string>

#include "pk_eval.h"
#include "packing.h"

namespace pk {

__device__ inline bool interiors_meet(const double (&ax)[MAXV], const double (&ay)[MAXV],
                                      const double (&bx)[MAXV], const double (&by)[MAXV]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int k = 0; k < MAXV; ++k) {
            const int k1 = (k + 1) % MAXV;
            const double ex = pass == 0 ? ax[k1] - ax[k] : bx[k1] - bx[k];
            const double ey = pass == 0 ? ay[k1] - ay[k] : by[k1] - by[k];
            if (ex == 0.0 && ey == 0.0) continue;
            const double nx = -ey, ny = ex;
            double mn1 = 1e300, mx1 = -1e300, mn2 = 1e300, mx2 = -1e300;
#pragma unroll
            for (int v = 0; v < MAXV; ++v) {
                const double p = ax[v] * nx + ay[v] * ny;
                mn1 = fmin(mn1, p);
                mx1 = fmax(mx1, p);
                const double q = bx[v] * nx + by[v] * ny;
                mn2 = fmin(mn2, q);
                mx2 = fmax(mx2, q);
            }
            if (mx1 <= mn2 || mx2 <= mn1) return false;
        }
    }
    return true;
}

template <typename T>
__global__ __launch_bounds__(256) void k_feasible(const T* __restrict__ xyt, const int* __restrict__ inst_off,
                                                  int n_trees, int n_ens, const GeomT<double>* __restrict__ geom,
                                                  int* __restrict__ out_pairs) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) char smem[];
    __shared__ GeomT<double> G;
    __shared__ int cnt;
    const int e = blockIdx.x;
    if (e >= n_ens) return;
    int off, N;
    if (inst_off) {
        off = inst_off[e];
        N = inst_off[e + 1] - off;
    } else {
        off = e * n_trees;
        N = n_trees;
    }
    load_geom(G, geom);
    double* px = reinterpret_cast<double*>(smem);
    double* py = px + N;
    double* pc = py + N;
    double* ps = pc + N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        px[i] = (double)xyt[3 * (off + i)];
        py[i] = (double)xyt[3 * (off + i) + 1];
        double s, c;
        sincos((double)xyt[3 * (off + i) + 2], &s, &c);
        pc[i] = c;
        ps[i] = s;
    }
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const long long npair = (long long)N * (N - 1) / 2;
    for (long long k = threadIdx.x; k < npair; k += blockDim.x) {
        // row-major upper triangle: i < j
        int i = (int)((2.0 * N - 1.0 - sqrt((2.0 * N - 1.0) * (2.0 * N - 1.0) - 8.0 * (double)k)) * 0.5);
        if (i < 0) i = 0;
        while ((long long)(i + 1) * (2 * N - i - 2) / 2 <= k) ++i;
        while (i > 0 && (long long)i * (2 * N - i - 1) / 2 > k) --i;
        const long long base = (long long)i * (2 * N - i - 1) / 2;
        const int j = i + 1 + (int)(k - base);
        const double dx = px[j] - px[i], dy = py[j] - py[i];
        if (dx * dx + dy * dy > G.cutoff_sq) continue;
        bool hit = false;
        for (int pi = 0; pi < G.P && !hit; ++pi) {
            double ax[MAXV], ay[MAXV];
#pragma unroll
            for (int v = 0; v < MAXV; ++v) {
                const double lx = G.px[pi * MAXV + v], ly = G.py[pi * MAXV + v];
                ax[v] = pc[i] * lx - ps[i] * ly + px[i];
                ay[v] = ps[i] * lx + pc[i] * ly + py[i];
            }
            for (int pj = 0; pj < G.P && !hit; ++pj) {
                double bx[MAXV], by[MAXV];
#pragma unroll
                for (int v = 0; v < MAXV; ++v) {
                    const double lx = G.px[pj * MAXV + v], ly = G.py[pj * MAXV + v];
                    bx[v] = pc[j] * lx - ps[j] * ly + px[j];
                    by[v] = ps[j] * lx + pc[j] * ly + py[j];
                }
                hit = interiors_meet(ax, ay, bx, by);
            }
        }
        if (hit) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) out_pairs[e] = cnt;
}

}  // namespace pk

using namespace pk;

extern "C" int pk_feasible(pk_geometry_t h, int dtype, const void* xyt, const int32_t* inst_off, int n_ens,
                           int n_trees, int32_t* out_pairs, void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g != nullptr, "pk_feasible: geometry is NULL");
    PK_REQUIRE(dtype == PK_F32 || dtype == PK_F64, "pk_feasible: bad dtype");
    PK_REQUIRE(n_ens >= 0 && n_trees >= 0 && n_trees <= 4096, "pk_feasible: bad size");
    if (n_ens == 0) return PK_OK;
    PK_REQUIRE(xyt && out_pairs, "pk_feasible: NULL argument");
    const size_t lds = sizeof(double) * 4 * (size_t)n_trees;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (dtype == PK_F32) {
        if (lds > 60000)
            PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_feasible<float>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_feasible<float>, dim3(n_ens), dim3(256), lds, st, (const float*)xyt, inst_off,
                           n_trees, n_ens, g->dd, out_pairs);
    } else {
        if (lds > 60000)
            PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_feasible<double>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_feasible<double>, dim3(n_ens), dim3(256), lds, st, (const double*)xyt,
                           inst_off, n_trees, n_ens, g->dd, out_pairs);
    }
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}
