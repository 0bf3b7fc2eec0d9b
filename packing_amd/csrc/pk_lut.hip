// Pose-space lookup table: handle, on-device builder (exact clipping of the polygon at the
// origin against the polygon at every grid pose) and the table-based overlap cost.
#include <algorithm>
#include <cstring>
#include <string>

#include "pk_eval.h"
#include "packing.h"

namespace pk {

struct Lut {
    LutDev d;
    float4* corners = nullptr;   // owned corner copy of the caller's table
    int device = 0;
};

// Corner copy of a table: entry (ix, iy, it), it = 0..nt, holds the values at (ix, iy),
// (ix, iy+1), (ix+1, iy), (ix+1, iy+1) and theta node min(it, nt - 1), upper indices
// clamped to the last node, so the eight corners of a cell are two adjacent entries, also
// at the upper theta end.
__global__ __launch_bounds__(256) void k_lut_corners(const float* __restrict__ vals, int nx, int ny, int nt,
                                                     float4* __restrict__ quad, uint32_t* __restrict__ bound) {
    const int ntc = nt + 1;
    uint32_t m[4] = {0u, 0u, 0u, 0u};   // max |v|, |step| along x, y, theta (float bits)
    const long long total = (long long)nx * ny * ntc;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int it = min((int)(q % ntc), nt - 1);
        const long long r = q / ntc;
        const int iy = (int)(r % ny), ix = (int)(r / ny);
        const int ix1 = min(ix + 1, nx - 1), iy1 = min(iy + 1, ny - 1);
        float4 e;
        e.x = vals[((long long)ix * ny + iy) * nt + it];
        e.y = vals[((long long)ix * ny + iy1) * nt + it];
        e.z = vals[((long long)ix1 * ny + iy) * nt + it];
        e.w = vals[((long long)ix1 * ny + iy1) * nt + it];
        quad[q] = e;
        if (q % ntc < nt) {   // each node (ix, iy, it) once: |v| and its steps to the next node
            const float vn = it + 1 < nt ? vals[((long long)ix * ny + iy) * nt + it + 1] : e.x;
            m[0] = max(m[0], __float_as_uint(fabsf(e.x)));
            m[1] = max(m[1], __float_as_uint(fabsf(e.z - e.x)));
            m[2] = max(m[2], __float_as_uint(fabsf(e.y - e.x)));
            m[3] = max(m[3], __float_as_uint(fabsf(vn - e.x)));
        }
    }
    // non-negative floats order as their bit patterns, and NaN above infinity
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t v = m[k];
        for (int o = WAVE / 2; o > 0; o /= 2) v = max(v, (uint32_t)__shfl_xor((int)v, o));
        if (threadIdx.x % WAVE == 0 && v) atomicMax(&bound[k], v);
    }
}

LutDev make_lut_dev(const void* h) { return reinterpret_cast<const Lut*>(h)->d; }

// One thread per grid node: area of (polygon at 0, theta 0) ∩ (polygon at (x, y, theta)).
// Equals what core/pack_cost.py:333-366 tabulates through the exact kernel.
__global__ __launch_bounds__(256) void k_lut_build(const GeomT<float>* __restrict__ geom,
                                                   const float* __restrict__ X, const float* __restrict__ Y,
                                                   const float* __restrict__ Th, int nx, int ny, int nt,
                                                   float* __restrict__ out) {
    __shared__ GeomT<float> G;
    load_geom(G, geom);
    __syncthreads();
    const long long total = (long long)nx * ny * nt;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int it = (int)(q % nt);
        const long long r = q / nt;
        const int iy = (int)(r % ny), ix = (int)(r / ny);
        const float dx = X[ix], dy = Y[iy];
        float s, c;
        sincosf(Th[it], &s, &c);
        float acc = 0.f;
        if (dx * dx + dy * dy <= G.cutoff_sq) {
            for (int pi = 0; pi < G.P; ++pi) {
                float ax[MAXV], ay[MAXV];
#pragma unroll
                for (int v = 0; v < MAXV; ++v) {
                    ax[v] = G.px[pi * MAXV + v];
                    ay[v] = G.py[pi * MAXV + v];
                }
                for (int pj = 0; pj < G.P; ++pj) {
                    const float cbx = c * G.ccx[pj] - s * G.ccy[pj] + dx - G.ccx[pi];
                    const float cby = s * G.ccx[pj] + c * G.ccy[pj] + dy - G.ccy[pi];
                    const float rr = G.cr[pi] + G.cr[pj];
                    if (cbx * cbx + cby * cby > rr * rr) continue;
                    float bx[MAXV], by[MAXV];
#pragma unroll
                    for (int v = 0; v < MAXV; ++v) {
                        bx[v] = c * G.px[pj * MAXV + v] - s * G.py[pj * MAXV + v] + dx;
                        by[v] = s * G.px[pj * MAXV + v] + c * G.py[pj * MAXV + v] + dy;
                    }
                    float gi[3], gj[3];
                    acc += area_pair<float, false>(ax, ay, bx, by, dx, dy, gi, gj);
                }
            }
        }
        out[q] = acc;
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_lut_overlap(const float* __restrict__ xyt,
                                                       const int* __restrict__ inst_off, int n_trees,
                                                       int n_ens, const GeomT<float>* __restrict__ geom,
                                                       LutDev lut, float* __restrict__ out_cost,
                                                       float* __restrict__ out_grads,
                                                       unsigned long long* __restrict__ stats) {
    extern __shared__ __align__(16) char smem[];
    __shared__ GeomT<float> G;
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
    __syncthreads();
    EvalLds<float> L = carve_eval_lds<float>(smem, N, 0, false, BLOCK / WAVE);
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        L.p[i].x = xyt[3 * (off + i) + 0];
        L.p[i].y = xyt[3 * (off + i) + 1];
        L.t[i] = xyt[3 * (off + i) + 2];
    }
    __syncthreads();
    stage_instance<float, BLOCK, false>(G, L);
    Evaluator<float, MODE_LUT, true, false, BLOCK> ev(G, L);
    ev.cax = ev.cay = ev.cbx = ev.cby = 0.f;
    ev.npd = 0;
    ev.only_self = 0;
    ev.lut = lut;
    ev.run();
    __syncthreads();
    if (threadIdx.x == 0) out_cost[e] = ev.cost_total();
    if (stats && threadIdx.x % WAVE == 0) {
        atomicAdd(&stats[0], (unsigned long long)ev.n_tp);
        atomicAdd(&stats[1], (unsigned long long)ev.n_pp);
    }
    if (out_grads)
        for (int q = threadIdx.x; q < 3 * N; q += BLOCK) out_grads[3 * off + q] = Acc<float>::get(L.g[q]);
}

}  // namespace pk

using namespace pk;

extern "C" {

int pk_lut_create(const float* vals, int nx, int ny, int nt, double x_min, double x_max, double y_min,
                  double y_max, double t_min, double t_max, int quad, int mirror, pk_lut_t* out) {
    PK_REQUIRE(out != nullptr && vals != nullptr, "pk_lut_create: NULL argument");
    PK_REQUIRE(nx >= 2 && ny >= 2 && nt >= 2, "pk_lut_create: every axis needs >= 2 nodes");
    PK_REQUIRE(x_max > x_min && y_max > y_min && t_max > t_min, "pk_lut_create: empty range");
    const long long nodes = (long long)nx * ny * (nt + 1);   // corner-copy entries
    PK_REQUIRE(nodes < (1ll << 32), "pk_lut_create: at most 2^32 - 1 entries nx ny (nt + 1)");
    PK_REQUIRE((long long)nx * ny < (1ll << 24) && nt + 1 < (1 << 24),
               "pk_lut_create: nx ny and nt + 1 must stay below 2^24");
    // a mirrored lookup indexes (-x, -theta): the theta range must be symmetric about 0, or a
    // relative angle inside [t_min, t_max] could land below the first theta node
    PK_REQUIRE(!mirror || (float)t_min == -(float)t_max,
               "pk_lut_create: a mirrored table needs a theta range symmetric about 0");
    // the copy lives on the table's device; the caller's writes to vals are complete
    // (whatever stream made them) before it is taken
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, vals) != hipSuccess) {
        (void)hipGetLastError();   // not a device allocation: clear the runtime's error
        set_error("pk_lut_create: vals must be device memory");
        return PK_EINVAL;
    }
    PK_REQUIRE(attr.type == hipMemoryTypeDevice || attr.isManaged,
               "pk_lut_create: vals must be device memory");
    int prev = 0;
    PK_CHECK_HIP(hipGetDevice(&prev));
    PK_CHECK_HIP(hipSetDevice(attr.device));
    float4* cq = nullptr;
    uint32_t* db = nullptr;
    uint32_t hb[4] = {0u, 0u, 0u, 0u};
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMalloc((void**)&cq, sizeof(float4) * (size_t)nodes);
    if (e == hipSuccess) e = hipMalloc((void**)&db, sizeof(hb));
    if (e == hipSuccess) e = hipMemcpy(db, hb, sizeof(hb), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int blocks = (int)std::min<long long>((nodes + 255) / 256, 65536);
        hipLaunchKernelGGL(k_lut_corners, dim3(blocks), dim3(256), 0, 0, vals, nx, ny, nt, cq, db);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(hb, db, sizeof(hb), hipMemcpyDeviceToHost);
    }
    if (db) (void)hipFree(db);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        if (cq) (void)hipFree(cq);
        set_error(std::string("pk_lut_create: ") + hipGetErrorString(e));
        return PK_EHIP;
    }
    // Pair terms are summed in fixed point with |term| < FIX_TERM_MAX.  A lookup's value
    // is a convex combination of nodes and its slopes of node steps, so with |v| <= B and
    // steps <= Dx, Dy, Dt the terms are at most: value B (B^2 squared), centre gradient
    // Dx sx + Dy sy, the first tree's angle gradient Dx sx Ym + Dy sy Xm + Dt st (Xm, Ym
    // the largest |x|, |y| in range), the second's Dt st; squaring multiplies slopes by 2B.
    {
        float f[4];
        std::memcpy(f, hb, sizeof(f));
        const double B = f[0], sx = (nx - 1) / (x_max - x_min), sy = (ny - 1) / (y_max - y_min),
                     st = (nt - 1) / (t_max - t_min);
        const double k = quad ? 2.0 * B : 1.0;
        const double gx = k * f[1] * sx, gy = k * f[2] * sy, gt = k * f[3] * st;
        const double Xm = std::max(std::fabs(x_min), std::fabs(x_max)), Ym = std::max(std::fabs(y_min), std::fabs(y_max));
        const double worst = std::max(std::max(quad ? B * B : B, gx + gy), std::max(gx * Ym + gy * Xm + gt, gt));
        if (!(worst <= (double)FIX_TERM_MAX)) {   // also NaN or infinite values
            (void)hipFree(cq);
            set_error("pk_lut_create: table values or slopes too large: a lookup term could reach " +
                      std::to_string(worst) + " (the fixed-point pair sums take |term| <= 2000)");
            return PK_ERANGE;
        }
    }
    Lut* L = new Lut();
    L->corners = cq;
    L->device = attr.device;
    L->d.vals = vals;
    L->d.corners = cq;
    L->d.nx = nx;
    L->d.ny = ny;
    L->d.nt = nt;
    L->d.x_min = (float)x_min;
    L->d.x_max = (float)x_max;
    L->d.x_lo = mirror ? (float)-x_max : (float)x_min;
    L->d.y_min = (float)y_min;
    L->d.y_max = (float)y_max;
    L->d.t_min = (float)t_min;
    L->d.t_max = (float)t_max;
    L->d.sx = (float)((nx - 1) / (x_max - x_min));
    L->d.sy = (float)((ny - 1) / (y_max - y_min));
    L->d.st = (float)((nt - 1) / (t_max - t_min));
    L->d.quad = quad ? 1 : 0;
    L->d.mirror = mirror ? 1 : 0;
    *out = reinterpret_cast<pk_lut_t>(L);
    return PK_OK;
}

int pk_lut_destroy(pk_lut_t h) {
    Lut* L = reinterpret_cast<Lut*>(h);
    if (L && L->corners) {
        int prev = 0;
        if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(L->device) == hipSuccess) {
            (void)hipFree(L->corners);
            (void)hipSetDevice(prev);
        }
    }
    delete L;
    return PK_OK;
}

int pk_lut_build(pk_geometry_t h, const float* X, int nx, const float* Y, int ny, const float* T, int nt,
                 float* out_vals, void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g && X && Y && T && out_vals, "pk_lut_build: NULL argument");
    PK_REQUIRE(nx >= 1 && ny >= 1 && nt >= 1, "pk_lut_build: empty grid");
    const long long total = (long long)nx * ny * nt;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_lut_build, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), g->df, X, Y, T, nx, ny, nt, out_vals);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

int pk_lut_overlap(pk_geometry_t h, pk_lut_t lut, const float* xyt, const int32_t* inst_off, int n_ens,
                   int n_trees, float* out_cost, float* out_grads, int block,
                   unsigned long long* stats, void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g && lut, "pk_lut_overlap: NULL geometry or table");
    PK_REQUIRE(n_ens >= 0 && n_trees >= 0 && n_trees <= 4096, "pk_lut_overlap: bad size");
    if (n_ens == 0) return PK_OK;
    PK_REQUIRE(xyt && out_cost, "pk_lut_overlap: xyt and out_cost are required");
    const LutDev d = reinterpret_cast<const Lut*>(lut)->d;
    const size_t lds = eval_lds_bytes<float>(n_trees, 0, false, MAX_WAVES);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int blk = block ? block : 256;
#define PK_LO(B)                                                                                   \
    {                                                                                              \
        auto kern = k_lut_overlap<B>;                                                              \
        if (lds > 60000)                                                                           \
            PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                  \
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        hipLaunchKernelGGL(kern, dim3(n_ens), dim3(B), lds, st, xyt, inst_off, n_trees, n_ens,      \
                           g->df, d, out_cost, out_grads, stats);                                  \
        PK_CHECK_HIP(hipGetLastError());                                                           \
        return PK_OK;                                                                              \
    }
    switch (blk) {
        case 64: PK_LO(64)
        case 128: PK_LO(128)
        case 256: PK_LO(256)
        default:
            set_error("pk_lut_overlap: block must be 64, 128 or 256");
            return PK_EINVAL;
    }
#undef PK_LO
}

}  // extern "C"
