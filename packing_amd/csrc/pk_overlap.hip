// Geometry handles, exact overlap / separation cost kernels and the boundary term,
// with their C-ABI entry points (declared in include/packing.h).
#include <math.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "pk_eval.h"
#include "packing.h"

namespace pk {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------
// The float cost+gradient kernels are held to 128 VGPRs (four waves per SIMD): with four
// 64-lane waves per instance that keeps four instances on every CU, so a batch of 1024
// instances is resident at once instead of running a second, one-wave-per-SIMD round.
// They fit without spilling; the other variants keep the compiler's own choice.
template <typename T, int MODE, bool GRAD, bool CRYSTAL, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu((sizeof(T) == 4 && GRAD && !CRYSTAL) ? 4 : 1))) void k_overlap(const T* __restrict__ xyt,
                                                   const int* __restrict__ inst_off, int n_trees,
                                                   int n_ens, const GeomT<T>* __restrict__ geom,
                                                   const T* __restrict__ crystal, int npd,
                                                   int only_self, T* __restrict__ out_cost,
                                                   T* __restrict__ out_grads,
                                                   unsigned long long* __restrict__ stats) {
    extern __shared__ __align__(16) char smem[];
    __shared__ GeomT<T> G;
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
    EvalLds<T> L = carve_eval_lds<T>(smem, N, G.NU, CRYSTAL, BLOCK / WAVE);
    L.P = G.P;
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        L.p[i].x = xyt[3 * (off + i) + 0];
        L.p[i].y = xyt[3 * (off + i) + 1];
        L.t[i] = xyt[3 * (off + i) + 2];
    }
    __syncthreads();
    stage_instance<T, BLOCK, true, !CRYSTAL>(G, L);
    Evaluator<T, MODE, GRAD, CRYSTAL, BLOCK> ev(G, L);
    if (CRYSTAL) {
        ev.cax = crystal[4 * e + 0];
        ev.cay = crystal[4 * e + 1];
        ev.cbx = crystal[4 * e + 2];
        ev.cby = crystal[4 * e + 3];
        ev.npd = npd;
        ev.only_self = only_self;
    } else {
        ev.cax = ev.cay = ev.cbx = ev.cby = T(0);
        ev.npd = 0;
        ev.only_self = 0;
    }
    ev.run();
    __syncthreads();
    if (threadIdx.x == 0) out_cost[e] = ev.cost_total();
    if (stats && threadIdx.x % WAVE == 0) {
        atomicAdd(&stats[0], (unsigned long long)ev.n_tp);
        atomicAdd(&stats[1], (unsigned long long)ev.n_pp);
    }
    if (GRAD)
        for (int q = threadIdx.x; q < 3 * N; q += BLOCK) out_grads[3 * off + q] = Acc<T>::get(L.g[q]);
}

// Boundary term per instance; cost and d/dh summed in fixed point (order independent).
template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_boundary(const T* __restrict__ xyt,
                                                    const int* __restrict__ inst_off, int n_trees,
                                                    int n_ens, const GeomT<T>* __restrict__ geom,
                                                    const T* __restrict__ h, T* __restrict__ out_cost,
                                                    T* __restrict__ out_grads,
                                                    T* __restrict__ out_grad_h) {
    __shared__ GeomT<T> G;
    __shared__ typename Acc<T>::type acc[4];
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
    if (threadIdx.x < 4) acc[threadIdx.x] = Acc<T>::zero();
    __syncthreads();
    const T half = h[3 * e] * T(0.5), hx = h[3 * e + 1], hy = h[3 * e + 2];
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        const T x = xyt[3 * (off + i)], y = xyt[3 * (off + i) + 1], t = xyt[3 * (off + i) + 2];
        T sn, cs;
        sincos_t(t, &sn, &cs);
        T g[3], gh[3];
        const T c = boundary_tree<T>(G, x, y, cs, sn, half, hx, hy, g, gh);
        Acc<T>::add_wide(&acc[0], c);
        Acc<T>::add_wide(&acc[1], gh[0]);
        Acc<T>::add_wide(&acc[2], gh[1]);
        Acc<T>::add_wide(&acc[3], gh[2]);
        if (out_grads) {
            out_grads[3 * (off + i)] = g[0];
            out_grads[3 * (off + i) + 1] = g[1];
            out_grads[3 * (off + i) + 2] = g[2];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out_cost[e] = Acc<T>::get(acc[0]);
        if (out_grad_h) {
            out_grad_h[3 * e] = Acc<T>::get(acc[1]);
            out_grad_h[3 * e + 1] = Acc<T>::get(acc[2]);
            out_grad_h[3 * e + 2] = Acc<T>::get(acc[3]);
        }
    }
}

// ---------------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------------
static int max_dyn_lds(int device) {
    static int cached[64];
    static std::once_flag once[64];
    if (device < 0 || device >= 64) return 65536;
    std::call_once(once[device], [&] {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
            v = 65536;
        cached[device] = v;
    });
    return cached[device];
}

// Launch a dynamic-LDS kernel, raising its LDS limit when needed.
template <typename K>
static int launch_lds(K kernel, int grid, int block, size_t lds, hipStream_t st,
                      size_t static_lds) {
    int dev = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    const int lim = max_dyn_lds(dev);
    if (lds + static_lds > (size_t)lim) {
        set_error("instance needs " + std::to_string(lds + static_lds) + " B of LDS, device allows " +
                  std::to_string(lim));
        return PK_ERANGE;
    }
    if (lds > 65536 - static_lds) {
        PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    return 0;
}

template <typename T, int MODE, bool GRAD, bool CRYSTAL, int BLOCK>
static int run_overlap(const Geometry* g, const void* xyt, const int* inst_off, int n_ens,
                       int n_trees, const void* crystal, int npd, int only_self, void* cost,
                       void* grads, unsigned long long* stats, hipStream_t st) {
    auto kern = k_overlap<T, MODE, GRAD, CRYSTAL, BLOCK>;
    const int NV = (sizeof(T) == 4) ? g->hf.NU : g->hd.NU;
    const size_t lds = eval_lds_bytes<T>(n_trees, NV, CRYSTAL, BLOCK / WAVE);
    int rc = launch_lds(kern, n_ens, BLOCK, lds, st, sizeof(GeomT<T>));
    if (rc) return rc;
    const GeomT<T>* dg = reinterpret_cast<const GeomT<T>*>(sizeof(T) == 4 ? (const void*)g->df
                                                                          : (const void*)g->dd);
    hipLaunchKernelGGL(kern, dim3(n_ens), dim3(BLOCK), lds, st, (const T*)xyt, inst_off, n_trees,
                       n_ens, dg, (const T*)crystal, npd, only_self, (T*)cost, (T*)grads, stats);
    PK_CHECK_HIP(hipGetLastError());
    return 0;
}

template <typename T, int BLOCK>
static int dispatch_overlap(const Geometry* g, const void* xyt, const int* inst_off, int n_ens,
                            int n_trees, int mode, const void* crystal, int npd, int only_self,
                            void* cost, void* grads, unsigned long long* stats, hipStream_t st) {
    const bool gr = grads != nullptr, cr = crystal != nullptr;
#define PK_OV(M, G_, C_)                                                                    \
    if (mode == M && gr == G_ && cr == C_)                                                  \
        return run_overlap<T, M, G_, C_, BLOCK>(g, xyt, inst_off, n_ens, n_trees, crystal, npd, \
                                                only_self, cost, grads, stats, st);
    PK_OV(MODE_AREA, true, false)
    PK_OV(MODE_AREA, false, false)
    PK_OV(MODE_SEP, true, false)
    PK_OV(MODE_SEP, false, false)
    PK_OV(MODE_AREA, true, true)
    PK_OV(MODE_AREA, false, true)
    PK_OV(MODE_SEP, true, true)
    PK_OV(MODE_SEP, false, true)
#undef PK_OV
    set_error("bad overlap mode");
    return PK_EINVAL;
}

}  // namespace pk

using namespace pk;

// Indices (ascending) of the outline vertices that are vertex 0 or lie on the boundary of
// the outline's convex hull (within 1e-9 of a hull edge's line); returns their count.
static int outline_hull_subset(const double* xy, int n, int* keep) {
    int idx[MAXO];
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx, idx + n, [&](int a, int b) {
        return xy[2 * a] < xy[2 * b] || (xy[2 * a] == xy[2 * b] && xy[2 * a + 1] < xy[2 * b + 1]);
    });
    auto cross = [&](int o, int a, int b) {
        return (xy[2 * a] - xy[2 * o]) * (xy[2 * b + 1] - xy[2 * o + 1]) -
               (xy[2 * a + 1] - xy[2 * o + 1]) * (xy[2 * b] - xy[2 * o]);
    };
    int hull[2 * MAXO], k = 0;   // Andrew's monotone chain, counter-clockwise, strict turns
    for (int i = 0; i < n; ++i) {
        while (k >= 2 && cross(hull[k - 2], hull[k - 1], idx[i]) <= 0) --k;
        hull[k++] = idx[i];
    }
    for (int i = n - 2, lo = k + 1; i >= 0; --i) {
        while (k >= lo && cross(hull[k - 2], hull[k - 1], idx[i]) <= 0) --k;
        hull[k++] = idx[i];
    }
    const int h = k - 1;   // the last point repeats the first
    int c = 0;
    for (int v = 0; v < n; ++v) {
        bool on = v == 0 || h < 3;
        for (int e = 0; e < h && !on; ++e) {
            const int a = hull[e], b = hull[e + 1];
            const double ex = xy[2 * b] - xy[2 * a], ey = xy[2 * b + 1] - xy[2 * a + 1];
            const double dist = cross(a, b, v) / std::sqrt(ex * ex + ey * ey);
            on = dist <= 1e-9;   // inside the hull every edge line has the point on its left
        }
        if (on) keep[c++] = v;
    }
    return c;
}

// ---------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------
extern "C" {

const char* pk_last_error(void) { return g_last_error.c_str(); }

int pk_version(void) { return PK_ABI_VERSION; }

int pk_geometry_create(const double* piece_xy, const int32_t* piece_nverts, int n_pieces,
                       int max_verts, const double* outline_xy, int n_outline, double radius,
                       pk_geometry_t* out) {
    PK_REQUIRE(out != nullptr, "pk_geometry_create: out is NULL");
    PK_REQUIRE(n_pieces >= 1 && n_pieces <= MAXP, "pk_geometry_create: 1 <= n_pieces <= 16");
    PK_REQUIRE(max_verts >= 3 && max_verts <= MAXV, "pk_geometry_create: 3 <= max_verts <= 4");
    PK_REQUIRE(n_outline >= 3 && n_outline <= MAXO, "pk_geometry_create: 3 <= n_outline <= 64");
    PK_REQUIRE(radius > 0, "pk_geometry_create: radius must be positive");
    // Pair terms are summed in fixed point with |term| < FIX_TERM_MAX.  With every vertex
    // within R of the origin a piece pair contributes at most pi R^2 of area, 2 pi R of
    // area per unit of translation and 2 pi R^2 per radian, or in separation mode a
    // squared depth (2R)^2 with derivatives up to 20 R^2: R <= 8 keeps all below 1300.
    for (int p = 0; p < n_pieces; ++p)
        for (int v = 0; v < piece_nverts[p] && v < max_verts; ++v) {
            const double x = piece_xy[(p * max_verts + v) * 2], y = piece_xy[(p * max_verts + v) * 2 + 1];
            PK_REQUIRE(std::sqrt(x * x + y * y) <= 8.0,
                       "pk_geometry_create: piece vertices must lie within 8 of the tree origin");
        }
    Geometry* g = new Geometry();
    std::memset(&g->hf, 0, sizeof(g->hf));
    std::memset(&g->hd, 0, sizeof(g->hd));
    g->hd.P = g->hf.P = n_pieces;
    g->hd.n_outline = g->hf.n_outline = n_outline;
    g->hd.radius = radius;
    g->hf.radius = (float)radius;
    g->hd.cutoff_sq = (2.0 * radius) * (2.0 * radius);
    g->hf.cutoff_sq = (float)g->hd.cutoff_sq;
    for (int p = 0; p < n_pieces; ++p) {
        const int n = piece_nverts[p];
        if (n < 3 || n > max_verts) {
            delete g;
            set_error("pk_geometry_create: piece vertex count out of range");
            return PK_EINVAL;
        }
        double mnx = 1e300, mxx = -1e300, mny = 1e300, mxy = -1e300;
        for (int v = 0; v < MAXV; ++v) {
            const int s = v < n ? v : n - 1;
            const double x = piece_xy[(p * max_verts + s) * 2], y = piece_xy[(p * max_verts + s) * 2 + 1];
            g->hd.px[p * MAXV + v] = x;
            g->hd.py[p * MAXV + v] = y;
            g->hf.px[p * MAXV + v] = (float)x;
            g->hf.py[p * MAXV + v] = (float)y;
            mnx = std::min(mnx, x); mxx = std::max(mxx, x);
            mny = std::min(mny, y); mxy = std::max(mxy, y);
        }
        const double cx = 0.5 * (mnx + mxx), cy = 0.5 * (mny + mxy);
        double r = 0;
        for (int v = 0; v < n; ++v) {
            const double dx = piece_xy[(p * max_verts + v) * 2] - cx;
            const double dy = piece_xy[(p * max_verts + v) * 2 + 1] - cy;
            r = std::max(r, std::sqrt(dx * dx + dy * dy));
        }
        r = r * (1.0 + 1e-6) + 1e-6;   // conservative against float rounding
        g->hd.ccx[p] = cx; g->hd.ccy[p] = cy; g->hd.cr[p] = r;
        g->hf.ccx[p] = (float)cx; g->hf.ccy[p] = (float)cy; g->hf.cr[p] = (float)r;
    }
    // distinct piece vertices (exact equality of the double coordinates)
    int nu = 0;
    for (int p = 0; p < n_pieces; ++p) {
        uint32_t word = 0;
        for (int v = 0; v < MAXV; ++v) {
            const double x = g->hd.px[p * MAXV + v], y = g->hd.py[p * MAXV + v];
            int u = 0;
            while (u < nu && !(g->hd.ux[u] == x && g->hd.uy[u] == y)) ++u;
            if (u == nu) {
                g->hd.ux[nu] = x;
                g->hd.uy[nu] = y;
                g->hf.ux[nu] = (float)x;
                g->hf.uy[nu] = (float)y;
                ++nu;
            }
            word |= (uint32_t)u << (8 * v);
        }
        g->hd.pv[p] = g->hf.pv[p] = word;
    }
    g->hd.NU = g->hf.NU = nu;
    // The boundary term is the largest squared distance of an outline vertex outside the
    // square: a convex function of the vertex, whose maximum over the outline is attained
    // on the boundary of its convex hull, and (for a positive maximum) never first at a
    // vertex strictly inside the hull.  The device keeps vertex 0 (the arg-max when no
    // vertex is outside) and every vertex on the hull's boundary, collinear ones included,
    // in their original order, so the strict first-wins arg-max selects the same vertex:
    // 5 of the tree's 15 vertices.  Padded to MAXO by repeating the last kept vertex
    // (boundary_tree reads whole blocks of 8; a repeat never wins its strict arg-max).
    int keep[MAXO];
    const int n_keep = outline_hull_subset(outline_xy, n_outline, keep);
    g->hd.n_outline = g->hf.n_outline = n_keep;
    for (int v = 0; v < MAXO; ++v) {
        const int w = keep[v < n_keep ? v : n_keep - 1];
        g->hd.ox[v] = outline_xy[2 * w];
        g->hd.oy[v] = outline_xy[2 * w + 1];
        g->hf.ox[v] = (float)outline_xy[2 * w];
        g->hf.oy[v] = (float)outline_xy[2 * w + 1];
    }
    hipError_t e1 = hipGetDevice(&g->device);
    hipError_t e2 = hipMalloc((void**)&g->df, sizeof(GeomT<float>));
    hipError_t e3 = hipMalloc((void**)&g->dd, sizeof(GeomT<double>));
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        set_error("pk_geometry_create: device allocation failed");
        if (g->df) (void)hipFree(g->df);
        if (g->dd) (void)hipFree(g->dd);
        delete g;
        return PK_ENOMEM;
    }
    PK_CHECK_HIP(hipMemcpy(g->df, &g->hf, sizeof(GeomT<float>), hipMemcpyHostToDevice));
    PK_CHECK_HIP(hipMemcpy(g->dd, &g->hd, sizeof(GeomT<double>), hipMemcpyHostToDevice));
    *out = reinterpret_cast<pk_geometry_t>(g);
    return PK_OK;
}

int pk_geometry_destroy(pk_geometry_t h) {
    Geometry* g = reinterpret_cast<Geometry*>(h);
    if (!g) return PK_OK;
    if (g->df) (void)hipFree(g->df);
    if (g->dd) (void)hipFree(g->dd);
    delete g;
    return PK_OK;
}

int pk_overlap(pk_geometry_t h, int dtype, const void* xyt, const int32_t* inst_off, int n_ens,
               int n_trees, int mode, const void* crystal, int n_periodic, int only_self,
               void* out_cost, void* out_grads, int block, unsigned long long* stats,
               void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g != nullptr, "pk_overlap: geometry is NULL");
    PK_REQUIRE(dtype == PK_F32 || dtype == PK_F64, "pk_overlap: dtype must be PK_F32 or PK_F64");
    PK_REQUIRE(mode == PK_MODE_AREA || mode == PK_MODE_SEP, "pk_overlap: bad mode");
    PK_REQUIRE(n_ens >= 0 && n_trees >= 0, "pk_overlap: negative size");
    PK_REQUIRE(n_trees <= 4096, "pk_overlap: at most 4096 trees per instance");
    PK_REQUIRE(n_periodic >= 0 && n_periodic <= 7, "pk_overlap: 0 <= n_periodic <= 7");
    if (n_ens == 0) return PK_OK;
    PK_REQUIRE(xyt && out_cost, "pk_overlap: xyt and out_cost are required");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (dtype == PK_F32) {
        // library's choice: four waves per instance, each walking a quarter of the trees
        // (measured on MI355X at 30, 50, 100, 150 and 200 trees: fastest or tied with 128
        // everywhere, 1.4x faster than one wave at 100 trees); results are the same bits
        // for every block
        if (block == 0) block = 256;
        switch (block) {
            case 64:
                return dispatch_overlap<float, 64>(g, xyt, inst_off, n_ens, n_trees, mode, crystal,
                                                   n_periodic, only_self, out_cost, out_grads, stats, st);
            case 128:
                return dispatch_overlap<float, 128>(g, xyt, inst_off, n_ens, n_trees, mode, crystal,
                                                    n_periodic, only_self, out_cost, out_grads, stats, st);
            case 256:
                return dispatch_overlap<float, 256>(g, xyt, inst_off, n_ens, n_trees, mode, crystal,
                                                    n_periodic, only_self, out_cost, out_grads, stats, st);
            default:
                set_error("pk_overlap: block must be 64, 128 or 256");
                return PK_EINVAL;
        }
    }
    return dispatch_overlap<double, 256>(g, xyt, inst_off, n_ens, n_trees, mode, crystal, n_periodic,
                                         only_self, out_cost, out_grads, stats, st);
}

int pk_boundary(pk_geometry_t h, int dtype, const void* xyt, const int32_t* inst_off, int n_ens,
                int n_trees, const void* hbox, void* out_cost, void* out_grads, void* out_grad_h,
                void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g != nullptr, "pk_boundary: geometry is NULL");
    PK_REQUIRE(dtype == PK_F32 || dtype == PK_F64, "pk_boundary: bad dtype");
    PK_REQUIRE(n_ens >= 0 && n_trees >= 0, "pk_boundary: negative size");
    if (n_ens == 0) return PK_OK;
    PK_REQUIRE(xyt && hbox && out_cost, "pk_boundary: xyt, h and out_cost are required");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (dtype == PK_F32) {
        hipLaunchKernelGGL((k_boundary<float, 256>), dim3(n_ens), dim3(256), 0, st, (const float*)xyt,
                           inst_off, n_trees, n_ens, g->df, (const float*)hbox, (float*)out_cost,
                           (float*)out_grads, (float*)out_grad_h);
    } else {
        hipLaunchKernelGGL((k_boundary<double, 256>), dim3(n_ens), dim3(256), 0, st,
                           (const double*)xyt, inst_off, n_trees, n_ens, g->dd, (const double*)hbox,
                           (double*)out_cost, (double*)out_grads, (double*)out_grad_h);
    }
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

int pk_device_info(int* max_lds_per_block, int* n_cu, int* clock_khz) {
    int dev = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    if (max_lds_per_block) *max_lds_per_block = max_dyn_lds(dev);
    if (n_cu) PK_CHECK_HIP(hipDeviceGetAttribute(n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    if (clock_khz) PK_CHECK_HIP(hipDeviceGetAttribute(clock_khz, hipDeviceAttributeClockRate, dev));
    return PK_OK;
}

}  // extern "C"
