// Separation table builder: the signed separation distance between the polygon at the
// origin (theta 0) and a copy at (x, y, theta), in double, for every node of a grid --
// the content core/pack_cost.py:1170-1200 tabulates through pack_minkowski
// (core/pack_minkowski.py:33-266), which computes it with shapely on the CPU.
//
// The zone where the copy's centre makes the two overlap is the union over piece pairs
// (i, j) of the convex hull of {a - R(theta) b : a a vertex of piece i, b of piece j}.
// The value at a point is its distance to the union's boundary, positive inside (the
// penetration depth) and negative outside (the clearance).  Per theta node:
//   k_sep_hulls     one thread per piece pair: the hull (monotone chain, in double);
//   k_sep_segments  one thread per hull edge: the parts of the edge not inside another
//                   hull (open intervals cut from the edge) -- the union's boundary;
//   k_sep_eval      one thread per (x, y) node: inside any hull?  distance to the
//                   nearest boundary part.
#include <algorithm>
#include <string>

#include "pk_common.h"
#include "packing.h"

namespace pk {

constexpr int SEP_HV = 16;       // hull vertices (sum of two 4-gons: at most 8; duplicates pruned)
constexpr int SEP_SMAX = 4096;   // boundary parts per theta node
constexpr int SEP_CUTS = 48;     // disjoint covered stretches kept per edge

struct SepSeg {
    double ax, ay, bx, by;
};

__device__ inline double sep_cross(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}

// One thread per (theta node in the chunk, piece pair).
__global__ __launch_bounds__(256) void k_sep_hulls(const GeomT<double>* __restrict__ G, const float* __restrict__ T,
                                                   int t0, int nt_chunk, double2* __restrict__ hulls,
                                                   int* __restrict__ hull_n) {
    const int PP = G->P * G->P;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nt_chunk * PP) return;
    const int tl = q / PP, k = q % PP;
    const int i = k / G->P, j = k % G->P;
    double s, c;
    sincos((double)T[t0 + tl], &s, &c);
    double px[MAXV * MAXV], py[MAXV * MAXV];
    int n = 0;
#pragma unroll
    for (int a = 0; a < MAXV; ++a)
#pragma unroll
        for (int b = 0; b < MAXV; ++b) {
            const double bx = c * G->px[j * MAXV + b] - s * G->py[j * MAXV + b];
            const double by = s * G->px[j * MAXV + b] + c * G->py[j * MAXV + b];
            px[n] = G->px[i * MAXV + a] - bx;
            py[n] = G->py[i * MAXV + a] - by;
            ++n;
        }
    // insertion sort by (x, y), then the two monotone chains
    for (int u = 1; u < n; ++u) {
        const double vx = px[u], vy = py[u];
        int w = u - 1;
        while (w >= 0 && (px[w] > vx || (px[w] == vx && py[w] > vy))) {
            px[w + 1] = px[w];
            py[w + 1] = py[w];
            --w;
        }
        px[w + 1] = vx;
        py[w + 1] = vy;
    }
    // drop repeated points (padded piece vertices, or a_1 - b_1 = a_2 - b_2): a repeat of
    // the rightmost point would open the upper chain with a zero-length edge, whose
    // inside test (cross > 0) no point passes
    {
        int w = 1;
        for (int u = 1; u < n; ++u)
            if (!(fabs(px[u] - px[w - 1]) <= 1e-13 && fabs(py[u] - py[w - 1]) <= 1e-13)) {
                px[w] = px[u];
                py[w] = py[u];
                ++w;
            }
        n = w;
    }
    double hx[2 * MAXV * MAXV + 1], hy[2 * MAXV * MAXV + 1];
    int m = 0;
    for (int u = 0; u < n; ++u) {
        while (m >= 2 && sep_cross(hx[m - 2], hy[m - 2], hx[m - 1], hy[m - 1], px[u], py[u]) <= 0.0) --m;
        hx[m] = px[u];
        hy[m] = py[u];
        ++m;
    }
    for (int u = n - 2, lo = m + 1; u >= 0; --u) {
        while (m >= lo && sep_cross(hx[m - 2], hy[m - 2], hx[m - 1], hy[m - 1], px[u], py[u]) <= 0.0) --m;
        hx[m] = px[u];
        hy[m] = py[u];
        ++m;
    }
    m = m > 1 ? m - 1 : m;
    if (m > SEP_HV) m = 0;   // cannot happen for 4-gons (at most 8 vertices); guards the buffer
    double2* H = hulls + ((size_t)tl * PP + k) * SEP_HV;
    for (int u = 0; u < m; ++u) H[u] = make_double2(hx[u], hy[u]);
    hull_n[tl * PP + k] = m;
}

// Open interval (lo, hi) of u + s (v - u) strictly inside hull l.
__device__ inline bool sep_seg_in_hull(const double2* __restrict__ H, int m, double ux, double uy, double vx,
                                       double vy, double& lo, double& hi) {
    if (m < 3) return false;
    double L = -1e300, U = 1e300;
    for (int a = 0; a < m; ++a) {
        const double2 w0 = H[a], w1 = H[a + 1 < m ? a + 1 : 0];
        const double f0 = sep_cross(w0.x, w0.y, w1.x, w1.y, ux, uy);
        const double f1 = sep_cross(w0.x, w0.y, w1.x, w1.y, vx, vy);
        const double df = f1 - f0;
        if (df == 0.0) {
            if (!(f0 > 0.0)) return false;
        } else {
            const double r = -f0 / df;
            if (df > 0.0) L = fmax(L, r);
            else U = fmin(U, r);
        }
    }
    if (!(L < U)) return false;
    lo = L;
    hi = U;
    return true;
}

// One thread per (theta node in the chunk, hull, edge): uncovered parts of the edge.
__global__ __launch_bounds__(256) void k_sep_segments(int PP, int nt_chunk, const double2* __restrict__ hulls,
                                                      const int* __restrict__ hull_n, SepSeg* __restrict__ segs,
                                                      int* __restrict__ seg_n, int* __restrict__ overflow) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nt_chunk * PP * SEP_HV) return;
    const int tl = q / (PP * SEP_HV), r = q % (PP * SEP_HV);
    const int k = r / SEP_HV, a = r % SEP_HV;
    const double2* Hs = hulls + (size_t)tl * PP * SEP_HV;
    const int* hn = hull_n + tl * PP;
    const int m = hn[k];
    if (m < 3 || a >= m) return;
    const double2 u = Hs[k * SEP_HV + a], v = Hs[k * SEP_HV + (a + 1 < m ? a + 1 : 0)];
    // disjoint covered stretches, sorted by start (open intervals merged when they overlap)
    double clo[SEP_CUTS], chi[SEP_CUTS];
    int nc = 0;
    for (int l = 0; l < PP; ++l) {
        double lo, hi;
        if (l == k || !sep_seg_in_hull(Hs + l * SEP_HV, hn[l], u.x, u.y, v.x, v.y, lo, hi)) continue;
        if (!(hi > 0.0 && lo < 1.0)) continue;
        // merge with every stretch it overlaps, then insert in order
        int w = 0;
        for (int c = 0; c < nc; ++c) {
            if (clo[c] < hi && lo < chi[c]) {
                lo = fmin(lo, clo[c]);
                hi = fmax(hi, chi[c]);
            } else {
                clo[w] = clo[c];
                chi[w] = chi[c];
                ++w;
            }
        }
        nc = w;
        if (nc == SEP_CUTS) {
            atomicExch(overflow, 1);
            return;
        }
        int p = nc;
        while (p > 0 && clo[p - 1] > lo) {
            clo[p] = clo[p - 1];
            chi[p] = chi[p - 1];
            --p;
        }
        clo[p] = lo;
        chi[p] = hi;
        ++nc;
    }
    SepSeg* S = segs + (size_t)tl * SEP_SMAX;
    double s0 = 0.0;
    for (int c = 0; c <= nc && s0 <= 1.0; ++c) {
        const double s1 = c < nc ? fmin(clo[c], 1.0) : 1.0;
        if (c < nc ? s1 > s0 : s1 >= s0) {
            const int at = atomicAdd(&seg_n[tl], 1);
            if (at < SEP_SMAX) {
                SepSeg g;
                g.ax = u.x + s0 * (v.x - u.x);
                g.ay = u.y + s0 * (v.y - u.y);
                g.bx = u.x + s1 * (v.x - u.x);
                g.by = u.y + s1 * (v.y - u.y);
                S[at] = g;
            } else {
                atomicExch(overflow, 1);
            }
        }
        if (c < nc && chi[c] > s0) s0 = chi[c];
    }
}

// One thread per (x, y) node of one theta node (blockIdx.y).
__global__ __launch_bounds__(256) void k_sep_eval(const float* __restrict__ X, int nx, const float* __restrict__ Y,
                                                  int ny, int nt, int t0, int PP, const double2* __restrict__ hulls,
                                                  const int* __restrict__ hull_n, const SepSeg* __restrict__ segs,
                                                  const int* __restrict__ seg_n, float* __restrict__ out) {
    const int tl = blockIdx.y;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)nx * ny) return;
    const int ix = (int)(q / ny), iy = (int)(q % ny);
    const double x = X[ix], y = Y[iy];
    const double2* Hs = hulls + (size_t)tl * PP * SEP_HV;
    const int* hn = hull_n + tl * PP;
    bool inside = false;
    for (int k = 0; k < PP && !inside; ++k) {
        const int m = hn[k];
        if (m < 3) continue;
        bool in = true;
        for (int a = 0; a < m && in; ++a) {
            const double2 w0 = Hs[k * SEP_HV + a], w1 = Hs[k * SEP_HV + (a + 1 < m ? a + 1 : 0)];
            in = sep_cross(w0.x, w0.y, w1.x, w1.y, x, y) > 0.0;
        }
        inside = in;
    }
    const SepSeg* S = segs + (size_t)tl * SEP_SMAX;
    const int ns = min(seg_n[tl], SEP_SMAX);
    double best = 1e300;
    for (int e = 0; e < ns; ++e) {
        const SepSeg g = S[e];
        const double ex = g.bx - g.ax, ey = g.by - g.ay, L = ex * ex + ey * ey;
        double t = L > 0.0 ? ((x - g.ax) * ex + (y - g.ay) * ey) / L : 0.0;
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const double dx = x - (g.ax + t * ex), dy = y - (g.ay + t * ey);
        best = fmin(best, dx * dx + dy * dy);
    }
    const double d = sqrt(best);
    out[((long long)ix * ny + iy) * nt + (t0 + tl)] = (float)(inside ? d : -d);
}

__global__ void k_sep_zero(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace pk

using namespace pk;

extern "C" {

int pk_sep_build(pk_geometry_t h, const float* X, int nx, const float* Y, int ny, const float* T, int nt,
                 float* out_vals, void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g && X && Y && T && out_vals, "pk_sep_build: NULL argument");
    PK_REQUIRE(nx >= 1 && ny >= 1 && nt >= 1, "pk_sep_build: empty grid");
    PK_REQUIRE((long long)nx * ny < (1ll << 31), "pk_sep_build: too many (x, y) nodes");
    const int PP = g->hd.P * g->hd.P;
    const int chunk = std::min(nt, 128);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    double2* hulls = nullptr;
    int* ints = nullptr;
    SepSeg* segs = nullptr;
    const size_t hb = sizeof(double2) * (size_t)chunk * PP * SEP_HV;
    const size_t nints = (size_t)chunk * PP + chunk + 1;
    hipError_t e = hipMalloc((void**)&hulls, hb);
    if (e == hipSuccess) e = hipMalloc((void**)&ints, sizeof(int) * nints);
    if (e == hipSuccess) e = hipMalloc((void**)&segs, sizeof(SepSeg) * (size_t)chunk * SEP_SMAX);
    int overflow = 0;
    for (int t0 = 0; e == hipSuccess && t0 < nt; t0 += chunk) {
        const int nc = std::min(chunk, nt - t0);
        int* hull_n = ints;
        int* seg_n = ints + (size_t)chunk * PP;
        int* ovf = seg_n + chunk;
        hipLaunchKernelGGL(k_sep_zero, dim3(1), dim3(256), 0, st, seg_n, chunk + 1);
        hipLaunchKernelGGL(k_sep_hulls, dim3((nc * PP + 255) / 256), dim3(256), 0, st, g->dd, T, t0, nc, hulls,
                           hull_n);
        hipLaunchKernelGGL(k_sep_segments, dim3((nc * PP * SEP_HV + 255) / 256), dim3(256), 0, st, PP, nc, hulls,
                           hull_n, segs, seg_n, ovf);
        const long long nxy = (long long)nx * ny;
        hipLaunchKernelGGL(k_sep_eval, dim3((unsigned)((nxy + 255) / 256), nc), dim3(256), 0, st, X, nx, Y, ny, nt,
                           t0, PP, hulls, hull_n, segs, seg_n, out_vals);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&overflow, ovf, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (overflow) break;
    }
    if (segs) (void)hipFree(segs);
    if (ints) (void)hipFree(ints);
    if (hulls) (void)hipFree(hulls);
    if (e != hipSuccess) {
        set_error(std::string("pk_sep_build: ") + hipGetErrorString(e));
        return PK_EHIP;
    }
    if (overflow) {
        set_error("pk_sep_build: the zone's boundary has more parts than the builder holds");
        return PK_ERANGE;
    }
    return PK_OK;
}

}  // extern "C"
