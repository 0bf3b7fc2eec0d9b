// Per-instance evaluation of the exact overlap / separation cost and its gradient.
//
// One workgroup evaluates one packing instance.  Stages (all in LDS):
//   1. poses -> sincos -> piece vertices rotated about their tree centre, and a hashed
//      cell grid (cell side >= the pair cutoff);
//   2. per wave, candidate tree pairs from each tree's own and forward-neighbour cells,
//      tested against the bounding-circle cutoff (2R)^2 and compacted with ballots into
//      a wave-local queue;
//   3. the queue drains 64 entries at a time: table lookups, or piece pairs of the tree
//      pairs tested against piece bounding circles into a second wave-local queue;
//   4. one lane per piece pair: exact clipped area (or SAT depth) and the gradient for
//      BOTH trees, accumulated as int64 fixed point (order independent), so the result
//      does not depend on the enumeration order or the workgroup size.
// Periodic images (crystal axes) use block-wide lists over all pairs and lattice cells.
//
// Semantics follow core/pack_cuda.py:142-546: self pairs skipped, cutoff
// |dp|^2 <= (2R)^2, total = sum over unordered pairs (the reference halves its ordered
// sum), separation cost = sum of sep^2 with sep from the SAT of
// core/pack_cuda_primitives.py:544-828, periodic images as core/pack_cuda.py:332-364
// (self images give cost only).  The area itself is computed differently from the
// reference (Sutherland-Hodgman with reverse mode there; Green's theorem over clipped
// edges with shape-derivative gradients here): identical value, no dynamic arrays.
#pragma once
#include "pk_common.h"

namespace pk {

enum { MODE_AREA = 0, MODE_SEP = 1, MODE_LUT = 2 };

// Pose-space lookup table in device memory: vals[ix][iy][it], theta fastest.
struct LutDev {
    const float* vals;     // vals[ix][iy][it], theta fastest (the caller's table)
    const float4* corners; // corners[ix][iy][it], it <= nt: values at (ix, iy), (ix, iy+1),
                           // (ix+1, iy), (ix+1, iy+1) and theta node min(it, nt - 1), upper
                           // indices clamped (library's copy, nt + 1 entries per (ix, iy))
    int nx, ny, nt;
    float x_lo, x_min, x_max, y_min, y_max, t_min, t_max;   // x_lo = range check low end
    float sx, sy, st;                                       // (n - 1) / (max - min)
    int quad;      // per-pair max(0, v)^2
    int mirror;    // table covers x >= 0; x < 0 looked up as (-x, -theta)
};

// One table lookup split in two so that the loads of several lookups can be in flight
// together: lut_fetch computes the cell and issues the loads, lut_combine interpolates.
// The eight corners of a cell are the four (x, y) corners of the corner copy at theta nodes
// it and it + 1: two adjacent 16-byte entries, one cache line in most cases (the caller's
// table would need four lines).  The copy repeats the last theta node once more, so a
// lookup at the upper theta end (it = nt - 1) reads node nt - 1 for both theta corners,
// the reference's clamped upper index, with no selection.
struct LutTap {
    float4 e0, e1;
    float fx, fy, ft;
    bool flip, in;
};

__device__ inline void lut_fetch(const LutDev& T_, float x, float y, float th, LutTap& t) {
    t.in = !(x < T_.x_lo || x > T_.x_max || y < T_.y_min || y > T_.y_max || th < T_.t_min || th > T_.t_max);
    t.flip = T_.mirror && x < 0.f;
    const float xt = t.flip ? -x : x, tt = t.flip ? -th : th;
    const float gx = (xt - T_.x_min) * T_.sx, gy = (y - T_.y_min) * T_.sy, gt = (tt - T_.t_min) * T_.st;
    int ix0 = (int)floorf(gx), iy0 = (int)floorf(gy), it0 = (int)floorf(gt);
    if (!t.in) ix0 = iy0 = it0 = 0;   // any cell of the table; the values are not used
    const int ix1 = min(ix0 + 1, T_.nx - 1), iy1 = min(iy0 + 1, T_.ny - 1);
    t.fx = gx - ix0;
    t.fy = gy - iy0;
    t.ft = gt - it0;
    (void)ix1;
    (void)iy1;   // the corner copy holds the clamped upper corners
    // 32-bit entry index from 24-bit products (pk_lut_create keeps nx ny and nt + 1 below
    // 2^24 and the copy below 2^32 entries, e.g. 900^3 nodes), scaled to a 64-bit address
    // by one shift-add
    const uint32_t idx = __umul24(__umul24((uint32_t)ix0, (uint32_t)T_.ny) + (uint32_t)iy0,
                                  (uint32_t)T_.nt + 1u) + (uint32_t)it0;
    const float4* __restrict__ p = T_.corners + idx;
    t.e0 = p[0];
    t.e1 = p[1];
}

// Trilinear value and analytic gradient w.r.t. (x, y, theta) of the table
// (core/pack_cuda_lut.py:298-415): outside the grid range 0; clamped upper corner.
__device__ inline float lut_combine(const LutDev& T_, const LutTap& t, float (&g)[3]) {
    g[0] = g[1] = g[2] = 0.f;
    if (!t.in) return 0.f;
    const float fx = t.fx, fy = t.fy, ft = t.ft;
    const float v000 = t.e0.x, v001 = t.e1.x;
    const float v010 = t.e0.y, v011 = t.e1.y;
    const float v100 = t.e0.z, v101 = t.e1.z;
    const float v110 = t.e0.w, v111 = t.e1.w;
    const float v00 = v000 * (1.f - ft) + v001 * ft, v01 = v010 * (1.f - ft) + v011 * ft;
    const float v10 = v100 * (1.f - ft) + v101 * ft, v11 = v110 * (1.f - ft) + v111 * ft;
    const float v0 = v00 * (1.f - fy) + v01 * fy, v1 = v10 * (1.f - fy) + v11 * fy;
    const float val = v0 * (1.f - fx) + v1 * fx;
    const float d0 = (v001 - v000) * (1.f - fy) + (v011 - v010) * fy;
    const float d1 = (v101 - v100) * (1.f - fy) + (v111 - v110) * fy;
    g[0] = (v1 - v0) * T_.sx;
    g[1] = ((1.f - fx) * (v01 - v00) + fx * (v11 - v10)) * T_.sy;
    g[2] = (d0 * (1.f - fx) + d1 * fx) * T_.st;
    if (t.flip) { g[0] = -g[0]; g[2] = -g[2]; }
    return val;
}

__device__ inline float lut_lookup(const LutDev& T_, float x, float y, float th, float (&g)[3]) {
    LutTap t;
    lut_fetch(T_, x, y, th, t);
    return lut_combine(T_, t, g);
}

// wrap to [-pi, pi] (core/pack_cuda_lut.py:289-293)
__device__ inline float wrap_pi(float t) {
    const float PI = 3.14159265358979323846f;
    while (t > PI) t -= 2.f * PI;
    while (t < -PI) t += 2.f * PI;
    return t;
}

// Copy the polygon tables into LDS (all threads of the block).
template <typename T>
__device__ inline void load_geom(GeomT<T>& G, const GeomT<T>* __restrict__ src) {
    const int* s = reinterpret_cast<const int*>(src);
    int* d = reinterpret_cast<int*>(&G);
    for (int q = threadIdx.x; q < (int)(sizeof(GeomT<T>) / sizeof(int)); q += blockDim.x) d[q] = s[q];
}

constexpr int TP_CAP = 1024;   // block-wide tree-pair list capacity (periodic images)
constexpr int PP_CAP = 2048;   // block-wide piece-pair list capacity (periodic images)
constexpr int QCAP = 160;      // per-wave queue capacity (entries), tree pairs and piece pairs
constexpr int XQ = 128;        // exact modes: per-wave tree-pair, piece-item and piece-pair
                               // queues and the wave's share of the final tree pairs, XQ
                               // entries each (each holds at most 127)
constexpr int LUT_UMAX = 4;    // most table lookups per lane per drain (the queue holds
                               // 2 QCAP >= 64 LUT_UMAX + 63 entries)
static_assert(2 * QCAP >= WAVE * LUT_UMAX + WAVE - 1, "table queue capacity");
constexpr int MAX_WAVES = 4;   // workgroups of at most 256 threads

// Cell grid for candidate pairs: hashed buckets of square cells whose side is at least
// the pair cutoff distance, so every pair within the cutoff lies in the same or an
// adjacent cell.
__host__ __device__ inline int grid_buckets(int N) {
    int h = 64;
    while (h < 2 * N) h <<= 1;
    return h;
}

// A tree's centre and the cosine and sine of its angle: one 16-byte LDS read (two in
// double) gives all four, a pair's centre difference two 8-byte reads.
template <typename T>
struct alignas(16) Pose {
    T x, y, c, s;
};

// LDS view for one instance.
template <typename T>
struct EvalLds {
    using A = typename Acc<T>::type;
    A* g;          // [N*3] gradient accumulators
    A* cost;       // [1]   cost accumulator
    Pose<T>* p;    // [N]   centres, cos and sin of the angles
    T* t;          // [N]   angles
    T* v;          // [N*NV*2] distinct piece vertices rotated about the tree centre,
                   // (x, y) interleaved: vertex u of tree i at v[2 (i NV + u)]
    int* head;     // [H]   grid: first tree of each bucket (-1: empty)
    int* next;     // [N]   grid: next tree in the same bucket
    uint32_t* cell;   // [N] grid: biased cell coordinates cx | cy << 16
    uint32_t* wq;  // [waves * wqs] per-wave queue space: table mode one queue of 2 QCAP
                   // entries; exact modes tree pairs, piece items, piece pairs and the
                   // shared-out final tree pairs (XQ each)
    uint32_t* tp;  // [TP_CAP] periodic images: tree-pair entries i | j << 12 | cell << 24
    uint32_t* pp;  // [PP_CAP] periodic images: piece-pair entries tp index | pi << 16 | pj << 20
    uint32_t* nb;  // [nb_cap] neighbour list (relaxation): pairs a | b << 12 within the
                   // cutoff plus a skin at the last rebuild; wcnt[8] holds its length
    int* wcnt;     // [16] per-wave ballot counts; [8] neighbour-list length; [12 + w] wave w's
                   // tree pairs left for the shared final expansion
    double* wsum;  // [MAX_WAVES] per-wave cost partials (double precision)
    int N, P, NV, H, nb_cap, wqs;
};

// Entries of queue space per wave: exact modes (NV > 0) three queues of XQ, table mode 2 QCAP.
__host__ __device__ inline int wave_queue_entries(int NV) { return NV > 0 ? 4 * XQ : 2 * QCAP; }

// Neighbour-list capacity for N trees (pairs).
__host__ __device__ inline int nb_capacity(int N) { return 6 * N + 64; }

__host__ __device__ inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bytes of dynamic LDS for one instance of N trees with NV distinct piece vertices per
// tree (NV = 0: no vertices).  `periodic` selects the block-wide lists of the
// periodic-image path instead of the grid.
template <typename T>
__host__ __device__ inline size_t eval_lds_bytes(int N, int NV, bool periodic = false, int waves = MAX_WAVES,
                                                 bool list = false) {
    size_t b = 0;
    b += sizeof(typename Acc<T>::type) * (3 * (size_t)N + 1);
    b = align_up(b, 16);
    b += (sizeof(Pose<T>) + sizeof(T)) * (size_t)N;
    b = align_up(b, 16);
    b += sizeof(T) * 2 * (size_t)N * NV;
    b = align_up(b, 16);
    if (periodic) {
        b += sizeof(uint32_t) * (TP_CAP + PP_CAP);
    } else {
        b += sizeof(int) * ((size_t)grid_buckets(N) + 2 * (size_t)N);
        b = align_up(b, 16);
        b += sizeof(uint32_t) * waves * wave_queue_entries(NV);
        if (list) b += sizeof(uint32_t) * (size_t)nb_capacity(N);
    }
    b += sizeof(int) * 16;
    b += sizeof(double) * MAX_WAVES;
    return align_up(b, 16);
}

template <typename T>
__device__ inline EvalLds<T> carve_eval_lds(char* base, int N, int NV, bool periodic = false,
                                            int waves = MAX_WAVES, bool list = false) {
    EvalLds<T> L;
    size_t o = 0;
    L.g = reinterpret_cast<typename Acc<T>::type*>(base + o);
    L.cost = L.g + 3 * N;
    o += sizeof(typename Acc<T>::type) * (3 * (size_t)N + 1);
    o = align_up(o, 16);
    L.p = reinterpret_cast<Pose<T>*>(base + o);
    L.t = reinterpret_cast<T*>(L.p + N);
    o += (sizeof(Pose<T>) + sizeof(T)) * (size_t)N;
    o = align_up(o, 16);
    L.v = reinterpret_cast<T*>(base + o);
    o += sizeof(T) * 2 * (size_t)N * NV;
    o = align_up(o, 16);
    L.H = grid_buckets(N);
    if (periodic) {
        L.tp = reinterpret_cast<uint32_t*>(base + o);
        L.pp = L.tp + TP_CAP;
        L.wcnt = reinterpret_cast<int*>(L.pp + PP_CAP);
        L.head = L.next = nullptr;
        L.cell = L.wq = L.nb = nullptr;
        L.nb_cap = 0;
    } else {
        L.head = reinterpret_cast<int*>(base + o);
        L.next = L.head + L.H;
        L.cell = reinterpret_cast<uint32_t*>(L.next + N);
        o += sizeof(int) * ((size_t)L.H + 2 * (size_t)N);
        o = align_up(o, 16);
        L.wq = reinterpret_cast<uint32_t*>(base + o);
        L.nb = L.wq + waves * wave_queue_entries(NV);
        L.nb_cap = list ? nb_capacity(N) : 0;
        L.wcnt = reinterpret_cast<int*>(L.nb + L.nb_cap);
        L.tp = L.pp = nullptr;
    }
    L.wsum = reinterpret_cast<double*>(L.wcnt + 16);
    L.N = N; L.NV = NV; L.P = 0; L.wqs = wave_queue_entries(NV);
    return L;
}

// ---------------------------------------------------------------------------------
// Geometry of one piece pair
// ---------------------------------------------------------------------------------

// Clip every edge of polygon p against the half-planes of convex polygon q (both CCW,
// padded to MAXV by repeated vertices).  Edge k keeps the parameter range
// [t0[k], t1[k]] of p_k + t (p_{k+1} - p_k); t1 <= t0 means nothing is inside.
// Exactly collinear edge pairs are counted once: on p's side only when `keep_same`
// is set and the two edges run the same way.
//
// The side-of-line values f = cross(q_{m+1} - q_m, p - q_m) are evaluated in double
// with separately rounded products (no fused multiply-add): for float vertices the
// differences are exact and a truly collinear point gives exactly 0, so both polygons
// are classified consistently (A's edges against B's lines and B's against A's).
// Green's sum then describes one closed boundary even at exact contacts.
template <typename T>
__host__ __device__ inline void clip_edges(const T (&px)[MAXV], const T (&py)[MAXV], const T (&qx)[MAXV],
                                  const T (&qy)[MAXV], bool keep_same, T (&t0)[MAXV],
                                  T (&t1)[MAXV]) {
#pragma clang fp contract(off)
    // Half-plane by half-plane (m outer): only one half-plane's double line and the four
    // side values of p's vertices against it are live at a time.  Each edge's bounds see
    // the half-planes in the order m = 0, 1, ..., as an edge-outer loop would.
    double pxd[MAXV], pyd[MAXV];
    T ex[MAXV], ey[MAXV];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        const int k1 = (k + 1) % MAXV;
        pxd[k] = (double)px[k];
        pyd[k] = (double)py[k];
        ex[k] = px[k1] - px[k];
        ey[k] = py[k1] - py[k];
        t0[k] = T(0);
        t1[k] = ((ex[k] == T(0)) & (ey[k] == T(0))) ? T(0) : T(1);
    }
#pragma unroll
    for (int m = 0; m < MAXV; ++m) {
        const int m1 = (m + 1) % MAXV;
        const T qex = qx[m1] - qx[m], qey = qy[m1] - qy[m];
        const double dqx = (double)qx[m], dqy = (double)qy[m];
        const double dex = (double)qx[m1] - dqx, dey = (double)qy[m1] - dqy;
        const bool qdeg = (dex == 0.0) & (dey == 0.0);
        T f[MAXV];
#pragma unroll
        for (int v = 0; v < MAXV; ++v) {
            const double a = dex * (pyd[v] - dqy);
            const double b = dey * (pxd[v] - dqx);
            const double val = a - b;
            f[v] = qdeg ? T(1) : (T)val;
        }
#pragma unroll
        for (int k = 0; k < MAXV; ++k) {
            const int k1 = (k + 1) % MAXV;
            const T f0 = f[k], f1 = f[k1];
            const bool coll = (f0 == T(0)) & (f1 == T(0));
            const bool same = keep_same & (ex[k] * qex + ey[k] * qey > T(0));
            const bool out = ((f0 < T(0)) & (f1 < T(0))) | (coll & !same);
            const T tt = f0 * fast_rcp(f0 - f1);
            const bool leave = (f0 >= T(0)) & (f1 < T(0));
            const bool enter = (f0 < T(0)) & (f1 >= T(0));
            T hi = t1[k], lo = t0[k];
            hi = leave ? fmin(hi, tt) : hi;
            lo = enter ? fmax(lo, tt) : lo;
            hi = out ? T(-1) : hi;
            t0[k] = lo;
            t1[k] = hi;
        }
#ifdef __HIP_DEVICE_COMPILE__
        // one half-plane at a time: interleaving them buys no latency the other waves do
        // not hide, and holds enough values to push the exact kernels past 128 registers
        // (three waves per SIMD instead of four)
        __builtin_amdgcn_sched_barrier(0);
#endif
    }
}

// Sums over the clipped edges of polygon p: twice the signed area term
// sum dt * cross(p_k, e_k), the chain displacement, the segment midpoints, and the
// shape derivative of the area w.r.t. a rigid motion of p about centre (cx, cy):
//   d/dx = sum dt e_y,  d/dy = -sum dt e_x,  d/dtheta = -sum dt (mid - c) . e
template <typename T, bool GRAD>
__host__ __device__ inline void edge_sums(const T (&px)[MAXV], const T (&py)[MAXV], const T (&t0)[MAXV],
                                 const T (&t1)[MAXV], T cx, T cy, T& a2, T& Dx, T& Dy, T& mx,
                                 T& my, int& cnt, T (&g)[3]) {
#pragma unroll
    for (int k = 0; k < MAXV; ++k) {
        const int k1 = (k + 1) % MAXV;
        const T ex = px[k1] - px[k], ey = py[k1] - py[k];
        const bool valid = t1[k] > t0[k];
        const T dt = valid ? t1[k] - t0[k] : T(0);
        const T tm = T(0.5) * (t0[k] + t1[k]);
        const T midx = px[k] + tm * ex, midy = py[k] + tm * ey;
        a2 += dt * (px[k] * ey - py[k] * ex);
        Dx += dt * ex;
        Dy += dt * ey;
        mx += valid ? midx : T(0);
        my += valid ? midy : T(0);
        cnt += valid ? 1 : 0;
        if (GRAD) {
            g[0] += dt * ey;
            g[1] -= dt * ex;
            g[2] -= dt * ((midx - cx) * ex + (midy - cy) * ey);
        }
    }
}

// Area of A ∩ B (A about the origin = centre of tree i, B about (djx, djy) = centre of
// tree j) and its gradient w.r.t. both poses.  The Green's-theorem sum is taken about
// the mean of the clipped segment midpoints, which keeps near-collinear contacts (where
// rounding can count an edge on one side only) accurate to rounding.
template <typename T, bool GRAD>
__host__ __device__ inline T area_pair(const T (&ax)[MAXV], const T (&ay)[MAXV], const T (&bx)[MAXV],
                              const T (&by)[MAXV], T djx, T djy, T (&gi)[3], T (&gj)[3]) {
    T t0[MAXV], t1[MAXV];
    T a2 = 0, Dx = 0, Dy = 0, mx = 0, my = 0;
    int cnt = 0;
    clip_edges<T>(ax, ay, bx, by, true, t0, t1);
    edge_sums<T, GRAD>(ax, ay, t0, t1, T(0), T(0), a2, Dx, Dy, mx, my, cnt, gi);
    clip_edges<T>(bx, by, ax, ay, false, t0, t1);
    edge_sums<T, GRAD>(bx, by, t0, t1, djx, djy, a2, Dx, Dy, mx, my, cnt, gj);
    if (cnt > 0) {
        const T inv = T(1) / T(cnt);
        const T ox = mx * inv, oy = my * inv;
        a2 -= ox * Dy - oy * Dx;
    }
    return T(0.5) * a2;
}

// SAT penetration depth of two convex pieces and its gradient w.r.t. both poses
// (value and critical-vertex choice as core/pack_cuda_primitives.py:544-828, axes of
// A first; gradient for B derived the same way with the roles of the axes swapped).
template <typename T, bool GRAD>
__host__ __device__ inline T sat_pair(const T (&ax)[MAXV], const T (&ay)[MAXV], const T (&bx)[MAXV],
                             const T (&by)[MAXV], T djx, T djy, T (&gi)[3], T (&gj)[3]) {
    T best = T(1e30);
    bool separated = false;
    bool case1 = true, from_a = true;
    T bnx = 0, bny = 0, vx = 0, vy = 0, wx = 0, wy = 0;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int k = 0; k < MAXV; ++k) {
            const int k1 = (k + 1) % MAXV;
            const T ex = pass == 0 ? ax[k1] - ax[k] : bx[k1] - bx[k];
            const T ey = pass == 0 ? ay[k1] - ay[k] : by[k1] - by[k];
            T nx = -ey, ny = ex;
            const T len = sqrt(nx * nx + ny * ny);
            const bool skip = len < T(1e-12);
            const T il = skip ? T(0) : T(1) / len;
            nx *= il;
            ny *= il;
            T mn1 = ax[0] * nx + ay[0] * ny, mx1 = mn1;
            int imn1 = 0, imx1 = 0;
#pragma unroll
            for (int v = 1; v < MAXV; ++v) {
                const T p = ax[v] * nx + ay[v] * ny;
                if (p < mn1) { mn1 = p; imn1 = v; }
                if (p > mx1) { mx1 = p; imx1 = v; }
            }
            T mn2 = bx[0] * nx + by[0] * ny, mx2 = mn2;
            int imn2 = 0, imx2 = 0;
#pragma unroll
            for (int v = 1; v < MAXV; ++v) {
                const T p = bx[v] * nx + by[v] * ny;
                if (p < mn2) { mn2 = p; imn2 = v; }
                if (p > mx2) { mx2 = p; imx2 = v; }
            }
            if (!skip) {
                separated |= (mx1 < mn2) | (mx2 < mn1);
                const T pen1 = mx1 - mn2, pen2 = mx2 - mn1;
                const T pen = pen1 < pen2 ? pen1 : pen2;
                if (pen < best) {
                    best = pen;
                    bnx = nx; bny = ny;
                    from_a = pass == 0;
                    case1 = pen1 <= pen2;
                    const int iv = case1 ? imx1 : imn1;
                    const int iw = case1 ? imn2 : imx2;
                    // select by unrolled compare (no dynamic register indexing)
#pragma unroll
                    for (int v = 0; v < MAXV; ++v) {
                        if (v == iv) { vx = ax[v]; vy = ay[v]; }
                        if (v == iw) { wx = bx[v]; wy = by[v]; }
                    }
                }
            }
        }
    }
    if (separated || !(best > T(1e-12))) return T(0);
    if (GRAD) {
        const T sg = case1 ? T(1) : T(-1);
        const T dvx = vx - wx, dvy = vy - wy;
        const T axis = sg * (dvx * (-bny) + dvy * bnx);       // s (v - w) . J n
        gi[0] += sg * bnx;
        gi[1] += sg * bny;
        gi[2] += sg * ((-vy) * bnx + vx * bny) + (from_a ? axis : T(0));
        const T rwx = wx - djx, rwy = wy - djy;
        gj[0] -= sg * bnx;
        gj[1] -= sg * bny;
        gj[2] += -sg * ((-rwy) * bnx + rwx * bny) + (from_a ? T(0) : axis);
    }
    return best;
}

// ---------------------------------------------------------------------------------
// Workgroup-level pipeline
// ---------------------------------------------------------------------------------

// Ordered block-wide append.  Every thread of the block must call it (uniform).
// Returns this thread's slot offset (valid only if pred) and the block total.
template <int BLOCK>
__device__ inline int block_append(bool pred, int* wcnt, int& total) {
    constexpr int NW = BLOCK / WAVE;
    const uint64_t m = __ballot(pred);
    const int wave = threadIdx.x / WAVE;
    const int rank = __popcll(m & lanemask_lt());
    if (__lane_id() == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int c = wcnt[w];
        off += w < wave ? c : 0;
        tot += c;
    }
    total = tot;
    return off + rank;
}

struct Crystal {
    float ax, ay, bx, by;
    int n_periodic;
    int only_self;
};

// Grid cells: coordinates scaled by 1/side, clamped to +-30000 cells and biased by 2^15
// into 16 bits each.  Stencil s = 0 is the tree's own cell, s = 1..4 the forward
// neighbours (1,-1), (1,0), (1,1), (0,1): together each unordered pair of adjacent
// cells is visited from exactly one side.
template <typename T>
__device__ inline uint32_t cell_key(T fx, T fy) {
    fx = fmin(fmax(fx, T(-30000)), T(30000));
    fy = fmin(fmax(fy, T(-30000)), T(30000));
    const uint32_t cx = (uint32_t)((int)floor(fx) + 32768);
    const uint32_t cy = (uint32_t)((int)floor(fy) + 32768);
    return cx | (cy << 16);
}
__device__ inline uint32_t cell_step(uint32_t key, int s) {
    const int dx = s == 4 ? 0 : 1;
    const int dy = s == 1 ? -1 : (s == 2 ? 0 : 1);
    return key + (uint32_t)dx + (uint32_t)(dy * 65536);
}
__device__ inline int cell_bucket(uint32_t key, int H) {
    uint32_t h = (key & 0xffffu) * 0x9E3779B1u + (key >> 16) * 0x7FEB352Du;
    h ^= h >> 16;
    return (int)(h & (uint32_t)(H - 1));
}
// Reciprocal cell side: the side is 1.01 x the pair cutoff distance, so two trees
// within the cutoff (coordinates and products rounded) never lie two cells apart.
template <typename T>
__device__ inline T grid_inv_side(const GeomT<T>& G) {
    return T(1) / (T(1.01) * sqrt(G.cutoff_sq));
}

// LU: table lookups per lane per queue drain, all in flight together (more hide more of
// the gather latency where a wave has its SIMD to itself, at the cost of registers).
template <typename T, int MODE, bool GRAD, bool CRYSTAL, int BLOCK, int LU = 2>
struct Evaluator {
    static_assert(LU >= 1 && LU <= LUT_UMAX, "lookups per lane");
    using AccT = Acc<T>;
    using A = typename AccT::type;
    static constexpr int NW = BLOCK / WAVE;
    static_assert(NW <= MAX_WAVES, "at most 4 waves per workgroup");
    const GeomT<T>& G;
    EvalLds<T>& L;
    T cax, cay, cbx, cby;   // crystal axes
    int npd, only_self;
    LutDev lut;             // MODE_LUT only
    // statistics, this wave's share (summed over waves by the caller): tree pairs
    // within the cutoff, and piece pairs clipped or table lookups made
    long long n_tp, n_pp;
    A lane_cost;            // this lane's share of the cost, added to L.cost by run()
    T cut_list;             // squared neighbour-list radius (list_mode 1)

    __device__ Evaluator(const GeomT<T>& g, EvalLds<T>& l)
        : G(g), L(l), n_tp(0), n_pp(0), lane_cost(AccT::zero()), cut_list(T(0)) {}

    // Relative vector from tree i to the (shifted) tree j of a periodic tree-pair entry.
    __device__ inline void rel(uint32_t e, int& i, int& j, T& dx, T& dy, bool& self) const {
        i = e & 0xfff;
        j = (e >> 12) & 0xfff;
        dx = L.p[j].x - L.p[i].x;
        dy = L.p[j].y - L.p[i].y;
        self = i == j;
        if (CRYSTAL) {
            const int nc = 2 * npd + 1;
            const int cell = (int)(e >> 24);
            const int cx = cell / nc - npd, cy = cell % nc - npd;
            dx += T(cx) * cax + T(cy) * cbx;
            dy += T(cx) * cay + T(cy) * cby;
        }
    }

    // Exact cost of piece pi of tree i against piece pj of tree j shifted by (dx, dy)
    // relative to tree i; cost to this lane, gradient to both trees.
    __device__ inline void piece_pair(int i, int j, int pi, int pj, T dx, T dy, bool self) {
        T ax[MAXV], ay[MAXV], bx[MAXV], by[MAXV];
        const uint32_t wa = G.pv[pi], wb = G.pv[pj];
        const T* va = L.v + 2 * i * L.NV;
        const T* vb = L.v + 2 * j * L.NV;
#pragma unroll
        for (int v = 0; v < MAXV; ++v) {
            const int ua = (wa >> (8 * v)) & 0xff, ub = (wb >> (8 * v)) & 0xff;
            ax[v] = va[2 * ua];
            ay[v] = va[2 * ua + 1];
            bx[v] = vb[2 * ub] + dx;
            by[v] = vb[2 * ub + 1] + dy;
        }
        T gi[3] = {0, 0, 0}, gj[3] = {0, 0, 0};
        T c;
        if (MODE == MODE_AREA) {
            c = area_pair<T, GRAD>(ax, ay, bx, by, dx, dy, gi, gj);
        } else {
            const T sep = sat_pair<T, GRAD>(ax, ay, bx, by, dx, dy, gi, gj);
            c = sep * sep;
#pragma unroll
            for (int k = 0; k < 3; ++k) { gi[k] *= T(2) * sep; gj[k] *= T(2) * sep; }
        }
        // Self images (periodic only) count twice when only they are summed and carry
        // no gradient, as core/pack_cuda.py:349,452-462.
        if (CRYSTAL && self && only_self) c *= T(2);
        lane_cost += AccT::local(c);
        if (GRAD && !self) {
            AccT::add(&L.g[3 * i + 0], gi[0]);
            AccT::add(&L.g[3 * i + 1], gi[1]);
            AccT::add(&L.g[3 * i + 2], gi[2]);
            AccT::add(&L.g[3 * j + 0], gj[0]);
            AccT::add(&L.g[3 * j + 1], gj[1]);
            AccT::add(&L.g[3 * j + 2], gj[2]);
        }
    }

    // Table lookup for tree pair (i, j) (core/pack_cuda_lut.py:420-547): relative pose
    // of j in i's frame, value (optionally max(0,v)^2) and gradient to both trees.
    struct LutPair {
        int i, j;
        float cr, sr, dxl, dyl;
        LutTap tap;
    };
    // Poses of a pair as read from LDS (loaded for every lookup of a drain before any is
    // used, so that the LDS round trips overlap).
    struct LutPose {
        float xi, yi, ti, ci, si, xj, yj, tj;
    };
    __device__ inline void lut_load(int i, int j, LutPose& o) const {
        const Pose<T> a = L.p[i];
        o.xi = a.x;
        o.yi = a.y;
        o.ci = a.c;
        o.si = a.s;
        o.ti = L.t[i];
        o.xj = L.p[j].x;
        o.yj = L.p[j].y;
        o.tj = L.t[j];
    }
    __device__ inline void lut_begin(int i, int j, const LutPose& o, LutPair& q) const {
        q.i = i;
        q.j = j;
        const T dx = o.xj - o.xi, dy = o.yj - o.yi;
        q.cr = o.ci;
        q.sr = o.si;
        q.dxl = q.cr * dx + q.sr * dy;
        q.dyl = -q.sr * dx + q.cr * dy;
        const float dth = wrap_pi(o.tj - o.ti);
        lut_fetch(lut, q.dxl, q.dyl, dth, q.tap);
    }
    __device__ inline void lut_end(const LutPair& q) {
        float gl[3];
        float v = lut_combine(lut, q.tap, gl);
        if (lut.quad) {
            const float k = v > 0.f ? 2.f * v : 0.f;
            gl[0] *= k; gl[1] *= k; gl[2] *= k;
            v = v > 0.f ? v * v : 0.f;
        }
        lane_cost += AccT::local(v);
        if (GRAD) {
            const int i = q.i, j = q.j;
            const float cr = q.cr, sr = q.sr;
            // the centre gradients of the pair are exact opposites: tree i's are the
            // negated fixed-point values of tree j's (the conversion rounds symmetrically)
            const A gx = AccT::local(gl[0] * cr - gl[1] * sr), gy = AccT::local(gl[0] * sr + gl[1] * cr);
            AccT::add_raw(&L.g[3 * i + 0], AccT::neg(gx));
            AccT::add_raw(&L.g[3 * i + 1], AccT::neg(gy));
            AccT::add(&L.g[3 * i + 2], gl[0] * q.dyl - gl[1] * q.dxl - gl[2]);
            AccT::add_raw(&L.g[3 * j + 0], gx);
            AccT::add_raw(&L.g[3 * j + 1], gy);
            AccT::add(&L.g[3 * j + 2], gl[2]);
        }
    }
    // Lookups for cnt <= 64 LU queue entries, LU per lane, with the table loads of all of
    // them in flight before any is interpolated.
    __device__ inline void lut_pairs(const uint32_t* q, int cnt) {
        const int lane = threadIdx.x % WAVE;
        // queue entries, then the poses of every pair (a lane without an entry reads tree
        // 0's, never used), then the table gathers, then the interpolations
        uint32_t e[LU];
#pragma unroll
        for (int u = 0; u < LU; ++u) e[u] = lane + WAVE * u < cnt ? q[lane + WAVE * u] : 0u;
        LutPose o[LU];
#pragma unroll
        for (int u = 0; u < LU; ++u) lut_load(e[u] & 0xfff, (e[u] >> 12) & 0xfff, o[u]);
        LutPair pr[LU];
#pragma unroll
        for (int u = 0; u < LU; ++u)
            if (lane + WAVE * u < cnt) lut_begin(e[u] & 0xfff, (e[u] >> 12) & 0xfff, o[u], pr[u]);
#pragma unroll
        for (int u = 0; u < LU; ++u)
            if (lane + WAVE * u < cnt) lut_end(pr[u]);
        n_pp += cnt;
    }

    // ---------------------------------------------------------------- grid path
    // Piece-pair queue entry: i | j << 12 | pi << 24 | pj << 28.
    __device__ inline void clip_entry(uint32_t e) {
        const int i = e & 0xfff, j = (e >> 12) & 0xfff, pi = (e >> 24) & 0xf, pj = e >> 28;
        piece_pair(i, j, pi, pj, L.p[j].x - L.p[i].x, L.p[j].y - L.p[i].y, false);
    }

    // True when piece pi of tree i and piece pj of tree j (shifted by dx, dy) are separated
    // along the line through their bounding-circle centres (direction ddx, ddy) by more than
    // 1e-4 radius: such a pair is far from contact, clips to nothing and has no SAT depth,
    // so it contributes exactly zero and needs no clipping.
    __device__ inline bool centre_axis_separated(int i, int j, int pi, int pj, T dx, T dy, T ddx,
                                                 T ddy) const {
        const uint32_t wa = G.pv[pi], wb = G.pv[pj];
        const T* va = L.v + 2 * i * L.NV;
        const T* vb = L.v + 2 * j * L.NV;
        T amax = -T(1e30), bmin = T(1e30);
#pragma unroll
        for (int v = 0; v < MAXV; ++v) {
            const int ua = (wa >> (8 * v)) & 0xff, ub = (wb >> (8 * v)) & 0xff;
            amax = fmax(amax, va[2 * ua] * ddx + va[2 * ua + 1] * ddy);
            bmin = fmin(bmin, (vb[2 * ub] + dx) * ddx + (vb[2 * ub + 1] + dy) * ddy);
        }
        const T gap = bmin - amax, mu = T(1e-4) * G.radius;
        return gap > T(0) && gap * gap > mu * mu * (ddx * ddx + ddy * ddy);
    }

    // Pieces pj of tree j against the items of iq (tree pair q of tq | piece pi of tree i
    // << 16) held by lanes below cnt: the item's poses and piece pi's circle are read once
    // per run of pj (the same pj on every lane: broadcast reads); a run ends when the
    // piece-pair queue holds 64 pairs, which are clipped before the lane state is read again
    // (so that no lane state is live across the clip).
    __device__ __attribute__((always_inline)) inline void expand_items(const uint32_t* tq, const uint32_t* iq,
                                                                        int cnt, uint32_t* pq, int& pn) {
        const int lane = threadIdx.x % WAVE;
        const int P = L.P;
        const bool act = lane < cnt;
        const uint32_t it = iq[act ? lane : 0];
        int pj = 0;
        while (pj < P) {
            {
                const int pi = (int)(it >> 16);
                const uint32_t e = tq[it & 0xffff];
                const int i = e & 0xfff, j = (e >> 12) & 0xfff;
                const T dx = L.p[j].x - L.p[i].x, dy = L.p[j].y - L.p[i].y;
                const T ci = L.p[i].c, si = L.p[i].s, cj = L.p[j].c, sj = L.p[j].s;
                const T cxa = ci * G.ccx[pi] - si * G.ccy[pi];
                const T cya = si * G.ccx[pi] + ci * G.ccy[pi];
                const T ra = G.cr[pi];
                for (; pj < P;) {
                    const T bx = G.ccx[pj], by = G.ccy[pj];
                    const T cxb = cj * bx - sj * by + dx;
                    const T cyb = sj * bx + cj * by + dy;
                    const T rr = ra + G.cr[pj];
                    const T ddx = cxb - cxa, ddy = cyb - cya;
                    bool pred = act && ddx * ddx + ddy * ddy <= rr * rr;
                    if (pred) pred = !centre_axis_separated(i, j, pi, pj, dx, dy, ddx, ddy);
                    const uint64_t m = __ballot(pred);
                    if (pred) pq[pn + __popcll(m & lanemask_lt())] = e | ((uint32_t)pi << 24) | ((uint32_t)pj << 28);
                    pn += __popcll(m);
                    ++pj;
                    if (pn >= WAVE) break;
                }
            }
            if (pn >= WAVE) {
                wave_sync();
                clip_entry(pq[pn - WAVE + lane]);
                pn -= WAVE;
                n_pp += WAVE;
                wave_sync();
            }
        }
    }

    // Expand the tn tree pairs of tq into the piece pairs whose bounding circles meet,
    // clipping 64 at a time from this wave's piece-pair queue.  First one lane per (tree
    // pair, piece pi of its first tree): a piece whose circle misses the second tree's
    // bounding circle (which holds every piece of that tree) meets none of its pieces and is
    // dropped; the others are compacted into the item queue iq and expanded 64 at a time
    // over the pieces of the second tree.  Dropped pieces would clip to exactly zero, so
    // the results are unchanged.
    __device__ inline void expand(const uint32_t* tq, int tn, uint32_t* iq, uint32_t* pq, int& pn) {
        const int lane = threadIdx.x % WAVE;
        const int P = L.P;
        const float iP = 1.f / (float)P;
        const int items = tn * P;
        int in_ = 0, u0 = 0;
        // one call site of expand_items (the clip is inlined once): each pass tests the next
        // 64 items, then expands 64 queued items, or the rest once every item is tested
        while (u0 < items || in_ > 0) {
            const int u = u0 + lane;
            bool pred = false;
            uint32_t item = 0;
            if (u < items) {
                // exact: u < 2^15 and (u + 1/2) / P is at least 1/(2 P) from an integer
                const int q = (int)(((float)u + 0.5f) * iP);
                const int pi = u - q * P;
                const uint32_t e = tq[q];
                const int i = e & 0xfff, j = (e >> 12) & 0xfff;
                const T dx = L.p[j].x - L.p[i].x, dy = L.p[j].y - L.p[i].y;
                const T ci = L.p[i].c, si = L.p[i].s;
                const T ox = ci * G.ccx[pi] - si * G.ccy[pi] - dx;
                const T oy = si * G.ccx[pi] + ci * G.ccy[pi] - dy;
                // tree j's bounding radius with a margin far above the rounding of the
                // centre and circle coordinates (read here: nothing is live across a clip)
                const T r = G.cr[pi] + (G.radius * T(1.0 + 1e-5) + T(1e-5));
                pred = ox * ox + oy * oy <= r * r;
                item = (uint32_t)q | ((uint32_t)pi << 16);
            }
            const uint64_t m = __ballot(pred);
            if (pred) iq[in_ + __popcll(m & lanemask_lt())] = item;
            in_ += __popcll(m);
            u0 += WAVE;
            if (in_ >= WAVE || (u0 >= items && in_ > 0)) {
                const int c = in_ < WAVE ? in_ : WAVE;
                wave_sync();
                expand_items(tq, iq + in_ - c, c, pq, pn);
                in_ -= c;
                wave_sync();
            }
        }
    }

    // Ballot-append a candidate to the wave's tree-pair queue; drain 64 entries when it
    // holds at least 64 (table lookups, or expansion into piece pairs).  Wave-uniform.
    // Exact modes, last: expand what is left (the expansion's one call site, so the clip is
    // inlined once); in four-wave workgroups the pairs left in all waves' queues are split
    // evenly over the waves first (which wave found a pair depends on where its trees lie,
    // and the busiest wave bounds the workgroup): each wave copies its share into its
    // fourth queue region and expands it there.  That step has a workgroup barrier, so
    // every wave makes the last call; the tree-pair queues and counts are used again only
    // after the caller's barrier.
    __device__ inline void push(bool pred, uint32_t entry, uint32_t* tq, int& tn, uint32_t* pq, int& pn,
                                bool last = false) {
        const int lane = threadIdx.x % WAVE;
        const uint64_t m = __ballot(pred);
        if (m) {
            if (pred) tq[tn + __popcll(m & lanemask_lt())] = entry;
            tn += __popcll(m);
            wave_sync();
        }
        if (MODE == MODE_LUT) {
            // the table queue uses both halves of the wave's queue space (2 QCAP entries)
            if (tn >= LU * WAVE) {
                lut_pairs(tq + tn - LU * WAVE, LU * WAVE);
                tn -= LU * WAVE;
                wave_sync();
            }
        } else if (tn >= WAVE || last) {
            const uint32_t* src = tq;
            int cnt = tn;
            if (NW == MAX_WAVES && last) {   // four-wave workgroups (the registers of the
                                             // two-wave register kernels have no room for it)
                if (lane == 0) L.wcnt[12 + threadIdx.x / WAVE] = tn;
                __syncthreads();
                int base[MAX_WAVES + 1];
                base[0] = 0;
#pragma unroll
                for (int w = 0; w < MAX_WAVES; ++w)   // wave-uniform
                    base[w + 1] = base[w] + (w < NW ? __builtin_amdgcn_readfirstlane(L.wcnt[12 + w]) : 0);
                const int total = base[NW], share = (total + NW - 1) / NW;
                const int lo = min(total, (int)(threadIdx.x / WAVE) * share);
                cnt = min(total, lo + share) - lo;
                uint32_t* gq = tq + 3 * XQ;
                for (int k = lane; k < cnt; k += WAVE) {
                    const int g = lo + k;
                    int qoff = g;   // position of pair g in the queue space
#pragma unroll
                    for (int v = 1; v < MAX_WAVES; ++v)
                        if (v < NW && g >= base[v]) qoff = v * L.wqs + (g - base[v]);
                    gq[k] = L.wq[qoff];
                }
                wave_sync();
                src = gq;
            }
            expand(src, cnt, tq + XQ, pq, pn);
            tn = 0;
            wave_sync();
        }
    }

    // Candidate pairs wave by wave, into the wave-local queue; no workgroup barrier until
    // the caller's.  list_mode 0: from the cell grid (each lane walks the stencil buckets
    // of one tree); 1: the same, also storing pairs within the list radius (cut_list) as
    // the neighbour list; 2: from the neighbour list.  Every mode passes exactly the pairs
    // within the cutoff to the queue, so the result does not depend on the mode.
    __device__ inline void run_grid(int list_mode) {
        const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
        uint32_t* tq = L.wq + wave * L.wqs;
        uint32_t* pq = tq + 2 * XQ;   // exact modes; the item queue is tq + XQ
        int tn = 0, pn = 0;
        const int N = L.N;
        if (list_mode == 2) {
            // four rows of 64 list entries per step, all loads issued before the tests
            constexpr int U = 4;
            const int nl = L.wcnt[8];
            for (int base = wave * WAVE; base < nl; base += U * BLOCK) {
                uint32_t e[U];
                bool pred[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = base + u * BLOCK + lane;
                    e[u] = k < nl ? L.nb[k] : 0u;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int k = base + u * BLOCK + lane;
                    const int i = e[u] & 0xfff, j = (e[u] >> 12) & 0xfff;
                    const T dx = L.p[j].x - L.p[i].x, dy = L.p[j].y - L.p[i].y;
                    pred[u] = k < nl && dx * dx + dy * dy <= G.cutoff_sq;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    n_tp += __popcll(__ballot(pred[u]));
                    push(pred[u], e[u], tq, tn, pq, pn);
                }
            }
        } else {
            // each wave walks a contiguous share of the trees (ceil(N / waves)), so that
            // every wave of a workgroup has candidates even when N < 64 x waves
            const int share = (N + NW - 1) / NW;
            const int lo = wave * share, hi = min(N, lo + share);
            for (int base = lo; base < hi; base += WAVE) {
                const int i = base + lane;
                bool alive = i < hi;
                uint32_t ci = 0, target = 0;
                T xi = T(0), yi = T(0);
                int s = 0, j = -1;
                if (alive) {
                    ci = L.cell[i];
                    target = ci;
                    xi = L.p[i].x;
                    yi = L.p[i].y;
                    j = L.head[cell_bucket(ci, L.H)];
                }
                while (true) {
                    bool pred = false, pl = false;
                    uint32_t entry = 0;
                    if (alive) {
                        while (j < 0 && s < 4) {
                            ++s;
                            target = cell_step(ci, s);
                            j = L.head[cell_bucket(target, L.H)];
                        }
                        if (j < 0) {
                            alive = false;
                        } else {
                            const int jj = j;
                            j = L.next[jj];
                            if (L.cell[jj] == target && (s > 0 || jj > i)) {
                                const T dx = L.p[jj].x - xi, dy = L.p[jj].y - yi;
                                const T d2 = dx * dx + dy * dy;
                                pred = d2 <= G.cutoff_sq;
                                pl = d2 <= cut_list;
                                const int a = i < jj ? i : jj, b = i < jj ? jj : i;
                                entry = (uint32_t)a | ((uint32_t)b << 12);
                            }
                        }
                    }
                    if (list_mode == 1) {
                        const uint64_t ml = __ballot(pl);
                        if (ml) {
                            int off = 0;
                            if (lane == 0) off = atomicAdd(&L.wcnt[8], __popcll(ml));
                            off = __shfl(off, 0);
                            const int q = off + __popcll(ml & lanemask_lt());
                            if (pl && q < L.nb_cap) L.nb[q] = entry;
                        }
                    }
                    n_tp += __popcll(__ballot(pred));
                    push(pred, entry, tq, tn, pq, pn);
                    if (__ballot(alive) == 0) break;
                }
            }
        }
        if (MODE == MODE_LUT) {
            lut_pairs(tq, tn);
        } else {
            push(false, 0u, tq, tn, pq, pn, true);
            if (lane < pn) clip_entry(pq[lane]);
            n_pp += pn;
        }
    }

    // ----------------------------------------------------- periodic-image path
    __device__ inline void flush_pp(int n) {
        for (int q = threadIdx.x; q < n; q += BLOCK) {
            const uint32_t e = L.pp[q];
            const uint32_t te = L.tp[e & 0xffff];
            const int pi = (e >> 16) & 0xf, pj = (e >> 20) & 0xf;
            int i, j;
            T dx, dy;
            bool self;
            rel(te, i, j, dx, dy, self);
            piece_pair(i, j, pi, pj, dx, dy, self);
        }
        if (threadIdx.x < WAVE) n_pp += n;
        __syncthreads();
    }

    // Expand the tree-pair list into piece pairs (circle test), clipping as the
    // piece-pair list fills.
    __device__ inline void flush_tp(int ntp) {
        const int PP = L.P * L.P;
        const int nitems = ntp * PP;
        int npp = 0;
        for (int base = 0; base < nitems; base += BLOCK) {
            const int u = base + threadIdx.x;
            bool pred = false;
            uint32_t entry = 0;
            if (u < nitems) {
                const int q = u / PP;
                const int r = u - q * PP;
                const int pi = r / L.P, pj = r - pi * L.P;
                int i, j;
                T dx, dy;
                bool self;
                rel(L.tp[q], i, j, dx, dy, self);
                const T cxa = L.p[i].c * G.ccx[pi] - L.p[i].s * G.ccy[pi];
                const T cya = L.p[i].s * G.ccx[pi] + L.p[i].c * G.ccy[pi];
                const T cxb = L.p[j].c * G.ccx[pj] - L.p[j].s * G.ccy[pj] + dx;
                const T cyb = L.p[j].s * G.ccx[pj] + L.p[j].c * G.ccy[pj] + dy;
                const T rr = G.cr[pi] + G.cr[pj];
                const T ddx = cxb - cxa, ddy = cyb - cya;
                pred = ddx * ddx + ddy * ddy <= rr * rr;
                entry = (uint32_t)q | ((uint32_t)pi << 16) | ((uint32_t)pj << 20);
            }
            int total;
            const int slot = block_append<BLOCK>(pred, L.wcnt, total);
            if (npp + total > PP_CAP) {
                flush_pp(npp);
                npp = 0;
            }
            if (pred) L.pp[npp + slot] = entry;
            npp += total;
            __syncthreads();
        }
        flush_pp(npp);
    }

    // Pairs i <= j times lattice cells; a tree's own images only for the positive half
    // of the cells (image s and -s have equal overlap).
    __device__ inline void run_periodic() {
        const int N = L.N;
        int ntp = 0;
        const int nc = 2 * npd + 1, ncell = nc * nc, center = (ncell - 1) / 2;
        const long long npair = (long long)N * (N + 1) / 2;
        const long long ncand = npair * ncell;
        for (long long base = 0; base < ncand; base += BLOCK) {
            const long long k = base + threadIdx.x;
            bool pred = false;
            uint32_t entry = 0;
            if (k < ncand) {
                const int cell = (int)(k % ncell);
                long long q = k / ncell;
                int i = 0;
                while (q >= N - i) { q -= N - i; ++i; }
                const int j = i + (int)q;
                const bool self = i == j;
                bool ok = self ? cell > center : !only_self;
                const uint32_t e = (uint32_t)i | ((uint32_t)j << 12) | ((uint32_t)cell << 24);
                int ii, jj;
                T dx, dy;
                bool sf;
                rel(e, ii, jj, dx, dy, sf);
                pred = ok && (dx * dx + dy * dy <= G.cutoff_sq);
                entry = e;
            }
            int total;
            const int slot = block_append<BLOCK>(pred, L.wcnt, total);
            if (ntp + total > TP_CAP) {
                flush_tp(ntp);
                ntp = 0;
            }
            if (pred) L.tp[ntp + slot] = entry;
            ntp += total;
            if (threadIdx.x < WAVE) n_tp += total;
            __syncthreads();
        }
        flush_tp(ntp);
    }

    // Evaluate every candidate pair; the caller synchronises the workgroup afterwards and
    // reads the cost with cost_total().  list_mode as run_grid (neighbour list: relaxation
    // only).  Float costs are summed in fixed point; double costs per wave in a fixed lane
    // order, then over the waves in order, so both are reproducible.
    __device__ inline void run(int list_mode = 0) {
        if (CRYSTAL)
            run_periodic();
        else
            run_grid(list_mode);
        if constexpr (sizeof(T) == 4) {
            if (lane_cost != AccT::zero()) AccT::add_raw(L.cost, lane_cost);
        } else {
            const double w = wave_sum(lane_cost);
            if (threadIdx.x % WAVE == 0) L.wsum[threadIdx.x / WAVE] = w;
        }
    }

    // Total cost, after the workgroup barrier that follows run().
    __device__ inline T cost_total() const {
        if constexpr (sizeof(T) == 4) {
            return AccT::get(*L.cost);
        } else {
            double c = 0.0;
            for (int w = 0; w < NW; ++w) c += L.wsum[w];
            return c;
        }
    }
};

// Stage poses (centres in L.p, angles in L.t) into sincos, rotated piece vertices (VERTS) and the
// cell grid (GRID); clear the accumulators.  Ends with a barrier.
template <typename T, int BLOCK, bool VERTS = true, bool GRID = true>
__device__ inline void stage_instance(const GeomT<T>& G, EvalLds<T>& L) {
    const int N = L.N, NV = L.NV;
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        T sn, cs;
        sincos_t(L.t[i], &sn, &cs);
        L.p[i].c = cs;
        L.p[i].s = sn;
    }
    for (int q = threadIdx.x; q < 3 * N + 1; q += BLOCK) L.g[q] = Acc<T>::zero();
    if (GRID)
        for (int h = threadIdx.x; h < L.H; h += BLOCK) L.head[h] = -1;
    __syncthreads();
    if (VERTS) {
        const int nv = N * NV;
        for (int q = threadIdx.x; q < nv; q += BLOCK) {
            const int i = q / NV;
            const int r = q - i * NV;
            const T lx = G.ux[r], ly = G.uy[r];
            L.v[2 * q] = L.p[i].c * lx - L.p[i].s * ly;
            L.v[2 * q + 1] = L.p[i].s * lx + L.p[i].c * ly;
        }
    }
    if (GRID) {
        const T inv = grid_inv_side(G);
        for (int i = threadIdx.x; i < N; i += BLOCK) {
            const uint32_t key = cell_key(L.p[i].x * inv, L.p[i].y * inv);
            L.cell[i] = key;
            L.next[i] = atomicExch(&L.head[cell_bucket(key, L.H)], i);
        }
    }
    if (VERTS || GRID) __syncthreads();
}

// A tree whose bounding circle (radius plus a margin far above the rounding of the vertex
// positions in boundary_tree) lies inside the square has no vertex outside: its boundary
// cost and gradients are zero, as the vertex loop would find.
template <typename T>
__device__ inline bool boundary_inside(const GeomT<T>& G, T x, T y, T half, T hx, T hy) {
    const T rx = fabs(x - hx), ry = fabs(y - hy);
    const T margin = G.radius * T(1e-4) + T(1e-5) * (T(1) + fabs(x) + fabs(y) + fabs(hx) + fabs(hy) + half);
    return rx + G.radius + margin <= half && ry + G.radius + margin <= half;
}

// Boundary term of one tree (core/pack_cuda.py:550-642): max over outline vertices of
// the squared distance outside the square h = [size, ox, oy]; gradients w.r.t. the pose
// and h at the arg-max vertex (strict >, first vertex on ties, vertex 0 and zero if
// inside).  G's outline holds vertex 0 and the vertices on the outline's convex hull, where
// that maximum lies (pk_geometry_create), padded with its last vertex, so whole blocks of
// 8 are read (one LDS round trip each); the arg-max vertex's index is carried.  CHECK:
// test the bounding circle first (callers that already know the tree is near the
// boundary skip it).
template <typename T, bool CHECK = true>
__device__ inline T boundary_tree(const GeomT<T>& G, T x, T y, T c, T s, T half, T hx, T hy,
                                  T (&g)[3], T (&gh)[3]) {
    if (CHECK && boundary_inside(G, x, y, half, hx, hy)) {
        g[0] = g[1] = g[2] = T(0);
        gh[0] = gh[1] = gh[2] = T(0);
        return T(0);
    }
    T best = T(0);
    int bi = 0;
    const int no = G.n_outline;
    for (int v0 = 0; v0 < no; v0 += 8) {
        T lxs[8], lys[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            lxs[u] = G.ox[v0 + u];
            lys[u] = G.oy[v0 + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T vx = c * lxs[u] - s * lys[u] + x - hx, vy = s * lxs[u] + c * lys[u] + y - hy;
            const T ax = fabs(vx), ay = fabs(vy);
            const T dx = ax > half ? ax - half : T(0), dy = ay > half ? ay - half : T(0);
            const T d2 = dx * dx + dy * dy;
            if (d2 > best) {
                best = d2;
                bi = v0 + u;
            }
        }
    }
    const T lx = G.ox[bi], ly = G.oy[bi];
    const T vx = c * lx - s * ly + x - hx, vy = s * lx + c * ly + y - hy;
    const T ax = fabs(vx), ay = fabs(vy);
    const T dx = ax > half ? ax - half : T(0), dy = ay > half ? ay - half : T(0);
    const T gvx = dx > T(0) ? T(2) * dx * (vx >= T(0) ? T(1) : T(-1)) : T(0);
    const T gvy = dy > T(0) ? T(2) * dy * (vy >= T(0) ? T(1) : T(-1)) : T(0);
    g[0] = gvx;
    g[1] = gvy;
    g[2] = gvx * (-s * lx - c * ly) + gvy * (c * lx - s * ly);
    gh[0] = -(dx + dy);
    gh[1] = -gvx;
    gh[2] = -gvy;
    return best;
}

}  // namespace pk
