// Fused relaxation engine: cost evaluation + L-BFGS iterations of one packing instance
// per workgroup, for a whole (ragged) batch in one launch per chunk of iterations.
//
// Objective (core/pack_dynamics.py:45-52,76-123 with fixed h): state x = [xyt (N,3) | h (3)],
//   cost = w_ov * overlap(area | sep | table) + w_bd * boundary + w_area * h0^2 (free h only);
// gradient w.r.t. h zeroed when h is fixed (core/pack_cost.py:178-188).
// Optimizer: core/lbfgs_torch_parallel.py:411-729 without line search, applied to each
// instance on its own (a batch of one), so an instance's result never depends on what
// else is in the launch: the two-loop direction (computed in the equivalent compact
// form), first step lr*min(1, 1/sum|g|), max-step clamp, directional-derivative stop,
// history append (a ring of slots) only from iteration 2 with y.s > 1e-10, and the
// reference's convergence tests.
//
// Two kernels: k_relax, one workgroup per instance with the vectors (and the history
// when it fits) in LDS and the L-BFGS arithmetic on wave 0; k_relax_w, one or two waves
// per instance with the state and history in registers.  Every reduction has a fixed
// order, so all give the same bits for any workgroup size and may continue each other's
// runs.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "pk_eval.h"
#include "packing.h"

namespace pk {

struct LbSlot {
    double loss, loss0, H;
    double ro[PK_MAX_HISTORY];
    int n_iter, n_eval, hist_len, head, status, reason, pad0, pad1;
};
static_assert(sizeof(LbSlot) == PK_LBFGS_SLOT_BYTES, "slot layout");

enum { ST_FRESH = 0, ST_RUN = 1, ST_DONE = 2 };

// reasons (bit set when an instance stops)
enum {
    R_GRAD = 1, R_EVAL = 2, R_PARAM = 4, R_LOSS = 8, R_REL = 16, R_INCREASE = 32, R_DIR = 64,
    R_ITER = 128
};

struct RelaxArgs {
    pk_relax_params p;
    const GeomT<float>* geom;
    LutDev lut;
};

__host__ __device__ inline int pad64(int n) { return (n + WAVE - 1) / WAVE * WAVE; }

// Neighbour-list skin: the list holds the pairs within cutoff + skin and stays valid while
// twice the accumulated centre displacement is below the skin.  0.75 of the polygon radius
// (measured on the ragged table workload: 0.35 R 1.25 ms per launch of the N <= 200
// bucket, 0.5 R 1.25, 0.75 R 1.17, 1.0 R and 1.25 R 1.17): longer lists, fewer rebuilds.
__device__ inline float list_skin(const GeomT<float>& G) { return 0.75f * G.radius; }

// Component j of the h gradient after an evaluation (relax_eval's formula), from the
// boundary sums; x0 = h[0].
__device__ inline float h_gradient(const pk_relax_params& p, const unsigned long long* bacc, int j, float x0) {
    if (p.fixed_h) return 0.f;
    if (j == 0) return p.w_boundary * Acc<float>::get(bacc[1]) + 2.f * p.w_area * x0;
    return p.w_boundary * Acc<float>::get(bacc[1 + j]);
}

// Cost and gradient of the state x into g (all threads); returns the loss.  list_mode:
// 1 rebuilds the neighbour list while evaluating, 2 evaluates from it (no cell grid).
// Each tree's poses, sincos and boundary term are handled by the same thread, so the
// staging needs no barrier between them.  HW: g receives the whole gradient: the weighted
// overlap and boundary terms of every tree, and (thread 0, visible to its own wave on
// return) the h gradient g[3N..3N+2].  Without HW the caller forms the gradient itself: g
// receives only the unweighted boundary gradient of the trees near the boundary (the
// caller keeps every other entry of g at zero), the overlap gradient stays in L.g and the
// h gradient is left to h_gradient.
template <int MODE, int BLOCK, int LU = 2, bool HW = true>
__device__ __forceinline__ double relax_eval(const GeomT<float>& G, EvalLds<float>& L, const RelaxArgs& A,
                                    const float* x, float* g, int N,
                                    unsigned long long* bacc /* [4] boundary accumulators */,
                                    long long& st_tp, long long& st_pp, int list_mode, float cut_list) {
    const pk_relax_params& p = A.p;
    const bool grid = list_mode != 2;
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        const float ti = x[3 * i + 2];
        L.t[i] = ti;
        float sn, cs;
        sincos_t(ti, &sn, &cs);
        L.p[i] = Pose<float>{x[3 * i], x[3 * i + 1], cs, sn};
    }
    for (int q = threadIdx.x; q < 3 * N + 1; q += BLOCK) L.g[q] = 0ull;
    if (grid)
        for (int h = threadIdx.x; h < L.H; h += BLOCK) L.head[h] = -1;
    if (threadIdx.x < 4) bacc[threadIdx.x] = 0ull;
    if (threadIdx.x == 0 && list_mode == 1) L.wcnt[8] = 0;
    __syncthreads();
    if (MODE != MODE_LUT) {
        const int NV = L.NV, nv = N * NV;
        for (int q = threadIdx.x; q < nv; q += BLOCK) {
            const int i = q / NV;
            const int r = q - i * NV;
            const float lx = G.ux[r], ly = G.uy[r];
            L.v[2 * q] = L.p[i].c * lx - L.p[i].s * ly;
            L.v[2 * q + 1] = L.p[i].s * lx + L.p[i].c * ly;
        }
    }
    if (grid) {
        const float inv = grid_inv_side(G);
        for (int i = threadIdx.x; i < N; i += BLOCK) {
            const uint32_t key = cell_key(L.p[i].x * inv, L.p[i].y * inv);
            L.cell[i] = key;
            L.next[i] = atomicExch(&L.head[cell_bucket(key, L.H)], i);
        }
    }
    if (MODE != MODE_LUT || grid) __syncthreads();
    Evaluator<float, MODE, true, false, BLOCK, LU> ev(G, L);
    ev.cax = ev.cay = ev.cbx = ev.cby = 0.f;
    ev.npd = 0;
    ev.only_self = 0;
    ev.lut = A.lut;
    ev.cut_list = cut_list;
    ev.run(list_mode);
    // The boundary pass below reads the fixed-point sums only with HW (without, it reads the
    // poses, staged before the last barrier, and writes near trees' entries of g only), and
    // reuses this wave's queue space, which in the exact modes of four-wave workgroups the
    // other waves read while they expand their share of all waves' tree pairs (so exact
    // modes keep the barrier); in table mode only this wave does.
    if (HW || MODE != MODE_LUT)
        __syncthreads();
    else
        wave_sync();   // the wave's last queue reads before the boundary pass reuses the queue
    st_tp += ev.n_tp;
    st_pp += ev.n_pp;
    const float half = x[3 * N] * 0.5f, hx = x[3 * N + 1], hy = x[3 * N + 2];
    unsigned long long lb[4] = {0ull, 0ull, 0ull, 0ull};
    // Boundary term: trees well inside the square (most of them) have zero boundary cost and
    // gradient and are finished in the first pass; the others are compacted into the wave's
    // queue space (free after the pair evaluation) and evaluated 64 at a time, so the
    // vertex loop runs on full rows.
    {
        const int lane = threadIdx.x % WAVE;
        uint32_t* bq = L.wq + (threadIdx.x / WAVE) * L.wqs;
        int bn = 0;
        auto finish = [&](int i) {
            float gb[3], gh[3];
            const float cb = boundary_tree<float, false>(G, L.p[i].x, L.p[i].y, L.p[i].c, L.p[i].s, half, hx, hy, gb, gh);
            lb[0] += Acc<float>::wide(cb);
            lb[1] += Acc<float>::wide(gh[0]);
            lb[2] += Acc<float>::wide(gh[1]);
            lb[3] += Acc<float>::wide(gh[2]);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                g[3 * i + k] = HW ? p.w_overlap * Acc<float>::get(L.g[3 * i + k]) + p.w_boundary * gb[k] : gb[k];
        };
        for (int base = threadIdx.x - lane; base < N; base += BLOCK) {
            const int i = base + lane;
            bool near = false;
            if (i < N) {
                near = !boundary_inside(G, L.p[i].x, L.p[i].y, half, hx, hy);
                if (HW && !near) {
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        g[3 * i + k] = p.w_overlap * Acc<float>::get(L.g[3 * i + k]) + p.w_boundary * 0.f;
                }
            }
            const uint64_t m = __ballot(near);
            if (near) bq[bn + __popcll(m & lanemask_lt())] = (uint32_t)i;
            bn += __popcll(m);
            wave_sync();
            if (bn >= WAVE) {
                finish((int)bq[bn - WAVE + lane]);
                bn -= WAVE;
                wave_sync();
            }
        }
        if (lane < bn) finish((int)bq[lane]);
    }
    // integer sums: exact in any order
    if (__ballot((lb[0] | lb[1] | lb[2] | lb[3]) != 0ull)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) Acc<float>::add_raw(&bacc[k], lb[k]);
    }
    __syncthreads();
    const double c_ov = Acc<float>::get(*L.cost), c_bd = Acc<float>::get(bacc[0]);
    double loss = (double)p.w_overlap * c_ov + (double)p.w_boundary * c_bd;
    // the h gradient is read by wave 0, which thread 0 belongs to, after this returns
    if (HW && threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < 3; ++j) g[3 * N + j] = h_gradient(p, bacc, j, x[3 * N]);
    }
    if (!p.fixed_h) loss += (double)p.w_area * (double)x[3 * N] * (double)x[3 * N];
    return loss;
}

// Dynamic LDS of one workgroup: evaluator, x g g_prev d, the small L-BFGS matrices S^T Y and
// Y^T Y (history x history doubles each), and (hl) the history s, y; the vectors padded to a
// multiple of 64 floats.
size_t relax_lds_bytes(int mode, int n_max, int NV, int m, bool hl, int block) {
    const int Pe = mode == MODE_LUT ? 0 : NV;
    size_t b = eval_lds_bytes<float>(n_max, Pe, false, block / WAVE, true);
    const size_t np = (size_t)pad64(3 * n_max + 3);
    b += sizeof(float) * 4 * np;
    b += sizeof(double) * 2 * (size_t)m * m;
    b = align_up(b, 16);
    if (hl) b += sizeof(float) * 2 * (size_t)m * np;
    return align_up(b, 16);
}

// History rows: physical slot offsets into s and y (row stride np in LDS, n in global
// memory, where elements beyond n belong to the next row and read as zero).
template <bool HL>
struct Hist {
    float* s;
    float* y;
    int stride, n;
    __device__ inline float S(int off, int k) const { return (HL || k < n) ? s[off + k] : 0.f; }
    __device__ inline float Y(int off, int k) const { return (HL || k < n) ? y[off + k] : 0.f; }
};

// ---------------------------------------------------------------------------------
// L-BFGS arithmetic shared by both relaxation kernels.  Vectors are walked element by
// element, k = lane + 64 r with r increasing; per-lane partial sums are float fused
// multiply-adds (at most a few terms per lane), and every wave-wide sum adds the 64
// partials through the butterflies of pk_common.h (in float for history dot products, in
// double for the step quantities), whose result does not depend on the list a value
// sits in.  Both kernels therefore produce the same bits, wherever the vectors live.
// Dot products with history rows are separate sums, so the order the rows come in
// (chronological in k_relax, by ring slot in k_relax_w) does not change them; the
// direction adds its history terms in ring-slot order in both kernels.
// ---------------------------------------------------------------------------------

// Canonical order of every sum over a state vector, in every kernel and layout: element k
// is row k / 64 of lane k % 64; rows 0-4 form group A and rows 5 and above group B; per
// lane each group's terms are chained in increasing row order; each group is summed over
// the 64 lanes by one butterfly; the two group totals are added, A + B (B = +0 when the
// vector has no row beyond 4).  A workgroup of two waves can then hold an instance, one
// group per wave, and produce the bits of one wave (or one LDS pass) holding both.
constexpr int ROWS_A = 5;
constexpr int ELEMS_A = ROWS_A * WAVE;

constexpr int pow2_ceil(int k) { return k <= 1 ? 1 : 2 * pow2_ceil((k + 1) / 2); }

// Ring slot of chronological row i: (head + i) mod m for 0 <= head, i < m, without an
// integer division.
__host__ __device__ inline int ring(int head, int i, int m) {
    const int v = head + i;
    return v >= m ? v - m : v;
}

template <int KH>
struct Lb {
    static constexpr int KD = 2 * KH;   // dots: i = s_i.v, KH + i = y_i.v

    __device__ static inline void dots(float (&acc)[KD], const float (&sv)[KH], const float (&yv)[KH],
                                       float v, int nv) {
#pragma unroll
        for (int i = 0; i < KH; ++i) {
            if (i < nv) {
                acc[i] = fmaf(sv[i], v, acc[i]);
                acc[KH + i] = fmaf(yv[i], v, acc[KH + i]);
            }
        }
    }
    // y.s and y.y of a pair (per-lane partials; summed over the wave in double)
    __device__ static inline void pair(float (&a)[2], float yk, float sk) {
        a[0] = fmaf(yk, sk, a[0]);
        a[1] = fmaf(yk, yk, a[1]);
    }
    // d_k = -gamma g_k + sum_i (cs_i s_i,k + cy_i y_i,k), in float, in this order, over all
    // KH rows: rows at or beyond the used count are zero, as are their coefficients.
    __device__ static inline float dir(float gamf, float gk, const float (&cs)[KH], const float (&cy)[KH],
                                       const float (&sv)[KH], const float (&yv)[KH]) {
        float dk = -gamf * gk;
#pragma unroll
        for (int i = 0; i < KH; ++i) {
            dk += cs[i] * sv[i];
            dk += cy[i] * yv[i];
        }
        return dk;
    }

    // Direction coefficients of the compact L-BFGS form (Byrd, Nocedal and Schnabel 1994,
    // eq. 3.1): with a = S^T g, b = Y^T g (r: the dots butterfly), R = upper part of S^T Y,
    // D = diag(S^T Y), gamma = H_diag:
    //   u = R^-1 a,  w = R^-T ((D + gamma Y^T Y) u - gamma b),  H g = gamma g + S w - gamma Y u,
    // the same direction as the two-loop recursion of core/lbfgs_torch_parallel.py:532-572.
    // The small matrices are indexed by ring slot, (head + i) mod m for chronological i.  The
    // hl x hl triangular solves run one row per lane.  Returns, in lane i < hl, the
    // coefficients of d = -gamma g + sum_i cs_i s_i + cy_i y_i.  a, b: s_i.g and y_i.g of
    // this lane's row i = lane mod KH.
    __device__ static inline void solve(float r, int hl, int head, int m, double gam, const double* SY,
                                        const double* YY, const double* ro, float& cs_l, float& cy_l) {
        const int li = __lane_id() & (KH - 1);
        solve_ab(__shfl(r, many_src<KD>(li)), __shfl(r, many_src<KD>(KH + li)), hl, head, m, gam, SY, YY, ro,
                 cs_l, cy_l);
    }
    __device__ static inline void solve_ab(float a, float b, int hl, int head, int m, double gam,
                                           const double* SY, const double* YY, const double* ro, float& cs_l,
                                           float& cy_l) {
        const int lane = __lane_id();
        const double a_l = (double)a;
        const double b_l = (double)b;
        const bool row = lane < hl;
        const int pl = ring(head, lane < m ? lane : 0, m);
        // all matrix entries this lane needs, loaded before the dependent chains: row pl of
        // S^T Y and Y^T Y, column pl of S^T Y, and 1 / s_j.y_j
        double syr[KH], yyr[KH], syc[KH], rj[KH];
#pragma unroll
        for (int j = 0; j < KH; ++j) {
            const int pj = ring(head, j < m ? j : 0, m);
            syr[j] = SY[pl * m + pj];
            yyr[j] = YY[pl * m + pj];
            syc[j] = SY[pj * m + pl];
            rj[j] = ro[pj];
        }
        const double dl = SY[pl * m + pl];
        // Lane i's running value is read once, at step j = i; the updates of lanes at or
        // beyond j (outside the triangle) come after that read or belong to unused rows, so
        // they need no mask.
        double acc = a_l, u_l = 0.0, zacc = 0.0;
#pragma unroll
        for (int j = KH - 1; j >= 0; --j) {
            if (j < hl) {
                const double uj = lane_d(acc, j) * rj[j];
                if (lane == j) u_l = uj;
                acc -= syr[j] * uj;
                zacc += yyr[j] * uj;
            }
        }
        const double z_l = row ? dl * u_l + gam * zacc - gam * b_l : 0.0;
        double accw = z_l, w_l = 0.0;
#pragma unroll
        for (int j = 0; j < KH; ++j) {
            if (j < hl) {
                const double wj = lane_d(accw, j) * rj[j];
                if (lane == j) w_l = wj;
                accw -= syc[j] * wj;
            }
        }
        cs_l = row ? (float)(-w_l) : 0.f;
        cy_l = row ? (float)(gam * u_l) : 0.f;
    }

    // Accept the pair into ring slot q: new column of S^T Y and row of Y^T Y from r (the
    // dots butterfly of y_new over the hl rows in use; a dropped oldest row is harmless),
    // the diagonal from the pair's own sums, H_diag, ro and the ring counters.
    __device__ static inline void store(float r, double ys, double yy, int hl, int head, int m, int q,
                                        double* SY, double* YY, LbSlot& S) {
        const int li = __lane_id() & (KH - 1);
        store_v(__shfl(r, many_src<KD>(li)), __shfl(r, many_src<KD>(KH + li)), ys, yy, hl, head, m, q, SY, YY,
                S);
    }
    // The same with this lane's s_i.y_new and y_i.y_new (row i = lane) given.
    __device__ static inline void store_v(float sy_l, float yy_l, double ys, double yy, int hl, int head, int m,
                                          int q, double* SY, double* YY, LbSlot& S) {
        const int lane = __lane_id();
        const double syc = (double)sy_l;
        const double yyr = (double)yy_l;
        if (lane < hl) {
            const int pi = ring(head, lane, m);
            if (pi != q) {
                SY[pi * m + q] = syc;
                YY[pi * m + q] = yyr;
                YY[q * m + pi] = yyr;
            }
        }
        if (lane == 0) {
            SY[q * m + q] = ys;
            YY[q * m + q] = yy;
            S.H = ys / yy;
            S.ro[q] = 1.0 / ys;
            if (S.hist_len < m) S.hist_len++;
            else S.head = ring(S.head, 1, m);
        }
    }

    // Chronological index of ring slot p, (p - head) mod m, for p < m; p itself beyond.
    __device__ static inline int chron(int p, int head, int m) {
        return p < m ? (p >= head ? p - head : p - head + m) : p;
    }
    // Direction coefficients by ring slot: lane p receives the coefficient of chronological
    // row chron(p) (zero for rows not in use, as solve leaves them).
    __device__ static inline float to_slot(float c_l, int head, int m) {
        return __shfl(c_l, chron(__lane_id(), head, m));
    }
    // Column of ring slot p from the dots butterfly r of y_p against every slot (value i:
    // s_i.y_p, KH + i: y_i.y_p) and the pair sums of (s_p, y_p): the entries of the slots
    // older than p, as store() wrote them.
    __device__ static inline void rebuild_col_slot(float r, double dsy, double dyy, int p, int hl, int head,
                                                   int m, double* SY, double* YY) {
        const int lane = __lane_id();
        const int li = lane & (KH - 1);
        const double sy = (double)__shfl(r, many_src<KD>(li));
        const double yy = (double)__shfl(r, many_src<KD>(KH + li));
        if (lane < hl && chron(lane, head, m) < chron(p, head, m)) {
            SY[lane * m + p] = sy;
            YY[lane * m + p] = yy;
            YY[p * m + lane] = yy;
        }
        if (lane == 0) {
            SY[p * m + p] = dsy;
            YY[p * m + p] = dyy;
        }
    }

    // Column j (chronological) of S^T Y (upper) and Y^T Y from the dots butterfly r of y_j
    // and the pair sums of (s_j, y_j), as store() wrote them.
    __device__ static inline void rebuild_col(float r, double dsy, double dyy, int j, int head, int m,
                                              double* SY, double* YY) {
        const int lane = __lane_id();
        const int li = lane & (KH - 1);
        const double sy = (double)__shfl(r, many_src<KD>(li));
        const double yy = (double)__shfl(r, many_src<KD>(KH + li));
        const int pj = ring(head, j, m);
        if (lane < j) {
            const int pi = ring(head, lane, m);
            SY[pi * m + pj] = sy;
            YY[pi * m + pj] = yy;
            YY[pj * m + pi] = yy;
        }
        if (lane == 0) {
            SY[pj * m + pj] = dsy;
            YY[pj * m + pj] = dyy;
        }
    }
};

// History rows of element k for the first nv entries of off (row offsets), from LDS or
// global memory; entries at or beyond nv read as zero.
template <int KH, bool HL>
__device__ inline void hist_elem(const Hist<HL>& H, const int (&off)[KH], int k, int nv, float (&sv)[KH],
                                 float (&yv)[KH]) {
#pragma unroll
    for (int i = 0; i < KH; ++i) {
        sv[i] = i < nv ? H.S(off[i], k) : 0.f;
        yv[i] = i < nv ? H.Y(off[i], k) : 0.f;
    }
}
template <int KH, bool HL>
__device__ inline void hist_offsets(const Hist<HL>& H, int head, int m, int (&off)[KH]) {
#pragma unroll
    for (int i = 0; i < KH; ++i) off[i] = ring(head, i < m ? i : 0, m) * H.stride;
}

// Dots of v with the first nv history rows (LDS / global history).  Wave 0 only.
template <int KH, bool HL, typename VF>
__device__ inline float lb_dots(int np, int nv, int head, int m, const Hist<HL>& H, VF v) {
    int off[KH];
    hist_offsets<KH, HL>(H, head, m, off);
    auto group = [&](int k0, int k1) {
        float acc[Lb<KH>::KD];
#pragma unroll
        for (int i = 0; i < Lb<KH>::KD; ++i) acc[i] = 0.f;
        for (int k = k0 + (int)threadIdx.x; k < k1; k += WAVE) {
            float sv[KH], yv[KH];
            hist_elem<KH, HL>(H, off, k, nv, sv, yv);
            Lb<KH>::dots(acc, sv, yv, v(k), nv);
        }
        return wave_sum_many_f<Lb<KH>::KD>(acc);
    };
    const float ra = group(0, np < ELEMS_A ? np : ELEMS_A);
    const float rb = np > ELEMS_A ? group(ELEMS_A, np) : 0.f;
    return ra + rb;
}

// y.s and y.y (or s.y, y.y of a stored pair) summed in the canonical order.  Wave 0 only.
template <typename PF>
__device__ inline double lb_pair_sums(int np, PF pf) {
    float pa[2] = {0.f, 0.f}, pb[2] = {0.f, 0.f};
    const int ka = np < ELEMS_A ? np : ELEMS_A;
    for (int k = threadIdx.x; k < ka; k += WAVE) pf(k, pa);
    for (int k = ELEMS_A + threadIdx.x; k < np; k += WAVE) pf(k, pb);
    const double ra = wave_sum_many<2>(pa);
    const double rb = np > ELEMS_A ? wave_sum_many<2>(pb) : 0.0;
    return ra + rb;
}

// Rebuild S^T Y and Y^T Y from the stored history rows (the same sums as the incremental
// updates, so a call that starts from stored history continues bit for bit).  Wave 0 only.
template <int KH, bool HL>
__device__ inline void lb_rebuild(int np, int hl, int head, int m, const Hist<HL>& H, double* SY,
                                  double* YY) {
    for (int j = 0; j < hl; ++j) {
        const int oj = ring(head, j, m) * H.stride;
        const float r = lb_dots<KH, HL>(np, j, head, m, H, [&](int k) { return H.Y(oj, k); });
        const double rp = lb_pair_sums(np, [&](int k, float (&a)[2]) { Lb<KH>::pair(a, H.Y(oj, k), H.S(oj, k)); });
        Lb<KH>::rebuild_col(r, lane_d(rp, many_src<2>(0)), lane_d(rp, many_src<2>(1)), j, head, m, SY, YY);
    }
}

// Direction pass: d written to LDS, with g.d and sum |g| (one butterfly), max |d| and
// max |x| (another).  The history terms are added in ring-slot order.  Wave 0 only.
template <int KH, bool HL>
__device__ inline void lb_direction(int np, int hl, int head, int m, float gamf, const Hist<HL>& H,
                                    const float* g, const float* x, float* d, float cs_l, float cy_l,
                                    double& gtd, double& gsum, double& dmax, double& xmax) {
    // rows by ring slot (slots at or beyond hl unused), coefficients permuted to match
    const float cs_s = Lb<KH>::to_slot(cs_l, head, m), cy_s = Lb<KH>::to_slot(cy_l, head, m);
    float cs[KH], cy[KH];
#pragma unroll
    for (int i = 0; i < KH; ++i) {
        cs[i] = lane_f(cs_s, i);
        cy[i] = lane_f(cy_s, i);
    }
    int off[KH];
#pragma unroll
    for (int i = 0; i < KH; ++i) off[i] = (i < m ? i : 0) * H.stride;
    double sa[2] = {0.0, 0.0}, sb[2] = {0.0, 0.0}, mv[2] = {0.0, 0.0};
    auto body = [&](int k, double (&sv2)[2]) {
        float sv[KH], yv[KH];
        hist_elem<KH, HL>(H, off, k, hl, sv, yv);
        const float gk = g[k];
        const float dk = Lb<KH>::dir(gamf, gk, cs, cy, sv, yv);
        d[k] = dk;
        sv2[0] += (double)gk * (double)dk;
        sv2[1] += fabs((double)gk);
        mv[0] = fmax(mv[0], fabs((double)dk));
        mv[1] = fmax(mv[1], fabs((double)x[k]));
    };
    const int ka = np < ELEMS_A ? np : ELEMS_A;
    for (int k = threadIdx.x; k < ka; k += WAVE) body(k, sa);
    for (int k = ELEMS_A + threadIdx.x; k < np; k += WAVE) body(k, sb);
    const double rsa = wave_sum_many<2>(sa);
    const double rs = rsa + (np > ELEMS_A ? wave_sum_many<2>(sb) : 0.0);
    const double rm = wave_max_many<2>(mv);
    gtd = lane_d(rs, many_src<2>(0));
    gsum = lane_d(rs, many_src<2>(1));
    dmax = lane_d(rm, many_src<2>(0));
    xmax = lane_d(rm, many_src<2>(1));
}

// HL: the history s_hist / y_hist is copied into LDS at the start of the call and back at
// the end; otherwise the L-BFGS passes read it from global memory (row stride n there,
// padded stride np in LDS).  KH: history rows the unrolled passes handle (history <= KH).
template <int MODE, int BLOCK, bool HL, int KH>
__global__ __launch_bounds__(BLOCK) void k_relax(RelaxArgs A) {
    extern __shared__ __align__(16) char smem[];
    __shared__ GeomT<float> G;
    __shared__ LbSlot S;
    __shared__ unsigned long long bacc[4];
    __shared__ int stop;
    __shared__ double t_step;
    __shared__ int eval_mode;     // neighbour list: 1 rebuild, 2 reuse
    __shared__ int nb_valid;      // the list holds every pair within the list radius
    __shared__ double vdisp;      // bound on any tree's displacement since the rebuild
    const pk_relax_params& p = A.p;
    const int b = blockIdx.x;
    if (b >= p.n_launch) return;
    const int slot = p.order ? p.order[b] : b;
    const int N = p.ntree[slot];
    // LDS, the row count and the tile were sized for n_tree_max: a larger count is refused
    // (status -1) rather than run past them
    if (N <= 0 || N > p.n_tree_max) {
        if (N > 0 && threadIdx.x == 0 && p.status_out) p.status_out[slot] = -1;
        return;
    }
    LbSlot* gS = reinterpret_cast<LbSlot*>(p.state) + slot;
    if (threadIdx.x == 0) S = *gS;
    load_geom(G, A.geom);
    __syncthreads();
    if (S.status == ST_DONE) return;

    const int n = 3 * N + 3, np = pad64(n), m = p.history;
    const long long xoff = p.xoff[slot];
    float* __restrict__ X = p.x + xoff;
    float* __restrict__ Gg = p.g + xoff;
    float* __restrict__ Sh = p.s_hist + (long long)m * xoff;
    float* __restrict__ Yh = p.y_hist + (long long)m * xoff;

    EvalLds<float> L = carve_eval_lds<float>(smem, N, MODE == MODE_LUT ? 0 : G.NU, false, BLOCK / WAVE, true);
    L.P = G.P;
    char* vb = smem + eval_lds_bytes<float>(N, MODE == MODE_LUT ? 0 : G.NU, false, BLOCK / WAVE, true);
    float* x = reinterpret_cast<float*>(vb);   // [np] each, zero beyond n
    float* g = x + np;
    float* gp = g + np;
    float* d = gp + np;
    double* SY = reinterpret_cast<double*>(d + np);   // [m][m] s_a . y_b by physical slot
    double* YY = SY + m * m;                            // [m][m] y_a . y_b
    Hist<HL> H;
    H.n = n;
    if (HL) {
        H.s = reinterpret_cast<float*>(reinterpret_cast<char*>(SY) + align_up(sizeof(double) * 2 * m * m, 16));
        H.y = H.s + m * np;
        H.stride = np;
    } else {
        H.s = Sh;
        H.y = Yh;
        H.stride = n;
    }
    for (int k = threadIdx.x; k < np; k += BLOCK) {
        const bool in = k < n;
        x[k] = in ? X[k] : 0.f;
        g[k] = in ? Gg[k] : 0.f;
        gp[k] = 0.f;
        d[k] = 0.f;
    }
    if (HL) {
        const bool run = S.status == ST_RUN;
        for (int q = threadIdx.x; q < m * np; q += BLOCK) {
            const int r = q / np, k = q - r * np;
            const bool in = run && k < n;
            H.s[q] = in ? Sh[(long long)r * n + k] : 0.f;
            H.y[q] = in ? Yh[(long long)r * n + k] : 0.f;
        }
    }
    __syncthreads();
    long long st_tp = 0, st_pp = 0;
    const bool w0 = threadIdx.x < WAVE;
    const int lane = threadIdx.x;
    // Neighbour list: pairs within cutoff + skin at the rebuild.  Two trees approach by at
    // most twice the largest displacement, so the list stays complete while
    // 2 * vdisp < skin; it is rebuilt at 0.9 skin (margin for rounding) and at every call.
    const float skin = list_skin(G);
    const float cut_list = (sqrtf(G.cutoff_sq) + skin) * (sqrtf(G.cutoff_sq) + skin);
    if (threadIdx.x == 0) {
        nb_valid = 0;
        vdisp = 0.0;
    }
    if (w0 && S.status == ST_RUN && S.hist_len > 0) lb_rebuild<KH, HL>(np, S.hist_len, S.head, m, H, SY, YY);

    if (S.status == ST_FRESH) {
        const double loss = relax_eval<MODE, BLOCK>(G, L, A, x, g, N, bacc, st_tp, st_pp, 1, cut_list);
        if (w0) {
            if (lane == 0) nb_valid = L.wcnt[8] <= L.nb_cap;
            double gm = 0.0;
            for (int k = lane; k < np; k += WAVE) gm = fmax(gm, fabs((double)g[k]));
            gm = wave_max(gm);
            if (lane == 0) {
                S.loss = loss;
                S.loss0 = loss;
                S.H = 1.0;
                S.n_eval = 1;
                S.n_iter = 0;
                S.hist_len = 0;
                S.head = 0;
                S.reason = 0;
                S.status = (gm <= (double)p.tol_grad) ? ST_DONE : ST_RUN;
                if (S.status == ST_DONE) S.reason = R_GRAD;
            }
        }
        __syncthreads();
    }

    for (int it = 0; it < p.iters; ++it) {
        if (S.status == ST_DONE) break;
        // ---- direction, step length, directional-derivative test, step (wave 0) ----
        if (w0) {
            const int niter = S.n_iter + 1;
            const int hl = niter > 1 ? S.hist_len : 0;
            const int head = S.head;
            float cs_l = 0.f, cy_l = 0.f;
            if (hl > 0) {
                const float r = lb_dots<KH, HL>(np, hl, head, m, H, [&](int k) { return g[k]; });
                Lb<KH>::solve(r, hl, head, m, S.H, SY, YY, S.ro, cs_l, cy_l);
            }
            // iteration 1: d = -g; later: d = -H_diag g + history terms
            const float gamf = niter > 1 ? (float)S.H : 1.f;
            double gtd, gsum, dmax, xmax;
            lb_direction<KH, HL>(np, hl, head, m, gamf, H, g, x, d, cs_l, cy_l, gtd, gsum, dmax, xmax);
            double t = niter == 1 ? fmin(1.0, 1.0 / gsum) * (double)p.lr : (double)p.lr;
            // step clamp: max |t d| <= max_step
            const double pc = dmax * fabs(t);
            if (p.max_step > 0.f && pc > (double)p.max_step) t *= (double)p.max_step / pc;
            const bool es = gtd > -(double)p.tol_change;
            if (!es) {
                const float tf = (float)t;
                for (int k = lane; k < np; k += WAVE) {
                    const float dk = d[k];
                    gp[k] = g[k];
                    x[k] = x[k] + tf * dk;
                }
            }
            if (lane == 0) {
                // largest coordinate change |t d|, plus the rounding of the float update (half an
                // ulp of the largest coordinate, doubled); a centre moves by at most sqrt(2)
                // times that
                const double step = dmax * fabs(t);
                const double xm = xmax + step;
                vdisp += 1.4142136 * (step * (1.0 + 1e-6) + xm * 1.2e-7);
                eval_mode = (nb_valid && 2.0 * vdisp < 0.9 * (double)skin) ? 2 : 1;
                if (eval_mode == 1) vdisp = 0.0;
                t_step = t;
                stop = es;
                if (es) {
                    S.status = ST_DONE;
                    S.reason |= R_DIR;
                } else {
                    S.n_iter = niter;
                }
            }
        }
        __syncthreads();
        if (stop) break;
        const double loss_prev = S.loss;
        double loss = loss_prev;
        const bool last = S.n_iter == p.max_iter;
        if (!last) {
            const int lm = eval_mode;
            loss = relax_eval<MODE, BLOCK>(G, L, A, x, g, N, bacc, st_tp, st_pp, lm, cut_list);
            if (threadIdx.x == 0 && lm == 1) nb_valid = L.wcnt[8] <= L.nb_cap;
        }
        // ---- history update and convergence tests (wave 0) ----
        if (w0) {
            const double t = t_step;
            const float tf = (float)t;
            const int niter = S.n_iter;
            const int hl = S.hist_len, head = S.head;
            // history pairs are stored from iteration 2 on (core/lbfgs_torch_parallel.py:636-690)
            double mu[2] = {0.0, 0.0};
            const double rp = lb_pair_sums(np, [&](int k, float (&a)[2]) {
                const float gk = g[k];
                const float sk = d[k] * tf;
                Lb<KH>::pair(a, gk - gp[k], sk);
                mu[0] = fmax(mu[0], fabs((double)gk));
                mu[1] = fmax(mu[1], fabs((double)sk));
            });
            const double rmu = wave_max_many<2>(mu);
            const double ys = lane_d(rp, many_src<2>(0)), yy = lane_d(rp, many_src<2>(1));
            const double gm = lane_d(rmu, many_src<2>(0)), pc = lane_d(rmu, many_src<2>(1));
            if (niter > 1 && ys > 1e-10) {
                const float r = lb_dots<KH, HL>(np, hl, head, m, H, [&](int k) { return g[k] - gp[k]; });
                const int q = hl < m ? ring(head, hl, m) : head;
                const int oq = q * H.stride;
                for (int k = lane; k < np; k += WAVE) {
                    if (HL || k < n) {
                        const float gk = g[k];
                        H.y[oq + k] = gk - gp[k];
                        H.s[oq + k] = d[k] * tf;
                    }
                }
                Lb<KH>::store(r, ys, yy, hl, head, m, q, SY, YY, S);
            }
            if (lane == 0) {
                S.n_eval += 1;
                S.loss = loss;
                int rr = 0;
                if (gm <= (double)p.tol_grad) rr |= R_GRAD;
                if (S.n_eval >= p.max_eval) rr |= R_EVAL;
                if (pc <= (double)p.tol_change) rr |= R_PARAM;
                const double lc = fabs(loss - loss_prev);
                if (lc < (double)p.tol_change) rr |= R_LOSS;
                if (p.tol_rel > 0.f && niter > m && lc / (fabs(loss_prev) + 1e-12) < (double)p.tol_rel)
                    rr |= R_REL;
                if (p.stop_on_increase && loss > S.loss0) rr |= R_INCREASE;
                if (niter >= p.max_iter) rr |= R_ITER;
                if (rr) {
                    S.status = ST_DONE;
                    S.reason |= rr;
                }
            }
        }
        __syncthreads();
    }
    // ---- write back ----
    for (int k = threadIdx.x; k < n; k += BLOCK) {
        X[k] = x[k];
        Gg[k] = g[k];
    }
    if (HL) {
        for (int q = threadIdx.x; q < m * n; q += BLOCK) {
            const int r = q / n, k = q - r * n;
            Sh[q] = H.s[r * np + k];
            Yh[q] = H.y[r * np + k];
        }
    }
    if (threadIdx.x == 0) {
        *gS = S;
        if (p.status_out) p.status_out[slot] = S.status == ST_DONE ? 1 : 0;
        if (p.loss_out) p.loss_out[slot] = (float)S.loss;
    }
    if (p.stats && threadIdx.x % WAVE == 0) {
        atomicAdd(&p.stats[0], (unsigned long long)st_tp);
        atomicAdd(&p.stats[1], (unsigned long long)st_pp);
    }
}

// Waves per SIMD the register kernels are compiled for: two for up to 5 rows per wave (the
// register budget of 256 then holds the state, the history and the evaluator), one above.
template <int RW>
constexpr int wave_occupancy() { return RW <= 5 ? 2 : 1; }
// Table lookups per lane in flight at once: four for a wave holding 9 or 10 rows (one wave
// per SIMD whatever the choice), two below, where the registers saved let a second wave
// share the SIMD.
template <int RW>
constexpr int wave_lookups() { return RW >= 9 ? 4 : 2; }
// Register rows per wave of k_relax_w: all R rows with one wave; with two, wave 0 holds
// group A (rows 0-4) and wave 1 group B (rows 5-9, zero beyond the state).
template <int R, int NWV>
constexpr int rows_per_wave() { return NWV == 1 ? R : ROWS_A; }
// Occupancy a register kernel is compiled for: that of its rows per wave, except the
// exact-area evaluator from 4 rows per wave, whose clipping state does not fit the
// two-wave register budget next to the history (it runs one wave per SIMD, without
// spills).
template <int MODE, int R, int NWV>
constexpr int kernel_occupancy() {
    return (MODE == MODE_AREA && rows_per_wave<R, NWV>() >= 4) ? 1 : wave_occupancy<rows_per_wave<R, NWV>()>();
}

// Totals exchanged between the two waves of a k_relax_w workgroup: each wave writes its
// group's butterfly totals (one word per value and lane), one barrier, then reads the
// other wave's.  Two buffers are used in turn, so one barrier per exchange suffices: a
// buffer is written again only after the barrier of the next exchange, which the other
// wave reaches only after its reads of this one.
constexpr int XW = 6;   // words per lane and exchange
struct Xch {
    float* buf;    // [2 buffers][2 waves][XW][64]
    int par, nput, nget;
    __device__ inline float* at(int wv, int i) const {
        return buf + ((par * 2 + wv) * XW + i) * WAVE + __lane_id();
    }
    __device__ inline void put(float v) { *at(threadIdx.x / WAVE, nput++) = v; }
    __device__ inline void put(double v) {
        put(__int_as_float(__double2loint(v)));
        put(__int_as_float(__double2hiint(v)));
    }
    __device__ inline void sync() { __syncthreads(); }
    __device__ inline float getf() { return *at(1 - (int)(threadIdx.x / WAVE), nget++); }
    __device__ inline double getd() {
        const int lo = __float_as_int(getf());
        const int hi = __float_as_int(getf());
        return __hiloint2double(hi, lo);
    }
    __device__ inline void done() {
        par ^= 1;
        nput = nget = 0;
    }
};

// Group totals of K float partial lists, a (group A, or this wave's group) and b (group B
// of a one-wave kernel): one butterfly each, added A + B.  Two waves: a's total is this
// wave's share; the caller exchanges and adds.
template <int NWV, int RW, int K>
__device__ inline float group_sum_f(float (&a)[K], float (&b)[K]) {
    const float ra = wave_sum_many_f<K>(a);
    if (NWV == 2) return ra;
    const float rb = RW > ROWS_A ? wave_sum_many_f<K>(b) : 0.f;
    return ra + rb;
}
template <int NWV, int RW, typename V, int K>
__device__ inline double group_sum_d(V (&a)[K], V (&b)[K]) {
    const double ra = wave_sum_many<K>(a);
    if (NWV == 2) return ra;
    const double rb = RW > ROWS_A ? wave_sum_many<K>(b) : 0.0;
    return ra + rb;
}
// Row r of this wave belongs to group B (one-wave kernel, r >= 5).
template <int NWV>
__device__ constexpr bool in_b(int r) { return NWV == 1 && r >= ROWS_A; }

// Evaluation of the register state x into g for k_relax_w (through the LDS vectors xs, gs;
// this wave's rows start at row0).
template <int MODE, int RW, int NWV>
__device__ __forceinline__ double wave_evaluate(const GeomT<float>& G, EvalLds<float>& L, const RelaxArgs& A,
                                                const float (&x)[RW], float (&g)[RW], float* xs, float* gs,
                                                int row0, int N, int n, unsigned long long* bacc, long long& st_tp,
                                                long long& st_pp, int lm, float cut_list, int& nb_valid) {
    const int lane = threadIdx.x % WAVE;
#pragma unroll
    for (int r = 0; r < RW; ++r) xs[lane + WAVE * (row0 + r)] = x[r];
    __syncthreads();
    const double loss = relax_eval<MODE, WAVE * NWV, wave_lookups<RW>(), false>(G, L, A, xs, gs, N, bacc,
                                                                                st_tp, st_pp, lm, cut_list);
    // relax_eval ends with a barrier after the last write of gs and of the boundary sums.
    // Each lane forms its own elements: the overlap gradient from the fixed-point sums (read
    // in element order, no bank conflicts) plus the boundary gradient of near trees, left
    // in gs (zero elsewhere; cleared again here), and the h gradient for element 3N + j.
    const int n3 = n - 3;
    const pk_relax_params& p = A.p;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int k = lane + WAVE * (row0 + r);
        if (k < n3) {
            const float gb = gs[k];
            g[r] = p.w_overlap * Acc<float>::get(L.g[k]) + p.w_boundary * gb;
            gs[k] = 0.f;
        } else {
            g[r] = k < n ? h_gradient(p, bacc, k - n3, x[r]) : 0.f;
        }
    }
    if (lm == 1) nb_valid = L.wcnt[8] <= L.nb_cap;
    return loss;
}

// One or two waves per instance, with the state vectors (x, g, g_prev, d) and the L-BFGS
// history in registers: element k of a state vector is row k / 64 of lane k % 64 (rows
// beyond the state are zero); one wave holds all R rows, or two waves hold group A (rows
// 0-4) and group B (rows 5-9) of an instance of up to 212 trees, halving each wave's
// registers so that two waves share a SIMD.  History row i = ring slot i (a new pair
// overwrites one slot; unused slots are zero).  Only the evaluator's tables and the small
// matrices live in LDS.  Every sum follows the canonical group order, so all variants
// and k_relax give the same bits and may continue each other's runs.  With two waves
// both compute the small-matrix solve and every scalar decision from the same exchanged
// totals; wave 0 alone writes the optimizer record and the matrices.
template <int MODE, int R, int KH, int NWV>
__global__ __launch_bounds__(WAVE * NWV, (kernel_occupancy<MODE, R, NWV>())) void k_relax_w(RelaxArgs A) {
    constexpr int RW = rows_per_wave<R, NWV>();
    constexpr int RT = RW * NWV;   // rows of the LDS staging vectors
    static_assert(NWV == 1 || NWV == 2, "one or two waves per instance");
    static_assert(NWV == 1 || R <= 2 * ROWS_A, "two waves hold at most 10 rows");
    extern __shared__ __align__(16) char smem[];
    __shared__ GeomT<float> G;
    __shared__ LbSlot S;
    __shared__ unsigned long long bacc[4];
    const pk_relax_params& p = A.p;
    const int b = blockIdx.x;
    if (b >= p.n_launch) return;
    const int slot = p.order ? p.order[b] : b;
    const int N = p.ntree[slot];
    // LDS, the row count and the tile were sized for n_tree_max: a larger count is refused
    // (status -1) rather than run past them
    if (N <= 0 || N > p.n_tree_max) {
        if (N > 0 && threadIdx.x == 0 && p.status_out) p.status_out[slot] = -1;
        return;
    }
    LbSlot* gS = reinterpret_cast<LbSlot*>(p.state) + slot;
    const int lane = NWV == 1 ? (int)threadIdx.x : (int)threadIdx.x % WAVE;
    const int wave = NWV == 1 ? 0 : (int)threadIdx.x / WAVE;
    const int row0 = wave * RW;
    const bool lead = threadIdx.x == 0;   // writes the optimizer record
    if (lead) S = *gS;
    load_geom(G, A.geom);
    __syncthreads();
    if (S.status == ST_DONE) return;

    const int n = 3 * N + 3, m = p.history;
    const long long xoff = p.xoff[slot];
    float* __restrict__ X = p.x + xoff;
    float* __restrict__ Gg = p.g + xoff;
    float* __restrict__ Sh = p.s_hist + (long long)m * xoff;
    float* __restrict__ Yh = p.y_hist + (long long)m * xoff;

    const int Pe = MODE == MODE_LUT ? 0 : G.NU;
    EvalLds<float> L = carve_eval_lds<float>(smem, N, Pe, false, NWV, true);
    L.P = G.P;
    char* vb = smem + eval_lds_bytes<float>(N, Pe, false, NWV, true);
    float* xs = reinterpret_cast<float*>(vb);      // [64 RT] evaluator input
    float* gs = xs + WAVE * RT;                    // [64 RT] evaluator output
    double* SY = reinterpret_cast<double*>(gs + WAVE * RT);
    double* YY = SY + m * m;
    Xch X2;
    X2.buf = reinterpret_cast<float*>(YY + m * m);
    X2.par = X2.nput = X2.nget = 0;

    float x[RW], g[RW], gp[RW], d[RW];
    float hs[KH][RW], hy[KH][RW];
    const bool run0 = S.status == ST_RUN;
    const int hl0 = run0 ? S.hist_len : 0, head0 = S.head;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int k = lane + WAVE * (row0 + r);
        const bool in = k < n;
        x[r] = in ? X[k] : 0.f;
        g[r] = in ? Gg[k] : 0.f;
        gp[r] = 0.f;
        d[r] = 0.f;
        gs[k] = 0.f;
#pragma unroll
        for (int i = 0; i < KH; ++i) {
            const bool use = in && i < hl0;
            const long long o = use ? (long long)i * n + k : 0;
            hs[i][r] = use ? Sh[o] : 0.f;
            hy[i][r] = use ? Yh[o] : 0.f;
        }
    }
    // rebuild S^T Y and Y^T Y from the stored history, one ring slot (column) at a time
#pragma unroll
    for (int pc = 0; pc < KH; ++pc) {
        if (pc < hl0) {
            float aa[Lb<KH>::KD], ab[Lb<KH>::KD];
#pragma unroll
            for (int i = 0; i < Lb<KH>::KD; ++i) aa[i] = ab[i] = 0.f;
            float pa[2] = {0.f, 0.f}, pb[2] = {0.f, 0.f};
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                float sv[KH], yv[KH];
#pragma unroll
                for (int i = 0; i < KH; ++i) {
                    sv[i] = hs[i][r];
                    yv[i] = hy[i][r];
                }
                if (in_b<NWV>(r)) {
                    Lb<KH>::dots(ab, sv, yv, hy[pc][r], KH);
                    Lb<KH>::pair(pb, hy[pc][r], hs[pc][r]);
                } else {
                    Lb<KH>::dots(aa, sv, yv, hy[pc][r], KH);
                    Lb<KH>::pair(pa, hy[pc][r], hs[pc][r]);
                }
            }
            float rr = group_sum_f<NWV, RW>(aa, ab);
            double rp = group_sum_d<NWV, RW>(pa, pb);
            if (NWV == 2) {
                X2.put(rr);
                X2.put(rp);
                X2.sync();
                rr = rr + X2.getf();
                rp = rp + X2.getd();
                X2.done();
            }
            if (wave == 0)
                Lb<KH>::rebuild_col_slot(rr, lane_d(rp, many_src<2>(0)), lane_d(rp, many_src<2>(1)), pc, hl0,
                                         head0, m, SY, YY);
        }
    }
    if (NWV == 2) __syncthreads();   // the matrices wave 0 wrote, before either wave reads them

    long long st_tp = 0, st_pp = 0;
    const float skin = list_skin(G);
    const float cut_list = (sqrtf(G.cutoff_sq) + skin) * (sqrtf(G.cutoff_sq) + skin);
    int nb_valid = 0;
    double vdisp = 0.0;

    if (S.status == ST_FRESH) {
        const double loss = wave_evaluate<MODE, RW, NWV>(G, L, A, x, g, xs, gs, row0, N, n, bacc, st_tp, st_pp,
                                                         1, cut_list, nb_valid);
        double gm = 0.0;
#pragma unroll
        for (int r = 0; r < RW; ++r) gm = fmax(gm, fabs((double)g[r]));
        gm = wave_max(gm);
        if (NWV == 2) {
            X2.put(gm);
            X2.sync();
            gm = fmax(gm, X2.getd());
            X2.done();
        }
        if (lead) {
            S.loss = loss;
            S.loss0 = loss;
            S.H = 1.0;
            S.n_eval = 1;
            S.n_iter = 0;
            S.hist_len = 0;
            S.head = 0;
            S.reason = 0;
            S.status = (gm <= (double)p.tol_grad) ? ST_DONE : ST_RUN;
            if (S.status == ST_DONE) S.reason = R_GRAD;
        }
        __syncthreads();
    }

    // History rows live in registers by ring slot (row i of hs, hy = slot i; slots at or
    // beyond the used count are zero), so a new pair overwrites one slot and nothing moves.
    // s_i.g and y_i.g of this lane's chronological history row (lane mod KH) for the next
    // direction: from the stored history here, then from the pass after each evaluation.
    const int li = lane & (KH - 1);
    float a_l = 0.f, b_l = 0.f;
    // rows of this lane that hold a tree centre coordinate (x or y, not theta or h):
    // k = lane + 64 row with k < 3N and k mod 3 != 2 (64 = 1 mod 3)
    unsigned cmask = 0;
#pragma unroll
    for (int r = 0; r < RW; ++r)
        if (lane + WAVE * (row0 + r) < 3 * N && (lane + row0 + r) % 3 != 2) cmask |= 1u << r;
    if (S.status == ST_RUN && hl0 > 0) {
        float aa[Lb<KH>::KD], ab[Lb<KH>::KD];
#pragma unroll
        for (int i = 0; i < Lb<KH>::KD; ++i) aa[i] = ab[i] = 0.f;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            float sv[KH], yv[KH];
#pragma unroll
            for (int i = 0; i < KH; ++i) {
                sv[i] = hs[i][r];
                yv[i] = hy[i][r];
            }
            if (in_b<NWV>(r))
                Lb<KH>::dots(ab, sv, yv, g[r], KH);
            else
                Lb<KH>::dots(aa, sv, yv, g[r], KH);
        }
        float rr = group_sum_f<NWV, RW>(aa, ab);
        if (NWV == 2) {
            X2.put(rr);
            X2.sync();
            rr = rr + X2.getf();
            X2.done();
        }
        const int ps = ring(head0, li < m ? li : 0, m);
        a_l = __shfl(rr, many_src<Lb<KH>::KD>(ps));
        b_l = __shfl(rr, many_src<Lb<KH>::KD>(KH + ps));
    }

    for (int it = 0; it < p.iters; ++it) {
        if (S.status == ST_DONE) break;
        // ---- direction, step length, directional-derivative test, step ----
        // (every read of the record precedes the exchange below, every write follows it)
        const int niter = S.n_iter + 1;
        const int hl = niter > 1 ? S.hist_len : 0;
        const int head = S.head;
        float cs_l = 0.f, cy_l = 0.f;
        if (hl > 0) Lb<KH>::solve_ab(a_l, b_l, hl, head, m, S.H, SY, YY, S.ro, cs_l, cy_l);
        const float gamf = niter > 1 ? (float)S.H : 1.f;
        const float cs_s = Lb<KH>::to_slot(cs_l, head, m), cy_s = Lb<KH>::to_slot(cy_l, head, m);
        float cs[KH], cy[KH];
#pragma unroll
        for (int i = 0; i < KH; ++i) {
            cs[i] = lane_f(cs_s, i);
            cy[i] = lane_f(cy_s, i);
        }
        double sa[2] = {0.0, 0.0}, sb[2] = {0.0, 0.0};
        float mv[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            float sv[KH], yv[KH];
#pragma unroll
            for (int i = 0; i < KH; ++i) {
                sv[i] = hs[i][r];
                yv[i] = hy[i][r];
            }
            const float gk = g[r];
            const float dk = Lb<KH>::dir(gamf, gk, cs, cy, sv, yv);
            d[r] = dk;
            if (in_b<NWV>(r)) {
                sb[0] += (double)gk * (double)dk;
                sb[1] += fabs((double)gk);
            } else {
                sa[0] += (double)gk * (double)dk;
                sa[1] += fabs((double)gk);
            }
            mv[0] = fmaxf(mv[0], fabsf(dk));
            const bool c = (cmask >> r) & 1u;
            mv[1] = fmaxf(mv[1], c ? fabsf(dk) : 0.f);
            mv[2] = fmaxf(mv[2], c ? fabsf(x[r]) : 0.f);
        }
        double rs = group_sum_d<NWV, RW>(sa, sb);
        float rm = wave_max_many_f<4>(mv);
        if (NWV == 2) {
            X2.put(rs);
            X2.put(rm);
            X2.sync();
            rs = rs + X2.getd();
            rm = fmaxf(rm, X2.getf());
            X2.done();
        }
        const double gtd = lane_d(rs, many_src<2>(0)), gsum = lane_d(rs, many_src<2>(1));
        const double dmax = (double)lane_f(rm, many_src<4>(0));
        const double dcen = (double)lane_f(rm, many_src<4>(1)), xcen = (double)lane_f(rm, many_src<4>(2));
        double t = niter == 1 ? fmin(1.0, 1.0 / gsum) * (double)p.lr : (double)p.lr;
        const double pc0 = dmax * fabs(t);
        if (p.max_step > 0.f && pc0 > (double)p.max_step) t *= (double)p.max_step / pc0;
        const bool es = gtd > -(double)p.tol_change;
        if (es) {
            if (lead) {
                S.status = ST_DONE;
                S.reason |= R_DIR;
            }
            __syncthreads();
            break;
        }
        const float tf = (float)t;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            gp[r] = g[r];
            x[r] = x[r] + tf * d[r];
        }
        // neighbour-list validity as in k_relax, from the centre coordinates only (theta and
        // h steps move no centre)
        const double step = dcen * fabs(t);
        vdisp += 1.4142136 * (step * (1.0 + 1e-6) + (xcen + step) * 1.2e-7);
        int lm = (nb_valid && 2.0 * vdisp < 0.9 * (double)skin) ? 2 : 1;
        if (lm == 1) vdisp = 0.0;
        // (S.n_iter is read again only at the next iteration, after its barriers)
        if (lead) S.n_iter = niter;
        const double loss_prev = S.loss;
        double loss = loss_prev;
        if (niter != p.max_iter)
            loss = wave_evaluate<MODE, RW, NWV>(G, L, A, x, g, xs, gs, row0, N, n, bacc, st_tp, st_pp, lm,
                                                cut_list, nb_valid);
        // read before the exchange below, after which wave 0 may update the record
        const int hl_now = S.hist_len;
        // ---- history update and convergence tests ----
        // One pass over the history rows i in use gives s_i.g and y_i.g for the next
        // direction and s_i.y_new and y_i.y_new for the new S^T Y column and Y^T Y row; the
        // new pair's own products s.g, y.g, y.s, y.y; and max |g|, max |s|.  Each sum has
        // the per-lane partials and the reduction tree of the separate passes of k_relax.
        constexpr int KF = 4 * KH;
        float fa[KF], fb[KF], na[4] = {0.f, 0.f, 0.f, 0.f}, nb[4] = {0.f, 0.f, 0.f, 0.f};
        float pa[2] = {0.f, 0.f}, pb[2] = {0.f, 0.f}, mu[2] = {0.f, 0.f};
#pragma unroll
        for (int i = 0; i < KF; ++i) fa[i] = fb[i] = 0.f;
        float yn[RW], sn[RW];
        auto post_row = [&](int r, float (&acc)[KF], float (&an)[4], float (&pp)[2]) {
            const float gk = g[r];
            yn[r] = gk - gp[r];
            sn[r] = d[r] * tf;
            Lb<KH>::pair(pp, yn[r], sn[r]);
            mu[0] = fmaxf(mu[0], fabsf(gk));
            mu[1] = fmaxf(mu[1], fabsf(sn[r]));
#pragma unroll
            for (int i = 0; i < KH; ++i) {
                acc[i] = fmaf(hs[i][r], gk, acc[i]);
                acc[KH + i] = fmaf(hy[i][r], gk, acc[KH + i]);
                acc[2 * KH + i] = fmaf(hs[i][r], yn[r], acc[2 * KH + i]);
                acc[3 * KH + i] = fmaf(hy[i][r], yn[r], acc[3 * KH + i]);
            }
            an[0] = fmaf(sn[r], gk, an[0]);
            an[1] = fmaf(yn[r], gk, an[1]);
        };
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            if (in_b<NWV>(r))
                post_row(r, fb, nb, pb);
            else
                post_row(r, fa, na, pa);
        }
        float rf = group_sum_f<NWV, RW>(fa, fb);
        float rn = group_sum_f<NWV, RW>(na, nb);
        double rp = group_sum_d<NWV, RW>(pa, pb);
        float rmu = wave_max_many_f<2>(mu);
        if (NWV == 2) {
            X2.put(rf);
            X2.put(rn);
            X2.put(rp);
            X2.put(rmu);
            X2.sync();
            rf = rf + X2.getf();
            rn = rn + X2.getf();
            rp = rp + X2.getd();
            rmu = fmaxf(rmu, X2.getf());
            X2.done();
        }
        const double ys = lane_d(rp, many_src<2>(0)), yy = lane_d(rp, many_src<2>(1));
        const double gm = (double)lane_f(rmu, many_src<2>(0)), pc = (double)lane_f(rmu, many_src<2>(1));
        const bool acc_ok = niter > 1 && ys > 1e-10;
        const bool full = hl_now >= m;
        const int q = full ? head : hl_now;   // slot of the new pair (the oldest when full)
        if (acc_ok) {
            const int pl = ring(head, li < m ? li : 0, m);
            if (wave == 0)
                Lb<KH>::store_v(__shfl(rf, many_src<KF>(2 * KH + pl)), __shfl(rf, many_src<KF>(3 * KH + pl)), ys,
                                yy, hl_now, head, m, q, SY, YY, S);
#pragma unroll
            for (int i = 0; i < KH; ++i) {
                if (i == q) {
#pragma unroll
                    for (int r = 0; r < RW; ++r) {
                        hs[i][r] = sn[r];
                        hy[i][r] = yn[r];
                    }
                }
            }
        }
        {
            // chronological row li of the history the next direction sees, by ring slot
            const int head_n = acc_ok && full ? ring(head, 1, m) : head;
            const int ps = ring(head_n, li < m ? li : 0, m);
            const bool is_new = acc_ok && ps == q;
            const float an_s = lane_f(rn, many_src<4>(0)), an_y = lane_f(rn, many_src<4>(1));
            const float ao = __shfl(rf, many_src<KF>(ps)), bo = __shfl(rf, many_src<KF>(KH + ps));
            a_l = is_new ? an_s : ao;
            b_l = is_new ? an_y : bo;
        }
        if (lead) {
            S.n_eval += 1;
            S.loss = loss;
            int rr = 0;
            if (gm <= (double)p.tol_grad) rr |= R_GRAD;
            if (S.n_eval >= p.max_eval) rr |= R_EVAL;
            if (pc <= (double)p.tol_change) rr |= R_PARAM;
            const double lc = fabs(loss - loss_prev);
            if (lc < (double)p.tol_change) rr |= R_LOSS;
            if (p.tol_rel > 0.f && niter > m && lc / (fabs(loss_prev) + 1e-12) < (double)p.tol_rel)
                rr |= R_REL;
            if (p.stop_on_increase && loss > S.loss0) rr |= R_INCREASE;
            if (niter >= p.max_iter) rr |= R_ITER;
            if (rr) {
                S.status = ST_DONE;
                S.reason |= rr;
            }
        }
        __syncthreads();
    }
    // ---- write back ----
    const int hl_f = S.hist_len;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int k = lane + WAVE * (row0 + r);
        if (k < n) {
            X[k] = x[r];
            Gg[k] = g[r];
#pragma unroll
            for (int i = 0; i < KH; ++i) {
                if (i < hl_f) {
                    const long long o = (long long)i * n + k;
                    Sh[o] = hs[i][r];
                    Yh[o] = hy[i][r];
                }
            }
        }
    }
    if (lead) {
        *gS = S;
        if (p.status_out) p.status_out[slot] = S.status == ST_DONE ? 1 : 0;
        if (p.loss_out) p.loss_out[slot] = (float)S.loss;
    }
    if (p.stats && lane == 0) {
        atomicAdd(&p.stats[0], (unsigned long long)st_tp);
        atomicAdd(&p.stats[1], (unsigned long long)st_pp);
    }
}

// Dynamic LDS of k_relax_w: evaluator (NWV waves), evaluator input/output vectors, small
// matrices, and (two waves) the exchange buffers.
size_t relax_w_lds_bytes(int mode, int n_max, int NV, int m, int R, int nwv) {
    const int Pe = mode == MODE_LUT ? 0 : NV;
    const int rt = nwv == 2 ? 2 * ROWS_A : R;
    size_t b = eval_lds_bytes<float>(n_max, Pe, false, nwv, true);
    b += sizeof(float) * 2 * WAVE * (size_t)rt;
    b += sizeof(double) * 2 * (size_t)m * m;
    if (nwv == 2) b += sizeof(float) * 2 * 2 * XW * WAVE;
    return align_up(b, 16);
}

// Occupancy query (pk_relax_occupancy): while set, the launch helpers below store the
// resident workgroups per CU of the kernel and LDS size they would launch, and launch nothing.
static thread_local int* t_occupancy = nullptr;

template <int MODE, int BLOCK, bool HL>
static int launch_relax(const Geometry* g, const RelaxArgs& A, size_t lds, hipStream_t st) {
    const void* kern = A.p.history <= 8 ? reinterpret_cast<const void*>(k_relax<MODE, BLOCK, HL, 8>)
                                        : reinterpret_cast<const void*>(k_relax<MODE, BLOCK, HL, 16>);
    PK_CHECK_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (t_occupancy) {
        PK_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(t_occupancy, kern, BLOCK, lds));
        return 0;
    }
    if (A.p.history <= 8)
        hipLaunchKernelGGL((k_relax<MODE, BLOCK, HL, 8>), dim3(A.p.n_launch), dim3(BLOCK), lds, st, A);
    else
        hipLaunchKernelGGL((k_relax<MODE, BLOCK, HL, 16>), dim3(A.p.n_launch), dim3(BLOCK), lds, st, A);
    PK_CHECK_HIP(hipGetLastError());
    return 0;
}

template <int MODE, int BLOCK>
static int launch_relax(const Geometry* g, const RelaxArgs& A, hipStream_t st) {
    int dev = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    int lim = 0;
    PK_CHECK_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t stat = sizeof(GeomT<float>) + sizeof(LbSlot) + 256;
    const int NV = g->hf.NU, m = A.p.history, nmax = A.p.n_tree_max;
    const size_t lds_g = relax_lds_bytes(MODE, nmax, NV, m, false, BLOCK);
    const size_t lds_l = relax_lds_bytes(MODE, nmax, NV, m, true, BLOCK);
    const bool fits_l = lds_l + stat <= (size_t)lim;
    if (A.p.hist_lds == 1 && !fits_l) {
        set_error("pk_relax_step: history of " + std::to_string(nmax) + " trees does not fit in LDS");
        return PK_ERANGE;
    }
    if (lds_g + stat > (size_t)lim) {
        set_error("pk_relax_step: instance of " + std::to_string(nmax) + " trees needs " +
                  std::to_string(lds_g + stat) + " B of LDS, device allows " + std::to_string(lim));
        return PK_ERANGE;
    }
    if (A.p.hist_lds != 0 && fits_l) return launch_relax<MODE, BLOCK, true>(g, A, lds_l, st);
    return launch_relax<MODE, BLOCK, false>(g, A, lds_g, st);
}

// Register kernels: rows of the state vector per lane, and the largest supported count.
constexpr int W_RMAX = 10;
constexpr int W_KH = 8;   // history rows held in registers
inline int wave_rows(int n_tree_max) { return (3 * n_tree_max + 3 + WAVE - 1) / WAVE; }

template <int MODE, int R, int KH, int NWV>
static int launch_relax_w_rk(const Geometry* g, const RelaxArgs& A, hipStream_t st) {
    int dev = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    int lim = 0;
    PK_CHECK_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t lds = relax_w_lds_bytes(MODE, A.p.n_tree_max, g->hf.NU, A.p.history, R, NWV);
    const size_t stat = sizeof(GeomT<float>) + sizeof(LbSlot) + 64;
    if (lds + stat > (size_t)lim) {
        set_error("pk_relax_step: instance of " + std::to_string(A.p.n_tree_max) + " trees needs " +
                  std::to_string(lds + stat) + " B of LDS, device allows " + std::to_string(lim));
        return PK_ERANGE;
    }
    auto kern = k_relax_w<MODE, R, KH, NWV>;
    PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (t_occupancy) {
        PK_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(t_occupancy, reinterpret_cast<const void*>(kern),
                                                                  WAVE * NWV, lds));
        return 0;
    }
    hipLaunchKernelGGL(kern, dim3(A.p.n_launch), dim3(WAVE * NWV), lds, st, A);
    PK_CHECK_HIP(hipGetLastError());
    return 0;
}

template <int MODE, int R>
static int launch_relax_w_r(const Geometry* g, const RelaxArgs& A, hipStream_t st) {
    return launch_relax_w_rk<MODE, R, W_KH, 1>(g, A, st);
}

// nwv 1: one wave holding R rows; 2: two waves holding 5 rows each (any R <= 10).
template <int MODE>
static int launch_relax_w(const Geometry* g, const RelaxArgs& A, int rows, int nwv, hipStream_t st) {
    if (nwv == 2) return launch_relax_w_rk<MODE, 2 * ROWS_A, W_KH, 2>(g, A, st);
    switch (rows) {
        case 1: return launch_relax_w_r<MODE, 1>(g, A, st);
        case 2: return launch_relax_w_r<MODE, 2>(g, A, st);
        case 3: return launch_relax_w_r<MODE, 3>(g, A, st);
        case 4: return launch_relax_w_r<MODE, 4>(g, A, st);
        case 5: return launch_relax_w_r<MODE, 5>(g, A, st);
        case 6: return launch_relax_w_r<MODE, 6>(g, A, st);
        case 7: return launch_relax_w_r<MODE, 7>(g, A, st);
        case 8: return launch_relax_w_r<MODE, 8>(g, A, st);
        case 9: return launch_relax_w_r<MODE, 9>(g, A, st);
        default: return launch_relax_w_r<MODE, 10>(g, A, st);
    }
}

LutDev make_lut_dev(const void* lut_handle);

// Kernel choice and launch shared by pk_relax_step and pk_relax_run_noalloc.
static int relax_dispatch(const Geometry* g, const RelaxArgs& A, int block, hipStream_t st) {
    const pk_relax_params* p = &A.p;
    const int rows = wave_rows(p->n_tree_max);
    int blk = block;
    const bool wave_ok = rows <= W_RMAX && p->history <= W_KH;
    // library's choice: registers where they fit, one wave up to 3 rows and two above
    // (measured on MI355X in table mode: two waves 1.15x faster at 5 rows, 1.37x at 10;
    // one wave faster at 3)
    if (blk == 0) blk = wave_ok ? (rows <= 3 ? 1 : 2) : 256;
    if (blk == 1 || blk == 2) {
        if (!wave_ok) {
            set_error("pk_relax_step: blocks 1 and 2 (register state, one or two waves per instance) "
                      "support up to " + std::to_string((W_RMAX * WAVE - 3) / 3) + " trees and history " +
                      std::to_string(W_KH));
            return PK_ERANGE;
        }
        if (p->mode == PK_MODE_AREA) return launch_relax_w<MODE_AREA>(g, A, rows, blk, st);
        if (p->mode == PK_MODE_SEP) return launch_relax_w<MODE_SEP>(g, A, rows, blk, st);
        return launch_relax_w<MODE_LUT>(g, A, rows, blk, st);
    }
#define PK_RL(M)                                                          \
    switch (blk) {                                                        \
        case 64: return launch_relax<M, 64>(g, A, st);                    \
        case 128: return launch_relax<M, 128>(g, A, st);                  \
        case 256: return launch_relax<M, 256>(g, A, st);                  \
        default: set_error("pk_relax_step: block must be 0, 1, 2, 64, 128 or 256"); \
            return PK_EINVAL;                                             \
    }
    if (p->mode == PK_MODE_AREA) { PK_RL(MODE_AREA) }
    if (p->mode == PK_MODE_SEP) { PK_RL(MODE_SEP) }
    PK_RL(MODE_LUT)
#undef PK_RL
}

// ---------------------------------------------------------------------------------
// Allocation-free relaxation: the optimizer records, gradients and L-BFGS history of one
// batch live in a static device array (one copy per GPU, held until the process ends).
// Host code never passes the array's name to a kernel: its device address is taken once
// per device with hipGetSymbolAddress and kept in scratch_dev.
// ---------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PK_RELAX_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex scratch_mu;
static void* scratch_dev[SCRATCH_MAX_DEVICES] = {};

static int scratch_address(void** out) {
    int dev = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        set_error("pk_relax_run_noalloc: device index out of range");
        return PK_EINVAL;
    }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_dev[dev] == nullptr) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr) {
            set_error(std::string("pk_relax_run_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return PK_EHIP;
        }
        scratch_dev[dev] = ptr;
    }
    *out = scratch_dev[dev];
    return PK_OK;
}

// Fresh optimizer records for the launched slots (order[b], b < n_launch; order NULL: b).
__global__ __launch_bounds__(256) void k_fresh_slots(LbSlot* __restrict__ rec, const int* __restrict__ order,
                                                     int n_launch) {
    constexpr int W = sizeof(LbSlot) / sizeof(uint32_t);
    const long long total = (long long)n_launch * W;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(q / W), w = (int)(q % W);
        const int s = order ? order[b] : b;
        reinterpret_cast<uint32_t*>(rec + s)[w] = 0u;
    }
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace pk

using namespace pk;

extern "C" {

int pk_relax_slot_bytes(void) { return (int)sizeof(LbSlot); }

int pk_relax_occupancy(pk_geometry_t h, int mode, int n_tree_max, int history, int hist_lds, int block,
                       int* blocks_per_cu) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g != nullptr && blocks_per_cu != nullptr, "pk_relax_occupancy: NULL argument");
    PK_REQUIRE(mode == PK_MODE_AREA || mode == PK_MODE_SEP || mode == PK_MODE_LUT, "pk_relax_occupancy: bad mode");
    PK_REQUIRE(history >= 1 && history <= PK_MAX_HISTORY, "pk_relax_occupancy: 1 <= history <= 16");
    PK_REQUIRE(n_tree_max >= 1 && n_tree_max <= 4096, "pk_relax_occupancy: bad n_tree_max");
    PK_REQUIRE(hist_lds >= -1 && hist_lds <= 1, "pk_relax_occupancy: hist_lds must be -1, 0 or 1");
    RelaxArgs A;
    std::memset(&A, 0, sizeof(A));
    A.p.mode = mode;
    A.p.n_tree_max = n_tree_max;
    A.p.history = history;
    A.p.hist_lds = hist_lds;
    A.geom = g->df;
    int nb = 0;
    t_occupancy = &nb;
    const int rc = relax_dispatch(g, A, block, nullptr);
    t_occupancy = nullptr;
    if (rc != PK_OK) return rc;
    *blocks_per_cu = nb;
    return PK_OK;
}

int pk_relax_lds_bytes(pk_geometry_t h, int mode, int n_tree_max, int history, int hist_lds, int block) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    if (!g || n_tree_max < 1 || history < 1) return PK_EINVAL;
    if (block == 1 || block == 2)
        return (int)relax_w_lds_bytes(mode, n_tree_max, g->hf.NU, history, wave_rows(n_tree_max), block);
    if (block != 64 && block != 128 && block != 256) return PK_EINVAL;
    return (int)relax_lds_bytes(mode, n_tree_max, g->hf.NU, history, hist_lds != 0, block);
}

int pk_relax_step(pk_geometry_t h, const pk_relax_params* p, void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g != nullptr && p != nullptr, "pk_relax_step: NULL argument");
    PK_REQUIRE(p->mode == PK_MODE_AREA || p->mode == PK_MODE_SEP || p->mode == PK_MODE_LUT,
               "pk_relax_step: bad mode");
    PK_REQUIRE(p->history >= 1 && p->history <= PK_MAX_HISTORY, "pk_relax_step: 1 <= history <= 16");
    PK_REQUIRE(p->n_tree_max >= 1 && p->n_tree_max <= 4096, "pk_relax_step: bad n_tree_max");
    PK_REQUIRE(p->iters >= 0 && p->max_iter >= 0, "pk_relax_step: negative iteration count");
    PK_REQUIRE(p->mode != PK_MODE_LUT || p->lut != nullptr, "pk_relax_step: table mode needs a table");
    PK_REQUIRE(p->hist_lds >= -1 && p->hist_lds <= 1, "pk_relax_step: hist_lds must be -1, 0 or 1");
    if (p->n_launch <= 0 || p->iters == 0) return PK_OK;
    PK_REQUIRE(p->xoff && p->ntree && p->x && p->g && p->s_hist && p->y_hist && p->state,
               "pk_relax_step: missing state array");
    RelaxArgs A;
    A.p = *p;
    if (A.p.max_eval <= 0) A.p.max_eval = p->max_iter * 5 / 4;
    A.geom = g->df;
    std::memset(&A.lut, 0, sizeof(A.lut));
    if (p->mode == PK_MODE_LUT) A.lut = make_lut_dev(p->lut);
    return relax_dispatch(g, A, p->block, reinterpret_cast<hipStream_t>(stream));
}

int pk_relax_run_noalloc(pk_geometry_t h, const pk_relax_params* p, int n_state, void* stream) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    PK_REQUIRE(g != nullptr && p != nullptr, "pk_relax_run_noalloc: NULL argument");
    PK_REQUIRE(p->mode == PK_MODE_AREA || p->mode == PK_MODE_SEP || p->mode == PK_MODE_LUT,
               "pk_relax_run_noalloc: bad mode");
    PK_REQUIRE(p->history >= 1 && p->history <= PK_MAX_HISTORY, "pk_relax_run_noalloc: 1 <= history <= 16");
    PK_REQUIRE(p->n_tree_max >= 1 && p->n_tree_max <= 4096, "pk_relax_run_noalloc: bad n_tree_max");
    PK_REQUIRE(p->max_iter >= 0, "pk_relax_run_noalloc: negative iteration count");
    PK_REQUIRE(p->mode != PK_MODE_LUT || p->lut != nullptr, "pk_relax_run_noalloc: table mode needs a table");
    PK_REQUIRE(p->hist_lds >= -1 && p->hist_lds <= 1, "pk_relax_run_noalloc: hist_lds must be -1, 0 or 1");
    PK_REQUIRE(p->g == nullptr && p->s_hist == nullptr && p->y_hist == nullptr && p->state == nullptr,
               "pk_relax_run_noalloc: g, s_hist, y_hist and state live in the library's scratch; pass NULL");
    PK_REQUIRE(p->n_slots >= 0 && n_state >= 0 && p->n_launch <= p->n_slots,
               "pk_relax_run_noalloc: bad n_slots, n_launch or n_state");
    const size_t rec_b = round256((size_t)p->n_slots * sizeof(LbSlot));
    const size_t vec_b = round256((size_t)n_state * sizeof(float));
    const size_t his_b = round256((size_t)p->history * (size_t)n_state * sizeof(float));
    const size_t need = rec_b + vec_b + 2 * his_b;
    if (need > (size_t)PK_RELAX_SCRATCH_BYTES) {
        set_error("pk_relax_run_noalloc: the batch needs " + std::to_string(need) +
                  " B of scratch (records, gradient, history), the library holds " +
                  std::to_string((size_t)PK_RELAX_SCRATCH_BYTES) + " B; split it or use pk_relax_step");
        return PK_ERANGE;
    }
    if (p->n_launch <= 0 || p->max_iter == 0) return PK_OK;
    PK_REQUIRE(p->xoff && p->ntree && p->x, "pk_relax_run_noalloc: missing x, xoff or ntree");
    void* base = nullptr;
    const int rc = scratch_address(&base);
    if (rc != PK_OK) return rc;
    char* b = static_cast<char*>(base);
    RelaxArgs A;
    A.p = *p;
    A.p.state = b;
    A.p.g = reinterpret_cast<float*>(b + rec_b);
    A.p.s_hist = reinterpret_cast<float*>(b + rec_b + vec_b);
    A.p.y_hist = reinterpret_cast<float*>(b + rec_b + vec_b + his_b);
    A.p.iters = p->max_iter;   // fresh slots, run to the end
    if (A.p.max_eval <= 0) A.p.max_eval = p->max_iter * 5 / 4;
    A.geom = g->df;
    std::memset(&A.lut, 0, sizeof(A.lut));
    if (p->mode == PK_MODE_LUT) A.lut = make_lut_dev(p->lut);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int blocks = std::max(1, std::min((p->n_launch * (int)(sizeof(LbSlot) / 4) + 255) / 256, 1024));
    hipLaunchKernelGGL(k_fresh_slots, dim3(blocks), dim3(256), 0, st, reinterpret_cast<LbSlot*>(b), p->order,
                       p->n_launch);
    PK_CHECK_HIP(hipGetLastError());
    return relax_dispatch(g, A, p->block, st);
}

}  // extern "C"
