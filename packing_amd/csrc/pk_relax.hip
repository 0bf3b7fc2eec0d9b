// Fused relaxation engine: cost evaluation + L-BFGS iterations of one packing instance
// per workgroup, for a whole (ragged) batch in one launch per chunk of iterations.
//
// Objective (core/pack_dynamics.py:45-52,76-123 with fixed h): state x = [xyt (N,3) | h (3)],
//   cost = w_ov * overlap(area | sep | table) + w_bd * boundary + w_area * h0^2 (free h only);
// gradient w.r.t. h zeroed when h is fixed (core/pack_cost.py:178-188).
// Optimizer: core/lbfgs_torch_parallel.py:411-729 without line search, applied to each
// instance on its own (a batch of one), so an instance's result never depends on what
// else is in the launch: two-loop recursion, first step lr*min(1, 1/sum|g|), max-step
// clamp, directional-derivative stop, history append/shift only from iteration 2 with
// y.s > 1e-10, and the reference's convergence tests.
//
// All vector reductions run on wave 0 in a fixed order (lane-strided partial sums in
// double, then a fixed DPP / lane-read tree), so results are independent of the
// workgroup size.  The history lives in LDS for the duration of a call when it fits.
#include <cstring>
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

struct F2 {
    float a, b;
};
struct F3 {
    float a, b, c;
};

// Wave-0 vector loop over k = lane + 64 r, r = 0 .. np/64 - 1 increasing (per-lane sums keep
// a fixed order), in blocks of WB elements per lane whose loads are issued before any of
// the block's stores: load(k) -> value, then op(k, value).  np is a multiple of 64; the
// vectors are zero beyond their true length.
constexpr int WB = 8;
template <typename LoadF, typename OpF>
__device__ inline void w0_for(int np, LoadF load, OpF op) {
    const int lane = threadIdx.x;
    const int R = np / WAVE;
    using V = decltype(load(0));
    for (int r0 = 0; r0 < R; r0 += WB) {
        V v[WB];
#pragma unroll
        for (int u = 0; u < WB; ++u)
            if (r0 + u < R) v[u] = load(lane + WAVE * (r0 + u));
#pragma unroll
        for (int u = 0; u < WB; ++u)
            if (r0 + u < R) op(lane + WAVE * (r0 + u), v[u]);
    }
}

__host__ __device__ inline int pad64(int n) { return (n + WAVE - 1) / WAVE * WAVE; }

// Cost and gradient of the state x into g (all threads); returns the loss.  list_mode:
// 1 rebuilds the neighbour list while evaluating, 2 evaluates from it.
template <int MODE, int BLOCK>
__device__ inline double relax_eval(const GeomT<float>& G, EvalLds<float>& L, const RelaxArgs& A,
                                    const float* x, float* g, int N,
                                    unsigned long long* bacc /* [4] boundary accumulators */,
                                    long long& st_tp, long long& st_pp, int list_mode, float cut_list) {
    const pk_relax_params& p = A.p;
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        L.x[i] = x[3 * i];
        L.y[i] = x[3 * i + 1];
        L.t[i] = x[3 * i + 2];
    }
    if (threadIdx.x < 4) bacc[threadIdx.x] = 0ull;
    if (threadIdx.x == 0 && list_mode == 1) L.wcnt[8] = 0;
    __syncthreads();
    stage_instance<float, BLOCK, MODE != MODE_LUT, true>(G, L);
    Evaluator<float, MODE, true, false, BLOCK> ev(G, L);
    ev.cax = ev.cay = ev.cbx = ev.cby = 0.f;
    ev.npd = 0;
    ev.only_self = 0;
    ev.lut = A.lut;
    ev.cut_list = cut_list;
    ev.run(list_mode);
    __syncthreads();
    st_tp += ev.n_tp;
    st_pp += ev.n_pp;
    const float half = x[3 * N] * 0.5f, hx = x[3 * N + 1], hy = x[3 * N + 2];
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        float gb[3], gh[3];
        const float cb = boundary_tree<float>(G, L.x[i], L.y[i], L.c[i], L.s[i], half, hx, hy, gb, gh);
        Acc<float>::add(&bacc[0], cb);
        Acc<float>::add(&bacc[1], gh[0]);
        Acc<float>::add(&bacc[2], gh[1]);
        Acc<float>::add(&bacc[3], gh[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            g[3 * i + k] = p.w_overlap * Acc<float>::get(L.g[3 * i + k]) + p.w_boundary * gb[k];
    }
    __syncthreads();
    const double c_ov = Acc<float>::get(*L.cost), c_bd = Acc<float>::get(bacc[0]);
    double loss = (double)p.w_overlap * c_ov + (double)p.w_boundary * c_bd;
    if (threadIdx.x == 0) {
        if (p.fixed_h) {
            g[3 * N] = g[3 * N + 1] = g[3 * N + 2] = 0.f;
        } else {
            g[3 * N] = p.w_boundary * Acc<float>::get(bacc[1]) + 2.f * p.w_area * x[3 * N];
            g[3 * N + 1] = p.w_boundary * Acc<float>::get(bacc[2]);
            g[3 * N + 2] = p.w_boundary * Acc<float>::get(bacc[3]);
        }
    }
    if (!p.fixed_h) loss += (double)p.w_area * (double)x[3 * N] * (double)x[3 * N];
    __syncthreads();
    return loss;
}

// Dynamic LDS of one workgroup: evaluator, x g g_prev d, and (hl) the history s, y; the
// vectors padded to a multiple of 64 floats.
size_t relax_lds_bytes(int mode, int n_max, int P, int m, bool hl, int block) {
    const int Pe = mode == MODE_LUT ? 0 : P;
    size_t b = eval_lds_bytes<float>(n_max, Pe, false, block / WAVE, true);
    const size_t np = (size_t)pad64(3 * n_max + 3);
    b += sizeof(float) * 4 * np;
    if (hl) b += sizeof(float) * 2 * (size_t)m * np;
    return align_up(b, 16);
}

// HL: the history s_hist / y_hist is copied into LDS at the start of the call and back at
// the end; otherwise the two-loop recursion reads it from global memory (row stride n
// there, padded stride np in LDS).
template <int MODE, int BLOCK, bool HL>
__global__ __launch_bounds__(BLOCK) void k_relax(RelaxArgs A) {
    extern __shared__ __align__(16) char smem[];
    __shared__ GeomT<float> G;
    __shared__ LbSlot S;
    __shared__ double al[PK_MAX_HISTORY];
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
    if (N <= 0) return;
    LbSlot* gS = reinterpret_cast<LbSlot*>(p.state) + slot;
    if (threadIdx.x == 0) S = *gS;
    load_geom(G, A.geom);
    __syncthreads();
    if (S.status == ST_DONE) return;

    const int n = 3 * N + 3, np = pad64(n), m = p.history;
    const int hstride = HL ? np : n;   // history row stride
    const long long xoff = p.xoff[slot];
    float* __restrict__ X = p.x + xoff;
    float* __restrict__ Gg = p.g + xoff;
    float* __restrict__ Sh = p.s_hist + (long long)m * xoff;
    float* __restrict__ Yh = p.y_hist + (long long)m * xoff;

    EvalLds<float> L = carve_eval_lds<float>(smem, N, MODE == MODE_LUT ? 0 : G.P, false, BLOCK / WAVE, true);
    L.P = G.P;
    char* vb = smem + eval_lds_bytes<float>(N, MODE == MODE_LUT ? 0 : G.P, false, BLOCK / WAVE, true);
    float* x = reinterpret_cast<float*>(vb);   // [np] each, zero beyond n
    float* g = x + np;
    float* gp = g + np;
    float* d = gp + np;
    float* hs = d + np;          // LDS history (HL): [m][np] s, then [m][np] y
    float* hy = hs + m * np;
    if (!HL) {
        hs = Sh;
        hy = Yh;
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
            hs[q] = in ? Sh[(long long)r * n + k] : 0.f;
            hy[q] = in ? Yh[(long long)r * n + k] : 0.f;
        }
    }
    __syncthreads();
    // history element k of row ph (zero beyond n)
    auto hrow = [&](const float* base, int ph, int k) -> float {
        if (HL) return base[ph * hstride + k];
        return k < n ? base[ph * hstride + k] : 0.f;
    };
    long long st_tp = 0, st_pp = 0;
    const bool w0 = threadIdx.x < WAVE;
    const int lane = threadIdx.x;
    // Neighbour list: pairs within cutoff + skin at the rebuild.  Two trees approach by at
    // most twice the largest displacement, so the list stays complete while
    // 2 * vdisp < skin; it is rebuilt at 0.9 skin (margin for rounding) and at every call.
    const float skin = 0.5f * G.radius;
    const float cut_list = (sqrtf(G.cutoff_sq) + skin) * (sqrtf(G.cutoff_sq) + skin);
    if (threadIdx.x == 0) {
        nb_valid = 0;
        vdisp = 0.0;
    }

    if (S.status == ST_FRESH) {
        const double loss = relax_eval<MODE, BLOCK>(G, L, A, x, g, N, bacc, st_tp, st_pp, 1, cut_list);
        if (w0) {
            if (lane == 0) nb_valid = L.wcnt[8] <= L.nb_cap;
            double gm = 0.0;
            w0_for(np, [&](int k) { return g[k]; }, [&](int, float v) { gm = fmax(gm, fabs((double)v)); });
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
        // ---- direction, step length, directional-derivative test (wave 0) ----
        // Two-loop recursion with each axpy fused with the next dot product (same values
        // and summation order as separate passes):
        //   q = -g; for i = hl-1..0: a_i = ro_i s_i.q, q -= a_i y_i;  r = H q;
        //   for i = 0..hl-1: b_i = ro_i y_i.r, r += (a_i - b_i) s_i;  d = r.
        if (w0) {
            const int niter = S.n_iter + 1;
            const int hl = niter > 1 ? S.hist_len : 0;
            const float hd = (float)S.H;
            const int head = S.head;
            double acc = 0.0;
            if (hl > 0) {
                const int ps = (head + hl - 1) % m;
                w0_for(np, [&](int k) { return F2{g[k], hrow(hs, ps, k)}; }, [&](int k, F2 v) {
                    const float q = -v.a;
                    d[k] = q;
                    acc += (double)v.b * (double)q;
                });
            } else if (niter > 1) {
                w0_for(np, [&](int k) { return g[k]; }, [&](int k, float v) { d[k] = (-v) * hd; });
            } else {
                w0_for(np, [&](int k) { return g[k]; }, [&](int k, float v) { d[k] = -v; });
            }
            for (int i = hl - 1; i >= 0; --i) {
                const int ph = (head + i) % m;
                const double a_i = wave_sum(acc) * S.ro[ph];
                al[i] = a_i;   // every lane writes the same value and reads its own write
                const float af = (float)a_i;
                acc = 0.0;
                if (i > 0) {
                    const int pn = (head + i - 1) % m;
                    w0_for(np, [&](int k) { return F3{d[k], hrow(hy, ph, k), hrow(hs, pn, k)}; },
                           [&](int k, F3 v) {
                               const float q = v.a - v.b * af;
                               d[k] = q;
                               acc += (double)v.c * (double)q;
                           });
                } else {
                    // last step of the first loop, the H scaling and the first dot of the
                    // second loop (y of slot i = 0 both times)
                    w0_for(np, [&](int k) { return F2{d[k], hrow(hy, ph, k)}; }, [&](int k, F2 v) {
                        const float q = v.a - v.b * af;
                        const float r = q * hd;
                        d[k] = r;
                        acc += (double)v.b * (double)r;
                    });
                }
            }
            for (int i = 0; i < hl; ++i) {
                const int ph = (head + i) % m;
                const double be = wave_sum(acc) * S.ro[ph];
                const float c = (float)(al[i] - be);
                acc = 0.0;
                if (i + 1 < hl) {
                    const int pn = (head + i + 1) % m;
                    w0_for(np, [&](int k) { return F3{d[k], hrow(hs, ph, k), hrow(hy, pn, k)}; },
                           [&](int k, F3 v) {
                               const float r = v.a + v.b * c;
                               d[k] = r;
                               acc += (double)v.c * (double)r;
                           });
                } else {
                    w0_for(np, [&](int k) { return F2{d[k], hrow(hs, ph, k)}; },
                           [&](int k, F2 v) { d[k] = v.a + v.b * c; });
                }
            }
            double t;
            if (niter == 1) {
                double gs = 0.0;
                w0_for(np, [&](int k) { return g[k]; }, [&](int, float v) { gs += fabs((double)v); });
                gs = wave_sum(gs);
                t = fmin(1.0, 1.0 / gs) * (double)p.lr;
            } else {
                t = (double)p.lr;
            }
            // step clamp and directional derivative in one pass
            const double t_pre = t;
            double pc = 0.0, gtd = 0.0;
            w0_for(np, [&](int k) { return F2{g[k], d[k]}; }, [&](int, F2 v) {
                pc = fmax(pc, fabs((double)v.b * t));
                gtd += (double)v.a * (double)v.b;
            });
            gtd = wave_sum(gtd);
            pc = wave_max(pc);
            if (p.max_step > 0.f && pc > (double)p.max_step) t *= (double)p.max_step / pc;
            const bool es = gtd > -(double)p.tol_change;
            double xm = 0.0;
            if (!es) {
                const float tf = (float)t;
                w0_for(np, [&](int k) { return F3{x[k], d[k], g[k]}; }, [&](int k, F3 v) {
                    gp[k] = v.c;
                    const float xn = v.a + tf * v.b;
                    x[k] = xn;
                    xm = fmax(xm, fabs((double)xn));
                });
                xm = wave_max(xm);
            }
            if (lane == 0) {
                // largest coordinate change: pc scaled to the final t, plus the rounding of
                // the float update (half an ulp of the largest coordinate, doubled); a
                // centre moves by at most sqrt(2) times that
                const double step = pc * (t / t_pre);
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
            double ys = 0.0, yy = 0.0, gm = 0.0, pc = 0.0;
            w0_for(np, [&](int k) { return F3{g[k], gp[k], d[k]}; }, [&](int, F3 v) {
                const double yk = (double)(v.a - v.b);
                const double sk = (double)(v.c * tf);
                ys += yk * sk;
                yy += yk * yk;
                gm = fmax(gm, fabs((double)v.a));
                pc = fmax(pc, fabs(sk));
            });
            gm = wave_max(gm);
            pc = wave_max(pc);
            if (niter > 1) {
                ys = wave_sum(ys);
                yy = wave_sum(yy);
                if (ys > 1e-10) {
                    const int slotp = S.hist_len < m ? (S.head + S.hist_len) % m : S.head;
                    float* so = hs + slotp * hstride;
                    float* yo = hy + slotp * hstride;
                    w0_for(np, [&](int k) { return F3{g[k], gp[k], d[k]}; }, [&](int k, F3 v) {
                        if (HL || k < n) {
                            yo[k] = v.a - v.b;
                            so[k] = v.c * tf;
                        }
                    });
                    if (lane == 0) {
                        S.H = ys / yy;
                        S.ro[slotp] = 1.0 / ys;
                        if (S.hist_len < m) S.hist_len++;
                        else S.head = (S.head + 1) % m;
                    }
                }
            }
            if (lane == 0) {
                S.n_eval += 1;
                S.loss = loss;
                int r = 0;
                if (gm <= (double)p.tol_grad) r |= R_GRAD;
                if (S.n_eval >= p.max_eval) r |= R_EVAL;
                if (pc <= (double)p.tol_change) r |= R_PARAM;
                const double lc = fabs(loss - loss_prev);
                if (lc < (double)p.tol_change) r |= R_LOSS;
                if (p.tol_rel > 0.f && niter > m && lc / (fabs(loss_prev) + 1e-12) < (double)p.tol_rel)
                    r |= R_REL;
                if (p.stop_on_increase && loss > S.loss0) r |= R_INCREASE;
                if (niter >= p.max_iter) r |= R_ITER;
                if (r) {
                    S.status = ST_DONE;
                    S.reason |= r;
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
            Sh[q] = hs[r * np + k];
            Yh[q] = hy[r * np + k];
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

template <int MODE, int BLOCK, bool HL>
static int launch_relax(const Geometry* g, const RelaxArgs& A, size_t lds, hipStream_t st) {
    auto kern = k_relax<MODE, BLOCK, HL>;
    PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(A.p.n_launch), dim3(BLOCK), lds, st, A);
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
    const int P = g->hf.P, m = A.p.history, nmax = A.p.n_tree_max;
    const size_t lds_g = relax_lds_bytes(MODE, nmax, P, m, false, BLOCK);
    const size_t lds_l = relax_lds_bytes(MODE, nmax, P, m, true, BLOCK);
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

LutDev make_lut_dev(const void* lut_handle);

}  // namespace pk

using namespace pk;

extern "C" {

int pk_relax_slot_bytes(void) { return (int)sizeof(LbSlot); }

int pk_relax_lds_bytes(pk_geometry_t h, int mode, int n_tree_max, int history, int hist_lds, int block) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    if (!g || n_tree_max < 1 || history < 1) return PK_EINVAL;
    if (block != 64 && block != 128 && block != 256) return PK_EINVAL;
    return (int)relax_lds_bytes(mode, n_tree_max, g->hf.P, history, hist_lds != 0, block);
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
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int blk = p->block ? p->block : 256;
#define PK_RL(M)                                                          \
    switch (blk) {                                                        \
        case 64: return launch_relax<M, 64>(g, A, st);                    \
        case 128: return launch_relax<M, 128>(g, A, st);                  \
        case 256: return launch_relax<M, 256>(g, A, st);                  \
        default: set_error("pk_relax_step: block must be 64, 128 or 256"); \
            return PK_EINVAL;                                             \
    }
    if (p->mode == PK_MODE_AREA) { PK_RL(MODE_AREA) }
    if (p->mode == PK_MODE_SEP) { PK_RL(MODE_SEP) }
    PK_RL(MODE_LUT)
#undef PK_RL
}

}  // extern "C"
