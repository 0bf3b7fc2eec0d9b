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
// double, then an xor butterfly), so results are independent of the workgroup size.
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

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// lane-strided dot product over wave 0; every lane gets lane 0's value
__device__ inline double w0_dot(const float* a, const float* b, int n) {
    const int lane = threadIdx.x;
    double acc = 0.0;
    for (int k = lane; k < n; k += WAVE) acc += (double)a[k] * (double)b[k];
    acc = wave_sum(acc);
    return __shfl(acc, 0);
}

template <int MODE, int BLOCK>
__device__ inline double relax_eval(const GeomT<float>& G, EvalLds<float>& L, const RelaxArgs& A,
                                    const float* x, float* g, int N,
                                    unsigned long long* bacc /* [4] boundary accumulators */,
                                    long long& st_tp, long long& st_pp) {
    const pk_relax_params& p = A.p;
    for (int i = threadIdx.x; i < N; i += BLOCK) {
        L.x[i] = x[3 * i];
        L.y[i] = x[3 * i + 1];
        L.t[i] = x[3 * i + 2];
    }
    if (threadIdx.x < 4) bacc[threadIdx.x] = 0ull;
    __syncthreads();
    stage_instance<float, BLOCK, MODE != MODE_LUT>(G, L);
    Evaluator<float, MODE, true, false, BLOCK> ev(G, L);
    ev.cax = ev.cay = ev.cbx = ev.cby = 0.f;
    ev.npd = 0;
    ev.only_self = 0;
    ev.lut = A.lut;
    ev.run();
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

size_t relax_lds_bytes(int mode, int n_max, int P) {
    const int Pe = mode == MODE_LUT ? 0 : P;
    size_t b = eval_lds_bytes<float>(n_max, Pe);
    b = align_up(b, 16);
    b += sizeof(float) * 4 * (size_t)(3 * n_max + 3);
    return align_up(b, 16);
}

template <int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_relax(RelaxArgs A) {
    extern __shared__ __align__(16) char smem[];
    __shared__ GeomT<float> G;
    __shared__ LbSlot S;
    __shared__ double al[PK_MAX_HISTORY];
    __shared__ unsigned long long bacc[4];
    __shared__ int stop;
    __shared__ double t_step;
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

    const int n = 3 * N + 3, m = p.history;
    const long long xoff = p.xoff[slot];
    float* __restrict__ X = p.x + xoff;
    float* __restrict__ Gg = p.g + xoff;
    float* __restrict__ Sh = p.s_hist + (long long)m * xoff;
    float* __restrict__ Yh = p.y_hist + (long long)m * xoff;

    EvalLds<float> L = carve_eval_lds<float>(smem, N, MODE == MODE_LUT ? 0 : G.P);
    L.P = G.P;
    char* vb = smem + align_up(eval_lds_bytes<float>(N, MODE == MODE_LUT ? 0 : G.P), 16);
    float* x = reinterpret_cast<float*>(vb);
    float* g = x + n;
    float* gp = g + n;
    float* d = gp + n;
    for (int k = threadIdx.x; k < n; k += BLOCK) {
        x[k] = X[k];
        g[k] = Gg[k];
    }
    __syncthreads();
    long long st_tp = 0, st_pp = 0;
    const bool w0 = threadIdx.x < WAVE;
    const int lane = threadIdx.x;

    if (S.status == ST_FRESH) {
        const double loss = relax_eval<MODE, BLOCK>(G, L, A, x, g, N, bacc, st_tp, st_pp);
        if (w0) {
            double gm = 0.0;
            for (int k = lane; k < n; k += WAVE) gm = fmax(gm, fabs((double)g[k]));
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
        if (w0) {
            const int niter = S.n_iter + 1;
            if (niter == 1) {
                for (int k = lane; k < n; k += WAVE) d[k] = -g[k];
            } else {
                for (int k = lane; k < n; k += WAVE) d[k] = -g[k];
                const int hl = S.hist_len;
                for (int i = hl - 1; i >= 0; --i) {
                    const int ph = (S.head + i) % m;
                    const float* si = Sh + (long long)ph * n;
                    const float* yi = Yh + (long long)ph * n;
                    const double a_i = w0_dot(si, d, n) * S.ro[ph];
                    al[i] = a_i;   // every lane writes the same value and reads its own write
                    const float af = (float)a_i;
                    for (int k = lane; k < n; k += WAVE) d[k] = d[k] - yi[k] * af;
                }
                const float hd = (float)S.H;
                for (int k = lane; k < n; k += WAVE) d[k] = d[k] * hd;
                for (int i = 0; i < hl; ++i) {
                    const int ph = (S.head + i) % m;
                    const float* si = Sh + (long long)ph * n;
                    const float* yi = Yh + (long long)ph * n;
                    const double be = w0_dot(yi, d, n) * S.ro[ph];
                    const float c = (float)(al[i] - be);
                    for (int k = lane; k < n; k += WAVE) d[k] = d[k] + si[k] * c;
                }
            }
            double t;
            if (niter == 1) {
                double gs = 0.0;
                for (int k = lane; k < n; k += WAVE) gs += fabs((double)g[k]);
                gs = __shfl(wave_sum(gs), 0);
                t = fmin(1.0, 1.0 / gs) * (double)p.lr;
            } else {
                t = (double)p.lr;
            }
            if (p.max_step > 0.f) {
                double pc = 0.0;
                for (int k = lane; k < n; k += WAVE) pc = fmax(pc, fabs((double)d[k] * t));
                pc = __shfl(wave_max(pc), 0);
                if (pc > (double)p.max_step) t *= (double)p.max_step / pc;
            }
            const double gtd = w0_dot(g, d, n);
            const bool es = gtd > -(double)p.tol_change;
            if (!es) {
                const float tf = (float)t;
                for (int k = lane; k < n; k += WAVE) {
                    gp[k] = g[k];
                    x[k] = x[k] + tf * d[k];
                }
            }
            if (lane == 0) {
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
        if (!last) loss = relax_eval<MODE, BLOCK>(G, L, A, x, g, N, bacc, st_tp, st_pp);
        // ---- history update and convergence tests (wave 0) ----
        if (w0) {
            const double t = t_step;
            const float tf = (float)t;
            const int niter = S.n_iter;
            if (niter > 1) {
                double ys = 0.0, yy = 0.0;
                for (int k = lane; k < n; k += WAVE) {
                    const double yk = (double)(g[k] - gp[k]);
                    const double sk = (double)(d[k] * tf);
                    ys += yk * sk;
                    yy += yk * yk;
                }
                ys = __shfl(wave_sum(ys), 0);
                yy = __shfl(wave_sum(yy), 0);
                if (ys > 1e-10) {
                    int slotp;
                    if (S.hist_len < m) {
                        slotp = (S.head + S.hist_len) % m;
                    } else {
                        slotp = S.head;
                    }
                    float* so = Sh + (long long)slotp * n;
                    float* yo = Yh + (long long)slotp * n;
                    for (int k = lane; k < n; k += WAVE) {
                        yo[k] = g[k] - gp[k];
                        so[k] = d[k] * tf;
                    }
                    if (lane == 0) {
                        S.H = ys / yy;
                        S.ro[slotp] = 1.0 / ys;
                        if (S.hist_len < m) S.hist_len++;
                        else S.head = (S.head + 1) % m;
                    }
                }
            }
            double gm = 0.0, pc = 0.0;
            for (int k = lane; k < n; k += WAVE) {
                gm = fmax(gm, fabs((double)g[k]));
                pc = fmax(pc, fabs((double)(d[k] * tf)));
            }
            gm = __shfl(wave_max(gm), 0);
            pc = __shfl(wave_max(pc), 0);
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
    if (threadIdx.x == 0) {
        *gS = S;
        if (p.status_out) p.status_out[slot] = S.status == ST_DONE ? 1 : 0;
        if (p.loss_out) p.loss_out[slot] = (float)S.loss;
        if (p.stats) {
            atomicAdd(&p.stats[0], (unsigned long long)st_tp);
            atomicAdd(&p.stats[1], (unsigned long long)st_pp);
        }
    }
}

template <int MODE, int BLOCK>
static int launch_relax(const Geometry* g, const RelaxArgs& A, hipStream_t st) {
    auto kern = k_relax<MODE, BLOCK>;
    const size_t lds = relax_lds_bytes(MODE, A.p.n_tree_max, g->hf.P);
    int dev = 0;
    PK_CHECK_HIP(hipGetDevice(&dev));
    int lim = 0;
    PK_CHECK_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t stat = sizeof(GeomT<float>) + sizeof(LbSlot) + 256;
    if (lds + stat > (size_t)lim) {
        set_error("pk_relax_step: instance of " + std::to_string(A.p.n_tree_max) + " trees needs " +
                  std::to_string(lds + stat) + " B of LDS, device allows " + std::to_string(lim));
        return PK_ERANGE;
    }
    PK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(A.p.n_launch), dim3(BLOCK), lds, st, A);
    PK_CHECK_HIP(hipGetLastError());
    return 0;
}

LutDev make_lut_dev(const void* lut_handle);

}  // namespace pk

using namespace pk;

extern "C" {

int pk_relax_slot_bytes(void) { return (int)sizeof(LbSlot); }

int pk_relax_lds_bytes(pk_geometry_t h, int mode, int n_tree_max) {
    const Geometry* g = reinterpret_cast<const Geometry*>(h);
    if (!g) return PK_EINVAL;
    return (int)relax_lds_bytes(mode, n_tree_max, g->hf.P);
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
