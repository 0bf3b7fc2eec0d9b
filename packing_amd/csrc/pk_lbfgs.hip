// Device-side batched L-BFGS for an arbitrary objective evaluated by the caller between
// launches: the drop-in for core/lbfgs_torch_parallel.py:411-729 (no line search) with
// its batch-coupled semantics.  One wave per system.  Per iteration the host issues
//   pk_lbfgs_direction(it)  history update + convergence tests of iteration it-1, then
//                           the two-loop direction, step length, max-step clamp and the
//                           directional-derivative test of iteration it;
//   pk_lbfgs_step(it)       x += t d for every system, unless no system was active after
//                           the convergence tests or after the directional test (the
//                           reference's loop exit / break), in which case the run stops;
//   func(x)                 (skipped on the last iteration, as the reference does).
// Stop decisions are counters in device memory, so the host never waits on the GPU.
#include <cstring>
#include <string>

#include "pk_common.h"
#include "packing.h"

namespace pk {

struct LbSys {
    double H, t, loss_prev, loss0;
    int hist_len, head, active, n_eval;   // n_eval: objective evaluations (line search)
};
static_assert(sizeof(LbSys) == PK_LBFGS_SYS_BYTES, "system record layout");

// min / max that propagate NaN, as torch.minimum / maximum / clamp / Tensor.max do (the
// line search meets NaN by design: its bracketing passes the same point twice to the
// cubic, giving 0 * inf), so that every test sees the value the reference sees.
__device__ inline double nmin(double a, double b) { return (a != a || b != b) ? a + b : (a < b ? a : b); }
__device__ inline double nmax(double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

__device__ inline double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return __shfl(v, 0);
}
__device__ inline double wmax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off));
    return __shfl(v, 0);
}

template <typename T>
__device__ inline double wdot(const T* a, const T* b, int n) {
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += WAVE) s += (double)a[k] * (double)b[k];
    return wsum(s);
}

template <typename T>
__global__ __launch_bounds__(WAVE) void k_lb_begin(pk_lbfgs_batch B) {
    const int a = blockIdx.x;
    if (a >= B.M) return;
    const T* g = reinterpret_cast<const T*>(B.g) + (size_t)a * B.n;
    double gm = 0.0;
    for (int k = threadIdx.x; k < B.n; k += WAVE) gm = nmax(gm, fabs((double)g[k]));
    gm = wmax(gm);
    if (threadIdx.x == 0) {
        LbSys* S = reinterpret_cast<LbSys*>(B.sys) + a;
        const double l = (double)reinterpret_cast<const T*>(B.loss)[a];
        S->H = 1.0;
        S->t = 0.0;
        S->loss_prev = l;
        S->loss0 = l;
        S->hist_len = 0;
        S->head = 0;
        S->n_eval = 1;
        S->active = !(gm <= B.tol_grad);
        if (S->active) atomicAdd(&B.counters[1], 1);   // active after iteration 0
    }
}

template <typename T>
__global__ __launch_bounds__(WAVE) void k_lb_direction(pk_lbfgs_batch B, int it) {
    const int a = blockIdx.x;
    if (a >= B.M || B.counters[0]) return;
    const int n = B.n, m = B.history, lane = threadIdx.x;
    LbSys* S = reinterpret_cast<LbSys*>(B.sys) + a;
    T* g = reinterpret_cast<T*>(B.g) + (size_t)a * n;
    T* d = reinterpret_cast<T*>(B.d) + (size_t)a * n;
    T* gp = reinterpret_cast<T*>(B.g_prev) + (size_t)a * n;
    T* Sh = reinterpret_cast<T*>(B.s_hist) + (size_t)a * m * n;
    T* Yh = reinterpret_cast<T*>(B.y_hist) + (size_t)a * m * n;
    double* ro = reinterpret_cast<double*>(B.ro) + (size_t)a * m;
    double* al = reinterpret_cast<double*>(B.al) + (size_t)a * m;
    const double loss = (double)reinterpret_cast<const T*>(B.loss)[a];
    int active = S->active;
    int hl = S->hist_len, head = S->head;
    double H = S->H;
    if (it > 1) {
        // ---- history update of iteration it-1 (core/lbfgs_torch_parallel.py:636-690) ----
        const int prev = it - 1;
        const double t = S->t;
        if (prev > 1) {
            double ys = 0.0, yy = 0.0;
            for (int k = lane; k < n; k += WAVE) {
                const double yk = (double)(g[k] - gp[k]);
                const double sk = (double)(d[k] * (T)t);
                ys += yk * sk;
                yy += yk * yk;
            }
            ys = wsum(ys);
            yy = wsum(yy);
            if (active && ys > 1e-10) {
                H = ys / yy;
                int slot;
                if (hl < m) {
                    slot = (head + hl) % m;
                    hl += 1;
                } else {
                    slot = head;
                    head = (head + 1) % m;
                }
                T* so = Sh + (size_t)slot * n;
                T* yo = Yh + (size_t)slot * n;
                for (int k = lane; k < n; k += WAVE) {
                    yo[k] = g[k] - gp[k];
                    so[k] = d[k] * (T)t;
                }
                if (lane == 0) ro[slot] = 1.0 / ys;
            }
        }
        // ---- convergence tests of iteration it-1 (:692-724) ----
        double gm = 0.0, pc = 0.0;
        for (int k = lane; k < n; k += WAVE) {
            gm = nmax(gm, fabs((double)g[k]));
            pc = nmax(pc, fabs((double)(d[k] * (T)t)));
        }
        gm = wmax(gm);
        pc = wmax(pc);
        if (gm <= B.tol_grad) active = 0;
        // evaluations so far: one per iteration without a line search, counted with one
        if ((B.line_search ? S->n_eval : 1 + prev) >= B.max_eval) active = 0;
        if (pc <= B.tol_change) active = 0;
        const double lc = fabs(loss - S->loss_prev);
        if (lc < B.tol_change) active = 0;
        if (B.tol_rel > 0 && prev > m && lc / (fabs(S->loss_prev) + 1e-12) < B.tol_rel) active = 0;
        if (B.stop_on_increase && loss > S->loss0) active = 0;
        if (lane == 0 && active) atomicAdd(&B.counters[1 + 2 * prev], 1);
    }
    // ---- direction of iteration it (:527-572) ----
    for (int k = lane; k < n; k += WAVE) d[k] = -g[k];
    if (it > 1) {
        for (int i = hl - 1; i >= 0; --i) {
            const int ph = (head + i) % m;
            const double a_i = wdot<T>(Sh + (size_t)ph * n, d, n) * ro[ph];
            if (lane == 0) al[i] = a_i;
            const T af = (T)a_i;
            const T* yi = Yh + (size_t)ph * n;
            for (int k = lane; k < n; k += WAVE) d[k] = d[k] - yi[k] * af;
        }
        const T hd = (T)H;
        for (int k = lane; k < n; k += WAVE) d[k] = d[k] * hd;
        for (int i = 0; i < hl; ++i) {
            const int ph = (head + i) % m;
            const double be = wdot<T>(Yh + (size_t)ph * n, d, n) * ro[ph];
            const T c = (T)(__shfl(lane == 0 ? al[i] : 0.0, 0) - be);
            const T* si = Sh + (size_t)ph * n;
            for (int k = lane; k < n; k += WAVE) d[k] = d[k] + si[k] * c;
        }
    } else {
        H = 1.0;
    }
    // ---- step length, clamp, directional derivative (:574-604) ----
    for (int k = lane; k < n; k += WAVE) gp[k] = g[k];
    double t;
    if (it == 1) {
        double gs = 0.0;
        for (int k = lane; k < n; k += WAVE) gs += fabs((double)g[k]);
        gs = wsum(gs);
        t = nmin(1.0, 1.0 / gs) * B.lr;
        if (!active) t = 0.0;
    } else {
        t = B.lr;
    }
    if (B.max_step > 0) {
        double pc = 0.0;
        for (int k = lane; k < n; k += WAVE) pc = nmax(pc, fabs((double)d[k] * t));
        pc = wmax(pc);
        if (pc > B.max_step) t *= B.max_step / pc;
    }
    const double gtd = wdot<T>(g, d, n);
    if (gtd > -B.tol_change) active = 0;
    if (lane == 0) {
        S->t = t;
        S->H = H;
        S->hist_len = hl;
        S->head = head;
        S->active = active;
        S->loss_prev = loss;
        if (active) atomicAdd(&B.counters[2 + 2 * it], 1);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_lb_step(pk_lbfgs_batch B, int it) {
    if (B.counters[0]) return;
    if (B.counters[1 + 2 * (it - 1)] == 0 || B.counters[2 + 2 * it] == 0) {
        if (threadIdx.x == 0) B.counters[0] = 1;
        return;
    }
    const LbSys* S = reinterpret_cast<const LbSys*>(B.sys);
    T* x = reinterpret_cast<T*>(B.x);
    const T* d = reinterpret_cast<const T*>(B.d);
    const long long total = (long long)B.M * B.n;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(q / B.n);
        x[q] = x[q] + (T)S[a].t * d[q];
    }
}

// ---------------------------------------------------------------------------------
// Strong-Wolfe line search (core/lbfgs_torch_parallel.py:30-407).  The search of system a
// is on x_a + t d_a; its scalars are double; the bracket holds two points (t, f, g, g.d).
// Constants of the reference's call (it passes none): c1 1e-4, c2 0.9, tolerance_change
// 1e-9 on the bracket width times max |d|, 25 search iterations shared by the batch.
// ---------------------------------------------------------------------------------
constexpr double LS_C1 = 1e-4, LS_C2 = 0.9, LS_TOL = 1e-9;

struct LsSys {
    double t, t_prev, f0, gtd0, f_prev, gtd_prev, f_new, gtd_new, dnorm, t_new;
    double bt[2], bf[2], bgtd[2];
    int size, done, active, low, insuf, evals, in_ls, pad;
};
static_assert(sizeof(LsSys) == PK_LBFGS_LS_BYTES, "line-search record layout");

// Minimiser of the cubic through (x1, f1, g1), (x2, f2, g2), clamped to [lo, hi]; the
// midpoint when the cubic has no real minimiser (core/lbfgs_torch_parallel.py:30-66).
__device__ inline double ls_cubic(double x1, double f1, double g1, double x2, double f2, double g2, double lo,
                                  double hi) {
    const double d1 = g1 + g2 - 3.0 * (f1 - f2) / (x1 - x2);
    const double d2s = d1 * d1 - g1 * g2;
    if (!(d2s >= 0.0)) return (lo + hi) / 2.0;
    const double d2 = sqrt(d2s);
    const double mp = x1 <= x2 ? x2 - (x2 - x1) * ((g2 + d2 - d1) / (g2 - g1 + 2.0 * d2))
                               : x1 - (x1 - x2) * ((g1 + d2 - d1) / (g1 - g2 + 2.0 * d2));
    return nmin(nmax(mp, lo), hi);   // torch.clamp(min=lo, max=hi): max first, then min
}

template <typename T>
struct LsView {
    const pk_lbfgs_batch& B;
    const pk_lbfgs_ls& Q;
    int a, n;
    __device__ LsView(const pk_lbfgs_batch& b, const pk_lbfgs_ls& q, int sys) : B(b), Q(q), a(sys), n(b.n) {}
    __device__ LsSys* rec() const { return reinterpret_cast<LsSys*>(Q.sys) + a; }
    __device__ T* x() const { return reinterpret_cast<T*>(B.x) + (size_t)a * n; }
    __device__ T* d() const { return reinterpret_cast<T*>(B.d) + (size_t)a * n; }
    __device__ T* g0() const { return reinterpret_cast<T*>(B.g) + (size_t)a * n; }
    __device__ T* xt() const { return reinterpret_cast<T*>(Q.x) + (size_t)a * n; }
    __device__ T* gn() const { return reinterpret_cast<T*>(Q.g) + (size_t)a * n; }
    __device__ T* gp() const { return reinterpret_cast<T*>(Q.g_prev) + (size_t)a * n; }
    __device__ T* bg(int k) const { return reinterpret_cast<T*>(Q.bg) + ((size_t)a * 2 + k) * n; }
    __device__ double fn() const { return (double)reinterpret_cast<const T*>(Q.f)[a]; }
    __device__ void copy(T* dst, const T* src) const {
        for (int k = threadIdx.x; k < n; k += WAVE) dst[k] = src[k];
    }
    // trial point x + t d, flagged for the next objective call
    __device__ void trial(double t) const {
        const T tf = (T)t;
        T* xt_ = xt();
        const T* x_ = x();
        const T* d_ = d();
        for (int k = threadIdx.x; k < n; k += WAVE) xt_[k] = x_[k] + tf * d_[k];
    }
};

__device__ inline void ls_flag(const pk_lbfgs_ls& Q, int a, int f) {
    if (threadIdx.x == 0) {
        Q.flag[a] = f;
        if (f) atomicAdd(Q.count, 1);
    }
}

template <typename T>
__global__ __launch_bounds__(WAVE) void k_ls_begin(pk_lbfgs_batch B, pk_lbfgs_ls Q) {
    const int a = blockIdx.x;
    if (a >= B.M) return;
    const LbSys* S = reinterpret_cast<const LbSys*>(B.sys) + a;
    LsView<T> V(B, Q, a);
    LsSys* R = V.rec();
    const int in_ls = S->active;
    const double gtd0 = wdot<T>(V.g0(), V.d(), V.n);
    double dn = 0.0;
    for (int k = threadIdx.x; k < V.n; k += WAVE) dn = nmax(dn, fabs((double)V.d()[k]));
    dn = wmax(dn);
    if (in_ls) {
        V.copy(V.gp(), V.g0());
        V.trial(S->t);
    }
    if (threadIdx.x == 0) {
        LsSys r;
        memset(&r, 0, sizeof(r));
        r.in_ls = in_ls;
        r.t = S->t;
        r.f0 = (double)reinterpret_cast<const T*>(B.loss)[a];
        r.gtd0 = gtd0;
        r.dnorm = dn;
        r.f_prev = r.f0;
        r.gtd_prev = gtd0;
        r.active = in_ls;
        r.evals = in_ls ? 1 : 0;
        *R = r;
    }
    ls_flag(Q, a, in_ls);
}

// g.d at the rows just evaluated
template <typename T>
__global__ __launch_bounds__(WAVE) void k_ls_gtd(pk_lbfgs_batch B, pk_lbfgs_ls Q) {
    const int a = blockIdx.x;
    if (a >= B.M || !Q.flag[a]) return;
    LsView<T> V(B, Q, a);
    const double v = wdot<T>(V.gn(), V.d(), V.n);
    if (threadIdx.x == 0) V.rec()->gtd_new = v;
}

// One bracketing iteration (core/lbfgs_torch_parallel.py:137-220): the tests on the last
// trial, then the next trial for systems still searching.
template <typename T>
__global__ __launch_bounds__(WAVE) void k_ls_bracket(pk_lbfgs_batch B, pk_lbfgs_ls Q, int ls_iter) {
    const int a = blockIdx.x;
    if (a >= B.M) return;
    LsView<T> V(B, Q, a);
    LsSys r = *V.rec();
    if (!(r.in_ls && r.active)) {
        ls_flag(Q, a, 0);
        return;
    }
    __syncthreads();   // every lane holds r before lane 0 writes the record
    r.f_new = V.fn();
    const bool armijo_fail = r.f_new > r.f0 + LS_C1 * r.t * r.gtd0;
    const bool not_dec = ls_iter > 1 && r.f_new >= r.f_prev;
    int set2 = 0;
    if (armijo_fail || not_dec) {
        set2 = 1;
    } else if (fabs(r.gtd_new) <= -LS_C2 * r.gtd0) {
        r.size = 1;
        r.bt[0] = r.t;
        r.bf[0] = r.f_new;
        V.copy(V.bg(0), V.gn());
        r.done = 1;
        r.active = 0;
    } else if (r.gtd_new >= 0.0) {
        set2 = 1;
    }
    if (set2) {
        r.size = 2;
        r.bt[0] = r.t_prev;
        r.bt[1] = r.t;
        r.bf[0] = r.f_prev;
        r.bf[1] = r.f_new;
        r.bgtd[0] = r.gtd_prev;
        r.bgtd[1] = r.gtd_new;
        V.copy(V.bg(0), V.gp());
        V.copy(V.bg(1), V.gn());
        r.active = 0;
    }
    if (r.active) {
        const double lo = r.t + 0.01 * (r.t - r.t_prev), hi = r.t * 10.0;
        const double tn = ls_cubic(r.t_prev, r.f_prev, r.gtd_prev, r.t, r.f_new, r.gtd_new, lo, hi);
        r.t = tn;
        r.t_prev = tn;   // the reference copies t after replacing it
        r.f_prev = r.f_new;
        r.gtd_prev = r.gtd_new;
        V.copy(V.gp(), V.gn());
        V.trial(tn);
        r.evals += 1;
    }
    if (threadIdx.x == 0) *V.rec() = r;
    ls_flag(Q, a, r.active);
}

// End of bracketing (:222-245): a system still unbracketed after the last search
// iteration takes [0, t]; the zoom starts with low = 1 when f0 <= f1, as the reference
// initialises it.
template <typename T>
__global__ __launch_bounds__(WAVE) void k_ls_zoom_begin(pk_lbfgs_batch B, pk_lbfgs_ls Q, int ls_iter, int max_ls) {
    const int a = blockIdx.x;
    if (a >= B.M) return;
    LsView<T> V(B, Q, a);
    LsSys r = *V.rec();
    __syncthreads();
    if (!r.in_ls) {
        ls_flag(Q, a, 0);
        return;
    }
    if (r.size == 0 && ls_iter == max_ls) {
        r.size = 2;
        r.bt[0] = 0.0;
        r.bt[1] = r.t;
        r.bf[0] = r.f0;
        r.bf[1] = V.fn();
        V.copy(V.bg(0), V.g0());
        V.copy(V.bg(1), V.gn());
    }
    r.active = r.size == 2 && !r.done;
    r.low = r.bf[0] <= r.bf[1] ? 1 : 0;
    r.insuf = 0;
    if (threadIdx.x == 0) *V.rec() = r;
    ls_flag(Q, a, r.active);
}

// Zoom trial (:247-300): width test, cubic step inside the bracket, moved off a bracket end
// when too close (0.1 of the width) and either past it or after insufficient progress.
template <typename T>
__global__ __launch_bounds__(WAVE) void k_ls_zoom_trial(pk_lbfgs_batch B, pk_lbfgs_ls Q) {
    const int a = blockIdx.x;
    if (a >= B.M) return;
    LsView<T> V(B, Q, a);
    LsSys r = *V.rec();
    __syncthreads();
    if (!(r.in_ls && r.active)) {
        ls_flag(Q, a, 0);
        return;
    }
    const double width = fabs(r.bt[1] - r.bt[0]);
    if (width * r.dnorm < LS_TOL) {
        r.active = 0;
        if (threadIdx.x == 0) *V.rec() = r;
        ls_flag(Q, a, 0);
        return;
    }
    const double bmin = nmin(r.bt[0], r.bt[1]), bmax = nmax(r.bt[0], r.bt[1]);
    double tn = ls_cubic(r.bt[0], r.bf[0], r.bgtd[0], r.bt[1], r.bf[1], r.bgtd[1], bmin, bmax);
    const double eps = 0.1 * width;
    const double dist = nmin(bmax - tn, tn - bmin);
    const bool close = dist < eps;
    const bool adj = close && (r.insuf || tn >= bmax || tn <= bmin);
    if (adj) {
        tn = fabs(tn - bmax) < fabs(tn - bmin) ? bmax - eps : bmin + eps;
        r.insuf = 0;
    } else {
        r.insuf = close ? 1 : 0;
    }
    r.t_new = tn;
    V.trial(tn);
    r.evals += 1;
    if (threadIdx.x == 0) *V.rec() = r;
    ls_flag(Q, a, 1);
}

// Zoom update after the trial's evaluation (:305-380).
template <typename T>
__global__ __launch_bounds__(WAVE) void k_ls_zoom_update(pk_lbfgs_batch B, pk_lbfgs_ls Q) {
    const int a = blockIdx.x;
    if (a >= B.M) return;
    LsView<T> V(B, Q, a);
    LsSys r = *V.rec();
    __syncthreads();
    if (!(r.in_ls && r.active && Q.flag[a])) {
        ls_flag(Q, a, 0);
        return;
    }
    const double fnew = V.fn(), tn = r.t_new;
    const int lo = r.low, hi = 1 - r.low;
    const bool todo1 = fnew > r.f0 + LS_C1 * tn * r.gtd0 || fnew >= r.bf[lo];
    if (todo1) {
        r.bt[hi] = tn;
        r.bf[hi] = fnew;
        r.bgtd[hi] = r.gtd_new;
        V.copy(V.bg(hi), V.gn());
        r.low = r.bf[0] <= r.bf[1] ? 0 : 1;
    } else {
        if (fabs(r.gtd_new) <= -LS_C2 * r.gtd0) {
            r.done = 1;
            r.active = 0;
        } else if (r.gtd_new * (r.bt[hi] - r.bt[lo]) >= 0.0) {
            r.bt[hi] = r.bt[lo];
            r.bf[hi] = r.bf[lo];
            r.bgtd[hi] = r.bgtd[lo];
            V.copy(V.bg(hi), V.bg(lo));
        }
        r.bt[lo] = tn;
        r.bf[lo] = fnew;
        r.bgtd[lo] = r.gtd_new;
        V.copy(V.bg(lo), V.gn());
    }
    if (threadIdx.x == 0) *V.rec() = r;
    ls_flag(Q, a, r.active);
}

// Accepted step (:382-407, :606-622): bracket point 0 after a Wolfe hit in bracketing,
// else the low point; x += t d, loss and gradient of that point, evaluations counted.
template <typename T>
__global__ __launch_bounds__(WAVE) void k_ls_end(pk_lbfgs_batch B, pk_lbfgs_ls Q) {
    const int a = blockIdx.x;
    if (a >= B.M) return;
    LsView<T> V(B, Q, a);
    const LsSys r = *V.rec();
    if (!r.in_ls) return;
    const int k0 = r.size == 1 ? 0 : r.low;
    const double t = r.bt[k0];
    const T tf = (T)t;
    T* x = V.x();
    const T* d = V.d();
    T* g = V.g0();
    const T* gb = V.bg(k0);
    for (int k = threadIdx.x; k < V.n; k += WAVE) {
        x[k] = x[k] + tf * d[k];
        g[k] = gb[k];
    }
    if (threadIdx.x == 0) {
        LbSys* S = reinterpret_cast<LbSys*>(B.sys) + a;
        S->t = t;
        S->n_eval += r.evals;
        reinterpret_cast<T*>(B.loss)[a] = (T)r.bf[k0];
    }
}

__global__ void k_ls_zero(int32_t* c) { *c = 0; }

}  // namespace pk

using namespace pk;

extern "C" {

static int lb_check(const pk_lbfgs_batch* b) {
    PK_REQUIRE(b != nullptr, "pk_lbfgs: NULL batch");
    PK_REQUIRE(b->dtype == PK_F32 || b->dtype == PK_F64, "pk_lbfgs: dtype must be PK_F32 or PK_F64");
    PK_REQUIRE(b->M >= 0 && b->n >= 0 && b->history >= 1, "pk_lbfgs: bad sizes");
    return PK_OK;
}

int pk_lbfgs_sys_bytes(void) { return (int)sizeof(LbSys); }

int pk_lbfgs_begin(const pk_lbfgs_batch* b, void* stream) {
    int rc = lb_check(b);
    if (rc || b->M == 0) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (b->dtype == PK_F32) hipLaunchKernelGGL(k_lb_begin<float>, dim3(b->M), dim3(WAVE), 0, st, *b);
    else hipLaunchKernelGGL(k_lb_begin<double>, dim3(b->M), dim3(WAVE), 0, st, *b);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

int pk_lbfgs_direction(const pk_lbfgs_batch* b, int it, void* stream) {
    int rc = lb_check(b);
    if (rc || b->M == 0) return rc;
    PK_REQUIRE(it >= 1, "pk_lbfgs_direction: iterations count from 1");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (b->dtype == PK_F32) hipLaunchKernelGGL(k_lb_direction<float>, dim3(b->M), dim3(WAVE), 0, st, *b, it);
    else hipLaunchKernelGGL(k_lb_direction<double>, dim3(b->M), dim3(WAVE), 0, st, *b, it);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

static int ls_check(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q) {
    int rc = lb_check(b);
    if (rc) return rc;
    PK_REQUIRE(b->line_search == 1, "pk_lbfgs_ls: the batch was not set up for a line search");
    PK_REQUIRE(q && q->f && q->g && q->x && q->g_prev && q->bg && q->sys && q->flag && q->count,
               "pk_lbfgs_ls: missing buffer");
    return PK_OK;
}

// Launch kernel K<float|double> over the M systems, after clearing the flagged-row count
// with a one-thread kernel (no memset node, so the sequence can be captured in a graph).

#define PK_LS_LAUNCH(K, ...)                                                                  \
    do {                                                                                      \
        int _rc = ls_check(b, q);                                                             \
        if (_rc || b->M == 0) return _rc;                                                     \
        hipStream_t _st = reinterpret_cast<hipStream_t>(stream);                              \
        hipLaunchKernelGGL(k_ls_zero, dim3(1), dim3(1), 0, _st, q->count);                    \
        if (b->dtype == PK_F32)                                                               \
            hipLaunchKernelGGL(K<float>, dim3(b->M), dim3(WAVE), 0, _st, *b, *q, ##__VA_ARGS__); \
        else                                                                                  \
            hipLaunchKernelGGL(K<double>, dim3(b->M), dim3(WAVE), 0, _st, *b, *q, ##__VA_ARGS__); \
        PK_CHECK_HIP(hipGetLastError());                                                      \
        return PK_OK;                                                                         \
    } while (0)

int pk_lbfgs_ls_begin(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q, void* stream) { PK_LS_LAUNCH(k_ls_begin); }
int pk_lbfgs_ls_bracket(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q, int ls_iter, void* stream) {
    PK_LS_LAUNCH(k_ls_bracket, ls_iter);
}
int pk_lbfgs_ls_zoom_begin(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q, int ls_iter, void* stream) {
    PK_LS_LAUNCH(k_ls_zoom_begin, ls_iter, 25);
}
int pk_lbfgs_ls_zoom_trial(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q, void* stream) {
    PK_LS_LAUNCH(k_ls_zoom_trial);
}
int pk_lbfgs_ls_zoom_update(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q, void* stream) {
    PK_LS_LAUNCH(k_ls_zoom_update);
}
int pk_lbfgs_ls_end(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q, void* stream) {
    int rc = ls_check(b, q);
    if (rc || b->M == 0) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (b->dtype == PK_F32) hipLaunchKernelGGL(k_ls_end<float>, dim3(b->M), dim3(WAVE), 0, st, *b, *q);
    else hipLaunchKernelGGL(k_ls_end<double>, dim3(b->M), dim3(WAVE), 0, st, *b, *q);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}
int pk_lbfgs_ls_gtd(const pk_lbfgs_batch* b, const pk_lbfgs_ls* q, void* stream) {
    int rc = ls_check(b, q);
    if (rc || b->M == 0) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (b->dtype == PK_F32) hipLaunchKernelGGL(k_ls_gtd<float>, dim3(b->M), dim3(WAVE), 0, st, *b, *q);
    else hipLaunchKernelGGL(k_ls_gtd<double>, dim3(b->M), dim3(WAVE), 0, st, *b, *q);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

int pk_lbfgs_step(const pk_lbfgs_batch* b, int it, void* stream) {
    int rc = lb_check(b);
    if (rc || b->M == 0) return rc;
    const long long total = (long long)b->M * b->n;
    long long blocks = (total + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (b->dtype == PK_F32) hipLaunchKernelGGL(k_lb_step<float>, dim3((unsigned)blocks), dim3(256), 0, st, *b, it);
    else hipLaunchKernelGGL(k_lb_step<double>, dim3((unsigned)blocks), dim3(256), 0, st, *b, it);
    PK_CHECK_HIP(hipGetLastError());
    return PK_OK;
}

}  // extern "C"
