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
    int hist_len, head, active, pad;
};
static_assert(sizeof(LbSys) == PK_LBFGS_SYS_BYTES, "system record layout");

__device__ inline double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return __shfl(v, 0);
}
__device__ inline double wmax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
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
    for (int k = threadIdx.x; k < B.n; k += WAVE) gm = fmax(gm, fabs((double)g[k]));
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
            gm = fmax(gm, fabs((double)g[k]));
            pc = fmax(pc, fabs((double)(d[k] * (T)t)));
        }
        gm = wmax(gm);
        pc = wmax(pc);
        if (gm <= B.tol_grad) active = 0;
        if (1 + prev >= B.max_eval) active = 0;
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
        t = fmin(1.0, 1.0 / gs) * B.lr;
        if (!active) t = 0.0;
    } else {
        t = B.lr;
    }
    if (B.max_step > 0) {
        double pc = 0.0;
        for (int k = lane; k < n; k += WAVE) pc = fmax(pc, fabs((double)d[k] * t));
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
