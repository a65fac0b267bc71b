// pcg.hip -- Jacobi-preconditioned CG on the device, MFEM CGSolver semantics,
// and the mixed-precision solver built on it.
//
// CGSolver::Mult (linalg/solvers.cpp:869-1050) with OperatorJacobiSmoother
// (:401-453) on the DIAG_ONE ConstrainedOperator (linalg/operator.cpp:601-662).
// MFEM runs ~12 vector passes and two host round-trips (Dot -> device sync ->
// host sum, general/reducers.hpp:575-584) per iteration.  Here every scalar
// stays on the device; one iteration is
//   update : alpha = nom/den; x += alpha d; r -= alpha z; z = dinv r;
//            partials of (r, z)                     -> fin_update: betanom,
//            convergence test, beta = betanom/nom
//   dir    : d = z + beta d
//   apply  : element kernel on d (constrained entries read as 0)
//   scatter: z = A d (DIAG_ONE rows copy d), partials of (d, z) -> fin_den
// Dot products: each block reduces its chunk of kChunk entries in a fixed
// tree and stores the partial; a one-block finalize kernel sums the partials
// in index order.  No block waits for or fences against another (a
// device-scope fence writes back and invalidates the whole L2 on this GPU),
// and the result does not depend on timing.  The vector type T is double or
// float; scalars and dot accumulation are double.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdlib.h>

#include "internal.hpp"
#include "row_sum.hpp"

namespace h1pa {

struct PcgState {
    double nom, nom0, den, betanom, beta, r0;
    int iter;        // MFEM's i
    int in_loop;     // 0 before the first update, then 1
    int max_iter;
    int done;
    int converged;
    int final_iter;
    int nonfinite;   // a (B r, r) or (A d, d) was not finite: MFEM_VERIFY fails
};

namespace kern {
namespace {

constexpr int kThreads = 256;

// Chunking of a vector of n entries: `per` entries per thread, kThreads * per
// per block.  per grows from 1 to 8 with n so that large vectors get ~2048
// blocks (few partials to sum) and small ones still spread over the CUs.
struct Chunks {
    unsigned blocks;
    int per;
};
inline Chunks chunks(int64_t n) {
    int64_t per = (n + (int64_t)kThreads * 2048 - 1) / ((int64_t)kThreads * 2048);
    per = std::min<int64_t>(std::max<int64_t>(per, 1), 8);
    const int64_t c = kThreads * per;
    return {(unsigned)std::max<int64_t>((n + c - 1) / c, 1), (int)per};
}

// The entries of this block's chunk owned by this thread (per of them,
// kThreads apart); `per` is a kernel argument.
#define H1PA_CHUNK(i, n)                                                                       \
    for (int64_t i = (int64_t)blockIdx.x * kThreads * per + threadIdx.x,                       \
                 i##_end = min((int64_t)(n), (int64_t)(blockIdx.x + 1) * kThreads * per);      \
         i < i##_end; i += kThreads)

// The same entries taken kGrp at a time: a kernel loads every operand of a
// group at the clamped indices g.lt[k] before the group's first store, so the
// loads of a group go out together instead of one entry's round trip after
// another; stores and sums then run over the g.ok entries in k order, which
// is the H1PA_CHUNK order (dot products come out bit-identical).  Slots at or
// past `per` belong to no entry of this thread (small vectors have per < kGrp):
// kernels issue no loads for them and skip them, as every such slot is not ok.
constexpr int kGrp = 4;
struct Grp {
    int64_t t[kGrp], lt[kGrp];
    bool ok[kGrp];
};
__device__ __forceinline__ Grp grp_of(int64_t b, int64_t end) {
    Grp g;
#pragma unroll
    for (int k = 0; k < kGrp; ++k) {
        const int64_t t = b + (int64_t)k * kThreads;
        g.t[k] = t;
        g.ok[k] = t < end;
        g.lt[k] = t < end ? t : end - 1;
    }
    return g;
}
#define H1PA_GROUPS(g, n)                                                                      \
    for (int64_t g##_b = (int64_t)blockIdx.x * kThreads * per + threadIdx.x,                   \
                 g##_e = min((int64_t)(n), (int64_t)(blockIdx.x + 1) * kThreads * per);        \
         g##_b < g##_e; g##_b += (int64_t)kGrp * kThreads)                                     \
        for (const Grp g = grp_of(g##_b, g##_e), *g##_1 = &g; g##_1; g##_1 = nullptr)

// Block sum in a fixed tree order, returned to every thread.
__device__ inline double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// partials[k][blockIdx.x] = block sum of acc[k]
template <int NV>
__device__ inline void store_partials(const double (&acc)[NV], double *partials, double *sh) {
    for (int k = 0; k < NV; ++k) {
        const double v = block_sum(acc[k], sh);
        if (threadIdx.x == 0) partials[(int64_t)k * gridDim.x + blockIdx.x] = v;
    }
}

// One block: tot[k] = sum over b < nb of partials[k][b], in a fixed order.
// Eight independent loads in flight per thread (the partials are read once,
// latency rather than bandwidth bounds this).
template <int NV>
__device__ inline void sum_partials(const double *partials, int nb, double *sh,
                                    double (&tot)[NV]) {
    constexpr int U8 = 8;
    for (int k = 0; k < NV; ++k) {
        const double *p = partials + (int64_t)k * nb;
        double v[U8];
#pragma unroll
        for (int u = 0; u < U8; ++u) v[u] = 0.0;
        for (int b0 = 0; b0 < nb; b0 += U8 * kThreads) {
#pragma unroll
            for (int u = 0; u < U8; ++u) {
                const int b = b0 + u * kThreads + threadIdx.x;
                if (b < nb) v[u] += p[b];
            }
        }
        double w = 0.0;
#pragma unroll
        for (int u = 0; u < U8; ++u) w += v[u];
        tot[k] = block_sum(w, sh);
    }
}

template <typename T>
__global__ void k_jacobi_dinv(int64_t n, const double *__restrict__ diag,
                              const uint8_t *__restrict__ mask, T *__restrict__ dinv) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    dinv[i] = (T)((mask && mask[i]) ? 1.0 : 1.0 / diag[i]);
}

// ---- FP64 / FP32 Jacobi-PCG -------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_pcg_init(int64_t n, int per, const T *__restrict__ b, const T *__restrict__ dinv, T *__restrict__ x,
           T *__restrict__ r, T *__restrict__ z, T *__restrict__ d, double *__restrict__ partials) {
    __shared__ double sh[kThreads];
    double acc[1] = {0.0};
    H1PA_GROUPS(g, n) {
        T vb[kGrp], vi[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            vb[k] = b[g.lt[k]];
            vi[k] = dinv[g.lt[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;
            const int64_t i = g.t[k];
            const T ri = vb[k];
            const T zi = vi[k] * ri;
            r[i] = ri;
            x[i] = T(0);
            z[i] = zi;
            d[i] = zi;
            acc[0] += (double)zi * (double)ri;
        }
    }
    store_partials<1>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_pcg_fin_init(int nb, const double *__restrict__ partials, PcgState *st, double rel_tol,
               double abs_tol) {
    __shared__ double sh[kThreads];
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    const double nom = tot[0];
    st->nom = st->nom0 = nom;
    st->iter = 1;
    st->in_loop = 0;
    double r0 = nom * rel_tol * rel_tol;
    if (abs_tol * abs_tol > r0) r0 = abs_tol * abs_tol;
    st->r0 = r0;
    st->final_iter = 0;
    st->converged = 0;
    st->done = 0;
    st->nonfinite = 0;
    st->betanom = nom;
    if (!isfinite(nom)) {  // MFEM_VERIFY(IsFinite(nom)) in CGSolver::Mult
        st->done = 1;
        st->nonfinite = 1;
    } else if (nom < 0.0) {  // the preconditioner is not positive definite
        st->done = 1;
    } else if (nom <= r0) {
        st->done = 1;
        st->converged = 1;
    }
}

// z = A d from the element output, DIAG_ONE rows copy d; partials of (d, z).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_pcg_scatter(int64_t n, int per, int64_t bb, const int *__restrict__ offsets,
              const int *__restrict__ indices, const T *__restrict__ ye,
              const uint8_t *__restrict__ mask, const T *__restrict__ d, T *__restrict__ z,
              double *__restrict__ partials, const PcgState *st, const int *__restrict__ perm,
              int64_t nperm) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double acc[1] = {0.0};
    H1PA_GROUPS(g, n) {
        // the group's rows, then their offsets, masks and d values, as batches
        int64_t ii[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            ii[k] = g.lt[k] < nperm ? perm[g.lt[k]] : g.lt[k];
        }
        T vd[kGrp];
        int ob[kGrp], oe[kGrp];
        bool m[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            vd[k] = d[ii[k]];
            ob[k] = offsets[ii[k]];
            oe[k] = offsets[ii[k] + 1];
            m[k] = mask && mask[ii[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;  // (every slot past `per` is not ok)
            T v;
            if (g.t[k] >= bb) {  // interior DOF: written by the element kernel
                v = z[ii[k]];
            } else {
                v = m[k] ? vd[k] : row_sum(indices, ob[k], oe[k], ye);
            }
            if (g.t[k] < bb) z[ii[k]] = v;
            acc[0] += (double)vd[k] * (double)v;
        }
    }
    store_partials<1>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_pcg_fin_den(int nb, const double *__restrict__ partials, PcgState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    st->den = tot[0];
    if (!isfinite(tot[0])) {  // MFEM_VERIFY(IsFinite(den))
        st->done = 1;
        st->nonfinite = 1;
        st->final_iter = st->in_loop ? st->iter : 0;
    } else if (tot[0] == 0.0) {  // CGSolver: den <= 0 and den == 0 ends the loop
        st->done = 1;
        st->final_iter = st->in_loop ? st->iter : 0;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_pcg_update(int64_t n, int per, const T *__restrict__ dinv, T *__restrict__ x, T *__restrict__ r,
             T *__restrict__ z, const T *__restrict__ d, double *__restrict__ partials,
             const PcgState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    const T alpha = (T)(st->nom / st->den);
    double acc[1] = {0.0};
    H1PA_GROUPS(g, n) {
        T vd[kGrp], vx[kGrp], vr[kGrp], vz[kGrp], vi[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            const int64_t j = g.lt[k];
            vd[k] = d[j];
            vx[k] = x[j];
            vr[k] = r[j];
            vz[k] = z[j];
            vi[k] = dinv[j];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;
            const int64_t i = g.t[k];
            x[i] = vx[k] + alpha * vd[k];
            const T ri = vr[k] - alpha * vz[k];
            r[i] = ri;
            const T zi = vi[k] * ri;
            z[i] = zi;
            acc[0] += (double)ri * (double)zi;
        }
    }
    store_partials<1>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_pcg_fin_update(int nb, const double *__restrict__ partials, PcgState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    const double betanom = tot[0];
    const int i = st->iter;
    st->in_loop = 1;
    st->betanom = betanom;
    if (!isfinite(betanom)) {  // MFEM_VERIFY(IsFinite(betanom))
        st->done = 1;
        st->nonfinite = 1;
        st->final_iter = i;
    } else if (betanom < 0.0) {
        st->done = 1;
        st->final_iter = i;
    } else if (betanom <= st->r0) {
        st->done = 1;
        st->converged = 1;
        st->final_iter = i;
    } else if (i + 1 > st->max_iter) {
        st->done = 1;
        st->final_iter = st->max_iter;
    } else {
        st->beta = betanom / st->nom;
        st->nom = betanom;
        st->iter = i + 1;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_pcg_dir(int64_t n, int per, const T *__restrict__ z, T *__restrict__ d, const PcgState *st) {
    if (st->done) return;
    const T beta = (T)st->beta;
    H1PA_GROUPS(g, n) {
        T vz[kGrp], vd[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            vz[k] = z[g.lt[k]];
            vd[k] = d[g.lt[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k)
            if (g.ok[k]) d[g.t[k]] = vz[k] + beta * vd[k];
    }
}

__global__ void k_to_float(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// FP64 -> BF16: round to FP32 (nearest even), then keep the upper 16 bits
// with round-to-nearest-even on the dropped half (NaN stays quiet NaN).
__global__ void k_to_bf16(int64_t n, const double *__restrict__ a, uint16_t *__restrict__ b) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float f = (float)a[i];
    const uint32_t u = __float_as_uint(f);
    uint16_t h;
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
        h = (uint16_t)((u >> 16) | 0x40u);
    else
        h = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    b[i] = h;
}

// E->L set over all rows in reduced precision (no constraints).
template <typename T>
__global__ void k_scatter_set(int64_t n, const int *__restrict__ offsets,
                              const int *__restrict__ indices, const T *__restrict__ ye,
                              T *__restrict__ y, const int *__restrict__ perm, int64_t nperm) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = t < nperm ? perm[t] : t;
    y[i] = row_sum(indices, offsets[i], offsets[i + 1], ye);
}

// ---- mixed precision: CG with reliable updates -----------------------------
// The FP32 inner CG runs on its recursive residual r and accumulates its
// solution in xs.  Whenever the preconditioned norm (r, z) has fallen by
// delta^2 since the last refinement (or reaches the stop level, or the
// iteration limit), a refinement step replaces it with FP64 quantities:
// x += xs, xs = 0, r64 = b - A x (FP64 operator), r = (float) r64 -- and the
// search direction d is kept, so the Krylov space is not restarted.  The
// stop test is MFEM's (B r, r) <= rel_tol^2 (B b, b) evaluated on the FP64
// true residual.
struct IrState {
    double rho, rho_ref, den, beta;  // inner CG, sums in FP64
    double rho64, rho64_0, r2, b2;   // FP64 true residual
    double tol2, delta2;
    int iter, outer, max_inner, max_outer;
    int ru;   // 1: refinement step due this iteration
    int nru;  // !ru: skip flag of the FP64 element kernel
    int done, converged;
    int nonfinite;  // a dot product was not finite (CGSolver's MFEM_VERIFY)
};

__global__ void __launch_bounds__(kThreads)
k_ir_init(int64_t n, int per, const double *__restrict__ b, const double *__restrict__ dinv64,
          const float *__restrict__ dinv, double *__restrict__ x, double *__restrict__ r64,
          float *__restrict__ r, float *__restrict__ xs, float *__restrict__ z,
          float *__restrict__ d, double *__restrict__ partials) {
    __shared__ double sh[kThreads];
    double acc[3] = {0.0, 0.0, 0.0};
    H1PA_CHUNK(i, n) {
        const double bi = b[i];
        x[i] = 0.0;
        r64[i] = bi;
        const float ri = (float)bi;
        const float zi = dinv[i] * ri;
        r[i] = ri;
        xs[i] = 0.f;
        z[i] = zi;
        d[i] = zi;
        acc[0] += (double)ri * (double)zi;
        acc[1] += bi * dinv64[i] * bi;
        acc[2] += bi * bi;
    }
    store_partials<3>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_ir_fin_init(int nb, const double *__restrict__ partials, IrState *st) {
    __shared__ double sh[kThreads];
    double tot[3];
    sum_partials<3>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    st->rho = st->rho_ref = tot[0];
    st->rho64 = st->rho64_0 = tot[1];
    st->r2 = st->b2 = tot[2];
    st->iter = 0;
    st->outer = 0;
    st->ru = 0;
    st->nru = 1;
    st->converged = 0;
    st->done = 0;
    st->nonfinite = 0;
    if (!isfinite(tot[0]) || !isfinite(tot[1])) {
        st->done = 1;
        st->nonfinite = 1;
    } else if (tot[1] < 0.0) {  // the preconditioner is not positive definite
        st->done = 1;
    } else if (tot[1] == 0.0) {  // b = 0 on the free DOFs: x = 0 is the solution
        st->done = 1;
        st->converged = 1;
    }
}

// q = A d (FP32; DIAG_ONE rows copy d), partials of (d, q).
__global__ void __launch_bounds__(kThreads)
k_ir_scatter(int64_t n, int per, int64_t bb, const int *__restrict__ offsets,
             const int *__restrict__ indices, const float *__restrict__ ye,
             const uint8_t *__restrict__ mask, const float *__restrict__ d,
             float *__restrict__ q, double *__restrict__ partials, const IrState *st,
             const int *__restrict__ perm, int64_t nperm) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double acc[1] = {0.0};
    H1PA_GROUPS(g, n) {
        int64_t ii[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            ii[k] = g.lt[k] < nperm ? perm[g.lt[k]] : g.lt[k];
        }
        float vd[kGrp];
        int ob[kGrp], oe[kGrp];
        bool m[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            vd[k] = d[ii[k]];
            ob[k] = offsets[ii[k]];
            oe[k] = offsets[ii[k] + 1];
            m[k] = mask && mask[ii[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;  // (every slot past `per` is not ok)
            float v;
            if (g.t[k] >= bb) {  // interior DOF: written by the element kernel
                v = q[ii[k]];
            } else {
                v = m[k] ? vd[k] : row_sum(indices, ob[k], oe[k], ye);
            }
            if (g.t[k] < bb) q[ii[k]] = v;
            acc[0] += (double)vd[k] * (double)v;
        }
    }
    store_partials<1>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_ir_fin_den(int nb, const double *__restrict__ partials, IrState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    st->den = tot[0];
    if (!isfinite(tot[0])) {
        st->done = 1;
        st->nonfinite = 1;
    }
}

__global__ void __launch_bounds__(kThreads)
k_ir_update(int64_t n, int per, const float *__restrict__ dinv, float *__restrict__ xs,
            float *__restrict__ r, float *__restrict__ z, const float *__restrict__ q,
            const float *__restrict__ d, double *__restrict__ partials, const IrState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    const double den = st->den;
    const float alpha = den != 0.0 ? (float)(st->rho / den) : 0.f;
    double acc[1] = {0.0};
    H1PA_GROUPS(g, n) {
        float vs[kGrp], vd[kGrp], vr[kGrp], vq[kGrp], vi[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            const int64_t j = g.lt[k];
            vs[k] = xs[j];
            vd[k] = d[j];
            vr[k] = r[j];
            vq[k] = q[j];
            vi[k] = dinv[j];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;
            const int64_t i = g.t[k];
            xs[i] = vs[k] + alpha * vd[k];
            const float ri = vr[k] - alpha * vq[k];
            r[i] = ri;
            const float zi = vi[k] * ri;
            z[i] = zi;
            acc[0] += (double)ri * (double)zi;
        }
    }
    store_partials<1>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_ir_fin_update(int nb, const double *__restrict__ partials, IrState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    const double rho_new = tot[0];
    st->iter += 1;
    if (!isfinite(rho_new)) {
        st->done = 1;
        st->nonfinite = 1;
        return;
    }
    if (rho_new < 0.0) {  // the preconditioner is not positive definite
        st->done = 1;
        return;
    }
    const bool trigger = st->den == 0.0 || rho_new <= st->delta2 * st->rho_ref ||
                         rho_new <= st->tol2 * st->rho64_0 || st->iter >= st->max_inner;
    if (trigger) {
        st->ru = 1;
        st->nru = 0;
    } else {
        st->beta = rho_new / st->rho;
        st->rho = rho_new;
    }
}

// Refinement step, part 1: x += xs, xs = 0 (grid-stride; cheap when skipped).
__global__ void __launch_bounds__(kThreads)
k_ir_ru_x(int64_t n, float *__restrict__ xs, double *__restrict__ x, const IrState *st) {
    if (!st->ru) return;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        x[i] += (double)xs[i];
        xs[i] = 0.f;
    }
}

// Refinement step, part 2 (after the FP64 element kernel on x):
// r64 = b - A x; partials of (B r64, r64), ||r64||^2 and, with
// r = (float) r64, z = dinv r, of (r, z).
__global__ void __launch_bounds__(kThreads)
k_ir_ru_res(int64_t n, int per, int64_t bb, const int *__restrict__ offsets,
            const int *__restrict__ indices, const double *__restrict__ ye64,
            const uint8_t *__restrict__ mask,
            const double *__restrict__ b, const double *__restrict__ x,
            const double *__restrict__ dinv64, const float *__restrict__ dinv,
            double *__restrict__ r64, float *__restrict__ r, float *__restrict__ z,
            double *__restrict__ partials, const IrState *st, const int *__restrict__ perm,
            int64_t nperm) {
    __shared__ double sh[kThreads];
    if (!st->ru) return;
    double acc[3] = {0.0, 0.0, 0.0};
    H1PA_CHUNK(t, n) {
        const int64_t i = t < nperm ? perm[t] : t;
        double y;
        if (t >= bb) {  // interior DOF: (A x)_i written into r64 by the element kernel
            y = r64[i];
        } else if (mask && mask[i]) {
            y = x[i];
        } else {
            y = row_sum(indices, offsets[i], offsets[i + 1], ye64);
        }
        const double ri = b[i] - y;
        r64[i] = ri;
        const float rf = (float)ri;
        const float zf = dinv[i] * rf;
        r[i] = rf;
        z[i] = zf;
        acc[0] += (double)rf * (double)zf;
        acc[1] += ri * dinv64[i] * ri;
        acc[2] += ri * ri;
    }
    store_partials<3>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_ir_fin_ru(int nb, const double *__restrict__ partials, IrState *st) {
    __shared__ double sh[kThreads];
    if (!st->ru) return;
    double tot[3];
    sum_partials<3>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    st->outer += 1;
    st->rho64 = tot[1];
    st->r2 = tot[2];
    st->ru = 0;
    st->nru = 1;
    if (!isfinite(tot[0]) || !isfinite(tot[1])) {
        st->done = 1;
        st->nonfinite = 1;
    } else if (tot[1] < 0.0) {
        st->done = 1;
    } else if (tot[1] <= st->tol2 * st->rho64_0) {
        st->done = 1;
        st->converged = 1;
    } else if (st->outer >= st->max_outer || st->iter >= st->max_inner) {
        st->done = 1;
    } else {
        st->beta = st->rho > 0.0 ? tot[0] / st->rho : 0.0;
        st->rho = tot[0];
        st->rho_ref = tot[0];
    }
}

__global__ void __launch_bounds__(kThreads)
k_ir_dir(int64_t n, int per, const float *__restrict__ z, float *__restrict__ d, const IrState *st) {
    if (st->done) return;
    const float beta = (float)st->beta;
    H1PA_GROUPS(g, n) {
        float vz[kGrp], vd[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            vz[k] = z[g.lt[k]];
            vd[k] = d[g.lt[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k)
            if (g.ok[k]) d[g.t[k]] = vz[k] + beta * vd[k];
    }
}

// ---- Chebyshev smoother (OperatorChebyshevSmoother::Mult, solvers.cpp:630-656)
// and the CG steps around it.  `stop` (nullable) ends the work once a solve
// is done.

// t = x * dinv; y = 0 + c0 t  (k = 0 of the smoother)
__global__ void __launch_bounds__(kThreads)
k_cheb_first(int64_t n, int per, const double *__restrict__ dinv, const double *__restrict__ x,
             double *__restrict__ t, double *__restrict__ y, double c0, const int *stop) {
    if (stop && *stop) return;
    H1PA_GROUPS(g, n) {
        double vx[kGrp], vi[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            vx[k] = x[g.lt[k]];
            vi[k] = dinv[g.lt[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;
            const double ti = vx[k] * vi[k];
            t[g.t[k]] = ti;
            y[g.t[k]] = 0.0 + c0 * ti;
        }
    }
}

// k >= 1: u = A t (constrained: interior rows written by the element kernel
// into u, DIAG_ONE rows copy t); t = u * dinv; y += ck t
__global__ void __launch_bounds__(kThreads)
k_cheb_scatter(int64_t n, int per, int64_t bb, const int *__restrict__ offsets,
               const int *__restrict__ indices, const double *__restrict__ ye,
               const uint8_t *__restrict__ mask, const double *__restrict__ dinv,
               const double *__restrict__ u, double *__restrict__ t, double *__restrict__ y,
               double ck, const int *stop, const int *__restrict__ perm, int64_t nperm) {
    if (stop && *stop) return;
    H1PA_GROUPS(g, n) {
        int64_t ii[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            ii[k] = g.lt[k] < nperm ? perm[g.lt[k]] : g.lt[k];
        }
        double vi[kGrp], vy[kGrp], vt[kGrp];
        int ob[kGrp], oe[kGrp];
        bool m[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            vi[k] = dinv[ii[k]];
            vy[k] = y[ii[k]];
            vt[k] = t[ii[k]];
            ob[k] = offsets[ii[k]];
            oe[k] = offsets[ii[k] + 1];
            m[k] = mask && mask[ii[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;  // (every slot past `per` is not ok)
            double v;
            if (g.t[k] >= bb) v = u[ii[k]];
            else v = m[k] ? vt[k] : row_sum(indices, ob[k], oe[k], ye);
            const double ti = v * vi[k];
            t[ii[k]] = ti;
            y[ii[k]] = vy[k] + ck * ti;
        }
    }
}

// CGSolver::Mult start without a fused preconditioner: r = b, x = 0.
__global__ void __launch_bounds__(kThreads)
k_cg_start(int64_t n, int per, const double *__restrict__ b, double *__restrict__ x,
           double *__restrict__ r) {
    H1PA_CHUNK(i, n) {
        r[i] = b[i];
        x[i] = 0.0;
    }
}

// d = z (the preconditioned r); partials of (d, r)
__global__ void __launch_bounds__(kThreads)
k_cg_first_dir(int64_t n, int per, const double *__restrict__ z, const double *__restrict__ r,
               double *__restrict__ d, double *__restrict__ partials) {
    __shared__ double sh[kThreads];
    double acc[1] = {0.0};
    H1PA_CHUNK(i, n) {
        const double di = z[i];
        d[i] = di;
        acc[0] += di * r[i];
    }
    store_partials<1>(acc, partials, sh);
}

// x += alpha d, r -= alpha q
__global__ void __launch_bounds__(kThreads)
k_cg_update_xr(int64_t n, int per, double *__restrict__ x, double *__restrict__ r,
               const double *__restrict__ q, const double *__restrict__ d, const PcgState *st) {
    if (st->done) return;
    const double alpha = st->nom / st->den;
    H1PA_GROUPS(g, n) {
        double vx[kGrp], vd[kGrp], vr[kGrp], vq[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            const int64_t j = g.lt[k];
            vx[k] = x[j];
            vd[k] = d[j];
            vr[k] = r[j];
            vq[k] = q[j];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (!g.ok[k]) continue;
            x[g.t[k]] = vx[k] + alpha * vd[k];
            r[g.t[k]] = vr[k] - alpha * vq[k];
        }
    }
}

// partials of (a, b) for a solve still running (stop null or 0)
__global__ void __launch_bounds__(kThreads)
k_dot(int64_t n, int per, const double *__restrict__ a, const double *__restrict__ b,
      double *__restrict__ partials, const int *stop) {
    __shared__ double sh[kThreads];
    if (stop && *stop) return;
    double acc[1] = {0.0};
    H1PA_GROUPS(g, n) {
        double va[kGrp], vb[kGrp];
#pragma unroll
        for (int k = 0; k < kGrp; ++k) {
            if (k >= per) break;  // past this thread's entries
            va[k] = a[g.lt[k]];
            vb[k] = b[g.lt[k]];
        }
#pragma unroll
        for (int k = 0; k < kGrp; ++k)
            if (g.ok[k]) acc[0] += va[k] * vb[k];
    }
    store_partials<1>(acc, partials, sh);
}

__global__ void __launch_bounds__(kThreads)
k_dot_fin(int nb, const double *__restrict__ partials, double *__restrict__ out) {
    __shared__ double sh[kThreads];
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x == 0) out[0] = tot[0];
}

// v *= m
__global__ void __launch_bounds__(kThreads)
k_scale(int64_t n, int per, double *__restrict__ v, double m) {
    H1PA_CHUNK(i, n) v[i] *= m;
}

// ---- PowerMethod::EstimateLargestEigenvalue (linalg/operator.cpp:886-935)
// with its state on the device, so the steps run without host round trips.
struct PowerState {
    double nrm, eig;
    int done;
};

// nrm = (v, v) from the partials (first dot of a step)
__global__ void __launch_bounds__(kThreads)
k_power_nrm(int nb, const double *__restrict__ partials, PowerState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x == 0) st->nrm = tot[0];
}

// v *= 1 / sqrt(nrm)
__global__ void __launch_bounds__(kThreads)
k_power_normalize(int64_t n, int per, double *__restrict__ v, const PowerState *st) {
    if (st->done) return;
    const double m = 1.0 / sqrt(st->nrm);
    H1PA_CHUNK(i, n) v[i] *= m;
}

// eigenvalue = (v, D^-1 A v); stop when it changed by less than tol (relative)
__global__ void __launch_bounds__(kThreads)
k_power_eig(int nb, const double *__restrict__ partials, PowerState *st, double tol) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double tot[1];
    sum_partials<1>(partials, nb, sh, tot);
    if (threadIdx.x != 0) return;
    const double enew = tot[0];
    const double diff = fabs((enew - st->eig) / st->eig);
    st->eig = enew;
    if (diff < tol) st->done = 1;
}

// y = 0 + dinv u (OperatorJacobiSmoother::Mult with iterative_mode off)
__global__ void __launch_bounds__(kThreads)
k_jacobi_mult(int64_t n, int per, const double *__restrict__ dinv, const double *__restrict__ u,
              double *__restrict__ y) {
    H1PA_CHUNK(i, n) y[i] = 0.0 + dinv[i] * u[i];
}

// y += a z (Vector::Add, linalg/vector.cpp:436-475)
__global__ void k_axpy(int64_t n, double a, const double *__restrict__ z, double *__restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * z[i];
}

__global__ void k_form_x(int64_t n, const uint8_t *__restrict__ mask, const double *__restrict__ x,
                         double *__restrict__ X) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) X[i] = mask[i] ? x[i] : 0.0;
}

__global__ void k_form_b(int64_t n, const uint8_t *__restrict__ mask, const double *__restrict__ X,
                         const double *__restrict__ b, const double *__restrict__ z,
                         double *__restrict__ B) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) B[i] = mask[i] ? X[i] : b[i] - z[i];
}

}  // namespace

void jacobi_dinv(int64_t n, const double *diag, const uint8_t *mask, double *dinv, hipStream_t s) {
    hipLaunchKernelGGL(k_jacobi_dinv<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                       diag, mask, dinv);
    H1PA_HIP(hipGetLastError());
}

void jacobi_dinv_f32(int64_t n, const double *diag, const uint8_t *mask, float *dinv,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_jacobi_dinv<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                       diag, mask, dinv);
    H1PA_HIP(hipGetLastError());
}

void axpy(int64_t n, double a, const double *z, double *y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, z, y);
    H1PA_HIP(hipGetLastError());
}

void form_x(int64_t n, const uint8_t *mask, const double *x, double *X, hipStream_t s) {
    hipLaunchKernelGGL(k_form_x, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, mask, x, X);
    H1PA_HIP(hipGetLastError());
}

void form_b(int64_t n, const uint8_t *mask, const double *X, const double *b, const double *z,
            double *B, hipStream_t s) {
    hipLaunchKernelGGL(k_form_b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, mask, X, b,
                       z, B);
    H1PA_HIP(hipGetLastError());
}

}  // namespace kern

namespace {

// Element part of the constrained apply, d -> A_E d, E-vector in native
// order (read back through indices_nat); with yL, the interior DOFs go
// straight to yL.  Stops early once `stop` is set.

// Interior DOFs may bypass the E-vector when no constrained DOF is interior:
// the mask is absent or is the space's own boundary mask.
int64_t direct_rows(const Space &sp, const uint8_t *mask) {
    return (!mask || mask == sp.ess_mask.p) ? sp.restr.bubble_begin : sp.ndofs;
}
template <typename T>
void pcg_element_apply(Form &f, int prec, const uint8_t *mask, const T *d, T *ye, T *yL,
                       const int *stop, hipStream_t s) {
    Space &sp = *f.space;
    const bool diff = f.desc.integrators & H1PA_INTEG_DIFFUSION;
    const bool mass = f.desc.integrators & H1PA_INTEG_MASS;
    kern::EOutT<T> out;
    out.ye = ye;
    out.l2n = sp.restr.lex2nat.p;
    out.stop = stop;
    if (yL) {  // interior DOFs straight to the L-vector
        out.yL = yL;
        out.bb = sp.restr.bubble_begin;
    }
    if constexpr (sizeof(T) == 8) {
        // FP64 vectors; prec = BF16_PA takes the diffusion PA data from the
        // form's BF16 shadow (the applies of a mixed-precision cycle)
        if (diff && prec == H1PA_PREC_BF16_PA) {
            H1PA_CHECK(f.lp_kind == 1 && f.pa_diff_lp.p, H1PA_EINVAL,
                       "BF16-PA apply: the form has no BF16 shadow");
            kern::diffusion_apply_d_bf16(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_diff_lp.p,
                                         sp.restr.gather_map.p, mask, d, out, false, s);
        } else if (diff) {
            kern::diffusion_apply(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_diff.p,
                                  sp.restr.gather_map.p, mask, d, out, false, s);
        }
        if (mass)
            kern::mass_apply(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_mass.p, sp.restr.gather_map.p,
                             mask, d, out, diff, s);
    } else {
        const int kind = prec == H1PA_PREC_BF16_PA ? 1 : 0;
        if (diff)
            kern::diffusion_apply_lp(sp.p, f.q1d, sp.ne, f.basis32.p, f.pa_diff_lp.p, kind,
                                     sp.restr.gather_map.p, mask, d, out, false, s);
        if (mass)
            kern::mass_apply_lp(sp.p, f.q1d, sp.ne, f.basis32.p, f.pa_mass_lp.p, kind,
                                sp.restr.gather_map.p, mask, d, out, diff, s);
    }
}

// Jacobi-PCG on the constrained operator with vectors of type T; dinv in T.
template <typename T>
void pcg_run(Form &f, int prec, const uint8_t *mask, const T *dinv, const T *b, T *x, T *ye,
             const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, hipStream_t s,
             PcgWork *ws = nullptr) {
    Space &sp = *f.space;
    const int64_t n = sp.ndofs;
    const kern::Chunks ch = kern::chunks(n);
    const unsigned nb = ch.blocks;
    const int per = ch.per;
    // work vectors: the caller's (FP64, kept between solves) or this solve's own
    StreamBuf<T> r_own, z_own, d_own;
    StreamBuf<double> part_own;
    StreamBuf<PcgState> st_own;
    T *r = nullptr, *z = nullptr, *d = nullptr;
    double *partials = nullptr;
    PcgState *st = nullptr;
    if constexpr (sizeof(T) == 8) {
        if (ws) {
            if (ws->r.n < (size_t)n) {
                ws->r.alloc(n);
                ws->z.alloc(n);
                ws->d.alloc(n);
            }
            if (ws->partials.n < nb) ws->partials.alloc(nb);
            if (ws->state.n < sizeof(PcgState)) ws->state.alloc(sizeof(PcgState));
            r = ws->r.p;
            z = ws->z.p;
            d = ws->d.p;
            partials = ws->partials.p;
            st = reinterpret_cast<PcgState *>(ws->state.p);
        }
    }
    if (!r) {
        r_own.alloc(n, s);
        z_own.alloc(n, s);
        d_own.alloc(n, s);
        part_own.alloc(nb, s);
        st_own.alloc(1, s);
        r = r_own.p;
        z = z_own.p;
        d = d_own.p;
        partials = part_own.p;
        st = st_own.p;
    }
    PcgState h{};
    h.max_iter = o.max_iter;
    H1PA_HIP(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    const int check = o.check_every > 0 ? o.check_every : 8;
    const int *stop = &st->done;
    const dim3 grid(nb), blk(kern::kThreads), one(1);

    auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(kern::k_pcg_init<T>, grid, blk, 0, s, n, per, b, dinv, x, r, z, d,
                       partials);
    hipLaunchKernelGGL(kern::k_pcg_fin_init, one, blk, 0, s, (int)nb, partials, st,
                       o.rel_tol, o.abs_tol);
    const int64_t bb = direct_rows(sp, mask);
    auto apply = [&]() {
        pcg_element_apply<T>(f, prec, mask, d, ye, bb < n ? z : nullptr, stop, s);
        hipLaunchKernelGGL(kern::k_pcg_scatter<T>, grid, blk, 0, s, n, per, bb, sp.restr.offsets.p,
                           sp.restr.indices_nat.p, ye, mask, d, z, partials, st,
                           sp.restr.perm.p, sp.restr.nperm);
        hipLaunchKernelGGL(kern::k_pcg_fin_den, one, blk, 0, s, (int)nb, partials, st);
    };
    apply();
    H1PA_HIP(hipGetLastError());
    for (int it = 1; it <= o.max_iter; ++it) {
        hipLaunchKernelGGL(kern::k_pcg_update<T>, grid, blk, 0, s, n, per, dinv, x, r, z, d,
                           partials, st);
        hipLaunchKernelGGL(kern::k_pcg_fin_update, one, blk, 0, s, (int)nb, partials, st);
        hipLaunchKernelGGL(kern::k_pcg_dir<T>, grid, blk, 0, s, n, per, z, d, st);
        apply();
        H1PA_HIP(hipGetLastError());
        if (it % check == 0 || it == o.max_iter) {
            H1PA_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
            H1PA_HIP(hipStreamSynchronize(s));
            if (h.done) break;
        }
    }
    H1PA_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
    H1PA_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    stats.converged = h.converged;
    stats.iterations = h.done ? h.final_iter : o.max_iter;
    stats.initial_norm = h.nom0 >= 0.0 ? std::sqrt(h.nom0) : h.nom0;
    stats.final_norm = h.betanom >= 0.0 ? std::sqrt(h.betanom) : h.betanom;
    stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    H1PA_CHECK(!h.nonfinite, H1PA_ENONFINITE,
               "PCG: (B r, r) or (A d, d) is not finite (CGSolver: MFEM_VERIFY(IsFinite))");
}

}  // namespace

namespace {
// Pick the reduced-precision diffusion variant for (kind, p, Q) once per
// process on the first large enough form: 2 warm-up + 5 timed launches each.
// H1PA_DIFF_VARIANT=<v> forces variant v (0, 1, 4, 5, 6, 7 or 8).
bool g_lp_tuned[2][10][11] = {};
std::mutex g_lp_tune_mutex;  // forms may be set up on several host threads
void autotune_lp(Form &f, int kind, hipStream_t s) {
    Space &sp = *f.space;
    const int p = sp.p, Q = f.q1d;
    if (const char *v = std::getenv("H1PA_DIFF_VARIANT")) {
        const int fv = std::atoi(v);
        H1PA_CHECK(fv == 0 || fv == 1 || (fv >= 4 && fv < 9), H1PA_EINVAL,
                   "H1PA_DIFF_VARIANT: reduced-precision variants are 0, 1, 4, 5, 6, 7, 8");
        kern::set_diffusion_variant_lp(kind, p, Q, fv);
        return;
    }
    const char *at = std::getenv("H1PA_AUTOTUNE");
    if ((at && at[0] == '0') || (sp.ne < 20000 && sp.ndofs < 500000) || p < 1 || p > 8 || Q > 10)
        return;
    std::lock_guard<std::mutex> lock(g_lp_tune_mutex);
    if (g_lp_tuned[kind][p][Q]) return;
    g_lp_tuned[kind][p][Q] = true;
    DevBuf<float> x(sp.ndofs);  // once per (p, Q) and process: its free may wait
    H1PA_HIP(hipMemsetAsync(x.p, 0, sizeof(float) * sp.ndofs, s));
    hipEvent_t e0, e1;
    H1PA_HIP(hipEventCreate(&e0));
    H1PA_HIP(hipEventCreate(&e1));
    kern::EOutT<float> out;
    out.ye = f.ework32.p;
    out.l2n = sp.restr.lex2nat.p;
    int best = 0;
    float best_ms = 1e30f;
    for (int v : {0, 1, 4, 5, 6, 7, 8}) {
        for (int w = 0; w < 2; ++w)
            kern::diffusion_apply_lp_variant(v, p, Q, sp.ne, f.basis32.p, f.pa_diff_lp.p, kind,
                                             sp.restr.gather_map.p, nullptr, x.p, out, false, s);
        H1PA_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 5; ++r)
            kern::diffusion_apply_lp_variant(v, p, Q, sp.ne, f.basis32.p, f.pa_diff_lp.p, kind,
                                             sp.restr.gather_map.p, nullptr, x.p, out, false, s);
        H1PA_HIP(hipEventRecord(e1, s));
        H1PA_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        H1PA_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    kern::set_diffusion_variant_lp(kind, p, Q, best);
}
}  // namespace

void require_no_batch_refs(const Form &f, const char *what) {
    const int n = f.batch_refs.load();
    H1PA_CHECK(n == 0, H1PA_EINVAL,
               std::string(what) + ": the form's device data are in use by " + std::to_string(n) +
                   " batch solve(s) not yet retired (h1pa_batch_poll) or multigrid "
                   "hierarchies not yet destroyed (h1pa_multigrid_destroy)");
}

void ensure_low_precision(Form &f, int precision, hipStream_t s) {
    const int kind = precision == H1PA_PREC_BF16_PA ? 1 : 0;
    if (f.lp_kind == kind) return;
    // replacing the shadow frees the buffers live batch slots read
    if (f.lp_kind >= 0) require_no_batch_refs(f, "reduced-precision shadow of another kind");
    Space &sp = *f.space;
    auto convert = [&](const DevBuf<double> &src, DevBuf<uint32_t> &dst) {
        if (!src.p) {
            dst.release();
            return;
        }
        const int64_t n = (int64_t)src.n;
        const unsigned g = (unsigned)((n + 255) / 256);
        if (kind == 0) {
            dst.alloc((size_t)n);
            hipLaunchKernelGGL(kern::k_to_float, dim3(g), dim3(256), 0, s, n, src.p,
                               reinterpret_cast<float *>(dst.p));
        } else {
            dst.alloc((size_t)(n + 1) / 2);
            hipLaunchKernelGGL(kern::k_to_bf16, dim3(g), dim3(256), 0, s, n, src.p,
                               reinterpret_cast<uint16_t *>(dst.p));
        }
        H1PA_HIP(hipGetLastError());
    };
    if (!f.basis32.p) {
        const int64_t nb = (int64_t)f.basis.n;
        f.basis32.alloc((size_t)nb);
        hipLaunchKernelGGL(kern::k_to_float, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s,
                           nb, f.basis.p, f.basis32.p);
        H1PA_HIP(hipGetLastError());
    }
    f.lp_kind = -1;
    convert(f.pa_diff, f.pa_diff_lp);
    convert(f.pa_mass, f.pa_mass_lp);
    if (f.ework32.n < (size_t)(sp.ne * sp.dof)) f.ework32.alloc((size_t)(sp.ne * sp.dof));
    f.lp_kind = kind;
    if (f.pa_diff_lp.p) autotune_lp(f, kind, s);
}

void form_mult_lp(Form &f, int precision, const float *x, float *y, hipStream_t s) {
    ensure_low_precision(f, precision, s);
    Space &sp = *f.space;
    pcg_element_apply<float>(f, precision, nullptr, x, f.ework32.p, nullptr, nullptr, s);
    hipLaunchKernelGGL(kern::k_scatter_set<float>, dim3((unsigned)((sp.ndofs + 255) / 256)),
                       dim3(256), 0, s, sp.ndofs, sp.restr.offsets.p, sp.restr.indices_nat.p,
                       f.ework32.p, y, sp.restr.perm.p, sp.restr.nperm);
    H1PA_HIP(hipGetLastError());
}

void pcg_solve(Form &f, const uint8_t *mask, const double *diag, const double *b, double *x,
               const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, hipStream_t s) {
    Space &sp = *f.space;
    StreamBuf<double> dinv(sp.ndofs, s);
    kern::jacobi_dinv(sp.ndofs, diag, mask, dinv.p, s);
    pcg_run<double>(f, H1PA_PREC_FP64, mask, dinv.p, b, x, f.ework0.p, o, stats, s);
}

void pcg_solve_dinv(Form &f, const uint8_t *mask, const double *dinv, const double *b, double *x,
                    const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, PcgWork &w, hipStream_t s,
                    int prec) {
    pcg_run<double>(f, prec, mask, dinv, b, x, f.ework0.p, o, stats, s, &w);
}

// u = A v on the DIAG_ONE constrained operator; interior rows go straight
// into u when the mask allows it.
void constrained_apply(Form &f, const uint8_t *mask, const double *v, double *u,
                       const int *stop, hipStream_t s, int prec) {
    Space &sp = *f.space;
    const int64_t bb = direct_rows(sp, mask);
    pcg_element_apply<double>(f, prec, mask, v, f.ework0.p, bb < sp.ndofs ? u : nullptr, stop,
                              s);
    kern::scatter_constrained(bb, sp.restr.offsets.p, sp.restr.indices_nat.p, f.ework0.p, mask,
                              v, u, sp.restr.perm.p, sp.restr.nperm, s);
}

// ---- OperatorChebyshevSmoother ----------------------------------------------
namespace {

// Vector::Randomize(seed) (linalg/vector.cpp:955-966): srand(seed), then
// rand() / (RAND_MAX + 1) per entry, drawn on the device.  glibc's rand() on
// its 128-byte state is the additive generator h[k] = h[k-31] + h[k-3]
// (mod 2^32) returning h[k] >> 1; the library's reentrant interface
// (initstate_r, with a private state, so the caller's rand() sequence is
// left alone) gives the 31 words h[k-31 .. k-1] before the first draw.  The
// recurrence is linear over the integers mod 2^32, so the state c L draws
// later is A^L applied c times (A the 31 x 31 step matrix, A^L by repeated
// squaring): the host computes one start state per chunk of L draws and each
// device thread runs the recurrence over its chunk -- the same draws as n
// calls of rand(), without generating and uploading n values on the host.
using Mat31 = std::array<uint32_t, 31 * 31>;

Mat31 mat31_mul(const Mat31 &a, const Mat31 &b) {
    Mat31 c{};
    for (int i = 0; i < 31; ++i)
        for (int k = 0; k < 31; ++k) {
            const uint32_t aik = a[31 * i + k];
            if (!aik) continue;
            for (int j = 0; j < 31; ++j) c[31 * i + j] += aik * b[31 * k + j];
        }
    return c;
}

__global__ void k_glibc_rand(int64_t n, int64_t L, const uint32_t *__restrict__ start,
                             double *__restrict__ v) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t i = c * L;
    if (i >= n) return;
    const int64_t e = i + L < n ? i + L : n;
    uint32_t h[31];  // h[j] = h[k - 31 + j] at the start of each block of 31 draws
#pragma unroll
    for (int j = 0; j < 31; ++j) h[j] = start[31 * c + j];
    for (; i < e; i += 31) {
#pragma unroll
        for (int j = 0; j < 31; ++j) {
            h[j] += h[(j + 28) % 31];  // h[k] = h[k - 31] + h[k - 3]
            if (i + j < e) v[i + j] = (double)(h[j] >> 1) * (1.0 / 2147483648.0);
        }
    }
}

void randomize_like_mfem(int seed, int64_t n, double *v, DevBuf<uint32_t> &d_starts,
                         hipStream_t s) {
    if (n <= 0) return;
    struct random_data rd;
    std::memset(&rd, 0, sizeof(rd));
    char state[128];
    H1PA_CHECK(initstate_r((unsigned)seed, state, sizeof(state), &rd) == 0, H1PA_EINVAL,
               "chebyshev: initstate_r failed");
    const int f = (int)(rd.fptr - rd.state), r = (int)(rd.rptr - rd.state);
    H1PA_CHECK(rd.rand_type == 3 && rd.rand_deg == 31 && rd.rand_sep == 3 &&
                   ((f - r) % 31 + 31) % 31 == 3,
               H1PA_EUNSUP, "chebyshev: unexpected C library generator");
    std::array<uint32_t, 31> x;  // h[k - 31 .. k - 1]
    for (int j = 0; j < 31; ++j) x[j] = (uint32_t)rd.state[(f + j) % 31];
    // chunks of L draws, at most 4096 of them
    const int64_t L = std::max<int64_t>(31 * 64, (n + 4095) / 4096);
    const int64_t T = (n + L - 1) / L;
    Mat31 step{}, AL{};
    for (int i = 0; i < 30; ++i) step[31 * i + i + 1] = 1;
    step[31 * 30 + 0] = 1;
    step[31 * 30 + 28] = 1;
    for (int i = 0; i < 31; ++i) AL[31 * i + i] = 1;
    for (int64_t e = L; e; e >>= 1) {  // AL = step^L
        if (e & 1) AL = mat31_mul(AL, step);
        if (e > 1) step = mat31_mul(step, step);
    }
    std::vector<uint32_t> starts((size_t)T * 31);
    for (int64_t c = 0; c < T; ++c) {
        std::copy(x.begin(), x.end(), starts.begin() + 31 * c);
        std::array<uint32_t, 31> y{};
        for (int i = 0; i < 31; ++i) {
            uint32_t acc = 0;
            for (int k = 0; k < 31; ++k) acc += AL[31 * i + k] * x[k];
            y[i] = acc;
        }
        x = y;
    }
    d_starts.upload(starts.data(), starts.size(), s);
    hipLaunchKernelGGL(k_glibc_rand, dim3((unsigned)((T + 63) / 64)), dim3(64), 0, s, n, L,
                       d_starts.p, v);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));  // the host copy `starts` is freed on return
}

// Coefficients of OperatorChebyshevSmoother::Setup (solvers.cpp:552-628):
// 1 - lambda p(lambda) is the Chebyshev polynomial of the first kind of the
// given order on [0.3, 1.2] max_eig, scaled to 1 at lambda = 0.
void chebyshev_coeffs(int order, double max_eig, double *c) {
    const double upper = 1.2 * max_eig, lower = 0.3 * max_eig;
    const double th = 0.5 * (upper + lower), de = 0.5 * (upper - lower);
    switch (order) {
        case 1: c[0] = 1.0 / th; break;
        case 2: {
            const double a = 1.0 / (std::pow(de, 2) - 2 * std::pow(th, 2));
            c[0] = -4 * th * a;
            c[1] = 2 * a;
            break;
        }
        case 3: {
            const double a = 3 * std::pow(de, 2), b2 = std::pow(th, 2);
            const double g = 1.0 / (-4 * std::pow(th, 3) + th * a);
            c[0] = g * (a - 12 * b2);
            c[1] = 12 / (a - 4 * b2);
            c[2] = -4 * g;
            break;
        }
        case 4: {
            const double d2 = std::pow(de, 2), t2 = std::pow(th, 2), e8 = 8 * d2;
            const double g = 1.0 / (std::pow(de, 4) + 8 * std::pow(th, 4) - t2 * e8);
            c[0] = g * (32 * std::pow(th, 3) - 16 * th * d2);
            c[1] = g * (-48 * t2 + e8);
            c[2] = 32 * th * g;
            c[3] = -8 * g;
            break;
        }
        case 5: {
            const double d4x5 = 5 * std::pow(de, 4), t4 = std::pow(th, 4), t2 = std::pow(th, 2);
            const double d2 = std::pow(de, 2), d2x60 = 60 * d2, d2x20 = 20 * d2;
            const double g = 1.0 / (16 * std::pow(th, 5) - std::pow(th, 3) * d2x20 + th * d4x5);
            const double t2x160 = 160 * t2;
            const double h = 1.0 / (d4x5 + 16 * t4 - t2 * d2x20);
            c[0] = g * (d4x5 + 80 * t4 - t2 * d2x60);
            c[1] = h * (d2x60 - t2x160);
            c[2] = g * (-d2x20 + t2x160);
            c[3] = -80 * h;
            c[4] = 16 * g;
            break;
        }
        default: throw Error{H1PA_EUNSUP, "chebyshev: order must be 1..5"};
    }
}

// y = sum_k c_k (D^-1 A)^k D^-1 x
void chebyshev_apply(Chebyshev &c, const double *x, double *y, const int *stop, hipStream_t s) {
    Form &f = *c.f;
    Space &sp = *f.space;
    const int64_t n = sp.ndofs;
    const kern::Chunks ch = kern::chunks(n);
    const int64_t bb = direct_rows(sp, c.mask);
    hipLaunchKernelGGL(kern::k_cheb_first, dim3(ch.blocks), dim3(kern::kThreads), 0, s, n, ch.per,
                       c.dinv.p, x, c.t.p, y, c.coeffs[0], stop);
    for (int k = 1; k < c.order; ++k) {
        pcg_element_apply<double>(f, c.prec, c.mask, c.t.p, f.ework0.p,
                                  bb < n ? c.u.p : nullptr, stop, s);
        hipLaunchKernelGGL(kern::k_cheb_scatter, dim3(ch.blocks), dim3(kern::kThreads), 0, s, n,
                           ch.per, bb, sp.restr.offsets.p, sp.restr.indices_nat.p, f.ework0.p,
                           c.mask, c.dinv.p, c.u.p, c.t.p, y, c.coeffs[k], stop, sp.restr.perm.p,
                           sp.restr.nperm);
    }
    H1PA_HIP(hipGetLastError());
}

}  // namespace

void chebyshev_setup(Chebyshev &c, const double *d_diag, double max_eig, int power_iterations,
                     double power_tolerance, int power_seed, hipStream_t s) {
    H1PA_CHECK(c.order >= 1 && c.order <= 5, H1PA_EUNSUP, "chebyshev: order must be 1..5");
    Form &f = *c.f;
    Space &sp = *f.space;
    const int64_t n = sp.ndofs;
    c.dinv.alloc(n);
    c.t.alloc(n);
    c.u.alloc(n);
    if (!d_diag) {  // the assembled diagonal, through u (free until the first Mult)
        operator_diagonal(f, c.u.p, s);
        d_diag = c.u.p;
    }
    kern::jacobi_dinv(n, d_diag, c.mask, c.dinv.p, s);
    if (max_eig > 0.0) {
        c.max_eig = max_eig;
    } else {
        // PowerMethod::EstimateLargestEigenvalue (linalg/operator.cpp:886-935)
        // on diag^-1 A from Vector::Randomize(power_seed)
        H1PA_CHECK(power_seed != 0, H1PA_EINVAL, "chebyshev: invalid seed");
        H1PA_CHECK(power_iterations >= 1, H1PA_EINVAL, "chebyshev: power_iterations must be >= 1");
        // each step: v /= |v|, u = A v, w = D^-1 u, eig = (v, w), v <- w; the
        // state (norm, estimate, stop) stays on the device and every launch
        // after the stop does nothing, so the host waits once, at the end
        const kern::Chunks ch = kern::chunks(n);
        // v0 in t, v1 in w, the state in the partials' tail
        c.w.alloc(n);
        c.partials.alloc(ch.blocks + (sizeof(kern::PowerState) + 7) / 8);
        double *const partials = c.partials.p;
        kern::PowerState *const st = reinterpret_cast<kern::PowerState *>(c.partials.p + ch.blocks);
        double *const v0 = c.t.p, *const v1 = c.w.p;
        randomize_like_mfem(power_seed, n, v0, c.starts, s);
        const kern::PowerState h0{0.0, 1.0, 0};
        H1PA_HIP(hipMemcpyAsync(st, &h0, sizeof(h0), hipMemcpyHostToDevice, s));
        const int *stop = &st->done;
        const dim3 grid(ch.blocks), blk(kern::kThreads), one(1);
        double *a = v0, *b = v1;
        for (int it = 0; it < power_iterations; ++it) {
            hipLaunchKernelGGL(kern::k_dot, grid, blk, 0, s, n, ch.per, a, a, partials, stop);
            hipLaunchKernelGGL(kern::k_power_nrm, one, blk, 0, s, (int)ch.blocks, partials, st);
            hipLaunchKernelGGL(kern::k_power_normalize, grid, blk, 0, s, n, ch.per, a, st);
            constrained_apply(*c.f, c.mask, a, c.u.p, stop, s, c.prec);
            hipLaunchKernelGGL(kern::k_jacobi_mult, grid, blk, 0, s, n, ch.per, c.dinv.p, c.u.p,
                               b);
            hipLaunchKernelGGL(kern::k_dot, grid, blk, 0, s, n, ch.per, a, b, partials, stop);
            hipLaunchKernelGGL(kern::k_power_eig, one, blk, 0, s, (int)ch.blocks, partials, st,
                               power_tolerance);
            H1PA_HIP(hipGetLastError());
            std::swap(a, b);
        }
        kern::PowerState hs{};
        H1PA_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
        H1PA_HIP(hipStreamSynchronize(s));
        c.max_eig = hs.eig;
    }
    chebyshev_coeffs(c.order, c.max_eig, c.coeffs);
}

void chebyshev_mult(Chebyshev &c, const double *x, double *y, hipStream_t s) {
    chebyshev_apply(c, x, y, nullptr, s);
}

// CGSolver::Mult (solvers.cpp:869-1050, iterative_mode = false) with a
// preconditioner z = M r on f's DIAG_ONE constrained operator: the Jacobi-PCG
// loop with z = M r computed by M instead of fused into the update.  The host
// reads the state every `check` iterations.
void pcg_preconditioned(Form &f, const uint8_t *mask, const Precond &M, const double *b,
                        double *x, const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, int check,
                        hipStream_t s, PcgWork *ws) {
    Space &sp = *f.space;
    const int64_t n = sp.ndofs;
    const kern::Chunks ch = kern::chunks(n);
    const unsigned nb = ch.blocks;
    const int per = ch.per;
    // work vectors: the caller's (kept between solves: no allocation, so no
    // device-wide wait in a free) or this solve's own
    PcgWork own;
    PcgWork &w = ws ? *ws : own;
    if (w.r.n < (size_t)n) {
        w.r.alloc(n);
        w.q.alloc(n);
        w.d.alloc(n);
        w.z.alloc(n);
    }
    if (w.partials.n < nb) w.partials.alloc(nb);
    if (w.state.n < sizeof(PcgState)) w.state.alloc(sizeof(PcgState));
    double *r = w.r.p, *q = w.q.p, *d = w.d.p, *z = w.z.p, *partials = w.partials.p;
    PcgState *st = reinterpret_cast<PcgState *>(w.state.p);
    PcgState h{};
    h.max_iter = o.max_iter;
    H1PA_HIP(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    check = check > 0 ? check : 8;
    const int *stop = &st->done;
    const dim3 grid(nb), blk(kern::kThreads), one(1);
    const int64_t bb = direct_rows(sp, mask);

    auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(kern::k_cg_start, grid, blk, 0, s, n, per, b, x, r);
    M(r, z, stop, s);
    hipLaunchKernelGGL(kern::k_cg_first_dir, grid, blk, 0, s, n, per, z, r, d, partials);
    hipLaunchKernelGGL(kern::k_pcg_fin_init, one, blk, 0, s, (int)nb, partials, st, o.rel_tol,
                       o.abs_tol);
    auto apply = [&]() {  // q = A d, den = (d, q)
        pcg_element_apply<double>(f, H1PA_PREC_FP64, mask, d, f.ework0.p,
                                  bb < n ? q : nullptr, stop, s);
        hipLaunchKernelGGL(kern::k_pcg_scatter<double>, grid, blk, 0, s, n, per, bb,
                           sp.restr.offsets.p, sp.restr.indices_nat.p, f.ework0.p, mask, d,
                           q, partials, st, sp.restr.perm.p, sp.restr.nperm);
        hipLaunchKernelGGL(kern::k_pcg_fin_den, one, blk, 0, s, (int)nb, partials, st);
    };
    apply();
    H1PA_HIP(hipGetLastError());
    for (int it = 1; it <= o.max_iter; ++it) {
        hipLaunchKernelGGL(kern::k_cg_update_xr, grid, blk, 0, s, n, per, x, r, q, d, st);
        M(r, z, stop, s);
        hipLaunchKernelGGL(kern::k_dot, grid, blk, 0, s, n, per, r, z, partials, stop);
        hipLaunchKernelGGL(kern::k_pcg_fin_update, one, blk, 0, s, (int)nb, partials, st);
        hipLaunchKernelGGL(kern::k_pcg_dir<double>, grid, blk, 0, s, n, per, z, d, st);
        apply();
        H1PA_HIP(hipGetLastError());
        if (it % check == 0 || it == o.max_iter) {
            H1PA_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
            H1PA_HIP(hipStreamSynchronize(s));
            if (h.done) break;
        }
    }
    H1PA_HIP(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
    H1PA_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    stats.converged = h.converged;
    stats.iterations = h.done ? h.final_iter : o.max_iter;
    stats.initial_norm = h.nom0 >= 0.0 ? std::sqrt(h.nom0) : h.nom0;
    stats.final_norm = h.betanom >= 0.0 ? std::sqrt(h.betanom) : h.betanom;
    stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    H1PA_CHECK(!h.nonfinite, H1PA_ENONFINITE,
               "PCG: (B r, r) or (A d, d) is not finite (CGSolver: MFEM_VERIFY(IsFinite))");
}

// CGSolver with the Chebyshev smoother as preconditioner.
void pcg_chebyshev(Chebyshev &c, const double *b, double *x, const h1pa_pcg_opts &o,
                   h1pa_pcg_stats &stats, hipStream_t s) {
    const Precond M = [&c](const double *r, double *z, const int *stop, hipStream_t st) {
        chebyshev_apply(c, r, z, stop, st);
    };
    pcg_preconditioned(*c.f, c.mask, M, b, x, o, stats, o.check_every, s, nullptr);
}

void ir_solve(Form &f, const uint8_t *mask, const double *diag, const double *b, double *x,
              const h1pa_ir_opts &o, h1pa_ir_stats &stats, hipStream_t s) {
    Space &sp = *f.space;
    const int64_t n = sp.ndofs;
    const unsigned nb256 = (unsigned)((n + 255) / 256);
    const kern::Chunks ch = kern::chunks(n);
    const unsigned nb = ch.blocks;
    const int per = ch.per;
    const unsigned nbx = std::min<unsigned>(nb, 1024);
    ensure_low_precision(f, o.inner_precision, s);
    StreamBuf<float> dinv(n, s), r(n, s), xs(n, s), z(n, s), d(n, s), q(n, s);
    StreamBuf<double> dinv64(n, s), r64(n, s);
    StreamBuf<double> partials(3 * (size_t)nb, s);
    StreamBuf<kern::IrState> st(1, s);
    hipLaunchKernelGGL(kern::k_jacobi_dinv<float>, dim3(nb256), dim3(256), 0, s, n, diag, mask,
                       dinv.p);
    hipLaunchKernelGGL(kern::k_jacobi_dinv<double>, dim3(nb256), dim3(256), 0, s, n, diag, mask,
                       dinv64.p);
    kern::IrState h{};
    h.tol2 = o.rel_tol * o.rel_tol;
    h.delta2 = o.delta * o.delta;
    h.max_inner = o.max_inner;
    h.max_outer = o.max_outer;
    H1PA_HIP(hipMemcpyAsync(st.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    const int *done = &st.p->done;
    const int *nru = &st.p->nru;
    const int64_t bb = direct_rows(sp, mask);
    const dim3 grid(nb), blk(kern::kThreads), one(1);
    auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(kern::k_ir_init, grid, blk, 0, s, n, per, b, dinv64.p, dinv.p, x, r64.p, r.p,
                       xs.p, z.p, d.p, partials.p);
    hipLaunchKernelGGL(kern::k_ir_fin_init, one, blk, 0, s, (int)nb, partials.p, st.p);
    H1PA_HIP(hipGetLastError());
    const int limit = o.max_inner + o.max_outer + 1;
    for (int it = 1; it <= limit; ++it) {
        // q = A d with the reduced-precision operator
        pcg_element_apply<float>(f, o.inner_precision, mask, d.p, f.ework32.p,
                                 bb < n ? q.p : nullptr, done, s);
        hipLaunchKernelGGL(kern::k_ir_scatter, grid, blk, 0, s, n, per, bb, sp.restr.offsets.p,
                           sp.restr.indices_nat.p, f.ework32.p, mask, d.p, q.p, partials.p, st.p,
                           sp.restr.perm.p, sp.restr.nperm);
        hipLaunchKernelGGL(kern::k_ir_fin_den, one, blk, 0, s, (int)nb, partials.p, st.p);
        hipLaunchKernelGGL(kern::k_ir_update, grid, blk, 0, s, n, per, dinv.p, xs.p, r.p, z.p, q.p,
                           d.p, partials.p, st.p);
        hipLaunchKernelGGL(kern::k_ir_fin_update, one, blk, 0, s, (int)nb, partials.p, st.p);
        hipLaunchKernelGGL(kern::k_ir_ru_x, dim3(nbx), blk, 0, s, n, xs.p, x, st.p);
        pcg_element_apply<double>(f, H1PA_PREC_FP64, mask, x, f.ework0.p,
                                  bb < n ? r64.p : nullptr, nru, s);
        hipLaunchKernelGGL(kern::k_ir_ru_res, grid, blk, 0, s, n, per, bb, sp.restr.offsets.p,
                           sp.restr.indices_nat.p, f.ework0.p, mask, b, x, dinv64.p, dinv.p,
                           r64.p, r.p, z.p, partials.p, st.p, sp.restr.perm.p, sp.restr.nperm);
        hipLaunchKernelGGL(kern::k_ir_fin_ru, one, blk, 0, s, (int)nb, partials.p, st.p);
        hipLaunchKernelGGL(kern::k_ir_dir, grid, blk, 0, s, n, per, z.p, d.p, st.p);
        H1PA_HIP(hipGetLastError());
        if (it % 8 == 0 || it == limit) {
            H1PA_HIP(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
            H1PA_HIP(hipStreamSynchronize(s));
            if (h.done) break;
        }
    }
    H1PA_HIP(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
    H1PA_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    stats.converged = h.converged;
    stats.outer_iterations = h.outer;
    stats.inner_iterations = h.iter;
    stats.final_norm = std::sqrt(std::max(h.rho64, 0.0));
    stats.rel_residual = h.b2 > 0 ? std::sqrt(h.r2 / h.b2) : 0.0;
    stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    H1PA_CHECK(!h.nonfinite, H1PA_ENONFINITE,
               "mixed solve: a residual or search-direction product is not finite");
}

}  // namespace h1pa
