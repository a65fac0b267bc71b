// batch.hip -- many independent solves advanced together.
//
// No MFEM counterpart: MFEM solves one system per CGSolver::Mult
// (linalg/solvers.cpp:869-1050).  Here every live solve ("slot") owns its
// vectors and a small device state machine; one replay of a captured HIP
// graph advances all of them by K iterations:
//
//   init      slots in INIT: x = 0, r = b, z = dinv r, d = z, (z, r)
//   K x {     element kernels, one launch per (p, Q) group, over all slots'
//               elements (a work list of (slot, first element) blocks)
//             scatter   z = A d (DIAG_ONE rows copy d), den = (d, z)
//             update    x += a d, r -= a z, z = dinv r, (r, z), stop test
//             [mixed, last iteration of the replay: refinement step of the
//              slots whose inner residual fell by delta since their last one
//              (they wait in PH_RU until then): x += xs, r64 = b - A x in
//              FP64, r = (float) r64]
//             dir       d = z + b d                                       }
//
// FP64 mode is MFEM's CGSolver on each slot.  Mixed mode is an FP32 inner CG
// (FP32 or BF16-stored PA data) with FP64 reliable updates, stopping on
// CGSolver's measure taken on the FP64 true residual.  Every kernel skips,
// block-uniformly, the slots that are not in the phase it serves.  Dot
// products are reduced per slot over fixed chunks of that slot's DOFs: each
// chunk's block stores its partial, and a finalize kernel (one block per
// running slot) sums them in index order and applies the step's scalar
// logic.  So a slot's iterates do not depend on which other slots share the
// batch: a solve gives bitwise the same result alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "apply_body.hpp"
#include "row_sum.hpp"

namespace h1pa {

enum : int {
    PH_INNER = 0,   // iterations running
    PH_RU = 1,      // mixed: FP64 refinement step due this iteration
    PH_DONE = 2,    // finished (converged or not)
    PH_INIT = 4,    // freshly added
    PH_EMPTY = 5    // slot retired
};

struct SlotState {
    double nom, nom0, den, betanom, beta, r0;  // PCG scalars (mixed: nom = inner (r, z))
    double rho_ref, rho64, rho64_0, r2, b2;    // mixed: refinement bookkeeping
    int iter, in_loop, max_iter, converged, final_iter;
    int phase, outer;
    int nonfinite;  // a (B r, r) or (A d, d) was not finite: retired with a failed status
};

struct SlotDesc {
    const void *pa;        // inner PA data (FP64, FP32 or BF16)
    const double *pa64;    // FP64 PA data (refinement residual)
    const int *gmap, *offsets, *indices, *perm, *l2n;
    const uint8_t *mask;
    const double *b;
    double *x;
    void *r, *z, *d, *dinv, *xi, *ye;  // inner vectors (type T of the batch); xi: FP64
                                       // mode the solution x, mixed the update xs
    double *r64, *dinv64, *ye64;       // mixed: FP64 residual, Jacobi, E-vector
    int64_t ne, n, nperm;
    int64_t bb;                        // first row written directly by the element kernel
                                       // (interior DOFs; n when none)
    int part0, nchunks;
    double rel_tol, abs_tol;           // CGSolver settings (mixed: rel_tol on the FP64 measure)
    double delta;                      // mixed: refinement trigger
    int max_outer;
    int pad_;
};

struct VecTask {
    int slot, chunk;
};
struct ElemTask {
    int slot;
    int e0;
};

namespace kern {
namespace {

constexpr int kT = 256;          // threads per vector block
constexpr int kPer = 4;          // DOFs per thread per chunk
constexpr int kChunk = kT * kPer;

// Block sum of v over the kT threads, the same value in every thread: a
// butterfly within each 64-lane wave (lane 0's total is kept), then the four
// wave totals in a fixed order.
static_assert(kT == 256, "bsum assumes four 64-lane waves");
__device__ inline double bsum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

// partials[NV (part0 + chunk) + k] = block sum of acc[k]: plain stores, no
// fence (a device-scope fence writes back and invalidates the L2 here); the
// finalize kernels read them after the kernel boundary.
template <int NV>
__device__ inline void store_chunk(const double (&acc)[NV], const SlotDesc &sd, int chunk,
                                   double *partials, double *sh) {
    for (int k = 0; k < NV; ++k) {
        const double v = bsum(acc[k], sh);
        if (threadIdx.x == 0) partials[(int64_t)NV * (sd.part0 + chunk) + k] = v;
    }
}

// Finalize block: the slot's totals, chunk partials summed in index order.
template <int NV>
__device__ inline void slot_totals(const SlotDesc &sd, const double *partials, double *sh,
                                   double (&tot)[NV]) {
    const double *p = partials + (int64_t)NV * sd.part0;
    constexpr int U8 = 8;  // independent loads in flight per thread
    for (int k = 0; k < NV; ++k) {
        double v[U8];
#pragma unroll
        for (int u = 0; u < U8; ++u) v[u] = 0.0;
        for (int i0 = 0; i0 < sd.nchunks; i0 += U8 * kT) {
#pragma unroll
            for (int u = 0; u < U8; ++u) {
                const int i = i0 + u * kT + threadIdx.x;
                if (i < sd.nchunks) v[u] += p[(int64_t)NV * i + k];
            }
        }
        double w = 0.0;
#pragma unroll
        for (int u = 0; u < U8; ++u) w += v[u];
        tot[k] = bsum(w, sh);
    }
}

template <typename T>
__device__ inline T *vp(void *p) { return reinterpret_cast<T *>(p); }

// The kPer DOFs t[k] = chunk kChunk + k kT + tid of this thread's share of a
// chunk (ok[k]: t[k] < n).  Kernels load every operand of all kPer DOFs at
// the clamped index lt[k] (in bounds for every k) before the first store, so
// the loads of a share go out together instead of one DOF's round trip after
// another; stores and sums then run over the ok entries in k order.
struct Share {
    int64_t t[kPer], lt[kPer];
    bool ok[kPer];
};
__device__ __forceinline__ Share share_of(int chunk, int64_t n) {
    Share h;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t t = (int64_t)chunk * kChunk + k * kT + threadIdx.x;
        h.t[k] = t;
        h.ok[k] = t < n;
        h.lt[k] = t < n ? t : n - 1;
    }
    return h;
}

// The DOFs of chunk tk.chunk that this thread owns (kPer of them, kT apart),
// one after another (kernels whose DOFs branch on their kind).
#define H1PA_CHUNK_LOOP(i)                                                                 \
    for (int64_t i = (int64_t)tk.chunk * kChunk + threadIdx.x,                             \
                 i##_end = min(sd.n, (int64_t)(tk.chunk + 1) * kChunk);                    \
         i < i##_end; i += kT)

// Descriptor pointer of type T as a global-memory pointer.
template <typename T>
__device__ __forceinline__ T *gv(void *p) { return global_ptr(reinterpret_cast<T *>(p)); }
template <typename T>
__device__ __forceinline__ const T *gv(const void *p) {
    return global_ptr(reinterpret_cast<const T *>(p));
}

// Producer kernels run one block per (slot, chunk) work item and skip the
// slots not in the phase they serve; each finalize kernel runs one block per
// running slot (runlist) and applies the scalar logic of that step.
#define H1PA_TASK_PROLOGUE(PHASE_OK)                                              \
    const VecTask tk = tasks[blockIdx.x];                                         \
    const SlotDesc &sd = desc[tk.slot];                                           \
    const int ph_ = states[tk.slot].phase;                                        \
    if (!(PHASE_OK)) return;
#define H1PA_FIN_PROLOGUE(PHASE_OK)                                               \
    const int slot = runlist[blockIdx.x];                                         \
    const SlotDesc &sd = desc[slot];                                              \
    SlotState *st = states + slot;                                                \
    const int ph_ = st->phase;                                                    \
    if (!(PHASE_OK)) return;

// ---- FP64 mode: CGSolver on each slot ---------------------------------------
__global__ void __launch_bounds__(kT)
kb_init(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
        const SlotState *__restrict__ states, double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_TASK_PROLOGUE(ph_ == PH_INIT)
    double *r = gv<double>(sd.r), *z = gv<double>(sd.z), *d = gv<double>(sd.d);
    double *x = global_ptr(sd.x);
    const double *dinv = gv<double>(sd.dinv), *b = global_ptr(sd.b);
    const Share h = share_of(tk.chunk, sd.n);
    double vb[kPer], vi[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        vb[k] = b[h.lt[k]];
        vi[k] = dinv[h.lt[k]];
    }
    double acc[1] = {0.0};
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (!h.ok[k]) continue;
        const int64_t i = h.t[k];
        const double ri = vb[k];
        const double zi = vi[k] * ri;
        r[i] = ri;
        x[i] = 0.0;
        z[i] = zi;
        d[i] = zi;
        acc[0] += zi * ri;
    }
    store_chunk<1>(acc, sd, tk.chunk, partials, sh);
}

__global__ void __launch_bounds__(kT)
kb_fin_init(const int *__restrict__ runlist, const SlotDesc *__restrict__ desc,
            SlotState *__restrict__ states, const double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_FIN_PROLOGUE(ph_ == PH_INIT)
    double tot[1];
    slot_totals<1>(sd, partials, sh, tot);
    if (threadIdx.x != 0) return;
    const double nom = tot[0];
    st->nom = st->nom0 = st->betanom = nom;
    st->iter = 1;
    st->in_loop = 0;
    st->final_iter = 0;
    double r0 = nom * sd.rel_tol * sd.rel_tol;
    if (sd.abs_tol * sd.abs_tol > r0) r0 = sd.abs_tol * sd.abs_tol;
    st->r0 = r0;
    int next = PH_INNER;
    if (!isfinite(nom)) {  // MFEM_VERIFY(IsFinite(nom)) in CGSolver::Mult
        st->nonfinite = 1;
        next = PH_DONE;
    } else if (nom < 0.0) {
        next = PH_DONE;
    } else if (nom <= r0) {
        st->converged = 1;
        next = PH_DONE;
    }
    st->phase = next;
}

// ---- mixed mode: FP32 inner CG with FP64 reliable updates ---------------------
__global__ void __launch_bounds__(kT)
kbm_init(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
         const SlotState *__restrict__ states, double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_TASK_PROLOGUE(ph_ == PH_INIT)
    float *r = gv<float>(sd.r), *z = gv<float>(sd.z), *d = gv<float>(sd.d);
    float *xs = gv<float>(sd.xi);
    const float *dinv = gv<float>(sd.dinv);
    double *x = global_ptr(sd.x), *r64 = global_ptr(sd.r64);
    const double *b = global_ptr(sd.b), *dinv64 = global_ptr(sd.dinv64);
    const Share h = share_of(tk.chunk, sd.n);
    double vb[kPer], vi64[kPer];
    float vi[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        vb[k] = b[h.lt[k]];
        vi[k] = dinv[h.lt[k]];
        vi64[k] = dinv64[h.lt[k]];
    }
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (!h.ok[k]) continue;
        const int64_t i = h.t[k];
        const double bi = vb[k];
        x[i] = 0.0;
        r64[i] = bi;
        const float ri = (float)bi;
        const float zi = vi[k] * ri;
        r[i] = ri;
        xs[i] = 0.f;
        z[i] = zi;
        d[i] = zi;
        acc[0] += (double)ri * (double)zi;
        acc[1] += bi * vi64[k] * bi;
        acc[2] += bi * bi;
    }
    store_chunk<3>(acc, sd, tk.chunk, partials, sh);
}

__global__ void __launch_bounds__(kT)
kbm_fin_init(const int *__restrict__ runlist, const SlotDesc *__restrict__ desc,
             SlotState *__restrict__ states, const double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_FIN_PROLOGUE(ph_ == PH_INIT)
    double tot[3];
    slot_totals<3>(sd, partials, sh, tot);
    if (threadIdx.x != 0) return;
    st->nom = st->rho_ref = tot[0];
    st->rho64 = st->rho64_0 = tot[1];
    st->r2 = st->b2 = tot[2];
    st->iter = 0;
    st->outer = 0;
    st->converged = 0;
    if (!isfinite(tot[0]) || !isfinite(tot[1])) {
        st->nonfinite = 1;
        st->phase = PH_DONE;
    } else if (tot[1] < 0.0) {  // the preconditioner is not positive definite
        st->phase = PH_DONE;
    } else if (tot[1] == 0.0) {  // b = 0 on the free DOFs: x = 0 is the solution
        st->converged = 1;
        st->phase = PH_DONE;
    } else {
        st->phase = PH_INNER;
    }
}

// Element kernel over a group's work list.  RU = true: the FP64 residual
// apply (x -> ye64) of mixed slots in PH_RU; otherwise d -> ye for slots in
// PH_INNER.  force = 1 ignores the phase (timing only).
template <typename T, typename TP, int D, int Q, int NB, int U, int K, bool RU, int W>
__global__ void __launch_bounds__(NB * Q * Q, W)
kb_diffusion(const ElemTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
             const SlotState *__restrict__ states, const T *__restrict__ basis, int force) {
    const ElemTask tk = tasks[blockIdx.x];
    const SlotDesc &sd = desc[tk.slot];
    if (!force && states[tk.slot].phase != (RU ? PH_RU : PH_INNER)) return;
    __shared__ T smem[NB * diffusion_buf<D, Q>()];
    EOutT<T> out;
    out.l2n = global_ptr(sd.l2n);
    if (sd.bb < sd.n) {  // interior DOFs straight to the L-vector (z, or r64 for RU)
        out.yL = RU ? gv<T>(sd.r64) : gv<T>(sd.z);
        out.bb = sd.bb;
    }
    const int *gmap = global_ptr(sd.gmap);
    const uint8_t *mask = sd.mask ? global_ptr(sd.mask) : nullptr;
    if constexpr (RU) {
        out.ye = gv<T>(sd.ye64);
        diffusion_elements<T, TP, D, Q, NB, U, K>(tk.e0, sd.ne, basis, gv<TP>(sd.pa64), gmap, mask,
                                                  gv<T>(sd.x), out, 0, smem);
    } else {
        out.ye = gv<T>(sd.ye);
        diffusion_elements<T, TP, D, Q, NB, U, K>(tk.e0, sd.ne, basis, gv<TP>(sd.pa), gmap, mask,
                                                  gv<T>(sd.d), out, 0, smem);
    }
}

// z = A d for PH_INNER slots (DIAG_ONE rows copy d); partials of (d, z).
template <typename T>
__global__ void __launch_bounds__(kT)
kb_scatter(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
           const SlotState *__restrict__ states, double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_TASK_PROLOGUE(ph_ == PH_INNER)
    const T *ye = gv<const T>(sd.ye), *d = gv<const T>(sd.d);
    T *z = gv<T>(sd.z);
    const int *perm = sd.perm ? global_ptr(sd.perm) : nullptr;
    const int *off = global_ptr(sd.offsets), *ind = global_ptr(sd.indices);
    const uint8_t *mask = sd.mask ? global_ptr(sd.mask) : nullptr;
    // the share's rows, then their offsets, masks and d values, as batches
    const Share h = share_of(tk.chunk, sd.n);
    int64_t ii[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) ii[k] = h.lt[k] < sd.nperm ? perm[h.lt[k]] : h.lt[k];
    T vd[kPer];
    int ob[kPer], oe[kPer];
    bool m[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        vd[k] = d[ii[k]];
        ob[k] = off[ii[k]];
        oe[k] = off[ii[k] + 1];
        m[k] = mask && mask[ii[k]];
    }
    double acc[1] = {0.0};
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (!h.ok[k]) continue;
        T v;
        if (h.t[k] >= sd.bb) {  // interior DOF: written by the element kernel
            v = z[ii[k]];
        } else {
            v = m[k] ? vd[k] : row_sum(ind, ob[k], oe[k], ye);
            z[ii[k]] = v;
        }
        acc[0] += (double)vd[k] * (double)v;
    }
    store_chunk<1>(acc, sd, tk.chunk, partials, sh);
}

template <bool MIXED>
__global__ void __launch_bounds__(kT)
kb_fin_den(const int *__restrict__ runlist, const SlotDesc *__restrict__ desc,
           SlotState *__restrict__ states, const double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_FIN_PROLOGUE(ph_ == PH_INNER)
    double tot[1];
    slot_totals<1>(sd, partials, sh, tot);
    if (threadIdx.x != 0) return;
    st->den = tot[0];
    if (!isfinite(tot[0])) {  // MFEM_VERIFY(IsFinite(den))
        st->nonfinite = 1;
        st->final_iter = st->in_loop ? st->iter : 0;
        st->phase = PH_DONE;
    } else if (tot[0] == 0.0) {
        if (MIXED) {
            st->phase = PH_RU;  // inner residual is zero: check the true one
        } else {            // CGSolver: den == 0 ends the loop
            st->final_iter = st->in_loop ? st->iter : 0;
            st->phase = PH_DONE;
        }
    }
}

__global__ void __launch_bounds__(kT)
kb_update(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
          const SlotState *__restrict__ states, double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_TASK_PROLOGUE(ph_ == PH_INNER)
    const SlotState *st = states + tk.slot;
    const double alpha = st->nom / st->den;
    double *x = global_ptr(sd.x), *r = gv<double>(sd.r), *z = gv<double>(sd.z);
    const double *d = gv<double>(sd.d), *dinv = gv<double>(sd.dinv);
    const Share h = share_of(tk.chunk, sd.n);
    double vx[kPer], vd[kPer], vr[kPer], vz[kPer], vi[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t j = h.lt[k];
        vx[k] = x[j];
        vd[k] = d[j];
        vr[k] = r[j];
        vz[k] = z[j];
        vi[k] = dinv[j];
    }
    double acc[1] = {0.0};
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (!h.ok[k]) continue;
        const int64_t i = h.t[k];
        x[i] = vx[k] + alpha * vd[k];
        const double ri = vr[k] - alpha * vz[k];
        r[i] = ri;
        const double zi = vi[k] * ri;
        z[i] = zi;
        acc[0] += ri * zi;
    }
    store_chunk<1>(acc, sd, tk.chunk, partials, sh);
}

__global__ void __launch_bounds__(kT)
kb_fin_update(const int *__restrict__ runlist, const SlotDesc *__restrict__ desc,
              SlotState *__restrict__ states, const double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_FIN_PROLOGUE(ph_ == PH_INNER)
    double tot[1];
    slot_totals<1>(sd, partials, sh, tot);
    if (threadIdx.x != 0) return;
    const double betanom = tot[0];
    const int i = st->iter;
    st->in_loop = 1;
    st->betanom = betanom;
    if (!isfinite(betanom)) {  // MFEM_VERIFY(IsFinite(betanom))
        st->nonfinite = 1;
        st->final_iter = i;
        st->phase = PH_DONE;
    } else if (betanom < 0.0) {
        st->final_iter = i;
        st->phase = PH_DONE;
    } else if (betanom <= st->r0) {
        st->converged = 1;
        st->final_iter = i;
        st->phase = PH_DONE;
    } else if (i + 1 > st->max_iter) {
        st->final_iter = st->max_iter;
        st->phase = PH_DONE;
    } else {
        st->beta = betanom / st->nom;
        st->nom = betanom;
        st->iter = i + 1;
    }
}

__global__ void __launch_bounds__(kT)
kbm_update(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
           const SlotState *__restrict__ states, double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_TASK_PROLOGUE(ph_ == PH_INNER)
    const SlotState *st = states + tk.slot;
    const float alpha = (float)(st->nom / st->den);
    float *xs = gv<float>(sd.xi), *r = gv<float>(sd.r), *z = gv<float>(sd.z);
    const float *d = gv<float>(sd.d), *dinv = gv<float>(sd.dinv);
    const Share h = share_of(tk.chunk, sd.n);
    float vx[kPer], vd[kPer], vr[kPer], vz[kPer], vi[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t j = h.lt[k];
        vx[k] = xs[j];
        vd[k] = d[j];
        vr[k] = r[j];
        vz[k] = z[j];
        vi[k] = dinv[j];
    }
    double acc[1] = {0.0};
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (!h.ok[k]) continue;
        const int64_t i = h.t[k];
        xs[i] = vx[k] + alpha * vd[k];
        const float ri = vr[k] - alpha * vz[k];
        r[i] = ri;
        const float zi = vi[k] * ri;
        z[i] = zi;
        acc[0] += (double)ri * (double)zi;
    }
    store_chunk<1>(acc, sd, tk.chunk, partials, sh);
}

__global__ void __launch_bounds__(kT)
kbm_fin_update(const int *__restrict__ runlist, const SlotDesc *__restrict__ desc,
               SlotState *__restrict__ states, const double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_FIN_PROLOGUE(ph_ == PH_INNER)
    double tot[1];
    slot_totals<1>(sd, partials, sh, tot);
    if (threadIdx.x != 0) return;
    const double rho = tot[0];
    st->iter += 1;
    if (!isfinite(rho)) {
        st->nonfinite = 1;
        st->phase = PH_DONE;
        return;
    }
    if (rho < 0.0) {  // the preconditioner is not positive definite
        st->phase = PH_DONE;
        return;
    }
    const double tol2 = sd.rel_tol * sd.rel_tol;
    if (rho <= sd.delta * sd.delta * st->rho_ref || rho <= tol2 * st->rho64_0 ||
        st->iter >= st->max_iter) {
        st->phase = PH_RU;
    } else {
        st->beta = rho / st->nom;
        st->nom = rho;
    }
}

template <typename T>
__global__ void __launch_bounds__(kT)
kb_dir(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
       const SlotState *__restrict__ states) {
    H1PA_TASK_PROLOGUE(ph_ == PH_INNER)
    const T beta = (T)states[tk.slot].beta;
    const T *z = gv<T>(sd.z);
    T *d = gv<T>(sd.d);
    const Share h = share_of(tk.chunk, sd.n);
    T vz[kPer], vd[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        vz[k] = z[h.lt[k]];
        vd[k] = d[h.lt[k]];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (h.ok[k]) d[h.t[k]] = vz[k] + beta * vd[k];
}

// Refinement step, part 1 (PH_RU): x += xs, xs = 0.
__global__ void __launch_bounds__(kT)
kbm_ru_x(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
         const SlotState *__restrict__ states) {
    H1PA_TASK_PROLOGUE(ph_ == PH_RU)
    float *xs = gv<float>(sd.xi);
    double *x = global_ptr(sd.x);
    const Share h = share_of(tk.chunk, sd.n);
    double vx[kPer];
    float vs[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        vx[k] = x[h.lt[k]];
        vs[k] = xs[h.lt[k]];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (!h.ok[k]) continue;
        x[h.t[k]] = vx[k] + (double)vs[k];
        xs[h.t[k]] = 0.f;
    }
}

// Refinement step, part 2 (PH_RU, after the FP64 element kernel on x):
// r64 = b - A x, r = (float) r64, z = dinv r; partials of (r, z), (B r64, r64)
// and ||r64||^2.
__global__ void __launch_bounds__(kT)
kbm_ru_res(const VecTask *__restrict__ tasks, const SlotDesc *__restrict__ desc,
           const SlotState *__restrict__ states, double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_TASK_PROLOGUE(ph_ == PH_RU)
    float *r = vp<float>(sd.r), *z = vp<float>(sd.z);
    const float *dinv = vp<float>(sd.dinv);
    double acc[3] = {0.0, 0.0, 0.0};
    H1PA_CHUNK_LOOP(t) {
        const int64_t i = t < sd.nperm ? sd.perm[t] : t;
        double y;
        if (t >= sd.bb) {  // interior DOF: (A x)_i written into r64 by the element kernel
            y = sd.r64[i];
        } else if (sd.mask && sd.mask[i]) {
            y = sd.x[i];
        } else {
            y = row_sum(sd.indices, sd.offsets[i], sd.offsets[i + 1], sd.ye64);
        }
        const double ri = sd.b[i] - y;
        sd.r64[i] = ri;
        const float rf = (float)ri;
        const float zf = dinv[i] * rf;
        r[i] = rf;
        z[i] = zf;
        acc[0] += (double)rf * (double)zf;
        acc[1] += ri * sd.dinv64[i] * ri;
        acc[2] += ri * ri;
    }
    store_chunk<3>(acc, sd, tk.chunk, partials, sh);
}

__global__ void __launch_bounds__(kT)
kbm_fin_ru(const int *__restrict__ runlist, const SlotDesc *__restrict__ desc,
           SlotState *__restrict__ states, const double *__restrict__ partials) {
    __shared__ double sh[kT];
    H1PA_FIN_PROLOGUE(ph_ == PH_RU)
    double tot[3];
    slot_totals<3>(sd, partials, sh, tot);
    if (threadIdx.x != 0) return;
    st->outer += 1;
    st->rho64 = tot[1];
    st->r2 = tot[2];
    if (!isfinite(tot[0]) || !isfinite(tot[1])) {
        st->nonfinite = 1;
        st->phase = PH_DONE;
    } else if (tot[1] < 0.0) {
        st->phase = PH_DONE;
    } else if (tot[1] <= sd.rel_tol * sd.rel_tol * st->rho64_0) {
        st->converged = 1;
        st->phase = PH_DONE;
    } else if (st->outer >= sd.max_outer || st->iter >= st->max_iter) {
        st->phase = PH_DONE;
    } else {
        st->beta = st->nom > 0.0 ? tot[0] / st->nom : 0.0;
        st->nom = tot[0];
        st->rho_ref = tot[0];
        st->phase = PH_INNER;
    }
}

// ---- element-kernel dispatch over (D, Q, variant) --------------------------
template <typename T, typename TP, int D, int Q, int NB, int U, int K, bool RU, int W = 1>
void launch_kb(const ElemTask *tasks, int ntasks, const SlotDesc *desc, const SlotState *st,
               const T *basis, int force, hipStream_t s) {
    hipLaunchKernelGGL((kb_diffusion<T, TP, D, Q, NB, U, K, RU, W>), dim3((unsigned)ntasks),
                       dim3(NB * Q * Q), 0, s, tasks, desc, st, basis, force);
}

template <typename T, int D, int Q>
constexpr int nb_max() { return nb_for(Q, 3 * D * Q * Q, sizeof(T)); }
template <typename T, int D, int Q>
constexpr int nb_half() { return nb_max<T, D, Q>() / 2 > 0 ? nb_max<T, D, Q>() / 2 : 1; }

// Elements per block of variant v for value type T (same table as the
// single-problem kernels).
template <typename T, int D, int Q>
int nb_of(int v) { return (v == 2 || v == 3 || v == 8) ? nb_half<T, D, Q>() : nb_max<T, D, Q>(); }

template <typename T, typename TP, int D, int Q, bool RU>
void launch_kb_variant(int v, const ElemTask *tasks, int ntasks, const SlotDesc *desc,
                       const SlotState *st, const T *basis, int force, hipStream_t s) {
    constexpr int NM = nb_max<T, D, Q>(), NH = nb_half<T, D, Q>();
    constexpr int OW = lds_waves<T, D, Q, NM>();
    if constexpr (sizeof(T) == 8) {
        switch (v) {
            case 1: launch_kb<T, TP, D, Q, NM, 2, 0, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 2: launch_kb<T, TP, D, Q, NH, 1, 0, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 3: launch_kb<T, TP, D, Q, NH, 2, 0, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 4: launch_kb<T, TP, D, Q, NM, 1, 2, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 5: launch_kb<T, TP, D, Q, NM, 1, kAllPlanes, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 6: launch_kb<T, TP, D, Q, NM, 1, 0, RU, OW>(tasks, ntasks, desc, st, basis, force, s); return;
            case 7: launch_kb<T, TP, D, Q, NM, 1, 2, RU, OW>(tasks, ntasks, desc, st, basis, force, s); return;
            case 8: launch_kb<T, TP, D, Q, NH, 1, 2, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            default: launch_kb<T, TP, D, Q, NM, 1, 0, RU>(tasks, ntasks, desc, st, basis, force, s); return;
        }
    } else {
        switch (v) {
            case 1: launch_kb<T, TP, D, Q, NM, 2, 0, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 4: launch_kb<T, TP, D, Q, NM, 1, 2, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 5: launch_kb<T, TP, D, Q, NM, 1, kAllPlanes, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            case 6: launch_kb<T, TP, D, Q, NM, 1, 0, RU, OW>(tasks, ntasks, desc, st, basis, force, s); return;
            case 7: launch_kb<T, TP, D, Q, NM, 1, 2, RU, OW>(tasks, ntasks, desc, st, basis, force, s); return;
            case 8: launch_kb<T, TP, D, Q, NH, 1, 2, RU>(tasks, ntasks, desc, st, basis, force, s); return;
            default: launch_kb<T, TP, D, Q, NM, 1, 0, RU>(tasks, ntasks, desc, st, basis, force, s); return;
        }
    }
}

}  // namespace
}  // namespace kern

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

// Variants valid for the inner type (FP32 applies have 0, 1, 4, 5, 6, 7, 8).
bool variant_ok(int prec, int v) {
    return prec == H1PA_PREC_FP64 ? (v >= 0 && v < 9) : (v == 0 || v == 1 || (v >= 4 && v < 9));
}

#define H1PA_BATCH_DQ(X) X(2, 3) X(3, 4) X(4, 5) X(5, 6) X(6, 7) X(7, 8) X(8, 9) X(9, 10)

// Elements per block of variant v: FP64 kernels (fp64 = true) or FP32 ones.
int nb_of_dq(int D, int Q, int v, bool fp64) {
#define X(DD, QQ) \
    if (D == DD && Q == QQ) \
        return fp64 ? kern::nb_of<double, DD, QQ>(v) : kern::nb_of<float, DD, QQ>(v);
    H1PA_BATCH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "batch: unsupported (p, q1d)"};
}

void launch_group(int prec, bool ru, int D, int Q, int v, const ElemTask *tasks, int ntasks,
                  const SlotDesc *desc, const SlotState *st, const void *basis, int force,
                  hipStream_t s) {
    if (ntasks <= 0) return;
#define X(DD, QQ)                                                                              \
    if (D == DD && Q == QQ) {                                                                  \
        if (ru || prec == H1PA_PREC_FP64) {                                                    \
            if (ru)                                                                            \
                kern::launch_kb_variant<double, double, DD, QQ, true>(                         \
                    v, tasks, ntasks, desc, st, (const double *)basis, force, s);              \
            else                                                                               \
                kern::launch_kb_variant<double, double, DD, QQ, false>(                        \
                    v, tasks, ntasks, desc, st, (const double *)basis, force, s);              \
        } else if (prec == H1PA_PREC_FP32) {                                                   \
            kern::launch_kb_variant<float, float, DD, QQ, false>(v, tasks, ntasks, desc, st,   \
                                                                 (const float *)basis, force, s); \
        } else {                                                                               \
            kern::launch_kb_variant<float, bf16_t, DD, QQ, false>(                             \
                v, tasks, ntasks, desc, st, (const float *)basis, force, s);                   \
        }                                                                                      \
        H1PA_HIP(hipGetLastError());                                                           \
        return;                                                                                \
    }
    H1PA_BATCH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "batch: unsupported (p, q1d)"};
}

}  // namespace

struct BatchSlot {
    h1pa_form *form = nullptr;
    int64_t tag = 0;
    int group = -1;
    int64_t n = 0, ne = 0;
    DevBuf<uint8_t> mask;
    DevBuf<double> v64[5];  // FP64 mode: r z d dinv ye;  mixed: r64 dinv64 ye64
    DevBuf<float> v32[5];   // mixed: r z d dinv xs
    SlotDesc desc{};
    SlotState init{};       // state the slot starts from
    hipEvent_t ready = nullptr;  // pending slot: its setup work on the caller's stream
    size_t bytes = 0;
    bool live = false;      // added and not yet polled
    ~BatchSlot() {
        if (ready) (void)hipEventDestroy(ready);
    }
};

struct BatchGroup {
    int p = 0, q1d = 0, variant = 0;
    int nb = 1;    // elements per block of the inner kernels (the batch's type)
    int nb64 = 1;  // of the FP64 refinement kernels of a mixed batch
    // the group's own copies of the basis (forms may be destroyed before the
    // batch is): FP64, and FP32 for a mixed batch
    DevBuf<double> basis64_own;
    DevBuf<float> basis32_own;
    const void *basis = nullptr;      // inner basis (type of the batch)
    const double *basis64 = nullptr;  // FP64 basis (refinement residual)
    DevBuf<ElemTask> tasks;
    int ntasks = 0;
    DevBuf<ElemTask> tasks64;  // mixed: work list of the FP64 refinement kernels
    int ntasks64 = 0;
    bool tuned = false;
};

}  // namespace h1pa

struct h1pa_batch {
    h1pa_batch_opts o{};
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<h1pa::BatchSlot>> slots;
    std::vector<h1pa::BatchGroup> groups;
    h1pa::DevBuf<h1pa::SlotDesc> d_desc;
    h1pa::DevBuf<h1pa::SlotState> d_state;
    std::vector<h1pa::SlotState> h_state;
    h1pa::DevBuf<double> partials;
    h1pa::DevBuf<h1pa::VecTask> vtasks;
    int nvtasks = 0;
    h1pa::DevBuf<int> runlist;  // slots with work items, one finalize block each
    int nrun = 0;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t replays = 0, captures = 0;
    // Every device address and launch size the captured replay holds in its
    // kernel nodes, taken at capture; a launch of the cached graph checks
    // them against the batch's current buffers.
    std::vector<uintptr_t> baked;
    // solves added since the last launch; merged into `slots` by the next one
    std::vector<std::unique_ptr<h1pa::BatchSlot>> pending;
    bool in_flight = false;
    h1pa::SlotState *h_pinned = nullptr;  // replay's state copy (pinned host memory)
    size_t h_pinned_n = 0;

    ~h1pa_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        // solves never retired by poll release their hold on the forms
        for (auto &sl : pending)
            if (sl && sl->live) --sl->form->f.batch_refs;
        for (auto &sl : slots)
            if (sl && sl->live) --sl->form->f.batch_refs;
        pending.clear();
        slots.clear();
        if (h_pinned) (void)hipHostFree(h_pinned);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace h1pa {
using Batch = h1pa_batch;

namespace {

bool mixed_mode(const Batch &B) { return B.o.precision != H1PA_PREC_FP64; }

int group_for(Batch &B, Form &f) {
    const int p = f.space->p;
    for (size_t g = 0; g < B.groups.size(); ++g)
        if (B.groups[g].p == p && B.groups[g].q1d == f.q1d) return (int)g;
    BatchGroup G;
    G.p = p;
    G.q1d = f.q1d;
    G.variant = B.o.precision == H1PA_PREC_FP64
                    ? kern::get_diffusion_variant(p, f.q1d)
                    : kern::get_diffusion_variant_lp(B.o.precision == H1PA_PREC_BF16_PA, p, f.q1d);
    if (!variant_ok(B.o.precision, G.variant)) G.variant = 0;
    G.nb = nb_of_dq(p + 1, f.q1d, G.variant, B.o.precision == H1PA_PREC_FP64);
    G.nb64 = nb_of_dq(p + 1, f.q1d, G.variant, true);
    B.groups.push_back(std::move(G));
    return (int)B.groups.size() - 1;
}

bool running_phase(int ph) { return ph == PH_INIT || ph == PH_INNER || ph == PH_RU; }

void drop_graph(Batch &B) {
    if (B.exec) {
        H1PA_HIP(hipGraphExecDestroy(B.exec));
        B.exec = nullptr;
    }
    if (B.graph) {
        H1PA_HIP(hipGraphDestroy(B.graph));
        B.graph = nullptr;
    }
}

// Work list of group g: (slot, first element) per block, in slot and element
// order (consecutive blocks take neighbouring elements of one slot).
void element_tasks(const Batch &B, int g, int nb, std::vector<ElemTask> &out) {
    out.clear();
    for (size_t k = 0; k < B.slots.size(); ++k) {
        const BatchSlot &sl = *B.slots[k];
        if (sl.group != g || !sl.live || !running_phase(B.h_state[k].phase)) continue;
        for (int64_t e = 0; e < sl.ne; e += nb) out.push_back({(int)k, (int)e});
    }
}

// Rebuild descriptors, work lists and partial slices for the running slots;
// the captured graph is dropped.
void rebuild(Batch &B) {
    drop_graph(B);
    const int ns = (int)B.slots.size();
    std::vector<SlotDesc> desc(ns);
    std::vector<VecTask> vt;
    std::vector<int> run;
    int part = 0;
    for (int k = 0; k < ns; ++k) {
        BatchSlot &sl = *B.slots[k];
        desc[k] = sl.desc;
        desc[k].part0 = 0;
        desc[k].nchunks = 0;
        if (!sl.live || !running_phase(B.h_state[k].phase)) continue;
        const int nch = (int)((sl.n + kern::kChunk - 1) / kern::kChunk);
        desc[k].part0 = part;
        desc[k].nchunks = nch;
        part += nch;
        for (int c = 0; c < nch; ++c) vt.push_back({k, c});
        run.push_back(k);
    }
    for (int k = 0; k < ns; ++k) B.slots[k]->desc = desc[k];
    if (B.d_desc.n < (size_t)ns) B.d_desc.alloc(std::max<size_t>(ns, 2 * B.d_desc.n));
    B.d_desc.upload(desc.data(), ns, B.stream);
    if (B.partials.n < 3 * (size_t)part + 3) B.partials.alloc(3 * (size_t)part + 3);
    B.nvtasks = (int)vt.size();
    if (B.vtasks.n < vt.size()) B.vtasks.alloc(vt.size());
    B.vtasks.upload(vt.data(), vt.size(), B.stream);
    B.nrun = (int)run.size();
    if (B.runlist.n < run.size()) B.runlist.alloc(run.size());
    B.runlist.upload(run.data(), run.size(), B.stream);
    std::vector<ElemTask> et;
    for (size_t g = 0; g < B.groups.size(); ++g) {
        BatchGroup &G = B.groups[g];
        element_tasks(B, (int)g, G.nb, et);
        G.ntasks = (int)et.size();
        if (G.tasks.n < et.size()) G.tasks.alloc(et.size());
        G.tasks.upload(et.data(), et.size(), B.stream);
        if (mixed_mode(B)) {
            element_tasks(B, (int)g, G.nb64, et);
            G.ntasks64 = (int)et.size();
            if (G.tasks64.n < et.size()) G.tasks64.alloc(et.size());
            G.tasks64.upload(et.data(), et.size(), B.stream);
        }
    }
    H1PA_HIP(hipStreamSynchronize(B.stream));
    B.dirty = false;
}

// Pick each group's element-kernel variant by timing every candidate on the
// group's current slots (phase ignored; the E-vectors written are scratch
// here).  Variants with other elements-per-block get their own work list.
void autotune(Batch &B) {
    if (!B.o.autotune) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    H1PA_HIP(hipEventCreate(&e0));
    H1PA_HIP(hipEventCreate(&e1));
    std::vector<ElemTask> et;
    DevBuf<ElemTask> tmp;
    try {
        for (size_t g = 0; g < B.groups.size(); ++g) {
            BatchGroup &G = B.groups[g];
            if (G.tuned || G.ntasks < 256) continue;
            G.tuned = true;
            float best = 1e30f;
            int bestv = G.variant;
            for (int v = 0; v < 9; ++v) {
                if (!variant_ok(B.o.precision, v)) continue;
                const int nb = nb_of_dq(G.p + 1, G.q1d, v, B.o.precision == H1PA_PREC_FP64);
                element_tasks(B, (int)g, nb, et);
                tmp.upload(et.data(), et.size(), B.stream);
                const int nt = (int)et.size();
                for (int rep = 0; rep < 2; ++rep)
                    launch_group(B.o.precision, false, G.p + 1, G.q1d, v, tmp.p, nt, B.d_desc.p,
                                 B.d_state.p, G.basis, 1, B.stream);
                H1PA_HIP(hipEventRecord(e0, B.stream));
                for (int rep = 0; rep < 5; ++rep)
                    launch_group(B.o.precision, false, G.p + 1, G.q1d, v, tmp.p, nt, B.d_desc.p,
                                 B.d_state.p, G.basis, 1, B.stream);
                H1PA_HIP(hipEventRecord(e1, B.stream));
                H1PA_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                H1PA_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bestv = v;
                }
            }
            if (bestv != G.variant) {
                G.variant = bestv;
                G.nb = nb_of_dq(G.p + 1, G.q1d, bestv, B.o.precision == H1PA_PREC_FP64);
                G.nb64 = nb_of_dq(G.p + 1, G.q1d, bestv, true);
                B.dirty = true;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// The device addresses and sizes a replay's kernel launches take as
// arguments (enqueue_replay), in a fixed order.
std::vector<uintptr_t> replay_arguments(const Batch &B) {
    std::vector<uintptr_t> a{(uintptr_t)B.vtasks.p, (uintptr_t)B.runlist.p,
                             (uintptr_t)B.d_desc.p, (uintptr_t)B.d_state.p,
                             (uintptr_t)B.partials.p, (uintptr_t)B.nvtasks, (uintptr_t)B.nrun};
    for (const BatchGroup &G : B.groups) {
        a.push_back((uintptr_t)G.tasks.p);
        a.push_back((uintptr_t)G.ntasks);
        a.push_back((uintptr_t)G.tasks64.p);
        a.push_back((uintptr_t)G.ntasks64);
        a.push_back((uintptr_t)G.basis);
        a.push_back((uintptr_t)G.basis64);
        a.push_back((uintptr_t)G.variant);
    }
    return a;
}

void enqueue_replay(Batch &B, hipStream_t s) {
    const unsigned nv = (unsigned)B.nvtasks;
    if (nv == 0) return;
    const int prec = B.o.precision;
    const dim3 grid(nv), blk(kern::kT), fin((unsigned)B.nrun);
    const VecTask *vt = B.vtasks.p;
    const int *rl = B.runlist.p;
    const SlotDesc *ds = B.d_desc.p;
    SlotState *st = B.d_state.p;
    double *pt = B.partials.p;
    auto elements = [&](bool ru) {
        for (BatchGroup &G : B.groups) {
            if (ru)
                launch_group(prec, true, G.p + 1, G.q1d, G.variant, G.tasks64.p, G.ntasks64, ds,
                             st, G.basis64, 0, s);
            else
                launch_group(prec, false, G.p + 1, G.q1d, G.variant, G.tasks.p, G.ntasks, ds, st,
                             G.basis, 0, s);
        }
    };
    if (!mixed_mode(B)) {
        hipLaunchKernelGGL(kern::kb_init, grid, blk, 0, s, vt, ds, st, pt);
        hipLaunchKernelGGL(kern::kb_fin_init, fin, blk, 0, s, rl, ds, st, pt);
        for (int it = 0; it < B.o.iters_per_replay; ++it) {
            elements(false);
            hipLaunchKernelGGL(kern::kb_scatter<double>, grid, blk, 0, s, vt, ds, st, pt);
            hipLaunchKernelGGL(kern::kb_fin_den<false>, fin, blk, 0, s, rl, ds, st, pt);
            hipLaunchKernelGGL(kern::kb_update, grid, blk, 0, s, vt, ds, st, pt);
            hipLaunchKernelGGL(kern::kb_fin_update, fin, blk, 0, s, rl, ds, st, pt);
            hipLaunchKernelGGL(kern::kb_dir<double>, grid, blk, 0, s, vt, ds, st);
        }
    } else {
        hipLaunchKernelGGL(kern::kbm_init, grid, blk, 0, s, vt, ds, st, pt);
        hipLaunchKernelGGL(kern::kbm_fin_init, fin, blk, 0, s, rl, ds, st, pt);
        for (int it = 0; it < B.o.iters_per_replay; ++it) {
            elements(false);
            hipLaunchKernelGGL(kern::kb_scatter<float>, grid, blk, 0, s, vt, ds, st, pt);
            hipLaunchKernelGGL(kern::kb_fin_den<true>, fin, blk, 0, s, rl, ds, st, pt);
            hipLaunchKernelGGL(kern::kbm_update, grid, blk, 0, s, vt, ds, st, pt);
            hipLaunchKernelGGL(kern::kbm_fin_update, fin, blk, 0, s, rl, ds, st, pt);
            if (it == B.o.iters_per_replay - 1) {
                // refinement steps once per replay: a slot whose inner residual
                // fell by delta waits in PH_RU (skipped by the inner kernels)
                // until here, instead of every iteration launching the
                // refinement kernels over all slots
                hipLaunchKernelGGL(kern::kbm_ru_x, grid, blk, 0, s, vt, ds, st);
                elements(true);
                hipLaunchKernelGGL(kern::kbm_ru_res, grid, blk, 0, s, vt, ds, st, pt);
                hipLaunchKernelGGL(kern::kbm_fin_ru, fin, blk, 0, s, rl, ds, st, pt);
            }
            hipLaunchKernelGGL(kern::kb_dir<float>, grid, blk, 0, s, vt, ds, st);
        }
    }
    H1PA_HIP(hipGetLastError());
}

}  // namespace

void batch_destroy(Batch *B) { delete B; }

hipStream_t batch_stream(const Batch &B) { return B.stream; }

Batch *batch_create(const h1pa_batch_opts &o) {
    H1PA_CHECK(o.precision == H1PA_PREC_FP64 || o.precision == H1PA_PREC_FP32 ||
                   o.precision == H1PA_PREC_BF16_PA,
               H1PA_EINVAL, "batch: unknown precision");
    H1PA_CHECK(o.iters_per_replay >= 1 && o.iters_per_replay <= 256, H1PA_EINVAL,
               "batch: iters_per_replay must be in [1, 256]");
    if (o.precision != H1PA_PREC_FP64)
        H1PA_CHECK(o.max_outer >= 1 && o.max_inner >= 1 && o.delta > 0.0 && o.delta < 1.0,
                   H1PA_EINVAL,
                   "batch: mixed precision needs max_outer, max_inner >= 1 and 0 < delta < 1");
    std::unique_ptr<Batch> B(new Batch);
    B->o = o;
    H1PA_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    return B.release();
}

// Set a solve up on the caller's stream `cs` (shadow PA data, constraint
// mask, Jacobi diagonal) and queue it; the next launch merges it.  Nothing
// here touches data that an in-flight replay uses, and nothing frees device
// memory (a free would wait for the replay).
int batch_add(Batch &B, h1pa_form *form, const uint8_t *mask_src, const double *b, double *x,
              const h1pa_pcg_opts &po, int64_t tag, hipStream_t cs) {
    Form &f = form->f;
    Space &sp = *f.space;
    H1PA_CHECK(f.desc.integrators == H1PA_INTEG_DIFFUSION, H1PA_EUNSUP,
               "batch: diffusion-only forms are supported");
    H1PA_CHECK(f.q1d == sp.p + 2, H1PA_EUNSUP, "batch: q1d must be p + 2");
    H1PA_CHECK(po.max_iter >= 1, H1PA_EINVAL, "batch: max_iter must be >= 1");
    const bool mixed = mixed_mode(B);
    if (mixed) ensure_low_precision(f, B.o.precision, cs);
    std::unique_ptr<BatchSlot> sl(new BatchSlot);
    sl->form = form;
    sl->tag = tag;
    sl->n = sp.ndofs;
    sl->ne = sp.ne;
    sl->group = group_for(B, f);
    BatchGroup &G = B.groups[sl->group];
    if (!G.basis64_own.p) {  // a new group: not referenced by any captured replay yet
        G.basis64_own.alloc(f.basis.n);
        H1PA_HIP(hipMemcpyAsync(G.basis64_own.p, f.basis.p, f.basis.n * sizeof(double),
                                hipMemcpyDeviceToDevice, cs));
        G.basis64 = G.basis64_own.p;
        G.basis = G.basis64_own.p;
        if (mixed) {
            G.basis32_own.alloc(f.basis32.n);
            H1PA_HIP(hipMemcpyAsync(G.basis32_own.p, f.basis32.p, f.basis32.n * sizeof(float),
                                    hipMemcpyDeviceToDevice, cs));
            G.basis = G.basis32_own.p;
        }
    }
    const int64_t n = sp.ndofs, nE = sp.ne * sp.dof;
    if (mask_src) {
        sl->mask.alloc(n);
        H1PA_HIP(hipMemcpyAsync(sl->mask.p, mask_src, n, hipMemcpyDeviceToDevice, cs));
    }
    SlotDesc &d = sl->desc;
    d.gmap = sp.restr.gather_map.p;
    d.offsets = sp.restr.offsets.p;
    d.indices = sp.restr.indices_nat.p;
    d.perm = sp.restr.perm.p;
    d.nperm = sp.restr.nperm;
    d.l2n = sp.restr.lex2nat.p;
    d.mask = sl->mask.p;
    d.b = b;
    d.x = x;
    d.ne = sp.ne;
    d.n = n;
    // interior DOFs bypass the E-vector unless a constrained DOF may be interior
    d.bb = (!mask_src || mask_src == sp.ess_mask.p) ? sp.restr.bubble_begin : n;
    d.pa64 = f.pa_diff.p;
    d.rel_tol = po.rel_tol;
    d.abs_tol = po.abs_tol;
    size_t bytes = sl->mask.n;
    if (!mixed) {
        for (int k = 0; k < 4; ++k) sl->v64[k].alloc(n);
        sl->v64[4].alloc(nE);
        d.r = sl->v64[0].p;
        d.z = sl->v64[1].p;
        d.d = sl->v64[2].p;
        d.dinv = sl->v64[3].p;
        d.ye = sl->v64[4].p;
        d.xi = x;
        d.pa = f.pa_diff.p;
        // the diagonal passes through r, which the first replay overwrites
        operator_diagonal(f, sl->v64[0].p, cs);
        kern::jacobi_dinv(n, sl->v64[0].p, sl->mask.p, sl->v64[3].p, cs);
        bytes += (4 * n + nE) * 8;
    } else {
        for (int k = 0; k < 5; ++k) sl->v32[k].alloc(n);
        sl->v64[0].alloc(n);
        sl->v64[1].alloc(n);
        sl->v64[2].alloc(nE);
        d.r = sl->v32[0].p;
        d.z = sl->v32[1].p;
        d.d = sl->v32[2].p;
        d.dinv = sl->v32[3].p;
        d.xi = sl->v32[4].p;
        d.r64 = sl->v64[0].p;
        d.dinv64 = sl->v64[1].p;
        // the FP32 E-vector lives in the FP64 one: they are never live together
        d.ye = sl->v64[2].p;
        d.ye64 = sl->v64[2].p;
        d.pa = f.pa_diff_lp.p;
        d.delta = B.o.delta;
        d.max_outer = B.o.max_outer;
        // the diagonal passes through r64, which the first replay overwrites
        operator_diagonal(f, sl->v64[0].p, cs);
        kern::jacobi_dinv(n, sl->v64[0].p, sl->mask.p, sl->v64[1].p, cs);
        kern::jacobi_dinv_f32(n, sl->v64[0].p, sl->mask.p, sl->v32[3].p, cs);
        bytes += 5 * n * 4 + 2 * n * 8 + nE * 8;
    }
    sl->bytes = bytes;
    sl->live = true;
    ++f.batch_refs;  // the descriptor holds f's PA pointers until poll retires it
    sl->init.phase = PH_INIT;
    sl->init.max_iter = mixed ? B.o.max_inner : po.max_iter;
    H1PA_HIP(hipEventCreateWithFlags(&sl->ready, hipEventDisableTiming));
    H1PA_HIP(hipEventRecord(sl->ready, cs));
    B.pending.push_back(std::move(sl));
    return (int)B.pending.size() - 1;
}

namespace {

// Merge the pending solves (nothing in flight): the batch stream waits for
// their setup, each takes a retired position or a new one.
void merge_pending(Batch &B) {
    if (B.pending.empty()) return;
    hipStream_t s = B.stream;
    for (auto &up : B.pending) {
        H1PA_HIP(hipStreamWaitEvent(s, up->ready, 0));
        int k = -1;
        for (size_t i = 0; i < B.slots.size(); ++i)
            if (!B.slots[i]->live && B.h_state[i].phase == PH_EMPTY) {
                k = (int)i;
                break;
            }
        if (k < 0) {
            k = (int)B.slots.size();
            if (B.d_state.n < (size_t)k + 1) {
                DevBuf<SlotState> nst(std::max<size_t>(64, 2 * ((size_t)k + 1)));
                if (k > 0)
                    H1PA_HIP(hipMemcpyAsync(nst.p, B.d_state.p, sizeof(SlotState) * k,
                                            hipMemcpyDeviceToDevice, s));
                H1PA_HIP(hipStreamSynchronize(s));
                B.d_state = std::move(nst);
            }
            B.slots.push_back(nullptr);
            B.h_state.push_back(SlotState{});
        }
        B.h_state[k] = up->init;
        B.slots[k] = std::move(up);
    }
    B.pending.clear();
    H1PA_HIP(hipStreamSynchronize(s));
    // Nothing is in flight, so the device states equal h_state: upload them
    // all with one synchronous copy.  (An asynchronous copy from h_state,
    // issued inside the loop above, could still be reading the vector when a
    // push_back moved it.)
    H1PA_HIP(hipMemcpy(B.d_state.p, B.h_state.data(), sizeof(SlotState) * B.slots.size(),
                       hipMemcpyHostToDevice));
    B.dirty = true;
}

int running_slots(const Batch &B) {
    int live = (int)B.pending.size();
    for (size_t k = 0; k < B.slots.size(); ++k)
        if (B.slots[k]->live && running_phase(B.h_state[k].phase)) ++live;
    return live;
}

}  // namespace

// Launch one replay (asynchronously) for the running solves; returns false
// when nothing runs.
bool batch_launch(Batch &B) {
    H1PA_CHECK(!B.in_flight, H1PA_EINVAL, "batch: a replay is in flight (call wait first)");
    hipStream_t s = B.stream;
    merge_pending(B);
    if (running_slots(B) == 0) return false;
    // compact once a third of the work list belongs to finished slots
    if (!B.dirty) {
        int64_t dead = 0, all = 0;
        for (size_t k = 0; k < B.slots.size(); ++k) {
            const int nch = B.slots[k]->desc.nchunks;
            all += nch;
            if (!running_phase(B.h_state[k].phase)) dead += nch;
        }
        if (all > 0 && 3 * dead > all) B.dirty = true;
    }
    if (B.dirty) {
        rebuild(B);
        autotune(B);
        if (B.dirty) rebuild(B);  // a variant with another block size was chosen
    }
    if (B.o.use_graph) {
        if (!B.exec) {
            H1PA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_replay(B, s);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            H1PA_HIP(hipStreamEndCapture(s, &B.graph));
            H1PA_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
            B.baked = replay_arguments(B);
            ++B.captures;
        } else {
            // every buffer the captured kernels address must still be the one
            // they were captured with (buffers move only in rebuild, which
            // drops the graph): a stale graph is refused, never launched
            H1PA_CHECK(replay_arguments(B) == B.baked, H1PA_EINVAL,
                       "batch: the captured replay refers to buffers that have been replaced");
        }
        H1PA_HIP(hipGraphLaunch(B.exec, s));
    } else {
        enqueue_replay(B, s);
    }
    if (B.h_pinned_n < B.slots.size()) {
        if (B.h_pinned) H1PA_HIP(hipHostFree(B.h_pinned));
        B.h_pinned = nullptr;
        B.h_pinned_n = 0;
        const size_t cap = std::max<size_t>(64, 2 * B.slots.size());
        H1PA_HIP(hipHostMalloc((void **)&B.h_pinned, cap * sizeof(SlotState), hipHostMallocDefault));
        B.h_pinned_n = cap;
    }
    H1PA_HIP(hipMemcpyAsync(B.h_pinned, B.d_state.p, sizeof(SlotState) * B.slots.size(),
                            hipMemcpyDeviceToHost, s));
    B.in_flight = true;
    return true;
}

// Wait for the replay in flight (if any); returns the number of running solves.
int batch_wait(Batch &B) {
    if (B.in_flight) {
        H1PA_HIP(hipStreamSynchronize(B.stream));
        std::copy(B.h_pinned, B.h_pinned + B.slots.size(), B.h_state.begin());
        B.in_flight = false;
        ++B.replays;
    }
    return running_slots(B);
}

bool batch_idle(Batch &B) {
    if (!B.in_flight) return true;
    const hipError_t e = hipStreamQuery(B.stream);
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: keep it out of later checks
        return false;
    }
    H1PA_HIP(e);
    return true;
}

// Runs up to `max_replays` replays (stops early once no slot is running).
// Returns the number of slots still running.
int batch_run(Batch &B, int max_replays) {
    batch_wait(B);
    for (int rp = 0; rp < max_replays; ++rp) {
        if (!batch_launch(B)) break;
        batch_wait(B);
    }
    return running_slots(B);
}

// Finished slots not yet reported: fills up to `cap` results and retires them.
int batch_poll(Batch &B, int cap, int64_t *tags, h1pa_batch_result *res) {
    batch_wait(B);
    int n = 0;
    const bool mixed = mixed_mode(B);
    for (size_t k = 0; k < B.slots.size() && n < cap; ++k) {
        BatchSlot &sl = *B.slots[k];
        const SlotState &st = B.h_state[k];
        if (!sl.live || st.phase != PH_DONE) continue;
        if (tags) tags[n] = sl.tag;
        if (res) {
            h1pa_batch_result &r = res[n];
            r.converged = st.converged;
            r.iterations = mixed ? st.iter : st.final_iter;
            r.outer_iterations = mixed ? st.outer : 0;
            r.final_norm = mixed ? std::sqrt(std::max(st.rho64, 0.0))
                                 : (st.betanom >= 0 ? std::sqrt(st.betanom) : st.betanom);
            r.rel_residual = mixed ? (st.b2 > 0 ? std::sqrt(st.r2 / st.b2) : 0.0) : -1.0;
            r.status = st.nonfinite ? H1PA_ENONFINITE : (st.converged ? H1PA_OK : H1PA_ENOTCONV);
        }
        ++n;
        // retire: free the slot's vectors, keep its position
        sl.live = false;
        --sl.form->f.batch_refs;
        for (auto &v : sl.v64) v.release();
        for (auto &v : sl.v32) v.release();
        sl.mask.release();
        sl.bytes = 0;
        B.h_state[k].phase = PH_EMPTY;
    }
    // nothing is in flight (batch_wait above): one synchronous state upload
    if (n)
        H1PA_HIP(hipMemcpy(B.d_state.p, B.h_state.data(), sizeof(SlotState) * B.slots.size(),
                           hipMemcpyHostToDevice));
    return n;
}

void batch_info(const Batch &B, int64_t *info) {
    int64_t live = (int64_t)B.pending.size(), running = live, bytes = 0;
    for (const auto &p : B.pending) bytes += (int64_t)p->bytes;
    for (size_t k = 0; k < B.slots.size(); ++k) {
        const BatchSlot &sl = *B.slots[k];
        if (!sl.live) continue;
        ++live;
        if (running_phase(B.h_state[k].phase)) ++running;
        bytes += (int64_t)sl.bytes;
    }
    info[0] = live;
    info[1] = running;
    info[2] = bytes;
    info[3] = B.replays;
    info[4] = B.captures;
    info[5] = (int64_t)B.groups.size();
}

}  // namespace h1pa
