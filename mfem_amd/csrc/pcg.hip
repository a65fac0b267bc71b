// pcg.hip -- Jacobi-preconditioned CG on the device, MFEM CGSolver semantics.
//
// CGSolver::Mult (linalg/solvers.cpp:869-1050) with OperatorJacobiSmoother
// (:401-453) on the DIAG_ONE ConstrainedOperator (linalg/operator.cpp:601-662).
// MFEM runs ~12 vector passes and two host round-trips (Dot -> device sync ->
// host sum, general/reducers.hpp:575-584) per iteration.  Here one iteration
// is four launches with every scalar kept on the device:
//   update : alpha = nom/den; x += alpha d; r -= alpha z; z = dinv r;
//            betanom = (r, z); convergence test; beta = betanom/nom
//   dir    : d = z + beta d
//   apply  : element kernel on d (constrained entries read as 0)
//   scatter: z = A d (DIAG_ONE rows copy d), den = (d, z)
// Dot products reduce deterministically: fixed grid, fixed in-block tree,
// and the last block to finish sums the block partials in index order.
#include <chrono>
#include <cmath>

#include "internal.hpp"

namespace h1pa {

struct PcgState {
    double nom, nom0, den, betanom, beta, r0;
    int iter;        // MFEM's i
    int in_loop;     // 0 before the first update, then 1
    int max_iter;
    int done;
    int converged;
    int final_iter;
    unsigned counter;
};

namespace kern {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

inline unsigned vec_blocks(int64_t n) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b > kMaxBlocks) b = kMaxBlocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// Block sum in a fixed tree order; result valid in thread 0.
__device__ inline double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}

// Writes this block's partial; returns true in the one block that finishes
// last, which then holds the fixed-order total in `total` (thread 0 valid).
__device__ inline bool reduce_partials(double part, double *partials, unsigned *counter,
                                       double *sh, double &total) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = part;
        __threadfence();
        const unsigned prev = atomicAdd(counter, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    const volatile double *vp = partials;
    double v = 0.0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += kThreads) v += vp[i];
    total = block_sum(v, sh);
    if (threadIdx.x == 0) *counter = 0u;
    return true;
}

__global__ void k_jacobi_dinv(int64_t n, const double *__restrict__ diag,
                              const uint8_t *__restrict__ mask, double *__restrict__ dinv) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    dinv[i] = (mask && mask[i]) ? 1.0 : 1.0 / diag[i];
}

__global__ void __launch_bounds__(kThreads)
k_pcg_init(int64_t n, const double *__restrict__ b, const double *__restrict__ dinv,
           double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
           double *__restrict__ d, double *__restrict__ partials, PcgState *st, double rel_tol,
           double abs_tol) {
    __shared__ double sh[kThreads];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double ri = b[i];
        const double zi = dinv[i] * ri;
        r[i] = ri;
        x[i] = 0.0;
        z[i] = zi;
        d[i] = zi;
        acc += zi * ri;
    }
    const double part = block_sum(acc, sh);
    double nom;
    if (reduce_partials(part, partials, &st->counter, sh, nom) && threadIdx.x == 0) {
        st->nom = st->nom0 = nom;
        st->iter = 1;
        st->in_loop = 0;
        double r0 = nom * rel_tol * rel_tol;
        if (abs_tol * abs_tol > r0) r0 = abs_tol * abs_tol;
        st->r0 = r0;
        st->final_iter = 0;
        st->converged = 0;
        st->done = 0;
        st->betanom = nom;
        if (nom < 0.0) {
            st->done = 1;
        } else if (nom <= r0) {
            st->done = 1;
            st->converged = 1;
        }
    }
}

// z = A d from the element output, DIAG_ONE rows copy d; den = (d, z).
__global__ void __launch_bounds__(kThreads)
k_pcg_scatter(int64_t n, const int *__restrict__ offsets, const int *__restrict__ indices,
              const double *__restrict__ ye, const uint8_t *__restrict__ mask,
              const double *__restrict__ d, double *__restrict__ z, double *__restrict__ partials,
              PcgState *st, const int *__restrict__ perm, int64_t nperm) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    double acc = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)kThreads + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * kThreads) {
        const int64_t i = t < nperm ? perm[t] : t;
        double v;
        if (mask && mask[i]) {
            v = d[i];
        } else {
            v = 0.0;
            for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
                const int ix = indices[j];
                v += ix >= 0 ? ye[ix] : -ye[-1 - ix];
            }
        }
        z[i] = v;
        acc += d[i] * v;
    }
    const double part = block_sum(acc, sh);
    double den;
    if (reduce_partials(part, partials, &st->counter, sh, den) && threadIdx.x == 0) {
        st->den = den;
        if (den == 0.0) {  // CGSolver: den <= 0 and den == 0 ends the loop
            st->done = 1;
            st->final_iter = st->in_loop ? st->iter : 0;
        }
    }
}

__global__ void __launch_bounds__(kThreads)
k_pcg_update(int64_t n, const double *__restrict__ dinv, double *__restrict__ x,
             double *__restrict__ r, double *__restrict__ z, const double *__restrict__ d,
             double *__restrict__ partials, PcgState *st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    const double alpha = st->nom / st->den;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double di = d[i];
        x[i] += alpha * di;
        const double ri = r[i] - alpha * z[i];
        r[i] = ri;
        const double zi = dinv[i] * ri;
        z[i] = zi;
        acc += ri * zi;
    }
    const double part = block_sum(acc, sh);
    double betanom;
    if (reduce_partials(part, partials, &st->counter, sh, betanom) && threadIdx.x == 0) {
        const int i = st->iter;
        st->in_loop = 1;
        st->betanom = betanom;
        if (betanom < 0.0) {
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
}

__global__ void __launch_bounds__(kThreads)
k_pcg_dir(int64_t n, const double *__restrict__ z, double *__restrict__ d, const PcgState *st) {
    if (st->done) return;
    const double beta = st->beta;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        d[i] = z[i] + beta * d[i];
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
    hipLaunchKernelGGL(k_jacobi_dinv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, diag,
                       mask, dinv);
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

// Element part of the constrained apply into f.ework0 (d -> A_E d).
static void pcg_element_apply(Form &f, const uint8_t *mask, const double *d, hipStream_t s) {
    Space &sp = *f.space;
    const bool diff = f.desc.integrators & H1PA_INTEG_DIFFUSION;
    const bool mass = f.desc.integrators & H1PA_INTEG_MASS;
    kern::EOut out;  // E-vector in native order, read back through indices_nat
    out.ye = f.ework0.p;
    out.l2n = sp.restr.lex2nat.p;
    if (diff)
        kern::diffusion_apply(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_diff.p, sp.restr.gather_map.p,
                              mask, d, out, false, s);
    if (mass)
        kern::mass_apply(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_mass.p, sp.restr.gather_map.p, mask, d,
                         out, diff, s);
}

void pcg_solve(Form &f, const uint8_t *mask, const double *diag, const double *b, double *x,
               const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, hipStream_t s) {
    Space &sp = *f.space;
    const int64_t n = sp.ndofs;
    DevBuf<double> r(n), z(n), d(n), dinv(n), partials(kern::kMaxBlocks);
    DevBuf<PcgState> st(1);
    PcgState h{};
    h.max_iter = o.max_iter;
    H1PA_HIP(hipMemcpyAsync(st.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    kern::jacobi_dinv(n, diag, mask, dinv.p, s);
    const unsigned nb = kern::vec_blocks(n);
    const int check = o.check_every > 0 ? o.check_every : 8;

    auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(kern::k_pcg_init, dim3(nb), dim3(kern::kThreads), 0, s, n, b, dinv.p, x, r.p,
                       z.p, d.p, partials.p, st.p, o.rel_tol, o.abs_tol);
    auto apply = [&]() {
        pcg_element_apply(f, mask, d.p, s);
        hipLaunchKernelGGL(kern::k_pcg_scatter, dim3(nb), dim3(kern::kThreads), 0, s, n,
                           sp.restr.offsets.p, sp.restr.indices_nat.p, f.ework0.p, mask, d.p, z.p,
                           partials.p, st.p, sp.restr.perm.p, sp.restr.nperm);
    };
    apply();
    H1PA_HIP(hipGetLastError());
    for (int it = 1; it <= o.max_iter; ++it) {
        hipLaunchKernelGGL(kern::k_pcg_update, dim3(nb), dim3(kern::kThreads), 0, s, n, dinv.p, x,
                           r.p, z.p, d.p, partials.p, st.p);
        hipLaunchKernelGGL(kern::k_pcg_dir, dim3(nb), dim3(kern::kThreads), 0, s, n, z.p, d.p, st.p);
        apply();
        H1PA_HIP(hipGetLastError());
        if (it % check == 0 || it == o.max_iter) {
            H1PA_HIP(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
            H1PA_HIP(hipStreamSynchronize(s));
            if (h.done) break;
        }
    }
    H1PA_HIP(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
    H1PA_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    stats.converged = h.converged;
    stats.iterations = h.done ? h.final_iter : o.max_iter;
    stats.initial_norm = h.nom0 >= 0.0 ? std::sqrt(h.nom0) : h.nom0;
    stats.final_norm = h.betanom >= 0.0 ? std::sqrt(h.betanom) : h.betanom;
    stats.seconds = std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace h1pa
