// apply.hip -- sum-factorized PA apply and diagonal kernels for H1 hexahedra.
//
// Math: y_e = sum_c (B/G)^T(x) (B/G)^T(y) (B/G)^T(z) [ D_e (grad u_e) ]_c with
// grad u_e = (G B B, B G B, B B G) u_e, D_e the symmetric 3x3 PA data at each
// quadrature point (fem/integ/bilininteg_diffusion_kernels.hpp:989-1221); the
// mass operator is B^T B^T B^T [ v_e B B B u_e ] (bilininteg_mass_kernels.hpp:
// 815-1049).
//
// CDNA4 mapping (one Q x Q thread plane per element, NB elements per block):
//  * x-lines: thread (dy,dz) holds u(:,dy,dz) in registers and produces the
//    whole Q-line of B u and G u -- the inputs are read once, not once per
//    output;
//  * y-lines: thread (qx,dz) reads its line from LDS (one transpose);
//  * z-lines: thread (qx,qy) keeps the z-line in registers, so the z
//    contraction, the pointwise 3x3 product and the backward z contraction
//    never touch LDS;
//  * every 1D contraction uses the even-odd split of the centro-symmetric B
//    and centro-antisymmetric G (apply_body.hpp): half the multiply-adds, and
//    the z-stage walks plane pairs (q, Q-1-q);
//  * B and G enter the FMAs as scalar (SGPR) operands.  The contracted index
//    is unrolled, the output-line loop is not: each iteration needs one row
//    of the even-odd tables, loaded with wide scalar loads, so the SGPR file
//    never spills;
//  * one LDS buffer of max(2 D^2 Q, 3 D Q^2) values per element, reused
//    across the stages with a barrier between the reads and the writes.
//
// basis layout: [ B(q,d) q-fastest | G q-fastest | B d-fastest | G d-fastest |
//                 even-odd tables (EO<D, Q> in apply_body.hpp) ]; the diagonal
// kernels read the first four blocks.
#include "apply_body.hpp"

namespace h1pa {
namespace kern {

namespace {

template <typename T, typename TP, int D, int Q, int NB, int U, int K, int W>
__global__ void __launch_bounds__(NB * Q * Q, W)
k_diffusion_apply(int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
                  const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
                  const T *__restrict__ x, const EOutT<T> out, int accumulate) {
    if (out.stop && *out.stop) return;  // solver already finished (block-uniform)
    __shared__ T smem[NB * diffusion_buf<D, Q>()];
    diffusion_elements<T, TP, D, Q, NB, U, K>((int64_t)blockIdx.x * NB, ne, basis, pa,
                                              gmap, mask, x, out, accumulate, smem);
}

template <typename T, typename TP, int D, int Q, int NB>
__global__ void __launch_bounds__(NB * Q * Q)
k_mass_apply(int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
             const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
             const T *__restrict__ x, const EOutT<T> out, int accumulate) {
    if (out.stop && *out.stop) return;
    __shared__ T smem[NB * D * Q * Q];
    mass_elements<T, TP, D, Q, NB>((int64_t)blockIdx.x * NB, ne, basis, pa, gmap, mask,
                                   x, out, accumulate, smem);
}

// Diagonal of the element diffusion matrices (PADiffusionDiagonal3D,
// bilininteg_diffusion_kernels.hpp:268-367), all nine (i,j) pairs grouped by
// the 1D factor in x: GG (O11), BB (O22, O33, 2 O23), GB (2 O12, 2 O13).
// NB elements per block (Q^2 threads each); the z loop is unrolled so that
// the PA planes' loads can be issued ahead of their use.
template <int D, int Q, int NB>
__global__ void __launch_bounds__(NB * Q * Q)
k_diffusion_diag(int64_t ne, const double *__restrict__ basis, const double *__restrict__ pa,
                 const EOut out, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q;
    constexpr int DQQ = D * Q * Q, DDQ = D * D * Q;
    __shared__ double smem[NB * 6 * DQQ];
    const double *__restrict__ Bq = basis;
    const double *__restrict__ Gq = basis + Q * D;
    const double *__restrict__ Bd = basis + 2 * Q * D;
    const double *__restrict__ Gd = basis + 3 * Q * D;
    const int le = threadIdx.x / QQ, t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = (int64_t)blockIdx.x * NB + le;
    const bool valid = e < ne;
    double *s = smem + le * 6 * DQQ;
    if (valid) {   // z: thread (qx, qy)
        const double *p = pa + (int64_t)QQQ * 6 * e + a + Q * b;
        double acc[6][D];
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int dz = 0; dz < D; ++dz) acc[k][dz] = 0.0;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) {
            const double *br = Bd + D * qz, *gr = Gd + D * qz;
            const double O11 = p[QQQ * 0 + QQ * qz], O12 = p[QQQ * 1 + QQ * qz],
                         O13 = p[QQQ * 2 + QQ * qz], O22 = p[QQQ * 3 + QQ * qz],
                         O23 = p[QQQ * 4 + QQ * qz], O33 = p[QQQ * 5 + QQ * qz];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                const double bb = br[dz] * br[dz], gg = gr[dz] * gr[dz], gb = gr[dz] * br[dz];
                acc[0][dz] = fma(bb, O11, acc[0][dz]);  // -> x: GG
                acc[1][dz] = fma(bb, O22, acc[1][dz]);  // -> x: BB, y: GG
                acc[2][dz] = fma(gg, O33, acc[2][dz]);  // -> x: BB, y: BB
                acc[3][dz] = fma(gb, O23, acc[3][dz]);  // -> x: BB, y: GB
                acc[4][dz] = fma(bb, O12, acc[4][dz]);  // -> x: GB, y: GB
                acc[5][dz] = fma(gb, O13, acc[5][dz]);  // -> x: GB, y: BB
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int dz = 0; dz < D; ++dz) s[k * DQQ + a + Q * (b + Q * dz)] = acc[k][dz];
    }
    __syncthreads();
    double ygg[D], ybb[D], ygb[D];
    const bool acty = valid && b < D;
    if (acty) {   // y: thread (qx, dz)
#pragma unroll
        for (int dy = 0; dy < D; ++dy) ygg[dy] = ybb[dy] = ygb[dy] = 0.0;
#pragma unroll 1
        for (int qy = 0; qy < Q; ++qy) {
            const double *br = Bd + D * qy, *gr = Gd + D * qy;
            const int id = a + Q * (qy + Q * b);
            const double A11 = s[0 * DQQ + id], A22 = s[1 * DQQ + id], A33 = s[2 * DQQ + id],
                         A23 = s[3 * DQQ + id], A12 = s[4 * DQQ + id], A13 = s[5 * DQQ + id];
#pragma unroll
            for (int dy = 0; dy < D; ++dy) {
                const double bb = br[dy] * br[dy], gg = gr[dy] * gr[dy], gb = gr[dy] * br[dy];
                ygg[dy] = fma(bb, A11, ygg[dy]);
                ybb[dy] = fma(gg, A22, fma(bb, A33, fma(2.0 * gb, A23, ybb[dy])));
                ygb[dy] = fma(2.0 * gb, A12, fma(2.0 * bb, A13, ygb[dy]));
            }
        }
    }
    __syncthreads();
    if (acty) {
#pragma unroll
        for (int dy = 0; dy < D; ++dy) {
            s[0 * DDQ + a + Q * (dy + D * b)] = ygg[dy];
            s[1 * DDQ + a + Q * (dy + D * b)] = ybb[dy];
            s[2 * DDQ + a + Q * (dy + D * b)] = ygb[dy];
        }
    }
    __syncthreads();
    if (valid && a < D && b < D) {   // x: thread (dy, dz)
        double yv[D];
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            const double *bc = Bq + Q * dx, *gc = Gq + Q * dx;
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) {
                const int id = qx + Q * (a + D * b);
                v = fma(gc[qx] * gc[qx], s[0 * DDQ + id], v);
                v = fma(bc[qx] * bc[qx], s[1 * DDQ + id], v);
                v = fma(gc[qx] * bc[qx], s[2 * DDQ + id], v);
            }
            yv[dx] = v;
        }
        store_line<D, double>(out, e, a, b, yv, accumulate);
    }
}

template <int D, int Q, int NB>
__global__ void __launch_bounds__(NB * Q * Q)
k_mass_diag(int64_t ne, const double *__restrict__ basis, const double *__restrict__ pa,
            const EOut out, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q;
    constexpr int DQQ = D * Q * Q;
    __shared__ double smem[NB * DQQ];
    const double *__restrict__ Bq = basis;
    const double *__restrict__ Bd = basis + 2 * Q * D;
    const int le = threadIdx.x / QQ, t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = (int64_t)blockIdx.x * NB + le;
    const bool valid = e < ne;
    double *s = smem + le * DQQ;
    if (valid) {
        const double *p = pa + (int64_t)QQQ * e + a + Q * b;
        double acc[D];
#pragma unroll
        for (int dz = 0; dz < D; ++dz) acc[dz] = 0.0;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) {
            const double *br = Bd + D * qz;
            const double v = p[QQ * qz];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) acc[dz] = fma(br[dz] * br[dz], v, acc[dz]);
        }
#pragma unroll
        for (int dz = 0; dz < D; ++dz) s[a + Q * (b + Q * dz)] = acc[dz];
    }
    __syncthreads();
    double y2[D];
    const bool acty = valid && b < D;
    if (acty) {
#pragma unroll
        for (int dy = 0; dy < D; ++dy) y2[dy] = 0.0;
#pragma unroll 1
        for (int qy = 0; qy < Q; ++qy) {
            const double *br = Bd + D * qy;
            const double v = s[a + Q * (qy + Q * b)];
#pragma unroll
            for (int dy = 0; dy < D; ++dy) y2[dy] = fma(br[dy] * br[dy], v, y2[dy]);
        }
    }
    __syncthreads();
    if (acty)
#pragma unroll
        for (int dy = 0; dy < D; ++dy) s[a + Q * (dy + D * b)] = y2[dy];
    __syncthreads();
    if (valid && a < D && b < D) {
        double yv[D];
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            const double *bc = Bq + Q * dx;
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v = fma(bc[qx] * bc[qx], s[qx + Q * (a + D * b)], v);
            yv[dx] = v;
        }
        store_line<D, double>(out, e, a, b, yv, accumulate);
    }
}

// ---- dispatch over (D, Q) and variants -------------------------------------
// W: minimum waves per SIMD the compiler must allow (1: no bound; the
// occupancy-capped variants pass lds_waves()).
template <typename T, typename TP, int D, int Q, int NB, int U, int K, int W = 1>
void launch_diffusion_v(int64_t ne, const T *basis, const TP *pa, const int *gmap,
                        const uint8_t *mask, const T *x, const EOutT<T> &out, bool acc,
                        hipStream_t s) {
    const unsigned grid = (unsigned)((ne + NB - 1) / NB);
    hipLaunchKernelGGL((k_diffusion_apply<T, TP, D, Q, NB, U, K, W>), dim3(grid), dim3(NB * Q * Q),
                       0, s, ne, basis, pa, gmap, mask, x, out, acc ? 1 : 0);
}

// Variants of the diffusion apply: (elements per block, line-loop unroll,
// PA-data pipeline depth: 0, 2 planes or all of them; 6 and 7 are 0 and 4
// with the registers capped so that as many blocks stay resident as the LDS
// allows; 8 is 4 with half the elements per block).  Every variant performs
// the same arithmetic per element in the same order, so they give
// bit-identical results and differ only in speed.
constexpr int kDiffVariants = 9;
template <int D, int Q>
void launch_diffusion(int variant, int64_t ne, const double *basis, const double *pa,
                      const int *gmap, const uint8_t *mask, const double *x, const EOut &out,
                      bool acc, hipStream_t s) {
    using T = double;
    using TP = double;
    constexpr int NBmax = nb_for(Q, 3 * D * Q * Q, sizeof(T));
    constexpr int NBhalf = NBmax / 2 > 0 ? NBmax / 2 : 1;
    switch (variant) {
        case 1: launch_diffusion_v<T, TP, D, Q, NBmax, 2, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 2: launch_diffusion_v<T, TP, D, Q, NBhalf, 1, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 3: launch_diffusion_v<T, TP, D, Q, NBhalf, 2, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 4: launch_diffusion_v<T, TP, D, Q, NBmax, 1, 2>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 5: launch_diffusion_v<T, TP, D, Q, NBmax, 1, kAllPlanes>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 6: launch_diffusion_v<T, TP, D, Q, NBmax, 1, 0, lds_waves<T, D, Q, NBmax>()>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 7: launch_diffusion_v<T, TP, D, Q, NBmax, 1, 2, lds_waves<T, D, Q, NBmax>()>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 8: launch_diffusion_v<T, TP, D, Q, NBhalf, 1, 2>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        default: launch_diffusion_v<T, TP, D, Q, NBmax, 1, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
    }
}
template <typename T, typename TP, int D, int Q>
void launch_mass(int64_t ne, const T *basis, const TP *pa, const int *gmap, const uint8_t *mask,
                 const T *x, const EOutT<T> &out, bool acc, hipStream_t s) {
    constexpr int NB = nb_for(Q, D * Q * Q, sizeof(T));
    const unsigned grid = (unsigned)((ne + NB - 1) / NB);
    hipLaunchKernelGGL((k_mass_apply<T, TP, D, Q, NB>), dim3(grid), dim3(NB * Q * Q), 0, s, ne,
                       basis, pa, gmap, mask, x, out, acc ? 1 : 0);
}

// Reduced-precision (FP32 vectors and arithmetic; PA data stored as FP32 or
// BF16) diffusion apply, compiled for the MFEM-default Q = D + 1 only.
template <typename TP, int D, int Q>
void launch_diffusion_lp(int variant, int64_t ne, const float *basis, const TP *pa,
                         const int *gmap, const uint8_t *mask, const float *x,
                         const EOutT<float> &out, bool acc, hipStream_t s) {
    constexpr int NBmax = nb_for(Q, 3 * D * Q * Q, sizeof(float));
    constexpr int NBhalf = NBmax / 2 > 0 ? NBmax / 2 : 1;
    if (variant == 8)
        launch_diffusion_v<float, TP, D, Q, NBhalf, 1, 2>(ne, basis, pa, gmap, mask, x, out, acc, s);
    else if (variant == 1)
        launch_diffusion_v<float, TP, D, Q, NBmax, 2, 0>(ne, basis, pa, gmap, mask, x, out, acc, s);
    else if (variant == 4)
        launch_diffusion_v<float, TP, D, Q, NBmax, 1, 2>(ne, basis, pa, gmap, mask, x, out, acc, s);
    else if (variant == 5)
        launch_diffusion_v<float, TP, D, Q, NBmax, 1, kAllPlanes>(ne, basis, pa, gmap, mask, x,
                                                                  out, acc, s);
    else if (variant == 6)
        launch_diffusion_v<float, TP, D, Q, NBmax, 1, 0, lds_waves<float, D, Q, NBmax>()>(
            ne, basis, pa, gmap, mask, x, out, acc, s);
    else if (variant == 7)
        launch_diffusion_v<float, TP, D, Q, NBmax, 1, 2, lds_waves<float, D, Q, NBmax>()>(
            ne, basis, pa, gmap, mask, x, out, acc, s);
    else
        launch_diffusion_v<float, TP, D, Q, NBmax, 1, 0>(ne, basis, pa, gmap, mask, x, out, acc, s);
}
template <int D, int Q>
void launch_diff_diag(int64_t ne, const double *basis, const double *pa, const EOut &out,
                      bool acc, hipStream_t s) {
    constexpr int NB = nb_for(Q, 6 * D * Q * Q, sizeof(double));
    hipLaunchKernelGGL((k_diffusion_diag<D, Q, NB>), dim3((unsigned)((ne + NB - 1) / NB)),
                       dim3(NB * Q * Q), 0, s, ne, basis, pa, out, acc ? 1 : 0);
}
template <int D, int Q>
void launch_mass_diag(int64_t ne, const double *basis, const double *pa, const EOut &out,
                      bool acc, hipStream_t s) {
    constexpr int NB = nb_for(Q, D * Q * Q, sizeof(double));
    hipLaunchKernelGGL((k_mass_diag<D, Q, NB>), dim3((unsigned)((ne + NB - 1) / NB)),
                       dim3(NB * Q * Q), 0, s, ne, basis, pa, out, acc ? 1 : 0);
}

// (D, Q) pairs compiled: Q = D (q = p+1) and Q = D + 1 (q = p+2, MFEM default).
#define H1PA_FOR_EACH_DQ(X) \
    X(2, 2) X(2, 3) X(3, 3) X(3, 4) X(4, 4) X(4, 5) X(5, 5) X(5, 6) X(6, 6) X(6, 7) \
    X(7, 7) X(7, 8) X(8, 8) X(8, 9) X(9, 9) X(9, 10)

// Selected diffusion variant per (D, Q); filled by the autotuner, which may
// run on another host thread (a setup thread) while launches read the table.
std::atomic<int> g_diff_variant[11][11] = {};
// Reduced-precision applies: per PA storage kind (0 FP32, 1 BF16) and (D, Q).
std::atomic<int> g_diff_variant_lp[2][11][11] = {};

}  // namespace

bool apply_supported(int p, int q1d) { return p >= 1 && p <= 8 && (q1d == p + 1 || q1d == p + 2); }

int diffusion_variant_count() { return kDiffVariants; }
void set_diffusion_variant(int p, int q1d, int v) {
    if (p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10) g_diff_variant[p + 1][q1d].store(v);
}
int get_diffusion_variant(int p, int q1d) {
    return (p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10) ? g_diff_variant[p + 1][q1d].load() : 0;
}

void diffusion_apply(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                     const int *gmap, const uint8_t *mask, const double *x, const EOut &out,
                     bool accumulate, hipStream_t s) {
    diffusion_apply_variant(get_diffusion_variant(p, q1d), p, q1d, ne, basis, pa, gmap, mask, x,
                            out, accumulate, s);
}

void diffusion_apply_variant(int variant, int p, int q1d, int64_t ne, const double *basis,
                             const double *pa, const int *gmap, const uint8_t *mask,
                             const double *x, const EOut &out, bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                           \
    if (D == DD && q1d == QQ) {                                                             \
        launch_diffusion<DD, QQ>(variant, ne, basis, pa, gmap, mask, x, out, accumulate, s);     \
        H1PA_HIP(hipGetLastError());                                                        \
        return;                                                                             \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "diffusion apply: unsupported (p, q1d)"};
}

void mass_apply(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                const int *gmap, const uint8_t *mask, const double *x, const EOut &out,
                bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_mass<double, double, DD, QQ>(ne, basis, pa, gmap, mask, x, out, accumulate, s); \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "mass apply: unsupported (p, q1d)"};
}

void diffusion_diagonal(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                        const EOut &out, bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_diff_diag<DD, QQ>(ne, basis, pa, out, accumulate, s);                       \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "diffusion diagonal: unsupported (p, q1d)"};
}

void mass_diagonal(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                   const EOut &out, bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_mass_diag<DD, QQ>(ne, basis, pa, out, accumulate, s);                       \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "mass diagonal: unsupported (p, q1d)"};
}

#define H1PA_FOR_EACH_DQ_DEFAULT(X) \
    X(2, 3) X(3, 4) X(4, 5) X(5, 6) X(6, 7) X(7, 8) X(8, 9) X(9, 10)

void set_diffusion_variant_lp(int pa_kind, int p, int q1d, int v) {
    if ((pa_kind == 0 || pa_kind == 1) && p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10)
        g_diff_variant_lp[pa_kind][p + 1][q1d].store(v);
}
int get_diffusion_variant_lp(int pa_kind, int p, int q1d) {
    return ((pa_kind == 0 || pa_kind == 1) && p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10)
               ? g_diff_variant_lp[pa_kind][p + 1][q1d].load()
               : 0;
}

void diffusion_apply_lp(int p, int q1d, int64_t ne, const float *basis, const void *pa,
                        int pa_kind, const int *gmap, const uint8_t *mask, const float *x,
                        const EOutT<float> &out, bool accumulate, hipStream_t s) {
    diffusion_apply_lp_variant(get_diffusion_variant_lp(pa_kind, p, q1d), p, q1d, ne, basis, pa,
                               pa_kind, gmap, mask, x, out, accumulate, s);
}

void diffusion_apply_lp_variant(int v, int p, int q1d, int64_t ne, const float *basis,
                                const void *pa, int pa_kind, const int *gmap,
                                const uint8_t *mask, const float *x, const EOutT<float> &out,
                                bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        if (pa_kind == 1)                                                                  \
            launch_diffusion_lp<bf16_t, DD, QQ>(v, ne, basis, (const bf16_t *)pa, gmap,    \
                                                mask, x, out, accumulate, s);              \
        else                                                                               \
            launch_diffusion_lp<float, DD, QQ>(v, ne, basis, (const float *)pa, gmap, mask, \
                                               x, out, accumulate, s);                     \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ_DEFAULT(X)
#undef X
    throw Error{H1PA_EUNSUP, "reduced-precision diffusion apply needs q1d = p + 2"};
}

void mass_apply_lp(int p, int q1d, int64_t ne, const float *basis, const void *pa, int pa_kind,
                   const int *gmap, const uint8_t *mask, const float *x, const EOutT<float> &out,
                   bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        if (pa_kind == 1)                                                                  \
            launch_mass<float, bf16_t, DD, QQ>(ne, basis, (const bf16_t *)pa, gmap, mask, x, \
                                               out, accumulate, s);                        \
        else                                                                               \
            launch_mass<float, float, DD, QQ>(ne, basis, (const float *)pa, gmap, mask, x, \
                                              out, accumulate, s);                         \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ_DEFAULT(X)
#undef X
    throw Error{H1PA_EUNSUP, "reduced-precision mass apply needs q1d = p + 2"};
}

// FP64 vectors and arithmetic with BF16-stored PA data (the applies of a
// mixed-precision multigrid cycle): one variant, the PA planes requested two
// ahead, compiled for the default Q = D + 1.
void diffusion_apply_d_bf16(int p, int q1d, int64_t ne, const double *basis, const void *pa,
                            const int *gmap, const uint8_t *mask, const double *x,
                            const EOut &out, bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        constexpr int NB = nb_for(QQ, 3 * DD * QQ * QQ, sizeof(double));                   \
        launch_diffusion_v<double, bf16_t, DD, QQ, NB, 1, 2>(ne, basis, (const bf16_t *)pa, \
                                                             gmap, mask, x, out, accumulate, s); \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ_DEFAULT(X)
#undef X
    throw Error{H1PA_EUNSUP, "BF16-PA apply needs q1d = p + 2"};
}

}  // namespace kern
}  // namespace h1pa
