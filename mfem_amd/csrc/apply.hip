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
//  * the PA data of a thread's (qx,qy) column is loaded at kernel entry, so
//    its HBM latency overlaps the x/y stages;
//  * B and G are read with uniform addresses (scalar loads, SGPR operands of
//    v_fma_f64);
//  * one LDS buffer of max(2 D^2 Q, 3 D Q^2) doubles per element, reused
//    across the stages with a barrier between the reads and the writes.
#include "internal.hpp"

namespace h1pa {
namespace kern {

namespace {

constexpr int nb_for(int Q, int buf_doubles) {
    int nb = 256 / (Q * Q);
    if (nb < 1) nb = 1;
    while (nb > 1 && nb * buf_doubles * 8 > 48 * 1024) --nb;
    return nb;
}

// Read x at E-DOF `i` through the gather map (L-vector input) or directly
// (E-vector input); constrained L-DOFs read as zero.
__device__ __forceinline__ double load_x(const int *__restrict__ gmap,
                                         const uint8_t *__restrict__ mask,
                                         const double *__restrict__ x, int64_t i) {
    if (!gmap) return x[i];
    const int g = gmap[i];
    const int j = g >= 0 ? g : -1 - g;
    if (mask && mask[j]) return 0.0;
    const double v = x[j];
    return g >= 0 ? v : -v;
}

template <int D, int Q, int NB, bool PF>
__global__ void __launch_bounds__(NB * Q * Q)
k_diffusion_apply(int64_t ne, const double *__restrict__ Bm, const double *__restrict__ Gm,
                  const double *__restrict__ pa, const int *__restrict__ gmap,
                  const uint8_t *__restrict__ mask, const double *__restrict__ x,
                  double *__restrict__ ye, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DDQ = D * D * Q, DQQ = D * Q * Q;
    constexpr int BUF = (2 * DDQ > 3 * DQQ) ? 2 * DDQ : 3 * DQQ;
    __shared__ double smem[NB * BUF];
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = (int64_t)blockIdx.x * NB + le;
    const bool valid = e < ne;
    double *s = smem + le * BUF;
#define B_(q, d) Bm[(q) + Q * (d)]
#define G_(q, d) Gm[(q) + Q * (d)]

    // PA data of column (a,b): 6 comps x Q points.  PF: all issued at entry
    // so their latency overlaps the x/y stages; otherwise read per qz.
    const double *pcol = pa + (int64_t)QQQ * 6 * (valid ? e : 0) + a + Q * b;
    double Dq[6][PF ? Q : 1];
    if (PF && valid) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int qz = 0; qz < (PF ? Q : 1); ++qz) Dq[c][qz] = pcol[QQQ * c + QQ * qz];
    }

    // ---- forward x: thread (dy, dz) = (a, b)
    if (valid && a < D && b < D) {
        double u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) u[dx] = load_x(gmap, mask, x, base + dx);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
            double bu = 0.0, gu = 0.0;
#pragma unroll
            for (int dx = 0; dx < D; ++dx) {
                bu += B_(qx, dx) * u[dx];
                gu += G_(qx, dx) * u[dx];
            }
            s[qx + Q * (a + D * b)] = bu;
            s[DDQ + qx + Q * (a + D * b)] = gu;
        }
    }
    __syncthreads();

    // ---- forward y: thread (qx, dz) = (a, b)
    {
        double r0[D], r1[D];
        const bool act = valid && b < D;
        if (act) {
#pragma unroll
            for (int dy = 0; dy < D; ++dy) {
                r0[dy] = s[a + Q * (dy + D * b)];
                r1[dy] = s[DDQ + a + Q * (dy + D * b)];
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int qy = 0; qy < Q; ++qy) {
                double bb = 0.0, gb = 0.0, bg = 0.0;
#pragma unroll
                for (int dy = 0; dy < D; ++dy) {
                    bb += B_(qy, dy) * r0[dy];
                    gb += G_(qy, dy) * r0[dy];
                    bg += B_(qy, dy) * r1[dy];
                }
                s[0 * DQQ + a + Q * (qy + Q * b)] = bg;  // -> d/dx
                s[1 * DQQ + a + Q * (qy + Q * b)] = gb;  // -> d/dy
                s[2 * DQQ + a + Q * (qy + Q * b)] = bb;  // -> d/dz
            }
        }
    }
    __syncthreads();

    // ---- z in registers: thread (qx, qy) = (a, b)
    {
        double z0[D], z1[D], z2[D];
        if (valid) {
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                z0[dz] = s[0 * DQQ + a + Q * (b + Q * dz)];
                z1[dz] = s[1 * DQQ + a + Q * (b + Q * dz)];
                z2[dz] = s[2 * DQQ + a + Q * (b + Q * dz)];
            }
        }
        __syncthreads();
        if (valid) {
            double T0[D], T1[D], T2[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) T0[dz] = T1[dz] = T2[dz] = 0.0;
#pragma unroll
            for (int qz = 0; qz < Q; ++qz) {
                double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) {
                    g0 += B_(qz, dz) * z0[dz];
                    g1 += B_(qz, dz) * z1[dz];
                    g2 += G_(qz, dz) * z2[dz];
                }
                double d6[6];
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    d6[c] = PF ? Dq[c][PF ? qz : 0] : pcol[QQQ * c + QQ * qz];
                const double O0 = d6[0] * g0 + d6[1] * g1 + d6[2] * g2;
                const double O1 = d6[1] * g0 + d6[3] * g1 + d6[4] * g2;
                const double O2 = d6[2] * g0 + d6[4] * g1 + d6[5] * g2;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) {
                    T0[dz] += B_(qz, dz) * O0;
                    T1[dz] += B_(qz, dz) * O1;
                    T2[dz] += G_(qz, dz) * O2;
                }
            }
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                s[0 * DQQ + a + Q * (b + Q * dz)] = T0[dz];
                s[1 * DQQ + a + Q * (b + Q * dz)] = T1[dz];
                s[2 * DQQ + a + Q * (b + Q * dz)] = T2[dz];
            }
        }
    }
    __syncthreads();

    // ---- backward y: thread (qx, dz) = (a, b)
    {
        double t0[Q], t1[Q], t2[Q];
        const bool act = valid && b < D;
        if (act) {
#pragma unroll
            for (int qy = 0; qy < Q; ++qy) {
                t0[qy] = s[0 * DQQ + a + Q * (qy + Q * b)];
                t1[qy] = s[1 * DQQ + a + Q * (qy + Q * b)];
                t2[qy] = s[2 * DQQ + a + Q * (qy + Q * b)];
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int dy = 0; dy < D; ++dy) {
                double sa = 0.0, sb = 0.0;
#pragma unroll
                for (int qy = 0; qy < Q; ++qy) {
                    sa += B_(qy, dy) * t0[qy];
                    sb += G_(qy, dy) * t1[qy] + B_(qy, dy) * t2[qy];
                }
                s[a + Q * (dy + D * b)] = sa;
                s[DDQ + a + Q * (dy + D * b)] = sb;
            }
        }
    }
    __syncthreads();

    // ---- backward x: thread (dy, dz) = (a, b)
    if (valid && a < D && b < D) {
        double sa[Q], sb[Q];
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
            sa[qx] = s[qx + Q * (a + D * b)];
            sb[qx] = s[DDQ + qx + Q * (a + D * b)];
        }
        double *yp = ye + (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v += G_(qx, dx) * sa[qx] + B_(qx, dx) * sb[qx];
            if (accumulate) yp[dx] += v;
            else yp[dx] = v;
        }
    }
#undef B_
#undef G_
}

template <int D, int Q>
__global__ void __launch_bounds__(nb_for(Q, D * Q * Q) * Q * Q)
k_mass_apply(int64_t ne, const double *__restrict__ Bm, const double *__restrict__ pa,
             const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
             const double *__restrict__ x, double *__restrict__ ye, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DDQ = D * D * Q, DQQ = D * Q * Q;
    constexpr int BUF = DQQ;
    constexpr int NB = nb_for(Q, DQQ);
    __shared__ double smem[NB * BUF];
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = (int64_t)blockIdx.x * NB + le;
    const bool valid = e < ne;
    double *s = smem + le * BUF;
#define B_(q, d) Bm[(q) + Q * (d)]
    double vq[Q];
    if (valid) {
        const double *p = pa + (int64_t)QQQ * e + a + Q * b;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) vq[qz] = p[QQ * qz];
    }
    if (valid && a < D && b < D) {
        double u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) u[dx] = load_x(gmap, mask, x, base + dx);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
            double bu = 0.0;
#pragma unroll
            for (int dx = 0; dx < D; ++dx) bu += B_(qx, dx) * u[dx];
            s[qx + Q * (a + D * b)] = bu;
        }
    }
    __syncthreads();
    {
        double r[D];
        const bool act = valid && b < D;
        if (act)
#pragma unroll
            for (int dy = 0; dy < D; ++dy) r[dy] = s[a + Q * (dy + D * b)];
        __syncthreads();
        if (act)
#pragma unroll
            for (int qy = 0; qy < Q; ++qy) {
                double v = 0.0;
#pragma unroll
                for (int dy = 0; dy < D; ++dy) v += B_(qy, dy) * r[dy];
                s[a + Q * (qy + Q * b)] = v;
            }
    }
    __syncthreads();
    {
        double z[D];
        if (valid)
#pragma unroll
            for (int dz = 0; dz < D; ++dz) z[dz] = s[a + Q * (b + Q * dz)];
        __syncthreads();
        if (valid) {
            double T[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) T[dz] = 0.0;
#pragma unroll
            for (int qz = 0; qz < Q; ++qz) {
                double v = 0.0;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) v += B_(qz, dz) * z[dz];
                v *= vq[qz];
#pragma unroll
                for (int dz = 0; dz < D; ++dz) T[dz] += B_(qz, dz) * v;
            }
#pragma unroll
            for (int dz = 0; dz < D; ++dz) s[a + Q * (b + Q * dz)] = T[dz];
        }
    }
    __syncthreads();
    {
        double tq[Q];
        const bool act = valid && b < D;
        if (act)
#pragma unroll
            for (int qy = 0; qy < Q; ++qy) tq[qy] = s[a + Q * (qy + Q * b)];
        __syncthreads();
        if (act)
#pragma unroll
            for (int dy = 0; dy < D; ++dy) {
                double v = 0.0;
#pragma unroll
                for (int qy = 0; qy < Q; ++qy) v += B_(qy, dy) * tq[qy];
                s[a + Q * (dy + D * b)] = v;
            }
    }
    __syncthreads();
    if (valid && a < D && b < D) {
        double sx[Q];
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) sx[qx] = s[qx + Q * (a + D * b)];
        double *yp = ye + (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v += B_(qx, dx) * sx[qx];
            if (accumulate) yp[dx] += v;
            else yp[dx] = v;
        }
    }
#undef B_
    (void)DDQ;
}

// Diagonal of the element diffusion matrices (PADiffusionDiagonal3D,
// bilininteg_diffusion_kernels.hpp:268-367), all nine (i,j) pairs grouped by
// the 1D factor in x: GG (O11), BB (O22, O33, 2 O23), GB (2 O12, 2 O13).
template <int D, int Q>
__global__ void __launch_bounds__(Q * Q)
k_diffusion_diag(int64_t ne, const double *__restrict__ Bm, const double *__restrict__ Gm,
                 const double *__restrict__ pa, double *__restrict__ ye, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DQQ = D * Q * Q, DDQ = D * D * Q;
    __shared__ double s[6 * DQQ];
    const int a = threadIdx.x % Q, b = threadIdx.x / Q;
    const int64_t e = blockIdx.x;
#define B_(q, d) Bm[(q) + Q * (d)]
#define G_(q, d) Gm[(q) + Q * (d)]
    {   // z: thread (qx, qy)
        const double *p = pa + (int64_t)QQQ * 6 * e + a + Q * b;
        double acc[6][D];
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int dz = 0; dz < D; ++dz) acc[k][dz] = 0.0;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) {
            const double O11 = p[QQQ * 0 + QQ * qz], O12 = p[QQQ * 1 + QQ * qz],
                         O13 = p[QQQ * 2 + QQ * qz], O22 = p[QQQ * 3 + QQ * qz],
                         O23 = p[QQQ * 4 + QQ * qz], O33 = p[QQQ * 5 + QQ * qz];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                const double bb = B_(qz, dz) * B_(qz, dz), gg = G_(qz, dz) * G_(qz, dz),
                             gb = G_(qz, dz) * B_(qz, dz);
                acc[0][dz] += bb * O11;  // -> x: GG
                acc[1][dz] += bb * O22;  // -> x: BB, y: GG
                acc[2][dz] += gg * O33;  // -> x: BB, y: BB
                acc[3][dz] += gb * O23;  // -> x: BB, y: GB
                acc[4][dz] += bb * O12;  // -> x: GB, y: GB
                acc[5][dz] += gb * O13;  // -> x: GB, y: BB
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int dz = 0; dz < D; ++dz) s[k * DQQ + a + Q * (b + Q * dz)] = acc[k][dz];
    }
    __syncthreads();
    double ygg[D], ybb[D], ygb[D];
    const bool acty = b < D;
    if (acty) {   // y: thread (qx, dz)
#pragma unroll
        for (int dy = 0; dy < D; ++dy) ygg[dy] = ybb[dy] = ygb[dy] = 0.0;
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) {
            const int id = a + Q * (qy + Q * b);
            const double A11 = s[0 * DQQ + id], A22 = s[1 * DQQ + id], A33 = s[2 * DQQ + id],
                         A23 = s[3 * DQQ + id], A12 = s[4 * DQQ + id], A13 = s[5 * DQQ + id];
#pragma unroll
            for (int dy = 0; dy < D; ++dy) {
                const double bb = B_(qy, dy) * B_(qy, dy), gg = G_(qy, dy) * G_(qy, dy),
                             gb = G_(qy, dy) * B_(qy, dy);
                ygg[dy] += bb * A11;
                ybb[dy] += gg * A22 + bb * A33 + 2.0 * gb * A23;
                ygb[dy] += 2.0 * gb * A12 + 2.0 * bb * A13;
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
    if (a < D && b < D) {   // x: thread (dy, dz)
        double *yp = ye + (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) {
                const int id = qx + Q * (a + D * b);
                v += G_(qx, dx) * G_(qx, dx) * s[0 * DDQ + id] +
                     B_(qx, dx) * B_(qx, dx) * s[1 * DDQ + id] +
                     G_(qx, dx) * B_(qx, dx) * s[2 * DDQ + id];
            }
            if (accumulate) yp[dx] += v;
            else yp[dx] = v;
        }
    }
#undef B_
#undef G_
}

template <int D, int Q>
__global__ void __launch_bounds__(Q * Q)
k_mass_diag(int64_t ne, const double *__restrict__ Bm, const double *__restrict__ pa,
            double *__restrict__ ye, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DQQ = D * Q * Q, DDQ = D * D * Q;
    __shared__ double s[DQQ];
    const int a = threadIdx.x % Q, b = threadIdx.x / Q;
    const int64_t e = blockIdx.x;
#define B_(q, d) Bm[(q) + Q * (d)]
    {
        const double *p = pa + (int64_t)QQQ * e + a + Q * b;
        double acc[D];
#pragma unroll
        for (int dz = 0; dz < D; ++dz) acc[dz] = 0.0;
#pragma unroll
        for (int qz = 0; qz < Q; ++qz) {
            const double v = p[QQ * qz];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) acc[dz] += B_(qz, dz) * B_(qz, dz) * v;
        }
#pragma unroll
        for (int dz = 0; dz < D; ++dz) s[a + Q * (b + Q * dz)] = acc[dz];
    }
    __syncthreads();
    double y2[D];
    const bool acty = b < D;
    if (acty) {
#pragma unroll
        for (int dy = 0; dy < D; ++dy) y2[dy] = 0.0;
#pragma unroll
        for (int qy = 0; qy < Q; ++qy) {
            const double v = s[a + Q * (qy + Q * b)];
#pragma unroll
            for (int dy = 0; dy < D; ++dy) y2[dy] += B_(qy, dy) * B_(qy, dy) * v;
        }
    }
    __syncthreads();
    if (acty)
#pragma unroll
        for (int dy = 0; dy < D; ++dy) s[a + Q * (dy + D * b)] = y2[dy];
    __syncthreads();
    if (a < D && b < D) {
        double *yp = ye + (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v += B_(qx, dx) * B_(qx, dx) * s[qx + Q * (a + D * b)];
            if (accumulate) yp[dx] += v;
            else yp[dx] = v;
        }
    }
#undef B_
    (void)DDQ;
}

// ---- dispatch over (D, Q) -----------------------------------------------------
template <int D, int Q, int NB, bool PF>
void launch_diffusion_v(int64_t ne, const double *B, const double *G, const double *pa,
                        const int *gmap, const uint8_t *mask, const double *x, double *ye,
                        bool acc, hipStream_t s) {
    const unsigned grid = (unsigned)((ne + NB - 1) / NB);
    hipLaunchKernelGGL((k_diffusion_apply<D, Q, NB, PF>), dim3(grid), dim3(NB * Q * Q), 0, s, ne, B,
                       G, pa, gmap, mask, x, ye, acc ? 1 : 0);
}

// Variants of the diffusion apply: (elements per block, prefetch).  Every
// variant performs the same arithmetic per element in the same order, so
// they give bit-identical results and only differ in speed.
constexpr int kDiffVariants = 4;
template <int D, int Q>
void launch_diffusion(int variant, int64_t ne, const double *B, const double *G,
                      const double *pa, const int *gmap, const uint8_t *mask, const double *x,
                      double *ye, bool acc, hipStream_t s) {
    constexpr int NBmax = nb_for(Q, 3 * D * Q * Q);
    constexpr int NBhalf = NBmax / 2 > 0 ? NBmax / 2 : 1;
    switch (variant) {
        case 1: launch_diffusion_v<D, Q, NBmax, false>(ne, B, G, pa, gmap, mask, x, ye, acc, s); break;
        case 2: launch_diffusion_v<D, Q, NBhalf, true>(ne, B, G, pa, gmap, mask, x, ye, acc, s); break;
        case 3: launch_diffusion_v<D, Q, NBhalf, false>(ne, B, G, pa, gmap, mask, x, ye, acc, s); break;
        default: launch_diffusion_v<D, Q, NBmax, true>(ne, B, G, pa, gmap, mask, x, ye, acc, s); break;
    }
}
template <int D, int Q>
void launch_mass(int64_t ne, const double *B, const double *pa, const int *gmap,
                 const uint8_t *mask, const double *x, double *ye, bool acc, hipStream_t s) {
    constexpr int NB = nb_for(Q, D * Q * Q);
    const unsigned grid = (unsigned)((ne + NB - 1) / NB);
    hipLaunchKernelGGL((k_mass_apply<D, Q>), dim3(grid), dim3(NB * Q * Q), 0, s, ne, B, pa, gmap,
                       mask, x, ye, acc ? 1 : 0);
}
template <int D, int Q>
void launch_diff_diag(int64_t ne, const double *B, const double *G, const double *pa, double *ye,
                      bool acc, hipStream_t s) {
    hipLaunchKernelGGL((k_diffusion_diag<D, Q>), dim3((unsigned)ne), dim3(Q * Q), 0, s, ne, B, G,
                       pa, ye, acc ? 1 : 0);
}
template <int D, int Q>
void launch_mass_diag(int64_t ne, const double *B, const double *pa, double *ye, bool acc,
                      hipStream_t s) {
    hipLaunchKernelGGL((k_mass_diag<D, Q>), dim3((unsigned)ne), dim3(Q * Q), 0, s, ne, B, pa, ye,
                       acc ? 1 : 0);
}

// (D, Q) pairs compiled: Q = D (q = p+1) and Q = D + 1 (q = p+2, MFEM default).
#define H1PA_FOR_EACH_DQ(X) \
    X(2, 2) X(2, 3) X(3, 3) X(3, 4) X(4, 4) X(4, 5) X(5, 5) X(5, 6) X(6, 6) X(6, 7) \
    X(7, 7) X(7, 8) X(8, 8) X(8, 9) X(9, 9) X(9, 10)

}  // namespace

bool apply_supported(int p, int q1d) { return p >= 1 && p <= 8 && (q1d == p + 1 || q1d == p + 2); }

namespace {
// Selected diffusion variant per (D, Q); filled by the autotuner.
int g_diff_variant[11][11] = {};
}  // namespace

int diffusion_variant_count() { return kDiffVariants; }
void set_diffusion_variant(int p, int q1d, int v) {
    if (p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10) g_diff_variant[p + 1][q1d] = v;
}
int get_diffusion_variant(int p, int q1d) {
    return (p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10) ? g_diff_variant[p + 1][q1d] : 0;
}

void diffusion_apply(int p, int q1d, int64_t ne, const double *B, const double *G,
                     const double *pa, const int *gmap, const uint8_t *mask, const double *x,
                     double *ye, bool accumulate, hipStream_t s) {
    diffusion_apply_variant(get_diffusion_variant(p, q1d), p, q1d, ne, B, G, pa, gmap, mask, x, ye,
                            accumulate, s);
}

void diffusion_apply_variant(int variant, int p, int q1d, int64_t ne, const double *B,
                             const double *G, const double *pa, const int *gmap,
                             const uint8_t *mask, const double *x, double *ye, bool accumulate,
                             hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_diffusion<DD, QQ>(variant, ne, B, G, pa, gmap, mask, x, ye, accumulate, s); \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "diffusion apply: unsupported (p, q1d)"};
}

void mass_apply(int p, int q1d, int64_t ne, const double *B, const double *pa, const int *gmap,
                const uint8_t *mask, const double *x, double *ye, bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_mass<DD, QQ>(ne, B, pa, gmap, mask, x, ye, accumulate, s);                  \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "mass apply: unsupported (p, q1d)"};
}

void diffusion_diagonal(int p, int q1d, int64_t ne, const double *B, const double *G,
                        const double *pa, double *ye, bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_diff_diag<DD, QQ>(ne, B, G, pa, ye, accumulate, s);                         \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "diffusion diagonal: unsupported (p, q1d)"};
}

void mass_diagonal(int p, int q1d, int64_t ne, const double *B, const double *pa, double *ye,
                   bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_mass_diag<DD, QQ>(ne, B, pa, ye, accumulate, s);                            \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "mass diagonal: unsupported (p, q1d)"};
}

}  // namespace kern
}  // namespace h1pa
