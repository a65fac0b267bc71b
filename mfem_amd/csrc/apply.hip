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
//  * B and G enter the FMAs as scalar (SGPR) operands.  The contracted index
//    is unrolled, the output-line loop is not: each iteration needs one row
//    (or column) of B and G -- at most 2*10 doubles, loaded with wide scalar
//    loads from a contiguous copy -- so the SGPR file never spills;
//  * one LDS buffer of max(2 D^2 Q, 3 D Q^2) doubles per element, reused
//    across the stages with a barrier between the reads and the writes.
//
// basis layout (4*Q*D doubles): [ B(q,d) q-fastest | G q-fastest |
//                                 B d-fastest      | G d-fastest ].
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

// Store the E-DOF (dx, a, b) of element e.  With out.yL != nullptr the
// interior ("bubble") DOFs, which no other element touches, go straight into
// the L-vector at yL[bb + e (D-2)^3 + lex interior index] (MFEM's bubble
// block, fem/fespace.cpp:3518-3527); everything else goes to the E-vector, at
// its lexicographic position or, with out.l2n, at its native position.
template <int D>
__device__ __forceinline__ void store_out(const EOut &out, int64_t e, int a, int b, int dx,
                                          double v, int accumulate) {
    constexpr int P = D - 1, NI = D - 2;
    double *dst;
    if (out.yL && a >= 1 && a < P && b >= 1 && b < P && dx >= 1 && dx < P) {
        dst = out.yL + out.bb + e * (int64_t)(NI * NI * NI) + (dx - 1) + NI * ((a - 1) + NI * (b - 1));
    } else {
        const int lex = dx + D * (a + D * b);
        dst = out.ye + (int64_t)(D * D * D) * e + (out.l2n ? out.l2n[lex] : lex);
    }
    if (accumulate) *dst += v;
    else *dst = v;
}

// U: unroll factor of the output-line loops (1 or 2).
// K: PA-data pipeline depth.  0: each z-plane's 6 values are loaded in its
//    own iteration.  K > 0: the first K planes are loaded at kernel entry
//    (their latency hides behind the x/y stages) and plane qz + K is issued
//    while plane qz is computed, from a K x 6 register ring.
template <int D, int Q, int NB, int U, int K>
__global__ void __launch_bounds__(NB * Q * Q)
k_diffusion_apply(int64_t ne, const double *__restrict__ basis, const double *__restrict__ pa,
                  const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
                  const double *__restrict__ x, const EOut out, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DDQ = D * D * Q, DQQ = D * Q * Q;
    constexpr int BUF = (2 * DDQ > 3 * DQQ) ? 2 * DDQ : 3 * DQQ;
    __shared__ double smem[NB * BUF];
    const double *__restrict__ Bq = basis;              // B[q + Q d]
    const double *__restrict__ Gq = basis + Q * D;
    const double *__restrict__ Bd = basis + 2 * Q * D;  // B[d + D q]
    const double *__restrict__ Gd = basis + 3 * Q * D;
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = (int64_t)blockIdx.x * NB + le;
    const bool valid = e < ne;
    double *s = smem + le * BUF;
    constexpr int KK = K < Q ? K : Q;
    const double *pcol = pa + (int64_t)QQQ * 6 * (valid ? e : 0) + a + Q * b;
    double ring[KK > 0 ? KK : 1][6];
    if constexpr (KK > 0) {
        if (valid) {
#pragma unroll
            for (int j = 0; j < KK; ++j)
#pragma unroll
                for (int c = 0; c < 6; ++c) ring[j][c] = pcol[QQQ * c + QQ * j];
        }
    }

    // ---- forward x: thread (dy, dz) = (a, b)
    if (valid && a < D && b < D) {
        double u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) u[dx] = load_x(gmap, mask, x, base + dx);
        double *out = s + Q * (a + D * b);
#pragma unroll U
        for (int qx = 0; qx < Q; ++qx) {
            const double *br = Bd + D * qx, *gr = Gd + D * qx;
            double bu = 0.0, gu = 0.0;
#pragma unroll
            for (int dx = 0; dx < D; ++dx) {
                bu = fma(br[dx], u[dx], bu);
                gu = fma(gr[dx], u[dx], gu);
            }
            out[qx] = bu;
            out[DDQ + qx] = gu;
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
#pragma unroll U
            for (int qy = 0; qy < Q; ++qy) {
                const double *br = Bd + D * qy, *gr = Gd + D * qy;
                double bb = 0.0, gb = 0.0, bg = 0.0;
#pragma unroll
                for (int dy = 0; dy < D; ++dy) {
                    bb = fma(br[dy], r0[dy], bb);
                    gb = fma(gr[dy], r0[dy], gb);
                    bg = fma(br[dy], r1[dy], bg);
                }
                double *o = s + a + Q * (qy + Q * b);
                o[0 * DQQ] = bg;  // -> d/dx
                o[1 * DQQ] = gb;  // -> d/dy
                o[2 * DQQ] = bb;  // -> d/dz
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
                const double *i = s + a + Q * (b + Q * dz);
                z0[dz] = i[0 * DQQ];
                z1[dz] = i[1 * DQQ];
                z2[dz] = i[2 * DQQ];
            }
        }
        __syncthreads();
        if (valid && KK > 0) {
            double T0[D], T1[D], T2[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) T0[dz] = T1[dz] = T2[dz] = 0.0;
#pragma unroll 1
            for (int qb = 0; qb < Q; qb += (KK > 0 ? KK : 1)) {
#pragma unroll
                for (int j = 0; j < (KK > 0 ? KK : 1); ++j) {
                    const int qz = qb + j;
                    if (qz < Q) {
                        const double *br = Bd + D * qz, *gr = Gd + D * qz;
                        const double d11 = ring[j][0], d21 = ring[j][1], d31 = ring[j][2];
                        const double d22 = ring[j][3], d32 = ring[j][4], d33 = ring[j][5];
                        if (qz + KK < Q) {
#pragma unroll
                            for (int c = 0; c < 6; ++c)
                                ring[j][c] = pcol[QQQ * c + QQ * (qz + KK)];
                        }
                        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
                        for (int dz = 0; dz < D; ++dz) {
                            g0 = fma(br[dz], z0[dz], g0);
                            g1 = fma(br[dz], z1[dz], g1);
                            g2 = fma(gr[dz], z2[dz], g2);
                        }
                        const double O0 = fma(d11, g0, fma(d21, g1, d31 * g2));
                        const double O1 = fma(d21, g0, fma(d22, g1, d32 * g2));
                        const double O2 = fma(d31, g0, fma(d32, g1, d33 * g2));
#pragma unroll
                        for (int dz = 0; dz < D; ++dz) {
                            T0[dz] = fma(br[dz], O0, T0[dz]);
                            T1[dz] = fma(br[dz], O1, T1[dz]);
                            T2[dz] = fma(gr[dz], O2, T2[dz]);
                        }
                    }
                }
            }
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                double *o = s + a + Q * (b + Q * dz);
                o[0 * DQQ] = T0[dz];
                o[1 * DQQ] = T1[dz];
                o[2 * DQQ] = T2[dz];
            }
        } else if (valid) {
            double T0[D], T1[D], T2[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) T0[dz] = T1[dz] = T2[dz] = 0.0;
#pragma unroll U
            for (int qz = 0; qz < Q; ++qz) {
                const double *br = Bd + D * qz, *gr = Gd + D * qz;
                const double *pq = pcol + QQ * qz;
                const double d11 = pq[0 * QQQ], d21 = pq[1 * QQQ], d31 = pq[2 * QQQ];
                const double d22 = pq[3 * QQQ], d32 = pq[4 * QQQ], d33 = pq[5 * QQQ];
                double g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) {
                    g0 = fma(br[dz], z0[dz], g0);
                    g1 = fma(br[dz], z1[dz], g1);
                    g2 = fma(gr[dz], z2[dz], g2);
                }
                const double O0 = fma(d11, g0, fma(d21, g1, d31 * g2));
                const double O1 = fma(d21, g0, fma(d22, g1, d32 * g2));
                const double O2 = fma(d31, g0, fma(d32, g1, d33 * g2));
#pragma unroll
                for (int dz = 0; dz < D; ++dz) {
                    T0[dz] = fma(br[dz], O0, T0[dz]);
                    T1[dz] = fma(br[dz], O1, T1[dz]);
                    T2[dz] = fma(gr[dz], O2, T2[dz]);
                }
            }
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                double *o = s + a + Q * (b + Q * dz);
                o[0 * DQQ] = T0[dz];
                o[1 * DQQ] = T1[dz];
                o[2 * DQQ] = T2[dz];
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
                const double *i = s + a + Q * (qy + Q * b);
                t0[qy] = i[0 * DQQ];
                t1[qy] = i[1 * DQQ];
                t2[qy] = i[2 * DQQ];
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll U
            for (int dy = 0; dy < D; ++dy) {
                const double *bc = Bq + Q * dy, *gc = Gq + Q * dy;
                double sa = 0.0, sb = 0.0;
#pragma unroll
                for (int qy = 0; qy < Q; ++qy) {
                    sa = fma(bc[qy], t0[qy], sa);
                    sb = fma(gc[qy], t1[qy], sb);
                    sb = fma(bc[qy], t2[qy], sb);
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
        const double *in = s + Q * (a + D * b);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
            sa[qx] = in[qx];
            sb[qx] = in[DDQ + qx];
        }
#pragma unroll U
        for (int dx = 0; dx < D; ++dx) {
            const double *bc = Bq + Q * dx, *gc = Gq + Q * dx;
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) {
                v = fma(gc[qx], sa[qx], v);
                v = fma(bc[qx], sb[qx], v);
            }
            store_out<D>(out, e, a, b, dx, v, accumulate);
        }
    }
}

template <int D, int Q, int NB>
__global__ void __launch_bounds__(NB * Q * Q)
k_mass_apply(int64_t ne, const double *__restrict__ basis, const double *__restrict__ pa,
             const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
             const double *__restrict__ x, const EOut out, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DQQ = D * Q * Q;
    constexpr int BUF = DQQ;
    __shared__ double smem[NB * BUF];
    const double *__restrict__ Bq = basis;
    const double *__restrict__ Bd = basis + 2 * Q * D;
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = (int64_t)blockIdx.x * NB + le;
    const bool valid = e < ne;
    double *s = smem + le * BUF;
    if (valid && a < D && b < D) {
        double u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) u[dx] = load_x(gmap, mask, x, base + dx);
#pragma unroll 1
        for (int qx = 0; qx < Q; ++qx) {
            const double *br = Bd + D * qx;
            double v = 0.0;
#pragma unroll
            for (int dx = 0; dx < D; ++dx) v = fma(br[dx], u[dx], v);
            s[qx + Q * (a + D * b)] = v;
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
#pragma unroll 1
            for (int qy = 0; qy < Q; ++qy) {
                const double *br = Bd + D * qy;
                double v = 0.0;
#pragma unroll
                for (int dy = 0; dy < D; ++dy) v = fma(br[dy], r[dy], v);
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
            const double *pcol = pa + (int64_t)QQQ * e + a + Q * b;
            double T[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) T[dz] = 0.0;
#pragma unroll 1
            for (int qz = 0; qz < Q; ++qz) {
                const double *br = Bd + D * qz;
                double v = 0.0;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) v = fma(br[dz], z[dz], v);
                v *= pcol[QQ * qz];
#pragma unroll
                for (int dz = 0; dz < D; ++dz) T[dz] = fma(br[dz], v, T[dz]);
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
#pragma unroll 1
            for (int dy = 0; dy < D; ++dy) {
                const double *bc = Bq + Q * dy;
                double v = 0.0;
#pragma unroll
                for (int qy = 0; qy < Q; ++qy) v = fma(bc[qy], tq[qy], v);
                s[a + Q * (dy + D * b)] = v;
            }
    }
    __syncthreads();
    if (valid && a < D && b < D) {
        double sx[Q];
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) sx[qx] = s[qx + Q * (a + D * b)];
#pragma unroll 1
        for (int dx = 0; dx < D; ++dx) {
            const double *bc = Bq + Q * dx;
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v = fma(bc[qx], sx[qx], v);
            store_out<D>(out, e, a, b, dx, v, accumulate);
        }
    }
}

// Diagonal of the element diffusion matrices (PADiffusionDiagonal3D,
// bilininteg_diffusion_kernels.hpp:268-367), all nine (i,j) pairs grouped by
// the 1D factor in x: GG (O11), BB (O22, O33, 2 O23), GB (2 O12, 2 O13).
template <int D, int Q>
__global__ void __launch_bounds__(Q * Q)
k_diffusion_diag(int64_t ne, const double *__restrict__ basis, const double *__restrict__ pa,
                 double *__restrict__ ye, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DQQ = D * Q * Q, DDQ = D * D * Q;
    __shared__ double s[6 * DQQ];
    const double *__restrict__ Bq = basis;
    const double *__restrict__ Gq = basis + Q * D;
    const double *__restrict__ Bd = basis + 2 * Q * D;
    const double *__restrict__ Gd = basis + 3 * Q * D;
    const int a = threadIdx.x % Q, b = threadIdx.x / Q;
    const int64_t e = blockIdx.x;
    {   // z: thread (qx, qy)
        const double *p = pa + (int64_t)QQQ * 6 * e + a + Q * b;
        double acc[6][D];
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int dz = 0; dz < D; ++dz) acc[k][dz] = 0.0;
#pragma unroll 1
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
    const bool acty = b < D;
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
    if (a < D && b < D) {   // x: thread (dy, dz)
        double *yp = ye + (int64_t)DDD * e + D * (a + D * b);
#pragma unroll 1
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
            if (accumulate) yp[dx] += v;
            else yp[dx] = v;
        }
    }
}

template <int D, int Q>
__global__ void __launch_bounds__(Q * Q)
k_mass_diag(int64_t ne, const double *__restrict__ basis, const double *__restrict__ pa,
            double *__restrict__ ye, int accumulate) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DQQ = D * Q * Q;
    __shared__ double s[DQQ];
    const double *__restrict__ Bq = basis;
    const double *__restrict__ Bd = basis + 2 * Q * D;
    const int a = threadIdx.x % Q, b = threadIdx.x / Q;
    const int64_t e = blockIdx.x;
    {
        const double *p = pa + (int64_t)QQQ * e + a + Q * b;
        double acc[D];
#pragma unroll
        for (int dz = 0; dz < D; ++dz) acc[dz] = 0.0;
#pragma unroll 1
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
    const bool acty = b < D;
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
    if (a < D && b < D) {
        double *yp = ye + (int64_t)DDD * e + D * (a + D * b);
#pragma unroll 1
        for (int dx = 0; dx < D; ++dx) {
            const double *bc = Bq + Q * dx;
            double v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v = fma(bc[qx] * bc[qx], s[qx + Q * (a + D * b)], v);
            if (accumulate) yp[dx] += v;
            else yp[dx] = v;
        }
    }
}

// ---- dispatch over (D, Q) and variants -------------------------------------
template <int D, int Q, int NB, int U, int K>
void launch_diffusion_v(int64_t ne, const double *basis, const double *pa, const int *gmap,
                        const uint8_t *mask, const double *x, const EOut &out, bool acc,
                        hipStream_t s) {
    const unsigned grid = (unsigned)((ne + NB - 1) / NB);
    hipLaunchKernelGGL((k_diffusion_apply<D, Q, NB, U, K>), dim3(grid), dim3(NB * Q * Q), 0, s,
                       ne, basis, pa, gmap, mask, x, out, acc ? 1 : 0);
}

// Variants of the diffusion apply: (elements per block, line-loop unroll,
// PA-data pipeline depth).  Every variant performs the same arithmetic per element
// in the same order, so they give bit-identical results and differ only in
// speed.
constexpr int kDiffVariants = 6;
template <int D, int Q>
void launch_diffusion(int variant, int64_t ne, const double *basis, const double *pa,
                      const int *gmap, const uint8_t *mask, const double *x, const EOut &out,
                      bool acc, hipStream_t s) {
    constexpr int NBmax = nb_for(Q, 3 * D * Q * Q);
    constexpr int NBhalf = NBmax / 2 > 0 ? NBmax / 2 : 1;
    switch (variant) {
        case 1: launch_diffusion_v<D, Q, NBmax, 2, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 2: launch_diffusion_v<D, Q, NBhalf, 1, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 3: launch_diffusion_v<D, Q, NBhalf, 2, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 4: launch_diffusion_v<D, Q, NBmax, 1, 2>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        case 5: launch_diffusion_v<D, Q, NBmax, 1, 3>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
        default: launch_diffusion_v<D, Q, NBmax, 1, 0>(ne, basis, pa, gmap, mask, x, out, acc, s); break;
    }
}
template <int D, int Q>
void launch_mass(int64_t ne, const double *basis, const double *pa, const int *gmap,
                 const uint8_t *mask, const double *x, const EOut &out, bool acc,
                 hipStream_t s) {
    constexpr int NB = nb_for(Q, D * Q * Q);
    const unsigned grid = (unsigned)((ne + NB - 1) / NB);
    hipLaunchKernelGGL((k_mass_apply<D, Q, NB>), dim3(grid), dim3(NB * Q * Q), 0, s, ne, basis, pa,
                       gmap, mask, x, out, acc ? 1 : 0);
}
template <int D, int Q>
void launch_diff_diag(int64_t ne, const double *basis, const double *pa, double *ye, bool acc,
                      hipStream_t s) {
    hipLaunchKernelGGL((k_diffusion_diag<D, Q>), dim3((unsigned)ne), dim3(Q * Q), 0, s, ne, basis,
                       pa, ye, acc ? 1 : 0);
}
template <int D, int Q>
void launch_mass_diag(int64_t ne, const double *basis, const double *pa, double *ye, bool acc,
                      hipStream_t s) {
    hipLaunchKernelGGL((k_mass_diag<D, Q>), dim3((unsigned)ne), dim3(Q * Q), 0, s, ne, basis, pa,
                       ye, acc ? 1 : 0);
}

// (D, Q) pairs compiled: Q = D (q = p+1) and Q = D + 1 (q = p+2, MFEM default).
#define H1PA_FOR_EACH_DQ(X) \
    X(2, 2) X(2, 3) X(3, 3) X(3, 4) X(4, 4) X(4, 5) X(5, 5) X(5, 6) X(6, 6) X(6, 7) \
    X(7, 7) X(7, 8) X(8, 8) X(8, 9) X(9, 9) X(9, 10)

// Selected diffusion variant per (D, Q); filled by the autotuner.
int g_diff_variant[11][11] = {};

}  // namespace

bool apply_supported(int p, int q1d) { return p >= 1 && p <= 8 && (q1d == p + 1 || q1d == p + 2); }

int diffusion_variant_count() { return kDiffVariants; }
void set_diffusion_variant(int p, int q1d, int v) {
    if (p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10) g_diff_variant[p + 1][q1d] = v;
}
int get_diffusion_variant(int p, int q1d) {
    return (p >= 1 && p <= 8 && q1d >= 1 && q1d <= 10) ? g_diff_variant[p + 1][q1d] : 0;
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
        launch_mass<DD, QQ>(ne, basis, pa, gmap, mask, x, out, accumulate, s);             \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "mass apply: unsupported (p, q1d)"};
}

void diffusion_diagonal(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                        double *ye, bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_diff_diag<DD, QQ>(ne, basis, pa, ye, accumulate, s);                        \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "diffusion diagonal: unsupported (p, q1d)"};
}

void mass_diagonal(int p, int q1d, int64_t ne, const double *basis, const double *pa, double *ye,
                   bool accumulate, hipStream_t s) {
    if (ne <= 0) return;
    const int D = p + 1;
#define X(DD, QQ)                                                                          \
    if (D == DD && q1d == QQ) {                                                            \
        launch_mass_diag<DD, QQ>(ne, basis, pa, ye, accumulate, s);                        \
        H1PA_HIP(hipGetLastError());                                                       \
        return;                                                                            \
    }
    H1PA_FOR_EACH_DQ(X)
#undef X
    throw Error{H1PA_EUNSUP, "mass diagonal: unsupported (p, q1d)"};
}

}  // namespace kern
}  // namespace h1pa
