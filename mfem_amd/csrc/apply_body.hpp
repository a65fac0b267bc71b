// apply_body.hpp -- device bodies of the sum-factorized PA applies, shared by
// the single-problem kernels (apply.hip) and the batched kernels (batch.hip).
// See apply.hip for the math and the CDNA4 mapping.
#pragma once

#include "internal.hpp"

namespace h1pa {
namespace kern {

constexpr int nb_for(int Q, int buf_doubles) {
    int nb = 256 / (Q * Q);
    if (nb < 1) nb = 1;
    while (nb > 1 && nb * buf_doubles * 8 > 48 * 1024) --nb;
    return nb;
}

// Work-groups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8), and each XCD has its own L2.  xcd_block(b, G) renumbers the blocks
// so that every XCD processes one contiguous range of the grid: elements that
// are neighbours in the (space-filling-curve) element order, and so share
// DOFs, run on the same XCD and find each other's x lines in its L2.  The map
// is a bijection of [0, G).
__host__ __device__ inline int64_t xcd_block(int64_t b, int64_t G) {
    constexpr int64_t X = 8;
    const int64_t q = G / X, r = G % X, k = b % X, j = b / X;
    return k * q + (k < r ? k : r) + j;
}

// PA-data element -> compute type (bf16 is stored as its upper 16 bits).
template <typename T, typename TP>
__device__ __forceinline__ T to_t(TP v) { return static_cast<T>(v); }
template <>
__device__ __forceinline__ float to_t<float, bf16_t>(bf16_t v) {
    return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
}

// Read x at E-DOF `i` through the gather map (L-vector input) or directly
// (E-vector input); constrained L-DOFs read as zero.
template <typename T>
__device__ __forceinline__ T load_x(const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
                                    const T *__restrict__ x, int64_t i) {
    if (!gmap) return x[i];
    const int g = gmap[i];
    const int j = g >= 0 ? g : -1 - g;
    if (mask && mask[j]) return T(0);
    const T v = x[j];
    return g >= 0 ? v : -v;
}

// Store the E-DOF (dx, a, b) of element e.  With out.yL != nullptr the
// interior ("bubble") DOFs, which no other element touches, go straight into
// the L-vector at yL[bb + e (D-2)^3 + lex interior index] (MFEM's bubble
// block, fem/fespace.cpp:3518-3527); everything else goes to the E-vector, at
// its lexicographic position or, with out.l2n, at its native position.
template <int D, typename T>
__device__ __forceinline__ void store_out(const EOutT<T> &out, int64_t e, int a, int b, int dx,
                                          T v, int accumulate) {
    constexpr int P = D - 1, NI = D - 2;
    T *dst;
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
// LDS doubles (or floats) per element of the diffusion body.
template <int D, int Q>
constexpr int diffusion_buf() {
    return (2 * D * D * Q > 3 * D * Q * Q) ? 2 * D * D * Q : 3 * D * Q * Q;
}

// The elements e0 + (threadIdx.x / Q^2), NB per work-group, of one problem.
template <typename T, typename TP, int D, int Q, int NB, int U, int K>
__device__ __forceinline__ void
diffusion_elements(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
                   const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
                   const T *__restrict__ x, const EOutT<T> &out, int accumulate,
                   T *__restrict__ smem) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DDQ = D * D * Q, DQQ = D * Q * Q;
    constexpr int BUF = diffusion_buf<D, Q>();
    const T *__restrict__ Bq = basis;              // B[q + Q d]
    const T *__restrict__ Gq = basis + Q * D;
    const T *__restrict__ Bd = basis + 2 * Q * D;  // B[d + D q]
    const T *__restrict__ Gd = basis + 3 * Q * D;
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = e0 + le;
    const bool valid = e < ne;
    T *s = smem + le * BUF;
    constexpr int KK = K < Q ? K : Q;
    const TP *pcol = pa + (int64_t)QQQ * 6 * (valid ? e : 0) + a + Q * b;
    T ring[KK > 0 ? KK : 1][6];
    if constexpr (KK > 0) {
        if (valid) {
#pragma unroll
            for (int j = 0; j < KK; ++j)
#pragma unroll
                for (int c = 0; c < 6; ++c) ring[j][c] = to_t<T>(pcol[QQQ * c + QQ * j]);
        }
    }

    // ---- forward x: thread (dy, dz) = (a, b)
    if (valid && a < D && b < D) {
        T u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) u[dx] = load_x(gmap, mask, x, base + dx);
        T *out = s + Q * (a + D * b);
#pragma unroll U
        for (int qx = 0; qx < Q; ++qx) {
            const T *br = Bd + D * qx, *gr = Gd + D * qx;
            T bu = 0.0, gu = 0.0;
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
        T r0[D], r1[D];
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
                const T *br = Bd + D * qy, *gr = Gd + D * qy;
                T bb = 0.0, gb = 0.0, bg = 0.0;
#pragma unroll
                for (int dy = 0; dy < D; ++dy) {
                    bb = fma(br[dy], r0[dy], bb);
                    gb = fma(gr[dy], r0[dy], gb);
                    bg = fma(br[dy], r1[dy], bg);
                }
                T *o = s + a + Q * (qy + Q * b);
                o[0 * DQQ] = bg;  // -> d/dx
                o[1 * DQQ] = gb;  // -> d/dy
                o[2 * DQQ] = bb;  // -> d/dz
            }
        }
    }
    __syncthreads();

    // ---- z in registers: thread (qx, qy) = (a, b)
    {
        T z0[D], z1[D], z2[D];
        if (valid) {
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                const T *i = s + a + Q * (b + Q * dz);
                z0[dz] = i[0 * DQQ];
                z1[dz] = i[1 * DQQ];
                z2[dz] = i[2 * DQQ];
            }
        }
        __syncthreads();
        if (valid && KK > 0) {
            T T0[D], T1[D], T2[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) T0[dz] = T1[dz] = T2[dz] = 0.0;
#pragma unroll 1
            for (int qb = 0; qb < Q; qb += (KK > 0 ? KK : 1)) {
#pragma unroll
                for (int j = 0; j < (KK > 0 ? KK : 1); ++j) {
                    const int qz = qb + j;
                    if (qz < Q) {
                        const T *br = Bd + D * qz, *gr = Gd + D * qz;
                        const T d11 = ring[j][0], d21 = ring[j][1], d31 = ring[j][2];
                        const T d22 = ring[j][3], d32 = ring[j][4], d33 = ring[j][5];
                        if (qz + KK < Q) {
#pragma unroll
                            for (int c = 0; c < 6; ++c)
                                ring[j][c] = to_t<T>(pcol[QQQ * c + QQ * (qz + KK)]);
                        }
                        T g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
                        for (int dz = 0; dz < D; ++dz) {
                            g0 = fma(br[dz], z0[dz], g0);
                            g1 = fma(br[dz], z1[dz], g1);
                            g2 = fma(gr[dz], z2[dz], g2);
                        }
                        const T O0 = fma(d11, g0, fma(d21, g1, d31 * g2));
                        const T O1 = fma(d21, g0, fma(d22, g1, d32 * g2));
                        const T O2 = fma(d31, g0, fma(d32, g1, d33 * g2));
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
                T *o = s + a + Q * (b + Q * dz);
                o[0 * DQQ] = T0[dz];
                o[1 * DQQ] = T1[dz];
                o[2 * DQQ] = T2[dz];
            }
        } else if (valid) {
            T T0[D], T1[D], T2[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) T0[dz] = T1[dz] = T2[dz] = 0.0;
#pragma unroll U
            for (int qz = 0; qz < Q; ++qz) {
                const T *br = Bd + D * qz, *gr = Gd + D * qz;
                const TP *pq = pcol + QQ * qz;
                const T d11 = to_t<T>(pq[0 * QQQ]), d21 = to_t<T>(pq[1 * QQQ]), d31 = to_t<T>(pq[2 * QQQ]);
                const T d22 = to_t<T>(pq[3 * QQQ]), d32 = to_t<T>(pq[4 * QQQ]), d33 = to_t<T>(pq[5 * QQQ]);
                T g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) {
                    g0 = fma(br[dz], z0[dz], g0);
                    g1 = fma(br[dz], z1[dz], g1);
                    g2 = fma(gr[dz], z2[dz], g2);
                }
                const T O0 = fma(d11, g0, fma(d21, g1, d31 * g2));
                const T O1 = fma(d21, g0, fma(d22, g1, d32 * g2));
                const T O2 = fma(d31, g0, fma(d32, g1, d33 * g2));
#pragma unroll
                for (int dz = 0; dz < D; ++dz) {
                    T0[dz] = fma(br[dz], O0, T0[dz]);
                    T1[dz] = fma(br[dz], O1, T1[dz]);
                    T2[dz] = fma(gr[dz], O2, T2[dz]);
                }
            }
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                T *o = s + a + Q * (b + Q * dz);
                o[0 * DQQ] = T0[dz];
                o[1 * DQQ] = T1[dz];
                o[2 * DQQ] = T2[dz];
            }
        }
    }
    __syncthreads();

    // ---- backward y: thread (qx, dz) = (a, b)
    {
        T t0[Q], t1[Q], t2[Q];
        const bool act = valid && b < D;
        if (act) {
#pragma unroll
            for (int qy = 0; qy < Q; ++qy) {
                const T *i = s + a + Q * (qy + Q * b);
                t0[qy] = i[0 * DQQ];
                t1[qy] = i[1 * DQQ];
                t2[qy] = i[2 * DQQ];
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll U
            for (int dy = 0; dy < D; ++dy) {
                const T *bc = Bq + Q * dy, *gc = Gq + Q * dy;
                T sa = 0.0, sb = 0.0;
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
        T sa[Q], sb[Q];
        const T *in = s + Q * (a + D * b);
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) {
            sa[qx] = in[qx];
            sb[qx] = in[DDQ + qx];
        }
#pragma unroll U
        for (int dx = 0; dx < D; ++dx) {
            const T *bc = Bq + Q * dx, *gc = Gq + Q * dx;
            T v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) {
                v = fma(gc[qx], sa[qx], v);
                v = fma(bc[qx], sb[qx], v);
            }
            store_out<D, T>(out, e, a, b, dx, v, accumulate);
        }
    }
}

template <typename T, typename TP, int D, int Q, int NB>
__device__ __forceinline__ void
mass_elements(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
              const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
              const T *__restrict__ x, const EOutT<T> &out, int accumulate,
              T *__restrict__ smem) {
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DQQ = D * Q * Q;
    constexpr int BUF = DQQ;
    const T *__restrict__ Bq = basis;
    const T *__restrict__ Bd = basis + 2 * Q * D;
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = e0 + le;
    const bool valid = e < ne;
    T *s = smem + le * BUF;
    if (valid && a < D && b < D) {
        T u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
#pragma unroll
        for (int dx = 0; dx < D; ++dx) u[dx] = load_x(gmap, mask, x, base + dx);
#pragma unroll 1
        for (int qx = 0; qx < Q; ++qx) {
            const T *br = Bd + D * qx;
            T v = 0.0;
#pragma unroll
            for (int dx = 0; dx < D; ++dx) v = fma(br[dx], u[dx], v);
            s[qx + Q * (a + D * b)] = v;
        }
    }
    __syncthreads();
    {
        T r[D];
        const bool act = valid && b < D;
        if (act)
#pragma unroll
            for (int dy = 0; dy < D; ++dy) r[dy] = s[a + Q * (dy + D * b)];
        __syncthreads();
        if (act)
#pragma unroll 1
            for (int qy = 0; qy < Q; ++qy) {
                const T *br = Bd + D * qy;
                T v = 0.0;
#pragma unroll
                for (int dy = 0; dy < D; ++dy) v = fma(br[dy], r[dy], v);
                s[a + Q * (qy + Q * b)] = v;
            }
    }
    __syncthreads();
    {
        T z[D];
        if (valid)
#pragma unroll
            for (int dz = 0; dz < D; ++dz) z[dz] = s[a + Q * (b + Q * dz)];
        __syncthreads();
        if (valid) {
            const TP *pcol = pa + (int64_t)QQQ * e + a + Q * b;
            T tz[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) tz[dz] = 0.0;
#pragma unroll 1
            for (int qz = 0; qz < Q; ++qz) {
                const T *br = Bd + D * qz;
                T v = 0.0;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) v = fma(br[dz], z[dz], v);
                v *= to_t<T>(pcol[QQ * qz]);
#pragma unroll
                for (int dz = 0; dz < D; ++dz) tz[dz] = fma(br[dz], v, tz[dz]);
            }
#pragma unroll
            for (int dz = 0; dz < D; ++dz) s[a + Q * (b + Q * dz)] = tz[dz];
        }
    }
    __syncthreads();
    {
        T tq[Q];
        const bool act = valid && b < D;
        if (act)
#pragma unroll
            for (int qy = 0; qy < Q; ++qy) tq[qy] = s[a + Q * (qy + Q * b)];
        __syncthreads();
        if (act)
#pragma unroll 1
            for (int dy = 0; dy < D; ++dy) {
                const T *bc = Bq + Q * dy;
                T v = 0.0;
#pragma unroll
                for (int qy = 0; qy < Q; ++qy) v = fma(bc[qy], tq[qy], v);
                s[a + Q * (dy + D * b)] = v;
            }
    }
    __syncthreads();
    if (valid && a < D && b < D) {
        T sx[Q];
#pragma unroll
        for (int qx = 0; qx < Q; ++qx) sx[qx] = s[qx + Q * (a + D * b)];
#pragma unroll 1
        for (int dx = 0; dx < D; ++dx) {
            const T *bc = Bq + Q * dx;
            T v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v = fma(bc[qx], sx[qx], v);
            store_out<D, T>(out, e, a, b, dx, v, accumulate);
        }
    }
}

}  // namespace kern
}  // namespace h1pa
