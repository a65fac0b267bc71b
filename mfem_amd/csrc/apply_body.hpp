// apply_body.hpp -- device bodies of the sum-factorized PA applies, shared by
// the single-problem kernels (apply.hip) and the batched kernels (batch.hip).
// See apply.hip for the math and the CDNA4 mapping.
#pragma once

#include "internal.hpp"

namespace h1pa {
namespace kern {

// A pointer read from a device-side descriptor, marked as pointing to global
// memory.  The compiler cannot tell where such a pointer points and emits
// flat loads and stores for it, which also count against the LDS wait
// counter: every LDS wait of an element body would then wait for the PA
// prefetch too.  The empty asm hides the integer from the optimizer, so the
// cast to the global address space is not folded away.
template <typename T>
__device__ __forceinline__ T *global_ptr(T *p) {
    uint64_t u = reinterpret_cast<uint64_t>(p);
    asm("" : "+s"(u));
    return (T *)(__attribute__((address_space(1))) T *)u;
}

// PA-data pipeline depth (template parameter K of the element bodies) that
// loads every plane of the element before the forward contractions start.
constexpr int kAllPlanes = 16;

// Fraction of the lanes of its 64-wide waves that nb elements of Q^2 threads use.
constexpr double lane_use(int nb, int QQ) {
    const int t = nb * QQ;
    return (double)t / (64.0 * ((t + 63) / 64));
}

// Elements per work-group for an element body of `buf` LDS values of `bytes`
// each: about 256 threads within 48 KiB of LDS; where that leaves more than
// a tenth of the lanes idle (Q = 9, 10: 81 or 100 threads per element) and
// the values are 4-byte, the count with the best lane use within 1024 threads
// and 64 KiB of LDS.  FP64 bodies keep the 256-thread count: their larger LDS
// footprint per element makes the wider blocks cost more resident waves than
// the idle lanes do (p = 8 FP64 apply measured 23.2k -> 17.9k MDOF/s with 3
// instead of 2 elements per block).
constexpr int nb_for(int Q, int buf, int bytes) {
    const int QQ = Q * Q;
    int nb = 256 / QQ;
    if (nb < 1) nb = 1;
    while (nb > 1 && nb * buf * bytes > 48 * 1024) --nb;
    const double u0 = lane_use(nb, QQ);
    if (bytes == 4 && u0 < 0.9) {
        int best = nb;
        double bu = u0;
        for (int n = nb + 1; n * QQ <= 1024 && n * buf * bytes <= 64 * 1024; ++n)
            if (lane_use(n, QQ) > bu + 1e-9) {
                best = n;
                bu = lane_use(n, QQ);
            }
        if (bu > u0 + 0.05) nb = best;
    }
    return nb;
}

// Load one PA value as the compute type (bf16 is stored as the upper 16 bits
// of an FP32).  PA data is streamed once per apply, so with NT it is loaded
// non-temporally and does not push the gathered x-values (re-read by up to 8
// neighbouring elements) out of the L2.
template <typename T, bool NT, typename TP>
__device__ __forceinline__ T load_pa(const TP *p) {
    if constexpr (NT) return static_cast<T>(__builtin_nontemporal_load(p));
    else return static_cast<T>(*p);
}
template <typename T, bool NT>
__device__ __forceinline__ float load_pa(const bf16_t *p) {
    const uint16_t *u = reinterpret_cast<const uint16_t *>(p);
    uint16_t b;
    if constexpr (NT) b = __builtin_nontemporal_load(u);
    else b = *u;
    return __uint_as_float(static_cast<uint32_t>(b) << 16);
}

// Whether the PA planes of a Q^3 element are streamed non-temporally.  A
// plane of one component is Q^2 values read by one load per thread; when it
// does not fill whole 128-byte cache lines, neighbouring planes share lines,
// and a non-temporal load makes the second plane fetch the shared line
// again.  Measured (MI355X): non-temporal is 10-15% faster from Q = 8 on and
// 10-15% slower at Q = 6 (288-byte FP64 planes).
template <int Q, typename TP>
constexpr bool stream_pa() { return Q >= 8 || (Q * Q * sizeof(TP)) % 128 == 0; }

// Read the D consecutive E-DOFs base .. base+D-1 (one x-line of an element)
// through the gather map (L-vector input) or directly (E-vector input);
// constrained L-DOFs read as zero.  All index loads are issued first, then all
// value and mask loads, with no branch between them: a line costs two memory
// round trips, not D dependent (index, value) pairs.
// SAVE: the line's D gather indices are also stored to gsave (LDS).
template <int D, typename T, bool SAVE = false>
__device__ __forceinline__ void load_line(const int *__restrict__ gmap,
                                          const uint8_t *__restrict__ mask,
                                          const T *__restrict__ x, int64_t base, T (&u)[D],
                                          int *__restrict__ gsave = nullptr) {
    if (!gmap) {
#pragma unroll
        for (int d = 0; d < D; ++d) u[d] = x[base + d];
        return;
    }
    int g[D];
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = gmap[base + d];
    if constexpr (SAVE) {
#pragma unroll
        for (int d = 0; d < D; ++d) gsave[d] = g[d];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = x[g[d] >= 0 ? g[d] : -1 - g[d]];
    if (mask) {
        uint8_t m[D];
#pragma unroll
        for (int d = 0; d < D; ++d) m[d] = mask[g[d] >= 0 ? g[d] : -1 - g[d]];
#pragma unroll
        for (int d = 0; d < D; ++d) u[d] = m[d] ? T(0) : (g[d] >= 0 ? u[d] : -u[d]);
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) u[d] = g[d] >= 0 ? u[d] : -u[d];
    }
}

// Store the D E-DOFs (0..D-1, a, b) of element e held in y.  With
// out.yL != nullptr the interior ("bubble") DOFs, which no other element
// touches, go straight into the L-vector at yL[bb + e (D-2)^3 + lex interior
// index] (MFEM's bubble block, fem/fespace.cpp:3518-3527); everything else
// goes to the E-vector, at its lexicographic position or, with out.l2n, at
// its native position.  The native positions and, when accumulating, the old
// values are loaded as one batch before the first store, so the stores of a
// line wait on no memory round trip of their own.
template <int D, typename T>
__device__ __forceinline__ void store_line(const EOutT<T> &out, int64_t e, int a, int b,
                                           const T (&y)[D], int accumulate) {
    constexpr int P = D - 1, NI = D - 2;
    const int lex0 = D * (a + D * b);
    int pos[D];
    if (out.l2n) {
#pragma unroll
        for (int dx = 0; dx < D; ++dx) pos[dx] = out.l2n[lex0 + dx];
    } else {
#pragma unroll
        for (int dx = 0; dx < D; ++dx) pos[dx] = lex0 + dx;
    }
    const bool inner = out.yL && a >= 1 && a < P && b >= 1 && b < P;
    T *const ye = out.ye + (int64_t)(D * D * D) * e;
    T *const yb = inner ? out.yL + out.bb + e * (int64_t)(NI * NI * NI) + NI * ((a - 1) + NI * (b - 1)) - 1
                        : nullptr;
    T *dst[D];
#pragma unroll
    for (int dx = 0; dx < D; ++dx) dst[dx] = (inner && dx >= 1 && dx < P) ? yb + dx : ye + pos[dx];
    if (accumulate) {
        T old[D];
#pragma unroll
        for (int dx = 0; dx < D; ++dx) old[dx] = *dst[dx];
#pragma unroll
        for (int dx = 0; dx < D; ++dx) *dst[dx] = old[dx] + y[dx];
    } else {
#pragma unroll
        for (int dx = 0; dx < D; ++dx) *dst[dx] = y[dx];
    }
}

// Even-odd form of the 1D operators.  B (GLL Lagrange basis at the GL
// points) is centro-symmetric, B(Q-1-q, D-1-d) = B(q, d), and G = B' is
// centro-antisymmetric, G(Q-1-q, D-1-d) = -G(q, d).  Splitting a line u into
// e_d = u_d + u_{D-1-d}, o_d = u_d - u_{D-1-d} (d < D/2) and the middle value
// m (D odd), and a row q of M into Me = (M(q,d) + M(q,D-1-d))/2,
// Mo = (M(q,d) - M(q,D-1-d))/2, Mm = M(q, D/2):
//   A = sum Me e + Mm m,  C = sum Mo o,
//   (M u)_q = A + C,  (M u)_{Q-1-q} = s (A - C)           (s = +1 B, -1 G)
// and for the transpose, with P_q = v_q + s v_{Q-1-q}, N_q = v_q - s v_{Q-1-q}
// (q < Q/2; both equal v_{Q/2} at the middle q of an odd Q):
//   E_d = sum_q Me(q,d) P_q,  F_d = sum_q Mo(q,d) N_q,
//   (M^T v)_d = E_d + F_d,  (M^T v)_{D-1-d} = E_d - F_d,  (M^T v)_{D/2} = sum_q Mm(q) P_q.
// Each 1D contraction then costs about half the multiply-adds.  The tables
// follow the four plain blocks of the basis array (eo_tables below).
template <int D, int Q>
struct EO {
    static constexpr int HD = D / 2, OD = D % 2, HQ = Q / 2, OQ = Q % 2;
    static constexpr int R = HQ + OQ;        // row pairs (+ middle row)
    static constexpr int W = 2 * HD + OD;    // forward row: Me | Mo | Mm
    static constexpr int RB = 2 * R;         // backward row: Me column | Mo column
    static constexpr int FB = 4 * Q * D;     // offsets in the basis array
    static constexpr int FG = FB + R * W;
    static constexpr int BB = FG + R * W;
    static constexpr int BG = BB + (HD + OD) * RB;
    static constexpr int SIZE = BG + (HD + OD) * RB;  // total basis length
};

// U: unroll factor of the output-line loops (1 or 2).
// K: PA-data pipeline depth in z-planes.  0: the 12 values of a plane pair
//    are loaded in its own iteration.  K > 0: (K+1)/2 plane pairs are loaded
//    at kernel entry (their latency hides behind the x/y stages) and pair
//    j + (K+1)/2 is issued while pair j is computed.
// LDS doubles (or floats) per element of the diffusion body.
template <int D, int Q>
constexpr int diffusion_buf() {
    return (2 * D * D * Q > 3 * D * Q * Q) ? 2 * D * D * Q : 3 * D * Q * Q;
}

// Waves per SIMD that the LDS of an NB-element diffusion block allows: a CU's
// 160 KiB holds floor(160 KiB / block LDS) blocks, each putting ceil(waves/4)
// waves on every SIMD.  Used as the minimum-waves launch bound of the
// occupancy-capped variants, so that registers, not LDS, stop limiting the
// resident blocks (the compiler then keeps the VGPR count within 512 / W).
template <typename T, int D, int Q, int NB>
constexpr int lds_waves() {
    const int lds = NB * diffusion_buf<D, Q>() * (int)sizeof(T);
    const int blocks = (160 * 1024) / lds;
    const int per_simd = ((NB * Q * Q + 63) / 64 + 3) / 4;
    const int w = blocks * per_simd;
    return w > 8 ? 8 : (w < 1 ? 1 : w);
}

// The elements e0 + (threadIdx.x / Q^2), NB per work-group, of one problem.
// OM (output mode) 0: the element result goes to `out` (store_line).  OM 1
// (the column-sweep apply): e0 is this thread's element itself, valid when
// ne != 0; the element's D^3 results are left in LDS at s + 2 D^2 Q in
// lexicographic order (a region the last stage does not read) and its gather
// indices at sg + le D^3; `out` is not used.
template <typename T, typename TP, int D, int Q, int NB, int U, int K, int OM = 0>
__device__ __forceinline__ void
diffusion_elements_eo(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
                   const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
                   const T *__restrict__ x, const EOutT<T> &out, int accumulate,
                   T *__restrict__ smem, int *__restrict__ sg = nullptr) {
    using E = EO<D, Q>;
    constexpr int HD = E::HD, OD = E::OD, HQ = E::HQ, R = E::R, W = E::W, RB = E::RB;
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DDQ = D * D * Q, DQQ = D * Q * Q;
    constexpr int BUF = diffusion_buf<D, Q>();
    const T *__restrict__ FBt = basis + E::FB;  // forward rows of B
    const T *__restrict__ FGt = basis + E::FG;  // forward rows of G
    const T *__restrict__ BBt = basis + E::BB;  // backward rows of B
    const T *__restrict__ BGt = basis + E::BG;  // backward rows of G
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = OM ? e0 : e0 + le;
    const bool valid = OM ? ne != 0 : e < ne;
    T *s = smem + le * BUF;
    constexpr int KP = K > 0 ? ((K + 1) / 2 < R ? (K + 1) / 2 : R) : 0;  // pairs ahead
    const TP *pcol = pa + (int64_t)QQQ * 6 * (valid ? e : 0) + a + Q * b;
    // the 12 PA values of plane pair j: planes j and Q-1-j (the same plane
    // twice at the middle of an odd Q)
    auto load_pair = [&](int j, T (&r)[12]) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            r[c] = load_pa<T, stream_pa<Q, TP>()>(pcol + QQQ * c + QQ * j);
            r[6 + c] = load_pa<T, stream_pa<Q, TP>()>(pcol + QQQ * c + QQ * (Q - 1 - j));
        }
    };
    T ring[KP > 0 ? KP : 1][12];
    if constexpr (KP > 0) {
        if (valid) {
#pragma unroll
            for (int j = 0; j < KP; ++j) load_pair(j, ring[j]);
        }
    }

    // ---- forward x: thread (dy, dz) = (a, b); u -> B u, G u along x
    if (valid && a < D && b < D) {
        T ue[HD > 0 ? HD : 1], uo[HD > 0 ? HD : 1], um = T(0);
        {
            T u[D];
            const int64_t base = (int64_t)DDD * e + D * (a + D * b);
            load_line<D, T, OM == 1>(gmap, mask, x, base, u,
                                     OM == 1 ? sg + le * DDD + D * (a + D * b) : nullptr);
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                ue[d] = u[d] + u[D - 1 - d];
                uo[d] = u[d] - u[D - 1 - d];
            }
            if constexpr (OD) um = u[HD];
        }
        T *o = s + Q * (a + D * b);
#pragma unroll U
        for (int q = 0; q < R; ++q) {
            const T *fb = FBt + W * q, *fg = FGt + W * q;
            T Ab = OD ? fb[2 * HD] * um : T(0), Ag = OD ? fg[2 * HD] * um : T(0);
            T Cb = T(0), Cg = T(0);
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                Ab = fma(fb[d], ue[d], Ab);
                Cb = fma(fb[HD + d], uo[d], Cb);
                Ag = fma(fg[d], ue[d], Ag);
                Cg = fma(fg[HD + d], uo[d], Cg);
            }
            o[q] = Ab + Cb;
            o[DDQ + q] = Ag + Cg;
            if (q < HQ) {
                o[Q - 1 - q] = Ab - Cb;
                o[DDQ + Q - 1 - q] = Cg - Ag;
            }
        }
    }
    __syncthreads();

    // ---- forward y: thread (qx, dz) = (a, b); lines r0 = B_x u, r1 = G_x u
    {
        T r0e[HD > 0 ? HD : 1], r0o[HD > 0 ? HD : 1], r1e[HD > 0 ? HD : 1], r1o[HD > 0 ? HD : 1];
        T r0m = T(0), r1m = T(0);
        const bool act = valid && b < D;
        if (act) {
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                const T p0 = s[a + Q * (d + D * b)], q0 = s[a + Q * (D - 1 - d + D * b)];
                const T p1 = s[DDQ + a + Q * (d + D * b)], q1 = s[DDQ + a + Q * (D - 1 - d + D * b)];
                r0e[d] = p0 + q0;
                r0o[d] = p0 - q0;
                r1e[d] = p1 + q1;
                r1o[d] = p1 - q1;
            }
            if constexpr (OD) {
                r0m = s[a + Q * (HD + D * b)];
                r1m = s[DDQ + a + Q * (HD + D * b)];
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll U
            for (int q = 0; q < R; ++q) {
                const T *fb = FBt + W * q, *fg = FGt + W * q;
                T A0 = OD ? fb[2 * HD] * r0m : T(0), G0 = OD ? fg[2 * HD] * r0m : T(0);
                T A1 = OD ? fb[2 * HD] * r1m : T(0);
                T C0 = T(0), H0 = T(0), C1 = T(0);
#pragma unroll
                for (int d = 0; d < HD; ++d) {
                    A0 = fma(fb[d], r0e[d], A0);
                    C0 = fma(fb[HD + d], r0o[d], C0);
                    G0 = fma(fg[d], r0e[d], G0);
                    H0 = fma(fg[HD + d], r0o[d], H0);
                    A1 = fma(fb[d], r1e[d], A1);
                    C1 = fma(fb[HD + d], r1o[d], C1);
                }
                // bb = B_y B_x u (-> d/dz), gb = G_y B_x u (-> d/dy), bg = B_y G_x u (-> d/dx)
                T *o = s + a + Q * (q + Q * b);
                o[0 * DQQ] = A1 + C1;
                o[1 * DQQ] = G0 + H0;
                o[2 * DQQ] = A0 + C0;
                if (q < HQ) {
                    T *o2 = s + a + Q * (Q - 1 - q + Q * b);
                    o2[0 * DQQ] = A1 - C1;
                    o2[1 * DQQ] = H0 - G0;
                    o2[2 * DQQ] = A0 - C0;
                }
            }
        }
    }
    __syncthreads();

    // ---- z in registers: thread (qx, qy) = (a, b); plane pairs (j, Q-1-j)
    {
        T z0e[HD > 0 ? HD : 1], z0o[HD > 0 ? HD : 1], z1e[HD > 0 ? HD : 1], z1o[HD > 0 ? HD : 1];
        T z2e[HD > 0 ? HD : 1], z2o[HD > 0 ? HD : 1];
        T z0m = T(0), z1m = T(0), z2m = T(0);
        if (valid) {
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                const T *i = s + a + Q * (b + Q * d);
                const T *k = s + a + Q * (b + Q * (D - 1 - d));
                z0e[d] = i[0 * DQQ] + k[0 * DQQ];
                z0o[d] = i[0 * DQQ] - k[0 * DQQ];
                z1e[d] = i[1 * DQQ] + k[1 * DQQ];
                z1o[d] = i[1 * DQQ] - k[1 * DQQ];
                z2e[d] = i[2 * DQQ] + k[2 * DQQ];
                z2o[d] = i[2 * DQQ] - k[2 * DQQ];
            }
            if constexpr (OD) {
                const T *i = s + a + Q * (b + Q * HD);
                z0m = i[0 * DQQ];
                z1m = i[1 * DQQ];
                z2m = i[2 * DQQ];
            }
        }
        __syncthreads();
        if (valid) {
            // backward accumulators in even-odd form: E (Me), F (Mo), M (Mm)
            T E0[HD > 0 ? HD : 1], F0[HD > 0 ? HD : 1], E1[HD > 0 ? HD : 1], F1[HD > 0 ? HD : 1];
            T E2[HD > 0 ? HD : 1], F2[HD > 0 ? HD : 1];
            T M0 = T(0), M1 = T(0), M2 = T(0);
#pragma unroll
            for (int d = 0; d < HD; ++d) E0[d] = F0[d] = E1[d] = F1[d] = E2[d] = F2[d] = T(0);
            // with a prefetch ring, fully unrolled: the ring slots become
            // plain values, so no loop-carried copy waits for a prefetch that
            // is still in flight
            constexpr int ZU = KP > 0 ? R : 1;
#pragma unroll ZU
            for (int jb = 0; jb < R; jb += (KP > 0 ? KP : 1)) {
#pragma unroll
                for (int jj = 0; jj < (KP > 0 ? KP : 1); ++jj) {
                    const int j = jb + jj;
                    if (j < R) {
                        T pd[12];
                        if constexpr (KP > 0) {
#pragma unroll
                            for (int c = 0; c < 12; ++c) pd[c] = ring[jj][c];
                            if (j + KP < R) load_pair(j + KP, ring[jj]);
                        } else {
                            load_pair(j, pd);
                        }
                        const T *fb = FBt + W * j, *fg = FGt + W * j;
                        T A0 = OD ? fb[2 * HD] * z0m : T(0), A1 = OD ? fb[2 * HD] * z1m : T(0);
                        T A2 = OD ? fg[2 * HD] * z2m : T(0);
                        T C0 = T(0), C1 = T(0), C2 = T(0);
#pragma unroll
                        for (int d = 0; d < HD; ++d) {
                            A0 = fma(fb[d], z0e[d], A0);
                            C0 = fma(fb[HD + d], z0o[d], C0);
                            A1 = fma(fb[d], z1e[d], A1);
                            C1 = fma(fb[HD + d], z1o[d], C1);
                            A2 = fma(fg[d], z2e[d], A2);
                            C2 = fma(fg[HD + d], z2o[d], C2);
                        }
                        // gradient at plane j and at plane Q-1-j
                        const T g0 = A0 + C0, g1 = A1 + C1, g2 = A2 + C2;
                        const T h0 = A0 - C0, h1 = A1 - C1, h2 = C2 - A2;
                        const T O0 = fma(pd[0], g0, fma(pd[1], g1, pd[2] * g2));
                        const T O1 = fma(pd[1], g0, fma(pd[3], g1, pd[4] * g2));
                        const T O2 = fma(pd[2], g0, fma(pd[4], g1, pd[5] * g2));
                        T P0, N0, P1, N1, P2, N2;
                        if (j < HQ) {
                            const T Q0 = fma(pd[6], h0, fma(pd[7], h1, pd[8] * h2));
                            const T Q1 = fma(pd[7], h0, fma(pd[9], h1, pd[10] * h2));
                            const T Q2 = fma(pd[8], h0, fma(pd[10], h1, pd[11] * h2));
                            P0 = O0 + Q0;  // B: s = +1
                            N0 = O0 - Q0;
                            P1 = O1 + Q1;
                            N1 = O1 - Q1;
                            P2 = O2 - Q2;  // G: s = -1
                            N2 = O2 + Q2;
                        } else {  // middle plane of an odd Q
                            P0 = N0 = O0;
                            P1 = N1 = O1;
                            P2 = N2 = O2;
                        }
#pragma unroll
                        for (int d = 0; d < HD; ++d) {
                            E0[d] = fma(fb[d], P0, E0[d]);
                            F0[d] = fma(fb[HD + d], N0, F0[d]);
                            E1[d] = fma(fb[d], P1, E1[d]);
                            F1[d] = fma(fb[HD + d], N1, F1[d]);
                            E2[d] = fma(fg[d], P2, E2[d]);
                            F2[d] = fma(fg[HD + d], N2, F2[d]);
                        }
                        if constexpr (OD) {
                            M0 = fma(fb[2 * HD], P0, M0);
                            M1 = fma(fb[2 * HD], P1, M1);
                            M2 = fma(fg[2 * HD], P2, M2);
                        }
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                T *o = s + a + Q * (b + Q * d);
                T *k = s + a + Q * (b + Q * (D - 1 - d));
                o[0 * DQQ] = E0[d] + F0[d];
                k[0 * DQQ] = E0[d] - F0[d];
                o[1 * DQQ] = E1[d] + F1[d];
                k[1 * DQQ] = E1[d] - F1[d];
                o[2 * DQQ] = E2[d] + F2[d];
                k[2 * DQQ] = E2[d] - F2[d];
            }
            if constexpr (OD) {
                T *o = s + a + Q * (b + Q * HD);
                o[0 * DQQ] = M0;
                o[1 * DQQ] = M1;
                o[2 * DQQ] = M2;
            }
        }
    }
    __syncthreads();

    // ---- backward y: thread (qx, dz) = (a, b)
    //      sa = B_y^T t0,  sb = G_y^T t1 + B_y^T t2
    {
        T P0[R], N0[R], P1[R], N1[R], P2[R], N2[R];
        const bool act = valid && b < D;
        if (act) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const T *i = s + a + Q * (j + Q * b);
                if (j < HQ) {
                    const T *k = s + a + Q * (Q - 1 - j + Q * b);
                    P0[j] = i[0 * DQQ] + k[0 * DQQ];
                    N0[j] = i[0 * DQQ] - k[0 * DQQ];
                    P1[j] = i[1 * DQQ] - k[1 * DQQ];  // G
                    N1[j] = i[1 * DQQ] + k[1 * DQQ];
                    P2[j] = i[2 * DQQ] + k[2 * DQQ];
                    N2[j] = i[2 * DQQ] - k[2 * DQQ];
                } else {
                    P0[j] = N0[j] = i[0 * DQQ];
                    P1[j] = N1[j] = i[1 * DQQ];
                    P2[j] = N2[j] = i[2 * DQQ];
                }
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll U
            for (int d = 0; d < HD + OD; ++d) {
                const T *bb = BBt + RB * d, *bg = BGt + RB * d;
                T Ea = T(0), Fa = T(0), Eb = T(0), Fb = T(0);
                if (d < HD) {
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        Ea = fma(bb[j], P0[j], Ea);
                        Fa = fma(bb[R + j], N0[j], Fa);
                        Eb = fma(bg[j], P1[j], Eb);
                        Fb = fma(bg[R + j], N1[j], Fb);
                        Eb = fma(bb[j], P2[j], Eb);
                        Fb = fma(bb[R + j], N2[j], Fb);
                    }
                    s[a + Q * (d + D * b)] = Ea + Fa;
                    s[a + Q * (D - 1 - d + D * b)] = Ea - Fa;
                    s[DDQ + a + Q * (d + D * b)] = Eb + Fb;
                    s[DDQ + a + Q * (D - 1 - d + D * b)] = Eb - Fb;
                } else {  // middle d of an odd D
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        Ea = fma(bb[j], P0[j], Ea);
                        Eb = fma(bg[j], P1[j], Eb);
                        Eb = fma(bb[j], P2[j], Eb);
                    }
                    s[a + Q * (d + D * b)] = Ea;
                    s[DDQ + a + Q * (d + D * b)] = Eb;
                }
            }
        }
    }
    __syncthreads();

    // ---- backward x: thread (dy, dz) = (a, b); v = G_x^T sa + B_x^T sb
    if (valid && a < D && b < D) {
        T Pa[R], Na[R], Pb[R], Nb[R];
        const T *in = s + Q * (a + D * b);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j < HQ) {
                const T sa0 = in[j], sa1 = in[Q - 1 - j];
                const T sb0 = in[DDQ + j], sb1 = in[DDQ + Q - 1 - j];
                Pa[j] = sa0 - sa1;  // G
                Na[j] = sa0 + sa1;
                Pb[j] = sb0 + sb1;  // B
                Nb[j] = sb0 - sb1;
            } else {
                Pa[j] = Na[j] = in[j];
                Pb[j] = Nb[j] = in[DDQ + j];
            }
        }
        T yv[D];
#pragma unroll
        for (int d = 0; d < HD + OD; ++d) {
            const T *bb = BBt + RB * d, *bg = BGt + RB * d;
            T Ev = T(0), Fv = T(0);
            if (d < HD) {
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    Ev = fma(bg[j], Pa[j], Ev);
                    Fv = fma(bg[R + j], Na[j], Fv);
                    Ev = fma(bb[j], Pb[j], Ev);
                    Fv = fma(bb[R + j], Nb[j], Fv);
                }
                yv[d] = Ev + Fv;
                yv[D - 1 - d] = Ev - Fv;
            } else {
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    Ev = fma(bg[j], Pa[j], Ev);
                    Ev = fma(bb[j], Pb[j], Ev);
                }
                yv[d] = Ev;
            }
        }
        if constexpr (OM == 0) {
            store_line<D, T>(out, e, a, b, yv, accumulate);
        } else {
            T *o = s + 2 * DDQ + D * (a + D * b);
#pragma unroll
            for (int dx = 0; dx < D; ++dx) o[dx] = yv[dx];
        }
    }
}

template <typename T, typename TP, int D, int Q, int NB>
__device__ __forceinline__ void
mass_elements_eo(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
              const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
              const T *__restrict__ x, const EOutT<T> &out, int accumulate,
              T *__restrict__ smem) {
    using E = EO<D, Q>;
    constexpr int HD = E::HD, OD = E::OD, HQ = E::HQ, R = E::R, W = E::W, RB = E::RB;
    constexpr int QQ = Q * Q, QQQ = Q * Q * Q, DDD = D * D * D;
    constexpr int DQQ = D * Q * Q;
    constexpr int BUF = DQQ;
    const T *__restrict__ FBt = basis + E::FB;
    const T *__restrict__ BBt = basis + E::BB;
    const int le = threadIdx.x / QQ;
    const int t = threadIdx.x - le * QQ;
    const int a = t % Q, b = t / Q;
    const int64_t e = e0 + le;
    const bool valid = e < ne;
    T *s = smem + le * BUF;
    // the column's Q PA values are loaded first, so that their round trip
    // overlaps the gather and the x and y stages instead of the z stage
    T pz[Q];
    if (valid) {
        const TP *pcol = pa + (int64_t)QQQ * e + a + Q * b;
#pragma unroll
        for (int q = 0; q < Q; ++q) pz[q] = load_pa<T, stream_pa<Q, TP>()>(pcol + QQ * q);
    }
    // forward 1D contraction of a line held in even-odd form, B only
    auto fwd = [&](const T (&le_)[HD > 0 ? HD : 1], const T (&lo_)[HD > 0 ? HD : 1], T lm, int q,
                   T &lo, T &hi) {
        const T *fb = FBt + W * q;
        T A = OD ? fb[2 * HD] * lm : T(0), C = T(0);
#pragma unroll
        for (int d = 0; d < HD; ++d) {
            A = fma(fb[d], le_[d], A);
            C = fma(fb[HD + d], lo_[d], C);
        }
        lo = A + C;
        hi = A - C;
    };
    // x: thread (dy, dz)
    if (valid && a < D && b < D) {
        T ue[HD > 0 ? HD : 1], uo[HD > 0 ? HD : 1], um = T(0);
        {
            T u[D];
            const int64_t base = (int64_t)DDD * e + D * (a + D * b);
            load_line<D>(gmap, mask, x, base, u);
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                ue[d] = u[d] + u[D - 1 - d];
                uo[d] = u[d] - u[D - 1 - d];
            }
            if constexpr (OD) um = u[HD];
        }
#pragma unroll 1
        for (int q = 0; q < R; ++q) {
            T lo, hi;
            fwd(ue, uo, um, q, lo, hi);
            s[q + Q * (a + D * b)] = lo;
            if (q < HQ) s[Q - 1 - q + Q * (a + D * b)] = hi;
        }
    }
    __syncthreads();
    // y: thread (qx, dz)
    {
        T re[HD > 0 ? HD : 1], ro[HD > 0 ? HD : 1], rm = T(0);
        const bool act = valid && b < D;
        if (act) {
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                const T p = s[a + Q * (d + D * b)], k = s[a + Q * (D - 1 - d + D * b)];
                re[d] = p + k;
                ro[d] = p - k;
            }
            if constexpr (OD) rm = s[a + Q * (HD + D * b)];
        }
        __syncthreads();
        if (act)
#pragma unroll 1
            for (int q = 0; q < R; ++q) {
                T lo, hi;
                fwd(re, ro, rm, q, lo, hi);
                s[a + Q * (q + Q * b)] = lo;
                if (q < HQ) s[a + Q * (Q - 1 - q + Q * b)] = hi;
            }
    }
    __syncthreads();
    // z: thread (qx, qy); pairs of planes
    {
        T ze[HD > 0 ? HD : 1], zo[HD > 0 ? HD : 1], zm = T(0);
        if (valid) {
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                const T p = s[a + Q * (b + Q * d)], k = s[a + Q * (b + Q * (D - 1 - d))];
                ze[d] = p + k;
                zo[d] = p - k;
            }
            if constexpr (OD) zm = s[a + Q * (b + Q * HD)];
        }
        __syncthreads();
        if (valid) {
            T Ez[HD > 0 ? HD : 1], Fz[HD > 0 ? HD : 1], Mz = T(0);
#pragma unroll
            for (int d = 0; d < HD; ++d) Ez[d] = Fz[d] = T(0);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                T lo, hi;
                fwd(ze, zo, zm, j, lo, hi);
                const T v0 = lo * pz[j];
                T P, N;
                if (j < HQ) {
                    const T v1 = hi * pz[Q - 1 - j];
                    P = v0 + v1;
                    N = v0 - v1;
                } else {
                    P = N = v0;
                }
                const T *fb = FBt + W * j;
#pragma unroll
                for (int d = 0; d < HD; ++d) {
                    Ez[d] = fma(fb[d], P, Ez[d]);
                    Fz[d] = fma(fb[HD + d], N, Fz[d]);
                }
                if constexpr (OD) Mz = fma(fb[2 * HD], P, Mz);
            }
#pragma unroll
            for (int d = 0; d < HD; ++d) {
                s[a + Q * (b + Q * d)] = Ez[d] + Fz[d];
                s[a + Q * (b + Q * (D - 1 - d))] = Ez[d] - Fz[d];
            }
            if constexpr (OD) s[a + Q * (b + Q * HD)] = Mz;
        }
    }
    __syncthreads();
    // backward: line of Q values -> D values with B^T
    auto bwd = [&](const T (&P)[R], const T (&N)[R], int d, T &lo, T &hi) {
        const T *bb = BBt + RB * d;
        T Ev = T(0), Fv = T(0);
        if (d < HD) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                Ev = fma(bb[j], P[j], Ev);
                Fv = fma(bb[R + j], N[j], Fv);
            }
            lo = Ev + Fv;
            hi = Ev - Fv;
        } else {
#pragma unroll
            for (int j = 0; j < R; ++j) Ev = fma(bb[j], P[j], Ev);
            lo = hi = Ev;
        }
    };
    // y^T: thread (qx, dz)
    {
        T P[R], N[R];
        const bool act = valid && b < D;
        if (act) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const T v0 = s[a + Q * (j + Q * b)];
                if (j < HQ) {
                    const T v1 = s[a + Q * (Q - 1 - j + Q * b)];
                    P[j] = v0 + v1;
                    N[j] = v0 - v1;
                } else {
                    P[j] = N[j] = v0;
                }
            }
        }
        __syncthreads();
        if (act)
#pragma unroll 1
            for (int d = 0; d < HD + OD; ++d) {
                T lo, hi;
                bwd(P, N, d, lo, hi);
                s[a + Q * (d + D * b)] = lo;
                if (d < HD) s[a + Q * (D - 1 - d + D * b)] = hi;
            }
    }
    __syncthreads();
    // x^T: thread (dy, dz)
    if (valid && a < D && b < D) {
        T P[R], N[R];
        const T *in = s + Q * (a + D * b);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j < HQ) {
                P[j] = in[j] + in[Q - 1 - j];
                N[j] = in[j] - in[Q - 1 - j];
            } else {
                P[j] = N[j] = in[j];
            }
        }
        T yv[D];
#pragma unroll
        for (int d = 0; d < HD + OD; ++d) {
            T lo, hi;
            bwd(P, N, d, lo, hi);
            yv[d] = lo;
            if (d < HD) yv[D - 1 - d] = hi;
        }
        store_line<D, T>(out, e, a, b, yv, accumulate);
    }
}

// Direct form (no even-odd split), used where occupancy matters more than
// VALU issue: FP64 at p <= 5.
// U: unroll factor of the output-line loops (1 or 2).
// K: PA-data pipeline depth.  0: each z-plane's 6 values are loaded in its
//    own iteration.  K > 0: the first K planes are loaded at kernel entry
//    (their latency hides behind the x/y stages) and plane qz + K is issued
//    while plane qz is computed, from a K x 6 register ring.
// The elements e0 + (threadIdx.x / Q^2), NB per work-group, of one problem.
// OM (output mode) 0: the element result goes to `out` (store_line).  OM 1
// (the column-sweep apply): e0 is this thread's element itself, valid when
// ne != 0; the element's D^3 results are left in LDS at s + 2 D^2 Q in
// lexicographic order (a region the last stage does not read) and its gather
// indices at sg + le D^3; `out` is not used.
template <typename T, typename TP, int D, int Q, int NB, int U, int K, int OM = 0>
__device__ __forceinline__ void
diffusion_elements_direct(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
                   const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
                   const T *__restrict__ x, const EOutT<T> &out, int accumulate,
                   T *__restrict__ smem, int *__restrict__ sg = nullptr) {
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
    const int64_t e = OM ? e0 : e0 + le;
    const bool valid = OM ? ne != 0 : e < ne;
    T *s = smem + le * BUF;
    constexpr int KK = K < Q ? K : Q;
    const TP *pcol = pa + (int64_t)QQQ * 6 * (valid ? e : 0) + a + Q * b;
    T ring[KK > 0 ? KK : 1][6];
    if constexpr (KK > 0) {
        if (valid) {
#pragma unroll
            for (int j = 0; j < KK; ++j)
#pragma unroll
                for (int c = 0; c < 6; ++c) ring[j][c] = load_pa<T, stream_pa<Q, TP>()>(pcol + QQQ * c + QQ * j);
        }
    }

    // ---- forward x: thread (dy, dz) = (a, b)
    if (valid && a < D && b < D) {
        T u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
        load_line<D, T, OM == 1>(gmap, mask, x, base, u,
                                     OM == 1 ? sg + le * DDD + D * (a + D * b) : nullptr);
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
            // fully unrolled: the ring slots become plain values, so no
            // loop-carried copy waits for a prefetch that is still in flight
#pragma unroll Q
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
                                ring[j][c] = load_pa<T, stream_pa<Q, TP>()>(pcol + QQQ * c + QQ * (qz + KK));
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
                const T d11 = load_pa<T, stream_pa<Q, TP>()>(pq + 0 * QQQ), d21 = load_pa<T, stream_pa<Q, TP>()>(pq + 1 * QQQ), d31 = load_pa<T, stream_pa<Q, TP>()>(pq + 2 * QQQ);
                const T d22 = load_pa<T, stream_pa<Q, TP>()>(pq + 3 * QQQ), d32 = load_pa<T, stream_pa<Q, TP>()>(pq + 4 * QQQ), d33 = load_pa<T, stream_pa<Q, TP>()>(pq + 5 * QQQ);
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
        T yv[D];
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            const T *bc = Bq + Q * dx, *gc = Gq + Q * dx;
            T v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) {
                v = fma(gc[qx], sa[qx], v);
                v = fma(bc[qx], sb[qx], v);
            }
            yv[dx] = v;
        }
        if constexpr (OM == 0) {
            store_line<D, T>(out, e, a, b, yv, accumulate);
        } else {
            T *o = s + 2 * DDQ + D * (a + D * b);
#pragma unroll
            for (int dx = 0; dx < D; ++dx) o[dx] = yv[dx];
        }
    }
}

template <typename T, typename TP, int D, int Q, int NB>
__device__ __forceinline__ void
mass_elements_direct(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
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
    // the column's Q PA values first (their round trip overlaps the gather
    // and the x and y stages)
    T pz[Q];
    if (valid) {
        const TP *pcol = pa + (int64_t)QQQ * e + a + Q * b;
#pragma unroll
        for (int q = 0; q < Q; ++q) pz[q] = load_pa<T, stream_pa<Q, TP>()>(pcol + QQ * q);
    }
    if (valid && a < D && b < D) {
        T u[D];
        const int64_t base = (int64_t)DDD * e + D * (a + D * b);
        load_line<D>(gmap, mask, x, base, u);
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
            T tz[D];
#pragma unroll
            for (int dz = 0; dz < D; ++dz) tz[dz] = 0.0;
#pragma unroll
            for (int qz = 0; qz < Q; ++qz) {
                const T *br = Bd + D * qz;
                T v = 0.0;
#pragma unroll
                for (int dz = 0; dz < D; ++dz) v = fma(br[dz], z[dz], v);
                v *= pz[qz];
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
        T yv[D];
#pragma unroll
        for (int dx = 0; dx < D; ++dx) {
            const T *bc = Bq + Q * dx;
            T v = 0.0;
#pragma unroll
            for (int qx = 0; qx < Q; ++qx) v = fma(bc[qx], sx[qx], v);
            yv[dx] = v;
        }
        store_line<D, T>(out, e, a, b, yv, accumulate);
    }
}


// Even-odd where the kernels are VALU-bound (every FP32/BF16 instantiation,
// FP64 from p = 6); the direct form keeps more waves resident where HBM
// bounds the FP64 kernel (p <= 5).  Both give the same operator to rounding.
template <typename T, int D>
constexpr bool use_even_odd() { return sizeof(T) == 4 || D >= 7; }

template <typename T, typename TP, int D, int Q, int NB, int U, int K, int OM = 0>
__device__ __forceinline__ void
diffusion_elements(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
                   const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
                   const T *__restrict__ x, const EOutT<T> &out, int accumulate,
                   T *__restrict__ smem, int *__restrict__ sg = nullptr) {
    if constexpr (use_even_odd<T, D>())
        diffusion_elements_eo<T, TP, D, Q, NB, U, K, OM>(e0, ne, basis, pa, gmap, mask, x, out,
                                                         accumulate, smem, sg);
    else
        diffusion_elements_direct<T, TP, D, Q, NB, U, K, OM>(e0, ne, basis, pa, gmap, mask, x,
                                                             out, accumulate, smem, sg);
}

template <typename T, typename TP, int D, int Q, int NB>
__device__ __forceinline__ void
mass_elements(int64_t e0, int64_t ne, const T *__restrict__ basis, const TP *__restrict__ pa,
              const int *__restrict__ gmap, const uint8_t *__restrict__ mask,
              const T *__restrict__ x, const EOutT<T> &out, int accumulate,
              T *__restrict__ smem) {
    if constexpr (use_even_odd<T, D>())
        mass_elements_eo<T, TP, D, Q, NB>(e0, ne, basis, pa, gmap, mask, x, out, accumulate, smem);
    else
        mass_elements_direct<T, TP, D, Q, NB>(e0, ne, basis, pa, gmap, mask, x, out, accumulate,
                                              smem);
}

}  // namespace kern
}  // namespace h1pa
