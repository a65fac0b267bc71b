// qinterp.hip -- QuadratureInterpolator::Mult for H1 E-vectors on hexahedra
// (fem/quadinterpolator.cpp:226-341, tensor kernels fem/qinterp/eval.hpp and
// fem/qinterp/grad.hpp:233-375; determinants as fem/qinterp/det.hpp).
//
// One work-group per element with a Q x Q thread plane: each component of
// the element's lexicographic E-vector block (D^3 values) is staged in LDS
// and contracted along x, y, z (B for values, B/G for the three reference
// derivatives), the z stage in registers per (qx, qy) column.  Physical
// derivatives apply J^-T from the mesh's geometric factors at the same
// points; determinants (vdim = 3) are taken of the 3 x 3 reference
// derivative matrix in a second, per-point pass.
#include <hip/hip_runtime.h>

#include <vector>

#include "internal.hpp"

namespace h1pa {
namespace kern {
namespace {

// Q-vector offsets of value (c) and derivative (c, d) at point q of element e.
__device__ __forceinline__ int64_t val_at(int by_vdim, int64_t q, int c, int64_t e, int64_t nq,
                                          int vdim) {
    return by_vdim ? c + vdim * (q + nq * e) : q + nq * (c + vdim * e);
}
__device__ __forceinline__ int64_t der_at(int by_vdim, int64_t q, int c, int d, int64_t e,
                                          int64_t nq, int vdim) {
    return by_vdim ? c + vdim * (d + 3 * (q + nq * e)) : q + nq * (c + vdim * (d + 3 * e));
}

template <int D, int Q>
__global__ void __launch_bounds__(Q * Q)
k_qinterp(int64_t ne, int vdim, int want_val, int want_der, int phys, int by_vdim,
          const double *__restrict__ basis, const double *__restrict__ J,
          const double *__restrict__ x, double *__restrict__ val, double *__restrict__ der) {
    constexpr int QQ = Q * Q, NQ = Q * Q * Q, DDD = D * D * D;
    __shared__ double sB[Q * D], sG[Q * D];       // B[q + Q d]
    __shared__ double X[DDD];                      // X[dx + D (dy + D dz)]
    __shared__ double A0[D * D * Q], A1[D * D * Q];  // [qx + Q (dy + D dz)]
    __shared__ double C0[D * QQ], C1[D * QQ], C2[D * QQ];  // [qx + Q (qy + Q dz)]
    const int64_t e = blockIdx.x;
    const int t = threadIdx.x;
    const int a = t % Q, b = t / Q;
    for (int i = t; i < Q * D; i += QQ) {
        sB[i] = basis[i];
        sG[i] = basis[Q * D + i];
    }
    for (int c = 0; c < vdim; ++c) {
        const double *xe = x + (int64_t)DDD * (c + (int64_t)vdim * e);
        for (int i = t; i < DDD; i += QQ) X[i] = xe[i];
        __syncthreads();
        // x: u = sum_dx X B(qx, dx), v = sum_dx X G(qx, dx) for rows (dy, dz)
        for (int r = b; r < D * D; r += Q) {
            double u = 0.0, v = 0.0;
#pragma unroll
            for (int dx = 0; dx < D; ++dx) {
                const double in = X[dx + D * r];
                u = fma(in, sB[a + Q * dx], u);
                v = fma(in, sG[a + Q * dx], v);
            }
            A0[a + Q * r] = u;
            A1[a + Q * r] = v;
        }
        __syncthreads();
        // y: thread (qx, qy) over dz
        for (int dz = 0; dz < D; ++dz) {
            double u = 0.0, v = 0.0, w = 0.0;
#pragma unroll
            for (int dy = 0; dy < D; ++dy) {
                const double a0 = A0[a + Q * (dy + D * dz)], a1 = A1[a + Q * (dy + D * dz)];
                u = fma(a1, sB[b + Q * dy], u);  // d/dx
                v = fma(a0, sG[b + Q * dy], v);  // d/dy
                w = fma(a0, sB[b + Q * dy], w);  // value, d/dz
            }
            C0[a + Q * (b + Q * dz)] = u;
            C1[a + Q * (b + Q * dz)] = v;
            C2[a + Q * (b + Q * dz)] = w;
        }
        __syncthreads();
        // z: thread (qx, qy), all qz
        double c0[D], c1[D], c2[D];
#pragma unroll
        for (int dz = 0; dz < D; ++dz) {
            c0[dz] = C0[a + Q * (b + Q * dz)];
            c1[dz] = C1[a + Q * (b + Q * dz)];
            c2[dz] = C2[a + Q * (b + Q * dz)];
        }
        for (int qz = 0; qz < Q; ++qz) {
            double gx = 0.0, gy = 0.0, gz = 0.0, vv = 0.0;
#pragma unroll
            for (int dz = 0; dz < D; ++dz) {
                const double bz = sB[qz + Q * dz], g = sG[qz + Q * dz];
                gx = fma(c0[dz], bz, gx);
                gy = fma(c1[dz], bz, gy);
                gz = fma(c2[dz], g, gz);
                vv = fma(c2[dz], bz, vv);
            }
            const int64_t q = a + Q * (b + Q * qz);
            if (want_val) val[val_at(by_vdim, q, c, e, NQ, vdim)] = vv;
            if (want_der) {
                if (phys) {
                    // grad_x = J^-T grad_xi with J(r, c) = dx_r / dxi_c = j[r + 3 c]:
                    // (J^-T)(r, c) = C(r, c) / det J, C the cofactor matrix
                    double j[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) j[k] = J[q + NQ * (k + 9 * e)];
                    const double c00 = j[4] * j[8] - j[7] * j[5];
                    const double c01 = j[7] * j[2] - j[1] * j[8];
                    const double c02 = j[1] * j[5] - j[4] * j[2];
                    const double c10 = j[6] * j[5] - j[3] * j[8];
                    const double c11 = j[0] * j[8] - j[6] * j[2];
                    const double c12 = j[3] * j[2] - j[0] * j[5];
                    const double c20 = j[3] * j[7] - j[6] * j[4];
                    const double c21 = j[6] * j[1] - j[0] * j[7];
                    const double c22 = j[0] * j[4] - j[3] * j[1];
                    const double s = 1.0 / (j[0] * c00 + j[3] * c01 + j[6] * c02);
                    const double U = (c00 * gx + c01 * gy + c02 * gz) * s;
                    const double V = (c10 * gx + c11 * gy + c12 * gz) * s;
                    const double W = (c20 * gx + c21 * gy + c22 * gz) * s;
                    gx = U;
                    gy = V;
                    gz = W;
                }
                der[der_at(by_vdim, q, c, 0, e, NQ, vdim)] = gx;
                der[der_at(by_vdim, q, c, 1, e, NQ, vdim)] = gy;
                der[der_at(by_vdim, q, c, 2, e, NQ, vdim)] = gz;
            }
        }
        __syncthreads();
    }
}

// det of the 3 x 3 reference derivative matrix M(c, d) (byNODES derivatives)
__global__ void k_qi_det(int64_t nq, int64_t ne, const double *__restrict__ der,
                         double *__restrict__ det) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nq * ne) return;
    const int64_t q = i % nq, e = i / nq;
    double m[9];  // m[c + 3 d]
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = der[q + nq * (k + 9 * e)];
    det[i] = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[3] * (m[1] * m[8] - m[2] * m[7]) +
             m[6] * (m[1] * m[5] - m[2] * m[4]);
}

template <int D, int Q>
void launch_qi(int64_t ne, int vdim, int want_val, int want_der, int phys, int by_vdim,
               const double *basis, const double *J, const double *x, double *val, double *der,
               hipStream_t s) {
    hipLaunchKernelGGL((k_qinterp<D, Q>), dim3((unsigned)ne), dim3(Q * Q), 0, s, ne, vdim,
                       want_val, want_der, phys, by_vdim, basis, J, x, val, der);
}

}  // namespace

bool qinterp_supported(int p, int q1d) {
    return p >= 1 && p <= 8 && q1d >= p + 1 && q1d <= p + 3 && q1d <= 10;
}

void qinterp(Space &sp, int q1d, int vdim, unsigned flags, int by_vdim, const double *x,
             double *val, double *der, double *det, hipStream_t s) {
    const int p = sp.p, D = p + 1, Q = q1d;
    const int64_t ne = sp.ne, nq = (int64_t)Q * Q * Q;
    const bool want_val = flags & H1PA_QI_VALUES;
    const bool phys = flags & H1PA_QI_PHYSICAL_DERIVATIVES;
    const bool want_der = (flags & H1PA_QI_DERIVATIVES) || phys;
    const bool want_det = flags & H1PA_QI_DETERMINANTS;
    if (ne == 0) return;
    // basis at the Gauss-Legendre points
    std::vector<double> xq(Q), wq(Q), h(2 * Q * D);
    gauss_legendre(Q, xq.data(), wq.data());
    lagrange_basis(p, Q, xq.data(), h.data(), h.data() + Q * D);
    StreamBuf<double> basis;
    basis.upload(h.data(), h.size(), s);
    StreamBuf<double> J;
    if (phys) {
        sp.mesh->ensure_nodes(s);
        J.alloc(nq * 9 * ne, s);
        geometric_factors(*sp.mesh, q1d, nullptr, J.p, nullptr, s);
    }
    // determinants need byNODES reference derivatives
    StreamBuf<double> dtmp;
    const double *dref = nullptr;
    if (want_det) {
        if ((flags & H1PA_QI_DERIVATIVES) && !by_vdim) {
            dref = der;
        } else {
            dtmp.alloc(nq * 9 * ne, s);
            dref = dtmp.p;
        }
    }
    auto run = [&](int wv, int wd, int ph, int bv, double *dv) {
#define X(DD, QQ)                                                                         \
    if (D == DD && Q == QQ) {                                                             \
        launch_qi<DD, QQ>(ne, vdim, wv, wd, ph, bv, basis.p, J.p, x, val, dv, s);         \
        H1PA_HIP(hipGetLastError());                                                      \
        return;                                                                           \
    }
        X(2, 2) X(2, 3) X(2, 4) X(3, 3) X(3, 4) X(3, 5) X(4, 4) X(4, 5) X(4, 6) X(5, 5)
        X(5, 6) X(5, 7) X(6, 6) X(6, 7) X(6, 8) X(7, 7) X(7, 8) X(7, 9) X(8, 8) X(8, 9)
        X(8, 10) X(9, 9) X(9, 10)
#undef X
        throw Error{H1PA_EUNSUP, "qinterp: unsupported (p, q1d)"};
    };
    if (want_val || want_der) run(want_val, want_der, phys, by_vdim, der);
    if (want_det) {
        if (dref != der) run(0, 1, 0, 0, dtmp.p);
        hipLaunchKernelGGL(k_qi_det, dim3((unsigned)((nq * ne + 255) / 256)), dim3(256), 0, s, nq,
                           ne, dref, det);
        H1PA_HIP(hipGetLastError());
    }
    // the temporaries are freed at return; hipFree waits for the kernels
}

}  // namespace kern
}  // namespace h1pa
