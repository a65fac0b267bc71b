// geometry.hip -- mesh nodes, geometric factors and PA setup on the device.
//
// The setup is one fused pass per (quadrature point, element): Jacobian from
// the element nodes (QuadratureInterpolator derivatives, fem/qinterp/grad.hpp:
// 233-375), adj(J) and det(J), the coefficient at the physical point, and the
// PA data (PADiffusionSetup3D, fem/integ/bilininteg_diffusion_kernels.cpp:
// 262-386; MassIntegrator::AssemblePA, fem/integ/bilininteg_mass_pa.cpp:76-88).
// J is never stored: at 32M DOFs and p = 1 it alone would be 61 GB.
#include <cmath>

#include "internal.hpp"

namespace h1pa {
namespace kern {

namespace {

constexpr int kBlock = 256;
inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct GeomBasis {          // geometry basis at the Q Gauss-Legendre points
    int Q, Dg;
    double B[10 * 10];      // B[q + Q*d]
    double G[10 * 10];
    double w[10];           // 1D Gauss-Legendre weights
};

GeomBasis make_geom_basis(int Q, int pg) {
    GeomBasis gb{};
    gb.Q = Q;
    gb.Dg = pg + 1;
    double x[16], w[16];
    gauss_legendre(Q, x, w);
    lagrange_basis(pg, Q, x, gb.B, gb.G);
    for (int i = 0; i < Q; ++i) gb.w[i] = w[i];
    return gb;
}

struct Fourier8 {           // g(x) = sum_m a cos(2 pi k_m . x + phi_m)
    double k[8][3];
    double phi[8];
    double a;
};

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

Fourier8 make_fourier(uint64_t seed) {
    // 8 modes; integer wave vectors in [-3,3]^3 \ {0}, phases U[0, 2 pi),
    // amplitude 0.5 / sqrt(8) each (root sum of squares 0.5; RMS of g
    // 0.5 / sqrt(2) for distinct modes).
    Fourier8 f{};
    uint64_t ctr = 0;
    auto draw = [&]() { return (double)(splitmix64(seed ^ splitmix64(ctr++)) >> 11) * (1.0 / 9007199254740992.0); };
    for (int m = 0; m < 8; ++m) {
        int kk[3];
        do {
            for (int c = 0; c < 3; ++c) kk[c] = (int)std::floor(draw() * 7.0) - 3;
        } while (kk[0] == 0 && kk[1] == 0 && kk[2] == 0);
        for (int c = 0; c < 3; ++c) f.k[m][c] = kk[c];
        f.phi[m] = 6.283185307179586 * draw();
    }
    f.a = 0.5 / std::sqrt(8.0);
    return f;
}

// x[i] = lo + (hi - lo) u_i, u_i = (splitmix64(seed ^ splitmix64(i)) >> 11) 2^-53.
__global__ void k_randomize(int64_t n, double *__restrict__ x, uint64_t seed, double lo,
                            double hi) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u = (double)(splitmix64(seed ^ splitmix64((uint64_t)i)) >> 11) *
                     (1.0 / 9007199254740992.0);
    x[i] = lo + (hi - lo) * u;
}

// ---- nodes -------------------------------------------------------------------
__global__ void k_nodes_trilinear(int64_t ne, const int *__restrict__ topo,
                                  const double *__restrict__ xyz, double *__restrict__ X) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne * 8) return;
    const int64_t e = i / 8;
    const int l = (int)(i % 8);
    // lex (a,b,c) -> MFEM hex vertex (p = 1 dof_map)
    const int lex2v[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    const int v = topo[29 * e + lex2v[l]];
    for (int c = 0; c < 3; ++c) X[l + 8 * (c + 3 * e)] = xyz[3 * (int64_t)v + c];
}

__global__ void k_nodes_curved(int64_t ne, int Dg, const int *__restrict__ topo,
                               const double *__restrict__ xyz, const double *__restrict__ gll,
                               double amp, double *__restrict__ X) {
    const int nd = Dg * Dg * Dg;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne * nd) return;
    const int64_t e = i / nd;
    const int l = (int)(i - e * nd);
    const double s[3] = {gll[l % Dg], gll[(l / Dg) % Dg], gll[l / (Dg * Dg)]};
    const int lex2v[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    double P[3] = {0.0, 0.0, 0.0};
    for (int c8 = 0; c8 < 8; ++c8) {
        const double N = ((c8 & 1) ? s[0] : 1.0 - s[0]) * ((c8 & 2) ? s[1] : 1.0 - s[1]) *
                         ((c8 & 4) ? s[2] : 1.0 - s[2]);
        const int v = topo[29 * e + lex2v[c8]];
        for (int c = 0; c < 3; ++c) P[c] += N * xyz[3 * (int64_t)v + c];
    }
    const double tp = 6.283185307179586476925286766559;
    const double xp = P[0] + amp * sin(tp * P[1]) * sin(tp * P[2]);
    const double yp = P[1] + amp * sin(tp * P[2]) * sin(tp * P[0]);
    const double zp = P[2] + amp * sin(tp * P[0]) * sin(tp * P[1]);
    X[l + nd * (0 + 3 * e)] = xp;
    X[l + nd * (1 + 3 * e)] = yp;
    X[l + nd * (2 + 3 * e)] = zp;
}

// Jacobian J(r, c) = dx_r / dxi_c and the physical point at quadrature point
// (qx, qy, qz) of element e, from nodes X[Dg^3][3][ne].
__device__ inline void eval_geom(const GeomBasis &gb, const double *__restrict__ X, int64_t e,
                                 int qx, int qy, int qz, double J[3][3], double P[3]) {
    const int Dg = gb.Dg, Q = gb.Q, nd = Dg * Dg * Dg;
    for (int r = 0; r < 3; ++r) {
        P[r] = 0.0;
        for (int c = 0; c < 3; ++c) J[r][c] = 0.0;
    }
    for (int k = 0; k < Dg; ++k) {
        const double bz = gb.B[qz + Q * k], gz = gb.G[qz + Q * k];
        for (int j = 0; j < Dg; ++j) {
            const double by = gb.B[qy + Q * j], gy = gb.G[qy + Q * j];
            for (int i = 0; i < Dg; ++i) {
                const double bx = gb.B[qx + Q * i], gx = gb.G[qx + Q * i];
                const int l = i + Dg * (j + Dg * k);
                const double w0 = gx * by * bz, w1 = bx * gy * bz, w2 = bx * by * gz,
                             wv = bx * by * bz;
                for (int r = 0; r < 3; ++r) {
                    const double xv = X[l + nd * (r + 3 * e)];
                    J[r][0] += w0 * xv;
                    J[r][1] += w1 * xv;
                    J[r][2] += w2 * xv;
                    P[r] += wv * xv;
                }
            }
        }
    }
}

__device__ inline double coeff_value(int kind, double cval, const double *__restrict__ cdata,
                                     const Fourier8 &fm, int64_t e, int q, int nq,
                                     const double P[3]) {
    const double tp = 6.283185307179586476925286766559;
    switch (kind) {
        case H1PA_COEFF_SINCOS:
            return 1.0 + 0.5 * sin(tp * P[0]) * cos(tp * P[1]) * sin(tp * P[2]);
        case H1PA_COEFF_PER_ELEMENT: return cdata[e];
        case H1PA_COEFF_QUADRATURE: return cdata[q + (int64_t)nq * e];
        case H1PA_COEFF_FOURIER: {
            double g = 0.0;
            for (int m = 0; m < 8; ++m)
                g += fm.a * cos(tp * (fm.k[m][0] * P[0] + fm.k[m][1] * P[1] + fm.k[m][2] * P[2]) +
                                fm.phi[m]);
            return exp(g);
        }
        default: return cval;
    }
}

__global__ void k_geometric_factors(int64_t ne, GeomBasis gb, const double *__restrict__ X,
                                    double *__restrict__ Xq, double *__restrict__ Jq,
                                    double *__restrict__ detJ) {
    const int Q = gb.Q, nq = Q * Q * Q;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne * nq) return;
    const int64_t e = i / nq;
    const int q = (int)(i - e * nq);
    double J[3][3], P[3];
    eval_geom(gb, X, e, q % Q, (q / Q) % Q, q / (Q * Q), J, P);
    if (Xq)
        for (int r = 0; r < 3; ++r) Xq[q + (int64_t)nq * (r + 3 * e)] = P[r];
    if (Jq)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) Jq[q + (int64_t)nq * (r + 3 * (c + 3 * e))] = J[r][c];
    if (detJ)
        detJ[q + (int64_t)nq * e] = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                                    J[1][0] * (J[0][1] * J[2][2] - J[2][1] * J[0][2]) +
                                    J[2][0] * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
}

__global__ void k_pa_setup(int64_t ne, GeomBasis gb, const double *__restrict__ X, int kind,
                           double cval, const double *__restrict__ cdata, Fourier8 fm,
                           double *__restrict__ Dd, double *__restrict__ Mv) {
    const int Q = gb.Q, nq = Q * Q * Q;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne * nq) return;
    const int64_t e = i / nq;
    const int q = (int)(i - e * nq);
    const int qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
    double J[3][3], P[3];
    eval_geom(gb, X, e, qx, qy, qz, J, P);
    const double W = gb.w[qx] * gb.w[qy] * gb.w[qz];
    const double k = coeff_value(kind, cval, cdata, fm, e, q, nq, P);
    const double J11 = J[0][0], J21 = J[1][0], J31 = J[2][0];
    const double J12 = J[0][1], J22 = J[1][1], J32 = J[2][1];
    const double J13 = J[0][2], J23 = J[1][2], J33 = J[2][2];
    const double detJ = J11 * (J22 * J33 - J32 * J23) - J21 * (J12 * J33 - J32 * J13) +
                        J31 * (J12 * J23 - J22 * J13);
    if (Mv) Mv[q + (int64_t)nq * e] = W * k * detJ;
    if (Dd) {
        // (1/detJ) adj(J) adj(J)^T, symmetric storage 11,21,31,22,32,33
        const double A11 = J22 * J33 - J23 * J32, A12 = J32 * J13 - J12 * J33,
                     A13 = J12 * J23 - J22 * J13;
        const double A21 = J31 * J23 - J21 * J33, A22 = J11 * J33 - J13 * J31,
                     A23 = J21 * J13 - J11 * J23;
        const double A31 = J21 * J32 - J31 * J22, A32 = J31 * J12 - J11 * J32,
                     A33 = J11 * J22 - J12 * J21;
        const double s = W / detJ * k;
        double *d = Dd + (int64_t)nq * 6 * e + q;
        d[0 * nq] = s * (A11 * A11 + A12 * A12 + A13 * A13);
        d[1 * nq] = s * (A11 * A21 + A12 * A22 + A13 * A23);
        d[2 * nq] = s * (A11 * A31 + A12 * A32 + A13 * A33);
        d[3 * nq] = s * (A21 * A21 + A22 * A22 + A23 * A23);
        d[4 * nq] = s * (A21 * A31 + A22 * A32 + A23 * A33);
        d[5 * nq] = s * (A31 * A31 + A32 * A32 + A33 * A33);
    }
}

// DomainLFIntegrator with constant f, Q = p + 1 (fem/lininteg.cpp:51-79):
// one block per element; detJ at the Q^3 points staged in LDS first.
// f at quadrature point q of element e: fq[q + nq e] when fq is given (a
// QuadratureFunction, points x fastest), else the constant f.
__global__ void k_domain_lf(int64_t ne, int D, GeomBasis gb, const double *__restrict__ Bs,
                            const double *__restrict__ X, double f,
                            const double *__restrict__ fq, double *__restrict__ bE) {
    __shared__ double wdet[1000];
    const int Q = gb.Q, nq = Q * Q * Q, nd = D * D * D;
    const int64_t e = blockIdx.x;
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
        const int qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
        double J[3][3], P[3];
        eval_geom(gb, X, e, qx, qy, qz, J, P);
        const double detJ = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                            J[1][0] * (J[0][1] * J[2][2] - J[2][1] * J[0][2]) +
                            J[2][0] * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
        wdet[q] = gb.w[qx] * gb.w[qy] * gb.w[qz] * detJ * (fq ? fq[q + (int64_t)nq * e] : f);
    }
    __syncthreads();
    for (int l = threadIdx.x; l < nd; l += blockDim.x) {
        const int dx = l % D, dy = (l / D) % D, dz = l / (D * D);
        double s = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
            s += wdet[q] * Bs[qx + Q * dx] * Bs[qy + Q * dy] * Bs[qz + Q * dz];
        }
        bE[l + (int64_t)nd * e] = s;
    }
}

}  // namespace

void nodes_trilinear(const Mesh &m, double *d_nodes, hipStream_t s) {
    StreamBuf<double> xyz;
    xyz.upload(m.xyz.data(), m.xyz.size(), s);
    hipLaunchKernelGGL(k_nodes_trilinear, dim3(nblocks(m.ne * 8)), dim3(kBlock), 0, s, m.ne,
                       m.d_topo.p, xyz.p, d_nodes);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));
}

void nodes_curved(const Mesh &m, int pg, double amp, double *d_nodes, hipStream_t s) {
    const int Dg = pg + 1;
    double z[16], w[16];
    gauss_lobatto(Dg, z, w);
    StreamBuf<double> xyz, gll;
    xyz.upload(m.xyz.data(), m.xyz.size(), s);
    gll.upload(z, Dg, s);
    hipLaunchKernelGGL(k_nodes_curved, dim3(nblocks(m.ne * Dg * Dg * Dg)), dim3(kBlock), 0, s, m.ne,
                       Dg, m.d_topo.p, xyz.p, gll.p, amp, d_nodes);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));
}

void geometric_factors(const Mesh &m, int q1d, double *X, double *J, double *detJ,
                       hipStream_t s) {
    H1PA_CHECK(q1d >= 1 && q1d <= 10, H1PA_EUNSUP, "q1d must be in [1,10]");
    const GeomBasis gb = make_geom_basis(q1d, m.node_order);
    const int64_t n = m.ne * q1d * q1d * q1d;
    hipLaunchKernelGGL(k_geometric_factors, dim3(nblocks(n)), dim3(kBlock), 0, s, m.ne, gb,
                       m.d_nodes.p, X, J, detJ);
    H1PA_HIP(hipGetLastError());
}

void pa_setup(const Mesh &m, const Form &f, hipStream_t s) {
    const int Q = f.q1d;
    H1PA_CHECK(Q >= 1 && Q <= 10, H1PA_EUNSUP, "q1d must be in [1,10]");
    const GeomBasis gb = make_geom_basis(Q, m.node_order);
    const Fourier8 fm = make_fourier(f.desc.coeff_seed);
    const int64_t n = m.ne * Q * Q * Q;
    hipLaunchKernelGGL(k_pa_setup, dim3(nblocks(n)), dim3(kBlock), 0, s, m.ne, gb, m.d_nodes.p,
                       f.desc.coeff_kind, f.desc.coeff_value, f.desc.d_coeff, fm, f.pa_diff.p,
                       f.pa_mass.p);
    H1PA_HIP(hipGetLastError());
}

void randomize(int64_t n, double *x, uint64_t seed, double lo, double hi, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_randomize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, seed,
                       lo, hi);
    H1PA_HIP(hipGetLastError());
}

void domain_lf(const Mesh &m, int p, double f, const double *fq, double *bE, hipStream_t s) {
    const int Q = p + 1, D = p + 1;
    const GeomBasis gb = make_geom_basis(Q, m.node_order);
    double x[16], w[16], B[256], G[256];
    gauss_legendre(Q, x, w);
    lagrange_basis(p, Q, x, B, G);
    StreamBuf<double> Bs;
    Bs.upload(B, Q * D, s);
    hipLaunchKernelGGL(k_domain_lf, dim3((unsigned)m.ne), dim3(256), 0, s, m.ne, D, gb, Bs.p,
                       m.d_nodes.p, f, fq, bE);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));
}

}  // namespace kern
}  // namespace h1pa
