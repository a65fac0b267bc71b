// basis.cpp -- 1D Gauss-Legendre / Gauss-Lobatto rules, the GLL Lagrange basis
// and the lexicographic dof_map of the H1 hexahedron.
#include <cmath>

#include "internal.hpp"

namespace h1pa {

namespace {
constexpr long double kPi = 3.14159265358979323846264338327950288L;

// P_n(x) and P_n'(x) and P_{n-1}(x) by the three-term recurrence.
void legendre(int n, long double x, long double &pn, long double &dpn, long double &pnm1) {
    long double a = 1.0L, b = x;
    if (n == 0) { pn = 1.0L; dpn = 0.0L; pnm1 = 0.0L; return; }
    for (int j = 2; j <= n; ++j) {
        const long double c = ((2 * j - 1) * x * b - (j - 1) * a) / j;
        a = b;
        b = c;
    }
    pn = b;
    pnm1 = a;
    dpn = n * (x * b - a) / (x * x - 1.0L);
}
}  // namespace

// Gauss-Legendre points/weights on [0,1], ascending
// (QuadratureFunctions1D::GaussLegendre, fem/intrules.cpp:620-705).
void gauss_legendre(int n, double *x, double *w) {
    if (n == 1) { x[0] = 0.5; w[0] = 1.0; return; }
    for (int i = 1; i <= (n + 1) / 2; ++i) {
        long double z = std::cos(kPi * (i - 0.25L) / (n + 0.5L));
        long double pn, dpn, pm;
        for (int it = 0; it < 64; ++it) {
            legendre(n, z, pn, dpn, pm);
            const long double dz = pn / dpn;
            z -= dz;
            if (std::fabs(dz) < 1e-19L) break;
        }
        legendre(n, z, pn, dpn, pm);
        const long double xi = (1.0L - z) / 2.0L;
        x[i - 1] = static_cast<double>(xi);
        x[n - i] = static_cast<double>(1.0L - xi);
        w[i - 1] = w[n - i] = static_cast<double>(1.0L / ((1.0L - z * z) * dpn * dpn));
    }
}

// Gauss-Lobatto points/weights on [0,1] (fem/intrules.cpp:708-830).
void gauss_lobatto(int n, double *x, double *w) {
    if (n == 1) { x[0] = 0.5; w[0] = 1.0; return; }
    const int N = n - 1;
    x[0] = 0.0;
    x[N] = 1.0;
    w[0] = w[N] = 1.0 / (N * (N + 1.0));
    for (int i = 1; i <= N / 2; ++i) {
        long double z = -std::cos(kPi * i / N);
        long double pn, dpn, pm;
        for (int it = 0; it < 64; ++it) {
            legendre(N, z, pn, dpn, pm);
            const long double d2 = (2.0L * z * dpn - N * (N + 1.0L) * pn) / (1.0L - z * z);
            const long double dz = dpn / d2;
            z -= dz;
            if (std::fabs(dz) < 1e-19L) break;
        }
        legendre(N, z, pn, dpn, pm);
        x[i] = static_cast<double>((1.0L + z) / 2.0L);
        x[N - i] = static_cast<double>((1.0L - z) / 2.0L);
        w[i] = w[N - i] = static_cast<double>(1.0L / (N * (N + 1.0L) * pn * pn));
    }
    if (N % 2 == 0) x[N / 2] = 0.5;
}

// Lagrange basis on the p+1 GLL nodes evaluated at xq (Poly_1D::Basis,
// fem/fe/fe_base.cpp:1811-2010): B[q + n*d] = l_d(x_q), G = l_d'(x_q).
void lagrange_basis(int p, int n, const double *xq, double *B, double *G) {
    double z[16], wz[16];
    gauss_lobatto(p + 1, z, wz);
    for (int d = 0; d <= p; ++d) {
        long double denom = 1.0L;
        for (int k = 0; k <= p; ++k)
            if (k != d) denom *= (long double)z[d] - z[k];
        for (int q = 0; q < n; ++q) {
            const long double x = xq[q];
            long double prod = 1.0L, dsum = 0.0L;
            // l_d(x) = prod_{k!=d}(x - z_k) / denom; l_d' = sum_m prod_{k!=d,m}(x - z_k) / denom
            for (int k = 0; k <= p; ++k) {
                if (k == d) continue;
                dsum = dsum * (x - z[k]) + prod;
                prod *= (x - z[k]);
            }
            B[q + n * d] = static_cast<double>(prod / denom);
            G[q + n * d] = static_cast<double>(dsum / denom);
        }
    }
}

// TensorBasisElement H1 dof_map, 3D (fem/fe/fe_base.cpp:2557-2674):
// dof_map[lexicographic index] = native (vertex, edge, face, interior) index.
void h1_dof_map(int p, int *dm) {
    const int p1 = p + 1;
    auto L = [p1](int i, int j, int k) { return i + p1 * (j + p1 * k); };
    const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int v = 0; v < 8; ++v) dm[L(corner[v][0] * p, corner[v][1] * p, corner[v][2] * p)] = v;
    int o = 8;
    // 12 edges in Geometry::Constants<CUBE>::Edges order; each runs along +axis
    // from its first vertex: (axis, fixed coordinates of the other two axes).
    const int edge_axis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
    const int edge_fix[12][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0},
                                 {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {0, 0, 1},
                                 {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
    for (int e = 0; e < 12; ++e)
        for (int i = 1; i < p; ++i) {
            int c[3] = {edge_fix[e][0] * p, edge_fix[e][1] * p, edge_fix[e][2] * p};
            c[edge_axis[e]] = i;
            dm[L(c[0], c[1], c[2])] = o++;
        }
    // 6 faces in FaceVert order; (i, j) are the face-local loop indices.
    for (int f = 0; f < 6; ++f)
        for (int j = 1; j < p; ++j)
            for (int i = 1; i < p; ++i) {
                int l;
                switch (f) {
                    case 0: l = L(i, p - j, 0); break;
                    case 1: l = L(i, 0, j); break;
                    case 2: l = L(p, i, j); break;
                    case 3: l = L(p - i, p, j); break;
                    case 4: l = L(0, p - i, j); break;
                    default: l = L(i, j, p); break;
                }
                dm[l] = o++;
            }
    for (int k = 1; k < p; ++k)
        for (int j = 1; j < p; ++j)
            for (int i = 1; i < p; ++i) dm[L(i, j, k)] = o++;
}

// ---- even-odd tables ---------------------------------------------------------
// For M in (B, G), M(q, d) = M[q + Q d]:
//   forward, R rows q < R = Q/2 + Q%2:  Me(q, 0..HD) | Mo(q, 0..HD) | Mm(q)
//   backward, rows d < HD:  Me(0..R, d) | Mo(0..R, d);  row HD (D odd): Mm(0..R) | 0
// with Me = (M(q,d) + M(q,D-1-d)) / 2, Mo = (M(q,d) - M(q,D-1-d)) / 2,
// Mm(q) = M(q, D/2); HD = D/2.
int even_odd_size(int D, int Q) {
    const int HD = D / 2, OD = D % 2, R = Q / 2 + Q % 2, W = 2 * HD + OD;
    return 2 * R * W + 2 * (HD + OD) * 2 * R;
}

int even_odd_tables(int D, int Q, const double *B, const double *G, double *out) {
    const int HD = D / 2, OD = D % 2, R = Q / 2 + Q % 2, W = 2 * HD + OD, RB = 2 * R;
    double *o = out;
    const double *Ms[2] = {B, G};
    for (const double *M : Ms) {  // forward rows
        for (int q = 0; q < R; ++q) {
            double *row = o + W * q;
            for (int d = 0; d < HD; ++d) {
                const double m0 = M[q + Q * d], m1 = M[q + Q * (D - 1 - d)];
                row[d] = 0.5 * (m0 + m1);
                row[HD + d] = 0.5 * (m0 - m1);
            }
            if (OD) row[2 * HD] = M[q + Q * HD];
        }
        o += R * W;
    }
    for (const double *M : Ms) {  // backward rows
        for (int d = 0; d < HD + OD; ++d) {
            double *row = o + RB * d;
            for (int q = 0; q < R; ++q) {
                if (d < HD) {
                    const double m0 = M[q + Q * d], m1 = M[q + Q * (D - 1 - d)];
                    row[q] = 0.5 * (m0 + m1);
                    row[R + q] = 0.5 * (m0 - m1);
                } else {
                    row[q] = M[q + Q * HD];
                    row[R + q] = 0.0;
                }
            }
        }
        o += (HD + OD) * RB;
    }
    return (int)(o - out);
}

}  // namespace h1pa
