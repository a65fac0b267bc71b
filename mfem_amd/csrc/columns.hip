// columns.hip -- the fused L->L PA diffusion apply on Cartesian meshes, by
// column sweeps: y = A x with no E-vector in HBM.
//
// MFEM's PABilinearFormExtension::Mult (fem/bilinearform_ext.cpp:487-560)
// gathers x into an E-vector, applies each element (SmemPADiffusionApply3D,
// fem/integ/bilininteg_diffusion_kernels.hpp:989-1221) into a second E-vector
// and scatters that with ElementRestriction::MultTranspose
// (fem/restriction.cpp:159-187).  The result is the same operator; here the
// element results are summed in LDS instead:
//
//  * a work-group owns a column of A x B elements in x, y and sweeps a
//    segment of its element layers along z, one layer at a time (the
//    element body is the same as the element kernels');
//  * after a layer, each point of the layer's (A P + 1) x (B P + 1) x P grid
//    of DOFs is the sum, in a fixed order, of the layer elements holding it,
//    plus (bottom plane) the carried top plane of the layer below.  A point no
//    other work-group touches is complete and written to y[L] at once; the
//    top plane is carried in LDS to the next layer;
//  * points on the column's sides (shared with the neighbouring columns) and
//    on the segment's end planes (shared with the neighbouring segments) are
//    written as partial sums to an interface buffer, side by side, and a
//    second pass adds each shared point's partials in a fixed order.
//
// Every sum has a fixed order, so the result is deterministic run to run; it
// differs from MFEM's summation order only by rounding.  The element order of
// the mesh does not matter (elements are found by grid position), so the
// L-vector numbering is MFEM's and the gather map is the space's.
#include "apply_body.hpp"

namespace h1pa {
namespace kern {

namespace {

// Interface slots of one work-group, one region per side: x-left and x-right
// columns [plane][y], y-front and y-back rows [plane][x], bottom and top
// planes [y][x].  A point on several sides goes to the first of them in this
// order; the writer and the second pass use the same rule.
struct Sides {
    int px, py, np;  // (A P + 1), (B P + 1), planes of the segment (ls P + 1)
    __device__ __forceinline__ int64_t size() const {
        return 2 * (int64_t)np * (px + py) + 2 * (int64_t)px * py;
    }
    // the slot of local point (xq, yq, lp) with the side flags of the point
    __device__ __forceinline__ int64_t slot(int xq, int yq, int lp, bool xl, bool xr, bool yf,
                                            bool yb, bool zb) const {
        const int64_t cx = (int64_t)np * py, cy = (int64_t)np * px;
        if (xl) return (int64_t)lp * py + yq;
        if (xr) return cx + (int64_t)lp * py + yq;
        if (yf) return 2 * cx + (int64_t)lp * px + xq;
        if (yb) return 2 * cx + cy + (int64_t)lp * px + xq;
        if (zb) return 2 * cx + 2 * cy + (int64_t)yq * px + xq;
        return 2 * cx + 2 * cy + (int64_t)px * py + (int64_t)yq * px + xq;
    }
};

template <int D, int A, int B>
__device__ __forceinline__ Sides sides_of(const ColumnGrid &g) {
    constexpr int P = D - 1;
    return Sides{A * P + 1, B * P + 1, g.ls * P + 1};
}

template <int D, int Q, int A, int B, int U, int K>
__global__ void __launch_bounds__(A * B * Q * Q)
k_diffusion_columns(const ColumnGrid g, const double *__restrict__ basis,
                    const double *__restrict__ pa, const int *__restrict__ gmap,
                    const double *__restrict__ x, double *__restrict__ y,
                    double *__restrict__ iface) {
    using T = double;
    constexpr int P = D - 1, NB = A * B, QQ = Q * Q, DDD = D * D * D, DDQ = D * D * Q;
    constexpr int BUF = diffusion_buf<D, Q>();
    constexpr int PX = A * P + 1, PY = B * P + 1;
    static_assert(2 * DDQ + DDD <= BUF, "element result does not fit beside the last stage");
    __shared__ T smem[NB * BUF];
    __shared__ int sg[NB * DDD];
    __shared__ T carry[2][PX * PY];
    const int wg = blockIdx.x;
    const int seg = wg % g.nseg, col = wg / g.nseg;
    const int cx = col % g.npx, cy = col / g.npx;
    const int av = min(A, g.nx - cx * A), bv = min(B, g.ny - cy * B);
    const int le = threadIdx.x / QQ, li = le % A, lj = le / A;
    const bool colv = li < av && lj < bv;
    const int gi = cx * A + li, gj = cy * B + lj;
    const int k0 = seg * g.ls, k1 = min(g.nz, k0 + g.ls);
    const int PXv = av * P + 1, PYv = bv * P + 1;
    const int NX = g.nx * P, NY = g.ny * P, NZ = g.nz * P;
    const int X0 = cx * A * P, Y0 = cy * B * P;
    const Sides sd = sides_of<D, A, B>(g);
    double *__restrict__ ifw = iface + (int64_t)wg * sd.size();
    for (int t = threadIdx.x; t < PX * PY; t += blockDim.x) carry[0][t] = 0.0;
    // the element body's first barrier orders these stores before any read
    const EOutT<T> none{};
    for (int k = k0; k < k1; ++k) {
        const int64_t e = colv ? g.elem_at[((int64_t)k * g.ny + gj) * g.nx + gi] : 0;
        diffusion_elements<T, T, D, Q, NB, U, K, 1>(e, colv ? 1 : 0, basis, pa, gmap, nullptr, x,
                                                    none, 0, smem, sg);
        __syncthreads();
        const int l = k - k0;
        const T *__restrict__ cin = carry[l & 1];
        T *__restrict__ cout = carry[(l + 1) & 1];
        const bool last = k == k1 - 1;
        const int npts = (P + 1) * PXv * PYv;
        for (int t = threadIdx.x; t < npts; t += blockDim.x) {
            const int xq = t % PXv, r = t / PXv, yq = r % PYv, zq = r / PYv;
            // this layer's elements holding the point: i in [ilo, ihi], j in [jlo, jhi]
            const int ih = xq / P, jh = yq / P;
            const int ilo = (xq % P == 0 && ih > 0) ? ih - 1 : ih, ihi = min(ih, av - 1);
            const int jlo = (yq % P == 0 && jh > 0) ? jh - 1 : jh, jhi = min(jh, bv - 1);
            T sum = 0.0;
            for (int j = jlo; j <= jhi; ++j)
                for (int i = ilo; i <= ihi; ++i)
                    sum += smem[(j * A + i) * BUF + 2 * DDQ + (xq - i * P) +
                                D * ((yq - j * P) + D * zq)];
            if (zq == P && !last) {
                cout[xq + PX * yq] = sum;
                continue;
            }
            if (zq == 0) sum = cin[xq + PX * yq] + sum;
            const int GX = X0 + xq, GY = Y0 + yq, GZ = k * P + zq;
            const bool xl = xq == 0 && GX > 0, xr = xq == PXv - 1 && GX < NX;
            const bool yf = yq == 0 && GY > 0, yb = yq == PYv - 1 && GY < NY;
            const bool zb = zq == 0 && k == k0 && GZ > 0, zt = zq == P && GZ < NZ;
            if (xl || xr || yf || yb || zb || zt) {
                ifw[sd.slot(xq, yq, l * P + zq, xl, xr, yf, yb, zb)] = sum;
            } else {
                const int L = sg[(jlo * A + ilo) * DDD + (xq - ilo * P) +
                                 D * ((yq - jlo * P) + D * zq)];
                y[L] = sum;
            }
        }
        __syncthreads();
    }
}

// The second pass: y[L] of every point shared between work-groups is the sum
// of its partials, taken in (segment, column y, column x) order.  Threads
// enumerate the interior column x-boundaries, then the y-boundaries (points
// not already on an x-boundary), then the segment planes (points on neither).
template <int D, int A, int B>
__global__ void k_columns_interface(const ColumnGrid g, const int *__restrict__ gmap,
                                    const double *__restrict__ iface, double *__restrict__ y,
                                    int64_t c1, int64_t c2, int64_t c3) {
    constexpr int P = D - 1, DDD = D * D * D;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int NX = g.nx * P, NY = g.ny * P, NZ = g.nz * P;
    const int WX = A * P, WY = B * P, WZ = g.ls * P;
    int GX, GY, GZ;
    if (t < c1) {  // x-boundary m = 1 .. npx-1; GY fastest
        int64_t r = t;
        GY = (int)(r % (NY + 1));
        r /= (NY + 1);
        GZ = (int)(r % (NZ + 1));
        GX = (int)(r / (NZ + 1) + 1) * WX;
    } else if (t < c1 + c2) {  // y-boundaries; GX fastest
        int64_t r = t - c1;
        GX = (int)(r % (NX + 1));
        r /= (NX + 1);
        GZ = (int)(r % (NZ + 1));
        GY = (int)(r / (NZ + 1) + 1) * WY;
        if (GX % WX == 0 && GX > 0 && GX < NX) return;
    } else if (t < c1 + c2 + c3) {  // segment planes; GX fastest
        int64_t r = t - c1 - c2;
        GX = (int)(r % (NX + 1));
        r /= (NX + 1);
        GY = (int)(r % (NY + 1));
        GZ = (int)(r / (NY + 1) + 1) * WZ;
        if ((GX % WX == 0 && GX > 0 && GX < NX) || (GY % WY == 0 && GY > 0 && GY < NY)) return;
    } else {
        return;
    }
    const int cxh = min(GX / WX, g.npx - 1), cyh = min(GY / WY, g.npy - 1);
    const int sh = min(GZ / WZ, g.nseg - 1);
    const int cxl = (GX % WX == 0 && GX > 0 && GX < NX) ? cxh - 1 : cxh;
    const int cyl = (GY % WY == 0 && GY > 0 && GY < NY) ? cyh - 1 : cyh;
    const int sl = (GZ % WZ == 0 && GZ > 0 && GZ < NZ) ? sh - 1 : sh;
    const Sides sd = sides_of<D, A, B>(g);
    const int64_t S = sd.size();
    double sum = 0.0;
    for (int s = sl; s <= sh; ++s)
        for (int cy = cyl; cy <= cyh; ++cy)
            for (int cx = cxl; cx <= cxh; ++cx) {
                const int64_t w = ((int64_t)cy * g.npx + cx) * g.nseg + s;
                const int av = min(A, g.nx - cx * A), bv = min(B, g.ny - cy * B);
                const int k0 = s * g.ls;
                const int xq = GX - cx * WX, yq = GY - cy * WY, lp = GZ - k0 * P;
                const bool xl = xq == 0 && GX > 0, xr = xq == av * P && GX < NX;
                const bool yf = yq == 0 && GY > 0, yb = yq == bv * P && GY < NY;
                const bool zb = lp == 0 && GZ > 0;
                sum += iface[w * S + sd.slot(xq, yq, lp, xl, xr, yf, yb, zb)];
            }
    const int i = min(GX / P, g.nx - 1), j = min(GY / P, g.ny - 1), kk = min(GZ / P, g.nz - 1);
    const int64_t e = g.elem_at[((int64_t)kk * g.ny + j) * g.nx + i];
    y[gmap[e * DDD + (GX - i * P) + D * ((GY - j * P) + D * (GZ - kk * P))]] = sum;
}

// Column cross-section (A x B elements) per order: about 256 threads (Q^2 per
// element) within the element body's LDS.
template <int D>
struct ColumnShape;
template <> struct ColumnShape<2> { static constexpr int A = 4, B = 7; };
template <> struct ColumnShape<3> { static constexpr int A = 4, B = 4; };
template <> struct ColumnShape<4> { static constexpr int A = 2, B = 5; };
template <> struct ColumnShape<5> { static constexpr int A = 2, B = 4; };
template <> struct ColumnShape<6> { static constexpr int A = 2, B = 3; };
template <> struct ColumnShape<7> { static constexpr int A = 2, B = 2; };
template <> struct ColumnShape<8> { static constexpr int A = 1, B = 3; };
template <> struct ColumnShape<9> { static constexpr int A = 1, B = 2; };

// parts: 1 the column sweep, 2 the interface pass, 3 both (in that order).
template <int D, int Q>
void launch_columns(const ColumnGrid &g, const double *basis, const double *pa, const int *gmap,
                    const double *x, double *y, double *iface, int parts, hipStream_t s) {
    constexpr int A = ColumnShape<D>::A, B = ColumnShape<D>::B;
    const unsigned nwg = (unsigned)(g.npx * g.npy * g.nseg);
    // K = 2 PA planes ahead (the pipeline of the element kernels' variant 4)
    if (parts & 1)
        hipLaunchKernelGGL((k_diffusion_columns<D, Q, A, B, 1, 2>), dim3(nwg),
                           dim3(A * B * Q * Q), 0, s, g, basis, pa, gmap, x, y, iface);
    if (!(parts & 2)) return;
    constexpr int P = D - 1;
    const int64_t NX = (int64_t)g.nx * P, NY = (int64_t)g.ny * P, NZ = (int64_t)g.nz * P;
    const int64_t c1 = (int64_t)(g.npx - 1) * (NY + 1) * (NZ + 1);
    const int64_t c2 = (int64_t)(g.npy - 1) * (NX + 1) * (NZ + 1);
    const int64_t c3 = (int64_t)(g.nseg - 1) * (NX + 1) * (NY + 1);
    const int64_t n = c1 + c2 + c3;
    if (n > 0)
        hipLaunchKernelGGL((k_columns_interface<D, A, B>), dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, s, g, gmap, iface, y, c1, c2, c3);
}

}  // namespace

bool columns_supported(int p, int q1d) { return p >= 1 && p <= 8 && q1d == p + 2; }

void column_shape(int p, int &A, int &B) {
    switch (p + 1) {
#define X(DD) \
    case DD: A = ColumnShape<DD>::A; B = ColumnShape<DD>::B; return;
        X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9)
#undef X
        default: A = B = 0;
    }
}

int64_t column_interface_size(int p, const ColumnGrid &g) {
    int A = 0, B = 0;
    column_shape(p, A, B);
    const int64_t P = p, px = A * P + 1, py = B * P + 1, np = g.ls * P + 1;
    return (int64_t)g.npx * g.npy * g.nseg * (2 * np * (px + py) + 2 * px * py);
}

void diffusion_columns(int p, const ColumnGrid &g, const double *basis, const double *pa,
                       const int *gmap, const double *x, double *y, double *iface, int parts,
                       hipStream_t s) {
    switch (p + 1) {
#define X(DD)                                                                   \
    case DD:                                                                    \
        launch_columns<DD, DD + 1>(g, basis, pa, gmap, x, y, iface, parts, s);  \
        break;
        X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9)
#undef X
        default: throw Error{H1PA_EUNSUP, "column apply: unsupported order"};
    }
    H1PA_HIP(hipGetLastError());
}

}  // namespace kern
}  // namespace h1pa
