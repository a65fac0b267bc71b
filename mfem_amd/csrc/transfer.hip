// transfer.hip -- p-refinement transfer between two H1 spaces on one mesh
// (TensorProductPRefinementTransferOperator, fem/transfer.cpp:2619-2990).
//
// MFEM gathers the low-order L-vector into an E-vector, interpolates each
// element's values to the high-order nodes (B = low-order basis at the
// high-order GLL nodes, three 1D contractions), multiplies by the high
// restriction's BooleanMask (1 at the first E-entry of each L-DOF) and sums
// the E-vector into the high L-vector; the transpose runs the same steps
// backwards.  Here the mask is a byte per E-entry, and since exactly one
// masked entry of each high L-DOF is nonzero, the prolongation writes that
// entry's value straight to the L-vector (no high E-vector, no sum: the sum
// of one value and zeros is that value).  The restriction gathers the masked
// high values, contracts them with B^T into the low E-vector (lexicographic)
// and the low restriction's deterministic E->L sum finishes it.
//
// The contractions keep MFEM's association (x first, then y, then z, each an
// ascending sum), so results differ from MFEM's only by fused multiply-add
// rounding.
#include "internal.hpp"

namespace h1pa {
namespace kern {

namespace {

constexpr int kThreads = 128;
constexpr int kMaxD = 9;

// emask[E] = 1 when E-index E is the first entry (indices order) of its
// L-DOF's row: ElementRestriction::BooleanMask (fem/restriction.cpp:255-290)
// as one byte per entry.
__global__ void k_first_entry_mask(int64_t nE, const int *__restrict__ gmap,
                                   const int *__restrict__ offsets,
                                   const int *__restrict__ indices, uint8_t *__restrict__ emask) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nE) return;
    const int g = gmap[i];
    const int row = g >= 0 ? g : -1 - g;
    const int first = indices[offsets[row]];
    emask[i] = (first >= 0 ? first : -1 - first) == i ? 1 : 0;
}

// One work-group per element: low E-values (gathered) -> high node values.
__global__ void __launch_bounds__(kThreads)
k_prolong(int64_t ne, int Dl, int Dh, const double *__restrict__ B, const int *__restrict__ gmap_l,
          const int *__restrict__ gmap_h, const uint8_t *__restrict__ emask_h,
          const double *__restrict__ xl, double *__restrict__ yh) {
    __shared__ double u[kMaxD * kMaxD * kMaxD];   // low values      [dz][dy][dx]
    __shared__ double t1[kMaxD * kMaxD * kMaxD];  // x contracted    [dz][dy][qx]
    __shared__ double t2[kMaxD * kMaxD * kMaxD];  // x, y contracted [dz][qy][qx]
    __shared__ double Bs[kMaxD * kMaxD];          // B[qh + Dh dl]
    const int64_t e = blockIdx.x;
    if (e >= ne) return;
    const int nl = Dl * Dl * Dl, nh = Dh * Dh * Dh;
    for (int i = threadIdx.x; i < Dh * Dl; i += blockDim.x) Bs[i] = B[i];
    for (int i = threadIdx.x; i < nl; i += blockDim.x) {
        const int g = gmap_l[e * nl + i];
        u[i] = g >= 0 ? xl[g] : -xl[-1 - g];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Dh * Dl * Dl; i += blockDim.x) {
        const int qx = i % Dh, r = i / Dh;  // r = dy + Dl dz
        double v = 0.0;
        for (int dx = 0; dx < Dl; ++dx) v += Bs[qx + Dh * dx] * u[dx + Dl * r];
        t1[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Dh * Dh * Dl; i += blockDim.x) {
        const int qx = i % Dh, qy = (i / Dh) % Dh, dz = i / (Dh * Dh);
        double v = 0.0;
        for (int dy = 0; dy < Dl; ++dy) v += Bs[qy + Dh * dy] * t1[qx + Dh * (dy + Dl * dz)];
        t2[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nh; i += blockDim.x) {
        const int64_t E = e * nh + i;
        if (!emask_h[E]) continue;
        const int qxy = i % (Dh * Dh), qz = i / (Dh * Dh);
        double v = 0.0;
        for (int dz = 0; dz < Dl; ++dz) v += Bs[qz + Dh * dz] * t2[qxy + Dh * Dh * dz];
        const int g = gmap_h[E];
        if (g >= 0) yh[g] = v;
        else yh[-1 - g] = -v;
    }
}

// One work-group per element: masked high values (gathered) -> low E-vector.
__global__ void __launch_bounds__(kThreads)
k_restrict(int64_t ne, int Dl, int Dh, const double *__restrict__ B, const int *__restrict__ gmap_h,
           const uint8_t *__restrict__ emask_h, const double *__restrict__ xh,
           double *__restrict__ yle) {
    __shared__ double u[kMaxD * kMaxD * kMaxD];   // masked high values [qz][qy][qx]
    __shared__ double t1[kMaxD * kMaxD * kMaxD];  // x contracted       [qz][qy][dx]
    __shared__ double t2[kMaxD * kMaxD * kMaxD];  // x, y contracted    [qz][dy][dx]
    __shared__ double Bs[kMaxD * kMaxD];
    const int64_t e = blockIdx.x;
    if (e >= ne) return;
    const int nl = Dl * Dl * Dl, nh = Dh * Dh * Dh;
    for (int i = threadIdx.x; i < Dh * Dl; i += blockDim.x) Bs[i] = B[i];
    for (int i = threadIdx.x; i < nh; i += blockDim.x) {
        const int64_t E = e * nh + i;
        const int g = gmap_h[E];
        const double v = g >= 0 ? xh[g] : -xh[-1 - g];
        u[i] = emask_h[E] ? v : 0.0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Dl * Dh * Dh; i += blockDim.x) {
        const int dx = i % Dl, r = i / Dl;  // r = qy + Dh qz
        double v = 0.0;
        for (int qx = 0; qx < Dh; ++qx) v += Bs[qx + Dh * dx] * u[qx + Dh * r];
        t1[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Dl * Dl * Dh; i += blockDim.x) {
        const int dx = i % Dl, dy = (i / Dl) % Dl, qz = i / (Dl * Dl);
        double v = 0.0;
        for (int qy = 0; qy < Dh; ++qy) v += Bs[qy + Dh * dy] * t1[dx + Dl * (qy + Dh * qz)];
        t2[i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nl; i += blockDim.x) {
        const int dxy = i % (Dl * Dl), dz = i / (Dl * Dl);
        double v = 0.0;
        for (int qz = 0; qz < Dh; ++qz) v += Bs[qz + Dh * dz] * t2[dxy + Dl * Dl * qz];
        yle[e * nl + i] = v;
    }
}

}  // namespace

void first_entry_mask(int64_t nE, const int *gmap, const int *offsets, const int *indices,
                      uint8_t *emask, hipStream_t s) {
    if (nE <= 0) return;
    hipLaunchKernelGGL(k_first_entry_mask, dim3((unsigned)((nE + 255) / 256)), dim3(256), 0, s, nE,
                       gmap, offsets, indices, emask);
    H1PA_HIP(hipGetLastError());
}

void transfer_prolong(int64_t ne, int pl, int ph, const double *B, const int *gmap_l,
                      const int *gmap_h, const uint8_t *emask_h, const double *xl, double *yh,
                      hipStream_t s) {
    H1PA_CHECK(pl >= 1 && ph <= kMaxD - 1 && pl < ph, H1PA_EUNSUP, "transfer: orders");
    if (ne <= 0) return;
    hipLaunchKernelGGL(k_prolong, dim3((unsigned)ne), dim3(kThreads), 0, s, ne, pl + 1, ph + 1, B,
                       gmap_l, gmap_h, emask_h, xl, yh);
    H1PA_HIP(hipGetLastError());
}

void transfer_restrict_e(int64_t ne, int pl, int ph, const double *B, const int *gmap_h,
                         const uint8_t *emask_h, const double *xh, double *yle, hipStream_t s) {
    H1PA_CHECK(pl >= 1 && ph <= kMaxD - 1 && pl < ph, H1PA_EUNSUP, "transfer: orders");
    if (ne <= 0) return;
    hipLaunchKernelGGL(k_restrict, dim3((unsigned)ne), dim3(kThreads), 0, s, ne, pl + 1, ph + 1, B,
                       gmap_h, emask_h, xh, yle);
    H1PA_HIP(hipGetLastError());
}

}  // namespace kern
}  // namespace h1pa
