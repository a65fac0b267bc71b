// restriction.hip -- ElementRestriction on the device.
//
//  * gather map: one thread per E-DOF evaluates MFEM's element DOF numbering
//    (vertex / edge / face / bubble blocks, orientation permutations) straight
//    from the compact element topology, already permuted to lexicographic
//    order -- FiniteElementSpace::GetElementDofs (fem/fespace.cpp:3428-3535)
//    fused with the dof_map step of the ElementRestriction constructor
//    (fem/restriction.cpp:60-98).
//  * transpose table: integer count -> exclusive scan -> slot fill -> per-row
//    sort.  The slot fill uses integer atomics, the per-row sort makes the
//    result independent of their order, so offsets/indices come out
//    bit-identical to MFEM's sequential build (restriction.cpp:73-113).
//  * E->L: one thread per L-DOF sums its E-entries in table order -- no
//    floating-point atomics, deterministic (restriction.cpp:159-228).
#include "internal.hpp"
#include "row_sum.hpp"


namespace h1pa {
namespace kern {

namespace {

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

__device__ inline int quad_dof_ord(int ori, int o, int pm1) {
    // H1_FECollection::QuadDofOrd (fem/fe_coll.cpp:1920-1936)
    const int pm2 = pm1 - 1, i = o % pm1, j = o / pm1;
    switch (ori & 7) {
        case 0: return i + j * pm1;
        case 1: return j + i * pm1;
        case 2: return j + (pm2 - i) * pm1;
        case 3: return (pm2 - i) + j * pm1;
        case 4: return (pm2 - i) + (pm2 - j) * pm1;
        case 5: return (pm2 - j) + (pm2 - i) * pm1;
        case 6: return (pm2 - j) + i * pm1;
        default: return i + (pm2 - j) * pm1;
    }
}

__global__ void k_build_gather(int64_t ne, int p, int dof, const int *__restrict__ topo,
                               const int8_t *__restrict__ ori, const int *__restrict__ dof_map,
                               int64_t nvdofs, int64_t nedofs, int64_t nfdofs,
                               int *__restrict__ elem_dof, int *__restrict__ gather_map) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne * dof) return;
    const int64_t e = i / dof;
    const int lex = (int)(i - e * dof);
    const int n = dof_map[lex];
    const int pm1 = p - 1, nf = pm1 * pm1;
    const int *t = topo + 29 * e;
    const int8_t *o = ori + 18 * e;
    int64_t g;
    if (n < 8) {
        g = t[n];
    } else if (n < 8 + 12 * pm1) {
        const int k = n - 8, ed = k / pm1, j = k % pm1;
        g = nvdofs + (int64_t)t[8 + ed] * pm1 + (o[ed] > 0 ? j : pm1 - 1 - j);
    } else if (n < 8 + 12 * pm1 + 6 * nf) {
        const int k = n - 8 - 12 * pm1, f = k / nf, j = k % nf;
        g = nvdofs + nedofs + (int64_t)t[20 + f] * nf + quad_dof_ord(o[12 + f], j, pm1);
    } else {
        g = nvdofs + nedofs + nfdofs + e * (int64_t)nf * pm1 + (n - (8 + 12 * pm1 + 6 * nf));
    }
    gather_map[i] = (int)g;
    elem_dof[e * dof + n] = (int)g;
}

__global__ void k_gather_from_table(int64_t ne, int dof, const int *__restrict__ elem_dof,
                                    const int *__restrict__ dof_map, int *__restrict__ gmap) {
    // restriction.cpp:91-110 (signed entries: -1-i encodes a sign flip)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne * dof) return;
    const int64_t e = i / dof;
    const int d = (int)(i - e * dof);
    const int sdid = dof_map ? dof_map[d] : 0;
    const int did = dof_map ? (sdid >= 0 ? sdid : -1 - sdid) : d;
    const int sgid = elem_dof[dof * e + did];
    const int gid = sgid >= 0 ? sgid : -1 - sgid;
    const bool plus = (sgid >= 0 && sdid >= 0) || (sgid < 0 && sdid < 0);
    gmap[i] = plus ? gid : -1 - gid;
}

__global__ void k_count(int64_t n, const int *__restrict__ gmap, int *__restrict__ cnt) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = gmap[i];
    atomicAdd(&cnt[(g >= 0 ? g : -1 - g) + 1], 1);
}

__global__ void k_fill(int64_t n, const int *__restrict__ gmap, int *__restrict__ cursor,
                       int *__restrict__ indices) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = gmap[i];
    const bool plus = g >= 0;
    const int gid = plus ? g : -1 - g;
    const int pos = atomicAdd(&cursor[gid], 1);
    indices[pos] = plus ? (int)i : -1 - (int)i;
}

__global__ void k_sort_rows(int64_t ndofs, const int *__restrict__ offsets, int *__restrict__ idx) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ndofs) return;
    const int b = offsets[i], e = offsets[i + 1];
    auto key = [](int v) { return v >= 0 ? v : -1 - v; };
    for (int a = b + 1; a < e; ++a) {
        const int v = idx[a], kv = key(v);
        int c = a - 1;
        while (c >= b && key(idx[c]) > kv) {
            idx[c + 1] = idx[c];
            --c;
        }
        idx[c + 1] = v;
    }
}

// ---- inclusive scan of int32 (three passes) --------------------------------
constexpr int kScanThreads = 256, kScanItems = 16, kScanTile = kScanThreads * kScanItems;

__device__ inline int block_inclusive_scan(int v, int *sh) {
    // Hillis-Steele over 256 threads in LDS.
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    return sh[t];
}

__global__ void k_scan_tiles(int *__restrict__ a, int64_t n, int *__restrict__ tile_sums) {
    __shared__ int sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int loc[kScanItems];
    int s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        s += (i < n) ? a[i] : 0;
        loc[k] = s;
    }
    const int incl = block_inclusive_scan(s, sh);
    const int excl = incl - s;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        if (i < n) a[i] = loc[k] + excl;
    }
    if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = incl;
}

__global__ void k_scan_sums(int *__restrict__ sums, int64_t n) {
    // one block: sequential carry across chunks of 256
    __shared__ int sh[kScanThreads];
    int carry = 0;
    for (int64_t c = 0; c < n; c += kScanThreads) {
        const int64_t i = c + threadIdx.x;
        const int v = i < n ? sums[i] : 0;
        const int incl = block_inclusive_scan(v, sh);
        if (i < n) sums[i] = incl + carry;
        __syncthreads();
        carry += sh[kScanThreads - 1];
        __syncthreads();
    }
}

__global__ void k_scan_add(int *__restrict__ a, int64_t n, const int *__restrict__ tile_sums) {
    if (blockIdx.x == 0) return;
    const int add = tile_sums[blockIdx.x - 1];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
        const int64_t i = base + k;
        if (i < n) a[i] += add;
    }
}

void inclusive_scan(int *a, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    StreamBuf<int> sums(tiles, s);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, a, n, sums.p);
    if (tiles > 1) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, sums.p, tiles);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, a, n, sums.p);
    }
    H1PA_HIP(hipGetLastError());
}

// DOFs on the element's boundary faces (topo[26]: bit j set when local face j
// -- z = 0, y = 0, x = p, y = p, x = 0, z = p in Geometry::Constants<CUBE>
// order -- is a boundary face).
__global__ void k_boundary_mark(int64_t ne, int p, const int *__restrict__ topo,
                                const int *__restrict__ gmap, uint8_t *__restrict__ mask) {
    const int D = p + 1, dof = D * D * D;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= ne * dof) return;
    const int64_t e = i / dof;
    const int lex = (int)(i - e * dof);
    const int a = lex % D, b = (lex / D) % D, c = lex / (D * D);
    const int f = topo[29 * e + 26];
    const bool on = ((f & 1) && c == 0) || ((f & 2) && b == 0) || ((f & 4) && a == p) ||
                    ((f & 8) && b == p) || ((f & 16) && a == 0) || ((f & 32) && c == p);
    if (on) {
        const int g = gmap[i];
        mask[g >= 0 ? g : -1 - g] = 1;
    }
}

__global__ void k_mask_to_int(int64_t n, const uint8_t *__restrict__ m, int *__restrict__ o) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) o[i] = m[i] ? 1 : 0;
}

__global__ void k_compact(int64_t n, const uint8_t *__restrict__ m, const int *__restrict__ incl,
                          int *__restrict__ list) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && m[i]) list[incl[i] - 1] = (int)i;
}

__global__ void k_list_to_mask(int64_t nl, const int *__restrict__ list, uint8_t *__restrict__ m,
                               int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const int j = list[i];
    if (j >= 0 && j < n) m[j] = 1;  // out-of-range entries are ignored
}

// ---- L <-> E ---------------------------------------------------------------
// One E-entry per thread (any alignment).
__global__ void k_restrict_mult_1(int64_t nE, const int *__restrict__ gmap,
                                  const double *__restrict__ x, double *__restrict__ y, bool abs_) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nE) return;
    const int g = gmap[i];
    const bool plus = g >= 0;
    const double v = x[plus ? g : -1 - g];
    y[i] = (plus || abs_) ? v : -v;
}

// P consecutive E-entries per thread (P a multiple of 4; 4 is used: 8 was
// measured 20-30% slower): 16-byte index loads, P gathers issued together,
// 16-byte stores (the tail thread goes entry by entry).  Needs 16-byte
// aligned gmap and y.  NT: the index loads and E-vector stores, each touched
// once, go non-temporal.
typedef int gv4i __attribute__((ext_vector_type(4)));
typedef double gv2d __attribute__((ext_vector_type(2)));
template <int P, bool NT>
__global__ void __launch_bounds__(kBlock)
k_restrict_mult(int64_t nE, const int *__restrict__ gmap, const double *__restrict__ x,
                double *__restrict__ y, bool abs_) {
    const int64_t i0 = P * ((int64_t)blockIdx.x * kBlock + threadIdx.x);
    if (i0 >= nE) return;
    if (i0 + P <= nE) {
        int gg[P];
#pragma unroll
        for (int k = 0; k < P; k += 4) {
            const gv4i *src = reinterpret_cast<const gv4i *>(gmap + i0 + k);
            const gv4i g = NT ? __builtin_nontemporal_load(src) : *src;
            gg[k] = g.x;
            gg[k + 1] = g.y;
            gg[k + 2] = g.z;
            gg[k + 3] = g.w;
        }
        double v[P];
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = x[gg[k] >= 0 ? gg[k] : -1 - gg[k]];
#pragma unroll
        for (int k = 0; k < P; ++k) v[k] = (gg[k] >= 0 || abs_) ? v[k] : -v[k];
#pragma unroll
        for (int k = 0; k < P; k += 2) {
            gv2d *dst = reinterpret_cast<gv2d *>(y + i0 + k);
            const gv2d w = {v[k], v[k + 1]};
            if (NT)
                __builtin_nontemporal_store(w, dst);
            else
                *dst = w;
        }
        return;
    }
    for (int64_t i = i0; i < nE; ++i) {
        const int g = gmap[i];
        const bool plus = g >= 0;
        const double v = x[plus ? g : -1 - g];
        y[i] = (plus || abs_) ? v : -v;
    }
}

// E -> L: R rows per thread (rows blk R kBlock + r kBlock + tid), every
// table and value load of the R rows issued before the first sum.
template <int R>
__global__ void __launch_bounds__(kBlock)
k_restrict_mult_t(int64_t nrows, const int *__restrict__ offsets, const int *__restrict__ indices,
                  const double *__restrict__ xe, double *__restrict__ y, int mode,
                  const int *__restrict__ perm, int64_t nperm) {
    const int64_t t0 = (int64_t)blockIdx.x * (R * kBlock) + threadIdx.x;
    int64_t ii[R];
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t t = t0 + r * kBlock;
        ok[r] = t < nrows;
        const int64_t tl = ok[r] ? t : nrows - 1;
        ii[r] = tl < nperm ? perm[tl] : tl;
    }
    int b[R], n[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        b[r] = offsets[ii[r]];
        n[r] = offsets[ii[r] + 1] - b[r];
    }
    double v[R];
    if (mode == 3) {  // left inverse: the row's last entry
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int ix = n[r] > 0 ? indices[b[r] + n[r] - 1] : 0;
            v[r] = n[r] > 0 ? (ix >= 0 ? xe[ix] : -xe[-1 - ix]) : 0.0;
        }
    } else {
        rows_sum<R>(indices, b, n, xe, v, mode == 2);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!ok[r]) continue;
        if (mode == 1) y[ii[r]] += v[r];
        else y[ii[r]] = v[r];
    }
}

// E -> L for MFEM's bubble block of an H1 space (interior DOFs, one E-entry
// each, fem/fespace.cpp:3518-3527): row bb + e NI^3 + j holds the interior
// DOF j (x fastest) of element e, at lexicographic position (jx+1, jy+1,
// jz+1) of the element's E-entries -- no table is read.  The sum of the row
// is MFEM's (0 + x; the left inverse takes x itself).
__global__ void __launch_bounds__(kBlock)
k_bubble_mult_t(int64_t nrows, int NI, int D, int64_t bb, const double *__restrict__ xe,
                double *__restrict__ y, int mode) {
    const int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (r >= nrows) return;
    const int ni3 = NI * NI * NI;
    const int64_t e = r / ni3;
    const int j = (int)(r - e * ni3);
    const int jx = j % NI, jy = (j / NI) % NI, jz = j / (NI * NI);
    const double v = xe[e * (int64_t)(D * D * D) + (jx + 1) + D * ((jy + 1) + D * (jz + 1))];
    double t = 0.0;
    t = t + v;
    if (mode == 3) y[bb + r] = v;
    else if (mode == 1) y[bb + r] += t;
    else y[bb + r] = t;
}

// ElementRestriction::BooleanMask (fem/restriction.cpp:255-290): the first
// E-entry of each L-DOF's row (in indices order) is 1, the others 0.  A signed
// entry -1-k stands for E-index k (MFEM's loop indexes with the raw value,
// which its H1 tables never make negative).
__global__ void k_boolean_mask(int64_t nrows, const int *__restrict__ offsets,
                               const int *__restrict__ indices, double *__restrict__ ye) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int b = offsets[i], e = offsets[i + 1];
    for (int j = b; j < e; ++j) {
        const int ix = indices[j];
        ye[ix >= 0 ? ix : -1 - ix] = j == b ? 1.0 : 0.0;
    }
}

__global__ void k_scatter_constrained(int64_t nrows, const int *__restrict__ offsets,
                                      const int *__restrict__ indices,
                                      const double *__restrict__ xe,
                                      const uint8_t *__restrict__ mask,
                                      const double *__restrict__ xin, double *__restrict__ y,
                                      const int *__restrict__ perm, int64_t nperm) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const int64_t i = t < nperm ? perm[t] : t;
    if (mask && mask[i]) {
        y[i] = xin[i];
        return;
    }
    const int b = offsets[i], e = offsets[i + 1];
    y[i] = row_sum(indices, b, e, xe);
}

__global__ void k_native_indices(int64_t nE, int dof, const int *__restrict__ indices,
                                 const int *__restrict__ l2n, int *__restrict__ out) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= nE) return;
    const int v = indices[j];
    const int lid = v >= 0 ? v : -1 - v;
    const int e = lid / dof, l = lid - e * dof;
    const int n = e * dof + l2n[l];
    out[j] = v >= 0 ? n : -1 - n;
}

// flag[i] = 1 if E-index i is the first contribution of an L-DOF below nv
__global__ void k_first_flag(int64_t nE, const int *__restrict__ gmap,
                             const int *__restrict__ offsets, const int *__restrict__ indices,
                             int64_t nv, int *__restrict__ flag) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nE) return;
    const int g = gmap[i];
    const int gid = g >= 0 ? g : -1 - g;
    int f = 0;
    if (gid < nv) {
        const int first = indices[offsets[gid]];
        f = ((first >= 0 ? first : -1 - first) == i) ? 1 : 0;
    }
    flag[i] = f;
}

__global__ void k_first_place(int64_t nE, const int *__restrict__ gmap,
                              const int *__restrict__ incl, int *__restrict__ perm) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nE) return;
    const int prev = i > 0 ? incl[i - 1] : 0;
    if (incl[i] != prev) {
        const int g = gmap[i];
        perm[prev] = g >= 0 ? g : -1 - g;
    }
}

}  // namespace

void build_gather_map(const Mesh &m, int p, int64_t nvdofs, int64_t nedofs, int64_t nfdofs,
                      int *elem_dof, int *gather_map, hipStream_t s) {
    const int D = p + 1, dof = D * D * D;
    std::vector<int> dm(dof);
    h1_dof_map(p, dm.data());
    StreamBuf<int> d_dm;
    d_dm.upload(dm.data(), dof, s);
    const int64_t n = m.ne * dof;
    hipLaunchKernelGGL(k_build_gather, dim3(nblocks(n)), dim3(kBlock), 0, s, m.ne, p, dof,
                       m.d_topo.p, m.d_ori.p, d_dm.p, nvdofs, nedofs, nfdofs, elem_dof, gather_map);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));
}

void gather_from_table(int64_t ne, int dof, const int *elem_dof, const int *dof_map, int *gmap,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_gather_from_table, dim3(nblocks(ne * dof)), dim3(kBlock), 0, s, ne, dof,
                       elem_dof, dof_map, gmap);
    H1PA_HIP(hipGetLastError());
}

void build_transpose(int64_t ne, int dof, int64_t ndofs, const int *gmap, int *offsets,
                     int *indices, hipStream_t s) {
    const int64_t n = ne * dof;
    H1PA_HIP(hipMemsetAsync(offsets, 0, sizeof(int) * (ndofs + 1), s));
    hipLaunchKernelGGL(k_count, dim3(nblocks(n)), dim3(kBlock), 0, s, n, gmap, offsets);
    H1PA_HIP(hipGetLastError());
    inclusive_scan(offsets + 1, ndofs, s);  // offsets[i] = start of row i
    StreamBuf<int> cursor(ndofs + 1, s);
    H1PA_HIP(hipMemcpyAsync(cursor.p, offsets, sizeof(int) * (ndofs + 1), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kBlock), 0, s, n, gmap, cursor.p, indices);
    hipLaunchKernelGGL(k_sort_rows, dim3(nblocks(ndofs)), dim3(kBlock), 0, s, ndofs, offsets,
                       indices);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));
}

void boundary_mask(const Mesh &m, int p, const int *gmap, uint8_t *mask, int64_t ndofs,
                   hipStream_t s) {
    const int D = p + 1;
    const int64_t n = m.ne * D * D * D;
    H1PA_HIP(hipMemsetAsync(mask, 0, ndofs, s));
    hipLaunchKernelGGL(k_boundary_mark, dim3(nblocks(n)), dim3(kBlock), 0, s, m.ne, p,
                       m.d_topo.p, gmap, mask);
    H1PA_HIP(hipGetLastError());
}

int64_t mask_to_list(const uint8_t *mask, int64_t n, DevBuf<int> &list, hipStream_t s) {
    StreamBuf<int> incl(n, s);
    hipLaunchKernelGGL(k_mask_to_int, dim3(nblocks(n)), dim3(kBlock), 0, s, n, mask, incl.p);
    H1PA_HIP(hipGetLastError());
    inclusive_scan(incl.p, n, s);
    int total = 0;
    H1PA_HIP(hipMemcpyAsync(&total, incl.p + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    H1PA_HIP(hipStreamSynchronize(s));
    list.alloc(total > 0 ? total : 1);
    hipLaunchKernelGGL(k_compact, dim3(nblocks(n)), dim3(kBlock), 0, s, n, mask, incl.p, list.p);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));
    return total;
}

void list_to_mask(const int *list, int64_t nl, uint8_t *mask, int64_t n, hipStream_t s) {
    H1PA_HIP(hipMemsetAsync(mask, 0, n, s));
    if (nl > 0)
        hipLaunchKernelGGL(k_list_to_mask, dim3(nblocks(nl)), dim3(kBlock), 0, s, nl, list, mask, n);
    H1PA_HIP(hipGetLastError());
}

void restrict_mult(int64_t nE, const int *gmap, const double *x, double *y, bool abs_,
                   hipStream_t s) {
    // non-temporal index loads and E-vector stores: 0.242 -> 0.210 ms at p = 4
    // (79^3), 0.198 -> 0.172 ms at p = 6, the L-vector reads keeping the caches
    if (((reinterpret_cast<uintptr_t>(gmap) | reinterpret_cast<uintptr_t>(y)) & 15) == 0)
        hipLaunchKernelGGL((k_restrict_mult<4, true>), dim3(nblocks((nE + 3) / 4)), dim3(kBlock), 0,
                           s, nE, gmap, x, y, abs_);
    else
        hipLaunchKernelGGL(k_restrict_mult_1, dim3(nblocks(nE)), dim3(kBlock), 0, s, nE, gmap, x,
                           y, abs_);
    H1PA_HIP(hipGetLastError());
}

void native_indices(int64_t nE, int dof, const int *indices, const int *lex2nat, int *out,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_native_indices, dim3(nblocks(nE)), dim3(kBlock), 0, s, nE, dof, indices,
                       lex2nat, out);
    H1PA_HIP(hipGetLastError());
}

void first_contribution_order(int64_t nE, const int *gmap, const int *offsets,
                              const int *indices, int64_t nv, DevBuf<int> &perm, hipStream_t s) {
    StreamBuf<int> incl(nE, s);
    hipLaunchKernelGGL(k_first_flag, dim3(nblocks(nE)), dim3(kBlock), 0, s, nE, gmap, offsets,
                       indices, nv, incl.p);
    H1PA_HIP(hipGetLastError());
    inclusive_scan(incl.p, nE, s);
    int total = 0;
    H1PA_HIP(hipMemcpyAsync(&total, incl.p + nE - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    H1PA_HIP(hipStreamSynchronize(s));
    H1PA_CHECK(total == nv, H1PA_EINVAL, "first-contribution order: every vertex DOF needs an entry");
    perm.alloc(nv > 0 ? nv : 1);
    hipLaunchKernelGGL(k_first_place, dim3(nblocks(nE)), dim3(kBlock), 0, s, nE, gmap, incl.p,
                       perm.p);
    H1PA_HIP(hipGetLastError());
    H1PA_HIP(hipStreamSynchronize(s));
}

void restrict_mult_transpose(int64_t nrows, const int *offsets, const int *indices,
                             const double *xe, double *y, int mode, const int *perm,
                             int64_t nperm, hipStream_t s) {
    if (nrows <= 0) return;
    // two rows per thread: measured against one and four rows per thread at
    // p = 2, 4, 6, 8 (32M, 32M, 24M, 33M DOFs), fastest at every order
    hipLaunchKernelGGL(k_restrict_mult_t<2>, dim3(nblocks((nrows + 1) / 2)), dim3(kBlock), 0, s,
                       nrows, offsets, indices, xe, y, mode, perm, perm ? nperm : 0);
    H1PA_HIP(hipGetLastError());
}

void bubble_mult_transpose(int64_t ne, int D, int64_t bb, const double *xe, double *y, int mode,
                           hipStream_t s) {
    const int NI = D - 2;
    if (NI <= 0 || ne <= 0) return;
    const int64_t nrows = ne * (int64_t)(NI * NI * NI);
    hipLaunchKernelGGL(k_bubble_mult_t, dim3(nblocks(nrows)), dim3(kBlock), 0, s, nrows, NI, D, bb,
                       xe, y, mode);
    H1PA_HIP(hipGetLastError());
}

void boolean_mask(int64_t nrows, const int *offsets, const int *indices, double *ye,
                  hipStream_t s) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_boolean_mask, dim3(nblocks(nrows)), dim3(kBlock), 0, s, nrows, offsets,
                       indices, ye);
    H1PA_HIP(hipGetLastError());
}

void scatter_constrained(int64_t nrows, const int *offsets, const int *indices, const double *xe,
                         const uint8_t *mask, const double *xin, double *y, const int *perm,
                         int64_t nperm, hipStream_t s) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_scatter_constrained, dim3(nblocks(nrows)), dim3(kBlock), 0, s, nrows,
                       offsets, indices, xe, mask, xin, y, perm, perm ? nperm : 0);
    H1PA_HIP(hipGetLastError());
}

}  // namespace kern
}  // namespace h1pa
