// row_sum.hpp -- the E -> L reduction of one L-DOF row, shared by the
// restriction transpose (restriction.hip) and the solver scatter kernels
// (pcg.hip, batch.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace h1pa {
namespace kern {

// Rows of n <= R entries, 1 <= n: R index loads, then R value loads, all
// unconditional (entries past n re-read entry n-1), then the n adds in order.
template <int R, typename T>
__device__ __forceinline__ T row_sum_batch(const int *__restrict__ indices, int b, int n,
                                           const T *__restrict__ xe, bool unsigned_) {
    int ix[R];
#pragma unroll
    for (int k = 0; k < R; ++k) ix[k] = indices[b + (k < n ? k : n - 1)];
    T v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = xe[ix[k] >= 0 ? ix[k] : -1 - ix[k]];
    T s = T(0);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const T t = (ix[k] >= 0 || unsigned_) ? v[k] : -v[k];
        s = k < n ? s + t : s;
    }
    return s;
}

// Sum of the E-vector entries listed in indices[b, e) (ElementRestriction::
// TAddMultTranspose, fem/restriction.cpp:159-187): an entry encoded as -1-k
// enters as -xe[k] (as +xe[k] when `unsigned_` is set, AbsMultTranspose at
// :203), and the entries are added in table order, so the result is MFEM's
// bit for bit.
//
// A row's first entries are read as one batch of 2, 4 or 8 -- the smallest
// that holds every row of the wave (face, edge and vertex rows of a Cartesian
// hexahedral mesh; in the fused operator's row order the vertex rows come
// first, so its waves are mostly of one kind, while the element-major order of
// the lexicographic scatters mixes them and takes the batch of 8): every index
// load is issued before any value load, with no branch between them, so a row
// costs two dependent memory round trips instead of two per entry.  Entries
// past the 8th continue in a loop.
template <typename T>
__device__ __forceinline__ T row_sum(const int *__restrict__ indices, int b, int e,
                                     const T *__restrict__ xe, bool unsigned_ = false) {
    const int n = e - b;
    if (n <= 0) return T(0);
    T s;
    if (!__ballot(n > 2)) s = row_sum_batch<2>(indices, b, n, xe, unsigned_);
    else if (!__ballot(n > 4)) s = row_sum_batch<4>(indices, b, n, xe, unsigned_);
    else s = row_sum_batch<8>(indices, b, n < 8 ? n : 8, xe, unsigned_);
    for (int j = b + 8; j < e; ++j) {
        const int k = indices[j];
        s += k >= 0 ? xe[k] : (unsigned_ ? xe[-1 - k] : -xe[-1 - k]);
    }
    return s;
}

// R rows at once (rows_sum): every index load of the R rows, then every value
// load, then the sums -- R times the loads in flight of one row_sum, each
// row's entries still added in table order.  Needs n[r] >= 1; the loads are
// unconditional (a predicated load per entry costs more than it saves).
template <int R, int NB, typename T>
__device__ __forceinline__ void rows_sum_batch(const int *__restrict__ indices, const int (&b)[R],
                                               const int (&n)[R], const T *__restrict__ xe,
                                               bool unsigned_, T (&s)[R]) {
    int ix[R][NB];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < NB; ++k) ix[r][k] = indices[b[r] + (k < n[r] ? k : n[r] - 1)];
    T v[R][NB];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < NB; ++k) v[r][k] = xe[ix[r][k] >= 0 ? ix[r][k] : -1 - ix[r][k]];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const T t = (ix[r][k] >= 0 || unsigned_) ? v[r][k] : -v[r][k];
            acc = k < n[r] ? acc + t : acc;
        }
        s[r] = acc;
    }
}

// s[r] = the sum of row r's entries indices[b[r], b[r] + n[r]) in table order
// (as row_sum), for R rows of one thread.  A wave holding an empty row takes
// row_sum row by row.
template <int R, typename T>
__device__ __forceinline__ void rows_sum(const int *__restrict__ indices, const int (&b)[R],
                                         const int (&n)[R], const T *__restrict__ xe, T (&s)[R],
                                         bool unsigned_ = false) {
    int nmin = n[0], nmax = n[0];
#pragma unroll
    for (int r = 1; r < R; ++r) {
        nmin = n[r] < nmin ? n[r] : nmin;
        nmax = n[r] > nmax ? n[r] : nmax;
    }
    if (__ballot(nmin <= 0)) {
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = row_sum(indices, b[r], b[r] + n[r], xe, unsigned_);
        return;
    }
    if (!__ballot(nmax > 2)) rows_sum_batch<R, 2>(indices, b, n, xe, unsigned_, s);
    else if (!__ballot(nmax > 4)) rows_sum_batch<R, 4>(indices, b, n, xe, unsigned_, s);
    else rows_sum_batch<R, 8>(indices, b, n, xe, unsigned_, s);
    if (__ballot(nmax > 8)) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            for (int j = b[r] + 8; j < b[r] + n[r]; ++j) {
                const int k = indices[j];
                s[r] += k >= 0 ? xe[k] : (unsigned_ ? xe[-1 - k] : -xe[-1 - k]);
            }
    }
}

}  // namespace kern
}  // namespace h1pa
