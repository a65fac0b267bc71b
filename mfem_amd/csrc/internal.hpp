// internal.hpp -- shared declarations of the h1pa library (host C++ + HIP).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <functional>
#include <memory>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/h1pa.h"

namespace h1pa {

// ---- errors ---------------------------------------------------------------
void set_error(const std::string &msg);
struct Error {
    int code;
    std::string msg;
};
#define H1PA_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::h1pa::Error{H1PA_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)
#define H1PA_CHECK(cond, code, msg)                                                      \
    do {                                                                                 \
        if (!(cond)) throw ::h1pa::Error{(code), (msg)};                                 \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// bfloat16 storage: the upper 16 bits of an IEEE float.
struct bf16_t {
    uint16_t bits;
};

// ---- device buffer ---------------------------------------------------------
// Device blocks for DevBuf (devmem.cpp).  With the block cache off (the
// default) these are hipMalloc and hipFree.  With it on
// (h1pa_set_device_cache), a freed block is kept on a per-device list instead
// of hipFree, which synchronizes the whole device, and an allocation reuses a
// kept block of at least its size (at most 1.25x); the caller then must not
// release an object while device work on its buffers is pending.
void *device_alloc(size_t bytes, size_t &cap, int &dev);
void device_free(void *p, size_t cap, int dev);
void device_cache_set(int64_t max_bytes);  // 0: off (kept blocks are freed)
int64_t device_cache_bytes();               // bytes kept, all devices

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    size_t cap = 0;  // bytes of the block behind p
    int dev = 0;     // device it was allocated on
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), cap(o.cap), dev(o.dev) {
        o.p = nullptr;
        o.n = 0;
        o.cap = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            cap = o.cap;
            dev = o.dev;
            o.p = nullptr;
            o.n = 0;
            o.cap = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        p = static_cast<T *>(device_alloc(count * sizeof(T), cap, dev));
        n = count;
    }
    void release() {
        if (p) device_free(p, cap, dev);
        p = nullptr;
        n = 0;
        cap = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        if (count) H1PA_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// Device memory ordered on one stream (hipMallocAsync / hipFreeAsync from the
// device's default pool, which keeps freed blocks for reuse): for
// temporaries that only that stream uses.  Releasing one does not wait: a
// hipFree synchronizes the whole device, which stalls a set-up or a solve
// running beside others on their own streams.
void keep_stream_pool();  // the default pool keeps freed memory (once per device)
template <typename T>
struct StreamBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;
    StreamBuf() = default;
    StreamBuf(size_t count, hipStream_t st) { alloc(count, st); }
    StreamBuf(const StreamBuf &) = delete;
    StreamBuf &operator=(const StreamBuf &) = delete;
    ~StreamBuf() { release(); }
    void alloc(size_t count, hipStream_t st) {
        release();
        s = st;
        if (count == 0) return;
        keep_stream_pool();
        if (hipMallocAsync(reinterpret_cast<void **>(&p), count * sizeof(T), st) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw Error{H1PA_ENOMEM, "hipMallocAsync of " + std::to_string(count * sizeof(T)) +
                                         " bytes failed"};
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFreeAsync(p, s);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t st) {
        if (count > n || st != s) alloc(count, st);
        if (count) H1PA_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
};

// ---- 1D rules and basis (host, long double) --------------------------------
void gauss_legendre(int n, double *x, double *w);
void gauss_lobatto(int n, double *x, double *w);
// B[q + n*d] = l_d(x_q), G likewise; Lagrange basis on p+1 GLL nodes.
void lagrange_basis(int p, int n, const double *xq, double *B, double *G);
void h1_dof_map(int p, int *dof_map);  // lex -> native (fem/fe/fe_base.cpp:2557-2674)
// Even-odd tables of B and G (Q x D, q fastest) appended to the packed basis
// after its 4 Q D plain entries; returns the number of doubles written to
// out (the layout the apply kernels read, EO<D, Q> in apply_body.hpp).
int even_odd_tables(int D, int Q, const double *B, const double *G, double *out);
int even_odd_size(int D, int Q);

// ---- mesh (host topology, device coordinates) -------------------------------
struct Mesh {
    int nx = 0, ny = 0, nz = 0;
    int64_t nv = 0, ne = 0, nedges = 0, nfaces = 0;
    std::vector<int> ev, ee, eeo, ef, efo;   // host topology
    std::vector<int> ecoord;                 // [ne][3] grid position of each element
    std::vector<double> xyz;                 // host vertex coordinates [nv][3]
    // geometry: order-pg nodes (pg = 1: vertices), device E-vector [Dg^3][3][ne]
    int node_order = 1;
    double curve_amp = 0.0;
    DevBuf<double> d_nodes;
    bool nodes_valid = false;
    bool nodes_external = false;             // d_nodes given by the caller (set_nodes)
    std::vector<uint8_t> bface;              // [ne] bit j: local face j is on the boundary
    bool from_file = false;                  // read from an MFEM mesh file (no grid)
    DevBuf<int> d_elem_at;  // Cartesian meshes: element at each grid position (built on use)
    h1pa_mesh *handle = nullptr;  // the C handle holding this mesh
    DevBuf<int> d_topo;   // [ne][29] vertex, edge, face ids, boundary-face bits, 2 unused
    DevBuf<int8_t> d_ori; // [ne][12+6] edge (+-1) and face (0..7) orientations
    void build_topology(bool sfc);
    // ReadMFEMMesh for "MFEM mesh v1.0" hexahedral text; sets ev, xyz and the
    // topology (first-encounter edges/faces, boundary from the file's list).
    void load_mfem_v10(const char *text, size_t len);
    // Mesh from vertex, element and boundary-quad arrays (+ FinalizeTopology).
    void from_arrays(int64_t n_v, const double *pts, int64_t n_el, const int *evi,
                     int64_t n_bdr, const int *bq);
    void ensure_device(hipStream_t s);
    void ensure_nodes(hipStream_t s);
};

// Column decomposition of a Cartesian mesh for the fused apply
// (columns.hip): npx x npy columns of A x B elements, each swept in nseg
// segments of ls element layers; elem_at[(k ny + j) nx + i] is the element
// at grid position (i, j, k).
struct ColumnGrid {
    int nx = 0, ny = 0, nz = 0;
    int npx = 0, npy = 0, nseg = 0, ls = 0;
    const int *elem_at = nullptr;
};

// ---- restriction -----------------------------------------------------------
struct Restriction {
    int64_t ne = 0, dof = 0, ndofs = 0;
    DevBuf<int> gather_map, offsets, indices;
    // Scatter order of the first `nperm` L-DOFs (the vertex block): sorted by
    // their first contributing E-index, so that a wave of the E->L pass reads
    // E-entries of neighbouring elements.  Empty: natural order.
    DevBuf<int> perm;
    int64_t nperm = 0;
    // Scatter order of the first `nperm_lex` L-DOFs (every row below the
    // bubble block) for E-vectors in the lexicographic layout, by first
    // contributing E-index: rows that read the same element's cache lines run
    // together, so each line is fetched about once.  Empty: natural order.
    DevBuf<int> perm_lex;
    int64_t nperm_lex = 0;
    // First L-DOF of the bubble block (interior DOFs, one E-entry each) or
    // ndofs when the numbering has no such block.
    int64_t bubble_begin = 0;
    // Internal E-vector layout of the fused operator: per element in native
    // (vertex, edge, face, bubble) order, so the entries of one edge or face
    // are contiguous for the E->L pass.  lex2nat[lex] = native position;
    // indices_nat = indices re-addressed to that layout (same order).
    DevBuf<int> lex2nat, indices_nat;
};

struct Space {
    Mesh *mesh = nullptr;
    int p = 1;
    int64_t ne = 0, dof = 0, ndofs = 0;
    int64_t nvdofs = 0, nedofs = 0, nfdofs = 0;
    DevBuf<int> elem_dof;   // native order [ne][dof]
    DevBuf<int> ess;        // boundary true DOFs, ascending
    int64_t n_ess = 0;
    Restriction restr;
    DevBuf<uint8_t> ess_mask;  // [ndofs] 1 on boundary DOFs
};

struct Form {
    Space *space = nullptr;
    h1pa_form_desc desc{};
    int q1d = 0;
    DevBuf<double> basis;      // [B q-fastest | G | B d-fastest | G d-fastest]
    DevBuf<double> pa_diff;    // [Q^3][6][ne]
    DevBuf<double> pa_mass;    // [Q^3][ne]
    DevBuf<double> ework0, ework1;  // E-vector scratch
    DevBuf<double> lwork0, lwork1;  // L-vector scratch
    DevBuf<uint8_t> cmask;          // constraint mask of the last ess list passed in
    bool assembled = false;
    // Reduced-precision shadows for the inner solver (built on first use from
    // the FP64 data): basis in FP32, PA data in FP32 or BF16 (lp_kind 0 / 1).
    DevBuf<float> basis32;
    DevBuf<uint32_t> pa_diff_lp, pa_mass_lp;  // raw storage, 4-byte aligned
    int lp_kind = -1;
    DevBuf<float> ework32;
    // Batch solves (added, not yet retired by poll) and multigrid
    // hierarchies that hold raw pointers to this form's PA data: while
    // nonzero, calls that would free or move those buffers (Assemble, a
    // reduced-precision shadow of the other kind, destroy) are refused.
    std::atomic<int> batch_refs{0};
    // fused column apply (Cartesian meshes, diffusion only, q1d = p + 2)
    ColumnGrid cols;
    DevBuf<double> iface;
    int use_columns = -1;  // -1 not decided yet, 0 no, 1 yes
};
// Refuses (H1PA_EINVAL) when live batch solves or multigrid hierarchies
// reference f's device data.
void require_no_batch_refs(const Form &f, const char *what);

// ---- kernel launchers (HIP translation units) ------------------------------
namespace kern {
void build_gather_map(const Mesh &m, int p, int64_t nvdofs, int64_t nedofs, int64_t nfdofs,
                      int *elem_dof_native, int *gather_map, hipStream_t s);
void gather_from_table(int64_t ne, int dof, const int *elem_dof, const int *dof_map,
                       int *gather_map, hipStream_t s);
void build_transpose(int64_t ne, int dof, int64_t ndofs, const int *gather_map, int *offsets,
                     int *indices, hipStream_t s);
void boundary_mask(const Mesh &m, int p, const int *gather_map, uint8_t *mask, int64_t ndofs,
                   hipStream_t s);
int64_t mask_to_list(const uint8_t *mask, int64_t n, DevBuf<int> &list, hipStream_t s);
void list_to_mask(const int *list, int64_t nlist, uint8_t *mask, int64_t n, hipStream_t s);

void restrict_mult(int64_t nE, const int *gmap, const double *x, double *y, bool abs_,
                   hipStream_t s);
// indices_nat[j] = indices[j] with E-index e*dof + lex moved to e*dof + lex2nat[lex].
void native_indices(int64_t nE, int dof, const int *indices, const int *lex2nat, int *out,
                    hipStream_t s);
// First-contribution order of the L-DOFs [0, nv): perm[k] = k-th vertex DOF.
void first_contribution_order(int64_t nE, const int *gmap, const int *offsets,
                              const int *indices, int64_t nv, DevBuf<int> &perm, hipStream_t s);
// E->L over rows [0, nrows), rows t < nperm taken as perm[t].
// mode 0 set, 1 add, 2 abs set, 3 left inverse
void restrict_mult_transpose(int64_t nrows, const int *offsets, const int *indices,
                             const double *xe, double *y, int mode, const int *perm,
                             int64_t nperm, hipStream_t s);
// The bubble block of an H1 space's restriction (rows bb .. bb + ne (D-2)^3)
// from a lexicographic E-vector, without the transpose table; same modes.
void bubble_mult_transpose(int64_t ne, int D, int64_t bb, const double *xe, double *y, int mode,
                           hipStream_t s);
// p-refinement transfer (transfer.hip): emask[E] = 1 at the first E-entry
// (indices order) of each L-DOF; prolongation low L -> high L (every high
// L-DOF written); restriction high L -> low E-vector (lexicographic), to be
// summed by the low restriction's E->L pass.  B[qh + (ph+1) dl]: the order-pl
// basis at the order-ph GLL nodes.
void first_entry_mask(int64_t nE, const int *gmap, const int *offsets, const int *indices,
                      uint8_t *emask, hipStream_t s);
void transfer_prolong(int64_t ne, int pl, int ph, const double *B, const int *gmap_l,
                      const int *gmap_h, const uint8_t *emask_h, const double *xl, double *yh,
                      hipStream_t s);
void transfer_restrict_e(int64_t ne, int pl, int ph, const double *B, const int *gmap_h,
                         const uint8_t *emask_h, const double *xh, double *yle, hipStream_t s);
// BooleanMask: ye[first E-entry of each row] = 1, every other E-entry 0.
void boolean_mask(int64_t nrows, const int *offsets, const int *indices, double *ye,
                  hipStream_t s);
// E->L set with constraint handling: y[i] = mask[i] ? xin[i] : sum (DIAG_ONE)
void scatter_constrained(int64_t nrows, const int *offsets, const int *indices,
                         const double *xe, const uint8_t *mask, const double *xin, double *y,
                         const int *perm, int64_t nperm, hipStream_t s);

void perturb_vertices(Mesh &m, double amp, uint64_t seed);  // host
void nodes_trilinear(const Mesh &m, double *d_nodes, hipStream_t s);
void nodes_curved(const Mesh &m, int pg, double amp, double *d_nodes, hipStream_t s);
// QuadratureInterpolator::Mult on an E-vector [D^3][vdim][ne] of the space
// at the q1d^3 Gauss-Legendre points (H1PA_QI_* flags; byVDIM layout when
// by_vdim).
bool qinterp_supported(int p, int q1d);
void qinterp(Space &sp, int q1d, int vdim, unsigned flags, int by_vdim, const double *x,
             double *val, double *der, double *det, hipStream_t s);
void geometric_factors(const Mesh &m, int q1d, double *X, double *J, double *detJ,
                       hipStream_t s);
void pa_setup(const Mesh &m, const Form &f, hipStream_t s);
// DomainLFIntegrator element vectors at q1d = p + 1: constant f, or f given
// per quadrature point (fq[q + (p+1)^3 e]) when fq is not null.
void domain_lf(const Mesh &m, int p, double f, const double *fq, double *bE, hipStream_t s);
void randomize(int64_t n, double *x, uint64_t seed, double lo, double hi, hipStream_t s);

// E->E: ye = A_E xe (overwrite) or += ; x may be gathered from an L-vector
// through gmap (gmap != nullptr) with constrained entries masked to zero.
// basis = [B q-fastest | G q-fastest | B d-fastest | G d-fastest], 4*Q*D doubles.
// Output of an element kernel: the E-vector ye (lexicographic, or native
// order when l2n != nullptr) and, when yL != nullptr, interior (bubble) DOFs
// written straight to the L-vector yL at MFEM's bubble block starting at bb.
template <typename T>
struct EOutT {
    T *ye = nullptr;
    T *yL = nullptr;
    int64_t bb = 0;
    const int *l2n = nullptr;
    const int *stop = nullptr;  // device flag: when nonzero the launch does nothing
};
using EOut = EOutT<double>;
// Reduced-precision applies: FP32 vectors/arithmetic, PA data as FP32
// (pa_kind 0) or BF16 (pa_kind 1, FP32 accumulation).
void diffusion_apply_lp(int p, int q1d, int64_t ne, const float *basis, const void *pa,
                        int pa_kind, const int *gmap, const uint8_t *mask, const float *x,
                        const EOutT<float> &out, bool accumulate, hipStream_t s);
// Variants of the reduced-precision diffusion apply: 0, 1 (unroll 2), 4 (PA
// prefetch 2 planes); bit-identical results.
void diffusion_apply_lp_variant(int variant, int p, int q1d, int64_t ne, const float *basis,
                                const void *pa, int pa_kind, const int *gmap,
                                const uint8_t *mask, const float *x, const EOutT<float> &out,
                                bool accumulate, hipStream_t s);
void set_diffusion_variant_lp(int pa_kind, int p, int q1d, int variant);
int get_diffusion_variant_lp(int pa_kind, int p, int q1d);
// FP64 vectors with BF16-stored diffusion PA data (q1d = p + 2).
void diffusion_apply_d_bf16(int p, int q1d, int64_t ne, const double *basis, const void *pa,
                            const int *gmap, const uint8_t *mask, const double *x,
                            const EOut &out, bool accumulate, hipStream_t s);
void mass_apply_lp(int p, int q1d, int64_t ne, const float *basis, const void *pa, int pa_kind,
                   const int *gmap, const uint8_t *mask, const float *x, const EOutT<float> &out,
                   bool accumulate, hipStream_t s);
void diffusion_apply(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                     const int *gmap, const uint8_t *mask, const double *x, const EOut &out,
                     bool accumulate, hipStream_t s);
void mass_apply(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                const int *gmap, const uint8_t *mask, const double *x, const EOut &out,
                bool accumulate, hipStream_t s);
// Element diagonals into an element output (same layouts as the applies).
void diffusion_diagonal(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                        const EOut &out, bool accumulate, hipStream_t s);
void mass_diagonal(int p, int q1d, int64_t ne, const double *basis, const double *pa,
                   const EOut &out, bool accumulate, hipStream_t s);
bool apply_supported(int p, int q1d);
// Fused L->L diffusion apply by column sweeps (Cartesian meshes, q1d = p + 2):
// y = A x written whole; iface holds column_interface_size(p, g) doubles.
bool columns_supported(int p, int q1d);
void column_shape(int p, int &A, int &B);
int64_t column_interface_size(int p, const ColumnGrid &g);
// parts: 1 the column sweep, 2 the interface pass, 3 both.
void diffusion_columns(int p, const ColumnGrid &g, const double *basis, const double *pa,
                       const int *gmap, const double *x, double *y, double *iface, int parts,
                       hipStream_t s);
// Speed variants of the diffusion apply (bit-identical results).
int diffusion_variant_count();
void set_diffusion_variant(int p, int q1d, int variant);
int get_diffusion_variant(int p, int q1d);
void diffusion_apply_variant(int variant, int p, int q1d, int64_t ne, const double *basis,
                             const double *pa, const int *gmap, const uint8_t *mask,
                             const double *x, const EOut &out, bool accumulate, hipStream_t s);
void jacobi_dinv(int64_t n, const double *diag, const uint8_t *mask, double *dinv, hipStream_t s);
void jacobi_dinv_f32(int64_t n, const double *diag, const uint8_t *mask, float *dinv,
                     hipStream_t s);
void form_x(int64_t n, const uint8_t *mask, const double *x, double *X, hipStream_t s);
// y += a z on n entries
void axpy(int64_t n, double a, const double *z, double *y, hipStream_t s);
void form_b(int64_t n, const uint8_t *mask, const double *X, const double *b, const double *z,
            double *B, hipStream_t s);
}  // namespace kern

// OperatorChebyshevSmoother on a form's DIAG_ONE constrained operator.
struct Chebyshev {
    Form *f = nullptr;
    const uint8_t *mask = nullptr;  // constraint mask (the space's or own.p)
    DevBuf<uint8_t> own;
    int order = 0;
    int prec = H1PA_PREC_FP64;  // BF16_PA: applies with the form's BF16 PA shadow
    double max_eig = 0.0;
    double coeffs[5] = {0, 0, 0, 0, 0};
    DevBuf<double> dinv, t, u;
    // power-method work kept with the smoother: freeing device memory waits
    // for the whole device, which stalls a set-up that runs beside solves
    DevBuf<double> w, partials;
    DevBuf<uint32_t> starts;  // the start vector's generator states
};
// dinv from d_diag (null: the form's assembled diagonal), the eigenvalue
// estimate (max_eig > 0: given; else the power method), the coefficients.
void chebyshev_setup(Chebyshev &c, const double *d_diag, double max_eig, int power_iterations,
                     double power_tolerance, int power_seed, hipStream_t s);
void chebyshev_mult(Chebyshev &c, const double *x, double *y, hipStream_t s);
void pcg_chebyshev(Chebyshev &c, const double *b, double *x, const h1pa_pcg_opts &o,
                   h1pa_pcg_stats &stats, hipStream_t s);

void pcg_solve(Form &f, const uint8_t *mask, const double *diag, const double *b, double *x,
               const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, hipStream_t s);
// Work vectors of a CG solve kept by the caller between solves (multigrid's
// coarse solves allocate nothing once they have the size).
struct PcgWork {
    DevBuf<double> r, z, d, q, partials;
    DevBuf<uint8_t> state;
};
// Jacobi-PCG with the inverse diagonal given and the caller's work vectors.
// prec = H1PA_PREC_BF16_PA applies the operator with the form's BF16 PA shadow
// (FP64 vectors and arithmetic).
void pcg_solve_dinv(Form &f, const uint8_t *mask, const double *dinv, const double *b, double *x,
                    const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, PcgWork &w, hipStream_t s,
                    int prec = H1PA_PREC_FP64);
// u = A v on f's DIAG_ONE constrained operator (mask: the constrained DOFs;
// stop: a device flag that skips the work when set, or null; prec as above).
void constrained_apply(Form &f, const uint8_t *mask, const double *v, double *u,
                       const int *stop, hipStream_t s, int prec = H1PA_PREC_FP64);
// A preconditioner z = M r on the device (stop as above).
using Precond = std::function<void(const double *r, double *z, const int *stop, hipStream_t s)>;
// CGSolver::Mult (linalg/solvers.cpp:869-1050, iterative_mode = false) with
// SetPreconditioner(M) on f's DIAG_ONE constrained operator; the host reads
// the solver state every `check` iterations.
// ws: work vectors kept by the caller (null: allocated for this solve).
void pcg_preconditioned(Form &f, const uint8_t *mask, const Precond &M, const double *b,
                        double *x, const h1pa_pcg_opts &o, h1pa_pcg_stats &stats, int check,
                        hipStream_t s, PcgWork *ws);
// Builds f's reduced-precision shadows for `precision` (H1PA_PREC_FP32 or
// H1PA_PREC_BF16_PA) unless they already match.
void ensure_low_precision(Form &f, int precision, hipStream_t s);
void ir_solve(Form &f, const uint8_t *mask, const double *diag, const double *b, double *x,
              const h1pa_ir_opts &o, h1pa_ir_stats &stats, hipStream_t s);
// Reduced-precision L->L apply of the unconstrained operator (FP32 vectors).
void form_mult_lp(Form &f, int precision, const float *x, float *y, hipStream_t s);
// Jacobi diagonal of the unconstrained operator (L-vector, set).
void operator_diagonal(Form &f, double *diag, hipStream_t s);
// ElementRestriction::MultTranspose family (mode 0 set, 1 add, 2 abs, 3 left
// inverse) from a lexicographic E-vector; bubble_block: R is an H1 space's
// restriction, whose interior-DOF block is written by position.
void restriction_scatter(const Restriction &R, bool bubble_block, const double *x, double *y,
                         int mode, hipStream_t s);

// ---- p-multigrid (transfer.hip, multigrid.cpp) -----------------------------
// TensorProductPRefinementTransferOperator between two spaces of one mesh.
struct Transfer {
    Space *lo = nullptr, *hi = nullptr;
    DevBuf<double> B;         // low-order basis at the high-order GLL nodes
    DevBuf<uint8_t> emask_h;  // BooleanMask of the high restriction, one byte per E-entry
    DevBuf<double> eL;        // low E-vector of the restriction
    Transfer(Space &lo, Space &hi, hipStream_t s);
    void mult(const double *xl, double *yh, hipStream_t s);             // prolongation
    void mult_transpose(const double *xh, double *yl, hipStream_t s);   // restriction
};
struct MgLevel {
    int prec = H1PA_PREC_FP64;        // PA data of the cycle's applies on this level
    h1pa_space *space_own = nullptr;  // levels below the finest: made here
    h1pa_form *form_own = nullptr;
    Form *f = nullptr;
    Space *sp = nullptr;
    const uint8_t *mask = nullptr;        // the level's boundary DOFs
    std::unique_ptr<Chebyshev> smoother;  // levels >= 1
    DevBuf<double> X, Y, R, Z;
};
struct Multigrid {
    h1pa_mg_opts o{};
    std::vector<MgLevel> lv;                   // 0 = coarsest
    std::vector<std::unique_ptr<Transfer>> P;  // P[l]: level l -> l + 1
    DevBuf<double> coarse_dinv;                // ones: CG without a preconditioner
    PcgWork coarse_work, outer_work;           // kept between cycles and solves
    int64_t cycles = 0, coarse_solves = 0, coarse_iterations = 0;
    Form *fine_ref = nullptr;  // the caller's fine form, held (batch_refs) while this lives
    ~Multigrid();
};
Multigrid *multigrid_create(Form &fine, int nlevels, const int *orders, const h1pa_mg_opts &o,
                            hipStream_t s);
void multigrid_mult(Multigrid &mg, const double *x, double *y, hipStream_t s);
void pcg_multigrid(Multigrid &mg, const double *b, double *x, const h1pa_pcg_opts &o,
                   h1pa_pcg_stats &stats, hipStream_t s);
void multigrid_info(const Multigrid &mg, int64_t *info);
double multigrid_max_eig(const Multigrid &mg, int level);

// ---- batched solves (batch.hip) --------------------------------------------
h1pa_batch *batch_create(const h1pa_batch_opts &o);
void batch_destroy(h1pa_batch *B);
hipStream_t batch_stream(const h1pa_batch &B);
// mask: the solve's constraint mask (copied) or null; the setup work runs on
// `cs`; the solve joins at the next launch.  Returns its index among the
// pending solves.
int batch_add(h1pa_batch &B, h1pa_form *form, const uint8_t *mask, const double *b, double *x,
              const h1pa_pcg_opts &po, int64_t tag, hipStream_t cs);
bool batch_launch(h1pa_batch &B);
int batch_wait(h1pa_batch &B);
// True when no replay is in flight or the one in flight has finished (a
// batch_wait would return at once); never blocks.
bool batch_idle(h1pa_batch &B);
int batch_run(h1pa_batch &B, int max_replays);
int batch_poll(h1pa_batch &B, int cap, int64_t *tags, h1pa_batch_result *res);
void batch_info(const h1pa_batch &B, int64_t *info);
// Device bytes held by a form, its space and its mesh.
int64_t form_device_bytes(const Form &f);

}  // namespace h1pa

struct h1pa_mesh { h1pa::Mesh m; };
struct h1pa_space { h1pa::Space s; h1pa_restriction *r = nullptr; };
struct h1pa_restriction {
    h1pa::Restriction own;                  // maps built by h1pa_restriction_create
    const h1pa::Restriction *view = nullptr; // or the space's maps (borrowed)
    bool owned_by_space = false;
    const h1pa::Restriction &R() const { return view ? *view : own; }
};
struct h1pa_form { h1pa::Form f; };
struct h1pa_chebyshev { h1pa::Chebyshev c; };
struct h1pa_transfer { std::unique_ptr<h1pa::Transfer> t; };
struct h1pa_multigrid { std::unique_ptr<h1pa::Multigrid> mg; };
