// capi.cpp -- extern "C" entry points of libh1pa (declared in include/h1pa.h).
//
// Every entry point catches h1pa::Error (and std::exception) and turns it into
// a status code plus a thread-local message; nothing throws across the ABI.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <exception>
#include <new>

#include "internal.hpp"

namespace h1pa {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace h1pa

using namespace h1pa;

#define API_BEGIN try {
#define API_END                                                     \
    }                                                               \
    catch (const h1pa::Error &e) {                                  \
        set_error(e.msg);                                           \
        return e.code;                                              \
    }                                                               \
    catch (const std::bad_alloc &) {                                \
        set_error("host allocation failed");                        \
        return H1PA_ENOMEM;                                         \
    }                                                               \
    catch (const std::exception &e) {                               \
        set_error(e.what());                                        \
        return H1PA_EINVAL;                                         \
    }                                                               \
    set_error("");                                                  \
    return H1PA_OK;

#define REQUIRE(cond, msg) H1PA_CHECK(cond, H1PA_EINVAL, msg)

extern "C" {

const char *h1pa_last_error(void) { return g_last_error.c_str(); }
const char *h1pa_version(void) { return "h1pa 0.1 (gfx950)"; }

int h1pa_set_device(int device) {
    API_BEGIN
    H1PA_HIP(hipSetDevice(device));
    API_END
}

int h1pa_stream_synchronize(void *stream) {
    API_BEGIN
    H1PA_HIP(hipStreamSynchronize(as_stream(stream)));
    API_END
}

// ---- mesh -----------------------------------------------------------------
int h1pa_mesh_create_cartesian(int nx, int ny, int nz, int sfc, h1pa_mesh **mesh) {
    API_BEGIN
    REQUIRE(mesh, "mesh: null output");
    REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "mesh: nx, ny, nz must be >= 1");
    REQUIRE((int64_t)(nx + 1) * (ny + 1) * (nz + 1) < (1LL << 31), "mesh: too many vertices");
    auto *m = new h1pa_mesh();
    m->m.handle = m;
    m->m.nx = nx;
    m->m.ny = ny;
    m->m.nz = nz;
    try {
        m->m.build_topology(sfc != 0);
    } catch (...) {
        delete m;
        throw;
    }
    *mesh = m;
    API_END
}

int h1pa_mesh_load(const char *text, int64_t len, h1pa_mesh **mesh) {
    API_BEGIN
    REQUIRE(text && mesh && len > 0, "mesh_load: null argument");
    auto m = std::make_unique<h1pa_mesh>();
    m->m.handle = m.get();
    m->m.load_mfem_v10(text, (size_t)len);
    *mesh = m.release();
    API_END
}

int h1pa_mesh_create(int64_t nv, const double *xyz, int64_t ne, const int *ev, int64_t nbdr,
                     const int *bdr, h1pa_mesh **mesh) {
    API_BEGIN
    REQUIRE(xyz && ev && mesh && (nbdr == 0 || bdr), "mesh_create: null argument");
    REQUIRE(nv >= 8 && nv < (1LL << 31), "mesh_create: vertex count must be in [8, 2^31)");
    REQUIRE(ne >= 1 && ne < (1LL << 28), "mesh_create: element count must be in [1, 2^28)");
    REQUIRE(nbdr >= 0 && nbdr <= 6 * ne, "mesh_create: boundary count must be in [0, 6 ne]");
    for (int64_t i = 0; i < 8 * ne; ++i)
        REQUIRE(ev[i] >= 0 && ev[i] < nv, "mesh_create: element vertex index out of range");
    for (int64_t i = 0; i < 4 * nbdr; ++i)
        REQUIRE(bdr[i] >= 0 && bdr[i] < nv, "mesh_create: boundary vertex index out of range");
    for (int64_t i = 0; i < 3 * nv; ++i)
        REQUIRE(std::isfinite(xyz[i]), "mesh_create: vertex coordinates must be finite");
    auto m = std::make_unique<h1pa_mesh>();
    m->m.handle = m.get();
    m->m.from_arrays(nv, xyz, ne, ev, nbdr, bdr);
    *mesh = m.release();
    API_END
}

int h1pa_mesh_set_nodes(h1pa_mesh *mesh, int order, const double *d_nodes, void *stream) {
    API_BEGIN
    REQUIRE(mesh && d_nodes, "mesh_set_nodes: null argument");
    REQUIRE(order >= 1 && order <= 8, "mesh_set_nodes: order must be in [1, 8]");
    Mesh &m = mesh->m;
    hipStream_t s = as_stream(stream);
    const int64_t Dg = order + 1, n = Dg * Dg * Dg * 3 * m.ne;
    m.ensure_device(s);
    if (m.d_nodes.n != (size_t)n) m.d_nodes.alloc(n);
    H1PA_HIP(hipMemcpyAsync(m.d_nodes.p, d_nodes, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    m.node_order = order;
    m.curve_amp = 0.0;
    m.nodes_valid = true;
    m.nodes_external = true;
    API_END
}

int h1pa_mesh_destroy(h1pa_mesh *mesh) {
    API_BEGIN
    delete mesh;
    API_END
}

int h1pa_mesh_perturb(h1pa_mesh *mesh, double amp, uint64_t seed) {
    API_BEGIN
    REQUIRE(mesh, "mesh: null");
    REQUIRE(amp >= 0.0 && amp < 0.5, "perturbation amplitude must be in [0, 0.5)");
    kern::perturb_vertices(mesh->m, amp, seed);
    API_END
}

int h1pa_mesh_set_curvature(h1pa_mesh *mesh, int order, double amp) {
    API_BEGIN
    REQUIRE(mesh, "mesh: null");
    REQUIRE(order >= 1 && order <= 8, "curvature order must be in [1, 8]");
    mesh->m.node_order = order;
    mesh->m.curve_amp = amp;
    mesh->m.nodes_valid = false;
    mesh->m.nodes_external = false;
    API_END
}

int h1pa_mesh_counts(const h1pa_mesh *mesh, int64_t *c) {
    API_BEGIN
    REQUIRE(mesh && c, "mesh_counts: null argument");
    c[0] = mesh->m.nv;
    c[1] = mesh->m.ne;
    c[2] = mesh->m.nedges;
    c[3] = mesh->m.nfaces;
    API_END
}

int h1pa_mesh_topology(const h1pa_mesh *mesh, int *ev, int *ee, int *eeo, int *ef, int *efo) {
    API_BEGIN
    REQUIRE(mesh, "mesh: null");
    const Mesh &m = mesh->m;
    if (ev) std::memcpy(ev, m.ev.data(), sizeof(int) * m.ev.size());
    if (ee) std::memcpy(ee, m.ee.data(), sizeof(int) * m.ee.size());
    if (eeo) std::memcpy(eeo, m.eeo.data(), sizeof(int) * m.eeo.size());
    if (ef) std::memcpy(ef, m.ef.data(), sizeof(int) * m.ef.size());
    if (efo) std::memcpy(efo, m.efo.data(), sizeof(int) * m.efo.size());
    API_END
}

int h1pa_mesh_vertices(const h1pa_mesh *mesh, double *xyz) {
    API_BEGIN
    REQUIRE(mesh && xyz, "mesh_vertices: null argument");
    std::memcpy(xyz, mesh->m.xyz.data(), sizeof(double) * mesh->m.xyz.size());
    API_END
}

int h1pa_mesh_geometric_factors(const h1pa_mesh *mesh, int q1d, unsigned flags, double *X,
                                double *J, double *detJ, void *stream) {
    API_BEGIN
    REQUIRE(mesh, "mesh: null");
    hipStream_t s = as_stream(stream);
    Mesh &m = const_cast<h1pa_mesh *>(mesh)->m;
    m.ensure_nodes(s);
    kern::geometric_factors(m, q1d, (flags & H1PA_QI_VALUES) ? X : nullptr,
                            (flags & H1PA_QI_DERIVATIVES) ? J : nullptr,
                            (flags & H1PA_QI_DETERMINANTS) ? detJ : nullptr, s);
    API_END
}

int h1pa_space_qinterp(const h1pa_space *space, int q1d, int vdim, unsigned flags, int layout,
                       const double *e, double *val, double *der, double *det, void *stream) {
    API_BEGIN
    REQUIRE(space && e, "space_qinterp: null argument");
    Space &sp = const_cast<h1pa_space *>(space)->s;
    REQUIRE(kern::qinterp_supported(sp.p, q1d), "space_qinterp: q1d must be in [p+1, min(p+3, 10)]");
    REQUIRE(vdim >= 1 && vdim <= 3, "space_qinterp: vdim must be 1, 2 or 3");
    REQUIRE(layout == H1PA_QI_BY_NODES || layout == H1PA_QI_BY_VDIM,
            "space_qinterp: layout must be H1PA_QI_BY_NODES or H1PA_QI_BY_VDIM");
    REQUIRE((flags & ~15u) == 0, "space_qinterp: unknown flag");
    REQUIRE(!((flags & H1PA_QI_DERIVATIVES) && (flags & H1PA_QI_PHYSICAL_DERIVATIVES)),
            "space_qinterp: DERIVATIVES and PHYSICAL_DERIVATIVES are exclusive");
    REQUIRE(!(flags & H1PA_QI_VALUES) || val, "space_qinterp: VALUES needs d_val");
    REQUIRE(!(flags & (H1PA_QI_DERIVATIVES | H1PA_QI_PHYSICAL_DERIVATIVES)) || der,
            "space_qinterp: derivatives need d_der");
    REQUIRE(!(flags & H1PA_QI_DETERMINANTS) || (det && vdim == 3),
            "space_qinterp: DETERMINANTS needs d_det and vdim = 3");
    hipStream_t s = as_stream(stream);
    kern::qinterp(sp, q1d, vdim, flags, layout == H1PA_QI_BY_VDIM, e, val, der, det, s);
    H1PA_HIP(hipStreamSynchronize(s));
    API_END
}

// ---- space ----------------------------------------------------------------
int h1pa_space_create(h1pa_mesh *mesh, int order, h1pa_space **space) {
    API_BEGIN
    REQUIRE(mesh && space, "space_create: null argument");
    REQUIRE(order >= 1 && order <= 8, "order must be in [1, 8]");
    Mesh &m = mesh->m;
    auto *sp = new h1pa_space();
    try {
        Space &s = sp->s;
        s.mesh = &m;
        s.p = order;
        s.ne = m.ne;
        const int64_t D = order + 1, pm1 = order - 1;
        s.dof = D * D * D;
        s.nvdofs = m.nv;
        s.nedofs = m.nedges * pm1;
        s.nfdofs = m.nfaces * pm1 * pm1;
        s.ndofs = s.nvdofs + s.nedofs + s.nfdofs + m.ne * pm1 * pm1 * pm1;
        REQUIRE(s.ndofs < (1LL << 31) && s.ne * s.dof < (1LL << 31), "space too large for int32 maps");
        hipStream_t st = nullptr;
        m.ensure_device(st);
        s.elem_dof.alloc(s.ne * s.dof);
        Restriction &r = s.restr;
        r.ne = s.ne;
        r.dof = s.dof;
        r.ndofs = s.ndofs;
        r.gather_map.alloc(s.ne * s.dof);
        r.offsets.alloc(s.ndofs + 1);
        r.indices.alloc(s.ne * s.dof);
        kern::build_gather_map(m, order, s.nvdofs, s.nedofs, s.nfdofs, s.elem_dof.p,
                               r.gather_map.p, st);
        kern::build_transpose(s.ne, (int)s.dof, s.ndofs, r.gather_map.p, r.offsets.p, r.indices.p, st);
        s.ess_mask.alloc(s.ndofs);
        kern::boundary_mask(m, order, r.gather_map.p, s.ess_mask.p, s.ndofs, st);
        s.n_ess = kern::mask_to_list(s.ess_mask.p, s.ndofs, s.ess, st);
        // vertex block scattered in first-contribution order; bubble block
        // (interior DOFs) written directly by the element kernels
        kern::first_contribution_order(s.ne * s.dof, r.gather_map.p, r.offsets.p, r.indices.p,
                                       s.nvdofs, r.perm, st);
        r.nperm = s.nvdofs;
        r.bubble_begin = s.nvdofs + s.nedofs + s.nfdofs;
        // non-bubble rows in first-contribution order (bubble rows too were
        // measured slower at p = 4, 6, 8 -- 0.394 vs 0.360 ms at p = 4, 32M
        // DOFs -- and faster only at p = 2)
        r.nperm_lex = r.bubble_begin;
        kern::first_contribution_order(s.ne * s.dof, r.gather_map.p, r.offsets.p, r.indices.p,
                                       r.nperm_lex, r.perm_lex, st);
        {
            std::vector<int> dm(s.dof);
            h1_dof_map(order, dm.data());  // lex -> native
            r.lex2nat.upload(dm.data(), dm.size(), st);
            r.indices_nat.alloc(s.ne * s.dof);
            kern::native_indices(s.ne * s.dof, (int)s.dof, r.indices.p, r.lex2nat.p,
                                 r.indices_nat.p, st);
            H1PA_HIP(hipStreamSynchronize(st));
        }
        auto *rh = new h1pa_restriction();
        rh->owned_by_space = true;
        sp->r = rh;
    } catch (...) {
        delete sp;
        throw;
    }
    *space = sp;
    API_END
}

int h1pa_space_destroy(h1pa_space *space) {
    API_BEGIN
    if (space) delete space->r;
    delete space;
    API_END
}

int h1pa_space_info(const h1pa_space *space, int64_t *info) {
    API_BEGIN
    REQUIRE(space && info, "space_info: null argument");
    info[0] = space->s.p;
    info[1] = space->s.ne;
    info[2] = space->s.dof;
    info[3] = space->s.ndofs;
    API_END
}

int h1pa_space_element_dofs(const h1pa_space *space, const int **d) {
    API_BEGIN
    REQUIRE(space && d, "space_element_dofs: null argument");
    *d = space->s.elem_dof.p;
    API_END
}

int h1pa_space_boundary_dofs(const h1pa_space *space, const int **d_ess, int64_t *n) {
    API_BEGIN
    REQUIRE(space && d_ess && n, "space_boundary_dofs: null argument");
    *d_ess = space->s.ess.p;
    *n = space->s.n_ess;
    API_END
}

}  // extern "C"

extern "C" {

int h1pa_space_restriction(h1pa_space *space, h1pa_restriction **r) {
    API_BEGIN
    REQUIRE(space && r, "space_restriction: null argument");
    // The borrowed handle is a view of the space's device maps.
    space->r->view = &space->s.restr;
    *r = space->r;
    API_END
}

// ---- restriction --------------------------------------------------------------
int h1pa_restriction_create(int ne, int dof, int ndofs, const int *h_elem_dof, const int *h_dof_map,
                            h1pa_restriction **restriction) {
    API_BEGIN
    REQUIRE(restriction && h_elem_dof, "restriction_create: null argument");
    REQUIRE(ne >= 0 && dof >= 1 && ndofs >= 0, "restriction_create: bad sizes");
    for (int64_t i = 0; i < (int64_t)ne * dof; ++i) {
        const int v = h_elem_dof[i], g = v >= 0 ? v : -1 - v;
        REQUIRE(g < ndofs, "restriction_create: element DOF out of range");
    }
    if (h_dof_map) {
        std::vector<char> seen(dof, 0);
        for (int d = 0; d < dof; ++d) {
            const int v = h_dof_map[d], g = v >= 0 ? v : -1 - v;
            REQUIRE(g < dof && !seen[g], "restriction_create: dof_map is not a permutation");
            seen[g] = 1;
        }
    }
    auto *r = new h1pa_restriction();
    try {
        Restriction &R = r->own;
        R.ne = ne;
        R.dof = dof;
        R.ndofs = ndofs;
        hipStream_t s = nullptr;
        StreamBuf<int> ed, dm;
        ed.upload(h_elem_dof, (size_t)ne * dof, s);
        if (h_dof_map) dm.upload(h_dof_map, dof, s);
        R.gather_map.alloc((size_t)ne * dof);
        R.offsets.alloc(ndofs + 1);
        R.indices.alloc((size_t)ne * dof);
        kern::gather_from_table(ne, dof, ed.p, h_dof_map ? dm.p : nullptr, R.gather_map.p, s);
        kern::build_transpose(ne, dof, ndofs, R.gather_map.p, R.offsets.p, R.indices.p, s);
        R.nperm = 0;
        R.bubble_begin = ndofs;
    } catch (...) {
        delete r;
        throw;
    }
    *restriction = r;
    API_END
}

int h1pa_restriction_destroy(h1pa_restriction *r) {
    API_BEGIN
    if (r && r->owned_by_space) {
        throw Error{H1PA_EINVAL, "restriction is owned by its space"};
    }
    delete r;
    API_END
}

int h1pa_restriction_maps(const h1pa_restriction *r, const int **g, const int **o, const int **i) {
    API_BEGIN
    REQUIRE(r, "restriction: null");
    if (g) *g = r->R().gather_map.p;
    if (o) *o = r->R().offsets.p;
    if (i) *i = r->R().indices.p;
    API_END
}

int h1pa_restriction_mult(const h1pa_restriction *r, const double *x, double *y, void *stream) {
    API_BEGIN
    REQUIRE(r && x && y, "restriction_mult: null argument");
    kern::restrict_mult(r->R().ne * r->R().dof, r->R().gather_map.p, x, y, false, as_stream(stream));
    API_END
}

int h1pa_restriction_abs_mult(const h1pa_restriction *r, const double *x, double *y, void *stream) {
    API_BEGIN
    REQUIRE(r && x && y, "restriction_abs_mult: null argument");
    kern::restrict_mult(r->R().ne * r->R().dof, r->R().gather_map.p, x, y, true, as_stream(stream));
    API_END
}

int h1pa_restriction_mult_transpose(const h1pa_restriction *r, const double *x, double *y,
                                    void *stream) {
    API_BEGIN
    REQUIRE(r && x && y, "restriction_mult_transpose: null argument");
    restriction_scatter(r->R(), r->view != nullptr, x, y, 0, as_stream(stream));
    API_END
}

int h1pa_restriction_add_mult_transpose(const h1pa_restriction *r, const double *x, double *y,
                                        double a, void *stream) {
    API_BEGIN
    REQUIRE(r && x && y, "restriction_add_mult_transpose: null argument");
    REQUIRE(a == 1.0, "General coefficient case is not yet supported!");
    restriction_scatter(r->R(), r->view != nullptr, x, y, 1, as_stream(stream));
    API_END
}

int h1pa_restriction_abs_mult_transpose(const h1pa_restriction *r, const double *x, double *y,
                                        void *stream) {
    API_BEGIN
    REQUIRE(r && x && y, "restriction_abs_mult_transpose: null argument");
    restriction_scatter(r->R(), r->view != nullptr, x, y, 2, as_stream(stream));
    API_END
}

int h1pa_restriction_mult_left_inverse(const h1pa_restriction *r, const double *x, double *y,
                                       void *stream) {
    API_BEGIN
    REQUIRE(r && x && y, "restriction_mult_left_inverse: null argument");
    restriction_scatter(r->R(), r->view != nullptr, x, y, 3, as_stream(stream));
    API_END
}

int h1pa_restriction_boolean_mask(const h1pa_restriction *r, double *ye, void *stream) {
    API_BEGIN
    REQUIRE(r && ye, "restriction_boolean_mask: null argument");
    kern::boolean_mask(r->R().ndofs, r->R().offsets.p, r->R().indices.p, ye, as_stream(stream));
    API_END
}

// ---- form -------------------------------------------------------------------
int h1pa_form_create(h1pa_space *space, const h1pa_form_desc *desc, h1pa_form **form) {
    API_BEGIN
    REQUIRE(space && desc && form, "form_create: null argument");
    REQUIRE(desc->integrators != 0 &&
                (desc->integrators & ~(unsigned)(H1PA_INTEG_MASS | H1PA_INTEG_DIFFUSION)) == 0,
            "form_create: integrators must be a non-empty subset of MASS|DIFFUSION");
    REQUIRE(desc->coeff_kind >= H1PA_COEFF_CONSTANT && desc->coeff_kind <= H1PA_COEFF_FOURIER,
            "form_create: unknown coefficient kind");
    REQUIRE(!(desc->coeff_kind == H1PA_COEFF_PER_ELEMENT || desc->coeff_kind == H1PA_COEFF_QUADRATURE) ||
                desc->d_coeff, "form_create: coefficient data required");
    const int p = space->s.p;
    const int q1d = desc->q1d > 0 ? desc->q1d : p + 2;
    H1PA_CHECK(kern::apply_supported(p, q1d), H1PA_EUNSUP, "form_create: q1d must be p+1 or p+2");
    auto *f = new h1pa_form();
    f->f.space = &space->s;
    f->f.desc = *desc;
    f->f.q1d = q1d;
    *form = f;
    API_END
}

int h1pa_set_device_cache(int64_t max_bytes) {
    API_BEGIN
    REQUIRE(max_bytes >= 0, "set_device_cache: max_bytes must be >= 0");
    device_cache_set(max_bytes);
    API_END
}

int h1pa_device_cache_bytes(int64_t *bytes) {
    API_BEGIN
    REQUIRE(bytes, "device_cache_bytes: null argument");
    *bytes = device_cache_bytes();
    API_END
}

int h1pa_form_destroy(h1pa_form *form) {
    API_BEGIN
    if (form) require_no_batch_refs(form->f, "form_destroy");
    delete form;
    API_END
}

}  // extern "C"

namespace {
// Per-(p, q1d) choice among the diffusion-apply variants, made once per
// process on the first form large enough to time reliably.  All variants
// compute bit-identical results.  H1PA_DIFF_VARIANT=<v> forces a variant,
// H1PA_AUTOTUNE=0 keeps the default.
// g_tuned and the tuning itself are serialized: forms may be assembled on
// several host threads at once.
bool g_tuned[10][11] = {};
std::mutex g_tune_mutex;

void autotune_diffusion(Form &f, hipStream_t s) {
    Space &sp = *f.space;
    const int p = sp.p, Q = f.q1d;
    if (const char *v = std::getenv("H1PA_DIFF_VARIANT")) {
        const int fv = std::atoi(v);
        H1PA_CHECK(fv >= 0 && fv < kern::diffusion_variant_count(), H1PA_EINVAL,
                   "H1PA_DIFF_VARIANT: no such diffusion-apply variant");
        kern::set_diffusion_variant(p, Q, fv);
        return;
    }
    const char *at = std::getenv("H1PA_AUTOTUNE");
    // large enough to time reliably: >= 20000 elements or >= 500k DOFs
    if ((at && at[0] == '0') || (sp.ne < 20000 && sp.ndofs < 500000)) return;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (g_tuned[p][Q]) return;
    // x and an L-vector for the interior DOFs; once per (p, Q) and process
    DevBuf<double> x(2 * sp.ndofs);
    H1PA_HIP(hipMemsetAsync(x.p, 0, sizeof(double) * 2 * sp.ndofs, s));
    hipEvent_t e0, e1;
    H1PA_HIP(hipEventCreate(&e0));
    H1PA_HIP(hipEventCreate(&e1));
    // the Mult's own output: native layout, interior DOFs straight to an L-vector
    kern::EOut out;
    out.ye = f.ework0.p;
    out.l2n = sp.restr.lex2nat.p;
    if (sp.restr.bubble_begin < sp.ndofs) {
        out.yL = x.p + sp.ndofs;
        out.bb = sp.restr.bubble_begin;
    }
    // three interleaved rounds (1 warm-up + 4 timed launches per variant each),
    // so that a drift of the clock or of the other work on the device does
    // not pick the variant that happened to run in a quiet moment
    const int nv = kern::diffusion_variant_count();
    std::vector<float> total(nv, 0.f);
    for (int round = 0; round < 3; ++round)
        for (int v = 0; v < nv; ++v) {
            kern::diffusion_apply_variant(v, p, Q, sp.ne, f.basis.p, f.pa_diff.p,
                                          sp.restr.gather_map.p, nullptr, x.p, out, false, s);
            H1PA_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < 4; ++r)
                kern::diffusion_apply_variant(v, p, Q, sp.ne, f.basis.p, f.pa_diff.p,
                                              sp.restr.gather_map.p, nullptr, x.p, out, false, s);
            H1PA_HIP(hipEventRecord(e1, s));
            H1PA_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            H1PA_HIP(hipEventElapsedTime(&ms, e0, e1));
            total[v] += ms;
        }
    const int best = (int)(std::min_element(total.begin(), total.end()) - total.begin());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    kern::set_diffusion_variant(p, Q, best);
    g_tuned[p][Q] = true;
}
}  // namespace

extern "C" {

int h1pa_form_assemble(h1pa_form *form, void *stream) {
    API_BEGIN
    REQUIRE(form, "form: null");
    Form &f = form->f;
    require_no_batch_refs(f, "form_assemble");
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    sp.mesh->ensure_nodes(s);
    const int Q = f.q1d, D = sp.p + 1;
    std::vector<double> xq(Q), wq(Q), B(Q * D), G(Q * D);
    gauss_legendre(Q, xq.data(), wq.data());
    lagrange_basis(sp.p, Q, xq.data(), B.data(), G.data());
    // packed basis: B, G with q fastest, then their transposes (d fastest),
    // then the even-odd tables of the apply kernels
    std::vector<double> pack(4 * Q * D + even_odd_size(D, Q));
    for (int q = 0; q < Q; ++q)
        for (int d = 0; d < D; ++d) {
            pack[q + Q * d] = B[q + Q * d];
            pack[Q * D + q + Q * d] = G[q + Q * d];
            pack[2 * Q * D + d + D * q] = B[q + Q * d];
            pack[3 * Q * D + d + D * q] = G[q + Q * d];
        }
    even_odd_tables(D, Q, B.data(), G.data(), pack.data() + 4 * Q * D);
    f.basis.upload(pack.data(), pack.size(), s);
    const int64_t nq = (int64_t)Q * Q * Q;
    if (f.desc.integrators & H1PA_INTEG_DIFFUSION) f.pa_diff.alloc(nq * 6 * sp.ne);
    else f.pa_diff.release();
    if (f.desc.integrators & H1PA_INTEG_MASS) f.pa_mass.alloc(nq * sp.ne);
    else f.pa_mass.release();
    kern::pa_setup(*sp.mesh, f, s);
    f.ework0.alloc(sp.ne * sp.dof);
    H1PA_HIP(hipStreamSynchronize(s));
    f.assembled = true;
    if (f.desc.integrators & H1PA_INTEG_DIFFUSION) autotune_diffusion(f, s);
    API_END
}

}  // extern "C"

namespace {
void require_assembled(const Form &f) {
    H1PA_CHECK(f.assembled, H1PA_EINVAL, "form: Assemble() has not been called");
}

// E-part of the operator: ework0 = sum of integrators applied to x.  `gmap`
// null means x is already an E-vector.
// yL != nullptr selects the fused layout: bubble DOFs straight into yL and
// the other E-entries in native order (read back through indices_nat).
void element_apply(Form &f, const int *gmap, const uint8_t *mask, const double *x, double *ye,
                   double *yL, bool accumulate, hipStream_t s) {
    kern::EOut out;
    out.ye = ye;
    if (yL) {
        out.yL = yL;
        out.bb = f.space->restr.bubble_begin;
        out.l2n = f.space->restr.lex2nat.p;
    }
    Space &sp = *f.space;
    const bool diff = f.desc.integrators & H1PA_INTEG_DIFFUSION;
    const bool mass = f.desc.integrators & H1PA_INTEG_MASS;
    if (diff)
        kern::diffusion_apply(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_diff.p, gmap, mask, x, out,
                              accumulate, s);
    if (mass)
        kern::mass_apply(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_mass.p, gmap, mask, x, out,
                         accumulate || diff, s);
}

// Whether Mult takes the fused column apply (columns.hip): a Cartesian mesh
// (grid positions known), one diffusion integrator at q1d = p + 2, and
// H1PA_COLUMNS=1 in the environment.  Off by default: measured on the box it
// is slower than the element kernel + scatter at every order but p = 2 (32M
// DOFs, p = 4: 2.73 vs 1.40 ms per Mult).  The grid and the interface buffer
// are set up on the first call.
bool columns_ready(Form &f, hipStream_t s) {
    if (f.use_columns == 0) return false;
    Space &sp = *f.space;
    Mesh &m = *sp.mesh;
    if (f.use_columns < 0) {
        const char *env = std::getenv("H1PA_COLUMNS");
        const bool ok = env && env[0] == '1' && !m.from_file && m.nx > 0 &&
                        f.desc.integrators == H1PA_INTEG_DIFFUSION &&
                        kern::columns_supported(sp.p, f.q1d) && sp.ne == (int64_t)m.nx * m.ny * m.nz;
        if (!ok) {
            f.use_columns = 0;
            return false;
        }
        if (!m.d_elem_at.p) {
            std::vector<int> at(m.ne);
            for (int64_t e = 0; e < m.ne; ++e)
                at[((int64_t)m.ecoord[3 * e + 2] * m.ny + m.ecoord[3 * e + 1]) * m.nx +
                   m.ecoord[3 * e]] = (int)e;
            m.d_elem_at.upload(at.data(), at.size(), s);
            H1PA_HIP(hipStreamSynchronize(s));
        }
        int A = 0, B = 0;
        kern::column_shape(sp.p, A, B);
        ColumnGrid &g = f.cols;
        g.nx = m.nx;
        g.ny = m.ny;
        g.nz = m.nz;
        g.npx = (m.nx + A - 1) / A;
        g.npy = (m.ny + B - 1) / B;
        // about 10 work-groups per CU: enough segments per column to fill the
        // chip, as few as that allows (each segment boundary is shared)
        const int64_t cols = (int64_t)g.npx * g.npy;
        int nseg = (int)std::min<int64_t>(m.nz, std::max<int64_t>(1, (2560 + cols - 1) / cols));
        g.ls = (m.nz + nseg - 1) / nseg;
        g.nseg = (m.nz + g.ls - 1) / g.ls;
        g.elem_at = m.d_elem_at.p;
        f.iface.alloc((size_t)kern::column_interface_size(sp.p, g));
        f.use_columns = 1;
    }
    return true;
}

const uint8_t *constraint_mask(Form &f, const int *d_ess, int64_t n_ess, hipStream_t s) {
    Space &sp = *f.space;
    if (d_ess == sp.ess.p && n_ess == sp.n_ess) return sp.ess_mask.p;
    // rebuilt on every call: the caller may have changed the list in place
    if (!f.cmask.p) f.cmask.alloc(sp.ndofs);
    kern::list_to_mask(d_ess, n_ess, f.cmask.p, sp.ndofs, s);
    return f.cmask.p;
}

}  // namespace

// MultTranspose family on a lexicographic E-vector.  A space's restriction
// (bubble_block: MFEM's interior-DOF block from bubble_begin on) sums its
// non-bubble rows through the transpose table in element-major order and
// writes the bubble block by position; other restrictions use the table for
// every row.
void h1pa::restriction_scatter(const Restriction &R, bool bubble_block, const double *x,
                               double *y, int mode, hipStream_t s) {
    if (bubble_block && R.bubble_begin < R.ndofs) {
        int D = 1;
        while ((int64_t)D * D * D < R.dof) ++D;
        H1PA_CHECK((int64_t)D * D * D == R.dof && R.nperm_lex == R.bubble_begin &&
                       R.bubble_begin + R.ne * (int64_t)(D - 2) * (D - 2) * (D - 2) == R.ndofs,
                   H1PA_EINVAL, "restriction: unexpected bubble block");
        kern::restrict_mult_transpose(R.bubble_begin, R.offsets.p, R.indices.p, x, y, mode,
                                      R.perm_lex.p, R.nperm_lex, s);
        kern::bubble_mult_transpose(R.ne, D, R.bubble_begin, x, y, mode, s);
        return;
    }
    kern::restrict_mult_transpose(R.ndofs, R.offsets.p, R.indices.p, x, y, mode, R.perm_lex.p,
                                  R.nperm_lex, s);
}

// The default pool of each device keeps the blocks StreamBuf frees (release
// threshold at its maximum), so a temporary's hipMallocAsync after the first
// reuses memory instead of going back to the driver.
void h1pa::keep_stream_pool() {
    static std::atomic<uint64_t> done{0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return;
    }
    const uint64_t bit = uint64_t(1) << dev;
    if (done.load(std::memory_order_acquire) & bit) return;
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
        uint64_t threshold = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
    }
    (void)hipGetLastError();
    done.fetch_or(bit, std::memory_order_acq_rel);
}

// Jacobi diagonal of the unconstrained operator into diag (L-vector): the
// element diagonals in the fused layout (bubble DOFs straight into diag, the
// rest in native order), then AbsMultTranspose of the other rows.
void h1pa::operator_diagonal(Form &f, double *diag, hipStream_t s) {
    Space &sp = *f.space;
    const bool diff = f.desc.integrators & H1PA_INTEG_DIFFUSION;
    const bool mass = f.desc.integrators & H1PA_INTEG_MASS;
    kern::EOut out;
    out.ye = f.ework0.p;
    out.yL = diag;
    out.bb = sp.restr.bubble_begin;
    out.l2n = sp.restr.lex2nat.p;
    if (diff) kern::diffusion_diagonal(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_diff.p, out, false, s);
    if (mass) kern::mass_diagonal(sp.p, f.q1d, sp.ne, f.basis.p, f.pa_mass.p, out, diff, s);
    kern::restrict_mult_transpose(sp.restr.bubble_begin, sp.restr.offsets.p,
                                  sp.restr.indices_nat.p, f.ework0.p, diag, 2, sp.restr.perm.p,
                                  sp.restr.nperm, s);
}

int64_t h1pa::form_device_bytes(const Form &f) {
    const Space &sp = *f.space;
    const Mesh &m = *sp.mesh;
    int64_t b = 0;
    b += f.basis.n * 8 + f.pa_diff.n * 8 + f.pa_mass.n * 8 + f.ework0.n * 8 + f.ework1.n * 8;
    b += f.lwork0.n * 8 + f.lwork1.n * 8 + f.cmask.n + f.basis32.n * 4 + f.iface.n * 8;
    b += f.pa_diff_lp.n * 4 + f.pa_mass_lp.n * 4 + f.ework32.n * 4;
    const Restriction &r = sp.restr;
    b += (r.gather_map.n + r.offsets.n + r.indices.n + r.perm.n + r.perm_lex.n + r.lex2nat.n +
          r.indices_nat.n) * 4;
    b += (sp.elem_dof.n + sp.ess.n) * 4 + sp.ess_mask.n;
    b += m.d_nodes.n * 8 + m.d_topo.n * 4 + m.d_ori.n + m.d_elem_at.n * 4;
    return b;
}

namespace {
}  // namespace

extern "C" {

int h1pa_form_mult(h1pa_form *form, const double *x, double *y, void *stream) {
    API_BEGIN
    REQUIRE(form && x && y, "form_mult: null argument");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    if (columns_ready(f, s)) {
        kern::diffusion_columns(sp.p, f.cols, f.basis.p, f.pa_diff.p, sp.restr.gather_map.p, x, y,
                                f.iface.p, 3, s);
    } else {
        element_apply(f, sp.restr.gather_map.p, nullptr, x, f.ework0.p, y, false, s);
        kern::restrict_mult_transpose(sp.restr.bubble_begin, sp.restr.offsets.p,
                                      sp.restr.indices_nat.p, f.ework0.p, y, 0, sp.restr.perm.p,
                                      sp.restr.nperm, s);
    }
    API_END
}

int h1pa_form_add_mult(h1pa_form *form, const double *x, double *y, double a, void *stream) {
    API_BEGIN
    REQUIRE(form && x && y, "form_add_mult: null argument");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    if (!f.lwork0.p) f.lwork0.alloc(sp.ndofs);
    if (columns_ready(f, s)) {
        kern::diffusion_columns(sp.p, f.cols, f.basis.p, f.pa_diff.p, sp.restr.gather_map.p, x,
                                f.lwork0.p, f.iface.p, 3, s);
    } else {
        element_apply(f, sp.restr.gather_map.p, nullptr, x, f.ework0.p, f.lwork0.p, false, s);
        kern::restrict_mult_transpose(sp.restr.bubble_begin, sp.restr.offsets.p,
                                      sp.restr.indices_nat.p, f.ework0.p, f.lwork0.p, 0,
                                      sp.restr.perm.p, sp.restr.nperm, s);
    }
    kern::axpy(sp.ndofs, a, f.lwork0.p, y, s);
    API_END
}

}  // extern "C"

// ---- the apply that allocates nothing ---------------------------------------
// Scratch E-vector of h1pa_form_mult_noalloc: a file-scope device array (one
// copy per GPU, held until the process ends).  Host code never passes the
// array's name to a kernel: its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and kept in g_scratch.
namespace {
constexpr int64_t kScratchDoubles = int64_t(8) << 20;  // 64 MiB
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
double *scratch_on_current_device() {
    int dev = 0;
    H1PA_HIP(hipGetDevice(&dev));
    H1PA_CHECK(dev >= 0 && dev < kMaxDevices, H1PA_EINVAL, "form_mult_noalloc: device id too large");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        H1PA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        H1PA_CHECK(ptr != nullptr, H1PA_EHIP, "form_mult_noalloc: no address for the scratch array");
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    return g_scratch[dev];
}
}  // namespace

extern "C" {

int h1pa_form_mult_noalloc(h1pa_form *form, const double *x, double *y, void *stream) {
    API_BEGIN
    REQUIRE(form && x && y, "form_mult_noalloc: null argument");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    REQUIRE(sp.ne * sp.dof <= kScratchDoubles,
            "form_mult_noalloc: the E-vector (" + std::to_string(sp.ne * sp.dof) +
                " entries) exceeds the scratch of " + std::to_string(kScratchDoubles) +
                " entries");
    hipStream_t s = as_stream(stream);
    double *ye = scratch_on_current_device();
    // the element path of h1pa_form_mult, with its E-vector in the scratch:
    // every E-entry the scatter reads is written first (accumulate = false)
    element_apply(f, sp.restr.gather_map.p, nullptr, x, ye, y, false, s);
    kern::restrict_mult_transpose(sp.restr.bubble_begin, sp.restr.offsets.p,
                                  sp.restr.indices_nat.p, ye, y, 0, sp.restr.perm.p,
                                  sp.restr.nperm, s);
    API_END
}

int h1pa_form_mult_timed(h1pa_form *form, const double *x, double *y, int steps, double *ms_total,
                         double *ms_element, double *ms_scatter, void *stream) {
    API_BEGIN
    REQUIRE(form && x && y && steps >= 1, "form_mult_timed: bad argument");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    std::vector<hipEvent_t> ev(3 * (size_t)steps);
    for (auto &e : ev) H1PA_HIP(hipEventCreate(&e));
    try {
        const bool cols = columns_ready(f, s);
        for (int k = 0; k < steps; ++k) {
            H1PA_HIP(hipEventRecord(ev[3 * k], s));
            if (cols)
                kern::diffusion_columns(sp.p, f.cols, f.basis.p, f.pa_diff.p,
                                        sp.restr.gather_map.p, x, y, f.iface.p, 1, s);
            else
                element_apply(f, sp.restr.gather_map.p, nullptr, x, f.ework0.p, y, false, s);
            H1PA_HIP(hipEventRecord(ev[3 * k + 1], s));
            if (cols)
                kern::diffusion_columns(sp.p, f.cols, f.basis.p, f.pa_diff.p,
                                        sp.restr.gather_map.p, x, y, f.iface.p, 2, s);
            else
                kern::restrict_mult_transpose(sp.restr.bubble_begin, sp.restr.offsets.p,
                                              sp.restr.indices_nat.p, f.ework0.p, y, 0,
                                              sp.restr.perm.p, sp.restr.nperm, s);
            H1PA_HIP(hipEventRecord(ev[3 * k + 2], s));
        }
        H1PA_HIP(hipEventSynchronize(ev.back()));
        double te = 0.0, ts = 0.0;
        for (int k = 0; k < steps; ++k) {
            float a = 0.f, b = 0.f;
            H1PA_HIP(hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]));
            H1PA_HIP(hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]));
            te += a;
            ts += b;
        }
        float tot = 0.f;
        H1PA_HIP(hipEventElapsedTime(&tot, ev[0], ev.back()));
        if (ms_total) *ms_total = tot;
        if (ms_element) *ms_element = te / steps;
        if (ms_scatter) *ms_scatter = ts / steps;
    } catch (...) {
        for (auto &e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    API_END
}

int h1pa_form_add_mult_e(h1pa_form *form, const double *xE, double *yE, void *stream) {
    API_BEGIN
    REQUIRE(form && xE && yE, "form_add_mult_e: null argument");
    Form &f = form->f;
    require_assembled(f);
    element_apply(f, nullptr, nullptr, xE, yE, nullptr, true, as_stream(stream));
    API_END
}

int h1pa_form_assemble_diagonal(h1pa_form *form, double *diag, void *stream) {
    API_BEGIN
    REQUIRE(form && diag, "form_assemble_diagonal: null argument");
    Form &f = form->f;
    require_assembled(f);
    operator_diagonal(f, diag, as_stream(stream));
    API_END
}

int h1pa_form_pa_data(const h1pa_form *form, const double **dd, const double **dm) {
    API_BEGIN
    REQUIRE(form, "form: null");
    if (dd) *dd = form->f.pa_diff.p;
    if (dm) *dm = form->f.pa_mass.p;
    API_END
}

int h1pa_form_info(const h1pa_form *form, int64_t *info) {
    API_BEGIN
    REQUIRE(form && info, "form_info: null argument");
    info[0] = form->f.q1d;
    info[1] = form->f.desc.integrators;
    info[2] = form->f.space->ndofs;
    info[3] = (form->f.desc.integrators & H1PA_INTEG_DIFFUSION)
                  ? kern::get_diffusion_variant(form->f.space->p, form->f.q1d)
                  : -1;
    API_END
}

int h1pa_form_mult_path(h1pa_form *form, int *path, void *stream) {
    API_BEGIN
    REQUIRE(form && path, "form_mult_path: null argument");
    require_assembled(form->f);
    *path = columns_ready(form->f, as_stream(stream)) ? 1 : 0;
    API_END
}

int h1pa_form_constrained_mult(h1pa_form *form, const int *d_ess, int64_t n_ess, const double *x,
                               double *y, void *stream) {
    API_BEGIN
    REQUIRE(form && x && y && (n_ess == 0 || d_ess), "form_constrained_mult: null argument");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    const uint8_t *mask = n_ess > 0 ? constraint_mask(f, d_ess, n_ess, s) : nullptr;
    // interior DOFs go straight to y unless the constraint list may hold some
    const bool direct = !mask || mask == sp.ess_mask.p;
    element_apply(f, sp.restr.gather_map.p, mask, x, f.ework0.p, direct ? y : nullptr, false, s);
    kern::scatter_constrained(direct ? sp.restr.bubble_begin : sp.ndofs, sp.restr.offsets.p,
                              direct ? sp.restr.indices_nat.p : sp.restr.indices.p, f.ework0.p,
                              mask, x, y, direct ? sp.restr.perm.p : sp.restr.perm_lex.p,
                              direct ? sp.restr.nperm : sp.restr.nperm_lex, s);
    API_END
}

int h1pa_form_linear_system(h1pa_form *form, const int *d_ess, int64_t n_ess, const double *x,
                            const double *b, double *X, double *B, void *stream) {
    API_BEGIN
    REQUIRE(form && x && b && X && B && (n_ess == 0 || d_ess), "form_linear_system: null argument");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    StreamBuf<uint8_t> zero_mask;
    const uint8_t *mask;
    if (n_ess > 0) {
        mask = constraint_mask(f, d_ess, n_ess, s);
    } else {
        zero_mask.alloc(sp.ndofs, s);
        H1PA_HIP(hipMemsetAsync(zero_mask.p, 0, sp.ndofs, s));
        mask = zero_mask.p;
    }
    StreamBuf<double> z(sp.ndofs, s);
    kern::form_x(sp.ndofs, mask, x, X, s);  // X = x on ess, 0 elsewhere (= w)
    element_apply(f, sp.restr.gather_map.p, nullptr, X, f.ework0.p, z.p, false, s);
    kern::restrict_mult_transpose(sp.restr.bubble_begin, sp.restr.offsets.p,
                                  sp.restr.indices_nat.p, f.ework0.p, z.p, 0, sp.restr.perm.p,
                                  sp.restr.nperm, s);
    kern::form_b(sp.ndofs, mask, X, b, z.p, B, s);
    H1PA_HIP(hipStreamSynchronize(s));
    API_END
}

int h1pa_vector_add(double *y, double a, const double *x, int64_t n, void *stream) {
    API_BEGIN
    REQUIRE(n >= 0 && (n == 0 || (x && y)), "vector_add: bad argument");
    kern::axpy(n, a, x, y, as_stream(stream));
    API_END
}

int h1pa_vector_randomize(double *x, int64_t n, uint64_t seed, double lo, double hi,
                          void *stream) {
    API_BEGIN
    REQUIRE(n >= 0 && (n == 0 || x), "vector_randomize: bad argument");
    kern::randomize(n, x, seed, lo, hi, as_stream(stream));
    API_END
}

int h1pa_space_domain_lf(h1pa_space *space, double fval, double *b, void *stream) {
    return h1pa_space_domain_lf_qf(space, fval, nullptr, b, stream);
}

int h1pa_space_domain_lf_qf(h1pa_space *space, double fval, const double *d_fq, double *b,
                            void *stream) {
    API_BEGIN
    REQUIRE(space && b, "space_domain_lf: null argument");
    Space &sp = space->s;
    hipStream_t s = as_stream(stream);
    sp.mesh->ensure_nodes(s);
    StreamBuf<double> bE(sp.ne * sp.dof, s);
    kern::domain_lf(*sp.mesh, sp.p, fval, d_fq, bE.p, s);
    restriction_scatter(sp.restr, true, bE.p, b, 0, s);
    H1PA_HIP(hipStreamSynchronize(s));
    API_END
}

int h1pa_pcg_solve(h1pa_form *form, const int *d_ess, int64_t n_ess, const double *b, double *x,
                   const h1pa_pcg_opts *opts, h1pa_pcg_stats *stats, void *stream) {
    API_BEGIN
    REQUIRE(form && b && x && opts && stats && (n_ess == 0 || d_ess), "pcg_solve: null argument");
    REQUIRE(opts->max_iter >= 1, "pcg_solve: max_iter must be >= 1");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    const uint8_t *mask = n_ess > 0 ? constraint_mask(f, d_ess, n_ess, s) : nullptr;
    StreamBuf<double> diag(sp.ndofs, s);
    operator_diagonal(f, diag.p, s);
    pcg_solve(f, mask, diag.p, b, x, *opts, *stats, s);
    if (!stats->converged) {
        set_error("PCG: No convergence!");
        return H1PA_ENOTCONV;
    }
    API_END
}

int h1pa_chebyshev_create(h1pa_form *form, const double *d_diag, const int *d_ess,
                          int64_t n_ess, int order, double max_eig_estimate, int power_iterations, double power_tolerance,
                          int power_seed, h1pa_chebyshev **cheb, void *stream) {
    API_BEGIN
    REQUIRE(form && cheb && (n_ess == 0 || d_ess), "chebyshev_create: null argument");
    REQUIRE(order >= 1 && order <= 5, "chebyshev_create: order must be 1..5");
    Form &f = form->f;
    require_assembled(f);
    REQUIRE(f.desc.integrators != 0, "chebyshev_create: form has no integrator");
    Space &sp = *f.space;
    hipStream_t s = as_stream(stream);
    auto h = std::make_unique<h1pa_chebyshev>();
    Chebyshev &c = h->c;
    c.f = &f;
    c.order = order;
    if (n_ess == 0) {
        c.mask = nullptr;
    } else if (d_ess == sp.ess.p && n_ess == sp.n_ess) {
        c.mask = sp.ess_mask.p;
    } else {  // own copy: the form's scratch mask is rebuilt by later calls
        c.own.alloc(sp.ndofs);
        kern::list_to_mask(d_ess, n_ess, c.own.p, sp.ndofs, s);
        c.mask = c.own.p;
    }
    chebyshev_setup(c, d_diag, max_eig_estimate, power_iterations, power_tolerance, power_seed,
                    s);
    H1PA_HIP(hipStreamSynchronize(s));
    *cheb = h.release();
    API_END
}

int h1pa_chebyshev_destroy(h1pa_chebyshev *cheb) {
    API_BEGIN
    delete cheb;
    API_END
}

int h1pa_chebyshev_info(const h1pa_chebyshev *cheb, double *max_eig, double *coeffs) {
    API_BEGIN
    REQUIRE(cheb, "chebyshev_info: null argument");
    if (max_eig) *max_eig = cheb->c.max_eig;
    if (coeffs)
        for (int k = 0; k < cheb->c.order; ++k) coeffs[k] = cheb->c.coeffs[k];
    API_END
}

int h1pa_chebyshev_mult(h1pa_chebyshev *cheb, const double *x, double *y, void *stream) {
    API_BEGIN
    REQUIRE(cheb && x && y, "chebyshev_mult: null argument");
    chebyshev_mult(cheb->c, x, y, as_stream(stream));
    API_END
}

int h1pa_pcg_solve_chebyshev(h1pa_chebyshev *cheb, const double *b, double *x,
                             const h1pa_pcg_opts *opts, h1pa_pcg_stats *stats, void *stream) {
    API_BEGIN
    REQUIRE(cheb && b && x && opts && stats, "pcg_solve_chebyshev: null argument");
    REQUIRE(opts->max_iter >= 1, "pcg_solve_chebyshev: max_iter must be >= 1");
    pcg_chebyshev(cheb->c, b, x, *opts, *stats, as_stream(stream));
    if (!stats->converged) {
        set_error("PCG: No convergence!");
        return H1PA_ENOTCONV;
    }
    API_END
}

int h1pa_form_mult_f32(h1pa_form *form, int precision, const float *x, float *y, void *stream) {
    API_BEGIN
    REQUIRE(form && x && y, "form_mult_f32: null argument");
    REQUIRE(precision == H1PA_PREC_FP32 || precision == H1PA_PREC_BF16_PA,
            "form_mult_f32: precision must be H1PA_PREC_FP32 or H1PA_PREC_BF16_PA");
    Form &f = form->f;
    require_assembled(f);
    H1PA_CHECK(f.q1d == f.space->p + 2, H1PA_EUNSUP,
               "form_mult_f32: reduced-precision applies support q1d = p + 2 only");
    form_mult_lp(f, precision, x, y, as_stream(stream));
    API_END
}

int h1pa_ir_solve(h1pa_form *form, const int *d_ess, int64_t n_ess, const double *b, double *x,
                  const h1pa_ir_opts *opts, h1pa_ir_stats *stats, void *stream) {
    API_BEGIN
    REQUIRE(form && b && x && opts && stats && (n_ess == 0 || d_ess), "ir_solve: null argument");
    REQUIRE(opts->inner_precision == H1PA_PREC_FP32 || opts->inner_precision == H1PA_PREC_BF16_PA,
            "ir_solve: inner_precision must be H1PA_PREC_FP32 or H1PA_PREC_BF16_PA");
    REQUIRE(opts->max_outer >= 1 && opts->max_inner >= 1,
            "ir_solve: max_outer and max_inner must be >= 1");
    REQUIRE(opts->rel_tol > 0.0 && opts->delta > 0.0 && opts->delta < 1.0,
            "ir_solve: need rel_tol > 0 and 0 < delta < 1");
    Form &f = form->f;
    require_assembled(f);
    Space &sp = *f.space;
    H1PA_CHECK(f.q1d == sp.p + 2, H1PA_EUNSUP,
               "ir_solve: reduced-precision applies support q1d = p + 2 only");
    hipStream_t s = as_stream(stream);
    const uint8_t *mask = n_ess > 0 ? constraint_mask(f, d_ess, n_ess, s) : nullptr;
    StreamBuf<double> diag(sp.ndofs, s);
    operator_diagonal(f, diag.p, s);
    ir_solve(f, mask, diag.p, b, x, *opts, *stats, s);
    if (!stats->converged) {
        set_error("IR: No convergence!");
        return H1PA_ENOTCONV;
    }
    API_END
}

int h1pa_batch_create(const h1pa_batch_opts *opts, h1pa_batch **batch) {
    API_BEGIN
    REQUIRE(opts && batch, "batch_create: null argument");
    *batch = batch_create(*opts);
    API_END
}

int h1pa_batch_destroy(h1pa_batch *batch) {
    API_BEGIN
    batch_destroy(batch);
    API_END
}

int h1pa_batch_add(h1pa_batch *batch, h1pa_form *form, const int *d_ess, int64_t n_ess,
                   const double *b, double *x, const h1pa_pcg_opts *opts, int64_t tag,
                   int *slot, void *stream) {
    API_BEGIN
    REQUIRE(batch && form && b && x && opts && (n_ess == 0 || d_ess), "batch_add: null argument");
    Form &f = form->f;
    require_assembled(f);
    hipStream_t cs = as_stream(stream);
    const uint8_t *mask = n_ess > 0 ? constraint_mask(f, d_ess, n_ess, cs) : nullptr;
    const int k = batch_add(*batch, form, mask, b, x, *opts, tag, cs);
    if (slot) *slot = k;
    API_END
}

int h1pa_batch_launch(h1pa_batch *batch, int *launched) {
    API_BEGIN
    REQUIRE(batch, "batch_launch: null argument");
    const bool l = batch_launch(*batch);
    if (launched) *launched = l ? 1 : 0;
    API_END
}

int h1pa_batch_wait(h1pa_batch *batch, int *n_running) {
    API_BEGIN
    REQUIRE(batch, "batch_wait: null argument");
    const int r = batch_wait(*batch);
    if (n_running) *n_running = r;
    API_END
}

int h1pa_batch_query(h1pa_batch *batch, int *idle) {
    API_BEGIN
    REQUIRE(batch && idle, "batch_query: null argument");
    *idle = batch_idle(*batch) ? 1 : 0;
    API_END
}

int h1pa_batch_run(h1pa_batch *batch, int max_replays, int *n_running) {
    API_BEGIN
    REQUIRE(batch && max_replays >= 0, "batch_run: bad argument");
    const int r = batch_run(*batch, max_replays);
    if (n_running) *n_running = r;
    API_END
}

int h1pa_batch_poll(h1pa_batch *batch, int cap, int64_t *tags, h1pa_batch_result *results,
                    int *n) {
    API_BEGIN
    REQUIRE(batch && n && cap >= 0, "batch_poll: bad argument");
    *n = batch_poll(*batch, cap, tags, results);
    API_END
}

int h1pa_batch_info(const h1pa_batch *batch, int64_t *info) {
    API_BEGIN
    REQUIRE(batch && info, "batch_info: null argument");
    batch_info(*batch, info);
    API_END
}

int h1pa_form_device_bytes(const h1pa_form *form, int64_t *bytes) {
    API_BEGIN
    REQUIRE(form && bytes, "form_device_bytes: null argument");
    *bytes = form_device_bytes(form->f);
    API_END
}


// ---- p-refinement transfer and p-multigrid ------------------------------------
int h1pa_transfer_create(h1pa_space *low, h1pa_space *high, h1pa_transfer **transfer) {
    API_BEGIN
    REQUIRE(low && high && transfer, "transfer_create: null argument");
    auto t = std::make_unique<h1pa_transfer>();
    t->t.reset(new Transfer(low->s, high->s, nullptr));
    *transfer = t.release();
    API_END
}

int h1pa_transfer_destroy(h1pa_transfer *transfer) {
    API_BEGIN
    delete transfer;
    API_END
}

int h1pa_transfer_mult(h1pa_transfer *transfer, const double *xl, double *yh, void *stream) {
    API_BEGIN
    REQUIRE(transfer && xl && yh, "transfer_mult: null argument");
    transfer->t->mult(xl, yh, as_stream(stream));
    API_END
}

int h1pa_transfer_mult_transpose(h1pa_transfer *transfer, const double *xh, double *yl,
                                 void *stream) {
    API_BEGIN
    REQUIRE(transfer && xh && yl, "transfer_mult_transpose: null argument");
    transfer->t->mult_transpose(xh, yl, as_stream(stream));
    API_END
}

int h1pa_multigrid_create(h1pa_form *fine, int nlevels, const int *orders,
                          const h1pa_mg_opts *opts, h1pa_multigrid **mg, void *stream) {
    API_BEGIN
    REQUIRE(fine && orders && opts && mg, "multigrid_create: null argument");
    auto m = std::make_unique<h1pa_multigrid>();
    m->mg.reset(multigrid_create(fine->f, nlevels, orders, *opts, as_stream(stream)));
    *mg = m.release();
    API_END
}

int h1pa_multigrid_destroy(h1pa_multigrid *mg) {
    API_BEGIN
    delete mg;
    API_END
}

int h1pa_multigrid_mult(h1pa_multigrid *mg, const double *x, double *y, void *stream) {
    API_BEGIN
    REQUIRE(mg && x && y, "multigrid_mult: null argument");
    multigrid_mult(*mg->mg, x, y, as_stream(stream));
    API_END
}

int h1pa_multigrid_info(const h1pa_multigrid *mg, int64_t *info, double *max_eig) {
    API_BEGIN
    REQUIRE(mg && info, "multigrid_info: null argument");
    multigrid_info(*mg->mg, info);
    if (max_eig)
        for (int l = 0; l < (int)info[0]; ++l) max_eig[l] = multigrid_max_eig(*mg->mg, l);
    API_END
}

int h1pa_pcg_solve_multigrid(h1pa_multigrid *mg, const double *b, double *x,
                             const h1pa_pcg_opts *opts, h1pa_pcg_stats *stats, void *stream) {
    API_BEGIN
    REQUIRE(mg && b && x && opts && stats, "pcg_solve_multigrid: null argument");
    REQUIRE(opts->max_iter >= 1, "pcg_solve_multigrid: max_iter must be >= 1");
    *stats = h1pa_pcg_stats{};
    pcg_multigrid(*mg->mg, b, x, *opts, *stats, as_stream(stream));
    if (!stats->converged) {
        set_error("PCG: No convergence!");
        return H1PA_ENOTCONV;
    }
    API_END
}

}  // extern "C"
