// capi_mesh_arrays.cpp -- a C++ caller of include/h1pa.h that hands the
// library its own mesh, as an MFEM caller holding a mfem::Mesh would:
// vertex, element and boundary arrays (Mesh(vertices, ..., boundary, ...),
// mesh/mesh.cpp:4891-4930) and, optionally, high-order nodes as the
// lexicographic E-vector of the nodal space (GeometricFactors::Compute,
// mesh/mesh.cpp:15445-15497).  It builds the H1 space and a PA diffusion
// form (sin/cos coefficient) on that mesh and writes the restriction maps,
// boundary DOFs, y = A x, the Jacobi diagonal and the BooleanMask E-vector.
//
//   capi_mesh_arrays <dir>
//
// dir/in.bin : int64 {nv, ne, nbdr, order, node_order (0: straight), nx}
//              double xyz[nv][3]; int32 ev[ne][8]; int32 bdr[nbdr][4];
//              double nodes[(node_order+1)^3][3][ne] (node_order > 0);
//              double x[nx] (nx must equal the space's DOF count)
// dir/out.bin: int64 {ndofs, dof, n_ess}
//              int32 gather[ne dof], offsets[ndofs+1], indices[ne dof], ess[n_ess]
//              double y[ndofs], diag[ndofs], mask[ne dof]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "h1pa.h"

namespace {

void check(int rc, const char *what) {
    if (rc != H1PA_OK) {
        std::fprintf(stderr, "capi_mesh_arrays: %s failed (%d): %s\n", what, rc, h1pa_last_error());
        std::exit(1);
    }
}

void hcheck(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "capi_mesh_arrays: %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

template <typename T>
void read(FILE *f, std::vector<T> &v, size_t n) {
    v.resize(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) {
        std::fprintf(stderr, "capi_mesh_arrays: short input\n");
        std::exit(1);
    }
}

template <typename T>
void write(FILE *f, const T *p, size_t n) {
    if (n && std::fwrite(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "capi_mesh_arrays: short write\n");
        std::exit(1);
    }
}

template <typename T>
std::vector<T> to_host(const T *d, size_t n) {
    std::vector<T> h(n);
    if (n) hcheck(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "copy to host");
    return h;
}

template <typename T>
T *to_device(const std::vector<T> &h) {
    T *d = nullptr;
    hcheck(hipMalloc(&d, (h.empty() ? 1 : h.size()) * sizeof(T)), "hipMalloc");
    if (!h.empty())
        hcheck(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "copy to device");
    return d;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: capi_mesh_arrays <dir>\n");
        return 2;
    }
    const std::string dir = argv[1];
    FILE *in = std::fopen((dir + "/in.bin").c_str(), "rb");
    if (!in) {
        std::fprintf(stderr, "capi_mesh_arrays: cannot open %s/in.bin\n", dir.c_str());
        return 1;
    }
    std::vector<int64_t> hdr;
    read(in, hdr, 6);
    const int64_t nv = hdr[0], ne = hdr[1], nbdr = hdr[2], nx = hdr[5];
    const int order = (int)hdr[3], node_order = (int)hdr[4];
    std::vector<double> xyz, nodes, x;
    std::vector<int> ev, bdr;
    read(in, xyz, 3 * nv);
    read(in, ev, 8 * ne);
    read(in, bdr, 4 * nbdr);
    if (node_order > 0) {
        const size_t dg = node_order + 1;
        read(in, nodes, dg * dg * dg * 3 * ne);
    }
    read(in, x, nx);
    std::fclose(in);

    // the caller's mesh, handed over as arrays
    h1pa_mesh *mesh = nullptr;
    check(h1pa_mesh_create(nv, xyz.data(), ne, ev.data(), nbdr, bdr.data(), &mesh), "mesh_create");
    double *d_nodes = nullptr;
    if (node_order > 0) {
        d_nodes = to_device(nodes);
        check(h1pa_mesh_set_nodes(mesh, node_order, d_nodes, nullptr), "mesh_set_nodes");
    }
    h1pa_space *space = nullptr;
    check(h1pa_space_create(mesh, order, &space), "space_create");
    int64_t info[4];
    check(h1pa_space_info(space, info), "space_info");
    const int64_t dof = info[2], ndofs = info[3];
    if (nx != ndofs) {
        std::fprintf(stderr, "capi_mesh_arrays: x has %lld entries, the space %lld DOFs\n",
                     (long long)nx, (long long)ndofs);
        return 1;
    }
    h1pa_restriction *R = nullptr;
    check(h1pa_space_restriction(space, &R), "space_restriction");
    const int *d_g = nullptr, *d_o = nullptr, *d_i = nullptr;
    check(h1pa_restriction_maps(R, &d_g, &d_o, &d_i), "restriction_maps");
    const int *d_ess = nullptr;
    int64_t n_ess = 0;
    check(h1pa_space_boundary_dofs(space, &d_ess, &n_ess), "boundary_dofs");

    h1pa_form_desc desc{};
    desc.integrators = H1PA_INTEG_DIFFUSION;
    desc.coeff_kind = H1PA_COEFF_SINCOS;
    h1pa_form *form = nullptr;
    check(h1pa_form_create(space, &desc, &form), "form_create");
    check(h1pa_form_assemble(form, nullptr), "form_assemble");
    double *d_x = to_device(x);
    double *d_y = nullptr, *d_diag = nullptr, *d_mask = nullptr;
    hcheck(hipMalloc(&d_y, ndofs * sizeof(double)), "hipMalloc");
    hcheck(hipMalloc(&d_diag, ndofs * sizeof(double)), "hipMalloc");
    hcheck(hipMalloc(&d_mask, ne * dof * sizeof(double)), "hipMalloc");
    check(h1pa_form_mult(form, d_x, d_y, nullptr), "form_mult");
    check(h1pa_form_assemble_diagonal(form, d_diag, nullptr), "assemble_diagonal");
    check(h1pa_restriction_boolean_mask(R, d_mask, nullptr), "boolean_mask");
    check(h1pa_stream_synchronize(nullptr), "synchronize");

    FILE *out = std::fopen((dir + "/out.bin").c_str(), "wb");
    if (!out) {
        std::fprintf(stderr, "capi_mesh_arrays: cannot write %s/out.bin\n", dir.c_str());
        return 1;
    }
    const int64_t ohdr[3] = {ndofs, dof, n_ess};
    write(out, ohdr, 3);
    write(out, to_host(d_g, ne * dof).data(), ne * dof);
    write(out, to_host(d_o, ndofs + 1).data(), ndofs + 1);
    write(out, to_host(d_i, ne * dof).data(), ne * dof);
    write(out, to_host(d_ess, n_ess).data(), n_ess);
    write(out, to_host(d_y, ndofs).data(), ndofs);
    write(out, to_host(d_diag, ndofs).data(), ndofs);
    write(out, to_host(d_mask, ne * dof).data(), ne * dof);
    std::fclose(out);

    hcheck(hipFree(d_x), "hipFree");
    hcheck(hipFree(d_y), "hipFree");
    hcheck(hipFree(d_diag), "hipFree");
    hcheck(hipFree(d_mask), "hipFree");
    if (d_nodes) hcheck(hipFree(d_nodes), "hipFree");
    check(h1pa_form_destroy(form), "form_destroy");
    check(h1pa_space_destroy(space), "space_destroy");
    check(h1pa_mesh_destroy(mesh), "mesh_destroy");
    std::printf("capi_mesh_arrays: ne=%lld ndofs=%lld ok\n", (long long)ne, (long long)ndofs);
    return 0;
}
