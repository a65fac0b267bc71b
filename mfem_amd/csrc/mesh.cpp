// mesh.cpp -- hexahedral meshes with MFEM's element, edge and face numbering:
// Cartesian meshes built in O(NE) with direct-indexed structured arrays, and
// meshes read from MFEM's v1.0 mesh files with hash tables.
//
// MFEM numbers edges and faces by first encounter while it walks the elements
// in order and each element's local edges/faces in Geometry::Constants<CUBE>
// order (DSTable/STable3D pushes, mesh/mesh.cpp:8526-8549, :8996-9080).  On a
// Cartesian grid every edge and face has a structured address, so the same
// numbering falls out of one pass with an "already numbered?" array instead of
// hash tables.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <sstream>
#include <string>
#include <unordered_map>

#include "internal.hpp"

namespace h1pa {

namespace {

// Generalized Hilbert curve of an a x b x c box (mesh/ncmesh.cpp:5435-5597):
// the sub-box splits and their order fix MFEM's element numbering.
struct Gilbert {
    std::vector<int> &out;
    static int sg(int v) { return (v > 0) - (v < 0); }
    void emit(int x, int y, int z) {
        out.push_back(x);
        out.push_back(y);
        out.push_back(z);
    }
    void run(int x, int y, int z, int ax, int ay, int az, int bx, int by, int bz, int cx, int cy,
             int cz) {
        const int w = std::abs(ax + ay + az), h = std::abs(bx + by + bz), d = std::abs(cx + cy + cz);
        const int dax = sg(ax), day = sg(ay), daz = sg(az);
        const int dbx = sg(bx), dby = sg(by), dbz = sg(bz);
        const int dcx = sg(cx), dcy = sg(cy), dcz = sg(cz);
        if (h == 1 && d == 1) {
            for (int i = 0; i < w; ++i) emit(x + i * dax, y + i * day, z + i * daz);
            return;
        }
        if (w == 1 && d == 1) {
            for (int i = 0; i < h; ++i) emit(x + i * dbx, y + i * dby, z + i * dbz);
            return;
        }
        if (w == 1 && h == 1) {
            for (int i = 0; i < d; ++i) emit(x + i * dcx, y + i * dcy, z + i * dcz);
            return;
        }
        int ax2 = ax / 2, ay2 = ay / 2, az2 = az / 2;
        int bx2 = bx / 2, by2 = by / 2, bz2 = bz / 2;
        int cx2 = cx / 2, cy2 = cy / 2, cz2 = cz / 2;
        // keep the halves even where the side allows it
        if ((std::abs(ax2 + ay2 + az2) & 1) && w > 2) { ax2 += dax; ay2 += day; az2 += daz; }
        if ((std::abs(bx2 + by2 + bz2) & 1) && h > 2) { bx2 += dbx; by2 += dby; bz2 += dbz; }
        if ((std::abs(cx2 + cy2 + cz2) & 1) && d > 2) { cx2 += dcx; cy2 += dcy; cz2 += dcz; }
        if (2 * w > 3 * h && 2 * w > 3 * d) {  // long along a: two halves
            run(x, y, z, ax2, ay2, az2, bx, by, bz, cx, cy, cz);
            run(x + ax2, y + ay2, z + az2, ax - ax2, ay - ay2, az - az2, bx, by, bz, cx, cy, cz);
        } else if (3 * h > 4 * d) {  // keep c whole
            run(x, y, z, bx2, by2, bz2, cx, cy, cz, ax2, ay2, az2);
            run(x + bx2, y + by2, z + bz2, ax, ay, az, bx - bx2, by - by2, bz - bz2, cx, cy, cz);
            run(x + (ax - dax) + (bx2 - dbx), y + (ay - day) + (by2 - dby),
                z + (az - daz) + (bz2 - dbz), -bx2, -by2, -bz2, cx, cy, cz, -(ax - ax2),
                -(ay - ay2), -(az - az2));
        } else if (3 * d > 4 * h) {  // keep b whole
            run(x, y, z, cx2, cy2, cz2, ax2, ay2, az2, bx, by, bz);
            run(x + cx2, y + cy2, z + cz2, ax, ay, az, bx, by, bz, cx - cx2, cy - cy2, cz - cz2);
            run(x + (ax - dax) + (cx2 - dcx), y + (ay - day) + (cy2 - dcy),
                z + (az - daz) + (cz2 - dcz), -cx2, -cy2, -cz2, -(ax - ax2), -(ay - ay2),
                -(az - az2), bx, by, bz);
        } else {  // split all three
            run(x, y, z, bx2, by2, bz2, cx2, cy2, cz2, ax2, ay2, az2);
            run(x + bx2, y + by2, z + bz2, cx, cy, cz, ax2, ay2, az2, bx - bx2, by - by2,
                bz - bz2);
            run(x + (bx2 - dbx) + (cx - dcx), y + (by2 - dby) + (cy - dcy),
                z + (bz2 - dbz) + (cz - dcz), ax, ay, az, -bx2, -by2, -bz2, -(cx - cx2),
                -(cy - cy2), -(cz - cz2));
            run(x + (ax - dax) + bx2 + (cx - dcx), y + (ay - day) + by2 + (cy - dcy),
                z + (az - daz) + bz2 + (cz - dcz), -cx, -cy, -cz, -(ax - ax2), -(ay - ay2),
                -(az - az2), bx - bx2, by - by2, bz - bz2);
            run(x + (ax - dax) + (bx2 - dbx), y + (ay - day) + (by2 - dby),
                z + (az - daz) + (bz2 - dbz), -bx2, -by2, -bz2, cx2, cy2, cz2, -(ax - ax2),
                -(ay - ay2), -(az - az2));
        }
    }
};

// Hex local edges: axis and the corner (dx,dy,dz) the edge starts from
// (Geometry::Constants<CUBE>::Edges, fem/geom.cpp:1020-1025: every edge runs
// from its lower to its higher corner along one axis).
const int kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
const int kEdgeStart[12][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {1, 0, 1},
                               {0, 1, 1}, {0, 0, 1}, {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
const int kEdgeVerts[12][2] = {{0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
                               {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// Hex local faces: normal axis, offset along it, and the vertex cycle
// (FaceVert, fem/geom.cpp:1030-1035).
const int kFaceAxis[6] = {2, 1, 0, 1, 0, 2};
const int kFaceOff[6] = {0, 0, 1, 1, 0, 1};
const int kFaceVerts[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5},
                              {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};

// Mesh::GetQuadOrientation (mesh/mesh.cpp:7586-7630).
int quad_orientation(const int *base, const int *test) {
    int i = 0;
    while (i < 4 && test[i] != base[0]) ++i;
    return (test[(i + 1) % 4] == base[1]) ? 2 * i : 2 * i + 1;
}

}  // namespace

void Mesh::build_topology(bool sfc) {
    ne = (int64_t)nx * ny * nz;
    nv = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    ecoord.resize(3 * ne);
    if (sfc) {
        std::vector<int> c;
        c.reserve(3 * ne);
        Gilbert g{c};
        if (nx >= ny && nx >= nz) g.run(0, 0, 0, nx, 0, 0, 0, ny, 0, 0, 0, nz);
        else if (ny >= nx && ny >= nz) g.run(0, 0, 0, 0, ny, 0, nx, 0, 0, 0, 0, nz);
        else g.run(0, 0, 0, 0, 0, nz, nx, 0, 0, 0, ny, 0);
        H1PA_CHECK((int64_t)c.size() == 3 * ne, H1PA_EINVAL, "space-filling curve size mismatch");
        std::copy(c.begin(), c.end(), ecoord.begin());
    } else {
        int64_t k = 0;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x, ++k) {
                    ecoord[3 * k] = x; ecoord[3 * k + 1] = y; ecoord[3 * k + 2] = z;
                }
    }
    const int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
    auto vtx = [&](int64_t x, int64_t y, int64_t z) { return (int)(x + sx * y + sxy * z); };
    ev.resize(8 * ne);
    ee.resize(12 * ne);
    eeo.resize(12 * ne);
    ef.resize(6 * ne);
    efo.resize(6 * ne);
    static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int64_t e = 0; e < ne; ++e) {
        const int x = ecoord[3 * e], y = ecoord[3 * e + 1], z = ecoord[3 * e + 2];
        for (int v = 0; v < 8; ++v)
            ev[8 * e + v] = vtx(x + corner[v][0], y + corner[v][1], z + corner[v][2]);
    }
    // Structured edge address: axis a, start vertex (x,y,z); sized per axis.
    const int64_t n_ax[3] = {(int64_t)nx * (ny + 1) * (nz + 1), (int64_t)(nx + 1) * ny * (nz + 1),
                             (int64_t)(nx + 1) * (ny + 1) * nz};
    const int64_t ext[3][3] = {{nx, ny + 1, nz + 1}, {nx + 1, ny, nz + 1}, {nx + 1, ny + 1, nz}};
    std::vector<int> eid[3];
    for (int a = 0; a < 3; ++a) eid[a].assign(n_ax[a], -1);
    int next = 0;
    for (int64_t e = 0; e < ne; ++e) {
        const int x = ecoord[3 * e], y = ecoord[3 * e + 1], z = ecoord[3 * e + 2];
        for (int j = 0; j < 12; ++j) {
            const int a = kEdgeAxis[j];
            const int64_t ix = x + kEdgeStart[j][0], iy = y + kEdgeStart[j][1], iz = z + kEdgeStart[j][2];
            const int64_t s = ix + ext[a][0] * (iy + ext[a][1] * iz);
            int &id = eid[a][s];
            if (id < 0) id = next++;
            ee[12 * e + j] = id;
            const int va = ev[8 * e + kEdgeVerts[j][0]], vb = ev[8 * e + kEdgeVerts[j][1]];
            eeo[12 * e + j] = va < vb ? 1 : -1;
        }
    }
    nedges = next;
    // Structured face address: normal axis a, lower corner (x,y,z).
    const int64_t f_ax[3] = {(int64_t)(nx + 1) * ny * nz, (int64_t)nx * (ny + 1) * nz,
                             (int64_t)nx * ny * (nz + 1)};
    const int64_t fext[3][3] = {{nx + 1, ny, nz}, {nx, ny + 1, nz}, {nx, ny, nz + 1}};
    std::vector<int> fid[3];
    std::vector<int64_t> first;  // (element * 6 + local face) that numbered the face
    for (int a = 0; a < 3; ++a) fid[a].assign(f_ax[a], -1);
    first.reserve(3 * ne + 3 * (int64_t)nx * ny);
    next = 0;
    for (int64_t e = 0; e < ne; ++e) {
        const int x = ecoord[3 * e], y = ecoord[3 * e + 1], z = ecoord[3 * e + 2];
        for (int j = 0; j < 6; ++j) {
            const int a = kFaceAxis[j];
            int64_t c[3] = {x, y, z};
            c[a] += kFaceOff[j];
            const int64_t s = c[0] + fext[a][0] * (c[1] + fext[a][1] * c[2]);
            int &id = fid[a][s];
            if (id < 0) {
                id = next++;
                first.push_back(6 * e + j);
                efo[6 * e + j] = 0;
            } else {
                const int64_t e1 = first[id] / 6;
                const int j1 = (int)(first[id] % 6);
                int base[4], test[4];
                for (int k = 0; k < 4; ++k) {
                    base[k] = ev[8 * e1 + kFaceVerts[j1][k]];
                    test[k] = ev[8 * e + kFaceVerts[j][k]];
                }
                efo[6 * e + j] = quad_orientation(base, test);
            }
            ef[6 * e + j] = id;
        }
    }
    nfaces = next;
    bface.assign(ne, 0);
    const int n_axis[3] = {nx, ny, nz};
    for (int64_t e = 0; e < ne; ++e)
        for (int j = 0; j < 6; ++j) {
            const int a = kFaceAxis[j], c = ecoord[3 * e + a];
            if ((kFaceOff[j] == 0 && c == 0) || (kFaceOff[j] == 1 && c == n_axis[a] - 1))
                bface[e] |= (uint8_t)(1u << j);
        }
    xyz.resize(3 * nv);
    int64_t v = 0;
    for (int z = 0; z <= nz; ++z)
        for (int y = 0; y <= ny; ++y)
            for (int x = 0; x <= nx; ++x, ++v) {
                xyz[3 * v] = (double)x / nx;
                xyz[3 * v + 1] = (double)y / ny;
                xyz[3 * v + 2] = (double)z / nz;
            }
}

// ---- MFEM mesh files ----------------------------------------------------------
namespace {

// Whitespace tokens of an MFEM mesh text; '#' starts a comment to the end of
// the line (skip_comment_lines in mesh/mesh_readers.cpp).
struct Tokens {
    std::istringstream in;
    explicit Tokens(const std::string &t) : in(t) {}
    std::string next() {
        std::string w;
        while (in >> w) {
            if (w[0] != '#') return w;
            std::string rest;
            std::getline(in, rest);
        }
        throw Error{H1PA_EINVAL, "mesh file: unexpected end of input"};
    }
    long long integer(const char *what) {
        const std::string w = next();
        char *end = nullptr;
        const long long v = std::strtoll(w.c_str(), &end, 10);
        H1PA_CHECK(end && *end == '\0', H1PA_EINVAL, std::string("mesh file: bad ") + what);
        return v;
    }
    double real() {
        const std::string w = next();
        char *end = nullptr;
        const double v = std::strtod(w.c_str(), &end);
        H1PA_CHECK(end && *end == '\0', H1PA_EINVAL, "mesh file: bad vertex coordinate");
        return v;
    }
    void expect(const char *word) {
        H1PA_CHECK(next() == word, H1PA_EINVAL, std::string("mesh file: expected '") + word + "'");
    }
};

struct QuadKey {
    std::array<int, 4> v;
    bool operator==(const QuadKey &o) const { return v == o.v; }
};
struct QuadHash {
    size_t operator()(const QuadKey &k) const {
        uint64_t h = 1469598103934665603ULL;
        for (int x : k.v) h = (h ^ (uint64_t)(uint32_t)x) * 1099511628211ULL;
        return (size_t)h;
    }
};
QuadKey quad_key(int a, int b, int c, int d) {
    QuadKey k{{a, b, c, d}};
    std::sort(k.v.begin(), k.v.end());
    return k;
}

}  // namespace

// Mesh::ReadMFEMMesh (mesh/mesh_readers.cpp:40-130), version 1.0, for
// conforming meshes of hexahedra (geometry 5) with a boundary of quadrilaterals
// (geometry 3) and straight vertex coordinates; a high-order "nodes" section,
// other element types and other dimensions are rejected.  The elements keep
// the file order.  Topology as FinalizeTopology makes it: faces and edges
// numbered by first encounter over the elements in order and their local
// faces/edges in Geometry::Constants<CUBE> order (STable3D / DSTable pushes,
// mesh/mesh.cpp:8526-8549, :8996-9080), face orientation from the first
// element's vertex cycle (GetQuadOrientation), edge orientation from the
// vertex order.  The essential boundary is every face of the file's boundary
// list (all attributes marked, as GetBoundaryTrueDofs does).
void Mesh::load_mfem_v10(const char *text, size_t len) {
    const std::string t(text, len);
    {
        std::istringstream first(t);
        std::string line;
        std::getline(first, line);
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
        H1PA_CHECK(line == "MFEM mesh v1.0", H1PA_EUNSUP,
                   "mesh file: only 'MFEM mesh v1.0' is supported");
    }
    Tokens tk(t.substr(t.find('\n') + 1));
    tk.expect("dimension");
    H1PA_CHECK(tk.integer("dimension") == 3, H1PA_EUNSUP, "mesh file: only 3D meshes are supported");
    tk.expect("elements");
    const long long n_el = tk.integer("element count");
    H1PA_CHECK(n_el >= 1 && n_el < (1LL << 28), H1PA_EINVAL, "mesh file: bad element count");
    // indices stay 64-bit until the vertex count is known and they are range
    // checked: a narrowing cast first could wrap 2^32 + v into [0, n_v)
    std::vector<long long> ev64(8 * (size_t)n_el);
    for (long long e = 0; e < n_el; ++e) {
        (void)tk.integer("attribute");
        H1PA_CHECK(tk.integer("geometry") == 5, H1PA_EUNSUP,
                   "mesh file: only hexahedral elements (geometry 5) are supported");
        for (int k = 0; k < 8; ++k) ev64[8 * e + k] = tk.integer("vertex index");
    }
    tk.expect("boundary");
    const long long n_bdr = tk.integer("boundary count");
    H1PA_CHECK(n_bdr >= 0 && n_bdr <= 6 * n_el, H1PA_EINVAL, "mesh file: bad boundary count");
    std::vector<long long> bq64(4 * (size_t)n_bdr);
    for (long long b = 0; b < n_bdr; ++b) {
        (void)tk.integer("attribute");
        H1PA_CHECK(tk.integer("geometry") == 3, H1PA_EUNSUP,
                   "mesh file: only quadrilateral boundary elements (geometry 3) are supported");
        for (int k = 0; k < 4; ++k) bq64[4 * b + k] = tk.integer("vertex index");
    }
    tk.expect("vertices");
    const long long n_v = tk.integer("vertex count");
    H1PA_CHECK(n_v >= 8 && n_v < (1LL << 31), H1PA_EINVAL, "mesh file: bad vertex count");
    const std::string sd = tk.next();
    H1PA_CHECK(sd != "nodes", H1PA_EUNSUP,
               "mesh file: curved meshes (a 'nodes' section) are not supported");
    H1PA_CHECK(sd == "3", H1PA_EUNSUP, "mesh file: vertices must have 3 coordinates");
    std::vector<double> pts(3 * (size_t)n_v);
    for (auto &c : pts) c = tk.real();
    for (long long v : ev64)
        H1PA_CHECK(v >= 0 && v < n_v, H1PA_EINVAL, "mesh file: vertex index out of range");
    for (long long v : bq64)
        H1PA_CHECK(v >= 0 && v < n_v, H1PA_EINVAL, "mesh file: vertex index out of range");
    const std::vector<int> evi(ev64.begin(), ev64.end()), bq(bq64.begin(), bq64.end());
    from_arrays(n_v, pts.data(), n_el, evi.data(), n_bdr, bq.data());
}

// Mesh(vertices, nv, element_indices, HEXAHEDRON, ..., boundary_indices,
// SQUARE, ...) + FinalizeTopology (mesh/mesh.cpp:4891-4930): elements keep
// the given order; faces and edges are numbered by first encounter
// (GetElementToFaceTable mesh/mesh.cpp:8996, GetElementToEdgeTable :8551);
// the boundary is the given quadrilateral list.  Indices are range checked
// by the callers.
void Mesh::from_arrays(int64_t n_v, const double *pts, int64_t n_el, const int *evi,
                       int64_t n_bdr, const int *bq) {
    xyz.assign(pts, pts + 3 * n_v);
    ev.assign(evi, evi + 8 * n_el);
    nx = ny = nz = 0;
    nv = n_v;
    ne = n_el;
    from_file = true;
    node_order = 1;
    curve_amp = 0.0;
    nodes_external = false;
    d_topo.release();
    d_ori.release();
    d_elem_at.release();
    ecoord.assign(3 * ne, 0);
    // an inverted element would flip the sign of every derivative: refuse it
    // (trilinear Jacobian at the element centre)
    for (int64_t e = 0; e < ne; ++e) {
        double J[3][3] = {};
        static const int sgn[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                      {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        for (int k = 0; k < 8; ++k)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) J[r][c] += 0.25 * sgn[k][c] * xyz[3 * ev[8 * e + k] + r];
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        H1PA_CHECK(det > 0.0, H1PA_EUNSUP, "mesh: inverted or degenerate element " +
                                               std::to_string(e));
    }
    // faces
    ef.resize(6 * ne);
    efo.resize(6 * ne);
    std::unordered_map<QuadKey, int, QuadHash> fmap;
    fmap.reserve(4 * ne);
    std::vector<std::array<int, 4>> fverts;  // first element's vertex cycle of each face
    std::vector<int> fcount;
    for (int64_t e = 0; e < ne; ++e)
        for (int j = 0; j < 6; ++j) {
            const int *cyc = kFaceVerts[j];
            std::array<int, 4> vv;
            for (int k = 0; k < 4; ++k) vv[k] = ev[8 * e + cyc[k]];
            auto it = fmap.emplace(quad_key(vv[0], vv[1], vv[2], vv[3]), (int)fverts.size());
            const int id = it.first->second;
            if (it.second) {
                fverts.push_back(vv);
                fcount.push_back(1);
                efo[6 * e + j] = 0;
            } else {
                H1PA_CHECK(++fcount[id] <= 2, H1PA_EINVAL, "mesh: face shared by 3 elements");
                efo[6 * e + j] = quad_orientation(fverts[id].data(), vv.data());
            }
            ef[6 * e + j] = id;
        }
    nfaces = (int64_t)fverts.size();
    // edges
    ee.resize(12 * ne);
    eeo.resize(12 * ne);
    std::unordered_map<uint64_t, int> emap;
    emap.reserve(4 * ne);
    for (int64_t e = 0; e < ne; ++e)
        for (int j = 0; j < 12; ++j) {
            const int va = ev[8 * e + kEdgeVerts[j][0]], vb = ev[8 * e + kEdgeVerts[j][1]];
            const uint64_t key = (uint64_t)std::min(va, vb) * (uint64_t)nv + (uint64_t)std::max(va, vb);
            auto it = emap.emplace(key, (int)emap.size());
            ee[12 * e + j] = it.first->second;
            eeo[12 * e + j] = va < vb ? 1 : -1;
        }
    nedges = (int64_t)emap.size();
    // boundary faces: the given list
    std::vector<uint8_t> on(nfaces, 0);
    for (long long b = 0; b < n_bdr; ++b) {
        auto it = fmap.find(quad_key(bq[4 * b], bq[4 * b + 1], bq[4 * b + 2], bq[4 * b + 3]));
        H1PA_CHECK(it != fmap.end(), H1PA_EINVAL,
                   "mesh: boundary element " + std::to_string(b) + " is not an element face");
        on[it->second] = 1;
    }
    bface.assign(ne, 0);
    for (int64_t e = 0; e < ne; ++e)
        for (int j = 0; j < 6; ++j)
            if (on[ef[6 * e + j]]) bface[e] |= (uint8_t)(1u << j);
    nodes_valid = false;
}

void Mesh::ensure_device(hipStream_t s) {
    if (d_topo.p) return;
    // [ne][29]: 8 vertices, 12 edges, 6 faces, 3 grid coordinates
    std::vector<int> topo(29 * ne);
    std::vector<int8_t> ori(18 * ne);
    for (int64_t e = 0; e < ne; ++e) {
        int *t = topo.data() + 29 * e;
        for (int k = 0; k < 8; ++k) t[k] = ev[8 * e + k];
        for (int k = 0; k < 12; ++k) {
            t[8 + k] = ee[12 * e + k];
            ori[18 * e + k] = (int8_t)eeo[12 * e + k];
        }
        for (int k = 0; k < 6; ++k) {
            t[20 + k] = ef[6 * e + k];
            ori[18 * e + 12 + k] = (int8_t)efo[6 * e + k];
        }
        t[26] = bface[e];
        t[27] = t[28] = 0;
    }
    d_topo.upload(topo.data(), topo.size(), s);
    d_ori.upload(ori.data(), ori.size(), s);
    H1PA_HIP(hipStreamSynchronize(s));
}

void Mesh::ensure_nodes(hipStream_t s) {
    if (nodes_valid) return;
    H1PA_CHECK(!nodes_external, H1PA_EINVAL, "mesh: the nodes given by set_nodes are gone");
    const int64_t Dg = node_order + 1;
    d_nodes.alloc(Dg * Dg * Dg * 3 * ne);
    ensure_device(s);
    if (node_order == 1 && curve_amp == 0.0) kern::nodes_trilinear(*this, d_nodes.p, s);
    else kern::nodes_curved(*this, node_order, curve_amp, d_nodes.p, s);
    nodes_valid = true;
}

namespace kern {

namespace {
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
}  // namespace

void perturb_vertices(Mesh &m, double amp, uint64_t seed) {
    H1PA_CHECK(!m.from_file, H1PA_EUNSUP, "perturb: only Cartesian meshes can be perturbed");
    // Interior vertices move by U(-amp h, amp h) per coordinate; the uniform
    // draw for coordinate c of vertex v is splitmix64(seed ^ splitmix64(3v+c)).
    const double h[3] = {1.0 / m.nx, 1.0 / m.ny, 1.0 / m.nz};
    int64_t v = 0;
    for (int z = 0; z <= m.nz; ++z)
        for (int y = 0; y <= m.ny; ++y)
            for (int x = 0; x <= m.nx; ++x, ++v) {
                if (x == 0 || y == 0 || z == 0 || x == m.nx || y == m.ny || z == m.nz) continue;
                for (int c = 0; c < 3; ++c) {
                    const uint64_t r = splitmix64(seed ^ splitmix64(3ULL * (uint64_t)v + c));
                    const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
                    m.xyz[3 * v + c] += amp * h[c] * (2.0 * u - 1.0);
                }
            }
    m.nodes_valid = false;
    m.nodes_external = false;  // the nodes follow the moved vertices again
}

}  // namespace kern
}  // namespace h1pa
