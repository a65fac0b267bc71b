/*
 * h1pa_oracle.c -- CPU restatement of MFEM's H1 partial-assembly path.
 *
 * TEST INFRASTRUCTURE ONLY (see h1pa_oracle.h for the pinning strategy).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this; the product in mfem_amd/ never does.
 *
 * The loops follow MFEM's host execution of the cited kernels (MFEM's
 * MFEM_FOREACH_THREAD becomes a plain loop on the host,
 * general/backends.hpp:64-70).  Element loops are OpenMP-parallel so the
 * same code serves as the multi-core CPU baseline.
 */
#include "h1pa_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Sets the OpenMP thread count of the following calls (the CPU baseline
 * times the apply on all cores and on one). */
void orc_set_num_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

/* ======================================================================
 * Generalized Hilbert curve on an nx x ny x nz grid.
 * Follows mesh/ncmesh.cpp:5372-5375 (sgn), :5435-5597 (HilbertSfc3D) and
 * :5616-5640 (GridSfcOrdering3D).  The recursion is the published
 * "gilbert3d" generalized Hilbert curve; the split rules, the even-step
 * preference and the order of the sub-blocks must match exactly because they
 * fix MFEM's element numbering.
 * ====================================================================== */
static int sgn_i(int x) { return (x < 0) ? -1 : (x > 0) ? 1 : 0; }

typedef struct { int *c; long n; } sfc_out;

static void sfc_push(sfc_out *o, int x, int y, int z)
{
    o->c[3 * o->n + 0] = x;
    o->c[3 * o->n + 1] = y;
    o->c[3 * o->n + 2] = z;
    o->n++;
}

static void hilbert3d(int x, int y, int z, int ax, int ay, int az,
                      int bx, int by, int bz, int cx, int cy, int cz,
                      sfc_out *o)
{
    const int w = abs(ax + ay + az), h = abs(bx + by + bz), d = abs(cx + cy + cz);
    const int dax = sgn_i(ax), day = sgn_i(ay), daz = sgn_i(az);
    const int dbx = sgn_i(bx), dby = sgn_i(by), dbz = sgn_i(bz);
    const int dcx = sgn_i(cx), dcy = sgn_i(cy), dcz = sgn_i(cz);

    if (h == 1 && d == 1) {
        for (int i = 0; i < w; i++, x += dax, y += day, z += daz) sfc_push(o, x, y, z);
        return;
    }
    if (w == 1 && d == 1) {
        for (int i = 0; i < h; i++, x += dbx, y += dby, z += dbz) sfc_push(o, x, y, z);
        return;
    }
    if (w == 1 && h == 1) {
        for (int i = 0; i < d; i++, x += dcx, y += dcy, z += dcz) sfc_push(o, x, y, z);
        return;
    }

    int ax2 = ax / 2, ay2 = ay / 2, az2 = az / 2;
    int bx2 = bx / 2, by2 = by / 2, bz2 = bz / 2;
    int cx2 = cx / 2, cy2 = cy / 2, cz2 = cz / 2;
    const int w2 = abs(ax2 + ay2 + az2), h2 = abs(bx2 + by2 + bz2), d2 = abs(cx2 + cy2 + cz2);

    if ((w2 & 1) && w > 2) { ax2 += dax; ay2 += day; az2 += daz; }
    if ((h2 & 1) && h > 2) { bx2 += dbx; by2 += dby; bz2 += dbz; }
    if ((d2 & 1) && d > 2) { cx2 += dcx; cy2 += dcy; cz2 += dcz; }

    if (2 * w > 3 * h && 2 * w > 3 * d) {
        hilbert3d(x, y, z, ax2, ay2, az2, bx, by, bz, cx, cy, cz, o);
        hilbert3d(x + ax2, y + ay2, z + az2, ax - ax2, ay - ay2, az - az2,
                  bx, by, bz, cx, cy, cz, o);
    } else if (3 * h > 4 * d) {
        hilbert3d(x, y, z, bx2, by2, bz2, cx, cy, cz, ax2, ay2, az2, o);
        hilbert3d(x + bx2, y + by2, z + bz2, ax, ay, az, bx - bx2, by - by2, bz - bz2,
                  cx, cy, cz, o);
        hilbert3d(x + (ax - dax) + (bx2 - dbx), y + (ay - day) + (by2 - dby),
                  z + (az - daz) + (bz2 - dbz), -bx2, -by2, -bz2, cx, cy, cz,
                  -(ax - ax2), -(ay - ay2), -(az - az2), o);
    } else if (3 * d > 4 * h) {
        hilbert3d(x, y, z, cx2, cy2, cz2, ax2, ay2, az2, bx, by, bz, o);
        hilbert3d(x + cx2, y + cy2, z + cz2, ax, ay, az, bx, by, bz,
                  cx - cx2, cy - cy2, cz - cz2, o);
        hilbert3d(x + (ax - dax) + (cx2 - dcx), y + (ay - day) + (cy2 - dcy),
                  z + (az - daz) + (cz2 - dcz), -cx2, -cy2, -cz2,
                  -(ax - ax2), -(ay - ay2), -(az - az2), bx, by, bz, o);
    } else {
        hilbert3d(x, y, z, bx2, by2, bz2, cx2, cy2, cz2, ax2, ay2, az2, o);
        hilbert3d(x + bx2, y + by2, z + bz2, cx, cy, cz, ax2, ay2, az2,
                  bx - bx2, by - by2, bz - bz2, o);
        hilbert3d(x + (bx2 - dbx) + (cx - dcx), y + (by2 - dby) + (cy - dcy),
                  z + (bz2 - dbz) + (cz - dcz), ax, ay, az, -bx2, -by2, -bz2,
                  -(cx - cx2), -(cy - cy2), -(cz - cz2), o);
        hilbert3d(x + (ax - dax) + bx2 + (cx - dcx), y + (ay - day) + by2 + (cy - dcy),
                  z + (az - daz) + bz2 + (cz - dcz), -cx, -cy, -cz,
                  -(ax - ax2), -(ay - ay2), -(az - az2), bx - bx2, by - by2, bz - bz2, o);
        hilbert3d(x + (ax - dax) + (bx2 - dbx), y + (ay - day) + (by2 - dby),
                  z + (az - daz) + (bz2 - dbz), -bx2, -by2, -bz2, cx2, cy2, cz2,
                  -(ax - ax2), -(ay - ay2), -(az - az2), o);
    }
}

void orc_grid_sfc3d(int nx, int ny, int nz, int *coords)
{
    sfc_out o = {coords, 0};
    if (nx >= ny && nx >= nz)
        hilbert3d(0, 0, 0, nx, 0, 0, 0, ny, 0, 0, 0, nz, &o);
    else if (ny >= nx && ny >= nz)
        hilbert3d(0, 0, 0, 0, ny, 0, nx, 0, 0, 0, 0, nz, &o);
    else
        hilbert3d(0, 0, 0, 0, 0, nz, nx, 0, 0, 0, ny, 0, &o);
}

/* ======================================================================
 * Mesh + topology.
 * ====================================================================== */
/* fem/geom.cpp:1020-1036: Constants<Geometry::CUBE>::Edges, FaceVert. */
static const int HEX_EDGES[12][2] = {
    {0, 1}, {1, 2}, {3, 2}, {0, 3}, {4, 5}, {5, 6},
    {7, 6}, {4, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
static const int HEX_FACES[6][4] = {
    {3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5},
    {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};

void orc_mesh_vertices(int nx, int ny, int nz, double *xyz)
{
    /* mesh/mesh.cpp:3870-3882: coord = (real_t) i / n, x fastest. */
    long v = 0;
    for (int z = 0; z <= nz; z++)
        for (int y = 0; y <= ny; y++)
            for (int x = 0; x <= nx; x++, v++) {
                xyz[3 * v + 0] = (double)x / nx;
                xyz[3 * v + 1] = (double)y / ny;
                xyz[3 * v + 2] = (double)z / nz;
            }
}

static unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static double u01(unsigned long long seed, unsigned long long ctr)
{
    /* 53 random bits -> [0,1) */
    return (double)(splitmix64(seed ^ splitmix64(ctr)) >> 11) * (1.0 / 9007199254740992.0);
}

void orc_perturb_vertices(int nx, int ny, int nz, double amp, unsigned long long seed, double *xyz)
{
    const double h[3] = {1.0 / nx, 1.0 / ny, 1.0 / nz};
    long v = 0;
    for (int z = 0; z <= nz; z++)
        for (int y = 0; y <= ny; y++)
            for (int x = 0; x <= nx; x++, v++) {
                if (x == 0 || y == 0 || z == 0 || x == nx || y == ny || z == nz) continue;
                for (int c = 0; c < 3; c++) {
                    const double u = u01(seed, 3ULL * (unsigned long long)v + c);
                    xyz[3 * v + c] += amp * h[c] * (2.0 * u - 1.0);
                }
            }
}

/* DSTable (general/table.cpp:612-645): per-row singly linked lists, new
 * entries numbered in push order. */
typedef struct { int *head; int *col, *idx, *prev; int n; } dstable;

static int ds_push(dstable *t, int a, int b)
{
    const int r = a <= b ? a : b, c = a <= b ? b : a;
    for (int n = t->head[r]; n >= 0; n = t->prev[n])
        if (t->col[n] == c) return t->idx[n];
    const int n = t->n;
    t->col[n] = c;
    t->idx[n] = n;
    t->prev[n] = t->head[r];
    t->head[r] = n;
    t->n++;
    return n;
}

/* STable3D (general/stable3d.cpp:33-95, :139-161). */
typedef struct { int *head; int *col, *flr, *num, *prev; int n; } stable3d;

static void sort3(int *r, int *c, int *f)
{
    int t;
    if (*r > *c) { t = *r; *r = *c; *c = t; }
    if (*c > *f) { t = *c; *c = *f; *f = t; }
    if (*r > *c) { t = *r; *r = *c; *c = t; }
}

static int st_push(stable3d *t, int r, int c, int f)
{
    sort3(&r, &c, &f);
    for (int n = t->head[r]; n >= 0; n = t->prev[n])
        if (t->col[n] == c && t->flr[n] == f) return t->num[n];
    const int n = t->n;
    t->col[n] = c;
    t->flr[n] = f;
    t->num[n] = n;
    t->prev[n] = t->head[r];
    t->head[r] = n;
    t->n++;
    return n;
}

static int st_push4(stable3d *t, int r, int c, int f, int q)
{
    /* drop the largest of the four vertex numbers */
    int i = 0, mx = r;
    if (mx < c) { mx = c; i = 1; }
    if (mx < f) { mx = f; i = 2; }
    if (mx < q) { i = 3; }
    switch (i) {
    case 0: return st_push(t, c, f, q);
    case 1: return st_push(t, r, f, q);
    case 2: return st_push(t, r, c, q);
    default: return st_push(t, r, c, f);
    }
}

/* Mesh::GetQuadOrientation (mesh/mesh.cpp:7586-7630). */
static int quad_orientation(const int *base, const int *test)
{
    int i;
    for (i = 0; i < 4; i++)
        if (test[i] == base[0]) break;
    if (test[(i + 1) % 4] == base[1]) return 2 * i;
    return 2 * i + 1;
}

static int orc_mesh_topology(orc_mesh *m);

int orc_mesh_cartesian(int nx, int ny, int nz, int sfc, orc_mesh *m)
{
    memset(m, 0, sizeof(*m));
    m->nx = nx; m->ny = ny; m->nz = nz;
    m->nv = (nx + 1) * (ny + 1) * (nz + 1);
    m->ne = nx * ny * nz;
    const int ne = m->ne;
    m->ev = (int *)malloc(sizeof(int) * 8 * (size_t)ne);
    m->ee = (int *)malloc(sizeof(int) * 12 * (size_t)ne);
    m->eeo = (int *)malloc(sizeof(int) * 12 * (size_t)ne);
    m->ef = (int *)malloc(sizeof(int) * 6 * (size_t)ne);
    m->efo = (int *)malloc(sizeof(int) * 6 * (size_t)ne);
    if (!m->ev || !m->ee || !m->eeo || !m->ef || !m->efo) return -1;

#define VTX(X, Y, Z) ((X) + ((Y) + (Z) * (ny + 1)) * (nx + 1))
    /* mesh/mesh.cpp:3905-3958 */
    if (sfc) {
        int *c = (int *)malloc(sizeof(int) * 3 * (size_t)ne);
        if (!c) return -1;
        orc_grid_sfc3d(nx, ny, nz, c);
        for (int k = 0; k < ne; k++) {
            const int x = c[3 * k], y = c[3 * k + 1], z = c[3 * k + 2];
            int *v = m->ev + 8 * (size_t)k;
            v[0] = VTX(x, y, z);         v[1] = VTX(x + 1, y, z);
            v[2] = VTX(x + 1, y + 1, z); v[3] = VTX(x, y + 1, z);
            v[4] = VTX(x, y, z + 1);     v[5] = VTX(x + 1, y, z + 1);
            v[6] = VTX(x + 1, y + 1, z + 1); v[7] = VTX(x, y + 1, z + 1);
        }
        free(c);
    } else {
        int k = 0;
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++, k++) {
                    int *v = m->ev + 8 * (size_t)k;
                    v[0] = VTX(x, y, z);         v[1] = VTX(x + 1, y, z);
                    v[2] = VTX(x + 1, y + 1, z); v[3] = VTX(x, y + 1, z);
                    v[4] = VTX(x, y, z + 1);     v[5] = VTX(x + 1, y, z + 1);
                    v[6] = VTX(x + 1, y + 1, z + 1); v[7] = VTX(x, y + 1, z + 1);
                }
    }
#undef VTX
    return orc_mesh_topology(m);
}

/* Faces and edges of the elements in m->ev (m->nv, m->ne set). */
static int orc_mesh_topology(orc_mesh *m)
{
    const int ne = m->ne, nv = m->nv;
    /* Faces first (FinalizeTopology, mesh/mesh.cpp:3700-3703):
     * GetElementToFaceTable (:8996-9080) then GenerateFaces (:8768+). */
    {
        const size_t cap = 6 * (size_t)ne;
        stable3d t;
        t.head = (int *)malloc(sizeof(int) * (size_t)nv);
        t.col = (int *)malloc(sizeof(int) * cap);
        t.flr = (int *)malloc(sizeof(int) * cap);
        t.num = (int *)malloc(sizeof(int) * cap);
        t.prev = (int *)malloc(sizeof(int) * cap);
        t.n = 0;
        if (!t.head || !t.col || !t.flr || !t.num || !t.prev) return -1;
        for (int i = 0; i < nv; i++) t.head[i] = -1;
        for (int i = 0; i < ne; i++) {
            const int *v = m->ev + 8 * (size_t)i;
            for (int j = 0; j < 6; j++) {
                const int *fv = HEX_FACES[j];
                m->ef[6 * (size_t)i + j] = st_push4(&t, v[fv[0]], v[fv[1]], v[fv[2]], v[fv[3]]);
            }
        }
        m->nfaces = t.n;
        free(t.head); free(t.col); free(t.flr); free(t.num); free(t.prev);

        /* AddQuadFaceElement (mesh/mesh.cpp:8741-8766): the first element to
         * meet a face stores its vertex order (orientation 0); the second gets
         * GetQuadOrientation(face vertices, its own face vertices). */
        int *fverts = (int *)malloc(sizeof(int) * 4 * (size_t)m->nfaces);
        m->f_elem1 = (int *)malloc(sizeof(int) * (size_t)m->nfaces);
        if (!fverts || !m->f_elem1) return -1;
        for (int f = 0; f < m->nfaces; f++) m->f_elem1[f] = -1;
        for (int i = 0; i < ne; i++) {
            const int *v = m->ev + 8 * (size_t)i;
            for (int j = 0; j < 6; j++) {
                const int gf = m->ef[6 * (size_t)i + j];
                const int *fv = HEX_FACES[j];
                int vv[4] = {v[fv[0]], v[fv[1]], v[fv[2]], v[fv[3]]};
                if (m->f_elem1[gf] < 0) {
                    m->f_elem1[gf] = i;
                    memcpy(fverts + 4 * (size_t)gf, vv, sizeof(vv));
                    m->efo[6 * (size_t)i + j] = 0;
                } else {
                    m->efo[6 * (size_t)i + j] = quad_orientation(fverts + 4 * (size_t)gf, vv);
                }
            }
        }
        free(fverts);
    }

    /* Edges: GetVertexToVertexTable (:8526-8549) pushes (element, local edge)
     * pairs in order; GetElementEdges (:8044-8073) gives the orientation. */
    {
        const size_t cap = 12 * (size_t)ne;
        dstable t;
        t.head = (int *)malloc(sizeof(int) * (size_t)nv);
        t.col = (int *)malloc(sizeof(int) * cap);
        t.idx = (int *)malloc(sizeof(int) * cap);
        t.prev = (int *)malloc(sizeof(int) * cap);
        t.n = 0;
        if (!t.head || !t.col || !t.idx || !t.prev) return -1;
        for (int i = 0; i < nv; i++) t.head[i] = -1;
        for (int i = 0; i < ne; i++) {
            const int *v = m->ev + 8 * (size_t)i;
            for (int j = 0; j < 12; j++) {
                const int a = v[HEX_EDGES[j][0]], b = v[HEX_EDGES[j][1]];
                m->ee[12 * (size_t)i + j] = ds_push(&t, a, b);
                m->eeo[12 * (size_t)i + j] = (a < b) ? 1 : -1;
            }
        }
        m->nedges = t.n;
        free(t.head); free(t.col); free(t.idx); free(t.prev);
    }
    return 0;
}

static int cmp_int4(const void *a, const void *b)
{
    const int *x = (const int *)a, *y = (const int *)b;
    for (int k = 0; k < 4; k++)
        if (x[k] != y[k]) return x[k] < y[k] ? -1 : 1;
    return 0;
}

static void sort4(int *v)
{
    for (int i = 1; i < 4; i++)
        for (int j = i; j > 0 && v[j - 1] > v[j]; j--) {
            const int t = v[j]; v[j] = v[j - 1]; v[j - 1] = t;
        }
}

int orc_mesh_from_elements(int nv, int ne, const int *ev, int nb, const int *bquads,
                           orc_mesh *m)
{
    memset(m, 0, sizeof(*m));
    m->nv = nv;
    m->ne = ne;
    m->ev = (int *)malloc(sizeof(int) * 8 * (size_t)ne);
    m->ee = (int *)malloc(sizeof(int) * 12 * (size_t)ne);
    m->eeo = (int *)malloc(sizeof(int) * 12 * (size_t)ne);
    m->ef = (int *)malloc(sizeof(int) * 6 * (size_t)ne);
    m->efo = (int *)malloc(sizeof(int) * 6 * (size_t)ne);
    if (!m->ev || !m->ee || !m->eeo || !m->ef || !m->efo) return -1;
    memcpy(m->ev, ev, sizeof(int) * 8 * (size_t)ne);
    if (orc_mesh_topology(m) != 0) return -1;
    /* boundary: the faces listed by the file (GetBoundaryTrueDofs marks every
     * boundary attribute); each face's sorted vertex set, looked up by bisection */
    const int nf = m->nfaces;
    int *key = (int *)malloc(sizeof(int) * 5 * (size_t)nf);
    m->f_bdr = (char *)calloc((size_t)nf, 1);
    if (!key || !m->f_bdr) return -1;
    for (int i = 0; i < ne; i++)
        for (int j = 0; j < 6; j++) {
            const int f = m->ef[6 * (size_t)i + j];
            int *k = key + 5 * (size_t)f;
            for (int c = 0; c < 4; c++) k[c] = ev[8 * (size_t)i + HEX_FACES[j][c]];
            sort4(k);
            k[4] = f;
        }
    qsort(key, (size_t)nf, sizeof(int) * 5, cmp_int4);
    int rc = 0;
    for (int b = 0; b < nb && rc == 0; b++) {
        int q[4] = {bquads[4 * b], bquads[4 * b + 1], bquads[4 * b + 2], bquads[4 * b + 3]};
        sort4(q);
        int lo = 0, hi = nf - 1, found = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) / 2, c = cmp_int4(q, key + 5 * (size_t)mid);
            if (c == 0) { found = key[5 * (size_t)mid + 4]; break; }
            if (c < 0) hi = mid - 1; else lo = mid + 1;
        }
        if (found < 0) rc = -2;
        else m->f_bdr[found] = 1;
    }
    free(key);
    return rc;
}

void orc_mesh_free(orc_mesh *m)
{
    free(m->ev); free(m->ee); free(m->eeo); free(m->ef); free(m->efo); free(m->f_elem1);
    free(m->f_bdr);
    memset(m, 0, sizeof(*m));
}

/* ======================================================================
 * H1 space numbering.
 * ====================================================================== */
void orc_h1_dof_map(int p, int *dof_map)
{
    /* fem/fe/fe_base.cpp:2557-2674 (dims == 3): dof_map[lex] = native. */
    const int p1 = p + 1;
#define L(i, j, k) ((i) + ((j) + (k) * p1) * p1)
    dof_map[L(0, 0, 0)] = 0; dof_map[L(p, 0, 0)] = 1;
    dof_map[L(p, p, 0)] = 2; dof_map[L(0, p, 0)] = 3;
    dof_map[L(0, 0, p)] = 4; dof_map[L(p, 0, p)] = 5;
    dof_map[L(p, p, p)] = 6; dof_map[L(0, p, p)] = 7;
    int o = 8;
    for (int i = 1; i < p; i++) dof_map[L(i, 0, 0)] = o++;   /* (0,1) */
    for (int i = 1; i < p; i++) dof_map[L(p, i, 0)] = o++;   /* (1,2) */
    for (int i = 1; i < p; i++) dof_map[L(i, p, 0)] = o++;   /* (3,2) */
    for (int i = 1; i < p; i++) dof_map[L(0, i, 0)] = o++;   /* (0,3) */
    for (int i = 1; i < p; i++) dof_map[L(i, 0, p)] = o++;   /* (4,5) */
    for (int i = 1; i < p; i++) dof_map[L(p, i, p)] = o++;   /* (5,6) */
    for (int i = 1; i < p; i++) dof_map[L(i, p, p)] = o++;   /* (7,6) */
    for (int i = 1; i < p; i++) dof_map[L(0, i, p)] = o++;   /* (4,7) */
    for (int i = 1; i < p; i++) dof_map[L(0, 0, i)] = o++;   /* (0,4) */
    for (int i = 1; i < p; i++) dof_map[L(p, 0, i)] = o++;   /* (1,5) */
    for (int i = 1; i < p; i++) dof_map[L(p, p, i)] = o++;   /* (2,6) */
    for (int i = 1; i < p; i++) dof_map[L(0, p, i)] = o++;   /* (3,7) */
    for (int j = 1; j < p; j++) for (int i = 1; i < p; i++) dof_map[L(i, p - j, 0)] = o++;
    for (int j = 1; j < p; j++) for (int i = 1; i < p; i++) dof_map[L(i, 0, j)] = o++;
    for (int j = 1; j < p; j++) for (int i = 1; i < p; i++) dof_map[L(p, i, j)] = o++;
    for (int j = 1; j < p; j++) for (int i = 1; i < p; i++) dof_map[L(p - i, p, j)] = o++;
    for (int j = 1; j < p; j++) for (int i = 1; i < p; i++) dof_map[L(0, p - i, j)] = o++;
    for (int j = 1; j < p; j++) for (int i = 1; i < p; i++) dof_map[L(i, j, p)] = o++;
    for (int k = 1; k < p; k++)
        for (int j = 1; j < p; j++)
            for (int i = 1; i < p; i++) dof_map[L(i, j, k)] = o++;
#undef L
}

/* H1_FECollection::QuadDofOrd (fem/fe_coll.cpp:1920-1936). */
static int quad_dof_ord(int ori, int o, int pm1)
{
    const int pm2 = pm1 - 1;
    const int i = o % pm1, j = o / pm1;
    switch (ori) {
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

int orc_h1_elem_dofs(const orc_mesh *m, int p, int *elem_dof)
{
    /* FiniteElementSpace::Construct (fem/fespace.cpp:2818-2895) block sizes
     * and GetElementDofs (:3428-3535) per-element order. */
    const int pm1 = p - 1;
    const int nd = (p + 1) * (p + 1) * (p + 1);
    const long nvdofs = m->nv;
    const long nedofs = (long)m->nedges * pm1;
    const long nfdofs = (long)m->nfaces * pm1 * pm1;
    const long nb = (long)pm1 * pm1 * pm1;
    for (long e = 0; e < m->ne; e++) {
        int *d = elem_dof + e * nd;
        int o = 0;
        for (int i = 0; i < 8; i++) d[o++] = m->ev[8 * e + i];
        for (int i = 0; i < 12; i++) {
            const long ebase = (long)m->ee[12 * e + i] * pm1;
            const int orient = m->eeo[12 * e + i];
            for (int j = 0; j < pm1; j++) {
                /* SegDofOrd (fem/fe_coll.cpp:1826-1831) */
                const int ind = orient > 0 ? j : (pm1 - 1 - j);
                d[o++] = (int)(nvdofs + ebase + ind);
            }
        }
        for (int i = 0; i < 6; i++) {
            const long fbase = (long)m->ef[6 * e + i] * pm1 * pm1;
            const int ori = m->efo[6 * e + i] % 8;
            for (int j = 0; j < pm1 * pm1; j++)
                d[o++] = (int)(nvdofs + nedofs + fbase + quad_dof_ord(ori, j, pm1));
        }
        const long bbase = e * nb + nvdofs + nedofs + nfdofs;
        for (long j = 0; j < nb; j++) d[o++] = (int)(bbase + j);
    }
    return (int)(nvdofs + nedofs + nfdofs + nb * m->ne);
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int orc_h1_boundary_dofs(const orc_mesh *m, int p, const int *elem_dof, int *ess, int max_ess)
{
    /* DOFs on boundary faces of the unit cube (every boundary attribute
     * marked; GetEssentialTrueDofs, fem/fespace.cpp:550-680). */
    const int p1 = p + 1, nd = p1 * p1 * p1;
    int *dof_map = (int *)malloc(sizeof(int) * nd);
    orc_h1_dof_map(p, dof_map);
    int ndofs = 0;
    for (long e = 0; e < (long)m->ne * nd; e++)
        if (elem_dof[e] + 1 > ndofs) ndofs = elem_dof[e] + 1;
    char *mark = (char *)calloc((size_t)ndofs, 1);
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    /* file meshes: DOFs of the listed boundary faces; local face j of the
     * hex is z = 0, y = 0, x = p, y = p, x = 0, z = p (fem/geom.cpp:1030-1035) */
    for (long e = 0; m->f_bdr && e < m->ne; e++)
        for (int j = 0; j < 6; j++) {
            if (!m->f_bdr[m->ef[6 * e + j]]) continue;
            for (int k = 0; k < p1; k++)
                for (int jj = 0; jj < p1; jj++)
                    for (int i = 0; i < p1; i++) {
                        const int on = (j == 0 && k == 0) || (j == 1 && jj == 0) ||
                                       (j == 2 && i == p) || (j == 3 && jj == p) ||
                                       (j == 4 && i == 0) || (j == 5 && k == p);
                        if (on) mark[elem_dof[e * nd + dof_map[i + p1 * (jj + p1 * k)]]] = 1;
                    }
        }
    for (long e = 0; !m->f_bdr && e < m->ne; e++) {
        const int v0 = m->ev[8 * e];
        const int x = v0 % (nx + 1), y = (v0 / (nx + 1)) % (ny + 1), z = v0 / ((nx + 1) * (ny + 1));
        for (int k = 0; k < p1; k++)
            for (int j = 0; j < p1; j++)
                for (int i = 0; i < p1; i++) {
                    const int on = (x == 0 && i == 0) || (x == nx - 1 && i == p) ||
                                   (y == 0 && j == 0) || (y == ny - 1 && j == p) ||
                                   (z == 0 && k == 0) || (z == nz - 1 && k == p);
                    if (on) mark[elem_dof[e * nd + dof_map[i + p1 * (j + p1 * k)]]] = 1;
                }
    }
    int n = 0;
    for (int i = 0; i < ndofs; i++)
        if (mark[i]) {
            if (n < max_ess) ess[n] = i;
            n++;
        }
    free(mark);
    free(dof_map);
    if (n <= max_ess) qsort(ess, (size_t)n, sizeof(int), cmp_int);
    return n;
}

/* ======================================================================
 * ElementRestriction (fem/restriction.cpp:26-228).
 * ====================================================================== */
void orc_restriction_build(int ne, int dof, int ndofs, const int *elem_dof,
                           const int *dof_map, int *gather_map, int *offsets,
                           int *indices)
{
    /* restriction.cpp:73-113 */
    for (int i = 0; i <= ndofs; i++) offsets[i] = 0;
    for (long e = 0; e < ne; e++)
        for (int d = 0; d < dof; d++) {
            const int sgid = elem_dof[dof * e + d];
            const int gid = sgid >= 0 ? sgid : -1 - sgid;
            ++offsets[gid + 1];
        }
    for (int i = 1; i <= ndofs; i++) offsets[i] += offsets[i - 1];
    for (long e = 0; e < ne; e++)
        for (int d = 0; d < dof; d++) {
            const int sdid = dof_map ? dof_map[d] : 0;
            const int did = dof_map ? (sdid >= 0 ? sdid : -1 - sdid) : d;
            const int sgid = elem_dof[dof * e + did];
            const int gid = sgid >= 0 ? sgid : -1 - sgid;
            const int lid = (int)(dof * e + d);
            const int plus = (sgid >= 0 && sdid >= 0) || (sgid < 0 && sdid < 0);
            gather_map[lid] = plus ? gid : -1 - gid;
            indices[offsets[gid]++] = plus ? lid : -1 - lid;
        }
    for (int i = ndofs; i > 0; --i) offsets[i] = offsets[i - 1];
    offsets[0] = 0;
}

void orc_restriction_mult(int ne, int dof, const int *gather_map,
                          const double *x, double *y)
{
    /* restriction.cpp:116-136 (vdim = 1) */
    const long n = (long)ne * dof;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        const int gid = gather_map[i];
        const int plus = gid >= 0;
        const int j = plus ? gid : -1 - gid;
        y[i] = plus ? x[j] : -x[j];
    }
}

void orc_restriction_mult_transpose(int ndofs, int dof, const int *offsets,
                                    const int *indices, const double *x,
                                    double *y, int mode)
{
    /* restriction.cpp:159-228: one L-DOF at a time, summing its E-entries
     * in the order of the transpose table. */
    (void)dof;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ndofs; i++) {
        double v = 0.0;
        for (int j = offsets[i]; j < offsets[i + 1]; j++) {
            const int idx = indices[j] >= 0 ? indices[j] : -1 - indices[j];
            if (mode == 2 || indices[j] >= 0) v += x[idx];
            else v += -x[idx];
        }
        if (mode == 1) y[i] += v;
        else y[i] = v;
    }
}

/* ======================================================================
 * p-refinement transfer (fem/transfer.cpp:2619-2990).
 * ====================================================================== */
void orc_transfer_basis(int pl, int ph, double *B)
{
    /* the high-order space's nodes in lexicographic order are the tensor
     * GLL points (transfer.cpp:2641-2650); the low-order basis evaluated
     * there gives the DofToQuad B of :2653-2658 */
    double xh[16], wh[16], G[16 * 16];
    orc_gauss_lobatto(ph + 1, xh, wh);
    orc_basis_at(pl, ph + 1, xh, B, G);
}

void orc_transfer_prolong(int ne, int pl, int ph, const int *gmap_l, const int *offsets_h,
                          const int *indices_h, int ndofs_h, const double *mask_h,
                          const double *xl, double *yh)
{
    /* Prolongation3D (transfer.cpp:2741-2818), between the two restrictions */
    const int D = pl + 1, Q = ph + 1;
    const long nl = (long)ne * D * D * D, nh = (long)ne * Q * Q * Q;
    double B[16 * 16];
    orc_transfer_basis(pl, ph, B);
    double *xe = (double *)malloc(sizeof(double) * (size_t)nl);
    double *ye = (double *)malloc(sizeof(double) * (size_t)nh);
    orc_restriction_mult(ne, D * D * D, gmap_l, xl, xe);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; e++) {
        const double *x = xe + (long)e * D * D * D;
        double *y = ye + (long)e * Q * Q * Q;
        for (int i = 0; i < Q * Q * Q; i++) y[i] = 0.0;
        for (int dz = 0; dz < D; dz++) {
            double sxy[16][16];
            for (int qy = 0; qy < Q; qy++)
                for (int qx = 0; qx < Q; qx++) sxy[qy][qx] = 0.0;
            for (int dy = 0; dy < D; dy++) {
                double sx[16];
                for (int qx = 0; qx < Q; qx++) sx[qx] = 0.0;
                for (int dx = 0; dx < D; dx++) {
                    const double v = x[dx + D * (dy + D * dz)];
                    for (int qx = 0; qx < Q; qx++) sx[qx] += B[qx + Q * dx] * v;
                }
                for (int qy = 0; qy < Q; qy++) {
                    const double wy = B[qy + Q * dy];
                    for (int qx = 0; qx < Q; qx++) sxy[qy][qx] += wy * sx[qx];
                }
            }
            for (int qz = 0; qz < Q; qz++) {
                const double wz = B[qz + Q * dz];
                for (int qy = 0; qy < Q; qy++)
                    for (int qx = 0; qx < Q; qx++) y[qx + Q * (qy + Q * qz)] += wz * sxy[qy][qx];
            }
        }
        const double *m = mask_h + (long)e * Q * Q * Q;
        for (int i = 0; i < Q * Q * Q; i++) y[i] *= m[i];
    }
    orc_restriction_mult_transpose(ndofs_h, Q * Q * Q, offsets_h, indices_h, ye, yh, 0);
    free(xe);
    free(ye);
}

void orc_transfer_restrict(int ne, int pl, int ph, const int *gmap_h, const int *offsets_l,
                           const int *indices_l, int ndofs_l, const double *mask_h,
                           const double *xh, double *yl)
{
    /* Restriction3D (transfer.cpp:2867-2938), between the two restrictions */
    const int D = pl + 1, Q = ph + 1;
    const long nl = (long)ne * D * D * D, nh = (long)ne * Q * Q * Q;
    double B[16 * 16];
    orc_transfer_basis(pl, ph, B);  /* Bt(d, q) = B(q, d) */
    double *xe = (double *)malloc(sizeof(double) * (size_t)nh);
    double *ye = (double *)malloc(sizeof(double) * (size_t)nl);
    orc_restriction_mult(ne, Q * Q * Q, gmap_h, xh, xe);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; e++) {
        const double *x = xe + (long)e * Q * Q * Q;
        const double *m = mask_h + (long)e * Q * Q * Q;
        double *y = ye + (long)e * D * D * D;
        for (int i = 0; i < D * D * D; i++) y[i] = 0.0;
        for (int qz = 0; qz < Q; qz++) {
            double sxy[16][16];
            for (int dy = 0; dy < D; dy++)
                for (int dx = 0; dx < D; dx++) sxy[dy][dx] = 0.0;
            for (int qy = 0; qy < Q; qy++) {
                double sx[16];
                for (int dx = 0; dx < D; dx++) sx[dx] = 0.0;
                for (int qx = 0; qx < Q; qx++) {
                    const int i = qx + Q * (qy + Q * qz);
                    const double v = m[i] * x[i];
                    for (int dx = 0; dx < D; dx++) sx[dx] += B[qx + Q * dx] * v;
                }
                for (int dy = 0; dy < D; dy++) {
                    const double wy = B[qy + Q * dy];
                    for (int dx = 0; dx < D; dx++) sxy[dy][dx] += wy * sx[dx];
                }
            }
            for (int dz = 0; dz < D; dz++) {
                const double wz = B[qz + Q * dz];
                for (int dy = 0; dy < D; dy++)
                    for (int dx = 0; dx < D; dx++) y[dx + D * (dy + D * dz)] += wz * sxy[dy][dx];
            }
        }
    }
    orc_restriction_mult_transpose(ndofs_l, D * D * D, offsets_l, indices_l, ye, yl, 0);
    free(xe);
    free(ye);
}

/* ======================================================================
 * 1D quadrature and basis.
 * ====================================================================== */
static void legendre_ld(int n, long double x, long double *pn, long double *dpn, long double *pn1)
{
    /* P_n, P_n' and P_{n-1} by the three-term recurrence. */
    long double p0 = 1.0L, p1 = x;
    if (n == 0) { *pn = 1.0L; *dpn = 0.0L; *pn1 = 0.0L; return; }
    for (int j = 2; j <= n; j++) {
        const long double p2 = ((2 * j - 1) * x * p1 - (j - 1) * p0) / j;
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *pn1 = p0;
    *dpn = n * (x * p1 - p0) / (x * x - 1.0L);
}

void orc_gauss_legendre(int n, double *x, double *w)
{
    /* fem/intrules.cpp:620-705: Newton on P_n from cos(pi (i-1/4)/(n+1/2)),
     * mapped to (0,1); done in long double. */
    const int m = (n + 1) / 2;
    for (int i = 1; i <= m; i++) {
        long double z = cosl(3.14159265358979323846264338327950288L * (i - 0.25L) / (n + 0.5L));
        long double pn, dpn, pn1;
        for (int it = 0; it < 100; it++) {
            legendre_ld(n, z, &pn, &dpn, &pn1);
            const long double dz = pn / dpn;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        legendre_ld(n, z, &pn, &dpn, &pn1);
        const long double xi = (1.0L - z) / 2.0L;
        x[i - 1] = (double)xi;
        x[n - i] = (double)(1.0L - xi);
        const long double wi = 1.0L / (4.0L * xi * (1.0L - xi) * dpn * dpn);
        w[i - 1] = w[n - i] = (double)wi;
    }
    if (n == 1) { x[0] = 0.5; w[0] = 1.0; }
}

void orc_gauss_lobatto(int n, double *x, double *w)
{
    /* fem/intrules.cpp:708-830: endpoints plus the zeros of P'_{n-1}. */
    const int N = n - 1;
    if (n == 1) { x[0] = 0.5; w[0] = 1.0; return; }
    x[0] = 0.0; x[N] = 1.0;
    w[0] = w[N] = 1.0 / (N * (N + 1.0));
    for (int i = 1; i <= N / 2; i++) {
        /* Chebyshev-Gauss-Lobatto guess, Newton on q = P_N'. */
        long double z = -cosl(3.14159265358979323846264338327950288L * i / N);
        for (int it = 0; it < 100; it++) {
            long double pn, dpn, pn1;
            legendre_ld(N, z, &pn, &dpn, &pn1);
            /* (1-z^2) P_N'' = 2 z P_N' - N(N+1) P_N */
            const long double d2 = (2.0L * z * dpn - N * (N + 1.0L) * pn) / (1.0L - z * z);
            const long double dz = dpn / d2;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        long double pn, dpn, pn1;
        legendre_ld(N, z, &pn, &dpn, &pn1);
        const long double xi = (1.0L + z) / 2.0L;
        x[i] = (double)xi;
        x[N - i] = (double)(1.0L - xi);
        w[i] = w[N - i] = (double)(1.0L / (N * (N + 1.0L) * pn * pn));
    }
    if (N % 2 == 0) x[N / 2] = 0.5;
}

void orc_basis_at(int p, int n, const double *xq, double *B, double *G)
{
    /* Lagrange basis on the p+1 GLL nodes (Poly_1D::Basis, fem/fe/fe_base.cpp:
     * 1811-1873, evaluated as in :1936-2010); B[q + n*d], G[q + n*d]. */
    double z[16], wz[16];
    orc_gauss_lobatto(p + 1, z, wz);
    for (int q = 0; q < n; q++) {
        const long double x = xq[q];
        for (int d = 0; d <= p; d++) {
            long double l = 1.0L, dl = 0.0L;
            for (int k = 0; k <= p; k++) {
                if (k == d) continue;
                const long double den = (long double)z[d] - z[k];
                /* product rule, accumulating the derivative */
                dl = dl * (x - z[k]) / den + l / den;
                l = l * (x - z[k]) / den;
            }
            B[q + n * d] = (double)l;
            G[q + n * d] = (double)dl;
        }
    }
}

void orc_basis(int p, int Q, double *B, double *G)
{
    double xq[16], wq[16];
    orc_gauss_legendre(Q, xq, wq);
    orc_basis_at(p, Q, xq, B, G);
}

void orc_tensor_weights(int Q, double *W)
{
    double x[16], w[16];
    orc_gauss_legendre(Q, x, w);
    for (int k = 0; k < Q; k++)
        for (int j = 0; j < Q; j++)
            for (int i = 0; i < Q; i++) W[i + Q * (j + Q * k)] = w[i] * w[j] * w[k];
}

/* ======================================================================
 * Geometry.
 * ====================================================================== */
void orc_nodes_trilinear(const orc_mesh *m, const double *xyz, double *X)
{
    /* lex (i,j,k) in {0,1}^3 -> native vertex via the p = 1 dof_map */
    int dm[8];
    orc_h1_dof_map(1, dm);
    for (long e = 0; e < m->ne; e++)
        for (int l = 0; l < 8; l++) {
            const int v = m->ev[8 * e + dm[l]];
            for (int c = 0; c < 3; c++) X[l + 8 * (c + 3 * e)] = xyz[3 * (long)v + c];
        }
}

void orc_nodes_curved(const orc_mesh *m, const double *xyz, int pg, double amp, double *X)
{
    /* Mesh::SetCurvature(pg) interpolates the trilinear map at the GLL nodes
     * (mesh/mesh.cpp:7211); Mesh::Transform (:14056) then maps every node. */
    const int Dg = pg + 1, nd = Dg * Dg * Dg;
    const double twopi = 6.283185307179586476925286766559;
    double z[16], wz[16];
    orc_gauss_lobatto(Dg, z, wz);
    int dm[8];
    orc_h1_dof_map(1, dm);
    for (long e = 0; e < m->ne; e++) {
        double V[8][3];
        for (int l = 0; l < 8; l++)
            for (int c = 0; c < 3; c++) V[l][c] = xyz[3 * (long)m->ev[8 * e + dm[l]] + c];
        for (int k = 0; k < Dg; k++)
            for (int j = 0; j < Dg; j++)
                for (int i = 0; i < Dg; i++) {
                    const double s[3] = {z[i], z[j], z[k]};
                    double P[3] = {0, 0, 0};
                    for (int l = 0; l < 8; l++) {
                        const double N = ((l & 1) ? s[0] : 1 - s[0]) *
                                         ((l & 2) ? s[1] : 1 - s[1]) *
                                         ((l & 4) ? s[2] : 1 - s[2]);
                        for (int c = 0; c < 3; c++) P[c] += N * V[l][c];
                    }
                    const double xp = P[0] + amp * sin(twopi * P[1]) * sin(twopi * P[2]);
                    const double yp = P[1] + amp * sin(twopi * P[2]) * sin(twopi * P[0]);
                    const double zp = P[2] + amp * sin(twopi * P[0]) * sin(twopi * P[1]);
                    const int l = i + Dg * (j + Dg * k);
                    X[l + nd * (0 + 3 * e)] = xp;
                    X[l + nd * (1 + 3 * e)] = yp;
                    X[l + nd * (2 + 3 * e)] = zp;
                }
    }
}

void orc_geom(int ne, int pg, int Q, const double *X, double *J, double *P)
{
    /* QuadratureInterpolator derivatives/values by nodes
     * (fem/qinterp/grad.hpp:233-375, eval.hpp): three 1D contractions. */
    const int Dg = pg + 1, nd = Dg * Dg * Dg, nq = Q * Q * Q;
    /* stage buffers hold Q Dg^2 and Q^2 Dg values: sized for Q, Dg <= 11 */
    double B[256], G[256];
    if (Q > 11 || Dg > 11) abort();
    orc_basis(pg, Q, B, G);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double t1[2][1331], t2[3][1331];
        for (int c = 0; c < 3; c++) {
            const double *x = X + nd * (c + 3 * e);
            /* contract x: t1[0]=B, t1[1]=G over dx -> (qx,dy,dz) */
            for (int dz = 0; dz < Dg; dz++)
                for (int dy = 0; dy < Dg; dy++)
                    for (int qx = 0; qx < Q; qx++) {
                        double sb = 0, sg = 0;
                        for (int dx = 0; dx < Dg; dx++) {
                            const double v = x[dx + Dg * (dy + Dg * dz)];
                            sb += B[qx + Q * dx] * v;
                            sg += G[qx + Q * dx] * v;
                        }
                        t1[0][qx + Q * (dy + Dg * dz)] = sb;
                        t1[1][qx + Q * (dy + Dg * dz)] = sg;
                    }
            /* contract y: (qx,qy,dz): BB, BG(y-deriv), GB(x-deriv) */
            for (int dz = 0; dz < Dg; dz++)
                for (int qy = 0; qy < Q; qy++)
                    for (int qx = 0; qx < Q; qx++) {
                        double bb = 0, bgy = 0, gxb = 0;
                        for (int dy = 0; dy < Dg; dy++) {
                            const double u = t1[0][qx + Q * (dy + Dg * dz)];
                            const double g = t1[1][qx + Q * (dy + Dg * dz)];
                            bb += B[qy + Q * dy] * u;
                            bgy += G[qy + Q * dy] * u;
                            gxb += B[qy + Q * dy] * g;
                        }
                        t2[0][qx + Q * (qy + Q * dz)] = bb;
                        t2[1][qx + Q * (qy + Q * dz)] = bgy;
                        t2[2][qx + Q * (qy + Q * dz)] = gxb;
                    }
            for (int qz = 0; qz < Q; qz++)
                for (int qy = 0; qy < Q; qy++)
                    for (int qx = 0; qx < Q; qx++) {
                        double val = 0, d0 = 0, d1 = 0, d2 = 0;
                        for (int dz = 0; dz < Dg; dz++) {
                            const int id = qx + Q * (qy + Q * dz);
                            val += B[qz + Q * dz] * t2[0][id];
                            d0 += B[qz + Q * dz] * t2[2][id];
                            d1 += B[qz + Q * dz] * t2[1][id];
                            d2 += G[qz + Q * dz] * t2[0][id];
                        }
                        const long q = qx + Q * (qy + Q * qz);
                        if (J) {
                            J[q + nq * (c + 3 * (0 + 3 * e))] = d0;
                            J[q + nq * (c + 3 * (1 + 3 * e))] = d1;
                            J[q + nq * (c + 3 * (2 + 3 * e))] = d2;
                        }
                        if (P) P[q + nq * (c + 3 * e)] = val;
                    }
        }
    }
}

/* The seeded log-normal coefficient of the batched and served workloads
 * (no MFEM counterpart; MFEM would project it as a FunctionCoefficient,
 * fem/coefficient.cpp:52-72): k = exp(g), g(x) = sum_m a cos(2 pi k_m . x +
 * phi_m) over 8 modes.  Mode m draws its integer wave vector k_m uniformly
 * from [-3,3]^3 \ {0} (three draws floor(7u) - 3, redrawn while all are
 * zero), then its phase phi_m = 2 pi u, all from the stream u01(seed, ctr),
 * ctr = 0, 1, 2, ...; every mode has amplitude a = 0.5 / sqrt(8): the
 * root sum of squares of the 8 amplitudes is the workload's 0.5, g has RMS
 * 0.5 / sqrt(2) for distinct modes, and |g| <= 8 a = sqrt(2). */
void orc_fourier_modes(unsigned long long seed, int *kv, double *phi, double *amp)
{
    unsigned long long ctr = 0;
    for (int m = 0; m < 8; m++) {
        int k[3];
        do {
            for (int c = 0; c < 3; c++) k[c] = (int)floor(u01(seed, ctr++) * 7.0) - 3;
        } while (k[0] == 0 && k[1] == 0 && k[2] == 0);
        for (int c = 0; c < 3; c++) kv[3 * m + c] = k[c];
        phi[m] = 6.283185307179586 * u01(seed, ctr++);
    }
    *amp = 0.5 / sqrt(8.0);
}

/* k = exp(sum_{m < nmodes} amp cos(2 pi kv_m . x + phi_m)) at the physical
 * points P[nq][3][ne] into C[nq][ne]. */
void orc_coeff_fourier(int nmodes, const int *kv, const double *phi, double amp, int nq, int ne,
                       const double *P, double *C)
{
    const double twopi = 6.283185307179586476925286766559;
    for (long e = 0; e < ne; e++)
        for (int q = 0; q < nq; q++) {
            const double x = P[q + (long)nq * (0 + 3 * e)];
            const double y = P[q + (long)nq * (1 + 3 * e)];
            const double z = P[q + (long)nq * (2 + 3 * e)];
            double g = 0.0;
            for (int m = 0; m < nmodes; m++)
                g += amp * cos(twopi * (kv[3 * m] * x + kv[3 * m + 1] * y + kv[3 * m + 2] * z) +
                               phi[m]);
            C[q + (long)nq * e] = exp(g);
        }
}

void orc_coeff_sincos(int nq, int ne, const double *P, double *C)
{
    /* FunctionCoefficient projected at the physical quadrature points
     * (CoefficientVector::Project, fem/coefficient.cpp:2058-2076). */
    const double twopi = 6.283185307179586476925286766559;
    for (long e = 0; e < ne; e++)
        for (int q = 0; q < nq; q++) {
            const double x = P[q + (long)nq * (0 + 3 * e)];
            const double y = P[q + (long)nq * (1 + 3 * e)];
            const double z = P[q + (long)nq * (2 + 3 * e)];
            C[q + (long)nq * e] = 1.0 + 0.5 * sin(twopi * x) * cos(twopi * y) * sin(twopi * z);
        }
}

/* ======================================================================
 * PA setup.
 * ====================================================================== */
void orc_diffusion_setup(int Q, int ne, const double *W, const double *J,
                         const double *C, int c_size, double *Dd)
{
    /* bilininteg_diffusion_kernels.cpp:262-386, scalar coefficient branch. */
    const int nq = Q * Q * Q;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++)
        for (int q = 0; q < nq; q++) {
#define JJ(r, c) J[q + (long)nq * ((r) + 3 * ((c) + 3 * e))]
            const double J11 = JJ(0, 0), J21 = JJ(1, 0), J31 = JJ(2, 0);
            const double J12 = JJ(0, 1), J22 = JJ(1, 1), J32 = JJ(2, 1);
            const double J13 = JJ(0, 2), J23 = JJ(1, 2), J33 = JJ(2, 2);
#undef JJ
            const double detJ = J11 * (J22 * J33 - J32 * J23) -
                                J21 * (J12 * J33 - J32 * J13) +
                                J31 * (J12 * J23 - J22 * J13);
            const double w_detJ = W[q] / detJ;
            const double A11 = (J22 * J33) - (J23 * J32);
            const double A12 = (J32 * J13) - (J12 * J33);
            const double A13 = (J12 * J23) - (J22 * J13);
            const double A21 = (J31 * J23) - (J21 * J33);
            const double A22 = (J11 * J33) - (J13 * J31);
            const double A23 = (J21 * J13) - (J11 * J23);
            const double A31 = (J21 * J32) - (J31 * J22);
            const double A32 = (J31 * J12) - (J11 * J32);
            const double A33 = (J11 * J22) - (J12 * J21);
            const double k = c_size == 1 ? C[0] : C[q + (long)nq * e];
            double *d = Dd + (long)nq * 6 * e;
            d[q + nq * 0] = w_detJ * k * (A11 * A11 + A12 * A12 + A13 * A13);
            d[q + nq * 1] = w_detJ * k * (A11 * A21 + A12 * A22 + A13 * A23);
            d[q + nq * 2] = w_detJ * k * (A11 * A31 + A12 * A32 + A13 * A33);
            d[q + nq * 3] = w_detJ * k * (A21 * A21 + A22 * A22 + A23 * A23);
            d[q + nq * 4] = w_detJ * k * (A21 * A31 + A22 * A32 + A23 * A33);
            d[q + nq * 5] = w_detJ * k * (A31 * A31 + A32 * A32 + A33 * A33);
        }
}

void orc_mass_setup(int Q, int ne, const double *W, const double *J,
                    const double *C, int c_size, double *Mv)
{
    /* bilininteg_mass_pa.cpp:76-88: v = W * k * detJ */
    const int nq = Q * Q * Q;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++)
        for (int q = 0; q < nq; q++) {
#define JJ(r, c) J[q + (long)nq * ((r) + 3 * ((c) + 3 * e))]
            const double detJ = JJ(0, 0) * (JJ(1, 1) * JJ(2, 2) - JJ(2, 1) * JJ(1, 2)) -
                                JJ(1, 0) * (JJ(0, 1) * JJ(2, 2) - JJ(2, 1) * JJ(0, 2)) +
                                JJ(2, 0) * (JJ(0, 1) * JJ(1, 2) - JJ(1, 1) * JJ(0, 2));
#undef JJ
            const double k = c_size == 1 ? C[0] : C[q + (long)nq * e];
            Mv[q + (long)nq * e] = W[q] * k * detJ;
        }
}

/* ======================================================================
 * PA apply (E -> E, accumulate).
 * ====================================================================== */
void orc_diffusion_apply(int D1, int Q1, int ne, const double *B, const double *G,
                         const double *Dd, const double *x, double *y)
{
    /* SmemPADiffusionApply3D (bilininteg_diffusion_kernels.hpp:989-1221) run
     * as host loops: x, y, z forward contractions, symmetric 3x3 pointwise
     * product, z, y, x backward contractions. */
    const int D = D1, Q = Q1, nd = D * D * D, nq = Q * Q * Q;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double DDQ0[1000], DDQ1[1000], DQQ0[1000], DQQ1[1000], DQQ2[1000];
        double QQQ0[1000], QQQ1[1000], QQQ2[1000];
        const double *X = x + (long)nd * e;
        double *Y = y + (long)nd * e;
        const double *d = Dd + (long)nq * 6 * e;
        for (int dz = 0; dz < D; dz++)
            for (int dy = 0; dy < D; dy++)
                for (int qx = 0; qx < Q; qx++) {
                    double u = 0, v = 0;
                    for (int dx = 0; dx < D; dx++) {
                        const double s = X[dx + D * (dy + D * dz)];
                        u += s * B[qx + Q * dx];
                        v += s * G[qx + Q * dx];
                    }
                    DDQ0[qx + Q * (dy + D * dz)] = u;
                    DDQ1[qx + Q * (dy + D * dz)] = v;
                }
        for (int dz = 0; dz < D; dz++)
            for (int qy = 0; qy < Q; qy++)
                for (int qx = 0; qx < Q; qx++) {
                    double u = 0, v = 0, w = 0;
                    for (int dy = 0; dy < D; dy++) {
                        u += DDQ1[qx + Q * (dy + D * dz)] * B[qy + Q * dy];
                        v += DDQ0[qx + Q * (dy + D * dz)] * G[qy + Q * dy];
                        w += DDQ0[qx + Q * (dy + D * dz)] * B[qy + Q * dy];
                    }
                    DQQ0[qx + Q * (qy + Q * dz)] = u;
                    DQQ1[qx + Q * (qy + Q * dz)] = v;
                    DQQ2[qx + Q * (qy + Q * dz)] = w;
                }
        for (int qz = 0; qz < Q; qz++)
            for (int qy = 0; qy < Q; qy++)
                for (int qx = 0; qx < Q; qx++) {
                    double u = 0, v = 0, w = 0;
                    for (int dz = 0; dz < D; dz++) {
                        u += DQQ0[qx + Q * (qy + Q * dz)] * B[qz + Q * dz];
                        v += DQQ1[qx + Q * (qy + Q * dz)] * B[qz + Q * dz];
                        w += DQQ2[qx + Q * (qy + Q * dz)] * G[qz + Q * dz];
                    }
                    const int q = qx + Q * (qy + Q * qz);
                    const double O11 = d[q + nq * 0], O12 = d[q + nq * 1], O13 = d[q + nq * 2];
                    const double O22 = d[q + nq * 3], O23 = d[q + nq * 4], O33 = d[q + nq * 5];
                    QQQ0[q] = O11 * u + O12 * v + O13 * w;
                    QQQ1[q] = O12 * u + O22 * v + O23 * w;
                    QQQ2[q] = O13 * u + O23 * v + O33 * w;
                }
        /* backward: z */
        for (int dz = 0; dz < D; dz++)
            for (int qy = 0; qy < Q; qy++)
                for (int qx = 0; qx < Q; qx++) {
                    double u = 0, v = 0, w = 0;
                    for (int qz = 0; qz < Q; qz++) {
                        const int q = qx + Q * (qy + Q * qz);
                        u += QQQ0[q] * B[qz + Q * dz];
                        v += QQQ1[q] * B[qz + Q * dz];
                        w += QQQ2[q] * G[qz + Q * dz];
                    }
                    DQQ0[qx + Q * (qy + Q * dz)] = u;
                    DQQ1[qx + Q * (qy + Q * dz)] = v;
                    DQQ2[qx + Q * (qy + Q * dz)] = w;
                }
        /* backward: y */
        for (int dz = 0; dz < D; dz++)
            for (int dy = 0; dy < D; dy++)
                for (int qx = 0; qx < Q; qx++) {
                    double u = 0, v = 0, w = 0;
                    for (int qy = 0; qy < Q; qy++) {
                        u += DQQ0[qx + Q * (qy + Q * dz)] * B[qy + Q * dy];
                        v += DQQ1[qx + Q * (qy + Q * dz)] * G[qy + Q * dy];
                        w += DQQ2[qx + Q * (qy + Q * dz)] * B[qy + Q * dy];
                    }
                    DDQ0[qx + Q * (dy + D * dz)] = u;
                    DDQ1[qx + Q * (dy + D * dz)] = v + w;
                }
        /* backward: x */
        for (int dz = 0; dz < D; dz++)
            for (int dy = 0; dy < D; dy++)
                for (int dx = 0; dx < D; dx++) {
                    double u = 0, v = 0;
                    for (int qx = 0; qx < Q; qx++) {
                        u += DDQ0[qx + Q * (dy + D * dz)] * G[qx + Q * dx];
                        v += DDQ1[qx + Q * (dy + D * dz)] * B[qx + Q * dx];
                    }
                    Y[dx + D * (dy + D * dz)] += u + v;
                }
    }
}

void orc_mass_apply(int D1, int Q1, int ne, const double *B, const double *Mv,
                    const double *x, double *y)
{
    /* SmemPAMassApply3D (bilininteg_mass_kernels.hpp:815-1049) as host loops. */
    const int D = D1, Q = Q1, nd = D * D * D, nq = Q * Q * Q;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double a[1000], b[1000], c[1000];
        const double *X = x + (long)nd * e;
        double *Y = y + (long)nd * e;
        const double *v = Mv + (long)nq * e;
        for (int dz = 0; dz < D; dz++)
            for (int dy = 0; dy < D; dy++)
                for (int qx = 0; qx < Q; qx++) {
                    double s = 0;
                    for (int dx = 0; dx < D; dx++) s += X[dx + D * (dy + D * dz)] * B[qx + Q * dx];
                    a[qx + Q * (dy + D * dz)] = s;
                }
        for (int dz = 0; dz < D; dz++)
            for (int qy = 0; qy < Q; qy++)
                for (int qx = 0; qx < Q; qx++) {
                    double s = 0;
                    for (int dy = 0; dy < D; dy++) s += a[qx + Q * (dy + D * dz)] * B[qy + Q * dy];
                    b[qx + Q * (qy + Q * dz)] = s;
                }
        for (int qz = 0; qz < Q; qz++)
            for (int qy = 0; qy < Q; qy++)
                for (int qx = 0; qx < Q; qx++) {
                    double s = 0;
                    for (int dz = 0; dz < D; dz++) s += b[qx + Q * (qy + Q * dz)] * B[qz + Q * dz];
                    const int q = qx + Q * (qy + Q * qz);
                    c[q] = s * v[q];
                }
        for (int dz = 0; dz < D; dz++)
            for (int qy = 0; qy < Q; qy++)
                for (int qx = 0; qx < Q; qx++) {
                    double s = 0;
                    for (int qz = 0; qz < Q; qz++) s += c[qx + Q * (qy + Q * qz)] * B[qz + Q * dz];
                    b[qx + Q * (qy + Q * dz)] = s;
                }
        for (int dz = 0; dz < D; dz++)
            for (int dy = 0; dy < D; dy++)
                for (int qx = 0; qx < Q; qx++) {
                    double s = 0;
                    for (int qy = 0; qy < Q; qy++) s += b[qx + Q * (qy + Q * dz)] * B[qy + Q * dy];
                    a[qx + Q * (dy + D * dz)] = s;
                }
        for (int dz = 0; dz < D; dz++)
            for (int dy = 0; dy < D; dy++)
                for (int dx = 0; dx < D; dx++) {
                    double s = 0;
                    for (int qx = 0; qx < Q; qx++) s += a[qx + Q * (dy + D * dz)] * B[qx + Q * dx];
                    Y[dx + D * (dy + D * dz)] += s;
                }
    }
}

/* ======================================================================
 * PA diagonal (E accumulate).
 * ====================================================================== */
void orc_diffusion_diag(int D1, int Q1, int ne, const double *B, const double *G,
                        const double *Dd, double *y)
{
    /* PADiffusionDiagonal3D (bilininteg_diffusion_kernels.hpp:268-367). */
    const int D = D1, Q = Q1, nd = D * D * D, nq = Q * Q * Q;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double QQD[1000], QDD[1000];
        double *Y = y + (long)nd * e;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const int k = j >= i ? 3 - (3 - i) * (2 - i) / 2 + j : 3 - (3 - j) * (2 - j) / 2 + i;
                for (int qx = 0; qx < Q; qx++)
                    for (int qy = 0; qy < Q; qy++)
                        for (int dz = 0; dz < D; dz++) {
                            double s = 0;
                            for (int qz = 0; qz < Q; qz++) {
                                const int q = qx + (qy + qz * Q) * Q;
                                const double O = Dd[q + (long)nq * (k + 6 * e)];
                                const double Bz = B[qz + Q * dz], Gz = G[qz + Q * dz];
                                const double L = i == 2 ? Gz : Bz, R = j == 2 ? Gz : Bz;
                                s += L * O * R;
                            }
                            QQD[qx + Q * (qy + Q * dz)] = s;
                        }
                for (int qx = 0; qx < Q; qx++)
                    for (int dz = 0; dz < D; dz++)
                        for (int dy = 0; dy < D; dy++) {
                            double s = 0;
                            for (int qy = 0; qy < Q; qy++) {
                                const double By = B[qy + Q * dy], Gy = G[qy + Q * dy];
                                const double L = i == 1 ? Gy : By, R = j == 1 ? Gy : By;
                                s += L * QQD[qx + Q * (qy + Q * dz)] * R;
                            }
                            QDD[qx + Q * (dy + D * dz)] = s;
                        }
                for (int dz = 0; dz < D; dz++)
                    for (int dy = 0; dy < D; dy++)
                        for (int dx = 0; dx < D; dx++) {
                            double s = 0;
                            for (int qx = 0; qx < Q; qx++) {
                                const double Bx = B[qx + Q * dx], Gx = G[qx + Q * dx];
                                const double L = i == 0 ? Gx : Bx, R = j == 0 ? Gx : Bx;
                                s += L * QDD[qx + Q * (dy + D * dz)] * R;
                            }
                            Y[dx + D * (dy + D * dz)] += s;
                        }
            }
    }
}

void orc_mass_diag(int D1, int Q1, int ne, const double *B, const double *Mv, double *y)
{
    /* PAMassAssembleDiagonal3D (bilininteg_mass_kernels.hpp:261-330). */
    const int D = D1, Q = Q1, nd = D * D * D, nq = Q * Q * Q;
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double QQD[1000], QDD[1000];
        const double *v = Mv + (long)nq * e;
        double *Y = y + (long)nd * e;
        for (int qx = 0; qx < Q; qx++)
            for (int qy = 0; qy < Q; qy++)
                for (int dz = 0; dz < D; dz++) {
                    double s = 0;
                    for (int qz = 0; qz < Q; qz++)
                        s += B[qz + Q * dz] * B[qz + Q * dz] * v[qx + Q * (qy + Q * qz)];
                    QQD[qx + Q * (qy + Q * dz)] = s;
                }
        for (int qx = 0; qx < Q; qx++)
            for (int dz = 0; dz < D; dz++)
                for (int dy = 0; dy < D; dy++) {
                    double s = 0;
                    for (int qy = 0; qy < Q; qy++)
                        s += B[qy + Q * dy] * B[qy + Q * dy] * QQD[qx + Q * (qy + Q * dz)];
                    QDD[qx + Q * (dy + D * dz)] = s;
                }
        for (int dz = 0; dz < D; dz++)
            for (int dy = 0; dy < D; dy++)
                for (int dx = 0; dx < D; dx++) {
                    double s = 0;
                    for (int qx = 0; qx < Q; qx++)
                        s += B[qx + Q * dx] * B[qx + Q * dx] * QDD[qx + Q * (dy + D * dz)];
                    Y[dx + D * (dy + D * dz)] += s;
                }
    }
}

/* ======================================================================
 * L-vector operator, constrained operator, Jacobi, CG.
 * ====================================================================== */
void orc_pa_mult(const orc_pa_op *op, const double *x, double *y)
{
    /* PABilinearFormExtension::MultInternal (fem/bilinearform_ext.cpp:487-560):
     * gather, localY = 0, each integrator adds, MultTranspose. */
    const long nE = (long)op->ne * op->D1 * op->D1 * op->D1;
    const int nd = op->D1 * op->D1 * op->D1;
    double *xe = (double *)malloc(sizeof(double) * nE);
    double *ye = (double *)calloc((size_t)nE, sizeof(double));
    orc_restriction_mult(op->ne, nd, op->gather_map, x, xe);
    if (op->Dd) orc_diffusion_apply(op->D1, op->Q1, op->ne, op->B, op->G, op->Dd, xe, ye);
    if (op->Mv) orc_mass_apply(op->D1, op->Q1, op->ne, op->B, op->Mv, xe, ye);
    orc_restriction_mult_transpose(op->ndofs, nd, op->offsets, op->indices, ye, y, 0);
    free(xe);
    free(ye);
}

void orc_pa_diagonal(const orc_pa_op *op, double *diag)
{
    /* PABilinearFormExtension::AssembleDiagonal (bilinearform_ext.cpp:370-454) */
    const long nE = (long)op->ne * op->D1 * op->D1 * op->D1;
    const int nd = op->D1 * op->D1 * op->D1;
    double *ye = (double *)calloc((size_t)nE, sizeof(double));
    if (op->Dd) orc_diffusion_diag(op->D1, op->Q1, op->ne, op->B, op->G, op->Dd, ye);
    if (op->Mv) orc_mass_diag(op->D1, op->Q1, op->ne, op->B, op->Mv, ye);
    orc_restriction_mult_transpose(op->ndofs, nd, op->offsets, op->indices, ye, diag, 2);
    free(ye);
}

void orc_constrained_mult(const orc_pa_op *op, const double *x, double *y)
{
    /* ConstrainedOperator::ConstrainedMult, DIAG_ONE (operator.cpp:601-662) */
    const int n = op->ndofs;
    double *z = (double *)malloc(sizeof(double) * n);
    memcpy(z, x, sizeof(double) * n);
    for (int i = 0; i < op->n_ess; i++) z[op->ess[i]] = 0.0;
    orc_pa_mult(op, z, y);
    for (int i = 0; i < op->n_ess; i++) y[op->ess[i]] = x[op->ess[i]];
    free(z);
}

void orc_form_linear_system(const orc_pa_op *op, double *x, const double *b,
                            double *X, double *Bv)
{
    /* Operator::FormLinearSystem (operator.cpp:129-145) with copy_interior = 0
     * and ConstrainedOperator::EliminateRHS (:574-599). */
    const int n = op->ndofs;
    double *w = (double *)calloc((size_t)n, sizeof(double));
    double *z = (double *)malloc(sizeof(double) * n);
    char *is_ess = (char *)calloc((size_t)n, 1);
    for (int i = 0; i < op->n_ess; i++) is_ess[op->ess[i]] = 1;
    for (int i = 0; i < n; i++) X[i] = is_ess[i] ? x[i] : 0.0;
    for (int i = 0; i < op->n_ess; i++) w[op->ess[i]] = X[op->ess[i]];
    orc_pa_mult(op, w, z);
    for (int i = 0; i < n; i++) Bv[i] = b[i] - z[i];
    for (int i = 0; i < op->n_ess; i++) Bv[op->ess[i]] = X[op->ess[i]];
    free(w); free(z); free(is_ess);
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int orc_pcg(const orc_pa_op *op, const double *diag, const double *b, double *x,
            double rel_tol, double abs_tol, int max_iter, int *iters,
            double *final_norm, double *initial_norm)
{
    /* OperatorJacobiSmoother::Setup/Mult (solvers.cpp:401-453, damping 1)
     * and CGSolver::Mult with iterative_mode = false (solvers.cpp:869-1050),
     * on the DIAG_ONE constrained operator. */
    const int n = op->ndofs;
    double *dinv = (double *)malloc(sizeof(double) * n);
    double *r = (double *)malloc(sizeof(double) * n);
    double *d = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) dinv[i] = 1.0 / diag[i];
    for (int i = 0; i < op->n_ess; i++) dinv[op->ess[i]] = 1.0;

    int converged = 0, final_iter = 0;
    double nom, nom0, den, betanom = 0.0, alpha, beta, r0;
    memcpy(r, b, sizeof(double) * n);
    for (int i = 0; i < n; i++) x[i] = 0.0;
    for (int i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    memcpy(d, z, sizeof(double) * n);
    nom0 = nom = dot(n, d, r);
    *initial_norm = nom0 >= 0.0 ? sqrt(nom0) : nom0;
    if (nom < 0.0) {
        *iters = 0; *final_norm = nom; converged = 0;
        goto done;
    }
    r0 = nom * rel_tol * rel_tol;
    if (abs_tol * abs_tol > r0) r0 = abs_tol * abs_tol;
    if (nom <= r0) {
        *iters = 0; *final_norm = sqrt(nom); converged = 1;
        goto done;
    }
    orc_constrained_mult(op, d, z);
    den = dot(n, z, d);
    if (den <= 0.0 && den == 0.0) {
        *iters = 0; *final_norm = sqrt(nom); converged = 0;
        goto done;
    }
    final_iter = max_iter;
    for (int i = 1;;) {
        alpha = nom / den;
        for (int k = 0; k < n; k++) x[k] += alpha * d[k];
        for (int k = 0; k < n; k++) r[k] -= alpha * z[k];
        for (int k = 0; k < n; k++) z[k] = dinv[k] * r[k];
        betanom = dot(n, r, z);
        if (betanom < 0.0) { converged = 0; final_iter = i; break; }
        if (betanom <= r0) { converged = 1; final_iter = i; break; }
        if (++i > max_iter) break;
        beta = betanom / nom;
        for (int k = 0; k < n; k++) d[k] = z[k] + beta * d[k];
        orc_constrained_mult(op, d, z);
        den = dot(n, d, z);
        if (den <= 0.0 && den == 0.0) { final_iter = i; break; }
        nom = betanom;
    }
    *iters = final_iter;
    *final_norm = sqrt(betanom);
done:
    free(dinv); free(r); free(d); free(z);
    return converged;
}

/* ======================================================================
 * RHS: DomainLFIntegrator with a constant f (fem/lininteg.cpp:51-79).
 * ====================================================================== */
void orc_domain_lf(int p, int ne, int pg, const double *X, double f, double *bE)
{
    orc_domain_lf_qf(p, ne, pg, X, f, NULL, bE);
}

void orc_domain_lf_qf(int p, int ne, int pg, const double *X, double f, const double *fq,
                      double *bE)
{
    /* DomainLFIntegrator::AssembleRHSElementVect (fem/lininteg.cpp:51-81) with
     * the coefficient's value at quadrature point q of element e: fq[q + nq e]
     * (a QuadratureFunction, points x fastest), or the constant f (fq NULL) */
    const int D = p + 1, Q = p + 1, nd = D * D * D, nq = Q * Q * Q;
    double B[256], G[256], W[1000];
    double *J = (double *)malloc(sizeof(double) * (size_t)nq * 9);
    orc_basis(p, Q, B, G);
    orc_tensor_weights(Q, W);
    const int Dg = pg + 1;
    const long ndg = (long)Dg * Dg * Dg;
    for (long e = 0; e < ne; e++) {
        orc_geom(1, pg, Q, X + ndg * 3 * e, J, NULL);
        for (int q = 0; q < nq; q++) {
#define JJ(r, c) J[q + (long)nq * ((r) + 3 * (c))]
            const double detJ = JJ(0, 0) * (JJ(1, 1) * JJ(2, 2) - JJ(2, 1) * JJ(1, 2)) -
                                JJ(1, 0) * (JJ(0, 1) * JJ(2, 2) - JJ(2, 1) * JJ(0, 2)) +
                                JJ(2, 0) * (JJ(0, 1) * JJ(1, 2) - JJ(1, 1) * JJ(0, 2));
#undef JJ
            const int qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
            const double val = W[q] * detJ * (fq ? fq[q + (long)nq * e] : f);
            for (int dz = 0; dz < D; dz++)
                for (int dy = 0; dy < D; dy++)
                    for (int dx = 0; dx < D; dx++)
                        bE[dx + D * (dy + D * dz) + nd * e] +=
                            val * B[qx + Q * dx] * B[qy + Q * dy] * B[qz + Q * dz];
        }
    }
    free(J);
}

/* ======================================================================
 * Chebyshev smoother and power method.
 * ====================================================================== */

/* srand(seed) then n calls of rand() for glibc's default generator (the
 * additive feedback generator of its random_r.c, TYPE_3: degree 31,
 * separation 3), which Vector::Randomize (linalg/vector.cpp:955-966) draws
 * from.  Published form: r[0] = seed (0 -> 1); r[i] = 16807 r[i-1] mod
 * (2^31 - 1) for i = 1..30 (Schrage's method, negative results + 2^31 - 1);
 * r[i] = r[i-31] for i = 31..33; r[i] = r[i-31] + r[i-3] mod 2^32 from
 * i = 34 on; the k-th output is r[k + 344] >> 1. */
void orc_glibc_rand(unsigned seed, int n, int *out)
{
    uint32_t r[34];
    int32_t w = (int32_t)(seed == 0 ? 1u : seed);
    r[0] = (uint32_t)w;
    for (int i = 1; i < 31; i++) {
        const long hi = w / 127773, lo = w % 127773;
        long word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        w = (int32_t)word;
        r[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    /* ring of the last 34 values: r[i] lives at i % 34 */
    for (long i = 34; i < 344 + (long)n; i++) {
        const uint32_t v = r[(i - 31) % 34] + r[(i - 3) % 34];
        r[i % 34] = v;
        if (i >= 344) out[i - 344] = (int)(v >> 1);
    }
}

/* Vector::Randomize(seed): data[i] = rand() / (RAND_MAX + 1.0), RAND_MAX =
 * 2^31 - 1 (linalg/vector.hpp:61-77). */
void orc_randomize(unsigned seed, int n, double *v)
{
    int *o = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    orc_glibc_rand(seed, n, o);
    for (int i = 0; i < n; i++) v[i] = (double)o[i] / 2147483648.0;
    free(o);
}

/* OperatorChebyshevSmoother::Setup coefficients (linalg/solvers.cpp:552-628):
 * bounds 0.3 and 1.2 times max_eig, theta and delta their mid-point and
 * half-width; order 1..5.  Returns 0 on an unsupported order. */
int orc_chebyshev_coeffs(int order, double max_eig, double *c)
{
    const double upper = 1.2 * max_eig, lower = 0.3 * max_eig;
    const double th = 0.5 * (upper + lower), de = 0.5 * (upper - lower);
    switch (order - 1) {
    case 0:
        c[0] = 1.0 / th;
        return 1;
    case 1: {
        const double t0 = 1.0 / (pow(de, 2) - 2 * pow(th, 2));
        c[0] = -4 * th * t0;
        c[1] = 2 * t0;
        return 1;
    }
    case 2: {
        const double t0 = 3 * pow(de, 2), t1 = pow(th, 2);
        const double t2 = 1.0 / (-4 * pow(th, 3) + th * t0);
        c[0] = t2 * (t0 - 12 * t1);
        c[1] = 12 / (t0 - 4 * t1);
        c[2] = -4 * t2;
        return 1;
    }
    case 3: {
        const double t0 = pow(de, 2), t1 = pow(th, 2), t2 = 8 * t0;
        const double t3 = 1.0 / (pow(de, 4) + 8 * pow(th, 4) - t1 * t2);
        c[0] = t3 * (32 * pow(th, 3) - 16 * th * t0);
        c[1] = t3 * (-48 * t1 + t2);
        c[2] = 32 * th * t3;
        c[3] = -8 * t3;
        return 1;
    }
    case 4: {
        const double t0 = 5 * pow(de, 4), t1 = pow(th, 4), t2 = pow(th, 2), t3 = pow(de, 2);
        const double t4 = 60 * t3, t5 = 20 * t3;
        const double t6 = 1.0 / (16 * pow(th, 5) - pow(th, 3) * t5 + th * t0);
        const double t7 = 160 * t2;
        const double t8 = 1.0 / (t0 + 16 * t1 - t2 * t5);
        c[0] = t6 * (t0 + 80 * t1 - t2 * t4);
        c[1] = t8 * (t4 - t7);
        c[2] = t6 * (-t5 + t7);
        c[3] = -80 * t8;
        c[4] = 16 * t6;
        return 1;
    }
    default:
        return 0;
    }
}

/* Jacobi inverse with ess rows 1 (OperatorJacobiSmoother::Setup, damping 1,
 * solvers.cpp:401-425; the same in OperatorChebyshevSmoother::Setup). */
static void jacobi_inverse(const orc_pa_op *op, const double *diag, double *dinv)
{
    for (int i = 0; i < op->ndofs; i++) dinv[i] = 1.0 / diag[i];
    for (int i = 0; i < op->n_ess; i++) dinv[op->ess[i]] = 1.0;
}

/* PowerMethod::EstimateLargestEigenvalue (linalg/operator.cpp:886-935) on
 * diag^-1 A (ProductOperator of OperatorJacobiSmoother and the constrained
 * operator, solvers.cpp:496-511); v0 holds the start vector on entry. */
double orc_power_method(const orc_pa_op *op, const double *diag, double *v0, int steps,
                        double tol)
{
    const int n = op->ndofs;
    double *dinv = (double *)malloc(sizeof(double) * n);
    double *v1 = (double *)malloc(sizeof(double) * n);
    double *u = (double *)malloc(sizeof(double) * n);
    double *a = v0, *b = v1;
    jacobi_inverse(op, diag, dinv);
    double eig = 1.0;
    for (int it = 0; it < steps; it++) {
        const double nrm = dot(n, a, a);
        const double m = 1.0 / sqrt(nrm);
        for (int i = 0; i < n; i++) a[i] *= m;
        orc_constrained_mult(op, a, u);
        for (int i = 0; i < n; i++) b[i] = 0.0 + dinv[i] * u[i];
        const double enew = dot(n, a, b);
        const double diff = fabs((enew - eig) / eig);
        eig = enew;
        double *t = a; a = b; b = t;
        if (diff < tol) break;
    }
    if (a != v0) memcpy(v0, a, sizeof(double) * n);
    free(dinv); free(v1); free(u);
    return eig;
}

/* OperatorChebyshevSmoother::Mult (solvers.cpp:630-656):
 * y = sum_k c_k (diag^-1 A)^k diag^-1 x on the constrained operator. */
void orc_chebyshev_mult(const orc_pa_op *op, const double *diag, int order, const double *c,
                        const double *x, double *y)
{
    const int n = op->ndofs;
    double *dinv = (double *)malloc(sizeof(double) * n);
    double *res = (double *)malloc(sizeof(double) * n);
    double *h = (double *)malloc(sizeof(double) * n);
    jacobi_inverse(op, diag, dinv);
    memcpy(res, x, sizeof(double) * n);
    for (int i = 0; i < n; i++) y[i] = 0.0;
    for (int k = 0; k < order; k++) {
        if (k > 0) {
            orc_constrained_mult(op, res, h);
            memcpy(res, h, sizeof(double) * n);
        }
        for (int i = 0; i < n; i++) res[i] *= dinv[i];
        for (int i = 0; i < n; i++) y[i] += c[k] * res[i];
    }
    free(dinv); free(res); free(h);
}

/* CGSolver::Mult (solvers.cpp:869-1050, iterative_mode = false) with the
 * Chebyshev smoother as preconditioner, on the constrained operator. */
int orc_pcg_chebyshev(const orc_pa_op *op, const double *diag, int order, const double *c,
                      const double *b, double *x, double rel_tol, double abs_tol, int max_iter,
                      int *iters, double *final_norm, double *initial_norm)
{
    const int n = op->ndofs;
    double *r = (double *)malloc(sizeof(double) * n);
    double *d = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n);
    int converged = 0, final_iter = 0;
    double nom, nom0, den, betanom = 0.0, alpha, beta, r0;
    memcpy(r, b, sizeof(double) * n);
    for (int i = 0; i < n; i++) x[i] = 0.0;
    orc_chebyshev_mult(op, diag, order, c, r, z);
    memcpy(d, z, sizeof(double) * n);
    nom0 = nom = dot(n, d, r);
    *initial_norm = nom0 >= 0.0 ? sqrt(nom0) : nom0;
    if (nom < 0.0) {
        *iters = 0; *final_norm = nom; converged = 0;
        goto done;
    }
    r0 = nom * rel_tol * rel_tol;
    if (abs_tol * abs_tol > r0) r0 = abs_tol * abs_tol;
    if (nom <= r0) {
        *iters = 0; *final_norm = sqrt(nom); converged = 1;
        goto done;
    }
    orc_constrained_mult(op, d, z);
    den = dot(n, z, d);
    if (den <= 0.0 && den == 0.0) {
        *iters = 0; *final_norm = sqrt(nom); converged = 0;
        goto done;
    }
    final_iter = max_iter;
    for (int i = 1;;) {
        alpha = nom / den;
        for (int k = 0; k < n; k++) x[k] += alpha * d[k];
        for (int k = 0; k < n; k++) r[k] -= alpha * z[k];
        orc_chebyshev_mult(op, diag, order, c, r, z);
        betanom = dot(n, r, z);
        if (betanom < 0.0) { converged = 0; final_iter = i; break; }
        if (betanom <= r0) { converged = 1; final_iter = i; break; }
        if (++i > max_iter) break;
        beta = betanom / nom;
        for (int k = 0; k < n; k++) d[k] = z[k] + beta * d[k];
        orc_constrained_mult(op, d, z);
        den = dot(n, d, z);
        if (den <= 0.0 && den == 0.0) { final_iter = i; break; }
        nom = betanom;
    }
    *iters = final_iter;
    *final_norm = sqrt(betanom);
done:
    free(r); free(d); free(z);
    return converged;
}
