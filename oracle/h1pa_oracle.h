/*
 * h1pa_oracle.h -- CPU restatement of MFEM's H1 partial-assembly (PA) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in mfem_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function restates one piece of the reference algorithm and cites the
 * reference file:line it follows (paths relative to the mfem/mfem tree).
 * Layouts follow MFEM's Reshape convention: first index fastest.
 *
 * How each part is pinned (MFEM itself cannot be compiled here: every MFEM
 * source includes config/_config.hpp, which only MFEM's own build system
 * generates):
 *   - SFC order: permutation + unit-step (face-adjacent) path + endpoint
 *     invariants of the generalized Hilbert curve; exact 2^k Hilbert cases.
 *   - edge/face/DOF numbering: first-encounter monotonicity, multiplicity
 *     census {1,2,4,8}, boundary-DOF count, and geometric conformity (every
 *     E-DOF mapped to one L-DOF sits at the same physical point).
 *   - quadrature, basis: closed forms (numpy leggauss, polynomial exactness,
 *     partition of unity, derivative identities).
 *   - setup/apply/diagonal: brute-force dense element matrices (no sum
 *     factorization), closed forms on Cartesian meshes, A*1 = 0, 1'M1 = |Omega|,
 *     symmetry.
 *   - PCG: exact solution of a manufactured problem and the textbook CG
 *     recurrence checked against numpy's dense solve.
 */
#ifndef H1PA_ORACLE_H
#define H1PA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- space-filling curve (mesh/ncmesh.cpp:5435-5634) ------------------- */
/* coords[3*k+0..2] = (x,y,z) of the k-th cell along the curve. */
void orc_grid_sfc3d(int nx, int ny, int nz, int *coords);

/* ---- Cartesian hex mesh + topology --------------------------------------
 * mesh/mesh.cpp:3841-4121 (Make3D, sfc_ordering=true), :8526-8549 (edges),
 * :8996-9080 (faces), :8741-8766 (face orientation), :8044-8073 (edge
 * orientation), general/table.cpp:623-645 (DSTable), general/stable3d.cpp
 * (STable3D::Push4). */
typedef struct {
    int nx, ny, nz;
    int nv, ne, nedges, nfaces;
    int *ev;    /* [ne][8]   element vertices, MFEM hex order          */
    int *ee;    /* [ne][12]  element edges                             */
    int *eeo;   /* [ne][12]  edge orientation (+1 / -1)                */
    int *ef;    /* [ne][6]   element faces                             */
    int *efo;   /* [ne][6]   face orientation 0..7                     */
    int *f_elem1; /* [nfaces] first element that met the face            */
    char *f_bdr;  /* [nfaces] 1 on boundary faces (file meshes; NULL for
                     Cartesian meshes, whose boundary is the cube's)     */
} orc_mesh;

int  orc_mesh_cartesian(int nx, int ny, int nz, int sfc, orc_mesh *m);
/* A mesh given by its hexahedra (MFEM vertex order, file order kept) and its
 * boundary quadrilaterals, as Mesh::ReadMFEMMesh + FinalizeTopology build it
 * (mesh/mesh_readers.cpp:40-130); nx = ny = nz = 0.  Returns -1 when out of
 * memory, -2 when a boundary quadrilateral is not an element face. */
int  orc_mesh_from_elements(int nv, int ne, const int *ev, int nb, const int *bquads,
                            orc_mesh *m);
void orc_mesh_free(orc_mesh *m);
/* Unit-cube vertex coordinates, x fastest, [nv][3] (mesh/mesh.cpp:3870-3882). */
void orc_mesh_vertices(int nx, int ny, int nz, double *xyz);
/* Seeded perturbation of interior vertices by U(-amp*h, amp*h) per coordinate,
 * counter-based splitmix64 (synthetic input definition, shared with the product). */
void orc_perturb_vertices(int nx, int ny, int nz, double amp, unsigned long long seed, double *xyz);

/* ---- H1 space (fem/fespace.cpp:2769-2900, :3428-3535;
 *                 fem/fe_coll.cpp:1741-1960, :2087-2107) ------------------- */
/* Element-to-DOF table in native (vertex, edge, face, bubble) order.
 * elem_dof[e*(p+1)^3 + i]; returns the number of DOFs. */
int  orc_h1_elem_dofs(const orc_mesh *m, int p, int *elem_dof);
/* Lexicographic -> native map (fem/fe/fe_base.cpp:2557-2674). */
void orc_h1_dof_map(int p, int *dof_map);
/* Boundary true DOFs (all 6 attributes), ascending (fem/fespace.cpp:550-680). */
int  orc_h1_boundary_dofs(const orc_mesh *m, int p, const int *elem_dof, int *ess, int max_ess);

/* ---- ElementRestriction (fem/restriction.cpp:26-228) ------------------- */
void orc_restriction_build(int ne, int dof, int ndofs, const int *elem_dof,
                           const int *dof_map, int *gather_map, int *offsets,
                           int *indices);
void orc_restriction_mult(int ne, int dof, const int *gather_map,
                          const double *x, double *y);
/* mode 0: MultTranspose (set), 1: AddMultTranspose, 2: AbsMultTranspose */
void orc_restriction_mult_transpose(int ndofs, int dof, const int *offsets,
                                    const int *indices, const double *x,
                                    double *y, int mode);

/* ---- 1D rules and basis (fem/intrules.cpp:620-830;
 *                          fem/fe/fe_base.cpp:1811-2010, :2697-2740) ------ */
void orc_gauss_legendre(int n, double *x, double *w);
void orc_gauss_lobatto(int n, double *x, double *w);
/* B[q + Q*d] = l_d(x_q), G = l_d'(x_q): Lagrange basis on the p+1 GLL nodes
 * evaluated at the Q Gauss-Legendre points. */
void orc_basis(int p, int Q, double *B, double *G);
/* Same basis evaluated at arbitrary points x[0..n-1]. */
void orc_basis_at(int p, int n, const double *x, double *B, double *G);

/* ---- Geometry (mesh/mesh.cpp:15445-15497, fem/qinterp/grad.hpp:233-375) -- */
/* Element node E-vector X[Dg^3][3][ne] of a trilinear mesh (Dg=2) built from
 * the vertex coordinates; lex order. */
void orc_nodes_trilinear(const orc_mesh *m, const double *xyz, double *X);
/* Order-pg nodes: each lex GLL node of each element is the trilinear point
 * mapped by the cfg3 curving map x' = x + a sin(2 pi y) sin(2 pi z) (cyclic). */
void orc_nodes_curved(const orc_mesh *m, const double *xyz, int pg, double amp, double *X);
/* Jacobians J[Q^3][3][3][ne] (J(q,r,c,e) = dx_r/dxi_c) and physical points
 * P[Q^3][3][ne] from nodes of order pg. */
void orc_geom(int ne, int pg, int Q, const double *X, double *J, double *P);

/* ---- coefficients (fem/coefficient.cpp:52-72, :2058-2130) --------------- */
/* k(x) = 1 + 0.5 sin(2 pi x) cos(2 pi y) sin(2 pi z) at P[nq][3][ne] -> C[nq][ne]. */
void orc_coeff_sincos(int nq, int ne, const double *P, double *C);
void orc_fourier_modes(unsigned long long seed, int *kv, double *phi, double *amp);
void orc_coeff_fourier(int nmodes, const int *kv, const double *phi, double amp, int nq, int ne,
                       const double *P, double *C);

/* ---- PA setup (fem/integ/bilininteg_diffusion_kernels.cpp:262-386;
 *                fem/integ/bilininteg_mass_pa.cpp:25-90) ------------------ */
/* W[Q^3] tensor weights x fastest (fem/intrules.cpp:68-96). */
void orc_tensor_weights(int Q, double *W);
/* C has size 1 (constant) or Q^3*ne.  Dd[Q^3][6][ne], comps 11,21,31,22,32,33. */
void orc_diffusion_setup(int Q, int ne, const double *W, const double *J,
                         const double *C, int c_size, double *Dd);
/* Mv[Q^3][ne] = W * k * detJ. */
void orc_mass_setup(int Q, int ne, const double *W, const double *J,
                    const double *C, int c_size, double *Mv);

/* ---- PA apply, E->E accumulate
 * (fem/integ/bilininteg_diffusion_kernels.hpp:989-1221,
 *  fem/integ/bilininteg_mass_kernels.hpp:815-1049) ------------------------ */
void orc_diffusion_apply(int D1, int Q1, int ne, const double *B, const double *G,
                         const double *Dd, const double *x, double *y);
void orc_mass_apply(int D1, int Q1, int ne, const double *B, const double *Mv,
                    const double *x, double *y);

/* ---- PA diagonal, E accumulate
 * (fem/integ/bilininteg_diffusion_kernels.hpp:268-367,
 *  fem/integ/bilininteg_mass_kernels.hpp:261-419) ------------------------- */
void orc_diffusion_diag(int D1, int Q1, int ne, const double *B, const double *G,
                        const double *Dd, double *y);
void orc_mass_diag(int D1, int Q1, int ne, const double *B, const double *Mv,
                   double *y);

/* ---- L-vector operator + constrained Jacobi-PCG
 * (fem/bilinearform_ext.cpp:370-454, :487-677; linalg/operator.cpp:129-145,
 *  :526-662; linalg/solvers.cpp:324-453, :869-1050) ----------------------- */
typedef struct {
    int D1, Q1, ne, ndofs;
    const int *gather_map, *offsets, *indices;
    const double *B, *G;
    const double *Dd;   /* diffusion PA data or NULL */
    const double *Mv;   /* mass PA data or NULL      */
    int n_ess;
    const int *ess;     /* essential (constrained) DOFs */
} orc_pa_op;

/* y = A x, unconstrained L->L PA operator. */
void orc_pa_mult(const orc_pa_op *op, const double *x, double *y);
/* L-vector diagonal (E diag then AbsMultTranspose). */
void orc_pa_diagonal(const orc_pa_op *op, double *diag);
/* ConstrainedOperator::Mult with DIAG_ONE. */
void orc_constrained_mult(const orc_pa_op *op, const double *x, double *y);
/* FormLinearSystem with x on essential DOFs: B = b - A x_ess, B[ess] = x[ess],
 * X[interior] = 0. */
void orc_form_linear_system(const orc_pa_op *op, double *x, const double *b,
                            double *X, double *Bv);
/* Jacobi-preconditioned CG with CGSolver semantics; dinv from the diagonal.
 * Returns converged flag; iters, final_norm (sqrt of (Br,r)), initial norm. */
int  orc_pcg(const orc_pa_op *op, const double *diag, const double *b, double *x,
             double rel_tol, double abs_tol, int max_iter, int *iters,
             double *final_norm, double *initial_norm);

/* ---- Chebyshev smoother (linalg/solvers.cpp:455-656) and power method
 *      (linalg/operator.cpp:886-935) ------------------------------------- */
/* srand(seed); n x rand() of glibc's default generator. */
void orc_glibc_rand(unsigned seed, int n, int *out);
/* Vector::Randomize(seed) (linalg/vector.cpp:955-966). */
void orc_randomize(unsigned seed, int n, double *v);
/* Coefficients of OperatorChebyshevSmoother::Setup; returns 0 for order > 5. */
int  orc_chebyshev_coeffs(int order, double max_eig, double *c);
/* Largest eigenvalue estimate of diag^-1 A (constrained), v0 = start vector. */
double orc_power_method(const orc_pa_op *op, const double *diag, double *v0, int steps,
                        double tol);
/* OperatorChebyshevSmoother::Mult. */
void orc_chebyshev_mult(const orc_pa_op *op, const double *diag, int order, const double *c,
                        const double *x, double *y);
/* CGSolver preconditioned by the Chebyshev smoother; as orc_pcg. */
int  orc_pcg_chebyshev(const orc_pa_op *op, const double *diag, int order, const double *c,
                       const double *b, double *x, double rel_tol, double abs_tol, int max_iter,
                       int *iters, double *final_norm, double *initial_norm);

/* ---- p-refinement transfer between two H1 spaces on one mesh
 *      (TensorProductPRefinementTransferOperator, fem/transfer.cpp:2619-2990) -- */
/* B[qh + (ph+1) dl] = l_dl(xi_qh): the order-pl GLL Lagrange basis at the
 * order-ph GLL nodes (lexicographic), transfer.cpp:2641-2660. */
void orc_transfer_basis(int pl, int ph, double *B);
/* Mult (transfer.cpp:2941-2965): y_h = R_h^T (mask_h o (B x B x B) R_l x_l);
 * mask_h is ElementRestriction::BooleanMask of R_h (restriction.cpp:255-290). */
void orc_transfer_prolong(int ne, int pl, int ph, const int *gmap_l, const int *offsets_h,
                          const int *indices_h, int ndofs_h, const double *mask_h,
                          const double *xl, double *yh);
/* MultTranspose (transfer.cpp:2967-2990): y_l = R_l^T (Bt x Bt x Bt)(mask_h o R_h x_h). */
void orc_transfer_restrict(int ne, int pl, int ph, const int *gmap_h, const int *offsets_l,
                           const int *indices_l, int ndofs_l, const double *mask_h,
                           const double *xh, double *yl);

/* ---- RHS: DomainLFIntegrator (fem/lininteg.cpp:51-79) ------------------ */
/* b_E[D^3][ne] += sum_q w_q detJ f(x_q) phi(x_q), Q = p+1 GL points,
 * constant f. */
void orc_domain_lf(int p, int ne, int pg, const double *X, double f, double *bE);
/* The same with f given at the (p+1)^3 points of every element, fq[q + nq e]
 * (f used when fq is NULL). */
void orc_domain_lf_qf(int p, int ne, int pg, const double *X, double f, const double *fq,
                      double *bE);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int  orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
