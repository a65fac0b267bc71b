/*
 * h1pa.h -- C ABI of the MI355X-native H1 partial-assembly (PA) operator path.
 *
 * Drop-in boundary for MFEM's matrix-free H1 path on hexahedra:
 *   BilinearForm::SetAssemblyLevel(AssemblyLevel::PARTIAL)  (fem/bilinearform.cpp:109-136)
 *   BilinearForm::Assemble / Operator::Mult / AssembleDiagonal
 *                                                 (fem/bilinearform_ext.cpp:332-454, :487-677)
 *   ElementRestriction                            (fem/restriction.hpp:40-108, restriction.cpp:26-290)
 *   QuadratureInterpolator (geometric factors)    (fem/quadinterpolator.cpp:226-341)
 *   ConstrainedOperator + OperatorJacobiSmoother + CGSolver
 *                                                 (linalg/operator.cpp:129-145, :526-662;
 *                                                  linalg/solvers.cpp:324-453, :869-1050)
 *
 * Conventions
 *   - Every function returns an int status (H1PA_OK = 0); h1pa_last_error()
 *     returns a thread-local message for the last failure.  No exception
 *     crosses this boundary (MFEM: MFEM_VERIFY/MFEM_ABORT, general/error.hpp:104-166).
 *   - Vector arguments named d_* are DEVICE pointers (FP64 unless stated);
 *     arguments named h_* are HOST pointers.  Layouts follow MFEM: L-vectors
 *     are indexed by global DOF; E-vectors are [dof][ne] with dof =
 *     (p+1)^3 in lexicographic order (ElementDofOrdering::LEXICOGRAPHIC).
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *     Calls on one handle are not re-entrant; distinct handles may run on
 *     distinct streams.
 */
#ifndef H1PA_H
#define H1PA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
    H1PA_OK = 0,
    H1PA_EINVAL = 1,    /* bad argument                          */
    H1PA_ENOMEM = 2,    /* host or device allocation failed      */
    H1PA_EHIP = 3,      /* HIP runtime error                     */
    H1PA_ENOTCONV = 4,  /* solver did not converge               */
    H1PA_EUNSUP = 5,    /* unsupported order / shape             */
    H1PA_ENONFINITE = 6 /* a solver's (B r, r) or (A d, d) was not finite
                           (CGSolver's MFEM_VERIFY(IsFinite(...)) aborts,
                           linalg/solvers.cpp:895, :935, :970, :1011)     */
};

typedef struct h1pa_mesh h1pa_mesh;
typedef struct h1pa_space h1pa_space;
typedef struct h1pa_restriction h1pa_restriction;
typedef struct h1pa_form h1pa_form;

/* Last error message of the calling thread ("" if none). */
const char *h1pa_last_error(void);
/* Library version string. */
const char *h1pa_version(void);
/* Select the HIP device for subsequent handle creation (Device::Configure
 * with "hip:<id>", general/device.cpp:191, which calls hipSetDevice at :460). */
int h1pa_set_device(int device);
/* Block until all work the library queued on `stream` is done
 * (MFEM_STREAM_SYNC / MFEM_DEVICE_SYNC, general/backends.hpp:48-51). */
int h1pa_stream_synchronize(void *stream);

/* ---- mesh: Mesh::MakeCartesian3D(nx, ny, nz, HEXAHEDRON, 1, 1, 1, sfc)
 *      (mesh/mesh.cpp:3841-4121, :4786-4794) -------------------------------- */
int h1pa_mesh_create_cartesian(int nx, int ny, int nz, int sfc_ordering, h1pa_mesh **mesh);
/* Mesh(istream) / Mesh::ReadMFEMMesh (mesh/mesh_readers.cpp:40-130) on the text
 * of an "MFEM mesh v1.0" file (len bytes, not NUL-terminated): a conforming
 * mesh of hexahedra (geometry 5) with a quadrilateral boundary list (geometry
 * 3) and straight vertex coordinates.  Elements keep the file order; edges and
 * faces are numbered as FinalizeTopology numbers them (first encounter); the
 * boundary DOFs of spaces on it are those of the file's boundary faces (all
 * attributes).  A 'nodes' section, other element types, inverted elements and
 * other dimensions give H1PA_EUNSUP. */
int h1pa_mesh_load(const char *text, int64_t len, h1pa_mesh **mesh);
/* Mesh(vertices, nv, element_indices, Geometry::CUBE, attributes, ne,
 * boundary_indices, Geometry::SQUARE, bdr_attributes, nbdr, 3, 3), which ends
 * in FinalizeTopology (mesh/mesh.cpp:4891-4930): h_xyz[nv][3] vertex
 * coordinates, h_ev[ne][8] element vertices in MFEM's hexahedron order
 * (Geometry::CUBE, fem/geom.cpp), h_bdr[nbdr][4] boundary quadrilaterals
 * (each must be an element face).  Elements keep the given order; edges and
 * faces are numbered by first encounter, as FinalizeTopology numbers them;
 * the boundary DOFs of spaces on the mesh are those of the listed faces (MFEM
 * marks every boundary attribute essential in the solves here, so the
 * attributes are not needed).  Host inputs, copied.  Inverted elements give
 * H1PA_EUNSUP. */
int h1pa_mesh_create(int64_t nv, const double *h_xyz, int64_t ne, const int *h_ev, int64_t nbdr,
                     const int *h_bdr, h1pa_mesh **mesh);
/* High-order mesh nodes (Mesh::SetCurvature(order) then writing the Nodes,
 * or Mesh::SetNodalGridFunction, mesh/mesh.cpp:7200-7240): the nodes of an
 * H1_FECollection(order, 3) (Gauss-Lobatto) vector space, given as the
 * lexicographic E-vector d_nodes[(order+1)^3][3][ne] -- what
 * GetElementRestriction(LEXICOGRAPHIC)->Mult(*mesh.GetNodes(), Enodes) gives
 * and what GeometricFactors::Compute interpolates (mesh/mesh.cpp:15445-15497).
 * Copied on `stream`; forms assembled afterwards use them (MFEM: forms
 * assembled after Mesh::DeleteGeometricFactors).  h1pa_mesh_set_curvature
 * or h1pa_mesh_perturb replaces them by nodes computed from the vertices. */
int h1pa_mesh_set_nodes(h1pa_mesh *mesh, int order, const double *d_nodes, void *stream);
int h1pa_mesh_destroy(h1pa_mesh *mesh);
/* Perturb interior vertices by U(-amp*h, amp*h) per coordinate (counter-based
 * splitmix64 with `seed`); boundary vertices stay on the cube faces.  No MFEM
 * function does this; MFEM code moves vertices with Mesh::Transform
 * (mesh/mesh.cpp:14056). */
int h1pa_mesh_perturb(h1pa_mesh *mesh, double amp, uint64_t seed);
/* Give the mesh order-`order` nodes (Mesh::SetCurvature, mesh/mesh.cpp:7211)
 * mapped by x' = x + amp sin(2 pi y) sin(2 pi z) and its cyclic permutations
 * (Mesh::Transform, mesh/mesh.cpp:14056).  amp = 0 keeps the trilinear map. */
int h1pa_mesh_set_curvature(h1pa_mesh *mesh, int order, double amp);
/* counts[0..3] = NV, NE, NEdges, NFaces. */
int h1pa_mesh_counts(const h1pa_mesh *mesh, int64_t *counts);
/* Host copies of the topology (NULL to skip): element vertices [ne][8],
 * edges [ne][12], edge orientations (+-1) [ne][12], faces [ne][6], face
 * orientations [ne][6]  (Mesh::GetElementVertices/Edges/Faces,
 * mesh/mesh.cpp:8044-8073, :8318-8340). */
int h1pa_mesh_topology(const h1pa_mesh *mesh, int *h_ev, int *h_ee, int *h_eeo,
                       int *h_ef, int *h_efo);
/* Host copy of the vertex coordinates [nv][3]. */
int h1pa_mesh_vertices(const h1pa_mesh *mesh, double *h_xyz);

/* ---- H1 space: FiniteElementSpace(mesh, H1_FECollection(p, 3, GaussLobatto))
 *      (fem/fespace.cpp:2769-2900, :3428-3535; fem/fe_coll.cpp:1741-2107) ---- */
int h1pa_space_create(h1pa_mesh *mesh, int order, h1pa_space **space);
int h1pa_space_destroy(h1pa_space *space);
/* info[0..3] = order, NE, dof per element, NDofs (= true VSize). */
int h1pa_space_info(const h1pa_space *space, int64_t *info);
/* Element-to-DOF table in MFEM's native element order
 * (FiniteElementSpace::GetElementToDofTable), device [ne][dof]. */
int h1pa_space_element_dofs(const h1pa_space *space, const int **d_elem_dof);
/* Essential true DOFs for all boundary attributes, ascending
 * (FiniteElementSpace::GetBoundaryTrueDofs, fem/fespace.cpp:668-680). */
int h1pa_space_boundary_dofs(const h1pa_space *space, const int **d_ess, int64_t *n_ess);
/* Borrowed ElementRestriction of the space (LEXICOGRAPHIC;
 * FiniteElementSpace::GetElementRestriction, fem/fespace.cpp:1476-1505). */
int h1pa_space_restriction(h1pa_space *space, h1pa_restriction **restriction);

/* ---- ElementRestriction (fem/restriction.hpp:40-108) ---------------------- */
/* Build from MFEM's own element-to-DOF table and lexicographic dof_map, as the
 * ElementRestriction constructor does (fem/restriction.cpp:26-114).  Signed
 * entries follow MFEM's -1-i convention.  h_dof_map = NULL gives the NATIVE
 * element ordering (no reordering, restriction.cpp:60-98).  Host inputs. */
int h1pa_restriction_create(int ne, int dof, int ndofs, const int *h_elem_dof,
                            const int *h_dof_map, h1pa_restriction **restriction);
/* Releases what h1pa_restriction_create made (~ElementRestriction). */
int h1pa_restriction_destroy(h1pa_restriction *restriction);
/* Device pointers of gather_map [ne*dof], offsets [ndofs+1], indices [ne*dof]
 * (GatherMap/Offsets/Indices, fem/restriction.hpp:105-107). */
int h1pa_restriction_maps(const h1pa_restriction *restriction, const int **d_gather_map,
                          const int **d_offsets, const int **d_indices);
/* L->E: y[i%dof, i/dof] = +-x[gather_map[i]]   (restriction.cpp:116-136) */
int h1pa_restriction_mult(const h1pa_restriction *r, const double *d_x, double *d_y, void *stream);
/* L->E without signs (restriction.cpp:138-157) */
int h1pa_restriction_abs_mult(const h1pa_restriction *r, const double *d_x, double *d_y, void *stream);
/* E->L set, deterministic transpose-table sum (restriction.cpp:159-193) */
int h1pa_restriction_mult_transpose(const h1pa_restriction *r, const double *d_x, double *d_y,
                                    void *stream);
/* E->L add; a must be 1.0 as in MFEM (restriction.cpp:195-201) */
int h1pa_restriction_add_mult_transpose(const h1pa_restriction *r, const double *d_x,
                                        double *d_y, double a, void *stream);
/* E->L set without signs (restriction.cpp:203-228) */
int h1pa_restriction_abs_mult_transpose(const h1pa_restriction *r, const double *d_x,
                                        double *d_y, void *stream);
/* E->L taking the last E-entry of each L-DOF (restriction.cpp:230-253) */
int h1pa_restriction_mult_left_inverse(const h1pa_restriction *r, const double *d_x,
                                       double *d_y, void *stream);

/* E-vector mask (restriction.cpp:255-290): d_ye[first E-entry of each L-DOF
 * in indices order] = 1, every other entry 0 -- the weights that make an E->L
 * sum count each L-DOF once (TensorProductPRefinementTransferOperator,
 * fem/transfer.cpp:2682). */
int h1pa_restriction_boolean_mask(const h1pa_restriction *r, double *d_ye, void *stream);

/* ---- quadrature interpolation of the mesh nodes (geometric factors)
 *      GeometricFactors::Compute (mesh/mesh.cpp:15445-15497) via
 *      QuadratureInterpolator::Mult (fem/quadinterpolator.cpp:226-341) ------- */
enum {
    H1PA_QI_VALUES = 1,               /* QuadratureInterpolator::VALUES               */
    H1PA_QI_DERIVATIVES = 2,          /* ::DERIVATIVES (reference coordinates)        */
    H1PA_QI_DETERMINANTS = 4,         /* ::DETERMINANTS                               */
    H1PA_QI_PHYSICAL_DERIVATIVES = 8  /* ::PHYSICAL_DERIVATIVES (J^-T applied)        */
};
enum { H1PA_QI_BY_NODES = 0, H1PA_QI_BY_VDIM = 1 };  /* QVectorLayout */
/* Evaluate the mesh nodes at the q1d^3 Gauss-Legendre points: X[nq][3][ne],
 * J[nq][3][3][ne] (J(q,r,c,e) = dx_r/dxi_c), detJ[nq][ne]; NULL skips
 * (Mesh::GetGeometricFactors, mesh/mesh.cpp:958; GeometricFactors::Compute,
 * :15445-15497). */
int h1pa_mesh_geometric_factors(const h1pa_mesh *mesh, int q1d, unsigned flags,
                                double *d_X, double *d_J, double *d_detJ, void *stream);

/* QuadratureInterpolator::Mult (fem/quadinterpolator.cpp:226-341, tensor
 * kernels fem/qinterp/eval.hpp, grad.hpp:233-375, det.hpp) for an E-vector of
 * the space: d_e is lexicographic [D^3][vdim][ne] (the ElementRestriction
 * output of a vdim-component field), evaluated at the q1d^3 Gauss-Legendre
 * points, q1d in [p+1, min(p+3, 10)].  Outputs (NULL where not requested), nq = q1d^3:
 *   VALUES       byNODES val[nq][vdim][ne],       byVDIM val[vdim][nq][ne]
 *   DERIVATIVES or PHYSICAL_DERIVATIVES (not both)
 *                byNODES der[nq][vdim][3][ne],    byVDIM der[vdim][3][nq][ne]
 *   DETERMINANTS (vdim = 3) det[nq][ne] of the reference derivative matrix.
 * Physical derivatives are J^-T times the reference ones, J from the mesh's
 * geometric factors at the same points. */
int h1pa_space_qinterp(const h1pa_space *space, int q1d, int vdim, unsigned flags, int layout,
                       const double *d_e, double *d_val, double *d_der, double *d_det,
                       void *stream);

/* ---- PA bilinear form ---------------------------------------------------- */
enum { H1PA_INTEG_MASS = 1, H1PA_INTEG_DIFFUSION = 2 };
enum {
    H1PA_COEFF_CONSTANT = 0,    /* ConstantCoefficient(value)                         */
    H1PA_COEFF_SINCOS = 1,      /* 1 + 0.5 sin(2 pi x) cos(2 pi y) sin(2 pi z)        */
    H1PA_COEFF_PER_ELEMENT = 2, /* PWConstCoefficient-like: d_data[ne] device values  */
    H1PA_COEFF_QUADRATURE = 3,  /* QuadratureFunction: d_data[q1d^3][ne] device       */
    H1PA_COEFF_FOURIER = 4      /* exp(sum of 8 seeded Fourier modes, amplitude 0.5)  */
};
typedef struct {
    unsigned integrators;   /* H1PA_INTEG_* bitmask                         */
    int coeff_kind;         /* H1PA_COEFF_*                                  */
    double coeff_value;     /* constant value                                */
    const double *d_coeff;  /* device data for PER_ELEMENT / QUADRATURE      */
    uint64_t coeff_seed;    /* seed for FOURIER                              */
    int q1d;                /* 0 -> MFEM default p+2                          */
} h1pa_form_desc;

/* BilinearForm a(fes); a.SetAssemblyLevel(PARTIAL); a.AddDomainIntegrator(...) */
int h1pa_form_create(h1pa_space *space, const h1pa_form_desc *desc, h1pa_form **form);
int h1pa_form_destroy(h1pa_form *form);
/* a.Assemble(): geometric factors, coefficient and PA data on the device
 * (DiffusionIntegrator::AssemblePA, fem/integ/bilininteg_diffusion_pa.cpp:108-171;
 *  MassIntegrator::AssemblePA, fem/integ/bilininteg_mass_pa.cpp:25-90). */
int h1pa_form_assemble(h1pa_form *form, void *stream);
/* Operator::Mult: y = A x on L-vectors, y overwritten
 * (PABilinearFormExtension::Mult, fem/bilinearform_ext.cpp:487-560). */
int h1pa_form_mult(h1pa_form *form, const double *d_x, double *d_y, void *stream);
/* Operator::AddMult: y += a A x (PABilinearFormExtension inherits
 * Operator::AddMult, linalg/operator.cpp:51-56: A x into a form-owned
 * L-vector, then y += a z).  The mass and diffusion operators here are
 * symmetric, so MultTranspose / AddMultTranspose are h1pa_form_mult /
 * h1pa_form_add_mult (DiffusionIntegrator::AddMultTransposePA,
 * fem/integ/bilininteg_diffusion_pa.cpp:95-106; MassIntegrator's,
 * bilininteg_mass_pa.cpp:230-234). */
int h1pa_form_add_mult(h1pa_form *form, const double *d_x, double *d_y, double a, void *stream);
/* AddMultPA on E-vectors: yE += A_E xE (fem/bilininteg.hpp:74-81). */
int h1pa_form_add_mult_e(h1pa_form *form, const double *d_xE, double *d_yE, void *stream);
/* Operator::AssembleDiagonal: L-vector diagonal, set
 * (PABilinearFormExtension::AssembleDiagonal, fem/bilinearform_ext.cpp:370-454). */
int h1pa_form_assemble_diagonal(h1pa_form *form, double *d_diag, void *stream);
/* Device pointers of the PA data: diffusion D[q1d^3][6][ne] (comps 11,21,31,
 * 22,32,33) and mass v[q1d^3][ne]; NULL where the integrator is absent. */
int h1pa_form_pa_data(const h1pa_form *form, const double **d_diffusion, const double **d_mass);
/* info[0..3] = q1d, integrators, ndofs, selected diffusion-apply variant (-1: none) */
int h1pa_form_info(const h1pa_form *form, int64_t *info);
/* Which kernels h1pa_form_mult runs (no MFEM counterpart; MFEM's Mult is
 * gather, element apply, scatter, fem/bilinearform_ext.cpp:487-560):
 * *path = 1 the fused column apply (a diffusion form on a Cartesian mesh at
 * q1d = p + 2, opt-in with H1PA_COLUMNS=1 in the environment: element results
 * summed in LDS, no E-vector), 0 the element kernels followed by the E->L
 * scatter (the default).  Sets the column apply up on first
 * use, on `stream`. */
int h1pa_form_mult_path(h1pa_form *form, int *path, void *stream);
/* No counterpart in MFEM.  For callers that cannot allocate device memory
 * during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): y = A x as h1pa_form_mult, through the
 * element kernels and the E->L scatter, with the E-vector in a static device
 * array of the library (64 MiB per GPU, held until the process ends) instead
 * of form-owned memory; no allocation of any kind.  Forms whose E-vector has
 * more than 8388608 entries (ne (p+1)^3) give H1PA_EINVAL.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (other streams or
 * threads) race and must be serialized by the caller.  The form must be
 * assembled.  Same operator as h1pa_form_mult to rounding (1e-13 relative). */
int h1pa_form_mult_noalloc(h1pa_form *form, const double *d_x, double *d_y, void *stream);
/* Benchmark helper (no MFEM counterpart): `steps` back-to-back h1pa_form_mult
 * calls with HIP events recorded on `stream` around each launch.  Returns the
 * wall time of the whole region and the average duration of the element
 * kernel(s) (the column sweep on the fused path) and of the E->L scatter (the
 * interface pass), in milliseconds. */
int h1pa_form_mult_timed(h1pa_form *form, const double *d_x, double *d_y, int steps,
                         double *ms_total, double *ms_element, double *ms_scatter, void *stream);

/* ---- essential BCs and Jacobi-PCG --------------------------------------- */
/* ConstrainedOperator (DIAG_ONE) Mult with the ess list (linalg/operator.cpp:601-662). */
int h1pa_form_constrained_mult(h1pa_form *form, const int *d_ess, int64_t n_ess,
                               const double *d_x, double *d_y, void *stream);
/* Operator::FormLinearSystem with copy_interior = 0 (linalg/operator.cpp:129-145,
 * :574-599): X = x on ess, 0 elsewhere; B = b - A X_ess; B[ess] = X[ess]. */
int h1pa_form_linear_system(h1pa_form *form, const int *d_ess, int64_t n_ess,
                            const double *d_x, const double *d_b, double *d_X, double *d_B,
                            void *stream);
/* Vector::Add (linalg/vector.cpp:326): d_y += a d_x on n device entries. */
int h1pa_vector_add(double *d_y, double a, const double *d_x, int64_t n, void *stream);
/* Vector::Randomize (linalg/vector.cpp:955-966) on the device with a
 * counter-based generator: d_x[i] = lo + (hi - lo) u_i, u_i =
 * (splitmix64(seed ^ splitmix64(i)) >> 11) 2^-53, reproducible for any n. */
int h1pa_vector_randomize(double *d_x, int64_t n, uint64_t seed, double lo, double hi,
                          void *stream);
/* DomainLFIntegrator(ConstantCoefficient(f)) assembled into an L-vector with
 * q1d = p+1 (fem/lininteg.cpp:51-79, LinearForm::Assemble). */
int h1pa_space_domain_lf(h1pa_space *space, double f, double *d_b, void *stream);
/* DomainLFIntegrator(QuadratureFunctionCoefficient) on the device path
 * (DomainLFIntegrator::AssembleDevice, fem/integ/lininteg_domain.cpp:48-59:
 * default rule of order 2p, CoefficientVector::Project referencing the
 * QuadratureFunction, fem/coefficient.cpp:2058-2075): the same with f given at
 * the (p+1)^3 Gauss-Legendre points of every element, d_fq[q + (p+1)^3 e],
 * points x fastest -- the points of h1pa_mesh_geometric_factors at
 * q1d = p + 1, whose X[q + nq (c + 3 e)] gives their coordinates, so a field
 * evaluated there can be passed; d_fq = NULL uses the constant f. */
int h1pa_space_domain_lf_qf(h1pa_space *space, double f, const double *d_fq, double *d_b,
                            void *stream);

typedef struct {
    double rel_tol;     /* CGSolver::SetRelTol (preconditioned-norm criterion) */
    double abs_tol;     /* CGSolver::SetAbsTol                                 */
    int max_iter;       /* CGSolver::SetMaxIter                                */
    int check_every;    /* iterations per host convergence check (0 -> 8)      */
} h1pa_pcg_opts;
typedef struct {
    int converged;
    int iterations;       /* CGSolver::GetNumIterations                    */
    double initial_norm;  /* sqrt((B r0, r0))                              */
    double final_norm;    /* sqrt((B r, r)), CGSolver::GetFinalNorm        */
    double seconds;       /* wall time of the iteration loop               */
} h1pa_pcg_stats;
/* OperatorJacobiSmoother(a, ess) (linalg/solvers.cpp:401-453) + CGSolver
 * (:869-1050) on the DIAG_ONE ConstrainedOperator (linalg/operator.cpp:601-662),
 * iterative_mode = false (x is overwritten; initial guess 0). */
int h1pa_pcg_solve(h1pa_form *form, const int *d_ess, int64_t n_ess, const double *d_b,
                   double *d_x, const h1pa_pcg_opts *opts, h1pa_pcg_stats *stats,
                   void *stream);

/* ---- Chebyshev smoother ---------------------------------------------------
 * OperatorChebyshevSmoother (linalg/solvers.hpp:502-585, solvers.cpp:455-656)
 * on the DIAG_ONE constrained operator of a form, with the Jacobi diagonal
 * of AssembleDiagonal (ess rows 1), and PowerMethod::EstimateLargestEigenvalue
 * (linalg/operator.cpp:886-935) for its eigenvalue bound. */
typedef struct h1pa_chebyshev h1pa_chebyshev;
/* d_diag: the L-vector diagonal (MFEM's `d`; NULL -> AssembleDiagonal of the
 * form), read during the call.  order 1..5.  max_eig_estimate > 0 is used as
 * given (the constructor taking max_eig_estimate); otherwise the largest
 * eigenvalue of diag^-1 A is estimated by power_iterations power steps
 * (relative tolerance power_tolerance) from Vector::Randomize(power_seed) --
 * the C library's srand/rand sequence, as MFEM draws it; MFEM's defaults are
 * 10, 1e-8, 12345.  The essential list is copied; the form must outlive the
 * smoother. */
int h1pa_chebyshev_create(h1pa_form *form, const double *d_diag, const int *d_ess,
                          int64_t n_ess, int order, double max_eig_estimate,
                          int power_iterations, double power_tolerance, int power_seed,
                          h1pa_chebyshev **cheb, void *stream);
/* ~OperatorChebyshevSmoother. */
int h1pa_chebyshev_destroy(h1pa_chebyshev *cheb);
/* The eigenvalue estimate and the order polynomial coefficients (coeffs has
 * room for order values; either pointer may be NULL) -- the values the
 * constructor computes, linalg/solvers.cpp:546-617. */
int h1pa_chebyshev_info(const h1pa_chebyshev *cheb, double *max_eig, double *coeffs);
/* OperatorChebyshevSmoother::Mult (linalg/solvers.cpp:619-656):
 * y = sum_k c_k (diag^-1 A)^k diag^-1 x. */
int h1pa_chebyshev_mult(h1pa_chebyshev *cheb, const double *d_x, double *d_y, void *stream);
/* CGSolver with SetPreconditioner(smoother) on the smoother's constrained
 * operator (linalg/solvers.cpp:869-1050), iterative_mode = false; as
 * h1pa_pcg_solve otherwise. */
int h1pa_pcg_solve_chebyshev(h1pa_chebyshev *cheb, const double *d_b, double *d_x,
                             const h1pa_pcg_opts *opts, h1pa_pcg_stats *stats, void *stream);

/* ---- p-refinement transfer and p-multigrid ---------------------------------
 * TensorProductPRefinementTransferOperator (fem/transfer.cpp:2619-2990)
 * between two H1 spaces of one mesh (low order below high order). */
typedef struct h1pa_transfer h1pa_transfer;
int h1pa_transfer_create(h1pa_space *low, h1pa_space *high, h1pa_transfer **transfer);
int h1pa_transfer_destroy(h1pa_transfer *transfer);
/* Mult (transfer.cpp:2941-2965): d_yh = R_h^T (mask_h o (B x B x B) R_l d_xl),
 * B the low-order basis at the high-order nodes, mask_h the BooleanMask of
 * the high restriction: every high DOF gets the interpolated value of the
 * first element holding it.  d_xl: low L-vector, d_yh: high L-vector (set). */
int h1pa_transfer_mult(h1pa_transfer *transfer, const double *d_xl, double *d_yh, void *stream);
/* MultTranspose (transfer.cpp:2967-2990): d_yl = R_l^T (Bt x Bt x Bt)(mask_h o R_h d_xh). */
int h1pa_transfer_mult_transpose(h1pa_transfer *transfer, const double *d_xh, double *d_yl,
                                 void *stream);

/* Multigrid with order-refined levels of the fine form's mesh: the
 * GeometricMultigrid of a FiniteElementSpaceHierarchy built with
 * AddOrderRefinedLevel (fem/multigrid.hpp:135-200,
 * fem/fespacehierarchy.cpp:86-99), levels set up as ex26's
 * DiffusionMultigrid (examples/ex26.cpp:42-110): a PA form with the fine
 * form's integrators and coefficient on every level (default rule
 * q1d = p + 2), the DIAG_ONE constrained operator of each level's boundary
 * DOFs, OperatorChebyshevSmoother on levels >= 1 and CGSolver without a
 * preconditioner (abs_tol 0) on the coarsest. */
typedef struct h1pa_multigrid h1pa_multigrid;
typedef struct {
    int smoother_order;      /* OperatorChebyshevSmoother order, 1..5 (ex26: 2)      */
    int pre_smoothing;       /* SetCycleType(type, pre, post) (ex26: 1, 1)           */
    int post_smoothing;
    int w_cycle;             /* 0: CycleType::VCYCLE, 1: CycleType::WCYCLE           */
    double coarse_rel_tol;   /* coarse CGSolver::SetRelTol (ex26: sqrt(1e-4))        */
    int coarse_max_iter;     /* coarse CGSolver::SetMaxIter (ex26: 200)              */
    int power_iterations;    /* smoothers' eigenvalue estimate (MFEM: 10, 1e-8, 12345) */
    double power_tolerance;
    int power_seed;
    int cycle_precision;     /* H1PA_PREC_FP64, or H1PA_PREC_BF16_PA: the cycle's
                                operator applies (smoothing, residuals, coarse CG)
                                read BF16 copies of each level's PA data with FP64
                                vectors and arithmetic -- a mixed-precision
                                preconditioner inside the FP64 CG (no MFEM
                                counterpart; diffusion forms at q1d = p + 2)    */
    int coarse_jacobi;       /* 0: the coarse CG has no preconditioner (ex26);
                                1: OperatorJacobiSmoother on the coarse level    */
} h1pa_mg_opts;
/* orders[0 .. nlevels-1]: increasing, orders[nlevels-1] = the fine space's
 * order (2 <= nlevels <= 8).  The fine form must be assembled; the hierarchy
 * holds it until h1pa_multigrid_destroy (meanwhile h1pa_form_assemble,
 * h1pa_form_destroy and a reduced-precision shadow of another kind on it
 * give H1PA_EINVAL, as for a form in a live batch solve).  A
 * QuadratureFunction coefficient (tied to the fine rule) gives H1PA_EUNSUP.
 * The coarser spaces, forms, smoothers and transfers are made and assembled
 * here, on `stream`. */
int h1pa_multigrid_create(h1pa_form *fine, int nlevels, const int *orders,
                          const h1pa_mg_opts *opts, h1pa_multigrid **mg, void *stream);
int h1pa_multigrid_destroy(h1pa_multigrid *mg);
/* MultigridBase::Mult (fem/multigrid.cpp:106-145): d_y = one cycle applied to
 * d_x (fine L-vectors), starting from y = 0. */
int h1pa_multigrid_mult(h1pa_multigrid *mg, const double *d_x, double *d_y, void *stream);
/* info[0..3] = levels, cycles run, coarse solves, coarse CG iterations (sum);
 * max_eig (may be NULL): [nlevels] the smoothers' eigenvalue estimates (0 on
 * the coarsest level). */
int h1pa_multigrid_info(const h1pa_multigrid *mg, int64_t *info, double *max_eig);
/* CGSolver with SetPreconditioner(multigrid) (linalg/solvers.cpp:869-1050)
 * on the fine form's DIAG_ONE constrained operator of its boundary DOFs,
 * iterative_mode = false; as h1pa_pcg_solve otherwise. */
int h1pa_pcg_solve_multigrid(h1pa_multigrid *mg, const double *d_b, double *d_x,
                             const h1pa_pcg_opts *opts, h1pa_pcg_stats *stats, void *stream);

/* ---- mixed precision ------------------------------------------------------
 * No MFEM counterpart: MFEM's precision is one build-wide real_t
 * (config/config.hpp:41-48) and it has no mixed-precision PA or PCG.  The
 * operator data of the reduced-precision applies are shadows of the FP64 PA
 * data made by h1pa_form_assemble's output (FP64 -> FP32 round-to-nearest-even,
 * BF16 from that FP32 value, round-to-nearest-even). */
enum {
    H1PA_PREC_FP64 = 0,     /* FP64 PA data, vectors and arithmetic                  */
    H1PA_PREC_FP32 = 1,     /* FP32 PA data, basis, vectors and arithmetic           */
    H1PA_PREC_BF16_PA = 2   /* BF16-stored PA data, FP32 basis/vectors/arithmetic    */
};
/* y = A x in reduced precision (precision FP32 or BF16_PA): FP32 L-vectors
 * on the device, same operator as h1pa_form_mult. */
int h1pa_form_mult_f32(h1pa_form *form, int precision, const float *d_x, float *d_y,
                       void *stream);
typedef struct {
    int inner_precision;   /* H1PA_PREC_FP32 or H1PA_PREC_BF16_PA                    */
    double rel_tol;        /* stop: (B r, r) <= rel_tol^2 (B b, b) on the FP64 true
                              residual r = b - A x, B = Jacobi (CGSolver's measure)  */
    double delta;          /* refinement trigger: a step is taken whenever the inner
                              residual's B-norm has fallen by this factor (0 < delta < 1) */
    int max_outer;         /* refinement steps                                       */
    int max_inner;         /* inner iterations in total                              */
} h1pa_ir_opts;
typedef struct {
    int converged;
    int outer_iterations;  /* refinement steps taken                                 */
    int inner_iterations;  /* inner CG iterations                                    */
    double final_norm;     /* sqrt((B r, r)) of the FP64 true residual               */
    double rel_residual;   /* ||b - A x||_2 / ||b||_2 with the FP64 operator          */
    double seconds;
} h1pa_ir_stats;
/* Mixed-precision solve of the DIAG_ONE ConstrainedOperator system of the ess
 * list: an inner Jacobi-CG with FP32 vectors and the FP32 or BF16-stored PA
 * operator, wrapped in FP64 iterative refinement by reliable updates: at each
 * refinement step x += xs (FP64), r = b - A x (FP64 operator) and the inner
 * residual is replaced by (float) r while the search direction is kept.
 * d_b and d_x are FP64 L-vectors; b's ess entries are the values x takes
 * there (as after h1pa_form_linear_system). */
int h1pa_ir_solve(h1pa_form *form, const int *d_ess, int64_t n_ess, const double *d_b,
                  double *d_x, const h1pa_ir_opts *opts, h1pa_ir_stats *stats, void *stream);

/* ---- batched solves -------------------------------------------------------
 * No MFEM counterpart (MFEM runs one CGSolver::Mult at a time,
 * linalg/solvers.cpp:869-1050).  A batch advances many independent
 * Jacobi-PCG (or mixed-precision refinement) solves together: one replay of
 * a captured HIP graph runs `iters_per_replay` iterations of every running
 * solve, with one element-kernel launch per (p, q1d) group over all solves'
 * elements.  Each solve's arithmetic is independent of the others, so a
 * solve gives bitwise the same result alone and in any mix.  The batch owns
 * its own stream; it is not re-entrant.  Forms must be diffusion-only with
 * q1d = p + 2; each form, d_b and d_x must stay valid until the solve has
 * been returned by h1pa_batch_poll. */
typedef struct h1pa_batch h1pa_batch;
typedef struct {
    int precision;         /* H1PA_PREC_FP64: plain PCG (CGSolver semantics);
                              H1PA_PREC_FP32 / BF16_PA: as h1pa_ir_solve       */
    int iters_per_replay;  /* iterations per graph replay (1..256)              */
    int use_graph;         /* 1: capture and replay a HIP graph; 0: plain launches */
    int autotune;          /* 1: time the element-kernel variants per group      */
    double delta;          /* mixed: refinement trigger (as h1pa_ir_opts.delta)  */
    int max_outer;         /* mixed: refinement steps                            */
    int max_inner;         /* mixed: inner iterations in total                   */
} h1pa_batch_opts;
typedef struct {
    int converged;
    int iterations;        /* PCG iterations (mixed: inner iterations, summed)  */
    int outer_iterations;  /* mixed: refinement steps; FP64: 0                  */
    double final_norm;     /* FP64: sqrt((B r, r)); mixed: ||b - A x||_2        */
    double rel_residual;   /* mixed: ||b - A x||_2 / ||b||_2; FP64: -1          */
    int status;            /* H1PA_OK (converged), H1PA_ENOTCONV (max_iter, or
                              (B r, r) < 0 or (A d, d) = 0 as CGSolver stops),
                              H1PA_ENONFINITE (a non-finite product: the solve
                              is retired, the others go on)                  */
} h1pa_batch_result;
/* No MFEM counterpart (MFEM runs one CGSolver::Mult at a time). */
int h1pa_batch_create(const h1pa_batch_opts *opts, h1pa_batch **batch);
int h1pa_batch_destroy(h1pa_batch *batch);
/* Add the solve of the DIAG_ONE constrained system (ess list) A x = b.
 * opts->rel_tol / abs_tol / max_iter are the CGSolver settings in FP64 mode;
 * in mixed mode opts->rel_tol is h1pa_ir_opts.rel_tol (abs_tol, max_iter
 * unused: the batch's max_outer / max_inner apply).
 * `tag` is returned by h1pa_batch_poll; *slot receives the solve's index
 * among those added since the last launch.  The solve's setup (Jacobi
 * diagonal, reduced-precision data) runs on `stream`, after the work already
 * queued there (the form's assembly, d_b); the batch waits for it. */
int h1pa_batch_add(h1pa_batch *batch, h1pa_form *form, const int *d_ess, int64_t n_ess,
                   const double *d_b, double *d_x, const h1pa_pcg_opts *opts, int64_t tag,
                   int *slot, void *stream);
/* Run up to max_replays replays, stopping early when no solve is running;
 * *n_running receives the number of solves still running. */
int h1pa_batch_run(h1pa_batch *batch, int max_replays, int *n_running);
/* Asynchronous form of one replay: launch returns at once (*launched = 0 when
 * nothing runs); h1pa_batch_add may be called while it is in flight (the new
 * solve joins at the next launch); wait blocks until the replay is done. */
int h1pa_batch_launch(h1pa_batch *batch, int *launched);
int h1pa_batch_wait(h1pa_batch *batch, int *n_running);
/* *idle = 1 when no replay is in flight or the one in flight has finished,
 * so that h1pa_batch_wait would return at once; never blocks.  A serving loop
 * polls it between admissions to start the next replay without delay (no
 * MFEM counterpart: CGSolver::Mult is synchronous). */
int h1pa_batch_query(h1pa_batch *batch, int *idle);
/* Return (and retire) up to `cap` finished solves: tags[i], results[i]. */
int h1pa_batch_poll(h1pa_batch *batch, int cap, int64_t *tags, h1pa_batch_result *results,
                    int *n);
/* info[0..5] = solves added and not yet polled, solves running, device bytes
 * held by the batch's vectors, replays, graph captures, (p, q1d) groups. */
int h1pa_batch_info(const h1pa_batch *batch, int64_t *info);
/* Device bytes held by a form, its space and its mesh (admission control; no
 * MFEM counterpart). */
int h1pa_form_device_bytes(const h1pa_form *form, int64_t *bytes);
/* Block cache for the library's device memory (MFEM's pooled device memory
 * types, general/mem_manager.hpp:52-55, DEVICE_UMPIRE, play this part there).
 * max_bytes > 0: freed blocks, up to max_bytes per device, are kept and
 * reused by later allocations of a similar size instead of going back through
 * hipFree, which synchronizes the whole device -- for a caller that releases
 * objects while other work runs on other streams.  The caller must then not
 * destroy an object while device work on its buffers is still pending (with
 * the cache off, hipFree waits for that work).  A lower max_bytes frees kept
 * blocks down to it; max_bytes = 0 (the default) turns the cache off and frees
 * what it kept.  *bytes: the bytes kept now.  (Temporaries inside a call are
 * stream-ordered, hipMallocAsync / hipFreeAsync on the call's stream, from
 * the device's default memory pool, whose release threshold the library
 * raises to its maximum so that freed blocks are reused.) */
int h1pa_set_device_cache(int64_t max_bytes);
int h1pa_device_cache_bytes(int64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif /* H1PA_H */
