// multigrid.cpp -- p-multigrid on the PA operators: order-refined levels of
// one mesh, Chebyshev smoothing, a CG coarse solve, and CG preconditioned by
// the V-cycle.
//
// MFEM's pieces, restated for this library's device kernels:
//   MultigridBase::Mult / Cycle / SmoothingStep   fem/multigrid.cpp:106-240
//   GeometricMultigrid with FiniteElementSpaceHierarchy::AddOrderRefinedLevel
//     (fem/fespacehierarchy.cpp:86-99), whose prolongations are
//     TensorProductPRefinementTransferOperator (fem/transfer.cpp:2619-2990)
//   the level set-up of examples/ex26.cpp:42-110 (DiffusionMultigrid): PA
//     forms on every level, OperatorChebyshevSmoother on the DIAG_ONE
//     constrained operators of levels >= 1, CGSolver without a
//     preconditioner on the coarsest level
// Every level's constrained DOFs are its space's boundary DOFs (ess_bdr = all
// attributes).  Vectors live on the device; the coarse solve reads its state
// on the host every few iterations, which also orders the cycle's launches.
#include <memory>
#include <vector>

#include "internal.hpp"

namespace h1pa {

namespace {

constexpr int kBlock = 256;
inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__global__ void k_fill(int64_t n, double v, double *__restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

// r = x - r (r holds A y): MFEM's R = X; R += -1 A Y
__global__ void k_sub_from(int64_t n, const double *__restrict__ x, double *__restrict__ r) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) r[i] = x[i] + (-1.0) * r[i];
}

void fill(int64_t n, double v, double *y, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kBlock), 0, s, n, v, y);
    H1PA_HIP(hipGetLastError());
}

void check_rc(int rc) {
    if (rc != H1PA_OK) throw Error{rc, h1pa_last_error()};
}

}  // namespace

// ---- TensorProductPRefinementTransferOperator -------------------------------
Transfer::Transfer(Space &lo_, Space &hi_, hipStream_t s) : lo(&lo_), hi(&hi_) {
    H1PA_CHECK(lo->mesh == hi->mesh, H1PA_EINVAL, "transfer: the spaces must share one mesh");
    H1PA_CHECK(lo->p < hi->p, H1PA_EINVAL, "transfer: the low order must be below the high order");
    const int Dh = hi->p + 1;
    double xh[16], wh[16], Bh[16 * 16], Gh[16 * 16];
    gauss_lobatto(Dh, xh, wh);
    lagrange_basis(lo->p, Dh, xh, Bh, Gh);  // B[qh + Dh dl]
    B.upload(Bh, (size_t)Dh * (lo->p + 1), s);
    const int64_t nEh = hi->ne * hi->dof;
    emask_h.alloc(nEh);
    kern::first_entry_mask(nEh, hi->restr.gather_map.p, hi->restr.offsets.p, hi->restr.indices.p,
                           emask_h.p, s);
    eL.alloc(lo->ne * lo->dof);
    H1PA_HIP(hipStreamSynchronize(s));
}

void Transfer::mult(const double *xl, double *yh, hipStream_t s) {
    kern::transfer_prolong(lo->ne, lo->p, hi->p, B.p, lo->restr.gather_map.p,
                           hi->restr.gather_map.p, emask_h.p, xl, yh, s);
}

void Transfer::mult_transpose(const double *xh, double *yl, hipStream_t s) {
    kern::transfer_restrict_e(lo->ne, lo->p, hi->p, B.p, hi->restr.gather_map.p, emask_h.p, xh,
                              eL.p, s);
    restriction_scatter(lo->restr, true, eL.p, yl, 0, s);
}

// ---- GeometricMultigrid with order-refined levels ---------------------------
Multigrid::~Multigrid() {
    if (fine_ref) --fine_ref->batch_refs;
    P.clear();
    for (auto it = lv.rbegin(); it != lv.rend(); ++it) {
        it->smoother.reset();
        if (it->form_own) (void)h1pa_form_destroy(it->form_own);
        if (it->space_own) (void)h1pa_space_destroy(it->space_own);
    }
}

Multigrid *multigrid_create(Form &fine, int nlevels, const int *orders, const h1pa_mg_opts &o,
                            hipStream_t s) {
    Space &fs = *fine.space;
    H1PA_CHECK(fine.assembled, H1PA_EINVAL, "multigrid: assemble the fine form first");
    H1PA_CHECK(nlevels >= 2 && nlevels <= 8, H1PA_EINVAL, "multigrid: 2..8 levels");
    H1PA_CHECK(orders[nlevels - 1] == fs.p, H1PA_EINVAL,
               "multigrid: the last order must be the fine space's order");
    for (int l = 0; l < nlevels; ++l) {
        H1PA_CHECK(orders[l] >= 1 && orders[l] <= 8, H1PA_EINVAL, "multigrid: orders in 1..8");
        if (l) H1PA_CHECK(orders[l] > orders[l - 1], H1PA_EINVAL, "multigrid: orders must increase");
    }
    H1PA_CHECK(fine.desc.coeff_kind != H1PA_COEFF_QUADRATURE, H1PA_EUNSUP,
               "multigrid: a QuadratureFunction coefficient belongs to the fine rule only");
    H1PA_CHECK(o.smoother_order >= 1 && o.smoother_order <= 5, H1PA_EINVAL,
               "multigrid: smoother order 1..5");
    H1PA_CHECK(o.pre_smoothing >= 0 && o.post_smoothing >= 0 &&
                   o.pre_smoothing + o.post_smoothing >= 1,
               H1PA_EINVAL, "multigrid: smoothing steps");
    H1PA_CHECK(o.coarse_max_iter >= 1 && o.coarse_rel_tol >= 0.0, H1PA_EINVAL,
               "multigrid: coarse solver settings");
    H1PA_CHECK(fs.mesh->handle, H1PA_EINVAL, "multigrid: the mesh has no handle");
    H1PA_CHECK(o.cycle_precision == H1PA_PREC_FP64 || o.cycle_precision == H1PA_PREC_BF16_PA,
               H1PA_EINVAL, "multigrid: cycle precision must be FP64 or BF16_PA");
    if (o.cycle_precision == H1PA_PREC_BF16_PA)
        H1PA_CHECK((fine.desc.integrators & H1PA_INTEG_DIFFUSION) && fine.q1d == fs.p + 2,
                   H1PA_EUNSUP, "multigrid: a BF16 cycle needs a diffusion form at q1d = p + 2");
    std::unique_ptr<Multigrid> mg(new Multigrid);
    mg->o = o;
    mg->lv.resize(nlevels);
    for (int l = 0; l < nlevels; ++l) {
        MgLevel &L = mg->lv[l];
        if (l == nlevels - 1) {
            L.f = &fine;
        } else {
            // BilinearForm on the order-refined space, the fine form's
            // integrators and coefficient, the level's default rule
            check_rc(h1pa_space_create(fs.mesh->handle, orders[l], &L.space_own));
            h1pa_form_desc d = fine.desc;
            d.q1d = 0;
            check_rc(h1pa_form_create(L.space_own, &d, &L.form_own));
            check_rc(h1pa_form_assemble(L.form_own, s));
            L.f = &L.form_own->f;
        }
        L.sp = L.f->space;
        L.mask = L.sp->ess_mask.p;
        L.prec = o.cycle_precision;
        if (L.prec == H1PA_PREC_BF16_PA) ensure_low_precision(*L.f, H1PA_PREC_BF16_PA, s);
        const int64_t n = L.sp->ndofs;
        if (l > 0) {
            // OperatorChebyshevSmoother(A_l, diag, ess, order) (ex26.cpp:99-104)
            L.smoother.reset(new Chebyshev);
            L.smoother->f = L.f;
            L.smoother->mask = L.mask;
            L.smoother->order = o.smoother_order;
            L.smoother->prec = L.prec;
            chebyshev_setup(*L.smoother, nullptr, 0.0, o.power_iterations, o.power_tolerance,
                            o.power_seed, s);
            L.R.alloc(n);
            L.Z.alloc(n);
        }
        if (l < nlevels - 1) {
            L.X.alloc(n);
            L.Y.alloc(n);
        }
    }
    for (int l = 0; l + 1 < nlevels; ++l)
        mg->P.emplace_back(new Transfer(*mg->lv[l].sp, *mg->lv[l + 1].sp, s));
    // CGSolver on the coarsest level: without a preconditioner (Jacobi with
    // dinv = 1 gives z = r, the same iteration), or with OperatorJacobiSmoother
    MgLevel &C = mg->lv[0];
    mg->coarse_dinv.alloc(C.sp->ndofs);
    if (o.coarse_jacobi) {  // the diagonal through the coarse level's X (free until a cycle)
        operator_diagonal(*C.f, C.X.p, s);
        kern::jacobi_dinv(C.sp->ndofs, C.X.p, C.mask, mg->coarse_dinv.p, s);
    } else {
        fill(C.sp->ndofs, 1.0, mg->coarse_dinv.p, s);
    }
    H1PA_HIP(hipStreamSynchronize(s));
    // the finest level reads the caller's form: hold it (Assemble, destroy
    // or a reduced-precision shadow of the other kind are refused meanwhile)
    mg->fine_ref = &fine;
    ++fine.batch_refs;
    return mg.release();
}

namespace {

// r = x - A y on level l
void residual(MgLevel &L, const double *x, const double *y, double *r, hipStream_t s) {
    constrained_apply(*L.f, L.mask, y, r, nullptr, s, L.prec);
    const int64_t n = L.sp->ndofs;
    hipLaunchKernelGGL(k_sub_from, dim3(nblocks(n)), dim3(kBlock), 0, s, n, x, r);
    H1PA_HIP(hipGetLastError());
}

// SmoothingStep (fem/multigrid.cpp:147-177), not from zero: y += S (x - A y)
void smoothing_step(MgLevel &L, const double *x, double *y, hipStream_t s) {
    residual(L, x, y, L.R.p, s);
    chebyshev_mult(*L.smoother, L.R.p, L.Z.p, s);  // the smoother is symmetric: S^T = S
    kern::axpy(L.sp->ndofs, 1.0, L.Z.p, y, s);
}

// MultigridBase::Cycle (fem/multigrid.cpp:179-240)
void cycle(Multigrid &mg, int l, const double *x, double *y, hipStream_t s) {
    MgLevel &L = mg.lv[l];
    if (l == 0) {  // coarse solve: SmoothingStep(0, zero = true)
        h1pa_pcg_opts co{mg.o.coarse_rel_tol, 0.0, mg.o.coarse_max_iter, 4};
        h1pa_pcg_stats st{};
        pcg_solve_dinv(*L.f, L.mask, mg.coarse_dinv.p, x, y, co, st, mg.coarse_work, s, L.prec);
        mg.coarse_iterations += st.iterations;
        ++mg.coarse_solves;
        return;
    }
    for (int i = 0; i < mg.o.pre_smoothing; ++i) {
        if (!mg.o.w_cycle && i == 0) chebyshev_mult(*L.smoother, x, y, s);  // from y = 0
        else smoothing_step(L, x, y, s);
    }
    MgLevel &C = mg.lv[l - 1];
    residual(L, x, y, L.R.p, s);
    mg.P[l - 1]->mult_transpose(L.R.p, C.X.p, s);
    fill(C.sp->ndofs, 0.0, C.Y.p, s);
    cycle(mg, l - 1, C.X.p, C.Y.p, s);
    if (mg.o.w_cycle) cycle(mg, l - 1, C.X.p, C.Y.p, s);
    mg.P[l - 1]->mult(C.Y.p, L.Z.p, s);
    kern::axpy(L.sp->ndofs, 1.0, L.Z.p, y, s);
    for (int i = 0; i < mg.o.post_smoothing; ++i) smoothing_step(L, x, y, s);
}

}  // namespace

// MultigridBase::Mult (fem/multigrid.cpp:106-145): y = 0, one cycle.
void multigrid_mult(Multigrid &mg, const double *x, double *y, hipStream_t s) {
    const int top = (int)mg.lv.size() - 1;
    fill(mg.lv[top].sp->ndofs, 0.0, y, s);
    cycle(mg, top, x, y, s);
    ++mg.cycles;
}

// CGSolver with SetPreconditioner(multigrid) on the fine DIAG_ONE operator.
// Each cycle already waits on the host for its coarse solve, so the outer
// state is read after every iteration (no cycle runs past convergence).
void pcg_multigrid(Multigrid &mg, const double *b, double *x, const h1pa_pcg_opts &o,
                   h1pa_pcg_stats &stats, hipStream_t s) {
    MgLevel &F = mg.lv.back();
    const Precond M = [&mg](const double *r, double *z, const int *, hipStream_t st) {
        multigrid_mult(mg, r, z, st);
    };
    pcg_preconditioned(*F.f, F.mask, M, b, x, o, stats, 1, s, &mg.outer_work);
}

void multigrid_info(const Multigrid &mg, int64_t *info) {
    info[0] = (int64_t)mg.lv.size();
    info[1] = mg.cycles;
    info[2] = mg.coarse_solves;
    info[3] = mg.coarse_iterations;
}

double multigrid_max_eig(const Multigrid &mg, int level) {
    const MgLevel &L = mg.lv[level];
    return L.smoother ? L.smoother->max_eig : 0.0;
}

}  // namespace h1pa
