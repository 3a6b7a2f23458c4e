// Multigrid hierarchy for the Wilson-clover operator: test-vector setup,
// aggregation and Galerkin coarse operators on every level, the K-cycle
// preconditioner, the V-cycle, and the FP64 outer FGMRES solve.
//
// Reference semantics:
//   include/MultigridPreconditioner.h:165-211   initialSetup (test vectors on every level)
//   include/MultigridPreconditioner.h:253-309   operator(): restrict, K-cycle wrapper
//                                               FGMRES (or V-cycle recursion), prolongate,
//                                               post-smoothing GMRES
//   include/MultigridPreconditioner.h:566-641   coarsest level: GMRES
//   include/MultigridPreconditioner.h:458-477   recreateOperators: Aggregation::Create +
//                                               CoarsenOperator
//   include/BasisVectors.h:74-149               Generate (test-vector setup: solve
//                                               A psi = v from psi = 0, v <- psi)
//   include/VCycleRework.h:84-189               NonHermitianVCyclePreconditioner
//   src/Test_wilsonclover_mg.cc:140-151         FGMRES(D) preconditioned by the hierarchy
//
// Precision: the outer FGMRES and its true residuals are FP64; the
// preconditioner (smoothers, coarse operators, Krylov wrappers) runs in the
// hierarchy precision (FP64 "plain form" or FP32).  Galerkin operators are
// always built in FP64 and rounded for the FP32 hierarchy.
#include <algorithm>
#include <chrono>
#include <string>
#include <memory>
#include <type_traits>
#include <vector>

#include "aggregation.hpp"
#include "bkrylov.hpp"
#include "coarse_op.hpp"
#include "fine_op.hpp"
#include "dkrylov.hpp"
#include "krylov.hpp"

namespace mgc {

template <typename R> constexpr int prec_of() { return sizeof(R) == 8 ? MGC_F64 : MGC_F32; }

// Re-orthogonalisation threshold of the FP64 outer FGMRES: a second
// projection only when one classical projection removed more than
// 1 - 1e-4 of |w|^2.  With a strong preconditioner each new direction lies
// mostly in the span of the basis, so the default 1/2 re-projects at nearly
// every step (every one of the 43 outer steps of a 48^4 Wilson-clover solve);
// up to this threshold one FP64 projection leaves the new vector orthogonal
// to the basis to ~1e-14 and its Pythagorean norm accurate to ~1e-12.
constexpr double OUTER_REORTH_ETA = 1e-4;
// The same for the GMRES smoothers, K-cycle wrappers and coarsest solvers of
// the preconditioner: 1e-4 in FP64; 1e-2 in FP32, where one projection then
// leaves the few-step smoother basis orthogonal to ~1e-6 and its norms
// accurate to ~1e-5 (the inner products are FP64 sums over FP32 vectors).
template <typename R> constexpr double hier_reorth_eta() { return sizeof(R) == 8 ? 1e-4 : 1e-2; }

static void ok(int st) {
  if (st != MGC_OK) throw Error(st, mgc_last_error());
}

// The fine operator of a hierarchy: a caller's callback operator, or the
// library's Wilson-clover operator (whose fused kernel then builds the first
// coarse operator).  M applies the whole operator in the precision asked
// for, converting through FP64 staging buffers when the callback does not
// take that precision, and in calls of at most max_nrhs right-hand sides.
struct FineSrc {
  mgc_fine_linop lop{};
  mgc_fine_op* fop = nullptr;
  Lat4 lat{};
  int V = 0;
  DevBuf<cplx<double>> st_in, st_out;  // staging for precisions the callback lacks
  size_t st_cap = 0;

  void M(const void* in, void* out, int prec, int nrhs, hipStream_t s) {
    // vectors are [site][rhs][12]: a batch cannot be split along rhs in place
    MGC_CHECK(nrhs <= lop.max_nrhs, MGC_ERR_ARG,
              "the fine operator takes fewer right-hand sides per call than the batch");
    if (lop.precisions & (1 << prec)) {
      call(9, prec, in, out, nrhs, s);
      return;
    }
    MGC_CHECK(prec == MGC_F32 && (lop.precisions & (1 << MGC_F64)), MGC_ERR_PRECISION,
              "the fine operator takes neither this precision nor F64");
    const size_t n = (size_t)V * 12 * nrhs;
    if (n > st_cap) {
      st_in.alloc(n);
      st_out.alloc(n);
      st_cap = n;
    }
    vconvert<float, double>((const cplx<float>*)in, st_in.p, n, 1.0, s);
    call(9, MGC_F64, st_in.p, st_out.p, nrhs, s);
    vconvert<double, float>(st_out.p, (cplx<float>*)out, n, 1.0, s);
  }
  void call(int point, int prec, const void* in, void* out, int nrhs, hipStream_t s) {
    const int st = lop.apply(lop.ctx, point, prec, in, out, nrhs, (void*)s);
    if (st != MGC_OK)
      throw Error(st < 0 ? st : MGC_ERR_ARG,
                  "the fine operator callback failed (status " + std::to_string(st) + ")");
  }
};

template <typename R>
struct FineLin : LinOp<R> {
  FineSrc* src;
  explicit FineLin(FineSrc* f) : src(f) { this->n = (size_t)f->V * 12; }
  void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override {
    src->M(in, out, prec_of<R>(), 1, s);
  }
};

template <typename R>
struct CoarseLin : LinOp<R> {
  mgc_coarse_op* op;
  explicit CoarseLin(mgc_coarse_op* o) : op(o) { this->n = (size_t)o->V * o->N; }
  void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override {
    ok(mgc_coarse_op_M(op, in, out, (void*)s));
  }
};

// the solver types of include/Helpers.h:38-56 (unpreconditioned but FGMRES)
template <typename R>
std::unique_ptr<Solver<R>> make_solver(int type, size_t n, int restart, double eta = 0.5) {
  switch (type) {
    case MGC_SOLVER_GMRES:
    case MGC_SOLVER_FGMRES: {
      auto g = std::make_unique<Gmres<R>>(n, restart, type == MGC_SOLVER_FGMRES);
      g->set_reorth_eta(eta);
      return g;
    }
    case MGC_SOLVER_BICGSTAB: return std::make_unique<BiCgStab<R>>(n);
    case MGC_SOLVER_MR: return std::make_unique<MinRes<R>>(n);
  }
  throw Error(MGC_ERR_ARG, "unknown solver type " + std::to_string(type));
}

// one level of the hierarchy in precision R
template <typename R>
struct Level {
  int index = 0;
  size_t n = 0;                        // vector length at this level
  std::unique_ptr<LinOp<R>> op;        // this level's operator
  mgc_aggregation* agg = nullptr;      // to the next level (null on the coarsest)
  mgc_coarse_op* cop64 = nullptr;      // next level's operator, FP64 (Galerkin build)
  mgc_coarse_op* copR = nullptr;       // next level's operator in R (== cop64 for FP64)
  std::unique_ptr<Solver<R>> smoother;  // smoother / coarsest solver at this level
  std::unique_ptr<Solver<R>> wrapper;   // K-cycle FGMRES on the next level
  DevBuf<cplx<R>> tv;                   // test vectors [nb][n], kept for refinement
  cplx<R>* rc = nullptr;               // next-level residual / correction
  cplx<R>* ec = nullptr;
  cplx<R>* t1 = nullptr;               // level temporaries (V-cycle)
  cplx<R>* t2 = nullptr;
  cplx<R>* t3 = nullptr;
  ~Level() {
    if (agg) (void)mgc_aggregation_destroy(agg);
    if (copR && copR != cop64) (void)mgc_coarse_op_destroy(copR);
    if (cop64) (void)mgc_coarse_op_destroy(cop64);
    void* b[] = {rc, ec, t1, t2, t3};
    for (void* p : b)
      if (p) (void)hipFree(p);
  }
};

// batched operators: nrhs right-hand sides per apply, vectors [site][rhs][dof]
template <typename R>
struct BFineLin : BLinOp<R> {
  FineSrc* src;
  BFineLin(FineSrc* f, int nrhs) : src(f) {
    this->sh.n = (size_t)f->V * 12 * nrhs;
    this->sh.nrhs = nrhs;
    this->sh.dof = 12;
  }
  void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override {
    src->M(in, out, prec_of<R>(), this->sh.nrhs, s);
  }
};

// coarse levels of the batched solve: the multi-rhs MFMA stencil with FP32
// vectors, in FP32 or in FP16/BF16 operands (quantised on the fly)
struct BCoarseLin : BLinOp<float> {
  mgc_coarse_op* op;
  int mprec;
  BCoarseLin(mgc_coarse_op* o, int nrhs, int mp) : op(o), mprec(mp) {
    sh.n = (size_t)o->V * o->N * nrhs;
    sh.nrhs = nrhs;
    sh.dof = o->N;
  }
  void apply(const cplx<float>* in, cplx<float>* out, hipStream_t s) override {
    ok(mgc_coarse_op_M_mrhs(op, in, nullptr, MGC_F32, out, sh.nrhs, mprec, 0, (void*)s));
  }
};

struct MGBase {
  virtual ~MGBase() {}
  virtual void setup(hipStream_t s) = 0;
  // FP64 in/out wrappers
  virtual void precondition64(const cplx<double>* in, cplx<double>* out, hipStream_t s) = 0;
  virtual void vcycle64(const cplx<double>* in, cplx<double>* out, int max_iter, double tol,
                        hipStream_t s) = 0;
  virtual mgc_coarse_op* coarse_op(int level) = 0;
  virtual mgc_aggregation* aggregation(int level) = 0;
  // batched preconditioner for nrhs right-hand sides (FP32 hierarchy only)
  virtual void prepare_batch(int nrhs, int coarse_prec, hipStream_t s) = 0;
  virtual void precondition64_batch(const cplx<double>* in, cplx<double>* out, hipStream_t s) = 0;
  int precond_applies = 0;
};

template <typename R>
struct MG : MGBase {
  FineSrc* fsrc;
  mgc_mg_params p;
  std::vector<std::unique_ptr<Level<R>>> L;
  Blas bl;
  cplx<R>* in_r = nullptr;
  cplx<R>* out_r = nullptr;
  bool ready = false;

  // K-cycle preconditioner of level l as a Prec (used by the level l-1 wrapper)
  struct LevelPrec : Prec<R> {
    MG* mg;
    int l;
    LevelPrec(MG* m, int lv) : mg(m), l(lv) {}
    void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override { mg->kcycle(l, in, out, s); }
  };
  std::vector<std::unique_ptr<LevelPrec>> precs;

  // graph mode (p.graph != 0): device slots holding the current in / out
  // pointers, |in|^2, the pinned staging of the slots, and the captured graph
  DevBuf<void*> slots;
  DevBuf<double> d_nin;
  void** h_slots = nullptr;
  hipEvent_t slots_done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  MG(FineSrc* f, const mgc_mg_params& pp, hipStream_t s) : fsrc(f), p(pp), bl(s) {}
  ~MG() override {
    free_batch();
    free_graph();
    L.clear();
    if (in_r) (void)hipFree(in_r);
    if (out_r) (void)hipFree(out_r);
    if (h_slots) (void)hipHostFree(h_slots);
    if (slots_done) (void)hipEventDestroy(slots_done);
  }
  void free_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  // restrict, coarse solve (K-cycle wrapper or coarsest GMRES), prolongate,
  // post-smooth: MultiGridPreconditioner::operator()
  void kcycle(int l, const cplx<R>* in, cplx<R>* out, hipStream_t s) {
    Level<R>& lv = *L[l];
    Level<R>& nx = *L[l + 1];
    agg_transfer(lv.agg, in, lv.rc, prec_of<R>(), 1, true, 0, 8, s);
    if (l + 1 == p.nlevels - 1) {
      nx.smoother->solve(bl, *nx.op, nullptr, lv.rc, lv.ec, true, p.coarsest_tol, p.coarsest_iters);
    } else {
      lv.wrapper->solve(bl, *nx.op, precs[l + 1].get(), lv.rc, lv.ec, true, p.wrapper_tol,
                        p.wrapper_iters);
    }
    agg_transfer(lv.agg, lv.ec, out, prec_of<R>(), 1, false, 0, 8, s);
    lv.smoother->solve(bl, *lv.op, nullptr, in, out, false, p.smoother_tol, p.smoother_iters);
  }

  // Orthonormalise the set of nb vectors [nb][n] over the whole lattice
  // (Gram-Schmidt with a second projection, as the reference's
  // orthogonalise(V) of include/BasisVectors.h:152-156 does for the set)
  void orthonormalise_set(cplx<R>* V, int nb, size_t n, hipStream_t s) {
    std::vector<cplx<double>> h(nb);
    for (int i = 0; i < nb; ++i) {
      cplx<R>* v = V + (size_t)i * n;
      for (int rep = 0; rep < 2 && i > 0; ++rep) {
        vmultidot<R>(bl, V, n, i, v, n, h.data());
        for (int j = 0; j < i; ++j) h[j] = {-h[j].re, -h[j].im};
        vmultiaxpy<R>(bl, h.data(), V, n, i, v, n);
      }
      const double nv = std::sqrt(vnorm2<R>(bl, v, n));
      MGC_CHECK(nv > 0.0 && std::isfinite(nv), MGC_ERR_STATE,
                "test vectors linearly dependent or not finite in the setup");
      vscale<R>(1.0 / nv, v, n, s);
    }
  }

  // Test vectors of level l (MGBasisVectors::Generate, include/BasisVectors.h:74-149):
  // Gaussian, then setup_iters rounds of [pre-orthonormalise,] v_k <- A^-1 v_k
  // approximately (test vectors: from 0 with rhs v_k; null vectors: A psi = 0
  // from psi = v_k) [, post-orthonormalise].
  void generate(int l, cplx<R>* tv, hipStream_t s) {
    Level<R>& lv = *L[l];
    const size_t n = lv.n;
    const int nb = p.nb[l];
    ok(mgc_random_gaussian(tv, (size_t)nb * n, prec_of<R>(), p.seed, 1000 + l, (void*)s));
    auto solver = make_solver<R>(p.setup_solver, n, std::max(1, std::min(p.setup_restart, 63)));
    const bool nullv = p.setup_basis_type == MGC_BASIS_NULL_VECTORS;
    const int nbatch = setup_batch(l);
    DevBuf<cplx<R>> src(n), e(n);
    for (int it = 0; it < p.setup_iters; ++it) {
      if (p.setup_pre_orthonormalise) orthonormalise_set(tv, nb, n, s);
      for (int k = 0; k < nb; ++k) {
        cplx<R>* v = tv + (size_t)k * n;
        const double nv = std::sqrt(vnorm2<R>(bl, v, n));
        MGC_CHECK(nv > 0.0 && std::isfinite(nv), MGC_ERR_STATE,
                  "test vector with zero or non-finite norm in the setup");
        vscale<R>(1.0 / nv, v, n, s);
        if (nbatch > 1) continue;  // solved below, nbatch at a time
        if (!nullv) {
          vcopy<R>(v, src.p, n, s);
          solver->solve(bl, *lv.op, nullptr, src.p, v, true, p.setup_tol, p.setup_solve_iters);
        } else if (p.setup_solver == MGC_SOLVER_BICGSTAB) {
          // rhs 0: BiCGSTAB relaxes psi = v for exactly setup_solve_iters steps
          vzero<R>(src.p, n, s);
          solver->solve(bl, *lv.op, nullptr, src.p, v, false, p.setup_tol, p.setup_solve_iters);
        } else {
          // A psi = 0 from psi = v as psi = v + e with A e = -A v, to setup_tol of |A v|
          lv.op->apply(v, src.p, s);
          vscale<R>(-1.0, src.p, n, s);
          solver->solve(bl, *lv.op, nullptr, src.p, e.p, true, p.setup_tol, p.setup_solve_iters);
          vaxpy<R>({1.0, 0.0}, e.p, v, n, s);
        }
      }
      if (nbatch > 1) generate_batched(l, tv, nbatch, s);
      if (p.setup_post_orthonormalise) orthonormalise_set(tv, nb, n, s);
    }
  }

  // Right-hand sides per lockstep batch of the set-up solves on level l: 1
  // (one at a time) unless setup_nrhs asks for more and the level has a
  // multi-rhs operator (the fine operator up to its max_nrhs; a coarse level
  // of an FP32 hierarchy, through the FP32 multi-rhs stencil)
  int setup_batch(int l) const {
    int b = std::min(p.setup_nrhs, std::min(p.nb[l], (int)BBlas::RMAX));
    if (l == 0) b = std::min(b, fsrc->lop.max_nrhs);
    if (l > 0 && !std::is_same<R, float>::value) b = 1;
    return std::max(b, 1);
  }

  std::unique_ptr<BLinOp<R>> batched_op(int l, int nr) {
    if (l == 0) return std::make_unique<BFineLin<R>>(fsrc, nr);
    if constexpr (std::is_same<R, float>::value)
      return std::make_unique<BCoarseLin>(L[l - 1]->copR, nr, MGC_F32);
    throw Error(MGC_ERR_PRECISION, "no multi-rhs operator on an FP64 coarse level");
  }

  std::unique_ptr<BSolver<R>> batched_solver(const BShape& sh) {
    switch (p.setup_solver) {
      case MGC_SOLVER_BICGSTAB: return std::make_unique<BBiCgStab<R>>(sh);
      case MGC_SOLVER_GMRES:
      case MGC_SOLVER_FGMRES:
        return std::make_unique<BGmres<R>>(sh, std::max(1, std::min(p.setup_restart, 63)),
                                           p.setup_solver == MGC_SOLVER_FGMRES);
      case MGC_SOLVER_MR: return std::make_unique<BMinRes<R>>(sh);
    }
    throw Error(MGC_ERR_ARG, "unknown solver type " + std::to_string(p.setup_solver));
  }

  // The solves of one test-vector round (the inner loop of Generate,
  // include/BasisVectors.h:89-139) on nb vectors that are already normalised,
  // in lockstep batches of up to nbatch: each batch is gathered into one
  // multi-rhs vector [site][rhs][dof], solved by the batched form of the
  // set-up solver (per rhs the recurrence and stopping rules of the single-rhs
  // solver) and scattered back.  Batches are as even as they can be, larger
  // ones first.
  void generate_batched(int l, cplx<R>* tv, int nbatch, hipStream_t s) {
    Level<R>& lv = *L[l];
    const size_t n = lv.n;
    const int nb = p.nb[l];
    const int dof = l == 0 ? 12 : L[l - 1]->copR->N;
    const bool nullv = p.setup_basis_type == MGC_BASIS_NULL_VECTORS;
    const int nchunk = (nb + nbatch - 1) / nbatch;
    DevBuf<cplx<R>> xb((size_t)nbatch * n), bp((size_t)nbatch * n), eb;
    if (nullv && p.setup_solver != MGC_SOLVER_BICGSTAB) eb.alloc((size_t)nbatch * n);
    BBlas bb(s);
    std::unique_ptr<BLinOp<R>> op;
    std::unique_ptr<BSolver<R>> solver;
    int k0 = 0;
    for (int c = 0; c < nchunk; ++c) {
      const int nr = nb / nchunk + (c < nb % nchunk ? 1 : 0);
      if (!op || op->sh.nrhs != nr) {
        solver.reset();  // the previous batch size's workspaces first
        op = batched_op(l, nr);
        op->sh.n = n * nr;
        op->sh.nrhs = nr;
        op->sh.dof = dof;
        solver = batched_solver(op->sh);
      }
      const BShape& sh = op->sh;
      cplx<R>* set = tv + (size_t)k0 * n;
      if (!nullv) {
        // A psi = v from psi = 0
        bpack<R>(set, n, bp.p, sh, s);
        solver->solve(bb, *op, nullptr, bp.p, xb.p, true, p.setup_tol, p.setup_solve_iters);
      } else if (p.setup_solver == MGC_SOLVER_BICGSTAB) {
        // rhs 0: BiCGSTAB relaxes psi = v for exactly setup_solve_iters steps
        bpack<R>(set, n, xb.p, sh, s);
        vzero<R>(bp.p, sh.n, s);
        solver->solve(bb, *op, nullptr, bp.p, xb.p, false, p.setup_tol, p.setup_solve_iters);
      } else {
        // A psi = 0 from psi = v as psi = v + e with A e = -A v
        bpack<R>(set, n, xb.p, sh, s);
        op->apply(xb.p, bp.p, s);
        vscale<R>(-1.0, bp.p, sh.n, s);
        solver->solve(bb, *op, nullptr, bp.p, eb.p, true, p.setup_tol, p.setup_solve_iters);
        vaxpy<R>({1.0, 0.0}, eb.p, xb.p, sh.n, s);
      }
      bunpack<R>(xb.p, set, n, sh, s);
      k0 += nr;
    }
    MGC_HIP(hipStreamSynchronize(s));  // the batch buffers are freed on return
  }

  // Aggregation::Create (copy + block orthonormalisation) from the level's
  // test vectors
  void install_basis(int l, const cplx<R>* tv, hipStream_t s) {
    Level<R>& lv = *L[l];
    const size_t total = (size_t)p.nb[l] * lv.n;
    DevBuf<cplx<double>> tv64(total);
    vconvert<R, double>(tv, tv64.p, total, 1.0, s);
    ok(mgc_aggregation_set_basis(lv.agg, tv64.p, (void*)s));
    ok(mgc_aggregation_orthogonalise(lv.agg, 2, (void*)s));
    MGC_HIP(hipStreamSynchronize(s));  // tv64 is freed on return
  }

  // CoarsenOperator of level l onto l+1 (FP64), rounded into the FP32 copy
  void coarsen_level(int l, hipStream_t s) {
    Level<R>& lv = *L[l];
    if (l == 0 && fsrc->fop)
      ok(mgc_coarsen_operator(lv.cop64, fsrc->fop, lv.agg, p.coarsen_mode, (void*)s));
    else if (l == 0)
      ok(mgc_coarsen_operator_linop(lv.cop64, &fsrc->lop, lv.agg, p.coarsen_mode, (void*)s));
    else
      ok(mgc_coarsen_coarse_operator(lv.cop64, L[l - 1]->cop64, lv.agg, p.coarsen_mode,
                                     (void*)s));
    if (lv.copR != lv.cop64) {
      const size_t cnt = (size_t)9 * lv.cop64->V * lv.cop64->N * lv.cop64->N;
      DevBuf<cplx<double>> tmp(cnt);
      ok(mgc_coarse_op_get_links(lv.cop64, MGC_LAYOUT_FLAT, tmp.p, (void*)s));
      ok(mgc_coarse_op_set_links(lv.copR, MGC_LAYOUT_FLAT, tmp.p, 0, (void*)s));
      MGC_HIP(hipStreamSynchronize(s));
    }
  }

  void setup(hipStream_t s) override {
    MGC_CHECK(p.nlevels >= 2 && p.nlevels <= 4, MGC_ERR_ARG, "nlevels must be 2..4");
    ready = false;
    L.clear();
    precs.clear();
    // level 0: the fine operator
    {
      auto lv = std::make_unique<Level<R>>();
      lv->index = 0;
      lv->op = std::make_unique<FineLin<R>>(fsrc);
      lv->n = lv->op->n;
      L.push_back(std::move(lv));
    }
    int dims[4];
    for (int d = 0; d < 4; ++d) dims[d] = fsrc->lat.L[d];
    int Ef = 12;
    const bool keep_tv = p.setup_refine_iters > 0;
    for (int l = 0; l + 1 < p.nlevels; ++l) {
      Level<R>& lv = *L[l];
      const int nb = p.nb[l];
      int cd[4];
      for (int d = 0; d < 4; ++d) {
        MGC_CHECK(p.block[l][d] >= 1 && dims[d] % p.block[l][d] == 0, MGC_ERR_SHAPE,
                  "block does not divide the level lattice");
        cd[d] = dims[d] / p.block[l][d];
      }
      ok(mgc_aggregation_create_ex(&lv.agg, dims, p.block[l], nb, Ef));
      lv.tv.alloc((size_t)nb * lv.n);
      generate(l, lv.tv.p, s);
      install_basis(l, lv.tv.p, s);
      if (!keep_tv) lv.tv.reset();
      // Galerkin operator of the next level (FP64, rounded for an FP32 hierarchy)
      ok(mgc_coarse_op_create(&lv.cop64, cd, 2 * nb, MGC_F64, 0));
      if (prec_of<R>() == MGC_F64) lv.copR = lv.cop64;
      else ok(mgc_coarse_op_create(&lv.copR, cd, 2 * nb, MGC_F32, 0));
      coarsen_level(l, s);
      auto nx = std::make_unique<Level<R>>();
      nx->index = l + 1;
      nx->op = std::make_unique<CoarseLin<R>>(lv.copR);
      nx->n = nx->op->n;
      MGC_HIP(hipMalloc(&lv.rc, nx->n * sizeof(cplx<R>)));
      MGC_HIP(hipMalloc(&lv.ec, nx->n * sizeof(cplx<R>)));
      for (int d = 0; d < 4; ++d) dims[d] = cd[d];
      Ef = 2 * nb;
      L.push_back(std::move(nx));
    }
    // solvers: smoother (GMRES or MR), K-cycle FGMRES wrapper, coarsest
    // solver; in graph mode device-scalar GMRES / FGMRES with fixed step
    // counts (smoother_iters, wrapper_restart, coarsest_restart)
    free_graph();
    for (int l = 0; l < p.nlevels; ++l) {
      Level<R>& lv = *L[l];
      if (p.graph) {
        auto dg = [&](size_t n, int m, bool flexible) {
          auto g = std::make_unique<DGmres<R>>(n, m, flexible);
          g->set_reorth_eta(hier_reorth_eta<R>());
          return g;
        };
        if (l + 1 < p.nlevels) {
          lv.smoother = dg(lv.n, p.smoother_iters, false);
          if (l + 2 < p.nlevels) lv.wrapper = dg(L[l + 1]->n, p.wrapper_restart, true);
        } else {
          lv.smoother = dg(lv.n, p.coarsest_restart, false);
        }
      } else if (l + 1 < p.nlevels) {
        lv.smoother = make_solver<R>(p.smoother_solver, lv.n, p.smoother_restart,
                                     hier_reorth_eta<R>());
        if (l + 2 < p.nlevels)
          lv.wrapper = make_solver<R>(MGC_SOLVER_FGMRES, L[l + 1]->n, p.wrapper_restart,
                                      hier_reorth_eta<R>());
      } else {
        lv.smoother = make_solver<R>(p.coarsest_solver, lv.n, p.coarsest_restart,
                                     hier_reorth_eta<R>());
      }
      MGC_HIP(hipMalloc(&lv.t1, lv.n * sizeof(cplx<R>)));
      MGC_HIP(hipMalloc(&lv.t2, lv.n * sizeof(cplx<R>)));
      MGC_HIP(hipMalloc(&lv.t3, lv.n * sizeof(cplx<R>)));
    }
    for (int l = 0; l < p.nlevels; ++l) precs.push_back(std::make_unique<LevelPrec>(this, l));
    if (!in_r) MGC_HIP(hipMalloc(&in_r, L[0]->n * sizeof(cplx<R>)));
    if (!out_r) MGC_HIP(hipMalloc(&out_r, L[0]->n * sizeof(cplx<R>)));
    MGC_HIP(hipStreamSynchronize(s));
    ready = true;
    if (keep_tv) refine(s);
  }

  // Refinement setup (include/MultigridPreconditioner.h:213-251, the
  // Wuppertal form): on each level l, setup_refine_iters rounds of
  // v_k <- K-cycle_l(v_k) with the current hierarchy; with recursive updates
  // the coarse correction of that K-cycle becomes test vector k of level l+1.
  // After each round the operators of level l and every coarser level are
  // rebuilt from the current test vectors.
  void refine(hipStream_t s) {
    for (int l = 0; l + 1 < p.nlevels; ++l) {
      Level<R>& lv = *L[l];
      const size_t n = lv.n;
      const int nb = p.nb[l];
      DevBuf<cplx<R>> out(n);
      for (int it = 0; it < p.setup_refine_iters; ++it) {
        if (p.setup_pre_orthonormalise) orthonormalise_set(lv.tv.p, nb, n, s);
        for (int k = 0; k < nb; ++k) {
          cplx<R>* v = lv.tv.p + (size_t)k * n;
          kcycle(l, v, out.p, s);
          vcopy<R>(out.p, v, n, s);
          if (p.setup_recursive_updates && l + 2 < p.nlevels && k < p.nb[l + 1])
            vcopy<R>(lv.ec, L[l + 1]->tv.p + (size_t)k * L[l + 1]->n, L[l + 1]->n, s);
        }
        if (p.setup_post_orthonormalise) orthonormalise_set(lv.tv.p, nb, n, s);
        for (int m = l; m + 1 < p.nlevels; ++m) {
          if (m == l || p.setup_recursive_updates) install_basis(m, L[m]->tv.p, s);
          coarsen_level(m, s);
        }
      }
    }
    MGC_HIP(hipStreamSynchronize(s));
  }

  // the preconditioner on the vectors named by the device slots, with no host
  // synchronisation: |in|^2 on the device, in / |in| rounded to R, the
  // K-cycle, the result scaled back by |in|
  void precondition_dev(hipStream_t s) {
    const size_t n = L[0]->n;
    auto pin = reinterpret_cast<const cplx<double>* const*>(slots.p);
    auto pout = reinterpret_cast<cplx<double>* const*>(slots.p + 1);
    vnorm2_dev<double>(bl, nullptr, pin, n, d_nin.p);
    vconvert_dev<double, R>(nullptr, in_r, pin, nullptr, n, d_nin.p, 1, s);
    kcycle(0, in_r, out_r, s);
    vconvert_dev<R, double>(out_r, nullptr, nullptr, pout, n, d_nin.p, 2, s);
  }

  // graph mode: 1 = the device-scalar K-cycle issued eagerly, 2 = the same
  // captured into a HIP graph at the first call (after one eager application
  // that also makes every first-use allocation) and replayed from then on
  void precondition_graph(const cplx<double>* in, cplx<double>* out, hipStream_t s) {
    MGC_CHECK(s != nullptr, MGC_ERR_ARG,
              "graph mode needs a created stream (the null stream cannot be captured)");
    if (!slots.p) {
      slots.alloc(2);
      d_nin.alloc(2);
      MGC_HIP(hipHostMalloc((void**)&h_slots, 2 * sizeof(void*), hipHostMallocDefault));
      MGC_HIP(hipEventCreateWithFlags(&slots_done, hipEventDisableTiming));
      MGC_HIP(hipEventRecord(slots_done, s));
    }
    MGC_HIP(hipEventSynchronize(slots_done));  // the last upload has been read
    h_slots[0] = const_cast<cplx<double>*>(in);
    h_slots[1] = out;
    MGC_HIP(hipMemcpyAsync(slots.p, h_slots, 2 * sizeof(void*), hipMemcpyHostToDevice, s));
    MGC_HIP(hipEventRecord(slots_done, s));
    if (p.graph == 1) {
      precondition_dev(s);
      return;
    }
    if (!gexec) {
      precondition_dev(s);  // first-use allocations happen outside the capture
      MGC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        precondition_dev(s);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      MGC_HIP(hipStreamEndCapture(s, &graph));
      MGC_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }
    MGC_HIP(hipGraphLaunch(gexec, s));
  }

  void precondition64(const cplx<double>* in, cplx<double>* out, hipStream_t s) override {
    MGC_CHECK(ready, MGC_ERR_STATE, "multigrid setup not run");
    if (p.graph) {
      precondition_graph(in, out, s);
      ++precond_applies;
      return;
    }
    const size_t n = L[0]->n;
    // normalise the residual before rounding it to the hierarchy precision
    const double nin = std::sqrt(vnorm2<double>(bl, in, n));
    if (nin == 0.0) {
      vzero<double>(out, n, s);
      return;
    }
    vconvert<double, R>(in, in_r, n, 1.0 / nin, s);
    kcycle(0, in_r, out_r, s);
    vconvert<R, double>(out_r, out, n, nin, s);
    ++precond_applies;
  }

  // NonHermitianVCyclePreconditioner::operator() at level 0 (VCycleRework.h:84-189)
  void vcycle64(const cplx<double>* in, cplx<double>* out, int max_iter, double tol,
                hipStream_t s) override {
    MGC_CHECK(ready, MGC_ERR_STATE, "multigrid setup not run");
    Level<R>& lv = *L[0];
    const size_t n = lv.n;
    const double nin = std::sqrt(vnorm2<double>(bl, in, n));
    if (nin == 0.0) {
      vzero<double>(out, n, s);
      return;
    }
    cplx<R>* r = in_r;
    cplx<R>* x = out_r;
    cplx<R>* delta = lv.t1;
    cplx<R>* tmp = lv.t2;
    vconvert<double, R>(in, r, n, 1.0 / nin, s);
    vzero<R>(x, n, s);
    const double r2in = vnorm2<R>(bl, r, n);
    const double rsq = tol * tol * r2in;
    auto correct = [&]() {  // x += delta, r -= A delta
      vaxpy<R>({1.0, 0.0}, delta, x, n, s);
      lv.op->apply(delta, tmp, s);
      vaxpy<R>({-1.0, 0.0}, tmp, r, n, s);
    };
    for (int it = 0; it < max_iter; ++it) {
      lv.smoother->solve(bl, *lv.op, nullptr, r, delta, true, p.smoother_tol, p.smoother_iters);
      correct();
      agg_transfer(lv.agg, r, lv.rc, prec_of<R>(), 1, true, 0, 8, s);
      Level<R>& nx = *L[1];
      if (p.nlevels == 2)
        nx.smoother->solve(bl, *nx.op, nullptr, lv.rc, lv.ec, true, p.coarsest_tol,
                           p.coarsest_iters);
      else
        lv.wrapper->solve(bl, *nx.op, precs[1].get(), lv.rc, lv.ec, true, p.wrapper_tol,
                          p.wrapper_iters);
      agg_transfer(lv.agg, lv.ec, delta, prec_of<R>(), 1, false, 0, 8, s);
      correct();
      lv.smoother->solve(bl, *lv.op, nullptr, r, delta, true, p.smoother_tol, p.smoother_iters);
      correct();
      if (tol > 0.0 && vnorm2<R>(bl, r, n) <= rsq) break;
    }
    vconvert<R, double>(x, out, n, nin, s);
  }

  // ---- batched K-cycle (nrhs right-hand sides in lockstep) ----------------
  struct BLevelState {
    std::unique_ptr<BLinOp<R>> op;
    std::unique_ptr<BSolver<R>> smoother;
    std::unique_ptr<BSolver<R>> wrapper;
    cplx<R>* rc = nullptr;  // next-level residual / correction, batched
    cplx<R>* ec = nullptr;
    ~BLevelState() {
      if (rc) (void)hipFree(rc);
      if (ec) (void)hipFree(ec);
    }
  };
  struct BLevelPrec : BPrec<R> {
    MG* mg;
    int l;
    BLevelPrec(MG* m, int lv) : mg(m), l(lv) {}
    void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override {
      mg->kcycle_b(l, in, out, s);
    }
  };
  int bn = 0, bprec = -1;
  std::vector<std::unique_ptr<BLevelState>> BL;
  std::vector<std::unique_ptr<BLevelPrec>> bprecs;
  std::unique_ptr<BBlas> bbl;
  cplx<R>* bin = nullptr;
  cplx<R>* bout = nullptr;
  // graph mode: per-rhs |in|^2 on the device and the captured batched K-cycle
  DevBuf<cplx<double>> bnin;
  hipGraph_t bgraph = nullptr;
  hipGraphExec_t bgexec = nullptr;

  void free_batch() {
    if (bgexec) (void)hipGraphExecDestroy(bgexec);
    if (bgraph) (void)hipGraphDestroy(bgraph);
    bgexec = nullptr;
    bgraph = nullptr;
    BL.clear();
    bprecs.clear();
    if (bin) (void)hipFree(bin);
    if (bout) (void)hipFree(bout);
    bin = bout = nullptr;
    bn = 0;
  }

  void prepare_batch(int nrhs, int coarse_prec, hipStream_t s) override {
    MGC_CHECK(ready, MGC_ERR_STATE, "multigrid setup not run");
    MGC_CHECK(prec_of<R>() == MGC_F32, MGC_ERR_PRECISION,
              "the batched solve needs the FP32 hierarchy (multi-rhs MFMA coarse levels)");
    MGC_CHECK(coarse_prec == MGC_F32 || coarse_prec == MGC_F16 || coarse_prec == MGC_BF16,
              MGC_ERR_PRECISION, "coarse precision must be F32, F16 or BF16");
    if (bn == nrhs && bprec == coarse_prec) return;
    free_batch();
    if (!bbl) bbl = std::make_unique<BBlas>(s);
    for (int l = 0; l < p.nlevels; ++l) {
      auto st = std::make_unique<BLevelState>();
      if (l == 0) {
        st->op = std::make_unique<BFineLin<R>>(fsrc, nrhs);
      } else {
        if constexpr (std::is_same<R, float>::value)
          st->op = std::make_unique<BCoarseLin>(L[l - 1]->copR, nrhs, coarse_prec);
      }
      const BShape& sh = st->op->sh;
      auto bsolver = [&](int type, int restart) -> std::unique_ptr<BSolver<R>> {
        if (type == MGC_SOLVER_MR) return std::make_unique<BMinRes<R>>(sh);
        MGC_CHECK(type == MGC_SOLVER_GMRES, MGC_ERR_ARG,
                  "the batched solve takes GMRES or MR smoothers and coarsest solvers");
        if (p.graph) {  // device-scalar, fixed step counts (checked GMRES in graph mode)
          auto g = std::make_unique<DBGmres<R>>(sh, restart, false);
          g->set_reorth_eta(hier_reorth_eta<R>());
          return g;
        }
        auto g = std::make_unique<BGmres<R>>(sh, restart, false);
        g->set_reorth_eta(hier_reorth_eta<R>());
        return g;
      };
      if (l + 1 < p.nlevels) {
        st->smoother = bsolver(p.smoother_solver, p.graph ? p.smoother_iters : p.smoother_restart);
        const size_t nc = L[l + 1]->n * (size_t)nrhs;
        MGC_HIP(hipMalloc(&st->rc, nc * sizeof(cplx<R>)));
        MGC_HIP(hipMalloc(&st->ec, nc * sizeof(cplx<R>)));
      } else {
        st->smoother = bsolver(p.coarsest_solver, p.coarsest_restart);
      }
      BL.push_back(std::move(st));
    }
    for (int l = 0; l + 2 < p.nlevels; ++l) {
      if (p.graph) {
        auto g = std::make_unique<DBGmres<R>>(BL[l + 1]->op->sh, p.wrapper_restart, true);
        g->set_reorth_eta(hier_reorth_eta<R>());
        BL[l]->wrapper = std::move(g);
      } else {
        auto g = std::make_unique<BGmres<R>>(BL[l + 1]->op->sh, p.wrapper_restart, true);
        g->set_reorth_eta(hier_reorth_eta<R>());
        BL[l]->wrapper = std::move(g);
      }
    }
    for (int l = 0; l < p.nlevels; ++l) bprecs.push_back(std::make_unique<BLevelPrec>(this, l));
    MGC_HIP(hipMalloc(&bin, BL[0]->op->sh.n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&bout, BL[0]->op->sh.n * sizeof(cplx<R>)));
    // one apply of every coarse level on zero vectors here, so that the
    // multi-rhs variant tuner and the quantisation workspace of each shape
    // run at preparation and not inside the first solve
    for (int l = 1; l < p.nlevels; ++l) {
      vzero<R>(BL[l - 1]->rc, BL[l]->op->sh.n, s);
      BL[l]->op->apply(BL[l - 1]->rc, BL[l - 1]->ec, s);
    }
    MGC_HIP(hipStreamSynchronize(s));
    bn = nrhs;
    bprec = coarse_prec;
  }

  // MultiGridPreconditioner::operator() for nrhs right-hand sides at once
  void kcycle_b(int l, const cplx<R>* in, cplx<R>* out, hipStream_t s) {
    BLevelState& lv = *BL[l];
    BLevelState& nx = *BL[l + 1];
    agg_transfer(L[l]->agg, in, lv.rc, prec_of<R>(), bn, true, 0, 8, s);
    if (l + 1 == p.nlevels - 1)
      nx.smoother->solve(*bbl, *nx.op, nullptr, lv.rc, lv.ec, true, p.coarsest_tol,
                         p.coarsest_iters);
    else
      lv.wrapper->solve(*bbl, *nx.op, bprecs[l + 1].get(), lv.rc, lv.ec, true, p.wrapper_tol,
                        p.wrapper_iters);
    agg_transfer(L[l]->agg, lv.ec, out, prec_of<R>(), bn, false, 0, 8, s);
    lv.smoother->solve(*bbl, *lv.op, nullptr, in, out, false, p.smoother_tol, p.smoother_iters);
  }

  // graph mode: the per-rhs norms and conversions on the device around the
  // batched K-cycle on bin / bout, which graph = 2 captures into a HIP graph
  // at the first call (after one eager application that makes every
  // first-use allocation) and replays from then on
  void precondition_batch_graph(const cplx<double>* in, cplx<double>* out, hipStream_t s) {
    MGC_CHECK(s != nullptr, MGC_ERR_ARG,
              "graph mode needs a created stream (the null stream cannot be captured)");
    const BShape sh64 = BL[0]->op->sh;
    if (!bnin.p) bnin.alloc(BBlas::RMAX);
    bdots_dev<double>(*bbl, in, 0, 0, in, sh64, true, bnin.p);  // |in|^2 per rhs
    bconvert_dev<double, R>(in, bin, sh64, bnin.p, 1, s);         // in / |in|
    if (p.graph == 1) {
      kcycle_b(0, bin, bout, s);
    } else {
      if (!bgexec) {
        kcycle_b(0, bin, bout, s);  // first-use allocations happen outside the capture
        MGC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          kcycle_b(0, bin, bout, s);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        MGC_HIP(hipStreamEndCapture(s, &bgraph));
        MGC_HIP(hipGraphInstantiate(&bgexec, bgraph, nullptr, nullptr, 0));
      }
      MGC_HIP(hipGraphLaunch(bgexec, s));
    }
    bconvert_dev<R, double>(bout, out, sh64, bnin.p, 2, s);  // * |in|
  }

  void precondition64_batch(const cplx<double>* in, cplx<double>* out, hipStream_t s) override {
    MGC_CHECK(bn > 0, MGC_ERR_STATE, "batched preconditioner not prepared");
    if (p.graph) {
      precondition_batch_graph(in, out, s);
      ++precond_applies;
      return;
    }
    BShape sh64 = BL[0]->op->sh;
    std::vector<double> nin(bn), inv(bn);
    bnorm2<double>(*bbl, in, sh64, nin.data());
    for (int r = 0; r < bn; ++r) {
      nin[r] = std::sqrt(nin[r]);
      inv[r] = nin[r] > 0.0 ? 1.0 / nin[r] : 0.0;
    }
    bconvert<double, R>(*bbl, in, bin, sh64, inv.data());
    kcycle_b(0, bin, bout, s);
    bconvert<R, double>(*bbl, bout, out, sh64, nin.data());
    ++precond_applies;
  }

  mgc_coarse_op* coarse_op(int level) override {
    MGC_CHECK(ready && level >= 1 && level < p.nlevels, MGC_ERR_ARG, "no such coarse level");
    return L[level - 1]->cop64;
  }
  mgc_aggregation* aggregation(int level) override {
    MGC_CHECK(ready && level >= 0 && level + 1 < p.nlevels, MGC_ERR_ARG, "no such aggregation");
    return L[level]->agg;
  }
};

struct PrecAdapter : Prec<double> {
  MGBase* mg;
  explicit PrecAdapter(MGBase* m) : mg(m) {}
  void apply(const cplx<double>* in, cplx<double>* out, hipStream_t s) override {
    mg->precondition64(in, out, s);
  }
};

struct BPrecAdapter : BPrec<double> {
  MGBase* mg;
  explicit BPrecAdapter(MGBase* m) : mg(m) {}
  void apply(const cplx<double>* in, cplx<double>* out, hipStream_t s) override {
    mg->precondition64_batch(in, out, s);
  }
};

}  // namespace mgc

struct mgc_mg {
  mgc::FineSrc fine;
  mgc_mg_params params;
  std::unique_ptr<mgc::MGBase> impl;
  hipStream_t stream = nullptr;
};

using namespace mgc;

extern "C" {

int mgc_mg_params_default(mgc_mg_params* p) {
  return guarded([&] {
    MGC_CHECK(p, MGC_ERR_ARG, "null params");
    // include/MultigridParams.h:95-133 (3 levels, blocks 4^4 then 2^4, one
    // initial setup round of BiCGSTAB 1e-3 / 100 with pre- and
    // post-orthonormalisation of test vectors, no refinement, K-cycle,
    // FGMRES wrapper 0.1 / 10 / 5, GMRES smoother 1e-3 / 4 / 4, coarsest
    // GMRES 5e-2 / 50 / 10); nb = 20 is the default NBASIS = 40 of
    // src/Test_wilsonclover_mg.cc:39-40 (20 vectors, chirally doubled)
    p->nlevels = 3;
    for (int l = 0; l < 3; ++l)
      for (int d = 0; d < 4; ++d) p->block[l][d] = l == 0 ? 4 : 2;
    p->nb[0] = p->nb[1] = p->nb[2] = 20;
    p->precision = MGC_F32;
    p->coarsen_mode = MGC_COARSEN_DIRSAVE;
    p->setup_iters = 1;
    p->setup_solve_iters = 100;
    p->setup_tol = 1e-3;
    p->setup_solver = MGC_SOLVER_BICGSTAB;
    p->setup_restart = 20;
    p->setup_basis_type = MGC_BASIS_TEST_VECTORS;
    p->setup_pre_orthonormalise = 1;
    p->setup_post_orthonormalise = 1;
    p->setup_refine_iters = 0;
    p->setup_recursive_updates = 1;
    p->smoother_solver = MGC_SOLVER_GMRES;
    p->coarsest_solver = MGC_SOLVER_GMRES;
    p->graph = 0;
    p->setup_nrhs = 0;
    p->smoother_tol = 1e-3;
    p->smoother_iters = 4;
    p->smoother_restart = 4;
    p->wrapper_tol = 1e-1;
    p->wrapper_iters = 10;
    p->wrapper_restart = 5;
    p->coarsest_tol = 5e-2;
    p->coarsest_iters = 50;
    p->coarsest_restart = 10;
    p->seed = 1ull | (2ull << 16) | (3ull << 32) | (4ull << 48);
  });
}

static void check_params(const mgc_mg_params* params) {
  MGC_CHECK(params->precision == MGC_F64 || params->precision == MGC_F32, MGC_ERR_PRECISION,
            "hierarchy precision must be F64 or F32");
  auto known = [](int t) {
    return t == MGC_SOLVER_MR || t == MGC_SOLVER_GMRES || t == MGC_SOLVER_FGMRES ||
           t == MGC_SOLVER_BICGSTAB;
  };
  MGC_CHECK(known(params->setup_solver) && known(params->coarsest_solver), MGC_ERR_ARG,
            "unknown setup or coarsest solver type");
  MGC_CHECK(params->smoother_solver == MGC_SOLVER_GMRES ||
                params->smoother_solver == MGC_SOLVER_MR,
            MGC_ERR_ARG, "the smoother is GMRES or MR");
  MGC_CHECK(params->setup_basis_type == MGC_BASIS_TEST_VECTORS ||
                params->setup_basis_type == MGC_BASIS_NULL_VECTORS,
            MGC_ERR_ARG, "unknown basis-vector type");
  MGC_CHECK(params->setup_refine_iters >= 0 && params->setup_iters >= 0, MGC_ERR_ARG,
            "negative setup iteration count");
  MGC_CHECK(params->graph >= 0 && params->graph <= 2, MGC_ERR_ARG, "graph must be 0, 1 or 2");
  MGC_CHECK(params->setup_nrhs >= 0 && params->setup_nrhs <= 64, MGC_ERR_ARG,
            "setup_nrhs must be 0..64");
  MGC_CHECK(!params->graph || (params->smoother_solver == MGC_SOLVER_GMRES &&
                               params->coarsest_solver == MGC_SOLVER_GMRES),
            MGC_ERR_ARG, "graph mode runs GMRES smoothers and a GMRES coarsest solver");
}

int mgc_mg_create_linop(mgc_mg** out, const mgc_fine_linop* lop, const mgc_mg_params* params) {
  return guarded([&] {
    MGC_CHECK(out && lop && lop->apply && params, MGC_ERR_ARG, "null argument");
    MGC_CHECK(lop->precisions & (1 << MGC_F64), MGC_ERR_PRECISION,
              "the fine operator must apply in F64 (outer solve, CoarsenOperator)");
    MGC_CHECK(lop->max_nrhs >= 1, MGC_ERR_ARG, "max_nrhs < 1");
    for (int d = 0; d < 4; ++d) MGC_CHECK(lop->dims[d] >= 1, MGC_ERR_SHAPE, "lattice extent < 1");
    check_params(params);
    auto* m = new mgc_mg();
    m->fine.lop = *lop;
    for (int d = 0; d < 4; ++d) m->fine.lat.L[d] = lop->dims[d];
    m->fine.V = m->fine.lat.volume();
    m->params = *params;
    *out = m;
  });
}

int mgc_mg_create(mgc_mg** out, mgc_fine_op* fop, const mgc_mg_params* params) {
  return guarded([&] {
    MGC_CHECK(out && fop && params, MGC_ERR_ARG, "null argument");
    MGC_CHECK(fop->gauge_set, MGC_ERR_STATE, "fine gauge field not set");
    check_params(params);
    auto* m = new mgc_mg();
    ok(mgc_fine_op_linop(fop, &m->fine.lop));
    m->fine.fop = fop;
    m->fine.lat = fop->lat;
    m->fine.V = fop->V;
    m->params = *params;
    *out = m;
  });
}

int mgc_mg_destroy(mgc_mg* m) {
  return guarded([&] { delete m; });
}

int mgc_mg_setup(mgc_mg* m, void* stream) {
  return guarded([&] {
    MGC_CHECK(m, MGC_ERR_ARG, "null handle");
    hipStream_t s = as_stream(stream);
    m->impl.reset();
    if (m->params.precision == MGC_F64)
      m->impl = std::make_unique<MG<double>>(&m->fine, m->params, s);
    else
      m->impl = std::make_unique<MG<float>>(&m->fine, m->params, s);
    m->stream = s;
    m->impl->setup(s);
  });
}

int mgc_mg_precondition(mgc_mg* m, const void* in, void* out, void* stream) {
  return guarded([&] {
    MGC_CHECK(m && m->impl && in && out, MGC_ERR_STATE, "setup not run or null vector");
    MGC_CHECK(as_stream(stream) == m->stream, MGC_ERR_ARG, "use the stream given to setup");
    m->impl->precondition64((const cplx<double>*)in, (cplx<double>*)out, m->stream);
  });
}

int mgc_mg_vcycle(mgc_mg* m, const void* in, void* out, int max_iter, double tol, void* stream) {
  return guarded([&] {
    MGC_CHECK(m && m->impl && in && out, MGC_ERR_STATE, "setup not run or null vector");
    MGC_CHECK(as_stream(stream) == m->stream, MGC_ERR_ARG, "use the stream given to setup");
    MGC_CHECK(max_iter >= 1, MGC_ERR_ARG, "max_iter < 1");
    m->impl->vcycle64((const cplx<double>*)in, (cplx<double>*)out, max_iter, tol, m->stream);
  });
}

int mgc_mg_solve(mgc_mg* m, const void* b, void* x, double tol, int maxit, int restart,
                 mgc_solve_stats* stats, void* stream) {
  return guarded([&] {
    MGC_CHECK(m && m->impl && b && x, MGC_ERR_STATE, "setup not run or null vector");
    MGC_CHECK(as_stream(stream) == m->stream, MGC_ERR_ARG, "use the stream given to setup");
    hipStream_t s = m->stream;
    Blas bl(s);
    FineLin<double> A(&m->fine);
    PrecAdapter P(m->impl.get());
    Gmres<double> outer(A.n, restart, true);
    outer.set_reorth_eta(OUTER_REORTH_ETA);
    MGC_HIP(hipStreamSynchronize(s));
    const int pa0 = m->impl->precond_applies;
    const auto t0 = std::chrono::steady_clock::now();
    const KrylovStats st =
        outer.solve(bl, A, &P, (const cplx<double>*)b, (cplx<double>*)x, true, tol, maxit);
    MGC_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
      stats->iterations = st.iters;
      stats->converged = st.converged ? 1 : 0;
      stats->true_residual = st.rel_res;
      stats->precond_applies = m->impl->precond_applies - pa0;
      stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
  });
}

int mgc_mg_solve_mrhs(mgc_mg* m, const void* b, void* x, int nrhs, int coarse_prec, double tol,
                      int maxit, int restart, mgc_solve_stats* stats, void* stream) {
  return guarded([&] {
    MGC_CHECK(m && m->impl && b && x, MGC_ERR_STATE, "setup not run or null vector");
    MGC_CHECK(as_stream(stream) == m->stream, MGC_ERR_ARG, "use the stream given to setup");
    MGC_CHECK(nrhs >= 1 && nrhs <= BBlas::RMAX, MGC_ERR_ARG, "nrhs must be 1..64");
    hipStream_t s = m->stream;
    m->impl->prepare_batch(nrhs, coarse_prec, s);
    BBlas bl(s);
    MGC_CHECK(nrhs <= m->fine.lop.max_nrhs, MGC_ERR_ARG,
              "nrhs exceeds the right-hand sides the fine operator takes per call");
    BFineLin<double> A(&m->fine, nrhs);
    BPrecAdapter P(m->impl.get());
    BGmres<double> outer(A.sh, restart, true);
    outer.set_reorth_eta(OUTER_REORTH_ETA);
    MGC_HIP(hipStreamSynchronize(s));
    const int pa0 = m->impl->precond_applies;
    const auto t0 = std::chrono::steady_clock::now();
    const BKrylovStats st =
        outer.solve(bl, A, &P, (const cplx<double>*)b, (cplx<double>*)x, true, tol, maxit);
    MGC_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats)
      for (int r = 0; r < nrhs; ++r) {
        stats[r].iterations = st.iters_rhs[r];
        stats[r].converged = st.converged[r];
        stats[r].true_residual = st.rel_res[r];
        stats[r].precond_applies = m->impl->precond_applies - pa0;
        stats[r].seconds = std::chrono::duration<double>(t1 - t0).count();
      }
  });
}

int mgc_mg_coarse_op(mgc_mg* m, int level, mgc_coarse_op** cop) {
  return guarded([&] {
    MGC_CHECK(m && m->impl && cop, MGC_ERR_STATE, "setup not run or null output");
    *cop = m->impl->coarse_op(level);
  });
}

int mgc_mg_aggregation(mgc_mg* m, int level, mgc_aggregation** agg) {
  return guarded([&] {
    MGC_CHECK(m && m->impl && agg, MGC_ERR_STATE, "setup not run or null output");
    *agg = m->impl->aggregation(level);
  });
}

int mgc_fine_gmres_solve(mgc_fine_op* fop, const void* b, void* x, double tol, int maxit,
                         int restart, mgc_solve_stats* stats, void* stream) {
  return guarded([&] {
    MGC_CHECK(fop && b && x, MGC_ERR_ARG, "null argument");
    hipStream_t s = as_stream(stream);
    Blas bl(s);
    FineSrc src;
    ok(mgc_fine_op_linop(fop, &src.lop));
    src.fop = fop;
    src.lat = fop->lat;
    src.V = fop->V;
    FineLin<double> A(&src);
    Gmres<double> solver(A.n, restart, false);
    const auto t0 = std::chrono::steady_clock::now();
    const KrylovStats st =
        solver.solve(bl, A, nullptr, (const cplx<double>*)b, (cplx<double>*)x, true, tol, maxit);
    MGC_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
      stats->iterations = st.iters;
      stats->converged = st.converged ? 1 : 0;
      stats->true_residual = st.rel_res;
      stats->precond_applies = 0;
      stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
  });
}

}  // extern "C"
