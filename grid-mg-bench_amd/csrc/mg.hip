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
#include <chrono>
#include <memory>
#include <type_traits>
#include <vector>

#include "aggregation.hpp"
#include "bkrylov.hpp"
#include "coarse_op.hpp"
#include "fine_op.hpp"
#include "krylov.hpp"

namespace mgc {

template <typename R> constexpr int prec_of() { return sizeof(R) == 8 ? MGC_F64 : MGC_F32; }

static void ok(int st) {
  if (st != MGC_OK) throw Error(st, mgc_last_error());
}

template <typename R>
struct FineLin : LinOp<R> {
  mgc_fine_op* op;
  explicit FineLin(mgc_fine_op* o) : op(o) { this->n = (size_t)o->V * 12; }
  void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override {
    ok(mgc_fine_op_M(op, in, out, prec_of<R>(), 1, (void*)s));
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

// one level of the hierarchy in precision R
template <typename R>
struct Level {
  int index = 0;
  size_t n = 0;                        // vector length at this level
  std::unique_ptr<LinOp<R>> op;        // this level's operator
  mgc_aggregation* agg = nullptr;      // to the next level (null on the coarsest)
  mgc_coarse_op* cop64 = nullptr;      // next level's operator, FP64 (Galerkin build)
  mgc_coarse_op* copR = nullptr;       // next level's operator in R (== cop64 for FP64)
  std::unique_ptr<Gmres<R>> smoother;  // GMRES smoother / coarsest solver at this level
  std::unique_ptr<Gmres<R>> wrapper;   // K-cycle FGMRES on the next level
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
  mgc_fine_op* op;
  BFineLin(mgc_fine_op* o, int nrhs) : op(o) {
    this->sh.n = (size_t)o->V * 12 * nrhs;
    this->sh.nrhs = nrhs;
    this->sh.dof = 12;
  }
  void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override {
    ok(mgc_fine_op_M(op, in, out, prec_of<R>(), this->sh.nrhs, (void*)s));
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
  mgc_fine_op* fop;
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

  MG(mgc_fine_op* f, const mgc_mg_params& pp, hipStream_t s) : fop(f), p(pp), bl(s) {}
  ~MG() override {
    free_batch();
    L.clear();
    if (in_r) (void)hipFree(in_r);
    if (out_r) (void)hipFree(out_r);
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

  void setup(hipStream_t s) override {
    MGC_CHECK(p.nlevels >= 2 && p.nlevels <= 4, MGC_ERR_ARG, "nlevels must be 2..4");
    L.clear();
    precs.clear();
    // level 0: the fine operator
    {
      auto lv = std::make_unique<Level<R>>();
      lv->index = 0;
      lv->op = std::make_unique<FineLin<R>>(fop);
      lv->n = lv->op->n;
      L.push_back(std::move(lv));
    }
    int dims[4];
    for (int d = 0; d < 4; ++d) dims[d] = fop->lat.L[d];
    int Ef = 12;
    mgc_coarse_op* prev64 = nullptr;
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
      // test vectors: Gaussian, then setup_iters rounds of psi = A^{-1} v (approx.)
      const size_t n = lv.n;
      cplx<R>* tv = nullptr;
      cplx<double>* tv64 = nullptr;
      MGC_HIP(hipMalloc(&tv, (size_t)nb * n * sizeof(cplx<R>)));
      MGC_HIP(hipMalloc(&tv64, (size_t)nb * n * sizeof(cplx<double>)));
      try {
        ok(mgc_random_gaussian(tv, (size_t)nb * n, prec_of<R>(), p.seed, 1000 + l, (void*)s));
        Gmres<R> setup_solver(n, std::max(1, std::min(p.setup_solve_iters, 32)), false);
        cplx<R>* src = nullptr;
        MGC_HIP(hipMalloc(&src, n * sizeof(cplx<R>)));
        for (int it = 0; it < p.setup_iters; ++it)
          for (int k = 0; k < nb; ++k) {
            cplx<R>* v = tv + (size_t)k * n;
            const double nv = std::sqrt(vnorm2<R>(bl, v, n));
            vconvert<R, R>(v, src, n, 1.0 / nv, s);
            setup_solver.solve(bl, *lv.op, nullptr, src, v, true, p.setup_tol, p.setup_solve_iters);
          }
        (void)hipFree(src);
        vconvert<R, double>(tv, tv64, (size_t)nb * n, 1.0, s);
        ok(mgc_aggregation_set_basis(lv.agg, tv64, (void*)s));
        ok(mgc_aggregation_orthogonalise(lv.agg, 2, (void*)s));
      } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(tv);
        (void)hipFree(tv64);
        throw;
      }
      MGC_HIP(hipStreamSynchronize(s));
      (void)hipFree(tv);
      (void)hipFree(tv64);
      // Galerkin operator of the next level, FP64
      ok(mgc_coarse_op_create(&lv.cop64, cd, 2 * nb, MGC_F64, 0));
      if (l == 0) ok(mgc_coarsen_operator(lv.cop64, fop, lv.agg, p.coarsen_mode, (void*)s));
      else ok(mgc_coarsen_coarse_operator(lv.cop64, prev64, lv.agg, p.coarsen_mode, (void*)s));
      if (prec_of<R>() == MGC_F64) {
        lv.copR = lv.cop64;
      } else {
        const int N = 2 * nb;
        const int Vc = cd[0] * cd[1] * cd[2] * cd[3];
        ok(mgc_coarse_op_create(&lv.copR, cd, N, MGC_F32, 0));
        cplx<double>* tmp = nullptr;
        MGC_HIP(hipMalloc(&tmp, (size_t)9 * Vc * N * N * sizeof(cplx<double>)));
        const int st1 = mgc_coarse_op_get_links(lv.cop64, MGC_LAYOUT_FLAT, tmp, (void*)s);
        const int st2 = st1 == MGC_OK
                            ? mgc_coarse_op_set_links(lv.copR, MGC_LAYOUT_FLAT, tmp, 0, (void*)s)
                            : st1;
        (void)hipStreamSynchronize(s);
        (void)hipFree(tmp);
        ok(st2);
      }
      prev64 = lv.cop64;
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
    // solvers
    for (int l = 0; l < p.nlevels; ++l) {
      Level<R>& lv = *L[l];
      if (l + 1 < p.nlevels) {
        lv.smoother = std::make_unique<Gmres<R>>(lv.n, p.smoother_restart, false);
        if (l + 2 < p.nlevels)
          lv.wrapper = std::make_unique<Gmres<R>>(L[l + 1]->n, p.wrapper_restart, true);
      } else {
        lv.smoother = std::make_unique<Gmres<R>>(lv.n, p.coarsest_restart, false);
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
  }

  void precondition64(const cplx<double>* in, cplx<double>* out, hipStream_t s) override {
    MGC_CHECK(ready, MGC_ERR_STATE, "multigrid setup not run");
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
    std::unique_ptr<BGmres<R>> smoother, wrapper;
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

  void free_batch() {
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
        st->op = std::make_unique<BFineLin<R>>(fop, nrhs);
      } else {
        if constexpr (std::is_same<R, float>::value)
          st->op = std::make_unique<BCoarseLin>(L[l - 1]->copR, nrhs, coarse_prec);
      }
      const BShape& sh = st->op->sh;
      if (l + 1 < p.nlevels) {
        st->smoother = std::make_unique<BGmres<R>>(sh, p.smoother_restart, false);
        const size_t nc = L[l + 1]->n * (size_t)nrhs;
        MGC_HIP(hipMalloc(&st->rc, nc * sizeof(cplx<R>)));
        MGC_HIP(hipMalloc(&st->ec, nc * sizeof(cplx<R>)));
      } else {
        st->smoother = std::make_unique<BGmres<R>>(sh, p.coarsest_restart, false);
      }
      BL.push_back(std::move(st));
    }
    for (int l = 0; l + 2 < p.nlevels; ++l)
      BL[l]->wrapper = std::make_unique<BGmres<R>>(BL[l + 1]->op->sh, p.wrapper_restart, true);
    for (int l = 0; l < p.nlevels; ++l) bprecs.push_back(std::make_unique<BLevelPrec>(this, l));
    MGC_HIP(hipMalloc(&bin, BL[0]->op->sh.n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&bout, BL[0]->op->sh.n * sizeof(cplx<R>)));
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

  void precondition64_batch(const cplx<double>* in, cplx<double>* out, hipStream_t s) override {
    MGC_CHECK(bn > 0, MGC_ERR_STATE, "batched preconditioner not prepared");
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
  mgc_fine_op* fop = nullptr;
  mgc_mg_params params;
  std::unique_ptr<mgc::MGBase> impl;
  hipStream_t stream = nullptr;
};

using namespace mgc;

extern "C" {

int mgc_mg_params_default(mgc_mg_params* p) {
  return guarded([&] {
    MGC_CHECK(p, MGC_ERR_ARG, "null params");
    // include/MultigridParams.h:95-133 (3 levels, K-cycle, FGMRES wrapper
    // 0.1 / 10 / 5, GMRES smoother 1e-3 / 4 / 4, coarsest GMRES 5e-2 / 50 / 10)
    p->nlevels = 3;
    for (int l = 0; l < 3; ++l)
      for (int d = 0; d < 4; ++d) p->block[l][d] = l == 0 ? 4 : 2;
    p->nb[0] = p->nb[1] = p->nb[2] = 16;
    p->precision = MGC_F32;
    p->coarsen_mode = MGC_COARSEN_DIRSAVE;
    p->setup_iters = 1;
    p->setup_solve_iters = 8;
    p->setup_tol = 1e-3;
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

int mgc_mg_create(mgc_mg** out, mgc_fine_op* fop, const mgc_mg_params* params) {
  return guarded([&] {
    MGC_CHECK(out && fop && params, MGC_ERR_ARG, "null argument");
    MGC_CHECK(fop->gauge_set, MGC_ERR_STATE, "fine gauge field not set");
    MGC_CHECK(params->precision == MGC_F64 || params->precision == MGC_F32, MGC_ERR_PRECISION,
              "hierarchy precision must be F64 or F32");
    auto* m = new mgc_mg();
    m->fop = fop;
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
      m->impl = std::make_unique<MG<double>>(m->fop, m->params, s);
    else
      m->impl = std::make_unique<MG<float>>(m->fop, m->params, s);
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
    FineLin<double> A(m->fop);
    PrecAdapter P(m->impl.get());
    Gmres<double> outer(A.n, restart, true);
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
    BFineLin<double> A(m->fop, nrhs);
    BPrecAdapter P(m->impl.get());
    BGmres<double> outer(A.sh, restart, true);
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
    FineLin<double> A(fop);
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
