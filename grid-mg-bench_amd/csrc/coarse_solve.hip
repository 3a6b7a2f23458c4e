// Coarse solves on several lattices at once, with one multi-lattice launch
// of the multi-rhs stencil (mgc_coarse_op_M_mrhs_batched) for every operator
// application of all of them.  This is the batched form of the coarsest
// solve of the reference's preconditioner (a GMRES on each coarse operator,
// include/MultigridPreconditioner.h:566-641 and include/VCycleRework.h:84-189,
// one lattice at a time there).
//
// Each lattice k runs its own lockstep batched GMRES over its nrhs columns
// (exact-size vectors [V_k][nrhs][N] FP32, so the Krylov BLAS moves no more
// than a solve of that lattice alone) on its own host thread and stream.
// Their operator applications meet at a rendezvous: every live lattice
// deposits its input and output and records an event on its stream; the last
// to arrive makes a launch stream wait on all those events, quantises the
// inputs (FP16 / BF16), issues the one batched launch, and records a
// completion event that every lattice's stream waits on before it goes on.
// A lattice whose columns have all converged (or that stopped at maxit)
// leaves the rendezvous, and the remaining ones go on batching without it.
// The lattices' host-side Krylov steps (one small device-to-host read per
// Arnoldi step each) overlap one another on their threads.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bkrylov.hpp"
#include "coarse_op.hpp"
#include "dkrylov.hpp"

namespace mgc {

constexpr int CS_MAX = 8;  // lattices per batch (the batched stencil's limit)

static void cs_ok(int st) {
  if (st != MGC_OK) throw Error(st, mgc_last_error());
}

// The meeting point of the lattices' operator applications.
class LaunchRendezvous {
 public:
  LaunchRendezvous(const std::vector<mgc_coarse_op*>& ops, int nrhs, int mprec)
      : ops_(ops), K_(int(ops.size())), nrhs_(nrhs), mprec_(mprec), in_(K_), out_(K_),
        present_(K_, 0), ready_(K_, nullptr) {
    for (int k = 0; k < K_; ++k)
      MGC_HIP(hipEventCreateWithFlags(&ready_[k], hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) MGC_HIP(hipEventCreateWithFlags(&done_[i], hipEventDisableTiming));
    MGC_HIP(hipStreamCreateWithFlags(&ls_, hipStreamNonBlocking));
    if (mprec != MGC_F32) {
      for (int k = 0; k < K_; ++k) {
        const size_t rows = (size_t)ops[k]->V * nrhs;
        q_[k].alloc(rows * ops[k]->N * 4);
        qs_[k].alloc(rows);
      }
    }
  }
  ~LaunchRendezvous() {
    for (hipEvent_t e : ready_)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : done_)
      if (e) (void)hipEventDestroy(e);
    if (ls_) (void)hipStreamDestroy(ls_);
  }
  LaunchRendezvous(const LaunchRendezvous&) = delete;
  LaunchRendezvous& operator=(const LaunchRendezvous&) = delete;

  // start of a solve: every lattice takes part
  void open() {
    std::lock_guard<std::mutex> l(m_);
    live_ = K_;
    arrived_ = 0;
    std::fill(present_.begin(), present_.end(), 0);
    error_.clear();
  }

  // lattice k: y = M_k x (FP32 [V_k][nrhs][N]), ordered on its stream sk
  void apply(int k, const cplx<float>* x, cplx<float>* y, hipStream_t sk) {
    MGC_HIP(hipEventRecord(ready_[k], sk));
    std::unique_lock<std::mutex> l(m_);
    in_[k] = x;
    out_[k] = y;
    present_[k] = 1;
    const long g = gen_;
    if (++arrived_ == live_) launch_locked();
    else cv_.wait(l, [&] { return gen_ != g; });
    const std::string err = error_;
    l.unlock();
    if (!err.empty()) throw Error(MGC_ERR_HIP, "batched coarse apply failed: " + err);
    // the generation after g + 1 needs this lattice to arrive again, so
    // done_[g & 1] is not recorded anew before this wait is queued
    MGC_HIP(hipStreamWaitEvent(sk, done_[g & 1], 0));
  }

  // run f under the rendezvous lock: library calls from a lattice's thread
  // that may reach the stencil's tuning tables (shared by all threads) go
  // one at a time, as the batched launches do
  template <typename F>
  void serialised(F&& f) {
    std::lock_guard<std::mutex> l(m_);
    f();
  }

  // a lattice has finished its solve (or failed): the others go on without it
  void leave() {
    std::lock_guard<std::mutex> l(m_);
    --live_;
    if (live_ > 0 && arrived_ == live_) launch_locked();
  }

 private:
  void launch_locked() {
    try {
      std::vector<mgc_coarse_op*> ops;
      std::vector<const void*> ins;
      std::vector<const float*> scs;
      std::vector<void*> outs;
      for (int k = 0; k < K_; ++k) {
        if (!present_[k]) continue;
        MGC_HIP(hipStreamWaitEvent(ls_, ready_[k], 0));
        ops.push_back(ops_[k]);
        outs.push_back(out_[k]);
        if (mprec_ == MGC_F32) {
          ins.push_back(in_[k]);
        } else {
          cs_ok(mgc_quantize(in_[k], (size_t)ops_[k]->V * nrhs_, ops_[k]->N, mprec_, q_[k].p,
                             qs_[k].p, (void*)ls_));
          ins.push_back(q_[k].p);
          scs.push_back(qs_[k].p);
        }
      }
      cs_ok(mgc_coarse_op_M_mrhs_batched(ops.data(), int(ops.size()), ins.data(),
                                         mprec_ == MGC_F32 ? nullptr : scs.data(), mprec_,
                                         outs.data(), nrhs_, mprec_, 0, (void*)ls_));
      MGC_HIP(hipEventRecord(done_[gen_ & 1], ls_));
    } catch (const std::exception& e) {
      error_ = e.what();
    }
    arrived_ = 0;
    std::fill(present_.begin(), present_.end(), 0);
    ++gen_;
    cv_.notify_all();
  }

  std::vector<mgc_coarse_op*> ops_;
  int K_, nrhs_, mprec_;
  std::mutex m_;
  std::condition_variable cv_;
  int live_ = 0, arrived_ = 0;
  long gen_ = 0;
  std::vector<const cplx<float>*> in_;
  std::vector<cplx<float>*> out_;
  std::vector<char> present_;
  std::vector<hipEvent_t> ready_;
  hipEvent_t done_[2] = {nullptr, nullptr};
  hipStream_t ls_ = nullptr;
  DevBuf<unsigned char> q_[CS_MAX];  // quantised inputs (FP16 / BF16 stencil)
  DevBuf<float> qs_[CS_MAX];
  std::string error_;
};

// lattice k's operator for its solver: the rendezvous
struct RendezvousOp : BLinOp<float> {
  LaunchRendezvous* rv;
  int k;
  RendezvousOp(LaunchRendezvous* r, int kk, const mgc_coarse_op* op, int nrhs) : rv(r), k(kk) {
    sh.n = (size_t)op->V * op->N * nrhs;
    sh.nrhs = nrhs;
    sh.dof = op->N;
  }
  void apply(const cplx<float>* in, cplx<float>* out, hipStream_t s) override {
    rv->apply(k, in, out, s);
  }
};

// one lattice's solver state
struct LatticeSolve {
  mgc_coarse_op* op;
  hipStream_t s = nullptr;
  hipEvent_t fin = nullptr;
  std::unique_ptr<BBlas> bl;
  std::unique_ptr<RendezvousOp> A;
  std::unique_ptr<BGmres<float>> solver;
  DevBuf<cplx<float>> r;
  LatticeSolve(LaunchRendezvous* rv, int k, mgc_coarse_op* o, int nrhs, int restart) : op(o) {
    MGC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    MGC_HIP(hipEventCreateWithFlags(&fin, hipEventDisableTiming));
    bl = std::make_unique<BBlas>(s);
    A = std::make_unique<RendezvousOp>(rv, k, o, nrhs);
    solver = std::make_unique<BGmres<float>>(A->sh, restart, false);
    r.alloc(A->sh.n);
  }
  ~LatticeSolve() {
    solver.reset();
    A.reset();
    bl.reset();
    if (fin) (void)hipEventDestroy(fin);
    if (s) (void)hipStreamDestroy(s);
  }
  LatticeSolve(const LatticeSolve&) = delete;
  LatticeSolve& operator=(const LatticeSolve&) = delete;
};

// A prepared batch: the rendezvous and every lattice's solver (Krylov basis,
// BLAS workspaces, stream), kept from one solve to the next.
struct CoarseBatch {
  hipStream_t s;
  int device = 0;
  int nrhs, mprec;
  std::unique_ptr<LaunchRendezvous> rv;
  std::vector<std::unique_ptr<LatticeSolve>> lat;
  hipEvent_t start = nullptr;
  CoarseBatch(mgc_coarse_op* const* ops, int count, int nr, int mp, int restart, hipStream_t st)
      : s(st), nrhs(nr), mprec(mp) {
    MGC_HIP(hipGetDevice(&device));
    std::vector<mgc_coarse_op*> v(ops, ops + count);
    rv = std::make_unique<LaunchRendezvous>(v, nr, mp);
    for (int k = 0; k < count; ++k)
      lat.push_back(std::make_unique<LatticeSolve>(rv.get(), k, ops[k], nr, restart));
    MGC_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
  }
  ~CoarseBatch() {
    lat.clear();
    rv.reset();
    if (start) (void)hipEventDestroy(start);
  }
  CoarseBatch(const CoarseBatch&) = delete;
  CoarseBatch& operator=(const CoarseBatch&) = delete;
};

// Fixed-step form (no counterpart in the reference): every lattice runs the
// same number of device-scalar GMRES steps from x = 0 (DBGmres in phases),
// all of them in lockstep on one stream with one multi-lattice stencil launch
// per step and no host synchronisation, so that graph = 2 captures the whole
// solve into a HIP graph at the first call (after one eager solve that makes
// every first-use allocation and tunes the launch) and replays it after;
// graph = 1 issues the same work eagerly.  The right-hand sides and
// solutions pass through staging vectors, copied outside the graph.
struct FixedCoarseBatch {
  hipStream_t s;
  int nrhs, mprec, steps, graph;
  std::vector<mgc_coarse_op*> ops;
  std::unique_ptr<BBlas> bl;
  std::vector<std::unique_ptr<DBGmres<float>>> sol;
  DevBuf<cplx<float>> bst[CS_MAX], xst[CS_MAX], r[CS_MAX];
  DevBuf<unsigned char> q[CS_MAX];
  DevBuf<float> qs[CS_MAX];
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  FixedCoarseBatch(mgc_coarse_op* const* o, int count, int nr, int mp, int st, int gm,
                   hipStream_t str)
      : s(str), nrhs(nr), mprec(mp), steps(st), graph(gm), ops(o, o + count) {
    bl = std::make_unique<BBlas>(s);
    for (int k = 0; k < count; ++k) {
      BShape sh;
      sh.n = (size_t)ops[k]->V * ops[k]->N * nr;
      sh.nrhs = nr;
      sh.dof = ops[k]->N;
      sol.push_back(std::make_unique<DBGmres<float>>(sh, st, false));
      bst[k].alloc(sh.n);
      xst[k].alloc(sh.n);
      r[k].alloc(sh.n);
      if (mp != MGC_F32) {
        const size_t rows = (size_t)ops[k]->V * nr;
        q[k].alloc(rows * ops[k]->N * 4);
        qs[k].alloc(rows);
      }
    }
  }
  ~FixedCoarseBatch() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
  }
  FixedCoarseBatch(const FixedCoarseBatch&) = delete;
  FixedCoarseBatch& operator=(const FixedCoarseBatch&) = delete;

  // every lattice's operator on its input, in one launch
  void apply_all(const cplx<float>* const* in, cplx<float>* const* out) {
    const int K = int(ops.size());
    const void* src[CS_MAX];
    const float* sc[CS_MAX] = {};
    void* dst[CS_MAX];
    for (int k = 0; k < K; ++k) {
      src[k] = in[k];
      dst[k] = out[k];
      if (mprec != MGC_F32) {
        cs_ok(mgc_quantize(in[k], (size_t)ops[k]->V * nrhs, ops[k]->N, mprec, q[k].p, qs[k].p,
                           (void*)s));
        src[k] = q[k].p;
        sc[k] = qs[k].p;
      }
    }
    cs_ok(mgc_coarse_op_M_mrhs_batched(ops.data(), K, src, mprec == MGC_F32 ? nullptr : sc,
                                       mprec, dst, nrhs, mprec, 0, (void*)s));
  }
  // the solve on the staging vectors: no host synchronisation
  void run() {
    const int K = int(ops.size());
    for (int k = 0; k < K; ++k) sol[k]->begin(*bl, bst[k].p, xst[k].p, true);
    const cplx<float>* in[CS_MAX];
    cplx<float>* out[CS_MAX];
    for (int j = 0; j < steps; ++j) {
      for (int k = 0; k < K; ++k) {
        in[k] = sol[k]->direction(j, nullptr, s);
        out[k] = sol[k]->product();
      }
      apply_all(in, out);
      for (int k = 0; k < K; ++k) sol[k]->arnoldi(*bl, j);
    }
    for (int k = 0; k < K; ++k) sol[k]->finish(*bl, xst[k].p);
  }
  void solve(const void* const* b, void* const* x, double tol, mgc_solve_stats* stats) {
    const int K = int(ops.size());
    MGC_CHECK(b && x && stats, MGC_ERR_ARG, "null argument");
    MGC_CHECK(tol > 0.0, MGC_ERR_ARG, "tol > 0 needed");
    for (int k = 0; k < K; ++k) MGC_CHECK(b[k] && x[k], MGC_ERR_ARG, "null vector in the batch");
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < K; ++k)
      MGC_HIP(hipMemcpyAsync(bst[k].p, b[k], sol[k]->shape().n * sizeof(cplx<float>),
                             hipMemcpyDeviceToDevice, s));
    if (graph == 1) {
      run();
    } else {
      if (!ge) {
        run();  // first-use allocations and the launch tuning happen outside the capture
        MGC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          run();
        } catch (...) {
          hipGraph_t gg = nullptr;
          (void)hipStreamEndCapture(s, &gg);
          if (gg) (void)hipGraphDestroy(gg);
          throw;
        }
        MGC_HIP(hipStreamEndCapture(s, &g));
        MGC_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      }
      MGC_HIP(hipGraphLaunch(ge, s));
    }
    for (int k = 0; k < K; ++k)
      MGC_HIP(hipMemcpyAsync(x[k], xst[k].p, sol[k]->shape().n * sizeof(cplx<float>),
                             hipMemcpyDeviceToDevice, s));
    MGC_HIP(hipStreamSynchronize(s));
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // b - M x with the stencil of the solve, per rhs
    for (int k = 0; k < K; ++k) {
      const BShape& sh = sol[k]->shape();
      std::vector<double> bn(nrhs), rn(nrhs);
      bnorm2<float>(*bl, bst[k].p, sh, bn.data());
      cs_ok(mgc_coarse_op_M_mrhs(ops[k], xst[k].p, nullptr, MGC_F32, r[k].p, nrhs, mprec, 0,
                                 (void*)s));
      vsub<float>(bst[k].p, r[k].p, r[k].p, sh.n, s);
      bnorm2<float>(*bl, r[k].p, sh, rn.data());
      for (int c = 0; c < nrhs; ++c) {
        mgc_solve_stats& o = stats[k * nrhs + c];
        o.iterations = steps;
        o.true_residual = bn[c] > 0.0 ? std::sqrt(rn[c] / bn[c]) : 0.0;
        o.converged = o.true_residual <= tol ? 1 : 0;
        o.precond_applies = 0;
        o.seconds = secs;
      }
    }
  }
};

}  // namespace mgc

struct mgc_coarse_batch {
  std::unique_ptr<mgc::CoarseBatch> impl;
  std::unique_ptr<mgc::FixedCoarseBatch> fixed;
};

using namespace mgc;

static void check_batch_ops(mgc_coarse_op* const* ops, int count, int nrhs, int mrhs_prec,
                            int restart) {
  MGC_CHECK(ops, MGC_ERR_ARG, "null argument");
  MGC_CHECK(count >= 1 && count <= CS_MAX, MGC_ERR_ARG, "count must be 1..8");
  MGC_CHECK(nrhs >= 1 && nrhs <= BBlas::RMAX, MGC_ERR_ARG, "nrhs must be 1..64");
  MGC_CHECK(mrhs_prec == MGC_F32 || mrhs_prec == MGC_F16 || mrhs_prec == MGC_BF16,
            MGC_ERR_PRECISION, "multi-rhs precision must be F32, F16 or BF16");
  MGC_CHECK(restart >= 1 && restart < BBlas::KMAX, MGC_ERR_ARG, "1 <= restart < 64 needed");
  for (int k = 0; k < count; ++k) {
    MGC_CHECK(ops[k], MGC_ERR_ARG, "null lattice in the batch");
    MGC_CHECK(ops[k]->links_set, MGC_ERR_STATE, "links not set");
    MGC_CHECK(ops[k]->N == ops[0]->N, MGC_ERR_SHAPE,
              "every lattice in a batch needs the same nbasis");
    MGC_CHECK(ops[k]->N >= 8 && ops[k]->N <= 64 && ops[k]->N % 8 == 0, MGC_ERR_SHAPE,
              "multi-rhs stencil supports nbasis 8, 16, ..., 64");
    for (int j = 0; j < k; ++j)
      MGC_CHECK(ops[j] != ops[k], MGC_ERR_ARG, "a lattice appears twice in the batch");
  }
}

static void batch_solve(CoarseBatch& cb, const void* const* b, void* const* x, double tol,
                        int maxit, mgc_solve_stats* stats) {
  const int count = int(cb.lat.size()), nrhs = cb.nrhs;
  MGC_CHECK(b && x && stats, MGC_ERR_ARG, "null argument");
  MGC_CHECK(tol > 0.0 && maxit >= 1, MGC_ERR_ARG, "tol > 0 and maxit >= 1 needed");
  for (int k = 0; k < count; ++k)
    MGC_CHECK(b[k] && x[k], MGC_ERR_ARG, "null vector in the batch");
  const auto t0 = std::chrono::steady_clock::now();
  // the lattices' streams start after the caller's earlier work
  MGC_HIP(hipEventRecord(cb.start, cb.s));
  for (auto& L : cb.lat) MGC_HIP(hipStreamWaitEvent(L->s, cb.start, 0));
  cb.rv->open();
  std::vector<std::string> err(count);
  std::vector<BKrylovStats> st(count);
  std::vector<std::vector<double>> bn(count), rn(count);
  auto run = [&](int k) {
    LatticeSolve& L = *cb.lat[k];
    bool left = false;
    try {
      MGC_HIP(hipSetDevice(cb.device));
      const auto* bk = static_cast<const cplx<float>*>(b[k]);
      auto* xk = static_cast<cplx<float>*>(x[k]);
      st[k] = L.solver->solve(*L.bl, *L.A, nullptr, bk, xk, true, tol, maxit);
      cb.rv->leave();
      left = true;
      // b - M x with the stencil of the solve, per rhs
      bn[k].resize(nrhs);
      rn[k].resize(nrhs);
      bnorm2<float>(*L.bl, bk, L.A->sh, bn[k].data());
      cb.rv->serialised([&] {
        cs_ok(mgc_coarse_op_M_mrhs(L.op, xk, nullptr, MGC_F32, L.r.p, nrhs, cb.mprec, 0,
                                   (void*)L.s));
      });
      vsub<float>(bk, L.r.p, L.r.p, L.A->sh.n, L.s);
      bnorm2<float>(*L.bl, L.r.p, L.A->sh, rn[k].data());
      MGC_HIP(hipEventRecord(L.fin, L.s));
    } catch (const std::exception& e) {
      err[k] = e.what();
      if (!left) cb.rv->leave();
    }
  };
  std::vector<std::thread> th;
  for (int k = 1; k < count; ++k) th.emplace_back(run, k);
  run(0);
  for (auto& t : th) t.join();
  for (int k = 0; k < count; ++k)
    if (!err[k].empty()) throw Error(MGC_ERR_HIP, "lattice " + std::to_string(k) + ": " + err[k]);
  for (auto& L : cb.lat) MGC_HIP(hipStreamWaitEvent(cb.s, L->fin, 0));
  MGC_HIP(hipStreamSynchronize(cb.s));
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int k = 0; k < count; ++k)
    for (int r = 0; r < nrhs; ++r) {
      mgc_solve_stats& o = stats[k * nrhs + r];
      o.iterations = st[k].converged[r] ? st[k].iters_rhs[r] : st[k].iters;
      o.converged = st[k].converged[r];
      o.precond_applies = 0;
      o.true_residual = bn[k][r] > 0.0 ? std::sqrt(rn[k][r] / bn[k][r]) : 0.0;
      o.seconds = secs;
    }
}

extern "C" {

int mgc_coarse_batch_create(mgc_coarse_batch** out, mgc_coarse_op* const* ops, int count,
                            int nrhs, int mrhs_prec, int restart, void* stream) {
  return guarded([&] {
    MGC_CHECK(out, MGC_ERR_ARG, "null output handle");
    *out = nullptr;
    check_batch_ops(ops, count, nrhs, mrhs_prec, restart);
    auto h = std::make_unique<mgc_coarse_batch>();
    h->impl = std::make_unique<CoarseBatch>(ops, count, nrhs, mrhs_prec, restart,
                                            as_stream(stream));
    *out = h.release();
  });
}

int mgc_coarse_batch_create_fixed(mgc_coarse_batch** out, mgc_coarse_op* const* ops, int count,
                                  int nrhs, int mrhs_prec, int steps, int graph, void* stream) {
  return guarded([&] {
    MGC_CHECK(out, MGC_ERR_ARG, "null output handle");
    *out = nullptr;
    check_batch_ops(ops, count, nrhs, mrhs_prec, steps);
    MGC_CHECK(graph == 1 || graph == 2, MGC_ERR_ARG, "graph must be 1 or 2");
    MGC_CHECK(graph == 1 || stream != nullptr, MGC_ERR_ARG,
              "graph mode needs a created stream (the null stream cannot be captured)");
    auto h = std::make_unique<mgc_coarse_batch>();
    h->fixed = std::make_unique<FixedCoarseBatch>(ops, count, nrhs, mrhs_prec, steps, graph,
                                                  as_stream(stream));
    *out = h.release();
  });
}

int mgc_coarse_batch_solve(mgc_coarse_batch* cb, const void* const* b, void* const* x, double tol,
                           int maxit, mgc_solve_stats* stats, void* stream) {
  return guarded([&] {
    MGC_CHECK(cb && (cb->impl || cb->fixed), MGC_ERR_ARG, "null batch handle");
    if (cb->fixed) {
      MGC_CHECK(as_stream(stream) == cb->fixed->s, MGC_ERR_ARG, "use the stream given to create");
      cb->fixed->solve(b, x, tol, stats);
      return;
    }
    MGC_CHECK(as_stream(stream) == cb->impl->s, MGC_ERR_ARG, "use the stream given to create");
    batch_solve(*cb->impl, b, x, tol, maxit, stats);
  });
}

int mgc_coarse_batch_destroy(mgc_coarse_batch* cb) {
  return guarded([&] { delete cb; });
}

int mgc_coarse_solve_batched(mgc_coarse_op* const* ops, int count, const void* const* b,
                             void* const* x, int nrhs, int mrhs_prec, double tol, int maxit,
                             int restart, mgc_solve_stats* stats, void* stream) {
  return guarded([&] {
    check_batch_ops(ops, count, nrhs, mrhs_prec, restart);
    CoarseBatch cb(ops, count, nrhs, mrhs_prec, restart, as_stream(stream));
    batch_solve(cb, b, x, tol, maxit, stats);
  });
}

}  // extern "C"
