// mgc.hpp -- header-only C++ host API over the C ABI of libmgc.so.
//
// Mirrors the reference's operator interfaces so that code written against
// them reads the same:
//   CoarsenedMatrix  include/CoarsenedMatrixRework.h:38-613 (M, Mdag, Mdiag, Mdir,
//                    MdirAll, Dhop, DhopOE, DhopEO, Meooe, MeooeDag, Mooee,
//                    MooeeInv, MooeeDag, MooeeInvDag, CoarsenOperator, ConstEE)
//   Aggregation      include/AggregationRework.h:33-147 (ProjectToSubspace,
//                    PromoteFromSubspace, Orthogonalise, CheckOrthogonal, Create)
//   MultiGridPreconditioner  include/MultigridPreconditioner.h:33-42, 137-309
//                    (initialSetup, operator())
//   NonHermitianVCyclePreconditioner include/VCycleRework.h:34-189
// Fields are device buffers owned by the caller (DeviceField); every call is
// asynchronous on the handle's stream.  Errors (the reference asserts) throw
// mgc::Exception with the library's status code and message.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/mgc.h"

namespace mgc {

struct Exception : std::runtime_error {
  int code;
  Exception(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int status) {
  if (status != MGC_OK) throw Exception(status, mgc_last_error());
}

struct Coordinate {
  int v[4];
  int volume() const { return v[0] * v[1] * v[2] * v[3]; }
};

// a device buffer of `count` complex numbers of precision `prec`
class DeviceField {
 public:
  DeviceField() = default;
  DeviceField(size_t count, int prec) : count_(count), prec_(prec) {
    check(mgc_malloc(&ptr_, count * bytes_per_complex(prec)));
  }
  ~DeviceField() {
    if (ptr_) (void)mgc_free(ptr_);
  }
  DeviceField(DeviceField&& o) noexcept { *this = std::move(o); }
  DeviceField& operator=(DeviceField&& o) noexcept {
    std::swap(ptr_, o.ptr_);
    std::swap(count_, o.count_);
    std::swap(prec_, o.prec_);
    return *this;
  }
  DeviceField(const DeviceField&) = delete;
  DeviceField& operator=(const DeviceField&) = delete;

  void* data() { return ptr_; }
  const void* data() const { return ptr_; }
  size_t count() const { return count_; }
  int precision() const { return prec_; }
  size_t bytes() const { return count_ * bytes_per_complex(prec_); }
  void upload(const void* host, void* stream = nullptr) {
    check(mgc_memcpy_h2d(ptr_, host, bytes(), stream));
  }
  void download(void* host, void* stream = nullptr) const {
    check(mgc_memcpy_d2h(host, ptr_, bytes(), stream));
  }
  void zero(void* stream = nullptr) { check(mgc_fill_zero(ptr_, bytes(), stream)); }
  void gaussian(uint64_t seed, uint64_t stream_id, void* stream = nullptr) {
    check(mgc_random_gaussian(ptr_, count_, prec_, seed, stream_id, stream));
  }
  static size_t bytes_per_complex(int prec) {
    return prec == MGC_F64 ? 16 : prec == MGC_F32 ? 8 : 4;
  }

 private:
  void* ptr_ = nullptr;
  size_t count_ = 0;
  int prec_ = MGC_F64;
};

class WilsonCloverOperator {
 public:
  WilsonCloverOperator(Coordinate dims, double mass, double csw, bool antiperiodic_t = false,
                       void* stream = nullptr)
      : dims_(dims), stream_(stream) {
    check(mgc_fine_op_create(&h_, dims.v, mass, csw, antiperiodic_t ? 1 : 0));
  }
  ~WilsonCloverOperator() {
    if (h_) (void)mgc_fine_op_destroy(h_);
  }
  WilsonCloverOperator(const WilsonCloverOperator&) = delete;
  WilsonCloverOperator& operator=(const WilsonCloverOperator&) = delete;

  void ImportGauge(const DeviceField& U) { check(mgc_fine_op_set_gauge(h_, U.data(), stream_)); }
  void M(const DeviceField& in, DeviceField& out) {
    check(mgc_fine_op_M(h_, in.data(), out.data(), in.precision(), 1, stream_));
  }
  void Mdag(const DeviceField& in, DeviceField& out) {
    check(mgc_fine_op_Mdag(h_, in.data(), out.data(), in.precision(), 1, stream_));
  }
  void OpDir(const DeviceField& in, DeviceField& out, int dir, int disp) {
    check(mgc_fine_op_Mdir(h_, in.data(), out.data(), dir, disp, in.precision(), 1, stream_));
  }
  void OpDiag(const DeviceField& in, DeviceField& out) {
    check(mgc_fine_op_Mdiag(h_, in.data(), out.data(), in.precision(), 1, stream_));
  }
  mgc_fine_op* handle() { return h_; }
  Coordinate dims() const { return dims_; }

 private:
  mgc_fine_op* h_ = nullptr;
  Coordinate dims_;
  void* stream_;
};

class Aggregation {
 public:
  // Aggregation(CoarseGrid, FineGrid, cb): here the fine lattice and the block
  Aggregation(Coordinate fine, Coordinate block, int nb, int site_elems = 12,
              void* stream = nullptr)
      : nb_(nb), ef_(site_elems), fine_(fine), stream_(stream) {
    check(mgc_aggregation_create_ex(&h_, fine.v, block.v, nb, site_elems));
    for (int d = 0; d < 4; ++d) coarse_.v[d] = fine.v[d] / block.v[d];
  }
  ~Aggregation() {
    if (h_) (void)mgc_aggregation_destroy(h_);
  }
  Aggregation(const Aggregation&) = delete;
  Aggregation& operator=(const Aggregation&) = delete;

  // Create(MGBasisVectors): copy the nb fine fields [nb][site][Ef] then orthogonalise
  void Create(const DeviceField& basis, int passes = 2) {
    check(mgc_aggregation_set_basis(h_, basis.data(), stream_));
    Orthogonalise(0, passes);
  }
  void SetSubspace(const DeviceField& basis) {
    check(mgc_aggregation_set_basis(h_, basis.data(), stream_));
  }
  void GetSubspace(DeviceField& basis) const {
    check(mgc_aggregation_get_basis(h_, basis.data(), stream_));
  }
  void Orthogonalise(int checkOrthogonality = 1, int passes = 2) {
    check(mgc_aggregation_orthogonalise(h_, passes, stream_));
    if (checkOrthogonality) CheckOrthogonal();
  }
  std::vector<double> CheckOrthogonal() {
    std::vector<double> errs(nb_);
    check(mgc_aggregation_check_orthogonal(h_, errs.data(), stream_));
    return errs;
  }
  void ProjectToSubspace(DeviceField& coarse, const DeviceField& fine) const {
    check(mgc_aggregation_project(h_, fine.data(), coarse.data(), fine.precision(), 1, stream_));
  }
  void PromoteFromSubspace(const DeviceField& coarse, DeviceField& fine) const {
    check(mgc_aggregation_promote(h_, coarse.data(), fine.data(), coarse.precision(), 1,
                                  stream_));
  }
  int nbasis() const { return 2 * nb_; }
  Coordinate CoarseGrid() const { return coarse_; }
  Coordinate FineGrid() const { return fine_; }
  mgc_aggregation* handle() { return h_; }

 private:
  mgc_aggregation* h_ = nullptr;
  int nb_, ef_;
  Coordinate fine_, coarse_;
  void* stream_;
};

class CoarsenedMatrix {
 public:
  CoarsenedMatrix(Coordinate dims, int nbasis, int prec = MGC_F64, int hermitian = 0,
                  void* stream = nullptr)
      : dims_(dims), n_(nbasis), prec_(prec), stream_(stream) {
    check(mgc_coarse_op_create(&h_, dims.v, nbasis, prec, hermitian));
  }
  ~CoarsenedMatrix() {
    if (h_) (void)mgc_coarse_op_destroy(h_);
  }
  CoarsenedMatrix(const CoarsenedMatrix&) = delete;
  CoarsenedMatrix& operator=(const CoarsenedMatrix&) = delete;

  // the public link members A[p] / Y_[p]: FLAT [9][site][N][N] complex double
  void SetLinks(const DeviceField& links, int layout = MGC_LAYOUT_FLAT, bool invert = true) {
    check(mgc_coarse_op_set_links(h_, layout, links.data(), invert ? 1 : 0, stream_));
  }
  void GetLinks(DeviceField& links, int layout = MGC_LAYOUT_FLAT) const {
    check(mgc_coarse_op_get_links(h_, layout, links.data(), stream_));
  }
  int ConstEE() const { return 0; }
  void M(const DeviceField& in, DeviceField& out) { check(mgc_coarse_op_M(h_, in.data(), out.data(), stream_)); }
  void Mdag(const DeviceField& in, DeviceField& out) { check(mgc_coarse_op_Mdag(h_, in.data(), out.data(), stream_)); }
  void Mdiag(const DeviceField& in, DeviceField& out) { check(mgc_coarse_op_Mdiag(h_, in.data(), out.data(), stream_)); }
  void Mdir(const DeviceField& in, DeviceField& out, int dir, int disp) {
    check(mgc_coarse_op_Mdir(h_, in.data(), out.data(), dir, disp, stream_));
  }
  // out holds 8 consecutive coarse vectors
  void MdirAll(const DeviceField& in, DeviceField& out8) {
    check(mgc_coarse_op_MdirAll(h_, in.data(), out8.data(), stream_));
  }
  void Dhop(const DeviceField& in, DeviceField& out, int dag) {
    check(mgc_coarse_op_Dhop(h_, in.data(), out.data(), dag, stream_));
  }
  void DhopOE(const DeviceField& in, DeviceField& out, int dag) {
    check(mgc_coarse_op_DhopOE(h_, in.data(), out.data(), dag, stream_));
  }
  void DhopEO(const DeviceField& in, DeviceField& out, int dag) {
    check(mgc_coarse_op_DhopEO(h_, in.data(), out.data(), dag, stream_));
  }
  void Meooe(const DeviceField& in, DeviceField& out, int in_cb) {
    check(mgc_coarse_op_Meooe(h_, in.data(), out.data(), in_cb, stream_));
  }
  void MeooeDag(const DeviceField& in, DeviceField& out, int in_cb) {
    check(mgc_coarse_op_MeooeDag(h_, in.data(), out.data(), in_cb, stream_));
  }
  void Mooee(const DeviceField& in, DeviceField& out, int cb = MGC_FULL) {
    check(mgc_coarse_op_Mooee(h_, in.data(), out.data(), cb, stream_));
  }
  void MooeeInv(const DeviceField& in, DeviceField& out, int cb = MGC_FULL) {
    check(mgc_coarse_op_MooeeInv(h_, in.data(), out.data(), cb, stream_));
  }
  void MooeeDag(const DeviceField& in, DeviceField& out, int cb = MGC_FULL) {
    check(mgc_coarse_op_MooeeDag(h_, in.data(), out.data(), cb, stream_));
  }
  void MooeeInvDag(const DeviceField& in, DeviceField& out, int cb = MGC_FULL) {
    check(mgc_coarse_op_MooeeInvDag(h_, in.data(), out.data(), cb, stream_));
  }
  // multi-rhs M: in/out [site][nrhs][N], FP32 complex in and out
  void M_mrhs(const DeviceField& in, DeviceField& out, int nrhs, int mrhs_prec = MGC_F16,
              int dag = 0) {
    check(mgc_coarse_op_M_mrhs(h_, in.data(), nullptr, MGC_F32, out.data(), nrhs, mrhs_prec, dag,
                               stream_));
  }
  void CoarsenOperator(WilsonCloverOperator& linop, Aggregation& agg,
                       int mode = MGC_COARSEN_DIRSAVE) {
    check(mgc_coarsen_operator(h_, linop.handle(), agg.handle(), mode, stream_));
  }
  void CoarsenOperator(CoarsenedMatrix& linop, Aggregation& agg, int mode = MGC_COARSEN_DIRSAVE) {
    check(mgc_coarsen_coarse_operator(h_, linop.h_, agg.handle(), mode, stream_));
  }
  void ConstructRemainingLinks(int disp_have = -1) {
    check(mgc_coarse_op_construct_remaining_links(h_, disp_have, stream_));
  }
  Coordinate Grid() const { return dims_; }
  int nbasis() const { return n_; }
  mgc_coarse_op* handle() { return h_; }

 private:
  mgc_coarse_op* h_ = nullptr;
  Coordinate dims_;
  int n_, prec_;
  void* stream_;
};

// multi-rhs M of several coarse lattices (one nbasis) in one launch; inputs
// already in mrhs_prec (scales required for F16/BF16, from mgc_quantize)
inline void M_mrhs_batched(const std::vector<CoarsenedMatrix*>& ops,
                           const std::vector<const void*>& in,
                           const std::vector<const float*>& in_scales,
                           const std::vector<void*>& out, int nrhs, int mrhs_prec, int dag = 0,
                           void* stream = nullptr) {
  std::vector<mgc_coarse_op*> h;
  for (CoarsenedMatrix* op : ops) h.push_back(op->handle());
  check(mgc_coarse_op_M_mrhs_batched(h.data(), (int)h.size(), in.data(),
                                     in_scales.empty() ? nullptr : in_scales.data(), mrhs_prec,
                                     out.data(), nrhs, mrhs_prec, dag, stream));
}

// Coarse solves of several lattices of one nbasis in one lockstep batch
// (mgc_coarse_batch_*, no counterpart in the reference, which solves one
// lattice at a time): GMRES(restart) to a tolerance, or with steps > 0 a
// fixed number of GMRES steps, captured into a HIP graph (graph = 2, a
// created stream) or issued eagerly (graph = 1).  Every b and x is
// [site][nrhs][N] complex FP32; the operators must outlive the batch.
class CoarseSolveBatch {
 public:
  CoarseSolveBatch(const std::vector<CoarsenedMatrix*>& ops, int nrhs, int mrhs_prec = MGC_F32,
                   int restart = 20, void* stream = nullptr, int steps = 0, int graph = 2)
      : nrhs_(nrhs), count_((int)ops.size()), stream_(stream) {
    std::vector<mgc_coarse_op*> h;
    for (CoarsenedMatrix* op : ops) h.push_back(op->handle());
    if (steps > 0)
      check(mgc_coarse_batch_create_fixed(&h_, h.data(), count_, nrhs, mrhs_prec, steps, graph,
                                          stream));
    else
      check(mgc_coarse_batch_create(&h_, h.data(), count_, nrhs, mrhs_prec, restart, stream));
  }
  ~CoarseSolveBatch() {
    if (h_) (void)mgc_coarse_batch_destroy(h_);
  }
  CoarseSolveBatch(const CoarseSolveBatch&) = delete;
  CoarseSolveBatch& operator=(const CoarseSolveBatch&) = delete;

  // one entry per (lattice, rhs), lattice-major
  std::vector<mgc_solve_stats> Solve(const std::vector<const DeviceField*>& b,
                                     const std::vector<DeviceField*>& x, double tol,
                                     int maxit = 300) {
    if ((int)b.size() != count_ || (int)x.size() != count_)
      throw Exception(MGC_ERR_ARG, "one b and one x per lattice");
    std::vector<const void*> bp;
    std::vector<void*> xp;
    for (int k = 0; k < count_; ++k) {
      bp.push_back(b[k]->data());
      xp.push_back(x[k]->data());
    }
    std::vector<mgc_solve_stats> st((size_t)count_ * nrhs_);
    check(mgc_coarse_batch_solve(h_, bp.data(), xp.data(), tol, maxit, st.data(), stream_));
    return st;
  }

 private:
  mgc_coarse_batch* h_ = nullptr;
  int nrhs_, count_;
  void* stream_;
};

class MultiGridPreconditioner {
 public:
  MultiGridPreconditioner(WilsonCloverOperator& fine, const mgc_mg_params& params,
                          void* stream = nullptr)
      : stream_(stream) {
    check(mgc_mg_create(&h_, fine.handle(), &params));
  }
  ~MultiGridPreconditioner() {
    if (h_) (void)mgc_mg_destroy(h_);
  }
  MultiGridPreconditioner(const MultiGridPreconditioner&) = delete;
  MultiGridPreconditioner& operator=(const MultiGridPreconditioner&) = delete;

  void initialSetup() { check(mgc_mg_setup(h_, stream_)); }
  void operator()(const DeviceField& in, DeviceField& out) {
    check(mgc_mg_precondition(h_, in.data(), out.data(), stream_));
  }
  // NonHermitianVCyclePreconditioner::operator() with maxIter / tolerance
  void VCycle(const DeviceField& in, DeviceField& out, int maxIter, double tolerance = 0.0) {
    check(mgc_mg_vcycle(h_, in.data(), out.data(), maxIter, tolerance, stream_));
  }
  // FGMRES(D) x = b preconditioned by this hierarchy
  mgc_solve_stats Solve(const DeviceField& b, DeviceField& x, double tol, int maxit = 1000,
                        int restart = 20) {
    mgc_solve_stats st{};
    check(mgc_mg_solve(h_, b.data(), x.data(), tol, maxit, restart, &st, stream_));
    return st;
  }
  // the same for nrhs right-hand sides in lockstep, b and x [site][rhs][12];
  // coarse levels on the multi-rhs MFMA stencil in coarse_prec
  std::vector<mgc_solve_stats> SolveMRHS(const DeviceField& b, DeviceField& x, int nrhs,
                                         double tol, int coarse_prec = MGC_F16,
                                         int maxit = 1000, int restart = 20) {
    std::vector<mgc_solve_stats> st((size_t)nrhs);
    check(mgc_mg_solve_mrhs(h_, b.data(), x.data(), nrhs, coarse_prec, tol, maxit, restart,
                            st.data(), stream_));
    return st;
  }

 private:
  mgc_mg* h_ = nullptr;
  void* stream_;
};

}  // namespace mgc
