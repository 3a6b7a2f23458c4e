// Device-scalar (F)GMRES for the K-cycle in graph mode: a fixed number m of
// Arnoldi steps from one starting residual, with every scalar (norms,
// Hessenberg column, complex Givens rotations, the triangular solve) kept in
// device memory and updated by single-workgroup kernels.  It makes no host
// synchronisation and no allocation after construction, so a preconditioner
// built from it can be captured once into a HIP graph and replayed.
//
// Arithmetic per step is that of the host-driven Gmres: classical
// Gram-Schmidt with the second projection taken when one projection removed
// more than 1 - eta of |w|^2, the new norm by Pythagoras.  The test runs on
// the device (k_dg_reorth_test) and the second projection's kernels read its
// flag, so the decision is replayed with the graph; without the fused pass
// (more than 24 basis vectors) the second projection is always taken.
// A step whose new norm is not positive marks the Krylov space exhausted;
// the solution then uses the columns before it.
#pragma once
#include <complex>

#include "bkrylov.hpp"
#include "krylov.hpp"

namespace mgc {

// device state of one solve
struct DGmresState {
  cplx<double> H[(Blas::KMAX + 1) * Blas::KMAX];  // [i][j], row stride KMAX
  cplx<double> sn[Blas::KMAX];
  double cs[Blas::KMAX];
  cplx<double> g[Blas::KMAX + 1];
  cplx<double> negy[Blas::KMAX];  // -y, for x += sum y_i Z_i by the multi-subtract
  double r2[2];                   // |r0|^2 (from vnorm2_dev)
  int jmax;                       // Arnoldi columns in use
  int skip;                       // 1: this step's second projection is not taken
};

// Re-orthogonalisation test of step j on the device: with h1 at dh1[0..j]
// and |w|^2 at dh1[j+1].re, a second projection is needed when
// |w|^2 - |h1|^2 < eta |w|^2.  Without it, dh2 takes h1 and |w|^2 (so that
// vsubnorm_dev(dh2) forms v = (w - V h1) / |w'|) and dh1 is zeroed (so that
// the Hessenberg column dh1 + dh2 is h1); the skipped pass leaves dh2 alone.
static __global__ void k_dg_reorth_test(DGmresState* st, cplx<double>* dh1, cplx<double>* dh2, int j,
                                 double eta) {
  if (threadIdx.x != 0) return;
  double sum1 = 0.0;
  for (int i = 0; i <= j; ++i) sum1 += dh1[i].re * dh1[i].re + dh1[i].im * dh1[i].im;
  const double wn2 = dh1[j + 1].re;
  const int skip = wn2 - sum1 >= eta * wn2 ? 1 : 0;
  st->skip = skip;
  if (skip) {
    for (int i = 0; i <= j; ++i) {
      dh2[i] = dh1[i];
      dh1[i] = {0.0, 0.0};
    }
    dh2[j + 1] = dh1[j + 1];
  }
}

static __global__ void k_dg_begin(DGmresState* st) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i <= Blas::KMAX; ++i) st->g[i] = {0.0, 0.0};
  const double b2 = st->r2[0];
  st->g[0] = {b2 > 0.0 ? sqrt(b2) : 0.0, 0.0};
  st->jmax = b2 > 0.0 ? Blas::KMAX : 0;  // no Krylov space for a zero residual
}

// column j of the Hessenberg matrix from the two projections (h1 at dh1[0..j],
// h2 at dh2[0..j], the new norm at dh2[j+1].im), the earlier rotations, and
// the rotation that zeroes H[j+1][j]
static __global__ void k_dg_step(DGmresState* st, const cplx<double>* dh1, const cplx<double>* dh2,
                          int j) {
  if (threadIdx.x != 0) return;
  typedef std::complex<double> C;
  if (j >= st->jmax) return;
  const int ld = Blas::KMAX;
  C col[Blas::KMAX + 1];
  for (int i = 0; i <= j; ++i) col[i] = C(dh1[i].re + dh2[i].re, dh1[i].im + dh2[i].im);
  const double hn = dh2[j + 1].im;
  for (int i = 0; i < j; ++i) {
    const C a = col[i], b = col[i + 1];
    const C s(st->sn[i].re, st->sn[i].im);
    col[i] = st->cs[i] * a + s * b;
    col[i + 1] = -std::conj(s) * a + st->cs[i] * b;
  }
  const C a = col[j];
  const double bb = hn > 0.0 ? hn : 0.0;
  const double t = sqrt(std::norm(a) + bb * bb);
  double c;
  C s;
  if (std::abs(a) == 0.0) {
    c = 0.0;
    s = 1.0;
  } else {
    c = std::abs(a) / t;
    s = (a / std::abs(a)) * (bb / t);
  }
  col[j] = c * a + s * bb;
  st->cs[j] = c;
  st->sn[j] = {s.real(), s.imag()};
  const C gj(st->g[j].re, st->g[j].im);
  const C g1 = -std::conj(s) * gj, g0 = c * gj;
  st->g[j] = {g0.real(), g0.imag()};
  st->g[j + 1] = {g1.real(), g1.imag()};
  for (int i = 0; i <= j; ++i) st->H[i * ld + j] = {col[i].real(), col[i].imag()};
  if (!(hn > 0.0)) st->jmax = j + 1;  // the next vector is zero: stop using columns
}

// y = H^-1 g over the columns in use (m at most), stored negated
static __global__ void k_dg_solve(DGmresState* st, int m) {
  if (threadIdx.x != 0) return;
  typedef std::complex<double> C;
  const int ld = Blas::KMAX;
  const int jm = st->jmax < m ? st->jmax : m;
  C y[Blas::KMAX];
  for (int i = jm - 1; i >= 0; --i) {
    C acc(st->g[i].re, st->g[i].im);
    for (int k = i + 1; k < jm; ++k) acc -= C(st->H[i * ld + k].re, st->H[i * ld + k].im) * y[k];
    const C d(st->H[i * ld + i].re, st->H[i * ld + i].im);
    y[i] = std::abs(d) > 0.0 ? acc / d : C(0.0);
  }
  for (int i = 0; i < m; ++i) st->negy[i] = i < jm ? cplx<double>{-y[i].real(), -y[i].imag()}
                                                  : cplx<double>{0.0, 0.0};
}

template <typename R>
class DGmres : public Solver<R> {
 public:
  DGmres(size_t n, int m, bool flexible) : n_(n), m_(m), flexible_(flexible) {
    MGC_CHECK(m >= 1 && m + 1 <= Blas::KMAX, MGC_ERR_ARG, "Arnoldi steps out of range");
    V_.alloc((size_t)(m + 1) * n);
    if (flexible) Z_.alloc((size_t)m * n);
    w_.alloc(n);
    dh_.alloc((size_t)2 * (Blas::KMAX + 2));
    st_.alloc(1);
  }

  // m steps from x (or 0); tol and maxit are not used: the step count is fixed
  KrylovStats solve(Blas& bl, LinOp<R>& A, Prec<R>* P, const cplx<R>* b, cplx<R>* x, bool x_zero,
                    double, int) override {
    hipStream_t s = bl.s;
    const size_t n = n_;
    cplx<R>* V = V_.p;
    KrylovStats st;
    if (x_zero) {
      vzero<R>(x, n, s);
      vcopy<R>(b, V, n, s);
    } else {
      A.apply(x, w_.p, s);
      ++st.applies;
      vsub<R>(b, w_.p, V, n, s);
    }
    vnorm2_dev<R>(bl, V, nullptr, n, st_.p->r2);
    k_dg_begin<<<1, 64, 0, s>>>(st_.p);
    MGC_LAUNCH_CHECK();
    vconvert_dev<R, R>(V, V, nullptr, nullptr, n, st_.p->r2, 1, s);  // v0 = r / |r|
    cplx<double>* dh1 = dh_.p;
    for (int j = 0; j < m_; ++j) {
      cplx<R>* vj = V + (size_t)j * n;
      const cplx<R>* z = vj;
      if (flexible_) {
        cplx<R>* zj = Z_.p + (size_t)j * n;
        if (P) P->apply(vj, zj, s);
        else vcopy<R>(vj, zj, n, s);
        z = zj;
      }
      A.apply(z, w_.p, s);
      ++st.applies;
      cplx<double>* dh2 = dh_.p + (j + 2);
      vmultidot_dev<R>(bl, V, n, j + 1, w_.p, n, true, dh1);
      if (j + 1 <= 24) {
        k_dg_reorth_test<<<1, 64, 0, s>>>(st_.p, dh1, dh2, j, eta_);
        MGC_LAUNCH_CHECK();
        vmultisub_dot_dev<R>(bl, dh1, V, n, j + 1, w_.p, n, dh2, &st_.p->skip);
      } else {
        vmultisub_dev<R>(dh1, V, n, j + 1, w_.p, n, s);
        vmultidot_dev<R>(bl, V, n, j + 1, w_.p, n, true, dh2);
      }
      vsubnorm_dev<R>(dh2, V, n, j + 1, w_.p, V + (size_t)(j + 1) * n, n, s);
      k_dg_step<<<1, 64, 0, s>>>(st_.p, dh1, dh2, j);
      MGC_LAUNCH_CHECK();
      ++st.iters;
    }
    k_dg_solve<<<1, 64, 0, s>>>(st_.p, m_);
    MGC_LAUNCH_CHECK();
    vmultisub_dev<R>(st_.p->negy, flexible_ ? Z_.p : V, n, m_, x, n, s);  // x += sum y_i Z_i
    return st;
  }

  // re-orthogonalisation threshold eta, as Gmres::set_reorth_eta
  void set_reorth_eta(double eta) {
    MGC_CHECK(eta > 0.0 && eta <= 1.0, MGC_ERR_ARG, "re-orthogonalisation threshold in (0, 1]");
    eta_ = eta;
  }

 private:
  size_t n_;
  int m_;
  bool flexible_;
  double eta_ = 0.5;
  DevBuf<cplx<R>> V_, Z_, w_;
  DevBuf<cplx<double>> dh_;
  DevBuf<DGmresState> st_;
};

// ---------------------------------------------------------------------------
// Batched device-scalar (F)GMRES: nrhs systems on [site][rhs][dof] vectors,
// a fixed number m of Arnoldi steps each, the per-rhs scalars (Hessenberg
// column, Givens rotations, triangular solve) in device memory and updated
// by single-workgroup kernels with one thread per rhs, so that a batched
// K-cycle built from it can be captured into a HIP graph.  Per rhs the
// arithmetic of DGmres; the second projection of a step runs for the whole
// batch when some rhs still in use needs it (as BGmres decides it on the
// host), decided on the device.
// ---------------------------------------------------------------------------

// g = (|r0|, 0, ...), and no Krylov space for a zero residual
static __global__ void k_dbg_begin(cplx<double>* g, int* jmax, const cplx<double>* r2, int nr, int m) {
  const int r = threadIdx.x;
  if (r >= nr) return;
  for (int i = 0; i <= m; ++i) g[(size_t)r * (m + 1) + i] = {0.0, 0.0};
  const double b2 = r2[r].re;
  g[(size_t)r * (m + 1)] = {b2 > 0.0 ? sqrt(b2) : 0.0, 0.0};
  jmax[r] = b2 > 0.0 ? m : 0;
}

// re-orthogonalisation test of step j for the batch (h1 at dh1[i][r],
// |w|^2_r at dh1[j+1][r].re): skip the second projection unless some rhs in
// use has |w|^2 - |h1|^2 < eta |w|^2; when skipped, dh2 takes h1 and |w|^2
// and dh1 is zeroed, as k_dg_reorth_test does
static __global__ void k_dbg_reorth_test(cplx<double>* dh1, cplx<double>* dh2, const int* jmax, int j,
                                  double eta, int nr, int* skip) {
  const int r = threadIdx.x;
  int need = 0;
  if (r < nr && j < jmax[r]) {
    double sum1 = 0.0;
    for (int i = 0; i <= j; ++i) {
      const cplx<double> h = dh1[(size_t)i * nr + r];
      sum1 += h.re * h.re + h.im * h.im;
    }
    const double wn2 = dh1[(size_t)(j + 1) * nr + r].re;
    need = wn2 - sum1 >= eta * wn2 ? 0 : 1;
  }
  const int any = __syncthreads_or(need);
  if (threadIdx.x == 0) *skip = any ? 0 : 1;
  if (!any && r < nr) {
    for (int i = 0; i <= j; ++i) {
      dh2[(size_t)i * nr + r] = dh1[(size_t)i * nr + r];
      dh1[(size_t)i * nr + r] = {0.0, 0.0};
    }
    dh2[(size_t)(j + 1) * nr + r] = dh1[(size_t)(j + 1) * nr + r];
  }
}

// column j of each rhs's Hessenberg matrix (h1 + h2, the new norm at
// dh2[j+1][r].im), the earlier rotations and the new one, as k_dg_step
static __global__ void k_dbg_step(cplx<double>* H, double* cs, cplx<double>* sn, cplx<double>* g,
                           int* jmax, const cplx<double>* dh1, const cplx<double>* dh2, int j,
                           int nr, int m) {
  typedef std::complex<double> C;
  const int r = threadIdx.x;
  if (r >= nr || j >= jmax[r]) return;
  C col[Blas::KMAX + 1];
  for (int i = 0; i <= j; ++i) {
    const cplx<double> a = dh1[(size_t)i * nr + r], b = dh2[(size_t)i * nr + r];
    col[i] = C(a.re + b.re, a.im + b.im);
  }
  const double hn = dh2[(size_t)(j + 1) * nr + r].im;
  double* csr = cs + (size_t)r * m;
  cplx<double>* snr = sn + (size_t)r * m;
  for (int i = 0; i < j; ++i) {
    const C a = col[i], b = col[i + 1];
    const C sv(snr[i].re, snr[i].im);
    col[i] = csr[i] * a + sv * b;
    col[i + 1] = -std::conj(sv) * a + csr[i] * b;
  }
  const C a = col[j];
  const double bb = hn > 0.0 ? hn : 0.0;
  const double t = sqrt(std::norm(a) + bb * bb);
  double c;
  C sv;
  if (std::abs(a) == 0.0) {
    c = 0.0;
    sv = 1.0;
  } else {
    c = std::abs(a) / t;
    sv = (a / std::abs(a)) * (bb / t);
  }
  col[j] = c * a + sv * bb;
  csr[j] = c;
  snr[j] = {sv.real(), sv.imag()};
  cplx<double>* gr = g + (size_t)r * (m + 1);
  const C gj(gr[j].re, gr[j].im);
  const C g1 = -std::conj(sv) * gj, g0 = c * gj;
  gr[j] = {g0.real(), g0.imag()};
  gr[j + 1] = {g1.real(), g1.imag()};
  cplx<double>* Hr = H + (size_t)r * (m + 1) * m;
  for (int i = 0; i <= j; ++i) Hr[(size_t)i * m + j] = {col[i].real(), col[i].imag()};
  if (!(hn > 0.0)) jmax[r] = j + 1;  // the next vector is zero: stop using columns
}

// per rhs y = H^-1 g over its columns in use, stored negated as negy[i][r]
static __global__ void k_dbg_solve(const cplx<double>* H, const cplx<double>* g, const int* jmax,
                            cplx<double>* negy, int nr, int m) {
  typedef std::complex<double> C;
  const int r = threadIdx.x;
  if (r >= nr) return;
  const int jm = jmax[r] < m ? jmax[r] : m;
  const cplx<double>* Hr = H + (size_t)r * (m + 1) * m;
  const cplx<double>* gr = g + (size_t)r * (m + 1);
  C y[Blas::KMAX];
  for (int i = jm - 1; i >= 0; --i) {
    C acc(gr[i].re, gr[i].im);
    for (int k = i + 1; k < jm; ++k)
      acc -= C(Hr[(size_t)i * m + k].re, Hr[(size_t)i * m + k].im) * y[k];
    const C d(Hr[(size_t)i * m + i].re, Hr[(size_t)i * m + i].im);
    y[i] = std::abs(d) > 0.0 ? acc / d : C(0.0);
  }
  for (int i = 0; i < m; ++i)
    negy[(size_t)i * nr + r] = i < jm ? cplx<double>{-y[i].real(), -y[i].imag()}
                                      : cplx<double>{0.0, 0.0};
}

template <typename R>
class DBGmres : public BSolver<R> {
 public:
  DBGmres(const BShape& sh, int m, bool flexible) : sh_(sh), m_(m), flexible_(flexible) {
    MGC_CHECK(m >= 1 && m + 1 <= Blas::KMAX, MGC_ERR_ARG, "Arnoldi steps out of range");
    MGC_CHECK(sh.nrhs >= 1 && sh.nrhs <= 64, MGC_ERR_ARG, "batch of 1..64 rhs");
    const int nr = sh.nrhs;
    V_.alloc((size_t)(m + 1) * sh.n);
    if (flexible) Z_.alloc((size_t)m * sh.n);
    w_.alloc(sh.n);
    dh_.alloc((size_t)(2 * (m + 1) + 2) * nr);
    H_.alloc((size_t)nr * (m + 1) * m);
    sn_.alloc((size_t)nr * m);
    cs_.alloc((size_t)nr * m);
    g_.alloc((size_t)nr * (m + 1));
    negy_.alloc((size_t)m * nr);
    r2_.alloc(nr);
    jmax_.alloc(nr);
    skip_.alloc(1);
  }
  // re-orthogonalisation threshold eta, as Gmres::set_reorth_eta
  void set_reorth_eta(double eta) {
    MGC_CHECK(eta > 0.0 && eta <= 1.0, MGC_ERR_ARG, "re-orthogonalisation threshold in (0, 1]");
    eta_ = eta;
  }

  // m steps from x (or 0) for every rhs; tol and maxit are not used
  BKrylovStats solve(BBlas& bl, BLinOp<R>& A, BPrec<R>* P, const cplx<R>* b, cplx<R>* x,
                     bool x_zero, double, int) override {
    const int nr = sh_.nrhs;
    BKrylovStats st;
    st.rel_res.assign(nr, 0.0);
    st.converged.assign(nr, 0);
    st.iters_rhs.assign(nr, m_);
    if (!x_zero) {
      A.apply(x, w_.p, bl.s);
      ++st.applies;
    }
    begin(bl, b, x, x_zero);
    for (int j = 0; j < m_; ++j) {
      A.apply(direction(j, P, bl.s), w_.p, bl.s);
      ++st.applies;
      arnoldi(bl, j);
      ++st.iters;
    }
    finish(bl, x);
    return st;
  }

  // The same solve in phases, for a driver that applies the operators of
  // several solvers in one launch: begin (with w = A x already formed unless
  // x_zero), then per step j the operator applied to direction(j) into
  // product() and arnoldi(j), then finish.
  void begin(BBlas& bl, const cplx<R>* b, cplx<R>* x, bool x_zero) {
    hipStream_t s = bl.s;
    const size_t n = sh_.n;
    cplx<R>* V = V_.p;
    if (x_zero) {
      vzero<R>(x, n, s);
      vcopy<R>(b, V, n, s);
    } else {
      vsub<R>(b, w_.p, V, n, s);
    }
    bdots_dev<R>(bl, V, n, 0, V, sh_, true, r2_.p);  // |r0|^2 per rhs
    k_dbg_begin<<<1, 64, 0, s>>>(g_.p, jmax_.p, r2_.p, sh_.nrhs, m_);
    MGC_LAUNCH_CHECK();
    bscale_dev<R>(V, sh_, r2_.p, s);  // v0 = r0 / |r0|
  }
  const cplx<R>* direction(int j, BPrec<R>* P, hipStream_t s) {
    cplx<R>* vj = V_.p + (size_t)j * sh_.n;
    if (!flexible_) return vj;
    cplx<R>* zj = Z_.p + (size_t)j * sh_.n;
    if (P) P->apply(vj, zj, s);
    else vcopy<R>(vj, zj, sh_.n, s);
    return zj;
  }
  cplx<R>* product() { return w_.p; }
  void arnoldi(BBlas& bl, int j) {
    hipStream_t s = bl.s;
    const size_t n = sh_.n;
    const int nr = sh_.nrhs;
    cplx<R>* V = V_.p;
    cplx<double>* dh1 = dh_.p;
    cplx<double>* dh2 = dh_.p + (size_t)(j + 2) * nr;
    bdots_dev<R>(bl, V, n, j + 1, w_.p, sh_, true, dh1);
    if (j + 1 <= 24) {
      k_dbg_reorth_test<<<1, 64, 0, s>>>(dh1, dh2, jmax_.p, j, eta_, nr, skip_.p);
      MGC_LAUNCH_CHECK();
      bmultisub_dots_dev<R>(bl, dh1, V, n, j + 1, w_.p, sh_, dh2, skip_.p);
    } else {
      bmultisub_dev<R>(dh1, V, n, j + 1, w_.p, sh_, s);
      bdots_dev<R>(bl, V, n, j + 1, w_.p, sh_, true, dh2);
    }
    bsubnorm_dev<R>(dh2, V, n, j + 1, w_.p, V + (size_t)(j + 1) * n, sh_, s);
    k_dbg_step<<<1, 64, 0, s>>>(H_.p, cs_.p, sn_.p, g_.p, jmax_.p, dh1, dh2, j, nr, m_);
    MGC_LAUNCH_CHECK();
  }
  void finish(BBlas& bl, cplx<R>* x) {
    k_dbg_solve<<<1, 64, 0, bl.s>>>(H_.p, g_.p, jmax_.p, negy_.p, sh_.nrhs, m_);
    MGC_LAUNCH_CHECK();
    bmultisub_dev<R>(negy_.p, flexible_ ? Z_.p : V_.p, sh_.n, m_, x, sh_, bl.s);  // x += Z y
  }
  int steps() const { return m_; }
  const BShape& shape() const { return sh_; }

 private:
  BShape sh_;
  int m_;
  bool flexible_;
  double eta_ = 0.5;
  DevBuf<cplx<R>> V_, Z_, w_;
  DevBuf<cplx<double>> dh_, H_, sn_, g_, negy_, r2_;
  DevBuf<double> cs_;
  DevBuf<int> jmax_, skip_;
};

}  // namespace mgc
