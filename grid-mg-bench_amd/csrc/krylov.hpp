// Krylov solvers for the multigrid hierarchy: restarted (F)GMRES with
// classical Gram-Schmidt and selective re-orthogonalisation, complex Givens
// rotations, and the true residual recomputed at every restart.  Templated
// on the vector precision.
//
// One Arnoldi step: h1 = V^H w together with |w|^2 in one fused multi-dot,
// read by the host (one synchronisation).  If |w'|^2 = |w|^2 - |h1|^2 is at
// least half of |w|^2 one projection is enough and v = (w - V h1) / |w'| is
// formed in one fused pass (the Daniel-Gragg-Kaufman-Stewart test with
// threshold 1/sqrt(2)); otherwise the projection is repeated (CGS2) and the
// new norm taken from Pythagoras, |w''|^2 = |w'|^2 - |h2|^2, or directly when
// that cancels by more than a quarter.  The coefficients stay in device
// memory between the passes.
//
// Reference use: the FGMRES outer solver and K-cycle wrapper, GMRES smoother
// and coarsest-level solver of include/MultigridPreconditioner.h:253-309,
// 622-641 (created by include/Helpers.h:38-56 with the defaults of
// include/MultigridParams.h:95-133).
#pragma once
#include <cmath>
#include <complex>
#include <vector>

#include <memory>

#include "blas.hpp"
#include "bkrylov.hpp"

namespace mgc {

template <typename R>
struct LinOp {
  size_t n = 0;  // complex elements of a vector
  virtual ~LinOp() {}
  virtual void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) = 0;
};

// out = P in (the initial content of out is ignored)
template <typename R>
struct Prec {
  virtual ~Prec() {}
  virtual void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) = 0;
};

struct KrylovStats {
  int iters = 0;
  int applies = 0;
  double rel_res = 0.0;
  bool converged = false;
};

// A solver of A x = b (x the initial guess unless x_zero), stopping at
// ||b - A x|| <= tol ||b|| or maxit iterations.
template <typename R>
struct Solver {
  virtual ~Solver() {}
  virtual KrylovStats solve(Blas& bl, LinOp<R>& A, Prec<R>* P, const cplx<R>* b, cplx<R>* x,
                            bool x_zero, double tol, int maxit) = 0;
};

template <typename R>
class Gmres : public Solver<R> {
 public:
  Gmres(size_t n, int m, bool flexible) : n_(n), m_(m), flexible_(flexible) {
    MGC_CHECK(m >= 1 && m + 1 <= Blas::KMAX, MGC_ERR_ARG, "restart length out of range");
    MGC_HIP(hipMalloc(&V_, (size_t)(m + 1) * n * sizeof(cplx<R>)));
    if (flexible) MGC_HIP(hipMalloc(&Z_, (size_t)m * n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&w_, n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&r_, n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&d_h_, (size_t)(2 * Blas::KMAX + 3) * sizeof(cplx<double>)));
  }
  ~Gmres() {
    void* b[] = {V_, Z_, w_, r_, d_h_};
    for (void* p : b)
      if (p) (void)hipFree(p);
  }
  Gmres(const Gmres&) = delete;
  Gmres& operator=(const Gmres&) = delete;

  // Solve A x = b to ||b - A x|| <= tol ||b|| or maxit iterations.  x is the
  // initial guess unless x_zero.
  KrylovStats solve(Blas& bl, LinOp<R>& A, Prec<R>* P, const cplx<R>* b, cplx<R>* x, bool x_zero,
                    double tol, int maxit) override {
    typedef std::complex<double> C;
    hipStream_t s = bl.s;
    const size_t n = n_;
    KrylovStats st;
    const double bnorm = std::sqrt(vnorm2<R>(bl, b, n));
    if (bnorm == 0.0) {
      vzero<R>(x, n, s);
      st.converged = true;
      return st;
    }
    if (x_zero) {
      vzero<R>(x, n, s);
      vcopy<R>(b, r_, n, s);
    } else {
      A.apply(x, w_, s);
      ++st.applies;
      vsub<R>(b, w_, r_, n, s);
    }
    double beta = std::sqrt(vnorm2<R>(bl, r_, n));
    std::vector<C> H((size_t)(m_ + 1) * m_), g(m_ + 1), cs_s(m_);
    std::vector<double> cs_c(m_);
    std::vector<cplx<double>> h(m_ + 1), y(m_);
    while (true) {
      st.rel_res = beta / bnorm;
      if (st.rel_res <= tol) {
        st.converged = true;
        break;
      }
      if (st.iters >= maxit) break;
      vconvert<R, R>(r_, V_, n, 1.0 / beta, s);
      std::fill(g.begin(), g.end(), C(0.0));
      g[0] = beta;
      int j = 0;
      for (; j < m_;) {
        cplx<R>* vj = V_ + (size_t)j * n;
        const cplx<R>* z = vj;
        if (flexible_) {
          cplx<R>* zj = Z_ + (size_t)j * n;
          if (P) P->apply(vj, zj, s);
          else vcopy<R>(vj, zj, n, s);
          z = zj;
        }
        A.apply(z, w_, s);
        ++st.applies;
        // Classical Gram-Schmidt with device-resident coefficients and
        // selective re-orthogonalisation: h1 = V^H w and |w|^2 at
        // d_h_[0 .. j+1]; the second projection (h2, |w'|^2 at
        // d_h_[j+2 .. 2j+3]) only when |w'|^2 = |w|^2 - |h1|^2 < eta |w|^2.
        cplx<double>* dh1 = d_h_;
        cplx<double>* dh2 = d_h_ + (j + 2);
        cplx<R>* vnext = V_ + (size_t)(j + 1) * n;
        vmultidot_dev<R>(bl, V_, n, j + 1, w_, n, true, dh1);
        fetch(bl, reinterpret_cast<const double*>(dh1), 2 * (j + 2));
        double sum1 = 0.0;
        for (int i = 0; i <= j; ++i) {
          h[i] = {bl.h_res[2 * i], bl.h_res[2 * i + 1]};
          sum1 += h[i].re * h[i].re + h[i].im * h[i].im;
        }
        const double wn2 = bl.h_res[2 * (j + 1)];
        double hn;
        if (wn2 - sum1 >= eta_ * wn2) {
          // one projection suffices: v = (w - V h1) / sqrt(|w|^2 - |h1|^2)
          vsubnorm_dev<R>(dh1, V_, n, j + 1, w_, vnext, n, s);
          hn = std::sqrt(wn2 - sum1);
        } else {
          // w' = w - V h1, then h2 = V^H w' and |w'|^2 (one fused pass when j < 24)
          if (!vmultisub_dot_dev<R>(bl, dh1, V_, n, j + 1, w_, n, dh2)) {
            vmultisub_dev<R>(dh1, V_, n, j + 1, w_, n, s);
            vmultidot_dev<R>(bl, V_, n, j + 1, w_, n, true, dh2);
          }
          vsubnorm_dev<R>(dh2, V_, n, j + 1, w_, vnext, n, s);   // v = (w' - V h2) / hn
          fetch(bl, reinterpret_cast<const double*>(dh2), 2 * (j + 2));
          double sum2 = 0.0;
          for (int i = 0; i <= j; ++i) {
            const double ar = bl.h_res[2 * i], ai = bl.h_res[2 * i + 1];
            h[i].re += ar;
            h[i].im += ai;
            sum2 += ar * ar + ai * ai;
          }
          const double wn2b = bl.h_res[2 * (j + 1)];
          hn = wn2b - sum2 > 0.0 ? std::sqrt(wn2b - sum2) : 0.0;
          if (!(hn > 0.0) || sum2 > 0.25 * wn2b) {
            // strong cancellation: form w'' = w' - V h2 and take its norm directly
            vmultisub_dev<R>(dh2, V_, n, j + 1, w_, n, s);
            hn = std::sqrt(vnorm2<R>(bl, w_, n));
            if (hn > 0.0) vconvert<R, R>(w_, vnext, n, 1.0 / hn, s);
          }
        }
        for (int i = 0; i <= j; ++i) H[(size_t)i * m_ + j] = C(h[i].re, h[i].im);
        H[(size_t)(j + 1) * m_ + j] = hn;
        // previous rotations on column j
        for (int i = 0; i < j; ++i) {
          const C a = H[(size_t)i * m_ + j], bb = H[(size_t)(i + 1) * m_ + j];
          H[(size_t)i * m_ + j] = cs_c[i] * a + cs_s[i] * bb;
          H[(size_t)(i + 1) * m_ + j] = -std::conj(cs_s[i]) * a + cs_c[i] * bb;
        }
        // new rotation zeroing H[j+1][j]
        const C a = H[(size_t)j * m_ + j];
        const double bb = std::abs(H[(size_t)(j + 1) * m_ + j]);
        const double t = std::sqrt(std::norm(a) + bb * bb);
        if (std::abs(a) == 0.0) {
          cs_c[j] = 0.0;
          cs_s[j] = 1.0;
        } else {
          cs_c[j] = std::abs(a) / t;
          cs_s[j] = (a / std::abs(a)) * (bb / t);
        }
        H[(size_t)j * m_ + j] = cs_c[j] * a + cs_s[j] * bb;
        H[(size_t)(j + 1) * m_ + j] = 0.0;
        g[j + 1] = -std::conj(cs_s[j]) * g[j];
        g[j] = cs_c[j] * g[j];
        ++st.iters;
        ++j;
        if (std::abs(g[j]) <= tol * bnorm || st.iters >= maxit || hn == 0.0) break;
      }
      // back substitution H y = g
      std::vector<C> yy(j);
      for (int i = j - 1; i >= 0; --i) {
        C acc = g[i];
        for (int k = i + 1; k < j; ++k) acc -= H[(size_t)i * m_ + k] * yy[k];
        yy[i] = acc / H[(size_t)i * m_ + i];
      }
      for (int i = 0; i < j; ++i) y[i] = {yy[i].real(), yy[i].imag()};
      vmultiaxpy<R>(bl, y.data(), flexible_ ? Z_ : V_, n, j, x, n);
      // true residual
      A.apply(x, w_, s);
      ++st.applies;
      vsub<R>(b, w_, r_, n, s);
      beta = std::sqrt(vnorm2<R>(bl, r_, n));
    }
    return st;
  }

  size_t n() const { return n_; }
  // Re-orthogonalisation threshold eta (default 1/2, the Daniel-Gragg-
  // Kaufman-Stewart test).  One classical projection leaves the new vector
  // orthogonal to the basis to about u / sqrt(eta) (u the unit roundoff) and
  // its Pythagorean norm accurate to about u / eta relative, so FP64 solves
  // can take a small eta and skip most second projections.
  void set_reorth_eta(double eta) {
    MGC_CHECK(eta > 0.0 && eta <= 1.0, MGC_ERR_ARG, "re-orthogonalisation threshold in (0, 1]");
    eta_ = eta;
  }

 private:
  size_t n_;
  int m_;
  bool flexible_;
  double eta_ = 0.5;
  cplx<R>* V_ = nullptr;
  cplx<R>* Z_ = nullptr;
  cplx<R>* w_ = nullptr;
  cplx<R>* r_ = nullptr;
  cplx<double>* d_h_ = nullptr;  // device-resident CGS2 coefficients
};

// Unpreconditioned BiCGSTAB (van der Vorst), the setup solver of
// include/MultigridParams.h:106-107 and a selectable solver type of
// include/Helpers.h:38-56: the lockstep BBiCgStab with one right-hand side,
// so the recurrence, the stopping rules (tolerance, maxit, breakdowns) and
// the fused update passes (18 vector passes per step) are those of the
// batched set-up.  With b = 0 (the null-vector setup,
// include/BasisVectors.h:99-105: solve A psi = 0 from psi = v) there is no
// relative tolerance to meet, so exactly maxit iterations relax the initial
// guess.
template <typename R>
class BiCgStab : public Solver<R> {
 public:
  explicit BiCgStab(size_t n) : n_(n) {
    sh_.n = n;
    sh_.nrhs = 1;
    sh_.dof = n % 2 == 0 ? 2 : 1;  // even: the FP32 reductions take element pairs
    solver_ = std::make_unique<BBiCgStab<R>>(sh_);
  }

  KrylovStats solve(Blas& bl, LinOp<R>& A, Prec<R>*, const cplx<R>* b, cplx<R>* x, bool x_zero,
                    double tol, int maxit) override {
    if (!bbl_ || bbl_->s != bl.s) bbl_ = std::make_unique<BBlas>(bl.s);
    struct One : BLinOp<R> {
      LinOp<R>* A;
      void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) override { A->apply(in, out, s); }
    } one;
    one.sh = sh_;
    one.A = &A;
    const BKrylovStats b1 = solver_->solve(*bbl_, one, nullptr, b, x, x_zero, tol, maxit);
    KrylovStats st;
    st.iters = b1.iters_rhs[0];
    st.applies = b1.applies;
    st.rel_res = b1.rel_res[0];
    st.converged = b1.converged[0] != 0;
    return st;
  }

 private:
  size_t n_;
  BShape sh_;
  std::unique_ptr<BBiCgStab<R>> solver_;
  std::unique_ptr<BBlas> bbl_;
};

// Minimal residual iteration (over-relaxation 1), the MR solver type of
// include/Helpers.h:38-56: x += a r, r -= a A r with a = <A r, r> / |A r|^2.
template <typename R>
class MinRes : public Solver<R> {
 public:
  explicit MinRes(size_t n) : n_(n) {
    MGC_HIP(hipMalloc(&r_, n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&ar_, n * sizeof(cplx<R>)));
  }
  ~MinRes() {
    if (r_) (void)hipFree(r_);
    if (ar_) (void)hipFree(ar_);
  }
  MinRes(const MinRes&) = delete;
  MinRes& operator=(const MinRes&) = delete;

  KrylovStats solve(Blas& bl, LinOp<R>& A, Prec<R>*, const cplx<R>* b, cplx<R>* x, bool x_zero,
                    double tol, int maxit) override {
    hipStream_t s = bl.s;
    const size_t n = n_;
    KrylovStats st;
    const double b2 = vnorm2<R>(bl, b, n);
    if (!(b2 > 0.0)) {
      vzero<R>(x, n, s);
      st.converged = true;
      return st;
    }
    if (x_zero) {
      vzero<R>(x, n, s);
      vcopy<R>(b, r_, n, s);
    } else {
      A.apply(x, ar_, s);
      ++st.applies;
      vsub<R>(b, ar_, r_, n, s);
    }
    double r2 = vnorm2<R>(bl, r_, n);
    const double stop = tol * tol * b2;
    while (true) {
      st.rel_res = std::sqrt(r2 / b2);
      if (r2 <= stop) {
        st.converged = true;
        break;
      }
      if (st.iters >= maxit) break;
      A.apply(r_, ar_, s);
      ++st.applies;
      const cplx<double> num = vdot<R>(bl, ar_, r_, n);
      const double den = vnorm2<R>(bl, ar_, n);
      if (!(den > 0.0)) break;
      const cplx<double> a = {num.re / den, num.im / den};
      vaxpy<R>(a, r_, x, n, s);
      vaxpy<R>({-a.re, -a.im}, ar_, r_, n, s);
      r2 = vnorm2<R>(bl, r_, n);
      ++st.iters;
    }
    return st;
  }

 private:
  size_t n_;
  cplx<R>* r_ = nullptr;
  cplx<R>* ar_ = nullptr;
};

}  // namespace mgc
