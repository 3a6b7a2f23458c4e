// Batched restarted (F)GMRES: nrhs independent systems A x_r = b_r advanced in
// lockstep, so that every operator application and every BLAS pass serves
// all right-hand sides at once.  Each rhs keeps its own Hessenberg matrix,
// complex Givens rotations and convergence test: once rhs r has converged
// (|g_{j+1}| <= tol |b_r|) or broken down, the Krylov columns computed
// afterwards are not used for its update, exactly as if it had been solved
// alone and stopped there.  Classical Gram-Schmidt with selective
// re-orthogonalisation and device-resident coefficients, as the single-rhs
// solver (the second projection runs for the whole batch when any live rhs
// needs it).
#pragma once
#include <cmath>
#include <complex>
#include <vector>

#include "bblas.hpp"

namespace mgc {

template <typename R>
struct BLinOp {
  BShape sh;
  virtual ~BLinOp() {}
  virtual void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) = 0;
};

template <typename R>
struct BPrec {
  virtual ~BPrec() {}
  virtual void apply(const cplx<R>* in, cplx<R>* out, hipStream_t s) = 0;
};

struct BKrylovStats {
  int iters = 0;    // lockstep iterations
  int applies = 0;  // batched operator applications
  std::vector<double> rel_res;
  std::vector<int> converged;
  std::vector<int> iters_rhs;  // iteration at which each rhs converged (or iters)
};

template <typename R>
struct BSolver {
  virtual ~BSolver() {}
  virtual BKrylovStats solve(BBlas& bl, BLinOp<R>& A, BPrec<R>* P, const cplx<R>* b, cplx<R>* x,
                             bool x_zero, double tol, int maxit) = 0;
};

template <typename R>
class BGmres : public BSolver<R> {
 public:
  BGmres(const BShape& sh, int m, bool flexible) : sh_(sh), m_(m), flexible_(flexible) {
    MGC_CHECK(m >= 1 && 2 * m + 3 <= 2 * BBlas::KMAX + 3, MGC_ERR_ARG,
              "restart length out of range");
    MGC_CHECK(sh.nrhs >= 1 && sh.nrhs <= BBlas::RMAX, MGC_ERR_ARG, "batch of 1..64 rhs");
    const size_t n = sh.n;
    MGC_HIP(hipMalloc(&V_, (size_t)(m + 1) * n * sizeof(cplx<R>)));
    if (flexible) MGC_HIP(hipMalloc(&Z_, (size_t)m * n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&w_, n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&r_, n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&d_h_, (size_t)(2 * BBlas::KMAX + 3) * sh.nrhs * sizeof(cplx<double>)));
  }
  ~BGmres() {
    void* b[] = {V_, Z_, w_, r_, d_h_};
    for (void* p : b)
      if (p) (void)hipFree(p);
  }
  BGmres(const BGmres&) = delete;
  BGmres& operator=(const BGmres&) = delete;

  BKrylovStats solve(BBlas& bl, BLinOp<R>& A, BPrec<R>* P, const cplx<R>* b, cplx<R>* x,
                     bool x_zero, double tol, int maxit) override {
    typedef std::complex<double> C;
    hipStream_t s = bl.s;
    const size_t n = sh_.n;
    const int nr = sh_.nrhs;
    const int m = m_;
    BKrylovStats st;
    st.rel_res.assign(nr, 0.0);
    st.converged.assign(nr, 0);
    st.iters_rhs.assign(nr, 0);
    std::vector<double> bnorm(nr), beta(nr), tmp(nr);
    bnorm2<R>(bl, b, sh_, tmp.data());
    for (int r = 0; r < nr; ++r) bnorm[r] = std::sqrt(tmp[r]);
    if (x_zero) {
      vzero<R>(x, n, s);
      vcopy<R>(b, r_, n, s);
    } else {
      A.apply(x, w_, s);
      ++st.applies;
      vsub<R>(b, w_, r_, n, s);
    }
    bnorm2<R>(bl, r_, sh_, tmp.data());
    for (int r = 0; r < nr; ++r) beta[r] = std::sqrt(tmp[r]);
    std::vector<C> H((size_t)nr * (m + 1) * m), g((size_t)nr * (m + 1)), sn((size_t)nr * m);
    std::vector<double> cs((size_t)nr * m);
    std::vector<int> jr(nr);      // Krylov columns used by rhs r in this cycle
    std::vector<char> live(nr);   // rhs still iterating in this cycle
    std::vector<cplx<double>> y((size_t)m * nr);
    std::vector<double> scale(nr);
    auto Hc = [&](int r, int i, int j) -> C& { return H[((size_t)r * (m + 1) + i) * m + j]; };
    while (true) {
      bool all_done = true;
      for (int r = 0; r < nr; ++r) {
        st.rel_res[r] = bnorm[r] > 0.0 ? beta[r] / bnorm[r] : 0.0;
        const bool done = bnorm[r] == 0.0 || st.rel_res[r] <= tol;
        if (done && !st.converged[r]) {
          st.converged[r] = 1;
          st.iters_rhs[r] = st.iters;
        }
        live[r] = !done;
        all_done = all_done && done;
        jr[r] = 0;
      }
      if (all_done || st.iters >= maxit) break;
      for (int r = 0; r < nr; ++r) scale[r] = live[r] && beta[r] > 0.0 ? 1.0 / beta[r] : 0.0;
      bconvert<R, R>(bl, r_, V_, sh_, scale.data());
      std::fill(g.begin(), g.end(), C(0.0));
      for (int r = 0; r < nr; ++r) g[(size_t)r * (m + 1)] = live[r] ? beta[r] : 0.0;
      int j = 0;
      for (; j < m;) {
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
        // classical Gram-Schmidt with selective re-orthogonalisation (as the
        // single-rhs solver): h1 and |w|^2 at rows 0 .. j+1 of d_h_; the
        // second projection at rows j+2 .. 2j+3 when some live rhs has
        // |w'|^2 = |w|^2 - |h1|^2 < eta |w|^2
        cplx<double>* dh1 = d_h_;
        cplx<double>* dh2 = d_h_ + (size_t)(j + 2) * nr;
        cplx<R>* vnext = V_ + (size_t)(j + 1) * n;
        bdots_dev<R>(bl, V_, n, j + 1, w_, sh_, true, dh1);
        bfetch(bl, reinterpret_cast<const double*>(dh1), 2 * (j + 2) * nr);
        std::vector<double> hn(nr), sum1(nr, 0.0), wn2(nr);
        bool reorth = false;
        for (int r = 0; r < nr; ++r) {
          for (int i = 0; i <= j; ++i) {
            const double* a1 = bl.h_res + 2 * ((size_t)i * nr + r);
            Hc(r, i, j) = C(a1[0], a1[1]);
            sum1[r] += a1[0] * a1[0] + a1[1] * a1[1];
          }
          wn2[r] = bl.h_res[2 * ((size_t)(j + 1) * nr + r)];
          hn[r] = wn2[r] - sum1[r] > 0.0 ? std::sqrt(wn2[r] - sum1[r]) : 0.0;
          if (live[r] && wn2[r] - sum1[r] < eta_ * wn2[r]) reorth = true;
        }
        if (!reorth) {
          bsubnorm_dev<R>(dh1, V_, n, j + 1, w_, vnext, sh_, s);
        } else {
          // w' = w - V h1, then h2 = V^H w' and |w'|^2 (one fused pass when j < 24)
          if (!bmultisub_dots_dev<R>(bl, dh1, V_, n, j + 1, w_, sh_, dh2)) {
            bmultisub_dev<R>(dh1, V_, n, j + 1, w_, sh_, s);
            bdots_dev<R>(bl, V_, n, j + 1, w_, sh_, true, dh2);
          }
          bsubnorm_dev<R>(dh2, V_, n, j + 1, w_, vnext, sh_, s);
          bfetch(bl, reinterpret_cast<const double*>(dh2), 2 * (j + 2) * nr);
          bool redo = false;
          std::vector<double> sum2(nr, 0.0), wn2b(nr);
          for (int r = 0; r < nr; ++r) {
            for (int i = 0; i <= j; ++i) {
              const double* a2 = bl.h_res + 2 * ((size_t)i * nr + r);
              Hc(r, i, j) += C(a2[0], a2[1]);
              sum2[r] += a2[0] * a2[0] + a2[1] * a2[1];
            }
            wn2b[r] = bl.h_res[2 * ((size_t)(j + 1) * nr + r)];
            hn[r] = wn2b[r] - sum2[r] > 0.0 ? std::sqrt(wn2b[r] - sum2[r]) : 0.0;
            if (live[r] && wn2b[r] > 0.0 && (!(hn[r] > 0.0) || sum2[r] > 0.25 * wn2b[r]))
              redo = true;
          }
          if (redo) {
            // strong cancellation for some rhs: form w'' and take its norms directly
            bmultisub_dev<R>(dh2, V_, n, j + 1, w_, sh_, s);
            bnorm2<R>(bl, w_, sh_, tmp.data());
            for (int r = 0; r < nr; ++r) {
              hn[r] = std::sqrt(tmp[r]);
              scale[r] = hn[r] > 0.0 ? 1.0 / hn[r] : 0.0;
            }
            bconvert<R, R>(bl, w_, vnext, sh_, scale.data());
          }
        }
        ++st.iters;
        ++j;
        bool any_live = false;
        for (int r = 0; r < nr; ++r) {
          if (!live[r]) continue;
          const int jj = j - 1;
          Hc(r, jj + 1, jj) = hn[r];
          for (int i = 0; i < jj; ++i) {
            const C a = Hc(r, i, jj), bb = Hc(r, i + 1, jj);
            Hc(r, i, jj) = cs[(size_t)r * m + i] * a + sn[(size_t)r * m + i] * bb;
            Hc(r, i + 1, jj) = -std::conj(sn[(size_t)r * m + i]) * a + cs[(size_t)r * m + i] * bb;
          }
          const C a = Hc(r, jj, jj);
          const double bb = std::abs(Hc(r, jj + 1, jj));
          const double t = std::sqrt(std::norm(a) + bb * bb);
          if (std::abs(a) == 0.0) {
            cs[(size_t)r * m + jj] = 0.0;
            sn[(size_t)r * m + jj] = 1.0;
          } else {
            cs[(size_t)r * m + jj] = std::abs(a) / t;
            sn[(size_t)r * m + jj] = (a / std::abs(a)) * (bb / t);
          }
          Hc(r, jj, jj) = cs[(size_t)r * m + jj] * a + sn[(size_t)r * m + jj] * bb;
          Hc(r, jj + 1, jj) = 0.0;
          C* gr = &g[(size_t)r * (m + 1)];
          gr[jj + 1] = -std::conj(sn[(size_t)r * m + jj]) * gr[jj];
          gr[jj] = cs[(size_t)r * m + jj] * gr[jj];
          jr[r] = j;
          if (std::abs(gr[j]) <= tol * bnorm[r] || hn[r] == 0.0 || Hc(r, jj, jj) == 0.0)
            live[r] = 0;  // converged or broken down: freeze its columns
          any_live = any_live || live[r];
        }
        if (!any_live || st.iters >= maxit) break;
      }
      // per-rhs back substitution and one batched update of every x_r
      std::fill(y.begin(), y.end(), cplx<double>{0.0, 0.0});
      for (int r = 0; r < nr; ++r) {
        const int k = jr[r];
        std::vector<C> yy(k);
        for (int i = k - 1; i >= 0; --i) {
          C acc = g[(size_t)r * (m + 1) + i];
          for (int q = i + 1; q < k; ++q) acc -= Hc(r, i, q) * yy[q];
          yy[i] = Hc(r, i, i) != 0.0 ? acc / Hc(r, i, i) : C(0.0);
        }
        for (int i = 0; i < k; ++i) y[(size_t)i * nr + r] = {yy[i].real(), yy[i].imag()};
      }
      bmultiaxpy<R>(bl, y.data(), flexible_ ? Z_ : V_, n, j, x, sh_);
      A.apply(x, w_, s);
      ++st.applies;
      vsub<R>(b, w_, r_, n, s);
      bnorm2<R>(bl, r_, sh_, tmp.data());
      for (int r = 0; r < nr; ++r) beta[r] = std::sqrt(tmp[r]);
    }
    for (int r = 0; r < nr; ++r)
      if (!st.converged[r]) st.iters_rhs[r] = st.iters;
    return st;
  }

  const BShape& shape() const { return sh_; }

  // re-orthogonalisation threshold eta, as Gmres::set_reorth_eta
  void set_reorth_eta(double eta) {
    MGC_CHECK(eta > 0.0 && eta <= 1.0, MGC_ERR_ARG, "re-orthogonalisation threshold in (0, 1]");
    eta_ = eta;
  }

 private:
  BShape sh_;
  int m_;
  bool flexible_;
  double eta_ = 0.5;
  cplx<R>* V_ = nullptr;
  cplx<R>* Z_ = nullptr;
  cplx<R>* w_ = nullptr;
  cplx<R>* r_ = nullptr;
  cplx<double>* d_h_ = nullptr;  // [2*KMAX+3][nrhs] CGS2 coefficients
};

// Minimal residual iteration for nrhs systems in lockstep (the MR smoother
// of include/Helpers.h:38-56 on the multi-rhs grid): per rhs
// a_r = <A r, r>_r / |A r|^2_r, x_r += a_r r_r, r_r -= a_r (A r)_r; a rhs that
// has met its tolerance keeps a_r = 0, so its iterate is that of a solve of
// its own.
template <typename R>
class BMinRes : public BSolver<R> {
 public:
  explicit BMinRes(const BShape& sh) : sh_(sh) {
    MGC_CHECK(sh.nrhs >= 1 && sh.nrhs <= BBlas::RMAX, MGC_ERR_ARG, "batch of 1..64 rhs");
    MGC_HIP(hipMalloc(&r_, sh.n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&ar_, sh.n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&d_out_, (size_t)2 * sh.nrhs * sizeof(cplx<double>)));
  }
  ~BMinRes() {
    void* b[] = {r_, ar_, d_out_};
    for (void* p : b)
      if (p) (void)hipFree(p);
  }
  BMinRes(const BMinRes&) = delete;
  BMinRes& operator=(const BMinRes&) = delete;

  BKrylovStats solve(BBlas& bl, BLinOp<R>& A, BPrec<R>*, const cplx<R>* b, cplx<R>* x,
                     bool x_zero, double tol, int maxit) override {
    hipStream_t s = bl.s;
    const size_t n = sh_.n;
    const int nr = sh_.nrhs;
    BKrylovStats st;
    st.rel_res.assign(nr, 0.0);
    st.converged.assign(nr, 0);
    st.iters_rhs.assign(nr, 0);
    std::vector<double> b2(nr), r2(nr);
    bnorm2<R>(bl, b, sh_, b2.data());
    if (x_zero) {
      vzero<R>(x, n, s);
      vcopy<R>(b, r_, n, s);
    } else {
      A.apply(x, ar_, s);
      ++st.applies;
      vsub<R>(b, ar_, r_, n, s);
    }
    bnorm2<R>(bl, r_, sh_, r2.data());
    std::vector<cplx<double>> c(nr);
    while (true) {
      bool all = true;
      for (int r = 0; r < nr; ++r) {
        st.rel_res[r] = b2[r] > 0.0 ? std::sqrt(r2[r] / b2[r]) : 0.0;
        if (!st.converged[r] && (b2[r] == 0.0 || r2[r] <= tol * tol * b2[r])) {
          st.converged[r] = 1;
          st.iters_rhs[r] = st.iters;
        }
        all = all && st.converged[r];
      }
      if (all || st.iters >= maxit) break;
      A.apply(r_, ar_, s);
      ++st.applies;
      // <A r, r>_r and |A r|^2_r
      bdots_dev<R>(bl, ar_, n, 1, r_, sh_, false, d_out_);
      bdots_dev<R>(bl, ar_, n, 1, ar_, sh_, false, d_out_ + nr);
      bfetch(bl, reinterpret_cast<const double*>(d_out_), 4 * nr);
      for (int r = 0; r < nr; ++r) {
        const double* num = bl.h_res + 2 * r;
        const double den = bl.h_res[2 * (nr + r)];
        c[r] = (st.converged[r] || !(den > 0.0)) ? cplx<double>{0.0, 0.0}
                                                  : cplx<double>{num[0] / den, num[1] / den};
      }
      bmultiaxpy<R>(bl, c.data(), r_, n, 1, x, sh_);
      for (int r = 0; r < nr; ++r) c[r] = {-c[r].re, -c[r].im};
      bmultiaxpy<R>(bl, c.data(), ar_, n, 1, r_, sh_);
      bnorm2<R>(bl, r_, sh_, r2.data());
      ++st.iters;
    }
    for (int r = 0; r < nr; ++r)
      if (!st.converged[r]) st.iters_rhs[r] = st.iters;
    return st;
  }

 private:
  BShape sh_;
  cplx<R>* r_ = nullptr;
  cplx<R>* ar_ = nullptr;
  cplx<double>* d_out_ = nullptr;
};

// BiCGSTAB for nrhs systems in lockstep: per rhs the recurrence of the
// single-rhs BiCgStab (the BiCGSTAB solver type of include/Helpers.h:51-53),
// with its stopping rules per rhs: |r| <= tol |b|, maxit, a breakdown
// (rho, omega or <rh, v> zero), and for b = 0 (null vectors relaxed from x)
// exactly maxit steps with no convergence test.  A rhs that has stopped keeps
// its x and residual while the others go on.  Three host reductions per step
// ((<rh, r>, |r|^2), <rh, v>, (<s, t>, |t|^2)) serve the whole batch; the
// first of them is formed in the pass that updates x and r.
template <typename R>
class BBiCgStab : public BSolver<R> {
 public:
  explicit BBiCgStab(const BShape& sh) : sh_(sh) {
    MGC_CHECK(sh.nrhs >= 1 && sh.nrhs <= BBlas::RMAX, MGC_ERR_ARG, "batch of 1..64 rhs");
    // r and rh adjacent, so that one two-row reduction gives <r, r> and <rh, r>
    MGC_HIP(hipMalloc(&w_, (size_t)6 * sh.n * sizeof(cplx<R>)));
    MGC_HIP(hipMalloc(&d_out_, (size_t)3 * sh.nrhs * sizeof(cplx<double>)));
  }
  ~BBiCgStab() {
    if (w_) (void)hipFree(w_);
    if (d_out_) (void)hipFree(d_out_);
  }
  BBiCgStab(const BBiCgStab&) = delete;
  BBiCgStab& operator=(const BBiCgStab&) = delete;

  BKrylovStats solve(BBlas& bl, BLinOp<R>& A, BPrec<R>*, const cplx<R>* b, cplx<R>* x,
                     bool x_zero, double tol, int maxit) override {
    typedef std::complex<double> C;
    hipStream_t st_ = bl.s;
    const size_t n = sh_.n;
    const int nr = sh_.nrhs;
    cplx<R>*r = w_, *rh = w_ + n, *p = w_ + 2 * n, *v = w_ + 3 * n, *sv = w_ + 4 * n,
           *t = w_ + 5 * n;
    BKrylovStats st;
    st.rel_res.assign(nr, 0.0);
    st.converged.assign(nr, 0);
    st.iters_rhs.assign(nr, 0);
    std::vector<double> b2(nr), r2(nr);
    std::vector<int> live(nr, 1), relax(nr);
    std::vector<C> rho(nr, 1.0), alpha(nr, 1.0), omega(nr, 1.0), rho1(nr);
    std::vector<cplx<double>> c(3 * (size_t)nr);
    bnorm2<R>(bl, b, sh_, b2.data());
    for (int q = 0; q < nr; ++q) relax[q] = !(b2[q] > 0.0);
    if (x_zero) {
      vzero<R>(x, n, st_);
      vcopy<R>(b, r, n, st_);
      for (int q = 0; q < nr; ++q)
        if (relax[q]) {  // b = 0 from x = 0: x = 0 at once
          live[q] = 0;
          st.converged[q] = 1;
        }
    } else {
      A.apply(x, t, st_);
      ++st.applies;
      vsub<R>(b, t, r, n, st_);
    }
    vcopy<R>(r, rh, n, st_);
    vzero<R>(p, n, st_);
    vzero<R>(v, n, st_);
    auto cd = [](C z) { return cplx<double>{z.real(), z.imag()}; };
    auto dc = [](const double* h) { return C(h[0], h[1]); };
    // (<r, r>, <rh, r>) per rhs: computed here the first time, by the fused
    // update of x and r afterwards
    auto fetch_r = [&](bool compute) {
      if (compute) bdots_dev<R>(bl, r, n, 2, r, sh_, false, d_out_);
      bfetch(bl, reinterpret_cast<const double*>(d_out_), 4 * nr);
      for (int q = 0; q < nr; ++q) {
        r2[q] = bl.h_res[2 * q];
        rho1[q] = dc(bl.h_res + 2 * (nr + q));
      }
    };
    auto flags = [&]() {
      for (int q = 0; q < nr; ++q) c[2 * (size_t)nr + q] = {live[q] ? 1.0 : 0.0, 0.0};
    };
    fetch_r(true);
    while (true) {
      bool any = false;
      for (int q = 0; q < nr; ++q) {
        if (!live[q]) continue;
        st.rel_res[q] = relax[q] ? std::sqrt(r2[q]) : std::sqrt(r2[q] / b2[q]);
        st.iters_rhs[q] = st.iters;
        if (!relax[q] && r2[q] <= tol * tol * b2[q]) {
          st.converged[q] = 1;
          live[q] = 0;
        } else if (st.iters >= maxit || std::abs(rho1[q]) == 0.0 || std::abs(omega[q]) == 0.0) {
          live[q] = 0;  // iteration limit or breakdown
        }
        any = any || live[q];
      }
      if (!any) break;
      for (int q = 0; q < nr; ++q) {
        if (!live[q]) continue;
        const C beta = (rho1[q] / rho[q]) * (alpha[q] / omega[q]);
        rho[q] = rho1[q];
        c[q] = cd(beta);
        c[nr + q] = cd(omega[q]);
      }
      flags();
      bbicg_p<R>(bl, c.data(), r, v, p, sh_);  // p = r + beta (p - omega v)
      A.apply(p, v, st_);
      bdots_dev<R>(bl, rh, n, 1, v, sh_, false, d_out_);
      bfetch(bl, reinterpret_cast<const double*>(d_out_), 2 * nr);
      for (int q = 0; q < nr; ++q) {
        if (!live[q]) continue;
        const C rhv = dc(bl.h_res + 2 * q);
        if (std::abs(rhv) == 0.0) {
          live[q] = 0;  // breakdown: x and r as they were
          continue;
        }
        alpha[q] = rho[q] / rhv;
        c[q] = cd(alpha[q]);
      }
      flags();
      bbicg_s<R>(bl, c.data(), r, v, sv, sh_);  // s = r - alpha v
      A.apply(sv, t, st_);
      st.applies += 2;
      // (<t, t>, <s, t>) per rhs: omega = <t, s> / |t|^2 = conj(<s, t>) / |t|^2
      bdots_dev<R>(bl, sv, n, 1, t, sh_, true, d_out_);
      bfetch(bl, reinterpret_cast<const double*>(d_out_), 4 * nr);
      for (int q = 0; q < nr; ++q) {
        if (!live[q]) continue;
        const C ts = std::conj(dc(bl.h_res + 2 * q));
        const double tt = bl.h_res[2 * (nr + q)];
        omega[q] = tt > 0.0 ? ts / tt : C(0.0);
        c[q] = cd(alpha[q]);
        c[nr + q] = cd(omega[q]);
      }
      // x += alpha p + omega s; r = s - omega t; (<r, r>, <rh, r>) in the same pass
      bbicg_xr_dots<R>(bl, c.data(), p, sv, t, rh, x, r, sh_, d_out_);
      ++st.iters;
      fetch_r(false);
    }
    return st;
  }

 private:
  BShape sh_;
  cplx<R>* w_ = nullptr;  // r, rh, p, v, s, t
  cplx<double>* d_out_ = nullptr;
};

}  // namespace mgc
