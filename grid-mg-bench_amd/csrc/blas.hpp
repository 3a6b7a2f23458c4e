// Complex BLAS-1 on device vectors with FP64 reductions, for the Krylov
// solvers of the multigrid hierarchy.  All calls are stream-ordered.  The
// host-result reductions (vdot, vnorm2, vmultidot) synchronise the stream;
// the *_dev forms keep their scalars in device memory and do not.
#pragma once
#include "common.hpp"

namespace mgc {

struct Blas {
  hipStream_t s = nullptr;
  double* d_part = nullptr;  // partial sums [kmax + 1][NPART][2]
  double* d_res = nullptr;   // final results [kmax][2]
  double* h_res = nullptr;   // pinned host copy
  cplx<double>* d_coef = nullptr;  // coefficients for multi-axpy [kmax]
  cplx<double>* h_coef = nullptr;  // pinned staging of host coefficients [kmax]
  hipEvent_t coef_done = nullptr;  // the last upload of h_coef has been read
  static constexpr int NPART = 2048;
  static constexpr int KMAX = 64;
  explicit Blas(hipStream_t st);
  ~Blas();
  Blas(const Blas&) = delete;
  Blas& operator=(const Blas&) = delete;
};

template <typename R> void vzero(cplx<R>* x, size_t n, hipStream_t s);
template <typename R> void vcopy(const cplx<R>* x, cplx<R>* y, size_t n, hipStream_t s);
template <typename Rs, typename Rd>
void vconvert(const cplx<Rs>* x, cplx<Rd>* y, size_t n, double scale, hipStream_t s);
// y += a x
template <typename R> void vaxpy(cplx<double> a, const cplx<R>* x, cplx<R>* y, size_t n,
                                 hipStream_t s);
// y = x + a y
template <typename R> void vxpay(const cplx<R>* x, cplx<double> a, cplx<R>* y, size_t n,
                                 hipStream_t s);
// x *= a (real)
template <typename R> void vscale(double a, cplx<R>* x, size_t n, hipStream_t s);
// z = x - y
template <typename R> void vsub(const cplx<R>* x, const cplx<R>* y, cplx<R>* z, size_t n,
                                hipStream_t s);
// <x, y> = sum conj(x) y
template <typename R> cplx<double> vdot(Blas& b, const cplx<R>* x, const cplx<R>* y, size_t n);
template <typename R> double vnorm2(Blas& b, const cplx<R>* x, size_t n);
// h[j] = <V_j, w> for j < k, V_j = V + j*ldv
template <typename R> void vmultidot(Blas& b, const cplx<R>* V, size_t ldv, int k,
                                     const cplx<R>* w, size_t n, cplx<double>* h);
// w += sum_j c[j] V_j  (c on the host; uploaded through pinned memory, no sync)
template <typename R> void vmultiaxpy(Blas& b, const cplx<double>* c, const cplx<R>* V, size_t ldv,
                                      int k, cplx<R>* w, size_t n);

// Device-scalar forms (no host synchronisation):
// d_h[j] = <V_j, w> for j < k, and, if with_norm, d_h[k].re = |w|^2
template <typename R> void vmultidot_dev(Blas& b, const cplx<R>* V, size_t ldv, int k,
                                         const cplx<R>* w, size_t n, bool with_norm,
                                         cplx<double>* d_h);
// w -= sum_j d_c[j] V_j
template <typename R> void vmultisub_dev(const cplx<double>* d_c, const cplx<R>* V, size_t ldv,
                                         int k, cplx<R>* w, size_t n, hipStream_t s);
// CGS2 tail: hn = sqrt(d_h2[k].re - sum_{j<k} |d_h2[j]|^2) (the norm of w after
// the second projection, by Pythagoras from the norm before it), v = w / hn
// (v = 0 if hn <= 0); hn is stored in d_h2[k].im
template <typename R> void vnormalize_dev(const cplx<R>* w, cplx<R>* v, size_t n,
                                          cplx<double>* d_h2, int k, hipStream_t s);
// Fused CGS2 tail: v = (w - sum_j d_h2[j] V_j) / hn with hn = sqrt(d_h2[k].re - sum_j |d_h2[j]|^2)
// (Pythagoras, as vnormalize_dev; v = 0 if hn <= 0; hn stored in d_h2[k].im);
// w is left unchanged
template <typename R> void vsubnorm_dev(cplx<double>* d_h2, const cplx<R>* V, size_t ldv, int k,
                                        const cplx<R>* w, cplx<R>* v, size_t n, hipStream_t s);
// Fused re-orthogonalisation pass: w -= sum_j d_c[j] V_j, then d_h[j] =
// <V_j, w> for j < k and d_h[k].re = |w|^2 of the updated w, with the same
// results as vmultisub_dev followed by vmultidot_dev(..., true, d_h) but one
// read of V from memory.  Returns false (and does nothing) for k > 24.  With
// skip, a device flag read by the kernels, a nonzero *skip makes the pass
// change nothing (w and d_h as they were): a device-side decision that a
// captured graph can replay.
template <typename R> bool vmultisub_dot_dev(Blas& b, const cplx<double>* d_c, const cplx<R>* V,
                                             size_t ldv, int k, cplx<R>* w, size_t n,
                                             cplx<double>* d_h, const int* skip = nullptr);
// y = x * a with a from the device scalar d_s[0]: mode 0 a = d_s[0], mode 1
// a = 1/sqrt(d_s[0]), mode 2 a = sqrt(d_s[0]) (a = 0 where d_s[0] <= 0 in modes
// 1, 2); px / py, when not null, are device slots holding the vector
// pointers (read by the kernel, so a captured graph can take new vectors)
template <typename Rs, typename Rd>
void vconvert_dev(const cplx<Rs>* x, cplx<Rd>* y, const cplx<Rs>* const* px, cplx<Rd>* const* py,
                  size_t n, const double* d_s, int mode, hipStream_t s);
// d_res[0] = |x|^2 (d_res[1] = 0) in device memory, x direct or through the slot px
template <typename R>
void vnorm2_dev(Blas& b, const cplx<R>* x, const cplx<R>* const* px, size_t n, double* d_res);
// copy count doubles from device to the pinned h_res and wait for them
void fetch(Blas& b, const double* d_src, int count);

}  // namespace mgc
