// Batched complex BLAS-1 for nrhs independent systems solved in lockstep.
// A batched vector is [site][rhs][dof] (the multi-rhs layout of the ABI):
// element i belongs to right-hand side r(i) = (i % (nrhs*dof)) / dof.  Every
// reduction returns one FP64 result per rhs; coefficient arrays hold one
// entry per rhs ([j][nrhs] for the multi-vector forms).
//
// The reduction kernels launch a multiple of nrhs*dof threads, so each thread
// only ever touches elements of one rhs; a workgroup's partial sums are added
// per rhs in LDS in a fixed order, and the per-workgroup partials in a second
// fixed-order pass, so results are reproducible run to run.
#pragma once
#include "blas.hpp"

namespace mgc {

struct BShape {
  size_t n = 0;  // complex elements of one batched vector (sites * nrhs * dof)
  int nrhs = 1;
  int dof = 1;
  int cols() const { return nrhs * dof; }
};

struct BBlas {
  hipStream_t s = nullptr;
  double* d_part = nullptr;  // [rows][nrhs][nblk][2]
  size_t part_cap = 0;       // doubles
  double* h_res = nullptr;   // pinned, [rows][nrhs][2]
  cplx<double>* d_coef = nullptr;  // device coefficients [KMAX+1][nrhs]
  cplx<double>* h_coef = nullptr;  // pinned staging of host coefficients
  hipEvent_t coef_done = nullptr;
  static constexpr int KMAX = 64;
  static constexpr int RMAX = 64;  // rhs per batch
  BBlas(hipStream_t st);
  ~BBlas();
  BBlas(const BBlas&) = delete;
  BBlas& operator=(const BBlas&) = delete;
};

// per-rhs reductions (synchronise the stream; results [row][r] in h_res)
template <typename R>
void bdots_dev(BBlas& b, const cplx<R>* V, size_t ldv, int k, const cplx<R>* w,
               const BShape& sh, bool with_norm, cplx<double>* d_out);
// copy count doubles to b.h_res and wait
void bfetch(BBlas& b, const double* d_src, int count);
template <typename R> void bnorm2(BBlas& b, const cplx<R>* x, const BShape& sh, double* out);

// Fused re-orthogonalisation pass: w -= sum_j c[j][r] V_j, then d_out[j][r] =
// <V_j, w>_r for j < k and d_out[k][r].re = |w|^2_r, as bmultisub_dev followed
// by bdots_dev(..., true, d_out) but with one read of V; false (nothing done)
// for k > 24.  A nonzero device flag *skip makes the pass change nothing.
template <typename R>
bool bmultisub_dots_dev(BBlas& b, const cplx<double>* d_c, const cplx<R>* V, size_t ldv, int k,
                        cplx<R>* w, const BShape& sh, cplx<double>* d_out,
                        const int* skip = nullptr);
// w -= sum_j c[j][r] V_j (c on the device, [k][nrhs])
template <typename R>
void bmultisub_dev(const cplx<double>* d_c, const cplx<R>* V, size_t ldv, int k, cplx<R>* w,
                   const BShape& sh, hipStream_t s);
// w += sum_j c[j][r] V_j (c on the host, [k][nrhs], uploaded through pinned memory)
template <typename R>
void bmultiaxpy(BBlas& b, const cplx<double>* c, const cplx<R>* V, size_t ldv, int k,
                cplx<R>* w, const BShape& sh);
// v = (w - sum_j h2[j][r] V_j) / hn_r with hn_r = sqrt(h2[k][r].re - sum_j |h2[j][r]|^2)
// (0 if hn_r <= 0); hn_r stored in h2[k][r].im
template <typename R>
void bsubnorm_dev(cplx<double>* d_h2, const cplx<R>* V, size_t ldv, int k, const cplx<R>* w,
                  cplx<R>* v, const BShape& sh, hipStream_t s);
// y = x * a[r] with precision change (a on the host, nrhs reals)
template <typename Rs, typename Rd>
void bconvert(BBlas& b, const cplx<Rs>* x, cplx<Rd>* y, const BShape& sh, const double* a);

// BiCGSTAB updates with per-rhs coefficients c [3][nrhs] on the host (rows 0
// and 1 as named, row 2 .re = 1 for a rhs still iterating, 0 for one that
// has stopped: its vectors are left as they are), each one pass:
//   bbicg_p : p = r + c0 (p - c1 v)
//   bbicg_s : s = r - c0 v
//   bbicg_xr: x += c0 p + c1 s; r = s - c1 t
template <typename R>
void bbicg_p(BBlas& b, const cplx<double>* c, const cplx<R>* r, const cplx<R>* v, cplx<R>* p,
             const BShape& sh);
template <typename R>
void bbicg_s(BBlas& b, const cplx<double>* c, const cplx<R>* r, const cplx<R>* v, cplx<R>* s,
             const BShape& sh);
template <typename R>
void bbicg_xr(BBlas& b, const cplx<double>* c, const cplx<R>* p, const cplx<R>* s,
              const cplx<R>* t, cplx<R>* x, cplx<R>* r, const BShape& sh);
// bbicg_xr, then per rhs <r, r> and <rh, r> of the updated r into d_out rows
// 0 and 1 ([2][nrhs] complex, on the device; a stopped rhs gets zeros), in
// the same pass (bdots_dev's reduction over the updated r, without a second
// read of r)
template <typename R>
void bbicg_xr_dots(BBlas& b, const cplx<double>* c, const cplx<R>* p, const cplx<R>* s,
                   const cplx<R>* t, const cplx<R>* rh, cplx<R>* x, cplx<R>* r,
                   const BShape& sh, cplx<double>* d_out);

// per-rhs scaling by a device factor from r2 ([nrhs] complex, .re used, as
// bdots_dev's |w|^2 row leaves it), with no host synchronisation:
//   bconvert_dev: y = x / sqrt(r2) (mode 1, 0 where r2 <= 0) or x * sqrt(r2)
//                 (mode 2), with precision change
//   bscale_dev:   x /= sqrt(r2) in place (0 where r2 <= 0)
template <typename Rs, typename Rd>
void bconvert_dev(const cplx<Rs>* x, cplx<Rd>* y, const BShape& sh, const cplx<double>* d_r2,
                  int mode, hipStream_t s);
template <typename R>
void bscale_dev(cplx<R>* x, const BShape& sh, const cplx<double>* d_r2, hipStream_t s);

// a set of sh.nrhs vectors [rhs][ld] (ld = sites * dof) into one batched
// vector [site][rhs][dof], and back
template <typename R>
void bpack(const cplx<R>* set, size_t ld, cplx<R>* out, const BShape& sh, hipStream_t s);
template <typename R>
void bunpack(const cplx<R>* in, cplx<R>* set, size_t ld, const BShape& sh, hipStream_t s);

}  // namespace mgc
