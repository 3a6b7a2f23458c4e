// Internal state of a coarse operator handle (CoarsenedMatrix).
#pragma once
#include "common.hpp"

// Internal link storage (precision `prec`):
//   links      [site][point 9][col N][row N]   column-major per (site, point),
//              so one site's 9 matrices are one contiguous stream and a
//              wavefront reads a column's rows with 16-byte coalesced loads.
//   self_dag   [site][col][row]  = adj(A_8)
//   self_inv   [site][col][row]  = A_8^{-1}         (after invert)
//   self_invdag[site][col][row]  = adj(A_8^{-1})
//   mrhs_links half/fp32 copies with per-(site, point) scales for the
//              multi-rhs MFMA kernels (built lazily, see coarse_mrhs.hip)
struct mgc_coarse_op {
  mgc::Lat4 lat;
  int V = 0;
  int N = 0;
  int prec = MGC_F64;
  int hermitian = 0;
  bool links_set = false;
  bool inv_set = false;

  void* links = nullptr;
  void* self_dag = nullptr;
  void* self_inv = nullptr;
  void* self_invdag = nullptr;

  int* nbr_full = nullptr;   // [V][8]   full-lattice neighbour index
  int* nbr_half = nullptr;   // [V][8]   half-lattice index (x>>1) of the neighbour
  int* cb_sites = nullptr;   // [2][V/2] full index of half-lattice site h of parity cb

  // multi-rhs operand (links re-laid for MFMA), built from `links` on demand
  void* mrhs_links = nullptr;
  float* mrhs_scales = nullptr;  // [site][9] (half precisions)
  int mrhs_prec = -1;
  // FP32 -> F16/BF16 quantised input of the multi-rhs apply (grown on demand,
  // kept for later calls): q_in [site][rhs][N] half pairs, q_scale [site][rhs]
  void* q_in = nullptr;
  float* q_scale = nullptr;
  size_t q_rows = 0;  // rows (site x rhs) the workspace holds

  size_t elem_bytes() const { return 2 * mgc::prec_real_bytes(prec); }
  size_t vec_bytes() const { return (size_t)V * N * elem_bytes(); }
};

namespace mgc {
// single-rhs stencil launcher shared by the apply calls
void stencil_launch(const mgc_coarse_op* op, const void* links, size_t site_stride,
                    size_t point_stride, const void* in, void* out, const int* out_site,
                    int n_out, const int* nbr, unsigned mask, int dag, hipStream_t s);
}  // namespace mgc
