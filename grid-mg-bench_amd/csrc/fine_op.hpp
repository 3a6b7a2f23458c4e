// Internal state of the fine Wilson-clover operator handle.
#pragma once
#include "common.hpp"

// Convention (chiral basis, g5 = g_x g_y g_z g_t = diag(1,1,-1,-1)):
//   (D psi)(x) = [(4+m) + C(x)] psi(x)
//                - 1/2 sum_mu [ (1 - g_mu) U_mu(x) psi(x+mu) + (1 + g_mu) U_mu(x-mu)^dag psi(x-mu) ]
//   C(x) = -(csw/2) sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x), sigma = (i/2)[g_mu, g_nu],
//   F = (Q - Q^dag)/(8i), Q = the four plaquette leaves ("clover") at x.
// C commutes with g5, so (4+m) + C is stored as two 6x6 blocks per site.
struct mgc_fine_op {
  mgc::Lat4 lat;
  int V = 0;
  double mass = 0.0, csw = 0.0;
  int antiperiodic_t = 0;
  bool gauge_set = false;
  mgc::cplx<double>* gauge64 = nullptr;  // [4][V][3][3], boundary phase folded in
  mgc::cplx<double>* diag64 = nullptr;   // [V][2][6][6]
  // site-fastest copies read by the apply kernel, so that consecutive lanes
  // (consecutive sites) read consecutive addresses:
  int sv = 0;  // site stride of the site-fastest copies: V + 32, so that the 36
               // streams of one site do not all start in the same L2 set
  mgc::cplx<double>* gsoa64 = nullptr;   // [4][9][sv]  (mu, row*3+col, site)
  mgc::cplx<double>* dsoa64 = nullptr;   // [2][18][sv]  (chirality, packed slot, site):
                                         // each 6x6 diagonal block is hermitian; slots
                                         // 0..2 hold the real diagonal as pairs
                                         // (d0,d1) (d2,d3) (d4,d5), slots 3..17 the
                                         // upper entries (a<b) in row order
  mgc::cplx<float>* gsoa32 = nullptr;
  mgc::cplx<float>* dsoa32 = nullptr;
  int* nbr = nullptr;                    // [V][8]
};

namespace mgc {
// gamma_mu acting on spin s: (g_mu psi)_s = phase[mu][s] * psi_{col[mu][s]}
struct GammaTab {
  int col[4][4];
  float re[4][4], im[4][4];
};
__host__ __device__ inline GammaTab gamma_table() {
  // x: (0,3,i) (1,2,i) (2,1,-i) (3,0,-i); y: (0,3,-1) (1,2,1) (2,1,1) (3,0,-1)
  // z: (0,2,i) (1,3,-i) (2,0,-i) (3,1,i);  t: (0,2,1) (1,3,1) (2,0,1) (3,1,1)
  GammaTab g = {{{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {2, 3, 0, 1}},
                {{0, 0, 0, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}, {1, 1, 1, 1}},
                {{1, 1, -1, -1}, {0, 0, 0, 0}, {1, -1, -1, 1}, {0, 0, 0, 0}}};
  return g;
}
}  // namespace mgc
