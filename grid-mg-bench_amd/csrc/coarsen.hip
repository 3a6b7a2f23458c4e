// Galerkin coarse-operator construction (CoarsenOperator) and the
// gamma5-hermiticity link reconstruction, FP64.
//
// Reference semantics:
//   include/CoarsenedMatrixBaseline.h:575-705   CoarsenOperator, inner/outer-mask form
//   include/CoarsenedMatrixUpstreamImprovedDirsave.h:835-932  backward links + self only,
//   include/CoarsenedMatrixUpstreamImprovedDirsave.h:999-1032 ConstructRemainingLinks
//   include/CoarsenedMatrixKernelsRework.h:141-167 shiftLinks (the converse direction)
//   include/CoarsenedMatrixRework.h:640-668     InvertSelfStencilPoint (done by set_links)
//
// With the chirally doubled basis v_{l*nb+c} = P_l u_c the links are
//   A_p(xc)[k*nb+c][l*nb+c'] = sum_{x in B(xc), outer_p} sum_{e in k} conj(u_c(x,e)) W_p^l(x,e,c')
//   W_p^l(x) = H_p(x) P_l u_{c'}(x + d_p),  H_p = -1/2 (1 - disp g_mu) (U_mu(x) | U_mu(x-mu)^dag)
// and the self link uses the block-restricted operator D_BB (diagonal term plus
// every hop that stays inside the block), which is the sum of the reference's
// Mdiag and inner-mask projections.
//
// One workgroup per (coarse site, point).  Per stage of SB fine sites the
// workgroup builds W for both chiralities in LDS (items (l, c', colour)), stages
// the block-panel rows of u, and every thread accumulates a TC x TC register
// tile of one chirality block (k, l) of the N x N link in FP64.
#include <algorithm>
#include <string>
#include <type_traits>

#include "aggregation.hpp"
#include "coarse_op.hpp"
#include "fine_op.hpp"

namespace mgc {

constexpr int CO_SB = 4;  // fine sites per stage

// Workgroup b -> (coarse site, point).  The points of one coarse site are
// consecutive in the dispatch sequence of one XCD (b % 8 is the XCD), so the
// workgroups that read one block panel run together on one L2; with a coarse
// volume that is not a multiple of 8 they are simply consecutive.
__device__ __forceinline__ void coarsen_work(int b, int np, int Vc, int& xc, int& pi) {
  if (Vc % 8 == 0) {
    const int i = b >> 3;
    xc = (i / np) * 8 + (b & 7);
    pi = i % np;
  } else {
    xc = b / np;
    pi = b % np;
  }
}

// Geometry of one (coarse site, point) work item of the construction.
struct CoGeom {
  const cplx<double>* basis;  // [xc][xb][e 12][nb] panels
  const cplx<double>* gauge;  // [mu][site][3][3]
  const cplx<double>* diag;   // [site][2][6][6]
  Lat4 fine, coarse, blk;
  int Vf, Vb;
  int xc, p, mu, disp, nsite;
  int cc[4], bs[4], fs[4], cs[4];

  __device__ CoGeom(const cplx<double>* b, const cplx<double>* g, const cplx<double>* d, Lat4 f,
                    Lat4 c, Lat4 bl, int vf, int vb, int x, int pt)
      : basis(b), gauge(g), diag(d), fine(f), coarse(c), blk(bl), Vf(vf), Vb(vb), xc(x), p(pt) {
    coarse.coords(xc, cc);
    bs[0] = 1; bs[1] = blk.L[0]; bs[2] = bs[1] * blk.L[1]; bs[3] = bs[2] * blk.L[2];
    fs[0] = 1; fs[1] = fine.L[0]; fs[2] = fs[1] * fine.L[1]; fs[3] = fs[2] * fine.L[2];
    cs[0] = 1; cs[1] = coarse.L[0]; cs[2] = cs[1] * coarse.L[1]; cs[3] = cs[2] * coarse.L[2];
    mu = p & 3;
    disp = p < 4 ? 1 : -1;
    // sites of this point: outer face of the block for p < 8, the whole block for p = 8
    // (blk.L[mu] by a select: a runtime index would put the geometry in scratch memory)
    const int Lmu = mu == 0 ? blk.L[0] : mu == 1 ? blk.L[1] : mu == 2 ? blk.L[2] : blk.L[3];
    nsite = (p == 8) ? Vb : Vb / Lmu;
  }
  // face site f: the coordinates other than mu run lexicographically over the
  // face, coordinate mu is the block's last (+mu) or first (-mu) layer
  __device__ int face_site(int f) const {
    if (p == 8) return f;
    int r = f, idx = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      int cd;
      if (d == mu) {
        cd = disp > 0 ? blk.L[d] - 1 : 0;
      } else {
        cd = r % blk.L[d];
        r /= blk.L[d];
      }
      idx += cd * stride;
      stride *= blk.L[d];
    }
    return idx;
  }
};

// One W item: W^lw(xb)[e, cp] for all 12 spin-colour elements, where
// W^l = H_p P_l u_cp for a hop (p < 8) and D_BB P_l u_cp (inner hops plus the
// diagonal) for p = 8.  The chirality lw is a runtime value, so that lanes
// of both chiralities run one instruction stream: ws[sl*3 + ci] holds spin
// 2 lw + sl (chirality lw), wo[k*3 + ci] spin 2 (1 - lw) + k, and the
// spin-dependent constants are chosen by selects between compile-time values
// (an index computed from lw would put the accumulators in scratch memory).
// The hops are taken HB at a time: the loads of the batch (link, and the
// neighbour's chirality-lw basis values) are issued before any is used.
// One hop per batch (HB = 1) measured fastest: with two the loads take ~230
// registers and one wavefront per SIMD; links staged in LDS for the
// wavefront's site cost a barrier per stage and ran slower.
template <int NB, int HB>
__device__ __forceinline__ void coarsen_w_lane(const CoGeom& g, const GammaTab& G, int xb, int lw,
                                               int cp, cplx<double> ws[6], cplx<double> wo[6]) {
  constexpr int ROW = 12 * NB;
  const int p = g.p;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    ws[e] = {0.0, 0.0};
    wo[e] = {0.0, 0.0};
  }
  // block and fine coordinates of this site
  int bc[4], x = 0;
  {
    int r = xb;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      bc[d] = r % g.blk.L[d];
      r /= g.blk.L[d];
      x += (g.cc[d] * g.blk.L[d] + bc[d]) * g.fs[d];
    }
  }
  const int e0 = 6 * lw;  // first element of chirality lw
#pragma unroll
  for (int q0 = 0; q0 < 8; q0 += HB) {
    if (p != 8 && (p < q0 || p >= q0 + HB)) continue;  // uniform: a hop point uses one hop
    cplx<double> L[HB][9], B[HB][6];  // link U (row-major), basis (sl, colour)
    bool use[HB];
#pragma unroll
    for (int h = 0; h < HB; ++h) {
      const int q = q0 + h;
      const int qm = q & 3;
      const int disp = q < 4 ? 1 : -1;
      const bool inner = (unsigned)(bc[qm] + disp) < (unsigned)g.blk.L[qm];
      // self: inner hops only, by the reference's coordinate mask
      // (x_mu mod b != b-1 for +mu, != 0 for -mu)
      use[h] = p == 8 ? inner : q == p;
      if (!use[h]) continue;
      // fine neighbour y = x + disp*mu (periodic) and its panel row
      const int fq = g.cc[qm] * g.blk.L[qm] + bc[qm] + disp;
      const int y = x + disp * g.fs[qm] -
                    (fq < 0 ? -g.fine.L[qm] * g.fs[qm]
                            : fq >= g.fine.L[qm] ? g.fine.L[qm] * g.fs[qm] : 0);
      int ry;
      if (inner) {
        ry = g.xc * g.Vb + xb + disp * g.bs[qm];
      } else {
        const int cq = g.cc[qm] + disp;
        const int xn = g.xc + disp * g.cs[qm] -
                       (cq < 0 ? -g.coarse.L[qm] * g.cs[qm]
                               : cq >= g.coarse.L[qm] ? g.coarse.L[qm] * g.cs[qm] : 0);
        ry = xn * g.Vb + xb - disp * (g.blk.L[qm] - 1) * g.bs[qm];
      }
      const cplx<double>* U = g.gauge + ((size_t)qm * g.Vf + (q < 4 ? x : y)) * 9;
#pragma unroll
      for (int k = 0; k < 9; ++k) L[h][k] = U[k];
      const cplx<double>* uy = g.basis + (size_t)ry * ROW + (size_t)e0 * NB + cp;
#pragma unroll
      for (int k = 0; k < 6; ++k) B[h][k] = uy[k * NB];  // element e0 + k = 3 sl + j
    }
#pragma unroll
    for (int h = 0; h < HB; ++h) {
      if (!use[h]) continue;
      const int q = q0 + h;
      const int qm = q & 3;
      const double sg = q < 4 ? 1.0 : -1.0;
#pragma unroll
      for (int sl = 0; sl < 2; ++sl) {
        // gamma image of spin 2 lw + sl: spin s2 of the other chirality,
        // local index k = s2 - 2 (1 - lw), phase of s2
        const int s2a = G.col[qm][sl], s2b = G.col[qm][2 + sl];  // lw = 0, lw = 1
        const int ka = s2a - 2, kb = s2b;
        const cplx<double> pa = {(double)G.re[qm][s2a], (double)G.im[qm][s2a]};
        const cplx<double> pb = {(double)G.re[qm][s2b], (double)G.im[qm][s2b]};
        const cplx<double> ph = lw ? pb : pa;
        const int kimg = lw ? kb : ka;
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) {
          cplx<double> chi = {0.0, 0.0};
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            // row ci of U (q < 4) or of U^dag (q >= 4)
            const cplx<double> l = q < 4 ? L[h][ci * 3 + j]
                                         : cplx<double>{L[h][j * 3 + ci].re, -L[h][j * 3 + ci].im};
            cmac(chi, l, B[h][sl * 3 + j]);
          }
          // same-chirality part: -1/2 chi
          ws[sl * 3 + ci].re -= 0.5 * chi.re;
          ws[sl * 3 + ci].im -= 0.5 * chi.im;
          // gamma image: + sg/2 * phase(s2) * chi at spin s2
          const cplx<double> t = cmul(ph, chi);
          const cplx<double> v = {0.5 * sg * t.re, 0.5 * sg * t.im};
          if (ka == kb) {  // the same local spin for both chiralities (compile time)
            wo[ka * 3 + ci].re += v.re;
            wo[ka * 3 + ci].im += v.im;
          } else {
            const double f0 = kimg == 0 ? 1.0 : 0.0;
            wo[ci].re += f0 * v.re;
            wo[ci].im += f0 * v.im;
            wo[3 + ci].re += (1.0 - f0) * v.re;
            wo[3 + ci].im += (1.0 - f0) * v.im;
          }
        }
      }
    }
  }
  if (p == 8) {  // diagonal (mass + clover) term, block-diagonal in chirality
    const cplx<double>* D = g.diag + ((size_t)x * 2 + lw) * 36;
    const cplx<double>* ux = g.basis + ((size_t)g.xc * g.Vb + xb) * ROW + (size_t)e0 * NB + cp;
    cplx<double> v[6];
#pragma unroll
    for (int b = 0; b < 6; ++b) v[b] = ux[b * NB];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      cplx<double> a = {0.0, 0.0};
#pragma unroll
      for (int b = 0; b < 6; ++b) cmac(a, D[r * 6 + b], v[b]);
      ws[r].re += a.re;
      ws[r].im += a.im;
    }
  }
}

// FP64 vector form: one workgroup per (coarse site, point).  Per stage of
// CO_SB fine sites the workgroup builds W for both chiralities in LDS (items
// (l, c', colour)), stages the block-panel rows of u, and every thread
// accumulates a TC x TC register tile of one chirality block (k, l) of the
// N x N link.
template <int NB, int TC>
__global__ __launch_bounds__(4 * (NB / TC) * (NB / TC)) void k_coarsen(
    const cplx<double>* __restrict__ basis, const cplx<double>* __restrict__ gauge,
    const cplx<double>* __restrict__ diag, Lat4 fine, Lat4 coarse, Lat4 blk, int Vf, int Vb,
    const int* __restrict__ plist, int np, int Vc, cplx<double>* __restrict__ links) {
  constexpr int NT1 = NB / TC;
  constexpr int THREADS = 4 * NT1 * NT1;
  constexpr int ROW = 12 * NB;  // complex per panel row group (one fine site)
  __shared__ cplx<double> sU[CO_SB][12][NB];
  __shared__ cplx<double> sW[CO_SB][2][12][NB];
  const GammaTab G = gamma_table();

  int xc, pi;
  coarsen_work((int)blockIdx.x, np, Vc, xc, pi);
  const CoGeom g(basis, gauge, diag, fine, coarse, blk, Vf, Vb, xc, plist[pi]);
  const int tid = threadIdx.x;
  const int kl = tid / (NT1 * NT1);
  const int k = kl >> 1, l = kl & 1;
  const int rem = tid % (NT1 * NT1);
  const int cb = rem / NT1, pb = rem % NT1;

  cplx<double> acc[TC][TC];
#pragma unroll
  for (int i = 0; i < TC; ++i)
#pragma unroll
    for (int j = 0; j < TC; ++j) acc[i][j] = {0.0, 0.0};

  for (int s0 = 0; s0 < g.nsite; s0 += CO_SB) {
    // ---- stage u rows of the SB sites
    for (int idx = tid; idx < CO_SB * ROW; idx += THREADS) {
      const int s = idx / ROW, r = idx - s * ROW;
      // past the face end: load block site 0 (in bounds) and stage zero
      const bool ok = s0 + s < g.nsite;
      const int xb = ok ? g.face_site(s0 + s) : 0;
      const cplx<double> v = basis[((size_t)xc * Vb + xb) * ROW + r];
      (&sU[s][0][0])[r] = ok ? v : cplx<double>{0.0, 0.0};
    }
    // ---- W items (s, l', c'), all 12 elements each
    for (int it = tid; it < CO_SB * 2 * NB; it += THREADS) {
      const int s = it / (2 * NB);
      const int r = it - s * 2 * NB;
      const int lw = r / NB, cp = r - lw * NB;
      cplx<double> ws[6], wo[6];
      if (s0 + s < g.nsite) {
        coarsen_w_lane<NB, 1>(g, G, g.face_site(s0 + s), lw, cp, ws, wo);
      } else {
#pragma unroll
        for (int e = 0; e < 6; ++e) ws[e] = wo[e] = {0.0, 0.0};
      }
#pragma unroll
      for (int e = 0; e < 6; ++e) {
        sW[s][lw][6 * lw + e][cp] = ws[e];
        sW[s][lw][6 - 6 * lw + e][cp] = wo[e];
      }
    }
    __syncthreads();
    // ---- accumulate the (k, l) block: acc[c][c'] += sum_{e in k} conj(u(e,c)) W^l(e,c')
#pragma unroll
    for (int s = 0; s < CO_SB; ++s)
#pragma unroll
      for (int e6 = 0; e6 < 6; ++e6) {
        const int e = 6 * k + e6;
        cplx<double> a[TC], b[TC];
#pragma unroll
        for (int i = 0; i < TC; ++i) a[i] = sU[s][e][cb * TC + i];
#pragma unroll
        for (int j = 0; j < TC; ++j) b[j] = sW[s][l][e][pb * TC + j];
#pragma unroll
        for (int i = 0; i < TC; ++i)
#pragma unroll
          for (int j = 0; j < TC; ++j) cmac_conj(acc[i][j], a[i], b[j]);
      }
    __syncthreads();
  }
  // ---- write A_p(xc) rows k*NB + c, columns l*NB + c'
  constexpr int N = 2 * NB;
  cplx<double>* Ap = links + ((size_t)g.p * Vc + xc) * N * N;
#pragma unroll
  for (int i = 0; i < TC; ++i)
#pragma unroll
    for (int j = 0; j < TC; ++j)
      Ap[(size_t)(k * NB + cb * TC + i) * N + l * NB + pb * TC + j] = acc[i][j];
}

// FP64 MFMA form (v_mfma_f64_16x16x4_f64), one workgroup of four wavefronts
// per (coarse site, point), wavefront w owning the chirality block
// (k, l) = (w / 2, w % 2) of the N x N link.  Each block is the complex
// product C = sum_{s, e in k} conj(u(s,e))^T W^l(s,e), run as one real GEMM
// with the complex index folded into K (the scheme of the multi-rhs stencil):
//   A'[c][2m+t]   = (Re u, -Im u)(m, c)           for t = 0, 1 (m = (s, e))
//   B'[2m+t][2c'+v] = v = 0: (Re w, -Im w),  v = 1: (Im w, Re w)
// so that C'[c][2c'+v] = (Re, Im) C[c][c'].  Rows and columns past NB are zero
// padding of the 16 x 16 tiles and are never stored.  The staged u and W are
// kept as separate real and imaginary planes in LDS, so every lane's operand
// is one 8-byte read.  W is built on the vector ALU as in the FP64 form.
// MFMA 16x16x4 f64 lane layout: A[i = l%16][k = l/16], B[k = l/16][j = l%16],
// D[i = 4r + l/16][j = l%16] for result r < 4 (rows interleaved, unlike the
// FP32 16x16x4 form's 4(l/16) + r).
constexpr int CM_SB = 4;  // fine sites per stage
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <int NB>
__global__ __launch_bounds__(256) void k_coarsen_mfma(
    const cplx<double>* __restrict__ basis, const cplx<double>* __restrict__ gauge,
    const cplx<double>* __restrict__ diag, Lat4 fine, Lat4 coarse, Lat4 blk, int Vf, int Vb,
    const int* __restrict__ plist, int np, int Vc, cplx<double>* __restrict__ links) {
  constexpr int MT = (NB + 15) / 16;   // m-tiles (rows c)
  constexpr int NTL = (NB + 7) / 8;    // n-tiles (columns 2c'+v)
  constexpr int NBP = NB + 1;          // LDS row pitch (doubles) against bank conflicts
  constexpr int ROW = 12 * NB;
  static_assert(CM_SB == 4 && NB <= 32, "one stage site per wavefront, 2 NB lanes per site");
  __shared__ double sUr[CM_SB][12][NBP], sUi[CM_SB][12][NBP];      // u, with -Im u
  __shared__ double sWr[CM_SB][2][12][NBP], sWi[CM_SB][2][12][NBP];  // W
  const GammaTab G = gamma_table();

  int xc, pi;
  coarsen_work((int)blockIdx.x, np, Vc, xc, pi);
  const CoGeom g(basis, gauge, diag, fine, coarse, blk, Vf, Vb, xc, plist[pi]);
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  const int kb = w >> 1, lb = w & 1;  // chirality block of this wavefront
  const int lk = l >> 4;              // k index of this lane within a k-step
  const int t = lk & 1;               // real (0) or imaginary (1) slot of K
  // A operand: row c = mt*16 + l%16, from the plane of t
  const double* aplane = t ? &sUi[0][0][0] : &sUr[0][0][0];
  // B operand: column n = nt*16 + l%16, c' = n/2, v = n%2: plane Re for
  // t == v, else Im, negated for (t, v) = (1, 0)
  const int v = l & 1;
  const double* bplane = (t ^ v) ? &sWi[0][0][0][0] : &sWr[0][0][0][0];
  const double bsign = (t == 1 && v == 0) ? -1.0 : 1.0;

  doublex4 acc[MT][NTL];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt) acc[mt][nt] = {0.0, 0.0, 0.0, 0.0};

  for (int s0 = 0; s0 < g.nsite; s0 += CM_SB) {
    // ---- stage u rows of the SB sites as real and (negated) imaginary planes
    for (int idx = tid; idx < CM_SB * ROW; idx += 256) {
      const int s = idx / ROW, r = idx - s * ROW;
      const int e = r / NB, c = r - e * NB;
      const bool ok = s0 + s < g.nsite;
      const int xb = ok ? g.face_site(s0 + s) : 0;
      const cplx<double> u = basis[((size_t)xc * Vb + xb) * ROW + r];
      sUr[s][e][c] = ok ? u.re : 0.0;
      sUi[s][e][c] = ok ? -u.im : 0.0;
    }
    // ---- W: wavefront w builds site w of the stage, lane l = (lw, c') with
    // lw = l / NB (both chiralities in one instruction stream; the u staging
    // above is not read here, so no barrier separates the two)
    {
      const int s = __builtin_amdgcn_readfirstlane(tid >> 6);
      const int lw = l / NB, cp = l - lw * NB;
      if (l < 2 * NB) {
        cplx<double> ws[6], wo[6];
        if (s0 + s < g.nsite) {
          coarsen_w_lane<NB, 1>(g, G, g.face_site(s0 + s), lw, cp, ws, wo);
        } else {
#pragma unroll
          for (int e = 0; e < 6; ++e) ws[e] = wo[e] = {0.0, 0.0};
        }
#pragma unroll
        for (int e = 0; e < 6; ++e) {
          sWr[s][lw][6 * lw + e][cp] = ws[e].re;
          sWi[s][lw][6 * lw + e][cp] = ws[e].im;
          sWr[s][lw][6 - 6 * lw + e][cp] = wo[e].re;
          sWi[s][lw][6 - 6 * lw + e][cp] = wo[e].im;
        }
      }
    }
    __syncthreads();
    // ---- the block's GEMM: per site three k-steps of two elements each
#pragma unroll 1
    for (int s = 0; s < CM_SB; ++s)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int e = 6 * kb + 2 * q + (lk >> 1);  // element of this lane's K slot
        double a[MT], b[NTL];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const int c = mt * 16 + (l & 15);
          a[mt] = (NB % 16 == 0 || c < NB) ? aplane[(s * 12 + e) * NBP + c] : 0.0;
        }
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) {
          const int cq = (nt * 16 + (l & 15)) >> 1;
          b[nt] = (NB % 8 == 0 || cq < NB)
                      ? bsign * bplane[((s * 2 + lb) * 12 + e) * NBP + cq]
                      : 0.0;
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int nt = 0; nt < NTL; ++nt)
            acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
      }
    __syncthreads();
  }
  // ---- write A_p(xc) rows k*NB + c, columns l*NB + c' (component v)
  constexpr int N = 2 * NB;
  double* Ap = reinterpret_cast<double*>(links + ((size_t)g.p * Vc + xc) * N * N);
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt) {
      const int n = nt * 16 + (l & 15);
      const int cq = n >> 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int c = mt * 16 + 4 * r + lk;
        if ((NB % 16 == 0 || c < NB) && (NB % 8 == 0 || cq < NB))
          Ap[((size_t)(kb * NB + c) * N + lb * NB + cq) * 2 + v] = acc[mt][nt][r];
      }
    }
}

// A_want(x) = sigma o adj(A_have(x + disp_want * mu)), FLAT layout, in place
__global__ void k_reconstruct(cplx<double>* __restrict__ links, Lat4 lat, int N, int disp_have,
                              const int* __restrict__ nbr) {
  const int V = lat.volume();
  const size_t total = (size_t)4 * V * N * N;
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const int nb = N / 2;
  for (; t < total; t += stride) {
    const int c = (int)(t % N);
    size_t q = t / N;
    const int r = (int)(q % N);
    q /= N;
    const int x = (int)(q % V);
    const int mu = (int)(q / V);
    const int ph = disp_have > 0 ? mu : mu + 4;  // point we have
    const int pw = disp_have > 0 ? mu + 4 : mu;  // point we want
    // neighbour in direction disp_want = -disp_have: the want-point's own neighbour
    const int y = nbr[(size_t)x * 8 + pw];
    cplx<double> a = links[(((size_t)ph * V + y) * N + c) * N + r];
    a.im = -a.im;
    if ((r / nb + c / nb) & 1) { a.re = -a.re; a.im = -a.im; }
    links[(((size_t)pw * V + x) * N + r) * N + c] = a;
  }
}

template <int NB, int TC>
static void launch_coarsen_t(const mgc_aggregation* agg, const mgc_fine_op* fop, const int* plist,
                             int np, cplx<double>* links, hipStream_t s) {
  constexpr int THREADS = 4 * (NB / TC) * (NB / TC);
  k_coarsen<NB, TC><<<agg->Vc * np, THREADS, 0, s>>>(agg->basis64, fop->gauge64, fop->diag64,
                                                      agg->fine, agg->coarse, agg->blk, agg->Vf,
                                                      agg->Vb, plist, np, agg->Vc, links);
  MGC_LAUNCH_CHECK();
}

template <int NB>
static void launch_coarsen_mfma_t(const mgc_aggregation* agg, const mgc_fine_op* fop,
                                  const int* plist, int np, cplx<double>* links, hipStream_t s) {
  k_coarsen_mfma<NB><<<agg->Vc * np, 256, 0, s>>>(agg->basis64, fop->gauge64, fop->diag64,
                                                  agg->fine, agg->coarse, agg->blk, agg->Vf,
                                                  agg->Vb, plist, np, agg->Vc, links);
  MGC_LAUNCH_CHECK();
}

// build kernel: 0 = FP64 vector form, 1 = FP64 MFMA form; -1 = the default (MFMA)
static int g_coarsen_variant = -1;

static void launch_coarsen(const mgc_aggregation* agg, const mgc_fine_op* fop, const int* plist,
                           int np, cplx<double>* links, hipStream_t s) {
  if (g_coarsen_variant != 0) {
    switch (agg->nb) {
#define CASE(n)                                                    \
  case n:                                                          \
    launch_coarsen_mfma_t<n>(agg, fop, plist, np, links, s);       \
    return;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11)
      CASE(12) CASE(13) CASE(14) CASE(15) CASE(16) CASE(17) CASE(18) CASE(19) CASE(20) CASE(21)
      CASE(22) CASE(23) CASE(24) CASE(25) CASE(26) CASE(27) CASE(28) CASE(29) CASE(30) CASE(31)
      CASE(32)
#undef CASE
    }
    throw Error(MGC_ERR_SHAPE, "CoarsenOperator supports nb from 1 to 32");
  }
  switch (agg->nb) {
    case 2: launch_coarsen_t<2, 1>(agg, fop, plist, np, links, s); return;
    case 4: launch_coarsen_t<4, 1>(agg, fop, plist, np, links, s); return;
    case 8: launch_coarsen_t<8, 2>(agg, fop, plist, np, links, s); return;
    case 12: launch_coarsen_t<12, 3>(agg, fop, plist, np, links, s); return;
    case 16: launch_coarsen_t<16, 2>(agg, fop, plist, np, links, s); return;
    case 24: launch_coarsen_t<24, 3>(agg, fop, plist, np, links, s); return;
    case 32: launch_coarsen_t<32, 4>(agg, fop, plist, np, links, s); return;
    case 6: launch_coarsen_t<6, 1>(agg, fop, plist, np, links, s); return;
    case 10: launch_coarsen_t<10, 2>(agg, fop, plist, np, links, s); return;
    case 14: launch_coarsen_t<14, 2>(agg, fop, plist, np, links, s); return;
    case 18: launch_coarsen_t<18, 2>(agg, fop, plist, np, links, s); return;
    case 20: launch_coarsen_t<20, 2>(agg, fop, plist, np, links, s); return;
    case 22: launch_coarsen_t<22, 2>(agg, fop, plist, np, links, s); return;
    case 26: launch_coarsen_t<26, 2>(agg, fop, plist, np, links, s); return;
    case 28: launch_coarsen_t<28, 4>(agg, fop, plist, np, links, s); return;
    case 30: launch_coarsen_t<30, 3>(agg, fop, plist, np, links, s); return;
  }
  throw Error(MGC_ERR_SHAPE,
              "the vector form of CoarsenOperator supports nb = 2 and every even nb from 4 to 32");
}

static void reconstruct(cplx<double>* links, const mgc_coarse_op* cop, int disp_have,
                        hipStream_t s) {
  const size_t total = (size_t)4 * cop->V * cop->N * cop->N;
  k_reconstruct<<<grid_for(total, 256, 16384), 256, 0, s>>>(links, cop->lat, cop->N, disp_have,
                                                            cop->nbr_full);
  MGC_LAUNCH_CHECK();
}

static void set_links_or_throw(mgc_coarse_op* cop, const void* links, int invert, hipStream_t s) {
  const int st = mgc_coarse_op_set_links(cop, MGC_LAYOUT_FLAT, links, invert, (void*)s);
  if (st != MGC_OK) throw Error(st, mgc_last_error());
}

// links[p][xc][j][n0 + r] (=, or += when accumulate) proj[xc][r][j], r < nr
__global__ void k_scatter_columns(cplx<double>* __restrict__ links,
                                  const cplx<double>* __restrict__ proj, int Vc, int N, int p,
                                  int n0, int nr, int accumulate) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= (size_t)Vc * nr * N) return;
  const int j = (int)(t % N);
  const size_t q = t / N;
  const int r = (int)(q % nr);
  const size_t xc = q / nr;
  cplx<double>& dst = links[(((size_t)p * Vc + xc) * N + j) * N + n0 + r];
  const cplx<double> v = proj[t];
  if (accumulate) dst = cadd(dst, v);
  else dst = v;
}

static void check_call(int st) {
  if (st != MGC_OK) throw Error(st, mgc_last_error());
}

// Galerkin product of an operator given point by point, in the inner/outer-
// mask form of include/CoarsenedMatrixBaseline.h:609-675: for a batch of nr
// doubled basis columns v_n, W_p = (stencil point p of the operator) v_n for
// p = 0..8 (OpDir / OpDiag), then A_p(xc)[j][n] = sum over the outer sites of
// p of <v_j, W_p> and A_8(xc)[j][n] = sum over the inner sites of every p and
// all sites of W_8, FP64.  apply(p, in, out, nr) acts on FP64 fields
// [site][nr][Ef].  DIRSAVE skips the forward points and rebuilds their links
// by gamma5-hermiticity.
template <typename Apply>
static void coarsen_by_points(mgc_coarse_op* cop2, mgc_aggregation* agg, int mode, int batch,
                              Apply&& apply, hipStream_t s) {
  const int N2 = cop2->N, Vc2 = cop2->V, Vf = agg->Vf, Ef = agg->Ef;
  batch = std::max(1, std::min(batch, N2));
  DevBuf<cplx<double>> links((size_t)9 * Vc2 * N2 * N2), phi((size_t)Vf * batch * Ef),
      mphi((size_t)Vf * batch * Ef), proj((size_t)Vc2 * batch * N2);
  fill_zero(links.p, (size_t)9 * Vc2 * N2 * N2 * sizeof(cplx<double>), s);
  for (int n0 = 0; n0 < N2; n0 += batch) {
    const int nr = std::min(batch, N2 - n0);
    const unsigned g = (unsigned)(((size_t)Vc2 * nr * N2 + 255) / 256);
    agg_extract_doubled(agg, n0, nr, phi.p, s);
    for (int p = 0; p < 9; ++p) {
      apply(p, phi.p, mphi.p, nr);
      // outer sites: link p (DIRSAVE rebuilds the forward links instead)
      if (p < 8 && !(mode == MGC_COARSEN_DIRSAVE && p < 4)) {
        agg_transfer(agg, mphi.p, proj.p, MGC_F64, nr, true, 2, p, s);
        k_scatter_columns<<<g, 256, 0, s>>>(links.p, proj.p, Vc2, N2, p, n0, nr, 0);
        MGC_LAUNCH_CHECK();
      }
      // inner sites of every hop, and every site of the diagonal: the self link
      agg_transfer(agg, mphi.p, proj.p, MGC_F64, nr, true, 1, p, s);
      k_scatter_columns<<<g, 256, 0, s>>>(links.p, proj.p, Vc2, N2, 8, n0, nr, 1);
      MGC_LAUNCH_CHECK();
    }
  }
  if (mode == MGC_COARSEN_DIRSAVE) reconstruct(links.p, cop2, -1, s);
  set_links_or_throw(cop2, links.p, 1, s);
  MGC_HIP(hipStreamSynchronize(s));  // the buffers are freed on return
}

// CoarsenOperator of a coarse operator: one basis column at a time (the
// single-rhs coarse apply)
static void coarsen_coarse(mgc_coarse_op* cop2, mgc_coarse_op* cop1, mgc_aggregation* agg,
                           int mode, hipStream_t s) {
  coarsen_by_points(cop2, agg, mode, 1,
                    [&](int p, const void* in, void* out, int) {
                      if (p == 8) check_call(mgc_coarse_op_Mdiag(cop1, in, out, (void*)s));
                      else
                        check_call(mgc_coarse_op_Mdir(cop1, in, out, p & 3, p < 4 ? 1 : -1,
                                                      (void*)s));
                    },
                    s);
}

// CoarsenOperator of a caller's fine operator, through its point callback
static void coarsen_linop(mgc_coarse_op* cop, const mgc_fine_linop& lop, mgc_aggregation* agg,
                          int mode, hipStream_t s) {
  const int batch = std::max(1, std::min(lop.max_nrhs, 16));
  coarsen_by_points(cop, agg, mode, batch,
                    [&](int p, const void* in, void* out, int nr) {
                      const int st = lop.apply(lop.ctx, p, MGC_F64, in, out, nr, (void*)s);
                      MGC_CHECK(st == MGC_OK, st < 0 ? st : MGC_ERR_ARG,
                                "the fine operator callback failed (point " +
                                    std::to_string(p) + ", status " + std::to_string(st) + ")");
                    },
                    s);
}

}  // namespace mgc

using namespace mgc;

extern "C" {

int mgc_coarsen_operator(mgc_coarse_op* cop, mgc_fine_op* fop, mgc_aggregation* agg, int mode,
                         void* stream) {
  return guarded([&] {
    MGC_CHECK(cop && fop && agg, MGC_ERR_ARG, "null handle");
    MGC_CHECK(mode == MGC_COARSEN_FULL || mode == MGC_COARSEN_DIRSAVE, MGC_ERR_ARG, "bad mode");
    MGC_CHECK(agg->basis_set, MGC_ERR_STATE, "aggregation basis not set");
    MGC_CHECK(fop->gauge_set, MGC_ERR_STATE, "fine gauge field not set");
    MGC_CHECK(agg->Ef == 12, MGC_ERR_SHAPE, "the aggregation must act on Dirac (12-element) sites");
    for (int d = 0; d < 4; ++d) {
      MGC_CHECK(fop->lat.L[d] == agg->fine.L[d], MGC_ERR_SHAPE, "fine lattices differ");
      MGC_CHECK(cop->lat.L[d] == agg->coarse.L[d], MGC_ERR_SHAPE, "coarse lattices differ");
    }
    MGC_CHECK(cop->N == 2 * agg->nb, MGC_ERR_SHAPE, "coarse nbasis must be 2 * nb");
    hipStream_t s = as_stream(stream);
    const int N = cop->N, Vc = cop->V;
    cplx<double>* links = nullptr;
    int* plist = nullptr;
    MGC_HIP(hipMalloc(&links, (size_t)9 * Vc * N * N * sizeof(cplx<double>)));
    MGC_HIP(hipMalloc(&plist, 9 * sizeof(int)));
    try {
      const int full[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
      const int dsave[5] = {4, 5, 6, 7, 8};
      const int* pl = mode == MGC_COARSEN_FULL ? full : dsave;
      const int np = mode == MGC_COARSEN_FULL ? 9 : 5;
      MGC_HIP(hipMemcpyAsync(plist, pl, np * sizeof(int), hipMemcpyHostToDevice, s));
      launch_coarsen(agg, fop, plist, np, links, s);
      if (mode == MGC_COARSEN_DIRSAVE) reconstruct(links, cop, -1, s);
      set_links_or_throw(cop, links, 1, s);
    } catch (...) {
      (void)hipStreamSynchronize(s);
      (void)hipFree(links);
      (void)hipFree(plist);
      throw;
    }
    MGC_HIP(hipStreamSynchronize(s));
    MGC_HIP(hipFree(links));
    MGC_HIP(hipFree(plist));
  });
}

int mgc_coarsen_set_variant(int variant) {
  return guarded([&] {
    MGC_CHECK(variant >= -1 && variant <= 1, MGC_ERR_ARG, "unknown CoarsenOperator variant");
    g_coarsen_variant = variant;
  });
}

int mgc_coarsen_operator_linop(mgc_coarse_op* cop, const mgc_fine_linop* lop,
                               mgc_aggregation* agg, int mode, void* stream) {
  return guarded([&] {
    MGC_CHECK(cop && lop && agg && lop->apply, MGC_ERR_ARG, "null handle or callback");
    MGC_CHECK(mode == MGC_COARSEN_FULL || mode == MGC_COARSEN_DIRSAVE, MGC_ERR_ARG, "bad mode");
    MGC_CHECK(mode == MGC_COARSEN_FULL || lop->gamma5_hermitian, MGC_ERR_ARG,
              "DIRSAVE rebuilds forward links by gamma5-hermiticity: the operator must be "
              "gamma5-hermitian");
    MGC_CHECK(lop->precisions & (1 << MGC_F64), MGC_ERR_PRECISION,
              "CoarsenOperator needs an operator that applies in F64");
    MGC_CHECK(lop->max_nrhs >= 1, MGC_ERR_ARG, "max_nrhs < 1");
    MGC_CHECK(agg->basis_set, MGC_ERR_STATE, "aggregation basis not set");
    MGC_CHECK(agg->Ef == 12, MGC_ERR_SHAPE, "the aggregation must act on Dirac (12-element) sites");
    for (int d = 0; d < 4; ++d) {
      MGC_CHECK(lop->dims[d] == agg->fine.L[d], MGC_ERR_SHAPE, "fine lattices differ");
      MGC_CHECK(cop->lat.L[d] == agg->coarse.L[d], MGC_ERR_SHAPE, "coarse lattices differ");
    }
    MGC_CHECK(cop->N == 2 * agg->nb, MGC_ERR_SHAPE, "coarse nbasis must be 2 * nb");
    coarsen_linop(cop, *lop, agg, mode, as_stream(stream));
  });
}

int mgc_coarsen_coarse_operator(mgc_coarse_op* cop2, mgc_coarse_op* cop1, mgc_aggregation* agg,
                                int mode, void* stream) {
  return guarded([&] {
    MGC_CHECK(cop2 && cop1 && agg, MGC_ERR_ARG, "null handle");
    MGC_CHECK(mode == MGC_COARSEN_FULL || mode == MGC_COARSEN_DIRSAVE, MGC_ERR_ARG, "bad mode");
    MGC_CHECK(agg->basis_set, MGC_ERR_STATE, "aggregation basis not set");
    MGC_CHECK(cop1->links_set, MGC_ERR_STATE, "level operator links not set");
    MGC_CHECK(cop1->prec == MGC_F64, MGC_ERR_PRECISION,
              "the operator being coarsened must be an F64 operator");
    MGC_CHECK(agg->Ef == cop1->N, MGC_ERR_SHAPE, "aggregation site elements must equal nbasis");
    for (int d = 0; d < 4; ++d) {
      MGC_CHECK(cop1->lat.L[d] == agg->fine.L[d], MGC_ERR_SHAPE, "fine lattices differ");
      MGC_CHECK(cop2->lat.L[d] == agg->coarse.L[d], MGC_ERR_SHAPE, "coarse lattices differ");
    }
    MGC_CHECK(cop2->N == 2 * agg->nb, MGC_ERR_SHAPE, "coarse nbasis must be 2 * nb");
    coarsen_coarse(cop2, cop1, agg, mode, as_stream(stream));
  });
}

int mgc_coarse_op_construct_remaining_links(mgc_coarse_op* cop, int disp_have, void* stream) {
  return guarded([&] {
    MGC_CHECK(cop, MGC_ERR_ARG, "null handle");
    MGC_CHECK(cop->links_set, MGC_ERR_STATE, "links not set");
    MGC_CHECK(disp_have == 1 || disp_have == -1, MGC_ERR_ARG, "disp_have must be +1 or -1");
    hipStream_t s = as_stream(stream);
    const int N = cop->N, V = cop->V;
    cplx<double>* links = nullptr;
    MGC_HIP(hipMalloc(&links, (size_t)9 * V * N * N * sizeof(cplx<double>)));
    const int st = mgc_coarse_op_get_links(cop, MGC_LAYOUT_FLAT, links, stream);
    if (st != MGC_OK) {
      (void)hipFree(links);
      throw Error(st, mgc_last_error());
    }
    try {
      reconstruct(links, cop, disp_have, s);
      set_links_or_throw(cop, links, cop->inv_set ? 1 : 0, s);
    } catch (...) {
      (void)hipStreamSynchronize(s);
      (void)hipFree(links);
      throw;
    }
    MGC_HIP(hipStreamSynchronize(s));
    MGC_HIP(hipFree(links));
  });
}

}  // extern "C"
