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
#include <type_traits>

#include "aggregation.hpp"
#include "coarse_op.hpp"
#include "fine_op.hpp"

namespace mgc {

constexpr int CO_SB = 2;  // fine sites per stage

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
  const int p = plist[pi];
  // coarse coordinates of this block, and lexicographic strides
  int cc[4];
  coarse.coords(xc, cc);
  const int bs[4] = {1, blk.L[0], blk.L[0] * blk.L[1], blk.L[0] * blk.L[1] * blk.L[2]};
  const int fs[4] = {1, fine.L[0], fine.L[0] * fine.L[1], fine.L[0] * fine.L[1] * fine.L[2]};
  const int cs[4] = {1, coarse.L[0], coarse.L[0] * coarse.L[1],
                     coarse.L[0] * coarse.L[1] * coarse.L[2]};
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

  // sites of this point: outer face of the block for p < 8, the whole block for p = 8
  const int mu = p & 3;
  const int disp = p < 4 ? 1 : -1;
  const int nsite = (p == 8) ? Vb : Vb / blk.L[mu];
  // face site f: the coordinates other than mu run lexicographically over the
  // face, coordinate mu is the block's last (+mu) or first (-mu) layer
  auto face_site = [&](int f) {
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
  };

  for (int s0 = 0; s0 < nsite; s0 += CO_SB) {
    // ---- stage u rows of the SB sites
    for (int idx = tid; idx < CO_SB * ROW; idx += THREADS) {
      const int s = idx / ROW, r = idx - s * ROW;
      // past the face end: load block site 0 (in bounds) and stage zero
      const bool ok = s0 + s < nsite;
      const int xb = ok ? face_site(s0 + s) : 0;
      const cplx<double> v = basis[((size_t)xc * Vb + xb) * ROW + r];
      (&sU[s][0][0])[r] = ok ? v : cplx<double>{0.0, 0.0};
    }
    // ---- W items (s, l', c', colour i)
    for (int it = tid; it < CO_SB * 2 * NB * 3; it += THREADS) {
      const int s = it / (2 * NB * 3);
      int r = it - s * 2 * NB * 3;
      const int lw = r / (NB * 3);
      r -= lw * NB * 3;
      const int cp = r / 3, ci = r - cp * 3;
      cplx<double> w[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};  // spins 0..3 at colour ci
      // the body is instantiated per chirality LW so that every spin index is
      // a compile-time constant (no indexed private arrays)
      auto build_w = [&](auto lw_c) {
        constexpr int LW = decltype(lw_c)::value;
        const int xb = face_site(s0 + s);
        // block and fine coordinates of this site
        int bc[4], x = 0;
        {
          int r = xb;
#pragma unroll
          for (int d = 0; d < 4; ++d) {
            bc[d] = r % blk.L[d];
            r /= blk.L[d];
            x += (cc[d] * blk.L[d] + bc[d]) * fs[d];
          }
        }
        // hopping contributions; unrolled so that the spin indices below are
        // compile-time constants (no indexed private arrays)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (p != 8 && q != p) continue;
          const int qm = q & 3;
          const int disp = q < 4 ? 1 : -1;
          const bool inner = (unsigned)(bc[qm] + disp) < (unsigned)blk.L[qm];
          // self: inner hops only, by the reference's coordinate mask
          // (x_mu mod b != b-1 for +mu, != 0 for -mu)
          if (p == 8 && !inner) continue;
          // fine neighbour y = x + disp*mu (periodic) and its panel row
          const int fq = cc[qm] * blk.L[qm] + bc[qm] + disp;
          const int y = x + disp * fs[qm] -
                        (fq < 0 ? -fine.L[qm] * fs[qm] : fq >= fine.L[qm] ? fine.L[qm] * fs[qm] : 0);
          int ry;
          if (inner) {
            ry = xc * Vb + xb + disp * bs[qm];
          } else {
            const int cq = cc[qm] + disp;
            const int xn = xc + disp * cs[qm] -
                           (cq < 0 ? -coarse.L[qm] * cs[qm]
                                   : cq >= coarse.L[qm] ? coarse.L[qm] * cs[qm] : 0);
            ry = xn * Vb + xb - disp * (blk.L[qm] - 1) * bs[qm];
          }
          const double sg = q < 4 ? 1.0 : -1.0;
          cplx<double> L[3];  // row ci of the link matrix
          if (q < 4) {
            const cplx<double>* U = gauge + ((size_t)qm * Vf + x) * 9;
#pragma unroll
            for (int j = 0; j < 3; ++j) L[j] = U[ci * 3 + j];
          } else {
            const cplx<double>* U = gauge + ((size_t)qm * Vf + y) * 9;
#pragma unroll
            for (int j = 0; j < 3; ++j) L[j] = {U[j * 3 + ci].re, -U[j * 3 + ci].im};
          }
          const cplx<double>* uy = basis + (size_t)ry * ROW;
#pragma unroll
          for (int sl = 0; sl < 2; ++sl) {
            const int sp = 2 * LW + sl;
            cplx<double> chi = {0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 3; ++j) cmac(chi, L[j], uy[(sp * 3 + j) * NB + cp]);
            // same-chirality part: -1/2 chi
            w[sp].re -= 0.5 * chi.re;
            w[sp].im -= 0.5 * chi.im;
            // gamma image: spin s2 = col(sp), + sg/2 * phase(s2) * chi
            const int s2 = G.col[qm][sp];
            const cplx<double> ph = {(double)G.re[qm][s2], (double)G.im[qm][s2]};
            const cplx<double> t = cmul(ph, chi);
            w[s2].re += 0.5 * sg * t.re;
            w[s2].im += 0.5 * sg * t.im;
          }
        }
        if (p == 8) {  // diagonal (mass + clover) term, block-diagonal in chirality
          const cplx<double>* D = diag + ((size_t)x * 2 + LW) * 36;
          const cplx<double>* ux = basis + ((size_t)xc * Vb + xb) * ROW;
#pragma unroll
          for (int sl = 0; sl < 2; ++sl) {
            cplx<double> a = {0.0, 0.0};
#pragma unroll
            for (int b = 0; b < 6; ++b) cmac(a, D[(sl * 3 + ci) * 6 + b], ux[(6 * LW + b) * NB + cp]);
            w[2 * LW + sl].re += a.re;
            w[2 * LW + sl].im += a.im;
          }
        }
      };
      if (s0 + s < nsite) {
        if (lw == 0) build_w(std::integral_constant<int, 0>{});
        else build_w(std::integral_constant<int, 1>{});
      }
#pragma unroll
      for (int sp = 0; sp < 4; ++sp) sW[s][lw][sp * 3 + ci][cp] = w[sp];
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
  cplx<double>* Ap = links + ((size_t)p * Vc + xc) * N * N;
#pragma unroll
  for (int i = 0; i < TC; ++i)
#pragma unroll
    for (int j = 0; j < TC; ++j)
      Ap[(size_t)(k * NB + cb * TC + i) * N + l * NB + pb * TC + j] = acc[i][j];
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

static void launch_coarsen(const mgc_aggregation* agg, const mgc_fine_op* fop, const int* plist,
                           int np, cplx<double>* links, hipStream_t s) {
  switch (agg->nb) {
    case 2: launch_coarsen_t<2, 1>(agg, fop, plist, np, links, s); return;
    case 4: launch_coarsen_t<4, 1>(agg, fop, plist, np, links, s); return;
    case 8: launch_coarsen_t<8, 2>(agg, fop, plist, np, links, s); return;
    case 12: launch_coarsen_t<12, 3>(agg, fop, plist, np, links, s); return;
    case 16: launch_coarsen_t<16, 2>(agg, fop, plist, np, links, s); return;
    case 24: launch_coarsen_t<24, 3>(agg, fop, plist, np, links, s); return;
    case 32: launch_coarsen_t<32, 4>(agg, fop, plist, np, links, s); return;
  }
  throw Error(MGC_ERR_SHAPE, "CoarsenOperator supports nb in {2, 4, 8, 12, 16, 24, 32}");
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

// links[p][xc][j][i] (=, or += when accumulate) proj[xc][j]
__global__ void k_scatter_column(cplx<double>* __restrict__ links, const cplx<double>* __restrict__ proj,
                                 int Vc, int N, int p, int i, int accumulate) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= (size_t)Vc * N) return;
  const size_t xc = t / N, j = t % N;
  cplx<double>& dst = links[(((size_t)p * Vc + xc) * N + j) * N + i];
  const cplx<double> v = proj[t];
  if (accumulate) dst = cadd(dst, v);
  else dst = v;
}

static void check_call(int st) {
  if (st != MGC_OK) throw Error(st, mgc_last_error());
}

// Galerkin product of a coarse operator on the next level's basis, in the
// inner/outer-mask form of include/CoarsenedMatrixBaseline.h:609-675 (one
// doubled basis vector at a time: Mdir_p v_i, masked projections, links column
// i), FP64.
static void coarsen_coarse(mgc_coarse_op* cop2, mgc_coarse_op* cop1, mgc_aggregation* agg,
                           int mode, hipStream_t s) {
  const int N2 = cop2->N, Vc2 = cop2->V, Vf = agg->Vf, Ef = agg->Ef, nb = agg->nb;
  cplx<double>* links = nullptr;
  cplx<double>* phi = nullptr;
  cplx<double>* mphi = nullptr;
  cplx<double>* proj = nullptr;
  auto cleanup = [&] {
    (void)hipStreamSynchronize(s);
    void* bufs[] = {links, phi, mphi, proj};
    for (void* b : bufs)
      if (b) (void)hipFree(b);
  };
  try {
    MGC_HIP(hipMalloc(&links, (size_t)9 * Vc2 * N2 * N2 * sizeof(cplx<double>)));
    MGC_HIP(hipMalloc(&phi, (size_t)Vf * Ef * sizeof(cplx<double>)));
    MGC_HIP(hipMalloc(&mphi, (size_t)Vf * Ef * sizeof(cplx<double>)));
    MGC_HIP(hipMalloc(&proj, (size_t)Vc2 * N2 * sizeof(cplx<double>)));
    fill_zero(links, (size_t)9 * Vc2 * N2 * N2 * sizeof(cplx<double>), s);
    const unsigned g = (unsigned)(((size_t)Vc2 * N2 + 255) / 256);
    for (int i = 0; i < N2; ++i) {
      const int l = i / nb, c = i % nb;
      agg_extract_basis(agg, c, l, phi, MGC_F64, s);
      for (int p = 0; p < 9; ++p) {
        if (p == 8) check_call(mgc_coarse_op_Mdiag(cop1, phi, mphi, (void*)s));
        else check_call(mgc_coarse_op_Mdir(cop1, phi, mphi, p & 3, p < 4 ? 1 : -1, (void*)s));
        // every point's inner part feeds the self link; DIRSAVE rebuilds the
        // forward links instead of projecting them
        if (p < 8 && !(mode == MGC_COARSEN_DIRSAVE && p < 4)) {
          agg_transfer(agg, mphi, proj, MGC_F64, 1, true, 2, p, s);
          k_scatter_column<<<g, 256, 0, s>>>(links, proj, Vc2, N2, p, i, 0);
          MGC_LAUNCH_CHECK();
        }
        agg_transfer(agg, mphi, proj, MGC_F64, 1, true, 1, p, s);
        k_scatter_column<<<g, 256, 0, s>>>(links, proj, Vc2, N2, 8, i, 1);
        MGC_LAUNCH_CHECK();
      }
    }
    if (mode == MGC_COARSEN_DIRSAVE) reconstruct(links, cop2, -1, s);
    set_links_or_throw(cop2, links, 1, s);
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
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
