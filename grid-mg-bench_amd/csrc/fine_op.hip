// Fine Wilson-clover operator on the GPU (the caller-side input of
// CoarsenOperator and the smoother's operator), and the SU(3) gauge
// generator for synthetic inputs.
//
// Reference use: the fine LinearOperator's OpDir / OpDiag / Op called by
// CoarsenOperator (include/CoarsenedMatrixBaseline.h:628-633,
// include/CoarsenedMatrixUpstream.h:843-844, 890-891) and by the multigrid
// smoother (include/MultigridPreconditioner.h:295-304); Wilson-clover with
// m = -0.1, csw = 1.0 as in src/Test_coarse_red_black.cc:99-100.
#include "fine_op.hpp"

#include <vector>

namespace mgc {

// ---------------------------------------------------------------------------
// gauge generator: U = exp(i eps H), H = traceless Hermitian part of a 3x3
// complex Gaussian (entries from the counter-based generator, element index
// (mu*V + x)*9 + 3i + j).  Exponential by scaling and squaring of a degree-18
// Taylor polynomial in FP64.
// ---------------------------------------------------------------------------
__host__ __device__ uint64_t splitmix64_g(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ inline cplx<double> gauss_at(uint64_t key, uint64_t i) {
  auto unif = [&](uint64_t ctr) {
    uint64_t b = splitmix64_g(key ^ splitmix64_g(ctr));
    return ((double)(b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  const double u1 = unif(2 * i), u2 = unif(2 * i + 1);
  const double r = sqrt(-log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  return {r * c, r * s};
}

__device__ inline void m3mul(const cplx<double>* a, const cplx<double>* b, cplx<double>* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx<double> s = {0.0, 0.0};
      for (int k = 0; k < 3; ++k) cmac(s, a[i * 3 + k], b[k * 3 + j]);
      c[i * 3 + j] = s;
    }
}

__global__ void k_random_gauge(cplx<double>* __restrict__ U, size_t nmat, double eps,
                               uint64_t key) {
  size_t m = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (m >= nmat) return;
  cplx<double> a[9], X[9], E[9], T[9];
  for (int k = 0; k < 9; ++k) a[k] = gauss_at(key, m * 9 + k);
  // H = (a + a^dag)/2 - tr/3 ; X = i eps H
  double tr = 0.0;
  for (int i = 0; i < 3; ++i) tr += a[i * 4].re;
  tr /= 3.0;
  double nrm = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx<double> h = {0.5 * (a[i * 3 + j].re + a[j * 3 + i].re),
                        0.5 * (a[i * 3 + j].im - a[j * 3 + i].im)};
      if (i == j) h.re -= tr;
      X[i * 3 + j] = {-eps * h.im, eps * h.re};
      nrm += X[i * 3 + j].re * X[i * 3 + j].re + X[i * 3 + j].im * X[i * 3 + j].im;
    }
  nrm = sqrt(nrm);
  int sq = 0;
  while (nrm > 0.25) { nrm *= 0.5; ++sq; }
  const double f = ldexp(1.0, -sq);
  for (int k = 0; k < 9; ++k) { X[k].re *= f; X[k].im *= f; }
  // Horner: E = 1 + X(1 + X/2 (1 + X/3 (...)))
  for (int k = 0; k < 9; ++k) E[k] = {(k % 4 == 0) ? 1.0 : 0.0, 0.0};
  for (int n = 18; n >= 1; --n) {
    m3mul(X, E, T);
    for (int k = 0; k < 9; ++k) {
      E[k] = {T[k].re / n, T[k].im / n};
      if (k % 4 == 0) E[k].re += 1.0;
    }
  }
  for (int s = 0; s < sq; ++s) {
    m3mul(E, E, T);
    for (int k = 0; k < 9; ++k) E[k] = T[k];
  }
  for (int k = 0; k < 9; ++k) U[m * 9 + k] = E[k];
}

// ---------------------------------------------------------------------------
// clover term: (4+m) + C(x) as two 6x6 chirality blocks per site (FP64)
// ---------------------------------------------------------------------------
__device__ inline void m3dag(const cplx<double>* a, cplx<double>* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[i * 3 + j] = {a[j * 3 + i].re, -a[j * 3 + i].im};
}
__device__ inline void m3load(const cplx<double>* g, int V, int mu, int x, cplx<double>* o) {
  const cplx<double>* s = g + ((size_t)mu * V + x) * 9;
  for (int k = 0; k < 9; ++k) o[k] = s[k];
}

__global__ void k_clover(const cplx<double>* __restrict__ U, Lat4 lat, double mass, double csw,
                         cplx<double>* __restrict__ diag) {
  const int V = lat.volume();
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= V) return;
  const GammaTab G = gamma_table();
  cplx<double> D[144];
  for (int i = 0; i < 144; ++i) D[i] = {0.0, 0.0};
  for (int i = 0; i < 12; ++i) D[i * 13].re = 4.0 + mass;
  if (csw != 0.0) {
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        const int xpm = lat.shift(x, mu, +1), xpn = lat.shift(x, nu, +1);
        const int xmm = lat.shift(x, mu, -1), xmn = lat.shift(x, nu, -1);
        cplx<double> Q[9], A[9], B[9], T1[9], T2[9];
        for (int k = 0; k < 9; ++k) Q[k] = {0.0, 0.0};
        // leaf 1
        m3load(U, V, mu, x, A); m3load(U, V, nu, xpm, B); m3mul(A, B, T1);
        m3load(U, V, mu, xpn, A); m3dag(A, B); m3mul(T1, B, T2);
        m3load(U, V, nu, x, A); m3dag(A, B); m3mul(T2, B, T1);
        for (int k = 0; k < 9; ++k) Q[k] = cadd(Q[k], T1[k]);
        // leaf 2
        m3load(U, V, mu, lat.shift(xmm, nu, +1), A); m3dag(A, B);
        m3load(U, V, nu, x, A); m3mul(A, B, T1);
        m3load(U, V, nu, xmm, A); m3dag(A, B); m3mul(T1, B, T2);
        m3load(U, V, mu, xmm, A); m3mul(T2, A, T1);
        for (int k = 0; k < 9; ++k) Q[k] = cadd(Q[k], T1[k]);
        // leaf 3
        const int xmmmn = lat.shift(xmm, nu, -1);
        m3load(U, V, mu, xmm, A); m3dag(A, T1);
        m3load(U, V, nu, xmmmn, A); m3dag(A, B); m3mul(T1, B, T2);
        m3load(U, V, mu, xmmmn, A); m3mul(T2, A, T1);
        m3load(U, V, nu, xmn, A); m3mul(T1, A, T2);
        for (int k = 0; k < 9; ++k) Q[k] = cadd(Q[k], T2[k]);
        // leaf 4
        m3load(U, V, nu, xmn, A); m3dag(A, T1);
        m3load(U, V, mu, xmn, A); m3mul(T1, A, T2);
        m3load(U, V, nu, lat.shift(xmn, mu, +1), A); m3mul(T2, A, T1);
        m3load(U, V, mu, x, A); m3dag(A, B); m3mul(T1, B, T2);
        for (int k = 0; k < 9; ++k) Q[k] = cadd(Q[k], T2[k]);
        // F = (Q - Q^dag)/(8i)
        cplx<double> F[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const double zr = Q[i * 3 + j].re - Q[j * 3 + i].re;
            const double zi = Q[i * 3 + j].im + Q[j * 3 + i].im;
            F[i * 3 + j] = {zi / 8.0, -zr / 8.0};
          }
        // sigma_{mu nu}[r][c] = (i/2)(g_mu g_nu - g_nu g_mu)[r][c]
        for (int r = 0; r < 4; ++r) {
          // (g_mu g_nu)[r][c] nonzero at c = col_nu[col_mu[r]]
          const int k1 = G.col[mu][r], c1 = G.col[nu][k1];
          const cplx<double> p1 = cmul(cplx<double>{G.re[mu][r], G.im[mu][r]},
                                       cplx<double>{G.re[nu][k1], G.im[nu][k1]});
          const int k2 = G.col[nu][r], c2 = G.col[mu][k2];
          const cplx<double> p2 = cmul(cplx<double>{G.re[nu][r], G.im[nu][r]},
                                       cplx<double>{G.re[mu][k2], G.im[mu][k2]});
          // both products land on the same column for mu != nu
          const cplx<double> s = {0.5 * -(p1.im - p2.im), 0.5 * (p1.re - p2.re)};  // (i/2)(p1-p2)
          const int c = c1;
          (void)c2;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              const cplx<double> v = cmul(s, F[i * 3 + j]);
              D[(r * 3 + i) * 12 + c * 3 + j].re += -0.5 * csw * v.re;
              D[(r * 3 + i) * 12 + c * 3 + j].im += -0.5 * csw * v.im;
            }
        }
      }
  }
  // store the two chirality blocks
  for (int l = 0; l < 2; ++l)
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b)
        diag[(((size_t)x * 2 + l) * 6 + a) * 6 + b] = D[(6 * l + a) * 12 + 6 * l + b];
}

__global__ void k_antiperiodic(cplx<double>* __restrict__ U, Lat4 lat) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= lat.volume()) return;
  int c[4];
  lat.coords(x, c);
  if (c[3] != lat.L[3] - 1) return;
  cplx<double>* u = U + ((size_t)3 * lat.volume() + x) * 9;
  for (int k = 0; k < 9; ++k) { u[k].re = -u[k].re; u[k].im = -u[k].im; }
}

// ---------------------------------------------------------------------------
// Wilson-clover apply: thread per (site, rhs); vectors [V][nrhs][12]; links
// U [4][9][V] and packed hermitian diagonal blocks diag [2][18][V], site index
// fastest, so a wavefront's link and clover loads are contiguous.  mask bit p < 8 selects
// hopping point p, bit 8 the diagonal; dag applies g5 D g5 (= D^dag)
// ---------------------------------------------------------------------------
// Workgroup -> sites.  A workgroup covers a tile of S consecutive sites (all
// nrhs of each).  Consecutive workgroups are dispatched to the eight XCDs in
// turn, each with its own L2, so with xcd = 1 XCD k receives the time slices
// [k*TL, (k+1)*TL) and walks them t fastest, then z within strips of ZS, then
// the S-site groups g of an xy plane, then the z strips.  The t, z and y
// neighbours a tile reads were then fetched into the same L2 by a workgroup
// a few steps earlier (1, TL and TL*ZS steps).  With xcd = 0 the tiles are
// simply consecutive (any lattice).
struct WilsonTiles {
  int S;       // sites per workgroup
  int xcd;     // 1: XCD-local ordering below, 0: tile b = blockIdx
  int G, TL, ZS, L2, P;  // groups per xy plane, t slices per XCD, z strip, L_z, L_x*L_y
};

__device__ __forceinline__ int wilson_tile_site(const WilsonTiles& w, int b) {
  if (!w.xcd) return b * w.S;
  const int i = b >> 3;
  const int tl = i % w.TL;
  int r = i / w.TL;
  const int zi = r % w.ZS;
  r /= w.ZS;
  const int g = r % w.G;
  const int zs = r / w.G;
  const int t = (b & 7) * w.TL + tl;
  const int z = zs * w.ZS + zi;
  return g * w.S + w.P * (z + w.L2 * t);
}

// Clover term of one (site, rhs): acc += D(x) v with the packed hermitian
// blocks of chirality l; dk(l, slot) returns packed slot `slot` of block l,
// sp(e) element e of the site's spinor.
template <typename R, typename DK, typename SP>
__device__ __forceinline__ void wilson_diag_apply(cplx<R>* acc, DK&& dk, SP&& sp) {
#pragma unroll
  for (int l = 0; l < 2; ++l) {
    cplx<R> v[6];
#pragma unroll
    for (int b = 0; b < 6; ++b) v[b] = sp(6 * l + b);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const cplx<R> dd = dk(l, a);  // (d_{2a}, d_{2a+1})
      acc[6 * l + 2 * a].re += dd.re * v[2 * a].re;
      acc[6 * l + 2 * a].im += dd.re * v[2 * a].im;
      acc[6 * l + 2 * a + 1].re += dd.im * v[2 * a + 1].re;
      acc[6 * l + 2 * a + 1].im += dd.im * v[2 * a + 1].im;
    }
    int slot = 3;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a + 1; b < 6; ++b, ++slot) {
        const cplx<R> u = dk(l, slot);
        cmac(acc[6 * l + a], u, v[b]);       // upper: D_ab v_b
        cmac_conj(acc[6 * l + b], u, v[a]);  // lower: conj(D_ab) v_a
      }
  }
}

// The clover term with the blocks read from the site-strided SoA field
// (packed slot s of block l at diag[(l * 18 + s) * sv + x]).
template <typename R, typename SP>
__device__ __forceinline__ void wilson_diag_term(cplx<R>* acc, const cplx<R>* __restrict__ diag,
                                                 int x, int sv, SP&& sp) {
  wilson_diag_apply<R>(acc, [&](int l, int s) { return diag[((size_t)l * 18 + s) * sv + x]; },
                       sp);
}

// Hopping term p with the rank-2 spin projector: for the upper spins s = 0, 1
// (g_mu maps them to s' = col(s) in {2, 3}),
//   h_s = psi_s - sg ph(s) psi_s',  (1 - sg g_mu) psi at s' = -sg ph(s') h_s,
// so only the half spinor h is multiplied by the SU(3) link.  p is a
// compile-time constant at every call (unrolled loops), so every spin index
// is too.  sp(e) returns element e of the neighbour's spinor psi(y); lk(k)
// returns element k (row-major) of the link U_mu at x (p < 4) or at y.
template <typename R, typename LK, typename SP>
__device__ __forceinline__ void wilson_hop_apply(cplx<R>* acc, const GammaTab& G, int p, LK&& lk,
                                                 SP&& sp) {
  const int mu = p & 3;
  const R sg = p < 4 ? R(1) : R(-1);
  cplx<R> h[2][3];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int cs = G.col[mu][s];
    const cplx<R> ph = {(R)(sg * G.re[mu][s]), (R)(sg * G.im[mu][s])};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const cplx<R> a = sp(s * 3 + j), b = sp(cs * 3 + j);
      const cplx<R> t = cmul(ph, b);
      h[s][j] = {a.re - t.re, a.im - t.im};
    }
  }
  // chi_s = L h_s, one colour row i at a time; each chi_s(i) goes into acc
  // as soon as it is formed (every acc element gets one update per hop)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    cplx<R> L[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const cplx<R> l = lk(p < 4 ? i * 3 + j : j * 3 + i);
      L[j] = p < 4 ? l : cplx<R>{l.re, -l.im};
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      cplx<R> chi = {R(0), R(0)};
#pragma unroll
      for (int j = 0; j < 3; ++j) cmac(chi, L[j], h[s][j]);
      const int cs = G.col[mu][s];
      // lower spin cs: -1/2 * (-sg ph(cs)) chi_s = (sg/2) ph(cs) chi_s
      const cplx<R> ph = {(R)(0.5 * sg * G.re[mu][cs]), (R)(0.5 * sg * G.im[mu][cs])};
      acc[s * 3 + i].re -= R(0.5) * chi.re;
      acc[s * 3 + i].im -= R(0.5) * chi.im;
      const cplx<R> t = cmul(ph, chi);
      acc[cs * 3 + i].re += t.re;
      acc[cs * 3 + i].im += t.im;
    }
  }
}

// One colour row c of the clover and hopping terms, for a thread that holds
// the four spin components of colour c of one (site, rhs): acc[s] is element
// s * 3 + c.  Every element gets the same updates in the same order as in
// wilson_diag_apply and wilson_hop_apply, so the results are bit-identical;
// c is uniform over a wavefront, so the tests on it are scalar branches.
template <typename R, typename DK, typename SP>
__device__ __forceinline__ void wilson_diag_apply_row(cplx<R>* acc, int c, DK&& dk, SP&& sp) {
  const int e0 = c, e1 = 3 + c;  // this thread's elements of each block
#pragma unroll
  for (int l = 0; l < 2; ++l) {
    cplx<R>& A0 = acc[2 * l];
    cplx<R>& A1 = acc[2 * l + 1];
    cplx<R> v[6];
#pragma unroll
    for (int b = 0; b < 6; ++b) v[b] = sp(6 * l + b);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const bool lo0 = 2 * a == e0, hi0 = 2 * a + 1 == e0;
      const bool lo1 = 2 * a == e1, hi1 = 2 * a + 1 == e1;
      if (!(lo0 || hi0 || lo1 || hi1)) continue;
      const cplx<R> dd = dk(l, a);
      if (lo0) { A0.re += dd.re * v[2 * a].re; A0.im += dd.re * v[2 * a].im; }
      if (hi0) { A0.re += dd.im * v[2 * a + 1].re; A0.im += dd.im * v[2 * a + 1].im; }
      if (lo1) { A1.re += dd.re * v[2 * a].re; A1.im += dd.re * v[2 * a].im; }
      if (hi1) { A1.re += dd.im * v[2 * a + 1].re; A1.im += dd.im * v[2 * a + 1].im; }
    }
    int slot = 3;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a + 1; b < 6; ++b, ++slot) {
        const bool u0 = a == e0, u1 = a == e1, w0 = b == e0, w1 = b == e1;
        if (!(u0 || u1 || w0 || w1)) continue;
        const cplx<R> u = dk(l, slot);
        if (u0) cmac(A0, u, v[b]);
        if (u1) cmac(A1, u, v[b]);
        if (w0) cmac_conj(A0, u, v[a]);
        if (w1) cmac_conj(A1, u, v[a]);
      }
  }
}
// lk(j): entry j of row c of U_mu(x) (p < 4) or of column c of U_mu(y)
template <typename R, typename LK, typename SP>
__device__ __forceinline__ void wilson_hop_apply_row(cplx<R>* acc, const GammaTab& G, int p,
                                                     LK&& lk, SP&& sp) {
  const int mu = p & 3;
  const R sg = p < 4 ? R(1) : R(-1);
  cplx<R> h[2][3];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int cs = G.col[mu][s];
    const cplx<R> ph = {(R)(sg * G.re[mu][s]), (R)(sg * G.im[mu][s])};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const cplx<R> a = sp(s * 3 + j), b = sp(cs * 3 + j);
      const cplx<R> t = cmul(ph, b);
      h[s][j] = {a.re - t.re, a.im - t.im};
    }
  }
  cplx<R> L[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const cplx<R> l = lk(j);
    L[j] = p < 4 ? l : cplx<R>{l.re, -l.im};
  }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    cplx<R> chi = {R(0), R(0)};
#pragma unroll
    for (int j = 0; j < 3; ++j) cmac(chi, L[j], h[s][j]);
    const int cs = G.col[mu][s];
    const cplx<R> ph = {(R)(0.5 * sg * G.re[mu][cs]), (R)(0.5 * sg * G.im[mu][cs])};
    acc[s].re -= R(0.5) * chi.re;
    acc[s].im -= R(0.5) * chi.im;
    const cplx<R> t = cmul(ph, chi);
    acc[cs].re += t.re;
    acc[cs].im += t.im;
  }
}

// The hopping term with the link read from the site-strided SoA field
// (element k of U_mu at x in U[(mu * 9 + k) * sv + x]).
template <typename R, typename SP>
__device__ __forceinline__ void wilson_hop_term(cplx<R>* acc, const cplx<R>* __restrict__ U,
                                                int x, int y, int sv, const GammaTab& G, int p,
                                                SP&& sp) {
  const cplx<R>* u = U + (size_t)(p & 3) * 9 * sv + (p < 4 ? x : y);
  wilson_hop_apply<R>(acc, G, p, [&](int k) { return u[(size_t)k * sv]; }, sp);
}

template <typename R>
__global__ __launch_bounds__(256) void k_wilson(const cplx<R>* __restrict__ U,
                                                const cplx<R>* __restrict__ diag,
                                                const cplx<R>* __restrict__ in,
                                                cplx<R>* __restrict__ out,
                                                const int* __restrict__ nbr, int V, int sv,
                                                int nrhs, unsigned mask, int dag,
                                                WilsonTiles wt) {
  const int x = wilson_tile_site(wt, (int)blockIdx.x) + (int)threadIdx.x / nrhs;
  const int rr = (int)threadIdx.x % nrhs;
  if (x >= V) return;
  const GammaTab G = gamma_table();
  cplx<R> acc[12];
#pragma unroll
  for (int e = 0; e < 12; ++e) acc[e] = {R(0), R(0)};
  auto ld = [&](int site, int e) {
    cplx<R> v = in[((size_t)site * nrhs + rr) * 12 + e];
    if (dag && e >= 6) { v.re = -v.re; v.im = -v.im; }
    return v;
  };
  if (mask & 0x100u) wilson_diag_term<R>(acc, diag, x, sv, [&](int e) { return ld(x, e); });
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    if (!((mask >> p) & 1u)) continue;
    const int y = nbr[(size_t)x * 8 + p];
    wilson_hop_term<R>(acc, U, x, y, sv, G, p, [&](int e) { return ld(y, e); });
  }
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    cplx<R> v = acc[e];
    if (dag && e >= 6) { v.re = -v.re; v.im = -v.im; }
    out[((size_t)x * nrhs + rr) * 12 + e] = v;
  }
}

// LDS position of chunk e of staged record q (records padded to RS chunks)
template <int CPS, int RS>
__device__ __forceinline__ int rec_pos(int q, int e) {
  static_assert(RS > CPS, "padded records");
  return q * RS + e;
}
// Cooperative copy of T records of CPS 16-byte chunks into LDS: chunk
// c = t + k*T of record q = c / CPS (tile site q / nrhs, rhs q % nrhs), part
// e = c % CPS, is read from record (sites[q / nrhs], rhs) of the field and
// stored at position rec_pos(q, e).
template <int CPS, int RS>
__device__ __forceinline__ void wilson_copy(uint4* s_rec, const uint4* __restrict__ in4,
                                            const int* sites, int t, int T, int nrhs) {
#pragma unroll
  for (int k = 0; k < CPS; ++k) {
    const int c = t + k * T, q = c / CPS, e = c - q * CPS, s = q / nrhs, r = q - s * nrhs;
    s_rec[rec_pos<CPS, RS>(q, e)] = in4[((size_t)sites[s] * nrhs + r) * CPS + e];
  }
}

// LDS-gathered form of the apply.  Each (site, rhs) spinor record of 12
// complex is contiguous in memory, so a workgroup takes a hop's neighbour
// records for all its threads as one cooperative copy of 16-byte chunks in
// which consecutive lanes read consecutive chunks (coalesced, unlike a
// thread reading its own record with a stride of one record per lane).  The
// chunks go through registers into LDS, each thread then reads its own
// record there.  Hops run self first, then 0..7; the neighbour sites of the
// next hop are staged in LDS one hop ahead, and a hop's link loads are
// issued before its record copy.  Records are padded by
// one chunk (13 for FP64, 7 for FP32), so that the records that
// neighbouring lanes read fall in different banks.  The result leaves through LDS in the
// same coalesced way.  Threads past the lattice copy site 0 and do no
// arithmetic.  Per-thread arithmetic is that of k_wilson.
template <typename R>
__global__ __launch_bounds__(256) void k_wilson_lds(const cplx<R>* __restrict__ U,
                                                    const cplx<R>* __restrict__ diag,
                                                    const cplx<R>* __restrict__ in,
                                                    cplx<R>* __restrict__ out,
                                                    const int* __restrict__ nbr, int V, int sv,
                                                    int nrhs, unsigned mask, int dag,
                                                    WilsonTiles wt) {
  constexpr int CPS = 12 * (int)sizeof(cplx<R>) / 16;  // 16-byte chunks per record
  constexpr int RS = CPS + 1;  // record stride in chunks
  __shared__ uint4 s_rec[256 * RS];
  __shared__ int s_y[2][256];
  const int T = (int)blockDim.x;  // S sites x nrhs
  const int S = T / nrhs;
  const int t = (int)threadIdx.x;
  const int x0 = wilson_tile_site(wt, (int)blockIdx.x);
  const int x = x0 + t / nrhs;  // thread t is record t: site t / nrhs, rhs t % nrhs
  const bool live = x < V;
  const GammaTab G = gamma_table();

  auto pos = [&](int q, int e) { return rec_pos<CPS, RS>(q, e); };
  auto hop_site = [&](int s, int h) {  // h = 8: the site itself
    const int xs = x0 + s;
    if (xs >= V) return 0;
    return h == 8 ? xs : nbr[(size_t)xs * 8 + h];
  };
  // next active hop after stage i (stages: 0 = self, 1 + p = hop p), -1 if none
  auto next_hop = [&](int i) {
    for (int k = i + 1; k < 9; ++k) {
      const int h = k == 0 ? 8 : k - 1;
      if ((mask >> h) & 1u) return h;
    }
    return -1;
  };
  const uint4* in4 = reinterpret_cast<const uint4*>(in);
  auto rec = [&](int e) {  // element e of this thread's staged record
    cplx<R> v = reinterpret_cast<const cplx<R>*>(&s_rec[pos(t, e * CPS / 12)])[(e * CPS) % 12 / 6];
    if (dag && e >= 6) { v.re = -v.re; v.im = -v.im; }
    return v;
  };

  cplx<R> acc[12];
#pragma unroll
  for (int e = 0; e < 12; ++e) acc[e] = {R(0), R(0)};
  int cb = 0;  // s_y buffer of the current hop
  const int h0 = next_hop(-1);
  if (h0 >= 0) {
    if (t < S) s_y[0][t] = hop_site(t, h0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int h = i == 0 ? 8 : i - 1;
    if (!((mask >> h) & 1u)) continue;
    const int hn = next_hop(i);
    __syncthreads();  // every thread is done with the previous hop's records
    // The loads this stage needs besides the records (this hop's link, 9
    // complex, or in the self stage the first 9 clover slots, and the next
    // hop's neighbour sites) are issued before the record copy, so that
    // their latencies overlap its own instead of following the barrier:
    // FP32 8-12% faster, FP64 3% (252 registers, the same two wavefronts per
    // SIMD as without).
    const int y = s_y[cb][t / nrhs];  // written before the previous stage's last barrier
    cplx<R> lk[9];  // the hop's link, or in the self stage clover slots 0..8 of block 0
    if (live) {
      const cplx<R>* u = h == 8 ? diag + x : U + (size_t)(h & 3) * 9 * sv + (h < 4 ? x : y);
#pragma unroll
      for (int k = 0; k < 9; ++k) lk[k] = u[(size_t)k * sv];
    }
    const int ynext = hn >= 0 && t < S ? hop_site(t, hn) : 0;
    wilson_copy<CPS, RS>(s_rec, in4, s_y[cb], t, T, nrhs);
    if (hn >= 0 && t < S) s_y[cb ^ 1][t] = ynext;
    __syncthreads();
    cb ^= 1;
    if (live) {
      if (h == 8)
        wilson_diag_apply<R>(acc, [&](int l, int s) {
          return l == 0 && s < 9 ? lk[s] : diag[((size_t)l * 18 + s) * sv + x];
        }, rec);
      else
        wilson_hop_apply<R>(acc, G, h, [&](int k) { return lk[k]; }, rec);
    }
  }
  // result: records through LDS, then coalesced stores of whole chunks
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    cplx<R> v = acc[e];
    if (dag && e >= 6) { v.re = -v.re; v.im = -v.im; }
    reinterpret_cast<cplx<R>*>(&s_rec[pos(t, e * CPS / 12)])[(e * CPS) % 12 / 6] = v;
  }
  __syncthreads();
  uint4* out4 = reinterpret_cast<uint4*>(out);
#pragma unroll
  for (int k = 0; k < CPS; ++k) {
    const int c = t + k * T, q = c / CPS, e = c - q * CPS, s = q / nrhs, r = q - s * nrhs;
    if (x0 + s < V) out4[((size_t)(x0 + s) * nrhs + r) * CPS + e] = s_rec[pos(q, e)];
  }
}

// Multi-rhs form with the next hop's loads in flight during the current
// hop's arithmetic (nrhs >= 5: LPT = 1 link entry per thread for nrhs >= 9,
// when a tile has at most 28 sites, else LPT = 2, at most 51 sites; the
// hop's 9 x S link entries are loaded by the first threads).  The tile's 8 x S
// neighbour sites are read once, coalesced, into LDS.  Each stage then
// stores the records and link entries that were loaded during the previous
// stage into LDS, issues the loads of the next stage into registers (six
// 16-byte record chunks and one link entry per thread), and computes from
// LDS; the link of a site is loaded once for all its rhs instead of by each
// rhs's thread.  The self stage's "link" is clover slots 0..8 of block 0.
// Record layout, output path and per-thread arithmetic are those of
// k_wilson_lds.
template <typename R, int LPT>
__global__ __launch_bounds__(256) void k_wilson_pf(const cplx<R>* __restrict__ U,
                                                   const cplx<R>* __restrict__ diag,
                                                   const cplx<R>* __restrict__ in,
                                                   cplx<R>* __restrict__ out,
                                                   const int* __restrict__ nbr, int V, int sv,
                                                   int nrhs, unsigned mask, int dag,
                                                   WilsonTiles wt) {
  constexpr int CPS = 12 * (int)sizeof(cplx<R>) / 16;
  constexpr int RS = CPS + 1;
  constexpr int SMAX = LPT == 1 ? 28 : 51;  // sites per tile at nrhs >= 9 or >= 5
  __shared__ uint4 s_rec[256 * RS];
  __shared__ cplx<R> s_lk[9 * SMAX];  // [k][site]
  __shared__ int s_nb[9][SMAX];       // [hop][site], hop 8 = the site itself
  const int T = (int)blockDim.x;
  const int S = T / nrhs;
  const int t = (int)threadIdx.x;
  const int x0 = wilson_tile_site(wt, (int)blockIdx.x);
  const int sl = t / nrhs;
  const int x = x0 + sl;
  const bool live = x < V;
  const GammaTab G = gamma_table();

  for (int j = t; j < 8 * S; j += T) {  // j = s * 8 + h: consecutive lanes, consecutive entries
    const int s = j >> 3, h = j & 7;
    s_nb[h][s] = x0 + s < V ? nbr[(size_t)(x0 + s) * 8 + h] : 0;
  }
  if (t < S) s_nb[8][t] = x0 + t < V ? x0 + t : 0;
  auto next_hop = [&](int i) {
    for (int k = i + 1; k < 9; ++k) {
      const int h = k == 0 ? 8 : k - 1;
      if ((mask >> h) & 1u) return h;
    }
    return -1;
  };
  // native 4-vectors: copies of uint4 (a union type) are memcpys that keep
  // the staging array out of registers
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4* in4 = reinterpret_cast<const u32x4*>(in);
  u32x4* s_rec4 = reinterpret_cast<u32x4*>(s_rec);
  u32x4 nrec[CPS];
  cplx<R> nlk[LPT];
  // loads of hop h into registers: record chunks c = t + k*T of the tile's
  // records, and link entries j = t + m*T (k = j / S, site j % S) for j < 9 S
  auto issue = [&](int h) {
#pragma unroll
    for (int k = 0; k < CPS; ++k) {
      const int c = t + k * T, q = c / CPS, e = c - q * CPS, s = q / nrhs, r = q - s * nrhs;
      nrec[k] = in4[((size_t)s_nb[h][s] * nrhs + r) * CPS + e];
    }
#pragma unroll
    for (int m = 0; m < LPT; ++m) {
      const int j = t + m * T;
      if (j < 9 * S) {
        const int k = j / S, s = j - k * S;
        const cplx<R>* u = h == 8 ? diag + s_nb[8][s]
                                  : U + (size_t)(h & 3) * 9 * sv + (h < 4 ? s_nb[8][s] : s_nb[h][s]);
        nlk[m] = u[(size_t)k * sv];
      }
    }
  };
  auto rec = [&](int e) {
    cplx<R> v = reinterpret_cast<const cplx<R>*>(&s_rec[rec_pos<CPS, RS>(t, e * CPS / 12)])
        [(e * CPS) % 12 / 6];
    if (dag && e >= 6) { v.re = -v.re; v.im = -v.im; }
    return v;
  };

  cplx<R> acc[12];
#pragma unroll
  for (int e = 0; e < 12; ++e) acc[e] = {R(0), R(0)};
  __syncthreads();  // s_nb
  const int h0 = next_hop(-1);
  if (h0 >= 0) issue(h0);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int h = i == 0 ? 8 : i - 1;
    if (!((mask >> h) & 1u)) continue;
    const int hn = next_hop(i);
    __syncthreads();  // every thread is done with the previous hop's LDS data
#pragma unroll
    for (int k = 0; k < CPS; ++k) {
      const int c = t + k * T;
      s_rec4[rec_pos<CPS, RS>(c / CPS, c % CPS)] = nrec[k];
    }
#pragma unroll
    for (int m = 0; m < LPT; ++m)
      if (t + m * T < 9 * S) s_lk[t + m * T] = nlk[m];
    if (hn >= 0) issue(hn);
    __syncthreads();
    if (live) {
      const cplx<R>* lk = s_lk + sl;
      if (h == 8)
        wilson_diag_apply<R>(acc, [&](int l, int s) {
          return l == 0 && s < 9 ? lk[s * S] : diag[((size_t)l * 18 + s) * sv + x];
        }, rec);
      else
        wilson_hop_apply<R>(acc, G, h, [&](int k) { return lk[k * S]; }, rec);
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    cplx<R> v = acc[e];
    if (dag && e >= 6) { v.re = -v.re; v.im = -v.im; }
    reinterpret_cast<cplx<R>*>(&s_rec[rec_pos<CPS, RS>(t, e * CPS / 12)])[(e * CPS) % 12 / 6] = v;
  }
  __syncthreads();
  uint4* out4 = reinterpret_cast<uint4*>(out);
#pragma unroll
  for (int k = 0; k < CPS; ++k) {
    const int c = t + k * T, q = c / CPS, e = c - q * CPS, s = q / nrhs, r = q - s * nrhs;
    if (x0 + s < V) out4[((size_t)(x0 + s) * nrhs + r) * CPS + e] = s_rec[rec_pos<CPS, RS>(q, e)];
  }
}

// Single-rhs form with three threads per site: wavefront c of a 192-thread
// workgroup computes colour row c of 64 sites (tile of 64 sites).  A thread
// keeps 4 accumulators instead of 12 and loads one row (or column) of each
// hop's link instead of all nine entries, so it needs far fewer registers
// than k_wilson_lds's thread per site, and more wavefronts fit on a SIMD.
// The records are staged in LDS as in k_wilson_lds (each read by the three
// threads of the site); hops run self first, then 0..7, with the next hop's
// neighbour sites staged one hop ahead and the link loads issued before the
// record copy.
template <typename R>
__global__ __launch_bounds__(192) void k_wilson_row(const cplx<R>* __restrict__ U,
                                                    const cplx<R>* __restrict__ diag,
                                                    const cplx<R>* __restrict__ in,
                                                    cplx<R>* __restrict__ out,
                                                    const int* __restrict__ nbr, int V, int sv,
                                                    unsigned mask, int dag, WilsonTiles wt) {
  constexpr int CPS = 12 * (int)sizeof(cplx<R>) / 16;
  constexpr int RS = CPS + 1;
  constexpr int S = 64, T = 192, NCH = S * CPS / T;  // record chunks per thread
  static_assert(S * CPS % T == 0, "whole chunks per thread");
  __shared__ uint4 s_rec[S * RS];
  __shared__ int s_y[2][S];
  const int t = (int)threadIdx.x;
  const int c = __builtin_amdgcn_readfirstlane(t >> 6);  // colour row of this wavefront
  const int sl = t & 63;
  const int x0 = wilson_tile_site(wt, (int)blockIdx.x);
  const int x = x0 + sl;
  const bool live = x < V;
  const GammaTab G = gamma_table();

  auto hop_site = [&](int s, int h) {  // h = 8: the site itself
    const int xs = x0 + s;
    if (xs >= V) return 0;
    return h == 8 ? xs : nbr[(size_t)xs * 8 + h];
  };
  auto next_hop = [&](int i) {
    for (int k = i + 1; k < 9; ++k) {
      const int h = k == 0 ? 8 : k - 1;
      if ((mask >> h) & 1u) return h;
    }
    return -1;
  };
  const uint4* in4 = reinterpret_cast<const uint4*>(in);
  auto rec = [&](int e) {  // element e of this site's staged record
    cplx<R> v = reinterpret_cast<const cplx<R>*>(&s_rec[rec_pos<CPS, RS>(sl, e * CPS / 12)])
        [(e * CPS) % 12 / 6];
    if (dag && e >= 6) { v.re = -v.re; v.im = -v.im; }
    return v;
  };

  cplx<R> acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = {R(0), R(0)};
  int cb = 0;
  const int h0 = next_hop(-1);
  if (h0 >= 0) {
    if (t < S) s_y[0][t] = hop_site(t, h0);
    __syncthreads();
  }
  cplx<R> lk[3];
  // one stage's loads: this hop's link row (or column), then the records of
  // its neighbour sites into LDS, the next hop's sites staged one ahead
  auto begin_stage = [&](int h, int hn) __attribute__((always_inline)) {
    __syncthreads();  // every thread is done with the previous hop's records
    const int y = s_y[cb][sl];
    if (live && h != 8) {
      const cplx<R>* u = U + (size_t)(h & 3) * 9 * sv + (h < 4 ? x : y);
#pragma unroll
      for (int j = 0; j < 3; ++j) lk[j] = u[(size_t)(h < 4 ? c * 3 + j : j * 3 + c) * sv];
    }
    const int ynext = hn >= 0 && t < S ? hop_site(t, hn) : 0;
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int cc = t + k * T, q = cc / CPS, e = cc - q * CPS;
      s_rec[rec_pos<CPS, RS>(q, e)] = in4[(size_t)s_y[cb][q] * CPS + e];
    }
    if (hn >= 0 && t < S) s_y[cb ^ 1][t] = ynext;
    __syncthreads();
    cb ^= 1;
  };
  // the self stage apart from the hop loop, so that the loop body stays
  // small enough to be unrolled (every hop index a compile-time constant)
  if ((mask >> 8) & 1u) {
    begin_stage(8, next_hop(0));
    if (live)
      wilson_diag_apply_row<R>(acc, c, [&](int l, int s) {
        return diag[((size_t)l * 18 + s) * sv + x];
      }, rec);
  }
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    if (!((mask >> p) & 1u)) continue;
    begin_stage(p, next_hop(p + 1));
    if (live) wilson_hop_apply_row<R>(acc, G, p, [&](int j) { return lk[j]; }, rec);
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int e = s * 3 + c;
    cplx<R> v = acc[s];
    if (dag && s >= 2) { v.re = -v.re; v.im = -v.im; }
    reinterpret_cast<cplx<R>*>(&s_rec[rec_pos<CPS, RS>(sl, e * CPS / 12)])[(e * CPS) % 12 / 6] = v;
  }
  __syncthreads();
  uint4* out4 = reinterpret_cast<uint4*>(out);
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    const int cc = t + k * T, q = cc / CPS, e = cc - q * CPS;
    if (x0 + q < V) out4[(size_t)(x0 + q) * CPS + e] = s_rec[rec_pos<CPS, RS>(q, e)];
  }
}

// [4][V][9] -> [4][9][sv] and [V][2][36] -> [2][18][sv] (with precision change);
// the site stride sv is V plus a pad
template <typename Rd>
__global__ void k_gauge_soa(const cplx<double>* __restrict__ g, cplx<Rd>* __restrict__ o, int V,
                            int sv) {
  const size_t n = (size_t)36 * V;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t x = t % V, mk = t / V, mu = mk / 9, k = mk % 9;
    const cplx<double> v = g[(mu * V + x) * 9 + k];
    o[mk * sv + x] = {(Rd)v.re, (Rd)v.im};
  }
}
// packed hermitian blocks: [V][2][6][6] -> [2][18][V] (see fine_op.hpp); the
// upper triangle is kept, the lower one is its conjugate by construction
// (F = (Q - Q^dag)/(8i) and sigma_{mu nu} are hermitian)
template <typename Rd>
__global__ void k_diag_soa(const cplx<double>* __restrict__ d, cplx<Rd>* __restrict__ o, int V,
                           int sv) {
  const size_t n = (size_t)36 * V;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t x = t % V, ls = t / V;
    const int l = (int)(ls / 18), slot = (int)(ls % 18);
    const cplx<double>* D = d + (x * 2 + l) * 36;
    cplx<double> v;
    if (slot < 3) {
      v = {D[(2 * slot) * 7].re, D[(2 * slot + 1) * 7].re};
    } else {
      int q = slot - 3, a = 0;
      while (q >= 5 - a) {  // row a holds 5 - a upper entries
        q -= 5 - a;
        ++a;
      }
      v = D[a * 6 + a + 1 + q];
    }
    o[ls * sv + x] = {(Rd)v.re, (Rd)v.im};
  }
}

// (re)build the site-fastest copies the apply reads, in precision Rd
template <typename Rd>
static void build_soa(mgc_fine_op* op, cplx<Rd>* gs, cplx<Rd>* ds, hipStream_t s) {
  k_gauge_soa<Rd><<<grid_for((size_t)36 * op->V, 256, 16384), 256, 0, s>>>(op->gauge64, gs,
                                                                           op->V, op->sv);
  k_diag_soa<Rd><<<grid_for((size_t)36 * op->V, 256, 16384), 256, 0, s>>>(op->diag64, ds, op->V,
                                                                          op->sv);
  MGC_LAUNCH_CHECK();
}

static void free_fine(mgc_fine_op* op) {
  if (!op) return;
  void* bufs[] = {op->gauge64, op->diag64, op->gsoa64, op->dsoa64,
                  op->gsoa32,  op->dsoa32, op->nbr};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete op;
}

// Tile geometry of the apply: S sites of nrhs threads each (at most 256
// threads), XCD-local ordering where the lattice allows it.
// Tiles of exactly S sites (S a power of two), XCD-local where the lattice
// allows it.
static WilsonTiles wilson_tiles_fixed(const Lat4& L, int S) {
  WilsonTiles w{};
  w.S = S;
  w.P = L.L[0] * L.L[1];
  w.L2 = L.L[2];
  w.xcd = (L.L[3] % 8 == 0 && w.P % S == 0) ? 1 : 0;
  if (!w.xcd) return w;
  w.G = w.P / S;
  w.TL = L.L[3] / 8;
  w.ZS = (L.L[2] % 4 == 0) ? 4 : (L.L[2] % 2 == 0 ? 2 : 1);
  return w;
}

static WilsonTiles wilson_tiles(const Lat4& L, int nrhs) {
  WilsonTiles w{};
  const int P = L.L[0] * L.L[1];
  int S = 1;
  while (2 * S * nrhs <= 256 && P % (2 * S) == 0) S *= 2;
  w.S = S;
  w.P = P;
  w.L2 = L.L[2];
  w.xcd = (L.L[3] % 8 == 0 && P % S == 0) ? 1 : 0;
  if (!w.xcd) {
    // consecutive tiles of as many whole sites as fit 256 threads
    w.S = 256 / nrhs;
    return w;
  }
  w.G = P / S;
  w.TL = L.L[3] / 8;
  w.ZS = (L.L[2] % 4 == 0) ? 4 : (L.L[2] % 2 == 0 ? 2 : 1);
  return w;
}

// apply kernel: 0 = k_wilson (each thread reads its own records), 1 =
// k_wilson_lds (records gathered through LDS), 2 = k_wilson_pf for FP32 with
// nrhs >= 5 (k_wilson_lds otherwise), 3 = k_wilson_row for one rhs (as 2
// otherwise); -1 = the default, as 3
static int g_wilson_variant = -1;
static bool wilson_row(int nrhs) {
  return nrhs == 1 && (g_wilson_variant == 3 || g_wilson_variant == -1);
}

static void fine_apply(mgc_fine_op* op, const void* in, void* out, int prec, int nrhs,
                       unsigned mask, int dag, hipStream_t s) {
  MGC_CHECK(op, MGC_ERR_ARG, "null fine operator");
  MGC_CHECK(op->gauge_set, MGC_ERR_STATE, "gauge field not set");
  MGC_CHECK(in && out && in != out, MGC_ERR_ARG, "null or aliased vector");
  MGC_CHECK(nrhs >= 1, MGC_ERR_ARG, "nrhs < 1");
  MGC_CHECK(nrhs <= 256, MGC_ERR_ARG, "fine operator: nrhs > 256");
  const WilsonTiles wt = wilson_tiles(op->lat, nrhs);
  const unsigned g = (unsigned)((op->V + wt.S - 1) / wt.S);
  const unsigned threads = (unsigned)(wt.S * nrhs);
  const bool lds = g_wilson_variant != 0;
  const WilsonTiles wr = wilson_tiles_fixed(op->lat, 64);
  const unsigned gr = (unsigned)((op->V + 63) / 64);
  if (prec == MGC_F64 && wilson_row(nrhs)) {
    k_wilson_row<double><<<gr, 192, 0, s>>>(op->gsoa64, op->dsoa64, (const cplx<double>*)in,
                                            (cplx<double>*)out, op->nbr, op->V, op->sv, mask,
                                            dag, wr);
  } else if (prec == MGC_F64) {
    auto k = lds ? k_wilson_lds<double> : k_wilson<double>;
    k<<<g, threads, 0, s>>>(op->gsoa64, op->dsoa64, (const cplx<double>*)in, (cplx<double>*)out,
                            op->nbr, op->V, op->sv, nrhs, mask, dag, wt);
  } else if (prec == MGC_F32) {
    if (!op->gsoa32 || !op->dsoa32) {
      // both copies or neither: a failed second allocation leaves the handle
      // as it was, so a later FP32 apply retries instead of using a null copy
      const size_t bytes = (size_t)36 * op->sv * sizeof(cplx<float>);
      cplx<float>* g32 = nullptr;
      cplx<float>* d32 = nullptr;
      MGC_HIP(hipMalloc(&g32, bytes));
      const hipError_t e = hipMalloc(&d32, bytes);
      if (e != hipSuccess) {
        (void)hipFree(g32);
        MGC_HIP(e);
      }
      build_soa<float>(op, g32, d32, s);
      op->gsoa32 = g32;
      op->dsoa32 = d32;
    }
    if (wilson_row(nrhs)) {
      k_wilson_row<float><<<gr, 192, 0, s>>>(op->gsoa32, op->dsoa32, (const cplx<float>*)in,
                                             (cplx<float>*)out, op->nbr, op->V, op->sv, mask,
                                             dag, wr);
      MGC_LAUNCH_CHECK();
      return;
    }
    auto k = !lds ? k_wilson<float>
             : g_wilson_variant == 1 || nrhs < 5 ? k_wilson_lds<float>
             : nrhs >= 9                        ? k_wilson_pf<float, 1>
                                                : k_wilson_pf<float, 2>;
    k<<<g, threads, 0, s>>>(op->gsoa32, op->dsoa32, (const cplx<float>*)in, (cplx<float>*)out,
                            op->nbr, op->V, op->sv, nrhs, mask, dag, wt);
  } else {
    throw Error(MGC_ERR_PRECISION, "fine operator: F64 or F32");
  }
  MGC_LAUNCH_CHECK();
}

}  // namespace mgc

using namespace mgc;

extern "C" {

// the built-in operator behind the callback interface: points 0..7 = Mdir,
// 8 = Mdiag, 9 = M
static int builtin_linop_apply(void* ctx, int point, int prec, const void* in, void* out,
                               int nrhs, void* stream) {
  mgc_fine_op* op = static_cast<mgc_fine_op*>(ctx);
  if (point >= 0 && point < 8)
    return mgc_fine_op_Mdir(op, in, out, point_dir(point), point_disp(point), prec, nrhs, stream);
  if (point == 8) return mgc_fine_op_Mdiag(op, in, out, prec, nrhs, stream);
  if (point == 9) return mgc_fine_op_M(op, in, out, prec, nrhs, stream);
  set_last_error("fine operator callback: point must be in 0..9");
  return MGC_ERR_ARG;
}

int mgc_fine_op_linop(mgc_fine_op* op, mgc_fine_linop* out) {
  return guarded([&] {
    MGC_CHECK(op && out, MGC_ERR_ARG, "null argument");
    out->apply = builtin_linop_apply;
    out->ctx = op;
    for (int d = 0; d < 4; ++d) out->dims[d] = op->lat.L[d];
    out->precisions = (1 << MGC_F64) | (1 << MGC_F32);
    out->max_nrhs = 256;
    out->gamma5_hermitian = 1;
  });
}

int mgc_fine_op_set_variant(int variant) {
  return guarded([&] {
    MGC_CHECK(variant >= -1 && variant <= 3, MGC_ERR_ARG, "unknown fine-operator kernel variant");
    g_wilson_variant = variant;
  });
}

int mgc_random_gauge(void* gauge_f64, const int dims[4], double eps, uint64_t seed,
                     uint64_t stream_id, void* stream) {
  return guarded([&] {
    MGC_CHECK(gauge_f64 && dims, MGC_ERR_ARG, "null argument");
    size_t V = 1;
    for (int d = 0; d < 4; ++d) V *= (size_t)dims[d];
    const size_t nmat = 4 * V;
    const uint64_t key = splitmix64_g(splitmix64_g(seed) ^ (stream_id * 0xD1B54A32D192ED03ull));
    k_random_gauge<<<(unsigned)((nmat + 127) / 128), 128, 0, as_stream(stream)>>>(
        (cplx<double>*)gauge_f64, nmat, eps, key);
    MGC_LAUNCH_CHECK();
  });
}

int mgc_fine_op_create(mgc_fine_op** out, const int dims[4], double mass, double csw,
                       int antiperiodic_t) {
  mgc_fine_op* op = nullptr;
  int st = guarded([&] {
    MGC_CHECK(out && dims, MGC_ERR_ARG, "null argument");
    op = new mgc_fine_op();
    for (int d = 0; d < 4; ++d) {
      MGC_CHECK(dims[d] >= 1, MGC_ERR_SHAPE, "extent < 1");
      op->lat.L[d] = dims[d];
    }
    op->V = op->lat.volume();
    op->sv = op->V + 32;
    op->mass = mass;
    op->csw = csw;
    op->antiperiodic_t = antiperiodic_t ? 1 : 0;
    std::vector<int> nb((size_t)op->V * 8);
    for (int x = 0; x < op->V; ++x)
      for (int p = 0; p < 8; ++p) nb[(size_t)x * 8 + p] = op->lat.shift(x, p & 3, p < 4 ? 1 : -1);
    MGC_HIP(hipMalloc(&op->nbr, nb.size() * sizeof(int)));
    MGC_HIP(hipMemcpy(op->nbr, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice));
    MGC_HIP(hipMalloc(&op->gauge64, (size_t)4 * op->V * 9 * sizeof(cplx<double>)));
    MGC_HIP(hipMalloc(&op->diag64, (size_t)op->V * 72 * sizeof(cplx<double>)));
    MGC_HIP(hipMalloc(&op->gsoa64, (size_t)36 * op->sv * sizeof(cplx<double>)));
    MGC_HIP(hipMalloc(&op->dsoa64, (size_t)36 * op->sv * sizeof(cplx<double>)));
    *out = op;
  });
  if (st != MGC_OK) free_fine(op);
  return st;
}

int mgc_fine_op_destroy(mgc_fine_op* op) {
  return guarded([&] { free_fine(op); });
}

int mgc_fine_op_set_gauge(mgc_fine_op* op, const void* gauge_f64, void* stream) {
  return guarded([&] {
    MGC_CHECK(op && gauge_f64, MGC_ERR_ARG, "null argument");
    hipStream_t s = as_stream(stream);
    MGC_HIP(hipMemcpyAsync(op->gauge64, gauge_f64, (size_t)4 * op->V * 9 * sizeof(cplx<double>),
                           hipMemcpyDeviceToDevice, s));
    if (op->antiperiodic_t) {
      k_antiperiodic<<<(op->V + 255) / 256, 256, 0, s>>>(op->gauge64, op->lat);
      MGC_LAUNCH_CHECK();
    }
    k_clover<<<(op->V + 63) / 64, 64, 0, s>>>(op->gauge64, op->lat, op->mass, op->csw,
                                              op->diag64);
    MGC_LAUNCH_CHECK();
    build_soa<double>(op, op->gsoa64, op->dsoa64, s);
    if (op->gsoa32) build_soa<float>(op, op->gsoa32, op->dsoa32, s);  // refresh the FP32 copies
    op->gauge_set = true;
  });
}

int mgc_fine_op_get_diag(const mgc_fine_op* op, void* out_f64, void* stream) {
  return guarded([&] {
    MGC_CHECK(op && out_f64, MGC_ERR_ARG, "null argument");
    MGC_CHECK(op->gauge_set, MGC_ERR_STATE, "gauge field not set");
    MGC_HIP(hipMemcpyAsync(out_f64, op->diag64, (size_t)op->V * 72 * sizeof(cplx<double>),
                           hipMemcpyDeviceToDevice, as_stream(stream)));
  });
}

int mgc_fine_op_M(mgc_fine_op* op, const void* in, void* out, int prec, int nrhs, void* stream) {
  return guarded([&] { fine_apply(op, in, out, prec, nrhs, 0x1FFu, 0, as_stream(stream)); });
}
int mgc_fine_op_Mdag(mgc_fine_op* op, const void* in, void* out, int prec, int nrhs,
                     void* stream) {
  return guarded([&] { fine_apply(op, in, out, prec, nrhs, 0x1FFu, 1, as_stream(stream)); });
}
int mgc_fine_op_Mdiag(mgc_fine_op* op, const void* in, void* out, int prec, int nrhs,
                      void* stream) {
  return guarded([&] { fine_apply(op, in, out, prec, nrhs, 0x100u, 0, as_stream(stream)); });
}
int mgc_fine_op_Mdir(mgc_fine_op* op, const void* in, void* out, int dir, int disp, int prec,
                     int nrhs, void* stream) {
  return guarded([&] {
    MGC_CHECK(disp == 1 || disp == -1, MGC_ERR_ARG, "Mdir needs disp = +1 or -1");
    MGC_CHECK(dir >= 0 && dir < 4, MGC_ERR_ARG, "dir in 0..3");
    const int p = disp == 1 ? dir : dir + 4;
    fine_apply(op, in, out, prec, nrhs, 1u << p, 0, as_stream(stream));
  });
}

}  // extern "C"
