/*
 * wilson.c -- FP64 CPU Wilson-clover fine operator and the Galerkin coarse
 * operator construction.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Fine operator (this build's documented convention; Grid's own is not
 * available here, so the fine operator itself is pinned by the identities
 * M = Mdiag + sum_p Mdir_p (MultigridPreconditioner.h:317-350), gamma5-
 * hermiticity and a numpy restatement in the tests):
 *   (D psi)(x) = [(4+m) + C(x)] psi(x)
 *                - 1/2 sum_mu [ (1 - g_mu) U_mu(x) psi(x+mu)
 *                             + (1 + g_mu) U_mu(x-mu)^dag psi(x-mu) ]
 *   C(x) = -(csw/2) sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
 *   sigma_{mu nu} = (i/2)[g_mu, g_nu],  F = (Q - Q^dag)/(8i),
 *   Q = sum of the four plaquette leaves at x in the (mu, nu) plane.
 *   Chiral gamma basis, g5 = g_x g_y g_z g_t = diag(1,1,-1,-1).
 *
 * Coarse operator: CoarsenOperator in the inner/outer-mask form of
 * include/CoarsenedMatrixBaseline.h:575-705 (valid for odd coarse extents):
 *   for each doubled basis vector i:  for each point p:
 *     Mphi = Mdir_p v_i (p < 8) or Mdiag v_i (p = 8)
 *     A_p(:, i)  = R_outer_p Mphi        (p < 8)
 *     A_8(:, i) += R_inner_p Mphi
 * with the chirally doubled basis v_{l*nb + c} = P_l u_c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct { double re, im; } cd;
static inline cd mk(double r, double i) { cd z = {r, i}; return z; }
static inline cd mul(cd a, cd b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cd add(cd a, cd b) { return mk(a.re + b.re, a.im + b.im); }
static inline cd sub(cd a, cd b) { return mk(a.re - b.re, a.im - b.im); }
static inline cd cj(cd a) { return mk(a.re, -a.im); }
static inline cd scl(double s, cd a) { return mk(s * a.re, s * a.im); }

static int vol(const int d[4]) { return d[0] * d[1] * d[2] * d[3]; }

/* gamma matrices, chiral basis: gamma[mu][row][col] */
static cd G[4][4][4];
static int G_init = 0;
static void init_gamma(void) {
  if (G_init) return;
  memset(G, 0, sizeof(G));
  const cd I = mk(0, 1), mI = mk(0, -1), one = mk(1, 0), mone = mk(-1, 0);
  /* x */
  G[0][0][3] = I; G[0][1][2] = I; G[0][2][1] = mI; G[0][3][0] = mI;
  /* y */
  G[1][0][3] = mone; G[1][1][2] = one; G[1][2][1] = one; G[1][3][0] = mone;
  /* z */
  G[2][0][2] = I; G[2][1][3] = mI; G[2][2][0] = mI; G[2][3][1] = I;
  /* t */
  G[3][0][2] = one; G[3][1][3] = one; G[3][2][0] = one; G[3][3][1] = one;
  G_init = 1;
}

void orc_gamma(int mu, double* out) {
  init_gamma();
  memcpy(out, G[mu], sizeof(G[mu]));
}

/* 3x3 complex helpers, row-major */
static void m3mul(const cd* a, const cd* b, cd* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cd s = mk(0, 0);
      for (int k = 0; k < 3; ++k) s = add(s, mul(a[i * 3 + k], b[k * 3 + j]));
      c[i * 3 + j] = s;
    }
}
static void m3dag(const cd* a, cd* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[i * 3 + j] = cj(a[j * 3 + i]);
}
static const cd* link(const double* gauge, int V, int mu, int x) {
  return (const cd*)gauge + ((size_t)mu * V + x) * 9;
}

/* clover field + mass term, stored as [V][12][12] (spin-major s*3+c) */
void orc_wilson_diag(const int d[4], const double* gauge, double mass, double csw, double* diag) {
  init_gamma();
  const int V = vol(d);
  cd* D = (cd*)diag;
#pragma omp parallel for schedule(static)
  for (int x = 0; x < V; ++x) {
    cd* Dx = D + (size_t)x * 144;
    for (int i = 0; i < 144; ++i) Dx[i] = mk(0, 0);
    for (int i = 0; i < 12; ++i) Dx[i * 12 + i] = mk(4.0 + mass, 0);
    if (csw == 0.0) continue;
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        const int xpm = orc_site_shift(d, x, mu, +1), xpn = orc_site_shift(d, x, nu, +1);
        const int xmm = orc_site_shift(d, x, mu, -1), xmn = orc_site_shift(d, x, nu, -1);
        const int xmmpn = orc_site_shift(d, xmm, nu, +1);
        const int xmmmn = orc_site_shift(d, xmm, nu, -1);
        const int xmnpm = orc_site_shift(d, xmn, mu, +1);
        cd Q[9], t1[9], t2[9], a[9], b[9];
        for (int i = 0; i < 9; ++i) Q[i] = mk(0, 0);
        /* leaf 1: U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
        m3mul(link(gauge, V, mu, x), link(gauge, V, nu, xpm), t1);
        m3dag(link(gauge, V, mu, xpn), a);
        m3mul(t1, a, t2);
        m3dag(link(gauge, V, nu, x), a);
        m3mul(t2, a, t1);
        for (int i = 0; i < 9; ++i) Q[i] = add(Q[i], t1[i]);
        /* leaf 2: U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu) */
        m3dag(link(gauge, V, mu, xmmpn), a);
        m3mul(link(gauge, V, nu, x), a, t1);
        m3dag(link(gauge, V, nu, xmm), a);
        m3mul(t1, a, t2);
        m3mul(t2, link(gauge, V, mu, xmm), t1);
        for (int i = 0; i < 9; ++i) Q[i] = add(Q[i], t1[i]);
        /* leaf 3: U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu) */
        m3dag(link(gauge, V, mu, xmm), a);
        m3dag(link(gauge, V, nu, xmmmn), b);
        m3mul(a, b, t1);
        m3mul(t1, link(gauge, V, mu, xmmmn), t2);
        m3mul(t2, link(gauge, V, nu, xmn), t1);
        for (int i = 0; i < 9; ++i) Q[i] = add(Q[i], t1[i]);
        /* leaf 4: U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu) U_mu(x)^dag */
        m3dag(link(gauge, V, nu, xmn), a);
        m3mul(a, link(gauge, V, mu, xmn), t1);
        m3mul(t1, link(gauge, V, nu, xmnpm), t2);
        m3dag(link(gauge, V, mu, x), a);
        m3mul(t2, a, t1);
        for (int i = 0; i < 9; ++i) Q[i] = add(Q[i], t1[i]);
        /* F = (Q - Q^dag)/(8i) */
        cd F[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            cd z = sub(Q[i * 3 + j], cj(Q[j * 3 + i]));
            F[i * 3 + j] = mk(z.im / 8.0, -z.re / 8.0); /* z / (8i) */
          }
        /* sigma = (i/2)(g_mu g_nu - g_nu g_mu) */
        cd S[4][4];
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) {
            cd s = mk(0, 0);
            for (int k = 0; k < 4; ++k)
              s = add(s, sub(mul(G[mu][r][k], G[nu][k][c]), mul(G[nu][r][k], G[mu][k][c])));
            S[r][c] = mul(mk(0, 0.5), s);
          }
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c)
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) {
                cd v = scl(-0.5 * csw, mul(S[r][c], F[i * 3 + j]));
                Dx[(r * 3 + i) * 12 + c * 3 + j] = add(Dx[(r * 3 + i) * 12 + c * 3 + j], v);
              }
      }
  }
}

/* out(x) = -1/2 (1 - disp*g_mu) L psi(x + disp*mu), L = U_mu(x) or U_mu(x-mu)^dag */
void orc_wilson_dir(const int d[4], const double* gauge, const double* in, double* out, int mu,
                    int disp) {
  init_gamma();
  const int V = vol(d);
  const cd* psi = (const cd*)in;
  cd* o = (cd*)out;
#pragma omp parallel for schedule(static)
  for (int x = 0; x < V; ++x) {
    const int y = orc_site_shift(d, x, mu, disp);
    cd L[9];
    if (disp == +1) memcpy(L, link(gauge, V, mu, x), sizeof(L));
    else m3dag(link(gauge, V, mu, y), L);
    cd chi[12];
    for (int s = 0; s < 4; ++s)
      for (int i = 0; i < 3; ++i) {
        cd acc = mk(0, 0);
        for (int j = 0; j < 3; ++j) acc = add(acc, mul(L[i * 3 + j], psi[(size_t)y * 12 + s * 3 + j]));
        chi[s * 3 + i] = acc;
      }
    for (int s = 0; s < 4; ++s)
      for (int i = 0; i < 3; ++i) {
        cd g = mk(0, 0);
        for (int t = 0; t < 4; ++t) g = add(g, mul(G[mu][s][t], chi[t * 3 + i]));
        cd v = sub(chi[s * 3 + i], scl((double)disp, g));
        o[(size_t)x * 12 + s * 3 + i] = scl(-0.5, v);
      }
  }
}

void orc_wilson_diag_apply(int V, const double* diag, const double* in, double* out) {
  const cd* D = (const cd*)diag;
  const cd* psi = (const cd*)in;
  cd* o = (cd*)out;
#pragma omp parallel for schedule(static)
  for (int x = 0; x < V; ++x)
    for (int r = 0; r < 12; ++r) {
      cd acc = mk(0, 0);
      for (int c = 0; c < 12; ++c) acc = add(acc, mul(D[(size_t)x * 144 + r * 12 + c], psi[(size_t)x * 12 + c]));
      o[(size_t)x * 12 + r] = acc;
    }
}

void orc_wilson_M(const int d[4], const double* gauge, const double* diag, const double* in,
                  double* out) {
  const int V = vol(d);
  double* tmp = (double*)malloc(sizeof(double) * 24 * (size_t)V);
  orc_wilson_diag_apply(V, diag, in, out);
  for (int p = 0; p < 8; ++p) {
    int mu, disp;
    orc_point_dir_disp(p, &mu, &disp);
    orc_wilson_dir(d, gauge, in, tmp, mu, disp);
    for (size_t i = 0; i < 24 * (size_t)V; ++i) out[i] += tmp[i];
  }
  free(tmp);
}

/* CoarsenOperator, Baseline mask form.  basis: nb split vectors [nb][Vf][12];
 * links out: FLAT [9][Vc][N][N] with N = 2*nb */
void orc_coarsen_operator(const int fd[4], const int blk[4], int nb, const double* gauge,
                          const double* diag, const double* basis, double* links) {
  int cdims[4];
  for (int k = 0; k < 4; ++k) cdims[k] = fd[k] / blk[k];
  const int Vf = vol(fd), Vc = vol(cdims), N = 2 * nb;
  cd* A = (cd*)links;
  memset(A, 0, sizeof(cd) * 9 * (size_t)Vc * N * N);
  double* phi = (double*)malloc(sizeof(double) * 24 * (size_t)Vf);
  double* Mphi = (double*)malloc(sizeof(double) * 24 * (size_t)Vf);
  double* proj = (double*)malloc(sizeof(double) * 2 * (size_t)Vc * N);
  for (int i = 0; i < N; ++i) {
    const int l = i / nb, c = i % nb;
    /* v_i = P_l u_c */
    const double* u = basis + 2 * (size_t)c * Vf * 12;
    for (size_t x = 0; x < (size_t)Vf; ++x)
      for (int e = 0; e < 12; ++e) {
        const int keep = (e / 6) == l;
        phi[2 * (x * 12 + e)] = keep ? u[2 * (x * 12 + e)] : 0.0;
        phi[2 * (x * 12 + e) + 1] = keep ? u[2 * (x * 12 + e) + 1] : 0.0;
      }
    for (int p = 0; p < 9; ++p) {
      int mu, disp;
      orc_point_dir_disp(p, &mu, &disp);
      if (p == 8) orc_wilson_diag_apply(Vf, diag, phi, Mphi);
      else orc_wilson_dir(fd, gauge, phi, Mphi, mu, disp);
      if (p < 8) {
        orc_project(fd, blk, nb, basis, Mphi, proj, 2, p); /* outer */
        const cd* pr = (const cd*)proj;
        for (int xc = 0; xc < Vc; ++xc)
          for (int j = 0; j < N; ++j) A[(((size_t)p * Vc + xc) * N + j) * N + i] = pr[(size_t)xc * N + j];
      }
      orc_project(fd, blk, nb, basis, Mphi, proj, 1, p); /* inner */
      const cd* pr = (const cd*)proj;
      for (int xc = 0; xc < Vc; ++xc)
        for (int j = 0; j < N; ++j) {
          cd* a = &A[(((size_t)8 * Vc + xc) * N + j) * N + i];
          *a = add(*a, pr[(size_t)xc * N + j]);
        }
    }
  }
  free(phi);
  free(Mphi);
  free(proj);
}
