/*
 * coarse.c -- FP64 CPU restatement of the coarse operator (CoarsenedMatrix).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   include/CoarsenedMatrixBaseline.h:314-344   M (site loop, dense N x N matvec per point)
 *   include/CoarsenedMatrixUpstream.h:609-621    Mdag = G5C M G5C (Galerkin)
 *   include/CoarsenedMatrixUpstream.h:622-742    MdirCalc / MdirAll / Mdir / Mdiag
 *   include/CoarsenedMatrixBaseline.h:439-554    Dhop, DhopOE, DhopEO (no self point)
 *   include/CoarsenedMatrixBaseline.h:468-489    Mooee family on the self link
 *   include/CoarsenedMatrixBaseline.h:735-768    InvertSelfStencilPoint
 *   include/Layout_converters.h:33-89            flat <-> split link layout
 *   include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h:767-794  multi-rhs M
 *   include/CoarsenedMatrixUpstreamImprovedDirsave.h:999-1032        ConstructRemainingLinks
 *   include/CoarsenedMatrixKernelsRework.h:141-167                   shiftLinks
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cd;

static inline cd cd_mul(cd a, cd b) { cd r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cd cd_add(cd a, cd b) { cd r = {a.re + b.re, a.im + b.im}; return r; }
static inline cd cd_conj(cd a) { cd r = {a.re, -a.im}; return r; }

static int volume(const int d[4]) { return d[0] * d[1] * d[2] * d[3]; }

int orc_num_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

/* threads of the following parallel regions (the CPU-baseline timing runs
   the same work on all host threads and on one) */
void orc_set_num_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

/* ---------------- counter-based Gaussian generator ---------------- */
static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static double rng_uniform(uint64_t key, uint64_t ctr) {
  uint64_t b = splitmix64(key ^ splitmix64(ctr));
  return ((double)(b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
void orc_gaussian(double* out, size_t n, uint64_t seed, uint64_t stream) {
  uint64_t key = splitmix64(splitmix64(seed) ^ (stream * 0xD1B54A32D192ED03ull));
  const double two_pi = 6.283185307179586476925286766559;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    double u1 = rng_uniform(key, 2 * (uint64_t)i);
    double u2 = rng_uniform(key, 2 * (uint64_t)i + 1);
    double r = sqrt(-log(u1));
    out[2 * i] = r * cos(two_pi * u2);
    out[2 * i + 1] = r * sin(two_pi * u2);
  }
}

/* ---------------- geometry ---------------- */
/* directions-faster point order, include/GeometryRework.h:90-97 */
int orc_point_from_dir_disp(int dir, int disp) {
  if (dir == 0 && disp == 0) return 8;
  return (1 - disp) / 2 * 4 + dir;
}
void orc_point_dir_disp(int p, int* dir, int* disp) {
  /* include/GeometryRework.h:109-125 */
  if (p == 8) { *dir = 0; *disp = 0; return; }
  *dir = p % 4;
  *disp = p < 4 ? +1 : -1;
}
static void coords(const int d[4], int x, int c[4]) {
  for (int k = 0; k < 4; ++k) { c[k] = x % d[k]; x /= d[k]; }
}
static int lex(const int d[4], const int c[4]) {
  return c[0] + d[0] * (c[1] + d[1] * (c[2] + d[2] * c[3]));
}
int orc_site_shift(const int d[4], int x, int dir, int disp) {
  int c[4];
  coords(d, x, c);
  c[dir] = ((c[dir] + disp) % d[dir] + d[dir]) % d[dir];
  return lex(d, c);
}
int orc_parity(const int d[4], int x) {
  int c[4];
  coords(d, x, c);
  return (c[0] + c[1] + c[2] + c[3]) & 1;
}
int orc_half_index(const int d[4], int x) {
  /* count the sites of the same parity that precede x lexicographically */
  int par = orc_parity(d, x), h = 0;
  for (int y = 0; y < x; ++y) h += (orc_parity(d, y) == par);
  return h;
}

/* ---------------- coarse stencil ---------------- */
static inline const cd* link_at(const double* links, int V, int N, int p, int x) {
  return (const cd*)links + ((size_t)p * V + x) * N * N;
}

static void site_apply(int V, int N, const double* links, const cd* in_full_lookup_base,
                       int x, const int* nbr_idx, unsigned mask, int dag, cd* out_row,
                       cd* tmp) {
  /* out(x) = sum_p A_p(x) g(in)(nbr_p), g = G5C when dag */
  for (int r = 0; r < N; ++r) out_row[r].re = out_row[r].im = 0.0;
  for (int p = 0; p < 9; ++p) {
    if (!((mask >> p) & 1u)) continue;
    const cd* v = in_full_lookup_base + (size_t)nbr_idx[p] * N;
    for (int c = 0; c < N; ++c) {
      tmp[c] = v[c];
      if (dag && c >= N / 2) { tmp[c].re = -tmp[c].re; tmp[c].im = -tmp[c].im; }
    }
    const cd* A = link_at(links, V, N, p, x);
    for (int r = 0; r < N; ++r) {
      cd acc = {0.0, 0.0};
      for (int c = 0; c < N; ++c) acc = cd_add(acc, cd_mul(A[(size_t)r * N + c], tmp[c]));
      out_row[r] = cd_add(out_row[r], acc);
    }
  }
  if (dag)
    for (int r = N / 2; r < N; ++r) { out_row[r].re = -out_row[r].re; out_row[r].im = -out_row[r].im; }
}

void orc_coarse_apply(const int d[4], int N, const double* links, const double* in, double* out,
                      unsigned mask, int dag) {
  const int V = volume(d);
#pragma omp parallel
  {
    cd* tmp = (cd*)malloc(sizeof(cd) * N);
#pragma omp for schedule(static)
    for (int x = 0; x < V; ++x) {
      int nb[9];
      for (int p = 0; p < 8; ++p) {
        int dir, disp;
        orc_point_dir_disp(p, &dir, &disp);
        nb[p] = orc_site_shift(d, x, dir, disp);
      }
      nb[8] = x;
      site_apply(V, N, links, (const cd*)in, x, nb, mask, dag, (cd*)out + (size_t)x * N, tmp);
    }
    free(tmp);
  }
}

static int* half_table(const int d[4]) {
  /* full site -> half index, by enumeration */
  const int V = volume(d);
  int* h = (int*)malloc(sizeof(int) * V);
  int cnt[2] = {0, 0};
  for (int x = 0; x < V; ++x) {
    int par = orc_parity(d, x);
    h[x] = cnt[par]++;
  }
  return h;
}

void orc_coarse_apply_cb(const int d[4], int N, const double* links, const double* in_half,
                         double* out_half, int out_cb, unsigned mask, int dag) {
  const int V = volume(d);
  int* h = half_table(d);
  /* list of output sites */
  int* sites = (int*)malloc(sizeof(int) * V);
  int n_out = 0;
  for (int x = 0; x < V; ++x)
    if (orc_parity(d, x) == out_cb) sites[n_out++] = x;
#pragma omp parallel
  {
    cd* tmp = (cd*)malloc(sizeof(cd) * N);
#pragma omp for schedule(static)
    for (int o = 0; o < n_out; ++o) {
      const int x = sites[o];
      int nb[9];
      for (int p = 0; p < 8; ++p) {
        int dir, disp;
        orc_point_dir_disp(p, &dir, &disp);
        nb[p] = h[orc_site_shift(d, x, dir, disp)];
      }
      nb[8] = h[x];
      site_apply(V, N, links, (const cd*)in_half, x, nb, mask, dag,
                 (cd*)out_half + (size_t)h[x] * N, tmp);
    }
    free(tmp);
  }
  free(sites);
  free(h);
}

void orc_site_matrix_apply(const int d[4], int N, const double* mats, const double* in,
                           double* out, int cb, int adjoint) {
  const int V = volume(d);
  int* h = half_table(d);
#pragma omp parallel for schedule(static)
  for (int x = 0; x < V; ++x) {
    if (cb >= 0 && orc_parity(d, x) != cb) continue;
    const int io = cb >= 0 ? h[x] : x;
    const cd* A = (const cd*)mats + (size_t)x * N * N;
    const cd* v = (const cd*)in + (size_t)io * N;
    cd* w = (cd*)out + (size_t)io * N;
    for (int r = 0; r < N; ++r) {
      cd acc = {0.0, 0.0};
      for (int c = 0; c < N; ++c) {
        cd a = adjoint ? cd_conj(A[(size_t)c * N + r]) : A[(size_t)r * N + c];
        acc = cd_add(acc, cd_mul(a, v[c]));
      }
      w[r] = acc;
    }
  }
  free(h);
}

/* Gauss-Jordan elimination on the augmented matrix [A | 1] with partial
 * pivoting (the textbook algorithm; the reference calls Eigen's .inverse(),
 * include/CoarsenedMatrixBaseline.h:759) */
int orc_invert(int N, const double* Ain, double* Aout) {
  const int W = 2 * N;
  cd* T = (cd*)malloc(sizeof(cd) * (size_t)N * W);
  const cd* A = (const cd*)Ain;
  for (int r = 0; r < N; ++r)
    for (int c = 0; c < W; ++c) {
      cd z = {0.0, 0.0};
      if (c < N) z = A[(size_t)r * N + c];
      else if (c - N == r) z.re = 1.0;
      T[(size_t)r * W + c] = z;
    }
  for (int k = 0; k < N; ++k) {
    int piv = k;
    double best = -1.0;
    for (int r = k; r < N; ++r) {
      cd v = T[(size_t)r * W + k];
      double m = v.re * v.re + v.im * v.im;
      if (m > best) { best = m; piv = r; }
    }
    if (!(best > 0.0)) { free(T); return -1; }
    if (piv != k)
      for (int c = 0; c < W; ++c) {
        cd t = T[(size_t)k * W + c];
        T[(size_t)k * W + c] = T[(size_t)piv * W + c];
        T[(size_t)piv * W + c] = t;
      }
    cd p = T[(size_t)k * W + k];
    double den = p.re * p.re + p.im * p.im;
    cd pinv = {p.re / den, -p.im / den};
    for (int c = 0; c < W; ++c) T[(size_t)k * W + c] = cd_mul(T[(size_t)k * W + c], pinv);
    for (int r = 0; r < N; ++r) {
      if (r == k) continue;
      cd f = T[(size_t)r * W + k];
      if (f.re == 0.0 && f.im == 0.0) continue;
      for (int c = 0; c < W; ++c) {
        cd t = cd_mul(f, T[(size_t)k * W + c]);
        T[(size_t)r * W + c].re -= t.re;
        T[(size_t)r * W + c].im -= t.im;
      }
    }
  }
  cd* out = (cd*)Aout;
  for (int r = 0; r < N; ++r)
    for (int c = 0; c < N; ++c) out[(size_t)r * N + c] = T[(size_t)r * W + N + c];
  free(T);
  return 0;
}

int orc_invert_self(const int d[4], int N, const double* links, double* inv) {
  const int V = volume(d);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int x = 0; x < V; ++x) {
    const double* A = (const double*)link_at(links, V, N, 8, x);
    if (orc_invert(N, A, inv + 2 * (size_t)x * N * N) != 0) bad |= 1;
  }
  return bad ? -1 : 0;
}

/* include/Layout_converters.h:33-89: n = s * Nc_c + c */
void orc_links_flat_to_split(int V, int N, const double* flat, double* split) {
  const int nb = N / 2;
  const cd* f = (const cd*)flat;
  cd* s = (cd*)split;
  for (int p = 0; p < 9; ++p)
    for (int x = 0; x < V; ++x)
      for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
          int s1 = r / nb, c1 = r % nb, s2 = c / nb, c2 = c % nb;
          s[(((((size_t)p * V + x) * 2 + s1) * 2 + s2) * nb + c1) * nb + c2] =
              f[(((size_t)p * V + x) * N + r) * N + c];
        }
}
void orc_links_split_to_flat(int V, int N, const double* split, double* flat) {
  const int nb = N / 2;
  const cd* s = (const cd*)split;
  cd* f = (cd*)flat;
  for (int p = 0; p < 9; ++p)
    for (int x = 0; x < V; ++x)
      for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
          int s1 = r / nb, c1 = r % nb, s2 = c / nb, c2 = c % nb;
          f[(((size_t)p * V + x) * N + r) * N + c] =
              s[(((((size_t)p * V + x) * 2 + s1) * 2 + s2) * nb + c1) * nb + c2];
        }
}

/* 5d layout sF = sU * Ls + s5: the links of coarse site sU serve every rhs */
void orc_coarse_mrhs_M(const int d[4], int N, int nrhs, const double* links, const double* in,
                       double* out) {
  const int V = volume(d);
#pragma omp parallel for schedule(static)
  for (int x = 0; x < V; ++x) {
    for (int k = 0; k < nrhs; ++k) {
      cd* w = (cd*)out + ((size_t)x * nrhs + k) * N;
      for (int r = 0; r < N; ++r) w[r].re = w[r].im = 0.0;
      for (int p = 0; p < 9; ++p) {
        int dir, disp, y;
        orc_point_dir_disp(p, &dir, &disp);
        y = (p == 8) ? x : orc_site_shift(d, x, dir, disp);
        const cd* v = (const cd*)in + ((size_t)y * nrhs + k) * N;
        const cd* A = link_at(links, V, N, p, x);
        for (int r = 0; r < N; ++r) {
          cd acc = {0.0, 0.0};
          for (int c = 0; c < N; ++c) acc = cd_add(acc, cd_mul(A[(size_t)r * N + c], v[c]));
          w[r] = cd_add(w[r], acc);
        }
      }
    }
  }
}

/* A_{want}(x) = sigma o adj(A_{have}(x + disp_want * mu)), sigma = -1 on the
 * off-diagonal chirality blocks */
void orc_construct_remaining_links(const int d[4], int N, double* links, int disp_have) {
  const int V = volume(d);
  const int nb = N / 2;
  const int disp_want = -disp_have;
  for (int mu = 0; mu < 4; ++mu) {
    const int ph = orc_point_from_dir_disp(mu, disp_have);
    const int pw = orc_point_from_dir_disp(mu, disp_want);
#pragma omp parallel for schedule(static)
    for (int x = 0; x < V; ++x) {
      /* Cshift(tmp, mu, disp_want)(x) = tmp(x + disp_want * mu) */
      const int y = orc_site_shift(d, x, mu, disp_want);
      const cd* Ah = link_at(links, V, N, ph, y);
      cd* Aw = (cd*)links + ((size_t)pw * V + x) * N * N;
      for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
          cd a = cd_conj(Ah[(size_t)c * N + r]);
          if ((r / nb + c / nb) % 2 == 1) { a.re = -a.re; a.im = -a.im; }
          Aw[(size_t)r * N + c] = a;
        }
    }
  }
}
