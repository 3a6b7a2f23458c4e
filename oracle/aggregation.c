/*
 * aggregation.c -- FP64 CPU restatement of the block transfer operators.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   include/CoarseningLookupTableRework.h:121-179   LUT populate (mask, reverse table)
 *   include/AggregationKernelsRework.h:49-94         aggregateProject (chirality split)
 *   include/AggregationKernelsRework.h:159-205       aggregatePromote (chirality split)
 *   include/AggregationKernelsRework.h:266-341       aggregateOrthogonalise (fused MGS)
 *   include/CoarsenedMatrixUpstreamImprovedDirsaveLut.h:70-239  LUT project/axpy/MGS
 *                                                    on the chirally doubled basis
 *   include/Benchmark_helpers.h:156-181, 203-213     chiral doubling and its undo
 *   include/AggregationRework.h:91-108               CheckOrthogonal
 *   include/CoarsenedMatrixBaseline.h:609-675        CoarsenOperator (mask form), here on a
 *                                                    coarse operator (coarse-of-coarse level)
 *
 * Fine fields are [site][Ef]: Ef = 12 for a Dirac field ([spin 4][colour 3]),
 * Ef = N for a coarse field.  The chirality of element e is e / (Ef/2)
 * (spin / 2 for gamma5 = diag(1,1,-1,-1); the first / second half of a
 * coarse vector).  Split basis: nb fine fields; coarse vectors
 * [coarse site][2*nb] with n = k*nb + i.  Doubled basis: 2*nb fine fields of
 * definite chirality; coarse vector index n.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct { double re, im; } cd;

static int vol(const int d[4]) { return d[0] * d[1] * d[2] * d[3]; }
static void coords(const int d[4], int x, int c[4]) {
  for (int k = 0; k < 4; ++k) { c[k] = x % d[k]; x /= d[k]; }
}
static int lex(const int d[4], const int c[4]) {
  return c[0] + d[0] * (c[1] + d[1] * (c[2] + d[2] * c[3]));
}

/* mask kinds: 0 = every site; 1 = inner sites of point p; 2 = outer sites of
 * point p (include/CoarsenedMatrixRework.h:437-453) */
static int mask_value(const int fd[4], const int blk[4], int sf, int kind, int p) {
  if (kind == 0) return 1;
  int dir, disp, c[4];
  orc_point_dir_disp(p, &dir, &disp);
  coords(fd, sf, c);
  int inner;
  if (disp == 0) inner = 1;
  else if (disp == +1) inner = (c[dir] % blk[dir]) != blk[dir] - 1;
  else inner = (c[dir] % blk[dir]) != 0;
  return kind == 1 ? inner : !inner;
}

int orc_lut_populate(const int fd[4], const int blk[4], int kind, int point, int64_t* lut,
                     int64_t* sizes, int64_t* reverse) {
  int cd_[4];
  int bv = 1;
  for (int d = 0; d < 4; ++d) {
    if (fd[d] % blk[d] != 0) return -1;
    cd_[d] = fd[d] / blk[d];
    bv *= blk[d];
  }
  const int Vc = vol(cd_);
  for (int sc = 0; sc < Vc; ++sc) {
    int cc[4];
    coords(cd_, sc, cc);
    int count = 0;
    for (int sb = 0; sb < bv; ++sb) {
      int cb[4], cf[4];
      coords(blk, sb, cb);
      for (int d = 0; d < 4; ++d) cf[d] = cc[d] * blk[d] + cb[d];
      int sf = lex(fd, cf);
      if (mask_value(fd, blk, sf, kind, point)) lut[(size_t)sc * bv + count++] = sf;
      reverse[sf] = sc;
    }
    sizes[sc] = count;
  }
  return 0;
}

static int coarse_of(const int fd[4], const int blk[4], int sf) {
  int c[4], cd_[4];
  coords(fd, sf, c);
  for (int d = 0; d < 4; ++d) { cd_[d] = fd[d] / blk[d]; c[d] /= blk[d]; }
  return lex(cd_, c);
}

/* c(sc)(k, i) = sum_{sf in lut(sc)} sum_{e: chirality k} <basis_i(sf)(e), f(sf)(e)> */
void orc_project_ef(const int fd[4], const int blk[4], int nb, int Ef, const double* basis,
                    const double* fine, double* coarse, int kind, int point) {
  const int Vf = vol(fd), half = Ef / 2;
  int cd_[4];
  for (int d = 0; d < 4; ++d) cd_[d] = fd[d] / blk[d];
  const int Vc = vol(cd_);
  const cd* B = (const cd*)basis;
  const cd* f = (const cd*)fine;
  cd* c = (cd*)coarse;
  memset(c, 0, sizeof(cd) * (size_t)Vc * 2 * nb);
  for (int sf = 0; sf < Vf; ++sf) {
    if (!mask_value(fd, blk, sf, kind, point)) continue;
    const int sc = coarse_of(fd, blk, sf);
    for (int i = 0; i < nb; ++i) {
      const cd* u = B + ((size_t)i * Vf + sf) * Ef;
      for (int k = 0; k < 2; ++k) {
        cd acc = {0, 0};
        for (int e = k * half; e < (k + 1) * half; ++e) {
          cd a = u[e], b = f[(size_t)sf * Ef + e];
          acc.re += a.re * b.re + a.im * b.im;
          acc.im += a.re * b.im - a.im * b.re;
        }
        cd* t = c + (size_t)sc * 2 * nb + k * nb + i;
        t->re += acc.re;
        t->im += acc.im;
      }
    }
  }
}

void orc_project(const int fd[4], const int blk[4], int nb, const double* basis, const double* fine,
                 double* coarse, int kind, int point) {
  orc_project_ef(fd, blk, nb, 12, basis, fine, coarse, kind, point);
}

/* f(sf)(e) = sum_i c(sc)(chirality(e), i) basis_i(sf)(e) */
void orc_promote_ef(const int fd[4], const int blk[4], int nb, int Ef, const double* basis,
                    const double* coarse, double* fine) {
  const int Vf = vol(fd), half = Ef / 2;
  const cd* B = (const cd*)basis;
  const cd* c = (const cd*)coarse;
  cd* f = (cd*)fine;
  for (int sf = 0; sf < Vf; ++sf) {
    const int sc = coarse_of(fd, blk, sf);
    for (int e = 0; e < Ef; ++e) {
      const int k = e / half;
      cd acc = {0, 0};
      for (int i = 0; i < nb; ++i) {
        cd a = c[(size_t)sc * 2 * nb + k * nb + i], b = B[((size_t)i * Vf + sf) * Ef + e];
        acc.re += a.re * b.re - a.im * b.im;
        acc.im += a.re * b.im + a.im * b.re;
      }
      f[(size_t)sf * Ef + e] = acc;
    }
  }
}

void orc_promote(const int fd[4], const int blk[4], int nb, const double* basis,
                 const double* coarse, double* fine) {
  orc_promote_ef(fd, blk, nb, 12, basis, coarse, fine);
}

/* per block and chirality: modified Gram-Schmidt, v -= <u,v> u for u < v with
 * the updated v, then v *= <v,v>^{-1/2}; `passes` sweeps */
void orc_orthogonalise_ef(const int fd[4], const int blk[4], int nb, int Ef, double* basis,
                          int passes) {
  const int Vf = vol(fd), half = Ef / 2;
  int cd_[4], bv = 1;
  for (int d = 0; d < 4; ++d) { cd_[d] = fd[d] / blk[d]; bv *= blk[d]; }
  const int Vc = vol(cd_);
  cd* B = (cd*)basis;
  int64_t* lut = (int64_t*)malloc(sizeof(int64_t) * (size_t)Vc * bv);
  int64_t* sizes = (int64_t*)malloc(sizeof(int64_t) * Vc);
  int64_t* rev = (int64_t*)malloc(sizeof(int64_t) * Vf);
  orc_lut_populate(fd, blk, 0, 8, lut, sizes, rev);
  for (int pass = 0; pass < passes; ++pass) {
#pragma omp parallel for schedule(static)
    for (int sc = 0; sc < Vc; ++sc) {
      const int64_t* L = lut + (size_t)sc * bv;
      for (int v = 0; v < nb; ++v) {
        for (int u = 0; u < v; ++u) {
          cd alpha[2] = {{0, 0}, {0, 0}};
          for (int q = 0; q < sizes[sc]; ++q) {
            const cd* pu = B + ((size_t)u * Vf + L[q]) * Ef;
            const cd* pv = B + ((size_t)v * Vf + L[q]) * Ef;
            for (int e = 0; e < Ef; ++e) {
              const int k = e / half;
              alpha[k].re += pu[e].re * pv[e].re + pu[e].im * pv[e].im;
              alpha[k].im += pu[e].re * pv[e].im - pu[e].im * pv[e].re;
            }
          }
          for (int q = 0; q < sizes[sc]; ++q) {
            const cd* pu = B + ((size_t)u * Vf + L[q]) * Ef;
            cd* pv = B + ((size_t)v * Vf + L[q]) * Ef;
            for (int e = 0; e < Ef; ++e) {
              const cd a = alpha[e / half];
              pv[e].re -= a.re * pu[e].re - a.im * pu[e].im;
              pv[e].im -= a.re * pu[e].im + a.im * pu[e].re;
            }
          }
        }
        double nrm[2] = {0, 0};
        for (int q = 0; q < sizes[sc]; ++q) {
          const cd* pv = B + ((size_t)v * Vf + L[q]) * Ef;
          for (int e = 0; e < Ef; ++e) nrm[e / half] += pv[e].re * pv[e].re + pv[e].im * pv[e].im;
        }
        nrm[0] = pow(nrm[0], -0.5);
        nrm[1] = pow(nrm[1], -0.5);
        for (int q = 0; q < sizes[sc]; ++q) {
          cd* pv = B + ((size_t)v * Vf + L[q]) * Ef;
          for (int e = 0; e < Ef; ++e) {
            pv[e].re *= nrm[e / half];
            pv[e].im *= nrm[e / half];
          }
        }
      }
    }
  }
  free(lut);
  free(sizes);
  free(rev);
}

void orc_orthogonalise(const int fd[4], const int blk[4], int nb, double* basis, int passes) {
  orc_orthogonalise_ef(fd, blk, nb, 12, basis, passes);
}

/* ---------- doubled (Upstream) forms: 2*nb full vectors ---------- */
void orc_chiral_double(int Vf, int nb, const double* in, double* out) {
  /* v_n = 1/2 (1 + g5) v, v_{n+nb} = 1/2 (1 - g5) v with g5 = diag(1,1,-1,-1) */
  const cd* u = (const cd*)in;
  cd* o = (cd*)out;
  for (int n = 0; n < nb; ++n)
    for (int x = 0; x < Vf; ++x)
      for (int e = 0; e < 12; ++e) {
        const cd z = u[((size_t)n * Vf + x) * 12 + e];
        const cd zero = {0, 0};
        o[((size_t)n * Vf + x) * 12 + e] = e < 6 ? z : zero;
        o[((size_t)(n + nb) * Vf + x) * 12 + e] = e < 6 ? zero : z;
      }
}

/* Grid blockProject on N doubled vectors: c(sc)(n) = sum_x <v_n(x), f(x)> */
void orc_project_doubled(const int fd[4], const int blk[4], int N, const double* basis,
                         const double* fine, double* coarse) {
  const int Vf = vol(fd);
  int cd_[4];
  for (int d = 0; d < 4; ++d) cd_[d] = fd[d] / blk[d];
  const cd* B = (const cd*)basis;
  const cd* f = (const cd*)fine;
  cd* c = (cd*)coarse;
  memset(c, 0, sizeof(cd) * (size_t)vol(cd_) * N);
  for (int x = 0; x < Vf; ++x) {
    const int sc = coarse_of(fd, blk, x);
    for (int n = 0; n < N; ++n) {
      cd acc = {0, 0};
      for (int e = 0; e < 12; ++e) {
        cd a = B[((size_t)n * Vf + x) * 12 + e], b = f[(size_t)x * 12 + e];
        acc.re += a.re * b.re + a.im * b.im;
        acc.im += a.re * b.im - a.im * b.re;
      }
      c[(size_t)sc * N + n].re += acc.re;
      c[(size_t)sc * N + n].im += acc.im;
    }
  }
}

/* blockLutedOrthonormalise (Lut.h:176-239): MGS over all N doubled vectors per block */
void orc_orthogonalise_doubled(const int fd[4], const int blk[4], int N, double* basis,
                               int passes) {
  const int Vf = vol(fd);
  cd* B = (cd*)basis;
  int cd_[4];
  for (int d = 0; d < 4; ++d) cd_[d] = fd[d] / blk[d];
  const int Vc = vol(cd_);
  for (int pass = 0; pass < passes; ++pass)
    for (int sc = 0; sc < Vc; ++sc)
      for (int v = 0; v < N; ++v) {
        for (int u = 0; u < v; ++u) {
          cd a = {0, 0};
          for (int x = 0; x < Vf; ++x) {
            if (coarse_of(fd, blk, x) != sc) continue;
            for (int e = 0; e < 12; ++e) {
              cd pu = B[((size_t)u * Vf + x) * 12 + e], pv = B[((size_t)v * Vf + x) * 12 + e];
              a.re += pu.re * pv.re + pu.im * pv.im;
              a.im += pu.re * pv.im - pu.im * pv.re;
            }
          }
          for (int x = 0; x < Vf; ++x) {
            if (coarse_of(fd, blk, x) != sc) continue;
            for (int e = 0; e < 12; ++e) {
              cd pu = B[((size_t)u * Vf + x) * 12 + e];
              cd* pv = &B[((size_t)v * Vf + x) * 12 + e];
              pv->re -= a.re * pu.re - a.im * pu.im;
              pv->im -= a.re * pu.im + a.im * pu.re;
            }
          }
        }
        double nrm = 0;
        for (int x = 0; x < Vf; ++x) {
          if (coarse_of(fd, blk, x) != sc) continue;
          for (int e = 0; e < 12; ++e) {
            cd pv = B[((size_t)v * Vf + x) * 12 + e];
            nrm += pv.re * pv.re + pv.im * pv.im;
          }
        }
        nrm = pow(nrm, -0.5);
        for (int x = 0; x < Vf; ++x) {
          if (coarse_of(fd, blk, x) != sc) continue;
          for (int e = 0; e < 12; ++e) {
            B[((size_t)v * Vf + x) * 12 + e].re *= nrm;
            B[((size_t)v * Vf + x) * 12 + e].im *= nrm;
          }
        }
      }
}

/* CoarsenOperator of a coarse operator (links1 FLAT [9][V1][N1][N1]) on the
 * split basis of the next level (nb2 vectors of N1 elements per site):
 * for each doubled vector i and point p, Mphi = A_p v_i, outer projection to
 * A2_p(:, i), inner projection added to A2_8(:, i).  links2: [9][V2][N2][N2]. */
void orc_coarsen_coarse_operator(const int d1[4], const int blk[4], int nb2, int N1,
                                 const double* links1, const double* basis, double* links2) {
  int d2[4];
  for (int k = 0; k < 4; ++k) d2[k] = d1[k] / blk[k];
  const int V1 = vol(d1), V2 = vol(d2), N2 = 2 * nb2, half = N1 / 2;
  cd* A = (cd*)links2;
  memset(A, 0, sizeof(cd) * 9 * (size_t)V2 * N2 * N2);
  double* phi = (double*)malloc(sizeof(double) * 2 * (size_t)V1 * N1);
  double* Mphi = (double*)malloc(sizeof(double) * 2 * (size_t)V1 * N1);
  double* proj = (double*)malloc(sizeof(double) * 2 * (size_t)V2 * N2);
  for (int i = 0; i < N2; ++i) {
    const int l = i / nb2, c = i % nb2;
    const double* u = basis + 2 * (size_t)c * V1 * N1;
    for (size_t x = 0; x < (size_t)V1; ++x)
      for (int e = 0; e < N1; ++e) {
        const int keep = (e / half) == l;
        phi[2 * (x * N1 + e)] = keep ? u[2 * (x * N1 + e)] : 0.0;
        phi[2 * (x * N1 + e) + 1] = keep ? u[2 * (x * N1 + e) + 1] : 0.0;
      }
    for (int p = 0; p < 9; ++p) {
      orc_coarse_apply(d1, N1, links1, phi, Mphi, 1u << p, 0);
      if (p < 8) {
        orc_project_ef(d1, blk, nb2, N1, basis, Mphi, proj, 2, p);
        const cd* pr = (const cd*)proj;
        for (int xc = 0; xc < V2; ++xc)
          for (int j = 0; j < N2; ++j) A[(((size_t)p * V2 + xc) * N2 + j) * N2 + i] = pr[(size_t)xc * N2 + j];
      }
      orc_project_ef(d1, blk, nb2, N1, basis, Mphi, proj, 1, p);
      const cd* pr = (const cd*)proj;
      for (int xc = 0; xc < V2; ++xc)
        for (int j = 0; j < N2; ++j) {
          cd* a = &A[(((size_t)8 * V2 + xc) * N2 + j) * N2 + i];
          a->re += pr[(size_t)xc * N2 + j].re;
          a->im += pr[(size_t)xc * N2 + j].im;
        }
    }
  }
  free(phi);
  free(Mphi);
  free(proj);
}
