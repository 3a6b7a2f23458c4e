/*
 * oracle.h -- CPU FP64 restatement of the reference's multigrid coarse-level
 * algorithms.  TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check of
 * __graft_entry__.py and by the CPU-baseline leg of bench.py, never by the
 * library (grid-mg-bench_amd/) it checks.
 *
 * Conventions are those of include/mgc.h: complex numbers interleaved (re, im)
 * as double[2], lattice sites lexicographic with dimension 0 fastest, stencil
 * points 0..3 = +mu, 4..7 = -mu, 8 = self (include/GeometryRework.h:70-78).
 */
#ifndef MGC_ORACLE_H
#define MGC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* number of OpenMP threads the oracle's parallel loops use */
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* ---- counter-based Gaussian generator (same definition as libmgc) ---- */
void orc_gaussian(double* out, size_t n_complex, uint64_t seed, uint64_t stream);

/* ---- geometry (include/GeometryRework.h:65-125) ---- */
int orc_point_from_dir_disp(int dir, int disp);
void orc_point_dir_disp(int point, int* dir, int* disp);
int orc_site_shift(const int dims[4], int x, int dir, int disp);
int orc_parity(const int dims[4], int x);
/* half-lattice index of full site x among the sites of its parity, counted in
 * lexicographic order (Grid's red-black grid) */
int orc_half_index(const int dims[4], int x);

/* ---- coarse stencil, FLAT links [9][V][N][N] row-major ---- */
void orc_coarse_apply(const int dims[4], int N, const double* links, const double* in,
                      double* out, unsigned mask, int dag);
void orc_coarse_apply_cb(const int dims[4], int N, const double* links, const double* in_half,
                         double* out_half, int out_cb, unsigned mask, int dag);
/* per-site N x N matrices [V][N][N] row-major applied on the full (cb = -1)
 * or a half lattice; adjoint != 0 applies the Hermitian conjugate */
void orc_site_matrix_apply(const int dims[4], int N, const double* mats, const double* in,
                           double* out, int cb, int adjoint);
/* A_inv = A^{-1} (N x N row-major), Gauss-Jordan with partial pivoting;
 * returns 0 on success, -1 when A is singular */
int orc_invert(int N, const double* A, double* A_inv);
int orc_invert_self(const int dims[4], int N, const double* links, double* inv);
void orc_links_flat_to_split(int V, int N, const double* flat, double* split);
void orc_links_split_to_flat(int V, int N, const double* split, double* flat);
/* multi-rhs M on vectors [V][nrhs][N] (include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h:736-804) */
void orc_coarse_mrhs_M(const int dims[4], int N, int nrhs, const double* links, const double* in,
                       double* out);
/* gamma5-hermiticity link reconstruction: disp_have = -1 builds the forward
 * links from the backward ones (include/CoarsenedMatrixUpstreamImprovedDirsave.h:999-1032),
 * disp_have = +1 the converse (include/CoarsenedMatrixKernelsRework.h:141-167) */
void orc_construct_remaining_links(const int dims[4], int N, double* links, int disp_have);

/* ---- block transfer (aggregation.c) ---- */
/* LUT: lut[sc][0..sizes[sc]) = fine sites of block sc with mask != 0
 * (kind 0 = all, 1 = inner of point, 2 = outer of point); reverse[sf] = sc */
int orc_lut_populate(const int fine_dims[4], const int block[4], int kind, int point,
                     int64_t* lut, int64_t* sizes, int64_t* reverse);
void orc_project(const int fine_dims[4], const int block[4], int nb, const double* basis,
                 const double* fine, double* coarse, int kind, int point);
void orc_promote(const int fine_dims[4], const int block[4], int nb, const double* basis,
                 const double* coarse, double* fine);
void orc_orthogonalise(const int fine_dims[4], const int block[4], int nb, double* basis,
                       int passes);
/* the same on fields of Ef elements per site (chirality = element / (Ef/2)) */
void orc_project_ef(const int fine_dims[4], const int block[4], int nb, int Ef,
                    const double* basis, const double* fine, double* coarse, int kind, int point);
void orc_promote_ef(const int fine_dims[4], const int block[4], int nb, int Ef,
                    const double* basis, const double* coarse, double* fine);
void orc_orthogonalise_ef(const int fine_dims[4], const int block[4], int nb, int Ef,
                          double* basis, int passes);
void orc_coarsen_coarse_operator(const int dims1[4], const int block[4], int nb2, int N1,
                                 const double* links1, const double* basis, double* links2);
void orc_chiral_double(int Vf, int nb, const double* in, double* out);
void orc_project_doubled(const int fine_dims[4], const int block[4], int N, const double* basis,
                         const double* fine, double* coarse);
void orc_orthogonalise_doubled(const int fine_dims[4], const int block[4], int N, double* basis,
                               int passes);

/* ---- fine Wilson-clover operator and coarse operator construction (wilson.c) ---- */
void orc_gamma(int mu, double* out4x4);
/* (4+m) + clover term per site, [V][12][12] */
void orc_wilson_diag(const int dims[4], const double* gauge, double mass, double csw, double* diag);
void orc_wilson_dir(const int dims[4], const double* gauge, const double* in, double* out, int mu,
                    int disp);
void orc_wilson_diag_apply(int V, const double* diag, const double* in, double* out);
void orc_wilson_M(const int dims[4], const double* gauge, const double* diag, const double* in,
                  double* out);
void orc_coarsen_operator(const int fine_dims[4], const int block[4], int nb, const double* gauge,
                          const double* diag, const double* basis, double* links);

#ifdef __cplusplus
}
#endif
#endif
