/*
 * mgc.h -- C ABI of the MI355X multigrid coarse-level library (libmgc.so).
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an `int` status (MGC_OK == 0, negative on error; the message of the
 * last error on the calling thread is returned by mgc_last_error()).  No C++
 * exception crosses this boundary.
 *
 * Memory and streams
 *   - Field pointers passed to the apply/transfer calls are DEVICE pointers
 *     owned by the caller.  The library only reads `in` and writes `out`.
 *   - Every call that launches work takes a HIP stream (`void*`, a
 *     hipStream_t; NULL is the null stream) and is asynchronous on it.
 *   - A handle is used by one host thread at a time; distinct handles are
 *     independent.  A handle keeps device workspaces (e.g. the partial sums
 *     of a split projection), so work of one handle issued on different
 *     streams must not overlap on the device.
 *
 * Layouts (all complex numbers are interleaved re,im)
 *   - Lattice sites are lexicographic with dimension 0 fastest:
 *       x = x0 + L0*(x1 + L1*(x2 + L2*x3)).
 *   - A half-checkerboard field of parity cb holds the sites with
 *     (x0+x1+x2+x3)%2 == cb, site h = x>>1 (L0 must be even).
 *   - Coarse vector: [site][N].  Multi-rhs coarse vector: [site][rhs][N]
 *     (rhs index between site and basis index, as the 5d rhs grid of the
 *     reference).
 *   - Coarse links, FLAT layout: [point][site][row][col], N x N row-major
 *     per (point, site).  SPLIT layout: [point][site][s1][s2][c1][c2] with
 *     Ns_c = 2 chiralities and row n = s*(N/2)+c.
 *   - Stencil points: 0..3 = +x,+y,+z,+t; 4..7 = -x,-y,-z,-t; 8 = self.
 *     out(x) = sum_p A_p(x) in(x + delta_p), periodic.
 *   - Fine spinor field: [site][spin 4][colour 3].  Gauge field:
 *     [mu 4][site][3][3] row-major.  Basis (Aggregation subspace):
 *     [nb][fine site][12], one fine field per basis vector.
 */
#ifndef MGC_H
#define MGC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status codes, precisions, kinds                                    */
/* ------------------------------------------------------------------ */
enum {
  MGC_OK = 0,
  MGC_ERR_ARG = -1,       /* null pointer / bad enum / bad value        */
  MGC_ERR_SHAPE = -2,     /* lattice / block / nbasis not conformable   */
  MGC_ERR_PRECISION = -3, /* precision not supported by this call       */
  MGC_ERR_HIP = -4,       /* a HIP runtime call failed                  */
  MGC_ERR_STATE = -5,     /* call made before the data it needs was set */
  MGC_ERR_SINGULAR = -6   /* a self link could not be inverted          */
};

enum { MGC_F64 = 0, MGC_F32 = 1, MGC_F16 = 2, MGC_BF16 = 3 };

enum { MGC_LAYOUT_FLAT = 0, MGC_LAYOUT_SPLIT = 1 };

/* parity argument of the checkerboarded calls */
enum { MGC_EVEN = 0, MGC_ODD = 1, MGC_FULL = -1 };

typedef struct mgc_coarse_op mgc_coarse_op;

/* ------------------------------------------------------------------ */
/* runtime helpers (device memory, streams, events)                   */
/* ------------------------------------------------------------------ */
const char* mgc_last_error(void);
int mgc_version(void);
int mgc_device_count(int* count);
int mgc_device_name(char* buf, int len);
int mgc_malloc(void** ptr, size_t bytes);
int mgc_free(void* ptr);
int mgc_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int mgc_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int mgc_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int mgc_fill_zero(void* ptr, size_t bytes, void* stream);
/* device-to-device copy by a kernel (16-byte aligned), e.g. for calibrating
 * hardware byte counters against a copy of known size */
int mgc_device_copy(void* dst, const void* src, size_t bytes, void* stream);
int mgc_stream_create(void** stream);
int mgc_stream_destroy(void* stream);
int mgc_stream_synchronize(void* stream);
int mgc_device_synchronize(void);
int mgc_event_create(void** event);
int mgc_event_destroy(void* event);
int mgc_event_record(void* event, void* stream);
int mgc_event_synchronize(void* event);
int mgc_event_elapsed_ms(float* ms, void* start, void* stop);

/* ------------------------------------------------------------------ */
/* synthetic inputs: counter-based Gaussian RNG                        */
/* ------------------------------------------------------------------ */
/* Fill n complex numbers (precision prec) with independent re, im ~
 * N(0, 1/2) from the counter-based generator keyed by (seed, stream);
 * element i of the output is a pure function of (seed, stream, i).
 * Replaces Grid's gaussian(GridParallelRNG&, Field&) as used by
 * src/Benchmark_coarse_operator.cc:151 (random inputs of the benchmarks). */
int mgc_random_gaussian(void* out, size_t n_complex, int prec, uint64_t seed,
                        uint64_t stream, void* hip_stream);

/* ------------------------------------------------------------------ */
/* coarse operator (CoarsenedMatrix)                                  */
/* ------------------------------------------------------------------ */
/* Constructor.  Replaces CoarsenedMatrix(GridCartesian&, GridRedBlackCartesian&,
 * int hermitian) of include/CoarsenedMatrixBaseline.h:557 and
 * include/CoarsenedMatrixRework.h:82.  dims = coarse lattice extents,
 * nbasis = N (even), prec = storage/compute precision of the links and of
 * the single-rhs vectors (MGC_F64 or MGC_F32; MGC_F16/MGC_BF16 operators
 * serve the multi-rhs call only). */
int mgc_coarse_op_create(mgc_coarse_op** op, const int dims[4], int nbasis,
                         int prec, int hermitian);
int mgc_coarse_op_destroy(mgc_coarse_op* op);
int mgc_coarse_op_info(const mgc_coarse_op* op, int dims[4], int* nbasis,
                       int* prec, int* volume);

/* Install the 9 link fields from a device buffer of complex doubles in the
 * FLAT or SPLIT layout.  Replaces the direct writes of the public link
 * members A[p] / Y_[p] (src/Benchmark_coarse_operator.cc:121-127) and the
 * layout converters of include/Layout_converters.h:33-89.  Also computes the
 * self-link inverses (InvertSelfStencilPoint, include/CoarsenedMatrixRework.h:640)
 * when invert != 0. */
int mgc_coarse_op_set_links(mgc_coarse_op* op, int layout, const void* links_f64,
                            int invert, void* stream);
/* Export the links (complex double, FLAT or SPLIT layout). */
int mgc_coarse_op_get_links(const mgc_coarse_op* op, int layout, void* links_f64,
                            void* stream);
/* Export the inverse self links [site][N][N] row-major (complex double). */
int mgc_coarse_op_get_self_inverse(const mgc_coarse_op* op, void* out_f64, void* stream);

/* out = M in.  include/CoarsenedMatrixUpstream.h:546-607,
 * include/CoarsenedMatrixRework.h:109-163, include/CoarsenedMatrixBaseline.h:314-344. */
int mgc_coarse_op_M(mgc_coarse_op* op, const void* in, void* out, void* stream);
/* out = Mdag in (G5C M G5C for Galerkin, M for hermitian).
 * include/CoarsenedMatrixUpstream.h:609-621, include/CoarsenedMatrixRework.h:165-176. */
int mgc_coarse_op_Mdag(mgc_coarse_op* op, const void* in, void* out, void* stream);
/* out = A_8 in.  include/CoarsenedMatrixUpstream.h:738-742. */
int mgc_coarse_op_Mdiag(mgc_coarse_op* op, const void* in, void* out, void* stream);
/* out = A_p in(x+delta_p) for the point of (dir, disp); disp = 0 means self.
 * include/CoarsenedMatrixUpstream.h:704-736, include/CoarsenedMatrixRework.h:239-286. */
int mgc_coarse_op_Mdir(mgc_coarse_op* op, const void* in, void* out, int dir, int disp,
                       void* stream);
/* out[p] (8 consecutive coarse vectors) = A_p in(x+delta_p), p = 0..7.
 * include/CoarsenedMatrixUpstream.h:691-703. */
int mgc_coarse_op_MdirAll(mgc_coarse_op* op, const void* in, void* out8, void* stream);
/* out = sum_{p<8} A_p in(x+delta_p) (dag: G5C Dhop G5C), full lattice.
 * include/CoarsenedMatrixRework.h:210-217, 288-399. */
int mgc_coarse_op_Dhop(mgc_coarse_op* op, const void* in, void* out, int dag, void* stream);
/* Half-lattice hopping: in Even -> out Odd (DhopOE), in Odd -> out Even (DhopEO).
 * include/CoarsenedMatrixRework.h:219-237. */
int mgc_coarse_op_DhopOE(mgc_coarse_op* op, const void* in_even, void* out_odd, int dag,
                         void* stream);
int mgc_coarse_op_DhopEO(mgc_coarse_op* op, const void* in_odd, void* out_even, int dag,
                         void* stream);
/* Meooe / MeooeDag dispatch on the parity of `in`.  include/CoarsenedMatrixRework.h:186-200. */
int mgc_coarse_op_Meooe(mgc_coarse_op* op, const void* in, void* out, int in_cb, void* stream);
int mgc_coarse_op_MeooeDag(mgc_coarse_op* op, const void* in, void* out, int in_cb,
                           void* stream);
/* Self-link applications on the full lattice (cb = MGC_FULL) or a half
 * lattice (cb = MGC_EVEN/MGC_ODD).  include/CoarsenedMatrixRework.h:202-208, 617-638. */
int mgc_coarse_op_Mooee(mgc_coarse_op* op, const void* in, void* out, int cb, void* stream);
int mgc_coarse_op_MooeeInv(mgc_coarse_op* op, const void* in, void* out, int cb, void* stream);
int mgc_coarse_op_MooeeDag(mgc_coarse_op* op, const void* in, void* out, int cb, void* stream);
int mgc_coarse_op_MooeeInvDag(mgc_coarse_op* op, const void* in, void* out, int cb,
                              void* stream);

/* ------------------------------------------------------------------ */
/* multi-rhs coarse stencil on MFMA                                    */
/* ------------------------------------------------------------------ */
/* out[site][nrhs][N] (complex FP32) = M in (dag: G5C M G5C), all rhs in one
 * launch, links read once for every rhs.  Replaces the 5d-grid M of
 * include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h:736-804 (rhs index
 * between site and basis index, links shared over rhs).
 *   mrhs_prec = MGC_F32  : FP32 MFMA, in must be complex FP32.
 *   mrhs_prec = MGC_F16 / MGC_BF16 : half-precision MFMA with FP32
 *     accumulation; links carry one FP32 scale per (site, point); in is
 *     either complex FP32 (quantised on the fly) or already half precision
 *     (in_prec == mrhs_prec) with in_scales[site][nrhs] from mgc_quantize.
 * nbasis must be a multiple of 8 up to 64 (MGC_ERR_SHAPE otherwise); other
 * than 16, 32, 48 and 64 the operand is zero-padded to the next multiple of 16.
 * The operator keeps a swizzled copy of its links for the last mrhs_prec
 * used (built from the F64/F32 links on first use or by prepare_mrhs), and a
 * quantisation workspace for FP32 input to an F16/BF16 apply, grown on the
 * first call that needs more and kept for later calls. */
int mgc_coarse_op_prepare_mrhs(mgc_coarse_op* op, int mrhs_prec, void* stream);
int mgc_coarse_op_M_mrhs(mgc_coarse_op* op, const void* in, const float* in_scales, int in_prec,
                         void* out, int nrhs, int mrhs_prec, int dag, void* stream);
/* The multi-rhs M of up to 8 coarse lattices of any volume and one nbasis in
 * a single launch (one workgroup per site of every lattice): lattice l maps
 * in[l] (and in_scales[l] for F16/BF16) to out[l] exactly as
 * mgc_coarse_op_M_mrhs would.  Inputs must already be in mrhs_prec
 * (in_prec == mrhs_prec; mgc_quantize makes half-precision ones).  Replaces
 * one call of the MRHS M per coarse lattice
 * (include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h:736-804), so that
 * the small coarse levels of several solves fill the device together. */
int mgc_coarse_op_M_mrhs_batched(mgc_coarse_op* const* ops, int count, const void* const* in,
                                 const float* const* in_scales, int in_prec, void* const* out,
                                 int nrhs, int mrhs_prec, int dag, void* stream);
/* The multi-rhs M of mgc_coarse_op_M_mrhs for complex FP32 input, calling no
 * device allocator: no counterpart in the reference.  It is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  For
 * mrhs_prec MGC_F16 / MGC_BF16 the quantised input is staged in a 64 MiB
 * scratch array of the library that every device holds until the process
 * ends, so sites x nrhs x (4 nbasis + 4) bytes must not exceed 64 MiB
 * (MGC_ERR_SHAPE otherwise); MGC_F32 needs no scratch.  The operand must be
 * prepared for mrhs_prec beforehand (mgc_coarse_op_prepare_mrhs, else
 * MGC_ERR_STATE), and a shape the autotuner has not timed runs the default
 * kernel variant.  All calls on one device share its scratch: calls that
 * overlap on one device race. */
int mgc_coarse_op_M_mrhs_noalloc(mgc_coarse_op* op, const void* in_f32, void* out, int nrhs,
                                 int mrhs_prec, int dag, void* stream);
/* Kernel-variant control of the multi-rhs stencil.  By default the first
 * call for a (precision, nbasis, column-tile, lattice) shape times every
 * variant on its own operands and keeps the fastest (autotuner).  variant >= 0 forces a
 * variant (0 = LDS-staged operand, 1 = LDS-staged operand with non-temporal
 * link loads, 2 = register operand with two m-tiles per wavefront, 3 and 4 =
 * variants 0 and 1 with sites dealt to the eight XCDs in contiguous ranges,
 * 5 = persistent workgroups walking the sites, 6 = persistent with
 * non-temporal link loads and the XCD-contiguous order, 7 = input copied
 * straight to LDS with two m-tiles per wavefront, 8 = variant 7 with
 * non-temporal link loads and the XCD-contiguous order, 9 = variant 1 with
 * each XCD taking a 4d block of the lattice (halves of x, y and z, all t),
 * 10 = variant 9 with the input and link loads issued two points ahead;
 * variants 2, 7 and 8 are built for nbasis 32 and 64 and run as 0, 3 and 4
 * for the other nbasis values; the batched apply
 * uses variants 0, 1 and 3 to 6 and tunes itself when another one is
 * forced); -1 restores autotuning.  get_variant reports the forced variant,
 * else the one tuned last for that (precision, nbasis, rhs count) on any
 * lattice (-1 = not tuned yet). */
int mgc_mrhs_set_variant(int variant);
int mgc_mrhs_get_variant(int mrhs_prec, int nbasis, int nrhs, int* variant);
/* rows x n complex FP32 -> rows x n complex F16/BF16 with scales[row] =
 * max |component| of the row (values divided by it). */
int mgc_quantize(const void* in_f32, size_t rows, int n, int prec, void* out, float* scales,
                 void* stream);

/* ------------------------------------------------------------------ */
/* block transfer (Aggregation), chirality-split form: nb basis        */
/* vectors, coarse vectors of N = 2*nb with n = k*nb + i, k = spin/2    */
/* ------------------------------------------------------------------ */
typedef struct mgc_aggregation mgc_aggregation;

/* Aggregation(GridBase* coarse, GridBase* fine, int cb, int fast)
 * (include/AggregationRework.h:64-73); block = fine/coarse extents, which
 * must divide the fine lattice (subdivides, :71). */
int mgc_aggregation_create(mgc_aggregation** agg, const int fine_dims[4], const int block[4],
                           int nb);
/* Same for a "fine" level whose sites carry site_elems (even) components with
 * chirality = element / (site_elems/2): site_elems = 12 for a Dirac field,
 * = N for a coarse field being coarsened again (coarse-of-coarse aggregation,
 * src/Benchmark_create_coarse_coarse_operator.cc:240-244). */
int mgc_aggregation_create_ex(mgc_aggregation** agg, const int fine_dims[4], const int block[4],
                              int nb, int site_elems);
int mgc_aggregation_destroy(mgc_aggregation* agg);
int mgc_aggregation_info(const mgc_aggregation* agg, int coarse_dims[4], int* nb,
                         int* fine_volume, int* coarse_volume);
/* Copy nb fine fields [nb][fine site][12] (complex double) into the subspace:
 * the copy step of Create (include/AggregationRework.h:133-146) and the
 * writable Subspace() (:75). */
int mgc_aggregation_set_basis(mgc_aggregation* agg, const void* basis_f64, void* stream);
int mgc_aggregation_get_basis(const mgc_aggregation* agg, void* basis_f64, void* stream);
/* Block orthonormalisation per (block, chirality), FP64: Cholesky QR in two
 * rounds (see mgc_aggregation_set_ortho_method), or `passes` sweeps over the
 * columns (each column's inner products reduced together, projected again
 * when its norm drops below 1/sqrt(2)); for a basis of full rank the result is
 * the modified Gram-Schmidt one of the reference to rounding.
 * Orthogonalise (include/AggregationRework.h:80-89) -> aggregateOrthogonalise
 * (include/AggregationKernelsRework.h:266-341). */
int mgc_aggregation_orthogonalise(mgc_aggregation* agg, int passes, void* stream);
/* Method of mgc_aggregation_orthogonalise, process-wide: 0 (default) =
 * Cholesky QR in two rounds (three streaming reads of each block panel and
 * one write; a panel whose columns are nearly dependent, pivot^2 < 1e-14 of
 * its column norm^2, falls back to the column sweep; the sweep alone for nb >
 * 32), 1 = the column sweep everywhere.  Both give the reference's Q of a
 * full-rank panel to rounding.  No counterpart in the reference: it serves
 * tests and measurements. */
int mgc_aggregation_set_ortho_method(int method);
/* CheckOrthogonal (include/AggregationRework.h:91-108): errs_host[i] =
 * norm2(e_i - R v_i) summed over coarse sites (host array of nb doubles). */
int mgc_aggregation_check_orthogonal(mgc_aggregation* agg, double* errs_host, void* stream);
/* ProjectToSubspace: coarse[Vc][nrhs][N] = R fine[Vf][nrhs][12]
 * (include/AggregationRework.h:110-124, include/AggregationKernelsRework.h:96-157).
 * prec = MGC_F64 or MGC_F32 vectors (FP32 uses an FP32 copy of the basis). */
int mgc_aggregation_project(mgc_aggregation* agg, const void* fine, void* coarse, int prec,
                            int nrhs, void* stream);
/* ProjectToSubspace with a lookup table of masked sites
 * (include/AggregationRework.h:117-124, include/CoarsenedMatrixRework.h:437-453):
 * mask_kind 0 = all sites, 1 = inner sites of `point`, 2 = outer sites. */
int mgc_aggregation_project_masked(mgc_aggregation* agg, const void* fine, void* coarse,
                                   int prec, int nrhs, int mask_kind, int point, void* stream);
/* PromoteFromSubspace: fine = P coarse
 * (include/AggregationRework.h:126-131, include/AggregationKernelsRework.h:207-264). */
int mgc_aggregation_promote(mgc_aggregation* agg, const void* coarse, void* fine, int prec,
                            int nrhs, void* stream);

/* ------------------------------------------------------------------ */
/* fine Wilson-clover operator (input of CoarsenOperator, smoother)    */
/* ------------------------------------------------------------------ */
/* Convention (chiral basis, g5 = g_x g_y g_z g_t = diag(1,1,-1,-1)):
 *   D psi(x) = [(4+m) + C(x)] psi(x) - 1/2 sum_mu [(1-g_mu) U_mu(x) psi(x+mu)
 *                                                 + (1+g_mu) U_mu(x-mu)^dag psi(x-mu)]
 *   C = -(csw/2) sum_{mu<nu} sigma_{mu nu} F_{mu nu} (clover leaves).
 * Plays the role of the fine LinearOperatorBase the reference's
 * CoarsenOperator takes (OpDir / OpDiag / Op, include/CoarsenedMatrixBaseline.h:628-633)
 * with the Wilson-clover parameters of src/Test_coarse_red_black.cc:99-100.
 * Vectors: [site][nrhs][12] complex in precision prec (F64 or F32). */
typedef struct mgc_fine_op mgc_fine_op;

int mgc_fine_op_create(mgc_fine_op** op, const int dims[4], double mass, double csw,
                       int antiperiodic_t);
int mgc_fine_op_destroy(mgc_fine_op* op);
/* gauge field [mu][site][3][3] complex double (device); builds the clover term */
int mgc_fine_op_set_gauge(mgc_fine_op* op, const void* gauge_f64, void* stream);
/* (4+m) + C(x) as [site][2 chiralities][6][6] complex double (device) */
int mgc_fine_op_get_diag(const mgc_fine_op* op, void* out_f64, void* stream);
/* vectors [site][rhs][spin 4][colour 3], prec MGC_F64 or MGC_F32, 1 <= nrhs <= 256 */
int mgc_fine_op_M(mgc_fine_op* op, const void* in, void* out, int prec, int nrhs, void* stream);
int mgc_fine_op_Mdag(mgc_fine_op* op, const void* in, void* out, int prec, int nrhs,
                     void* stream);
int mgc_fine_op_Mdiag(mgc_fine_op* op, const void* in, void* out, int prec, int nrhs,
                      void* stream);
int mgc_fine_op_Mdir(mgc_fine_op* op, const void* in, void* out, int dir, int disp, int prec,
                     int nrhs, void* stream);
/* Kernel-variant control of the fine apply (all handles, process-wide):
 * 0 = each thread reads its own spinor records, 1 = each hop's records
 * gathered by the workgroup through LDS with coalesced loads, 2 = as 1 with
 * each hop's link shared through LDS by the rhs of a site and the next hop's
 * loads issued before the current hop's arithmetic (FP32 with nrhs >= 5; 1
 * otherwise), 3 = as 2, but with one rhs three threads per site, one colour
 * row each; -1 restores the default (3).  All give bit-identical results.
 * No counterpart in the reference: it serves tests and measurements. */
int mgc_fine_op_set_variant(int variant);
/* SU(3) gauge field U = exp(i eps H), H the traceless Hermitian part of a
 * Gaussian 3x3 (counter-based generator, element (mu*V + x)*9 + 3i + j);
 * the synthetic-gauge recipe of the multigrid configurations (eps = 0.4). */
int mgc_random_gauge(void* gauge_f64, const int dims[4], double eps, uint64_t seed,
                     uint64_t stream_id, void* stream);

/* ------------------------------------------------------------------ */
/* a caller's fine operator (the reference's LinearOperatorBase)       */
/* ------------------------------------------------------------------ */
/* The reference's CoarsenOperator and multigrid take any fine linear
 * operator (CoarsenOperator(GridBase*, LinearOperatorBase<FineField>&,
 * Aggregation&), include/CoarsenedMatrixRework.h:401-403; the MdagM operator
 * of include/MultigridPreconditioner.h:123-124).  A caller hands its operator
 * to the library as a callback:
 *   apply(ctx, point, prec, in, out, nrhs, stream)
 *     point 0..7 : out = the hopping term of that stencil point (OpDir with
 *                  (dir, disp) of the point order, include/CoarsenedMatrixBaseline.h:628-633)
 *     point 8    : out = the site-diagonal term (OpDiag)
 *     point 9    : out = the whole operator (Op = sum of points 0..8)
 *   in, out: device fields [site][nrhs][spin 4][colour 3], complex, in
 *   precision prec (MGC_F64 or MGC_F32, among `precisions`), nrhs <= max_nrhs;
 *   the work is enqueued on `stream` (or completed before returning); the
 *   return value is 0 on success, anything else is reported as an error.
 * The callback is called from the thread that called the library. */
typedef int (*mgc_linop_apply_fn)(void* ctx, int point, int prec, const void* in, void* out,
                                  int nrhs, void* stream);
typedef struct mgc_fine_linop {
  mgc_linop_apply_fn apply;
  void* ctx;
  int dims[4];          /* fine lattice extents */
  int precisions;       /* bit mask of accepted precisions: (1 << MGC_F64) | (1 << MGC_F32) */
  int max_nrhs;         /* largest nrhs of one call (>= 1) */
  int gamma5_hermitian; /* nonzero: D^dag = g5 D g5 (needed by MGC_COARSEN_DIRSAVE) */
} mgc_fine_linop;

/* Fill a callback operator that applies the library's own Wilson-clover
 * operator (F64 and F32, up to 256 rhs, gamma5-hermitian). */
int mgc_fine_op_linop(mgc_fine_op* op, mgc_fine_linop* out);

/* ------------------------------------------------------------------ */
/* coarse operator construction                                        */
/* ------------------------------------------------------------------ */
enum { MGC_COARSEN_FULL = 0, MGC_COARSEN_DIRSAVE = 1 };

/* CoarsenOperator(FineGrid, linop, Aggregation): the Galerkin links
 * A_p = R_outer D_p P, A_8 = R D_BB P of the fine operator on the
 * aggregation's (chirally doubled) basis, in FP64, then installed in `cop`
 * with its self-link inverses.  FULL computes all nine points
 * (include/CoarsenedMatrixBaseline.h:575-705); DIRSAVE computes the backward
 * points and the self link and rebuilds the forward ones by
 * gamma5-hermiticity (include/CoarsenedMatrixUpstreamImprovedDirsave.h:835-932,
 * 999-1032).  cop must have the aggregation's coarse lattice and N = 2*nb. */
int mgc_coarsen_operator(mgc_coarse_op* cop, mgc_fine_op* fop, mgc_aggregation* agg, int mode,
                         void* stream);
/* CoarsenOperator(FineGrid, linop, Aggregation) of a caller's operator
 * (include/CoarsenedMatrixRework.h:401-611, in the inner/outer-mask form of
 * include/CoarsenedMatrixBaseline.h:609-675): for batches of up to
 * min(max_nrhs, 16) doubled basis vectors the callback applies every stencil
 * point (FULL: 0..8; DIRSAVE: 0..8 as well, the forward points for their
 * inner sites only, forward links rebuilt by gamma5-hermiticity), and the
 * masked block projections of the results give the links in FP64, installed
 * in `cop` with their self-link inverses.  The caller's operator is used in
 * F64 only. */
int mgc_coarsen_operator_linop(mgc_coarse_op* cop, const mgc_fine_linop* lop,
                               mgc_aggregation* agg, int mode, void* stream);
/* Kernel of mgc_coarsen_operator's build (process-wide): 0 = FP64 on the
 * vector ALU, 1 = FP64 MFMA (v_mfma_f64_16x16x4_f64), -1 = the default (1).
 * Both give the same links to FP64 rounding.  No counterpart in the
 * reference: it serves tests and measurements. */
int mgc_coarsen_set_variant(int variant);
/* CoarsenOperator of a coarse operator (the next level of the hierarchy,
 * src/Benchmark_create_coarse_coarse_operator.cc:250-381): cop1 (F64) on the
 * fine side, agg created with site_elems = cop1's nbasis, result in cop2. */
int mgc_coarsen_coarse_operator(mgc_coarse_op* cop2, mgc_coarse_op* cop1, mgc_aggregation* agg,
                                int mode, void* stream);
/* ConstructRemainingLinks (include/CoarsenedMatrixUpstreamImprovedDirsave.h:999-1032,
 * disp_have = -1) and shiftLinks (include/CoarsenedMatrixKernelsRework.h:141-167,
 * disp_have = +1): A_want(x) = sigma o adj(A_have(x + disp_want mu)), sigma = -1
 * on the off-diagonal chirality blocks. */
int mgc_coarse_op_construct_remaining_links(mgc_coarse_op* op, int disp_have, void* stream);

/* ------------------------------------------------------------------ */
/* multigrid hierarchy, K-cycle / V-cycle, outer solve                 */
/* ------------------------------------------------------------------ */
/* Solver types (MGUtilizedSolverType, include/MultigridParams.h:44-49, same
 * values) and basis-vector types (MGBasisVectorsType, :36-39). */
enum { MGC_SOLVER_MR = 1, MGC_SOLVER_GMRES = 2, MGC_SOLVER_FGMRES = 3, MGC_SOLVER_BICGSTAB = 4 };
enum { MGC_BASIS_NULL_VECTORS = 1, MGC_BASIS_TEST_VECTORS = 2 };

/* Parameters of the hierarchy (MultiGridParams, include/MultigridParams.h:73-139;
 * mgc_mg_params_default fills the reference defaults of :95-133). */
typedef struct mgc_mg_params {
  int nlevels;             /* 2..4 */
  int block[3][4];         /* aggregation block of each coarsening */
  int nb[3];               /* basis vectors of each coarsening (N = 2*nb) */
  int precision;           /* hierarchy precision: MGC_F64 (plain) or MGC_F32 */
  int coarsen_mode;        /* MGC_COARSEN_FULL or MGC_COARSEN_DIRSAVE */
  int setup_iters;         /* test-vector rounds (setupInitialIter) */
  int setup_solve_iters;   /* setup solver iterations per test vector and round (maxIter) */
  double setup_tol;
  double smoother_tol;     /* post-smoother GMRES */
  int smoother_iters;
  int smoother_restart;
  double wrapper_tol;      /* K-cycle FGMRES on the coarse levels */
  int wrapper_iters;
  int wrapper_restart;
  double coarsest_tol;     /* coarsest-level GMRES */
  int coarsest_iters;
  int coarsest_restart;
  unsigned long long seed; /* test-vector generator seed */
  /* setup and solver types (the per-level vectors of the reference, one
   * value for every level here) */
  int setup_solver;        /* setupSolverType: MGC_SOLVER_BICGSTAB (default), _GMRES or _MR */
  int setup_restart;       /* restart length of a GMRES setup solver (default 20) */
  int setup_basis_type;    /* setupBasisVectorsType: MGC_BASIS_TEST_VECTORS (default):
                              solve A psi = v from psi = 0; MGC_BASIS_NULL_VECTORS: relax
                              A psi = 0 from psi = v for setup_solve_iters iterations */
  int setup_pre_orthonormalise;  /* setupPreOrthonormalise (default 1) */
  int setup_post_orthonormalise; /* setupPostOrthonormalise (default 1) */
  int setup_refine_iters;  /* setupRefinementIter (default 0): rounds of v <- K-cycle(v) */
  int setup_recursive_updates;   /* setupRecursiveUpdates (default 1) */
  int smoother_solver;     /* smootherSolverType: MGC_SOLVER_GMRES (default) or _MR */
  int coarsest_solver;     /* coarsestSolverType: MGC_SOLVER_GMRES (default), _BICGSTAB, _MR */
  int graph;               /* K-cycle of mgc_mg_precondition / mgc_mg_solve and of the batched
                              mgc_mg_solve_mrhs (no counterpart in
                              the reference): 0 (default) = host-driven Krylov solvers;
                              1 = device-scalar GMRES with fixed step counts (smoother_iters
                              smoother steps, wrapper_restart FGMRES steps, coarsest_restart
                              coarsest steps) and no host synchronisation; 2 = the same
                              captured once into a HIP graph and replayed (bit-equal to 1).
                              1 and 2 need GMRES smoothers and a created (non-null) stream */
  int setup_nrhs;          /* test vectors solved together in the setup (no counterpart in the
                              reference, whose Generate solves them one at a time):
                              0 or 1 (default) = one at a time; k = 2..64 = in lockstep
                              batches of up to k through the multi-rhs operator (the fine
                              operator up to its max_nrhs; coarse levels of an MGC_F32
                              hierarchy through the FP32 multi-rhs stencil; coarse levels of
                              an MGC_F64 hierarchy stay one at a time).  Per vector the set-up
                              solver's own recurrence and stopping rules, so the basis equals
                              the one-at-a-time basis up to rounding.  Device memory: about
                              8 k level vectors while a level's batches run (BiCGSTAB) */
} mgc_mg_params;

typedef struct mgc_solve_stats {
  int iterations;          /* outer (F)GMRES iterations */
  int converged;
  int precond_applies;
  double true_residual;    /* ||b - D x|| / ||b||, recomputed in FP64 */
  double seconds;          /* wall time of the solve */
} mgc_solve_stats;

typedef struct mgc_mg mgc_mg;

int mgc_mg_params_default(mgc_mg_params* params);
/* MultiGridPreconditioner over a fine Wilson-clover operator
 * (include/MultigridPreconditioner.h:137-163, createMGInstance :665). */
int mgc_mg_create(mgc_mg** mg, mgc_fine_op* fop, const mgc_mg_params* params);
/* The same hierarchy over a caller's fine operator (the reference's
 * MultiGridPreconditioner takes any fine LinearOperatorBase,
 * include/MultigridPreconditioner.h:123-124): the operator must apply in F64
 * (outer FGMRES, true residuals, CoarsenOperator through the point callback);
 * an FP32 hierarchy applies it in F32 where `precisions` allows, else through
 * F64 with conversions.  The callback stays in use until mgc_mg_destroy. */
int mgc_mg_create_linop(mgc_mg** mg, const mgc_fine_linop* lop, const mgc_mg_params* params);
int mgc_mg_destroy(mgc_mg* mg);
/* initialSetup (include/MultigridPreconditioner.h:165-211): test vectors on
 * every level, aggregation + block orthonormalisation, FP64 Galerkin
 * operators (rounded to FP32 for an FP32 hierarchy).  Later calls on the
 * hierarchy must use the same stream. */
int mgc_mg_setup(mgc_mg* mg, void* stream);
/* one preconditioner application, FP64 vectors [site][12]
 * (MultiGridPreconditioner::operator(), include/MultigridPreconditioner.h:253-309) */
int mgc_mg_precondition(mgc_mg* mg, const void* in, void* out, void* stream);
/* NonHermitianVCyclePreconditioner::operator() (include/VCycleRework.h:84-189):
 * pre-smooth, coarse correction, post-smooth, up to max_iter cycles or
 * ||r|| <= tol ||in|| (tol = 0: exactly max_iter cycles). */
int mgc_mg_vcycle(mgc_mg* mg, const void* in, void* out, int max_iter, double tol, void* stream);
/* FGMRES(D) x = b preconditioned by the hierarchy (src/Test_wilsonclover_mg.cc:140-151),
 * FP64 outer iteration, zero initial guess, stop at ||b - D x|| <= tol ||b||. */
int mgc_mg_solve(mgc_mg* mg, const void* b, void* x, double tol, int maxit, int restart,
                 mgc_solve_stats* stats, void* stream);
/* The same solve for nrhs right-hand sides at once (1 <= nrhs <= 64): b and x
 * are [site][rhs][12] FP64.  The outer FGMRES runs the nrhs systems in
 * lockstep (one batched operator application and one batched BLAS pass per
 * step for all of them, each rhs with its own Hessenberg matrix, rotations
 * and stopping point), and the K-cycle applies the whole batch level by
 * level: multi-rhs fine operator and transfers, and the coarse levels
 * through the multi-rhs MFMA stencil (mgc_coarse_op_M_mrhs) with operands in
 * coarse_prec (MGC_F32, MGC_F16 or MGC_BF16; FP32 accumulation).  Needs an
 * FP32 hierarchy whose coarse nbasis is a multiple of 8 up to 64.  stats
 * holds nrhs entries:
 * per rhs its outer iteration count at convergence and its true residual;
 * precond_applies and seconds are those of the whole batch.  Replaces one
 * solve per rhs with the reference's solver on the multi-rhs (5d) grid
 * (src/Benchmark_mrhs_coarsening.cc, include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h). */
int mgc_mg_solve_mrhs(mgc_mg* mg, const void* b, void* x, int nrhs, int coarse_prec, double tol,
                      int maxit, int restart, mgc_solve_stats* stats, void* stream);
/* Coarse solves on up to 8 lattices at once (one nbasis, any volumes): for
 * each lattice k and rhs r, M_k x_k[., r] = b_k[., r] with b_k, x_k complex
 * FP32 [V_k][nrhs][nbasis] on the device (x_k: output, zero initial guess),
 * nrhs <= 64.  Each lattice runs a lockstep restarted GMRES(restart) over its
 * nrhs columns, each column with its own Krylov coefficients and stopping
 * test (relative residual <= tol, at most maxit iterations), on a host thread
 * and stream of its own; the lattices' operator applications meet, and each
 * meeting is a single mgc_coarse_op_M_mrhs_batched launch over the lattices
 * still iterating, in mrhs_prec (F32, or F16 / BF16 with FP32
 * accumulation), so the small lattices ride along with the large one.  The
 * result of every lattice equals that of a batch of that lattice alone, bit
 * for bit.  Each lattice may appear once.  The batched form of the
 * coarsest GMRES of the reference's preconditioner
 * (include/MultigridPreconditioner.h:566-641), which solves one lattice at a
 * time.  stats[k * nrhs + r]: iterations until that column converged,
 * converged, true_residual = ||b - M x|| / ||b|| recomputed with the
 * multi-rhs stencil in mrhs_prec, seconds of the whole call (precond_applies
 * 0). */
int mgc_coarse_solve_batched(mgc_coarse_op* const* ops, int count, const void* const* b,
                             void* const* x, int nrhs, int mrhs_prec, double tol, int maxit,
                             int restart, mgc_solve_stats* stats, void* stream);
/* The same solve with its state kept between calls (operator workspaces, the
 * padded vectors and the Krylov basis of GMRES(restart)), for a caller that
 * solves on the same lattices repeatedly (the coarsest level of a K-cycle):
 * create once, solve with new b / x as often as needed on the stream given to
 * create, destroy.  The operators must outlive the handle. */
typedef struct mgc_coarse_batch mgc_coarse_batch;
int mgc_coarse_batch_create(mgc_coarse_batch** out, mgc_coarse_op* const* ops, int count,
                            int nrhs, int mrhs_prec, int restart, void* stream);
int mgc_coarse_batch_solve(mgc_coarse_batch* cb, const void* const* b, void* const* x, double tol,
                           int maxit, mgc_solve_stats* stats, void* stream);
int mgc_coarse_batch_destroy(mgc_coarse_batch* cb);
/* A prepared batch with a fixed step count (no counterpart in the reference,
 * whose coarsest solves stop at a tolerance): every lattice runs `steps`
 * GMRES steps from x = 0 (1 <= steps < 64), all lattices in lockstep on the
 * stream with one multi-lattice stencil launch per step and no host
 * synchronisation.  graph = 2 captures the whole solve into a HIP graph at
 * the first mgc_coarse_batch_solve (after one eager solve) and replays it
 * after, so the stream must be a created one; graph = 1 runs the same work
 * eagerly (the same results, bit for bit).  mgc_coarse_batch_solve then
 * copies b into the handle's staging vectors and x out of them around the
 * graph, ignores maxit, reports iterations = steps and true_residual
 * ||b - M x|| / ||b|| per (lattice, rhs), and converged = true_residual <= tol. */
int mgc_coarse_batch_create_fixed(mgc_coarse_batch** out, mgc_coarse_op* const* ops, int count,
                                  int nrhs, int mrhs_prec, int steps, int graph, void* stream);
/* the FP64 coarse operator of level >= 1 and the aggregation of level >= 0 (owned by mg) */
int mgc_mg_coarse_op(mgc_mg* mg, int level, mgc_coarse_op** cop);
int mgc_mg_aggregation(mgc_mg* mg, int level, mgc_aggregation** agg);
/* unpreconditioned restarted GMRES on the fine operator, FP64 (baseline solve) */
int mgc_fine_gmres_solve(mgc_fine_op* fop, const void* b, void* x, double tol, int maxit,
                         int restart, mgc_solve_stats* stats, void* stream);

/* ------------------------------------------------------------------ */
/* on-disk formats of the reference's drivers                          */
/* ------------------------------------------------------------------ */
/* The driver parameters besides the multigrid ones (MGTestOtherParams,
 * include/MultigridParams.h:140-176; mgc_test_params_default fills its
 * defaults). */
typedef struct mgc_test_params {
  double outer_tol;        /* outerSolverTol */
  int outer_maxit;         /* outerSolverMaxIter */
  int outer_restart;       /* outerSolverRestartLength */
  double mass_setup;       /* massSetup */
  double mass_solve;       /* massSolve */
  double csw;              /* csw */
  char config[1024];       /* config: gauge file, "foo" = a random field */
  char source_type[32];    /* sourceType: ones, random, gaussian */
  int antiperiodic_t;      /* useAntiPeriodicBC */
} mgc_test_params;
int mgc_test_params_default(mgc_test_params* test);
/* Read the reference drivers' XML parameter file (Grid's XML serialisation
 * of MGTestParams: <grid><Params><mg>..</mg><test>..</test></Params></grid>,
 * read by src/Test_wilsonclover_mg.cc:51-58; a bare <Multigrid> element of
 * MultiGridParams is read too).  Members missing from the file keep the
 * defaults; per-level vectors must hold one value for all levels (this
 * build's parameters are per hierarchy), and the cycling strategy must be
 * KCYCLE with an FGMRES wrapper.  Either output may be NULL. */
int mgc_params_read_xml(const char* path, mgc_mg_params* mg, mgc_test_params* test);
/* Write the same file (test may be NULL: its defaults are written). */
int mgc_params_write_xml(const char* path, const mgc_mg_params* mg, const mgc_test_params* test);

enum { MGC_GAUGE_NERSC = 0, MGC_GAUGE_ILDG = 1 };
/* Read a gauge configuration in the NERSC format (4D_SU3_GAUGE two-row or
 * 4D_SU3_GAUGE_3x3, IEEE32/64 big or little endian, header plaquette, link
 * trace and checksum verified) or in ILDG (LIME records ildg-format and
 * ildg-binary-data, big-endian, precision 32 or 64), as
 * src/Test_wilsonclover_mg.cc:94-101 reads a configuration with IldgReader.
 * dims receives the lattice extents; gauge_host_f64 (host memory, NULL to
 * read only the header) receives [mu][site][3][3] complex double, at least
 * 4 * volume * 9 * 16 bytes of capacity; plaquette (may be NULL) the average
 * plaquette Re tr / 3. */
int mgc_gauge_read(const char* path, int dims[4], void* gauge_host_f64, size_t capacity_bytes,
                   double* plaquette);
/* Write a host gauge field [mu][site][3][3] complex double as a NERSC
 * (4D_SU3_GAUGE_3x3, big endian, with plaquette, link trace and checksum) or
 * ILDG file in 32- or 64-bit precision. */
int mgc_gauge_write(const char* path, int format, const int dims[4], const void* gauge_host_f64,
                    int precision_bits);

#ifdef __cplusplus
}
#endif
#endif /* MGC_H */
