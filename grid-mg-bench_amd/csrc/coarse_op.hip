// Coarse operator (CoarsenedMatrix) for a single right-hand side:
// handle lifetime, link installation, self-link inversion, and the
// link-streaming 9-point stencil behind M, Mdag, Mdir, MdirAll, Mdiag, Dhop,
// DhopOE/EO, Meooe(Dag) and the Mooee family.
//
// Reference semantics: include/CoarsenedMatrixUpstream.h:546-742 (M, Mdag,
// MdirCalc, MdirAll, Mdir, Mdiag), include/CoarsenedMatrixRework.h:109-399
// (split-layout M, Dhop*, Meooe*, Mooee*), include/CoarsenedMatrixRework.h:640-679
// (InvertSelfStencilPoint, FillHalfCbs), include/Layout_converters.h:33-89.
#include "coarse_op.hpp"

#include <vector>
#include <cstring>

namespace mgc {

// ---------------------------------------------------------------------------
// stencil kernel
//
// One workgroup per output site.  Thread t owns VR consecutive rows
// (rg = t % RT) of the columns c = s, s+S, ... (s = t / RT) of every point's
// N x N matrix.  Because links are stored column-major per (site, point), a
// wavefront's 16-byte loads cover whole contiguous column runs: a site's
// 9*N*N link entries are streamed exactly once, fully coalesced.  The 9
// neighbour vectors (N complex each, with the Galerkin G5C sign folded in for
// the dagger) are staged once in LDS and read as broadcasts; the S partial
// row sums are combined through LDS at the end.
// ---------------------------------------------------------------------------
template <typename R, int N>
struct StencilCfg {
  static constexpr int VR = 16 / (2 * (int)sizeof(R));  // complex per 16-byte load
  static constexpr int RT = N / VR;                      // threads along a column
  static constexpr int S = (256 / RT) > 0 ? (256 / RT) : 1;
  static constexpr int BLOCK = RT * S;
  static constexpr int CI = (N + S - 1) / S;             // columns per thread per point
};

template <typename R> struct Load16;
template <> struct Load16<double> { typedef double __attribute__((ext_vector_type(2))) T; };
template <> struct Load16<float> { typedef float __attribute__((ext_vector_type(4))) T; };

template <typename R, int N>
__global__ __launch_bounds__((StencilCfg<R, N>::BLOCK)) void k_stencil(
    const cplx<R>* __restrict__ links, size_t site_stride, size_t point_stride,
    const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
    const int* __restrict__ out_site, const int* __restrict__ nbr, unsigned mask, int dag) {
  using C = StencilCfg<R, N>;
  typedef typename Load16<R>::T V16;
  __shared__ cplx<R> s_in[9][N];
  __shared__ cplx<R> s_red[C::S][N];

  const int o = blockIdx.x;
  const int site = out_site ? out_site[o] : o;
  const int tid = threadIdx.x;

  for (int idx = tid; idx < 9 * N; idx += C::BLOCK) {
    const int p = idx / N;
    const int c = idx - p * N;
    cplx<R> v = {R(0), R(0)};
    if ((mask >> p) & 1u) {
      const int j = (p == 8) ? o : nbr[(size_t)site * 8 + p];
      v = in[(size_t)j * N + c];
      if (dag && c >= N / 2) { v.re = -v.re; v.im = -v.im; }
    }
    s_in[p][c] = v;
  }
  __syncthreads();

  const int rg = tid % C::RT;
  const int s = tid / C::RT;
  cplx<R> acc[C::VR];
#pragma unroll
  for (int v = 0; v < C::VR; ++v) acc[v] = {R(0), R(0)};

  const cplx<R>* A = links + (size_t)site * site_stride;
#pragma unroll 1
  for (int p = 0; p < 9; ++p) {
    if (!((mask >> p) & 1u)) continue;
    const cplx<R>* Ap = A + (size_t)p * point_stride;
    V16 buf[C::CI];
#pragma unroll
    for (int k = 0; k < C::CI; ++k) {
      const int c = s + k * C::S;
      if ((N % C::S) == 0 || c < N)
        buf[k] = *reinterpret_cast<const V16*>(Ap + (size_t)c * N + rg * C::VR);
    }
#pragma unroll
    for (int k = 0; k < C::CI; ++k) {
      const int c = s + k * C::S;
      if ((N % C::S) == 0 || c < N) {
        const cplx<R> x = s_in[p][c];
#pragma unroll
        for (int v = 0; v < C::VR; ++v) {
          cplx<R> a = {buf[k][2 * v], buf[k][2 * v + 1]};
          cmac(acc[v], a, x);
        }
      }
    }
  }
#pragma unroll
  for (int v = 0; v < C::VR; ++v) s_red[s][rg * C::VR + v] = acc[v];
  __syncthreads();
  for (int r = tid; r < N; r += C::BLOCK) {
    cplx<R> t = s_red[0][r];
#pragma unroll
    for (int q = 1; q < C::S; ++q) t = cadd(t, s_red[q][r]);
    if (dag && r >= N / 2) { t.re = -t.re; t.im = -t.im; }
    out[(size_t)o * N + r] = t;
  }
}

template <typename R, int N>
static void launch_stencil_t(const void* links, size_t site_stride, size_t point_stride,
                             const void* in, void* out, const int* out_site, int n_out,
                             const int* nbr, unsigned mask, int dag, hipStream_t s) {
  using C = StencilCfg<R, N>;
  k_stencil<R, N><<<n_out, C::BLOCK, 0, s>>>((const cplx<R>*)links, site_stride, point_stride,
                                             (const cplx<R>*)in, (cplx<R>*)out, out_site, nbr,
                                             mask, dag);
  MGC_LAUNCH_CHECK();
}

#define MGC_FOR_EACH_N(X)                                                                  \
  X(2) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) \
  X(34) X(36) X(38) X(40) X(42) X(44) X(46) X(48) X(50) X(52) X(54) X(56) X(58) X(60) X(62)   \
  X(64)

template <typename R>
static void launch_stencil_r(int N, const void* links, size_t ss, size_t ps, const void* in,
                             void* out, const int* out_site, int n_out, const int* nbr,
                             unsigned mask, int dag, hipStream_t s) {
  switch (N) {
#define CASE(n)                                                                         \
  case n:                                                                               \
    launch_stencil_t<R, n>(links, ss, ps, in, out, out_site, n_out, nbr, mask, dag, s); \
    return;
    MGC_FOR_EACH_N(CASE)
#undef CASE
  }
  throw Error(MGC_ERR_SHAPE, "nbasis not supported by the single-rhs stencil");
}

void stencil_launch(const mgc_coarse_op* op, const void* links, size_t site_stride,
                    size_t point_stride, const void* in, void* out, const int* out_site,
                    int n_out, const int* nbr, unsigned mask, int dag, hipStream_t s) {
  if (n_out <= 0) return;
  if (op->prec == MGC_F64)
    launch_stencil_r<double>(op->N, links, site_stride, point_stride, in, out, out_site, n_out,
                             nbr, mask, dag, s);
  else if (op->prec == MGC_F32)
    launch_stencil_r<float>(op->N, links, site_stride, point_stride, in, out, out_site, n_out,
                            nbr, mask, dag, s);
  else
    throw Error(MGC_ERR_PRECISION, "single-rhs applies need an F64 or F32 operator");
}

// ---------------------------------------------------------------------------
// link installation: external FLAT/SPLIT complex-double links -> internal
// column-major storage in the operator precision, plus an FP64 row-major
// copy of the self links for the inversion
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t ext_link_index(int layout, int V, int N, int p, int x, int r,
                                                 int c) {
  if (layout == MGC_LAYOUT_FLAT) return (((size_t)p * V + x) * N + r) * N + c;
  const int nb = N / 2;
  const int s1 = r / nb, c1 = r - s1 * nb, s2 = c / nb, c2 = c - s2 * nb;
  return (((((size_t)p * V + x) * 2 + s1) * 2 + s2) * nb + c1) * nb + c2;
}

template <typename R>
__global__ void k_import_links(const cplx<double>* __restrict__ src, int layout, int V, int N,
                               cplx<R>* __restrict__ links, cplx<R>* __restrict__ self_dag,
                               cplx<double>* __restrict__ self_rm) {
  const size_t total = (size_t)V * 9 * N * N;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    // destination order [x][p][c][r]
    const int r = (int)(i % N);
    size_t t = i / N;
    const int c = (int)(t % N);
    t /= N;
    const int p = (int)(t % 9);
    const int x = (int)(t / 9);
    const cplx<double> a = src[ext_link_index(layout, V, N, p, x, r, c)];
    links[i] = {(R)a.re, (R)a.im};
    if (p == 8) {
      // adj(A8) column-major: element (row r, col c) = conj(A8(c, r))
      const cplx<double> b = src[ext_link_index(layout, V, N, 8, x, c, r)];
      self_dag[((size_t)x * N + c) * N + r] = {(R)b.re, (R)(-b.im)};
      self_rm[((size_t)x * N + r) * N + c] = a;
    }
  }
}

// The same import, one workgroup per (site, point) matrix through LDS: the
// source is read along its rows (c fastest, contiguous in both layouts) and
// the column-major destination written along its columns (r fastest), so
// both sides are coalesced.  For N <= 64.
template <typename R>
__global__ __launch_bounds__(256) void k_import_links_tiled(
    const cplx<double>* __restrict__ src, int layout, int V, int N, cplx<R>* __restrict__ links,
    cplx<R>* __restrict__ self_dag, cplx<double>* __restrict__ self_rm) {
  __shared__ cplx<double> sA[64][65];
  const int x = blockIdx.x / 9, p = blockIdx.x - x * 9;
  const int t = threadIdx.x;
  for (int idx = t; idx < N * N; idx += 256) {
    const int r = idx / N, c = idx - r * N;
    sA[r][c] = src[ext_link_index(layout, V, N, p, x, r, c)];
  }
  __syncthreads();
  cplx<R>* dst = links + ((size_t)x * 9 + p) * N * N;  // [c][r]
  for (int idx = t; idx < N * N; idx += 256) {
    const int c = idx / N, r = idx - c * N;
    const cplx<double> a = sA[r][c];
    dst[idx] = {(R)a.re, (R)a.im};
  }
  if (p == 8) {
    // adj(A8) column-major: element (row r, col c) = conj(A8(c, r)), at [c][r]
    for (int idx = t; idx < N * N; idx += 256) {
      const int c = idx / N, r = idx - c * N;
      const cplx<double> b = sA[c][r];
      self_dag[(size_t)x * N * N + idx] = {(R)b.re, (R)(-b.im)};
      self_rm[(size_t)x * N * N + idx] = sA[c][r];  // row-major: idx = row c, column r
    }
  }
}

template <typename R>
__global__ void k_export_links(const cplx<R>* __restrict__ links, int layout, int V, int N,
                               cplx<double>* __restrict__ dst) {
  const size_t total = (size_t)V * 9 * N * N;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    const int r = (int)(i % N);
    size_t t = i / N;
    const int c = (int)(t % N);
    t /= N;
    const int p = (int)(t % 9);
    const int x = (int)(t / 9);
    const cplx<R> a = links[i];
    dst[ext_link_index(layout, V, N, p, x, r, c)] = {(double)a.re, (double)a.im};
  }
}

// ---------------------------------------------------------------------------
// batched in-place Gauss-Jordan inversion with partial (row) pivoting, FP64,
// one workgroup per site, the N x N matrix resident in LDS.  Replaces the
// host-side Eigen inverse of include/CoarsenedMatrixRework.h:640-668 and
// include/CoarsenedMatrixBaseline.h:735-768.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gauss_jordan(cplx<double>* __restrict__ mats, int N,
                                                      int* __restrict__ singular) {
  extern __shared__ cplx<double> sm[];  // N*N matrix followed by N column buffer
  cplx<double>* M = sm;
  cplx<double>* col = sm + N * N;
  __shared__ int s_perm[128];
  __shared__ int s_piv;
  __shared__ int s_bad;
  const int tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * N * N;
  for (int i = tid; i < N * N; i += blockDim.x) M[i] = mats[base + i];
  if (tid == 0) s_bad = 0;
  __syncthreads();

  for (int k = 0; k < N; ++k) {
    if (tid < 64) {  // wavefront 0 finds the pivot row
      double best = -1.0;
      int bi = k;
      for (int i = k + tid; i < N; i += 64) {
        const cplx<double> v = M[i * N + k];
        const double m = v.re * v.re + v.im * v.im;
        if (m > best) { best = m; bi = i; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      if (tid == 0) {
        s_piv = bi;
        s_perm[k] = bi;
        if (!(best > 0.0)) s_bad = 1;
      }
    }
    __syncthreads();
    const int pr = s_piv;
    if (pr != k)
      for (int c = tid; c < N; c += blockDim.x) {
        const cplx<double> t = M[k * N + c];
        M[k * N + c] = M[pr * N + c];
        M[pr * N + c] = t;
      }
    __syncthreads();
    const cplx<double> piv = M[k * N + k];
    const double den = piv.re * piv.re + piv.im * piv.im;
    const cplx<double> pinv = den > 0.0 ? cplx<double>{piv.re / den, -piv.im / den}
                                        : cplx<double>{0.0, 0.0};
    for (int i = tid; i < N; i += blockDim.x) col[i] = M[i * N + k];
    __syncthreads();
    for (int c = tid; c < N; c += blockDim.x)
      M[k * N + c] = (c == k) ? pinv : cmul(M[k * N + c], pinv);
    __syncthreads();
    for (int idx = tid; idx < N * N; idx += blockDim.x) {
      const int i = idx / N, c = idx - i * N;
      if (i == k) continue;
      const cplx<double> f = col[i];
      if (c == k) {
        const cplx<double> t = cmul(f, pinv);
        M[idx] = {-t.re, -t.im};
      } else {
        const cplx<double> t = cmul(f, M[k * N + c]);
        M[idx] = {M[idx].re - t.re, M[idx].im - t.im};
      }
    }
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = N - 1; k >= 0; --k) {
    const int pr = s_perm[k];
    if (pr != k)
      for (int i = tid; i < N; i += blockDim.x) {
        const cplx<double> t = M[i * N + k];
        M[i * N + k] = M[i * N + pr];
        M[i * N + pr] = t;
      }
    __syncthreads();
  }
  for (int i = tid; i < N * N; i += blockDim.x) mats[base + i] = M[i];
  if (tid == 0 && s_bad) atomicOr(singular, 1);
}

// The same Gauss-Jordan inversion for N <= 64 with the matrix in registers:
// the 256 threads hold it as a 16 x 16 grid of 4 x 4 tiles (thread t: rows
// 4 (t / 16) .., columns 4 (t % 16) ..), zero-padded to 64 x 64 with an
// identity block.  Per step k the pivot column and the normalised pivot row
// pass through LDS; every element gets exactly the operations of
// k_gauss_jordan (pivot = first largest |M(i,k)|^2, row k times the pivot's
// inverse, M(i,c) -= M(i,k) M(k,c)), so both give the same bits.  The row
// interchanges are undone as one column permutation at the store.
__global__ __launch_bounds__(256) void k_gauss_jordan_reg(cplx<double>* __restrict__ mats, int N,
                                                          int* __restrict__ singular) {
  __shared__ cplx<double> s_col[2][64];  // pivot column before the step (double-buffered)
  __shared__ cplx<double> s_row[64];     // normalised pivot row
  __shared__ cplx<double> s_rk[64], s_rp[64];  // rows k and pr for the interchange
  __shared__ int s_perm[64];
  __shared__ int s_pos[64];
  __shared__ int s_piv, s_bad;
  const int t = threadIdx.x;
  const int ti = t >> 4, tj = t & 15;
  const size_t base = (size_t)blockIdx.x * N * N;
  cplx<double> m[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 4 * ti + r, j = 4 * tj + c;
      m[r][c] = (i < N && j < N) ? mats[base + (size_t)i * N + j]
                                 : cplx<double>{i == j ? 1.0 : 0.0, 0.0};
    }
  if (t == 0) s_bad = 0;
  for (int k = 0; k < N; ++k) {
    const int kb = k >> 2, kr = k & 3;
    cplx<double>* col = s_col[k & 1];
    // a. the pivot column to LDS
    if (tj == kb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cplx<double> v = m[r][0];
#pragma unroll
        for (int c = 1; c < 4; ++c)
          if (c == kr) v = m[r][c];
        col[4 * ti + r] = v;
      }
    __syncthreads();
    // b. wavefront 0 finds the pivot row (first largest |v|^2 at or below k)
    if (t < 64) {
      double best = -1.0;
      int bi = k;
      if (t >= k && t < N) {
        const cplx<double> v = col[t];
        best = v.re * v.re + v.im * v.im;
        bi = t;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (t == 0) {
        s_piv = bi;
        s_perm[k] = bi;
        if (!(best > 0.0)) s_bad = 1;
        const cplx<double> a = col[k];  // the pivot column follows the interchange
        col[k] = col[bi];
        col[bi] = a;
      }
    }
    __syncthreads();
    const int pr = s_piv;
    // c. interchange rows k and pr
    if (pr != k) {
      const int pb = pr >> 2, prr = pr & 3;
      if (ti == kb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (r == kr)
#pragma unroll
            for (int c = 0; c < 4; ++c) s_rk[4 * tj + c] = m[r][c];
      if (ti == pb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (r == prr)
#pragma unroll
            for (int c = 0; c < 4; ++c) s_rp[4 * tj + c] = m[r][c];
      __syncthreads();
      if (ti == kb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (r == kr)
#pragma unroll
            for (int c = 0; c < 4; ++c) m[r][c] = s_rp[4 * tj + c];
      if (ti == pb)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (r == prr)
#pragma unroll
            for (int c = 0; c < 4; ++c) m[r][c] = s_rk[4 * tj + c];
    }
    const cplx<double> piv = col[k];
    const double den = piv.re * piv.re + piv.im * piv.im;
    const cplx<double> pinv = den > 0.0 ? cplx<double>{piv.re / den, -piv.im / den}
                                        : cplx<double>{0.0, 0.0};
    // d. normalise row k, publish it
    if (ti == kb)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r == kr)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const int j = 4 * tj + c;
            m[r][c] = (j == k) ? pinv : cmul(m[r][c], pinv);
            s_row[j] = m[r][c];
          }
    __syncthreads();
    // e. eliminate column k from every other row
    cplx<double> f[4], rk[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) f[r] = col[4 * ti + r];
#pragma unroll
    for (int c = 0; c < 4; ++c) rk[c] = s_row[4 * tj + c];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (4 * ti + r == k) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (4 * tj + c == k) {
          const cplx<double> q = cmul(f[r], pinv);
          m[r][c] = {-q.re, -q.im};
        } else {
          const cplx<double> q = cmul(f[r], rk[c]);
          m[r][c] = {m[r][c].re - q.re, m[r][c].im - q.im};
        }
      }
    }
  }
  // the row interchanges, undone last first as column interchanges, make one
  // permutation: output column j = current column s_pos[j]
  __syncthreads();
  if (t == 0) {
    for (int j = 0; j < N; ++j) s_pos[j] = j;
    for (int k = N - 1; k >= 0; --k) {
      const int pr = s_perm[k];
      const int a = s_pos[k];
      s_pos[k] = s_pos[pr];
      s_pos[pr] = a;
    }
    // s_perm <- inverse: output column of current column c
    for (int j = 0; j < N; ++j) s_perm[s_pos[j]] = j;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 4 * ti + r, j = 4 * tj + c;
      if (i < N && j < N) mats[base + (size_t)i * N + s_perm[j]] = m[r][c];
    }
  if (t == 0 && s_bad) atomicOr(singular, 1);
}

// row-major FP64 -> column-major (prec) plus its adjoint
template <typename R>
__global__ void k_store_inverse(const cplx<double>* __restrict__ rm, int V, int N,
                                cplx<R>* __restrict__ inv, cplx<R>* __restrict__ invdag) {
  const size_t total = (size_t)V * N * N;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    const int r = (int)(i % N);
    size_t t = i / N;
    const int c = (int)(t % N);
    const size_t x = t / N;
    const cplx<double> a = rm[(x * N + r) * N + c];   // inv(r, c)
    const cplx<double> b = rm[(x * N + c) * N + r];   // inv(c, r)
    inv[i] = {(R)a.re, (R)a.im};
    invdag[i] = {(R)b.re, (R)(-b.im)};
  }
}

template <typename R>
__global__ void k_export_inverse(const cplx<R>* __restrict__ inv, int V, int N,
                                 cplx<double>* __restrict__ rm) {
  const size_t total = (size_t)V * N * N;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    const int r = (int)(i % N);
    size_t t = i / N;
    const int c = (int)(t % N);
    const size_t x = t / N;
    rm[(x * N + r) * N + c] = {(double)inv[i].re, (double)inv[i].im};
  }
}

void invert_self_rowmajor(cplx<double>* rm, int V, int N, hipStream_t s) {
  MGC_CHECK(N <= 96, MGC_ERR_SHAPE, "self-link inversion supports nbasis <= 96");
  int* d_flag = nullptr;
  MGC_HIP(hipMalloc(&d_flag, sizeof(int)));
  fill_zero(d_flag, sizeof(int), s);
  if (N <= 64) {
    k_gauss_jordan_reg<<<V, 256, 0, s>>>(rm, N, d_flag);
  } else {
    const size_t shmem = (size_t)(N * N + N) * sizeof(cplx<double>);
    if (shmem > 64 * 1024)
      MGC_HIP(hipFuncSetAttribute((const void*)k_gauss_jordan,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    k_gauss_jordan<<<V, 256, shmem, s>>>(rm, N, d_flag);
  }
  MGC_LAUNCH_CHECK();
  int flag = 0;
  MGC_HIP(hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
  MGC_HIP(hipStreamSynchronize(s));
  MGC_HIP(hipFree(d_flag));
  MGC_CHECK(!flag, MGC_ERR_SINGULAR, "a self link is singular");
}

static void free_op(mgc_coarse_op* op) {
  if (!op) return;
  void* bufs[] = {op->links,    op->self_dag, op->self_inv,   op->self_invdag, op->nbr_full,
                  op->nbr_half, op->cb_sites, op->mrhs_links, op->mrhs_scales,
                  op->q_in,     op->q_scale};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete op;
}

static void check_op(const mgc_coarse_op* op) {
  MGC_CHECK(op, MGC_ERR_ARG, "null operator handle");
  MGC_CHECK(op->links_set, MGC_ERR_STATE, "links not set");
}

static bool full_dims_even(const mgc_coarse_op* op) { return (op->lat.L[0] % 2) == 0; }

}  // namespace mgc

using namespace mgc;

extern "C" {

int mgc_coarse_op_create(mgc_coarse_op** out, const int dims[4], int nbasis, int prec,
                         int hermitian) {
  mgc_coarse_op* op = nullptr;
  int st = guarded([&] {
    MGC_CHECK(out && dims, MGC_ERR_ARG, "null argument");
    for (int d = 0; d < 4; ++d) MGC_CHECK(dims[d] >= 1, MGC_ERR_SHAPE, "lattice extent < 1");
    MGC_CHECK(nbasis >= 2 && nbasis % 2 == 0 && nbasis <= 128, MGC_ERR_SHAPE,
              "nbasis must be even and in 2..128");
    MGC_CHECK(prec >= MGC_F64 && prec <= MGC_BF16, MGC_ERR_PRECISION, "unknown precision");
    op = new mgc_coarse_op();
    for (int d = 0; d < 4; ++d) op->lat.L[d] = dims[d];
    op->V = op->lat.volume();
    op->N = nbasis;
    op->prec = prec;
    op->hermitian = hermitian ? 1 : 0;
    const int V = op->V;
    // neighbour tables on the host
    std::vector<int> nf((size_t)V * 8), nh((size_t)V * 8), cbs(V + 1, 0);
    for (int x = 0; x < V; ++x)
      for (int p = 0; p < 8; ++p) {
        const int y = op->lat.shift(x, point_dir(p), point_disp(p));
        nf[(size_t)x * 8 + p] = y;
        nh[(size_t)x * 8 + p] = y >> 1;
      }
    if (dims[0] % 2 == 0) {
      int cnt[2] = {0, 0};
      for (int x = 0; x < V; ++x) {
        int c[4];
        op->lat.coords(x, c);
        const int par = (c[0] + c[1] + c[2] + c[3]) & 1;
        cbs[(size_t)par * (V / 2) + cnt[par]++] = x;
      }
    }
    MGC_HIP(hipMalloc(&op->nbr_full, nf.size() * sizeof(int)));
    MGC_HIP(hipMalloc(&op->nbr_half, nh.size() * sizeof(int)));
    MGC_HIP(hipMalloc(&op->cb_sites, (size_t)(V + 1) * sizeof(int)));
    MGC_HIP(hipMemcpy(op->nbr_full, nf.data(), nf.size() * sizeof(int), hipMemcpyHostToDevice));
    MGC_HIP(hipMemcpy(op->nbr_half, nh.data(), nh.size() * sizeof(int), hipMemcpyHostToDevice));
    MGC_HIP(hipMemcpy(op->cb_sites, cbs.data(), (size_t)(V + 1) * sizeof(int),
                      hipMemcpyHostToDevice));
    const size_t eb = op->elem_bytes();
    MGC_HIP(hipMalloc(&op->links, (size_t)V * 9 * nbasis * nbasis * eb));
    MGC_HIP(hipMalloc(&op->self_dag, (size_t)V * nbasis * nbasis * eb));
    *out = op;
  });
  if (st != MGC_OK) free_op(op);
  return st;
}

int mgc_coarse_op_destroy(mgc_coarse_op* op) {
  return guarded([&] { free_op(op); });
}

int mgc_coarse_op_info(const mgc_coarse_op* op, int dims[4], int* nbasis, int* prec,
                       int* volume) {
  return guarded([&] {
    MGC_CHECK(op, MGC_ERR_ARG, "null operator handle");
    if (dims)
      for (int d = 0; d < 4; ++d) dims[d] = op->lat.L[d];
    if (nbasis) *nbasis = op->N;
    if (prec) *prec = op->prec;
    if (volume) *volume = op->V;
  });
}

int mgc_coarse_op_set_links(mgc_coarse_op* op, int layout, const void* links_f64, int invert,
                            void* stream) {
  return guarded([&] {
    MGC_CHECK(op && links_f64, MGC_ERR_ARG, "null argument");
    MGC_CHECK(layout == MGC_LAYOUT_FLAT || layout == MGC_LAYOUT_SPLIT, MGC_ERR_ARG,
              "unknown link layout");
    hipStream_t s = as_stream(stream);
    const int V = op->V, N = op->N;
    cplx<double>* rm = nullptr;
    MGC_HIP(hipMalloc(&rm, (size_t)V * N * N * sizeof(cplx<double>)));
    const size_t total = (size_t)V * 9 * N * N;
    const unsigned g = grid_for(total, 256, 16384);
    const auto* lf = (const cplx<double>*)links_f64;
    if (op->prec == MGC_F64 && N <= 64)
      k_import_links_tiled<double><<<V * 9, 256, 0, s>>>(lf, layout, V, N, (cplx<double>*)op->links,
                                                         (cplx<double>*)op->self_dag, rm);
    else if (op->prec == MGC_F32 && N <= 64)
      k_import_links_tiled<float><<<V * 9, 256, 0, s>>>(lf, layout, V, N, (cplx<float>*)op->links,
                                                        (cplx<float>*)op->self_dag, rm);
    else if (op->prec == MGC_F64)
      k_import_links<double><<<g, 256, 0, s>>>(lf, layout, V, N, (cplx<double>*)op->links,
                                               (cplx<double>*)op->self_dag, rm);
    else if (op->prec == MGC_F32)
      k_import_links<float><<<g, 256, 0, s>>>(lf, layout, V, N, (cplx<float>*)op->links,
                                              (cplx<float>*)op->self_dag, rm);
    else
      throw Error(MGC_ERR_PRECISION,
                  "half-precision operators take their links through an F32/F64 operator");
    MGC_LAUNCH_CHECK();
    op->links_set = true;
    op->inv_set = false;
    if (op->mrhs_links) {
      MGC_HIP(hipFree(op->mrhs_links));
      op->mrhs_links = nullptr;
      op->mrhs_prec = -1;
    }
    if (invert) {
      try {
        invert_self_rowmajor(rm, V, N, s);
      } catch (...) {
        (void)hipFree(rm);
        throw;
      }
      const size_t eb = op->elem_bytes();
      if (!op->self_inv) MGC_HIP(hipMalloc(&op->self_inv, (size_t)V * N * N * eb));
      if (!op->self_invdag) MGC_HIP(hipMalloc(&op->self_invdag, (size_t)V * N * N * eb));
      const unsigned g2 = grid_for((size_t)V * N * N, 256, 16384);
      if (op->prec == MGC_F64)
        k_store_inverse<double><<<g2, 256, 0, s>>>(rm, V, N, (cplx<double>*)op->self_inv,
                                                   (cplx<double>*)op->self_invdag);
      else
        k_store_inverse<float><<<g2, 256, 0, s>>>(rm, V, N, (cplx<float>*)op->self_inv,
                                                  (cplx<float>*)op->self_invdag);
      MGC_LAUNCH_CHECK();
      op->inv_set = true;
    }
    MGC_HIP(hipStreamSynchronize(s));
    MGC_HIP(hipFree(rm));
  });
}

int mgc_coarse_op_get_links(const mgc_coarse_op* op, int layout, void* links_f64,
                            void* stream) {
  return guarded([&] {
    check_op(op);
    MGC_CHECK(links_f64, MGC_ERR_ARG, "null output");
    MGC_CHECK(layout == MGC_LAYOUT_FLAT || layout == MGC_LAYOUT_SPLIT, MGC_ERR_ARG,
              "unknown link layout");
    hipStream_t s = as_stream(stream);
    const size_t total = (size_t)op->V * 9 * op->N * op->N;
    const unsigned g = grid_for(total, 256, 16384);
    if (op->prec == MGC_F64)
      k_export_links<double><<<g, 256, 0, s>>>((const cplx<double>*)op->links, layout, op->V,
                                               op->N, (cplx<double>*)links_f64);
    else if (op->prec == MGC_F32)
      k_export_links<float><<<g, 256, 0, s>>>((const cplx<float>*)op->links, layout, op->V,
                                              op->N, (cplx<double>*)links_f64);
    else
      throw Error(MGC_ERR_PRECISION, "no single-rhs links in a half-precision operator");
    MGC_LAUNCH_CHECK();
  });
}

int mgc_coarse_op_get_self_inverse(const mgc_coarse_op* op, void* out_f64, void* stream) {
  return guarded([&] {
    check_op(op);
    MGC_CHECK(op->inv_set, MGC_ERR_STATE, "self-link inverse not computed");
    MGC_CHECK(out_f64, MGC_ERR_ARG, "null output");
    hipStream_t s = as_stream(stream);
    const unsigned g = grid_for((size_t)op->V * op->N * op->N, 256, 16384);
    if (op->prec == MGC_F64)
      k_export_inverse<double><<<g, 256, 0, s>>>((const cplx<double>*)op->self_inv, op->V,
                                                 op->N, (cplx<double>*)out_f64);
    else
      k_export_inverse<float><<<g, 256, 0, s>>>((const cplx<float>*)op->self_inv, op->V,
                                                op->N, (cplx<double>*)out_f64);
    MGC_LAUNCH_CHECK();
  });
}

// ----------------------------- applies -------------------------------------
static void apply_full(mgc_coarse_op* op, const void* in, void* out, unsigned mask, int dag,
                       hipStream_t s) {
  check_op(op);
  MGC_CHECK(in && out, MGC_ERR_ARG, "null vector");
  MGC_CHECK(in != out, MGC_ERR_ARG, "in and out must not alias");
  const size_t NN = (size_t)op->N * op->N;
  stencil_launch(op, op->links, 9 * NN, NN, in, out, nullptr, op->V, op->nbr_full, mask, dag, s);
}

int mgc_coarse_op_M(mgc_coarse_op* op, const void* in, void* out, void* stream) {
  return guarded([&] { apply_full(op, in, out, 0x1FFu, 0, as_stream(stream)); });
}
int mgc_coarse_op_Mdag(mgc_coarse_op* op, const void* in, void* out, void* stream) {
  return guarded([&] {
    MGC_CHECK(op, MGC_ERR_ARG, "null operator handle");
    apply_full(op, in, out, 0x1FFu, op->hermitian ? 0 : 1, as_stream(stream));
  });
}
int mgc_coarse_op_Mdiag(mgc_coarse_op* op, const void* in, void* out, void* stream) {
  return guarded([&] { apply_full(op, in, out, 1u << 8, 0, as_stream(stream)); });
}
int mgc_coarse_op_Mdir(mgc_coarse_op* op, const void* in, void* out, int dir, int disp,
                       void* stream) {
  return guarded([&] {
    const int p = point_from_dir_disp(dir, disp);
    apply_full(op, in, out, 1u << p, 0, as_stream(stream));
  });
}
int mgc_coarse_op_MdirAll(mgc_coarse_op* op, const void* in, void* out8, void* stream) {
  return guarded([&] {
    check_op(op);
    MGC_CHECK(out8, MGC_ERR_ARG, "null output");
    for (int p = 0; p < 8; ++p)
      apply_full(op, in, (char*)out8 + (size_t)p * op->vec_bytes(), 1u << p, 0,
                 as_stream(stream));
  });
}
int mgc_coarse_op_Dhop(mgc_coarse_op* op, const void* in, void* out, int dag, void* stream) {
  return guarded([&] {
    MGC_CHECK(op, MGC_ERR_ARG, "null operator handle");
    apply_full(op, in, out, 0xFFu, dag ? 1 : 0, as_stream(stream));
  });
}

static void dhop_cb(mgc_coarse_op* op, const void* in, void* out, int out_cb, int dag,
                    hipStream_t s) {
  check_op(op);
  MGC_CHECK(in && out && in != out, MGC_ERR_ARG, "null or aliased vector");
  MGC_CHECK(full_dims_even(op), MGC_ERR_SHAPE, "checkerboarding needs an even x extent");
  const size_t NN = (size_t)op->N * op->N;
  const int H = op->V / 2;
  stencil_launch(op, op->links, 9 * NN, NN, in, out, op->cb_sites + (size_t)out_cb * H, H,
                 op->nbr_half, 0xFFu, dag, s);
}

int mgc_coarse_op_DhopOE(mgc_coarse_op* op, const void* in_even, void* out_odd, int dag,
                         void* stream) {
  return guarded([&] { dhop_cb(op, in_even, out_odd, MGC_ODD, dag ? 1 : 0, as_stream(stream)); });
}
int mgc_coarse_op_DhopEO(mgc_coarse_op* op, const void* in_odd, void* out_even, int dag,
                         void* stream) {
  return guarded([&] { dhop_cb(op, in_odd, out_even, MGC_EVEN, dag ? 1 : 0, as_stream(stream)); });
}
int mgc_coarse_op_Meooe(mgc_coarse_op* op, const void* in, void* out, int in_cb, void* stream) {
  return guarded([&] {
    MGC_CHECK(in_cb == MGC_EVEN || in_cb == MGC_ODD, MGC_ERR_ARG, "Meooe needs a half field");
    dhop_cb(op, in, out, 1 - in_cb, 0, as_stream(stream));
  });
}
int mgc_coarse_op_MeooeDag(mgc_coarse_op* op, const void* in, void* out, int in_cb,
                           void* stream) {
  return guarded([&] {
    MGC_CHECK(op, MGC_ERR_ARG, "null operator handle");
    MGC_CHECK(in_cb == MGC_EVEN || in_cb == MGC_ODD, MGC_ERR_ARG, "MeooeDag needs a half field");
    dhop_cb(op, in, out, 1 - in_cb, 1, as_stream(stream));
  });
}

// which: 0 = A8, 1 = A8^-1, 2 = A8^dag, 3 = (A8^-1)^dag
static void mooee(mgc_coarse_op* op, const void* in, void* out, int cb, int which,
                  hipStream_t s) {
  check_op(op);
  MGC_CHECK(in && out && in != out, MGC_ERR_ARG, "null or aliased vector");
  MGC_CHECK(cb == MGC_FULL || cb == MGC_EVEN || cb == MGC_ODD, MGC_ERR_ARG, "bad parity");
  const size_t NN = (size_t)op->N * op->N;
  const void* L = nullptr;
  size_t ss = NN;
  switch (which) {
    case 0:
      // the self point inside the 9-point block: offset by point 8, stride 9*NN
      L = (const char*)op->links + 8 * NN * op->elem_bytes();
      ss = 9 * NN;
      break;
    case 1: L = op->self_inv; break;
    case 2: L = op->self_dag; break;
    case 3: L = op->self_invdag; break;
  }
  if (which == 1 || which == 3)
    MGC_CHECK(op->inv_set, MGC_ERR_STATE, "self-link inverse not computed");
  if (cb == MGC_FULL) {
    stencil_launch(op, L, ss, 0, in, out, nullptr, op->V, op->nbr_full, 1u << 8, 0, s);
  } else {
    MGC_CHECK(full_dims_even(op), MGC_ERR_SHAPE, "checkerboarding needs an even x extent");
    const int H = op->V / 2;
    stencil_launch(op, L, ss, 0, in, out, op->cb_sites + (size_t)cb * H, H, op->nbr_half,
                   1u << 8, 0, s);
  }
}

int mgc_coarse_op_Mooee(mgc_coarse_op* op, const void* in, void* out, int cb, void* stream) {
  return guarded([&] { mooee(op, in, out, cb, 0, as_stream(stream)); });
}
int mgc_coarse_op_MooeeInv(mgc_coarse_op* op, const void* in, void* out, int cb, void* stream) {
  return guarded([&] { mooee(op, in, out, cb, 1, as_stream(stream)); });
}
int mgc_coarse_op_MooeeDag(mgc_coarse_op* op, const void* in, void* out, int cb, void* stream) {
  return guarded([&] { mooee(op, in, out, cb, 2, as_stream(stream)); });
}
int mgc_coarse_op_MooeeInvDag(mgc_coarse_op* op, const void* in, void* out, int cb,
                              void* stream) {
  return guarded([&] { mooee(op, in, out, cb, 3, as_stream(stream)); });
}

}  // extern "C"
