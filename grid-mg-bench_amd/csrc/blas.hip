// Complex BLAS-1 kernels with FP64 reductions (grid-stride elementwise ops,
// two-stage block reductions: NPART partial sums per result, then one block).
#include "blas.hpp"

#include <type_traits>

// No floating-point contraction in this file: every a * b + c rounds the
// product and the sum separately, so the fused passes and the two-kernel forms
// they replace (same expressions, same element order) give the same bits in
// FP64 as well.  With contraction left to the compiler, each kernel's choice
// of which products to fold into an FMA differed.  The kernels are bound by
// memory traffic; the extra FP64 instructions do not show.
#pragma clang fp contract(off)

namespace mgc {

Blas::Blas(hipStream_t st) : s(st) {
  MGC_HIP(hipMalloc(&d_part, (size_t)(KMAX + 1) * NPART * 2 * sizeof(double)));
  MGC_HIP(hipMalloc(&d_res, (size_t)(KMAX + 1) * 2 * sizeof(double)));
  MGC_HIP(hipMalloc(&d_coef, (size_t)KMAX * sizeof(cplx<double>)));
  MGC_HIP(hipHostMalloc(&h_res, (size_t)4 * (KMAX + 1) * 2 * sizeof(double),
                        hipHostMallocDefault));
  MGC_HIP(hipHostMalloc(&h_coef, (size_t)KMAX * sizeof(cplx<double>), hipHostMallocDefault));
  MGC_HIP(hipEventCreateWithFlags(&coef_done, hipEventDisableTiming));
}
Blas::~Blas() {
  if (d_part) (void)hipFree(d_part);
  if (d_res) (void)hipFree(d_res);
  if (d_coef) (void)hipFree(d_coef);
  if (h_res) (void)hipHostFree(h_res);
  if (h_coef) (void)hipHostFree(h_coef);
  if (coef_done) (void)hipEventDestroy(coef_done);
}

static unsigned ew_grid(size_t n) { return grid_for(n, 256, 65536); }

template <typename R>
__global__ void k_vzero(cplx<R>* x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = {R(0), R(0)};
}
template <typename R>
__global__ void k_vcopy(const cplx<R>* x, cplx<R>* y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
template <typename Rs, typename Rd>
__global__ void k_vconvert(const cplx<Rs>* x, cplx<Rd>* y, size_t n, double a) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = {(Rd)((double)x[i].re * a), (Rd)((double)x[i].im * a)};
}
template <typename R>
__global__ void k_vaxpy(double ar, double ai, const cplx<R>* x, cplx<R>* y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const cplx<R> v = x[i];
    y[i].re += (R)(ar * v.re - ai * v.im);
    y[i].im += (R)(ar * v.im + ai * v.re);
  }
}
template <typename R>
__global__ void k_vxpay(const cplx<R>* x, double ar, double ai, cplx<R>* y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const cplx<R> v = y[i];
    y[i] = {(R)(x[i].re + ar * v.re - ai * v.im), (R)(x[i].im + ar * v.im + ai * v.re)};
  }
}
template <typename R>
__global__ void k_vscale(double a, cplx<R>* x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = {(R)(a * x[i].re), (R)(a * x[i].im)};
}
template <typename R>
__global__ void k_vsub(const cplx<R>* x, const cplx<R>* y, cplx<R>* z, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    z[i] = {x[i].re - y[i].re, x[i].im - y[i].im};
}

__device__ inline double2 block_sum2(double a, double b) {
  __shared__ double sh[2 * 16];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[2 * w] = a;
    sh[2 * w + 1] = b;
  }
  __syncthreads();
  double ra = 0, rb = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      ra += sh[2 * k];
      rb += sh[2 * k + 1];
    }
  return make_double2(ra, rb);
}

// partial[j][blockIdx.x] = sum over the block's share of conj(V_j) w (or |x|^2)
template <typename R, bool NORM>
__global__ __launch_bounds__(256) void k_dot_part(const cplx<R>* V, size_t ldv,
                                                  const cplx<R>* w, size_t n, double* part) {
  const int j = blockIdx.y;
  const cplx<R>* v = V + (size_t)j * ldv;
  double re = 0.0, im = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double ar = v[i].re, ai = v[i].im;
    if (NORM) {
      re += ar * ar + ai * ai;
    } else {
      const double br = w[i].re, bi = w[i].im;
      re += ar * br + ai * bi;
      im += ar * bi - ai * br;
    }
  }
  const double2 t = block_sum2(re, im);
  if (threadIdx.x == 0) {
    part[((size_t)j * gridDim.x + blockIdx.x) * 2] = t.x;
    part[((size_t)j * gridDim.x + blockIdx.x) * 2 + 1] = t.y;
  }
}

// k_dot_final, skipped (whole grid, nothing written) when *skip is nonzero
__global__ __launch_bounds__(256) void k_dot_final_unless(const int* __restrict__ skip,
                                                          const double* part, int nparts,
                                                          double* res) {
  if (*skip) return;
  const int j = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    re += part[((size_t)j * nparts + i) * 2];
    im += part[((size_t)j * nparts + i) * 2 + 1];
  }
  const double2 t = block_sum2(re, im);
  if (threadIdx.x == 0) {
    res[2 * j] = t.x;
    res[2 * j + 1] = t.y;
  }
}

__global__ __launch_bounds__(256) void k_dot_final(const double* part, int nparts, double* res) {
  const int j = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    re += part[((size_t)j * nparts + i) * 2];
    im += part[((size_t)j * nparts + i) * 2 + 1];
  }
  const double2 t = block_sum2(re, im);
  if (threadIdx.x == 0) {
    res[2 * j] = t.x;
    res[2 * j + 1] = t.y;
  }
}

// Partial sums of <V_j, w> for j < k and, for row j == k when launched with
// k + 1 rows, of |w|^2.  blockIdx.y covers rows [MD_ROWS*y, MD_ROWS*(y+1)):
// each thread loads w[i] once for all of its rows, so w is read once per
// MD_ROWS rows instead of once per row.
constexpr int MD_ROWS = 8;
template <typename R>
__global__ __launch_bounds__(256) void k_mdot_part(const cplx<R>* V, size_t ldv, int k, int rows,
                                                   const cplx<R>* w, size_t n, double* part) {
  const int j0 = blockIdx.y * MD_ROWS;
  const int nj = min(MD_ROWS, rows - j0);
  double re[MD_ROWS], im[MD_ROWS];
#pragma unroll
  for (int jj = 0; jj < MD_ROWS; ++jj) re[jj] = im[jj] = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double br = w[i].re, bi = w[i].im;
#pragma unroll
    for (int jj = 0; jj < MD_ROWS; ++jj) {
      if (jj < nj) {
        const int j = j0 + jj;
        if (j == k) {
          re[jj] += br * br + bi * bi;
        } else {
          const cplx<R> a = V[(size_t)j * ldv + i];
          const double ar = a.re, ai = a.im;
          re[jj] += ar * br + ai * bi;
          im[jj] += ar * bi - ai * br;
        }
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < MD_ROWS; ++jj) {
    if (jj < nj) {  // uniform over the block
      const double2 t = block_sum2(re[jj], im[jj]);
      if (threadIdx.x == 0) {
        part[((size_t)(j0 + jj) * gridDim.x + blockIdx.x) * 2] = t.x;
        part[((size_t)(j0 + jj) * gridDim.x + blockIdx.x) * 2 + 1] = t.y;
      }
      __syncthreads();  // block_sum2's LDS is reused by the next row
    }
  }
}

template <typename R>
__global__ void k_vmultisub(const cplx<double>* c, const cplx<R>* V, size_t ldv, int k,
                            cplx<R>* w, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double re = w[i].re, im = w[i].im;
    for (int j = 0; j < k; ++j) {
      const cplx<R> v = V[(size_t)j * ldv + i];
      re -= c[j].re * v.re - c[j].im * v.im;
      im -= c[j].re * v.im + c[j].im * v.re;
    }
    w[i] = {(R)re, (R)im};
  }
}

template <typename R>
__global__ void k_vnormalize(const cplx<R>* w, cplx<R>* v, size_t n, cplx<double>* h2, int k) {
  __shared__ double s_inv;
  if (threadIdx.x == 0) {
    double q = h2[k].re;
    for (int j = 0; j < k; ++j) q -= h2[j].re * h2[j].re + h2[j].im * h2[j].im;
    const double hn = q > 0.0 ? sqrt(q) : 0.0;
    s_inv = hn > 0.0 ? 1.0 / hn : 0.0;
    if (blockIdx.x == 0) h2[k].im = hn;
  }
  __syncthreads();
  const double a = s_inv;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    v[i] = {(R)(a * w[i].re), (R)(a * w[i].im)};
}

// k_mdot_part for complex FP32 vectors, two elements per 16-byte load (n,
// ldv even and the pointers 16-byte aligned); the same partial sums per
// thread in the same element order
__global__ __launch_bounds__(256) void k_mdot_part_f2(const cplx<float>* V, size_t ldv, int k,
                                                      int rows, const cplx<float>* w, size_t n,
                                                      double* part) {
  const int j0 = blockIdx.y * MD_ROWS;
  const int nj = min(MD_ROWS, rows - j0);
  double re[MD_ROWS], im[MD_ROWS];
#pragma unroll
  for (int jj = 0; jj < MD_ROWS; ++jj) re[jj] = im[jj] = 0.0;
  const size_t n2 = n / 2;
  for (size_t i2 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i2 < n2;
       i2 += (size_t)gridDim.x * blockDim.x) {
    const float4 b = reinterpret_cast<const float4*>(w)[i2];
#pragma unroll
    for (int jj = 0; jj < MD_ROWS; ++jj) {
      if (jj < nj) {
        const int j = j0 + jj;
        if (j == k) {
          re[jj] += (double)b.x * b.x + (double)b.y * b.y;
          re[jj] += (double)b.z * b.z + (double)b.w * b.w;
        } else {
          const float4 a = reinterpret_cast<const float4*>(V + (size_t)j * ldv)[i2];
          re[jj] += (double)a.x * b.x + (double)a.y * b.y;
          im[jj] += (double)a.x * b.y - (double)a.y * b.x;
          re[jj] += (double)a.z * b.z + (double)a.w * b.w;
          im[jj] += (double)a.z * b.w - (double)a.w * b.z;
        }
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < MD_ROWS; ++jj) {
    if (jj < nj) {
      const double2 t = block_sum2(re[jj], im[jj]);
      if (threadIdx.x == 0) {
        part[((size_t)(j0 + jj) * gridDim.x + blockIdx.x) * 2] = t.x;
        part[((size_t)(j0 + jj) * gridDim.x + blockIdx.x) * 2 + 1] = t.y;
      }
      __syncthreads();
    }
  }
}

static bool pairable(const void* a, const void* b, size_t ldv, size_t n) {
  return n % 2 == 0 && ldv % 2 == 0 && ((uintptr_t)a % 16) == 0 && ((uintptr_t)b % 16) == 0;
}

// Fused re-orthogonalisation pass: w <- w - sum_j c[j] V_j (stored in the
// vector precision), and the partial sums of <V_j, w> for j < k and of |w|^2
// (row k) of the updated w, as k_vmultisub followed by k_mdot_part /
// k_mdot_part_f2 would produce them (same grid, same per-thread element
// order and arithmetic, so the same bits), reading V once from memory.  With
// KT > 8 rows, the first MSUB_STAGED values of a thread's element are kept in
// that thread's own LDS slots between the two passes (no thread reads another
// thread's slots, so no barrier): the window of V a resident grid touches
// between the two reads (k x 16 B x 256 threads per workgroup) outgrows L2 at
// k = 16..24, and the second read would go back to memory.  Rows j >=
// MSUB_STAGED, and all rows when KT <= 8, are read again (an L2 hit).
// Sixteen staged rows (64 KB of LDS, two workgroups per CU) ran faster at
// k = 17..24 than eight, and as fast at k = 9..16.  KT >= k.
constexpr int MSUB_STAGED = 16;
template <typename R, int KT, bool PAIRED>
__global__ __launch_bounds__(256) void k_msub_mdot(const cplx<double>* __restrict__ c,
                                                   const cplx<R>* V, size_t ldv, int k,
                                                   cplx<R>* w, size_t n, double* part,
                                                   const int* __restrict__ skip) {
  if (skip && *skip) return;  // uniform: the whole grid returns, w and part untouched
  constexpr int KS = KT > 8 ? MSUB_STAGED : 0;
  using Slot = typename std::conditional<PAIRED, float4, cplx<R>>::type;
  __shared__ Slot s_v[KS > 0 ? KS : 1][256];
  __shared__ cplx<double> s_c[KT];
  for (int j = threadIdx.x; j < k; j += blockDim.x) s_c[j] = c[j];
  __syncthreads();
  double re[KT + 1], im[KT];
#pragma unroll
  for (int j = 0; j <= KT; ++j) re[j] = 0.0;
#pragma unroll
  for (int j = 0; j < KT; ++j) im[j] = 0.0;
  if constexpr (!PAIRED) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
      double xr = w[i].re, xi = w[i].im;
#pragma unroll
      for (int j = 0; j < KT; ++j)
        if (j < k) {
          const cplx<R> v = V[(size_t)j * ldv + i];
          if (j < KS) s_v[j][threadIdx.x] = v;
          xr -= s_c[j].re * v.re - s_c[j].im * v.im;
          xi -= s_c[j].re * v.im + s_c[j].im * v.re;
        }
      const cplx<R> b = {(R)xr, (R)xi};
      w[i] = b;
      // re-read V_j (from LDS or memory) below instead of holding k loaded
      // values in registers
      asm volatile("" ::: "memory");
      const double br = b.re, bi = b.im;
#pragma unroll
      for (int j = 0; j < KT; ++j)
        if (j < k) {
          const cplx<R> a = j < KS ? s_v[j][threadIdx.x] : V[(size_t)j * ldv + i];
          const double ar = a.re, ai = a.im;
          re[j] += ar * br + ai * bi;
          im[j] += ar * bi - ai * br;
        }
      re[KT] += br * br + bi * bi;
    }
  } else {
    // FP32 pairs: elements 2 i2 and 2 i2 + 1 (one 16-byte load per row)
    const size_t n2 = n / 2;
    for (size_t i2 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i2 < n2;
         i2 += (size_t)gridDim.x * blockDim.x) {
      const float4 wv = reinterpret_cast<const float4*>(w)[i2];
      double x0r = wv.x, x0i = wv.y, x1r = wv.z, x1i = wv.w;
#pragma unroll
      for (int j = 0; j < KT; ++j)
        if (j < k) {
          const float4 v = reinterpret_cast<const float4*>(V + (size_t)j * ldv)[i2];
          if (j < KS) s_v[j][threadIdx.x] = v;
          x0r -= s_c[j].re * v.x - s_c[j].im * v.y;
          x0i -= s_c[j].re * v.y + s_c[j].im * v.x;
          x1r -= s_c[j].re * v.z - s_c[j].im * v.w;
          x1i -= s_c[j].re * v.w + s_c[j].im * v.z;
        }
      const float4 b = make_float4((float)x0r, (float)x0i, (float)x1r, (float)x1i);
      reinterpret_cast<float4*>(w)[i2] = b;
      asm volatile("" ::: "memory");
#pragma unroll
      for (int j = 0; j < KT; ++j)
        if (j < k) {
          const float4 a = j < KS ? s_v[j][threadIdx.x]
                                  : reinterpret_cast<const float4*>(V + (size_t)j * ldv)[i2];
          re[j] += (double)a.x * b.x + (double)a.y * b.y;
          im[j] += (double)a.x * b.y - (double)a.y * b.x;
          re[j] += (double)a.z * b.z + (double)a.w * b.w;
          im[j] += (double)a.z * b.w - (double)a.w * b.z;
        }
      re[KT] += (double)b.x * b.x + (double)b.y * b.y;
      re[KT] += (double)b.z * b.z + (double)b.w * b.w;
    }
  }
#pragma unroll
  for (int j = 0; j <= KT; ++j) {
    if (j < k || j == KT) {  // uniform over the block
      const int row = j < k ? j : k;
      const double2 t = block_sum2(re[j], j < KT ? im[j] : 0.0);
      if (threadIdx.x == 0) {
        part[((size_t)row * gridDim.x + blockIdx.x) * 2] = t.x;
        part[((size_t)row * gridDim.x + blockIdx.x) * 2 + 1] = t.y;
      }
      __syncthreads();  // block_sum2's LDS is reused by the next row
    }
  }
}

// v = (w - sum_j h2[j] V_j) / hn, hn = sqrt(h2[k].re - sum_j |h2[j]|^2)
template <typename R>
__global__ __launch_bounds__(256) void k_subnorm(cplx<double>* h2, const cplx<R>* __restrict__ V,
                                                 size_t ldv, int k, const cplx<R>* __restrict__ w,
                                                 cplx<R>* __restrict__ v, size_t n) {
  __shared__ cplx<double> s_h[Blas::KMAX];
  __shared__ double s_inv;
  for (int j = threadIdx.x; j < k; j += blockDim.x) s_h[j] = h2[j];
  if (threadIdx.x == 0) {
    double q = h2[k].re;
    for (int j = 0; j < k; ++j) q -= h2[j].re * h2[j].re + h2[j].im * h2[j].im;
    const double hn = q > 0.0 ? sqrt(q) : 0.0;
    s_inv = hn > 0.0 ? 1.0 / hn : 0.0;
    if (blockIdx.x == 0) h2[k].im = hn;  // other blocks read only h2[k].re
  }
  __syncthreads();
  const double a = s_inv;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double xr = w[i].re, xi = w[i].im;
    for (int j = 0; j < k; ++j) {
      const cplx<R> b = V[(size_t)j * ldv + i];
      xr -= s_h[j].re * b.re - s_h[j].im * b.im;
      xi -= s_h[j].re * b.im + s_h[j].im * b.re;
    }
    v[i] = {(R)(a * xr), (R)(a * xi)};
  }
}

template <typename R>
__global__ void k_vmultiaxpy(const cplx<double>* c, const cplx<R>* V, size_t ldv, int k,
                             cplx<R>* w, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double re = w[i].re, im = w[i].im;
    for (int j = 0; j < k; ++j) {
      const cplx<R> v = V[(size_t)j * ldv + i];
      re += c[j].re * v.re - c[j].im * v.im;
      im += c[j].re * v.im + c[j].im * v.re;
    }
    w[i] = {(R)re, (R)im};
  }
}

// y = x * a with a = d_s[0] (scale_mode 0), 1 / sqrt(d_s[0]) (1) or
// sqrt(d_s[0]) (2), 0 where the scalar is not positive; x and y given
// directly or through device pointer slots (px, py non-null)
template <typename Rs, typename Rd>
__global__ void k_vconvert_dev(const cplx<Rs>* x, cplx<Rd>* y, const cplx<Rs>* const* px,
                               cplx<Rd>* const* py, size_t n, const double* d_s, int mode) {
  const double v = d_s[0];
  const double a = mode == 0 ? v : !(v > 0.0) ? 0.0 : mode == 1 ? 1.0 / sqrt(v) : sqrt(v);
  if (px) x = *px;
  if (py) y = *py;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = {(Rd)(x[i].re * a), (Rd)(x[i].im * a)};
}

// |x|^2 into res[0] (res[1] = 0) with x read through a device pointer slot
template <typename R>
__global__ __launch_bounds__(256) void k_norm_part_ind(const cplx<R>* const* px, size_t n,
                                                       double* part) {
  const cplx<R>* x = *px;
  double re = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    re += (double)x[i].re * x[i].re + (double)x[i].im * x[i].im;
  const double2 t = block_sum2(re, 0.0);
  if (threadIdx.x == 0) {
    part[(size_t)blockIdx.x * 2] = t.x;
    part[(size_t)blockIdx.x * 2 + 1] = t.y;
  }
}

template <typename R> void vzero(cplx<R>* x, size_t n, hipStream_t s) {
  k_vzero<R><<<ew_grid(n), 256, 0, s>>>(x, n);
  MGC_LAUNCH_CHECK();
}
template <typename R> void vcopy(const cplx<R>* x, cplx<R>* y, size_t n, hipStream_t s) {
  k_vcopy<R><<<ew_grid(n), 256, 0, s>>>(x, y, n);
  MGC_LAUNCH_CHECK();
}
template <typename Rs, typename Rd>
void vconvert(const cplx<Rs>* x, cplx<Rd>* y, size_t n, double scale, hipStream_t s) {
  k_vconvert<Rs, Rd><<<ew_grid(n), 256, 0, s>>>(x, y, n, scale);
  MGC_LAUNCH_CHECK();
}
template <typename R> void vaxpy(cplx<double> a, const cplx<R>* x, cplx<R>* y, size_t n,
                                 hipStream_t s) {
  k_vaxpy<R><<<ew_grid(n), 256, 0, s>>>(a.re, a.im, x, y, n);
  MGC_LAUNCH_CHECK();
}
template <typename R> void vxpay(const cplx<R>* x, cplx<double> a, cplx<R>* y, size_t n,
                                 hipStream_t s) {
  k_vxpay<R><<<ew_grid(n), 256, 0, s>>>(x, a.re, a.im, y, n);
  MGC_LAUNCH_CHECK();
}
template <typename R> void vscale(double a, cplx<R>* x, size_t n, hipStream_t s) {
  k_vscale<R><<<ew_grid(n), 256, 0, s>>>(a, x, n);
  MGC_LAUNCH_CHECK();
}
template <typename R> void vsub(const cplx<R>* x, const cplx<R>* y, cplx<R>* z, size_t n,
                                hipStream_t s) {
  k_vsub<R><<<ew_grid(n), 256, 0, s>>>(x, y, z, n);
  MGC_LAUNCH_CHECK();
}

static int nparts_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > (size_t)Blas::NPART) g = Blas::NPART;
  if (g < 1) g = 1;
  return (int)g;
}

template <typename R, bool NORM>
static void reduce(Blas& b, const cplx<R>* V, size_t ldv, int k, const cplx<R>* w, size_t n) {
  MGC_CHECK(k >= 1 && k <= Blas::KMAX, MGC_ERR_ARG, "too many simultaneous reductions");
  const int np = nparts_for(n);
  k_dot_part<R, NORM><<<dim3(np, k), 256, 0, b.s>>>(V, ldv, w, n, b.d_part);
  MGC_LAUNCH_CHECK();
  k_dot_final<<<k, 256, 0, b.s>>>(b.d_part, np, b.d_res);
  MGC_LAUNCH_CHECK();
  MGC_HIP(hipMemcpyAsync(b.h_res, b.d_res, (size_t)k * 2 * sizeof(double), hipMemcpyDeviceToHost,
                         b.s));
  MGC_HIP(hipStreamSynchronize(b.s));
}

template <typename R> cplx<double> vdot(Blas& b, const cplx<R>* x, const cplx<R>* y, size_t n) {
  reduce<R, false>(b, x, 0, 1, y, n);
  return {b.h_res[0], b.h_res[1]};
}
template <typename R> double vnorm2(Blas& b, const cplx<R>* x, size_t n) {
  reduce<R, true>(b, x, 0, 1, x, n);
  return b.h_res[0];
}
template <typename R> void vmultidot(Blas& b, const cplx<R>* V, size_t ldv, int k,
                                     const cplx<R>* w, size_t n, cplx<double>* h) {
  MGC_CHECK(k >= 1 && k <= Blas::KMAX, MGC_ERR_ARG, "too many simultaneous reductions");
  // the multi-row reduction reads w once per MD_ROWS rows
  vmultidot_dev<R>(b, V, ldv, k, w, n, false, reinterpret_cast<cplx<double>*>(b.d_res));
  fetch(b, b.d_res, 2 * k);
  for (int j = 0; j < k; ++j) h[j] = {b.h_res[2 * j], b.h_res[2 * j + 1]};
}
template <typename R> void vmultiaxpy(Blas& b, const cplx<double>* c, const cplx<R>* V, size_t ldv,
                                      int k, cplx<R>* w, size_t n) {
  MGC_CHECK(k >= 0 && k <= Blas::KMAX, MGC_ERR_ARG, "too many vectors");
  if (k == 0) return;
  // the previous upload must have been read before the staging is rewritten
  MGC_HIP(hipEventSynchronize(b.coef_done));
  for (int j = 0; j < k; ++j) b.h_coef[j] = c[j];
  MGC_HIP(hipMemcpyAsync(b.d_coef, b.h_coef, (size_t)k * sizeof(cplx<double>),
                         hipMemcpyHostToDevice, b.s));
  MGC_HIP(hipEventRecord(b.coef_done, b.s));
  k_vmultiaxpy<R><<<ew_grid(n), 256, 0, b.s>>>(b.d_coef, V, ldv, k, w, n);
  MGC_LAUNCH_CHECK();
}

template <typename R> void vmultidot_dev(Blas& b, const cplx<R>* V, size_t ldv, int k,
                                         const cplx<R>* w, size_t n, bool with_norm,
                                         cplx<double>* d_h) {
  MGC_CHECK(k >= 0 && k <= Blas::KMAX && (k > 0 || with_norm), MGC_ERR_ARG,
            "too many simultaneous reductions");
  const int rows = k + (with_norm ? 1 : 0);
  const dim3 g(nparts_for(std::is_same<R, float>::value ? n / 2 : n), (rows + MD_ROWS - 1) / MD_ROWS);
  const int np = (int)g.x;
  if constexpr (std::is_same<R, float>::value) {
    if (pairable(V, w, ldv, n))
      k_mdot_part_f2<<<g, 256, 0, b.s>>>(V, ldv, k, rows, w, n, b.d_part);
    else
      k_mdot_part<R><<<g, 256, 0, b.s>>>(V, ldv, k, rows, w, n, b.d_part);
  } else {
    k_mdot_part<R><<<g, 256, 0, b.s>>>(V, ldv, k, rows, w, n, b.d_part);
  }
  MGC_LAUNCH_CHECK();
  k_dot_final<<<rows, 256, 0, b.s>>>(b.d_part, np, reinterpret_cast<double*>(d_h));
  MGC_LAUNCH_CHECK();
}
template <typename R> void vmultisub_dev(const cplx<double>* d_c, const cplx<R>* V, size_t ldv,
                                         int k, cplx<R>* w, size_t n, hipStream_t s) {
  if (k == 0) return;
  k_vmultisub<R><<<ew_grid(n), 256, 0, s>>>(d_c, V, ldv, k, w, n);
  MGC_LAUNCH_CHECK();
}
template <typename R> void vnormalize_dev(const cplx<R>* w, cplx<R>* v, size_t n,
                                          cplx<double>* d_h2, int k, hipStream_t s) {
  k_vnormalize<R><<<ew_grid(n), 256, 0, s>>>(w, v, n, d_h2, k);
  MGC_LAUNCH_CHECK();
}
template <typename R> void vsubnorm_dev(cplx<double>* d_h2, const cplx<R>* V, size_t ldv, int k,
                                        const cplx<R>* w, cplx<R>* v, size_t n, hipStream_t s) {
  MGC_CHECK(k >= 1 && k <= Blas::KMAX, MGC_ERR_ARG, "fused CGS2 pass: k out of range");
  k_subnorm<R><<<ew_grid(n), 256, 0, s>>>(d_h2, V, ldv, k, w, v, n);
  MGC_LAUNCH_CHECK();
}
template <typename R, int KT>
static void launch_msub_mdot(Blas& b, const cplx<double>* d_c, const cplx<R>* V, size_t ldv,
                             int k, cplx<R>* w, size_t n, int np, const int* skip) {
  const dim3 g(np);
  if constexpr (std::is_same<R, float>::value) {
    if (pairable(V, w, ldv, n)) {
      k_msub_mdot<R, KT, true><<<g, 256, 0, b.s>>>(d_c, V, ldv, k, w, n, b.d_part, skip);
      return;
    }
  }
  k_msub_mdot<R, KT, false><<<g, 256, 0, b.s>>>(d_c, V, ldv, k, w, n, b.d_part, skip);
}

template <typename R> bool vmultisub_dot_dev(Blas& b, const cplx<double>* d_c, const cplx<R>* V,
                                             size_t ldv, int k, cplx<R>* w, size_t n,
                                             cplx<double>* d_h, const int* skip) {
  if (k < 1 || k > 24) return false;
  // the grid of vmultidot_dev, so that the partial sums are the same
  const bool paired = std::is_same<R, float>::value && pairable(V, w, ldv, n);
  const int np = nparts_for(paired ? n / 2 : n);
  if (k <= 8) launch_msub_mdot<R, 8>(b, d_c, V, ldv, k, w, n, np, skip);
  else if (k <= 16) launch_msub_mdot<R, 16>(b, d_c, V, ldv, k, w, n, np, skip);
  else launch_msub_mdot<R, 24>(b, d_c, V, ldv, k, w, n, np, skip);
  MGC_LAUNCH_CHECK();
  if (skip)
    k_dot_final_unless<<<k + 1, 256, 0, b.s>>>(skip, b.d_part, np, reinterpret_cast<double*>(d_h));
  else
    k_dot_final<<<k + 1, 256, 0, b.s>>>(b.d_part, np, reinterpret_cast<double*>(d_h));
  MGC_LAUNCH_CHECK();
  return true;
}

template <typename Rs, typename Rd>
void vconvert_dev(const cplx<Rs>* x, cplx<Rd>* y, const cplx<Rs>* const* px, cplx<Rd>* const* py,
                  size_t n, const double* d_s, int mode, hipStream_t s) {
  k_vconvert_dev<Rs, Rd><<<ew_grid(n), 256, 0, s>>>(x, y, px, py, n, d_s, mode);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void vnorm2_dev(Blas& b, const cplx<R>* x, const cplx<R>* const* px, size_t n, double* d_res) {
  const int np = nparts_for(n);
  if (px) k_norm_part_ind<R><<<np, 256, 0, b.s>>>(px, n, b.d_part);
  else k_dot_part<R, true><<<dim3(np, 1), 256, 0, b.s>>>(x, 0, x, n, b.d_part);
  MGC_LAUNCH_CHECK();
  k_dot_final<<<1, 256, 0, b.s>>>(b.d_part, np, d_res);
  MGC_LAUNCH_CHECK();
}

void fetch(Blas& b, const double* d_src, int count) {
  MGC_CHECK(count >= 0 && count <= 4 * (Blas::KMAX + 1) * 2, MGC_ERR_ARG, "fetch too large");
  MGC_HIP(hipMemcpyAsync(b.h_res, d_src, (size_t)count * sizeof(double), hipMemcpyDeviceToHost,
                         b.s));
  MGC_HIP(hipStreamSynchronize(b.s));
}

#define INST(R)                                                                              \
  template void vzero<R>(cplx<R>*, size_t, hipStream_t);                                     \
  template void vcopy<R>(const cplx<R>*, cplx<R>*, size_t, hipStream_t);                     \
  template void vaxpy<R>(cplx<double>, const cplx<R>*, cplx<R>*, size_t, hipStream_t);       \
  template void vxpay<R>(const cplx<R>*, cplx<double>, cplx<R>*, size_t, hipStream_t);       \
  template void vscale<R>(double, cplx<R>*, size_t, hipStream_t);                            \
  template void vsub<R>(const cplx<R>*, const cplx<R>*, cplx<R>*, size_t, hipStream_t);      \
  template cplx<double> vdot<R>(Blas&, const cplx<R>*, const cplx<R>*, size_t);              \
  template double vnorm2<R>(Blas&, const cplx<R>*, size_t);                                  \
  template void vmultidot<R>(Blas&, const cplx<R>*, size_t, int, const cplx<R>*, size_t,     \
                             cplx<double>*);                                                 \
  template void vmultiaxpy<R>(Blas&, const cplx<double>*, const cplx<R>*, size_t, int,       \
                              cplx<R>*, size_t);                                             \
  template void vmultidot_dev<R>(Blas&, const cplx<R>*, size_t, int, const cplx<R>*, size_t, \
                                 bool, cplx<double>*);                                       \
  template void vmultisub_dev<R>(const cplx<double>*, const cplx<R>*, size_t, int, cplx<R>*, \
                                 size_t, hipStream_t);                                       \
  template void vnormalize_dev<R>(const cplx<R>*, cplx<R>*, size_t, cplx<double>*, int,      \
                                  hipStream_t);                                              \
  template void vsubnorm_dev<R>(cplx<double>*, const cplx<R>*, size_t, int, const cplx<R>*,   \
                                cplx<R>*, size_t, hipStream_t);                              \
  template bool vmultisub_dot_dev<R>(Blas&, const cplx<double>*, const cplx<R>*, size_t, int,  \
                                     cplx<R>*, size_t, cplx<double>*, const int*);
INST(double)
INST(float)
#undef INST
template void vnorm2_dev<double>(Blas&, const cplx<double>*, const cplx<double>* const*, size_t,
                                 double*);
template void vnorm2_dev<float>(Blas&, const cplx<float>*, const cplx<float>* const*, size_t,
                                double*);
#define INSTC(Rs, Rd)                                                                           \
  template void vconvert_dev<Rs, Rd>(const cplx<Rs>*, cplx<Rd>*, const cplx<Rs>* const*,       \
                                     cplx<Rd>* const*, size_t, const double*, int, hipStream_t);
INSTC(double, float)
INSTC(float, double)
INSTC(double, double)
INSTC(float, float)
#undef INSTC
template void vconvert<double, float>(const cplx<double>*, cplx<float>*, size_t, double,
                                      hipStream_t);
template void vconvert<float, double>(const cplx<float>*, cplx<double>*, size_t, double,
                                      hipStream_t);
template void vconvert<double, double>(const cplx<double>*, cplx<double>*, size_t, double,
                                       hipStream_t);
template void vconvert<float, float>(const cplx<float>*, cplx<float>*, size_t, double,
                                     hipStream_t);

}  // namespace mgc
