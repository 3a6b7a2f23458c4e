// Batched complex BLAS-1 with per-rhs FP64 reductions (see bblas.hpp).
#include "bblas.hpp"

#include <algorithm>
#include <type_traits>
#include <vector>

// No floating-point contraction in this file: every a * b + c rounds the
// product and the sum separately, so the fused passes and the two-kernel forms
// they replace (same expressions, same element order) give the same bits in
// FP64 as well.  With contraction left to the compiler, each kernel's choice
// of which products to fold into an FMA differed.  The kernels are bound by
// memory traffic; the extra FP64 instructions do not show.
#pragma clang fp contract(off)

namespace mgc {

BBlas::BBlas(hipStream_t st) : s(st) {
  MGC_HIP(hipHostMalloc(&h_res, (size_t)(2 * KMAX + 3) * RMAX * 2 * sizeof(double),
                        hipHostMallocDefault));
  MGC_HIP(hipMalloc(&d_coef, (size_t)(KMAX + 1) * RMAX * sizeof(cplx<double>)));
  MGC_HIP(hipHostMalloc(&h_coef, (size_t)(KMAX + 1) * RMAX * sizeof(cplx<double>),
                        hipHostMallocDefault));
  MGC_HIP(hipEventCreateWithFlags(&coef_done, hipEventDisableTiming));
}
BBlas::~BBlas() {
  if (d_part) (void)hipFree(d_part);
  if (h_res) (void)hipHostFree(h_res);
  if (d_coef) (void)hipFree(d_coef);
  if (h_coef) (void)hipHostFree(h_coef);
  if (coef_done) (void)hipEventDestroy(coef_done);
}

static unsigned bew_grid(size_t n) { return grid_for(n, 256, 65536); }

static size_t gcd_sz(size_t a, size_t b) {
  while (b) {
    const size_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// reduction grid: a multiple of C/gcd(256, C) blocks (C = columns per site,
// or half of them when a thread takes element pairs), so that the total
// thread count is a multiple of C and each thread keeps one rhs
static int bred_blocks(const BShape& sh, bool pairs) {
  const size_t C = (size_t)sh.cols() / (pairs ? 2 : 1);
  const size_t n = sh.n / (pairs ? 2 : 1);
  const size_t base = C / gcd_sz(256, C);  // blocks per lcm(256, C) threads
  size_t want = std::min<size_t>(2048, (n + 255) / 256);
  want = std::max<size_t>(want, base);
  return (int)((want + base - 1) / base * base);
}

constexpr int BD_ROWS = 8;

// i % C along a grid-stride walk i = i0, i0 + T, i0 + 2T, ...: one 64-bit
// division at the start instead of one per element (C = nrhs * dof < 2^31).
struct ModWalk {
  unsigned m, step, c;
  __device__ ModWalk(size_t i0, size_t T, int C)
      : m((unsigned)(i0 % (size_t)C)), step((unsigned)(T % (size_t)C)), c((unsigned)C) {}
  __device__ void next() {
    m += step;
    if (m >= c) m -= c;
  }
};

// rows [BD_ROWS*y, ...): <V_j, w>_r for j < k, |w|^2_r for j == k
template <typename R>
__global__ __launch_bounds__(256) void k_bdot_part(const cplx<R>* __restrict__ V, size_t ldv,
                                                   int k, int rows,
                                                   const cplx<R>* __restrict__ w, size_t n,
                                                   int C, int dof, int nrhs,
                                                   double* __restrict__ part) {
  __shared__ double sm[BD_ROWS][256][2];
  __shared__ int s_rt[256];  // the rhs of each thread's elements
  const int j0 = blockIdx.y * BD_ROWS;
  const int nj = min(BD_ROWS, rows - j0);
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  s_rt[threadIdx.x] = (int)((g % C) / dof);
  double re[BD_ROWS], im[BD_ROWS];
#pragma unroll
  for (int jj = 0; jj < BD_ROWS; ++jj) re[jj] = im[jj] = 0.0;
  for (size_t i = g; i < n; i += T) {
    const double br = w[i].re, bi = w[i].im;
#pragma unroll
    for (int jj = 0; jj < BD_ROWS; ++jj) {
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
  for (int jj = 0; jj < BD_ROWS; ++jj) {
    sm[jj][threadIdx.x][0] = re[jj];
    sm[jj][threadIdx.x][1] = im[jj];
  }
  __syncthreads();
  // per-rhs sums of this block, threads in ascending order
  for (int q = threadIdx.x; q < nj * nrhs; q += blockDim.x) {
    const int jj = q / nrhs, rr = q % nrhs;
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      if (s_rt[t] == rr) {
        sr += sm[jj][t][0];
        si += sm[jj][t][1];
      }
    }
    double* o = part + ((((size_t)(j0 + jj) * nrhs + rr) * gridDim.x) + blockIdx.x) * 2;
    o[0] = sr;
    o[1] = si;
  }
}

// FP32 form taking two consecutive elements (one 16-byte load; dof is even,
// so a pair never straddles two rhs): thread t keeps rhs ((2t) % C) / dof
__global__ __launch_bounds__(256) void k_bdot_part_f2(const cplx<float>* __restrict__ V,
                                                      size_t ldv, int k, int rows,
                                                      const cplx<float>* __restrict__ w,
                                                      size_t n, int C, int dof, int nrhs,
                                                      double* __restrict__ part) {
  __shared__ double sm[BD_ROWS][256][2];
  __shared__ int s_rt[256];  // the rhs of each thread's element pairs
  const int j0 = blockIdx.y * BD_ROWS;
  const int nj = min(BD_ROWS, rows - j0);
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n2 = n / 2;
  s_rt[threadIdx.x] = (int)((2 * (g % (size_t)(C / 2))) / dof);
  double re[BD_ROWS], im[BD_ROWS];
#pragma unroll
  for (int jj = 0; jj < BD_ROWS; ++jj) re[jj] = im[jj] = 0.0;
  for (size_t i2 = g; i2 < n2; i2 += T) {
    const float4 b = reinterpret_cast<const float4*>(w)[i2];
#pragma unroll
    for (int jj = 0; jj < BD_ROWS; ++jj) {
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
  for (int jj = 0; jj < BD_ROWS; ++jj) {
    sm[jj][threadIdx.x][0] = re[jj];
    sm[jj][threadIdx.x][1] = im[jj];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nj * nrhs; q += blockDim.x) {
    const int jj = q / nrhs, rr = q % nrhs;
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      if (s_rt[t] == rr) {
        sr += sm[jj][t][0];
        si += sm[jj][t][1];
      }
    }
    double* o = part + ((((size_t)(j0 + jj) * nrhs + rr) * gridDim.x) + blockIdx.x) * 2;
    o[0] = sr;
    o[1] = si;
  }
}

// Fused re-orthogonalisation pass: w -= sum_j c[j][r] V_j (stored in the
// vector precision), then the per-rhs partials of <V_j, w>_r for j < k and of
// |w|^2_r (row k) for the updated w, on the grid and with the per-thread
// element order of k_bdot_part / k_bdot_part_f2 (so the same partials), with
// V read once from memory.  With KT > 8 rows, the first BMSUB_STAGED values
// of a thread's element are kept in that thread's own LDS slots between the
// two passes (no thread reads another thread's slots, so no barrier): the
// window of V a resident grid touches between the two reads (k x 16 B x 256
// threads per workgroup) outgrows L2 at k = 16..24, and the second read would
// go back to memory.  Rows j >= BMSUB_STAGED, and all rows when KT <= 8, are
// read again (an L2 hit).  The staging slots and the block sums share one LDS
// buffer: the sums are written only after the barrier that ends the element
// loop for every thread.  Eight staged rows fill exactly the 32 KB the block
// sums take, so occupancy is that of the unstaged kernel; sixteen (64 KB, two
// workgroups per CU) ran slower than none.  KT >= k.
constexpr int BMSUB_STAGED = 8;
template <typename R, int KT, bool PAIRED>
__global__ __launch_bounds__(256) void k_bmsub_dot(const cplx<double>* __restrict__ c,
                                                   const cplx<R>* V, size_t ldv, int k,
                                                   cplx<R>* w, size_t n, int C, int dof, int nrhs,
                                                   double* __restrict__ part,
                                                   const int* __restrict__ skip) {
  if (skip && *skip) return;  // uniform: the whole grid returns, w and part untouched
  constexpr int KS = KT > 8 ? BMSUB_STAGED : 0;
  using Slot = typename std::conditional<PAIRED, float4, cplx<R>>::type;
  constexpr size_t SUM_BYTES = sizeof(double) * BD_ROWS * 256 * 2;
  constexpr size_t STAGE_BYTES = sizeof(Slot) * (KS > 0 ? KS : 1) * 256;
  __shared__ __align__(16) unsigned char s_raw[SUM_BYTES > STAGE_BYTES ? SUM_BYTES : STAGE_BYTES];
  double(&sm)[BD_ROWS][256][2] = *reinterpret_cast<double(*)[BD_ROWS][256][2]>(s_raw);
  Slot(*s_v)[256] = reinterpret_cast<Slot(*)[256]>(s_raw);
  __shared__ int s_rt[256];  // the rhs of each thread's first element
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  s_rt[threadIdx.x] = PAIRED ? (int)((2 * (g % (size_t)(C / 2))) / dof) : (int)((g % C) / dof);
  double re[KT + 1], im[KT];
#pragma unroll
  for (int j = 0; j <= KT; ++j) re[j] = 0.0;
#pragma unroll
  for (int j = 0; j < KT; ++j) im[j] = 0.0;
  if constexpr (!PAIRED) {
    ModWalk mw(g, T, C);
    for (size_t i = g; i < n; i += T, mw.next()) {
      const int r = (int)(mw.m / (unsigned)dof);
      double xr = w[i].re, xi = w[i].im;
#pragma unroll
      for (int j = 0; j < KT; ++j)
        if (j < k) {
          const cplx<double> cj = c[(size_t)j * nrhs + r];
          const cplx<R> v = V[(size_t)j * ldv + i];
          if (j < KS) s_v[j][threadIdx.x] = v;
          xr += -1.0 * (cj.re * v.re - cj.im * v.im);
          xi += -1.0 * (cj.re * v.im + cj.im * v.re);
        }
      const cplx<R> b = {(R)xr, (R)xi};
      w[i] = b;
      asm volatile("" ::: "memory");  // re-read V_j (LDS or memory) instead of holding it
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
    const size_t n2 = n / 2;
    const int C2 = C / 2;
    ModWalk mw(g, T, C2);
    for (size_t i2 = g; i2 < n2; i2 += T, mw.next()) {
      const int r = (int)((2 * mw.m) / (unsigned)dof);
      const float4 wv = reinterpret_cast<const float4*>(w)[i2];
      double x0r = wv.x, x0i = wv.y, x1r = wv.z, x1i = wv.w;
#pragma unroll
      for (int j = 0; j < KT; ++j)
        if (j < k) {
          const cplx<double> cj = c[(size_t)j * nrhs + r];
          const float4 v = reinterpret_cast<const float4*>(V + (size_t)j * ldv)[i2];
          if (j < KS) s_v[j][threadIdx.x] = v;
          x0r += -1.0 * (cj.re * v.x - cj.im * v.y);
          x0i += -1.0 * (cj.re * v.y + cj.im * v.x);
          x1r += -1.0 * (cj.re * v.z - cj.im * v.w);
          x1i += -1.0 * (cj.re * v.w + cj.im * v.z);
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
  // per-rhs sums of this block in groups of BD_ROWS rows (row k, |w|^2, is
  // the accumulator re[KT]), threads in ascending order as in k_bdot_part
  const int rows = k + 1;
#pragma unroll
  for (int j0 = 0; j0 <= KT; j0 += BD_ROWS) {
    if (j0 < rows) {  // uniform over the block
      __syncthreads();
#pragma unroll
      for (int jj = 0; jj < BD_ROWS; ++jj) {
        const int j = j0 + jj;  // group row; row k takes re[KT]
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int q = 0; q <= KT; ++q)
          if (q == j) {
            a = q < k ? re[q] : re[KT];
            b = q < k && q < KT ? im[q] : 0.0;
          }
        if (j == k) {
          a = re[KT];
          b = 0.0;
        }
        sm[jj][threadIdx.x][0] = a;
        sm[jj][threadIdx.x][1] = b;
      }
      __syncthreads();
      const int nj = min(BD_ROWS, rows - j0);
      for (int q = threadIdx.x; q < nj * nrhs; q += blockDim.x) {
        const int jj = q / nrhs, rr = q % nrhs;
        double sr = 0.0, si = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
          if (s_rt[t] == rr) {
            sr += sm[jj][t][0];
            si += sm[jj][t][1];
          }
        }
        double* o = part + ((((size_t)(j0 + jj) * nrhs + rr) * gridDim.x) + blockIdx.x) * 2;
        o[0] = sr;
        o[1] = si;
      }
    }
  }
}

// res[q] = sum over nblk partials of part[q][*], q = row*nrhs + r
__device__ __forceinline__ void bdot_final_body(const double* __restrict__ part, int nblk,
                                                double* __restrict__ res) {
  __shared__ double sh[2][256];
  const int q = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    re += part[((size_t)q * nblk + i) * 2];
    im += part[((size_t)q * nblk + i) * 2 + 1];
  }
  sh[0][threadIdx.x] = re;
  sh[1][threadIdx.x] = im;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[2 * q] = sh[0][0];
    res[2 * q + 1] = sh[1][0];
  }
}

__global__ __launch_bounds__(256) void k_bdot_final(const double* __restrict__ part, int nblk,
                                                    double* __restrict__ res) {
  bdot_final_body(part, nblk, res);
}

// k_bdot_final, skipped (whole grid, nothing written) when *skip is nonzero
__global__ __launch_bounds__(256) void k_bdot_final_unless(const int* __restrict__ skip,
                                                           const double* __restrict__ part,
                                                           int nblk, double* __restrict__ res) {
  if (*skip) return;
  bdot_final_body(part, nblk, res);
}

template <typename R>
__global__ void k_bmultisub(const cplx<double>* __restrict__ c, const cplx<R>* __restrict__ V,
                            size_t ldv, int k, cplx<R>* __restrict__ w, size_t n, int C, int dof,
                            int nrhs, double sign) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const int r = (int)(mw.m / (unsigned)dof);
    double re = w[i].re, im = w[i].im;
    for (int j = 0; j < k; ++j) {
      const cplx<double> cj = c[(size_t)j * nrhs + r];
      const cplx<R> v = V[(size_t)j * ldv + i];
      re += sign * (cj.re * v.re - cj.im * v.im);
      im += sign * (cj.re * v.im + cj.im * v.re);
    }
    w[i] = {(R)re, (R)im};
  }
}

template <typename R>
__global__ void k_bsubnorm(cplx<double>* h2, const cplx<R>* __restrict__ V, size_t ldv, int k,
                           const cplx<R>* __restrict__ w, cplx<R>* __restrict__ v, size_t n,
                           int C, int dof, int nrhs) {
  __shared__ double s_inv[BBlas::RMAX];
  for (int r = threadIdx.x; r < nrhs; r += blockDim.x) {
    double q = h2[(size_t)k * nrhs + r].re;
    for (int j = 0; j < k; ++j) {
      const cplx<double> h = h2[(size_t)j * nrhs + r];
      q -= h.re * h.re + h.im * h.im;
    }
    const double hn = q > 0.0 ? sqrt(q) : 0.0;
    s_inv[r] = hn > 0.0 ? 1.0 / hn : 0.0;
    if (blockIdx.x == 0) h2[(size_t)k * nrhs + r].im = hn;  // other blocks read only .re
  }
  __syncthreads();
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const int r = (int)(mw.m / (unsigned)dof);
    double re = w[i].re, im = w[i].im;
    for (int j = 0; j < k; ++j) {
      const cplx<double> cj = h2[(size_t)j * nrhs + r];
      const cplx<R> b = V[(size_t)j * ldv + i];
      re -= cj.re * b.re - cj.im * b.im;
      im -= cj.re * b.im + cj.im * b.re;
    }
    const double a = s_inv[r];
    v[i] = {(R)(a * re), (R)(a * im)};
  }
}

template <typename Rs, typename Rd>
__global__ void k_bconvert(const cplx<Rs>* __restrict__ x, cplx<Rd>* __restrict__ y, size_t n,
                           int C, int dof, const cplx<double>* __restrict__ a) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const double s = a[mw.m / (unsigned)dof].re;
    y[i] = {(Rd)((double)x[i].re * s), (Rd)((double)x[i].im * s)};
  }
}

static void check_shape(const BShape& sh) {
  MGC_CHECK(sh.nrhs >= 1 && sh.nrhs <= BBlas::RMAX, MGC_ERR_ARG, "batch of 1..64 rhs");
  MGC_CHECK(sh.dof >= 1 && sh.n % (size_t)sh.cols() == 0, MGC_ERR_SHAPE,
            "batched vector length is not sites x nrhs x dof");
}

template <typename R>
void bdots_dev(BBlas& b, const cplx<R>* V, size_t ldv, int k, const cplx<R>* w,
               const BShape& sh, bool with_norm, cplx<double>* d_out) {
  check_shape(sh);
  MGC_CHECK(k >= 0 && k <= BBlas::KMAX && (k > 0 || with_norm), MGC_ERR_ARG,
            "batched reduction: bad row count");
  const int rows = k + (with_norm ? 1 : 0);
  bool pairs = false;
  if constexpr (std::is_same<R, float>::value)
    pairs = sh.dof % 2 == 0 && ldv % 2 == 0 && ((uintptr_t)V % 16) == 0 &&
            ((uintptr_t)w % 16) == 0;
  const int nblk = bred_blocks(sh, pairs);
  const size_t need = (size_t)rows * sh.nrhs * nblk * 2;
  if (need > b.part_cap) {
    if (b.d_part) MGC_HIP(hipFree(b.d_part));
    b.d_part = nullptr;
    b.part_cap = 0;
    const size_t cap = std::max(need, (size_t)(BBlas::KMAX + 1) * sh.nrhs * nblk * 2);
    MGC_HIP(hipMalloc(&b.d_part, cap * sizeof(double)));
    b.part_cap = cap;
  }
  const dim3 grid(nblk, (rows + BD_ROWS - 1) / BD_ROWS);
  if constexpr (std::is_same<R, float>::value) {
    if (pairs)
      k_bdot_part_f2<<<grid, 256, 0, b.s>>>(V, ldv, k, rows, w, sh.n, sh.cols(), sh.dof,
                                            sh.nrhs, b.d_part);
    else
      k_bdot_part<R><<<grid, 256, 0, b.s>>>(V, ldv, k, rows, w, sh.n, sh.cols(), sh.dof,
                                            sh.nrhs, b.d_part);
  } else {
    k_bdot_part<R><<<grid, 256, 0, b.s>>>(V, ldv, k, rows, w, sh.n, sh.cols(), sh.dof, sh.nrhs,
                                          b.d_part);
  }
  MGC_LAUNCH_CHECK();
  k_bdot_final<<<rows * sh.nrhs, 256, 0, b.s>>>(b.d_part, nblk, reinterpret_cast<double*>(d_out));
  MGC_LAUNCH_CHECK();
}

void bfetch(BBlas& b, const double* d_src, int count) {
  MGC_CHECK(count >= 0 && count <= (2 * BBlas::KMAX + 3) * BBlas::RMAX * 2, MGC_ERR_ARG,
            "fetch too large");
  MGC_HIP(hipMemcpyAsync(b.h_res, d_src, (size_t)count * sizeof(double), hipMemcpyDeviceToHost,
                         b.s));
  MGC_HIP(hipStreamSynchronize(b.s));
}

template <typename R> void bnorm2(BBlas& b, const cplx<R>* x, const BShape& sh, double* out) {
  bdots_dev<R>(b, x, 0, 0, x, sh, true, b.d_coef + (size_t)BBlas::KMAX * BBlas::RMAX);
  bfetch(b, reinterpret_cast<const double*>(b.d_coef + (size_t)BBlas::KMAX * BBlas::RMAX),
         2 * sh.nrhs);
  for (int r = 0; r < sh.nrhs; ++r) out[r] = b.h_res[2 * r];
}

template <typename R, int KT>
static void launch_bmsub_dot(BBlas& b, const cplx<double>* d_c, const cplx<R>* V, size_t ldv,
                             int k, cplx<R>* w, const BShape& sh, bool pairs, int nblk,
                             const int* skip) {
  if constexpr (std::is_same<R, float>::value) {
    if (pairs) {
      k_bmsub_dot<R, KT, true><<<nblk, 256, 0, b.s>>>(d_c, V, ldv, k, w, sh.n, sh.cols(), sh.dof,
                                                        sh.nrhs, b.d_part, skip);
      return;
    }
  }
  k_bmsub_dot<R, KT, false><<<nblk, 256, 0, b.s>>>(d_c, V, ldv, k, w, sh.n, sh.cols(), sh.dof,
                                                     sh.nrhs, b.d_part, skip);
}

template <typename R>
bool bmultisub_dots_dev(BBlas& b, const cplx<double>* d_c, const cplx<R>* V, size_t ldv, int k,
                        cplx<R>* w, const BShape& sh, cplx<double>* d_out, const int* skip) {
  check_shape(sh);
  if (k < 1 || k > 24) return false;
  bool pairs = false;
  if constexpr (std::is_same<R, float>::value)
    pairs = sh.dof % 2 == 0 && ldv % 2 == 0 && ((uintptr_t)V % 16) == 0 &&
            ((uintptr_t)w % 16) == 0;
  const int nblk = bred_blocks(sh, pairs);
  const int rows = k + 1;
  const size_t need = (size_t)rows * sh.nrhs * nblk * 2;
  if (need > b.part_cap) {
    if (b.d_part) MGC_HIP(hipFree(b.d_part));
    b.d_part = nullptr;
    b.part_cap = 0;
    const size_t cap = std::max(need, (size_t)(BBlas::KMAX + 1) * sh.nrhs * nblk * 2);
    MGC_HIP(hipMalloc(&b.d_part, cap * sizeof(double)));
    b.part_cap = cap;
  }
  if (k <= 8) launch_bmsub_dot<R, 8>(b, d_c, V, ldv, k, w, sh, pairs, nblk, skip);
  else if (k <= 16) launch_bmsub_dot<R, 16>(b, d_c, V, ldv, k, w, sh, pairs, nblk, skip);
  else launch_bmsub_dot<R, 24>(b, d_c, V, ldv, k, w, sh, pairs, nblk, skip);
  MGC_LAUNCH_CHECK();
  if (skip)
    k_bdot_final_unless<<<rows * sh.nrhs, 256, 0, b.s>>>(skip, b.d_part, nblk,
                                                         reinterpret_cast<double*>(d_out));
  else
    k_bdot_final<<<rows * sh.nrhs, 256, 0, b.s>>>(b.d_part, nblk,
                                                  reinterpret_cast<double*>(d_out));
  MGC_LAUNCH_CHECK();
  return true;
}

template <typename R>
void bmultisub_dev(const cplx<double>* d_c, const cplx<R>* V, size_t ldv, int k, cplx<R>* w,
                   const BShape& sh, hipStream_t s) {
  check_shape(sh);
  if (k == 0) return;
  k_bmultisub<R><<<bew_grid(sh.n), 256, 0, s>>>(d_c, V, ldv, k, w, sh.n, sh.cols(), sh.dof,
                                                sh.nrhs, -1.0);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bmultiaxpy(BBlas& b, const cplx<double>* c, const cplx<R>* V, size_t ldv, int k,
                cplx<R>* w, const BShape& sh) {
  check_shape(sh);
  MGC_CHECK(k >= 0 && k <= BBlas::KMAX, MGC_ERR_ARG, "too many vectors");
  if (k == 0) return;
  MGC_HIP(hipEventSynchronize(b.coef_done));
  std::copy(c, c + (size_t)k * sh.nrhs, b.h_coef);
  MGC_HIP(hipMemcpyAsync(b.d_coef, b.h_coef, (size_t)k * sh.nrhs * sizeof(cplx<double>),
                         hipMemcpyHostToDevice, b.s));
  MGC_HIP(hipEventRecord(b.coef_done, b.s));
  k_bmultisub<R><<<bew_grid(sh.n), 256, 0, b.s>>>(b.d_coef, V, ldv, k, w, sh.n, sh.cols(),
                                                  sh.dof, sh.nrhs, 1.0);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bsubnorm_dev(cplx<double>* d_h2, const cplx<R>* V, size_t ldv, int k, const cplx<R>* w,
                  cplx<R>* v, const BShape& sh, hipStream_t s) {
  check_shape(sh);
  k_bsubnorm<R><<<bew_grid(sh.n), 256, 0, s>>>(d_h2, V, ldv, k, w, v, sh.n, sh.cols(), sh.dof,
                                               sh.nrhs);
  MGC_LAUNCH_CHECK();
}

template <typename Rs, typename Rd>
void bconvert(BBlas& b, const cplx<Rs>* x, cplx<Rd>* y, const BShape& sh, const double* a) {
  check_shape(sh);
  MGC_HIP(hipEventSynchronize(b.coef_done));
  for (int r = 0; r < sh.nrhs; ++r) b.h_coef[r] = {a[r], 0.0};
  MGC_HIP(hipMemcpyAsync(b.d_coef, b.h_coef, (size_t)sh.nrhs * sizeof(cplx<double>),
                         hipMemcpyHostToDevice, b.s));
  MGC_HIP(hipEventRecord(b.coef_done, b.s));
  k_bconvert<Rs, Rd><<<bew_grid(sh.n), 256, 0, b.s>>>(x, y, sh.n, sh.cols(), sh.dof, b.d_coef);
  MGC_LAUNCH_CHECK();
}

// ---- BiCGSTAB updates for nrhs systems in lockstep ----
// Coefficients [3][nrhs] on the device: rows 0 and 1 per kernel below, row 2
// (.re) 1 for a rhs still iterating; the vectors of a stopped rhs are left
// untouched, so a breakdown's infinities never reach them.

// p = r + beta (p - omega v); c0 = beta, c1 = omega
template <typename R>
__global__ void k_bbicg_p(const cplx<double>* __restrict__ c, const cplx<R>* __restrict__ r,
                          const cplx<R>* __restrict__ v, cplx<R>* __restrict__ p, size_t n, int C,
                          int dof, int nrhs) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const int q = (int)(mw.m / (unsigned)dof);
    if (c[2 * nrhs + q].re == 0.0) continue;
    const cplx<double> be = c[q], om = c[nrhs + q];
    const cplx<R> vi = v[i], pi = p[i], ri = r[i];
    const double tr = (double)pi.re - (om.re * vi.re - om.im * vi.im);
    const double ti = (double)pi.im - (om.re * vi.im + om.im * vi.re);
    p[i] = {(R)(ri.re + (be.re * tr - be.im * ti)), (R)(ri.im + (be.re * ti + be.im * tr))};
  }
}

// s = r - alpha v; c0 = alpha
template <typename R>
__global__ void k_bbicg_s(const cplx<double>* __restrict__ c, const cplx<R>* __restrict__ r,
                          const cplx<R>* __restrict__ v, cplx<R>* __restrict__ s, size_t n, int C,
                          int dof, int nrhs) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const int q = (int)(mw.m / (unsigned)dof);
    if (c[2 * nrhs + q].re == 0.0) continue;
    const cplx<double> al = c[q];
    const cplx<R> vi = v[i], ri = r[i];
    s[i] = {(R)(ri.re - (al.re * vi.re - al.im * vi.im)),
            (R)(ri.im - (al.re * vi.im + al.im * vi.re))};
  }
}

// x += alpha p + omega s; r = s - omega t; c0 = alpha, c1 = omega
template <typename R>
__global__ void k_bbicg_xr(const cplx<double>* __restrict__ c, const cplx<R>* __restrict__ p,
                           const cplx<R>* __restrict__ s, const cplx<R>* __restrict__ t,
                           cplx<R>* __restrict__ x, cplx<R>* __restrict__ r, size_t n, int C,
                           int dof, int nrhs) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const int q = (int)(mw.m / (unsigned)dof);
    if (c[2 * nrhs + q].re == 0.0) continue;
    const cplx<double> al = c[q], om = c[nrhs + q];
    const cplx<R> pi = p[i], si = s[i], ti = t[i], xi = x[i];
    const double xr = (double)xi.re + (al.re * pi.re - al.im * pi.im) +
                      (om.re * si.re - om.im * si.im);
    const double xm = (double)xi.im + (al.re * pi.im + al.im * pi.re) +
                      (om.re * si.im + om.im * si.re);
    x[i] = {(R)xr, (R)xm};
    r[i] = {(R)(si.re - (om.re * ti.re - om.im * ti.im)),
            (R)(si.im - (om.re * ti.im + om.im * ti.re))};
  }
}

// bbicg_xr fused with the reductions of the next step: x += alpha p + omega s,
// r = s - omega t, then per rhs <r, r> (row 0) and <rh, r> (row 1) of the
// updated r, on the reduction grid and with the block sums of k_bdot_part
// (one element per thread per step).  A stopped rhs keeps x and r and adds
// nothing to its sums.
template <typename R>
__global__ __launch_bounds__(256) void k_bbicg_xr_dots(
    const cplx<double>* __restrict__ c, const cplx<R>* __restrict__ p,
    const cplx<R>* __restrict__ s, const cplx<R>* __restrict__ t, const cplx<R>* __restrict__ rh,
    cplx<R>* __restrict__ x, cplx<R>* __restrict__ r, size_t n, int C, int dof, int nrhs,
    double* __restrict__ part) {
  __shared__ double sm[2][256][2];
  __shared__ int s_rt[256];
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = (int)((g % C) / dof);  // this thread's rhs, the same for all its elements
  s_rt[threadIdx.x] = q;
  double rr = 0.0, hr = 0.0, hi = 0.0;
  if (c[2 * nrhs + q].re != 0.0) {
    const cplx<double> al = c[q], om = c[nrhs + q];
    for (size_t i = g; i < n; i += T) {
      const cplx<R> pi = p[i], si = s[i], ti = t[i], xi = x[i];
      const double xr = (double)xi.re + (al.re * pi.re - al.im * pi.im) +
                        (om.re * si.re - om.im * si.im);
      const double xm = (double)xi.im + (al.re * pi.im + al.im * pi.re) +
                        (om.re * si.im + om.im * si.re);
      x[i] = {(R)xr, (R)xm};
      const cplx<R> rn = {(R)(si.re - (om.re * ti.re - om.im * ti.im)),
                          (R)(si.im - (om.re * ti.im + om.im * ti.re))};
      r[i] = rn;
      const double br = rn.re, bi = rn.im;
      rr += br * br + bi * bi;
      const cplx<R> a = rh[i];
      const double ar = a.re, ai = a.im;
      hr += ar * br + ai * bi;
      hi += ar * bi - ai * br;
    }
  }
  sm[0][threadIdx.x][0] = rr;
  sm[0][threadIdx.x][1] = 0.0;
  sm[1][threadIdx.x][0] = hr;
  sm[1][threadIdx.x][1] = hi;
  __syncthreads();
  // per-rhs sums of this block, threads in ascending order (as k_bdot_part)
  for (int o = threadIdx.x; o < 2 * nrhs; o += blockDim.x) {
    const int jj = o / nrhs, rq = o % nrhs;
    double sr = 0.0, si2 = 0.0;
    for (int k = 0; k < (int)blockDim.x; ++k) {
      if (s_rt[k] == rq) {
        sr += sm[jj][k][0];
        si2 += sm[jj][k][1];
      }
    }
    double* out = part + ((((size_t)jj * nrhs + rq) * gridDim.x) + blockIdx.x) * 2;
    out[0] = sr;
    out[1] = si2;
  }
}

// i -> (i / C, i % C) along a grid-stride walk, one 64-bit division at the start
struct DivWalk {
  size_t q, qstep;
  unsigned m, step, c;
  __device__ DivWalk(size_t i0, size_t T, int C)
      : q(i0 / (size_t)C), qstep(T / (size_t)C), m((unsigned)(i0 % (size_t)C)),
        step((unsigned)(T % (size_t)C)), c((unsigned)C) {}
  __device__ void next() {
    q += qstep;
    m += step;
    if (m >= c) {
      m -= c;
      ++q;
    }
  }
};

// a set of nrhs vectors [rhs][ld] (each sites x dof) <-> one batched vector
// [site][rhs][dof]; i walks the batched vector
template <typename R, bool UNPACK>
__global__ void k_bpack(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, size_t ld,
                        size_t n, int C, int dof) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  DivWalk dw(g, T, C);
  for (size_t i = g; i < n; i += T, dw.next()) {
    const unsigned q = dw.m / (unsigned)dof;
    const size_t f = (size_t)q * ld + dw.q * (size_t)dof + (dw.m - q * (unsigned)dof);
    if (UNPACK)
      out[f] = in[i];
    else
      out[i] = in[f];
  }
}

// y = x * f(r2[q].re) per rhs q with a device factor: mode 1 f = 1/sqrt
// (0 where r2 <= 0), mode 2 f = sqrt; precision changes from Rs to Rd
template <typename Rs, typename Rd>
__global__ void k_bconvert_dev(const cplx<Rs>* __restrict__ x, cplx<Rd>* __restrict__ y,
                               size_t n, int C, int dof, const cplx<double>* __restrict__ r2,
                               int mode) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const double q = r2[mw.m / (unsigned)dof].re;
    const double f = mode == 1 ? (q > 0.0 ? 1.0 / sqrt(q) : 0.0) : (q > 0.0 ? sqrt(q) : 0.0);
    y[i] = {(Rd)((double)x[i].re * f), (Rd)((double)x[i].im * f)};
  }
}

// x *= 1 / sqrt(r2[q].re) per rhs q in place (0 where r2 <= 0)
template <typename R>
__global__ void k_bscale_dev(cplx<R>* x, size_t n, int C, int dof,
                             const cplx<double>* __restrict__ r2) {
  const size_t T = (size_t)gridDim.x * blockDim.x;
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  ModWalk mw(g, T, C);
  for (size_t i = g; i < n; i += T, mw.next()) {
    const double q = r2[mw.m / (unsigned)dof].re;
    const double f = q > 0.0 ? 1.0 / sqrt(q) : 0.0;
    const cplx<R> v = x[i];
    x[i] = {(R)((double)v.re * f), (R)((double)v.im * f)};
  }
}

template <typename Rs, typename Rd>
void bconvert_dev(const cplx<Rs>* x, cplx<Rd>* y, const BShape& sh, const cplx<double>* d_r2,
                  int mode, hipStream_t s) {
  check_shape(sh);
  k_bconvert_dev<Rs, Rd><<<bew_grid(sh.n), 256, 0, s>>>(x, y, sh.n, sh.cols(), sh.dof, d_r2,
                                                        mode);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bscale_dev(cplx<R>* x, const BShape& sh, const cplx<double>* d_r2, hipStream_t s) {
  check_shape(sh);
  k_bscale_dev<R><<<bew_grid(sh.n), 256, 0, s>>>(x, sh.n, sh.cols(), sh.dof, d_r2);
  MGC_LAUNCH_CHECK();
}

static void upload_coef(BBlas& b, const cplx<double>* c, int rows, int nrhs) {
  MGC_HIP(hipEventSynchronize(b.coef_done));
  std::copy(c, c + (size_t)rows * nrhs, b.h_coef);
  MGC_HIP(hipMemcpyAsync(b.d_coef, b.h_coef, (size_t)rows * nrhs * sizeof(cplx<double>),
                         hipMemcpyHostToDevice, b.s));
  MGC_HIP(hipEventRecord(b.coef_done, b.s));
}

template <typename R>
void bbicg_p(BBlas& b, const cplx<double>* c, const cplx<R>* r, const cplx<R>* v, cplx<R>* p,
             const BShape& sh) {
  check_shape(sh);
  upload_coef(b, c, 3, sh.nrhs);
  k_bbicg_p<R><<<bew_grid(sh.n), 256, 0, b.s>>>(b.d_coef, r, v, p, sh.n, sh.cols(), sh.dof,
                                                sh.nrhs);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bbicg_s(BBlas& b, const cplx<double>* c, const cplx<R>* r, const cplx<R>* v, cplx<R>* s,
             const BShape& sh) {
  check_shape(sh);
  upload_coef(b, c, 3, sh.nrhs);
  k_bbicg_s<R><<<bew_grid(sh.n), 256, 0, b.s>>>(b.d_coef, r, v, s, sh.n, sh.cols(), sh.dof,
                                                sh.nrhs);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bbicg_xr(BBlas& b, const cplx<double>* c, const cplx<R>* p, const cplx<R>* s,
              const cplx<R>* t, cplx<R>* x, cplx<R>* r, const BShape& sh) {
  check_shape(sh);
  upload_coef(b, c, 3, sh.nrhs);
  k_bbicg_xr<R><<<bew_grid(sh.n), 256, 0, b.s>>>(b.d_coef, p, s, t, x, r, sh.n, sh.cols(),
                                                 sh.dof, sh.nrhs);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bbicg_xr_dots(BBlas& b, const cplx<double>* c, const cplx<R>* p, const cplx<R>* s,
                   const cplx<R>* t, const cplx<R>* rh, cplx<R>* x, cplx<R>* r,
                   const BShape& sh, cplx<double>* d_out) {
  check_shape(sh);
  upload_coef(b, c, 3, sh.nrhs);
  const int nblk = bred_blocks(sh, false);
  const size_t need = (size_t)2 * sh.nrhs * nblk * 2;
  if (need > b.part_cap) {
    if (b.d_part) MGC_HIP(hipFree(b.d_part));
    b.d_part = nullptr;
    b.part_cap = 0;
    MGC_HIP(hipMalloc(&b.d_part, need * sizeof(double)));
    b.part_cap = need;
  }
  k_bbicg_xr_dots<R><<<nblk, 256, 0, b.s>>>(b.d_coef, p, s, t, rh, x, r, sh.n, sh.cols(), sh.dof,
                                            sh.nrhs, b.d_part);
  MGC_LAUNCH_CHECK();
  k_bdot_final<<<2 * sh.nrhs, 256, 0, b.s>>>(b.d_part, nblk, reinterpret_cast<double*>(d_out));
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bpack(const cplx<R>* set, size_t ld, cplx<R>* out, const BShape& sh, hipStream_t s) {
  check_shape(sh);
  MGC_CHECK(ld * sh.nrhs == sh.n, MGC_ERR_SHAPE, "vector set does not match the batch");
  k_bpack<R, false><<<bew_grid(sh.n), 256, 0, s>>>(set, out, ld, sh.n, sh.cols(), sh.dof);
  MGC_LAUNCH_CHECK();
}

template <typename R>
void bunpack(const cplx<R>* in, cplx<R>* set, size_t ld, const BShape& sh, hipStream_t s) {
  check_shape(sh);
  MGC_CHECK(ld * sh.nrhs == sh.n, MGC_ERR_SHAPE, "vector set does not match the batch");
  k_bpack<R, true><<<bew_grid(sh.n), 256, 0, s>>>(in, set, ld, sh.n, sh.cols(), sh.dof);
  MGC_LAUNCH_CHECK();
}

#define INST(R)                                                                               \
  template void bdots_dev<R>(BBlas&, const cplx<R>*, size_t, int, const cplx<R>*,             \
                             const BShape&, bool, cplx<double>*);                             \
  template void bnorm2<R>(BBlas&, const cplx<R>*, const BShape&, double*);                    \
  template bool bmultisub_dots_dev<R>(BBlas&, const cplx<double>*, const cplx<R>*, size_t, int, \
                                      cplx<R>*, const BShape&, cplx<double>*, const int*);     \
  template void bmultisub_dev<R>(const cplx<double>*, const cplx<R>*, size_t, int, cplx<R>*,  \
                                 const BShape&, hipStream_t);                                 \
  template void bmultiaxpy<R>(BBlas&, const cplx<double>*, const cplx<R>*, size_t, int,       \
                              cplx<R>*, const BShape&);                                       \
  template void bsubnorm_dev<R>(cplx<double>*, const cplx<R>*, size_t, int, const cplx<R>*,   \
                                cplx<R>*, const BShape&, hipStream_t);
INST(double)
INST(float)
#undef INST
#define INST(R)                                                                               \
  template void bbicg_p<R>(BBlas&, const cplx<double>*, const cplx<R>*, const cplx<R>*,       \
                           cplx<R>*, const BShape&);                                          \
  template void bbicg_s<R>(BBlas&, const cplx<double>*, const cplx<R>*, const cplx<R>*,       \
                           cplx<R>*, const BShape&);                                          \
  template void bbicg_xr<R>(BBlas&, const cplx<double>*, const cplx<R>*, const cplx<R>*,      \
                            const cplx<R>*, cplx<R>*, cplx<R>*, const BShape&);               \
  template void bbicg_xr_dots<R>(BBlas&, const cplx<double>*, const cplx<R>*, const cplx<R>*, \
                                 const cplx<R>*, const cplx<R>*, cplx<R>*, cplx<R>*,           \
                                 const BShape&, cplx<double>*);                               \
  template void bscale_dev<R>(cplx<R>*, const BShape&, const cplx<double>*, hipStream_t);     \
  template void bpack<R>(const cplx<R>*, size_t, cplx<R>*, const BShape&, hipStream_t);       \
  template void bunpack<R>(const cplx<R>*, cplx<R>*, size_t, const BShape&, hipStream_t);
INST(double)
INST(float)
#undef INST
template void bconvert<double, float>(BBlas&, const cplx<double>*, cplx<float>*, const BShape&,
                                      const double*);
template void bconvert<float, double>(BBlas&, const cplx<float>*, cplx<double>*, const BShape&,
                                      const double*);
template void bconvert<double, double>(BBlas&, const cplx<double>*, cplx<double>*,
                                       const BShape&, const double*);
template void bconvert_dev<double, float>(const cplx<double>*, cplx<float>*, const BShape&,
                                          const cplx<double>*, int, hipStream_t);
template void bconvert_dev<float, double>(const cplx<float>*, cplx<double>*, const BShape&,
                                          const cplx<double>*, int, hipStream_t);
template void bconvert_dev<double, double>(const cplx<double>*, cplx<double>*, const BShape&,
                                           const cplx<double>*, int, hipStream_t);
template void bconvert<float, float>(BBlas&, const cplx<float>*, cplx<float>*, const BShape&,
                                     const double*);

}  // namespace mgc
