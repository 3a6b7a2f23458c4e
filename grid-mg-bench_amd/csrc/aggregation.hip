// Block transfer operators (Aggregation): basis import/export, block
// projection (restriction R), promotion (prolongation P), block modified
// Gram-Schmidt in FP64, the orthogonality check, and the masked projections
// of the coarse-operator build.
//
// Reference semantics (chirality-split form, Ns_c = 2):
//   include/AggregationKernelsRework.h:96-157   aggregateProjectFast
//   include/AggregationKernelsRework.h:207-264  aggregatePromoteFast
//   include/AggregationKernelsRework.h:266-341  aggregateOrthogonalise
//   include/AggregationRework.h:64-146          class Aggregation
//   include/CoarseningLookupTableRework.h:121-179 (block <-> fine site tables,
//                                               inner/outer masks of a point)
#include "aggregation.hpp"

#include <algorithm>
#include <vector>

namespace mgc {

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------
__global__ void k_import_basis(const cplx<double>* __restrict__ src, int nb, int Vf, int Ef,
                               size_t total, const int* __restrict__ blk_fine,
                               cplx<double>* __restrict__ dst) {
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < total; t += stride) {
    const int i = (int)(t % nb);
    size_t r = t / nb;
    const int e = (int)(r % Ef);
    r /= Ef;  // r = xc*Vb + xb
    const int sf = blk_fine[r];
    dst[t] = src[((size_t)i * Vf + sf) * Ef + e];
  }
}

__global__ void k_export_basis(const cplx<double>* __restrict__ src, int nb, int Vf, int Ef,
                               size_t total, const int* __restrict__ blk_fine,
                               cplx<double>* __restrict__ dst) {
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < total; t += stride) {
    const int i = (int)(t % nb);
    size_t r = t / nb;
    const int e = (int)(r % Ef);
    r /= Ef;
    const int sf = blk_fine[r];
    dst[((size_t)i * Vf + sf) * Ef + e] = src[t];
  }
}

// fine field [Vf][Ef] = chirality-l part of basis vector i (l = -1: both)
template <typename R>
__global__ void k_extract_basis(const cplx<double>* __restrict__ U, int nb, int Ef, int i, int l,
                                size_t total, const int* __restrict__ blk_fine,
                                cplx<R>* __restrict__ out) {
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < total; t += stride) {  // t = (xc*Vb + xb)*Ef + e
    const int e = (int)(t % Ef);
    const int sf = blk_fine[t / Ef];
    cplx<double> v = U[t * nb + i];
    if (l >= 0 && (2 * e >= Ef) != (l == 1)) v = {0.0, 0.0};
    out[(size_t)sf * Ef + e] = {(R)v.re, (R)v.im};
  }
}

// fine fields [Vf][nr][Ef] holding the doubled basis columns n = n0 .. n0+nr-1:
// column n is the chirality-(n / nb) part of basis vector n % nb
__global__ void k_extract_doubled(const cplx<double>* __restrict__ U, int nb, int Ef, int n0,
                                  int nr, size_t total, const int* __restrict__ blk_fine,
                                  cplx<double>* __restrict__ out) {
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < total; t += stride) {  // t = ((xc*Vb + xb)*nr + r)*Ef + e
    const int e = (int)(t % Ef);
    const size_t q = t / Ef;
    const int r = (int)(q % nr);
    const size_t rr = q / nr;  // panel row xc*Vb + xb
    const int n = n0 + r, l = n / nb, c = n % nb;
    cplx<double> v = U[(rr * Ef + e) * nb + c];
    if ((2 * e >= Ef) != (l == 1)) v = {0.0, 0.0};
    out[((size_t)blk_fine[rr] * nr + r) * Ef + e] = v;
  }
}

void agg_extract_doubled(mgc_aggregation* a, int n0, int nr, void* out, hipStream_t s) {
  MGC_CHECK(a && a->basis_set, MGC_ERR_STATE, "aggregation basis not set");
  MGC_CHECK(n0 >= 0 && nr >= 1 && n0 + nr <= 2 * a->nb, MGC_ERR_ARG, "basis columns out of range");
  const size_t total = (size_t)a->Vc * a->Vb * nr * a->Ef;
  k_extract_doubled<<<grid_for(total, 256, 16384), 256, 0, s>>>(
      a->basis64, a->nb, a->Ef, n0, nr, total, a->blk_fine, (cplx<double>*)out);
  MGC_LAUNCH_CHECK();
}

__global__ void k_basis_to_f32(const cplx<double>* __restrict__ src, cplx<float>* __restrict__ dst,
                               size_t n) {
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) dst[t] = {(float)src[t].re, (float)src[t].im};
}

const cplx<float>* agg_basis32(mgc_aggregation* a, hipStream_t s) {
  if (!a->basis32_valid) {
    const size_t n = (size_t)a->Vc * a->Vb * a->Ef * a->nb;
    if (!a->basis32) MGC_HIP(hipMalloc(&a->basis32, n * sizeof(cplx<float>)));
    k_basis_to_f32<<<grid_for(n, 256, 16384), 256, 0, s>>>(a->basis64, a->basis32, n);
    MGC_LAUNCH_CHECK();
    a->basis32_valid = true;
  }
  return a->basis32;
}

// Site mask of a block site: kind 0 = all, 1 = inner of point p, 2 = outer of
// point p (outer for +mu: x_mu mod b = b-1; for -mu: x_mu mod b = 0).
__device__ inline bool site_in_mask(const Lat4& blk, int xb, int kind, int p) {
  if (kind == 0 || p == 8) return kind != 2;
  int c[4];
  blk.coords(xb, c);
  const int mu = p & 3;
  const bool outer = c[mu] == (p < 4 ? blk.L[mu] - 1 : 0);
  return kind == 2 ? outer : !outer;
}

// ---------------------------------------------------------------------------
// project: c(xc)(rhs)(k, i) = sum_{xb in mask, e in k} conj(U(xc,xb,e,i)) f(fine(xc,xb))(rhs)(e)
// One workgroup per (coarse site, chunk of RC right-hand sides).  Thread
// (i = t % nb, g = t / nb) walks rows q = g, g+G, ... of the block panel:
// the nb threads of a group read one row's nb basis entries (contiguous) and
// the G groups consecutive rows, so a wavefront streams the panel coalesced;
// f(q) is a same-address broadcast load.  Group partial sums meet in LDS.
// When the coarse lattice has few sites, blockIdx.z splits each panel into
// gridDim.z row ranges of rows_per rows; the workgroups then write partial
// sums to part[((xc * gridDim.y + y) * gridDim.z + z)][RC][2][nb], which
// k_project_reduce adds in a fixed order.
// With W = 2 (FP32, even nb) a thread takes the two adjacent basis columns
// i = 2 i2, 2 i2 + 1 of each row in one 16-byte load, as the FP64 form reads
// one column in 16 bytes; each column's sum keeps the row order q, q + G, ...
// ---------------------------------------------------------------------------
constexpr int RC = 4;  // right-hand sides per workgroup

// W basis entries of one row from p (W = 2: FP32 only, p 16-byte aligned)
template <typename R, int W>
__device__ __forceinline__ void project_row(const cplx<R>* p, cplx<R> (&o)[W]) {
  if constexpr (W == 1) {
    o[0] = p[0];
  } else {
    static_assert(W == 2 && sizeof(R) == 4, "paired columns are an FP32 form");
    const float4 t = *reinterpret_cast<const float4*>(p);
    o[0] = {t.x, t.y};
    o[1] = {t.z, t.w};
  }
}

// RCT = right-hand sides a workgroup actually holds (1 for a single rhs)
template <typename R, int RCT, int W = 1>
__global__ __launch_bounds__(256) void k_project(const cplx<R>* __restrict__ U,
                                                 const cplx<R>* __restrict__ f,
                                                 cplx<R>* __restrict__ c,
                                                 const int* __restrict__ blk_fine, int nb, int Vb,
                                                 int Ef, int nrhs, Lat4 blk, int kind, int point,
                                                 int rows_per, cplx<R>* __restrict__ part) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* s_red = reinterpret_cast<cplx<R>*>(smem_raw);  // [G][RC][2][nb]
  const int xc = blockIdx.x;
  const int r0 = blockIdx.y * RC;
  const int nr = min(RC, nrhs - r0);
  const int tid = threadIdx.x;
  const int nw = nb / W;  // threads per row
  const int G = blockDim.x / nw;
  const int i = (tid % nw) * W, g = tid / nw;  // first column of this thread
  const int half = Ef / 2;
  cplx<R> acc[RCT][2][W];
#pragma unroll
  for (int r = 0; r < RCT; ++r)
#pragma unroll
    for (int w = 0; w < W; ++w) acc[r][0][w] = acc[r][1][w] = {R(0), R(0)};
  if (g < G) {
    const int rows = Vb * Ef;
    const cplx<R>* Ub = U + (size_t)xc * rows * nb;
    const int* bf = blk_fine + (size_t)xc * Vb;
    const int qbeg = blockIdx.z * rows_per, qend = min(rows, qbeg + rows_per);
    int q = qbeg + g;
    if (kind == 0) {
      // unmasked: PU rows per step, their loads issued together (rows are
      // still accumulated in the order q, q + G, q + 2G, ...)
      constexpr int PU = (sizeof(R) * W == 4 ? 8 : 4) * (RCT == 1 ? 2 : 1);
      for (; q + (PU - 1) * G < qend; q += PU * G) {
        cplx<R> uu[PU][W];
        const cplx<R>* fp[PU];
        bool lo[PU];
#pragma unroll
        for (int v = 0; v < PU; ++v) {
          const int qq = q + v * G;
          const int xb = qq / Ef, e = qq - xb * Ef;
          lo[v] = e < half;
          project_row<R, W>(Ub + (size_t)qq * nb + i, uu[v]);
          fp[v] = f + ((size_t)bf[xb] * nrhs + r0) * Ef + e;
        }
        cplx<R> xv[PU][RCT];
#pragma unroll
        for (int v = 0; v < PU; ++v)
#pragma unroll
          for (int r = 0; r < RCT; ++r) xv[v][r] = r < nr ? fp[v][(size_t)r * Ef] : cplx<R>{R(0), R(0)};
#pragma unroll
        for (int v = 0; v < PU; ++v)
#pragma unroll
          for (int r = 0; r < RCT; ++r)
            if (r < nr)
#pragma unroll
              for (int w = 0; w < W; ++w) {
                if (lo[v]) cmac_conj(acc[r][0][w], uu[v][w], xv[v][r]);
                else cmac_conj(acc[r][1][w], uu[v][w], xv[v][r]);
              }
      }
    }
    for (; q < qend; q += G) {
      const int xb = q / Ef, e = q - xb * Ef;
      if (kind != 0 && !site_in_mask(blk, xb, kind, point)) continue;
      const int k = e >= half;
      cplx<R> u[W];
      project_row<R, W>(Ub + (size_t)q * nb + i, u);
      const cplx<R>* fr = f + ((size_t)bf[xb] * nrhs + r0) * Ef + e;
#pragma unroll
      for (int r = 0; r < RCT; ++r)
        if (r < nr) {
          const cplx<R> x = fr[(size_t)r * Ef];
#pragma unroll
          for (int w = 0; w < W; ++w) {
            if (k == 0) cmac_conj(acc[r][0][w], u[w], x);
            else cmac_conj(acc[r][1][w], u[w], x);
          }
        }
    }
#pragma unroll
    for (int r = 0; r < RCT; ++r)
      if (r < nr)
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int w = 0; w < W; ++w)
            s_red[(((size_t)g * RC + r) * 2 + k) * nb + i + w] = acc[r][k][w];
  }
  __syncthreads();
  for (int o = tid; o < nr * 2 * nb; o += blockDim.x) {
    const int r = o / (2 * nb), kn = o - r * 2 * nb;  // kn = k*nb + i
    cplx<R> t = {R(0), R(0)};
    for (int gg = 0; gg < G; ++gg) t = cadd(t, s_red[((size_t)gg * RC + r) * 2 * nb + kn]);
    if (gridDim.z == 1)
      c[((size_t)xc * nrhs + r0 + r) * 2 * nb + kn] = t;
    else
      part[(((size_t)xc * gridDim.y + blockIdx.y) * gridDim.z + blockIdx.z) * RC * 2 * nb + o] = t;
  }
}

// c(xc)(y*RC + r)(kn) = sum over the S row ranges of the partial sums.  One
// workgroup per (xc, y): thread (o, zg) sums z = zg, zg + ZG, ... for output
// o, and the ZG group sums are added in order in LDS.
template <typename R>
__global__ __launch_bounds__(256) void k_project_reduce(const cplx<R>* __restrict__ part,
                                                        cplx<R>* __restrict__ c, int nchunk,
                                                        int S, int nrhs, int nb) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* s_acc = reinterpret_cast<cplx<R>*>(smem_raw);  // [ZG][RC*2*nb]
  const int no = RC * 2 * nb;
  const size_t xy = blockIdx.x;
  const int y = (int)(xy % nchunk), xc = (int)(xy / nchunk);
  if (no > (int)blockDim.x) {  // wide outputs: one thread per output, z in order
    for (int oo = threadIdx.x; oo < no; oo += blockDim.x) {
      cplx<R> t = {R(0), R(0)};
      for (int z = 0; z < S; ++z) t = cadd(t, part[(xy * S + z) * no + oo]);
      const int r = oo / (2 * nb), kn = oo - r * 2 * nb;
      if (y * RC + r < nrhs) c[((size_t)xc * nrhs + y * RC + r) * 2 * nb + kn] = t;
    }
    return;
  }
  const int ZG = (int)blockDim.x / no;
  const int o = threadIdx.x % no, zg = threadIdx.x / no;
  cplx<R> acc = {R(0), R(0)};
  if (zg < ZG)
    for (int z = zg; z < S; z += ZG) acc = cadd(acc, part[(xy * S + z) * no + o]);
  if (zg < ZG) s_acc[zg * no + o] = acc;
  __syncthreads();
  for (int oo = threadIdx.x; oo < no; oo += blockDim.x) {
    cplx<R> t = {R(0), R(0)};
    for (int g = 0; g < ZG; ++g) t = cadd(t, s_acc[g * no + oo]);
    const int r = oo / (2 * nb), kn = oo - r * 2 * nb;
    if (y * RC + r < nrhs) c[((size_t)xc * nrhs + y * RC + r) * 2 * nb + kn] = t;
  }
}

// ---------------------------------------------------------------------------
// promote: f(fine(xc,xb))(rhs)(e) = sum_i c(xc)(rhs)(chirality(e), i) U(xc,xb,e,i)
// One workgroup per (coarse site, chunk of RC right-hand sides).  The block's
// basis panel (rows q = (xb, e), nb entries each, row-major) is streamed in
// tiles of 256/LPR rows: each tile is read with contiguous loads
// (consecutive threads take consecutive panel entries), stored transposed in
// LDS ([i][row], one row of padding against bank conflicts), and each group of
// LPR adjacent lanes then sums one row, nb/LPR entries each, joined by lane
// exchanges.  LPR = 2, or 4 where a 128-row tile would not fit in 64 KB.
// ---------------------------------------------------------------------------
template <typename R, int LPR, int TPT>
__global__ __launch_bounds__(256) void k_promote(const cplx<R>* __restrict__ U,
                                                 const cplx<R>* __restrict__ c,
                                                 cplx<R>* __restrict__ f,
                                                 const int* __restrict__ blk_fine, int nb, int Vb,
                                                 int Ef, int nrhs, int rows_per) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* s_c = reinterpret_cast<cplx<R>*>(smem_raw);  // [RC][2][nb]
  constexpr int PR = 256 / LPR;                         // rows per tile
  constexpr int LD = PR + 1;
  cplx<R>* s_u = s_c + RC * 2 * nb;                     // [nb][LD]
  const int xc = blockIdx.x;
  const int r0 = blockIdx.y * RC;
  const int nr = min(RC, nrhs - r0);
  const int tid = threadIdx.x;
  const int half = Ef / 2;
  for (int o = tid; o < nr * 2 * nb; o += blockDim.x)
    s_c[o] = c[((size_t)xc * nrhs + r0) * 2 * nb + o];
  const int rows = Vb * Ef;
  const cplx<R>* Ub = U + (size_t)xc * rows * nb;
  const int* bf = blk_fine + (size_t)xc * Vb;
  const int row_l = tid / LPR, h = tid % LPR;
  const int nh = (nb + LPR - 1) / LPR;
  const int i0 = min(nb, h * nh), i1 = min(nb, i0 + nh);
  // per-thread walk of the tile load (t = tid + 256 k): row t / nb, entry t % nb
  const int step_r = 256 / nb, step_i = 256 % nb;
  // this workgroup's row range (blockIdx.z splits the panel on small lattices)
  const int qbeg = blockIdx.z * rows_per, qend = min(rows, qbeg + rows_per);
  // the tile's entries are loaded into registers one tile ahead, so the
  // next tile's loads are in flight while the current one is computed;
  // TPT >= ceil(PR * nb / 256) entries per thread
  // (real and imaginary parts in separate arrays: kept in registers)
  R pre_re[TPT], pre_im[TPT];
  auto fetch = [&](int qt) __attribute__((always_inline)) {
    const int nrow_t = min(PR, qend - qt);
    const cplx<R>* src = Ub + (size_t)qt * nb;
    // past the tile end the last entry is loaded again (never stored)
    const int last = nrow_t * nb - 1;
#pragma unroll
    for (int k2 = 0; k2 < TPT; ++k2) {
      const cplx<R> v = src[min(tid + 256 * k2, last)];
      pre_re[k2] = v.re;
      pre_im[k2] = v.im;
    }
  };
  if (qbeg < qend) fetch(qbeg);
  for (int q0 = qbeg; q0 < qend; q0 += PR) {
    const int nrow = min(PR, qend - q0);
    __syncthreads();  // the previous tile has been consumed (and s_c is visible)
    {
      int rl = tid / nb, i = tid - rl * nb;
#pragma unroll
      for (int k2 = 0; k2 < TPT; ++k2) {
        const int t = tid + 256 * k2;
        if (t < nrow * nb) s_u[i * LD + rl] = cplx<R>{pre_re[k2], pre_im[k2]};
        rl += step_r;
        i += step_i;
        if (i >= nb) {
          i -= nb;
          ++rl;
        }
      }
    }
    __syncthreads();
    // the next tile (after the last one, the current tile again: L2 hits,
    // never stored; keeps the prefetch unconditional)
    fetch(q0 + PR < qend ? q0 + PR : q0);
    cplx<R> acc[RC];
#pragma unroll
    for (int r = 0; r < RC; ++r) acc[r] = {R(0), R(0)};
    const bool active = row_l < nrow;
    int k = 0;
    if (active) {
      const int q = q0 + row_l;
      k = (q % Ef) >= half;
      for (int i = i0; i < i1; ++i) {
        const cplx<R> u = s_u[i * LD + row_l];
#pragma unroll
        for (int r = 0; r < RC; ++r)
          if (r < nr) cmac(acc[r], s_c[(r * 2 + k) * nb + i], u);
      }
    }
#pragma unroll
    for (int off = 1; off < LPR; off <<= 1)
#pragma unroll
      for (int r = 0; r < RC; ++r) {
        acc[r].re += __shfl_xor(acc[r].re, off);
        acc[r].im += __shfl_xor(acc[r].im, off);
      }
    if (active && h == 0) {
      const int q = q0 + row_l;
      const int xb = q / Ef, e = q - xb * Ef;
#pragma unroll
      for (int r = 0; r < RC; ++r)
        if (r < nr) f[((size_t)bf[xb] * nrhs + r0 + r) * Ef + e] = acc[r];
    }
  }
}

// ---------------------------------------------------------------------------
// block Gram-Schmidt, FP64.  One workgroup per (coarse site, chirality);
// thread t owns panel rows t, t+T, ... (at most RPT rows) and only ever
// touches its own rows, so the only communication is workgroup-wide
// reductions.  Each pass sweeps the columns v = 0 .. nb-1: all inner
// products <u, v> with the finished columns u < v are reduced together
// (classical Gram-Schmidt, one reduction of v values instead of v scalar
// reductions), v -= sum_u <u, v> u, and the projection is repeated when the
// norm fell below 1/sqrt(2) of its value before it (the
// Daniel-Gragg-Kaufman-Stewart test); then v /= |v|.  For a panel of full
// rank this gives the Q factor of the modified Gram-Schmidt sweep of the
// reference (the unique one with a positive real R diagonal) to rounding,
// with at least its orthogonality.
// ---------------------------------------------------------------------------
constexpr int MGS_THREADS = 512;
constexpr int MGS_RPT = 8;
constexpr int MGS_NBMAX = 64;

// per-wavefront sums of (a, b) for each of n values, lane 0 of wave w
// writing s_part[w][i]; the caller syncs and adds the wavefronts
__device__ inline void wave_sum2(double a, double b, double* out) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  if ((threadIdx.x & 63) == 0) {
    out[0] = a;
    out[1] = b;
  }
}

// flags != null: only the panels whose flag is set (the fallback of the
// Cholesky form below)
__global__ __launch_bounds__(MGS_THREADS) void k_orthogonalise(cplx<double>* __restrict__ U,
                                                                int nb, int Vb, int Ef,
                                                                int passes,
                                                                const int* __restrict__ flags) {
  constexpr int NW = MGS_THREADS / 64;
  // s_part[w][u][2]: wavefront partial sums; row MGS_NBMAX carries |v|^2
  __shared__ double s_part[NW][MGS_NBMAX + 1][2];
  __shared__ cplx<double> s_alpha[MGS_NBMAX];
  __shared__ double s_n2;
  if (flags && !flags[blockIdx.x]) return;  // uniform over the workgroup
  const int xc = blockIdx.x >> 1;
  const int k = blockIdx.x & 1;
  const int half = Ef / 2;
  const int Rk = Vb * half;
  const int w = threadIdx.x >> 6;
  size_t rowoff[MGS_RPT];
  int nrow = 0;
#pragma unroll
  for (int j = 0; j < MGS_RPT; ++j) {
    const int r = threadIdx.x + j * MGS_THREADS;
    if (r < Rk) nrow = j + 1;
    const int rr = r < Rk ? r : 0;
    const int xb = rr / half, e = half * k + rr % half;
    rowoff[j] = agg_row(nb, Vb, Ef, xc, xb, e);
  }
  // sum of this workgroup's s_part rows [0, cnt) into s_alpha (and row
  // MGS_NBMAX into s_n2); ends with a barrier
  auto finish = [&](int cnt, bool with_norm) {
    __syncthreads();
    for (int u = threadIdx.x; u < cnt; u += MGS_THREADS) {
      double re = 0.0, im = 0.0;
      for (int ww = 0; ww < NW; ++ww) {
        re += s_part[ww][u][0];
        im += s_part[ww][u][1];
      }
      s_alpha[u] = {re, im};
    }
    if (with_norm && threadIdx.x == MGS_THREADS - 1) {
      double n2 = 0.0;
      for (int ww = 0; ww < NW; ++ww) n2 += s_part[ww][MGS_NBMAX][0];
      s_n2 = n2;
    }
    __syncthreads();
  };
  for (int pass = 0; pass < passes; ++pass) {
    for (int v = 0; v < nb; ++v) {
      cplx<double> vr[MGS_RPT];
#pragma unroll
      for (int j = 0; j < MGS_RPT; ++j)
        vr[j] = j < nrow ? U[rowoff[j] + v] : cplx<double>{0.0, 0.0};
      double n2v = 0.0;  // |v|^2 before the projection
      {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < MGS_RPT; ++j) a += vr[j].re * vr[j].re + vr[j].im * vr[j].im;
        wave_sum2(a, 0.0, s_part[w][MGS_NBMAX]);
      }
      for (int rep = 0; rep < 2 && v > 0; ++rep) {
        // <u, v> for every finished column u < v, four columns' loads at a time
        for (int u0 = 0; u0 < v; u0 += 4) {
          cplx<double> d[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
          for (int j = 0; j < MGS_RPT; ++j)
            if (j < nrow) {
              cplx<double> uu[4];
#pragma unroll
              for (int q = 0; q < 4; ++q) uu[q] = U[rowoff[j] + min(u0 + q, v - 1)];
#pragma unroll
              for (int q = 0; q < 4; ++q) cmac_conj(d[q], uu[q], vr[j]);
            }
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (u0 + q < v) wave_sum2(d[q].re, d[q].im, s_part[w][u0 + q]);
        }
        finish(v, rep == 0);
        if (rep == 0) n2v = s_n2;
        // v -= sum_u <u, v> u, and |v|^2 after it
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < MGS_RPT; ++j) {
          if (j < nrow) {
            double xr = vr[j].re, xi = vr[j].im;
            for (int u0 = 0; u0 < v; u0 += 4) {
              cplx<double> uu[4];
#pragma unroll
              for (int q = 0; q < 4; ++q) uu[q] = U[rowoff[j] + min(u0 + q, v - 1)];
#pragma unroll
              for (int q = 0; q < 4; ++q)
                if (u0 + q < v) {
                  const cplx<double> al = s_alpha[u0 + q];
                  xr -= al.re * uu[q].re - al.im * uu[q].im;
                  xi -= al.re * uu[q].im + al.im * uu[q].re;
                }
            }
            vr[j] = {xr, xi};
          }
          a += vr[j].re * vr[j].re + vr[j].im * vr[j].im;
        }
        wave_sum2(a, 0.0, s_part[w][MGS_NBMAX]);
        finish(0, true);
        // one projection suffices unless the norm dropped below 1/sqrt(2)
        if (s_n2 >= 0.5 * n2v) break;
        n2v = s_n2;
      }
      if (v == 0) finish(0, true);
      const double sc = 1.0 / sqrt(s_n2);
#pragma unroll
      for (int j = 0; j < MGS_RPT; ++j)
        if (j < nrow) U[rowoff[j] + v] = {vr[j].re * sc, vr[j].im * sc};
      __syncthreads();  // column v is final before the next column reads it
    }
  }
}

// ---------------------------------------------------------------------------
// Block orthonormalisation by Cholesky QR in two rounds (CholQR2), one
// workgroup per (block, chirality) panel V (rows = block sites x Ef/2
// elements, nb columns):
//   read 1 (k_cholqr_gram, round 0):  G1 = V^H V,  R1 = chol(G1),  T1 = R1^-1
//   read 2 (k_cholqr_gram, round 1):  V1 = V T1 on the fly (not stored),
//                                     G2 = V1^H V1, R2 = chol(G2), T2 = R2^-1
//   read 3 (k_cholqr_apply):          V <- (V T1) T2  (in place, the only write)
// V1 is formed by the same code in rounds 1 and 3, so the second round
// corrects exactly the V1 that the apply produces (applying T1 T2 at once
// would lose that: the columns of V T1 carry errors of eps * kappa(V)).
// For a panel of full rank Q = V T is the unique QR factor with a positive
// real R diagonal, which is what the reference's modified Gram-Schmidt
// computes (include/AggregationKernelsRework.h:266-341), here to O(kappa eps).
// A panel whose Cholesky pivot drops below CHOLQR_PIVOT_FLOOR of its column
// norm (nearly dependent columns, kappa > ~1e7) is flagged, left untouched by
// the apply, and orthonormalised by the column sweep instead.
// T1 and T2 are nb x nb complex per panel in a workspace of the handle; the
// panel is streamed in chunks of CQ_CH rows through LDS, and the products run
// on FP64 MFMA (below).
// ---------------------------------------------------------------------------
// 1: always the column sweep (mgc_aggregation_set_ortho_method)
static int g_ortho_sweep_only = 0;
constexpr int CQ_CH = 32;         // panel rows per staged chunk
constexpr int CQ_NBMAX = 32;      // widest panel of the Cholesky form
constexpr double CHOLQR_PIVOT_FLOOR = 1e-14;  // on |pivot|^2 / G_jj

__device__ __forceinline__ size_t cq_row_off(int r, int nb, int Vb, int Ef, int xc, int k) {
  const int half = Ef / 2;
  return agg_row(nb, Vb, Ef, xc, r / half, half * k + r % half);
}

// Cholesky G = R^H R (upper R, real positive diagonal) of the nb x nb matrix
// in sG (overwritten by R), and sX <- R^-1 (upper triangular).  Run by the
// whole workgroup; the factorisation itself by wavefront 0.
template <int NS>
__device__ void cq_factor(cplx<double> (*sG)[NS], cplx<double> (*sX)[NS],
                          int nb, int* flag) {
  const int t = threadIdx.x;
  __shared__ int s_bad;
  if (t == 0) s_bad = 0;
  __syncthreads();
  if (t < 64) {
    // right-looking factorisation, lane k owns column k
    for (int j = 0; j < nb; ++j) {
      const double gjj = sG[j][j].re;
      double d = gjj;
      for (int i = 0; i < j; ++i) d -= sG[i][j].re * sG[i][j].re + sG[i][j].im * sG[i][j].im;
      const bool bad = !(d > CHOLQR_PIVOT_FLOOR * gjj) || !(gjj > 0.0);
      const double rjj = bad ? 1.0 : sqrt(d);
      if (bad && t == 0) s_bad = 1;
      // row j of R: R[j][k] = (G[j][k] - sum_{i<j} conj(R[i][j]) R[i][k]) / R[j][j]
      for (int k = j + 1 + t; k < nb; k += 64) {
        cplx<double> a = sG[j][k];
        for (int i = 0; i < j; ++i) {
          const cplx<double> rij = sG[i][j], rik = sG[i][k];
          a.re -= rij.re * rik.re + rij.im * rik.im;
          a.im -= rij.re * rik.im - rij.im * rik.re;
        }
        sG[j][k] = {a.re / rjj, a.im / rjj};
      }
      // one wavefront: the fences keep its LDS writes ahead of the next reads
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      if (t == 0) sG[j][j] = {rjj, 0.0};
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    // X = R^-1 (upper), lane j owns column j, from the diagonal upwards
    for (int j = t; j < nb; j += 64) {
      for (int i = nb - 1; i > j; --i) sX[i][j] = {0.0, 0.0};
      sX[j][j] = {1.0 / sG[j][j].re, 0.0};
      for (int i = j - 1; i >= 0; --i) {
        cplx<double> a = {0.0, 0.0};
        for (int k = i + 1; k <= j; ++k) cmac(a, sG[i][k], sX[k][j]);
        const double inv = 1.0 / sG[i][i].re;
        sX[i][j] = {-a.re * inv, -a.im * inv};
      }
    }
  }
  __syncthreads();
  if (s_bad && t == 0) *flag = 1;
}

// The Gram sums and the products V T run on FP64 MFMA (v_mfma_f64_16x16x4_f64)
// over chunks of CQ_CH panel rows staged in LDS as interleaved doubles
// (row m: Re v(m,0), Im v(m,0), Re v(m,1), ...), the columns zero-padded to
// NP = 16 or 32.  A complex product is one real GEMM with the complex index
// folded into K (K = 2i + t, t = 0: real, 1: imaginary part):
//   X = V T:    A'[m][2i+t] = the stored row itself,
//               B'[2i+t][2j+u] = (t,u) (0,0): Re T, (1,0): -Im T, (0,1): Im T, (1,1): Re T
//   G = V^H V:  A'[c][2m+t] = t ? -Im v(m,c) : Re v(m,c),  B'[2m+t][2c'+u] as B' above with v
// so the FP64 accumulator holds Re and Im of the result interleaved along the
// columns, in the layout of a stored row.  Lane layout of the 16x16x4 f64
// form: A[i = l%16][k = l/16], B[k = l/16][j = l%16], D[i = 4r + l/16][j = l%16].
typedef double cq_dx4 __attribute__((ext_vector_type(4)));

template <int NP>
struct CqShape {
  static constexpr int MT = NP / 16;            // 16-row tiles of the Gram
  static constexpr int NTL = NP / 8;            // 16-column tiles (2 NP real columns)
  static constexpr int WAVES = NP == 32 ? 4 : 2;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int GTPW = MT * NTL / WAVES;  // Gram tiles per wavefront
  static constexpr int PER = CQ_CH * NP / THREADS;  // staged complex per thread and chunk
  static constexpr int LDP = 2 * NP + 2;  // LDS row pitch in doubles (conflict-free A reads)
  static constexpr int NS = NP + 1;       // row stride (complex) of the factorisation's matrices
};

// chunk rows r0 .. r0+CQ_CH-1 of panel (xc, k) into registers (zero past the
// panel's rows and columns)
template <int NP>
__device__ __forceinline__ void cq_fetch(const cplx<double>* __restrict__ U, cplx<double>* reg,
                                         int r0, int rows, int nb, int Vb, int Ef, int xc, int k) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < CqShape<NP>::PER; ++q) {
    const int idx = t + q * CqShape<NP>::THREADS, r = idx / NP, i = idx % NP;
    reg[q] = (r0 + r < rows && i < nb) ? U[cq_row_off(r0 + r, nb, Vb, Ef, xc, k) + i]
                                       : cplx<double>{0.0, 0.0};
  }
}
template <int NP>
__device__ __forceinline__ void cq_stage(double* sV, const cplx<double>* reg) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < CqShape<NP>::PER; ++q) {
    const int idx = t + q * CqShape<NP>::THREADS, r = idx / NP, i = idx % NP;
    sV[r * CqShape<NP>::LDP + 2 * i] = reg[q].re;
    sV[r * CqShape<NP>::LDP + 2 * i + 1] = reg[q].im;
  }
}

// T (nb x nb complex, row-major, upper triangular) into LDS as interleaved
// doubles, zero-padded to NP x NP
template <int NP>
__device__ __forceinline__ void cq_load_T(double* sT, const cplx<double>* __restrict__ T, int nb) {
  for (int idx = threadIdx.x; idx < NP * NP; idx += CqShape<NP>::THREADS) {
    const int i = idx / NP, j = idx % NP;
    const cplx<double> v = (i < nb && j < nb) ? T[i * nb + j] : cplx<double>{0.0, 0.0};
    sT[i * CqShape<NP>::LDP + 2 * j] = v.re;
    sT[i * CqShape<NP>::LDP + 2 * j + 1] = v.im;
  }
}

// dst = src T for the CQ_CH staged rows (T upper triangular: output column
// tile nt needs K steps 0 .. 4 nt + 3 only; the skipped terms are exact
// zeros).  Wavefront w computes row tile w / (WAVES / 2) and two column
// tiles chosen so that every wavefront runs the same number of steps.
template <int NP>
__device__ __forceinline__ void cq_times_T(const double* src, const double* sT, double* dst) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lk = l >> 4, t = lk & 1;
  int mt, nts[2];
  if (NP == 32) {
    mt = w >> 1;
    nts[0] = (w & 1) ? 1 : 0;
    nts[1] = (w & 1) ? 2 : 3;
  } else {
    mt = w;
    nts[0] = 0;
    nts[1] = 1;
  }
  const int m = mt * 16 + (l & 15);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int nt = nts[q];
    const int n = nt * 16 + (l & 15), cq = n >> 1, u = n & 1;
    const double bsign = (t == 1 && u == 0) ? -1.0 : 1.0;
    cq_dx4 acc = {0.0, 0.0, 0.0, 0.0};
    const int ksteps = 4 * (nt + 1);
    for (int ks = 0; ks < ksteps; ++ks) {
      const double a = src[m * CqShape<NP>::LDP + 4 * ks + lk];
      const int i = 2 * ks + (lk >> 1);
      const double b = bsign * sT[i * CqShape<NP>::LDP + 2 * cq + (t ^ u)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) dst[(mt * 16 + 4 * r + lk) * CqShape<NP>::LDP + n] = acc[r];
  }
}

// acc += src^H src over the CQ_CH staged rows, for this wavefront's Gram tiles
// (tiles wholly below the diagonal are skipped: the factorisation reads the
// upper triangle only)
template <int NP>
__device__ __forceinline__ void cq_gram_acc(const double* src, cq_dx4* acc) {
  typedef CqShape<NP> S;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lk = l >> 4, t = lk & 1;
  const double asign = t ? -1.0 : 1.0;
#pragma unroll
  for (int q = 0; q < S::GTPW; ++q) {
    const int tile = w * S::GTPW + q, mt = tile / S::NTL, nt = tile % S::NTL;
    if (mt * 16 >= (nt + 1) * 8) continue;  // every c > every c' of the tile
    const int c = mt * 16 + (l & 15);
    const int n = nt * 16 + (l & 15), cq = n >> 1, u = n & 1;
    const double bsign = (t == 1 && u == 0) ? -1.0 : 1.0;
#pragma unroll 4
    for (int ks = 0; ks < CQ_CH / 2; ++ks) {
      const int m = 2 * ks + (lk >> 1);
      const double a = asign * src[m * S::LDP + 2 * c + t];
      const double b = bsign * src[m * S::LDP + 2 * cq + (t ^ u)];
      acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
    }
  }
}

// One Gram round over every panel (blockIdx.x = 2 * coarse site + chirality),
// writing R^-1 of the round to T[round].  Round 1 skips flagged panels and
// forms the rows of V T1 (cq_times_T, the same code as the apply) before the
// Gram sum.  The next chunk's loads are issued before the current chunk's
// products.
template <int NP>
__global__ __launch_bounds__(CqShape<NP>::THREADS) void k_cholqr_gram(
    const cplx<double>* __restrict__ U, cplx<double>* __restrict__ Tg, int* __restrict__ flags,
    int nb, int Vb, int Ef, int round) {
  typedef CqShape<NP> S;
  static_assert(CQ_CH * S::LDP >= 2 * NP * S::NS, "factorisation matrices fit the staging");
  __shared__ double sV[CQ_CH * S::LDP];  // staged rows; also R^-1 of the factorisation
  __shared__ double sW[CQ_CH * S::LDP];  // V T1; also the Gram matrix
  __shared__ double sT[NP * S::LDP];
  const int panel = blockIdx.x;
  const int xc = panel >> 1, k = panel & 1;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (round == 1 && flags[panel]) return;  // uniform over the workgroup
  const size_t tstride = (size_t)gridDim.x * nb * nb;  // T1 block, then T2 block
  if (round == 1) cq_load_T<NP>(sT, Tg + (size_t)panel * nb * nb, nb);
  const int rows = Vb * (Ef / 2);
  cq_dx4 acc[S::GTPW];
#pragma unroll
  for (int q = 0; q < S::GTPW; ++q) acc[q] = {0.0, 0.0, 0.0, 0.0};
  // two register sets: the loads of chunk c + 2 are issued as chunk c is
  // staged, so each chunk's loads have two chunks of products to arrive
  cplx<double> ra[S::PER], rb[S::PER];
  cq_fetch<NP>(U, ra, 0, rows, nb, Vb, Ef, xc, k);
  cq_fetch<NP>(U, rb, CQ_CH, rows, nb, Vb, Ef, xc, k);
  auto chunk = [&](int r0, cplx<double>* reg) {
    __syncthreads();  // previous chunk consumed
    cq_stage<NP>(sV, reg);
    __syncthreads();
    if (r0 + 2 * CQ_CH < rows) cq_fetch<NP>(U, reg, r0 + 2 * CQ_CH, rows, nb, Vb, Ef, xc, k);
    const double* src = sV;
    if (round == 1) {
      cq_times_T<NP>(sV, sT, sW);
      __syncthreads();
      src = sW;
    }
    cq_gram_acc<NP>(src, acc);
  };
  for (int r0 = 0; r0 < rows; r0 += 2 * CQ_CH) {
    chunk(r0, ra);
    if (r0 + CQ_CH < rows) chunk(r0 + CQ_CH, rb);
  }
  __syncthreads();
  auto sG = reinterpret_cast<cplx<double>(*)[S::NS]>(sW);
#pragma unroll
  for (int q = 0; q < S::GTPW; ++q) {
    const int tile = w * S::GTPW + q, mt = tile / S::NTL, nt = tile % S::NTL;
    const int n = nt * 16 + (l & 15), cq = n >> 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = mt * 16 + 4 * r + (l >> 4);
      if (c < nb && cq < nb) reinterpret_cast<double*>(&sG[c][cq])[n & 1] = acc[q][r];
    }
  }
  __syncthreads();
  auto sX = reinterpret_cast<cplx<double>(*)[S::NS]>(sV);
  cq_factor(sG, sX, nb, &flags[panel]);
  cplx<double>* Tp = Tg + round * tstride + (size_t)panel * nb * nb;
  for (int idx = threadIdx.x; idx < nb * nb; idx += S::THREADS) Tp[idx] = sX[idx / nb][idx % nb];
}

// V <- (V T1) T2 in place on every panel that is not flagged
template <int NP>
__global__ __launch_bounds__(CqShape<NP>::THREADS) void k_cholqr_apply(
    cplx<double>* __restrict__ U, const cplx<double>* __restrict__ Tg,
    const int* __restrict__ flags, int nb, int Vb, int Ef) {
  typedef CqShape<NP> S;
  __shared__ double sV[CQ_CH * S::LDP];
  __shared__ double sW[CQ_CH * S::LDP];
  __shared__ double sT[NP * S::LDP];
  __shared__ double sT2[NP * S::LDP];
  const int panel = blockIdx.x;
  const int xc = panel >> 1, k = panel & 1;
  const int t = threadIdx.x;
  if (flags[panel]) return;
  const cplx<double>* T1 = Tg + (size_t)panel * nb * nb;
  cq_load_T<NP>(sT, T1, nb);
  cq_load_T<NP>(sT2, T1 + (size_t)gridDim.x * nb * nb, nb);
  const int rows = Vb * (Ef / 2);
  // loads two chunks ahead, as the Gram rounds (every chunk's rows are read
  // before any of them is written back, so the in-place update is safe)
  cplx<double> ra[S::PER], rb[S::PER];
  cq_fetch<NP>(U, ra, 0, rows, nb, Vb, Ef, xc, k);
  cq_fetch<NP>(U, rb, CQ_CH, rows, nb, Vb, Ef, xc, k);
  auto chunk = [&](int r0, cplx<double>* reg) {
    __syncthreads();  // previous chunk stored
    cq_stage<NP>(sV, reg);
    __syncthreads();
    if (r0 + 2 * CQ_CH < rows) cq_fetch<NP>(U, reg, r0 + 2 * CQ_CH, rows, nb, Vb, Ef, xc, k);
    cq_times_T<NP>(sV, sT, sW);
    __syncthreads();
    cq_times_T<NP>(sW, sT2, sV);
    __syncthreads();
    // the chunk's rows back in place, whole complex values per thread
    for (int idx = t; idx < CQ_CH * nb; idx += S::THREADS) {
      const int r = idx / nb, i = idx % nb;
      if (r0 + r < rows)
        U[cq_row_off(r0 + r, nb, Vb, Ef, xc, k) + i] = {sV[r * CqShape<NP>::LDP + 2 * i],
                                                        sV[r * CqShape<NP>::LDP + 2 * i + 1]};
    }
  };
  for (int r0 = 0; r0 < rows; r0 += 2 * CQ_CH) {
    chunk(r0, ra);
    if (r0 + CQ_CH < rows) chunk(r0 + CQ_CH, rb);
  }
}

// errs[i] += sum_{k, j} |delta_ij - <U_j, U_i>_{block,k}|^2 (CheckOrthogonal)
__global__ __launch_bounds__(256) void k_check_orthogonal(const cplx<double>* __restrict__ U,
                                                          int nb, int Vb, int Ef,
                                                          double* __restrict__ errs) {
  const int xc = blockIdx.x >> 1;
  const int k = blockIdx.x & 1;
  const int half = Ef / 2;
  for (int pair = threadIdx.x; pair < nb * nb; pair += blockDim.x) {
    const int i = pair / nb, j = pair % nb;
    cplx<double> g = {0.0, 0.0};
    for (int r = 0; r < Vb * half; ++r) {
      const size_t off = agg_row(nb, Vb, Ef, xc, r / half, half * k + r % half);
      cmac_conj(g, U[off + j], U[off + i]);
    }
    const double dr = (i == j ? 1.0 : 0.0) - g.re;
    atomicAdd(&errs[i], dr * dr + g.im * g.im);
  }
}

static void free_agg(mgc_aggregation* a) {
  if (!a) return;
  void* bufs[] = {a->basis64, a->basis32, a->blk_fine, a->fine_row, a->part, a->cq_T,
                  a->cq_flag};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete a;
}

static void check_agg(const mgc_aggregation* a) {
  MGC_CHECK(a, MGC_ERR_ARG, "null aggregation handle");
  MGC_CHECK(a->basis_set, MGC_ERR_STATE, "basis not set");
}

void agg_transfer(mgc_aggregation* a, const void* in, void* out, int prec, int nrhs,
                  bool project, int kind, int point, hipStream_t s) {
  check_agg(a);
  MGC_CHECK(in && out && in != out, MGC_ERR_ARG, "null or aliased vector");
  MGC_CHECK(nrhs >= 1, MGC_ERR_ARG, "nrhs < 1");
  const int nchunk = (nrhs + RC - 1) / RC;
  const int nb = a->nb;
  const int rows = a->Vb * a->Ef;
  // split each panel over S workgroups when the coarse lattice alone gives
  // fewer than ~1024 workgroups (at least 64 rows per workgroup)
  int S = 1;
  if ((size_t)a->Vc * nchunk < 1024) {
    S = (int)((1024 + (size_t)a->Vc * nchunk - 1) / ((size_t)a->Vc * nchunk));
    S = std::min(S, std::max(1, rows / 64));
  }
  const int rows_per = (rows + S - 1) / S;
  S = (rows + rows_per - 1) / rows_per;
  const dim3 grid(a->Vc, nchunk, S);
  const size_t es = prec == MGC_F64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
  if (project) {
    MGC_CHECK(prec == MGC_F64 || prec == MGC_F32, MGC_ERR_PRECISION,
              "project: F64 or F32 vectors");
    const int block = (256 / nb) * nb;
    const size_t sh = (size_t)(block / nb) * RC * 2 * nb * es;
    void* part = nullptr;
    if (S > 1) {
      const size_t need = (size_t)a->Vc * nchunk * S * RC * 2 * nb * es;
      if (need > a->part_bytes) {
        if (a->part) MGC_HIP(hipFree(a->part));
        a->part = nullptr;
        a->part_bytes = 0;
        MGC_HIP(hipMalloc(&a->part, need));
        a->part_bytes = need;
      }
      part = a->part;
    }
    const int nred = a->Vc * nchunk;  // reduction workgroups
    const size_t red_sh = (size_t)256 * es;  // ZG * (RC*2*nb) <= 256 entries
    if (prec == MGC_F64) {
      if (nrhs == 1)
        k_project<double, 1><<<grid, block, sh, s>>>(a->basis64, (const cplx<double>*)in,
                                                     (cplx<double>*)out, a->blk_fine, nb, a->Vb,
                                                     a->Ef, nrhs, a->blk, kind, point, rows_per,
                                                     (cplx<double>*)part);
      else
        k_project<double, RC><<<grid, block, sh, s>>>(a->basis64, (const cplx<double>*)in,
                                                      (cplx<double>*)out, a->blk_fine, nb, a->Vb,
                                                      a->Ef, nrhs, a->blk, kind, point, rows_per,
                                                      (cplx<double>*)part);
      if (S > 1)
        k_project_reduce<double><<<nred, 256, red_sh, s>>>((const cplx<double>*)part,
                                                            (cplx<double>*)out, nchunk, S, nrhs,
                                                            nb);
    } else {
      const cplx<float>* b32 = agg_basis32(a, s);
      // even nb: two adjacent columns per thread in one 16-byte load
      const bool paired = nb % 2 == 0;
      const int block2 = (256 / (nb / 2 > 0 ? nb / 2 : 1)) * (nb / 2);
      const size_t sh2 = (size_t)(paired ? block2 / (nb / 2) : 0) * RC * 2 * nb * es;
      if (paired && nrhs == 1)
        k_project<float, 1, 2><<<grid, block2, sh2, s>>>(b32, (const cplx<float>*)in,
                                                         (cplx<float>*)out, a->blk_fine, nb,
                                                         a->Vb, a->Ef, nrhs, a->blk, kind, point,
                                                         rows_per, (cplx<float>*)part);
      else if (paired)
        k_project<float, RC, 2><<<grid, block2, sh2, s>>>(b32, (const cplx<float>*)in,
                                                          (cplx<float>*)out, a->blk_fine, nb,
                                                          a->Vb, a->Ef, nrhs, a->blk, kind, point,
                                                          rows_per, (cplx<float>*)part);
      else if (nrhs == 1)
        k_project<float, 1><<<grid, block, sh, s>>>(b32, (const cplx<float>*)in,
                                                    (cplx<float>*)out, a->blk_fine, nb, a->Vb,
                                                    a->Ef, nrhs, a->blk, kind, point, rows_per,
                                                    (cplx<float>*)part);
      else
        k_project<float, RC><<<grid, block, sh, s>>>(b32, (const cplx<float>*)in,
                                                     (cplx<float>*)out, a->blk_fine, nb, a->Vb,
                                                     a->Ef, nrhs, a->blk, kind, point, rows_per,
                                                     (cplx<float>*)part);
      if (S > 1)
        k_project_reduce<float><<<nred, 256, red_sh, s>>>((const cplx<float>*)part,
                                                           (cplx<float>*)out, nchunk, S, nrhs,
                                                           nb);
    }
  } else {
    auto lds = [&](int lpr) { return ((size_t)RC * 2 * nb + (size_t)nb * (256 / lpr + 1)) * es; };
    // FP32 with nb <= 32: one lane per row, 256-row tiles (half the barriers
    // per panel); otherwise two (or four) lanes per row
    const int lpr = prec == MGC_F32 && nb <= 32 && lds(1) <= 65536 ? 1
                    : lds(2) <= 65536                               ? 2
                                                                    : 4;
    const size_t sh = lds(lpr);
    MGC_CHECK(sh <= 65536, MGC_ERR_SHAPE, "promote: nbasis too large for the LDS tile");
    // register prefetch depth: entries per thread of one tile, rounded up to a power of 2
    const int need = (256 / lpr * nb + 255) / 256;
    const int tpt = need <= 2 ? 2 : need <= 4 ? 4 : need <= 8 ? 8 : need <= 16 ? 16 : 32;
#define PROMOTE_T(R, LPR, T, basis)                                                            \
  k_promote<R, LPR, T><<<grid, 256, sh, s>>>(basis, (const cplx<R>*)in, (cplx<R>*)out,          \
                                             a->blk_fine, nb, a->Vb, a->Ef, nrhs, rows_per)
#define PROMOTE(R, LPR, basis)                           \
  switch (tpt) {                                         \
    case 2: PROMOTE_T(R, LPR, 2, basis); break;          \
    case 4: PROMOTE_T(R, LPR, 4, basis); break;          \
    case 8: PROMOTE_T(R, LPR, 8, basis); break;          \
    case 16: PROMOTE_T(R, LPR, 16, basis); break;        \
    default: PROMOTE_T(R, LPR, 32, basis); break;        \
  }
    if (prec == MGC_F64) {
      if (lpr == 2) {
        PROMOTE(double, 2, a->basis64)
      } else {
        PROMOTE(double, 4, a->basis64)
      }
    } else if (prec == MGC_F32) {
      if (lpr == 1) {
        PROMOTE(float, 1, agg_basis32(a, s))
      } else if (lpr == 2) {
        PROMOTE(float, 2, agg_basis32(a, s))
      } else {
        PROMOTE(float, 4, agg_basis32(a, s))
      }
    } else {
      throw Error(MGC_ERR_PRECISION, "promote: F64 or F32 vectors");
    }
#undef PROMOTE
#undef PROMOTE_T
  }
  MGC_LAUNCH_CHECK();
}

void agg_extract_basis(mgc_aggregation* a, int i, int l, void* out, int prec, hipStream_t s) {
  check_agg(a);
  const size_t total = (size_t)a->Vc * a->Vb * a->Ef;
  const unsigned g = grid_for(total, 256, 16384);
  if (prec == MGC_F64)
    k_extract_basis<double><<<g, 256, 0, s>>>(a->basis64, a->nb, a->Ef, i, l, total, a->blk_fine,
                                              (cplx<double>*)out);
  else
    k_extract_basis<float><<<g, 256, 0, s>>>(a->basis64, a->nb, a->Ef, i, l, total, a->blk_fine,
                                             (cplx<float>*)out);
  MGC_LAUNCH_CHECK();
}

}  // namespace mgc

using namespace mgc;

extern "C" {

int mgc_aggregation_create_ex(mgc_aggregation** out, const int fine_dims[4], const int block[4],
                              int nb, int site_elems) {
  mgc_aggregation* a = nullptr;
  int st = guarded([&] {
    MGC_CHECK(out && fine_dims && block, MGC_ERR_ARG, "null argument");
    MGC_CHECK(nb >= 1 && nb <= 64, MGC_ERR_SHAPE, "nb must be in 1..64");
    MGC_CHECK(site_elems >= 2 && site_elems % 2 == 0 && site_elems <= 256, MGC_ERR_SHAPE,
              "site elements must be even (two chiralities) and <= 256");
    a = new mgc_aggregation();
    a->Ef = site_elems;
    for (int d = 0; d < 4; ++d) {
      MGC_CHECK(fine_dims[d] >= 1 && block[d] >= 1, MGC_ERR_SHAPE, "extent < 1");
      MGC_CHECK(fine_dims[d] % block[d] == 0, MGC_ERR_SHAPE,
                "block size must divide the fine lattice (subdivides)");
      a->fine.L[d] = fine_dims[d];
      a->blk.L[d] = block[d];
      a->coarse.L[d] = fine_dims[d] / block[d];
    }
    a->Vf = a->fine.volume();
    a->Vc = a->coarse.volume();
    a->Vb = a->blk.volume();
    a->nb = nb;
    std::vector<int> bf((size_t)a->Vc * a->Vb);
    for (int xc = 0; xc < a->Vc; ++xc) {
      int cc[4];
      a->coarse.coords(xc, cc);
      for (int xb = 0; xb < a->Vb; ++xb) {
        int cb[4], cf[4];
        a->blk.coords(xb, cb);
        for (int d = 0; d < 4; ++d) cf[d] = cc[d] * block[d] + cb[d];
        bf[(size_t)xc * a->Vb + xb] = a->fine.index(cf);
      }
    }
    std::vector<int> fr((size_t)a->Vf);
    for (size_t r = 0; r < bf.size(); ++r) fr[bf[r]] = (int)r;
    MGC_HIP(hipMalloc(&a->blk_fine, bf.size() * sizeof(int)));
    MGC_HIP(hipMemcpy(a->blk_fine, bf.data(), bf.size() * sizeof(int), hipMemcpyHostToDevice));
    MGC_HIP(hipMalloc(&a->fine_row, fr.size() * sizeof(int)));
    MGC_HIP(hipMemcpy(a->fine_row, fr.data(), fr.size() * sizeof(int), hipMemcpyHostToDevice));
    MGC_HIP(hipMalloc(&a->basis64, (size_t)a->Vf * a->Ef * nb * sizeof(cplx<double>)));
    *out = a;
  });
  if (st != MGC_OK) free_agg(a);
  return st;
}

int mgc_aggregation_create(mgc_aggregation** out, const int fine_dims[4], const int block[4],
                           int nb) {
  return mgc_aggregation_create_ex(out, fine_dims, block, nb, 12);
}

int mgc_aggregation_destroy(mgc_aggregation* a) {
  return guarded([&] { free_agg(a); });
}

int mgc_aggregation_info(const mgc_aggregation* a, int coarse_dims[4], int* nb, int* fine_volume,
                         int* coarse_volume) {
  return guarded([&] {
    MGC_CHECK(a, MGC_ERR_ARG, "null aggregation handle");
    if (coarse_dims)
      for (int d = 0; d < 4; ++d) coarse_dims[d] = a->coarse.L[d];
    if (nb) *nb = a->nb;
    if (fine_volume) *fine_volume = a->Vf;
    if (coarse_volume) *coarse_volume = a->Vc;
  });
}

int mgc_aggregation_set_basis(mgc_aggregation* a, const void* basis_f64, void* stream) {
  return guarded([&] {
    MGC_CHECK(a && basis_f64, MGC_ERR_ARG, "null argument");
    const size_t total = (size_t)a->Vf * a->Ef * a->nb;
    hipStream_t s = as_stream(stream);
    k_import_basis<<<grid_for(total, 256, 16384), 256, 0, s>>>(
        (const cplx<double>*)basis_f64, a->nb, a->Vf, a->Ef, total, a->blk_fine, a->basis64);
    MGC_LAUNCH_CHECK();
    a->basis_set = true;
    a->basis32_valid = false;
  });
}

int mgc_aggregation_get_basis(const mgc_aggregation* a, void* basis_f64, void* stream) {
  return guarded([&] {
    check_agg(a);
    MGC_CHECK(basis_f64, MGC_ERR_ARG, "null output");
    const size_t total = (size_t)a->Vf * a->Ef * a->nb;
    k_export_basis<<<grid_for(total, 256, 16384), 256, 0, as_stream(stream)>>>(
        a->basis64, a->nb, a->Vf, a->Ef, total, a->blk_fine, (cplx<double>*)basis_f64);
    MGC_LAUNCH_CHECK();
  });
}

int mgc_aggregation_orthogonalise(mgc_aggregation* a, int passes, void* stream) {
  return guarded([&] {
    check_agg(a);
    MGC_CHECK(passes >= 0, MGC_ERR_ARG, "passes < 0");
    MGC_CHECK(a->Vb * a->Ef / 2 <= MGS_THREADS * MGS_RPT, MGC_ERR_SHAPE,
              "block too large for the in-register Gram-Schmidt (|block| * site elements / 2 "
              "<= 4096)");
    if (passes == 0) return;
    hipStream_t s = as_stream(stream);
    const int panels = 2 * a->Vc;
    if (a->nb > CQ_NBMAX || g_ortho_sweep_only) {
      k_orthogonalise<<<panels, MGS_THREADS, 0, s>>>(a->basis64, a->nb, a->Vb, a->Ef, passes,
                                                     nullptr);
      MGC_LAUNCH_CHECK();
      a->basis32_valid = false;
      return;
    }
    if (!a->cq_T) {
      MGC_HIP(hipMalloc(&a->cq_T, (size_t)2 * panels * a->nb * a->nb * sizeof(cplx<double>)));
      const hipError_t e = hipMalloc(&a->cq_flag, (size_t)panels * sizeof(int));
      if (e != hipSuccess) {
        (void)hipFree(a->cq_T);
        a->cq_T = nullptr;
        MGC_HIP(e);
      }
    }
    fill_zero(a->cq_flag, (size_t)panels * sizeof(int), s);
    const bool wide = a->nb > 16;
    for (int round = 0; round < 2; ++round) {
      if (wide)
        k_cholqr_gram<32><<<panels, CqShape<32>::THREADS, 0, s>>>(a->basis64, a->cq_T, a->cq_flag,
                                                                 a->nb, a->Vb, a->Ef, round);
      else
        k_cholqr_gram<16><<<panels, CqShape<16>::THREADS, 0, s>>>(a->basis64, a->cq_T, a->cq_flag,
                                                                 a->nb, a->Vb, a->Ef, round);
      MGC_LAUNCH_CHECK();
    }
    if (wide)
      k_cholqr_apply<32><<<panels, CqShape<32>::THREADS, 0, s>>>(a->basis64, a->cq_T, a->cq_flag,
                                                                a->nb, a->Vb, a->Ef);
    else
      k_cholqr_apply<16><<<panels, CqShape<16>::THREADS, 0, s>>>(a->basis64, a->cq_T, a->cq_flag,
                                                                a->nb, a->Vb, a->Ef);
    MGC_LAUNCH_CHECK();
    // panels whose Cholesky broke down: the column sweep (most workgroups exit at once)
    k_orthogonalise<<<panels, MGS_THREADS, 0, s>>>(a->basis64, a->nb, a->Vb, a->Ef, passes,
                                                   a->cq_flag);
    MGC_LAUNCH_CHECK();
    a->basis32_valid = false;
  });
}

int mgc_aggregation_set_ortho_method(int method) {
  return guarded([&] {
    MGC_CHECK(method == 0 || method == 1, MGC_ERR_ARG, "method must be 0 or 1");
    g_ortho_sweep_only = method;
  });
}

int mgc_aggregation_check_orthogonal(mgc_aggregation* a, double* errs_host, void* stream) {
  return guarded([&] {
    check_agg(a);
    MGC_CHECK(errs_host, MGC_ERR_ARG, "null output");
    hipStream_t s = as_stream(stream);
    double* d = nullptr;
    MGC_HIP(hipMalloc(&d, a->nb * sizeof(double)));
    fill_zero(d, a->nb * sizeof(double), s);
    k_check_orthogonal<<<2 * a->Vc, 256, 0, s>>>(a->basis64, a->nb, a->Vb, a->Ef, d);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipMemcpyAsync(errs_host, d, a->nb * sizeof(double), hipMemcpyDeviceToHost, s);
    hipError_t e3 = hipStreamSynchronize(s);
    (void)hipFree(d);
    MGC_HIP(e1);
    MGC_HIP(e2);
    MGC_HIP(e3);
  });
}

int mgc_aggregation_project(mgc_aggregation* a, const void* fine, void* coarse, int prec,
                            int nrhs, void* stream) {
  return guarded([&] { agg_transfer(a, fine, coarse, prec, nrhs, true, 0, 8, as_stream(stream)); });
}

int mgc_aggregation_project_masked(mgc_aggregation* a, const void* fine, void* coarse, int prec,
                                   int nrhs, int mask_kind, int point, void* stream) {
  return guarded([&] {
    MGC_CHECK(mask_kind >= 0 && mask_kind <= 2, MGC_ERR_ARG, "mask kind must be 0, 1 or 2");
    MGC_CHECK(point >= 0 && point <= 8, MGC_ERR_ARG, "point must be in 0..8");
    agg_transfer(a, fine, coarse, prec, nrhs, true, mask_kind, point, as_stream(stream));
  });
}

int mgc_aggregation_promote(mgc_aggregation* a, const void* coarse, void* fine, int prec,
                            int nrhs, void* stream) {
  return guarded([&] {
    agg_transfer(a, coarse, fine, prec, nrhs, false, 0, 8, as_stream(stream));
  });
}

}  // extern "C"
