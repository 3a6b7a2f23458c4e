// Multi-right-hand-side coarse stencil on MFMA (FP32, FP16, BF16 operands,
// FP32 accumulation), plus the FP16/BF16 input quantiser: host side (operand
// preparation, autotuner, entry points).  The kernels are in mrhs_kernels.hpp,
// instantiated per nbasis in mrhs_n*.hip.
//
// Reference semantics: include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h:736-804
// (M on a 5d grid whose local 5th dimension is the rhs index, links shared
// across rhs: sU = sF / Ls) and :806-820 (Mdag = G5C M G5C).
//
// Per coarse site x the apply is the complex GEMM
//     C_x[N x nrhs] = sum_p A_p(x)[N x N] . B_{x+d_p}[N x nrhs].
// It runs as ONE real GEMM per point with the complex index folded into K:
//     A'[r][2c+t] = (Re, Im)(A[r][c])        (the row-major complex row itself)
//     B'[2c+t][2j+u] = u=0: ( Re b, -Im b)   u=1: (Im b, Re b)   with b = B[c][j]
// so that C'[r][2j+u] = (Re, Im)(C[r][j]): the FP32 accumulator tile is already
// the interleaved complex output and no flop is wasted (8 real flops per
// complex multiply-add, as counted).
//
// Work decomposition: one workgroup per coarse site, one wavefront per
// 16-row m-tile (N/16 wavefronts), all nrhs columns per wavefront in NT tiles
// of 16 columns (v_mfma_f32_16x16x4_f32 or v_mfma_f32_16x16x32_{f16,bf16}).
// Links are pre-swizzled into MFMA fragment order [site][point][m-tile]
// [k-chunk][lane][16 B], so each wavefront load is one contiguous 1 KB
// request and every link byte is read from HBM exactly once.  The neighbour's
// B' fragments are staged in LDS once per point and read by every wavefront
// with ds_read_b128.  For FP16/BF16 each (site, point) link block carries an
// FP32 scale (block max -> 1) and each (site, rhs) input column an FP32
// scale; a point's product is accumulated in its own FP32 tile and folded
// into the result as C += s_link(x,p) * s_in(x+d_p, j) * T_p.
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "mrhs_kernels.hpp"

namespace mgc {

// ---------------------------------------------------------------------------
// link preparation: per-(site, point) scale and the swizzled operand
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_link_scale(const cplx<R>* __restrict__ links, int N,
                                                    float* __restrict__ scales, int unit) {
  // one workgroup per (site, point); links internal [x][p][c][r]
  const size_t base = (size_t)blockIdx.x * N * N;
  float m = 0.f;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) {
    const cplx<R> a = links[base + i];
    m = fmaxf(m, fmaxf(fabsf((float)a.re), fabsf((float)a.im)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  __shared__ float s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = fmaxf(fmaxf(s[0], s[1]), fmaxf(s[2], s[3]));
    if (unit || !(t > 0.f)) t = 1.f;
    scales[blockIdx.x] = t;
  }
}

template <int KIND, typename R>
__global__ void k_link_swizzle(const cplx<R>* __restrict__ links, int N,
                               const float* __restrict__ scales, uint4* __restrict__ out,
                               size_t total) {
  // out index: ((((x*9 + p)*NMT + mt)*NKC + kc)*64 + lane)
  using M = MK<KIND>;
  typedef typename M::T T;
  const int NMT = mrhs_mtiles(N), NKC = 2 * NMT * 16 / M::KC;
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < total; t += stride) {
    const int lane = (int)(t % 64);
    size_t q = t / 64;
    const int kc = (int)(q % NKC);
    q /= NKC;
    const int mt = (int)(q % NMT);
    const size_t xp = q / NMT;  // x*9 + p
    const float inv = 1.f / scales[xp];
    const int r = mt * 16 + (lane & 15);
    const int k0 = kc * M::KC + (lane >> 4) * M::KCL;
    T v[M::KCL];
#pragma unroll
    for (int e = 0; e < M::KCL; ++e) {
      const int k = k0 + e;
      const int c = k >> 1;
      // column-major (c, r); zero in the padding rows and columns
      const cplx<R> a = (r < N && c < N) ? links[(xp * N + c) * N + r] : cplx<R>{R(0), R(0)};
      const float f = (float)((k & 1) ? a.im : a.re) * inv;
      v[e] = (T)f;
    }
    uint4 w;
    __builtin_memcpy(&w, v, 16);
    out[t] = w;
  }
}

// ---------------------------------------------------------------------------
// input quantiser: x[row][N] (FP32 complex) -> half[row][N] + scale[row],
// scale = max |component| over the row (1 for an all-zero row)
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(256) void k_quantize(const float2* __restrict__ in, int N,
                                                  size_t rows, void* __restrict__ out,
                                                  float* __restrict__ scales) {
  typedef typename MK<KIND>::T T;
  // one wavefront per row, lanes over N (any N)
  const size_t row = blockIdx.x * (size_t)(blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  const float2* r = in + row * N;
  float m = 0.f;
  for (int c = lane; c < N; c += 64) m = fmaxf(m, fmaxf(fabsf(r[c].x), fabsf(r[c].y)));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  const float s = m > 0.f ? m : 1.f;
  const float inv = 1.f / s;
  T* o = reinterpret_cast<T*>(out) + row * 2 * N;
  for (int c = lane; c < N; c += 64) {
    o[2 * c] = (T)(r[c].x * inv);
    o[2 * c + 1] = (T)(r[c].y * inv);
  }
  if (lane == 0) scales[row] = s;
}

// The same for even N <= 64 and aligned rows, with the same results (the
// maximum does not depend on the order it is taken in):
template <int KIND>
__global__ __launch_bounds__(256) void k_quantize_pairs(const float2* __restrict__ in, int N,
                                                        size_t rows, void* __restrict__ out,
                                                        float* __restrict__ scales) {
  typedef typename MK<KIND>::T T;
  // half a wavefront per row: lane c2 (0..31) takes the complex pair
  // (2 c2, 2 c2 + 1) of the row in one 16-byte load (N even, at most 64),
  // so a wavefront reads two whole rows contiguously
  const size_t row = blockIdx.x * (size_t)(blockDim.x / 32) + (threadIdx.x >> 5);
  const int c2 = threadIdx.x & 31;
  const bool in_row = row < rows;  // the half-wave of a missing row still joins the shuffles
  const bool has = in_row && 2 * c2 < N;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (has) v = reinterpret_cast<const float4*>(in + row * N)[c2];
  float m = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
  for (int off = 16; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));  // within the half
  if (!has) {
    if (in_row && c2 == 0) scales[row] = m > 0.f ? m : 1.f;
    return;
  }
  const float s = m > 0.f ? m : 1.f;
  const float inv = 1.f / s;
  T q[4] = {(T)(v.x * inv), (T)(v.y * inv), (T)(v.z * inv), (T)(v.w * inv)};
  uint2 w;
  __builtin_memcpy(&w, q, 8);
  reinterpret_cast<uint2*>(reinterpret_cast<T*>(out) + row * 2 * N)[c2] = w;
  if (c2 == 0) scales[row] = s;
}

// runtime nbasis -> the instantiation of that nbasis
static void launch_any(int mp, int variant, int N, int NT, int V, int dims, const uint4* A,
                       const float* as, const void* in, const float* is, float* out,
                       const int* nbr, int nrhs, int rhs0, int nr, int dag, hipStream_t s) {
  switch (N) {
#define CASE(n)                                                                                \
  case n:                                                                                      \
    mrhs_launch<n>(mp, variant, NT, V, dims, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s); \
    return;
    MGC_MRHS_FOR_EACH_N(CASE)
#undef CASE
  }
  throw Error(MGC_ERR_SHAPE, "multi-rhs stencil supports nbasis 8, 16, ..., 64");
}

static void launch_batched(int mp, int N, int variant, int NT, int blocks, const MrhsBatch& bt,
                           int nrhs, int rhs0, int nr, hipStream_t s) {
  switch (N) {
#define CASE(n)                                                                  \
  case n:                                                                        \
    mrhs_launch_batched<n>(mp, variant, NT, blocks, bt, nrhs, rhs0, nr, s);      \
    return;
    MGC_MRHS_FOR_EACH_N(CASE)
#undef CASE
  }
  throw Error(MGC_ERR_SHAPE, "multi-rhs stencil supports nbasis 8, 16, ..., 64");
}

// variants the tuner times: all eleven for nbasis 32 and 64, the LDS-staged
// ones elsewhere; the batched launch has the LDS-staged ones without the
// XCD-block order
static const int ALL_VARIANTS[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
static const int LDS_VARIANTS[8] = {0, 1, 3, 4, 5, 6, 9, 10};
static const int BATCH_VARIANTS[6] = {0, 1, 3, 4, 5, 6};

// packed lattice extents for the XCD-block site order (0 where it does not
// apply: odd or large extents of the first three dimensions, or a site
// count that is not a multiple of 8)
static int xblk_dims(const Lat4& L) {
  for (int d = 0; d < 4; ++d)
    if (L.L[d] > 127) return 0;
  for (int d = 0; d < 3; ++d)
    if (L.L[d] % 2) return 0;
  return L.L[0] | (L.L[1] << 8) | (L.L[2] << 16) | (L.L[3] << 24);
}

// Autotuner: the first launch of a (precision, nbasis, column tiles) shape
// times every variant on the caller's own operands and keeps the fastest for
// the rest of the process.  Two rounds over all variants (the first also
// brings the clocks up), five launches each after one warm-up launch, HIP
// events on the caller's stream; a variant's time is its better round.  The
// outputs are rewritten by the final launch.
// Tuned choices are kept per (precision, nbasis, column tiles, site count,
// XCD-block geometry): the persistent and XCD-ordered variants depend on the
// lattice, so a choice made on one lattice is not reused on another.
typedef std::tuple<int, int, int, int, int> TuneKey;
static std::map<TuneKey, int> g_mrhs_choice;                   // -> variant
static std::map<std::tuple<int, int, int>, int> g_mrhs_last;  // (prec, N, NT) -> last tuned
static int g_mrhs_forced = -1;  // >= 0: use this variant (set by mgc_mrhs_set_variant)

template <typename F>
static int time_variants(const int* list, int count, F&& launch, hipStream_t s) {
  constexpr int ROUNDS = 2, REPS = 5, MAXV = 16;
  hipEvent_t e0, e1;
  MGC_HIP(hipEventCreate(&e0));
  MGC_HIP(hipEventCreate(&e1));
  float best_of[MAXV];
  for (int v = 0; v < count && v < MAXV; ++v) best_of[v] = 1e30f;
  for (int round = 0; round < ROUNDS; ++round) {
    for (int v = 0; v < count && v < MAXV; ++v) {
      launch(list[v]);  // warm-up
      MGC_HIP(hipEventRecord(e0, s));
      for (int k = 0; k < REPS; ++k) launch(list[v]);
      MGC_HIP(hipEventRecord(e1, s));
      MGC_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      MGC_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_of[v]) best_of[v] = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int bestv = 0;
  for (int v = 1; v < count && v < MAXV; ++v)
    if (best_of[v] < best_of[bestv]) bestv = v;
  return list[bestv];
}

static int tuned_variant(int mp, int N, int NT, int V, int dims, const uint4* A,
                         const float* as, const void* in, const float* is, float* out,
                         const int* nbr, int nrhs, int rhs0, int nr, int dag, hipStream_t s) {
  if (g_mrhs_forced >= 0) return g_mrhs_forced;
  const TuneKey key(mp, N, NT, V, dims);
  const auto hit = g_mrhs_choice.find(key);
  if (hit != g_mrhs_choice.end()) return hit->second;
  const bool all = mrhs_all_variants(N);
  const int v = time_variants(
      all ? ALL_VARIANTS : LDS_VARIANTS, all ? 11 : 8,
      [&](int v) {
        launch_any(mp, v, N, NT, V, dims, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);
      },
      s);
  g_mrhs_choice[key] = v;
  g_mrhs_last[std::make_tuple(mp, N, NT)] = v;
  return v;
}

static std::map<TuneKey, int> g_batch_choice;  // key site count = all lattices together

static int tuned_batched(int mp, int N, int NT, int blocks, const MrhsBatch& bt, int nrhs,
                         int rhs0, int nr, hipStream_t s) {
  if (g_mrhs_forced >= 0 && g_mrhs_forced != 2 && g_mrhs_forced < 7) return g_mrhs_forced;
  const TuneKey key(mp, N, NT, blocks, bt.count);
  const auto hit = g_batch_choice.find(key);
  if (hit != g_batch_choice.end()) return hit->second;
  const int v = time_variants(
      BATCH_VARIANTS, 6,
      [&](int v) { launch_batched(mp, N, v, NT, blocks, bt, nrhs, rhs0, nr, s); }, s);
  g_batch_choice[key] = v;
  return v;
}

static int tiles_for(int nr) {
  const int need = (2 * nr + 15) / 16;
  const int opts[] = {1, 2, 3, 4, 6, 8};
  for (int o : opts)
    if (o >= need) return o;
  return 8;
}

// build (or rebuild) the swizzled link operand for precision mp
static void prepare_mrhs(mgc_coarse_op* op, int mp, hipStream_t s) {
  if (op->mrhs_links && op->mrhs_prec == mp) return;
  const int N = op->N, V = op->V;
  MGC_CHECK(mrhs_supported(N), MGC_ERR_SHAPE,
            "multi-rhs stencil supports nbasis 8, 16, 24, ..., 64 (multiples of 8)");
  const int NP = mrhs_mtiles(N) * 16;  // padded operand size
  if (op->mrhs_links) {
    MGC_HIP(hipFree(op->mrhs_links));
    op->mrhs_links = nullptr;
  }
  const size_t eb = mp == MGC_F32 ? 4 : 2;
  MGC_HIP(hipMalloc(&op->mrhs_links, (size_t)V * 9 * NP * NP * 2 * eb));
  if (!op->mrhs_scales) MGC_HIP(hipMalloc(&op->mrhs_scales, (size_t)V * 9 * sizeof(float)));
  const int unit = mp == MGC_F32;
  if (op->prec == MGC_F64)
    k_link_scale<double><<<V * 9, 256, 0, s>>>((const cplx<double>*)op->links, N, op->mrhs_scales,
                                               unit);
  else
    k_link_scale<float><<<V * 9, 256, 0, s>>>((const cplx<float>*)op->links, N, op->mrhs_scales,
                                              unit);
  MGC_LAUNCH_CHECK();
  const int KC = mp == MGC_F32 ? 16 : 32;
  const size_t total = (size_t)V * 9 * (NP / 16) * (2 * NP / KC) * 64;
  const unsigned g = grid_for(total, 256, 16384);
  uint4* o = (uint4*)op->mrhs_links;
#define SWZ(KIND, R)                                                                        \
  k_link_swizzle<KIND, R><<<g, 256, 0, s>>>((const cplx<R>*)op->links, N, op->mrhs_scales, o, \
                                            total)
  if (op->prec == MGC_F64) {
    if (mp == MGC_F32) SWZ(MGC_F32, double);
    else if (mp == MGC_F16) SWZ(MGC_F16, double);
    else SWZ(MGC_BF16, double);
  } else {
    if (mp == MGC_F32) SWZ(MGC_F32, float);
    else if (mp == MGC_F16) SWZ(MGC_F16, float);
    else SWZ(MGC_BF16, float);
  }
#undef SWZ
  MGC_LAUNCH_CHECK();
  op->mrhs_prec = mp;
}

static void quantize(const void* in, size_t rows, int N, int prec, void* out, float* scales,
                     hipStream_t s) {
  const bool pairs = N % 2 == 0 && N <= 64 && (uintptr_t)in % 16 == 0 && (uintptr_t)out % 8 == 0;
  const unsigned g = (unsigned)(pairs ? (rows + 7) / 8 : (rows + 3) / 4);
  if (prec == MGC_F16 && pairs)
    k_quantize_pairs<MGC_F16><<<g, 256, 0, s>>>((const float2*)in, N, rows, out, scales);
  else if (prec == MGC_BF16 && pairs)
    k_quantize_pairs<MGC_BF16><<<g, 256, 0, s>>>((const float2*)in, N, rows, out, scales);
  else if (prec == MGC_F16)
    k_quantize<MGC_F16><<<g, 256, 0, s>>>((const float2*)in, N, rows, out, scales);
  else if (prec == MGC_BF16)
    k_quantize<MGC_BF16><<<g, 256, 0, s>>>((const float2*)in, N, rows, out, scales);
  else
    throw Error(MGC_ERR_PRECISION, "quantize: target must be F16 or BF16");
  MGC_LAUNCH_CHECK();
}

// Scratch of the no-allocation multi-rhs apply: the quantised FP32 input
// (sites x rhs rows of half-precision complex, then one FP32 scale per row).
// One copy per device, held until the process ends; every call on a device
// shares it, so calls that overlap on one device race.  Host code uses the
// device address from hipGetSymbolAddress (scratch_ptr), never the symbol.
constexpr size_t MRHS_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[MRHS_SCRATCH_BYTES];
constexpr int MRHS_MAX_DEVICES = 64;
static void* g_scratch_ptr[MRHS_MAX_DEVICES];
static std::mutex g_scratch_mutex;

// device address of this device's scratch, looked up once per device on the
// first call made with that device current
static void* scratch_ptr() {
  int dev = 0;
  MGC_HIP(hipGetDevice(&dev));
  MGC_CHECK(dev >= 0 && dev < MRHS_MAX_DEVICES, MGC_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    MGC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MGC_CHECK(ptr, MGC_ERR_HIP, "no device address for the multi-rhs scratch");
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

static void check_mrhs_args(const mgc_coarse_op* op, const void* in, const void* out, int nrhs,
                            int mrhs_prec) {
  MGC_CHECK(op && in && out && in != out, MGC_ERR_ARG, "null or aliased argument");
  MGC_CHECK(op->links_set, MGC_ERR_STATE, "links not set");
  MGC_CHECK(nrhs >= 1, MGC_ERR_ARG, "nrhs < 1");
  MGC_CHECK(mrhs_prec == MGC_F32 || mrhs_prec == MGC_F16 || mrhs_prec == MGC_BF16,
            MGC_ERR_PRECISION, "multi-rhs precision must be F32, F16 or BF16");
}

// The multi-rhs apply on a prepared operand.  FP32 input to an F16/BF16
// apply is quantised into qbuf / qsc (rows = sites x nrhs) first.  With
// tune = false an untuned shape runs the default variant instead of timing
// all of them (no events, no host synchronisation).
static void mrhs_apply(mgc_coarse_op* op, const void* in, const float* in_scales, int in_prec,
                       void* out, int nrhs, int mrhs_prec, int dag, hipStream_t s, bool tune,
                       void* qbuf, float* qsc) {
  const int N = op->N, V = op->V;
  const void* src = in;
  const float* sc = in_scales;
  if (mrhs_prec == MGC_F32) {
    MGC_CHECK(in_prec == MGC_F32, MGC_ERR_PRECISION, "F32 multi-rhs apply takes F32 vectors");
  } else if (in_prec == MGC_F32) {
    quantize(in, (size_t)V * nrhs, N, mrhs_prec, qbuf, qsc, s);
    src = qbuf;
    sc = qsc;
  } else {
    MGC_CHECK(in_prec == mrhs_prec, MGC_ERR_PRECISION,
              "half-precision input must match the multi-rhs precision");
    MGC_CHECK(in_scales, MGC_ERR_ARG, "half-precision input needs its scales");
  }
  const int dg = (dag && !op->hermitian) ? 1 : 0;
  const uint4* A = (const uint4*)op->mrhs_links;
  const int dims = xblk_dims(op->lat);
  for (int r0 = 0; r0 < nrhs; r0 += 64) {
    const int nr = nrhs - r0 < 64 ? nrhs - r0 : 64;
    const int NT = tiles_for(nr);
    int v;
    if (tune) {
      v = tuned_variant(mrhs_prec, N, NT, V, dims, A, op->mrhs_scales, src, sc, (float*)out,
                        op->nbr_full, nrhs, r0, nr, dg, s);
    } else if (g_mrhs_forced >= 0) {
      v = g_mrhs_forced;
    } else {
      const auto hit = g_mrhs_choice.find(TuneKey(mrhs_prec, N, NT, V, dims));
      v = hit != g_mrhs_choice.end() ? hit->second : MRHS_DEFAULT_VARIANT;
    }
    launch_any(mrhs_prec, v, N, NT, V, dims, A, op->mrhs_scales, src, sc, (float*)out,
               op->nbr_full, nrhs, r0, nr, dg, s);
  }
}

}  // namespace mgc

using namespace mgc;

extern "C" {

int mgc_quantize(const void* in_f32, size_t rows, int n, int prec, void* out, float* scales,
                 void* stream) {
  return guarded([&] {
    MGC_CHECK(in_f32 && out && scales, MGC_ERR_ARG, "null argument");
    MGC_CHECK(n >= 1, MGC_ERR_ARG, "n < 1");
    if (rows == 0) return;
    quantize(in_f32, rows, n, prec, out, scales, as_stream(stream));
  });
}

int mgc_mrhs_set_variant(int variant) {
  return guarded([&] {
    MGC_CHECK(variant >= -1 && variant < MRHS_VARIANTS, MGC_ERR_ARG, "unknown kernel variant");
    g_mrhs_forced = variant;
  });
}

int mgc_mrhs_get_variant(int mrhs_prec, int nbasis, int nrhs, int* variant) {
  return guarded([&] {
    MGC_CHECK(variant, MGC_ERR_ARG, "null output");
    MGC_CHECK(mrhs_prec >= MGC_F32 && mrhs_prec <= MGC_BF16, MGC_ERR_PRECISION, "bad precision");
    const int nr = nrhs < 64 ? nrhs : 64;
    if (g_mrhs_forced >= 0) {
      *variant = g_mrhs_forced;
      return;
    }
    const auto hit = g_mrhs_last.find(std::make_tuple(mrhs_prec, nbasis, tiles_for(nr)));
    *variant = hit == g_mrhs_last.end() ? -1 : hit->second;
  });
}

int mgc_coarse_op_prepare_mrhs(mgc_coarse_op* op, int mrhs_prec, void* stream) {
  return guarded([&] {
    MGC_CHECK(op, MGC_ERR_ARG, "null operator handle");
    MGC_CHECK(op->links_set, MGC_ERR_STATE, "links not set");
    MGC_CHECK(mrhs_prec == MGC_F32 || mrhs_prec == MGC_F16 || mrhs_prec == MGC_BF16,
              MGC_ERR_PRECISION, "multi-rhs precision must be F32, F16 or BF16");
    prepare_mrhs(op, mrhs_prec, as_stream(stream));
  });
}

int mgc_coarse_op_M_mrhs(mgc_coarse_op* op, const void* in, const float* in_scales, int in_prec,
                         void* out, int nrhs, int mrhs_prec, int dag, void* stream) {
  return guarded([&] {
    check_mrhs_args(op, in, out, nrhs, mrhs_prec);
    hipStream_t s = as_stream(stream);
    prepare_mrhs(op, mrhs_prec, s);
    void* qbuf = nullptr;
    float* qsc = nullptr;
    if (mrhs_prec != MGC_F32 && in_prec == MGC_F32) {
      // quantisation workspace of the operator, grown on demand and kept
      const size_t rows = (size_t)op->V * nrhs;
      if (rows > op->q_rows) {
        if (op->q_in) MGC_HIP(hipFree(op->q_in));
        if (op->q_scale) MGC_HIP(hipFree(op->q_scale));
        op->q_in = nullptr;
        op->q_scale = nullptr;
        op->q_rows = 0;
        MGC_HIP(hipMalloc(&op->q_in, rows * op->N * 4));
        MGC_HIP(hipMalloc((void**)&op->q_scale, rows * sizeof(float)));
        op->q_rows = rows;
      }
      qbuf = op->q_in;
      qsc = op->q_scale;
    }
    mrhs_apply(op, in, in_scales, in_prec, out, nrhs, mrhs_prec, dag, s, true, qbuf, qsc);
  });
}

int mgc_coarse_op_M_mrhs_noalloc(mgc_coarse_op* op, const void* in_f32, void* out, int nrhs,
                                 int mrhs_prec, int dag, void* stream) {
  return guarded([&] {
    check_mrhs_args(op, in_f32, out, nrhs, mrhs_prec);
    MGC_CHECK(op->mrhs_links && op->mrhs_prec == mrhs_prec, MGC_ERR_STATE,
              "multi-rhs operand not prepared for this precision "
              "(call mgc_coarse_op_prepare_mrhs first)");
    void* qbuf = nullptr;
    float* qsc = nullptr;
    if (mrhs_prec != MGC_F32) {
      const size_t rows = (size_t)op->V * nrhs;
      const size_t qbytes = (rows * op->N * 4 + 255) / 256 * 256;
      MGC_CHECK(qbytes + rows * sizeof(float) <= MRHS_SCRATCH_BYTES, MGC_ERR_SHAPE,
                "sites x nrhs x (4 nbasis + 4) bytes exceed the 64 MiB scratch of the "
                "no-allocation multi-rhs apply");
      char* base = (char*)scratch_ptr();
      qbuf = base;
      qsc = (float*)(base + qbytes);
    }
    mrhs_apply(op, in_f32, nullptr, MGC_F32, out, nrhs, mrhs_prec, dag, as_stream(stream), false,
               qbuf, qsc);
  });
}

int mgc_coarse_op_M_mrhs_batched(mgc_coarse_op* const* ops, int count, const void* const* in,
                                 const float* const* in_scales, int in_prec, void* const* out,
                                 int nrhs, int mrhs_prec, int dag, void* stream) {
  return guarded([&] {
    MGC_CHECK(ops && in && out, MGC_ERR_ARG, "null argument");
    MGC_CHECK(count >= 1 && count <= MRHS_BATCH_MAX, MGC_ERR_ARG, "count must be 1..8");
    MGC_CHECK(nrhs >= 1, MGC_ERR_ARG, "nrhs < 1");
    MGC_CHECK(mrhs_prec == MGC_F32 || mrhs_prec == MGC_F16 || mrhs_prec == MGC_BF16,
              MGC_ERR_PRECISION, "multi-rhs precision must be F32, F16 or BF16");
    MGC_CHECK(in_prec == mrhs_prec, MGC_ERR_PRECISION,
              "batched apply takes vectors already in the multi-rhs precision");
    hipStream_t s = as_stream(stream);
    MrhsBatch bt{};
    bt.count = count;
    const int N = ops[0] ? ops[0]->N : 0;
    long long blocks = 0;
    for (int l = 0; l < count; ++l) {
      mgc_coarse_op* op = ops[l];
      MGC_CHECK(op && in[l] && out[l] && in[l] != out[l], MGC_ERR_ARG,
                "null or aliased argument in the batch");
      MGC_CHECK(op->links_set, MGC_ERR_STATE, "links not set");
      MGC_CHECK(op->N == N, MGC_ERR_SHAPE, "every lattice in a batch needs the same nbasis");
      MGC_CHECK(mrhs_prec == MGC_F32 || (in_scales && in_scales[l]), MGC_ERR_ARG,
                "half-precision input needs its scales");
      prepare_mrhs(op, mrhs_prec, s);
      bt.A[l] = (const uint4*)op->mrhs_links;
      bt.ascale[l] = op->mrhs_scales;
      bt.in[l] = in[l];
      bt.in_scale[l] = mrhs_prec == MGC_F32 ? nullptr : in_scales[l];
      bt.out[l] = (float*)out[l];
      bt.nbr[l] = op->nbr_full;
      bt.dag[l] = (dag && !op->hermitian) ? 1 : 0;
      bt.first[l] = (int)blocks;
      blocks += op->V;
    }
    MGC_CHECK(blocks <= 0x7fffffffLL, MGC_ERR_SHAPE, "batch too large for one launch");
    bt.first[count] = (int)blocks;
    for (int r0 = 0; r0 < nrhs; r0 += 64) {
      const int nr = nrhs - r0 < 64 ? nrhs - r0 : 64;
      const int NT = tiles_for(nr);
      const int v = tuned_batched(mrhs_prec, N, NT, (int)blocks, bt, nrhs, r0, nr, s);
      launch_batched(mrhs_prec, N, v, NT, (int)blocks, bt, nrhs, r0, nr, s);
    }
  });
}

}  // extern "C"
