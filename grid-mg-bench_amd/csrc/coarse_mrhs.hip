// Multi-right-hand-side coarse stencil on MFMA (FP32, FP16, BF16 operands,
// FP32 accumulation), plus the FP16/BF16 input quantiser.
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
#include "coarse_op.hpp"

namespace mgc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int KIND> struct MK;
template <> struct MK<MGC_F32> {
  static constexpr int KC = 16;   // K elements per chunk (4 MFMA k-steps of 4)
  static constexpr int KCL = 4;   // K elements per lane per chunk
  typedef float T;
};
template <> struct MK<MGC_F16> {
  static constexpr int KC = 32;
  static constexpr int KCL = 8;
  typedef _Float16 T;
};
template <> struct MK<MGC_BF16> {
  static constexpr int KC = 32;
  static constexpr int KCL = 8;
  typedef __bf16 T;
};

__device__ inline floatx4 mfma_chunk_f32(uint4 a, uint4 b, floatx4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.x), __uint_as_float(b.x), c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.y), __uint_as_float(b.y), c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.z), __uint_as_float(b.z), c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a.w), __uint_as_float(b.w), c, 0, 0, 0);
  return c;
}

template <int KIND>
__device__ inline floatx4 mfma_chunk(uint4 a, uint4 b, floatx4 c) {
  if constexpr (KIND == MGC_F32) {
    return mfma_chunk_f32(a, b, c);
  } else if constexpr (KIND == MGC_F16) {
    halfx8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, c, 0, 0, 0);
  } else {
    bf16x8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
  }
}

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
  const int NMT = N / 16, NKC = 2 * N / M::KC;
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
      const cplx<R> a = links[(xp * N + c) * N + r];  // column-major (c, r)
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
  // one wavefront per row, lanes over N (N <= 64... generic loop)
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

// ---------------------------------------------------------------------------
// the multi-rhs stencil
// ---------------------------------------------------------------------------
// Slot s of the staged B' operand (kc = (s/64)/NT, nt = (s/64)%NT, lane = s%64)
// holds B'[k][n] for k = kc*KC + (lane/16)*KCL + e (e < KCL), n = nt*16 + lane%16,
// i.e. the KCL/2 complex entries c0.. of input column j = n/2, permuted into
// (Re, -Im) for n even and (Im, Re) for n odd.  For the half precisions every
// point's column is pre-multiplied by f(p, j) = s_link(x,p) s_in(y_p, j) / R_j
// with R_j = max_p s_link s_in (so |f| <= 1 and nothing overflows), and the
// accumulated tile is multiplied by R_j once at the end: all nine points then
// accumulate into a single FP32 tile.
template <int KIND, int N, int NT>
__global__ __launch_bounds__(N / 16 * 64) void k_mrhs(
    const uint4* __restrict__ A, const float* __restrict__ ascale, const void* __restrict__ in_,
    const float* __restrict__ in_scale, float* __restrict__ out, const int* __restrict__ nbr,
    int nrhs, int rhs0, int nr, int dag) {
  using M = MK<KIND>;
  typedef typename M::T T;
  constexpr int NMT = N / 16;
  constexpr int NKC = 2 * N / M::KC;
  constexpr int THREADS = NMT * 64;
  constexpr int NSLOT = NKC * NT * 64;
  constexpr int SPT = NSLOT / THREADS;  // staged slots per thread
  constexpr int NCOL = NT * 8;          // rhs columns covered by the tiles
  constexpr bool SCALED = KIND != MGC_F32;
  __shared__ uint4 sB[NSLOT];
  __shared__ float s_fac[SCALED ? 9 : 1][NCOL];
  __shared__ float s_R[NCOL];

  const int x = blockIdx.x;
  const int w = threadIdx.x >> 6;
  const int l = threadIdx.x & 63;
  const T* in = reinterpret_cast<const T*>(in_);  // [site][nrhs][N] complex (T pairs)

  int ys[9];
#pragma unroll
  for (int p = 0; p < 8; ++p) ys[p] = nbr[(size_t)x * 8 + p];
  ys[8] = x;

  if constexpr (SCALED) {
    for (int j = threadIdx.x; j < NCOL; j += THREADS) {
      float v[9], R = 0.f;
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        v[p] = j < nr ? ascale[(size_t)x * 9 + p] * in_scale[(size_t)ys[p] * nrhs + rhs0 + j] : 0.f;
        R = fmaxf(R, v[p]);
      }
      const float inv = R > 0.f ? 1.f / R : 0.f;
#pragma unroll
      for (int p = 0; p < 9; ++p) s_fac[p][j] = v[p] * inv;
      s_R[j] = R;
    }
  }

  // per-thread staging geometry (independent of the point)
  int soff[SPT];  // element offset of the 16-byte source chunk within a site, -1 if padding
  int scol[SPT];
#pragma unroll
  for (int q = 0; q < SPT; ++q) {
    const int s = threadIdx.x + q * THREADS;
    const int lane = s & 63, nt = (s >> 6) % NT, kc = (s >> 6) / NT;
    const int n = nt * 16 + (lane & 15);
    const int j = n >> 1;
    const int c0 = (kc * M::KC + (lane >> 4) * M::KCL) >> 1;
    soff[q] = j < nr ? ((rhs0 + j) * N + c0) * 2 : -1;
    scol[q] = n;
  }

  uint4 a_n[NKC], b_n[SPT];
  auto issue = [&](int p) {
    const uint4* Ap = A + ((((size_t)x * 9 + p) * NMT + w) * NKC) * 64 + l;
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) a_n[kc] = Ap[(size_t)kc * 64];
    const T* base = in + (size_t)ys[p] * nrhs * N * 2;
#pragma unroll
    for (int q = 0; q < SPT; ++q)
      if (soff[q] >= 0) b_n[q] = *reinterpret_cast<const uint4*>(base + soff[q]);
      else b_n[q] = make_uint4(0u, 0u, 0u, 0u);
  };

  floatx4 C[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) C[nt] = {0.f, 0.f, 0.f, 0.f};

  issue(0);
#pragma unroll 1
  for (int p = 0; p < 9; ++p) {
    uint4 a[NKC], b[SPT];
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) a[kc] = a_n[kc];
#pragma unroll
    for (int q = 0; q < SPT; ++q) b[q] = b_n[q];

    __syncthreads();  // previous point's MFMA readers are done with sB
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
      const int s = threadIdx.x + q * THREADS;
      const int n = scol[q];
      const int u = n & 1;
      const int c0 = (soff[q] >= 0) ? ((soff[q] >> 1) % N) : 0;
      float f = 1.f;
      if constexpr (SCALED) f = s_fac[p][n >> 1];
      T raw[M::KCL], v[M::KCL];
      __builtin_memcpy(raw, &b[q], 16);
#pragma unroll
      for (int h = 0; h < M::KCL / 2; ++h) {
        float re = (float)raw[2 * h], im = (float)raw[2 * h + 1];
        const float sg = (dag && (c0 + h) >= N / 2) ? -f : f;
        re *= sg;
        im *= sg;
        v[2 * h] = (T)(u ? im : re);
        v[2 * h + 1] = (T)(u ? re : -im);
      }
      uint4 wv;
      __builtin_memcpy(&wv, v, 16);
      sB[s] = wv;
    }
    __syncthreads();
    if (p < 8) issue(p + 1);  // next point's loads overlap this point's MFMAs

#pragma unroll
    for (int kc = 0; kc < NKC; ++kc)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        C[nt] = mfma_chunk<KIND>(a[kc], sB[(kc * NT + nt) * 64 + l], C[nt]);
  }
  // D[m][n]: lane l holds rows 4*(l/16)+i, column l%16
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int n = nt * 16 + (l & 15);
    const int j = n >> 1, u = n & 1;
    if (j < nr) {
      float R = 1.f;
      if constexpr (SCALED) R = s_R[j];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = w * 16 + 4 * (l >> 4) + i;
        float v = C[nt][i] * R;
        if (dag && r >= N / 2) v = -v;
        out[(((size_t)x * nrhs + rhs0 + j) * N + r) * 2 + u] = v;
      }
    }
  }
}

// Register-only variant: no LDS and no barrier.  Each wavefront owns MPW
// consecutive m-tiles of one site and builds the B' fragments it needs in
// registers straight from 16-byte loads of the neighbour's input (lanes l and
// l^1 read the same chunk), so wavefronts stream independently and the
// hardware overlaps one wavefront's loads with another's MFMAs.  The B'
// values, scale folding and MFMA order are those of k_mrhs.
template <int KIND, int N, int NT, int MPW>
__global__ __launch_bounds__(N / 16 / MPW * 64) void k_mrhs_reg(
    const uint4* __restrict__ A, const float* __restrict__ ascale, const void* __restrict__ in_,
    const float* __restrict__ in_scale, float* __restrict__ out, const int* __restrict__ nbr,
    int nrhs, int rhs0, int nr, int dag) {
  using M = MK<KIND>;
  typedef typename M::T T;
  constexpr int NMT = N / 16;
  constexpr int NKC = 2 * N / M::KC;
  constexpr bool SCALED = KIND != MGC_F32;
  const int x = blockIdx.x;
  const int w = threadIdx.x >> 6;
  const int l = threadIdx.x & 63;
  const T* in = reinterpret_cast<const T*>(in_);

  int ys[9];
#pragma unroll
  for (int p = 0; p < 8; ++p) ys[p] = nbr[(size_t)x * 8 + p];
  ys[8] = x;
  // this lane's columns: n = nt*16 + l%16, rhs j = n/2, part u = n%2
  const int u = l & 1;
  float invR[NT], Rj[NT];
  int jj[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int j = (nt * 16 + (l & 15)) >> 1;
    jj[nt] = j;
    float R = 0.f;
    if constexpr (SCALED) {
      if (j < nr)
#pragma unroll
        for (int p = 0; p < 9; ++p)
          R = fmaxf(R, ascale[(size_t)x * 9 + p] * in_scale[(size_t)ys[p] * nrhs + rhs0 + j]);
    } else {
      R = 1.f;
    }
    invR[nt] = R > 0.f ? 1.f / R : 0.f;
    Rj[nt] = R;
  }
  const int cl = (l >> 4) * M::KCL / 2;  // this lane's first complex within a k-chunk

  floatx4 C[MPW][NT];
#pragma unroll
  for (int mt = 0; mt < MPW; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) C[mt][nt] = {0.f, 0.f, 0.f, 0.f};

#pragma unroll 1
  for (int p = 0; p < 9; ++p) {
    const uint4* Ap = A + (((size_t)x * 9 + p) * NMT + w * MPW) * NKC * 64 + l;
    const T* base = in + (size_t)ys[p] * nrhs * N * 2;
    const float sl = SCALED ? ascale[(size_t)x * 9 + p] : 1.f;
    float f[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      f[nt] = 1.f;
      if constexpr (SCALED)
        f[nt] = jj[nt] < nr ? sl * in_scale[(size_t)ys[p] * nrhs + rhs0 + jj[nt]] * invR[nt] : 0.f;
    }
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      uint4 a[MPW];
#pragma unroll
      for (int mt = 0; mt < MPW; ++mt) a[mt] = Ap[((size_t)mt * NKC + kc) * 64];
      const int c0 = kc * M::KC / 2 + cl;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        uint4 raw4 = make_uint4(0u, 0u, 0u, 0u);
        if (jj[nt] < nr)
          raw4 = *reinterpret_cast<const uint4*>(base + (((size_t)(rhs0 + jj[nt])) * N + c0) * 2);
        T raw[M::KCL], v[M::KCL];
        __builtin_memcpy(raw, &raw4, 16);
#pragma unroll
        for (int h = 0; h < M::KCL / 2; ++h) {
          const float sg = (dag && (c0 + h) >= N / 2) ? -f[nt] : f[nt];
          const float re = (float)raw[2 * h] * sg, im = (float)raw[2 * h + 1] * sg;
          v[2 * h] = (T)(u ? im : re);
          v[2 * h + 1] = (T)(u ? re : -im);
        }
        uint4 b;
        __builtin_memcpy(&b, v, 16);
#pragma unroll
        for (int mt = 0; mt < MPW; ++mt) C[mt][nt] = mfma_chunk<KIND>(a[mt], b, C[mt][nt]);
      }
    }
  }
#pragma unroll
  for (int mt = 0; mt < MPW; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int j = jj[nt];
      if (j < nr) {
        const float R = Rj[nt];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = (w * MPW + mt) * 16 + 4 * (l >> 4) + i;
          float val = C[mt][nt][i] * R;
          if (dag && r >= N / 2) val = -val;
          out[(((size_t)x * nrhs + rhs0 + j) * N + r) * 2 + u] = val;
        }
      }
    }
}

// kernel variants: 0 = LDS-staged B', 1 = registers with 1 m-tile per
// wavefront, 2 = registers with 2 m-tiles per wavefront
constexpr int MRHS_VARIANTS = 3;

template <int KIND, int N, int NT>
static void launch_variant(int variant, int V, const uint4* A, const float* as, const void* in,
                           const float* is, float* out, const int* nbr, int nrhs, int rhs0, int nr,
                           int dag, hipStream_t s) {
  const dim3 g(V);
  if (variant == 0)
    k_mrhs<KIND, N, NT><<<g, N / 16 * 64, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr, dag);
  else if (variant == 1)
    k_mrhs_reg<KIND, N, NT, 1><<<g, N / 16 * 64, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr,
                                                        dag);
  else
    k_mrhs_reg<KIND, N, NT, 2><<<g, N / 32 * 64, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr,
                                                        dag);
}

template <int KIND, int N>
static void launch_mrhs_n(int variant, int NT, int V, const uint4* A, const float* as,
                          const void* in, const float* is, float* out, const int* nbr, int nrhs,
                          int rhs0, int nr, int dag, hipStream_t s) {
  switch (NT) {
#define CASE(nt)                                                                              \
  case nt:                                                                                    \
    launch_variant<KIND, N, nt>(variant, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s); \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(6) CASE(8)
#undef CASE
    default:
      throw Error(MGC_ERR_ARG, "internal: unsupported column tile count");
  }
  MGC_LAUNCH_CHECK();
}

template <int KIND>
static void launch_mrhs(int variant, int N, int NT, int V, const uint4* A, const float* as,
                        const void* in, const float* is, float* out, const int* nbr, int nrhs,
                        int rhs0, int nr, int dag, hipStream_t s) {
  if (N == 32)
    launch_mrhs_n<KIND, 32>(variant, NT, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);
  else if (N == 64)
    launch_mrhs_n<KIND, 64>(variant, NT, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);
  else
    throw Error(MGC_ERR_SHAPE, "multi-rhs stencil supports nbasis 32 and 64");
}

static void launch_any(int mp, int variant, int N, int NT, int V, const uint4* A, const float* as,
                       const void* in, const float* is, float* out, const int* nbr, int nrhs,
                       int rhs0, int nr, int dag, hipStream_t s) {
  if (mp == MGC_F32)
    launch_mrhs<MGC_F32>(variant, N, NT, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);
  else if (mp == MGC_F16)
    launch_mrhs<MGC_F16>(variant, N, NT, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);
  else
    launch_mrhs<MGC_BF16>(variant, N, NT, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);
}

// Autotuner: the first launch of a (precision, nbasis, column tiles) shape
// times every variant on the caller's own operands (three launches each,
// HIP events on the caller's stream; the outputs are rewritten by the final
// launch) and keeps the fastest for the rest of the process.
static int g_mrhs_choice[4][2][9];  // [prec][N==64][NT] -> variant + 1 (0 = untuned)
static int g_mrhs_forced = -1;      // >= 0: use this variant (set by mgc_mrhs_set_variant)

static int tuned_variant(int mp, int N, int NT, int V, const uint4* A, const float* as,
                         const void* in, const float* is, float* out, const int* nbr, int nrhs,
                         int rhs0, int nr, int dag, hipStream_t s) {
  if (g_mrhs_forced >= 0) return g_mrhs_forced;
  int& slot = g_mrhs_choice[mp][N == 64 ? 1 : 0][NT];
  if (slot) return slot - 1;
  hipEvent_t e0, e1;
  MGC_HIP(hipEventCreate(&e0));
  MGC_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bestv = 0;
  for (int v = 0; v < MRHS_VARIANTS; ++v) {
    launch_any(mp, v, N, NT, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);  // warm-up
    MGC_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < 3; ++k)
      launch_any(mp, v, N, NT, V, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, s);
    MGC_HIP(hipEventRecord(e1, s));
    MGC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    MGC_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      bestv = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  slot = bestv + 1;
  return bestv;
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
  MGC_CHECK(N == 32 || N == 64, MGC_ERR_SHAPE, "multi-rhs stencil supports nbasis 32 and 64");
  if (op->mrhs_links) {
    MGC_HIP(hipFree(op->mrhs_links));
    op->mrhs_links = nullptr;
  }
  const size_t eb = mp == MGC_F32 ? 4 : 2;
  MGC_HIP(hipMalloc(&op->mrhs_links, (size_t)V * 9 * N * N * 2 * eb));
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
  const size_t total = (size_t)V * 9 * (N / 16) * (2 * N / KC) * 64;
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
  const unsigned g = (unsigned)((rows + 3) / 4);
  if (prec == MGC_F16)
    k_quantize<MGC_F16><<<g, 256, 0, s>>>((const float2*)in, N, rows, out, scales);
  else if (prec == MGC_BF16)
    k_quantize<MGC_BF16><<<g, 256, 0, s>>>((const float2*)in, N, rows, out, scales);
  else
    throw Error(MGC_ERR_PRECISION, "quantize: target must be F16 or BF16");
  MGC_LAUNCH_CHECK();
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
    *variant = g_mrhs_forced >= 0 ? g_mrhs_forced
                                  : g_mrhs_choice[mrhs_prec][nbasis == 64 ? 1 : 0][tiles_for(nr)] - 1;
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
    MGC_CHECK(op && in && out && in != out, MGC_ERR_ARG, "null or aliased argument");
    MGC_CHECK(op->links_set, MGC_ERR_STATE, "links not set");
    MGC_CHECK(nrhs >= 1, MGC_ERR_ARG, "nrhs < 1");
    MGC_CHECK(mrhs_prec == MGC_F32 || mrhs_prec == MGC_F16 || mrhs_prec == MGC_BF16,
              MGC_ERR_PRECISION, "multi-rhs precision must be F32, F16 or BF16");
    hipStream_t s = as_stream(stream);
    prepare_mrhs(op, mrhs_prec, s);
    const int N = op->N, V = op->V;
    const void* src = in;
    const float* sc = in_scales;
    void* qbuf = nullptr;
    float* qsc = nullptr;
    if (mrhs_prec == MGC_F32) {
      MGC_CHECK(in_prec == MGC_F32, MGC_ERR_PRECISION, "F32 multi-rhs apply takes F32 vectors");
    } else if (in_prec == MGC_F32) {
      // quantise the FP32 input on the fly (per-(site, rhs) scales)
      MGC_HIP(hipMallocAsync(&qbuf, (size_t)V * nrhs * N * 4, s));
      MGC_HIP(hipMallocAsync((void**)&qsc, (size_t)V * nrhs * sizeof(float), s));
      quantize(in, (size_t)V * nrhs, N, mrhs_prec, qbuf, qsc, s);
      src = qbuf;
      sc = qsc;
    } else {
      MGC_CHECK(in_prec == mrhs_prec, MGC_ERR_PRECISION,
                "half-precision input must match the multi-rhs precision");
      MGC_CHECK(in_scales, MGC_ERR_ARG, "half-precision input needs its scales");
    }
    const int dg = (dag && !op->hermitian) ? 1 : 0;
    for (int r0 = 0; r0 < nrhs; r0 += 64) {
      const int nr = nrhs - r0 < 64 ? nrhs - r0 : 64;
      const int NT = tiles_for(nr);
      const uint4* A = (const uint4*)op->mrhs_links;
      const int v = tuned_variant(mrhs_prec, N, NT, V, A, op->mrhs_scales, src, sc, (float*)out,
                                  op->nbr_full, nrhs, r0, nr, dg, s);
      launch_any(mrhs_prec, v, N, NT, V, A, op->mrhs_scales, src, sc, (float*)out, op->nbr_full,
                 nrhs, r0, nr, dg, s);
    }
    if (qbuf) {
      MGC_HIP(hipFreeAsync(qbuf, s));
      MGC_HIP(hipFreeAsync(qsc, s));
    }
  });
}

}  // extern "C"
