// Multi-right-hand-side coarse stencil on MFMA: the kernels and their
// launchers, shared by the translation units that instantiate them for one
// nbasis each (mrhs_n*.hip), so that the variants of every nbasis compile in
// parallel.  The host side (operand preparation, autotuner, entry points) is
// coarse_mrhs.hip.
//
// Reference semantics: include/CoarsenedMatrixUpstreamImprovedDirsaveLutMRHS.h:736-804
// (M on a 5d grid whose local 5th dimension is the rhs index, links shared
// across rhs: sU = sF / Ls) and :806-820 (Mdag = G5C M G5C).
//
// nbasis N is any multiple of 8 up to 64.  The link operand is padded with
// zero rows and columns to the next multiple of 16 (mrhs_mtiles), so a padded
// K entry of B' (read from the next input column, or the zeros of a buffer
// range check) always meets a zero link entry, and rows past N are never
// stored.  N % 8 == 0 keeps each lane's chunk of KCL/2 complex entries on one
// side of N/2, where the Mdag sign changes.
#pragma once
#include <type_traits>

#include "coarse_op.hpp"

namespace mgc {

// 16-row m-tiles of an nbasis-N operand (rows and K padded to 16 * tiles)
__host__ __device__ constexpr int mrhs_mtiles(int N) { return (N + 15) / 16; }
// nbasis for which every kernel variant is built (the others run the
// LDS-staged variants only)
constexpr bool mrhs_all_variants(int N) { return N == 32 || N == 64; }
// nbasis values with a multi-rhs instantiation
constexpr bool mrhs_supported(int N) { return N >= 8 && N <= 64 && N % 8 == 0; }

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
// the multi-rhs stencil
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Link fragment loads.  The links are read exactly once, the neighbour
// inputs by up to nine workgroups, so the link stream should not displace
// the inputs in L2.  LAUX selects the cache policy of the link loads: 0 = a
// plain global load; otherwise a buffer load over the site's link block with
// the policy bits LAUX (2 = nt: non-temporal).
constexpr int MRHS_BUF_DWORD3 = 0x00020000;  // buffer resource word 3 (raw, 32-bit format)

template <int LAUX>
struct LinkSrc {
  const uint4* base;  // plain loads: this wavefront's lane pointer into the site block
  __amdgpu_buffer_rsrc_t rsrc;
  __device__ LinkSrc(const uint4* site_block, int bytes, int lane_off) {
    if constexpr (LAUX == 0) {
      base = site_block + lane_off;
    } else {
      rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4*>(site_block), (short)0, bytes,
                                               MRHS_BUF_DWORD3);
      base = nullptr;
    }
  }
  // fragment at element offset e (uint4 units) past this lane's position
  __device__ __forceinline__ uint4 load(int lane_off, int e) const {
    if constexpr (LAUX == 0) {
      return base[e];
    } else {
      const u32x4 v =
          __builtin_amdgcn_raw_buffer_load_b128(rsrc, (lane_off + e) * 16, 0, LAUX);
      return make_uint4(v.x, v.y, v.z, v.w);
    }
  }
};

// One 16-byte B' fragment from the stored input chunk with bit operations
// only: u = 0 gives (Re, -Im), u = 1 gives (Im, Re) for every complex entry,
// and neg negates both parts (the G5C sign of Mdag).
template <int KIND>
__device__ __forceinline__ uint4 bprime(uint4 r, int u, bool neg) {
  constexpr unsigned S = 0x80000000u;
  uint4 o;
  if constexpr (KIND == MGC_F32) {
    // words: Re0, Im0, Re1, Im1
    const unsigned mn = neg ? S : 0u;
    const unsigned mi = (u ? 0u : S) ^ mn;
    o.x = (u ? r.y : r.x) ^ mn;
    o.y = (u ? r.x : r.y) ^ mi;
    o.z = (u ? r.w : r.z) ^ mn;
    o.w = (u ? r.z : r.w) ^ mi;
  } else {
    // one complex entry per word: Re in the low half, Im in the high half
    const unsigned m = (u ? 0u : S) ^ (neg ? 0x80008000u : 0u);
    o.x = (u ? __builtin_rotateright32(r.x, 16) : r.x) ^ m;
    o.y = (u ? __builtin_rotateright32(r.y, 16) : r.y) ^ m;
    o.z = (u ? __builtin_rotateright32(r.z, 16) : r.z) ^ m;
    o.w = (u ? __builtin_rotateright32(r.w, 16) : r.w) ^ m;
  }
  return o;
}

// Order in which a workgroup visits the nine points: self, -t, +x, -x, +y,
// -y, +z, -z, +t.  Workgroups of one XCD run nearly in step, so consecutive
// steps read largely the same neighbour inputs (shifted by one lattice
// spacing) while they are still in L2.  With sites walked t slowest on each
// XCD, the -t inputs were read by the workgroups that ran just before, and
// the +t inputs are read again by the ones that run just after.
__device__ __forceinline__ int mrhs_point(int q) {
  return q == 0 ? 8 : q == 1 ? 7 : q == 8 ? 3 : ((q - 2) >> 1) + (((q - 2) & 1) ? 4 : 0);
}

// B' words of one staged 16-byte chunk for FP16/BF16 with a per-lane byte
// selector and sign mask: sel = identity (u = 0) or half swap (u = 1), and
// m = (u = 0 ? sign of the high half : 0) ^ (both signs for the G5C sign of
// Mdag).  The result is bit-identical to bprime().
__device__ __forceinline__ uint4 bprime_perm(uint4 r, unsigned sel, unsigned m) {
  uint4 o;
  o.x = __builtin_amdgcn_perm(r.x, r.x, sel) ^ m;
  o.y = __builtin_amdgcn_perm(r.y, r.y, sel) ^ m;
  o.z = __builtin_amdgcn_perm(r.z, r.z, sel) ^ m;
  o.w = __builtin_amdgcn_perm(r.w, r.w, sel) ^ m;
  return o;
}

// LDS-staged form for one site x.  Slot s of the staged B' operand
// (kc = (s/64)/NT, nt = (s/64)%NT, lane = s%64) holds B'[k][n] for
// k = kc*KC + (lane/16)*KCL + e (e < KCL), n = nt*16 + lane%16: the KCL/2
// complex entries c0.. of input column j = n/2.  The first eight points run
// as a rolled loop in which every load is issued and the ninth is peeled, so
// the memory counters the compiler waits on are static.  Neighbour inputs are
// read with unconditional buffer loads over the neighbour's site block (a
// column j >= nr is never stored, see the staging geometry below).  The next
// point's input loads are issued as soon as this point's chunks are staged,
// before the barrier that releases the MFMAs.  For FP16/BF16 the point's
// product T_p is accumulated from zero and folded into the FP32 result as
// C += s_link(x,p) * s_in(x+d_p, j) * T_p.
template <int KIND, int N, int NT, int LAUX, bool D2 = false>
__device__ __forceinline__ void mrhs_lds_site(
    const int x, const uint4* __restrict__ A, const float* __restrict__ ascale,
    const void* __restrict__ in_, const float* __restrict__ in_scale, float* __restrict__ out,
    const int* __restrict__ nbr, int nrhs, int rhs0, int nr, int dag) {
  using M = MK<KIND>;
  typedef typename M::T T;
  constexpr int NMT = mrhs_mtiles(N);   // 16-row m-tiles, the last one zero-padded
  constexpr int NKC = 2 * NMT * 16 / M::KC;
  constexpr int THREADS = NMT * 64;
  constexpr int NSLOT = NKC * NT * 64;
  constexpr int SPT = NSLOT / THREADS;  // staged slots per thread
  constexpr int NCOL = NT * 8;          // rhs columns covered by the tiles
  constexpr bool SCALED = KIND != MGC_F32;
  static_assert(NCOL <= THREADS, "one thread per column for the scale table");
  __shared__ uint4 sB[NSLOT];
  __shared__ float s_fac[SCALED ? 9 : 1][NCOL];

  __shared__ int s_ys[NMT][8];  // each wavefront's copy of the neighbour sites

  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  const int u = l & 1;  // every slot of this thread has lane l, so column parity l%2
  const char* in = reinterpret_cast<const char*>(in_);  // [site][nrhs][N] complex (T pairs)
  const int site_bytes = nrhs * N * 2 * (int)sizeof(T);

  // neighbour sites: scalar for the scale table, and a per-wavefront LDS
  // copy from which the point loop reads the next neighbour (written and
  // read by the same wavefront, so no barrier is involved)
  int ys[9];
#pragma unroll
  for (int p = 0; p < 8; ++p) ys[p] = nbr[(size_t)x * 8 + p];
  ys[8] = x;
  if (l < 8) s_ys[w][l] = nbr[(size_t)x * 8 + l];

  if constexpr (SCALED) {
    if (threadIdx.x < NCOL) {
      const int j = threadIdx.x;
#pragma unroll
      for (int p = 0; p < 9; ++p)
        s_fac[p][j] =
            j < nr ? ascale[(size_t)x * 9 + p] * in_scale[(size_t)ys[p] * nrhs + rhs0 + j] : 0.f;
    }
  }

  // Staging geometry.  Slot q of this thread reads the 16-byte chunk at
  // byte lane_b + slot_b(q) of the neighbour's site block: lane_b depends on
  // the lane only, slot_b(q) and the slot's k-chunk (hence its Mdag sign) on
  // the wavefront only.  A column j >= nr reads another column of the same
  // site (rhs0 + j < nrhs) or, past the block, the zeros that the buffer
  // range check returns; it only reaches output columns that are not stored.
  constexpr int ROWB = N * 2 * (int)sizeof(T);  // bytes of one input column
  const int lane_b = ((l & 15) >> 1) * ROWB + (l >> 4) * M::KCL * (int)sizeof(T);
  auto slot_nt = [&](int q) { return (w + q * NMT) % NT; };
  auto slot_kc = [&](int q) { return (w + q * NMT) / NT; };
  auto slot_b = [&](int q) {
    return (rhs0 + slot_nt(q) * 8) * ROWB + slot_kc(q) * M::KC * (int)sizeof(T);
  };
  // Mdag sign of the slot's entries (c >= N/2): uniform over a chunk when N
  // is a multiple of the chunk, else per lane (N % 8 == 0 keeps every lane's
  // KCL/2 entries on one side)
  auto slot_neg = [&](int q) {
    if constexpr (N % M::KC == 0) return dag && slot_kc(q) * M::KC >= N;
    else return dag && slot_kc(q) * M::KC + (l >> 4) * M::KCL >= N;
  };
  const unsigned ssel = u ? 0x01000302u : 0x03020100u;
  const unsigned smask0 = u ? 0u : 0x80000000u;

  const int lane_off = w * NKC * 64 + l;
  const LinkSrc<LAUX> src(A + (size_t)x * 9 * NMT * NKC * 64, 9 * NMT * NKC * 64 * 16, lane_off);
  // 16-byte input chunks of neighbour y into dst (one per staged slot)
  auto issue_b = [&](uint4* dst, int y) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char*>(in + (size_t)y * site_bytes), (short)0, site_bytes, MRHS_BUF_DWORD3);
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
      const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, lane_b, slot_b(q), 0);
      dst[q] = make_uint4(v.x, v.y, v.z, v.w);
    }
  };

  floatx4 C[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) C[nt] = {0.f, 0.f, 0.f, 0.f};

  // One point: stage the landed input chunks b as B', issue the input loads
  // of neighbour yl into b (the point LOOK points ahead), release the MFMAs;
  // the link fragment a is reloaded for point pl as soon as this point's
  // MFMAs have consumed it.
  auto point = [&](int q9, uint4* b, uint4* a, int pl, int yl, auto more_c) {
    constexpr bool more = decltype(more_c)::value;
    const int p = mrhs_point(q9);
    if (q9 > 0) __syncthreads();  // previous point's MFMA readers are done with sB
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
      if constexpr (SCALED)
        sB[threadIdx.x + q * THREADS] =
            bprime_perm(b[q], ssel, smask0 ^ (slot_neg(q) ? 0x80008000u : 0u));
      else
        sB[threadIdx.x + q * THREADS] = bprime<KIND>(b[q], u, slot_neg(q));
    }
    if constexpr (more) issue_b(b, yl);  // later point's input loads overlap this point's MFMAs
    __syncthreads();
    const int anext = (pl * NMT * NKC) * 64;

    if constexpr (SCALED) {
      floatx4 Tp[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) Tp[nt] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < NKC; ++kc) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          Tp[nt] = mfma_chunk<KIND>(a[kc], sB[(kc * NT + nt) * 64 + l], Tp[nt]);
        if constexpr (more) a[kc] = src.load(lane_off, anext + kc * 64);
      }
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const float f = s_fac[p][(nt * 16 + (l & 15)) >> 1];
#pragma unroll
        for (int i = 0; i < 4; ++i) C[nt][i] = fmaf(f, Tp[nt][i], C[nt][i]);
      }
    } else {
#pragma unroll
      for (int kc = 0; kc < NKC; ++kc) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          C[nt] = mfma_chunk<KIND>(a[kc], sB[(kc * NT + nt) * 64 + l], C[nt]);
        if constexpr (more) a[kc] = src.load(lane_off, anext + kc * 64);
      }
    }
  };
  auto load_a = [&](uint4* a, int pt) {
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) a[kc] = src.load(lane_off, (pt * NMT * NKC + kc) * 64);
  };
  const std::integral_constant<bool, true> yes{};
  const std::integral_constant<bool, false> no{};
  // neighbour of visiting step q (q >= 1: never the self point)
  auto nb_of = [&](int q) { return __builtin_amdgcn_readfirstlane(s_ys[w][mrhs_point(q)]); };

  if constexpr (!D2) {
    // one point ahead: one set of input and link registers
    uint4 a[NKC], b[SPT];
    load_a(a, 8);
    issue_b(b, x);  // the self point comes first
#pragma unroll 1
    for (int q9 = 0; q9 < 8; ++q9) point(q9, b, a, mrhs_point(q9 + 1), nb_of(q9 + 1), yes);
    point(8, b, a, 0, 0, no);
  } else {
    // two points ahead: two register sets, the point loop unrolled by two so
    // that every set is named at compile time; the loads of point q + 2 are
    // issued while point q is staged, and wait behind those of point q + 1
    uint4 a0[NKC], b0[SPT], a1[NKC], b1[SPT];
    load_a(a0, 8);
    issue_b(b0, x);
    load_a(a1, mrhs_point(1));
    issue_b(b1, nb_of(1));
#pragma unroll 1
    for (int q9 = 0; q9 < 6; q9 += 2) {
      point(q9, b0, a0, mrhs_point(q9 + 2), nb_of(q9 + 2), yes);
      point(q9 + 1, b1, a1, mrhs_point(q9 + 3), nb_of(q9 + 3), yes);
    }
    point(6, b0, a0, mrhs_point(8), nb_of(8), yes);
    point(7, b1, a1, 0, 0, no);
    point(8, b0, a0, 0, 0, no);
  }

  // D[m][n]: lane l holds rows 4*(l/16)+i, column l%16
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int j = (nt * 16 + (l & 15)) >> 1;
    if (j < nr) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = w * 16 + 4 * (l >> 4) + i;
        float v = C[nt][i];
        if (dag && r >= N / 2) v = -v;
        if (N % 16 == 0 || r < N) out[(((size_t)x * nrhs + rhs0 + j) * N + r) * 2 + u] = v;
      }
    }
  }
}

// Workgroup -> site order.  Consecutive workgroups are dispatched to the
// eight XCDs in turn, each with its own L2.  With xchunk = V/8 > 0, XCD k
// receives the contiguous site range [k*xchunk, (k+1)*xchunk) in lexicographic
// order (for a lattice with L3 % 8 == 0 a set of whole time slices), so that
// most neighbour inputs a workgroup reads were fetched into the same L2 by a
// workgroup that ran shortly before; xchunk = 0 keeps the identity order.
__device__ __forceinline__ int site_of_block(int b, int xchunk) {
  return xchunk > 0 ? (b & 7) * xchunk + (b >> 3) : b;
}

// Occupancy floor of the LDS-staged kernels: four wavefronts per SIMD (the
// register budget 128) where the body fits it without spilling, as measured
// from the compiler's resource report; no floor for the widest tiles.
template <int KIND, int NT>
constexpr int mrhs_min_waves() {
  return (KIND == MGC_F32 ? NT <= 4 : NT <= 6) ? 4 : 1;
}

template <int KIND, int N, int NT, int LAUX>
__global__ __launch_bounds__(mrhs_mtiles(N) * 64, (mrhs_min_waves<KIND, NT>())) void k_mrhs(
    const uint4* __restrict__ A, const float* __restrict__ ascale, const void* __restrict__ in_,
    const float* __restrict__ in_scale, float* __restrict__ out, const int* __restrict__ nbr,
    int nrhs, int rhs0, int nr, int dag, int xchunk) {
  mrhs_lds_site<KIND, N, NT, LAUX>(site_of_block((int)blockIdx.x, xchunk), A, ascale, in_,
                                  in_scale, out, nbr, nrhs, rhs0, nr, dag);
}

// XCD-block site order: XCD k = b mod 8 takes the sites with x, y and z in
// halves (k & 1, k >> 1 & 1, k >> 2 & 1) of the lattice and every t, walked
// x fastest.  Against contiguous site ranges (whole t slices for 8^4) this
// shrinks the neighbour inputs that an XCD reads from other XCDs' ranges.
// dims packs L0 | L1 << 8 | L2 << 16 | L3 << 24 (L0, L1, L2 even).
__device__ __forceinline__ int site_of_block_xblk(int b, int dims) {
  const int L0 = dims & 255, L1 = (dims >> 8) & 255, L2 = (dims >> 16) & 255;
  const int h0 = L0 >> 1, h1 = L1 >> 1, h2 = L2 >> 1;
  const int k = b & 7, i = b >> 3;
  const int x0 = (k & 1) * h0 + i % h0;
  const int x1 = ((k >> 1) & 1) * h1 + (i / h0) % h1;
  const int x2 = ((k >> 2) & 1) * h2 + (i / (h0 * h1)) % h2;
  const int x3 = i / (h0 * h1 * h2);
  return x0 + L0 * (x1 + L1 * (x2 + L2 * x3));
}

template <int KIND, int N, int NT, int LAUX>
__global__ __launch_bounds__(mrhs_mtiles(N) * 64, (mrhs_min_waves<KIND, NT>())) void k_mrhs_xblk(
    const uint4* __restrict__ A, const float* __restrict__ ascale, const void* __restrict__ in_,
    const float* __restrict__ in_scale, float* __restrict__ out, const int* __restrict__ nbr,
    int nrhs, int rhs0, int nr, int dag, int dims) {
  mrhs_lds_site<KIND, N, NT, LAUX>(site_of_block_xblk((int)blockIdx.x, dims), A, ascale, in_,
                                  in_scale, out, nbr, nrhs, rhs0, nr, dag);
}

// Occupancy floor of the two-points-ahead form, which holds a second set of
// input and link registers: three wavefronts per SIMD (the register budget
// 168) where the body fits it, as measured from the compiler's resource
// report; no floor otherwise.
template <int KIND, int NT>
constexpr int mrhs_min_waves_d2() {
  return (KIND == MGC_F32 ? NT <= 2 : NT <= 6) ? 3 : 1;
}

// XCD-block site order with the input and link loads issued two points
// ahead (variant 10): arithmetic and order of variant 9, more bytes in
// flight per workgroup at one wavefront per SIMD less.
template <int KIND, int N, int NT, int LAUX>
__global__ __launch_bounds__(mrhs_mtiles(N) * 64, (mrhs_min_waves_d2<KIND, NT>())) void k_mrhs_xblk2(
    const uint4* __restrict__ A, const float* __restrict__ ascale, const void* __restrict__ in_,
    const float* __restrict__ in_scale, float* __restrict__ out, const int* __restrict__ nbr,
    int nrhs, int rhs0, int nr, int dag, int dims, int xchunk) {
  const int x = dims ? site_of_block_xblk((int)blockIdx.x, dims)
                     : site_of_block((int)blockIdx.x, xchunk);
  mrhs_lds_site<KIND, N, NT, LAUX, true>(x, A, ascale, in_, in_scale, out, nbr, nrhs, rhs0, nr,
                                         dag);
}

// Register form: no LDS and no barrier.  Each wavefront owns MPW consecutive
// m-tiles of one site and builds the B' fragments it needs in registers
// straight from 16-byte loads of the neighbour's input (lanes l and l^1 read
// the same chunk), so wavefronts stream independently.  B' values, scale
// folding and MFMA order are those of the LDS-staged form.
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
  const int u = l & 1;
  const T* in = reinterpret_cast<const T*>(in_);

  int ys[9];
#pragma unroll
  for (int p = 0; p < 8; ++p) ys[p] = nbr[(size_t)x * 8 + p];
  ys[8] = x;
  int jj[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) jj[nt] = (nt * 16 + (l & 15)) >> 1;
  const int cl = (l >> 4) * M::KCL / 2;  // this lane's first complex within a k-chunk

  floatx4 C[MPW][NT];
#pragma unroll
  for (int mt = 0; mt < MPW; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) C[mt][nt] = {0.f, 0.f, 0.f, 0.f};

#pragma unroll 1
  for (int q9 = 0; q9 < 9; ++q9) {
    const int p = mrhs_point(q9);
    const uint4* Ap = A + (((size_t)x * 9 + p) * NMT + w * MPW) * NKC * 64 + l;
    const T* base = in + (size_t)ys[p] * nrhs * N * 2;
    floatx4 Tp[MPW][NT];
#pragma unroll
    for (int mt = 0; mt < MPW; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) Tp[mt][nt] = SCALED ? floatx4{0.f, 0.f, 0.f, 0.f} : C[mt][nt];
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      uint4 a[MPW];
#pragma unroll
      for (int mt = 0; mt < MPW; ++mt) a[mt] = Ap[((size_t)mt * NKC + kc) * 64];
      const int c0 = kc * M::KC / 2 + cl;
      const bool neg = dag && c0 >= N / 2;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        uint4 raw = make_uint4(0u, 0u, 0u, 0u);
        if (jj[nt] < nr)
          raw = *reinterpret_cast<const uint4*>(base + (((size_t)(rhs0 + jj[nt])) * N + c0) * 2);
        const uint4 b = bprime<KIND>(raw, u, neg);
#pragma unroll
        for (int mt = 0; mt < MPW; ++mt) Tp[mt][nt] = mfma_chunk<KIND>(a[mt], b, Tp[mt][nt]);
      }
    }
    if constexpr (SCALED) {
      const float sl = ascale[(size_t)x * 9 + p];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const float f =
            jj[nt] < nr ? sl * in_scale[(size_t)ys[p] * nrhs + rhs0 + jj[nt]] : 0.f;
#pragma unroll
        for (int mt = 0; mt < MPW; ++mt)
#pragma unroll
          for (int i = 0; i < 4; ++i) C[mt][nt][i] = fmaf(f, Tp[mt][nt][i], C[mt][nt][i]);
      }
    } else {
#pragma unroll
      for (int mt = 0; mt < MPW; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) C[mt][nt] = Tp[mt][nt];
    }
  }
#pragma unroll
  for (int mt = 0; mt < MPW; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int j = jj[nt];
      if (j < nr) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = (w * MPW + mt) * 16 + 4 * (l >> 4) + i;
          float val = C[mt][nt][i];
          if (dag && r >= N / 2) val = -val;
          out[(((size_t)x * nrhs + rhs0 + j) * N + r) * 2 + u] = val;
        }
      }
    }
}

// Direct-to-LDS form: two m-tiles per wavefront (N/32 wavefronts), the
// neighbour's input copied unchanged into LDS by global_load_lds_dwordx4
// (no registers, no per-slot transform), double-buffered so that one barrier
// per point suffices, and B' built from the raw 16-byte chunk at read time
// (lanes l and l^1 read the same chunk).  Each B' fragment read from LDS
// feeds two MFMAs, so LDS reads per flop are a quarter of the LDS-staged
// form's.  Raw layout per buffer: row j (input column rhs0 + j) holds the CPR
// 16-byte chunks of its N complex entries plus one pad chunk, so the eight
// rows a 16-lane group reads fall in different bank groups.  Columns j >= nr
// are copied from row 0 of the same site (in bounds) and only ever reach
// output columns that are not stored.  Arithmetic order per output element
// is that of the other forms.
template <int KIND, int N, int NT>
struct MrhsDmaGeom {
  using M = MK<KIND>;
  static constexpr int WAVES = N / 32;
  static constexpr int THREADS = WAVES * 64;
  static constexpr int NCOL = NT * 8;
  static constexpr int CPR = N * 2 * (int)sizeof(typename M::T) / 16;  // chunks per row
  static constexpr int RSL = CPR + 1;                                    // slots per row
  static constexpr int SLOTS = NCOL * RSL;
  static constexpr int IPW = (SLOTS + THREADS - 1) / THREADS;  // copies per wavefront
  static constexpr int BUF = IPW * THREADS;                    // slots per buffer
};

template <int KIND, int N, int NT, int LAUX>
__global__ __launch_bounds__(N / 32 * 64) void k_mrhs_dma(
    const uint4* __restrict__ A, const float* __restrict__ ascale, const void* __restrict__ in_,
    const float* __restrict__ in_scale, float* __restrict__ out, const int* __restrict__ nbr,
    int nrhs, int rhs0, int nr, int dag, int xchunk) {
  using M = MK<KIND>;
  using G = MrhsDmaGeom<KIND, N, NT>;
  typedef typename M::T T;
  constexpr int NMT = N / 16;
  constexpr int NKC = 2 * N / M::KC;
  constexpr bool SCALED = KIND != MGC_F32;
  static_assert(G::NCOL <= G::THREADS, "one thread per column for the scale table");
  __shared__ uint4 sraw[2][G::BUF];
  __shared__ float s_fac[SCALED ? 9 : 1][G::NCOL];

  const int x = site_of_block((int)blockIdx.x, xchunk);
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  const int u = l & 1;
  const char* in = reinterpret_cast<const char*>(in_);
  constexpr size_t ROWB = (size_t)N * 2 * sizeof(T);  // bytes of one input column

  int ys[9];
#pragma unroll
  for (int p = 0; p < 8; ++p) ys[p] = nbr[(size_t)x * 8 + p];
  ys[8] = x;
  if constexpr (SCALED) {
    if ((int)threadIdx.x < G::NCOL) {
      const int j = threadIdx.x;
#pragma unroll
      for (int p = 0; p < 9; ++p)
        s_fac[p][j] =
            j < nr ? ascale[(size_t)x * 9 + p] * in_scale[(size_t)ys[p] * nrhs + rhs0 + j] : 0.f;
    }
  }
  // byte offset, within a site's input block, of the chunk this lane copies
  // in its i-th copy (rows >= nr and pad slots copy row 0, chunk 0)
  auto src_off = [&](int i) -> int {
    const int slot = (i * G::WAVES + w) * 64 + l;
    const int j = slot / G::RSL, c16 = slot - j * G::RSL;
    return (j < nr && c16 < G::CPR) ? (int)((rhs0 + j) * ROWB) + c16 * 16 : (int)(rhs0 * ROWB);
  };
  auto issue = [&](int p, int bf) {
    const char* site = in + (size_t)ys[p] * nrhs * ROWB;
#pragma unroll
    for (int i = 0; i < G::IPW; ++i)
      __builtin_amdgcn_global_load_lds(
          (const void*)(site + src_off(i)),
          (__attribute__((address_space(3))) void*)&sraw[bf][(i * G::WAVES + w) * 64], 16, 0, 0);
  };

  // one link source per m-tile of this wavefront (lane positions lo0, lo1)
  const int lo0 = (2 * w) * NKC * 64 + l, lo1 = (2 * w + 1) * NKC * 64 + l;
  const uint4* site_links = A + (size_t)x * 9 * NMT * NKC * 64;
  const LinkSrc<LAUX> src0(site_links, 9 * NMT * NKC * 64 * 16, lo0);
  const LinkSrc<LAUX> src1(site_links, 9 * NMT * NKC * 64 * 16, lo1);
  uint4 a0[NKC], a1[NKC];
#pragma unroll
  for (int kc = 0; kc < NKC; ++kc) {
    a0[kc] = src0.load(lo0, (8 * NMT * NKC + kc) * 64);
    a1[kc] = src1.load(lo1, (8 * NMT * NKC + kc) * 64);
  }
  issue(mrhs_point(0), 0);

  floatx4 C0[NT], C1[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) C0[nt] = C1[nt] = floatx4{0.f, 0.f, 0.f, 0.f};

#pragma unroll 1
  for (int q9 = 0; q9 < 9; ++q9) {
    const int p = mrhs_point(q9);
    const int bf = q9 & 1;
    // this point's copies (and link fragments) have landed; every wavefront
    // is done reading the other buffer
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const bool more = q9 < 8;
    const int pn = more ? mrhs_point(q9 + 1) : 0;
    if (more) issue(pn, bf ^ 1);
    const int anext = pn * NMT * NKC * 64;
    floatx4 T0[NT], T1[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      T0[nt] = SCALED ? floatx4{0.f, 0.f, 0.f, 0.f} : C0[nt];
      T1[nt] = SCALED ? floatx4{0.f, 0.f, 0.f, 0.f} : C1[nt];
    }
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      const int chunk = kc * 4 + (l >> 4);
      const int c0 = (kc * M::KC + (l >> 4) * M::KCL) >> 1;
      const bool neg = dag && c0 >= N / 2;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int j = nt * 8 + ((l & 15) >> 1);
        const uint4 b = bprime<KIND>(sraw[bf][j * G::RSL + chunk], u, neg);
        T0[nt] = mfma_chunk<KIND>(a0[kc], b, T0[nt]);
        T1[nt] = mfma_chunk<KIND>(a1[kc], b, T1[nt]);
      }
      if (more) {
        a0[kc] = src0.load(lo0, anext + kc * 64);
        a1[kc] = src1.load(lo1, anext + kc * 64);
      }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      if constexpr (SCALED) {
        const float f = s_fac[p][(nt * 16 + (l & 15)) >> 1];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          C0[nt][i] = fmaf(f, T0[nt][i], C0[nt][i]);
          C1[nt][i] = fmaf(f, T1[nt][i], C1[nt][i]);
        }
      } else {
        C0[nt] = T0[nt];
        C1[nt] = T1[nt];
      }
    }
  }
  // D[m][n]: lane l holds rows 4*(l/16)+i of each m-tile, column l%16
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int j = (nt * 16 + (l & 15)) >> 1;
    if (j < nr) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = (2 * w + h) * 16 + 4 * (l >> 4) + i;
          float v = h ? C1[nt][i] : C0[nt][i];
          if (dag && r >= N / 2) v = -v;
          out[(((size_t)x * nrhs + rhs0 + j) * N + r) * 2 + u] = v;
        }
      }
    }
  }
}

// Several coarse lattices of (possibly) different volume and one nbasis in a
// single launch: block b belongs to the lattice whose site range
// [first[l], first[l+1]) contains b.
constexpr int MRHS_BATCH_MAX = 8;
struct MrhsBatch {
  const uint4* A[MRHS_BATCH_MAX];
  const float* ascale[MRHS_BATCH_MAX];
  const void* in[MRHS_BATCH_MAX];
  const float* in_scale[MRHS_BATCH_MAX];
  float* out[MRHS_BATCH_MAX];
  const int* nbr[MRHS_BATCH_MAX];
  int dag[MRHS_BATCH_MAX];
  int first[MRHS_BATCH_MAX + 1];
  int count;
};

template <int KIND, int N, int NT, int LAUX>
__global__ __launch_bounds__(mrhs_mtiles(N) * 64, (mrhs_min_waves<KIND, NT>())) void k_mrhs_batched(MrhsBatch bt, int nrhs, int rhs0,
                                                              int nr, int xchunk) {
  const int b = site_of_block((int)blockIdx.x, xchunk);
  int l = 0;
  while (l + 1 < bt.count && b >= bt.first[l + 1]) ++l;
  mrhs_lds_site<KIND, N, NT, LAUX>(b - bt.first[l], bt.A[l], bt.ascale[l], bt.in[l],
                                  bt.in_scale[l], bt.out[l], bt.nbr[l], nrhs, rhs0, nr,
                                  bt.dag[l]);
}

// Persistent form: a grid of about one resident workgroup per slot (CUs x
// occupancy) walks the sites of every lattice in the batch, iteration
// it = blockIdx.x + k * gridDim.x.  The loads of the next site's first point
// are issued during the current site's last point, so a workgroup's memory
// pipeline does not drain between sites.  Arithmetic per site is that of
// mrhs_lds_site.
template <int KIND, int NT>
constexpr int mrhs_persistent_min_waves() {
  return KIND == MGC_F32 ? (NT <= 3 ? 3 : 1) : (NT <= 3 ? 4 : NT <= 6 ? 3 : 1);
}

template <int KIND, int N, int NT, int LAUX>
__global__ __launch_bounds__(mrhs_mtiles(N) * 64, (mrhs_persistent_min_waves<KIND, NT>())) void k_mrhs_persistent(
    MrhsBatch bt, int nrhs, int rhs0, int nr, int xchunk) {
  using M = MK<KIND>;
  typedef typename M::T T;
  constexpr int NMT = mrhs_mtiles(N);
  constexpr int NKC = 2 * NMT * 16 / M::KC;
  constexpr int THREADS = NMT * 64;
  constexpr int NSLOT = NKC * NT * 64;
  constexpr int SPT = NSLOT / THREADS;
  constexpr int NCOL = NT * 8;
  constexpr int SITE_FRAGS = 9 * NMT * NKC * 64;  // uint4 fragments of one site's links
  constexpr bool SCALED = KIND != MGC_F32;
  static_assert(NCOL <= THREADS, "one thread per column for the scale table");
  __shared__ uint4 sB[NSLOT];
  __shared__ float s_fac[2][SCALED ? 9 : 1][NCOL];

  const int total = bt.first[bt.count];
  int it = blockIdx.x;
  if (it >= total) return;
  const int w = threadIdx.x >> 6;
  const int l = threadIdx.x & 63;
  const int u = l & 1;
  const int lane_off = w * NKC * 64 + l;

  auto decode = [&](int i, int& lat, int& x) {
    const int b = site_of_block(i, xchunk);
    lat = 0;
    while (lat + 1 < bt.count && b >= bt.first[lat + 1]) ++lat;
    x = b - bt.first[lat];
  };
  auto load_ys = [&](int lat, int x, int* ys) {
#pragma unroll
    for (int p = 0; p < 8; ++p) ys[p] = bt.nbr[lat][(size_t)x * 8 + p];
    ys[8] = x;
  };
  auto fill_fac = [&](int buf, int lat, int x, const int* ys) {
    if constexpr (SCALED) {
      if (threadIdx.x < NCOL) {
        const int j = threadIdx.x;
#pragma unroll
        for (int p = 0; p < 9; ++p)
          s_fac[buf][p][j] = j < nr ? bt.ascale[lat][(size_t)x * 9 + p] *
                                          bt.in_scale[lat][(size_t)ys[p] * nrhs + rhs0 + j]
                                    : 0.f;
      }
    }
  };

  int soff[SPT];
  bool shi[SPT];  // chunk in the lower chirality half
#pragma unroll
  for (int q = 0; q < SPT; ++q) {
    const int s = threadIdx.x + q * THREADS;
    const int nt = (s >> 6) % NT, kc = (s >> 6) / NT;
    const int j = (nt * 16 + (l & 15)) >> 1;
    const int c0 = (kc * M::KC + (l >> 4) * M::KCL) >> 1;
    soff[q] = (j < nr && c0 < N) ? ((rhs0 + j) * N + c0) * 2 : -1;
    shi[q] = c0 >= N / 2;
  }
  uint4 a[NKC], b_n[SPT];
  auto issue_b = [&](const T* base) {
#pragma unroll
    for (int q = 0; q < SPT; ++q)
      if (soff[q] >= 0) b_n[q] = *reinterpret_cast<const uint4*>(base + soff[q]);
      else b_n[q] = make_uint4(0u, 0u, 0u, 0u);
  };

  int cl, cx;
  decode(it, cl, cx);
  int ys[9];
  load_ys(cl, cx, ys);
  int buf = 0;
  fill_fac(0, cl, cx, ys);
  LinkSrc<LAUX> src(bt.A[cl] + (size_t)cx * SITE_FRAGS, SITE_FRAGS * 16, lane_off);
#pragma unroll
  for (int kc = 0; kc < NKC; ++kc) a[kc] = src.load(lane_off, (8 * NMT * NKC + kc) * 64);
  const T* cin = reinterpret_cast<const T*>(bt.in[cl]);
  issue_b(cin + (size_t)ys[8] * nrhs * N * 2);

  floatx4 C[NT];
  int nl = cl, nx = cx;
  int nys[9];
  while (true) {
    const int nit = it + (int)gridDim.x;
    const bool has_next = nit < total;
    const int cdag = bt.dag[cl];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) C[nt] = {0.f, 0.f, 0.f, 0.f};

#pragma unroll 1
    for (int q9 = 0; q9 < 9; ++q9) {
      const int p = mrhs_point(q9);
      uint4 b[SPT];
#pragma unroll
      for (int q = 0; q < SPT; ++q) b[q] = b_n[q];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < SPT; ++q)
        sB[threadIdx.x + q * THREADS] = bprime<KIND>(b[q], u, cdag && shi[q]);
      __syncthreads();
      // where the next step's operands come from: this site's next point, or
      // the next site's first point (the self point)
      const T* bnext = cin;
      int ynext = 0, anext = 8 * NMT * NKC * 64;
      bool more = true;
      if (q9 < 8) {
        const int pn = mrhs_point(q9 + 1);
        ynext = ys[pn];
        anext = pn * NMT * NKC * 64;
      } else if (has_next) {
        decode(nit, nl, nx);
        load_ys(nl, nx, nys);
        fill_fac(buf ^ 1, nl, nx, nys);
        src = LinkSrc<LAUX>(bt.A[nl] + (size_t)nx * SITE_FRAGS, SITE_FRAGS * 16, lane_off);
        cin = reinterpret_cast<const T*>(bt.in[nl]);
        bnext = cin;
        ynext = nys[8];
      } else {
        more = false;
      }
      if (more) issue_b(bnext + (size_t)ynext * nrhs * N * 2);
      auto reload = [&](int kc) {
        if (more) a[kc] = src.load(lane_off, anext + kc * 64);
      };
      if constexpr (SCALED) {
        floatx4 Tp[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) Tp[nt] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
#pragma unroll
          for (int nt = 0; nt < NT; ++nt)
            Tp[nt] = mfma_chunk<KIND>(a[kc], sB[(kc * NT + nt) * 64 + l], Tp[nt]);
          reload(kc);
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const float f = s_fac[buf][p][(nt * 16 + (l & 15)) >> 1];
#pragma unroll
          for (int i = 0; i < 4; ++i) C[nt][i] = fmaf(f, Tp[nt][i], C[nt][i]);
        }
      } else {
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
#pragma unroll
          for (int nt = 0; nt < NT; ++nt)
            C[nt] = mfma_chunk<KIND>(a[kc], sB[(kc * NT + nt) * 64 + l], C[nt]);
          reload(kc);
        }
      }
    }
    float* out = bt.out[cl];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int j = (nt * 16 + (l & 15)) >> 1;
      if (j < nr) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = w * 16 + 4 * (l >> 4) + i;
          float v = C[nt][i];
          if (cdag && r >= N / 2) v = -v;
          if (N % 16 == 0 || r < N) out[(((size_t)cx * nrhs + rhs0 + j) * N + r) * 2 + u] = v;
        }
      }
    }
    if (!has_next) break;
    it = nit;
    cl = nl;
    cx = nx;
#pragma unroll
    for (int p = 0; p < 9; ++p) ys[p] = nys[p];
    buf ^= 1;
  }
}

// kernel variants: 0 = LDS-staged B', 1 = LDS-staged B' with non-temporal
// link loads, 2 = register B' with two m-tiles per wavefront, 3 and 4 =
// variants 0 and 1 with the XCD-contiguous site order (identity order when
// the site count is not a multiple of 8), 5 = persistent LDS-staged form,
// 6 = persistent with non-temporal link loads and the XCD-contiguous order,
// 7 = direct-to-LDS input with two m-tiles per wavefront, 8 = variant 7 with
// non-temporal link loads and the XCD-contiguous order, 9 = variant 1 with
// the XCD-block site order (variant 4 where the lattice does not allow it),
// 10 = variant 9 with the loads two points ahead
constexpr int MRHS_VARIANTS = 11;
// variant used without tuning: the LDS-staged form, non-temporal links, XCD-block
// site order (the contiguous XCD order where the lattice does not allow it)
constexpr int MRHS_DEFAULT_VARIANT = 9;
static_assert(MRHS_VARIANTS <= 16, "the autotuner times at most 16 variants");
static int variant_laux(int v) {
  return (v == 1 || v == 4 || v == 6 || v == 8 || v == 9 || v == 10) ? 2 : 0;
}
static int variant_xchunk(int v, int blocks) {
  return ((v == 3 || v == 4 || v == 6 || v == 8) && blocks % 8 == 0) ? blocks / 8 : 0;
}

// persistent grid: every resident workgroup slot of the device, a multiple of
// 8 so that a workgroup keeps its XCD across iterations
template <typename K>
static int persistent_grid(K kernel, int threads, int sites) {
  static int slots = 0;  // per kernel instantiation
  if (!slots) {
    int dev = 0, cus = 0, per_cu = 0;
    MGC_HIP(hipGetDevice(&dev));
    MGC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    MGC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    slots = cus * (per_cu > 0 ? per_cu : 1);
    slots = slots >= 8 ? slots / 8 * 8 : slots;
  }
  return sites < slots ? sites : slots;
}

template <int KIND, int N, int NT>
static void launch_persistent(int variant, const MrhsBatch& bt, int nrhs, int rhs0, int nr,
                              hipStream_t s) {
  const int total = bt.first[bt.count];
  const int xc = variant_xchunk(variant, total);
  const int b = mrhs_mtiles(N) * 64;
  if (variant_laux(variant) == 2) {
    auto k = k_mrhs_persistent<KIND, N, NT, 2>;
    k<<<persistent_grid(k, b, total), b, 0, s>>>(bt, nrhs, rhs0, nr, xc);
  } else {
    auto k = k_mrhs_persistent<KIND, N, NT, 0>;
    k<<<persistent_grid(k, b, total), b, 0, s>>>(bt, nrhs, rhs0, nr, xc);
  }
}

template <int KIND, int N, int NT>
static void launch_variant(int variant, int V, int dims, const uint4* A, const float* as,
                           const void* in, const float* is, float* out, const int* nbr, int nrhs,
                           int rhs0, int nr, int dag, hipStream_t s) {
  // the register and direct-to-LDS forms (2, 7, 8) are built for nbasis 32
  // and 64 only; elsewhere they run as the LDS-staged forms 0, 3 and 4
  if constexpr (!mrhs_all_variants(N))
    variant = variant == 2 ? 0 : variant == 7 ? 3 : variant == 8 ? 4 : variant;
  const dim3 g(V);
  const int xc = variant_xchunk(variant, V);
  const int b = mrhs_mtiles(N) * 64;
  if (variant == 5 || variant == 6) {
    MrhsBatch bt{};
    bt.A[0] = A;
    bt.ascale[0] = as;
    bt.in[0] = in;
    bt.in_scale[0] = is;
    bt.out[0] = out;
    bt.nbr[0] = nbr;
    bt.dag[0] = dag;
    bt.first[0] = 0;
    bt.first[1] = V;
    bt.count = 1;
    launch_persistent<KIND, N, NT>(variant, bt, nrhs, rhs0, nr, s);
    return;
  }
  if constexpr (mrhs_all_variants(N)) {
    if (variant == 2) {
      k_mrhs_reg<KIND, N, NT, 2><<<g, N / 32 * 64, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0,
                                                          nr, dag);
      return;
    }
  }
  if (variant == 10) {
    k_mrhs_xblk2<KIND, N, NT, 2><<<g, b, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr, dag,
                                                  dims, dims ? 0 : variant_xchunk(4, V));
    return;
  }
  if (variant == 9) {
    if (dims) {
      k_mrhs_xblk<KIND, N, NT, 2><<<g, b, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr, dag,
                                                   dims);
    } else {
      k_mrhs<KIND, N, NT, 2><<<g, b, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr, dag,
                                              variant_xchunk(4, V));
    }
    return;
  }
  if constexpr (mrhs_all_variants(N)) {
    if (variant == 7 || variant == 8) {
      if (variant_laux(variant) == 2)
        k_mrhs_dma<KIND, N, NT, 2><<<g, N / 32 * 64, 0, s>>>(A, as, in, is, out, nbr, nrhs,
                                                            rhs0, nr, dag, xc);
      else
        k_mrhs_dma<KIND, N, NT, 0><<<g, N / 32 * 64, 0, s>>>(A, as, in, is, out, nbr, nrhs,
                                                            rhs0, nr, dag, xc);
      return;
    }
  }
  if (variant_laux(variant) == 2)
    k_mrhs<KIND, N, NT, 2><<<g, b, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, xc);
  else
    k_mrhs<KIND, N, NT, 0><<<g, b, 0, s>>>(A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, xc);
}

template <int KIND, int N>
static void launch_mrhs_n(int variant, int NT, int V, int dims, const uint4* A, const float* as,
                          const void* in, const float* is, float* out, const int* nbr, int nrhs,
                          int rhs0, int nr, int dag, hipStream_t s) {
  switch (NT) {
#define CASE(nt)                                                                              \
  case nt:                                                                                    \
    launch_variant<KIND, N, nt>(variant, V, dims, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag, \
                                s);                                                           \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(6) CASE(8)
#undef CASE
    default:
      throw Error(MGC_ERR_ARG, "internal: unsupported column tile count");
  }
  MGC_LAUNCH_CHECK();
}

// batched launch over several lattices: the LDS-staged variants
template <int KIND, int N, int NT>
static void launch_batched_nt(int variant, int blocks, const MrhsBatch& bt, int nrhs, int rhs0,
                              int nr, hipStream_t s) {
  const int xc = variant_xchunk(variant, blocks);
  const int b = mrhs_mtiles(N) * 64;
  if (variant == 5 || variant == 6)
    launch_persistent<KIND, N, NT>(variant, bt, nrhs, rhs0, nr, s);
  else if (variant_laux(variant) == 2)
    k_mrhs_batched<KIND, N, NT, 2><<<blocks, b, 0, s>>>(bt, nrhs, rhs0, nr, xc);
  else
    k_mrhs_batched<KIND, N, NT, 0><<<blocks, b, 0, s>>>(bt, nrhs, rhs0, nr, xc);
}
template <int KIND, int N>
static void launch_batched_n(int variant, int NT, int blocks, const MrhsBatch& bt, int nrhs,
                             int rhs0, int nr, hipStream_t s) {
  switch (NT) {
#define CASE(nt)                                                            \
  case nt:                                                                  \
    launch_batched_nt<KIND, N, nt>(variant, blocks, bt, nrhs, rhs0, nr, s); \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(6) CASE(8)
#undef CASE
    default:
      throw Error(MGC_ERR_ARG, "internal: unsupported column tile count");
  }
  MGC_LAUNCH_CHECK();
}

// Entry points of one nbasis, instantiated in mrhs_n<N>.hip: the single
// lattice launch of `variant` and the batched launch over several lattices.
template <int N>
void mrhs_launch(int mp, int variant, int NT, int V, int dims, const uint4* A, const float* as,
                 const void* in, const float* is, float* out, const int* nbr, int nrhs,
                 int rhs0, int nr, int dag, hipStream_t s) {
  if (mp == MGC_F32)
    launch_mrhs_n<MGC_F32, N>(variant, NT, V, dims, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag,
                              s);
  else if (mp == MGC_F16)
    launch_mrhs_n<MGC_F16, N>(variant, NT, V, dims, A, as, in, is, out, nbr, nrhs, rhs0, nr, dag,
                              s);
  else
    launch_mrhs_n<MGC_BF16, N>(variant, NT, V, dims, A, as, in, is, out, nbr, nrhs, rhs0, nr,
                               dag, s);
}

template <int N>
void mrhs_launch_batched(int mp, int variant, int NT, int blocks, const MrhsBatch& bt, int nrhs,
                         int rhs0, int nr, hipStream_t s) {
  if (mp == MGC_F32)
    launch_batched_n<MGC_F32, N>(variant, NT, blocks, bt, nrhs, rhs0, nr, s);
  else if (mp == MGC_F16)
    launch_batched_n<MGC_F16, N>(variant, NT, blocks, bt, nrhs, rhs0, nr, s);
  else
    launch_batched_n<MGC_BF16, N>(variant, NT, blocks, bt, nrhs, rhs0, nr, s);
}

#define MGC_MRHS_FOR_EACH_N(X) X(8) X(16) X(24) X(32) X(40) X(48) X(56) X(64)

#define MGC_MRHS_DECLARE(n)                                                                   \
  extern template void mrhs_launch<n>(int, int, int, int, int, const uint4*, const float*,    \
                                      const void*, const float*, float*, const int*, int, int, \
                                      int, int, hipStream_t);                                   \
  extern template void mrhs_launch_batched<n>(int, int, int, int, const MrhsBatch&, int, int,  \
                                              int, hipStream_t);
MGC_MRHS_FOR_EACH_N(MGC_MRHS_DECLARE)
#undef MGC_MRHS_DECLARE

}  // namespace mgc
