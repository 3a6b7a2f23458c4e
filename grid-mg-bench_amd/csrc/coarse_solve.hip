// Coarse solves on several lattices at once: one lockstep batched GMRES over
// the right-hand sides of every lattice, whose operator application is a
// single multi-lattice launch of the multi-rhs stencil
// (mgc_coarse_op_M_mrhs_batched).  This is the batched form of the coarsest
// solve of the reference's preconditioner (a GMRES on each coarse operator,
// include/MultigridPreconditioner.h:566-641 and include/VCycleRework.h:84-189,
// one lattice at a time there).
//
// The solver's vectors hold all lattices side by side as extra columns:
// [Vmax][K * nrhs][N] FP32, lattice k in columns k*nrhs .. k*nrhs + nrhs - 1
// of its first V_k sites and zero in the padding sites beyond.  The batched
// BLAS then reduces per (lattice, rhs) column, and the padding stays zero
// (the operator writes zeros there).  Each application gathers every
// lattice's columns into its own [V_k][nrhs][N] input (quantised to FP16 /
// BF16 with one scale per (site, rhs) row for a half-precision stencil), runs
// the one batched launch, and scatters the outputs back.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>
#include <vector>

#include "bkrylov.hpp"
#include "coarse_op.hpp"

namespace mgc {

constexpr int CS_MAX = 8;  // lattices per batched solve (the batched stencil's limit)

static void cs_ok(int st) {
  if (st != MGC_OK) throw Error(st, mgc_last_error());
}
static unsigned cs_grid(size_t n) { return grid_for(n, 256, 8192); }

struct CsLayout {
  int K, nrhs, N, Vmax;
  int V[CS_MAX];
  int first[CS_MAX + 1];  // prefix sums of V_k * nrhs (rows of the dense inputs)
  void* dense[CS_MAX];    // per-lattice [V_k][nrhs][N]: FP32, or half with scales
  float* scale[CS_MAX];   // per-lattice [V_k][nrhs] (half precision only)
};

__device__ __forceinline__ int cs_lattice(const CsLayout& L, int row) {
  int k = 0;
#pragma unroll
  for (int i = 1; i < CS_MAX; ++i)
    if (i < L.K && row >= L.first[i]) k = i;
  return k;
}

// padded columns -> per-lattice dense rows, one wavefront per (lattice, site,
// rhs) row; KIND = MGC_F32 copies, MGC_F16 / MGC_BF16 quantise the row with
// scale = max |component| (1 for a zero row), as the stencil's own quantiser
template <int KIND>
__global__ __launch_bounds__(256) void k_cs_gather(const float2* __restrict__ padded, CsLayout L) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= L.first[L.K]) return;
  const int k = cs_lattice(L, row);
  const int local = row - L.first[k];
  const int x = local / L.nrhs, r = local - x * L.nrhs;
  const float2* src = padded + ((size_t)x * L.K * L.nrhs + (size_t)k * L.nrhs + r) * L.N;
  if (KIND == MGC_F32) {
    float2* dst = reinterpret_cast<float2*>(L.dense[k]) + (size_t)local * L.N;
    for (int c = lane; c < L.N; c += 64) dst[c] = src[c];
    return;
  }
  typedef typename std::conditional<KIND == MGC_F16, _Float16, __bf16>::type T;
  float m = 0.f;
  for (int c = lane; c < L.N; c += 64) m = fmaxf(m, fmaxf(fabsf(src[c].x), fabsf(src[c].y)));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  const float s = m > 0.f ? m : 1.f;
  const float inv = 1.f / s;
  T* o = reinterpret_cast<T*>(L.dense[k]) + (size_t)local * 2 * L.N;
  for (int c = lane; c < L.N; c += 64) {
    o[2 * c] = (T)(src[c].x * inv);
    o[2 * c + 1] = (T)(src[c].y * inv);
  }
  if (lane == 0) L.scale[k][local] = s;
}

// per-lattice FP32 outputs -> padded columns, zeros in the padding sites;
// one thread per padded element
__global__ __launch_bounds__(256) void k_cs_scatter(const CsLayout L, float2* __restrict__ padded,
                                                   const float2* const* __restrict__ src) {
  const size_t cols = (size_t)L.K * L.nrhs;
  const size_t total = (size_t)L.Vmax * cols * L.N;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % L.N);
    const size_t q = i / L.N;
    const int col = (int)(q % cols);
    const int x = (int)(q / cols);
    const int k = col / L.nrhs, r = col - k * L.nrhs;
    padded[i] = x < L.V[k] ? src[k][((size_t)x * L.nrhs + r) * L.N + c] : float2{0.f, 0.f};
  }
}

// the K coarse operators on the padded layout, one batched stencil launch
struct MultiLatticeOp : BLinOp<float> {
  std::vector<mgc_coarse_op*> ops;
  int mprec;
  CsLayout L{};
  DevBuf<unsigned char> dense_in;
  DevBuf<float> scales;
  DevBuf<cplx<float>> dense_out;
  DevBuf<const float2*> out_ptrs;
  std::vector<const void*> in_v;
  std::vector<const float*> sc_v;
  std::vector<void*> out_v;

  MultiLatticeOp(mgc_coarse_op* const* o, int K, int nrhs, int mp, hipStream_t s)
      : ops(o, o + K), mprec(mp) {
    L.K = K;
    L.nrhs = nrhs;
    L.N = o[0]->N;
    L.Vmax = 0;
    L.first[0] = 0;
    for (int k = 0; k < K; ++k) {
      L.V[k] = o[k]->V;
      L.Vmax = std::max(L.Vmax, o[k]->V);
      L.first[k + 1] = L.first[k] + o[k]->V * nrhs;
    }
    sh.nrhs = K * nrhs;
    sh.dof = L.N;
    sh.n = (size_t)L.Vmax * sh.nrhs * L.N;
    const size_t rows = (size_t)L.first[K];
    const size_t elem = mp == MGC_F32 ? 8 : 4;  // bytes per complex in the stencil's input
    dense_in.alloc(rows * L.N * elem);
    if (mp != MGC_F32) scales.alloc(rows);
    dense_out.alloc(rows * L.N);
    std::vector<const float2*> hp(K);
    for (int k = 0; k < K; ++k) {
      L.dense[k] = dense_in.p + (size_t)L.first[k] * L.N * elem;
      L.scale[k] = mp == MGC_F32 ? nullptr : scales.p + L.first[k];
      hp[k] = reinterpret_cast<const float2*>(dense_out.p + (size_t)L.first[k] * L.N);
      in_v.push_back(L.dense[k]);
      sc_v.push_back(L.scale[k]);
      out_v.push_back(dense_out.p + (size_t)L.first[k] * L.N);
    }
    out_ptrs.alloc(K);
    MGC_HIP(hipMemcpyAsync(out_ptrs.p, hp.data(), K * sizeof(const float2*),
                           hipMemcpyHostToDevice, s));
    MGC_HIP(hipStreamSynchronize(s));
  }

  void gather(const cplx<float>* padded, hipStream_t s) {
    const int rows = L.first[L.K];
    const int g = (rows + 3) / 4;
    const float2* p = reinterpret_cast<const float2*>(padded);
    if (mprec == MGC_F32) k_cs_gather<MGC_F32><<<g, 256, 0, s>>>(p, L);
    else if (mprec == MGC_F16) k_cs_gather<MGC_F16><<<g, 256, 0, s>>>(p, L);
    else k_cs_gather<MGC_BF16><<<g, 256, 0, s>>>(p, L);
    MGC_LAUNCH_CHECK();
  }
  void scatter(cplx<float>* padded, hipStream_t s) {
    k_cs_scatter<<<cs_grid(sh.n), 256, 0, s>>>(L, reinterpret_cast<float2*>(padded),
                                               out_ptrs.p);
    MGC_LAUNCH_CHECK();
  }
  void apply(const cplx<float>* in, cplx<float>* out, hipStream_t s) override {
    gather(in, s);
    cs_ok(mgc_coarse_op_M_mrhs_batched(ops.data(), L.K, in_v.data(),
                                    mprec == MGC_F32 ? nullptr : sc_v.data(), mprec, out_v.data(),
                                    L.nrhs, mprec, 0, (void*)s));
    scatter(out, s);
  }
};

}  // namespace mgc

using namespace mgc;

extern "C" {

int mgc_coarse_solve_batched(mgc_coarse_op* const* ops, int count, const void* const* b,
                             void* const* x, int nrhs, int mrhs_prec, double tol, int maxit,
                             int restart, mgc_solve_stats* stats, void* stream) {
  return guarded([&] {
    MGC_CHECK(ops && b && x && stats, MGC_ERR_ARG, "null argument");
    MGC_CHECK(count >= 1 && count <= CS_MAX, MGC_ERR_ARG, "count must be 1..8");
    MGC_CHECK(nrhs >= 1 && count * nrhs <= BBlas::RMAX, MGC_ERR_ARG,
              "count * nrhs must be 1..64 (columns of one lockstep batch)");
    MGC_CHECK(mrhs_prec == MGC_F32 || mrhs_prec == MGC_F16 || mrhs_prec == MGC_BF16,
              MGC_ERR_PRECISION, "multi-rhs precision must be F32, F16 or BF16");
    MGC_CHECK(tol > 0.0 && maxit >= 1 && restart >= 1 && restart < BBlas::KMAX, MGC_ERR_ARG,
              "tol > 0, maxit >= 1 and 1 <= restart < 64 needed");
    for (int k = 0; k < count; ++k) {
      MGC_CHECK(ops[k] && b[k] && x[k], MGC_ERR_ARG, "null lattice in the batch");
      MGC_CHECK(ops[k]->links_set, MGC_ERR_STATE, "links not set");
      MGC_CHECK(ops[k]->N == ops[0]->N, MGC_ERR_SHAPE,
                "every lattice in a batch needs the same nbasis");
      MGC_CHECK(ops[k]->N >= 8 && ops[k]->N <= 64 && ops[k]->N % 8 == 0, MGC_ERR_SHAPE,
                "multi-rhs stencil supports nbasis 8, 16, ..., 64");
    }
    hipStream_t s = as_stream(stream);
    const auto t0 = std::chrono::steady_clock::now();
    MultiLatticeOp A(ops, count, nrhs, mrhs_prec, s);
    const size_t n = A.sh.n;
    DevBuf<cplx<float>> pb(n), px(n), pr(n);
    // right-hand sides into the padded layout: the dense FP32 inputs are the
    // scatter's sources
    {
      std::vector<const float2*> hp(count);
      for (int k = 0; k < count; ++k) hp[k] = reinterpret_cast<const float2*>(b[k]);
      DevBuf<const float2*> dp(count);
      MGC_HIP(hipMemcpyAsync(dp.p, hp.data(), count * sizeof(const float2*),
                             hipMemcpyHostToDevice, s));
      k_cs_scatter<<<cs_grid(n), 256, 0, s>>>(A.L, reinterpret_cast<float2*>(pb.p), dp.p);
      MGC_LAUNCH_CHECK();
      MGC_HIP(hipStreamSynchronize(s));
    }
    BBlas bl(s);
    BGmres<float> solver(A.sh, restart, false);
    const BKrylovStats st = solver.solve(bl, A, nullptr, pb.p, px.p, true, tol, maxit);
    // residual b - M x recomputed with the FP32 stencil (the F16 / BF16 solve's
    // operator rounding shows here), per (lattice, rhs) column
    std::vector<double> bn(A.sh.nrhs), rn(A.sh.nrhs);
    bnorm2<float>(bl, pb.p, A.sh, bn.data());
    {
      MultiLatticeOp A32(ops, count, nrhs, MGC_F32, s);
      A32.apply(px.p, pr.p, s);
      vsub<float>(pb.p, pr.p, pr.p, n, s);
      bnorm2<float>(bl, pr.p, A.sh, rn.data());
    }
    // solutions back to the callers' [V_k][nrhs][N] FP32 vectors
    {
      CsLayout Lx = A.L;
      for (int k = 0; k < count; ++k) Lx.dense[k] = x[k];
      k_cs_gather<MGC_F32><<<(Lx.first[count] + 3) / 4, 256, 0, s>>>(
          reinterpret_cast<const float2*>(px.p), Lx);
      MGC_LAUNCH_CHECK();
    }
    MGC_HIP(hipStreamSynchronize(s));
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int c = 0; c < A.sh.nrhs; ++c) {
      mgc_solve_stats& o = stats[c];
      o.iterations = st.converged[c] ? st.iters_rhs[c] : st.iters;
      o.converged = st.converged[c];
      o.precond_applies = 0;
      o.true_residual = bn[c] > 0.0 ? std::sqrt(rn[c] / bn[c]) : 0.0;
      o.seconds = secs;
    }
  });
}

}  // extern "C"
