// Bit-equality of the fused re-orthogonalisation passes with the two-kernel
// forms they replace, on the GPU:
//   vmultisub_dot_dev  ==  vmultisub_dev, then vmultidot_dev(with_norm)
//   bmultisub_dots_dev ==  bmultisub_dev, then bdots_dev(with_norm)
// for FP64 and FP32 vectors (16-byte paired and unpaired loads), and basis
// sizes k on both sides of every kernel template (8, 16, 24) and of the rows
// the fused kernels keep in LDS between their two passes.  The updated w and
// every inner product / norm must agree to the bit.  k = 25 must be declined
// with nothing written.  Prints "ALL PASSED" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bblas.hpp"
#include "blas.hpp"

using namespace mgc;

namespace {

int g_fail = 0;

#define HIPOK(x)                                                               \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                   \
      std::exit(2);                                                            \
    }                                                                          \
  } while (0)

template <typename T> struct Dev {
  T* p = nullptr;
  size_t n = 0;
  explicit Dev(size_t count) : n(count) { HIPOK(hipMalloc(&p, count * sizeof(T))); }
  ~Dev() { (void)hipFree(p); }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  void put(const std::vector<T>& h) { HIPOK(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice)); }
  std::vector<T> get() const {
    std::vector<T> h(n);
    HIPOK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

template <typename R> std::vector<cplx<R>> random_vec(size_t n, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<cplx<R>> v(n);
  for (auto& z : v) z = {(R)u(g), (R)u(g)};
  return v;
}

template <typename T> bool same_bits(const std::vector<T>& a, const std::vector<T>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0;
}

void report(const char* what, const char* prec, size_t n, size_t ldv, int k, int extra, bool ok) {
  std::printf("%-8s %-4s n=%zu ldv=%zu k=%d %s%d : %s\n", what, prec, n, ldv, k,
              extra >= 0 ? "nrhs*dof=" : "", extra >= 0 ? extra : 0, ok ? "ok" : "FAIL");
  std::fflush(stdout);
  if (!ok) ++g_fail;
}

// single-rhs: n complex elements per vector, basis rows ldv apart
template <typename R>
void single_case(Blas& b, hipStream_t s, const char* prec, size_t n, size_t ldv, int k,
                 uint64_t seed) {
  Dev<cplx<R>> V((size_t)k * ldv), wf(n), wu(n);
  Dev<cplx<double>> c(k), hf(k + 1), hu(k + 1);
  V.put(random_vec<R>((size_t)k * ldv, seed));
  const auto w0 = random_vec<R>(n, seed + 1);
  wf.put(w0);
  wu.put(w0);
  // small coefficients, as in a second Gram-Schmidt pass
  auto cc = random_vec<double>(k, seed + 2);
  for (auto& z : cc) z = {z.re * 0.05, z.im * 0.05};
  c.put(cc);
  const std::vector<cplx<double>> marker(k + 1, cplx<double>{-7.0, -7.0});
  hf.put(marker);
  const bool took = vmultisub_dot_dev<R>(b, c.p, V.p, ldv, k, wf.p, n, hf.p);
  if (k > 24) {
    HIPOK(hipStreamSynchronize(s));
    report("single", prec, n, ldv, k, -1, !took && same_bits(wf.get(), w0) && same_bits(hf.get(), marker));
    return;
  }
  vmultisub_dev<R>(c.p, V.p, ldv, k, wu.p, n, s);
  vmultidot_dev<R>(b, V.p, ldv, k, wu.p, n, true, hu.p);
  HIPOK(hipStreamSynchronize(s));
  auto a = hf.get(), bb = hu.get();
  a.resize(k + 1);
  bb.resize(k + 1);
  report("single", prec, n, ldv, k, -1, took && same_bits(wf.get(), wu.get()) && same_bits(a, bb));
}

// batched: sites x nrhs x dof elements, basis rows ldv apart
template <typename R>
void batched_case(BBlas& b, hipStream_t s, const char* prec, size_t sites, int nrhs, int dof,
                  size_t pad, int k, uint64_t seed) {
  BShape sh;
  sh.n = sites * (size_t)nrhs * dof;
  sh.nrhs = nrhs;
  sh.dof = dof;
  const size_t ldv = sh.n + pad;
  Dev<cplx<R>> V((size_t)k * ldv), wf(sh.n), wu(sh.n);
  Dev<cplx<double>> c((size_t)k * nrhs), hf((size_t)(k + 1) * nrhs), hu((size_t)(k + 1) * nrhs);
  V.put(random_vec<R>((size_t)k * ldv, seed));
  const auto w0 = random_vec<R>(sh.n, seed + 1);
  wf.put(w0);
  wu.put(w0);
  auto cc = random_vec<double>((size_t)k * nrhs, seed + 2);
  for (auto& z : cc) z = {z.re * 0.05, z.im * 0.05};
  c.put(cc);
  const std::vector<cplx<double>> marker((size_t)(k + 1) * nrhs, cplx<double>{-7.0, -7.0});
  hf.put(marker);
  const bool took = bmultisub_dots_dev<R>(b, c.p, V.p, ldv, k, wf.p, sh, hf.p);
  if (k > 24) {
    HIPOK(hipStreamSynchronize(s));
    report("batched", prec, sh.n, ldv, k, nrhs * dof,
           !took && same_bits(wf.get(), w0) && same_bits(hf.get(), marker));
    return;
  }
  bmultisub_dev<R>(c.p, V.p, ldv, k, wu.p, sh, s);
  bdots_dev<R>(b, V.p, ldv, k, wu.p, sh, true, hu.p);
  HIPOK(hipStreamSynchronize(s));
  report("batched", prec, sh.n, ldv, k, nrhs * dof,
         took && same_bits(wf.get(), wu.get()) && same_bits(hf.get(), hu.get()));
}

}  // namespace

int main() {
  try {
    hipStream_t s;
    HIPOK(hipStreamCreate(&s));
    {
      Blas b(s);
      BBlas bb(s);
      const int ks[] = {1, 8, 9, 16, 17, 24, 25};
      // ~2 grid-stride iterations per thread at the 2048-workgroup grid
      const size_t n_odd = 1048583, n_even = 1048582;
      uint64_t seed = 1000;
      for (int k : ks) {
        single_case<double>(b, s, "f64", n_odd, n_odd, k, seed += 10);
        single_case<float>(b, s, "f32", n_even, n_even + 6, k, seed += 10);  // paired loads
        single_case<float>(b, s, "f32", n_odd, n_odd, k, seed += 10);        // unpaired
      }
      // batched: the fine-level shape (12 rhs x 12 spinor components), a
      // coarse-level shape (nbasis 32) and an odd dof (unpaired FP32 loads)
      for (int k : ks) {
        batched_case<double>(bb, s, "f64", 6000, 12, 12, 0, k, seed += 10);
        batched_case<float>(bb, s, "f32", 6000, 12, 12, 4, k, seed += 10);
        batched_case<double>(bb, s, "f64", 2401, 12, 32, 0, k, seed += 10);
        batched_case<float>(bb, s, "f32", 2401, 12, 32, 0, k, seed += 10);
        batched_case<float>(bb, s, "f32", 20000, 5, 3, 0, k, seed += 10);
      }
    }
    HIPOK(hipStreamDestroy(s));
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 3;
  }
  if (g_fail) {
    std::printf("%d case(s) FAILED\n", g_fail);
    return 1;
  }
  std::printf("ALL PASSED\n");
  return 0;
}
