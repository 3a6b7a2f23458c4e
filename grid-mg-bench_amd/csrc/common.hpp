// Shared device/host helpers of libmgc: status handling, complex types,
// lattice indexing.  Everything here is gfx950-only HIP.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <stdexcept>

#include "../../include/mgc.h"

namespace mgc {

// --------------------------------------------------------------------------
// error handling: internal code throws mgc::Error, every extern "C" entry
// point catches and converts to a status code (no exception crosses the ABI)
// --------------------------------------------------------------------------
struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
// zero `bytes` of device memory with a kernel (graph-capture safe)
void fill_zero(void* ptr, size_t bytes, hipStream_t s);

#define MGC_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::mgc::Error(MGC_ERR_HIP, std::string(#call) + ": " +             \
                                          hipGetErrorString(e_));              \
  } while (0)

#define MGC_CHECK(cond, code, msg)                                             \
  do {                                                                         \
    if (!(cond)) throw ::mgc::Error((code), (msg));                            \
  } while (0)

#define MGC_LAUNCH_CHECK() MGC_HIP(hipGetLastError())

template <class F>
inline int guarded(F&& f) {
  try {
    f();
    return MGC_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MGC_ERR_ARG;
  }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// device allocation freed when it goes out of scope (also on an exception)
template <typename T>
struct DevBuf {
  T* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  void alloc(size_t count) {
    reset();
    MGC_HIP(hipMalloc((void**)&p, count * sizeof(T)));
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  ~DevBuf() { reset(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  T* get() const { return p; }
};

// --------------------------------------------------------------------------
// complex helpers (interleaved re,im)
// --------------------------------------------------------------------------
template <typename R> struct cplx { R re, im; };

template <typename R>
__device__ __host__ inline cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__device__ __host__ inline cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }

// acc += a*b
template <typename R>
__device__ inline void cmac(cplx<R>& acc, cplx<R> a, cplx<R> b) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
}
// acc += conj(a)*b
template <typename R>
__device__ inline void cmac_conj(cplx<R>& acc, cplx<R> a, cplx<R> b) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(-a.im, b.re, acc.im);
}

inline size_t prec_real_bytes(int prec) {
  switch (prec) {
    case MGC_F64: return 8;
    case MGC_F32: return 4;
    case MGC_F16: return 2;
    case MGC_BF16: return 2;
  }
  throw Error(MGC_ERR_PRECISION, "unknown precision");
}

// --------------------------------------------------------------------------
// lattice indexing (dimension 0 fastest)
// --------------------------------------------------------------------------
struct Lat4 {
  int L[4];
  __host__ __device__ int volume() const { return L[0] * L[1] * L[2] * L[3]; }
  __host__ __device__ void coords(int x, int c[4]) const {
    c[0] = x % L[0]; x /= L[0];
    c[1] = x % L[1]; x /= L[1];
    c[2] = x % L[2]; x /= L[2];
    c[3] = x;
  }
  __host__ __device__ int index(const int c[4]) const {
    return c[0] + L[0] * (c[1] + L[1] * (c[2] + L[2] * c[3]));
  }
  // neighbour of x along dir with displacement disp (+1/-1), periodic
  __host__ __device__ int shift(int x, int dir, int disp) const {
    int c[4];
    coords(x, c);
    c[dir] = (c[dir] + disp + L[dir]) % L[dir];
    return index(c);
  }
};

// stencil point order ("directions faster"): p<4 -> (dir p, +1), p<8 -> (dir p-4, -1), 8 -> self
__host__ __device__ inline int point_dir(int p) { return p == 8 ? 0 : (p & 3); }
__host__ __device__ inline int point_disp(int p) { return p == 8 ? 0 : (p < 4 ? +1 : -1); }
inline int point_from_dir_disp(int dir, int disp) {
  MGC_CHECK(disp == -1 || disp == 0 || disp == 1, MGC_ERR_ARG, "disp must be -1, 0 or +1");
  if (disp == 0) {
    MGC_CHECK(dir == 0, MGC_ERR_ARG, "self point needs dir == 0");
    return 8;
  }
  MGC_CHECK(dir >= 0 && dir < 4, MGC_ERR_ARG, "dir must be in 0..3");
  return (1 - disp) / 2 * 4 + dir;
}

// number of blocks for a 1d grid-stride launch
inline unsigned grid_for(size_t n, unsigned block, unsigned cap = 65535u * 8u) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace mgc
