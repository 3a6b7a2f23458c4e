// Runtime helpers of the C ABI: last-error string, device memory, streams,
// events, and the counter-based Gaussian generator used for synthetic inputs.
#include "common.hpp"

#include <cstring>
#include <cstdio>
#include <cmath>

namespace mgc {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// zero fill kernel (used instead of hipMemsetAsync so that callers may
// capture the library's work into a HIP graph)
// ---------------------------------------------------------------------------
__global__ void k_fill_zero_u32(uint32_t* p, size_t n_words) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n_words; i += stride) p[i] = 0u;
}
__global__ void k_fill_zero_u8(uint8_t* p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

void fill_zero(void* ptr, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  size_t words = bytes / 4;
  if (words) {
    k_fill_zero_u32<<<grid_for(words, 256, 16384), 256, 0, s>>>((uint32_t*)ptr, words);
    MGC_LAUNCH_CHECK();
  }
  size_t rest = bytes - words * 4;
  if (rest) {
    k_fill_zero_u8<<<1, 64, 0, s>>>((uint8_t*)ptr + words * 4, rest);
    MGC_LAUNCH_CHECK();
  }
}

// ---------------------------------------------------------------------------
// counter-based generator: splitmix64 finaliser over (seed, stream, counter).
// uniform u in (0,1) from the top 53 bits; Box-Muller pairs give re, im.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__host__ __device__ inline uint64_t rng_key(uint64_t seed, uint64_t stream) {
  return splitmix64(splitmix64(seed) ^ (stream * 0xD1B54A32D192ED03ull));
}
__host__ __device__ inline double rng_uniform(uint64_t key, uint64_t ctr) {
  uint64_t b = splitmix64(key ^ splitmix64(ctr));
  return ((double)(b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

template <typename R>
__global__ void k_gaussian(cplx<R>* out, size_t n, uint64_t key) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    double u1 = rng_uniform(key, 2 * i);
    double u2 = rng_uniform(key, 2 * i + 1);
    double r = sqrt(-log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    out[i] = {(R)(r * c), (R)(r * s)};
  }
}

}  // namespace mgc

using namespace mgc;

extern "C" {

const char* mgc_last_error(void) { return g_last_error.c_str(); }
int mgc_version(void) { return 10000; }

int mgc_device_count(int* count) {
  return guarded([&] {
    MGC_CHECK(count, MGC_ERR_ARG, "null count");
    MGC_HIP(hipGetDeviceCount(count));
  });
}
int mgc_device_name(char* buf, int len) {
  return guarded([&] {
    MGC_CHECK(buf && len > 0, MGC_ERR_ARG, "bad buffer");
    int dev = 0;
    MGC_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MGC_HIP(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount);
  });
}
int mgc_malloc(void** ptr, size_t bytes) {
  return guarded([&] {
    MGC_CHECK(ptr, MGC_ERR_ARG, "null ptr");
    MGC_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  });
}
int mgc_free(void* ptr) {
  return guarded([&] { if (ptr) MGC_HIP(hipFree(ptr)); });
}
int mgc_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    MGC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    MGC_HIP(hipStreamSynchronize(as_stream(stream)));
  });
}
int mgc_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    MGC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    MGC_HIP(hipStreamSynchronize(as_stream(stream)));
  });
}
int mgc_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    MGC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  });
}
int mgc_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    MGC_CHECK(dst && src, MGC_ERR_ARG, "null pointer");
    MGC_CHECK(bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0,
              MGC_ERR_ARG, "device_copy needs 16-byte aligned sizes and pointers");
    const size_t n = bytes / 16;
    k_copy_u4<<<grid_for(n, 256, 16384), 256, 0, as_stream(stream)>>>((const uint4*)src,
                                                                      (uint4*)dst, n);
    MGC_LAUNCH_CHECK();
  });
}

int mgc_fill_zero(void* ptr, size_t bytes, void* stream) {
  return guarded([&] {
    MGC_CHECK(ptr || bytes == 0, MGC_ERR_ARG, "null ptr");
    fill_zero(ptr, bytes, as_stream(stream));
  });
}
int mgc_stream_create(void** stream) {
  return guarded([&] {
    MGC_CHECK(stream, MGC_ERR_ARG, "null stream ptr");
    hipStream_t s;
    MGC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
  });
}
int mgc_stream_destroy(void* stream) {
  return guarded([&] { if (stream) MGC_HIP(hipStreamDestroy(as_stream(stream))); });
}
int mgc_stream_synchronize(void* stream) {
  return guarded([&] { MGC_HIP(hipStreamSynchronize(as_stream(stream))); });
}
int mgc_device_synchronize(void) {
  return guarded([&] { MGC_HIP(hipDeviceSynchronize()); });
}
int mgc_event_create(void** event) {
  return guarded([&] {
    MGC_CHECK(event, MGC_ERR_ARG, "null event ptr");
    hipEvent_t e;
    MGC_HIP(hipEventCreate(&e));
    *event = (void*)e;
  });
}
int mgc_event_destroy(void* event) {
  return guarded([&] { if (event) MGC_HIP(hipEventDestroy((hipEvent_t)event)); });
}
int mgc_event_record(void* event, void* stream) {
  return guarded([&] { MGC_HIP(hipEventRecord((hipEvent_t)event, as_stream(stream))); });
}
int mgc_event_synchronize(void* event) {
  return guarded([&] { MGC_HIP(hipEventSynchronize((hipEvent_t)event)); });
}
int mgc_event_elapsed_ms(float* ms, void* start, void* stop) {
  return guarded([&] {
    MGC_CHECK(ms, MGC_ERR_ARG, "null ms");
    MGC_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  });
}

int mgc_random_gaussian(void* out, size_t n_complex, int prec, uint64_t seed,
                        uint64_t stream, void* hip_stream) {
  return guarded([&] {
    MGC_CHECK(out || n_complex == 0, MGC_ERR_ARG, "null out");
    if (n_complex == 0) return;
    uint64_t key = rng_key(seed, stream);
    hipStream_t s = as_stream(hip_stream);
    unsigned g = grid_for(n_complex, 256, 16384);
    if (prec == MGC_F64)
      k_gaussian<double><<<g, 256, 0, s>>>((cplx<double>*)out, n_complex, key);
    else if (prec == MGC_F32)
      k_gaussian<float><<<g, 256, 0, s>>>((cplx<float>*)out, n_complex, key);
    else
      throw Error(MGC_ERR_PRECISION, "random_gaussian: F64 or F32 only");
    MGC_LAUNCH_CHECK();
  });
}

}  // extern "C"
