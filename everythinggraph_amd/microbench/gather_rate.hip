// gather_rate -- random 4-byte gathers per second from an array of a given
// size: the ceiling of the pull kernels, whose time goes into gathering one
// float per stored entry from a vertex-indexed vector.
//
//   bin/gather_rate [array_floats ...]
//
// For each size, 2^27 indices uniform in [0, size) (generated and checked on
// the host) are streamed 16 per lane, as k_pull does, and the gathered values
// are summed per thread so that nothing is optimised away.  Prints the best of
// three timed launches (HIP events) and the rate in gathers per second.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#define CK(x)                                                                    \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      exit(1);                                                                   \
    }                                                                            \
  } while (0)

constexpr int kDepth = 16;

__global__ __launch_bounds__(256) void k_gather(const uint32_t *__restrict__ idx,
                                                const float *__restrict__ x, uint32_t m,
                                                float *__restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t T = gridDim.x * blockDim.x;
  float acc = 0.f;
  for (uint32_t base = t; base < m; base += T * kDepth) {
    uint32_t u[kDepth];
#pragma unroll
    for (int q = 0; q < kDepth; q++) {
      const uint32_t k = base + q * T;
      u[q] = k < m ? __builtin_nontemporal_load(idx + k) : 0u;
    }
    float v[kDepth];
#pragma unroll
    for (int q = 0; q < kDepth; q++) v[q] = x[u[q]];
#pragma unroll
    for (int q = 0; q < kDepth; q++) acc += v[q];
  }
  out[t] = acc;
}

int main(int argc, char **argv) {
  const uint32_t m = 1u << 27;
  const int grid = 256 * 16;
  std::vector<uint32_t> sizes;
  for (int i = 1; i < argc; i++) sizes.push_back((uint32_t)strtoul(argv[i], nullptr, 10));
  if (sizes.empty()) sizes = {8192, 262144, 1048576, 4194304, 16777216, 67108864};
  uint32_t *idx;
  float *out;
  CK(hipMalloc(&idx, (size_t)m * 4));
  CK(hipMalloc(&out, (size_t)grid * 256 * 4));
  std::vector<uint32_t> h(m);
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  for (uint32_t n : sizes) {
    if (n == 0 || n > (1u << 28)) {
      fprintf(stderr, "size %u out of range (1 .. 2^28 floats)\n", n);
      return 2;
    }
    uint64_t s = 12345;
    for (uint32_t i = 0; i < m; i++) {
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      h[i] = (uint32_t)((s >> 33) % n);
    }
    float *x;
    CK(hipMalloc(&x, (size_t)n * 4));
    CK(hipMemset(x, 0, (size_t)n * 4));
    CK(hipMemcpy(idx, h.data(), (size_t)m * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, 0, idx, x, m, out);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < 3; r++) {
      CK(hipEventRecord(a, 0));
      hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, 0, idx, x, m, out);
      CK(hipEventRecord(b, 0));
      CK(hipEventSynchronize(b));
      float ms;
      CK(hipEventElapsedTime(&ms, a, b));
      if (ms < best) best = ms;
    }
    printf("array %10u floats (%8.2f MB): %u gathers in %.3f ms = %.1f G gathers/s\n", n,
           n * 4.0 / 1e6, m, best, m / (best * 1e-3) / 1e9);
    fflush(stdout);
    CK(hipFree(x));
  }
  return 0;
}
