// ceilings -- device rates that bound the binned PageRank phases and the
// batched PPR SpMM:
//  * stream shapes over 2 GiB buffers: 16-byte reads, 16-byte and 4-byte
//    writes, a 16-byte copy, and the "2-byte offset in, 4-byte value out"
//    map of the binned scatter, once with one u16 load per lane and once with
//    8-byte u16x4 loads passed through an LDS stage (lane-contiguous stores);
//  * random 256-byte row gathers (64 FP32 columns, a wave per row, eight rows
//    in flight per wave, indices streamed as the SpMM streams its column
//    indices) from tables of 2^14 .. 2^24 rows: L2-resident, MALL-sized and
//    HBM-sized working sets.
//
//   bin/ceilings
//
// Every figure is the best of five launches timed with HIP events; bytes are
// the bytes each kernel reads plus writes.
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

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));

__global__ void k_read16(const u32x4 *__restrict__ a, size_t n, uint32_t *__restrict__ out) {
  uint32_t x = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const u32x4 v = __builtin_nontemporal_load(a + i);
    x ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (x == 0x9E3779B9u) out[0] = x;  // keeps the loads
}

__global__ void k_write16(u32x4 *__restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = u32x4{(uint32_t)i, 1u, 2u, 3u};
}

__global__ void k_write4(uint32_t *__restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = (uint32_t)i;
}

__global__ void k_copy16(const u32x4 *__restrict__ a, u32x4 *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = __builtin_nontemporal_load(a + i);
}

// n positions: one u16 read and one u32 written per lane
__global__ void k_map_u16(const uint16_t *__restrict__ a, uint32_t *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i] * 3u;
}

// nq quads: u16x4 loads, a per-wave LDS stage, lane-contiguous u32 stores
// (256-thread blocks: four waves, 256 positions per wave and step)
__global__ __launch_bounds__(256) void k_map_u16x4(const u16x4 *__restrict__ a,
                                                   uint32_t *__restrict__ b, size_t nq) {
  __shared__ u16x4 stage[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint16_t *st = reinterpret_cast<const uint16_t *>(stage[w]);
  const size_t wave = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
  const size_t nw = ((size_t)gridDim.x * blockDim.x) >> 6;
  for (size_t blk = wave; blk * 64 < nq; blk += nw) {
    const size_t qi = blk * 64 + lane;
    stage[w][lane] = qi < nq ? __builtin_nontemporal_load(a + qi) : u16x4{0, 0, 0, 0};
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < 4; q++) {
      const size_t p = blk * 256 + q * 64 + lane;
      if (p < nq * 4) b[p] = st[q * 64 + lane] * 3u;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// m gathers of 256-byte rows: a wave per 64 indices, lane q reads column q
__global__ __launch_bounds__(256) void k_rows(const uint32_t *__restrict__ idx, size_t m,
                                              const float *__restrict__ C,
                                              float *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const size_t wave = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
  const size_t nw = ((size_t)gridDim.x * blockDim.x) >> 6;
  float acc = 0.f;
  for (size_t b = wave * 64; b < m; b += nw * 64) {
    const uint32_t my = b + lane < m ? __builtin_nontemporal_load(idx + b + lane) : 0u;
    const int cnt = (int)(m - b < 64 ? m - b : 64);
    for (int j = 0; j < cnt; j += 8) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const uint32_t u = __shfl(my, (j + t) & 63);
        v[t] = j + t < cnt ? C[(size_t)u * 64 + lane] : 0.f;
      }
#pragma unroll
      for (int t = 0; t < 8; t++) acc += v[t];
    }
  }
  if (acc == 1234.5f) out[0] = acc;
}

template <typename F>
static float best_ms(F launch) {
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  float best = 1e30f;
  for (int r = 0; r < 5; r++) {
    CK(hipEventRecord(e0, 0));
    launch();
    CK(hipGetLastError());
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) best = ms;
  }
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  return best;
}

int main() {
  const size_t bytes = (size_t)2 << 30;
  void *A, *B;
  uint32_t *out;
  CK(hipMalloc(&A, bytes));
  CK(hipMalloc(&B, bytes));
  CK(hipMalloc(&out, 4));
  CK(hipMemset(A, 1, bytes));
  CK(hipMemset(B, 1, bytes));
  const dim3 grid(4096), blk(256);
  printf("stream shapes (2 GiB buffers, grid %u x %u):\n", grid.x, blk.x);
  struct Shape {
    const char *name;
    double moved;  // bytes read + written
    float ms;
  } sh[6];
  sh[0] = {"read 16 B/lane", (double)bytes, best_ms([&] {
             hipLaunchKernelGGL(k_read16, grid, blk, 0, 0, (const u32x4 *)A, bytes / 16, out);
           })};
  sh[1] = {"write 16 B/lane", (double)bytes, best_ms([&] {
             hipLaunchKernelGGL(k_write16, grid, blk, 0, 0, (u32x4 *)A, bytes / 16);
           })};
  sh[2] = {"write 4 B/lane", (double)bytes, best_ms([&] {
             hipLaunchKernelGGL(k_write4, grid, blk, 0, 0, (uint32_t *)A, bytes / 4);
           })};
  // copy: 1 GiB read + 1 GiB written
  sh[3] = {"copy 16 B/lane", (double)bytes, best_ms([&] {
             hipLaunchKernelGGL(k_copy16, grid, blk, 0, 0, (const u32x4 *)A, (u32x4 *)B,
                                bytes / 32);
           })};
  // bytes / 8 positions: 2 B read (bytes / 4 of A) and 4 B written (bytes / 2 of B)
  sh[4] = {"u16 in / u32 out, 1 per lane", bytes / 8 * 6.0, best_ms([&] {
             hipLaunchKernelGGL(k_map_u16, grid, blk, 0, 0, (const uint16_t *)A, (uint32_t *)B,
                                bytes / 8);
           })};
  sh[5] = {"u16 in / u32 out, u16x4 loads + LDS stage", bytes / 8 * 6.0, best_ms([&] {
             hipLaunchKernelGGL(k_map_u16x4, grid, blk, 0, 0, (const u16x4 *)A, (uint32_t *)B,
                                bytes / 32);
           })};
  for (const Shape &s : sh)
    printf("  %-44s %8.3f ms %8.1f GB/s\n", s.name, s.ms, s.moved / (s.ms * 1e-3) / 1e9);
  CK(hipFree(B));

  // row gathers: A holds up to 2^24 rows of 256 B (4 GiB)
  CK(hipFree(A));
  const size_t max_rows = (size_t)1 << 24;
  const size_t m = (size_t)1 << 27;
  float *C;
  uint32_t *idx;
  CK(hipMalloc(&C, max_rows * 256));
  CK(hipMalloc(&idx, m * 4));
  CK(hipMemset(C, 0, max_rows * 256));
  std::vector<uint32_t> h(m);
  printf("random 256-byte row gathers (2^27 per launch):\n");
  uint64_t s = 88172645463325252ull;
  for (int lg = 14; lg <= 24; lg += 2) {
    const size_t rows = (size_t)1 << lg;
    for (size_t i = 0; i < m; i++) {
      s ^= s << 13;
      s ^= s >> 7;
      s ^= s << 17;
      h[i] = (uint32_t)(s % rows);  // < rows <= max_rows: every gather in bounds
    }
    CK(hipMemcpy(idx, h.data(), m * 4, hipMemcpyHostToDevice));
    const float ms = best_ms([&] {
      hipLaunchKernelGGL(k_rows, grid, blk, 0, 0, idx, m, (const float *)C, (float *)out);
    });
    printf("  table 2^%d rows (%6zu MiB): %8.3f ms %6.1f G rows/s %8.1f GB/s of rows\n", lg,
           rows * 256 >> 20, ms, m / (ms * 1e-3) / 1e9, m * 256.0 / (ms * 1e-3) / 1e9);
  }
  CK(hipFree(C));
  CK(hipFree(idx));
  CK(hipFree(out));
  return 0;
}
