// spmm_layouts -- row-gather rate of the batched PPR iterate layouts, with
// sources drawn from the entry distribution of an R-MAT graph in degree order
// (hubs first), as the SpMM gathers them:
//  * row-major: one 256-byte row of 64 FP32 columns per source, gathered with
//    1, 2 or 4 floats per lane (64, 32 or 16 lanes per row);
//  * column-blocked: four 16-column blocks of 64-byte rows, gathered as four
//    passes, or in one launch whose workgroups take a block by XCD (the two
//    XCDs x and x + 4 gather block x % 4, each over half the entries: every
//    L2 then caches 64-byte rows of one block only), or in one launch that
//    spreads the blocks over the XCDs;
// and whether a buffer written by one kernel comes back faster than from HBM
// when the next kernel reads it (retention in the MALL), by buffer size.
//
//   bin/spmm_layouts [scale]      (default 24: n = 2^24, 16 * n records, -u)
//
// The distribution comes from libeg: the R-MAT generator and the degree
// download of an undirected graph of that scale; entry e of the sampled
// stream is source v with probability deg(v) / m.  Every figure is the best
// of three launches after one warm-up, timed with HIP events.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <functional>
#include <vector>

#include "eg.h"

#define CK(x)                                                                    \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      exit(1);                                                                   \
    }                                                                            \
  } while (0)

constexpr uint32_t kSamples = 1u << 27;  // gathered entries per layout
constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Degree runs of the descending degree sequence: run r covers vertices
// [run_v[r], run_v[r+1]) of degree run_d[r] (> 0) and entries [run_e[r], ...).
__global__ void k_sample(const uint64_t *__restrict__ run_e, const uint32_t *__restrict__ run_v,
                         const uint32_t *__restrict__ run_d, uint32_t nruns, uint64_t m,
                         uint32_t n, uint32_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kSamples) return;
  const uint64_t p = mix64(i + 0x5EEDull) % m;
  uint32_t lo = 0, hi = nruns - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (run_e[mid] <= p) lo = mid; else hi = mid - 1;
  }
  const uint32_t v = run_v[lo] + (uint32_t)((p - run_e[lo]) / run_d[lo]);
  out[i] = v < n ? v : n - 1;
}

// Row-major 64-column rows, VEC floats per lane: a wave takes 64 sampled
// sources, 64 / (64 / VEC) rows per load instruction, D loads in flight.
template <int D, int VEC>
__global__ __launch_bounds__(kBlock) void k_rows(const uint32_t *__restrict__ idx,
                                                 const float *__restrict__ T,
                                                 float *__restrict__ out) {
  constexpr int kLanesPerRow = 64 / VEC;
  constexpr int kPer = 64 / (64 / kLanesPerRow);  // sources per slot per 64
  typedef float fv __attribute__((ext_vector_type(VEC)));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / kLanesPerRow, q = lane % kLanesPerRow;
  float acc = 0.f;
  for (uint32_t base = (blockIdx.x * 4 + wave) * 64; base < kSamples; base += gridDim.x * 256) {
    const uint32_t my = idx[base + lane];
    for (int j = 0; j < kPer; j += D) {
      fv v[D];
#pragma unroll
      for (int t = 0; t < D; t++) {
        const uint32_t u = __shfl(my, slot * kPer + j + t);
        v[t] = *reinterpret_cast<const fv *>(T + (uint64_t)u * 64 + q * VEC);
      }
#pragma unroll
      for (int t = 0; t < D; t++)
#pragma unroll
        for (int c = 0; c < VEC; c++) acc += v[t][c];
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = acc;
}

// Column-blocked: block b of T holds 16 columns of every source (64-byte
// rows).  MODE 0: block blk over all samples; MODE 1: block (xcd % 4), half
// (xcd / 4) of the samples; MODE 2: the same eight (block, half) classes dealt
// to workgroups regardless of their XCD.  Workgroup b runs on XCD b % 8.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_blocks(const uint32_t *__restrict__ idx,
                                                   const float *__restrict__ T, uint32_t n,
                                                   float *__restrict__ out, uint32_t blk) {
  constexpr int D = 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane >> 4, q = lane & 15;
  uint32_t b, lo, hi, wg, nwg;
  if (MODE == 0) {
    b = blk;
    lo = 0;
    hi = kSamples;
    wg = blockIdx.x;
    nwg = gridDim.x;
  } else {
    const uint32_t x = blockIdx.x & 7, g = blockIdx.x >> 3, ng = gridDim.x >> 3;
    const uint32_t cls = MODE == 1 ? x : (g & 7);
    b = cls & 3;
    lo = (cls >> 2) * (kSamples / 2);
    hi = lo + kSamples / 2;
    wg = MODE == 1 ? g : (g >> 3) * 8 + x;
    nwg = MODE == 1 ? ng : (ng >> 3) * 8;
  }
  const float *Tb = T + (uint64_t)b * n * 16;
  float acc = 0.f;
  for (uint32_t base = lo + (wg * 4 + wave) * 64; base < hi; base += nwg * 256) {
    const uint32_t my = idx[base + lane];
    for (int j = 0; j < 16; j += D) {
      float v[D];
#pragma unroll
      for (int t = 0; t < D; t++) v[t] = Tb[(uint64_t)__shfl(my, slot * 16 + j + t) * 16 + q];
#pragma unroll
      for (int t = 0; t < D; t++) acc += v[t];
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = acc;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void k_fill(f32x4 *__restrict__ p, uint64_t n4, float v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    p[i] = f32x4{v, v, v, (float)i};
}

__global__ void k_sum(const f32x4 *__restrict__ p, uint64_t n4, float *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  f32x4 a = f32x4{0.f, 0.f, 0.f, 0.f};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    a += p[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = a.x + a.y + a.z + a.w;
}

static float best_of(hipEvent_t e0, hipEvent_t e1, const std::function<void()> &run) {
  run();  // warm-up
  float best = 1e30f;
  for (int r = 0; r < 3; r++) {
    CK(hipEventRecord(e0, 0));
    run();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, ms);
  }
  return best;
}

int main(int argc, char **argv) {
  const int scale = argc > 1 ? atoi(argv[1]) : 24;
  if (scale < 16 || scale > 26) {
    fprintf(stderr, "scale must be in [16, 26]\n");
    return 2;
  }
  const uint32_t n = 1u << scale;
  // entry distribution of the undirected R-MAT graph, hubs first
  std::vector<uint32_t> deg(n);
  {
    eg_graph *g = nullptr;
    eg_gen_params p = {};
    p.kind = 0;
    p.scale = scale;
    p.m_in = 16ull << scale;
    p.seed = 1;
    p.a = 0.57;
    p.b = 0.19;
    p.c = 0.19;
    if (eg_graph_create(n, EG_UNDIRECTED | EG_SYMMETRIC, EG_MODE_ADJ_SORT, 0, &g) != EG_OK ||
        eg_graph_generate(g, &p) != EG_OK || eg_graph_get_degrees(g, deg.data(), nullptr) != EG_OK) {
      fprintf(stderr, "libeg: %s\n", eg_last_error());
      return 1;
    }
    eg_graph_free(g);
  }
  std::sort(deg.begin(), deg.end(), std::greater<uint32_t>());
  std::vector<uint64_t> run_e;
  std::vector<uint32_t> run_v, run_d;
  uint64_t m = 0;
  for (uint32_t v = 0; v < n && deg[v] > 0; v++) {
    if (v == 0 || deg[v] != deg[v - 1]) {
      run_e.push_back(m);
      run_v.push_back(v);
      run_d.push_back(deg[v]);
    }
    m += deg[v];
  }
  const uint32_t nruns = (uint32_t)run_e.size();
  printf("R-MAT %d -u: n = %u, m = %llu stored entries, %u degree runs; %u sampled entries\n",
         scale, n, (unsigned long long)m, nruns, kSamples);
  printf("share of entries on the first sources:");
  {
    uint64_t c = 0;
    uint32_t next = 1u << 10;
    for (uint32_t v = 0; v < n; v++) {
      c += deg[v];
      if (v + 1 == next) {
        printf(" 2^%d %.3f", __builtin_ctz(next), (double)c / (double)m);
        next <<= 2;
      }
    }
    printf("\n");
  }
  uint64_t *d_e;
  uint32_t *d_v, *d_d, *d_idx;
  float *d_T, *d_out;
  CK(hipMalloc(&d_e, nruns * sizeof(uint64_t)));
  CK(hipMalloc(&d_v, nruns * sizeof(uint32_t)));
  CK(hipMalloc(&d_d, nruns * sizeof(uint32_t)));
  CK(hipMemcpy(d_e, run_e.data(), nruns * sizeof(uint64_t), hipMemcpyHostToDevice));
  CK(hipMemcpy(d_v, run_v.data(), nruns * sizeof(uint32_t), hipMemcpyHostToDevice));
  CK(hipMemcpy(d_d, run_d.data(), nruns * sizeof(uint32_t), hipMemcpyHostToDevice));
  CK(hipMalloc(&d_idx, (size_t)kSamples * sizeof(uint32_t)));
  CK(hipMalloc(&d_T, (size_t)n * 64 * sizeof(float)));
  CK(hipMemset(d_T, 0, (size_t)n * 64 * sizeof(float)));
  const uint32_t grid = 8192;  // workgroups of the row-major kernels
  CK(hipMalloc(&d_out, (size_t)grid * 4 * kBlock * sizeof(float)));
  hipLaunchKernelGGL(k_sample, dim3(kSamples / kBlock), dim3(kBlock), 0, 0, d_e, d_v, d_d, nruns,
                     m, n, d_idx);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));

  printf("k = 64 row gathers (%u sources, 256 bytes each when row-major):\n", kSamples);
  auto report = [&](const char *name, float ms) {
    printf("  %-52s %7.3f ms  %5.1f G rows/s  %6.0f GB/s of rows\n", name, ms,
           kSamples / (ms * 1e-3) / 1e9, kSamples * 256.0 / (ms * 1e-3) / 1e9);
    fflush(stdout);
  };
  report("row-major, 1 float per lane (64 lanes per row)", best_of(e0, e1, [&] {
           hipLaunchKernelGGL((k_rows<8, 1>), dim3(grid), dim3(kBlock), 0, 0, d_idx, d_T, d_out);
         }));
  report("row-major, 2 floats per lane (32 lanes per row)", best_of(e0, e1, [&] {
           hipLaunchKernelGGL((k_rows<8, 2>), dim3(grid), dim3(kBlock), 0, 0, d_idx, d_T, d_out);
         }));
  report("row-major, 4 floats per lane (16 lanes per row)", best_of(e0, e1, [&] {
           hipLaunchKernelGGL((k_rows<8, 4>), dim3(grid), dim3(kBlock), 0, 0, d_idx, d_T, d_out);
         }));
  report("16-column blocks, one pass per block", best_of(e0, e1, [&] {
           for (uint32_t b = 0; b < 4; b++)
             hipLaunchKernelGGL((k_blocks<0>), dim3(grid), dim3(kBlock), 0, 0, d_idx, d_T, n,
                                d_out, b);
         }));
  report("16-column blocks, one launch, block by XCD", best_of(e0, e1, [&] {
           hipLaunchKernelGGL((k_blocks<1>), dim3(grid * 4), dim3(kBlock), 0, 0, d_idx, d_T, n,
                              d_out, 0u);
         }));
  report("16-column blocks, one launch, blocks across XCDs", best_of(e0, e1, [&] {
           hipLaunchKernelGGL((k_blocks<2>), dim3(grid * 4), dim3(kBlock), 0, 0, d_idx, d_T, n,
                              d_out, 0u);
         }));

  // retention of written data: the iterate buffer (4 GiB at scale 24) doubles as both buffers
  const uint64_t tbytes = (uint64_t)n * 64 * sizeof(float);
  const uint64_t flush4 = (2048ull << 20) / 16;
  const uint64_t sizes_mib[] = {64, 128, 192, 256, 384, 512, 1024};
  if (tbytes >= (3072ull << 20))
    printf("write, then read back at once, vs read after 2 GiB of other traffic:\n");
  for (uint64_t mib : sizes_mib) {
    if ((mib << 20) + (2048ull << 20) > tbytes) break;
    const uint64_t n4 = (mib << 20) / 16;
    f32x4 *buf = reinterpret_cast<f32x4 *>(d_T);
    f32x4 *fl = buf + (tbytes / 16 - flush4);  // the last 2 GiB, disjoint from buf[0, n4)
    float w_ms = 1e30f, rw_ms = 1e30f, cold_ms = 1e30f;
    for (int r = 0; r < 4; r++) {
      hipLaunchKernelGGL(k_fill, dim3(4096), dim3(kBlock), 0, 0, fl, flush4, 2.f);
      hipLaunchKernelGGL(k_sum, dim3(4096), dim3(kBlock), 0, 0, fl, flush4, d_out);
      CK(hipEventRecord(e0, 0));
      hipLaunchKernelGGL(k_fill, dim3(4096), dim3(kBlock), 0, 0, buf, n4, 1.f + r);
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float a;
      CK(hipEventElapsedTime(&a, e0, e1));
      CK(hipEventRecord(e0, 0));
      hipLaunchKernelGGL(k_sum, dim3(4096), dim3(kBlock), 0, 0, buf, n4, d_out);
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float b;
      CK(hipEventElapsedTime(&b, e0, e1));
      hipLaunchKernelGGL(k_fill, dim3(4096), dim3(kBlock), 0, 0, fl, flush4, 3.f);
      hipLaunchKernelGGL(k_sum, dim3(4096), dim3(kBlock), 0, 0, fl, flush4, d_out);
      CK(hipEventRecord(e0, 0));
      hipLaunchKernelGGL(k_sum, dim3(4096), dim3(kBlock), 0, 0, buf, n4, d_out);
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float c;
      CK(hipEventElapsedTime(&c, e0, e1));
      if (r) {  // the first round warms up
        w_ms = std::min(w_ms, a);
        rw_ms = std::min(rw_ms, b);
        cold_ms = std::min(cold_ms, c);
      }
    }
    const double bytes = (double)(mib << 20);
    printf("  %5llu MiB: write %6.0f GB/s, read right after %6.0f GB/s, read cold %6.0f GB/s\n",
           (unsigned long long)mib, bytes / (w_ms * 1e-3) / 1e9, bytes / (rw_ms * 1e-3) / 1e9,
           bytes / (cold_ms * 1e-3) / 1e9);
    fflush(stdout);
  }
  CK(hipFree(d_e));
  CK(hipFree(d_v));
  CK(hipFree(d_d));
  CK(hipFree(d_idx));
  CK(hipFree(d_T));
  CK(hipFree(d_out));
  printf("done\n");
  return 0;
}
