// g3a-gather: the random-access ceiling that bounds FM-index seeding on one
// GPU.  Threads read rows of a table at random (16-byte rows, as the narrow
// Occ lines of seeding, or 64-byte rows), eight independent reads in flight
// per thread, and the useful bytes per second are reported:
//   g3a-gather [table_MiB]    -> one JSON line on stdout
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

__device__ __forceinline__ uint64_t mix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

template <int kVec>
__global__ void __launch_bounds__(256) k_gather(const uint4 *table, uint64_t rows, int iters, uint32_t *sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t st[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) st[u] = mix(tid * 8 + u + 1);
  uint32_t acc = 0;
  for (int it = 0; it < iters; ++it) {
    uint4 v[8][kVec];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      st[u] = mix(st[u]);
      const uint64_t r = st[u] % rows;
#pragma unroll
      for (int c = 0; c < kVec; ++c) v[u][c] = table[r * kVec + c];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int c = 0; c < kVec; ++c) acc ^= v[u][c].x ^ v[u][c].y ^ v[u][c].z ^ v[u][c].w;
  }
  if (acc == 0x9e3779b9u) sink[0] = acc;  // keeps the loads; never true for a zeroed table
}

#define CK(x)                                                                 \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::fprintf(stderr, "g3a-gather: %s: %s\n", #x, hipGetErrorString(e_)); \
      return 1;                                                               \
    }                                                                         \
  } while (0)

template <int kVec>
int run(const uint4 *table, uint64_t bytes, uint32_t *sink, int n_cu, double *gbps) {
  const uint64_t rows = bytes / (16 * kVec);
  const int blocks = n_cu * 8, threads = 256, iters = 256;
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  hipLaunchKernelGGL(k_gather<kVec>, dim3(blocks), dim3(threads), 0, 0, table, rows, 4, sink);  // warm-up
  CK(hipEventRecord(a, 0));
  hipLaunchKernelGGL(k_gather<kVec>, dim3(blocks), dim3(threads), 0, 0, table, rows, iters, sink);
  CK(hipEventRecord(b, 0));
  CK(hipEventSynchronize(b));
  CK(hipGetLastError());
  float ms = 0;
  CK(hipEventElapsedTime(&ms, a, b));
  const double useful = (double)blocks * threads * iters * 8 * 16 * kVec;
  *gbps = useful / (ms * 1e-3) / 1e9;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const uint64_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
  const uint64_t bytes = mib << 20;
  int n_cu = 0;
  CK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, 0));
  uint4 *table = nullptr;
  uint32_t *sink = nullptr;
  CK(hipMalloc(&table, bytes));
  CK(hipMalloc(&sink, 4));
  CK(hipMemset(table, 0, bytes));
  double g16 = 0, g64 = 0;
  if (run<1>(table, bytes, sink, n_cu, &g16) || run<4>(table, bytes, sink, n_cu, &g64)) return 1;
  std::printf("{\"16B_rows_GBps\": %.1f, \"64B_rows_GBps\": %.1f, \"table_bytes\": %llu, \"what\": \"random rows of "
              "a table in HBM, 8 independent reads in flight per thread, %d blocks of 256 threads; useful bytes per "
              "second\"}\n",
              g16, g64, (unsigned long long)bytes, n_cu * 8);
  CK(hipFree(table));
  CK(hipFree(sink));
  return 0;
}
