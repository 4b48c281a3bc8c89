// Device-wide exclusive scan of uint32 (three passes: block sums, scan of the
// block sums in one block, rescan with offsets). HBM-bound; used for the
// order-preserving compactions of marching cubes.
#include "qj_internal.h"

namespace qj {

constexpr int kScanThreads = 256, kScanItems = 4, kScanChunk = kScanThreads * kScanItems;

// Exclusive scan of one value per thread over the block; *total = block sum.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* total) {
  __shared__ uint32_t s_wave[kScanThreads / 64 + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kScanThreads / 64; i++) {
      const uint32_t c = s_wave[i];
      s_wave[i] = t;
      t += c;
    }
    s_wave[kScanThreads / 64] = t;
  }
  __syncthreads();
  const uint32_t excl = x - v + s_wave[wave];
  *total = s_wave[kScanThreads / 64];
  __syncthreads();
  return excl;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                              uint32_t* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; k++)
    if (base + k < n) s += in[base + k];
  uint32_t total;
  (void)block_exclusive(s, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One block scans the np partials in place (exclusive) and writes the grand total.
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(uint32_t* __restrict__ partial, int64_t np,
                                                                uint32_t* __restrict__ total_out) {
  const int64_t per = (np + kScanThreads - 1) / kScanThreads;
  const int64_t b = (int64_t)threadIdx.x * per;
  uint32_t s = 0;
  for (int64_t i = b; i < b + per && i < np; i++) s += partial[i];
  uint32_t total;
  uint32_t run = block_exclusive(s, &total);
  for (int64_t i = b; i < b + per && i < np; i++) {
    const uint32_t v = partial[i];
    partial[i] = run;
    run += v;
  }
  if (threadIdx.x == 0 && total_out) *total_out = total;
}

// in and out may alias (in-place scan): each thread reads its items before writing them.
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* in, int64_t n,
                                                             const uint32_t* __restrict__ partial, uint32_t* out) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive(s, &total) + partial[blockIdx.x];
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

size_t scan_tmp_bytes(int64_t n) {
  const int64_t nb = (n + kScanChunk - 1) / kScanChunk;
  return sizeof(uint32_t) * (size_t)(nb + 1);
}

qj_status scan_u32(const uint32_t* d_in, uint32_t* d_out, int64_t n, uint32_t* d_total, void* d_tmp,
                   size_t tmp_bytes, hipStream_t s) {
  if (n <= 0) {
    if (d_total) QJ_HIP(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
    return QJ_OK;
  }
  const int64_t nb = (n + kScanChunk - 1) / kScanChunk;
  if (tmp_bytes < scan_tmp_bytes(n) || nb > INT32_MAX) return QJ_ERR_CAPACITY;
  uint32_t* partial = (uint32_t*)d_tmp;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, s, d_in, n, partial);
  QJ_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, s, partial, nb, d_total);
  QJ_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, s, d_in, n, partial, d_out);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
}

}  // namespace qj
