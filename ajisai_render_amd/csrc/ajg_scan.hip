// ajg_scan.hip — device-wide u32 prefix sum (inclusive or exclusive) used by
// the PLOC merge compaction (the std::inclusive_scan of
// ploc_aggregate.cpp:470-471), the radix-sort digit offsets and the
// fat-node index map. Blocks of 256 threads scan 1024 elements; block sums
// are scanned recursively, then added back.
#include "ajg_internal.h"

namespace ajg {

static constexpr int kScanThreads = 256;
static constexpr int kScanItems = 4;
static constexpr uint32_t kScanTile = kScanThreads * kScanItems;

// Block-wide exclusive scan of one value per thread (64-wide waves).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wave_sums[kScanThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wave_sums[w] = x;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / 64; ++i) {
    if (i < w) wave_off += wave_sums[i];
    tot += wave_sums[i];
  }
  __syncthreads();
  *total = tot;
  return wave_off + x - v;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_tile(const uint32_t *in, uint32_t *out, uint32_t n, int inclusive,
            uint32_t *block_sums) {
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint32_t idx = base + i;
    v[i] = idx < n ? in[idx] : 0u;
    local += v[i];
  }
  uint32_t total;
  uint32_t off = block_excl_scan(local, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint32_t idx = base + i;
    uint32_t inc = off + v[i];
    if (idx < n) out[idx] = inclusive ? inc : off;
    off = inc;
  }
  if (block_sums && threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_add(uint32_t *out, uint32_t n, const uint32_t *block_offsets) {
  const uint32_t add = block_offsets[blockIdx.x];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint32_t idx = base + i;
    if (idx < n) out[idx] += add;
  }
}

size_t scan_tmp_words(uint32_t n) {
  size_t words = 0;
  uint32_t m = n;
  while (m > kScanTile) {
    m = (m + kScanTile - 1) / kScanTile;
    words += 2 * (size_t)m + 64;
  }
  return words + 64;
}

void scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, bool inclusive,
              uint32_t *tmp, hipStream_t s) {
  if (n == 0) return;
  const uint32_t blocks = (n + kScanTile - 1) / kScanTile;
  if (blocks == 1) {
    hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kScanThreads), 0, s, in, out,
                       n, inclusive ? 1 : 0, (uint32_t *)nullptr);
    return;
  }
  uint32_t *sums = tmp;
  uint32_t *sums_scanned = tmp + blocks;
  uint32_t *rest = tmp + 2 * (size_t)blocks + 64;
  hipLaunchKernelGGL(k_scan_tile, dim3(blocks), dim3(kScanThreads), 0, s, in,
                     out, n, inclusive ? 1 : 0, sums);
  scan_u32(sums, sums_scanned, blocks, false, rest, s);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks), dim3(kScanThreads), 0, s, out,
                     n, (const uint32_t *)sums_scanned);
}

}  // namespace ajg
