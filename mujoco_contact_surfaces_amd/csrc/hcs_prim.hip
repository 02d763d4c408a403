/*
 * hcs_prim.hip -- the engine's own integer primitives for gfx950: an
 * exclusive scan of unsigned counts (survivor slots of the contact step,
 * fan and sample offsets of the sensors) and a stable LSD radix sort of
 * 32- or 64-bit keys with int32 values (Morton codes of the LBVH build,
 * polygon keys of the sensors).
 *
 * Scan (reduce-then-scan, three launches, no atomics, no waits): tiles of
 * 2048 counts (256 threads x 8), a per-tile sum, one block scanning the
 * tile sums, then every tile scanned again with its offset.  Within a block
 * each thread sums its 8 contiguous counts, a wave scans the thread sums
 * with shuffles, and the four wave totals are scanned in LDS.
 *
 * Radix sort: 8-bit digits from begin_bit; per digit pass, a per-tile
 * histogram (LDS atomics), an exclusive scan of the histograms laid out
 * digit-major (so tile t's keys of digit d start after every key of a
 * smaller digit and after digit d's keys of tiles < t), then a scatter
 * that ranks each key among the tile's keys of its digit in original order:
 * 8 rounds of 256 keys (key j*256 + t of the tile in round j), the rank in
 * the wave from ballots over the 8 digit bits, the rank among earlier waves
 * and earlier rounds from LDS counters.  Passes ping-pong between the
 * output and a temporary copy so the last one lands in the output.
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "hcs_prim.h"

namespace hcs {

namespace {

constexpr int kThreads = 256;
constexpr int kPer = 8;
constexpr int kTile = kThreads * kPer;
constexpr int kSumThreads = 1024;
constexpr int kMaxSumPer = 64; /* at most this many tile sums per thread of the middle pass */

__device__ inline unsigned wave_incl_scan(unsigned v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

/* Exclusive scan of v over the block (blockDim.x threads, a multiple of
 * 64); *total = the block's sum. */
__device__ inline unsigned block_excl_scan(unsigned v, unsigned* lds, unsigned* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned inc = wave_incl_scan(v);
  __syncthreads();
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  unsigned before = 0, all = 0;
  for (int k = 0; k < nw; ++k) {
    const unsigned t = lds[k];
    before += k < w ? t : 0u;
    all += t;
  }
  *total = all;
  return before + inc - v;
}

__global__ __launch_bounds__(kThreads) void k_scan_reduce(const unsigned* __restrict__ in, long long n,
                                                         unsigned* __restrict__ sums) {
  __shared__ unsigned lds[kThreads / 64];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
  unsigned s = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) s += base + j < n ? in[base + j] : 0u;
  unsigned total;
  (void)block_excl_scan(s, lds, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

/* One block: exclusive scan of ntiles sums in place, each thread a
 * contiguous run of `per` sums (read once to sum, again to write). */
__global__ __launch_bounds__(kSumThreads) void k_scan_sums(unsigned* __restrict__ sums, int ntiles, int per) {
  __shared__ unsigned lds[kSumThreads / 64];
  const int b0 = threadIdx.x * per;
  unsigned s = 0;
  for (int j = 0; j < per; ++j) s += b0 + j < ntiles ? sums[b0 + j] : 0u;
  unsigned total;
  unsigned run = block_excl_scan(s, lds, &total);
  for (int j = 0; j < per && b0 + j < ntiles; ++j) {
    const unsigned v = sums[b0 + j];
    sums[b0 + j] = run;
    run += v;
  }
}

/* in == out is allowed: every thread of a tile loads its counts before the
 * block scan's barriers and writes only after them. */
__global__ __launch_bounds__(kThreads) void k_scan_apply(const unsigned* in, long long n,
                                                        const unsigned* __restrict__ sums, unsigned* out) {
  __shared__ unsigned lds[kThreads / 64];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
  unsigned v[kPer];
  unsigned s = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    v[j] = base + j < n ? in[base + j] : 0u;
    s += v[j];
  }
  unsigned total;
  unsigned run = sums[blockIdx.x] + block_excl_scan(s, lds, &total);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

template <class K>
__global__ __launch_bounds__(kThreads) void k_radix_hist(const K* __restrict__ keys, long long n, int shift,
                                                        unsigned dmask, int ntiles, unsigned* __restrict__ hist) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kTile;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const long long i = base + (long long)j * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & dmask], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(const K* __restrict__ kin, const int* __restrict__ vin,
                                                           long long n, int shift, unsigned dmask, int ntiles,
                                                           const unsigned* __restrict__ offs, K* __restrict__ kout,
                                                           int* __restrict__ vout) {
  __shared__ unsigned base[256], run[256], cnt[kThreads / 64][256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  base[t] = offs[(size_t)t * ntiles + blockIdx.x];
  run[t] = 0;
  for (int k = 0; k < kThreads / 64; ++k) cnt[k][t] = 0;
  __syncthreads();
  const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const long long tile = (long long)blockIdx.x * kTile;
  for (int j = 0; j < kPer; ++j) {
    const long long i = tile + (long long)j * kThreads + t;
    const bool valid = i < n;
    K key = valid ? kin[i] : (K)0;
    const unsigned d = (unsigned)(key >> shift) & dmask;
    /* lanes of this wave holding the same digit */
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const unsigned rank = (unsigned)__popcll(peers & lt);
    if (valid && rank == 0) cnt[w][d] = (unsigned)__popcll(peers);
    __syncthreads();
    if (valid) {
      unsigned before = 0;
      for (int k = 0; k < w; ++k) before += cnt[k][d];
      const unsigned pos = base[d] + run[d] + before + rank;
      kout[pos] = key;
      vout[pos] = vin[i];
    }
    __syncthreads();
    unsigned add = 0;
    for (int k = 0; k < kThreads / 64; ++k) add += cnt[k][t], cnt[k][t] = 0;
    run[t] += add;
    __syncthreads();
  }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

size_t scan_temp_bytes(long long n) {
  const long long tiles = std::max(1LL, (n + kTile - 1) / kTile);
  return align256(sizeof(unsigned) * (size_t)tiles);
}

int exclusive_scan_u32(hipStream_t s, const unsigned* in, unsigned* out, long long n, void* tmp, size_t tmp_bytes) {
  if (n <= 0) return 0;
  const long long tiles = (n + kTile - 1) / kTile;
  if (tiles > (long long)kSumThreads * kMaxSumPer || tmp_bytes < scan_temp_bytes(n)) return -1;
  const int per = (int)((tiles + kSumThreads - 1) / kSumThreads);
  unsigned* sums = (unsigned*)tmp;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)tiles), dim3(kThreads), 0, s, in, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kSumThreads), 0, s, sums, (int)tiles, per);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(kThreads), 0, s, in, n, (const unsigned*)sums, out);
  return 0;
}

size_t sort_temp_bytes(long long n, int key_bytes) {
  const long long tiles = std::max(1LL, (n + kTile - 1) / kTile);
  const long long nh = 256 * tiles;
  return align256((size_t)key_bytes * n) + align256(sizeof(int) * (size_t)n) + align256(sizeof(unsigned) * nh) +
         scan_temp_bytes(nh);
}

template <class K>
static int radix_sort(hipStream_t s, const K* kin, K* kout, const int* vin, int* vout, long long n, int begin_bit,
                      int end_bit, void* tmp, size_t tmp_bytes) {
  if (n <= 0) return 0;
  if (tmp_bytes < sort_temp_bytes(n, sizeof(K)) || begin_bit < 0 || end_bit > 8 * (int)sizeof(K) ||
      begin_bit >= end_bit)
    return -1;
  const long long tiles = (n + kTile - 1) / kTile;
  const long long nh = 256 * tiles;
  if (nh > 0x7fffffffLL) return -1;
  char* p = (char*)tmp;
  K* kalt = (K*)p;
  p += align256(sizeof(K) * (size_t)n);
  int* valt = (int*)p;
  p += align256(sizeof(int) * (size_t)n);
  unsigned* hist = (unsigned*)p;
  p += align256(sizeof(unsigned) * (size_t)nh);
  void* stmp = p;
  const int passes = (end_bit - begin_bit + 7) / 8;
  const K* ks = kin;
  const int* vs = vin;
  for (int q = 0; q < passes; ++q) {
    /* the last pass writes the output */
    const bool to_out = ((passes - 1 - q) % 2) == 0;
    K* kd = to_out ? kout : kalt;
    int* vd = to_out ? vout : valt;
    const int shift = begin_bit + 8 * q;
    const unsigned dmask = (1u << std::min(8, end_bit - shift)) - 1u; /* only the requested bits */
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_radix_hist<K>), dim3((unsigned)tiles), dim3(kThreads), 0, s, ks, n, shift,
                       dmask, (int)tiles, hist);
    if (exclusive_scan_u32(s, hist, hist, nh, stmp, scan_temp_bytes(nh))) return -1;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_radix_scatter<K>), dim3((unsigned)tiles), dim3(kThreads), 0, s, ks, vs, n,
                       shift, dmask, (int)tiles, (const unsigned*)hist, kd, vd);
    ks = kd;
    vs = vd;
  }
  return 0;
}

int radix_sort_pairs_u32(hipStream_t s, const unsigned* kin, unsigned* kout, const int* vin, int* vout, long long n,
                         int begin_bit, int end_bit, void* tmp, size_t tmp_bytes) {
  return radix_sort<unsigned>(s, kin, kout, vin, vout, n, begin_bit, end_bit, tmp, tmp_bytes);
}

int radix_sort_pairs_u64(hipStream_t s, const unsigned long long* kin, unsigned long long* kout, const int* vin,
                         int* vout, long long n, int begin_bit, int end_bit, void* tmp, size_t tmp_bytes) {
  return radix_sort<unsigned long long>(s, kin, kout, vin, vout, n, begin_bit, end_bit, tmp, tmp_bytes);
}

}  // namespace hcs
