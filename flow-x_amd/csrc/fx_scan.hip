// Device-wide exclusive scan of int32 counts (reduce -> scan partials ->
// scan tiles; up to kFoldTiles tiles the last step scans the partials too).
// Used for cell offsets (counting sort) and surface compaction.
#include "fx_common.h"

namespace fx {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048
constexpr int kMaxPartials = 1 << 20;  // one block scans these

size_t scan_partials(long long n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

// Inclusive scan of one value per thread across a 256-thread block (4 waves
// of 64), returns the inclusive prefix; *total receives the block sum.
__device__ __forceinline__ int block_inclusive_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  int add = 0;
  for (int w = 0; w < wave; ++w) add += lds[w];
  int tot = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  *total = tot;
  return v + add;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t* __restrict__ in,
                                                               long long n,
                                                               int32_t* __restrict__ partial) {
  __shared__ int lds[4];
  long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long q = base + i;
    if (q < n) s += in[q];
  }
  int tot;
  block_inclusive_scan(s, lds, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Exclusive scan of partial[0..nb) in place by one block; partial[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_partials(int32_t* __restrict__ partial, int nb) {
  __shared__ int lds[16];
  const int per = (nb + 1023) / 1024;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int begin = threadIdx.x * per;
  int s = 0;
  for (int i = 0; i < per; ++i)
    if (begin + i < nb) s += partial[begin + i];
  int v = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  int add = 0;
  for (int w = 0; w < wave; ++w) add += lds[w];
  int run = v + add - s;  // exclusive prefix of this thread's chunk
  for (int i = 0; i < per; ++i) {
    if (begin + i < nb) {
      int c = partial[begin + i];
      partial[begin + i] = run;
      run += c;
    }
  }
  if (threadIdx.x == 1023) partial[nb] = v + add;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_final(const int32_t* __restrict__ in,
                                                              int32_t* __restrict__ out,
                                                              long long n,
                                                              const int32_t* __restrict__ partial,
                                                              int nb) {
  __shared__ int lds[4];
  long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long q = base + i;
    v[i] = q < n ? in[q] : 0;
    s += v[i];
  }
  int tot;
  int incl = block_inclusive_scan(s, lds, &tot);
  int run = partial[blockIdx.x] + incl - s;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long q = base + i;
    if (q < n) out[q] = run;
    run += v[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partial[nb];
}

// k_scan_final that scans the tile sums itself: tile b adds up partial[0..b)
// (the last tile all of them, for the total), so no launch scans the
// partials in between.  Up to kFoldTiles tiles (each block then reads at
// most kFoldTiles / 256 sums per thread).
constexpr int kFoldTiles = 2048;

__global__ __launch_bounds__(kScanThreads) void k_scan_final_fold(const int32_t* __restrict__ in,
                                                                   int32_t* __restrict__ out, long long n,
                                                                   const int32_t* __restrict__ partial,
                                                                   int nb) {
  __shared__ int lds[4];
  const int b = (int)blockIdx.x;
  const bool last = b == nb - 1;
  int before = 0, all = 0;
  for (int i = (int)threadIdx.x; i < (last ? nb : b); i += kScanThreads) {
    const int v = partial[i];
    if (i < b) before += v;
    all += v;
  }
  int prefix, total;
  block_inclusive_scan(before, lds, &prefix);
  block_inclusive_scan(all, lds, &total);
  long long base = (long long)b * kScanTile + (long long)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long q = base + i;
    v[i] = q < n ? in[q] : 0;
    s += v[i];
  }
  int tot;
  int incl = block_inclusive_scan(s, lds, &tot);
  int run = prefix + incl - s;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    long long q = base + i;
    if (q < n) out[q] = run;
    run += v[i];
  }
  if (last && threadIdx.x == 0) out[n] = total;
}

// Up to kSmallScanMax items, one block scans everything in one launch: a
// small scene's cell offsets then cost one launch instead of three.  The
// counts pass through LDS so that global loads and stores stay coalesced
// (item i * 1024 + thread); each thread then scans per = ceil(n / 1024)
// consecutive items in LDS.  One pad word per 32 items spreads those
// stride-per reads over the banks.  Up to 1024 items (one per thread) the
// direct loads are coalesced already and skip the staging.
constexpr int kSmallScanItems = 16;
constexpr long long kSmallScanMax = 1024LL * kSmallScanItems;

__device__ __forceinline__ int scan_slot(int j) { return j + (j >> 5); }

__global__ __launch_bounds__(1024) void k_scan_small(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                     int n) {
  __shared__ int data[kSmallScanMax + kSmallScanMax / 32];
  __shared__ int lds[16];
  const int per = (n + 1023) / 1024;  // <= kSmallScanItems
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int begin = threadIdx.x * per;
  int s = 0;
  if (per == 1) {
    s = begin < n ? in[begin] : 0;  // one item per thread: already coalesced
  } else {
#pragma unroll
    for (int i = 0; i < kSmallScanItems; ++i) {
      const int j = i * 1024 + (int)threadIdx.x;
      if (j < n) data[scan_slot(j)] = in[j];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSmallScanItems; ++i)
      if (i < per && begin + i < n) s += data[scan_slot(begin + i)];
  }
  int incl = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(incl, d, 64);
    if (lane >= d) incl += u;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  for (int w = 0; w < wave; ++w) incl += lds[w];
  int run = incl - s;
  if (per == 1) {
    if (begin < n) out[begin] = run;
  } else {
#pragma unroll
    for (int i = 0; i < kSmallScanItems; ++i) {
      if (i < per && begin + i < n) {
        const int v = data[scan_slot(begin + i)];
        data[scan_slot(begin + i)] = run;
        run += v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSmallScanItems; ++i) {
      const int j = i * 1024 + (int)threadIdx.x;
      if (j < n) out[j] = data[scan_slot(j)];
    }
  }
  if (threadIdx.x == 1023) out[n] = incl;
}

int exclusive_scan_i32(const int32_t* counts, int32_t* out, long long n, int32_t* partial,
                       hipStream_t st) {
  if (n <= kSmallScanMax) {
    FX_LAUNCH(k_scan_small, dim3(1), dim3(1024), st, counts, out, (int)n);
    return check_launch("exclusive_scan_i32");
  }
  long long nb = (n + kScanTile - 1) / kScanTile;
  if (nb < 1) nb = 1;
  if (nb > kMaxPartials) return fail(FX_ERANGE, "scan of %lld items exceeds %d tiles", n, kMaxPartials);
  FX_LAUNCH(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), st, counts, n, partial);
  if (nb <= kFoldTiles) {
    FX_LAUNCH(k_scan_final_fold, dim3((unsigned)nb), dim3(kScanThreads), st, counts, out, n,
              (const int32_t*)partial, (int)nb);
    return check_launch("exclusive_scan_i32");
  }
  FX_LAUNCH(k_scan_partials, dim3(1), dim3(1024), st, partial, (int)nb);
  FX_LAUNCH(k_scan_final, dim3((unsigned)nb), dim3(kScanThreads), st, counts, out, n, (const int32_t*)partial,
            (int)nb);
  return check_launch("exclusive_scan_i32");
}

}  // namespace fx
