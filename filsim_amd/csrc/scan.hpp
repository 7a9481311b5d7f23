// scan.hpp -- device-wide exclusive scan of int counts (counting sorts of
// the spread bins and the barrier cells): block sums, one-block scan of the
// block sums, then block-local scans with the block offsets; arrays of up
// to 16Ki counts in one workgroup instead.
#pragma once

#include "common.hpp"

namespace fs {

constexpr int kScanBlock = 1024;
constexpr int kScanMaxBlocks = 65536;  // up to 64Mi elements

namespace {  // each including translation unit gets its own kernels

__global__ void __launch_bounds__(256)
    k_scan_blocksum(const int *__restrict__ cnt, int n, int *__restrict__ bsum) {
  __shared__ int red[256];
  const int b = blockIdx.x;
  const int lo = b * kScanBlock;
  int s = 0;
  for (int i = lo + threadIdx.x; i < min(n, lo + kScanBlock); i += 256) s += cnt[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[b] = red[0];
}

// Exclusive scan in place of nb block sums, 1024 at a time with a running
// carry; bsum[nb] = total.
__global__ void __launch_bounds__(1024) k_scan_top(int *__restrict__ bsum, int nb) {
  __shared__ int s[1024];
  const int t = threadIdx.x;
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + t;
    s[t] = (i < nb) ? bsum[i] : 0;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int v = (t >= off) ? s[t - off] : 0;
      __syncthreads();
      s[t] += v;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + ((t == 0) ? 0 : s[t - 1]);
    carry += s[1023];
    __syncthreads();  // s is rewritten by the next chunk
  }
  if (t == 0) bsum[nb] = carry;
}

__global__ void __launch_bounds__(1024)
    k_scan_apply(const int *__restrict__ cnt, int n,
                 const int *__restrict__ bsum, int nb, int *__restrict__ start) {
  __shared__ int s[kScanBlock];
  const int b = blockIdx.x, t = threadIdx.x;
  const int i = b * kScanBlock + t;
  s[t] = (i < n) ? cnt[i] : 0;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int v = (t >= off) ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  if (i < n) start[i] = bsum[b] + ((t == 0) ? 0 : s[t - 1]);
  if (b == 0 && t == 0) start[n] = bsum[nb];
}

// One-workgroup exclusive scan for short count arrays (n <= kScanOneMax):
// thread t holds the run [16 t, 16 t + 16) in registers (four 16-byte loads
// issued together), the 1024 run sums are scanned by wavefront prefix sums
// and one more over the 16 wavefront totals, then each thread writes its
// run's starts. With clear != 0 the counts cnt[0..n] (including the extra
// counter cnt[n]) are zeroed after they are read, so the next counting pass
// needs no fill kernel. act: loop-control flag (nullptr: always run; 0: the
// launch returns at once). cnt and start: 16-byte aligned.
constexpr int kScanOneMax = 16 * 1024;
__device__ __forceinline__ void d_scan_one(const int *__restrict__ act,
                                           int *__restrict__ cnt, int n,
                                           int *__restrict__ start, int clear) {
  if (act && !*act) return;
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int lo = 16 * t;
  const bool full = lo + 16 <= n;
  int v[16];
  if (full) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int4 c = reinterpret_cast<const int4 *>(cnt + lo)[q];
      v[4 * q] = c.x;
      v[4 * q + 1] = c.y;
      v[4 * q + 2] = c.z;
      v[4 * q + 3] = c.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = lo + j < n ? cnt[lo + j] : 0;
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += v[j];
  int x = s;  // inclusive prefix over the wavefront's runs
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    int u = lane < 16 ? wsum[lane] : 0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const int y = __shfl_up(u, off, 64);
      if (lane >= off) u += y;
    }
    if (lane < 16) wsum[lane] = u;
  }
  __syncthreads();
  int base = (w ? wsum[w - 1] : 0) + x - s;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int c = v[j];
    v[j] = base;
    base += c;
  }
  if (full) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      reinterpret_cast<int4 *>(start + lo)[q] =
          int4{v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
      if (clear) reinterpret_cast<int4 *>(cnt + lo)[q] = int4{0, 0, 0, 0};
    }
  } else {
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (lo + j < n) {
        start[lo + j] = v[j];
        if (clear) cnt[lo + j] = 0;
      }
  }
  if (t == 1023) start[n] = wsum[15];
  if (clear && t == 0) cnt[n] = 0;
}
__global__ void __launch_bounds__(1024)
    k_scan_one(const int *__restrict__ act, int *__restrict__ cnt, int n,
               int *__restrict__ start, int clear) {
  d_scan_one(act, cnt, n, start, clear);
}

}  // namespace

// start[0..n] = exclusive scan of cnt[0..n) in one workgroup (n <=
// kScanOneMax); clear != 0 also zeroes cnt[0..n].
inline int device_scan_one(const int *act, int *cnt, int n, int *start,
                           int clear, hipStream_t st) {
  FS_CHECK(n >= 1 && n <= kScanOneMax, FS_ERR_ARG,
           "one-workgroup scan of %d elements", n);
  hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, st, act, cnt, n, start,
                     clear);
  return FS_OK;
}

// start[0..n] = exclusive scan of cnt[0..n); scratch: >= n/1024 + 2 ints.
inline int device_scan(const int *cnt, int n, int *start, int *scratch,
                       hipStream_t st) {
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  FS_CHECK(nb >= 1 && nb <= kScanMaxBlocks, FS_ERR_ARG,
           "scan of %d elements too large", n);
  hipLaunchKernelGGL(k_scan_blocksum, dim3(nb), dim3(256), 0, st, cnt, n,
                     scratch);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, scratch, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, st, cnt, n,
                     scratch, nb, start);
  return FS_OK;
}

}  // namespace fs
