// scan.hpp -- device-wide exclusive scan of int counts (counting sorts of
// the spread bins and the barrier cells): block sums, one-block scan of the
// block sums, then block-local scans with the block offsets.
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

}  // namespace

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
