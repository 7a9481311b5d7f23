// fcm.hip -- Force Coupling Method mobility solve for gfx950.
//
// Replaces Fcm::mobility_solve (FCMfunctions.cpp:175-464). One solve runs
//   prepare   per-segment 1D Gaussian tables + stencil bin   (gaussian_setup,
//             FCMfunctions.cpp:549-593; GA_setup :536-547)
//   bin       counting sort of segments by the grid tile holding their
//             stencil origin (deterministic order inside a bin)
//   spread    gather form of mono_and_dipole_force_distribution
//             (:596-646): one workgroup per grid tile sums, in a fixed
//             order, every segment whose stencil covers the tile, then
//             writes each grid point once (no atomics, no zeroing pass)
//   fft_z_r2c real->half-complex along z, two real rows per complex FFT
//   fft_y     strided complex FFT along y (forward)
//   fft_x_prj forward FFT along x, the Stokes projection of flow_solve
//             (:648-736), inverse FFT along x -- three components resident
//             in LDS, so the spectrum makes one round trip through HBM
//   fft_y     inverse along y
//   fft_z_c2r half-complex->real along z
//   interp    particle_velocities + particle_rotations (:740-808) fused:
//             one wavefront per segment, segments visited in bin order
// Grids: real R[3][nx][ny][nz], half spectrum C[3][nx][ny][nz/2+1] (complex).
// Precision: FS_FP64 runs every stage in double; FS_FP32 runs them in float
// with float accumulation (stencil offsets are formed from the FP64
// positions before rounding).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "fcm_impl.hpp"
#include "fft.hpp"

namespace fs {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kMaxNgd = 24;     // largest Gaussian support handled
constexpr int kCandChunk = 16;  // segments staged in LDS per spread chunk
constexpr int kMaxBinsPerTile = 64;

// ---------------------------------------------------------------------
// prepare: tables [seg][9][ngd] (ga x,y,z | gd x,y,z | gr x,y,z), stencil
// origin (wrapped index of support point 0), bin, spread coefficients.
// ---------------------------------------------------------------------
struct GaussConst {
  double dx, anorm, anorm2, anormdip, anormdip2, sigmadipsq;
  int ngd, ngdh, nx, ny, nz;
};

template <typename T>
__global__ void k_prepare(int np, const double *__restrict__ Y,
                          const double *__restrict__ F,
                          const double *__restrict__ Tq, GaussConst g,
                          int tx, int ty, int tz, int nbx, int nby,
                          T *__restrict__ tab, int *__restrict__ origin,
                          T *__restrict__ coef, int *__restrict__ bin,
                          int *__restrict__ bin_count, int *__restrict__ slot) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  const int nn[3] = {g.nx, g.ny, g.nz};
  int o[3];
  T *ts = tab + (size_t)s * 9 * g.ngd;
  for (int d = 0; d < 3; ++d) {
    double y = Y[3 * s + d];
    int c = (int)(y / g.dx);  // truncation, as FCMfunctions.cpp:565-567
    int x0 = c - g.ngdh + 1;
    int w = x0 % nn[d];
    if (w < 0) w += nn[d];
    o[d] = w;
    for (int i = 0; i < g.ngd; ++i) {
      int xg = x0 + i;
      double xx = ((double)xg) * g.dx - y;
      double e1 = g.anorm * exp(-xx * xx / g.anorm2);
      double e2 = g.anormdip * exp(-xx * xx / g.anormdip2);
      double gr = -xx / g.sigmadipsq;
      ts[(0 + d) * g.ngd + i] = (T)e1;
      ts[(3 + d) * g.ngd + i] = (T)e2;
      ts[(6 + d) * g.ngd + i] = (T)gr;
    }
  }
  origin[3 * s + 0] = o[0];
  origin[3 * s + 1] = o[1];
  origin[3 * s + 2] = o[2];
  // spread coefficients: F and half torque (GA_setup: GA = T/2 pattern)
  for (int d = 0; d < 3; ++d) {
    coef[6 * s + d] = (T)F[3 * s + d];
    coef[6 * s + 3 + d] = (T)(0.5 * Tq[3 * s + d]);
  }
  int b = (o[0] / tx) + nbx * ((o[1] / ty) + nby * (o[2] / tz));
  bin[s] = b;
  slot[s] = atomicAdd(&bin_count[b], 1);
}

__global__ void k_fill_int(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Exclusive scan of n counts into start[0..n] by one workgroup of 1024.
__global__ void k_scan(const int *__restrict__ cnt, int n,
                       int *__restrict__ start) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int lo = min(n, t * per), hi = min(n, lo + per);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t == 0) ? 0 : part[t - 1];
  for (int i = lo; i < hi; ++i) {
    start[i] = run;
    run += cnt[i];
  }
  if (t == 1023) start[n] = part[1023];
}

__global__ void k_scatter(int np, const int *__restrict__ bin,
                          const int *__restrict__ slot,
                          const int *__restrict__ start, int *__restrict__ ids) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  ids[start[bin[s]] + slot[s]] = s;
}

// Ascending segment order inside each bin (atomic slots are unordered).
__global__ void k_sort_bins(int nbins, const int *__restrict__ start,
                            int *__restrict__ ids) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int lo = start[b], hi = start[b + 1];
  for (int i = lo + 1; i < hi; ++i) {
    int v = ids[i];
    int j = i - 1;
    while (j >= lo && ids[j] > v) {
      ids[j + 1] = ids[j];
      --j;
    }
    ids[j + 1] = v;
  }
}

// ---------------------------------------------------------------------
// spread (gather): one workgroup of 256 per tile of tx*ty*tz <= 1024 points
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
    k_spread(SpreadGeom G, const T *__restrict__ tab,
             const int *__restrict__ origin, const T *__restrict__ coef,
             const int *__restrict__ start, const int *__restrict__ ids,
             T *__restrict__ R) {
  __shared__ int s_bstart[kMaxBinsPerTile + 1];
  __shared__ int s_bfirst[kMaxBinsPerTile];
  __shared__ int s_ntot;
  __shared__ T s_tab[kCandChunk][9 * kMaxNgd];
  __shared__ int s_org[kCandChunk][3];
  __shared__ T s_coef[kCandChunk][6];

  const int ntiles = G.nbx * G.nby * G.nbz;
  // XCD-aware order: consecutive tiles land on the same XCD's L2.
  int bid = blockIdx.x;
  if ((ntiles & 7) == 0) bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
  const int bx = bid % G.nbx;
  const int by = (bid / G.nbx) % G.nby;
  const int bz = bid / (G.nbx * G.nby);
  const int ngd = G.ngd;

  // bins that can reach this tile: origins in [tile - (ngd-1), tile_end)
  const int cx = min(G.hx + 1, G.nbx), cy = min(G.hy + 1, G.nby),
            cz = min(G.hz + 1, G.nbz);
  const int nbins = cx * cy * cz;
  if (threadIdx.x < nbins) {
    int t = threadIdx.x;
    int dz = t / (cx * cy), dy = (t / cx) % cy, dx = t % cx;
    int ix = ((bx - G.hx + dx) % G.nbx + G.nbx) % G.nbx;
    int iy = ((by - G.hy + dy) % G.nby + G.nby) % G.nby;
    int iz = ((bz - G.hz + dz) % G.nbz + G.nbz) % G.nbz;
    int b = ix + G.nbx * (iy + G.nby * iz);
    s_bfirst[t] = start[b];
    s_bstart[t + 1] = start[b + 1] - start[b];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_bstart[0] = 0;
    for (int t = 0; t < nbins; ++t) s_bstart[t + 1] += s_bstart[t];
    s_ntot = s_bstart[nbins];
  }
  __syncthreads();
  const int ntot = s_ntot;

  const int npts = G.tx * G.ty * G.tz;
  constexpr int PPT = 4;
  T acc[PPT][3];
  int px[PPT], py[PPT], pz[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    acc[k][0] = acc[k][1] = acc[k][2] = (T)0;
    int p = threadIdx.x + k * 256;
    int lz = p % G.tz, ly = (p / G.tz) % G.ty, lx = p / (G.tz * G.ty);
    px[k] = bx * G.tx + lx;
    py[k] = by * G.ty + ly;
    pz[k] = bz * G.tz + lz;
  }

  for (int c0 = 0; c0 < ntot; c0 += kCandChunk) {
    const int nc = min(kCandChunk, ntot - c0);
    // stage candidate tables
    for (int e = threadIdx.x; e < nc * 9 * ngd; e += 256) {
      int c = e / (9 * ngd), r = e % (9 * ngd);
      int gc = c0 + c;
      int t = 0;
      while (gc >= s_bstart[t + 1]) ++t;
      int s = ids[s_bfirst[t] + gc - s_bstart[t]];
      s_tab[c][r] = tab[(size_t)s * 9 * ngd + r];
      if (r < 3) s_org[c][r] = origin[3 * s + r];
      if (r < 6) s_coef[c][r] = coef[6 * s + r];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      if (threadIdx.x + k * 256 >= npts) break;
      for (int c = 0; c < nc; ++c) {
        int ox = px[k] - s_org[c][0];
        if (ox < 0) ox += G.nx;
        int oy = py[k] - s_org[c][1];
        if (oy < 0) oy += G.ny;
        int oz = pz[k] - s_org[c][2];
        if (oz < 0) oz += G.nz;
        if (ox >= ngd || oy >= ngd || oz >= ngd) continue;
        const T *tb = s_tab[c];
        T temp = tb[0 * ngd + ox] * tb[1 * ngd + oy] * tb[2 * ngd + oz];
        T temp2 = tb[3 * ngd + ox] * tb[4 * ngd + oy] * tb[5 * ngd + oz];
        T xx = tb[6 * ngd + ox], yy = tb[7 * ngd + oy], zz = tb[8 * ngd + oz];
        const T *cf = s_coef[c];
        acc[k][0] += cf[0] * temp + (cf[5] * yy - cf[4] * zz) * temp2;
        acc[k][1] += cf[1] * temp + (cf[3] * zz - cf[5] * xx) * temp2;
        acc[k][2] += cf[2] * temp + (cf[4] * xx - cf[3] * yy) * temp2;
      }
    }
    __syncthreads();
  }
  const size_t N = (size_t)G.nx * G.ny * G.nz;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (threadIdx.x + k * 256 >= npts) break;
    size_t o = ((size_t)px[k] * G.ny + py[k]) * G.nz + pz[k];
    R[o] = acc[k][0];
    R[N + o] = acc[k][1];
    R[2 * N + o] = acc[k][2];
  }
}

// ---------------------------------------------------------------------
// z passes: P row pairs per workgroup, LDS layout [pair][n]
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
    k_fft_z_r2c(const T *__restrict__ R, cx<T> *__restrict__ C, FftPlan pl,
                const cx<T> *__restrict__ twg, long nrows, int P) {
  extern __shared__ unsigned char smem_raw[];
  const int n = pl.n, nh = n / 2 + 1;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *a = tw + n;
  cx<T> *b = a + (size_t)P * n;
  const int comp = blockIdx.y;
  const long npairs = (nrows + 1) / 2;
  const long pair0 = (long)blockIdx.x * P;
  const T *Rc = R + (size_t)comp * nrows * n;
  cx<T> *Cc = C + (size_t)comp * nrows * nh;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
    int m = e % n, p = e / n;
    long pr = pair0 + p;
    long ra = 2 * pr, rb = 2 * pr + 1;
    T va = (pr < npairs) ? Rc[(size_t)ra * n + m] : (T)0;
    T vb = (pr < npairs && rb < nrows) ? Rc[(size_t)rb * n + m] : (T)0;
    a[p * n + m] = {va, vb};
  }
  __syncthreads();
  cx<T> *res = lds_fft(a, b, pl, P, 1, n, tw, false);
  for (int e = threadIdx.x; e < P * nh; e += blockDim.x) {
    int k = e % nh, p = e / nh;
    long pr = pair0 + p;
    if (pr >= npairs) continue;
    long ra = 2 * pr, rb = 2 * pr + 1;
    cx<T> z = res[p * n + k];
    cx<T> w = res[p * n + (k == 0 ? 0 : n - k)];
    const T half = (T)0.5;
    Cc[(size_t)ra * nh + k] = {half * (z.x + w.x), half * (z.y - w.y)};
    if (rb < nrows) Cc[(size_t)rb * nh + k] = {half * (z.y + w.y), half * (w.x - z.x)};
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_fft_z_c2r(const cx<T> *__restrict__ C, T *__restrict__ R, FftPlan pl,
                const cx<T> *__restrict__ twg, long nrows, int P) {
  extern __shared__ unsigned char smem_raw[];
  const int n = pl.n, nh = n / 2 + 1;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *a = tw + n;
  cx<T> *b = a + (size_t)P * n;
  const int comp = blockIdx.y;
  const long npairs = (nrows + 1) / 2;
  const long pair0 = (long)blockIdx.x * P;
  T *Rc = R + (size_t)comp * nrows * n;
  const cx<T> *Cc = C + (size_t)comp * nrows * nh;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  for (int e = threadIdx.x; e < P * nh; e += blockDim.x) {
    int k = e % nh, p = e / nh;
    long pr = pair0 + p;
    long ra = 2 * pr, rb = 2 * pr + 1;
    cx<T> A = (pr < npairs) ? Cc[(size_t)ra * nh + k] : cx<T>{0, 0};
    cx<T> B = (pr < npairs && rb < nrows) ? Cc[(size_t)rb * nh + k] : cx<T>{0, 0};
    if (k == 0 || 2 * k == n) {  // DC / Nyquist of a real line are real
      A.y = 0;
      B.y = 0;
    }
    a[p * n + k] = {A.x - B.y, A.y + B.x};
    if (k > 0 && 2 * k < n) a[p * n + n - k] = {A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  cx<T> *res = lds_fft(a, b, pl, P, 1, n, tw, true);
  for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
    int m = e % n, p = e / n;
    long pr = pair0 + p;
    if (pr >= npairs) continue;
    long ra = 2 * pr, rb = 2 * pr + 1;
    cx<T> z = res[p * n + m];
    Rc[(size_t)ra * n + m] = z.x;
    if (rb < nrows) Rc[(size_t)rb * n + m] = z.y;
  }
}

// ---------------------------------------------------------------------
// y pass: one workgroup per (comp, x, kz chunk of B columns)
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
    k_fft_y(cx<T> *__restrict__ C, FftPlan pl, const cx<T> *__restrict__ twg,
            int nx, int nh, int B, int inverse) {
  extern __shared__ unsigned char smem_raw[];
  const int n = pl.n;  // ny
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *a = tw + n;
  cx<T> *b = a + (size_t)B * n;
  const int nchunk = (nh + B - 1) / B;
  const int chunk = blockIdx.x % nchunk;
  const int x = blockIdx.x / nchunk;
  const int comp = blockIdx.y;
  const int kz0 = chunk * B;
  cx<T> *base = C + ((size_t)comp * nx + x) * n * nh;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
    int col = e % B, m = e / B;
    int kz = kz0 + col;
    a[m * B + col] = (kz < nh) ? base[(size_t)m * nh + kz] : cx<T>{0, 0};
  }
  __syncthreads();
  cx<T> *res = lds_fft(a, b, pl, B, B, 1, tw, inverse != 0);
  for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
    int col = e % B, m = e / B;
    int kz = kz0 + col;
    if (kz < nh) base[(size_t)m * nh + kz] = res[m * B + col];
  }
}

// ---------------------------------------------------------------------
// x pass + projection: one workgroup per (y, kz chunk), all 3 components
// ---------------------------------------------------------------------
struct ProjConst {
  int nx, ny, nz, nh;
  double kfx, kfy, kfz;  // 2 pi / L per axis
  double scale;          // 1 / (N mu)
};

template <typename T>
__global__ void __launch_bounds__(256)
    k_fft_x_project(cx<T> *__restrict__ C, FftPlan pl,
                    const cx<T> *__restrict__ twg, ProjConst pc, int B) {
  extern __shared__ unsigned char smem_raw[];
  const int n = pl.n;  // nx
  const int nh = pc.nh;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf[4];
  buf[0] = tw + n;
  for (int i = 1; i < 4; ++i) buf[i] = buf[i - 1] + (size_t)B * n;
  const int nchunk = (nh + B - 1) / B;
  const int chunk = blockIdx.x % nchunk;
  const int y = blockIdx.x / nchunk;
  const int kz0 = chunk * B;
  const size_t comp_stride = (size_t)pc.nx * pc.ny * nh;
  const size_t xs = (size_t)pc.ny * nh;
  cx<T> *base = C + (size_t)y * nh;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  for (int c = 0; c < 3; ++c)
    for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
      int col = e % B, m = e / B;
      int kz = kz0 + col;
      buf[c][m * B + col] =
          (kz < nh) ? base[c * comp_stride + m * xs + kz] : cx<T>{0, 0};
    }
  __syncthreads();
  cx<T> *cur[3];
  cx<T> *spare = buf[3];
  for (int c = 0; c < 3; ++c) {
    cx<T> *r = lds_fft(buf[c], spare, pl, B, B, 1, tw, false);
    if (r == spare) {
      spare = buf[c];
    }
    cur[c] = r;
  }
  // Stokes projection (flow_solve, FCMfunctions.cpp:648-736)
  const T q2 = (T)(((y <= pc.ny / 2) ? y : y - pc.ny) * pc.kfy);
  for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
    int col = e % B, m = e / B;
    int kz = kz0 + col;
    T q1 = (T)(((m <= pc.nx / 2) ? m : m - pc.nx) * pc.kfx);
    T q3 = (T)(kz * pc.kfz);
    T qq = q2 * q2 + q1 * q1 + (T)1e-18;
    T norm = (T)1 / (qq + q3 * q3);
    T sc = (T)pc.scale;
    cx<T> f1 = cur[0][m * B + col], f2 = cur[1][m * B + col],
          f3 = cur[2][m * B + col];
    cx<T> u1, u2, u3;
    T kf = (q1 * f1.x + q2 * f2.x + q3 * f3.x) * norm;
    u1.x = norm * (f1.x - q1 * kf) * sc;
    u2.x = norm * (f2.x - q2 * kf) * sc;
    u3.x = norm * (f3.x - q3 * kf) * sc;
    kf = (q1 * f1.y + q2 * f2.y + q3 * f3.y) * norm;
    u1.y = norm * (f1.y - q1 * kf) * sc;
    u2.y = norm * (f2.y - q2 * kf) * sc;
    u3.y = norm * (f3.y - q3 * kf) * sc;
    if (m == 0 && y == 0 && kz == 0) u1 = u2 = u3 = cx<T>{0, 0};
    cur[0][m * B + col] = u1;
    cur[1][m * B + col] = u2;
    cur[2][m * B + col] = u3;
  }
  __syncthreads();
  for (int c = 0; c < 3; ++c) {
    cx<T> *r = lds_fft(cur[c], spare, pl, B, B, 1, tw, true);
    if (r == spare) spare = cur[c];
    cur[c] = r;
  }
  for (int c = 0; c < 3; ++c)
    for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
      int col = e % B, m = e / B;
      int kz = kz0 + col;
      if (kz < nh) base[c * comp_stride + m * xs + kz] = cur[c][m * B + col];
    }
}

// ---------------------------------------------------------------------
// interpolation: one 64-lane wavefront per segment (bin order)
// ---------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_interp(int np, SpreadGeom G, double dx3, const T *__restrict__ tab,
             const int *__restrict__ origin, const int *__restrict__ ids,
             const T *__restrict__ R, double *__restrict__ V,
             double *__restrict__ W) {
  __shared__ T s_tab[4][9 * kMaxNgd];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wave;
  const bool valid = w < np;
  const int s = valid ? ids[w] : 0;
  const int ngd = G.ngd;
  if (valid)
    for (int r = lane; r < 9 * ngd; r += 64)
      s_tab[wave][r] = tab[(size_t)s * 9 * ngd + r];
  const int ox = valid ? origin[3 * s] : 0, oy = valid ? origin[3 * s + 1] : 0,
            oz = valid ? origin[3 * s + 2] : 0;
  __syncthreads();
  const T *tb = s_tab[wave];
  const size_t N = (size_t)G.nx * G.ny * G.nz;
  T v0 = 0, v1 = 0, v2 = 0, w0 = 0, w1 = 0, w2 = 0;
  const int npts = valid ? ngd * ngd * ngd : 0;
  for (int p = lane; p < npts; p += 64) {
    int k = p % ngd, j = (p / ngd) % ngd, i = p / (ngd * ngd);
    int ii = ox + i;
    if (ii >= G.nx) ii -= G.nx;
    int jj = oy + j;
    if (jj >= G.ny) jj -= G.ny;
    int kk = oz + k;
    if (kk >= G.nz) kk -= G.nz;
    size_t o = ((size_t)ii * G.ny + jj) * G.nz + kk;
    T ux = R[o], uy = R[N + o], uz = R[2 * N + o];
    T temp = tb[0 * ngd + i] * tb[1 * ngd + j] * tb[2 * ngd + k];
    T temp2 = tb[3 * ngd + i] * tb[4 * ngd + j] * tb[5 * ngd + k];
    T xx = tb[6 * ngd + i], yy = tb[7 * ngd + j], zz = tb[8 * ngd + k];
    v0 += ux * temp;
    v1 += uy * temp;
    v2 += uz * temp;
    w0 += (uz * yy - uy * zz) * temp2;
    w1 += (ux * zz - uz * xx) * temp2;
    w2 += (uy * xx - ux * yy) * temp2;
  }
  v0 = wave_sum(v0);
  v1 = wave_sum(v1);
  v2 = wave_sum(v2);
  w0 = wave_sum(w0);
  w1 = wave_sum(w1);
  w2 = wave_sum(w2);
  if (valid && lane == 0) {
    V[3 * s + 0] = dx3 * (double)v0;
    V[3 * s + 1] = dx3 * (double)v1;
    V[3 * s + 2] = dx3 * (double)v2;
    W[3 * s + 0] = -0.5 * dx3 * (double)w0;
    W[3 * s + 1] = -0.5 * dx3 * (double)w1;
    W[3 * s + 2] = -0.5 * dx3 * (double)w2;
  }
}

__global__ void k_friction(int n3, const double *__restrict__ F,
                           const double *__restrict__ Tq, double a1, double a2,
                           double *__restrict__ V, double *__restrict__ W) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) {
    V[i] = a1 * F[i];
    W[i] = a2 * Tq[i];
  }
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static bool good_size(int n) {
  if (n < 2 || n > 1024 || (n & 1)) return false;
  FftPlan p;
  return make_plan(n, &p);
}

static void twiddles_host(int n, std::vector<double> &re,
                          std::vector<double> &im) {
  re.resize(n);
  im.resize(n);
  for (int m = 0; m < n; ++m) {
    long double ang = 2.0L * (long double)kPi * (long double)m / (long double)n;
    re[m] = (double)cosl(ang);
    im[m] = -(double)sinl(ang);
  }
}

template <typename T>
static int upload_tw(dbuf &b, int n) {
  std::vector<double> re, im;
  twiddles_host(n, re, im);
  std::vector<cx<T>> h(n);
  for (int m = 0; m < n; ++m) h[m] = {(T)re[m], (T)im[m]};
  FS_TRY(b.alloc(sizeof(cx<T>) * n));
  FS_HIP(hipMemcpy(b.p, h.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
  return FS_OK;
}

int FcmImpl::init(const fs_fcm_params &pp, int maxnp, int prec) {
  p = pp;
  max_np = maxnp;
  precision = prec;
  FS_CHECK(prec == FS_FP64 || prec == FS_FP32, FS_ERR_ARG,
           "precision must be FS_FP64 or FS_FP32");
  FS_CHECK(good_size(p.nx) && good_size(p.ny) && good_size(p.nz), FS_ERR_ARG,
           "grid %dx%dx%d: each size must be even, <= 1024 and 2^a 3^b 5^c",
           p.nx, p.ny, p.nz);
  FS_CHECK(p.ngd >= 2 && p.ngd <= kMaxNgd && p.ngd <= p.nx &&
               p.ngd <= p.ny && p.ngd <= p.nz,
           FS_ERR_ARG, "ngd=%d must be in [2,%d] and <= every grid size",
           p.ngd, kMaxNgd);
  FS_CHECK(p.dx > 0 && p.sigma > 0 && p.sigma_dip > 0 && p.mu > 0 && p.a > 0,
           FS_ERR_ARG, "dx, sigma, sigma_dip, mu, a must be positive");
  FS_CHECK(maxnp >= 1, FS_ERR_ARG, "max_np must be >= 1");
  make_plan(p.nx, &plx);
  make_plan(p.ny, &ply);
  make_plan(p.nz, &plz);
  tsz = (prec == FS_FP64) ? 8 : 4;
  const size_t N = (size_t)p.nx * p.ny * p.nz;
  const size_t NC = (size_t)p.nx * p.ny * (p.nz / 2 + 1);
  FS_TRY(R.alloc(3 * N * tsz));
  FS_TRY(C.alloc(3 * NC * 2 * tsz));
  if (prec == FS_FP64) {
    FS_TRY(upload_tw<double>(twx, p.nx));
    FS_TRY(upload_tw<double>(twy, p.ny));
    FS_TRY(upload_tw<double>(twz, p.nz));
  } else {
    FS_TRY(upload_tw<float>(twx, p.nx));
    FS_TRY(upload_tw<float>(twy, p.ny));
    FS_TRY(upload_tw<float>(twz, p.nz));
  }
  FS_TRY(tab.alloc((size_t)maxnp * 9 * p.ngd * tsz));
  FS_TRY(origin.alloc((size_t)maxnp * 3 * sizeof(int)));
  FS_TRY(coef.alloc((size_t)maxnp * 6 * tsz));
  FS_TRY(bin.alloc((size_t)maxnp * sizeof(int)));
  FS_TRY(slot.alloc((size_t)maxnp * sizeof(int)));
  FS_TRY(ids.alloc((size_t)maxnp * sizeof(int)));
  tune.fft_batch = (prec == FS_FP64) ? 4 : 8;
  tune.spread_tile = 8;
  tune.interp_wave = 1;
  FS_TRY(set_geometry());
  for (int i = 0; i < 8; ++i) FS_HIP(hipEventCreate(&ev[i]));
  // FFT kernels stage up to 160 KB of LDS per workgroup
  const int lds_max = 160 * 1024;
  if (prec == FS_FP64) {
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_z_r2c<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_z_c2r<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_y<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_x_project<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  } else {
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_z_r2c<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_z_c2r<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_y<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
    FS_HIP(hipFuncSetAttribute((const void *)k_fft_x_project<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  }
  return FS_OK;
}

static int largest_divisor_le(int n, int t) {
  for (int d = std::min(n, t); d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

int FcmImpl::set_geometry() {
  int t = tune.spread_tile;
  FS_CHECK(t == 4 || t == 8 || t == 16, FS_ERR_ARG,
           "spread_tile must be 4, 8 or 16");
  // target tile (x, y, z); z is the contiguous axis
  int tgt[3] = {8, 8, 8};
  if (t == 4) tgt[0] = tgt[1] = 4, tgt[2] = 32;
  if (t == 16) tgt[2] = 16;
  const int n[3] = {p.nx, p.ny, p.nz};
  int td[3], bins = 0;
  for (int attempt = 0; attempt < 6; ++attempt) {
    bins = 1;
    for (int d = 0; d < 3; ++d) {
      td[d] = largest_divisor_le(n[d], tgt[d]);
      int nb = n[d] / td[d];
      int h = (p.ngd - 1 + td[d] - 1) / td[d];
      bins *= std::min(h + 1, nb);
    }
    if (bins <= kMaxBinsPerTile && td[0] * td[1] * td[2] <= 1024) break;
    // grow the smallest tile edge and retry
    int dmin = 0;
    for (int d = 1; d < 3; ++d)
      if (td[d] < td[dmin]) dmin = d;
    tgt[dmin] *= 2;
  }
  FS_CHECK(bins <= kMaxBinsPerTile && td[0] * td[1] * td[2] <= 1024,
           FS_ERR_ARG, "no spread tile fits grid %dx%dx%d with ngd=%d", p.nx,
           p.ny, p.nz, p.ngd);
  geom.nx = p.nx;
  geom.ny = p.ny;
  geom.nz = p.nz;
  geom.ngd = p.ngd;
  geom.tx = td[0];
  geom.ty = td[1];
  geom.tz = td[2];
  geom.nbx = p.nx / td[0];
  geom.nby = p.ny / td[1];
  geom.nbz = p.nz / td[2];
  geom.hx = (p.ngd - 1 + td[0] - 1) / td[0];
  geom.hy = (p.ngd - 1 + td[1] - 1) / td[1];
  geom.hz = (p.ngd - 1 + td[2] - 1) / td[2];
  nbins = geom.nbx * geom.nby * geom.nbz;
  FS_TRY(bin_count.alloc((size_t)nbins * sizeof(int)));
  FS_TRY(bin_start.alloc((size_t)(nbins + 1) * sizeof(int)));
  return FS_OK;
}

void FcmImpl::release() {
  R.release();
  C.release();
  twx.release();
  twy.release();
  twz.release();
  tab.release();
  origin.release();
  coef.release();
  bin.release();
  slot.release();
  ids.release();
  bin_count.release();
  bin_start.release();
  for (int i = 0; i < 8; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
}

template <typename T>
int FcmImpl::run(int np, const double *Y, const double *F, const double *Tq,
                 double *V, double *W, hipStream_t st) {
  GaussConst g;
  g.dx = p.dx;
  double sigmadipsq = p.sigma_dip * p.sigma_dip;
  g.sigmadipsq = sigmadipsq;
  g.anormdip = 1.0 / std::sqrt(2.0 * kPi * sigmadipsq);
  g.anormdip2 = 2.0 * sigmadipsq;
  double sigmasq = p.sigma * p.sigma;
  g.anorm = 1.0 / std::sqrt(2.0 * kPi * sigmasq);
  g.anorm2 = 2.0 * sigmasq;
  g.ngd = p.ngd;
  g.ngdh = p.ngd / 2;
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = p.nz;

  auto mark = [&](int i) {
    if (timing) (void)hipEventRecord(ev[i], st);
  };
  mark(0);
  // bin counting sort
  hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(nbins, 256)), dim3(256), 0, st,
                     bin_count.as<int>(), nbins, 0);
  hipLaunchKernelGGL(k_prepare<T>, dim3(ceil_div(np, 128)), dim3(128), 0, st,
                     np, Y, F, Tq, g, geom.tx, geom.ty, geom.tz, geom.nbx,
                     geom.nby, tab.as<T>(), origin.as<int>(), coef.as<T>(),
                     bin.as<int>(), bin_count.as<int>(), slot.as<int>());
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, bin_count.as<int>(),
                     nbins, bin_start.as<int>());
  hipLaunchKernelGGL(k_scatter, dim3(ceil_div(np, 256)), dim3(256), 0, st, np,
                     bin.as<int>(), slot.as<int>(), bin_start.as<int>(),
                     ids.as<int>());
  hipLaunchKernelGGL(k_sort_bins, dim3(ceil_div(nbins, 256)), dim3(256), 0, st,
                     nbins, bin_start.as<int>(), ids.as<int>());
  mark(1);
  const int ntiles = nbins;
  hipLaunchKernelGGL(k_spread<T>, dim3(ntiles), dim3(256), 0, st, geom,
                     tab.as<T>(), origin.as<int>(), coef.as<T>(),
                     bin_start.as<int>(), ids.as<int>(), R.as<T>());
  FS_LAUNCHED();
  mark(2);
  const long nrows = (long)p.nx * p.ny;
  const int nh = p.nz / 2 + 1;
  // z: pairs per block so that LDS stays <= 64 KB
  int P = std::max(1, std::min(16, (int)(65536 / (2 * sizeof(cx<T>) * p.nz))));
  size_t shz = sizeof(cx<T>) * (p.nz + 2 * (size_t)P * p.nz);
  long npairs = (nrows + 1) / 2;
  hipLaunchKernelGGL(k_fft_z_r2c<T>, dim3(ceil_div(npairs, P), 3), dim3(256),
                     shz, st, R.as<T>(), C.as<cx<T>>(), plz, twz.as<cx<T>>(),
                     nrows, P);
  FS_LAUNCHED();
  const int B = tune.fft_batch;
  const int nchunk = (nh + B - 1) / B;
  size_t shy = sizeof(cx<T>) * (p.ny + 2 * (size_t)B * p.ny);
  hipLaunchKernelGGL(k_fft_y<T>, dim3(p.nx * nchunk, 3), dim3(256), shy, st,
                     C.as<cx<T>>(), ply, twy.as<cx<T>>(), p.nx, nh, B, 0);
  FS_LAUNCHED();
  mark(3);
  ProjConst pc;
  pc.nx = p.nx;
  pc.ny = p.ny;
  pc.nz = p.nz;
  pc.nh = nh;
  pc.kfx = 2.0 * kPi / (p.nx * p.dx);
  pc.kfy = 2.0 * kPi / (p.ny * p.dx);
  pc.kfz = 2.0 * kPi / (p.nz * p.dx);
  pc.scale = 1.0 / ((double)p.nx * p.ny * p.nz * p.mu);
  size_t shx = sizeof(cx<T>) * (p.nx + 4 * (size_t)B * p.nx);
  hipLaunchKernelGGL(k_fft_x_project<T>, dim3(p.ny * nchunk), dim3(256), shx,
                     st, C.as<cx<T>>(), plx, twx.as<cx<T>>(), pc, B);
  FS_LAUNCHED();
  mark(4);
  hipLaunchKernelGGL(k_fft_y<T>, dim3(p.nx * nchunk, 3), dim3(256), shy, st,
                     C.as<cx<T>>(), ply, twy.as<cx<T>>(), p.nx, nh, B, 1);
  hipLaunchKernelGGL(k_fft_z_c2r<T>, dim3(ceil_div(npairs, P), 3), dim3(256),
                     shz, st, C.as<cx<T>>(), R.as<T>(), plz, twz.as<cx<T>>(),
                     nrows, P);
  FS_LAUNCHED();
  mark(5);
  double dx3 = p.dx * p.dx * p.dx;
  hipLaunchKernelGGL(k_interp<T>, dim3(ceil_div(np, 4)), dim3(256), 0, st, np,
                     geom, dx3, tab.as<T>(), origin.as<int>(), ids.as<int>(),
                     R.as<T>(), V, W);
  FS_LAUNCHED();
  mark(6);
  return FS_OK;
}

int FcmImpl::mobility(int np, const double *Y, const double *F,
                      const double *Tq, double *V, double *W, hipStream_t st) {
  FS_CHECK(np >= 0, FS_ERR_ARG, "np < 0");
  FS_CHECK(np <= max_np, FS_ERR_SIZE, "np=%d exceeds max_np=%d", np, max_np);
  FS_CHECK(np == 0 || (Y && F && Tq && V && W), FS_ERR_ARG, "null array");
  if (np == 0) return FS_OK;
  last_np = np;
  if (precision == FS_FP64) return run<double>(np, Y, F, Tq, V, W, st);
  return run<float>(np, Y, F, Tq, V, W, st);
}

int FcmImpl::friction(int np, const double *F, const double *Tq, double *V,
                      double *W, hipStream_t st) {
  FS_CHECK(np >= 0 && np <= max_np, FS_ERR_SIZE, "bad np=%d", np);
  if (np == 0) return FS_OK;
  double a1 = 1. / (6 * kPi * p.a * p.mu);
  double a2 = 1. / (8 * kPi * p.a * p.a * p.a * p.mu);
  hipLaunchKernelGGL(k_friction, dim3(ceil_div(3 * np, 256)), dim3(256), 0,
                     st, 3 * np, F, Tq, a1, a2, V, W);
  FS_LAUNCHED();
  return FS_OK;
}

int FcmImpl::stage_times(float *ms7) {
  for (int i = 0; i < 7; ++i) ms7[i] = 0;
  if (!timing) return FS_OK;
  FS_HIP(hipEventSynchronize(ev[6]));
  float t[6];
  for (int i = 0; i < 6; ++i) FS_HIP(hipEventElapsedTime(&t[i], ev[i], ev[i + 1]));
  // buckets: make-zero, gauss, spread, fft fwd, flow, fft bwd, interp
  ms7[0] = 0.f;     // no zeroing pass: the spread writes every point once
  ms7[1] = t[0];    // tables + bin sort
  ms7[2] = t[1];    // spread
  ms7[3] = t[2];    // z r2c + y forward
  ms7[4] = t[3];    // x forward + projection + x inverse (fused)
  ms7[5] = t[4];    // y inverse + z c2r
  ms7[6] = t[5];    // interpolation
  return FS_OK;
}

}  // namespace fs

// ---------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------
using fs::FcmImpl;

struct fs_fcm {
  FcmImpl impl;
};

extern "C" int fs_fcm_create(const fs_fcm_params *p, int max_np, int precision,
                             fs_fcm **out) {
  FS_CHECK(p && out, FS_ERR_ARG, "null argument");
  *out = nullptr;
  fs_fcm *h = new fs_fcm();
  int rc = h->impl.init(*p, max_np, precision);
  if (rc != FS_OK) {
    h->impl.release();
    delete h;
    return rc;
  }
  *out = h;
  return FS_OK;
}

extern "C" int fs_fcm_destroy(fs_fcm *h) {
  if (!h) return FS_OK;
  (void)hipDeviceSynchronize();
  h->impl.release();
  delete h;
  return FS_OK;
}

extern "C" int fs_fcm_mobility(fs_fcm *h, int np, const double *Y,
                               const double *F, const double *T, double *V,
                               double *W, void *stream) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.mobility(np, Y, F, T, V, W, (hipStream_t)stream);
}

extern "C" int fs_fcm_mobility_host(fs_fcm *h, int np, const double *Y,
                                    const double *F, const double *T,
                                    double *V, double *W) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(np >= 0 && np <= h->impl.max_np, FS_ERR_SIZE, "bad np=%d", np);
  if (np == 0) return FS_OK;
  fs::dbuf d;
  FS_TRY(d.alloc(sizeof(double) * 15 * (size_t)np));
  double *dY = d.as<double>(), *dF = dY + 3 * np, *dT = dF + 3 * np,
         *dV = dT + 3 * np, *dW = dV + 3 * np;
  size_t b = sizeof(double) * 3 * np;
  int rc = FS_OK;
  if (hipMemcpy(dY, Y, b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dF, F, b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dT, T, b, hipMemcpyHostToDevice) != hipSuccess) {
    fs::set_error("host to device copy failed");
    rc = FS_ERR_HIP;
  }
  if (rc == FS_OK) rc = h->impl.mobility(np, dY, dF, dT, dV, dW, nullptr);
  if (rc == FS_OK && (hipMemcpy(V, dV, b, hipMemcpyDeviceToHost) != hipSuccess ||
                      hipMemcpy(W, dW, b, hipMemcpyDeviceToHost) != hipSuccess)) {
    fs::set_error("device to host copy failed");
    rc = FS_ERR_HIP;
  }
  d.release();
  return rc;
}

extern "C" int fs_fcm_friction(fs_fcm *h, int np, const double *F,
                               const double *T, double *V, double *W,
                               void *stream) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.friction(np, F, T, V, W, (hipStream_t)stream);
}

extern "C" int fs_fcm_set_timing(fs_fcm *h, int on) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  h->impl.timing = on ? 1 : 0;
  return FS_OK;
}

extern "C" int fs_fcm_stage_times(fs_fcm *h, float *ms7) {
  FS_CHECK(h && ms7, FS_ERR_ARG, "null argument");
  return h->impl.stage_times(ms7);
}

extern "C" int fs_fcm_get_tuning(fs_fcm *h, fs_fcm_tuning *t) {
  FS_CHECK(h && t, FS_ERR_ARG, "null argument");
  *t = h->impl.tune;
  return FS_OK;
}

extern "C" int fs_fcm_set_tuning(fs_fcm *h, const fs_fcm_tuning *t) {
  FS_CHECK(h && t, FS_ERR_ARG, "null argument");
  fs_fcm_tuning old = h->impl.tune;
  fs_fcm_tuning nt = old;
  if (t->fft_batch > 0) nt.fft_batch = t->fft_batch;
  if (t->spread_tile > 0) nt.spread_tile = t->spread_tile;
  if (t->interp_wave >= 0) nt.interp_wave = t->interp_wave;
  FS_CHECK(nt.fft_batch == 1 || nt.fft_batch == 2 || nt.fft_batch == 4 ||
               nt.fft_batch == 8 || nt.fft_batch == 16,
           FS_ERR_ARG, "fft_batch must be 1,2,4,8 or 16");
  // the fused x pass holds 4*B*nx complex in LDS: cap at 160 KB
  size_t lds = (size_t)(h->impl.tsz * 2) * (h->impl.p.nx + 4 * (size_t)nt.fft_batch * h->impl.p.nx);
  FS_CHECK(lds <= 160 * 1024, FS_ERR_ARG, "fft_batch %d needs %zu B of LDS",
           nt.fft_batch, lds);
  h->impl.tune = nt;
  int rc = h->impl.set_geometry();
  if (rc != FS_OK) {
    h->impl.tune = old;
    h->impl.set_geometry();
  }
  return rc;
}
