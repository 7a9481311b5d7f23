// fcm.hip -- Force Coupling Method mobility solve for gfx950.
//
// Replaces Fcm::mobility_solve (FCMfunctions.cpp:175-464). One solve runs
//   prepare   per-segment 1D Gaussian tables + stencil bin   (gaussian_setup,
//             FCMfunctions.cpp:549-593; GA_setup :536-547)
//   bin       counting sort of segments by the grid tile holding their
//             stencil origin (deterministic order inside a bin); on the
//             matrix-core path each tile then gets the records of the
//             segments whose stencil meets it (k_tile_lists)
//   spread    gather form of mono_and_dipole_force_distribution
//             (:596-646): one workgroup per grid tile sums, in a fixed
//             order, every segment whose stencil covers the tile, then
//             writes each grid point once (no atomics, no zeroing pass);
//             by default as per-tile matrix-core products (k_spread_mfma,
//             both precisions), else with the vector gather (k_spread)
//   fft_z_r2c real->half-complex along z, two real rows per complex FFT
//   fft_y     strided complex FFT along y (forward)
//   fft_x_prj forward FFT along x, the Stokes projection of flow_solve
//             (:648-736), inverse FFT along x -- three components resident
//             in LDS, so the spectrum makes one round trip through HBM
//   fft_y     inverse along y
//   fft_z_c2r half-complex->real along z
//   interp    particle_velocities + particle_rotations (:740-808) fused:
//             on the matrix cores per tile (k_interp_mfma: the transpose
//             of the tile spread, partial sums added per segment in a
//             fixed order), else one wavefront per segment in bin order
// Grids: real R[3][nx][ny][nz], half spectrum C[3][nx][ny][nz/2+1] (complex).
// Precision: FS_FP64 runs every stage in double; FS_FP32 runs them in float
// with float accumulation (stencil offsets are formed from the FP64
// positions before rounding).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "fcm_impl.hpp"
#include "fft.hpp"
#include "scan.hpp"

namespace fs {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kMaxNgd = 24;     // largest Gaussian support handled
constexpr int kCandChunk = 16;  // segments staged in LDS per spread chunk
constexpr int kMaxBinsPerTile = 64;
constexpr int kCoef = 8;  // spread coefficients per segment (6 used)

// ---------------------------------------------------------------------
// prepare: tables [seg][9][ngd] (ga x,y,z | gd x,y,z | gr x,y,z), stencil
// origin (wrapped index of support point 0), bin, spread coefficients.
// ---------------------------------------------------------------------
struct GaussConst {
  double dx, anorm, anorm2, anormdip, anormdip2, sigmadipsq;
  int ngd, ngdh, nx, ny, nz;
  // FP32 weights: g = anorm 2^(-xx^2 e2), gD = anormdip 2^(-xx^2 ed2),
  // e2 = log2(e) / anorm2, ed2 = log2(e) / anormdip2
  double e2, ed2;
  double nrsd;  // -1 / sigmadipsq
};

// What a tile needs of one segment: per axis the FP64 offset xx0 = (grid
// coordinate of stencil point 0) * dx - y (point i lies at xx0 + i dx), its
// id, stencil origin (10 bits per axis, x lowest) and spread coefficients
// (force, half torque: GA_setup).
template <typename T>
struct alignas(16) SegRec {
  double xx0[3];
  int id, org;
  T cf[6];
};
constexpr int kTileCap = 32;  // records in a tile's slab; the rest: pool
inline size_t rec_bytes(int tsz) {
  return tsz == 8 ? sizeof(SegRec<double>) : sizeof(SegRec<float>);
}

// Grid cell holding coordinate y: truncation of y/dx, as
// FCMfunctions.cpp:565-567. A coordinate that is not finite or lies more
// than 2^30 cells out (a diverged iterate) gets cell s instead: no
// saturating conversion or signed overflow downstream, and such segments
// spread over many bins rather than piling into one (the per-bin sort is
// quadratic in the bin's population). Its tables are then 0 or NaN.
__device__ __forceinline__ int stencil_cell(double y, double dx, int s) {
  const double cy = y / dx;
  return fabs(cy) < 1073741824.0 ? (int)cy : s;
}

// Bin of tile (bx, by, bz): z fastest, so the bin-sorted segment list runs
// along the grid's contiguous axis and segments next to each other in it
// share cache lines of the real grids (interpolation reads them in list
// order). The spread enumerates its candidate bins by tile offset, so its
// summation order does not depend on this numbering.
__device__ __forceinline__ int spread_bin(int bx, int by, int bz, int nbx,
                                          int nby, int nbz) {
  return bz + nbz * (by + nby * bx);
}

template <typename T>
__global__ void k_prepare_tab(const int *__restrict__ act, int np, const double *__restrict__ Y,
                              GaussConst g, T *__restrict__ tab) {
  if (act && !*act) return;  // batched sweep: suspension converged
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)np * 3 * g.ngd) return;
  const int i = (int)(t % g.ngd), d = (int)((t / g.ngd) % 3),
            s = (int)(t / (3 * g.ngd));
  const double y = Y[3 * s + d];
  const int c = stencil_cell(y, g.dx, s);
  const int xg = c - g.ngdh + 1 + i;
  const double xx = ((double)xg) * g.dx - y;
  T *ts = tab + (size_t)s * 9 * g.ngd;
  ts[(0 + d) * g.ngd + i] = (T)(g.anorm * exp(-xx * xx / g.anorm2));
  ts[(3 + d) * g.ngd + i] = (T)(g.anormdip * exp(-xx * xx / g.anormdip2));
  ts[(6 + d) * g.ngd + i] = (T)(-xx / g.sigmadipsq);
}

template <typename T, typename GC = GaussConst>
__device__ __forceinline__ void d_prepare(const int *__restrict__ act, int np, const double *__restrict__ Y,
                          const double *__restrict__ F,
                          const double *__restrict__ Tq, GC g,
                          int tx, int ty, int tz, int nbx, int nby, int nbz,
                          int *__restrict__ origin, T *__restrict__ coef,
                          int *__restrict__ bin, int *__restrict__ bin_count,
                          int *__restrict__ slot) {
  if (act && !*act) return;  // batched sweep: suspension converged
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  const int nn[3] = {g.nx, g.ny, g.nz};
  int o[3];
  for (int d = 0; d < 3; ++d) {
    int c = stencil_cell(Y[3 * s + d], g.dx, s);
    int w = (c - g.ngdh + 1) % nn[d];
    if (w < 0) w += nn[d];
    o[d] = w;
    origin[3 * s + d] = w;
  }
  // spread coefficients: F and half torque (GA_setup: GA = T/2 pattern),
  // kCoef per segment so a record is whole 16-byte words
  for (int d = 0; d < 3; ++d) {
    coef[kCoef * s + d] = (T)F[3 * s + d];
    coef[kCoef * s + 3 + d] = (T)(0.5 * Tq[3 * s + d]);
  }
  coef[kCoef * s + 6] = (T)0;
  coef[kCoef * s + 7] = (T)0;
  int b = spread_bin(o[0] / tx, o[1] / ty, o[2] / tz, nbx, nby, nbz);
  bin[s] = b;
  slot[s] = atomicAdd(&bin_count[b], 1);
}
template <typename T>
__global__ void k_prepare(const int *__restrict__ act, int np, const double *__restrict__ Y,
                          const double *__restrict__ F,
                          const double *__restrict__ Tq, GaussConst g,
                          int tx, int ty, int tz, int nbx, int nby, int nbz,
                          int *__restrict__ origin, T *__restrict__ coef,
                          int *__restrict__ bin, int *__restrict__ bin_count,
                          int *__restrict__ slot) {
  d_prepare<T>(act, np, Y, F, Tq, g, tx, ty, tz, nbx, nby, nbz, origin, coef, bin, bin_count, slot);
}
// arguments of one suspension for k_prepare_multi
template <typename T>
struct A_prepare {
  const int *act;
  int np;
  const double *Y;
  const double *F;
  const double *Tq;
  GaussConst g;
  int tx;
  int ty;
  int tz;
  int nbx;
  int nby;
  int nbz;
  int *origin;
  T *coef;
  int *bin;
  int *bin_count;
  int *slot;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T>
__global__ void k_prepare_multi(const A_prepare<T> *__restrict__ a) {
  const A_prepare<T> &A = a[blockIdx.z];
  d_prepare<T, const GaussConst &>(A.act, A.np, A.Y, A.F, A.Tq, A.g, A.tx, A.ty, A.tz, A.nbx, A.nby, A.nbz, A.origin, A.coef, A.bin, A.bin_count, A.slot);
}

__device__ __forceinline__ void d_fill_int(const int *__restrict__ act, int *p, int n, int v) {
  if (act && !*act) return;  // batched sweep: suspension converged
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_int(const int *__restrict__ act, int *p, int n, int v) {
  d_fill_int(act, p, n, v);
}
// arguments of one suspension for k_fill_int_multi
struct A_fill_int {
  const int *act;
  int *p;
  int n;
  int v;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
__global__ void k_fill_int_multi(const A_fill_int *__restrict__ a) {
  const A_fill_int &A = a[blockIdx.z];
  d_fill_int(A.act, A.p, A.n, A.v);
}

__device__ __forceinline__ void d_scatter(int np, const int *__restrict__ bin,
                          const int *__restrict__ slot,
                          const int *__restrict__ start, int *__restrict__ ids) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  ids[start[bin[s]] + slot[s]] = s;
}
__global__ void k_scatter(int np, const int *__restrict__ bin,
                          const int *__restrict__ slot,
                          const int *__restrict__ start, int *__restrict__ ids) {
  d_scatter(np, bin, slot, start, ids);
}
// arguments of one suspension for k_scatter_multi
struct A_scatter {
  int np;
  const int *bin;
  const int *slot;
  const int *start;
  int *ids;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
__global__ void k_scatter_multi(const A_scatter *__restrict__ a) {
  const A_scatter &A = a[blockIdx.z];
  d_scatter(A.np, A.bin, A.slot, A.start, A.ids);
}

// Ascending segment order inside each bin (atomic slots are unordered);
// srec and sorg (the matrix-core path; may be null) receive each segment's
// record and packed stencil origin in the same order.
template <typename T>
__device__ __forceinline__ void d_sort_bins(const int *__restrict__ act, int nbins,
                            const int *__restrict__ start,
                            int *__restrict__ ids,
                            const int *__restrict__ origin,
                            const double *__restrict__ Y,
                            const double *__restrict__ F,
                            const double *__restrict__ Tq, double dx,
                            int ngdh, SegRec<T> *__restrict__ srec,
                            int *__restrict__ sorg) {
  if (act && !*act) return;  // batched sweep: suspension converged
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
  if (srec)
    for (int i = lo; i < hi; ++i) {
      const int sid = ids[i];
      SegRec<T> r;
      for (int d = 0; d < 3; ++d) {
        const double y = Y[3 * sid + d];
        r.xx0[d] = ((double)(stencil_cell(y, dx, sid) - ngdh + 1)) * dx - y;
        r.cf[d] = (T)F[3 * sid + d];
        r.cf[3 + d] = (T)(0.5 * Tq[3 * sid + d]);
      }
      r.id = sid;
      r.org = origin[3 * sid] | (origin[3 * sid + 1] << 10) |
              (origin[3 * sid + 2] << 20);
      srec[i] = r;
      sorg[i] = r.org;
    }
}
template <typename T>
__global__ void k_sort_bins(const int *__restrict__ act, int nbins,
                            const int *__restrict__ start,
                            int *__restrict__ ids,
                            const int *__restrict__ origin,
                            const double *__restrict__ Y,
                            const double *__restrict__ F,
                            const double *__restrict__ Tq, double dx,
                            int ngdh, SegRec<T> *__restrict__ srec,
                            int *__restrict__ sorg) {
  d_sort_bins<T>(act, nbins, start, ids, origin, Y, F, Tq, dx, ngdh, srec, sorg);
}
// arguments of one suspension for k_sort_bins_multi
template <typename T>
struct A_sort_bins {
  const int *act;
  int nbins;
  const int *start;
  int *ids;
  const int *origin;
  const double *Y;
  const double *F;
  const double *Tq;
  double dx;
  int ngdh;
  SegRec<T> *srec;
  int *sorg;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T>
__global__ void k_sort_bins_multi(const A_sort_bins<T> *__restrict__ a) {
  const A_sort_bins<T> &A = a[blockIdx.z];
  d_sort_bins<T>(A.act, A.nbins, A.start, A.ids, A.origin, A.Y, A.F, A.Tq, A.dx, A.ngdh, A.srec, A.sorg);
}

// ---------------------------------------------------------------------
// spread (gather): one workgroup per grid tile of tx * ty * TZ points, one
// thread per (x, y) column owning TZ consecutive z points. Candidates (the
// segments whose stencil origin lies in the bins that can reach the tile)
// are staged in LDS in a fixed order (bins, then ascending segment id) with
// their z overlap precomputed; each thread tests x and y once per
// candidate and accumulates the z run in registers. Every grid point is
// written once: no atomics, no zeroing pass, deterministic.
// ---------------------------------------------------------------------
template <typename T, int TZ>
__global__ void __launch_bounds__(256)
    k_spread(const int *__restrict__ act, SpreadGeom G, const T *__restrict__ tab,
             const int *__restrict__ origin, const T *__restrict__ coef,
             const int *__restrict__ start, const int *__restrict__ ids,
             T *__restrict__ R) {
  if (act && !*act) return;  // batched sweep: suspension converged
  __shared__ int s_bstart[kMaxBinsPerTile + 1];
  __shared__ int s_bfirst[kMaxBinsPerTile];
  __shared__ int s_ntot;
  // candidate tables during the sum, the output tile during the store
  constexpr int kTabSz = kCandChunk * 9 * kMaxNgd;
  constexpr int kBufSz = kTabSz > 256 * TZ ? kTabSz : 256 * TZ;
  __shared__ T s_buf[kBufSz];
  __shared__ int s_org[kCandChunk][2];
  __shared__ int s_z[kCandChunk][3];  // zlo, zhi, stencil z index of z0
  __shared__ T s_coef[kCandChunk][6];
  __shared__ int s_cid[kCandChunk];

  const int ntiles = G.nbx * G.nby * G.nbz;
  // XCD-aware order: consecutive tiles land on the same XCD's L2.
  int bid = blockIdx.x;
  if ((ntiles & 7) == 0) bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
  const int bx = bid % G.nbx;
  const int by = (bid / G.nbx) % G.nby;
  const int bz = bid / (G.nbx * G.nby);
  const int ngd = G.ngd;

  const int cx = min(G.hx + 1, G.nbx), cy = min(G.hy + 1, G.nby),
            cz = min(G.hz + 1, G.nbz);
  const int nbins = cx * cy * cz;
  // the workgroup may have fewer threads (tx*ty) than bins: stride
  for (int t = threadIdx.x; t < nbins; t += blockDim.x) {
    int dz = t / (cx * cy), dy = (t / cx) % cy, dx = t % cx;
    int ix = ((bx - G.hx + dx) % G.nbx + G.nbx) % G.nbx;
    int iy = ((by - G.hy + dy) % G.nby + G.nby) % G.nby;
    int iz = ((bz - G.hz + dz) % G.nbz + G.nbz) % G.nbz;
    int b = spread_bin(ix, iy, iz, G.nbx, G.nby, G.nbz);
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
  const int lx = threadIdx.x / G.ty, ly = threadIdx.x % G.ty;
  const bool active = lx < G.tx;
  const int px = bx * G.tx + lx, py = by * G.ty + ly, pz0 = bz * TZ;
  T acc[TZ][3];
#pragma unroll
  for (int k = 0; k < TZ; ++k) acc[k][0] = acc[k][1] = acc[k][2] = (T)0;

  for (int c0 = 0; c0 < ntot; c0 += kCandChunk) {
    const int nc = min(kCandChunk, ntot - c0);
    // candidate ids, origins, z overlap and coefficients: one thread each
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const int gc = c0 + c;
      int t = 0;
      while (gc >= s_bstart[t + 1]) ++t;
      const int s = ids[s_bfirst[t] + gc - s_bstart[t]];
      s_cid[c] = s;
      s_org[c][0] = origin[3 * s + 0];
      s_org[c][1] = origin[3 * s + 1];
      for (int r = 0; r < 6; ++r) s_coef[c][r] = coef[kCoef * s + r];
      // z overlap of the stencil [oz, oz+ngd) with [pz0, pz0+TZ) (mod nz;
      // nz >= TZ + ngd so at most one interval)
      const int oz = origin[3 * s + 2];
      int o = pz0 - oz;
      if (o < 0) o += G.nz;
      int zlo = 0, zhi = 0, izb = 0;
      if (o < ngd) {
        zhi = min(TZ, ngd - o);
        izb = o;
      } else if (o + TZ > G.nz) {
        zlo = G.nz - o;
        zhi = min(TZ, G.nz - o + ngd);
        izb = o - G.nz;
      }
      s_z[c][0] = zlo;
      s_z[c][1] = zhi;
      s_z[c][2] = izb;
    }
    __syncthreads();
    // tables: flat copy with independent loads in flight. e / tl via a
    // float reciprocal is exact here: (e + 0.5) / tl stays >= 0.5 / tl from
    // an integer and e < 2^13.
    {
      const int tl = 9 * ngd, ne = nc * tl;
      const float inv_tl = 1.0f / (float)tl;
#pragma unroll 4
      for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const int c = (int)(((float)e + 0.5f) * inv_tl);
        const int r = e - c * tl;
        s_buf[c * (9 * kMaxNgd) + r] = tab[(size_t)s_cid[c] * tl + r];
      }
    }
    __syncthreads();
    if (active) {
      for (int c = 0; c < nc; ++c) {
        int ix = px - s_org[c][0];
        if (ix < 0) ix += G.nx;
        int iy = py - s_org[c][1];
        if (iy < 0) iy += G.ny;
        if (ix >= ngd || iy >= ngd) continue;
        const int zlo = s_z[c][0], zhi = s_z[c][1], izb = s_z[c][2];
        const T *tb = s_buf + c * (9 * kMaxNgd);
        const T gxy = tb[0 * ngd + ix] * tb[1 * ngd + iy];
        const T dxy = tb[3 * ngd + ix] * tb[4 * ngd + iy];
        const T xx = tb[6 * ngd + ix], yy = tb[7 * ngd + iy];
        const T *cf = s_coef[c];
        // constant parts of the dipole terms (g = T/2 pattern, GA_setup)
        const T ax = cf[5] * yy, ay = -cf[5] * xx, az = cf[4] * xx - cf[3] * yy;
#pragma unroll
        for (int k = 0; k < TZ; ++k) {
          if (k >= zlo && k < zhi) {
            const int iz = izb + k;
            const T temp = gxy * tb[2 * ngd + iz];
            const T temp2 = dxy * tb[5 * ngd + iz];
            const T zz = tb[8 * ngd + iz];
            acc[k][0] += cf[0] * temp + (ax - cf[4] * zz) * temp2;
            acc[k][1] += cf[1] * temp + (cf[3] * zz + ay) * temp2;
            acc[k][2] += cf[2] * temp + az * temp2;
          }
        }
      }
    }
    __syncthreads();
  }
  // stage each component's tile in LDS (the table buffer is free now) and
  // write it with consecutive threads on consecutive z: runs of TZ values
  // instead of one address per lane per store
  const size_t N = (size_t)G.nx * G.ny * G.nz;
  T *stage = s_buf;
  const int nrow = G.tx * G.ty;
  for (int c = 0; c < 3; ++c) {
    if (active) {
#pragma unroll
      for (int k = 0; k < TZ; ++k) stage[threadIdx.x * TZ + k] = acc[k][c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nrow * TZ; e += blockDim.x) {
      const int row = e / TZ, k = e - row * TZ;
      const int rx = bx * G.tx + row / G.ty, ry = by * G.ty + row % G.ty;
      R[(size_t)c * N + ((size_t)rx * G.ny + ry) * G.nz + pz0 + k] = stage[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------
// Matrix-core spread and interpolation work on 8 x 8 x 16 grid tiles.
// Per solve, k_tile_lists gives every tile the records of the segments
// whose stencil meets it; both kernels then load a tile's records in one
// round trip and form the Gaussian weights of those segments at the tile's
// own 8 + 8 + 16 axis points in LDS (no per-segment tables).
// ---------------------------------------------------------------------

// Whether a stencil of ngd points starting at o overlaps the te points
// starting at p0 on a periodic axis of n points (o, p0 in [0, n)).
__device__ __forceinline__ bool stencil_meets(int o, int p0, int te, int ngd,
                                             int n) {
  int d = o - p0;
  if (d < 0) d += n;
  return d < te || d > n - ngd;
}

// Block id -> tile (bx, by, bz): consecutive block ids on one XCD (blocks
// are dealt to the 8 XCDs round-robin) take a contiguous run of tiles, and
// z-neighbouring tiles, which share 128-byte lines of the grid, are next to
// each other in it when nbz is even.
__device__ __forceinline__ void tile_of_block(const SpreadGeom &G, int bid,
                                              int *bx, int *by, int *bz) {
  const int ntiles = G.nbx * G.nby * G.nbz;
  if ((ntiles & 7) == 0) bid = (bid & 7) * (ntiles >> 3) + (bid >> 3);
  if ((G.nbz & 1) == 0) {
    const int zl = bid & 1, r = bid >> 1;
    *bx = r % G.nbx;
    *by = (r / G.nbx) % G.nby;
    *bz = 2 * (r / (G.nbx * G.nby)) + zl;
  } else {
    *bx = bid % G.nbx;
    *by = (bid / G.nbx) % G.nby;
    *bz = bid / (G.nbx * G.nby);
  }
}

// One wavefront per tile (tile = 4 * block id + wave, x fastest): the
// segments of the bins that can reach the tile whose stencil meets it, in
// (neighbouring bin, position in bin) order; the test reads the packed
// origins (sorg), only kept segments' records are copied. The first
// kTileCap records go to the tile's slab, the rest to a block of the pool
// reserved with one atomic per tile (where the block lies varies between
// runs, the order inside it does not); tcnt = how many, tovf = the pool
// block. The pool holds np times the tiles one stencil can meet, a bound on
// the records of all tiles.
template <typename T, typename SG = SpreadGeom>
__device__ __forceinline__ void d_tile_lists(const int *__restrict__ act, SG G,
                 const int *__restrict__ start, const int *__restrict__ sorg,
                 const SegRec<T> *__restrict__ srec, SegRec<T> *__restrict__ slab,
                 int *__restrict__ tcnt, int *__restrict__ tovf,
                 SegRec<T> *__restrict__ pool, int *__restrict__ pool_top) {
  if (act && !*act) return;  // batched sweep: suspension converged
  __shared__ int s_pre[4][64], s_first[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntiles = G.nbx * G.nby * G.nbz;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= ntiles) return;  // whole wave; no workgroup barrier below
  int *pre = s_pre[wave], *first = s_first[wave];
  const int bx = tile % G.nbx, by = (tile / G.nbx) % G.nby,
            bz = tile / (G.nbx * G.nby);
  const int cx = min(G.hx + 1, G.nbx), cy = min(G.hy + 1, G.nby),
            cz = min(G.hz + 1, G.nbz);
  const int nbins = cx * cy * cz;  // <= 64
  int n = 0, f = 0;
  if (lane < nbins) {
    const int dz = lane / (cx * cy), dy = (lane / cx) % cy, dx = lane % cx;
    const int ix = ((bx - G.hx + dx) % G.nbx + G.nbx) % G.nbx;
    const int iy = ((by - G.hy + dy) % G.nby + G.nby) % G.nby;
    const int iz = ((bz - G.hz + dz) % G.nbz + G.nbz) % G.nbz;
    const int b = spread_bin(ix, iy, iz, G.nbx, G.nby, G.nbz);
    f = start[b];
    n = start[b + 1] - f;
  }
  int incl = n;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  const int total = __shfl(incl, 63, 64);
  pre[lane] = incl - n;
  first[lane] = f;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int px0 = bx * 8, py0 = by * 8, pz0 = bz * 16;
  int kept = 0, base = 0;
  // pass 0 counts and fills the slab; pass 1 (a tile with more than
  // kTileCap records) fills its pool block
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      if (kept <= kTileCap) break;
      int b0 = 0;
      if (lane == 0) b0 = atomicAdd(pool_top, kept - kTileCap);
      base = __shfl(b0, 0, 64);
      kept = 0;
    }
    for (int j0 = 0; j0 < total; j0 += 64) {
      const int j = j0 + lane;
      bool keep = false;
      int pos = 0;
      if (j < total) {
        // the last bin whose first list position is <= j holds entry j
        int lo = 0, hi = nbins - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (pre[mid] <= j)
            lo = mid;
          else
            hi = mid - 1;
        }
        pos = first[lo] + j - pre[lo];
        const int org = sorg[pos];
        keep = stencil_meets(org & 1023, px0, 8, G.ngd, G.nx) &&
               stencil_meets((org >> 10) & 1023, py0, 8, G.ngd, G.ny) &&
               stencil_meets(org >> 20, pz0, 16, G.ngd, G.nz);
      }
      const unsigned long long kb = __ballot(keep);
      const int k = kept + __popcll(kb & ((1ull << lane) - 1ull));
      if (keep) {
        if (pass == 0 && k < kTileCap)
          slab[(size_t)tile * kTileCap + k] = srec[pos];
        else if (pass == 1 && k >= kTileCap)
          pool[base + k - kTileCap] = srec[pos];
      }
      kept += __popcll(kb);
    }
  }
  if (lane == 0) {
    tcnt[tile] = kept;
    tovf[tile] = base;
  }
}
template <typename T>
__global__ void __launch_bounds__(256)
    k_tile_lists(const int *__restrict__ act, SpreadGeom G,
                 const int *__restrict__ start, const int *__restrict__ sorg,
                 const SegRec<T> *__restrict__ srec, SegRec<T> *__restrict__ slab,
                 int *__restrict__ tcnt, int *__restrict__ tovf,
                 SegRec<T> *__restrict__ pool, int *__restrict__ pool_top) {
  d_tile_lists<T>(act, G, start, sorg, srec, slab, tcnt, tovf, pool, pool_top);
}
// arguments of one suspension for k_tile_lists_multi
template <typename T>
struct A_tile_lists {
  const int *act;
  SpreadGeom G;
  const int *start;
  const int *sorg;
  const SegRec<T> *srec;
  SegRec<T> *slab;
  int *tcnt;
  int *tovf;
  SegRec<T> *pool;
  int *pool_top;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T>
__global__ void __launch_bounds__(256)
    k_tile_lists_multi(const A_tile_lists<T> *__restrict__ a) {
  const A_tile_lists<T> &A = a[blockIdx.z];
  d_tile_lists<T, const SpreadGeom &>(A.act, A.G, A.start, A.sorg, A.srec, A.slab, A.tcnt, A.tovf, A.pool, A.pool_top);
}

// Stages records [r0, r0 + n) of tile `tile` into s_rec: r0 = 0 copies the
// whole slab (issued before the count is known; entries past the count are
// never used), r0 >= kTileCap the tile's pool block. 256 threads.
template <typename T>
__device__ __forceinline__ void stage_tile_records(
    const SegRec<T> *__restrict__ slab, const SegRec<T> *__restrict__ pool,
    int tile, int ovf, int r0, int n, SegRec<T> *s_rec) {
  constexpr int W = sizeof(SegRec<T>) / 16;
  static_assert(kTileCap * W <= 256, "one 16-byte word per thread");
  const int4 *src =
      r0 == 0 ? reinterpret_cast<const int4 *>(slab + (size_t)tile * kTileCap)
              : reinterpret_cast<const int4 *>(pool + ovf + r0 - kTileCap);
  const int nw = (r0 == 0 ? kTileCap : n) * W;
  if ((int)threadIdx.x < nw)
    reinterpret_cast<int4 *>(s_rec)[threadIdx.x] = src[threadIdx.x];
}

// A tile's record count and this thread's 16-byte word of its slab, loaded
// ahead of use: the persistent tile kernels fetch the next tile's while
// they work on the current one.
template <typename T>
struct TilePrefetch {
  static constexpr int W = sizeof(SegRec<T>) / 16;
  int4 w;
  int cnt;
  __device__ __forceinline__ void load(const SegRec<T> *__restrict__ slab,
                                       const int *__restrict__ tcnt, int tile) {
    if ((int)threadIdx.x < kTileCap * W)
      w = reinterpret_cast<const int4 *>(slab + (size_t)tile * kTileCap)[threadIdx.x];
    cnt = tcnt[tile];
  }
  __device__ __forceinline__ void store(SegRec<T> *s_rec) const {
    if ((int)threadIdx.x < kTileCap * W)
      reinterpret_cast<int4 *>(s_rec)[threadIdx.x] = w;
  }
};

// Gaussian weights of one segment at a tile's axis points. Point a: x
// (a < 8), y (8 <= a < 16), z (16 <= a < 32) of the tile starting at
// (px0, py0, pz0). w[a] = g (monopole), w[32 + a] = gD (dipole), w[64 + a]
// = the gradient factor -xx / sigma_dip^2, zero off the stencil; xx is
// formed in FP64. FP64 uses the expressions of k_prepare_tab; FP32 rounds
// the exponent, formed in FP64 in base 2, before the hardware exp2.
constexpr int kTileW = 104;  // weights (96) + coefficients (8) per record
template <typename T>
__device__ __forceinline__ void tile_weights(const GaussConst &g,
                                             const SegRec<T> &r, int a, int px0,
                                             int py0, int pz0, T *w) {
  const int d = a < 8 ? 0 : (a < 16 ? 1 : 2);
  const int p = d == 0 ? px0 + a : (d == 1 ? py0 + a - 8 : pz0 + a - 16);
  const int n = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
  int i = p - ((r.org >> (10 * d)) & 1023);
  if (i < 0) i += n;
  T ga = 0, gd = 0, gr = 0;
  if (i < g.ngd) {
    const double x0 = d == 0 ? r.xx0[0] : (d == 1 ? r.xx0[1] : r.xx0[2]);
    const double xx = x0 + (double)i * g.dx;
    if constexpr (sizeof(T) == 8) {
      ga = g.anorm * exp(-xx * xx / g.anorm2);
      gd = g.anormdip * exp(-xx * xx / g.anormdip2);
      gr = -xx / g.sigmadipsq;
    } else {
      const double x2 = xx * xx;
      ga = (float)g.anorm * __builtin_amdgcn_exp2f((float)(-x2 * g.e2));
      gd = (float)g.anormdip * __builtin_amdgcn_exp2f((float)(-x2 * g.ed2));
      gr = (float)(xx * g.nrsd);
    }
  }
  w[a] = ga;
  w[32 + a] = gd;
  w[64 + a] = gr;
}

// Weights (and, with CF, coefficients) of records s_rec[c0 .. c0 + nc) into
// s_w[0 .. KC); rows nc .. KC are zero. 256 threads.
template <typename T, int KC, bool CF>
__device__ __forceinline__ void tile_chunk_weights(const GaussConst &g,
                                                   const SegRec<T> *s_rec,
                                                   int c0, int nc, int px0,
                                                   int py0, int pz0,
                                                   T (*s_w)[kTileW]) {
#pragma unroll
  for (int e = threadIdx.x; e < KC * 32; e += 256) {
    const int c = e >> 5, a = e & 31;
    if (c < nc) {
      tile_weights<T>(g, s_rec[c0 + c], a, px0, py0, pz0, s_w[c]);
    } else {
      s_w[c][a] = 0;
      s_w[c][32 + a] = 0;
      s_w[c][64 + a] = 0;
    }
  }
  if constexpr (CF) {
    for (int e = threadIdx.x; e < KC * 8; e += 256) {
      const int c = e >> 3, k = e & 7;
      s_w[c][96 + k] = (c < nc && k < 6) ? s_rec[c0 + c].cf[k] : (T)0;
    }
  }
}

// ---------------------------------------------------------------------
// spread on the matrix cores: one workgroup per 8 x 8 x 16 tile,
// four wavefronts of 16 (x, y) points each. Per component the tile is
//   R[xy][z] = sum over candidates c and terms t of A_t[xy][c] * B_t[c][z]
// with the force term   A_0 = F * gx gy,            B_0 = gz,
// and the torque terms  A_1 = (ax, ay, az) gdx gdy,  B_1 = gdz,
//                       A_2 = (-G_y, G_x, 0) gdx gdy, B_2 = gdz * grad_z
// (ax = G_z grad_y, ay = -G_z grad_x, az = G_y grad_x - G_x grad_y, G the
// half torque: the same sum as k_spread). Each MFMA 16x16x4 step takes four
// candidates of one term; lane l supplies A[xy = l % 16][c = l / 16] and
// B[c = l / 16][z = l % 16] and holds R[xy = Mfma<T>::row(l, r)][z = l % 16].
// Weights are zero off a candidate's stencil and for the unused rows of a
// round. Candidates are summed in tile-list order: no atomics, no zeroing
// pass, every grid point written once.
// ---------------------------------------------------------------------
// Candidates per round (FP32 32, FP64 16: the same 13 KB of weights).
template <typename T>
constexpr int kMfmaChunk = sizeof(T) == 8 ? 16 : 32;

// One 16x16x4 matrix-core step D += A B in the real type, and the row of
// D held in accumulator element r of lane l (column l % 16): 4 (l / 16) + r
// for FP32, 4 r + l / 16 for FP64 (the instructions' output layouts).
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) {
    return 4 * (lane >> 4) + r;
  }
};
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) {
    return 4 * r + (lane >> 4);
  }
};

template <typename T, typename SG = SpreadGeom, typename GC = GaussConst>
__device__ __forceinline__ void d_spread_mfma(const int *__restrict__ act, SG G, GC g,
                  const SegRec<T> *__restrict__ slab,
                  const SegRec<T> *__restrict__ pool,
                  const int *__restrict__ tcnt, const int *__restrict__ tovf,
                  T *__restrict__ R) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int KC = kMfmaChunk<T>;
  if (act && !*act) return;  // batched sweep: suspension converged
  __shared__ SegRec<T> s_rec[kTileCap];
  __shared__ T s_w[KC + 1][kTileW];  // row KC stays zero

  int bx, by, bz;
  tile_of_block(G, blockIdx.x, &bx, &by, &bz);
  const int tile = bx + G.nbx * (by + G.nby * bz);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, i16 = lane & 15;
  const int xy = wave * 16 + i16;  // this lane's A row
  const int lx = xy >> 3, ly = xy & 7, lz = i16;
  const int px0 = bx * 8, py0 = by * 8, pz0 = bz * 16;
  // the slab is copied while the count is still on its way
  stage_tile_records<T>(slab, pool, tile, 0, 0, 0, s_rec);
  for (int e = threadIdx.x; e < kTileW; e += 256) s_w[KC][e] = 0;
  const int cnt = tcnt[tile];
  const int ovf = cnt > kTileCap ? tovf[tile] : 0;
  v4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0}, acc2 = {0, 0, 0, 0};

  for (int r0 = 0; r0 < cnt; r0 += kTileCap) {
    const int nb = min(kTileCap, cnt - r0);
    if (r0 > 0) {
      __syncthreads();  // the last round still reads s_rec
      stage_tile_records<T>(slab, pool, tile, ovf, r0, nb, s_rec);
    }
    __syncthreads();
    for (int c0 = 0; c0 < nb; c0 += KC) {
      const int nc = min(KC, nb - c0);
      tile_chunk_weights<T, KC, true>(g, s_rec, c0, nc, px0, py0, pz0, s_w);
      __syncthreads();
      // this wave's rows lie in x columns 2 wave and 2 wave + 1: candidates
      // with no stencil point there (zero monopole and dipole weights in
      // both) add exact zeros to them and are skipped
      const T *wl = s_w[lane < nc ? lane : KC];
      unsigned long long m = __ballot(
          lane < nc && (wl[2 * wave] != (T)0 || wl[2 * wave + 1] != (T)0 ||
                        wl[32 + 2 * wave] != (T)0 || wl[33 + 2 * wave] != (T)0));
      const int nw = __popcll(m);
      for (int c4 = 0; c4 < nw; c4 += 4) {
        // lane group q takes the (c4 + q)-th candidate of the wave's list
        // (the zero row past its end)
        int c = KC;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int ck = m ? (int)__builtin_ctzll(m) : KC;
          if (k == q) c = ck;
          m &= m - 1;
        }
        const T *w = s_w[c];
        const T *cf = w + 96;
        const T gxy = w[lx] * w[8 + ly], dxy = w[32 + lx] * w[40 + ly];
        const T xx = w[64 + lx], yy = w[72 + ly];
        const T gz = w[16 + lz], gdz = w[48 + lz], grz = w[80 + lz];
        const T ax = cf[5] * yy, ay = -cf[5] * xx, az = cf[4] * xx - cf[3] * yy;
        const T b2 = gdz * grz;
        acc0 = Mfma<T>::step(cf[0] * gxy, gz, acc0);
        acc1 = Mfma<T>::step(cf[1] * gxy, gz, acc1);
        acc2 = Mfma<T>::step(cf[2] * gxy, gz, acc2);
        acc0 = Mfma<T>::step(ax * dxy, gdz, acc0);
        acc1 = Mfma<T>::step(ay * dxy, gdz, acc1);
        acc2 = Mfma<T>::step(az * dxy, gdz, acc2);
        acc0 = Mfma<T>::step(-cf[4] * dxy, b2, acc0);
        acc1 = Mfma<T>::step(cf[3] * dxy, b2, acc1);
      }
      __syncthreads();  // the next round overwrites the weights
    }
  }
  // lane holds R[xy = wave*16 + Mfma<T>::row(lane, r)][z = i16]: 16 lanes
  // write 16 consecutive z of one (x, y) column
  const size_t N = (size_t)G.nx * G.ny * G.nz;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int oxy = wave * 16 + Mfma<T>::row(lane, r);
    const size_t o = ((size_t)(px0 + (oxy >> 3)) * G.ny + py0 + (oxy & 7)) *
                         G.nz + pz0 + i16;
    R[o] = acc0[r];
    R[N + o] = acc1[r];
    R[2 * N + o] = acc2[r];
  }
}
template <typename T>
__global__ void __launch_bounds__(256)
    k_spread_mfma(const int *__restrict__ act, SpreadGeom G, GaussConst g,
                  const SegRec<T> *__restrict__ slab,
                  const SegRec<T> *__restrict__ pool,
                  const int *__restrict__ tcnt, const int *__restrict__ tovf,
                  T *__restrict__ R) {
  d_spread_mfma<T>(act, G, g, slab, pool, tcnt, tovf, R);
}
// arguments of one suspension for k_spread_mfma_multi
template <typename T>
struct A_spread_mfma {
  const int *act;
  SpreadGeom G;
  GaussConst g;
  const SegRec<T> *slab;
  const SegRec<T> *pool;
  const int *tcnt;
  const int *tovf;
  T *R;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T>
__global__ void __launch_bounds__(256)
    k_spread_mfma_multi(const A_spread_mfma<T> *__restrict__ a) {
  const A_spread_mfma<T> &A = a[blockIdx.z];
  d_spread_mfma<T, const SpreadGeom &, const GaussConst &>(A.act, A.G, A.g, A.slab, A.pool, A.tcnt, A.tovf, A.R);
}
// ---------------------------------------------------------------------
// FFT passes. Half spectrum layout C[c][kz][y][x] (x fastest): the z pass
// writes it transposed so the x pass reads contiguous lines and the y pass
// reads runs of consecutive x.
// ---------------------------------------------------------------------

// Kernels of the FFT passes are instantiated on LOGN: n = 2^LOGN known at
// compile time (LOGN > 0) or taken from the plan (LOGN = 0).
template <int LOGN>
__device__ __forceinline__ int fft_len(const FftPlan &pl) {
  if constexpr (LOGN > 0)
    return 1 << LOGN;
  else
    return pl.n;
}
template <int LOGN>
__device__ __forceinline__ int div_len(int e, const FftPlan &pl) {  // e / n
  if constexpr (LOGN > 0)
    return e >> LOGN;
  else
    return pl.dn.q(e);
}

// 16-byte global accesses of reals (float4 / double2); p must be 16-byte
// aligned.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4, logn = 2;
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2, logn = 1;
};
template <typename T>
__device__ __forceinline__ void ld16(const T *p, T *out) {
  using V = typename Vec16<T>::type;
  const V v = *reinterpret_cast<const V *>(p);
  __builtin_memcpy(out, &v, 16);
}
// Non-temporal 16-byte load, for data read once that the solve does not
// read again before overwriting it (the z passes' inputs): the hint keeps it
// from displacing the grid the next pass reads from the Infinity Cache
// (measured at 256^3 FP32: spread, z passes and forward y pass 0.039 ms
// faster per solve; the interpolation, which shares 128-byte lines between
// neighbouring tiles through L2, loads normally).
typedef int i32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ void ld16_nt(const T *p, T *out) {
  const i32x4 v = __builtin_nontemporal_load(reinterpret_cast<const i32x4 *>(p));
  __builtin_memcpy(out, &v, 16);
}
template <typename T>
__device__ __forceinline__ void st16(T *p, const T *in) {
  using V = typename Vec16<T>::type;
  V v;
  __builtin_memcpy(&v, in, 16);
  *reinterpret_cast<V *>(p) = v;
}

// Complex values per 16-byte access: 2 (FP32) or 1 (FP64).
template <typename T>
constexpr int kCx16 = 16 / (int)sizeof(cx<T>);
template <typename T>
__device__ __forceinline__ void ld16c(const cx<T> *p, cx<T> *out) {
  ld16(reinterpret_cast<const T *>(p), reinterpret_cast<T *>(out));
}
template <typename T>
__device__ __forceinline__ void st16c(cx<T> *p, const cx<T> *in) {
  st16(reinterpret_cast<T *>(p), reinterpret_cast<const T *>(in));
}

// z forward: XB = 2P consecutive x rows at one y, as P complex sequences
// (row 2p in the real part, row 2p+1 in the imaginary part).
template <typename T, int E, int LOGN, typename PL = FftPlan>
__device__ __forceinline__ void d_fft_z_r2c(const int *__restrict__ act, const T *__restrict__ R, cx<T> *__restrict__ C, PL pl,
                const cx<T> *__restrict__ twg, int nx, int ny, int P,
                FastDiv dXB, FastDiv dP) {
  if (act && !*act) return;  // batched sweep: suspension converged
  extern __shared__ unsigned char smem_raw[];
  const int n = fft_len<LOGN>(pl), nh = n / 2 + 1, XB = 2 * P;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *a = tw + n;
  const int comp = blockIdx.y;
  const int nxb = nx / XB;
  const int x0 = (blockIdx.x % nxb) * XB, y = blockIdx.x / nxb;
  const int pn = fft_padded_len(n);
  const T *Rc = R + (size_t)comp * nx * ny * n;
  cx<T> *Cc = C + (size_t)comp * nh * ny * nx;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  const T half = (T)0.5;
  if constexpr (LOGN > 0) {
    // 16-byte loads of both rows of a sequence (VW consecutive z each);
    // z is a multiple of VW <= 4, so the VW values share one pad group
    constexpr int VW = Vec16<T>::n, LV = LOGN - Vec16<T>::logn;
    for (int e = threadIdx.x; e < (P << LV); e += blockDim.x) {
      const int p = e >> LV, z = (e & ((1 << LV) - 1)) * VW;
      const T *r0 = Rc + ((size_t)(x0 + 2 * p) * ny + y) * n + z;
      T u0[VW], u1[VW];
      ld16_nt(r0, u0);
      ld16_nt(r0 + (size_t)ny * n, u1);
      cx<T> *dst = a + p * pn + fft_pad(z);
#pragma unroll
      for (int j = 0; j < VW; ++j) dst[j] = cx<T>{u0[j], u1[j]};
    }
    __syncthreads();
    lds_fft<T, E, true, false, LOGN>(a, pl, P, 1, pn, tw);
    // both rows of a sequence from the same two spectrum values, stored
    // as one 16-byte pair (FP32) or two (FP64) of consecutive x
    for (int e = threadIdx.x; e < nh * P; e += blockDim.x) {
      const int kz = dP.q(e), p = e - kz * P;
      const cx<T> z = a[p * pn + fft_pad(kz)];
      const cx<T> w = a[p * pn + fft_pad((n - kz) & (n - 1))];
      const T o[4] = {half * (z.x + w.x), half * (z.y - w.y),
                      half * (z.y + w.y), half * (w.x - z.x)};
      T *dst = reinterpret_cast<T *>(Cc + ((size_t)kz * ny + y) * nx + x0 + 2 * p);
      st16(dst, o);
      if constexpr (VW == 2) st16(dst + 2, o + 2);
    }
  } else {
    for (int e = threadIdx.x; e < XB * n; e += blockDim.x) {
      int i = div_len<LOGN>(e, pl), z = e - i * n;
      T v = Rc[((size_t)(x0 + i) * ny + y) * n + z];
      if (i & 1)
        a[(i >> 1) * pn + fft_pad(z)].y = v;
      else
        a[(i >> 1) * pn + fft_pad(z)].x = v;
    }
    __syncthreads();
    lds_fft<T, E, true, false, LOGN>(a, pl, P, 1, pn, tw);
    for (int e = threadIdx.x; e < nh * XB; e += blockDim.x) {
      int kz = dXB.q(e), i = e - kz * XB;
      int p = i >> 1;
      cx<T> z = a[p * pn + fft_pad(kz)];
      cx<T> w = a[p * pn + fft_pad(kz == 0 ? 0 : n - kz)];
      cx<T> o = (i & 1) ? cx<T>{half * (z.y + w.y), half * (w.x - z.x)}
                        : cx<T>{half * (z.x + w.x), half * (z.y - w.y)};
      Cc[((size_t)kz * ny + y) * nx + x0 + i] = o;
    }
  }
}
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_z_r2c(const int *__restrict__ act, const T *__restrict__ R, cx<T> *__restrict__ C, FftPlan pl,
                const cx<T> *__restrict__ twg, int nx, int ny, int P,
                FastDiv dXB, FastDiv dP) {
  d_fft_z_r2c<T, E, LOGN>(act, R, C, pl, twg, nx, ny, P, dXB, dP);
}
// arguments of one suspension for k_fft_z_r2c_multi
template <typename T, int E, int LOGN>
struct A_fft_z_r2c {
  const int *act;
  const T *R;
  cx<T> *C;
  FftPlan pl;
  const cx<T> *twg;
  int nx;
  int ny;
  int P;
  FastDiv dXB;
  FastDiv dP;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_z_r2c_multi(const A_fft_z_r2c<T, E, LOGN> *__restrict__ a) {
  const A_fft_z_r2c<T, E, LOGN> &A = a[blockIdx.z];
  d_fft_z_r2c<T, E, LOGN, const FftPlan &>(A.act, A.R, A.C, A.pl, A.twg, A.nx, A.ny, A.P, A.dXB, A.dP);
}

// z inverse: the mirror of k_fft_z_r2c.
template <typename T, int E, int LOGN, typename PL = FftPlan>
__device__ __forceinline__ void d_fft_z_c2r(const int *__restrict__ act, const cx<T> *__restrict__ C, T *__restrict__ R, PL pl,
                const cx<T> *__restrict__ twg, int nx, int ny, int P,
                FastDiv dP) {
  if (act && !*act) return;  // batched sweep: suspension converged
  extern __shared__ unsigned char smem_raw[];
  const int n = fft_len<LOGN>(pl), nh = n / 2 + 1, XB = 2 * P;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *a = tw + n;
  const int comp = blockIdx.y;
  const int nxb = nx / XB;
  const int x0 = (blockIdx.x % nxb) * XB, y = blockIdx.x / nxb;
  const int pn = fft_padded_len(n);
  T *Rc = R + (size_t)comp * nx * ny * n;
  const cx<T> *Cc = C + (size_t)comp * nh * ny * nx;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  for (int e = threadIdx.x; e < nh * P; e += blockDim.x) {
    int k = dP.q(e), p = e - k * P;
    const cx<T> *src = Cc + ((size_t)k * ny + y) * nx + x0 + 2 * p;
    cx<T> A, B;
    if constexpr (LOGN > 0) {  // x0 + 2p even: the pair is 16-byte aligned
      T u[4];
      ld16_nt(reinterpret_cast<const T *>(src), u);
      if constexpr (Vec16<T>::n == 2) ld16_nt(reinterpret_cast<const T *>(src) + 2, u + 2);
      A = {u[0], u[1]};
      B = {u[2], u[3]};
    } else {
      A = src[0];
      B = src[1];
    }
    if (k == 0 || 2 * k == n) {  // DC / Nyquist of a real line are real
      A.y = 0;
      B.y = 0;
    }
    a[p * pn + fft_pad(k)] = {A.x - B.y, A.y + B.x};
    if (k > 0 && 2 * k < n) a[p * pn + fft_pad(n - k)] = {A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  lds_fft<T, E, true, true, LOGN>(a, pl, P, 1, pn, tw);
  if constexpr (LOGN > 0) {  // 16-byte stores of both rows of a sequence
    constexpr int VW = Vec16<T>::n, LV = LOGN - Vec16<T>::logn;
    for (int e = threadIdx.x; e < (P << LV); e += blockDim.x) {
      const int p = e >> LV, z = (e & ((1 << LV) - 1)) * VW;
      const cx<T> *src = a + p * pn + fft_pad(z);
      T u0[VW], u1[VW];
#pragma unroll
      for (int j = 0; j < VW; ++j) {
        u0[j] = src[j].x;
        u1[j] = src[j].y;
      }
      T *r0 = Rc + ((size_t)(x0 + 2 * p) * ny + y) * n + z;
      st16(r0, u0);
      st16(r0 + (size_t)ny * n, u1);
    }
  } else {
    for (int e = threadIdx.x; e < XB * n; e += blockDim.x) {
      int i = div_len<LOGN>(e, pl), z = e - i * n;
      cx<T> v = a[(i >> 1) * pn + fft_pad(z)];
      Rc[((size_t)(x0 + i) * ny + y) * n + z] = (i & 1) ? v.y : v.x;
    }
  }
}
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_z_c2r(const int *__restrict__ act, const cx<T> *__restrict__ C, T *__restrict__ R, FftPlan pl,
                const cx<T> *__restrict__ twg, int nx, int ny, int P,
                FastDiv dP) {
  d_fft_z_c2r<T, E, LOGN>(act, C, R, pl, twg, nx, ny, P, dP);
}
// arguments of one suspension for k_fft_z_c2r_multi
template <typename T, int E, int LOGN>
struct A_fft_z_c2r {
  const int *act;
  const cx<T> *C;
  T *R;
  FftPlan pl;
  const cx<T> *twg;
  int nx;
  int ny;
  int P;
  FastDiv dP;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_z_c2r_multi(const A_fft_z_c2r<T, E, LOGN> *__restrict__ a) {
  const A_fft_z_c2r<T, E, LOGN> &A = a[blockIdx.z];
  d_fft_z_c2r<T, E, LOGN, const FftPlan &>(A.act, A.C, A.R, A.pl, A.twg, A.nx, A.ny, A.P, A.dP);
}

// y pass: one workgroup per (comp, kz, run of B consecutive x).
template <typename T, int E, int LOGN, typename PL = FftPlan>
__device__ __forceinline__ void d_fft_y(const int *__restrict__ act, cx<T> *__restrict__ C, PL pl, const cx<T> *__restrict__ twg,
            int nx, int nh, int B, int inverse, FastDiv dB) {
  if (act && !*act) return;  // batched sweep: suspension converged
  extern __shared__ unsigned char smem_raw[];
  const int n = fft_len<LOGN>(pl);  // ny
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *a = tw + n;
  const int nxb = nx / B;
  const int x0 = (blockIdx.x % nxb) * B, kz = blockIdx.x / nxb;
  const int comp = blockIdx.y;
  cx<T> *base = C + ((size_t)comp * nh + kz) * n * nx + x0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  // 16-byte accesses of VC consecutive x (x0 and B multiples of VC: the
  // run starts 16-byte aligned)
  constexpr int VC = kCx16<T>;
  const bool vec = LOGN > 0 && B % VC == 0;
  if (vec) {
    for (int e = threadIdx.x * VC; e < B * n; e += blockDim.x * VC) {
      int m = dB.q(e), col = e - m * B;
      ld16c(base + (size_t)m * nx + col, a + fft_rowpad(m) * B + col);
    }
  } else {
    for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
      int m = dB.q(e), col = e - m * B;
      a[fft_rowpad(m) * B + col] = base[(size_t)m * nx + col];
    }
  }
  __syncthreads();
  if (inverse)
    lds_fft<T, E, false, true, LOGN>(a, pl, B, B, 1, tw, dB);
  else
    lds_fft<T, E, false, false, LOGN>(a, pl, B, B, 1, tw, dB);
  if (vec) {
    for (int e = threadIdx.x * VC; e < B * n; e += blockDim.x * VC) {
      int m = dB.q(e), col = e - m * B;
      st16c(base + (size_t)m * nx + col, a + fft_rowpad(m) * B + col);
    }
  } else {
    for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
      int m = dB.q(e), col = e - m * B;
      base[(size_t)m * nx + col] = a[fft_rowpad(m) * B + col];
    }
  }
}
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_y(const int *__restrict__ act, cx<T> *__restrict__ C, FftPlan pl, const cx<T> *__restrict__ twg,
            int nx, int nh, int B, int inverse, FastDiv dB) {
  d_fft_y<T, E, LOGN>(act, C, pl, twg, nx, nh, B, inverse, dB);
}
// arguments of one suspension for k_fft_y_multi
template <typename T, int E, int LOGN>
struct A_fft_y {
  const int *act;
  cx<T> *C;
  FftPlan pl;
  const cx<T> *twg;
  int nx;
  int nh;
  int B;
  int inverse;
  FastDiv dB;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_y_multi(const A_fft_y<T, E, LOGN> *__restrict__ a) {
  const A_fft_y<T, E, LOGN> &A = a[blockIdx.z];
  d_fft_y<T, E, LOGN, const FftPlan &>(A.act, A.C, A.pl, A.twg, A.nx, A.nh, A.B, A.inverse, A.dB);
}

struct ProjConst {
  int nx, ny, nz, nh;
  double kfx, kfy, kfz;  // 2 pi / L per axis
  double scale;          // 1 / (N mu)
};

// x pass + Stokes projection + inverse x pass: one workgroup per (kz, Q
// consecutive y), the three components' contiguous x lines in LDS.
template <typename T, int E, int LOGN, typename PL = FftPlan, typename PC = ProjConst>
__device__ __forceinline__ void d_fft_x_project(const int *__restrict__ act, cx<T> *__restrict__ C, PL pl,
                    const cx<T> *__restrict__ twg, PC pc, int Q,
                    FastDiv dQ) {
  if (act && !*act) return;  // batched sweep: suspension converged
  extern __shared__ unsigned char smem_raw[];
  const int n = fft_len<LOGN>(pl);  // nx
  const int nh = pc.nh;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *a = tw + n;  // [3][Q][padded n]
  const int pn = fft_padded_len(n);
  const int nyb = pc.ny / Q;
  const int y0 = (blockIdx.x % nyb) * Q, kz = blockIdx.x / nyb;
  const size_t comp_stride = (size_t)nh * pc.ny * n;
  cx<T> *base = C + ((size_t)kz * pc.ny + y0) * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
  const int nseq = 3 * Q;
  constexpr int VC = kCx16<T>;  // x even: VC values in one pad group
  if constexpr (LOGN > 0) {
    for (int e = threadIdx.x * VC; e < nseq * n; e += blockDim.x * VC) {
      int sq = e >> LOGN, x = e & (n - 1);
      int c = dQ.q(sq), ql = sq - c * Q;
      ld16c(base + c * comp_stride + (size_t)ql * n + x, a + sq * pn + fft_pad(x));
    }
  } else {
    for (int e = threadIdx.x; e < nseq * n; e += blockDim.x) {
      int sq = div_len<LOGN>(e, pl), x = e - sq * n;
      int c = dQ.q(sq), ql = sq - c * Q;
      a[sq * pn + fft_pad(x)] = base[c * comp_stride + (size_t)ql * n + x];
    }
  }
  __syncthreads();
  lds_fft<T, E, true, false, LOGN>(a, pl, nseq, 1, pn, tw);
  // flow_solve (FCMfunctions.cpp:648-736)
  const T q3 = (T)(kz * pc.kfz);
  const T sc = (T)pc.scale;
  for (int e = threadIdx.x; e < Q * n; e += blockDim.x) {
    int ql = div_len<LOGN>(e, pl), x = e - ql * n;
    int y = y0 + ql;
    T q1 = (T)(((x <= pc.nx / 2) ? x : x - pc.nx) * pc.kfx);
    T q2 = (T)(((y <= pc.ny / 2) ? y : y - pc.ny) * pc.kfy);
    T qq = q2 * q2 + q1 * q1 + (T)1e-18;
    T norm = (T)1 / (qq + q3 * q3);
    const int px = fft_pad(x);
    cx<T> f1 = a[(0 * Q + ql) * pn + px], f2 = a[(1 * Q + ql) * pn + px],
          f3 = a[(2 * Q + ql) * pn + px];
    cx<T> u1, u2, u3;
    T kf = (q1 * f1.x + q2 * f2.x + q3 * f3.x) * norm;
    u1.x = norm * (f1.x - q1 * kf) * sc;
    u2.x = norm * (f2.x - q2 * kf) * sc;
    u3.x = norm * (f3.x - q3 * kf) * sc;
    kf = (q1 * f1.y + q2 * f2.y + q3 * f3.y) * norm;
    u1.y = norm * (f1.y - q1 * kf) * sc;
    u2.y = norm * (f2.y - q2 * kf) * sc;
    u3.y = norm * (f3.y - q3 * kf) * sc;
    if (x == 0 && y == 0 && kz == 0) u1 = u2 = u3 = cx<T>{0, 0};
    a[(0 * Q + ql) * pn + px] = u1;
    a[(1 * Q + ql) * pn + px] = u2;
    a[(2 * Q + ql) * pn + px] = u3;
  }
  __syncthreads();
  lds_fft<T, E, true, true, LOGN>(a, pl, nseq, 1, pn, tw);
  if constexpr (LOGN > 0) {
    for (int e = threadIdx.x * VC; e < nseq * n; e += blockDim.x * VC) {
      int sq = e >> LOGN, x = e & (n - 1);
      int c = dQ.q(sq), ql = sq - c * Q;
      st16c(base + c * comp_stride + (size_t)ql * n + x, a + sq * pn + fft_pad(x));
    }
  } else {
    for (int e = threadIdx.x; e < nseq * n; e += blockDim.x) {
      int sq = div_len<LOGN>(e, pl), x = e - sq * n;
      int c = dQ.q(sq), ql = sq - c * Q;
      base[c * comp_stride + (size_t)ql * n + x] = a[sq * pn + fft_pad(x)];
    }
  }
}
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_x_project(const int *__restrict__ act, cx<T> *__restrict__ C, FftPlan pl,
                    const cx<T> *__restrict__ twg, ProjConst pc, int Q,
                    FastDiv dQ) {
  d_fft_x_project<T, E, LOGN>(act, C, pl, twg, pc, Q, dQ);
}
// arguments of one suspension for k_fft_x_project_multi
template <typename T, int E, int LOGN>
struct A_fft_x_project {
  const int *act;
  cx<T> *C;
  FftPlan pl;
  const cx<T> *twg;
  ProjConst pc;
  int Q;
  FastDiv dQ;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T, int E, int LOGN>
__global__ void __launch_bounds__(1024)
    k_fft_x_project_multi(const A_fft_x_project<T, E, LOGN> *__restrict__ a) {
  const A_fft_x_project<T, E, LOGN> &A = a[blockIdx.z];
  d_fft_x_project<T, E, LOGN, const FftPlan &, const ProjConst &>(A.act, A.C, A.pl, A.twg, A.pc, A.Q, A.dQ);
}

// ---------------------------------------------------------------------
// interpolation: one 64-lane wavefront per segment (segments in bin order
// so neighbours share L2 lines); each lane owns one z of the stencil and a
// share of its (x, y) columns, then the wavefront reduces.
// ---------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T, int NGD>
__global__ void __launch_bounds__(256)
    k_interp(const int *__restrict__ act, int np, SpreadGeom G, double dx3, const T *__restrict__ tab,
             const int *__restrict__ origin, const int *__restrict__ ids,
             const T *__restrict__ R, double *__restrict__ V,
             double *__restrict__ W) {
  if (act && !*act) return;  // batched sweep: suspension converged
  constexpr int NC = NGD * NGD;          // (x, y) columns of a stencil
  constexpr int NP = NC * NGD;           // stencil points
  constexpr int PER = (NP + 63) / 64;    // points per lane
  __shared__ T s_gxy[4][NC], s_dxy[4][NC];
  __shared__ T s_z[4][3 * NGD];          // g_z, gD_z, grad_z
  __shared__ T s_xx[4][NGD], s_yy[4][NGD];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // XCD-aware order: the 8 XCDs take contiguous runs of the bin-sorted
  // segment list, so neighbouring stencils share an L2
  const int nblk = (np + 3) / 4;
  const int q8 = (nblk + 7) / 8;
  const int b = (blockIdx.x & 7) * q8 + (blockIdx.x >> 3);
  const int w = b * 4 + wave;
  const bool valid = (b < nblk) && (w < np);
  const int s = valid ? ids[w] : 0;
  const T *ts = tab + (size_t)s * 9 * NGD;
  if (valid) {
    for (int c = lane; c < NC; c += 64) {
      int i = c / NGD, j = c - (c / NGD) * NGD;
      s_gxy[wave][c] = ts[0 * NGD + i] * ts[1 * NGD + j];
      s_dxy[wave][c] = ts[3 * NGD + i] * ts[4 * NGD + j];
    }
    for (int k = lane; k < NGD; k += 64) {
      s_z[wave][k] = ts[2 * NGD + k];
      s_z[wave][NGD + k] = ts[5 * NGD + k];
      s_z[wave][2 * NGD + k] = ts[8 * NGD + k];
      s_xx[wave][k] = ts[6 * NGD + k];
      s_yy[wave][k] = ts[7 * NGD + k];
    }
  }
  const int ox = valid ? origin[3 * s] : 0, oy = valid ? origin[3 * s + 1] : 0,
            oz = valid ? origin[3 * s + 2] : 0;
  __syncthreads();
  const size_t N = (size_t)G.nx * G.ny * G.nz;
  // NGD lanes per (x, y) column, KL columns per pass: lane (cl, k) keeps one
  // z (its weights and wrapped z are constants) and walks the columns
  // cl, cl + KL, ...; consecutive lanes still read consecutive z. The
  // stencil sums factor per lane into 9 running sums scaled by the lane's
  // z weights at the end:
  //   V   = g_z   sum gx gy u
  //   W_x = gd_z (sum d y u_z - z sum d u_y), d = gdx gdy, x, y, z the
  //   W_y = gd_z (z sum d u_x - sum d x u_z)   dipole gradient factors
  //   W_z = gd_z (sum d x u_y - sum d y u_x)
  constexpr int KL = 64 / NGD;
  constexpr int NPASS = (NC + KL - 1) / KL;
  constexpr int KLI = KL / NGD, KLJ = KL % NGD;
  constexpr int CH = 6;  // passes whose loads are in flight together
  const int cl = lane / NGD, kz = lane - cl * NGD;
  const bool on = valid && cl < KL;
  int zw = oz + (on ? kz : 0);
  if (zw >= G.nz) zw -= G.nz;
  const T gz = on ? s_z[wave][kz] : (T)0, gdz = on ? s_z[wave][NGD + kz] : (T)0,
          zk = on ? s_z[wave][2 * NGD + kz] : (T)0;
  int ci = cl / NGD, cj = cl % NGD;  // the next pass's column (i, j)
  T sg0 = 0, sg1 = 0, sg2 = 0, sd0 = 0, sd1 = 0;
  T sx1 = 0, sx2 = 0, sy0 = 0, sy2 = 0;
  for (int t0 = 0; t0 < NPASS; t0 += CH) {
    T ux[CH], uy[CH], uz[CH];
    int ic[CH], jc[CH];
    bool okc[CH];
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      const bool ok = on && t0 + t < NPASS && ci < NGD;
      int ii = ox + (ok ? ci : 0);
      if (ii >= G.nx) ii -= G.nx;
      int jj = oy + (ok ? cj : 0);
      if (jj >= G.ny) jj -= G.ny;
      const size_t o = ((size_t)ii * G.ny + jj) * G.nz + zw;
      ux[t] = ok ? R[o] : (T)0;
      uy[t] = ok ? R[N + o] : (T)0;
      uz[t] = ok ? R[2 * N + o] : (T)0;
      ic[t] = ok ? ci : 0;
      jc[t] = ok ? cj : 0;
      okc[t] = ok;
      ci += KLI;
      cj += KLJ;
      if (cj >= NGD) {
        cj -= NGD;
        ++ci;
      }
    }
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      const int c = ic[t] * NGD + jc[t];
      const T gxy = okc[t] ? s_gxy[wave][c] : (T)0;
      const T dxy = okc[t] ? s_dxy[wave][c] : (T)0;
      const T tx = dxy * s_xx[wave][ic[t]], ty = dxy * s_yy[wave][jc[t]];
      sg0 += gxy * ux[t];
      sg1 += gxy * uy[t];
      sg2 += gxy * uz[t];
      sd0 += dxy * ux[t];
      sd1 += dxy * uy[t];
      sx1 += tx * uy[t];
      sx2 += tx * uz[t];
      sy0 += ty * ux[t];
      sy2 += ty * uz[t];
    }
  }
  T v0 = gz * sg0, v1 = gz * sg1, v2 = gz * sg2;
  T w0 = gdz * (sy2 - zk * sd1);
  T w1 = gdz * (zk * sd0 - sx2);
  T w2 = gdz * (sx1 - sy0);
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

// ---------------------------------------------------------------------
// interpolation on the matrix cores: the transpose of k_spread_mfma. One
// workgroup per 8 x 8 x 16 tile reads the tile's three velocity grids once
// (each lane four consecutive z of one (x, y) column per component) and,
// for the segments of its tile list (up to 32 per round; each wave takes
// those with a stencil point in its rows, 16 per group), contracts them
// with each segment's weights:
//   P_m[xy][c] = sum_z u_d[xy][z] b_m,c[z]   (per wave 16 xy rows x 16
//   segments, v_mfma 16x16x4 over z; 8 products m: u_0..2 with g_z,
//   u_0..2 with gD_z, u_0..1 with gD_z grad_z)
// then weights each (x, y) row with the segment's x and y factors, sums
// the lane's rows, the four lane groups and the four waves in a fixed
// order, and stores the tile's 6 partial sums per segment in
// part[segment][component][slot] (slot = the tile's offset from the
// segment's origin tile). k_interp_sum adds a segment's slots in slot
// order: no atomics, results independent of scheduling.
// ---------------------------------------------------------------------
// Tiles of te points that a stencil of ngd points can meet along an axis.
__host__ __device__ constexpr int stencil_tiles(int ngd, int te) {
  return (ngd + te - 2) / te + 1;
}
template <int NGD>
constexpr int kInterpSlots = stencil_tiles(NGD, 8) * stencil_tiles(NGD, 8) *
                             stencil_tiles(NGD, 16);

// Four consecutive reals at p (16-byte aligned).
template <typename T>
__device__ __forceinline__ void ld4(const T *p, T *out) {
  ld16(p, out);
  if constexpr (sizeof(T) == 8) ld16(p + 2, out + 2);
}

template <typename T, int NGD, typename SG = SpreadGeom, typename GC = GaussConst>
__device__ __forceinline__ void d_interp_mfma(const int *__restrict__ act, SG G, GC g,
                  const SegRec<T> *__restrict__ slab,
                  const SegRec<T> *__restrict__ pool,
                  const int *__restrict__ tcnt, const int *__restrict__ tovf,
                  const T *__restrict__ R, T *__restrict__ part) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int KC = kTileCap;  // candidates per round
  constexpr int SX = stencil_tiles(NGD, 8), SZ = stencil_tiles(NGD, 16);
  constexpr int NSLOT = kInterpSlots<NGD>;
  if (act && !*act) return;  // batched sweep: suspension converged
  __shared__ SegRec<T> s_rec[kTileCap];
  __shared__ T s_w[KC + 1][kTileW];  // row KC stays zero
  __shared__ T s_red[4][KC][6];
  __shared__ int s_wl[4][KC];  // each wave's candidate list

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, i16 = lane & 15;
  const int ntiles = G.nbx * G.nby * G.nbz;
  const size_t N = (size_t)G.nx * G.ny * G.nz;
  for (int e = threadIdx.x; e < kTileW; e += 256) s_w[KC][e] = 0;
  // persistent: block b takes tiles b, b + gridDim.x, ... (gridDim.x <=
  // ntiles, a multiple of 8 so each block's tiles stay on its XCD's run)
  int v = blockIdx.x;
  TilePrefetch<T> pf;
  {
    int bx, by, bz;
    tile_of_block(G, v, &bx, &by, &bz);
    pf.load(slab, tcnt, bx + G.nbx * (by + G.nby * bz));
  }
  for (; v < ntiles; v += gridDim.x) {
    int bx, by, bz;
    tile_of_block(G, v, &bx, &by, &bz);
    const int tile = bx + G.nbx * (by + G.nby * bz);
    const int px0 = bx * 8, py0 = by * 8, pz0 = bz * 16;
    // A operands: u_d[xy = wave*16 + i16][z = pz0 + 4q + k] for K step k
    // (lane group q supplies z 4q .. 4q + 3); in flight during the weights
    T u[3][4];
    {
      const int xy = wave * 16 + i16;
      const size_t o = ((size_t)(px0 + (xy >> 3)) * G.ny + py0 + (xy & 7)) * G.nz +
                       pz0 + 4 * q;
      ld4(R + o, u[0]);
      ld4(R + N + o, u[1]);
      ld4(R + 2 * N + o, u[2]);
    }
    const int cnt = pf.cnt;
    __syncthreads();  // the last tile is done with s_rec, s_w and s_red
    pf.store(s_rec);
    if (v + (int)gridDim.x < ntiles) {
      int nx_, ny_, nz_;
      tile_of_block(G, v + gridDim.x, &nx_, &ny_, &nz_);
      pf.load(slab, tcnt, nx_ + G.nbx * (ny_ + G.nby * nz_));
    }
    const int ovf = cnt > kTileCap ? tovf[tile] : 0;

    for (int r0 = 0; r0 < cnt; r0 += kTileCap) {
      const int nc = min(kTileCap, cnt - r0);
      if (r0 > 0) {
        __syncthreads();  // the last round still reads s_rec
        stage_tile_records<T>(slab, pool, tile, ovf, r0, nc, s_rec);
      }
      __syncthreads();
      tile_chunk_weights<T, KC, false>(g, s_rec, 0, nc, px0, py0, pz0, s_w);
      __syncthreads();
      // this wave's rows lie in x columns 2 wave and 2 wave + 1: it handles
      // only the candidates with a stencil point there (non-zero monopole or
      // dipole weight), 16 per group, and leaves zeros for the others
      {
        const T *wl = s_w[lane < nc ? lane : KC];
        const bool mine = lane < nc && (wl[2 * wave] != (T)0 ||
                                        wl[2 * wave + 1] != (T)0 ||
                                        wl[32 + 2 * wave] != (T)0 ||
                                        wl[33 + 2 * wave] != (T)0);
        const unsigned long long mm = __ballot(mine);
        if (mine) s_wl[wave][__popcll(mm & ((1ull << lane) - 1ull))] = lane;
        for (int e = lane; e < KC * 6; e += 64) (&s_red[wave][0][0])[e] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int nw = __popcll(mm);
        for (int g0 = 0; g0 < nw; g0 += 16) {
          // B operands: this lane's column is the wave's candidate g0 + i16
          const int c = g0 + i16 < nw ? s_wl[wave][g0 + i16] : KC;
          const T *w = s_w[c];
          const v4 z4 = {0, 0, 0, 0};
          v4 pg0 = z4, pg1 = z4, pg2 = z4, pd0 = z4, pd1 = z4, pd2 = z4,
             pr0 = z4, pr1 = z4;
  #pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int z = 4 * q + k;
            const T gz = w[16 + z], gdz = w[48 + z];
            const T grz = gdz * w[80 + z];
            pg0 = Mfma<T>::step(u[0][k], gz, pg0);
            pg1 = Mfma<T>::step(u[1][k], gz, pg1);
            pg2 = Mfma<T>::step(u[2][k], gz, pg2);
            pd0 = Mfma<T>::step(u[0][k], gdz, pd0);
            pd1 = Mfma<T>::step(u[1][k], gdz, pd1);
            pd2 = Mfma<T>::step(u[2][k], gdz, pd2);
            pr0 = Mfma<T>::step(u[0][k], grz, pr0);
            pr1 = Mfma<T>::step(u[1][k], grz, pr1);
          }
          // lane holds P[xy = wave*16 + row(lane, r)][candidate c]: weight
          // each row with the candidate's x/y factors
          T v0 = 0, v1 = 0, v2 = 0, w0 = 0, w1 = 0, w2 = 0;
  #pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int rxy = wave * 16 + Mfma<T>::row(lane, r);
            const int x = rxy >> 3, y = rxy & 7;
            const T gxy = w[x] * w[8 + y], dxy = w[32 + x] * w[40 + y];
            const T xx = w[64 + x], yy = w[72 + y];
            v0 += gxy * pg0[r];
            v1 += gxy * pg1[r];
            v2 += gxy * pg2[r];
            w0 += dxy * (yy * pd2[r] - pr1[r]);
            w1 += dxy * (pr0[r] - xx * pd2[r]);
            w2 += dxy * (xx * pd1[r] - yy * pd0[r]);
          }
          // the four lane groups (rows) of the wave
  #pragma unroll
          for (int off = 16; off < 64; off <<= 1) {
            v0 += __shfl_xor(v0, off, 64);
            v1 += __shfl_xor(v1, off, 64);
            v2 += __shfl_xor(v2, off, 64);
            w0 += __shfl_xor(w0, off, 64);
            w1 += __shfl_xor(w1, off, 64);
            w2 += __shfl_xor(w2, off, 64);
          }
          if (q == 0 && c < KC) {
            T *sr = s_red[wave][c];
            sr[0] = v0;
            sr[1] = v1;
            sr[2] = v2;
            sr[3] = w0;
            sr[4] = w1;
            sr[5] = w2;
          }
        }
      }
      __syncthreads();
      // the four waves in a fixed order
      if (threadIdx.x < KC * 6) {
        const int c = threadIdx.x / 6, m = threadIdx.x - 6 * c;
        if (c < nc) {
          const T sum = ((s_red[0][c][m] + s_red[1][c][m]) + s_red[2][c][m]) +
                        s_red[3][c][m];
          const SegRec<T> &rc = s_rec[c];
          int sx = bx - (rc.org & 1023) / 8, sy = by - ((rc.org >> 10) & 1023) / 8,
              sz = bz - (rc.org >> 20) / 16;
          if (sx < 0) sx += G.nbx;
          if (sy < 0) sy += G.nby;
          if (sz < 0) sz += G.nbz;
          part[((size_t)rc.id * 6 + m) * NSLOT + (sx * SX + sy) * SZ + sz] = sum;
        }
      }
    }
  }
}
template <typename T, int NGD>
__global__ void __launch_bounds__(256)
    k_interp_mfma(const int *__restrict__ act, SpreadGeom G, GaussConst g,
                  const SegRec<T> *__restrict__ slab,
                  const SegRec<T> *__restrict__ pool,
                  const int *__restrict__ tcnt, const int *__restrict__ tovf,
                  const T *__restrict__ R, T *__restrict__ part) {
  d_interp_mfma<T, NGD>(act, G, g, slab, pool, tcnt, tovf, R, part);
}
// arguments of one suspension for k_interp_mfma_multi
template <typename T, int NGD>
struct A_interp_mfma {
  const int *act;
  SpreadGeom G;
  GaussConst g;
  const SegRec<T> *slab;
  const SegRec<T> *pool;
  const int *tcnt;
  const int *tovf;
  const T *R;
  T *part;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T, int NGD>
__global__ void __launch_bounds__(256)
    k_interp_mfma_multi(const A_interp_mfma<T, NGD> *__restrict__ a) {
  const A_interp_mfma<T, NGD> &A = a[blockIdx.z];
  d_interp_mfma<T, NGD, const SpreadGeom &, const GaussConst &>(A.act, A.G, A.g, A.slab, A.pool, A.tcnt, A.tovf, A.R, A.part);
}

// One thread per (segment, component): adds the segment's tile partials in
// slot order (only the tiles its stencil meets, the ones k_interp_mfma
// wrote) and scales them: V = dx^3 sum, Omega = -dx^3 / 2 sum
// (particle_velocities / particle_rotations, FCMfunctions.cpp:740-808).
template <typename T, int NGD, typename SG = SpreadGeom>
__device__ __forceinline__ void d_interp_sum(const int *__restrict__ act, int np, SG G,
                             double dx3, const int *__restrict__ origin,
                             const T *__restrict__ part,
                             double *__restrict__ V, double *__restrict__ W) {
  constexpr int SX = stencil_tiles(NGD, 8), SZ = stencil_tiles(NGD, 16);
  constexpr int NSLOT = kInterpSlots<NGD>;
  if (act && !*act) return;  // batched sweep: suspension converged
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = e / 6, m = e - 6 * s;
  if (s >= np) return;
  const int ox = origin[3 * s], oy = origin[3 * s + 1], oz = origin[3 * s + 2];
  // every slot is read (in range); the ones no tile wrote are not added
  T v[NSLOT];
  const T *pp = part + ((size_t)s * 6 + m) * NSLOT;
#pragma unroll
  for (int k = 0; k < NSLOT; ++k) v[k] = pp[k];
  double acc = 0;
#pragma unroll
  for (int sx = 0; sx < SX; ++sx) {
    int tx = ox / 8 + sx;
    if (tx >= G.nbx) tx -= G.nbx;
    const bool mx = sx < G.nbx && stencil_meets(ox, tx * 8, 8, NGD, G.nx);
#pragma unroll
    for (int sy = 0; sy < SX; ++sy) {
      int ty = oy / 8 + sy;
      if (ty >= G.nby) ty -= G.nby;
      const bool my = sy < G.nby && stencil_meets(oy, ty * 8, 8, NGD, G.ny);
#pragma unroll
      for (int sz = 0; sz < SZ; ++sz) {
        int tz = oz / 16 + sz;
        if (tz >= G.nbz) tz -= G.nbz;
        const bool mz = sz < G.nbz && stencil_meets(oz, tz * 16, 16, NGD, G.nz);
        if (mx && my && mz) acc += (double)v[(sx * SX + sy) * SZ + sz];
      }
    }
  }
  if (m < 3)
    V[3 * s + m] = dx3 * acc;
  else
    W[3 * s + m - 3] = -0.5 * dx3 * acc;
}
template <typename T, int NGD>
__global__ void k_interp_sum(const int *__restrict__ act, int np, SpreadGeom G,
                             double dx3, const int *__restrict__ origin,
                             const T *__restrict__ part,
                             double *__restrict__ V, double *__restrict__ W) {
  d_interp_sum<T, NGD>(act, np, G, dx3, origin, part, V, W);
}
// arguments of one suspension for k_interp_sum_multi
template <typename T, int NGD>
struct A_interp_sum {
  const int *act;
  int np;
  SpreadGeom G;
  double dx3;
  const int *origin;
  const T *part;
  double *V;
  double *W;
};
// one launch for a group of suspensions with equal launch shapes:
// blockIdx.z selects the suspension
template <typename T, int NGD>
__global__ void k_interp_sum_multi(const A_interp_sum<T, NGD> *__restrict__ a) {
  const A_interp_sum<T, NGD> &A = a[blockIdx.z];
  d_interp_sum<T, NGD, const SpreadGeom &>(A.act, A.np, A.G, A.dx3, A.origin, A.part, A.V, A.W);
}

__global__ void k_friction(const int *__restrict__ act, int n3, const double *__restrict__ F,
                           const double *__restrict__ Tq, double a1, double a2,
                           double *__restrict__ V, double *__restrict__ W) {
  if (act && !*act) return;  // batched sweep: suspension converged
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) {
    V[i] = a1 * F[i];
    W[i] = a2 * Tq[i];
  }
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
// Calls f(std::integral_constant<int, E>) for the register tile E of a plan.
template <typename F>
static bool with_E(int E, F f) {
  switch (E) {
    case 6: f(std::integral_constant<int, 6>{}); return true;
    case 8: f(std::integral_constant<int, 8>{}); return true;
    case 10: f(std::integral_constant<int, 10>{}); return true;
    case 12: f(std::integral_constant<int, 12>{}); return true;
    case 16: f(std::integral_constant<int, 16>{}); return true;
    case 20: f(std::integral_constant<int, 20>{}); return true;
    case 24: f(std::integral_constant<int, 24>{}); return true;
    case 30: f(std::integral_constant<int, 30>{}); return true;
  }
  return false;
}

static int largest_divisor_le(int n, int t);

// Register tile E and compile-time length of an FFT plan's kernels: n = 2^LOGN
// for n in {128, 256, 512} with E = 8 or 16, provided the plan's radices are
// those of the compile-time rule (p2_radix); otherwise LOGN = 0 (run-time
// plan).
static bool p2_plan_matches(const FftPlan &pl, int logn) {
  int ls = 0, s = 0;
  while (ls < logn) {
    int rem = logn - ls;
    int r = (rem >= 4 && pl.E >= 16) ? 16 : (rem >= 3 && pl.E >= 8) ? 8
            : (rem >= 2)                ? 4
                                        : 2;
    if (s >= pl.nst || pl.rad[s] != r) return false;
    ++s;
    ls += (r == 16 ? 4 : r == 8 ? 3 : r == 4 ? 2 : 1);
  }
  return s == pl.nst;
}

template <typename F>
static bool with_EL(const FftPlan &pl, F f) {
  using L0 = std::integral_constant<int, 0>;
  int logn = 0;
  if (pl.n == 128) logn = 7;
  if (pl.n == 256) logn = 8;
  if (pl.n == 512) logn = 9;
  if (logn && (pl.E == 8 || pl.E == 16) && p2_plan_matches(pl, logn)) {
    auto go = [&](auto e) {
      switch (logn) {
        case 7: f(e, std::integral_constant<int, 7>{}); break;
        case 8: f(e, std::integral_constant<int, 8>{}); break;
        default: f(e, std::integral_constant<int, 9>{}); break;
      }
    };
    if (pl.E == 8)
      go(std::integral_constant<int, 8>{});
    else
      go(std::integral_constant<int, 16>{});
    return true;
  }
  return with_E(pl.E, [&](auto e) { f(e, L0{}); });
}

template <typename F>
static bool with_TZ(int tz, F f) {
  switch (tz) {
    case 4: f(std::integral_constant<int, 4>{}); return true;
    case 5: f(std::integral_constant<int, 5>{}); return true;
    case 6: f(std::integral_constant<int, 6>{}); return true;
    case 8: f(std::integral_constant<int, 8>{}); return true;
    case 10: f(std::integral_constant<int, 10>{}); return true;
    case 12: f(std::integral_constant<int, 12>{}); return true;
    case 16: f(std::integral_constant<int, 16>{}); return true;
  }
  return false;
}

template <typename F>
static bool with_NGD(int ngd, F f) {
  switch (ngd) {
    case 6: f(std::integral_constant<int, 6>{}); return true;
    case 8: f(std::integral_constant<int, 8>{}); return true;
    case 10: f(std::integral_constant<int, 10>{}); return true;
    case 12: f(std::integral_constant<int, 12>{}); return true;
    case 16: f(std::integral_constant<int, 16>{}); return true;
    case 20: f(std::integral_constant<int, 20>{}); return true;
    case 24: f(std::integral_constant<int, 24>{}); return true;
  }
  return false;
}

template <typename T>
static int set_fft_lds_attr(const FftPlan &plx, const FftPlan &ply,
                            const FftPlan &plz) {
  const int lds_max = 160 * 1024;
  bool ok = true;
  ok &= with_EL(plz, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    (void)hipFuncSetAttribute((const void *)k_fft_z_r2c<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void *)k_fft_z_c2r<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void *)k_fft_z_r2c_multi<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void *)k_fft_z_c2r_multi<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  });
  ok &= with_EL(ply, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    (void)hipFuncSetAttribute((const void *)k_fft_y<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void *)k_fft_y_multi<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  });
  ok &= with_EL(plx, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    (void)hipFuncSetAttribute((const void *)k_fft_x_project<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    (void)hipFuncSetAttribute((const void *)k_fft_x_project_multi<T, EE, LL>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
  });
  FS_CHECK(ok, FS_ERR_ARG, "no FFT register tile for these sizes");
  return FS_OK;
}
static bool good_size(int n) {
  if (n < 2 || n > 1024 || (n & 1)) return false;
  FftPlan p;
  return make_plan(n, &p, 8);
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
  FS_CHECK((p.ngd == 6 || p.ngd == 8 || p.ngd == 10 || p.ngd == 12 ||
            p.ngd == 16 || p.ngd == 20 || p.ngd == 24) &&
               p.ngd + 4 <= p.nx && p.ngd + 4 <= p.ny && p.ngd + 4 <= p.nz,
           FS_ERR_ARG,
           "ngd=%d must be one of 6,8,10,12,16,20,24 and ngd+4 <= every grid "
           "size", p.ngd);
  FS_CHECK(p.dx > 0 && p.sigma > 0 && p.sigma_dip > 0 && p.mu > 0 && p.a > 0,
           FS_ERR_ARG, "dx, sigma, sigma_dip, mu, a must be positive");
  FS_CHECK(maxnp >= 1, FS_ERR_ARG, "max_np must be >= 1");
  // radix 16 only for the FP32 fused x pass (measured at 256^3: x pass
  // 162 -> 146 us with 4 lines per workgroup; z/y passes and FP64 slower)
  make_plan(p.nx, &plx, prec == FS_FP32 ? 16 : 8);
  make_plan(p.ny, &ply, 8);
  make_plan(p.nz, &plz, 8);
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
  FS_TRY(coef.alloc((size_t)maxnp * kCoef * tsz));
  FS_TRY(srec.alloc((size_t)std::max(maxnp, 1) * rec_bytes(tsz)));
  FS_TRY(sorg.alloc((size_t)std::max(maxnp, 1) * sizeof(int)));
  FS_TRY(bin.alloc((size_t)maxnp * sizeof(int)));
  FS_TRY(slot.alloc((size_t)maxnp * sizeof(int)));
  FS_TRY(ids.alloc((size_t)maxnp * sizeof(int)));
  tune.fft_batch = (prec == FS_FP64) ? 8 : 16;
  tune.spread_tile = 16;
  tune.spread_zrun = 8;
  tune.spread_mfma = 1;
  tune.fft_xlines = (prec == FS_FP32) ? 4 : 2;
  FS_TRY(set_geometry());
  for (int i = 0; i < 8; ++i) FS_HIP(hipEventCreate(&ev[i]));
  for (int i = 0; i < 11; ++i) FS_HIP(hipEventCreate(&kev[i]));
  if (prec == FS_FP64)
    FS_TRY(set_fft_lds_attr<double>(plx, ply, plz));
  else
    FS_TRY(set_fft_lds_attr<float>(plx, ply, plz));
  return FS_OK;
}

static int largest_divisor_le(int n, int t) {
  for (int d = std::min(n, t); d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

// Blocks of a persistent tile kernel (256 threads): as many as are resident
// on the device at once, a multiple of 8, at most one per tile.
template <typename K>
static int persistent_blocks(K kernel, int ntiles) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) !=
          hipSuccess ||
      per < 1 || cus < 1)
    return ntiles;
  long b = (long)cus * per;
  b = std::max(8L, b - b % 8);
  return (int)std::min<long>(b, ntiles);
}

int FcmImpl::set_geometry() {
  int t = tune.spread_tile;
  FS_CHECK(t == 4 || t == 8 || t == 16, FS_ERR_ARG,
           "spread_tile must be 4, 8 or 16");
  FS_CHECK(tune.spread_zrun >= 4 && tune.spread_zrun <= 16, FS_ERR_ARG,
           "spread_zrun must be in [4, 16]");
  // tile z extent: the largest instantiated length <= spread_zrun dividing
  // nz with nz >= run + ngd
  static const int runs[] = {16, 12, 10, 8, 6, 5, 4};
  int tz = 0;
  for (int r : runs)
    if (r <= tune.spread_zrun && p.nz % r == 0 && p.nz >= r + p.ngd) {
      tz = r;
      break;
    }
  FS_CHECK(tz > 0, FS_ERR_ARG, "no spread z run divides nz=%d", p.nz);
  int tx = largest_divisor_le(p.nx, t), ty = largest_divisor_le(p.ny, t);
  // matrix-core spread: fixed 8 x 8 x 16 tiles (four 16-point MFMA rows)
  mfma_spread = tune.spread_mfma == 1 &&
                p.nx % 8 == 0 && p.ny % 8 == 0 && p.nz % 16 == 0 &&
                p.nz >= 16 + p.ngd;
  if (mfma_spread) {
    tx = ty = 8;
    tz = 16;
  }
  int bins = 0;
  for (int attempt = 0; attempt < 8; ++attempt) {
    int hx = (p.ngd - 1 + tx - 1) / tx, hy = (p.ngd - 1 + ty - 1) / ty,
        hz = (p.ngd - 1 + tz - 1) / tz;
    bins = std::min(hx + 1, p.nx / tx) * std::min(hy + 1, p.ny / ty) *
           std::min(hz + 1, p.nz / tz);
    if (bins <= kMaxBinsPerTile || mfma_spread) break;
    int nx2 = largest_divisor_le(p.nx, 2 * tx), ny2 = largest_divisor_le(p.ny, 2 * ty);
    if (nx2 * ny2 > 256) break;
    tx = nx2;
    ty = ny2;
  }
  FS_CHECK(bins <= kMaxBinsPerTile && tx * ty <= 256, FS_ERR_ARG,
           "no spread tile fits grid %dx%dx%d with ngd=%d", p.nx, p.ny, p.nz,
           p.ngd);
  geom.nx = p.nx;
  geom.ny = p.ny;
  geom.nz = p.nz;
  geom.ngd = p.ngd;
  geom.tx = tx;
  geom.ty = ty;
  geom.tz = tz;
  geom.nbx = p.nx / tx;
  geom.nby = p.ny / ty;
  geom.nbz = p.nz / tz;
  geom.hx = (p.ngd - 1 + tx - 1) / tx;
  geom.hy = (p.ngd - 1 + ty - 1) / ty;
  geom.hz = (p.ngd - 1 + tz - 1) / tz;
  nbins = geom.nbx * geom.nby * geom.nbz;
  if (mfma_spread) {
    const size_t nslot = (size_t)stencil_tiles(p.ngd, 8) *
                         stencil_tiles(p.ngd, 8) * stencil_tiles(p.ngd, 16);
    FS_TRY(ipart.alloc((size_t)max_np * 6 * tsz * nslot));
    FS_TRY(tslab.alloc((size_t)nbins * kTileCap * rec_bytes(tsz)));
    FS_TRY(tcnt.alloc((size_t)nbins * sizeof(int)));
    FS_TRY(tovf.alloc((size_t)nbins * sizeof(int)));
    FS_TRY(tpool.alloc((size_t)max_np * nslot * rec_bytes(tsz)));
    with_NGD(p.ngd, [&](auto ng) {
      constexpr int NG = decltype(ng)::value;
      grid_interp = tsz == 8 ? persistent_blocks(k_interp_mfma<double, NG>, nbins)
                             : persistent_blocks(k_interp_mfma<float, NG>, nbins);
    });
  }
  // one more counter after the bins: the tile lists' pool top
  FS_TRY(bin_count.alloc((size_t)(nbins + 1) * sizeof(int)));
  FS_HIP(hipMemset(bin_count.p, 0, (size_t)(nbins + 1) * sizeof(int)));
  FS_TRY(bin_start.alloc((size_t)(nbins + 1) * sizeof(int)));
  FS_TRY(scan_tmp.alloc((size_t)(nbins / kScanBlock + 4) * sizeof(int)));
  // FFT pass geometry: z rows per workgroup (even divisor of nx), y-pass run
  // of consecutive x, x-pass lines per workgroup; workgroups stay <= 1024
  const size_t cs = 2 * tsz, lds_max = 160 * 1024;
  zP = 0;
  for (int xb = std::min(16, p.nx); xb >= 2; --xb)
    if (xb % 2 == 0 && p.nx % xb == 0 && p.nz / plz.E * (xb / 2) <= 1024 &&
        cs * (p.nz + (size_t)(xb / 2) * fft_padded_len(p.nz)) <= lds_max) {
      zP = xb / 2;
      break;
    }
  yB = largest_divisor_le(p.nx, tune.fft_batch);
  while (yB > 1 && p.ny / ply.E * yB > 1024) yB = largest_divisor_le(p.nx, yB - 1);
  xQ = largest_divisor_le(p.ny, tune.fft_xlines);
  while (xQ > 1 && (p.nx / plx.E * 3 * xQ > 1024 ||
                    cs * (p.nx + 3 * (size_t)xQ * fft_padded_len(p.nx)) > lds_max))
    xQ = largest_divisor_le(p.ny, xQ - 1);
  FS_CHECK(zP >= 1 && p.ny / ply.E * yB <= 1024 && p.nx / plx.E * 3 * xQ <= 1024,
           FS_ERR_ARG, "no FFT workgroup shape fits %dx%dx%d", p.nx, p.ny, p.nz);
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
  srec.release();
  sorg.release();
  tslab.release();
  tcnt.release();
  tovf.release();
  tpool.release();
  ipart.release();
  bin_count.release();
  bin_start.release();
  scan_tmp.release();
  for (int i = 0; i < 8; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  for (int i = 0; i < 11; ++i)
    if (kev[i]) (void)hipEventDestroy(kev[i]);
  for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
  prof_ev.clear();
  prof_used = 0;
}

// Gaussian constants of the spread, tables and interpolation.
static GaussConst gauss_const(const fs_fcm_params &p) {
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
  g.e2 = 1.4426950408889634074 / g.anorm2;
  g.ed2 = 1.4426950408889634074 / g.anormdip2;
  g.nrsd = -1.0 / sigmadipsq;
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = p.nz;
  return g;
}

// Constants of the Stokes projection (flow_solve).
static ProjConst proj_const(const fs_fcm_params &p) {
  ProjConst pc;
  pc.nx = p.nx;
  pc.ny = p.ny;
  pc.nz = p.nz;
  pc.nh = p.nz / 2 + 1;
  pc.kfx = 2.0 * kPi / (p.nx * p.dx);
  pc.kfy = 2.0 * kPi / (p.ny * p.dx);
  pc.kfz = 2.0 * kPi / (p.nz * p.dx);
  pc.scale = 1.0 / ((double)p.nx * p.ny * p.nz * p.mu);
  return pc;
}

template <typename T>
int FcmImpl::run(int np, const double *Y, const double *F, const double *Tq,
                 double *V, double *W, hipStream_t st) {
  const GaussConst g = gauss_const(p);

  hipEvent_t *pset = nullptr;  // this solve's profiling events
  if (prof) {
    while (prof_ev.size() < prof_used + 7) {
      hipEvent_t e = nullptr;
      FS_HIP(hipEventCreate(&e));
      prof_ev.push_back(e);
    }
    pset = prof_ev.data() + prof_used;
    prof_used += 7;
  }
  auto mark = [&](int i) {
    if (timing) (void)hipEventRecord(ev[i], st);
    if (pset) (void)hipEventRecord(pset[i], st);
  };
  // kernel boundaries (timing only); a boundary not recorded in this solve
  // marks a kernel the path does not have
  for (int i = 0; i < 11; ++i) kev_used[i] = false;
  auto kmark = [&](int i) {
    if (timing) {
      (void)hipEventRecord(kev[i], st);
      kev_used[i] = true;
    }
  };
  mark(0);
  kmark(0);
  // bin counting sort. Up to kScanOneMax bins the counts are scanned by one
  // workgroup that also zeroes them (and the pool counter after them) for
  // the next solve: they start zeroed (set_geometry) and every solve leaves
  // them so, so no fill pass is needed.
  const bool scan_one = nbins <= kScanOneMax;
  if (!scan_one)
    hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(nbins + 1, 256)), dim3(256), 0, st,
                       act, bin_count.as<int>(), nbins + 1, 0);
  if (!mfma_spread)  // the matrix-core kernels form their weights per tile
    hipLaunchKernelGGL(k_prepare_tab<T>, dim3(ceil_div((long)np * 3 * p.ngd, 256)),
                       dim3(256), 0, st, act, np, Y, g, tab.as<T>());
  hipLaunchKernelGGL(k_prepare<T>, dim3(ceil_div(np, 64)), dim3(64), 0, st, act, np,
                     Y, F, Tq, g, geom.tx, geom.ty, geom.tz, geom.nbx, geom.nby,
                     geom.nbz, origin.as<int>(), coef.as<T>(), bin.as<int>(),
                     bin_count.as<int>(), slot.as<int>());
  if (scan_one)
    FS_TRY(device_scan_one(act, bin_count.as<int>(), nbins, bin_start.as<int>(), 1, st));
  else
    FS_TRY(device_scan(bin_count.as<int>(), nbins, bin_start.as<int>(),
                       scan_tmp.as<int>(), st));
  hipLaunchKernelGGL(k_scatter, dim3(ceil_div(np, 256)), dim3(256), 0, st, np,
                     bin.as<int>(), slot.as<int>(), bin_start.as<int>(),
                     ids.as<int>());
  hipLaunchKernelGGL(k_sort_bins<T>, dim3(ceil_div(nbins, 64)), dim3(64), 0, st,
                     act, nbins, bin_start.as<int>(), ids.as<int>(),
                     (const int *)origin.as<int>(), Y, F, Tq, p.dx, p.ngd / 2,
                     mfma_spread ? srec.as<SegRec<T>>() : (SegRec<T> *)nullptr,
                     mfma_spread ? sorg.as<int>() : (int *)nullptr);
  const int ntiles = nbins;
  kmark(1);
  if (mfma_spread)
    hipLaunchKernelGGL(k_tile_lists<T>, dim3(ceil_div(ntiles, 4)), dim3(256), 0, st,
                       act, geom, bin_start.as<int>(), (const int *)sorg.as<int>(),
                       srec.as<SegRec<T>>(),
                       tslab.as<SegRec<T>>(), tcnt.as<int>(), tovf.as<int>(),
                       tpool.as<SegRec<T>>(), bin_count.as<int>() + nbins);
  mark(1);
  if (mfma_spread) kmark(2);
  bool okz = true;
  if (mfma_spread)
    hipLaunchKernelGGL(k_spread_mfma<T>, dim3(ntiles), dim3(256), 0, st, act, geom,
                       g, tslab.as<SegRec<T>>(), tpool.as<SegRec<T>>(),
                       tcnt.as<int>(), tovf.as<int>(), R.as<T>());
  if (!mfma_spread)
    okz = with_TZ(geom.tz, [&](auto z) {
      constexpr int TZ = decltype(z)::value;
      hipLaunchKernelGGL((k_spread<T, TZ>), dim3(ntiles), dim3(geom.tx * geom.ty),
                         0, st, act, geom, tab.as<T>(), origin.as<int>(), coef.as<T>(),
                         bin_start.as<int>(), ids.as<int>(), R.as<T>());
    });
  FS_CHECK(okz, FS_ERR_ARG, "spread dispatch failed");
  FS_LAUNCHED();
  mark(2);
  kmark(3);
  const int nh = p.nz / 2 + 1;
  const size_t csz = sizeof(cx<T>);
  bool ok = with_EL(plz, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    hipLaunchKernelGGL((k_fft_z_r2c<T, EE, LL>), dim3(p.ny * (p.nx / (2 * zP)), 3),
                       dim3(p.nz / EE * zP),
                       csz * (p.nz + (size_t)zP * fft_padded_len(p.nz)), st, act,
                       R.as<T>(), C.as<cx<T>>(), plz, twz.as<cx<T>>(), p.nx,
                       p.ny, zP, make_div(2 * zP), make_div(zP));
  });
  kmark(4);
  ok &= with_EL(ply, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    hipLaunchKernelGGL((k_fft_y<T, EE, LL>), dim3(nh * (p.nx / yB), 3),
                       dim3(p.ny / EE * yB), csz * (p.ny + (size_t)yB * fft_rowpad_len(p.ny)), st, act,
                       C.as<cx<T>>(), ply, twy.as<cx<T>>(), p.nx, nh, yB, 0,
                       make_div(yB));
  });
  FS_CHECK(ok, FS_ERR_ARG, "FFT dispatch failed");
  FS_LAUNCHED();
  mark(3);
  kmark(5);
  const ProjConst pc = proj_const(p);
  ok = with_EL(plx, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    hipLaunchKernelGGL((k_fft_x_project<T, EE, LL>), dim3(nh * (p.ny / xQ)),
                       dim3(p.nx / EE * 3 * xQ),
                       csz * (p.nx + 3 * (size_t)xQ * fft_padded_len(p.nx)), st, act,
                       C.as<cx<T>>(),
                       plx, twx.as<cx<T>>(), pc, xQ, make_div(xQ));
  });
  FS_CHECK(ok, FS_ERR_ARG, "FFT dispatch failed");
  FS_LAUNCHED();
  mark(4);
  kmark(6);
  ok = with_EL(ply, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    hipLaunchKernelGGL((k_fft_y<T, EE, LL>), dim3(nh * (p.nx / yB), 3),
                       dim3(p.ny / EE * yB), csz * (p.ny + (size_t)yB * fft_rowpad_len(p.ny)), st, act,
                       C.as<cx<T>>(), ply, twy.as<cx<T>>(), p.nx, nh, yB, 1,
                       make_div(yB));
  });
  kmark(7);
  ok &= with_EL(plz, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    hipLaunchKernelGGL((k_fft_z_c2r<T, EE, LL>), dim3(p.ny * (p.nx / (2 * zP)), 3),
                       dim3(p.nz / EE * zP),
                       csz * (p.nz + (size_t)zP * fft_padded_len(p.nz)), st, act,
                       C.as<cx<T>>(), R.as<T>(), plz, twz.as<cx<T>>(), p.nx,
                       p.ny, zP, make_div(zP));
  });
  FS_CHECK(ok, FS_ERR_ARG, "FFT dispatch failed");
  FS_LAUNCHED();
  mark(5);
  kmark(8);
  double dx3 = p.dx * p.dx * p.dx;
  bool okn = with_NGD(p.ngd, [&](auto ng) {
    constexpr int NG = decltype(ng)::value;
    if (mfma_spread) {
      hipLaunchKernelGGL((k_interp_mfma<T, NG>), dim3(grid_interp), dim3(256), 0,
                         st, act, geom, g, tslab.as<SegRec<T>>(),
                         tpool.as<SegRec<T>>(), tcnt.as<int>(), tovf.as<int>(),
                         R.as<T>(), ipart.as<T>());
      kmark(9);
      hipLaunchKernelGGL((k_interp_sum<T, NG>), dim3(ceil_div(6L * np, 256)),
                         dim3(256), 0, st, act, np, geom, dx3,
                         origin.as<int>(), ipart.as<T>(), V, W);
    } else {
      hipLaunchKernelGGL((k_interp<T, NG>),
                         dim3(8 * ceil_div(ceil_div(np, 4), 8)), dim3(256), 0,
                         st, act, np, geom, dx3, tab.as<T>(), origin.as<int>(),
                         ids.as<int>(), R.as<T>(), V, W);
    }
  });
  FS_CHECK(okn, FS_ERR_ARG, "interp dispatch failed");
  FS_LAUNCHED();
  mark(6);
  kmark(10);
  return FS_OK;
}

// ---------------------------------------------------------------------
// Batched sweep: the mobility solves of a group of suspensions with one
// launch per FCM stage. Members have equal launch shapes (same precision,
// grid, Gaussian support and tile/FFT shapes), so every stage's grid is the
// same for each member and blockIdx.z selects the member; the kernels are
// the single-solve kernels' bodies (the k_*_multi wrappers), so each
// member's results are bit-identical to its own solve. Segment counts may
// differ: segment-indexed stages launch for the largest and return past a
// member's own count. The argument tables are built once for fixed
// Y/F/T/V/W and loop-control pointers.
// ---------------------------------------------------------------------
struct A_scan_small {
  const int *act;
  int *cnt;
  int n;
  int *start;
};
// start[0..n] = exclusive scan of cnt[0..n) with one workgroup per member;
// up to kScanOneMax bins the single-solve scan, which leaves cnt[0..n]
// zeroed as the single-solve path expects, else 1024-element chunks with a
// running carry
__global__ void __launch_bounds__(1024)
    k_scan_small_multi(const A_scan_small *__restrict__ a) {
  const A_scan_small &A = a[blockIdx.z];
  if (A.n <= kScanOneMax) {
    d_scan_one(A.act, A.cnt, A.n, A.start, 1);
    return;
  }
  if (A.act && !*A.act) return;  // batched sweep: suspension converged
  __shared__ int s[1024];
  const int t = threadIdx.x;
  int carry = 0;
  for (int base = 0; base < A.n; base += 1024) {
    const int i = base + t;
    s[t] = (i < A.n) ? A.cnt[i] : 0;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int v = (t >= off) ? s[t - off] : 0;
      __syncthreads();
      s[t] += v;
      __syncthreads();
    }
    if (i < A.n) A.start[i] = carry + ((t == 0) ? 0 : s[t - 1]);
    carry += s[1023];
    __syncthreads();  // s is rewritten by the next chunk
  }
  if (t == 0) A.start[A.n] = carry;
}

bool fcm_same_shapes(const FcmImpl &a, const FcmImpl &b) {
  return a.precision == b.precision && a.p.nx == b.p.nx && a.p.ny == b.p.ny &&
         a.p.nz == b.p.nz && a.p.ngd == b.p.ngd && a.mfma_spread && b.mfma_spread &&
         a.zP == b.zP && a.yB == b.yB && a.xQ == b.xQ &&
         a.grid_interp == b.grid_interp && a.nbins == b.nbins;
}

namespace {

// One pass over the stages: with `host` set, appends each stage's argument
// table (256-byte aligned) and records its offset; otherwise launches each
// stage with its table in G.tab.
template <typename T>
int group_stages(FcmGroup &G, std::vector<char> *host, hipStream_t st) {
  const int S = (int)G.m.size();
  const FcmImpl &f0 = *G.m[0].fcm;
  const fs_fcm_params &p = f0.p;
  int npmax = 0;
  for (auto &m : G.m) npmax = std::max(npmax, m.np);
  const int nbins = f0.nbins, ntiles = f0.nbins, nh = p.nz / 2 + 1;
  const size_t csz = sizeof(cx<T>);
  size_t k = 0;  // stage index
  // table of stage k: build it (host) or return its device address
  auto table = [&](auto make) -> const void * {
    using A = decltype(make(G.m[0]));
    if (host) {
      size_t at = (host->size() + 255) & ~(size_t)255;
      host->resize(at + sizeof(A) * S);
      for (int i = 0; i < S; ++i) {
        A v = make(G.m[i]);
        std::memcpy(host->data() + at + sizeof(A) * i, &v, sizeof(A));
      }
      G.off.push_back(at);
      ++k;
      return nullptr;
    }
    return static_cast<const char *>(G.tab.p) + G.off[k++];
  };
  const GaussConst g = gauss_const(p);
  const ProjConst pc = proj_const(p);
  const double dx3 = p.dx * p.dx * p.dx;
  {
    auto t = table([&](const FcmGroupMember &m) {
      return A_fill_int{m.act, m.fcm->bin_count.as<int>(), nbins + 1, 0};
    });
    if (!host)
      hipLaunchKernelGGL(k_fill_int_multi, dim3(ceil_div(nbins + 1, 256), 1, S),
                         dim3(256), 0, st, (const A_fill_int *)t);
  }
  {
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_prepare<T>{m.act, m.np, m.Y, m.F, m.T, g, f.geom.tx, f.geom.ty,
                          f.geom.tz, f.geom.nbx, f.geom.nby, f.geom.nbz,
                          f.origin.as<int>(), f.coef.as<T>(), f.bin.as<int>(),
                          f.bin_count.as<int>(), f.slot.as<int>()};
    });
    if (!host)
      hipLaunchKernelGGL(k_prepare_multi<T>, dim3(ceil_div(npmax, 64), 1, S),
                         dim3(64), 0, st, (const A_prepare<T> *)t);
  }
  {
    auto t = table([&](const FcmGroupMember &m) {
      return A_scan_small{m.act, m.fcm->bin_count.as<int>(), nbins,
                          m.fcm->bin_start.as<int>()};
    });
    if (!host)
      hipLaunchKernelGGL(k_scan_small_multi, dim3(1, 1, S), dim3(1024), 0, st,
                         (const A_scan_small *)t);
  }
  {
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_scatter{m.np, f.bin.as<int>(), f.slot.as<int>(),
                       f.bin_start.as<int>(), f.ids.as<int>()};
    });
    if (!host)
      hipLaunchKernelGGL(k_scatter_multi, dim3(ceil_div(npmax, 256), 1, S),
                         dim3(256), 0, st, (const A_scatter *)t);
  }
  {
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_sort_bins<T>{m.act, nbins, f.bin_start.as<int>(), f.ids.as<int>(),
                            f.origin.as<int>(), m.Y, m.F, m.T, p.dx, p.ngd / 2,
                            f.srec.as<SegRec<T>>(), f.sorg.as<int>()};
    });
    if (!host)
      hipLaunchKernelGGL(k_sort_bins_multi<T>, dim3(ceil_div(nbins, 64), 1, S),
                         dim3(64), 0, st, (const A_sort_bins<T> *)t);
  }
  {
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_tile_lists<T>{m.act, f.geom, f.bin_start.as<int>(), f.sorg.as<int>(),
                             f.srec.as<SegRec<T>>(), f.tslab.as<SegRec<T>>(),
                             f.tcnt.as<int>(), f.tovf.as<int>(),
                             f.tpool.as<SegRec<T>>(), f.bin_count.as<int>() + nbins};
    });
    if (!host)
      hipLaunchKernelGGL(k_tile_lists_multi<T>, dim3(ceil_div(ntiles, 4), 1, S),
                         dim3(256), 0, st, (const A_tile_lists<T> *)t);
  }
  {
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_spread_mfma<T>{m.act, f.geom, g, f.tslab.as<SegRec<T>>(),
                              f.tpool.as<SegRec<T>>(), f.tcnt.as<int>(),
                              f.tovf.as<int>(), f.R.as<T>()};
    });
    if (!host)
      hipLaunchKernelGGL(k_spread_mfma_multi<T>, dim3(ntiles, 1, S), dim3(256), 0,
                         st, (const A_spread_mfma<T> *)t);
  }
  bool ok = with_EL(f0.plz, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_fft_z_r2c<T, EE, LL>{m.act, f.R.as<T>(), f.C.as<cx<T>>(), f.plz,
                                    f.twz.as<cx<T>>(), p.nx, p.ny, f.zP,
                                    make_div(2 * f.zP), make_div(f.zP)};
    });
    if (!host)
      hipLaunchKernelGGL((k_fft_z_r2c_multi<T, EE, LL>),
                         dim3(p.ny * (p.nx / (2 * f0.zP)), 3, S),
                         dim3(p.nz / EE * f0.zP),
                         csz * (p.nz + (size_t)f0.zP * fft_padded_len(p.nz)), st,
                         (const A_fft_z_r2c<T, EE, LL> *)t);
  });
  for (int inv = 0; inv < 2 && ok; ++inv) {
    ok &= with_EL(f0.ply, [&](auto e, auto l) {
      constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
      auto t = table([&](const FcmGroupMember &m) {
        const FcmImpl &f = *m.fcm;
        return A_fft_y<T, EE, LL>{m.act, f.C.as<cx<T>>(), f.ply, f.twy.as<cx<T>>(),
                                  p.nx, nh, f.yB, inv, make_div(f.yB)};
      });
      if (!host)
        hipLaunchKernelGGL((k_fft_y_multi<T, EE, LL>),
                           dim3(nh * (p.nx / f0.yB), 3, S), dim3(p.ny / EE * f0.yB),
                           csz * (p.ny + (size_t)f0.yB * fft_rowpad_len(p.ny)), st,
                           (const A_fft_y<T, EE, LL> *)t);
    });
    if (inv == 0)
      ok &= with_EL(f0.plx, [&](auto e, auto l) {
        constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
        auto t = table([&](const FcmGroupMember &m) {
          const FcmImpl &f = *m.fcm;
          return A_fft_x_project<T, EE, LL>{m.act, f.C.as<cx<T>>(), f.plx,
                                            f.twx.as<cx<T>>(), pc, f.xQ,
                                            make_div(f.xQ)};
        });
        if (!host)
          hipLaunchKernelGGL((k_fft_x_project_multi<T, EE, LL>),
                             dim3(nh * (p.ny / f0.xQ), 1, S),
                             dim3(p.nx / EE * 3 * f0.xQ),
                             csz * (p.nx + 3 * (size_t)f0.xQ * fft_padded_len(p.nx)),
                             st, (const A_fft_x_project<T, EE, LL> *)t);
      });
  }
  ok &= with_EL(f0.plz, [&](auto e, auto l) {
    constexpr int EE = decltype(e)::value, LL = decltype(l)::value;
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_fft_z_c2r<T, EE, LL>{m.act, f.C.as<cx<T>>(), f.R.as<T>(), f.plz,
                                    f.twz.as<cx<T>>(), p.nx, p.ny, f.zP,
                                    make_div(f.zP)};
    });
    if (!host)
      hipLaunchKernelGGL((k_fft_z_c2r_multi<T, EE, LL>),
                         dim3(p.ny * (p.nx / (2 * f0.zP)), 3, S),
                         dim3(p.nz / EE * f0.zP),
                         csz * (p.nz + (size_t)f0.zP * fft_padded_len(p.nz)), st,
                         (const A_fft_z_c2r<T, EE, LL> *)t);
  });
  ok &= with_NGD(p.ngd, [&](auto ng) {
    constexpr int NG = decltype(ng)::value;
    auto t = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_interp_mfma<T, NG>{m.act, f.geom, g, f.tslab.as<SegRec<T>>(),
                                  f.tpool.as<SegRec<T>>(), f.tcnt.as<int>(),
                                  f.tovf.as<int>(), f.R.as<T>(), f.ipart.as<T>()};
    });
    if (!host)
      hipLaunchKernelGGL((k_interp_mfma_multi<T, NG>), dim3(f0.grid_interp, 1, S),
                         dim3(256), 0, st, (const A_interp_mfma<T, NG> *)t);
    auto t2 = table([&](const FcmGroupMember &m) {
      const FcmImpl &f = *m.fcm;
      return A_interp_sum<T, NG>{m.act, m.np, f.geom, dx3, f.origin.as<int>(),
                                 f.ipart.as<T>(), m.V, m.W};
    });
    if (!host)
      hipLaunchKernelGGL((k_interp_sum_multi<T, NG>),
                         dim3(ceil_div(6L * npmax, 256), 1, S), dim3(256), 0, st,
                         (const A_interp_sum<T, NG> *)t2);
  });
  FS_CHECK(ok, FS_ERR_ARG, "group FCM dispatch failed");
  if (!host) FS_LAUNCHED();
  return FS_OK;
}

}  // namespace

int FcmGroup::build() {
  FS_CHECK(!m.empty(), FS_ERR_ARG, "empty FCM group");
  for (auto &x : m)
    FS_CHECK(fcm_same_shapes(*m[0].fcm, *x.fcm) && x.np >= 1 &&
                 x.np <= x.fcm->max_np,
             FS_ERR_ARG, "FCM group members differ in launch shapes or sizes");
  std::vector<char> host;
  off.clear();
  int rc = m[0].fcm->precision == FS_FP64 ? group_stages<double>(*this, &host, nullptr)
                                          : group_stages<float>(*this, &host, nullptr);
  FS_TRY(rc);
  FS_TRY(tab.alloc(host.size()));
  FS_HIP(hipMemcpy(tab.p, host.data(), host.size(), hipMemcpyHostToDevice));
  return FS_OK;
}

int FcmGroup::solve(hipStream_t st) {
  FS_CHECK(tab.p && !off.empty(), FS_ERR_STATE, "FCM group not built");
  return m[0].fcm->precision == FS_FP64 ? group_stages<double>(*this, nullptr, st)
                                        : group_stages<float>(*this, nullptr, st);
}

// add the stage times of every solve since the last call into acc7
// (buckets of stage_times); the caller has synchronised the stream
int FcmImpl::prof_collect(double *acc7) {
  for (size_t k = 0; k + 7 <= prof_used; k += 7) {
    float t[6];
    for (int i = 0; i < 6; ++i)
      FS_HIP(hipEventElapsedTime(&t[i], prof_ev[k + i], prof_ev[k + i + 1]));
    acc7[1] += t[0];
    acc7[2] += t[1];
    acc7[3] += t[2];
    acc7[4] += t[3];
    acc7[5] += t[4];
    acc7[6] += t[5];
  }
  prof_used = 0;
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
                     st, act, 3 * np, F, Tq, a1, a2, V, W);
  FS_LAUNCHED();
  return FS_OK;
}

int FcmImpl::kernel_times(float *ms10) {
  for (int i = 0; i < 10; ++i) ms10[i] = 0;
  if (!timing || !kev_used[10]) return FS_OK;
  FS_HIP(hipEventSynchronize(kev[10]));
  auto span = [&](int a, int b, float *out) -> int {
    FS_HIP(hipEventElapsedTime(out, kev[a], kev[b]));
    return FS_OK;
  };
  const int sp0 = kev_used[2] ? 2 : 1;   // spread starts after the lists
  const int ie = kev_used[9] ? 9 : 10;   // interpolation ends before its sums
  FS_TRY(span(0, 1, &ms10[0]));
  if (kev_used[2]) FS_TRY(span(1, 2, &ms10[1]));
  FS_TRY(span(sp0, 3, &ms10[2]));
  for (int k = 3; k < 8; ++k) FS_TRY(span(k, k + 1, &ms10[k]));
  FS_TRY(span(8, ie, &ms10[8]));
  if (kev_used[9]) FS_TRY(span(9, 10, &ms10[9]));
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

int FcmImpl::set_tuning(const fs_fcm_tuning &t) {
  fs_fcm_tuning old = tune;
  fs_fcm_tuning nt = old;
  if (t.fft_batch > 0) nt.fft_batch = t.fft_batch;
  if (t.spread_tile > 0) nt.spread_tile = t.spread_tile;
  if (t.spread_zrun > 0) nt.spread_zrun = t.spread_zrun;
  if (t.fft_xlines > 0) nt.fft_xlines = t.fft_xlines;
  if (t.spread_mfma >= 0) nt.spread_mfma = t.spread_mfma != 0;
  FS_CHECK(nt.fft_batch >= 1 && nt.fft_batch <= 64 && nt.fft_xlines >= 1 &&
               nt.fft_xlines <= 16,
           FS_ERR_ARG, "fft_batch must be in [1,64], fft_xlines in [1,16]");
  // LDS: y pass holds ny + B*(row-padded ny) complex, x pass nx + 3Q*(padded nx)
  size_t cs = (size_t)tsz * 2;
  size_t lds = std::max(cs * (p.ny + (size_t)nt.fft_batch * fft_rowpad_len(p.ny)),
                        cs * (p.nx + 3 * (size_t)nt.fft_xlines *
                                         fft_padded_len(p.nx)));
  FS_CHECK(lds <= 160 * 1024, FS_ERR_ARG, "tuning needs %zu B of LDS", lds);
  tune = nt;
  int rc = set_geometry();
  if (rc != FS_OK) {
    tune = old;
    set_geometry();
  }
  return rc;
}

// synthetic segments for the autotuner: positions uniform in the box,
// forces and torques uniform in [-1, 1] (splitmix64 of the element index)
__global__ void k_synth(int np, double Lx, double Ly, double Lz,
                        double *__restrict__ Y, double *__restrict__ F,
                        double *__restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 9 * np) return;
  unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull +
                         0x2545F4914F6CDD1Dull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  const int arr = i / (3 * np), k = i - arr * 3 * np, d = k % 3;
  if (arr == 0)
    Y[k] = u * (d == 0 ? Lx : d == 1 ? Ly : Lz);
  else
    (arr == 1 ? F : T)[k] = 2.0 * u - 1.0;
}

int FcmImpl::autotune(int np, int reps, fs_fcm_tuning *best) {
  FS_CHECK(np >= 1 && np <= max_np, FS_ERR_SIZE, "np=%d outside [1, %d]", np,
           max_np);
  FS_CHECK(reps >= 1, FS_ERR_ARG, "reps must be >= 1");
  dbuf buf;
  FS_TRY(buf.alloc(sizeof(double) * 15 * (size_t)np));
  double *Y = buf.as<double>(), *F = Y + 3 * (size_t)np,
         *T = F + 3 * (size_t)np, *V = T + 3 * (size_t)np,
         *W = V + 3 * (size_t)np;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FS_HIP(hipStreamCreate(&st));
  int rc = FS_OK;
  auto cleanup = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  };
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    FS_CHECK(false, FS_ERR_HIP, "event creation failed");
  }
  hipLaunchKernelGGL(k_synth, dim3(ceil_div(9 * np, 256)), dim3(256), 0, st, np,
                     p.nx * p.dx, p.ny * p.dx, p.nz * p.dx, Y, F, T);
  auto time_it = [&](float *ms) -> int {
    FS_TRY(mobility(np, Y, F, T, V, W, st));  // warm-up
    FS_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) FS_TRY(mobility(np, Y, F, T, V, W, st));
    FS_HIP(hipEventRecord(e1, st));
    FS_HIP(hipEventSynchronize(e1));
    FS_HIP(hipEventElapsedTime(ms, e0, e1));
    return FS_OK;
  };
  fs_fcm_tuning cur = tune;
  float tbest = 0;
  rc = time_it(&tbest);
  static const int mfma[] = {0, 1}, tiles[] = {4, 8, 16}, zruns[] = {4, 8, 16},
                   batches[] = {4, 8, 16, 32}, xlines[] = {1, 2, 4, 8};
  struct Field {
    int fs_fcm_tuning::*m;
    const int *v;
    int n;
  } fields[] = {{&fs_fcm_tuning::spread_mfma, mfma, 2},
                {&fs_fcm_tuning::spread_tile, tiles, 3},
                {&fs_fcm_tuning::spread_zrun, zruns, 3},
                {&fs_fcm_tuning::fft_batch, batches, 4},
                {&fs_fcm_tuning::fft_xlines, xlines, 4}};
  for (const Field &fd : fields) {
    if (rc != FS_OK) break;
    for (int k = 0; k < fd.n && rc == FS_OK; ++k) {
      if (fd.v[k] == cur.*fd.m) continue;
      fs_fcm_tuning trial = cur;
      trial.*fd.m = fd.v[k];
      if (set_tuning(trial) != FS_OK) continue;  // shape does not fit
      float t = 0;
      rc = time_it(&t);
      if (rc == FS_OK && t < tbest) {
        tbest = t;
        cur = tune;
      }
    }
    if (set_tuning(cur) != FS_OK) rc = FS_ERR_STATE;
  }
  cleanup();
  if (rc != FS_OK) return rc;
  if (best) *best = tune;
  set_error("");
  return FS_OK;
}

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

extern "C" int fs_fcm_kernel_times(fs_fcm *h, float *ms10) {
  FS_CHECK(h && ms10, FS_ERR_ARG, "null handle or output");
  return h->impl.kernel_times(ms10);
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
  return h->impl.set_tuning(*t);
}

extern "C" int fs_fcm_autotune(fs_fcm *h, int np, int reps, fs_fcm_tuning *best) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.autotune(np, reps, best);
}
