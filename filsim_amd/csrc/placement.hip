// placement.hip -- initial configurations: straight filaments with heads
// uniform in the periodic box and orientations uniform on the sphere,
// rejected while any two segments of different filaments are closer than
// min_sep. The reference places filaments one after another, testing each
// trial against every placed segment (Filament.cpp:8-153,
// checkOverlapWithPeriodic; filament_initialisation_functions.cpp:631-650).
//
// Both entry points draw from one counter-based stream (splitmix64 from the
// seed; trial k uses draws 11k+1 .. 11k+11: head 3 uniforms, quaternion 4
// Box-Muller normals) and accept the same trials, so they return the same
// bits:
//   fs_make_filaments      sequential, hashed cell grid on the host;
//   fs_make_filaments_gpu  windows of trials tested in parallel on the GPU
//                          against a counting-sorted cell list of the
//                          placed segments, then walked in trial order on
//                          the host against the window's own acceptances.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "scan.hpp"

namespace fs {
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr int kDrawsPerTrial = 11;

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += kGolden);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller, two draws
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

// Trial k: n segment positions (cand, 3n) and the unit quaternion.
void trial(uint64_t seed, long k, const double box[3], int n, double dL,
           double *cand, double qq[4]) {
  Rng rng{seed + (uint64_t)k * kDrawsPerTrial * kGolden};
  double head[3];
  for (int d = 0; d < 3; ++d) head[d] = rng.uniform() * box[d];
  double nn = 0;
  for (int c = 0; c < 4; ++c) {
    qq[c] = rng.normal();
    nn += qq[c] * qq[c];
  }
  nn = std::sqrt(nn);
  for (int c = 0; c < 4; ++c) qq[c] /= nn;
  // tangent = R(q) e_x (UnitQuaternion.hpp:341-352)
  const double t[3] = {1 - 2 * qq[2] * qq[2] - 2 * qq[3] * qq[3],
                       2 * (qq[1] * qq[2] + qq[0] * qq[3]),
                       2 * (qq[1] * qq[3] - qq[0] * qq[2])};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) cand[3 * i + d] = head[d] + (i * dL) * t[d];
}

// Squared minimum-image distance (VecPeriodicX, spring_link_functions.cpp:
// 131-138), the same operations on host and device.
__host__ __device__ inline double periodic_r2(const double *a, const double *b,
                                              const double L[3]) {
#pragma clang fp contract(off)
  double r2 = 0;
  for (int d = 0; d < 3; ++d) {
    double x = a[d] - b[d];
    x -= L[d] * floor(x / L[d] + 0.5);
    r2 += x * x;
  }
  return r2;
}

// Cell geometry: cells at least min_sep wide, so the 27 cells around a
// point hold every point closer than min_sep.
struct CellGeom {
  double L[3];
  int M[3];
  double cw[3];
  __host__ __device__ void coords(const double *x, int *c) const {
    for (int d = 0; d < 3; ++d) {
      double w = x[d] - L[d] * floor(x[d] / L[d]);
      int ci = (int)(w / cw[d]);
      c[d] = ci < 0 ? 0 : (ci > M[d] - 1 ? M[d] - 1 : ci);
    }
  }
  __host__ __device__ int wrap(int c, int d) const {
    return ((c % M[d]) + M[d]) % M[d];
  }
};

CellGeom make_geom(const double box[3], double min_sep, int max_cells) {
  CellGeom G;
  for (int d = 0; d < 3; ++d) {
    G.L[d] = box[d];
    G.M[d] = std::max(1, (int)(box[d] / std::max(min_sep, 1e-12)));
    G.M[d] = std::min(G.M[d], max_cells);
    G.cw[d] = box[d] / G.M[d];
  }
  return G;
}

// Host grid of placed points (hashed cells).
struct HostGrid {
  CellGeom G;
  const double *pts = nullptr;  // 3 per point, indexed by the stored ids
  std::unordered_map<long, std::vector<int>> cells;
  long key(int a, int b, int c) const {
    return ((long)a * G.M[1] + b) * G.M[2] + c;
  }
  void insert(const double *x, int id) {
    int c[3];
    G.coords(x, c);
    cells[key(c[0], c[1], c[2])].push_back(id);
  }
  bool overlaps(const double *x, double ms2) const {
    int c[3];
    G.coords(x, c);
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int e = -1; e <= 1; ++e) {
          auto it = cells.find(key(G.wrap(c[0] + a, 0), G.wrap(c[1] + b, 1),
                                   G.wrap(c[2] + e, 2)));
          if (it == cells.end()) continue;
          for (int j : it->second)
            if (periodic_r2(x, pts + 3 * j, G.L) < ms2) return true;
        }
    return false;
  }
};

// ---- device: cell list of the placed segments and the window test ----

__global__ void k_pl_fill(int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_pl_bin(int np, const double *__restrict__ pts, CellGeom G,
                         int *__restrict__ cell, int *__restrict__ cnt,
                         int *__restrict__ slot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int c[3];
  G.coords(pts + 3 * i, c);
  const int k = (c[0] * G.M[1] + c[1]) * G.M[2] + c[2];
  cell[i] = k;
  slot[i] = atomicAdd(&cnt[k], 1);
}

__global__ void k_pl_scatter(int np, const int *__restrict__ cell,
                             const int *__restrict__ slot,
                             const int *__restrict__ start,
                             int *__restrict__ ids) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) ids[start[cell[i]] + slot[i]] = i;
}

// flag[t] = 1 when any segment of trial t (n segments) is closer than
// min_sep to a placed segment. Stores of 1 only: concurrent writers agree.
__global__ void k_pl_test(int nseg, int n, const double *__restrict__ cand,
                          CellGeom G, double ms2,
                          const int *__restrict__ start,
                          const int *__restrict__ ids,
                          const double *__restrict__ pts,
                          int *__restrict__ flag) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const double x[3] = {cand[3 * s], cand[3 * s + 1], cand[3 * s + 2]};
  int c[3];
  G.coords(x, c);
  bool hit = false;
  for (int a = -1; a <= 1 && !hit; ++a)
    for (int b = -1; b <= 1 && !hit; ++b)
      for (int e = -1; e <= 1 && !hit; ++e) {
        const int k = (G.wrap(c[0] + a, 0) * G.M[1] + G.wrap(c[1] + b, 1)) *
                          G.M[2] +
                      G.wrap(c[2] + e, 2);
        for (int j = start[k]; j < start[k + 1]; ++j)
          if (periodic_r2(x, pts + 3 * (size_t)ids[j], G.L) < ms2) {
            hit = true;
            break;
          }
      }
  if (hit) flag[s / n] = 1;
}

}  // namespace
}  // namespace fs

using namespace fs;

extern "C" int fs_make_filaments(const double box[3], int nfil, const int *nw,
                                 double dL, double min_sep,
                                 unsigned long long seed, int max_tries,
                                 double *X, double *q) {
  FS_CHECK(box && nw && X && q && nfil >= 1 && dL > 0 && min_sep >= 0 &&
               max_tries >= 1,
           FS_ERR_ARG, "bad argument");
  for (int f = 0; f < nfil; ++f)
    FS_CHECK(nw[f] >= 1, FS_ERR_ARG, "filament %d has %d segments", f, nw[f]);
  HostGrid H;
  H.G = make_geom(box, min_sep, 4096);
  H.pts = X;
  const double ms2 = min_sep * min_sep;
  long k = 0;  // trials drawn so far
  int g0 = 0;
  std::vector<double> cand;
  for (int f = 0; f < nfil; ++f) {
    const int n = nw[f];
    cand.resize(3 * (size_t)n);
    double qq[4];
    bool ok = false;
    for (int tr = 0; tr < max_tries && !ok; ++tr, ++k) {
      trial(seed, k, box, n, dL, cand.data(), qq);
      ok = true;
      for (int i = 0; i < n && ok; ++i)
        if (H.overlaps(&cand[3 * i], ms2)) ok = false;
    }
    FS_CHECK(ok, FS_ERR_ARG,
             "could not place filament %d without overlap in %d tries", f,
             max_tries);
    for (int i = 0; i < n; ++i) {
      const int g = g0 + i;
      for (int d = 0; d < 3; ++d) X[3 * g + d] = cand[3 * i + d];
      for (int c = 0; c < 4; ++c) q[4 * g + c] = qq[c];
      H.insert(&X[3 * g], g);
    }
    g0 += n;
  }
  return FS_OK;
}

extern "C" int fs_make_filaments_gpu(const double box[3], int nfil,
                                     const int *nw, double dL, double min_sep,
                                     unsigned long long seed, int max_tries,
                                     double *X, double *q) {
  FS_CHECK(box && nw && X && q && nfil >= 1 && dL > 0 && min_sep >= 0 &&
               max_tries >= 1,
           FS_ERR_ARG, "bad argument");
  long npt = 0;
  int maxnw = 0;
  for (int f = 0; f < nfil; ++f) {
    FS_CHECK(nw[f] >= 1, FS_ERR_ARG, "filament %d has %d segments", f, nw[f]);
    npt += nw[f];
    maxnw = std::max(maxnw, nw[f]);
  }
  FS_CHECK(npt < (1L << 30), FS_ERR_SIZE, "too many segments");
  // at most 400 cells per axis: the cell-count scan stays under 64Mi cells
  const CellGeom G = make_geom(box, min_sep, 400);
  const long ncell = (long)G.M[0] * G.M[1] * G.M[2];
  const double ms2 = min_sep * min_sep;
  const int K = 4096;  // trials per window
  hipStream_t st = nullptr;
  dbuf pts, cell, slot, cnt, start, ids, tmp, cand, flag;
  struct Guard {
    dbuf *b[9];
    hipStream_t *st;
    ~Guard() {
      if (*st) {
        (void)hipStreamSynchronize(*st);
        (void)hipStreamDestroy(*st);
      }
      for (dbuf *p : b) p->release();
    }
  } guard{{&pts, &cell, &slot, &cnt, &start, &ids, &tmp, &cand, &flag}, &st};
  FS_HIP(hipStreamCreate(&st));
  FS_TRY(pts.alloc(sizeof(double) * 3 * (size_t)npt));
  FS_TRY(cell.alloc(sizeof(int) * (size_t)npt));
  FS_TRY(slot.alloc(sizeof(int) * (size_t)npt));
  FS_TRY(ids.alloc(sizeof(int) * (size_t)npt));
  FS_TRY(cnt.alloc(sizeof(int) * (size_t)ncell));
  FS_TRY(start.alloc(sizeof(int) * (size_t)(ncell + 1)));
  FS_TRY(tmp.alloc(sizeof(int) * (size_t)(ncell / kScanBlock + 4)));
  FS_TRY(cand.alloc(sizeof(double) * 3 * (size_t)K * maxnw));
  FS_TRY(flag.alloc(sizeof(int) * K));
  // empty cell list
  hipLaunchKernelGGL(k_pl_fill, dim3(ceil_div(ncell + 1, 256)), dim3(256), 0,
                     st, start.as<int>(), (int)ncell + 1);
  FS_LAUNCHED();

  std::vector<double> hc(3 * (size_t)K * maxnw), hq(4 * (size_t)K);
  std::vector<int> hflag(K);
  HostGrid W;  // segments placed in the current window
  W.G = G;
  W.pts = X;
  long k = 0, kf = 0;  // next trial; first trial of filament f
  int f = 0, g0 = 0, placed = 0;  // placed: segments in the device list
  while (f < nfil) {
    const int n = nw[f];
    // the window's trials, all shaped for filament f's segment count
    for (int j = 0; j < K; ++j)
      trial(seed, k + j, box, n, dL, &hc[3 * (size_t)j * n], &hq[4 * j]);
    FS_HIP(hipMemcpyAsync(cand.p, hc.data(), sizeof(double) * 3 * K * n,
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pl_fill, dim3(ceil_div(K, 256)), dim3(256), 0, st,
                       flag.as<int>(), K);
    if (placed > 0)
      hipLaunchKernelGGL(k_pl_test, dim3(ceil_div((long)K * n, 256)), dim3(256),
                         0, st, K * n, n, cand.as<double>(), G, ms2,
                         start.as<int>(), ids.as<int>(), pts.as<double>(),
                         flag.as<int>());
    FS_LAUNCHED();
    FS_HIP(hipMemcpyAsync(hflag.data(), flag.p, sizeof(int) * K,
                          hipMemcpyDeviceToHost, st));
    FS_HIP(hipStreamSynchronize(st));
    // walk the window in trial order, as the sequential loop would
    W.cells.clear();
    const int gw = g0;  // first segment placed in this window
    int j = 0;
    for (; j < K && f < nfil; ++j) {
      FS_CHECK(k + j - kf < max_tries, FS_ERR_ARG,
               "could not place filament %d without overlap in %d tries", f,
               max_tries);
      if (hflag[j]) continue;
      const double *cj = &hc[3 * (size_t)j * n];
      bool ok = true;
      for (int i = 0; i < n && ok; ++i)
        if (W.overlaps(cj + 3 * i, ms2)) ok = false;
      if (!ok) continue;
      for (int i = 0; i < n; ++i) {
        const int g = g0 + i;
        for (int d = 0; d < 3; ++d) X[3 * g + d] = cj[3 * i + d];
        for (int c = 0; c < 4; ++c) q[4 * g + c] = hq[4 * j + c];
        W.insert(&X[3 * g], g);
      }
      g0 += n;
      ++f;
      kf = k + j + 1;
      if (f < nfil && nw[f] != n) {  // later trials have the wrong shape
        ++j;
        break;
      }
    }
    k += j;
    // add the window's segments to the device cell list (rebuilt whole)
    if (g0 > gw) {
      FS_HIP(hipMemcpyAsync(pts.as<double>() + 3 * (size_t)gw, X + 3 * (size_t)gw,
                            sizeof(double) * 3 * (size_t)(g0 - gw),
                            hipMemcpyHostToDevice, st));
      placed = g0;
      hipLaunchKernelGGL(k_pl_fill, dim3(ceil_div(ncell, 256)), dim3(256), 0,
                         st, cnt.as<int>(), (int)ncell);
      hipLaunchKernelGGL(k_pl_bin, dim3(ceil_div(placed, 256)), dim3(256), 0,
                         st, placed, pts.as<double>(), G, cell.as<int>(),
                         cnt.as<int>(), slot.as<int>());
      FS_TRY(device_scan(cnt.as<int>(), (int)ncell, start.as<int>(),
                         tmp.as<int>(), st));
      hipLaunchKernelGGL(k_pl_scatter, dim3(ceil_div(placed, 256)), dim3(256),
                         0, st, placed, cell.as<int>(), slot.as<int>(),
                         start.as<int>(), ids.as<int>());
      FS_LAUNCHED();
    }
  }
  FS_HIP(hipStreamSynchronize(st));
  return FS_OK;
}
