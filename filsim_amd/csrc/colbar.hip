// colbar.hip -- linked-list collision barrier for gfx950.
//
// Replaces struct LinkedList (CollisionBarrierFilament.hpp:19-57).
//   link   LinkedList::link (CollisionBarrierFilament.cpp:370-433): the same
//          trunc-based wrap and (int) binning per segment, then a counting
//          sort by cell; inside a cell segments are kept in descending index
//          order, which is the order of the reference's LIFO head/list
//          chain, so head[] and list[] are reproduced exactly.
//   apply  apply_collision_barrier (:24-271): one thread per segment walks
//          the 27 cells around the cell it was linked into (the reference's
//          own cell + 13-cell half stencil, seen from both ends of each
//          pair) and sums the capped quartic repulsion on itself. Pair
//          forces are exactly antisymmetric, so the totals equal the
//          reference's up to summation order. The minimum squared distance
//          over all visited pairs is reduced exactly (bit-pattern atomicMin).
// The cell list is built once per timestep and reused stale by every
// Broyden iteration (main.cpp:394 vs :565), as in the reference.
// Requires M_X, M_Y, M_Z >= 3 so the 27-cell stencil has no repeated cell
// (the reference's half stencil double-counts otherwise).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "colbar_impl.hpp"
#include "common.hpp"
#include "scan.hpp"

namespace fs {

__global__ void k_cb_fill(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Cell of one unwrapped position (CollisionBarrierFilament.cpp:393-395,
// 407-409); contraction off so every product rounds as in the reference.
// A coordinate that is not finite or lies 10^9 boxes out (a diverged
// iterate; the reference prints "segfault imminent", :398-404) is binned
// into cell 0 of that axis, so no conversion leaves the int range.
__device__ int cb_cell(const double *x, CbGeom g) {
#pragma clang fp contract(off)
  const int large_number = 100000;
  int c[3];
  const double L[3] = {g.Lx, g.Ly, g.Lz};
  const int M[3] = {g.MX, g.MY, g.MZ};
  for (int d = 0; d < 3; ++d) {
    double xr = x[d];
    if (!(fabs(xr / L[d]) < 1e9)) {
      c[d] = 0;
      continue;
    }
    xr = xr - L[d] * ((double)((int)(xr / L[d] + large_number) - large_number));
    int ci = (int)((xr / L[d]) * (double)M[d]);
    if (ci >= M[d]) ci = M[d] - 1;
    if (ci < 0) ci = 0;
    c[d] = ci;
  }
  return c[0] + c[1] * g.MX + c[2] * g.MX * g.MY;
}

__global__ void k_cb_bin(int np, const double *__restrict__ Xs, CbGeom g,
                         int *__restrict__ cell, int *__restrict__ cnt,
                         int *__restrict__ slot) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  int c = cb_cell(Xs + 3 * s, g);
  cell[s] = c;
  slot[s] = atomicAdd(&cnt[c], 1);
}

__global__ void k_cb_scatter(int np, const int *__restrict__ cell,
                             const int *__restrict__ slot,
                             const int *__restrict__ start,
                             int *__restrict__ ids) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  ids[start[cell[s]] + slot[s]] = s;
}

// Descending order inside each cell = LIFO chain order of the reference.
__global__ void k_cb_sort(int ncell, const int *__restrict__ start,
                          int *__restrict__ ids) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int lo = start[c], hi = start[c + 1];
  for (int i = lo + 1; i < hi; ++i) {
    int v = ids[i];
    int j = i - 1;
    while (j >= lo && ids[j] < v) {
      ids[j + 1] = ids[j];
      --j;
    }
    ids[j + 1] = v;
  }
}

// Pair mask of the barrier variants: apply_collision_barrier_not_swimmer
// (CollisionBarrierFilament.cpp:495, 585: both or neither in filament 0)
// and apply_collision_barrier_not_test_particles (:754, 840: neither of
// the last two segments). A masked pair is not visited at all (no force,
// no minimum distance), as in the reference.
__device__ __forceinline__ bool cb_mask_in(const CbConst &c, int i, int j,
                                           int ifil, int jfil) {
  if (c.mask == FS_BARRIER_NOT_SWIMMER) return (ifil == 0) == (jfil == 0);
  if (c.mask == FS_BARRIER_NOT_TEST_PARTICLES)
    return i < c.np - 2 && j < c.np - 2;
  return true;
}

// One pair seen from segment i (CollisionBarrierFilament.cpp:80-150).
// Returns true when the pair is not an excluded neighbour.
__device__ __forceinline__ bool cb_pair(const double *yi, const double *yj,
                                        int ifil, int iseg, int jfil, int jseg,
                                        const CbConst &c, double &rsq,
                                        double f[3]) {
#pragma clang fp contract(off)
  f[0] = f[1] = f[2] = 0.0;
  if (!(!(jfil == ifil && abs(jseg - iseg) == 1) || c.nt < c.ifs)) return false;
  double xij = yi[0] - yj[0], yij = yi[1] - yj[1], zij = yi[2] - yj[2];
  xij = xij - c.Lx * floor(xij / c.Lx + 0.5);
  yij = yij - c.Ly * floor(yij / c.Ly + 0.5);
  zij = zij - c.Lz * floor(zij / c.Lz + 0.5);
  double rijsq = xij * xij + yij * yij + zij * zij;
  rsq = rijsq;
  if (rijsq < c.chiRnm2) {
    if (rijsq < c.cap2) {
      double rij = sqrt(rijsq);
      xij = xij / rij * c.cap;
      yij = yij / rij * c.cap;
      zij = zij / rij * c.cap;
      rijsq = c.cap2;
    }
    double prefac = c.chiRnm2 - rijsq;
    prefac /= (c.chiRnm2 - c.Rnm2);
    prefac *= prefac;
    prefac *= prefac;
    prefac *= c.FS;
    f[0] = prefac * xij;
    f[1] = prefac * yij;
    f[2] = prefac * zij;
  }
  return true;
}

__device__ __forceinline__ int cb_wrap(int v, int m) {
  v %= m;
  return v < 0 ? v + m : v;
}

// mode 0: F[i] += sum of pair forces, min distance; mode 1: count pairs
// (i<j) inside the cutoff into *npairs and write them when room.
// kCbLanes lanes per segment split its 27 cells (lane q: cells q,
// q + kCbLanes, ..., in the reference's dz, dy, dx order) and combine their
// partial forces pairwise by xor distance 1, 2 (deterministic); more
// wavefronts in flight for the dependent cell/id/position loads than one
// thread per segment.
constexpr int kCbLanes = 4;

__global__ void __launch_bounds__(256)
    k_cb_apply(const int *__restrict__ act, int np,
               const double *__restrict__ Y, CbGeom g, CbConst c,
               const int *__restrict__ cell, const int *__restrict__ start,
               const int *__restrict__ ids, const int *__restrict__ sfil,
               const int *__restrict__ sseg, double *__restrict__ F,
               unsigned long long *__restrict__ min_bits, int mode,
               int *__restrict__ npairs, int max_pairs,
               int *__restrict__ pairs) {
  if (act && !*act) return;  // batched sweep: suspension converged
  __shared__ double s_min[256];
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt / kCbLanes, q = gt % kCbLanes;
  double mymin = 100000.0;
  double fx = 0, fy = 0, fz = 0;
  if (i < np) {
    const double yi[3] = {Y[3 * i], Y[3 * i + 1], Y[3 * i + 2]};
    const int ifil = sfil[i], iseg = sseg[i];
    const int ci = cell[i];
    const int cx = ci % g.MX, cy = (ci / g.MX) % g.MY, cz = ci / (g.MX * g.MY);
    // the lane's cell ranges first, all loads in flight together (most
    // cells are empty), then the pairs in the same cell order
    constexpr int kCells = (27 + kCbLanes - 1) / kCbLanes;
    int lo[kCells], hi[kCells];
#pragma unroll
    for (int t = 0; t < kCells; ++t) {
      const int k = q + kCbLanes * t;
      lo[t] = hi[t] = 0;
      if (k < 27) {
        const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
        const int nc = cb_wrap(cx + dx, g.MX) +
                       g.MX * (cb_wrap(cy + dy, g.MY) + g.MY * cb_wrap(cz + dz, g.MZ));
        lo[t] = start[nc];
        hi[t] = start[nc + 1];
      }
    }
#pragma unroll
    for (int t = 0; t < kCells; ++t) {
      for (int e = lo[t]; e < hi[t]; ++e) {
        int j = ids[e];
        if (j == i) continue;
        const int jfil = sfil[j];
        if (!cb_mask_in(c, i, j, ifil, jfil)) continue;
        double yj[3] = {Y[3 * j], Y[3 * j + 1], Y[3 * j + 2]};
        double rsq, f[3];
        if (!cb_pair(yi, yj, ifil, iseg, jfil, sseg[j], c, rsq, f))
          continue;
        mymin = fmin(mymin, rsq);
        if (rsq < c.chiRnm2) {
          if (mode == 0) {
            fx += f[0];
            fy += f[1];
            fz += f[2];
          } else if (i < j) {
            int kk = atomicAdd(npairs, 1);
            if (kk < max_pairs) {
              pairs[2 * kk] = i;
              pairs[2 * kk + 1] = j;
            }
          }
        }
      }
    }
  }
  // every lane of the wavefront takes part in the shuffles
#pragma unroll
  for (int m = 1; m < kCbLanes; m <<= 1) {
    fx += __shfl_xor(fx, m, 64);
    fy += __shfl_xor(fy, m, 64);
    fz += __shfl_xor(fz, m, 64);
  }
  if (i < np && q == 0 && mode == 0) {
    F[3 * i] += fx;
    F[3 * i + 1] += fy;
    F[3 * i + 2] += fz;
  }
  s_min[threadIdx.x] = mymin;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && mode == 0) {
    // positive doubles order like their bit patterns
    atomicMin(min_bits, (unsigned long long)__double_as_longlong(s_min[0]));
  }
}

__global__ void k_cb_init_min(const int *__restrict__ act,
                              unsigned long long *m, unsigned long long v) {
  if (act && !*act) return;
  *m = v;
}

// Cell grid of LinkedList::initialise (CollisionBarrierFilament.cpp:
// 274-289): M = int(L / 2.5a) cells per axis.
int ColbarImpl::geometry(const fs_colbar_params &pp, int n) {
  p = pp;
  np = n;
  FS_CHECK(n >= 0, FS_ERR_ARG, "np < 0");
  FS_CHECK(p.box[0] > 0 && p.box[1] > 0 && p.box[2] > 0 && p.a > 0,
           FS_ERR_ARG, "box and radius must be positive");
  double Rref = 2.5 * p.a;
  g.Lx = p.box[0];
  g.Ly = p.box[1];
  g.Lz = p.box[2];
  FS_CHECK(p.box[0] / Rref < 2048 && p.box[1] / Rref < 2048 &&
               p.box[2] / Rref < 2048,
           FS_ERR_SIZE, "box too large for the cell list (2047 cells per axis)");
  g.MX = (int)(p.box[0] / Rref);
  g.MY = (int)(p.box[1] / Rref);
  g.MZ = (int)(p.box[2] / Rref);
  FS_CHECK(g.MX >= 3 && g.MY >= 3 && g.MZ >= 3, FS_ERR_ARG,
           "box too small for the cell list: %d x %d x %d cells (need >= 3)",
           g.MX, g.MY, g.MZ);
  const long nc = (long)g.MX * g.MY * g.MZ;
  FS_CHECK(nc < (long)kScanBlock * kScanMaxBlocks, FS_ERR_SIZE,
           "%ld barrier cells exceed the cell-count scan", nc);
  ncell = (int)nc;
  return FS_OK;
}

int ColbarImpl::init(const fs_colbar_params &pp, int n, const int *seg_fil,
                     const int *seg_idx) {
  FS_TRY(geometry(pp, n));
  FS_TRY(cell.alloc(sizeof(int) * std::max(n, 1)));
  FS_TRY(slot.alloc(sizeof(int) * std::max(n, 1)));
  FS_TRY(ids.alloc(sizeof(int) * std::max(n, 1)));
  FS_TRY(cnt.alloc(sizeof(int) * ncell));
  FS_TRY(start.alloc(sizeof(int) * (ncell + 1)));
  FS_TRY(scan_tmp.alloc(sizeof(int) * (ncell / kScanBlock + 4)));
  FS_TRY(sfil.alloc(sizeof(int) * std::max(n, 1)));
  FS_TRY(sseg.alloc(sizeof(int) * std::max(n, 1)));
  FS_TRY(minb.alloc(sizeof(unsigned long long) + sizeof(int)));
  if (n > 0) {
    FS_HIP(hipMemcpy(sfil.p, seg_fil, sizeof(int) * n, hipMemcpyHostToDevice));
    FS_HIP(hipMemcpy(sseg.p, seg_idx, sizeof(int) * n, hipMemcpyHostToDevice));
  }
  linked = false;
  return FS_OK;
}

void ColbarImpl::release() {
  cell.release();
  slot.release();
  ids.release();
  cnt.release();
  start.release();
  scan_tmp.release();
  sfil.release();
  sseg.release();
  minb.release();
}

int ColbarImpl::link(const double *Xs, hipStream_t st) {
  FS_CHECK(np == 0 || Xs, FS_ERR_ARG, "null Xs");
  hipLaunchKernelGGL(k_cb_fill, dim3(ceil_div(ncell, 256)), dim3(256), 0, st,
                     cnt.as<int>(), ncell, 0);
  if (np > 0)
    hipLaunchKernelGGL(k_cb_bin, dim3(ceil_div(np, 256)), dim3(256), 0, st, np,
                       Xs, g, cell.as<int>(), cnt.as<int>(), slot.as<int>());
  FS_TRY(device_scan(cnt.as<int>(), ncell, start.as<int>(),
                     scan_tmp.as<int>(), st));
  if (np > 0)
    hipLaunchKernelGGL(k_cb_scatter, dim3(ceil_div(np, 256)), dim3(256), 0, st,
                       np, cell.as<int>(), slot.as<int>(), start.as<int>(),
                       ids.as<int>());
  hipLaunchKernelGGL(k_cb_sort, dim3(ceil_div(ncell, 256)), dim3(256), 0, st,
                     ncell, start.as<int>(), ids.as<int>());
  FS_LAUNCHED();
  linked = true;
  return FS_OK;
}

CbConst ColbarImpl::consts(int nt) const {
  CbConst c;
  c.Lx = p.box[0];
  c.Ly = p.box[1];
  c.Lz = p.box[2];
  double Rnm = 2 * p.a;
  c.Rnm2 = Rnm * Rnm;
  c.chiRnm2 = 1.21 * c.Rnm2;
  c.FS = p.fs_over_2a;
  c.cap = p.cap;
  c.cap2 = p.cap * p.cap;
  c.nt = nt;
  c.ifs = p.initial_friction_steps;
  c.mask = mask;
  c.np = np;
  return c;
}

int ColbarImpl::apply(const double *Y, int nt, double *F, double *min_dist2,
                      hipStream_t st) {
  FS_CHECK(linked, FS_ERR_STATE, "apply before link");
  FS_CHECK(np == 0 || (Y && F), FS_ERR_ARG, "null array");
  double init = 100000.0;
  unsigned long long ib;
  memcpy(&ib, &init, sizeof ib);
  hipLaunchKernelGGL(k_cb_init_min, dim3(1), dim3(1), 0, st, act,
                     minb.as<unsigned long long>(), ib);
  if (np > 0)
    hipLaunchKernelGGL(k_cb_apply, dim3(ceil_div((long)np * kCbLanes, 256)), dim3(256), 0, st,
                       act, np, Y, g, consts(nt), cell.as<int>(), start.as<int>(),
                       ids.as<int>(), sfil.as<int>(), sseg.as<int>(), F,
                       minb.as<unsigned long long>(), 0, (int *)nullptr, 0,
                       (int *)nullptr);
  FS_LAUNCHED();
  if (min_dist2) {
    unsigned long long hb;
    FS_HIP(hipMemcpyAsync(&hb, minb.p, sizeof hb, hipMemcpyDeviceToHost, st));
    FS_HIP(hipStreamSynchronize(st));
    memcpy(min_dist2, &hb, sizeof hb);
  }
  return FS_OK;
}

int ColbarImpl::get_lists(int *head, int *list) {
  FS_CHECK(linked, FS_ERR_STATE, "get_lists before link");
  FS_HIP(hipDeviceSynchronize());
  std::vector<int> hs(ncell + 1), hid(std::max(np, 1));
  FS_HIP(hipMemcpy(hs.data(), start.p, sizeof(int) * (ncell + 1),
                   hipMemcpyDeviceToHost));
  if (np > 0)
    FS_HIP(hipMemcpy(hid.data(), ids.p, sizeof(int) * np, hipMemcpyDeviceToHost));
  for (int c = 0; c < ncell; ++c) {
    head[c] = (hs[c] < hs[c + 1]) ? hid[hs[c]] : -1;
    for (int e = hs[c]; e < hs[c + 1]; ++e)
      list[hid[e]] = (e + 1 < hs[c + 1]) ? hid[e + 1] : -1;
  }
  return FS_OK;
}

int ColbarImpl::pairs(const double *Y, int nt, int max_pairs, int *out,
                      int *npairs) {
  FS_CHECK(linked, FS_ERR_STATE, "pairs before link");
  FS_CHECK(max_pairs >= 0 && npairs && (max_pairs == 0 || out), FS_ERR_ARG,
           "bad pair buffer");
  dbuf d;
  FS_TRY(d.alloc(sizeof(int) * (2 * (size_t)max_pairs + 1)));
  int *dn = d.as<int>(), *dp = dn + 1;
  hipLaunchKernelGGL(k_cb_fill, dim3(1), dim3(1), 0, 0, dn, 1, 0);
  if (np > 0)
    hipLaunchKernelGGL(k_cb_apply, dim3(ceil_div((long)np * kCbLanes, 256)), dim3(256), 0, 0,
                       (const int *)nullptr, np,
                       Y, g, consts(nt), cell.as<int>(), start.as<int>(),
                       ids.as<int>(), sfil.as<int>(), sseg.as<int>(),
                       (double *)nullptr, (unsigned long long *)nullptr, 1, dn,
                       max_pairs, dp);
  FS_LAUNCHED();
  int n = 0;
  FS_HIP(hipMemcpy(&n, dn, sizeof(int), hipMemcpyDeviceToHost));
  *npairs = n;
  int m = std::min(n, max_pairs);
  std::vector<int> h(2 * (size_t)std::max(m, 1));
  if (m > 0)
    FS_HIP(hipMemcpy(h.data(), dp, sizeof(int) * 2 * m, hipMemcpyDeviceToHost));
  std::vector<std::pair<int, int>> v(m);
  for (int k = 0; k < m; ++k) v[k] = {h[2 * k], h[2 * k + 1]};
  std::sort(v.begin(), v.end());
  for (int k = 0; k < m; ++k) {
    out[2 * k] = v[k].first;
    out[2 * k + 1] = v[k].second;
  }
  d.release();
  return FS_OK;
}

// ---------------------------------------------------------------------
// No-allocation barrier: link + apply with the cell list in a file-scope
// device array instead of buffers of a handle. No counterpart in the
// reference; for callers that cannot allocate device memory.
// ---------------------------------------------------------------------
constexpr size_t kCbScratchBytes = size_t(32) << 20;  // per GPU, process life
static __device__ __align__(256) unsigned char scratch[kCbScratchBytes];

__global__ void k_cb_store_min(const unsigned long long *__restrict__ min_bits,
                               double *__restrict__ out) {
  *out = __longlong_as_double((long long)*min_bits);
}

namespace {

constexpr int kCbMaxDevices = 64;
std::mutex g_cb_scratch_mu;
unsigned char *g_cb_scratch[kCbMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current. The array's host-side name is a
// stand-in, not a device address: kernels get only this pointer.
int cb_scratch_ptr(unsigned char **out) {
  int dev = -1;
  FS_HIP(hipGetDevice(&dev));
  FS_CHECK(dev >= 0 && dev < kCbMaxDevices, FS_ERR_ARG,
           "device %d beyond the %d the barrier scratch table holds", dev,
           kCbMaxDevices);
  std::lock_guard<std::mutex> lock(g_cb_scratch_mu);
  if (!g_cb_scratch[dev]) {
    void *ptr = nullptr;
    FS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    FS_CHECK(ptr != nullptr, FS_ERR_HIP,
             "no device address for the barrier scratch on device %d", dev);
    g_cb_scratch[dev] = static_cast<unsigned char *>(ptr);
  }
  *out = g_cb_scratch[dev];
  return FS_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of the scratch areas for ncell cells and np segments; the
// total is returned.
struct CbScratchLayout {
  size_t minb, cnt, start, scan, cell, slot, ids, total;
  CbScratchLayout(long ncell, long np) {
    size_t o = 0;
    minb = o;
    o += align256(sizeof(unsigned long long));
    cnt = o;
    o += align256(sizeof(int) * (size_t)ncell);
    start = o;
    o += align256(sizeof(int) * (size_t)(ncell + 1));
    scan = o;
    o += align256(sizeof(int) * (size_t)(ncell / kScanBlock + 4));
    cell = o;
    o += align256(sizeof(int) * (size_t)std::max(np, 1L));
    slot = o;
    o += align256(sizeof(int) * (size_t)std::max(np, 1L));
    ids = o;
    o += align256(sizeof(int) * (size_t)std::max(np, 1L));
    total = o;
  }
};

}  // namespace

int colbar_barrier_noalloc(const fs_colbar_params *pp, int np,
                           const int *seg_fil, const int *seg_idx,
                           const double *Xs, const double *Y, int nt, int mask,
                           double *F, double *min_dist2, hipStream_t st) {
  FS_CHECK(pp, FS_ERR_ARG, "null parameters");
  FS_CHECK(mask == FS_BARRIER_ALL || mask == FS_BARRIER_NOT_SWIMMER ||
               mask == FS_BARRIER_NOT_TEST_PARTICLES,
           FS_ERR_ARG, "unknown barrier mask %d", mask);
  ColbarImpl cb;  // geometry and constants only: its buffers stay empty
  FS_TRY(cb.geometry(*pp, np));
  cb.mask = mask;
  FS_CHECK(np == 0 || (seg_fil && seg_idx && Xs && Y && F), FS_ERR_ARG,
           "null array");
  const CbScratchLayout L(cb.ncell, np);
  FS_CHECK(L.total <= kCbScratchBytes, FS_ERR_SIZE,
           "%d cells and %d segments need %zu bytes of barrier scratch; the "
           "no-allocation entry point has %zu",
           cb.ncell, np, L.total, kCbScratchBytes);
  unsigned char *base = nullptr;
  FS_TRY(cb_scratch_ptr(&base));
  unsigned long long *minb = reinterpret_cast<unsigned long long *>(base + L.minb);
  int *cnt = reinterpret_cast<int *>(base + L.cnt);
  int *start = reinterpret_cast<int *>(base + L.start);
  int *scan = reinterpret_cast<int *>(base + L.scan);
  int *cell = reinterpret_cast<int *>(base + L.cell);
  int *slot = reinterpret_cast<int *>(base + L.slot);
  int *ids = reinterpret_cast<int *>(base + L.ids);
  // link: the cell counts start from zero on every call (fill kernel, so a
  // captured graph replays it)
  hipLaunchKernelGGL(k_cb_fill, dim3(ceil_div(cb.ncell, 256)), dim3(256), 0,
                     st, cnt, cb.ncell, 0);
  if (np > 0)
    hipLaunchKernelGGL(k_cb_bin, dim3(ceil_div(np, 256)), dim3(256), 0, st, np,
                       Xs, cb.g, cell, cnt, slot);
  FS_TRY(device_scan(cnt, cb.ncell, start, scan, st));
  if (np > 0)
    hipLaunchKernelGGL(k_cb_scatter, dim3(ceil_div(np, 256)), dim3(256), 0, st,
                       np, cell, slot, start, ids);
  hipLaunchKernelGGL(k_cb_sort, dim3(ceil_div(cb.ncell, 256)), dim3(256), 0,
                     st, cb.ncell, start, ids);
  // apply
  double init = 100000.0;
  unsigned long long ib;
  memcpy(&ib, &init, sizeof ib);
  hipLaunchKernelGGL(k_cb_init_min, dim3(1), dim3(1), 0, st,
                     (const int *)nullptr, minb, ib);
  if (np > 0)
    hipLaunchKernelGGL(k_cb_apply, dim3(ceil_div((long)np * kCbLanes, 256)),
                       dim3(256), 0, st, (const int *)nullptr, np, Y, cb.g,
                       cb.consts(nt), cell, start, ids, seg_fil, seg_idx, F,
                       minb, 0, (int *)nullptr, 0, (int *)nullptr);
  if (min_dist2)
    hipLaunchKernelGGL(k_cb_store_min, dim3(1), dim3(1), 0, st, minb,
                       min_dist2);
  FS_LAUNCHED();
  return FS_OK;
}

}  // namespace fs

using fs::ColbarImpl;

extern "C" int fs_colbar_barrier_noalloc(const fs_colbar_params *p, int np,
                                         const int *seg_fil,
                                         const int *seg_idx, const double *Xs,
                                         const double *Y, int nt, int mask,
                                         double *F, double *min_dist2,
                                         void *stream) {
  return fs::colbar_barrier_noalloc(p, np, seg_fil, seg_idx, Xs, Y, nt, mask,
                                    F, min_dist2, (hipStream_t)stream);
}

struct fs_colbar {
  ColbarImpl impl;
};

extern "C" int fs_colbar_create(const fs_colbar_params *p, int np,
                                const int *seg_fil, const int *seg_idx,
                                fs_colbar **out) {
  FS_CHECK(p && out && (np == 0 || (seg_fil && seg_idx)), FS_ERR_ARG,
           "null argument");
  *out = nullptr;
  fs_colbar *h = new fs_colbar();
  int rc = h->impl.init(*p, np, seg_fil, seg_idx);
  if (rc != FS_OK) {
    h->impl.release();
    delete h;
    return rc;
  }
  *out = h;
  return FS_OK;
}

extern "C" int fs_colbar_destroy(fs_colbar *h) {
  if (!h) return FS_OK;
  (void)hipDeviceSynchronize();
  h->impl.release();
  delete h;
  return FS_OK;
}

extern "C" int fs_colbar_set_mask(fs_colbar *h, int mask) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(mask == FS_BARRIER_ALL || mask == FS_BARRIER_NOT_SWIMMER ||
               mask == FS_BARRIER_NOT_TEST_PARTICLES,
           FS_ERR_ARG, "unknown barrier mask %d", mask);
  h->impl.mask = mask;
  return FS_OK;
}

extern "C" int fs_colbar_dims(fs_colbar *h, int *m3) {
  FS_CHECK(h && m3, FS_ERR_ARG, "null argument");
  m3[0] = h->impl.g.MX;
  m3[1] = h->impl.g.MY;
  m3[2] = h->impl.g.MZ;
  return FS_OK;
}

extern "C" int fs_colbar_link(fs_colbar *h, const double *Xs, void *stream) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.link(Xs, (hipStream_t)stream);
}

extern "C" int fs_colbar_apply(fs_colbar *h, const double *Y, int nt,
                               double *F, double *min_dist2, void *stream) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.apply(Y, nt, F, min_dist2, (hipStream_t)stream);
}

extern "C" int fs_colbar_get_lists(fs_colbar *h, int *head, int *list) {
  FS_CHECK(h && head && list, FS_ERR_ARG, "null argument");
  return h->impl.get_lists(head, list);
}

extern "C" int fs_colbar_pairs(fs_colbar *h, const double *Y, int nt,
                               int max_pairs, int *pairs, int *npairs) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.pairs(Y, nt, max_pairs, pairs, npairs);
}

extern "C" int fs_colbar_link_host(fs_colbar *h, const double *Xs) {
  FS_CHECK(h && (h->impl.np == 0 || Xs), FS_ERR_ARG, "null argument");
  int np = h->impl.np;
  if (np == 0) return h->impl.link(nullptr, nullptr);
  fs::dbuf d;
  FS_TRY(d.alloc(sizeof(double) * 3 * (size_t)np));
  int rc = FS_OK;
  if (hipMemcpy(d.p, Xs, sizeof(double) * 3 * np, hipMemcpyHostToDevice) !=
      hipSuccess) {
    fs::set_error("host to device copy failed");
    rc = FS_ERR_HIP;
  }
  if (rc == FS_OK) rc = h->impl.link(d.as<double>(), nullptr);
  if (rc == FS_OK && hipDeviceSynchronize() != hipSuccess) {
    fs::set_error("link failed");
    rc = FS_ERR_HIP;
  }
  d.release();
  return rc;
}

extern "C" int fs_colbar_apply_host(fs_colbar *h, const double *Y, int nt,
                                    double *F, double *min_dist2) {
  FS_CHECK(h && (h->impl.np == 0 || (Y && F)), FS_ERR_ARG, "null argument");
  int np = h->impl.np;
  fs::dbuf d;
  FS_TRY(d.alloc(sizeof(double) * 6 * (size_t)std::max(np, 1)));
  double *dY = d.as<double>(), *dF = dY + 3 * np;
  size_t b = sizeof(double) * 3 * np;
  int rc = FS_OK;
  if (np > 0 && (hipMemcpy(dY, Y, b, hipMemcpyHostToDevice) != hipSuccess ||
                 hipMemcpy(dF, F, b, hipMemcpyHostToDevice) != hipSuccess)) {
    fs::set_error("host to device copy failed");
    rc = FS_ERR_HIP;
  }
  if (rc == FS_OK) rc = h->impl.apply(dY, nt, dF, min_dist2, nullptr);
  if (rc == FS_OK && np > 0 &&
      hipMemcpy(F, dF, b, hipMemcpyDeviceToHost) != hipSuccess) {
    fs::set_error("device to host copy failed");
    rc = FS_ERR_HIP;
  }
  d.release();
  return rc;
}
