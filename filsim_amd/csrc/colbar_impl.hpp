// colbar_impl.hpp -- host-side state of one collision barrier (fs_colbar).
#pragma once

#include "common.hpp"

namespace fs {

struct CbGeom {
  double Lx, Ly, Lz;
  int MX, MY, MZ;
};

struct CbConst {
  double Lx, Ly, Lz, Rnm2, chiRnm2, FS, cap, cap2;
  int nt, ifs;
  int mask, np;  // pair mask (FS_BARRIER_*) and segment count
};

struct ColbarImpl {
  fs_colbar_params p{};
  int np = 0;
  // device flag read by the pair-force kernels of apply(): 0 = skip (a
  // converged suspension of a batched sweep); nullptr = always run
  const int *act = nullptr;
  int ncell = 0;
  CbGeom g{};
  bool linked = false;
  int mask = FS_BARRIER_ALL;
  dbuf cell, slot, ids, cnt, start, scan_tmp, sfil, sseg, minb;

  // cell grid and sizes only (no allocation)
  int geometry(const fs_colbar_params &pp, int n);
  int init(const fs_colbar_params &pp, int n, const int *seg_fil,
           const int *seg_idx);
  void release();
  int link(const double *Xs, hipStream_t st);
  int apply(const double *Y, int nt, double *F, double *min_dist2,
            hipStream_t st);
  int get_lists(int *head, int *list);
  int pairs(const double *Y, int nt, int max_pairs, int *out, int *npairs);
  CbConst consts(int nt) const;
};

}  // namespace fs
