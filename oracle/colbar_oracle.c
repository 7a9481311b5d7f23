/*
 * colbar_oracle.c -- FP64 CPU restatement of the linked-list collision
 * barrier.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows CollisionBarrierFilament.cpp:
 *   initialise (cell counts)   :274-304
 *   icell / bulkmap            :307-343
 *   link (LIFO head/list)      :370-433
 *   apply_collision_barrier    :24-271 (same traversal and accumulation
 *                              order, so forces match the reference's)
 *   the _not_swimmer (:441-696) and _not_test_particles (:698-945)
 *   variants                   the same walk with a pair mask
 * The MPI cell-range split (:284-289) collapses to one rank owning every
 * cell; the Allreduce MIN/SUM (:250,256) become the local values.
 *
 * Parity pins (tests/test_oracle_colbar.py): the pair set equals a
 * brute-force O(N^2) minimum-image enumeration (or_colbar_pairs_brute);
 * Newton's third law (sum of forces = 0); closed-form pair force on both
 * sides of the distance cap; head/list LIFO order by construction; a masked
 * pair set equals the brute-force set filtered by the mask's rule, and the
 * masked forces equal those of a barrier over the kept pairs alone.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_colbar {
  fs_colbar_params p;
  int np;
  int mask; /* FS_BARRIER_* */
  int M[3];
  int ncell;
  int *map, *head, *list;
  int *sfil, *sseg;
  double *Ftot; /* F_total of the last apply (zero before the first) */
};

/* icell (CollisionBarrierFilament.cpp:307-315) */
static int icell(int MX, int MY, int MZ, int x, int y, int z) {
  double cellix = (double)MX, celliy = (double)MY;
  double beta = fmod((double)(x + MX), (double)MX) +
                fmod((double)(y + MY), (double)MY) * cellix +
                fmod((double)(z + MZ), (double)MZ) * cellix * celliy;
  return (int)beta;
}

/* bulkmap: 13-cell half stencil (CollisionBarrierFilament.cpp:318-343) */
static void bulkmap(int *map, int MX, int MY, int MZ) {
  static const int off[13][3] = {{1, 0, 0},  {1, 1, 0},  {0, 1, 0},
                                 {-1, 1, 0}, {1, 0, -1}, {1, 1, -1},
                                 {0, 1, -1}, {-1, 1, -1}, {1, 0, 1},
                                 {1, 1, 1},  {0, 1, 1},  {-1, 1, 1},
                                 {0, 0, 1}};
  for (int iz = 0; iz < MZ; iz++)
    for (int iy = 0; iy < MY; iy++)
      for (int ix = 0; ix < MX; ix++) {
        int imap = icell(MX, MY, MZ, ix, iy, iz) * 13;
        for (int n = 0; n < 13; ++n)
          map[imap + n] = icell(MX, MY, MZ, ix + off[n][0], iy + off[n][1],
                                iz + off[n][2]);
      }
}

or_colbar *or_colbar_create(const fs_colbar_params *p, int np,
                            const int *seg_fil, const int *seg_idx) {
  or_colbar *h = (or_colbar *)calloc(1, sizeof(or_colbar));
  h->p = *p;
  h->np = np;
  double Rref = 2.5 * p->a;
  for (int d = 0; d < 3; ++d) h->M[d] = (int)(p->box[d] / Rref);
  h->ncell = h->M[0] * h->M[1] * h->M[2];
  h->map = (int *)malloc(sizeof(int) * 13 * h->ncell);
  h->head = (int *)malloc(sizeof(int) * h->ncell);
  h->list = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  for (int c = 0; c < h->ncell; ++c) h->head[c] = -1;
  bulkmap(h->map, h->M[0], h->M[1], h->M[2]);
  h->sfil = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  h->sseg = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  memcpy(h->sfil, seg_fil, sizeof(int) * np);
  memcpy(h->sseg, seg_idx, sizeof(int) * np);
  h->Ftot = (double *)calloc(3 * (size_t)(np > 0 ? np : 1), sizeof(double));
  return h;
}

void or_colbar_destroy(or_colbar *h) {
  if (!h) return;
  free(h->map);
  free(h->head);
  free(h->list);
  free(h->sfil);
  free(h->sseg);
  free(h->Ftot);
  free(h);
}

void or_colbar_dims(or_colbar *h, int *m3) {
  for (int d = 0; d < 3; ++d) m3[d] = h->M[d];
}

/* Cell of one unwrapped position: trunc-based wrap then (int) binning
 * (CollisionBarrierFilament.cpp:393-395, 407-409). A wrapped coordinate that
 * rounds to exactly L would index cell M (out of range in the reference);
 * it is clamped to M-1 here and in the device code. */
static int cell_of(const or_colbar *h, const double *x) {
  const int large_number = 100000;
  int c[3];
  for (int d = 0; d < 3; ++d) {
    double L = h->p.box[d];
    double xr = x[d];
    xr = xr - L * ((double)((int)(xr / L + large_number) - large_number));
    int ci = (int)((xr / L) * (double)h->M[d]);
    if (ci >= h->M[d]) ci = h->M[d] - 1;
    if (ci < 0) ci = 0;
    c[d] = ci;
  }
  return c[0] + c[1] * h->M[0] + c[2] * h->M[0] * h->M[1];
}

void or_colbar_link(or_colbar *h, const double *Xs) {
  for (int i = 0; i < h->ncell; i++) h->head[i] = -1;
  for (int g = 0; g < h->np; ++g) {
    int index = cell_of(h, Xs + 3 * g);
    h->list[g] = h->head[index];
    h->head[index] = g;
  }
}

void or_colbar_get_lists(or_colbar *h, int *head, int *list) {
  memcpy(head, h->head, sizeof(int) * h->ncell);
  memcpy(list, h->list, sizeof(int) * h->np);
}

typedef struct {
  double Lx, Ly, Lz, Rnm2, chiRnm2, FS, cap, cap2;
  int nt, ifs;
} pairconst;

void or_colbar_set_mask(or_colbar *h, int mask) { h->mask = mask; }

/* The variants' pair masks, tested before anything else of a pair:
 * apply_collision_barrier_not_swimmer (CollisionBarrierFilament.cpp:495,
 * 585), apply_collision_barrier_not_test_particles (:754, 840). */
static int mask_in(const or_colbar *h, int i, int j) {
  if (h->mask == FS_BARRIER_NOT_SWIMMER)
    return (h->sfil[i] == 0) == (h->sfil[j] == 0);
  if (h->mask == FS_BARRIER_NOT_TEST_PARTICLES)
    return i < h->np - 2 && j < h->np - 2;
  return 1;
}

static pairconst make_pc(const fs_colbar_params *p, int nt) {
  pairconst c;
  c.Lx = p->box[0];
  c.Ly = p->box[1];
  c.Lz = p->box[2];
  double Rnm = 2 * p->a;
  c.Rnm2 = Rnm * Rnm;
  c.chiRnm2 = 1.21 * c.Rnm2;
  c.FS = p->fs_over_2a;
  c.cap = p->cap;
  c.cap2 = p->cap * p->cap;
  c.nt = nt;
  c.ifs = p->initial_friction_steps;
  return c;
}

/* One pair (CollisionBarrierFilament.cpp:80-150): returns 1 if the pair is
 * considered (not an excluded neighbour), writes rijsq and, when inside the
 * cutoff, the force on i in f[3] (else f = 0). */
static int pair_force(const pairconst *c, const double *yi, const double *yj,
                      int ifil, int iseg, int jfil, int jseg, double *rsq_out,
                      double *f) {
  f[0] = f[1] = f[2] = 0.0;
  if (!(!(jfil == ifil && abs(jseg - iseg) == 1) || c->nt < c->ifs)) return 0;
  double xij = yi[0] - yj[0], yij = yi[1] - yj[1], zij = yi[2] - yj[2];
  xij = xij - c->Lx * floor(xij / c->Lx + 0.5);
  yij = yij - c->Ly * floor(yij / c->Ly + 0.5);
  zij = zij - c->Lz * floor(zij / c->Lz + 0.5);
  double rijsq = xij * xij + yij * yij + zij * zij;
  *rsq_out = rijsq;
  if (rijsq < c->chiRnm2) {
    if (rijsq < c->cap2) {
      double rij = sqrt(rijsq);
      xij = xij / rij * c->cap;
      yij = yij / rij * c->cap;
      zij = zij / rij * c->cap;
      rijsq = c->cap2;
    }
    double prefac = c->chiRnm2 - rijsq;
    prefac /= (c->chiRnm2 - c->Rnm2);
    prefac *= prefac;
    prefac *= prefac;
    prefac *= c->FS;
    f[0] = prefac * xij;
    f[1] = prefac * yij;
    f[2] = prefac * zij;
  }
  return 1;
}

double or_colbar_apply(or_colbar *h, const double *Y, int nt, double *F) {
  pairconst c = make_pc(&h->p, nt);
  double *Fn = (double *)calloc(3 * (size_t)(h->np > 0 ? h->np : 1),
                                sizeof(double));
  double min_dist2 = 100000;
  for (int ic = 0; ic < h->ncell; ic++) {
    int i = h->head[ic];
    while (i > -1) {
      double fxi = Fn[3 * i], fyi = Fn[3 * i + 1], fzi = Fn[3 * i + 2];
      /* rest of own cell, then the 13 half-stencil neighbours */
      for (int pass = 0; pass < 14; ++pass) {
        int j = (pass == 0) ? h->list[i] : h->head[h->map[13 * ic + pass - 1]];
        while (j > -1) {
          double rsq, f[3];
          if (mask_in(h, i, j) &&
              pair_force(&c, Y + 3 * i, Y + 3 * j, h->sfil[i], h->sseg[i],
                         h->sfil[j], h->sseg[j], &rsq, f)) {
            if (rsq < min_dist2) min_dist2 = rsq;
            if (rsq < c.chiRnm2) {
              fxi += f[0];
              fyi += f[1];
              fzi += f[2];
              Fn[3 * j] -= f[0];
              Fn[3 * j + 1] -= f[1];
              Fn[3 * j + 2] -= f[2];
            }
          }
          j = h->list[j];
        }
      }
      Fn[3 * i] = fxi;
      Fn[3 * i + 1] = fyi;
      Fn[3 * i + 2] = fzi;
      i = h->list[i];
    }
  }
  /* F_total of CollisionBarrierFilament.cpp:681-686, kept as the reference
   * keeps linked_list.F_total, then added to the FCM forces */
  for (int i = 0; i < 3 * h->np; ++i) {
    h->Ftot[i] = Fn[i];
    F[i] += Fn[i];
  }
  free(Fn);
  return min_dist2;
}

void or_colbar_last_force(or_colbar *h, double *out) {
  memcpy(out, h->Ftot, sizeof(double) * 3 * (size_t)h->np);
}

static int cmp_pair(const void *a, const void *b) {
  const int *x = (const int *)a, *y = (const int *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}

static void push_pair(int *pairs, int max_pairs, int *n, int i, int j) {
  if (*n < max_pairs) {
    pairs[2 * *n] = i < j ? i : j;
    pairs[2 * *n + 1] = i < j ? j : i;
  }
  ++*n;
}

int or_colbar_pairs(or_colbar *h, const double *Y, int nt, int max_pairs,
                    int *pairs) {
  pairconst c = make_pc(&h->p, nt);
  int n = 0;
  for (int ic = 0; ic < h->ncell; ic++) {
    for (int i = h->head[ic]; i > -1; i = h->list[i]) {
      for (int pass = 0; pass < 14; ++pass) {
        int j = (pass == 0) ? h->list[i] : h->head[h->map[13 * ic + pass - 1]];
        for (; j > -1; j = h->list[j]) {
          double rsq, f[3];
          if (mask_in(h, i, j) &&
              pair_force(&c, Y + 3 * i, Y + 3 * j, h->sfil[i], h->sseg[i],
                         h->sfil[j], h->sseg[j], &rsq, f) &&
              rsq < c.chiRnm2)
            push_pair(pairs, max_pairs, &n, i, j);
        }
      }
    }
  }
  if (n <= max_pairs) qsort(pairs, n, 2 * sizeof(int), cmp_pair);
  return n;
}

int or_colbar_pairs_brute(const fs_colbar_params *p, int np, const double *Y,
                          const int *seg_fil, const int *seg_idx, int nt,
                          int max_pairs, int *pairs) {
  pairconst c = make_pc(p, nt);
  int n = 0;
  for (int i = 0; i < np; ++i)
    for (int j = i + 1; j < np; ++j) {
      double rsq, f[3];
      if (pair_force(&c, Y + 3 * i, Y + 3 * j, seg_fil[i], seg_idx[i],
                     seg_fil[j], seg_idx[j], &rsq, f) &&
          rsq < c.chiRnm2)
        push_pair(pairs, max_pairs, &n, i, j);
    }
  if (n <= max_pairs) qsort(pairs, n, 2 * sizeof(int), cmp_pair);
  return n;
}
