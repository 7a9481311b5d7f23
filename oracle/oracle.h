/*
 * oracle.h -- clean-room FP64 CPU restatement of FilSim's FCM mobility step
 * and implicit filament timestep.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker
 * (or the timed CPU baseline). The product path never calls it.
 *
 * It follows the reference loop for loop; each function cites the file:line
 * it restates. Parity pins (what checks each function) are listed in the
 * header of every .c file.
 */
#ifndef FILSIM_ORACLE_H
#define FILSIM_ORACLE_H

#include "../include/filsim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (replaces FFTW 2.1.5; restated from the DFT definition) ---- */
/* In-place complex DFT of length n (any n; fastest for 2^a 3^b 5^c),
 * sign -1 forward, +1 inverse, unnormalised. data: 2n doubles (re,im). */
void or_fft1d(int n, double *data, int sign);
/* 3D real->half-complex forward transform of r[nx][ny][nz] into
 * c[nx][ny][nz/2+1] (interleaved re,im), unnormalised. */
void or_rfft3d(int nx, int ny, int nz, const double *r, double *c);
/* Inverse: half-complex -> real, unnormalised (C2R ignores the imaginary
 * parts of the kz=0 and kz=nz/2 lines after the x/y transforms). */
void or_irfft3d(int nx, int ny, int nz, const double *c, double *r);

/* ---- FCM ---- */
/* Whole mobility solve: Fcm::mobility_solve (FCMfunctions.cpp:175-464). */
void or_fcm_mobility(const fs_fcm_params *p, int np, const double *Y,
                     const double *F, const double *T, double *V, double *W);
/* Pieces, for stage-by-stage checks. grid buffers: 3 x nx*ny*nz. */
void or_fcm_spread(const fs_fcm_params *p, int np, const double *Y,
                   const double *F, const double *T, double *f3);
/* flow_solve on half spectra (3 x nx*ny*(nz/2+1) complex), in place. */
void or_fcm_project(const fs_fcm_params *p, double *c3);
void or_fcm_interp(const fs_fcm_params *p, int np, const double *Y,
                   const double *u3, double *V, double *W);
void or_fcm_friction(const fs_fcm_params *p, int np, const double *F,
                     const double *T, double *V, double *W);
/* Per-stage wall time (seconds) of the last or_fcm_mobility call:
 * gauss+spread, fft forward, flow solve, fft backward, interpolation. */
void or_fcm_last_times(double *t5);
/* Threads of the OpenMP loops (OMP_NUM_THREADS or every core). */
int or_num_threads(void);

/* ---- Collision barrier ---- */
typedef struct or_colbar or_colbar;
or_colbar *or_colbar_create(const fs_colbar_params *p, int np,
                            const int *seg_fil, const int *seg_idx);
void or_colbar_destroy(or_colbar *h);
void or_colbar_dims(or_colbar *h, int *m3);
void or_colbar_link(or_colbar *h, const double *Xs);
void or_colbar_get_lists(or_colbar *h, int *head, int *list);
/* F += barrier forces; returns min_dist2. */
double or_colbar_apply(or_colbar *h, const double *Y, int nt, double *F);
/* linked_list.F_total: the barrier force of the last apply (np x 3). */
void or_colbar_last_force(or_colbar *h, double *out);
/* Pair mask (FS_BARRIER_*) of later apply / pairs calls. */
void or_colbar_set_mask(or_colbar *h, int mask);
/* Pairs within cutoff via the cell list (i<j sorted); returns count
 * (writes at most max_pairs). */
int or_colbar_pairs(or_colbar *h, const double *Y, int nt, int max_pairs,
                    int *pairs);
/* Same set by brute-force O(N^2) minimum-image enumeration. */
int or_colbar_pairs_brute(const fs_colbar_params *p, int np, const double *Y,
                          const int *seg_fil, const int *seg_idx, int nt,
                          int max_pairs, int *pairs);

/* ---- Filament system ---- */
typedef struct or_sim or_sim;
#define OR_NBUCKETS 12
/* Seconds of the last or_sim_timestep(_trial) per profiler bucket
 * (profilers.hpp:24-55): assign F, build Jinv, collision barrier, FCM,
 * error check, solve Jinv, timestep total, FCM spread (make_zero +
 * gaussian_setup + force_distribution), FFT forward, flow solve, FFT
 * backward, particle velocities/rotations. */
void or_sim_bucket_times(or_sim *h, double *out);
or_sim *or_sim_create(const fs_sim_params *p, int nfil, const int *nw,
                      const fs_fil_props *props);
void or_sim_destroy(or_sim *h);
void or_sim_sizes(or_sim *h, int *nfil, int *np, int *nlam);
void or_sim_set_state(or_sim *h, const fs_state_ptrs *s);
void or_sim_get_state(or_sim *h, fs_state_ptrs *s);
/* One whole timestep (main.cpp:284-754); returns Broyden iterations. */
int or_sim_timestep(or_sim *h, int nt);
/* The same without the final step (main_gaitreset.cpp trial passes). */
int or_sim_timestep_trial(or_sim *h, int nt);
/* Barrier of later timesteps: on = 0 skips it; mask: FS_BARRIER_*. */
void or_sim_set_barrier(or_sim *h, int on, int mask);
/* Pieces. */
void or_sim_initial_guess(or_sim *h);
void or_sim_forces(or_sim *h, int nt);
int or_sim_set_spring_links(or_sim *h, int n, const int *links, double k,
                            double len);
void or_sim_fcm_arrays(or_sim *h, double **Y, double **F, double **T,
                       double **V, double **W);
void or_sim_build_jacobian(or_sim *h, int nt);
void or_sim_get_jacobian(or_sim *h, int f, double *out);
void or_sim_get_jinv(or_sim *h, int f, double *out);
void or_sim_jsolve(or_sim *h, const double *b, double *x, int iter);
void or_sim_broyden_update(or_sim *h, const double *E, const double *Enew,
                           const double *update, int iter);
void or_sim_update(or_sim *h, const double *update);
int or_sim_residual(or_sim *h, int nt, const double *V, const double *W,
                    double *err, double *maxabs);
void or_sim_step(or_sim *h, int nt);
void or_sim_mobility(or_sim *h, int nt);
double or_sim_barrier(or_sim *h, int nt, int link);
/* F_total of the sim's last barrier application (zero if none). */
void or_sim_barrier_force(or_sim *h, double *out);

/* ---- Small closed-form helpers exposed for tests ---- */
void or_quat_expm(double *q, const double *v);        /* q <- exp(v) q   */
void or_quat_midpoint(const double *p1, const double *p2, double *out);
void or_quat_tangent(const double *q, double *t);
void or_quat_rotate(const double *q, double *v);
void or_dexpinv(const double *u, const double *w, double *out);
/* General dense inverse with partial pivoting (LU); returns 0 if singular. */
int or_inverse(int n, const double *A, double *Ainv);

#ifdef __cplusplus
}
#endif
#endif
