/*
 * filsim.h -- C-ABI of the MI355X (gfx950) FilSim hot path.
 *
 * One shared library (libfilsim_amd.so) exports everything below. Every
 * entry point returns an int status: FS_OK (0) on success, a negative FS_ERR_*
 * code otherwise; fs_last_error() gives the message of the last failure on the
 * calling thread. Handles are not thread-safe: one handle, one host thread.
 *
 * Pointer conventions
 *   - "device" arguments are HBM pointers (hipMalloc / torch.cuda tensors).
 *   - "host" arguments are ordinary host memory.
 *   - Per-segment arrays are Np x 3 row-major doubles, the layout of the
 *     reference's double** Y/F/T/V/W blocks (c_array_functions.cpp:12-21,
 *     FCMfunctions.hpp:13-17): &Y[0][0] may be passed directly.
 *   - stream arguments are hipStream_t passed as void* (NULL = null stream).
 *
 * Reference call surface each group replaces is cited per function.
 */
#ifndef FILSIM_H
#define FILSIM_H

#ifdef __cplusplus
extern "C" {
#endif

#define FS_OK 0
#define FS_ERR_ARG (-1)    /* invalid argument or size                  */
#define FS_ERR_HIP (-2)    /* a HIP runtime call failed                 */
#define FS_ERR_SIZE (-3)   /* input larger than the handle was built for */
#define FS_ERR_STATE (-4)  /* call out of order (e.g. apply before link) */

#define FS_FP64 0 /* spread/FFT/projection/interpolation in FP64           */
#define FS_FP32 1 /* the same four stages in FP32 with FP32 accumulation   */

/* ------------------------------------------------------------------ */
/* Library                                                             */
/* ------------------------------------------------------------------ */
const char *fs_last_error(void);
int fs_version(void);
/* Number of visible HIP devices (0 when none); never faults. */
int fs_device_count(int *count);

/* ------------------------------------------------------------------ */
/* FCM mobility (replaces struct Fcm, FCMfunctions.hpp:7-59)            */
/* ------------------------------------------------------------------ */
typedef struct fs_fcm_params {
  int nx, ny, nz;   /* NPTS_X/Y/Z, periodic grid (config.hpp:73-75)         */
  int ngd;          /* Gaussian support points per dimension (config.hpp:76)*/
  double a;         /* segment radius                                       */
  double mu;        /* viscosity                                            */
  double dx;        /* grid spacing (multi_filament_header.hpp:48)          */
  double sigma;     /* force envelope width  (multi_filament_header.hpp:57) */
  double sigma_dip; /* dipole envelope width (multi_filament_header.hpp:47) */
} fs_fcm_params;

typedef struct fs_fcm fs_fcm;

/* Fcm::initialise (FCMfunctions.cpp:10-168): grids, wavenumbers, tables.
 * Sizes must factor as 2^a 3^b 5^c with nx,ny,nz <= 1024 (even). */
int fs_fcm_create(const fs_fcm_params *p, int max_np, int precision,
                  fs_fcm **out);
/* Fcm::free_memory (FCMfunctions.cpp:468-519). */
int fs_fcm_destroy(fs_fcm *h);
/* Fcm::mobility_solve (FCMfunctions.cpp:175-464): V,W from Y,F,T (device). */
int fs_fcm_mobility(fs_fcm *h, int np, const double *Y, const double *F,
                    const double *T, double *V, double *W, void *stream);
/* Same with host arrays (copies in and out; synchronous). */
int fs_fcm_mobility_host(fs_fcm *h, int np, const double *Y, const double *F,
                         const double *T, double *V, double *W);
/* Fcm::friction_mobility_solve (FCMfunctions.cpp:865-880) (device). */
int fs_fcm_friction(fs_fcm *h, int np, const double *F, const double *T,
                    double *V, double *W, void *stream);
/* Per-stage device time of the last solve in ms, in the order of the
 * reference profiler buckets (profilers.hpp:38-51): make-zero, gaussian
 * setup, force distribution, FFT forward, flow solve, FFT backward,
 * particle velocities+rotations. Stages fused into one kernel report their
 * combined time in the first bucket and 0 in the others. Requires
 * fs_fcm_set_timing(h,1) before the solve. */
int fs_fcm_set_timing(fs_fcm *h, int on);
int fs_fcm_stage_times(fs_fcm *h, float *ms7);
/* Per-kernel device time of the last solve in ms (no reference counterpart;
 * finer than fs_fcm_stage_times): [0] tables and bin sort, [1] tile lists,
 * [2] spread, [3] z R2C, [4] y forward, [5] x FFT + projection + x inverse,
 * [6] y inverse, [7] z C2R, [8] interpolation, [9] interpolation sums (0
 * for a kernel the solve's path does not have). Requires
 * fs_fcm_set_timing(h,1) before the solve. */
int fs_fcm_kernel_times(fs_fcm *h, float *ms10);
/* Launch-shape tunables (tile/batch shapes); values <= 0 keep current. */
typedef struct fs_fcm_tuning {
  int fft_batch;    /* columns per workgroup in the strided FFT passes */
  int spread_tile;  /* spread tile edge in x and y (4, 8 or 16)        */
  int spread_zrun;  /* largest spread tile extent in z (4..16)         */
  int fft_xlines;   /* x lines (per component) per workgroup, x pass  */
  int spread_mfma;  /* 1 spreads and interpolates with matrix-core
                       products on 8 x 8 x 16 tiles (grids with nx, ny
                       multiples of 8, nz of 16, nz >= 16 + ngd), 0 with
                       the vector gather and one wavefront per segment. In
                       fs_fcm_set_tuning a negative value keeps the current
                       setting, as do values <= 0 of the other fields. */
} fs_fcm_tuning;
int fs_fcm_get_tuning(fs_fcm *h, fs_fcm_tuning *t);
int fs_fcm_set_tuning(fs_fcm *h, const fs_fcm_tuning *t);
/* Autotuner (no reference counterpart; the reference has no launch
 * shapes): times `reps` mobility solves of np synthetic segments (uniform
 * in the box, forces/torques uniform in [-1, 1]) for each candidate of each
 * tunable in turn (spread_mfma, spread_tile, spread_zrun, fft_batch,
 * fft_xlines), keeps
 * the fastest, leaves it set and returns it in *best (may be NULL). Results
 * of every candidate agree to rounding (the FP64 path to 1e-12). */
int fs_fcm_autotune(fs_fcm *h, int np, int reps, fs_fcm_tuning *best);

/* ------------------------------------------------------------------ */
/* Collision barrier (replaces struct LinkedList,                      */
/* CollisionBarrierFilament.hpp:19-57)                                 */
/* ------------------------------------------------------------------ */
typedef struct fs_colbar_params {
  double box[3];        /* LfcmBox_x/y/z = NPTS*dx                         */
  double a;             /* segment radius (Rnm = 2a)                       */
  double fs_over_2a;    /* Barrier_FS_over_2a                              */
  double cap;           /* Barrier_distance_cap                            */
  int initial_friction_steps; /* adjacent pairs are repelled while nt < it */
} fs_colbar_params;

typedef struct fs_colbar fs_colbar;

/* LinkedList::initialise + set_global_segment_number
 * (CollisionBarrierFilament.cpp:274-367). seg_fil/seg_idx: host, np each. */
int fs_colbar_create(const fs_colbar_params *p, int np, const int *seg_fil,
                     const int *seg_idx, fs_colbar **out);
int fs_colbar_destroy(fs_colbar *h);
/* M_X, M_Y, M_Z cell counts (CollisionBarrierFilament.cpp:277-280). */
int fs_colbar_dims(fs_colbar *h, int *m3);
/* LinkedList::link (CollisionBarrierFilament.cpp:370-433) from Xs (device,
 * np x 3, unwrapped positions). */
int fs_colbar_link(fs_colbar *h, const double *Xs, void *stream);
/* LinkedList::apply_collision_barrier (CollisionBarrierFilament.cpp:24-271):
 * F (device, np x 3) += barrier forces computed from Y (device) with the
 * list built by the last link. min_dist2 (host, may be NULL) receives the
 * minimum squared distance over all visited pairs. */
int fs_colbar_apply(fs_colbar *h, const double *Y, int nt, double *F,
                    double *min_dist2, void *stream);
/* Host-array variants (copy in/out; synchronous): the shim's LinkedList. */
int fs_colbar_link_host(fs_colbar *h, const double *Xs);
int fs_colbar_apply_host(fs_colbar *h, const double *Y, int nt, double *F,
                         double *min_dist2);
/* The reference's head[ncell] and list[np] arrays (host out). */
int fs_colbar_get_lists(fs_colbar *h, int *head, int *list);
/* Pair masks: which pairs apply, pairs and the minimum distance visit.
 * FS_BARRIER_ALL: apply_collision_barrier (CollisionBarrierFilament.cpp:
 * 24-271), the default. FS_BARRIER_NOT_SWIMMER:
 * apply_collision_barrier_not_swimmer (:441-696), no pairs between
 * filament 0 and the others (seg_fil == 0 against seg_fil != 0).
 * FS_BARRIER_NOT_TEST_PARTICLES: apply_collision_barrier_not_test_particles
 * (:698-945), no pairs with either of the last two segments (np-2, np-1).
 * The mask holds for every later apply / pairs call. */
#define FS_BARRIER_ALL 0
#define FS_BARRIER_NOT_SWIMMER 1
#define FS_BARRIER_NOT_TEST_PARTICLES 2
int fs_colbar_set_mask(fs_colbar *h, int mask);
/* Unordered pairs (i<j) inside the cutoff r^2 < 1.21(2a)^2 found through the
 * cell list, sorted by (i,j); pairs: host, 2*max_pairs ints. */
int fs_colbar_pairs(fs_colbar *h, const double *Y, int nt, int max_pairs,
                    int *pairs, int *npairs);

/* No-allocation barrier. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). One call does what
 * fs_colbar_link(Xs) followed by fs_colbar_apply(Y, nt, F) does on a handle
 * created with the same parameters and pair mask (FS_BARRIER_*), with the
 * same kernels and results, but keeps the cell list in a 32 MiB device
 * array of the library instead of buffers it allocates: each GPU the call
 * runs on holds that array until the process ends. It calls no hipMalloc,
 * hipFree or allocating library routine and no host synchronisation.
 * seg_fil, seg_idx (np ints each: filament and segment index of every
 * segment), Xs, Y, F (np x 3) and min_dist2 (one double, may be NULL;
 * receives the minimum squared distance over the visited pairs) are device
 * pointers. Returns FS_ERR_SIZE when the cell list of the box (M = int(L /
 * 2.5a) cells per axis) and np segments need more than 32 MiB (about 3.8M
 * cells with 240,000 segments). Every call on one GPU shares its array:
 * calls that overlap on one GPU (other streams, other host threads) race. */
int fs_colbar_barrier_noalloc(const fs_colbar_params *p, int np,
                              const int *seg_fil, const int *seg_idx,
                              const double *Xs, const double *Y, int nt,
                              int mask, double *F, double *min_dist2,
                              void *stream);

/* ------------------------------------------------------------------ */
/* Filament system + implicit timestep (replaces class Filament,       */
/* filamentJsolver, ErrorCheckFCM and the Broyden loop of main.cpp)    */
/* ------------------------------------------------------------------ */
typedef struct fs_sim_params {
  fs_fcm_params fcm;
  int precision;            /* FS_FP64 or FS_FP32 for the FCM stages       */
  double dL;                /* segment spacing DL                          */
  double dt;                /* time step                                    */
  double tol;               /* Broyden tolerance TOL (infinity norm)       */
  int broyden_maxiter;      /* broyden_maxiter                              */
  int impl_euler_steps;     /* ImplEulerSteps                               */
  int initial_friction_steps; /* initialFrictionSteps                       */
  double weight_per_length[3]; /* gravity force per unit length            */
  int barrier_on;
  double barrier_fs_over_2a;
  double barrier_cap;
  int sperm_decay;          /* SpermCurvatureDecay                          */
  double strain_twist0;     /* STRAINTWISTZ: static twist strain            */
} fs_sim_params;

/* Per-filament properties (Filament.hpp:27-44). */
typedef struct fs_fil_props {
  double kap;        /* myKAP: bending modulus                               */
  double c;          /* myC: twist modulus                                   */
  double T0;         /* preferred-curvature amplitude (curvature/L_long)     */
  double omega;      /* AngFrequency                                         */
  double wavenumber; /* Wavenumber                                           */
  double phase;      /* Phase                                                */
  double alpha;      /* Alpha (helix, sin part)                              */
  double beta;       /* Beta (planar, cos part)                              */
  double radius;     /* bead radius                                          */
} fs_fil_props;

typedef struct fs_system fs_system;

/* Filament constructors + Jsolver(Np) + LinkedList/Fcm initialise
 * (main.cpp:230-261). nw: host, segments per filament (each >= 2). */
int fs_system_create(const fs_sim_params *p, int nfil, const int *nw,
                     const fs_fil_props *props, fs_system **out);
int fs_system_destroy(fs_system *h);
int fs_system_sizes(fs_system *h, int *nfil, int *np, int *nlam);
/* State in and out (host). Per bead: X,Xs,Xt,U,Us,Ut (np x 3), q,qs,qt
 * (np x 4, scalar first); per link: lam,lam1,lam2 (nlam x 3, links of one
 * filament contiguous). Any pointer may be NULL to skip that field. */
typedef struct fs_state_ptrs {
  double *X, *Xs, *Xt, *U, *Us, *Ut, *q, *qs, *qt, *lam, *lam1, *lam2;
  double *F, *TAU;
} fs_state_ptrs;
int fs_system_set_state(fs_system *h, const fs_state_ptrs *s);
int fs_system_get_state(fs_system *h, fs_state_ptrs *s);
/* Device pointers of the FCM arrays the system drives (np x 3 each). */
int fs_system_fcm_arrays(fs_system *h, double **Y, double **F, double **T,
                         double **V, double **W);
/* linked_list.F_total (CollisionBarrierFilament.hpp:32, filled at
 * CollisionBarrierFilament.cpp:681-686): the collision-barrier force alone,
 * which the last step added to F, into F_total (device, np x 3): the barrier
 * recomputed on the current Y with the cell list linked at the step's start
 * (each segment's pair sum is the one the step added). Zero when the
 * barrier is off. */
int fs_system_barrier_force(fs_system *h, int nt, double *F_total);
/* filaments_copy = filaments; filaments.resize(1) of the gait-reset-and-
 * remove driver (main_gaitresetandremove.cpp:331-337): copies every state
 * array of src's filaments first .. first + (dst's filament count) - 1 into
 * dst, device to device. Segment counts must match; the two systems should
 * share their parameters (dst is typically a one-filament system of src's
 * filament 0 with the same FCM grid and box). */
int fs_system_copy_filaments(fs_system *dst, fs_system *src, int first);
/* Filament::getCOM (Filament.hpp:248-255) of filament fil: the mean of its
 * bead positions X (main_gaitreset.cpp:760-772 prints its change over dt). */
int fs_system_com(fs_system *h, int fil, double com[3]);
/* The swimmer speed printout of main_gaitresetandremove.cpp:781-787: the
 * mean of V over filament fil's segments, accumulated in float. */
int fs_system_mean_velocity(fs_system *h, int fil, double v[3]);

/* One whole implicit timestep nt (main.cpp:284-754) on the device: initial
 * guess, forces, Jacobian, link, barrier, mobility, Broyden loop, step.
 * *iters receives the Broyden iteration count. */
int fs_system_timestep(fs_system *h, int nt, int *iters);

/* The same step decomposed, one call per reference call site. */
/* Filament::initialGuess + unRobotArmify (main.cpp:334-336). */
int fs_fil_initial_guess(fs_system *h);
/* setZeroForcesTorques + applyElasticTorques + applyConstraintForcesTorques
 * + applyExternalForcesTorques + Fcm::assign_filament_data
 * (main.cpp:337-361, Filament.hpp:271-402, FCMfunctions.cpp:882-918). */
int fs_fil_forces(fs_system *h, int nt);
/* filamentJsolver::buildJacobian (FilamentJacobianSolver.hpp:36-47):
 * per-filament approximate Jacobian and its inverse; resets the history. */
int fs_fil_build_jacobian(fs_system *h, int nt);
/* Copy filament f's inverse Jacobian (6Nw x 6Nw row-major) to host. */
int fs_fil_get_jinv(fs_system *h, int f, double *out);
/* Copy filament f's Jacobian before inversion (host out). */
int fs_fil_get_jacobian(fs_system *h, int f, double *out);
/* filamentJsolver::solve (FilamentJacobianSolver.hpp:49-73): x = H b with
 * the first iter history columns. b, x: device, 6Np. */
int fs_fil_jsolve(fs_system *h, const double *b, double *x, int iter);
/* Bad-Broyden column update (main.cpp:640-663): stores column iter from
 * dE = Enew - E and the last update (device, 6Np each). */
int fs_fil_broyden_update(fs_system *h, const double *E, const double *Enew,
                          const double *update, int iter);
/* Filament::update + unRobotArmify (main.cpp:528-531); update: device 6Np. */
int fs_fil_update(fs_system *h, const double *update);
/* ErrorCheckFCM (MobilityConstraints.cpp:447-770): err (device, 6Np) from
 * V, W (device); *check = 1 if any |err| > tol; *maxabs = max |err|. */
int fs_fil_residual(fs_system *h, int nt, const double *V, const double *W,
                    double *err, int *check, double *maxabs);
/* Filament::step (Filament.hpp:455-478, Bead.hpp:266-276). */
int fs_fil_step(fs_system *h, int nt);
/* Mobility on the system's Y/F/T into its V/W (the reference's
 * fcm.mobility_solve or friction_mobility_solve, chosen by nt). */
int fs_system_mobility(fs_system *h, int nt);
/* Collision barrier on the system's Y/F (link = 1 rebuilds the list). */
int fs_system_barrier(fs_system *h, int nt, int link, double *min_dist2);
/* Barrier of the following timesteps (main_gaitreset.cpp:431-458, 606-622:
 * off, FS_BARRIER_NOT_SWIMMER or the full barrier per pass): on = 0 skips
 * it, on = 1 applies it with the pair mask (FS_BARRIER_*). on = 1 needs a
 * system created with barrier_on = 1; not allowed while the system belongs
 * to a batch (its launches are captured). */
int fs_system_set_barrier(fs_system *h, int on, int mask);
/* The gait-reset drivers' trial pass (main_gaitreset.cpp:310-812 with
 * gait_reset_i < 2): fs_system_timestep without the final Filament::step.
 * X, Xt, U, Ut, q, qt, lam1 and lam2 are left as they were; the iterate
 * Xs, Us, qs, F, TAU and V, W hold the trial's result, and so does lam,
 * which the next step's initial guess does not reset (Filament.hpp:
 * 226-236), so the next pass of the same nt starts from it as in the
 * reference. Plain path only (no warm start attached). */
int fs_system_timestep_trial(fs_system *h, int nt, int *iters);

/* Outcome of the last fs_system_timestep / fs_system_timestep_trial, which
 * the reference reports on stdout only. converged: 1 when the final
 * residual met TOL, 0 when the loop stopped at broyden_maxiter (the step
 * is committed anyway, as main.cpp:503, 676-679 does). exploded: 1 when a
 * residual of the step exceeded 1e6 in the infinity norm ("Prepare for
 * explosion", MobilityConstraints.cpp:764-767). nonfinite: 1 when a
 * residual held a NaN or an infinity; the reference's test |e| > TOL never
 * counts a NaN, and neither does TOL here, so a step whose residual is NaN
 * throughout reports converged = 1 with nonfinite = 1. res_inf: infinity
 * norm of the final residual over its finite entries. A state that is not
 * finite usually stops the step earlier: its approximate Jacobian is
 * singular and fs_system_timestep returns FS_ERR_ARG. Pointers may be
 * NULL. */
int fs_system_step_status(fs_system *h, int *converged, int *exploded,
                          int *nonfinite, double *res_inf);

/* Deterministic initial configuration (the reference's overlap-rejected
 * placement, Filament.cpp:8-153, filament_initialisation_functions.cpp:
 * 631-650): nfil straight filaments of nw[f] segments, heads uniform in the
 * box, orientation uniform on the sphere, each retried (up to max_tries
 * times) while any of its segments is closer than min_sep to a segment
 * already placed (periodic). Writes host X (np x 3) and q (np x 4); seed:
 * 64-bit. fs_make_filaments_gpu tests windows of trials on the current GPU
 * and returns the same bits as fs_make_filaments. */
int fs_make_filaments(const double box[3], int nfil, const int *nw,
                      double dL, double min_sep, unsigned long long seed,
                      int max_tries, double *X, double *q);
int fs_make_filaments_gpu(const double box[3], int nfil, const int *nw,
                          double dL, double min_sep, unsigned long long seed,
                          int max_tries, double *X, double *q);

/* ------------------------------------------------------------------ */
/* Learned warm start (no reference counterpart: replaces the linear   */
/* extrapolation of Bead::initialGuess, Bead.hpp:202-217, by an MLP    */
/* correction, kept only when it lowers the residual's infinity norm)  */
/* ------------------------------------------------------------------ */
typedef struct fs_mlp fs_mlp;
/* MLP n_in -> width -> width -> width -> n_out (ReLU), BF16 weights on
 * MFMA with FP32 accumulation, FP32 master weights, Adam. */
int fs_mlp_create(int n_in, int width, int n_out, unsigned long long seed,
                  fs_mlp **out);
int fs_mlp_destroy(fs_mlp *h);
int fs_mlp_num_params(fs_mlp *h, int *n);
/* Optional fixed linear term: y_n += x_n S^T with S (host, n_out x n_in
 * row-major, FP32, stored in BF16), e.g. a ridge regression of the
 * normalised targets on the normalised features; the layers are then
 * trained on what it leaves. NULL removes it. */
int fs_mlp_set_skip(fs_mlp *h, const float *S);
/* Feature/target normalisation: x_n = (x - mu)/sigma, y = tau * y_n. */
int fs_mlp_set_norm(fs_mlp *h, const float *mu, const float *sigma,
                    const float *tau);
/* X (device, batch x n_in) -> Y (device, batch x n_out), fp32. */
int fs_mlp_forward(fs_mlp *h, int batch, const float *X, float *Y,
                   void *stream);
/* One Adam step on the mean squared normalised error against T (device,
 * batch x n_out); *loss (host, may be NULL) receives the loss. */
int fs_mlp_train_step(fs_mlp *h, int batch, const float *X, const float *T,
                      float lr, float *loss, void *stream);
/* Loss gradient w.r.t. all parameters (host out, fp32) packed as
 * W1,b1,W2,b2,W3,b3,W4,b4 with W_l row-major [out][in]. */
int fs_mlp_gradients(fs_mlp *h, int batch, const float *X, const float *T,
                     float *grad, int n_grad, float *loss);
int fs_mlp_get_params(fs_mlp *h, float *p, int n);
int fs_mlp_set_params(fs_mlp *h, const float *p, int n);

/* Attach a network to a system (NULL detaches). Needs one segment count
 * Nw for every filament; the network maps 24 Nw - 4 features to 6 Nw
 * outputs (fs_system_warmstart_dims). */
int fs_system_set_warmstart(fs_system *h, fs_mlp *m);
int fs_system_warmstart_dims(fs_system *h, int *n_in, int *n_out);
/* Record (features, target) of every following timestep into device
 * buffers X (capacity x n_in) and T (capacity x n_out), one row per
 * filament per step, until capacity rows are filled. NULL stops. */
int fs_system_ws_record(fs_system *h, float *X, float *T, int capacity);
int fs_system_ws_count(fs_system *h, int *count);
/* How often the extrapolated guess is solved beside the network's (no
 * reference counterpart). every = 1 (the default): on every step, and the
 * guess with the smaller residual infinity norm starts the Broyden
 * iteration. every = k > 1: on every k-th step (and the first after this
 * call or after attaching a network); on the other steps only the network's
 * guess is solved, and it is kept when its residual is below the last
 * measured residual of the extrapolated guess, else the extrapolated guess
 * is solved and used. Saves one mobility solve per step while the network
 * keeps winning. */
int fs_system_set_warmstart_check(fs_system *h, int every);
/* Broyden history carried across timesteps (no reference counterpart; the
 * reference restarts the history with each step's new Jacobian,
 * FilamentJacobianSolver.hpp:40-41). max_cols = 0 (the default): restart
 * every step. max_cols > 0 (at most 192): a step starts from the new
 * approximate inverse plus the rank-one corrections the previous steps
 * accumulated; when they exceed max_cols columns the newest max_cols / 2
 * are kept. Every step still ends at the same residual tolerance. Setting
 * the state clears the carried history. The history buffers grow to
 * broyden_maxiter + max_cols columns of 6 Np doubles (two of them) on the
 * first call that needs it; not while the system belongs to a batch. */
int fs_system_set_history_carry(fs_system *h, int max_cols);
/* Whether the last timestep started from the network's guess, and the
 * infinity norms of the residuals of both candidate guesses (res_plain is
 * NaN when the extrapolated guess was not solved). */
int fs_system_last_step_info(fs_system *h, int *used_mlp, double *res_plain,
                             double *res_mlp);
/* Per-timestep profile in the reference profiler's buckets
 * (profilers.hpp:24-57), milliseconds of device time measured with HIP
 * events on the system's stream, for the last fs_system_timestep with
 * profiling on. Order (14): FCM, Assign F, Bld Jinv, ErrorChk, Slv Jinv,
 * Total, ColBarri, FCMMakeZ, FCMGauss, FCMForcD, FCMFFTFw, FCMFFTBk,
 * FCMFlowS, FCMParVR. Here FCMMakeZ is 0 (the spread writes every point
 * once), FCMGauss holds the Gaussian tables and the bin sort, FCMFFTFw the
 * z and y forward passes, FCMFlowS the fused x forward FFT + projection + x
 * inverse FFT and FCMFFTBk the y and z inverse passes. Profiling adds event
 * records (and one synchronisation per step), not kernels; not allowed in a
 * batch. */
int fs_system_set_profiling(fs_system *h, int on);
int fs_system_profile(fs_system *h, double *ms, int n);
/* Autotune the system's FCM solver for its segment count (fs_fcm_autotune). */
int fs_system_autotune(fs_system *h, int reps, fs_fcm_tuning *best);
/* Set the launch shapes of the system's FCM solver (fs_fcm_set_tuning). */
int fs_system_set_tuning(fs_system *h, const fs_fcm_tuning *t);

/* ---- Spring-link networks (struct spring_links, spring_link_functions.hpp:
 * 18-39). n links as (filament A, bead B, filament C, bead D) rows; each
 * timestep's forces (main.cpp:354-357, 540-543) add, in link order,
 * f = -k (|d| - L) d / |d| to bead (A, B) and -f to bead (C, D), with d the
 * minimum-image difference of their previous positions Xs
 * (apply_spring_link_forces, spring_link_functions.cpp:24-45). n = 0
 * removes them. Set them before a batch is created from the system. The
 * file functions keep the reference's "-springlinks.dat" layout
 * (spring_link_functions.cpp:50-101); read with links = NULL for the count.
 */
int fs_system_set_spring_links(fs_system *h, int n, const int *links, double k,
                               double natural_length);
int fs_io_write_spring_links(const char *path, int n, const int *links);
int fs_io_read_spring_links(const char *path, int *n, int *links, int cap);

/* ---- Trajectory output and restart (file_functions.cpp, Filament.hpp
 * writeData/writeBackupData/continue_from_checkpoint). Text files keep the
 * reference's layout (space separated, fixed notation, 16 decimals):
 *   .dat  appends "nt " + per bead Xs(3) q(4) (file_functions.cpp:150-161)
 *   .bak  "nt " + per filament: per bead X Xt q qt U Ut, per link lam lam1
 *         lam2 (file_functions.cpp:211-222); reading sets Xs=X, qs=q, Us=U
 *         (Filament.hpp:691-759)
 *   .par  parameter file whose line 2 holds the filament count and line 5
 *         the bead numbers (file_functions.cpp:6-81, 242-280)
 * The binary checkpoint keeps every state array bit for bit. Host arrays in
 * the fs_state_ptrs layout. */
int fs_io_write_dat(const char *path, int nt, int nf, const int *nw,
                    const double *X, const double *q);
int fs_io_write_bak(const char *path, int nt, int nf, const int *nw,
                    const fs_state_ptrs *s);
int fs_io_read_bak(const char *path, int nf, const int *nw, fs_state_ptrs *s,
                   int *nt);
int fs_io_write_par(const char *path, const char *name, const char *tag, int nf,
                    const int *nw, const double *bending_factors,
                    const fs_sim_params *p, double kap, double c,
                    int time_steps, int plot_steps);
int fs_io_read_par(const char *path, int *nf, int *nw, int cap);
int fs_io_write_checkpoint(const char *path, int nt, int nf, const int *nw,
                           const fs_state_ptrs *s);
int fs_io_read_checkpoint(const char *path, int nf, const int *nw,
                          fs_state_ptrs *s, int *nt);
/* The same for a system's state (copied to or from the device). */
int fs_system_write_dat(fs_system *h, const char *path, int nt);
int fs_system_write_bak(fs_system *h, const char *path, int nt);
int fs_system_read_bak(fs_system *h, const char *path, int *nt);
int fs_system_write_checkpoint(fs_system *h, const char *path, int nt);
int fs_system_read_checkpoint(fs_system *h, const char *path, int *nt);

/* ---- Swimmer and force output (SaveExtendedSwimmerData /
 * SaveExtendedForcesData, the gait-reset files; file_functions.cpp:102-145,
 * 164-208, Filament.hpp:509-650), in the text layout above:
 *   swimmer velocity  "nt " + filament 0 per bead Xs(3) qs(4), then per
 *                     segment V(3) W(3) F(3)
 *   swimmer collision "nt " + filament 0 per segment F_total(3)
 *   forces            "nt " + per filament per segment F(3) T(3), filament
 *                     f reading rows (first segment of f) / 6 + i as
 *                     Filament::writeForceData does (filament 0: its own)
 * At nt = 0 every velocity and force is written as the integer 0, and the
 * arrays may be NULL. The name writers start a new file with "it " and the
 * column names of filament 0 (nw0 segments). Host arrays; the fs_system_
 * versions write the system's last step or trial pass (F_total from
 * fs_system_barrier_force). */
int fs_io_write_swimmer_velocity_names(const char *path, int nw0);
int fs_io_write_swimmer_collision_names(const char *path, int nw0);
int fs_io_write_force_names(const char *path, int nw0);
int fs_io_write_swimmer_velocity(const char *path, int nt, int nw0,
                                 const double *Xs, const double *qs,
                                 const double *V, const double *W,
                                 const double *F);
int fs_io_write_swimmer_collision(const char *path, int nt, int nw0,
                                  const double *F_total);
int fs_io_write_forces(const char *path, int nt, int nf, const int *nw,
                       const double *F, const double *T);
int fs_system_write_swimmer_velocity(fs_system *h, const char *path, int nt);
int fs_system_write_swimmer_collision(fs_system *h, const char *path, int nt);
int fs_system_write_forces(fs_system *h, const char *path, int nt);

/* ---- Batched sweep (no reference counterpart: the reference runs one
 * process per sweep point, batch_run.py:60-229). Independent systems of any
 * grid and filament counts advance together: each system's timestep is
 * captured once as hipGraphs (begin, one Broyden iteration, end) with the
 * loop state on the device, and one parent graph per phase launches every
 * system. A batched system reproduces its unbatched fs_system_timestep bit
 * for bit. Systems need initial_friction_steps == 0, no warm start and no
 * carried Broyden history; they must outlive the batch and must not be
 * stepped on their own meanwhile. */
typedef struct fs_batch fs_batch;
int fs_batch_create(fs_system **systems, int n, fs_batch **out);
int fs_batch_destroy(fs_batch *b);
int fs_batch_size(fs_batch *b, int *n);
/* One timestep of every system; iters[i] = Broyden iterations of system i
 * (may be NULL). */
int fs_batch_timestep(fs_batch *b, int nt, int *iters);
/* Launch mode: lanes in [1, 16] launches each system's own graphs
 * round-robin on that many streams so independent systems overlap on the
 * hardware queues (default min(n, 4)); 0 launches one parent graph per
 * phase. Results do not depend on the mode. */
int fs_batch_set_lanes(fs_batch *b, int lanes);
/* Fused FCM stages (off by default; available when every system runs the
 * matrix-core tile path and initial_friction_steps == 0): in the begin and
 * iteration phases the mobility solves of all systems run as one launch
 * per FCM stage for each group of systems with equal launch shapes
 * (precision, grid, support, tile and FFT shapes), between the systems'
 * own graphs for the work before and after the solve (on the lanes).
 * Results do not depend on it; FS_ERR_ARG if on and a system cannot
 * join. */
int fs_batch_set_fused(fs_batch *b, int on);

#ifdef __cplusplus
}
#endif
#endif /* FILSIM_H */
