// system_api.cpp -- C-ABI of the filament system (fs_system_*, fs_fil_*)
// and the deterministic initial-configuration generator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "system_impl.hpp"
#include "warmstart.hpp"

using fs::SystemImpl;

struct fs_system {
  SystemImpl impl;
  fs::WarmStart ws;
  int batches = 0;  // batches holding this system (captured launches)
};

namespace {

int copy_field(double *dst, const double *src, size_t n, hipMemcpyKind k) {
  if (!dst || !src || n == 0) return FS_OK;
  FS_HIP(hipMemcpy(dst, src, sizeof(double) * n, k));
  return FS_OK;
}

}  // namespace

extern "C" int fs_system_create(const fs_sim_params *p, int nfil, const int *nw,
                                const fs_fil_props *props, fs_system **out) {
  FS_CHECK(p && nw && props && out, FS_ERR_ARG, "null argument");
  *out = nullptr;
  fs_system *h = new fs_system();
  int rc = h->impl.init(*p, nfil, nw, props);
  if (rc != FS_OK) {
    h->impl.release();
    delete h;
    return rc;
  }
  *out = h;
  return FS_OK;
}

extern "C" int fs_system_destroy(fs_system *h) {
  if (!h) return FS_OK;
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  h->ws.release();
  h->impl.release();
  delete h;
  return FS_OK;
}

extern "C" int fs_system_sizes(fs_system *h, int *nfil, int *np, int *nlam) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  if (nfil) *nfil = h->impl.nf;
  if (np) *np = h->impl.np;
  if (nlam) *nlam = h->impl.nlam;
  return FS_OK;
}

extern "C" int fs_system_set_state(fs_system *h, const fs_state_ptrs *s) {
  FS_CHECK(h && s, FS_ERR_ARG, "null argument");
  FS_HIP(hipStreamSynchronize(h->impl.st));
  auto &d = h->impl.d;
  size_t n3 = 3 * (size_t)h->impl.np, n4 = 4 * (size_t)h->impl.np,
         l3 = 3 * (size_t)h->impl.nlam;
  auto K = hipMemcpyHostToDevice;
  FS_TRY(copy_field(d.X, s->X, n3, K));
  FS_TRY(copy_field(d.Xs, s->Xs, n3, K));
  FS_TRY(copy_field(d.Xt, s->Xt, n3, K));
  FS_TRY(copy_field(d.U, s->U, n3, K));
  FS_TRY(copy_field(d.Us, s->Us, n3, K));
  FS_TRY(copy_field(d.Ut, s->Ut, n3, K));
  FS_TRY(copy_field(d.q, s->q, n4, K));
  FS_TRY(copy_field(d.qs, s->qs, n4, K));
  FS_TRY(copy_field(d.qt, s->qt, n4, K));
  FS_TRY(copy_field(d.lam, s->lam, l3, K));
  FS_TRY(copy_field(d.lam1, s->lam1, l3, K));
  FS_TRY(copy_field(d.lam2, s->lam2, l3, K));
  FS_TRY(copy_field(d.F, s->F, n3, K));
  FS_TRY(copy_field(d.TAU, s->TAU, n3, K));
  h->impl.hist_cols = 0;  // a new state: no history carried in
  return FS_OK;
}

extern "C" int fs_system_get_state(fs_system *h, fs_state_ptrs *s) {
  FS_CHECK(h && s, FS_ERR_ARG, "null argument");
  FS_HIP(hipStreamSynchronize(h->impl.st));
  auto &d = h->impl.d;
  size_t n3 = 3 * (size_t)h->impl.np, n4 = 4 * (size_t)h->impl.np,
         l3 = 3 * (size_t)h->impl.nlam;
  auto K = hipMemcpyDeviceToHost;
  FS_TRY(copy_field(s->X, d.X, n3, K));
  FS_TRY(copy_field(s->Xs, d.Xs, n3, K));
  FS_TRY(copy_field(s->Xt, d.Xt, n3, K));
  FS_TRY(copy_field(s->U, d.U, n3, K));
  FS_TRY(copy_field(s->Us, d.Us, n3, K));
  FS_TRY(copy_field(s->Ut, d.Ut, n3, K));
  FS_TRY(copy_field(s->q, d.q, n4, K));
  FS_TRY(copy_field(s->qs, d.qs, n4, K));
  FS_TRY(copy_field(s->qt, d.qt, n4, K));
  FS_TRY(copy_field(s->lam, d.lam, l3, K));
  FS_TRY(copy_field(s->lam1, d.lam1, l3, K));
  FS_TRY(copy_field(s->lam2, d.lam2, l3, K));
  FS_TRY(copy_field(s->F, d.F, n3, K));
  FS_TRY(copy_field(s->TAU, d.TAU, n3, K));
  return FS_OK;
}

extern "C" int fs_system_fcm_arrays(fs_system *h, double **Y, double **F,
                                    double **T, double **V, double **W) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  if (Y) *Y = h->impl.d.Ya;
  if (F) *F = h->impl.d.Fa;
  if (T) *T = h->impl.d.Ta;
  if (V) *V = h->impl.d.V;
  if (W) *W = h->impl.d.W;
  return FS_OK;
}

extern "C" int fs_system_barrier_force(fs_system *h, int nt, double *F_total) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.barrier_force(nt, F_total);
}

// filaments_copy = filaments; filaments.resize(1) (main_gaitresetandremove.
// cpp:331-334), as a device-to-device copy of every state array of source
// filaments first .. first + dst's count - 1.
extern "C" int fs_system_copy_filaments(fs_system *dst, fs_system *src,
                                        int first) {
  FS_CHECK(dst && src, FS_ERR_ARG, "null handle");
  FS_CHECK(dst != src, FS_ERR_ARG, "source and destination are one system");
  SystemImpl &D = dst->impl, &S = src->impl;
  FS_CHECK(first >= 0 && first + D.nf <= S.nf, FS_ERR_ARG,
           "filaments %d..%d are outside the source's %d", first,
           first + D.nf - 1, S.nf);
  for (int f = 0; f < D.nf; ++f)
    FS_CHECK(D.hnw[f] == S.hnw[first + f], FS_ERR_ARG,
             "filament %d has %d segments, source filament %d has %d", f,
             D.hnw[f], first + f, S.hnw[first + f]);
  FS_HIP(hipStreamSynchronize(S.st));
  FS_HIP(hipStreamSynchronize(D.st));
  const size_t b = S.hboff[first], l = S.hloff[first];
  const size_t n3 = 3 * (size_t)D.np, n4 = 4 * (size_t)D.np,
               l3 = 3 * (size_t)D.nlam;
  auto &d = D.d;
  auto &s = S.d;
  auto K = hipMemcpyDeviceToDevice;
  FS_TRY(copy_field(d.X, s.X + 3 * b, n3, K));
  FS_TRY(copy_field(d.Xs, s.Xs + 3 * b, n3, K));
  FS_TRY(copy_field(d.Xt, s.Xt + 3 * b, n3, K));
  FS_TRY(copy_field(d.U, s.U + 3 * b, n3, K));
  FS_TRY(copy_field(d.Us, s.Us + 3 * b, n3, K));
  FS_TRY(copy_field(d.Ut, s.Ut + 3 * b, n3, K));
  FS_TRY(copy_field(d.q, s.q + 4 * b, n4, K));
  FS_TRY(copy_field(d.qs, s.qs + 4 * b, n4, K));
  FS_TRY(copy_field(d.qt, s.qt + 4 * b, n4, K));
  FS_TRY(copy_field(d.lam, s.lam + 3 * l, l3, K));
  FS_TRY(copy_field(d.lam1, s.lam1 + 3 * l, l3, K));
  FS_TRY(copy_field(d.lam2, s.lam2 + 3 * l, l3, K));
  FS_TRY(copy_field(d.F, s.F + 3 * b, n3, K));
  FS_TRY(copy_field(d.TAU, s.TAU + 3 * b, n3, K));
  return FS_OK;
}

// Filament::getCOM (Filament.hpp:248-255): the mean of the filament's bead
// positions X, summed in bead order and divided by the bead count.
extern "C" int fs_system_com(fs_system *h, int fil, double com[3]) {
  FS_CHECK(h && com, FS_ERR_ARG, "null argument");
  SystemImpl &S = h->impl;
  FS_CHECK(fil >= 0 && fil < S.nf, FS_ERR_ARG, "no filament %d", fil);
  const int nw = S.hnw[fil];
  std::vector<double> x(3 * (size_t)nw);
  FS_HIP(hipStreamSynchronize(S.st));
  FS_TRY(copy_field(x.data(), S.d.X + 3 * (size_t)S.hboff[fil], x.size(),
                    hipMemcpyDeviceToHost));
  double c[3] = {0, 0, 0};
  for (int n = 0; n < nw; ++n)
    for (int k = 0; k < 3; ++k) c[k] += x[3 * n + k];
  for (int k = 0; k < 3; ++k) com[k] = c[k] / nw;
  return FS_OK;
}

// The swimmer's mean segment velocity of main_gaitresetandremove.cpp:781-
// 787: float v += V[s] / n over the filament's n segments of the last
// mobility solve (each term in double, the sum rounded to float).
extern "C" int fs_system_mean_velocity(fs_system *h, int fil, double v[3]) {
  FS_CHECK(h && v, FS_ERR_ARG, "null argument");
  SystemImpl &S = h->impl;
  FS_CHECK(fil >= 0 && fil < S.nf, FS_ERR_ARG, "no filament %d", fil);
  const int nw = S.hnw[fil];
  std::vector<double> x(3 * (size_t)nw);
  FS_HIP(hipStreamSynchronize(S.st));
  FS_TRY(copy_field(x.data(), S.d.V + 3 * (size_t)S.hboff[fil], x.size(),
                    hipMemcpyDeviceToHost));
  float a[3] = {0.f, 0.f, 0.f};
  for (int n = 0; n < nw; ++n)
    for (int k = 0; k < 3; ++k) a[k] = (float)(a[k] + x[3 * n + k] / nw);
  for (int k = 0; k < 3; ++k) v[k] = a[k];
  return FS_OK;
}

namespace {

// One timestep (commit = 1) or gait-reset trial pass (commit = 0), with the
// profiler's per-step buckets when profiling is on.
int timestep_profiled(fs_system *h, int nt, int *iters, bool commit) {
  SystemImpl &S = h->impl;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (S.prof) {
    S.prof_reset();
    t0 = S.prof_event();
    t1 = S.prof_event();
    if (t0) FS_HIP(hipEventRecord(t0, S.st));
  }
  int rc = !commit ? S.timestep_trial(nt, iters)
           : h->ws.attached() ? h->ws.timestep(S, nt, iters)
                              : S.timestep(nt, iters);
  if (rc != FS_OK || !S.prof) return rc;
  if (t1) FS_HIP(hipEventRecord(t1, S.st));
  FS_HIP(hipStreamSynchronize(S.st));
  FS_TRY(S.prof_collect());
  float total = 0;
  if (t0 && t1) FS_HIP(hipEventElapsedTime(&total, t0, t1));
  S.pms[SystemImpl::P_TOTAL] = total;
  return FS_OK;
}

}  // namespace

extern "C" int fs_system_timestep(fs_system *h, int nt, int *iters) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return timestep_profiled(h, nt, iters, true);
}

extern "C" int fs_system_set_spring_links(fs_system *h, int n,
                                          const int *links, double k,
                                          double natural_length) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  return h->impl.set_spring_links(n, links, k, natural_length);
}

extern "C" int fs_system_set_profiling(fs_system *h, int on) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  FS_HIP(hipStreamSynchronize(h->impl.st));
  h->impl.prof = on ? 1 : 0;
  h->impl.fcm.prof = on ? 1 : 0;
  h->impl.prof_reset();
  return FS_OK;
}

extern "C" int fs_system_profile(fs_system *h, double *ms, int n) {
  FS_CHECK(h && ms && n >= SystemImpl::P_N, FS_ERR_ARG,
           "need room for %d buckets", (int)SystemImpl::P_N);
  for (int i = 0; i < SystemImpl::P_N; ++i) ms[i] = h->impl.pms[i];
  return FS_OK;
}

namespace fs {
struct MlpImpl;
MlpImpl *mlp_impl(fs_mlp *h);
}  // namespace fs

extern "C" int fs_system_set_barrier(fs_system *h, int on, int mask) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  return h->impl.set_barrier(on, mask);
}

extern "C" int fs_system_timestep_trial(fs_system *h, int nt, int *iters) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(!h->ws.attached(), FS_ERR_STATE,
           "trial steps run on the plain path (detach the warm start)");
  return timestep_profiled(h, nt, iters, false);
}

extern "C" int fs_system_set_warmstart(fs_system *h, fs_mlp *m) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  return h->ws.attach(h->impl, fs::mlp_impl(m));
}

extern "C" int fs_system_set_warmstart_check(fs_system *h, int every) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(every >= 1, FS_ERR_ARG, "check interval must be >= 1");
  h->ws.check_every = every;
  h->ws.since_check = 0;
  h->ws.ref_plain = -1;
  return FS_OK;
}

extern "C" int fs_system_set_history_carry(fs_system *h, int max_cols) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  return h->impl.set_history_carry(max_cols);
}

extern "C" int fs_system_autotune(fs_system *h, int reps, fs_fcm_tuning *best) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  return h->impl.fcm.autotune(h->impl.np, reps, best);
}

// ---- trajectory files of a system (io.cpp holds the formats)
namespace {

struct HostState {
  std::vector<double> a3[6], a4[3], al[3];
  fs_state_ptrs p{};
  explicit HostState(const SystemImpl &S) {
    for (auto &v : a3) v.resize(3 * (size_t)S.np);
    for (auto &v : a4) v.resize(4 * (size_t)S.np);
    for (auto &v : al) v.resize(3 * (size_t)std::max(S.nlam, 1));
    p.X = a3[0].data();
    p.Xs = a3[1].data();
    p.Xt = a3[2].data();
    p.U = a3[3].data();
    p.Us = a3[4].data();
    p.Ut = a3[5].data();
    p.q = a4[0].data();
    p.qs = a4[1].data();
    p.qt = a4[2].data();
    p.lam = al[0].data();
    p.lam1 = al[1].data();
    p.lam2 = al[2].data();
  }
};

}  // namespace

extern "C" int fs_system_write_dat(fs_system *h, const char *path, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  HostState hs(h->impl);
  FS_TRY(fs_system_get_state(h, &hs.p));
  return fs_io_write_dat(path, nt, h->impl.nf, h->impl.hnw.data(), hs.p.X,
                         hs.p.q);
}

// The swimmer (filament 0) and force rows of the system's last step or
// trial pass: Xs, qs of the iterate, V, W and F of the last mobility solve.
extern "C" int fs_system_write_swimmer_velocity(fs_system *h, const char *path,
                                                int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  SystemImpl &S = h->impl;
  const size_t n = S.hnw[0];
  std::vector<double> xs(3 * n), qs(4 * n), v(3 * n), w(3 * n), f(3 * n);
  FS_HIP(hipStreamSynchronize(S.st));
  auto K = hipMemcpyDeviceToHost;
  FS_TRY(copy_field(xs.data(), S.d.Xs, 3 * n, K));
  FS_TRY(copy_field(qs.data(), S.d.qs, 4 * n, K));
  FS_TRY(copy_field(v.data(), S.d.V, 3 * n, K));
  FS_TRY(copy_field(w.data(), S.d.W, 3 * n, K));
  FS_TRY(copy_field(f.data(), S.d.Fa, 3 * n, K));
  return fs_io_write_swimmer_velocity(path, nt, (int)n, xs.data(), qs.data(),
                                      v.data(), w.data(), f.data());
}

extern "C" int fs_system_write_swimmer_collision(fs_system *h,
                                                 const char *path, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  SystemImpl &S = h->impl;
  fs::dbuf d;
  FS_TRY(d.alloc(sizeof(double) * 3 * (size_t)S.np));
  FS_TRY(S.barrier_force(nt, d.as<double>()));
  std::vector<double> ft(3 * (size_t)S.hnw[0]);
  FS_TRY(copy_field(ft.data(), d.as<double>(), ft.size(),
                    hipMemcpyDeviceToHost));
  return fs_io_write_swimmer_collision(path, nt, S.hnw[0], ft.data());
}

extern "C" int fs_system_write_forces(fs_system *h, const char *path, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  SystemImpl &S = h->impl;
  const size_t n3 = 3 * (size_t)S.np;
  std::vector<double> f(n3), t(n3);
  FS_HIP(hipStreamSynchronize(S.st));
  FS_TRY(copy_field(f.data(), S.d.Fa, n3, hipMemcpyDeviceToHost));
  FS_TRY(copy_field(t.data(), S.d.Ta, n3, hipMemcpyDeviceToHost));
  return fs_io_write_forces(path, nt, S.nf, S.hnw.data(), f.data(), t.data());
}

extern "C" int fs_system_write_bak(fs_system *h, const char *path, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  HostState hs(h->impl);
  FS_TRY(fs_system_get_state(h, &hs.p));
  return fs_io_write_bak(path, nt, h->impl.nf, h->impl.hnw.data(), &hs.p);
}

extern "C" int fs_system_read_bak(fs_system *h, const char *path, int *nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  HostState hs(h->impl);
  FS_TRY(fs_io_read_bak(path, h->impl.nf, h->impl.hnw.data(), &hs.p, nt));
  return fs_system_set_state(h, &hs.p);
}

extern "C" int fs_system_write_checkpoint(fs_system *h, const char *path,
                                          int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  HostState hs(h->impl);
  FS_TRY(fs_system_get_state(h, &hs.p));
  return fs_io_write_checkpoint(path, nt, h->impl.nf, h->impl.hnw.data(),
                                &hs.p);
}

extern "C" int fs_system_read_checkpoint(fs_system *h, const char *path,
                                         int *nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  HostState hs(h->impl);
  FS_TRY(fs_io_read_checkpoint(path, h->impl.nf, h->impl.hnw.data(), &hs.p,
                               nt));
  return fs_system_set_state(h, &hs.p);
}

extern "C" int fs_system_set_tuning(fs_system *h, const fs_fcm_tuning *t) {
  FS_CHECK(h && t, FS_ERR_ARG, "null argument");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  return h->impl.fcm.set_tuning(*t);
}

extern "C" int fs_system_warmstart_dims(fs_system *h, int *n_in, int *n_out) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  for (int f = 1; f < h->impl.nf; ++f)
    FS_CHECK(h->impl.hnw[f] == h->impl.hnw[0], FS_ERR_ARG,
             "the warm start needs one segment count for every filament");
  if (n_in) *n_in = fs::ws_inputs(h->impl.hnw[0]);
  if (n_out) *n_out = fs::ws_outputs(h->impl.hnw[0]);
  return FS_OK;
}

extern "C" int fs_system_ws_record(fs_system *h, float *X, float *T,
                                   int capacity) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  FS_CHECK(h->batches == 0, FS_ERR_STATE,
           "the system belongs to a batch (its launches are captured)");
  return h->ws.record(h->impl, X, T, capacity);
}

extern "C" int fs_system_ws_count(fs_system *h, int *count) {
  FS_CHECK(h && count, FS_ERR_ARG, "null argument");
  *count = h->ws.rec_count;
  return FS_OK;
}

extern "C" int fs_system_last_step_info(fs_system *h, int *used_mlp,
                                        double *res_plain, double *res_mlp) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  if (used_mlp) *used_mlp = h->ws.used_mlp;
  if (res_plain) *res_plain = h->ws.tracking ? h->ws.res_plain : h->impl.last_res0;
  if (res_mlp) *res_mlp = h->ws.res_mlp;
  return FS_OK;
}

extern "C" int fs_system_step_status(fs_system *h, int *converged,
                                     int *exploded, int *nonfinite,
                                     double *res_inf) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  if (converged) *converged = h->impl.step_check ? 0 : 1;
  if (exploded) *exploded = h->impl.step_exploded;
  if (nonfinite) *nonfinite = h->impl.step_nonfinite;
  if (res_inf) *res_inf = h->impl.last_res;
  return FS_OK;
}

extern "C" int fs_fil_initial_guess(fs_system *h) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.initial_guess();
}

extern "C" int fs_fil_forces(fs_system *h, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.forces(nt);
}

extern "C" int fs_fil_build_jacobian(fs_system *h, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.build_jacobian(nt);
}

static int get_matrix(fs_system *h, int f, double *out, bool inverse) {
  FS_CHECK(h && out, FS_ERR_ARG, "null argument");
  FS_CHECK(f >= 0 && f < h->impl.nf, FS_ERR_ARG, "bad filament %d", f);
  if (inverse) return h->impl.jinv_dense(f, out);
  FS_HIP(hipStreamSynchronize(h->impl.st));
  int n = 6 * h->impl.hnw[f];
  FS_HIP(hipMemcpy(out, h->impl.Jb.as<double>() + h->impl.hjoff[f],
                   sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost));
  return FS_OK;
}

extern "C" int fs_fil_get_jinv(fs_system *h, int f, double *out) {
  return get_matrix(h, f, out, true);
}

extern "C" int fs_fil_get_jacobian(fs_system *h, int f, double *out) {
  return get_matrix(h, f, out, false);
}

extern "C" int fs_fil_jsolve(fs_system *h, const double *b, double *x,
                             int iter) {
  FS_CHECK(h && b && x, FS_ERR_ARG, "null argument");
  return h->impl.jsolve(b, x, iter);
}

extern "C" int fs_fil_broyden_update(fs_system *h, const double *E,
                                     const double *Enew, const double *update,
                                     int iter) {
  FS_CHECK(h && E && Enew && update, FS_ERR_ARG, "null argument");
  return h->impl.broyden_update(E, Enew, update, iter);
}

extern "C" int fs_fil_update(fs_system *h, const double *update) {
  FS_CHECK(h && update, FS_ERR_ARG, "null argument");
  return h->impl.update(update);
}

extern "C" int fs_fil_residual(fs_system *h, int nt, const double *V,
                               const double *W, double *err, int *check,
                               double *maxabs) {
  FS_CHECK(h && V && W && err, FS_ERR_ARG, "null argument");
  return h->impl.residual(nt, V, W, err, check, maxabs);
}

extern "C" int fs_fil_step(fs_system *h, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.step(nt);
}

extern "C" int fs_system_mobility(fs_system *h, int nt) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.mobility(nt);
}

extern "C" int fs_system_barrier(fs_system *h, int nt, int link,
                                 double *min_dist2) {
  FS_CHECK(h, FS_ERR_ARG, "null handle");
  return h->impl.barrier(nt, link, min_dist2);
}

// ---------------------------------------------------------------------
// Batched sweep: many independent suspensions advanced together. No
// reference counterpart (the reference runs one process per sweep point,
// batch_run.py:60-229). Each suspension's timestep is captured once as
// three graphs (begin, one Broyden iteration, end) with its loop state on
// the device; a parent graph per phase holds every suspension as an
// independent child node. A timestep launches begin, then the iteration
// graph until no suspension is still iterating (one status read per
// launch, not one per suspension and iteration), then end.
// ---------------------------------------------------------------------
struct fs_batch {
  std::vector<fs_system *> sys;
  hipStream_t st = nullptr;
  hipGraph_t g[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t ge[3] = {nullptr, nullptr, nullptr};
  // lanes > 0: each suspension's own executable graphs, launched round-robin
  // on `lanes` streams so independent suspensions overlap on the hardware
  // queues; lanes == 0: the parent graph of each phase
  int lanes = 0;
  std::vector<hipGraphExec_t> sge[3];
  std::vector<hipStream_t> lst;
  std::vector<hipEvent_t> lev;
  hipEvent_t ev0 = nullptr;
  int *dnt = nullptr, *dstat = nullptr;  // device: step number, 3 ints each
  int *hnt = nullptr, *hstat = nullptr;  // pinned host mirrors
  int maxiter = 0;
  // fused mode (begin and iteration phases): each suspension's phase is
  // captured in two halves around its mobility solve, and the solves of
  // all suspensions run in between as one launch per FCM stage per group
  // of equal launch shapes (FcmGroup), captured in one graph per phase
  bool fusable = false, fused = false;
  std::vector<hipGraphExec_t> spre[2], spost[2];
  hipGraphExec_t gfcm[2] = {nullptr, nullptr};
  std::vector<fs::FcmGroup> groups[2];
  hipEvent_t ev1 = nullptr;
};

namespace {

void batch_release(fs_batch *b) {
  for (int k = 0; k < 3; ++k) {
    if (b->ge[k]) (void)hipGraphExecDestroy(b->ge[k]);
    if (b->g[k]) (void)hipGraphDestroy(b->g[k]);
    for (hipGraphExec_t e : b->sge[k]) (void)hipGraphExecDestroy(e);
  }
  for (int k = 0; k < 2; ++k) {
    for (hipGraphExec_t e : b->spre[k]) (void)hipGraphExecDestroy(e);
    for (hipGraphExec_t e : b->spost[k]) (void)hipGraphExecDestroy(e);
    if (b->gfcm[k]) (void)hipGraphExecDestroy(b->gfcm[k]);
    for (fs::FcmGroup &g : b->groups[k]) g.release();
  }
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (hipStream_t s2 : b->lst) {
    (void)hipStreamSynchronize(s2);
    (void)hipStreamDestroy(s2);
  }
  for (hipEvent_t e : b->lev) (void)hipEventDestroy(e);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->st) (void)hipStreamDestroy(b->st);
  if (b->dnt) (void)hipFree(b->dnt);
  if (b->dstat) (void)hipFree(b->dstat);
  if (b->hnt) (void)hipHostFree(b->hnt);
  if (b->hstat) (void)hipHostFree(b->hstat);
  for (fs_system *h : b->sys) h->impl.set_device_control(nullptr, nullptr);
  delete b;
}

int batch_capture(fs_batch *b, int kind) {
  FS_HIP(hipGraphCreate(&b->g[kind], 0));
  for (size_t i = 0; i < b->sys.size(); ++i) {
    SystemImpl &S = b->sys[i]->impl;
    FS_HIP(hipStreamSynchronize(S.st));
    FS_CHECK(!S.prof, FS_ERR_STATE, "system %d: profiling is on", (int)i);
    S.set_device_control(nullptr, b->dnt);
    FS_HIP(hipStreamBeginCapture(S.st, hipStreamCaptureModeThreadLocal));
    int *status = b->dstat + 3 * i;
    int rc = kind == 0   ? S.begin_step_dev(status)
             : kind == 1 ? S.iteration_dev(status)
                         : S.end_step_dev(status);
    hipGraph_t child = nullptr;
    hipError_t e = hipStreamEndCapture(S.st, &child);
    S.set_device_control(nullptr, nullptr);
    if (rc != FS_OK) {
      if (child) (void)hipGraphDestroy(child);
      return rc;
    }
    FS_CHECK(e == hipSuccess && child, FS_ERR_HIP, "graph capture failed: %s",
             hipGetErrorString(e));
    hipGraphNode_t node;
    hipError_t e2 = hipGraphAddChildGraphNode(&node, b->g[kind], nullptr, 0, child);
    hipGraphExec_t ex = nullptr;
    hipError_t e3 = hipGraphInstantiate(&ex, child, nullptr, nullptr, 0);
    (void)hipGraphDestroy(child);
    FS_CHECK(e2 == hipSuccess && e3 == hipSuccess, FS_ERR_HIP,
             "adding a child graph failed: %s",
             hipGetErrorString(e2 != hipSuccess ? e2 : e3));
    b->sge[kind].push_back(ex);
  }
  FS_HIP(hipGraphInstantiate(&b->ge[kind], b->g[kind], nullptr, nullptr, 0));
  return FS_OK;
}

// captures one half (part 0: before the mobility solve, 1: after it) of
// phase `kind` (0 begin, 1 iteration) of every suspension
int batch_capture_half(fs_batch *b, int kind, int part) {
  auto &out = part == 0 ? b->spre[kind] : b->spost[kind];
  for (size_t i = 0; i < b->sys.size(); ++i) {
    SystemImpl &S = b->sys[i]->impl;
    FS_HIP(hipStreamSynchronize(S.st));
    S.set_device_control(nullptr, b->dnt);
    FS_HIP(hipStreamBeginCapture(S.st, hipStreamCaptureModeThreadLocal));
    int *status = b->dstat + 3 * i;
    int rc = kind == 0 ? S.begin_step_dev(status, part)
                       : S.iteration_dev(status, part);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(S.st, &gr);
    S.set_device_control(nullptr, nullptr);
    if (rc != FS_OK) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    FS_CHECK(e == hipSuccess && gr, FS_ERR_HIP, "graph capture failed: %s",
             hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    hipError_t e2 = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    FS_CHECK(e2 == hipSuccess, FS_ERR_HIP, "graph instantiation failed: %s",
             hipGetErrorString(e2));
    out.push_back(ex);
  }
  return FS_OK;
}

// groups of equal FCM launch shapes for phase `kind` (loop control: none
// in the begin phase, each suspension's iterating flag in the iteration
// phase), and the graph that solves them all on b->st
int batch_build_fcm(fs_batch *b, int kind) {
  auto &groups = b->groups[kind];
  for (fs_system *h : b->sys) {
    SystemImpl &S = h->impl;
    fs::FcmGroupMember m;
    m.fcm = &S.fcm;
    m.act = kind == 0 ? nullptr : S.ctl.as<int>() + 1;
    m.np = S.np;
    m.Y = S.d.Ya;
    m.F = S.d.Fa;
    m.T = S.d.Ta;
    m.V = S.d.V;
    m.W = S.d.W;
    fs::FcmGroup *g = nullptr;
    for (fs::FcmGroup &x : groups)
      if (fs::fcm_same_shapes(*x.m[0].fcm, S.fcm)) g = &x;
    if (!g) {
      groups.emplace_back();
      g = &groups.back();
    }
    g->m.push_back(m);
  }
  for (fs::FcmGroup &g : groups) FS_TRY(g.build());
  FS_HIP(hipStreamSynchronize(b->st));
  FS_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  int rc = FS_OK;
  for (fs::FcmGroup &g : groups)
    if (rc == FS_OK) rc = g.solve(b->st);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(b->st, &gr);
  if (rc != FS_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  FS_CHECK(e == hipSuccess && gr, FS_ERR_HIP, "graph capture failed: %s",
           hipGetErrorString(e));
  hipError_t e2 = hipGraphInstantiate(&b->gfcm[kind], gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  FS_CHECK(e2 == hipSuccess, FS_ERR_HIP, "graph instantiation failed: %s",
           hipGetErrorString(e2));
  return FS_OK;
}

// whether every suspension can join a fused FCM group: the matrix-core
// tile path, and the FCM mobility (not the friction start-up) in every step
bool batch_fusable(fs_batch *b) {
  for (fs_system *h : b->sys) {
    const SystemImpl &S = h->impl;
    if (!S.fcm.mfma_spread || S.p.initial_friction_steps != 0) return false;
  }
  return true;
}

int batch_status(fs_batch *b) {
  FS_HIP(hipMemcpyAsync(b->hstat, b->dstat, 3 * sizeof(int) * b->sys.size(),
                        hipMemcpyDeviceToHost, b->st));
  FS_HIP(hipStreamSynchronize(b->st));
  return FS_OK;
}

}  // namespace

extern "C" int fs_batch_set_lanes(fs_batch *b, int lanes);

extern "C" int fs_batch_create(fs_system **sys, int n, fs_batch **out) {
  FS_CHECK(sys && out && n >= 1, FS_ERR_ARG, "bad argument");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    FS_CHECK(sys[i], FS_ERR_ARG, "null system %d", i);
    FS_CHECK(!sys[i]->ws.attached(), FS_ERR_STATE,
             "system %d has a warm start attached or recording", i);
    FS_CHECK(sys[i]->impl.p.initial_friction_steps == 0, FS_ERR_ARG,
             "system %d: batched sweeps need initial_friction_steps == 0", i);
    FS_CHECK(sys[i]->impl.hist_cap == 0, FS_ERR_STATE,
             "system %d carries its Broyden history across steps (batched "
             "sweeps restart it every step)", i);
    for (int j = 0; j < i; ++j)
      FS_CHECK(sys[j] != sys[i], FS_ERR_ARG, "system %d listed twice", i);
  }
  fs_batch *b = new fs_batch;
  b->sys.assign(sys, sys + n);
  int rc = FS_OK;
  auto fail = [&](int r) {
    batch_release(b);
    return r;
  };
  if (hipStreamCreate(&b->st) != hipSuccess ||
      hipMalloc((void **)&b->dnt, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&b->dstat, 3 * sizeof(int) * n) != hipSuccess ||
      hipHostMalloc((void **)&b->hnt, sizeof(int), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&b->hstat, 3 * sizeof(int) * n,
                    hipHostMallocDefault) != hipSuccess) {
    fs::set_error("batch allocation failed");
    return fail(FS_ERR_HIP);
  }
  for (int i = 0; i < n; ++i)
    b->maxiter = std::max(b->maxiter, sys[i]->impl.p.broyden_maxiter);
  for (int k = 0; k < 3 && rc == FS_OK; ++k) rc = batch_capture(b, k);
  if (rc != FS_OK) return fail(rc);
  b->fusable = batch_fusable(b);
  if (b->fusable) {
    for (int k = 0; k < 2 && rc == FS_OK; ++k) {
      rc = batch_capture_half(b, k, 0);
      if (rc == FS_OK) rc = batch_capture_half(b, k, 1);
      if (rc == FS_OK) rc = batch_build_fcm(b, k);
    }
    if (rc != FS_OK) return fail(rc);
  }
  // measured on a 64-suspension sweep: 4 streams 2.4x sequential stepping,
  // one parent graph 0.8x (its child graphs ran one after another)
  rc = fs_batch_set_lanes(b, std::min(n, 4));
  if (rc != FS_OK) return fail(rc);
  for (int i = 0; i < n; ++i) ++sys[i]->batches;
  *out = b;
  return FS_OK;
}

extern "C" int fs_batch_destroy(fs_batch *b) {
  if (!b) return FS_OK;
  for (fs_system *s : b->sys) --s->batches;
  batch_release(b);
  return FS_OK;
}

namespace {

// one phase for every suspension, ordered after everything already on b->st
// and before what follows on it
int batch_phase(fs_batch *b, int kind) {
  if (b->fused && kind < 2) {
    // halves before the solves on the lanes, the fused FCM graph on b->st,
    // halves after the solves on the lanes
    const int lanes = std::max(1, b->lanes);
    FS_HIP(hipEventRecord(b->ev0, b->st));
    for (int l = 0; l < lanes; ++l) FS_HIP(hipStreamWaitEvent(b->lst[l], b->ev0, 0));
    for (size_t i = 0; i < b->sys.size(); ++i)
      FS_HIP(hipGraphLaunch(b->spre[kind][i], b->lst[i % lanes]));
    for (int l = 0; l < lanes; ++l) {
      FS_HIP(hipEventRecord(b->lev[l], b->lst[l]));
      FS_HIP(hipStreamWaitEvent(b->st, b->lev[l], 0));
    }
    FS_HIP(hipGraphLaunch(b->gfcm[kind], b->st));
    FS_HIP(hipEventRecord(b->ev1, b->st));
    for (int l = 0; l < lanes; ++l) FS_HIP(hipStreamWaitEvent(b->lst[l], b->ev1, 0));
    for (size_t i = 0; i < b->sys.size(); ++i)
      FS_HIP(hipGraphLaunch(b->spost[kind][i], b->lst[i % lanes]));
    for (int l = 0; l < lanes; ++l) {
      FS_HIP(hipEventRecord(b->lev[l], b->lst[l]));
      FS_HIP(hipStreamWaitEvent(b->st, b->lev[l], 0));
    }
    return FS_OK;
  }
  if (b->lanes == 0) {
    FS_HIP(hipGraphLaunch(b->ge[kind], b->st));
    return FS_OK;
  }
  FS_HIP(hipEventRecord(b->ev0, b->st));
  for (int l = 0; l < b->lanes; ++l)
    FS_HIP(hipStreamWaitEvent(b->lst[l], b->ev0, 0));
  for (size_t i = 0; i < b->sys.size(); ++i)
    FS_HIP(hipGraphLaunch(b->sge[kind][i], b->lst[i % b->lanes]));
  for (int l = 0; l < b->lanes; ++l) {
    FS_HIP(hipEventRecord(b->lev[l], b->lst[l]));
    FS_HIP(hipStreamWaitEvent(b->st, b->lev[l], 0));
  }
  return FS_OK;
}

}  // namespace

extern "C" int fs_batch_set_fused(fs_batch *b, int on) {
  FS_CHECK(b, FS_ERR_ARG, "null batch");
  FS_CHECK(!on || b->fusable, FS_ERR_ARG,
           "fused FCM stages need every suspension on the matrix-core tile "
           "path and with initial_friction_steps == 0");
  FS_HIP(hipStreamSynchronize(b->st));
  b->fused = on != 0;
  return FS_OK;
}

extern "C" int fs_batch_set_lanes(fs_batch *b, int lanes) {
  FS_CHECK(b && lanes >= 0 && lanes <= 16, FS_ERR_ARG,
           "lanes must be in [0, 16]");
  FS_HIP(hipStreamSynchronize(b->st));
  const int want = std::max(lanes, 1);  // the fused mode uses at least one
  while ((int)b->lst.size() < want) {
    hipStream_t s2 = nullptr;
    hipEvent_t e = nullptr;
    FS_HIP(hipStreamCreate(&s2));
    b->lst.push_back(s2);
    FS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->lev.push_back(e);
  }
  if (!b->ev0) FS_HIP(hipEventCreateWithFlags(&b->ev0, hipEventDisableTiming));
  if (!b->ev1) FS_HIP(hipEventCreateWithFlags(&b->ev1, hipEventDisableTiming));
  b->lanes = lanes;
  return FS_OK;
}

extern "C" int fs_batch_timestep(fs_batch *b, int nt, int *iters) {
  FS_CHECK(b, FS_ERR_ARG, "null batch");
  const int n = (int)b->sys.size();
  *b->hnt = nt;
  FS_HIP(hipMemcpyAsync(b->dnt, b->hnt, sizeof(int), hipMemcpyHostToDevice,
                        b->st));
  FS_TRY(batch_phase(b, 0));
  FS_TRY(batch_status(b));
  for (int i = 0; i < n; ++i)
    FS_CHECK(b->hstat[3 * i + 2] == 0, FS_ERR_ARG,
             "system %d: approximate Jacobian is singular", i);
  for (int it = 0; it < b->maxiter; ++it) {
    bool any = false;
    for (int i = 0; i < n; ++i) any = any || b->hstat[3 * i + 1] != 0;
    if (!any) break;
    FS_TRY(batch_phase(b, 1));
    FS_TRY(batch_status(b));
  }
  FS_TRY(batch_phase(b, 2));
  FS_HIP(hipStreamSynchronize(b->st));
  if (iters)
    for (int i = 0; i < n; ++i) iters[i] = b->hstat[3 * i];
  return FS_OK;
}

extern "C" int fs_batch_size(fs_batch *b, int *n) {
  FS_CHECK(b && n, FS_ERR_ARG, "bad argument");
  *n = (int)b->sys.size();
  return FS_OK;
}
