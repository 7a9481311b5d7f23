// io.cpp -- trajectory output and restart files of a filament system.
//
// Text formats follow the reference byte for byte in layout (space
// separated, fixed notation with 16 decimals):
//   .dat  append "nt " + per bead Xs(3) q(4) + "\n"
//         (Filament::writeData, Filament.hpp:547-562; save_data_to_file,
//         file_functions.cpp:150-161)
//   .bak  overwrite "nt " + per filament: per bead X Xt q qt U Ut (20
//         values), then per link lam lam1 lam2 (9 values), no newline
//         (writeBackupData, Filament.hpp:658-687; save_backup_data_to_file,
//         file_functions.cpp:211-222); read back as continue_from_checkpoint
//         does (Filament.hpp:691-759, file_functions.cpp:228-319): Xs = X,
//         qs = q, Us = U
//   .par  header line, value line, blank, "BeadNumbers", counts, blank,
//         "BendingModuli are KAP*", factors (save_parameter_values_to_file,
//         file_functions.cpp:6-81); the restart reader takes the filament
//         count (3rd value of line 2) and the bead numbers (line 5)
// The binary checkpoint (.fsck) stores every state array bit for bit, which
// the 16-decimal text cannot.
// The swimmer files (velocity, collision) and the force file follow
// save_swimmer_velocity_data_to_file / save_swimmer_collision_data_to_file /
// save_force_data_to_file and their column-name writers.
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "common.hpp"
// filsim.h comes with common.hpp

namespace {

struct File {
  FILE *f = nullptr;
  File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// one value as the reference's ostream with precision(16), fixed, showpoint
inline void put(FILE *f, double v) { std::fprintf(f, "%.16f ", v); }

// next space/newline separated token as a double (stod semantics: leading
// whitespace skipped, whole token must be a number)
bool get_double(FILE *f, double *v) {
  char buf[128];
  if (std::fscanf(f, "%127s", buf) != 1) return false;
  char *end = nullptr;
  errno = 0;
  *v = std::strtod(buf, &end);
  return end && *end == '\0' && errno == 0;
}

int check_layout(int nf, const int *nw) {
  FS_CHECK(nf >= 1 && nw, FS_ERR_ARG, "bad filament layout");
  for (int f = 0; f < nf; ++f)
    FS_CHECK(nw[f] >= 2, FS_ERR_ARG, "filament %d has %d segments", f, nw[f]);
  return FS_OK;
}

}  // namespace

extern "C" int fs_io_write_dat(const char *path, int nt, int nf, const int *nw,
                               const double *X, const double *q) {
  FS_CHECK(path && X && q, FS_ERR_ARG, "null argument");
  FS_TRY(check_layout(nf, nw));
  File o(path, "a");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f, "%d ", nt);
  size_t b = 0;
  for (int f = 0; f < nf; ++f)
    for (int i = 0; i < nw[f]; ++i, ++b) {
      for (int d = 0; d < 3; ++d) put(o.f, X[3 * b + d]);
      for (int c = 0; c < 4; ++c) put(o.f, q[4 * b + c]);
    }
  std::fprintf(o.f, "\n");
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

// ---- swimmer and force output (file_functions.cpp:102-145, 164-208;
// Filament.hpp:509-650). The column names go to a new file; each data call
// appends "nt " and one row. At nt = 0 the reference writes the integer 0
// (no decimals) in place of every velocity and force.
namespace {

int write_names(const char *path, int nw0,
                std::initializer_list<std::initializer_list<const char *>> g) {
  FS_CHECK(path, FS_ERR_ARG, "null path");
  FS_CHECK(nw0 >= 1, FS_ERR_ARG, "bad segment count %d", nw0);
  File o(path, "w");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f, "it ");
  for (const auto &names : g)
    for (int i = 0; i < nw0; ++i)
      for (const char *n : names) std::fprintf(o.f, "%s ", n);
  std::fprintf(o.f, "\n");
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

void zeros(FILE *f, size_t n) {
  for (size_t k = 0; k < n; ++k) std::fprintf(f, "0 ");
}

}  // namespace

extern "C" int fs_io_write_swimmer_velocity_names(const char *path, int nw0) {
  return write_names(path, nw0,
                     {{"X", "Y", "Z", "qr", "qi", "qj", "qk"},
                      {"Vx", "Vy", "Vz", "Wx", "Wy", "Wz", "Fx", "Fy", "Fz"}});
}

extern "C" int fs_io_write_swimmer_collision_names(const char *path,
                                                   int nw0) {
  return write_names(path, nw0, {{"Fx", "Fy", "Fz"}});
}

extern "C" int fs_io_write_force_names(const char *path, int nw0) {
  return write_names(path, nw0, {{"Fx", "Fy", "Fz", "Tx", "Ty", "Tz"}});
}

extern "C" int fs_io_write_swimmer_velocity(const char *path, int nt, int nw0,
                                            const double *Xs, const double *qs,
                                            const double *V, const double *W,
                                            const double *F) {
  FS_CHECK(path && Xs && qs, FS_ERR_ARG, "null argument");
  FS_CHECK(nt == 0 || (V && W && F), FS_ERR_ARG, "null V, W or F");
  FS_CHECK(nw0 >= 1, FS_ERR_ARG, "bad segment count %d", nw0);
  File o(path, "a");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f, "%d ", nt);
  for (int i = 0; i < nw0; ++i) {
    for (int d = 0; d < 3; ++d) put(o.f, Xs[3 * i + d]);
    for (int c = 0; c < 4; ++c) put(o.f, qs[4 * i + c]);
  }
  if (nt != 0) {
    for (int i = 0; i < nw0; ++i)
      for (const double *a : {V, W, F})
        for (int d = 0; d < 3; ++d) put(o.f, a[3 * i + d]);
  } else {
    zeros(o.f, 9 * (size_t)nw0);
  }
  std::fprintf(o.f, "\n");
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

extern "C" int fs_io_write_swimmer_collision(const char *path, int nt, int nw0,
                                             const double *F_total) {
  FS_CHECK(path && (nt == 0 || F_total), FS_ERR_ARG, "null argument");
  FS_CHECK(nw0 >= 1, FS_ERR_ARG, "bad segment count %d", nw0);
  File o(path, "a");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f, "%d ", nt);
  if (nt != 0) {
    for (int i = 0; i < 3 * nw0; ++i) put(o.f, F_total[i]);
  } else {
    zeros(o.f, 3 * (size_t)nw0);
  }
  std::fprintf(o.f, "\n");
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

// Filament::writeForceData reads rows forceLengthFilled / 6 + i, where
// forceLengthFilled is the filament's first segment (Filament.cpp:3-5,
// filament_initialisation_functions.cpp:52): filament 0 gets its own rows,
// filament f > 0 the nw[f] rows from (its first segment) / 6 on, as the
// reference writes them.
extern "C" int fs_io_write_forces(const char *path, int nt, int nf,
                                  const int *nw, const double *F,
                                  const double *T) {
  FS_CHECK(path && (nt == 0 || (F && T)), FS_ERR_ARG, "null argument");
  FS_TRY(check_layout(nf, nw));
  File o(path, "a");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f, "%d ", nt);
  size_t off = 0;
  for (int f = 0; f < nf; ++f) {
    if (nt != 0) {
      const size_t r0 = off / 6;
      for (int i = 0; i < nw[f]; ++i) {
        for (int d = 0; d < 3; ++d) put(o.f, F[3 * (r0 + i) + d]);
        for (int d = 0; d < 3; ++d) put(o.f, T[3 * (r0 + i) + d]);
      }
    } else {
      zeros(o.f, 6 * (size_t)nw[f]);
    }
    off += nw[f];
  }
  std::fprintf(o.f, "\n");
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

extern "C" int fs_io_write_bak(const char *path, int nt, int nf, const int *nw,
                               const fs_state_ptrs *s) {
  FS_CHECK(path && s && s->X && s->Xt && s->q && s->qt && s->U && s->Ut &&
               s->lam && s->lam1 && s->lam2,
           FS_ERR_ARG, "null argument");
  FS_TRY(check_layout(nf, nw));
  File o(path, "w");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f, "%d ", nt);
  size_t b = 0, l = 0;
  for (int f = 0; f < nf; ++f) {
    for (int i = 0; i < nw[f]; ++i, ++b) {
      for (int d = 0; d < 3; ++d) put(o.f, s->X[3 * b + d]);
      for (int d = 0; d < 3; ++d) put(o.f, s->Xt[3 * b + d]);
      for (int c = 0; c < 4; ++c) put(o.f, s->q[4 * b + c]);
      for (int c = 0; c < 4; ++c) put(o.f, s->qt[4 * b + c]);
      for (int d = 0; d < 3; ++d) put(o.f, s->U[3 * b + d]);
      for (int d = 0; d < 3; ++d) put(o.f, s->Ut[3 * b + d]);
    }
    for (int i = 0; i < nw[f] - 1; ++i, ++l) {
      for (int d = 0; d < 3; ++d) put(o.f, s->lam[3 * l + d]);
      for (int d = 0; d < 3; ++d) put(o.f, s->lam1[3 * l + d]);
      for (int d = 0; d < 3; ++d) put(o.f, s->lam2[3 * l + d]);
    }
  }
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

extern "C" int fs_io_read_bak(const char *path, int nf, const int *nw,
                              fs_state_ptrs *s, int *nt) {
  FS_CHECK(path && s && nt && s->X && s->Xt && s->q && s->qt && s->U &&
               s->Ut && s->lam && s->lam1 && s->lam2,
           FS_ERR_ARG, "null argument");
  FS_TRY(check_layout(nf, nw));
  File in(path, "r");
  FS_CHECK(in.f, FS_ERR_ARG, "cannot open %s", path);
  double v = 0;
  FS_CHECK(get_double(in.f, &v) && v == (double)(int)v, FS_ERR_ARG,
           "%s: bad step number", path);
  *nt = (int)v;
  size_t b = 0, l = 0;
  auto rd = [&](double *dst, int n) -> bool {
    for (int k = 0; k < n; ++k)
      if (!get_double(in.f, dst + k)) return false;
    return true;
  };
  for (int f = 0; f < nf; ++f) {
    for (int i = 0; i < nw[f]; ++i, ++b) {
      bool ok = rd(s->X + 3 * b, 3) && rd(s->Xt + 3 * b, 3) &&
                rd(s->q + 4 * b, 4) && rd(s->qt + 4 * b, 4) &&
                rd(s->U + 3 * b, 3) && rd(s->Ut + 3 * b, 3);
      FS_CHECK(ok, FS_ERR_ARG, "%s: truncated at filament %d bead %d", path,
               f, i);
      if (s->Xs) std::memcpy(s->Xs + 3 * b, s->X + 3 * b, 3 * sizeof(double));
      if (s->qs) std::memcpy(s->qs + 4 * b, s->q + 4 * b, 4 * sizeof(double));
      if (s->Us) std::memcpy(s->Us + 3 * b, s->U + 3 * b, 3 * sizeof(double));
    }
    for (int i = 0; i < nw[f] - 1; ++i, ++l) {
      bool ok = rd(s->lam + 3 * l, 3) && rd(s->lam1 + 3 * l, 3) &&
                rd(s->lam2 + 3 * l, 3);
      FS_CHECK(ok, FS_ERR_ARG, "%s: truncated at filament %d link %d", path,
               f, i);
    }
  }
  return FS_OK;
}

extern "C" int fs_io_write_par(const char *path, const char *name,
                               const char *tag, int nf, const int *nw,
                               const double *bending_factors,
                               const fs_sim_params *p, double kap, double c,
                               int time_steps, int plot_steps) {
  FS_CHECK(path && p, FS_ERR_ARG, "null argument");
  FS_TRY(check_layout(nf, nw));
  File o(path, "w");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f,
               "SimulationName SimulationTag Nsw Nworm a ah DL Sp4 Kap C MU "
               "curvature K0 TimeSteps StepsPerPeriod TOL gmres_tol "
               "broyden_maxiter gmres_maxiter LfcmBox_x plot_steps "
               "STRAINTWISTX STRAINTWISTY STRAINTWISTZ Kap_NetworkFilaments "
               "LfcmBox_z LfcmBox_y B StepsPerSettlingTime dt "
               "initialFrictionSteps EnableSpringLinks "
               "ConnectedNodesNumSwimmingFilaments ConnectedNodesNumNodes "
               "ConnectedNodesConnectionRadius ConnectedNodesMinNodeSpacing "
               "ConnectedNodesSpacingAwayFromNode NetworkSpringConstant "
               "NetworkSpringNaturalLength \n");
  int nworm = 0;
  for (int f = 0; f < nf; ++f) nworm = nworm > nw[f] ? nworm : nw[f];
  const double a = p->fcm.a;
  // fields with no counterpart here (swimmer number Sp4, curvature, K0,
  // steps per period, spring links, networks) are written as 0
  std::fprintf(o.f,
               "%s %s %d %d %g %g %g %g %g %g %g %g %g %d %g %g %d %d %d %g "
               "%d %g %g %g %g %g %g %g %g %g %d %d %d %d %g %g %g %g %g \n\n",
               name ? name : "filsim", tag ? tag : "run", nf, nworm, a, a,
               p->dL, 0.0, kap, c, p->fcm.mu, 0.0, 0.0, time_steps, 0.0,
               p->tol, 0 /* gmres_tol: the int 0, file_functions.cpp:42 */,
               p->broyden_maxiter, 0,
               p->fcm.nx * p->fcm.dx, plot_steps, 0.0, 0.0, p->strain_twist0,
               0.0,
               p->fcm.nz * p->fcm.dx, p->fcm.ny * p->fcm.dx, 0.0, 0.0, p->dt,
               p->initial_friction_steps, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0);
  std::fprintf(o.f, "BeadNumbers\n");
  for (int f = 0; f < nf; ++f) std::fprintf(o.f, f + 1 < nf ? "%d " : "%d", nw[f]);
  std::fprintf(o.f, "\n\nBendingModuli are KAP*\n");
  for (int f = 0; f < nf; ++f)
    std::fprintf(o.f, f + 1 < nf ? "%g " : "%g",
                 bending_factors ? bending_factors[f] : 1.0);
  std::fprintf(o.f, "\n");
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

// Filament count and bead numbers from a .par file (as the restart reader
// of file_functions.cpp:242-280). nw may be NULL to query the count.
extern "C" int fs_io_read_par(const char *path, int *nf, int *nw, int cap) {
  FS_CHECK(path && nf, FS_ERR_ARG, "null argument");
  File in(path, "r");
  FS_CHECK(in.f, FS_ERR_ARG, "cannot open %s", path);
  std::vector<std::string> lines;
  char buf[1 << 16];
  while (std::fgets(buf, sizeof buf, in.f)) {
    std::string s(buf);
    FS_CHECK(s.empty() || s.back() == '\n' || std::feof(in.f), FS_ERR_ARG,
             "%s: line too long", path);
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
    lines.push_back(s);
  }
  FS_CHECK(lines.size() >= 5, FS_ERR_ARG, "%s: too few lines", path);
  char n1[256], n2[256];
  int count = 0;
  FS_CHECK(std::sscanf(lines[1].c_str(), "%255s %255s %d", n1, n2, &count) == 3 &&
               count >= 1,
           FS_ERR_ARG, "%s: bad filament count", path);
  *nf = count;
  if (!nw) return FS_OK;
  FS_CHECK(cap >= count, FS_ERR_SIZE, "%s: %d filaments, room for %d", path,
           count, cap);
  const char *p = lines[4].c_str();
  for (int f = 0; f < count; ++f) {
    char *end = nullptr;
    long v = std::strtol(p, &end, 10);
    FS_CHECK(end != p && v >= 2, FS_ERR_ARG, "%s: bad bead number %d", path, f);
    nw[f] = (int)v;
    p = end;
  }
  return FS_OK;
}

// ---- spring links (spring_links::save_spring_links_to_file and
// read_spring_links_from_file, spring_link_functions.cpp:50-101): a header
// line, the link count, then one "A B C D" row per link (filament A bead B
// to filament C bead D).
extern "C" int fs_io_write_spring_links(const char *path, int n,
                                        const int *links) {
  FS_CHECK(path && n >= 0 && (n == 0 || links), FS_ERR_ARG, "bad argument");
  File o(path, "w");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  std::fprintf(o.f,
               "Spring link data file. Number of links written first, then "
               "springs between {filament A, bead B} and {filament C, bead D} "
               "in the form 'A B C D'.\n%d\n",
               n);
  for (int l = 0; l < n; ++l)
    std::fprintf(o.f, "%d %d %d %d\n", links[4 * l], links[4 * l + 1],
                 links[4 * l + 2], links[4 * l + 3]);
  FS_CHECK(!std::ferror(o.f), FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

// links may be NULL to query the count.
extern "C" int fs_io_read_spring_links(const char *path, int *n, int *links,
                                       int cap) {
  FS_CHECK(path && n, FS_ERR_ARG, "null argument");
  File in(path, "r");
  FS_CHECK(in.f, FS_ERR_ARG, "cannot open %s", path);
  int c;
  while ((c = std::fgetc(in.f)) != EOF && c != '\n') {
  }
  int count = -1;
  FS_CHECK(c == '\n' && std::fscanf(in.f, "%d", &count) == 1 && count >= 0,
           FS_ERR_ARG, "%s: bad link count", path);
  *n = count;
  if (!links) return FS_OK;
  FS_CHECK(cap >= count, FS_ERR_SIZE, "%s: %d links, room for %d", path, count,
           cap);
  for (int l = 0; l < count; ++l)
    FS_CHECK(std::fscanf(in.f, "%d %d %d %d", links + 4 * l, links + 4 * l + 1,
                         links + 4 * l + 2, links + 4 * l + 3) == 4,
             FS_ERR_ARG, "%s: bad link %d", path, l);
  return FS_OK;
}

// ---- binary checkpoint: "FSCK", version, nt, nf, np, nlam, nw[nf], then
// X Xs Xt U Us Ut (np*3 each), q qs qt (np*4 each), lam lam1 lam2 (nlam*3
// each) as raw little-endian doubles
namespace {
constexpr uint32_t kMagic = 0x4b435346u;  // "FSCK"
constexpr uint32_t kVersion = 1;
}  // namespace

extern "C" int fs_io_write_checkpoint(const char *path, int nt, int nf,
                                      const int *nw, const fs_state_ptrs *s) {
  FS_CHECK(path && s, FS_ERR_ARG, "null argument");
  FS_TRY(check_layout(nf, nw));
  size_t np = 0;
  for (int f = 0; f < nf; ++f) np += nw[f];
  const size_t nl = np - nf;
  const double *arr3[6] = {s->X, s->Xs, s->Xt, s->U, s->Us, s->Ut};
  const double *arr4[3] = {s->q, s->qs, s->qt};
  const double *arrl[3] = {s->lam, s->lam1, s->lam2};
  for (auto a : arr3) FS_CHECK(a, FS_ERR_ARG, "null state array");
  for (auto a : arr4) FS_CHECK(a, FS_ERR_ARG, "null state array");
  for (auto a : arrl) FS_CHECK(a, FS_ERR_ARG, "null state array");
  File o(path, "wb");
  FS_CHECK(o.f, FS_ERR_ARG, "cannot open %s", path);
  const int64_t hdr[6] = {kMagic, kVersion, nt, nf, (int64_t)np, (int64_t)nl};
  bool ok = std::fwrite(hdr, sizeof hdr, 1, o.f) == 1 &&
            std::fwrite(nw, sizeof(int), nf, o.f) == (size_t)nf;
  for (auto a : arr3) ok = ok && std::fwrite(a, 8, 3 * np, o.f) == 3 * np;
  for (auto a : arr4) ok = ok && std::fwrite(a, 8, 4 * np, o.f) == 4 * np;
  for (auto a : arrl) ok = ok && std::fwrite(a, 8, 3 * nl, o.f) == 3 * nl;
  FS_CHECK(ok, FS_ERR_ARG, "write to %s failed", path);
  return FS_OK;
}

extern "C" int fs_io_read_checkpoint(const char *path, int nf, const int *nw,
                                     fs_state_ptrs *s, int *nt) {
  FS_CHECK(path && s && nt, FS_ERR_ARG, "null argument");
  FS_TRY(check_layout(nf, nw));
  File in(path, "rb");
  FS_CHECK(in.f, FS_ERR_ARG, "cannot open %s", path);
  int64_t hdr[6];
  FS_CHECK(std::fread(hdr, sizeof hdr, 1, in.f) == 1 && hdr[0] == kMagic &&
               hdr[1] == kVersion,
           FS_ERR_ARG, "%s: not a checkpoint of this version", path);
  FS_CHECK(hdr[3] == nf, FS_ERR_ARG, "%s: %" PRId64 " filaments, system has %d",
           path, hdr[3], nf);
  std::vector<int> fnw(nf);
  FS_CHECK(std::fread(fnw.data(), sizeof(int), nf, in.f) == (size_t)nf,
           FS_ERR_ARG, "%s: truncated", path);
  size_t np = 0;
  for (int f = 0; f < nf; ++f) {
    FS_CHECK(fnw[f] == nw[f], FS_ERR_ARG,
             "%s: filament %d has %d segments, system has %d", path, f, fnw[f],
             nw[f]);
    np += nw[f];
  }
  const size_t nl = np - nf;
  FS_CHECK(hdr[4] == (int64_t)np && hdr[5] == (int64_t)nl, FS_ERR_ARG,
           "%s: size mismatch", path);
  double *arr3[6] = {s->X, s->Xs, s->Xt, s->U, s->Us, s->Ut};
  double *arr4[3] = {s->q, s->qs, s->qt};
  double *arrl[3] = {s->lam, s->lam1, s->lam2};
  for (auto a : arr3) FS_CHECK(a, FS_ERR_ARG, "null state array");
  for (auto a : arr4) FS_CHECK(a, FS_ERR_ARG, "null state array");
  for (auto a : arrl) FS_CHECK(a, FS_ERR_ARG, "null state array");
  bool ok = true;
  for (auto a : arr3) ok = ok && std::fread(a, 8, 3 * np, in.f) == 3 * np;
  for (auto a : arr4) ok = ok && std::fread(a, 8, 4 * np, in.f) == 4 * np;
  for (auto a : arrl) ok = ok && std::fread(a, 8, 3 * nl, in.f) == 3 * nl;
  FS_CHECK(ok, FS_ERR_ARG, "%s: truncated", path);
  *nt = (int)hdr[2];
  return FS_OK;
}
