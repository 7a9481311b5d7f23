// filsim_shim.hpp -- header-only C++ layer over filsim.h that keeps the
// reference's call surface, so the driver (main.cpp) swaps its solver by
// changing includes, not call sites.
//
//   filsim::Fcm          struct Fcm          (FCMfunctions.hpp:7-59)
//   filsim::LinkedList   struct LinkedList   (CollisionBarrierFilament.hpp:19-57)
//   filsim::Filaments    the filaments vector + filamentJsolver + the
//                        per-filament calls of main.cpp:334-352, 518-547,
//                        640-663, 704-707 (FilamentJacobianSolver.hpp,
//                        Filament.hpp), batched over every filament
//   filsim::ErrorCheckFCM (MobilityConstraints.cpp:447)
//   filsim::spring_links struct spring_links (spring_link_functions.hpp:18-39)
//   filsim::save_*_to_file  the trajectory, backup, swimmer and force
//                        writers of file_functions.cpp
//
// Differences from the reference, all forced by running on one GPU:
//   - sizes are runtime parameters (fs_fcm_params / fs_sim_params) instead
//     of config.hpp macros; Np is fixed at initialise;
//   - mpi_broadcast() is a no-op (one process owns the device state);
//   - errors throw filsim::Error (the reference calls MPI_Abort or prints).
// The host arrays Y/F/T/V/W are double** over one contiguous Np x 3 block,
// like double_array_2D (c_array_functions.cpp:12-21).
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "filsim.h"

namespace filsim {

struct Error : std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

inline void check(int rc, const char *what) {
  if (rc != FS_OK)
    throw Error(std::string(what) + ": " + fs_last_error());
}

// Np x 3 contiguous array with double** row pointers.
struct Array2D {
  std::vector<double> data;
  std::vector<double *> rows;
  void allocate(int n, int m) {
    data.assign((size_t)n * m, 0.0);
    rows.resize(n);
    for (int i = 0; i < n; ++i) rows[i] = data.data() + (size_t)i * m;
  }
  double **ptr() { return rows.data(); }
};

// struct Fcm: host-array interface; each solve copies Y/F/T in and V/W out.
struct Fcm {
  double **Y = nullptr, **F = nullptr, **T = nullptr, **V = nullptr,
         **W = nullptr;
  int Np = 0;
  fs_fcm_params params{};
  int precision = FS_FP64;
  fs_fcm *h = nullptr;
  Array2D aY, aF, aT, aV, aW;

  Fcm() = default;
  Fcm(const fs_fcm_params &p, int np, int prec = FS_FP64)
      : Np(np), params(p), precision(prec) {}

  // Fcm::initialise (FCMfunctions.cpp:10-168)
  void initialise(int myrank = 0) {
    (void)myrank;
    check(fs_fcm_create(&params, Np, precision, &h), "fs_fcm_create");
    aY.allocate(Np, 3);
    aF.allocate(Np, 3);
    aT.allocate(Np, 3);
    aV.allocate(Np, 3);
    aW.allocate(Np, 3);
    Y = aY.ptr();
    F = aF.ptr();
    T = aT.ptr();
    V = aV.ptr();
    W = aW.ptr();
  }
  // Fcm::mobility_solve (FCMfunctions.cpp:175-464)
  void mobility_solve() {
    check(fs_fcm_mobility_host(h, Np, &Y[0][0], &F[0][0], &T[0][0], &V[0][0],
                               &W[0][0]),
          "fs_fcm_mobility_host");
  }
  // Fcm::friction_mobility_solve (FCMfunctions.cpp:865-880)
  void friction_mobility_solve() {
    const double PI = 3.14159265358979323846264338327950288;
    double a1 = 1. / (6 * PI * params.a * params.mu);
    double a2 = 1. / (8 * PI * params.a * params.a * params.a * params.mu);
    for (int i = 0; i < Np; ++i)
      for (int d = 0; d < 3; ++d) {
        V[i][d] = a1 * F[i][d];
        W[i][d] = a2 * T[i][d];
      }
  }
  // Fcm::assign_filament_data (FCMfunctions.cpp:882-918): Y = fmod(Xs, L),
  // F, T from the driver's own filament objects. Any container of objects
  // with length() and beads[i].Xs(d), .F(d), .TAU(d) (the reference's
  // Filament and Bead) works; it is read in the caller's translation unit,
  // so this header needs none of the caller's libraries.
  template <class FilamentVector>
  void assign_filament_data(const FilamentVector &filaments) {
    const double L[3] = {params.nx * params.dx, params.ny * params.dx,
                         params.nz * params.dx};
    int j = 0;
    for (const auto &fil : filaments)
      for (int i = 0; i < (int)fil.length(); ++i, ++j) {
        if (j >= Np) throw Error("assign_filament_data: more segments than Np");
        for (int d = 0; d < 3; ++d) {
          Y[j][d] = std::fmod(fil.beads[i].Xs(d), L[d]);
          F[j][d] = fil.beads[i].F(d);
          T[j][d] = fil.beads[i].TAU(d);
        }
      }
    if (j != Np) throw Error("assign_filament_data: fewer segments than Np");
  }
  void mpi_broadcast() {}
  // Fcm::free_memory (FCMfunctions.cpp:468-519)
  void free_memory() {
    if (h) fs_fcm_destroy(h);
    h = nullptr;
  }
};

// struct LinkedList over host positions (copied to the device per call).
struct LinkedList {
  double min_dist2 = 999999;
  fs_colbar *h = nullptr;
  int Np = 0;
  fs_colbar_params params{};

  LinkedList() = default;
  // the box, radius and barrier constants the reference takes from
  // config.hpp, for the filament-vector initialise below
  explicit LinkedList(const fs_colbar_params &p) : params(p) {}

  // LinkedList::initialise + set_global_segment_number
  // (CollisionBarrierFilament.cpp:274-367)
  void initialise(const fs_colbar_params &p, const std::vector<int> &seg_fil,
                  const std::vector<int> &seg_idx) {
    params = p;
    Np = (int)seg_fil.size();
    check(fs_colbar_create(&p, Np, seg_fil.data(), seg_idx.data(), &h),
          "fs_colbar_create");
  }
  // The reference's own call, linked_list.initialise(myrank, totalnodes,
  // filaments) (main.cpp:251), with the parameters given at construction:
  // segment g of filament i_fil, index i_seg, in filament order (:345-367).
  template <class FilamentVector>
  void initialise(int myrank, int totalnodes, const FilamentVector &filaments) {
    (void)myrank;
    (void)totalnodes;
    std::vector<int> seg_fil, seg_idx;
    int i_fil = 0;
    for (const auto &fil : filaments) {
      for (int i_seg = 0; i_seg < (int)fil.length(); ++i_seg) {
        seg_fil.push_back(i_fil);
        seg_idx.push_back(i_seg);
      }
      ++i_fil;
    }
    initialise(params, seg_fil, seg_idx);
  }
  // LinkedList::link (CollisionBarrierFilament.cpp:370-433), from the
  // unwrapped positions Xs (Np x 3)
  void link(double **Xs) { check(fs_colbar_link_host(h, &Xs[0][0]), "link"); }
  // The reference's own call, linked_list.link(filaments) (main.cpp:394):
  // the beads' Xs in filament order.
  template <class FilamentVector>
  void link(const FilamentVector &filaments) {
    std::vector<double> xs;
    xs.reserve(3 * (size_t)Np);
    for (const auto &fil : filaments)
      for (int i = 0; i < (int)fil.length(); ++i)
        for (int d = 0; d < 3; ++d) xs.push_back(fil.beads[i].Xs(d));
    if ((int)xs.size() != 3 * Np) throw Error("link: segment count != Np");
    check(fs_colbar_link_host(h, xs.data()), "link");
  }
  void mpi_broadcast() {}
  // LinkedList::apply_collision_barrier (CollisionBarrierFilament.cpp:24-271)
  void apply_collision_barrier(Fcm &fcm, int nt) {
    masked(fcm, nt, FS_BARRIER_ALL);
  }
  // apply_collision_barrier_not_swimmer (:441-696)
  void apply_collision_barrier_not_swimmer(Fcm &fcm, int nt) {
    masked(fcm, nt, FS_BARRIER_NOT_SWIMMER);
  }
  // apply_collision_barrier_not_test_particles (:698-945)
  void apply_collision_barrier_not_test_particles(Fcm &fcm, int nt) {
    masked(fcm, nt, FS_BARRIER_NOT_TEST_PARTICLES);
  }
  void masked(Fcm &fcm, int nt, int mask) {
    check(fs_colbar_set_mask(h, mask), "fs_colbar_set_mask");
    check(fs_colbar_apply_host(h, &fcm.Y[0][0], nt, &fcm.F[0][0], &min_dist2),
          "apply_collision_barrier");
  }
  void free_memory() {
    if (h) fs_colbar_destroy(h);
    h = nullptr;
  }
};

// The filament set with its device-resident state: the reference's
// std::vector<Filament> + filamentJsolver, one handle.
struct Filaments {
  fs_system *h = nullptr;
  int nfil = 0, Np = 0, Nlam = 0;
  std::vector<int> nw;  // segments per filament

  void initialise(const fs_sim_params &p, const std::vector<int> &nw_,
                  const std::vector<fs_fil_props> &props) {
    check(fs_system_create(&p, (int)nw_.size(), nw_.data(), props.data(), &h),
          "fs_system_create");
    check(fs_system_sizes(h, &nfil, &Np, &Nlam), "fs_system_sizes");
    nw = nw_;
  }
  void set_state(const fs_state_ptrs &s) {
    check(fs_system_set_state(h, &s), "fs_system_set_state");
  }
  void get_state(fs_state_ptrs &s) {
    check(fs_system_get_state(h, &s), "fs_system_get_state");
  }
  // for nn: initialGuess(); unRobotArmify();  (main.cpp:334-336)
  void initialGuess() { check(fs_fil_initial_guess(h), "initialGuess"); }
  // setZeroForcesTorques .. applyExternalForcesTorques + assign_filament_data
  void applyForcesTorques(int nt) { check(fs_fil_forces(h, nt), "forces"); }
  // Jsolver.buildJacobian(filaments, nt) (main.cpp:369)
  void buildJacobian(int nt) {
    check(fs_fil_build_jacobian(h, nt), "buildJacobian");
  }
  // Jsolver.solve(filaments, b, x, iter) (device vectors of 6 Np)
  void solve(const double *b, double *x, int iter) {
    check(fs_fil_jsolve(h, b, x, iter), "solve");
  }
  // bad-Broyden columns (main.cpp:640-663)
  void broydenUpdate(const double *E, const double *Enew, const double *upd,
                     int iter) {
    check(fs_fil_broyden_update(h, E, Enew, upd, iter), "broydenUpdate");
  }
  // for nn: update(Update); unRobotArmify();  (main.cpp:528-531)
  void update(const double *upd) { check(fs_fil_update(h, upd), "update"); }
  // for nn: step(nt)  (main.cpp:704-707)
  void step(int nt) { check(fs_fil_step(h, nt), "step"); }
  // fcm.mobility_solve() / friction_mobility_solve() on the device arrays
  void mobility(int nt) { check(fs_system_mobility(h, nt), "mobility"); }
  // linked_list.link + apply_collision_barrier on the device arrays
  double barrier(int nt, bool link) {
    double md = 0;
    check(fs_system_barrier(h, nt, link ? 1 : 0, &md), "barrier");
    return md;
  }
  // the whole of main.cpp:284-754 for one nt
  int timestep(int nt) {
    int it = 0;
    check(fs_system_timestep(h, nt, &it), "timestep");
    return it;
  }
  // main_gaitreset.cpp:310-812: pass gait_reset_i of timestep nt (0: no
  // barrier, 1: no swimmer/network pairs, both trial passes without
  // Filament::step; 2: the full barrier and the committed step)
  int gaitResetPass(int nt, int gait_reset_i) {
    int it = 0;
    if (gait_reset_i == 0)
      check(fs_system_set_barrier(h, 0, FS_BARRIER_ALL), "set_barrier");
    else
      check(fs_system_set_barrier(
                h, 1, gait_reset_i == 1 ? FS_BARRIER_NOT_SWIMMER : FS_BARRIER_ALL),
            "set_barrier");
    if (gait_reset_i < 2)
      check(fs_system_timestep_trial(h, nt, &it), "timestep_trial");
    else
      check(fs_system_timestep(h, nt, &it), "timestep");
    return it;
  }
  // main_gaitresetandremove.cpp:328-339, 458-480, 775-820: pass
  // gait_reset_i of timestep nt. 0: every filament, no barrier, trial.
  // 1: filament 0 alone, no barrier, trial, on `swimmer` (a one-filament
  // system of filament 0's segments and properties on the same grid, made
  // with initialise) after it takes filament 0's state (filaments_copy =
  // filaments; filaments.resize(1); Np = its segment count, where the
  // reference writes 15). 2: every filament, full barrier, committed step.
  // This system is not touched by pass 1, so pass 2 starts from the state
  // pass 0 left, as the reference's filaments = filaments_copy restores it.
  // Spring links of a network stay on this system (the reference's pass 1
  // would apply them to filaments it removed).
  int gaitResetAndRemovePass(int nt, int gait_reset_i, Filaments &swimmer) {
    int it = 0;
    if (gait_reset_i == 1) {
      check(fs_system_copy_filaments(swimmer.h, h, 0), "copy_filaments");
      check(fs_system_set_barrier(swimmer.h, 0, FS_BARRIER_ALL),
            "set_barrier");
      check(fs_system_timestep_trial(swimmer.h, nt, &it), "timestep_trial");
      return it;
    }
    check(fs_system_set_barrier(h, gait_reset_i == 2 ? 1 : 0, FS_BARRIER_ALL),
          "set_barrier");
    if (gait_reset_i == 0)
      check(fs_system_timestep_trial(h, nt, &it), "timestep_trial");
    else
      check(fs_system_timestep(h, nt, &it), "timestep");
    return it;
  }
  // filaments[fil].getCOM() (Filament.hpp:248-255)
  std::vector<double> getCOM(int fil) {
    std::vector<double> c(3);
    check(fs_system_com(h, fil, c.data()), "getCOM");
    return c;
  }
  // average_v of main_gaitresetandremove.cpp:781-787 (filament fil's mean
  // segment velocity, float accumulation)
  std::vector<double> meanVelocity(int fil) {
    std::vector<double> v(3);
    check(fs_system_mean_velocity(h, fil, v.data()), "meanVelocity");
    return v;
  }
  void free_memory() {
    if (h) fs_system_destroy(h);
    h = nullptr;
  }
};

// file_functions.cpp writers over a system's device state (the reference's
// V, W, F, T and linked_list.F_total arguments are the system's own arrays
// of its last step or trial pass; at nt = 0 only the positions are read).
inline void save_data_to_file(const std::string &name, Filaments &f, int nt) {
  check(fs_system_write_dat(f.h, name.c_str(), nt), "save_data_to_file");
}
inline void save_backup_data_to_file(const std::string &name, Filaments &f,
                                     int nt) {
  check(fs_system_write_bak(f.h, name.c_str(), nt),
        "save_backup_data_to_file");
}
inline void save_data_column_names_to_swimmer_velocity_file(
    const std::string &name, Filaments &f) {
  check(fs_io_write_swimmer_velocity_names(name.c_str(), f.nw.at(0)),
        "save_data_column_names_to_swimmer_velocity_file");
}
inline void save_data_column_names_to_swimmer_collision_file(
    const std::string &name, Filaments &f) {
  check(fs_io_write_swimmer_collision_names(name.c_str(), f.nw.at(0)),
        "save_data_column_names_to_swimmer_collision_file");
}
inline void save_data_column_names_to_forces_file(const std::string &name,
                                                  Filaments &f) {
  check(fs_io_write_force_names(name.c_str(), f.nw.at(0)),
        "save_data_column_names_to_forces_file");
}
inline void save_swimmer_velocity_data_to_file(const std::string &name,
                                               Filaments &f, int nt) {
  check(fs_system_write_swimmer_velocity(f.h, name.c_str(), nt),
        "save_swimmer_velocity_data_to_file");
}
inline void save_swimmer_collision_data_to_file(const std::string &name,
                                                Filaments &f, int nt) {
  check(fs_system_write_swimmer_collision(f.h, name.c_str(), nt),
        "save_swimmer_collision_data_to_file");
}
inline void save_force_data_to_file(const std::string &name, Filaments &f,
                                    int nt) {
  check(fs_system_write_forces(f.h, name.c_str(), nt),
        "save_force_data_to_file");
}

// struct spring_links: the link table {A, B, C, D} (filament A bead B to
// filament C bead D) and its "-springlinks.dat" file. The reference calls
// apply_spring_link_forces after every force assembly (main.cpp:354-357,
// 540-543); here attach() hands the table to the device once and
// Filaments::applyForcesTorques / timestep apply it from then on.
struct spring_links {
  std::vector<std::vector<int>> links;

  int size() const { return (int)links.size(); }
  void read_spring_links_from_file(const std::string &path) {
    int n = 0;
    check(fs_io_read_spring_links(path.c_str(), &n, nullptr, 0),
          "read_spring_links_from_file");
    std::vector<int> flat(4 * (size_t)n);
    if (n > 0)
      check(fs_io_read_spring_links(path.c_str(), &n, flat.data(), n),
            "read_spring_links_from_file");
    links.assign(n, std::vector<int>(4));
    for (int l = 0; l < n; ++l)
      for (int c = 0; c < 4; ++c) links[l][c] = flat[4 * l + c];
  }
  void save_spring_links_to_file(const std::string &path) const {
    std::vector<int> flat = this->flat();
    check(fs_io_write_spring_links(path.c_str(), size(), flat.data()),
          "save_spring_links_to_file");
  }
  // NetworkSpringConstant, NetworkSpringNaturalLength (config_template.hpp:
  // 117-118)
  void attach(Filaments &f, double k, double natural_length) const {
    std::vector<int> flat = this->flat();
    check(fs_system_set_spring_links(f.h, size(), flat.data(), k,
                                     natural_length),
          "spring_links::attach");
  }
  // Same as attach: the forces are then added in every force assembly.
  void apply_spring_link_forces(Filaments &f, double k,
                                double natural_length) const {
    attach(f, k, natural_length);
  }

 private:
  std::vector<int> flat() const {
    std::vector<int> v;
    v.reserve(4 * links.size());
    for (const auto &r : links) v.insert(v.end(), r.begin(), r.begin() + 4);
    return v;
  }
};

// ErrorCheckFCM (MobilityConstraints.cpp:447): Error (device, 6 Np) from the
// device V/W; returns check.
inline bool ErrorCheckFCM(Filaments &f, double *Error, const double *V,
                          const double *W, int nt, double *maxabs = nullptr) {
  int chk = 0;
  check(fs_fil_residual(f.h, nt, V, W, Error, &chk, maxabs), "ErrorCheckFCM");
  return chk != 0;
}

}  // namespace filsim
