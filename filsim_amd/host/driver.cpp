// driver.cpp -- the reference driver's time loop (main.cpp:284-754) written
// against filsim_shim.hpp: the same call sequence per timestep, with the
// solver running on one MI355X. Built with the host compiler only (the
// C-ABI has no HIP types).
//
//   filsim_driver [grid nfil nw steps precision seed tol out.bin calls.bin]
//
// Prints one row per step (Broyden iterations, step wall time) like the
// reference profiler (profilers.cpp:46-65) and, when out.bin is given,
// writes the final X (np x 3) and q (np x 4) as raw doubles. With calls.bin
// it then replays the reference's own call sites on its filament container
// (main.cpp:251 linked_list.initialise(myrank, totalnodes, filaments), 361
// fcm.assign_filament_data(filaments), 394 linked_list.link(filaments),
// then apply_collision_barrier and mobility_solve) with a stand-in for
// class Filament holding the final Xs (filament 1 moved beside filament 0),
// F, TAU, and writes Xs, F, TAU (input) then F after the barrier, V and W
// (np x 3 each).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/filsim_shim.hpp"

// Stand-in for the reference's Filament / Bead (Filament.hpp, Bead.hpp) with
// the members the call sites read: beads[i].Xs(d), .F(d), .TAU(d), length().
struct Vec3 {
  double v[3];
  double operator()(int d) const { return v[d]; }
};
struct Bead {
  Vec3 Xs, F, TAU;
};
struct Filament {
  std::vector<Bead> beads;
  int length() const { return (int)beads.size(); }
};

int main(int argc, char **argv) {
  int grid = argc > 1 ? atoi(argv[1]) : 64;
  int nfil = argc > 2 ? atoi(argv[2]) : 16;
  int nw = argc > 3 ? atoi(argv[3]) : 15;
  int steps = argc > 4 ? atoi(argv[4]) : 5;
  int precision = argc > 5 ? atoi(argv[5]) : FS_FP64;
  unsigned long long seed = argc > 6 ? strtoull(argv[6], nullptr, 10) : 1;
  double tol = argc > 7 ? atof(argv[7]) : 1e-4;
  const char *out = argc > 8 ? argv[8] : nullptr;
  const char *calls = argc > 9 ? argv[9] : nullptr;
  const double PI = 3.14159265358979323846264338327950288;

  // configuration (config.hpp / multi_filament_header.hpp as runtime data)
  fs_sim_params sp{};
  sp.fcm.nx = sp.fcm.ny = sp.fcm.nz = grid;
  sp.fcm.ngd = 12;
  sp.fcm.a = 1.0;
  sp.fcm.mu = 1.0;
  sp.fcm.dx = 1.0 / 1.86;
  sp.fcm.sigma = 1.0 / std::sqrt(PI);
  sp.fcm.sigma_dip = 1.0 / std::pow(36.0 * PI, 1.0 / 6.0);
  sp.precision = precision;
  sp.dL = 2.2;
  sp.dt = 0.01;
  sp.tol = tol;
  sp.broyden_maxiter = 100;
  sp.impl_euler_steps = 1;
  sp.initial_friction_steps = 0;
  sp.weight_per_length[2] = -1.0;
  sp.barrier_on = 1;
  sp.barrier_fs_over_2a = 15 * PI;
  sp.barrier_cap = 2.0;
  std::vector<int> nws(nfil, nw);
  std::vector<fs_fil_props> props(nfil);
  for (auto &p : props) {
    p = fs_fil_props{};
    p.kap = p.c = 1.0;
    p.beta = 1.0;
    p.radius = 1.0;
  }
  // filament_position_initialisation (seeded, overlap-rejected at 2.2a)
  int np = nfil * nw;
  std::vector<double> X(3 * (size_t)np), q(4 * (size_t)np);
  double box[3] = {grid * sp.fcm.dx, grid * sp.fcm.dx, grid * sp.fcm.dx};
  filsim::check(fs_make_filaments(box, nfil, nws.data(), sp.dL, 2.2, seed,
                                  100000, X.data(), q.data()),
                "fs_make_filaments");

  filsim::Filaments filaments;
  filaments.initialise(sp, nws, props);
  fs_state_ptrs s{};
  s.X = s.Xs = s.Xt = X.data();
  s.q = s.qs = s.qt = q.data();
  filaments.set_state(s);

  printf(" Timestep  (Bro) [ Total  ]\n");
  for (int nt = 1; nt <= steps; ++nt) {
    auto t0 = std::chrono::steady_clock::now();
    int iter = filaments.timestep(nt);
    double ms = std::chrono::duration<double, std::milli>(
                    std::chrono::steady_clock::now() - t0)
                    .count();
    printf(" %4d/%-4d (%3d) [%8.2fms]\n", nt, steps, iter, ms);
  }
  std::vector<double> Xs(3 * (size_t)np), Fb(3 * (size_t)np),
      Tb(3 * (size_t)np);
  fs_state_ptrs o{};
  o.X = X.data();
  o.q = q.data();
  o.Xs = Xs.data();
  o.F = Fb.data();
  o.TAU = Tb.data();
  filaments.get_state(o);
  if (out) {
    FILE *f = fopen(out, "wb");
    if (!f) return 1;
    fwrite(X.data(), sizeof(double), X.size(), f);
    fwrite(q.data(), sizeof(double), q.size(), f);
    fclose(f);
  }
  filaments.free_memory();
  if (calls) {
    // filament 1 translated so its first bead lies 1.9a from filament 0's:
    // inside the barrier's cutoff, so the barrier acts
    if (nfil > 1)
      for (int d = 0; d < 3; ++d) {
        const double sh = Xs[d] + (d == 0 ? 1.9 : 0.0) - Xs[3 * (size_t)nw + d];
        for (int i = 0; i < nw; ++i) Xs[3 * ((size_t)nw + i) + d] += sh;
      }
    std::vector<Filament> fils(nfil);
    for (int f = 0; f < nfil; ++f) {
      fils[f].beads.resize(nw);
      for (int i = 0; i < nw; ++i)
        for (int d = 0; d < 3; ++d) {
          const size_t g = 3 * ((size_t)f * nw + i) + d;
          fils[f].beads[i].Xs.v[d] = Xs[g];
          fils[f].beads[i].F.v[d] = Fb[g];
          fils[f].beads[i].TAU.v[d] = Tb[g];
        }
    }
    fs_colbar_params cp{};
    for (int d = 0; d < 3; ++d) cp.box[d] = box[d];
    cp.a = sp.fcm.a;
    cp.fs_over_2a = sp.barrier_fs_over_2a;
    cp.cap = sp.barrier_cap;
    cp.initial_friction_steps = sp.initial_friction_steps;
    filsim::LinkedList linked_list(cp);
    linked_list.initialise(0, 1, fils);                 // main.cpp:251
    filsim::Fcm fcm(sp.fcm, np, precision);
    fcm.initialise(0);                                  // main.cpp:255-261
    fcm.assign_filament_data(fils);                     // main.cpp:361
    linked_list.link(fils);                             // main.cpp:394
    linked_list.apply_collision_barrier(fcm, steps);    // main.cpp:400
    fcm.mobility_solve();                               // main.cpp:433-457
    FILE *f = fopen(calls, "wb");
    if (!f) return 1;
    fwrite(Xs.data(), sizeof(double), Xs.size(), f);
    fwrite(Fb.data(), sizeof(double), Fb.size(), f);
    fwrite(Tb.data(), sizeof(double), Tb.size(), f);
    fwrite(&fcm.F[0][0], sizeof(double), 3 * (size_t)np, f);
    fwrite(&fcm.V[0][0], sizeof(double), 3 * (size_t)np, f);
    fwrite(&fcm.W[0][0], sizeof(double), 3 * (size_t)np, f);
    fclose(f);
    linked_list.free_memory();
    fcm.free_memory();
  }
  return 0;
}
