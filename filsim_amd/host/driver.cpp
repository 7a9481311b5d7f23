// driver.cpp -- the reference driver's time loop (main.cpp:284-754) written
// against filsim_shim.hpp: the same call sequence per timestep, with the
// solver running on one MI355X. Built with the host compiler only (the
// C-ABI has no HIP types).
//
//   filsim_driver [grid nfil nw steps precision seed tol out.bin]
//
// Prints one row per step (Broyden iterations, step wall time) like the
// reference profiler (profilers.cpp:46-65) and, when out.bin is given,
// writes the final X (np x 3) and q (np x 4) as raw doubles.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/filsim_shim.hpp"

int main(int argc, char **argv) {
  int grid = argc > 1 ? atoi(argv[1]) : 64;
  int nfil = argc > 2 ? atoi(argv[2]) : 16;
  int nw = argc > 3 ? atoi(argv[3]) : 15;
  int steps = argc > 4 ? atoi(argv[4]) : 5;
  int precision = argc > 5 ? atoi(argv[5]) : FS_FP64;
  unsigned long long seed = argc > 6 ? strtoull(argv[6], nullptr, 10) : 1;
  double tol = argc > 7 ? atof(argv[7]) : 1e-4;
  const char *out = argc > 8 ? argv[8] : nullptr;
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
  fs_state_ptrs o{};
  o.X = X.data();
  o.q = q.data();
  filaments.get_state(o);
  if (out) {
    FILE *f = fopen(out, "wb");
    if (!f) return 1;
    fwrite(X.data(), sizeof(double), X.size(), f);
    fwrite(q.data(), sizeof(double), q.size(), f);
    fclose(f);
  }
  filaments.free_memory();
  return 0;
}
