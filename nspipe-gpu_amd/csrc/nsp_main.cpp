// nsp_run: host driver over the C ABI, the counterpart of the reference's
// main() (Main_Code/src/main.cu:11-56).  Parameters come from the command
// line instead of head.h #defines:
//
//   nsp_run --nr 48 --nt 64 --nz 257 --lz 100 --re 1850 --dt 0.01 --steps 1000
//           [--restart] [--a_p 0.3] [--mixed] [--device 0] [--out_every N]
//           [--seed S --rms R]
//
// Without --restart it starts from the reference initial condition (or the
// seeded one with --seed); with --restart it reads ur_s.bin, ut_s.bin,
// uz_s.bin (mod_io_nonlinear.cu:426-429).  Every --out_every steps it prints
// step, time, err and the friction-file values (t, u_cl, u_tau) like
// io_friction.txt rows (mod_io_nonlinear.cu:474); at the end it writes
// ur.bin, ut.bin, uz.bin (main.cu:42-44) and the wall-clock time.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nsp.h"

static void die(const char* what, nsp_status st) {
  std::fprintf(stderr, "nsp_run: %s failed (status %d): %s\n", what, int(st), nsp_last_error());
  std::exit(1);
}

int main(int argc, char** argv) {
  int nr = 48, nt = 64, nz = 257, steps = 10, device = 0, out_every = 50, mixed = 0, restart = 0;
  double lz = 100.0, re = 1850.0, dt = 0.01, a_p = 0.3, rms = -1.0;
  unsigned long long seed = 0;
  for (int a = 1; a < argc; a++) {
    const std::string k = argv[a];
    auto next = [&]() -> const char* {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", k.c_str());
        std::exit(2);
      }
      return argv[++a];
    };
    if (k == "--nr") nr = std::atoi(next());
    else if (k == "--nt") nt = std::atoi(next());
    else if (k == "--nz") nz = std::atoi(next());
    else if (k == "--lz") lz = std::atof(next());
    else if (k == "--re") re = std::atof(next());
    else if (k == "--dt") dt = std::atof(next());
    else if (k == "--steps") steps = std::atoi(next());
    else if (k == "--a_p") a_p = std::atof(next());
    else if (k == "--device") device = std::atoi(next());
    else if (k == "--out_every") out_every = std::atoi(next());
    else if (k == "--mixed") mixed = 1;
    else if (k == "--restart") restart = 1;
    else if (k == "--seed") seed = std::strtoull(next(), nullptr, 10);
    else if (k == "--rms") rms = std::atof(next());
    else {
      std::fprintf(stderr, "unknown option %s\n", k.c_str());
      return 2;
    }
  }
  nsp_params p;
  nsp_status st = nsp_default_params(&p, nr, nt, nz, lz, re, dt);
  if (st) die("nsp_default_params", st);
  p.precision = mixed ? NSP_MIXED : NSP_FP64;
  nsp_handle* h = nullptr;
  std::printf("nsPipe on MI355X: NR=%d NT=%d NZ=%d LZ=%g Re=%g dt=%g (%s)\n", nr, nt, nz, lz, re, dt,
              mixed ? "mixed" : "fp64");
  if ((st = nsp_create(&p, device, &h))) die("nsp_create", st);
  if (restart) {
    if ((st = nsp_read_state_files(h, "ur_s.bin", "ut_s.bin", "uz_s.bin"))) die("reading restart", st);
  } else if (rms >= 0.0) {
    if ((st = nsp_seeded_ic(h, seed, 4, 4, 0, rms))) die("nsp_seeded_ic", st);
  } else {
    if ((st = nsp_reference_ic(h, a_p))) die("nsp_reference_ic", st);
  }
  if ((st = nsp_autotune(h, 2))) die("nsp_autotune", st);
  const auto t0 = std::chrono::steady_clock::now();
  double t = 0.0;
  for (int s = 0; s < steps; s++) {
    nsp_step_info info;
    if ((st = nsp_step(h, 1, &info))) die("nsp_step", st);
    t += dt;
    if (info.flags & NSP_FLAG_NAN) {
      std::fprintf(stderr, "step %d: convergence measure is NaN\n", s + 1);
      break;
    }
    if (out_every > 0 && (s + 1) % out_every == 0) {
      double d[7];
      if ((st = nsp_diagnostics(h, d))) die("nsp_diagnostics", st);
      std::printf(" step %d of %d  err=%e  n_it=%d  %.8e  %.8e  %.8e\n", s + 1, steps, info.err, info.n_it, t, d[0],
                  d[1]);
    }
  }
  if ((st = nsp_synchronize(h))) die("nsp_synchronize", st);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if ((st = nsp_write_state_files(h, "ur.bin", "ut.bin", "uz.bin"))) die("writing state", st);
  std::printf("Total time = %.3fs (%.4f steps/s)\n", secs, steps / secs);
  nsp_destroy(h);
  return 0;
}
