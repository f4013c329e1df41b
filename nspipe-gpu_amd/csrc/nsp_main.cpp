// nsp_run: host driver over the C ABI, the counterpart of the reference's
// main() (Main_Code/src/main.cu:11-56).  Parameters come from the command
// line instead of head.h #defines:
//
//   nsp_run --nr 48 --nt 64 --nz 257 --lz 100 --re 1850 --dt 0.01 --steps 1000
//           [--restart | --restart_dir DIR] [--save_dir DIR] [--io_dir DIR]
//           [--dt_frc 0.5 --dt_qcr 0.5 --dt_mnp 0.5] [--a_p 0.3] [--mixed]
//           [--device 0] [--out_every N] [--seed S --rms R]
//
// Without a restart it starts from the reference initial condition (or the
// seeded one with --seed); --restart reads ur_s.bin, ut_s.bin, uz_s.bin
// (mod_io_nonlinear.cu:426-429); --restart_dir reads a restart directory
// (state plus time and step).  With --io_dir it writes io_friction.txt,
// io_qcross.txt, io_ucross.txt and io_meanpr.txt there, each when the time
// since its last row reaches its --dt_* interval (mod_integrate.cu:131-151,
// head.h:33-40).  Every --out_every steps it prints step, err and the
// friction values (t, u_cl, u_tau).  At the end it writes ur.bin, ut.bin,
// uz.bin (main.cu:42-44), or a restart directory with --save_dir, and the
// wall-clock time.
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
  double dt_frc = 0.5, dt_qcr = 0.5, dt_mnp = 0.5;
  std::string restart_dir, save_dir, io_dir;
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
    else if (k == "--restart_dir") restart_dir = next();
    else if (k == "--save_dir") save_dir = next();
    else if (k == "--io_dir") io_dir = next();
    else if (k == "--dt_frc") dt_frc = std::atof(next());
    else if (k == "--dt_qcr") dt_qcr = std::atof(next());
    else if (k == "--dt_mnp") dt_mnp = std::atof(next());
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
  double t = 0.0;
  long long step0 = 0;
  if (!restart_dir.empty()) {
    if ((st = nsp_read_restart(h, restart_dir.c_str(), &t, &step0))) die("reading restart directory", st);
    std::printf("restart at t=%.8e, step %lld\n", t, step0);
  } else if (restart) {
    if ((st = nsp_read_state_files(h, "ur_s.bin", "ut_s.bin", "uz_s.bin"))) die("reading restart", st);
  } else if (rms >= 0.0) {
    if ((st = nsp_seeded_ic(h, seed, 4, 4, 0, rms))) die("nsp_seeded_ic", st);
  } else {
    if ((st = nsp_reference_ic(h, a_p))) die("nsp_reference_ic", st);
  }
  if ((st = nsp_autotune(h, 2))) die("nsp_autotune", st);
  if (!io_dir.empty() && (st = nsp_io_open(h, io_dir.c_str()))) die("nsp_io_open", st);
  double t_frc = 0.0, t_qcr = 0.0, t_mnp = 0.0;
  int done = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; s++) {
    nsp_step_info info;
    if ((st = nsp_step(h, 1, &info))) die("nsp_step", st);
    done = s + 1;
    t += dt;
    t_frc += dt;
    t_qcr += dt;
    t_mnp += dt;
    if (!io_dir.empty()) {
      int mask = 0;
      if (t_frc >= dt_frc) { mask |= NSP_IO_FRICTION; t_frc = 0.0; }
      if (t_qcr >= dt_qcr) { mask |= NSP_IO_CROSS; t_qcr = 0.0; }
      if (t_mnp >= dt_mnp) { mask |= NSP_IO_MEANPR; t_mnp = 0.0; }
      if (mask && (st = nsp_io_write(h, t, mask))) die("nsp_io_write", st);
    }
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
  if (!save_dir.empty()) {
    if ((st = nsp_write_restart(h, save_dir.c_str(), t, step0 + done))) die("writing restart directory", st);
  } else if ((st = nsp_write_state_files(h, "ur.bin", "ut.bin", "uz.bin"))) {
    die("writing state", st);
  }
  if (!io_dir.empty()) nsp_io_close(h);
  std::printf("Total time = %.3fs (%.4f steps/s)\n", secs, done / secs);
  nsp_destroy(h);
  return 0;
}
