// Internal declarations shared by the host and HIP sources of libnsp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nsp.h"

namespace nsp {

constexpr double PI2 = 6.283185307179586;  // head.h:12
constexpr int STEN = 7;                    // head.h:44
constexpr int IW = 3;                      // head.h:45
constexpr int MAXRAD = 16;

// ----------------------------------------------------------------- tables
// Host FP64 radial tables (init_fd, get_dr0, get_dr1: mod_radialFD.cu:199-358).
struct RadialTables {
  std::vector<double> r, D, Lu, Lp, rdr, dr0, dr1;
};
void build_radial_tables(int nr, RadialTables& T);
void radial_grid(int nr, double* r);
// [2][nrn][nro] radial interpolation of the remesh (half 0 even, 1 odd in r)
std::vector<double> remesh_matrix(const std::vector<double>& ro, const std::vector<double>& rn, bool mirror);


// ----------------------------------------------------------------- FFT plan
// In-place decimation-in-time mixed-radix plan; input is placed at perm[n].
struct FftPlan {
  int L;
  int nrad;
  int rad[MAXRAD];
};
bool make_fft_plan(int L, FftPlan& p, std::vector<int>& perm);

// complex pair aligned to its size, so loads and stores are single vector ops
template <typename T> struct alignas(2 * sizeof(T)) cpx { T x, y; };

// Geometry and constants shared by all kernels of one pipe.
struct Geom {
  int nr, nt, nz, modes;   // NR, NT, NZ, NT*NZ
  int ntp, nzf, nh;        // 3NT/2, 3(NZ-1), nzf/2
  size_t n;                // NR*NT*NZ (field stride)
  double kz0;              // 2*pi/LZ as the reference forms it (PI2/LZ)
  double re, dt, d_im, qdes;
  // Corrector launches of a captured time step: nonzero *skip means the step
  // has already converged and the launch returns at once.  Null elsewhere.
  const int* skip;
  // Levels a theta or z launch covers: [lv0, lv0 + nlv).  The whole column
  // (0, NR) except for the level chunks of the pipelined nonlinear stage.
  int lv0, nlv;
};

__device__ __forceinline__ bool step_done(const Geom& g) { return g.skip != nullptr && *g.skip != 0; }

// Transform intermediates W (theta-inverse output) and Y (z-pass output): per
// field and level a plane of NTP azimuthal rows by NZ axial columns, stored
// tile-major: tiles of 64 columns, each NTP rows of 64 contiguous values.
// A theta kernel's column tile then moves whole tile rows (512 bytes in FP32)
// instead of 64-byte pieces of rows 12 KB apart, and a z-pass row is NZ/64
// runs of 64 values.  Element (th, l) of a plane: th * 64 + tp_col(g, l).
__host__ __device__ __forceinline__ size_t tp_plane(const Geom& g) {
  return (size_t)g.ntp * (size_t)(((g.nz + 63) / 64) * 64);
}
__host__ __device__ __forceinline__ size_t tp_col(const Geom& g, int l) {
  return (size_t)(l >> 6) * ((size_t)g.ntp * 64) + (size_t)(l & 63);
}
// d_r u (dR, written by dr3, read by the theta inverse) and the stored N(u)
// (written by the theta forward, read by the sweep) are tile-major the same
// way, over planes of NT spectral rows: element (it, l) of a plane at
// it * 64 + dr_col(g, l); components dr_cs(g) values apart.  A sweep wave's
// 64 modes (one azimuthal row, one 64-mode chunk) are one tile row.
__host__ __device__ __forceinline__ size_t dr_plane(const Geom& g) {
  return (size_t)g.nt * (size_t)(((g.nz + 63) / 64) * 64);
}
__host__ __device__ __forceinline__ size_t dr_cs(const Geom& g) { return (size_t)g.nr * dr_plane(g); }
__host__ __device__ __forceinline__ size_t dr_col(const Geom& g, int l) {
  return (size_t)(l >> 6) * ((size_t)g.nt * 64) + (size_t)(l & 63);
}

// Workgroups are dealt to the 8 XCDs round robin; remap so that each XCD
// walks one contiguous range of logical blocks and neighbouring tiles and
// radial levels meet in the same L2.
__device__ __forceinline__ int xcd_block(int b, int G) {
  constexpr int NX = 8;
  const int q = G / NX, rmd = G % NX;
  const int x = b % NX, j = b / NX;
  return (x < rmd ? x * (q + 1) : rmd * (q + 1) + (x - rmd) * q) + j;
}

// A pointer read from a device table of launch arguments is a generic pointer
// to the compiler, and every access through it becomes a flat memory
// instruction.  Re-deriving it from the table pointer (a kernel argument, which
// the compiler knows to be global memory) by its byte distance lets the
// accesses use global instructions, as for pointers passed as kernel arguments.
// Only for pointers that are never null: the compiler takes the result of the
// pointer arithmetic as non-null (a null test of it becomes a test of base).
template <typename T>
__device__ __forceinline__ T* gmem(const void* base, T* p) {
  const long long off = (long long)(size_t)p - (long long)(size_t)base;
  return (T*)((char*)const_cast<void*>(base) + off);
}

// Sweep-side layout.  The buffers only the linear sweep reads and writes
// (column scratch SR/SP, explicit part E, LU factors F, homogeneous pressure
// data PH) are stored per sweep workgroup, so that one workgroup's radial
// column is a contiguous run of memory instead of NR lines spread over the
// whole field.  Workgroup lb = pair * nch + ch covers the axial chunk ch (64
// modes, nch = ceil(NZ/64)) of the azimuthal rows it = pair (wave 0) and
// NT - pair (wave 1; pair 0: the Nyquist row NT/2).  A workgroup has 128
// lanes, lane = 64 * wave + l % 64.
__host__ __device__ __forceinline__ int sw_nch(const Geom& g) { return (g.nz + 63) / 64; }
__host__ __device__ __forceinline__ size_t sw_nblk(const Geom& g) { return (size_t)(g.nt / 2) * sw_nch(g); }
// values of one component of a blocked field (>= NR*NT*NZ)
__host__ __device__ __forceinline__ size_t sw_field(const Geom& g) { return sw_nblk(g) * g.nr * 128; }
// Quantities that depend on |m| (uz and pressure factors, PH) are stored for
// the rows it <= NT/2 only: rows it < NT/2 in their workgroup's wave-0 lanes
// (64 per row), the Nyquist row in a region of its own after all workgroups.
// Index of row i, value c (of nc per row) of the mode k = it * NZ + l, it <= NT/2.
__host__ __device__ __forceinline__ size_t sw_abs_idx(const Geom& g, int nc, int i, int c, int k) {
  const int it = k / g.nz, l = k - it * g.nz, ch = l >> 6, lane = l & 63;
  const size_t blk = it < g.nt / 2 ? (size_t)it * sw_nch(g) + ch : sw_nblk(g) + ch;
  return ((blk * g.nr + i) * nc + c) * 64 + lane;
}
// doubles of PH: [workgroup or Nyquist chunk][NR][2][64]
__host__ __device__ __forceinline__ size_t sw_ph_count(const Geom& g) {
  return (sw_nblk(g) + sw_nch(g)) * g.nr * 2 * 64;
}

// Read-only tables seen through the constant address space: loads at
// wave-uniform addresses become scalar-cache loads instead of vector loads.
typedef const double __attribute__((address_space(4)))* ctab;

// Device tables of one pipe.
struct DevTables {
  double* r;     // [NR]
  double* D;     // [2][NR][7]
  double* Lu;    // [2][NR][7]
  double* Lp;    // [2][NR][7]
  double* rdr;   // [NR]
  double* dr0;   // [4]
  double* dr1;   // [7]
  double* uc;    // [NR] flux profile (mod_cublasflux.cu:69-76)
  double* qc;    // [1]  Qc = rdr . uc
};

// Influence-matrix data (mod_boundary.cu:200), +/- form.
struct DevInfluence {
  double2* uh;    // [3][N] homogeneous velocity solutions
  double2* uhP;   // [3][N] gradient of the homogeneous pressure solution
  double2* invc;  // [MODES][4] column 3 of the inverse influence matrix
};

// Outcome of one captured step of one pipe, written on the device: converged
// flag, corrector iterations, flags (st[3] unused); final err.
struct StepRec {
  int st[4];
  double err;
};

// One pipe of a batched nonlinear launch (ensembles): blockIdx.y (dr3: .z)
// selects the member, whose geometry, tables and fields replace the launch's
// own arguments.  Blocks beyond the member's extent return at once.
struct NlMember {
  Geom g;
  const double* r;  // [NR]
  const double* D;  // [2][NR][7]
  const double2* u; // input state
  double2* dR;      // d_r u (FP32 pairs when f32)
  void* W;          // theta-inverse intermediates (class precision)
  void* Y;          // z-forward intermediates
  double2* N;       // N(u) (FP32 pairs when f32)
  size_t ws;        // intermediate field stride
  int f32;          // mixed form
};
// A batched launch: members of one compiled-plan class, and the largest
// extents among them (grid sizes).
struct Batch {
  const NlMember* tab;
  int count, max_nr, max_nz, max_modes, max_nt;
};

// Streaming linear sweep (linear2.hip) argument block.
struct Lin2Args {
  Geom g;
  const double* r;      // [NR]
  const double* rdr;    // [NR]
  const double* uc;     // [NR] flux profile
  const double* qc;     // [1]
  const double* Dc;     // [2][NR][7] D weights of columns i-3..i+3 (0 outside the grid)
  const double* Luc;    // [2][NR][7] same for Lu
  const double* Lu;     // [2][NR][7] operator tables (factor recurrence on the fly)
  const double* Lp;     // [2][NR][7]
  const double* F;      // banded LU factors (sweep-side layout: fidx in linear2.hip)
  const float* F32;     // mixed form: the u+, u-, uz factors rounded to FP32 (same indices), or null
  int m32;              // |m| from which the sweep reads F32 instead of F
  int s32;              // mixed form: column scratch SR, SP and E of |m| >= m32 held as FP32 pairs
  int form64;           // 1: the pipe is of the FP64 form (selects its pass-D kernel, batched or not)
  const double* PH;     // dr p_h, p_h of the homogeneous pressure solution (sw_abs_idx, 2 per row)
  const double2* invc;  // [MODES][4] column 3 of the inverse influence matrix
  const double2* N;     // [3] nonlinear term (r/theta); float2 storage when n32
  int n32;              // mixed form: N(u) stored as FP32 pairs
  double2* E;           // [3] explicit part (+/-), written by the predictor (sweep-side layout)
  const double2* u;     // [3] predictor: u^n; corrector: u^k (r/theta)
  double2* out;         // [3] new iterate (r/theta)
  double2* SR;          // [3] scratch (sweep-side layout)
  double2* SP;          // [1] scratch (sweep-side layout)
  double2* DH;          // [1] refined form: div(H u^n) of the step (sweep-side layout), else null
  const double2* un;    // [3] u^n of the step (r/theta); the refined form's solves are for u - u^n
  int inc;              // refined form: increment solves (sweep_kernel INC = 1)
  unsigned long long* norm;
  int predictor;
  int ord;              // block order (NlVariant::s_ord)
};
// F32 (may be null): FP32 copies of the u+, u-, uz factor sets, same indices
void launch_factor(const Geom& g, const double* r, const double* Lu, const double* Lp, double* F, float* F32,
                   hipStream_t s);
size_t factor_count(const Geom& g);    // doubles of the stored LU factors
size_t factor32_count(const Geom& g);  // floats of the FP32 u+, u-, uz sets
void launch_sweep2(const Lin2Args& a, hipStream_t s);
// one launch per pass for `count` pipes (tab in device memory); max_nt/max_nz
// are the largest extents among them; fp64/mixed: some pipe is of that form
void launch_sweep_batch(const Lin2Args* tab, int count, int predictor, int max_nt, int max_nz, bool fp64,
                        bool mixed, hipStream_t s);
// centred stencil table: w-th weight of row i multiplies column i-3+w
std::vector<double> centred_table(int nr, const std::vector<double>& T);

// Kernel variant of one theta stage: alg 0 = LDS-scatter form (tl 4, 8 or 16
// columns, 256 threads; pad 1 = one LDS slot per 16 or 8 elements, pad 2 =
// rows of (NTP/R_last)*tl elements padded by tl slots; tl 16 dense only),
// alg 1 = register form ((tl, nthr) = (4, 192), (8, 256), (8, 384) or
// (16, 256)); ord = block
// order (0: radial level fastest, 1: column tile fastest).
struct ThetaVar {
  // alg 0 LDS-scatter form, 1 register form; tl columns per block; nthr threads;
  // pad (scatter form): 0 dense, 1 one slot per 16 (tl 4) or 8 (tl 8) elements,
  // 2 row padding, 3 swizzled rows; ord block order
  int alg, tl, nthr, pad, ord;
};
// Kernel variant of the compiled-plan path (chosen by the autotuner).
constexpr int kNlMaxChunks = 8;
struct NlVariant {
  ThetaVar inv, fwd;
  int z_alg;   // z pass: 0 LDS-scatter form, 1 register form (256 threads), 2 register form with the
               // inverse transforms in triples, 3 pair form (two real transforms as one full-length
               // complex transform), 4 pair form with two rows per workgroup
  int z_nthr;  // scatter form: threads per row (256 or 512)
  int z_pad;   // scatter form: 1 = one LDS slot per 16 elements, 2 = row padding as the register form;
               // triples: row padding of the three-column layout 0, 1 or R0 slots; pairs: row
               // padding 0, 1 or R0 slots
  int z_pf;    // register and pair forms: load the next transform's inputs before the LDS passes
  int s_ord;   // linear sweep: 1 = the chunks of an azimuthal row XCD-contiguous, 0 = dispatch order
  int nl_chunks;  // nonlinear stage: 1 = stage after stage; K > 1 = the theta inverse and the z pass
                  // in K level chunks on two streams, the z pass of chunk c overlapping the theta
                  // inverse of chunk c + 1
};
// Nonlinear-phase buffers and FFT data.
struct NlBufs {
  FftPlan pt, ph;      // theta (length NTP) and half-length z (length NZF/2) plans
  int* perm_t;         // [NTP]
  int* perm_h;         // [NH]
  double2* tw_t;       // [NTP] exp(-2 pi i e/NTP)
  double2* tw_h;       // [NH]  exp(-2 pi i e/NH)
  double2* tw_z;       // [NH+1] exp(-2 pi i k/NZF)
  double2* dR;         // [3][N] radial derivatives
  double2* W;          // [9][NR*NTP*NZ] theta-inverse intermediates
  double2* Y;          // [3][NR*NTP*NZ] z-forward intermediates
  size_t wstride;      // NR*NTP*NZ
  // compile-time plans (when the lengths have one): per-pass twiddles
  bool use_ct;
  double2* ct_tw_t;    // [NTP-1]
  double2* ct_tw_h;    // [NH-1] half-length z plan, then [NZF-1] of the pair form's plan if compiled
  // mixed form: FP32 transforms and products
  bool mixed;
  float2* W32;         // [9][NR*NTP*NZ]
  float2* Y32;         // [3][NR*NTP*NZ]
  float2* ct_tw_t32;
  float2* ct_tw_h32;
  float2* tw_z32;
  NlVariant var;
  // the second stream and the events of the pipelined nonlinear stage
  hipStream_t s2;
  hipEvent_t ev[kNlMaxChunks + 1];
};

// ----------------------------------------------------------------- profiler
// Per-kernel HIP-event timing on the launching stream (enabled on demand).
enum KernelId { K_DR3 = 0, K_THETA_INV, K_ZPASS, K_THETA_FWD, K_LIN_PRED, K_LIN_CORR, K_COUNT };
const char* kernel_name(int id);

struct Profiler {
  struct Pair {
    hipEvent_t b, e;
    int id;
  };
  bool on = false;
  std::vector<Pair> pending;  // recorded since the last collect
  std::vector<hipEvent_t> spare;
  double total_ms[K_COUNT] = {};
  long long launches[K_COUNT] = {};
  void reset() {
    for (int q = 0; q < K_COUNT; q++) { total_ms[q] = 0.0; launches[q] = 0; }
    recycle();
  }
  hipEvent_t get() {
    hipEvent_t e = nullptr;
    if (!spare.empty()) {
      e = spare.back();
      spare.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  }
  void begin(int id, hipStream_t s) {
    if (!on) return;
    Pair p{get(), get(), id};
    if (!p.b || !p.e) { on = false; return; }
    (void)hipEventRecord(p.b, s);
    pending.push_back(p);
  }
  void end(hipStream_t s) {
    if (!on || pending.empty()) return;
    (void)hipEventRecord(pending.back().e, s);
  }
  void recycle() {
    for (const Pair& p : pending) {
      spare.push_back(p.b);
      spare.push_back(p.e);
    }
    pending.clear();
  }
  // call after the stream has been synchronised
  void collect() {
    for (const Pair& p : pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, p.b, p.e) == hipSuccess) {
        total_ms[p.id] += ms;
        launches[p.id] += 1;
      }
    }
    recycle();
  }
  ~Profiler() {
    recycle();
    for (auto e : spare) (void)hipEventDestroy(e);
  }
};

// Compiled-plan support (nonlinear_ct64.hip).
bool ct_has_theta(int L);
bool ct_has_zhalf(int L);
bool ct_has_zreg(int L);
std::vector<int> ct_radices(int L, bool theta);
std::vector<int> ct_pair_radices(int nh);
// per-pass twiddles of a plan: for each pass (radix R, sub-length S) S*(R-1)
// entries exp(-2 pi i q k / (S R)) at [k*(R-1) + q-1]
std::vector<double2> pass_twiddles(const std::vector<int>& radices);
std::vector<double2> compact_twiddles(const std::vector<int>& radices);
template <typename T>
bool ct_theta_inv(const NlVariant& v, const Geom& g, const cpx<T>* tw, const double* r, const double2* u,
                  const double2* dR, cpx<T>* W, size_t ws, hipStream_t s, const Batch* bt = nullptr);
template <typename T>
bool ct_theta_fwd(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* Y, size_t ws, const double* D,
                  const double2* u, double2* N, hipStream_t s, const Batch* bt = nullptr);
template <typename T>
bool ct_zpass(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* twz, const double* r,
              const cpx<T>* W, size_t ws, cpx<T>* Y, hipStream_t s, const Batch* bt = nullptr);
// dR is stored as FP32 pairs when f32 (the mixed form); with a batch, one
// launch covers every member of bt.tab
void launch_dr3(const Geom& g, const double* D, const double2* u, double2* dR, bool f32, hipStream_t s,
                const Batch* bt = nullptr);

// Physical space (output.hip): component u (one field) through the padded
// inverse transform into out rows of `pitch` doubles (mode 0 values, 1
// squares, 2 add squares), Wt = NR*NTP*NZ complex scratch; cross-section
// integrals of F (NR*NTP rows) into q[NZF] with q1[NR*NZF] scratch.
void launch_physical(const Geom& g, const NlBufs& nb, const double2* u, double2* Wt, double* out, size_t pitch,
                     int mode, hipStream_t s);
void launch_cross(const Geom& g, const double* F, size_t pitch, const double* rdr, double* q1, double* q,
                  hipStream_t s);

// Launchers (HIP sources).
void launch_nonlinear(const Geom& g, const DevTables& t, const NlBufs& nb, const double2* u,
                      double2* N, hipStream_t s, Profiler* pf);
void launch_norm_reset(unsigned long long* norm, const int* skip, hipStream_t s);
void launch_widen(const float2* in, double2* out, size_t count, hipStream_t s);
// tile-major spectral field (three components, dr_cs apart; FP32 pairs when
// f32) to the reference layout [3][NR][NT][NZ] in FP64
void launch_untile(const Geom& g, const void* in, bool f32, double2* out, hipStream_t s);
void launch_add(double2* u, const double2* f, size_t count, hipStream_t s);
// new state = remesh of the old (both on this device); RM and band [nrn] of
// (first, last) nonzero columns in device memory
void launch_remesh(const Geom& gn, const Geom& go, const double* RM, const int2* band, const double2* uo,
                   double2* un, hipStream_t s);
// Ensemble bookkeeping, one thread per member: reset the records, reset the
// norms of members still iterating, and the loop test after corrector `it`.
void launch_ens_begin(StepRec* rec, int count, hipStream_t s);
void launch_ens_norm_reset(unsigned long long* const* norms, const StepRec* rec, int count, hipStream_t s);
void launch_ens_converge(unsigned long long* const* norms, StepRec* rec, int count, double tol, int it, int maxit,
                         hipStream_t s);
// Step bookkeeping of a captured step: st = {converged, corrector iterations, flags}.
void launch_step_begin(int* st, double* err, hipStream_t s);
void launch_converge(const unsigned long long* norm, int* st, double* err, double tol, int it, int maxit, hipStream_t s);
void launch_setup_influence(const Geom& g, const DevTables& t, const DevInfluence& b,
                            double2* scratch_p, double* scratch_U, double* PH, hipStream_t s);
void launch_diagnostics(const Geom& g, const DevTables& t, const double2* u, double* partial,
                        hipStream_t s);
// Allocation-free diagnostics (nsp_diagnostics_noalloc): block partials of at
// most kDiagMaxBlocks blocks of 256 modes in a static per-device array.
constexpr int kDiagMaxBlocks = 16384;
constexpr int kDiagMaxModes = kDiagMaxBlocks * 256;
hipError_t diag_scratch_init(int device);  // once per device, that device current
double* diag_scratch(int device);          // its address on that device (null before init)
void launch_diagnostics_noalloc(const Geom& g, const DevTables& t, const double2* u, double* part, double* out,
                                hipStream_t s);

}  // namespace nsp
