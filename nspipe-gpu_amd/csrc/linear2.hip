// Linear phase, streaming form: one thread per Fourier mode, four passes
// over the radial column with sliding register windows.
//
//   A (up)   rhs = E + d_im N (corrector) or N + CN part of u^n (predictor,
//            which also stores the explicit part E); to r/theta; divergence
//            (div_k, zero wall row); forward elimination of the pressure system
//   B (down) pressure back substitution (p overwrites its eliminated rhs)
//   C (up)   rhs -= grad p three rows behind (addGrad_k); to +/-; zero wall
//            rows (write_bcon_k); forward elimination of the u+, u-, uz systems
//   D (down) back substitution of u+, u-, uz together with the homogeneous
//            wall-unit solutions; the wall divergence of the four wall rows
//            fixes the influence coefficients (corBCon); a second walk from
//            the wall applies them row by row; to r/theta; convergence
//            measure (mcor_k) and output.  Mode (0,0) then takes modeZZ_k
//            and the flux correction (adj_flux).
//
// The banded LU factors of the four operators depend only on (m, kz), dt and
// Re, so they are computed once at setup with the reference's recurrence
// (LUsolve_*_k, mod_linear.cu:55-646) and stored per workgroup, 7 per row
// (fidx): every access of a wavefront is one contiguous line.  (Recomputing the L
// multipliers in the upward passes saves 3 of the 7 but needs the U window of
// three operators in registers, which halves occupancy: slower.)
// Workgroups pair the modes m and -m of one axial chunk: the u- factors at m
// are the u+ factors at -m (bit for bit: the operators differ only in the sign
// of the 2m/r^2 term) and the uz and pressure factors and the pressure
// homogeneous solution depend on |m|, so both waves read the same lines.
// Scratch SR (3 complex per row) and SP (1 complex per row) carry the column
// between passes.  Scratch, E, the factors and PH are stored per workgroup
// (sweep-side layout, nsp_internal.h): a workgroup's column is contiguous.
#include "nsp_internal.h"

namespace nsp {

struct ModeQ {
  int k, it, l, m;
  double kt, kz;
};

__device__ __forceinline__ ModeQ modeq(const Geom& g, int k) {
  ModeQ q;
  q.k = k;
  q.it = k / g.nz;
  q.l = k - q.it * g.nz;
  q.m = q.it < g.nt / 2 ? q.it : q.it - g.nt;
  q.kt = double(q.m);
  q.kz = g.kz0 * double(q.l);
  return q;
}

__device__ __forceinline__ void pm_of(double2& a, double2& b) {  // decoupleForward_k
  const double2 ur = a, ut = b;
  a = make_double2(ur.x - ut.y, ur.y + ut.x);
  b = make_double2(ur.x + ut.y, ur.y - ut.x);
}
__device__ __forceinline__ void rt_of(double2& a, double2& b) {  // decoupleBackward_k
  const double2 up = a, um = b;
  a = make_double2(0.5 * (up.x + um.x), 0.5 * (up.y + um.y));
  b = make_double2(0.5 * (up.y - um.y), -0.5 * (up.x - um.x));
}

// factor slot c of operator `kind` at row i: c 0..2 = L multipliers of
// columns i-3..i-1, 3 = 1/U_ii, 4..6 = U_{i,i+1..i+3}.  Only the factors the
// sweep reads are stored, in the sweep-side layout:
//   u+ (kind 0), every mode:          [workgroup][NR][7][128 lanes]
//   u- (kind 1), azimuthal row NT/2:  [chunk][NR][7][128] (lanes 64..127);
//                elsewhere u- at m is u+ at -m, the other wave of the workgroup
//   uz, p (kinds 2, 3), rows it <= NT/2 (they depend on |m|):
//                                     [workgroup or Nyquist chunk][NR][7][64] each
// Per row the u+/u- sets step by FR_UP = 7*128 doubles and per slot by 128,
// the uz/p sets by FR_ABS = 7*64 and 64.
constexpr int FR_UP = 7 * 128, FR_ABS = 7 * 64;
__host__ __device__ __forceinline__ size_t f_u1(const Geom& g) { return sw_nblk(g) * g.nr * FR_UP; }
__host__ __device__ __forceinline__ size_t f_abs(const Geom& g) {
  return f_u1(g) + (size_t)sw_nch(g) * g.nr * FR_UP;
}
__host__ __device__ __forceinline__ size_t f_absn(const Geom& g) {  // doubles of one |m| set
  return (sw_nblk(g) + sw_nch(g)) * g.nr * FR_ABS;
}
__host__ __device__ __forceinline__ size_t fidx(const Geom& g, int kind, int i, int c, int k) {
  const int it = k / g.nz, l = k - it * g.nz, nth = g.nt / 2;
  if (kind == 0) {
    const int pair = it < nth ? it : (it == nth ? 0 : g.nt - it), w = it < nth ? 0 : 1;
    const size_t lb = (size_t)pair * sw_nch(g) + (l >> 6);
    return ((lb * g.nr + i) * 7 + c) * 128 + w * 64 + (l & 63);
  }
  if (kind == 1) return f_u1(g) + (((size_t)(l >> 6) * g.nr + i) * 7 + c) * 128 + 64 + (l & 63);
  return f_abs(g) + (size_t)(kind - 2) * f_absn(g) + sw_abs_idx(g, 7, i, c, k);
}

// The operator rows of one mode (LUsolve_*_k): kind 0 u+, 1 u-, 2 uz
// (velocity Helmholtz, (1/dt) I - (d_im/Re)(Lu + diag)), 3 pressure (Lp + diag;
// mode (0,0) has the wall row p = 0).
struct OpMode {
  const double* LHS;
  int ip0, kind;
  bool vel, p00;
  double a, idt, kt, kz;
};
__device__ __forceinline__ OpMode op_mode(const Geom& g, const double* Lu, const double* Lp, int kind, int m, int l) {
  OpMode o;
  o.kind = kind;
  o.LHS = (kind == 3) ? Lp : Lu;
  const int fp = (kind <= 1) ? 1 : 0;
  o.ip0 = abs((abs(m) % 2) - fp) * STEN * g.nr;
  o.vel = kind < 3;
  o.p00 = (kind == 3) && (l == 0 && m == 0);
  o.a = g.d_im / g.re;
  o.idt = 1.0 / g.dt;
  o.kt = double(m);
  o.kz = g.kz0 * double(l);
  return o;
}

// Row i of the reference recurrence without the right-hand side.  Uw[s] is
// the U row i-3+s (s = 2: row i-1), diagonal first.  Out: Lc[c] = multiplier
// of column i-3+c (0 where the column is outside the band), Ui = U row i.
// NL = min(i, 3) is the number of columns left of the diagonal; it is a
// template parameter so that every register index is a constant.
// Contraction is off so every caller gets the same bits.
template <int NL>
__device__ __forceinline__ void factor_row_nl(const OpMode& o, int nr, int i, double ri, const double (&Uw)[IW][IW + 1],
                                              double (&Lc)[IW], double (&Ui)[IW + 1]) {
#pragma clang fp contract(off)
  const int right = min(i + IW, nr - 1);
  const int i0 = i + IW - right;
  const double r2 = ri * ri;
  const bool last = (i == nr - 1);
  const double* row = o.LHS + o.ip0 + STEN * i + i0;  // row[j] = entry of column left + j
  double L[IW];
#pragma unroll
  for (int j = 0; j < IW; j++) {
    L[j] = 0.0;
    if (j < NL) {
      double XM = row[j];
      if (!last) {
        if (o.vel) XM = -o.a * XM;
      } else if (o.p00) {
        XM = 0.0;
      }
      double Lj = XM;
      // U row of column left+h lives at Uw[IW - NL + h]
#pragma unroll
      for (int h = 0; h < IW; h++)
        if (h < j) Lj -= L[h] * Uw[IW - NL + h][j - h];
      Lj = Lj / Uw[IW - NL + j][0];
      L[j] = Lj;
    }
  }
#pragma unroll
  for (int j = 0; j <= IW; j++) {
    Ui[j] = 0.0;
    if (j < right - i + 1) {
      double XM = row[NL + j];
      if (!last) {
        if (j == 0) {
          const double kt = o.kt, kz = o.kz;
          if (o.kind == 0) XM += -(1.0 / r2) - (2.0 * kt / r2) - kz * kz - kt * kt / r2;
          else if (o.kind == 1) XM += -(1.0 / r2) + (2.0 * kt / r2) - kz * kz - kt * kt / r2;
          else XM += -kz * kz - kt * kt / r2;
          if (o.vel) XM = o.idt - o.a * XM;
        } else if (o.vel) {
          XM = -o.a * XM;
        }
      } else if (o.p00) {
        XM = (j == 0) ? 1.0 : 0.0;
      }
      const int hleft = (NL + j - IW) > 0 ? (NL + j - IW) : 0;
      double Uij = XM;
#pragma unroll
      for (int h = 0; h < IW; h++)
        if (h >= hleft && h < NL) Uij -= L[h] * Uw[IW - NL + h][j + NL - h];
      Ui[j] = Uij;
    }
  }
#pragma unroll
  for (int c = 0; c < IW; c++) Lc[c] = (c - (IW - NL) >= 0) ? L[(c - (IW - NL)) >= 0 ? c - (IW - NL) : 0] : 0.0;
}

__device__ __forceinline__ void factor_row(const OpMode& o, int nr, int i, double ri, const double (&Uw)[IW][IW + 1],
                                           double (&Lc)[IW], double (&Ui)[IW + 1]) {
  if (i >= IW) factor_row_nl<IW>(o, nr, i, ri, Uw, Lc, Ui);
  else if (i == 2) factor_row_nl<2>(o, nr, i, ri, Uw, Lc, Ui);
  else if (i == 1) factor_row_nl<1>(o, nr, i, ri, Uw, Lc, Ui);
  else factor_row_nl<0>(o, nr, i, ri, Uw, Lc, Ui);
}

__device__ __forceinline__ void shift_u(double (&Uw)[IW][IW + 1], const double (&Ui)[IW + 1]) {
#pragma unroll
  for (int s = 0; s < IW - 1; s++)
#pragma unroll
    for (int j = 0; j <= IW; j++) Uw[s][j] = Uw[s + 1][j];
#pragma unroll
  for (int j = 0; j <= IW; j++) Uw[IW - 1][j] = Ui[j];
}

// ------------------------------------------------------------------ factors
__global__ __launch_bounds__(256) void factor_kernel(Geom g, const double* __restrict__ r,
                                                     const double* __restrict__ Lu,
                                                     const double* __restrict__ Lp, double* __restrict__ F,
                                                     float* __restrict__ F32) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int kind = blockIdx.y;
  if (k >= g.modes) return;
  const ModeQ q = modeq(g, k);
  if (kind == 1 && q.it != g.nt / 2) return;  // stored factor sets only (fidx)
  if (kind >= 2 && q.it > g.nt / 2) return;
  const OpMode o = op_mode(g, Lu, Lp, kind, q.m, q.l);
  double Uw[IW][IW + 1];
#pragma unroll
  for (int s = 0; s < IW; s++)
#pragma unroll
    for (int j = 0; j <= IW; j++) Uw[s][j] = 0.0;
  for (int i = 0; i < g.nr; i++) {
    double Lc[IW], Ui[IW + 1];
    factor_row(o, g.nr, i, r[i], Uw, Lc, Ui);
    double fv[7];
#pragma unroll
    for (int c = 0; c < IW; c++) fv[c] = Lc[c];
    fv[3] = 1.0 / Ui[0];
#pragma unroll
    for (int j = 1; j <= IW; j++) fv[3 + j] = Ui[j];
#pragma unroll
    for (int c = 0; c < 7; c++) {
      F[fidx(g, kind, i, c, k)] = fv[c];
      if (F32 && kind <= 2) F32[fidx(g, kind, i, c, k)] = float(fv[c]);
    }
    shift_u(Uw, Ui);
  }
}

// Column scratch of one workgroup at FP64, or (mixed form, |m| >= m32) as
// FP32 pairs.  Both are indexed with the FP64 sweep-side indices; the FP32
// view is offset by the workgroup region's first FP64 index, so its values
// fill the first half of the workgroup's own region.
template <bool S32>
struct Scr;
template <>
struct Scr<false> {
  double2* p;
  __device__ static Scr at(double2* base, size_t) { return Scr{base}; }
  __device__ double2 ld(size_t i) const { return p[i]; }
  __device__ void st(size_t i, double2 v) const { p[i] = v; }
};
template <>
struct Scr<true> {
  float2* p;
  __device__ static Scr at(double2* base, size_t region0) { return Scr{reinterpret_cast<float2*>(base) + region0}; }
  __device__ double2 ld(size_t i) const {
    const float2 v = p[i];
    return make_double2(v.x, v.y);
  }
  __device__ void st(size_t i, double2 v) const { p[i] = make_float2(float(v.x), float(v.y)); }
};

// ------------------------------------------------------------------ sweep
// Workgroup = 2 waves: azimuthal rows it and NT - it (0 with NT/2) over one
// chunk of 64 axial modes.  One kernel per pass (PASS 0..3 = A..D): the
// column moves between passes through SR/SP in any case, and each pass then
// gets its own register allocation (and occupancy).
// B = 1: batched launch, one pipe per blockIdx.y with its arguments from tab
// (the single-pipe form B = 0 reads only its own argument block).
// HOLD (pass D): 1 keeps the four wall rows in registers across the
// coefficient solve, 0 walks them twice (fewer registers, one more wave per
// SIMD). The FP64 form is faster with the first, the mixed form with the
// second. The two compile to different instruction sequences (contraction),
// so a pipe's form fixes which one it runs, batched or not.
// INC = 1 (refined mixed form, single pipes): the velocity solves are for the
// increment delta = u^{k+1} - u^n, whose right-hand side is formed in FP64 from
// small quantities, so the FP32 column storage of |m| >= m32 rounds delta,
// not u (one step of iterative refinement from u^n):
//   A1 (PASS 4, predictor only): E' = (1/Re) L'u^n + (1 - d_im) N(u^n) and
//        Dh = div(H u^n) with H u^n = u^n/dt - (d_im/Re) L'u^n, L' = Lu + diag
//        (addLap_k's operator), for every row;
//   A:   rhs' = E' + d_im N(u^k) = rhs - H u^n; div(rhs) = div(rhs') + Dh;
//   C:   wall rows -u^n (so that u^n + delta vanishes there);
//   D:   back substitution gives delta; u^n + delta takes the place of the
//        solution from there on (wall divergence, correction, output).
template <int PASS, int PRED, int B, int HOLD = 0, int INC = 0>
__global__ __launch_bounds__(128) void sweep_kernel(Lin2Args A, const Lin2Args* tab) {
  if constexpr (B != 0) {
    A = tab[blockIdx.y];
    A.r = gmem(tab, A.r);
    A.rdr = gmem(tab, A.rdr);
    A.uc = gmem(tab, A.uc);
    A.qc = gmem(tab, A.qc);
    A.Dc = gmem(tab, A.Dc);
    A.Luc = gmem(tab, A.Luc);
    A.Lu = gmem(tab, A.Lu);
    A.Lp = gmem(tab, A.Lp);
    A.F = gmem(tab, A.F);
    if (A.F32) A.F32 = gmem(tab, A.F32);  // nullable: tested before re-deriving
    A.PH = gmem(tab, A.PH);
    A.invc = gmem(tab, A.invc);
    A.N = gmem(tab, A.N);
    A.E = gmem(tab, A.E);
    A.u = gmem(tab, A.u);
    A.out = gmem(tab, A.out);
    A.SR = gmem(tab, A.SR);
    A.SP = gmem(tab, A.SP);
    A.norm = gmem(tab, A.norm);
    if constexpr (PASS == 3) {  // members of the other form are the other launch's
      // The form is tested on a copy the compiler cannot see through: knowing
      // the form past this test would let it drop the other form's paths
      // and then contract the remaining arithmetic differently from the
      // single-pipe kernel, whose results the batched one must match bit for bit.
      int fp64 = __builtin_amdgcn_readfirstlane(A.form64 ? 1 : 0);
      asm volatile("" : "+s"(fp64));
      if (HOLD != fp64) return;
    }
  }
  const Geom& g = A.g;
  if (step_done(g)) return;
  const int nch = (g.nz + 63) / 64;
  // ord 1: the chunks of one azimuthal row, adjacent in memory, run in one XCD's L2
  const int lb = A.ord ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int pair = lb / nch, ch = lb - pair * nch;
  if (pair >= g.nt / 2) return;  // beyond this pipe's extent
  const int nth = g.nt / 2;
  const int it = (threadIdx.x < 64) ? pair : (pair == 0 ? nth : g.nt - pair);
  const int l = ch * 64 + (threadIdx.x & 63);
  if (l >= g.nz) return;
  const int k = it * g.nz + l;
  const ModeQ q = modeq(g, k);
  // sweep-side bases of this lane: factors u+ at this mode; u- = u+ at -m,
  // the other wave's lane (the Nyquist row has its own u- set, row 0 is its
  // own mirror); uz, p and the pressure homogeneous solution at |m|, the
  // wave-0 lane of the workgroup (Nyquist row: its own region)
  const int nr = g.nr;
  const int lane = threadIdx.x & 63;
  const bool nyq = pair == 0 && threadIdx.x >= 64;
  const size_t fb_up = (size_t)lb * nr * FR_UP + threadIdx.x;
  const size_t fb_um = nyq ? f_u1(g) + (size_t)ch * nr * FR_UP + threadIdx.x
                           : (pair == 0 ? fb_up : (size_t)lb * nr * FR_UP + (threadIdx.x ^ 64));
  const size_t ab = nyq ? sw_nblk(g) + ch : (size_t)lb;
  const size_t fb_uz = f_abs(g) + ab * nr * FR_ABS + lane;
  const size_t fb_p = fb_uz + f_absn(g);
  // the mixed form reads the u+, u-, uz factors as FP32 for |m| >= m32 (the
  // velocity Helmholtz operators, diagonally dominant through 1/dt); one |m|
  // per wave: rows pair and NT - pair, or the Nyquist row
  const bool f32w = __builtin_amdgcn_readfirstlane(A.F32 != nullptr && (nyq ? nth : pair) >= A.m32 ? 1 : 0) != 0;
  const size_t phb = ab * nr * 128 + lane;              // PH row i: phb + 128 i (+64: p_h)
  const size_t sb3 = (size_t)lb * nr * 384 + threadIdx.x;  // SR, E row i, component c: sb3 + 384 i + 128 c
  const size_t sb1 = (size_t)lb * nr * 128 + threadIdx.x;  // SP row i: sb1 + 128 i
  // N(u) is tile-major (dr_col): this mode's value of level j at nb0 + j npl,
  // components ncs apart; a wave's 64 modes are one tile row
  const size_t nb0 = (size_t)it * 64 + dr_col(g, l), npl = dr_plane(g), ncs = dr_cs(g);
  const size_t n = g.n;
  const int MS = g.modes;
  const double d_im = g.d_im;
  // Radial tables are read through the scalar cache: the parity of |m| is
  // the same for every lane of a wave (rows pair and NT - pair, or 0 and
  // NT/2), and the row index of a pass is uniform.
  const int par_u = __builtin_amdgcn_readfirstlane(abs((abs(q.m) % 2) - 1));  // parity table of ur, ut (u+, u-)
  const int par_p = 1 - par_u;                                                 // parity table of uz, p
  const ctab Dcu = ctab(A.Dc) + par_u * STEN * nr;
  const ctab Dcp = ctab(A.Dc) + par_p * STEN * nr;
  const ctab rt = ctab(A.r);
  const double2 z2 = make_double2(0.0, 0.0);
  // Column scratch precision of this workgroup: the mixed form holds SR, SP
  // and E of |m| >= m32 as FP32 pairs (the rhs, eliminated rows and pressure
  // of the fluctuation modes; the mean flow and low |m| stay FP64).  Both
  // waves have the same |m| (pair 0, rows 0 and NT/2, stays FP64).
  const bool s32b = A.s32 != 0 && pair >= A.m32;
  // The factor precision of this wave (F32W) and the storage of N(u) (N32,
  // FP32 in the mixed form) are compile-time branches too: run-time tests
  // inside the walks made the compiler drain outstanding loads before the
  // loads of each branch (two or three memory round trips per row).
  auto body = [&](auto s32_tag, auto f32_tag, auto n32_tag) {
  constexpr bool S32 = decltype(s32_tag)::value;
  constexpr bool F32W = decltype(f32_tag)::value;
  constexpr bool N32 = decltype(n32_tag)::value;
  const Scr<S32> SR = Scr<S32>::at(A.SR, (size_t)lb * nr * 384);
  const Scr<S32> SP = Scr<S32>::at(A.SP, (size_t)lb * nr * 128);
  const Scr<S32> E = Scr<S32>::at(A.E, (size_t)lb * nr * 384);

  // ============================== pass A1 (up, INC predictor): E', Dh
  if constexpr (PASS == 4) {
    const Scr<S32> DH = Scr<S32>::at(A.DH, (size_t)lb * nr * 128);
    double2 uw[3][7];      // u^n +/- rows t-6..t
    double2 Hr[7];         // (H u^n)_r rows i-3..i+3
    double2 Ht[4], Hz[4];  // (H u^n)_t, (H u^n)_z rows i..i+3
#pragma unroll
    for (int w = 0; w < 7; w++) {
      Hr[w] = z2;
#pragma unroll
      for (int c = 0; c < 3; c++) uw[c][w] = z2;
    }
#pragma unroll
    for (int w = 0; w < 4; w++) Ht[w] = Hz[w] = z2;
    const double ire = 1.0 / g.re, aim = d_im / g.re, idt = 1.0 / g.dt;
    for (int t = 0; t < nr + 6; t++) {
      double2 a = z2, b = z2, c = z2;
      {  // row t, or zero past the wall: the loads themselves stay in bounds
        const size_t h = (size_t)(t < nr ? t : nr - 1) * MS + k;
        const double2 a0 = A.u[h], b0 = A.u[n + h], c0 = A.u[2 * n + h];
        if (t < nr) {
          a = a0;
          b = b0;
          c = c0;
          pm_of(a, b);
        }
      }
#pragma unroll
      for (int w = 0; w < 6; w++) {
        uw[0][w] = uw[0][w + 1];
        uw[1][w] = uw[1][w + 1];
        uw[2][w] = uw[2][w + 1];
      }
      uw[0][6] = a;
      uw[1][6] = b;
      uw[2][6] = c;
      const int j = t - 3;  // row whose 7-point window is complete
      double2 H0 = z2, H1 = z2, H2 = z2;
      if (j >= 0 && j < nr) {
        const size_t hn = nb0 + (size_t)j * npl;  // tile-major N(u^n)
        double2 nn[3];
        if constexpr (N32) {
          const float2* N32 = reinterpret_cast<const float2*>(A.N);
          const float2 f0 = N32[hn], f1 = N32[ncs + hn], f2 = N32[2 * ncs + hn];
          nn[0] = make_double2(f0.x, f0.y);
          nn[1] = make_double2(f1.x, f1.y);
          nn[2] = make_double2(f2.x, f2.y);
        } else {
          nn[0] = A.N[hn];
          nn[1] = A.N[ncs + hn];
          nn[2] = A.N[2 * ncs + hn];
        }
        pm_of(nn[0], nn[1]);
        const double rj = rt[j], r2 = rj * rj;
        double2 hu[3];
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {  // L'u^n as in addLap_k (mod_linear.cu:12-51)
          const ctab Lc = ctab(A.Luc) + ((cc < 2) ? par_u : par_p) * STEN * nr + STEN * j;
          double px = 0.0, py = 0.0;
#pragma unroll
          for (int w = 0; w < 7; w++) {
            px += Lc[w] * uw[cc][w].x;
            py += Lc[w] * uw[cc][w].y;
          }
          const double2 aux = uw[cc][3];
          double de = -q.kz * q.kz - q.kt * q.kt / r2;
          if (cc == 0) de += -(1.0 / r2) - (2.0 * q.kt / r2);
          else if (cc == 1) de += -(1.0 / r2) + (2.0 * q.kt / r2);
          px += de * aux.x;
          py += de * aux.y;
          E.st(sb3 + (size_t)j * 384 + cc * 128,
               make_double2(ire * px + (1.0 - d_im) * nn[cc].x, ire * py + (1.0 - d_im) * nn[cc].y));
          hu[cc] = make_double2(idt * aux.x - aim * px, idt * aux.y - aim * py);
        }
        rt_of(hu[0], hu[1]);
        H0 = hu[0];
        H1 = hu[1];
        H2 = hu[2];
      }
#pragma unroll
      for (int w = 0; w < 6; w++) Hr[w] = Hr[w + 1];
      Hr[6] = H0;
#pragma unroll
      for (int w = 0; w < 3; w++) {
        Ht[w] = Ht[w + 1];
        Hz[w] = Hz[w + 1];
      }
      Ht[3] = H1;
      Hz[3] = H2;
      const int i = j - 3;
      if (i >= 0 && i < nr) {  // div_k (mod_radialFD.cu:67-96) of H u^n, wall row zero
        double2 dv = z2;
        if (i < nr - 1) {
          const ctab Dr = Dcu + STEN * i;
          double px = 0.0, py = 0.0;
#pragma unroll
          for (int w = 0; w < 7; w++) {
            px += Dr[w] * Hr[w].x;
            py += Dr[w] * Hr[w].y;
          }
          const double rK = rt[i];
          const double2 ur = Hr[3], ut = Ht[0], uz = Hz[0];
          px += ur.x / rK;
          py += ur.y / rK;
          px += -q.kt * ut.y / rK;
          py += q.kt * ut.x / rK;
          px += -q.kz * uz.y;
          py += q.kz * uz.x;
          dv = make_double2(px, py);
        }
        DH.st(sb1 + (size_t)i * 128, dv);
      }
    }
  }

  // ============================== pass A (up)
  if constexpr (PASS == 0) {
    constexpr bool XPRED = PRED && !INC;  // INC: the predictor's explicit part comes from A1
    constexpr int LAG = XPRED ? 6 : 3;
    const Scr<S32> DH = Scr<S32>::at(INC ? A.DH : A.SP, (size_t)lb * nr * 128);
    double2 uw[3][7];     // predictor: u^n +/- rows t-6..t
    double2 Rr[7];        // R_r rows i-3..i+3
    double2 Rt[4], Rz[4]; // R_t, R_z rows i..i+3
    double2 yp[3];        // pressure elimination rows i-3..i-1
#pragma unroll
    for (int w = 0; w < 7; w++) {
      Rr[w] = z2;
#pragma unroll
      for (int c = 0; c < 3; c++) uw[c][w] = z2;
    }
#pragma unroll
    for (int w = 0; w < 4; w++) Rt[w] = Rz[w] = z2;
#pragma unroll
    for (int w = 0; w < 3; w++) yp[w] = z2;
    const double cn = (1.0 - d_im) / g.re, idt = 1.0 / g.dt;
    for (int t = 0; t < nr + LAG; t++) {
      if (XPRED) {
        double2 a = z2, b = z2, c = z2;
        {  // row t, or zero past the wall: the loads themselves stay in bounds
          const size_t h = (size_t)(t < nr ? t : nr - 1) * MS + k;
          const double2 a0 = A.u[h], b0 = A.u[n + h], c0 = A.u[2 * n + h];
          if (t < nr) {
            a = a0;
            b = b0;
            c = c0;
            pm_of(a, b);
          }
        }
#pragma unroll
        for (int w = 0; w < 6; w++) {
          uw[0][w] = uw[0][w + 1];
          uw[1][w] = uw[1][w + 1];
          uw[2][w] = uw[2][w + 1];
        }
        uw[0][6] = a;
        uw[1][6] = b;
        uw[2][6] = c;
      }
      // rhs row j
      const int j = XPRED ? t - 3 : t;
      // pressure elimination row i = j - 3: its factors (and Dh) are loaded
      // before this iteration's stores, which the compiler would not move
      // loads across (one memory round trip per row instead of two)
      const int i = j - 3;
      const int ic = i < 0 ? 0 : (i < nr ? i : nr - 1);
      const double* fpi = A.F + fb_p + (size_t)ic * FR_ABS;
      const double fl0 = fpi[0], fl1 = fpi[64], fl2 = fpi[128];
      double2 dhi = z2;
      if constexpr (INC) dhi = DH.ld(sb1 + (size_t)ic * 128);
      double2 R0 = z2, R1 = z2, R2 = z2;
      if (j >= 0 && j < nr) {
        const size_t hn = nb0 + (size_t)j * npl;  // tile-major N(u)
        double2 n0, n1, n2;
        if constexpr (N32) {
          const float2* N32 = reinterpret_cast<const float2*>(A.N);
          const float2 f0 = N32[hn], f1 = N32[ncs + hn], f2 = N32[2 * ncs + hn];
          n0 = make_double2(f0.x, f0.y);
          n1 = make_double2(f1.x, f1.y);
          n2 = make_double2(f2.x, f2.y);
        } else {
          n0 = A.N[hn];
          n1 = A.N[ncs + hn];
          n2 = A.N[2 * ncs + hn];
        }
        pm_of(n0, n1);
        double2 rh[3];
        if (XPRED) {
          // addLap_k (mod_linear.cu:12-51) on u^n, then rhs = N + that
          const double rj = rt[j], r2 = rj * rj;
          const double2 nn[3] = {n0, n1, n2};
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const ctab Lc = ctab(A.Luc) + ((c < 2) ? par_u : par_p) * STEN * nr + STEN * j;
            double px = 0.0, py = 0.0;
#pragma unroll
            for (int w = 0; w < 7; w++) {
              px += Lc[w] * uw[c][w].x;
              py += Lc[w] * uw[c][w].y;
            }
            const double2 aux = uw[c][3];
            double de = -q.kz * q.kz - q.kt * q.kt / r2;
            if (c == 0) de += -(1.0 / r2) - (2.0 * q.kt / r2);
            else if (c == 1) de += -(1.0 / r2) + (2.0 * q.kt / r2);
            px += de * aux.x;
            py += de * aux.y;
            px *= cn;
            py *= cn;
            px += idt * aux.x;
            py += idt * aux.y;
            rh[c] = make_double2(nn[c].x + px, nn[c].y + py);
            E.st(sb3 + (size_t)j * 384 + c * 128,
                 make_double2(px + (1.0 - d_im) * nn[c].x, py + (1.0 - d_im) * nn[c].y));
          }
        } else {
          const size_t he = sb3 + (size_t)j * 384;
          const double2 e0 = E.ld(he), e1 = E.ld(he + 128), e2 = E.ld(he + 256);
          rh[0] = make_double2(e0.x + d_im * n0.x, e0.y + d_im * n0.y);
          rh[1] = make_double2(e1.x + d_im * n1.x, e1.y + d_im * n1.y);
          rh[2] = make_double2(e2.x + d_im * n2.x, e2.y + d_im * n2.y);
        }
        rt_of(rh[0], rh[1]);
        R0 = rh[0];
        R1 = rh[1];
        R2 = rh[2];
        const size_t hs = sb3 + (size_t)j * 384;
        SR.st(hs, R0);
        SR.st(hs + 128, R1);
        SR.st(hs + 256, R2);
      }
#pragma unroll
      for (int w = 0; w < 6; w++) Rr[w] = Rr[w + 1];
      Rr[6] = R0;
#pragma unroll
      for (int w = 0; w < 3; w++) {
        Rt[w] = Rt[w + 1];
        Rz[w] = Rz[w + 1];
      }
      Rt[3] = R1;
      Rz[3] = R2;
      // divergence and pressure elimination at row i = j - 3
      if (i >= 0 && i < nr) {
        double2 dv = z2;
        if (i < nr - 1) {  // div_k (mod_radialFD.cu:67-96), wall row zero
          const ctab Dr = Dcu + STEN * i;
          double px = 0.0, py = 0.0;
#pragma unroll
          for (int w = 0; w < 7; w++) {
            px += Dr[w] * Rr[w].x;
            py += Dr[w] * Rr[w].y;
          }
          const double rK = rt[i];
          const double2 ur = Rr[3], ut = Rt[0], uz = Rz[0];
          px += ur.x / rK;
          py += ur.y / rK;
          px += -q.kt * ut.y / rK;
          py += q.kt * ut.x / rK;
          px += -q.kz * uz.y;
          py += q.kz * uz.x;
          dv = make_double2(px, py);
          if constexpr (INC) {  // div(rhs) = div(rhs - H u^n) + div(H u^n)
            dv.x += dhi.x;
            dv.y += dhi.y;
          }
        }
        const double l0 = fl0, l1 = fl1, l2 = fl2;
        dv.x -= l0 * yp[0].x;
        dv.y -= l0 * yp[0].y;
        dv.x -= l1 * yp[1].x;
        dv.y -= l1 * yp[1].y;
        dv.x -= l2 * yp[2].x;
        dv.y -= l2 * yp[2].y;
        SP.st(sb1 + (size_t)i * 128, dv);
        yp[0] = yp[1];
        yp[1] = yp[2];
        yp[2] = dv;
      }
    }
  }

  // ============================== pass B (down)
  if constexpr (PASS == 1) {
    double2 pw[3];  // p rows t+1..t+3
#pragma unroll
    for (int w = 0; w < 3; w++) pw[w] = z2;
    for (int t = nr - 1; t >= 0; t--) {
      const double* fp = A.F + fb_p + (size_t)t * FR_ABS;
      const double ru = fp[3 * 64];
      const double u1 = fp[4 * 64], u2 = fp[5 * 64], u3 = fp[6 * 64];
      const size_t h = sb1 + (size_t)t * 128;
      double2 x = SP.ld(h);
      x.x -= u1 * pw[0].x;
      x.y -= u1 * pw[0].y;
      x.x -= u2 * pw[1].x;
      x.y -= u2 * pw[1].y;
      x.x -= u3 * pw[2].x;
      x.y -= u3 * pw[2].y;
      const double2 pn = make_double2(x.x * ru, x.y * ru);
      SP.st(h, pn);
      pw[2] = pw[1];
      pw[1] = pw[0];
      pw[0] = pn;
    }
  }

  // ============================== pass C (up)
  if constexpr (PASS == 2) {
    double2 pw[7];      // p rows j-3..j+3
    double2 yw[3][3];   // eliminated rows j-3..j-1 of u+, u-, uz
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int w = 0; w < 3; w++) yw[c][w] = z2;
#pragma unroll
    for (int w = 0; w < 7; w++) pw[w] = z2;
    for (int t = 0; t < nr + IW; t++) {
#pragma unroll
      for (int w = 0; w < 6; w++) pw[w] = pw[w + 1];
      {  // row t, or zero past the wall: the load itself stays in bounds
        const double2 pv = SP.ld(sb1 + (size_t)(t < nr ? t : nr - 1) * 128);
        pw[6] = (t < nr) ? pv : z2;
      }
      const int j = t - IW;
      if (j < 0) continue;
      // the L factors of the three solves at row j, loaded before the
      // row's stores (the compiler does not move loads across them)
      double Lc3[3][IW];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        // u+, u-: rows of FR_UP, slots of 128; uz: rows of FR_ABS, slots of 64
        const size_t fo = c == 0 ? fb_up + (size_t)j * FR_UP
                                 : (c == 1 ? fb_um + (size_t)j * FR_UP : fb_uz + (size_t)j * FR_ABS);
        const int fs = c == 2 ? 64 : 128;
        if constexpr (F32W) {
#pragma unroll
          for (int w = 0; w < IW; w++) Lc3[c][w] = A.F32[fo + w * fs];
        } else {
#pragma unroll
          for (int w = 0; w < IW; w++) Lc3[c][w] = A.F[fo + w * fs];
        }
      }
      // addGrad_k (mod_radialFD.cu:115-143)
      const size_t h = sb3 + (size_t)j * 384;
      const ctab Dr = Dcp + STEN * j;
      double px = 0.0, py = 0.0;
#pragma unroll
      for (int w = 0; w < 7; w++) {
        px += Dr[w] * pw[w].x;
        py += Dr[w] * pw[w].y;
      }
      const double rK = rt[j];
      const double2 ph = pw[3];
      double2 rh[3] = {SR.ld(h), SR.ld(h + 128), SR.ld(h + 256)};
      rh[0].x -= px;
      rh[0].y -= py;
      rh[1].x -= -q.kt * ph.y / rK;
      rh[1].y -= q.kt * ph.x / rK;
      rh[2].x -= -q.kz * ph.y;
      rh[2].y -= q.kz * ph.x;
      pm_of(rh[0], rh[1]);
      if (j == nr - 1) {  // write_bcon_k: wall rows zero (INC: delta = -u^n there)
        rh[0] = rh[1] = rh[2] = z2;
        if constexpr (INC) {
          const size_t hw = (size_t)j * MS + k;
          double2 a = A.un[hw], b = A.un[n + hw];
          pm_of(a, b);
          const double2 c = A.un[2 * n + hw];
          rh[0] = make_double2(-a.x, -a.y);
          rh[1] = make_double2(-b.x, -b.y);
          rh[2] = make_double2(-c.x, -c.y);
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {  // forward elimination
        const double* Lc = Lc3[c];
        double2 y = rh[c];
        y.x -= Lc[0] * yw[c][0].x;
        y.y -= Lc[0] * yw[c][0].y;
        y.x -= Lc[1] * yw[c][1].x;
        y.y -= Lc[1] * yw[c][1].y;
        y.x -= Lc[2] * yw[c][2].x;
        y.y -= Lc[2] * yw[c][2].y;
        SR.st(h + c * 128, y);
        yw[c][0] = yw[c][1];
        yw[c][1] = yw[c][2];
        yw[c][2] = y;
      }
    }
  }

  // ============================== pass D (down)
  if constexpr (PASS == 3) {
    double2 xw[3][3];   // x rows i+1..i+3
    double xh[3][3];    // homogeneous wall-unit solutions rows i+1..i+3
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int w = 0; w < 3; w++) {
        xw[c][w] = z2;
        xh[c][w] = 0.0;
      }
    auto backsub = [&](int i, double2* xo, double* ho) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t fo = c == 0 ? fb_up + (size_t)i * FR_UP
                                 : (c == 1 ? fb_um + (size_t)i * FR_UP : fb_uz + (size_t)i * FR_ABS);
        const int fs = c == 2 ? 64 : 128;
        double ru, u1, u2, u3;
        // FP32 factors by the body's compile-time flag (a run-time test
        // here made the scheduler serialise the FP32 loads of the mixed
        // form); the HOLD kernels (FP64 form) keep the run-time test, with
        // which they were measured 3-7% faster
        if (HOLD ? f32w : F32W) {
          ru = A.F32[fo + 3 * fs];
          u1 = A.F32[fo + 4 * fs];
          u2 = A.F32[fo + 5 * fs];
          u3 = A.F32[fo + 6 * fs];
        } else {
          ru = A.F[fo + 3 * fs];
          u1 = A.F[fo + 4 * fs];
          u2 = A.F[fo + 5 * fs];
          u3 = A.F[fo + 6 * fs];
        }
        double2 x = SR.ld(sb3 + (size_t)i * 384 + c * 128);
        x.x -= u1 * xw[c][0].x;
        x.y -= u1 * xw[c][0].y;
        x.x -= u2 * xw[c][1].x;
        x.y -= u2 * xw[c][1].y;
        x.x -= u3 * xw[c][2].x;
        x.y -= u3 * xw[c][2].y;
        x = make_double2(x.x * ru, x.y * ru);
        double hh = (i == nr - 1) ? 1.0 : 0.0;
        hh -= u1 * xh[c][0];
        hh -= u2 * xh[c][1];
        hh -= u3 * xh[c][2];
        hh *= ru;
        xw[c][2] = xw[c][1];
        xw[c][1] = xw[c][0];
        xw[c][0] = x;
        xh[c][2] = xh[c][1];
        xh[c][1] = xh[c][0];
        xh[c][0] = hh;
        xo[c] = x;
        ho[c] = hh;
      }
      if constexpr (INC) {  // the solution is u^n + delta (FP64)
        const size_t hu = (size_t)i * MS + k;
        double2 a = A.un[hu], b = A.un[n + hu];
        pm_of(a, b);
        const double2 c = A.un[2 * n + hu];
        xo[0] = make_double2(xo[0].x + a.x, xo[0].y + a.y);
        xo[1] = make_double2(xo[1].x + b.x, xo[1].y + b.y);
        xo[2] = make_double2(xo[2].x + c.x, xo[2].y + c.y);
      }
    };
    // wall divergence of the provisional solution (divWall_k on r/theta
    // form) from the four wall rows. Without HOLD the walk that emits the
    // solution repeats their back substitution (same arithmetic, same
    // values), which keeps the pass at 3 waves per SIMD instead of 2.
    double2 urw[4], utw = z2, uzw = z2;
    double2 xs[HOLD ? 4 : 1][3];  // HOLD: wall rows NR-1..NR-4
    double xhs[HOLD ? 4 : 1][3];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      double2 xo[3];
      double ho[3];
      backsub(nr - 1 - s, xo, ho);
      if constexpr (HOLD) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          xs[s][c] = xo[c];
          xhs[s][c] = ho[c];
        }
      }
      double2 a = xo[0], b = xo[1];
      rt_of(a, b);
      urw[s] = a;
      if (s == 0) {
        utw = b;
        uzw = xo[2];
      }
    }
    if constexpr (!HOLD) {
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int w = 0; w < 3; w++) {
          xw[c][w] = z2;
          xh[c][w] = 0.0;
        }
    }
    double2 dv;
    {
      const ctab Dr = Dcu + STEN * (nr - 1);
      double px = 0.0, py = 0.0;
#pragma unroll
      for (int w = 0; w < 4; w++) {  // columns NR-4..NR-1 = urw[3..0]
        px += Dr[w] * urw[3 - w].x;
        py += Dr[w] * urw[3 - w].y;
      }
      px += urw[0].x;
      py += urw[0].y;
      px += -q.kt * utw.y;
      py += q.kt * utw.x;
      px += -q.kz * uzw.y;
      py += q.kz * uzw.x;
      dv = make_double2(px, py);
    }
    double2 cf[4];  // compcoef_k
    {
      const double2 nd = make_double2(-dv.x, -dv.y);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const double2 im = A.invc[(size_t)k * 4 + j];
        cf[j] = make_double2(nd.x * im.x - nd.y * im.y, nd.x * im.y + nd.y * im.x);
      }
    }
    double num = 0.0, den = 0.0;
    // emit row i: pd = {dr p_h, p_h} and uc (the corrector's previous
    // iterate) come from a prefetched row, or are loaded here when null
    auto emit = [&](int i, const double2* xo, const double* ho, const double* pd = nullptr,
                    const double2* uc = nullptr) {
      const size_t h = (size_t)i * MS + k;
      // pressure-gradient homogeneous solution: (dr p_h -+ kt p_h / r, 0), (0, kz p_h)
      const double dph = pd ? pd[0] : A.PH[phb + (size_t)i * 128];
      const double ph = pd ? pd[1] : A.PH[phb + (size_t)i * 128 + 64];
      const double tq = q.kt * ph / rt[i];
      const double2 hp[3] = {make_double2(dph - tq, 0.0), make_double2(dph + tq, 0.0), make_double2(-0.0, q.kz * ph)};
      const double2 hv[3] = {make_double2(ho[0], 0.0), make_double2(ho[1], 0.0), make_double2(0.0, ho[2])};
      double2 o[3] = {xo[0], xo[1], xo[2]};
      rt_of(o[0], o[1]);  // decoupleBackward / decoupleForward round trip of corBCon
      pm_of(o[0], o[1]);
#pragma unroll
      for (int c = 0; c < 3; c++) {  // addBCon_k
        const double ax = cf[c].x * hv[c].x - cf[c].y * hv[c].y + cf[3].x * hp[c].x - cf[3].y * hp[c].y;
        const double ay = cf[c].x * hv[c].y + cf[c].y * hv[c].x + cf[3].x * hp[c].y + cf[3].y * hp[c].x;
        o[c] = make_double2(o[c].x + ax, o[c].y + ay);
      }
      rt_of(o[0], o[1]);
      if (!PRED && k != 0) {  // mcor_k; mode (0,0) is measured after its fix-up
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double2 u1 = uc ? uc[c] : A.u[c * n + h];
          const double dx = u1.x - o[c].x, dy = u1.y - o[c].y;
          const double a1 = dx * dx + dy * dy, a2 = u1.x * u1.x + u1.y * u1.y;
          num = (a1 > num || a1 != a1) ? a1 : num;
          den = (a2 > den || a2 != a2) ? a2 : den;
        }
      }
      A.out[h] = o[0];
      A.out[n + h] = o[1];
      A.out[2 * n + h] = o[2];
    };
    if constexpr (HOLD) {
#pragma unroll
      for (int s = 0; s < 4; s++) emit(nr - 1 - s, xs[s], xhs[s]);
      for (int i = nr - 5; i >= 0; i--) {
        double2 xo[3];
        double ho[3];
        backsub(i, xo, ho);
        emit(i, xo, ho);
      }
    } else {
      // The walk that emits the solution, software-pipelined: the loads of
      // row i - 1 (factors, column, p_h, previous iterate) are issued before
      // row i is solved and stored, so each row's memory latency overlaps
      // the previous row's arithmetic (the compiler does not move loads
      // across the stores of a row by itself).  Same arithmetic as backsub.
      using FT = std::conditional_t<F32W, float, double>;
      struct RowD {
        FT f[3][4];      // ru, u1, u2, u3 of the u+, u-, uz solves
        double2 x[3];    // eliminated column rows
        double pd[2];    // dr p_h, p_h
        double2 uc[3];   // corrector: previous iterate
        double2 un[3];   // INC: u^n (+/- form for the velocity pair)
      };
      auto loadrow = [&](int i, RowD& R) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const size_t fo = c == 0 ? fb_up + (size_t)i * FR_UP
                                   : (c == 1 ? fb_um + (size_t)i * FR_UP : fb_uz + (size_t)i * FR_ABS);
          const int fs = c == 2 ? 64 : 128;
#pragma unroll
          for (int w = 0; w < 4; w++) {
            if constexpr (F32W) R.f[c][w] = A.F32[fo + (3 + w) * fs];
            else R.f[c][w] = A.F[fo + (3 + w) * fs];
          }
          R.x[c] = SR.ld(sb3 + (size_t)i * 384 + c * 128);
        }
        R.pd[0] = A.PH[phb + (size_t)i * 128];
        R.pd[1] = A.PH[phb + (size_t)i * 128 + 64];
        const size_t h = (size_t)i * MS + k;
        if constexpr (!PRED) {
#pragma unroll
          for (int c = 0; c < 3; c++) R.uc[c] = A.u[c * n + h];
        }
        if constexpr (INC) {
#pragma unroll
          for (int c = 0; c < 3; c++) R.un[c] = A.un[c * n + h];
        }
      };
      auto solve = [&](int i, const RowD& R, double2* xo, double* ho) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double ru = R.f[c][0], u1 = R.f[c][1], u2 = R.f[c][2], u3 = R.f[c][3];
          double2 x = R.x[c];
          x.x -= u1 * xw[c][0].x;
          x.y -= u1 * xw[c][0].y;
          x.x -= u2 * xw[c][1].x;
          x.y -= u2 * xw[c][1].y;
          x.x -= u3 * xw[c][2].x;
          x.y -= u3 * xw[c][2].y;
          x = make_double2(x.x * ru, x.y * ru);
          double hh = (i == nr - 1) ? 1.0 : 0.0;
          hh -= u1 * xh[c][0];
          hh -= u2 * xh[c][1];
          hh -= u3 * xh[c][2];
          hh *= ru;
          xw[c][2] = xw[c][1];
          xw[c][1] = xw[c][0];
          xw[c][0] = x;
          xh[c][2] = xh[c][1];
          xh[c][1] = xh[c][0];
          xh[c][0] = hh;
          xo[c] = x;
          ho[c] = hh;
        }
        if constexpr (INC) {  // the solution is u^n + delta (FP64)
          double2 a = R.un[0], b = R.un[1];
          pm_of(a, b);
          const double2 c = R.un[2];
          xo[0] = make_double2(xo[0].x + a.x, xo[0].y + a.y);
          xo[1] = make_double2(xo[1].x + b.x, xo[1].y + b.y);
          xo[2] = make_double2(xo[2].x + c.x, xo[2].y + c.y);
        }
      };
      RowD cur, nxt;
      loadrow(nr - 1, cur);
      for (int i = nr - 1; i >= 0; i--) {
        loadrow(i > 0 ? i - 1 : 0, nxt);
        double2 xo[3];
        double ho[3];
        solve(i, cur, xo, ho);
        emit(i, xo, ho, cur.pd, PRED ? nullptr : cur.uc);
        cur = nxt;
      }
    }
    if (k == 0) {  // modeZZ_k, adj_flux (mod_integrate.cu:63-68, mod_cublasflux.cu:90-103)
      double qa = 0.0;
      for (int i = 0; i < nr; i++) {
        const size_t h = (size_t)i * MS;
        A.out[h] = z2;
        A.out[n + h].y = 0.0;
        A.out[2 * n + h].y = 0.0;
        qa += A.rdr[i] * A.out[2 * n + h].x;
      }
      const double corr = (g.qdes - qa) / A.qc[0];
      for (int i = 0; i < nr; i++) {
        const size_t h = (size_t)i * MS;
        A.out[2 * n + h].x += corr * A.uc[i];
        if (!PRED) {
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const double2 u1 = A.u[c * n + h], o = A.out[c * n + h];
            const double dx = u1.x - o.x, dy = u1.y - o.y;
            const double a1 = dx * dx + dy * dy, a2 = u1.x * u1.x + u1.y * u1.y;
            num = (a1 > num || a1 != a1) ? a1 : num;
            den = (a2 > den || a2 != a2) ? a2 : den;
          }
        }
      }
    }
    if (!PRED) {
      // num, den are >= 0 or NaN, whose bit patterns order as the values do
      // (NaN above every finite value and +inf). A full wave combines its
      // lanes' maxima and issues one atomic per value; a partly filled wave
      // (the last axial chunk) keeps one per lane.
      unsigned long long a = __double_as_longlong(num), b = __double_as_longlong(den);
      if (__builtin_amdgcn_read_exec() == ~0ull) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          const unsigned long long ao = __shfl_xor(a, o), bo = __shfl_xor(b, o);
          a = ao > a ? ao : a;
          b = bo > b ? bo : b;
        }
        if (lane == 0) {
          atomicMax(&A.norm[0], a);
          atomicMax(&A.norm[1], b);
        }
      } else {
        atomicMax(&A.norm[0], a);
        atomicMax(&A.norm[1], b);
      }
    }
  }
  };
  // the FP64 form has FP64 N(u), scratch and factors (s32b and f32w false)
  using T1 = std::true_type;
  using T0 = std::false_type;
  if (!A.n32) body(T0{}, T0{}, T0{});
  else if (s32b) {
    if (f32w) body(T1{}, T1{}, T1{});
    else body(T1{}, T0{}, T1{});
  } else {
    if (f32w) body(T0{}, T1{}, T1{});
    else body(T0{}, T0{}, T1{});
  }
}

size_t factor_count(const Geom& g) { return f_abs(g) + 2 * f_absn(g); }
size_t factor32_count(const Geom& g) { return f_abs(g) + f_absn(g); }

void launch_factor(const Geom& g, const double* r, const double* Lu, const double* Lp, double* F, float* F32,
                   hipStream_t s) {
  dim3 grid((g.modes + 255) / 256, 4);
  factor_kernel<<<grid, 256, 0, s>>>(g, r, Lu, Lp, F, F32);
}

// pass D runs as HOLD = 1 for the FP64-form pipes and as HOLD = 0 for the
// mixed-form ones (a batched launch of both forms makes both launches)
template <int PRED>
static void sweep_passes_inc(const Lin2Args& a, dim3 grid, hipStream_t s) {
  if (PRED) sweep_kernel<4, PRED, 0, 0, 1><<<grid, 128, 0, s>>>(a, nullptr);
  sweep_kernel<0, PRED, 0, 0, 1><<<grid, 128, 0, s>>>(a, nullptr);
  sweep_kernel<1, PRED, 0, 0, 1><<<grid, 128, 0, s>>>(a, nullptr);
  sweep_kernel<2, PRED, 0, 0, 1><<<grid, 128, 0, s>>>(a, nullptr);
  sweep_kernel<3, PRED, 0, 0, 1><<<grid, 128, 0, s>>>(a, nullptr);
}

template <int PRED, int B>
static void sweep_passes(const Lin2Args& a, const Lin2Args* tab, dim3 grid, bool fp64, bool mixed, hipStream_t s) {
  sweep_kernel<0, PRED, B><<<grid, 128, 0, s>>>(a, tab);
  sweep_kernel<1, PRED, B><<<grid, 128, 0, s>>>(a, tab);
  sweep_kernel<2, PRED, B><<<grid, 128, 0, s>>>(a, tab);
  if (fp64) sweep_kernel<3, PRED, B, 1><<<grid, 128, 0, s>>>(a, tab);
  if (mixed) sweep_kernel<3, PRED, B, 0><<<grid, 128, 0, s>>>(a, tab);
}

void launch_sweep2(const Lin2Args& a, hipStream_t s) {
  const dim3 grid((a.g.nt / 2) * ((a.g.nz + 63) / 64));
  if (a.inc) {
    if (a.predictor) sweep_passes_inc<1>(a, grid, s);
    else sweep_passes_inc<0>(a, grid, s);
    return;
  }
  const bool fp64 = a.form64 != 0;
  if (a.predictor) sweep_passes<1, 0>(a, nullptr, grid, fp64, !fp64, s);
  else sweep_passes<0, 0>(a, nullptr, grid, fp64, !fp64, s);
}

void launch_sweep_batch(const Lin2Args* tab, int count, int predictor, int max_nt, int max_nz, bool fp64,
                        bool mixed, hipStream_t s) {
  const dim3 grid((max_nt / 2) * ((max_nz + 63) / 64), count);
  Lin2Args dummy{};
  if (predictor) sweep_passes<1, 1>(dummy, tab, grid, fp64, mixed, s);
  else sweep_passes<0, 1>(dummy, tab, grid, fp64, mixed, s);
}

}  // namespace nsp
