// Linear-phase set-up and bookkeeping (the per-step solve itself is the
// streaming sweep of linear2.hip):
//   * setup_influence_kernel: the homogeneous wall-unit and pressure
//     solutions and the inverse influence matrix (setBCon,
//     mod_boundary.cu:205-255) and the flux profile (setCublasFlux,
//     mod_cublasflux.cu:56-80), one thread per Fourier mode (it, l) walking
//     its radial column with the reference's banded LU (LUsolve_*_k);
//   * the step record and convergence-test kernels of a captured step
//     (single pipe and ensemble), the mixed form's N(u) widening, additive
//     perturbations;
//   * diag_kernel: per-mode partial sums of the diagnostics.
// Column vectors of the set-up live in mode-fastest global arrays, so every
// access of a wavefront is one contiguous 1 KiB (FP64 complex) line.
#include <mutex>

#include "nsp_internal.h"

namespace nsp {

struct Mode {
  int k, it, l, m;
  double kt, kz;
};

__device__ __forceinline__ Mode mode_of(const Geom& g, int k) {
  Mode q;
  q.k = k;
  q.it = k / g.nz;
  q.l = k - q.it * g.nz;
  q.m = q.it < g.nt / 2 ? q.it : q.it - g.nt;
  q.kt = double(q.m);  // (PI2/LT)*it with LT = PI2 (head.h:21)
  q.kz = g.kz0 * double(q.l);
  return q;
}

__device__ __forceinline__ size_t at(const Geom& g, int i, int k) { return (size_t)i * g.modes + k; }

// Radial stencil  sum_j tab[ip0 + 7i + j + i0] f[left+j]  (derivr_k, mod_radialFD.cu:10-30)
__device__ __forceinline__ double2 stencil(const Geom& g, const double* tab, int ip0, int i,
                                           const double2* f, int k) {
  const int left = max(0, i - IW), right = min(i + IW, g.nr - 1);
  const int len = right + 1 - left, i0 = i + IW - right;
  double px = 0.0, py = 0.0;
  for (int j = 0; j < len; j++) {
    const double de = tab[ip0 + STEN * i + j + i0];
    const double2 a = f[at(g, left + j, k)];
    px += de * a.x;
    py += de * a.y;
  }
  return make_double2(px, py);
}

// Banded LU without pivoting (Doolittle, half band IW) fused with forward
// elimination, then back substitution: LUsolve_{up,um,uz,p}_k
// (mod_linear.cu:55-646).  kind 0: u+, 1: u-, 2: uz, 3: p.  x may alias rhs.
// U factor rows live in the mode-fastest scratch array Us.
__device__ void lu_solve(const Geom& g, const DevTables& t, double* Us, int kind, const Mode& q,
                         const double2* rhs, double2* x) {
  const double* LHS = (kind == 3) ? t.Lp : t.Lu;
  const int fp = (kind <= 1) ? 1 : 0;
  const int ip0 = abs((abs(q.m) % 2) - fp) * STEN * g.nr;
  const bool vel = kind < 3;
  const bool p00 = (kind == 3) && (q.l == 0 && q.m == 0);
  const double a = g.d_im / g.re, idt = 1.0 / g.dt;
  const double kt = q.kt, kz = q.kz;
  const int k = q.k;
  const int nr = g.nr;
#define UF(row, j) Us[(size_t)((row) * (1 + IW) + (j)) * g.modes + k]
  for (int i = 0; i < nr; i++) {
    const int left = max(0, i - IW), right = min(i + IW, nr - 1);
    const int i0 = i + IW - right;
    const int nl = i - left;
    const double r2 = t.r[i] * t.r[i];
    const bool last = (i == nr - 1);
    double2 xi = rhs[at(g, i, k)];
    double L[IW];
#pragma unroll
    for (int j = 0; j < IW; j++) {
      if (j < nl) {
        double XM = LHS[ip0 + j + i0 + STEN * i];
        if (!last) {
          if (vel) XM = -a * XM;
        } else if (p00) {
          XM = 0.0;
        }
        double Lj = XM;
#pragma unroll
        for (int h = 0; h < IW; h++)
          if (h < j) Lj -= L[h] * UF(h + left, j - h);
        Lj = Lj / UF(j + left, 0);
        L[j] = Lj;
        const double2 xj = x[at(g, j + left, k)];
        xi.x -= Lj * xj.x;
        xi.y -= Lj * xj.y;
      }
    }
    int coun = nl;
    for (int j = 0; j < right - i + 1; j++) {
      double XM = LHS[ip0 + STEN * i + i0 + coun];
      if (!last) {
        if (j == 0) {
          if (kind == 0) XM += -(1.0 / r2) - (2.0 * kt / r2) - kz * kz - kt * kt / r2;
          else if (kind == 1) XM += -(1.0 / r2) + (2.0 * kt / r2) - kz * kz - kt * kt / r2;
          else XM += -kz * kz - kt * kt / r2;
          if (vel) XM = idt - a * XM;
        } else if (vel) {
          XM = -a * XM;
        }
      } else if (p00) {
        XM = (j == 0) ? 1.0 : 0.0;
      }
      const int hleft = max(0, coun - IW);
      coun += 1;
      double Uij = XM;
#pragma unroll
      for (int h = 0; h < IW; h++)
        if (h >= hleft && h < nl) Uij -= L[h] * UF(h + left, j + nl - h);
      UF(i, j) = Uij;
    }
    x[at(g, i, k)] = xi;
  }
  {
    const int i = nr - 1;
    double2 xi = x[at(g, i, k)];
    const double d = UF(i, 0);
    xi.x = xi.x / d;
    xi.y = xi.y / d;
    x[at(g, i, k)] = xi;
  }
  for (int i = nr - 2; i > -1; i--) {
    const int right = min(i + IW, nr - 1);
    double2 xi = x[at(g, i, k)];
    for (int j = 1; j < right - i + 1; j++) {
      const double u = UF(i, j);
      const double2 xj = x[at(g, i + j, k)];
      xi.x -= u * xj.x;
      xi.y -= u * xj.y;
    }
    const double d = UF(i, 0);
    xi.x = xi.x / d;
    xi.y = xi.y / d;
    x[at(g, i, k)] = xi;
  }
#undef UF
}

// divWall_k (mod_radialFD.cu:34-63) on one column in r/theta form
__device__ __forceinline__ double2 div_wall(const Geom& g, const DevTables& t, const Mode& q,
                                            const double2* ur, const double2* ut, const double2* uz) {
  const int i = g.nr - 1;
  const int ip0 = abs((abs(q.m) % 2) - 1) * STEN * g.nr;
  double2 prod = stencil(g, t.D, ip0, i, ur, q.k);
  const size_t h = at(g, i, q.k);
  prod.x += ur[h].x;
  prod.y += ur[h].y;
  prod.x += -q.kt * ut[h].y;
  prod.y += q.kt * ut[h].x;
  prod.x += -q.kz * uz[h].y;
  prod.y += q.kz * uz[h].x;
  return prod;
}

__device__ __forceinline__ void to_pm(double2& a, double2& b) {  // decoupleForward_k
  const double2 ur = a, ut = b;
  a = make_double2(ur.x - ut.y, ur.y + ut.x);
  b = make_double2(ur.x + ut.y, ur.y - ut.x);
}
__device__ __forceinline__ void to_rt(double2& a, double2& b) {  // decoupleBackward_k
  const double2 up = a, um = b;
  a = make_double2(0.5 * (up.x + um.x), 0.5 * (up.y + um.y));
  b = make_double2(0.5 * (up.y - um.y), -0.5 * (up.x - um.x));
}
__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ void norm_reset_kernel(unsigned long long* norm, const int* skip) {
  if (skip && *skip) return;
  if (threadIdx.x < 2) norm[threadIdx.x] = 0ull;
}

// FP32 pairs to FP64 pairs (the mixed form's stored N(u), for nsp_nonlinear).
__global__ void widen_kernel(const float2* __restrict__ in, double2* __restrict__ out, size_t count) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < count) {
    const float2 v = in[q];
    out[q] = make_double2(v.x, v.y);
  }
}

// Tile-major spectral field to the reference layout (nsp_nonlinear's N(u)):
// one thread per value of the output.
__global__ void untile_kernel(Geom g, const void* __restrict__ in, int f32, double2* __restrict__ out) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * g.n) return;
  const int c = (int)(q / g.n);
  const size_t h = q - (size_t)c * g.n;
  const int i = (int)(h / g.modes);
  const int k = (int)(h - (size_t)i * g.modes);
  const int it = k / g.nz, l = k - it * g.nz;
  const size_t s = (size_t)c * dr_cs(g) + (size_t)i * dr_plane(g) + (size_t)it * 64 + dr_col(g, l);
  if (f32) {
    const float2 v = reinterpret_cast<const float2*>(in)[s];
    out[q] = make_double2(v.x, v.y);
  } else {
    out[q] = reinterpret_cast<const double2*>(in)[s];
  }
}

// u += f over `count` complex values (additive seeded perturbations).
__global__ void add_kernel(double2* __restrict__ u, const double2* __restrict__ f, size_t count) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < count) {
    const double2 a = u[q], b = f[q];
    u[q] = make_double2(a.x + b.x, a.y + b.y);
  }
}

// Ensemble forms of the three bookkeeping kernels below (one thread per member).
__global__ void ens_begin_kernel(StepRec* rec, int count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < count) {
    rec[j].st[0] = 0;
    rec[j].st[1] = 0;
    rec[j].st[2] = 0;
    rec[j].st[3] = 0;
    rec[j].err = 1e98;
  }
}
__global__ void ens_norm_reset_kernel(unsigned long long* const* norms, const StepRec* rec, int count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < count && !rec[j].st[0]) {
    norms[j][0] = 0ull;
    norms[j][1] = 0ull;
  }
}
__global__ void ens_converge_kernel(unsigned long long* const* norms, StepRec* rec, int count, double tol, int it,
                                    int maxit) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count || rec[j].st[0]) return;
  const unsigned long long* nm = norms[j];
  const double e = __longlong_as_double((long long)nm[0]) / __longlong_as_double((long long)nm[1]);
  rec[j].err = e;
  rec[j].st[1] = it;
  const bool nan = e != e;
  if (!(e > tol) || it >= maxit) {
    rec[j].st[0] = 1;
    rec[j].st[2] = (nan ? 1 : 0) | ((!nan && e > tol) ? 2 : 0);
  }
}

// Start of a captured step: not converged, no corrector yet.
__global__ void step_begin_kernel(int* st, double* err) {
  if (threadIdx.x == 0) {
    st[0] = 0;
    st[1] = 0;
    st[2] = 0;
    err[0] = 1e98;
  }
}

// After corrector iteration `it` (measucorr + the loop test of integrate(),
// mod_integrate.cu:116-131): err = max |ui - u|^2 / max |ui|^2; the loop
// goes on while err > tol and it < maxit, so a NaN err also ends it.
__global__ void converge_kernel(const unsigned long long* norm, int* st, double* err, double tol, int it,
                                int maxit) {
  if (threadIdx.x != 0 || st[0]) return;
  const double e = __longlong_as_double((long long)norm[0]) / __longlong_as_double((long long)norm[1]);
  err[0] = e;
  st[1] = it;
  const bool nan = e != e;
  if (!(e > tol) || it >= maxit) {
    st[0] = 1;
    st[2] = (nan ? 1 : 0) | ((!nan && e > tol) ? 2 : 0);
  }
}

// ------------------------------------------------------------------ setup
// setBCon (mod_boundary.cu:205-255) and setCublasFlux (mod_cublasflux.cu:56-80),
// one thread per mode.
__global__ __launch_bounds__(256) void setup_influence_kernel(Geom g, DevTables t, DevInfluence b,
                                                              double2* P, double* Us, double* PH) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.modes) return;
  const Mode q = mode_of(g, k);
  const size_t n = g.n;
  const int nr = g.nr;
  auto unit_rhs = [&](double2 wall) {
    for (int i = 0; i < nr - 1; i++) P[at(g, i, k)] = make_double2(0.0, 0.0);
    P[at(g, nr - 1, k)] = wall;
  };
  unit_rhs(make_double2(1.0, 0.0));
  lu_solve(g, t, Us, 0, q, P, b.uh);
  unit_rhs(make_double2(1.0, 0.0));
  lu_solve(g, t, Us, 1, q, P, b.uh + n);
  unit_rhs(make_double2(0.0, 1.0));
  lu_solve(g, t, Us, 2, q, P, b.uh + 2 * n);
  unit_rhs(make_double2(-1.0, 0.0));
  lu_solve(g, t, Us, 3, q, P, P);  // P now holds the homogeneous pressure p_h
  // wall divergences of the four homogeneous solutions (in r/theta form)
  const size_t hw = at(g, nr - 1, k);
  double2 d1, d2, d3, d4;
  {
    const int ip0 = abs((abs(q.m) % 2) - 1) * STEN * nr;
    const int i = nr - 1, left = i - IW;
    // DIV1: (u_h+, 0) -> r/theta: ur = up/2, ut = (up.y/2, -up.x/2)
    double px = 0, py = 0, sx = 0, sy = 0;
    for (int j = 0; j < IW + 1; j++) {
      const double de = t.D[ip0 + STEN * i + j + IW];
      double2 a = b.uh[at(g, left + j, k)], z0 = make_double2(0.0, 0.0);
      to_rt(a, z0);
      px += de * a.x;
      py += de * a.y;
      double2 c0 = make_double2(0.0, 0.0), bm = b.uh[n + at(g, left + j, k)];
      to_rt(c0, bm);
      sx += de * c0.x;
      sy += de * c0.y;
    }
    double2 a = b.uh[hw], z0 = make_double2(0.0, 0.0);
    to_rt(a, z0);
    d1 = make_double2(px + a.x - q.kt * z0.y, py + a.y + q.kt * z0.x);
    double2 c0 = make_double2(0.0, 0.0), bm = b.uh[n + hw];
    to_rt(c0, bm);
    d2 = make_double2(sx + c0.x - q.kt * bm.y, sy + c0.y + q.kt * bm.x);
    // DIV3: (0, 0, u_hz)
    const double2 uzw = b.uh[2 * n + hw];
    d3 = make_double2(0.0 + 0.0 - q.kz * uzw.y, 0.0 + 0.0 + q.kz * uzw.x);
  }
  // u_hP = grad p_h (derivr flag 0, derivt, derivz) in r/theta form
  {
    const int ip0 = abs((abs(q.m) % 2) - 0) * STEN * nr;
    const bool habs = k / g.nz <= g.nt / 2;  // the sweep reads PH at |m| only
    for (int i = 0; i < nr; i++) {
      const size_t h = at(g, i, k);
      const double2 dr = stencil(g, t.D, ip0, i, P, k);
      const double2 ph = P[h];
      if (habs) {  // p_h is real: its d/dr and value, stored for |m| (sweep-side layout)
        PH[sw_abs_idx(g, 2, i, 0, k)] = dr.x;
        PH[sw_abs_idx(g, 2, i, 1, k)] = ph.x;
      }
      b.uhP[h] = dr;
      b.uhP[n + h] = make_double2(-q.kt * ph.y / t.r[i], q.kt * ph.x / t.r[i]);
      b.uhP[2 * n + h] = make_double2(-q.kz * ph.y, q.kz * ph.x);
    }
    d4 = div_wall(g, t, q, b.uhP, b.uhP + n, b.uhP + 2 * n);
    for (int i = 0; i < nr; i++) {
      const size_t h = at(g, i, k);
      double2 a = b.uhP[h], c = b.uhP[n + h];
      to_pm(a, c);
      b.uhP[h] = a;
      b.uhP[n + h] = c;
    }
  }
  // column 3 of the inverse of [[1,0,0,a14],[0,1,0,a24],[0,0,i,a34],[a41,a42,a43,a44]]
  // (calc_invM_k, mod_boundary.cu:66-196: entries 3, 7, 11, 15)
  {
    const double2 a14 = b.uhP[hw], a24 = b.uhP[n + hw], a34 = b.uhP[2 * n + hw];
    const double2 a41 = d1, a42 = d2, a43 = d3, a44 = d4;
    double2 aux, Det;
    aux.x = a44.x - a24.x * a42.x + a24.y * a42.y - a14.x * a41.x + a14.y * a41.y;
    aux.y = a44.y - a24.x * a42.y - a24.y * a42.x - a14.x * a41.y - a14.y * a41.x;
    Det.x = -aux.y - a34.x * a43.x + a34.y * a43.y;
    Det.y = aux.x - a34.x * a43.y - a34.y * a43.x;
    const double D1 = Det.x * Det.x + Det.y * Det.y;
    Det.x = Det.x / D1;
    Det.y = -Det.y / D1;
    double2 IA[4];
    IA[0] = make_double2(a14.y, -a14.x);
    IA[1] = make_double2(a24.y, -a24.x);
    IA[2] = make_double2(-a34.x, -a34.y);
    IA[3] = make_double2(0.0, 1.0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double ax = IA[j].x * Det.x - IA[j].y * Det.y;
      const double ay = IA[j].x * Det.y + IA[j].y * Det.x;
      b.invc[(size_t)k * 4 + j] = make_double2(ax, ay);
    }
  }
  // flux profile: mode (0,0) uz-Helmholtz solution for rhs 1 with wall value 0
  if (k == 0) {
    for (int i = 0; i < nr - 1; i++) P[at(g, i, 0)] = make_double2(1.0, 0.0);
    P[at(g, nr - 1, 0)] = make_double2(0.0, 0.0);
    lu_solve(g, t, Us, 2, q, P, P);
    double qc = 0.0;
    for (int i = 0; i < nr; i++) {
      t.uc[i] = P[at(g, i, 0)].x;
      qc += t.rdr[i] * t.uc[i];
    }
    t.qc[0] = qc;
  }
}

// ------------------------------------------------------------------ diagnostics
// Per-mode partial sums: [0] KE, [1] |rdr|-weighted |div|^2, [2] max wall |u|,
// [3] |wall divergence|.
__device__ void diag_mode(const Geom& g, const DevTables& t, const double2* u, int k, double* out) {
  const Mode q = mode_of(g, k);
  const size_t n = g.n;
  const double w = (q.l == 0) ? 1.0 : 2.0;
  double ke = 0.0, d2 = 0.0;
  const int ip0 = abs((abs(q.m) % 2) - 1) * STEN * g.nr;
  for (int i = 0; i < g.nr; i++) {
    const size_t h = at(g, i, k);
    for (int c = 0; c < 3; c++) {
      const double2 a = u[c * n + h];
      ke += 0.5 * w * t.rdr[i] * (a.x * a.x + a.y * a.y);
    }
    const double rK = t.r[i];
    double2 prod = stencil(g, t.D, ip0, i, u, k);
    const double2 ur = u[h], ut = u[n + h], uz = u[2 * n + h];
    prod.x += ur.x / rK;
    prod.y += ur.y / rK;
    prod.x += -q.kt * ut.y / rK;
    prod.y += q.kt * ut.x / rK;
    prod.x += -q.kz * uz.y;
    prod.y += q.kz * uz.x;
    d2 += w * fabs(t.rdr[i]) * (prod.x * prod.x + prod.y * prod.y);
  }
  double wall = 0.0;
  for (int c = 0; c < 3; c++) {
    const double2 a = u[c * n + at(g, g.nr - 1, k)];
    wall = fmax(wall, sqrt(a.x * a.x + a.y * a.y));
  }
  const double2 dw = div_wall(g, t, q, u, u + n, u + 2 * n);
  out[0] = ke;
  out[1] = d2;
  out[2] = wall;
  out[3] = sqrt(dw.x * dw.x + dw.y * dw.y);
}

__global__ __launch_bounds__(256) void diag_kernel(Geom g, DevTables t, const double2* u, double* part) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.modes) return;
  diag_mode(g, t, u, k, part + (size_t)k * 4);
}

// Allocation-free diagnostics: the per-mode partials of a block of 256 modes
// are combined in the block (sums of [0], [1], maxima of [2], [3]) into one
// partial set per block, part[4 * blockIdx.x + j]; a single block then
// combines the block partials and adds the values of the mode (0,0) column
// (centre-line velocity, friction velocity, bulk flux: wrt_frc_k,
// mod_io_nonlinear.cu:40-52).  Every block writes its own partial set, so
// nothing has to start at zero.
constexpr int kDiagThreads = 256;
__device__ __forceinline__ void block_combine(double (*sm)[kDiagThreads], int tid) {
  for (int s = kDiagThreads / 2; s > 0; s >>= 1) {
    if (tid < s) {
      sm[0][tid] += sm[0][tid + s];
      sm[1][tid] += sm[1][tid + s];
      sm[2][tid] = fmax(sm[2][tid], sm[2][tid + s]);
      sm[3][tid] = fmax(sm[3][tid], sm[3][tid + s]);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kDiagThreads) void diag_block_kernel(Geom g, DevTables t, const double2* u,
                                                                   double* part) {
  __shared__ double sm[4][kDiagThreads];
  const int tid = threadIdx.x;
  const int k = blockIdx.x * kDiagThreads + tid;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (k < g.modes) diag_mode(g, t, u, k, v);
#pragma unroll
  for (int j = 0; j < 4; j++) sm[j][tid] = v[j];
  __syncthreads();
  block_combine(sm, tid);
  if (tid < 4) part[4 * (size_t)blockIdx.x + tid] = sm[tid][0];
}

__global__ __launch_bounds__(kDiagThreads) void diag_final_kernel(Geom g, DevTables t, const double2* u,
                                                                   const double* part, int nblk, double* out) {
  __shared__ double sm[4][kDiagThreads];
  const int tid = threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = tid; b < nblk; b += kDiagThreads) {
    v[0] += part[4 * (size_t)b];
    v[1] += part[4 * (size_t)b + 1];
    v[2] = fmax(v[2], part[4 * (size_t)b + 2]);
    v[3] = fmax(v[3], part[4 * (size_t)b + 3]);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) sm[j][tid] = v[j];
  __syncthreads();
  block_combine(sm, tid);
  if (tid == 0) {
    const double2* uz = u + 2 * g.n;  // mode (0,0) of uz at level i: uz[i * modes]
    double uc = 0.0, du = 0.0, q = 0.0;
    for (int i = 0; i < IW + 1; i++) uc += uz[(size_t)i * g.modes].x * t.dr0[i];
    for (int i = 0; i < STEN; i++) du += uz[(size_t)(g.nr - STEN + i) * g.modes].x * t.dr1[i];
    for (int i = 0; i < g.nr; i++) q += t.rdr[i] * uz[(size_t)i * g.modes].x;
    out[0] = uc;
    out[1] = sqrt(fabs(du) / g.re);
    out[2] = sm[0][0];
    out[3] = sqrt(sm[1][0]);
    out[4] = sm[2][0];
    out[5] = sm[3][0];
    out[6] = q;
  }
}

// The block partials of the allocation-free diagnostics live in this
// file-scope array, one copy per device, held until the process ends.  Host
// code never passes the array's name to a kernel (in host code it is only a
// stand-in for the symbol): diag_scratch_init obtains each device's address
// once, with that device current, and launchers use that address.
static __device__ __align__(256) double scratch[4 * kDiagMaxBlocks];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

hipError_t diag_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return hipSuccess;
}

double* diag_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ------------------------------------------------------------------ launchers
void launch_norm_reset(unsigned long long* norm, const int* skip, hipStream_t s) {
  norm_reset_kernel<<<1, 64, 0, s>>>(norm, skip);
}
void launch_add(double2* u, const double2* f, size_t count, hipStream_t s) {
  add_kernel<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(u, f, count);
}
void launch_ens_begin(StepRec* rec, int count, hipStream_t s) {
  ens_begin_kernel<<<(count + 63) / 64, 64, 0, s>>>(rec, count);
}
void launch_ens_norm_reset(unsigned long long* const* norms, const StepRec* rec, int count, hipStream_t s) {
  ens_norm_reset_kernel<<<(count + 63) / 64, 64, 0, s>>>(norms, rec, count);
}
void launch_ens_converge(unsigned long long* const* norms, StepRec* rec, int count, double tol, int it, int maxit,
                         hipStream_t s) {
  ens_converge_kernel<<<(count + 63) / 64, 64, 0, s>>>(norms, rec, count, tol, it, maxit);
}
void launch_untile(const Geom& g, const void* in, bool f32, double2* out, hipStream_t s) {
  untile_kernel<<<(unsigned)((3 * g.n + 255) / 256), 256, 0, s>>>(g, in, f32 ? 1 : 0, out);
}
void launch_widen(const float2* in, double2* out, size_t count, hipStream_t s) {
  widen_kernel<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(in, out, count);
}
void launch_step_begin(int* st, double* err, hipStream_t s) { step_begin_kernel<<<1, 64, 0, s>>>(st, err); }
void launch_converge(const unsigned long long* norm, int* st, double* err, double tol, int it, int maxit,
                     hipStream_t s) {
  converge_kernel<<<1, 64, 0, s>>>(norm, st, err, tol, it, maxit);
}
void launch_setup_influence(const Geom& g, const DevTables& t, const DevInfluence& b, double2* P, double* U,
                            double* PH, hipStream_t s) {
  setup_influence_kernel<<<(g.modes + 255) / 256, 256, 0, s>>>(g, t, b, P, U, PH);
}
void launch_diagnostics(const Geom& g, const DevTables& t, const double2* u, double* part, hipStream_t s) {
  diag_kernel<<<(g.modes + 255) / 256, 256, 0, s>>>(g, t, u, part);
}
void launch_diagnostics_noalloc(const Geom& g, const DevTables& t, const double2* u, double* part, double* out,
                                hipStream_t s) {
  const int nblk = (g.modes + kDiagThreads - 1) / kDiagThreads;  // <= kDiagMaxBlocks (checked by the caller)
  diag_block_kernel<<<nblk, kDiagThreads, 0, s>>>(g, t, u, part);
  diag_final_kernel<<<1, kDiagThreads, 0, s>>>(g, t, u, part, nblk, out);
}

}  // namespace nsp
