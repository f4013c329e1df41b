// nsp_oracle.cpp -- CPU FP64 restatement of the nsPipe-CUDA time step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
// implementation in nspipe-gpu_amd/.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load the library built from it, and only as
// the checker or the CPU baseline -- never as the thing measured or shipped.
//
// Every routine restates one routine of the reference (Mordered/nsPipe-GPU,
// Main_Code/src) with the same arithmetic order; each carries the file:line it
// follows.  Where the reference calls cuFFT or cuBLAS the oracle states the
// library's documented semantics instead (see fft_* and dot/idamax below).
//
// Pinning (what checks each part; see tests/test_oracle.py):
//   tables (init_fd, fd_weights, get_dr0)  : Fornberg exactness on monomials,
//        sum(rdr) == 1, rdr moments, dr0 interpolation exactness
//   FFT (fft2_inverse/forward)            : numpy.fft.irfft2 / rfft2
//   derivr/div/addGrad/addLap/derivt/z    : dense numpy operators built from
//        the tables, spectral identities
//   lusolve (4 kinds)                     : dense numpy.linalg.solve of the
//        assembled banded operator
//   setBCon / corBCon                     : numpy.linalg.inv of the 4x4
//        influence matrix, no-slip and wall divergence at round-off
//   adj_flux                              : bulk flux == Qdes to round-off
//   step                                  : rescaled Poiseuille is a discrete
//        fixed point; m=0 swirl eigenvector decays by the closed-form factor;
//        at m != 0 a perturbed Poiseuille step equals a dense numpy one-mode
//        step map at round-off, whose least-stable eigenvalue matches an
//        independent Chebyshev-collocation eigen-solve of linearised pipe flow
//   nonlinear                             : brute-force numpy pseudo-spectral
//        evaluation on a tiny grid
//   diagnostics                           : closed forms on alpha(1-r^2)
//        (u_cl, u_tau, flux), Parseval sum of the physical samples (KE),
//        analytic divergence and wall values of polynomial fields
//   reference_ic                          : numpy rfft2 of the init_per
//        samples, truncated as padBackward_k does
// Parity unpinned (no reference run is possible): cuFFT rounding and its
// handling of non-Hermitian kz=0 input, cublasDdot summation order, nvcc FMA
// contraction.  All are below the 1e-12 parity bar.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct cplx { double x, y; };

constexpr double PI2 = 6.283185307179586;   // head.h:12
constexpr double LT  = PI2;                 // head.h:21
constexpr int STEN = 7;                     // head.h:44
constexpr int IW   = STEN / 2;              // head.h:45

// ---------------------------------------------------------------------------
// Mixed-radix complex FFT (decimation in time), cuFFT sign convention:
// sign=-1 forward  X[k] = sum x[n] e^{-2 pi i nk/n},  sign=+1 inverse,
// no normalisation (mod_fft.cu:39,47; normalisation is applied separately).
// ---------------------------------------------------------------------------
struct FFT {
  int n = 0;
  std::vector<int> radices;
  std::vector<cplx> tw;  // e^{-2 pi i k/n}
  void init(int n_) {
    n = n_;
    radices.clear();
    int m = n;
    while (m % 4 == 0) { radices.push_back(4); m /= 4; }
    while (m % 2 == 0) { radices.push_back(2); m /= 2; }
    while (m % 3 == 0) { radices.push_back(3); m /= 3; }
    while (m % 5 == 0) { radices.push_back(5); m /= 5; }
    for (int p = 7; m > 1; p += 2) while (m % p == 0) { radices.push_back(p); m /= p; }
    tw.resize(n);
    for (int k = 0; k < n; k++) {
      double a = -PI2 * double(k) / double(n);
      tw[k] = {std::cos(a), std::sin(a)};
    }
  }
  inline cplx w(int e, int sign) const {
    cplx t = tw[e % n];
    return sign < 0 ? t : cplx{t.x, -t.y};
  }
  void rec(const cplx* in, size_t istr, cplx* out, int len, int sign, int twstride,
           int lvl) const {
    if (len == 1) { out[0] = in[0]; return; }
    int r = radices[lvl], m = len / r;
    for (int q = 0; q < r; q++)
      rec(in + q * istr, istr * r, out + q * m, m, sign, twstride * r, lvl + 1);
    cplx t[16], o[16];
    for (int k = 0; k < m; k++) {
      for (int q = 0; q < r; q++) {
        cplx v = out[q * m + k];
        cplx ww = w(q * k * twstride, sign);
        t[q] = {v.x * ww.x - v.y * ww.y, v.x * ww.y + v.y * ww.x};
      }
      for (int s = 0; s < r; s++) {
        cplx acc = {0.0, 0.0};
        for (int q = 0; q < r; q++) {
          cplx ww = w(((q * s) % r) * (n / r), sign);
          acc.x += t[q].x * ww.x - t[q].y * ww.y;
          acc.y += t[q].x * ww.y + t[q].y * ww.x;
        }
        o[s] = acc;
      }
      for (int s = 0; s < r; s++) out[k + s * m] = o[s];
    }
  }
  // out[0..n) = DFT of in[0], in[istr], ... ; out contiguous
  void exec(const cplx* in, size_t istr, cplx* out, int sign) const {
    rec(in, istr, out, n, sign, 1, 0);
  }
};

struct Params {
  int NR, NT, NZ;
  double LZ, Re, dt, d_im;
  int maxit;
  double tol, qdes;
};

struct Oracle {
  Params P;
  int NR, NT, NZ, NTP, NZP, NZF, MODES;
  size_t N;
  std::vector<double> r, D, Lu, Lp, rdr, dr0, dr1;
  // influence-matrix data (mod_boundary.cu:200)
  std::vector<cplx> uh[3], uhP[3], invM, coef, div1;
  std::vector<double> uc;  // mod_cublasflux.cu:69-76
  double Qc = 0.0;
  // state and work fields (main.cu:30-33, mod_integrate.cu:100-107)
  std::vector<cplx> u[3], rhs[3], uw[3], rhsw[3], ui[3], dtmp[3];
  std::vector<double> d1, d2;
  FFT fft_t, fft_z;
  int nthreads = 1;
  // > 0: nonlinear() evaluates only levels 0..sample_levels-1 (the others are
  // left untouched).  Used only to time bounded samples of the CPU baseline.
  int sample_levels = 0;

  // ------------------------------------------------------------------ index
  inline int mode_m(int k) const { int it = k / NZ; return it < NT / 2 ? it : it - NT; }
  inline int mode_l(int k) const { return k % NZ; }
  inline double kz_of(int l) const { return (PI2 / P.LZ) * double(l); }
  inline double kt_of(int m) const { return (PI2 / LT) * double(m); }

  // ------------------------------------------------- mod_radialFD.cu:365-395
  void fd_weights(double* c, const double* rau, int i0, int left, int right, int m) {
    int n = 1 + right - left;
    double c1, c2, c3, c4, c5, x0 = rau[i0 + IW];
    std::vector<double> x(n);
    for (int i = 0; i < STEN * STEN; i++) c[i] = 0.0;
    c[0] = 1.0; c1 = 1.0;
    for (int i = 0; i < n; i++) x[i] = rau[i + i0];
    c4 = x[0] - x0;
    for (int i = 1; i < n; i++) {
      int mn = std::min(i, m); c2 = 1.0; c5 = c4; c4 = x[i] - x0;
      for (int j = 0; j < i; j++) {
        c3 = x[i] - x[j]; c2 = c2 * c3;
        if (j == i - 1) {
          for (int k = mn; k > 0; k--)
            c[i + n * k] = c1 * (double(k) * c[i - 1 + n * (k - 1)] - c5 * c[i - 1 + n * k]) / c2;
          c[i] = -c1 * c5 * c[i - 1] / c2;
        }
        for (int k = mn; k > 0; k--)
          c[j + n * k] = (c4 * c[j + n * k] - double(k) * c[j + (k - 1) * n]) / c3;
        c[j] = c4 * c[j] / c3;
      }
      c1 = c2;
    }
  }

  // ------------------------------------------------- mod_radialFD.cu:199-312
  // Grid, parity-folded D/D2 tables, rdr quadrature, Lu/Lp.  The rdr loop of
  // the reference runs k = 0..len inclusive (mod_radialFD.cu:272) and for
  // k = len reads c[ir+len*len], past the 49 doubles it allocates; that term
  // carries dr^(len+1)/(len+1)! and is taken as zero here (loop stops at len-1).
  void init_fd() {
    int hD = STEN * NR;
    r.assign(NR, 0.0); D.assign(2 * hD, 0.0); Lu.assign(2 * hD, 0.0);
    Lp.assign(2 * hD, 0.0); rdr.assign(NR, 0.0);
    std::vector<double> D2(2 * hD, 0.0), rau(NR + IW), r_i(NR + 1), c(STEN * STEN + 8, 0.0);
    int N_p = NR + int(std::sqrt(double(NR)));
    double dr;
    for (int i = N_p - NR; i < N_p; i++) {
      int ir = N_p - i - 1;
      r[i - N_p + NR] = 0.5 * (1.0 + std::cos(0.5 * PI2 * (double(ir) / double(N_p))));
    }
    for (int i = 0; i < 10; i++) {
      dr = 1.5 * r[0] - 0.5 * r[1];
      for (int j = 0; j < NR; j++) r[j] = r[j] * (1.0 + dr) - dr;
    }
    for (int i = 0; i < NR; i++) r_i[i + 1] = r[i];
    r_i[0] = 0.0;
    for (int i = 0; i < NR + IW; i++) rau[i] = (i < IW) ? -r[IW - i - 1] : r[i - IW];
    int m = 2, left, left2, right, len, i0, coun = IW, ir;
    for (int i = 0; i < NR; i++) {
      left = i; right = std::min(IW + IW + i, NR + IW - 1); len = 1 + right - left;
      fd_weights(c.data(), rau.data(), i, left, right, m);
      if (i < IW) {
        for (int j = 0; j < len; j++) {
          if (j < coun) {
            ir = coun - j - 1;
            D[ir + STEN * i] += c[len + j];       D[hD + ir + STEN * i] -= c[len + j];
            D2[ir + STEN * i] += c[2 * len + j];  D2[hD + ir + STEN * i] -= c[2 * len + j];
          } else {
            ir = j - coun;
            D[ir + STEN * i] += c[len + j];       D[hD + ir + STEN * i] += c[len + j];
            D2[ir + STEN * i] += c[2 * len + j];  D2[hD + ir + STEN * i] += c[2 * len + j];
          }
        }
        coun -= 1;
      } else {
        i0 = STEN - len;
        for (int j = 0; j < len; j++) {
          ir = i0 + j;
          D[ir + STEN * i] += c[len + j];       D[hD + ir + STEN * i] += c[len + j];
          D2[ir + STEN * i] += c[2 * len + j];  D2[hD + ir + STEN * i] += c[2 * len + j];
        }
      }
    }
    double e1, e2; coun = IW;
    for (int i = 0; i < NR; i++) {
      left = i; right = std::min(IW + IW + i, NR + IW - 1); len = 1 + right - left;
      left2 = std::max(0, i - IW);
      fd_weights(c.data(), rau.data(), i, left, right, len - 1);
      i0 = 0; e1 = 1.0; e2 = 1.0;
      if (i < IW) {
        i0 = coun;
        for (int j = 0; j < coun; j++) {
          ir = coun + coun - 1 - j;
          for (int k = 0; k < len; k++) c[ir + k * len] += c[j + k * len];
        }
        coun -= 1;
      }
      for (int k = 0; k < len; k++) {  // reference: k < len+1 (see header note)
        e1 = e1 * (r_i[std::min(i + 2, NR)] - r_i[i + 1]) / (double(k + 1));
        e2 = e2 * (r_i[i] - r_i[i + 1]) / (double(k + 1));
        for (int j = 0; j < len - i0; j++) {
          ir = i0 + j;
          rdr[left2 + j] += (e1 - e2) * r[i] * c[ir + k * len];
        }
      }
    }
    for (int i = 0; i < NR; i++) {
      if (i < NR - 1) {
        for (int j = 0; j < STEN; j++) {
          ir = j + STEN * i;
          Lu[ir] = D2[ir] + D[ir] / r[i];  Lu[hD + ir] = D2[hD + ir] + D[hD + ir] / r[i];
          Lp[ir] = D2[ir] + D[ir] / r[i];  Lp[hD + ir] = D2[hD + ir] + D[hD + ir] / r[i];
        }
      } else {
        for (int j = 0; j < STEN; j++) {
          ir = j + STEN * i;
          if (j < STEN - 1) { Lu[ir] = 0.0; Lu[hD + ir] = 0.0; }
          else { Lu[ir] = 1.0; Lu[hD + ir] = 1.0; }
          Lp[ir] = D[ir]; Lp[hD + ir] = D[hD + ir];
        }
      }
    }
  }

  // ------------------------------------------------- mod_radialFD.cu:316-350
  void get_dr0() {
    int n = IW + 1, m = 0;
    double c1, c2, c3, c4, c5, x0 = 0.0;
    std::vector<double> c(n * (m + 1), 0.0), x(n);
    c[0] = 1.0; c1 = 1.0;
    for (int i = 0; i < n; i++) x[i] = r[i];
    c4 = x[0] - x0;
    for (int i = 1; i < n; i++) {
      int mn = std::min(i, m); c2 = 1.0; c5 = c4; c4 = x[i] - x0;
      for (int j = 0; j < i; j++) {
        c3 = x[i] - x[j]; c2 = c2 * c3;
        if (j == i - 1) {
          for (int k = mn; k > 0; k--)
            c[i + n * k] = c1 * (double(k) * c[i - 1 + n * (k - 1)] - c5 * c[i - 1 + n * k]) / c2;
          c[i] = -c1 * c5 * c[i - 1] / c2;
        }
        for (int k = mn; k > 0; k--)
          c[j + n * k] = (c4 * c[j + n * k] - double(k) * c[j + (k - 1) * n]) / c3;
        c[j] = c4 * c[j] / c3;
      }
      c1 = c2;
    }
    dr0.assign(c.begin(), c.begin() + n);
    // get_dr1 (mod_radialFD.cu:355-358): wall row of the even D table
    dr1.assign(D.begin() + (NR - 1) * STEN, D.begin() + NR * STEN);
  }

  // ------------------------------------------------- mod_radialFD.cu:10-30
  void derivr(cplx* dur, const cplx* uu, int flag) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NR; i++)
      for (int k = 0; k < MODES; k++) {
        int it = mode_m(k), left = std::max(0, i - IW), right = std::min(i + IW, NR - 1);
        int ip0 = std::abs((std::abs(it) % 2) - flag) * STEN * NR;
        int len = right + 1 - left, i0 = i + IW - right;
        cplx prod = {0.0, 0.0};
        for (int j = 0; j < len; j++) {
          double de = D[ip0 + STEN * i + j + i0];
          cplx a = uu[(size_t)(left + j) * MODES + k];
          prod.x += de * a.x; prod.y += de * a.y;
        }
        dur[(size_t)i * MODES + k] = prod;
      }
  }

  // ------------------------------------------------- mod_radialFD.cu:34-63
  void divWall(cplx* dv, const cplx* ur, const cplx* ut, const cplx* uz) const {
    int i = NR - 1;
    for (int k = 0; k < MODES; k++) {
      size_t h = (size_t)i * MODES + k;
      int l = mode_l(k), it = mode_m(k), left = i - IW, right = NR - 1;
      double kt = kt_of(it), kz = kz_of(l);
      int ip0 = std::abs((std::abs(it) % 2) - 1) * STEN * NR;
      int len = right + 1 - left, i0 = i + IW - right;
      cplx prod = {0.0, 0.0};
      for (int j = 0; j < len; j++) {
        double de = D[ip0 + STEN * i + j + i0];
        cplx a = ur[(size_t)(left + j) * MODES + k];
        prod.x += de * a.x; prod.y += de * a.y;
      }
      prod.x += ur[h].x;        prod.y += ur[h].y;
      prod.x += -kt * ut[h].y;  prod.y += kt * ut[h].x;
      prod.x += -kz * uz[h].y;  prod.y += kz * uz[h].x;
      dv[k] = prod;
    }
  }

  // ------------------------------------------------- mod_radialFD.cu:67-96
  void divergence(cplx* dv, const cplx* ur, const cplx* ut, const cplx* uz) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NR; i++)
      for (int k = 0; k < MODES; k++) {
        size_t h = (size_t)i * MODES + k;
        int l = mode_l(k), it = mode_m(k), left = std::max(0, i - IW), right = std::min(i + IW, NR - 1);
        double kt = kt_of(it), kz = kz_of(l);
        int ip0 = std::abs((std::abs(it) % 2) - 1) * STEN * NR;
        double rK = r[i];
        int len = right + 1 - left, i0 = i + IW - right;
        cplx prod = {0.0, 0.0};
        for (int j = 0; j < len; j++) {
          double de = D[ip0 + STEN * i + j + i0];
          cplx a = ur[(size_t)(left + j) * MODES + k];
          prod.x += de * a.x; prod.y += de * a.y;
        }
        prod.x += ur[h].x / rK;        prod.y += ur[h].y / rK;
        prod.x += -kt * ut[h].y / rK;  prod.y += kt * ut[h].x / rK;
        prod.x += -kz * uz[h].y;       prod.y += kz * uz[h].x;
        dv[h] = prod;
      }
  }

  // ------------------------------------------------- mod_radialFD.cu:100-111
  void pressGrad(cplx* duz, const cplx* uz) const {
    int iE = NR - 1, left = iE - IW, right = NR - 1, i0 = iE + IW - right, len = right + 1 - left;
    double du = 0.0;
    for (int j = 0; j < len; j++) du += D[STEN * iE + j + i0] * uz[(size_t)(left + j) * MODES].x;
    for (int i = 0; i < NR; i++) duz[(size_t)i * MODES].x -= (2.0 / P.Re) * du;
  }

  // ------------------------------------------------- mod_radialFD.cu:115-143
  void addGrad(cplx* rr, cplx* rt, cplx* rz, const cplx* p) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NR; i++)
      for (int k = 0; k < MODES; k++) {
        size_t h = (size_t)i * MODES + k;
        int l = mode_l(k), it = mode_m(k), left = std::max(0, i - IW), right = std::min(i + IW, NR - 1);
        double kt = kt_of(it), kz = kz_of(l);
        int ip0 = std::abs((std::abs(it) % 2) - 0) * STEN * NR;
        double rK = r[i];
        int len = right + 1 - left, i0 = i + IW - right;
        cplx prod = {0.0, 0.0};
        for (int j = 0; j < len; j++) {
          double de = D[ip0 + STEN * i + j + i0];
          cplx a = p[(size_t)(left + j) * MODES + k];
          prod.x += de * a.x; prod.y += de * a.y;
        }
        rr[h].x -= prod.x;              rr[h].y -= prod.y;
        rt[h].x -= -kt * p[h].y / rK;   rt[h].y -= kt * p[h].x / rK;
        rz[h].x -= -kz * p[h].y;        rz[h].y -= kz * p[h].x;
      }
  }

  // ------------------------------------------------- mod_deriv.cu:10-35
  void derivt(cplx* g, const cplx* pp) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NR; i++)
      for (int k = 0; k < MODES; k++) {
        size_t h = (size_t)i * MODES + k;
        double kt = kt_of(mode_m(k));
        cplx a = pp[h];
        g[h] = {-kt * a.y / r[i], kt * a.x / r[i]};
      }
  }
  void derivz(cplx* g, const cplx* pp) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NR; i++)
      for (int k = 0; k < MODES; k++) {
        size_t h = (size_t)i * MODES + k;
        double kz = kz_of(mode_l(k));
        cplx a = pp[h];
        g[h] = {-kz * a.y, kz * a.x};
      }
  }

  // ------------------------------------------------- mod_utils.cu:20-57
  void write_bcon(cplx* v, cplx b) const {
    for (int k = 0; k < MODES; k++) v[(size_t)(NR - 1) * MODES + k] = b;
  }
  void decoupleForward(cplx* ur, cplx* ut) const {
#pragma omp parallel for schedule(static)
    for (size_t h = 0; h < N; h++) {
      cplx a = ur[h], b = ut[h];
      ur[h] = {a.x - b.y, a.y + b.x};
      ut[h] = {a.x + b.y, a.y - b.x};
    }
  }
  void decoupleBackward(cplx* up, cplx* um) const {
#pragma omp parallel for schedule(static)
    for (size_t h = 0; h < N; h++) {
      cplx a = up[h], b = um[h];
      up[h] = {0.5 * (a.x + b.x), 0.5 * (a.y + b.y)};
      um[h] = {0.5 * (a.y - b.y), -0.5 * (a.x - b.x)};
    }
  }

  // ------------------------------------------------- mod_linear.cu:12-51
  void addLap(cplx* rh, const cplx* uu, const double* LHS, int flag) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NR; i++)
      for (int k = 0; k < MODES; k++) {
        size_t h = (size_t)i * MODES + k;
        int fp = std::abs(flag - 3) / 2;
        int it = mode_m(k), l = mode_l(k);
        int ip0 = std::abs((std::abs(it) % 2) - fp) * STEN * NR;
        double kz = kz_of(l), kt = kt_of(it);
        int left = std::max(0, i - IW), right = std::min(i + IW, NR - 1);
        int len = right + 1 - left, i0 = i + IW - right;
        cplx prod = {0.0, 0.0};
        for (int j = 0; j < len; j++) {
          double de = LHS[ip0 + STEN * i + j + i0];
          cplx a = uu[(size_t)(left + j) * MODES + k];
          prod.x += de * a.x; prod.y += de * a.y;
        }
        cplx a = uu[h];
        double r2 = r[i] * r[i], de = -kz * kz - kt * kt / r2;
        if (flag == 0) de += -(1.0 / r2) - (2.0 * kt / r2);
        else if (flag == 1) de += -(1.0 / r2) + (2.0 * kt / r2);
        prod.x += de * a.x; prod.y += de * a.y;
        prod.x *= (1.0 - P.d_im) / P.Re; prod.y *= (1.0 - P.d_im) / P.Re;
        prod.x += (1.0 / P.dt) * a.x; prod.y += (1.0 / P.dt) * a.y;
        rh[h].x += prod.x; rh[h].y += prod.y;
      }
  }
  void addLaplacian(cplx* const* rh, cplx* const* uu) const {  // mod_linear.cu:652-661
    addLap(rh[0], uu[0], Lu.data(), 0);
    addLap(rh[1], uu[1], Lu.data(), 1);
    addLap(rh[2], uu[2], Lu.data(), 2);
  }

  // ----------------------------------- mod_linear.cu:55-200, 204-349, 353-498, 502-646
  // kind 0: u+  1: u-  2: uz  3: p.  Doolittle banded LU (no pivoting, half
  // band IW) fused with forward elimination, then back substitution.
  void lusolve_mode(int kind, int k, const cplx* RHS, cplx* sol, double* U, cplx* x) const {
    const double* LHS = (kind == 3) ? Lp.data() : Lu.data();
    int it = mode_m(k), l = mode_l(k);
    int fp = (kind <= 1) ? 1 : 0;
    int ip0 = std::abs((std::abs(it) % 2) - fp) * STEN * NR;
    double kz = kz_of(l), kt = kt_of(it);
    bool vel = kind < 3;
    bool p00 = (kind == 3) && (std::abs(l) < 1 && std::abs(it) < 1);
    double L[1 + IW];
    const double a = P.d_im / P.Re, idt = 1.0 / P.dt;
    for (int i = 0; i < NR; i++) {
      int left = std::max(0, i - IW), right = std::min(i + IW, NR - 1), coun = 0, coun2;
      int i0 = i + IW - right;
      double r2 = r[i] * r[i];
      bool last = (i == NR - 1);
      x[i] = RHS[(size_t)i * MODES + k];
      for (int j = 0; j < i - left; j++) {
        double XM = LHS[ip0 + j + i0 + STEN * i];
        if (!last) { if (vel) XM = -a * XM; }
        else if (p00) XM = 0.0;
        L[j] = XM; coun += 1;
        if (j > 0) for (int h = 0; h < j; h++) L[j] -= L[h] * U[(1 + IW) * (h + left) + j - h];
        L[j] = L[j] / U[(j + left) * (1 + IW)];
        x[i].x -= L[j] * x[j + left].x; x[i].y -= L[j] * x[j + left].y;
      }
      coun2 = coun;
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
          XM = 0.0; if (j == 0) XM = 1.0;
        }
        U[j + (1 + IW) * i] = XM;
        int hleft = std::max(0, coun - IW); coun += 1;
        if (hleft < i - left)
          for (int h = hleft; h < i - left; h++)
            U[j + (1 + IW) * i] -= L[h] * U[(1 + IW) * (h + left) + j + coun2 - h];
      }
    }
    int i = NR - 1;
    x[i].x = x[i].x / U[i * (IW + 1)]; x[i].y = x[i].y / U[i * (IW + 1)];
    sol[(size_t)i * MODES + k] = x[i];
    for (i = NR - 2; i > -1; i--) {
      int right = std::min(i + IW, NR - 1);
      for (int j = 1; j < right - i + 1; j++) {
        x[i].x -= U[i * (IW + 1) + j] * x[i + j].x;
        x[i].y -= U[i * (IW + 1) + j] * x[i + j].y;
      }
      x[i].x = x[i].x / U[i * (IW + 1)]; x[i].y = x[i].y / U[i * (IW + 1)];
      sol[(size_t)i * MODES + k] = x[i];
    }
  }
  void lusolve(int kind, cplx* sol, const cplx* RHS) const {
#pragma omp parallel
    {
      std::vector<double> U((1 + IW) * NR);
      std::vector<cplx> x(NR);
#pragma omp for schedule(static)
      for (int k = 0; k < MODES; k++) lusolve_mode(kind, k, RHS, sol, U.data(), x.data());
    }
  }

  // ------------------------------------------------- mod_boundary.cu:66-196
  // Closed-form inverse of [[1,0,0,a14],[0,1,0,a24],[0,0,i,a34],[a41,a42,a43,a44]].
  static void cmul(cplx a, cplx b, cplx& o) { o = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
  void calc_invM(const cplx* d1, const cplx* d2, const cplx* d3, const cplx* d4) {
    for (int k = 0; k < MODES; k++) {
      size_t h = (size_t)(NR - 1) * MODES + k;
      cplx a14 = uhP[0][h], a24 = uhP[1][h], a34 = uhP[2][h], a44 = d4[k];
      cplx a41 = d1[k], a42 = d2[k], a43 = d3[k], Det, aux;
      cplx IA[16];
      aux.x = a44.x - a24.x * a42.x + a24.y * a42.y - a14.x * a41.x + a14.y * a41.y;
      aux.y = a44.y - a24.x * a42.y - a24.y * a42.x - a14.x * a41.y - a14.y * a41.x;
      Det.x = -aux.y - a34.x * a43.x + a34.y * a43.y;
      Det.y = aux.x - a34.x * a43.y - a34.y * a43.x;
      aux.x = a44.x - a24.x * a42.x + a24.y * a42.y;
      aux.y = a44.y - a24.x * a42.y - a24.y * a42.x;
      IA[0].x = -aux.y - a34.x * a43.x + a34.y * a43.y;
      IA[0].y = aux.x - a34.x * a43.y - a34.y * a43.x;
      IA[1].y = a14.x * a42.x - a14.y * a42.y;
      IA[1].x = -a14.x * a42.y - a14.y * a42.x;
      IA[2].x = a14.x * a43.x - a14.y * a43.y;
      IA[2].y = a14.x * a43.y + a14.y * a43.x;
      IA[3].x = a14.y;  IA[3].y = -a14.x;
      IA[4].y = a24.x * a41.x - a24.y * a41.y;
      IA[4].x = -a24.x * a41.y - a24.y * a41.x;
      aux.x = a44.x - a14.x * a41.x + a14.y * a41.y;
      aux.y = a44.y - a14.x * a41.y - a14.y * a41.x;
      IA[5].x = -aux.y - a34.x * a43.x + a34.y * a43.y;
      IA[5].y = aux.x - a34.x * a43.y - a34.y * a43.x;
      IA[6].x = a24.x * a43.x - a24.y * a43.y;
      IA[6].y = a24.x * a43.y + a24.y * a43.x;
      IA[7].x = a24.y;  IA[7].y = -a24.x;
      IA[8].x = a34.x * a41.x - a34.y * a41.y;
      IA[8].y = a34.x * a41.y + a34.y * a41.x;
      IA[9].x = a34.x * a42.x - a34.y * a42.y;
      IA[9].y = a34.x * a42.y + a34.y * a42.x;
      IA[10].x = a44.x - a14.x * a41.x + a14.y * a41.y - a24.x * a42.x + a24.y * a42.y;
      IA[10].y = a44.y - a14.x * a41.y - a14.y * a41.x - a24.x * a42.y - a24.y * a42.x;
      IA[11].x = -a34.x; IA[11].y = -a34.y;
      IA[12].x = a41.y;  IA[12].y = -a41.x;
      IA[13].x = a42.y;  IA[13].y = -a42.x;
      IA[14].x = -a43.x; IA[14].y = -a43.y;
      IA[15].x = 0.0;    IA[15].y = 1.0;
      double D1 = Det.x * Det.x + Det.y * Det.y;
      Det.x = Det.x / D1; Det.y = -Det.y / D1;
      for (int q = 0; q < 16; q++) {
        double ax = IA[q].x * Det.x - IA[q].y * Det.y;
        IA[q].y = IA[q].x * Det.y + IA[q].y * Det.x; IA[q].x = ax;
        invM[q + (size_t)k * 16] = IA[q];
      }
    }
  }

  // ------------------------------------------------- mod_boundary.cu:205-255
  void setBCon() {
    for (int c = 0; c < 3; c++) { uh[c].assign(N, {0, 0}); uhP[c].assign(N, {0, 0}); }
    invM.assign(16 * (size_t)MODES, {0, 0}); coef.assign(4 * (size_t)MODES, {0, 0});
    div1.assign(MODES, {0, 0});
    std::vector<cplx> rh(N), aux(N), d2v(MODES), d3v(MODES), d4v(MODES);
    cplx uni = {0, 0};
    std::fill(rh.begin(), rh.end(), uni); write_bcon(rh.data(), {1.0, 0.0});
    lusolve(0, uh[0].data(), rh.data());
    lusolve(1, uh[1].data(), rh.data());
    std::fill(rh.begin(), rh.end(), uni); write_bcon(rh.data(), {0.0, 1.0});
    lusolve(2, uh[2].data(), rh.data());
    std::fill(rh.begin(), rh.end(), uni); write_bcon(rh.data(), {-1.0, 0.0});
    lusolve(3, aux.data(), rh.data());
    // DIV1
    uhP[0] = uh[0]; std::fill(uhP[1].begin(), uhP[1].end(), uni); std::fill(uhP[2].begin(), uhP[2].end(), uni);
    decoupleBackward(uhP[0].data(), uhP[1].data());
    divWall(div1.data(), uhP[0].data(), uhP[1].data(), uhP[2].data());
    // DIV2
    uhP[1] = uh[1]; std::fill(uhP[0].begin(), uhP[0].end(), uni); std::fill(uhP[2].begin(), uhP[2].end(), uni);
    decoupleBackward(uhP[0].data(), uhP[1].data());
    divWall(d2v.data(), uhP[0].data(), uhP[1].data(), uhP[2].data());
    // DIV3
    uhP[2] = uh[2]; std::fill(uhP[0].begin(), uhP[0].end(), uni); std::fill(uhP[1].begin(), uhP[1].end(), uni);
    divWall(d3v.data(), uhP[0].data(), uhP[1].data(), uhP[2].data());
    std::fill(uhP[2].begin(), uhP[2].end(), uni);
    // DIV4 and u_hP
    derivr(uhP[0].data(), aux.data(), 0); derivt(uhP[1].data(), aux.data()); derivz(uhP[2].data(), aux.data());
    divWall(d4v.data(), uhP[0].data(), uhP[1].data(), uhP[2].data());
    decoupleForward(uhP[0].data(), uhP[1].data());
    calc_invM(div1.data(), d2v.data(), d3v.data(), d4v.data());
  }

  // ------------------------------------------------- mod_boundary.cu:269-278, 11-62
  void corBCon(cplx* const* v) {
    decoupleBackward(v[0], v[1]); divWall(div1.data(), v[0], v[1], v[2]); decoupleForward(v[0], v[1]);
    for (int k = 0; k < MODES; k++) {  // compcoef_k
      cplx dv = {-div1[k].x, -div1[k].y};
      for (int j = 0; j < 4; j++) {
        cplx IM = invM[3 + 4 * j + 16 * (size_t)k];
        coef[j + 4 * (size_t)k] = {dv.x * IM.x - dv.y * IM.y, dv.x * IM.y + dv.y * IM.x};
      }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NR; i++)  // addBCon_k
      for (int k = 0; k < MODES; k++) {
        size_t h = (size_t)i * MODES + k;
        cplx a0 = coef[0 + 4 * (size_t)k], a1 = coef[1 + 4 * (size_t)k];
        cplx a2 = coef[2 + 4 * (size_t)k], a3 = coef[3 + 4 * (size_t)k];
        cplx aj[3] = {a0, a1, a2};
        for (int c = 0; c < 3; c++) {
          cplx hh = uh[c][h], hp = uhP[c][h], ax;
          ax.x = aj[c].x * hh.x - aj[c].y * hh.y + a3.x * hp.x - a3.y * hp.y;
          ax.y = aj[c].x * hh.y + aj[c].y * hh.x + a3.x * hp.y + a3.y * hp.x;
          v[c][h].x += ax.x; v[c][h].y += ax.y;
        }
      }
  }

  // ------------------------------------------------- mod_cublasflux.cu:56-80, 90-103
  // cublasDdot is stated as the sequential sum (its order is unpinned).
  double dot_rdr(const double* v) const {
    double s = 0.0;
    for (int i = 0; i < NR; i++) s += rdr[i] * v[i];
    return s;
  }
  void setCublasFlux() {
    std::vector<cplx> aux(N, {1.0, 0.0}), uaux(N);
    write_bcon(aux.data(), {0.0, 0.0});
    lusolve(2, uaux.data(), aux.data());
    uc.resize(NR);
    for (int i = 0; i < NR; i++) uc[i] = uaux[(size_t)i * MODES].x;
    Qc = dot_rdr(uc.data());
  }
  void adj_flux(cplx* uz) const {
    std::vector<double> ua(NR);
    for (int i = 0; i < NR; i++) ua[i] = uz[(size_t)i * MODES].x;
    double Qa = dot_rdr(ua.data());
    double corr = (P.qdes - Qa) / Qc;
    for (int i = 0; i < NR; i++) uz[(size_t)i * MODES].x += corr * uc[i];
  }

  // ------------------------------------------------- mod_integrate.cu:63-68
  void modeZZ(cplx* const* v) const {
    for (int i = 0; i < NR; i++) {
      size_t h = (size_t)i * MODES;
      v[0][h].x = 0.0; v[0][h].y = 0.0; v[1][h].y = 0.0; v[2][h].y = 0.0;
    }
  }

  // --------------------------------------------- FFT semantics, mod_fft.cu:21-49
  // Padded spectral level: [NTP][NZP] complex; physical level: [NTP][NZF] real.
  // Inverse (Z2D): theta-pass (complex, sign +) then z C2R; at l=0 and l=NZF/2
  // only the real part is used (numpy irfft convention).  Forward (D2Z): z R2C
  // then theta-pass (sign -); unnormalised (normalize applied by caller).
  void fft2_inverse(const cplx* spec, double* phys, std::vector<cplx>& tmp, std::vector<cplx>& col,
                    std::vector<cplx>& full, int lmax) const {
    // lmax: columns l >= lmax are zero on input
    tmp.assign((size_t)NTP * NZP, {0.0, 0.0});
    col.resize(std::max(NTP, NZF));
    for (int l = 0; l < std::min(lmax, NZP); l++) {
      fft_t.exec(spec + l, NZP, col.data(), +1);
      for (int j = 0; j < NTP; j++) tmp[(size_t)j * NZP + l] = col[j];
    }
    full.resize(NZF);
    for (int j = 0; j < NTP; j++) {
      const cplx* row = &tmp[(size_t)j * NZP];
      full[0] = {row[0].x, 0.0};
      for (int l = 1; l < NZF / 2; l++) { full[l] = row[l]; full[NZF - l] = {row[l].x, -row[l].y}; }
      full[NZF / 2] = {row[NZF / 2].x, 0.0};
      fft_z.exec(full.data(), 1, col.data(), +1);
      for (int n = 0; n < NZF; n++) phys[(size_t)j * NZF + n] = col[n].x;
    }
  }
  void fft2_forward(const double* phys, cplx* spec, std::vector<cplx>& tmp, std::vector<cplx>& col,
                    std::vector<cplx>& full, int lmax) const {
    // only columns l < lmax of the output are produced (others left zero)
    tmp.assign((size_t)NTP * NZP, {0.0, 0.0});
    full.resize(std::max(NZF, NTP)); col.resize(std::max(NTP, NZF));
    for (int j = 0; j < NTP; j++) {
      for (int n = 0; n < NZF; n++) full[n] = {phys[(size_t)j * NZF + n], 0.0};
      fft_z.exec(full.data(), 1, col.data(), -1);
      for (int l = 0; l < NZP; l++) tmp[(size_t)j * NZP + l] = col[l];
    }
    for (size_t q = 0; q < (size_t)NTP * NZP; q++) spec[q] = {0.0, 0.0};
    for (int l = 0; l < std::min(lmax, NZP); l++) {
      fft_t.exec(tmp.data() + l, NZP, col.data(), -1);
      for (int j = 0; j < NTP; j++) spec[(size_t)j * NZP + l] = col[j];
    }
  }

  // ------------------------------------------------- mod_io_nonlinear.cu:173-214
  void padForward_level(cplx* pad, const cplx* ul) const {  // ul: one level [NT][NZ]
    for (int i = 0; i < NTP; i++)
      for (int k = 0; k < NZP; k++) {
        cplx aux = {0.0, 0.0};
        if ((i < NT / 2 || i > NT - 1) && k < NZ) {
          int ip = i < NT / 2 ? i : i - NT / 2;
          aux = ul[ip * NZ + k];
        }
        pad[(size_t)i * NZP + k] = aux;
      }
  }
  void padBackward_level(cplx* ul, const cplx* pad) const {
    for (int i = 0; i < NT; i++)
      for (int k = 0; k < NZ; k++) {
        int ip = i < NT / 2 ? i : NTP + (i - NT);
        cplx aux = pad[(size_t)ip * NZP + k];
        if (i == NT / 2 || k == NZ - 1) aux = {0.0, 0.0};
        ul[i * NZ + k] = aux;
      }
  }

  // ------------------------------------------------- mod_io_nonlinear.cu:380-409, 142-170
  void nonlinear(cplx* const* uu, cplx* const* du) {
    derivr(dtmp[0].data(), uu[0], 1); derivr(dtmp[1].data(), uu[1], 1); derivr(dtmp[2].data(), uu[2], 0);
    const double norm = (double)NTP * (2 * NZP - 2);
#pragma omp parallel
    {
      std::vector<cplx> lvl(MODES), pad((size_t)NTP * NZP), tmp, col, full;
      std::vector<double> ph(12 * (size_t)NTP * NZF);
      const int nlev = sample_levels > 0 ? std::min(sample_levels, NR) : NR;
#pragma omp for schedule(dynamic, 1)
      for (int i = 0; i < nlev; i++) {
        double rad = r[i];
        // fields: 0-2 u, 3-5 dr u, 6-8 dt u, 9-11 dz u
        for (int f = 0; f < 12; f++) {
          int c = f % 3, g = f / 3;
          const cplx* src = (g == 0) ? uu[c] : dtmp[c].data();
          for (int k = 0; k < MODES; k++) {
            size_t h = (size_t)i * MODES + k;
            cplx a = src[h];
            if (g == 2) { double kt = kt_of(mode_m(k)); cplx b = uu[c][h]; a = {-kt * b.y / r[i], kt * b.x / r[i]}; }
            if (g == 3) { double kz = kz_of(mode_l(k)); cplx b = uu[c][h]; a = {-kz * b.y, kz * b.x}; }
            lvl[k] = a;
          }
          padForward_level(pad.data(), lvl.data());
          fft2_inverse(pad.data(), &ph[(size_t)f * NTP * NZF], tmp, col, full, NZ);
        }
        size_t P2 = (size_t)NTP * NZF;
        double *urK = &ph[0], *utK = &ph[P2], *uzK = &ph[2 * P2];
        double *drurK = &ph[3 * P2], *drutK = &ph[4 * P2], *druzK = &ph[5 * P2];
        double *dturK = &ph[6 * P2], *dtutK = &ph[7 * P2], *dtuzK = &ph[8 * P2];
        double *dzurK = &ph[9 * P2], *dzutK = &ph[10 * P2], *dzuzK = &ph[11 * P2];
        for (size_t h = 0; h < P2; h++) {
          double ur = urK[h], ut = utK[h], uz = uzK[h];
          double udu_r = ur * drurK[h] + ut * dturK[h] + uz * dzurK[h] - ut * (ut / rad);
          double udu_t = ur * drutK[h] + ut * dtutK[h] + uz * dzutK[h] + ur * (ut / rad);
          double udu_z = ur * druzK[h] + ut * dtuzK[h] + uz * dzuzK[h];
          drurK[h] = -udu_r; drutK[h] = -udu_t; druzK[h] = -udu_z;
        }
        for (int c = 0; c < 3; c++) {
          fft2_forward(&ph[(size_t)(3 + c) * P2], pad.data(), tmp, col, full, NZ - 1);
          for (size_t q = 0; q < (size_t)NTP * NZP; q++) { pad[q].x = pad[q].x / norm; pad[q].y = pad[q].y / norm; }
          padBackward_level(lvl.data(), pad.data());
          std::memcpy(du[c] + (size_t)i * MODES, lvl.data(), sizeof(cplx) * MODES);
        }
      }
    }
    pressGrad(du[2], uu[2]);
  }

  // ------------------------------------------------- mod_integrate.cu:164-228
  void addRHS() {  // mod_integrate.cu:43-59
    for (int c = 0; c < 3; c++) {
      cplx* a = rhs[c].data(); const cplx* b = rhsw[c].data();
#pragma omp parallel for schedule(static)
      for (size_t h = 0; h < N; h++) {
        a[h].x = P.d_im * a[h].x + (1.0 - P.d_im) * b[h].x;
        a[h].y = P.d_im * a[h].y + (1.0 - P.d_im) * b[h].y;
      }
    }
  }
  void solve_common(cplx* const* uu, cplx* const* out) {
    // steps 3-13 of predictor/corrector; uu (state) is used as scratch (u.t=div, u.z=p)
    cplx* R[3] = {rhs[0].data(), rhs[1].data(), rhs[2].data()};
    addLaplacian(R, uw_ptrs());
    decoupleBackward(R[0], R[1]);
    divergence(uu[1], R[0], R[1], R[2]);
    write_bcon(uu[1], {0.0, 0.0}); lusolve(3, uu[2], uu[1]);
    addGrad(R[0], R[1], R[2], uu[2]);
    decoupleForward(R[0], R[1]);
    write_bcon(R[0], {0.0, 0.0}); lusolve(0, out[0], R[0]);
    write_bcon(R[1], {0.0, 0.0}); lusolve(1, out[1], R[1]);
    write_bcon(R[2], {0.0, 0.0}); lusolve(2, out[2], R[2]);
    corBCon(out);
    decoupleBackward(out[0], out[1]);
    modeZZ(out);
    adj_flux(out[2]);
  }
  cplx* uwp[3];
  cplx* const* uw_ptrs() { for (int c = 0; c < 3; c++) uwp[c] = uw[c].data(); return uwp; }
  void predictor() {
    decoupleForward(u[0].data(), u[1].data()); decoupleForward(rhs[0].data(), rhs[1].data());
    for (int c = 0; c < 3; c++) { uw[c] = u[c]; rhsw[c] = rhs[c]; }
    cplx* U3[3] = {u[0].data(), u[1].data(), u[2].data()};
    cplx* O3[3] = {ui[0].data(), ui[1].data(), ui[2].data()};
    solve_common(U3, O3);
  }
  void corrector() {
    decoupleForward(rhs[0].data(), rhs[1].data());
    addRHS();
    cplx* U3[3] = {u[0].data(), u[1].data(), u[2].data()};
    solve_common(U3, U3);
  }
  // mod_integrate.cu:9-40 + mod_cublasflux.cu:106-116.  cublasIdamax returns
  // the first maximum of |x|; for the non-negative entries of d1/d2 this is
  // the maximum value.  A NaN anywhere is reported as err = NaN (which ends
  // the corrector loop, as a NaN err does in the reference).
  double measucorr() {
    double num = 0.0, den = 0.0; bool nan = false;
    for (size_t h = 0; h < N; h++) {
      double n1 = 0.0, n2 = 0.0;
      for (int c = 0; c < 3; c++) {
        cplx a = ui[c][h], b = u[c][h];
        double dx = a.x - b.x, dy = a.y - b.y;
        n1 = std::max(n1, dx * dx + dy * dy);
        n2 = std::max(n2, a.x * a.x + a.y * a.y);
        if (std::isnan(dx) || std::isnan(dy) || std::isnan(a.x) || std::isnan(a.y)) nan = true;
      }
      num = std::max(num, n1); den = std::max(den, n2);
    }
    if (nan) return std::nan("");
    return num / den;
  }
  // mod_integrate.cu:112-131 (one pass of the time-step loop body)
  int step(double* err_out) {
    cplx* U3[3] = {u[0].data(), u[1].data(), u[2].data()};
    cplx* R3[3] = {rhs[0].data(), rhs[1].data(), rhs[2].data()};
    cplx* I3[3] = {ui[0].data(), ui[1].data(), ui[2].data()};
    nonlinear(U3, R3);
    predictor();
    double err0 = 1e98;
    int ite = 0;
    while (ite < P.maxit && err0 > P.tol) {
      ite += 1;
      nonlinear(I3, R3);
      corrector();
      err0 = measucorr();
      for (int c = 0; c < 3; c++) ui[c] = u[c];
    }
    if (err_out) *err_out = err0;
    return ite;
  }

  // ------------------------------------------------- mod_io_nonlinear.cu:424-449, 9-37
  void reference_ic(double A_P) {
    for (int c = 0; c < 3; c++) std::fill(u[c].begin(), u[c].end(), cplx{0.0, 0.0});
    for (int i = 0; i < NR; i++) u[2][(size_t)i * MODES].x = 1.0 - r[i] * r[i];
    const double norm = (double)NTP * (2 * NZP - 2);
    std::vector<double> pr((size_t)NTP * NZF), pt((size_t)NTP * NZF);
    std::vector<cplx> pad((size_t)NTP * NZP), tmp, col, full, lvl(MODES);
    for (int i = 0; i < NR; i++) {
      double rr = r[i];
      for (int it = 0; it < NTP; it++)
        for (int l = 0; l < NZF; l++) {
          double th = (double(it) / double(NTP)) * LT, z = double(l) / double(2 * NZP - 2);
          double ux = std::sin(0.5 * PI2 * z); ux = ux * ux; ux = std::exp(-10.0 * ux);
          double ur = (1.0 - rr * rr) * (1.0 - rr * rr);
          double ut = ur - 4.0 * rr * rr * (1.0 - rr * rr);
          ur = ur * ux * std::sin(th); ut = ut * ux * std::cos(th);
          pr[(size_t)it * NZF + l] = A_P * ur; pt[(size_t)it * NZF + l] = A_P * ut;
        }
      for (int c = 0; c < 2; c++) {
        fft2_forward(c == 0 ? pr.data() : pt.data(), pad.data(), tmp, col, full, NZP);
        for (auto& q : pad) { q.x = q.x / norm; q.y = q.y / norm; }
        padBackward_level(lvl.data(), pad.data());
        std::memcpy(u[c].data() + (size_t)i * MODES, lvl.data(), sizeof(cplx) * MODES);
      }
    }
  }

  void init(const Params& p) {
    P = p; NR = p.NR; NT = p.NT; NZ = p.NZ;
    NTP = 3 * NT / 2; NZP = 3 * (NZ - 1) / 2 + 1; NZF = 2 * NZP - 2;
    MODES = NT * NZ; N = (size_t)NR * MODES;
    fft_t.init(NTP); fft_z.init(NZF);
    init_fd(); get_dr0();
    for (int c = 0; c < 3; c++) {
      u[c].assign(N, {0, 0}); rhs[c].assign(N, {0, 0}); uw[c].assign(N, {0, 0});
      rhsw[c].assign(N, {0, 0}); ui[c].assign(N, {0, 0}); dtmp[c].assign(N, {0, 0});
    }
    setCublasFlux();
    setBCon();
  }
};

}  // namespace

// ============================================================================
// C API (ctypes).  Complex fields are interleaved (re,im) doubles in the
// reference layout double2[NR][NT][NZ], h = i*NT*NZ + it*NZ + l.
// ============================================================================
// ---------------------------------------------------------------------------
// Remesh tool (tools/Re_Mesh/src/mod_remesh.cu).  Pinning (tests/test_oracle.py):
// exact interpolation of even/odd polynomials of degree <= 5 on the rows the
// reference's column assignment leaves intact (every row with the mirror
// assignment), identity on an unchanged grid, spectral pad/truncate round trip.
// ---------------------------------------------------------------------------
namespace remesh {

// mod_remesh.cu:61-66 and 72-79 (the same grid as mod_radialFD.cu:211-216)
std::vector<double> grid(int nr) {
  std::vector<double> r(nr);
  int N_p = nr + int(std::sqrt(double(nr)));
  for (int i = N_p - nr; i < N_p; i++) {
    int ir = N_p - i - 1;
    r[i - N_p + nr] = 0.5 * (1.0 + std::cos(0.5 * PI2 * (double(ir) / double(N_p))));
  }
  for (int it = 0; it < 10; it++) {
    double dr = 1.5 * r[0] - 0.5 * r[1];
    for (int j = 0; j < nr; j++) r[j] = r[j] * (1.0 + dr) - dr;
  }
  return r;
}

// mod_remesh.cu:148-178 with m = 0 (the derivative loops are empty)
void fd_weights(double* c, const double* x, double x0, int n) {
  int m = 0;
  double c1, c2, c3, c4, c5;
  c[0] = 1.0; c1 = 1.0; c4 = x[0] - x0;
  for (int i = 1; i < n; i++) {
    int mn = std::min(i, m);
    (void)mn;
    c2 = 1.0; c5 = c4; c4 = x[i] - x0;
    for (int j = 0; j < i; j++) {
      c3 = x[i] - x[j]; c2 = c2 * c3;
      if (j == i - 1) c[i] = -c1 * c5 * c[i - 1] / c2;
      c[j] = c4 * c[j] / c3;
    }
    c1 = c2;
  }
}

// mod_remesh.cu:80-111; mirror = 1 puts a reflected point rau[p] in column
// IW-1-p (the reference uses IW-1-j for stencil position j)
std::vector<double> matrix(int NR, int Nrd, int mirror) {
  std::vector<double> r_o = grid(NR), r_h = grid(Nrd);
  std::vector<double> rau(NR + IW);
  for (int i = 0; i < NR + IW; i++) rau[i] = (i < IW) ? -r_o[IW - i - 1] : r_o[i - IW];
  std::vector<double> RM(2 * (size_t)NR * Nrd, 0.0);
  RM[(size_t)NR * Nrd - 1] = 1.0; RM[2 * (size_t)NR * Nrd - 1] = 1.0;
  for (int i = 0; i < Nrd - 1; i++) {
    double x0 = r_h[i];
    int jleft = 0;
    for (int j = 0; j < NR + IW; j++) if (rau[j] >= x0) { jleft = j; break; }
    jleft = jleft - IW;
    int jright = std::min(NR + IW - 1, jleft + 2 * IW - 1), n = jright - jleft + 1;
    std::vector<double> x(n), c(n, 0.0);
    for (int j = 0; j < n; j++) x[j] = rau[jleft + j];
    fd_weights(c.data(), x.data(), x0, n);
    int jrau = jleft;
    jleft = jleft - IW;
    int jaux = jleft;
    for (int j = 0; j < n; j++) {
      if (jaux < 0) {
        int jcol = mirror ? IW - 1 - (jrau + j) : IW - j - 1;
        jaux += 1;
        RM[(size_t)i * NR + jcol] += c[j]; RM[(size_t)NR * Nrd + (size_t)i * NR + jcol] -= c[j];
      } else {
        int jcol = jleft + j;
        RM[(size_t)i * NR + jcol] += c[j]; RM[(size_t)NR * Nrd + (size_t)i * NR + jcol] += c[j];
      }
    }
  }
  return RM;
}

}  // namespace remesh

extern "C" {

// Remesh matrix [2][nr_new][nr_old] (half 0: ur/ut of odd m, uz of even m).
void nspo_remesh_weights(int nr_old, int nr_new, int mirror, double* RM) {
  std::vector<double> M = remesh::matrix(nr_old, nr_new, mirror);
  std::copy(M.begin(), M.end(), RM);
}

// re_mesh (mod_remesh.cu:50-124): uo is [3][NR][NT][NZ] complex, un receives
// [3][Nrd][Ntd][Nzd] (fully overwritten; modes with no source are zero).
// rmsh_k (:8-30) scatters every old mode inside the new spectral range with
// the dense radial sum; modeZZ_k (:35-41) then clears ur(0,0) and the
// imaginary parts of ut(0,0), uz(0,0).
void nspo_remesh(int NR, int NT, int NZ, const double* uo_, int Nrd, int Ntd, int Nzd, double* un_,
                 int mirror) {
  const cplx* uo = (const cplx*)uo_;
  cplx* un = (cplx*)un_;
  std::vector<double> RM = remesh::matrix(NR, Nrd, mirror);
  const size_t no = (size_t)NR * NT * NZ, nn = (size_t)Nrd * Ntd * Nzd;
  for (size_t q = 0; q < 3 * nn; q++) un[q] = {0.0, 0.0};
  for (int c = 0; c < 3; c++) {
    int flag = c < 2 ? 1 : 0;
    for (int k_o = 0; k_o < NT * NZ; k_o++)
      for (int h_or = 0; h_or < Nrd; h_or++) {
        int h_oz = k_o % NZ, it = k_o / NZ, h_ot = it;
        h_ot = h_ot < NT / 2 ? h_ot : h_ot - NT;
        if (h_ot >= Ntd / 2 || h_ot < -Ntd / 2 || h_oz >= Nzd) continue;
        size_t ip0 = (size_t)std::abs((std::abs(h_ot) % 2) - flag) * NR * Nrd;
        if (h_ot < 0) h_ot = h_ot + Ntd;
        size_t h = (size_t)h_or * Ntd * Nzd + (size_t)h_ot * Nzd + h_oz;
        cplx du = {0.0, 0.0};
        for (int j = 0; j < NR; j++) {
          double rm = RM[(size_t)h_or * NR + j + ip0];
          const cplx a = uo[c * no + (size_t)j * NZ * NT + k_o];
          du.x += rm * a.x; du.y += rm * a.y;
        }
        un[c * nn + h] = du;
      }
  }
  for (int i = 0; i < Nrd; i++) {
    size_t h = (size_t)i * Ntd * Nzd;
    un[h] = {0.0, 0.0}; un[nn + h].y = 0.0; un[2 * nn + h].y = 0.0;
  }
}


void* nspo_create(int nr, int nt, int nz, double lz, double re, double dt, double d_im, int maxit,
                  double tol, double qdes) {
  if (nr < 2 * IW + 2 || nt < 2 || (nt % 2) || nz < 3 || !(nz % 2)) return nullptr;
  Oracle* o = new Oracle();
  Params p{nr, nt, nz, lz, re, dt, d_im, maxit, tol, qdes};
  o->init(p);
  return o;
}
void nspo_destroy(void* h) { delete (Oracle*)h; }

void nspo_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#endif
  (void)n;
}
// Timing samples only: nonlinear() evaluates levels 0..n-1 (n <= 0: all).
void nspo_set_sample_levels(void* h, int n) { ((Oracle*)h)->sample_levels = n; }

int nspo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// r[NR], D[2*7*NR], Lu[2*7*NR], Lp[2*7*NR], rdr[NR], dr0[4], dr1[7]
void nspo_tables(void* h, double* r, double* D, double* Lu, double* Lp, double* rdr, double* dr0,
                 double* dr1) {
  Oracle* o = (Oracle*)h;
  std::memcpy(r, o->r.data(), sizeof(double) * o->NR);
  std::memcpy(D, o->D.data(), sizeof(double) * o->D.size());
  std::memcpy(Lu, o->Lu.data(), sizeof(double) * o->Lu.size());
  std::memcpy(Lp, o->Lp.data(), sizeof(double) * o->Lp.size());
  std::memcpy(rdr, o->rdr.data(), sizeof(double) * o->NR);
  std::memcpy(dr0, o->dr0.data(), sizeof(double) * o->dr0.size());
  std::memcpy(dr1, o->dr1.data(), sizeof(double) * o->dr1.size());
}

void nspo_set_state(void* h, const double* ur, const double* ut, const double* uz) {
  Oracle* o = (Oracle*)h;
  const double* s[3] = {ur, ut, uz};
  for (int c = 0; c < 3; c++) std::memcpy(o->u[c].data(), s[c], sizeof(cplx) * o->N);
}
void nspo_get_state(void* h, double* ur, double* ut, double* uz) {
  Oracle* o = (Oracle*)h;
  double* s[3] = {ur, ut, uz};
  for (int c = 0; c < 3; c++) std::memcpy(s[c], o->u[c].data(), sizeof(cplx) * o->N);
}

// Reference initial condition (restart == 0): Poiseuille + A_P perturbation.
void nspo_reference_ic(void* h, double a_p) { ((Oracle*)h)->reference_ic(a_p); }

// du = N(u) for the three given fields (r/theta form), nonlinear() of the reference.
void nspo_nonlinear(void* h, const double* ur, const double* ut, const double* uz, double* dr,
                    double* dt, double* dz) {
  Oracle* o = (Oracle*)h;
  std::vector<cplx> a[3], b[3];
  const double* s[3] = {ur, ut, uz};
  for (int c = 0; c < 3; c++) {
    a[c].resize(o->N); b[c].resize(o->N);
    std::memcpy(a[c].data(), s[c], sizeof(cplx) * o->N);
  }
  cplx* A[3] = {a[0].data(), a[1].data(), a[2].data()};
  cplx* B[3] = {b[0].data(), b[1].data(), b[2].data()};
  o->nonlinear(A, B);
  double* d[3] = {dr, dt, dz};
  for (int c = 0; c < 3; c++) std::memcpy(d[c], b[c].data(), sizeof(cplx) * o->N);
}

// Single operators for unit tests.
void nspo_lusolve(void* h, int kind, const double* rhs, double* sol) {
  Oracle* o = (Oracle*)h;
  o->lusolve(kind, (cplx*)sol, (const cplx*)rhs);
}
void nspo_derivr(void* h, const double* u, double* du, int flag) {
  Oracle* o = (Oracle*)h;
  o->derivr((cplx*)du, (const cplx*)u, flag);
}
void nspo_divergence(void* h, const double* ur, const double* ut, const double* uz, double* dv) {
  Oracle* o = (Oracle*)h;
  o->divergence((cplx*)dv, (const cplx*)ur, (const cplx*)ut, (const cplx*)uz);
}
void nspo_divwall(void* h, const double* ur, const double* ut, const double* uz, double* dv) {
  Oracle* o = (Oracle*)h;
  o->divWall((cplx*)dv, (const cplx*)ur, (const cplx*)ut, (const cplx*)uz);
}
void nspo_addlap(void* h, double* rhs, const double* u, int flag) {
  Oracle* o = (Oracle*)h;
  o->addLap((cplx*)rhs, (const cplx*)u, o->Lu.data(), flag);
}
void nspo_addgrad(void* h, double* rr, double* rt, double* rz, const double* p) {
  Oracle* o = (Oracle*)h;
  o->addGrad((cplx*)rr, (cplx*)rt, (cplx*)rz, (const cplx*)p);
}
// corBCon on a field given in +/- form (modified in place).
void nspo_corbcon(void* h, double* up, double* um, double* uz) {
  Oracle* o = (Oracle*)h;
  cplx* v[3] = {(cplx*)up, (cplx*)um, (cplx*)uz};
  o->corBCon(v);
}
void nspo_adj_flux(void* h, double* uz) { ((Oracle*)h)->adj_flux((cplx*)uz); }
// invM column 3 (4 complex per mode) and the homogeneous fields.
void nspo_influence(void* h, double* invm_col3, double* uh_p, double* uh_m, double* uh_z,
                    double* uhp_p, double* uhp_m, double* uhp_z) {
  Oracle* o = (Oracle*)h;
  for (int k = 0; k < o->MODES; k++)
    for (int j = 0; j < 4; j++) ((cplx*)invm_col3)[4 * (size_t)k + j] = o->invM[3 + 4 * j + 16 * (size_t)k];
  double* d[6] = {uh_p, uh_m, uh_z, uhp_p, uhp_m, uhp_z};
  for (int c = 0; c < 3; c++) {
    if (d[c]) std::memcpy(d[c], o->uh[c].data(), sizeof(cplx) * o->N);
    if (d[c + 3]) std::memcpy(d[c + 3], o->uhP[c].data(), sizeof(cplx) * o->N);
  }
}
double nspo_flux_qc(void* h, double* uc) {
  Oracle* o = (Oracle*)h;
  if (uc) std::memcpy(uc, o->uc.data(), sizeof(double) * o->NR);
  return o->Qc;
}
// One padded level: spec [NTP][NZP] complex -> phys [NTP][3(NZ-1)] real and back.
void nspo_fft2_inverse(void* h, const double* spec, double* phys) {
  Oracle* o = (Oracle*)h;
  std::vector<cplx> tmp, col, full;
  o->fft2_inverse((const cplx*)spec, phys, tmp, col, full, o->NZP);
}
void nspo_fft2_forward(void* h, const double* phys, double* spec) {
  Oracle* o = (Oracle*)h;
  std::vector<cplx> tmp, col, full;
  o->fft2_forward(phys, (cplx*)spec, tmp, col, full, o->NZP);
}

// Advance nsteps time steps (integrate() loop body).  n_it[s] and err[s]
// receive the corrector count and final err of each step (may be NULL).
void nspo_step(void* h, int nsteps, int* n_it, double* err) {
  Oracle* o = (Oracle*)h;
  for (int s = 0; s < nsteps; s++) {
    double e;
    int it = o->step(&e);
    if (n_it) n_it[s] = it;
    if (err) err[s] = e;
  }
}

// Physical-space component comp (0 ur, 1 ut, 2 uz) of the current state:
// padForward + fftBackward (padForward_k, mod_io_nonlinear.cu:173-192; fftBackward, mod_fft.cu:45-49)
// per radial level, as the reference leaves it in u_pad: out[NR][NTP][2*NZP]
// doubles, NZF samples per row then the two padding doubles (zero here).
void nspo_physical(void* h, int comp, double* out) {
  Oracle* o = (Oracle*)h;
  const int NR = o->NR, NTP = o->NTP, NZP = o->NZP, NZF = o->NZF, MODES = o->MODES;
#pragma omp parallel
  {
    std::vector<cplx> pad((size_t)NTP * NZP), tmp, col, full;
    std::vector<double> ph((size_t)NTP * NZF);
#pragma omp for
    for (int i = 0; i < NR; i++) {
      o->padForward_level(pad.data(), o->u[comp].data() + (size_t)i * MODES);
      o->fft2_inverse(pad.data(), ph.data(), tmp, col, full, o->NZ);
      for (int j = 0; j < NTP; j++) {
        double* row = out + ((size_t)i * NTP + j) * (2 * NZP);
        for (int n = 0; n < NZF; n++) row[n] = ph[(size_t)j * NZF + n];
        row[NZF] = 0.0; row[NZF + 1] = 0.0;
      }
    }
  }
}

// Cross-section integrals of wrt_qcr/wrt_ucr (mod_io_nonlinear.cu:576-654):
// q0kernel_k (:55-66) ur^2+ut^2 or u0kernel_k (:68-79) uz^2 per physical
// point, q1kernel_k (:81-103) the periodic three-point rule
// Int_j = 0.5/(3 NTP) (f_{j-1} + 4 f_j + f_{j+1}), then intq
// (mod_cublasflux.cu:119-130): the sum over theta with weights qaux = 1 and the
// sum over r with rdr.  q[n], u2[n] for the NZF axial points (either NULL).
void nspo_cross(void* h, double* q, double* u2) {
  Oracle* o = (Oracle*)h;
  const int NR = o->NR, NTP = o->NTP, NZP = o->NZP, NZF = o->NZF;
  const size_t P = (size_t)NR * NTP * 2 * NZP;
  std::vector<double> a(P), b(P), f(P);
  for (int which = 0; which < 2; which++) {
    double* dst = which == 0 ? q : u2;
    if (!dst) continue;
    if (which == 0) {
      nspo_physical(h, 0, a.data());
      nspo_physical(h, 1, b.data());
      for (size_t k = 0; k < P; k++) f[k] = a[k] * a[k] + b[k] * b[k];
    } else {
      nspo_physical(h, 2, a.data());
      for (size_t k = 0; k < P; k++) f[k] = a[k] * a[k];
    }
    std::vector<double> q1((size_t)NR * 2 * NZP, 0.0);
    for (int i = 0; i < NR; i++)
      for (int l = 0; l < 2 * NZP; l++) {
        double acc = 0.0;
        for (int it = 0; it < NTP; it++) {
          int ia = it < 1 ? NTP - 1 : it - 1, ib = it > NTP - 2 ? 0 : it + 1;
          size_t ih = (size_t)i * NTP * 2 * NZP;
          double fa = f[ih + (size_t)ia * 2 * NZP + l], fab = f[ih + (size_t)it * 2 * NZP + l],
                 fb = f[ih + (size_t)ib * 2 * NZP + l];
          double Int = 0.5 * (1.0 / (3.0 * double(NTP))) * (fa + 4.0 * fab + fb);
          acc += 1.0 * Int;
        }
        q1[(size_t)i * 2 * NZP + l] = acc;
      }
    for (int l = 0; l < NZF; l++) {
      double acc = 0.0;
      for (int i = 0; i < NR; i++) acc += q1[(size_t)i * 2 * NZP + l] * o->rdr[i];
      dst[l] = acc;
    }
  }
}

// Diagnostics of the current state: out[0]=u_cl, out[1]=u_tau,
// out[2]=kinetic energy, out[3]=L2 divergence (weights |rdr|), out[4]=max |u| at the wall,
// out[5]=max |wall divergence|, out[6]=bulk flux.
void nspo_diagnostics(void* h, double* out) {
  Oracle* o = (Oracle*)h;
  int NR = o->NR, MODES = o->MODES;
  const cplx* uz = o->u[2].data();
  double uc = 0.0;  // mod_io_nonlinear.cu:45
  for (int i = 0; i < IW + 1; i++) uc += uz[(size_t)i * MODES].x * o->dr0[i];
  double du = 0.0;  // mod_io_nonlinear.cu:47-49
  for (int i = 0; i < STEN; i++) du += uz[(size_t)(NR - STEN + i) * MODES].x * o->dr1[i];
  du = std::sqrt(std::fabs(du) / o->P.Re);
  double ke = 0.0;
  for (int i = 0; i < NR; i++)
    for (int k = 0; k < MODES; k++) {
      double w = (o->mode_l(k) == 0) ? 1.0 : 2.0;
      size_t hh = (size_t)i * MODES + k;
      for (int c = 0; c < 3; c++) {
        cplx a = o->u[c][hh];
        ke += 0.5 * w * o->rdr[i] * (a.x * a.x + a.y * a.y);
      }
    }
  std::vector<cplx> dv(o->N), dw(MODES);
  o->divergence(dv.data(), o->u[0].data(), o->u[1].data(), o->u[2].data());
  double d2 = 0.0;
  for (int i = 0; i < NR; i++)
    for (int k = 0; k < MODES; k++) {
      double w = (o->mode_l(k) == 0) ? 1.0 : 2.0;
      cplx a = dv[(size_t)i * MODES + k];
      d2 += w * std::fabs(o->rdr[i]) * (a.x * a.x + a.y * a.y);
    }
  o->divWall(dw.data(), o->u[0].data(), o->u[1].data(), o->u[2].data());
  double wall = 0.0, wdiv = 0.0;
  for (int k = 0; k < MODES; k++) {
    for (int c = 0; c < 3; c++) {
      cplx a = o->u[c][(size_t)(NR - 1) * MODES + k];
      wall = std::max(wall, std::sqrt(a.x * a.x + a.y * a.y));
    }
    wdiv = std::max(wdiv, std::sqrt(dw[k].x * dw[k].x + dw[k].y * dw[k].y));
  }
  std::vector<double> ua(NR);
  for (int i = 0; i < NR; i++) ua[i] = uz[(size_t)i * MODES].x;
  out[0] = uc; out[1] = du; out[2] = ke; out[3] = std::sqrt(d2); out[4] = wall; out[5] = wdiv;
  out[6] = o->dot_rdr(ua.data());
}

}  // extern "C"
