// Host setup: radial grid, finite-difference tables, quadrature weights and
// FFT plans.  Runs once per pipe in FP64 on the host.
#include <algorithm>
#include <cmath>

#include "nsp_internal.h"

namespace nsp {

// Fornberg weights of derivatives 0..m at rau[i0+IW] from the points
// rau[i0 .. i0+right-left] (mod_radialFD.cu:365-395).  c[j + n*k] holds the
// weight of point j for derivative k, n = number of points.
static void fornberg(double* c, const double* rau, int i0, int left, int right, int m) {
  const int n = 1 + right - left;
  const double x0 = rau[i0 + IW];
  double x[STEN];
  for (int q = 0; q < STEN * STEN; q++) c[q] = 0.0;
  for (int q = 0; q < n; q++) x[q] = rau[q + i0];
  c[0] = 1.0;
  double c1 = 1.0, c4 = x[0] - x0;
  for (int i = 1; i < n; i++) {
    const int mn = std::min(i, m);
    double c2 = 1.0;
    const double c5 = c4;
    c4 = x[i] - x0;
    for (int j = 0; j < i; j++) {
      const double c3 = x[i] - x[j];
      c2 = c2 * c3;
      if (j == i - 1) {
        for (int k = mn; k > 0; k--)
          c[i + n * k] = c1 * (double(k) * c[i - 1 + n * (k - 1)] - c5 * c[i - 1 + n * k]) / c2;
        c[i] = -c1 * c5 * c[i - 1] / c2;
      }
      for (int k = mn; k > 0; k--) c[j + n * k] = (c4 * c[j + n * k] - double(k) * c[j + (k - 1) * n]) / c3;
      c[j] = c4 * c[j] / c3;
    }
    c1 = c2;
  }
}

// Cosine grid clustered to the wall, then stretched ten times so the axis sits
// midway to its mirror point (mod_radialFD.cu:211-216; the remesh tool builds
// the same grid, tools/Re_Mesh/src/mod_remesh.cu:61-79).
void radial_grid(int nr, double* r) {
  const int np = nr + int(std::sqrt(double(nr)));
  for (int i = np - nr; i < np; i++) {
    const int ir = np - i - 1;
    r[i - np + nr] = 0.5 * (1.0 + std::cos(0.5 * PI2 * (double(ir) / double(np))));
  }
  for (int it = 0; it < 10; it++) {
    const double dr = 1.5 * r[0] - 0.5 * r[1];
    for (int j = 0; j < nr; j++) r[j] = r[j] * (1.0 + dr) - dr;
  }
}

// Interpolation weights at x0 from the points x[0..n-1]: the derivative-free
// case of Fornberg's recurrence (tools/Re_Mesh/src/mod_remesh.cu:148-178).
static void lagrange_weights(double* c, const double* x, double x0, int n) {
  c[0] = 1.0;
  double c1 = 1.0, c4 = x[0] - x0;
  for (int i = 1; i < n; i++) {
    double c2 = 1.0;
    const double c5 = c4;
    c4 = x[i] - x0;
    for (int j = 0; j < i; j++) {
      const double c3 = x[i] - x[j];
      c2 = c2 * c3;
      if (j == i - 1) c[i] = -c1 * c5 * c[i - 1] / c2;
      c[j] = c4 * c[j] / c3;
    }
    c1 = c2;
  }
}

// Radial remesh matrix [2][nrn][nro] (tools/Re_Mesh/src/mod_remesh.cu:80-111):
// row i interpolates the old grid at rn[i] from 2*IW points of the old grid
// extended through the axis (rau[p] = -ro[IW-1-p] for p < IW).  Half 0 adds a
// reflected point's weight to its column (functions even in r), half 1
// subtracts it (odd).  The wall row copies the old wall value.  In the
// reference's column assignment a reflected point at stencil position j goes
// to column IW-1-j, which is its own column only when the stencil starts at
// rau[0]; mirror = true assigns column IW-1-p for the point rau[p] instead.
std::vector<double> remesh_matrix(const std::vector<double>& ro, const std::vector<double>& rn, bool mirror) {
  const int nro = (int)ro.size(), nrn = (int)rn.size();
  std::vector<double> RM(2 * (size_t)nro * nrn, 0.0);
  RM[(size_t)nro * nrn - 1] = 1.0;
  RM[2 * (size_t)nro * nrn - 1] = 1.0;
  std::vector<double> rau(nro + IW);
  for (int p = 0; p < nro + IW; p++) rau[p] = (p < IW) ? -ro[IW - p - 1] : ro[p - IW];
  double x[2 * IW], c[2 * IW];
  for (int i = 0; i < nrn - 1; i++) {
    const double x0 = rn[i];
    int first = nro + IW - 1;
    for (int p = 0; p < nro + IW; p++)
      if (rau[p] >= x0) {
        first = p;
        break;
      }
    const int a = first - IW;                        // first stencil point (extended index)
    const int b = std::min(nro + IW - 1, a + 2 * IW - 1);
    const int n = b - a + 1;
    for (int j = 0; j < n; j++) x[j] = rau[a + j];
    lagrange_weights(c, x, x0, n);
    const int s = a - IW;                            // its old-grid index (negative: reflected)
    for (int j = 0; j < n; j++) {
      double* row0 = &RM[(size_t)i * nro];
      double* row1 = &RM[(size_t)nro * nrn + (size_t)i * nro];
      if (s + j < 0) {
        const int col = mirror ? IW - 1 - (a + j) : IW - 1 - j;
        row0[col] += c[j];
        row1[col] -= c[j];
      } else {
        row0[s + j] += c[j];
        row1[s + j] += c[j];
      }
    }
  }
  return RM;
}

// Radial grid, parity-folded derivative tables, quadrature weights and the
// Laplacians of the velocity (Dirichlet wall row) and pressure (Neumann wall
// row): mod_radialFD.cu:199-312.  Table index: [parity][row][7], parity 0
// for even functions of r (ghost points folded with +), 1 for odd (-).
// The reference's quadrature loop runs one derivative order past the weights
// it computed (mod_radialFD.cu:272-277 reads c[ir+len*len], beyond its 49-double
// buffer); that term is of size dr^8/8! and is omitted here.
void build_radial_tables(int nr, RadialTables& T) {
  const int hD = STEN * nr;
  T.r.assign(nr, 0.0);
  T.D.assign(2 * hD, 0.0);
  T.Lu.assign(2 * hD, 0.0);
  T.Lp.assign(2 * hD, 0.0);
  T.rdr.assign(nr, 0.0);
  std::vector<double> D2(2 * hD, 0.0), rau(nr + IW), ri(nr + 1);
  double c[STEN * STEN];
  double* r = T.r.data();

  // Cosine grid clustered to the wall, then stretched so the axis sits
  // midway to its mirror point (mod_radialFD.cu:211-216).
  radial_grid(nr, r);
  ri[0] = 0.0;
  for (int i = 0; i < nr; i++) ri[i + 1] = r[i];
  for (int i = 0; i < nr + IW; i++) rau[i] = (i < IW) ? -r[IW - i - 1] : r[i - IW];

  // First and second derivative rows (mod_radialFD.cu:220-252).
  int ghosts = IW;
  for (int i = 0; i < nr; i++) {
    const int left = i, right = std::min(2 * IW + i, nr + IW - 1), len = 1 + right - left;
    fornberg(c, rau.data(), i, left, right, 2);
    if (i < IW) {
      for (int j = 0; j < len; j++) {
        if (j < ghosts) {  // ghost point -r_q folded onto r_q
          const int col = ghosts - j - 1;
          T.D[col + STEN * i] += c[len + j];
          T.D[hD + col + STEN * i] -= c[len + j];
          D2[col + STEN * i] += c[2 * len + j];
          D2[hD + col + STEN * i] -= c[2 * len + j];
        } else {
          const int col = j - ghosts;
          T.D[col + STEN * i] += c[len + j];
          T.D[hD + col + STEN * i] += c[len + j];
          D2[col + STEN * i] += c[2 * len + j];
          D2[hD + col + STEN * i] += c[2 * len + j];
        }
      }
      ghosts -= 1;
    } else {
      const int off = STEN - len;  // short near-wall rows are right-aligned
      for (int j = 0; j < len; j++) {
        const int col = off + j;
        T.D[col + STEN * i] += c[len + j];
        T.D[hD + col + STEN * i] += c[len + j];
        D2[col + STEN * i] += c[2 * len + j];
        D2[hD + col + STEN * i] += c[2 * len + j];
      }
    }
  }

  // Quadrature weights: rdr_j such that sum_j rdr_j f_j ~ 2 int_0^1 f r dr,
  // from Taylor expansions about each r_i over [r_{i-1}, r_{i+1}] with
  // r_{-1} = 0 and r_{NR} = r_{NR-1} (mod_radialFD.cu:253-280).
  ghosts = IW;
  for (int i = 0; i < nr; i++) {
    const int left = i, right = std::min(2 * IW + i, nr + IW - 1), len = 1 + right - left;
    const int left2 = std::max(0, i - IW);
    fornberg(c, rau.data(), i, left, right, len - 1);
    int off = 0;
    double e1 = 1.0, e2 = 1.0;
    if (i < IW) {
      off = ghosts;
      for (int j = 0; j < ghosts; j++) {
        const int col = 2 * ghosts - 1 - j;
        for (int k = 0; k < len; k++) c[col + k * len] += c[j + k * len];
      }
      ghosts -= 1;
    }
    for (int k = 0; k < len; k++) {
      e1 = e1 * (ri[std::min(i + 2, nr)] - ri[i + 1]) / (double(k + 1));
      e2 = e2 * (ri[i] - ri[i + 1]) / (double(k + 1));
      for (int j = 0; j < len - off; j++) T.rdr[left2 + j] += (e1 - e2) * r[i] * c[off + j + k * len];
    }
  }

  // Radial Laplacians (mod_radialFD.cu:281-302).
  for (int i = 0; i < nr; i++) {
    for (int j = 0; j < STEN; j++) {
      const int q = j + STEN * i;
      if (i < nr - 1) {
        T.Lu[q] = D2[q] + T.D[q] / r[i];
        T.Lu[hD + q] = D2[hD + q] + T.D[hD + q] / r[i];
        T.Lp[q] = D2[q] + T.D[q] / r[i];
        T.Lp[hD + q] = D2[hD + q] + T.D[hD + q] / r[i];
      } else {
        T.Lu[q] = (j < STEN - 1) ? 0.0 : 1.0;
        T.Lu[hD + q] = T.Lu[q];
        T.Lp[q] = T.D[q];
        T.Lp[hD + q] = T.D[hD + q];
      }
    }
  }

  // Interpolation weights to r = 0 from r[0..IW] (get_dr0, mod_radialFD.cu:316-350):
  // Lagrange weights l_j(0) = prod_{q != j} (0 - r_q)/(r_j - r_q), in the
  // reference's Fornberg recurrence order.
  {
    const int n = IW + 1;
    double cc[IW + 1], x[IW + 1];
    for (int q = 0; q < n; q++) { cc[q] = 0.0; x[q] = r[q]; }
    cc[0] = 1.0;
    double c1 = 1.0, c4 = x[0];
    for (int i = 1; i < n; i++) {
      double c2 = 1.0;
      const double c5 = c4;
      c4 = x[i];
      for (int j = 0; j < i; j++) {
        const double c3 = x[i] - x[j];
        c2 = c2 * c3;
        if (j == i - 1) cc[i] = -c1 * c5 * cc[i - 1] / c2;
        cc[j] = c4 * cc[j] / c3;
      }
      c1 = c2;
    }
    T.dr0.assign(cc, cc + n);
  }
  // Wall row of the even derivative table (get_dr1, mod_radialFD.cu:355-358).
  T.dr1.assign(T.D.begin() + (nr - 1) * STEN, T.D.begin() + nr * STEN);
}

// In-place DIT plan: radices 4, 2, 3, 5 in that order; perm[n] is the LDS
// position of input sample n (mixed-radix digit reversal: the last pass's
// radix digit of n selects the outermost block).
bool make_fft_plan(int L, FftPlan& p, std::vector<int>& perm) {
  p.L = L;
  p.nrad = 0;
  int m = L;
  const int order[4] = {4, 2, 3, 5};
  for (int o = 0; o < 4; o++)
    while (m % order[o] == 0) {
      if (p.nrad >= MAXRAD) return false;
      p.rad[p.nrad++] = order[o];
      m /= order[o];
    }
  if (m != 1) return false;
  perm.assign(L, 0);
  for (int n = 0; n < L; n++) {
    int pos = 0, rem = n, len = L;
    for (int q = p.nrad - 1; q >= 0; q--) {
      const int rq = p.rad[q];
      len /= rq;
      pos += (rem % rq) * len;
      rem /= rq;
    }
    perm[n] = pos;
  }
  return true;
}

std::vector<double> centred_table(int nr, const std::vector<double>& T) {
  std::vector<double> C(2 * STEN * nr, 0.0);
  for (int p = 0; p < 2; p++)
    for (int i = 0; i < nr; i++) {
      const int left = std::max(0, i - IW), right = std::min(i + IW, nr - 1);
      const int i0 = i + IW - right;
      for (int s = 0; s < right + 1 - left; s++) {
        const int w = left + s - (i - IW);
        C[p * STEN * nr + STEN * i + w] = T[p * STEN * nr + STEN * i + s + i0];
      }
    }
  return C;
}

std::vector<double2> pass_twiddles(const std::vector<int>& radices) {
  std::vector<double2> tw;
  int S = 1;
  for (int R : radices) {
    for (int k = 0; k < S; k++)
      for (int q = 1; q < R; q++) {
        const long long e = (long long)q * k;
        const double a = -PI2 * double(e % (S * R)) / double(S * R);
        tw.push_back(make_double2(std::cos(a), std::sin(a)));
      }
    S *= R;
  }
  return tw;
}

// For a plan whose last pass has radix 8 (sub-length S = L/8): w^1, w^2, w^4
// of w = exp(-2 pi i k / L) for each k < S, the entries the compact form of
// that pass reads (the other powers are products of these).  Empty for any
// other last radix.
std::vector<double2> compact_twiddles(const std::vector<int>& radices) {
  std::vector<double2> tw;
  if (radices.empty() || radices.back() != 8) return tw;
  long long L = 1;
  for (int R : radices) L *= R;
  const long long S = L / 8;
  for (long long k = 0; k < S; k++)
    for (int q : {1, 2, 4}) {
      const double a = -PI2 * double((q * k) % L) / double(L);
      tw.push_back(make_double2(std::cos(a), std::sin(a)));
    }
  return tw;
}

}  // namespace nsp
