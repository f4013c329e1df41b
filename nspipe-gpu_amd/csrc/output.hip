// State conversion and output kernels outside the time step: remesh onto
// another grid (tools/Re_Mesh), physical-space fields and the cross-section
// integrals of the reference's diagnostics files (mod_io_nonlinear.cu).
#include "fft_device.h"
#include "nsp_internal.h"

namespace nsp {

using C = cpx<double>;

// Remesh gather (rmsh_k, tools/Re_Mesh/src/mod_remesh.cu:8-29, then
// modeZZ_k :34-40).  One thread per (new mode, component): a mode present
// on the old grid gets, at every new radial row, the remesh row's weights
// times the old column, summed over the row's nonzero band in increasing
// column order; other modes are zero.  ur, ut take the odd-in-r half for even
// m, uz the even half (parity index |(|m| mod 2) - flag|, flag 1 for ur, ut).
__global__ __launch_bounds__(256) void remesh_kernel(Geom gn, Geom go, const double* __restrict__ RM,
                                                     const int2* __restrict__ band,
                                                     const double2* __restrict__ uo, double2* __restrict__ un) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (k >= gn.modes) return;
  const int it = k / gn.nz, l = k - it * gn.nz;
  const int m = it < gn.nt / 2 ? it : it - gn.nt;
  const bool src = m >= -go.nt / 2 && m < go.nt / 2 && l < go.nz;
  const int ko = src ? (m >= 0 ? m : m + go.nt) * go.nz + l : 0;
  const int par = abs((abs(m) % 2) - (c < 2 ? 1 : 0));
  const double* R = RM + (size_t)par * gn.nr * go.nr;
  const double2* col = uo + (size_t)c * go.n + ko;
  double2* dst = un + (size_t)c * gn.n + k;
  for (int i = 0; i < gn.nr; i++) {
    double2 v = make_double2(0.0, 0.0);
    if (src) {
      const int2 bd = band[i];
      for (int j = bd.x; j <= bd.y; j++) {
        const double w = R[(size_t)i * go.nr + j];
        const double2 a = col[(size_t)j * go.modes];
        v.x += w * a.x;
        v.y += w * a.y;
      }
    }
    if (k == 0) {  // mode (0,0): no radial velocity, real ut and uz
      if (c == 0) v = make_double2(0.0, 0.0);
      else v.y = 0.0;
    }
    dst[(size_t)i * gn.modes] = v;
  }
}

void launch_remesh(const Geom& gn, const Geom& go, const double* RM, const int2* band, const double2* uo,
                   double2* un, hipStream_t s) {
  const dim3 grid((gn.modes + 255) / 256, 3);
  remesh_kernel<<<grid, 256, 0, s>>>(gn, go, RM, band, uo, un);
}

// ------------------------------------------------------------------ physical space
// padForward_k (mod_io_nonlinear.cu:173-192) and the theta pass of the 2-D
// inverse transform (mod_fft.cu:43-49) for one spectral component: block
// (l-tile, i) transforms TL axial columns of level i; rows it < NT/2 go to
// padded row it, the others to it + NT/2, rows NT/2..NT-1 stay zero.
template <int TL>
__global__ __launch_bounds__(256) void phys_theta_kernel(Geom g, FftPlan P, const int* __restrict__ perm,
                                                         const C* __restrict__ tw, const double2* __restrict__ u,
                                                         double2* __restrict__ Wt) {
  extern __shared__ double2 smem[];
  C* buf = reinterpret_cast<C*>(smem);
  const int l0 = blockIdx.x * TL, i = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int half = g.nt / 2;
  const double2* src = u + (size_t)i * g.modes;
  for (int e = tid; e < g.ntp * TL; e += nthr) buf[e] = C{0.0, 0.0};
  __syncthreads();
  for (int e = tid; e < g.nt * TL; e += nthr) {
    const int it = e / TL, tl = e - it * TL, l = l0 + tl;
    if (l < g.nz) {
      const double2 v = src[(size_t)it * g.nz + l];
      const int j = it < half ? it : it + half;
      buf[perm[j] * TL + tl] = C{v.x, v.y};
    }
  }
  __syncthreads();
  fft_passes<double>(buf, P, TL, 1.0, tw, tid, nthr);
  double2* dst = Wt + (size_t)i * g.ntp * g.nz;
  for (int e = tid; e < g.ntp * TL; e += nthr) {
    const int j = e / TL, tl = e - j * TL, l = l0 + tl;
    if (l < g.nz) dst[(size_t)j * g.nz + l] = make_double2(buf[e].x, buf[e].y);
  }
}

// Axial real synthesis of one (i, theta) row of Wt: the NZF samples through a
// complex transform of length NH = NZF/2 on z[m] = x[2m] + i x[2m+1]
//   Z[k] = E + iO, E = X[k] + conj(X[NH-k]), O = (X[k] - conj(X[NH-k])) e^{+2 pi i k/NZF}
// (imaginary parts of X[0] ignored, X[k] = 0 for k >= NZ).  Row `row` of out
// (pitch doubles) receives the samples, then zeros up to the pitch; mode 1
// stores their squares instead, mode 2 adds the squares.
__global__ __launch_bounds__(256) void phys_z_kernel(Geom g, FftPlan P, const int* __restrict__ perm,
                                                     const C* __restrict__ tw, const C* __restrict__ twz,
                                                     const double2* __restrict__ Wt, double* __restrict__ out,
                                                     size_t pitch, int mode) {
  extern __shared__ double2 smem[];
  C* buf = reinterpret_cast<C*>(smem);
  const int row = blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nh = g.nh, nz = g.nz;
  const double2* X = Wt + (size_t)row * nz;
  for (int k = tid; k < nh; k += nthr) {
    const int kc = nh - k;
    C a{0.0, 0.0}, b{0.0, 0.0};
    if (k < nz) { const double2 v = X[k]; a = C{v.x, v.y}; }
    if (kc < nz) { const double2 v = X[kc]; b = C{v.x, v.y}; }
    if (k == 0) { a.y = 0.0; b.y = 0.0; }
    const C E{a.x + b.x, a.y - b.y};
    const C dlt{a.x - b.x, a.y + b.y};
    const C w = twz[k];
    const C O{dlt.x * w.x + dlt.y * w.y, dlt.y * w.x - dlt.x * w.y};  // dlt * conj(w)
    buf[perm[k]] = C{E.x - O.y, E.y + O.x};
  }
  __syncthreads();
  fft_passes<double>(buf, P, 1, 1.0, tw, tid, nthr);
  double* o = out + (size_t)row * pitch;
  for (size_t n = tid; n < pitch; n += nthr) {
    double v = 0.0;
    if (n < (size_t)g.nzf) {
      const C zc = buf[n >> 1];
      v = (n & 1) ? zc.y : zc.x;
    }
    if (mode == 0) o[n] = v;
    else if (mode == 1) o[n] = v * v;
    else o[n] += v * v;
  }
}

// q1kernel_k (mod_io_nonlinear.cu:81-103) and the theta sum of intq
// (mod_cublasflux.cu:119-130, weights qaux = 1): thread (n, i) sums the
// periodic three-point rule 0.5/(3 NTP) (f[j-1] + 4 f[j] + f[j+1]) over j.
__global__ __launch_bounds__(256) void cross_theta_kernel(Geom g, const double* __restrict__ F, size_t pitch,
                                                          double* __restrict__ q1) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (n >= g.nzf) return;
  const double* f = F + (size_t)i * g.ntp * pitch + n;
  const double w = 0.5 * (1.0 / (3.0 * double(g.ntp)));
  double acc = 0.0;
  for (int j = 0; j < g.ntp; j++) {
    const int ja = j < 1 ? g.ntp - 1 : j - 1, jb = j > g.ntp - 2 ? 0 : j + 1;
    acc += w * (f[(size_t)ja * pitch] + 4.0 * f[(size_t)j * pitch] + f[(size_t)jb * pitch]);
  }
  q1[(size_t)i * g.nzf + n] = acc;
}

// the radial sum of intq with the quadrature weights rdr
__global__ __launch_bounds__(256) void cross_r_kernel(Geom g, const double* __restrict__ q1,
                                                      const double* __restrict__ rdr, double* __restrict__ q) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nzf) return;
  double acc = 0.0;
  for (int i = 0; i < g.nr; i++) acc += q1[(size_t)i * g.nzf + n] * rdr[i];
  q[n] = acc;
}

void launch_physical(const Geom& g, const NlBufs& nb, const double2* u, double2* Wt, double* out, size_t pitch,
                     int mode, hipStream_t s) {
  const size_t col = size_t(g.ntp) * sizeof(double2);
  const dim3 tg((g.nz + 3) / 4, g.nr), tg2((g.nz + 1) / 2, g.nr), tg1(g.nz, g.nr);
  if (col * 4 <= 65536)
    phys_theta_kernel<4><<<tg, 256, col * 4, s>>>(g, nb.pt, nb.perm_t, (const C*)nb.tw_t, u, Wt);
  else if (col * 2 <= 65536)
    phys_theta_kernel<2><<<tg2, 256, col * 2, s>>>(g, nb.pt, nb.perm_t, (const C*)nb.tw_t, u, Wt);
  else
    phys_theta_kernel<1><<<tg1, 256, col, s>>>(g, nb.pt, nb.perm_t, (const C*)nb.tw_t, u, Wt);
  const size_t lz = size_t(g.nh) * sizeof(double2);
  (void)hipFuncSetAttribute((const void*)phys_z_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lz);
  phys_z_kernel<<<g.nr * g.ntp, 256, lz, s>>>(g, nb.ph, nb.perm_h, (const C*)nb.tw_h, (const C*)nb.tw_z, Wt, out,
                                             pitch, mode);
}

void launch_cross(const Geom& g, const double* F, size_t pitch, const double* rdr, double* q1, double* q,
                  hipStream_t s) {
  cross_theta_kernel<<<dim3((g.nzf + 255) / 256, g.nr), 256, 0, s>>>(g, F, pitch, q1);
  cross_r_kernel<<<(g.nzf + 255) / 256, 256, 0, s>>>(g, q1, rdr, q);
}

}  // namespace nsp
