// Pseudo-spectral nonlinear term -(u.grad)u with 3/2 dealiasing
// (nonlinear(), mod_io_nonlinear.cu:380-409), as four streaming kernels:
//
//   dr3        radial derivative of the three components (derivr_k, flags 1,1,0)
//   theta_inv  per (level, l-tile): pad to NTP azimuthal points and inverse
//              transform u, d_r u and d_theta u / r (9 fields) over theta
//   zpass      per (level, theta) row: 12 real inverse transforms of length
//              3(NZ-1) (d_z u formed on the fly), the convective products
//              of calcNonlinear_k, and 3 real forward transforms
//   theta_fwd  per (level, l-tile): forward theta transform, 1/(NTP*NZF)
//              normalisation, truncation to NT x NZ with the m=-NT/2 and
//              l=NZ-1 zeroing of padBackward_k, and the mean pressure
//              gradient of pressGrad_k on mode (0,0)
//
// The padded spectra of padForward_k never exist in memory: padding is
// applied while loading into LDS.  The 2-D transform is the cuFFT Z2D/D2Z
// pair of mod_fft.cu:21-49 (theta = outer dimension, z = real dimension);
// at l = 0 only the real part of the theta-transformed column enters the
// real inverse transform.
#include <algorithm>

#include "fft_device.h"

namespace nsp {

using C = cpx<double>;

__device__ __forceinline__ int m_of(int it, int nt) { return it < nt / 2 ? it : it - nt; }

// ------------------------------------------------------------------ d/dr
// One thread per (mode, component) walks r from the axis to the wall with the
// levels i-3..i+3 in a register ring, so every input value is read from HBM
// once (a 2-D grid over (mode, level) re-reads each level seven times, and a
// level of the three components does not fit in L2).  Levels up to PF ahead
// are loaded early so each thread keeps several loads in flight.  The sum
// runs over the same terms in the same order as derivr_k.
// B = 1: batched launch, one member per blockIdx.z with its arguments from tab.
template <int B>
__global__ __launch_bounds__(256) void dr3_kernel(Geom g, const double* __restrict__ D,
                                                  const double2* __restrict__ u,
                                                  void* __restrict__ dRv, int f32, const NlMember* tab) {
  if constexpr (B != 0) {
    const NlMember& mb = tab[blockIdx.z];
    g = mb.g;
    D = gmem(tab, mb.D);
    u = gmem(tab, mb.u);
    dRv = gmem(tab, mb.dR);
    f32 = mb.f32;
  }
  if (step_done(g)) return;
  constexpr int NB = 12, PF = 5;  // ring slots; levels i-3 .. i+3+PF live at step i
  static_assert(STEN + PF <= NB, "ring too small");
  // One wave per (azimuthal row, chunk of 64 axial modes): its reads of u are
  // 1 KB runs of one row, its writes whole 64-value rows of a tile of the
  // tile-major d_r u, and the parity of m is the same on every lane.
  const int nch = (g.nz + 63) >> 6;
  const int gc = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int it_k = gc / nch, l_k = (gc - it_k * nch) * 64 + (threadIdx.x & 63);
  const int c = blockIdx.y;
  if (it_k >= g.nt || l_k >= g.nz) return;
  const int k = it_k * g.nz + l_k;
  const int nr = g.nr;
  const int m = m_of(it_k, g.nt);
  // parity table of the wave's row, read through the scalar cache
  const int odd = __builtin_amdgcn_readfirstlane(abs((abs(m) % 2) - (c < 2 ? 1 : 0)));
  const ctab Dw = ctab(D) + odd * STEN * nr;
  const double2* src = u + (size_t)c * g.n + k;
  const size_t ld = g.modes;
  // output: tile-major d_r u (dr_col), levels dr_plane apart
  const size_t dbase = (size_t)c * dr_cs(g) + (size_t)it_k * 64 + dr_col(g, l_k);
  const size_t dld = dr_plane(g);
  double2 w[NB];  // level L in slot L % NB
#pragma unroll
  for (int L = 0; L < IW + PF; L++) w[L] = L < nr ? src[L * ld] : make_double2(0.0, 0.0);
  for (int ib = 0; ib < nr; ib += NB) {
#pragma unroll
    for (int s = 0; s < NB; s++) {
      const int i = ib + s;
      if (i < nr) {
        const int L = i + IW + PF;
        if (L < nr) w[(s + IW + PF) % NB] = src[L * ld];
        const int left = max(0, i - IW), right = min(i + IW, nr - 1);
        const int oi = STEN * i + (i + IW - right) - left;  // D[ip0 + oi + lev] = D[ip0 + 7i + j + i0]
        double px = 0.0, py = 0.0;
#pragma unroll
        for (int q = 0; q < STEN; q++) {
          const int lev = i - IW + q;
          if (lev >= 0 && lev < nr) {
            const double de = Dw[oi + lev];
            const double2 a = w[(s - IW + q + NB) % NB];
            px += de * a.x;
            py += de * a.y;
          }
        }
        if (f32)  // mixed form: d_r u only feeds the FP32 theta transforms
          reinterpret_cast<float2*>(dRv)[dbase + i * dld] = make_float2(float(px), float(py));
        else
          reinterpret_cast<double2*>(dRv)[dbase + i * dld] = make_double2(px, py);
      }
    }
  }
}

// ------------------------------------------------------------------ theta inverse
template <int TL>
__global__ __launch_bounds__(256) void theta_inv_kernel(Geom g, FftPlan P, const int* __restrict__ perm,
                                                        const C* __restrict__ tw,
                                                        const double* __restrict__ r,
                                                        const double2* __restrict__ u,
                                                        const double2* __restrict__ dR,
                                                        double2* __restrict__ W, size_t ws) {
  if (step_done(g)) return;
  extern __shared__ double2 smem[];
  C* buf = reinterpret_cast<C*>(smem);
  const int l0 = blockIdx.x * TL, i = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const double ri = r[i];
  const int half = g.nt / 2;
  for (int f = 0; f < 9; f++) {
    const int c = f % 3, kind = f / 3;  // kind 0: u, 1: d_r u, 2: d_theta u / r
    const double2* src = u + c * g.n + (size_t)i * g.modes;
    const double2* srd = dR + c * dr_cs(g) + (size_t)i * dr_plane(g);  // tile-major d_r u
    for (int e = tid; e < g.ntp * TL; e += nthr) buf[e] = C{0.0, 0.0};
    __syncthreads();
    for (int e = tid; e < g.nt * TL; e += nthr) {
      const int it = e / TL, tl = e - it * TL, l = l0 + tl;
      if (l < g.nz) {
        const double2 v = kind == 1 ? srd[(size_t)it * 64 + dr_col(g, l)] : src[(size_t)it * g.nz + l];
        C w{v.x, v.y};
        if (kind == 2) {  // derivt_k (mod_deriv.cu:10-21)
          const double kt = double(m_of(it, g.nt));
          w = C{-kt * v.y / ri, kt * v.x / ri};
        }
        const int j = it < half ? it : it + half;  // padForward_k row map
        buf[perm[j] * TL + tl] = w;
      }
    }
    __syncthreads();
    fft_passes<double>(buf, P, TL, 1.0, tw, tid, nthr);
    double2* dst = W + f * ws + (size_t)i * tp_plane(g);  // tile-major W
    for (int e = tid; e < g.ntp * TL; e += nthr) {
      const int j = e / TL, tl = e - j * TL, l = l0 + tl;
      if (l < g.nz) dst[(size_t)j * 64 + tp_col(g, l)] = make_double2(buf[e].x, buf[e].y);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ z pass
// Real transforms of length N = NZF through complex transforms of length
// NH = N/2 on z[m] = x[2m] + i x[2m+1].
//   inverse: Z[k] = E + iO, E = X[k] + conj(X[NH-k]),
//            O = (X[k] - conj(X[NH-k])) e^{+2 pi i k/N}
//   forward: E = (Z[k] + conj(Z[NH-k]))/2, O = (Z[k] - conj(Z[NH-k]))/(2i),
//            X[k] = E + e^{-2 pi i k/N} O
template <int MPT>
__global__ __launch_bounds__(256) void zpass_kernel(Geom g, FftPlan P, const int* __restrict__ perm,
                                                    const C* __restrict__ tw,
                                                    const C* __restrict__ twz,
                                                    const double* __restrict__ r,
                                                    const double2* __restrict__ W, size_t ws,
                                                    double2* __restrict__ Y) {
  if (step_done(g)) return;
  extern __shared__ double2 smem[];
  C* buf = reinterpret_cast<C*>(smem);
  const int row = blockIdx.x;
  const int i = row / g.ntp;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nh = g.nh, nz = g.nz, nzf = g.nzf;
  const double rad = r[i];
  // row th of level i in the tile-major planes of W, Y: value k at roff + tp_col(k)
  const size_t roff = (size_t)(row / g.ntp) * tp_plane(g) + (size_t)(row % g.ntp) * 64;

  // inverse real transform of spectral row X (entries l < nz, zero beyond);
  // dz: differentiate in z first (derivz_k, mod_deriv.cu:24-35)
  auto c2r = [&](const double2* X, bool dz) {
    for (int k = tid; k < nh; k += nthr) {
      const int kc = nh - k;
      C a{0.0, 0.0}, b{0.0, 0.0};
      if (k < nz) { const double2 v = X[tp_col(g, k)]; a = C{v.x, v.y}; }
      if (kc < nz) { const double2 v = X[tp_col(g, kc)]; b = C{v.x, v.y}; }
      if (k == 0) { a.y = 0.0; b.y = 0.0; }
      if (dz) {
        const double ka = g.kz0 * double(k), kb = g.kz0 * double(kc);
        a = C{-ka * a.y, ka * a.x};
        b = C{-kb * b.y, kb * b.x};
      }
      const C E{a.x + b.x, a.y - b.y};
      const C dlt{a.x - b.x, a.y + b.y};
      const C w = twz[k];
      const C O{dlt.x * w.x + dlt.y * w.y, dlt.y * w.x - dlt.x * w.y};  // dlt * conj(w)
      buf[perm[k]] = C{E.x - O.y, E.y + O.x};
    }
    __syncthreads();
    fft_passes<double>(buf, P, 1, 1.0, tw, tid, nthr);
  };
  auto sample = [&](int n) -> double {
    const C v = buf[n >> 1];
    return (n & 1) ? v.y : v.x;
  };

  double ur[MPT], ut[MPT], uz[MPT], acc[MPT];
  c2r(W + 0 * ws + roff, false);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = tid + q * nthr; ur[q] = n < nzf ? sample(n) : 0.0; }
  __syncthreads();
  c2r(W + 1 * ws + roff, false);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = tid + q * nthr; ut[q] = n < nzf ? sample(n) : 0.0; }
  __syncthreads();
  c2r(W + 2 * ws + roff, false);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = tid + q * nthr; uz[q] = n < nzf ? sample(n) : 0.0; }
  __syncthreads();

  for (int c = 0; c < 3; c++) {
    // calcNonlinear_k (mod_io_nonlinear.cu:142-170), summed in its order
    c2r(W + (3 + c) * ws + roff, false);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = tid + q * nthr; acc[q] = n < nzf ? ur[q] * sample(n) : 0.0; }
    __syncthreads();
    c2r(W + (6 + c) * ws + roff, false);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = tid + q * nthr; if (n < nzf) acc[q] += ut[q] * sample(n); }
    __syncthreads();
    c2r(W + c * ws + roff, true);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = tid + q * nthr; if (n < nzf) acc[q] += uz[q] * sample(n); }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MPT; q++) {
      const int n = tid + q * nthr;
      if (n < nzf) {
        double v = acc[q];
        if (c == 0) v = v - ut[q] * (ut[q] / rad);
        else if (c == 1) v = v + ur[q] * (ut[q] / rad);
        double* slot = reinterpret_cast<double*>(&buf[perm[n >> 1]]) + (n & 1);
        *slot = -v;
      }
    }
    __syncthreads();
    fft_passes<double>(buf, P, 1, -1.0, tw, tid, nthr);
    double2* dst = Y + c * ws + roff;
    for (int k = tid; k < nz; k += nthr) {
      C X{0.0, 0.0};
      if (k < nz - 1) {
        const C zk = buf[k], zc = buf[k == 0 ? 0 : nh - k];
        const C E{0.5 * (zk.x + zc.x), 0.5 * (zk.y - zc.y)};
        const C d{zk.x - zc.x, zk.y + zc.y};
        const C O{0.5 * d.y, -0.5 * d.x};
        const C w = twz[k];
        X = C{E.x + (w.x * O.x - w.y * O.y), E.y + (w.x * O.y + w.y * O.x)};
      }
      dst[tp_col(g, k)] = make_double2(X.x, X.y);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ theta forward
template <int TL>
__global__ __launch_bounds__(256) void theta_fwd_kernel(Geom g, FftPlan P, const int* __restrict__ perm,
                                                        const C* __restrict__ tw,
                                                        const double2* __restrict__ Y, size_t ws,
                                                        const double* __restrict__ D,
                                                        const double2* __restrict__ u,
                                                        double2* __restrict__ N) {
  if (step_done(g)) return;
  extern __shared__ double2 smem[];
  C* buf = reinterpret_cast<C*>(smem);
  const int l0 = blockIdx.x * TL, i = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int half = g.nt / 2;
  const double norm = double(g.ntp) * double(g.nzf);  // mod_fft.cu:41
  for (int c = 0; c < 3; c++) {
    const double2* src = Y + c * ws + (size_t)i * tp_plane(g);  // tile-major Y
    for (int e = tid; e < g.ntp * TL; e += nthr) {
      const int j = e / TL, tl = e - j * TL, l = l0 + tl;
      C v{0.0, 0.0};
      if (l < g.nz - 1) { const double2 a = src[(size_t)j * 64 + tp_col(g, l)]; v = C{a.x, a.y}; }
      buf[perm[j] * TL + tl] = v;
    }
    __syncthreads();
    fft_passes<double>(buf, P, TL, -1.0, tw, tid, nthr);
    for (int e = tid; e < g.nt * TL; e += nthr) {
      const int it = e / TL, tl = e - it * TL, l = l0 + tl;
      if (l < g.nz) {
        const int j = it < half ? it : it + half;  // padBackward_k row map
        C v = buf[j * TL + tl];
        v.x = v.x / norm;
        v.y = v.y / norm;
        if (it == half || l == g.nz - 1) v = C{0.0, 0.0};
        if (c == 2 && it == 0 && l == 0) {  // pressGrad_k (mod_radialFD.cu:100-111)
          double du = 0.0;
          for (int j2 = 0; j2 < IW + 1; j2++)
            du += D[STEN * (g.nr - 1) + j2 + IW] * u[2 * g.n + (size_t)(g.nr - 1 - IW + j2) * g.modes].x;
          v.x -= (2.0 / g.re) * du;
        }
        N[c * dr_cs(g) + (size_t)i * dr_plane(g) + (size_t)it * 64 + dr_col(g, l)] = make_double2(v.x, v.y);  // tile-major N(u)
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ launch
static int g_max_lds = 0;

static int max_lds() {
  if (!g_max_lds) {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 65536;
    g_max_lds = v;
  }
  return g_max_lds;
}

template <int TL>
static void run_theta(const Geom& g, const DevTables& t, const NlBufs& nb, const double2* u, double2* N,
                      hipStream_t s, bool inverse) {
  const size_t lds = size_t(g.ntp) * TL * sizeof(double2);
  dim3 grid((g.nz + TL - 1) / TL, g.nr);
  if (inverse) {
    (void)hipFuncSetAttribute((const void*)theta_inv_kernel<TL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    theta_inv_kernel<TL><<<grid, 256, lds, s>>>(g, nb.pt, nb.perm_t, (const C*)nb.tw_t, t.r, u, nb.dR, nb.W,
                                                nb.wstride);
  } else {
    (void)hipFuncSetAttribute((const void*)theta_fwd_kernel<TL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    theta_fwd_kernel<TL><<<grid, 256, lds, s>>>(g, nb.pt, nb.perm_t, (const C*)nb.tw_t, nb.Y, nb.wstride, t.D,
                                                u, N);
  }
}

static void theta_dispatch(const Geom& g, const DevTables& t, const NlBufs& nb, const double2* u, double2* N,
                           hipStream_t s, bool inverse) {
  const int lim = max_lds();
  const size_t col = size_t(g.ntp) * sizeof(double2);
  if (col * 16 <= (size_t)lim) run_theta<16>(g, t, nb, u, N, s, inverse);
  else if (col * 8 <= (size_t)lim) run_theta<8>(g, t, nb, u, N, s, inverse);
  else if (col * 4 <= (size_t)lim) run_theta<4>(g, t, nb, u, N, s, inverse);
  else if (col * 2 <= (size_t)lim) run_theta<2>(g, t, nb, u, N, s, inverse);
  else run_theta<1>(g, t, nb, u, N, s, inverse);
}

template <int MPT>
static void run_z(const Geom& g, const DevTables& t, const NlBufs& nb, hipStream_t s) {
  const size_t lds = size_t(g.nh) * sizeof(double2);
  (void)hipFuncSetAttribute((const void*)zpass_kernel<MPT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  zpass_kernel<MPT><<<g.nr * g.ntp, 256, lds, s>>>(g, nb.ph, nb.perm_h, (const C*)nb.tw_h, (const C*)nb.tw_z,
                                                    t.r, nb.W, nb.wstride, nb.Y);
}

const char* kernel_name(int id) {
  static const char* names[K_COUNT] = {"dr3_kernel", "theta_inv_kernel", "zpass_kernel", "theta_fwd_kernel",
                                       "linear_sweep_kernel(predictor)", "linear_sweep_kernel(corrector)"};
  return (id >= 0 && id < K_COUNT) ? names[id] : "";
}

void launch_dr3(const Geom& g, const double* D, const double2* u, double2* dR, bool f32, hipStream_t s,
                const Batch* bt) {
  // four (row, 64-mode chunk) waves per workgroup
  if (bt) {
    const int chunks = bt->max_nt * ((bt->max_nz + 63) / 64);
    dim3 gd((chunks + 3) / 4, 3, bt->count);
    dr3_kernel<1><<<gd, 256, 0, s>>>(g, D, u, dR, f32 ? 1 : 0, bt->tab);
  } else {
    const int chunks = g.nt * ((g.nz + 63) / 64);
    dim3 gd((chunks + 3) / 4, 3);
    dr3_kernel<0><<<gd, 256, 0, s>>>(g, D, u, dR, f32 ? 1 : 0, nullptr);
  }
}

void launch_nonlinear(const Geom& g, const DevTables& t, const NlBufs& nb, const double2* u, double2* N,
                      hipStream_t s, Profiler* pf) {
  if (nb.use_ct && !pf && nb.var.nl_chunks > 1 && nb.s2) {
    // Pipelined: the theta inverse of level chunk c + 1 (memory-bound) runs on
    // s beside the z pass of chunk c (VALU-bound) on s2; each z chunk waits
    // for its own theta chunk, the theta forward for the last z chunk.  The
    // chunks' planes of W and Y are disjoint.  Under stream capture the
    // events become graph edges (s2 joins the capture at its first wait).
    const NlVariant& v = nb.var;
    const int K = std::min(std::min(v.nl_chunks, kNlMaxChunks), g.nr);
    launch_dr3(g, t.D, u, nb.dR, nb.mixed, s);
    for (int c = 0; c < K; c++) {
      Geom gc = g;
      gc.lv0 = (int)((long long)c * g.nr / K);
      gc.nlv = (int)((long long)(c + 1) * g.nr / K) - gc.lv0;
      if (nb.mixed)
        ct_theta_inv<float>(v, gc, (const cpx<float>*)nb.ct_tw_t32, t.r, u, nb.dR, (cpx<float>*)nb.W32, nb.wstride, s);
      else
        ct_theta_inv<double>(v, gc, (const cpx<double>*)nb.ct_tw_t, t.r, u, nb.dR, (cpx<double>*)nb.W, nb.wstride,
                             s);
      (void)hipEventRecord(nb.ev[c], s);
      (void)hipStreamWaitEvent(nb.s2, nb.ev[c], 0);
      if (nb.mixed)
        ct_zpass<float>(v, gc, (const cpx<float>*)nb.ct_tw_h32, (const cpx<float>*)nb.tw_z32, t.r,
                        (const cpx<float>*)nb.W32, nb.wstride, (cpx<float>*)nb.Y32, nb.s2);
      else
        ct_zpass<double>(v, gc, (const cpx<double>*)nb.ct_tw_h, (const cpx<double>*)nb.tw_z, t.r,
                         (const cpx<double>*)nb.W, nb.wstride, (cpx<double>*)nb.Y, nb.s2);
    }
    (void)hipEventRecord(nb.ev[K], nb.s2);
    (void)hipStreamWaitEvent(s, nb.ev[K], 0);
    if (nb.mixed)
      ct_theta_fwd<float>(v, g, (const cpx<float>*)nb.ct_tw_t32, (const cpx<float>*)nb.Y32, nb.wstride, t.D, u, N, s);
    else
      ct_theta_fwd<double>(v, g, (const cpx<double>*)nb.ct_tw_t, (const cpx<double>*)nb.Y, nb.wstride, t.D, u, N, s);
    return;
  }
  if (nb.use_ct) {
    const NlVariant& v = nb.var;
    bool ok = true;
    if (pf) pf->begin(K_DR3, s);
    launch_dr3(g, t.D, u, nb.dR, nb.mixed, s);
    if (pf) pf->end(s);
    if (pf) pf->begin(K_THETA_INV, s);
    if (nb.mixed)
      ok &= ct_theta_inv<float>(v, g, (const cpx<float>*)nb.ct_tw_t32, t.r, u, nb.dR, (cpx<float>*)nb.W32,
                                nb.wstride, s);
    else
      ok &= ct_theta_inv<double>(v, g, (const cpx<double>*)nb.ct_tw_t, t.r, u, nb.dR, (cpx<double>*)nb.W,
                                 nb.wstride, s);
    if (pf) pf->end(s);
    if (pf) pf->begin(K_ZPASS, s);
    if (nb.mixed)
      ok &= ct_zpass<float>(v, g, (const cpx<float>*)nb.ct_tw_h32, (const cpx<float>*)nb.tw_z32, t.r,
                            (const cpx<float>*)nb.W32, nb.wstride, (cpx<float>*)nb.Y32, s);
    else
      ok &= ct_zpass<double>(v, g, (const cpx<double>*)nb.ct_tw_h, (const cpx<double>*)nb.tw_z, t.r,
                             (const cpx<double>*)nb.W, nb.wstride, (cpx<double>*)nb.Y, s);
    if (pf) pf->end(s);
    if (pf) pf->begin(K_THETA_FWD, s);
    if (nb.mixed)
      ok &= ct_theta_fwd<float>(v, g, (const cpx<float>*)nb.ct_tw_t32, (const cpx<float>*)nb.Y32, nb.wstride, t.D, u,
                                N, s);
    else
      ok &= ct_theta_fwd<double>(v, g, (const cpx<double>*)nb.ct_tw_t, (const cpx<double>*)nb.Y, nb.wstride, t.D, u,
                                 N, s);
    if (pf) pf->end(s);
    (void)ok;  // variants are validated when chosen
    return;
  }
  if (pf) pf->begin(K_DR3, s);
  launch_dr3(g, t.D, u, nb.dR, false, s);
  if (pf) pf->end(s);
  if (pf) pf->begin(K_THETA_INV, s);
  theta_dispatch(g, t, nb, u, N, s, true);
  if (pf) pf->end(s);
  if (pf) pf->begin(K_ZPASS, s);
  const int mpt = (g.nzf + 255) / 256;
  if (mpt <= 1) run_z<1>(g, t, nb, s);
  else if (mpt <= 2) run_z<2>(g, t, nb, s);
  else if (mpt <= 4) run_z<4>(g, t, nb, s);
  else if (mpt <= 6) run_z<6>(g, t, nb, s);
  else if (mpt <= 8) run_z<8>(g, t, nb, s);
  else if (mpt <= 12) run_z<12>(g, t, nb, s);
  else if (mpt <= 18) run_z<18>(g, t, nb, s);
  else if (mpt <= 24) run_z<24>(g, t, nb, s);
  else run_z<36>(g, t, nb, s);
  if (pf) pf->end(s);
  if (pf) pf->begin(K_THETA_FWD, s);
  theta_dispatch(g, t, nb, u, N, s, false);
  if (pf) pf->end(s);
}

}  // namespace nsp
