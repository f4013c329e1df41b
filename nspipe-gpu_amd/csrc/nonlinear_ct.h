// Nonlinear-term kernels with compile-time transform plans (see fft_ct.h).
// Same dataflow and arithmetic as nonlinear.hip (theta inverse -> z pass ->
// theta forward), with the transform precision T as a template parameter:
// T = double for the FP64 form, float for the FP32 transforms of the mixed
// form.  Lengths without a compiled plan use the generic kernels.
#pragma once
#include "fft_ct.h"

namespace nsp {

__device__ __forceinline__ int m_of_ct(int it, int nt) { return it < nt / 2 ? it : it - nt; }

// ------------------------------------------------------------------ block order
// Workgroups are dealt to the 8 XCDs round robin; remap so that each XCD
// walks one contiguous range of logical blocks and neighbouring tiles and
// radial levels meet in the same L2.
__device__ __forceinline__ int xcd_block(int b, int G) {
  constexpr int NX = 8;
  const int q = G / NX, rmd = G % NX;
  const int x = b % NX, j = b / NX;
  return (x < rmd ? x * (q + 1) : rmd * (q + 1) + (x - rmd) * q) + j;
}

// ------------------------------------------------------------------ theta inverse
// W: 9 intermediates [f][NR][NTP][NZ] in precision T: u, d_r u, d_theta u / r.
// FUSED: the radial derivative (derivr_k, mod_radialFD.cu:10-30, flags 1,1,0)
// is formed here from levels i-3..i+3 read through L2; otherwise it is read
// from dR (dr3_kernel).
template <typename T, int TL, int NTHR, int FUSED, typename PL>
__global__ __launch_bounds__(NTHR) void theta_inv_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ Dc,
                                                     const double2* __restrict__ u,
                                                     const double2* __restrict__ dR,
                                                     cpx<T>* __restrict__ W, size_t ws) {
  constexpr int NTP = PL::L;
  constexpr int NT = 2 * NTP / 3;
  constexpr int EPT = (NT * TL + NTHR - 1) / NTHR;
  __shared__ cpx<T> buf[NTP * TL];
  const int nr = g.nr, nz = g.nz;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = lb % nr, l0 = (lb / nr) * TL;
  const int tid = threadIdx.x;
  const double ri = r[i];
  constexpr int half = NT / 2;
  for (int c = 0; c < 3; c++) {
    const double2* src = u + c * g.n;
    double2 U[EPT], DR[EPT];
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const int e = tid + q * NTHR;
      const int it = e / TL, tl = e - it * TL, l = l0 + tl;
      U[q] = DR[q] = make_double2(0.0, 0.0);
      if (e < NT * TL && l < nz) {
        const size_t off = (size_t)it * nz + l;
        if (FUSED) {
          const int mm = it < half ? it : it - NT;
          const int par = abs((abs(mm) % 2) - (c < 2 ? 1 : 0));
          const double* Dr = Dc + (par * nr + i) * STEN;
          double px = 0.0, py = 0.0;
#pragma unroll
          for (int w = 0; w < STEN; w++) {
            const int lev = i - IW + w;
            if (lev >= 0 && lev < nr) {
              const double2 a = src[(size_t)lev * g.modes + off];
              px += Dr[w] * a.x;
              py += Dr[w] * a.y;
              if (w == IW) U[q] = a;
            }
          }
          DR[q] = make_double2(px, py);
        } else {
          U[q] = src[(size_t)i * g.modes + off];
          DR[q] = dR[c * g.n + (size_t)i * g.modes + off];
        }
      }
    }
    for (int kind = 0; kind < 3; kind++) {  // 0: u, 1: d_r u, 2: d_theta u / r
      for (int e = tid; e < (NTP - NT) * TL; e += NTHR) {  // padding rows
        const int j = half + e / TL, tl = e % TL;
        buf[PL::digitrev(j) * TL + tl] = cpx<T>{T(0), T(0)};
      }
#pragma unroll
      for (int q = 0; q < EPT; q++) {
        const int e = tid + q * NTHR;
        if (e < NT * TL) {
          const int it = e / TL, tl = e - it * TL;
          const double2 v = (kind == 1) ? DR[q] : U[q];
          cpx<T> w{T(v.x), T(v.y)};
          if (kind == 2) {  // derivt_k (mod_deriv.cu:10-21)
            const double kt = double(it < half ? it : it - NT);
            w = cpx<T>{T(-kt * v.y / ri), T(kt * v.x / ri)};
          }
          const int j = it < half ? it : it + half;  // padForward_k row map
          buf[PL::digitrev(j) * TL + tl] = w;
        }
      }
      __syncthreads();
      fft_run<T, TL, NTHR, +1>(PL{}, buf, tw, tid);
      cpx<T>* dst = W + (kind * 3 + c) * ws + (size_t)i * NTP * nz;
      for (int e = tid; e < NTP * TL; e += NTHR) {
        const int j = e / TL, tl = e - j * TL, l = l0 + tl;
        if (l < nz) dst[(size_t)j * nz + l] = buf[e];
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ z pass
template <typename T, int NTHR, int MPT, typename PL>
__global__ __launch_bounds__(NTHR) void zpass_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                 const cpx<T>* __restrict__ twz,
                                                 const double* __restrict__ r,
                                                 const cpx<T>* __restrict__ W, size_t ws,
                                                 cpx<T>* __restrict__ Y) {
  constexpr int NH = PL::L;
  constexpr int NZF = 2 * NH;
  __shared__ cpx<T> buf[NH];
  const int row = blockIdx.x;
  const int i = row / g.ntp;
  const int tid = threadIdx.x;
  const int nz = g.nz;
  const T rad = T(r[i]);
  const size_t roff = (size_t)row * nz;
  const T kz0 = T(g.kz0);

  auto c2r = [&](const cpx<T>* X, bool dz) {
    for (int k = tid; k < NH; k += NTHR) {
      const int kc = NH - k;
      cpx<T> a{T(0), T(0)}, b{T(0), T(0)};
      if (k < nz) a = X[k];
      if (kc < nz) b = X[kc];
      if (k == 0) { a.y = T(0); b.y = T(0); }
      if (dz) {  // derivz_k (mod_deriv.cu:24-35)
        const T ka = kz0 * T(k), kb = kz0 * T(kc);
        a = cpx<T>{-ka * a.y, ka * a.x};
        b = cpx<T>{-kb * b.y, kb * b.x};
      }
      const cpx<T> E{a.x + b.x, a.y - b.y};
      const cpx<T> dl{a.x - b.x, a.y + b.y};
      const cpx<T> O = cx_mulc(dl, twz[k]);
      buf[PL::digitrev(k)] = cpx<T>{E.x - O.y, E.y + O.x};
    }
    __syncthreads();
    fft_run<T, 1, NTHR, +1>(PL{}, buf, tw, tid);
  };
  auto sample = [&](int n) -> T {
    const cpx<T> v = buf[n >> 1];
    return (n & 1) ? v.y : v.x;
  };

  T ur[MPT], ut[MPT], uz[MPT], acc[MPT];
  c2r(W + 0 * ws + roff, false);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = tid + q * NTHR; ur[q] = n < NZF ? sample(n) : T(0); }
  __syncthreads();
  c2r(W + 1 * ws + roff, false);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = tid + q * NTHR; ut[q] = n < NZF ? sample(n) : T(0); }
  __syncthreads();
  c2r(W + 2 * ws + roff, false);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = tid + q * NTHR; uz[q] = n < NZF ? sample(n) : T(0); }
  __syncthreads();

  for (int c = 0; c < 3; c++) {
    // calcNonlinear_k (mod_io_nonlinear.cu:142-170), summed in its order
    c2r(W + (3 + c) * ws + roff, false);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = tid + q * NTHR; acc[q] = n < NZF ? ur[q] * sample(n) : T(0); }
    __syncthreads();
    c2r(W + (6 + c) * ws + roff, false);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = tid + q * NTHR; if (n < NZF) acc[q] += ut[q] * sample(n); }
    __syncthreads();
    c2r(W + c * ws + roff, true);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = tid + q * NTHR; if (n < NZF) acc[q] += uz[q] * sample(n); }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MPT; q++) {
      const int n = tid + q * NTHR;
      if (n < NZF) {
        T v = acc[q];
        if (c == 0) v = v - ut[q] * (ut[q] / rad);
        else if (c == 1) v = v + ur[q] * (ut[q] / rad);
        T* slot = reinterpret_cast<T*>(&buf[PL::digitrev(n >> 1)]) + (n & 1);
        *slot = -v;
      }
    }
    __syncthreads();
    fft_run<T, 1, NTHR, -1>(PL{}, buf, tw, tid);
    cpx<T>* dst = Y + c * ws + roff;
    for (int k = tid; k < nz; k += NTHR) {
      cpx<T> X{T(0), T(0)};
      if (k < nz - 1) {
        const cpx<T> zk = buf[k], zc = buf[k == 0 ? 0 : NH - k];
        const cpx<T> E{T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y - zc.y)};
        const cpx<T> O{T(0.5) * (zk.y + zc.y), T(-0.5) * (zk.x - zc.x)};
        const cpx<T> wo = cx_mul(twz[k], O);
        X = cpx<T>{E.x + wo.x, E.y + wo.y};
      }
      dst[k] = X;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ theta forward
template <typename T, int TL, int NTHR, typename PL>
__global__ __launch_bounds__(NTHR) void theta_fwd_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                     const cpx<T>* __restrict__ Y, size_t ws,
                                                     const double* __restrict__ D,
                                                     const double2* __restrict__ u,
                                                     double2* __restrict__ N) {
  constexpr int NTP = PL::L;
  __shared__ cpx<T> buf[NTP * TL];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int i = lb % g.nr, l0 = (lb / g.nr) * TL;
  const int tid = threadIdx.x;
  const int nt = g.nt, half = nt / 2, nz = g.nz;
  const double norm = double(NTP) * double(g.nzf);  // mod_fft.cu:41
  for (int c = 0; c < 3; c++) {
    const cpx<T>* src = Y + c * ws + (size_t)i * NTP * nz;
    for (int e = tid; e < NTP * TL; e += NTHR) {
      const int j = e / TL, tl = e - j * TL, l = l0 + tl;
      cpx<T> v{T(0), T(0)};
      if (l < nz - 1) v = src[(size_t)j * nz + l];
      buf[PL::digitrev(j) * TL + tl] = v;
    }
    __syncthreads();
    fft_run<T, TL, NTHR, -1>(PL{}, buf, tw, tid);
    for (int e = tid; e < nt * TL; e += NTHR) {
      const int it = e / TL, tl = e - it * TL, l = l0 + tl;
      if (l < nz) {
        const int j = it < half ? it : it + half;  // padBackward_k row map
        const cpx<T> w = buf[j * TL + tl];
        double vx = double(w.x) / norm, vy = double(w.y) / norm;
        if (it == half || l == nz - 1) vx = vy = 0.0;
        if (c == 2 && it == 0 && l == 0) {  // pressGrad_k (mod_radialFD.cu:100-111)
          double du = 0.0;
          for (int j2 = 0; j2 < IW + 1; j2++)
            du += D[STEN * (g.nr - 1) + j2 + IW] * u[2 * g.n + (size_t)(g.nr - 1 - IW + j2) * g.modes].x;
          vx -= (2.0 / g.re) * du;
        }
        N[c * g.n + ((size_t)i * nt + it) * nz + l] = make_double2(vx, vy);
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ plans
// Compiled plans: theta lengths NTP and half z lengths NZF/2 of the
// supported grids.
#define NSP_THETA_PLANS(X) \
  X(96, Plan<6, 16>)       \
  X(144, Plan<9, 16>)      \
  X(192, Plan<12, 16>)     \
  X(288, Plan<2, 9, 16>)   \
  X(576, Plan<4, 9, 16>)

#define NSP_ZHALF_PLANS(X)  \
  X(96, Plan<6, 16>)        \
  X(576, Plan<4, 9, 16>)    \
  X(768, Plan<3, 16, 16>)   \
  X(1536, Plan<6, 16, 16>)  \
  X(2304, Plan<9, 16, 16>)

// ------------------------------------------------------------------ stage launchers
// Return false when the requested variant is not compiled for this size.
template <typename T>
bool ct_theta_inv(const NlVariant& v, const Geom& g, const cpx<T>* tw, const double* r, const double* Dc,
                  const double2* u, const double2* dR, cpx<T>* W, size_t ws, hipStream_t s) {
  bool ok = false;
#define X(LL, ...)                                                                                      \
  if (g.ntp == LL) {                                                                                    \
    const int tl = v.tl_inv;                                                                            \
    if (LL * tl * (int)sizeof(cpx<T>) <= 163840) {                                                      \
      const int nb = ((g.nz + tl - 1) / tl) * g.nr;                                                     \
      if (tl == 4 && v.fused) { theta_inv_ct<T, 4, 256, 1, __VA_ARGS__><<<nb, 256, 0, s>>>(g, tw, r, Dc, u, dR, W, ws); ok = true; } \
      if (tl == 4 && !v.fused) { theta_inv_ct<T, 4, 256, 0, __VA_ARGS__><<<nb, 256, 0, s>>>(g, tw, r, Dc, u, dR, W, ws); ok = true; } \
      if (tl == 8 && v.fused) { theta_inv_ct<T, 8, 256, 1, __VA_ARGS__><<<nb, 256, 0, s>>>(g, tw, r, Dc, u, dR, W, ws); ok = true; } \
      if (tl == 8 && !v.fused) { theta_inv_ct<T, 8, 256, 0, __VA_ARGS__><<<nb, 256, 0, s>>>(g, tw, r, Dc, u, dR, W, ws); ok = true; } \
    }                                                                                                   \
  }
  NSP_THETA_PLANS(X)
#undef X
  return ok;
}

template <typename T>
bool ct_theta_fwd(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* Y, size_t ws, const double* D,
                  const double2* u, double2* N, hipStream_t s) {
  bool ok = false;
#define X(LL, ...)                                                                                      \
  if (g.ntp == LL) {                                                                                    \
    const int tl = v.tl_fwd;                                                                            \
    if (LL * tl * (int)sizeof(cpx<T>) <= 163840) {                                                      \
      const int nb = ((g.nz + tl - 1) / tl) * g.nr;                                                     \
      if (tl == 4) { theta_fwd_ct<T, 4, 256, __VA_ARGS__><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N); ok = true; } \
      if (tl == 8) { theta_fwd_ct<T, 8, 256, __VA_ARGS__><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N); ok = true; } \
    }                                                                                                   \
  }
  NSP_THETA_PLANS(X)
#undef X
  return ok;
}

template <typename T>
bool ct_zpass(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* twz, const double* r,
              const cpx<T>* W, size_t ws, cpx<T>* Y, hipStream_t s) {
  bool ok = false;
#define X(LL, ...)                                                                                      \
  if (g.nh == LL) {                                                                                     \
    if (v.z_nthr == 256) {                                                                              \
      zpass_ct<T, 256, (2 * LL + 255) / 256, __VA_ARGS__><<<g.nr * g.ntp, 256, 0, s>>>(g, tw, twz, r, W, ws, Y); \
      ok = true;                                                                                        \
    } else if (v.z_nthr == 512) {                                                                       \
      zpass_ct<T, 512, (2 * LL + 511) / 512, __VA_ARGS__><<<g.nr * g.ntp, 512, 0, s>>>(g, tw, twz, r, W, ws, Y); \
      ok = true;                                                                                        \
    }                                                                                                   \
  }
  NSP_ZHALF_PLANS(X)
#undef X
  return ok;
}

}  // namespace nsp
