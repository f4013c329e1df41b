// Nonlinear-term kernels with compile-time transform plans (see fft_ct.h).
// Same dataflow and arithmetic as nonlinear.hip (theta inverse -> z pass ->
// theta forward), with the transform precision T as a template parameter:
// T = double for the FP64 form, float for the FP32 transforms of the mixed
// form.  Lengths without a compiled plan use the generic kernels.
#include "fft_ct.h"

namespace nsp {

__device__ __forceinline__ int m_of_ct(int it, int nt) { return it < nt / 2 ? it : it - nt; }

// ------------------------------------------------------------------ theta inverse
// u, d_r u: FP64 state; W: 9 intermediates [f][NR][NTP][NZ] in precision T.
template <typename T, int TL, int NTHR, typename PL>
__global__ __launch_bounds__(NTHR) void theta_inv_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                     const double* __restrict__ r,
                                                     const double2* __restrict__ u,
                                                     const double2* __restrict__ dR,
                                                     cpx<T>* __restrict__ W, size_t ws) {
  constexpr int NTP = PL::L;
  __shared__ cpx<T> buf[NTP * TL];
  const int l0 = blockIdx.x * TL, i = blockIdx.y;
  const int tid = threadIdx.x;
  const double ri = r[i];
  const int nt = g.nt, half = nt / 2, nz = g.nz;
  for (int f = 0; f < 9; f++) {
    const int c = f % 3, kind = f / 3;  // 0: u, 1: d_r u, 2: d_theta u / r
    const double2* src = (kind == 1 ? dR : u) + c * g.n + (size_t)i * g.modes;
    // padding rows (NT/2 <= j < NT) are zero; every other slot is written below
    for (int e = tid; e < (NTP - nt) * TL; e += NTHR) {
      const int j = half + e / TL, tl = e % TL;
      buf[PL::digitrev(j) * TL + tl] = cpx<T>{T(0), T(0)};
    }
    for (int e = tid; e < nt * TL; e += NTHR) {
      const int it = e / TL, tl = e - it * TL, l = l0 + tl;
      cpx<T> w{T(0), T(0)};
      if (l < nz) {
        const double2 v = src[(size_t)it * nz + l];
        if (kind == 2) {  // derivt_k (mod_deriv.cu:10-21)
          const double kt = double(m_of_ct(it, nt));
          w = cpx<T>{T(-kt * v.y / ri), T(kt * v.x / ri)};
        } else {
          w = cpx<T>{T(v.x), T(v.y)};
        }
      }
      const int j = it < half ? it : it + half;  // padForward_k row map
      buf[PL::digitrev(j) * TL + tl] = w;
    }
    __syncthreads();
    fft_run<T, TL, NTHR, +1>(PL{}, buf, tw, tid);
    cpx<T>* dst = W + f * ws + (size_t)i * NTP * nz;
    for (int e = tid; e < NTP * TL; e += NTHR) {
      const int j = e / TL, tl = e - j * TL, l = l0 + tl;
      if (l < nz) dst[(size_t)j * nz + l] = buf[e];
    }
    __syncthreads();
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
  const int l0 = blockIdx.x * TL, i = blockIdx.y;
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

// ------------------------------------------------------------------ dispatch
// Compiled plans: theta lengths NTP and half z lengths NZF/2 of the
// supported grids.  TL columns per theta block keep LDS at <= 64 KiB.
#define NSP_THETA_PLANS(X) \
  X(96, Plan<6, 16>)       \
  X(144, Plan<9, 16>)      \
  X(192, Plan<12, 16>)     \
  X(288, Plan<2, 9, 16>)   \
  X(576, Plan<4, 9, 16>)   \
  X(1152, Plan<8, 9, 16>)

#define NSP_ZHALF_PLANS(X)  \
  X(96, Plan<6, 16>)        \
  X(576, Plan<4, 9, 16>)    \
  X(768, Plan<3, 16, 16>)   \
  X(1536, Plan<6, 16, 16>)  \
  X(2304, Plan<9, 16, 16>)

bool ct_has_theta(int L) {
#define X(LL, ...) if (L == LL) return true;
  NSP_THETA_PLANS(X)
#undef X
  return false;
}
bool ct_has_zhalf(int L) {
#define X(LL, ...) if (L == LL) return true;
  NSP_ZHALF_PLANS(X)
#undef X
  return false;
}
// radix list of a compiled plan (host side, for twiddle tables)
std::vector<int> ct_radices(int L, bool theta) {
#define X(LL, ...)                                                                 \
  if (L == LL) {                                                                   \
    using PP = __VA_ARGS__;                                                        \
    return std::vector<int>(PP::rad, PP::rad + PP::P);                             \
  }
  if (theta) {
    NSP_THETA_PLANS(X)
  } else {
    NSP_ZHALF_PLANS(X)
  }
#undef X
  return {};
}

template <int NTP>
constexpr int theta_tl(int bytes) {
  // largest power of two TL <= 16 with NTP*TL*bytes <= 64 KiB
  return (NTP * 16 * bytes <= 65536) ? 16 : (NTP * 8 * bytes <= 65536) ? 8 : (NTP * 4 * bytes <= 65536) ? 4 : 2;
}

template <typename T>
bool launch_nonlinear_ct(const Geom& g, const DevTables& t, const double2* u, const double2* dR, cpx<T>* W,
                         cpx<T>* Y, size_t ws, const CtTables<T>& tb, double2* N, hipStream_t s, Profiler* pf) {
  if (!ct_has_theta(g.ntp) || !ct_has_zhalf(g.nh)) return false;
  if (pf) pf->begin(K_THETA_INV, s);
#define X(LL, ...)                                                                                       \
  if (g.ntp == LL) {                                                                                     \
    constexpr int TL = theta_tl<LL>(sizeof(cpx<T>));                                                     \
    dim3 grid((g.nz + TL - 1) / TL, g.nr);                                                               \
    theta_inv_ct<T, TL, 256, __VA_ARGS__><<<grid, 256, 0, s>>>(g, tb.tw_t, t.r, u, dR, W, ws);                   \
  }
  NSP_THETA_PLANS(X)
#undef X
  if (pf) pf->end(s);
  if (pf) pf->begin(K_ZPASS, s);
#define X(LL, ...)                                                                                       \
  if (g.nh == LL) {                                                                                      \
    constexpr int MPT = (2 * LL + 255) / 256;                                                            \
    zpass_ct<T, 256, MPT, __VA_ARGS__><<<g.nr * g.ntp, 256, 0, s>>>(g, tb.tw_h, tb.tw_z, t.r, W, ws, Y);          \
  }
  NSP_ZHALF_PLANS(X)
#undef X
  if (pf) pf->end(s);
  if (pf) pf->begin(K_THETA_FWD, s);
#define X(LL, ...)                                                                                       \
  if (g.ntp == LL) {                                                                                     \
    constexpr int TL = theta_tl<LL>(sizeof(cpx<T>));                                                     \
    dim3 grid((g.nz + TL - 1) / TL, g.nr);                                                               \
    theta_fwd_ct<T, TL, 256, __VA_ARGS__><<<grid, 256, 0, s>>>(g, tb.tw_t, Y, ws, t.D, u, N);                     \
  }
  NSP_THETA_PLANS(X)
#undef X
  if (pf) pf->end(s);
  return true;
}

template bool launch_nonlinear_ct<double>(const Geom&, const DevTables&, const double2*, const double2*,
                                          cpx<double>*, cpx<double>*, size_t, const CtTables<double>&, double2*,
                                          hipStream_t, Profiler*);

}  // namespace nsp
