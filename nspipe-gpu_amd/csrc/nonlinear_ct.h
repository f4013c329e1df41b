// Nonlinear-term kernels with compile-time transform plans (see fft_ct.h).
// Same dataflow and arithmetic as nonlinear.hip (theta inverse -> z pass ->
// theta forward), with the transform precision T as a template parameter:
// T = double for the FP64 form, float for the FP32 transforms of the mixed
// form.  Lengths without a compiled plan use the generic kernels.
#pragma once
#include <type_traits>

#include "fft_ct.h"

namespace nsp {

__device__ __forceinline__ int m_of_ct(int it, int nt) { return it < nt / 2 ? it : it - nt; }

// A copy of x the compiler cannot see through.  Index arithmetic derived from
// it is recomputed where it is used instead of being hoisted out of the
// transform loops and kept live in registers (which cost occupancy).
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  __builtin_assume((unsigned)x < 1024u);
  return x;
}

// Raw buffer loads.  A load whose byte offset lies past num_records returns
// zero instead of touching memory, which gives the zero padding of a row
// beyond its stored modes without clamps or selects.  The resource is the
// gfx9-family layout: base address, stride 0, num_records, dword 3 = 32-bit
// data format without swizzle.  The loads are LLVM's buffer-load intrinsics
// bound by name (this compiler's b64 load builtin emits a 32-bit load).
typedef int nsp_i4v __attribute__((ext_vector_type(4)));
typedef float nsp_f2v __attribute__((ext_vector_type(2)));
typedef double nsp_d2v __attribute__((ext_vector_type(2)));
__device__ nsp_f2v nsp_raw_load_f32x2(nsp_i4v rsrc, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.buffer.load.v2f32");
__device__ nsp_d2v nsp_raw_load_f64x2(nsp_i4v rsrc, int voffset, int soffset, int aux) __asm(
    "llvm.amdgcn.raw.buffer.load.v2f64");
__device__ __forceinline__ nsp_i4v buf_rsrc(const void* base, int bytes) {
  const unsigned long long a = (unsigned long long)base;
  return nsp_i4v{(int)(unsigned)a, (int)(unsigned)(a >> 32), bytes, 0x00020000};
}
template <typename T>
__device__ __forceinline__ cpx<T> buf_load(nsp_i4v rs, int off) {
  if constexpr (sizeof(T) == 4) {
    const nsp_f2v v = nsp_raw_load_f32x2(rs, off, 0, 0);
    return cpx<T>{v[0], v[1]};
  } else {
    const nsp_d2v v = nsp_raw_load_f64x2(rs, off, 0, 0);
    return cpx<T>{v[0], v[1]};
  }
}

// ------------------------------------------------------------------ block order
// (level i, first column l0) of a theta block, levels counted from g.lv0.
// ord 0: level fastest;
// ord 1: column tile fastest, so the blocks an XCD runs together cover whole
// rows of one level (contiguous HBM pages, write lines filled in one L2).
__device__ __forceinline__ void theta_block(const Geom& g, int TL, int ord, int& i, int& l0) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  if (ord == 0) {
    i = g.lv0 + lb % g.nlv;
    l0 = (lb / g.nlv) * TL;
  } else {
    const int nt = (g.nz + TL - 1) / TL;
    i = g.lv0 + lb / nt;
    l0 = (lb % nt) * TL;
  }
}

// ------------------------------------------------------------------ theta inverse
// W: 9 intermediates [f][NR][NTP][NZ] in precision T: u, d_r u, d_theta u / r.
// The radial derivative (derivr_k, mod_radialFD.cu:10-30, flags 1,1,0) comes
// from dR (dr3_kernel, stored at precision T); forming it here from seven
// levels was slower.  Per component the transforms run in the order u,
// d_theta u / r (both from the u inputs), d_r u; one set of inputs is held
// in registers at a time (u at precision T, then d_r u), the next set's
// loads issued as soon as the previous one has been scattered into LDS, so
// that they are in flight during the transform passes.  The FP32 form is
// compiled for 4 waves per SIMD (two 512-thread workgroups per CU).
template <typename T, int TL, int NTHR, typename LY, typename PL, int BAT>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) void theta_inv_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                     const double* __restrict__ r,
                                                     const double2* __restrict__ u,
                                                     const cpx<T>* __restrict__ dR,
                                                     cpx<T>* __restrict__ W, size_t ws, int ord,
                                                     const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    r = gmem(tab, m.r);
    u = gmem(tab, m.u);
    dR = gmem(tab, reinterpret_cast<const cpx<T>*>(m.dR));
    W = gmem(tab, reinterpret_cast<cpx<T>*>(m.W));
    ws = m.ws;
  }
  if (step_done(g)) return;
  constexpr int NTP = PL::L;
  static_assert(64 % TL == 0, "a column tile lies in one 64-column tile of W, Y");
  constexpr int NT = 2 * NTP / 3;
  constexpr int EPT = (NT * TL + NTHR - 1) / NTHR;
  __shared__ cpx<T> buf[LY::size(NTP * TL)];
  // the pass twiddles (NTP - 1 entries) copied into LDS where that keeps the
  // workgroups per CU that the transform buffer allows (two up to 80 KB)
  constexpr int BUFB = LY::size(NTP * TL) * (int)sizeof(cpx<T>);
  constexpr int TWB = (NTP - 1) * (int)sizeof(cpx<T>);
  constexpr bool TW_LDS = BUFB <= 81920 ? BUFB + TWB <= 81920 : BUFB + TWB <= 163840;
  __shared__ cpx<T> twl[TW_LDS ? NTP - 1 : 1];
  const int nz = g.nz;
  int i, l0;
  theta_block(g, TL, ord, i, l0);
  if (i >= g.lv0 + g.nlv || l0 >= nz) return;  // beyond this member's (or level chunk's) extent
  const int tid = threadIdx.x;
  if constexpr (TW_LDS) {  // read after the barrier that follows the first scatter
    for (int e = tid; e < NTP - 1; e += NTHR) twl[e] = tw[e];
  }
  const double ri = r[i];
  // m u / r: the FP32 form multiplies by the reciprocal (the result is rounded
  // to FP32 anyway), the FP64 form divides as derivt_k does
  const double iri = 1.0 / ri;
  auto m_over_r = [&](double kt, double x) -> double {
    if constexpr (sizeof(T) == 4) return kt * x * iri;
    else return kt * x / ri;
  };
  constexpr int half = NT / 2;
  // digit-reversed LDS positions of element e of the inputs (padForward_k
  // row map included) and of element e of the zero padding rows; computed
  // where used (unsigned arithmetic, constant divisors) rather than held in
  // registers across the transforms
  auto pos = [&](int e) -> int {
    const unsigned it = (unsigned)e / TL, tl = (unsigned)e - it * TL;
    const unsigned j = it < (unsigned)half ? it : it + half;
    return LY::pos(PL::digitrev((int)j) * TL + (int)tl);
  };
  auto zpos = [&](int e) -> int {
    const unsigned it = (unsigned)e / TL, tl = (unsigned)e - it * TL;
    return LY::pos(PL::digitrev((int)(half + it)) * TL + (int)tl);
  };
  constexpr int EPZ = ((NTP - NT) * TL + NTHR - 1) / NTHR;  // padding-row elements per thread
  // A thread's elements share its column tl = tid % TL (NTHR is a multiple
  // of TL) and step by NTHR / TL rows: one column test, affine offsets, and
  // the loads of a component issued together without per-element branches.
  static_assert(NTHR % TL == 0, "a thread keeps one column");
  constexpr int FULL = (NT * TL) / NTHR;  // elements every thread has
  const int tcol = tid % TL;
  const bool lok = l0 + tcol < nz;
  const size_t o0 = (size_t)i * g.modes + (size_t)(tid / TL) * nz + (lok ? l0 + tcol : 0);
  const size_t ostep = (size_t)(NTHR / TL) * nz;
  // the same elements of the tile-major d_r u
  const size_t d0 = (size_t)i * dr_plane(g) + (size_t)(tid / TL) * 64 + dr_col(g, lok ? l0 + tcol : 0);
  constexpr int dstep = (NTHR / TL) * 64;
  // inputs of the current transforms: u of component c (rounded to T, which
  // is how every transform of the FP32 form takes them), or d_r u
  cpx<T> X[EPT];
  auto load = [&](int c, bool dr) {
    const double2* src = u + c * g.n;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const bool in = q < FULL || tid + q * NTHR < NT * TL;
      if (dr) {
        X[q] = dR[c * dr_cs(g) + (in ? d0 + q * dstep : d0)];
      } else {
        const double2 v = src[in ? o0 + q * ostep : o0];
        X[q] = cpx<T>{T(v.x), T(v.y)};
      }
      if ((q >= FULL && !in) || !lok) X[q] = cpx<T>{T(0), T(0)};  // rows past NT, columns past NZ
    }
  };
  load(0, false);
  for (int c = 0; c < 3; c++) {
#pragma unroll 1
    for (int kk = 0; kk < 3; kk++) {
      const int kind = kk == 0 ? 0 : (kk == 1 ? 2 : 1);  // 0: u, 1: d_r u, 2: d_theta u / r
      const int lt = opaque(tid);
#pragma unroll
      for (int q = 0; q < EPZ; q++) {  // padding rows
        const int e = lt + q * NTHR;
        if (e < (NTP - NT) * TL) buf[zpos(e)] = cpx<T>{T(0), T(0)};
      }
#pragma unroll
      for (int q = 0; q < EPT; q++) {
        const int e = lt + q * NTHR;
        if (q < FULL || e < NT * TL) {
          const int it = (int)((unsigned)e / TL);
          cpx<T> w = X[q];
          if (kind == 2) {  // derivt_k (mod_deriv.cu:10-21)
            const double kt = double(it < half ? it : it - NT);
            w = cpx<T>{T(m_over_r(-kt, double(X[q].y))), T(m_over_r(kt, double(X[q].x)))};
          }
          buf[pos(e)] = w;
        }
      }
      if (kind == 2) load(c, true);
      else if (kind == 1 && c < 2) load(c + 1, false);
      __syncthreads();
      fft_run_tw<T, TL, NTHR, +1, LY, TW_LDS ? NTP - 1 : 0>(PL{}, buf, tw, twl, lt);
      // rows lt / TL + q NTHR / TL of column tcol (tile-major W: rows of a
      // tile are 64 values apart): one pointer step per store
      constexpr int WFULL = (NTP * TL) / NTHR, WREM = (NTP * TL) % NTHR;
      constexpr int wstep = (NTHR / TL) * 64;
      if (lok) {
        cpx<T>* d = W + (kind * 3 + c) * ws + (size_t)i * tp_plane(g) + (size_t)(lt / TL) * 64 + tp_col(g, l0 + tcol);
#pragma unroll
        for (int q = 0; q < WFULL; q++) d[q * wstep] = buf[LY::pos(lt + q * NTHR)];
        if (WREM && lt < WREM) d[WFULL * wstep] = buf[LY::pos(lt + WFULL * NTHR)];
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ z pass
template <typename T, int NTHR, int MPT, typename LY, typename PL, int BAT>
__global__ __launch_bounds__(NTHR) void zpass_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                 const cpx<T>* __restrict__ twz,
                                                 const double* __restrict__ r,
                                                 const cpx<T>* __restrict__ W, size_t ws,
                                                 cpx<T>* __restrict__ Y, const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    r = gmem(tab, m.r);
    W = gmem(tab, reinterpret_cast<const cpx<T>*>(m.W));
    Y = gmem(tab, reinterpret_cast<cpx<T>*>(m.Y));
    ws = m.ws;
  }
  if (step_done(g)) return;
  constexpr int NH = PL::L;
  constexpr int NZF = 2 * NH;
  __shared__ cpx<T> buf[LY::size(NH)];
  const int row = g.lv0 * g.ntp + (int)blockIdx.x;
  if (row >= (g.lv0 + g.nlv) * g.ntp) return;  // beyond this member's (or level chunk's) extent
  const int i = row / g.ntp;
  const int tid = threadIdx.x;
  const int nz = g.nz;
  const T rad = T(r[i]);
  // x / r; the FP32 form multiplies by the reciprocal (within an ulp, without
  // the IEEE division sequence per sample)
  const T irad = T(1) / rad;
  auto over_r = [&](T x) -> T {
    if constexpr (sizeof(T) == 4) return x * irad;
    else return x / rad;
  };
  // row th of level i in the tile-major planes of W, Y: value k at roff + tp_col(k)
  const size_t roff = (size_t)i * tp_plane(g) + (size_t)(row - i * g.ntp) * 64;
  const T kz0 = T(g.kz0);

  auto c2r = [&](const cpx<T>* X, bool dz) {
    const int lt = opaque(tid);
    for (int k = lt; k < NH; k += NTHR) {
      const int kc = NH - k;
      cpx<T> a{T(0), T(0)}, b{T(0), T(0)};
      if (k < nz) a = X[tp_col(g, k)];
      if (kc < nz) b = X[tp_col(g, kc)];
      if (k == 0) { a.y = T(0); b.y = T(0); }
      if (dz) {  // derivz_k (mod_deriv.cu:24-35)
        const T ka = kz0 * T(k), kb = kz0 * T(kc);
        a = cpx<T>{-ka * a.y, ka * a.x};
        b = cpx<T>{-kb * b.y, kb * b.x};
      }
      const cpx<T> E{a.x + b.x, a.y - b.y};
      const cpx<T> dl{a.x - b.x, a.y + b.y};
      const cpx<T> O = cx_mulc(dl, twz[k]);
      buf[LY::pos(PL::digitrev(k))] = cpx<T>{E.x - O.y, E.y + O.x};
    }
    __syncthreads();
    fft_run<T, 1, NTHR, +1, LY>(PL{}, buf, tw, lt);
  };
  auto sample = [&](int n) -> T {  // n from an opaque thread index
    const cpx<T> v = buf[LY::pos(n >> 1)];
    return (n & 1) ? v.y : v.x;
  };

  T ur[MPT], ut[MPT], uz[MPT], acc[MPT];
  c2r(W + 0 * ws + roff, false);
  int lt = opaque(tid);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = lt + q * NTHR; ur[q] = n < NZF ? sample(n) : T(0); }
  __syncthreads();
  c2r(W + 1 * ws + roff, false);
  lt = opaque(tid);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = lt + q * NTHR; ut[q] = n < NZF ? sample(n) : T(0); }
  __syncthreads();
  c2r(W + 2 * ws + roff, false);
  lt = opaque(tid);
#pragma unroll
  for (int q = 0; q < MPT; q++) { const int n = lt + q * NTHR; uz[q] = n < NZF ? sample(n) : T(0); }
  __syncthreads();

#pragma unroll 1
  for (int c = 0; c < 3; c++) {
    // calcNonlinear_k (mod_io_nonlinear.cu:142-170), summed in its order
    c2r(W + (3 + c) * ws + roff, false);
    lt = opaque(tid);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = lt + q * NTHR; acc[q] = n < NZF ? ur[q] * sample(n) : T(0); }
    __syncthreads();
    c2r(W + (6 + c) * ws + roff, false);
    lt = opaque(tid);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = lt + q * NTHR; if (n < NZF) acc[q] += ut[q] * sample(n); }
    __syncthreads();
    c2r(W + c * ws + roff, true);
    lt = opaque(tid);
#pragma unroll
    for (int q = 0; q < MPT; q++) { const int n = lt + q * NTHR; if (n < NZF) acc[q] += uz[q] * sample(n); }
    __syncthreads();
    lt = opaque(tid);
#pragma unroll
    for (int q = 0; q < MPT; q++) {
      const int n = lt + q * NTHR;
      if (n < NZF) {
        T v = acc[q];
        if (c == 0) v = v - ut[q] * over_r(ut[q]);
        else if (c == 1) v = v + ur[q] * over_r(ut[q]);
        T* slot = reinterpret_cast<T*>(&buf[LY::pos(PL::digitrev(n >> 1))]) + (n & 1);
        *slot = -v;
      }
    }
    __syncthreads();
    fft_run<T, 1, NTHR, -1, LY>(PL{}, buf, tw, lt);
    cpx<T>* dst = Y + c * ws + roff;
    for (int k = lt; k < nz; k += NTHR) {
      cpx<T> X{T(0), T(0)};
      if (k < nz - 1) {
        const cpx<T> zk = buf[LY::pos(k)], zc = buf[LY::pos(k == 0 ? 0 : NH - k)];
        const cpx<T> E{T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y - zc.y)};
        const cpx<T> O{T(0.5) * (zk.y + zc.y), T(-0.5) * (zk.x - zc.x)};
        const cpx<T> wo = cx_mul(twz[k], O);
        X = cpx<T>{E.x + wo.x, E.y + wo.y};
      }
      dst[tp_col(g, k)] = X;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ z pass, register form
// Same transforms, products and arithmetic as zpass_ct, arranged so that no
// data passes through LDS in digit-reversed order:
//  * thread n0 < NH/R0 owns the complex samples k = n0 + (NH/R0) q, q < R0
//    (real samples 2k, 2k+1) of every transform of the row;
//  * the half-length C2R pre-processing of those k reads X[k] and X[NH-k]
//    straight from global memory (coalesced) and feeds the first radix-R0
//    pass in registers (first_pass_store); later passes run through LDS;
//  * the transform output is read back at the owned k, so the convective
//    products stay in registers, and they are exactly the first-pass inputs
//    of the forward transform of the row;
//  * with PF the 2 R0 inputs of the next transform are loaded before the
//    current one's LDS passes start, overlapping HBM latency with them.
// LDS rows of NH/R_last elements padded by PAD slots (bank spreading).
template <typename T, int NTHR, int PF, int PAD, typename PL, int BAT>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1))) void zpass2_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                  const cpx<T>* __restrict__ twz,
                                                  const double* __restrict__ r,
                                                  const cpx<T>* __restrict__ W, size_t ws,
                                                  cpx<T>* __restrict__ Y, const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    r = gmem(tab, m.r);
    W = gmem(tab, reinterpret_cast<const cpx<T>*>(m.W));
    Y = gmem(tab, reinterpret_cast<cpx<T>*>(m.Y));
    ws = m.ws;
  }
  if (step_done(g)) return;
  constexpr int NH = PL::L;
  constexpr int R0 = PL::rad[0];
  constexpr int NB1 = NH / R0;
  static_assert(NB1 <= NTHR, "one first-pass butterfly per thread");
  using LY = Lay<NH / PL::rad[PL::P - 1], PAD>;
  __shared__ cpx<T> buf[LY::size(NH)];
  const int row = g.lv0 * g.ntp + (int)blockIdx.x;
  if (row >= (g.lv0 + g.nlv) * g.ntp) return;  // beyond this member's (or level chunk's) extent
  const int i = row / g.ntp;
  const int tid = threadIdx.x;
  const bool own = NB1 == NTHR || tid < NB1;  // thread owns a first-pass butterfly
  const int nz = g.nz;
  const T rad = T(r[i]);
  // x / r; the FP32 form multiplies by the reciprocal (within an ulp, without
  // the IEEE division sequence per sample)
  const T irad = T(1) / rad;
  auto over_r = [&](T x) -> T {
    if constexpr (sizeof(T) == 4) return x * irad;
    else return x / rad;
  };
  // row th of level i in the tile-major planes of W, Y: value k at roff + tp_col(k)
  const int th = row - i * g.ntp;
  const size_t roff = (size_t)i * tp_plane(g) + (size_t)th * 64;
  // buffer range: from the row's first value to the end of its plane, past
  // every stored value of the row; modes past NZ load from an offset beyond
  // it, which the range check turns into zero
  const int rbytes = (int)((tp_plane(g) - (size_t)th * 64) * sizeof(cpx<T>));
  const int t64 = g.ntp * 64;  // values per 64-column tile
  const T kz0 = T(g.kz0);
  const cpx<T> zero{T(0), T(0)};

  // source field of inverse transform t: u_r, u_t, u_z, then for each
  // component c: d_r u_c, d_theta u_c / r, u_c (differentiated in z)
  auto src = [&](int t) -> const cpx<T>* {
    const int f = t < 3 ? t : ((t - 3) % 3 == 0 ? 3 + (t - 3) / 3 : ((t - 3) % 3 == 1 ? 6 + (t - 3) / 3 : (t - 3) / 3));
    return W + f * ws + roff;
  };
  // X[k] and X[NH-k] of the owned samples, zero beyond the stored NZ modes:
  // raw buffer loads of the row, whose modes past NZ take an offset past the
  // buffer range (OOB) and so load zero without touching memory.
  constexpr int CS = (int)sizeof(cpx<T>);
  constexpr int OOB = 0x7ffffff0;
  auto load = [&](const cpx<T>* X, cpx<T>* A, cpx<T>* B) {
    const int lt = opaque(tid);
    if (NB1 == NTHR || lt < NB1) {
      const nsp_i4v rs = buf_rsrc(X, rbytes);
      if constexpr (NB1 % 64 == 0 && NH % 64 == 0) {
        // k = lt + NB1 q and NH - k step by whole tiles: affine offsets
        const int ok0 = ((lt >> 6) * t64 + (lt & 63)) * CS;
        const int oc0 = (((NH >> 6) - ((lt + 63) >> 6)) * t64 + ((-lt) & 63)) * CS;
        const int dq = (NB1 / 64) * t64 * CS;
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int k = lt + NB1 * q, kc = NH - k;
          A[q] = buf_load<T>(rs, k < nz ? ok0 + q * dq : OOB);
          B[q] = buf_load<T>(rs, kc < nz ? oc0 - q * dq : OOB);
        }
      } else {
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int k = lt + NB1 * q, kc = NH - k;
          A[q] = buf_load<T>(rs, k < nz ? (int)tp_col(g, k) * CS : OOB);
          B[q] = buf_load<T>(rs, kc < nz ? (int)tp_col(g, kc) * CS : OOB);
        }
      }
    }
  };

  cpx<T> ur[R0], ut[R0], uz[R0];
  cpx<T> A[R0], B[R0];
  // split twiddles and LDS sample positions of the owned samples, the same
  // for all 12 transforms: kept in registers when that costs no occupancy
  // (the R0 = 9 plans already run at 2 waves per SIMD), else recomputed
  constexpr bool HOIST = R0 >= 9;
  cpx<T> wz[HOIST ? R0 : 1];
  int spos[HOIST ? R0 : 1];
  if constexpr (HOIST) {
#pragma unroll
    for (int q = 0; q < R0; q++) {
      wz[q] = own ? twz[tid + NB1 * q] : zero;
      spos[q] = LY::pos(tid + NB1 * q);
    }
  }
  // Inverse transform t of the inputs in A, B (DZ: differentiated in z);
  // leaves the physical values at the owned samples in sm.  With PF the
  // inputs of transform t + 1 are loaded into A, B meanwhile.  Every call
  // site is its own inlined copy, with DZ and the consumer of sm fixed, so
  // the per-transform branches and register copies of a loop over t vanish.
  auto inverse = [&](int t, auto dz_tag, cpx<T>(&sm)[R0]) {
    constexpr bool DZ = decltype(dz_tag)::value;
    const int lt = opaque(tid);
    cpx<T> nA[R0], nB[R0];
    if constexpr (PF) {
      if (t + 1 < 12) load(src(t + 1), nA, nB);
    }
    if (own) {
      cpx<T> v[R0];
#pragma unroll
      for (int q = 0; q < R0; q++) {  // half-length C2R pre-processing
        const int k = lt + NB1 * q, kc = NH - k;
        cpx<T> a = A[q], b = B[q];
        if (k == 0) { a.y = T(0); b.y = T(0); }
        if constexpr (DZ) {  // derivz_k (mod_deriv.cu:24-35)
          const T ka = kz0 * T(k), kb = kz0 * T(kc);
          a = cpx<T>{-ka * a.y, ka * a.x};
          b = cpx<T>{-kb * b.y, kb * b.x};
        }
        const cpx<T> E{a.x + b.x, a.y - b.y};
        const cpx<T> dl{a.x - b.x, a.y + b.y};
        cpx<T> w;
        if constexpr (HOIST) w = wz[q];
        else w = twz[k];
        const cpx<T> O = cx_mulc(dl, w);
        v[q] = cpx<T>{E.x - O.y, E.y + O.x};
      }
      first_pass_store<T, +1, LY, 1>(PL{}, v, buf, lt);
    }
    __syncthreads();
    fft_rest<T, NTHR, +1, LY>(PL{}, buf, tw, lt);
#pragma unroll
    for (int q = 0; q < R0; q++) {
      sm[q] = zero;
      if (own) {
        if constexpr (HOIST) sm[q] = buf[spos[q]];
        else sm[q] = buf[LY::pos(lt + NB1 * q)];
      }
    }
    __syncthreads();
    if constexpr (PF) {
      if (t + 1 < 12) {
#pragma unroll
        for (int q = 0; q < R0; q++) { A[q] = nA[q]; B[q] = nB[q]; }
      }
    } else {
      if (t + 1 < 12) load(src(t + 1), A, B);
    }
  };
  using NoDz = std::integral_constant<bool, false>;
  using Dz = std::integral_constant<bool, true>;

  load(src(0), A, B);
  // velocities: after three transforms ur, ut, uz hold transforms 0, 1, 2
#pragma unroll
  for (int q = 0; q < R0; q++) ut[q] = uz[q] = zero;
#pragma unroll 1
  for (int t = 0; t < 3; t++) {
    cpx<T> sm[R0];
    inverse(t, NoDz{}, sm);
#pragma unroll
    for (int q = 0; q < R0; q++) {
      ur[q] = ut[q];
      ut[q] = uz[q];
      uz[q] = sm[q];
    }
  }
  // calcNonlinear_k (mod_io_nonlinear.cu:142-170), summed in its order, per
  // component c: ur d_r u_c + ut d_theta u_c / r + uz d_z u_c (+ curvature)
#pragma unroll 1
  for (int c = 0; c < 3; c++) {
    const int t0 = 3 + 3 * c;
    cpx<T> acc[R0], sm[R0];
    inverse(t0, NoDz{}, sm);
#pragma unroll
    for (int q = 0; q < R0; q++) acc[q] = cpx<T>{ur[q].x * sm[q].x, ur[q].y * sm[q].y};
    inverse(t0 + 1, NoDz{}, sm);
#pragma unroll
    for (int q = 0; q < R0; q++) {
      acc[q].x += ut[q].x * sm[q].x;
      acc[q].y += ut[q].y * sm[q].y;
    }
    inverse(t0 + 2, Dz{}, sm);
    if (own) {
      cpx<T> v[R0];
#pragma unroll
      for (int q = 0; q < R0; q++) {
        T vx = acc[q].x + uz[q].x * sm[q].x, vy = acc[q].y + uz[q].y * sm[q].y;
        if (c == 0) {
          vx = vx - ut[q].x * over_r(ut[q].x);
          vy = vy - ut[q].y * over_r(ut[q].y);
        } else if (c == 1) {
          vx = vx + ur[q].x * over_r(ut[q].x);
          vy = vy + ur[q].y * over_r(ut[q].y);
        }
        v[q] = cpx<T>{-vx, -vy};
      }
      first_pass_store<T, -1, LY, 1>(PL{}, v, buf, opaque(tid));
    }
    __syncthreads();
    const int lt = opaque(tid);
    fft_rest<T, NTHR, -1, LY>(PL{}, buf, tw, lt);
    cpx<T>* dst = Y + c * ws + roff;
    for (int k = lt; k < nz; k += NTHR) {  // half-length R2C post-processing
      cpx<T> X = zero;
      if (k < nz - 1) {
        const cpx<T> zk = buf[LY::pos(k)], zc = buf[LY::pos(k == 0 ? 0 : NH - k)];
        const cpx<T> E{T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y - zc.y)};
        const cpx<T> O{T(0.5) * (zk.y + zc.y), T(-0.5) * (zk.x - zc.x)};
        const cpx<T> wo = cx_mul(twz[k], O);
        X = cpx<T>{E.x + wo.x, E.y + wo.y};
      }
      dst[tp_col(g, k)] = X;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ z pass, transforms in triples
// The register form with the twelve inverse transforms of a row taken three
// at a time through one LDS buffer of three interleaved columns (element n of
// column c at logical index c + 3n): the velocities (u_r, u_theta, u_z), then
// for each component c the three fields of its convective product
// (d_r u_c, d_theta u_c / r, d_z u_c).  A pass over the three columns spreads
// their 3 L/R butterflies over the threads: the radix-16 passes of 2304
// (144 butterflies) fill 432 of 512 thread slots in two rounds instead of 144
// of 256 in one, and a triple takes one barrier per pass instead of three.
// A component's product needs exactly its triple (no partial sums are kept
// between transforms).  The forward transform of each product runs alone, in
// the single-column layout LY of the same buffer.  Arithmetic per transform
// and per sample is that of zpass2_ct.
template <typename T, int NTHR, int PAD, int PAD3, typename PL, int BAT>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1))) void zpass3_ct(
    Geom g, const cpx<T>* __restrict__ tw, const cpx<T>* __restrict__ twz, const double* __restrict__ r,
    const cpx<T>* __restrict__ W, size_t ws, cpx<T>* __restrict__ Y, const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    r = gmem(tab, m.r);
    W = gmem(tab, reinterpret_cast<const cpx<T>*>(m.W));
    Y = gmem(tab, reinterpret_cast<cpx<T>*>(m.Y));
    ws = m.ws;
  }
  if (step_done(g)) return;
  constexpr int NH = PL::L;
  constexpr int R0 = PL::rad[0];
  constexpr int NB1 = NH / R0;
  static_assert(NB1 <= NTHR, "one first-pass butterfly per thread");
  constexpr int RL = NH / PL::rad[PL::P - 1];
  using LY = Lay<RL, PAD>;       // one column (forward transforms)
  using LY3 = Lay<3 * RL, PAD3>;  // three interleaved columns (inverse triples)
  constexpr int NBUF = LY3::size(3 * NH) > LY::size(NH) ? LY3::size(3 * NH) : LY::size(NH);
  __shared__ cpx<T> buf[NBUF];
  const int row = g.lv0 * g.ntp + (int)blockIdx.x;
  if (row >= (g.lv0 + g.nlv) * g.ntp) return;  // beyond this member's (or level chunk's) extent
  const int i = row / g.ntp;
  const int tid = threadIdx.x;
  const bool own = NB1 == NTHR || tid < NB1;
  const int nz = g.nz;
  const T rad = T(r[i]);
  const T irad = T(1) / rad;
  auto over_r = [&](T x) -> T {
    if constexpr (sizeof(T) == 4) return x * irad;
    else return x / rad;
  };
  const int th = row - i * g.ntp;
  const size_t roff = (size_t)i * tp_plane(g) + (size_t)th * 64;
  const int rbytes = (int)((tp_plane(g) - (size_t)th * 64) * sizeof(cpx<T>));
  const int t64 = g.ntp * 64;
  const T kz0 = T(g.kz0);
  const cpx<T> zero{T(0), T(0)};
  constexpr int CS = (int)sizeof(cpx<T>);
  constexpr int OOB = 0x7ffffff0;
  // X[k] and X[NH-k] of the owned samples of field f (zero past NZ: range-checked buffer loads)
  auto load = [&](int f, cpx<T>* A, cpx<T>* B) {
    const int lt = opaque(tid);
    if (NB1 == NTHR || lt < NB1) {
      const nsp_i4v rs = buf_rsrc(W + f * ws + roff, rbytes);
      if constexpr (NB1 % 64 == 0 && NH % 64 == 0) {
        const int ok0 = ((lt >> 6) * t64 + (lt & 63)) * CS;
        const int oc0 = (((NH >> 6) - ((lt + 63) >> 6)) * t64 + ((-lt) & 63)) * CS;
        const int dq = (NB1 / 64) * t64 * CS;
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int k = lt + NB1 * q, kc = NH - k;
          A[q] = buf_load<T>(rs, k < nz ? ok0 + q * dq : OOB);
          B[q] = buf_load<T>(rs, kc < nz ? oc0 - q * dq : OOB);
        }
      } else {
#pragma unroll
        for (int q = 0; q < R0; q++) {
          const int k = lt + NB1 * q, kc = NH - k;
          A[q] = buf_load<T>(rs, k < nz ? (int)tp_col(g, k) * CS : OOB);
          B[q] = buf_load<T>(rs, kc < nz ? (int)tp_col(g, kc) * CS : OOB);
        }
      }
    }
  };
  cpx<T> wz[R0];
#pragma unroll
  for (int q = 0; q < R0; q++) wz[q] = own ? twz[tid + NB1 * q] : zero;
  // C2R pre-processing of one column's owned samples and its first pass
  auto first = [&](cpx<T>* A, cpx<T>* B, bool dz, int col) {
    const int lt = opaque(tid);
    if (own) {
      cpx<T> v[R0];
#pragma unroll
      for (int q = 0; q < R0; q++) {
        const int k = lt + NB1 * q, kc = NH - k;
        cpx<T> a = A[q], b = B[q];
        if (k == 0) { a.y = T(0); b.y = T(0); }
        if (dz) {  // derivz_k (mod_deriv.cu:24-35)
          const T ka = kz0 * T(k), kb = kz0 * T(kc);
          a = cpx<T>{-ka * a.y, ka * a.x};
          b = cpx<T>{-kb * b.y, kb * b.x};
        }
        const cpx<T> E{a.x + b.x, a.y - b.y};
        const cpx<T> dl{a.x - b.x, a.y + b.y};
        const cpx<T> O = cx_mulc(dl, wz[q]);
        v[q] = cpx<T>{E.x - O.y, E.y + O.x};
      }
      first_pass_store<T, +1, LY3, 3>(PL{}, v, buf, lt, col);
    }
  };
  // Inverse transforms of fields f0, f1, f2 (the third differentiated in z
  // when DZ2); their physical values at the owned samples land in s0, s1, s2.
  // On entry A, B hold the loads of f0; on exit, those of fnext (if >= 0).
  cpx<T> A[R0], B[R0];
  auto triple = [&](int f0, int f1, int f2, bool dz2, int fnext, cpx<T>(&s0)[R0], cpx<T>(&s1)[R0],
                    cpx<T>(&s2)[R0]) {
    first(A, B, false, 0);
    load(f1, A, B);
    first(A, B, false, 1);
    load(f2, A, B);
    first(A, B, dz2, 2);
    __syncthreads();
    const int lt = opaque(tid);
    fft_rest_cols<T, 3, NTHR, +1, LY3>(PL{}, buf, tw, lt);
#pragma unroll
    for (int q = 0; q < R0; q++) {
      s0[q] = s1[q] = s2[q] = zero;
      if (own) {
        const int e = 3 * (lt + NB1 * q);
        s0[q] = buf[LY3::pos(e)];
        s1[q] = buf[LY3::pos(e + 1)];
        s2[q] = buf[LY3::pos(e + 2)];
      }
    }
    __syncthreads();
    if (fnext >= 0) load(fnext, A, B);
  };
  // field index in W: u_r, u_t, u_z (0-2), d_r u_c (3 + c), d_theta u_c / r (6 + c)
  cpx<T> ur[R0], ut[R0], uz[R0];
  load(0, A, B);
  triple(0, 1, 2, false, 3, ur, ut, uz);
#pragma unroll 1
  for (int c = 0; c < 3; c++) {
    cpx<T> sr[R0], st[R0], sz[R0];
    triple(3 + c, 6 + c, c, true, c < 2 ? 4 + c : -1, sr, st, sz);
    // calcNonlinear_k (mod_io_nonlinear.cu:142-170), summed in its order
    if (own) {
      cpx<T> v[R0];
#pragma unroll
      for (int q = 0; q < R0; q++) {
        T vx = ur[q].x * sr[q].x, vy = ur[q].y * sr[q].y;
        vx += ut[q].x * st[q].x;
        vy += ut[q].y * st[q].y;
        vx = vx + uz[q].x * sz[q].x;
        vy = vy + uz[q].y * sz[q].y;
        if (c == 0) {
          vx = vx - ut[q].x * over_r(ut[q].x);
          vy = vy - ut[q].y * over_r(ut[q].y);
        } else if (c == 1) {
          vx = vx + ur[q].x * over_r(ut[q].x);
          vy = vy + ur[q].y * over_r(ut[q].y);
        }
        v[q] = cpx<T>{-vx, -vy};
      }
      first_pass_store<T, -1, LY, 1>(PL{}, v, buf, opaque(tid));
    }
    __syncthreads();
    const int lt = opaque(tid);
    fft_rest<T, NTHR, -1, LY>(PL{}, buf, tw, lt);
    cpx<T>* dst = Y + c * ws + roff;
    for (int k = lt; k < nz; k += NTHR) {  // half-length R2C post-processing
      cpx<T> X = zero;
      if (k < nz - 1) {
        const cpx<T> zk = buf[LY::pos(k)], zc = buf[LY::pos(k == 0 ? 0 : NH - k)];
        const cpx<T> E{T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y - zc.y)};
        const cpx<T> O{T(0.5) * (zk.y + zc.y), T(-0.5) * (zk.x - zc.x)};
        const cpx<T> wo = cx_mul(twz[k], O);
        X = cpx<T>{E.x + wo.x, E.y + wo.y};
      }
      dst[tp_col(g, k)] = X;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ z pass, pair form
// The real inverse transforms of two fields a, b of a row run as one complex
// transform of the full length NZF: z = a + i b, from Z[k] = A[k] + i B[k]
// with A[NZF - k] = conj A[k] (the stored modes k < NZ, zero up to NZF/2;
// the l = 0 mode enters with its real part only, as in the C2R forms).  The
// real forward transforms of two products likewise, separated afterwards:
// A[k] = (Z[k] + conj Z[NZF - k]) / 2, B[k] = (Z[k] - conj Z[NZF - k]) / 2i.
// No even/odd split and no split twiddles; the LDS buffer holds one
// full-length column.  The thread with first-pass butterfly n0 owns the
// physical samples n0 + (NZF / R0) q, q < R0, of every field (the outputs it
// reads after an inverse transform are the inputs of its first forward
// pass).  A workgroup takes ROWS (1 or 2) rows, one after the other.  Per
// row: six inverse pairs, (u_r, u_t), (u_z, d_z u_r), (d_r u_r,
// d_theta u_r / r), (d_r u_t, d_theta u_t / r), (d_z u_t, d_z u_z), (d_r u_z,
// d_theta u_z / r), and the forward pair (N_r, N_t); N_z goes through a
// forward transform of its own (ROWS 1: 8 complex transforms per row) or
// together with the other row's N_z (ROWS 2: 7.5 per row).  The plan's pass
// twiddles follow the half-length plan's in tw (NZF/2 - 1 entries).  The FP32
// form without prefetch is compiled for 4 waves per SIMD (two 512-thread
// workgroups per CU).
template <typename T, int NTHR, int PAD, int PFE, int ROWS, typename PF, int BAT>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && !PFE && PF::rad[0] <= 9 ? 4 : 1))) void zpass4_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                  const cpx<T>* __restrict__ twz, const double* __restrict__ r,
                                                  const cpx<T>* __restrict__ W, size_t ws,
                                                  cpx<T>* __restrict__ Y, const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    r = gmem(tab, m.r);
    W = gmem(tab, reinterpret_cast<const cpx<T>*>(m.W));
    Y = gmem(tab, reinterpret_cast<cpx<T>*>(m.Y));
    ws = m.ws;
  }
  (void)twz;
  if (step_done(g)) return;
  constexpr int NF = PF::L;
  constexpr int NH = NF / 2;
  constexpr int R0 = PF::rad[0];
  constexpr int NB1 = NF / R0;
  static_assert(NB1 <= NTHR, "one first-pass butterfly per thread");
  using LY = Lay<NF / PF::rad[PF::P - 1], PAD>;
  __shared__ cpx<T> buf[LY::size(NF)];
  // LDS beside the transform buffer, in this order where a workgroup's total
  // stays within its share of a CU's 160 KB (half of it for the FP32 form
  // compiled for two workgroups per CU): the pass twiddles of every pass but
  // the last (TWL entries), the first row's N_z (two-row form), the compact
  // twiddles of a last radix-8 pass (TWN entries) and the load offsets.
  constexpr int CAP = sizeof(T) == 4 && !PFE && R0 <= 9 ? 81920 : 163840;
  constexpr int BUFB = LY::size(NF) * (int)sizeof(cpx<T>);
  constexpr int TWL = NF / PF::rad[PF::P - 1] - 1;
  constexpr bool TW_LDS = BUFB + TWL * (int)sizeof(cpx<T>) <= CAP;
  constexpr int B1 = BUFB + (TW_LDS ? TWL * (int)sizeof(cpx<T>) : 0);
  constexpr bool NZ_LDS = ROWS == 2 && B1 + R0 * NTHR * (int)sizeof(T) <= CAP;
  constexpr int B2 = B1 + (NZ_LDS ? R0 * NTHR * (int)sizeof(T) : 0);
  // a last pass of radix 8 reads the compact table that follows the plan's
  // NF - 1 pass twiddles
  constexpr int TWC = PF::rad[PF::P - 1] == 8 ? 1 : 0;
  constexpr int TWN = TWC ? 3 * (NF / 8) : 0;
  constexpr bool TWC_LDS = TWC && B2 + TWN * (int)sizeof(cpx<T>) <= CAP;
  constexpr int B3 = B2 + (TWC_LDS ? TWN * (int)sizeof(cpx<T>) : 0);
  constexpr bool OFF_LDS = B3 + R0 * NTHR * 4 <= CAP;
  __shared__ int offs[OFF_LDS ? R0 * NTHR : 1];
  __shared__ cpx<T> twl[TW_LDS ? TWL : 1];
  __shared__ cpx<T> twcl[TWC_LDS ? TWN : 1];
  const cpx<T>* __restrict__ twf = tw + (NH - 1);
  const cpx<T>* __restrict__ twc = twf + (NF - 1);
  const int nrows = (g.lv0 + g.nlv) * g.ntp;  // end of this member's (or level chunk's) rows
  const int row0 = g.lv0 * g.ntp + ROWS * (int)blockIdx.x;
  if (row0 >= nrows) return;
  const int nrow = ROWS == 2 && row0 + 1 < nrows ? 2 : 1;
  const int tid = threadIdx.x;
  const bool own = NB1 == NTHR || tid < NB1;
  const int nz = g.nz;
  const T kz0 = T(g.kz0);
  const cpx<T> zero{T(0), T(0)};
  constexpr int CS = (int)sizeof(cpx<T>);
  // the current row: level i, radius, row th of the level's tile-major planes
  T rad = T(1), irad = T(1);
  size_t roff = 0;
  int rbytes = 0;
  auto set_row = [&](int row) {
    const int i = row / g.ntp;
    rad = T(r[i]);
    irad = T(1) / rad;
    const int th = row - i * g.ntp;
    roff = (size_t)i * tp_plane(g) + (size_t)th * 64;
    rbytes = (int)((tp_plane(g) - (size_t)th * 64) * sizeof(cpx<T>));
  };
  // x / r: the FP32 form multiplies by the reciprocal, the FP64 form divides
  auto over_r = [&](T x) -> T {
    if constexpr (sizeof(T) == 4) return x * irad;
    else return x / rad;
  };
  // sample n = n0 + NB1 q reads mode m = n (n <= NZF/2) or NZF - n (conjugated)
  auto conj_of = [](int lt, int q) -> bool {
    if (NB1 * (q + 1) <= NH + 1) return false;  // every lane: n <= NH
    if (NB1 * q > NH) return true;              // every lane: n > NH
    return lt + NB1 * q > NH;
  };
  // the stored modes m of fields fa, fb at the owned samples, zero past NZ:
  // columns nz <= m < 64 ceil(nz / 64) are W's zero padding (W starts zeroed
  // and is written at m < nz only), columns beyond lie past the end of the
  // row's plane, where the buffer range check returns zero.  With NB1 and
  // NZF multiples of 64, the tile-major offsets of m = n and m = NZF - n are
  // affine in q.
  auto offset = [&](int lt, int q) -> int {
    const int n = lt + NB1 * q;
    const int t64 = g.ntp * 64;
    if constexpr (NB1 % 64 == 0 && NF % 64 == 0) {
      const int dq = (NB1 / 64) * t64 * CS;
      const int ok = ((lt >> 6) * t64 + (lt & 63)) * CS + q * dq;
      const int oc = (((NF >> 6) - ((lt + 63) >> 6)) * t64 + ((-lt) & 63)) * CS - q * dq;
      return conj_of(lt, q) ? oc : ok;
    } else {
      const int m = conj_of(lt, q) ? NF - n : n;
      return (int)tp_col(g, m) * CS;
    }
  };
  // (read after the barrier that follows the first pass of the first pair)
  if constexpr (TW_LDS) {
    for (int e = tid; e < TWL; e += NTHR) twl[e] = twf[e];
  }
  if constexpr (TWC_LDS) {
    for (int e = tid; e < TWN; e += NTHR) twcl[e] = twc[e];
  }
  // The offsets are the same for every pair and row of the workgroup: where
  // LDS has room they are computed once and read back per load.
  if constexpr (OFF_LDS) {
    if (own) {
#pragma unroll
      for (int q = 0; q < R0; q++) offs[q * NTHR + tid] = offset(tid, q);
    }
  }
  auto load = [&](int fa, int fb, cpx<T>* A, cpx<T>* B) {
    const int lt = opaque(tid);
    if (own) {
      const nsp_i4v ra = buf_rsrc(W + fa * ws + roff, rbytes);
      const nsp_i4v rb = buf_rsrc(W + fb * ws + roff, rbytes);
#pragma unroll
      for (int q = 0; q < R0; q++) {
        const int o = OFF_LDS ? offs[q * NTHR + lt] : offset(lt, q);
        A[q] = buf_load<T>(ra, o);
        B[q] = buf_load<T>(rb, o);
      }
    }
  };
  // z = a + i b at the owned samples (d_z applied first where asked) and the
  // first pass; sub0: field a enters without its l = 0 mode (added back to
  // the samples by the caller)
  auto first = [&](const cpx<T>* A, const cpx<T>* B, bool dza, bool dzb, bool sub0) {
    const int lt = opaque(tid);
    if (own) {
      cpx<T> v[R0];
#pragma unroll
      for (int q = 0; q < R0; q++) {
        const int n = lt + NB1 * q;
        const bool cj = conj_of(lt, q);
        cpx<T> a = A[q], b = B[q];
        if (q == 0 && n == 0) {
          a = cpx<T>{sub0 ? T(0) : a.x, T(0)};
          b.y = T(0);
        }
        if (dza || dzb) {  // derivz_k (mod_deriv.cu:24-35) on the stored mode
          const T km = kz0 * T(cj ? NF - n : n);
          if (dza) a = cpx<T>{-km * a.y, km * a.x};
          if (dzb) b = cpx<T>{-km * b.y, km * b.x};
        }
        v[q] = cj ? cpx<T>{a.x + b.y, b.x - a.y} : cpx<T>{a.x - b.y, a.y + b.x};
      }
      first_pass_store<T, +1, LY, 1>(PF{}, v, buf, lt);
    }
  };
  cpx<T> A[R0], B[R0];
  cpx<T> s[R0];
  // LDS position of output n = lt + NB1 q (0 <= lt < NTHR): with a padded
  // layout of rows of RW elements and NTHR <= RW, the row of n is that of
  // NB1 q or the next one, so one comparison replaces the division
  auto out_pos = [&](int lt, int q) -> int {
    constexpr int RW = NF / PF::rad[PF::P - 1];
    if constexpr (PAD != 0 && NTHR <= RW) {
      const int c = NB1 * q;
      return lt + c + (c / RW + (lt >= RW - c % RW ? 1 : 0)) * PAD;
    } else {
      return LY::pos(lt + NB1 * q);
    }
  };
  // inverse pair from A, B, then the loads of the next pair (if any): PFE 1
  // issues them right after the first pass (in flight during the LDS
  // passes), PFE 0 after the outputs are read (fewer live registers)
  auto inverse = [&](bool dza, bool dzb, int fa, int fb, bool sub0 = false) {
    first(A, B, dza, dzb, sub0);
    if (PFE && fa >= 0) load(fa, fb, A, B);
    __syncthreads();
    const int lt = opaque(tid);
    if constexpr (TWC_LDS) fft_rest_tw<T, NTHR, +1, LY, TW_LDS ? TWL : 0, TWC>(PF{}, buf, twf, twl, twcl, lt);
    else fft_rest_tw<T, NTHR, +1, LY, TW_LDS ? TWL : 0, TWC>(PF{}, buf, twf, twl, twc, lt);
#pragma unroll
    for (int q = 0; q < R0; q++) s[q] = own ? buf[out_pos(lt, q)] : zero;
    __syncthreads();
    if (!PFE && fa >= 0) load(fa, fb, A, B);
  };
  // forward pair of the products pa (+ i pb), separated into the Y rows da
  // (and db, if pb is given)
  auto forward = [&](const T* pa, const T* pb, cpx<T>* da, cpx<T>* db) {
    const int lt = opaque(tid);
    if (own) {
      cpx<T> v[R0];
#pragma unroll
      for (int q = 0; q < R0; q++) v[q] = cpx<T>{pa[q], pb ? pb[q] : T(0)};
      first_pass_store<T, -1, LY, 1>(PF{}, v, buf, lt);
    }
    __syncthreads();
    if constexpr (TWC_LDS) fft_rest_tw<T, NTHR, -1, LY, TW_LDS ? TWL : 0, TWC>(PF{}, buf, twf, twl, twcl, lt);
    else fft_rest_tw<T, NTHR, -1, LY, TW_LDS ? TWL : 0, TWC>(PF{}, buf, twf, twl, twc, lt);
    const int t64 = g.ntp * 64;
    for (int k = lt; k < nz; k += NTHR) {
      cpx<T> xa = zero, xb = zero;
      if (k < nz - 1) {
        const cpx<T> zk = buf[LY::pos(k)], zc = buf[LY::pos(k == 0 ? 0 : NF - k)];
        xa = cpx<T>{T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y - zc.y)};
        xb = cpx<T>{T(0.5) * (zk.y + zc.y), T(-0.5) * (zk.x - zc.x)};
      }
      const int o = (k >> 6) * t64 + (k & 63);  // tp_col(g, k)
      da[o] = xa;
      if (pb) db[o] = xb;
    }
    __syncthreads();
  };
  // field index in W: u_r, u_t, u_z (0-2), d_r u_c (3 + c), d_theta u_c / r (6 + c)
  T ur[R0], ut[R0], uz[R0], hz[R0], n0[R0], n1[R0];
  // Pairing.  A pair z = a + i b is rounded at the scale of the larger of
  // a and b, so the fields are paired by size: near the wall (distance d,
  // no slip, zero divergence) u_r, d_theta u_r / r and d_z u_r fall off as
  // d^2, u_t, u_z and their theta and z derivatives and d_r u_r as d, and
  // d_r u_t, d_r u_z stay finite.  The pairs (u_z, u_t), (u_r, d_z u_r),
  // (d_theta u_r / r, d_r u_r), (d_theta u_t / r, d_z u_t), (d_r u_z,
  // d_r u_t), (d_theta u_z / r, d_z u_z) mismatch only in the third, whose
  // d^2 member enters N_r alone, where u_t^2 / r (order d^2) dominates.  The
  // mean flow and its shear (the l = 0 modes of u_z and d_r u_z, constant
  // along the row and an order of magnitude above the fluctuations) are kept
  // out of their transforms and added to the samples afterwards.
  // N_z of the first row, transformed with the second row's: held in LDS
  // beside the transform buffer where both fit (each thread its own values,
  // so no barrier), which keeps it out of the registers the second row needs
  __shared__ T nza[NZ_LDS ? R0 * NTHR : 1];
  T nzr[NZ_LDS ? 1 : R0];
  cpx<T>* ya = nullptr;  // the first row's N_z destination
#pragma unroll 1
  for (int rr = 0; rr < nrow; rr++) {
    set_row(row0 + rr);
    const T m0z = W[2 * ws + roff].x, m0d = W[5 * ws + roff].x;  // l = 0 modes of u_z, d_r u_z
    load(2, 1, A, B);
    inverse(false, false, 0, 0, true);  // (u_z, u_t)
#pragma unroll
    for (int q = 0; q < R0; q++) { uz[q] = s[q].x + m0z; ut[q] = s[q].y; }
    inverse(false, true, 6, 3);  // (u_r, d_z u_r)
#pragma unroll
    for (int q = 0; q < R0; q++) { ur[q] = s[q].x; hz[q] = s[q].y; }
    // calcNonlinear_k (mod_io_nonlinear.cu:142-170), summed in its order
    inverse(false, false, 7, 1);  // (d_theta u_r / r, d_r u_r)
#pragma unroll
    for (int q = 0; q < R0; q++) {
      T v = ur[q] * s[q].y;
      v += ut[q] * s[q].x;
      v = v + uz[q] * hz[q];
      v = v - ut[q] * over_r(ut[q]);
      n0[q] = -v;
    }
    inverse(false, true, 5, 4);  // (d_theta u_t / r, d_z u_t)
#pragma unroll
    for (int q = 0; q < R0; q++) {
      n1[q] = ut[q] * s[q].x;
      hz[q] = s[q].y;
    }
    inverse(false, false, 8, 2, true);  // (d_r u_z, d_r u_t)
#pragma unroll
    for (int q = 0; q < R0; q++) {
      T v = ur[q] * s[q].y;
      v = v + n1[q];
      v = v + uz[q] * hz[q];
      v = v + ur[q] * over_r(ut[q]);
      n1[q] = -v;
      hz[q] = ur[q] * (s[q].x + m0d);  // the first term of N_z
    }
    forward(n0, n1, Y + roff, Y + ws + roff);
    inverse(false, true, -1, -1);  // (d_theta u_z / r, d_z u_z)
#pragma unroll
    for (int q = 0; q < R0; q++) {
      T v = hz[q];
      v += ut[q] * s[q].x;
      v = v + uz[q] * s[q].y;
      n0[q] = -v;
    }
    if (rr == 0 && nrow == 2) {  // held for the pair with the second row
#pragma unroll
      for (int q = 0; q < R0; q++) {
        if constexpr (NZ_LDS) nza[q * NTHR + tid] = n0[q];
        else nzr[q] = n0[q];
      }
      ya = Y + 2 * ws + roff;
    } else if (rr == 1) {
      T na[R0];
#pragma unroll
      for (int q = 0; q < R0; q++) {
        if constexpr (NZ_LDS) na[q] = nza[q * NTHR + tid];
        else na[q] = nzr[q];
      }
      forward(na, n0, ya, Y + 2 * ws + roff);
    } else {
      forward(n0, nullptr, Y + 2 * ws + roff, nullptr);
    }
  }
}

// ------------------------------------------------------------------ theta forward
template <typename T, int TL, int NTHR, typename LY, typename PL, int BAT>
__global__ __launch_bounds__(NTHR) void theta_fwd_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                     const cpx<T>* __restrict__ Y, size_t ws,
                                                     const double* __restrict__ D,
                                                     const double2* __restrict__ u,
                                                     double2* __restrict__ N, int ord, const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    Y = gmem(tab, reinterpret_cast<const cpx<T>*>(m.Y));
    ws = m.ws;
    D = gmem(tab, m.D);
    u = gmem(tab, m.u);
    N = gmem(tab, m.N);
  }
  if (step_done(g)) return;
  constexpr int NTP = PL::L;
  static_assert(64 % TL == 0, "a column tile lies in one 64-column tile of W, Y");
  __shared__ cpx<T> buf[LY::size(NTP * TL)];
  int i, l0;
  theta_block(g, TL, ord, i, l0);
  if (i >= g.lv0 + g.nlv || l0 >= g.nz) return;  // beyond this member's (or level chunk's) extent
  const int tid = threadIdx.x;
  const int nt = g.nt, half = nt / 2, nz = g.nz;
  const double norm = double(NTP) * double(g.nzf);  // mod_fft.cu:41
#pragma unroll 1
  for (int c = 0; c < 3; c++) {
    const int lt = opaque(tid);
    const cpx<T>* src = Y + c * ws + (size_t)i * tp_plane(g);  // tile-major Y
    for (int e = lt; e < NTP * TL; e += NTHR) {
      const int j = e / TL, tl = e - j * TL, l = l0 + tl;
      cpx<T> v{T(0), T(0)};
      if (l < nz - 1) v = src[(size_t)j * 64 + tp_col(g, l)];
      buf[LY::pos(PL::digitrev(j) * TL + tl)] = v;
    }
    __syncthreads();
    fft_run<T, TL, NTHR, -1, LY>(PL{}, buf, tw, lt);
    for (int e = lt; e < nt * TL; e += NTHR) {
      const int it = e / TL, tl = e - it * TL, l = l0 + tl;
      if (l < nz) {
        const int j = it < half ? it : it + half;  // padBackward_k row map
        const cpx<T> w = buf[LY::pos(j * TL + tl)];
        double vx = double(w.x) / norm, vy = double(w.y) / norm;
        if (it == half || l == nz - 1) vx = vy = 0.0;
        if (c == 2 && it == 0 && l == 0) {  // pressGrad_k (mod_radialFD.cu:100-111)
          double du = 0.0;
          for (int j2 = 0; j2 < IW + 1; j2++)
            du += D[STEN * (g.nr - 1) + j2 + IW] * u[2 * g.n + (size_t)(g.nr - 1 - IW + j2) * g.modes].x;
          vx -= (2.0 / g.re) * du;
        }
        // N(u) is stored at the transform precision (FP32 in the mixed form)
        reinterpret_cast<cpx<T>*>(N)[c * dr_cs(g) + (size_t)i * dr_plane(g) + (size_t)it * 64 + dr_col(g, l)] =
            cpx<T>{T(vx), T(vy)};  // tile-major N(u)
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ theta, register forms
// Same transforms and arithmetic as theta_inv_ct / theta_fwd_ct with the
// first pass fed from global memory and the last pass written to global
// memory from registers: per transform two LDS round trips instead of four,
// and no digit-reversed scatter.  Work item p = (butterfly n0, column):
// column = p % TL (lanes of a wave cover consecutive l), n0 = p / TL; it
// owns the rows j = n0 + (NTP/R0) q, q < R0.
//
// padForward_k row map: row j < NT/2 holds it = j, rows NT/2 <= j < NT are
// zero, row j >= NT holds it = j - NT/2 (NTP - NT = NT/2).
template <typename T, int TL, int NTHR, typename LY, typename PL, int BAT>
__global__ __launch_bounds__(NTHR) void theta_inv2_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                      const double* __restrict__ r,
                                                      const double2* __restrict__ u,
                                                      const cpx<T>* __restrict__ dR,
                                                      cpx<T>* __restrict__ W, size_t ws, int ord,
                                                      const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    r = gmem(tab, m.r);
    u = gmem(tab, m.u);
    dR = gmem(tab, reinterpret_cast<const cpx<T>*>(m.dR));
    W = gmem(tab, reinterpret_cast<cpx<T>*>(m.W));
    ws = m.ws;
  }
  if (step_done(g)) return;
  constexpr int NTP = PL::L;
  static_assert(64 % TL == 0, "a column tile lies in one 64-column tile of W, Y");
  constexpr int NT = 2 * NTP / 3;
  constexpr int half = NT / 2;
  constexpr int R0 = PL::rad[0];
  constexpr int NB1 = NTP / R0;
  constexpr int PAIRS = NB1 * TL;
  constexpr int PPT = (PAIRS + NTHR - 1) / NTHR;
  __shared__ cpx<T> buf[LY::size(NTP * TL)];
  const int nz = g.nz;
  int i, l0;
  theta_block(g, TL, ord, i, l0);
  if (i >= g.lv0 + g.nlv || l0 >= g.nz) return;  // beyond this member's (or level chunk's) extent
  const int tid = threadIdx.x;
  const double ri = r[i];
  // m u / r: the FP32 form multiplies by the reciprocal (the result is rounded
  // to FP32 anyway), the FP64 form divides as derivt_k does
  const double iri = 1.0 / ri;
  auto m_over_r = [&](double kt, double x) -> double {
    if constexpr (sizeof(T) == 4) return kt * x * iri;
    else return kt * x / ri;
  };
#pragma unroll 1
  for (int f = 0; f < 9; f++) {
    const int c = f % 3, kind = f / 3;  // kind 0: u, 1: d_r u, 2: d_theta u / r
    const int lt = opaque(tid);
    {
      const double2* src = u + c * g.n + (size_t)i * g.modes;
      const cpx<T>* srd = dR + c * dr_cs(g) + (size_t)i * dr_plane(g);  // tile-major d_r u
#pragma unroll
      for (int s = 0; s < PPT; s++) {
        const int p = lt + s * NTHR;
        if ((PAIRS % NTHR) == 0 || p < PAIRS) {
          const int col = p % TL, n0 = p / TL, l = l0 + col;
          cpx<T> v[R0];
#pragma unroll
          for (int q = 0; q < R0; q++) {
            const int j = n0 + NB1 * q;
            const int it = j < half ? j : (j >= NT ? j - half : -1);
            v[q] = cpx<T>{T(0), T(0)};
            if (it >= 0 && l < nz) {
              if (kind == 1) {
                v[q] = srd[(size_t)it * 64 + dr_col(g, l)];
              } else {
                const double2 a = src[(size_t)it * nz + l];
                if (kind == 2) {  // derivt_k (mod_deriv.cu:10-21)
                  const double kt = double(it < half ? it : it - NT);
                  v[q] = cpx<T>{T(m_over_r(-kt, a.y)), T(m_over_r(kt, a.x))};
                } else {
                  v[q] = cpx<T>{T(a.x), T(a.y)};
                }
              }
            }
          }
          first_pass_store<T, +1, LY, TL>(PL{}, v, buf, n0, col);
        }
      }
      __syncthreads();
      // tile-major W: the TL columns from l0 lie in one 64-column tile
      cpx<T>* dst = W + (kind * 3 + c) * ws + (size_t)i * tp_plane(g) + tp_col(g, l0);
      fft_rest_emit<T, TL, NTHR, +1, LY>(PL{}, buf, tw, lt, [&](int col, int j, cpx<T> w) {
        if (l0 + col < nz) dst[(size_t)j * 64 + col] = w;
      });
      __syncthreads();
    }
  }
}

// padBackward_k: output row j < NT/2 is it = j, row j >= NT is it = j - NT/2,
// rows NT/2 <= j < NT are dropped; it = NT/2 (m = -NT/2) and l = NZ-1 are
// zeroed.
template <typename T, int TL, int NTHR, typename LY, typename PL, int BAT>
__global__ __launch_bounds__(NTHR) void theta_fwd2_ct(Geom g, const cpx<T>* __restrict__ tw,
                                                      const cpx<T>* __restrict__ Y, size_t ws,
                                                      const double* __restrict__ D,
                                                      const double2* __restrict__ u,
                                                      double2* __restrict__ N, int ord, const NlMember* tab) {
  if constexpr (BAT != 0) {
    const NlMember& m = tab[blockIdx.y];
    g = m.g;
    Y = gmem(tab, reinterpret_cast<const cpx<T>*>(m.Y));
    ws = m.ws;
    D = gmem(tab, m.D);
    u = gmem(tab, m.u);
    N = gmem(tab, m.N);
  }
  if (step_done(g)) return;
  constexpr int NTP = PL::L;
  static_assert(64 % TL == 0, "a column tile lies in one 64-column tile of W, Y");
  constexpr int NT = 2 * NTP / 3;
  constexpr int half = NT / 2;
  constexpr int R0 = PL::rad[0];
  constexpr int NB1 = NTP / R0;
  constexpr int PAIRS = NB1 * TL;
  constexpr int PPT = (PAIRS + NTHR - 1) / NTHR;
  __shared__ cpx<T> buf[LY::size(NTP * TL)];
  // The pass twiddles (NTP - 1 entries) copied into LDS where that keeps
  // the workgroups per CU that the transform buffer allows (two up to 80 KB)
  constexpr int BUFB = LY::size(NTP * TL) * (int)sizeof(cpx<T>);
  constexpr int TWB = (NTP - 1) * (int)sizeof(cpx<T>);
  constexpr bool TW_LDS = BUFB <= 81920 ? BUFB + TWB <= 81920 : BUFB + TWB <= 163840;
  __shared__ cpx<T> twl[TW_LDS ? NTP - 1 : 1];
  int i, l0;
  theta_block(g, TL, ord, i, l0);
  if (i >= g.lv0 + g.nlv || l0 >= g.nz) return;  // beyond this member's (or level chunk's) extent
  const int tid = threadIdx.x;
  if constexpr (TW_LDS) {  // read after the barrier that follows the first pass
    for (int e = tid; e < NTP - 1; e += NTHR) twl[e] = tw[e];
  }
  const int nz = g.nz;
  const double norm = double(NTP) * double(g.nzf);  // mod_fft.cu:41
  // the FP32 form scales by the reciprocal (its result is rounded to FP32),
  // the FP64 form divides as the reference does
  const double inorm = 1.0 / norm;
  auto scale = [&](T x) -> double {
    if constexpr (sizeof(T) == 4) return double(x) * inorm;
    else return x / norm;
  };
  // A thread's first-pass pairs p = tid + s NTHR keep one column (NTHR is a
  // multiple of TL): one column test, and the loads of a pair issued
  // together with clamped offsets instead of per-element branches.
  static_assert(NTHR % TL == 0, "a thread keeps one column");
  const int tcol = tid % TL;
  const bool lok = l0 + tcol < nz - 1;  // l = NZ-1 and columns past NZ enter as zero
  // first-pass inputs of component c (tile-major Y); the next component's
  // are loaded while the current one's LDS passes run
  cpx<T> v[PPT][R0];
  auto load = [&](int c) {
    const int lt = opaque(tid);
    const cpx<T>* src = Y + c * ws + (size_t)i * tp_plane(g) + tp_col(g, lok ? l0 + tcol : 0);
#pragma unroll
    for (int s = 0; s < PPT; s++) {
      const int p = lt + s * NTHR;
      if ((PAIRS % NTHR) == 0 || p < PAIRS) {
        const int n0 = p / TL;
#pragma unroll
        for (int q = 0; q < R0; q++) v[s][q] = src[(size_t)(n0 + NB1 * q) * 64];
        if (!lok) {
#pragma unroll
          for (int q = 0; q < R0; q++) v[s][q] = cpx<T>{T(0), T(0)};
        }
      }
    }
  };
  load(0);
#pragma unroll 1
  for (int c = 0; c < 3; c++) {
    const int lt = opaque(tid);
#pragma unroll
    for (int s = 0; s < PPT; s++) {
      const int p = lt + s * NTHR;
      if ((PAIRS % NTHR) == 0 || p < PAIRS) first_pass_store<T, -1, LY, TL>(PL{}, v[s], buf, p / TL, p % TL);
    }
    __syncthreads();
    if (c + 1 < 3) load(c + 1);
    // N(u) is stored at the transform precision (FP32 in the mixed form)
    // tile-major N(u): the TL columns from l0 lie in one 64-column tile
    cpx<T>* dst = reinterpret_cast<cpx<T>*>(N) + c * dr_cs(g) + (size_t)i * dr_plane(g) + dr_col(g, l0);
    fft_rest_emit_tw<T, TL, NTHR, -1, LY, TW_LDS ? NTP - 1 : 0>(PL{}, buf, tw, twl, lt, [&](int col, int j, cpx<T> w) {
      const int l = l0 + col;
      if (l < nz && (j < half || j >= NT)) {
        const int it = j < half ? j : j - half;
        double vx = scale(w.x), vy = scale(w.y);
        if (it == half || l == nz - 1) vx = vy = 0.0;
        if (c == 2 && it == 0 && l == 0) {  // pressGrad_k (mod_radialFD.cu:100-111)
          double du = 0.0;
          for (int j2 = 0; j2 < IW + 1; j2++)
            du += D[STEN * (g.nr - 1) + j2 + IW] * u[2 * g.n + (size_t)(g.nr - 1 - IW + j2) * g.modes].x;
          vx -= (2.0 / g.re) * du;
        }
        dst[(unsigned)it * 64u + (unsigned)col] = cpx<T>{T(vx), T(vy)};
      }
    });
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
  X(576, Plan<4, 9, 16>)   \
  X(1152, Plan<8, 9, 16>)

// Full-length plans NZF = 2 NH of the pair form, keyed by NH (one first-pass
// butterfly per thread: NZF / R0 threads).
#define NSP_ZPAIR_PLANS(X)        \
  X(96, Plan<3, 8, 8>)            \
  X(576, Plan<9, 8, 16>)          \
  X(768, Plan<3, 8, 8, 8>)        \
  X(1536, Plan<6, 8, 8, 8>)       \
  X(2304, Plan<9, 8, 8, 8>)       \
  X(4608, Plan<18, 8, 8, 8>)

template <int NH>
struct ZPair {
  using P = void;
};
#define X(LL, ...)          \
  template <>               \
  struct ZPair<LL> {        \
    using P = __VA_ARGS__;  \
  };
NSP_ZPAIR_PLANS(X)
#undef X

#define NSP_ZHALF_PLANS(X)    \
  X(96, Plan<6, 16>)          \
  X(576, Plan<4, 9, 16>)      \
  X(768, Plan<3, 16, 16>)     \
  X(1536, Plan<6, 16, 16>)    \
  X(2304, Plan<9, 16, 16>)    \
  X(4608, Plan<9, 2, 16, 16>)

// ------------------------------------------------------------------ stage launchers
// A form is compiled for a plan only if its LDS buffer fits the 160 KB of a CU.
template <typename T, typename LY>
constexpr bool lds_fits(int n) {
  return LY::size(n) * (int)sizeof(cpx<T>) <= 163840;
}
// Variant dispatch for one compiled plan PL.  The LDS pad shift is 4 (one
// slot per 16 elements) for tile width 4 and 3 (one per 8) for width 8.
// Row padding of the theta scatter forms: rows of (NTP / R_last) positions x
// TL columns followed by TL unused slots (bank model of the scatter and pass
// patterns).
template <typename PL, int TL>
using ThetaRowPad = Lay<(PL::L / PL::rad[PL::P - 1]) * TL, TL>;

template <typename T, typename PL, int BAT>
bool theta_inv_plan(const ThetaVar& v, const Geom& g, const cpx<T>* tw, const double* r, const double2* u,
                    const double2* dR_, cpx<T>* W, size_t ws, hipStream_t s, const Batch* bt) {
  const cpx<T>* dR = reinterpret_cast<const cpx<T>*>(dR_);  // stored at precision T
  const int tl = v.tl;
  if (PL::L * tl * (int)sizeof(cpx<T>) > 163840) return false;
  const NlMember* tab = bt ? bt->tab : nullptr;
  const dim3 nb = bt ? dim3(((bt->max_nz + tl - 1) / tl) * bt->max_nr, bt->count) : dim3(((g.nz + tl - 1) / tl) * g.nlv);
  if (v.alg == 1) {
    if (tl == 4 && v.nthr == 192) { if constexpr (lds_fits<T, Dense>(PL::L * 4)) theta_inv2_ct<T, 4, 192, Dense, PL, BAT><<<nb, 192, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
    else if (tl == 8 && v.nthr == 256) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_inv2_ct<T, 8, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
    else if (tl == 8 && v.nthr == 384) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_inv2_ct<T, 8, 384, Dense, PL, BAT><<<nb, 384, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
    else if (tl == 16 && v.nthr == 256) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_inv2_ct<T, 16, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
    else return false;
    return true;
  }
  if (v.alg != 0) return false;
  using R4 = ThetaRowPad<PL, 4>;
  using R8 = ThetaRowPad<PL, 8>;
  using S4 = Swz<4, swz_bits<cpx<T>, 4>()>;  // pad 3: swizzled rows
  using S8 = Swz<8, swz_bits<cpx<T>, 8>()>;
  using S16 = Swz<16, swz_bits<cpx<T>, 16>()>;
  if (tl == 4 && v.pad == 0) { if constexpr (lds_fits<T, Dense>(PL::L * 4)) theta_inv_ct<T, 4, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 4 && v.pad == 1) { if constexpr (lds_fits<T, Lay<16, 1>>(PL::L * 4)) theta_inv_ct<T, 4, 256, Lay<16, 1>, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 4 && v.pad == 2) { if constexpr (lds_fits<T, R4>(PL::L * 4)) theta_inv_ct<T, 4, 256, R4, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 0) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_inv_ct<T, 8, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 1) { if constexpr (lds_fits<T, Lay<8, 1>>(PL::L * 8)) theta_inv_ct<T, 8, 256, Lay<8, 1>, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 2) { if constexpr (lds_fits<T, R8>(PL::L * 8)) theta_inv_ct<T, 8, 256, R8, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 0 && v.nthr == 512) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_inv_ct<T, 16, 512, Dense, PL, BAT><<<nb, 512, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 0 && v.nthr == 1024) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_inv_ct<T, 16, 1024, Dense, PL, BAT><<<nb, 1024, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 32 && v.pad == 0 && v.nthr == 1024) { if constexpr (lds_fits<T, Dense>(PL::L * 32)) theta_inv_ct<T, 32, 1024, Dense, PL, BAT><<<nb, 1024, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 0) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_inv_ct<T, 16, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 4 && v.pad == 3) { if constexpr (lds_fits<T, Dense>(PL::L * 4)) theta_inv_ct<T, 4, 256, S4, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 3) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_inv_ct<T, 8, 256, S8, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 3) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_inv_ct<T, 16, 256, S16, PL, BAT><<<nb, 256, 0, s>>>(g, tw, r, u, dR, W, ws, v.ord, tab); else return false; }
  else return false;
  return true;
}

template <typename T, typename PL, int BAT>
bool theta_fwd_plan(const ThetaVar& v, const Geom& g, const cpx<T>* tw, const cpx<T>* Y, size_t ws, const double* D,
                    const double2* u, double2* N, hipStream_t s, const Batch* bt) {
  const int tl = v.tl;
  if (PL::L * tl * (int)sizeof(cpx<T>) > 163840) return false;
  const NlMember* tab = bt ? bt->tab : nullptr;
  const dim3 nb = bt ? dim3(((bt->max_nz + tl - 1) / tl) * bt->max_nr, bt->count) : dim3(((g.nz + tl - 1) / tl) * g.nlv);
  if (v.alg == 1) {
    if (tl == 4 && v.nthr == 192) { if constexpr (lds_fits<T, Dense>(PL::L * 4)) theta_fwd2_ct<T, 4, 192, Dense, PL, BAT><<<nb, 192, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
    else if (tl == 8 && v.nthr == 256) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_fwd2_ct<T, 8, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
    else if (tl == 8 && v.nthr == 384) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_fwd2_ct<T, 8, 384, Dense, PL, BAT><<<nb, 384, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
    else if (tl == 16 && v.nthr == 256) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_fwd2_ct<T, 16, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
    else return false;
    return true;
  }
  if (v.alg != 0) return false;
  using R4 = ThetaRowPad<PL, 4>;
  using R8 = ThetaRowPad<PL, 8>;
  using S4 = Swz<4, swz_bits<cpx<T>, 4>()>;  // pad 3: swizzled rows
  using S8 = Swz<8, swz_bits<cpx<T>, 8>()>;
  using S16 = Swz<16, swz_bits<cpx<T>, 16>()>;
  if (tl == 4 && v.pad == 0) { if constexpr (lds_fits<T, Dense>(PL::L * 4)) theta_fwd_ct<T, 4, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 4 && v.pad == 1) { if constexpr (lds_fits<T, Lay<16, 1>>(PL::L * 4)) theta_fwd_ct<T, 4, 256, Lay<16, 1>, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 4 && v.pad == 2) { if constexpr (lds_fits<T, R4>(PL::L * 4)) theta_fwd_ct<T, 4, 256, R4, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 0) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_fwd_ct<T, 8, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 1) { if constexpr (lds_fits<T, Lay<8, 1>>(PL::L * 8)) theta_fwd_ct<T, 8, 256, Lay<8, 1>, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 2) { if constexpr (lds_fits<T, R8>(PL::L * 8)) theta_fwd_ct<T, 8, 256, R8, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 0 && v.nthr == 512) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_fwd_ct<T, 16, 512, Dense, PL, BAT><<<nb, 512, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 0 && v.nthr == 1024) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_fwd_ct<T, 16, 1024, Dense, PL, BAT><<<nb, 1024, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 0) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_fwd_ct<T, 16, 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 4 && v.pad == 3) { if constexpr (lds_fits<T, Dense>(PL::L * 4)) theta_fwd_ct<T, 4, 256, S4, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 8 && v.pad == 3) { if constexpr (lds_fits<T, Dense>(PL::L * 8)) theta_fwd_ct<T, 8, 256, S8, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else if (tl == 16 && v.pad == 3) { if constexpr (lds_fits<T, Dense>(PL::L * 16)) theta_fwd_ct<T, 16, 256, S16, PL, BAT><<<nb, 256, 0, s>>>(g, tw, Y, ws, D, u, N, v.ord, tab); else return false; }
  else return false;
  return true;
}

// z-pass row padding of the register form: R0 slots per row for the
// <R0, 16, 16> plans (bank model of the pass patterns), none otherwise.
template <typename PL>
constexpr int zpass2_pad() {
  return (PL::P == 3 && PL::rad[1] == 16 && PL::rad[2] == 16) ? PL::rad[0] : 0;
}

template <typename T, typename PL, int BAT>
bool zpass_plan(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* twz, const double* r,
                const cpx<T>* W, size_t ws, cpx<T>* Y, hipStream_t s, const Batch* bt) {
  constexpr int LL = PL::L;
  const NlMember* tab = bt ? bt->tab : nullptr;
  const dim3 nb = bt ? dim3(bt->max_nr * g.ntp, bt->count) : dim3(g.nlv * g.ntp);
  // register forms: one first-pass butterfly per thread, 256 threads (512
  // when the first pass has 512 butterflies, as for 4608)
  constexpr int NB1 = LL / PL::rad[0];
  constexpr int RT = NB1 <= 256 ? 256 : 512;
  if (v.z_alg == 1) {  // register form
    if constexpr (NB1 <= 512) {
      constexpr int PAD = zpass2_pad<PL>();
      if (v.z_pf)
        zpass2_ct<T, RT, 1, PAD, PL, BAT><<<nb, RT, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
      else
        zpass2_ct<T, RT, 0, PAD, PL, BAT><<<nb, RT, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
      return true;
    } else {
      return false;
    }
  }
  if (v.z_alg == 2) {  // triples of inverse transforms; z_pad: 3-column row padding 0, 1, R0
    if constexpr (NB1 <= 512 && lds_fits<T, Dense>(3 * LL + PL::rad[PL::P - 1] * PL::rad[0])) {
      constexpr int PAD = zpass2_pad<PL>();
      if (v.z_pad == 0)
        zpass3_ct<T, RT, PAD, 0, PL, BAT><<<nb, RT, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
      else if (v.z_pad == 1)
        zpass3_ct<T, RT, PAD, 1, PL, BAT><<<nb, RT, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
      else if (v.z_pad == 2)
        zpass3_ct<T, RT, PAD, PL::rad[0], PL, BAT><<<nb, RT, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
      else
        return false;
      return true;
    } else {
      return false;
    }
  }
  if (v.z_alg == 3 || v.z_alg == 4) {  // pairs of real transforms as full-length complex ones (4: two rows
                                       // per workgroup); z_pad: row padding 0, 1, R0
    using PP = typename ZPair<LL>::P;
    if constexpr (!std::is_void_v<PP>) {
      constexpr int NT4 = PP::L / PP::rad[0];
      constexpr int RL4 = PP::L / PP::rad[PP::P - 1];
      if constexpr (NT4 % 64 == 0 && NT4 <= 1024 && lds_fits<T, Lay<RL4, PP::rad[0]>>(PP::L)) {
        auto go = [&](auto pf, auto rows) {
          constexpr int PFE = decltype(pf)::value;
          constexpr int RW = decltype(rows)::value;
          const int nrw = bt ? bt->max_nr * g.ntp : g.nlv * g.ntp;
          const dim3 nbr = bt ? dim3((nrw + RW - 1) / RW, bt->count) : dim3((nrw + RW - 1) / RW);
          if (v.z_pad == 0)
            zpass4_ct<T, NT4, 0, PFE, RW, PP, BAT><<<nbr, NT4, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
          else if (v.z_pad == 1)
            zpass4_ct<T, NT4, 1, PFE, RW, PP, BAT><<<nbr, NT4, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
          else if (v.z_pad == 2)
            zpass4_ct<T, NT4, PP::rad[0], PFE, RW, PP, BAT><<<nbr, NT4, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
          else
            return false;
          return true;
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        if (v.z_alg == 3) return v.z_pf ? go(I1{}, I1{}) : go(I0{}, I1{});
        return v.z_pf ? go(I1{}, I2{}) : go(I0{}, I2{});
      }
    }
    return false;
  }
  if (v.z_alg != 0) return false;
  using RP = Lay<LL / PL::rad[PL::P - 1], zpass2_pad<PL>()>;
  constexpr bool t256 = (2 * LL + 255) / 256 <= 24;  // samples per thread of the 256-thread form
  if (v.z_nthr == 256 && !t256) return false;
  if (v.z_pad == 2) {
    if (v.z_nthr == 256) {
      if constexpr (t256)
        zpass_ct<T, 256, (2 * LL + 255) / 256, RP, PL, BAT><<<nb, 256, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
    } else if (v.z_nthr == 512)
      zpass_ct<T, 512, (2 * LL + 511) / 512, RP, PL, BAT><<<nb, 512, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
    else
      return false;
    return true;
  }
  if (v.z_nthr == 256 && !v.z_pad) {
    if constexpr (t256)
      zpass_ct<T, 256, (2 * LL + 255) / 256, Dense, PL, BAT><<<nb, 256, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
  } else if (v.z_nthr == 256) {
    if constexpr (t256)
      zpass_ct<T, 256, (2 * LL + 255) / 256, Lay<16, 1>, PL, BAT><<<nb, 256, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
  }
  else if (v.z_nthr == 512 && !v.z_pad)
    zpass_ct<T, 512, (2 * LL + 511) / 512, Dense, PL, BAT><<<nb, 512, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
  else if (v.z_nthr == 512)
    zpass_ct<T, 512, (2 * LL + 511) / 512, Lay<16, 1>, PL, BAT><<<nb, 512, 0, s>>>(g, tw, twz, r, W, ws, Y, tab);
  else
    return false;
  return true;
}

// Return false when the requested variant is not compiled for this size.
// BAT = 1: the batched kernels (one member per blockIdx.y, arguments from
// bt->tab); BAT = 0: the single-pipe kernels, whose arguments are only the
// launch's own (so every field pointer stays a global-memory pointer).
template <typename T, int BAT>
bool ct_theta_inv_b(const NlVariant& v, const Geom& g, const cpx<T>* tw, const double* r, const double2* u,
                    const double2* dR, cpx<T>* W, size_t ws, hipStream_t s, const Batch* bt) {
#define X(LL, ...) \
  if (g.ntp == LL) return theta_inv_plan<T, __VA_ARGS__, BAT>(v.inv, g, tw, r, u, dR, W, ws, s, bt);
  NSP_THETA_PLANS(X)
#undef X
  return false;
}

template <typename T, int BAT>
bool ct_theta_fwd_b(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* Y, size_t ws,
                    const double* D, const double2* u, double2* N, hipStream_t s, const Batch* bt) {
#define X(LL, ...) \
  if (g.ntp == LL) return theta_fwd_plan<T, __VA_ARGS__, BAT>(v.fwd, g, tw, Y, ws, D, u, N, s, bt);
  NSP_THETA_PLANS(X)
#undef X
  return false;
}

template <typename T, int BAT>
bool ct_zpass_b(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* twz, const double* r,
                const cpx<T>* W, size_t ws, cpx<T>* Y, hipStream_t s, const Batch* bt) {
#define X(LL, ...) \
  if (g.nh == LL) return zpass_plan<T, __VA_ARGS__, BAT>(v, g, tw, twz, r, W, ws, Y, s, bt);
  NSP_ZHALF_PLANS(X)
#undef X
  return false;
}

// The batched instantiations are compiled in their own translation units.
#define NSP_CT_EXTERN(T, B)                                                                                     \
  extern template bool ct_theta_inv_b<T, B>(const NlVariant&, const Geom&, const cpx<T>*, const double*,        \
                                            const double2*, const double2*, cpx<T>*, size_t, hipStream_t,      \
                                            const Batch*);                                                      \
  extern template bool ct_theta_fwd_b<T, B>(const NlVariant&, const Geom&, const cpx<T>*, const cpx<T>*, size_t, \
                                            const double*, const double2*, double2*, hipStream_t, const Batch*); \
  extern template bool ct_zpass_b<T, B>(const NlVariant&, const Geom&, const cpx<T>*, const cpx<T>*,            \
                                        const double*, const cpx<T>*, size_t, cpx<T>*, hipStream_t, const Batch*);
#define NSP_CT_INSTANTIATE(T, B)                                                                                \
  template bool ct_theta_inv_b<T, B>(const NlVariant&, const Geom&, const cpx<T>*, const double*, const double2*, \
                                     const double2*, cpx<T>*, size_t, hipStream_t, const Batch*);               \
  template bool ct_theta_fwd_b<T, B>(const NlVariant&, const Geom&, const cpx<T>*, const cpx<T>*, size_t,       \
                                     const double*, const double2*, double2*, hipStream_t, const Batch*);       \
  template bool ct_zpass_b<T, B>(const NlVariant&, const Geom&, const cpx<T>*, const cpx<T>*, const double*,    \
                                 const cpx<T>*, size_t, cpx<T>*, hipStream_t, const Batch*);

template <typename T>
bool ct_theta_inv(const NlVariant& v, const Geom& g, const cpx<T>* tw, const double* r, const double2* u,
                  const double2* dR, cpx<T>* W, size_t ws, hipStream_t s, const Batch* bt) {
  return bt ? ct_theta_inv_b<T, 1>(v, g, tw, r, u, dR, W, ws, s, bt)
            : ct_theta_inv_b<T, 0>(v, g, tw, r, u, dR, W, ws, s, bt);
}

template <typename T>
bool ct_theta_fwd(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* Y, size_t ws, const double* D,
                  const double2* u, double2* N, hipStream_t s, const Batch* bt) {
  return bt ? ct_theta_fwd_b<T, 1>(v, g, tw, Y, ws, D, u, N, s, bt)
            : ct_theta_fwd_b<T, 0>(v, g, tw, Y, ws, D, u, N, s, bt);
}

template <typename T>
bool ct_zpass(const NlVariant& v, const Geom& g, const cpx<T>* tw, const cpx<T>* twz, const double* r,
              const cpx<T>* W, size_t ws, cpx<T>* Y, hipStream_t s, const Batch* bt) {
  return bt ? ct_zpass_b<T, 1>(v, g, tw, twz, r, W, ws, Y, s, bt) : ct_zpass_b<T, 0>(v, g, tw, twz, r, W, ws, Y, s, bt);
}

}  // namespace nsp
