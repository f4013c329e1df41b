// Device FFT building blocks: small DFTs and an in-place mixed-radix
// decimation-in-time pass sequence over columns held in LDS.
//
// Data placement: sample n of a column must be stored at LDS position
// perm[n] (mixed-radix digit reversal, built by make_fft_plan) before the
// passes run; the transform then comes out in natural order.  Every pass
// combines r sub-transforms of length S into one of length S*r in place, so
// no staging buffer is needed and each thread may process any number of
// butterflies.  Twiddles come from a table tw[e] = exp(-2 pi i e / L).
#pragma once

#include "nsp_internal.h"

namespace nsp {

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by s*i (s = +1 or -1)
template <typename T>
__device__ __forceinline__ cpx<T> cmuli(cpx<T> a, T s) { return {-s * a.y, s * a.x}; }

// sign s: -1 forward (e^{-2 pi i nk/r}), +1 inverse
template <typename T>
__device__ __forceinline__ void dft2(cpx<T>* v) {
  const cpx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <typename T>
__device__ __forceinline__ void dft3(cpx<T>* v, T s) {
  const T h = T(-0.5), sn = s * T(0.86602540378443864676372317075294);
  const cpx<T> a = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
  const cpx<T> mm = {a.x + h * t1.x, a.y + h * t1.y};
  const cpx<T> rot = cmuli(t2, sn);
  v[0] = cadd(a, t1);
  v[1] = cadd(mm, rot);
  v[2] = csub(mm, rot);
}
template <typename T>
__device__ __forceinline__ void dft4(cpx<T>* v, T s) {
  const cpx<T> a = v[0], b = v[1], c = v[2], d = v[3];
  const cpx<T> ac = cadd(a, c), amc = csub(a, c), bd = cadd(b, d), bmd = cmuli(csub(b, d), s);
  v[0] = cadd(ac, bd);
  v[1] = cadd(amc, bmd);
  v[2] = csub(ac, bd);
  v[3] = csub(amc, bmd);
}
template <typename T>
__device__ __forceinline__ void dft5(cpx<T>* v, T s) {
  // cos(2pi/5), cos(4pi/5), sin(2pi/5), sin(4pi/5)
  const T c1 = T(0.30901699437494742410229341718282), c2 = T(-0.80901699437494742410229341718282);
  const T s1 = T(0.95105651629515357211643933337938), s2 = T(0.58778525229247312916870595463907);
  const cpx<T> a = v[0];
  const cpx<T> be = cadd(v[1], v[4]), bme = csub(v[1], v[4]);
  const cpx<T> cd = cadd(v[2], v[3]), cmd = csub(v[2], v[3]);
  const cpx<T> m1 = {a.x + c1 * be.x + c2 * cd.x, a.y + c1 * be.y + c2 * cd.y};
  const cpx<T> m2 = {a.x + c2 * be.x + c1 * cd.x, a.y + c2 * be.y + c1 * cd.y};
  const cpx<T> q1 = {s1 * bme.x + s2 * cmd.x, s1 * bme.y + s2 * cmd.y};
  const cpx<T> q2 = {s2 * bme.x - s1 * cmd.x, s2 * bme.y - s1 * cmd.y};
  const cpx<T> r1 = cmuli(q1, s), r2 = cmuli(q2, s);
  v[0] = cadd(a, cadd(be, cd));
  v[1] = cadd(m1, r1);
  v[4] = csub(m1, r1);
  v[2] = cadd(m2, r2);
  v[3] = csub(m2, r2);
}

// In-place passes over `ncol` columns: element n of column c lives at
// buf[c + n*ncol] (columns interleaved, so consecutive threads touch
// consecutive LDS words).  Caller has placed inputs at digit-reversed
// positions and synchronised.  Ends with a barrier.
template <typename T>
__device__ void fft_passes(cpx<T>* buf, const FftPlan& P, int ncol, T sign,
                           const cpx<T>* __restrict__ tw, int tid, int nthr) {
  int S = 1;
  for (int p = 0; p < P.nrad; ++p) {
    const int r = P.rad[p];
    const int nb = P.L / r;
    const int twm = P.L / (S * r);
    const int total = nb * ncol;
    for (int b = tid; b < total; b += nthr) {
      const int col = b % ncol;
      const int bb = b / ncol;
      const int k = bb % S;
      const int base = (bb - k) * r + k;  // (bb/S)*S*r + k
      cpx<T> v[5];
#pragma unroll
      for (int q = 0; q < 5; ++q)
        if (q < r) v[q] = buf[col + (base + q * S) * ncol];
#pragma unroll
      for (int q = 1; q < 5; ++q)
        if (q < r) {
          cpx<T> w = tw[q * k * twm];
          w.y = -sign * w.y;  // table holds e^{-i...}; sign=-1 keeps it
          v[q] = cmul(v[q], w);
        }
      switch (r) {
        case 2: dft2(v); break;
        case 3: dft3(v, sign); break;
        case 4: dft4(v, sign); break;
        default: dft5(v, sign); break;
      }
#pragma unroll
      for (int q = 0; q < 5; ++q)
        if (q < r) buf[col + (base + q * S) * ncol] = v[q];
    }
    __syncthreads();
    S *= r;
  }
}

}  // namespace nsp
