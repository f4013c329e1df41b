// Compile-time specialised FFT building blocks for LDS-resident transforms.
//
// Plan<R0, R1, ..., R_{P-1}>: length L = prod(R), in-place decimation in
// time.  Sample n is placed at LDS position Plan::digitrev(n); pass p (radix
// R_p, sub-length S_p = R_0...R_{p-1}) combines R_p sub-transforms of length
// S_p into one of length S_p*R_p; the result comes out in natural order.
// All radices, strides and divisors are compile-time constants, so index
// arithmetic reduces to shifts and multiply-high sequences and the butterfly
// loops unroll.  Butterflies of radix 6, 8, 9, 12, 16, 18 are composite
// in-register DFTs (R = R1*R2) with constant twiddles evaluated by the
// compiler.
//
// Twiddles between passes come from a compact table (host-built, see
// pass_twiddles) holding, for each pass in order, S*(R-1) entries
// tw[k*(R-1) + q-1] = exp(-2 pi i q k / (S R)), k < S, 1 <= q < R.
#pragma once

#include "nsp_internal.h"

namespace nsp {

// ---------------------------------------------------------------- constexpr trig
constexpr long double kPiL = 3.141592653589793238462643383279502884L;

constexpr long double ct_sin_small(long double x) {
  long double term = x, sum = x;
  for (int n = 1; n < 40; ++n) {
    term *= -x * x / ((2.0L * n) * (2.0L * n + 1.0L));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi m / R, exact to long double rounding, reduced to [-pi/4, pi/4]
constexpr long double ct_cos2pi(int m, int R) {
  m %= R;
  if (m < 0) m += R;
  // angle = 2 pi m / R; use octant reduction through sin/cos identities
  const long double a = 2.0L * kPiL * (long double)m / (long double)R;
  // cos(a) = sin(pi/2 - a); reduce into [-pi, pi]
  long double x = kPiL / 2.0L - a;
  while (x > kPiL) x -= 2.0L * kPiL;
  while (x < -kPiL) x += 2.0L * kPiL;
  if (x > kPiL / 2.0L) x = kPiL - x;
  if (x < -kPiL / 2.0L) x = -kPiL - x;
  return ct_sin_small(x);
}
constexpr long double ct_sin2pi(int m, int R) { return ct_cos2pi(m * 4 - R, 4 * R); }

// cos/sin(2 pi m / R) for m < R, evaluated by the compiler
template <int R>
struct TwTab {
  double c[R], s[R];
  constexpr TwTab() : c(), s() {
    for (int m = 0; m < R; ++m) {
      c[m] = double(ct_cos2pi(m, R));
      s[m] = double(ct_sin2pi(m, R));
    }
  }
};
template <int R>
inline constexpr TwTab<R> kTw{};

// ---------------------------------------------------------------- complex helpers
template <typename T>
__device__ __forceinline__ cpx<T> cx_mul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cx_mulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}

// ---------------------------------------------------------------- small DFTs
// In-place DFT of v[0..R) in natural order, sign SIGN (-1 forward, +1 inverse).
template <typename T, int R, int SIGN>
struct DFT;

template <typename T, int SIGN>
struct DFT<T, 2, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = {a.x + b.x, a.y + b.y};
    v[1] = {a.x - b.x, a.y - b.y};
  }
};
template <typename T, int SIGN>
struct DFT<T, 3, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) {
    constexpr T h = T(-0.5);
    constexpr T sn = T(SIGN) * T(0.86602540378443864676372317075294L);
    const cpx<T> a = v[0];
    const cpx<T> t1{v[1].x + v[2].x, v[1].y + v[2].y}, t2{v[1].x - v[2].x, v[1].y - v[2].y};
    const cpx<T> m{a.x + h * t1.x, a.y + h * t1.y};
    v[0] = {a.x + t1.x, a.y + t1.y};
    v[1] = {m.x - sn * t2.y, m.y + sn * t2.x};
    v[2] = {m.x + sn * t2.y, m.y - sn * t2.x};
  }
};
template <typename T, int SIGN>
struct DFT<T, 4, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) {
    constexpr T s = T(SIGN);
    const cpx<T> a = v[0], b = v[1], c = v[2], d = v[3];
    const cpx<T> ac{a.x + c.x, a.y + c.y}, amc{a.x - c.x, a.y - c.y};
    const cpx<T> bd{b.x + d.x, b.y + d.y};
    const cpx<T> bmd{-s * (b.y - d.y), s * (b.x - d.x)};  // s*i*(b-d)
    v[0] = {ac.x + bd.x, ac.y + bd.y};
    v[1] = {amc.x + bmd.x, amc.y + bmd.y};
    v[2] = {ac.x - bd.x, ac.y - bd.y};
    v[3] = {amc.x - bmd.x, amc.y - bmd.y};
  }
};
template <typename T, int SIGN>
struct DFT<T, 5, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) {
    constexpr T c1 = T(0.30901699437494742410229341718282L), c2 = T(-0.80901699437494742410229341718282L);
    constexpr T s1 = T(SIGN) * T(0.95105651629515357211643933337938L);
    constexpr T s2 = T(SIGN) * T(0.58778525229247312916870595463907L);
    const cpx<T> a = v[0];
    const cpx<T> be{v[1].x + v[4].x, v[1].y + v[4].y}, bme{v[1].x - v[4].x, v[1].y - v[4].y};
    const cpx<T> cd{v[2].x + v[3].x, v[2].y + v[3].y}, cmd{v[2].x - v[3].x, v[2].y - v[3].y};
    const cpx<T> m1{a.x + c1 * be.x + c2 * cd.x, a.y + c1 * be.y + c2 * cd.y};
    const cpx<T> m2{a.x + c2 * be.x + c1 * cd.x, a.y + c2 * be.y + c1 * cd.y};
    const cpx<T> q1{s1 * bme.x + s2 * cmd.x, s1 * bme.y + s2 * cmd.y};
    const cpx<T> q2{s2 * bme.x - s1 * cmd.x, s2 * bme.y - s1 * cmd.y};
    v[0] = {a.x + be.x + cd.x, a.y + be.y + cd.y};
    v[1] = {m1.x - q1.y, m1.y + q1.x};
    v[4] = {m1.x + q1.y, m1.y - q1.x};
    v[2] = {m2.x - q2.y, m2.y + q2.x};
    v[3] = {m2.x + q2.y, m2.y - q2.x};
  }
};

// Composite R = R1*R2: n = R2*n1 + n2, k = k1 + R1*k2,
// X[k] = sum_n2 W_R^{n2 k1} W_R2^{n2 k2} sum_n1 x[R2 n1 + n2] W_R1^{n1 k1}
template <typename T, int R1, int R2, int SIGN>
__device__ __forceinline__ void dft_composite(cpx<T>* v) {
  constexpr int R = R1 * R2;
  cpx<T> y[R];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    cpx<T> a[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[R2 * n1 + n2];
    DFT<T, R1, SIGN>::run(a);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) y[n2 * R1 + k1] = a[k1];
  }
  // constant twiddles W_R^{n2 k1} (compile-time values after unrolling)
#pragma unroll
  for (int n2 = 1; n2 < R2; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) {
      const int e = (n2 * k1) % R;
      const cpx<T> a = y[n2 * R1 + k1];
      if (4 * e == R) {  // W_R^e = -i: a swap and a sign (no products by 0 and 1; e is a constant after unrolling)
        y[n2 * R1 + k1] = {T(SIGN) * -a.y, T(SIGN) * a.x};
      } else {
        const T c = T(kTw<R>.c[e]);
        const T s = T(SIGN) * T(kTw<R>.s[e]);
        y[n2 * R1 + k1] = {a.x * c - a.y * s, a.x * s + a.y * c};
      }
    }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cpx<T> b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2 * R1 + k1];
    DFT<T, R2, SIGN>::run(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
  }
}
template <typename T, int SIGN> struct DFT<T, 6, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) { dft_composite<T, 3, 2, SIGN>(v); } };
template <typename T, int SIGN> struct DFT<T, 8, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) { dft_composite<T, 4, 2, SIGN>(v); } };
template <typename T, int SIGN> struct DFT<T, 9, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) { dft_composite<T, 3, 3, SIGN>(v); } };
template <typename T, int SIGN> struct DFT<T, 12, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) { dft_composite<T, 4, 3, SIGN>(v); } };
template <typename T, int SIGN> struct DFT<T, 16, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) { dft_composite<T, 4, 4, SIGN>(v); } };
template <typename T, int SIGN> struct DFT<T, 18, SIGN> {
  __device__ __forceinline__ static void run(cpx<T>* v) { dft_composite<T, 9, 2, SIGN>(v); } };

// ---------------------------------------------------------------- plans
template <int... R>
struct Plan {
  static constexpr int P = sizeof...(R);
  static constexpr int rad[P] = {R...};
  static constexpr int L = (1 * ... * R);
  // LDS position of input sample n (mixed-radix digit reversal)
  // (unsigned arithmetic: n >= 0, and the divisions by constants need no
  // sign corrections)
  __device__ __forceinline__ static int digitrev(int n_) {
    unsigned n = (unsigned)n_, pos = 0, len = L;
#pragma unroll
    for (int q = P - 1; q >= 0; --q) {
      len /= rad[q];
      pos += (n % rad[q]) * len;
      n /= rad[q];
    }
    return (int)pos;
  }
};

// LDS layout: logical element e at e + (e / W) * PAD, i.e. rows of W
// elements each followed by PAD unused slots (PAD = 0: dense).  Row padding
// turns the power-of-two strides of the passes into distinct banks.
template <int W, int PAD>
struct Lay {
  __device__ __forceinline__ static int pos(int e) {
    if constexpr (PAD == 0) return e;
    else return e + (e / W) * PAD;
  }
  static constexpr int size(int n) { return PAD == 0 ? n : n + ((n + W - 1) / W) * PAD; }
  // Positions of the elements first + q*SE, q < R, of a group that starts at
  // a multiple of SPAN: p0 + q*step.  Exact when the group stays in one row
  // (W % SPAN == 0) or steps by whole rows (SE % W == 0), with one division
  // instead of one per element; otherwise p0 = -1 (compute each position).
  template <int SE, int SPAN>
  __device__ __forceinline__ static void affine(int first, int& p0, int& step) {
    if constexpr (PAD == 0) {
      p0 = first;
      step = SE;
    } else if constexpr (W % SPAN == 0) {
      p0 = pos(first);
      step = SE;
    } else if constexpr (SE % W == 0) {
      p0 = pos(first);
      step = SE + (SE / W) * PAD;
    } else {
      p0 = -1;
      step = 0;
    }
  }
  template <int SE, int SPAN>
  static constexpr bool is_affine() { return PAD == 0 || W % SPAN == 0 || SE % W == 0; }
  // affine() for first = j SPAN + c with 0 <= c < SE: the row of first follows
  // from the group index j alone where a group lies in one row (j / (W /
  // SPAN)) or its SE-spaced elements start in row j (SPAN / W) (SE == W).
  template <int SE, int SPAN>
  __device__ __forceinline__ static void affine_g(int first, int j, int& p0, int& step) {
    if constexpr (PAD == 0) {
      p0 = first;
      step = SE;
    } else if constexpr (W % SPAN == 0) {
      p0 = first + (int)((unsigned)j / (unsigned)(W / SPAN)) * PAD;
      step = SE;
    } else if constexpr (SE == W) {
      p0 = first + j * (SPAN / W) * PAD;
      step = SE + PAD;
    } else {
      affine<SE, SPAN>(first, p0, step);
    }
  }
};
using Dense = Lay<1, 0>;

// Swizzled layout of W interleaved columns (W a power of two): logical
// element e = col + n W sits in row n, stored as row n ^ (((n >> B) ^
// (n >> 2B)) & (2^B - 1)).  Only the low B bits of the row index change, by
// a function of the higher bits: a permutation inside each aligned group of
// 2^B rows (no padding).  With 2^B rows per 256-byte bank period, the rows
// that the digit-reversed scatter and the pass patterns of the theta plans
// touch together land in different bank groups.  Not affine.
template <int W, int B>
struct Swz {
  __device__ __forceinline__ static int pos(int e) {
    if constexpr (B == 0) {
      return e;
    } else {
      constexpr int M = (1 << B) - 1;
      const int n = e / W, col = e - n * W;
      return (n ^ (((n >> B) ^ (n >> (2 * B))) & M)) * W + col;
    }
  }
  static constexpr int size(int n) { return n; }
  template <int SE, int SPAN>
  __device__ __forceinline__ static void affine(int first, int& p0, int& step) {
    p0 = first;
    step = SE;
  }
  template <int SE, int SPAN>
  static constexpr bool is_affine() { return B == 0; }
  template <int SE, int SPAN>
  __device__ __forceinline__ static void affine_g(int first, int, int& p0, int& step) {
    p0 = first;
    step = SE;
  }
};
// bits B of the swizzle for W columns of T: 2^B rows per 256 bytes
template <typename T, int W>
constexpr int swz_bits() {
  constexpr int g = 256 / (W * (int)sizeof(T));
  return g >= 8 ? 3 : (g >= 4 ? 2 : (g >= 2 ? 1 : 0));
}

// Butterfly b of one DIT pass (radix R, sub-length S) over NCOL interleaved
// columns: column col = b % NCOL, group j and offset k < S of b / NCOL = j S
// + k; its elements are n = base + q S, base = j R S + k.  The index
// arithmetic is unsigned (b >= 0): divisions by the constants need no sign
// corrections.
template <typename LY, int NCOL, int S, int R>
struct Group {
  int col, base, p0, st;  // p0 = -1: positions not affine in q
  unsigned k;
  __device__ __forceinline__ explicit Group(int b) {
    const unsigned ub = (unsigned)b;
    const unsigned bb = ub / NCOL;
    const unsigned j = bb / S;
    init((int)(ub % NCOL), j, bb - j * S);
  }
  // group j, offset k of column c given directly
  __device__ __forceinline__ Group(int c, unsigned j, unsigned k_) { init(c, j, k_); }
  __device__ __forceinline__ void init(int c, unsigned j, unsigned k_) {
    col = c;
    k = k_;
    __builtin_assume(k < (unsigned)S);
    base = (int)(j * (R * S) + k);
    if constexpr (LY::template is_affine<S * NCOL, S * R * NCOL>())
      LY::template affine_g<S * NCOL, S * R * NCOL>(col + base * NCOL, (int)j, p0, st);
    else
      p0 = -1, st = 0;
  }
  __device__ __forceinline__ int pos(int q) const {
    if constexpr (LY::template is_affine<S * NCOL, S * R * NCOL>()) return p0 + q * st;
    else return LY::pos(col + (base + q * S) * NCOL);
  }
};

// Loads the R elements of a butterfly, applies the inter-pass twiddles
// (at TWO) and the radix-R DFT; element q belongs at position gr.pos(q).
// The twiddles of a pass that ends within the first TWL entries are read
// from twl, an LDS copy of them (LDS reads instead of per-lane vector-cache
// reads, whose bandwidth the twiddle traffic of the transforms otherwise
// fills), the others from tw.  With TWC, a last pass of radix 8 reads only
// w^1, w^2, w^4 (w = exp(-2 pi i k / L)) from the compact table twc
// (3 entries per k, compact_twiddles) and forms w^3 = w w^2, w^5 = w w^4,
// w^6 = w^2 w^4, w^7 = w^3 w^4: 24 instead of 56 bytes per butterfly in
// FP32, for 4 complex products.
template <typename T, int L, int NCOL, int SIGN, typename LY, int TWL, int TWC, int S, int TWO, int R>
__device__ __forceinline__ void dit_butterfly(const cpx<T>* buf, const cpx<T>* __restrict__ tw, const cpx<T>* twl,
                                              const cpx<T>* twc, const Group<LY, NCOL, S, R>& gr, cpx<T>* v) {
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = buf[gr.pos(q)];
  if constexpr (TWC != 0 && S * R == L && R == 8) {
    const int o = 3 * (int)gr.k;
    cpx<T> w[8];
    w[1] = twc[o];
    w[2] = twc[o + 1];
    w[4] = twc[o + 2];
    w[3] = cx_mul(w[1], w[2]);
    w[5] = cx_mul(w[1], w[4]);
    w[6] = cx_mul(w[2], w[4]);
    w[7] = cx_mul(w[3], w[4]);
#pragma unroll
    for (int q = 1; q < R; ++q) v[q] = (SIGN < 0) ? cx_mul(v[q], w[q]) : cx_mulc(v[q], w[q]);
  } else if constexpr (S > 1) {
    const int o = TWO + (int)gr.k * (R - 1);
#pragma unroll
    for (int q = 1; q < R; ++q) {
      cpx<T> w;
      if constexpr (TWO + S * (R - 1) <= TWL) w = twl[o + q - 1];
      else w = tw[o + q - 1];
      v[q] = (SIGN < 0) ? cx_mul(v[q], w) : cx_mulc(v[q], w);
    }
  }
  DFT<T, R, SIGN>::run(v);
}

// Stores the R outputs of a butterfly back to its positions.
template <typename T, int NCOL, typename LY, int S, int R>
__device__ __forceinline__ void dit_store(cpx<T>* buf, const cpx<T>* v, const Group<LY, NCOL, S, R>& gr) {
#pragma unroll
  for (int q = 0; q < R; ++q) buf[gr.pos(q)] = v[q];
}

// One DIT pass and the remaining ones.  NCOL interleaved columns (element n of
// column c at logical index c + n*NCOL, stored at LY::pos of it); NTHR
// threads cooperate; ends with a barrier.
template <typename T, int L, int NCOL, int NTHR, int SIGN, typename LY, int TWL, int TWC, int S, int TWO, int R,
          int... Rest>
__device__ __forceinline__ void dit_passes(cpx<T>* buf, const cpx<T>* __restrict__ tw, const cpx<T>* twl,
                                           const cpx<T>* twc, int tid) {
  constexpr int TOT = (L / R) * NCOL;
  constexpr int ITER = (TOT + NTHR - 1) / NTHR;
  __builtin_assume((unsigned)tid < (unsigned)NTHR);
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int b = tid + it * NTHR;
    if ((TOT % NTHR) == 0 || b < TOT) {
      cpx<T> v[R];
      const Group<LY, NCOL, S, R> gr(b);
      dit_butterfly<T, L, NCOL, SIGN, LY, TWL, TWC, S, TWO, R>(buf, tw, twl, twc, gr, v);
      dit_store<T, NCOL, LY, S, R>(buf, v, gr);
    }
  }
  __syncthreads();
  if constexpr (sizeof...(Rest) > 0)
    dit_passes<T, L, NCOL, NTHR, SIGN, LY, TWL, TWC, S * R, TWO + S * (R - 1), Rest...>(buf, tw, twl, twc, tid);
}

// Like dit_passes, but the last pass hands its outputs to emit(col, k, value)
// (k = natural output index) instead of storing them; no barrier after it.
template <typename T, int L, int NCOL, int NTHR, int SIGN, typename LY, int TWL, int S, int TWO, int R, int... Rest,
          typename F>
__device__ __forceinline__ void dit_passes_emit(cpx<T>* buf, const cpx<T>* __restrict__ tw, const cpx<T>* twl, int tid,
                                                F&& emit) {
  __builtin_assume((unsigned)tid < (unsigned)NTHR);
  if constexpr (sizeof...(Rest) > 0) {
    constexpr int TOT = (L / R) * NCOL;
    constexpr int ITER = (TOT + NTHR - 1) / NTHR;
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int b = tid + it * NTHR;
      if ((TOT % NTHR) == 0 || b < TOT) {
        cpx<T> v[R];
        const Group<LY, NCOL, S, R> gr(b);
        dit_butterfly<T, L, NCOL, SIGN, LY, TWL, 0, S, TWO, R>(buf, tw, twl, nullptr, gr, v);
        dit_store<T, NCOL, LY, S, R>(buf, v, gr);
      }
    }
    __syncthreads();
    dit_passes_emit<T, L, NCOL, NTHR, SIGN, LY, TWL, S * R, TWO + S * (R - 1), Rest...>(buf, tw, twl, tid, emit);
  } else {
    static_assert(S * R == L, "last pass");
    constexpr int TOT = (L / R) * NCOL;
    constexpr int ITER = (TOT + NTHR - 1) / NTHR;
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
      const int b = tid + it * NTHR;
      if ((TOT % NTHR) == 0 || b < TOT) {
        cpx<T> v[R];
        const Group<LY, NCOL, S, R> gr(b);
        dit_butterfly<T, L, NCOL, SIGN, LY, TWL, 0, S, TWO, R>(buf, tw, twl, nullptr, gr, v);
#pragma unroll
        for (int q = 0; q < R; ++q) emit(gr.col, gr.base + q * S, v[q]);
      }
    }
  }
}

// Whole transform: samples placed at digit-reversed positions, result in
// natural order.
template <typename T, int NCOL, int NTHR, int SIGN, typename LY, int... R>
__device__ __forceinline__ void fft_run(Plan<R...>, cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid) {
  dit_passes<T, Plan<R...>::L, NCOL, NTHR, SIGN, LY, 0, 0, 1, 0, R...>(buf, tw, nullptr, nullptr, tid);
}
// fft_run with the first TWL twiddle entries read from their LDS copy twl.
template <typename T, int NCOL, int NTHR, int SIGN, typename LY, int TWL, int... R>
__device__ __forceinline__ void fft_run_tw(Plan<R...>, cpx<T>* buf, const cpx<T>* __restrict__ tw, const cpx<T>* twl,
                                           int tid) {
  dit_passes<T, Plan<R...>::L, NCOL, NTHR, SIGN, LY, TWL, 0, 1, 0, R...>(buf, tw, twl, nullptr, tid);
}

// First pass in registers: the thread with butterfly index n0 < L/R0 of
// column col holds the samples x[n0 + (L/R0) q], q < R0.  Their
// digit-reversed positions are Plan::digitrev(n0) + q, so the radix-R0
// butterfly (no twiddles at sub-length 1) runs on them directly and stores
// its outputs there; no scatter through LDS is needed.
template <typename T, int SIGN, typename LY, int NCOL, int R0, int... Rest>
__device__ __forceinline__ void first_pass_store(Plan<R0, Rest...>, cpx<T>* v, cpx<T>* buf, int n0, int col = 0) {
  DFT<T, R0, SIGN>::run(v);
  // digitrev(n0) = R0 * (n0's position in the plan of the remaining radices)
  unsigned j = 0;
  if constexpr (sizeof...(Rest) > 0) j = (unsigned)Plan<Rest...>::digitrev(n0);
  dit_store<T, NCOL, LY, 1, R0>(buf, v, Group<LY, NCOL, 1, R0>(col, j, 0u));
}
// Remaining passes after first_pass_store (and a barrier).
template <typename T, int NTHR, int SIGN, typename LY, int R0, int... Rest>
__device__ __forceinline__ void fft_rest(Plan<R0, Rest...>, cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid) {
  dit_passes<T, Plan<R0, Rest...>::L, 1, NTHR, SIGN, LY, 0, 0, R0, R0 - 1, Rest...>(buf, tw, nullptr, nullptr, tid);
}
// fft_rest with the first TWL twiddle entries read from their LDS copy twl
// and (TWC) the last radix-8 pass's from the compact table twc.
template <typename T, int NTHR, int SIGN, typename LY, int TWL, int TWC, int R0, int... Rest>
__device__ __forceinline__ void fft_rest_tw(Plan<R0, Rest...>, cpx<T>* buf, const cpx<T>* __restrict__ tw,
                                            const cpx<T>* twl, const cpx<T>* twc, int tid) {
  dit_passes<T, Plan<R0, Rest...>::L, 1, NTHR, SIGN, LY, TWL, TWC, R0, R0 - 1, Rest...>(buf, tw, twl, twc, tid);
}
// Remaining passes after first_pass_store of NCOL interleaved columns (and a
// barrier): the butterflies of all columns are spread over the threads.
template <typename T, int NCOL, int NTHR, int SIGN, typename LY, int R0, int... Rest>
__device__ __forceinline__ void fft_rest_cols(Plan<R0, Rest...>, cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid) {
  dit_passes<T, Plan<R0, Rest...>::L, NCOL, NTHR, SIGN, LY, 0, 0, R0, R0 - 1, Rest...>(buf, tw, nullptr, nullptr, tid);
}
// Remaining passes after first_pass_store, the last one emitting.
template <typename T, int NCOL, int NTHR, int SIGN, typename LY, int R0, int... Rest, typename F>
__device__ __forceinline__ void fft_rest_emit(Plan<R0, Rest...>, cpx<T>* buf, const cpx<T>* __restrict__ tw, int tid,
                                              F&& emit) {
  dit_passes_emit<T, Plan<R0, Rest...>::L, NCOL, NTHR, SIGN, LY, 0, R0, R0 - 1, Rest...>(buf, tw, nullptr, tid, emit);
}
// fft_rest_emit with the first TWL twiddle entries read from their LDS copy twl.
template <typename T, int NCOL, int NTHR, int SIGN, typename LY, int TWL, int R0, int... Rest, typename F>
__device__ __forceinline__ void fft_rest_emit_tw(Plan<R0, Rest...>, cpx<T>* buf, const cpx<T>* __restrict__ tw,
                                                 const cpx<T>* twl, int tid, F&& emit) {
  dit_passes_emit<T, Plan<R0, Rest...>::L, NCOL, NTHR, SIGN, LY, TWL, R0, R0 - 1, Rest...>(buf, tw, twl, tid, emit);
}

}  // namespace nsp
