// jsd_fft_ct.h - LDS FFT with the transform length and radices fixed at
// compile time (gfx950, FP64).
//
// Same algorithm and arithmetic as the run-time fft_lds of jsd_kernels.h
// (in-place mixed-radix decimation in time, input loaded in digit-reversed
// order, correctly-rounded twiddle table exp(-2 pi i m / p)); with every
// length, stride and radix a constant the index arithmetic reduces to
// multiply-shift sequences, the twiddle index needs no modulo
// (m * k1 * stride < p) and each stage's work loop has a constant trip count.
#pragma once
#include "jsd_kernels.h"

// radices of the stages, innermost (first DIT stage) first; 1 = no stage
template <int R0, int R1 = 1, int R2 = 1, int R3 = 1>
struct Rad {
  static constexpr int P = R0 * R1 * R2 * R3;
  static constexpr int NS = 1 + (R1 > 1) + (R2 > 1) + (R3 > 1);
  static constexpr int r(int s) { return s == 0 ? R0 : s == 1 ? R1 : s == 2 ? R2 : R3; }
  static constexpr int len(int k) { return k <= 0 ? 1 : r(k - 1) * len(k - 1); }  // product of r(0..k-1)
};

// LDS position -> input index of the DIT order over stages 0..K-1:
// order(len) = [r * sub[j] + n2 for n2 in 0..r-1 for j], r the outermost radix.
template <class C, int K>
__device__ __forceinline__ int dit_perm(int q) {
  if constexpr (K == 0) {
    return 0;
  } else {
    constexpr int r = C::r(K - 1);
    constexpr int Q = C::len(K - 1);
    return r * dit_perm<C, K - 1>(q % Q) + q / Q;
  }
}
// input index -> LDS position
template <class C, int K>
__device__ __forceinline__ int dit_iperm(int n) {
  if constexpr (K == 0) {
    return 0;
  } else {
    constexpr int r = C::r(K - 1);
    constexpr int Q = C::len(K - 1);
    return (n % r) * Q + dit_iperm<C, K - 1>(n / r);
  }
}

// Radix 3 and 5 through the symmetric sums and differences of the inputs
// (v_m + v_{R-m}, v_m - v_{R-m}): out_q and out_{R-q} share the cosine part
// and differ in the sign of the sine part, 12 real operations for radix 3
// and 36 for radix 5 instead of the 16 and 80 of dft_generic.  The constants
// are the correctly rounded cos / sin of 2 pi m / R of DftConst.
template <int SIGN>
__device__ __forceinline__ void dft3_sym(double2* v) {
  constexpr double c1 = DftConst<3>::c[1], s1 = (double)SIGN * DftConst<3>::s[1];
  const double2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
  const double2 a = make_double2(v[0].x + c1 * s.x, v[0].y + c1 * s.y);
  const double2 b = make_double2(-s1 * d.y, s1 * d.x);  // i sign sin(2 pi / 3) d
  v[0] = cadd(v[0], s);
  v[1] = cadd(a, b);
  v[2] = csub(a, b);
}
template <int SIGN>
__device__ __forceinline__ void dft5_sym(double2* v) {
  constexpr double c1 = DftConst<5>::c[1], c2 = DftConst<5>::c[2];
  constexpr double s1 = (double)SIGN * DftConst<5>::s[1], s2 = (double)SIGN * DftConst<5>::s[2];
  const double2 p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
  const double2 p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
  const double2 a1 = make_double2(v[0].x + c1 * p1.x + c2 * p2.x, v[0].y + c1 * p1.y + c2 * p2.y);
  const double2 a2 = make_double2(v[0].x + c2 * p1.x + c1 * p2.x, v[0].y + c2 * p1.y + c1 * p2.y);
  // i (s1 m1 + s2 m2) and i (s2 m1 - s1 m2)
  const double2 b1 = make_double2(-(s1 * m1.y + s2 * m2.y), s1 * m1.x + s2 * m2.x);
  const double2 b2 = make_double2(-(s2 * m1.y - s1 * m2.y), s2 * m1.x - s1 * m2.x);
  v[0] = cadd(cadd(v[0], p1), p2);
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_ct(double2* v) {
  constexpr double sign = (double)SIGN;
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3_sym<SIGN>(v);
  else if constexpr (R == 4) dft4(v, sign);
  else if constexpr (R == 5) dft5_sym<SIGN>(v);
  else if constexpr (R == 8) dft8(v, sign);
  else dft_generic<R>(v, sign);
}

// One DIT stage (radix R, partial length L) on NROWS rows;
// element (row, pos) at buf[pos * PS + row * RS].
template <int R, int P, int L, int NROWS, int PS, int RS, int NTH, int SIGN>
__device__ __forceinline__ void fft_stage_ct(double2* buf, const double2* __restrict__ tw) {
  constexpr int Lp = L / R;
  constexpr int nb = P / R;
  constexpr int stride = P / L;
  constexpr int total = nb * NROWS;
#pragma unroll 2
  for (int w = threadIdx.x; w < total; w += NTH) {
    const int row = w % NROWS;
    const int b = w / NROWS;
    const int g = b / Lp, k1 = b - g * Lp;
    const int base = g * L + k1;
    double2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      double2 x = buf[(base + m * Lp) * PS + row * RS];
      if (Lp > 1 && m > 0 && k1 > 0) {
        double2 t = tw[m * k1 * stride];
        if (SIGN > 0) t.y = -t.y;
        x = cmul(x, t);
      }
      v[m] = x;
    }
    dft_ct<R, SIGN>(v);
#pragma unroll
    for (int m = 0; m < R; ++m) buf[(base + m * Lp) * PS + row * RS] = v[m];
  }
}

template <class C, int S, int NROWS, int PS, int RS, int NTH, int SIGN>
__device__ __forceinline__ void fft_stages_ct(double2* buf, const double2* __restrict__ tw) {
  if constexpr (S < C::NS) {
    fft_stage_ct<C::r(S), C::P, C::len(S + 1), NROWS, PS, RS, NTH, SIGN>(buf, tw);
    __syncthreads();
    fft_stages_ct<C, S + 1, NROWS, PS, RS, NTH, SIGN>(buf, tw);
  }
}

// 16-point DFT in registers, natural order in and out (n = n1 + 4 n2,
// k = 4 k1 + k2: four radix-4 DFTs, twiddles w16^(n1 k2), four radix-4 DFTs).
// Twiddles: cos / sin of 2 pi m / 16, correctly rounded.
struct W16 {
  static constexpr double c[16] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898,
                                   0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867,
                                   -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898,
                                   0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867};
  static constexpr double s[16] = {0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867,
                                   1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898,
                                   0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867,
                                   -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898};
};

template <int SIGN>
__device__ __forceinline__ void dft16(double2* v) {
  constexpr double sign = (double)SIGN;
  double2 a[4][4];  // a[n1][k2]
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    double2 t[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    dft4(t, sign);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      const int m = n1 * k2;
      if (m == 0) a[n1][k2] = t[k2];
      else a[n1][k2] = cmul(t[k2], make_double2(W16::c[m], sign * W16::s[m]));
    }
  }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    double2 t[4] = {a[0][k2], a[1][k2], a[2][k2], a[3][k2]};
    dft4(t, sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[4 * k1 + k2] = t[k1];
  }
}

// dft16 for inputs with v[5..11] == 0 (v[4] may be nonzero): the spectrum of
// at most 129 stored modes (|k| <= 64) zero-padded to 256 and split as
// n = n1 + 16 n2.  The first-stage 4-point DFTs drop the additions of the
// known zeros (x + 0 and 0 - x are exact, so the result is bit-identical).
template <int SIGN>
__device__ __forceinline__ void dft16_pz(double2* v) {
  constexpr double sign = (double)SIGN;
  double2 a[4][4];  // a[n1][k2]
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    const double2 t0 = v[n1], t3 = v[n1 + 12];
    double2 b0, b1;
    if (n1 == 0) {
      b0 = cadd(v[4], t3);
      b1 = csub(v[4], t3);
    } else {
      b0 = t3;
      b1 = make_double2(-t3.x, -t3.y);
    }
    const double2 b1r = make_double2(-sign * b1.y, sign * b1.x);
    double2 t[4] = {cadd(t0, b0), cadd(t0, b1r), csub(t0, b0), csub(t0, b1r)};
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      const int m = n1 * k2;
      if (m == 0) a[n1][k2] = t[k2];
      else a[n1][k2] = cmul(t[k2], make_double2(W16::c[m], sign * W16::s[m]));
    }
  }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    double2 t[4] = {a[0][k2], a[1][k2], a[2][k2], a[3][k2]};
    dft4(t, sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[4 * k1 + k2] = t[k1];
  }
}

// Full in-place FFT of NROWS rows already in DIT order; SIGN -1 forward, +1 inverse.
template <class C, int NROWS, int PS, int RS, int NTH, int SIGN>
__device__ __forceinline__ void fft_lds_ct(double2* buf, const double2* __restrict__ tw) {
  fft_stages_ct<C, 0, NROWS, PS, RS, NTH, SIGN>(buf, tw);
}
