// jsd_kernels.h - device code of the channel-flow step (gfx950, FP64).
//
// Layout of every spectral field: (kx index i, y index l, kz index j) with
// kz fastest, complex128 as double2.  Stored wavenumbers are [0..K, -K..-1]
// in both periodic directions; K = M/2, KZ = Z/2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double dbl4 __attribute__((ext_vector_type(4)));

#define JSD_NTMAX 9      // 16-row MFMA tiles along y
#define JSD_NYMAX 132     // Chebyshev points (and 4-row K padding) per pencil
#define JSD_MAXRAD 8     // FFT stages

struct FftPlan {
  int p;               // length
  int nrad;            // number of stages
  int rad[JSD_MAXRAD]; // radices, innermost first
  const int* perm;     // LDS position -> input index (device)
  const int* iperm;    // input index -> LDS position (device)
  const double2* tw;   // forward twiddles exp(-2 pi i m / p) (device)
};

// ---------------------------------------------------------------------------
// small complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// (i k) * a  computed like numpy's (0 + i k)(a.x + i a.y)
__device__ __forceinline__ double2 cik(double k, double2 a) { return make_double2(-k * a.y, k * a.x); }
// Hermitian part 0.5 (a + conj(b)) where b is the value at -k
__device__ __forceinline__ double2 cherm(double2 a, double2 b) {
  return make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
}

// ---------------------------------------------------------------------------
// radix-R DFT on registers; sign = -1 forward, +1 inverse
// ---------------------------------------------------------------------------
template <int R> struct DftConst;
template <> struct DftConst<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.8660254037844386, -0.8660254037844386};
};
template <> struct DftConst<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475,
                                  0.30901699437494745};
  static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731,
                                  -0.9510565162951535};
};
template <> struct DftConst<7> {
  static constexpr double c[7] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191,
                                  -0.9009688679024191, -0.2225209339563144, 0.6234898018587335};
  static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581,
                                  -0.4338837391175581, -0.9749279121818236, -0.7818314824680298};
};

template <int R>
__device__ __forceinline__ void dft_generic(double2* v, double sign) {
  double2 out[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    double re = v[0].x, im = v[0].y;
#pragma unroll
    for (int m = 1; m < R; ++m) {
      const int e = (q * m) % R;
      const double c = DftConst<R>::c[e];
      const double s = sign * DftConst<R>::s[e];
      // v[m] * (c + i s)
      re += v[m].x * c - v[m].y * s;
      im += v[m].x * s + v[m].y * c;
    }
    out[q] = make_double2(re, im);
  }
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = out[q];
}

__device__ __forceinline__ void dft2(double2* v) {
  double2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
// radix 4: w = exp(sign * 2 pi i / 4) = sign * i
__device__ __forceinline__ void dft4(double2* v, double sign) {
  double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
  // b1 * (sign i)
  double2 b1r = make_double2(-sign * b1.y, sign * b1.x);
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1r);
  v[3] = csub(a1, b1r);
}
__device__ __forceinline__ void dft8(double2* v, double sign) {
  const double h = 0.7071067811865476;
  // split even / odd
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, sign);
  dft4(o, sign);
  // twiddles w8^q, q = 0..3, w8 = exp(sign 2 pi i / 8)
  double2 t1 = make_double2(h * (o[1].x - sign * o[1].y), h * (o[1].y + sign * o[1].x));
  double2 t2 = make_double2(-sign * o[2].y, sign * o[2].x);
  double2 t3 = make_double2(h * (-o[3].x - sign * o[3].y), h * (-o[3].y + sign * o[3].x));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], t1);
  v[5] = csub(e[1], t1);
  v[2] = cadd(e[2], t2);
  v[6] = csub(e[2], t2);
  v[3] = cadd(e[3], t3);
  v[7] = csub(e[3], t3);
}

// One DIT stage of radix R on nrows rows held in LDS.
// element (row, pos) lives at buf[pos * ps + row * rs].
template <int R>
__device__ void fft_stage(double2* buf, int nrows, int ps, int rs, int p, int L, const double2* __restrict__ tw,
                          double sign) {
  const int Lp = L / R;
  const int nb = p / R;
  const int stride = p / L;
  for (int w = threadIdx.x; w < nb * nrows; w += blockDim.x) {
    const int row = w % nrows;
    const int b = w / nrows;
    const int g = b / Lp, k1 = b - g * Lp;
    const int base = g * L + k1;
    double2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      double2 x = buf[(base + m * Lp) * ps + row * rs];
      if (m > 0 && k1 > 0) {
        // exp(sign 2 pi i m k1 / L); tw holds exp(-2 pi i idx / p)
        double2 t = tw[(m * k1 * stride) % p];
        if (sign > 0) t.y = -t.y;
        x = cmul(x, t);
      }
      v[m] = x;
    }
    if (R == 2) dft2(v);
    else if (R == 4) dft4(v, sign);
    else if (R == 8) dft8(v, sign);
    else if (R == 3) dft_generic<3>(v, sign);
    else if (R == 5) dft_generic<5>(v, sign);
    else if (R == 7) dft_generic<7>(v, sign);
#pragma unroll
    for (int m = 0; m < R; ++m) buf[(base + m * Lp) * ps + row * rs] = v[m];
  }
}

// Full in-place FFT of nrows rows (input already in DIT order).
__device__ void fft_lds(double2* buf, int nrows, int ps, int rs, const FftPlan& P, double sign) {
  int L = 1;
  for (int s = 0; s < P.nrad; ++s) {
    const int R = P.rad[s];
    L *= R;
    switch (R) {
      case 2: fft_stage<2>(buf, nrows, ps, rs, P.p, L, P.tw, sign); break;
      case 3: fft_stage<3>(buf, nrows, ps, rs, P.p, L, P.tw, sign); break;
      case 4: fft_stage<4>(buf, nrows, ps, rs, P.p, L, P.tw, sign); break;
      case 5: fft_stage<5>(buf, nrows, ps, rs, P.p, L, P.tw, sign); break;
      case 7: fft_stage<7>(buf, nrows, ps, rs, P.p, L, P.tw, sign); break;
      case 8: fft_stage<8>(buf, nrows, ps, rs, P.p, L, P.tw, sign); break;
      default: break;
    }
    __syncthreads();
  }
}
