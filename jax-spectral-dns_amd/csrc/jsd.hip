// jsd.hip - MI355X (gfx950) FP64 velocity-vorticity channel-flow step.
//
// One RK3/CN substep is a short pipeline of kernels over the spectral state
// (kx, y, kz layout, complex128):
//   1. k_ygemm<PLAIN>  : y-derivatives D u_c on FP64 MFMA (16x16x4 tiles).
//   2a. nonlinear FFT path (NS, nonlinear perturbation):
//        k_fft_x_inv  : Hermitian part + zero-pad + inverse x-FFT of two
//                       packed real fields per launch,
//        k_fft_z_mid  : inverse z-FFT of all packed fields, skew-symmetric
//                       products in physical space, forward z-FFT,
//        k_fft_x_fwd  : forward x-FFT, unpacking of the real pair,
//                       truncation with the reference's +-K convention.
//   2b. linearised path: k_nl_lin computes the same H spectrum without any
//        FFT (products with a y-only base flow are diagonal in (kx, kz)).
//   3. k_ygemm<HV>     : h_v = -D(ikx H0 + ikz H2) - k^2 H1, h_g, conv lines.
//   4. k_solve         : per-(|kx|,|kz|) implicit Helmholtz/Poisson solves,
//                        influence-matrix wall correction, u/w from
//                        continuity, (0,0) mean-flow block.
// Reference: jax_spectral_dns navier_stokes.py:1219-1775 (step),
// navier_stokes.py:48-261 and navier_stokes_perturbation*.py (NL terms),
// domain.py:716-764 / 1167-1242 (transforms).
#include "jsd_fft_ct.h"
#include "jsd_kernels.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/jsd.h"

#define JSD_NPROF 7  // kernel classes: ygemm, fft_x_inv, fft_z_mid, fft_x_fwd, ygemm_hv, solve, nl_lin

// ===========================================================================
// y-derivative GEMM on FP64 MFMA
// ===========================================================================
// For one kx row i the field is an (N x 2Z) real matrix X (rows y, columns
// kz re/im interleaved); out = D^o X.  Tiles: 16 rows (output y) x 16 columns
// per MFMA; the wave owns one 16-column strip and all row tiles; the WG
// (4 waves) owns 64 columns.  D^o is pre-tiled on the host into MFMA A
// fragments: DF[(lt*KS + ks)*64 + lane] = D^o[lt*16 + lane%16][ks*4 + lane/16].
enum { GEMM_PLAIN = 0, GEMM_HV = 1, GEMM_HVT = 2, GEMM_ACC = 3 };
#define GEMM_GK 8  // k-steps per LDS-staged chunk of A fragments

// Logical block of a launch: the single-run kernels pass blockIdx, the
// batched kernels (one launch for many runs) pass the block within its run.
struct Blk {
  int x, y, z;
  int nx;
};

struct YgemmArgs {
  const double* DF;  // D^o pre-tiled (NT row tiles x KS k-steps x 64 lanes)
  int NT, KS;
  const double* in0;  // PLAIN: input; HV: H0; HVT: hvbar; ACC: input
  const double* in1;  // HV: H1; HVT: hgbar
  const double* in2;  // HV: H2
  double* out0;       // PLAIN: out; HV: h_v; HVT: Hbar0; ACC: accumulated output
  double* out1;       // HV: h_g; HVT: Hbar1
  double* out2;       // HVT: Hbar2
  double* conv;       // HV: conv lines (2*N complex); HVT: convbar
  const double* kx;
  const double* kz;
  int M, N, Z;
  long comp_stride;
};

template <int MODE>
__device__ void ygemm_body(const YgemmArgs& G, Blk blk) {
  const double* __restrict__ DF = G.DF;
  const int NT = G.NT, KS = G.KS;
  const double* __restrict__ in0 = G.in0;
  const double* __restrict__ in1 = G.in1;
  const double* __restrict__ in2 = G.in2;
  double* __restrict__ out0 = G.out0;
  double* __restrict__ out1 = G.out1;
  double* __restrict__ out2 = G.out2;
  double* __restrict__ conv = G.conv;
  const double* __restrict__ kx = G.kx;
  const double* __restrict__ kz = G.kz;
  const int N = G.N, Z = G.Z;
  const long comp_stride = G.comp_stride;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int i = blk.x;
  const int comp = blk.z;
  const int C2 = 2 * Z;
  const int col = blk.y * 64 + wave * 16 + (lane & 15);
  const bool colok = col < C2;
  const int kr = lane >> 4;
  const size_t rowbase = (size_t)comp * comp_stride + (size_t)i * N * C2;
  const double kxv = kx[i];
  const int kzi = col >> 1;
  const double kzv = colok ? kz[kzi] : 0.0;
  const int par = col & 1;

  dbl4 acc[JSD_NTMAX];
#pragma unroll
  for (int t = 0; t < JSD_NTMAX; ++t) acc[t] = (dbl4){0.0, 0.0, 0.0, 0.0};

  auto loadB = [&](int ks) -> double {
    const int j = ks * 4 + kr;
    if (!colok || j >= N) return 0.0;
    const size_t e = rowbase + (size_t)j * C2;
    if (MODE != GEMM_HV) return in0[e + col];
    // X = (i kx) H0 + (i kz) H2 ; (i k)(a + ib) = (-k b, k a)
    const size_t c0 = e + (kzi << 1);
    if (par == 0) return (-kxv * in0[c0 + 1]) + (-kzv * in2[c0 + 1]);
    return (kxv * in0[c0]) + (kzv * in2[c0]);
  };

  // A fragments of GK consecutive k-steps for all row tiles are staged in LDS
  // once per workgroup (shared by its 4 waves); B values of the next chunk
  // are loaded while the current chunk's MFMAs run.
  __shared__ double sA[JSD_NTMAX * GEMM_GK * 64];
  double bcur[GEMM_GK], bnext[GEMM_GK];
#pragma unroll
  for (int kk = 0; kk < GEMM_GK; ++kk) bcur[kk] = loadB(kk);
  for (int kc = 0; kc < KS; kc += GEMM_GK) {
    __syncthreads();
    for (int it = threadIdx.x; it < NT * GEMM_GK * 64; it += blockDim.x) {
      const int lt = it / (GEMM_GK * 64);
      const int rem = it - lt * (GEMM_GK * 64);
      const int kk = rem >> 6;
      sA[it] = (kc + kk < KS) ? DF[((size_t)lt * KS + kc + kk) * 64 + (rem & 63)] : 0.0;
    }
#pragma unroll
    for (int kk = 0; kk < GEMM_GK; ++kk) bnext[kk] = loadB(kc + GEMM_GK + kk);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GEMM_GK; ++kk) {
#pragma unroll
      for (int lt = 0; lt < JSD_NTMAX; ++lt) {
        if (lt < NT) {
          const double a = sA[(lt * GEMM_GK + kk) * 64 + lane];
          acc[lt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bcur[kk], acc[lt], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int kk = 0; kk < GEMM_GK; ++kk) bcur[kk] = bnext[kk];
  }
  if (MODE == GEMM_HVT) {
    // transpose of the h_v / h_g / conv assembly: acc = D^T hvbar, Xbar = -acc;
    // Hbar0 = conj(i kx) Xbar + conj(i kz) hgbar (- convbar0), Hbar1 = (-kx^2 - kz^2) hvbar,
    // Hbar2 = conj(i kz) Xbar + conj(-i kx) hgbar (- convbar2).  The other
    // component of Xbar sits in the neighbouring lane (column col ^ 1).
#pragma unroll
    for (int lt = 0; lt < JSD_NTMAX; ++lt) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double own = -acc[lt][r];
        const double oth = -__shfl_xor(acc[lt][r], 1);
        const int l = lt * 16 + r * 4 + kr;
        if (lt >= NT || l >= N || !colok) continue;
        const double xr = par == 0 ? own : oth, xi = par == 0 ? oth : own;
        const size_t e = rowbase + (size_t)l * C2;
        const size_t c0 = e + (kzi << 1);
        const double gr = in1[c0], gi = in1[c0 + 1];
        double h0, h2;
        if (par == 0) {
          h0 = kxv * xi + kzv * gi;
          h2 = kzv * xi - kxv * gi;
        } else {
          h0 = -kxv * xr - kzv * gr;
          h2 = -kzv * xr + kxv * gr;
        }
        if (i == 0 && kzi == 0) {
          h0 -= conv[2 * l + par];
          h2 -= conv[2 * (N + l) + par];
        }
        const double hv = in0[e + col];
        out0[e + col] = h0;
        out1[e + col] = (-(kxv * kxv)) * hv + (-(kzv * kzv)) * hv;
        out2[e + col] = h2;
      }
    }
    return;
  }
  if (!colok) return;
#pragma unroll
  for (int lt = 0; lt < JSD_NTMAX; ++lt) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int l = lt * 16 + r * 4 + kr;  // f64 16x16x4: reg r of lane L is row 4r + L/16
      if (lt >= NT || l >= N) continue;
      const size_t e = rowbase + (size_t)l * C2;
      if (MODE == GEMM_PLAIN) {
        out0[e + col] = acc[lt][r];
      } else if (MODE == GEMM_ACC) {
        out0[e + col] += acc[lt][r];
      } else {
        // h_v = (-D X + (-kx^2) H1) + (-kz^2) H1   (navier_stokes.py:53-61)
        const double h1 = in1[e + col];
        out0[e + col] = (-acc[lt][r] + (-(kxv * kxv)) * h1) + (-(kzv * kzv)) * h1;
        // h_g = (i kz) H0 - (i kx) H2               (navier_stokes.py:62-64)
        const size_t c0 = e + (kzi << 1);
        double hg;
        if (par == 0) hg = (-kzv * in0[c0 + 1]) - (-kxv * in2[c0 + 1]);
        else hg = (kzv * in0[c0]) - (kxv * in2[c0]);
        out1[e + col] = hg;
        // conv_ns = -H on the (0,0) line for the mean-flow solve
        if (i == 0 && kzi == 0) {
          conv[2 * l + par] = -in0[e + col];
          conv[2 * (N + l) + par] = -in2[e + col];
        }
      }
    }
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_ygemm(YgemmArgs G) {
  ygemm_body<MODE>(G, Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x});
}

// ===========================================================================
// y-derivative on even/odd parity blocks
// ===========================================================================
// D is centro-antisymmetric (D[N-1-i][N-1-j] = -D[i][j]) and D^2
// centro-symmetric.  With e_h = (x_h + x_h')/2, o_h = (x_h - x_h')/2
// (h' = N-1-h, h < ceil(N/2)) the product splits into two blocks of at most
// 66 x 66: D gives even rows from the odd part and odd rows from the even
// part, D^2 even from even and odd from odd; half the MFMA work of the full
// Ny x Ny product.  Columns are the flattened (kx row, kz re/im) index, so no
// workgroup is left with a ragged tail; a wave owns 16 columns and all row
// tiles; each block's A fragments are staged in LDS once per workgroup.
#define YP_TMAX 5   // 16-row tiles per half (N <= 132)
#define YP_KMAX 17  // 4-row k-steps per half

struct YparArgs {
  const double* F;  // fragments: even-output block (ET tiles x KIE k-steps), then odd-output (OT x KIO)
  int ET, OT, KIE, KIO, KE, KO;  // tiles per output half; k-steps feeding each; k-steps of the even / odd input
  int nEf, nOf;                  // half sizes (walls included): ceil(N/2), floor(N/2)
  int anti;                      // 1: D (parities swap), 0: D^2
  const double* in;
  double* out;
  int M, N, Z;
  long comp_stride;  // doubles between components
  int acc;           // 1: out += the product (the adjoint's D^T accumulation); 0: out = the product
};

// Block bx of 32 flattened columns, component bz.  Waves 0, 1 compute the
// even output rows of columns 0..15 / 16..31, waves 2, 3 the odd rows of
// the same columns; each wave reads its block's A fragments straight from
// L2 (prefetched one k-step ahead) and its input parity from HBM.  The odd
// halves pass through LDS (sO: 80 x 32 doubles) to the even waves, which
// write out[h] = e + o and out[N-1-h] = e - o.
__device__ __forceinline__ void ypar_body(const YparArgs& G, int bx, int bz, double* sO) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kr = lane >> 4, cc = lane & 15;
  const int half = wave >> 1;               // 0: even output rows, 1: odd
  const int lc = (wave & 1) * 16 + cc;      // column within the block
  const int N = G.N, C2 = 2 * G.Z;
  const long gcol = (long)bx * 32 + lc;
  const bool ok = gcol < (long)G.M * C2;
  const long i = ok ? gcol / C2 : 0;
  const int c = ok ? (int)(gcol - i * C2) : 0;
  const size_t base = (size_t)bz * G.comp_stride + (size_t)i * N * C2 + c;
  const bool anti = G.anti != 0;
  // input parity feeding this half: D swaps parities, D^2 keeps them
  const int qin = anti ? 1 - half : half;
  const int T = half ? G.OT : G.ET;
  const int K = half ? G.KIO : G.KIE;
  const double* __restrict__ F = G.F + (half ? (size_t)G.ET * G.KIE * 64 : 0) + lane;
  double bx_[YP_KMAX];
#pragma unroll
  for (int ks = 0; ks < YP_KMAX; ++ks) {
    const int h = 4 * ks + kr;
    double v = 0.0;
    if (ok && ks < K && h < (qin ? G.nOf : G.nEf)) {
      const double xa = G.in[base + (size_t)h * C2];
      const int hm = N - 1 - h;
      if (hm == h) v = qin ? 0.0 : xa;
      else {
        const double xb = G.in[base + (size_t)hm * C2];
        v = qin ? 0.5 * (xa - xb) : 0.5 * (xa + xb);
      }
    }
    bx_[ks] = v;
  }
  dbl4 acc[YP_TMAX];
#pragma unroll
  for (int t = 0; t < YP_TMAX; ++t) acc[t] = (dbl4){0.0, 0.0, 0.0, 0.0};
  double an[YP_TMAX];
#pragma unroll
  for (int t = 0; t < YP_TMAX; ++t) an[t] = (t < T && K > 0) ? F[(size_t)(t * K) * 64] : 0.0;
#pragma unroll
  for (int ks = 0; ks < YP_KMAX; ++ks) {
    if (ks < K) {
      double a[YP_TMAX];
#pragma unroll
      for (int t = 0; t < YP_TMAX; ++t) a[t] = an[t];
#pragma unroll
      for (int t = 0; t < YP_TMAX; ++t) an[t] = (t < T && ks + 1 < K) ? F[(size_t)(t * K + ks + 1) * 64] : 0.0;
#pragma unroll
      for (int t = 0; t < YP_TMAX; ++t)
        if (t < T) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], bx_[ks], acc[t], 0, 0, 0);
    }
  }
  if (half == 1) {
#pragma unroll
    for (int t = 0; t < YP_TMAX; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int h = t * 16 + 4 * r + kr;
        if (t < T) sO[h * 32 + lc] = acc[t][r];
      }
  }
  __syncthreads();
  if (half == 1 || !ok) return;
#pragma unroll
  for (int t = 0; t < YP_TMAX; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int h = t * 16 + 4 * r + kr;  // f64 16x16x4: reg r of lane L is row 4r + L/16
      if (t >= T || h >= G.nEf) continue;
      const double e = acc[t][r];
      const double o = (h < G.nOf) ? sO[h * 32 + lc] : 0.0;
      if (G.acc) {
        G.out[base + (size_t)h * C2] += e + o;
        if (N - 1 - h != h) G.out[base + (size_t)(N - 1 - h) * C2] += e - o;
      } else {
        G.out[base + (size_t)h * C2] = e + o;
        if (N - 1 - h != h) G.out[base + (size_t)(N - 1 - h) * C2] = e - o;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_ypar(YparArgs G) {
  __shared__ double sO[YP_TMAX * 16 * 32];
  ypar_body(G, blockIdx.x, blockIdx.z, sO);
}

// Streaming form of the same product for the plain launches.  A persistent
// grid of 512-thread workgroups stages the operator in LDS once, then every
// wave walks its own sequence of 16-column units (flattened over components,
// kx rows and kz re/im) with no further barrier.  A unit computes both output
// halves from one read of its rows: with x_h, x_h' read for h < ceil(N/2),
// e = (x_h + x_h')/2 and o = (x_h - x_h')/2 feed the even and odd blocks
// (0.5 (x + x) = x exactly at the centre row, where o = 0).
//
// Operator layout (YsArgs::F, built by ystream_frags): for each of the KL
// k-steps, 8 MFMA fragments (even-output tiles 0..3, odd-output tiles 0..3,
// rows 0..63, zero where a block has no such tile or k-step), then after all
// k-steps the coefficients of row 64 of each half (Ny = 129, 130: a VALU dot
// product whose four per-row-group partial sums meet by lane shuffles instead
// of a fifth tile of 15/16 padding), [ks][half][kr].
//
// The k-steps of a wave's units form one stream: the row pairs YS_PD k-steps
// ahead, across unit boundaries, are in flight while the current k-step's
// MFMAs run, so HBM reads overlap the products.
#define YS_PD 4
#define YS_THREADS 512

struct YsArgs {
  const double* F;  // ystream_frags layout
  int KL;           // k-steps of a unit: ceil(ceil(N/2) / 4)
  int nEf, nOf;     // half sizes
  int xE, xO;       // row 64 present in the even / odd output half
  int anti;
  const double* in;
  double* out;
  int M, N, Z;
  int comp_stride;  // doubles between components
  int ncol;         // flattened columns over all components
  const double* zero;  // one 0.0 in device memory (the read of a lane outside the data)
};

__host__ __device__ inline size_t ystream_lds_doubles(int KL) { return (size_t)KL * (8 * 64 + 8); }

// FULL: ceil(N/2) >= 64, every tile row is a stored row (the bench sizes).
// The loop runs whole groups of YS_PD k-steps (the ring slots are fixed
// registers), rounded up: the extra positions read zeros and store nothing.
// ACC: out += the product (the adjoint's u_bar += D^T Du_bar); the centre
// row of odd N (h = N-1-h) is added once.
template <bool ANTI, bool FULL, bool ACC = false>
__global__ __launch_bounds__(YS_THREADS, 4) void k_ypar_stream(YsArgs G) {
  extern __shared__ double sF[];
  const int KL = G.KL;
  const int nlds = (int)ystream_lds_doubles(KL);
  for (int w = threadIdx.x; w < nlds; w += YS_THREADS) sF[w] = G.F[w];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kr = lane >> 4, cc = lane & 15;
  const int N = G.N, C2 = 2 * G.Z, nEf = G.nEf, nOf = G.nOf, ncol = G.ncol;
  const int MC2 = G.M * C2;
  const int nunit = (ncol + 15) / 16;
  const int W = gridDim.x * (YS_THREADS / 64);
  // Unit of round r of this wave.  Round 0 is block-major (a workgroup's
  // eight waves take eight neighbouring units, whose rows share 128-byte
  // lines, in one XCD's L2).  Later rounds hand out chunks of four
  // neighbouring units to waves 0-3 of every workgroup in turn, then to
  // waves 4-7, so a partial last round spreads over the SIMDs of all CUs
  // instead of doubling the work of the first workgroups' CUs.
  const int blk = blockIdx.x, G8 = gridDim.x;
  auto unit_of = [&](int r) -> int {
    if (r == 0) return blk * 8 + wave;
    return r * W + 4 * (G8 * (wave >> 2) + blk) + (wave & 3);
  };
  const int u0 = unit_of(0);
  const double* __restrict__ in = G.in;
  double* __restrict__ out = G.out;
  const double* __restrict__ sC = sF + (size_t)KL * 8 * 64 + kr;
  // element offset of this lane's column in unit u (its y = +1 row), or -1
  auto col_base = [&](int u) -> int {
    const int q = u * 16 + cc;
    if (u >= nunit || q >= ncol) return -1;
    const int comp = q / MC2, rem = q - comp * MC2;
    const int i = rem / C2;
    return comp * G.comp_stride + i * N * C2 + (rem - i * C2);
  };
  // rows 4 ks + kr and N-1-(4 ks + kr).  Branch-free: a lane outside the
  // data reads a zero word instead, so every position issues the same two
  // loads.  dep, dep2 (e and o, the two values computed from the slot being
  // refilled) enter the addresses through an empty asm, so the refill cannot
  // be scheduled before the slot's last use and takes the slot's registers:
  // no copies of in-flight values at the loop edge (which would wait for the
  // newest loads), and the waits before the MFMAs leave the newer prefetches
  // in flight.
  auto fetch = [&](int base, int ks, double dep, double dep2, double& xa, double& xb) {
    const bool v = base >= 0 && 4 * ks + kr < nEf;
    int oa = base + kr * C2 + 4 * ks * C2, ob = base + (N - 1 - kr) * C2 - 4 * ks * C2;
    __asm__("" : "+v"(oa), "+v"(ob) : "v"(dep), "v"(dep2));
    xa = *(v ? in + oa : G.zero);
    xb = *(v ? in + ob : G.zero);
  };
  const int nfull = nunit / W;
  const int nu = nfull + (unit_of(nfull) < nunit ? 1 : 0);
  const int npos = ((nu * KL + YS_PD - 1) / YS_PD) * YS_PD;
  // prefetch cursor (unit base, k-step) and compute cursor
  int fr = 0, fb = col_base(u0), fk = 0;
  double ra[YS_PD], rb[YS_PD];
#pragma unroll
  for (int j = 0; j < YS_PD; ++j) {
    fetch(fb, fk, 0.0, 0.0, ra[j], rb[j]);
    if (++fk == KL) {
      fk = 0;
      fb = col_base(unit_of(++fr));
    }
  }
  int cr = 0, cb = col_base(u0), ck = 0;
  dbl4 ae[4], ao[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) ae[t] = ao[t] = (dbl4){0.0, 0.0, 0.0, 0.0};
  double pe = 0.0, po = 0.0;
  for (int p = 0; p < npos; p += YS_PD) {
#pragma unroll
    for (int j = 0; j < YS_PD; ++j) {
      const double e = 0.5 * (ra[j] + rb[j]), o = 0.5 * (ra[j] - rb[j]);
      fetch(fb, fk, e, o, ra[j], rb[j]);
      if (++fk == KL) {
        fk = 0;
        fb = col_base(unit_of(++fr));
      }
      const double inE = ANTI ? o : e, inO = ANTI ? e : o;
      const double* __restrict__ f = sF + (size_t)ck * 512 + lane;
#pragma unroll
      for (int t = 0; t < 4; ++t) ae[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(f[t * 64], inE, ae[t], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < 4; ++t) ao[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(f[(4 + t) * 64], inO, ao[t], 0, 0, 0);
      pe = fma(sC[ck * 8], inE, pe);
      po = fma(sC[ck * 8 + 4], inO, po);
      if (++ck == KL) {
        // unit done: rows 0..63 from the tiles, row 64 from the dot products
        pe += __shfl_xor(pe, 16);
        pe += __shfl_xor(pe, 32);
        po += __shfl_xor(po, 16);
        po += __shfl_xor(po, 32);
        if (cb >= 0) {
          // row h = 16 t + 4 r + kr (f64 16x16x4: reg r of lane L is row
          // 4r + L/16) and its mirror N-1-h: lane pointers plus offsets that
          // are the same for every lane
          double* __restrict__ lp = out + cb + kr * C2;
          double* __restrict__ lm = out + cb + (N - 1 - kr) * C2;
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int h0 = t * 16 + 4 * r, h = h0 + kr;
              if (!FULL && h >= nEf) continue;
              const double ev = ae[t][r];
              const double od = h < nOf ? ao[t][r] : 0.0;
              // the centre row (h = N-1-h) has od = 0: both stores agree
              if (ACC) {
                lp[h0 * C2] += ev + od;
                if (2 * h != N - 1) lm[-h0 * C2] += ev - od;
              } else {
                lp[h0 * C2] = ev + od;
                lm[-h0 * C2] = ev - od;
              }
            }
          if (G.xE && kr == 0) {
            const double od = G.xO ? po : 0.0;
            if (ACC) {
              lp[64 * C2] += pe + od;
              if (2 * 64 != N - 1) lm[-64 * C2] += pe - od;
            } else {
              lp[64 * C2] = pe + od;
              lm[-64 * C2] = pe - od;
            }
          }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) ae[t] = ao[t] = (dbl4){0.0, 0.0, 0.0, 0.0};
        pe = po = 0.0;
        ck = 0;
        cb = col_base(unit_of(++cr));
      }
    }
  }
}

// ===========================================================================
// FFT passes (nonlinear path)
// ===========================================================================
// A WG of the x-passes owns QB kz "pairs" q = q0..q0+QB-1 (q <= KZ); slot c
// (0..2QB-1) holds column j = q (c < QB) or j = Z - q (c >= QB, q >= 1).
// The partner column -kz of slot c is slot c ^ QB (slot c itself for q = 0).
// Streaming accesses (nontemporal: no reuse expected from the caches) for the
// z pass's intermediates: each T row is read once and each V row written once.
typedef double jsd_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
  jsd_d2v t = {v.x, v.y};
  __builtin_nontemporal_store(t, (jsd_d2v*)p);
}
__device__ __forceinline__ double2 ld_nt(const double2* p) {
  jsd_d2v t = __builtin_nontemporal_load((const jsd_d2v*)p);
  return make_double2(t.x, t.y);
}

#define QB 4
#define NSLOT (2 * QB)

__device__ __forceinline__ int slot_col(int c, int q0, int Z, int KZ, bool* ok) {
  const int q = q0 + (c % QB);
  const bool neg = c >= QB;
  *ok = (q <= KZ) && (!neg || q >= 1);
  return neg ? Z - q : q;
}
// Column of stored kz index j in the rows of the transformed fields (T, V,
// the adjoint rows): negative kz sit ZCOL_SHIFT columns further right (the
// row pitch ZT >= Z + ZCOL_SHIFT), so that with Z = 1 (mod 4) a block of QB
// negative kz starts on a 64-byte boundary, as a block of positive kz does,
// and the x passes write and read whole 64-byte segments.
#define ZCOL_SHIFT 2
__device__ __forceinline__ int zcol(int j, int KZ) { return j > KZ ? j + ZCOL_SHIFT : j; }
__device__ __forceinline__ int slot_partner(int c, int q0) {
  const int q = q0 + (c % QB);
  return q == 0 ? c : (c ^ QB);
}

// XCD-aware decode of a 1-D grid of 8*ceil(total/8) blocks: blocks are
// dispatched round-robin over the 8 XCDs, so logical index g =
// (bid % 8) * per + bid / 8 puts consecutive g on one XCD (one L2).
__device__ __forceinline__ int xcd_logical() {
  const int per = (int)(gridDim.x / 8);
  return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}

enum { OP_ID = 0, OP_DX = 1, OP_DZ = 2, OP_ZERO = 3 };

__device__ __forceinline__ double2 apply_op(int op, double2 v, double kxv, double kzv) {
  if (op == OP_DX) return cik(kxv, v);
  if (op == OP_DZ) return cik(kzv, v);
  if (op == OP_ZERO) return make_double2(0.0, 0.0);
  return v;
}

// Inverse x pass: T[x][l][j] = sum_kx (Herm(opA A) + i Herm(opB B))[kx, l, j] e^{+i kx x}
// scaled by s_inv.  One launch covers all packed transforms (blockIdx.z).
// grid (nqb, N, ntrans), block 256.
#define JSD_MAXTR 12
struct XInvSet {
  const double2* A[JSD_MAXTR];
  const double2* B[JSD_MAXTR];
  double2* T[JSD_MAXTR];
  int opA[JSD_MAXTR];
  int opB[JSD_MAXTR];
};

__global__ __launch_bounds__(256) void k_fft_x_inv(XInvSet set, int ntr, int nqb, FftPlan PX,
                                                   const double* __restrict__ kx, const double* __restrict__ kz, int M,
                                                   int N, int Z, int ZT, double s_inv) {
  extern __shared__ double2 lds[];
  double2* buf = lds;  // [p][slot]
  // logical order (l, qb, tr), tr fastest: the transforms reading the same
  // velocity component run together on one XCD
  const int g = xcd_logical();
  if (g >= ntr * nqb * N) return;
  const int tr = g % ntr;
  const int q0 = ((g / ntr) % nqb) * QB;
  const int l = g / (ntr * nqb);
  const double2* __restrict__ A = set.A[tr];
  const double2* __restrict__ B = set.B[tr];
  const int opA = set.opA[tr], opB = set.opB[tr];
  const int K = M / 2, KZ = Z / 2;
  const int P = PX.p;
  for (int w = threadIdx.x; w < P * NSLOT; w += blockDim.x) {
    const int c = w % NSLOT;
    const int pos = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    double2 val = make_double2(0.0, 0.0);
    const int n = PX.perm[pos];  // bin index
    int i = -1;
    if (n <= K) i = n;
    else if (n >= P - K) i = n - P + M;
    if (ok && i >= 0) {
      const int pi = (M - i) % M;
      const int pj = (Z - j) % Z;
      const size_t e = ((size_t)i * N + l) * Z + j;
      const size_t ep = ((size_t)pi * N + l) * Z + pj;
      const double2 a = apply_op(opA, A[e], kx[i], kz[j]);
      const double2 ap = apply_op(opA, A[ep], kx[pi], kz[pj]);
      const double2 b = apply_op(opB, B[e], kx[i], kz[j]);
      const double2 bp = apply_op(opB, B[ep], kx[pi], kz[pj]);
      const double2 ha = cherm(a, ap);
      const double2 hb = cherm(b, bp);
      val = make_double2((ha.x - hb.y) * s_inv, (ha.y + hb.x) * s_inv);
    }
    buf[pos * NSLOT + c] = val;
  }
  __syncthreads();
  fft_lds(buf, NSLOT, NSLOT, 1, PX, +1.0);
  double2* __restrict__ T = set.T[tr];
  for (int w = threadIdx.x; w < P * NSLOT; w += blockDim.x) {
    const int c = w % NSLOT;
    const int x = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    if (ok) T[((size_t)x * N + l) * ZT + zcol(j, KZ)] = buf[x * NSLOT + c];
  }
}

// Physical-space product variants
enum { PROD_NS = 0, PROD_PERT = 1, PROD_DUAL = 2 };

// Skew-symmetric products at one physical point.  row[2c] = (u_c, d_x u_c),
// row[2c+1] = (d_y u_c, d_z u_c); the dual adds the forward field in rows
// 6..11 (the base flow U(y) is added here).  H gets the three components.
// row[t * ST] holds packed transform t at one physical point.
// The fields at a physical point from the six inverse rows (row t at
// row[t * ST]): (u_c, d_x u_c), (d_y u_c, d_z u_c) for c = 0, 1, 2, or with
// T5 (u_0, u_1), (d_z u_0, d_z u_1), (u_2, d_z u_2), (d_x u_0, d_x u_1),
// (d_x u_2, d_y u_0), (d_y u_1, d_y u_2) (zrow_src, zrow_post).
template <int ST, bool T5>
__device__ __forceinline__ void unpack_fields(const double2* row, double* u, double (*g)[3]) {
  if constexpr (T5) {
    const double2 r0 = row[0], r1 = row[ST], r2 = row[2 * ST], r3 = row[3 * ST], r4 = row[4 * ST], r5 = row[5 * ST];
    u[0] = r0.x;
    u[1] = r0.y;
    g[0][2] = r1.x;
    g[1][2] = r1.y;
    u[2] = r2.x;
    g[2][2] = r2.y;
    g[0][0] = r3.x;
    g[1][0] = r3.y;
    g[2][0] = r4.x;
    g[0][1] = r4.y;
    g[1][1] = r5.x;
    g[2][1] = r5.y;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 r0 = row[(2 * c) * ST];
      const double2 r1 = row[(2 * c + 1) * ST];
      u[c] = r0.x;
      g[c][0] = r0.y;
      g[c][1] = r1.x;
      g[c][2] = r1.y;
    }
  }
}

// The five-row layout (T5) of the transformed fields of one velocity:
// T0 = u_0 + i u_1, T1 = u_2 (+ i 0), T2 = d_x u_0 + i d_x u_1,
// T3 = d_x u_2 + i d_y u_0, T4 = d_y u_1 + i d_y u_2.  The z derivatives come
// from the z spectra: inverse row 1 = i kz T0 (d_z u_0 + i d_z u_1), inverse
// row 2 = (1 - kz) T1 = T1 + i (i kz T1) (u_2 + i d_z u_2, as T1's imaginary
// part is zero), so the six inverse rows are T0, i kz T0, (1 - kz) T1, T2,
// T3, T4: every lane forms its row from one load, with no other lane's data.
// The wavenumber of stored index j is kz1 times its signed index.
// (zrow_src gives the row of T that inverse row t reads; zrow_post applies
// the kz factor once all loads of the thread are in flight, so that no load
// waits inside the load loop.)
template <bool T5>
__device__ __forceinline__ int zrow_src(int t) {
  return T5 ? (t <= 1 ? 0 : t - 1) : t;
}
template <bool T5>
__device__ __forceinline__ void zrow_post(int t, int j, int Z, int KZ, double kz1, double2& a) {
  if constexpr (T5) {
    if (j < 0 || t > 2 || t == 0) return;
    const double k = kz1 * (double)(j <= KZ ? j : j - Z);
    a = t == 1 ? make_double2(-k * a.y, k * a.x) : make_double2((1.0 - k) * a.x, (1.0 - k) * a.y);
  }
}

template <int PROD, int ST = 1, bool T5 = false>
__device__ __forceinline__ void zmid_products(const double2* row, double U, double dU, double clin, double* H) {
  double u[3], g[3][3];
  unpack_fields<ST, T5>(row, u, g);
  const double div = g[0][0] + g[1][1] + g[2][2];
  if constexpr (PROD == PROD_NS) {
    // H = -(u.grad) u + 0.5 (div u) u   (navier_stokes.py:236-261)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double conv = u[0] * g[c][0] + u[1] * g[c][1] + u[2] * g[c][2];
      H[c] = -conv + 0.5 * (div * u[c]);
    }
  } else if constexpr (PROD == PROD_PERT) {
    // H = -(c u.grad u + U d_x u + v U' e_x) + 0.5 (c div u u + div u U e_x)
    // (navier_stokes_perturbation.py:160-239)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double conv = u[0] * g[c][0] + u[1] * g[c][1] + u[2] * g[c][2];
      double hel = clin * conv + U * g[c][0];
      double dv = clin * (div * u[c]);
      if (c == 0) {
        hel += u[1] * dU;
        dv += div * U;
      }
      H[c] = -hel + 0.5 * dv;
    }
  } else {
    // dual: rows 0..5 adjoint v, rows 6..11 the forward field W (plus U)
    double wv[3], gw[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 r0 = row[(6 + 2 * c) * ST];
      const double2 r1 = row[(7 + 2 * c) * ST];
      wv[c] = clin * r0.x;
      gw[c][0] = clin * r0.y;
      gw[c][1] = clin * r1.x;
      gw[c][2] = clin * r1.y;
    }
    wv[0] += U;
    gw[0][1] += dU;
    const double divw = gw[0][0] + gw[1][1] + gw[2][2];
    // H = -(W.grad) v + [sum_j v_j d_i W_j] + 0.25 ((div W) v + (div v) W)
    // (navier_stokes_perturbation_dual.py:214-294)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double t1 = wv[0] * g[c][0] + wv[1] * g[c][1] + wv[2] * g[c][2];
      const double t2 = u[0] * gw[0][c] + u[1] * gw[1][c] + u[2] * gw[2][c];
      H[c] = -t1 + t2 + 0.5 * (0.5 * (divw * u[c] + div * wv[c]));
    }
  }
}

// Middle z pass: inverse z-FFT of nt packed inputs, products, forward z-FFT of
// two packed outputs, for YB consecutive y planes of one x.  grid
// (PXlen, ceil(N/YB)), block 256.  LDS rows r = yb*nt + t hold element
// (pos, r) at lds[pos*R + r]; after the products the same memory holds the
// 2*YB output rows, written directly in DIT input order (iperm).
// Input rows (NS / perturbation): row 2c = (u_c, d_x u_c), row 2c+1 = (d_y u_c, d_z u_c).
// Dual adds rows 6+2c, 7+2c for the forward field (the base flow is added here).
#define ZMID_ITEMS 8
__global__ __launch_bounds__(256) void k_fft_z_mid(const double2* __restrict__ T, long tstride, int nt,
                                                   double2* __restrict__ V, long vstride, FftPlan PZ, int N, int Z,
                                                   int ZT, int prod, double clin, const double* __restrict__ Ub,
                                                   const double* __restrict__ dUb, int YB) {
  extern __shared__ double2 lds[];
  const int P = PZ.p;
  const int x = blockIdx.x;
  const int l0 = blockIdx.y * YB;
  const int nyb = min(YB, N - l0);
  const int KZ = Z / 2;
  const int R = nt * YB;
  // zero the padded band, then scatter the kept kz bins (read with consecutive
  // lanes on consecutive kz) into their DIT positions
  for (int w = threadIdx.x; w < P * R; w += blockDim.x) lds[w] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int w = threadIdx.x; w < Z * R; w += blockDim.x) {
    const int j = w % Z;
    const int r = w / Z;
    const int yb = r / nt, t = r - yb * nt;
    if (yb >= nyb) continue;
    const int n = (j <= KZ) ? j : P - (Z - j);
    lds[PZ.iperm[n] * R + r] = T[t * tstride + ((size_t)x * N + l0 + yb) * ZT + zcol(j, KZ)];
  }
  __syncthreads();
  fft_lds(lds, R, R, 1, PZ, +1.0);
  double Hs[ZMID_ITEMS][3];
#pragma unroll
  for (int q = 0; q < ZMID_ITEMS; ++q) {
    const int item = threadIdx.x + q * blockDim.x;
    if (item >= P * YB) continue;
    const int yb = item / P, z = item - yb * P;
    const int l = min(l0 + yb, N - 1);
    const double2* row = lds + z * R + yb * nt;
    const double U = Ub ? Ub[l] : 0.0;
    const double dU = dUb ? dUb[l] : 0.0;
    if (prod == PROD_NS) zmid_products<PROD_NS>(row, U, dU, clin, Hs[q]);
    else if (prod == PROD_PERT) zmid_products<PROD_PERT>(row, U, dU, clin, Hs[q]);
    else zmid_products<PROD_DUAL>(row, U, dU, clin, Hs[q]);
  }
  __syncthreads();
  const int R2 = 2 * YB;
#pragma unroll
  for (int q = 0; q < ZMID_ITEMS; ++q) {
    const int item = threadIdx.x + q * blockDim.x;
    if (item >= P * YB) continue;
    const int yb = item / P, z = item - yb * P;
    const int pos = PZ.iperm[z];
    lds[pos * R2 + 2 * yb] = make_double2(Hs[q][0], Hs[q][1]);
    lds[pos * R2 + 2 * yb + 1] = make_double2(Hs[q][2], 0.0);
  }
  __syncthreads();
  fft_lds(lds, R2, R2, 1, PZ, -1.0);
  for (int w = threadIdx.x; w < Z * R2; w += blockDim.x) {
    const int j = w % Z;
    const int r = w / Z;
    const int yb = r / 2, o = r % 2;
    if (yb >= nyb) continue;
    const int n = (j <= KZ) ? j : P - (Z - j);
    V[o * vstride + ((size_t)x * N + l0 + yb) * ZT + zcol(j, KZ)] = lds[n * R2 + r];
  }
}

// Forward x pass with unpacking and the reference's +-K storage convention
// (domain.py:724-742).  grid (nqb, N, 2), block 256.  Output t=0 writes
// H0, H1; t=1 writes H2.
__global__ __launch_bounds__(256) void k_fft_x_fwd(const double2* __restrict__ V, long vstride,
                                                   double2* __restrict__ H, long hstride, FftPlan PX, int M, int N,
                                                   int Z, int ZT, int nqb, double s_fwd) {
  extern __shared__ double2 lds[];
  double2* buf = lds;
  const int g = xcd_logical();
  if (g >= 2 * nqb * N) return;
  const int q0 = (g % nqb) * QB;
  const int t = (g / nqb) % 2;
  const int l = g / (2 * nqb);
  const int K = M / 2, KZ = Z / 2;
  const int P = PX.p;
  const double2* Vt = V + t * vstride;
  for (int w = threadIdx.x; w < P * NSLOT; w += blockDim.x) {
    const int c = w % NSLOT;
    const int pos = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    const int xi = PX.perm[pos];
    buf[pos * NSLOT + c] = ok ? Vt[((size_t)xi * N + l) * ZT + zcol(j, KZ)] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_lds(buf, NSLOT, NSLOT, 1, PX, -1.0);
  auto bin = [&](int i) { return i <= K ? i : P - (M - i); };
  for (int w = threadIdx.x; w < M * NSLOT; w += blockDim.x) {
    const int c = w % NSLOT;
    const int i = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    if (!ok) continue;
    int si = i, sc = c;
    if (i == K + 1 && j == KZ + 1) {
      si = K;
      sc = slot_partner(c, q0);
    } else if (i == K + 1) {
      sc = slot_partner(c, q0);
    } else if (j == KZ + 1) {
      si = (M - i) % M;
    }
    const int spi = (M - si) % M;
    const double2 Wv = buf[bin(si) * NSLOT + sc];
    const double2 Wp = buf[bin(spi) * NSLOT + slot_partner(sc, q0)];
    const size_t e = ((size_t)i * N + l) * Z + j;
    if (t == 0) {
      const double2 g0 = make_double2(0.5 * (Wv.x + Wp.x), 0.5 * (Wv.y - Wp.y));
      const double dr = Wv.x - Wp.x, di = Wv.y + Wp.y;  // W - conj(Wp)
      const double2 g1 = make_double2(0.5 * di, -0.5 * dr);
      H[e] = cscale(g0, s_fwd);
      H[hstride + e] = cscale(g1, s_fwd);
    } else {
      const double2 g2 = make_double2(0.5 * (Wv.x + Wp.x), 0.5 * (Wv.y - Wp.y));
      H[2 * hstride + e] = cscale(g2, s_fwd);
    }
  }
}

// ===========================================================================
// FFT passes with the transform length fixed at compile time
// ===========================================================================
// Same data flow and arithmetic as the three kernels above; C = Rad<...> is
// the radix list of the plan.  The host launches these for the lengths of
// the supported configurations and the run-time kernels otherwise.

template <class C>
__global__ __launch_bounds__(256) void k_fft_x_inv_t(XInvSet set, int ntr, int nqb, const double2* __restrict__ tw,
                                                     const double* __restrict__ kx, const double* __restrict__ kz,
                                                     int M, int N, int Z, int ZT, double s_inv) {
  extern __shared__ double2 lds[];
  constexpr int P = C::P;
  const int g = xcd_logical();
  if (g >= ntr * nqb * N) return;
  const int tr = g % ntr;
  const int q0 = ((g / ntr) % nqb) * QB;
  const int l = g / (ntr * nqb);
  const double2* __restrict__ A = set.A[tr];
  const double2* __restrict__ B = set.B[tr];
  const int opA = set.opA[tr], opB = set.opB[tr];
  const bool sameAB = (A == B);  // (u_c, d_x u_c): one field read once
  const int K = M / 2, KZ = Z / 2;
  // zero padding bins K+1 .. P-K-1 (3/2-rule grid beyond the stored modes)
  for (int w = threadIdx.x; w < (P - M) * NSLOT; w += 256) {
    const int c = w % NSLOT, n = K + 1 + w / NSLOT;
    lds[dit_iperm<C, C::NS>(n) * NSLOT + c] = make_double2(0.0, 0.0);
  }
  // Hermitian pairs: item (i <= K, slot c) reads (i, kz_c) and its partner
  // (-kx, -kz) = (M-i, slot c^QB) once and writes both bins, whose packed
  // values are Herm(a) + i Herm(b) and its conjugate-pair counterpart.  All
  // loads of a thread are issued before any is used.
  constexpr int IT = ((P + 2) / 3 * NSLOT + 255) / 256;  // K + 1 <= (P + 2) / 3
  const int nitem = (K + 1) * NSLOT;
  double2 ra[IT], rap[IT], rb[IT], rbp[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int w = threadIdx.x + 256 * k;
    const int c = w % NSLOT, i = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    ra[k] = rap[k] = rb[k] = rbp[k] = make_double2(0.0, 0.0);
    if (w < nitem && ok && !(i == 0 && c >= QB)) {
      const int pi = (M - i) % M, pj = (Z - j) % Z;
      const size_t e = ((size_t)i * N + l) * Z + j;
      const size_t ep = ((size_t)pi * N + l) * Z + pj;
      ra[k] = A[e];
      rap[k] = A[ep];
      if (!sameAB) {
        rb[k] = B[e];
        rbp[k] = B[ep];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int w = threadIdx.x + 256 * k;
    const int c = w % NSLOT, i = w / NSLOT;
    if (w >= nitem) continue;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    if (i == 0 && c >= QB && ok) continue;  // written by its partner item (0, c - QB)
    const int pi = (M - i) % M, pj = (Z - j) % Z;
    const int cp = slot_partner(c, q0);
    double2 val = make_double2(0.0, 0.0), valp = val;
    if (ok) {
      const double2 b0 = sameAB ? ra[k] : rb[k], bp0 = sameAB ? rap[k] : rbp[k];
      const double2 a = apply_op(opA, ra[k], kx[i], kz[j]);
      const double2 ap = apply_op(opA, rap[k], kx[pi], kz[pj]);
      const double2 b = apply_op(opB, b0, kx[i], kz[j]);
      const double2 bp = apply_op(opB, bp0, kx[pi], kz[pj]);
      const double2 ha = cherm(a, ap), hap = cherm(ap, a);
      const double2 hb = cherm(b, bp), hbp = cherm(bp, b);
      val = make_double2((ha.x - hb.y) * s_inv, (ha.y + hb.x) * s_inv);
      valp = make_double2((hap.x - hbp.y) * s_inv, (hap.y + hbp.x) * s_inv);
    }
    const int n = i, np_ = pi <= K ? pi : P - (M - pi);  // bins of kx index i and M - i
    lds[dit_iperm<C, C::NS>(n) * NSLOT + c] = val;
    if (!(pi == i && cp == c)) lds[dit_iperm<C, C::NS>(np_) * NSLOT + cp] = valp;
  }
  __syncthreads();
  fft_lds_ct<C, NSLOT, NSLOT, 1, 256, +1>(lds, tw);
  double2* __restrict__ T = set.T[tr];
  for (int w = threadIdx.x; w < P * NSLOT; w += 256) {
    const int c = w % NSLOT;
    const int x = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    if (ok) T[((size_t)x * N + l) * ZT + zcol(j, KZ)] = lds[x * NSLOT + c];
  }
}

// z pass for one (x, y) line: NT packed input rows, products, two packed
// output rows.  grid (PXlen, N), block zmid_threads(P).
__host__ __device__ constexpr int zmid_threads(int p) { return p >= 128 ? 192 : 64; }

template <class C, int NT, int PROD>
__global__ __launch_bounds__(zmid_threads(C::P)) void k_fft_z_mid_t(const double2* __restrict__ T, long tstride,
                                                                    double2* __restrict__ V, long vstride,
                                                                    const double2* __restrict__ tw, int N, int Z,
                                                                    int ZT, double clin, const double* __restrict__ Ub,
                                                                    const double* __restrict__ dUb) {
  extern __shared__ double2 lds[];
  constexpr int P = C::P;
  constexpr int NTH = zmid_threads(P);
  constexpr int ITEMS = (P + NTH - 1) / NTH;
  const int x = blockIdx.x;
  const int l = blockIdx.y;
  const int KZ = Z / 2;
  const size_t line = ((size_t)x * N + l) * ZT;
  for (int w = threadIdx.x; w < P * NT; w += NTH) lds[w] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int j = threadIdx.x; j < Z; j += NTH) {
    const int n = (j <= KZ) ? j : P - (Z - j);
    const int pos = dit_iperm<C, C::NS>(n);
    double2 vals[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) vals[t] = T[t * tstride + line + zcol(j, KZ)];
#pragma unroll
    for (int t = 0; t < NT; ++t) lds[pos * NT + t] = vals[t];
  }
  __syncthreads();
  fft_lds_ct<C, NT, NT, 1, NTH, +1>(lds, tw);
  const double U = Ub ? Ub[l] : 0.0;
  const double dU = dUb ? dUb[l] : 0.0;
  double Hs[ITEMS][3];
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const int z = threadIdx.x + q * NTH;
    if (z < P) zmid_products<PROD>(lds + z * NT, U, dU, clin, Hs[q]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const int z = threadIdx.x + q * NTH;
    if (z < P) {
      const int pos = dit_iperm<C, C::NS>(z);
      lds[pos * 2] = make_double2(Hs[q][0], Hs[q][1]);
      lds[pos * 2 + 1] = make_double2(Hs[q][2], 0.0);
    }
  }
  __syncthreads();
  fft_lds_ct<C, 2, 2, 1, NTH, -1>(lds, tw);
  for (int j = threadIdx.x; j < Z; j += NTH) {
    const int n = (j <= KZ) ? j : P - (Z - j);
    V[line + zcol(j, KZ)] = lds[n * 2];
    V[vstride + line + zcol(j, KZ)] = lds[n * 2 + 1];
  }
}

// z pass for P = 16 N1 (N1 = 16: P = 256; N1 = 8: P = 128) with the
// butterflies in registers.  With n = n1 + N1 n2 and k = 16 k1 + k2: step 1
// (one thread per row t and n1 < N1) reads x[n1 + N1 n2] straight from HBM
// (zero padding by index), takes the 16-point DFT over n2 and multiplies by
// wP^(n1 k2); step 2 (one thread per t and k2 < 16) takes the N1-point DFT
// over n1, giving the physical point z = 16 k1 + k2.  Two LDS exchanges per
// transform instead of three or four stages, no LDS zero fill.  The forward
// transform of the two packed product rows takes the same split, each DFT
// spread over lanes (below), and stores only the kept bins.  A workgroup
// takes LN neighbouring y lines of one x: with LN = 2 and 6 rows the 16-lane
// steps fill 3 of 4 waves.  T5: the six inverse rows are formed from the
// five-row layout of the transformed fields (zrow_src, zrow_post); kz1 is the
// z wavenumber of stored index 1.
#define Z16_LD 17  // padded row (in double2) of the [t][n1][k2] exchange layout

template <int NT, int LN = 1>
__host__ __device__ constexpr int z256_threads() {
  return NT * 16 * LN > 128 ? (LN == 1 ? NT * 16 : 256) : 128;
}

// PZ (P = 256 only): at most 129 stored kz (|kz| <= 64), so step 1 of the
// inverse sees zeros at n2 = 5..11 (dft16_pz).
template <int P, int NT, int PROD, bool PZ, int LN = 1, bool T5 = false>
__global__ __launch_bounds__((z256_threads<NT, LN>()), (P == 128 ? 4 : 3)) void k_fft_z_mid_r(
    const double2* __restrict__ T, long tstride, double2* __restrict__ V, long vstride,
    const double2* __restrict__ tw, int N, int Z, int ZT, double clin, const double* __restrict__ Ub,
    const double* __restrict__ dUb, double kz1) {
  static_assert(!T5 || NT == 6, "the five-row layout holds one velocity");
  static_assert(P == 128 || P == 256, "register z pass: P = 128 or 256");
  static_assert(!PZ || P == 256, "the pruned first stage assumes P = 256");
  constexpr int N1 = P / 16;
  constexpr int NTH = z256_threads<NT, LN>();
  static_assert(NT * 16 * LN <= NTH, "rows of the workgroup exceed its threads");
  constexpr int ITEMS = (LN * P + NTH - 1) / NTH;
  // per line: [t][n1][k2] (padded) for the exchanges, [t][z] for the products
  constexpr int REG = NT * N1 * Z16_LD > P * NT ? NT * N1 * Z16_LD : P * NT;
  __shared__ double2 lds[LN * REG];
  // grid (P'x, line groups) taken in dispatch order with the line fastest:
  // consecutive workgroups read neighbouring lines of one x (2 KB apart)
  const int bid = blockIdx.x + gridDim.x * blockIdx.y;
  const int x = bid / gridDim.y;
  const int l0 = (bid % gridDim.y) * LN;  // lines l0 .. l0+LN-1; those >= N load zeros and store nothing
  const int KZ = Z / 2;
  const int nneg = Z - KZ - 1;  // stored negative bins
  // lanes of the forward transforms (below): FL per packed row, line fli,
  // row fo, position fi in the row's DFT, quad / pair position fj; their
  // twiddles are loaded here, ahead of the inverse, so that no table load
  // waits inside the forward steps
  constexpr int FL = P / 4;
  static_assert(2 * FL * LN <= NTH, "forward lanes exceed the workgroup");
  const bool fwd = (int)threadIdx.x < 2 * FL * LN;
  const int fli = threadIdx.x / (2 * FL), fo = (threadIdx.x / FL) & 1;
  const int fr = threadIdx.x % FL;
  const int fj = fr / N1, fi = fr % N1;  // step 1: a = fi < N1, n0 = fj < 4
  const int f2d = fr & 15, f2j = fr >> 4;  // step 2: d = f2d < 16, f2j < FL / 16
  double2 fw16[4], fw1[4], fw2[4];  // w16^(fj k0); wP^(fi (fj + 4 k1)); step 2 twiddles
  if (fwd) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      fw16[k] = tw[N1 * fj * k];
      fw1[k] = tw[fi * (fj + 4 * k)];
      fw2[k] = tw[16 * f2j * k];  // w16^(a0 c0) (P = 256) or w8^(a0 c0) (P = 128)
    }
  }
  double2 v[16];
  // ---- inverse, step 1
  if (threadIdx.x < LN * NT * N1) {
    const int li = threadIdx.x / (NT * N1), r = threadIdx.x % (NT * N1);
    const int t = r / N1, n1 = r % N1;
    const int l = l0 + li;
    const size_t line = ((size_t)x * N + (l < N ? l : 0)) * ZT;
    const int ts = zrow_src<T5>(t);
    // the step's twiddle seeds: with P = 256 (nine row loads) they load
    // first and arrive with the rows; with P = 128 (sixteen row loads) the
    // registers are not there, and they load after the DFT
    double2 w1, w8;
    if constexpr (P == 256) {
      w1 = cconj(tw[n1]);
      w8 = cconj(tw[8 * n1]);
    }
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const int n = n1 + N1 * n2;
      int j = -1;
      if (n <= KZ) j = n;
      else if (n >= P - nneg) j = n - P + Z;
      v[n2] = (j >= 0 && l < N) ? ld_nt(&T[ts * tstride + line + zcol(j, KZ)]) : make_double2(0.0, 0.0);
    }
    if constexpr (T5) {
#pragma unroll
      for (int n2 = 0; n2 < 16; ++n2) {
        const int n = n1 + N1 * n2;
        const int j = n <= KZ ? n : (n >= P - nneg ? n - P + Z : -1);
        zrow_post<T5>(t, j, Z, KZ, kz1, v[n2]);
      }
    }
    if constexpr (PZ) dft16_pz<+1>(v);
    else dft16<+1>(v);
    // twiddles wP^(n1 k2) = w^k2 (w = wP^n1 from the table) as running
    // powers, restarted from the table at k2 = 8: two loads and two live
    // twiddles instead of fifteen (the powers agree with the table to ~8 ulp)
    {
      if constexpr (P != 256) {
        w1 = cconj(tw[n1]);
        w8 = cconj(tw[8 * n1]);
      }
      lds[li * REG + (t * N1 + n1) * Z16_LD] = v[0];
      double2 wk = w1;
#pragma unroll
      for (int k2 = 1; k2 < 16; ++k2) {
        if (k2 == 8) wk = w8;
        lds[li * REG + (t * N1 + n1) * Z16_LD + k2] = cmul(v[k2], wk);
        if (k2 != 7) wk = cmul(wk, w1);
      }
    }
  }
  __syncthreads();
  // ---- inverse, step 2
  if (threadIdx.x < LN * NT * 16) {
    const int li = threadIdx.x / (NT * 16), r = threadIdx.x % (NT * 16);
    const int t = r >> 4, k2 = r & 15;
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = lds[li * REG + (t * N1 + n1) * Z16_LD + k2];
    if constexpr (N1 == 16) dft16<+1>(v);
    else dft8(v, 1.0);
  }
  __syncthreads();
  if (threadIdx.x < LN * NT * 16) {
    const int li = threadIdx.x / (NT * 16), r = threadIdx.x % (NT * 16);
    const int t = r >> 4, k2 = r & 15;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[li * REG + t * P + 16 * k1 + k2] = v[k1];
  }
  __syncthreads();
  // ---- products at each physical point
  double Hs[ITEMS][3];
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const int zz = threadIdx.x + q * NTH;
    const int li = zz / P, z = zz % P;
    const int l = l0 + li < N ? l0 + li : N - 1;
    const double U = Ub ? Ub[l] : 0.0;
    const double dU = dUb ? dUb[l] : 0.0;
    if (zz < LN * P) zmid_products<PROD, P, T5>(lds + li * REG + z, U, dU, clin, Hs[q]);
  }
  __syncthreads();
  // packed rows (H0 + i H1) and (H2 + 0 i) at [o][z]
#pragma unroll
  for (int q = 0; q < ITEMS; ++q) {
    const int zz = threadIdx.x + q * NTH;
    const int li = zz / P, z = zz % P;
    if (zz < LN * P) {
      lds[li * REG + z] = make_double2(Hs[q][0], Hs[q][1]);
      lds[li * REG + P + z] = make_double2(Hs[q][2], 0.0);
    }
  }
  __syncthreads();
  // ---- forward, n = a + N1 b, k = 16 c + d:
  //   X[16 c + d] = sum_a wP^(a d) wN1^(a c) sum_b x[a + N1 b] w16^(b d).
  // Step 1, per a: the 16-point DFT over b on four lanes as 4 x 4
  // (b = n0 + 4 n1, d = k0 + 4 k1: a 4-point DFT over n1, the twiddle
  // w16^(n0 k0), an exchange inside the wave, a 4-point DFT over n0), then
  // wP^(a d).  Step 2, per d: the N1-point DFT over a, on four lanes as
  // 4 x 4 (N1 = 16) or on two lanes as 4 x 2 (N1 = 8).  One wave (P = 256)
  // or half a wave (P = 128) per transform (line fli, packed row fo).  Per
  // line: Y' [o][a][d] (N1 x 17 padded) from 0, the lane exchange [o][P]
  // from 2 N1 17.
  constexpr int FY = N1 * Z16_LD, FS = 2 * FY;
  static_assert(FS + 2 * P <= REG, "forward transform layout exceeds the workgroup");
  double2* const frow = lds + fli * REG + fo * P;
  double2* const fY = lds + fli * REG + fo * FY;
  double2* const fS = lds + fli * REG + FS + fo * P;
  double2 u4[4];
  // step 1, first half: lane (a = fi, n0 = fj) takes x[a + N1 (n0 + 4 n1)]
  if (fwd) {
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) u4[n1] = frow[fi + N1 * fj + 4 * N1 * n1];
    dft4(u4, -1.0);
#pragma unroll
    for (int k0 = 1; k0 < 4; ++k0) u4[k0] = cmul(u4[k0], fw16[k0]);
#pragma unroll
    for (int k0 = 0; k0 < 4; ++k0) fS[(fj * 4 + k0) * N1 + fi] = u4[k0];
  }
  __syncthreads();
  // step 1, second half: lane (a, k0 = fj) gives Y[a][d], d = k0 + 4 k1,
  // times wP^(a d)
  if (fwd) {
#pragma unroll
    for (int n0 = 0; n0 < 4; ++n0) u4[n0] = fS[(n0 * 4 + fj) * N1 + fi];
    dft4(u4, -1.0);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      const int d = fj + 4 * k1;
      fY[fi * Z16_LD + d] = fi * d == 0 ? u4[k1] : cmul(u4[k1], fw1[k1]);
    }
  }
  __syncthreads();
  if constexpr (N1 == 16) {
    // step 2, first half: lane (d, a0 = f2j) takes Y'[a0 + 4 a1][d]
    if (fwd) {
#pragma unroll
      for (int a1 = 0; a1 < 4; ++a1) u4[a1] = fY[(f2j + 4 * a1) * Z16_LD + f2d];
      dft4(u4, -1.0);
#pragma unroll
      for (int c0 = 1; c0 < 4; ++c0) u4[c0] = cmul(u4[c0], fw2[c0]);
#pragma unroll
      for (int c0 = 0; c0 < 4; ++c0) fS[(f2j * 4 + c0) * 16 + f2d] = u4[c0];
    }
    __syncthreads();
    // step 2, second half: lane (d, c0 = f2j) gives bins 16 (c0 + 4 c1) + d
    if (fwd) {
#pragma unroll
      for (int a0 = 0; a0 < 4; ++a0) u4[a0] = fS[(a0 * 4 + f2j) * 16 + f2d];
      dft4(u4, -1.0);
    }
  } else {
    // step 2, first half: lane (d, a0 = f2j < 2) takes Y'[a0 + 2 a1][d],
    // a 4-point DFT over a1 gives c0, times w8^(a0 c0)
    if (fwd) {
#pragma unroll
      for (int a1 = 0; a1 < 4; ++a1) u4[a1] = fY[(f2j + 2 * a1) * Z16_LD + f2d];
      dft4(u4, -1.0);
#pragma unroll
      for (int c0 = 1; c0 < 4; ++c0) u4[c0] = cmul(u4[c0], fw2[c0]);
#pragma unroll
      for (int c0 = 0; c0 < 4; ++c0) fS[(f2j * 4 + c0) * 16 + f2d] = u4[c0];
    }
    __syncthreads();
    // step 2, second half: lane (d, e = f2j) takes c0 = 2 e, 2 e + 1 of both
    // a0 and a 2-point DFT over a0 gives c1: u4[2 m + c1] holds bin
    // 16 (2 e + m + 4 c1) + d
    if (fwd) {
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        double2 t2[2] = {fS[(0 * 4 + 2 * f2j + m) * 16 + f2d], fS[(1 * 4 + 2 * f2j + m) * 16 + f2d]};
        dft2(t2);
        u4[2 * m] = t2[0];
        u4[2 * m + 1] = t2[1];
      }
    }
  }
  if (fwd) {
    const int l = l0 + fli;
    if (l < N) {
      const size_t line = ((size_t)x * N + l) * ZT;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // P = 256: c = f2j + 4 q; P = 128: c = 2 f2j + (q >> 1) + 4 (q & 1)
        const int c = N1 == 16 ? f2j + 4 * q : 2 * f2j + (q >> 1) + 4 * (q & 1);
        const int b = 16 * c + f2d;
        int j = -1;
        if (b <= KZ) j = b;
        else if (b >= P - nneg) j = b - P + Z;
        if (j >= 0) st_nt(&V[fo * vstride + line + zcol(j, KZ)], u4[q]);
      }
    }
  }
}

template <class C>
__global__ __launch_bounds__(256) void k_fft_x_fwd_t(const double2* __restrict__ V, long vstride,
                                                     double2* __restrict__ H, long hstride,
                                                     const double2* __restrict__ tw, int M, int N, int Z, int ZT,
                                                     int nqb, double s_fwd) {
  extern __shared__ double2 lds[];
  constexpr int P = C::P;
  // logical order (l, t, qb), qb fastest: neighbouring kz blocks share
  // cache lines and run on one XCD
  const int g = xcd_logical();
  if (g >= 2 * nqb * N) return;
  const int q0 = (g % nqb) * QB;
  const int t = (g / nqb) % 2;
  const int l = g / (2 * nqb);
  const int K = M / 2, KZ = Z / 2;
  const double2* Vt = V + t * vstride;
  // all loads of a thread are issued before the LDS stores
  constexpr int IT = (P * NSLOT + 255) / 256;
  double2 vv[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int w = threadIdx.x + 256 * k;
    const int c = w % NSLOT, pos = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    vv[k] = make_double2(0.0, 0.0);
    if (w < P * NSLOT && ok) vv[k] = Vt[((size_t)dit_perm<C, C::NS>(pos) * N + l) * ZT + zcol(j, KZ)];
  }
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int w = threadIdx.x + 256 * k;
    if (w < P * NSLOT) lds[w] = vv[k];  // w = pos * NSLOT + c
  }
  __syncthreads();
  fft_lds_ct<C, NSLOT, NSLOT, 1, 256, -1>(lds, tw);
  auto bin = [&](int i) { return i <= K ? i : P - (M - i); };
  for (int w = threadIdx.x; w < M * NSLOT; w += 256) {
    const int c = w % NSLOT;
    const int i = w / NSLOT;
    bool ok;
    const int j = slot_col(c, q0, Z, KZ, &ok);
    if (!ok) continue;
    int si = i, sc = c;
    if (i == K + 1 && j == KZ + 1) {
      si = K;
      sc = slot_partner(c, q0);
    } else if (i == K + 1) {
      sc = slot_partner(c, q0);
    } else if (j == KZ + 1) {
      si = (M - i) % M;
    }
    const int spi = (M - si) % M;
    const double2 Wv = lds[bin(si) * NSLOT + sc];
    const double2 Wp = lds[bin(spi) * NSLOT + slot_partner(sc, q0)];
    const size_t e = ((size_t)i * N + l) * Z + j;
    if (t == 0) {
      const double2 g0 = make_double2(0.5 * (Wv.x + Wp.x), 0.5 * (Wv.y - Wp.y));
      const double dr = Wv.x - Wp.x, di = Wv.y + Wp.y;  // W - conj(Wp)
      const double2 g1 = make_double2(0.5 * di, -0.5 * dr);
      H[e] = cscale(g0, s_fwd);
      H[hstride + e] = cscale(g1, s_fwd);
    } else {
      const double2 g2 = make_double2(0.5 * (Wv.x + Wp.x), 0.5 * (Wv.y - Wp.y));
      H[2 * hstride + e] = cscale(g2, s_fwd);
    }
  }
}

// lengths with compile-time kernels: x (P' >= 3K+1 from radices 2..8 without 7)
// and z (powers of two) of the usual mode counts
typedef Rad<8, 5> RX40;
typedef Rad<5, 5, 3> RX75;
typedef Rad<5, 5, 4> RX100;
typedef Rad<5, 5, 3, 2> RX150;
typedef Rad<8, 5, 4> RX160;
typedef Rad<8, 5, 5> RX200;
typedef Rad<8> RZ8;
typedef Rad<8, 2> RZ16;
typedef Rad<8, 4> RZ32;
typedef Rad<8, 8> RZ64;
typedef Rad<8, 8, 2> RZ128;
typedef Rad<8, 8, 4> RZ256;

#include "jsd_fft_x.inc"

template <class C>
static bool plan_is(const FftPlan& f) {
  if (f.p != C::P || f.nrad != C::NS) return false;
  for (int s = 0; s < C::NS; ++s)
    if (f.rad[s] != C::r(s)) return false;
  return true;
}
template <class F>
static bool visit_x(const FftPlan& f, F&& fn) {
  if (plan_is<RX200>(f)) return fn(RX200()), true;
  if (plan_is<RX40>(f)) return fn(RX40()), true;
  if (plan_is<RX75>(f)) return fn(RX75()), true;
  if (plan_is<RX100>(f)) return fn(RX100()), true;
  if (plan_is<RX150>(f)) return fn(RX150()), true;
  if (plan_is<RX160>(f)) return fn(RX160()), true;
  return false;
}
template <class F>
static bool visit_z(const FftPlan& f, F&& fn) {
  if (plan_is<RZ256>(f)) return fn(RZ256()), true;
  if (plan_is<RZ8>(f)) return fn(RZ8()), true;
  if (plan_is<RZ16>(f)) return fn(RZ16()), true;
  if (plan_is<RZ32>(f)) return fn(RZ32()), true;
  if (plan_is<RZ64>(f)) return fn(RZ64()), true;
  if (plan_is<RZ128>(f)) return fn(RZ128()), true;
  return false;
}

// ===========================================================================
// linearised nonlinear term without transforms
// ===========================================================================
// The physical field of a stored spectrum X is IFFT(Herm(X)); multiplying by a
// y-only profile commutes with the x-z transforms, and the forward transform
// of the product stores G[kx*sx, kz*sz] (sz = -1 on the kx = -K row, sx = -1
// on the kz = -K column).  Perturbation (navier_stokes_perturbation.py:160-239
// with linearise): G0 = -(U ikx u + U' v) + 0.5 U div, G1 = -U ikx v,
// G2 = -U ikx w.  Dual (navier_stokes_perturbation_dual.py:214-294 with
// linearise): G0 = -U ikx v0 + 0.25 U div v, G1 = -U ikx v1 + U' v0,
// G2 = -U ikx v2.
struct NlLinArgs {
  const double2* U0;  // state (3 components)
  long cs;
  const double2* DV;  // D v
  double2* H;         // out: H (3 components, stored convention)
  const double* kx;
  const double* kz;
  const double* Ub;
  const double* dUb;
  int M, N, Z, dual;
};

__device__ __forceinline__ void nl_lin_body(const NlLinArgs& A, Blk blk) {
  const double2* __restrict__ U0 = A.U0;
  const long cs = A.cs;
  const double2* __restrict__ DV = A.DV;
  double2* __restrict__ H = A.H;
  const double* __restrict__ kx = A.kx;
  const double* __restrict__ kz = A.kz;
  const double* __restrict__ Ub = A.Ub;
  const double* __restrict__ dUb = A.dUb;
  const int M = A.M, N = A.N, Z = A.Z, dual = A.dual;
  const size_t tot = (size_t)M * N * Z;
  const size_t e = (size_t)blk.x * 256 + threadIdx.x;
  if (e >= tot) return;
  const int j = e % Z;
  const int l = (e / Z) % N;
  const int i = e / ((size_t)Z * N);
  const int K = M / 2, KZ = Z / 2;
  int si = i, sj = j;
  if (i == K + 1 && j == KZ + 1) {
    si = K;
    sj = KZ;
  } else if (i == K + 1) {
    sj = (Z - j) % Z;
  } else if (j == KZ + 1) {
    si = (M - i) % M;
  }
  const int pi = (M - si) % M, pj = (Z - sj) % Z;
  const size_t s0 = ((size_t)si * N + l) * Z + sj;
  const size_t p0 = ((size_t)pi * N + l) * Z + pj;
  const double2 hu = cherm(U0[s0], U0[p0]);
  const double2 hv = cherm(U0[cs + s0], U0[cs + p0]);
  const double2 hw = cherm(U0[2 * cs + s0], U0[2 * cs + p0]);
  const double2 hdv = cherm(DV[s0], DV[p0]);
  const double kxv = kx[si], kzv = kz[sj];
  const double U = Ub[l], dU = dUb[l];
  const double2 dxu = cik(kxv, hu), dxv = cik(kxv, hv), dxw = cik(kxv, hw);
  const double2 div = cadd(cadd(dxu, hdv), cik(kzv, hw));
  double2 g0, g1, g2;
  if (!dual) {
    g0 = cadd(cscale(cadd(cscale(dxu, U), cscale(hv, dU)), -1.0), cscale(div, 0.5 * U));
    g1 = cscale(dxv, -U);
    g2 = cscale(dxw, -U);
  } else {
    g0 = cadd(cscale(dxu, -U), cscale(div, 0.25 * U));
    g1 = cadd(cscale(dxv, -U), cscale(hu, dU));
    g2 = cscale(dxw, -U);
  }
  H[e] = g0;
  H[cs + e] = g1;
  H[2 * cs + e] = g2;
}

__global__ __launch_bounds__(256) void k_nl_lin(NlLinArgs A) {
  nl_lin_body(A, Blk{(int)blockIdx.x, 0, 0, (int)gridDim.x});
}

// The h_g addend of coupling_term = False (navier_stokes_perturbation.py:
// 214-224): G = (i kz) hat(U'(y) no_hat(v)).  The transform pair about a
// y-only factor is U'(y) times the Hermitian part of v with the reference's
// +-K storage map (as in k_nl_lin); the z-derivative uses the stored kz.
// DUAL = false: G = d_z hat(U' v) with V = v (the perturbation step's h_g
// addend).  DUAL = true: G += hat(U' v_x) with V = v_x, G = H_1 (the dual
// step's h_v addend (d_x^2 + d_z^2) hat(U' v_x), which the solve forms from
// H_1).  hat(no_hat(.)) of the stored modes is the Hermitian map of cherm.
template <bool DUAL>
__global__ __launch_bounds__(256) void k_coupling_corr(const double2* __restrict__ V, const double* __restrict__ dUb,
                                                      const double* __restrict__ kz, double2* __restrict__ G, int M,
                                                      int N, int Z) {
  const size_t tot = (size_t)M * N * Z;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= tot) return;
  const int j = e % Z;
  const int l = (e / Z) % N;
  const int i = e / ((size_t)Z * N);
  const int K = M / 2, KZ = Z / 2;
  int si = i, sj = j;
  if (i == K + 1 && j == KZ + 1) {
    si = K;
    sj = KZ;
  } else if (i == K + 1) {
    sj = (Z - j) % Z;
  } else if (j == KZ + 1) {
    si = (M - i) % M;
  }
  const int pi = (M - si) % M, pj = (Z - sj) % Z;
  const double2 hv = cherm(V[((size_t)si * N + l) * Z + sj], V[((size_t)pi * N + l) * Z + pj]);
  if (DUAL) {
    const double2 c = cscale(hv, dUb[l]), g = G[e];
    G[e] = make_double2(g.x + c.x, g.y + c.y);
  } else {
    G[e] = cik(kz[j], cscale(hv, dUb[l]));
  }
}

__global__ __launch_bounds__(256) void k_add(const double2* __restrict__ A, double2* __restrict__ B, size_t n) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) B[e] = cadd(B[e], A[e]);
}

// ---------------------------------------------------------------------------
// source term, enforce_base and externally supplied nonlinear terms
// ---------------------------------------------------------------------------
// The source term of one step (get_source_term, navier_stokes.py:1697-1704):
// Q = curl(source)[1] = (i kz) s_x - (i kx) s_z; the reference then takes the
// kx rows 0 and 2 of Q (curl_source[0], curl_source[2]), differentiates them
// as (Ny, Nz) arrays that broadcast over kx, and keeps kx row 1 of the
// difference, so every kx row of h_v gains
//   hvs[l][j] = (i kz_j) Q[0][l][j] - (i kx_1) Q[2][l][j]
// and every kx row of h_g gains hgs[l][j] = Q[1][l][j].  The (0,0) mean-flow
// block gains source_x_00 = s_x[0][l][0] and source_z_00 = s_z[0][l][0]
// (navier_stokes.py:1461-1490).  One thread per (l, j).
__global__ __launch_bounds__(256) void k_source_prep(const double2* __restrict__ S, long cs,
                                                     const double* __restrict__ kx, const double* __restrict__ kz,
                                                     int N, int Z, double2* __restrict__ SRC,
                                                     double2* __restrict__ S00) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= N * Z) return;
  const int l = t / Z, j = t - (t / Z) * Z;
  double2 q[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const size_t e = ((size_t)m * N + l) * Z + j;
    q[m] = csub(cik(kz[j], S[e]), cik(kx[m], S[2 * cs + e]));
  }
  SRC[t] = csub(cik(kz[j], q[0]), cik(kx[1], q[2]));
  SRC[(size_t)N * Z + t] = q[1];
  if (j == 0) {
    S00[l] = S[(size_t)l * Z];
    S00[N + l] = S[2 * cs + (size_t)l * Z];
  }
}

struct TermArgs {
  double2* H;            // nonlinear term H (3 components), rewritten in place
  long cs;
  double2* HGS;          // out: the h_g addend the solve reads
  const double2* HGC;    // coupling_term = False addend of this substep (or NULL)
  const double2* EHV;    // external h_v, h_g, conv_ns (nonlinear_update_fn), or NULL
  const double2* EHG;
  const double2* ECV;
  const double2* SRC;    // source addends hvs, hgs (N x Z each), or NULL
  const double2* S00;    // source_x_00, source_z_00, or NULL
  const double* kx;
  const double* kz;
  int M, N, Z, enforce_base;
};

// Rewrites H so that the solve, which forms h_v = -d_y(i kx H0 + i kz H2) -
// k^2 H1 (for the (0,0) pencil: h_v = H1, the h1_00 mode), h_g = (i kz) H0 -
// (i kx) H2 + HGS and conv_ns(0,0) = -H(0,0), sees the requested terms:
//  * external terms: H0 = H2 = 0 off (0,0), H1 = -h_v / k^2, HGS = h_g,
//    H(0,0) = -conv_ns(0,0), H1(0,0) = h_v(0,0);
//  * a source: H1 -= hvs / k^2 (H1 += hvs at (0,0), whose v the reference
//    solves with the source's h_v too), HGS += hgs,
//    H(0,0) += (source_x_00, source_z_00);
//  * enforce_base: the (0,0) convective terms are dropped (navier_stokes.py:
//    1581-1590) before the source is added.
// The (0,0) pencil's h_g is unused (the mean-flow block replaces u, w there).
__global__ __launch_bounds__(256) void k_step_terms(TermArgs A) {
  const size_t tot = (size_t)A.M * A.N * A.Z;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= tot) return;
  const int Z = A.Z, N = A.N;
  const int j = e % Z;
  const int l = (e / Z) % N;
  const int i = e / ((size_t)Z * N);
  const double k2 = A.kx[i] * A.kx[i] + A.kz[j] * A.kz[j];
  const bool z00 = (i == 0 && j == 0);
  double2 h0, h1, h2, g;
  const double2 zero = make_double2(0.0, 0.0);
  if (A.EHV) {
    const double2 hv = A.EHV[e];
    // the (0,0) pencil's H1 carries its h_v itself (the solve's h1_00 mode)
    h1 = z00 ? hv : (k2 != 0.0 ? make_double2(-hv.x / k2, -hv.y / k2) : zero);
    h0 = h2 = zero;
    if (z00) {
      const double2 c0 = A.ECV[e], c2 = A.ECV[2 * A.cs + e];
      h0 = make_double2(-c0.x, -c0.y);
      h2 = make_double2(-c2.x, -c2.y);
    }
    g = A.EHG[e];
  } else {
    h0 = A.H[e];
    h1 = z00 ? zero : A.H[A.cs + e];  // h_v of the (0,0) pencil is zero without a source
    h2 = A.H[2 * A.cs + e];
    g = A.HGC ? A.HGC[e] : zero;
  }
  if (A.SRC) {
    const size_t t = (size_t)l * Z + j;
    const double2 hvs = A.SRC[t];
    if (z00)
      h1 = cadd(h1, hvs);
    else if (k2 != 0.0)
      h1 = csub(h1, make_double2(hvs.x / k2, hvs.y / k2));
    g = cadd(g, A.SRC[(size_t)N * Z + t]);
  }
  if (z00) {
    if (A.enforce_base) h0 = h2 = zero;
    if (A.S00) {
      h0 = cadd(h0, A.S00[l]);
      h2 = cadd(h2, A.S00[N + l]);
    }
  }
  A.H[e] = h0;
  A.H[A.cs + e] = h1;
  A.H[2 * A.cs + e] = h2;
  A.HGS[e] = g;
}

// out = scale ((i kx)^2 f + D^2 f + (i kz)^2 f) for ncomp components, with
// out holding D^2 f on entry (Domain.laplacian, domain.py:671-673).
__global__ __launch_bounds__(256) void k_laplacian_finish(const double2* __restrict__ F, double2* __restrict__ O,
                                                          const double* __restrict__ kx,
                                                          const double* __restrict__ kz, int M, int N, int Z,
                                                          size_t tot, double scale) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= tot) return;
  const size_t r = e % ((size_t)M * N * Z);
  const int j = r % Z;
  const int i = r / ((size_t)Z * N);
  const double2 f = F[e], d2 = O[e];
  const double ax = -(kx[i] * kx[i]), az = -(kz[j] * kz[j]);
  const double2 s = cadd(cadd(cscale(f, ax), d2), cscale(f, az));
  O[e] = cscale(s, scale);
}

// acc += scale * sum_i sum_j (d_j u_i)^2 with the complex squares of the
// spectral derivatives, as the reference's dissipation sum does
// (navier_stokes_perturbation_dual.py:1087-1104); DU holds d_y u_i.
__global__ __launch_bounds__(256) void k_dissipation_acc(const double2* __restrict__ U, const double2* __restrict__ DU,
                                                         long cs, const double* __restrict__ kx,
                                                         const double* __restrict__ kz, int M, int N, int Z,
                                                         double2* __restrict__ acc, double scale) {
  const size_t tot = (size_t)M * N * Z;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= tot) return;
  const int j = e % Z;
  const int i = e / ((size_t)Z * N);
  double2 s = acc[e];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 u = U[c * cs + e];
    const double2 d[3] = {cik(kx[i], u), DU[c * cs + e], cik(kz[j], u)};
#pragma unroll
    for (int q = 0; q < 3; ++q) s = cadd(s, cscale(cmul(d[q], d[q]), scale));
  }
  acc[e] = s;
}

// Volume integral of no_hat(f) (field.py:1456-1466): the x-z sums of the
// physical field are sqrt(Px Pz) s Re f(0, l, 0), then the y weights.
__global__ void k_mean_integral(const double2* __restrict__ F, const double* __restrict__ wy, int N, int Z,
                                double scale, double* __restrict__ out) {
  __shared__ double red[256];
  double a = 0.0;
  for (int l = threadIdx.x; l < N; l += blockDim.x) a += wy[l] * F[(size_t)l * Z].x;
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// ===========================================================================
// per-wavenumber implicit solve (navier_stokes.py:1248-1546)
// ===========================================================================
// One WG per distinct (|kx|, |kz|) = (a, b); it handles the up-to-four signed
// pencils (+-a, +-b).  Vectors live in LDS as [vec][y], vec = 2*pencil + re/im.
// Matrices are stored transposed (AT[j*N + i] = A[i][j]) so consecutive
// threads (rows i) read consecutive addresses.
// Per-group matvecs on FP64 MFMA.  The group's vectors live in LDS as
// [y][SLD] (16 columns: column 2p+r is pencil p, re/im r; a second set of 8
// for omega_y where needed).  Rows N..JSD_NYMAX-1 stay zero so the K padding
// of the 16x16x4 tiles reads zeros.
#define SLD 16

// Y[i][c] = sum_j A[i][j] X[j][c] for rows r0 <= i < r1 and columns c < ncol.
// AT is A transposed (AT[j*N + i]); a wave owns row tiles w, w+4, w+8.  The
// A fragments of the next MV_KC k-steps are loaded into registers while the
// current ones feed the MFMAs (the per-|k| tables stream from HBM).
#define MV_KC 4
__device__ void mv_mfma(const double* __restrict__ AT, int N, const double* X, double* Y, int ncol, int r0, int r1) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int NT = (N + 15) / 16, KS = (N + 3) / 4;
  const int kr = lane >> 4, cc = lane & 15;
  dbl4 acc[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) acc[q] = (dbl4){0.0, 0.0, 0.0, 0.0};
  const int nt = (NT - wave + 3) / 4;  // tiles wave, wave+4, wave+8 below NT
  int row[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) row[q] = (wave + 4 * q) * 16 + cc;
  double acur[MV_KC][3], anext[MV_KC][3];
  auto loadA = [&](int ks0, double (&dst)[MV_KC][3]) {
#pragma unroll
    for (int kk = 0; kk < MV_KC; ++kk) {
      const int k = (ks0 + kk) * 4 + kr;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        dst[kk][q] = (q < nt && k < N && row[q] < N) ? AT[(size_t)k * N + row[q]] : 0.0;
    }
  };
  loadA(0, acur);
  for (int ks0 = 0; ks0 < KS; ks0 += MV_KC) {
    if (ks0 + MV_KC < KS) loadA(ks0 + MV_KC, anext);
#pragma unroll
    for (int kk = 0; kk < MV_KC; ++kk) {
      if (ks0 + kk < KS) {
        const double b = X[((ks0 + kk) * 4 + kr) * SLD + cc];
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (q < nt) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(acur[kk][q], b, acc[q], 0, 0, 0);
      }
    }
#pragma unroll
    for (int kk = 0; kk < MV_KC; ++kk)
#pragma unroll
      for (int q = 0; q < 3; ++q) acur[kk][q] = anext[kk][q];
  }
  if (cc >= ncol) return;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (q < nt) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = (wave + 4 * q) * 16 + 4 * r + kr;
        if (i >= r0 && i < r1) Y[i * SLD + cc] = acc[q][r];
      }
    }
  }
}

#include "jsd_solve.inc"
#include "jsd_solve8.inc"
#include "jsd_solve8p.inc"

#include "jsd_adjoint.inc"
#include "jsd_adjoint_nl.inc"
#include "jsd_adjoint8p.inc"

// ===========================================================================
// energy: E = sum_c sum_l (w_l / 4) s^2/(Px Pz) sum_k |Herm(u_c)|^2
// ===========================================================================
__global__ __launch_bounds__(256) void k_energy_partial(const double2* __restrict__ U, long cs,
                                                        const double* __restrict__ wy, int M, int N, int Z,
                                                        double* __restrict__ partial) {
  __shared__ double red[256];
  const size_t tot = (size_t)3 * M * N * Z;
  double acc = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
    const int c = e / cs;
    const size_t r = e - (size_t)c * cs;
    const int j = r % Z;
    const int l = (r / Z) % N;
    const int i = r / ((size_t)Z * N);
    const size_t p = ((size_t)((M - i) % M) * N + l) * Z + (Z - j) % Z;
    const double2 h = cherm(U[e], U[(size_t)c * cs + p]);
    acc += wy[l] * (h.x * h.x + h.y * h.y);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_energy_final(const double* __restrict__ partial, int n, double scale, double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// ===========================================================================
// host side
// ===========================================================================
static thread_local std::string g_err;

#define JSD_FAIL(code, ...)                                 \
  do {                                                      \
    char _b[512];                                           \
    snprintf(_b, sizeof(_b), __VA_ARGS__);                  \
    g_err = _b;                                             \
    return code;                                            \
  } while (0)

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t _e = (x);                                                                           \
    if (_e != hipSuccess) JSD_FAIL(JSD_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(_e), \
                                   __FILE__, __LINE__);                                            \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};

struct jsd_plan {
  jsd_plan_desc d;
  int M, N, Z, K, KZ, Ka, Kb, NT, KS;
  int ZT;      // row pitch of the x-z intermediates T, V (Z rounded up to 8: 128-byte rows)
  size_t fs;   // complex elements per field
  double s_inv, s_fwd, dpdx_scale;
  std::vector<void*> allocs;
  // device tables
  double *DF1, *DF2, *DFT1, *D_R, *D2_R, *RB_R, *LI_R, *PO_R;  // adjoint: D^T tiles, row-major tables
  double *YF1, *YF2;  // parity-block fragments of D and D^2 for k_ypar
  YparArgs yp1, yp2;  // their block geometry
  YparArgs yp1t;      // D^T in the same parity form (adjoint plans; the batched D^T accumulation)
  double* YF1T;
  double *YS1, *YS2;  // the same operators in k_ypar_stream's layout
  double* YS1T;       // D^T in that layout (adjoint plans): u_bar += D^T Du_bar
  double* TF;         // eigen-form adjoint plans: transposed solve blocks (k_solve_T8p)
  int solveT8;        // 1: k_solve_T8p (default when TF is set); 0: k_solve_T + the h_v / h_g transpose
                      // (JSD_SOLVET=table, a diagnostic switch)
  YsArgs ys1t;
  int yst_grid;       // persistent grid of the accumulating k_ypar_stream (0: k_ygemm<GEMM_ACC> instead)
  double* YSZ;        // zeros read by its lanes outside the data
  YsArgs ys1, ys2;
  int ys_grid[2];     // persistent grid of k_ypar_stream for D, D^2 (0: k_ypar instead)
  int zreg;           // 1: register z passes for lengths 128 and 256 (default); 0: the LDS kernels
                      // (JSD_ZPASS=lds, a diagnostic switch)
  double kz1;         // the z wavenumber of stored index 1
  int t5;             // NS / perturbation with a register z pass: five packed x transforms (the z pass
                      // forms d_z u_c itself); 0 under JSD_T5=0 (a diagnostic switch)
  double *VA, *INFL;                                           // adjoint: v_a, inverse influence matrices
  int adjoint;         // the discrete-adjoint tables are present
  double2* SS;         // adjoint of a nonlinear step: the three recomputed substep states
  double2* TS;         // and the inverse x transforms of substeps 0 and 1, kept from the recomputation
  double *DEO, *DOE, *D2E, *D2O, *RBP, *LIP, *POP, *CW, *RW, *VAP, *VINV;  // parity blocks of the solve
  double *EVF, *EIF, *LAM;  // eigen-form solve: E, E^-1 fragments [even, odd], eigenvalues [2][PH]
  double *FF, *GF, *CWE, *RWE;  // E^-1 D and D E fragments, wall columns / rows in eigen coordinates
  double *VD2, *Dm;             // D^2 v_a parity tables with wall parts; D row-major (mean-flow derivatives)
  double2* HGC;                 // coupling_term = False: the h_g addend of the current substep
  double2 *HGS, *SRC, *S00;     // source term / enforce_base / external terms (allocated on first use)
  int dual_cc;                  // dual plan, coupling_term = False: H_1 += hat(U' v_x) before the solve
  double c_impl[3];         // beta_s dt / Re of the Helmholtz operators
  double *EH_R, *EG_R, *LAMR;  // adjoint eigen form: full-row E and E^-1 (row-major), eigenvalue per row
  double *RNS, *LNS, *kx, *kz, *Ub, *dUb, *wy;
  int nE, nO;
  int *permx, *permz, *ipermx, *ipermz;
  double2 *twx, *twz;
  FftPlan fpx, fpz;
  int generic_fft;  // 1: run-time-length FFT kernels only (JSD_FFT_GENERIC=1, a diagnostic switch)
  int solve8;       // eigen form, pencils per solve workgroup and waves: 2 (default) 8 pencils, 8 waves, one
                    // parity per wave; 1 8 pencils, 4 waves; 0 4 pencils (JSD_SOLVE=8p / 8 / 4 or JSD_SOLVE4=1,
                    // diagnostic switches)
  int xreg;         // x-pass kernels: 1 (default) the k_fft_x_*_r kernels, inverse inputs scattered through LDS;
                    // 2 the inverse's first stage in registers too; 0 the k_fft_x_*_t kernels (JSD_XPASS=0/1/2)
  // workspaces
  double2 *DU, *H, *HV, *HG, *HVo, *HGo, *CONV, *CONVo, *T, *V, *WF, *DWF;
  double* partial;
  double* eout;
  int npartial;
  // optional per-kernel-class timing with HIP events on the launch stream
  int prof_on;  // bit mask of the bracketed kernel classes
  std::vector<hipEvent_t> prof_ev[JSD_NPROF];
  size_t prof_used[JSD_NPROF];
};

static void prof_begin(jsd_plan* p, int cls, hipStream_t st) {
  if (!(p->prof_on >> cls & 1)) return;
  auto& v = p->prof_ev[cls];
  const size_t k = p->prof_used[cls];
  if (v.size() < k + 2) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    v.push_back(a);
    v.push_back(b);
  }
  (void)hipEventRecord(v[k], st);
}
static void prof_end(jsd_plan* p, int cls, hipStream_t st) {
  if (!(p->prof_on >> cls & 1)) return;
  auto& v = p->prof_ev[cls];
  const size_t k = p->prof_used[cls];
  if (v.size() < k + 2) return;
  (void)hipEventRecord(v[k + 1], st);
  p->prof_used[cls] = k + 2;
}

static int dalloc(jsd_plan* p, void** ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) JSD_FAIL(JSD_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  p->allocs.push_back(*ptr);
  return JSD_OK;
}

template <typename T>
static int upload(jsd_plan* p, T** dst, const T* src, size_t count) {
  int rc = dalloc(p, (void**)dst, count * sizeof(T));
  if (rc) return rc;
  HIPCHK(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return JSD_OK;
}

static void transpose_sq(const double* src, double* dst, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) dst[(size_t)j * n + i] = src[(size_t)i * n + j];
}

static int factor_plan(int p, std::vector<int>& rad) {
  static const int cand[6] = {8, 7, 5, 4, 3, 2};
  int rest = p;
  rad.clear();
  for (int c : cand)
    while (rest % c == 0 && rest > 1) {
      rad.push_back(c);
      rest /= c;
    }
  return rest == 1 && (int)rad.size() <= JSD_MAXRAD;
}

// DIT input order (outermost radix = last entry), same recursion as the host plan
static void dit_order(int len, const std::vector<int>& rs, int nrs, std::vector<int>& out) {
  if (len == 1) {
    out.assign(1, 0);
    return;
  }
  const int r = rs[nrs - 1];
  const int q = len / r;
  std::vector<int> sub;
  dit_order(q, rs, nrs - 1, sub);
  out.clear();
  for (int n2 = 0; n2 < r; ++n2)
    for (int n1 : sub) out.push_back(r * n1 + n2);
}

static int make_fft(jsd_plan* p, int len, FftPlan* fp, int** perm, int** iperm, double2** tw) {
  std::vector<int> rad;
  if (!factor_plan(len, rad)) JSD_FAIL(JSD_ERR_ARG, "transform length %d is not a product of 2,3,4,5,7,8", len);
  std::vector<int> order;
  dit_order(len, rad, (int)rad.size(), order);
  std::vector<double2> t(len);
  for (int m = 0; m < len; ++m) {
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)len;
    t[m] = make_double2((double)cosl(ang), -(double)sinl(ang));
  }
  std::vector<int> inv(len);
  for (int q = 0; q < len; ++q) inv[order[q]] = q;
  int rc = upload(p, perm, order.data(), len);
  if (rc) return rc;
  rc = upload(p, iperm, inv.data(), len);
  if (rc) return rc;
  rc = upload(p, tw, t.data(), len);
  if (rc) return rc;
  fp->p = len;
  fp->nrad = (int)rad.size();
  for (int s = 0; s < fp->nrad; ++s) fp->rad[s] = rad[s];
  fp->perm = *perm;
  fp->iperm = *iperm;
  fp->tw = *tw;
  return JSD_OK;
}

// fragment-tiled D^order for the MFMA A operand
static void tile_D(const double* Dm, int N, int NT, int KS, std::vector<double>& out) {
  out.assign((size_t)NT * KS * 64, 0.0);
  for (int lt = 0; lt < NT; ++lt)
    for (int ks = 0; ks < KS; ++ks)
      for (int lane = 0; lane < 64; ++lane) {
        const int r = lt * 16 + (lane & 15), c = ks * 4 + (lane >> 4);
        out[((size_t)lt * KS + ks) * 64 + lane] = (r < N && c < N) ? Dm[(size_t)r * N + c] : 0.0;
      }
}

// ---------------------------------------------------------------------------
// parity blocks of the wall-normal operators (see jsd_solve.inc)
// ---------------------------------------------------------------------------
// Half sizes for N points: interior l = 1..N-2 pairs with l' = N-1-l; the odd
// half has N/2 - 1 entries, the even half one more when N is odd (centre).
static inline int half_odd(int N) { return N / 2 - 1; }
static inline int half_even(int N) { return N / 2 - 1 + (N & 1); }

// Row l (1-based interior or 0 = wall pair) of the block of A mapping the
// half of parity qi: entry g couples to interior column m = g+1 and its mirror.
// A is symmetrised first: (A[r][c] + sgn A[N-1-r][N-1-c]) / 2.
static double sym_entry(const double* A, int N, int sgn, int r, int c) {
  return 0.5 * (A[(size_t)r * N + c] + sgn * A[(size_t)(N - 1 - r) * N + (N - 1 - c)]);
}
static double parity_entry(const double* A, int N, int sgn, int qi, int l, int g) {
  const int m = g + 1, mp = N - 1 - m;
  if (m == mp) return sym_entry(A, N, sgn, l, m);
  return sym_entry(A, N, sgn, l, m) + (qi ? -1.0 : 1.0) * sym_entry(A, N, sgn, l, mp);
}
// dense PH x PH block (rows: output half qo, columns: input half qi)
static void parity_dense(const double* A, int N, int sgn, int qi, int qo, double* blk) {
  const int no = qo ? half_odd(N) : half_even(N), ni = qi ? half_odd(N) : half_even(N);
  for (int h = 0; h < PH; ++h)
    for (int g = 0; g < PH; ++g) blk[h * PH + g] = (h < no && g < ni) ? parity_entry(A, N, sgn, qi, h + 1, g) : 0.0;
}
// pre-tiled MFMA A fragments of a dense PH x PH block
static void parity_tile(const double* blk, double* frag) {
  for (int t = 0; t < 4; ++t)
    for (int ks = 0; ks < 16; ++ks)
      for (int lane = 0; lane < 64; ++lane)
        frag[(t * 16 + ks) * 64 + lane] = blk[(t * 16 + (lane & 15)) * PH + ks * 4 + (lane >> 4)];
}
static void parity_frag(const double* A, int N, int sgn, int qi, int qo, double* frag) {
  std::vector<double> blk((size_t)PH * PH);
  parity_dense(A, N, sgn, qi, qo, blk.data());
  parity_tile(blk.data(), frag);
}

static const double RK_BET[3] = {37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0};

// Eigen-form solve tables.  The interior block of every implicit y operator
// is a function of the interior D^2, which on each parity half is
// E diag(lam) E^-1 (host precompute, desc->y_eig): the Helmholtz inverse is
// E diag(1 / ((1 + c k^2) - c lam)) E^-1 with c = beta_s |dt / Re| and the
// Poisson inverse E diag(1 / (lam - k^2)) E^-1.  The kernel then needs the
// two shared blocks instead of one lhs_inv and one poisson block per |k|.
// The factorisation is checked against the lhs_inv / poisson tables on a
// sample of wavenumbers and every substep before it replaces them.
static int eigen_tables(jsd_plan* p, const jsd_plan_desc* d, int N, size_t nk, const double* deo, const double* doe,
                        const double* cw, const double* rw, const double* rw2) {
  const int nE = p->nE, nO = p->nO;
  const double* f = d->y_eig;
  const double* Ee = f;
  const double* Eo = Ee + (size_t)nE * nE;
  const double* Xe = Eo + (size_t)nO * nO;
  const double* Xo = Xe + (size_t)nE * nE;
  const double* le = Xo + (size_t)nO * nO;
  const double* lo = le + nE;
  std::vector<double> blk((size_t)PH * PH), frag(2 * (size_t)PFRAG), ifrag(2 * (size_t)PFRAG), lam(2 * PH, -1.0);
  auto pad = [&](const double* src, int n, double* out) {
    std::fill(blk.begin(), blk.end(), 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) blk[(size_t)i * PH + j] = src[(size_t)i * n + j];
    parity_tile(blk.data(), out);
  };
  pad(Ee, nE, frag.data());
  pad(Eo, nO, frag.data() + PFRAG);
  pad(Xe, nE, ifrag.data());
  pad(Xo, nO, ifrag.data() + PFRAG);
  for (int h = 0; h < nE; ++h) lam[h] = le[h];
  for (int h = 0; h < nO; ++h) lam[PH + h] = lo[h];
  for (int h = 0; h < nE + nO; ++h) {
    const double l = h < nE ? le[h] : lo[h - nE];
    if (!(l < 0.0)) JSD_FAIL(JSD_ERR_ARG, "y_eig: eigenvalue %d of the interior D^2 is %g (must be negative)", h, l);
  }
  for (int s = 0; s < 3; ++s) p->c_impl[s] = RK_BET[s] * fabs(d->dt) / fabs(d->re_tau);
  // check: table application vs eigen form on a fixed pseudo-random vector
  std::vector<double> x(N, 0.0), yt(N), vt(N), ye(N), ve(N), he(PH), ho(PH), z(PH);
  uint64_t st = 0x9e3779b97f4a7c15ull;
  for (int l = 1; l < N - 1; ++l) {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    x[l] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  for (int h = 0; h < nE; ++h) {
    const int l = h + 1, lm = N - 1 - l;
    he[h] = l == lm ? x[l] : 0.5 * (x[l] + x[lm]);
    if (h < nO) ho[h] = 0.5 * (x[l] - x[lm]);
  }
  const size_t nn = (size_t)N * N;
  for (size_t k = 0; k < nk; ++k) {
    // the checked entries: every 13th |k| and the last (the host may fill
    // only these in eigen form, domain_math.check_indices)
    if (k % 13 != 0 && k != nk - 1) continue;
    const int a = (int)(k / p->Kb), b = (int)(k % p->Kb);
    const double k2 = d->kx[a] * d->kx[a] + d->kz[b] * d->kz[b];
    for (int s = 0; s < 3; ++s) {
      const double* L = d->lhs_inv + ((size_t)s * nk + k) * nn;
      const double* P = d->poisson + k * nn;
      for (int i = 0; i < N; ++i) {
        double acc = 0.0;
        for (int j = 0; j < N; ++j) acc += L[(size_t)i * N + j] * x[j];
        yt[i] = acc;
      }
      for (int i = 0; i < N; ++i) {
        double acc = 0.0;
        for (int j = 0; j < N; ++j) acc += P[(size_t)i * N + j] * yt[j];
        vt[i] = acc;
      }
      std::fill(ye.begin(), ye.end(), 0.0);
      std::fill(ve.begin(), ve.end(), 0.0);
      const double c = p->c_impl[s];
      for (int q = 0; q < 2; ++q) {
        const int n = q ? nO : nE;
        const double* E = q ? Eo : Ee;
        const double* X = q ? Xo : Xe;
        const double* lq = q ? lo : le;
        const double* in = q ? ho.data() : he.data();
        for (int i = 0; i < n; ++i) {
          double acc = 0.0;
          for (int j = 0; j < n; ++j) acc += X[(size_t)i * n + j] * in[j];
          z[i] = acc;
        }
        for (int i = 0; i < n; ++i) {
          double om = 0.0, v = 0.0;
          for (int j = 0; j < n; ++j) {
            const double dl = 1.0 / ((1.0 + c * k2) - c * lq[j]);
            om += E[(size_t)i * n + j] * dl * z[j];
            v += E[(size_t)i * n + j] * (dl / (lq[j] - k2)) * z[j];
          }
          const int l = i + 1, lm = N - 1 - l;
          const double sg = q ? -1.0 : 1.0;
          ye[l] += om;
          ve[l] += v;
          if (lm != l) {
            ye[lm] += sg * om;
            ve[lm] += sg * v;
          }
        }
      }
      double ey = 0.0, ny = 0.0, ev = 0.0, nv = 0.0;
      for (int l = 1; l < N - 1; ++l) {
        ey += (ye[l] - yt[l]) * (ye[l] - yt[l]);
        ny += yt[l] * yt[l];
        ev += (ve[l] - vt[l]) * (ve[l] - vt[l]);
        nv += vt[l] * vt[l];
      }
      const double ry = sqrt(ey / ny), rv = sqrt(ev / nv);
      if (!(ry < 1e-9) || !(rv < 1e-9))
        JSD_FAIL(JSD_ERR_ARG,
                 "y_eig does not reproduce the lhs_inv / poisson tables at |k| index %zu, substep %d "
                 "(rel. errors %.3g, %.3g): tables not built from this D, dt and Re",
                 k, s, ry, rv);
    }
  }
  // composite blocks of the eigen-coordinate solve (extended-precision sums):
  // F = E^-1 D (even coordinates <- odd rows: X_e DEO; odd <- even: X_o DOE),
  // G = D E (even rows <- odd coordinates: DEO E_o; odd <- even: DOE E_e),
  // the wall columns CW and wall rows RW of D, D^2 carried to the coordinates
  std::vector<double> ff(2 * (size_t)PFRAG), gf(2 * (size_t)PFRAG), cwe(4 * PH, 0.0), rwe(4 * PH, 0.0);
  auto prod = [&](const double* A, int lda, int ar, int ak, const double* B, int ldb, int bc, double* out) {
    std::fill(blk.begin(), blk.end(), 0.0);
    for (int i = 0; i < ar; ++i)
      for (int j = 0; j < bc; ++j) {
        long double acc = 0.0L;
        for (int k = 0; k < ak; ++k) acc += (long double)A[(size_t)i * lda + k] * B[(size_t)k * ldb + j];
        blk[(size_t)i * PH + j] = (double)acc;
      }
    parity_tile(blk.data(), out);
  };
  prod(Xe, nE, nE, nE, deo, PH, nO, ff.data());           // X_e (nE x nE) . DEO (nE x nO)
  prod(Xo, nO, nO, nO, doe, PH, nE, ff.data() + PFRAG);   // X_o (nO x nO) . DOE (nO x nE)
  prod(deo, PH, nE, nO, Eo, nO, nO, gf.data());           // DEO (nE x nO) . E_o (nO x nO)
  prod(doe, PH, nO, nE, Ee, nE, nE, gf.data() + PFRAG);   // DOE (nO x nE) . E_e (nE x nE)
  for (int q = 0; q < 2; ++q) {
    const int n = q ? nO : nE;
    const double* X = q ? Xo : Xe;
    for (int w = 0; w < 2; ++w) {
      const double* c = cw + (size_t)(2 * w + q) * PH;  // CW[q] (D), CW[2 + q] (D^2)
      for (int i = 0; i < n; ++i) {
        long double acc = 0.0L;
        for (int k = 0; k < n; ++k) acc += (long double)X[(size_t)i * n + k] * c[k];
        cwe[(size_t)(2 * w + q) * PH + i] = (double)acc;
      }
    }
  }
  for (int c = 0; c < nO; ++c) {
    long double acc = 0.0L;
    for (int h = 0; h < nO; ++h) acc += (long double)rw[h] * Eo[(size_t)h * nO + c];
    rwe[c] = (double)acc;
  }
  for (int c = 0; c < nE; ++c) {
    long double acc = 0.0L;
    for (int h = 0; h < nE; ++h) acc += (long double)rw[PH + h] * Ee[(size_t)h * nE + c];
    rwe[PH + c] = (double)acc;
  }
  // D^2 wall rows: RWE[2] = RW2[0] E_e (even wall part), RWE[3] = RW2[1] E_o
  for (int c = 0; c < nE; ++c) {
    long double acc = 0.0L;
    for (int h = 0; h < nE; ++h) acc += (long double)rw2[h] * Ee[(size_t)h * nE + c];
    rwe[2 * PH + c] = (double)acc;
  }
  for (int c = 0; c < nO; ++c) {
    long double acc = 0.0L;
    for (int h = 0; h < nO; ++h) acc += (long double)rw2[PH + h] * Eo[(size_t)h * nO + c];
    rwe[3 * PH + c] = (double)acc;
  }
  int rc = upload(p, &p->EVF, frag.data(), frag.size());
  if (!rc) rc = upload(p, &p->EIF, ifrag.data(), ifrag.size());
  if (!rc) rc = upload(p, &p->LAM, lam.data(), lam.size());
  if (!rc) rc = upload(p, &p->FF, ff.data(), ff.size());
  if (!rc) rc = upload(p, &p->GF, gf.data(), gf.size());
  if (!rc) rc = upload(p, &p->CWE, cwe.data(), cwe.size());
  if (!rc) rc = upload(p, &p->RWE, rwe.data(), rwe.size());
  // the transposed blocks of the eigen-form adjoint solve (k_solve_T8p):
  // [E^T][(D E)^T][E^-T][(E^-1 D)^T], each for output parity e, o; the
  // parity-swapping blocks transpose the block of the other output parity
  p->TF = nullptr;
  const bool adj = (d->kind == JSD_KIND_PERTURBATION && d->linearise) ||
                   (d->adjoint && (d->kind == JSD_KIND_NS || d->kind == JSD_KIND_PERTURBATION));
  if (!rc && adj) {
    std::vector<double> tf(8 * (size_t)PFRAG), dense((size_t)PH * PH), tr((size_t)PH * PH);
    auto tile_T = [&](const double* fr, double* out) {
      for (int t = 0; t < 4; ++t)
        for (int ks = 0; ks < 16; ++ks)
          for (int lane = 0; lane < 64; ++lane)
            dense[(size_t)(t * 16 + (lane & 15)) * PH + ks * 4 + (lane >> 4)] = fr[(t * 16 + ks) * 64 + lane];
      for (int i = 0; i < PH; ++i)
        for (int j = 0; j < PH; ++j) tr[(size_t)i * PH + j] = dense[(size_t)j * PH + i];
      parity_tile(tr.data(), out);
    };
    for (int q = 0; q < 2; ++q) {
      tile_T(frag.data() + q * PFRAG, tf.data() + (0 * 2 + q) * PFRAG);
      tile_T(gf.data() + (1 - q) * PFRAG, tf.data() + (1 * 2 + q) * PFRAG);
      tile_T(ifrag.data() + q * PFRAG, tf.data() + (2 * 2 + q) * PFRAG);
      tile_T(ff.data() + (1 - q) * PFRAG, tf.data() + (3 * 2 + q) * PFRAG);
    }
    rc = upload(p, &p->TF, tf.data(), tf.size());
  }
  return rc;
}

// The eigen form on full rows, for the adjoint solve (which works on full
// rows with row-major tables applied transposed): coordinate c of the
// interior (c < nE: even eigenvector c, else odd eigenvector c - nE) sits in
// row 1 + c, walls zero.  H maps coordinates to rows (H[l][1+c]: the
// eigenvector spread over l and its mirror, with the odd sign), G maps rows
// to coordinates (E^-1 applied to the parity halves e = (x_l + x_l')/2,
// o = (x_l - x_l')/2), so the interior Helmholtz inverse is H diag(dl) G.
static int eigen_tables_full(jsd_plan* p, const jsd_plan_desc* d, int N) {
  const int nE = p->nE, nO = p->nO;
  const double* Ee = d->y_eig;
  const double* Eo = Ee + (size_t)nE * nE;
  const double* Xe = Eo + (size_t)nO * nO;
  const double* Xo = Xe + (size_t)nE * nE;
  const double* le = Xo + (size_t)nO * nO;
  const double* lo = le + nE;
  const size_t nn = (size_t)N * N;
  std::vector<double> H(nn, 0.0), G(nn, 0.0), lam(N, -1.0);
  for (int h = 0; h < nE; ++h) {
    const int l = h + 1, lm = N - 1 - l;
    const bool centre = l == lm;
    for (int c = 0; c < nE; ++c) {
      H[(size_t)l * N + 1 + c] = Ee[(size_t)h * nE + c];
      if (!centre) H[(size_t)lm * N + 1 + c] = Ee[(size_t)h * nE + c];
      const double x = Xe[(size_t)c * nE + h];
      if (centre) {
        G[(size_t)(1 + c) * N + l] = x;
      } else {
        G[(size_t)(1 + c) * N + l] = 0.5 * x;
        G[(size_t)(1 + c) * N + lm] = 0.5 * x;
      }
    }
  }
  for (int h = 0; h < nO; ++h) {
    const int l = h + 1, lm = N - 1 - l;
    for (int c = 0; c < nO; ++c) {
      const int r = 1 + nE + c;
      H[(size_t)l * N + r] = Eo[(size_t)h * nO + c];
      H[(size_t)lm * N + r] = -Eo[(size_t)h * nO + c];
      const double x = Xo[(size_t)c * nO + h];
      G[(size_t)r * N + l] = 0.5 * x;
      G[(size_t)r * N + lm] = -0.5 * x;
    }
  }
  for (int c = 0; c < nE; ++c) lam[1 + c] = le[c];
  for (int c = 0; c < nO; ++c) lam[1 + nE + c] = lo[c];
  int rc = upload(p, &p->EH_R, H.data(), nn);
  if (!rc) rc = upload(p, &p->EG_R, G.data(), nn);
  if (!rc) rc = upload(p, &p->LAMR, lam.data(), (size_t)N);
  return rc;
}

// Fragments of the full-row parity blocks of A (walls included: half index h
// is row h and its mirror N-1-h), sgn = -1 for D (parities swap), +1 for D^2.
// k_ypar_stream's operator layout (see there) from the same parity-block
// entries as ypar_frags.
static void ystream_frags(const double* A, int N, int sgn, std::vector<double>& F, YsArgs& G) {
  const int nEf = (N + 1) / 2, nOf = N / 2;
  G.nEf = nEf;
  G.nOf = nOf;
  G.KL = (nEf + 3) / 4;
  G.xE = nEf > 64;
  G.xO = nOf > 64;
  G.anti = sgn < 0 ? 1 : 0;
  const int KL = G.KL;
  F.assign(ystream_lds_doubles(KL), 0.0);
  auto entry = [&](int h, int qi, int g) {
    const int m = g, mp = N - 1 - g;
    if (m == mp) return sym_entry(A, N, sgn, h, m);
    return sym_entry(A, N, sgn, h, m) + (qi ? -1.0 : 1.0) * sym_entry(A, N, sgn, h, mp);
  };
  for (int qo = 0; qo < 2; ++qo) {
    const int qi = G.anti ? 1 - qo : qo;
    const int no = qo ? nOf : nEf, ni = qi ? nOf : nEf;
    for (int ks = 0; ks < KL; ++ks) {
      for (int t = 0; t < 4; ++t)
        for (int lane = 0; lane < 64; ++lane) {
          const int h = t * 16 + (lane & 15), g = ks * 4 + (lane >> 4);
          F[((size_t)ks * 8 + qo * 4 + t) * 64 + lane] = (h < no && g < ni) ? entry(h, qi, g) : 0.0;
        }
      for (int kr = 0; kr < 4; ++kr) {
        const int g = ks * 4 + kr;
        F[(size_t)KL * 512 + ks * 8 + qo * 4 + kr] = (64 < no && g < ni) ? entry(64, qi, g) : 0.0;
      }
    }
  }
}

// Persistent grids of k_ypar_stream for D and D^2: resident workgroups per
// CU at its LDS size times the CUs.  Zero (k_ypar runs instead) for Ny < 25
// (fewer k-steps than the prefetch distance), Ny > 130, or when the device
// cannot hold the operator in LDS.
static int ypar_stream_setup(jsd_plan* p, const double* D, const double* D2) {
  p->ys_grid[0] = p->ys_grid[1] = 0;
  const int N = p->N;
  if ((N + 1) / 2 > 65) return JSD_OK;
  std::vector<double> f;
  ystream_frags(D, N, -1, f, p->ys1);
  int rc = upload(p, &p->YS1, f.data(), f.size());
  if (rc) return rc;
  ystream_frags(D2, N, +1, f, p->ys2);
  rc = upload(p, &p->YS2, f.data(), f.size());
  if (rc) return rc;
  const double z8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  rc = upload(p, &p->YSZ, z8, 8);
  if (rc) return rc;
  if (p->ys1.KL < YS_PD) return JSD_OK;
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const size_t sh = ystream_lds_doubles(p->ys1.KL) * sizeof(double);
  for (int o = 0; o < 2; ++o) {
    const bool full = p->ys1.nEf >= 64;
    const void* fn = o == 0 ? (full ? (const void*)k_ypar_stream<true, true> : (const void*)k_ypar_stream<true, false>)
                            : (full ? (const void*)k_ypar_stream<false, true> : (const void*)k_ypar_stream<false, false>);
    if (sh > 64 * 1024 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, YS_THREADS, sh) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    if (nb > 0 && ncu > 0) p->ys_grid[o] = nb * ncu;
  }
  return JSD_OK;
}

// D^T (centro-antisymmetric like D) for the accumulating streaming kernel of
// the adjoint; yst_grid = 0 leaves k_ygemm<GEMM_ACC> in place.
static int ypar_stream_setup_T(jsd_plan* p, const double* DT) {
  p->yst_grid = 0;
  p->YS1T = nullptr;
  const int N = p->N;
  if ((N + 1) / 2 > 65 || p->ys_grid[0] == 0) return JSD_OK;
  std::vector<double> f;
  ystream_frags(DT, N, -1, f, p->ys1t);
  int rc = upload(p, &p->YS1T, f.data(), f.size());
  if (rc) return rc;
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const size_t sh = ystream_lds_doubles(p->ys1t.KL) * sizeof(double);
  const bool full = p->ys1t.nEf >= 64;
  const void* fn = full ? (const void*)k_ypar_stream<true, true, true> : (const void*)k_ypar_stream<true, false, true>;
  if (sh > 64 * 1024 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh) != hipSuccess) {
    (void)hipGetLastError();
    return JSD_OK;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, YS_THREADS, sh) != hipSuccess) {
    (void)hipGetLastError();
    return JSD_OK;
  }
  if (nb > 0 && ncu > 0) p->yst_grid = nb * ncu;
  return JSD_OK;
}

static void ypar_frags(const double* A, int N, int sgn, std::vector<double>& F, YparArgs& G) {
  const int nEf = (N + 1) / 2, nOf = N / 2;
  G.nEf = nEf;
  G.nOf = nOf;
  G.KE = (nEf + 3) / 4;
  G.KO = (nOf + 3) / 4;
  G.ET = (nEf + 15) / 16;
  G.OT = (nOf + 15) / 16;
  G.anti = sgn < 0 ? 1 : 0;
  G.acc = 0;
  G.KIE = G.anti ? G.KO : G.KE;
  G.KIO = G.anti ? G.KE : G.KO;
  F.assign((size_t)(G.ET * G.KIE + G.OT * G.KIO) * 64, 0.0);
  auto entry = [&](int h, int qi, int g) {
    const int m = g, mp = N - 1 - g;
    if (m == mp) return sym_entry(A, N, sgn, h, m);
    return sym_entry(A, N, sgn, h, m) + (qi ? -1.0 : 1.0) * sym_entry(A, N, sgn, h, mp);
  };
  for (int qo = 0; qo < 2; ++qo) {
    const int qi = G.anti ? 1 - qo : qo;
    const int no = qo ? nOf : nEf, ni = qi ? nOf : nEf;
    const int T = qo ? G.OT : G.ET, K = qo ? G.KIO : G.KIE;
    double* f = F.data() + (qo ? (size_t)G.ET * G.KIE * 64 : 0);
    for (int t = 0; t < T; ++t)
      for (int ks = 0; ks < K; ++ks)
        for (int lane = 0; lane < 64; ++lane) {
          const int h = t * 16 + (lane & 15), g = ks * 4 + (lane >> 4);
          f[((size_t)t * K + ks) * 64 + lane] = (h < no && g < ni) ? entry(h, qi, g) : 0.0;
        }
  }
}

extern "C" int jsd_version(void) { return 1; }
extern "C" const char* jsd_last_error(void) { return g_err.c_str(); }

extern "C" int jsd_plan_destroy(jsd_plan* p) {
  if (!p) return JSD_OK;
  for (int c = 0; c < JSD_NPROF; ++c)
    for (hipEvent_t e : p->prof_ev[c]) (void)hipEventDestroy(e);
  for (void* a : p->allocs) (void)hipFree(a);
  delete p;
  return JSD_OK;
}

extern "C" int jsd_plan_create(const jsd_plan_desc* d, jsd_plan** out) {
  if (!d || !out) JSD_FAIL(JSD_ERR_ARG, "null argument");
  *out = nullptr;
  const int M = d->nxm, N = d->ny, Z = d->nzm;
  if (M < 1 || Z < 1 || (M % 2) == 0 || (Z % 2) == 0) JSD_FAIL(JSD_ERR_ARG, "mode counts must be odd (%d, %d)", M, Z);
  if (N < 4 || N > JSD_NYMAX) JSD_FAIL(JSD_ERR_ARG, "ny=%d outside [4, %d]", N, JSD_NYMAX);
  if (d->kind < 0 || d->kind > 2) JSD_FAIL(JSD_ERR_ARG, "unknown kind %d", d->kind);
  if (d->ppx < 3 * (M / 2) + 1 || d->ppz < 3 * (Z / 2) + 1)
    JSD_FAIL(JSD_ERR_ARG, "transform grid (%d, %d) cannot dealias (%d, %d) modes", d->ppx, d->ppz, M, Z);
  if (!d->D || !d->kx || !d->kz || !d->lhs_inv || !d->poisson || !d->v_a || !d->infl_inv || !d->rhs_base ||
      !d->rhs_ns || !d->lhs_inv_ns || !d->y_weights)
    JSD_FAIL(JSD_ERR_ARG, "missing operator table");
  if (d->kind != JSD_KIND_NS && (!d->base_u || !d->base_du)) JSD_FAIL(JSD_ERR_ARG, "missing base-flow profile");
  jsd_plan* p = new jsd_plan();
  p->d = *d;
  p->TF = nullptr;
  {
    const char* env = getenv("JSD_FFT_GENERIC");
    p->generic_fft = (env && env[0] == '1') ? 1 : 0;
    const char* envt = getenv("JSD_SOLVET");
    p->solveT8 = (envt && strcmp(envt, "table") == 0) ? 0 : 1;
    if (d->ny > 130) p->solveT8 = 0;
    const char* env4 = getenv("JSD_SOLVE4");
    const char* envs = getenv("JSD_SOLVE");
    p->solve8 = 2;
    if (envs && strcmp(envs, "8") == 0) p->solve8 = 1;
    if ((envs && strcmp(envs, "4") == 0) || (env4 && env4[0] == '1')) p->solve8 = 0;
    if (d->ny > 130) p->solve8 = std::min(p->solve8, 1);  // one interior row pair per thread
    const char* envx = getenv("JSD_XPASS");
    p->xreg = (envx && envx[0] >= '0' && envx[0] <= '2') ? envx[0] - '0' : 1;
    const char* envzp = getenv("JSD_ZPASS");
    p->zreg = (envzp && strcmp(envzp, "lds") == 0) ? 0 : 1;
  }
  p->M = M;
  p->N = N;
  p->Z = Z;
  p->K = M / 2;
  p->KZ = Z / 2;
  p->ZT = (Z + ZCOL_SHIFT + 7) & ~7;  // row pitch of the transformed fields (zcol)
  p->Ka = p->K + 1;
  p->Kb = p->KZ + 1;
  p->NT = (N + 15) / 16;
  p->KS = (N + 3) / 4;
  p->fs = (size_t)M * N * Z;
  const double sl = (d->lx / (2 * M_PI)) * (d->lz / (2 * M_PI));
  p->s_inv = sl / sqrt((double)d->px * (double)d->pz);
  p->s_fwd = sqrt((double)d->px * (double)d->pz) / (sl * (double)d->ppx * (double)d->ppz);
  p->dpdx_scale = pow((d->px * pow(2 * M_PI / d->lx, 2)) * (d->pz * pow(2 * M_PI / d->lz, 2)), 0.5);
  int rc = JSD_OK;
#define TRY(x)            \
  do {                    \
    rc = (x);             \
    if (rc) {             \
      jsd_plan_destroy(p); \
      return rc;          \
    }                     \
  } while (0)
  const size_t nn = (size_t)N * N;
  const size_t nk = (size_t)p->Ka * p->Kb;
  // D, D^2 (numpy matrix_power order: D @ D)
  std::vector<double> D2(nn, 0.0), tmp(nn);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < N; ++k) {
      const double a = d->D[(size_t)i * N + k];
      for (int j = 0; j < N; ++j) D2[(size_t)i * N + j] += a * d->D[(size_t)k * N + j];
    }
  std::vector<double> tiled;
  tile_D(d->D, N, p->NT, p->KS, tiled);
  TRY(upload(p, &p->DF1, tiled.data(), tiled.size()));
  tile_D(D2.data(), N, p->NT, p->KS, tiled);
  TRY(upload(p, &p->DF2, tiled.data(), tiled.size()));
  ypar_frags(d->D, N, -1, tiled, p->yp1);
  TRY(upload(p, &p->YF1, tiled.data(), tiled.size()));
  ypar_frags(D2.data(), N, +1, tiled, p->yp2);
  TRY(upload(p, &p->YF2, tiled.data(), tiled.size()));
  TRY(ypar_stream_setup(p, d->D, D2.data()));
  {
    // parity blocks: D (centro-antisymmetric), D^2, rhs_base[s], lhs_inv[s][k],
    // poisson[k] (centro-symmetric), wall rows/columns, v_a in parity form
    p->nE = half_even(N);
    p->nO = half_odd(N);
    if (p->nE > PH) JSD_FAIL(JSD_ERR_ARG, "Ny = %d exceeds the solver's %d-row parity blocks", N, 2 * PH + 2);
    std::vector<double> f(2 * (size_t)PFRAG);
    parity_frag(d->D, N, -1, 1, 0, f.data());
    TRY(upload(p, &p->DEO, f.data(), PFRAG));
    parity_frag(d->D, N, -1, 0, 1, f.data());
    TRY(upload(p, &p->DOE, f.data(), PFRAG));
    parity_frag(D2.data(), N, +1, 0, 0, f.data());
    parity_frag(D2.data(), N, +1, 1, 1, f.data() + PFRAG);
    TRY(upload(p, &p->D2E, f.data(), 2 * (size_t)PFRAG));
    p->D2O = p->D2E + PFRAG;
    {
      std::vector<double> rb(3 * 2 * (size_t)PFRAG);
      for (int s = 0; s < 3; ++s) {
        parity_frag(d->rhs_base + s * nn, N, +1, 0, 0, rb.data() + (size_t)s * 2 * PFRAG);
        parity_frag(d->rhs_base + s * nn, N, +1, 1, 1, rb.data() + (size_t)s * 2 * PFRAG + PFRAG);
      }
      TRY(upload(p, &p->RBP, rb.data(), rb.size()));
    }
    p->EVF = p->EIF = p->LAM = nullptr;
    p->FF = p->GF = p->CWE = p->RWE = nullptr;
    p->LIP = p->POP = nullptr;
    if (!d->y_eig) {
      std::vector<double> li(3 * nk * 2 * (size_t)PFRAG), po(nk * 2 * (size_t)PFRAG);
      for (size_t k = 0; k < 3 * nk; ++k) {
        parity_frag(d->lhs_inv + k * nn, N, +1, 0, 0, li.data() + k * 2 * PFRAG);
        parity_frag(d->lhs_inv + k * nn, N, +1, 1, 1, li.data() + k * 2 * PFRAG + PFRAG);
      }
      for (size_t k = 0; k < nk; ++k) {
        parity_frag(d->poisson + k * nn, N, +1, 0, 0, po.data() + k * 2 * PFRAG);
        parity_frag(d->poisson + k * nn, N, +1, 1, 1, po.data() + k * 2 * PFRAG + PFRAG);
      }
      TRY(upload(p, &p->LIP, li.data(), li.size()));
      TRY(upload(p, &p->POP, po.data(), po.size()));
    }
    // wall columns (input = wall pair (0, N-1)) and the wall row of D
    std::vector<double> cw(4 * PH, 0.0), rw(2 * PH, 0.0);
    for (int h = 0; h < p->nE; ++h) {
      const int l = h + 1;
      cw[0 * PH + h] = sym_entry(d->D, N, -1, l, 0) - sym_entry(d->D, N, -1, l, N - 1);
      cw[2 * PH + h] = sym_entry(D2.data(), N, +1, l, 0) + sym_entry(D2.data(), N, +1, l, N - 1);
      rw[1 * PH + h] = parity_entry(d->D, N, -1, 0, 0, h);
    }
    for (int h = 0; h < p->nO; ++h) {
      const int l = h + 1;
      cw[1 * PH + h] = sym_entry(d->D, N, -1, l, 0) + sym_entry(d->D, N, -1, l, N - 1);
      cw[3 * PH + h] = sym_entry(D2.data(), N, +1, l, 0) - sym_entry(D2.data(), N, +1, l, N - 1);
      rw[0 * PH + h] = parity_entry(d->D, N, -1, 1, 0, h);
    }
    TRY(upload(p, &p->CW, cw.data(), cw.size()));
    TRY(upload(p, &p->RW, rw.data(), rw.size()));
    // v_a[s][k] in parity form, its derivative and the diagonal influence inverses
    std::vector<double> deo((size_t)PH * PH), doe((size_t)PH * PH);
    parity_dense(d->D, N, -1, 1, 0, deo.data());
    parity_dense(d->D, N, -1, 0, 1, doe.data());
    // wall rows of D^2 over the interior parity halves: even wall part from
    // even input, odd wall part from odd input (D^2 keeps parities)
    std::vector<double> rw2(2 * PH, 0.0);
    for (int h = 0; h < p->nE; ++h) rw2[h] = parity_entry(D2.data(), N, +1, 0, 0, h);
    for (int h = 0; h < p->nO; ++h) rw2[PH + h] = parity_entry(D2.data(), N, +1, 1, 0, h);
    if (d->y_eig)
      TRY(eigen_tables(p, d, N, nk, deo.data(), doe.data(), cw.data(), rw.data(), rw2.data()));
    std::vector<double> vap(3 * nk * 4 * (size_t)PH, 0.0), vinv(3 * nk * 2);
    for (size_t k = 0; k < 3 * nk; ++k) {
      const double* va = d->v_a + k * N;
      double* o = vap.data() + k * 4 * PH;
      for (int h = 0; h < p->nE; ++h) {
        const int l = h + 1, lm = N - 1 - l;
        o[h] = (l == lm) ? va[l] : 0.5 * (va[l] + va[lm]);
        if (h < p->nO) o[PH + h] = 0.5 * (va[l] - va[lm]);
      }
      long double se = 0.0L, so = 0.0L;
      for (int h = 0; h < PH; ++h) {
        long double a1 = 0.0L, a2 = 0.0L;
        for (int g2 = 0; g2 < PH; ++g2) {
          a1 += (long double)deo[(size_t)h * PH + g2] * o[PH + g2];
          a2 += (long double)doe[(size_t)h * PH + g2] * o[g2];
        }
        o[2 * PH + h] = (double)a1;
        o[3 * PH + h] = (double)a2;
        se += (long double)rw[1 * PH + h] * o[h];
        so += (long double)rw[0 * PH + h] * o[PH + h];
      }
      if (se == 0.0L || so == 0.0L) JSD_FAIL(JSD_ERR_ARG, "singular influence matrix at table %zu", k);
      vinv[2 * k] = (double)(1.0L / se);
      vinv[2 * k + 1] = (double)(1.0L / so);
    }
    TRY(upload(p, &p->VAP, vap.data(), vap.size()));
    TRY(upload(p, &p->VINV, vinv.data(), vinv.size()));
    p->VD2 = p->Dm = nullptr;
    if (p->EVF) {
      // D^2 v_a[s][k] on the interior parity halves and its wall parts (the
      // eigen-form solve also returns the y-derivatives of the new state)
      std::vector<double> d2e((size_t)PH * PH), d2o((size_t)PH * PH), vd2(3 * nk * (2 * (size_t)PH + 2), 0.0);
      parity_dense(D2.data(), N, +1, 0, 0, d2e.data());
      parity_dense(D2.data(), N, +1, 1, 1, d2o.data());
      for (size_t k = 0; k < 3 * nk; ++k) {
        const double* o = vap.data() + k * 4 * PH;
        double* r = vd2.data() + k * (2 * PH + 2);
        long double we = 0.0L, wo = 0.0L;
        for (int h = 0; h < PH; ++h) {
          long double a1 = 0.0L, a2 = 0.0L;
          for (int g2 = 0; g2 < PH; ++g2) {
            a1 += (long double)d2e[(size_t)h * PH + g2] * o[g2];
            a2 += (long double)d2o[(size_t)h * PH + g2] * o[PH + g2];
          }
          r[h] = (double)a1;
          r[PH + h] = (double)a2;
          we += (long double)rw2[h] * o[h];
          wo += (long double)rw2[PH + h] * o[PH + h];
        }
        r[2 * PH] = (double)we;
        r[2 * PH + 1] = (double)wo;
      }
      TRY(upload(p, &p->VD2, vd2.data(), vd2.size()));
      TRY(upload(p, &p->Dm, d->D, nn));
    }
  }
  p->DFT1 = p->D_R = p->D2_R = p->RB_R = p->LI_R = p->PO_R = nullptr;
  p->EH_R = p->EG_R = p->LAMR = nullptr;
  p->YS1T = nullptr;
  p->yst_grid = 0;
  p->SS = nullptr;
  p->TS = nullptr;
  const bool adj = (d->kind == JSD_KIND_PERTURBATION && d->linearise) ||
                   (d->adjoint && (d->kind == JSD_KIND_NS || d->kind == JSD_KIND_PERTURBATION));
  p->adjoint = adj ? 1 : 0;
  if (adj) {
    // the discrete adjoint applies the transposes: row-major copies of the
    // tables (mv_mfma reads its operand transposed) and D^T fragments
    std::vector<double> DT(nn);
    transpose_sq(d->D, DT.data(), N);
    tile_D(DT.data(), N, p->NT, p->KS, tiled);
    TRY(upload(p, &p->DFT1, tiled.data(), tiled.size()));
    TRY(ypar_stream_setup_T(p, DT.data()));
    ypar_frags(DT.data(), N, -1, tiled, p->yp1t);  // D^T is centro-antisymmetric like D
    TRY(upload(p, &p->YF1T, tiled.data(), tiled.size()));
    TRY(upload(p, &p->D_R, d->D, nn));
    TRY(upload(p, &p->D2_R, D2.data(), nn));
    TRY(upload(p, &p->RB_R, d->rhs_base, 3 * nn));
    if (p->EVF) {
      TRY(eigen_tables_full(p, d, N));
    } else {
      TRY(upload(p, &p->LI_R, d->lhs_inv, 3 * nk * nn));
      TRY(upload(p, &p->PO_R, d->poisson, nk * nn));
    }
  }
  p->VA = p->INFL = nullptr;
  if (adj) {
    TRY(upload(p, &p->VA, d->v_a, 3 * nk * N));
    TRY(upload(p, &p->INFL, d->infl_inv, 3 * nk * 4));
  }
  {
    // mean-flow blocks transposed: the solve reads them column-wise (coalesced)
    std::vector<double> rt(3 * 4 * nn), lt(3 * 4 * nn);
    for (int s = 0; s < 3; ++s) {
      transpose_sq(d->rhs_ns + s * 4 * nn, rt.data() + s * 4 * nn, 2 * N);
      transpose_sq(d->lhs_inv_ns + s * 4 * nn, lt.data() + s * 4 * nn, 2 * N);
    }
    TRY(upload(p, &p->RNS, rt.data(), rt.size()));
    TRY(upload(p, &p->LNS, lt.data(), lt.size()));
  }
  TRY(upload(p, &p->kx, d->kx, M));
  TRY(upload(p, &p->kz, d->kz, Z));
  p->kz1 = Z > 1 ? d->kz[1] : 0.0;
  TRY(upload(p, &p->wy, d->y_weights, N));
  if (d->kind != JSD_KIND_NS) {
    TRY(upload(p, &p->Ub, d->base_u, N));
    TRY(upload(p, &p->dUb, d->base_du, N));
  } else {
    p->Ub = p->dUb = nullptr;
  }
  TRY(make_fft(p, d->ppx, &p->fpx, &p->permx, &p->ipermx, &p->twx));
  TRY(make_fft(p, d->ppz, &p->fpz, &p->permz, &p->ipermz, &p->twz));
  {
    const char* env5 = getenv("JSD_T5");
    p->t5 = !p->generic_fft && p->zreg && (plan_is<RZ128>(p->fpz) || plan_is<RZ256>(p->fpz)) &&
            d->kind != JSD_KIND_DUAL && !(env5 && env5[0] == '0');
  }
  // the five-row z pass forms kz of stored index j as kz1 times its signed
  // index: keep it to tables of that form (every table the library builds)
  for (int j = 0; j < Z && p->t5; ++j) {
    const double want = p->kz1 * (double)(j <= Z / 2 ? j : j - Z);
    if (fabs(d->kz[j] - want) > 1e-12 * fabs(p->kz1) * Z) p->t5 = 0;
  }
  const size_t fb = p->fs * sizeof(double2);
  TRY(dalloc(p, (void**)&p->DU, 3 * fb));
  p->HGC = nullptr;
  p->dual_cc = 0;
  if (d->no_coupling_term) {
    if (d->kind == JSD_KIND_PERTURBATION) {
      TRY(dalloc(p, (void**)&p->HGC, fb));
    } else if (d->kind == JSD_KIND_DUAL) {
      p->dual_cc = 1;
    } else {
      jsd_plan_destroy(p);
      JSD_FAIL(JSD_ERR_ARG, "no_coupling_term applies to perturbation and dual plans");
    }
  }
  TRY(dalloc(p, (void**)&p->H, 3 * fb));
  TRY(dalloc(p, (void**)&p->HV, fb));
  TRY(dalloc(p, (void**)&p->HG, fb));
  TRY(dalloc(p, (void**)&p->HVo, fb));
  TRY(dalloc(p, (void**)&p->HGo, fb));
  TRY(dalloc(p, (void**)&p->CONV, 2 * N * sizeof(double2)));
  TRY(dalloc(p, (void**)&p->CONVo, 2 * N * sizeof(double2)));
  const bool nonlin = (d->kind == JSD_KIND_NS) || !d->linearise;
  if (nonlin) {
    const int ntr = (d->kind == JSD_KIND_DUAL) ? 12 : 6;
    const size_t ts = (size_t)d->ppx * N * p->ZT;
    TRY(dalloc(p, (void**)&p->T, ntr * ts * sizeof(double2)));
    TRY(dalloc(p, (void**)&p->V, 2 * ts * sizeof(double2)));
  } else {
    p->T = p->V = nullptr;
  }
  if (d->kind == JSD_KIND_DUAL && !d->linearise) {
    TRY(dalloc(p, (void**)&p->DWF, 3 * fb));
  } else {
    p->DWF = nullptr;
  }
  p->WF = nullptr;
  p->npartial = 1024;
  TRY(dalloc(p, (void**)&p->partial, p->npartial * sizeof(double)));
  TRY(dalloc(p, (void**)&p->eout, sizeof(double)));
#undef TRY
  *out = p;
  return JSD_OK;
}

// ---------------------------------------------------------------------------
static YgemmArgs ygemm_args(const jsd_plan* p, const double* DF, const void* in0, const void* in1, const void* in2,
                            void* out0, void* out1, void* out2, void* conv, long comp_stride) {
  YgemmArgs G;
  G.DF = DF;
  G.NT = p->NT;
  G.KS = p->KS;
  G.in0 = (const double*)in0;
  G.in1 = (const double*)in1;
  G.in2 = (const double*)in2;
  G.out0 = (double*)out0;
  G.out1 = (double*)out1;
  G.out2 = (double*)out2;
  G.conv = (double*)conv;
  G.kx = p->kx;
  G.kz = p->kz;
  G.M = p->M;
  G.N = p->N;
  G.Z = p->Z;
  G.comp_stride = comp_stride;
  return G;
}
static dim3 ygemm_grid(const jsd_plan* p, int ncomp) { return dim3(p->M, (2 * p->Z + 63) / 64, ncomp); }

// D^order along y on the parity blocks: arguments, number of 64-column
// blocks and the dynamic LDS it needs (kept as the maximum in *lds)
static YparArgs ypar_args(const jsd_plan* p, const void* in, void* out, int order, unsigned* nblk, size_t* lds) {
  YparArgs G = order == 2 ? p->yp2 : p->yp1;
  G.F = order == 2 ? p->YF2 : p->YF1;
  G.in = (const double*)in;
  G.out = (double*)out;
  G.M = p->M;
  G.N = p->N;
  G.Z = p->Z;
  G.comp_stride = (long)(2 * p->fs);
  G.acc = 0;
  const long ncol = (long)p->M * 2 * p->Z;
  *nblk = (unsigned)((ncol + 31) / 32);
  *lds = std::max(*lds, (size_t)0);
  return G;
}

static int launch_ygemm_plain(jsd_plan* p, const double2* in, double2* out, int ncomp, int order, hipStream_t st) {
  unsigned nblk = 0;
  size_t lds = 0;
  const YparArgs G = ypar_args(p, in, out, order, &nblk, &lds);
  const int o = order == 2 ? 1 : 0;
  prof_begin(p, 0, st);
  if (p->ys_grid[o] > 0 && (long)ncomp * G.comp_stride < (1L << 31)) {
    YsArgs Y = o ? p->ys2 : p->ys1;
    Y.F = o ? p->YS2 : p->YS1;
    Y.in = G.in;
    Y.out = G.out;
    Y.M = p->M;
    Y.N = p->N;
    Y.Z = p->Z;
    Y.comp_stride = (int)G.comp_stride;
    Y.ncol = ncomp * p->M * 2 * p->Z;
    Y.zero = p->YSZ;
    const int units = (Y.ncol + 15) / 16;
    const int grid = std::min(p->ys_grid[o], (units + YS_THREADS / 64 - 1) / (YS_THREADS / 64));
    const size_t sh = ystream_lds_doubles(Y.KL) * sizeof(double);
    const bool full = Y.nEf >= 64;
    if (Y.anti && full)
      hipLaunchKernelGGL((k_ypar_stream<true, true>), dim3(grid), dim3(YS_THREADS), sh, st, Y);
    else if (Y.anti)
      hipLaunchKernelGGL((k_ypar_stream<true, false>), dim3(grid), dim3(YS_THREADS), sh, st, Y);
    else if (full)
      hipLaunchKernelGGL((k_ypar_stream<false, true>), dim3(grid), dim3(YS_THREADS), sh, st, Y);
    else
      hipLaunchKernelGGL((k_ypar_stream<false, false>), dim3(grid), dim3(YS_THREADS), sh, st, Y);
  } else {
    hipLaunchKernelGGL(k_ypar, dim3(nblk, 1, ncomp), dim3(256), lds, st, G);
  }
  prof_end(p, 0, st);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

static int launch_hv(jsd_plan* p, const double2* H, double2* hv, double2* hg, double2* conv, hipStream_t st) {
  const double* h = (const double*)H;
  prof_begin(p, 4, st);
  hipLaunchKernelGGL(k_ygemm<GEMM_HV>, ygemm_grid(p, 1), dim3(256), 0, st,
                     ygemm_args(p, p->DF1, h, h + 2 * p->fs, h + 4 * p->fs, hv, hg, nullptr, conv, 0L));
  prof_end(p, 4, st);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

static NlLinArgs nl_lin_args(const jsd_plan* p, const double2* U) {
  NlLinArgs A;
  A.U0 = U;
  A.cs = (long)p->fs;
  A.DV = p->DU;
  A.H = p->H;
  A.kx = p->kx;
  A.kz = p->kz;
  A.Ub = p->Ub;
  A.dUb = p->dUb;
  A.M = p->M;
  A.N = p->N;
  A.Z = p->Z;
  A.dual = p->d.kind == JSD_KIND_DUAL ? 1 : 0;
  return A;
}

// H (3 comps, stored convention) from the state U (and the forward state for a nonlinear dual)
// The register x kernels take even transforms as (A, d_x A) and odd ones as
// (A, d_z B); every set built here has that form.
static bool xset_alternates(const XInvSet& set, int ntr) {
  for (int t = 0; t < ntr; ++t) {
    if (set.opA[t] != OP_ID) return false;
    if ((t & 1) == 0 && (set.opB[t] != OP_DX || set.A[t] != set.B[t])) return false;
    if ((t & 1) == 1 && set.opB[t] != OP_DZ) return false;
  }
  return true;
}

// the five-row set of fill_set_u with t5
static bool xset_t5(const XInvSet& set, int ntr) {
  if (ntr != 5) return false;
  const int oA[5] = {OP_ID, OP_ID, OP_DX, OP_DX, OP_ID};
  const int oB[5] = {OP_ID, OP_ZERO, OP_DX, OP_ID, OP_ID};
  for (int t = 0; t < 5; ++t)
    if (set.opA[t] != oA[t] || set.opB[t] != oB[t]) return false;
  return true;
}

static void launch_x_inv_ct(jsd_plan* p, const XInvSet& set, int ntr, int nqb, dim3 grid, size_t lds, hipStream_t st,
                            bool* done) {
  const bool reg = p->xreg && xset_alternates(set, ntr);
  const bool reg5 = p->xreg && xset_t5(set, ntr);
  *done = visit_x(p->fpx, [&](auto c) {
    if (reg5)
      hipLaunchKernelGGL((k_fft_x_inv_r<decltype(c), false, true>), grid, dim3(256), lds, st, set, ntr, nqb,
                         p->fpx.tw, p->kx, p->kz, p->M, p->N, p->Z, p->ZT, p->s_inv);
    else if (reg && p->xreg == 2)
      hipLaunchKernelGGL((k_fft_x_inv_r<decltype(c), true>), grid, dim3(256), lds, st, set, ntr, nqb, p->fpx.tw, p->kx,
                         p->kz, p->M, p->N, p->Z, p->ZT, p->s_inv);
    else if (reg)
      hipLaunchKernelGGL((k_fft_x_inv_r<decltype(c), false>), grid, dim3(256), lds, st, set, ntr, nqb, p->fpx.tw,
                         p->kx, p->kz, p->M, p->N, p->Z, p->ZT, p->s_inv);
    else
      hipLaunchKernelGGL(k_fft_x_inv_t<decltype(c)>, grid, dim3(256), lds, st, set, ntr, nqb, p->fpx.tw, p->kx, p->kz,
                         p->M, p->N, p->Z, p->ZT, p->s_inv);
  });
}

// The packed inverse x transforms of a velocity U (D U in DU) from transform
// t0 on: six rows (u_c, d_x u_c), (D u_c, d_z u_c), or with t5 the five rows
// (u_0, u_1), (u_2, 0), (d_x u_0, d_x u_1), (d_x u_2, D u_0), (D u_1, D u_2),
// whose z pass forms the z derivatives from the z spectra of the first two
// (zrow_src, zrow_post).  Returns the number of rows.
static int fill_set_u(const jsd_plan* p, const double2* U, const double2* DU, XInvSet& set, int t0, bool t5) {
  const size_t fs = p->fs;
  if (t5) {
    const double2* A[5] = {U, U + 2 * fs, U, U + 2 * fs, DU + fs};
    const double2* B[5] = {U + fs, U + 2 * fs, U + fs, DU, DU + 2 * fs};
    const int oA[5] = {OP_ID, OP_ID, OP_DX, OP_DX, OP_ID};
    const int oB[5] = {OP_ID, OP_ZERO, OP_DX, OP_ID, OP_ID};
    for (int t = 0; t < 5; ++t) {
      set.A[t0 + t] = A[t];
      set.B[t0 + t] = B[t];
      set.opA[t0 + t] = oA[t];
      set.opB[t0 + t] = oB[t];
    }
    return 5;
  }
  for (int c = 0; c < 3; ++c) {
    set.A[t0 + 2 * c] = U + c * fs;
    set.opA[t0 + 2 * c] = OP_ID;
    set.B[t0 + 2 * c] = U + c * fs;
    set.opB[t0 + 2 * c] = OP_DX;
    set.A[t0 + 2 * c + 1] = DU + c * fs;
    set.opA[t0 + 2 * c + 1] = OP_ID;
    set.B[t0 + 2 * c + 1] = U + c * fs;
    set.opB[t0 + 2 * c + 1] = OP_DZ;
  }
  return 6;
}

static int nonlinear_H(jsd_plan* p, const double2* U, const double2* UF, hipStream_t st, bool have_du) {
  const jsd_plan_desc& d = p->d;
  const size_t fs = p->fs;
  const int M = p->M, N = p->N, Z = p->Z;
  const bool lin = (d.kind != JSD_KIND_NS) && d.linearise;
  if (p->HGC) {  // coupling_term = False: the h_g addend d_z hat(U' v)
    hipLaunchKernelGGL(k_coupling_corr<false>, dim3((unsigned)((fs + 255) / 256)), dim3(256), 0, st, U + fs, p->dUb,
                       p->kz, p->HGC, M, N, Z);
    HIPCHK(hipGetLastError());
  }
  if (lin) {
    int rc = have_du ? JSD_OK : launch_ygemm_plain(p, U + fs, p->DU, 1, 1, st);
    if (rc) return rc;
    const size_t tot = fs;
    prof_begin(p, 6, st);
    hipLaunchKernelGGL(k_nl_lin, dim3((tot + 255) / 256), dim3(256), 0, st, nl_lin_args(p, U));
    prof_end(p, 6, st);
    HIPCHK(hipGetLastError());
    return JSD_OK;
  }
  int rc = have_du ? JSD_OK : launch_ygemm_plain(p, U, p->DU, 3, 1, st);
  if (rc) return rc;
  const int nqb = (p->KZ + 1 + QB - 1) / QB;
  const int ZT = p->ZT;
  const size_t ts = (size_t)d.ppx * N * ZT;
  const size_t lds_x = (size_t)d.ppx * NSLOT * sizeof(double2);
  int prod = PROD_NS;
  double clin = 1.0;
  int ntr = 6;
  if (d.kind == JSD_KIND_PERTURBATION) prod = PROD_PERT;
  if (d.kind == JSD_KIND_DUAL) {
    prod = PROD_DUAL;
    ntr = 12;
    if (!UF) JSD_FAIL(JSD_ERR_ARG, "nonlinear dual step needs the forward state");
    rc = launch_ygemm_plain(p, UF, p->DWF, 3, 1, st);
    if (rc) return rc;
  }
  XInvSet set;
  // (u_c, d_x u_c) and (d_y u_c, d_z u_c), five rows with t5; the dual's
  // forward field W in six more
  const bool t5 = p->t5 && ntr == 6;
  fill_set_u(p, U, p->DU, set, 0, t5);
  if (ntr == 12) fill_set_u(p, UF, p->DWF, set, 6, false);
  if (t5) ntr = 5;
  for (int t = 0; t < ntr; ++t) set.T[t] = p->T + t * ts;
  const bool ct = !p->generic_fft;
  prof_begin(p, 1, st);
  const dim3 grid_xi(8 * ((ntr * nqb * N + 7) / 8));
  bool xi_ct = false;
  if (ct) launch_x_inv_ct(p, set, ntr, nqb, grid_xi, lds_x, st, &xi_ct);
  if (!xi_ct)
    hipLaunchKernelGGL(k_fft_x_inv, grid_xi, dim3(256), lds_x, st, set, ntr, nqb, p->fpx, p->kx, p->kz, M, N, Z, ZT,
                       p->s_inv);
  prof_end(p, 1, st);
  HIPCHK(hipGetLastError());
  prof_begin(p, 2, st);
  const bool z_256 = ct && p->zreg && plan_is<RZ256>(p->fpz);
  const bool z_128 = ct && p->zreg && plan_is<RZ128>(p->fpz);
  if (z_256 || z_128) {
    const dim3 grid(d.ppx, N), grid2(d.ppx, (N + 1) / 2);
    auto launch = [&](auto pc, auto pz) {
      constexpr int P = decltype(pc)::value;
      constexpr bool PZ = decltype(pz)::value;
      // five-row layout: one y line per 256-point workgroup (six two-wave
      // workgroups per CU measured 0.6 % faster than three of two lines)
      if (P == 256 && prod == PROD_NS && t5)
        hipLaunchKernelGGL((k_fft_z_mid_r<P, 6, PROD_NS, PZ, 1, true>), grid, dim3(z256_threads<6, 1>()), 0, st,
                           p->T, (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
      else if (P == 256 && prod == PROD_PERT && t5)
        hipLaunchKernelGGL((k_fft_z_mid_r<P, 6, PROD_PERT, PZ, 1, true>), grid, dim3(z256_threads<6, 1>()), 0, st,
                           p->T, (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
      else if (prod == PROD_NS && t5)
        hipLaunchKernelGGL((k_fft_z_mid_r<P, 6, PROD_NS, PZ, 2, true>), grid2, dim3(z256_threads<6, 2>()), 0, st,
                           p->T, (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
      else if (prod == PROD_PERT && t5)
        hipLaunchKernelGGL((k_fft_z_mid_r<P, 6, PROD_PERT, PZ, 2, true>), grid2, dim3(z256_threads<6, 2>()), 0, st,
                           p->T, (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
      else if (prod == PROD_NS)
        hipLaunchKernelGGL((k_fft_z_mid_r<P, 6, PROD_NS, PZ, 2>), grid2, dim3(z256_threads<6, 2>()), 0, st, p->T,
                           (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
      else if (prod == PROD_PERT)
        hipLaunchKernelGGL((k_fft_z_mid_r<P, 6, PROD_PERT, PZ, 2>), grid2, dim3(z256_threads<6, 2>()), 0, st, p->T,
                           (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
      else
        hipLaunchKernelGGL((k_fft_z_mid_r<P, 12, PROD_DUAL, PZ>), grid, dim3(z256_threads<12>()), 0, st, p->T,
                           (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
    };
    if (z_128)
      launch(std::integral_constant<int, 128>(), std::false_type());
    else if (Z <= 129)
      launch(std::integral_constant<int, 256>(), std::true_type());
    else
      launch(std::integral_constant<int, 256>(), std::false_type());
  }
  const bool z_ct = z_256 || z_128 || (ct && visit_z(p->fpz, [&](auto c) {
    using C = decltype(c);
    const size_t lds = (size_t)C::P * ntr * sizeof(double2);
    const dim3 grid(d.ppx, N), block(zmid_threads(C::P));
    if (prod == PROD_NS)
      hipLaunchKernelGGL((k_fft_z_mid_t<C, 6, PROD_NS>), grid, block, lds, st, p->T, (long)ts, p->V, (long)ts,
                         p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb);
    else if (prod == PROD_PERT)
      hipLaunchKernelGGL((k_fft_z_mid_t<C, 6, PROD_PERT>), grid, block, lds, st, p->T, (long)ts, p->V, (long)ts,
                         p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb);
    else
      hipLaunchKernelGGL((k_fft_z_mid_t<C, 12, PROD_DUAL>), grid, block, lds, st, p->T, (long)ts, p->V, (long)ts,
                         p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb);
  }));
  if (!z_ct) {
    // y planes per z-pass workgroup: LDS <= 80 KB and <= ZMID_ITEMS items per thread
    int YB = (int)((80 * 1024) / ((size_t)d.ppz * ntr * sizeof(double2)));
    YB = std::max(1, std::min(YB, std::min(4, (ZMID_ITEMS * 256) / d.ppz)));
    const size_t lds_z = (size_t)d.ppz * ntr * YB * sizeof(double2);
    hipLaunchKernelGGL(k_fft_z_mid, dim3(d.ppx, (N + YB - 1) / YB), dim3(256), lds_z, st, p->T, (long)ts, ntr, p->V,
                       (long)ts, p->fpz, N, Z, ZT, prod, clin, p->Ub, p->dUb, YB);
  }
  prof_end(p, 2, st);
  HIPCHK(hipGetLastError());
  prof_begin(p, 3, st);
  const dim3 grid_xf(8 * ((2 * nqb * N + 7) / 8));
  const bool xf_ct = ct && visit_x(p->fpx, [&](auto c) {
    if (p->xreg)
      hipLaunchKernelGGL(k_fft_x_fwd_r<decltype(c)>, grid_xf, dim3(256), lds_x, st, p->V, (long)ts, p->H, (long)fs,
                         p->fpx.tw, M, N, Z, ZT, nqb, p->s_fwd);
    else
      hipLaunchKernelGGL(k_fft_x_fwd_t<decltype(c)>, grid_xf, dim3(256), lds_x, st, p->V, (long)ts, p->H, (long)fs,
                         p->fpx.tw, M, N, Z, ZT, nqb, p->s_fwd);
  });
  if (!xf_ct)
    hipLaunchKernelGGL(k_fft_x_fwd, grid_xf, dim3(256), lds_x, st, p->V, (long)ts, p->H, (long)fs, p->fpx, M, N, Z,
                       ZT, nqb, p->s_fwd);
  prof_end(p, 3, st);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

static const double RK_GAM[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double RK_XI[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
static const double RK_ALP[3] = {29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0};

// Buffers of the old-term carry between substeps: the current substep writes
// HV/HG/CONV and reads HVo/HGo/CONVo (the previous substep's output).
struct Roles {
  double2 *HV, *HG, *CONV, *HVo, *HGo, *CONVo;
};
static Roles roles_of(const jsd_plan* p) { return Roles{p->HV, p->HG, p->CONV, p->HVo, p->HGo, p->CONVo}; }
static Roles roles_swapped(const Roles& r) { return Roles{r.HVo, r.HGo, r.CONVo, r.HV, r.HG, r.CONV}; }

static SolveArgs solve_args(const jsd_plan* p, double2* U, int s, double dPdx, const Roles& r) {
  const size_t nn = (size_t)p->N * p->N;
  const size_t nk = (size_t)p->Ka * p->Kb;
  SolveArgs S;
  S.U = U;
  S.cs = (long)p->fs;
  S.H = p->H;
  S.HV = r.HV;
  S.HG = r.HG;
  S.HVo = r.HVo;
  S.HGo = r.HGo;
  S.CONV = r.CONV;
  S.CONVo = r.CONVo;
  S.DEO = p->DEO;
  S.DOE = p->DOE;
  S.D2E = p->D2E;
  S.D2O = p->D2O;
  S.RBE = p->RBP + (size_t)s * 2 * PFRAG;
  S.RBO = S.RBE + PFRAG;
  S.LIP = p->LIP ? p->LIP + (size_t)s * nk * 2 * PFRAG : nullptr;
  S.POP = p->POP;
  S.EVF = p->EVF;
  S.EIF = p->EIF;
  S.FF = p->FF;
  S.GF = p->GF;
  S.CWE = p->CWE;
  S.RWE = p->RWE;
  S.LAM = p->LAM;
  S.VD2 = p->VD2 ? p->VD2 + (size_t)s * nk * (2 * PH + 2) : nullptr;
  S.Dm = p->Dm;
  S.DUo = nullptr;
  S.du_full = 0;
  S.HGC = p->HGC;
  S.c_impl = p->EVF ? p->c_impl[s] : 0.0;
  S.CW = p->CW;
  S.RW = p->RW;
  S.VAP = p->VAP + (size_t)s * nk * 4 * PH;
  S.VINV = p->VINV + (size_t)s * nk * 2;
  S.RNS = p->RNS + (size_t)s * 4 * nn;
  S.LNS = p->LNS + (size_t)s * 4 * nn;
  S.kx = p->kx;
  S.kz = p->kz;
  S.M = p->M;
  S.N = p->N;
  S.Z = p->Z;
  S.Ka = p->Ka;
  S.Kb = p->Kb;
  S.nE = p->nE;
  S.nO = p->nO;
  const double dt = p->d.dt;
  S.dt_gamma = dt * RK_GAM[s];
  S.dt_xi = dt * RK_XI[s];
  // alpha_s dt (-k^2)/Re enters as diag_alpha * (-k^2) with diag_alpha = alpha_s*dt/Re;
  // the tables were built with dt_tab/Re_tab = dt/Re (the dual flips both signs)
  S.diag_alpha = RK_ALP[s] * fabs(dt) / fabs(p->d.re_tau);
  S.dpdx = dPdx * p->dpdx_scale;
  S.use_old = s > 0;
  S.h1_00 = 0;
  return S;
}

// chain = true: p->DU already holds D U (the previous substep's solve wrote
// it) and this substep's solve writes D of its new state for the next one.
// Dual plans with coupling_term = False: H_1 += hat(U' v_x)
// (navier_stokes_perturbation_dual.py:269-278).
static int dual_coupling_corr(jsd_plan* p, const double2* U, hipStream_t st) {
  if (!p->dual_cc) return JSD_OK;
  hipLaunchKernelGGL(k_coupling_corr<true>, dim3((unsigned)((p->fs + 255) / 256)), dim3(256), 0, st, U, p->dUb,
                     p->kz, p->H + p->fs, p->M, p->N, p->Z);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

// What a substep takes beyond the library's own nonlinear term: externally
// supplied h_v, h_g, conv_ns (the nonlinear_update_fn plug-in), a source term
// (prepared by k_source_prep into p->SRC / p->S00) and enforce_base.
struct StepExtras {
  const double2 *hv, *hg, *conv;  // all set or all NULL
  bool src;
  int enforce_base;
};

static int substep(jsd_plan* p, double2* U, const double2* UF, int s, double dPdx, hipStream_t st,
                   bool have_du = false, bool write_du = false, const StepExtras* x = nullptr) {
  const bool ext = x && x->hv;
  int rc;
  if (!ext) {
    rc = nonlinear_H(p, U, UF, st, have_du);
    if (rc) return rc;
    rc = dual_coupling_corr(p, U, st);
    if (rc) return rc;
  }
  if (x) {
    TermArgs T;
    T.H = p->H;
    T.cs = (long)p->fs;
    T.HGS = p->HGS;
    T.HGC = ext ? nullptr : p->HGC;
    T.EHV = ext ? x->hv : nullptr;
    T.EHG = ext ? x->hg : nullptr;
    T.ECV = ext ? x->conv : nullptr;
    T.SRC = x->src ? p->SRC : nullptr;
    T.S00 = x->src ? p->S00 : nullptr;
    T.kx = p->kx;
    T.kz = p->kz;
    T.M = p->M;
    T.N = p->N;
    T.Z = p->Z;
    T.enforce_base = x->enforce_base;
    hipLaunchKernelGGL(k_step_terms, dim3((unsigned)((p->fs + 255) / 256)), dim3(256), 0, st, T);
    HIPCHK(hipGetLastError());
  }
  const bool s8 = p->EVF && p->solve8;
  const int groups = s8 ? solve_groups8(p->M, p->Z) : solve_groups(p->EVF != nullptr, p->M, p->Z, p->Ka, p->Kb);
  const int nblk = ((groups + 7) / 8) * 8;
  SolveArgs S = solve_args(p, U, s, dPdx, roles_of(p));
  if (x) {
    S.HGC = p->HGS;
    S.h1_00 = 1;
  }
  if (write_du) {
    S.DUo = p->DU;
    S.du_full = (p->d.kind == JSD_KIND_NS) || !p->d.linearise;  // linearised steps need D v only
  }
  prof_begin(p, 5, st);
  const bool s8p = s8 && p->solve8 == 2;
  if (!p->EVF)
    hipLaunchKernelGGL(k_solve, dim3(nblk), dim3(256), 0, st, S);
  else if (s8p && !S.DUo)
    hipLaunchKernelGGL(k_solve_eig8p<0>, dim3(nblk), dim3(512), 0, st, S);
  else if (s8p && S.du_full)
    hipLaunchKernelGGL(k_solve_eig8p<2>, dim3(nblk), dim3(512), 0, st, S);
  else if (s8p)
    hipLaunchKernelGGL(k_solve_eig8p<1>, dim3(nblk), dim3(512), 0, st, S);
  else if (s8 && !S.DUo)
    hipLaunchKernelGGL(k_solve_eig8<0>, dim3(nblk), dim3(256), 0, st, S);
  else if (s8 && S.du_full)
    hipLaunchKernelGGL(k_solve_eig8<2>, dim3(nblk), dim3(256), 0, st, S);
  else if (s8)
    hipLaunchKernelGGL(k_solve_eig8<1>, dim3(nblk), dim3(256), 0, st, S);
  else if (!S.DUo)
    hipLaunchKernelGGL(k_solve_eig<0>, dim3(nblk), dim3(256), 0, st, S);
  else if (S.du_full)
    hipLaunchKernelGGL(k_solve_eig<2>, dim3(nblk), dim3(256), 0, st, S);
  else
    hipLaunchKernelGGL(k_solve_eig<1>, dim3(nblk), dim3(256), 0, st, S);
  prof_end(p, 5, st);
  HIPCHK(hipGetLastError());
  std::swap(p->HV, p->HVo);
  std::swap(p->HG, p->HGo);
  std::swap(p->CONV, p->CONVo);
  return JSD_OK;
}

extern "C" int jsd_step(jsd_plan* p, void* vel_hat, double dPdx, int nsteps, void* stream) {
  if (!p || !vel_hat || nsteps < 0) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (p->d.kind == JSD_KIND_DUAL && !p->d.linearise)
    JSD_FAIL(JSD_ERR_STATE, "a nonlinear dual plan steps with jsd_dual_step");
  hipStream_t st = (hipStream_t)stream;
  // nonlinear steps in eigen form: the solves of substeps 1 and 2 also return
  // D u_c of their new state, so substeps 2 and 3 skip the y-derivative
  // kernel.  The chain stays inside a step, so a trajectory does not depend
  // on how its steps are grouped into calls (checkpointed recomputation
  // reproduces it bit for bit).
  const bool chain = p->EVF != nullptr;  // NS, perturbation and linearised dual plans
  for (int n = 0; n < nsteps; ++n)
    for (int s = 0; s < 3; ++s) {
      int rc = substep(p, (double2*)vel_hat, nullptr, s, dPdx, st, chain && s > 0, chain && s < 2);
      if (rc) return rc;
    }
  return JSD_OK;
}

extern "C" int jsd_dual_step(jsd_plan* p, void* adj_hat, const void* fwd_hat, void* stream) {
  if (!p || !adj_hat) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (p->d.kind != JSD_KIND_DUAL) JSD_FAIL(JSD_ERR_STATE, "not a dual plan");
  hipStream_t st = (hipStream_t)stream;
  for (int s = 0; s < 3; ++s) {
    int rc = substep(p, (double2*)adj_hat, (const double2*)fwd_hat, s, 0.0, st);
    if (rc) return rc;
  }
  return JSD_OK;
}

// Workspaces of the source term, enforce_base and external terms, allocated
// on the first call that needs them and kept by the plan.
static int extras_alloc(jsd_plan* p) {
  if (!p->EVF) JSD_FAIL(JSD_ERR_STATE, "source terms, enforce_base and external terms need the eigen-form solve");
  if (p->HGS) return JSD_OK;
  int rc = dalloc(p, (void**)&p->HGS, p->fs * sizeof(double2));
  if (rc) return rc;
  rc = dalloc(p, (void**)&p->SRC, 2 * (size_t)p->N * p->Z * sizeof(double2));
  if (rc) return rc;
  return dalloc(p, (void**)&p->S00, 2 * (size_t)p->N * sizeof(double2));
}

static int source_prep(jsd_plan* p, const double2* src, hipStream_t st) {
  if (p->M < 3) JSD_FAIL(JSD_ERR_ARG, "a source term needs at least 3 stored kx (the reference reads kx rows 0..2)");
  const int n = p->N * p->Z;
  hipLaunchKernelGGL(k_source_prep, dim3((n + 255) / 256), dim3(256), 0, st, src, (long)p->fs, p->kx, p->kz, p->N,
                     p->Z, p->SRC, p->S00);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

extern "C" int jsd_step_source(jsd_plan* p, void* vel_hat, const void* fwd_hat, double dPdx, const void* source_hat,
                               int enforce_base, void* stream) {
  if (!p || !vel_hat) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  const bool dual = p->d.kind == JSD_KIND_DUAL;
  if (!dual && fwd_hat) JSD_FAIL(JSD_ERR_ARG, "a forward state is taken by dual plans only");
  if (dual && !p->d.linearise && !fwd_hat) JSD_FAIL(JSD_ERR_ARG, "a nonlinear dual step needs the forward state");
  hipStream_t st = (hipStream_t)stream;
  int rc = extras_alloc(p);
  if (rc) return rc;
  if (source_hat) {
    rc = source_prep(p, (const double2*)source_hat, st);
    if (rc) return rc;
  }
  StepExtras x{nullptr, nullptr, nullptr, source_hat != nullptr, enforce_base ? 1 : 0};
  // the same derivative chain as jsd_step / jsd_dual_step
  const bool chain = !dual && p->EVF != nullptr;
  for (int s = 0; s < 3; ++s) {
    rc = substep(p, (double2*)vel_hat, (const double2*)fwd_hat, s, dPdx, st, chain && s > 0, chain && s < 2, &x);
    if (rc) return rc;
  }
  return JSD_OK;
}

extern "C" int jsd_substep_terms(jsd_plan* p, void* vel_hat, int s, double dPdx, const void* h_v, const void* h_g,
                                 const void* conv_ns, const void* source_hat, int enforce_base, void* stream) {
  if (!p || !vel_hat || !h_v || !h_g || !conv_ns || s < 0 || s > 2) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = extras_alloc(p);
  if (rc) return rc;
  if (source_hat) {
    rc = source_prep(p, (const double2*)source_hat, st);
    if (rc) return rc;
  }
  StepExtras x{(const double2*)h_v, (const double2*)h_g, (const double2*)conv_ns, source_hat != nullptr,
               enforce_base ? 1 : 0};
  return substep(p, (double2*)vel_hat, nullptr, s, dPdx, st, false, false, &x);
}

extern "C" int jsd_laplacian(jsd_plan* p, const void* in, void* out, int ncomp, double scale, void* stream) {
  if (!p || !in || !out || ncomp < 1 || in == out) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = launch_ygemm_plain(p, (const double2*)in, (double2*)out, ncomp, 2, st);
  if (rc) return rc;
  const size_t tot = (size_t)ncomp * p->fs;
  hipLaunchKernelGGL(k_laplacian_finish, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, (const double2*)in,
                     (double2*)out, p->kx, p->kz, p->M, p->N, p->Z, tot, scale);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

extern "C" int jsd_dissipation_accumulate(jsd_plan* p, const void* vel_hat, void* acc, double scale, void* stream) {
  if (!p || !vel_hat || !acc) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = launch_ygemm_plain(p, (const double2*)vel_hat, p->DU, 3, 1, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dissipation_acc, dim3((unsigned)((p->fs + 255) / 256)), dim3(256), 0, st,
                     (const double2*)vel_hat, p->DU, (long)p->fs, p->kx, p->kz, p->M, p->N, p->Z, (double2*)acc,
                     scale);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

extern "C" int jsd_volume_integral(jsd_plan* p, const void* f_hat, double* out, void* stream) {
  if (!p || !f_hat || !out) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  hipStream_t st = (hipStream_t)stream;
  const double sl = (p->d.lx / (2 * M_PI)) * (p->d.lz / (2 * M_PI));
  const double px = p->d.px, pz = p->d.pz;
  const double scale = (p->d.lx / px) * (p->d.lz / pz) * sqrt(px * pz) * sl;
  hipLaunchKernelGGL(k_mean_integral, dim3(1), dim3(256), 0, st, (const double2*)f_hat, p->wy, p->N, p->Z, scale,
                     p->eout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, p->eout, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return JSD_OK;
}

__global__ void k_negate(const double* __restrict__ x, double* __restrict__ y, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = -x[i];
}

extern "C" int jsd_nonlinear(jsd_plan* p, const void* vel_hat, const void* fwd_hat, void* h_v, void* h_g,
                             void* conv_ns, void* stream) {
  if (!p || !vel_hat || !h_v || !h_g || !conv_ns) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = nonlinear_H(p, (const double2*)vel_hat, (const double2*)fwd_hat, st, false);
  if (rc) return rc;
  // conv_ns = -H, before the dual coupling correction (which the reference
  // leaves out of conv_ns, navier_stokes_perturbation_dual.py:280-291)
  const size_t n = 6 * p->fs;
  hipLaunchKernelGGL(k_negate, dim3((n + 255) / 256), dim3(256), 0, st, (const double*)p->H, (double*)conv_ns, n);
  HIPCHK(hipGetLastError());
  rc = dual_coupling_corr(p, (const double2*)vel_hat, st);
  if (rc) return rc;
  rc = launch_hv(p, p->H, (double2*)h_v, (double2*)h_g, p->CONV, st);
  if (rc) return rc;
  if (p->HGC) {
    hipLaunchKernelGGL(k_add, dim3((unsigned)((p->fs + 255) / 256)), dim3(256), 0, st, (const double2*)p->HGC,
                       (double2*)h_g, p->fs);
    HIPCHK(hipGetLastError());
  }
  return JSD_OK;
}

extern "C" int jsd_energy(jsd_plan* p, const void* vel_hat, double* out, void* stream) {
  if (!p || !vel_hat || !out) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  hipStream_t st = (hipStream_t)stream;
  const double sl = (p->d.lx / (2 * M_PI)) * (p->d.lz / (2 * M_PI));
  const double scale = 0.25 * sl * sl / ((double)p->d.px * (double)p->d.pz);
  hipLaunchKernelGGL(k_energy_partial, dim3(p->npartial), dim3(256), 0, st, (const double2*)vel_hat, (long)p->fs,
                     p->wy, p->M, p->N, p->Z, p->partial);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, p->partial, p->npartial, scale, p->eout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, p->eout, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return JSD_OK;
}

// ---------------------------------------------------------------------------
// allocation-free energy: the partial sums live in a file-scope device array
// ---------------------------------------------------------------------------
// JSD_SCRATCH_DOUBLES partial sums (8 KB), one per workgroup of
// k_energy_partial: the same grid as jsd_energy, so the two agree bit for
// bit.  Every workgroup writes its entry before k_energy_final reads it
// (stream order), so nothing needs clearing; calls that overlap on one GPU
// share the array and race.
#define JSD_SCRATCH_DOUBLES 1024
#define JSD_MAX_DEVICES 64
static __device__ __align__(256) double scratch[JSD_SCRATCH_DOUBLES];
static double* g_scratch_dev[JSD_MAX_DEVICES];  // device address of scratch, per device

// The device address of scratch on the current device, looked up once per
// device (the host-side name is a stand-in, never passed to a kernel).
static int scratch_on_current_device(double** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= JSD_MAX_DEVICES) JSD_FAIL(JSD_ERR_ARG, "device %d beyond the %d the library tracks", dev,
                                                  JSD_MAX_DEVICES);
  double* ptr = __atomic_load_n(&g_scratch_dev[dev], __ATOMIC_ACQUIRE);
  if (!ptr) {
    void* a = nullptr;
    HIPCHK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    if (!a) JSD_FAIL(JSD_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    ptr = (double*)a;
    __atomic_store_n(&g_scratch_dev[dev], ptr, __ATOMIC_RELEASE);
  }
  *out = ptr;
  return JSD_OK;
}

extern "C" int jsd_energy_noalloc(const void* vel_hat, int nxm, int ny, int nzm, int px, int pz, double lx,
                                  double lz, const double* y_weights, double* out, void* stream) {
  if (!vel_hat || !y_weights || !out) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (nxm < 1 || nzm < 1 || (nxm & 1) == 0 || (nzm & 1) == 0)
    JSD_FAIL(JSD_ERR_ARG, "stored mode counts must be odd (got %d, %d)", nxm, nzm);
  if (ny < 3 || ny > JSD_NYMAX) JSD_FAIL(JSD_ERR_ARG, "ny = %d outside 3..%d", ny, JSD_NYMAX);
  if (px < 1 || pz < 1 || !(lx > 0.0) || !(lz > 0.0)) JSD_FAIL(JSD_ERR_ARG, "invalid physical grid or box");
  const size_t fs = (size_t)nxm * ny * nzm;
  if (3 * fs > ((size_t)1 << 40)) JSD_FAIL(JSD_ERR_ARG, "field of %zu elements beyond the supported size", fs);
  double* part = nullptr;
  int rc = scratch_on_current_device(&part);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int npartial = JSD_SCRATCH_DOUBLES;
  const double sl = (lx / (2 * M_PI)) * (lz / (2 * M_PI));
  const double scale = 0.25 * sl * sl / ((double)px * (double)pz);
  hipLaunchKernelGGL(k_energy_partial, dim3(npartial), dim3(256), 0, st, (const double2*)vel_hat, (long)fs,
                     y_weights, nxm, ny, nzm, part);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, part, npartial, scale, out);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

extern "C" int jsd_ydiff(jsd_plan* p, const void* in, void* out, int ncomp, int order, void* stream) {
  if (!p || !in || !out || ncomp < 1 || (order != 1 && order != 2)) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  return launch_ygemm_plain(p, (const double2*)in, (double2*)out, ncomp, order, (hipStream_t)stream);
}

extern "C" int jsd_profile(jsd_plan* p, int enable, double* ms, long* count) {
  if (!p) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (ms || count) {
    for (int c = 0; c < JSD_NPROF; ++c) {
      double tot = 0.0;
      const size_t used = p->prof_used[c];
      for (size_t k = 0; k + 1 < used; k += 2) {
        HIPCHK(hipEventSynchronize(p->prof_ev[c][k + 1]));
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, p->prof_ev[c][k], p->prof_ev[c][k + 1]));
        tot += t;
      }
      if (ms) ms[c] = tot;
      if (count) count[c] = (long)(used / 2);
    }
  }
  for (int c = 0; c < JSD_NPROF; ++c) p->prof_used[c] = 0;
  p->prof_on = enable & ((1 << JSD_NPROF) - 1);
  return JSD_OK;
}

extern "C" int jsd_set_base_flow(jsd_plan* p, const double* base_u, const double* base_du, void* stream) {
  if (!p || !base_u || !base_du) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (p->d.kind == JSD_KIND_NS || !p->Ub) JSD_FAIL(JSD_ERR_STATE, "the plan has no base flow");
  hipStream_t st = (hipStream_t)stream;
  const size_t b = (size_t)p->N * sizeof(double);
  HIPCHK(hipMemcpyAsync(p->Ub, base_u, b, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(p->dUb, base_du, b, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // the host arrays may go away after return
  return JSD_OK;
}

extern "C" int jsd_set_base_flow_device(jsd_plan* p, const double* base_u, const double* base_du, void* stream) {
  if (!p || !base_u || !base_du) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (p->d.kind == JSD_KIND_NS || !p->Ub) JSD_FAIL(JSD_ERR_STATE, "the plan has no base flow");
  hipStream_t st = (hipStream_t)stream;
  const size_t b = (size_t)p->N * sizeof(double);
  HIPCHK(hipMemcpyAsync(p->Ub, base_u, b, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(p->dUb, base_du, b, hipMemcpyDeviceToDevice, st));
  return JSD_OK;
}

// ---------------------------------------------------------------------------
// discrete adjoint of the linearised perturbation step
// ---------------------------------------------------------------------------
static NlLinTArgs nl_lin_T_args(const jsd_plan* p, double2* UB) {
  NlLinTArgs A;
  A.HB = p->H;
  A.cs = (long)p->fs;
  A.UB = UB;
  A.DVB = p->DU;
  A.kx = p->kx;
  A.kz = p->kz;
  A.Ub = p->Ub;
  A.dUb = p->dUb;
  A.M = p->M;
  A.N = p->N;
  A.Z = p->Z;
  return A;
}

static SolveTArgs solve_T_args(const jsd_plan* p, double2* UB, int s, bool has_carry, const Roles& r) {
  const size_t nn = (size_t)p->N * p->N;
  const size_t nk = (size_t)p->Ka * p->Kb;
  SolveTArgs S;
  S.UB = UB;
  S.cs = (long)p->fs;
  S.HVB = r.HV;
  S.HGB = r.HG;
  S.HVO = r.HVo;
  S.HGO = r.HGo;
  S.CB = r.CONV;
  S.CBO = r.CONVo;
  S.D_R = p->D_R;
  S.D2_R = p->D2_R;
  S.RB_R = p->RB_R + s * nn;
  S.LI_R = p->LI_R ? p->LI_R + (size_t)s * nk * nn : nullptr;
  S.PO_R = p->PO_R;
  S.EH_R = p->EH_R;
  S.EG_R = p->EG_R;
  S.LAMR = p->LAMR;
  S.c_impl = p->EH_R ? p->c_impl[s] : 0.0;
  S.VA = p->VA + (size_t)s * nk * p->N;
  S.INFL = p->INFL + (size_t)s * nk * 4;
  S.RNS = p->RNS + (size_t)s * 4 * nn;
  S.LNS = p->LNS + (size_t)s * 4 * nn;
  S.kx = p->kx;
  S.kz = p->kz;
  S.M = p->M;
  S.N = p->N;
  S.Z = p->Z;
  S.Ka = p->Ka;
  S.Kb = p->Kb;
  const double dt = p->d.dt;
  S.dt_gamma = dt * RK_GAM[s];
  S.dt_xi = dt * RK_XI[s];
  S.diag_alpha = RK_ALP[s] * fabs(dt) / fabs(p->d.re_tau);
  S.use_old = s > 0;
  S.has_carry = has_carry ? 1 : 0;
  return S;
}
static SolveT8Args solve_T8_args(const jsd_plan* p, double2* UB, int s, bool has_carry, const Roles& r) {
  const size_t nn = (size_t)p->N * p->N;
  const size_t nk = (size_t)p->Ka * p->Kb;
  SolveT8Args S;
  S.UB = UB;
  S.cs = (long)p->fs;
  S.HB = p->H;
  S.CVI = r.HV;
  S.CGI = r.HG;
  S.CVO = r.HVo;
  S.CGO = r.HGo;
  S.CB = r.CONV;
  S.CBO = r.CONVo;
  S.TF = p->TF;
  S.LAM = p->LAM;
  S.CWE = p->CWE;
  S.RWE = p->RWE;
  S.VAP = p->VAP + (size_t)s * nk * 4 * PH;
  S.VINV = p->VINV + (size_t)s * nk * 2;
  S.RNS = p->RNS + (size_t)s * 4 * nn;
  S.LNS = p->LNS + (size_t)s * 4 * nn;
  S.kx = p->kx;
  S.kz = p->kz;
  S.M = p->M;
  S.N = p->N;
  S.Z = p->Z;
  S.Kb = p->Kb;
  S.nE = p->nE;
  S.nO = p->nO;
  const double dt = p->d.dt;
  S.c_impl = p->c_impl[s];
  S.diag_alpha = RK_ALP[s] * fabs(dt) / fabs(p->d.re_tau);
  S.dt_gamma = dt * RK_GAM[s];
  S.dt_xi = dt * RK_XI[s];
  S.use_old = s > 0;
  S.has_carry = has_carry ? 1 : 0;
  return S;
}

// Hbar from (hvbar, hgbar, convbar): transposed h_v / h_g assembly
static YgemmArgs hvt_args(const jsd_plan* p, const Roles& r) {
  double* hb = (double*)p->H;
  return ygemm_args(p, p->DFT1, r.HV, r.HG, nullptr, hb, hb + 2 * p->fs, hb + 4 * p->fs, r.CONV, 0L);
}
// vbar += D^T Dvbar
static YgemmArgs acc_args(const jsd_plan* p, double2* UB) {
  return ygemm_args(p, p->DFT1, p->DU, nullptr, nullptr, UB + p->fs, nullptr, nullptr, nullptr, 0L);
}
// the same accumulation v_bar += D^T Dv_bar on the parity blocks of D^T (ypar_body)
static YparArgs acc_ypar_args(const jsd_plan* p, double2* UB, unsigned* nblk) {
  size_t lds = 0;
  YparArgs G = ypar_args(p, p->DU, UB + p->fs, 1, nblk, &lds);
  const YparArgs T = p->yp1t;
  G.F = p->YF1T;
  G.ET = T.ET;
  G.OT = T.OT;
  G.KIE = T.KIE;
  G.KIO = T.KIO;
  G.KE = T.KE;
  G.KO = T.KO;
  G.nEf = T.nEf;
  G.nOf = T.nOf;
  G.anti = T.anti;
  G.acc = 1;
  return G;
}

// The solve transpose of one substep and the h_v / h_g transpose: one
// eigen-form kernel writing Hbar (p->H), or the table form and k_ygemm<GEMM_HVT>.
static int solve_T_stage(jsd_plan* p, double2* UB, int s, bool has_carry, hipStream_t st) {
  const Roles r = roles_of(p);
  if (p->TF && p->solveT8) {
    const int nblk = ((solve_groups8(p->M, p->Z) + 7) / 8) * 8;
    hipLaunchKernelGGL(k_solve_T8p, dim3(nblk), dim3(512), 0, st, solve_T8_args(p, UB, s, has_carry, r));
    HIPCHK(hipGetLastError());
    return JSD_OK;
  }
  const int nblk = ((solve_groups(p->EH_R != nullptr, p->M, p->Z, p->Ka, p->Kb) + 7) / 8) * 8;
  hipLaunchKernelGGL(k_solve_T, dim3(nblk), dim3(256), 0, st, solve_T_args(p, UB, s, has_carry, r));
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_ygemm<GEMM_HVT>, ygemm_grid(p, 1), dim3(256), 0, st, hvt_args(p, r));
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

// out (ncomp components, component stride cstride doubles) += D^T in: the
// accumulating streaming kernel, or k_ygemm<GEMM_ACC>
static int launch_acc_T(jsd_plan* p, const double2* in, double2* out, int ncomp, long cstride, hipStream_t st) {
  if (p->yst_grid > 0 && (long)ncomp * cstride < (1L << 31)) {
    YsArgs Y = p->ys1t;
    Y.F = p->YS1T;
    Y.in = (const double*)in;
    Y.out = (double*)out;
    Y.M = p->M;
    Y.N = p->N;
    Y.Z = p->Z;
    Y.comp_stride = (int)cstride;
    Y.ncol = ncomp * p->M * 2 * p->Z;
    Y.zero = p->YSZ;
    const int units = (Y.ncol + 15) / 16;
    const int grid = std::min(p->yst_grid, (units + YS_THREADS / 64 - 1) / (YS_THREADS / 64));
    const size_t sh = ystream_lds_doubles(Y.KL) * sizeof(double);
    if (Y.nEf >= 64)
      hipLaunchKernelGGL((k_ypar_stream<true, true, true>), dim3(grid), dim3(YS_THREADS), sh, st, Y);
    else
      hipLaunchKernelGGL((k_ypar_stream<true, false, true>), dim3(grid), dim3(YS_THREADS), sh, st, Y);
  } else {
    hipLaunchKernelGGL(k_ygemm<GEMM_ACC>, ygemm_grid(p, ncomp), dim3(256), 0, st,
                       ygemm_args(p, p->DFT1, in, nullptr, nullptr, out, nullptr, nullptr, nullptr, cstride));
  }
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

static int adjoint_substep(jsd_plan* p, double2* UB, int s, bool has_carry, hipStream_t st) {
  int rc0 = solve_T_stage(p, UB, s, has_carry, st);
  if (rc0) return rc0;
  // state adjoint += transposed linearised product term; Dvbar -> DU[0]
  hipLaunchKernelGGL(k_nl_lin_T, dim3((p->fs + 255) / 256), dim3(256), 0, st, nl_lin_T_args(p, UB));
  HIPCHK(hipGetLastError());
  const int rc = launch_acc_T(p, p->DU, UB + p->fs, 1, 0L, st);  // vbar += D^T Dvbar
  if (rc) return rc;
  // carries of this substep's old-term usage feed the previous substep
  std::swap(p->HV, p->HVo);
  std::swap(p->HG, p->HGo);
  std::swap(p->CONV, p->CONVo);
  return JSD_OK;
}

// Forward part of the nonlinear term at state U: D u_c into DU and the six
// packed inverse x transforms into T (the inputs of the z pass).
static int forward_inverse_x(jsd_plan* p, const double2* U, hipStream_t st) {
  const size_t fs = p->fs;
  const int M = p->M, N = p->N, Z = p->Z, ZT = p->ZT;
  int rc = launch_ygemm_plain(p, U, p->DU, 3, 1, st);
  if (rc) return rc;
  const int nqb = (p->KZ + 1 + QB - 1) / QB;
  const size_t ts = (size_t)p->d.ppx * N * ZT;
  const size_t lds_x = (size_t)p->d.ppx * NSLOT * sizeof(double2);
  XInvSet set;
  const int ntr = fill_set_u(p, U, p->DU, set, 0, p->t5 != 0);
  for (int t = 0; t < ntr; ++t) set.T[t] = p->T + t * ts;
  const dim3 grid_xi(8 * ((ntr * nqb * N + 7) / 8));
  bool ok = false;
  launch_x_inv_ct(p, set, ntr, nqb, grid_xi, lds_x, st, &ok);
  if (!ok)
    hipLaunchKernelGGL(k_fft_x_inv, grid_xi, dim3(256), lds_x, st, set, ntr, nqb, p->fpx, p->kx, p->kz, M, N, Z, ZT,
                       p->s_inv);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

// UB += (dH/du at Us)^T Hbar, Hbar in p->H (overwritten); the transpose of the
// nonlinear term of the NS / nonlinear perturbation step.
// Tpre: the inverse x transforms at Us kept from the forward recomputation
// (NULL: transformed here).  The adjoint products overwrite them.
static int nl_adjoint_terms(jsd_plan* p, double2* UB, const double2* Us, double2* Tpre, hipStream_t st) {
  const jsd_plan_desc& d = p->d;
  const size_t fs = p->fs;
  const int M = p->M, N = p->N, Z = p->Z, ZT = p->ZT;
  const int nqb = (p->KZ + 1 + QB - 1) / QB;
  const size_t ts = (size_t)d.ppx * N * ZT;
  const size_t lds_x = (size_t)d.ppx * NSLOT * sizeof(double2);
  // (i) Hbar -> two packed rows of h_bar, into V
  {
    const dim3 grid(8 * ((2 * nqb * N + 7) / 8));
    const bool ok = visit_x(p->fpx, [&](auto c) {
      hipLaunchKernelGGL(k_adj_x_inv_t<decltype(c)>, grid, dim3(256), lds_x, st, p->H, (long)fs, p->V, (long)ts, nqb,
                         p->fpx.tw, M, N, Z, ZT, p->s_fwd);
    });
    if (!ok) JSD_FAIL(JSD_ERR_ARG, "x transform length %d has no adjoint kernel", d.ppx);
    HIPCHK(hipGetLastError());
  }
  // forward fields at the substep's state: T (and DU), unless kept from the recomputation
  double2* const T0 = p->T;
  if (Tpre) {
    p->T = Tpre;
  } else {
    int rc = forward_inverse_x(p, Us, st);
    if (rc) return rc;
  }
  struct Restore {
    jsd_plan* p;
    double2* t;
    ~Restore() { p->T = t; }
  } restore{p, T0};
  // (ii) products adjoint, in place over T
  {
    const int prod = d.kind == JSD_KIND_NS ? PROD_NS : PROD_PERT;
    const double clin = 1.0;
    const dim3 grid(d.ppx, N);
    const bool zr = !p->generic_fft && p->zreg && (plan_is<RZ128>(p->fpz) || plan_is<RZ256>(p->fpz));
    if (zr) {
      // register form, two y lines per workgroup
      const dim3 grid2(d.ppx, (N + 1) / 2);
      auto launch = [&](auto pc, auto pz) {
        constexpr int P = decltype(pc)::value;
        constexpr bool PZ = decltype(pz)::value;
        if (prod == PROD_NS && p->t5)
          hipLaunchKernelGGL((k_adj_z_r<P, PROD_NS, PZ, 2, true>), grid2, dim3(z256_threads<8, 2>()), 0, st, p->T,
                             (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
        else if (p->t5)
          hipLaunchKernelGGL((k_adj_z_r<P, PROD_PERT, PZ, 2, true>), grid2, dim3(z256_threads<8, 2>()), 0, st, p->T,
                             (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
        else if (prod == PROD_NS)
          hipLaunchKernelGGL((k_adj_z_r<P, PROD_NS, PZ, 2>), grid2, dim3(z256_threads<8, 2>()), 0, st, p->T,
                             (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
        else
          hipLaunchKernelGGL((k_adj_z_r<P, PROD_PERT, PZ, 2>), grid2, dim3(z256_threads<8, 2>()), 0, st, p->T,
                             (long)ts, p->V, (long)ts, p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb, p->kz1);
      };
      if (plan_is<RZ128>(p->fpz))
        launch(std::integral_constant<int, 128>(), std::false_type());
      else if (Z <= 129)
        launch(std::integral_constant<int, 256>(), std::true_type());
      else
        launch(std::integral_constant<int, 256>(), std::false_type());
    }
    const bool ok = zr || visit_z(p->fpz, [&](auto c) {
      using C = decltype(c);
      const size_t lds = (size_t)C::P * 8 * sizeof(double2);
      if (prod == PROD_NS)
        hipLaunchKernelGGL((k_adj_z_t<C, PROD_NS>), grid, dim3(256), lds, st, p->T, (long)ts, p->V, (long)ts,
                           p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb);
      else
        hipLaunchKernelGGL((k_adj_z_t<C, PROD_PERT>), grid, dim3(256), lds, st, p->T, (long)ts, p->V, (long)ts,
                           p->fpz.tw, N, Z, ZT, clin, p->Ub, p->dUb);
    });
    if (!ok) JSD_FAIL(JSD_ERR_ARG, "z transform length %d has no adjoint kernel", d.ppz);
    HIPCHK(hipGetLastError());
  }
  // (iii) forward x transforms, Herm and derivative factors: odd rows, then even
  if (p->xreg) {
    const bool ok = visit_x(p->fpx, [&](auto c) {
      hipLaunchKernelGGL(k_adj_x_fwd_r<decltype(c)>, dim3(8 * ((3 * nqb * N + 7) / 8)), dim3(256), lds_x, st, p->T,
                         (long)ts, UB, p->DU, (long)fs, p->fpx.tw, p->kx, p->kz, M, N, Z, ZT, nqb, p->s_inv);
    });
    if (!ok) JSD_FAIL(JSD_ERR_ARG, "x transform length %d has no adjoint kernel", d.ppx);
    HIPCHK(hipGetLastError());
  } else {
    for (int parity = 1; parity >= 0; --parity) {
      const bool ok = visit_x(p->fpx, [&](auto c) {
        hipLaunchKernelGGL(k_adj_x_fwd_t<decltype(c)>, dim3(nqb, N, 3), dim3(256), lds_x, st, p->T, (long)ts, parity,
                           UB, p->DU, p->H, (long)fs, p->fpx.tw, p->kx, p->kz, M, N, Z, ZT, p->s_inv);
      });
      if (!ok) JSD_FAIL(JSD_ERR_ARG, "x transform length %d has no adjoint kernel", d.ppx);
      HIPCHK(hipGetLastError());
    }
  }
  // (iv) ubar_c += D^T Dubar_c
  return launch_acc_T(p, p->DU, UB, 3, (long)(2 * fs), st);
}

static int adjoint_substep_nl(jsd_plan* p, double2* UB, const double2* Us, double2* Tpre, int s, bool has_carry,
                              hipStream_t st) {
  int rc = solve_T_stage(p, UB, s, has_carry, st);
  if (rc) return rc;
  rc = nl_adjoint_terms(p, UB, Us, Tpre, st);
  if (rc) return rc;
  std::swap(p->HV, p->HVo);
  std::swap(p->HG, p->HGo);
  std::swap(p->CONV, p->CONVo);
  return JSD_OK;
}

// the three substep states the nonlinear adjoint recomputes, kept by the plan
static int recompute_alloc(jsd_plan* p) {
  if (p->SS) return JSD_OK;
  int rc = dalloc(p, (void**)&p->SS, 9 * p->fs * sizeof(double2));
  if (rc) return rc;
  // inverse x transforms of substeps 0 and 1 (6 packed transforms each), when the
  // compile-time kernels run (their layout is the one nl_adjoint_terms reads)
  p->TS = nullptr;
  if (p->generic_fft) return JSD_OK;
  const size_t ts = (size_t)p->d.ppx * p->N * p->ZT;
  return dalloc(p, (void**)&p->TS, 2 * 6 * ts * sizeof(double2));
}

extern "C" int jsd_step_vjp_state(jsd_plan* p, const void* state, void* grad, double dPdx, void* stream) {
  if (!p || !state || !grad) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (p->d.kind == JSD_KIND_PERTURBATION && p->d.linearise) return jsd_step_vjp(p, grad, 1, stream);
  if (p->d.kind == JSD_KIND_DUAL || !p->adjoint)
    JSD_FAIL(JSD_ERR_STATE, "the nonlinear adjoint needs an NS or perturbation plan created with adjoint = 1");
  if (p->HGC) JSD_FAIL(JSD_ERR_STATE, "the discrete adjoint has no coupling_term = False transpose");
  hipStream_t st = (hipStream_t)stream;
  const size_t sb = 3 * p->fs * sizeof(double2);
  int rc0 = recompute_alloc(p);
  if (rc0) return rc0;
  // substep states 0, 1, 2 of the step
  double2* S0 = p->SS;
  double2* S1 = p->SS + 3 * p->fs;
  double2* S2 = p->SS + 6 * p->fs;
  HIPCHK(hipMemcpyAsync(S0, state, sb, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(S1, state, sb, hipMemcpyDeviceToDevice, st));
  // the recomputed substeps 0 and 1 leave their inverse x transforms (the
  // forward fields the adjoint of their nonlinear term is linearised about)
  // in TS instead of the plan's T
  const size_t ts = (size_t)p->d.ppx * p->N * p->ZT;
  double2* T0 = p->T;
  double2* Tk[3] = {p->TS, p->TS ? p->TS + 6 * ts : nullptr, nullptr};
  if (Tk[0]) p->T = Tk[0];
  int rc = substep(p, S1, nullptr, 0, dPdx, st);
  p->T = T0;
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(S2, S1, sb, hipMemcpyDeviceToDevice, st));
  if (Tk[1]) p->T = Tk[1];
  rc = substep(p, S2, nullptr, 1, dPdx, st);
  p->T = T0;
  if (rc) return rc;
  const double2* Ss[3] = {S0, S1, S2};
  for (int s = 2; s >= 0; --s) {
    rc = adjoint_substep_nl(p, (double2*)grad, Ss[s], Tk[s], s, s < 2, st);
    if (rc) return rc;
  }
  return JSD_OK;
}

// A step record: the substep states 1 and 2 of one nonlinear step and the
// inverse x transforms of substeps 0 and 1 (when the compile-time transform
// kernels run), written by the forward step and read by its adjoint, which
// then recomputes nothing of the step.  Each substep's slot holds six rows
// (REC_ROWS) whatever the forward layout: the adjoint writes its six adjoint
// rows over the kept transforms.
#define REC_ROWS 6
static size_t record_elems(const jsd_plan* p) {
  const size_t ts = (size_t)p->d.ppx * p->N * p->ZT;
  const size_t rows = p->generic_fft ? 0 : REC_ROWS;
  return 6 * p->fs + 2 * rows * ts;
}

static int record_check(const jsd_plan* p) {
  if (p->d.kind == JSD_KIND_DUAL || (p->d.kind == JSD_KIND_PERTURBATION && p->d.linearise) || !p->adjoint)
    JSD_FAIL(JSD_ERR_STATE, "step records are for NS and nonlinear perturbation plans created with adjoint = 1");
  if (p->HGC) JSD_FAIL(JSD_ERR_STATE, "the discrete adjoint has no coupling_term = False transpose");
  return JSD_OK;
}

extern "C" int jsd_step_record_bytes(jsd_plan* p, size_t* bytes) {
  if (!p || !bytes) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  int rc = record_check(p);
  if (rc) return rc;
  *bytes = record_elems(p) * sizeof(double2);
  return JSD_OK;
}

extern "C" int jsd_step_record(jsd_plan* p, void* vel_hat, void* record, double dPdx, void* stream) {
  if (!p || !vel_hat || !record) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  int rc = record_check(p);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  double2* U = (double2*)vel_hat;
  double2* R = (double2*)record;
  const size_t sb = 3 * p->fs * sizeof(double2);
  const size_t ts = (size_t)p->d.ppx * p->N * p->ZT;
  double2* const T0 = p->T;
  double2* Tk[2] = {nullptr, nullptr};
  if (!p->generic_fft) {
    Tk[0] = R + 6 * p->fs;
    Tk[1] = Tk[0] + REC_ROWS * ts;
  }
  // the substeps of jsd_step (the same D u chain), substeps 0 and 1 writing
  // their inverse x transforms into the record
  const bool chain = p->EVF != nullptr;
  for (int s = 0; s < 3; ++s) {
    if (s < 2 && Tk[s]) p->T = Tk[s];
    rc = substep(p, U, nullptr, s, dPdx, st, chain && s > 0, chain && s < 2);
    p->T = T0;
    if (rc) return rc;
    if (s < 2) HIPCHK(hipMemcpyAsync(R + 3 * s * p->fs, U, sb, hipMemcpyDeviceToDevice, st));
  }
  return JSD_OK;
}

extern "C" int jsd_step_vjp_recorded(jsd_plan* p, const void* state, const void* record, void* grad, double dPdx,
                                     void* stream) {
  if (!p || !state || !record || !grad) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  int rc = record_check(p);
  if (rc) return rc;
  (void)dPdx;  // the adjoint of a step does not depend on its pressure gradient
  hipStream_t st = (hipStream_t)stream;
  const double2* R = (const double2*)record;
  const size_t ts = (size_t)p->d.ppx * p->N * p->ZT;
  double2* Tk[3] = {nullptr, nullptr, nullptr};
  if (!p->generic_fft) {
    Tk[0] = (double2*)R + 6 * p->fs;
    Tk[1] = Tk[0] + REC_ROWS * ts;
  }
  const double2* Ss[3] = {(const double2*)state, R, R + 3 * p->fs};
  for (int s = 2; s >= 0; --s) {
    rc = adjoint_substep_nl(p, (double2*)grad, Ss[s], Tk[s], s, s < 2, st);
    if (rc) return rc;
  }
  return JSD_OK;
}

extern "C" int jsd_step_vjp(jsd_plan* p, void* grad, int nsteps, void* stream) {
  if (!p || !grad || nsteps < 0) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (p->d.kind != JSD_KIND_PERTURBATION || !p->d.linearise || !p->adjoint)
    JSD_FAIL(JSD_ERR_STATE, "the discrete adjoint is built for linearised perturbation plans");
  if (p->HGC) JSD_FAIL(JSD_ERR_STATE, "the discrete adjoint has no coupling_term = False transpose");
  hipStream_t st = (hipStream_t)stream;
  for (int n = 0; n < nsteps; ++n)
    for (int s = 2; s >= 0; --s) {
      int rc = adjoint_substep(p, (double2*)grad, s, s < 2, st);
      if (rc) return rc;
    }
  return JSD_OK;
}

extern "C" int jsd_energy_grad(jsd_plan* p, const void* vel_hat, void* out, double scale, void* stream) {
  if (!p || !vel_hat || !out) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  const double sl = (p->d.lx / (2 * M_PI)) * (p->d.lz / (2 * M_PI));
  const double c = scale * sl * sl / (2.0 * (double)p->d.px * (double)p->d.pz);
  const size_t tot = 3 * p->fs;
  hipLaunchKernelGGL(k_energy_grad, dim3((tot + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)vel_hat, (long)p->fs, (double2*)out, p->wy, p->M, p->N, p->Z, c);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

// ===========================================================================
// batches of independent linearised runs
// ===========================================================================
// Every stage kernel of the linearised step has a body that takes its
// arguments and logical block explicitly; k_batched walks the concatenated
// per-run grids (run r owns logical blocks [first_r, first_{r+1})) with a
// persistent grid of any size.

struct OpYgemmHVT {
  static constexpr int THREADS = 256;
  typedef YgemmArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) { ygemm_body<GEMM_HVT>(a, k); }
};
struct OpYgemmAcc {
  static constexpr int THREADS = 256;
  typedef YgemmArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) { ygemm_body<GEMM_ACC>(a, k); }
};
struct OpYpar {
  static constexpr int THREADS = 256;
  typedef YparArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) {
    __shared__ double sO[YP_TMAX * 16 * 32];
    ypar_body(a, k.x, k.z, sO);
  }
};
struct OpNlLin {
  static constexpr int THREADS = 256;
  typedef NlLinArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) { nl_lin_body(a, k); }
};
struct OpNlLinT {
  static constexpr int THREADS = 256;
  typedef NlLinTArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) { nl_lin_T_body(a, k); }
};
struct OpSolve {
  static constexpr int THREADS = 256;
  typedef SolveArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) {
    if (!a.EVF)
      solve_body(a, k.x);
    else if (!a.DUo)
      solve_body_eig<0>(a, k.x);
    else
      solve_body_eig<1>(a, k.x);  // batches hold linearised runs: D v only
  }
};
// the same solve with eight pencils per logical block (two column tiles per
// wave); eigen form only
struct OpSolve8 {
  static constexpr int THREADS = 256;
  typedef SolveArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) {
    if (!a.DUo)
      solve_body_eig8<0>(a, k.x);
    else
      solve_body_eig8<1>(a, k.x);
  }
};
struct OpSolveT {
  static constexpr int THREADS = 256;
  typedef SolveTArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) { solve_T_body(a, k.x); }
};
// eight pencils on eight waves, one parity per wave (eigen form)
struct OpSolve8p {
  static constexpr int THREADS = 512;
  typedef SolveArgs Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) {
    if (!a.DUo)
      solve_body_eig8p<0>(a, k.x);
    else
      solve_body_eig8p<1>(a, k.x);
  }
};
// the eigen-form transpose of the solve with the h_v / h_g transpose folded in
struct OpSolveT8p {
  static constexpr int THREADS = 512;
  typedef SolveT8Args Args;
  static __device__ __forceinline__ void run(const Args& a, Blk k) { solve_T_body8p(a, k.x); }
};

// runs[r] = (nx, ny, nz, first logical block of run r), ordered by first
template <class Op>
__global__ __launch_bounds__(Op::THREADS) void k_batched(const typename Op::Args* __restrict__ args,
                                                 const int4* __restrict__ runs, int nrun, int total) {
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    int lo = 0, hi = nrun - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (runs[mid].w <= b) lo = mid;
      else hi = mid - 1;
    }
    const int4 R = runs[lo];
    const int lb = b - R.w;
    Op::run(args[lo], Blk{lb % R.x, (lb / R.x) % R.y, lb / (R.x * R.y), R.x});
    __syncthreads();  // the next logical block reuses the LDS
  }
}

enum { BS_YG = 0, BS_NL, BS_SOLVE, BS_SOLVET, BS_HVT, BS_NLT, BS_ACC, BS_NKIND };

struct BatchStage {
  void* args = nullptr;
  int4* runs = nullptr;
  int total = 0;
  size_t lds = 0;  // dynamic LDS bytes of the stage kernel
};

struct jsd_batch {
  int nrun = 0;
  bool vjp = false;
  bool chain = false;  // every plan in eigen form: substeps 2, 3 take D v from the solve
  std::vector<jsd_plan*> plans;
  std::vector<double2*> state, adj;
  std::vector<void*> allocs;
  BatchStage yg, nl, solve[3];          // forward: y-GEMM (D v), linearised term, solve per substep
  BatchStage solve8[3];                 // the solve with eight pencils per logical block (eigen form)
  int tile[BS_NKIND];                   // tile shape per stage: pencils per solve workgroup (4, 8), else 0
  int waves[BS_NKIND];                  // and waves per workgroup (4, 8)
  bool adj8 = false;                    // adjoint: eigen-form transpose with the h_v / h_g transpose folded in
  BatchStage solveT8[3];
  BatchStage solveT[3], hvt[3], nlT, acc;  // adjoint, indexed by call order c (substep 2 - c)
  BatchStage accp;                         // the D^T accumulation on parity blocks (accpar)
  bool accpar = false;
  int cap[BS_NKIND];
  int total[BS_NKIND];
  hipStream_t cs = nullptr;
  hipGraphExec_t xf = nullptr, xb = nullptr;
  hipGraph_t gf = nullptr, gb = nullptr;
  // nonlinear runs: each run steps on its own stream, forked from and joined
  // back to the caller's stream, so the runs' stage kernels overlap
  bool concurrent = false;
  std::vector<hipStream_t> rs;
  std::vector<hipEvent_t> rev;
  hipEvent_t fork = nullptr;
};

static int bdalloc(jsd_batch* b, void** ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) JSD_FAIL(JSD_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  b->allocs.push_back(*ptr);
  return JSD_OK;
}

template <class A>
static int make_stage(jsd_batch* b, BatchStage& S, const std::vector<A>& args, const std::vector<dim3>& grids) {
  std::vector<int4> runs(args.size());
  int first = 0;
  for (size_t r = 0; r < args.size(); ++r) {
    runs[r] = make_int4((int)grids[r].x, (int)grids[r].y, (int)grids[r].z, first);
    first += (int)(grids[r].x * grids[r].y * grids[r].z);
  }
  S.total = first;
  int rc = bdalloc(b, &S.args, args.size() * sizeof(A));
  if (rc) return rc;
  rc = bdalloc(b, (void**)&S.runs, runs.size() * sizeof(int4));
  if (rc) return rc;
  HIPCHK(hipMemcpy(S.args, args.data(), args.size() * sizeof(A), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(S.runs, runs.data(), runs.size() * sizeof(int4), hipMemcpyHostToDevice));
  return JSD_OK;
}

template <class Op>
static int launch_stage(const jsd_batch* b, const BatchStage& S, int cap, hipStream_t st) {
  const int grid = std::max(1, std::min(S.total, cap));
  hipLaunchKernelGGL(k_batched<Op>, dim3(grid), dim3(Op::THREADS), S.lds, st, (const typename Op::Args*)S.args,
                     S.runs, b->nrun, S.total);
  HIPCHK(hipGetLastError());
  return JSD_OK;
}

static int batch_fwd_step(const jsd_batch* b, hipStream_t st) {
  for (int s = 0; s < 3; ++s) {
    // the solves of substeps 1 and 2 return D v of their new state (eigen form)
    int rc = (s > 0 && b->chain) ? JSD_OK : launch_stage<OpYpar>(b, b->yg, b->cap[BS_YG], st);
    if (!rc) rc = launch_stage<OpNlLin>(b, b->nl, b->cap[BS_NL], st);
    if (!rc) {
      if (b->tile[BS_SOLVE] == 8 && b->waves[BS_SOLVE] == 8)
        rc = launch_stage<OpSolve8p>(b, b->solve8[s], b->cap[BS_SOLVE], st);
      else if (b->tile[BS_SOLVE] == 8)
        rc = launch_stage<OpSolve8>(b, b->solve8[s], b->cap[BS_SOLVE], st);
      else
        rc = launch_stage<OpSolve>(b, b->solve[s], b->cap[BS_SOLVE], st);
    }
    if (rc) return rc;
  }
  return JSD_OK;
}

static int batch_bwd_step(const jsd_batch* b, hipStream_t st) {
  for (int c = 0; c < 3; ++c) {
    int rc;
    if (b->adj8) {
      rc = launch_stage<OpSolveT8p>(b, b->solveT8[c], b->cap[BS_SOLVET], st);
    } else {
      rc = launch_stage<OpSolveT>(b, b->solveT[c], b->cap[BS_SOLVET], st);
      if (!rc) rc = launch_stage<OpYgemmHVT>(b, b->hvt[c], b->cap[BS_HVT], st);
    }
    if (!rc) rc = launch_stage<OpNlLinT>(b, b->nlT, b->cap[BS_NLT], st);
    if (!rc) {
      if (b->accpar) rc = launch_stage<OpYpar>(b, b->accp, b->cap[BS_ACC], st);
      else rc = launch_stage<OpYgemmAcc>(b, b->acc, b->cap[BS_ACC], st);
    }
    if (rc) return rc;
  }
  return JSD_OK;
}

// time one stage kernel (all its substep variants) with a candidate grid size
template <class Op>
static float time_stage(const jsd_batch* b, const BatchStage* S, int nS, int cap, hipStream_t st) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 1e30f;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return 1e30f;
  }
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < nS; ++i)
      if (launch_stage<Op>(b, S[i], cap, st)) break;
    (void)hipEventRecord(e1, st);
    float ms = 1e30f;
    if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && rep > 0)
      best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// best grid of one stage kernel: one block per logical block, or persistent
// grids of 1..8 blocks per CU; returns the time of the best
template <class Op>
static float tune_grid(const jsd_batch* b, const BatchStage* S, int nS, hipStream_t st, int* cap) {
  const int cands[5] = {S[0].total, 256 * 8, 256 * 4, 256 * 2, 256};
  int best = S[0].total;
  float tbest = 1e30f;
  for (int c : cands) {
    if (c > S[0].total) continue;
    const float t = time_stage<Op>(b, S, nS, c, st);
    if (t < tbest * 0.97f) {
      tbest = t;
      best = c;
    }
  }
  *cap = best;
  return tbest;
}

template <class Op>
static void tune_stage(jsd_batch* b, int kind, const BatchStage* S, int nS, hipStream_t st) {
  tune_grid<Op>(b, S, nS, st, &b->cap[kind]);
  b->tile[kind] = 0;
}

// the solve: tile shape (4 pencils per workgroup, one 16-column MFMA tile
// per wave; or 8 pencils, two tiles per wave) and grid, timed together
// (JSD_BATCH_TILE=4 or 8 restricts the choice: a diagnostic switch).
static void tune_solve(jsd_batch* b, hipStream_t st) {
  const char* env = getenv("JSD_BATCH_TILE");  // "4", "8" (four waves) or "8p" (eight waves)
  const int only = env ? (strcmp(env, "8p") == 0 ? 88 : atoi(env)) : 0;
  int cap4 = 0, cap8 = 0;
  const float t4 = (only == 8 || only == 88) && b->chain ? 1e30f : tune_grid<OpSolve>(b, b->solve, 3, st, &cap4);
  b->cap[BS_SOLVE] = cap4;
  b->tile[BS_SOLVE] = 4;
  b->waves[BS_SOLVE] = 4;
  if (!b->chain || only == 4) return;  // eight pencils per workgroup: eigen form only
  const float t8 = only == 88 ? 1e30f : tune_grid<OpSolve8>(b, b->solve8, 3, st, &cap8);
  float tbest = t4;
  if (t8 < tbest) {
    tbest = t8;
    b->cap[BS_SOLVE] = cap8;
    b->tile[BS_SOLVE] = 8;
  }
  if (only == 8) return;
  int cap8p = 0;
  const float t8p = tune_grid<OpSolve8p>(b, b->solve8, 3, st, &cap8p);
  if (t8p < tbest) {
    b->cap[BS_SOLVE] = cap8p;
    b->tile[BS_SOLVE] = 8;
    b->waves[BS_SOLVE] = 8;
  }
}

extern "C" int jsd_batch_destroy(jsd_batch* b) {
  if (!b) return JSD_OK;
  for (hipStream_t s : b->rs) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->rev) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->xf) (void)hipGraphExecDestroy(b->xf);
  if (b->xb) (void)hipGraphExecDestroy(b->xb);
  if (b->gf) (void)hipGraphDestroy(b->gf);
  if (b->gb) (void)hipGraphDestroy(b->gb);
  if (b->cs) (void)hipStreamDestroy(b->cs);
  for (void* a : b->allocs) (void)hipFree(a);
  delete b;
  return JSD_OK;
}

static int capture(jsd_batch* b, bool fwd, hipGraph_t* g, hipGraphExec_t* x) {
  HIPCHK(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
  const int rc = fwd ? batch_fwd_step(b, b->cs) : batch_bwd_step(b, b->cs);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cs, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) JSD_FAIL(JSD_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  *g = graph;
  HIPCHK(hipGraphInstantiate(x, graph, nullptr, nullptr, 0));
  return JSD_OK;
}

extern "C" int jsd_batch_create(jsd_plan* const* plans, void* const* states, void* const* adjoints, int nrun,
                                void* stream, jsd_batch** out) {
  if (!plans || !states || nrun <= 0 || !out) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  *out = nullptr;
  hipStream_t st = (hipStream_t)stream;
  jsd_batch* b = new jsd_batch();
  b->nrun = nrun;
  b->vjp = adjoints != nullptr;
  // nonlinear perturbation runs step concurrently on their own streams
  {
    int nl = 0;
    for (int r = 0; r < nrun; ++r)
      if (plans[r] && plans[r]->d.kind == JSD_KIND_PERTURBATION && !plans[r]->d.linearise) ++nl;
    if (nl > 0 && nl < nrun) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_STATE, "a batch holds linearised runs or nonlinear perturbation runs, not both");
    }
    b->concurrent = nl == nrun;
  }
  if (b->concurrent) {
    for (int r = 0; r < nrun; ++r) {
      jsd_plan* p = plans[r];
      if (!states[r] || (b->vjp && !adjoints[r])) {
        jsd_batch_destroy(b);
        JSD_FAIL(JSD_ERR_ARG, "run %d: missing plan or buffer", r);
      }
      for (int q = 0; q < r; ++q)
        if (plans[q] == p || states[q] == states[r] || (b->vjp && adjoints[q] == adjoints[r]) ||
            (b->vjp && (adjoints[q] == states[r] || states[q] == adjoints[r]))) {
          jsd_batch_destroy(b);
          JSD_FAIL(JSD_ERR_ARG, "runs %d and %d share a plan or a buffer: each run needs its own", q, r);
        }
      if (b->vjp && (adjoints[r] == states[r] || !p->adjoint || p->HGC)) {
        jsd_batch_destroy(b);
        JSD_FAIL(JSD_ERR_STATE, "run %d: the adjoint needs its own buffer and a plan created with adjoint = 1 "
                                "(coupling term on)", r);
      }
      if (b->vjp) {
        const int rc = recompute_alloc(p);
        if (rc) {
          jsd_batch_destroy(b);
          return rc;
        }
      }
      b->plans.push_back(p);
      b->state.push_back((double2*)states[r]);
      b->adj.push_back(b->vjp ? (double2*)adjoints[r] : nullptr);
      hipStream_t s = nullptr;
      hipEvent_t e = nullptr;
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        if (s) (void)hipStreamDestroy(s);
        jsd_batch_destroy(b);
        JSD_FAIL(JSD_ERR_HIP, "stream or event creation failed");
      }
      b->rs.push_back(s);
      b->rev.push_back(e);
    }
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
      b->fork = nullptr;
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_HIP, "event creation failed");
    }
    for (int k = 0; k < BS_NKIND; ++k) b->cap[k] = b->total[k] = b->tile[k] = b->waves[k] = 0;
    *out = b;
    return JSD_OK;
  }
  for (int r = 0; r < nrun; ++r) {
    jsd_plan* p = plans[r];
    if (!p || !states[r] || (b->vjp && !adjoints[r])) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_ARG, "run %d: missing plan or buffer", r);
    }
    if (!p->d.linearise || p->d.kind == JSD_KIND_NS || p->HGC || p->dual_cc) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_STATE, "run %d: batches hold linearised perturbation or dual plans (coupling term on) only", r);
    }
    if (b->vjp && (p->d.kind != JSD_KIND_PERTURBATION || !p->adjoint)) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_STATE, "run %d: the adjoint is built for linearised perturbation plans", r);
    }
    // every run writes its own plan's workspaces inside each stage launch:
    // a repeated plan or buffer would race
    for (int q = 0; q < r; ++q)
      if (plans[q] == p || states[q] == states[r] || (b->vjp && adjoints[q] == adjoints[r]) ||
          (b->vjp && (adjoints[q] == states[r] || states[q] == adjoints[r]))) {
        jsd_batch_destroy(b);
        JSD_FAIL(JSD_ERR_ARG, "runs %d and %d share a plan or a buffer: each run needs its own", q, r);
      }
    if (b->vjp && adjoints[r] == states[r]) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_ARG, "run %d: state and adjoint share a buffer", r);
    }
    b->plans.push_back(p);
    b->state.push_back((double2*)states[r]);
    b->adj.push_back(b->vjp ? (double2*)adjoints[r] : nullptr);
  }
  b->chain = true;
  for (int r = 0; r < nrun; ++r) b->chain = b->chain && plans[r]->EVF != nullptr;
  b->adj8 = b->vjp && b->chain;
  for (int r = 0; r < nrun; ++r) b->adj8 = b->adj8 && plans[r]->TF != nullptr && plans[r]->solveT8;
#define BTRY(x)               \
  do {                        \
    int _rc = (x);            \
    if (_rc) {                \
      jsd_batch_destroy(b);   \
      return _rc;             \
    }                         \
  } while (0)
  // forward stages; the carry buffers alternate with the substep (any
  // assignment is valid at step boundaries: substep 0 reads no old terms)
  {
    std::vector<YparArgs> ya;
    std::vector<NlLinArgs> na;
    std::vector<dim3> yg, yp, ng, sg, sg8;
    size_t ylds = 0;
    for (int r = 0; r < nrun; ++r) {
      jsd_plan* p = b->plans[r];
      double2* U = b->state[r];
      unsigned nyb = 0;
      ya.push_back(ypar_args(p, U + p->fs, p->DU, 1, &nyb, &ylds));
      yp.push_back(dim3(nyb));
      yg.push_back(ygemm_grid(p, 1));
      na.push_back(nl_lin_args(p, U));
      ng.push_back(dim3((unsigned)((p->fs + 255) / 256)));
      sg.push_back(dim3(solve_groups(p->EVF != nullptr, p->M, p->Z, p->Ka, p->Kb)));
      sg8.push_back(dim3(solve_groups8(p->M, p->Z)));
    }
    BTRY(make_stage(b, b->yg, ya, yp));
    b->yg.lds = ylds;
    BTRY(make_stage(b, b->nl, na, ng));
    for (int s = 0; s < 3; ++s) {
      std::vector<SolveArgs> sa;
      for (int r = 0; r < nrun; ++r) {
        jsd_plan* p = b->plans[r];
        const Roles r0 = roles_of(p);
        SolveArgs S = solve_args(p, b->state[r], s, 0.0, (s & 1) ? roles_swapped(r0) : r0);
        if (b->chain && s < 2) {
          S.DUo = p->DU;
          S.du_full = 0;
        }
        sa.push_back(S);
      }
      BTRY(make_stage(b, b->solve[s], sa, sg));
      if (b->chain) BTRY(make_stage(b, b->solve8[s], sa, sg8));
    }
    if (b->vjp) {
      std::vector<NlLinTArgs> nta;
      std::vector<YgemmArgs> aa;
      for (int r = 0; r < nrun; ++r) {
        jsd_plan* p = b->plans[r];
        nta.push_back(nl_lin_T_args(p, b->adj[r]));
        aa.push_back(acc_args(p, b->adj[r]));
      }
      BTRY(make_stage(b, b->nlT, nta, ng));
      BTRY(make_stage(b, b->acc, aa, yg));
      b->accpar = !getenv("JSD_BATCH_ACC_GEMM");
      for (int r = 0; r < nrun; ++r) b->accpar = b->accpar && b->plans[r]->YF1T != nullptr;
      if (b->accpar) {
        std::vector<YparArgs> pa;
        std::vector<dim3> pg;
        for (int r = 0; r < nrun; ++r) {
          unsigned nyb = 0;
          pa.push_back(acc_ypar_args(b->plans[r], b->adj[r], &nyb));
          pg.push_back(dim3(nyb));
        }
        BTRY(make_stage(b, b->accp, pa, pg));
      }
      for (int c = 0; c < 3; ++c) {
        const int s = 2 - c;
        std::vector<SolveTArgs> ta;
        std::vector<YgemmArgs> ha;
        for (int r = 0; r < nrun; ++r) {
          jsd_plan* p = b->plans[r];
          const Roles r0 = roles_of(p);
          const Roles rr = (c & 1) ? roles_swapped(r0) : r0;
          ta.push_back(solve_T_args(p, b->adj[r], s, c > 0, rr));
          ha.push_back(hvt_args(p, rr));
        }
        if (b->adj8) {
          std::vector<SolveT8Args> t8;
          for (int r = 0; r < nrun; ++r) {
            jsd_plan* p = b->plans[r];
            const Roles r0 = roles_of(p);
            t8.push_back(solve_T8_args(p, b->adj[r], s, c > 0, (c & 1) ? roles_swapped(r0) : r0));
          }
          BTRY(make_stage(b, b->solveT8[c], t8, sg8));
        } else {
          BTRY(make_stage(b, b->solveT[c], ta, sg));
          BTRY(make_stage(b, b->hvt[c], ha, yg));
        }
      }
    }
  }
  b->total[BS_YG] = b->yg.total;
  b->total[BS_NL] = b->nl.total;
  b->total[BS_SOLVE] = b->solve[0].total;
  b->total[BS_SOLVET] = b->vjp ? (b->adj8 ? b->solveT8[0].total : b->solveT[0].total) : 0;
  b->total[BS_HVT] = b->vjp && !b->adj8 ? b->hvt[0].total : 0;
  b->total[BS_NLT] = b->vjp ? b->nlT.total : 0;
  b->total[BS_ACC] = b->vjp ? (b->accpar ? b->accp.total : b->acc.total) : 0;
  for (int k = 0; k < BS_NKIND; ++k) {
    b->cap[k] = std::max(1, b->total[k]);
    b->tile[k] = b->waves[k] = 0;
  }
  b->tile[BS_SOLVE] = 4;
  b->waves[BS_SOLVE] = 4;
  if (b->vjp) {
    b->tile[BS_SOLVET] = b->adj8 ? 8 : 4;
    b->waves[BS_SOLVET] = b->adj8 ? 8 : 4;
  }
  // autotune the persistent grid of each stage on the real data, then put
  // the states (and adjoints) back as they were
  {
    std::vector<size_t> bytes(nrun);
    size_t tot = 0;
    for (int r = 0; r < nrun; ++r) {
      bytes[r] = 3 * b->plans[r]->fs * sizeof(double2);
      tot += bytes[r];
    }
    void* save = nullptr;
    const size_t nsave = b->vjp ? 2 * tot : tot;
    if (hipMalloc(&save, nsave) != hipSuccess) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_NOMEM, "hipMalloc(%zu) failed for the autotuner", nsave);
    }
    size_t off = 0;
    int rc = JSD_OK;
    for (int r = 0; r < nrun && !rc; ++r) {
      if (hipMemcpyAsync((char*)save + off, b->state[r], bytes[r], hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = JSD_ERR_HIP;
      if (b->vjp && hipMemcpyAsync((char*)save + tot + off, b->adj[r], bytes[r], hipMemcpyDeviceToDevice, st) !=
                        hipSuccess)
        rc = JSD_ERR_HIP;
      off += bytes[r];
    }
    if (!rc) {
      tune_stage<OpYpar>(b, BS_YG, &b->yg, 1, st);
      tune_stage<OpNlLin>(b, BS_NL, &b->nl, 1, st);
      tune_solve(b, st);
      if (b->vjp) {
        if (b->adj8) {
          tune_grid<OpSolveT8p>(b, b->solveT8, 3, st, &b->cap[BS_SOLVET]);
        } else {
          tune_grid<OpSolveT>(b, b->solveT, 3, st, &b->cap[BS_SOLVET]);
          tune_stage<OpYgemmHVT>(b, BS_HVT, b->hvt, 3, st);
        }
        tune_stage<OpNlLinT>(b, BS_NLT, &b->nlT, 1, st);
        if (b->accpar) tune_stage<OpYpar>(b, BS_ACC, &b->accp, 1, st);
        else tune_stage<OpYgemmAcc>(b, BS_ACC, &b->acc, 1, st);
      }
    }
    off = 0;
    for (int r = 0; r < nrun && !rc; ++r) {
      if (hipMemcpyAsync(b->state[r], (char*)save + off, bytes[r], hipMemcpyDeviceToDevice, st) != hipSuccess)
        rc = JSD_ERR_HIP;
      if (b->vjp && hipMemcpyAsync(b->adj[r], (char*)save + tot + off, bytes[r], hipMemcpyDeviceToDevice, st) !=
                        hipSuccess)
        rc = JSD_ERR_HIP;
      off += bytes[r];
    }
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipFree(save);
    if (rc || es != hipSuccess) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_HIP, "autotuner copies failed");
    }
  }
  // one step forward (and adjoint) captured once, replayed per step
  {
    hipError_t e = hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking);
    if (e != hipSuccess) {
      jsd_batch_destroy(b);
      JSD_FAIL(JSD_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    BTRY(capture(b, true, &b->gf, &b->xf));
    if (b->vjp) BTRY(capture(b, false, &b->gb, &b->xb));
  }
#undef BTRY
  *out = b;
  return JSD_OK;
}

// Runs fn(r, run stream) for every run on its own stream, after the work
// already queued on st and before anything queued on st later.
template <class F>
static int batch_concurrent(jsd_batch* b, hipStream_t st, F&& fn) {
  HIPCHK(hipEventRecord(b->fork, st));
  for (int r = 0; r < b->nrun; ++r) {
    HIPCHK(hipStreamWaitEvent(b->rs[r], b->fork, 0));
    const int rc = fn(r, b->rs[r]);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b->rev[r], b->rs[r]));
  }
  for (int r = 0; r < b->nrun; ++r) HIPCHK(hipStreamWaitEvent(st, b->rev[r], 0));
  return JSD_OK;
}

extern "C" int jsd_batch_step(jsd_batch* b, int nsteps, void* stream) {
  if (!b || nsteps < 0) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (b->concurrent)
    return batch_concurrent(b, (hipStream_t)stream, [&](int r, hipStream_t s) {
      return jsd_step(b->plans[r], b->state[r], 0.0, nsteps, s);
    });
  for (int n = 0; n < nsteps; ++n) HIPCHK(hipGraphLaunch(b->xf, (hipStream_t)stream));
  return JSD_OK;
}

extern "C" int jsd_batch_step_vjp(jsd_batch* b, int nsteps, void* stream) {
  if (!b || nsteps < 0) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (!b->vjp) JSD_FAIL(JSD_ERR_STATE, "batch created without adjoint buffers");
  if (b->concurrent) JSD_FAIL(JSD_ERR_STATE, "the adjoint of nonlinear runs takes their forward states: "
                                             "jsd_batch_step_vjp_state");
  for (int n = 0; n < nsteps; ++n) HIPCHK(hipGraphLaunch(b->xb, (hipStream_t)stream));
  return JSD_OK;
}

extern "C" int jsd_batch_step_vjp_state(jsd_batch* b, const void* const* states, void* stream) {
  if (!b || !states) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  if (!b->vjp) JSD_FAIL(JSD_ERR_STATE, "batch created without adjoint buffers");
  if (!b->concurrent) return jsd_batch_step_vjp(b, 1, stream);  // the linearised transpose needs no state
  for (int r = 0; r < b->nrun; ++r)
    if (!states[r]) JSD_FAIL(JSD_ERR_ARG, "run %d: missing forward state", r);
  return batch_concurrent(b, (hipStream_t)stream, [&](int r, hipStream_t s) {
    return jsd_step_vjp_state(b->plans[r], states[r], b->adj[r], 0.0, s);
  });
}

extern "C" int jsd_batch_info(jsd_batch* b, int* grid, int* blocks) {
  if (!b) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  for (int k = 0; k < BS_NKIND; ++k) {
    const int total = (k == BS_SOLVE && b->tile[k] == 8) ? b->solve8[0].total : b->total[k];  // BS_SOLVET: total
                                                                                              // is of its kernel
    if (grid) grid[k] = std::min(b->cap[k], total);
    if (blocks) blocks[k] = total;
  }
  return JSD_OK;
}

extern "C" int jsd_batch_tiles(jsd_batch* b, int* pencils, int* waves) {
  if (!b || !pencils) JSD_FAIL(JSD_ERR_ARG, "invalid argument");
  for (int k = 0; k < BS_NKIND; ++k) {
    pencils[k] = b->tile[k];
    if (waves) waves[k] = b->waves[k];
  }
  return JSD_OK;
}
