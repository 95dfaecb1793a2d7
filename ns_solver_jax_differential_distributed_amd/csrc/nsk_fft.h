// nsk_fft.h - in-LDS complex FFT for lines of N = 2^m points (8 <= N <= 4096)
// on CDNA4 (gfx950), used by the row and column passes of the spectral
// pressure solve.
//
// Algorithm: Stockham autosort FFT with mixed radices 16/8/4.  Each stage
// reads R points with stride N/R from one LDS line, applies the inter-stage
// twiddles, does an R-point DFT in registers and writes the R results back
// to positions (i-k)*R + k + j*P of the same line.  All loads of a stage
// complete before a workgroup barrier and all stores happen after it, so
// one LDS line per transform suffices.  Lines are padded by one element
// per 16 (pad()) to spread the strided accesses of the first stages over
// the LDS banks.
//
// Twiddles come from a table of exp(-2 pi i m / 4096) computed in double
// precision on the device (g_tw); an N-point transform reads every
// (4096/N)-th entry.
#pragma once
#include <hip/hip_runtime.h>

namespace nsk {

constexpr int TW_N = 4096;
// One table per translation unit (internal linkage); filled by k_init_tw.
namespace {
__device__ float2 g_tw[TW_N];
}

__device__ __forceinline__ int pad(int m) { return m + (m >> 4); }

template <int N>
struct LineLen {
  static constexpr int NP = N + N / 16;
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// cos / sin of 2*pi*i/16, i = 0..7
__device__ __forceinline__ constexpr float w16c(int i) {
  return i == 0 ? 1.0f
       : i == 1 ? 0.923879532511286756f
       : i == 2 ? 0.707106781186547524f
       : i == 3 ? 0.382683432365089772f
       : i == 4 ? 0.0f
       : i == 5 ? -0.382683432365089772f
       : i == 6 ? -0.707106781186547524f
                : -0.923879532511286756f;
}
__device__ __forceinline__ constexpr float w16s(int i) {
  return i == 0 ? 0.0f
       : i == 1 ? 0.382683432365089772f
       : i == 2 ? 0.707106781186547524f
       : i == 3 ? 0.923879532511286756f
       : i == 4 ? 1.0f
       : i == 5 ? 0.923879532511286756f
       : i == 6 ? 0.707106781186547524f
                : 0.382683432365089772f;
}

// v * exp(DIR * 2 pi i q / R) for a compile-time q (after unrolling).
template <int R, int DIR>
__device__ __forceinline__ float2 rot_const(int q, float2 v) {
  const int idx = q * (16 / R);
  if (idx == 0) return v;
  if (idx == 4) return DIR < 0 ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
  const float c = w16c(idx);
  const float s = DIR < 0 ? -w16s(idx) : w16s(idx);
  return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

// In-register R-point DFT, natural order in and out:
// x[q] <- sum_j x[j] exp(DIR * 2 pi i j q / R).  Radix-2 decimation in time.
template <int R, int DIR>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    const float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int j = 0; j < R / 2; ++j) {
      e[j] = x[2 * j];
      o[j] = x[2 * j + 1];
    }
    dft<R / 2, DIR>(e);
    dft<R / 2, DIR>(o);
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      const float2 t = rot_const<R, DIR>(q, o[q]);
      x[q] = cadd(e[q], t);
      x[q + R / 2] = csub(e[q], t);
    }
  }
}

// Radix plan per transform length: up to three Stockham stages.
template <int N> struct Radix;
template <> struct Radix<4096> { static constexpr int S = 3, R0 = 16, R1 = 16, R2 = 16; };
template <> struct Radix<2048> { static constexpr int S = 3, R0 = 16, R1 = 16, R2 = 8; };
template <> struct Radix<1024> { static constexpr int S = 3, R0 = 16, R1 = 16, R2 = 4; };
template <> struct Radix<512> { static constexpr int S = 3, R0 = 8, R1 = 8, R2 = 8; };
template <> struct Radix<256> { static constexpr int S = 2, R0 = 16, R1 = 16, R2 = 1; };
template <> struct Radix<128> { static constexpr int S = 2, R0 = 16, R1 = 8, R2 = 1; };
template <> struct Radix<64> { static constexpr int S = 2, R0 = 8, R1 = 8, R2 = 1; };
template <> struct Radix<32> { static constexpr int S = 2, R0 = 8, R1 = 4, R2 = 1; };
template <> struct Radix<16> { static constexpr int S = 1, R0 = 16, R1 = 1, R2 = 1; };
template <> struct Radix<8> { static constexpr int S = 1, R0 = 8, R1 = 1, R2 = 1; };

// Per-thread twiddles of the second and third stages for a fixed mapping
// of (NL lines, T threads) to butterflies.  Loaded once per kernel when a
// workgroup transforms many line sets with the same mapping; the inverse
// transform uses the conjugates.
template <int N, int NL, int T>
struct TwCache {
  using Rd = Radix<N>;
  static constexpr int P1 = Rd::R0, P2 = Rd::R0 * Rd::R1;
  static constexpr int B1 = Rd::S > 1 ? (NL * (N / Rd::R1) + T - 1) / T : 1;
  static constexpr int B2 = Rd::S > 2 ? (NL * (N / Rd::R2) + T - 1) / T : 1;
  static constexpr int C1 = Rd::S > 1 ? Rd::R1 - 1 : 1;
  static constexpr int C2 = Rd::S > 2 ? Rd::R2 - 1 : 1;
  float2 w1[B1][C1];
  float2 w2[B2][C2];

  __device__ __forceinline__ void load(int tid) {
    if constexpr (Rd::S > 1) {
      constexpr int NB = N / Rd::R1, TOT = NL * NB;
#pragma unroll
      for (int q = 0; q < B1; ++q) {
        const int b = tid + q * T;
        const int k = (b % NB) & (P1 - 1);
#pragma unroll
        for (int j = 1; j <= C1; ++j)
          w1[q][j - 1] = (b < TOT) ? g_tw[(j * k) * (TW_N / (P1 * Rd::R1))] : make_float2(1.f, 0.f);
      }
    }
    if constexpr (Rd::S > 2) {
      constexpr int NB = N / Rd::R2, TOT = NL * NB;
#pragma unroll
      for (int q = 0; q < B2; ++q) {
        const int b = tid + q * T;
        const int k = (b % NB) & (P2 - 1);
#pragma unroll
        for (int j = 1; j <= C2; ++j)
          w2[q][j - 1] = (b < TOT) ? g_tw[(j * k) * (TW_N / (P2 * Rd::R2))] : make_float2(1.f, 0.f);
      }
    }
  }
};

// One Stockham stage of radix R on NL lines of length N held in LDS.
// Line l lives at buf + ((first + l) % MOD) * NP.  P = product of the radices
// of the earlier stages.  T threads cooperate; every thread must call this.
// tw (optional): this thread's twiddles [BPT][R-1] for the stage.
template <int N, int R, int P, int DIR, int NL, int T, int MOD, int BPT_TW = 0, int CT = 1>
__device__ __forceinline__ void fft_stage(float2* buf, int first, int tid,
                                          const float2 (*tw)[CT] = nullptr) {
  constexpr int NB = N / R;
  constexpr int TOT = NL * NB;
  constexpr int BPT = (TOT + T - 1) / T;
  constexpr int NP = LineLen<N>::NP;
  static_assert(BPT * R <= 48, "too many points per thread");
  static_assert(BPT_TW == 0 || (BPT_TW == BPT && CT == R - 1), "twiddle cache shape");
  float2 x[BPT][R];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * T;
    if ((TOT % T == 0) || b < TOT) {
      const int line = b / NB, i = b % NB;
      const float2* L = buf + ((first + line) % MOD) * NP;
#pragma unroll
      for (int j = 0; j < R; ++j) x[q][j] = L[pad(i + j * NB)];
      if constexpr (P > 1) {
        const int k = i & (P - 1);
#pragma unroll
        for (int j = 1; j < R; ++j) {
          float2 w;
          if constexpr (BPT_TW > 0)
            w = tw[q][j - 1];
          else
            w = g_tw[(j * k) * (TW_N / (P * R))];
          if (DIR > 0) w.y = -w.y;
          x[q][j] = cmul(x[q][j], w);
        }
      }
      dft<R, DIR>(x[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * T;
    if ((TOT % T == 0) || b < TOT) {
      const int line = b / NB, i = b % NB;
      const int k = i & (P - 1);
      float2* L = buf + ((first + line) % MOD) * NP;
      const int base = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) L[pad(base + j * P)] = x[q][j];
    }
  }
  __syncthreads();
}

// Full unnormalised transform of NL lines: DIR = -1 forward, +1 inverse.
template <int N, int DIR, int NL, int T, int MOD>
__device__ __forceinline__ void fft_lines(float2* buf, int first, int tid) {
  using Rd = Radix<N>;
  fft_stage<N, Rd::R0, 1, DIR, NL, T, MOD>(buf, first, tid);
  if constexpr (Rd::S > 1) fft_stage<N, Rd::R1, Rd::R0, DIR, NL, T, MOD>(buf, first, tid);
  if constexpr (Rd::S > 2) fft_stage<N, Rd::R2, Rd::R0 * Rd::R1, DIR, NL, T, MOD>(buf, first, tid);
}

// Same with this thread's twiddles from a TwCache built for (NL, T).
template <int N, int DIR, int NL, int T, int MOD>
__device__ __forceinline__ void fft_lines_tc(float2* buf, int first, int tid,
                                             const TwCache<N, NL, T>& tc) {
  using Rd = Radix<N>;
  using TC = TwCache<N, NL, T>;
  fft_stage<N, Rd::R0, 1, DIR, NL, T, MOD>(buf, first, tid);
  if constexpr (Rd::S > 1)
    fft_stage<N, Rd::R1, Rd::R0, DIR, NL, T, MOD, TC::B1, TC::C1>(buf, first, tid, tc.w1);
  if constexpr (Rd::S > 2)
    fft_stage<N, Rd::R2, Rd::R0 * Rd::R1, DIR, NL, T, MOD, TC::B2, TC::C2>(buf, first, tid, tc.w2);
}

}  // namespace nsk
