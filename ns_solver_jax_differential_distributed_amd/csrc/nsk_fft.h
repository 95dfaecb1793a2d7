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

// One Stockham stage of radix R on NL lines of length N held in LDS.
// Line l lives at buf + ((first + l) % MOD) * NP.  P = product of the radices
// of the earlier stages.  T threads cooperate; every thread must call this.
template <int N, int R, int P, int DIR, int NL, int T, int MOD>
__device__ __forceinline__ void fft_stage(float2* buf, int first, int tid) {
  constexpr int NB = N / R;
  constexpr int TOT = NL * NB;
  constexpr int BPT = (TOT + T - 1) / T;
  constexpr int NP = LineLen<N>::NP;
  static_assert(BPT * R <= 48, "too many points per thread");
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
          float2 w = g_tw[(j * k) * (TW_N / (P * R))];
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
  if constexpr (N == 4096) {
    fft_stage<N, 16, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 16, 16, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 16, 256, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 2048) {
    fft_stage<N, 16, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 16, 16, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 8, 256, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 1024) {
    fft_stage<N, 16, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 16, 16, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 4, 256, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 512) {
    fft_stage<N, 8, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 8, 8, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 8, 64, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 256) {
    fft_stage<N, 16, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 16, 16, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 128) {
    fft_stage<N, 16, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 8, 16, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 64) {
    fft_stage<N, 8, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 8, 8, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 32) {
    fft_stage<N, 8, 1, DIR, NL, T, MOD>(buf, first, tid);
    fft_stage<N, 4, 8, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 16) {
    fft_stage<N, 16, 1, DIR, NL, T, MOD>(buf, first, tid);
  } else if constexpr (N == 8) {
    fft_stage<N, 8, 1, DIR, NL, T, MOD>(buf, first, tid);
  } else {
    static_assert(N == 8, "unsupported N");
  }
}

}  // namespace nsk
