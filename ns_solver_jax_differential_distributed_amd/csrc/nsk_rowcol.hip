// nsk_rowcol.hip - row and column passes of the Chorin step for one grid
// size.  Compiled once per N with -DNSK_N=<N>; each object carries its own
// copy of the twiddle table (g_tw) and an initialiser for it.
//
// Data flow of one rollout step n (NS_solver.py:742-747 and 671-697):
//
//   row pass  (k_row, RK_STEP_ROLL): spectrum of u_n (row-spectral form)
//       -> inverse row FFT of rows x-1..x+1 -> u_n rows in LDS
//       -> objective partial (u_n . f_n, before the step)
//       -> predictor u*_n = u_n + dt (nu lap u_n + adv(u_n) + f_n)
//       -> forward row FFT -> spectrum of u*_n
//   column pass (k_col, COL_PROJECT): spectrum of u*_n
//       -> column FFT -> (I - s s^T/|k|^2) (the pressure projection of
//          NS_solver.py:691-697 in Fourier space: s = sin(k dx)/dx is the
//          symbol of the central difference, |k|^2 that of the spectral
//          Laplacian of SpectralPoissonSolver, k = 0 untouched)
//       -> inverse column FFT -> spectrum of u_{n+1}
//
// The state between steps never leaves the row-spectral form, so one step
// reads and writes 8 bytes per cell in each pass.
#include "nsk_fft.h"
#include "nsk_internal.h"

#ifndef NSK_N
#error "compile with -DNSK_N=<grid size>"
#endif

#define NSK_CAT2(a, b) a##b
#define NSK_CAT(a, b) NSK_CAT2(a, b)

namespace nsk {
namespace {

__global__ void k_init_tw() {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < TW_N) {
    double s, c;
    sincospi((double)m / (double)(TW_N / 2), &s, &c);
    g_tw[m] = make_float2((float)c, (float)(-s));
  }
}

// The hardware hands workgroup i to XCD i % 8; give each XCD a contiguous
// range of bands/columns so that neighbours share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  return (nb % 8 == 0) ? (bid % 8) * (nb / 8) + bid / 8 : bid;
}

// GaussianForcingField._x_component in float32, same operation order
// (NS_solver.py:621-627); rintf is round-half-to-even like jnp.round.
__device__ __forceinline__ float forcing_x(float A, float B, float t, float x,
                                           const nsk_plan& p) {
  const float two_pi = 6.28318530717958647692f;
  const float mu = p.mu0 + sinf(two_pi * A * t);
  const float cb = cosf(two_pi * B * t);
  const float sig = p.sigma0 + cb * cb;
  const float xm = x - mu;
  const float d = xm - rintf(xm);
  return p.C * expf(-(d * d) / (2.0f * sig * sig));
}

// d f_x / dA and d f_x / dB (round() has zero derivative).
__device__ inline void forcing_dtheta(float A, float B, float t, float x,
                                      const nsk_plan& p, double& dA, double& dB) {
  const double pi2 = 6.283185307179586476925;
  const double a = pi2 * (double)A * (double)t;
  const double bb = pi2 * (double)B * (double)t;
  const double mu = (double)p.mu0 + sin(a);
  const double cb = cos(bb);
  const double sig = (double)p.sigma0 + cb * cb;
  const double xm = (double)x - mu;
  const double d = xm - rint(xm);
  const double f = (double)p.C * exp(-(d * d) / (2.0 * sig * sig));
  dA = f * d / (sig * sig) * (pi2 * (double)t * cos(a));
  dB = f * d * d / (sig * sig * sig) * (-pi2 * (double)t * sin(2.0 * bb));
}

// Deterministic workgroup sum (fixed shuffle tree, then waves in order);
// the result is valid in thread 0.
template <int T>
__device__ __forceinline__ double block_sum(double v, double* red, int tid) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0) {
    for (int w = 0; w < T / 64; ++w) s += red[w];
  }
  __syncthreads();
  return s;
}

// Predictor u* = u + dt (nu lap u + adv(u) + f) at one cell from the packed
// pair (ux + i uy) at the centre c and the neighbours x-1 (l), x+1 (r),
// y-1 (d), y+1 (u).  Stencils of NS_solver.py:243-265, 474-477, 685-689.
template <int SCHEME>
__device__ __forceinline__ float2 predictor(float2 c, float2 l, float2 r, float2 d,
                                            float2 u, float fx, float fy, float nu,
                                            float dt, const nsk_plan& p) {
  const float lapx = (l.x + r.x + u.x + d.x - 4.0f * c.x) * p.inv_dx2;
  const float lapy = (l.y + r.y + u.y + d.y - 4.0f * c.y) * p.inv_dx2;
  float advx, advy;
  if constexpr (SCHEME == NSK_SCHEME_CENTRAL) {
    const float dxux = (r.x - l.x) * p.inv_2dx, dyux = (u.x - d.x) * p.inv_2dx;
    const float dxuy = (r.y - l.y) * p.inv_2dx, dyuy = (u.y - d.y) * p.inv_2dx;
    advx = -(c.x * dxux + c.y * dyux);
    advy = -(c.x * dxuy + c.y * dyuy);
  } else {
    const bool px = c.x >= 0.0f, py = c.y >= 0.0f;
    const float dxux = (px ? c.x - l.x : r.x - c.x) * p.inv_dx;
    const float dyux = (py ? c.x - d.x : u.x - c.x) * p.inv_dx;
    const float dxuy = (px ? c.y - l.y : r.y - c.y) * p.inv_dx;
    const float dyuy = (py ? c.y - d.y : u.y - c.y) * p.inv_dx;
    advx = -(c.x * dxux + c.y * dyux);
    advy = -(c.x * dxuy + c.y * dyuy);
  }
  return make_float2(c.x + dt * (nu * lapx + advx + fx), c.y + dt * (nu * lapy + advy + fy));
}

// Transpose of the predictor's Jacobian at state U applied to m:
//   m + dt (nu lap m + (dA/du)^T m)
// m: packed (mx + i my) at c,l,r,d,u; U: packed state at the same points.
template <int SCHEME>
__device__ __forceinline__ float2 predictor_vjp(float2 c, float2 l, float2 r, float2 d,
                                                float2 u, float2 Uc, float2 Ul, float2 Ur,
                                                float2 Ud, float2 Uu, float nu, float dt,
                                                const nsk_plan& p) {
  const float lapx = (l.x + r.x + u.x + d.x - 4.0f * c.x) * p.inv_dx2;
  const float lapy = (l.y + r.y + u.y + d.y - 4.0f * c.y) * p.inv_dx2;
  float ax, ay;
  if constexpr (SCHEME == NSK_SCHEME_CENTRAL) {
    // (.)_c = G_x(ux m_c) + G_y(uy m_c) - sum_d m_d G_c u_d
    const float h = p.inv_2dx;
    const float gx_mx = (Ur.x * r.x - Ul.x * l.x) * h, gy_mx = (Uu.y * u.x - Ud.y * d.x) * h;
    const float gx_my = (Ur.x * r.y - Ul.x * l.y) * h, gy_my = (Uu.y * u.y - Ud.y * d.y) * h;
    const float dxux = (Ur.x - Ul.x) * h, dxuy = (Ur.y - Ul.y) * h;
    const float dyux = (Uu.x - Ud.x) * h, dyuy = (Uu.y - Ud.y) * h;
    ax = gx_mx + gy_mx - (c.x * dxux + c.y * dxuy);
    ay = gx_my + gy_my - (c.x * dyux + c.y * dyuy);
  } else {
    // sides fixed by the forward signs: s = (u >= 0) selects the backward
    // difference; D^T w (j) = (s_j ? w_j : -w_j) - s_{j+1} w_{j+1} + !s_{j-1} w_{j-1}
    const float h = p.inv_dx;
    const bool sxl = Ul.x >= 0.0f, sxc = Uc.x >= 0.0f, sxr = Ur.x >= 0.0f;
    const bool syd = Ud.y >= 0.0f, syc = Uc.y >= 0.0f, syu = Uu.y >= 0.0f;
    const float Dx_ux = (sxc ? Uc.x - Ul.x : Ur.x - Uc.x) * h;
    const float Dx_uy = (sxc ? Uc.y - Ul.y : Ur.y - Uc.y) * h;
    const float Dy_ux = (syc ? Uc.x - Ud.x : Uu.x - Uc.x) * h;
    const float Dy_uy = (syc ? Uc.y - Ud.y : Uu.y - Uc.y) * h;
    float dtx[2], dty[2];
    const float mc[2] = {c.x, c.y}, ml[2] = {l.x, l.y}, mr[2] = {r.x, r.y};
    const float md[2] = {d.x, d.y}, mu[2] = {u.x, u.y};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const float wc = Uc.x * mc[k], wr = Ur.x * mr[k], wl = Ul.x * ml[k];
      dtx[k] = ((sxc ? wc : -wc) - (sxr ? wr : 0.0f) + (sxl ? 0.0f : wl)) * h;
      const float vc = Uc.y * mc[k], vu = Uu.y * mu[k], vd = Ud.y * md[k];
      dty[k] = ((syc ? vc : -vc) - (syu ? vu : 0.0f) + (syd ? 0.0f : vd)) * h;
    }
    ax = -dtx[0] - dty[0] - (c.x * Dx_ux + c.y * Dx_uy);
    ay = -dtx[1] - dty[1] - (c.x * Dy_ux + c.y * Dy_uy);
  }
  return make_float2(c.x + dt * (nu * lapx + ax), c.y + dt * (nu * lapy + ay));
}

// Split the forward FFT Z of (a + i b) into the half spectra of a and b and
// store them in the packed float4 layout.
template <int N>
__device__ __forceinline__ float4 split_pair(const float2* L, int c) {
  constexpr int HN = N / 2;
  if (c == 0) {
    const float2 z0 = L[pad(0)], zn = L[pad(HN)];
    return make_float4(z0.x, zn.x, z0.y, zn.y);
  }
  const float2 zk = L[pad(c)], zm = L[pad(N - c)];
  return make_float4(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y), 0.5f * (zk.y + zm.y),
                     0.5f * (zm.x - zk.x));
}

// Inverse of split_pair: expand one packed float4 of column c into Z(c) and
// Z(N-c) of (a + i b) (Hermitian symmetry of each real field).
template <int N>
__device__ __forceinline__ void join_pair(float2* L, int c, float4 h) {
  constexpr int HN = N / 2;
  if (c == 0) {
    L[pad(0)] = make_float2(h.x, h.z);
    L[pad(HN)] = make_float2(h.y, h.w);
  } else {
    L[pad(c)] = make_float2(h.x - h.w, h.y + h.z);
    L[pad(N - c)] = make_float2(h.x + h.w, h.z - h.y);
  }
}

// Row-pass helpers: G rows move between global memory and the LDS ring
// through registers, so the loads of the next rows can be issued before the
// current rows are processed.
template <int N, int G, int T>
struct SpecRows {
  static constexpr int HN = N / 2, PF = (G * HN + T - 1) / T;
  float4 v[PF];
  __device__ __forceinline__ void fetch(const float4* src, int rfirst, int tid) {
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int e = tid + q * T;
      if ((G * HN) % T == 0 || e < G * HN) {
        const int l = e / HN, c = e % HN;
        v[q] = src[(size_t)((rfirst + l) & (N - 1)) * HN + c];
      }
    }
  }
  // expand into the row FFT input Z of (a + i b) in ring slots
  template <int RING>
  __device__ __forceinline__ void commit(float2* ring, int sfirst, int tid) const {
    constexpr int NP = LineLen<N>::NP;
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int e = tid + q * T;
      if ((G * HN) % T == 0 || e < G * HN) {
        const int l = e / HN, c = e % HN;
        join_pair<N>(ring + ((sfirst + l) % RING) * NP, c, v[q]);
      }
    }
  }
};

template <int N, int G, int T>
struct RealRows {
  static constexpr int Q = N / 4, PF = (G * Q + T - 1) / T;
  float4 vx[PF], vy[PF];
  __device__ __forceinline__ void fetch(const float* src, bool two, int rfirst, int tid) {
    const size_t NN = (size_t)N * N;
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int e = tid + q * T;
      if ((G * Q) % T == 0 || e < G * Q) {
        const int l = e / Q, y4 = e % Q;
        const size_t off = (size_t)((rfirst + l) & (N - 1)) * N + 4 * y4;
        vx[q] = *reinterpret_cast<const float4*>(src + off);
        vy[q] = two ? *reinterpret_cast<const float4*>(src + NN + off) : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  }
  template <int RING>
  __device__ __forceinline__ void commit(float2* ring, int sfirst, int tid) const {
    constexpr int NP = LineLen<N>::NP;
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int e = tid + q * T;
      if ((G * Q) % T == 0 || e < G * Q) {
        const int l = e / Q, y = 4 * (e % Q);
        float2* L = ring + ((sfirst + l) % RING) * NP;
        L[pad(y)] = make_float2(vx[q].x, vy[q].x);
        L[pad(y + 1)] = make_float2(vx[q].y, vy[q].y);
        L[pad(y + 2)] = make_float2(vx[q].z, vy[q].z);
        L[pad(y + 3)] = make_float2(vx[q].w, vy[q].w);
      }
    }
  }
};

// The row pass.  A workgroup owns a band of G*ITER consecutive rows and
// marches through it G rows at a time.  The LDS ring holds RING = G + 2*HALO
// rows of real pairs (u_x, u_y): rows xb-HALO .. xb+G-1+HALO for the rows
// xb .. xb+G-1 being processed.  Outputs of the predictor go back into the
// slots of rows xb-1 .. xb+G-2, which no later row needs; after they have
// been transformed and stored, the next rows (prefetched into registers at
// the top of the iteration) take the same slots.  Row r sits in slot
// (r - x0 + HALO) % RING.
template <int N, int KIND, int SCHEME>
__global__ void __launch_bounds__(RowCfg<N>::T) k_row(RowArgs a) {
  constexpr int G = RowCfg<N>::G, ITER = RowCfg<N>::ITER, T = RowCfg<N>::T;
  constexpr int NP = LineLen<N>::NP, HN = N / 2;
  constexpr int SRC = (KIND == RK_REAL_TO_SPEC || KIND == RK_STEP_FROM_REAL || KIND == RK_STEP_FORCE)
                          ? SRC_REAL : SRC_SPEC;
  constexpr int OP = (KIND == RK_REAL_TO_SPEC || KIND == RK_SPEC_TO_REAL) ? OP_COPY
                     : (KIND == RK_ADJ ? OP_ADJ : OP_STEP);
  constexpr bool FFIELD = (KIND == RK_STEP_FORCE);
  constexpr bool ROLL = (KIND == RK_STEP_ROLL);
  constexpr int HALO = (OP == OP_COPY) ? 0 : 1;
  constexpr int RING = G + 2 * HALO;
  constexpr bool TWR = NSK_ROW_TWR && ITER > 1;  // twiddles reused across iterations
  constexpr int GD = (OP == OP_ADJ) ? G : 1;
  constexpr int PH = (G == 2) ? 2 : 1;      // stencil phases
  constexpr int GP = G / PH;                // rows per phase
  constexpr int RES = (GP * N + T - 1) / T;
  __shared__ float2 lds[RING * NP];
  __shared__ float s_f[G];
  __shared__ double s_dA[GD], s_dB[GD];
  __shared__ double s_red[T / 64];

  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int nb = gridDim.x;
  const int band = xcd_remap(blockIdx.x, nb);
  nsk_slot sl;
  if (a.slots != nullptr) {
    sl = a.slots[b];
  } else {
    sl.nu = sl.dt = sl.A = sl.B = 0.0f;
    sl.step = sl.nsteps = sl.pad1 = 0;
    sl.jseed = 1.0f;
  }
  const int n = sl.step + a.step_off;
  bool final_mode = false;
  if (a.require == REQ_ROLL) {
    if (n < 0 || n > sl.nsteps) return;
    final_mode = (n == sl.nsteps);
  } else if (a.require == REQ_ACTIVE) {
    if (n < 0 || n >= sl.nsteps) return;
  }
  const nsk_plan& p = a.plan;
  const float dt = sl.dt, nu = sl.nu;
  const float t = (a.t_explicit >= 0.0f) ? a.t_explicit : (float)n * dt;
  const size_t NN = (size_t)N * N;

  const float* srcR = (SRC == SRC_REAL) ? a.src_real + (long long)b * a.src_bstride : nullptr;
  const float4* srcS = (SRC == SRC_SPEC) ? a.src_spec + (size_t)b * N * HN : nullptr;
  float4* dstS = a.dst_spec ? a.dst_spec + (size_t)b * N * HN : nullptr;
  float* dstR = a.dst_real ? a.dst_real + (long long)b * a.dst_bstride : nullptr;
  const bool two = a.src_nc > 1;

  float* st = nullptr;
  if (OP == OP_STEP && a.ustore != nullptr && !final_mode) {
    const int rel = n - a.st_base;
    if (rel >= 0 && rel % a.st_every == 0 && rel / a.st_every < a.st_cap)
      st = a.ustore + (long long)b * a.st_bstride + (long long)(rel / a.st_every) * 2 * NN;
  }
  const float* us = nullptr;
  if (OP == OP_ADJ) us = a.ustate + (long long)b * a.us_bstride + (long long)(n - a.us_base) * 2 * NN;

  const int x0 = band * G * ITER;
  float2* ring = lds;
  double acc0 = 0.0, acc1 = 0.0;
  const double dx2 = p.dx * p.dx;

  TwCache<N, G, T> tc;
  if constexpr (TWR) tc.load(tid);

  // ---- prologue: rows x0-HALO .. x0+G-1+HALO into slots 0 .. RING-1
  if constexpr (SRC == SRC_SPEC) {
    for (int e = tid; e < RING * HN; e += T) {
      const int l = e / HN, c = e % HN;
      join_pair<N>(ring + l * NP, c, srcS[(size_t)((x0 - HALO + l) & (N - 1)) * HN + c]);
    }
    __syncthreads();
    if constexpr (HALO == 0 || G == 2) {
#pragma unroll
      for (int part = 0; part < RING / G; ++part) {
        if constexpr (TWR)
          fft_lines_tc<N, +1, G, T, RING>(ring, part * G, tid, tc);
        else
          fft_lines<N, +1, G, T, RING>(ring, part * G, tid);
      }
    } else {
      fft_lines<N, +1, RING, T, RING>(ring, 0, tid);
    }
  } else {
    const float* s1 = srcR + NN;
    for (int e = tid; e < RING * (N / 4); e += T) {
      const int l = e / (N / 4), y = 4 * (e % (N / 4));
      const size_t off = (size_t)((x0 - HALO + l) & (N - 1)) * N + y;
      const float4 vx = *reinterpret_cast<const float4*>(srcR + off);
      const float4 vy = two ? *reinterpret_cast<const float4*>(s1 + off) : make_float4(0.f, 0.f, 0.f, 0.f);
      float2* L = ring + l * NP;
      L[pad(y)] = make_float2(vx.x, vy.x);
      L[pad(y + 1)] = make_float2(vx.y, vy.y);
      L[pad(y + 2)] = make_float2(vx.z, vy.z);
      L[pad(y + 3)] = make_float2(vx.w, vy.w);
    }
    __syncthreads();
  }

#pragma unroll 1
  for (int it = 0; it < ITER; ++it) {
    const int xb = x0 + it * G;
    const bool has_next = it + 1 < ITER;
    const int s0 = (xb - x0) % RING;           // first slot freed by this iteration
    const int rnext = xb + G + HALO;           // first row loaded for the next one

    // ---- prefetch the next rows into registers
    SpecRows<N, G, T> pfs;
    RealRows<N, G, T> pfr;
    if (has_next) {
      if constexpr (SRC == SRC_SPEC)
        pfs.fetch(srcS, rnext, tid);
      else
        pfr.fetch(srcR, two, rnext, tid);
    }

    if constexpr (OP == OP_COPY) {
      if constexpr (SRC == SRC_SPEC) {
        const bool two_out = a.dst_nc > 1;
        for (int e = tid; e < G * N; e += T) {
          const int g = e / N, y = e % N;
          const float2 v = ring[((s0 + g) % RING) * NP + pad(y)];
          dstR[(size_t)(xb + g) * N + y] = v.x;
          if (two_out) dstR[NN + (size_t)(xb + g) * N + y] = v.y;
        }
      } else {
        if constexpr (TWR)
          fft_lines_tc<N, -1, G, T, RING>(ring, s0, tid, tc);
        else
          fft_lines<N, -1, G, T, RING>(ring, s0, tid);
        for (int e = tid; e < G * HN; e += T) {
          const int g = e / HN, c = e % HN;
          dstS[(size_t)(xb + g) * HN + c] = split_pair<N>(ring + ((s0 + g) % RING) * NP, c);
        }
      }
    } else if (ROLL && final_mode) {
      for (int e = tid; e < G * N; e += T) {
        const int g = e / N, y = e % N;
        const float2 v = ring[((s0 + 1 + g) % RING) * NP + pad(y)];
        dstR[(size_t)(xb + g) * N + y] = v.x;
        dstR[NN + (size_t)(xb + g) * N + y] = v.y;
      }
    } else {
      // ---- per-row forcing (f_x depends on x only, f_y = 0)
      if (tid < G) {
        const int x = xb + tid;
        const float xc = p.Lf * ((float)x / (float)N);
        s_f[tid] = FFIELD ? 0.0f : forcing_x(sl.A, sl.B, t, xc, p);
        if constexpr (OP == OP_ADJ) {
          double dA, dB;
          forcing_dtheta(sl.A, sl.B, t, xc, p, dA, dB);
          s_dA[tid] = dA;
          s_dB[tid] = dB;
        }
      }
      __syncthreads();

      // ---- stencil, PH phases of GP rows: outputs are held in registers
      // until the phase's reads are done, then written into the slots of
      // rows xb-1+g, which later rows no longer read.
      const float* F = FFIELD ? a.force + (long long)b * a.force_bstride : nullptr;
      const double seed = (double)a.jseed * (double)sl.jseed;
      const float scale_ke = (a.objective == NSK_OBJ_MEAN_KE)
                                 ? (float)(seed / ((double)sl.nsteps * (double)NN)) : 0.0f;
      const float jdir = (float)(seed * (double)dt * dx2);
#pragma unroll
      for (int ph = 0; ph < PH; ++ph) {
        float2 res[RES];
#pragma unroll
        for (int q = 0; q < RES; ++q) {
          const int e = tid + q * T;
          if ((GP * N) % T == 0 || e < GP * N) {
            const int g = ph * GP + e / N, y = e % N;
            const int x = xb + g;
            const float2* Lm = ring + ((s0 + g) % RING) * NP;      // row x-1
            const float2* Lc = ring + ((s0 + g + 1) % RING) * NP;  // row x
            const float2* Lp = ring + ((s0 + g + 2) % RING) * NP;  // row x+1
            const int ym = (y - 1) & (N - 1), yp = (y + 1) & (N - 1);
            const float2 c = Lc[pad(y)], l = Lm[pad(y)], r = Lp[pad(y)];
            const float2 d = Lc[pad(ym)], u = Lc[pad(yp)];
            if constexpr (OP == OP_STEP) {
              float fx = s_f[g], fy = 0.0f;
              if constexpr (FFIELD) {
                fx = F[(size_t)x * N + y];
                fy = F[NN + (size_t)x * N + y];
              } else {
                if (a.objective == NSK_OBJ_WORK)
                  acc0 += (double)fx * (double)c.x;
                else
                  acc0 += 0.5 * ((double)c.x * (double)c.x + (double)c.y * (double)c.y);
              }
              if (st) {
                st[(size_t)x * N + y] = c.x;
                st[NN + (size_t)x * N + y] = c.y;
              }
              res[q] = predictor<SCHEME>(c, l, r, d, u, fx, fy, nu, dt, p);
            } else {
              const int xm = (x - 1) & (N - 1), xp = (x + 1) & (N - 1);
              const float* U0 = us;
              const float* U1 = us + NN;
              const float2 Uc = make_float2(U0[(size_t)x * N + y], U1[(size_t)x * N + y]);
              const float2 Ul = make_float2(U0[(size_t)xm * N + y], U1[(size_t)xm * N + y]);
              const float2 Ur = make_float2(U0[(size_t)xp * N + y], U1[(size_t)xp * N + y]);
              const float2 Ud = make_float2(U0[(size_t)x * N + ym], U1[(size_t)x * N + ym]);
              const float2 Uu = make_float2(U0[(size_t)x * N + yp], U1[(size_t)x * N + yp]);
              float2 v = predictor_vjp<SCHEME>(c, l, r, d, u, Uc, Ul, Ur, Ud, Uu, nu, dt, p);
              double grow;
              if (a.objective == NSK_OBJ_WORK) {
                v.x += jdir * s_f[g];
                grow = seed * (double)dt * dx2 * (double)Uc.x + (double)dt * (double)c.x;
              } else {
                v.x += scale_ke * Uc.x;
                v.y += scale_ke * Uc.y;
                grow = (double)dt * (double)c.x;
              }
              acc0 += grow * s_dA[g];
              acc1 += grow * s_dB[g];
              res[q] = v;
            }
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RES; ++q) {
          const int e = tid + q * T;
          if ((GP * N) % T == 0 || e < GP * N) {
            const int g = ph * GP + e / N, y = e % N;
            ring[((s0 + g) % RING) * NP + pad(y)] = res[q];
          }
        }
        __syncthreads();
      }
      if constexpr (TWR)
        fft_lines_tc<N, -1, G, T, RING>(ring, s0, tid, tc);
      else
        fft_lines<N, -1, G, T, RING>(ring, s0, tid);
      for (int e = tid; e < G * HN; e += T) {
        const int g = e / HN, c = e % HN;
        dstS[(size_t)(xb + g) * HN + c] = split_pair<N>(ring + ((s0 + g) % RING) * NP, c);
      }
    }
    __syncthreads();

    // ---- next rows into the freed slots
    if (has_next) {
      if constexpr (SRC == SRC_SPEC)
        pfs.template commit<RING>(ring, s0, tid);
      else
        pfr.template commit<RING>(ring, s0, tid);
      __syncthreads();
      if constexpr (SRC == SRC_SPEC) {
        if constexpr (TWR)
          fft_lines_tc<N, +1, G, T, RING>(ring, s0, tid, tc);
        else
          fft_lines<N, +1, G, T, RING>(ring, s0, tid);
      }
    }
  }

  if constexpr (OP == OP_STEP && !FFIELD) {
    if (a.part != nullptr && !final_mode) {
      const double s = block_sum<T>(acc0, s_red, tid);
      if (tid == 0 && n < a.pcap) {
        const double scale = (a.objective == NSK_OBJ_WORK)
                                 ? (double)dt * dx2
                                 : 1.0 / ((double)sl.nsteps * (double)NN);
        a.part[((size_t)b * a.pcap + n) * nb + band] = s * scale;
      }
    }
  }
  if constexpr (OP == OP_ADJ) {
    if (a.part != nullptr) {
      const double sA = block_sum<T>(acc0, s_red, tid);
      const double sB = block_sum<T>(acc1, s_red, tid);
      if (tid == 0 && n < a.pcap) {
        double* q = a.part + (((size_t)b * a.pcap + n) * nb + band) * 2;
        q[0] = sA;
        q[1] = sB;
      }
    }
  }
}

// Spectral multiplier at mode (kx, ky) applied to the pair (X, Y).
template <int N, int MODE>
__device__ __forceinline__ void apply_mode(float2& X, float2& Y, int kx, int ky,
                                           const nsk_plan& p) {
  constexpr int HN = N / 2;
  const int kxs = kx < HN ? kx : kx - N;
  const int kys = ky < HN ? ky : ky - N;
  const bool zero = (kxs == 0 && kys == 0);
  const float k2 = (float)(kxs * kxs + kys * kys) * p.kscale;
  if constexpr (MODE == COL_PROJECT) {
    if (zero) return;
    const float sx = -g_tw[kx * (TW_N / N)].y * p.inv_dx;  // sin(2 pi kx / N) / dx
    const float sy = -g_tw[ky * (TW_N / N)].y * p.inv_dx;
    const float inv = 1.0f / k2;
    const float2 dot = make_float2(sx * X.x + sy * Y.x, sx * X.y + sy * Y.y);
    const float ax = sx * inv, ay = sy * inv;
    X = make_float2(X.x - ax * dot.x, X.y - ax * dot.y);
    Y = make_float2(Y.x - ay * dot.x, Y.y - ay * dot.y);
  } else if constexpr (MODE == COL_POISSON) {
    const float m = zero ? 0.0f : -1.0f / k2;
    X = make_float2(X.x * m, X.y * m);
    Y = make_float2(Y.x * m, Y.y * m);
  }
}

// The column pass.  Each workgroup handles blocks of W spectral columns
// (both fields: 2W lines of N points), blocks NWG apart, and prefetches the
// next block into registers while it transforms the current one.
#if NSK_COL_WPE > 0
#define NSK_COL_ATTR __attribute__((amdgpu_waves_per_eu(NSK_COL_WPE)))
#else
#define NSK_COL_ATTR
#endif
template <int N, int MODE>
__global__ void __launch_bounds__(ColCfg<N>::T) NSK_COL_ATTR k_col(ColArgs a) {
  constexpr int W = ColCfg<N>::W, T = ColCfg<N>::T, NL = 2 * W;
  constexpr int NP = LineLen<N>::NP, HN = N / 2, NBLK = HN / W;
  constexpr int PF = (N * W + T - 1) / T;
  __shared__ float2 lds[NL * NP];
  const int tid = threadIdx.x, b = blockIdx.y;
  const int nwg = gridDim.x;
  const int w0 = xcd_remap(blockIdx.x, nwg);
  if (a.require == REQ_ACTIVE) {
    const nsk_slot sl = a.slots[b];
    const int n = sl.step + a.step_off;
    if (n < 0 || n >= sl.nsteps) return;
  }
  float4* H = a.spec + (size_t)b * N * HN;
  const nsk_plan& p = a.plan;
  const float scale = 1.0f / ((float)N * (float)N);

  float4 pf[PF];
  auto fetch = [&](int kb) {
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int e = tid + q * T;
      if ((N * W) % T == 0 || e < N * W) {
        const int x = e / W, kw = e % W;
        pf[q] = H[(size_t)x * HN + kb * W + kw];
      }
    }
  };
  // one block per workgroup: load straight into LDS; otherwise stage the
  // next block in registers while the current one is transformed
  constexpr bool PERSIST = ColCfg<N>::NWG < NBLK;
  if (PERSIST && w0 < NBLK) fetch(w0);
#pragma unroll 1
  for (int kb = w0; kb < NBLK; kb += nwg) {
    if constexpr (PERSIST) {
#pragma unroll
      for (int q = 0; q < PF; ++q) {
        const int e = tid + q * T;
        if ((N * W) % T == 0 || e < N * W) {
          const int x = e / W, kw = e % W;
          lds[(2 * kw) * NP + pad(x)] = make_float2(pf[q].x, pf[q].y);
          lds[(2 * kw + 1) * NP + pad(x)] = make_float2(pf[q].z, pf[q].w);
        }
      }
    } else {
      for (int e = tid; e < N * W; e += T) {
        const int x = e / W, kw = e % W;
        const float4 h = H[(size_t)x * HN + kb * W + kw];
        lds[(2 * kw) * NP + pad(x)] = make_float2(h.x, h.y);
        lds[(2 * kw + 1) * NP + pad(x)] = make_float2(h.z, h.w);
      }
    }
    __syncthreads();
    if (PERSIST && kb + nwg < NBLK) fetch(kb + nwg);
    fft_lines<N, -1, NL, T, NL>(lds, 0, tid);
    for (int e = tid; e < W * N; e += T) {
      const int kw = e / N, kx = e % N;
      const int ky = kb * W + kw;
      float2* LX = lds + (2 * kw) * NP;
      float2* LY = LX + NP;
      if (ky == 0) {
        // column 0 carries ky = 0 (real part of the row data) and ky = N/2
        // (imaginary part); separate them, treat each, recombine.
        if (kx > HN) continue;
        const int kx2 = (N - kx) & (N - 1);
        const float2 x1 = LX[pad(kx)], x2 = LX[pad(kx2)];
        const float2 y1 = LY[pad(kx)], y2 = LY[pad(kx2)];
        float2 AX = make_float2(0.5f * (x1.x + x2.x), 0.5f * (x1.y - x2.y));
        float2 BX = make_float2(0.5f * (x1.y + x2.y), 0.5f * (x2.x - x1.x));
        float2 AY = make_float2(0.5f * (y1.x + y2.x), 0.5f * (y1.y - y2.y));
        float2 BY = make_float2(0.5f * (y1.y + y2.y), 0.5f * (y2.x - y1.x));
        apply_mode<N, MODE>(AX, AY, kx, 0, p);
        apply_mode<N, MODE>(BX, BY, kx, HN, p);
        LX[pad(kx)] = make_float2((AX.x - BX.y) * scale, (AX.y + BX.x) * scale);
        LY[pad(kx)] = make_float2((AY.x - BY.y) * scale, (AY.y + BY.x) * scale);
        if (kx2 != kx) {
          LX[pad(kx2)] = make_float2((AX.x + BX.y) * scale, (BX.x - AX.y) * scale);
          LY[pad(kx2)] = make_float2((AY.x + BY.y) * scale, (BY.x - AY.y) * scale);
        }
      } else {
        float2 X = LX[pad(kx)], Y = LY[pad(kx)];
        apply_mode<N, MODE>(X, Y, kx, ky, p);
        LX[pad(kx)] = make_float2(X.x * scale, X.y * scale);
        LY[pad(kx)] = make_float2(Y.x * scale, Y.y * scale);
      }
    }
    __syncthreads();
    fft_lines<N, +1, NL, T, NL>(lds, 0, tid);
    for (int e = tid; e < N * W; e += T) {
      const int x = e / W, kw = e % W;
      const float2 X = lds[(2 * kw) * NP + pad(x)], Y = lds[(2 * kw + 1) * NP + pad(x)];
      H[(size_t)x * HN + kb * W + kw] = make_float4(X.x, X.y, Y.x, Y.y);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// launchers for this N
// ---------------------------------------------------------------------------
constexpr int NN_ = NSK_N;

}  // namespace

namespace {
template <int KIND, int SCHEME>
hipError_t go_row(const RowArgs& a, int batch, hipStream_t s) {
  const int bands = NN_ / (RowCfg<NN_>::G * RowCfg<NN_>::ITER);
  hipLaunchKernelGGL((k_row<NN_, KIND, SCHEME>), dim3(bands, batch), dim3(RowCfg<NN_>::T), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t NSK_CAT(launch_row_, NSK_N)(int kind, int scheme, const RowArgs& a, int batch,
                                       hipStream_t s) {
  const bool up = scheme == NSK_SCHEME_UPWIND;
  switch (kind) {
    case RK_REAL_TO_SPEC: return go_row<RK_REAL_TO_SPEC, 0>(a, batch, s);
    case RK_SPEC_TO_REAL: return go_row<RK_SPEC_TO_REAL, 0>(a, batch, s);
    case RK_STEP_FROM_REAL:
      return up ? go_row<RK_STEP_FROM_REAL, 1>(a, batch, s) : go_row<RK_STEP_FROM_REAL, 0>(a, batch, s);
    case RK_STEP_FORCE:
      return up ? go_row<RK_STEP_FORCE, 1>(a, batch, s) : go_row<RK_STEP_FORCE, 0>(a, batch, s);
    case RK_STEP_ROLL:
      return up ? go_row<RK_STEP_ROLL, 1>(a, batch, s) : go_row<RK_STEP_ROLL, 0>(a, batch, s);
    case RK_ADJ:
      return up ? go_row<RK_ADJ, 1>(a, batch, s) : go_row<RK_ADJ, 0>(a, batch, s);
    default: return hipErrorInvalidValue;
  }
}

namespace {
template <int MODE>
hipError_t go_col(const ColArgs& a, int batch, hipStream_t s) {
  hipLaunchKernelGGL((k_col<NN_, MODE>), dim3(ColCfg<NN_>::NWG, batch), dim3(ColCfg<NN_>::T), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t NSK_CAT(launch_col_, NSK_N)(int mode, const ColArgs& a, int batch, hipStream_t s) {
  switch (mode) {
    case COL_PROJECT: return go_col<COL_PROJECT>(a, batch, s);
    case COL_POISSON: return go_col<COL_POISSON>(a, batch, s);
    case COL_IDENT: return go_col<COL_IDENT>(a, batch, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t NSK_CAT(launch_init_tw_, NSK_N)(hipStream_t s) {
  hipLaunchKernelGGL(k_init_tw, dim3(TW_N / 256), dim3(256), 0, s);
  return hipGetLastError();
}

}  // namespace nsk
