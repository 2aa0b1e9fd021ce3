// kernels_transform.hip -- filter, input and output transforms for gfx950.
//
// All three are HBM-bound streaming kernels: FP32 arithmetic, fp16 (or e4m3)
// operands written in the GEMM's blocked layout (wino_internal.h).
//   filter : reference K1 filter half / K7
//            (src/winograd_2x3_non_fused.cpp:249-301, src/winograd_4x3_non_fused.cpp:327-405)
//   input  : reference K1 input half / K5 / K6
//            (src/winograd_2x3_non_fused.cpp:302-377, src/winograd_4x3_non_fused.cpp:14-322)
//   output : reference K2 / K8
//            (src/winograd_2x3_non_fused.cpp:167-228, src/winograd_4x3_non_fused.cpp:410-526)
#include "wino_internal.h"
#include "wino_math.h"

namespace wino {

// ---------------------------------------------------------------- filter

// One thread per (c8 group, k, lane-in-group j): 8 consecutive threads write
// one 16-byte U entry, a wave writes 128 contiguous bytes per tile position.
template <int m>
__global__ __launch_bounds__(256) void k_filter_f16(const half_t* __restrict__ W,
                                                    half_t* __restrict__ U, uint32_t K, uint32_t C,
                                                    uint32_t Kp, uint32_t C8p) {
  constexpr int T = m + 2, T2 = T * T;
  const uint64_t idx = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t total = (uint64_t)Kp * C8p * 8u;
  if (idx >= total) return;
  const uint32_t j = (uint32_t)(idx & 7u);
  const uint64_t r = idx >> 3;
  const uint32_t k = (uint32_t)(r % Kp);
  const uint32_t c8 = (uint32_t)(r / Kp);
  const uint32_t c = c8 * 8u + j;
  float u[T2];
  if (k < K && c < C) {
    const half_t* src = W + ((uint64_t)k * C + c) * 9u;
    float g[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = (float)src[i];
    filter2d<m>(g, u);
  } else {
#pragma unroll
    for (int t = 0; t < T2; ++t) u[t] = 0.0f;
  }
  const uint64_t plane = (uint64_t)C8p * Kp * 8u;
  half_t* dst = U + ((uint64_t)c8 * Kp + k) * 8u + j;
#pragma unroll
  for (int t = 0; t < T2; ++t) dst[(uint64_t)t * plane] = (half_t)u[t];
}

// One 64-lane block per output channel k: pass 1 finds amax over c of every
// tile position (per-(t,k) scale, a power of two), pass 2 quantizes.
// PLANES = 2 also writes lo = e4m3((x - hi) * 16).
template <int m, int PLANES>
__global__ __launch_bounds__(64) void k_filter_f8(const half_t* __restrict__ W,
                                                  uint8_t* __restrict__ U, float* __restrict__ us,
                                                  uint32_t K, uint32_t C, uint32_t Kp,
                                                  uint32_t C32p) {
  constexpr int T = m + 2, T2 = T * T;
  const uint32_t k = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  float amax[T2];
#pragma unroll
  for (int t = 0; t < T2; ++t) amax[t] = 0.0f;
  if (k < K) {
    for (uint32_t c = lane; c < C; c += 64u) {
      const half_t* src = W + ((uint64_t)k * C + c) * 9u;
      float g[9], u[T2];
#pragma unroll
      for (int i = 0; i < 9; ++i) g[i] = (float)src[i];
      filter2d<m>(g, u);
#pragma unroll
      for (int t = 0; t < T2; ++t) amax[t] = fmaxf(amax[t], fabsf(u[t]));
    }
  }
  float inv[T2];
#pragma unroll
  for (int t = 0; t < T2; ++t) {
    float a = amax[t];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) a = fmaxf(a, __shfl_xor(a, o));
    const float s = pow2_scale(a);
    inv[t] = 1.0f / s;  // exact: s is a power of two
    if (lane == 0) us[(uint64_t)t * Kp + k] = s;
  }
  const uint64_t plane = (uint64_t)T2 * C32p * Kp * 32u;
  for (uint32_t c = lane; c < C32p * 32u; c += 64u) {
    float u[T2];
    if (k < K && c < C) {
      const half_t* src = W + ((uint64_t)k * C + c) * 9u;
      float g[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) g[i] = (float)src[i];
      filter2d<m>(g, u);
    } else {
#pragma unroll
      for (int t = 0; t < T2; ++t) u[t] = 0.0f;
    }
    const uint64_t base = ((uint64_t)(c >> 5) * Kp + k) * 32u + (c & 31u);
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const float x = u[t] * inv[t];
      const uint32_t hi = f32_to_e4m3(x);
      const uint64_t off = (uint64_t)t * C32p * Kp * 32u + base;
      U[off] = (uint8_t)hi;
      if (PLANES == 2) U[plane + off] = (uint8_t)f32_to_e4m3((x - e4m3_to_f32(hi)) * 16.0f);
    }
  }
}

// ----------------------------------------------------------- input amax

__global__ void k_amax_zero(float* amax) { *amax = 0.0f; }

// max |x| over n halves: 16-byte loads when x is 16-byte aligned, a block
// reduction in LDS and one atomic per block (non-negative floats order as
// their bit patterns).
template <bool VEC>
__global__ __launch_bounds__(256) void k_amax(const half_t* __restrict__ x, uint64_t n,
                                              float* __restrict__ amax) {
  __shared__ float red[4];
  float a = 0.0f;
  const uint64_t stride = (uint64_t)gridDim.x * 256u;
  const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  uint64_t tail = 0;
  if (VEC) {
    const uint64_t nv = n / 8;
    const half8* xv = (const half8*)x;
    for (uint64_t i = gid; i < nv; i += stride) {
      const half8 v = xv[i];
#pragma unroll
      for (int j = 0; j < 8; ++j) a = fmaxf(a, fabsf((float)v[j]));
    }
    tail = nv * 8;
  }
  for (uint64_t i = tail + gid; i < n; i += stride) a = fmaxf(a, fabsf((float)x[i]));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) a = fmaxf(a, __shfl_xor(a, o));
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    a = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    atomicMax((unsigned int*)amax, __float_as_uint(a));
  }
}

// ------------------------------------------------------------------ input

// A block stages CG channels of a window of WR x WC tiles (plus the 2-pixel
// halo, zero outside the image) into LDS with coalesced row reads, then each
// thread transforms one (tile, channel) pair from LDS; consecutive lanes are
// consecutive channels of one tile, so a wave writes contiguous V rows.
// PREC: 0 fp16 (CG = 8), 1 fp8, 2 fp8 hi/lo (CG = 32).
struct InputGeom {
  uint32_t N, C, H, W;
  uint32_t TH, TW;     // tile grid
  uint32_t WR, WC;     // window in tiles
  uint32_t nwr, nwc;   // windows per image
  uint32_t CGp;        // channel groups in the padded layout (Cp / CG)
  uint64_t Pp;
};

template <int m, int CG, int PREC>
__global__ __launch_bounds__(256) void k_input(const half_t* __restrict__ X, void* __restrict__ Vout,
                                               InputGeom g, const float* __restrict__ amax,
                                               float* __restrict__ vscale) {
  constexpr int T = m + 2, T2 = T * T;
  extern __shared__ half_t xs[];
  uint32_t b = blockIdx.x;
  const uint32_t wc = b % g.nwc;
  b /= g.nwc;
  const uint32_t wr = b % g.nwr;
  const uint32_t n = b / g.nwr;
  const uint32_t cg = blockIdx.y;
  const uint32_t th0 = wr * g.WR, tw0 = wc * g.WC;
  const uint32_t nth = min(g.WR, g.TH - th0), ntw = min(g.WC, g.TW - tw0);
  const int LH = (int)nth * m + 2, LW = (int)ntw * m + 2;
  const int y0 = (int)(th0 * m) - 1, x0 = (int)(tw0 * m) - 1;
  const int tid = threadIdx.x;

  // stage the window (zero halo / zero channels beyond C)
  const int per_ch = LH * LW;
  for (int e = tid; e < CG * per_ch; e += 256) {
    const int ch = e / per_ch;
    const int rem = e - ch * per_ch;
    const int row = rem / LW, col = rem - (rem / LW) * LW;
    const int y = y0 + row, x = x0 + col;
    const uint32_t c = cg * CG + ch;
    half_t v = (half_t)0.0f;
    if (c < g.C && y >= 0 && y < (int)g.H && x >= 0 && x < (int)g.W)
      v = X[(((uint64_t)n * g.C + c) * g.H + y) * g.W + x];
    xs[e] = v;
  }
  __syncthreads();

  float inv[T2];
  if (PREC != 0) {
    const float a = *amax;
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const float s = pow2_scale(Tf<m>::gain(t / T) * Tf<m>::gain(t % T) * a);
      inv[t] = 1.0f / s;
      if (blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) vscale[t] = s;
    }
  }

  const int cl = tid % CG;
  const int ntiles = (int)(nth * ntw);
  const uint64_t tplane = (uint64_t)g.CGp * g.Pp * CG;  // elements per tile position
  for (int i = tid / CG; i < ntiles; i += 256 / CG) {
    const int lth = i / (int)ntw, ltw = i - lth * (int)ntw;
    const half_t* src = xs + cl * per_ch + (lth * m) * LW + ltw * m;
    float d[T2], v[T2];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
      for (int c2 = 0; c2 < T; ++c2) d[a * T + c2] = (float)src[a * LW + c2];
    input2d<m>(d, v);
    const uint64_t p = ((uint64_t)n * g.TH + th0 + lth) * g.TW + tw0 + ltw;
    const uint64_t off = ((uint64_t)cg * g.Pp + p) * CG + cl;
    if (PREC == 0) {
      half_t* V = (half_t*)Vout;
#pragma unroll
      for (int t = 0; t < T2; ++t) V[t * tplane + off] = (half_t)v[t];
    } else {
      uint8_t* V = (uint8_t*)Vout;
#pragma unroll
      for (int t = 0; t < T2; ++t) {
        const float x = v[t] * inv[t];
        const uint32_t hi = f32_to_e4m3(x);
        V[t * tplane + off] = (uint8_t)hi;
        if (PREC == 2)
          V[(uint64_t)T2 * tplane + t * tplane + off] =
              (uint8_t)f32_to_e4m3((x - e4m3_to_f32(hi)) * 16.0f);
      }
    }
  }
}

// ----------------------------------------------------------------- output

// One thread per (k, tile); reads the T*T GEMM results of its tile (a wave
// reads contiguous runs of M along p) and writes the m x m output patch.
template <int m>
__global__ __launch_bounds__(256) void k_output(const half_t* __restrict__ M,
                                                half_t* __restrict__ Y, uint32_t K, uint32_t TH,
                                                uint32_t TW, uint64_t P, uint64_t Pp, uint32_t Kp,
                                                uint32_t OH, uint32_t OW) {
  constexpr int T = m + 2, T2 = T * T;
  const uint64_t idx = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (idx >= (uint64_t)K * P) return;
  const uint32_t k = (uint32_t)(idx / P);
  const uint64_t p = idx - (uint64_t)k * P;
  const uint32_t per_img = TH * TW;
  const uint32_t n = (uint32_t)(p / per_img);
  const uint32_t rem = (uint32_t)(p - (uint64_t)n * per_img);
  const uint32_t th = rem / TW, tw = rem - (rem / TW) * TW;
  float x[T2], y[m * m];
  const uint64_t plane = (uint64_t)Kp * Pp;
  const half_t* src = M + (uint64_t)k * Pp + p;
#pragma unroll
  for (int t = 0; t < T2; ++t) x[t] = (float)src[t * plane];
  output2d<m>(x, y);
  half_t* dst = Y + (((uint64_t)n * K + k) * OH + th * m) * OW + tw * m;
#pragma unroll
  for (int i = 0; i < m; ++i)
#pragma unroll
    for (int j = 0; j < m; ++j)
      if (th * m + i < OH && tw * m + j < OW) dst[i * OW + j] = (half_t)y[i * m + j];
}

// --------------------------------------------------------------- launchers

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

hipError_t launch_filter_transform(const Plan& pl, const _Float16* w, uint32_t K, uint32_t C,
                                   void* U, float* u_scale, hipStream_t st) {
  if (pl.precision == WINO_PREC_FP16) {
    const uint32_t C8p = pl.Cp / 8;
    const uint64_t total = (uint64_t)pl.Kp * C8p * 8u;
    dim3 grid(div_up(total, 256));
    if (pl.m == 2)
      hipLaunchKernelGGL(k_filter_f16<2>, grid, dim3(256), 0, st, w, (half_t*)U, K, C, pl.Kp, C8p);
    else
      hipLaunchKernelGGL(k_filter_f16<4>, grid, dim3(256), 0, st, w, (half_t*)U, K, C, pl.Kp, C8p);
  } else {
    const uint32_t C32p = pl.Cp / 32;
    dim3 grid(pl.Kp);
    const bool two = pl.precision == WINO_PREC_FP8X2;
    if (pl.m == 2) {
      if (two)
        hipLaunchKernelGGL((k_filter_f8<2, 2>), grid, dim3(64), 0, st, w, (uint8_t*)U, u_scale, K, C, pl.Kp, C32p);
      else
        hipLaunchKernelGGL((k_filter_f8<2, 1>), grid, dim3(64), 0, st, w, (uint8_t*)U, u_scale, K, C, pl.Kp, C32p);
    } else {
      if (two)
        hipLaunchKernelGGL((k_filter_f8<4, 2>), grid, dim3(64), 0, st, w, (uint8_t*)U, u_scale, K, C, pl.Kp, C32p);
      else
        hipLaunchKernelGGL((k_filter_f8<4, 1>), grid, dim3(64), 0, st, w, (uint8_t*)U, u_scale, K, C, pl.Kp, C32p);
    }
  }
  return hipGetLastError();
}

hipError_t launch_input_amax(const _Float16* x, uint64_t count, float* amax, hipStream_t st) {
  hipLaunchKernelGGL(k_amax_zero, dim3(1), dim3(1), 0, st, amax);
  uint32_t blocks = div_up(count, 256u * 8u * 4u);
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) blocks = 1;
  if (((uintptr_t)x & 15u) == 0)
    hipLaunchKernelGGL(k_amax<true>, dim3(blocks), dim3(256), 0, st, x, count, amax);
  else
    hipLaunchKernelGGL(k_amax<false>, dim3(blocks), dim3(256), 0, st, x, count, amax);
  return hipGetLastError();
}

// Window of WR x WC tiles whose CG-channel LDS image fits the budget.
static void input_window(const Plan& pl, int CG, uint32_t& WR, uint32_t& WC) {
  const uint64_t budget = 40u * 1024u;  // bytes of LDS per block
  WC = pl.tiles_w < 32u ? pl.tiles_w : 32u;
  for (;;) {
    const uint64_t lw = (uint64_t)WC * pl.m + 2;
    const uint64_t rows = budget / ((uint64_t)CG * 2u * lw);
    if (rows >= (uint64_t)pl.m + 2 || WC == 1) {
      uint64_t wr = rows > 2 ? (rows - 2) / pl.m : 1;
      if (wr < 1) wr = 1;
      // enough tiles for a few passes of the block, no more than the image
      const uint64_t want = (uint64_t)(4 * 256 / CG + WC - 1) / WC;
      if (wr > want) wr = want;
      if (wr > pl.tiles_h) wr = pl.tiles_h;
      WR = (uint32_t)(wr < 1 ? 1 : wr);
      return;
    }
    WC = (WC + 1) / 2;
  }
}

hipError_t launch_input_transform(const Plan& pl, const _Float16* x, uint32_t N, uint32_t C,
                                  uint32_t H, uint32_t W, void* V, const float* amax,
                                  float* v_scale, hipStream_t st) {
  const int CG = pl.precision == WINO_PREC_FP16 ? 8 : 32;
  InputGeom g;
  g.N = N;
  g.C = C;
  g.H = H;
  g.W = W;
  g.TH = pl.tiles_h;
  g.TW = pl.tiles_w;
  input_window(pl, CG, g.WR, g.WC);
  g.nwr = div_up(g.TH, g.WR);
  g.nwc = div_up(g.TW, g.WC);
  g.CGp = pl.Cp / CG;
  g.Pp = pl.Pp;
  const size_t lds = (size_t)CG * (g.WR * pl.m + 2) * (g.WC * pl.m + 2) * sizeof(half_t);
  dim3 grid((uint32_t)((uint64_t)N * g.nwr * g.nwc), div_up(C, CG));
  if (pl.precision == WINO_PREC_FP16) {
    if (pl.m == 2)
      hipLaunchKernelGGL((k_input<2, 8, 0>), grid, dim3(256), lds, st, x, V, g, amax, v_scale);
    else
      hipLaunchKernelGGL((k_input<4, 8, 0>), grid, dim3(256), lds, st, x, V, g, amax, v_scale);
  } else if (pl.precision == WINO_PREC_FP8) {
    if (pl.m == 2)
      hipLaunchKernelGGL((k_input<2, 32, 1>), grid, dim3(256), lds, st, x, V, g, amax, v_scale);
    else
      hipLaunchKernelGGL((k_input<4, 32, 1>), grid, dim3(256), lds, st, x, V, g, amax, v_scale);
  } else {
    if (pl.m == 2)
      hipLaunchKernelGGL((k_input<2, 32, 2>), grid, dim3(256), lds, st, x, V, g, amax, v_scale);
    else
      hipLaunchKernelGGL((k_input<4, 32, 2>), grid, dim3(256), lds, st, x, V, g, amax, v_scale);
  }
  return hipGetLastError();
}

hipError_t launch_output_transform(const Plan& pl, const _Float16* M, uint32_t N, uint32_t K,
                                   _Float16* y, hipStream_t st) {
  (void)N;
  dim3 grid(div_up((uint64_t)K * pl.P, 256));
  if (pl.m == 2)
    hipLaunchKernelGGL(k_output<2>, grid, dim3(256), 0, st, M, y, K, pl.tiles_h, pl.tiles_w, pl.P,
                       pl.Pp, pl.Kp, pl.oh, pl.ow);
  else
    hipLaunchKernelGGL(k_output<4>, grid, dim3(256), 0, st, M, y, K, pl.tiles_h, pl.tiles_w, pl.P,
                       pl.Pp, pl.Kp, pl.oh, pl.ow);
  return hipGetLastError();
}

}  // namespace wino
