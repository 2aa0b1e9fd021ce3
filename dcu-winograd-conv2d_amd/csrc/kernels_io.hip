// kernels_io.hip -- input transform V = B^T d B and output transform
// Y = A^T M A for gfx950, single-problem and grouped (batched) forms.
//
// Both are HBM-bound streaming kernels with FP32 arithmetic:
//   input  : reference K1 input half / K5 / K6
//            (src/winograd_2x3_non_fused.cpp:302-377, src/winograd_4x3_non_fused.cpp:14-322)
//   output : reference K2 / K8
//            (src/winograd_2x3_non_fused.cpp:167-228, src/winograd_4x3_non_fused.cpp:410-526)
#include "wino_internal.h"
#include "wino_math.h"

namespace wino {

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ------------------------------------------------------------------ input
//
// A block stages CG channels of a window of tiles (plus the 2-pixel halo,
// zero outside the image) into LDS: thread -> (row slot, column), rows
// advance incrementally, so a wave reads contiguous runs of input rows.
// Then each thread transforms one tile for CPT channels (2 for fp16, 4 for
// fp8) and stores CPT consecutive channel values at once: a wave writes
// 256 contiguous bytes of V per tile position.
// PREC: 0 fp16 (CG = 8), 1 fp8, 2 fp8 hi/lo (CG = 32).
template <int m, int CG, int PREC>
__device__ inline void input_block(const half_t* __restrict__ X, void* __restrict__ Vout,
                                   const InputGeom& g, const float* __restrict__ amax,
                                   float* __restrict__ vscale, uint32_t bx, uint32_t cg,
                                   half_t* xs, bool write_scale) {
  constexpr int T = m + 2, T2 = T * T;
  constexpr int CPT = PREC == 0 ? 2 : 4;  // channels per thread in the transform
  uint32_t b = bx;
  const uint32_t wc = b % g.nwc;
  b /= g.nwc;
  const uint32_t wr = b % g.nwr;
  const uint32_t n = b / g.nwr;
  const uint32_t th0 = wr * g.WR, tw0 = wc * g.WC;
  const uint32_t nth = min(g.WR, g.TH - th0), ntw = min(g.WC, g.TW - tw0);
  const int LH = (int)nth * m + 2, LW = (int)ntw * m + 2;
  const int y0 = (int)(th0 * m) - 1, x0 = (int)(tw0 * m) - 1;
  const int tid = threadIdx.x;

  // stage the window: rows (channel-major) x columns
  {
    const int rpi = 256 / LW;  // rows per pass (LW <= 256 by construction)
    const int col = tid % LW;
    int rr = tid / LW;
    int ch = rr / LH, lr = rr - ch * LH;
    const int x = x0 + col;
    const bool xin = x >= 0 && x < (int)g.W;
    const half_t* plane0 = X + ((uint64_t)n * g.C + cg * CG) * g.H * g.W;
    if (rr < rpi) {
      for (; ch < CG; rr += rpi) {
        const int y = y0 + lr;
        const uint32_t c = cg * CG + ch;
        half_t v = (half_t)0.0f;
        if (xin && c < g.C && y >= 0 && y < (int)g.H)
          v = plane0[((uint64_t)ch * g.H + y) * g.W + x];
        xs[(ch * LH + lr) * LW + col] = v;
        lr += rpi;
        while (lr >= LH) {
          lr -= LH;
          ++ch;
        }
      }
    }
  }
  __syncthreads();

  float inv[T2];
  if (PREC != 0) {
    const float a = *amax;
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const float s = pow2_scale(Tf<m>::gain(t / T) * Tf<m>::gain(t % T) * a);
      inv[t] = 1.0f / s;
      if (write_scale && tid == 0) vscale[t] = s;
    }
  }

  constexpr int LANES_PER_TILE = CG / CPT;  // 4 (fp16) or 8 (fp8)
  const int cq = tid % LANES_PER_TILE;       // channel quad/pair index
  const int per_ch = LH * LW;
  const int ntiles = (int)(nth * ntw);
  const uint64_t tplane = (uint64_t)g.CGp * g.Pp * CG;  // elements per tile position
  for (int i = tid / LANES_PER_TILE; i < ntiles; i += 256 / LANES_PER_TILE) {
    const int lth = i / (int)ntw, ltw = i - lth * (int)ntw;
    const uint64_t p = ((uint64_t)n * g.TH + th0 + lth) * g.TW + tw0 + ltw;
    const uint64_t off = ((uint64_t)cg * g.Pp + p) * CG + cq * CPT;
    if (PREC == 0) {
      half2_t out[T2];
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const half_t* src = xs + (cq * 2 + cc) * per_ch + (lth * m) * LW + ltw * m;
        float d[T2], v[T2];
#pragma unroll
        for (int a = 0; a < T; ++a)
#pragma unroll
          for (int c2 = 0; c2 < T; ++c2) d[a * T + c2] = (float)src[a * LW + c2];
        input2d<m>(d, v);
#pragma unroll
        for (int t = 0; t < T2; ++t) out[t][cc] = (half_t)v[t];
      }
      half_t* V = (half_t*)Vout;
#pragma unroll
      for (int t = 0; t < T2; ++t) *(half2_t*)(V + t * tplane + off) = out[t];
    } else {
      float vq[4][T2];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const half_t* src = xs + (cq * 4 + cc) * per_ch + (lth * m) * LW + ltw * m;
        float d[T2];
#pragma unroll
        for (int a = 0; a < T; ++a)
#pragma unroll
          for (int c2 = 0; c2 < T; ++c2) d[a * T + c2] = (float)src[a * LW + c2];
        input2d<m>(d, vq[cc]);
      }
      uint8_t* V = (uint8_t*)Vout;
#pragma unroll
      for (int t = 0; t < T2; ++t) {
        const float x4[4] = {vq[0][t] * inv[t], vq[1][t] * inv[t], vq[2][t] * inv[t],
                             vq[3][t] * inv[t]};
        if (PREC == 2) {
          uint32_t hi, lo;
          e4m3x4_split(x4, hi, lo);
          *(uint32_t*)(V + t * tplane + off) = hi;
          *(uint32_t*)(V + (uint64_t)T2 * tplane + t * tplane + off) = lo;
        } else {
          *(uint32_t*)(V + t * tplane + off) = e4m3x4(x4[0], x4[1], x4[2], x4[3]);
        }
      }
    }
  }
}

template <int m, int CG, int PREC>
__global__ __launch_bounds__(256) void k_input(const half_t* __restrict__ X, void* __restrict__ Vout,
                                               InputGeom g, const float* __restrict__ amax,
                                               float* __restrict__ vscale) {
  extern __shared__ half_t xs_dyn[];
  input_block<m, CG, PREC>(X, Vout, g, amax, vscale, blockIdx.x, blockIdx.y, xs_dyn,
                           blockIdx.x == 0 && blockIdx.y == 0);
}

// ----------------------------------------------------------------- output
//
// One thread per (k, pair of consecutive tiles): the T*T GEMM results of
// both tiles are read as 4-byte pairs (a wave reads contiguous runs of M),
// each tile's m x m patch is written with row-vector stores when in range.
// With split > 1 the GEMM left `split` FP32 partial sums [split][T2][Kp][Pp]
// in M32; they are added in a fixed order (deterministic).
template <int m>
__device__ inline void output_pair(const half_t* __restrict__ M, const float* __restrict__ M32,
                                   int split, half_t* __restrict__ Y, uint32_t K, uint32_t TH,
                                   uint32_t TW, uint64_t P, uint64_t Pp, uint32_t Kp,
                                   uint32_t OH, uint32_t OW, uint64_t idx) {
  constexpr int T = m + 2, T2 = T * T;
  const uint64_t pairs = (P + 1) / 2;
  if (idx >= (uint64_t)K * pairs) return;
  const uint32_t k = (uint32_t)(idx / pairs);
  const uint64_t p0 = 2 * (idx - (uint64_t)k * pairs);
  float x0[T2], x1[T2];
  const uint64_t plane = (uint64_t)Kp * Pp;
  if (split <= 1) {
    const half_t* src = M + (uint64_t)k * Pp + p0;  // Pp is even: 4-byte aligned pairs
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      const half2_t v = *(const half2_t*)(src + t * plane);
      x0[t] = (float)v[0];
      x1[t] = (float)v[1];
    }
  } else {
    const float* src = M32 + (uint64_t)k * Pp + p0;
#pragma unroll
    for (int t = 0; t < T2; ++t) {
      x0[t] = src[t * plane];
      x1[t] = src[t * plane + 1];
    }
    for (int sp = 1; sp < split; ++sp) {
      const float* s2 = src + (uint64_t)sp * T2 * plane;
#pragma unroll
      for (int t = 0; t < T2; ++t) {
        x0[t] += s2[t * plane];
        x1[t] += s2[t * plane + 1];
      }
    }
  }
  const uint32_t per_img = TH * TW;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint64_t p = p0 + h;
    if (p >= P) break;
    float y[m * m];
    output2d<m>(h ? x1 : x0, y);
    const uint32_t n = (uint32_t)(p / per_img);
    const uint32_t rem = (uint32_t)(p - (uint64_t)n * per_img);
    const uint32_t th = rem / TW, tw = rem - (rem / TW) * TW;
    half_t* dst = Y + (((uint64_t)n * K + k) * OH + th * m) * OW + tw * m;
    const bool full_w = tw * m + m <= OW && (OW % m) == 0;
#pragma unroll
    for (int i = 0; i < m; ++i) {
      if (th * m + i >= OH) break;
      if (full_w) {
        if (m == 4) {
          *(half4*)(dst + i * OW) = half4{(half_t)y[i * m], (half_t)y[i * m + 1],
                                          (half_t)y[i * m + 2], (half_t)y[i * m + 3]};
        } else {
          *(half2_t*)(dst + i * OW) = half2_t{(half_t)y[i * m], (half_t)y[i * m + 1]};
        }
      } else {
#pragma unroll
        for (int j = 0; j < m; ++j)
          if (tw * m + j < OW) dst[i * OW + j] = (half_t)y[i * m + j];
      }
    }
  }
}

template <int m>
__global__ __launch_bounds__(256) void k_output(const half_t* __restrict__ M,
                                                half_t* __restrict__ Y, uint32_t K, uint32_t TH,
                                                uint32_t TW, uint64_t P, uint64_t Pp, uint32_t Kp,
                                                uint32_t OH, uint32_t OW) {
  output_pair<m>(M, nullptr, 1, Y, K, TH, TW, P, Pp, Kp, OH, OW,
                 (uint64_t)blockIdx.x * 256u + threadIdx.x);
}

// --------------------------------------------------------------- geometry

void input_window(const Plan& pl, int CG, uint64_t budget, uint32_t& WR, uint32_t& WC) {
  WC = pl.tiles_w < 32u ? pl.tiles_w : 32u;
  for (;;) {
    const uint64_t lw = (uint64_t)WC * pl.m + 2;
    const uint64_t rows = budget / ((uint64_t)CG * 2u * lw);
    if ((rows >= (uint64_t)pl.m + 2 && lw <= 256) || WC == 1) {
      uint64_t wr = rows > 2 ? (rows - 2) / pl.m : 1;
      if (wr < 1) wr = 1;
      // enough tiles for ~4 passes of the block (64 tiles per pass for fp16,
      // 32 for fp8), no more than the image
      const uint64_t tpp = CG == 8 ? 64 : 32;
      const uint64_t want = (4 * tpp + WC - 1) / WC;
      if (wr > want) wr = want;
      if (wr > pl.tiles_h) wr = pl.tiles_h;
      WR = (uint32_t)(wr < 1 ? 1 : wr);
      return;
    }
    WC = (WC + 1) / 2;
  }
}

InputGeom input_geom(const Plan& pl, uint32_t N, uint32_t C, uint32_t H, uint32_t W,
                     uint64_t budget) {
  const int CG = pl.precision == WINO_PREC_FP16 ? 8 : 32;
  InputGeom g;
  g.N = N;
  g.C = C;
  g.H = H;
  g.W = W;
  g.TH = pl.tiles_h;
  g.TW = pl.tiles_w;
  input_window(pl, CG, budget, g.WR, g.WC);
  g.nwr = div_up(g.TH, g.WR);
  g.nwc = div_up(g.TW, g.WC);
  g.CGp = pl.Cp / CG;
  g.Pp = pl.Pp;
  return g;
}

static size_t input_lds(const Plan& pl, const InputGeom& g) {
  const int CG = pl.precision == WINO_PREC_FP16 ? 8 : 32;
  return (size_t)CG * (g.WR * pl.m + 2) * (g.WC * pl.m + 2) * sizeof(half_t);
}

// ------------------------------------------------------ single problem

hipError_t launch_input_transform(const Plan& pl, const _Float16* x, uint32_t N, uint32_t C,
                                  uint32_t H, uint32_t W, void* V, const float* amax,
                                  float* v_scale, hipStream_t st) {
  const InputGeom g = input_geom(pl, N, C, H, W, 40u * 1024u);
  const int CG = pl.precision == WINO_PREC_FP16 ? 8 : 32;
  const size_t lds = input_lds(pl, g);
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
  dim3 grid(div_up((uint64_t)K * ((pl.P + 1) / 2), 256));
  if (pl.m == 2)
    hipLaunchKernelGGL(k_output<2>, grid, dim3(256), 0, st, M, y, K, pl.tiles_h, pl.tiles_w, pl.P,
                       pl.Pp, pl.Kp, pl.oh, pl.ow);
  else
    hipLaunchKernelGGL(k_output<4>, grid, dim3(256), 0, st, M, y, K, pl.tiles_h, pl.tiles_w, pl.P,
                       pl.Pp, pl.Kp, pl.oh, pl.ow);
  return hipGetLastError();
}

// ------------------------------------------------------------- grouped
// Fixed grids; virtual block v is mapped to its problem by a scan over the
// problems' first-block offsets (batches hold at most a few hundred problems).

template <int PREC>
__global__ __launch_bounds__(256) void k_input_grouped(const GroupHeader* __restrict__ hdr,
                                                       const GProb* __restrict__ probs) {
  extern __shared__ half_t gxs[];
  constexpr int CG = PREC == 0 ? 8 : 32;
  const uint32_t total = hdr->in_blocks, np = hdr->nprob;
  for (uint32_t v = blockIdx.x; v < total; v += gridDim.x) {
    uint32_t p = 0;
    while (p + 1 < np && probs[p + 1].in_first <= v) ++p;
    const GProb& pr = probs[p];
    const uint32_t local = v - pr.in_first;
    const uint32_t bx = local % pr.in_gx, cg = local / pr.in_gx;
    if (pr.m == 2)
      input_block<2, CG, PREC>(pr.x, pr.V, pr.ig, pr.amax, pr.vs, bx, cg, gxs, local == 0);
    else
      input_block<4, CG, PREC>(pr.x, pr.V, pr.ig, pr.amax, pr.vs, bx, cg, gxs, local == 0);
    __syncthreads();  // xs is restaged by the next virtual block
  }
}

__global__ __launch_bounds__(256) void k_output_grouped(const GroupHeader* __restrict__ hdr,
                                                        const GProb* __restrict__ probs) {
  const uint32_t total = hdr->out_blocks, np = hdr->nprob;
  for (uint32_t v = blockIdx.x; v < total; v += gridDim.x) {
    uint32_t p = 0;
    while (p + 1 < np && probs[p + 1].out_first <= v) ++p;
    const GProb& pr = probs[p];
    const uint64_t idx = (uint64_t)(v - pr.out_first) * 256u + threadIdx.x;
    if (pr.m == 2)
      output_pair<2>(pr.M, pr.M32, (int)pr.split, pr.y, pr.K, pr.ig.TH, pr.ig.TW, pr.P, pr.Pp,
                     pr.Kp, pr.OH, pr.OW, idx);
    else
      output_pair<4>(pr.M, pr.M32, (int)pr.split, pr.y, pr.K, pr.ig.TH, pr.ig.TW, pr.P, pr.Pp,
                     pr.Kp, pr.OH, pr.OW, idx);
  }
}

hipError_t launch_input_grouped(int precision, const GroupHeader* hdr, const GProb* probs,
                                uint32_t grid, hipStream_t st) {
  if (precision == WINO_PREC_FP16)
    hipLaunchKernelGGL(k_input_grouped<0>, dim3(grid), dim3(256), kGroupInputLds, st, hdr, probs);
  else if (precision == WINO_PREC_FP8)
    hipLaunchKernelGGL(k_input_grouped<1>, dim3(grid), dim3(256), kGroupInputLds, st, hdr, probs);
  else
    hipLaunchKernelGGL(k_input_grouped<2>, dim3(grid), dim3(256), kGroupInputLds, st, hdr, probs);
  return hipGetLastError();
}

hipError_t launch_output_grouped(const GroupHeader* hdr, const GProb* probs, uint32_t grid,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_output_grouped, dim3(grid), dim3(256), 0, st, hdr, probs);
  return hipGetLastError();
}

}  // namespace wino
