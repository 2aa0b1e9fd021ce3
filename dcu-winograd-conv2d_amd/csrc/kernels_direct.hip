// kernels_direct.hip -- the general fallback: direct convolution as an
// implicit GEMM on MFMA, for every problem the Winograd paths do not compute
// (filter size other than 3x3, stride other than 1, padding other than 1).
//
// The reference accepted such problems and returned wrong results
// (src/conv2d.cpp:24-114 ignores r, s, u, v, p, q; include/common.h:76-86
// ignores the stride); its own CPU check computes them correctly
// (include/verfiy.h:5-44).  Here
//   Y[n][k][oy][ox] = sum over (c, kh, kw) of W[k][c][kh][kw] *
//                     X[n][c][oy*u - p + kh][ox*v - q + kw]   (zero outside)
// is a GEMM of the K x (C*R*S) filter matrix (KCRS rows, used in place) with
// the (C*R*S) x (N*OH*OW) im2col matrix, which is never materialised.
//
// Block: 256 threads = 4 waves stacked along the output channels, each wave
// TK x 4 MFMA tiles (v_mfma_f32_16x16x32_f16) = 16 TK channels x 64 output
// pixels; the block covers 64 TK channels x 64 pixels.  Per 32-element
// reduction step every wave loads its own filter rows straight into
// registers (16-byte loads when C*R*S is a multiple of 8), while the block
// gathers the 32 x 64 im2col slice into LDS (double-buffered, gathered two
// steps ahead): thread (wave w, lane l) gathers reduction elements
// 8w .. 8w+7 of the step for output pixel l.  The reduction index -> (c, kh,
// kw) arithmetic is the same for all lanes of a wave, so it runs on the
// scalar unit; a lane adds its pixel's offset and checks the image bounds.
// FP32 accumulation, one fp16 rounding at the end.
#include "wino_internal.h"
#include "wino_math.h"

namespace wino {

template <int TK, bool VEC>
__global__ __launch_bounds__(256) void k_direct_f16(const half_t* __restrict__ X,
                                                    const half_t* __restrict__ Wt,
                                                    half_t* __restrict__ Y, DirectGeom g) {
  constexpr int BK = 64 * TK;
  __shared__ half8 Bs[2][4][64];  // [buffer][8-element group of the step][output pixel]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int g4 = lane >> 4, r16 = lane & 15;
  const uint32_t k0 = blockIdx.y * BK + wave * 16 * TK;  // this wave's first channel
  const uint64_t q0 = (uint64_t)blockIdx.x * 64u;
  const uint32_t ohw = g.OH * g.OW, hw = g.H * g.W;
  const uint32_t nsteps = (g.RED + 31u) / 32u;

  // ---- filter rows of this wave: row k0 + 16 i + r16, elements 8 g4 .. 8 g4 + 7
  const half_t* wrow[TK];
  bool krow_ok[TK];
#pragma unroll
  for (int i = 0; i < TK; ++i) {
    const uint32_t k = k0 + 16 * i + r16;
    krow_ok[i] = k < g.K;
    wrow[i] = Wt + (uint64_t)(krow_ok[i] ? k : 0u) * g.RED + 8 * g4;
  }
  auto load_a = [&](uint32_t s, half8 (&a)[TK]) {
    const uint32_t j = s * 32u + 8u * g4;
#pragma unroll
    for (int i = 0; i < TK; ++i) {
      if constexpr (VEC) {
        a[i] = (krow_ok[i] && j < g.RED) ? *(const half8*)(wrow[i] + s * 32u) : half8{};
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          a[i][e] = (krow_ok[i] && j + e < g.RED) ? wrow[i][s * 32u + e] : (half_t)0;
      }
    }
  };

  // ---- im2col gather: this lane's output pixel
  const uint64_t qb = q0 + lane;
  const bool qb_ok = qb < g.NPIX;
  const uint32_t nb = qb_ok ? (uint32_t)(qb / ohw) : 0u;
  const uint32_t pix = qb_ok ? (uint32_t)(qb - (uint64_t)nb * ohw) : 0u;
  const uint32_t oy = pix / g.OW, ox = pix - oy * g.OW;
  const int ybase = (int)(oy * g.U) - (int)g.P, xbase = (int)(ox * g.V) - (int)g.Q;
  // element offset of (c, kh, kw) = pix_off + c*H*W + kh*W + kw (may start negative)
  const int64_t pix_off = (int64_t)nb * g.C * hw + (int64_t)ybase * g.W + xbase;
  auto gather = [&](uint32_t s, half8& b) {
    // reduction elements j .. j+7 (wave-uniform): j -> (c, kh, kw)
    const uint32_t j = s * 32u + 8u * wave;
    uint32_t c = j / g.RS;
    const uint32_t rem = j - c * g.RS;
    uint32_t kh = rem / g.S, kw = rem - kh * g.S;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int y = ybase + (int)kh, x = xbase + (int)kw;
      const bool ok = qb_ok && c < g.C && (uint32_t)y < g.H && (uint32_t)x < g.W;
      const int64_t off = pix_off + (int64_t)c * hw + (int64_t)kh * g.W + kw;
      b[e] = ok ? X[off] : (half_t)0;
      if (++kw == g.S) {
        kw = 0;
        if (++kh == g.R) {
          kh = 0;
          ++c;
        }
      }
    }
  };

  floatx4 acc[TK][4];
#pragma unroll
  for (int i = 0; i < TK; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  half8 a[TK], b;
  load_a(0, a);
  gather(0, b);
  Bs[0][wave][lane] = b;
  if (nsteps > 1) gather(1, b);
  __syncthreads();
  for (uint32_t s = 0; s < nsteps; ++s) {
    const int cur = s & 1;
    half8 an[TK];
    if (s + 1 < nsteps) load_a(s + 1, an);
    half8 fb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) fb[j] = Bs[cur][g4][j * 16 + r16];
    // D rows = output channels (4 g4 + e), D columns = output pixels (r16)
#pragma unroll
    for (int i = 0; i < TK; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], fb[j], acc[i][j], 0, 0, 0);
    if (s + 1 < nsteps) {
      Bs[cur ^ 1][wave][lane] = b;  // step s + 1
      if (s + 2 < nsteps) gather(s + 2, b);
#pragma unroll
      for (int i = 0; i < TK; ++i) a[i] = an[i];
    }
    __syncthreads();
  }

#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t q = q0 + j * 16 + r16;
    if (q >= g.NPIX) continue;
    const uint32_t n = (uint32_t)(q / ohw);
    const uint32_t pq = (uint32_t)(q - (uint64_t)n * ohw);
    half_t* yp = Y + (uint64_t)n * g.K * ohw + pq;
#pragma unroll
    for (int i = 0; i < TK; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t k = k0 + i * 16 + 4 * g4 + e;
        if (k < g.K) yp[(uint64_t)k * ohw] = (half_t)acc[i][j][e];
      }
  }
}

DirectGeom direct_geom(uint32_t N, uint32_t C, uint32_t H, uint32_t W, uint32_t K, uint32_t R,
                       uint32_t S, uint32_t U, uint32_t V, uint32_t P, uint32_t Q) {
  DirectGeom g;
  g.N = N;
  g.C = C;
  g.H = H;
  g.W = W;
  g.K = K;
  g.R = R;
  g.S = S;
  g.U = U;
  g.V = V;
  g.P = P;
  g.Q = Q;
  g.OH = (H + 2 * P - R) / U + 1;
  g.OW = (W + 2 * Q - S) / V + 1;
  g.RS = R * S;
  g.RED = C * R * S;
  g.NPIX = (uint64_t)N * g.OH * g.OW;
  return g;
}

// Channels per block: 256 (4 waves x 4 tiles) when K > 128, else 128 or 64.
int direct_block_k(uint32_t K) { return K > 128 ? 256 : K > 64 ? 128 : 64; }

hipError_t launch_direct(const DirectGeom& g, const _Float16* x, const _Float16* w, _Float16* y,
                         hipStream_t st) {
  const int bk = direct_block_k(g.K);
  dim3 grid((uint32_t)((g.NPIX + 63) / 64), (g.K + bk - 1) / bk);
  const bool vec = g.RED % 8 == 0;
#define WINO_DIRECT(TK)                                                                        \
  do {                                                                                         \
    if (vec)                                                                                   \
      hipLaunchKernelGGL((k_direct_f16<TK, true>), grid, dim3(256), 0, st, (const half_t*)x,   \
                         (const half_t*)w, (half_t*)y, g);                                     \
    else                                                                                       \
      hipLaunchKernelGGL((k_direct_f16<TK, false>), grid, dim3(256), 0, st, (const half_t*)x,  \
                         (const half_t*)w, (half_t*)y, g);                                     \
  } while (0)
  if (bk == 256)
    WINO_DIRECT(4);
  else if (bk == 128)
    WINO_DIRECT(2);
  else
    WINO_DIRECT(1);
#undef WINO_DIRECT
  return hipGetLastError();
}

}  // namespace wino
