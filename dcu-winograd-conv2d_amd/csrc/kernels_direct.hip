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
// the (C*R*S) x (N*OH*OW) im2col matrix, which is never materialised: each
// block gathers its 32-element reduction slice of the im2col columns into
// LDS with the padding applied, the next slice prefetched into registers.
// FP32 accumulation (v_mfma_f32_16x16x32_f16), one fp16 rounding at the end.
//
// Block: 256 threads = 4 waves in a 2 x 2 grid over 64 output channels x 64
// output pixels; wave tile 32 x 32 = 2 x 2 MFMA tiles.
#include "wino_internal.h"
#include "wino_math.h"

namespace wino {

__global__ __launch_bounds__(256) void k_direct_f16(const half_t* __restrict__ X,
                                                    const half_t* __restrict__ Wt,
                                                    half_t* __restrict__ Y, DirectGeom g) {
  __shared__ half8 As[4][64];  // [8-element group of the step][output channel]
  __shared__ half8 Bs[4][64];  // [8-element group of the step][output pixel]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wk = wave & 1, wp = wave >> 1;
  const int g4 = lane >> 4, r16 = lane & 15;
  const uint32_t k0 = blockIdx.y * 64u;
  const uint64_t q0 = (uint64_t)blockIdx.x * 64u;
  const uint32_t ohw = g.OH * g.OW;

  // staging: this thread fills group sg of channel row / pixel column srow
  const int sg = tid >> 6, srow = tid & 63;
  const uint32_t ka = k0 + srow;
  const bool ka_ok = ka < g.K;
  const half_t* wrow = Wt + (uint64_t)(ka_ok ? ka : 0u) * g.RED;
  const uint64_t qb = q0 + srow;
  const bool qb_ok = qb < g.NPIX;
  const uint32_t nb = qb_ok ? (uint32_t)(qb / ohw) : 0u;
  const uint32_t pix = qb_ok ? (uint32_t)(qb - (uint64_t)nb * ohw) : 0u;
  const uint32_t oy = pix / g.OW, ox = pix - oy * g.OW;
  const int ybase = (int)(oy * g.U) - (int)g.P, xbase = (int)(ox * g.V) - (int)g.Q;
  const half_t* xn = X + (uint64_t)nb * g.C * g.H * g.W;

  auto gather = [&](uint32_t j0, half8& a, half8& b) {
    const uint32_t j = j0 + 8u * sg;
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = (ka_ok && j + e < g.RED) ? wrow[j + e] : (half_t)0;
    // reduction index j -> (channel, filter row, filter column), then stepped
    uint32_t c = j / g.RS;
    const uint32_t rem = j - c * g.RS;
    uint32_t kh = rem / g.S, kw = rem - kh * g.S;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      half_t v = (half_t)0;
      const int y = ybase + (int)kh, x = xbase + (int)kw;
      if (qb_ok && c < g.C && y >= 0 && y < (int)g.H && x >= 0 && x < (int)g.W)
        v = xn[((uint64_t)c * g.H + (uint32_t)y) * g.W + (uint32_t)x];
      b[e] = v;
      if (++kw == g.S) {
        kw = 0;
        if (++kh == g.R) {
          kh = 0;
          ++c;
        }
      }
    }
  };

  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  half8 a, b;
  gather(0, a, b);
  for (uint32_t j0 = 0; j0 < g.RED; j0 += 32u) {
    __syncthreads();  // the previous step's fragments have been read
    As[sg][srow] = a;
    Bs[sg][srow] = b;
    __syncthreads();
    if (j0 + 32u < g.RED) gather(j0 + 32u, a, b);  // next slice, in flight during the MFMAs
    half8 fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) fa[i] = As[g4][wk * 32 + i * 16 + r16];
#pragma unroll
    for (int j = 0; j < 2; ++j) fb[j] = Bs[g4][wp * 32 + j * 16 + r16];
    // D rows = output channels (4 g4 + e), D columns = output pixels (r16)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[i], fb[j], acc[i][j], 0, 0, 0);
  }

#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint64_t q = q0 + wp * 32 + j * 16 + r16;
    if (q >= g.NPIX) continue;
    const uint32_t n = (uint32_t)(q / ohw);
    const uint32_t pq = (uint32_t)(q - (uint64_t)n * ohw);
    half_t* yp = Y + (uint64_t)n * g.K * ohw + pq;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t k = k0 + wk * 32 + i * 16 + 4 * g4 + e;
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

hipError_t launch_direct(const DirectGeom& g, const _Float16* x, const _Float16* w, _Float16* y,
                         hipStream_t st) {
  dim3 grid((uint32_t)((g.NPIX + 63) / 64), (g.K + 63) / 64);
  hipLaunchKernelGGL(k_direct_f16, grid, dim3(256), 0, st, (const half_t*)x, (const half_t*)w,
                     (half_t*)y, g);
  return hipGetLastError();
}

}  // namespace wino
