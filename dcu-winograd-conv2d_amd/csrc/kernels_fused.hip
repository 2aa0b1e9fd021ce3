// kernels_fused.hip -- fused F(2x3) Winograd conv for gfx950.
//
// Counterpart of the reference's winograd_2x3_kernel (K9,
// src/winograd_2x3_fused.cpp:10-369): input transform, the 16 per-position
// GEMMs and the output transform in one launch, transformed tiles never
// leaving LDS/registers.  Re-designed for CDNA4:
//  * the filter is pre-transformed once (resident U, fp16 [16][C/8][Kp][8]);
//    MFMA A-fragments are read straight from U (one 16-byte load per lane),
//    each U element is read once per block;
//  * a block owns a spatial region (NB images x RH tile rows x RW tile
//    columns, <= 64 tiles) and 32 output channels; the region's input is
//    staged into LDS with coalesced row reads, prefetched into registers one
//    32-channel step ahead;
//  * wave w owns tile positions t = 4w..4w+3 (row xi = w of the 4x4 grid),
//    so the A^T(.)A output transform's column pass runs in registers and only
//    2 of 4 partial rows go through LDS;
//  * two barriers per 32-channel step, each LDS region written only after
//    the barrier that retires its readers (fixes the reference's WAR race,
//    src/winograd_2x3_fused.cpp:60-284 with one barrier at :195);
//  * blocks are renumbered so that each XCD works on contiguous regions with
//    all their output-channel blocks back to back (the region's input is
//    then read from HBM once per XCD L2).
#include "wino_internal.h"
#include "wino_math.h"

namespace wino {

constexpr int kFusedMaxLd = 48;  // staged input halves per thread per step

struct FusedGeom {
  uint32_t N, C, H, W, K;
  uint32_t TH, TW, OH, OW;
  uint32_t RH, RW, NB;   // region shape
  uint32_t nrh, nrw;     // regions per image set
  uint32_t Kp, C8p;
  uint32_t kblocks;      // Kp / 32
  uint32_t nwork;        // regions * kblocks
  uint32_t LH, LW;       // staged rows / cols per image (2RH+2, 2RW+2)
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_fused_f23(
    const half_t* __restrict__ X, const half8* __restrict__ U, half_t* __restrict__ Y,
    FusedGeom g) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  half8* Vl = (half8*)smem;                    // [16 t][4 grp][64 tile]  64 KiB
  float* Rl = (float*)smem;                    // [4 xi][2 j][32 k][64 tile] (after loop)
  half_t* xs = (half_t*)(smem + 65536);        // [32 ch][NB][LH][LW]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int grp = lane >> 4, r16 = lane & 15;

  // XCD-contiguous renumbering (block b runs on XCD b % 8).
  const uint32_t nb = gridDim.x, q8 = nb / 8, r8 = nb % 8;
  const uint32_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const uint32_t w = xcd * q8 + (xcd < r8 ? xcd : r8) + slot;
  const uint32_t kb = w % g.kblocks;
  uint32_t region = w / g.kblocks;
  const uint32_t rw = region % g.nrw;
  region /= g.nrw;
  const uint32_t rh = region % g.nrh;
  const uint32_t nbi = region / g.nrh;
  const uint32_t n0 = nbi * g.NB, th0 = rh * g.RH, tw0 = rw * g.RW;
  const int y0 = (int)(2 * th0) - 1, x0 = (int)(2 * tw0) - 1;
  const int LH = (int)g.LH, LW = (int)g.LW;
  const int per_img = LH * LW;
  const int per_ch = (int)g.NB * per_img;
  const int stage_elems = 32 * per_ch;

  // staged-input prefetch (registers) for one 32-channel step
  half_t pf[kFusedMaxLd];
  auto prefetch = [&](int s) {
#pragma unroll
    for (int i = 0; i < kFusedMaxLd; ++i) {
      const int e = tid + i * 256;
      half_t v = (half_t)0.0f;
      if (e < stage_elems) {
        const int ch = e / per_ch;
        int rem = e - ch * per_ch;
        const int bi = rem / per_img;
        rem -= bi * per_img;
        const int row = rem / LW, col = rem - row * LW;
        const uint32_t c = (uint32_t)(32 * s + ch), n = n0 + bi;
        const int y = y0 + row, x = x0 + col;
        if (c < g.C && n < g.N && y >= 0 && y < (int)g.H && x >= 0 && x < (int)g.W)
          v = X[(((uint64_t)n * g.C + c) * g.H + y) * g.W + x];
      }
      pf[i] = v;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < kFusedMaxLd; ++i) {
      const int e = tid + i * 256;
      if (e < stage_elems) xs[e] = pf[i];
    }
  };

  // U fragments of this wave: t = 4*wave + tl, 2 row tiles of 16 channels k.
  // Issued right after the staging barrier; the transform phase hides them.
  half8 ua[4][2];
  auto uload = [&](int s) {
#pragma unroll
    for (int tl = 0; tl < 4; ++tl)
#pragma unroll
      for (int kt = 0; kt < 2; ++kt) {
        const uint32_t t = 4 * wave + tl;
        ua[tl][kt] =
            U[((uint64_t)t * g.C8p + 4 * s + grp) * g.Kp + kb * 32 + kt * 16 + r16];
      }
  };

  // this thread's tile in the transform phase
  const int it = tid & 63, ig = tid >> 6;  // tile, channel group (8 channels)
  const int rsz = (int)(g.RH * g.RW);
  const int ibi = it / rsz, ilr = (it % rsz) / (int)g.RW, ilc = it % (int)g.RW;
  const bool tvalid = it < (int)g.NB * rsz && n0 + ibi < g.N && th0 + ilr < g.TH &&
                      tw0 + ilc < g.TW;

  floatx4 acc[4][2][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[a][b][c] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int nsteps = (int)(g.C8p / 4);
  prefetch(0);
  for (int s = 0; s < nsteps; ++s) {
    stage();
    __syncthreads();
    uload(s);

    // input transform of 8 channels of one tile, 4 channels at a time
    // -> 16 x 8-byte V entries per half
    half4* Vl4 = (half4*)Vl;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      half4 vo[16];
      if (tvalid) {
        const half_t* src =
            xs + (8 * ig + 4 * hh) * per_ch + ibi * per_img + (2 * ilr) * LW + 2 * ilc;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
          float d[16], v[16];
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) d[a * 4 + b] = (float)src[cc * per_ch + a * LW + b];
          input2d<2>(d, v);
#pragma unroll
          for (int t = 0; t < 16; ++t) vo[t][cc] = (half_t)v[t];
        }
      } else {
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) vo[t][cc] = (half_t)0.0f;
      }
#pragma unroll
      for (int t = 0; t < 16; ++t) Vl4[((t * 4 + ig) * 64 + it) * 2 + hh] = vo[t];
    }
    __syncthreads();

    if (s + 1 < nsteps) prefetch(s + 1);

#pragma unroll
    for (int tl = 0; tl < 4; ++tl) {
      const int t = 4 * wave + tl;
      half8 b[4];
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) b[pt] = Vl[(t * 4 + grp) * 64 + pt * 16 + r16];
#pragma unroll
      for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int pt = 0; pt < 4; ++pt)
          acc[tl][kt][pt] =
              __builtin_amdgcn_mfma_f32_16x16x32_f16(ua[tl][kt], b[pt], acc[tl][kt][pt], 0, 0, 0);
    }
  }

  // column pass of A^T M A in registers: R[j] = sum_nu A^T[j][nu] M[xi=wave][nu]
  __syncthreads();  // all waves done reading Vl (Rl aliases it)
#pragma unroll
  for (int kt = 0; kt < 2; ++kt)
#pragma unroll
    for (int pt = 0; pt < 4; ++pt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float m0 = acc[0][kt][pt][e], m1 = acc[1][kt][pt][e], m2 = acc[2][kt][pt][e],
                    m3 = acc[3][kt][pt][e];
        const int k = kt * 16 + 4 * grp + e, p = pt * 16 + r16;
        Rl[((wave * 2 + 0) * 32 + k) * 64 + p] = m0 + m1 + m2;
        Rl[((wave * 2 + 1) * 32 + k) * 64 + p] = m1 - m2 - m3;
      }
  __syncthreads();

  // row pass and store: Y[i][j] = sum_xi A^T[i][xi] R[xi][j]
  for (int q = tid; q < 32 * 64; q += 256) {
    const int k = q >> 6, p = q & 63;
    const uint32_t kk = kb * 32 + k;
    const int bi = p / rsz, lr = (p % rsz) / (int)g.RW, lc = p % (int)g.RW;
    const uint32_t n = n0 + bi, th = th0 + lr, tw = tw0 + lc;
    if (kk >= g.K || p >= (int)g.NB * rsz || n >= g.N || th >= g.TH || tw >= g.TW) continue;
    float R[4][2];
#pragma unroll
    for (int xi = 0; xi < 4; ++xi)
#pragma unroll
      for (int j = 0; j < 2; ++j) R[xi][j] = Rl[((xi * 2 + j) * 32 + k) * 64 + p];
    half_t* dst = Y + (((uint64_t)n * g.K + kk) * g.OH + 2 * th) * g.OW + 2 * tw;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float y0v = R[0][j] + R[1][j] + R[2][j];
      const float y1v = R[1][j] - R[2][j] - R[3][j];
      if (2 * tw + j < g.OW) {
        dst[j] = (half_t)y0v;
        if (2 * th + 1 < g.OH) dst[g.OW + j] = (half_t)y1v;
      }
    }
  }
}

// Region of <= 64 tiles maximising useful lanes, within the staging budget.
static void fused_region(uint32_t N, uint32_t TH, uint32_t TW, uint32_t& RH, uint32_t& RW,
                         uint32_t& NB) {
  double best = -1.0;
  RH = RW = NB = 1;
  for (uint32_t rw = 1; rw <= (TW < 64 ? TW : 64); ++rw)
    for (uint32_t rh = 1; rh <= TH && rh * rw <= 64; ++rh) {
      const bool full = rh == TH && rw == TW;
      const uint32_t nbmax = full ? 64 / (rh * rw) : 1;
      for (uint32_t nbv = 1; nbv <= nbmax && nbv <= N; ++nbv) {
        const uint64_t elems = 32ull * nbv * (2 * rh + 2) * (2 * rw + 2);
        if (elems > 256ull * kFusedMaxLd) continue;
        const double blocks = (double)((N + nbv - 1) / nbv) * ((TH + rh - 1) / rh) *
                              ((TW + rw - 1) / rw);
        const double util = (double)N * TH * TW / (blocks * 64.0);
        // prefer less halo re-read on ties
        const double score = util - 1e-3 * (double)elems / (nbv * rh * rw * 32.0 * 16.0);
        if (score > best) {
          best = score;
          RH = rh;
          RW = rw;
          NB = nbv;
        }
      }
    }
}

hipError_t launch_fused_f23(const Plan& pl, const _Float16* x, const void* U, uint32_t N,
                            uint32_t C, uint32_t H, uint32_t W, uint32_t K, _Float16* y,
                            hipStream_t st) {
  FusedGeom g;
  g.N = N;
  g.C = C;
  g.H = H;
  g.W = W;
  g.K = K;
  g.TH = pl.tiles_h;
  g.TW = pl.tiles_w;
  g.OH = pl.oh;
  g.OW = pl.ow;
  fused_region(N, g.TH, g.TW, g.RH, g.RW, g.NB);
  g.nrh = (g.TH + g.RH - 1) / g.RH;
  g.nrw = (g.TW + g.RW - 1) / g.RW;
  g.Kp = pl.Kp;
  g.C8p = pl.Cp / 8;
  g.kblocks = pl.Kp / 32;
  g.LH = 2 * g.RH + 2;
  g.LW = 2 * g.RW + 2;
  const uint64_t regions = (uint64_t)((N + g.NB - 1) / g.NB) * g.nrh * g.nrw;
  g.nwork = (uint32_t)(regions * g.kblocks);
  const size_t lds = 65536 + (size_t)32 * g.NB * g.LH * g.LW * sizeof(half_t);
  (void)hipFuncSetAttribute((const void*)k_fused_f23, hipFuncAttributeMaxDynamicSharedMemorySize,
                      (int)lds);
  hipLaunchKernelGGL(k_fused_f23, dim3(g.nwork), dim3(256), lds, st, x, (const half8*)U, y, g);
  return hipGetLastError();
}

}  // namespace wino
