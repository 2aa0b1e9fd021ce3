// kernels_fused.hip -- fused F(2x3) Winograd conv for gfx950.
//
// Counterpart of the reference's winograd_2x3_kernel (K9,
// src/winograd_2x3_fused.cpp:10-369): input transform, the 16 per-position
// GEMMs and the output transform in one launch, transformed tiles never
// leaving LDS/registers.  Re-designed for CDNA4:
//  * the filter is pre-transformed once (resident U, fp16 [16][C/8][Kp][8]);
//    MFMA A-fragments are read straight from U (one 16-byte buffer load per
//    lane, no per-load address arithmetic), one 32-channel step ahead;
//  * a block owns a spatial region (NB images x RH tile rows x RW tile
//    columns, <= 32 tiles) and 64 output channels.  The region's input is
//    loaded two 32-channel steps ahead into registers (two alternating sets,
//    branch-free loads with per-lane offsets computed once) and written to
//    LDS at the top of its step;
//  * the input transform is FP32; its column pass reads the fp16 inputs
//    straight into v_fma_mix_f32 (input2d_f23_mix);
//  * 8 waves: wave w owns tile-position row xi = w % 4 (t = 4 xi .. 4 xi + 3)
//    for output channels (w / 4) * 32 .. +31, so the column pass of A^T M A
//    runs in registers and only its 2 partial rows go through LDS;
//  * two barriers per step; every LDS buffer is refilled only after the
//    barrier that retires its readers (the reference has one barrier per
//    step and a write-after-read race, src/winograd_2x3_fused.cpp:60-284,:195);
//  * blocks are renumbered so each XCD works on contiguous regions with all
//    their output-channel blocks back to back (the region's input is read
//    from HBM once per XCD L2).
#include "wino_internal.h"
#include "wino_math.h"

namespace wino {

constexpr int kFusedPB = 32;   // tiles per block
constexpr int kFusedKB = 64;   // output channels per block
constexpr int kFusedThreads = 512;


struct FusedGeom {
  uint32_t N, C, H, W, K;
  uint32_t TH, TW, OH, OW;
  uint32_t RH, RW, NB;   // region shape (tiles)
  uint32_t nrh, nrw;     // regions per image set
  uint32_t Kp, C8p;
  uint32_t kblocks;      // Kp / 64
  uint32_t LH;           // staged rows per image (2RH + 2)
  uint32_t LWh;          // staged row stride in halves
  uint32_t CU;           // load units per row (dwords if even W, else halves)
  uint32_t units;        // load units per channel (NB * LH * CU)
  uint64_t CHW;          // C * H * W
  uint64_t HW;           // H * W
  uint32_t ubytes;       // bytes of U (16 x Cp x Kp fp16, < 2^31)
  float one;             // 1.0f (see input2d_f23_mix)
};

// v = B^T d B of F(2x3) from fp16 inputs, in FP32: the column pass is
// fma(d_i, 1, +-d_j) on the fp16 values (v_fma_mix_f32 reads them straight
// from their packed registers, no separate conversion; one FP32 rounding,
// the same value as converting first and subtracting), the row pass plain
// FP32 adds.  Same result as input2d<2> on the converted inputs.
// `one` is 1.0 from a kernel argument: a literal 1 would let the compiler
// turn fma(x, 1, y) back into a conversion and an add.
__device__ inline void input2d_f23_mix(const half_t* d, float one, float* v) {
  float t[16];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float d0 = (float)d[0 * 4 + j], d1 = (float)d[1 * 4 + j], d2 = (float)d[2 * 4 + j],
                d3 = (float)d[3 * 4 + j];
    t[0 * 4 + j] = __builtin_fmaf(d0, one, -d2);
    t[1 * 4 + j] = __builtin_fmaf(d1, one, d2);
    t[2 * 4 + j] = __builtin_fmaf(d2, one, -d1);
    t[3 * 4 + j] = __builtin_fmaf(d1, one, -d3);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) Tf<2>::BT(t + 4 * a, v + 4 * a);
}

// EVENW: W even, so every staged row starts on a dword and is loaded as
// dwords (x from 2*tw0 - 2); otherwise as halves (x from 2*tw0 - 1).
template <bool EVENW>
__global__ __launch_bounds__(kFusedThreads) void k_fused_f23(const half_t* __restrict__ X,
                                                             const half8* __restrict__ U,
                                                             half_t* __restrict__ Y, FusedGeom g) {
  constexpr int PB = kFusedPB, KB = kFusedKB;
  constexpr int MAXP = EVENW ? 3 : 4;    // load passes of 64 units per channel
  // Channel slot stride in halves.  Two channels apart is 16 banks (mod 64)
  // apart, so the 4 channel pairs a wave's lanes read in the same instruction
  // spread over the banks.
  constexpr int CH_H = EVENW ? 400 : 272;
  extern __shared__ __attribute__((aligned(16))) char smem[];
  half8* Vl = (half8*)smem;                            // 2 x [16 t][4 grp][PB] = 64 KiB
  half_t* xs = (half_t*)(smem + 2 * 16 * 4 * PB * 16);  // [32 ch][CH_H]
  float* Rl = (float*)smem;                            // [4 xi][2 j][KB][PB] after loop

  const int tid = threadIdx.x, lane = tid & 63;
  // wave-uniform by construction; readfirstlane lets the compiler keep it
  // (and the per-channel buffer resources built from it) in SGPRs
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int grp = lane >> 4, r16 = lane & 15;
  const int xi = wave & 3, kh = wave >> 2;

  // XCD-contiguous renumbering (block b is dispatched to XCD b % 8)
  const uint32_t nb = gridDim.x, q8 = nb / 8, r8 = nb % 8;
  const uint32_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const uint32_t wk = xcd * q8 + (xcd < r8 ? xcd : r8) + slot;
  const uint32_t kb = wk % g.kblocks;
  uint32_t region = wk / g.kblocks;
  const uint32_t rw = region % g.nrw;
  region /= g.nrw;
  const uint32_t rh = region % g.nrh;
  const uint32_t n0 = (region / g.nrh) * g.NB, th0 = rh * g.RH, tw0 = rw * g.RW;
  const int y0 = 2 * (int)th0 - 1;
  const int x0 = EVENW ? 2 * (int)tw0 - 2 : 2 * (int)tw0 - 1;

  // ---- staging: unit u of a channel (row stride CU units) lands at LDS
  // unit u of that channel's slot.  Loads go through a buffer resource whose
  // base is the channel's plane (wave-uniform, in SGPRs) and whose range ends
  // at the end of the input: a unit outside the image gets an out-of-range
  // byte offset and the hardware returns 0 for it, as it does for every unit
  // of a channel >= C (empty range).  No per-load address arithmetic.
  int voff[MAXP];
#pragma unroll
  for (int j = 0; j < MAXP; ++j) {
    const int u = lane + 64 * j;
    voff[j] = 0x7ffffff0;
    if (u < (int)g.units) {
      const int cu = u % (int)g.CU;
      const int rr = u / (int)g.CU;  // row over NB*LH
      const int bi = rr / (int)g.LH, r = rr % (int)g.LH;
      const int y = y0 + r;
      const int x = x0 + (EVENW ? 2 : 1) * cu;
      const uint32_t n = n0 + bi;
      if (n < g.N && y >= 0 && y < (int)g.H && x >= 0 && x < (int)g.W)
        voff[j] = 2 * (int)((uint64_t)n * g.CHW + (uint64_t)y * g.W + x);
    }
  }
  const uint64_t total_b = 2ull * g.N * g.CHW;
  uint32_t pf[4][MAXP];
  auto prefetch = [&](int s) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // branch-free (scalar selects): a channel >= C gets base X, range 0
      const uint32_t c = 32u * s + 4u * wave + i;
      const bool in = c < g.C;
      const uint64_t cb = 2ull * (in ? c : 0u) * g.HW;
      const uint32_t range = in ? (uint32_t)(total_b - cb) : 0u;
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)((const char*)X + cb), 0, (int)range, 0x00020000);
#pragma unroll
      for (int j = 0; j < MAXP; ++j) {
        if constexpr (EVENW)
          pf[i][j] = __builtin_amdgcn_raw_buffer_load_b32(rs, voff[j], 0, 0);
        else
          pf[i][j] = __builtin_amdgcn_raw_buffer_load_b16(rs, voff[j], 0, 0);
      }
    }
  };
  // All 64 * MAXP units are written (the slot has room; units past the
  // region hold zeros), so staging is branch-free.
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      half_t* dst = xs + (4 * wave + i) * CH_H;
#pragma unroll
      for (int j = 0; j < MAXP; ++j) {
        if constexpr (EVENW)
          ((uint32_t*)dst)[lane + 64 * j] = pf[i][j];
        else
          dst[lane + 64 * j] = __builtin_bit_cast(half_t, (uint16_t)pf[i][j]);
      }
    }
  };

  // U fragments of this wave for one step: t = 4 xi + tl, k rows kh*32 + kt*16 + r16
  typedef half8 ufrag_t[4][2];
  half8 ua[4][2], ub[4][2];
  // U through one buffer resource: a per-lane byte offset (8-channel
  // group, output channel), the tile position and step in the wave-uniform
  // soffset, the 16-row half as an immediate.  U is below 2 GiB (checked by
  // the planner).
  const __amdgpu_buffer_rsrc_t urs =
      __builtin_amdgcn_make_buffer_rsrc((void*)U, 0, (int)g.ubytes, 0x00020000);
  const uint32_t uoff = ((uint32_t)grp * g.Kp + kb * KB + kh * 32 + r16) * 16u;
  auto uload = [&](int s, ufrag_t& dst) {
#pragma unroll
    for (int tl = 0; tl < 4; ++tl) {
      const uint32_t so = __builtin_amdgcn_readfirstlane(
          (((uint32_t)(4 * xi + tl) * g.C8p + 4u * (uint32_t)s) * g.Kp) * 16u);
#pragma unroll
      for (int kt = 0; kt < 2; ++kt)
        dst[tl][kt] = __builtin_bit_cast(
            half8, __builtin_amdgcn_raw_buffer_load_b128(urs, uoff + kt * 256u, so, 0));
    }
  };

  // transform phase: thread -> (tile it, channel pair cp = 4 * cgrp + q);
  // q runs fastest across lanes, so the 4 half2 V writes of one 16-byte
  // entry come from consecutive lanes (conflict-free).
  const int q = tid & 3, it = (tid >> 2) % PB, cgrp = tid / (4 * PB);
  const int cp = 4 * cgrp + q;  // channels 2cp, 2cp+1
  const int rsz = (int)(g.RH * g.RW);
  const int ibi = it / rsz, ilr = (it % rsz) / (int)g.RW, ilc = it % (int)g.RW;
  const bool tvalid = it < (int)g.NB * rsz && n0 + ibi < g.N && th0 + ilr < g.TH &&
                      tw0 + ilc < g.TW;
  // Dword holding the patch's first half (the high half when W is even).
  // A slot past the region's tiles transforms an arbitrary patch of its
  // channel: its V values only reach MFMA output columns of tiles that are
  // never stored, so the transform needs no branch.
  const int src0 = tvalid ? ((ibi * (int)g.LH + 2 * ilr) * (int)g.LWh) / 2 + ilc : 0;
  const int lw2 = (int)g.LWh >> 1;

  auto transform = [&](half8* Vout) {
    half2_t vo[16];
    {
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const uint32_t* src = (const uint32_t*)(xs + (2 * cp + cc) * CH_H) + src0;
        half_t d[16];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const uint32_t w0 = src[a * lw2], w1 = src[a * lw2 + 1];
          if constexpr (EVENW) {
            const uint32_t w2 = src[a * lw2 + 2];
            d[a * 4 + 0] = __builtin_bit_cast(half_t, (uint16_t)(w0 >> 16));
            d[a * 4 + 1] = __builtin_bit_cast(half_t, (uint16_t)w1);
            d[a * 4 + 2] = __builtin_bit_cast(half_t, (uint16_t)(w1 >> 16));
            d[a * 4 + 3] = __builtin_bit_cast(half_t, (uint16_t)w2);
          } else {
            d[a * 4 + 0] = __builtin_bit_cast(half_t, (uint16_t)w0);
            d[a * 4 + 1] = __builtin_bit_cast(half_t, (uint16_t)(w0 >> 16));
            d[a * 4 + 2] = __builtin_bit_cast(half_t, (uint16_t)w1);
            d[a * 4 + 3] = __builtin_bit_cast(half_t, (uint16_t)(w1 >> 16));
          }
        }
        float v[16];
        input2d_f23_mix(d, g.one, v);
#pragma unroll
        for (int t = 0; t < 16; ++t) vo[t][cc] = (half_t)v[t];
      }
    }
    // V entry (t, 8-channel group cgrp, tile it), halves 2q .. 2q+1
    half2_t* V2 = (half2_t*)Vout;
#pragma unroll
    for (int t = 0; t < 16; ++t) V2[((t * 4 + cgrp) * PB + it) * 4 + q] = vo[t];
  };

  floatx4 acc[4][2][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[a][b][c] = floatx4{0.f, 0.f, 0.f, 0.f};

  auto mfma_step = [&](const half8* Vin, const ufrag_t& uf) {
#pragma unroll
    for (int tl = 0; tl < 4; ++tl) {
      const int t = 4 * xi + tl;
      half8 b[2];
#pragma unroll
      for (int pt = 0; pt < 2; ++pt) b[pt] = Vin[(t * 4 + grp) * PB + pt * 16 + r16];
#pragma unroll
      for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int pt = 0; pt < 2; ++pt)
          acc[tl][kt][pt] =
              __builtin_amdgcn_mfma_f32_16x16x32_f16(uf[tl][kt], b[pt], acc[tl][kt][pt], 0, 0, 0);
    }
  };

  // Software pipeline over 32-channel steps.  Iteration s runs the MFMAs of
  // step s (V buffer s&1, U registers of step s) and the input transform of
  // step s+1 (into V buffer (s+1)&1); they are independent and sit in one
  // branch-free basic block, so their instructions interleave and the VALU
  // and the matrix cores work at the same time.  X(s+2) and U(s+1) are
  // loaded during iteration s.  The loop is unrolled by two so the U
  // register sets alternate without copies; an odd step count runs one
  // extra step whose channels are >= C (zeros, U of the last step).
  const int nsteps = (int)(g.C8p / 4);
  constexpr int VB = 16 * 4 * PB;  // half8 entries per V buffer
  prefetch(0);
  uload(0, ua);
  stage();
  __syncthreads();
  prefetch(1);
  transform(Vl);
  __syncthreads();
  auto iter = [&](int s, const ufrag_t& ucur, ufrag_t& unext) {
    stage();  // X(s+1) -> xs (readers of xs finished before the last barrier)
    __syncthreads();
    uload(s + 1 < nsteps ? s + 1 : nsteps - 1, unext);
    prefetch(s + 2);
    mfma_step(Vl + (s & 1) * VB, ucur);
    transform(Vl + ((s + 1) & 1) * VB);
    __syncthreads();
  };
  for (int s = 0; s < nsteps; s += 2) {
    iter(s, ua, ub);
    iter(s + 1, ub, ua);
  }

  // column pass of A^T M A in registers: R[j] = sum_nu A^T[j][nu] M[xi][nu]
  // (the loop ended with a barrier: Rl may overwrite V and xs)
#pragma unroll
  for (int kt = 0; kt < 2; ++kt)
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float m0 = acc[0][kt][pt][e], m1 = acc[1][kt][pt][e], m2 = acc[2][kt][pt][e],
                    m3 = acc[3][kt][pt][e];
        const int k = kh * 32 + kt * 16 + 4 * grp + e, p = pt * 16 + r16;
        Rl[((xi * 2 + 0) * KB + k) * PB + p] = m0 + m1 + m2;
        Rl[((xi * 2 + 1) * KB + k) * PB + p] = m1 - m2 - m3;
      }
  __syncthreads();

  // row pass and store: Y[i][j] = sum_xi A^T[i][xi] R[xi][j]
  for (int q = tid; q < KB * PB; q += kFusedThreads) {
    const int k = q / PB, p = q % PB;
    const uint32_t kk = kb * KB + k;
    const int bi = p / rsz, lr = (p % rsz) / (int)g.RW, lc = p % (int)g.RW;
    const uint32_t n = n0 + bi, th = th0 + lr, tw = tw0 + lc;
    if (kk >= g.K || p >= (int)g.NB * rsz || n >= g.N || th >= g.TH || tw >= g.TW) continue;
    float R[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int j = 0; j < 2; ++j) R[x][j] = Rl[((x * 2 + j) * KB + k) * PB + p];
    half_t* dst = Y + (((uint64_t)n * g.K + kk) * g.OH + 2 * th) * g.OW + 2 * tw;
    const float a0 = R[0][0] + R[1][0] + R[2][0], a1 = R[0][1] + R[1][1] + R[2][1];
    const float b0 = R[1][0] - R[2][0] - R[3][0], b1 = R[1][1] - R[2][1] - R[3][1];
    const bool row1 = 2 * th + 1 < g.OH;
    if (EVENW) {  // OW even: the 2-wide output pair is always in range and aligned
      *(half2_t*)dst = half2_t{(half_t)a0, (half_t)a1};
      if (row1) *(half2_t*)(dst + g.OW) = half2_t{(half_t)b0, (half_t)b1};
    } else {
      const bool col1 = 2 * tw + 1 < g.OW;
      dst[0] = (half_t)a0;
      if (col1) dst[1] = (half_t)a1;
      if (row1) {
        dst[g.OW] = (half_t)b0;
        if (col1) dst[g.OW + 1] = (half_t)b1;
      }
    }
  }
}

// Region of <= PB tiles maximising useful tile slots whose staged image
// (NB x (2RH+2) rows x CU units) fits one channel slot of the ring.
static bool fused_region(uint32_t N, uint32_t TH, uint32_t TW, bool evenw, uint32_t& RH,
                         uint32_t& RW, uint32_t& NB) {
  const uint32_t maxunits = evenw ? 192u : 256u;  // 64 x MAXP load units
  double best = -1.0;
  RH = RW = NB = 0;
  for (uint32_t rw = 1; rw <= TW && rw <= (uint32_t)kFusedPB; ++rw)
    for (uint32_t rh = 1; rh <= TH && rh * rw <= (uint32_t)kFusedPB; ++rh) {
      const bool full = rh == TH && rw == TW;
      const uint32_t nbmax = full ? kFusedPB / (rh * rw) : 1;
      for (uint32_t nbv = 1; nbv <= nbmax && nbv <= N; ++nbv) {
        const uint32_t cu = evenw ? rw + 2 : 2 * rw + 2;
        if (nbv * (2 * rh + 2) * cu > maxunits) continue;
        const double blocks = (double)((N + nbv - 1) / nbv) * ((TH + rh - 1) / rh) *
                              ((TW + rw - 1) / rw);
        const double util = (double)N * TH * TW / (blocks * kFusedPB);
        const double halo = (double)(nbv * (2 * rh + 2) * (2 * rw + 2)) / (nbv * rh * rw * 4.0);
        const double score = util - 1e-3 * halo;
        if (score > best) {
          best = score;
          RH = rh;
          RW = rw;
          NB = nbv;
        }
      }
    }
  return RH > 0;
}

uint64_t fused_zero_bytes() { return 256; }

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
  const bool evenw = (W % 2) == 0;
  // 32-bit buffer offsets: input below the out-of-range sentinel, U below 2 GiB
  if (2ull * N * C * H * W > (1ull << 30) || 32ull * pl.Cp * pl.Kp >= (1ull << 31))
    return hipErrorInvalidValue;
  if (!fused_region(N, g.TH, g.TW, evenw, g.RH, g.RW, g.NB)) return hipErrorInvalidValue;
  g.nrh = (g.TH + g.RH - 1) / g.RH;
  g.nrw = (g.TW + g.RW - 1) / g.RW;
  g.Kp = pl.Kp;
  g.C8p = pl.Cp / 8;
  g.kblocks = pl.Kp / kFusedKB;
  g.LH = 2 * g.RH + 2;
  g.CU = evenw ? g.RW + 2 : 2 * g.RW + 2;
  g.LWh = evenw ? 2 * g.CU : g.CU;
  g.units = g.NB * g.LH * g.CU;
  g.CHW = (uint64_t)C * H * W;
  g.HW = (uint64_t)H * W;
  g.ubytes = (uint32_t)(16ull * g.C8p * g.Kp * 16);
  g.one = 1.0f;
  const uint64_t regions = (uint64_t)((N + g.NB - 1) / g.NB) * g.nrh * g.nrw;
  const uint32_t blocks = (uint32_t)(regions * g.kblocks);
  const size_t stage_b = (size_t)2 * 16 * 4 * kFusedPB * 16 + (size_t)32 * 400 * 2;
  const size_t r_b = (size_t)4 * 2 * kFusedKB * kFusedPB * 4;
  const size_t lds = stage_b > r_b ? stage_b : r_b;
  if (evenw) {
    (void)hipFuncSetAttribute((const void*)k_fused_f23<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_fused_f23<true>, dim3(blocks), dim3(kFusedThreads), lds, st, x,
                       (const half8*)U, y, g);
  } else {
    (void)hipFuncSetAttribute((const void*)k_fused_f23<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_fused_f23<false>, dim3(blocks), dim3(kFusedThreads), lds, st, x,
                       (const half8*)U, y, g);
  }
  return hipGetLastError();
}

}  // namespace wino
