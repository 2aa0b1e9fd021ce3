// checker.hip -- error statement against an exact FP64 direct convolution,
// computed on the GPU for whole tensors (no operator path calls it).
//
// For a finished output y (fp16, NKHW) of a problem (any filter size,
// stride, padding), every output element is recomputed as an exact FP64
// correlation of the fp16 input and filter (products and sums in double,
// zero padding, the definition include/verfiy.h:5-44 states with its own
// fp16 product rounding removed) and compared with y.  Each block reduces
// sum (y - ref)^2, sum ref^2 and max |y - ref| over its elements into its
// own slot; the host adds the slots in a fixed order, so the statement is
// deterministic and no atomics are used.
#include <stdio.h>

#include <vector>

#include "wino_internal.h"
#include "wino_math.h"

namespace wino {

struct ErrPartial {
  double d2, r2, maxabs, pad;
};

__global__ __launch_bounds__(256) void k_conv_error_f64(const half_t* __restrict__ X,
                                                        const half_t* __restrict__ Wt,
                                                        const half_t* __restrict__ Y,
                                                        DirectGeom g, uint64_t per_block,
                                                        ErrPartial* __restrict__ out) {
  __shared__ double s_d2[256], s_r2[256], s_mx[256];
  const uint64_t total = g.NPIX * g.K;
  const uint64_t b0 = (uint64_t)blockIdx.x * per_block;
  const uint64_t b1 = b0 + per_block < total ? b0 + per_block : total;
  const uint32_t ohw = g.OH * g.OW;
  double d2 = 0.0, r2 = 0.0, mx = 0.0;
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += 256) {
    // i indexes NKHW
    const uint32_t ox = (uint32_t)(i % g.OW);
    const uint32_t oy = (uint32_t)((i / g.OW) % g.OH);
    const uint32_t k = (uint32_t)((i / ohw) % g.K);
    const uint32_t n = (uint32_t)(i / ((uint64_t)ohw * g.K));
    double acc = 0.0;
    for (uint32_t c = 0; c < g.C; ++c) {
      const half_t* xp = X + ((uint64_t)n * g.C + c) * g.H * g.W;
      const half_t* wp = Wt + ((uint64_t)k * g.C + c) * g.RS;
      for (uint32_t kh = 0; kh < g.R; ++kh) {
        const int y = (int)(oy * g.U + kh) - (int)g.P;
        if (y < 0 || y >= (int)g.H) continue;
        for (uint32_t kw = 0; kw < g.S; ++kw) {
          const int x = (int)(ox * g.V + kw) - (int)g.Q;
          if (x < 0 || x >= (int)g.W) continue;
          acc += (double)xp[(uint64_t)y * g.W + x] * (double)wp[kh * g.S + kw];
        }
      }
    }
    const double d = (double)Y[i] - acc;
    d2 += d * d;
    r2 += acc * acc;
    mx = fmax(mx, fabs(d));
  }
  s_d2[threadIdx.x] = d2;
  s_r2[threadIdx.x] = r2;
  s_mx[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_d2[threadIdx.x] += s_d2[threadIdx.x + w];
      s_r2[threadIdx.x] += s_r2[threadIdx.x + w];
      s_mx[threadIdx.x] = fmax(s_mx[threadIdx.x], s_mx[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = ErrPartial{s_d2[0], s_r2[0], s_mx[0], 0.0};
}

}  // namespace wino

using namespace wino;

extern "C" int wino_conv_error_f64(const problem_t* pr, double* max_abs, double* rel_l2) {
  if (!pr || !pr->in || !pr->weight || !pr->out || !max_abs || !rel_l2) return WINO_ERR_INVALID;
  int st = validate(pr);
  if (st) return st;
  const DirectGeom g = direct_geom(pr->n, pr->c, pr->h, pr->w, pr->k, pr->r, pr->s, pr->u,
                                   pr->v, pr->p, pr->q);
  const uint64_t total = g.NPIX * g.K;
  const uint32_t blocks = 4096;
  const uint64_t per_block = (total + blocks - 1) / blocks;
  ErrPartial* dev = nullptr;
  if (hipMalloc(&dev, sizeof(ErrPartial) * blocks) != hipSuccess) {
    (void)hipGetLastError();
    return WINO_ERR_ALLOC;
  }
  hipLaunchKernelGGL(k_conv_error_f64, dim3(blocks), dim3(256), 0, nullptr,
                     (const half_t*)pr->in, (const half_t*)pr->weight, (const half_t*)pr->out, g,
                     per_block, dev);
  std::vector<ErrPartial> host(blocks);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpy(host.data(), dev, sizeof(ErrPartial) * blocks, hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  if (e != hipSuccess) {
    fprintf(stderr, "winoconv: wino_conv_error_f64: %s\n", hipGetErrorString(e));
    return WINO_ERR_LAUNCH;
  }
  double d2 = 0.0, r2 = 0.0, mx = 0.0;
  for (const ErrPartial& p : host) {
    d2 += p.d2;
    r2 += p.r2;
    mx = mx > p.maxabs ? mx : p.maxabs;
  }
  *max_abs = mx;
  *rel_l2 = r2 > 0.0 ? sqrt(d2 / r2) : sqrt(d2);
  return WINO_OK;
}
