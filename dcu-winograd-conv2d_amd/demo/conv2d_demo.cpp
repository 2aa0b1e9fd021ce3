// conv2d_demo.cpp -- command-line harness with the reference's semantics
// (src/main.cpp:7-134): the same 11 positional arguments
//
//     conv2dfp16demo n c h w k r s u v p q [--algo auto|f43|f23|f23_fused|direct]
//                    [--precision fp16|fp8|fp8x2] [--iters N] [--no-verify]
//
// the same data (rand() % 255 / 255 for input and filter, src/main.cpp:35-43),
// the same flow through getParamsize / getkernelInfo / conv2d_fp16 (one
// warm-up call, then N timed calls between two events on the null stream,
// src/main.cpp:74-110) and the same check: a CPU direct correlation whose
// fp16 x fp16 products are rounded to fp16 and summed in double
// (include/verfiy.h:5-44), and an element fails when
// |gpu - cpu| / gpu > 0.01 or gpu is NaN/Inf (src/main.cpp:118-129).
//
// Differences: problems other than 3x3 stride 1 pad 1 run on the direct
// (implicit-GEMM) fallback instead of producing garbage; a problem the
// operator rejects is reported and exits with status 2; the verifier runs on
// all host cores (the reference used one); the workspace is released with
// wino_param_destroy.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "conv2d.h"
#include "winoconv.h"

static void usage() {
  fprintf(stderr,
          "usage: conv2dfp16demo n c h w k r s u v p q [--algo auto|f43|f23|f23_fused|direct]\n"
          "                      [--precision fp16|fp8|fp8x2] [--iters N] [--no-verify]\n");
}

#define HIP_OK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
              __LINE__);                                                            \
      exit(1);                                                                      \
    }                                                                               \
  } while (0)

// Direct correlation with zero padding, loop order n, k, oh, ow, c, r, s;
// each product rounded to fp16, sums in double, result stored as fp16.
static void conv_cpu(const _Float16* in, const _Float16* wt, _Float16* out, int n, int c, int h,
                     int w, int k, int r, int s, int u, int v, int p, int q) {
  const int oh = (h - r + 2 * p) / u + 1, ow = (w - s + 2 * q) / v + 1;
  const int jobs = n * k;
  unsigned nt = std::max(1u, std::thread::hardware_concurrency());
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t)
    pool.emplace_back([=]() {
      for (int job = (int)t; job < jobs; job += (int)nt) {
        const int b = job / k, kk = job % k;
        for (int i = 0; i < oh; ++i)
          for (int j = 0; j < ow; ++j) {
            double sum = 0.0;
            for (int cc = 0; cc < c; ++cc)
              for (int a = 0; a < r; ++a)
                for (int e = 0; e < s; ++e) {
                  const int y = i * u - p + a, x = j * v - q + e;
                  if (y < 0 || y >= h || x < 0 || x >= w) continue;
                  const float prod = (float)in[(((size_t)b * c + cc) * h + y) * w + x] *
                                     (float)wt[(((size_t)kk * c + cc) * r + a) * s + e];
                  sum += (double)(_Float16)prod;
                }
            out[(((size_t)b * k + kk) * oh + i) * ow + j] = (_Float16)sum;
          }
      }
    });
  for (auto& th : pool) th.join();
}

int main(int argc, char** argv) {
  if (argc < 12) {
    usage();
    return 1;
  }
  int dims[11];
  for (int i = 0; i < 11; ++i) dims[i] = atoi(argv[i + 1]);
  const int n = dims[0], c = dims[1], h = dims[2], w = dims[3], k = dims[4], r = dims[5],
            s = dims[6], u = dims[7], v = dims[8], p = dims[9], q = dims[10];
  int algo = WINO_ALGO_AUTO, precision = WINO_PREC_FP16, iters = 100;
  bool verify = true;
  for (int i = 12; i < argc; ++i) {
    if (!strcmp(argv[i], "--no-verify")) {
      verify = false;
    } else if (!strcmp(argv[i], "--iters") && i + 1 < argc) {
      iters = std::max(1, atoi(argv[++i]));
    } else if (!strcmp(argv[i], "--algo") && i + 1 < argc) {
      const char* a = argv[++i];
      algo = !strcmp(a, "f43") ? WINO_ALGO_F43
             : !strcmp(a, "f23") ? WINO_ALGO_F23
             : !strcmp(a, "f23_fused") ? WINO_ALGO_F23_FUSED
             : !strcmp(a, "direct") ? WINO_ALGO_DIRECT
                                    : WINO_ALGO_AUTO;
    } else if (!strcmp(argv[i], "--precision") && i + 1 < argc) {
      const char* a = argv[++i];
      precision = !strcmp(a, "fp8") ? WINO_PREC_FP8 : !strcmp(a, "fp8x2") ? WINO_PREC_FP8X2
                                                                            : WINO_PREC_FP16;
    } else {
      usage();
      return 1;
    }
  }
  if (n <= 0 || c <= 0 || h <= 0 || w <= 0 || k <= 0 || r <= 0 || s <= 0 || u <= 0 || v <= 0 ||
      p < 0 || q < 0 || h + 2 * p < r || w + 2 * q < s) {
    fprintf(stderr, "invalid problem\n");
    return 1;
  }
  const int outh = (h - r + 2 * p) / u + 1, outw = (w - s + 2 * q) / v + 1;
  const size_t nin = (size_t)n * c * h * w, nwt = (size_t)k * c * r * s,
               nout = (size_t)n * k * outh * outw;

  std::vector<_Float16> in(nin), wt(nwt), out_cpu(nout, (_Float16)0), out_gpu(nout, (_Float16)0);
  for (size_t i = 0; i < nin; ++i) in[i] = (_Float16)((rand() % 255) / 255.0);
  for (size_t i = 0; i < nwt; ++i) wt[i] = (_Float16)((rand() % 255) / 255.0);

  _Float16 *d_in, *d_wt, *d_out;
  HIP_OK(hipMalloc((void**)&d_in, nin * sizeof(_Float16)));
  HIP_OK(hipMalloc((void**)&d_wt, nwt * sizeof(_Float16)));
  HIP_OK(hipMalloc((void**)&d_out, nout * sizeof(_Float16)));
  HIP_OK(hipMemcpy(d_in, in.data(), nin * sizeof(_Float16), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_wt, wt.data(), nwt * sizeof(_Float16), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d_out, out_gpu.data(), nout * sizeof(_Float16), hipMemcpyHostToDevice));

  problem_t problem = {d_in, d_wt, d_out, (unsigned)n, (unsigned)c, (unsigned)h, (unsigned)w,
                       (unsigned)k, (unsigned)r, (unsigned)s, (unsigned)u, (unsigned)v,
                       (unsigned)p, (unsigned)q};
  int paramSize = 0;
  getParamsize(&problem, &paramSize);
  printf("paramsize:%d\n", paramSize);
  void* param = malloc((size_t)paramSize);
  kernelInfo_t kernelInfo;
  int st = getkernelInfo(&problem, &kernelInfo, param);
  if (st == WINO_OK && algo != WINO_ALGO_AUTO) st = wino_param_set_algo(param, algo);
  if (st == WINO_OK && precision != WINO_PREC_FP16) st = wino_param_set_precision(param, precision);
  if (st != WINO_OK) {
    printf("unsupported problem: %s\n", wino_status_string(st));
    wino_param_destroy(param);
    free(param);
    return 2;
  }
  wino_info_t info;
  wino_param_info(param, &info);
  static const char* names[] = {"F(4x3) non-fused", "F(2x3) non-fused", "F(2x3) fused",
                                "F(4x3) fused", "direct (implicit GEMM)"};
  static const char* precs[] = {"fp16", "fp8", "fp8x2"};
  printf("algorithm: %s, precision %s\n", info.algo >= 0 && info.algo < 5 ? names[info.algo] : "?",
         precs[info.precision]);

  conv2d_fp16(param);  // warm-up and result
  HIP_OK(hipMemcpy(out_gpu.data(), d_out, nout * sizeof(_Float16), hipMemcpyDeviceToHost));
  if (wino_param_status(param) != WINO_OK) {
    printf("conv2d_fp16 failed: %s\n", wino_status_string(wino_param_status(param)));
    return 1;
  }

  hipEvent_t start, stop;
  HIP_OK(hipEventCreate(&start));
  HIP_OK(hipEventCreate(&stop));
  HIP_OK(hipEventRecord(start, 0));
  for (int i = 0; i < iters; ++i) conv2d_fp16(param);
  HIP_OK(hipEventRecord(stop, 0));
  HIP_OK(hipEventSynchronize(stop));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, start, stop));
  printf("time: %f us\n", ms * 1000.0f / iters);
  const double gflop = 2.0 * n * k * outh * outw * (double)c * r * s / 1e9;
  printf("effective: %.1f TFLOP/s\n", gflop / (ms / iters * 1e-3) / 1e3);
  HIP_OK(hipEventDestroy(start));
  HIP_OK(hipEventDestroy(stop));
  wino_param_destroy(param);
  free(param);

  int error = 0;
  if (verify) {
    printf("===================start verfiy===================\n");
    conv_cpu(in.data(), wt.data(), out_cpu.data(), n, c, h, w, k, r, s, u, v, p, q);
    for (size_t i = 0; i < nout; ++i) {
      const float g = (float)out_gpu[i], cpu = (float)out_cpu[i];
      if (fabsf(g - cpu) / g > 0.01f || isnan(g) || isinf(g)) {
        printf("error, postion:%zu, gpuvalue:%f, cpuvalue:%f\n", i, g, cpu);
        ++error;
        break;
      }
    }
    printf("================finish,error:%d=========================\n", error);
  }
  HIP_OK(hipFree(d_in));
  HIP_OK(hipFree(d_wt));
  HIP_OK(hipFree(d_out));
  return error ? 3 : 0;
}
