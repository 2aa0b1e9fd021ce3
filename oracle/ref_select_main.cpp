// Runs the reference selector (src/conv2d.cpp:24-114, compiled unchanged from
// the reference tree) for one problem given as the reference harness's 11
// positional integers (src/main.cpp:9-19) and prints the chosen algorithm
// (include/common.h:140-144).  Each call is a fresh process, so the
// reference's process-global tile sizes (src/conv2d.cpp:3-8) start at their
// F(4x3) values.  The workspace hipMalloc fails harmlessly without a GPU.
// Test infrastructure only.
#include <cstdio>
#include <cstdlib>
#include "conv2d.h"
#define FLT_H 3L
#define FLT_W 3L
#define TILE_IN_H 6L
#define TILE_IN_W 6L
#define TILE_OUT_H 4L
#define TILE_OUT_W 4L
#include "common.h"

int main(int argc, char** argv) {
  if (argc < 12) return 2;
  problem_t pr = {};
  unsigned int* f[11] = {&pr.n, &pr.c, &pr.h, &pr.w, &pr.k, &pr.r, &pr.s, &pr.u, &pr.v, &pr.p, &pr.q};
  for (int i = 0; i < 11; ++i) *f[i] = (unsigned int)atoi(argv[i + 1]);
  int sz = 0;
  getParamsize(&pr, &sz);
  void* param = calloc(1, sz);
  kernelInfo_t ki;
  getkernelInfo(&pr, &ki, param);
  printf("%d %d\n", (int)((mykernelParamType*)param)->kernel, sz);
  return 0;
}
