// Exposes the reference CPU direct convolution (include/verfiy.h:5-44,
// compiled unchanged from the reference tree) with C linkage so that the
// oracle tests and the CPU baseline can call it.  Test infrastructure only.
#include "verfiy.h"

extern "C" void ref_conv2dcpu(_Float16* pin, _Float16* pwei, _Float16* pout, int n, int c, int h,
                              int w, int k, int r, int s, int u, int v, int p, int q) {
  conv2dcpu(pin, pwei, pout, n, c, h, w, k, r, s, u, v, p, q);
}
