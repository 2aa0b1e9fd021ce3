/*
 * conv2d.h -- drop-in operator header for the fp16 NCHW 3x3 conv2d path.
 *
 * Same names, argument order, struct layouts and return convention as the
 * reference operator API (include/conv2d.h:8-42 of DCU-Winograd-Conv2d):
 *   problem_t      include/conv2d.h:8-24   (72 bytes on x86-64)
 *   kernelInfo_t   include/conv2d.h:26-36  (40 bytes on x86-64)
 *   getParamsize   include/conv2d.h:39     (src/conv2d.cpp:16-21)
 *   getkernelInfo  include/conv2d.h:40     (src/conv2d.cpp:24-116)
 *   conv2d_fp16    include/conv2d.h:42     (src/conv2d.cpp:118-133)
 * All three entry points have C linkage here so that C, ctypes and
 * recompiled C++ harnesses bind the same symbols.
 *
 * Usage is the reference harness's (src/main.cpp:78-105): the caller fills a
 * problem_t with device pointers and shape, calls getParamsize, mallocs a
 * parameter blob of that size, calls getkernelInfo (which selects the
 * algorithm and allocates the device workspace), then calls conv2d_fp16 any
 * number of times.  Release the workspace with wino_param_destroy()
 * (winoconv.h) before freeing the blob.
 */
#ifndef WINO_CONV2D_FP16_FWD_HEADER
#define WINO_CONV2D_FP16_FWD_HEADER

#define __in__
#define __out__
#define __in_out__

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  _Float16* in;     /* input,  NCHW [n][c][h][w], device memory   */
  _Float16* weight; /* filter, KCRS [k][c][r][s], device memory   */
  _Float16* out;    /* output, NKHW [n][k][oh][ow], device memory */
  unsigned int n;   /* batch size                                 */
  unsigned int c;   /* input channels                             */
  unsigned int h;   /* input height                               */
  unsigned int w;   /* input width                                */
  unsigned int k;   /* output channels                            */
  unsigned int r;   /* filter height (3)                          */
  unsigned int s;   /* filter width (3)                           */
  unsigned int u;   /* vertical stride (1)                        */
  unsigned int v;   /* horizontal stride (1)                      */
  unsigned int p;   /* vertical padding (1)                       */
  unsigned int q;   /* horizontal padding (1)                     */
} problem_t;

typedef struct {
  unsigned int blockx;
  unsigned int blocky;
  unsigned int blockz;
  unsigned int threadx;
  unsigned int thready;
  unsigned int threadz;
  unsigned int dynmicLdsSize;
  void* kernelPtr;
} kernelInfo_t;

/* Writes the size of the opaque parameter blob; returns 0. */
int getParamsize(__in__ problem_t* problem, __out__ int* paramSize);

/* Validates the problem, selects the algorithm with the reference rule
 * (src/conv2d.cpp:83-114), plans tiles and allocates the device workspace
 * into the caller's blob.  Returns 0 on success, a WINO_ERR_* code from
 * winoconv.h otherwise (the reference always returned 0).  kernelInfo is
 * filled with the main kernel's launch shape (informational; the reference
 * left it untouched) and may be NULL. */
int getkernelInfo(__in__ problem_t* problem, __out__ kernelInfo_t* kernelInfo,
                  __in_out__ void* param);

/* Enqueues the convolution on the blob's stream (the null stream unless
 * wino_param_set_stream was called); asynchronous like the reference.
 * Errors are recorded in the blob (wino_param_status) and printed to
 * stderr instead of terminating the process. */
void conv2d_fp16(const void* param_ptr);

#ifdef __cplusplus
}
#endif
#endif
