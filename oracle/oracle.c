/*
 * oracle.c -- CPU restatement of the reference Winograd conv2d path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP operator:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The product path never calls it.
 *
 * Every function cites the reference file:line it restates.  Reference paths
 * are relative to the reference repository root.
 *
 * Contents
 *   h2d / d2h                 IEEE binary16 <-> double, round-to-nearest-even
 *   oracle_conv2d_verfiy      include/verfiy.h:5-44 (fp16-rounded products,
 *                             double sum in c->kh->kw order, fp16 store)
 *   oracle_conv2d_f64         exact FP64 direct correlation (no fp16 rounding
 *                             of products), optionally on a sample of outputs
 *   oracle_wino_*             FP64 Winograd stage model: U = G g G^T,
 *                             V = B^T d B, M[t] = U[t] V[t], Y = A^T M A with
 *                             the coefficient tables of
 *                             src/winograd_2x3_non_fused.cpp:120-160,302-377,184-227
 *                             src/winograd_4x3_non_fused.cpp:41-146,342-404,427-525
 *                             and the tile <-> (n, th, tw) mapping of
 *                             include/common.h:85-120
 *   oracle_select             src/conv2d.cpp:24-114 selector, 32-bit unsigned
 *                             arithmetic as in the reference, evaluated with
 *                             the fresh-process F(4x3) tile globals
 *   oracle_e4m3_*             OCP FP8 E4M3 (max 448, no inf) encode (RNE,
 *                             saturating) / decode
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ fp16 */

double oracle_h2d(uint16_t h) {
  int s = h >> 15, e = (h >> 10) & 0x1f, f = h & 0x3ff;
  double v;
  if (e == 0)
    v = ldexp((double)f, -24);
  else if (e == 31)
    v = f ? NAN : INFINITY;
  else
    v = ldexp((double)(f | 0x400), e - 25);
  return s ? -v : v;
}

/* Correctly rounded double -> binary16 (round to nearest, ties to even;
 * overflow to infinity), the semantics of a C cast to _Float16. */
uint16_t oracle_d2h(double x) {
  uint16_t sign = signbit(x) ? 0x8000 : 0;
  if (isnan(x)) return sign | 0x7e00;
  double a = fabs(x);
  if (a >= 65520.0) return sign | 0x7c00;
  if (a < 0x1p-14) {
    double r = nearbyint(a * 0x1p24); /* quantum 2^-24; 1024 encodes 2^-14 */
    return sign | (uint16_t)r;
  }
  int e;
  frexp(a, &e); /* a = f * 2^e, f in [0.5, 1) */
  int E = e - 1;
  double r = nearbyint(ldexp(a, 10 - E)); /* in [1024, 2048] */
  if (r >= 2048.0) {
    r = 1024.0;
    E++;
  }
  if (E > 15) return sign | 0x7c00;
  return sign | (uint16_t)(((E + 15) << 10) | ((int)r - 1024));
}

void oracle_h2d_array(const uint16_t* h, double* d, int64_t n) {
  for (int64_t i = 0; i < n; ++i) d[i] = oracle_h2d(h[i]);
}

void oracle_d2h_array(const double* d, uint16_t* h, int64_t n) {
  for (int64_t i = 0; i < n; ++i) h[i] = oracle_d2h(d[i]);
}

/* Round every element to the nearest binary16 value (kept as double). */
void oracle_round_f16(double* d, int64_t n) {
  for (int64_t i = 0; i < n; ++i) d[i] = oracle_h2d(oracle_d2h(d[i]));
}

/* ---------------------------------------------------- direct convolution */

/* Restates include/verfiy.h:5-44 (conv2dcpu).  Output size per verfiy.h:7-8.
 * Each product of two _Float16 operands is a _Float16 (rounded once; the
 * exact product of two binary16 values fits a double), it is widened to
 * double (verfiy.h:33) and summed in c -> kh -> kw order; the sum is cast
 * to _Float16 (verfiy.h:39).  Loop order n, k, i, j as verfiy.h:10-17. */
void oracle_conv2d_verfiy(const uint16_t* pin, const uint16_t* pwei, uint16_t* pout, int n, int c,
                          int h, int w, int k, int r, int s, int u, int v, int p, int q) {
  int oh = (h + 2 * p - r) / u + 1;
  int ow = (w + 2 * q - s) / v + 1;
  for (int nn = 0; nn < n; nn++)
    for (int kk = 0; kk < k; kk++)
      for (int i = 0; i < oh; i++)
        for (int j = 0; j < ow; j++) {
          double sum = 0.0;
          int posh = i * u - p;
          int posw = j * v - q;
          for (int cc = 0; cc < c; cc++)
            for (int kh = 0; kh < r; kh++)
              for (int kw = 0; kw < s; kw++) {
                int y = posh + kh, x = posw + kw;
                if (x >= 0 && y >= 0 && x < w && y < h) {
                  double a = oracle_h2d(pin[(int64_t)nn * c * h * w + (int64_t)cc * w * h + (int64_t)y * w + x]);
                  double b = oracle_h2d(pwei[(int64_t)kk * r * s * c + (int64_t)cc * r * s + kh * s + kw]);
                  sum += oracle_h2d(oracle_d2h(a * b));
                }
              }
          pout[(int64_t)nn * k * oh * ow + (int64_t)kk * oh * ow + (int64_t)i * ow + j] = oracle_d2h(sum);
        }
}

/* Exact FP64 direct correlation (products and sum in double, no rounding to
 * fp16), zero padding, general r, s, stride and padding; out is NKHW double.
 * Threads over (n, k) when built with OpenMP. */
void oracle_conv2d_f64(const uint16_t* pin, const uint16_t* pwei, double* pout, int n, int c, int h,
                       int w, int k, int r, int s, int u, int v, int p, int q) {
  int oh = (h + 2 * p - r) / u + 1;
  int ow = (w + 2 * q - s) / v + 1;
  double* wd = (double*)malloc(sizeof(double) * (size_t)k * c * r * s);
  for (int64_t i = 0; i < (int64_t)k * c * r * s; ++i) wd[i] = oracle_h2d(pwei[i]);
#pragma omp parallel
  {
    double* plane = (double*)malloc(sizeof(double) * (size_t)h * w);
#pragma omp for collapse(2) schedule(static)
    for (int nn = 0; nn < n; nn++)
      for (int kk = 0; kk < k; kk++) {
        double* o = pout + ((int64_t)nn * k + kk) * oh * ow;
        for (int64_t i = 0; i < (int64_t)oh * ow; ++i) o[i] = 0.0;
        for (int cc = 0; cc < c; cc++) {
          const uint16_t* xin = pin + ((int64_t)nn * c + cc) * h * w;
          for (int64_t i = 0; i < (int64_t)h * w; ++i) plane[i] = oracle_h2d(xin[i]);
          const double* wk = wd + ((int64_t)kk * c + cc) * r * s;
          for (int i = 0; i < oh; i++)
            for (int j = 0; j < ow; j++) {
              double acc = 0.0;
              for (int kh = 0; kh < r; kh++) {
                int y = i * u - p + kh;
                if (y < 0 || y >= h) continue;
                for (int kw = 0; kw < s; kw++) {
                  int x = j * v - q + kw;
                  if (x < 0 || x >= w) continue;
                  acc += plane[y * w + x] * wk[kh * s + kw];
                }
              }
              o[i * ow + j] += acc;
            }
        }
      }
    free(plane);
  }
  free(wd);
}

/* FP64 direct correlation at selected NKHW output indices only (the sampled
 * oracle used for large problems). */
void oracle_conv2d_f64_sample(const uint16_t* pin, const uint16_t* pwei, int n, int c, int h,
                              int w, int k, int r, int s, int u, int v, int p, int q,
                              const int64_t* idx, int64_t count, double* out) {
  int oh = (h + 2 * p - r) / u + 1;
  int ow = (w + 2 * q - s) / v + 1;
  (void)n;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < count; ++t) {
    int64_t id = idx[t];
    int j = (int)(id % ow);
    int i = (int)((id / ow) % oh);
    int kk = (int)((id / ((int64_t)ow * oh)) % k);
    int nn = (int)(id / ((int64_t)ow * oh * k));
    double acc = 0.0;
    for (int cc = 0; cc < c; cc++)
      for (int kh = 0; kh < r; kh++) {
        int y = i * u - p + kh;
        if (y < 0 || y >= h) continue;
        for (int kw = 0; kw < s; kw++) {
          int x = j * v - q + kw;
          if (x < 0 || x >= w) continue;
          acc += oracle_h2d(pin[((int64_t)nn * c + cc) * h * w + (int64_t)y * w + x]) *
                 oracle_h2d(pwei[((int64_t)kk * c + cc) * r * s + kh * s + kw]);
        }
      }
    out[t] = acc;
  }
}

/* ------------------------------------------------------ Winograd tables */

/* F(2x3): G (4x3), B^T (4x4), A^T (2x4).
 * G   : src/winograd_2x3_non_fused.cpp:120-160
 * B^T : src/winograd_2x3_non_fused.cpp:324-377
 * A^T : src/winograd_2x3_non_fused.cpp:184-227 */
static const double G2[4][3] = {{1, 0, 0}, {0.5, 0.5, 0.5}, {0.5, -0.5, 0.5}, {0, 0, 1}};
static const double BT2[4][4] = {{1, 0, -1, 0}, {0, 1, 1, 0}, {0, -1, 1, 0}, {0, 1, 0, -1}};
static const double AT2[2][4] = {{1, 1, 1, 0}, {0, 1, -1, -1}};

/* F(4x3): G (6x3), B^T (6x6), A^T (4x6).
 * G   : src/winograd_4x3_non_fused.cpp:342-404
 * B^T : src/winograd_4x3_non_fused.cpp:41-146
 * A^T : src/winograd_4x3_non_fused.cpp:427-525 */
static const double G4[6][3] = {{1.0 / 4, 0, 0},
                                {-1.0 / 6, -1.0 / 6, -1.0 / 6},
                                {-1.0 / 6, 1.0 / 6, -1.0 / 6},
                                {1.0 / 24, 1.0 / 12, 1.0 / 6},
                                {1.0 / 24, -1.0 / 12, 1.0 / 6},
                                {0, 0, 1}};
static const double BT4[6][6] = {{4, 0, -5, 0, 1, 0},  {0, -4, -4, 1, 1, 0}, {0, 4, -4, -1, 1, 0},
                                 {0, -2, -1, 2, 1, 0}, {0, 2, -1, -2, 1, 0}, {0, 4, 0, -5, 0, 1}};
static const double AT4[4][6] = {
    {1, 1, 1, 1, 1, 0}, {0, 1, -1, 2, -2, 0}, {0, 1, 1, 4, 4, 0}, {0, 1, -1, 8, -8, 1}};

static const double* tabG(int m) { return m == 2 ? &G2[0][0] : &G4[0][0]; }
static const double* tabBT(int m) { return m == 2 ? &BT2[0][0] : &BT4[0][0]; }
static const double* tabAT(int m) { return m == 2 ? &AT2[0][0] : &AT4[0][0]; }

/* Copies a coefficient table out (for tests that check the tables
 * themselves): which = 0 -> G (T x 3), 1 -> B^T (T x T), 2 -> A^T (m x T). */
int oracle_wino_table(int m, int which, double* out) {
  int T = m + 2;
  const double* src = which == 0 ? tabG(m) : which == 1 ? tabBT(m) : tabAT(m);
  int cnt = which == 0 ? T * 3 : which == 1 ? T * T : m * T;
  memcpy(out, src, sizeof(double) * cnt);
  return cnt;
}

/* U[t][k][c] = (G g G^T)[t] for every (k, c); t = xi * T + nu.
 * Layout U[T*T][K][C] as the reference (winograd_2x3_non_fused.cpp:296-299). */
void oracle_wino_filter(int m, const uint16_t* wt, int K, int C, double* U) {
  int T = m + 2;
  const double* G = tabG(m);
  int64_t KC = (int64_t)K * C;
  for (int64_t kc = 0; kc < KC; ++kc) {
    double g[9], tmp[6][3];
    for (int i = 0; i < 9; ++i) g[i] = oracle_h2d(wt[kc * 9 + i]);
    for (int a = 0; a < T; ++a)
      for (int j = 0; j < 3; ++j) {
        double acc = 0;
        for (int i = 0; i < 3; ++i) acc += G[a * 3 + i] * g[i * 3 + j];
        tmp[a][j] = acc;
      }
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b) {
        double acc = 0;
        for (int j = 0; j < 3; ++j) acc += tmp[a][j] * G[b * 3 + j];
        U[(int64_t)(a * T + b) * KC + kc] = acc;
      }
  }
}

/* Tile grid of include/common.h:85-93 for an Oh x Ow output: th = ceil(Oh/m),
 * tw = ceil(Ow/m), P = N * th * tw, tile p <-> (n, th, tw) row-major
 * (common.h:113-120). */
int64_t oracle_wino_tiles(int m, int N, int Oh, int Ow) {
  return (int64_t)N * ((Oh + m - 1) / m) * ((Ow + m - 1) / m);
}

/* V[t][c][p] = (B^T d B)[t], d the T x T window at rows th*m - pad.. and cols
 * tw*m - pad.., zero outside the image (winograd_4x3_non_fused.cpp:41-146). */
void oracle_wino_input(int m, const uint16_t* x, int N, int C, int H, int W, int pad, double* V) {
  int T = m + 2;
  const double* BT = tabBT(m);
  int Oh = H + 2 * pad - 2, Ow = W + 2 * pad - 2;
  int th_n = (Oh + m - 1) / m, tw_n = (Ow + m - 1) / m;
  int64_t P = (int64_t)N * th_n * tw_n;
  for (int c = 0; c < C; ++c)
    for (int64_t p = 0; p < P; ++p) {
      int n = (int)(p / ((int64_t)th_n * tw_n));
      int rem = (int)(p % ((int64_t)th_n * tw_n));
      int th = rem / tw_n, tw = rem % tw_n;
      double d[6][6], tmp[6][6];
      for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) {
          int yy = th * m + i - pad, xx = tw * m + j - pad;
          d[i][j] = (yy >= 0 && yy < H && xx >= 0 && xx < W)
                        ? oracle_h2d(x[((int64_t)n * C + c) * H * W + (int64_t)yy * W + xx])
                        : 0.0;
        }
      for (int a = 0; a < T; ++a)
        for (int j = 0; j < T; ++j) {
          double acc = 0;
          for (int i = 0; i < T; ++i) acc += BT[a * T + i] * d[i][j];
          tmp[a][j] = acc;
        }
      for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) {
          double acc = 0;
          for (int j = 0; j < T; ++j) acc += tmp[a][j] * BT[b * T + j];
          V[((int64_t)(a * T + b) * C + c) * P + p] = acc;
        }
    }
}

/* M[t][k][p] = sum_c U[t][k][c] * V[t][c][p] (the batched GEMM of
 * include/hep_sgemm.h:495-612 with the operand layouts of its call sites
 * winograd_4x3_non_fused.cpp:587-597). */
void oracle_wino_gemm(int T2, const double* U, const double* V, int K, int C, int64_t P,
                      double* M) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int t = 0; t < T2; ++t)
    for (int k = 0; k < K; ++k) {
      double* mrow = M + ((int64_t)t * K + k) * P;
      for (int64_t p = 0; p < P; ++p) mrow[p] = 0.0;
      for (int c = 0; c < C; ++c) {
        double uv = U[((int64_t)t * K + k) * C + c];
        const double* vrow = V + ((int64_t)t * C + c) * P;
        for (int64_t p = 0; p < P; ++p) mrow[p] += uv * vrow[p];
      }
    }
}

/* Y = A^T M A per (k, tile), written to NKHW with the partial-tile masking of
 * winograd_4x3_non_fused.cpp:513-525. */
void oracle_wino_output(int m, const double* M, int N, int K, int Oh, int Ow, double* Y) {
  int T = m + 2;
  const double* AT = tabAT(m);
  int th_n = (Oh + m - 1) / m, tw_n = (Ow + m - 1) / m;
  int64_t P = (int64_t)N * th_n * tw_n;
  for (int k = 0; k < K; ++k)
    for (int64_t p = 0; p < P; ++p) {
      int n = (int)(p / ((int64_t)th_n * tw_n));
      int rem = (int)(p % ((int64_t)th_n * tw_n));
      int th = rem / tw_n, tw = rem % tw_n;
      double mm[6][6], tmp[4][6];
      for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) mm[a][b] = M[((int64_t)(a * T + b) * K + k) * P + p];
      for (int i = 0; i < m; ++i)
        for (int b = 0; b < T; ++b) {
          double acc = 0;
          for (int a = 0; a < T; ++a) acc += AT[i * T + a] * mm[a][b];
          tmp[i][b] = acc;
        }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          double acc = 0;
          for (int b = 0; b < T; ++b) acc += tmp[i][b] * AT[j * T + b];
          int yy = th * m + i, xx = tw * m + j;
          if (yy < Oh && xx < Ow) Y[(((int64_t)n * K + k) * Oh + yy) * Ow + xx] = acc;
        }
    }
}

/* --------------------------------------------------------- FP8 E4M3 */

/* OCP E4M3 (bias 7, max 448, NaN = S.1111.111, no infinities). */
double oracle_e4m3_decode(uint8_t b) {
  int s = b >> 7, e = (b >> 3) & 0xf, f = b & 7;
  double v;
  if (e == 15 && f == 7) return NAN;
  if (e == 0)
    v = ldexp((double)f, -9);
  else
    v = ldexp((double)(f | 8), e - 10);
  return s ? -v : v;
}

/* Round to nearest even, saturate to +-448 (finite inputs). */
uint8_t oracle_e4m3_encode(double x) {
  uint8_t sign = signbit(x) ? 0x80 : 0;
  if (isnan(x)) return 0x7f;
  double a = fabs(x);
  if (a >= 448.0) return sign | 0x7e;
  if (a < 0x1p-6) {
    double r = nearbyint(a * 0x1p9); /* quantum 2^-9; 8 encodes 2^-6 */
    return sign | (uint8_t)r;
  }
  int e;
  frexp(a, &e);
  int E = e - 1;
  double r = nearbyint(ldexp(a, 3 - E)); /* in [8, 16] */
  if (r >= 16.0) {
    r = 8.0;
    E++;
  }
  if (E > 8) return sign | 0x7e;
  return sign | (uint8_t)(((E + 7) << 3) | ((int)r - 8));
}

/* x -> scale * e4m3(x / scale) for every element, in place. */
void oracle_quant_e4m3(double* d, int64_t n, double scale) {
  for (int64_t i = 0; i < n; ++i) d[i] = scale * oracle_e4m3_decode(oracle_e4m3_encode(d[i] / scale));
}

/* --------------------------------------------------------- selector */

/* Restates src/conv2d.cpp:24-114 (getkernelInfo) for a fresh process
 * (TILE_* globals still F(4x3), conv2d.cpp:3-8).  All sizes are 32-bit
 * unsigned as in the reference (conv2d.cpp:76-82); tile counts are 64-bit
 * (include/common.h:85-93) and truncated on assignment.
 * algo: 0 = F(4x3) non-fused, 1 = F(2x3) non-fused, 2 = F(2x3) fused
 * (include/common.h:140-144).  ws_bytes: the hipMalloc size; tiles: P of
 * the chosen algorithm (0 for fused). */
void oracle_select(uint32_t n, uint32_t c, uint32_t h, uint32_t w, uint32_t k, uint32_t r,
                   uint32_t s, uint32_t u, uint32_t v, uint32_t p, uint32_t q, int* algo,
                   uint64_t* ws_bytes, uint64_t* tiles) {
  uint32_t outh = (h - r + 2 * p) / u + 1;
  uint32_t outw = (w - s + 2 * q) / v + 1;
  uint64_t P4 = (uint64_t)n * ((outh + 3ull) / 4) * ((outw + 3ull) / 4);
  uint32_t image_size = n * c * h * w;
  uint32_t filter_size = k * c * r * s;
  uint32_t U_size = (uint32_t)(36ull * k * c);
  uint32_t V_size = (uint32_t)(36ull * P4 * c);
  uint32_t M_size = (uint32_t)(36ull * k * P4);
  if (image_size >= 10u * filter_size) {
    *algo = 2;
    *ws_bytes = 0;
    *tiles = 0;
  } else if (U_size > 4u * V_size) {
    uint64_t P2 = (uint64_t)n * ((outh + 1ull) / 2) * ((outw + 1ull) / 2);
    uint32_t U2 = (uint32_t)(16ull * k * c);
    uint32_t V2 = (uint32_t)(16ull * P2 * c);
    uint32_t M2 = (uint32_t)(16ull * k * P2);
    *algo = 1;
    *ws_bytes = (uint32_t)(2u * (U2 + V2 + M2));
    *tiles = P2;
  } else {
    *algo = 0;
    *ws_bytes = (uint32_t)(2u * (U_size + V_size + M_size));
    *tiles = P4;
  }
}

/* verfiy.h:5-44 semantics (fp16-rounded products, double sum in c->kh->kw
 * order, fp16 result) at selected NKHW output indices only. */
void oracle_conv2d_verfiy_sample(const uint16_t* pin, const uint16_t* pwei, int n, int c, int h,
                                 int w, int k, int r, int s, int u, int v, int p, int q,
                                 const int64_t* idx, int64_t count, uint16_t* out) {
  int oh = (h + 2 * p - r) / u + 1;
  int ow = (w + 2 * q - s) / v + 1;
  (void)n;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < count; ++t) {
    int64_t id = idx[t];
    int j = (int)(id % ow);
    int i = (int)((id / ow) % oh);
    int kk = (int)((id / ((int64_t)ow * oh)) % k);
    int nn = (int)(id / ((int64_t)ow * oh * k));
    double sum = 0.0;
    for (int cc = 0; cc < c; cc++)
      for (int kh = 0; kh < r; kh++)
        for (int kw = 0; kw < s; kw++) {
          int y = i * u - p + kh, x = j * v - q + kw;
          if (x >= 0 && y >= 0 && x < w && y < h) {
            double a = oracle_h2d(pin[((int64_t)nn * c + cc) * h * w + (int64_t)y * w + x]);
            double b = oracle_h2d(pwei[((int64_t)kk * c + cc) * r * s + kh * s + kw]);
            sum += oracle_h2d(oracle_d2h(a * b));
          }
        }
    out[t] = oracle_d2h(sum);
  }
}
