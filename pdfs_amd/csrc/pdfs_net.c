/*
 * Learned warm start of the FP32 FSM (no counterpart in the reference).
 *
 * Network: a 3-D conv encoder-decoder on the source mask at 1/4 resolution.
 * 4 levels of 16/32/64/64 channels, 3x3x3 convolutions with ReLU, 2x2x2
 * average pooling down, nearest x2 upsampling + skip concatenation up:
 *   e0a 16->16, e0b 16->16 = s0          (level 0, the 1/4 grid)
 *   e1a 16->32, e1b 32->32 = s1          (level 1)
 *   e2a 32->64, e2b 64->64 = s2          (level 2)
 *   e3a 64->64, e3b 64->64 = b3          (level 3)
 *   d2 [up b3 | s2] 128->64, d1 [up d2 | s1] 96->32, d0 [up d1 | s0] 48->16
 *   head 16->16 linear, FP32 out; channel 0 = scale * distance in cells at
 *   the coarse point (4z, 4y, 4x).
 * BF16 weights and activations, FP32 accumulation (MFMA), FP32 master
 * weights, gradients and Adam moments.
 *
 * Warm start: w = min(init, a * d + b * hmax), d the trilinear upsampling
 * of the prediction (times hmax). The FSM then only lowers values, so a
 * w that lies below the true solution somewhere leaves the field wrong
 * there; the FP64 residual judge detects it and the field is solved again
 * from its init (fallback).
 */
#include "pdfs.h"
#include "pdfs_internal.h"
#include "pdfs_nn.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { L_E0A, L_E0B, L_E1A, L_E1B, L_E2A, L_E2B, L_E3A, L_E3B, L_D2, L_D1, L_D0, L_HEAD, NL };
static const int CIN[NL] = {16, 16, 16, 32, 32, 64, 64, 64, 128, 96, 48, 16};
static const int COUT[NL] = {16, 16, 32, 32, 64, 64, 64, 64, 64, 32, 16, 16};

/* activations (BF16) of one forward, per buffer: level and channels */
enum {
  A_A0, A_S0, A_P0, A_A1, A_S1, A_P1, A_A2, A_S2, A_P2, A_A3, A_B3, A_C2, A_D2, A_C1, A_D1, A_C0,
  A_D0, NA
};
static const int A_LVL[NA] = {0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0};
static const int A_CH[NA] = {16, 16, 16, 32, 32, 32, 64, 64, 64, 64, 64, 128, 64, 96, 32, 48, 16};

struct pdfs_net {
  pdfs_ctx *c;
  hipStream_t st;
  long long woff[NL], boff[NL], nparam;
  float *P, *G, *M, *V; /* FP32 masters, gradients, Adam moments (device) */
  uint16_t *WB, *WT;    /* BF16 weights [27][Co][Ci] and transposes [27][Ci][Co] */
  int step;
  float scale, a, b;    /* target scale; upper-bound inflation */
  float cut;            /* predictions at or beyond cut (in h) keep the init; 0: none */
  /* workspace for (nb, D0) */
  long long cap_v0;
  int train_ws;
  uint16_t *act[NA];
  float *out;
  uint16_t *gpre, *gout;
  float *gA, *gB, *gs0, *gs1, *gs2, *loss;
  float *h_loss;
  /* warm-start scratch */
  uint16_t *mask;
  long long cap_mask;
  /* uneven-batch tables: 4 levels of nf + 1 segments, then nf + 1 fields */
  void *rag, *h_rag;
  size_t cap_rag;
};

#define HCK(call)                                                                    \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return pdfsi_fail(PDFS_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,        \
                        hipGetErrorString(e_));                                      \
  } while (0)
#define NCK(call)                                                                    \
  do {                                                                               \
    int r_ = (call);                                                                 \
    if (r_ != 0)                                                                     \
      return pdfsi_fail(r_ < 0 ? PDFS_EINVAL : PDFS_EHIP, "%s:%d %s failed (%d)",     \
                        __FILE__, __LINE__, #call, r_);                              \
  } while (0)

/* HIP call in a function that owns host allocations: record and leave via
 * its `out` label */
#define HGO(call)                                                                    \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      rc = pdfsi_fail(PDFS_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,          \
                      hipGetErrorString(e_));                                        \
      goto out;                                                                      \
    }                                                                                \
  } while (0)

static unsigned long long sm64(unsigned long long *s)
{
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double u01(unsigned long long *s) { return ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

static uint16_t f2bf_host(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static float bf2f_host(uint16_t h)
{
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static void free_ws(pdfs_net *n)
{
  for (int i = 0; i < NA; i++) {
    hipFree(n->act[i]);
    n->act[i] = NULL;
  }
  hipFree(n->out);
  hipFree(n->gpre);
  hipFree(n->gout);
  hipFree(n->gA);
  hipFree(n->gB);
  hipFree(n->gs0);
  hipFree(n->gs1);
  hipFree(n->gs2);
  n->out = NULL;
  n->gpre = n->gout = NULL;
  n->gA = n->gB = n->gs0 = n->gs1 = n->gs2 = NULL;
  n->cap_v0 = 0;
  n->train_ws = 0;
}

static int ensure_ws(pdfs_net *n, long long v0, int train)
{
  if (v0 <= n->cap_v0 && (!train || n->train_ws)) return 0;
  HCK(hipStreamSynchronize(n->st));
  if (v0 < n->cap_v0) v0 = n->cap_v0;
  free_ws(n);
  for (int i = 0; i < NA; i++) {
    size_t e = (size_t)(v0 >> (3 * A_LVL[i])) * A_CH[i];
    HCK(hipMalloc((void **)&n->act[i], e * sizeof(uint16_t)));
  }
  HCK(hipMalloc((void **)&n->out, (size_t)v0 * 16 * sizeof(float)));
  if (train) {
    HCK(hipMalloc((void **)&n->gpre, (size_t)v0 * 16 * sizeof(uint16_t)));
    HCK(hipMalloc((void **)&n->gout, (size_t)v0 * 16 * sizeof(uint16_t)));
    HCK(hipMalloc((void **)&n->gA, (size_t)v0 * 48 * sizeof(float)));
    HCK(hipMalloc((void **)&n->gB, (size_t)v0 * 48 * sizeof(float)));
    HCK(hipMalloc((void **)&n->gs0, (size_t)v0 * 16 * sizeof(float)));
    HCK(hipMalloc((void **)&n->gs1, (size_t)(v0 / 8) * 32 * sizeof(float)));
    HCK(hipMalloc((void **)&n->gs2, (size_t)(v0 / 64) * 64 * sizeof(float)));
  }
  n->cap_v0 = v0;
  n->train_ws = train;
  return 0;
}

/* refresh the BF16 weights and their transposes from the FP32 masters */
static int refresh_bf16(pdfs_net *n)
{
  for (int l = 0; l < NL; l++) {
    long long nw = 27LL * CIN[l] * COUT[l];
    NCK(pdfsnn_f32_to_bf16(n->P + n->woff[l], n->WB + n->woff[l], nw, n->st));
    NCK(pdfsnn_transpose_w(n->WB + n->woff[l], n->WT + n->woff[l], COUT[l], CIN[l], n->st));
  }
  return 0;
}

pdfs_net *pdfs_net_create(pdfs_ctx *c, unsigned long long seed)
{
  if (!c) {
    pdfsi_fail(PDFS_EINVAL, "pdfs_net_create: no context");
    return NULL;
  }
  pdfs_net *n = (pdfs_net *)calloc(1, sizeof *n);
  if (!n) return NULL;
  n->c = c;
  n->st = (hipStream_t)pdfs_ctx_stream(c);
  long long off = 0;
  for (int l = 0; l < NL; l++) {
    n->woff[l] = off;
    off += 27LL * CIN[l] * COUT[l];
  }
  for (int l = 0; l < NL; l++) {
    n->boff[l] = off;
    off += COUT[l];
  }
  n->nparam = off;
  n->scale = 1.f / 32.f;
  n->a = 1.f;
  n->b = 0.f;
  n->cut = 0.f;
  size_t bytes = (size_t)off * sizeof(float);
  float *h = (float *)malloc(bytes);
  if (!h || hipMalloc((void **)&n->P, bytes) != hipSuccess ||
      hipMalloc((void **)&n->G, bytes) != hipSuccess ||
      hipMalloc((void **)&n->M, bytes) != hipSuccess ||
      hipMalloc((void **)&n->V, bytes) != hipSuccess ||
      hipMalloc((void **)&n->WB, (size_t)off * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc((void **)&n->WT, (size_t)off * sizeof(uint16_t)) != hipSuccess ||
      hipHostMalloc((void **)&n->h_loss, sizeof(float), 0) != hipSuccess ||
      hipMalloc((void **)&n->loss, sizeof(float)) != hipSuccess) {
    free(h);
    pdfs_net_destroy(n);
    pdfsi_fail(PDFS_ENOMEM, "pdfs_net_create: allocation failed");
    return NULL;
  }
  /* He-normal initialisation (Box-Muller on a splitmix64 stream), zero bias */
  unsigned long long s = seed;
  for (int l = 0; l < NL; l++) {
    long long nw = 27LL * CIN[l] * COUT[l];
    double sd = sqrt(2.0 / (27.0 * CIN[l]));
    if (l == L_HEAD) sd *= 0.1;
    for (long long i = 0; i < nw; i++) {
      double u1 = u01(&s), u2 = u01(&s);
      h[n->woff[l] + i] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
    for (int i = 0; i < COUT[l]; i++) h[n->boff[l] + i] = 0.f;
  }
  int rc = hipMemcpy(n->P, h, bytes, hipMemcpyHostToDevice) != hipSuccess ||
           hipMemset(n->M, 0, bytes) != hipSuccess || hipMemset(n->V, 0, bytes) != hipSuccess ||
           hipMemset(n->G, 0, bytes) != hipSuccess;
  free(h);
  if (rc || refresh_bf16(n) || hipStreamSynchronize(n->st) != hipSuccess) {
    pdfs_net_destroy(n);
    pdfsi_fail(PDFS_EHIP, "pdfs_net_create: initialisation failed");
    return NULL;
  }
  return n;
}

void pdfs_net_destroy(pdfs_net *n)
{
  if (!n) return;
  /* no call on the context's stream here (the context may already be gone);
   * hipFree waits for the device's outstanding work itself */
  free_ws(n);
  hipFree(n->P);
  hipFree(n->G);
  hipFree(n->M);
  hipFree(n->V);
  hipFree(n->WB);
  hipFree(n->WT);
  hipFree(n->loss);
  hipFree(n->mask);
  hipFree(n->rag);
  free(n->h_rag);
  if (n->h_loss) hipHostFree(n->h_loss);
  free(n);
}

long long pdfs_net_num_params(const pdfs_net *n) { return n ? n->nparam : 0; }

int pdfs_net_get(pdfs_net *n, int which, float *host)
{
  if (!n || !host || which < 0 || which > 1) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_get: bad arguments");
  HCK(hipStreamSynchronize(n->st));
  HCK(hipMemcpy(host, which ? n->G : n->P, (size_t)n->nparam * sizeof(float),
                hipMemcpyDeviceToHost));
  return PDFS_OK;
}

int pdfs_net_set(pdfs_net *n, const float *host)
{
  if (!n || !host) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_set: bad arguments");
  HCK(hipStreamSynchronize(n->st));
  HCK(hipMemcpy(n->P, host, (size_t)n->nparam * sizeof(float), hipMemcpyHostToDevice));
  if (refresh_bf16(n)) return PDFS_EHIP;
  HCK(hipStreamSynchronize(n->st));
  return PDFS_OK;
}

void pdfs_net_get_calibration(const pdfs_net *n, float *scale, float *a, float *b)
{
  if (!n) return;
  if (scale) *scale = n->scale;
  if (a) *a = n->a;
  if (b) *b = n->b;
}

void pdfs_net_set_calibration(pdfs_net *n, float scale, float a, float b)
{
  if (!n) return;
  if (scale > 0) n->scale = scale;
  n->a = a;
  n->b = b;
}

float pdfs_net_get_cut(const pdfs_net *n)
{
  return n ? n->cut : 0.f;
}

void pdfs_net_set_cut(pdfs_net *n, float cut)
{
  if (n) n->cut = cut > 0.f ? cut : 0.f;
}

/* the cutoff in the units of the prediction d (h * hmax); none: +inf */
static float cut_units(const pdfs_net *n, double hmax)
{
  return n->cut > 0.f ? (float)(n->cut * hmax) : INFINITY;
}

/* File: "PDFSNET1", int32 NL, NL x (int32 cin, int32 cout), float32 scale,
 * a, b, cut (0: none), then per layer BF16 weights [27][cout][cin] and FP32 bias[cout].
 * Little-endian. */
int pdfs_net_save(pdfs_net *n, const char *path)
{
  if (!n || !path) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_save: bad arguments");
  float *h = (float *)malloc((size_t)n->nparam * sizeof(float));
  if (!h) return pdfsi_fail(PDFS_ENOMEM, "pdfs_net_save: out of memory");
  int rc = pdfs_net_get(n, 0, h);
  if (rc) {
    free(h);
    return rc;
  }
  FILE *f = fopen(path, "wb");
  if (!f) {
    free(h);
    return pdfsi_fail(PDFS_EINVAL, "pdfs_net_save: cannot open %s", path);
  }
  int ok = fwrite("PDFSNET1", 1, 8, f) == 8;
  int nl = NL;
  ok &= fwrite(&nl, 4, 1, f) == 1;
  for (int l = 0; l < NL; l++) {
    int d[2] = {CIN[l], COUT[l]};
    ok &= fwrite(d, 4, 2, f) == 2;
  }
  float cal[4] = {n->scale, n->a, n->b, n->cut};
  ok &= fwrite(cal, 4, 4, f) == 4;
  for (int l = 0; l < NL && ok; l++) {
    long long nw = 27LL * CIN[l] * COUT[l];
    for (long long i = 0; i < nw && ok; i++) {
      uint16_t b = f2bf_host(h[n->woff[l] + i]);
      ok &= fwrite(&b, 2, 1, f) == 1;
    }
    ok &= fwrite(h + n->boff[l], 4, (size_t)COUT[l], f) == (size_t)COUT[l];
  }
  ok &= fclose(f) == 0;
  free(h);
  return ok ? PDFS_OK : pdfsi_fail(PDFS_EINVAL, "pdfs_net_save: write failed");
}

int pdfs_net_load(pdfs_net *n, const char *path)
{
  if (!n || !path) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_load: bad arguments");
  FILE *f = fopen(path, "rb");
  if (!f) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_load: cannot open %s", path);
  float *h = (float *)malloc((size_t)n->nparam * sizeof(float));
  char magic[8];
  int nl = 0, ok = h != NULL;
  ok = ok && fread(magic, 1, 8, f) == 8 && memcmp(magic, "PDFSNET1", 8) == 0;
  ok = ok && fread(&nl, 4, 1, f) == 1 && nl == NL;
  for (int l = 0; l < NL && ok; l++) {
    int d[2];
    ok = fread(d, 4, 2, f) == 2 && d[0] == CIN[l] && d[1] == COUT[l];
  }
  float cal[4];
  ok = ok && fread(cal, 4, 4, f) == 4;
  for (int l = 0; l < NL && ok; l++) {
    long long nw = 27LL * CIN[l] * COUT[l];
    for (long long i = 0; i < nw && ok; i++) {
      uint16_t b;
      ok = fread(&b, 2, 1, f) == 1;
      h[n->woff[l] + i] = bf2f_host(b);
    }
    ok = ok && fread(h + n->boff[l], 4, (size_t)COUT[l], f) == (size_t)COUT[l];
  }
  fclose(f);
  int rc = ok ? pdfs_net_set(n, h) : pdfsi_fail(PDFS_EINVAL, "pdfs_net_load: %s is not a weights file of this network", path);
  free(h);
  if (rc == PDFS_OK) {
    n->scale = cal[0];
    n->a = cal[1];
    n->b = cal[2];
    n->cut = cal[3] > 0.f ? cal[3] : 0.f;
  }
  return rc;
}

/* ---- forward / backward ---- */
static int conv(pdfs_net *n, int l, const uint16_t *in, uint16_t *out, float *out_f32, int nb,
                int lvl, int Dz, int Dy, int Dx, int relu)
{
  NCK(pdfsnn_conv3(in, CIN[l], n->WB + n->woff[l], n->P + n->boff[l], COUT[l], out, out_f32, nb,
                   Dz >> lvl, Dy >> lvl, Dx >> lvl, relu, 0, n->st));
  return 0;
}

static int forward(pdfs_net *n, const uint16_t *x, int nb, int Dz, int Dy, int Dx)
{
  uint16_t **A = n->act;
  int rc;
#define CV(l, in, out, lvl) \
  if ((rc = conv(n, l, in, out, NULL, nb, lvl, Dz, Dy, Dx, 1))) return rc
#define POOL(in, out, lvl, C) \
  NCK(pdfsnn_pool2(in, out, nb, Dz >> (lvl), Dy >> (lvl), Dx >> (lvl), C, n->st))
#define UPCAT(src, Cs, skip, Ck, dst, lvl) \
  NCK(pdfsnn_up2_concat(src, Cs, skip, Ck, dst, nb, Dz >> (lvl), Dy >> (lvl), Dx >> (lvl), n->st))
  CV(L_E0A, x, A[A_A0], 0);
  CV(L_E0B, A[A_A0], A[A_S0], 0);
  POOL(A[A_S0], A[A_P0], 0, 16);
  CV(L_E1A, A[A_P0], A[A_A1], 1);
  CV(L_E1B, A[A_A1], A[A_S1], 1);
  POOL(A[A_S1], A[A_P1], 1, 32);
  CV(L_E2A, A[A_P1], A[A_A2], 2);
  CV(L_E2B, A[A_A2], A[A_S2], 2);
  POOL(A[A_S2], A[A_P2], 2, 64);
  CV(L_E3A, A[A_P2], A[A_A3], 3);
  CV(L_E3B, A[A_A3], A[A_B3], 3);
  UPCAT(A[A_B3], 64, A[A_S2], 64, A[A_C2], 2);
  CV(L_D2, A[A_C2], A[A_D2], 2);
  UPCAT(A[A_D2], 64, A[A_S1], 32, A[A_C1], 1);
  CV(L_D1, A[A_C1], A[A_D1], 1);
  UPCAT(A[A_D1], 32, A[A_S0], 16, A[A_C0], 0);
  CV(L_D0, A[A_C0], A[A_D0], 0);
  if ((rc = conv(n, L_HEAD, A[A_D0], NULL, n->out, nb, 0, Dz, Dy, Dx, 0))) return rc;
  return 0;
}

/* forward over an uneven batch: sg[k] (device) are the level-k segment
 * tables, nblk[k] / neblk[k] the convolution / element-wise blocks of level
 * k; the same layer sequence as forward() */
static int forward_rag(pdfs_net *n, const uint16_t *x, int nf, const pdfsnn_seg *const *sg,
                       const int *nblk, const int *neblk)
{
  uint16_t **A = n->act;
#define CVR(l, in, out, lvl)                                                                  \
  NCK(pdfsnn_conv3_rag(in, CIN[l], n->WB + n->woff[l], n->P + n->boff[l], COUT[l], out, NULL, \
                       sg[lvl], nf, nblk[lvl], 1, n->st))
#define POOLR(in, out, lvl, C) \
  NCK(pdfsnn_pool2_rag(in, out, sg[lvl], sg[(lvl) + 1], nf, neblk[(lvl) + 1], C, n->st))
#define UPCATR(src, Cs, skip, Ck, dst, lvl) \
  NCK(pdfsnn_up2_concat_rag(src, Cs, skip, Ck, dst, sg[lvl], sg[(lvl) + 1], nf, neblk[lvl], n->st))
  CVR(L_E0A, x, A[A_A0], 0);
  CVR(L_E0B, A[A_A0], A[A_S0], 0);
  POOLR(A[A_S0], A[A_P0], 0, 16);
  CVR(L_E1A, A[A_P0], A[A_A1], 1);
  CVR(L_E1B, A[A_A1], A[A_S1], 1);
  POOLR(A[A_S1], A[A_P1], 1, 32);
  CVR(L_E2A, A[A_P1], A[A_A2], 2);
  CVR(L_E2B, A[A_A2], A[A_S2], 2);
  POOLR(A[A_S2], A[A_P2], 2, 64);
  CVR(L_E3A, A[A_P2], A[A_A3], 3);
  CVR(L_E3B, A[A_A3], A[A_B3], 3);
  UPCATR(A[A_B3], 64, A[A_S2], 64, A[A_C2], 2);
  CVR(L_D2, A[A_C2], A[A_D2], 2);
  UPCATR(A[A_D2], 64, A[A_S1], 32, A[A_C1], 1);
  CVR(L_D1, A[A_C1], A[A_D1], 1);
  UPCATR(A[A_D1], 32, A[A_S0], 16, A[A_C0], 0);
  CVR(L_D0, A[A_C0], A[A_D0], 0);
  NCK(pdfsnn_conv3_rag(A[A_D0], CIN[L_HEAD], n->WB + n->woff[L_HEAD], n->P + n->boff[L_HEAD],
                       COUT[L_HEAD], NULL, n->out, sg[0], nf, nblk[0], 0, n->st));
#undef CVR
#undef POOLR
#undef UPCATR
  return 0;
}

/* one layer's backward: gy_pre (BF16, already through the ReLU mask) ->
 * dW, db and (optionally) gx = input gradient (FP32) */
static int layer_bwd(pdfs_net *n, int l, const uint16_t *x, const uint16_t *gy, float *gx, int nb,
                     int lvl, int Dz, int Dy, int Dx)
{
  const int dz = Dz >> lvl, dy = Dy >> lvl, dx = Dx >> lvl;
  NCK(pdfsnn_conv3_wgrad(x, CIN[l], gy, COUT[l], n->G + n->woff[l], n->G + n->boff[l], nb, dz,
                         dy, dx, n->st));
  if (gx)
    NCK(pdfsnn_conv3(gy, COUT[l], n->WT + n->woff[l], NULL, CIN[l], NULL, gx, nb, dz, dy, dx, 0, 1,
                     n->st));
  return 0;
}

static int backward(pdfs_net *n, const uint16_t *x, int nb, int Dz, int Dy, int Dx)
{
  uint16_t **A = n->act;
  const long long V0 = (long long)nb * Dz * Dy * Dx, V1 = V0 / 8, V2 = V0 / 64, V3 = V0 / 512;
  int rc;
#define LB(l, xin, gx, lvl) \
  if ((rc = layer_bwd(n, l, xin, n->gpre, gx, nb, lvl, Dz, Dy, Dx))) return rc
#define RB(g, y, N) NCK(pdfsnn_relu_bwd(g, y, n->gpre, N, n->st))
#define UB(gdst, Cs, Ck, dsrc, dskip, lvl) \
  NCK(pdfsnn_up2_concat_bwd(gdst, Cs, Ck, dsrc, dskip, nb, Dz >> (lvl), Dy >> (lvl), Dx >> (lvl), n->st))
#define PB(gc, gf, lvl, C) \
  NCK(pdfsnn_pool2_bwd(gc, gf, nb, Dz >> (lvl), Dy >> (lvl), Dx >> (lvl), C, 1, n->st))
  /* head (linear): gout -> dW, gd0 */
  if ((rc = layer_bwd(n, L_HEAD, A[A_D0], n->gout, n->gA, nb, 0, Dz, Dy, Dx))) return rc;
  RB(n->gA, A[A_D0], V0 * 16);
  LB(L_D0, A[A_C0], n->gB, 0);
  UB(n->gB, 32, 16, n->gA, n->gs0, 0);   /* gd1 (level 1), gs0 */
  RB(n->gA, A[A_D1], V1 * 32);
  LB(L_D1, A[A_C1], n->gB, 1);
  UB(n->gB, 64, 32, n->gA, n->gs1, 1);   /* gd2, gs1 */
  RB(n->gA, A[A_D2], V2 * 64);
  LB(L_D2, A[A_C2], n->gB, 2);
  UB(n->gB, 64, 64, n->gA, n->gs2, 2);   /* gb3, gs2 */
  RB(n->gA, A[A_B3], V3 * 64);
  LB(L_E3B, A[A_A3], n->gB, 3);
  RB(n->gB, A[A_A3], V3 * 64);
  LB(L_E3A, A[A_P2], n->gA, 3);
  PB(n->gA, n->gs2, 2, 64);
  RB(n->gs2, A[A_S2], V2 * 64);
  LB(L_E2B, A[A_A2], n->gB, 2);
  RB(n->gB, A[A_A2], V2 * 64);
  LB(L_E2A, A[A_P1], n->gA, 2);
  PB(n->gA, n->gs1, 1, 32);
  RB(n->gs1, A[A_S1], V1 * 32);
  LB(L_E1B, A[A_A1], n->gB, 1);
  RB(n->gB, A[A_A1], V1 * 32);
  LB(L_E1A, A[A_P0], n->gA, 1);
  PB(n->gA, n->gs0, 0, 16);
  RB(n->gs0, A[A_S0], V0 * 16);
  LB(L_E0B, A[A_A0], n->gB, 0);
  RB(n->gB, A[A_A0], V0 * 16);
  LB(L_E0A, x, NULL, 0);
  return 0;
}

static int check_dims(int nb, int Dz, int Dy, int Dx)
{
  if (nb < 1 || Dz < 8 || Dy < 8 || Dx < 8 || Dz % 8 || Dy % 8 || Dx % 8)
    return pdfsi_fail(PDFS_EINVAL, "network grid %dx%dx%dx%d: each side must be a multiple of 8",
                      nb, Dz, Dy, Dx);
  return 0;
}

int pdfs_net_forward_dev(pdfs_net *n, const uint16_t *mask, int nb, int Dz, int Dy, int Dx,
                         float *pred)
{
  if (!n || !mask) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_forward_dev: bad arguments");
  int rc;
  if ((rc = check_dims(nb, Dz, Dy, Dx))) return rc;
  const long long V0 = (long long)nb * Dz * Dy * Dx;
  if ((rc = ensure_ws(n, V0, 0))) return rc;
  if ((rc = forward(n, mask, nb, Dz, Dy, Dx))) return rc;
  if (pred)
    HCK(hipMemcpyAsync(pred, n->out, (size_t)V0 * 16 * sizeof(float), hipMemcpyDeviceToDevice,
                       n->st));
  HCK(hipStreamSynchronize(n->st));
  return PDFS_OK;
}

int pdfs_net_train_step_dev(pdfs_net *n, const uint16_t *mask, const float *target, int nb,
                            int Dz, int Dy, int Dx, float lr, float *loss)
{
  if (!n || !mask || !target) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_train_step_dev: bad arguments");
  int rc;
  if ((rc = check_dims(nb, Dz, Dy, Dx))) return rc;
  const long long V0 = (long long)nb * Dz * Dy * Dx;
  if ((rc = ensure_ws(n, V0, 1))) return rc;
  if ((rc = forward(n, mask, nb, Dz, Dy, Dx))) return rc;
  HCK(hipMemsetAsync(n->loss, 0, sizeof(float), n->st));
  HCK(hipMemsetAsync(n->G, 0, (size_t)n->nparam * sizeof(float), n->st));
  NCK(pdfsnn_mse(n->out, 16, target, V0, n->gout, n->loss, n->st));
  if ((rc = backward(n, mask, nb, Dz, Dy, Dx))) return rc;
  if (lr > 0.f) {
    n->step++;
    for (int l = 0; l < NL; l++) {
      long long nw = 27LL * CIN[l] * COUT[l];
      NCK(pdfsnn_adam(n->P + n->woff[l], n->G + n->woff[l], n->M + n->woff[l], n->V + n->woff[l],
                      nw, lr, 0.9f, 0.999f, 1e-8f, n->step, n->WB + n->woff[l],
                      n->WT + n->woff[l], COUT[l], CIN[l], n->st));
      NCK(pdfsnn_adam(n->P + n->boff[l], n->G + n->boff[l], n->M + n->boff[l], n->V + n->boff[l],
                      COUT[l], lr, 0.9f, 0.999f, 1e-8f, n->step, NULL, NULL, COUT[l], 1, n->st));
    }
  }
  HCK(hipMemcpyAsync(n->h_loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->st));
  HCK(hipStreamSynchronize(n->st));
  if (loss) *loss = *n->h_loss;
  return PDFS_OK;
}

/* coarse grid of a field: ceil(side / 4) rounded up to a multiple of 8 */
static int coarse(int side) { return ((side + 3) / 4 + 7) / 8 * 8; }

void pdfs_net_grid(int x, int y, int z, int *Dz, int *Dy, int *Dx)
{
  if (Dz) *Dz = coarse(z);
  if (Dy) *Dy = coarse(y);
  if (Dx) *Dx = coarse(x);
}

int pdfs_net_features_dev(pdfs_net *n, const float *init, int x, int y, int z, uint16_t *mask)
{
  if (!n || !init || !mask) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_features_dev: bad arguments");
  NCK(pdfsnn_mask_down4(init, x, y, z, mask, coarse(z), coarse(y), coarse(x), 1e6f, n->st));
  return PDFS_OK;
}

int pdfs_net_target_dev(pdfs_net *n, const float *u, int x, int y, int z, float *target)
{
  if (!n || !u || !target) return pdfsi_fail(PDFS_EINVAL, "pdfs_net_target_dev: bad arguments");
  NCK(pdfsnn_target_down4(u, x, y, z, target, coarse(z), coarse(y), coarse(x), n->scale, n->st));
  return PDFS_OK;
}

int pdfs_warm_start_f32_dev(pdfs_net *n, const float *init, float *w, int x, int y, int z,
                            double hmax)
{
  if (!n || !init || !w || x < 1 || y < 1 || z < 1 || pdfsi_bad_h(hmax, hmax, hmax))
    return pdfsi_fail(PDFS_EINVAL, "pdfs_warm_start_f32_dev: bad arguments");
  const int Dz = coarse(z), Dy = coarse(y), Dx = coarse(x);
  const long long V0 = (long long)Dz * Dy * Dx;
  if (V0 * 16 > n->cap_mask) {
    HCK(hipStreamSynchronize(n->st));
    hipFree(n->mask);
    n->mask = NULL;
    HCK(hipMalloc((void **)&n->mask, (size_t)V0 * 16 * sizeof(uint16_t)));
    n->cap_mask = V0 * 16;
  }
  int rc;
  if ((rc = ensure_ws(n, V0, 0))) return rc;
  NCK(pdfsnn_mask_down4(init, x, y, z, n->mask, Dz, Dy, Dx, 1e6f, n->st));
  if ((rc = forward(n, n->mask, 1, Dz, Dy, Dx))) return rc;
  NCK(pdfsnn_warm_up4(n->out, 16, Dz, Dy, Dx, (float)(hmax / n->scale), n->a,
                      (float)(n->b * hmax), cut_units(n, hmax), init, w, x, y, z, n->st));
  return PDFS_OK;
}

/* Warm starts of an uneven batch in one launch per layer: the fields'
 * coarse grids are packed level by level (segment tables), so every
 * convolution of the batch fills the GPU at once. Per field the result is
 * bit-identical to pdfs_warm_start_f32_dev. */
int pdfs_warm_start_batch_f32_dev(pdfs_net *n, int nf, const float *const *inits,
                                  float *const *ws, const int *dims, double hmax)
{
  if (!n || nf < 1 || !inits || !ws || !dims || pdfsi_bad_h(hmax, hmax, hmax))
    return pdfsi_fail(PDFS_EINVAL, "pdfs_warm_start_batch_f32_dev: bad arguments");
  const size_t segb = 4 * (size_t)(nf + 1) * sizeof(pdfsnn_seg);
  const size_t need = segb + (size_t)(nf + 1) * sizeof(pdfsnn_fld);
  if (need > n->cap_rag) {
    HCK(hipStreamSynchronize(n->st));
    hipFree(n->rag);
    free(n->h_rag);
    n->rag = NULL;
    n->h_rag = malloc(need);
    if (!n->h_rag) return pdfsi_fail(PDFS_ENOMEM, "pdfs_warm_start_batch_f32_dev: out of memory");
    HCK(hipMalloc(&n->rag, need));
    n->cap_rag = need;
  } else {
    /* the host tables may still feed an earlier copy */
    HCK(hipStreamSynchronize(n->st));
  }
  pdfsnn_seg *hs = (pdfsnn_seg *)n->h_rag;
  pdfsnn_fld *hf = (pdfsnn_fld *)((char *)n->h_rag + segb);
  long long pos[4] = {0, 0, 0, 0}, blk[4] = {0, 0, 0, 0}, eblk[4] = {0, 0, 0, 0}, fblk = 0;
  for (int f = 0; f < nf; f++) {
    const int x = dims[3 * f], y = dims[3 * f + 1], z = dims[3 * f + 2];
    if (!inits[f] || !ws[f] || x < 1 || y < 1 || z < 1)
      return pdfsi_fail(PDFS_EINVAL, "pdfs_warm_start_batch_f32_dev: field %d: bad pointer or dims", f);
    const int Dz = coarse(z), Dy = coarse(y), Dx = coarse(x);
    for (int k = 0; k < 4; k++) {
      const int dz = Dz >> k, dy = Dy >> k, dx = Dx >> k;
      pdfsnn_seg *g = &hs[k * (nf + 1) + f];
      g->pos = (int)pos[k];
      g->Dz = dz;
      g->Dy = dy;
      g->Dx = dx;
      g->blk = (int)blk[k];
      g->eblk = (int)eblk[k];
      pos[k] += (long long)dz * dy * dx;
      blk[k] += (long long)((dz + 3) / 4) * ((dy + 7) / 8) * ((dx + 7) / 8);
      eblk[k] += ((long long)dz * dy * dx + RAG_EP - 1) / RAG_EP;
    }
    hf[f].init = inits[f];
    hf[f].w = ws[f];
    hf[f].X = x;
    hf[f].Y = y;
    hf[f].Z = z;
    hf[f].blk = (int)fblk;
    fblk += ((long long)x * y * z + RAG_EF - 1) / RAG_EF;
  }
  if (pos[0] >= (1LL << 31) / 16 || blk[0] >= (1LL << 31) || fblk >= (1LL << 31))
    return pdfsi_fail(PDFS_EINVAL, "pdfs_warm_start_batch_f32_dev: batch too large for one launch");
  for (int k = 0; k < 4; k++) {
    pdfsnn_seg *g = &hs[k * (nf + 1) + nf];
    g->pos = (int)pos[k];
    g->Dz = g->Dy = g->Dx = 0;
    g->blk = (int)blk[k];
    g->eblk = (int)eblk[k];
  }
  memset(&hf[nf], 0, sizeof hf[nf]);
  hf[nf].blk = (int)fblk;
  const long long V0 = pos[0];
  int rc;
  if ((rc = ensure_ws(n, V0, 0))) return rc;
  if (V0 * 16 > n->cap_mask) {
    HCK(hipStreamSynchronize(n->st));
    hipFree(n->mask);
    n->mask = NULL;
    HCK(hipMalloc((void **)&n->mask, (size_t)V0 * 16 * sizeof(uint16_t)));
    n->cap_mask = V0 * 16;
  }
  HCK(hipMemcpyAsync(n->rag, n->h_rag, need, hipMemcpyHostToDevice, n->st));
  const pdfsnn_seg *dsg[4];
  int nblk[4], neblk[4];
  for (int k = 0; k < 4; k++) {
    dsg[k] = (const pdfsnn_seg *)n->rag + k * (nf + 1);
    nblk[k] = (int)blk[k];
    neblk[k] = (int)eblk[k];
  }
  const pdfsnn_fld *dfl = (const pdfsnn_fld *)((const char *)n->rag + segb);
  NCK(pdfsnn_mask_down4_rag(dfl, dsg[0], nf, neblk[0], n->mask, 1e6f, n->st));
  if ((rc = forward_rag(n, n->mask, nf, dsg, nblk, neblk))) return rc;
  NCK(pdfsnn_warm_up4_rag(n->out, 16, dsg[0], dfl, nf, (int)fblk, (float)(hmax / n->scale), n->a,
                          (float)(n->b * hmax), cut_units(n, hmax), n->st));
  return PDFS_OK;
}

/* Warm-started certified solve: field = warm start, FP32 passes until the
 * FP64 judge passes; if the judge still flags points (the warm start was
 * below the solution somewhere) or the pass cap is hit, the field is reset
 * to its init and solved plainly (a fallback). The FP64 polish (opts.polish)
 * runs once, on the accepted result: after the certified solve, or within
 * the plain re-solve. */
int pdfs_solve_warm_f32_dev(pdfs_ctx *c, pdfs_net *n, float *field, const float *init, int x,
                            int y, int z, double dx, double dy, double dz, const pdfs_opts *o,
                            pdfs_stats *st)
{
  if (!c || !n || !field || !init || x < 1 || y < 1 || z < 1 || pdfsi_bad_h(dx, dy, dz))
    return pdfsi_fail(PDFS_EINVAL, "pdfs_solve_warm_f32_dev: bad arguments");
  pdfs_opts oo;
  if (o)
    oo = *o;
  else
    pdfs_default_opts(&oo);
  hipStream_t s = (hipStream_t)pdfs_ctx_stream(c);
  hipEvent_t e0, e1;
  HCK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return pdfsi_fail(PDFS_EHIP, "pdfs_solve_warm_f32_dev: hipEventCreate failed");
  }
  double hmax = dx > dy ? dx : dy;
  if (dz > hmax) hmax = dz;
  float wms = 0.f;
  int rc = PDFS_OK;
  if (hipEventRecord(e0, s) != hipSuccess) rc = pdfsi_fail(PDFS_EHIP, "hipEventRecord failed");
  if (rc == PDFS_OK) rc = pdfs_warm_start_f32_dev(n, init, field, x, y, z, hmax);
  if (rc == PDFS_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
    rc = pdfsi_fail(PDFS_EHIP, "pdfs_solve_warm_f32_dev: event failed");
  if (rc == PDFS_OK) hipEventElapsedTime(&wms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  pdfs_opts ow = oo;
  ow.certify = 1;
  ow.judge = 1;
  ow.polish = 0;
  pdfs_stats sw;
  rc = pdfs_solve_f32_dev(c, field, init, x, y, z, dx, dy, dz, &ow, &sw);
  if (rc && rc != PDFS_ENOTCONV) return rc;
  sw.ms_warm = wms;
  if (rc == PDFS_ENOTCONV || sw.judge_flagged > 0) {
    HCK(hipMemcpyAsync(field, init, (size_t)x * y * z * sizeof(float), hipMemcpyDeviceToDevice, s));
    pdfs_stats sp;
    rc = pdfs_solve_f32_dev(c, field, init, x, y, z, dx, dy, dz, &oo, &sp);
    if (rc && rc != PDFS_ENOTCONV) return rc;
    sp.passes += sw.passes;
    sp.orderings += sw.orderings;
    sp.point_updates += sw.point_updates;
    sp.ms_sweeps += sw.ms_sweeps;
    sp.ms_judge += sw.ms_judge;
    sp.ms_sweep_kernels += sw.ms_sweep_kernels;
    sp.judge_runs += sw.judge_runs;
    sp.fallbacks = 1;
    sp.ms_warm = wms;
    sp.ms_convert += sw.ms_convert;
    sp.ms_total = sp.ms_sweeps + sp.ms_judge + sp.ms_polish + sp.ms_convert + wms;
    if (st) *st = sp;
    return rc;
  }
  /* fields swept in fixed point need no polish (pdfs_opts.storage) */
  if (oo.polish && sw.storage != PDFS_STORE_FIXED) {
    const int dims[3] = {x, y, z};
    rc = pdfsi_polish(c, 1, &field, &init, NULL, dims, dx, dy, dz, &oo, &sw);
    if (rc && rc != PDFS_ENOTCONV) return rc;
  }
  sw.ms_total = sw.ms_sweeps + sw.ms_judge + sw.ms_polish + sw.ms_convert + wms;
  if (st) *st = sw;
  return rc;
}

/* Batched warm-started certified solve: every field gets its warm start, the
 * batch runs in shared persistent launches with certified stop and a final
 * judge of the fields the passes did not certify; flagged (or unconverged)
 * fields are reset to their init and solved plainly in a second batch. The
 * FP64 polish (opts.polish) runs once per field, on the accepted results.
 * field_passes / field_fallback (optional) report per field. */
int pdfs_solve_batch_warm_f32_dev(pdfs_ctx *c, pdfs_net *n, int nf, float *const *fields,
                                  const float *const *inits, const int *dims, double dx,
                                  double dy, double dz, const pdfs_opts *o, pdfs_stats *st,
                                  int *field_passes, int *field_fallback)
{
  if (!c || !n || nf < 1 || !fields || !inits || !dims || pdfsi_bad_h(dx, dy, dz))
    return pdfsi_fail(PDFS_EINVAL, "pdfs_solve_batch_warm_f32_dev: bad arguments");
  pdfs_opts oo;
  if (o)
    oo = *o;
  else
    pdfs_default_opts(&oo);
  hipStream_t s = (hipStream_t)pdfs_ctx_stream(c);
  double hmax = dx > dy ? dx : dy;
  if (dz > hmax) hmax = dz;
  hipEvent_t e0, e1;
  HCK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return pdfsi_fail(PDFS_EHIP, "pdfs_solve_batch_warm_f32_dev: hipEventCreate failed");
  }
  int rc = 0;
  float wms = 0.f;
  pdfs_opts ow = oo;
  ow.certify = 1;
  ow.judge = 1;
  ow.polish = 0;
  pdfs_stats sw;
  memset(&sw, 0, sizeof sw);
  int *passes = (int *)calloc((size_t)nf, sizeof(int));
  int *bad = (int *)calloc((size_t)nf, sizeof(int));
  long long *fl = (long long *)calloc((size_t)nf, sizeof(long long));
  float **rf = (float **)calloc((size_t)nf, sizeof(float *));
  const float **ri = (const float **)calloc((size_t)nf, sizeof(float *));
  const float **gi = (const float **)calloc((size_t)nf, sizeof(float *));
  int *rd = (int *)calloc((size_t)3 * nf, sizeof(int));
  if (!passes || !bad || !fl || !rf || !ri || !gi || !rd) {
    rc = pdfsi_fail(PDFS_ENOMEM, "pdfs_solve_batch_warm_f32_dev: out of memory");
    goto out;
  }
  /* all warm starts in one launch per network layer */
  HGO(hipEventRecord(e0, s));
  if ((rc = pdfs_warm_start_batch_f32_dev(n, nf, inits, fields, dims, hmax))) goto out;
  HGO(hipEventRecord(e1, s));
  HGO(hipEventSynchronize(e1));
  hipEventElapsedTime(&wms, e0, e1);
  /* certified batch; fields the passes did not certify are judged once at
   * the end (field_flagged: their failing points, -1 if unconverged) */
  rc = pdfsi_batch_f32(c, nf, fields, inits, dims, dx, dy, dz, &ow, &sw, passes, fl);
  if (rc && rc != PDFS_ENOTCONV) goto out;
  rc = 0;
  int nbad = 0;
  for (int f = 0; f < nf; f++) {
    bad[f] = fl[f] != 0;
    nbad += bad[f];
    gi[f] = bad[f] ? NULL : inits[f];
  }
  sw.fallbacks = 0;
  sw.judge_flagged = 0;
  /* polish of the accepted fields (not of fields swept in fixed point) */
  if (oo.polish && nbad < nf && sw.storage != PDFS_STORE_FIXED) {
    rc = pdfsi_polish(c, nf, fields, gi, NULL, dims, dx, dy, dz, &oo, &sw);
    if (rc && rc != PDFS_ENOTCONV) goto out;
  }
  int rc_pol = rc;
  rc = 0;
  if (nbad) {
    int k = 0;
    for (int f = 0; f < nf; f++) {
      if (!bad[f]) continue;
      HGO(hipMemcpyAsync(fields[f], inits[f],
                         (size_t)dims[3 * f] * dims[3 * f + 1] * dims[3 * f + 2] * sizeof(float),
                         hipMemcpyDeviceToDevice, s));
      rf[k] = fields[f];
      ri[k] = inits[f];
      memcpy(rd + 3 * k, dims + 3 * f, 3 * sizeof(int));
      k++;
    }
    pdfs_opts op = oo;
    op.judge = 1;
    pdfs_stats sp;
    int *pp = (int *)calloc((size_t)k, sizeof(int));
    if (!pp) {
      rc = pdfsi_fail(PDFS_ENOMEM, "pdfs_solve_batch_warm_f32_dev: out of memory");
      goto out;
    }
    rc = pdfs_solve_batch_f32_dev(c, k, rf, ri, rd, dx, dy, dz, &op, &sp, pp, NULL);
    k = 0;
    for (int f = 0; f < nf; f++)
      if (bad[f]) passes[f] += pp[k++];
    free(pp);
    if (rc && rc != PDFS_ENOTCONV) goto out;
    sw.passes += sp.passes;
    sw.orderings += sp.orderings;
    sw.point_updates += sp.point_updates;
    sw.ms_sweeps += sp.ms_sweeps;
    sw.ms_judge += sp.ms_judge;
    sw.judge_runs += sp.judge_runs;
    sw.judge_flagged = sp.judge_flagged;
    if (sp.polish_passes > sw.polish_passes) sw.polish_passes = sp.polish_passes;
    sw.polish_orderings += sp.polish_orderings;
    if (sp.polish_maxdiff > sw.polish_maxdiff) sw.polish_maxdiff = sp.polish_maxdiff;
    sw.polish_skipped += sp.polish_skipped;
    sw.ms_polish += sp.ms_polish;
    sw.ms_convert += sp.ms_convert;
    sw.fallbacks = nbad;
  }
  if (rc == PDFS_OK) rc = rc_pol;
  sw.ms_warm = wms;
  sw.ms_total = sw.ms_sweeps + sw.ms_judge + sw.ms_polish + sw.ms_convert + wms;
  if (st) *st = sw;
  for (int f = 0; f < nf; f++) {
    if (field_passes) field_passes[f] = passes[f];
    if (field_fallback) field_fallback[f] = bad[f];
  }
out:
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  free(passes);
  free(bad);
  free(fl);
  free(rf);
  free(ri);
  free(gi);
  free(rd);
  return rc;
}
