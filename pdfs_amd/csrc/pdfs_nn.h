/* Launchers of the warm-start network's gfx950 kernels (pdfs_nn.hip).
 * Tensors are channels-last [n][z][y][x][C] with C a multiple of 16;
 * BF16 tensors are uint16_t (upper half of an IEEE float), weights are
 * [27][Cout][Cin] with tap = (dz*3 + dy)*3 + dx. All launches are
 * asynchronous on stream st and return 0 or a HIP error code (-1: shape
 * outside what the kernel supports). */
#ifndef PDFS_NN_H
#define PDFS_NN_H
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* out = act(conv3x3x3(in, W) + bias), zero padding; flip = 1 mirrors the taps
 * (with transposed weights [27][Cin][Cout] this is the input gradient).
 * Cin <= 128, Cout <= 128; out (BF16) and out_f32 each optional. */
int pdfsnn_conv3(const uint16_t *in, int Cin, const uint16_t *W, const float *bias, int Cout,
                 uint16_t *out, float *out_f32, int n, int Dz, int Dy, int Dx, int relu, int flip,
                 hipStream_t st);
/* dW[tap][co][ci] += sum_p dY[p][co] * X[p + shift(tap)][ci]; (Cout/16)(Cin/16) <= 32 */
/* dW += the weight gradient; db (or NULL) += the bias gradient */
int pdfsnn_conv3_wgrad(const uint16_t *X, int Cin, const uint16_t *dY, int Cout, float *dW,
                       float *db, int n, int Dz, int Dy, int Dx, hipStream_t st);
int pdfsnn_pool2(const uint16_t *in, uint16_t *out, int n, int Dz, int Dy, int Dx, int C,
                 hipStream_t st);
/* din (fine dims Dz,Dy,Dx) = or += dout / 8 */
int pdfsnn_pool2_bwd(const float *dout, float *din, int n, int Dz, int Dy, int Dx, int C,
                     int accumulate, hipStream_t st);
/* dst (fine, Cs+Ck channels) = [nearest-up2(src) | skip] */
int pdfsnn_up2_concat(const uint16_t *src, int Cs, const uint16_t *skip, int Ck, uint16_t *dst,
                      int n, int Dz, int Dy, int Dx, hipStream_t st);
/* dsrc = sum of the 8 children's first Cs channels; dskip = channels [Cs, Cs+Ck) */
int pdfsnn_up2_concat_bwd(const float *ddst, int Cs, int Ck, float *dsrc, float *dskip, int n,
                          int Dz, int Dy, int Dx, hipStream_t st);
/* out = y ? (y > 0 ? g : 0) : g, as BF16 */
int pdfsnn_relu_bwd(const float *g, const uint16_t *y, uint16_t *out, long N, hipStream_t st);
/* loss += mean (o[i*C] - t[i])^2 over the N samples; grad channel 0 = 2 (o - t) / N */
int pdfsnn_mse(const float *o, int C, const float *t, long N, uint16_t *grad, float *loss,
               hipStream_t st);
/* Adam step `step` (1-based) on FP32 masters; refreshes BF16 weights and their
 * [27][Ci][Co] transpose when wb != NULL */
int pdfsnn_adam(float *w, const float *gw, float *m, float *v, long N, float lr, float b1, float b2,
                float eps, int step, uint16_t *wb, uint16_t *wt, int Co, int Ci, hipStream_t st);
int pdfsnn_f32_to_bf16(const float *x, uint16_t *y, long N, hipStream_t st);
int pdfsnn_transpose_w(const uint16_t *wb, uint16_t *wt, int Co, int Ci, hipStream_t st);
/* fraction of known cells (init < far) per 4x4x4 block -> channel 0 of [Dz][Dy][Dx][16] */
int pdfsnn_mask_down4(const float *f, int X, int Y, int Z, uint16_t *m, int Dz, int Dy, int Dx,
                      float far, hipStream_t st);
/* t[z][y][x] = scale * u(4z, 4y, 4x) (clamped to the field) */
int pdfsnn_target_down4(const float *u, int X, int Y, int Z, float *t, int Dz, int Dy, int Dx,
                        float scale, hipStream_t st);
/* d = trilinear(pred) * inv_scale; w = min(init, a * max(d, 0) + b) on the
 * dense field, and w = init where d >= cut */
int pdfsnn_warm_up4(const float *pred, int C, int Dz, int Dy, int Dx, float inv_scale, float a,
                    float b, float cut, const float *init, float *w, int X, int Y, int Z,
                    hipStream_t st);

/* Uneven batches. The tensors of a batch's fields are packed one after the
 * other: at a network level, field f's tensor starts at position seg[f].pos
 * of the level's packed tensor and has dims (Dz, Dy, Dx); seg[f].blk is its
 * first block in a convolution launch and seg[f].eblk its first block in an
 * element-wise launch (RAG_EP positions per block). seg[nf] is a sentinel
 * holding the totals. The kernels compute, per field, exactly what their
 * single-field forms compute. */
#define RAG_EP 64
#define RAG_EF 4096
typedef struct {
  int pos, Dz, Dy, Dx, blk, eblk;
} pdfsnn_seg;
/* a field's dense ends of the warm start: init (read), w (written), dims,
 * its first block in the dense warm-start launch (RAG_EF points per block;
 * fld[nf].blk is the total) */
typedef struct {
  const float *init;
  float *w;
  int X, Y, Z, blk;
} pdfsnn_fld;
int pdfsnn_conv3_rag(const uint16_t *in, int Cin, const uint16_t *W, const float *bias, int Cout,
                     uint16_t *out, float *out_f32, const pdfsnn_seg *seg, int nf, int nblocks,
                     int relu, hipStream_t st);
int pdfsnn_pool2_rag(const uint16_t *in, uint16_t *out, const pdfsnn_seg *seg_in,
                     const pdfsnn_seg *seg_out, int nf, int eblocks_out, int C, hipStream_t st);
int pdfsnn_up2_concat_rag(const uint16_t *src, int Cs, const uint16_t *skip, int Ck,
                          uint16_t *dst, const pdfsnn_seg *seg_fine, const pdfsnn_seg *seg_coarse,
                          int nf, int eblocks_fine, hipStream_t st);
int pdfsnn_mask_down4_rag(const pdfsnn_fld *fld, const pdfsnn_seg *seg0, int nf, int eblocks0,
                          uint16_t *m, float far, hipStream_t st);
int pdfsnn_warm_up4_rag(const float *pred, int C, const pdfsnn_seg *seg0, const pdfsnn_fld *fld,
                        int nf, int fblocks, float inv_scale, float a, float b, float cut,
                        hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
