/*
 * Kernels of the warm-start network of libagdeblend.so (MI355X): a 3-D
 * U-Net over (space, space, time) patches of pseudo-deblended data, trained
 * on the GPU by these kernels, whose prediction becomes the starting model
 * of the inversion through agdx_warm_model / agdx_run_warm. No counterpart
 * in the reference (which always starts from the thresholded first
 * update); used by agdeblend_amd.warmstart.
 *
 * Tensors are device pointers. Activations: [n][d][h][w][c] bf16 with c a
 * multiple of 16; weights: [27][co][ci] bf16 (tap = (kd * 3 + kh) * 3 + kw);
 * masters, moments and gradients: fp32. stream: a hipStream_t or NULL.
 * Return 0 on success, 1 on failure (message on stderr).
 */
#ifndef AGDEBLEND_UNET_H
#define AGDEBLEND_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = conv3x3x3(x, wt) (+ bias if mode & 1), ReLU if mode & 2, times
 * (gate > 0) if mode & 4 (gate has y's layout) */
int agdu_conv3(void const *x, void const *wt, float const *bias, void const *gate, void *y,
               int N, int D, int H, int W, int CI, int CO, int mode, void *stream);
/* agdu_conv3 (mode bits 1 and 2) that also writes pool, the 2x2x2 max
 * pooling of y ([N][D/2][H/2][W/2][CO], the values of agdu_maxpool, no
 * arg); D, H, W even. Kernels whose blocks hold whole plane pairs only
 * (16 and 32 output channels). */
int agdu_conv3_pool(void const *x, void const *wt, float const *bias, void *y, void *pool,
                    int N, int D, int H, int W, int CI, int CO, int mode, void *stream);
/* agdu_conv3 on the concatenation [upsample(xu), x] without forming it:
 * xu is [N][D/2][H/2][W/2][CU] (nearest x2 upsampling; D, H, W even), x is
 * [N][D][H][W][CS]; wt is [27][CO][CU + CS] */
int agdu_conv3_cat(void const *xu, void const *x, void const *wt, float const *bias,
                   void const *gate, void *y, int N, int D, int H, int W, int CU, int CS,
                   int CO, int mode, void *stream);
/* Combined taps of the first CU input channels of wt ([27][CO][CI]) for
 * agdu_conv3_up: wcu is [48][CO][CU] bf16. A 3x3x3 convolution over a
 * nearest x2 upsampled input reads, along each of d and h, only two
 * coarse rows per output parity; the weights of taps that read the same
 * coarse row are summed (fp32, rounded once to bf16). */
int agdu_up_weights(void const *wt, void *wcu, int CI, int CO, int CU, void *stream);
/* agdu_conv3_cat with the xu channels on the combined taps wcu
 * (agdu_up_weights of wt): 12 instead of 27 matrix-core steps per 16 of
 * those channels. Equal to agdu_conv3_cat up to the rounding of the
 * combined weights to bf16. */
int agdu_conv3_up(void const *xu, void const *x, void const *wt, void const *wcu,
                  float const *bias, void const *gate, void *y, int N, int D, int H, int W,
                  int CU, int CS, int CO, int mode, void *stream);
/* First layer from a one-channel fp32 field:
 * y[v][co] = ReLU(bias[co] + sum_tap wt[tap][co][0] bf16(src[v + tap] scale)),
 * 16 output channels, wt [27][16][wci] (input channel 0 is used) */
int agdu_conv_in(float const *src, void const *wt, int wci, float const *bias, void *y, int N,
                 int D, int H, int W, float scale, void *stream);
/* Output channel 0 of a 16 -> 16 convolution, in fp32:
 * dst[v] = base[v] + unscale (bias[0] + sum_tap sum_ci wt[tap][0][ci] x[v + tap][ci])
 * (base may be NULL), wt [27][wco][16] */
int agdu_conv_out(void const *x, void const *wt, int wco, float const *bias, float const *base,
                  float *dst, int N, int D, int H, int W, float unscale, void *stream);
/* weights of the input gradient: wtt[tap][ci][co] = wt[26 - tap][co][ci] */
int agdu_flip_weights(void const *wt, void *wtt, int CI, int CO, void *stream);
/* floats of scratch the weight gradient needs (-1: no kernel for the
 * channel pair) */
int64_t agdu_wgrad_part_len(int N, int D, int H, int W, int CI, int CO);
/* dw[tap][co][ci] = sum_v dy[v][co] x[v + tap][ci] and, if db is not NULL,
 * db[co] = sum_v dy[v][co]; fp32, summed in a fixed order (deterministic).
 * part: agdu_wgrad_part_len floats */
int agdu_conv3_wgrad(void const *x, void const *dy, float *part, float *dw, float *db, int N,
                     int D, int H, int W, int CI, int CO, void *stream);
/* 2x2x2 max pooling, arg: winning offset per output element */
int agdu_maxpool(void const *x, void *y, void *arg, int N, int D, int H, int W, int C,
                 void *stream);
int agdu_maxpool_bwd(void const *dy, void const *arg, void *dx, int N, int D, int H, int W,
                     int C, void *stream);
/* nearest x2 upsampling into channels [c0, c0 + C) of a Cy-channel tensor */
int agdu_upsample(void const *x, void *y, int N, int D, int H, int W, int C, int Cy, int c0,
                  void *stream);
int agdu_upsample_bwd(void const *dy, void *dx, int N, int D, int H, int W, int C, int Cy,
                      int c0, void *stream);
/* y[v][cy0 + c] = x[v][cx0 + c] (mode 1: added to y; mode 2: zero where
 * gate[v][cy0 + c] <= 0) */
int agdu_copy_channels(void const *x, void *y, int64_t nvox, int Cx, int cx0, int Cy,
                       int cy0, int C, int mode, void const *gate, void *stream);
/* fp32 field -> channel 0 of a C-channel bf16 tensor (times scale), and back */
int agdu_pack_input(float const *src, void *x, int64_t nvox, int C, float scale,
                    void *stream);
int agdu_unpack_output(void const *y, float *dst, float const *base, int64_t nvox, int C,
                       float scale, void *stream);
/* L2 loss of channel 0 against t * tscale: per-block partial sums of the
 * squared error in part (agdu_loss_parts(nvox) floats) and the gradient
 * 2 (y - t) / nvox in dy */
int agdu_l2_loss(void const *y, float const *t, void *dy, float *part, int64_t nvox, int C,
                 float tscale, void *stream);
int64_t agdu_loss_parts(int64_t nvox);
/* Adam step number step (1-based) on fp32 masters; refreshes the bf16 copy */
int agdu_adam(float *w, float *m, float *v, float const *g, void *wb, int len, float lr,
              float b1, float b2, float eps, int step, void *stream);
int agdu_to_bf16(float const *a, void *b, int64_t n, void *stream);
int agdu_from_bf16(void const *b, float *a, int64_t n, void *stream);

#ifdef __cplusplus
}
#endif

#endif
