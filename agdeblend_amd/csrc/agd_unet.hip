// HIP kernels of the warm-start network (gfx950): a 3-D U-Net over
// (space, space, time) patches, channels-last, BF16 activations and weights,
// FP32 accumulation, trained on the GPU by these kernels.
//
// Layouts (row-major, last index fastest):
//   activations  [n][d][h][w][c]  bf16 (c a multiple of 16)
//   weights      [27][co][ci]     bf16, tap = (kd * 3 + kh) * 3 + kw
//   masters      fp32 copies of the weights and biases, Adam moments
//
// 3x3x3 convolutions (stride 1, zero padding 1) are implicit GEMMs on the
// matrix cores: one wavefront computes 16 consecutive voxels along w by all
// output channels with v_mfma_f32_16x16x16_bf16 (M = voxels, N = 16 output
// channels, K = 16 input channels of one tap per instruction). The input
// gradient is the same kernel run on the flipped, transposed weights; the
// weight gradient is a deterministic two-pass reduction (per-block FP32
// partial sums over voxel slabs, then a fixed-order sum).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <stdio.h>

#include "agdeblend_unet.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define UT 256

static int hip_ok(hipError_t e, char const *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "ERROR: HIP %s: %s\n", what, hipGetErrorString(e));
    return 1;
  }
  return 0;
}

static hipStream_t ustream(void *s) { return (hipStream_t)s; }

template <typename K, typename... A>
static int ulaunch(char const *name, K kernel, dim3 grid, dim3 block, void *stream,
                   A... args) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return 0;
  hipLaunchKernelGGL(kernel, grid, block, 0, ustream(stream), args...);
  return hip_ok(hipGetLastError(), name);
}

__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t const u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  uint32_t const r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

// ---------------------------------------------------------------------------
// Convolution (forward, or input gradient with flipped weights)
// ---------------------------------------------------------------------------
// grid: (ceil(W / 64), H, N * D); block: 256 = 4 waves, each 16 voxels along w.
// mode bits: 1 = add bias, 2 = ReLU on the output, 4 = multiply the output
// by (gate > 0) (the ReLU mask of the layer below, for input gradients).
template <int CI, int CO>
__global__ __launch_bounds__(UT) void k_conv3(
    uint16_t const *__restrict__ x, uint16_t const *__restrict__ wt,
    float const *__restrict__ bias, uint16_t const *__restrict__ gate,
    uint16_t *__restrict__ y, int D, int H, int W, int mode) {
  static_assert(CI % 16 == 0 && CO % 16 == 0, "channels in multiples of 16");
  constexpr int NT = CO / 16;
  int const lane = threadIdx.x & 63;
  int const wave = threadIdx.x >> 6;
  int const n = blockIdx.z / D;
  int const d = blockIdx.z - n * D;
  int const h = blockIdx.y;
  int const w0 = blockIdx.x * 64 + wave * 16;
  int const i = lane & 15;   // A row: voxel; B / D column: channel in the tile
  int const kq = lane >> 4;  // k quarter: channels 4 kq .. 4 kq + 3
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  int64_t const plane = (int64_t)H * W;
  uint16_t const *xn = x + (int64_t)n * D * plane * CI;
  for (int kd = 0; kd < 3; ++kd) {
    int const dd = d + kd - 1;
    bool const okd = dd >= 0 && dd < D;
    for (int kh = 0; kh < 3; ++kh) {
      int const hh = h + kh - 1;
      bool const okh = okd && hh >= 0 && hh < H;
#pragma unroll
      for (int kw = 0; kw < 3; ++kw) {
        int const ww = w0 + i + kw - 1;
        bool const ok = okh && ww >= 0 && ww < W;
        int const tap = (kd * 3 + kh) * 3 + kw;
        int64_t const vox = ((int64_t)(ok ? dd : 0) * H + (ok ? hh : 0)) * W + (ok ? ww : 0);
        uint16_t const *xa = xn + vox * CI + 4 * kq;
        uint16_t const *wb = wt + (int64_t)tap * CO * CI + (int64_t)i * CI + 4 * kq;
#pragma unroll
        for (int c = 0; c < CI / 16; ++c) {
          bf16x4 a = *(bf16x4 const *)(xa + 16 * c);
          if (!ok) a = bf16x4{0, 0, 0, 0};
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            bf16x4 const b = *(bf16x4 const *)(wb + (int64_t)t * 16 * CI + 16 * c);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
          }
        }
      }
    }
  }
  // D: lane holds voxels 4 kq + r (r < 4) of channel t * 16 + i
  int64_t const orow = (((int64_t)n * D + d) * H + h) * W;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int const co = t * 16 + i;
    float const b = (mode & 1) ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int const wv = w0 + 4 * kq + r;
      if (wv >= W) continue;
      int64_t const o = (orow + wv) * CO + co;
      float v = acc[t][r] + b;
      if (mode & 2) v = fmaxf(v, 0.f);
      if ((mode & 4) && !(bf2f(gate[o]) > 0.f)) v = 0.f;
      y[o] = f2bf(v);
    }
  }
}

// wt_t[tap][ci][co] = wt[26 - tap][co][ci]: the weights of the input
// gradient (a convolution of dy with the flipped kernel, CO -> CI channels)
__global__ void k_flip_weights(uint16_t const *__restrict__ wt, uint16_t *__restrict__ wtt,
                               int CI, int CO) {
  int const idx = blockIdx.x * blockDim.x + threadIdx.x;
  int const total = 27 * CI * CO;
  if (idx >= total) return;
  int const tap = idx / (CI * CO);
  int const r = idx - tap * CI * CO;
  int const ci = r / CO;
  int const co = r - ci * CO;
  wtt[idx] = wt[((int64_t)(26 - tap) * CO + co) * CI + ci];
}

// ---------------------------------------------------------------------------
// Weight and bias gradients: dW[tap][co][ci] = sum_v dy[v][co] x[v + tap][ci]
// ---------------------------------------------------------------------------
// grid: (27, nslab); block 256. Each block sums its voxel slab for one tap
// into part[slab][tap][co][ci] (fp32); threads own (co, ci) pairs.
__global__ __launch_bounds__(UT) void k_conv3_wgrad(
    uint16_t const *__restrict__ x, uint16_t const *__restrict__ dy, float *__restrict__ part,
    int N, int D, int H, int W, int CI, int CO, int slab) {
  __shared__ float sx[64 * 96];
  __shared__ float sd[64 * 64];
  int const tap = blockIdx.x;
  int const kd = tap / 9, kh = (tap / 3) % 3, kw = tap % 3;
  int64_t const nvox = (int64_t)N * D * H * W;
  int64_t const v0 = (int64_t)blockIdx.y * slab;
  int64_t const v1 = v0 + slab < nvox ? v0 + slab : nvox;
  int const npair = CI * CO;
  float acc[24];
#pragma unroll
  for (int q = 0; q < 24; ++q) acc[q] = 0.f;
  for (int64_t vb = v0; vb < v1; vb += 64) {
    int const nv = (int)(v1 - vb < 64 ? v1 - vb : 64);
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * CI; e += UT) {
      int const j = e / CI, ci = e - j * CI;
      float val = 0.f;
      if (j < nv) {
        int64_t const v = vb + j;
        int const w = (int)(v % W);
        int64_t r = v / W;
        int const h = (int)(r % H);
        r /= H;
        int const d = (int)(r % D);
        int const n = (int)(r / D);
        int const dd = d + kd - 1, hh = h + kh - 1, ww = w + kw - 1;
        if (dd >= 0 && dd < D && hh >= 0 && hh < H && ww >= 0 && ww < W)
          val = bf2f(x[((((int64_t)n * D + dd) * H + hh) * W + ww) * CI + ci]);
      }
      sx[j * CI + ci] = val;
    }
    for (int e = threadIdx.x; e < 64 * CO; e += UT) {
      int const j = e / CO, co = e - j * CO;
      sd[j * CO + co] = j < nv ? bf2f(dy[(vb + j) * CO + co]) : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 24; ++q) {
      int const p = threadIdx.x + q * UT;
      if (p < npair) {
        int const co = p / CI, ci = p - co * CI;
        float s = acc[q];
        for (int j = 0; j < nv; ++j) s += sd[j * CO + co] * sx[j * CI + ci];
        acc[q] = s;
      }
    }
  }
  float *out = part + ((int64_t)blockIdx.y * 27 + tap) * npair;
#pragma unroll
  for (int q = 0; q < 24; ++q) {
    int const p = threadIdx.x + q * UT;
    if (p < npair) out[p] = acc[q];
  }
}

// bias partials: part[slab][co] = sum over the slab of dy[v][co]
__global__ __launch_bounds__(UT) void k_bias_grad(uint16_t const *__restrict__ dy,
                                                  float *__restrict__ part, int64_t nvox,
                                                  int CO, int slab) {
  int64_t const v0 = (int64_t)blockIdx.x * slab;
  int64_t const v1 = v0 + slab < nvox ? v0 + slab : nvox;
  for (int co = threadIdx.x; co < CO; co += blockDim.x) {
    float s = 0.f;
    for (int64_t v = v0; v < v1; ++v) s += bf2f(dy[v * CO + co]);
    part[(int64_t)blockIdx.x * CO + co] = s;
  }
}

// out[i] = sum_s part[s][i], s in order
__global__ void k_reduce_parts(float const *__restrict__ part, float *__restrict__ out,
                               int nslab, int len) {
  int const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (int k = 0; k < nslab; ++k) s += part[(int64_t)k * len + i];
  out[i] = s;
}

// ---------------------------------------------------------------------------
// Pooling, upsampling, concatenation, loss, optimizer
// ---------------------------------------------------------------------------
// 2x2x2 max pool: y[n][d][h][w][c] over x[n][2d..][2h..][2w..][c]; arg holds
// the winning offset (first maximum in (a, b, e) order)
__global__ void k_maxpool(uint16_t const *__restrict__ x, uint16_t *__restrict__ y,
                          uint8_t *__restrict__ arg, int N, int D, int H, int W, int C) {
  int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int const Do = D / 2, Ho = H / 2, Wo = W / 2;
  int64_t const total = (int64_t)N * Do * Ho * Wo * C;
  if (idx >= total) return;
  int const c = (int)(idx % C);
  int64_t r = idx / C;
  int const w = (int)(r % Wo);
  r /= Wo;
  int const h = (int)(r % Ho);
  r /= Ho;
  int const d = (int)(r % Do);
  int const n = (int)(r / Do);
  float best = -INFINITY;
  int barg = 0;
  for (int o = 0; o < 8; ++o) {
    int const a = o >> 2, b = (o >> 1) & 1, e = o & 1;
    float const v =
        bf2f(x[((((int64_t)n * D + 2 * d + a) * H + 2 * h + b) * W + 2 * w + e) * C + c]);
    if (v > best) {
      best = v;
      barg = o;
    }
  }
  y[idx] = f2bf(best);
  arg[idx] = (uint8_t)barg;
}

__global__ void k_maxpool_bwd(uint16_t const *__restrict__ dy, uint8_t const *__restrict__ arg,
                              uint16_t *__restrict__ dx, int N, int D, int H, int W, int C) {
  // one thread per input element
  int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t const total = (int64_t)N * D * H * W * C;
  if (idx >= total) return;
  int const c = (int)(idx % C);
  int64_t r = idx / C;
  int const w = (int)(r % W);
  r /= W;
  int const h = (int)(r % H);
  r /= H;
  int const d = (int)(r % D);
  int const n = (int)(r / D);
  int const Do = D / 2, Ho = H / 2, Wo = W / 2;
  int64_t const o =
      ((((int64_t)n * Do + d / 2) * Ho + h / 2) * Wo + w / 2) * C + c;
  int const mine = ((d & 1) << 2) | ((h & 1) << 1) | (w & 1);
  dx[idx] = (arg[o] == mine) ? dy[o] : (uint16_t)0;
}

// nearest x2 upsampling of x [n][D][H][W][C] into the channels [c0, c0 + C)
// of y [n][2D][2H][2W][Cy]
__global__ void k_upsample(uint16_t const *__restrict__ x, uint16_t *__restrict__ y, int N,
                           int D, int H, int W, int C, int Cy, int c0) {
  int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t const total = (int64_t)N * 8 * D * H * W * C;
  if (idx >= total) return;
  int const c = (int)(idx % C);
  int64_t r = idx / C;
  int const w = (int)(r % (2 * W));
  r /= 2 * W;
  int const h = (int)(r % (2 * H));
  r /= 2 * H;
  int const d = (int)(r % (2 * D));
  int const n = (int)(r / (2 * D));
  y[((((int64_t)n * 2 * D + d) * 2 * H + h) * 2 * W + w) * Cy + c0 + c] =
      x[((((int64_t)n * D + d / 2) * H + h / 2) * W + w / 2) * C + c];
}

// gradient of the upsampling: dx[v][c] = sum of the 8 children's
// dy[..][c0 + c], in a fixed order; accumulated in fp32
__global__ void k_upsample_bwd(uint16_t const *__restrict__ dy, uint16_t *__restrict__ dx,
                               int N, int D, int H, int W, int C, int Cy, int c0) {
  int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t const total = (int64_t)N * D * H * W * C;
  if (idx >= total) return;
  int const c = (int)(idx % C);
  int64_t r = idx / C;
  int const w = (int)(r % W);
  r /= W;
  int const h = (int)(r % H);
  r /= H;
  int const d = (int)(r % D);
  int const n = (int)(r / D);
  float s = 0.f;
  for (int o = 0; o < 8; ++o) {
    int const a = o >> 2, b = (o >> 1) & 1, e = o & 1;
    s += bf2f(dy[((((int64_t)n * 2 * D + 2 * d + a) * 2 * H + 2 * h + b) * 2 * W + 2 * w + e) *
                     Cy + c0 + c]);
  }
  dx[idx] = f2bf(s);
}

// channel slice copy: y[v][cy0 + c] = x[v][cx0 + c], c < C, v < nvox
__global__ void k_copy_channels(uint16_t const *__restrict__ x, uint16_t *__restrict__ y,
                                int64_t nvox, int Cx, int cx0, int Cy, int cy0, int C,
                                int mode, uint16_t const *__restrict__ gate) {
  int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nvox * C) return;
  int64_t const v = idx / C;
  int const c = (int)(idx - v * C);
  uint16_t val = x[v * Cx + cx0 + c];
  // mode 1: add into y (fp32 sum, rounded once); mode 2: ReLU mask by gate
  if (mode == 1) val = f2bf(bf2f(val) + bf2f(y[v * Cy + cy0 + c]));
  if (mode == 2 && !(bf2f(gate[v * Cy + cy0 + c]) > 0.f)) val = 0;
  y[v * Cy + cy0 + c] = val;
}

// fp32 field -> channel 0 of a bf16 tensor with C channels (others zero)
__global__ void k_pack_input(float const *__restrict__ src, uint16_t *__restrict__ x,
                             int64_t nvox, int C, float scale) {
  int64_t const idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nvox * C) return;
  int64_t const v = idx / C;
  int const c = (int)(idx - v * C);
  x[idx] = c == 0 ? f2bf(src[v] * scale) : (uint16_t)0;
}

// channel 0 of a bf16 tensor -> fp32 field times scale (plus base if given)
__global__ void k_unpack_output(uint16_t const *__restrict__ y, float *__restrict__ dst,
                                float const *__restrict__ base, int64_t nvox, int C,
                                float scale) {
  int64_t const v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  float const o = bf2f(y[v * C]) * scale;
  dst[v] = base ? base[v] + o : o;
}

// L2 loss on channel 0: partial sums of (y - t)^2 per block, and
// dy = 2 (y - t) / nvox on channel 0 (zero elsewhere)
__global__ __launch_bounds__(UT) void k_l2_loss(uint16_t const *__restrict__ y,
                                               float const *__restrict__ t,
                                               uint16_t *__restrict__ dy,
                                               float *__restrict__ part, int64_t nvox, int C,
                                               float tscale) {
  __shared__ float red[UT];
  int64_t const v = (int64_t)blockIdx.x * UT + threadIdx.x;
  float e2 = 0.f;
  if (v < nvox) {
    float const e = bf2f(y[v * C]) - t[v] * tscale;
    e2 = e * e;
    dy[v * C] = f2bf(2.f * e / (float)nvox);
    for (int c = 1; c < C; ++c) dy[v * C + c] = 0;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = UT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on fp32 masters; writes the bf16 working copy (wb may be NULL)
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       float const *__restrict__ g, uint16_t *__restrict__ wb, int len,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  int const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float const gi = g[i];
  float const mi = b1 * m[i] + (1.f - b1) * gi;
  float const vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float const wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(float const *__restrict__ a, uint16_t *__restrict__ b, int64_t n) {
  int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

__global__ void k_from_bf16(uint16_t const *__restrict__ b, float *__restrict__ a, int64_t n) {
  int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = bf2f(b[i]);
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
#define CONV_CASES(X) \
  X(16, 16) X(16, 32) X(32, 32) X(32, 64) X(64, 64) X(96, 32) X(48, 16) X(32, 16) \
  X(64, 32) X(32, 96) X(16, 48) X(64, 96) X(16, 64)

static inline int64_t blocks_for(int64_t n, int t) { return (n + t - 1) / t; }

extern "C" {

int agdu_conv3(void const *x, void const *wt, float const *bias, void const *gate, void *y,
               int N, int D, int H, int W, int CI, int CO, int mode, void *stream) {
  if (N < 1 || D < 1 || H < 1 || W < 1 || H > 65535 || (int64_t)N * D > 65535) {
    fprintf(stderr, "ERROR: convolution shape %d x %d x %d x %d unsupported\n", N, D, H, W);
    return 1;
  }
  dim3 const grid((unsigned)((W + 63) / 64), (unsigned)H, (unsigned)(N * D));
#define CC(a, b)                                                                      \
  if (CI == a && CO == b)                                                             \
    return ulaunch("conv3", k_conv3<a, b>, grid, dim3(UT), stream, (uint16_t const *)x, \
                   (uint16_t const *)wt, bias, (uint16_t const *)gate, (uint16_t *)y, D, \
                   H, W, mode);
  CONV_CASES(CC)
#undef CC
  fprintf(stderr, "ERROR: no convolution kernel for %d -> %d channels\n", CI, CO);
  return 1;
}

int agdu_flip_weights(void const *wt, void *wtt, int CI, int CO, void *stream) {
  int const total = 27 * CI * CO;
  return ulaunch("flip_weights", k_flip_weights, dim3((unsigned)blocks_for(total, UT)),
                 dim3(UT), stream, (uint16_t const *)wt, (uint16_t *)wtt, CI, CO);
}

int64_t agdu_wgrad_slabs(int N, int D, int H, int W) {
  int64_t const nvox = (int64_t)N * D * H * W;
  return blocks_for(nvox, 4096);
}

int agdu_conv3_wgrad(void const *x, void const *dy, float *part, float *dw, int N, int D,
                     int H, int W, int CI, int CO, void *stream) {
  int64_t const nslab = agdu_wgrad_slabs(N, D, H, W);
  if (CI > 96 || CO > 64 || CI * CO > 24 * UT) {
    fprintf(stderr, "ERROR: weight gradient for %d -> %d channels unsupported\n", CI, CO);
    return 1;
  }
  if (ulaunch("conv3_wgrad", k_conv3_wgrad, dim3(27, (unsigned)nslab), dim3(UT), stream,
              (uint16_t const *)x, (uint16_t const *)dy, part, N, D, H, W, CI, CO, 4096))
    return 1;
  int const len = 27 * CI * CO;
  return ulaunch("reduce_parts", k_reduce_parts, dim3((unsigned)blocks_for(len, UT)),
                 dim3(UT), stream, (float const *)part, dw, (int)nslab, len);
}

int agdu_bias_grad(void const *dy, float *part, float *db, int64_t nvox, int CO,
                   void *stream) {
  int64_t const nslab = blocks_for(nvox, 4096);
  if (ulaunch("bias_grad", k_bias_grad, dim3((unsigned)nslab), dim3(64), stream,
              (uint16_t const *)dy, part, nvox, CO, 4096))
    return 1;
  return ulaunch("reduce_parts", k_reduce_parts, dim3((unsigned)blocks_for(CO, UT)),
                 dim3(UT), stream, (float const *)part, db, (int)nslab, CO);
}

int agdu_maxpool(void const *x, void *y, void *arg, int N, int D, int H, int W, int C,
                 void *stream) {
  if ((D | H | W) & 1) {
    fprintf(stderr, "ERROR: max pooling needs even sizes, got %d x %d x %d\n", D, H, W);
    return 1;
  }
  int64_t const total = (int64_t)N * (D / 2) * (H / 2) * (W / 2) * C;
  return ulaunch("maxpool", k_maxpool, dim3((unsigned)blocks_for(total, UT)), dim3(UT),
                 stream, (uint16_t const *)x, (uint16_t *)y, (uint8_t *)arg, N, D, H, W, C);
}

int agdu_maxpool_bwd(void const *dy, void const *arg, void *dx, int N, int D, int H, int W,
                     int C, void *stream) {
  if ((D | H | W) & 1) {
    fprintf(stderr, "ERROR: max pooling needs even sizes, got %d x %d x %d\n", D, H, W);
    return 1;
  }
  int64_t const total = (int64_t)N * D * H * W * C;
  return ulaunch("maxpool_bwd", k_maxpool_bwd, dim3((unsigned)blocks_for(total, UT)),
                 dim3(UT), stream, (uint16_t const *)dy, (uint8_t const *)arg,
                 (uint16_t *)dx, N, D, H, W, C);
}

int agdu_upsample(void const *x, void *y, int N, int D, int H, int W, int C, int Cy, int c0,
                  void *stream) {
  int64_t const total = (int64_t)N * 8 * D * H * W * C;
  return ulaunch("upsample", k_upsample, dim3((unsigned)blocks_for(total, UT)), dim3(UT),
                 stream, (uint16_t const *)x, (uint16_t *)y, N, D, H, W, C, Cy, c0);
}

int agdu_upsample_bwd(void const *dy, void *dx, int N, int D, int H, int W, int C, int Cy,
                      int c0, void *stream) {
  int64_t const total = (int64_t)N * D * H * W * C;
  return ulaunch("upsample_bwd", k_upsample_bwd, dim3((unsigned)blocks_for(total, UT)),
                 dim3(UT), stream, (uint16_t const *)dy, (uint16_t *)dx, N, D, H, W, C, Cy,
                 c0);
}

int agdu_copy_channels(void const *x, void *y, int64_t nvox, int Cx, int cx0, int Cy,
                       int cy0, int C, int mode, void const *gate, void *stream) {
  return ulaunch("copy_channels", k_copy_channels,
                 dim3((unsigned)blocks_for(nvox * C, UT)), dim3(UT), stream,
                 (uint16_t const *)x, (uint16_t *)y, nvox, Cx, cx0, Cy, cy0, C, mode,
                 (uint16_t const *)gate);
}

int agdu_pack_input(float const *src, void *x, int64_t nvox, int C, float scale,
                    void *stream) {
  return ulaunch("pack_input", k_pack_input, dim3((unsigned)blocks_for(nvox * C, UT)),
                 dim3(UT), stream, src, (uint16_t *)x, nvox, C, scale);
}

int agdu_unpack_output(void const *y, float *dst, float const *base, int64_t nvox, int C,
                       float scale, void *stream) {
  return ulaunch("unpack_output", k_unpack_output, dim3((unsigned)blocks_for(nvox, UT)),
                 dim3(UT), stream, (uint16_t const *)y, dst, base, nvox, C, scale);
}

int agdu_l2_loss(void const *y, float const *t, void *dy, float *part, int64_t nvox, int C,
                 float tscale, void *stream) {
  return ulaunch("l2_loss", k_l2_loss, dim3((unsigned)blocks_for(nvox, UT)), dim3(UT),
                 stream, (uint16_t const *)y, t, (uint16_t *)dy, part, nvox, C, tscale);
}

int64_t agdu_loss_parts(int64_t nvox) { return blocks_for(nvox, UT); }

int agdu_adam(float *w, float *m, float *v, float const *g, void *wb, int len, float lr,
              float b1, float b2, float eps, int step, void *stream) {
  float const c1 = 1.f - powf(b1, (float)step);
  float const c2 = 1.f - powf(b2, (float)step);
  return ulaunch("adam", k_adam, dim3((unsigned)blocks_for(len, UT)), dim3(UT), stream, w, m,
                 v, g, (uint16_t *)wb, len, lr, b1, b2, eps, c1, c2);
}

int agdu_to_bf16(float const *a, void *b, int64_t n, void *stream) {
  return ulaunch("to_bf16", k_to_bf16, dim3((unsigned)blocks_for(n, UT)), dim3(UT), stream,
                 a, (uint16_t *)b, n);
}

int agdu_from_bf16(void const *b, float *a, int64_t n, void *stream) {
  return ulaunch("from_bf16", k_from_bf16, dim3((unsigned)blocks_for(n, UT)), dim3(UT),
                 stream, (uint16_t const *)b, a, n);
}

}  // extern "C"
