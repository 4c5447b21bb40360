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
// output channels with v_mfma_f32_16x16x32_bf16 (M = voxels, N = 16 output
// channels, K = the 16 input channels of two taps per instruction). The input
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
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16], e < 8, as 8 bf16 in a u32x4; D as for K = 16
__device__ __forceinline__ f32x4 mfma32(u32x4 a, u32x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

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
// Block tile: DB consecutive d planes x CT_H rows (h) x CT_W voxels (w) of
// one n, all output channels; one wave per row, 4 matrix-core tiles of 16
// voxels each. Per 16-channel input chunk the block stages that chunk of the
// weights ([27][CO][16]) once and walks the DB planes with a ring of three
// input planes in LDS ((CT_H + 2) rows x (CT_W + 2) voxels x 16 channels
// each; one new plane per step), so the input is read (DB + 2) / DB times
// instead of three. Each matrix-core step takes two taps: K = 32 is the 16
// channels of tap 2 s (lanes 0-31) and of tap 2 s + 1 (lanes 32-63), 14
// steps for the 27 taps (the last one's second tap has zero weights); both
// operands are 16-byte LDS reads per lane. Staging loads are branch-free
// (clamped address, zero selected afterwards) so a thread's loads stay in
// flight together.
// grid: (ceil(W / CT_W), ceil(H / CT_H), N * ceil(D / DB)); block 256.
// mode bits: 1 = add bias, 2 = ReLU on the output, 4 = multiply the output
// by (gate > 0) (the ReLU mask of the layer below, for input gradients).
#define CT_W 64
#define CT_H 4

// LDS rows of 16 bf16 channels (32 bytes) with their four 8-byte quarters
// rotated by (row >> 2) & 3: the 16 lanes of a matrix-core operand read
// (16 consecutive rows, one quarter) then fall in 16 distinct bank pairs
// instead of 4-way conflicts
__device__ __forceinline__ int swz_q(int row, int q) { return row * 16 + ((q ^ ((row >> 2) & 3)) << 2); }
__device__ __forceinline__ bf16x4 ld_swz(uint16_t const *base, int row, int q) {
  return *(bf16x4 const *)(base + swz_q(row, q));
}
// store 8 channels (quarters 2 half, 2 half + 1) of row
__device__ __forceinline__ void st_swz(uint16_t *base, int row, int half, uint4 v) {
  *(uint2 *)(base + swz_q(row, 2 * half)) = make_uint2(v.x, v.y);
  *(uint2 *)(base + swz_q(row, 2 * half + 1)) = make_uint2(v.z, v.w);
}
// The K = 32 operands read 16 bytes (8 channels, one half of a row) per
// lane. The LDS serves 128 bytes (32 banks) per cycle, i.e. 8 lanes of a
// 16-byte read: 8 consecutive rows of 32 bytes, one half each, span 256
// bytes, and rows r and r + 4 fall on the same banks; swapping the two
// halves of a row when (row >> 2) & 1 sends them to different banks, so
// every 8 lanes cover the 32 banks once
__device__ __forceinline__ int swz_h(int row, int half) {
  return row * 16 + ((half ^ ((row >> 2) & 1)) << 3);
}
__device__ __forceinline__ u32x4 ld_h(uint16_t const *base, int row, int half) {
  return *(u32x4 const *)(base + swz_h(row, half));
}
__device__ __forceinline__ void st_h(uint16_t *base, int row, int half, uint4 v) {
  *(uint4 *)(base + swz_h(row, half)) = v;
}

template <int CO, int CU, bool UPC = false>
struct ConvDepth {
  // depth planes per block: the ring saves input reads, the accumulators
  // (DB x 4 x CO / 16 tiles) cost registers
  static constexpr int DB = CO <= 16 ? ((CU && !UPC) ? 2 : 4) : (CO <= 32 ? 2 : 1);
};

// waves per SIMD the registers are bounded for: two where the LDS allows two
// blocks per CU, else one (the 64-channel layers, d2a on combined taps)
template <int CO, bool UPC>
struct ConvWaves {
  static constexpr int W = (CO >= 64 || (UPC && CO >= 32)) ? 1 : 2;
};

// Combined taps of a 3x3x3 convolution over a nearest x2 upsampled input.
// Along one axis, output index o = 2 c + p reads input o + k - 1 (k = 0, 1,
// 2), i.e. coarse c + floor((p + k - 1) / 2): for p = 0 the coarse offsets
// are -1 (k = 0) and 0 (k = 1, 2); for p = 1 they are 0 (k = 0, 1) and +1
// (k = 2). So per output parity the three taps become two, taps 0 and 1+2
// (p = 0) or 0+1 and 2 (p = 1), and the two read the input rows o - 1 + p
// and o + p. Applied along d and h (whose parities are uniform per plane
// and per row of a block), the 27 taps become 4 (a, b) pairs x 3 w taps for
// each of the 4 (pd, ph) parities: 48 combined taps, entry
// ((pd * 2 + ph) * 4 + a * 2 + b) * 3 + kw.
__host__ __device__ constexpr int up_tap(int pd, int ph, int a, int b, int kw) {
  return (((pd * 2 + ph) * 4 + a * 2 + b) * 3) + kw;
}

// CU > 0: the input is the concatenation [upsample(xu), x] without being
// formed: input channels [0, CU) are xu's ([n][D/2][H/2][W/2][CU], nearest
// x2 upsampling, D, H, W even) and [CU, CI) are x's ([n][D][H][W][CI - CU]).
// UPC: the xu channels use the 48 combined taps wcu ([48][CO][CU], built by
// k_up_weights from wt) instead of wt's 27: per 16-channel chunk 12
// combined taps, 6 matrix-core steps of two, instead of 27 taps in 14.
template <int CI, int CO, int CU, bool UPC = false>
__global__ __launch_bounds__(UT, (ConvWaves<CO, UPC>::W)) void k_conv3(
    uint16_t const *__restrict__ x, uint16_t const *__restrict__ xu,
    uint16_t const *__restrict__ wt, uint16_t const *__restrict__ wcu,
    float const *__restrict__ bias, uint16_t const *__restrict__ gate,
    uint16_t *__restrict__ y, uint16_t *__restrict__ pool, int D, int H, int W, int mode) {
  static_assert(CI % 16 == 0 && CO % 16 == 0 && CU % 16 == 0 && CU < CI,
                "channels in multiples of 16");
  static_assert(!UPC || CU > 0, "combined taps serve the upsampled channels");
  constexpr int CS = CI - CU;  // channels of x
  constexpr int NT = CO / 16;
  constexpr int DB = ConvDepth<CO, CU, UPC>::DB;
  constexpr int PW = CT_W + 2, PH = CT_H + 2;
  constexpr int PLANE = PH * PW * 16;  // bf16 elements of one staged plane
  constexpr int NPI = PH * PW * 2;     // staging items (8 channels) per plane
  constexpr int NWI = 27 * CO * 2;
  constexpr int NWU = 48 * CO * 2;     // staging items of the combined taps
  constexpr int NTW = UPC ? 48 : 27;   // taps the weight region holds
  constexpr int OBS = CO + 4;  // output staging row: CO channels + 4 (bank spread)
  // output staging: 4 x 64 rows of OBS, then the pooling partials [2][32][CO]
  constexpr int OSM = 4 * 64 * OBS + 2 * 32 * CO;
  constexpr int SMEM = 3 * PLANE + NTW * CO * 16 > OSM ? 3 * PLANE + NTW * CO * 16 : OSM;
  __shared__ __align__(16) uint16_t smem[SMEM];
  uint16_t *const sx = smem;               // ring of 3 input planes
  uint16_t *const sw = smem + 3 * PLANE;   // weights of the chunk
  int const tid = threadIdx.x;
  int const lane = tid & 63;
  int const wave = tid >> 6;
  int const ndb = (D + DB - 1) / DB;
  int const n = blockIdx.z / ndb;
  int const d0 = (blockIdx.z - n * ndb) * DB;
  int const h0 = blockIdx.y * CT_H;
  int const w0 = blockIdx.x * CT_W;
  int const i = lane & 15;   // A row: voxel; B / D column: channel in the tile
  int const kq = lane >> 4;  // k group: channels 8 (kq & 1) .. + 7 of the step's tap kq >> 1
  f32x4 acc[DB][4][NT];
#pragma unroll
  for (int j = 0; j < DB; ++j)
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[j][m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  uint16_t const *const xn = x + (int64_t)n * D * H * W * CS;
  int64_t const plane = (int64_t)H * W * CS;
  int const Hu = H / 2, Wu = W / 2;
  uint16_t const *const xun = CU ? xu + (int64_t)n * (D / 2) * Hu * Wu * CU : nullptr;
  int64_t const planeu = (int64_t)Hu * Wu * CU;
  constexpr int NPK = (NPI + UT - 1) / UT;  // staging items per thread and plane
  // per staging item: 32-bit offset within a plane (the same for every
  // plane and chunk) and whether its (h, w) lies inside the volume; items
  // past the last repeat it (same value, same place: no branch, so the
  // compiler keeps the loads ahead of the stores)
  uint32_t ioff[NPK], iofu[NPK];
  bool iok[NPK];
#pragma unroll
  for (int k = 0; k < NPK; ++k) {
    int const e0 = tid + k * UT;
    int const e = e0 < NPI ? e0 : NPI - 1;
    int const half = e & 1;
    int const r = e >> 1;
    int const pw = r % PW, ph = r / PW;
    int const hh = h0 + ph - 1, ww = w0 + pw - 1;
    iok[k] = hh >= 0 && hh < H && ww >= 0 && ww < W;
    ioff[k] = iok[k] ? (uint32_t)(hh * W + ww) * (uint32_t)CS + (uint32_t)(half * 8)
                     : (uint32_t)(half * 8);
    iofu[k] = iok[k] ? (uint32_t)((hh >> 1) * Wu + (ww >> 1)) * (uint32_t)CU + (uint32_t)(half * 8)
                     : (uint32_t)(half * 8);
  }
  // global loads of plane dd of chunk c, raw (clamped addresses) ...
  auto load_plane = [&](int dd, int c, uint4 (&v)[NPK]) {
    bool const okd = dd >= 0 && dd < D;
    bool const up = CU > 0 && c < CU / 16;  // the same for the whole block
    uint16_t const *const src = up ? xun + (okd ? dd >> 1 : 0) * planeu + c * 16
                                   : xn + (okd ? dd : 0) * plane + (c - CU / 16) * 16;
#pragma unroll
    for (int k = 0; k < NPK; ++k) v[k] = *(uint4 const *)(src + (up ? iofu[k] : ioff[k]));
  };
  // ... and their store into ring slot (dd + 3) % 3, zero outside the
  // volume (a mask, not a branch; applied here so that the loads stay in
  // flight until the store)
  auto store_plane = [&](int dd, uint4 const (&v)[NPK]) {
    bool const okd = dd >= 0 && dd < D;
    uint16_t *const dst = sx + ((dd + 3) % 3) * PLANE;
#pragma unroll
    for (int k = 0; k < NPK; ++k) {
      int const e0 = tid + k * UT;
      int const e = e0 < NPI ? e0 : NPI - 1;
      uint32_t const msk = (okd && iok[k]) ? ~0u : 0u;
      st_h(dst, e >> 1, e & 1, make_uint4(v[k].x & msk, v[k].y & msk, v[k].z & msk, v[k].w & msk));
    }
  };
  // a chunk's first three planes and its weights ([27][CO][16]), loaded
  // into registers (PRE: while the previous chunk is computed; for the
  // 64-channel layers, whose registers and LDS allow one block per CU
  // either way), then stored
  constexpr bool PRE = CO >= 64;
  constexpr int NXW = (NWI + UT - 1) / UT;
  uint4 xv[3][NPK], xw[NXW];
  auto load_chunk = [&](int c) {
#pragma unroll
    for (int q = 0; q < 3; ++q) load_plane(d0 - 1 + q, c, xv[q]);
#pragma unroll
    for (int k = 0; k < NXW; ++k) {
      int const e0 = tid + k * UT;
      int const e = e0 < NWI ? e0 : NWI - 1;
      xw[k] = *(uint4 const *)(wt + (int64_t)(e >> 1) * CI + c * 16 + (e & 1) * 8);  // tap * CO + co
    }
  };
  auto store_chunk = [&]() {
#pragma unroll
    for (int k = 0; k < NXW; ++k) {
      int const e0 = tid + k * UT;
      int const e = e0 < NWI ? e0 : NWI - 1;
      st_h(sw, e >> 1, e & 1, xw[k]);  // past the end: the last item again
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) store_plane(d0 - 1 + q, xv[q]);
  };
  // operands of one matrix-core step (two taps): A for the 4 voxel tiles,
  // B for the NT output channel tiles
  struct Ops {
    u32x4 a[4];
    u32x4 b[NT];
  };
  // UPC: the upsampled channels, on the 48 combined taps ([48][CO][CU]) and
  // staged at their own (coarse) resolution, the whole block's region at
  // once: coarse planes d0/2 - 1 .. (d0 + DB)/2, rows h0/2 - 1 ..
  // (h0 + CT_H)/2, columns w0/2 - 1 .. (w0 + CT_W)/2 (d0, h0, w0 even),
  // row (cd * CPH + ch) * CPW + cw of the ring's LDS. They are computed
  // after the x channels, whose last planes' compute overlaps the loads of
  // the first coarse chunk.
  constexpr int CPW = CT_W / 2 + 2, CPH = CT_H / 2 + 2, CPD = DB / 2 + 2;
  constexpr int NCI = CPD * CPH * CPW * 2;  // staging items (8 channels)
  constexpr int NCK = (NCI + UT - 1) / UT;
  static_assert(!UPC || (DB % 2 == 0 && CT_H % 2 == 0 && NCI * 8 <= 3 * PLANE),
                "coarse staging: whole coarse planes and rows, within the ring");
  static_assert(!UPC || NCK <= 32, "inside-the-volume bits of the coarse items");
  constexpr int NCW = (NWU + UT - 1) / UT;
  // per coarse item: its offset (the same for every chunk; 0 outside the
  // volume) and, in cok, whether it lies inside the volume
  int64_t coff[UPC ? NCK : 1];
  uint32_t cok = 0;
#pragma unroll
  for (int k = 0; k < (UPC ? NCK : 0); ++k) {
    int const e0 = tid + k * UT;
    int const e = e0 < NCI ? e0 : NCI - 1;
    int const half = e & 1, r = e >> 1;
    int const cw = r % CPW, t = r / CPW;
    int const gd = d0 / 2 - 1 + t / CPH, gh = h0 / 2 - 1 + t % CPH;
    int const gw = w0 / 2 - 1 + cw;
    bool const ok = gd >= 0 && gd < D / 2 && gh >= 0 && gh < Hu && gw >= 0 && gw < Wu;
    coff[k] = (ok ? (int64_t)gd * planeu + (int64_t)(gh * Wu + gw) * CU : 0) + half * 8;
    cok |= (ok ? 1u : 0u) << k;
  }
  // a coarse chunk's region, loaded into registers while the previous
  // chunk is computed (its combined weights, the same for every block,
  // come from L2 when stored)
  uint4 cv[UPC ? NCK : 1];
  auto load_coarse = [&](int c) {
#pragma unroll
    for (int k = 0; k < (UPC ? NCK : 0); ++k) cv[k] = *(uint4 const *)(xun + coff[k] + c * 16);
  };
  // the coarse chunks (UPC), chunk 0's region already in cv
  auto coarse_part = [&]() {
    if constexpr (UPC) {
#pragma unroll 1
      for (int c = 0; c < CU / 16; ++c) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NCW; ++k) {
          int const e0 = tid + k * UT;
          int const e = e0 < NWU ? e0 : NWU - 1;
          uint4 const v = *(uint4 const *)(wcu + (int64_t)(e >> 1) * CU + c * 16 + (e & 1) * 8);
          st_h(sw, e >> 1, e & 1, v);  // combined tap * CO + co
        }
#pragma unroll
        for (int k = 0; k < NCK; ++k) {
          int const e0 = tid + k * UT;
          int const e = e0 < NCI ? e0 : NCI - 1;
          uint32_t const msk = ((cok >> k) & 1u) ? ~0u : 0u;
          st_h(sx, e >> 1, e & 1,
               make_uint4(cv[k].x & msk, cv[k].y & msk, cv[k].z & msk, cv[k].w & msk));
        }
        __syncthreads();
        if (c + 1 < CU / 16) load_coarse(c + 1);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the compute
        int const ph = (h0 + wave) & 1;
        // Streamed over the staged coarse planes: fine plane j (parity j & 1,
        // d0 even) reads coarse plane (j + 1) / 2 + a with its d-combination
        // a, so coarse plane cp feeds fine planes 2 (cp - a) - 1 and
        // 2 (cp - a) for a = 0, 1. Each A fragment (coarse plane cp, h
        // combination b, tap kw) is read once and used for those (up to
        // four) fine planes; K = 32 pairs two of the six (b, kw) positions.
        // Per fine plane the contributions keep their order (a = 0, then
        // a = 1; positions in order), so the sums are those of one plane at
        // a time.
#pragma unroll
        for (int cp = 0; cp < CPD; ++cp) {
#pragma unroll
          for (int sp = 0; sp < 3; ++sp) {
            int const u = 2 * sp + (kq >> 1);  // (b, kw) position of this lane
            int const b = u / 3, kw = u - 3 * b;
            int const v0 = (cp * CPH + (wave + 1) / 2 + b) * CPW + 1;
            u32x4 a4[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) a4[m] = ld_h(sx, v0 + ((16 * m + i + kw - 1) >> 1), kq & 1);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
              for (int jp = 0; jp < 2; ++jp) {
                int const j = 2 * (cp - a) - 1 + jp;
                if (j >= 0 && j < DB) {
                  int const pd = j & 1;
#pragma unroll
                  for (int t = 0; t < NT; ++t) {
                    u32x4 const bt = ld_h(sw, up_tap(pd, ph, a, b, kw) * CO + t * 16 + i, kq & 1);
#pragma unroll
                    for (int m = 0; m < 4; ++m) acc[j][m][t] = mfma32(a4[m], bt, acc[j][m][t]);
                  }
                }
              }
          }
        }
      }
    }
  };
  if constexpr (PRE) load_chunk(UPC ? CU / 16 : 0);
#pragma unroll 1
  for (int c = UPC ? CU / 16 : 0; c < CI / 16; ++c) {
    __syncthreads();
    if constexpr (!PRE) load_chunk(c);
    store_chunk();
    __syncthreads();
    // PRE: the next chunk's loads go out once this chunk's plane loads are
    // done (in the last plane steps), so that waiting for a plane never
    // waits for them
    bool const pre = PRE && c + 1 < CI / 16;
    if constexpr (UPC && DB >= 4) {
      // The x channels of d1a streamed over the input planes p = d0 - 1 ..
      // d0 + DB (in ring slot (p + 3) % 3): each A fragment (plane p, two
      // (kh, kw) positions of step s) is read once and used for the output
      // planes p + 1 - kd it feeds, with the weights of tap kd 9 + position
      // from LDS; plane p + 3 is loaded while plane p is computed and stored
      // into its slot.
#pragma unroll
      for (int jp = 0; jp < DB + 2; ++jp) {
        int const p = d0 - 1 + jp;
        uint16_t const *const pl = sx + ((p + 3) % 3) * PLANE;
        uint4 nxt[NPK];
        bool const more = jp + 3 < DB + 2;  // plane p + 3 still an input
        if (more)
          load_plane(p + 3, c, nxt);
        else if (jp == DB - 1 && pre)
          load_chunk(c + 1);
        else if (jp == DB - 1 && c + 1 == CI / 16)
          load_coarse(0);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the compute
#pragma unroll
        for (int st = 0; st < 5; ++st) {
          int const q = 2 * st + (kq >> 1);
          int const qq = q < 9 ? q : 8;
          int const kh = qq / 3, kw = qq - 3 * kh;
          unsigned const keep = q < 9 ? ~0u : 0u;
          int const v0 = (wave + kh) * PW + i + kw;
          u32x4 a4[4];
#pragma unroll
          for (int m = 0; m < 4; ++m) a4[m] = ld_h(pl, v0 + 16 * m, kq & 1);
#pragma unroll
          for (int kd = 0; kd < 3; ++kd) {
            int const oi = jp - kd;  // output plane d0 + oi = p + 1 - kd
            if (oi >= 0 && oi < DB) {
#pragma unroll
              for (int tt = 0; tt < NT; ++tt) {
                u32x4 const bt = ld_h(sw, (kd * 9 + qq) * CO + tt * 16 + i, kq & 1) & keep;
#pragma unroll
                for (int m = 0; m < 4; ++m) acc[oi][m][tt] = mfma32(a4[m], bt, acc[oi][m][tt]);
              }
            }
          }
        }
        if (more) {
          __syncthreads();  // every wave is done with plane p
          store_plane(p + 3, nxt);
          __syncthreads();
        }
      }
      continue;
    }
#pragma unroll
    for (int j = 0; j < DB; ++j) {
      int const d = d0 + j;
      // plane d + 2 is loaded while plane d is computed, then stored into
      // the slot of plane d - 1
      uint4 nxt[NPK];
      if (j + 1 < DB)
        load_plane(d + 2, c, nxt);
      else if (pre)
        load_chunk(c + 1);
      else if (UPC && c + 1 == CI / 16)
        load_coarse(0);
      __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the compute
      // operands of step st: this lane's tap 2 st + (kq >> 1) (the 28th,
      // past the last, reads tap 26 with zero weights), double-buffered so
      // the LDS reads of the next step overlap this step's MFMAs
      auto load_ops = [&](int st, Ops &o) {
        int const tl = 2 * st + (kq >> 1);
        int const t = tl < 27 ? tl : 26;
        int const kd = t / 9, kh = t / 3 - 3 * kd, kw = t - 3 * (t / 3);
        uint16_t const *const pl = sx + ((d + kd - 1 + 3) % 3) * PLANE;
        int const v0 = (wave + kh) * PW + i + kw;
        unsigned const keep = tl < 27 ? ~0u : 0u;
#pragma unroll
        for (int m = 0; m < 4; ++m) o.a[m] = ld_h(pl, v0 + 16 * m, kq & 1);
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) o.b[tt] = ld_h(sw, t * CO + tt * 16 + i, kq & 1) & keep;
      };
      auto mma_ops = [&](Ops const &o) {
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
#pragma unroll
          for (int m = 0; m < 4; ++m) acc[j][m][tt] = mfma32(o.a[m], o.b[tt], acc[j][m][tt]);
      };
      Ops opa, opb;
      load_ops(0, opa);
#pragma unroll 1
      for (int st = 1; st < 14; st += 2) {
        load_ops(st, opb);
        mma_ops(opa);
        if (st + 1 < 14) load_ops(st + 1, opa);
        mma_ops(opb);
      }
      if (j + 1 < DB) {
        __syncthreads();  // every wave is done with plane d - 1
        store_plane(d + 2, nxt);
        __syncthreads();
      }
    }
  }
  coarse_part();
  // Epilogue per plane: each wave writes its row (bias, ReLU applied) into
  // its own LDS rows [voxel][OBS] (lane (i, kq) holds voxels 4 kq + r of
  // channel t * 16 + i; the row padding puts the four kq groups in distinct
  // banks), then stores 16-byte pieces of 8 channels, coalesced.
  int const h = h0 + wave;
  uint16_t *const ob = smem + wave * 64 * OBS;
#pragma unroll
  for (int j = 0; j < DB; ++j) {
    int const d = d0 + j;
    __syncthreads();  // the LDS is free (j = 0: done with the operands)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      int const co = t * 16 + i;
      float const b = (mode & 1) ? bias[co] : 0.f;
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float v = acc[j][m][t][r] + b;
          if (mode & 2) v = fmaxf(v, 0.f);
          ob[(16 * m + 4 * kq + r) * OBS + co] = f2bf(v);
        }
    }
    __syncthreads();
    if (h < H && d < D) {
      int64_t const orow = (((int64_t)n * D + d) * H + h) * W;
      constexpr int C8 = CO / 8;
#pragma unroll
      for (int k = 0; k < (64 * C8 + 63) / 64; ++k) {
        int const e = lane + 64 * k;
        int const v = e / C8, c8 = e - v * C8;
        int const wv = w0 + v;
        if (e < 64 * C8 && wv < W) {
          uint2 const lo = *(uint2 const *)(ob + v * OBS + c8 * 8);
          uint2 const hi = *(uint2 const *)(ob + v * OBS + c8 * 8 + 4);
          uint4 val = make_uint4(lo.x, lo.y, hi.x, hi.y);
          int64_t const o = (orow + wv) * CO + c8 * 8;
          if (mode & 4) {
            uint4 const g = *(uint4 const *)(gate + o);
            uint32_t const gw[4] = {g.x, g.y, g.z, g.w};
            uint32_t vw[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              uint32_t const m0 = bf2f((uint16_t)(gw[q] & 0xffffu)) > 0.f ? 0xffffu : 0u;
              uint32_t const m1 = bf2f((uint16_t)(gw[q] >> 16)) > 0.f ? 0xffff0000u : 0u;
              vw[q] &= m0 | m1;
            }
            val = make_uint4(vw[0], vw[1], vw[2], vw[3]);
          }
          *(uint4 *)(y + o) = val;
        }
      }
    }
    if constexpr (DB % 2 == 0) {
      if (mode & 8) {
        // 2x2x2 max pooling of the output (D, H, W even; d0, h0, w0 even):
        // item (pooled row pr, column pc, 8 channels) scans its voxels in
        // (d, h, w) order with a strict >, as the pooling kernel does; the
        // even plane's partial (one of the bf16 values, so exact) waits in
        // LDS for the odd plane, handled by the same thread
        constexpr int C8 = CO / 8;
        uint16_t *const pm = smem + 4 * 64 * OBS;
        for (int e = tid; e < 2 * 32 * C8; e += UT) {
          int const c8 = e % C8, pc = (e / C8) % 32, pr = e / (C8 * 32);
          float best[8];
          if (j % 2 == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) best[k] = -INFINITY;
          } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) best[k] = bf2f(pm[(pr * 32 + pc) * CO + c8 * 8 + k]);
          }
#pragma unroll
          for (int o = 0; o < 4; ++o) {
            uint16_t const *const r =
                smem + (2 * pr + (o >> 1)) * 64 * OBS + (2 * pc + (o & 1)) * OBS + c8 * 8;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              float const f = bf2f(r[k]);
              if (f > best[k]) best[k] = f;
            }
          }
          if (j % 2 == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) pm[(pr * 32 + pc) * CO + c8 * 8 + k] = f2bf(best[k]);
          } else {
            int const dp = d / 2, hp = h0 / 2 + pr, wp = w0 / 2 + pc;
            if (d < D && 2 * hp < H && 2 * wp < W) {
              uint32_t q[4];
#pragma unroll
              for (int k = 0; k < 4; ++k)
                q[k] = (uint32_t)f2bf(best[2 * k]) | ((uint32_t)f2bf(best[2 * k + 1]) << 16);
              int64_t const po =
                  ((((int64_t)n * (D / 2) + dp) * (H / 2) + hp) * (W / 2) + wp) * CO + c8 * 8;
              *(uint4 *)(pool + po) = make_uint4(q[0], q[1], q[2], q[3]);
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 16 -> 16 channels, streaming over the input planes: each input plane is
// staged into LDS once (double-buffered) and each of its A fragments is read
// once and used for the three depth taps, i.e. for the three output planes
// the plane touches (kd = 0, 1, 2 into outputs p + 1, p, p - 1), with the
// weights of all 27 taps in registers. K = 32 pairs two (kh, kw) positions
// of one input plane: step s takes positions 2 s and 2 s + 1 (the tenth
// position, past the ninth, reads the ninth with zero weights). Per 16
// voxels and input plane this reads 5 A fragments from LDS and no B, where
// k_conv3 reads 14 A and 14 B fragments per 16 voxels and output plane.
// Block: CT_H rows x CT_W voxels of DB output planes; without pooling one
// row per wave, with it (POOL) two rows x 32 voxels per wave (wave w: rows
// 2 (w >> 1) + (m >> 1), voxels 32 (w & 1) + 16 (m & 1) + 0..15 of its
// 16-voxel tiles m), so that a lane holds every voxel of its 2 x 2 pooling
// windows (the w pair in one tile, the h pair in tiles m and m + 2).
// Epilogue per output plane: bias and ReLU (mode bits 1, 2), the wave's
// voxels through its own LDS rows as 16-byte coalesced stores; POOL (mode
// bit 8): the 2x2x2 max pooling (k_maxpool's scan order and strict >) in
// registers, the even plane's partial kept for the odd one (D, H, W even,
// DB even).
// grid: (ceil(W / CT_W), ceil(H / CT_H), N * ceil(D / DB)); block 256.
// ---------------------------------------------------------------------------
template <int DB, bool POOL>
__global__ __launch_bounds__(UT, 2) void k_conv3s(uint16_t const *__restrict__ x,
                                               uint16_t const *__restrict__ wt,
                                               float const *__restrict__ bias,
                                               uint16_t *__restrict__ y,
                                               uint16_t *__restrict__ pool, int D, int H, int W,
                                               int mode) {
  static_assert(DB % 2 == 0, "pooling pairs the planes of a block");
  constexpr int PW = CT_W + 2, PH = CT_H + 2;
  constexpr int PLANE = PH * PW * 16;  // bf16 elements of one staged plane
  constexpr int NPI = PH * PW * 2;     // staging items (8 channels) per plane
  constexpr int NPK = (NPI + UT - 1) / UT;
  constexpr int OBS = 16 + 4;               // output row: 16 channels + 4 (bank spread)
  constexpr int ROWS = CT_H * 64 * OBS;     // one output plane's rows
  static_assert(CT_H == 4 && CT_W == 64 && UT == 256, "four waves of 64 voxels");
  __shared__ __align__(16) uint16_t sx[2 * PLANE];
  __shared__ __align__(16) uint16_t so[2 * ROWS];
  int const tid = threadIdx.x;
  int const lane = tid & 63;
  int const wave = tid >> 6;
  int const ndb = (D + DB - 1) / DB;
  int const n = blockIdx.z / ndb;
  int const d0 = (blockIdx.z - n * ndb) * DB;
  int const h0 = blockIdx.y * CT_H;
  int const w0 = blockIdx.x * CT_W;
  int const i = lane & 15;   // A row: voxel; B / D column: output channel
  int const kq = lane >> 4;  // channels 8 (kq & 1) .. + 7 of position 2 s + (kq >> 1)
  uint16_t const *const xn = x + (int64_t)n * D * H * W * 16;
  int64_t const plane = (int64_t)H * W * 16;
  // the weights of every tap: bw[kd][s] = w[kd 9 + 2 s + (kq >> 1)][i][8 (kq & 1) ..]
  u32x4 bw[3][5];
#pragma unroll
  for (int kd = 0; kd < 3; ++kd)
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      int const q = 2 * s + (kq >> 1);
      u32x4 const v = *(u32x4 const *)(wt + ((int64_t)(kd * 9 + (q < 9 ? q : 8)) * 16 + i) * 16 +
                                       8 * (kq & 1));
      unsigned const keep = q < 9 ? ~0u : 0u;
      bw[kd][s] = v & keep;
    }
  // staging items: 32-bit offset within a plane and whether (h, w) lies
  // inside the volume; items past the last repeat it (no branch)
  uint32_t ioff[NPK];
  bool iok[NPK];
#pragma unroll
  for (int k = 0; k < NPK; ++k) {
    int const e0 = tid + k * UT;
    int const e = e0 < NPI ? e0 : NPI - 1;
    int const half = e & 1, r = e >> 1;
    int const pw = r % PW, ph = r / PW;
    int const hh = h0 + ph - 1, ww = w0 + pw - 1;
    iok[k] = hh >= 0 && hh < H && ww >= 0 && ww < W;
    ioff[k] = iok[k] ? (uint32_t)(hh * W + ww) * 16u + (uint32_t)(half * 8) : (uint32_t)(half * 8);
  }
  auto load_plane = [&](int dd, uint4 (&v)[NPK]) {
    bool const okd = dd >= 0 && dd < D;
    uint16_t const *const src = xn + (okd ? dd : 0) * plane;
#pragma unroll
    for (int k = 0; k < NPK; ++k) {
      uint32_t const msk = (okd && iok[k]) ? ~0u : 0u;
      uint4 const t = *(uint4 const *)(src + ioff[k]);
      v[k] = make_uint4(t.x & msk, t.y & msk, t.z & msk, t.w & msk);
    }
  };
  auto store_plane = [&](uint16_t *dst, uint4 const (&v)[NPK]) {
#pragma unroll
    for (int k = 0; k < NPK; ++k) {
      int const e0 = tid + k * UT;
      int const e = e0 < NPI ? e0 : NPI - 1;
      st_h(dst, e >> 1, e & 1, v[k]);  // past the end: the last item again
    }
  };
  f32x4 acc[DB][4];
#pragma unroll
  for (int j = 0; j < DB; ++j)
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[j][m] = f32x4{0.f, 0.f, 0.f, 0.f};
  // row and first voxel (within the block) of the wave's tile m
  auto trow = [&](int m) { return POOL ? 2 * (wave >> 1) + (m >> 1) : wave; };
  auto tcol = [&](int m) { return POOL ? 32 * (wave & 1) + 16 * (m & 1) : 16 * m; };
  // loaded once: a load in the loop would be issued after the next plane's
  // prefetch, and waiting for it would wait for the prefetch too
  float const b = (mode & 1) ? bias[i] : 0.f;
  // POOL: the even plane's pooling partials of this lane, per (tile column
  // mc, voxel pair rp) (one of the bf16 values, so exact)
  float pp[2][2] = {{-INFINITY, -INFINITY}, {-INFINITY, -INFINITY}};
  // output plane oi (d0 + oi): bias, ReLU, into the wave's rows of buffer
  // oi & 1, then stored by the wave itself; POOL: its pooling
  auto epilogue = [&](int oi) {
    int const d = d0 + oi;
    uint16_t *const ob = so + (oi & 1) * ROWS + wave * 64 * OBS;
    float fv[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = acc[oi][m][r] + b;
        if (mode & 2) v = fmaxf(v, 0.f);
        uint16_t const q = f2bf(v);
        fv[m][r] = bf2f(q);
        ob[(16 * m + 4 * kq + r) * OBS + i] = q;
      }
    __builtin_amdgcn_wave_barrier();
    if (d < D) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        int const e = lane + 64 * k;  // (voxel of the wave, 8-channel half)
        int const v = e >> 1, c8 = e & 1;
        int const hh = h0 + trow(v >> 4), ww = w0 + tcol(v >> 4) + (v & 15);
        if (hh < H && ww < W) {
          uint2 const lo = *(uint2 const *)(ob + v * OBS + c8 * 8);
          uint2 const hi = *(uint2 const *)(ob + v * OBS + c8 * 8 + 4);
          *(uint4 *)(y + ((((int64_t)n * D + d) * H + hh) * W + ww) * 16 + c8 * 8) =
              make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
      }
    }
    if constexpr (POOL) {
      // pooling window (tile column mc, voxel pair rp): rows m = mc, mc + 2
      // (h order), voxels 2 rp, 2 rp + 1 (w order), after the even plane's
      bool const odd = oi & 1;
#pragma unroll
      for (int mc = 0; mc < 2; ++mc)
#pragma unroll
        for (int rp = 0; rp < 2; ++rp) {
          float best = odd ? pp[mc][rp] : -INFINITY;
#pragma unroll
          for (int o = 0; o < 4; ++o) {
            float const f = fv[2 * (o >> 1) + mc][2 * rp + (o & 1)];
            if (f > best) best = f;
          }
          if (!odd) {
            pp[mc][rp] = best;
          } else {
            int const dp = d / 2, hp = h0 / 2 + (wave >> 1);
            int const wp = w0 / 2 + 16 * (wave & 1) + 8 * mc + 2 * kq + rp;
            if (d < D && 2 * hp < H && 2 * wp < W)
              pool[((((int64_t)n * (D / 2) + dp) * (H / 2) + hp) * (W / 2) + wp) * 16 + i] =
                  f2bf(best);
          }
        }
    }
  };
  // AH2: two planes in flight per block (nxt, nx2), since the loads a CU
  // keeps in flight set this layer's HBM rate (with pooling the registers
  // would spill: one plane)
  constexpr bool AH2 = !POOL;
  uint4 nxt[NPK], nx2[AH2 ? NPK : 1];
  {
    uint4 v0[NPK];
    load_plane(d0 - 1, v0);
    load_plane(d0, nxt);
    if constexpr (AH2) load_plane(d0 + 1, nx2);
    store_plane(sx, v0);
  }
  __syncthreads();
  // input plane p = d0 - 1 + j is in buffer j & 1; nxt holds plane p + 1,
  // stored into the other buffer (free: every wave passed the barrier after
  // reading it); then (AH2) nx2's plane p + 2 moves to nxt while plane p + 3
  // is loaded into nx2, or plane p + 2 is loaded into nxt
#pragma unroll
  for (int j = 0; j < DB + 2; ++j) {
    uint16_t const *const cur = sx + (j & 1) * PLANE;
    if (j + 1 < DB + 2) {
      store_plane(sx + ((j + 1) & 1) * PLANE, nxt);
      if constexpr (AH2) {
#pragma unroll
        for (int k = 0; k < NPK; ++k) nxt[k] = nx2[k];
        if (j + 3 < DB + 2) load_plane(d0 + 2 + j, nx2);
      } else if (j + 2 < DB + 2) {
        load_plane(d0 + 1 + j, nxt);
      }
      __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the compute
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      int const q = 2 * s + (kq >> 1);
      int const qq = q < 9 ? q : 8;
      int const kh = qq / 3, kw = qq - 3 * (qq / 3);
      u32x4 a[4];
#pragma unroll
      for (int m = 0; m < 4; ++m)
        a[m] = ld_h(cur, (trow(m) + kh) * PW + tcol(m) + i + kw, kq & 1);
#pragma unroll
      for (int kd = 0; kd < 3; ++kd) {
        int const oi = j - kd;  // output plane d0 + oi = p + 1 - kd
        if (oi >= 0 && oi < DB) {
#pragma unroll
          for (int m = 0; m < 4; ++m) acc[oi][m] = mfma32(a[m], bw[kd][s], acc[oi][m]);
        }
      }
    }
    if (j >= 2) epilogue(j - 2);
    __syncthreads();
  }
}

// Combined taps of the first CU input channels of wt ([27][CO][CI]):
// wcu[up_tap(pd, ph, a, b, kw)][co][cu] = bf16(sum over the d taps kd of
// (pd, a) and the h taps kh of (ph, b) of wt[(kd * 3 + kh) * 3 + kw][co][cu]),
// summed in fp32 in tap order. Taps of parity p and pair index a: p = 0:
// {0}, {1, 2}; p = 1: {0, 1}, {2}.
__global__ __launch_bounds__(UT) void k_up_weights(uint16_t const *__restrict__ wt,
                                                   uint16_t *__restrict__ wcu, int CI, int CO,
                                                   int CU) {
  int const n = 48 * CO * CU;
  for (int e = blockIdx.x * UT + threadIdx.x; e < n; e += gridDim.x * UT) {
    int const cu = e % CU;
    int const co = (e / CU) % CO;
    int const tap = e / (CU * CO);
    int const kw = tap % 3, ab = (tap / 3) % 4, pp = tap / 12;
    int const pd = pp >> 1, ph = pp & 1, a = ab >> 1, b = ab & 1;
    int const dlo = (pd == 0) ? (a == 0 ? 0 : 1) : (a == 0 ? 0 : 2);
    int const dhi = (pd == 0) ? (a == 0 ? 0 : 2) : (a == 0 ? 1 : 2);
    int const hlo = (ph == 0) ? (b == 0 ? 0 : 1) : (b == 0 ? 0 : 2);
    int const hhi = (ph == 0) ? (b == 0 ? 0 : 2) : (b == 0 ? 1 : 2);
    float acc = 0.f;
    for (int kd = dlo; kd <= dhi; ++kd)
      for (int kh = hlo; kh <= hhi; ++kh)
        acc += bf2f(wt[((int64_t)((kd * 3 + kh) * 3 + kw) * CO + co) * CI + cu]);
    wcu[e] = f2bf(acc);
  }
}

// ---------------------------------------------------------------------------
// First layer: one live input channel (the field), so the taps are the GEMM's
// K: D[voxel][co] = sum_tap bf16(field[voxel + tap] * scale) w[tap][co][0],
// K = 27 taps padded to 32 (one K = 32 step) instead of 27 steps of 16
// channels. Block: CT_H rows x CT_W voxels of FDB consecutive d planes, a
// ring of three bf16 planes of the field in LDS; output as in k_conv3
// (bias, ReLU, 16-byte stores through per-wave LDS rows).
// grid: (ceil(W / CT_W), ceil(H / CT_H), N * ceil(D / FDB)); block 256.
// ---------------------------------------------------------------------------
#define FDB 16
__global__ __launch_bounds__(UT) void k_conv_first(float const *__restrict__ src,
                                                   uint16_t const *__restrict__ wt, int wci,
                                                   float const *__restrict__ bias,
                                                   uint16_t *__restrict__ y, int D, int H,
                                                   int W, float scale) {
  constexpr int PW = CT_W + 2, PH = CT_H + 2, PLANE = PH * PW;
  constexpr int OBS = 20;
  __shared__ __align__(16) uint16_t sx[3 * PLANE];
  __shared__ __align__(16) uint16_t ob[4 * 64 * OBS];
  int const tid = threadIdx.x;
  int const lane = tid & 63, wave = tid >> 6;
  int const i = lane & 15, kq = lane >> 4;
  int const ndb = (D + FDB - 1) / FDB;
  int const n = blockIdx.z / ndb;
  int const d0 = (blockIdx.z - n * ndb) * FDB;
  int const h0 = blockIdx.y * CT_H, w0 = blockIdx.x * CT_W;
  float const *const fn = src + (int64_t)n * D * H * W;
  // B: taps 8 kq + e of output channel i (zero past tap 26): K = 32
  unsigned bh[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    int const t = 8 * kq + e;
    bh[e] = t < 27 ? (unsigned)wt[((int64_t)t * 16 + i) * wci] : 0u;
  }
  u32x4 const bw = {bh[0] | bh[1] << 16, bh[2] | bh[3] << 16, bh[4] | bh[5] << 16,
                    bh[6] | bh[7] << 16};
  // LDS offsets of this lane's A taps relative to its voxel (slot, row, col)
  int toff[8], tkd[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    int const t = 8 * kq + e;
    int const tt = t < 27 ? t : 13;  // padded taps read the centre (weight 0)
    tkd[e] = tt / 9;
    toff[e] = ((tt / 3) % 3) * PW + tt % 3;
  }
  // a plane's field values: raw loads at clamped offsets, converted and
  // zeroed outside the volume when stored, so that the next plane's loads
  // stay in flight while the current one is computed
  constexpr int NPK = (PLANE + UT - 1) / UT;
  int foff[NPK];
  bool fok[NPK];
#pragma unroll
  for (int k = 0; k < NPK; ++k) {
    int const e = min(tid + k * UT, PLANE - 1);
    int const ph = e / PW, pw = e - ph * PW;
    int const hh = h0 + ph - 1, ww = w0 + pw - 1;
    fok[k] = hh >= 0 && hh < H && ww >= 0 && ww < W;
    foff[k] = fok[k] ? hh * W + ww : 0;
  }
  auto load_plane = [&](int dd, float (&v)[NPK]) {
    float const *const src = fn + (int64_t)(dd >= 0 && dd < D ? dd : 0) * H * W;
#pragma unroll
    for (int k = 0; k < NPK; ++k) v[k] = src[foff[k]];
  };
  auto store_plane = [&](int dd, float const (&v)[NPK]) {
    uint16_t *const dst = sx + ((dd + 3) % 3) * PLANE;
    bool const okd = dd >= 0 && dd < D;
#pragma unroll
    for (int k = 0; k < NPK; ++k) {
      int const e = tid + k * UT;
      if (e < PLANE) dst[e] = (okd && fok[k]) ? f2bf(v[k] * scale) : (uint16_t)0;
    }
  };
  float const bsel = bias[i];
  // two planes in flight: nv holds plane d + 1 (stored at the top of the
  // iteration), nv2 plane d + 2 (moved to nv while plane d + 3 is loaded)
  float nv[NPK], nv2[NPK];
  load_plane(d0 - 1, nv);
  store_plane(d0 - 1, nv);
  load_plane(d0, nv);
  store_plane(d0, nv);
  load_plane(d0 + 1, nv);
  load_plane(d0 + 2, nv2);
  int const h = h0 + wave;
  uint16_t *const obw = ob + wave * 64 * OBS;
  for (int j = 0; j < FDB; ++j) {
    int const d = d0 + j;
    store_plane(d + 1, nv);
#pragma unroll
    for (int k = 0; k < NPK; ++k) nv[k] = nv2[k];
    if (j + 2 < FDB) load_plane(d + 3, nv2);
    __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the compute
    __syncthreads();
    f32x4 acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
      int const vb = wave * PW + 16 * m + i;  // the voxel's (row, col) in a plane
      unsigned ah[8];
#pragma unroll
      for (int e = 0; e < 8; ++e)
        ah[e] = sx[((d + tkd[e] - 1 + 3) % 3) * PLANE + vb + toff[e]];
      u32x4 const a = {ah[0] | ah[1] << 16, ah[2] | ah[3] << 16, ah[4] | ah[5] << 16,
                       ah[6] | ah[7] << 16};
      acc[m] = mfma32(a, bw, acc[m]);
    }
    // D: lane holds voxels 16 m + 4 kq + r of channel i
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        obw[(16 * m + 4 * kq + r) * OBS + i] = f2bf(fmaxf(acc[m][r] + bsel, 0.f));
    __syncthreads();  // the output rows are complete; the ring slot of d - 1 is free
    if (h < H && d < D) {
      int64_t const orow = (((int64_t)n * D + d) * H + h) * W;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        int const e = lane + 64 * k;  // (voxel, 8-channel half)
        int const v = e >> 1, c8 = e & 1;
        if (w0 + v < W) {
          uint2 const lo = *(uint2 const *)(obw + v * OBS + c8 * 8);
          uint2 const hi = *(uint2 const *)(obw + v * OBS + c8 * 8 + 4);
          *(uint4 *)(y + (orow + w0 + v) * 16 + c8 * 8) = make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Head: output channel 0 of a 16 -> 16 convolution, in fp32, plus the field:
// out[v] = base[v] + s (b0 + sum_tap sum_ci x[v + tap][ci] w[tap][0][ci]).
// For an output plane d, F[u][(kh, kw)] = sum_kd sum_ci x_{d+kd-1}[u][ci]
// w[(kd, kh, kw)][0][ci] over every staged voxel u (halo included) is a GEMM
// with M = voxels, K = (kd, ci) = 48 and N = 9 taps (padded to 16): two
// K = 32 MFMAs per 16 voxels (planes d - 1 and d, then plane d + 1 with the
// second half's weights zero). Then out[v] gathers F[v + (kh - 1, kw - 1)]
// [(kh, kw)], 9 terms. Block: CT_H rows x CT_W voxels of HDB consecutive d
// planes, the input planes in a ring of three (16-byte halves swizzled as in
// k_conv3), the next plane loaded while the current one is computed.
// grid: (ceil(W / CT_W), ceil(H / CT_H), N * ceil(D / HDB)); block 256.
// ---------------------------------------------------------------------------
#define HDB 16
__global__ __launch_bounds__(UT) void k_conv_head(uint16_t const *__restrict__ x,
                                                  uint16_t const *__restrict__ wt, int wco,
                                                  float const *__restrict__ bias,
                                                  float const *__restrict__ base,
                                                  float *__restrict__ dst, int D, int H, int W,
                                                  float unscale) {
  constexpr int PW = CT_W + 2, PH = CT_H + 2;
  constexpr int NV = PH * PW;           // staged voxels per plane (396)
  constexpr int NVT = (NV + 15) / 16;   // 16-voxel tiles (25)
  constexpr int PLANE = NV * 16;        // bf16 elements per staged plane
  constexpr int NPI = NV * 2;           // staging items of 8 channels
  constexpr int NPK = (NPI + UT - 1) / UT;
  // F tap-major, one row of NVP >= 16 NVT floats per tap: a lane's four
  // consecutive voxels of one tap are one 16-byte write, and NVP = 20 mod 32
  // spreads the eight lanes of a write cycle over distinct bank groups
  constexpr int NVP = 404;
  static_assert(NVP >= NVT * 16 && NVP % 32 == 20, "F rows");
  __shared__ __align__(16) uint16_t sx[3 * PLANE];
  __shared__ __align__(16) float se[9 * NVP];
  int const tid = threadIdx.x;
  int const lane = tid & 63, wave = tid >> 6;
  int const i = lane & 15, kq = lane >> 4;
  int const ndb = (D + HDB - 1) / HDB;
  int const n = blockIdx.z / ndb;
  int const d0 = (blockIdx.z - n * ndb) * HDB;
  int const h0 = blockIdx.y * CT_H, w0 = blockIdx.x * CT_W;
  uint16_t const *const xn = x + (int64_t)n * D * H * W * 16;
  int64_t const plane = (int64_t)H * W * 16;
  // B: lane (tap i, k group kq) holds input channels 8 (kq & 1) .. + 7 of
  // depth tap kd = kq >> 1 (bw0) and kd = 2 for kq < 2, zero for kq >= 2
  // (bw1), of output channel 0; taps i >= 9 zero
  u32x4 bw0, bw1;
  {
    unsigned h0w[8], h1w[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int const ci = 8 * (kq & 1) + e;
      h0w[e] = i < 9 ? (unsigned)wt[((int64_t)((kq >> 1) * 9 + i) * wco) * 16 + ci] : 0u;
      h1w[e] = (i < 9 && kq < 2) ? (unsigned)wt[((int64_t)(18 + i) * wco) * 16 + ci] : 0u;
    }
    bw0 = u32x4{h0w[0] | h0w[1] << 16, h0w[2] | h0w[3] << 16, h0w[4] | h0w[5] << 16,
                h0w[6] | h0w[7] << 16};
    bw1 = u32x4{h1w[0] | h1w[1] << 16, h1w[2] | h1w[3] << 16, h1w[4] | h1w[5] << 16,
                h1w[6] | h1w[7] << 16};
  }
  uint32_t ioff[NPK];
  bool iok[NPK];
#pragma unroll
  for (int k = 0; k < NPK; ++k) {
    int const e0 = tid + k * UT;
    int const e = e0 < NPI ? e0 : NPI - 1;
    int const half = e & 1, r = e >> 1;
    int const pw = r % PW, ph = r / PW;
    int const hh = h0 + ph - 1, ww = w0 + pw - 1;
    iok[k] = hh >= 0 && hh < H && ww >= 0 && ww < W;
    ioff[k] = iok[k] ? (uint32_t)(hh * W + ww) * 16u + (uint32_t)(half * 8) : (uint32_t)(half * 8);
  }
  // the loads are raw (clamped addresses); the halo and out-of-range planes
  // are zeroed when the plane is stored, so that a prefetch stays in flight
  // until then
  auto load_plane = [&](int dd, uint4 (&v)[NPK]) {
    bool const okd = dd >= 0 && dd < D;
    uint16_t const *const src = xn + (okd ? dd : 0) * plane;
#pragma unroll
    for (int k = 0; k < NPK; ++k) v[k] = *(uint4 const *)(src + ioff[k]);
  };
  auto store_plane = [&](int dd, uint4 const (&v)[NPK]) {
    bool const okd = dd >= 0 && dd < D;
    uint16_t *const dstp = sx + ((dd + 3) % 3) * PLANE;
#pragma unroll
    for (int k = 0; k < NPK; ++k) {
      int const e0 = tid + k * UT;
      int const e = e0 < NPI ? e0 : NPI - 1;
      uint32_t const msk = (okd && iok[k]) ? ~0u : 0u;
      st_h(dstp, e >> 1, e & 1, make_uint4(v[k].x & msk, v[k].y & msk, v[k].z & msk, v[k].w & msk));
    }
  };
  {
    uint4 v0[NPK], v1[NPK], v2[NPK];
    load_plane(d0 - 1, v0);
    load_plane(d0, v1);
    load_plane(d0 + 1, v2);
    store_plane(d0 - 1, v0);
    store_plane(d0, v1);
    store_plane(d0 + 1, v2);
  }
  int const row = tid >> 6, col = tid & 63;
  float const b0 = bias[0];
  constexpr int TPW = (NVT + UT / 64 - 1) / (UT / 64);  // tiles per wave
  for (int j = 0; j < HDB; ++j) {
    int const d = d0 + j;
    int const h = h0 + row, w = w0 + col;
    bool const live = h < H && w < W && d < D;
    int64_t const v = (((int64_t)n * D + (live ? d : 0)) * H + (live ? h : 0)) * W + (live ? w : 0);
    // the field value is loaded before the prefetch, so that waiting for
    // it does not wait for the prefetch
    float const bv = base ? base[v] : 0.f;
    uint4 nxt[NPK];
    if (j + 1 < HDB) load_plane(d + 2, nxt);
    __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the compute
    __syncthreads();  // the ring holds planes d - 1, d, d + 1; F is free
    {
      // lanes kq < 2 read plane d - 1 (first MFMA) and d + 1 (second), lanes
      // kq >= 2 plane d (first) and d + 1 again (second: zero weights)
      uint16_t const *const p1 = sx + ((kq < 2 ? d + 2 : d) % 3) * PLANE;
      uint16_t const *const p2 = sx + ((d + 1) % 3) * PLANE;
      u32x4 a1[TPW], a2[TPW];
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        int const u = (wave + q * (UT / 64)) * 16 + i;
        a1[q] = ld_h(p1, u < NV ? u : NV - 1, kq & 1);
        a2[q] = ld_h(p2, u < NV ? u : NV - 1, kq & 1);
      }
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        int const t = wave + q * (UT / 64);
        if (t < NVT) {
          f32x4 e = {0.f, 0.f, 0.f, 0.f};
          e = mfma32(a1[q], bw0, e);
          e = mfma32(a2[q], bw1, e);
          // D: lane holds voxels 16 t + 4 kq + r of tap i
          if (i < 9) *(f32x4 *)(se + i * NVP + t * 16 + 4 * kq) = e;
        }
      }
    }
    __syncthreads();  // F is complete; every wave is done with the ring
    if (j + 1 < HDB) store_plane(d + 2, nxt);  // into the slot of plane d - 1
    float acc = 0.f;
#pragma unroll
    for (int kh = 0; kh < 3; ++kh)
#pragma unroll
      for (int kw = 0; kw < 3; ++kw)
        acc += se[(kh * 3 + kw) * NVP + (row + kh) * PW + col + kw];
    if (live) {
      float const o = (acc + b0) * unscale;
      dst[v] = base ? bv + o : o;
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
// Weight and bias gradients: dW[tap][co][ci] = sum_v dy[v][co] x[v + tap][ci],
// db[co] = sum_v dy[v][co]
// ---------------------------------------------------------------------------
// A GEMM over voxels (K) on the matrix cores: for one tap, A[co][v] = dy and
// B[v][ci] = the shifted input. A block owns one 64-voxel w chunk of RB
// consecutive (n, d, h) rows and a group of taps (NKH = 3: the nine taps of
// one kd, grid.y = kd; NKH = 1: the three taps of one (kd, kh)). Per row it
// stages dy transposed (dys[co][v]) and the input rows transposed once per
// kw shift (xs[r][kw][ci][v]), so both operands are aligned 16-byte LDS
// reads (K = 32 voxels per matrix-core step); waves split the (tap, co
// tile, ci tile) products and the two 32-voxel chunks of the row. The block's partial sums (fixed-order
// reduction over its waves) go to part[block][27][CO][CI] (taps of its
// group only) and, for grid.y == 0, the bias partials to bpart[block][CO].
template <int CI, int CO, int NKH>
struct WgradShape {
  static constexpr int NTAP = NKH * 3;
  static constexpr int NTILE = NTAP * (CO / 16) * (CI / 16);
  static constexpr int TG = NTILE <= 64 ? 2 : 4;  // waves over tiles
  static constexpr int NCS = 4 / TG;               // chunk groups (of the two 32-voxel chunks)
  static constexpr int CPW = 2 / NCS;              // chunks per wave
  static constexpr int NPW = (NTILE + TG - 1) / TG;  // tiles per wave
  // LDS row: 64 voxels + 8: 16-byte aligned reads, and the eight rows of a
  // read cycle in distinct bank groups
  static constexpr int LR = 72;
  static constexpr int XS = NKH * 3 * CI * LR;                          // bf16 elements
  static constexpr int DS = CO * LR;
  static constexpr int RED = (NCS > 1 ? TG * NPW * 256 : 0) > 256       // floats
                                 ? (NCS > 1 ? TG * NPW * 256 : 0)
                                 : 256;
  static constexpr int LDS_BYTES =
      (2 * (XS + DS) > 4 * RED ? 2 * (XS + DS) : 4 * RED);
  static constexpr int NPART = 256 / CO;  // bias: voxel parts per channel
  static constexpr int VPP = 64 / NPART;
};

template <int CI, int CO, int NKH>
__global__ __launch_bounds__(UT) void k_conv3_wgrad(
    uint16_t const *__restrict__ x, uint16_t const *__restrict__ dy, float *__restrict__ part,
    float *__restrict__ bpart, int D, int H, int W, int64_t rows, int RB) {
  using S = WgradShape<CI, CO, NKH>;
  static_assert(CI % 16 == 0 && (CO == 16 || CO == 32 || CO == 64), "channel counts");
  constexpr int LR = S::LR;
  __shared__ __align__(16) unsigned char lds[S::LDS_BYTES];
  uint16_t *const xs = (uint16_t *)lds;         // [NKH][3][CI][LR]
  uint16_t *const dys = xs + S::XS;             // [CO][LR]
  int const tid = threadIdx.x;
  int const lane = tid & 63, wave = tid >> 6;
  int const i = lane & 15, kq = lane >> 4;
  int const tg = wave % S::TG, cs = wave / S::TG;
  int const nwc = (W + 63) / 64;
  int const wc = blockIdx.x % nwc;
  int64_t const rb = blockIdx.x / nwc;
  int const w0 = wc * 64;
  int const g = blockIdx.y;
  int const kd = NKH == 3 ? g : g / 3;
  int const kh0 = NKH == 3 ? 0 : g % 3;
  bool const do_bias = bpart != nullptr && g == 0;
  f32x4 acc[S::NPW];
#pragma unroll
  for (int t = 0; t < S::NPW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  int64_t const r0 = rb * RB;
  int64_t const r1 = r0 + RB < rows ? r0 + RB : rows;
  for (int64_t row = r0; row < r1; ++row) {
    int const h = (int)(row % H);
    int64_t const nd = row / H;
    int const d = (int)(nd % D);
    int64_t const n = nd / D;
    __syncthreads();
    // dy row, transposed: item = (voxel pair, 8-channel group)
    for (int e = tid; e < 32 * (CO / 8); e += UT) {
      int const v2 = e & 31, c8 = e >> 5;
      int const wv = w0 + 2 * v2;
      uint4 p0 = make_uint4(0, 0, 0, 0), p1 = p0;
      int64_t const base = (row * W + wv) * CO + c8 * 8;
      if (wv < W) p0 = *(uint4 const *)(dy + base);
      if (wv + 1 < W) p1 = *(uint4 const *)(dy + base + CO);
      uint32_t const a0[4] = {p0.x, p0.y, p0.z, p0.w}, a1[4] = {p1.x, p1.y, p1.z, p1.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        uint32_t const lo = (a0[k] & 0xffffu) | (a1[k] << 16);
        uint32_t const hi = (a0[k] >> 16) | (a1[k] & 0xffff0000u);
        *(uint32_t *)(dys + (c8 * 8 + 2 * k) * LR + 2 * v2) = lo;
        *(uint32_t *)(dys + (c8 * 8 + 2 * k + 1) * LR + 2 * v2) = hi;
      }
    }
    // input rows (dd, hh) for the group's kh, once per kw shift
    int const dd = d + kd - 1;
    for (int e = tid; e < NKH * 3 * 32 * (CI / 8); e += UT) {
      int const v2 = e & 31;
      int const c8 = (e >> 5) % (CI / 8);
      int const rk = (e >> 5) / (CI / 8);  // r * 3 + kw
      int const r = rk / 3, kw = rk - r * 3;
      int const hh = h + kh0 + r - 1;
      int const wv = w0 + 2 * v2 + kw - 1;
      bool const okrow = dd >= 0 && dd < D && hh >= 0 && hh < H;
      uint4 p0 = make_uint4(0, 0, 0, 0), p1 = p0;
      int64_t const base = (((n * D + dd) * H + hh) * (int64_t)W + wv) * CI + c8 * 8;
      if (okrow && wv >= 0 && wv < W) p0 = *(uint4 const *)(x + base);
      if (okrow && wv + 1 >= 0 && wv + 1 < W) p1 = *(uint4 const *)(x + base + CI);
      uint32_t const a0[4] = {p0.x, p0.y, p0.z, p0.w}, a1[4] = {p1.x, p1.y, p1.z, p1.w};
      uint16_t *const dst = xs + (rk * CI + c8 * 8) * LR + 2 * v2;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        *(uint32_t *)(dst + (2 * k) * LR) = (a0[k] & 0xffffu) | (a1[k] << 16);
        *(uint32_t *)(dst + (2 * k + 1) * LR) = (a0[k] >> 16) | (a1[k] & 0xffff0000u);
      }
    }
    __syncthreads();
    if (do_bias) {
      int const co = tid % CO, pt = tid / CO;
#pragma unroll
      for (int v = 0; v < S::VPP; v += 4) {
        bf16x4 const q = *(bf16x4 const *)(dys + co * LR + pt * S::VPP + v);
        bsum += (bf2f((uint16_t)q[0]) + bf2f((uint16_t)q[1])) +
                (bf2f((uint16_t)q[2]) + bf2f((uint16_t)q[3]));
      }
    }
#pragma unroll
    for (int cc = 0; cc < S::CPW; ++cc) {
      int const c = cs + cc * S::NCS;  // 32-voxel chunk
#pragma unroll
      for (int j = 0; j < S::NPW; ++j) {
        int const t = tg + j * S::TG;
        if (S::NTILE % S::TG != 0 && t >= S::NTILE) break;
        int const ct = t % (CO / 16);
        int const it = (t / (CO / 16)) % (CI / 16);
        int const tap = t / ((CO / 16) * (CI / 16));  // r * 3 + kw
        u32x4 const a = *(u32x4 const *)(dys + (ct * 16 + i) * LR + c * 32 + 8 * kq);
        u32x4 const b = *(u32x4 const *)(xs + (tap * CI + it * 16 + i) * LR + c * 32 + 8 * kq);
        acc[j] = mfma32(a, b, acc[j]);
      }
    }
  }
  // bias: fixed-order sum over the voxel parts
  float *const red = (float *)lds;
  if (do_bias) {
    __syncthreads();
    red[tid] = bsum;
    __syncthreads();
    if (tid < CO) {
      float b = 0.f;
      for (int pt = 0; pt < S::NPART; ++pt) b += red[pt * CO + tid];
      bpart[(int64_t)blockIdx.x * CO + tid] = b;
    }
  }
  // fixed-order sum over the chunk groups, then one partial per block
  for (int sgrp = 1; sgrp < S::NCS; ++sgrp) {
    __syncthreads();
    if (cs == sgrp) {
#pragma unroll
      for (int j = 0; j < S::NPW; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[(tg * S::NPW + j) * 256 + r * 64 + lane] = acc[j][r];
    }
    __syncthreads();
    if (cs == 0) {
#pragma unroll
      for (int j = 0; j < S::NPW; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[j][r] += red[(tg * S::NPW + j) * 256 + r * 64 + lane];
    }
  }
  if (cs == 0) {
    float *const out = part + (int64_t)blockIdx.x * 27 * CO * CI;
#pragma unroll
    for (int j = 0; j < S::NPW; ++j) {
      int const t = tg + j * S::TG;
      if (S::NTILE % S::TG != 0 && t >= S::NTILE) break;
      int const ct = t % (CO / 16);
      int const it = (t / (CO / 16)) % (CI / 16);
      int const tl = t / ((CO / 16) * (CI / 16));
      int const kh = kh0 + tl / 3, kw = tl % 3;
      int const tap = (kd * 3 + kh) * 3 + kw;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out[((int64_t)tap * CO + ct * 16 + 4 * kq + r) * CI + it * 16 + i] = acc[j][r];
    }
  }
}

// out[i] = sum_s part[s][i] for s < nslab, in order; grid.y splits the
// slabs into groups of `per` whose sums go to out[group][i] (a second pass
// with grid.y = 1 then sums the groups in order)
__global__ void k_reduce_parts(float const *__restrict__ part, float *__restrict__ out,
                               int nslab, int len, int per) {
  int const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  int const s0 = blockIdx.y * per;
  int const s1 = s0 + per < nslab ? s0 + per : nslab;
  float s = 0.f;
  for (int k = s0; k < s1; ++k) s += part[(int64_t)k * len + i];
  out[(int64_t)blockIdx.y * len + i] = s;
}

// ---------------------------------------------------------------------------
// Pooling, upsampling, concatenation, loss, optimizer
// ---------------------------------------------------------------------------
// Elementwise kernels move 8 channels (16 bytes) per thread; channel counts
// and offsets are multiples of 8 (checked on the host).
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint8_t u8x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ u16x8 ld8(uint16_t const *p) { return *(u16x8 const *)p; }
__device__ __forceinline__ void st8(uint16_t *p, u16x8 v) { *(u16x8 *)p = v; }

// 2x2x2 max pool: y[n][d][h][w][c] over x[n][2d..][2h..][2w..][c]; arg holds
// the winning offset (first maximum in (a, b, e) order)
__global__ void k_maxpool(uint16_t const *__restrict__ x, uint16_t *__restrict__ y,
                          uint8_t *__restrict__ arg, int N, int D, int H, int W, int C) {
  int const Do = D / 2, Ho = H / 2, Wo = W / 2, C8 = C / 8;
  int64_t const total = (int64_t)N * Do * Ho * Wo * C8;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int const c = (int)(idx % C8) * 8;
    int64_t r = idx / C8;
    int const w = (int)(r % Wo);
    r /= Wo;
    int const h = (int)(r % Ho);
    r /= Ho;
    int const d = (int)(r % Do);
    int const n = (int)(r / Do);
    float best[8];
    int barg[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      best[k] = -INFINITY;
      barg[k] = 0;
    }
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      int const a = o >> 2, b = (o >> 1) & 1, e = o & 1;
      u16x8 const v =
          ld8(x + ((((int64_t)n * D + 2 * d + a) * H + 2 * h + b) * W + 2 * w + e) * C + c);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        float const f = bf2f(v[k]);
        if (f > best[k]) {
          best[k] = f;
          barg[k] = o;
        }
      }
    }
    u16x8 out;
    u8x8 ar;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      out[k] = f2bf(best[k]);
      ar[k] = (uint8_t)barg[k];
    }
    st8(y + idx * 8, out);
    *(u8x8 *)(arg + idx * 8) = ar;
  }
}

// one thread per 8 channels of an input voxel
__global__ void k_maxpool_bwd(uint16_t const *__restrict__ dy, uint8_t const *__restrict__ arg,
                              uint16_t *__restrict__ dx, int N, int D, int H, int W, int C) {
  int const C8 = C / 8;
  int64_t const total = (int64_t)N * D * H * W * C8;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int const c = (int)(idx % C8) * 8;
    int64_t r = idx / C8;
    int const w = (int)(r % W);
    r /= W;
    int const h = (int)(r % H);
    r /= H;
    int const d = (int)(r % D);
    int const n = (int)(r / D);
    int const Do = D / 2, Ho = H / 2, Wo = W / 2;
    int64_t const o = ((((int64_t)n * Do + d / 2) * Ho + h / 2) * Wo + w / 2) * C + c;
    int const mine = ((d & 1) << 2) | ((h & 1) << 1) | (w & 1);
    u16x8 const g = ld8(dy + o);
    u8x8 const a = *(u8x8 const *)(arg + o);
    u16x8 out;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = a[k] == mine ? g[k] : (uint16_t)0;
    st8(dx + idx * 8, out);
  }
}

// nearest x2 upsampling of x [n][D][H][W][C] into the channels [c0, c0 + C)
// of y [n][2D][2H][2W][Cy]
__global__ void k_upsample(uint16_t const *__restrict__ x, uint16_t *__restrict__ y, int N,
                           int D, int H, int W, int C, int Cy, int c0) {
  int const C8 = C / 8;
  int64_t const total = (int64_t)N * 8 * D * H * W * C8;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int const c = (int)(idx % C8) * 8;
    int64_t r = idx / C8;
    int const w = (int)(r % (2 * W));
    r /= 2 * W;
    int const h = (int)(r % (2 * H));
    r /= 2 * H;
    int const d = (int)(r % (2 * D));
    int const n = (int)(r / (2 * D));
    st8(y + ((((int64_t)n * 2 * D + d) * 2 * H + h) * 2 * W + w) * Cy + c0 + c,
        ld8(x + ((((int64_t)n * D + d / 2) * H + h / 2) * W + w / 2) * C + c));
  }
}

// gradient of the upsampling: dx[v][c] = sum of the 8 children's
// dy[..][c0 + c], in a fixed order; accumulated in fp32
__global__ void k_upsample_bwd(uint16_t const *__restrict__ dy, uint16_t *__restrict__ dx,
                               int N, int D, int H, int W, int C, int Cy, int c0) {
  int const C8 = C / 8;
  int64_t const total = (int64_t)N * D * H * W * C8;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int const c = (int)(idx % C8) * 8;
    int64_t r = idx / C8;
    int const w = (int)(r % W);
    r /= W;
    int const h = (int)(r % H);
    r /= H;
    int const d = (int)(r % D);
    int const n = (int)(r / D);
    float s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = 0.f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      int const a = o >> 2, b = (o >> 1) & 1, e = o & 1;
      u16x8 const g = ld8(dy + ((((int64_t)n * 2 * D + 2 * d + a) * 2 * H + 2 * h + b) * 2 * W +
                                2 * w + e) * Cy + c0 + c);
#pragma unroll
      for (int k = 0; k < 8; ++k) s[k] += bf2f(g[k]);
    }
    u16x8 out;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = f2bf(s[k]);
    st8(dx + idx * 8, out);
  }
}

// channel slice copy: y[v][cy0 + c] = x[v][cx0 + c], c < C, v < nvox
// (mode 1: added to y, fp32 sum rounded once; mode 2: zero where
// gate[v][cy0 + c] <= 0)
__global__ void k_copy_channels(uint16_t const *__restrict__ x, uint16_t *__restrict__ y,
                                int64_t nvox, int Cx, int cx0, int Cy, int cy0, int C,
                                int mode, uint16_t const *__restrict__ gate) {
  int const C8 = C / 8;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nvox * C8;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t const v = idx / C8;
    int const c = (int)(idx - v * C8) * 8;
    u16x8 val = ld8(x + v * Cx + cx0 + c);
    if (mode == 1) {
      u16x8 const o = ld8(y + v * Cy + cy0 + c);
#pragma unroll
      for (int k = 0; k < 8; ++k) val[k] = f2bf(bf2f(val[k]) + bf2f(o[k]));
    } else if (mode == 2) {
      u16x8 const g = ld8(gate + v * Cy + cy0 + c);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (!(bf2f(g[k]) > 0.f)) val[k] = 0;
    }
    st8(y + v * Cy + cy0 + c, val);
  }
}

// fp32 field -> channel 0 of a bf16 tensor with C channels (others zero);
// one thread per voxel
__global__ void k_pack_input(float const *__restrict__ src, uint16_t *__restrict__ x,
                             int64_t nvox, int C, float scale) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (int64_t)gridDim.x * blockDim.x) {
    u16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    u16x8 first = z;
    first[0] = f2bf(src[v] * scale);
    st8(x + v * C, first);
    for (int c = 8; c < C; c += 8) st8(x + v * C + c, z);
  }
}

// channel 0 of a bf16 tensor -> fp32 field times scale (plus base if given)
__global__ void k_unpack_output(uint16_t const *__restrict__ y, float *__restrict__ dst,
                                float const *__restrict__ base, int64_t nvox, int C,
                                float scale) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (int64_t)gridDim.x * blockDim.x) {
    float const o = bf2f(y[v * C]) * scale;
    dst[v] = base ? base[v] + o : o;
  }
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
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = f2bf(a[i]);
}

__global__ void k_from_bf16(uint16_t const *__restrict__ b, float *__restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = bf2f(b[i]);
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
#define CONV_CASES(X) \
  X(16, 16) X(16, 32) X(32, 32) X(32, 64) X(64, 64) X(96, 32) X(48, 16) X(32, 16) \
  X(64, 32) X(32, 96) X(16, 48) X(64, 96) X(16, 64)

static inline int64_t blocks_for(int64_t n, int t) { return (n + t - 1) / t; }
// grid of the grid-stride elementwise kernels
static inline int64_t blocks_capped(int64_t n, int t) {
  int64_t const b = blocks_for(n, t);
  return b < (1 << 20) ? b : (1 << 20);
}

extern "C" {

// The kernels index within one d plane in 32-bit unsigned arithmetic (voxel
// index h * W + w in int, then times the channel count in uint32), so a
// plane's elements H * W * channels must stay below 2^32.
static int conv_shape_ok(int N, int D, int H, int W, int CI, int CO) {
  if (N < 1 || D < 1 || H < 1 || W < 1 || H > 65535 || (int64_t)N * D > 65535 ||
      (int64_t)H * W * (CI > CO ? CI : CO) >= ((int64_t)1 << 32)) {
    fprintf(stderr, "ERROR: convolution shape %d x %d x %d x %d unsupported\n", N, D, H, W);
    return 0;
  }
  return 1;
}

#define CONV_LAUNCH(a, b, u, p)                                                             \
  {                                                                                        \
    int const db = ConvDepth<b, u, p>::DB;                                                 \
    dim3 const grid((unsigned)((W + CT_W - 1) / CT_W), (unsigned)((H + CT_H - 1) / CT_H),  \
                    (unsigned)(N * ((D + db - 1) / db)));                                  \
    return ulaunch("conv3", k_conv3<a, b, u, p>, grid, dim3(UT), stream,                   \
                   (uint16_t const *)x, (uint16_t const *)xu, (uint16_t const *)wt,         \
                   (uint16_t const *)wcu, bias, (uint16_t const *)gate, (uint16_t *)y,       \
                   (uint16_t *)pool, D, H, W, mode);                                        \
  }

#define CONVS_DB 16
static int conv3s_launch(void const *x, void const *wt, float const *bias, void *y, void *pool,
                         int N, int D, int H, int W, int mode, void *stream) {
  dim3 const grid((unsigned)((W + CT_W - 1) / CT_W), (unsigned)((H + CT_H - 1) / CT_H),
                  (unsigned)(N * ((D + CONVS_DB - 1) / CONVS_DB)));
  if (mode & 8)
    return ulaunch("conv3s", k_conv3s<CONVS_DB, true>, grid, dim3(UT), stream,
                   (uint16_t const *)x, (uint16_t const *)wt, bias, (uint16_t *)y,
                   (uint16_t *)pool, D, H, W, mode);
  return ulaunch("conv3s", k_conv3s<CONVS_DB, false>, grid, dim3(UT), stream,
                 (uint16_t const *)x, (uint16_t const *)wt, bias, (uint16_t *)y,
                 (uint16_t *)pool, D, H, W, mode);
}

int agdu_conv3(void const *x, void const *wt, float const *bias, void const *gate, void *y,
               int N, int D, int H, int W, int CI, int CO, int mode, void *stream) {
  void const *const xu = NULL;
  void const *const wcu = NULL;
  void *const pool = NULL;
  mode &= 7;
  if (!conv_shape_ok(N, D, H, W, CI, CO)) return 1;
  // 16 -> 16 without the gradient's gate: the streaming form
  if (CI == 16 && CO == 16 && !(mode & 4))
    return conv3s_launch(x, wt, bias, y, NULL, N, D, H, W, mode, stream);
#define CC(a, b) \
  if (CI == a && CO == b) CONV_LAUNCH(a, b, 0, false)
  CONV_CASES(CC)
#undef CC
  fprintf(stderr, "ERROR: no convolution kernel for %d -> %d channels\n", CI, CO);
  return 1;
}

#define CAT_CASES(X) X(32, 16, 16) X(64, 32, 32)

/* agdu_conv3 that also writes the 2x2x2 max pooling of its output (the
   values agdu_maxpool gives, without the arg), for the layers whose blocks
   hold whole plane pairs */
int agdu_conv3_pool(void const *x, void const *wt, float const *bias, void *y, void *pool,
                    int N, int D, int H, int W, int CI, int CO, int mode, void *stream) {
  void const *const xu = NULL;
  void const *const wcu = NULL;
  void const *const gate = NULL;
  if (!conv_shape_ok(N, D, H, W, CI, CO)) return 1;
  if ((D | H | W) & 1 || pool == NULL) {
    fprintf(stderr, "ERROR: pooling needs even sizes and an output, got %d x %d x %d\n", D, H,
            W);
    return 1;
  }
  mode = (mode & 3) | 8;
  if (CI == 16 && CO == 16) return conv3s_launch(x, wt, bias, y, pool, N, D, H, W, mode, stream);
#define CP(a, b) \
  if (CI == a && CO == b && ConvDepth<b, 0>::DB % 2 == 0) CONV_LAUNCH(a, b, 0, false)
  CONV_CASES(CP)
#undef CP
  fprintf(stderr, "ERROR: no pooling convolution kernel for %d -> %d channels\n", CI, CO);
  return 1;
}

int agdu_conv3_cat(void const *xu, void const *x, void const *wt, float const *bias,
                   void const *gate, void *y, int N, int D, int H, int W, int CU, int CS,
                   int CO, int mode, void *stream) {
  void const *const wcu = NULL;
  void *const pool = NULL;
  mode &= 7;
  if (!conv_shape_ok(N, D, H, W, CU + CS, CO)) return 1;
  if ((D | H | W) & 1) {
    fprintf(stderr, "ERROR: upsampled concatenation needs even sizes, got %d x %d x %d\n", D,
            H, W);
    return 1;
  }
#define CT(u, c, o) \
  if (CU == u && CS == c && CO == o) CONV_LAUNCH(u + c, o, u, false)
  CAT_CASES(CT)
#undef CT
  fprintf(stderr, "ERROR: no concatenating convolution for %d + %d -> %d channels\n", CU, CS,
          CO);
  return 1;
}

int agdu_up_weights(void const *wt, void *wcu, int CI, int CO, int CU, void *stream) {
  if (CU < 16 || CU % 16 || CU >= CI || CO < 16 || CO % 16) {
    fprintf(stderr, "ERROR: combined taps for %d of %d -> %d channels unsupported\n", CU, CI,
            CO);
    return 1;
  }
  int64_t const n = (int64_t)48 * CO * CU;
  return ulaunch("up_weights", k_up_weights, dim3((unsigned)blocks_capped(n, UT)), dim3(UT),
                 stream, (uint16_t const *)wt, (uint16_t *)wcu, CI, CO, CU);
}

int agdu_conv3_up(void const *xu, void const *x, void const *wt, void const *wcu,
                  float const *bias, void const *gate, void *y, int N, int D, int H, int W,
                  int CU, int CS, int CO, int mode, void *stream) {
  if (!conv_shape_ok(N, D, H, W, CU + CS, CO)) return 1;
  if ((D | H | W) & 1) {
    fprintf(stderr, "ERROR: upsampled concatenation needs even sizes, got %d x %d x %d\n", D,
            H, W);
    return 1;
  }
  if (wcu == NULL) {
    fprintf(stderr, "ERROR: combined taps missing\n");
    return 1;
  }
  void *const pool = NULL;
  mode &= 7;
#define CT(u, c, o) \
  if (CU == u && CS == c && CO == o) CONV_LAUNCH(u + c, o, u, true)
  CAT_CASES(CT)
#undef CT
  fprintf(stderr, "ERROR: no concatenating convolution for %d + %d -> %d channels\n", CU, CS,
          CO);
  return 1;
}

int agdu_conv_in(float const *src, void const *wt, int wci, float const *bias, void *y, int N,
                 int D, int H, int W, float scale, void *stream) {
  if (!conv_shape_ok(N, D, H, W, 16, 16)) return 1;
  if (wci < 1) {
    fprintf(stderr, "ERROR: first-layer weights must have input channels, got %d\n", wci);
    return 1;
  }
  dim3 const grid((unsigned)((W + CT_W - 1) / CT_W), (unsigned)((H + CT_H - 1) / CT_H),
                  (unsigned)(N * ((D + FDB - 1) / FDB)));
  return ulaunch("conv_first", k_conv_first, grid, dim3(UT), stream, src,
                 (uint16_t const *)wt, wci, bias, (uint16_t *)y, D, H, W, scale);
}

int agdu_conv_out(void const *x, void const *wt, int wco, float const *bias, float const *base,
                  float *dst, int N, int D, int H, int W, float unscale, void *stream) {
  if (!conv_shape_ok(N, D, H, W, 16, 16)) return 1;
  if (wco < 1) {
    fprintf(stderr, "ERROR: head with %d output channels unsupported\n", wco);
    return 1;
  }
  dim3 const grid((unsigned)((W + CT_W - 1) / CT_W), (unsigned)((H + CT_H - 1) / CT_H),
                  (unsigned)(N * ((D + HDB - 1) / HDB)));
  return ulaunch("conv_head", k_conv_head, grid, dim3(UT), stream, (uint16_t const *)x,
                 (uint16_t const *)wt, wco, bias, base, dst, D, H, W, unscale);
}

int agdu_flip_weights(void const *wt, void *wtt, int CI, int CO, void *stream) {
  int const total = 27 * CI * CO;
  return ulaunch("flip_weights", k_flip_weights, dim3((unsigned)blocks_for(total, UT)),
                 dim3(UT), stream, (uint16_t const *)wt, (uint16_t *)wtt, CI, CO);
}

#define WGRAD_CASES(X) \
  X(16, 16, 3) X(16, 32, 3) X(32, 32, 3) X(48, 16, 3) X(32, 16, 3) X(32, 64, 1) X(64, 64, 1) \
  X(96, 32, 1) X(64, 32, 1) X(16, 64, 3)

// launch shape of the weight gradient: tap groups (grid.y), rows per block
// and blocks (grid.x); 0 when the channel pair has no kernel
static int wgrad_plan(int N, int D, int H, int W, int CI, int CO, int *groups, int *rb,
                      int64_t *nblk) {
  int nkh = 0;
#define WC(a, b, k) \
  if (CI == a && CO == b) nkh = k;
  WGRAD_CASES(WC)
#undef WC
  if (nkh == 0) return 0;
  *groups = nkh == 3 ? 3 : 9;
  int64_t const rows = (int64_t)N * D * H;
  int64_t const nwc = (W + 63) / 64;
  int64_t const target = 2048 / *groups;
  int64_t r = (rows * nwc + target - 1) / target;
  if (r < 1) r = 1;
  *rb = (int)r;
  *nblk = nwc * ((rows + r - 1) / r);
  return 1;
}

#define RED_PER 32

int64_t agdu_wgrad_part_len(int N, int D, int H, int W, int CI, int CO) {
  int groups, rb;
  int64_t nblk;
  if (!wgrad_plan(N, D, H, W, CI, CO, &groups, &rb, &nblk)) return -1;
  int64_t const ng = (nblk + RED_PER - 1) / RED_PER;
  return nblk * (27 * CI * CO + CO) + ng * (27 * CI * CO + CO);
}

int agdu_conv3_wgrad(void const *x, void const *dy, float *part, float *dw, float *db, int N,
                     int D, int H, int W, int CI, int CO, void *stream) {
  int groups, rb;
  int64_t nblk;
  if (N < 1 || D < 1 || H < 1 || W < 1 ||
      !wgrad_plan(N, D, H, W, CI, CO, &groups, &rb, &nblk) || nblk > 0x7fffffff) {
    fprintf(stderr, "ERROR: weight gradient for %d -> %d channels, %d x %d x %d x %d "
                    "unsupported\n", CI, CO, N, D, H, W);
    return 1;
  }
  int const len = 27 * CI * CO;
  float *const bpart = part + nblk * len;
  float *const gw = bpart + nblk * CO;
  int64_t const ng = (nblk + RED_PER - 1) / RED_PER;
  float *const gb = gw + ng * len;
  int64_t const rows = (int64_t)N * D * H;
  dim3 const grid((unsigned)nblk, (unsigned)groups);
  int rc = 1;
#define WC(a, b, k)                                                                        \
  if (CI == a && CO == b)                                                                  \
    rc = ulaunch("conv3_wgrad", k_conv3_wgrad<a, b, k>, grid, dim3(UT), stream,            \
                 (uint16_t const *)x, (uint16_t const *)dy, part, db ? bpart : (float *)0, \
                 D, H, W, rows, rb);
  WGRAD_CASES(WC)
#undef WC
  if (rc) return 1;
  // two-level fixed-order sums over the blocks
  if (ulaunch("reduce_parts", k_reduce_parts, dim3((unsigned)blocks_for(len, UT), (unsigned)ng),
              dim3(UT), stream, (float const *)part, gw, (int)nblk, len, RED_PER) ||
      ulaunch("reduce_parts", k_reduce_parts, dim3((unsigned)blocks_for(len, UT), 1), dim3(UT),
              stream, (float const *)gw, dw, (int)ng, len, (int)ng))
    return 1;
  if (db == NULL) return 0;
  if (ulaunch("reduce_parts", k_reduce_parts, dim3((unsigned)blocks_for(CO, UT), (unsigned)ng),
              dim3(UT), stream, (float const *)bpart, gb, (int)nblk, CO, RED_PER) ||
      ulaunch("reduce_parts", k_reduce_parts, dim3((unsigned)blocks_for(CO, UT), 1), dim3(UT),
              stream, (float const *)gb, db, (int)ng, CO, (int)ng))
    return 1;
  return 0;
}

int agdu_maxpool(void const *x, void *y, void *arg, int N, int D, int H, int W, int C,
                 void *stream) {
  if ((D | H | W) & 1) {
    fprintf(stderr, "ERROR: max pooling needs even sizes, got %d x %d x %d\n", D, H, W);
    return 1;
  }
  if (C % 8) {
    fprintf(stderr, "ERROR: channel count %d is not a multiple of 8\n", C);
    return 1;
  }
  int64_t const total = (int64_t)N * (D / 2) * (H / 2) * (W / 2) * (C / 8);
  return ulaunch("maxpool", k_maxpool, dim3((unsigned)blocks_capped(total, UT)), dim3(UT),
                 stream, (uint16_t const *)x, (uint16_t *)y, (uint8_t *)arg, N, D, H, W, C);
}

int agdu_maxpool_bwd(void const *dy, void const *arg, void *dx, int N, int D, int H, int W,
                     int C, void *stream) {
  if ((D | H | W) & 1) {
    fprintf(stderr, "ERROR: max pooling needs even sizes, got %d x %d x %d\n", D, H, W);
    return 1;
  }
  if (C % 8) {
    fprintf(stderr, "ERROR: channel count %d is not a multiple of 8\n", C);
    return 1;
  }
  int64_t const total = (int64_t)N * D * H * W * (C / 8);
  return ulaunch("maxpool_bwd", k_maxpool_bwd, dim3((unsigned)blocks_capped(total, UT)),
                 dim3(UT), stream, (uint16_t const *)dy, (uint8_t const *)arg,
                 (uint16_t *)dx, N, D, H, W, C);
}

int agdu_upsample(void const *x, void *y, int N, int D, int H, int W, int C, int Cy, int c0,
                  void *stream) {
  if ((C | Cy | c0) % 8 || c0 + C > Cy) {
    fprintf(stderr, "ERROR: upsampling channels %d at %d of %d unsupported\n", C, c0, Cy);
    return 1;
  }
  int64_t const total = (int64_t)N * 8 * D * H * W * (C / 8);
  return ulaunch("upsample", k_upsample, dim3((unsigned)blocks_capped(total, UT)), dim3(UT),
                 stream, (uint16_t const *)x, (uint16_t *)y, N, D, H, W, C, Cy, c0);
}

int agdu_upsample_bwd(void const *dy, void *dx, int N, int D, int H, int W, int C, int Cy,
                      int c0, void *stream) {
  if ((C | Cy | c0) % 8 || c0 + C > Cy) {
    fprintf(stderr, "ERROR: upsampling channels %d at %d of %d unsupported\n", C, c0, Cy);
    return 1;
  }
  int64_t const total = (int64_t)N * D * H * W * (C / 8);
  return ulaunch("upsample_bwd", k_upsample_bwd, dim3((unsigned)blocks_capped(total, UT)),
                 dim3(UT), stream, (uint16_t const *)dy, (uint16_t *)dx, N, D, H, W, C, Cy,
                 c0);
}

int agdu_copy_channels(void const *x, void *y, int64_t nvox, int Cx, int cx0, int Cy,
                       int cy0, int C, int mode, void const *gate, void *stream) {
  if ((Cx | cx0 | Cy | cy0 | C) % 8 || cx0 + C > Cx || cy0 + C > Cy) {
    fprintf(stderr, "ERROR: channel copy %d at %d of %d -> %d of %d unsupported\n", C, cx0, Cx,
            cy0, Cy);
    return 1;
  }
  return ulaunch("copy_channels", k_copy_channels,
                 dim3((unsigned)blocks_capped(nvox * (C / 8), UT)), dim3(UT), stream,
                 (uint16_t const *)x, (uint16_t *)y, nvox, Cx, cx0, Cy, cy0, C, mode,
                 (uint16_t const *)gate);
}

int agdu_pack_input(float const *src, void *x, int64_t nvox, int C, float scale,
                    void *stream) {
  if (C % 8) {
    fprintf(stderr, "ERROR: channel count %d is not a multiple of 8\n", C);
    return 1;
  }
  return ulaunch("pack_input", k_pack_input, dim3((unsigned)blocks_capped(nvox, UT)),
                 dim3(UT), stream, src, (uint16_t *)x, nvox, C, scale);
}

int agdu_unpack_output(void const *y, float *dst, float const *base, int64_t nvox, int C,
                       float scale, void *stream) {
  return ulaunch("unpack_output", k_unpack_output, dim3((unsigned)blocks_capped(nvox, UT)),
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
  return ulaunch("to_bf16", k_to_bf16, dim3((unsigned)blocks_capped(n, UT)), dim3(UT), stream,
                 a, (uint16_t *)b, n);
}

int agdu_from_bf16(void const *b, float *a, int64_t n, void *stream) {
  return ulaunch("from_bf16", k_from_bf16, dim3((unsigned)blocks_capped(n, UT)), dim3(UT),
                 stream, (uint16_t const *)b, a, n);
}

}  // extern "C"
