// HIP kernels of the MI355X deblending path (gfx950, wave64) and the thin
// C-ABI the C host calls (declared in agd_gpu.h).
//
// Operators (reference equivalents):
//   k_fwd_spatial  spatial inverse DFTs of fk_forward's C2R (fk.c:325-327),
//                  cropped to the window's traces
//   k_fwd_time     time C2R + taper + overlap-add of fk_forward (fk.c:329-372)
//   k_blend        deterministic gather form of blend_sum_forward +
//                  apply_mute + set_residual (blend.c:936-960, 862-875,
//                  model.c:1314-1341), and of the sum / mean / overwrite
//                  blends of agd_blend (blend.c:1007-1027, 1155-1260)
//   k_adj_time     blend_adjoint + time split/taper + R2C + x2 of fk_adjoint
//                  (blend.c:1262-1292, fk.c:388-451)
//   k_adj_spatial  spatial forward DFTs of fk_adjoint's R2C, fused with the
//                  model update and soft threshold (model.c:1149-1154,
//                  1392-1423) and the initial threshold max (model.c:1202)
//   k_wavelet      frequency-domain source-wavelet convolution
//                  (wavelet.c:239-301), as a circular convolution evaluated
//                  with a 7-smooth FFT length
//   k_output       set_output's overlap-add into the patches (model.c:1425)
//
// All FFTs are mixed-radix Stockham passes in LDS (radix 4, 2, 3 and a
// direct butterfly for any other prime), one wave64 group per block of
// lines. Sums that the reference performs serially are done here in the
// same order (gathers over precomputed, ordered entry lists), so no float
// atomics are used and results are deterministic.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "agd_fft.h"
#include "agd_gpu.h"

#define NTHR 256

static hipStream_t g_stream = nullptr;
static int g_timing = 0;

struct KTiming {
  double ms;
  long count;
};
static std::map<std::string, KTiming> g_times;
struct PendingEv {
  std::string name;
  hipEvent_t a, b;
};
static std::vector<PendingEv> g_pending;

static int hip_check(hipError_t e, char const *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "ERROR: HIP %s: %s\n", what, hipGetErrorString(e));
    return 1;
  }
  return 0;
}

static void flush_timing() {
  for (auto &p : g_pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(p.b);
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    auto &t = g_times[p.name];
    t.ms += ms;
    t.count += 1;
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  g_pending.clear();
}

// Launch helper: optional per-kernel event timing on the library stream.
template <typename K, typename... Args>
static int launch(char const *name, K kernel, dim3 grid, dim3 block, size_t lds,
                  Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return 0;
  if (g_timing) {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, g_stream);
  }
  hipLaunchKernelGGL(kernel, grid, block, lds, g_stream, args...);
  if (hip_check(hipGetLastError(), name)) return 1;
  if (g_timing) {
    (void)hipEventRecord(b, g_stream);
    g_pending.push_back({name, a, b});
    if (g_pending.size() > 4096) flush_timing();
  }
  return 0;
}

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
// roots[m] = exp(-2 pi i m / n); sign > 0 conjugates
__device__ __forceinline__ float2 rootf(float2 const *roots, int n, int e, int sign) {
  float2 w = roots[e % n];
  if (sign > 0) w.y = -w.y;
  return w;
}

// Blocks go to the eight XCDs in turn (block b runs on XCD b mod 8, each
// XCD with its own L2). This bijection of [0, n) gives XCD x a contiguous
// run of the n work items instead.
__device__ __forceinline__ int xcd_chunk(int b, int n) {
  int const q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

// ---------------------------------------------------------------------------
// Line geometries for LDS-resident batched FFTs
// ---------------------------------------------------------------------------
struct LinesContig {  // line l at l * ld
  int ld;
  __device__ __forceinline__ int base(int l) const { return l * ld; }
};

// lines along dim d of nplanes arrays [ext_0..ext_{nsp-1}] laid out with
// strides str[] (elements), plane stride pstr; dims other than d are
// enumerated over their active extents, last dim fastest
struct LinesND {
  int nsp, d, pstr, count;
  int ext[AGD_MAXSP];
  int str[AGD_MAXSP];
  __device__ __forceinline__ int base(int l) const {
    int const p = l / count;
    int r = l - p * count;
    int off = p * pstr;
    for (int e = nsp - 1; e >= 0; --e) {
      if (e == d) continue;
      int const c = r % ext[e];
      r /= ext[e];
      off += c * str[e];
    }
    return off;
  }
};

// One Stockham pass (radix R, Ns = product of the radices already applied)
// over nlines lines of length n; element j of line l at base(l) + j*stride.
// Radices with a register codelet (2, 3, 4, 5, 7, 8) are compile-time; each
// butterfly loads R - 1 twiddles exp(sign 2 pi i r k / (Ns R)) from the
// roots table of n (exponent r k n / (Ns R) < n, 32-bit).
template <int R, int S, class L>
__device__ void fft_pass_r(float2 const *__restrict__ src, float2 *__restrict__ dst, int n,
                           int Ns, float2 const *__restrict__ roots, int nlines, int stride,
                           L const &lines) {
  int const nr = n / R;
  int const total = nlines * nr;
  int const tw_step = n / (Ns * R);
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int const l = i / nr;
    int const j = i - l * nr;
    int const k = j % Ns;
    int const b = lines.base(l);
    int const ob = b + ((j / Ns) * Ns * R + k) * stride;
    float2 const *in = src + b + j * stride;
    int const istr = nr * stride;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[r * istr];
    if (k) {
      int const e1 = k * tw_step;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = roots[r * e1];
        if (S > 0) w.y = -w.y;
        v[r] = agdfft::c_mul(v[r], w);
      }
    }
    agdfft::butterfly<R, S>(v);
#pragma unroll
    for (int r = 0; r < R; ++r) dst[ob + r * Ns * stride] = v[r];
  }
}

// any other radix: direct DFT of the R inputs with table roots
template <int S, class L>
__device__ void fft_pass_any(float2 const *__restrict__ src, float2 *__restrict__ dst, int n,
                             int R, int Ns, float2 const *__restrict__ roots, int nlines,
                             int stride, L const &lines) {
  int const nr = n / R;
  int const total = nlines * nr;
  int const tw_step = n / (Ns * R);
  int const rstep = n / R;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int const l = i / nr;
    int const j = i - l * nr;
    int const k = j % Ns;
    int const b = lines.base(l);
    int const ob = b + ((j / Ns) * Ns * R + k) * stride;
    float2 const *in = src + b + j * stride;
    int const istr = nr * stride;
    for (int q = 0; q < R; ++q) {
      float2 acc = make_float2(0.f, 0.f);
      for (int r = 0; r < R; ++r) {
        float2 v = in[r * istr];
        int const e = (r * k * tw_step + ((q * r) % R) * rstep) % n;
        if (e) {
          float2 w = roots[e];
          if (S > 0) w.y = -w.y;
          v = agdfft::c_mul(v, w);
        }
        acc = agdfft::c_add(acc, v);
      }
      dst[ob + q * Ns * stride] = acc;
    }
  }
}

template <int S, class L>
__device__ void fft_pass_s(float2 const *__restrict__ src, float2 *__restrict__ dst, int n,
                           int R, int Ns, float2 const *roots, int nlines, int stride,
                           L const &lines) {
  switch (R) {
    case 2: fft_pass_r<2, S>(src, dst, n, Ns, roots, nlines, stride, lines); break;
    case 3: fft_pass_r<3, S>(src, dst, n, Ns, roots, nlines, stride, lines); break;
    case 4: fft_pass_r<4, S>(src, dst, n, Ns, roots, nlines, stride, lines); break;
    case 5: fft_pass_r<5, S>(src, dst, n, Ns, roots, nlines, stride, lines); break;
    case 7: fft_pass_r<7, S>(src, dst, n, Ns, roots, nlines, stride, lines); break;
    case 8: fft_pass_r<8, S>(src, dst, n, Ns, roots, nlines, stride, lines); break;
    default: fft_pass_any<S>(src, dst, n, R, Ns, roots, nlines, stride, lines); break;
  }
}

template <class L>
__device__ void fft_pass(float2 const *__restrict__ src, float2 *__restrict__ dst,
                         int n, int R, int Ns, int sign, float2 const *roots,
                         int nlines, int stride, L const &lines) {
  if (sign < 0)
    fft_pass_s<-1>(src, dst, n, R, Ns, roots, nlines, stride, lines);
  else
    fft_pass_s<+1>(src, dst, n, R, Ns, roots, nlines, stride, lines);
}

// Full FFT along the lines; returns the buffer holding the result. Ends
// with a barrier.
template <class L>
__device__ float2 *fft_lines(float2 *a, float2 *b, AgdFftPlan const &p,
                             float2 const *roots, int sign, int nlines, int stride,
                             L const &lines) {
  int Ns = 1;
  for (int s = 0; s < p.nstage; ++s) {
    fft_pass(a, b, p.n, p.radix[s], Ns, sign, roots, nlines, stride, lines);
    Ns *= p.radix[s];
    __syncthreads();
    float2 *t = a;
    a = b;
    b = t;
  }
  return a;
}

// ---------------------------------------------------------------------------
// Forward spatial transform: X[w][tw][k][S] -> U[w][tw][k][ntr]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTHR) void k_fwd_spatial(
    AgdWin const *__restrict__ wins, AgdTask const *__restrict__ tasks, int task0,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float2 const *__restrict__ X, float2 *__restrict__ U) {
  extern __shared__ float2 sm[];
  AgdTask const t = tasks[task0 + blockIdx.x];
  AgdWin const w = wins[t.win];
  int const S = w.S;
  int const nk = t.nk;
  int const nsp = w.nsp;
  float2 *A = sm;
  float2 *B = sm + nk * S;
  float2 const *src = X + w.x_off + ((int64_t)t.tw * w.Ft + t.k0) * S;
  for (int i = threadIdx.x; i < nk * S; i += blockDim.x) A[i] = src[i];
  __syncthreads();
  LinesND g;
  g.nsp = nsp;
  g.pstr = S;
  int st = 1;
  for (int e = nsp - 1; e >= 0; --e) {
    g.str[e] = st;
    st *= w.P[e];
  }
  for (int e = 0; e < nsp; ++e) g.ext[e] = w.P[e];
  for (int d = 0; d < nsp; ++d) {
    g.d = d;
    int c = 1;
    for (int e = 0; e < nsp; ++e)
      if (e != d) c *= g.ext[e];
    g.count = c;
    AgdFftPlan const &pl = plans[w.splan[d]];
    float2 *res = fft_lines(A, B, pl, roots + pl.roots_off, +1, nk * c, g.str[d], g);
    if (res != A) {
      B = A;
      A = res;
    }
    g.ext[d] = w.W[d];
  }
  int const ntr = w.ntr;
  for (int i = threadIdx.x; i < nk * ntr; i += blockDim.x) {
    int const k = i / ntr;
    int const tr = i - k * ntr;
    int r = tr, off = 0;
    for (int e = nsp - 1; e >= 0; --e) {
      int const c = r % w.W[e];
      r /= w.W[e];
      off += c * g.str[e];
    }
    U[w.u_off + ((int64_t)t.tw * w.Ft + t.k0 + k) * ntr + tr] = A[k * S + off];
  }
}

// ---------------------------------------------------------------------------
// Forward time transform: U -> traces T (C2R, taper, scale, overlap-add)
// LDS: H [tb][Ft] + A,B [tb][M] complex + acc [tb][Wt] float
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTHR) void k_fwd_time(
    AgdWin const *__restrict__ wins, int const *__restrict__ blocks, int block0, int tb,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float const *__restrict__ tapers, float const *__restrict__ tscale,
    float2 const *__restrict__ U, float *__restrict__ T) {
  extern __shared__ float2 sm[];
  int const bid = block0 + blockIdx.x;
  AgdWin const w = wins[blocks[TB_INTS * bid]];
  int const t0 = blocks[TB_INTS * bid + 1];
  int const nt = min(tb, w.ntr - t0);
  int const Pt = w.Pt, Ft = w.Ft, Wt = w.Wt, half = w.half;
  bool const even = (Pt % 2 == 0) && Pt >= 4;
  int const M = even ? Pt / 2 : Pt;
  float2 *H = sm;
  float2 *A = H + tb * Ft;
  float2 *B = A + tb * M;
  float *acc = (float *)(B + tb * M);
  float *nxt = acc + tb * Wt;
  AgdFftPlan const &pl = plans[w.tplan];
  float2 const *rt = roots + pl.roots_off;
  float2 const *rfull = roots + plans[w.tplan_full].roots_off;
  float const *tt = tapers + w.tt_off;
  for (int i = threadIdx.x; i < tb * Wt; i += blockDim.x) acc[i] = 0.f;
  LinesContig lc{M};
  for (int tw = 0; tw < w.ntw; ++tw) {
    float2 const *u = U + w.u_off + (int64_t)tw * Ft * w.ntr + t0;
    for (int i = threadIdx.x; i < Ft * tb; i += blockDim.x) {
      int const k = i / tb;
      int const j = i - k * tb;
      H[j * Ft + k] = j < nt ? u[(int64_t)k * w.ntr + j] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (even) {
      // half-length packing of a C2R; DC and Nyquist imaginary parts ignored
      for (int i = threadIdx.x; i < tb * M; i += blockDim.x) {
        int const j = i / M;
        int const k = i - j * M;
        float2 xk = H[j * Ft + k];
        float2 xc = H[j * Ft + (M - k)];
        if (k == 0) {
          xk.y = 0.f;
          xc.y = 0.f;
        }
        xc.y = -xc.y;
        float2 const e = cadd(xk, xc);
        float2 o = csub(xk, xc);
        o = cmul(o, rootf(rfull, Pt, k, +1));
        A[i] = make_float2(e.x - o.y, e.y + o.x);
      }
    } else {
      for (int i = threadIdx.x; i < tb * M; i += blockDim.x) {
        int const j = i / M;
        int const k = i - j * M;
        float2 v;
        if (k == 0) {
          v = make_float2(H[j * Ft].x, 0.f);
        } else if (k < Ft) {
          v = H[j * Ft + k];
        } else {
          float2 const c = H[j * Ft + (M - k)];
          v = make_float2(c.x, -c.y);
        }
        A[i] = v;
      }
    }
    __syncthreads();
    float2 *R = fft_lines(A, B, pl, rt, +1, tb, 1, lc);
    // accumulate this time window into the overlap buffer
    for (int i = threadIdx.x; i < tb * Wt; i += blockDim.x) {
      int const j = i / Wt;
      int const s = i - j * Wt;
      if (j >= nt) continue;
      float y;
      if (even) {
        float2 const z = R[j * M + (s >> 1)];
        y = (s & 1) ? z.y : z.x;
      } else {
        y = R[j * M + s].x;
      }
      acc[i] += y * tscale[w.trace0 + t0 + j] * tt[s];
    }
    __syncthreads();
    // the first `half` samples are complete: write them, and shift the rest
    // into the other overlap buffer
    for (int i = threadIdx.x; i < tb * Wt; i += blockDim.x) {
      int const j = i / Wt;
      int const s = i - j * Wt;
      if (s < half) {
        int const t = tw * half + s;
        if (j < nt && t < w.nconv) T[w.t_off + (int64_t)(t0 + j) * w.nconv + t] = acc[i];
      }
      nxt[i] = (s + half < Wt) ? acc[i + half] : 0.f;
    }
    __syncthreads();
    {
      float *tmp = acc;
      acc = nxt;
      nxt = tmp;
    }
  }
  // remaining samples after the last window
  for (int i = threadIdx.x; i < tb * (Wt - half); i += blockDim.x) {
    int const j = i / (Wt - half);
    int const s = i - j * (Wt - half);
    int const t = w.ntw * half + s;
    if (j < nt && t < w.nconv) T[w.t_off + (int64_t)(t0 + j) * w.nconv + t] = acc[j * Wt + s];
  }
  // zero the capacity beyond the data portion (wavelet input padding)
  int const tail0 = w.ntw * half + (Wt - half);
  for (int j = 0; j < nt; ++j) {
    for (int t = tail0 + threadIdx.x; t < w.nconv; t += blockDim.x)
      T[w.t_off + (int64_t)(t0 + j) * w.nconv + t] = 0.f;
  }
}

// ---------------------------------------------------------------------------
// Record gather (deterministic blend). One block per tile of the flat record;
// entries of a tile are listed in the reference's serial order.
// modes: 0 sum, 1 mean (taper weights), 2 overwrite, 3 sum - d (residual),
// 4 per-entry constant sum (stepsize fold)
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(NTHR) void k_blend(
    int64_t rec_len, int tile, int const *__restrict__ tile_ptr,
    int const *__restrict__ tile_ent, AgdEntry const *__restrict__ ent,
    float const *__restrict__ src, float const *__restrict__ cst,
    float const *__restrict__ taper, int taper_len,
    unsigned char const *__restrict__ mute, float const *__restrict__ d,
    float *__restrict__ out) {
  // the tile's entries, NTHR at a time, staged in LDS once per block (every
  // thread reads all of them; measured faster than each thread reading the
  // entry table, and than eight entries in flight)
  __shared__ int64_t s_rec[NTHR], s_src[NTHR];
  __shared__ int s_n[NTHR], s_id[NTHR];
  int const e0 = tile_ptr[blockIdx.x];
  int const e1 = tile_ptr[blockIdx.x + 1];
  int64_t const base = (int64_t)blockIdx.x * tile;
  // uniform loops: every thread takes part in every staging pass
  for (int ob = 0; ob < tile; ob += NTHR) {
    int const o = ob + (int)threadIdx.x;
    int64_t const s = base + o;
    bool const live = o < tile && s < rec_len;
    // loaded before the entry loop: a barrier does not wait for them
    bool const muted = live && mute != nullptr && mute[s];
    float const dv = MODE == 3 && live ? d[s] : 0.f;
    float acc = 0.f;
    float cnt = 0.f;
    for (int c0 = e0; c0 < e1; c0 += NTHR) {
      int const nc = min(NTHR, e1 - c0);
      __syncthreads();  // the previous pass's entries are no longer read
      if ((int)threadIdx.x < nc) {
        int const id = tile_ent[c0 + threadIdx.x];
        AgdEntry const e = ent[id];
        s_rec[threadIdx.x] = e.rec;
        s_src[threadIdx.x] = e.src + e.trace_start;
        s_n[threadIdx.x] = e.n;
        s_id[threadIdx.x] = id;
      }
      __syncthreads();
      // BG entries at a time: their loads are issued together (clamped
      // indices, values outside an entry's range unused), then added in the
      // entries' order, so the sums are those of the serial loop
      constexpr int BG = 16;
      for (int q = 0; q < nc; q += BG) {
        float v[BG], wg[BG];
        bool ok[BG];
#pragma unroll
        for (int u = 0; u < BG; ++u) {
          int const qq = min(q + u, nc - 1);
          int64_t const i = s - s_rec[qq];
          int const n = s_n[qq];
          ok[u] = live && (q + u < nc) && i >= 0 && i < n;
          int const ii = ok[u] ? (int)i : 0;
          wg[u] = 1.f;
          if (MODE == 4) {
            v[u] = cst[s_id[qq]];
          } else {
            v[u] = src[s_src[qq] + ii];
            if (MODE == 1) {
              int const to_end = n - 1 - ii;
              int const edge = ii < to_end ? ii : to_end;
              wg[u] = (taper_len > 1 && edge >= 0 && edge < taper_len) ? taper[edge] : 1.f;
            }
          }
        }
#pragma unroll
        for (int u = 0; u < BG; ++u) {
          if (!ok[u]) continue;
          if (MODE == 1) {
            acc += v[u] * wg[u];
            cnt += wg[u];
          } else if (MODE == 2) {
            acc = v[u];
          } else {
            acc += v[u];
          }
        }
      }
    }
    if (live) {
      if (MODE == 1 && cnt > 0.f) acc /= cnt;
      if (MODE == 3) {
        out[s] = muted ? 0.f : acc - dv;
      } else {
        out[s] = muted ? 0.f : acc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint time transform: record (or T2) -> V (taper, split, R2C, x2)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTHR) void k_adj_time(
    AgdWin const *__restrict__ wins, int const *__restrict__ blocks, int block0, int tb,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float const *__restrict__ tapers, float const *__restrict__ tscale,
    AgdEntry const *__restrict__ went, float const *__restrict__ rec,
    float const *__restrict__ Tin, float2 *__restrict__ V, int plain) {
  extern __shared__ float2 sm[];
  int const bid = block0 + blockIdx.x;
  AgdWin const w = wins[blocks[TB_INTS * bid]];
  int const t0 = blocks[TB_INTS * bid + 1];
  int const nt = min(tb, w.ntr - t0);
  int const Pt = w.Pt, Ft = w.Ft, Wt = w.Wt, half = w.half;
  bool const even = (Pt % 2 == 0) && Pt >= 4;
  int const M = even ? Pt / 2 : Pt;
  float2 *A = sm;
  float2 *B = A + tb * M;
  AgdFftPlan const &pl = plans[w.tplan];
  float2 const *rt = roots + pl.roots_off;
  float2 const *rfull = roots + plans[w.tplan_full].roots_off;
  float const *tt = tapers + w.tt_off;
  LinesContig lc{M};
  for (int tw = 0; tw < w.ntw; ++tw) {
    // real input, zero-padded to Pt, packed as complex pairs when even
    int const nreal = tb * Pt;
    for (int i = threadIdx.x; i < nreal; i += blockDim.x) {
      int const j = i / Pt;
      int const s = i - j * Pt;
      float v = 0.f;
      if (j < nt && s < Wt) {
        int const g = w.trace0 + t0 + j;
        int const t = tw * half + s;
        float x;
        if (Tin != nullptr) {
          x = Tin[w.t_off + (int64_t)(t0 + j) * w.nconv + t];
        } else {
          AgdEntry const e = went[g];
          int const i2 = t - e.trace_start;
          x = (i2 >= 0 && i2 < e.n) ? rec[e.rec + i2] : 0.f;
        }
        v = plain ? x * w.scaler : x * tscale[g] * tt[s];
      }
      if (even) {
        float *Af = (float *)A;
        Af[j * Pt + s] = v;  // (z[2m], z[2m+1]) pairs
      } else {
        A[i] = make_float2(v, 0.f);
      }
    }
    __syncthreads();
    float2 *R = fft_lines(A, B, pl, rt, -1, tb, 1, lc);
    float2 *out = (R == A) ? B : A;  // free buffer for the spectrum staging
    // post-process into the half spectrum and store (x2 on interior bins)
    int const kmax2 = (Pt - 1) / 2;
    for (int i = threadIdx.x; i < Ft * tb; i += blockDim.x) {
      int const k = i / tb;
      int const j = i - k * tb;
      if (j >= nt) continue;
      float2 X;
      if (even) {
        float2 const zk = R[j * M + (k % M)];
        float2 zc = R[j * M + ((M - k) % M)];
        zc.y = -zc.y;
        float2 const e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
        float2 const o = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
        X = cadd(e, cmul(o, rootf(rfull, Pt, k, -1)));
        if (k == 0 || k == M) X.y = 0.f;
      } else {
        X = R[j * M + k];
        if (k == 0) X.y = 0.f;
      }
      if (!plain && k >= 1 && k <= kmax2) {
        X.x *= 2.f;
        X.y *= 2.f;
      }
      V[w.u_off + ((int64_t)tw * Ft + k) * w.ntr + t0 + j] = X;
    }
    (void)out;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Adjoint spatial transform fused with the model update and shrink
// mode 0: X -= alpha*G, shrink(lamb); 1: X = -alpha*G, max |X|^2;
// mode 2: X -= alpha*G; 3: G only (X = G, used by operator tests)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTHR) void k_adj_spatial(
    AgdWin const *__restrict__ wins, AgdTask const *__restrict__ tasks, int task0,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float2 const *__restrict__ V, float2 *__restrict__ X, float alpha,
    float lamb, float const *__restrict__ lamb_dev, int mode,
    unsigned int *__restrict__ maxbits) {
  if (lamb_dev != nullptr) lamb = lamb_dev[0];
  extern __shared__ float2 sm[];
  AgdTask const t = tasks[task0 + blockIdx.x];
  AgdWin const w = wins[t.win];
  int const S = w.S;
  int const nk = t.nk;
  int const nsp = w.nsp;
  int const ntr = w.ntr;
  float2 *A = sm;
  float2 *B = sm + nk * S;
  LinesND g;
  g.nsp = nsp;
  g.pstr = S;
  int st = 1;
  for (int e = nsp - 1; e >= 0; --e) {
    g.str[e] = st;
    st *= w.P[e];
  }
  // Both ping-pong buffers start at zero: each dimension's transform only
  // writes the lines it processes, and the positions it skips are the zero
  // padding that the next dimension's transform reads.
  for (int i = threadIdx.x; i < 2 * nk * S; i += blockDim.x) A[i] = make_float2(0.f, 0.f);
  __syncthreads();
  for (int i = threadIdx.x; i < nk * ntr; i += blockDim.x) {
    int const k = i / ntr;
    int const tr = i - k * ntr;
    int r = tr, off = 0;
    for (int e = nsp - 1; e >= 0; --e) {
      int const c = r % w.W[e];
      r /= w.W[e];
      off += c * g.str[e];
    }
    A[k * S + off] = V[w.u_off + ((int64_t)t.tw * w.Ft + t.k0 + k) * ntr + tr];
  }
  __syncthreads();
  for (int e = 0; e < nsp; ++e) g.ext[e] = w.W[e];
  for (int d = nsp - 1; d >= 0; --d) {
    g.d = d;
    int c = 1;
    for (int e = 0; e < nsp; ++e)
      if (e != d) c *= g.ext[e];
    g.count = c;
    AgdFftPlan const &pl = plans[w.splan[d]];
    float2 *res = fft_lines(A, B, pl, roots + pl.roots_off, -1, nk * c, g.str[d], g);
    if (res != A) {
      B = A;
      A = res;
    }
    g.ext[d] = w.P[d];
  }
  float2 *x = X + w.x_off + ((int64_t)t.tw * w.Ft + t.k0) * S;
  float const l2 = lamb * lamb;
  float mx = 0.f;
  for (int i = threadIdx.x; i < nk * S; i += blockDim.x) {
    float2 const gv = A[i];
    float2 v;
    if (mode == 3) {
      v = gv;
    } else if (mode == 1) {
      v = make_float2(0.f - alpha * gv.x, 0.f - alpha * gv.y);
      mx = fmaxf(mx, v.x * v.x + v.y * v.y);
    } else {
      v = x[i];
      v.x -= alpha * gv.x;
      v.y -= alpha * gv.y;
      if (mode == 0) {
        float const a = v.x * v.x + v.y * v.y;
        if (a <= l2) {
          v = make_float2(0.f, 0.f);
        } else {
          float const ab = sqrtf(a);
          float const sh = (ab - lamb) / ab;
          v.x *= sh;
          v.y *= sh;
        }
      }
    }
    x[i] = v;
  }
  if (mode == 1) {
    // block max, then one atomic per block (max is order independent)
    __shared__ float red[NTHR / 64];
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmaxf(m, red[q]);
      atomicMax(maxbits, __float_as_uint(m));
    }
  }
}

__global__ void k_shrink(float2 *__restrict__ X, int64_t n, float lamb) {
  float const l2 = lamb * lamb;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float2 v = X[i];
    float const a = v.x * v.x + v.y * v.y;
    if (a <= l2) {
      v = make_float2(0.f, 0.f);
    } else {
      float const ab = sqrtf(a);
      float const sh = (ab - lamb) / ab;
      v.x *= sh;
      v.y *= sh;
    }
    X[i] = v;
  }
}

__global__ void k_max_bits(float const *__restrict__ v, int64_t n,
                           unsigned int *__restrict__ out) {
  float m = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, v[i]);
  __shared__ float red[NTHR / 64];
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float mm = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) mm = fmaxf(mm, red[q]);
    atomicMax(out, __float_as_uint(mm));
  }
}

// FP64 sum of squares: fixed-size partials (deterministic), summed on host
__global__ void k_norm2(float const *__restrict__ v, int64_t n,
                        double *__restrict__ partials) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double const x = (double)v[i];
    acc += x * x;
  }
  __shared__ double red[NTHR];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Source-wavelet convolution. One block per window trace.
// The reference applies C2R_N(R2C_N(x) * W) with N = nconv (wavelet.c:239);
// for the live traces this is the circular convolution of length N of x with
// h = C2R_N(W). It is evaluated with an FFT length M >= nconv + nfk - 1
// (7-smooth) against the precomputed spectrum Hm of h laid out for lags in
// (-nfk, N): forward y[t] = (IFFT_M(FFT_M(x) Hm))[t] for t < N (x supported
// on [0, nfk)), adjoint x[t] = (IFFT_M(FFT_M(y) conj(Hm)))[t] for t < nfk.
// Hm includes the 1/M normalisation. Non-live traces are transformed but not
// multiplied in the reference, which scales them by N (wavelet.c:273).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NTHR) void k_wavelet(
    AgdWin const *__restrict__ wins, int const *__restrict__ trace_win, int trace_first,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float2 const *__restrict__ wspec, int const *__restrict__ wsrc_off,
    int adjoint, AgdEntry const *__restrict__ went, float const *__restrict__ rec,
    float *__restrict__ T, float *__restrict__ T2) {
  extern __shared__ float2 sm[];
  int const g = trace_first + blockIdx.x;
  AgdWin const w = wins[trace_win[g]];
  int const tr = g - w.trace0;
  int const N = w.nconv;
  int const nfk = w.nfk;
  int const Mr = w.wM;  // real length
  int const M = Mr / 2;
  float *Tio = T + w.t_off + (int64_t)tr * N;
  float *Tout = T2 + w.t_off + (int64_t)tr * N;
  int const so = wsrc_off[g];
  AgdEntry const e = went[g];
  // input sample accessor
  int const nin = adjoint ? N : nfk;
  if (so < 0) {
    // not multiplied: C2R(R2C(x)) = N x
    float const fN = (float)N;
    if (!adjoint) {
      for (int t = threadIdx.x; t < N; t += blockDim.x) Tio[t] = t < nfk ? Tio[t] * fN : 0.f;
    } else {
      for (int t = threadIdx.x; t < N; t += blockDim.x) {
        int const i2 = t - e.trace_start;
        float const x = (i2 >= 0 && i2 < e.n) ? rec[e.rec + i2] : 0.f;
        Tout[t] = t < nfk ? x * fN : 0.f;
      }
    }
    return;
  }
  float2 *A = sm;
  float2 *B = A + M;
  float2 *Hs = B + M;  // half spectrum M+1
  float *Af = (float *)A;
  for (int t = threadIdx.x; t < Mr; t += blockDim.x) {
    float x = 0.f;
    if (t < nin) {
      if (!adjoint) {
        x = Tio[t];
      } else {
        int const i2 = t - e.trace_start;
        x = (i2 >= 0 && i2 < e.n) ? rec[e.rec + i2] : 0.f;
      }
    }
    Af[t] = x;
  }
  __syncthreads();
  AgdFftPlan const &pl = plans[w.wplan];
  float2 const *rt = roots + pl.roots_off;
  float2 const *rfull = roots + plans[w.wplan_full].roots_off;
  LinesContig lc{M};
  float2 *R = fft_lines(A, B, pl, rt, -1, 1, 1, lc);
  float2 const *hs = wspec + so;
  for (int k = threadIdx.x; k <= M; k += blockDim.x) {
    float2 const zk = R[k % M];
    float2 zc = R[(M - k) % M];
    zc.y = -zc.y;
    float2 const ev = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    float2 const o = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
    float2 X = cadd(ev, cmul(o, rootf(rfull, Mr, k, -1)));
    if (k == 0 || k == M) X.y = 0.f;
    float2 h = hs[k];
    if (adjoint) h.y = -h.y;
    Hs[k] = cmul(X, h);
  }
  __syncthreads();
  float2 *Z = (R == A) ? B : A;
  for (int k = threadIdx.x; k < M; k += blockDim.x) {
    float2 xk = Hs[k];
    float2 xc = Hs[M - k];
    if (k == 0) {
      xk.y = 0.f;
      xc.y = 0.f;
    }
    xc.y = -xc.y;
    float2 const ev = cadd(xk, xc);
    float2 o = csub(xk, xc);
    o = cmul(o, rootf(rfull, Mr, k, +1));
    Z[k] = make_float2(ev.x - o.y, ev.y + o.x);
  }
  __syncthreads();
  float2 *Y = fft_lines(Z, (Z == A) ? B : A, pl, rt, +1, 1, 1, lc);
  float const *Yf = (float const *)Y;
  if (!adjoint) {
    for (int t = threadIdx.x; t < N; t += blockDim.x) Tio[t] = Yf[t];
  } else {
    for (int t = threadIdx.x; t < N; t += blockDim.x) Tout[t] = t < nfk ? Yf[t] : 0.f;
  }
}

// ---------------------------------------------------------------------------
// Wavelet convolution with a compile-time transform length: the real length
// Mr = 2 MH of the window's convolution, MH the product of the radices Rs
// (the host picks Mr among the lengths built here when one is close to the
// smallest admissible length). One block of NT threads per window trace; the
// half-length complex FFT (z[n] = x[2n] + i x[2n+1]) runs as the radix
// passes of a Stockham transform through LDS tiles of MH values: in place
// (PP false: each pass reads every butterfly's inputs into registers, waits
// for the block, then writes) or between two tiles (PP true: one barrier
// per pass). One float2 of padding per 32 spreads the strided writes of the
// early passes over the banks; every pass length is a multiple of 32, so a
// butterfly's padded addresses are one base plus constant offsets.
// The forward's first pass reads its inputs straight from memory and the
// inverse's last pass writes straight to memory; between the two
// transforms the real-spectrum unpack, the product with the wavelet
// spectrum and the pack for the inverse run in place on the pairs
// (k, MH - k). Pruned passes: when the input has samples only in the first
// NB1 complex values (the forward reads n_times_fk samples), the first
// pass's DFT is its first input repeated; when only the first MH / RL
// complex outputs are needed (the adjoint keeps n_times_fk samples), the
// last pass computes only its output r = 0. The same arithmetic as
// k_wavelet otherwise. The adjoint writes only t < nfk: the rest of T2 is
// zeroed when it is allocated and never written.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wpad(int i) { return i + (i >> 5); }

template <int I, int... Rs>
constexpr int wav_radix() {
  constexpr int r[] = {Rs...};
  return r[I];
}
// A thread's twiddles of the passes 1 .. NP - 1 (pass 0 has none), kept in
// registers for the whole kernel: the inverse uses their conjugates (its
// pass p has the same Ns and butterflies as the forward's). wav_tw_off(p):
// the first of pass p in that array.
template <int NT, int... Rs>
constexpr int wav_tw_off(int p) {
  constexpr int r[] = {Rs...};
  constexpr int L = (Rs * ...);
  int o = 0;
  for (int i = 1; i < p; ++i) o += ((L / r[i] + NT - 1) / NT) * (r[i] - 1);
  return o;
}
template <int NT, int P, int Ns, int... Rs>
__device__ __forceinline__ void wav_tw_load(float2 *tw, float2 const *__restrict__ rt, int tid) {
  if constexpr (P < (int)sizeof...(Rs)) {
    constexpr int R = wav_radix<P, Rs...>(), L = (Rs * ...), NB = L / R;
    constexpr int Q = (NB + NT - 1) / NT, TS = L / (Ns * R), O = wav_tw_off<NT, Rs...>(P);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      int const j = min(tid + q * NT, NB - 1);
      int const k = j % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) tw[O + q * (R - 1) + r - 1] = rt[r * k * TS];  // < L
    }
    wav_tw_load<NT, P + 1, Ns * R, Rs...>(tw, rt, tid);
  }
}

// One Stockham pass (decimation in time, autosorting) of radix R over the
// L values in src, Ns the product of the radices before it: butterfly j
// takes src[j + r L/R], twiddles by exp(S 2 pi i r k / (Ns R)) (tw, for
// Ns > 1), k = j mod Ns, and writes dst[(j - k) R + k + r Ns] (dst == src
// unless PP).
template <int R, int Ns, int L, int S, int NT, bool PP>
__device__ __forceinline__ void wav_pass(float2 const *src, float2 *dst, float2 const *tw,
                                         int tid) {
  constexpr int NB = L / R, Q = (NB + NT - 1) / NT;
  static_assert(NB % 32 == 0, "padded offsets need pass lengths that are multiples of 32");
  float2 v[Q][R];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    int const j = min(tid + q * NT, NB - 1);
    int const b = wpad(j);
#pragma unroll
    for (int r = 0; r < R; ++r) v[q][r] = src[b + r * (NB + NB / 32)];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 t = tw[q * (R - 1) + r - 1];
        if (S > 0) t.y = -t.y;
        v[q][r] = cmul(v[q][r], t);
      }
    }
    agdfft::fft<R, S>(v[q]);
  }
  if (!PP) __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    int const j = tid + q * NT;
    if (j < NB) {
      int const k = j % Ns;
      int const w0 = (j - k) * R + k;
      if (Ns % 32 == 0) {
        int const b = wpad(w0);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[b + r * (Ns + Ns / 32)] = v[q][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) dst[wpad(w0 + r * Ns)] = v[q][r];
      }
    }
  }
  __syncthreads();
}

// the passes I .. E - 1 of the radix list Rs, from s0 (to s1 and back with PP)
template <int S, int NT, bool PP, int L, int Ns, int I, int E, int... Rs>
__device__ __forceinline__ void wav_passes(float2 *s0, float2 *s1, float2 const *tw, int tid) {
  if constexpr (I < E) {
    constexpr int R = wav_radix<I, Rs...>();
    wav_pass<R, Ns, L, S, NT, PP>(s0, PP ? s1 : s0, tw + (I > 0 ? wav_tw_off<NT, Rs...>(I) : 0),
                                  tid);
    wav_passes<S, NT, PP, L, Ns * R, I + 1, E, Rs...>(PP ? s1 : s0, PP ? s0 : s1, tw, tid);
  }
}

template <int NT, bool PP, int... Rs>
__global__ __launch_bounds__(NT) void k_wavelet_t(
    AgdWin const *__restrict__ wins, int const *__restrict__ trace_win, int trace_first,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float2 const *__restrict__ wspec, int const *__restrict__ wsrc_off, int adjoint,
    AgdEntry const *__restrict__ went, float const *__restrict__ rec, float *__restrict__ T,
    float *__restrict__ T2) {
  constexpr int NP = sizeof...(Rs);
  constexpr int MH = (Rs * ...), MP = MH + MH / 32;
  constexpr int R1 = wav_radix<0, Rs...>(), RL = wav_radix<NP - 1, Rs...>();
  constexpr int NB1 = MH / R1, NBL = MH / RL;  // NBL: the last pass's Ns
  static_assert(NB1 % 32 == 0 && NBL % 32 == 0, "pass lengths must be multiples of 32");
  __shared__ float2 sa[PP ? 2 * MP : MP];
  // with PP the first pass writes tile 0 and each later pass the other
  // tile: the forward transform ends in tile (NP - 1) % 2 (the spectrum
  // step works there in place), the inverse before its last pass in tile 0
  float2 *const sf = sa + (PP && (NP - 1) % 2 ? MP : 0);
  float2 *const sx = PP ? sa + (sf == sa ? MP : 0) : sa;  // the other tile
  float2 *const sg = sa;
  int const tid = threadIdx.x;
  int const g = trace_first + blockIdx.x;
  AgdWin const w = wins[trace_win[g]];
  int const tr = g - w.trace0;
  int const N = w.nconv;
  int const nfk = w.nfk;
  float *Tio = T + w.t_off + (int64_t)tr * N;
  float *Tout = T2 + w.t_off + (int64_t)tr * N;
  int const so = wsrc_off[g];
  AgdEntry const e = went[g];
  int const nin = adjoint ? N : nfk;
  if (so < 0) {
    // not multiplied: C2R(R2C(x)) = N x
    float const fN = (float)N;
    if (!adjoint) {
      for (int t = tid; t < N; t += NT) Tio[t] = t < nfk ? Tio[t] * fN : 0.f;
    } else {
      for (int t = tid; t < nfk; t += NT) {
        int const i2 = t - e.trace_start;
        float const x = (i2 >= 0 && i2 < e.n) ? rec[e.rec + i2] : 0.f;
        Tout[t] = x * fN;
      }
    }
    return;
  }
  float2 const *const rt = roots + plans[w.wplan].roots_off;          // e^{-2 pi i m / MH}
  float2 const *const rfull = roots + plans[w.wplan_full].roots_off;  // e^{-2 pi i m / (2 MH)}
  float2 const *const hs = wspec + so;
  // the twiddles of both transforms are loaded here, before the first
  // barrier (a barrier does not wait for global loads, so their latency
  // hides behind the first pass; loaded pass by pass they would each wait
  // behind a barrier)
  constexpr int NTW = wav_tw_off<NT, Rs...>(NP) > 0 ? wav_tw_off<NT, Rs...>(NP) : 1;
  float2 tw[NTW];
  wav_tw_load<NT, 1, R1, Rs...>(tw, rt, tid);
  // the trace's samples: one source (the trace buffer, or the record through
  // the blend entry), clamped index, value masked outside the range (no
  // branch around a load)
  float const *const src = adjoint ? rec : Tio;
  int64_t const sbase = adjoint ? e.rec : 0;
  int const sstart = adjoint ? e.trace_start : 0;
  int const scount = adjoint ? e.n : nin;
  auto sample = [&](int t) -> float {
    int const i2 = t - sstart;
    bool const ok = t < nin && (unsigned)i2 < (unsigned)scount;
    float const x = src[sbase + (ok ? i2 : 0)];
    return ok ? x : 0.f;
  };
  // forward transform, first pass from memory into sa (Ns = 1: no twiddles)
  {
    constexpr int Q = (NB1 + NT - 1) / NT;
    if (2 * NB1 >= nin) {
      // samples only in z[0, NB1): every output of butterfly j is z[j]
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        int const j = tid + q * NT;
        if (j < NB1) {
          float2 const x = make_float2(sample(2 * j), sample(2 * j + 1));
#pragma unroll
          for (int r = 0; r < R1; ++r) sa[wpad(j * R1 + r)] = x;
        }
      }
    } else {
      float2 v[Q][R1];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        int const j = min(tid + q * NT, NB1 - 1);
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          int const n = j + r * NB1;
          v[q][r] = make_float2(sample(2 * n), sample(2 * n + 1));
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        int const j = tid + q * NT;
        agdfft::fft<R1, -1>(v[q]);
        if (j < NB1) {
#pragma unroll
          for (int r = 0; r < R1; ++r) sa[wpad(j * R1 + r)] = v[q][r];
        }
      }
    }
  }
  __syncthreads();
  wav_passes<-1, NT, PP, MH, R1, 1, NP, Rs...>(sa, sa + (PP ? MP : 0), tw, tid);
  // spectrum: real unpack, product with the wavelet, pack for the inverse,
  // in place on the pairs (k, MH - k) of sf; wk = e^{-2 pi i k / (2 MH)}
  auto unpack = [&](float2 zk, float2 zo, float2 wk, float2 h, int k) {
    float2 const zc = make_float2(zo.x, -zo.y);
    float2 const ev = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
    float2 const o = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
    float2 X = cadd(ev, cmul(o, wk));
    if (k == 0 || k == MH) X.y = 0.f;
    if (adjoint) h.y = -h.y;
    return cmul(X, h);
  };
  auto pack = [&](float2 xk, float2 xo, float2 wk, int k) {
    if (k == 0) {
      xk.y = 0.f;
      xo.y = 0.f;
    }
    float2 const xc = make_float2(xo.x, -xo.y);
    float2 const ev = cadd(xk, xc);
    float2 const o = cmul(csub(xk, xc), make_float2(wk.x, -wk.y));
    return make_float2(ev.x - o.y, ev.y + o.x);
  };
  auto pair = [&](int k, float2 hk, float2 hc, float2 wk, bool write) {
    int const kc = MH - k;
    float2 const wc = make_float2(-wk.x, wk.y);  // e^{-2 pi i kc / (2 MH)}
    float2 const za = sf[wpad(k)];
    float2 const zb = sf[wpad(kc % MH)];
    float2 const xk = unpack(za, zb, wk, hk, k);
    float2 const xc = unpack(zb, za, wc, hc, kc);
    if (write) {
      sf[wpad(k)] = pack(xk, xc, wk, k);
      if (k != 0 && k != kc) sf[wpad(kc)] = pack(xc, xk, wc, kc);
    }
  };
  {
    // unrolled over the compile-time trip count with the index clamped for
    // the reads; a clamped thread writes nothing (its pair belongs to a
    // thread of another wave, which may already have written it)
    constexpr int KQ = (MH / 2 + 1 + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < KQ; ++q) {
      int const k0 = tid + q * NT;
      int const k = min(k0, MH / 2);
      pair(k, hs[k], hs[MH - k], rfull[k], k0 <= MH / 2);
    }
  }
  __syncthreads();
  wav_passes<+1, NT, PP, MH, 1, 0, NP - 1, Rs...>(sf, sx, tw, tid);
  // inverse transform, last pass from sg to memory: butterfly j < NBL (its
  // Ns) writes z[j + r NBL], i.e. the samples 2 (j + r NBL) and
  // 2 (j + r NBL) + 1
  {
    constexpr int Q = (NBL + NT - 1) / NT, TL = wav_tw_off<NT, Rs...>(NP - 1);
    int const nout = adjoint ? nfk : N;
    float *const dst = adjoint ? Tout : Tio;
    if (nout <= 2 * NBL) {
      // only the outputs r = 0: z[j] = sum_m twiddled inputs
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        int const j = tid + q * NT;
        if (2 * j < nout) {
          int const b = wpad(j);
          float2 acc = sg[b];
#pragma unroll
          for (int m = 1; m < RL; ++m) {
            float2 t = tw[TL + q * (RL - 1) + m - 1];  // e^{-2 pi i m j / MH}
            t.y = -t.y;
            acc = cadd(acc, cmul(sg[b + m * (NBL + NBL / 32)], t));
          }
          dst[2 * j] = acc.x;
          if (2 * j + 1 < nout) dst[2 * j + 1] = acc.y;
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        int const j = tid + q * NT;
        if (j < NBL) {
          int const b = wpad(j);
          float2 v[RL];
#pragma unroll
          for (int r = 0; r < RL; ++r) v[r] = sg[b + r * (NBL + NBL / 32)];
#pragma unroll
          for (int r = 1; r < RL; ++r) {
            float2 t = tw[TL + q * (RL - 1) + r - 1];  // e^{-2 pi i r j / MH}
            t.y = -t.y;
            v[r] = cmul(v[r], t);
          }
          agdfft::fft<RL, +1>(v);
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            int const t = 2 * (j + r * NBL);
            if (t < nout) dst[t] = v[r].x;
            if (t + 1 < nout) dst[t + 1] = v[r].y;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Output assembly and blend extraction
// ---------------------------------------------------------------------------
// Threshold of the next iteration of a captured loop: lamb_dev[0] =
// sched[idx], idx += 1 (one lane; vector stores only).
__global__ void k_next_lamb(float const *__restrict__ sched, int *__restrict__ idx,
                            float *__restrict__ lamb_dev) {
  int const lane = threadIdx.x;
  if (lane == 0) {
    int const i = idx[lane];
    lamb_dev[lane] = sched[i];
    idx[lane] = i + 1;
  }
}

__global__ __launch_bounds__(NTHR) void k_output(
    float const *__restrict__ T, int64_t const *__restrict__ o_off,
    int const *__restrict__ o_nt, int const *__restrict__ o_ptr,
    int64_t const *__restrict__ o_src, float *__restrict__ out) {
  int const i = blockIdx.x;
  int const c0 = o_ptr[i], c1 = o_ptr[i + 1];
  int const nt = o_nt[i];
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    float acc = 0.f;
    for (int c = c0; c < c1; ++c) acc += T[o_src[c] + t];
    out[o_off[i] + t] = acc;
  }
}

__global__ __launch_bounds__(NTHR) void k_extract(
    AgdEntry const *__restrict__ ent, int const *__restrict__ trace_len,
    float const *__restrict__ rec, float *__restrict__ out) {
  int const i = blockIdx.x;
  AgdEntry const e = ent[i];
  int const L = trace_len[i];
  for (int t = threadIdx.x; t < L; t += blockDim.x) {
    int const i2 = t - e.trace_start;
    out[e.src + t] = (i2 >= 0 && i2 < e.n) ? rec[e.rec + i2] : 0.f;
  }
}

// ---------------------------------------------------------------------------
// Size-specialised spatial transforms (one or two space dims).
// A task holds nk time-frequency planes of one (window, time window); each
// plane is P0 x P1 (P0 = 1 for one space dim) with W0 x W1 traces. The
// specialised sizes are multiples of 5, for which the window lengths are
// W = 4 P / 5 exactly (P = ceil(1.25 W)), so every loop bound is a
// compile-time constant. Column transforms (length P0) run from global
// memory into registers, row transforms (length P1) through one LDS tile
// holding the W0 trace rows of each plane (padded rows of P1 + 1); every
// FFT is a register codelet.
// ---------------------------------------------------------------------------
constexpr int TG = 16;  // traces per U / V tile of the specialised kernels

template <int P>
__host__ __device__ constexpr int win_len() {
  return P == 1 ? 1 : P * 4 / 5;
}

// row transforms of the W0 rows of each plane, in place: NIN inputs
// (zero padded to P1), NOUT outputs kept (S = +1: inverse of the forward
// operator)
template <int P0, int P1, int S, int NIN, int NOUT>
__device__ __forceinline__ void rows_fft(float2 *sm, int nk) {
  constexpr int LD = P1 + 1;
  constexpr int W0 = win_len<P0>();
  for (int i = threadIdx.x; i < nk * W0; i += blockDim.x) {
    float2 *row = sm + i * LD;
    float2 v[P1];
#pragma unroll
    for (int c = 0; c < P1; ++c) v[c] = (c < NIN) ? row[c] : make_float2(0.f, 0.f);
    agdfft::fft<P1, S>(v);
#pragma unroll
    for (int c = 0; c < NOUT; ++c) row[c] = v[c];
  }
}

// LDS rows of the nk planes <-> the window traces of the planes in U / V.
// Per time window U holds [ntr / ug][Ft][ug] (ug = 16: tiles of 16 traces
// for the specialised time kernels; ug = ntr: [Ft][ntr]); the element
// order below walks each tile's nk planes contiguously. Utw points at the
// task's time window, k0 is its first plane.
template <int P0, int P1, bool TILED, bool TO_LDS>
__device__ __forceinline__ void planes_copy(float2 *sm, float2 *Utw, int k0, int Ft, int nk) {
  constexpr int LD = P1 + 1;
  constexpr int W0 = win_len<P0>();
  constexpr int W1 = win_len<P1>();
  constexpr int ug = TILED ? TG : W0 * W1;
  constexpr int NB = 8;  // loads issued together per thread
  int const span = nk * ug;
  int const total = nk * W0 * W1;
  for (int base = 0; base < total; base += NB * NTHR) {
    float2 v[NB];
    int lp[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int const i = min(base + (int)threadIdx.x + b * NTHR, total - 1);
      int const g = i / span;
      int const rem = i - g * span;
      int const k = rem / ug;
      int const tr = g * ug + (rem - k * ug);
      int const row = tr / W1;
      lp[b] = (k * W0 + row) * LD + (tr - row * W1);
      float2 *u = Utw + ((int64_t)g * Ft + k0) * ug + rem;
      if (TO_LDS)
        v[b] = *u;
      else if (base + (int)threadIdx.x + b * NTHR < total)
        *u = sm[lp[b]];
    }
    if (TO_LDS) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (base + (int)threadIdx.x + b * NTHR < total) sm[lp[b]] = v[b];
    }
  }
}

template <int P0, int P1, bool TILED>
__global__ __launch_bounds__(NTHR) void k_fwd_spatial_t(
    AgdWin const *__restrict__ wins, AgdTask const *__restrict__ tasks, int task0,
    float2 const *__restrict__ X, float2 *__restrict__ U) {
  extern __shared__ float2 sm[];
  constexpr int LD = P1 + 1;
  constexpr int S = P0 * P1;
  constexpr int W0 = win_len<P0>();
  constexpr int W1 = win_len<P1>();
  AgdTask const t = tasks[task0 + blockIdx.x];
  AgdWin const &w = wins[t.win];
  int const nk = t.nk;
  float2 const *src = X + w.x_off + ((int64_t)t.tw * w.Ft + t.k0) * S;
  for (int i = threadIdx.x; i < nk * P1; i += blockDim.x) {
    int const k = i / P1;
    int const c1 = i - k * P1;
    float2 v[P0];
#pragma unroll
    for (int c0 = 0; c0 < P0; ++c0) v[c0] = src[k * S + c0 * P1 + c1];
    agdfft::fft<P0, +1>(v);
#pragma unroll
    for (int n0 = 0; n0 < W0; ++n0) sm[(k * W0 + n0) * LD + c1] = v[n0];
  }
  __syncthreads();
  rows_fft<P0, P1, +1, P1, W1>(sm, nk);
  __syncthreads();
  planes_copy<P0, P1, TILED, false>(sm, U + w.u_off + (int64_t)t.tw * w.Ft * (W0 * W1), t.k0,
                                    w.Ft, nk);
}

// Adjoint spatial transform of V and the model update, by MODE:
//   0  x <- S_lamb(x - alpha g)      (update and soft threshold)
//   1  x <- -alpha g, max |x|^2 into maxbits (first iteration)
//   2  x <- x - alpha g
//   3  x <- g
//   4  max |alpha g|^2 into maxbits, x untouched (the threshold scale of a
//      start whose model is written afterwards)
// With FUSE (modes 0 and 3) the new model's forward spatial transform, the
// next forward's first stage, is written to U.
template <int P0, int P1, bool TILED, int MODE, bool FUSE>
__global__ __launch_bounds__(NTHR) void k_adj_spatial_t(
    AgdWin const *__restrict__ wins, AgdTask const *__restrict__ tasks, int task0,
    float2 const *__restrict__ V, float2 *__restrict__ X, float2 *__restrict__ U,
    float alpha, float lamb, float const *__restrict__ lamb_dev,
    unsigned int *__restrict__ maxbits) {
  extern __shared__ float2 sm[];
  constexpr int LD = P1 + 1;
  constexpr int S = P0 * P1;
  constexpr int W0 = win_len<P0>();
  constexpr int W1 = win_len<P1>();
  if (lamb_dev != nullptr) lamb = lamb_dev[0];
  AgdTask const t = tasks[task0 + blockIdx.x];
  AgdWin const &w = wins[t.win];
  int const nk = t.nk;
  int64_t const uoff = w.u_off + (int64_t)t.tw * w.Ft * (W0 * W1);
  planes_copy<P0, P1, TILED, true>(sm, const_cast<float2 *>(V) + uoff, t.k0, w.Ft, nk);
  __syncthreads();
  rows_fft<P0, P1, -1, W1, P1>(sm, nk);
  __syncthreads();
  float2 *x = X + w.x_off + ((int64_t)t.tw * w.Ft + t.k0) * S;
  float mx = 0.f;
  for (int i = threadIdx.x; i < nk * P1; i += blockDim.x) {
    int const k = i / P1;
    int const c1 = i - k * P1;
    float2 v[P0];
#pragma unroll
    for (int r = 0; r < P0; ++r)
      v[r] = (r < W0) ? sm[(k * W0 + r) * LD + c1] : make_float2(0.f, 0.f);
    agdfft::fft<P0, -1>(v);
    float2 *xc = x + k * S + c1;
#pragma unroll
    for (int c0 = 0; c0 < P0; ++c0) {
      float2 const gv = v[c0];
      float2 nv;
      if constexpr (MODE == 3) {
        nv = gv;
      } else if constexpr (MODE == 1 || MODE == 4) {
        nv = make_float2(0.f - alpha * gv.x, 0.f - alpha * gv.y);
        mx = fmaxf(mx, nv.x * nv.x + nv.y * nv.y);
      } else {
        nv = xc[c0 * P1];
        nv.x -= alpha * gv.x;
        nv.y -= alpha * gv.y;
        if constexpr (MODE == 0) {
          // x (|x| - lamb) / |x| when |x| > lamb, else 0
          float const a = nv.x * nv.x + nv.y * nv.y;
          float const sh = fmaxf(1.f - lamb * __builtin_amdgcn_rsqf(a), 0.f);
          nv.x *= sh;
          nv.y *= sh;
        }
      }
      if constexpr (MODE != 4) xc[c0 * P1] = nv;
      v[c0] = nv;
    }
    if constexpr (FUSE) {
      agdfft::fft<P0, +1>(v);
#pragma unroll
      for (int n0 = 0; n0 < W0; ++n0) sm[(k * W0 + n0) * LD + c1] = v[n0];
    }
  }
  if constexpr (MODE == 1 || MODE == 4) {
    __shared__ float red[NTHR / 64];
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = red[0];
      for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmaxf(m, red[q]);
      atomicMax(maxbits, __float_as_uint(m));
    }
  }
  if constexpr (FUSE) {
    __syncthreads();
    rows_fft<P0, P1, +1, P1, W1>(sm, nk);
    __syncthreads();
    planes_copy<P0, P1, TILED, false>(sm, U + uoff, t.k0, w.Ft, nk);
  }
}

// ---------------------------------------------------------------------------
// Size-specialised time transforms of tiled windows. Pt = 2 M, M = M1 M2,
// and Wt = 8 M / 5 (Pt = ceil(1.25 Wt) with Pt a multiple of 5). One block
// serves a group of TG = 16 consecutive window traces, whose half spectra of
// one time window form one contiguous tile [Ft][TG] of U / V. Per time
// window: a two-phase FFT in one LDS buffer [TG][M + 1] (TG M2 register FFTs
// of length M1, twiddle, TG M1 register FFTs of length M2, in place), with
// the half-length packing of the real transform fused into the first phase
// (forward) or a last pass (adjoint). The overlap with the neighbouring time
// window stays in registers (each thread owns fixed samples of fixed
// traces), and the next time window's input is loaded into registers while
// the current one is transformed. A block serves a run [tw0, tw1) of its
// tile's time windows (all of them unless the host splits them, as it does
// for the window-chunk order); a forward block with tw0 > 0 transforms time
// window tw0 - 1 again for its carry, so every sample is the same sum as in
// one run over all time windows.
// ---------------------------------------------------------------------------
template <int M>
__host__ __device__ constexpr int time_threads() {
  return M >= 160 ? 256 : 128;
}
// dynamic LDS bytes of the two kernels (forward: LT selects the lean-tile
// variant, see k_fwd_time_t)
template <int M, bool LT>
__host__ __device__ constexpr size_t fwd_time_lds() {
  return LT ? (size_t)8 * (TG * (M + 1) + M + (M + 1)) + 4 * TG
            : (size_t)8 * (2 * TG * (M + 1) + M + (M + 1)) + (size_t)4 * TG * (4 * M / 5) +
                  4 * TG;
}
template <int M>
__host__ __device__ constexpr size_t adj_time_lds() {
  return (size_t)8 * (TG * (M + 1) + M + (M + 1)) + TG * 16;
}

// position of complex output n of the in-place two-phase FFT
template <int M1, int M2>
__device__ __forceinline__ int fft2_pos(int n) {
  return (n % M1) * M2 + n / M1;
}

// Two variants, chosen per time class by the autotuner:
//   LT = false: the input tile [Ft][TG] and the transform buffer [TG][LDL]
//     are separate LDS arrays, the overlap carry sits in LDS, and the next
//     time window's tile is loaded into registers while this one is
//     transformed;
//   LT = true ("lean tile"): one LDS tile serves as the input and, once the
//     first phase has read all of it into registers, as the transform
//     buffer; the carry stays in registers and the tile is loaded at the
//     start of its window: about half the LDS and fewer registers, so more
//     blocks share a CU and hide each other's loads.
template <int M1, int M2, bool LT>
__global__ __launch_bounds__(time_threads<M1 * M2>()) void k_fwd_time_t(
    AgdWin const *__restrict__ wins, int const *__restrict__ blocks, int block0,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float const *__restrict__ tapers, float const *__restrict__ tscale,
    float2 const *__restrict__ U, float *__restrict__ T) {
  constexpr int M = M1 * M2;
  constexpr int NT = time_threads<M>();
  constexpr int Ft = M + 1;
  constexpr int LDL = M + 1;
  constexpr int half = 4 * M / 5;
  constexpr int JS = NT / half;          // trace stride of a thread's outputs
  constexpr int NJ = TG / JS;            // traces per thread in the output pass
  constexpr int NCH = Ft * TG / 2;       // 16-byte chunks per tile
  constexpr int R = (NCH + NT - 1) / NT;
  constexpr int NI = TG * M2;            // first-phase items
  constexpr int ITEMS = (NI + NT - 1) / NT;
  static_assert(NT % half == 0 && TG % JS == 0, "output mapping");
  static_assert(!LT || ITEMS <= 2, "first-phase items per thread");
  extern __shared__ float2 sm[];
  float2 *IN = sm;                         // [Ft][TG] tile
  float2 *Z = LT ? sm : IN + Ft * TG;      // [TG][LDL]
  float2 *TW = Z + TG * LDL;               // twiddles [M2][M1]
  float2 *RP = TW + M;                     // exp(+2 pi i k / Pt), k <= M
  float *CA = (float *)(RP + Ft);          // [TG][half] overlap carry (LT: unused)
  float *SC = LT ? CA : CA + TG * half;    // [TG] per-trace scale
  int const bid = block0 + blockIdx.x;
  AgdWin const &w = wins[blocks[TB_INTS * bid]];
  int const t0 = blocks[TB_INTS * bid + 1];
  // the block's time windows [tw0, tw1); one that starts inside the window
  // first transforms time window tw0 - 1 again, for its overlap carry only
  int const tw0 = blocks[TB_INTS * bid + 2], tw1 = blocks[TB_INTS * bid + 3];
  int const twa = tw0 > 0 ? tw0 - 1 : 0;
  int const nconv = w.nconv, ntw = w.ntw;
  int const tid = threadIdx.x;
  {
    float2 const *rM = roots + plans[w.tplan].roots_off;
    float2 const *rP = roots + plans[w.tplan_full].roots_off;
    // TW[k2 M1 + n1] = exp(+2 pi i n1 k2 / M)
    for (int m = tid; m < M; m += NT) {
      float2 const r = rM[((m % M1) * (m / M1)) % M];
      TW[m] = make_float2(r.x, -r.y);
    }
    for (int k = tid; k < Ft; k += NT) RP[k] = make_float2(rP[k].x, -rP[k].y);
    if constexpr (!LT)
      for (int i = tid; i < TG * half; i += NT) CA[i] = 0.f;
    if (tid < TG) SC[tid] = tscale[w.trace0 + t0 + tid];
  }
  int const s = tid % half;
  int const j0 = tid / half;
  float const tap0 = tapers[w.tt_off + s];
  float const tap1 = tapers[w.tt_off + half + s];
  float ca[LT ? NJ : 1];  // LT: the overlap carry in registers
#pragma unroll
  for (int r = 0; r < (LT ? NJ : 1); ++r) ca[r] = 0.f;
  float *Tb = T + w.t_off + (int64_t)t0 * nconv;
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 const *tile = (f4 const *)(U + w.u_off + (int64_t)(t0 / TG) * Ft * TG);
  int64_t const tw_step = (int64_t)(w.ntr / TG) * Ft * TG / 2;  // float4 per time window
  // chunk r of this thread: min(tid + NT r, NCH - 1) (the clamped ones
  // repeat the last chunk and are not stored)
  f4 pre[LT ? 1 : R];
  if constexpr (!LT) {
#pragma unroll
    for (int r = 0; r < R; ++r) pre[r] = tile[(int64_t)twa * tw_step + min(tid + NT * r, NCH - 1)];
  }
  for (int tw = twa; tw < tw1; ++tw) {
    if constexpr (LT) {
      if (tw > twa) __syncthreads();  // the previous window's output pass has read Z
      f4 const *nt = tile + (int64_t)tw * tw_step;
      f4 ld[R];
#pragma unroll
      for (int r = 0; r < R; ++r) ld[r] = nt[min(tid + NT * r, NCH - 1)];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r + 1 < R || tid + NT * r < NCH) ((f4 *)IN)[tid + NT * r] = ld[r];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r + 1 < R || tid + NT * r < NCH) ((f4 *)IN)[tid + NT * r] = pre[r];
      if (tw + 1 < tw1) {
        f4 const *nt = tile + (int64_t)(tw + 1) * tw_step;
#pragma unroll
        for (int r = 0; r < R; ++r) pre[r] = nt[min(tid + NT * r, NCH - 1)];
      }
    }
    __syncthreads();
    // packing of the C2R (DC / Nyquist imaginary parts ignored), length-M1
    // FFTs over k1 of Z[k1 M2 + k2], twiddle
    auto pack = [&](int k2, int j, float2 (&v)[M1]) {
#pragma unroll
      for (int k1 = 0; k1 < M1; ++k1) {
        int const k = k1 * M2 + k2;
        float2 xk = IN[k * TG + j];
        float2 xc = IN[(M - k) * TG + j];
        if (k == 0) {
          xk.y = 0.f;
          xc.y = 0.f;
        }
        xc.y = -xc.y;
        float2 const e = agdfft::c_add(xk, xc);
        float2 const o = agdfft::c_mul(agdfft::c_sub(xk, xc), RP[k]);
        v[k1] = make_float2(e.x - o.y, e.y + o.x);
      }
    };
    auto phase1_out = [&](int k2, int j, float2 (&v)[M1]) {
      agdfft::fft<M1, +1>(v);
      float2 *zz = Z + j * LDL + k2;
      float2 const *tw = TW + k2 * M1;
#pragma unroll
      for (int n1 = 0; n1 < M1; ++n1)
        zz[n1 * M2] = (n1 == 0) ? v[n1] : agdfft::c_mul(v[n1], tw[n1]);
    };
    if constexpr (LT) {
      // every input of the tile read into registers before the tile, as Z,
      // is overwritten
      float2 v[ITEMS][M1];
#pragma unroll
      for (int q = 0; q < ITEMS; ++q) {
        int const i = min(tid + q * NT, NI - 1);
        pack(i / TG, i - (i / TG) * TG, v[q]);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < ITEMS; ++q) {
        int const i = tid + q * NT;
        if (i < NI) phase1_out(i / TG, i - (i / TG) * TG, v[q]);
      }
    } else {
      for (int i = tid; i < NI; i += NT) {
        float2 v[M1];
        pack(i / TG, i - (i / TG) * TG, v);
        phase1_out(i / TG, i - (i / TG) * TG, v);
      }
    }
    __syncthreads();
    for (int i = tid; i < TG * M1; i += NT) {
      int const n1 = i / TG;
      int const j = i - n1 * TG;
      float2 *zz = Z + j * LDL + n1 * M2;
      float2 v[M2];
#pragma unroll
      for (int k2 = 0; k2 < M2; ++k2) v[k2] = zz[k2];
      agdfft::fft<M2, +1>(v);
#pragma unroll
      for (int n2 = 0; n2 < M2; ++n2) zz[n2] = v[n2];
    }
    __syncthreads();
    // overlap-add: samples [0, half) complete this window's first half
    int const p0 = fft2_pos<M1, M2>(s >> 1);
    int const p1 = fft2_pos<M1, M2>((s + half) >> 1);
#pragma unroll
    for (int r = 0; r < NJ; ++r) {
      int const j = j0 + JS * r;
      float2 const z0 = Z[j * LDL + p0];
      float2 const z1 = Z[j * LDL + p1];
      float const y0 = (s & 1) ? z0.y : z0.x;
      float const y1 = (s & 1) ? z1.y : z1.x;
      float const sc = SC[j];
      if constexpr (LT) {
        if (tw >= tw0) Tb[j * nconv + tw * half + s] = ca[r] + y0 * sc * tap0;
        ca[r] = y1 * sc * tap1;
      } else {
        if (tw >= tw0) Tb[j * nconv + tw * half + s] = CA[j * half + s] + y0 * sc * tap0;
        CA[j * half + s] = y1 * sc * tap1;
      }
    }
  }
  if (tw1 < ntw) return;
  int const tl = ntw * half + s;
#pragma unroll
  for (int r = 0; r < NJ; ++r)
    if (tl < nconv)
      Tb[(j0 + JS * r) * nconv + tl] = LT ? ca[LT ? r : 0] : CA[(j0 + JS * r) * half + s];
  // zero the capacity beyond the data portion (wavelet input padding)
  for (int j = 0; j < TG; ++j)
    for (int t = (ntw + 1) * half + tid; t < nconv; t += NT) Tb[(int64_t)j * nconv + t] = 0.f;
}

// Adjoint time transform: record (or wavelet-adjoint traces) -> windowed,
// tapered time windows -> R2C (x2 interior bins) -> V tiles.
template <int M1, int M2>
__global__ __launch_bounds__(time_threads<M1 * M2>()) void k_adj_time_t(
    AgdWin const *__restrict__ wins, int const *__restrict__ blocks, int block0,
    AgdFftPlan const *__restrict__ plans, float2 const *__restrict__ roots,
    float const *__restrict__ tapers, float const *__restrict__ tscale,
    AgdEntry const *__restrict__ went, float const *__restrict__ rec,
    float const *__restrict__ Tin, float2 *__restrict__ V, int plain,
    int const *__restrict__ order) {
  constexpr int M = M1 * M2;
  constexpr int NT = time_threads<M>();
  constexpr int Ft = M + 1;
  constexpr int LDL = M + 1;
  constexpr int Pt = 2 * M;
  constexpr int half = 4 * M / 5;
  constexpr int Wt = 2 * half;
  constexpr int JS = NT / half;
  constexpr int NJ = TG / JS;
  constexpr int kmax2 = (Pt - 1) / 2;
  static_assert(NT % half == 0 && TG % JS == 0, "sample mapping");
  extern __shared__ float2 sm[];
  float2 *Z = sm;             // [TG][LDL]
  float2 *TW = Z + TG * LDL;  // twiddles [M2][M1]
  float2 *RP = TW + M;        // exp(-2 pi i k / Pt), k <= M
  // per trace: offset of its samples in src, index of its first sample
  // and count (the record through the blend entries, or a Tin trace)
  int64_t *erec = (int64_t *)(RP + Ft);
  int *ets = (int *)(erec + TG);
  int *en = ets + TG;
  // with an order: the class's blocks sorted by first trace, each XCD given
  // a contiguous run of them, so the (up to four) windows' blocks that read
  // the same traces' record samples meet in one L2
  int const bid = order ? order[block0 + xcd_chunk(blockIdx.x, gridDim.x)] : block0 + blockIdx.x;
  AgdWin const &w = wins[blocks[TB_INTS * bid]];
  int const t0 = blocks[TB_INTS * bid + 1];
  int const tw0 = blocks[TB_INTS * bid + 2], tw1 = blocks[TB_INTS * bid + 3];
  int const nconv = w.nconv;
  int const tid = threadIdx.x;
  {
    float2 const *rM = roots + plans[w.tplan].roots_off;
    float2 const *rP = roots + plans[w.tplan_full].roots_off;
    // TW[k2 M1 + n1] = exp(-2 pi i n1 k2 / M)
    for (int m = tid; m < M; m += NT) TW[m] = rM[((m % M1) * (m / M1)) % M];
    for (int k = tid; k < Ft; k += NT) RP[k] = rP[k];
    if (tid < TG) {
      if (Tin != nullptr) {
        erec[tid] = w.t_off + (int64_t)(t0 + tid) * nconv;
        ets[tid] = 0;
        en[tid] = nconv;
      } else {
        AgdEntry const e = went[w.trace0 + t0 + tid];
        erec[tid] = e.rec;
        ets[tid] = e.trace_start;
        en[tid] = e.n;
      }
    }
  }
  __syncthreads();
  int const s = tid % half;
  int const j0 = tid / half;
  float const tap0 = plain ? 1.f : tapers[w.tt_off + s];
  float const tap1 = plain ? 1.f : tapers[w.tt_off + half + s];
  float const two = plain ? 1.f : 2.f;
  // one uniform source; each sample load is unconditional (index clamped
  // into the trace's range) and zeroed outside it, so that the loads of a
  // time window stay in flight together
  float const *const src = Tin != nullptr ? Tin : rec;
  auto sample = [&](int j, int t) -> float {
    int const i2 = t - ets[j];
    bool const ok = (unsigned)i2 < (unsigned)en[j];
    float const x = src[erec[j] + (ok ? i2 : 0)];
    return ok ? x : 0.f;
  };
  float scl[NJ], cur[NJ], nxt[NJ];
#pragma unroll
  for (int r = 0; r < NJ; ++r) {
    int const j = j0 + JS * r;
    scl[r] = plain ? w.scaler : tscale[w.trace0 + t0 + j];
    cur[r] = sample(j, tw0 * half + s);
    nxt[r] = sample(j, (tw0 + 1) * half + s);
  }
  float2 *vt = V + w.u_off + (int64_t)(t0 / TG) * Ft * TG;
  int64_t const tw_step = (int64_t)(w.ntr / TG) * Ft * TG;
  for (int tw = tw0; tw < tw1; ++tw) {
#pragma unroll
    for (int r = 0; r < NJ; ++r) {
      float *zr = (float *)(Z + (j0 + JS * r) * LDL);
      zr[s] = cur[r] * scl[r] * tap0;
      zr[s + half] = nxt[r] * scl[r] * tap1;
      cur[r] = nxt[r];
    }
    for (int i = tid; i < TG * (Pt - Wt); i += NT) {
      int const j = i / (Pt - Wt);
      ((float *)(Z + j * LDL))[Wt + (i - j * (Pt - Wt))] = 0.f;
    }
    if (tw + 1 < tw1) {
#pragma unroll
      for (int r = 0; r < NJ; ++r) nxt[r] = sample(j0 + JS * r, (tw + 2) * half + s);
    }
    __syncthreads();
    for (int i = tid; i < TG * M2; i += NT) {
      int const k2 = i / TG;
      int const j = i - k2 * TG;
      float2 *zz = Z + j * LDL + k2;
      float2 v[M1];
#pragma unroll
      for (int k1 = 0; k1 < M1; ++k1) v[k1] = zz[k1 * M2];
      agdfft::fft<M1, -1>(v);
      float2 const *tw = TW + k2 * M1;
#pragma unroll
      for (int n1 = 0; n1 < M1; ++n1)
        zz[n1 * M2] = (n1 == 0) ? v[n1] : agdfft::c_mul(v[n1], tw[n1]);
    }
    __syncthreads();
    for (int i = tid; i < TG * M1; i += NT) {
      int const n1 = i / TG;
      int const j = i - n1 * TG;
      float2 *zz = Z + j * LDL + n1 * M2;
      float2 v[M2];
#pragma unroll
      for (int k2 = 0; k2 < M2; ++k2) v[k2] = zz[k2];
      agdfft::fft<M2, -1>(v);
#pragma unroll
      for (int n2 = 0; n2 < M2; ++n2) zz[n2] = v[n2];
    }
    __syncthreads();
    // half spectrum of the real window, x2 on the interior bins; the pair
    // (k, M - k) shares its two inputs. A thread takes two neighbouring
    // traces of a bin pair, so each of its stores is 16 bytes.
    float2 *vo = vt + (int64_t)tw * tw_step;
    for (int i = tid; i < (M / 2 + 1) * (TG / 2); i += NT) {
      int const k = i / (TG / 2);
      int const j0p = 2 * (i - k * (TG / 2));
      int const kc = (k == 0) ? 0 : M - k;
      float2 X1[2], X2[2];
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        int const j = j0p + jj;
        float2 const za = Z[j * LDL + fft2_pos<M1, M2>(k)];
        float2 const zb = Z[j * LDL + fft2_pos<M1, M2>(kc)];
        // X[k] from (za, conj zb), X[M - k] from (zb, conj za)
        float2 const e1 = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));
        float2 const o1 = make_float2(0.5f * (za.y + zb.y), -0.5f * (za.x - zb.x));
        X1[jj] = agdfft::c_add(e1, agdfft::c_mul(o1, RP[k]));
        X2[jj] = agdfft::c_add(make_float2(e1.x, -e1.y),
                               agdfft::c_mul(make_float2(o1.x, -o1.y), RP[M - k]));
        if (k == 0) {
          // DC, and Nyquist (RP[M] = -1), both real
          X1[jj].y = 0.f;
          X2[jj] = make_float2(za.x - za.y, 0.f);
        }
        if (k >= 1 && k <= kmax2) {
          X1[jj].x *= two;
          X1[jj].y *= two;
        }
        if (M - k >= 1 && M - k <= kmax2) {
          X2[jj].x *= two;
          X2[jj].y *= two;
        }
      }
      *(float4 *)(vo + k * TG + j0p) = make_float4(X1[0].x, X1[0].y, X1[1].x, X1[1].y);
      if (k != M - k)
        *(float4 *)(vo + (M - k) * TG + j0p) = make_float4(X2[0].x, X2[0].y, X2[1].x, X2[1].y);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Split-precision (bf16 x 3) spatial DFTs on the matrix cores, beside the
// FP32 register FFTs above (selected per problem, agdx_set_spatial_precision).
// A window's spatial transform of one plane is two complex matrix products
// with constant DFT matrices: forward U = A0 X A1^T (A[n][c] = e^{+2 pi i n c / P},
// n < W, c < P: the inverse DFT cropped to the window), adjoint
// G = B0 V B1^T (B[c][n] = e^{-2 pi i c n / P}: the DFT of the zero-padded
// traces). Each complex product is one real GEMM on stacked real and
// imaginary parts ([Yr; Yi] = [[Cr, -Ci]; [Ci, Cr]] [Zr; Zi]); each FP32
// operand v is split into bf16 parts h = bf16(v), l = bf16(v - h), and the
// product accumulates h.h + h.l + l.h in FP32 on v_mfma_f32_16x16x16_bf16
// (the dropped l.l term is ~2^-16 relative). Square windows, P in {20, 40},
// tiled U / V; one plane at a time per block, its operands in LDS.
// ---------------------------------------------------------------------------
namespace x3 {
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// v_mfma_f32_16x16x32_bf16: lane l holds A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16], e < 8; D[4 (l / 16) + r][l % 16], r < 4
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int const u = __float_as_uint(f);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // nearest even
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float((unsigned int)h << 16);
}
template <int N, typename V>
__device__ __forceinline__ void split(float const (&v)[N], V &hi, V &lo) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    unsigned short const h = f2bf(v[r]);
    hi[r] = (short)h;
    lo[r] = (short)f2bf(v[r] - bf2f(h));
  }
}
typedef short bf16x4 __attribute__((ext_vector_type(4)));
// coefficient of output part po (0 re, 1 im) from input part pi of the
// complex factor c: [[Cr, -Ci], [Ci, Cr]]
__device__ __forceinline__ float part(float2 c, int po, int pi) {
  return po == pi ? c.x : (pi == 1 ? -c.y : c.y);
}
// e^{SIGN 2 pi i m / P}
template <int P, int SIGN>
__device__ __forceinline__ float2 coef(int m) {
  float s, c;
  sincospif(2.f * (float)(m % P) / (float)P, &s, &c);
  return make_float2(c, SIGN * s);
}
__host__ __device__ constexpr int r16(int n) { return (n + 15) / 16 * 16; }

template <int P>
struct Shape {
  static constexpr int W = P * 4 / 5;
  static constexpr int KP = r16(P);
  static constexpr int KW = r16(W);
  static constexpr int BP = KP + 8;   // bf16 per LDS row of the split operands
  static constexpr int BL = KP * BP;  // bf16 of one (part, hi / lo) array
  static constexpr int FP = KP + 4;   // floats per LDS row of the fp32 output
  static constexpr int FL = KP * FP;  // floats of one part of the fp32 output
};

// The split data live in LDS as bf16 hi / lo arrays per part (real,
// imaginary): s[(part * 2 + lo) * BL + row * BP + col]. A split array is
// read 8 consecutive elements (16 bytes) at a time along its rows (the row
// pitch BP = KP + 8 puts the eight rows of a read cycle in distinct bank
// groups).
__device__ __forceinline__ bf16x8 ld8(uint16_t const *p) { return *(bf16x8 const *)p; }

// Y = Z C^T, C[o][q] = e^{SIGN 2 pi i o q / P} for o < NC, q < KC (else 0):
// Z split row-major ([row][q]), rows r16(MZ); Y written split and
// column-major ([o][row]) for the next product. Waves own 16-column tiles
// of the stacked output; the constant operands are built once (init).
template <int P, int MZ, int KC, int NC, int SIGN>
struct Right {
  static constexpr int BP = Shape<P>::BP, BL = Shape<P>::BL;
  static constexpr int KMZ = r16(MZ), KKC = r16(KC), KNC = r16(NC);
  static constexpr int NNT = 2 * KNC / 16, NKS = 2 * KKC / 32, NMT = KMZ / 16;
  static constexpr int TPW = (NNT + NTHR / 64 - 1) / (NTHR / 64);
  bf16x8 h[TPW][NKS], l[TPW][NKS];

  __device__ __forceinline__ void init() {
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const i = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      int const nt = wave + t * (NTHR / 64);
      int const nc = nt * 16 + i;
      int const po = nc / KNC, o = nc - po * KNC;
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        float v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          int const kk = ks * 32 + 8 * kq + r;
          int const pi = kk / KKC, q = kk - pi * KKC;
          float2 const c = (nt < NNT && o < NC && q < KC) ? coef<P, SIGN>(o * q)
                                                          : make_float2(0.f, 0.f);
          v[r] = part(c, po, pi);
        }
        split<8>(v, h[t][ks], l[t][ks]);
      }
    }
  }

  __device__ __forceinline__ void apply(uint16_t const *__restrict__ z,
                                        uint16_t *__restrict__ y) const {
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const i = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      int const nt = wave + t * (NTHR / 64);
      if (nt >= NNT) break;
      int const nc = nt * 16 + i;
      int const po = nc / KNC, o = nc - po * KNC;
      for (int mt = 0; mt < NMT; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        int const row = mt * 16 + i;
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
          int const kk = ks * 32 + 8 * kq;  // 8 elements of one part (KKC % 16 == 0)
          int const pi = kk / KKC, q = kk - pi * KKC;
          bf16x8 const ah = ld8(z + (pi * 2) * BL + row * BP + q);
          bf16x8 const al = ld8(z + (pi * 2 + 1) * BL + row * BP + q);
          acc = mma(ah, h[t][ks], acc);
          acc = mma(ah, l[t][ks], acc);
          acc = mma(al, h[t][ks], acc);
        }
        // D: rows mt 16 + 4 kq + r of column nc: 4 consecutive rows of the
        // column-major output
        float const v[4] = {acc[0], acc[1], acc[2], acc[3]};
        bf16x4 vh, vl;
        split<4>(v, vh, vl);
        int const off = o * BP + mt * 16 + 4 * kq;
        *(bf16x4 *)(y + (po * 2) * BL + off) = vh;
        *(bf16x4 *)(y + (po * 2 + 1) * BL + off) = vl;
      }
    }
  }
};

// Y = C Z, C[p][q] = e^{SIGN 2 pi i p q / P} for p < MC, q < KC (else 0):
// Z split column-major ([col][q]), columns r16(NZ); Y in fp32 row-major
// ([part][p][col], FP floats per row). Waves own 16-row tiles of the
// stacked output; the constant operands are built once (init).
template <int P, int MC, int KC, int NZ, int SIGN>
struct Left {
  static constexpr int BP = Shape<P>::BP, BL = Shape<P>::BL;
  static constexpr int FP = Shape<P>::FP, FL = Shape<P>::FL;
  static constexpr int KMC = r16(MC), KKC = r16(KC), KNZ = r16(NZ);
  static constexpr int NMT = 2 * KMC / 16, NKS = 2 * KKC / 32, NNT = KNZ / 16;
  static constexpr int TPW = (NMT + NTHR / 64 - 1) / (NTHR / 64);
  bf16x8 h[TPW][NKS], l[TPW][NKS];

  __device__ __forceinline__ void init() {
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const i = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      int const mt = wave + t * (NTHR / 64);
      int const mr = mt * 16 + i;
      int const po = mr / KMC, pr = mr - po * KMC;
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        float v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          int const kk = ks * 32 + 8 * kq + r;
          int const pi = kk / KKC, q = kk - pi * KKC;
          float2 const c = (mt < NMT && pr < MC && q < KC) ? coef<P, SIGN>(pr * q)
                                                           : make_float2(0.f, 0.f);
          v[r] = part(c, po, pi);
        }
        split<8>(v, h[t][ks], l[t][ks]);
      }
    }
  }

  __device__ __forceinline__ void apply(uint16_t const *__restrict__ z,
                                        float *__restrict__ y) const {
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const i = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      int const mt = wave + t * (NTHR / 64);
      if (mt >= NMT) break;
      for (int nt = 0; nt < NNT; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        int const col = nt * 16 + i;
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
          int const kk = ks * 32 + 8 * kq;  // 8 elements of one part (KKC % 16 == 0)
          int const pi = kk / KKC, q = kk - pi * KKC;
          bf16x8 const bh = ld8(z + (pi * 2) * BL + col * BP + q);
          bf16x8 const bl = ld8(z + (pi * 2 + 1) * BL + col * BP + q);
          acc = mma(h[t][ks], bh, acc);
          acc = mma(h[t][ks], bl, acc);
          acc = mma(l[t][ks], bh, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int const yr = mt * 16 + 4 * kq + r;
          int const yp = yr / KMC;
          y[yp * FL + (yr - yp * KMC) * FP + col] = acc[r];
        }
      }
    }
  }
};

// one complex value split into the four arrays at (row, col)
__device__ __forceinline__ void put_split(uint16_t *s, int BL, int off, float2 v) {
  unsigned short const rh = f2bf(v.x), ih = f2bf(v.y);
  s[off] = rh;
  s[BL + off] = f2bf(v.x - bf2f(rh));
  s[2 * BL + off] = ih;
  s[3 * BL + off] = f2bf(v.y - bf2f(ih));
}
}  // namespace x3

// X -> U (the forward spatial stage), bf16 x 3 DFT matrices:
// Y1 = X A1^T (columns), then U = A0 Y1 (rows)
template <int P>
__global__ __launch_bounds__(NTHR) void k_fwd_spatial_x3(
    AgdWin const *__restrict__ wins, AgdTask const *__restrict__ tasks, int task0,
    float2 const *__restrict__ X, float2 *__restrict__ U) {
  using S = x3::Shape<P>;
  constexpr int W = S::W, KP = S::KP, BP = S::BP, BL = S::BL, FP = S::FP, FL = S::FL;
  static_assert(2 * FL * 4 <= 4 * BL * 2, "the fp32 output fits the input's space");
  __shared__ __align__(16) uint16_t za[4 * BL];  // X split, row-major; then U (fp32)
  __shared__ __align__(16) uint16_t zb[4 * BL];  // Y1 split, column-major
  float *const zo = (float *)za;
  AgdTask const t = tasks[task0 + blockIdx.x];
  AgdWin const &w = wins[t.win];
  float2 const *const src = X + w.x_off + ((int64_t)t.tw * w.Ft + t.k0) * (P * P);
  float2 *const Utw = U + w.u_off + (int64_t)t.tw * w.Ft * (W * W);
  x3::Right<P, P, P, W, +1> a1;  // Y1 = X A1^T: rows c0, columns n1
  x3::Left<P, W, P, W, +1> a0;   // U = A0 Y1: rows n0, columns n1
  a1.init();
  a0.init();
  // plane k + 1 is loaded into registers while plane k is transformed
  constexpr int NE = (KP * KP + NTHR - 1) / NTHR;
  auto load = [&](int k, float2 (&v)[NE]) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      int const e = threadIdx.x + q * NTHR;
      int const r = e / KP, c = e - r * KP;
      bool const ok = e < KP * KP && r < P && c < P && k < t.nk;
      float2 const u = src[ok ? (int64_t)k * P * P + r * P + c : 0];
      v[q] = ok ? u : make_float2(0.f, 0.f);
    }
  };
  float2 cur[NE];
  load(0, cur);
  for (int k = 0; k < t.nk; ++k) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      int const e = threadIdx.x + q * NTHR;
      if (e < KP * KP) x3::put_split(za, BL, (e / KP) * BP + e % KP, cur[q]);
    }
    __syncthreads();
    load(k + 1, cur);
    a1.apply(za, zb);
    __syncthreads();
    a0.apply(zb, zo);
    __syncthreads();
    for (int tr = threadIdx.x; tr < W * W; tr += NTHR) {
      int const n0 = tr / W, n1 = tr - n0 * W;
      Utw[((int64_t)(tr / TG) * w.Ft + t.k0 + k) * TG + (tr % TG)] =
          make_float2(zo[n0 * FP + n1], zo[FL + n0 * FP + n1]);
    }
  }
}

// V -> G and the model update by MODE (as k_adj_spatial_t), bf16 x 3:
// R = V B1^T (columns), then G = B0 R (rows)
template <int P, int MODE>
__global__ __launch_bounds__(NTHR) void k_adj_spatial_x3(
    AgdWin const *__restrict__ wins, AgdTask const *__restrict__ tasks, int task0,
    float2 const *__restrict__ V, float2 *__restrict__ X, float alpha, float lamb,
    float const *__restrict__ lamb_dev, unsigned int *__restrict__ maxbits) {
  using S = x3::Shape<P>;
  constexpr int W = S::W, KW = S::KW, BP = S::BP, BL = S::BL, FP = S::FP, FL = S::FL;
  static_assert(2 * FL * 4 <= 4 * BL * 2, "the fp32 output fits the input's space");
  __shared__ __align__(16) uint16_t za[4 * BL];  // V split, row-major; then G (fp32)
  __shared__ __align__(16) uint16_t zb[4 * BL];  // R split, column-major
  float *const zo = (float *)za;
  if (lamb_dev != nullptr) lamb = lamb_dev[0];
  AgdTask const t = tasks[task0 + blockIdx.x];
  AgdWin const &w = wins[t.win];
  float2 const *const Vtw = V + w.u_off + (int64_t)t.tw * w.Ft * (W * W);
  float2 *const x = X + w.x_off + ((int64_t)t.tw * w.Ft + t.k0) * (P * P);
  x3::Right<P, W, W, P, -1> b1;  // R = V B1^T: rows n0, columns c1
  x3::Left<P, P, W, P, -1> b0;   // G = B0 R: rows c0, columns c1
  b1.init();
  b0.init();
  float mx = 0.f;
  // plane k + 1 is loaded into registers while plane k is transformed
  constexpr int NE = (KW * KW + NTHR - 1) / NTHR;
  auto load = [&](int k, float2 (&v)[NE]) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      int const e = threadIdx.x + q * NTHR;
      int const n0 = e / KW, n1 = e - n0 * KW;
      bool const ok = e < KW * KW && n0 < W && n1 < W && k < t.nk;
      int const tr = ok ? n0 * W + n1 : 0;
      float2 const u = Vtw[((int64_t)(tr / TG) * w.Ft + t.k0 + (ok ? k : 0)) * TG + (tr % TG)];
      v[q] = ok ? u : make_float2(0.f, 0.f);
    }
  };
  // the model values the update reads (modes 0 and 2), prefetched likewise
  constexpr int NX = (MODE == 0 || MODE == 2) ? (P * P + NTHR - 1) / NTHR : 1;
  auto load_x = [&](int k, float2 (&v)[NX]) {
    if constexpr (MODE == 0 || MODE == 2) {
#pragma unroll
      for (int q = 0; q < NX; ++q) {
        int const e = threadIdx.x + q * NTHR;
        bool const ok = e < P * P && k < t.nk;
        float2 const u = x[ok ? (int64_t)k * P * P + e : 0];
        v[q] = ok ? u : make_float2(0.f, 0.f);
      }
    }
  };
  float2 cur[NE], xcur[NX];
  load(0, cur);
  load_x(0, xcur);
  for (int k = 0; k < t.nk; ++k) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      int const e = threadIdx.x + q * NTHR;
      if (e < KW * KW) x3::put_split(za, BL, (e / KW) * BP + e % KW, cur[q]);
    }
    __syncthreads();
    load(k + 1, cur);
    float2 xnow[NX];
#pragma unroll
    for (int q = 0; q < NX; ++q) xnow[q] = xcur[q];
    load_x(k + 1, xcur);
    b1.apply(za, zb);
    __syncthreads();
    b0.apply(zb, zo);
    __syncthreads();
    float2 *const xk = x + (int64_t)k * P * P;
#pragma unroll
    for (int q = 0; q < (P * P + NTHR - 1) / NTHR; ++q) {
      int const e = threadIdx.x + q * NTHR;
      if (e >= P * P) break;
      int const c0 = e / P, c1 = e - c0 * P;
      float2 const gv = make_float2(zo[c0 * FP + c1], zo[FL + c0 * FP + c1]);
      float2 nv;
      if constexpr (MODE == 3) {
        nv = gv;
      } else if constexpr (MODE == 1) {
        nv = make_float2(0.f - alpha * gv.x, 0.f - alpha * gv.y);
        mx = fmaxf(mx, nv.x * nv.x + nv.y * nv.y);
      } else {
        nv = xnow[q];
        nv.x -= alpha * gv.x;
        nv.y -= alpha * gv.y;
        if constexpr (MODE == 0) {
          float const a = nv.x * nv.x + nv.y * nv.y;
          float const sh = fmaxf(1.f - lamb * __builtin_amdgcn_rsqf(a), 0.f);
          nv.x *= sh;
          nv.y *= sh;
        }
      }
      xk[e] = nv;
    }
  }
  if constexpr (MODE == 1) {
    __shared__ float red[NTHR / 64];
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = red[0];
      for (int q = 1; q < NTHR / 64; ++q) m = fmaxf(m, red[q]);
      atomicMax(maxbits, __float_as_uint(m));
    }
  }
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
// Sizes with specialised kernels; the host groups work by these keys
// (agd_host.c keeps the same lists).
#define SPATIAL_SIZES(X) \
  X(1, 10) X(1, 20) X(1, 40) X(10, 10) X(10, 20) X(10, 40) X(20, 10) X(20, 20) \
  X(20, 40) X(40, 10) X(40, 20) X(40, 40)
// (M1, M2) factorizations of the half length M = Pt / 2 with a kernel: the
// first of each M is the default, the others are the autotuner's candidates
#define TIME_SIZES(X)                                                            \
  X(5, 8) X(8, 5) X(4, 10) X(10, 8) X(8, 10) X(16, 5) X(10, 16) X(16, 10) X(20, 8) \
  X(8, 20) X(20, 16) X(16, 20) X(32, 10) X(10, 32)

// units of class c within the launch's window range: [*first, *first + n)
static inline int class_range(AgdDev const *g, AgdClass const &c, int *first) {
  int const a = c.wpos[g->w0], b = c.wpos[g->w1];
  *first = c.begin + a;
  return b - a;
}

// the specialised kernels may use up to the CU's 160 KB of LDS: set once per
// kernel instance, on its first launch
static int allow_lds(void const *f) {
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  (void)hipGetLastError();
  return 1;
}

template <int P0, int P1, bool TILED, int MODE, bool FUSE>
static int launch_adj_spatial(AgdDev const *g, AgdClass const &c, float alpha, float lamb) {
  static int const lds_ok = allow_lds((void const *)k_adj_spatial_t<P0, P1, TILED, MODE, FUSE>);
  (void)lds_ok;
  int first;
  int const n = class_range(g, c, &first);
  if (n <= 0) return 0;
  return launch("adj_spatial", k_adj_spatial_t<P0, P1, TILED, MODE, FUSE>, dim3(n),
                dim3(NTHR), c.lds, g->win, g->tasks, first, (float2 const *)g->V,
                (float2 *)g->X, (float2 *)g->U, alpha, lamb, (float const *)g->lamb_dev,
                g->scalar);
}

template <int P0, int P1, bool TILED>
static int adj_spatial_layout(AgdDev const *g, AgdClass const &c, float alpha, float lamb,
                              int mode, int fuse) {
  switch (mode) {
    case 0:
      return fuse ? launch_adj_spatial<P0, P1, TILED, 0, true>(g, c, alpha, lamb)
                  : launch_adj_spatial<P0, P1, TILED, 0, false>(g, c, alpha, lamb);
    case 1: return launch_adj_spatial<P0, P1, TILED, 1, false>(g, c, alpha, lamb);
    case 2: return launch_adj_spatial<P0, P1, TILED, 2, false>(g, c, alpha, lamb);
    case 4: return launch_adj_spatial<P0, P1, TILED, 4, false>(g, c, alpha, lamb);
    default:
      return fuse ? launch_adj_spatial<P0, P1, TILED, 3, true>(g, c, alpha, lamb)
                  : launch_adj_spatial<P0, P1, TILED, 3, false>(g, c, alpha, lamb);
  }
}

template <int P0, int P1>
static int adj_spatial_sized(AgdDev const *g, AgdClass const &c, float alpha, float lamb,
                             int mode, int fuse) {
  if (fuse && mode != 0 && mode != 3) {
    fprintf(stderr, "ERROR: fused update requested for mode %d\n", mode);
    return 1;
  }
  return c.tb ? adj_spatial_layout<P0, P1, true>(g, c, alpha, lamb, mode, fuse)
              : adj_spatial_layout<P0, P1, false>(g, c, alpha, lamb, mode, fuse);
}

template <int P0, int P1, bool TILED>
static int launch_fwd_spatial(AgdDev const *g, AgdClass const &c, float const *X) {
  static int const lds_ok = allow_lds((void const *)k_fwd_spatial_t<P0, P1, TILED>);
  (void)lds_ok;
  int first;
  int const n = class_range(g, c, &first);
  if (n <= 0) return 0;
  return launch("fwd_spatial", k_fwd_spatial_t<P0, P1, TILED>, dim3(n), dim3(NTHR), c.lds,
                g->win, g->tasks, first, (float2 const *)X, (float2 *)g->U);
}

template <int P0, int P1>
static int fwd_spatial_sized(AgdDev const *g, AgdClass const &c, float const *X) {
  return c.tb ? launch_fwd_spatial<P0, P1, true>(g, c, X)
              : launch_fwd_spatial<P0, P1, false>(g, c, X);
}

// ---------------------------------------------------------------------------
// Fixed scratch of the scratch-free entry point (agd_blend_noalloc): every
// device buffer of such a call is carved from this array, so the call makes
// no allocation. Each device holds its own copy from module load until the
// process ends. Host code never passes `scratch` itself (in host code the
// name is a host-side stand-in, not a device address): the device address
// of each device's copy is looked up once, with that device current, and
// kept in g_scratch_dev.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static void *g_scratch_dev[kMaxDevices];
static int g_device = -1;  // the device of g_stream

// device address of `scratch` on the current device (looked up once per device)
static int scratch_lookup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    fprintf(stderr, "ERROR: device %d is beyond the scratch table\n", dev);
    return 1;
  }
  if (g_scratch_dev[dev] != nullptr) return 0;
  void *p = nullptr;
  if (hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
    return 1;
  g_scratch_dev[dev] = p;
  return 0;
}

extern "C" {

int agdg_init(void) {
  if (g_stream != nullptr) return 0;
  int n = 0;
  if (hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount")) return 1;
  if (n < 1) {
    fprintf(stderr, "ERROR: no HIP device available\n");
    return 1;
  }
  if (hip_check(hipGetDevice(&g_device), "hipGetDevice")) return 1;
  if (scratch_lookup(g_device)) return 1;
  if (hip_check(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking),
                "hipStreamCreate"))
    return 1;
  // allow dynamic LDS up to the CU's 160 KB for the LDS-resident transforms
  {
    int const lds = 160 * 1024;
    (void)hipFuncSetAttribute((void const *)k_fwd_spatial,
                        hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((void const *)k_adj_spatial,
                        hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((void const *)k_fwd_time,
                        hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((void const *)k_adj_time,
                        hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((void const *)k_wavelet,
                        hipFuncAttributeMaxDynamicSharedMemorySize, lds);
#define ATTR_T(a, b)                                                              \
    (void)hipFuncSetAttribute((void const *)k_fwd_time_t<a, b, false>,            \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds);   \
    (void)hipFuncSetAttribute((void const *)k_fwd_time_t<a, b, true>,             \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds);   \
    (void)hipFuncSetAttribute((void const *)k_adj_time_t<a, b>,                   \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    TIME_SIZES(ATTR_T)
#undef ATTR_T
    (void)hipGetLastError();
  }
  return 0;
}

int agdg_scratch(void **ptr, size_t *bytes) {
  int dev = -1;
  if (agdg_init()) return 1;
  if (hip_check(hipGetDevice(&dev), "hipGetDevice")) return 1;
  if (dev != g_device) {
    fprintf(stderr, "ERROR: the library runs on device %d, the current device is %d\n",
            g_device, dev);
    return 1;
  }
  if (scratch_lookup(dev)) return 1;
  *ptr = g_scratch_dev[dev];
  *bytes = kScratchBytes;
  return 0;
}

void *agdg_alloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 16;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    fprintf(stderr, "ERROR: hipMalloc of %zu bytes failed\n", bytes);
    return nullptr;
  }
  return p;
}

void agdg_free(void *p) {
  if (p) (void)hipFree(p);
}

int agdg_h2d(void *dst, void const *src, size_t bytes) {
  if (bytes == 0) return 0;
  return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream),
                   "memcpy H2D") ||
         hip_check(hipStreamSynchronize(g_stream), "sync");
}

int agdg_d2h(void *dst, void const *src, size_t bytes) {
  if (bytes == 0) return 0;
  return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream),
                   "memcpy D2H") ||
         hip_check(hipStreamSynchronize(g_stream), "sync");
}

int agdg_zero(void *dst, size_t bytes) {
  if (bytes == 0) return 0;
  return hip_check(hipMemsetAsync(dst, 0, bytes, g_stream), "memset");
}

int agdg_sync(void) {
  int rc = hip_check(hipStreamSynchronize(g_stream), "stream sync");
  if (g_timing) flush_timing();
  return rc;
}

void *agdg_stream(void) { return (void *)g_stream; }

int agdg_can_fuse(AgdDev const *g) {
  if (g->precision != 0) return 0;  // the split-precision transforms are not fused
  for (int c = 0; c < g->n_sclass; ++c)
    if (g->sclass[c].p0 == 0) return 0;
  return 1;
}

// classes served by the split-precision transforms: square P in {20, 40},
// tiled U / V
static inline int x3_class(AgdDev const *g, AgdClass const &c) {
  return g->precision == 1 && c.tb && c.p0 == c.p1 && (c.p0 == 20 || c.p0 == 40);
}

int agdg_x3_classes(AgdDev const *g) {
  int n = 0;
  for (int ci = 0; ci < g->n_sclass; ++ci) {
    AgdClass c = g->sclass[ci];
    AgdDev h = *g;
    h.precision = 1;
    n += x3_class(&h, c);
  }
  return n;
}

int agdg_fwd_spatial(AgdDev const *g, float const *X) {
  for (int ci = 0; ci < g->n_sclass; ++ci) {
    AgdClass const &c = g->sclass[ci];
    bool done = false;
    if (x3_class(g, c)) {
      int first;
      int const n = class_range(g, c, &first);
      if (n > 0 &&
          (c.p0 == 20 ? launch("fwd_spatial_x3", k_fwd_spatial_x3<20>, dim3(n), dim3(NTHR), 0,
                               g->win, g->tasks, first, (float2 const *)X, (float2 *)g->U)
                      : launch("fwd_spatial_x3", k_fwd_spatial_x3<40>, dim3(n), dim3(NTHR), 0,
                               g->win, g->tasks, first, (float2 const *)X, (float2 *)g->U)))
        return 1;
      continue;
    }
#define FWD_CASE(a, b)                                                    \
    if (!done && c.p0 == a && c.p1 == b) {                                \
      if (fwd_spatial_sized<a, b>(g, c, X)) return 1;                     \
      done = true;                                                        \
    }
    SPATIAL_SIZES(FWD_CASE)
#undef FWD_CASE
    int first;
    int const n = class_range(g, c, &first);
    if (!done && n > 0 &&
        launch("fwd_spatial", k_fwd_spatial, dim3(n), dim3(NTHR), g->lds_spatial,
               g->win, g->tasks, first, g->plans, (float2 const *)g->roots,
               (float2 const *)X, (float2 *)g->U))
      return 1;
  }
  return 0;
}

int agdg_fwd_time(AgdDev const *g) {
  for (int ci = 0; ci < g->n_tclass; ++ci) {
    AgdClass const &c = g->tclass[ci];
    bool done = false;
    int first;
    int const n = class_range(g, c, &first);
#define FT_CASE(a, b)                                                                   \
    if (!done && c.p0 == a && c.p1 == b) {                                              \
      if (n > 0 && c.var == 0 &&                                                        \
          launch("fwd_time", k_fwd_time_t<a, b, false>, dim3(n),                        \
                 dim3(time_threads<a * b>()), fwd_time_lds<a * b, false>(), g->win,     \
                 g->tblocks, first, g->plans, (float2 const *)g->roots, g->tapers,      \
                 g->tscale, (float2 const *)g->U, g->T))                                \
        return 1;                                                                       \
      if (n > 0 && c.var != 0 &&                                                        \
          launch("fwd_time", k_fwd_time_t<a, b, true>, dim3(n),                         \
                 dim3(time_threads<a * b>()), fwd_time_lds<a * b, true>(), g->win,      \
                 g->tblocks, first, g->plans, (float2 const *)g->roots, g->tapers,      \
                 g->tscale, (float2 const *)g->U, g->T))                                \
        return 1;                                                                       \
      done = true;                                                                      \
    }
    TIME_SIZES(FT_CASE)
#undef FT_CASE
    if (!done && n > 0 &&
        launch("fwd_time", k_fwd_time, dim3(n), dim3(NTHR), g->lds_time, g->win,
               g->tblocks, first, c.tb, g->plans, (float2 const *)g->roots, g->tapers,
               g->tscale, (float2 const *)g->U, g->T))
      return 1;
  }
  return 0;
}

int agdg_blend_traces(AgdDev const *g, int resid) {
  return launch("blend", resid ? k_blend<3> : k_blend<0>, dim3(g->n_tiles), dim3(NTHR), 0,
                g->rec_len, g->tile, g->tile_ptr, g->tile_ent, g->went,
                (float const *)g->T, (float const *)nullptr, (float const *)nullptr, 0,
                g->mute, (float const *)g->d, g->r);
}

int agdg_stepsize_fold(AgdDev const *g) {
  if (launch("fold", k_blend<4>, dim3(g->n_tiles), dim3(NTHR), 0, g->rec_len, g->tile,
             g->tile_ptr, g->tile_ent, g->went, (float const *)nullptr, g->tconst,
             (float const *)nullptr, 0, (unsigned char const *)nullptr,
             (float const *)nullptr, g->r))
    return 1;
  int blocks = (int)((g->rec_len + NTHR - 1) / NTHR);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  return launch("max", k_max_bits, dim3(blocks), dim3(NTHR), 0, (float const *)g->r,
                g->rec_len, g->scalar + 1);
}

int agdg_adj_time(AgdDev const *g, float const *Tin, int plain) {
  for (int ci = 0; ci < g->n_tclass; ++ci) {
    AgdClass const &c = g->tclass[ci];
    bool done = false;
    int first;
    int const n = class_range(g, c, &first);
#define AT_CASE(a, b)                                                                   \
    if (!done && c.p0 == a && c.p1 == b) {                                              \
      if (n > 0 &&                                                                      \
          launch("adj_time", k_adj_time_t<a, b>, dim3(n),                                 \
                 dim3(time_threads<a * b>()), adj_time_lds<a * b>(), g->win, g->tblocks,   \
                 first, g->plans, (float2 const *)g->roots, g->tapers, g->tscale,        \
                 g->went, (float const *)g->r, Tin, (float2 *)g->V, plain,              \
                 first == c.begin && n == c.count ? g->tadj : (int const *)nullptr))    \
        return 1;                                                                       \
      done = true;                                                                      \
    }
    TIME_SIZES(AT_CASE)
#undef AT_CASE
    if (!done && n > 0 &&
        launch("adj_time", k_adj_time, dim3(n), dim3(NTHR), g->lds_time, g->win,
               g->tblocks, first, c.tb, g->plans, (float2 const *)g->roots, g->tapers,
               g->tscale, g->went, (float const *)g->r, Tin, (float2 *)g->V, plain))
      return 1;
  }
  return 0;
}

}  // extern "C"

extern "C" {

}  // extern "C"

template <int P>
static int launch_adj_x3(AgdDev const *g, AgdClass const &c, float alpha, float lamb,
                         int mode) {
  int first;
  int const n = class_range(g, c, &first);
  if (n <= 0) return 0;
#define X3_ARGS g->win, g->tasks, first, (float2 const *)g->V, (float2 *)g->X, alpha, lamb, \
                (float const *)g->lamb_dev, g->scalar
  switch (mode) {
    case 0: return launch("adj_spatial_x3", k_adj_spatial_x3<P, 0>, dim3(n), dim3(NTHR), 0, X3_ARGS);
    case 1: return launch("adj_spatial_x3", k_adj_spatial_x3<P, 1>, dim3(n), dim3(NTHR), 0, X3_ARGS);
    case 2: return launch("adj_spatial_x3", k_adj_spatial_x3<P, 2>, dim3(n), dim3(NTHR), 0, X3_ARGS);
    default: return launch("adj_spatial_x3", k_adj_spatial_x3<P, 3>, dim3(n), dim3(NTHR), 0, X3_ARGS);
  }
#undef X3_ARGS
}

extern "C" {

int agdg_adj_spatial(AgdDev const *g, float alpha, float lamb, int mode, int fuse) {
  for (int ci = 0; ci < g->n_sclass; ++ci) {
    AgdClass const &c = g->sclass[ci];
    bool done = false;
    if (x3_class(g, c)) {
      if (fuse) {
        fprintf(stderr, "ERROR: fused update requested with split-precision transforms\n");
        return 1;
      }
      // mode 4 as mode 1 here (the model is written, then overwritten)
      int const m = mode == 4 ? 1 : mode;
      if (c.p0 == 20 ? launch_adj_x3<20>(g, c, alpha, lamb, m)
                     : launch_adj_x3<40>(g, c, alpha, lamb, m))
        return 1;
      continue;
    }
#define AS_CASE(a, b)                                                       \
    if (!done && c.p0 == a && c.p1 == b) {                                  \
      if (adj_spatial_sized<a, b>(g, c, alpha, lamb, mode, fuse)) return 1; \
      done = true;                                                          \
    }
    SPATIAL_SIZES(AS_CASE)
#undef AS_CASE
    if (!done) {
      if (fuse) {
        fprintf(stderr, "ERROR: fused update requested for a generic window size\n");
        return 1;
      }
      int first;
      int const n = class_range(g, c, &first);
      if (n > 0 &&
          launch("adj_spatial", k_adj_spatial, dim3(n), dim3(NTHR), g->lds_spatial,
                 g->win, g->tasks, first, g->plans, (float2 const *)g->roots,
                 (float2 const *)g->V, (float2 *)g->X, alpha, lamb,
                 (float const *)g->lamb_dev, mode == 4 ? 1 : mode, g->scalar))
        return 1;
    }
  }
  return 0;
}

int agdg_shrink(AgdDev const *g, int64_t n_complex, float lamb) {
  int64_t blocks = (n_complex + NTHR - 1) / NTHR;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  return launch("shrink", k_shrink, dim3((unsigned)blocks), dim3(NTHR), 0,
                (float2 *)g->X, n_complex, lamb);
}

int agdg_wavelet(AgdDev const *g, int adjoint) {
  if (!g->use_wavelets) return 0;
  int const first = g->wtrace0[g->w0];
  int const n = g->wtrace0[g->w1] - first;
  if (n <= 0) return 0;
#define WAV_CASE(MH, NT, PP, ...)                                                           \
  case MH:                                                                                  \
    return launch(adjoint ? "wavelet_adj" : "wavelet_fwd", k_wavelet_t<NT, PP, __VA_ARGS__>, \
                  dim3(n), dim3(NT), 0, g->win, g->trace_win, first, g->plans,               \
                  (float2 const *)g->roots, (float2 const *)g->wspec, g->wsrc_off, adjoint, \
                  g->went, (float const *)g->r, g->T, g->T2);
  // one case per half length of AGD_WAVELET_HALF_LENGTHS
  switch (g->wavelet_mh) {
    WAV_CASE(512, 256, true, 8, 8, 8)
    WAV_CASE(768, 256, true, 8, 12, 8)
    WAV_CASE(1024, 256, true, 16, 8, 8)
    WAV_CASE(1536, 512, true, 8, 8, 3, 8)
    WAV_CASE(2048, 512, true, 8, 8, 4, 8)
    WAV_CASE(3072, 512, true, 8, 8, 6, 8)
    WAV_CASE(4096, 512, true, 8, 8, 8, 8)
    WAV_CASE(6144, 768, false, 8, 8, 12, 8)
    default:
      break;
  }
#undef WAV_CASE
  return launch(adjoint ? "wavelet_adj" : "wavelet_fwd", k_wavelet, dim3(n),
                dim3(NTHR), g->lds_wavelet, g->win, g->trace_win, first,
                g->plans, (float2 const *)g->roots, (float2 const *)g->wspec,
                g->wsrc_off, adjoint, g->went, (float const *)g->r, g->T, g->T2);
}

int agdg_norm2(AgdDev const *g, float const *v, int64_t n, double *out) {
  int blocks = g->n_partials;
  std::vector<double> h(blocks);
  if (launch("norm2", k_norm2, dim3(blocks), dim3(NTHR), 0, v, n, g->partials)) return 1;
  if (hip_check(hipMemcpyAsync(h.data(), g->partials, blocks * sizeof(double),
                               hipMemcpyDeviceToHost, g_stream),
                "memcpy") ||
      hip_check(hipStreamSynchronize(g_stream), "sync"))
    return 1;
  double s = 0.0;
  for (int i = 0; i < blocks; ++i) s += h[i];
  *out = s;
  return 0;
}

int agdg_output(float const *T, int n_out, int64_t const *o_off, int const *o_nt,
                int const *o_ptr, int64_t const *o_src, float *out) {
  return launch("output", k_output, dim3(n_out), dim3(NTHR), 0, T, o_off, o_nt, o_ptr,
                o_src, out);
}

int agdg_gather_resid(AgdGather const *gt, float const *src, float const *d, float *rec) {
  return launch("gather", k_blend<3>, dim3(gt->n_tiles), dim3(NTHR), 0, gt->rec_len, gt->tile,
                gt->tile_ptr, gt->tile_ent, gt->ent, src, (float const *)nullptr, gt->taper,
                gt->taper_len, gt->mute, d, rec);
}
int agdg_gather(AgdGather const *gt, float const *src, int mode, float *rec) {
  if (mode < 0 || mode > 2) return 1;
  return launch("gather", mode == 0 ? k_blend<0> : mode == 1 ? k_blend<1> : k_blend<2>,
                dim3(gt->n_tiles), dim3(NTHR), 0, gt->rec_len,
                gt->tile, gt->tile_ptr, gt->tile_ent, gt->ent, src,
                (float const *)nullptr, gt->taper, gt->taper_len, gt->mute,
                (float const *)nullptr, rec);
}

int agdg_next_lamb(AgdDev const *g) {
  return launch("next_lamb", k_next_lamb, dim3(1), dim3(64), 0, (float const *)g->lamb_sched,
                g->lamb_idx, g->lamb_dev);
}

int agdg_graph_begin(void) {
  return hip_check(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal),
                   "hipStreamBeginCapture");
}

int agdg_graph_end(void **exec) {
  hipGraph_t graph = nullptr;
  hipGraphExec_t ge = nullptr;
  *exec = nullptr;
  if (hip_check(hipStreamEndCapture(g_stream, &graph), "hipStreamEndCapture")) return 1;
  hipError_t const e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (hip_check(e, "hipGraphInstantiate")) return 1;
  *exec = (void *)ge;
  return 0;
}

int agdg_graph_launch(void *exec) {
  return hip_check(hipGraphLaunch((hipGraphExec_t)exec, g_stream), "hipGraphLaunch");
}

void agdg_graph_destroy(void *exec) {
  if (exec != nullptr) (void)hipGraphExecDestroy((hipGraphExec_t)exec);
}

int agdg_timing_enabled(void) { return g_timing; }

int agdg_extract(AgdEntry const *ent, int n_ent, int const *trace_len, float const *rec,
                 float *out) {
  return launch("extract", k_extract, dim3(n_ent), dim3(NTHR), 0, ent, trace_len, rec,
                out);
}

int agdg_timing(int enable) {
  if (!enable) {
    (void)hipStreamSynchronize(g_stream);
    flush_timing();
  }
  g_timing = enable;
  if (enable) g_times.clear();
  return 0;
}

int agdg_timing_get(char const *name, double *ms, long *count) {
  (void)hipStreamSynchronize(g_stream);
  flush_timing();
  auto it = g_times.find(name);
  if (it == g_times.end()) {
    *ms = 0.0;
    *count = 0;
    return 1;
  }
  *ms = it->second.ms;
  *count = it->second.count;
  return 0;
}

}  // extern "C"
