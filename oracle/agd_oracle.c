/*
 * ORACLE — CPU restatement of AGDeblend's deblending path.
 *
 * Test infrastructure only: it is loaded by tests/, by the smoke check and by
 * bench.py's cpu_baseline leg, as the checker and as the timed CPU baseline.
 * It is never linked into, called by or shipped with libagdeblend.so.
 *
 * Every function below restates the reference algorithm it cites (paths are
 * relative to the reference tree). Its FFTs come from oracle_fft.c, which
 * restates FFTW's conventions. Compiled twice: float (-DORACLE_FLOAT, the
 * reference's default AGD_TYPE) and double (the FP64 oracle).
 *
 * Parity pins (tests/test_oracle_*.py): the reference's own known-answer
 * tests (blend_test.cpp, fk_test.cpp, wavelet_test.cpp, model_test.cpp,
 * test.cpp), its dot-product tests, numpy.fft for every transform, and brute
 * force on small inputs.
 *
 * With -fopenmp the per-window transforms run in parallel; the blend
 * accumulation stays serial in window order, i.e. the order of the serial
 * reference (model.c:1086-1101), so results do not depend on thread count.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_DOUBLE
typedef double real;
#define RCOS cos
#define RSIN sin
#define RATAN2 atan2
#define REXP exp
#define RLOG log
#define RPOW pow
#define RSQRT sqrt
#define RCEIL ceil
#define OSYM(name) oracle64_##name
#else
typedef float real;
#define RCOS cosf
#define RSIN sinf
#define RATAN2 atan2f
#define REXP expf
#define RLOG logf
#define RPOW powf
#define RSQRT sqrtf
#define RCEIL ceilf
#define OSYM(name) oracle32_##name
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
/* The reference's own fallback value of pi (agdeblend.c:43-45) is unused
 * when math.h defines M_PI, as glibc does. */

#include "oracle_fft.c"

#define O_LIVE 0
#define O_BAD 1
#define O_MISSING 2
#define O_MAXC 4 /* AGD_MAX_N_COORDS */
#define O_PAD ((real)0.25)
#define O_TARGET_LAMB ((real)0.001)

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Intervals and channels (blend.c:19-342)                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
  long start, stop;
} oiv;

typedef struct {
  oiv *iv;
  int n;
  int cap;
} oivlist;

static int iv_overlap(oiv a, oiv b) { return a.start < b.stop && a.stop > b.start; }
static int iv_adjacent(oiv a, oiv b) { return a.start == b.stop || a.stop == b.start; }

static int ivl_find(oiv x, oivlist const *l, int from) {
  int i;
  for (i = from; i < l->n; i++) {
    if (iv_overlap(x, l->iv[i]) || iv_adjacent(x, l->iv[i])) return i;
  }
  return -1;
}

static void ivl_push(oivlist *l, oiv x) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 4;
    l->iv = (oiv *)realloc(l->iv, (size_t)l->cap * sizeof(oiv));
  }
  l->iv[l->n++] = x;
}

/* blend.c:113-151: extend the first overlapping-or-adjacent interval, then
 * absorb (and delete) later intervals that overlap or touch the NEW
 * interval; otherwise append. */
static void ivl_add(oivlist *l, oiv x) {
  int const first = ivl_find(x, l, 0);
  if (first < 0) {
    ivl_push(l, x);
    return;
  }
  {
    oiv *f = &l->iv[first];
    int from = first + 1;
    if (x.start < f->start) f->start = x.start;
    if (x.stop > f->stop) f->stop = x.stop;
    for (;;) {
      int const k = ivl_find(x, l, from);
      int j;
      if (k < 0) break;
      f = &l->iv[first];
      if (l->iv[k].start < f->start) f->start = l->iv[k].start;
      if (l->iv[k].stop > f->stop) f->stop = l->iv[k].stop;
      for (j = k; j < l->n - 1; j++) l->iv[j] = l->iv[j + 1];
      l->n--;
      from = k + 1;
    }
  }
}

typedef struct {
  size_t interval_start;
  int trace_start;
  int channel_idx;
  int interval_idx;
  int n_times;
} obc; /* BlendCoords, agdeblend.c:173-179 */

typedef struct {
  int n_channels;
  int *uniq;        /* unique channel values, first appearance order */
  oivlist *chan;    /* per channel intervals */
  obc *mute;        /* mute coords */
  int n_mute;
  size_t *rec_off;  /* flat offset of each (channel, interval) record */
  int *chan_iv0;    /* index of the first (channel,interval) of channel c */
  size_t rec_len;   /* total record samples */
} oblendcfg;

/* blend.c:176-223 */
static void blend_unique_channels(int np, int const *ntr, int const *const *ch,
                                  int const *const *tt, oblendcfg *cfg) {
  int p, t, u;
  int cap = 0;
  cfg->n_channels = 0;
  cfg->uniq = NULL;
  for (p = 0; p < np; p++) {
    for (t = 0; t < ntr[p]; t++) {
      if (tt[p][t] == O_MISSING) continue;
      for (u = 0; u < cfg->n_channels; u++)
        if (cfg->uniq[u] == ch[p][t]) break;
      if (u == cfg->n_channels) {
        if (cfg->n_channels == cap) {
          cap = cap ? 2 * cap : 8;
          cfg->uniq = (int *)realloc(cfg->uniq, (size_t)cap * sizeof(int));
        }
        cfg->uniq[cfg->n_channels++] = ch[p][t];
      }
    }
  }
}

/* blend.c:231-246 */
static int blend_channel_idx(int channel, int type, oblendcfg const *cfg) {
  int u;
  if (type == O_MISSING) return -1;
  for (u = 0; u < cfg->n_channels; u++)
    if (cfg->uniq[u] == channel) return u;
  return -1;
}

static void blend_layout(oblendcfg *cfg) {
  int c, i, k = 0;
  size_t off = 0;
  int total = 0;
  for (c = 0; c < cfg->n_channels; c++) total += cfg->chan[c].n;
  cfg->rec_off = (size_t *)malloc((size_t)(total + 1) * sizeof(size_t));
  cfg->chan_iv0 = (int *)malloc((size_t)(cfg->n_channels + 1) * sizeof(int));
  for (c = 0; c < cfg->n_channels; c++) {
    cfg->chan_iv0[c] = k;
    for (i = 0; i < cfg->chan[c].n; i++) {
      cfg->rec_off[k++] = off;
      off += (size_t)(cfg->chan[c].iv[i].stop - cfg->chan[c].iv[i].start);
    }
  }
  cfg->chan_iv0[cfg->n_channels] = k;
  cfg->rec_off[k] = off;
  cfg->rec_len = off;
}

/* blend.c:305-342, 585-691 */
static void blend_config(int np, int const *ntr, int const *ntimes,
                         long const *const *shot, int const *const *ch,
                         int const *const *tt, oblendcfg *cfg) {
  oivlist *mute_iv;
  int p, t, c, m, i;
  memset(cfg, 0, sizeof(*cfg));
  blend_unique_channels(np, ntr, ch, tt, cfg);
  cfg->chan = (oivlist *)calloc((size_t)(cfg->n_channels + 1), sizeof(oivlist));
  mute_iv = (oivlist *)calloc((size_t)(cfg->n_channels + 1), sizeof(oivlist));
  for (p = 0; p < np; p++) {
    for (t = 0; t < ntr[p]; t++) {
      int const cidx = blend_channel_idx(ch[p][t], tt[p][t], cfg);
      oiv x;
      x.start = shot[p][t];
      x.stop = shot[p][t] + ntimes[p];
      if (cidx < 0) continue;
      ivl_add(&cfg->chan[cidx], x);
      if (tt[p][t] == O_BAD) ivl_add(&mute_iv[cidx], x);
    }
  }
  /* mute coords: every (mute interval, channel interval) overlap */
  cfg->n_mute = 0;
  for (c = 0; c < cfg->n_channels; c++) cfg->n_mute += mute_iv[c].n;
  cfg->mute = (obc *)malloc((size_t)(cfg->n_mute + 1) * sizeof(obc));
  m = 0;
  for (c = 0; c < cfg->n_channels; c++) {
    int mi;
    for (mi = 0; mi < mute_iv[c].n; mi++) {
      oiv const mv = mute_iv[c].iv[mi];
      for (i = 0; i < cfg->chan[c].n; i++) {
        oiv const cv = cfg->chan[c].iv[i];
        if (iv_overlap(mv, cv)) {
          long const a = mv.start < cv.start ? cv.start : mv.start;
          long const b = mv.stop > cv.stop ? cv.stop : mv.stop;
          cfg->mute[m].channel_idx = c;
          cfg->mute[m].interval_idx = i;
          cfg->mute[m].interval_start = (size_t)(a - cv.start);
          cfg->mute[m].trace_start = 0;
          cfg->mute[m].n_times = (int)(b - a);
          m++;
        }
      }
    }
  }
  for (c = 0; c < cfg->n_channels; c++) free(mute_iv[c].iv);
  free(mute_iv);
  blend_layout(cfg);
}

static void blend_config_free(oblendcfg *cfg) {
  int c;
  for (c = 0; c < cfg->n_channels; c++) free(cfg->chan[c].iv);
  free(cfg->chan);
  free(cfg->uniq);
  free(cfg->mute);
  free(cfg->rec_off);
  free(cfg->chan_iv0);
  memset(cfg, 0, sizeof(*cfg));
}

/* blend.c:693-760 */
static void blend_coords(int ntr, int ntimes, long const *shot, int const *ch,
                         int const *tt, oblendcfg const *cfg, obc *out) {
  int t;
  for (t = 0; t < ntr; t++) {
    int const cidx = blend_channel_idx(ch[t], tt[t], cfg);
    oiv x;
    int i;
    x.start = shot[t];
    x.stop = shot[t] + ntimes;
    out[t].channel_idx = -1;
    out[t].interval_idx = 0;
    out[t].interval_start = 0;
    out[t].trace_start = 0;
    out[t].n_times = 0;
    if (tt[t] == O_MISSING || cidx < 0) continue;
    for (i = 0; i < cfg->chan[cidx].n; i++) {
      oiv const cv = cfg->chan[cidx].iv[i];
      if (iv_overlap(x, cv)) {
        long const a = x.start < cv.start ? cv.start : x.start;
        long const b = x.stop > cv.stop ? cv.stop : x.stop;
        out[t].channel_idx = cidx;
        out[t].interval_idx = i;
        out[t].interval_start = (size_t)(a - cv.start);
        out[t].trace_start = (int)(a - x.start);
        out[t].n_times = (int)(b - a);
        break;
      }
    }
  }
}

static real *rec_at(oblendcfg const *cfg, real *rec, obc const *b) {
  return rec + cfg->rec_off[cfg->chan_iv0[b->channel_idx] + b->interval_idx] +
         b->interval_start;
}

/* blend.c:936-960 */
static void blend_sum(real const *data, int ntr, int trace_len, obc const *bc,
                      oblendcfg const *cfg, real *rec) {
  int t, s;
  for (t = 0; t < ntr; t++) {
    if (bc[t].channel_idx < 0) continue;
    {
      real *dst = rec_at(cfg, rec, &bc[t]);
      real const *src = data + (size_t)t * trace_len + bc[t].trace_start;
      for (s = 0; s < bc[t].n_times; s++) dst[s] += src[s];
    }
  }
}

/* blend.c:1007-1027 */
static void blend_overwrite(real const *data, int ntr, int trace_len,
                            obc const *bc, oblendcfg const *cfg, real *rec) {
  int t;
  for (t = 0; t < ntr; t++) {
    if (bc[t].channel_idx < 0) continue;
    memcpy(rec_at(cfg, rec, &bc[t]),
           data + (size_t)t * trace_len + bc[t].trace_start,
           (size_t)bc[t].n_times * sizeof(real));
  }
}

/* blend.c:1155-1260 */
static void blend_mean(int np, real const *const *data, int const *ntr,
                       int const *trace_len, obc *const *bc,
                       oblendcfg const *cfg, int taper_length, real *rec) {
  real *count = (real *)calloc(cfg->rec_len + 1, sizeof(real));
  real *taper = NULL;
  size_t s;
  int p, t, i;
  if (taper_length > 1) {
    taper = (real *)malloc((size_t)taper_length * sizeof(real));
    for (i = 0; i < taper_length; i++) {
      taper[i] = RCOS((real)0.5 * (real)M_PI * (real)(taper_length - i) /
                      (real)(taper_length + 1));
    }
  }
  for (p = 0; p < np; p++) {
    for (t = 0; t < ntr[p]; t++) {
      obc const *b = &bc[p][t];
      if (b->channel_idx < 0) continue;
      {
        size_t const off =
            cfg->rec_off[cfg->chan_iv0[b->channel_idx] + b->interval_idx] +
            b->interval_start;
        real const *src = data[p] + (size_t)t * trace_len[p] + b->trace_start;
        for (i = 0; i < b->n_times; i++) {
          int const to_end = b->n_times - 1 - i;
          int const edge = i < to_end ? i : to_end;
          real const w =
              (taper_length > 1 && edge < taper_length) ? taper[edge] : (real)1;
          rec[off + i] += src[i] * w;
          count[off + i] += w;
        }
      }
    }
  }
  for (s = 0; s < cfg->rec_len; s++) {
    if (count[s] > (real)0) rec[s] /= count[s];
  }
  free(count);
  free(taper);
}

/* blend.c:862-875 */
static void blend_mute(oblendcfg const *cfg, real *rec) {
  int m;
  for (m = 0; m < cfg->n_mute; m++) {
    memset(rec_at(cfg, rec, &cfg->mute[m]), 0,
           (size_t)cfg->mute[m].n_times * sizeof(real));
  }
}

/* blend.c:1262-1292 */
static void blend_adjoint(real const *rec, int ntr, int trace_len,
                          obc const *bc, oblendcfg const *cfg, real *data) {
  int t;
  for (t = 0; t < ntr; t++) {
    real *dst = data + (size_t)t * trace_len;
    if (bc[t].channel_idx < 0) {
      memset(dst, 0, (size_t)trace_len * sizeof(real));
    } else {
      obc const *b = &bc[t];
      memset(dst, 0, (size_t)b->trace_start * sizeof(real));
      memcpy(dst + b->trace_start,
             rec + cfg->rec_off[cfg->chan_iv0[b->channel_idx] + b->interval_idx] +
                 b->interval_start,
             (size_t)b->n_times * sizeof(real));
      memset(dst + b->trace_start + b->n_times, 0,
             (size_t)(trace_len - (b->trace_start + b->n_times)) * sizeof(real));
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FK operator (fk.c)                                                        */
/* ------------------------------------------------------------------------ */

/* fk.c:28-73 */
static void taper_window(int len, int left, int right, real *w) {
  int const tl = len / 2;
  int i;
  for (i = 0; i < len; i++) w[i] = (real)1;
  if (left) {
    for (i = 0; i < tl; i++)
      w[i] = (real)0.5 *
             ((real)1 - RCOS((real)2 * (real)M_PI * (real)i / (real)(2 * tl)));
  }
  if (right) {
    for (i = 0; i < tl; i++)
      w[len - tl + i] =
          (real)0.5 * ((real)1 - RCOS((real)2 * (real)M_PI * (real)(tl + i) /
                                      (real)(2 * tl)));
  }
}

typedef struct {
  int patch;
  int ndim;           /* total dims including time */
  int coords[O_MAXC];
  int shape[O_MAXC];  /* space shape */
  int use_taper[2 * O_MAXC];
  int ntr;            /* traces in window */
  /* FK config, fk.c:75-169 */
  int W[O_MAXC + 1];  /* window shape incl. time */
  int P[O_MAXC + 1];  /* padded shape */
  int Ft;             /* P_t/2 + 1 */
  int ntw;
  int nfk, nconv;
  int npad_tr;        /* prod of padded space dims */
  int npad;           /* prod of all padded dims */
  int nfk_el;         /* npad_tr * Ft */
  real scaler;
  real *taper[O_MAXC + 1];
  real *trace_scale;  /* scaler * prod space tapers, fk.c:334-347 */
  /* blend */
  obc *bc;
  int *types;
  /* wavelet: per trace source config or -1 (NULL: no wavelets) */
  int *wsrc;
  size_t model_off;   /* complex elements */
} owin;

typedef struct {
  int orig;
  int n;
  ocpx *spec; /* n/2+1 */
} owsrc;

typedef struct {
  int n_patches;
  int *nd;            /* per patch: n_dims of its volume */
  int *ntr;           /* per patch traces */
  int *nt;            /* per patch data time samples */
  int *wt;            /* per patch time window */
  int *nfk, *nconv;
  int n_win;
  owin *win;
  size_t model_len;   /* complex elements */
  oblendcfg cfg;
  int use_wavelets;
  owsrc *wsrc;
  int n_wsrc;
  int max_ntr_conv;   /* max over windows of ntr*nconv */
  size_t max_work;
} omodel;

static void fk_setup(owin *w, int nfk, int nconv, int wt) {
  int const nd = w->ndim;
  int d;
  for (d = 0; d < nd - 1; d++) w->W[d] = w->shape[d];
  w->W[nd - 1] = wt;
  w->nfk = nfk;
  w->nconv = nconv;
  w->ntw = nfk / (wt / 2) - 1; /* fk.c:24-26 */
  w->npad_tr = 1;
  w->npad = 1;
  for (d = 0; d < nd; d++) {
    w->P[d] = (int)RCEIL((real)w->W[d] * ((real)1 + O_PAD));
    w->npad *= w->P[d];
    if (d < nd - 1) w->npad_tr *= w->P[d];
  }
  w->Ft = w->P[nd - 1] / 2 + 1;
  w->nfk_el = w->npad_tr * w->Ft;
  w->scaler = (real)1 / RSQRT((real)w->npad);
  for (d = 0; d < nd; d++) {
    int const left = d < nd - 1 ? w->use_taper[2 * d] : 1;
    int const right = d < nd - 1 ? w->use_taper[2 * d + 1] : 1;
    w->taper[d] = (real *)malloc((size_t)w->W[d] * sizeof(real));
    taper_window(w->W[d], left, right, w->taper[d]);
  }
  /* per-trace scale, same multiplication order as fk.c:343-347 */
  w->trace_scale = (real *)malloc((size_t)w->ntr * sizeof(real));
  {
    int c[O_MAXC] = {0};
    int t;
    for (t = 0; t < w->ntr; t++) {
      real s = w->scaler;
      for (d = nd - 2; d >= 0; d--) s *= w->taper[d][c[d]];
      w->trace_scale[t] = s;
      for (d = nd - 2; d >= 0; d--) {
        if (++c[d] == w->W[d]) c[d] = 0;
        else break;
      }
    }
  }
}

/* padded trace index of window trace t (row-major, last space dim fastest) */
static int fk_padded_trace(owin const *w, int t) {
  int d, idx = 0, stride = 1;
  for (d = w->ndim - 2; d >= 0; d--) {
    int const cd = t % w->W[d];
    t /= w->W[d];
    idx += cd * stride;
    stride *= w->P[d];
  }
  return idx;
}

/* fk.c:307-373. x: [ntw][npad_tr][Ft] complex; out: [ntr][nconv] real.
 * work: ntw*nfk_el complex + ntw*npad real + FFT scratch. */
static void fk_forward(owin const *w, ocpx const *x, real *out, ocpx *cwork,
                       real *rwork, ocpx *fwork) {
  int const nd = w->ndim;
  int const wt = w->W[nd - 1];
  int const half = wt / 2;
  real const *tt = w->taper[nd - 1];
  int tw, t, s;
  memcpy(cwork, x, (size_t)w->ntw * w->nfk_el * sizeof(ocpx));
  for (tw = 0; tw < w->ntw; tw++) {
    offt_c2r_nd(nd, w->P, cwork + (size_t)tw * w->nfk_el,
                rwork + (size_t)tw * w->npad, fwork);
  }
  memset(out, 0, (size_t)w->ntr * w->nconv * sizeof(real));
  for (t = 0; t < w->ntr; t++) {
    real const sc = w->trace_scale[t];
    int const pt = fk_padded_trace(w, t);
    real *o = out + (size_t)t * w->nconv;
    for (tw = 0; tw < w->ntw; tw++) {
      real const *y = rwork + (size_t)tw * w->npad + (size_t)pt * w->P[nd - 1];
      real *ow = o + tw * half;
      for (s = 0; s < wt; s++) ow[s] += y[s] * sc * tt[s];
    }
  }
}

/* fk.c:375-452. in: [ntr][nconv] real; g: [ntw][npad_tr][Ft] complex. */
static void fk_adjoint(owin const *w, real const *in, ocpx *g, real *rwork,
                       ocpx *fwork) {
  int const nd = w->ndim;
  int const wt = w->W[nd - 1];
  int const half = wt / 2;
  int const Pt = w->P[nd - 1];
  real const *tt = w->taper[nd - 1];
  int tw, t, s, k;
  memset(rwork, 0, (size_t)w->ntw * w->npad * sizeof(real));
  for (t = 0; t < w->ntr; t++) {
    real const sc = w->trace_scale[t];
    int const pt = fk_padded_trace(w, t);
    real const *i0 = in + (size_t)t * w->nconv;
    for (tw = 0; tw < w->ntw; tw++) {
      real *z = rwork + (size_t)tw * w->npad + (size_t)pt * Pt;
      real const *iw = i0 + tw * half;
      for (s = 0; s < wt; s++) z[s] = iw[s] * sc * tt[s];
    }
  }
  for (tw = 0; tw < w->ntw; tw++) {
    offt_r2c_nd(nd, w->P, rwork + (size_t)tw * w->npad,
                g + (size_t)tw * w->nfk_el, fwork);
  }
  for (tw = 0; tw < w->ntw; tw++) {
    for (t = 0; t < w->npad_tr; t++) {
      ocpx *row = g + (size_t)tw * w->nfk_el + (size_t)t * w->Ft;
      for (k = 1; k < 1 + (Pt - 1) / 2; k++) {
        row[k].re *= (real)2;
        row[k].im *= (real)2;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Wavelet (wavelet.c)                                                       */
/* ------------------------------------------------------------------------ */

/* wavelet.c:188-237: phase-only spectrum scaled by 1/n */
static void wavelet_spectrum(real const *wav, int len, int n, ocpx *spec,
                             ocpx *fwork) {
  real *tmp = (real *)calloc((size_t)n, sizeof(real));
  real const scale = (real)1 / (real)n;
  int k;
  memcpy(tmp, wav, (size_t)len * sizeof(real));
  offt_r2c(n, tmp, spec, fwork);
  for (k = 0; k <= n / 2; k++) {
    real const a = RATAN2(spec[k].im, spec[k].re);
    spec[k].re = RCOS(a) * scale;
    spec[k].im = RSIN(a) * scale;
  }
  free(tmp);
}

/* wavelet.c:239-301; buf: [ntr][n] real, in place */
static void wavelet_op(omodel const *m, owin const *w, real *buf, real dir,
                       ocpx *spec_work, ocpx *fwork) {
  int const n = w->nconv;
  int const nf = n / 2 + 1;
  int t, k;
  int any = 0;
  if (!m->use_wavelets || w->wsrc == NULL) return;
  for (t = 0; t < w->ntr; t++)
    if (w->wsrc[t] >= 0) any = 1;
  if (!any) return;
  for (t = 0; t < w->ntr; t++) {
    real *tr = buf + (size_t)t * n;
    offt_r2c(n, tr, spec_work, fwork);
    if (w->wsrc[t] >= 0) {
      ocpx const *wf = m->wsrc[w->wsrc[t]].spec;
      for (k = 0; k < nf; k++) {
        real const re = spec_work[k].re;
        real const im = spec_work[k].im;
        spec_work[k].re = re * wf[k].re - dir * im * wf[k].im;
        spec_work[k].im = dir * re * wf[k].im + im * wf[k].re;
      }
    }
    offt_c2r(n, spec_work, tr, fwork);
  }
}

/* ------------------------------------------------------------------------ */
/* Geometry (model.c, utils.c)                                               */
/* ------------------------------------------------------------------------ */

/* model.c:20-50 */
static void taper_from_neighbour(int vol, int nd, int const *pc, int ovol,
                                 int const *oc, int *ut) {
  int d;
  if (ovol != vol) return;
  for (d = 0; d < nd - 1; d++) {
    if (oc[d] != pc[d] - 1 && oc[d] != pc[d] && oc[d] != pc[d] + 1) return;
  }
  for (d = 0; d < nd - 1; d++) {
    if (oc[d] == pc[d] - 1) ut[2 * d] = 1;
    else if (oc[d] == pc[d] + 1) ut[2 * d + 1] = 1;
  }
}

static int n_win_dim(int L, int W) { return W == L ? 1 : L / (W / 2) - 1; }

/* model.c:691-745 (windows of every patch; live-trace deletion elsewhere) */
static int space_windows(int np, int const *volumes, int const *n_dims,
                         int const *const *wshape, int const *const *coords,
                         int const *const *shape, owin **out, int *n_out) {
  int p, q, d;
  int total = 0;
  int **ut = (int **)calloc((size_t)np, sizeof(int *));
  owin *win;
  int base = 0;
  for (p = 0; p < np; p++) {
    int const nd = n_dims[volumes[p]];
    int nw = 1;
    for (d = 0; d < nd - 1; d++) nw *= n_win_dim(shape[p][d], wshape[volumes[p]][d]);
    total += nw;
    ut[p] = (int *)calloc((size_t)(2 * (nd - 1) + 1), sizeof(int));
    for (q = 0; q < np; q++) {
      if (q == p) continue;
      taper_from_neighbour(volumes[p], nd, coords[p], volumes[q], coords[q],
                           ut[p]);
    }
  }
  /* model.c:551-580 */
  for (p = 0; p < np; p++) {
    int const nd = n_dims[volumes[p]];
    for (d = 0; d < nd - 1; d++) {
      if ((ut[p][2 * d] || ut[p][2 * d + 1]) && wshape[volumes[p]][d] % 2 == 1) {
        fprintf(stderr,
                "ERROR: window shapes may only be odd if there is no tapering. "
                "Patch %d touches another patch in dimension %d and so will be "
                "tapered, but has window shape %d.\n",
                p, d, wshape[volumes[p]][d]);
        for (q = 0; q < np; q++) free(ut[q]);
        free(ut);
        return 1;
      }
    }
  }
  win = (owin *)calloc((size_t)(total + 1), sizeof(owin));
  /* model.c:582-672 */
  for (p = 0; p < np; p++) {
    int const nd = n_dims[volumes[p]];
    int const *ws = wshape[volumes[p]];
    int nwd[O_MAXC];
    int nw = 1;
    int i;
    for (d = 0; d < nd - 1; d++) {
      nwd[d] = n_win_dim(shape[p][d], ws[d]);
      nw *= nwd[d];
    }
    for (i = 0; i < nw; i++) {
      win[base + i].patch = p;
      win[base + i].ndim = nd;
    }
    for (d = 0; d < nd - 1; d++) {
      int const tl = ws[d] / 2;
      int const n = nwd[d];
      int const bl = (shape[p][d] - tl) / n + tl;
      int const nl = shape[p][d] - tl - (bl - tl) * n;
      int start = 0;
      int k;
      for (k = 0; k < n; k++) {
        int const len = bl + (k < nl);
        /* every window whose coordinate in dim d is k */
        for (i = 0; i < nw; i++) {
          int rem = i, dd, cd = 0;
          for (dd = nd - 2; dd >= 0; dd--) {
            int const c = rem % nwd[dd];
            rem /= nwd[dd];
            if (dd == d) cd = c;
          }
          if (cd != k) continue;
          win[base + i].coords[d] = start;
          win[base + i].shape[d] = len;
          win[base + i].use_taper[2 * d] = (k == 0) ? ut[p][2 * d] : 1;
          win[base + i].use_taper[2 * d + 1] = (k == n - 1) ? ut[p][2 * d + 1] : 1;
        }
        start += len - tl;
      }
    }
    for (i = 0; i < nw; i++) {
      win[base + i].ntr = 1;
      for (d = 0; d < nd - 1; d++) win[base + i].ntr *= win[base + i].shape[d];
    }
    base += nw;
  }
  for (q = 0; q < np; q++) free(ut[q]);
  free(ut);
  *out = win;
  *n_out = total;
  return 0;
}

/* model.c:748-805: patch trace index of window trace t */
static int window_patch_trace(owin const *w, int const *pshape, int t) {
  int d, idx = 0, stride = 1;
  for (d = w->ndim - 2; d >= 0; d--) {
    int const c = t % w->shape[d];
    t /= w->shape[d];
    idx += (w->coords[d] + c) * stride;
    stride *= pshape[d];
  }
  return idx;
}

/* utils.c:72-123 */
static int volume_wavelet_lengths(int np, int const *volumes, int const *ntr,
                                  int const *const *tt, int const *wl,
                                  int const *const *widx, int **out) {
  int nv = 0;
  int p, v, t;
  for (p = 0; p < np; p++)
    if (volumes[p] > nv) nv = volumes[p];
  nv += 1;
  *out = (int *)calloc((size_t)nv, sizeof(int));
  if (widx == NULL) {
    for (v = 0; v < nv; v++) (*out)[v] = 1;
    return 0;
  }
  for (p = 0; p < np; p++) {
    v = volumes[p];
    if ((*out)[v] != 0) continue;
    for (t = 0; t < ntr[p]; t++) {
      if (tt[p][t] == O_LIVE) {
        (*out)[v] = wl[widx[p][t]];
        if ((*out)[v] <= 0) {
          fprintf(stderr, "ERROR: Wavelet lengths must be positive\n");
          return 1;
        }
        break;
      }
    }
  }
  for (v = 0; v < nv; v++)
    if ((*out)[v] == 0) (*out)[v] = 1;
  return 0;
}

static void model_free(omodel *m) {
  int w, d;
  if (m == NULL) return;
  for (w = 0; w < m->n_win; w++) {
    for (d = 0; d < O_MAXC + 1; d++) free(m->win[w].taper[d]);
    free(m->win[w].trace_scale);
    free(m->win[w].bc);
    free(m->win[w].types);
    free(m->win[w].wsrc);
  }
  free(m->win);
  for (w = 0; w < m->n_wsrc; w++) free(m->wsrc[w].spec);
  free(m->wsrc);
  free(m->nd);
  free(m->ntr);
  free(m->nt);
  free(m->wt);
  free(m->nfk);
  free(m->nconv);
  blend_config_free(&m->cfg);
  free(m);
}

/* model.c:932-1065 */
static omodel *model_new(int np, int const *volumes, int const *n_dims,
                         int const *const *wshape, int const *const *coords,
                         int const *const *shape, long const *const *shot,
                         int const *const *ch, int const *const *tt,
                         int const *wl, int const *const *widx,
                         real const *const *wav) {
  omodel *m = (omodel *)calloc(1, sizeof(omodel));
  int *vwl = NULL;
  long **shot_taper;
  int p, t, w, d, n_all;
  owin *all;
  m->n_patches = np;
  m->nd = (int *)malloc((size_t)np * sizeof(int));
  m->ntr = (int *)malloc((size_t)np * sizeof(int));
  m->nt = (int *)malloc((size_t)np * sizeof(int));
  m->wt = (int *)malloc((size_t)np * sizeof(int));
  m->nfk = (int *)malloc((size_t)np * sizeof(int));
  m->nconv = (int *)malloc((size_t)np * sizeof(int));
  for (p = 0; p < np; p++) {
    int const nd = n_dims[volumes[p]];
    m->nd[p] = nd;
    m->ntr[p] = 1;
    for (d = 0; d < nd - 1; d++) m->ntr[p] *= shape[p][d];
    m->nt[p] = shape[p][nd - 1];
    m->wt[p] = wshape[volumes[p]][nd - 1];
  }
  if (volume_wavelet_lengths(np, volumes, m->ntr, tt, wl, widx, &vwl)) {
    free(vwl);
    model_free(m);
    return NULL;
  }
  /* model.c:326-357 */
  for (p = 0; p < np; p++) {
    int const half = m->wt[p] / 2;
    int const lw = vwl[volumes[p]];
    int const pre = m->nt[p] - lw + 1;
    int const ntw = (pre + half - 1) / half + 1;
    m->nfk[p] = (ntw + 1) * half;
    m->nconv[p] = m->nfk[p] + lw - 1;
  }
  free(vwl);
  /* model.c:424-446 */
  shot_taper = (long **)malloc((size_t)np * sizeof(long *));
  for (p = 0; p < np; p++) {
    shot_taper[p] = (long *)malloc((size_t)(m->ntr[p] + 1) * sizeof(long));
    for (t = 0; t < m->ntr[p]; t++) shot_taper[p][t] = shot[p][t] - m->wt[p] / 2;
  }
  blend_config(np, m->ntr, m->nconv, (long const *const *)shot_taper, ch, tt,
               &m->cfg);
  if (space_windows(np, volumes, n_dims, wshape, coords, shape, &all, &n_all)) {
    for (p = 0; p < np; p++) free(shot_taper[p]);
    free(shot_taper);
    model_free(m);
    return NULL;
  }
  m->use_wavelets = widx != NULL;
  m->win = (owin *)calloc((size_t)(n_all + 1), sizeof(owin));
  m->n_win = 0;
  for (w = 0; w < n_all; w++) {
    owin wi = all[w];
    int const pp = wi.patch;
    long *wshot = (long *)malloc((size_t)wi.ntr * sizeof(long));
    int *wch = (int *)malloc((size_t)wi.ntr * sizeof(int));
    int *wtt = (int *)malloc((size_t)wi.ntr * sizeof(int));
    int *wwi = widx ? (int *)malloc((size_t)wi.ntr * sizeof(int)) : NULL;
    int live = 0;
    for (t = 0; t < wi.ntr; t++) {
      int const pt = window_patch_trace(&wi, shape[pp], t);
      wshot[t] = shot_taper[pp][pt];
      wch[t] = ch[pp][pt];
      wtt[t] = tt[pp][pt];
      if (wwi) wwi[t] = widx[pp][pt];
      if (wtt[t] == O_LIVE) live = 1;
    }
    if (!live) { /* model.c:999-1006 */
      free(wshot);
      free(wch);
      free(wtt);
      free(wwi);
      continue;
    }
    fk_setup(&wi, m->nfk[pp], m->nconv[pp], m->wt[pp]);
    /* wavelet.c:18-103 */
    if (wwi) {
      wi.wsrc = (int *)malloc((size_t)wi.ntr * sizeof(int));
      for (t = 0; t < wi.ntr; t++) {
        int s;
        if (wtt[t] != O_LIVE) {
          wi.wsrc[t] = -1;
          continue;
        }
        for (s = 0; s < m->n_wsrc; s++) {
          if (m->wsrc[s].orig == wwi[t] && m->wsrc[s].n == wi.nconv) break;
        }
        if (s == m->n_wsrc) {
          m->wsrc = (owsrc *)realloc(m->wsrc, (size_t)(m->n_wsrc + 1) * sizeof(owsrc));
          m->wsrc[s].orig = wwi[t];
          m->wsrc[s].n = wi.nconv;
          m->wsrc[s].spec = NULL;
          m->n_wsrc++;
        }
        wi.wsrc[t] = s;
      }
    }
    wi.bc = (obc *)malloc((size_t)wi.ntr * sizeof(obc));
    blend_coords(wi.ntr, wi.nconv, wshot, wch, wtt, &m->cfg, wi.bc);
    wi.types = wtt;
    free(wshot);
    free(wch);
    free(wwi);
    m->win[m->n_win++] = wi;
  }
  free(all);
  for (p = 0; p < np; p++) free(shot_taper[p]);
  free(shot_taper);
  /* model offsets and workspace */
  {
    size_t off = 0;
    size_t maxw = 0;
    for (w = 0; w < m->n_win; w++) {
      owin *wi = &m->win[w];
      size_t const nd_work = offt_nd_work_elems(wi->ndim, wi->P);
      size_t const wv_work = offt_real_work_elems(wi->nconv) + (size_t)(wi->nconv / 2 + 1);
      wi->model_off = off;
      off += (size_t)wi->ntw * wi->nfk_el;
      if (nd_work > maxw) maxw = nd_work;
      if (wv_work > maxw) maxw = wv_work;
      if (wi->ntr * wi->nconv > m->max_ntr_conv) m->max_ntr_conv = wi->ntr * wi->nconv;
    }
    m->model_len = off;
    m->max_work = maxw;
  }
  /* wavelet.c:188-237 */
  for (w = 0; w < m->n_wsrc; w++) {
    int const n = m->wsrc[w].n;
    ocpx *fw = (ocpx *)malloc((offt_real_work_elems(n) + 16) * sizeof(ocpx));
    m->wsrc[w].spec = (ocpx *)malloc((size_t)(n / 2 + 1) * sizeof(ocpx));
    wavelet_spectrum(wav[m->wsrc[w].orig], wl[m->wsrc[w].orig], n,
                     m->wsrc[w].spec, fw);
    free(fw);
  }
  return m;
}

/* ------------------------------------------------------------------------ */
/* Model operators (model.c:1067-1423)                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  ocpx *c;   /* complex work: ntw*nfk_el */
  real *r;   /* real work: ntw*npad */
  ocpx *f;   /* fft scratch */
  real *tr;  /* trace buffer ntr*nconv */
} oscratch;

static void scratch_alloc(omodel const *m, oscratch *s) {
  size_t cmax = 1, rmax = 1;
  int w;
  for (w = 0; w < m->n_win; w++) {
    owin const *wi = &m->win[w];
    size_t const c = (size_t)wi->ntw * wi->nfk_el;
    size_t const r = (size_t)wi->ntw * wi->npad;
    if (c > cmax) cmax = c;
    if (r > rmax) rmax = r;
  }
  s->c = (ocpx *)malloc(cmax * sizeof(ocpx));
  s->r = (real *)malloc(rmax * sizeof(real));
  s->f = (ocpx *)malloc((m->max_work + 64) * sizeof(ocpx));
  s->tr = (real *)malloc(((size_t)m->max_ntr_conv + 1) * sizeof(real));
}

static void scratch_free(oscratch *s) {
  free(s->c);
  free(s->r);
  free(s->f);
  free(s->tr);
}

static int n_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Forward of every window into per-window trace buffers (parallel), then
 * blend in window order. model.c:1067-1114 */
static void model_forward(omodel const *m, ocpx const *x, real *rec) {
  int const nth = n_threads();
  real **bufs = (real **)malloc((size_t)m->n_win * sizeof(real *));
  int w;
  memset(rec, 0, m->cfg.rec_len * sizeof(real));
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
  {
    oscratch s;
    scratch_alloc(m, &s);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (w = 0; w < m->n_win; w++) {
      owin const *wi = &m->win[w];
      real *buf = (real *)malloc((size_t)wi->ntr * wi->nconv * sizeof(real));
      fk_forward(wi, x + wi->model_off, buf, s.c, s.r, s.f);
      wavelet_op(m, wi, buf, (real)1, s.c, s.f);
      bufs[w] = buf;
    }
    scratch_free(&s);
  }
  for (w = 0; w < m->n_win; w++) {
    owin const *wi = &m->win[w];
    blend_sum(bufs[w], wi->ntr, wi->nconv, wi->bc, &m->cfg, rec);
    free(bufs[w]);
  }
  free(bufs);
  blend_mute(&m->cfg, rec);
}

/* model.c:1116-1157 */
static void model_adjoint_update(omodel const *m, real const *rec, real step,
                                 ocpx *x) {
  int const nth = n_threads();
  int w;
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
  {
    oscratch s;
    ocpx *g;
    size_t gmax = 1;
    int ww;
    for (ww = 0; ww < m->n_win; ww++) {
      size_t const c = (size_t)m->win[ww].ntw * m->win[ww].nfk_el;
      if (c > gmax) gmax = c;
    }
    g = (ocpx *)malloc(gmax * sizeof(ocpx));
    scratch_alloc(m, &s);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (w = 0; w < m->n_win; w++) {
      owin const *wi = &m->win[w];
      size_t const n = (size_t)wi->ntw * wi->nfk_el;
      ocpx *xw = x + wi->model_off;
      size_t i;
      blend_adjoint(rec, wi->ntr, wi->nconv, wi->bc, &m->cfg, s.tr);
      wavelet_op(m, wi, s.tr, (real)-1, s.c, s.f);
      fk_adjoint(wi, s.tr, g, s.r, s.f);
      for (i = 0; i < n; i++) {
        xw[i].re -= step * g[i].re;
        xw[i].im -= step * g[i].im;
      }
    }
    scratch_free(&s);
    free(g);
  }
}

/* model.c:1202-1227 */
static real model_initial_lamb(omodel const *m, ocpx const *x) {
  real l2 = 0;
  size_t i;
  for (i = 0; i < m->model_len; i++) {
    real const a = x[i].re * x[i].re + x[i].im * x[i].im;
    if (a > l2) l2 = a;
  }
  return RSQRT(l2);
}

/* model.c:1230-1312 */
static real model_stepsize(omodel const *m) {
  real *rec = (real *)calloc(m->cfg.rec_len + 1, sizeof(real));
  real factor = 1;
  size_t i;
  int w;
  for (w = 0; w < m->n_win; w++) {
    owin const *wi = &m->win[w];
    int const nd = wi->ndim;
    real *buf = (real *)malloc((size_t)wi->ntr * wi->nconv * sizeof(real));
    int c[O_MAXC] = {0};
    int t, s, d;
    for (t = 0; t < wi->ntr; t++) {
      real sc = 1;
      for (d = nd - 2; d >= 0; d--) sc *= wi->taper[d][c[d]];
      for (s = 0; s < wi->nconv; s++) buf[(size_t)t * wi->nconv + s] = sc;
      for (d = nd - 2; d >= 0; d--) {
        if (++c[d] == wi->W[d]) c[d] = 0;
        else break;
      }
    }
    blend_sum(buf, wi->ntr, wi->nconv, wi->bc, &m->cfg, rec);
    free(buf);
  }
  for (i = 0; i < m->cfg.rec_len; i++)
    if (rec[i] > factor) factor = rec[i];
  free(rec);
  return (real)0.5 / factor;
}

/* model.c:1392-1423 */
static void model_shrink(omodel const *m, real lamb, ocpx *x) {
  real const l2 = lamb * lamb;
  size_t i;
  for (i = 0; i < m->model_len; i++) {
    real const a = x[i].re * x[i].re + x[i].im * x[i].im;
    if (a <= l2) {
      x[i].re = 0;
      x[i].im = 0;
    } else {
      real const ab = RSQRT(a);
      real const sh = (ab - lamb) / ab;
      x[i].re *= sh;
      x[i].im *= sh;
    }
  }
}

/* blend.c:1294-1335: mean blend of the patch data, then mute */
static void model_blend_input(omodel const *m, real const *const *data,
                              long const *const *shot, int const *const *ch,
                              int const *const *tt, real *rec) {
  int const np = m->n_patches;
  obc **bc = (obc **)malloc((size_t)np * sizeof(obc *));
  int p;
  memset(rec, 0, m->cfg.rec_len * sizeof(real));
  for (p = 0; p < np; p++) {
    bc[p] = (obc *)malloc((size_t)(m->ntr[p] + 1) * sizeof(obc));
    blend_coords(m->ntr[p], m->nt[p], shot[p], ch[p], tt[p], &m->cfg, bc[p]);
  }
  blend_mean(np, data, m->ntr, m->nt, bc, &m->cfg, 0, rec);
  blend_mute(&m->cfg, rec);
  for (p = 0; p < np; p++) free(bc[p]);
  free(bc);
}

/* The patch data summed into the record (blend.c:936-960 on the patch
 * traces) and muted: the data as its own prediction of the record, the
 * warm-start guard's baseline (no counterpart in the reference). */
static void model_blend_sum_input(omodel const *m, real const *const *data,
                                  long const *const *shot, int const *const *ch,
                                  int const *const *tt, real *rec) {
  int p;
  memset(rec, 0, m->cfg.rec_len * sizeof(real));
  for (p = 0; p < m->n_patches; p++) {
    obc *bc = (obc *)malloc((size_t)(m->ntr[p] + 1) * sizeof(obc));
    blend_coords(m->ntr[p], m->nt[p], shot[p], ch[p], tt[p], &m->cfg, bc);
    blend_sum(data[p], m->ntr[p], m->nt[p], bc, &m->cfg, rec);
    free(bc);
  }
  blend_mute(&m->cfg, rec);
}

/* model.c:1425-1515 */
static void model_output(omodel const *m, ocpx const *x,
                         int const *const *shape, real *const *data) {
  int const nth = n_threads();
  real **bufs = (real **)malloc((size_t)m->n_win * sizeof(real *));
  int p, w;
  for (p = 0; p < m->n_patches; p++)
    memset(data[p], 0, (size_t)m->ntr[p] * m->nt[p] * sizeof(real));
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
  {
    oscratch s;
    scratch_alloc(m, &s);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (w = 0; w < m->n_win; w++) {
      owin const *wi = &m->win[w];
      real *buf = (real *)malloc((size_t)wi->ntr * wi->nconv * sizeof(real));
      fk_forward(wi, x + wi->model_off, buf, s.c, s.r, s.f);
      wavelet_op(m, wi, buf, (real)1, s.c, s.f);
      bufs[w] = buf;
    }
    scratch_free(&s);
  }
  for (w = 0; w < m->n_win; w++) {
    owin const *wi = &m->win[w];
    int const pp = wi->patch;
    int const half = wi->W[wi->ndim - 1] / 2;
    int t, s;
    for (t = 0; t < wi->ntr; t++) {
      int const pt = window_patch_trace(wi, shape[pp], t);
      real *dst = data[pp] + (size_t)pt * m->nt[pp];
      real const *src = bufs[w] + (size_t)t * wi->nconv + half;
      for (s = 0; s < m->nt[pp]; s++) dst[s] += src[s];
    }
    free(bufs[w]);
  }
  free(bufs);
}

/* ------------------------------------------------------------------------ */
/* Input validation (agdeblend.c:247-509, 651-732)                           */
/* ------------------------------------------------------------------------ */

static int chk_pos(int v, char const *name) {
  if (v <= 0) {
    fprintf(stderr, "ERROR: %s must be > 0\n", name);
    return 1;
  }
  return 0;
}

static int chk_nn(void const *p, char const *name) {
  if (p == NULL) {
    fprintf(stderr, "ERROR: %s must not be NULL\n", name);
    return 1;
  }
  return 0;
}

static int check_deblend(int np, int const *volumes, int const *n_dims,
                         int const *const *wshape, int const *const *coords,
                         int const *const *shape, long const *const *shot,
                         int const *const *ch, int const *const *tt,
                         int const *wl, int const *const *widx,
                         real const *const *wav, real f, int n_its,
                         real *const *data) {
  int p, d;
  if (chk_pos(np, "n_patches")) goto err;
  if (chk_pos(n_its, "n_its")) goto err;
  if (f <= (real)0 || f > (real)1) {
    fprintf(stderr, "ERROR: initial_factor must be in (0, 1]\n");
    goto err;
  }
  if (chk_nn(volumes, "volumes") || chk_nn(n_dims, "n_dims") ||
      chk_nn(wshape, "window_shapes") || chk_nn(coords, "coords") ||
      chk_nn(shape, "shapes") || chk_nn(shot, "shottimes") ||
      chk_nn(ch, "channels") || chk_nn(tt, "trace_types") ||
      chk_nn(data, "data"))
    goto err;
  for (p = 0; p < np; p++) {
    int const v = volumes[p];
    size_t wsize = 1, ntr = 1;
    int maxt;
    if (chk_nn(coords[p], "coords") || chk_nn(shot[p], "shottimes") ||
        chk_nn(ch[p], "channels") || chk_nn(tt[p], "trace_types") ||
        chk_nn(data[p], "data"))
      goto err;
    if (n_dims[v] <= 1) {
      fprintf(stderr, "ERROR: n_dims must be > 1\n");
      goto err;
    }
    if (n_dims[v] - 1 > O_MAXC) {
      fprintf(stderr,
              "ERROR: AGD_MAX_N_COORDS is currently set to %d, but %d space "
              "dimensions have been requested. Please increase "
              "AGD_MAX_N_COORDS in the source code and recompile\n",
              O_MAXC, n_dims[v] - 1);
      goto err;
    }
    if (chk_nn(wshape[v], "window_shapes") || chk_nn(shape[p], "shapes")) goto err;
    maxt = shape[p][n_dims[v] - 1] + wshape[v][n_dims[v] - 1];
    for (d = 0; d < n_dims[v]; d++) {
      if (wshape[v][d] < 2) {
        fprintf(stderr, "ERROR: window_shapes must be at least 2\n");
        goto err;
      }
      if (wshape[v][d] != shape[p][d] && wshape[v][d] % 2 != 0) {
        fprintf(stderr,
                "ERROR: window_shapes must be even if it doesn't cover the "
                "whole dimension\n");
        goto err;
      }
      if (wshape[v][d] > shape[p][d]) {
        fprintf(stderr, "ERROR: window_shape must not be bigger than patch shape\n");
        goto err;
      }
    }
    for (d = 0; d < n_dims[v] - 1; d++) {
      wsize *= (size_t)wshape[v][d];
      ntr *= (size_t)shape[p][d];
    }
    if (2 * wsize * (size_t)maxt >= INT_MAX) {
      fprintf(stderr, "ERROR: The window for volume %d is too big\n", v);
      goto err;
    }
    if (ntr >= INT_MAX) {
      fprintf(stderr, "ERROR: Patches must have fewer than %d traces\n", INT_MAX);
      goto err;
    }
  }
  if (wl != NULL || widx != NULL || wav != NULL) {
    int *ntr, *vwl = NULL;
    int nwav = 0;
    if (chk_nn(wl, "wavelet_lengths") || chk_nn(widx, "wavelet_idxs") ||
        chk_nn(wav, "wavelets"))
      goto err;
    ntr = (int *)malloc((size_t)np * sizeof(int));
    for (p = 0; p < np; p++) {
      ntr[p] = 0;
      for (d = 0; d < n_dims[volumes[p]] - 1; d++)
        ntr[p] = ntr[p] == 0 ? shape[p][d] : ntr[p] * shape[p][d];
    }
    if (volume_wavelet_lengths(np, volumes, ntr, tt, wl, widx, &vwl)) {
      free(vwl);
      free(ntr);
      goto err;
    }
    for (p = 0; p < np; p++) {
      int const v = volumes[p];
      int t;
      if (chk_nn(widx[p], "wavelet_idxs")) {
        free(vwl);
        free(ntr);
        goto err;
      }
      if (ntr[p] <= 0) continue;
      if (vwl[v] == 0) {
        for (t = 0; t < ntr[p]; t++)
          if (tt[p][t] == O_LIVE) break;
        if (t == ntr[p]) continue;
      }
      for (t = 0; t < ntr[p]; t++) {
        int wi;
        if (tt[p][t] != O_LIVE) continue;
        wi = widx[p][t];
        if (wi < 0) {
          fprintf(stderr, "ERROR: wavelet_idx must not be negative\n");
          free(vwl);
          free(ntr);
          goto err;
        }
        if (wi + 1 > nwav) nwav = wi + 1;
        if (wl[wi] != vwl[v]) {
          fprintf(stderr,
                  "ERROR: wavelets in the same volume must have the same length\n");
          free(vwl);
          free(ntr);
          goto err;
        }
      }
    }
    free(vwl);
    free(ntr);
    for (p = 0; p < nwav; p++)
      if (chk_nn(wav[p], "wavelets")) goto err;
  }
  return 0;
err:
  fprintf(stderr, "ERROR in agd_deblend inputs\n");
  return 1;
}

/* Warm-start model of a data estimate y (patch layout): per window and
 * time window, x = s R2C(pad(y_w)), with the window's samples taken like
 * set_output places them (trace buffer sample t <-> data time t - W_t/2,
 * model.c:1478-1484), no tapers and no x2. Then F x = taper . y_w, and
 * where the window tapers sum to one the output of x is y. Live-window
 * data only; wavelets are not deconvolved. (No reference counterpart: the
 * reference always starts from the thresholded first update,
 * agdeblend.c:584-590.) */
static void model_from_data(omodel const *m, real const *const *y,
                            int const *const *shape, ocpx *x) {
  int const nth = n_threads();
  int w;
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
  {
    oscratch sc;
    scratch_alloc(m, &sc);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (w = 0; w < m->n_win; w++) {
      owin const *wi = &m->win[w];
      int const nd = wi->ndim;
      int const wt = wi->W[nd - 1];
      int const half = wt / 2;
      int const Pt = wi->P[nd - 1];
      int const pp = wi->patch;
      int const nt = m->nt[pp];
      int tw, t, s;
      memset(sc.r, 0, (size_t)wi->ntw * wi->npad * sizeof(real));
      for (t = 0; t < wi->ntr; t++) {
        int const pt = window_patch_trace(wi, shape[pp], t);
        int const zt = fk_padded_trace(wi, t);
        real const *src = y[pp] + (size_t)pt * nt;
        for (tw = 0; tw < wi->ntw; tw++) {
          real *z = sc.r + (size_t)tw * wi->npad + (size_t)zt * Pt;
          for (s = 0; s < wt; s++) {
            int const ts = tw * half + s - half;
            z[s] = (ts >= 0 && ts < nt) ? src[ts] * wi->scaler : (real)0;
          }
        }
      }
      for (tw = 0; tw < wi->ntw; tw++)
        offt_r2c_nd(nd, wi->P, sc.r + (size_t)tw * wi->npad,
                    x + wi->model_off + (size_t)tw * wi->nfk_el, sc.f);
    }
    scratch_free(&sc);
  }
}

/* ------------------------------------------------------------------------ */
/* Exported entry points                                                     */
/* ------------------------------------------------------------------------ */

/* agdeblend.c:511-649. stats receives the first threshold and stepsize
 * when non-NULL. With warm (patch data, the layout of data) the start model
 * is model_from_data(warm) instead of the thresholded first update, unless
 * that model's residual is no smaller than the zero model's or than the
 * residual of the data itself summed into the record (then the zero start;
 * stats[2] = 1 in that case, else 0); the threshold schedule still
 * starts from the first update's max |x|. */
/* Threshold schedule (agdeblend.c:526-533, :623-627): lamb decays by
 * `decay` per iteration until first_linear_it, then falls linearly to zero
 * at the last iteration; slope is per unit of the first threshold. */
static void lamb_schedule(real f, int n_its, real *decay, int *first_linear_it,
                          real *slope) {
  *decay = REXP(RLOG(O_TARGET_LAMB / f) / (real)n_its);
  *first_linear_it =
      (int)((RLOG(*decay) * ((real)n_its - (real)1) + (real)1) / RLOG(*decay));
  *slope = f * RPOW(*decay, (real)*first_linear_it) /
           (real)(n_its - 1 - *first_linear_it);
}

static int deblend_impl(int np, int const *volumes, int const *n_dims,
                        int const *const *wshape, int const *const *coords,
                        int const *const *shape, long const *const *shot,
                        int const *const *ch, int const *const *tt,
                        int const *wl, int const *const *widx,
                        real const *const *wav, real f, int n_its,
                        int print_freq, real *const *data, double *stats,
                        real const *const *warm) {
  real decay, slope;
  int first_linear_it;
  omodel *m;
  ocpx *x;
  real *d, *r;
  real step, lamb;
  int it;
  if (check_deblend(np, volumes, n_dims, wshape, coords, shape, shot, ch, tt,
                    wl, widx, wav, f, n_its, data))
    goto err0;
  lamb_schedule(f, n_its, &decay, &first_linear_it, &slope);
  m = model_new(np, volumes, n_dims, wshape, coords, shape, shot, ch, tt, wl,
                widx, wav);
  if (m == NULL) goto err0;
  x = (ocpx *)calloc(m->model_len + 1, sizeof(ocpx));
  d = (real *)calloc(m->cfg.rec_len + 1, sizeof(real));
  r = (real *)calloc(m->cfg.rec_len + 1, sizeof(real));
  step = model_stepsize(m);
  if (step > (real)0.3) {
    printf("WARNING: There does not appear to be any overlap between shots\n");
  }
  model_blend_input(m, (real const *const *)data, shot, ch, tt, d);
  model_adjoint_update(m, d, step, x);
  lamb = model_initial_lamb(m, x);
  if (stats) {
    stats[0] = (double)lamb;
    stats[1] = (double)step;
  }
  slope *= lamb;
  lamb *= f;
  if (warm != NULL) {
    /* guard: keep the warm model only when it fits the data better than
     * the zero model and better than the data itself as a prediction of
     * the record, ||A x_warm - d|| < min(||d||, ||mute(sum(data)) - d||)
     * (FP64 sums); otherwise the zero start's thresholded first update */
    double r2 = 0, d2 = 0, in2 = 0;
    size_t i;
    int fallback;
    model_blend_sum_input(m, (real const *const *)data, shot, ch, tt, r);
    for (i = 0; i < m->cfg.rec_len; i++) {
      double const ri = (double)(r[i] - d[i]);
      in2 += ri * ri;
    }
    model_from_data(m, warm, shape, x);
    model_forward(m, x, r);
    for (i = 0; i < m->cfg.rec_len; i++) {
      double const ri = (double)(r[i] - d[i]);
      r2 += ri * ri;
      d2 += (double)d[i] * d[i];
    }
    fallback = !(r2 < d2 && r2 < in2);
    if (fallback) {
      printf("WARNING: the warm start fits the data no better than the zero start or the "
             "data itself (residual %g, zero start %g, data %g); starting from zero\n",
             sqrt(r2), sqrt(d2), sqrt(in2));
      memset(x, 0, (m->model_len + 1) * sizeof(ocpx));
      model_adjoint_update(m, d, step, x);
      model_shrink(m, lamb, x);
    }
    if (stats) stats[2] = (double)fallback;
  } else {
    model_shrink(m, lamb, x);
  }
  if (print_freq > 0) {
    double acc = 0;
    size_t i;
    for (i = 0; i < m->cfg.rec_len; i++) acc += (double)d[i] * d[i];
    printf("Iteration 0 Residual %f\n", sqrt(acc));
  }
  for (it = 1; it < n_its; it++) {
    size_t i;
    model_forward(m, x, r);
    for (i = 0; i < m->cfg.rec_len; i++) r[i] -= d[i];
    if (print_freq > 0 && it % print_freq == 0) {
      real acc = 0;
      for (i = 0; i < m->cfg.rec_len; i++) acc += r[i] * r[i];
      printf("Iteration %d Residual %f\n", it, (double)RSQRT(acc));
    }
    model_adjoint_update(m, r, step, x);
    if (it < first_linear_it) {
      lamb *= decay;
    } else {
      lamb = (real)(n_its - 1 - it) * slope;
    }
    model_shrink(m, lamb, x);
  }
  model_output(m, x, shape, data);
  free(x);
  free(d);
  free(r);
  model_free(m);
  return 0;
err0:
  fprintf(stderr, "ERROR in agd_deblend\n");
  return 1;
}

EXPORT int OSYM(deblend)(int np, int const *volumes, int const *n_dims,
                         int const *const *wshape, int const *const *coords,
                         int const *const *shape, long const *const *shot,
                         int const *const *ch, int const *const *tt,
                         int const *wl, int const *const *widx,
                         real const *const *wav, real f, int n_its,
                         int print_freq, real *const *data, double *stats) {
  return deblend_impl(np, volumes, n_dims, wshape, coords, shape, shot, ch, tt,
                      wl, widx, wav, f, n_its, print_freq, data, stats, NULL);
}

EXPORT int OSYM(deblend_warm)(int np, int const *volumes, int const *n_dims,
                              int const *const *wshape, int const *const *coords,
                              int const *const *shape, long const *const *shot,
                              int const *const *ch, int const *const *tt,
                              int const *wl, int const *const *widx,
                              real const *const *wav, real f, int n_its,
                              int print_freq, real *const *data, double *stats,
                              real const *const *warm) {
  return deblend_impl(np, volumes, n_dims, wshape, coords, shape, shot, ch, tt,
                      wl, widx, wav, f, n_its, print_freq, data, stats, warm);
}

/* agdeblend.c:651-732 */
static int check_blend(int np, int const *ntr, int const *nt,
                       long const *const *shot, int const *const *ch,
                       int const *const *tt, real *const *data, int mode,
                       int taper, int npo, int const *ntro, int const *nto,
                       long const *const *shoto, int const *const *cho,
                       int const *const *tto, real *const *datao) {
  int p, t;
  if (chk_pos(np, "n_patches")) goto err;
  if (chk_pos(np, "n_patches_out")) goto err; /* reference quirk: same value */
  if (chk_nn(ntr, "n_traces_per_volume") ||
      chk_nn(ntro, "n_traces_per_volume_out") ||
      chk_nn(nt, "n_times_per_volume") || chk_nn(nto, "n_times_per_volume_out") ||
      chk_nn(shot, "shottimes") || chk_nn(shoto, "shottimes_out") ||
      chk_nn(ch, "channels") || chk_nn(cho, "channels_out") ||
      chk_nn(tt, "trace_types") || chk_nn(tto, "trace_types_out") ||
      chk_nn(data, "data") || chk_nn(datao, "data_out"))
    goto err;
  for (p = 0; p < np; p++) {
    if (chk_pos(ntr[p], "n_traces_per_volume")) goto err;
    if (chk_pos(nt[p], "n_times_per_volume")) goto err;
    if (chk_nn(shot[p], "shottimes") || chk_nn(ch[p], "channels") ||
        chk_nn(tt[p], "trace_types") || chk_nn(data[p], "data"))
      goto err;
  }
  for (p = 0; p < npo; p++) {
    if (chk_pos(ntro[p], "n_traces_per_volume_out")) goto err;
    if (chk_pos(nto[p], "n_times_per_volume_out")) goto err;
    if (chk_nn(shoto[p], "shottimes_out") || chk_nn(cho[p], "channels_out") ||
        chk_nn(tto[p], "trace_types_out") || chk_nn(datao[p], "data_out"))
      goto err;
    for (t = 0; t < ntro[p]; t++) {
      if (tto[p][t] == O_BAD) {
        fprintf(stderr, "ERROR: Output trace types may not be AGDBad\n");
        goto err;
      }
    }
  }
  if (mode == 1 && taper < 0) {
    fprintf(stderr, "ERROR: taper_length must be >= 0 when blend_mode is AGDBlendMean\n");
    goto err;
  }
  return 0;
err:
  fprintf(stderr, "ERROR in agd_blend inputs\n");
  return 1;
}

/* agdeblend.c:800-938 */
EXPORT int OSYM(blend)(int np, int const *ntr, int const *nt,
                       long const *const *shot, int const *const *ch,
                       int const *const *tt, real *const *data, int mode,
                       int taper, int npo, int const *ntro, int const *nto,
                       long const *const *shoto, int const *const *cho,
                       int const *const *tto, real *const *datao) {
  oblendcfg cfg;
  obc **bi, **bo;
  real *rec;
  int p;
  if (check_blend(np, ntr, nt, shot, ch, tt, data, mode, taper, npo, ntro, nto,
                  shoto, cho, tto, datao)) {
    fprintf(stderr, "ERROR in agd_blend\n");
    return 1;
  }
  blend_config(np, ntr, nt, shot, ch, tt, &cfg);
  bi = (obc **)malloc((size_t)np * sizeof(obc *));
  bo = (obc **)malloc((size_t)(npo + 1) * sizeof(obc *));
  for (p = 0; p < np; p++) {
    bi[p] = (obc *)malloc((size_t)(ntr[p] + 1) * sizeof(obc));
    blend_coords(ntr[p], nt[p], shot[p], ch[p], tt[p], &cfg, bi[p]);
  }
  for (p = 0; p < npo; p++) {
    bo[p] = (obc *)malloc((size_t)(ntro[p] + 1) * sizeof(obc));
    blend_coords(ntro[p], nto[p], shoto[p], cho[p], tto[p], &cfg, bo[p]);
  }
  rec = (real *)calloc(cfg.rec_len + 1, sizeof(real));
  if (mode == 0) {
    for (p = 0; p < np; p++) blend_sum(data[p], ntr[p], nt[p], bi[p], &cfg, rec);
  } else if (mode == 2) {
    for (p = 0; p < np; p++)
      blend_overwrite(data[p], ntr[p], nt[p], bi[p], &cfg, rec);
  } else {
    blend_mean(np, (real const *const *)data, ntr, nt, bi, &cfg, taper, rec);
  }
  blend_mute(&cfg, rec);
  for (p = 0; p < npo; p++) blend_adjoint(rec, ntro[p], nto[p], bo[p], &cfg, datao[p]);
  for (p = 0; p < np; p++) free(bi[p]);
  for (p = 0; p < npo; p++) free(bo[p]);
  free(bi);
  free(bo);
  free(rec);
  blend_config_free(&cfg);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Handle API for tests: model geometry, tables and single operators         */
/* ------------------------------------------------------------------------ */

EXPORT void *OSYM(model_new)(int np, int const *volumes, int const *n_dims,
                             int const *const *wshape, int const *const *coords,
                             int const *const *shape, long const *const *shot,
                             int const *const *ch, int const *const *tt,
                             int const *wl, int const *const *widx,
                             real const *const *wav) {
  return model_new(np, volumes, n_dims, wshape, coords, shape, shot, ch, tt, wl,
                   widx, wav);
}

EXPORT void OSYM(model_free)(void *h) { model_free((omodel *)h); }
EXPORT int OSYM(model_n_windows)(void *h) { return ((omodel *)h)->n_win; }
EXPORT long OSYM(model_len)(void *h) { return (long)((omodel *)h)->model_len; }
EXPORT long OSYM(record_len)(void *h) { return (long)((omodel *)h)->cfg.rec_len; }
EXPORT int OSYM(n_channels)(void *h) { return ((omodel *)h)->cfg.n_channels; }

/* info layout (ints): patch, ndim, ntr, ntw, nfk, nconv, Ft, npad,
 * coords[4], shape[4], use_taper[8], W[5], P[5], model_off (low 31 bits),
 * n_traces_padded */
EXPORT void OSYM(model_window)(void *h, int w, int *info) {
  owin const *wi = &((omodel *)h)->win[w];
  int k;
  memset(info, 0, 40 * sizeof(int));
  info[0] = wi->patch;
  info[1] = wi->ndim;
  info[2] = wi->ntr;
  info[3] = wi->ntw;
  info[4] = wi->nfk;
  info[5] = wi->nconv;
  info[6] = wi->Ft;
  info[7] = wi->npad;
  for (k = 0; k < O_MAXC; k++) {
    info[8 + k] = wi->coords[k];
    info[12 + k] = wi->shape[k];
  }
  for (k = 0; k < 2 * O_MAXC; k++) info[16 + k] = wi->use_taper[k];
  for (k = 0; k < O_MAXC + 1; k++) {
    info[24 + k] = wi->W[k];
    info[29 + k] = wi->P[k];
  }
  info[34] = (int)wi->model_off;
  info[35] = wi->npad_tr;
}

EXPORT real OSYM(model_window_scaler)(void *h, int w) {
  return ((omodel *)h)->win[w].scaler;
}

EXPORT void OSYM(model_window_taper)(void *h, int w, int d, real *out) {
  owin const *wi = &((omodel *)h)->win[w];
  memcpy(out, wi->taper[d], (size_t)wi->W[d] * sizeof(real));
}

/* blend coords of window w: per trace 5 int64: channel, interval,
 * interval_start, trace_start, n_times */
EXPORT void OSYM(model_window_blend)(void *h, int w, int64_t *out) {
  owin const *wi = &((omodel *)h)->win[w];
  int t;
  for (t = 0; t < wi->ntr; t++) {
    out[5 * t + 0] = wi->bc[t].channel_idx;
    out[5 * t + 1] = wi->bc[t].interval_idx;
    out[5 * t + 2] = (int64_t)wi->bc[t].interval_start;
    out[5 * t + 3] = wi->bc[t].trace_start;
    out[5 * t + 4] = wi->bc[t].n_times;
  }
}

/* channel intervals: returns count; out (if not NULL) gets
 * (channel, start, stop) int64 triples in channel then interval order */
EXPORT int OSYM(model_intervals)(void *h, int64_t *out) {
  oblendcfg const *cfg = &((omodel *)h)->cfg;
  int c, i, k = 0;
  for (c = 0; c < cfg->n_channels; c++) {
    for (i = 0; i < cfg->chan[c].n; i++) {
      if (out) {
        out[3 * k + 0] = c;
        out[3 * k + 1] = cfg->chan[c].iv[i].start;
        out[3 * k + 2] = cfg->chan[c].iv[i].stop;
      }
      k++;
    }
  }
  return k;
}

/* mute coords: returns count; out (if not NULL) gets (channel, interval,
 * interval_start, n_times) int64 quadruples */
EXPORT int OSYM(model_mutes)(void *h, int64_t *out) {
  oblendcfg const *cfg = &((omodel *)h)->cfg;
  int m;
  if (out) {
    for (m = 0; m < cfg->n_mute; m++) {
      out[4 * m + 0] = cfg->mute[m].channel_idx;
      out[4 * m + 1] = cfg->mute[m].interval_idx;
      out[4 * m + 2] = (int64_t)cfg->mute[m].interval_start;
      out[4 * m + 3] = cfg->mute[m].n_times;
    }
  }
  return cfg->n_mute;
}

EXPORT void OSYM(model_unique_channels)(void *h, int *out) {
  oblendcfg const *cfg = &((omodel *)h)->cfg;
  memcpy(out, cfg->uniq, (size_t)cfg->n_channels * sizeof(int));
}

EXPORT void OSYM(model_forward)(void *h, real const *x, real *rec) {
  model_forward((omodel *)h, (ocpx const *)x, rec);
}

EXPORT void OSYM(model_adjoint_update)(void *h, real const *rec, real step,
                                       real *x) {
  model_adjoint_update((omodel *)h, rec, step, (ocpx *)x);
}

EXPORT real OSYM(model_stepsize)(void *h) { return model_stepsize((omodel *)h); }

EXPORT void OSYM(model_shrink)(void *h, real lamb, real *x) {
  model_shrink((omodel *)h, lamb, (ocpx *)x);
}

EXPORT void OSYM(model_blend_input)(void *h, real const *const *data,
                                    long const *const *shot,
                                    int const *const *ch, int const *const *tt,
                                    real *rec) {
  model_blend_input((omodel *)h, data, shot, ch, tt, rec);
}

EXPORT void OSYM(model_output)(void *h, real const *x, int const *const *shape,
                               real *const *data) {
  model_output((omodel *)h, (ocpx const *)x, shape, data);
}

EXPORT void OSYM(model_from_data)(void *h, real const *const *y,
                                  int const *const *shape, real *x) {
  model_from_data((omodel *)h, y, shape, (ocpx *)x);
}

/* FK forward / adjoint of window w alone (out: ntr*nconv real) */
EXPORT void OSYM(window_fk_forward)(void *h, int w, real const *xw, real *out) {
  omodel *m = (omodel *)h;
  oscratch s;
  scratch_alloc(m, &s);
  fk_forward(&m->win[w], (ocpx const *)xw, out, s.c, s.r, s.f);
  scratch_free(&s);
}

EXPORT void OSYM(window_fk_adjoint)(void *h, int w, real const *in, real *gw) {
  omodel *m = (omodel *)h;
  oscratch s;
  scratch_alloc(m, &s);
  fk_adjoint(&m->win[w], in, (ocpx *)gw, s.r, s.f);
  scratch_free(&s);
}

EXPORT void OSYM(window_wavelet)(void *h, int w, real dir, real *buf) {
  omodel *m = (omodel *)h;
  oscratch s;
  scratch_alloc(m, &s);
  wavelet_op(m, &m->win[w], buf, dir, s.c, s.f);
  scratch_free(&s);
}

/* The thresholds of iterations 0..n_its-1 for a first threshold lamb0
 * (the sequence deblend_impl applies) */
EXPORT void OSYM(lamb_schedule)(real f, int n_its, real lamb0, real *out) {
  real decay, slope, lamb;
  int first_linear_it, it;
  lamb_schedule(f, n_its, &decay, &first_linear_it, &slope);
  slope *= lamb0;
  lamb = lamb0 * f;
  out[0] = lamb;
  for (it = 1; it < n_its; it++) {
    if (it < first_linear_it) {
      lamb *= decay;
    } else {
      lamb = (real)(n_its - 1 - it) * slope;
    }
    out[it] = lamb;
  }
}

/* Stand-alone pieces for known-answer tests */
EXPORT int OSYM(add_intervals)(int n, long const *starts, long const *stops,
                               long *out_starts, long *out_stops) {
  oivlist l = {NULL, 0, 0};
  int i, k;
  for (i = 0; i < n; i++) {
    oiv x;
    x.start = starts[i];
    x.stop = stops[i];
    ivl_add(&l, x);
  }
  for (k = 0; k < l.n; k++) {
    out_starts[k] = l.iv[k].start;
    out_stops[k] = l.iv[k].stop;
  }
  k = l.n;
  free(l.iv);
  return k;
}

EXPORT void OSYM(taper_window)(int len, int left, int right, real *out) {
  taper_window(len, left, right, out);
}

EXPORT void OSYM(fft_c2c)(int n, int sign, real *data) {
  offt_plan const *p = offt_get(n);
  ocpx *work = (ocpx *)malloc((offt_work_elems(n) + 1) * sizeof(ocpx));
  offt_exec(p, sign, (ocpx *)data, work);
  free(work);
}

EXPORT void OSYM(fft_r2c)(int n, real const *in, real *out) {
  ocpx *work = (ocpx *)malloc((offt_real_work_elems(n) + 1) * sizeof(ocpx));
  offt_r2c(n, in, (ocpx *)out, work);
  free(work);
}

EXPORT void OSYM(fft_c2r)(int n, real const *in, real *out) {
  ocpx *work = (ocpx *)malloc((offt_real_work_elems(n) + 1) * sizeof(ocpx));
  offt_c2r(n, (ocpx const *)in, out, work);
  free(work);
}

EXPORT void OSYM(fft_c2r_nd)(int rank, int const *n, real *in, real *out) {
  ocpx *work = (ocpx *)malloc((offt_nd_work_elems(rank, n) + 1) * sizeof(ocpx));
  offt_c2r_nd(rank, n, (ocpx *)in, out, work);
  free(work);
}

EXPORT void OSYM(fft_r2c_nd)(int rank, int const *n, real const *in, real *out) {
  ocpx *work = (ocpx *)malloc((offt_nd_work_elems(rank, n) + 1) * sizeof(ocpx));
  offt_r2c_nd(rank, n, in, (ocpx *)out, work);
  free(work);
}

EXPORT void OSYM(wavelet_spectrum)(real const *wav, int len, int n, real *spec) {
  ocpx *work = (ocpx *)malloc((offt_real_work_elems(n) + 1) * sizeof(ocpx));
  wavelet_spectrum(wav, len, n, (ocpx *)spec, work);
  free(work);
}

EXPORT void OSYM(set_threads)(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

EXPORT int OSYM(get_threads)(void) { return n_threads(); }

/* Window trace -> patch trace map (model.c:748-805), for the
 * GetWindowArrays known answers. */
EXPORT void OSYM(model_window_traces)(void *h, int w, int const *pshape, int *out) {
  owin const *wi = &((omodel *)h)->win[w];
  int t;
  for (t = 0; t < wi->ntr; t++) out[t] = window_patch_trace(wi, pshape, t);
}

/* Stand-alone blend configuration (blend.c:629-760) for known-answer tests */
typedef struct {
  oblendcfg cfg;
  int np;
  int *ntr, *nt;
  obc **bc;
} oblendh;

EXPORT void *OSYM(blendcfg_new)(int np, int const *ntr, int const *nt,
                                long const *const *shot, int const *const *ch,
                                int const *const *tt) {
  oblendh *b = (oblendh *)calloc(1, sizeof(oblendh));
  int p;
  b->np = np;
  b->ntr = (int *)malloc((size_t)(np + 1) * sizeof(int));
  b->nt = (int *)malloc((size_t)(np + 1) * sizeof(int));
  b->bc = (obc **)calloc((size_t)(np + 1), sizeof(obc *));
  blend_config(np, ntr, nt, shot, ch, tt, &b->cfg);
  for (p = 0; p < np; p++) {
    b->ntr[p] = ntr[p];
    b->nt[p] = nt[p];
    b->bc[p] = (obc *)malloc((size_t)(ntr[p] + 1) * sizeof(obc));
    blend_coords(ntr[p], nt[p], shot[p], ch[p], tt[p], &b->cfg, b->bc[p]);
  }
  return b;
}

EXPORT void OSYM(blendcfg_free)(void *h) {
  oblendh *b = (oblendh *)h;
  int p;
  for (p = 0; p < b->np; p++) free(b->bc[p]);
  free(b->bc);
  free(b->ntr);
  free(b->nt);
  blend_config_free(&b->cfg);
  free(b);
}

EXPORT int OSYM(blendcfg_n_channels)(void *h) { return ((oblendh *)h)->cfg.n_channels; }
EXPORT long OSYM(blendcfg_record_len)(void *h) { return (long)((oblendh *)h)->cfg.rec_len; }

EXPORT void OSYM(blendcfg_unique)(void *h, int *out) {
  oblendcfg const *cfg = &((oblendh *)h)->cfg;
  memcpy(out, cfg->uniq, (size_t)cfg->n_channels * sizeof(int));
}

EXPORT int OSYM(blendcfg_intervals)(void *h, int64_t *out) {
  oblendcfg const *cfg = &((oblendh *)h)->cfg;
  int c, i, k = 0;
  for (c = 0; c < cfg->n_channels; c++) {
    for (i = 0; i < cfg->chan[c].n; i++) {
      if (out) {
        out[3 * k + 0] = c;
        out[3 * k + 1] = cfg->chan[c].iv[i].start;
        out[3 * k + 2] = cfg->chan[c].iv[i].stop;
      }
      k++;
    }
  }
  return k;
}

EXPORT int OSYM(blendcfg_mutes)(void *h, int64_t *out) {
  oblendcfg const *cfg = &((oblendh *)h)->cfg;
  int m;
  if (out) {
    for (m = 0; m < cfg->n_mute; m++) {
      out[4 * m + 0] = cfg->mute[m].channel_idx;
      out[4 * m + 1] = cfg->mute[m].interval_idx;
      out[4 * m + 2] = (int64_t)cfg->mute[m].interval_start;
      out[4 * m + 3] = cfg->mute[m].n_times;
    }
  }
  return cfg->n_mute;
}

static void bc_to_i64(obc const *b, int n, int64_t *out) {
  int t;
  for (t = 0; t < n; t++) {
    out[5 * t + 0] = b[t].channel_idx;
    out[5 * t + 1] = b[t].interval_idx;
    out[5 * t + 2] = (int64_t)b[t].interval_start;
    out[5 * t + 3] = b[t].trace_start;
    out[5 * t + 4] = b[t].n_times;
  }
}

EXPORT void OSYM(blendcfg_coords)(void *h, int p, int64_t *out) {
  oblendh *b = (oblendh *)h;
  bc_to_i64(b->bc[p], b->ntr[p], out);
}

/* coords of arbitrary traces against this configuration */
EXPORT void OSYM(blendcfg_coords_for)(void *h, int ntr, int nt,
                                      long const *shot, int const *ch,
                                      int const *tt, int64_t *out) {
  oblendh *b = (oblendh *)h;
  obc *bc = (obc *)malloc((size_t)(ntr + 1) * sizeof(obc));
  blend_coords(ntr, nt, shot, ch, tt, &b->cfg, bc);
  bc_to_i64(bc, ntr, out);
  free(bc);
}

/* mode 0 sum, 1 mean (with taper), 2 overwrite; record zeroed first; no mute */
EXPORT void OSYM(blendcfg_forward)(void *h, int mode, int taper,
                                   real const *const *data, real *rec) {
  oblendh *b = (oblendh *)h;
  int p;
  memset(rec, 0, b->cfg.rec_len * sizeof(real));
  if (mode == 0) {
    for (p = 0; p < b->np; p++) blend_sum(data[p], b->ntr[p], b->nt[p], b->bc[p], &b->cfg, rec);
  } else if (mode == 2) {
    for (p = 0; p < b->np; p++)
      blend_overwrite(data[p], b->ntr[p], b->nt[p], b->bc[p], &b->cfg, rec);
  } else {
    blend_mean(b->np, data, b->ntr, b->nt, b->bc, &b->cfg, taper, rec);
  }
}

EXPORT void OSYM(blendcfg_mute)(void *h, real *rec) { blend_mute(&((oblendh *)h)->cfg, rec); }

EXPORT void OSYM(blendcfg_adjoint)(void *h, real const *rec, int p, real *out) {
  oblendh *b = (oblendh *)h;
  blend_adjoint(rec, b->ntr[p], b->nt[p], b->bc[p], &b->cfg, out);
}
