/*
 * C host of libagdeblend.so: input validation, problem geometry, blend and
 * FFT tables, device upload, and the ISTA and blend drivers. All arithmetic
 * on model and record samples runs in the HIP kernels (agd_kernels.hip)
 * reached through agd_gpu.h; this file only builds integer tables, a few
 * per-trace constants and the scalar threshold schedule.
 *
 * Reference behaviour reproduced here (citations relative to the reference
 * tree):
 *   validation            agdeblend.c:247-509, 651-732 (same messages)
 *   schedule              agdeblend.c:526-533, 584-629 (float, same order)
 *   n_times_fk / conv     model.c:326-357
 *   window geometry       model.c:20-50, 449-745, 748-805, 999-1006
 *   FK configuration      fk.c:19-169
 *   blend configuration   blend.c:19-342, 585-760 (first-appearance channel
 *                         numbering, insertion-with-merge interval order)
 *   wavelet configuration wavelet.c:18-237, utils.c:72-123
 * The tables are bit-identical to the reference's integer state; tests
 * compare them with the oracle's.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <string.h>

#include "agd_gpu.h"
#include "agdeblend.h"
#include "agdeblend_gpu.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define EXPORT __attribute__((visibility("default")))
#define MAXSP AGD_MAXSP
#define LDS_BUDGET (64 * 1024)
#define LDS_MAX (160 * 1024)
#define GATHER_TILE 256

/* ------------------------------------------------------------------------ */
/* Small growable arrays                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  void *p;
  size_t n, cap, esz;
} vec;

static void vec_init(vec *v, size_t esz) {
  v->p = NULL;
  v->n = 0;
  v->cap = 0;
  v->esz = esz;
}

static void *vec_push(vec *v) {
  if (v->n == v->cap) {
    size_t const nc = v->cap ? 2 * v->cap : 64;
    void *q = realloc(v->p, nc * v->esz);
    if (q == NULL) return NULL;
    v->p = q;
    v->cap = nc;
  }
  return (char *)v->p + (v->n++) * v->esz;
}

static void vec_free(vec *v) {
  free(v->p);
  v->p = NULL;
  v->n = v->cap = 0;
}

/* ------------------------------------------------------------------------ */
/* Validation (agdeblend.c:247-509)                                          */
/* ------------------------------------------------------------------------ */

static int check_positive(int v, char const *name) {
  if (v <= 0) {
    fprintf(stderr, "ERROR: %s must be > 0\n", name);
    return 1;
  }
  return 0;
}

static int check_non_null(void const *p, char const *name) {
  if (p == NULL) {
    fprintf(stderr, "ERROR: %s must not be NULL\n", name);
    return 1;
  }
  return 0;
}

static int n_volumes_of(int np, int const *volumes) {
  int nv = 0, p;
  for (p = 0; p < np; p++)
    if (volumes[p] > nv) nv = volumes[p];
  return nv + 1;
}

/* product of the space dims, as utils.c:20-34 */
static int traces_of(int nd, int const *shape) {
  int n = 0, d;
  for (d = 0; d < nd - 1; d++) n = (n == 0) ? shape[d] : n * shape[d];
  return n;
}

/* utils.c:72-123 */
static int volume_wavelet_lengths(int np, int const *volumes, int const *ntr,
                                  enum AGDTraceType const *const *tt,
                                  int const *wl, int const *const *widx,
                                  int **out) {
  int const nv = n_volumes_of(np, volumes);
  int p, v, t;
  *out = (int *)calloc((size_t)nv, sizeof(int));
  if (*out == NULL) return 1;
  if (widx == NULL) {
    for (v = 0; v < nv; v++) (*out)[v] = 1;
    return 0;
  }
  for (p = 0; p < np; p++) {
    v = volumes[p];
    if ((*out)[v] != 0) continue;
    for (t = 0; t < ntr[p]; t++) {
      if (tt[p][t] == AGDLive) {
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

static int check_deblend_inputs(int np, int const *volumes, int const *n_dims,
                                int const *const *wshape,
                                int const *const *coords,
                                int const *const *shape,
                                long int const *const *shot,
                                int const *const *ch,
                                enum AGDTraceType const *const *tt,
                                int const *wl, int const *const *widx,
                                float const *const *wav, float f, int n_its,
                                float *const *data) {
  int p, d;
  if (check_positive(np, "n_patches")) goto err;
  if (check_positive(n_its, "n_its")) goto err;
  if (f <= 0.0f || f > 1.0f) {
    fprintf(stderr, "ERROR: initial_factor must be in (0, 1]\n");
    goto err;
  }
  if (check_non_null(volumes, "volumes")) goto err;
  if (check_non_null(n_dims, "n_dims")) goto err;
  if (check_non_null(wshape, "window_shapes")) goto err;
  if (check_non_null(coords, "coords")) goto err;
  if (check_non_null(shape, "shapes")) goto err;
  if (check_non_null(shot, "shottimes")) goto err;
  if (check_non_null(ch, "channels")) goto err;
  if (check_non_null(tt, "trace_types")) goto err;
  if (check_non_null(data, "data")) goto err;
  for (p = 0; p < np; p++) {
    int const v = volumes[p];
    size_t wsize = 1, ntr = 1;
    int maxt;
    if (check_non_null(coords[p], "coords")) goto err;
    if (check_non_null(shot[p], "shottimes")) goto err;
    if (check_non_null(ch[p], "channels")) goto err;
    if (check_non_null(tt[p], "trace_types")) goto err;
    if (check_non_null(data[p], "data")) goto err;
    if (n_dims[v] <= 1) {
      fprintf(stderr, "ERROR: n_dims must be > 1\n");
      goto err;
    }
    if (n_dims[v] - 1 > AGD_MAX_N_COORDS) {
      fprintf(stderr,
              "ERROR: AGD_MAX_N_COORDS is currently set to %d, "
              "but %d space dimensions have been requested. "
              "Please increase AGD_MAX_N_COORDS in the source code and "
              "recompile\n",
              AGD_MAX_N_COORDS, n_dims[v] - 1);
      goto err;
    }
    if (check_non_null(wshape[v], "window_shapes")) goto err;
    if (check_non_null(shape[p], "shapes")) goto err;
    maxt = shape[p][n_dims[v] - 1] + wshape[v][n_dims[v] - 1];
    for (d = 0; d < n_dims[v]; d++) {
      if (wshape[v][d] < 2) {
        fprintf(stderr, "ERROR: window_shapes must be at least 2\n");
        goto err;
      }
      if (wshape[v][d] != shape[p][d] && wshape[v][d] % 2 != 0) {
        fprintf(stderr,
                "ERROR: window_shapes must be even if it doesn't "
                "cover the whole dimension\n");
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
    int nwav = 0;
    int *ntr = NULL, *vwl = NULL;
    if (check_non_null(wl, "wavelet_lengths")) goto err;
    if (check_non_null(widx, "wavelet_idxs")) goto err;
    if (check_non_null(wav, "wavelets")) goto err;
    ntr = (int *)malloc((size_t)np * sizeof(int));
    if (ntr == NULL) goto err;
    for (p = 0; p < np; p++) ntr[p] = traces_of(n_dims[volumes[p]], shape[p]);
    if (volume_wavelet_lengths(np, volumes, ntr, tt, wl, widx, &vwl)) {
      free(vwl);
      free(ntr);
      goto err;
    }
    for (p = 0; p < np; p++) {
      int const v = volumes[p];
      int t;
      if (check_non_null(widx[p], "wavelet_idxs")) {
        free(vwl);
        free(ntr);
        goto err;
      }
      if (ntr[p] <= 0) continue;
      if (vwl[v] == 0) {
        for (t = 0; t < ntr[p]; t++)
          if (tt[p][t] == AGDLive) break;
        if (t == ntr[p]) continue;
      }
      for (t = 0; t < ntr[p]; t++) {
        int wi;
        if (tt[p][t] != AGDLive) continue;
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
      if (check_non_null(wav[p], "wavelets")) goto err;
  }
  return 0;
err:
  fprintf(stderr, "ERROR in agd_deblend inputs\n");
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Channels and intervals (blend.c:19-342, 585-760)                          */
/* ------------------------------------------------------------------------ */

/* open-addressing map channel value -> first-appearance index */
typedef struct {
  int *key;
  int *val;
  size_t cap;
} chmap;

static int chmap_init(chmap *m, size_t n) {
  size_t cap = 16;
  while (cap < 2 * n + 16) cap *= 2;
  m->cap = cap;
  m->key = (int *)malloc(cap * sizeof(int));
  m->val = (int *)malloc(cap * sizeof(int));
  if (m->key == NULL || m->val == NULL) return 1;
  memset(m->val, 0xff, cap * sizeof(int)); /* -1 = empty */
  return 0;
}

static size_t chmap_slot(chmap const *m, int k) {
  uint32_t h = (uint32_t)k * 2654435761u;
  size_t s = (size_t)(h ^ (h >> 16)) & (m->cap - 1);
  while (m->val[s] >= 0 && m->key[s] != k) s = (s + 1) & (m->cap - 1);
  return s;
}

static int chmap_get(chmap const *m, int k) {
  size_t const s = chmap_slot(m, k);
  return m->val[s];
}

static void chmap_free(chmap *m) {
  free(m->key);
  free(m->val);
}

typedef struct {
  long start, stop;
} Iv;

typedef struct {
  Iv *iv;
  int n, cap;
} IvList;

static int iv_touch(Iv a, Iv b) {
  return (a.start < b.stop && a.stop > b.start) || a.start == b.stop ||
         a.stop == b.start;
}

static int ivlist_push(IvList *l, Iv x) {
  if (l->n == l->cap) {
    int const nc = l->cap ? 2 * l->cap : 4;
    Iv *q = (Iv *)realloc(l->iv, (size_t)nc * sizeof(Iv));
    if (q == NULL) return 1;
    l->iv = q;
    l->cap = nc;
  }
  l->iv[l->n++] = x;
  return 0;
}

/* blend.c:113-151 */
static int ivlist_add(IvList *l, Iv x) {
  int first = -1, i;
  for (i = 0; i < l->n; i++) {
    if (iv_touch(x, l->iv[i])) {
      first = i;
      break;
    }
  }
  if (first < 0) return ivlist_push(l, x);
  if (x.start < l->iv[first].start) l->iv[first].start = x.start;
  if (x.stop > l->iv[first].stop) l->iv[first].stop = x.stop;
  i = first + 1;
  while (i < l->n) {
    if (iv_touch(x, l->iv[i])) {
      int j;
      if (l->iv[i].start < l->iv[first].start) l->iv[first].start = l->iv[i].start;
      if (l->iv[i].stop > l->iv[first].stop) l->iv[first].stop = l->iv[i].stop;
      for (j = i; j < l->n - 1; j++) l->iv[j] = l->iv[j + 1];
      l->n--;
      /* the reference resumes its search at the index after the deleted
       * interval, skipping the interval that moved into its place */
      i = i + 1;
    } else {
      i++;
    }
  }
  return 0;
}

typedef struct {
  int n_channels;
  int *uniq;
  chmap map;
  IvList *chan;
  /* flat record */
  int64_t *iv_off; /* per (channel, interval) in channel-major order */
  int *chan_iv0;
  int64_t rec_len;
  /* mutes: flat (start, length) */
  int64_t *mute_start;
  int *mute_len;
  int n_mute;
  int64_t *mute_iv; /* channel, interval, interval_start, length (4 each) */
} BlendCfg;

static int blendcfg_channel(BlendCfg const *c, int channel, int type) {
  if (type == AGDMissing) return -1;
  return chmap_get(&c->map, channel);
}

static void blendcfg_free(BlendCfg *c) {
  int k;
  if (c->chan)
    for (k = 0; k < c->n_channels; k++) free(c->chan[k].iv);
  free(c->chan);
  free(c->uniq);
  chmap_free(&c->map);
  free(c->iv_off);
  free(c->chan_iv0);
  free(c->mute_start);
  free(c->mute_len);
  free(c->mute_iv);
  memset(c, 0, sizeof(*c));
}

/* blend.c:629-691 with shottimes/lengths per patch */
static int blendcfg_build(int np, int const *ntr, int const *nt,
                          long int const *const *shot, int const *const *ch,
                          enum AGDTraceType const *const *tt, BlendCfg *c) {
  size_t total = 0;
  int p, t, k, i;
  IvList *mute = NULL;
  vec uniq;
  memset(c, 0, sizeof(*c));
  for (p = 0; p < np; p++) total += (size_t)ntr[p];
  if (chmap_init(&c->map, total)) goto err;
  vec_init(&uniq, sizeof(int));
  for (p = 0; p < np; p++) {
    for (t = 0; t < ntr[p]; t++) {
      size_t s;
      if (tt[p][t] == AGDMissing) continue;
      s = chmap_slot(&c->map, ch[p][t]);
      if (c->map.val[s] < 0) {
        int *u = (int *)vec_push(&uniq);
        if (u == NULL) goto err;
        *u = ch[p][t];
        c->map.key[s] = ch[p][t];
        c->map.val[s] = (int)(uniq.n - 1);
      }
    }
  }
  c->n_channels = (int)uniq.n;
  c->uniq = (int *)uniq.p;
  c->chan = (IvList *)calloc((size_t)c->n_channels + 1, sizeof(IvList));
  mute = (IvList *)calloc((size_t)c->n_channels + 1, sizeof(IvList));
  if (c->chan == NULL || mute == NULL) goto err;
  for (p = 0; p < np; p++) {
    for (t = 0; t < ntr[p]; t++) {
      int const cidx = blendcfg_channel(c, ch[p][t], tt[p][t]);
      Iv x;
      if (cidx < 0) continue;
      x.start = shot[p][t];
      x.stop = shot[p][t] + nt[p];
      if (ivlist_add(&c->chan[cidx], x)) goto err;
      if (tt[p][t] == AGDBad)
        if (ivlist_add(&mute[cidx], x)) goto err;
    }
  }
  /* flat record layout */
  {
    int niv = 0;
    int64_t off = 0;
    int q = 0;
    for (k = 0; k < c->n_channels; k++) niv += c->chan[k].n;
    c->iv_off = (int64_t *)malloc((size_t)(niv + 1) * sizeof(int64_t));
    c->chan_iv0 = (int *)malloc((size_t)(c->n_channels + 1) * sizeof(int));
    if (c->iv_off == NULL || c->chan_iv0 == NULL) goto err;
    for (k = 0; k < c->n_channels; k++) {
      c->chan_iv0[k] = q;
      for (i = 0; i < c->chan[k].n; i++) {
        c->iv_off[q++] = off;
        off += c->chan[k].iv[i].stop - c->chan[k].iv[i].start;
      }
    }
    c->chan_iv0[c->n_channels] = q;
    c->iv_off[q] = off;
    c->rec_len = off;
  }
  /* mute coordinates (blend.c:585-627) */
  {
    vec ms, ml, mi;
    vec_init(&ms, sizeof(int64_t));
    vec_init(&ml, sizeof(int));
    vec_init(&mi, 4 * sizeof(int64_t));
    for (k = 0; k < c->n_channels; k++) {
      int m;
      for (m = 0; m < mute[k].n; m++) {
        Iv const mv = mute[k].iv[m];
        for (i = 0; i < c->chan[k].n; i++) {
          Iv const cv = c->chan[k].iv[i];
          if (mv.start < cv.stop && mv.stop > cv.start) {
            long const a = mv.start < cv.start ? cv.start : mv.start;
            long const b = mv.stop > cv.stop ? cv.stop : mv.stop;
            int64_t *s = (int64_t *)vec_push(&ms);
            int *l = (int *)vec_push(&ml);
            int64_t *q4 = (int64_t *)vec_push(&mi);
            if (!s || !l || !q4) goto err;
            *s = c->iv_off[c->chan_iv0[k] + i] + (a - cv.start);
            *l = (int)(b - a);
            q4[0] = k;
            q4[1] = i;
            q4[2] = a - cv.start;
            q4[3] = b - a;
          }
        }
      }
    }
    c->n_mute = (int)ms.n;
    c->mute_start = (int64_t *)ms.p;
    c->mute_len = (int *)ml.p;
    c->mute_iv = (int64_t *)mi.p;
  }
  for (k = 0; k < c->n_channels; k++) free(mute[k].iv);
  free(mute);
  return 0;
err:
  fprintf(stderr, "ERROR in set_blend_config\n");
  if (mute) {
    for (k = 0; k < c->n_channels; k++) free(mute[k].iv);
    free(mute);
  }
  return 1;
}

/* blend.c:693-726: coordinates of one trace. Returns n = 0 when the trace
 * is not blended (channel -1). */
typedef struct {
  int channel, interval;
  int64_t interval_start;
  int trace_start, n;
  int64_t rec; /* flat record index of the first blended sample */
} TraceCoord;

static void trace_coord(BlendCfg const *c, long shot, int ntimes, int channel,
                        int type, TraceCoord *o) {
  int const cidx = blendcfg_channel(c, channel, type);
  int i;
  memset(o, 0, sizeof(*o));
  o->channel = -1;
  if (type == AGDMissing || cidx < 0) return;
  for (i = 0; i < c->chan[cidx].n; i++) {
    Iv const cv = c->chan[cidx].iv[i];
    long const s0 = shot, s1 = shot + ntimes;
    if (s0 < cv.stop && s1 > cv.start) {
      long const a = s0 < cv.start ? cv.start : s0;
      long const b = s1 > cv.stop ? cv.stop : s1;
      o->channel = cidx;
      o->interval = i;
      o->interval_start = a - cv.start;
      o->trace_start = (int)(a - s0);
      o->n = (int)(b - a);
      o->rec = c->iv_off[c->chan_iv0[cidx] + i] + o->interval_start;
      return;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FFT plans                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
  vec plans; /* AgdFftPlan */
  vec roots; /* float pairs */
} PlanSet;

static int planset_get(PlanSet *s, int n) {
  AgdFftPlan *pl = (AgdFftPlan *)s->plans.p;
  size_t i;
  int rem = n, f;
  AgdFftPlan np_;
  for (i = 0; i < s->plans.n; i++)
    if (pl[i].n == n) return (int)i;
  memset(&np_, 0, sizeof(np_));
  np_.n = n;
  while (rem % 4 == 0 && rem > 1) {
    np_.radix[np_.nstage++] = 4;
    rem /= 4;
  }
  while (rem % 2 == 0 && rem > 1) {
    np_.radix[np_.nstage++] = 2;
    rem /= 2;
  }
  for (f = 3; f <= rem; f += 2) {
    while (rem % f == 0) {
      if (np_.nstage >= AGD_MAXSTAGE) return -1;
      np_.radix[np_.nstage++] = f;
      rem /= f;
    }
  }
  np_.roots_off = (int)(s->roots.n);
  for (f = 0; f < n; f++) {
    float *r = (float *)vec_push(&s->roots);
    double const a = -2.0 * M_PI * (double)f / (double)n;
    if (r == NULL) return -1;
    r[0] = (float)cos(a);
    r[1] = (float)sin(a);
  }
  {
    AgdFftPlan *q = (AgdFftPlan *)vec_push(&s->plans);
    if (q == NULL) return -1;
    *q = np_;
  }
  return (int)(s->plans.n - 1);
}

/* smallest even M >= n whose half has no prime factor above 7 */
static int smooth_even(int n) {
  int m = n + (n & 1);
  for (;; m += 2) {
    int r = m / 2;
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    while (r % 5 == 0) r /= 5;
    while (r % 7 == 0) r /= 7;
    if (r == 1) return m;
  }
}

/* real length of the wavelet convolution of n >= nconv + nfk - 1 samples: a
 * length with a compile-time kernel when one is at most 25% longer than the
 * smallest even 7-smooth length, else that length (the generic kernel) */
static int wavelet_len(int n) {
  static int const half[] = AGD_WAVELET_HALF_LENGTHS;
  int const m = smooth_even(n);
  size_t i;
  for (i = 0; i < sizeof half / sizeof half[0]; i++)
    if (2 * half[i] >= n) return 4 * (2 * half[i]) <= 5 * m ? 2 * half[i] : m;
  return m;
}

/* ------------------------------------------------------------------------ */
/* Problem                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
  int patch;
  int nsp;
  int coords[MAXSP];
  int shape[MAXSP];
  int use_taper[2 * MAXSP];
  int ntr;
} SpaceWin;

typedef struct {
  int orig;
  int n;
  int off; /* complex offset in the spectra pool */
} WavSrc;

struct AgdProblem {
  int n_patches;
  int *nd, *ntr, *nt, *wt, *nfk, *nconv;
  int64_t *data_off; /* per patch offset in the flat patch-data buffer */
  int64_t data_len;
  int n_win;
  AgdWin *win;
  SpaceWin *sw;
  int n_wtr;
  TraceCoord *wcoord; /* per global window trace */
  int *wpatch_trace;  /* per global window trace: patch trace index */
  BlendCfg cfg;
  /* user data blend */
  TraceCoord *dcoord; /* per patch trace (global) */
  int n_dtr;
  /* device */
  AgdDev dev;
  AgdGather dgather;   /* user data -> record (mean) */
  void *dalloc[64];
  int n_dalloc;
  float *d_data;       /* flat patch data */
  int64_t *d_o_off;
  int *d_o_nt, *d_o_ptr;
  int64_t *d_o_src;
  AgdEntry *d_went_data; /* per global window trace: its patch trace in the
                            patch data (sample 0 at window sample W_t / 2) */
  int n_out;
  float stepsize;
  int have_data;
  double input_fit2;   /* ||mute(sum blend(data)) - d||^2 of the data set
                          last (FP64): the warm-start guard's baseline */
  double bytes_M, bytes_T, bytes_R;
  int use_wavelets;
  int host_only; /* tables only, no device (CPU tests of the geometry) */
  AgdClass *sclass, *tclass;
  int n_sclass, n_tclass;
  int fuse;      /* adjoint kernels also write the next forward's stage 1 */
  int *wtrace0;  /* per window: first global window trace; n_wtr at n_win */
  int *chunk;    /* window chunk boundaries [n_chunk + 1] */
  int n_chunk;
  int *tblk;     /* host copy of the time blocks [n_tblk][TB_INTS] */
  int n_tblk;
  float *lamb_cur;  /* device threshold read by captured iterations */
  int sched_cap;    /* capacity of dev.lamb_sched (floats) */
  int u_valid;   /* U holds the forward spatial transform of X */
  int t_valid;   /* T holds the forward time transform of that U (no wavelets) */
  int *nk_req;   /* per spatial class: planes per task (0: the default rule) */
  size_t u_len;  /* complex elements of U and of V */
  int64_t t_len; /* floats of the trace buffer T (and T2) */
  AgdEntry *d_went_ident; /* single-operator tests: each window trace read
                             whole from a trace-buffer layout (lazy) */
  int *d_trace_len;       /* per global window trace: nconv (lazy) */
};

/* Every change of X or U goes through here: U's state is v, and T no
 * longer holds the forward of the current model. */
static void set_u_valid(AgdProblem *p, int v) {
  p->u_valid = v;
  p->t_valid = 0;
}

static void *dev_alloc(AgdProblem *p, size_t bytes) {
  void *q;
  if (p->host_only) return (void *)p; /* placeholder, never dereferenced */
  if (p->n_dalloc >= 64) return NULL;
  q = agdg_alloc(bytes);
  if (q) p->dalloc[p->n_dalloc++] = q;
  return q;
}

static void *dev_upload(AgdProblem *p, void const *src, size_t bytes) {
  void *q = dev_alloc(p, bytes ? bytes : 16);
  if (q == NULL || p->host_only) return q;
  if (bytes && agdg_h2d(q, src, bytes)) return NULL;
  return q;
}

EXPORT void agdx_destroy(AgdProblem *p) {
  int i;
  if (p == NULL) return;
  if (!p->host_only) agdg_sync();
  for (i = 0; i < p->n_dalloc; i++) agdg_free(p->dalloc[i]);
  free(p->nd);
  free(p->ntr);
  free(p->nt);
  free(p->wt);
  free(p->nfk);
  free(p->nconv);
  free(p->data_off);
  free(p->win);
  free(p->sw);
  free(p->wcoord);
  free(p->wpatch_trace);
  free(p->dcoord);
  for (i = 0; i < p->n_sclass; i++) free((void *)p->sclass[i].wpos);
  free(p->nk_req);
  for (i = 0; i < p->n_tclass; i++) free((void *)p->tclass[i].wpos);
  free(p->sclass);
  free(p->tclass);
  free(p->wtrace0);
  free(p->chunk);
  free(p->tblk);
  blendcfg_free(&p->cfg);
  free(p);
}

/* model.c:20-50 */
static void taper_from_neighbour(int vol, int nsp, int const *pc, int ovol,
                                 int const *oc, int *ut) {
  int d;
  if (ovol != vol) return;
  for (d = 0; d < nsp; d++)
    if (oc[d] != pc[d] - 1 && oc[d] != pc[d] && oc[d] != pc[d] + 1) return;
  for (d = 0; d < nsp; d++) {
    if (oc[d] == pc[d] - 1) ut[2 * d] = 1;
    else if (oc[d] == pc[d] + 1) ut[2 * d + 1] = 1;
  }
}

static int n_win_dim(int L, int W) { return W == L ? 1 : L / (W / 2) - 1; }

/* model.c:691-745: all windows of all patches, in the reference's order */
static int build_space_windows(AgdProblem *p, int const *volumes,
                               int const *n_dims, int const *const *wshape,
                               int const *const *coords,
                               int const *const *shape, SpaceWin **out,
                               int *n_out) {
  int const np = p->n_patches;
  int total = 0, base = 0, a, b, d;
  int *ut = (int *)calloc((size_t)np * 2 * MAXSP + 1, sizeof(int));
  SpaceWin *sw;
  if (ut == NULL) return 1;
  for (a = 0; a < np; a++) {
    int const nsp = n_dims[volumes[a]] - 1;
    int nw = 1;
    for (d = 0; d < nsp; d++) nw *= n_win_dim(shape[a][d], wshape[volumes[a]][d]);
    total += nw;
  }
  /* neighbour tapers: patches with equal volume, hashed by volume would be
   * faster; patch counts are small in practice */
  for (a = 0; a < np; a++) {
    int const nsp = n_dims[volumes[a]] - 1;
    for (b = 0; b < np; b++) {
      if (a == b) continue;
      taper_from_neighbour(volumes[a], nsp, coords[a], volumes[b], coords[b],
                           ut + (size_t)a * 2 * MAXSP);
    }
  }
  for (a = 0; a < np; a++) {
    int const nsp = n_dims[volumes[a]] - 1;
    int const *u = ut + (size_t)a * 2 * MAXSP;
    for (d = 0; d < nsp; d++) {
      if ((u[2 * d] || u[2 * d + 1]) && wshape[volumes[a]][d] % 2 == 1) {
        fprintf(stderr,
                "ERROR: window shapes may only be odd if there is no tapering. "
                "Patch %d touches another patch in dimension %d and so will be "
                "tapered, but has window shape %d.\n",
                a, d, wshape[volumes[a]][d]);
        free(ut);
        return 1;
      }
    }
  }
  sw = (SpaceWin *)calloc((size_t)total + 1, sizeof(SpaceWin));
  if (sw == NULL) {
    free(ut);
    return 1;
  }
  for (a = 0; a < np; a++) {
    int const nsp = n_dims[volumes[a]] - 1;
    int const *ws = wshape[volumes[a]];
    int const *u = ut + (size_t)a * 2 * MAXSP;
    int nwd[MAXSP], stride[MAXSP];
    int nw = 1, i;
    for (d = 0; d < nsp; d++) {
      nwd[d] = n_win_dim(shape[a][d], ws[d]);
      nw *= nwd[d];
    }
    {
      int st = 1;
      for (d = nsp - 1; d >= 0; d--) {
        stride[d] = st;
        st *= nwd[d];
      }
    }
    for (i = 0; i < nw; i++) {
      sw[base + i].patch = a;
      sw[base + i].nsp = nsp;
    }
    for (d = 0; d < nsp; d++) {
      int const tl = ws[d] / 2;
      int const n = nwd[d];
      int const bl = (shape[a][d] - tl) / n + tl;
      int const nl = shape[a][d] - tl - (bl - tl) * n;
      int start = 0, k;
      for (k = 0; k < n; k++) {
        int const len = bl + (k < nl);
        for (i = 0; i < nw; i++) {
          if ((i / stride[d]) % n != k) continue;
          sw[base + i].coords[d] = start;
          sw[base + i].shape[d] = len;
          sw[base + i].use_taper[2 * d] = (k == 0) ? u[2 * d] : 1;
          sw[base + i].use_taper[2 * d + 1] = (k == n - 1) ? u[2 * d + 1] : 1;
        }
        start += len - tl;
      }
    }
    for (i = 0; i < nw; i++) {
      sw[base + i].ntr = 1;
      for (d = 0; d < nsp; d++) sw[base + i].ntr *= sw[base + i].shape[d];
    }
    base += nw;
  }
  free(ut);
  *out = sw;
  *n_out = total;
  return 0;
}

static int window_patch_trace(SpaceWin const *w, int const *pshape, int t) {
  int d, idx = 0, stride = 1;
  for (d = w->nsp - 1; d >= 0; d--) {
    int const c = t % w->shape[d];
    t /= w->shape[d];
    idx += (w->coords[d] + c) * stride;
    stride *= pshape[d];
  }
  return idx;
}

/* fk.c:28-73 */
static void taper_window(int len, int left, int right, float *w) {
  int const tl = len / 2;
  int i;
  for (i = 0; i < len; i++) w[i] = 1.0f;
  if (left)
    for (i = 0; i < tl; i++)
      w[i] = 0.5f * (1.0f - cosf(2.0f * (float)M_PI * (float)i / (float)(2 * tl)));
  if (right)
    for (i = 0; i < tl; i++)
      w[len - tl + i] =
          0.5f * (1.0f - cosf(2.0f * (float)M_PI * (float)(tl + i) / (float)(2 * tl)));
}

/* Per-record-tile entry lists in serial order: CSR */
static int build_tiles(int64_t rec_len, int tile, TraceCoord const *coords,
                       int n, int **ptr_out, int **ent_out, int *n_tiles_out) {
  int64_t const nt64 = (rec_len + tile - 1) / tile;
  int const nt = (int)nt64;
  int *cnt = (int *)calloc((size_t)nt + 2, sizeof(int));
  int *ptr, *ent;
  int i;
  int64_t tot = 0;
  if (cnt == NULL) return 1;
  for (i = 0; i < n; i++) {
    int64_t a, b, k;
    if (coords[i].n <= 0) continue;
    a = coords[i].rec / tile;
    b = (coords[i].rec + coords[i].n - 1) / tile;
    for (k = a; k <= b; k++) cnt[k]++;
  }
  ptr = (int *)malloc(((size_t)nt + 1) * sizeof(int));
  if (ptr == NULL) {
    free(cnt);
    return 1;
  }
  for (i = 0; i < nt; i++) {
    ptr[i] = (int)tot;
    tot += cnt[i];
  }
  ptr[nt] = (int)tot;
  if (tot > INT_MAX) {
    fprintf(stderr, "ERROR: too many blend entries\n");
    free(cnt);
    free(ptr);
    return 1;
  }
  ent = (int *)malloc(((size_t)tot + 1) * sizeof(int));
  if (ent == NULL) {
    free(cnt);
    free(ptr);
    return 1;
  }
  memset(cnt, 0, ((size_t)nt + 2) * sizeof(int));
  for (i = 0; i < n; i++) {
    int64_t a, b, k;
    if (coords[i].n <= 0) continue;
    a = coords[i].rec / tile;
    b = (coords[i].rec + coords[i].n - 1) / tile;
    for (k = a; k <= b; k++) ent[ptr[k] + cnt[k]++] = i;
  }
  free(cnt);
  *ptr_out = ptr;
  *ent_out = ent;
  *n_tiles_out = nt;
  return 0;
}

/* Phase-only wavelet spectrum of wavelet.c:188-237 (host, double), then
 * the convolution spectrum used by k_wavelet: h = C2R_N(W) laid out for
 * lags in (-nfk, N) on length M, transformed, times 1/M. */
static int wavelet_conv_spectrum(float const *wav, int len, int N, int nfk, int M,
                                 float *out /* (M/2+1) complex */) {
  int const nf = N / 2 + 1;
  double *Wr = (double *)malloc((size_t)nf * sizeof(double));
  double *Wi = (double *)malloc((size_t)nf * sizeof(double));
  double *h = (double *)malloc((size_t)N * sizeof(double));
  double *hb = (double *)calloc((size_t)M, sizeof(double));
  double *cs = (double *)malloc((size_t)(N > M ? N : M) * sizeof(double));
  double *sn = (double *)malloc((size_t)(N > M ? N : M) * sizeof(double));
  int k, t, j;
  if (!Wr || !Wi || !h || !hb || !cs || !sn) {
    free(Wr);
    free(Wi);
    free(h);
    free(hb);
    free(cs);
    free(sn);
    return 1;
  }
  for (t = 0; t < N; t++) {
    cs[t] = cos(2.0 * M_PI * (double)t / (double)N);
    sn[t] = sin(2.0 * M_PI * (double)t / (double)N);
  }
  for (k = 0; k < nf; k++) {
    double re = 0, im = 0, ang;
    long e = 0;
    for (t = 0; t < len && t < N; t++) {
      re += (double)wav[t] * cs[e];
      im -= (double)wav[t] * sn[e];
      e += k;
      if (e >= N) e -= N;
    }
    ang = atan2(im, re);
    Wr[k] = cos(ang) / (double)N;
    Wi[k] = sin(ang) / (double)N;
  }
  /* h = C2R_N(W): imag of DC and Nyquist ignored */
  for (t = 0; t < N; t++) {
    double acc = Wr[0];
    long e = 0;
    for (k = 1; k < nf; k++) {
      double c;
      e += t;
      if (e >= N) e -= N;
      c = Wr[k] * cs[e] - Wi[k] * sn[e];
      if (N % 2 == 0 && k == N / 2) acc += Wr[k] * cs[e];
      else acc += 2.0 * c;
    }
    h[t] = acc;
  }
  /* lags j in (-nfk, N) at index j mod M */
  for (j = -(nfk - 1); j < N; j++) {
    int const src = ((j % N) + N) % N;
    int const dst = ((j % M) + M) % M;
    hb[dst] = h[src];
  }
  for (t = 0; t < M; t++) {
    cs[t] = cos(2.0 * M_PI * (double)t / (double)M);
    sn[t] = sin(2.0 * M_PI * (double)t / (double)M);
  }
  for (k = 0; k <= M / 2; k++) {
    double re = 0, im = 0;
    long e = 0;
    for (t = 0; t < M; t++) {
      if (hb[t] != 0.0) {
        re += hb[t] * cs[e];
        im -= hb[t] * sn[e];
      }
      e += k;
      if (e >= M) e -= M;
    }
    out[2 * k] = (float)(re / (double)M);
    out[2 * k + 1] = (float)(im / (double)M);
  }
  free(Wr);
  free(Wi);
  free(h);
  free(hb);
  free(cs);
  free(sn);
  return 0;
}

/* Sizes with specialised kernels (must match SPATIAL_SIZES / TIME_SIZES in
 * agd_kernels.hip) */
static int const kSpatialSizes[][2] = {{1, 10},  {1, 20},  {1, 40},  {10, 10},
                                       {10, 20}, {10, 40}, {20, 10}, {20, 20},
                                       {20, 40}, {40, 10}, {40, 20}, {40, 40}};
static int const kTimeSizes[][2] = {{5, 8}, {10, 8}, {10, 16}, {20, 16}};
/* every (M1, M2) factorization with a kernel (the autotuner's candidates) */
static int const kTimeCands[][2] = {{5, 8},   {8, 5},   {4, 10},  {10, 8},  {8, 10},
                                    {16, 5},  {10, 16}, {16, 10}, {20, 8},  {8, 20},
                                    {20, 16}, {16, 20}, {32, 10}, {10, 32}};

static void spatial_key(AgdWin const *W, int *p0, int *p1) {
  size_t i;
  int a = 0, b = 0;
  if (W->nsp == 1) {
    a = 1;
    b = W->P[0];
  } else if (W->nsp == 2) {
    a = W->P[0];
    b = W->P[1];
  }
  *p0 = 0;
  *p1 = 0;
  /* the specialised kernels take the window lengths as W = 4 P / 5 */
  for (i = 0; i < (size_t)W->nsp; i++)
    if (5 * W->W[i] != 4 * W->P[i]) return;
  for (i = 0; i < sizeof(kSpatialSizes) / sizeof(kSpatialSizes[0]); i++)
    if (kSpatialSizes[i][0] == a && kSpatialSizes[i][1] == b) {
      *p0 = a;
      *p1 = b;
    }
}

/* window length of a specialised padded size (W = 4 P / 5; 1 for P = 1) */
static int spec_len(int P) { return P == 1 ? 1 : P * 4 / 5; }

/* specialised time sizes of a window (Pt = 2 M1 M2 and Wt = 4 Pt / 5),
 * regardless of the layout */
static void time_sizes(AgdWin const *W, int *m1, int *m2) {
  size_t i;
  *m1 = 0;
  *m2 = 0;
  if (W->Pt % 2 || 5 * W->Wt != 4 * W->Pt) return;
  for (i = 0; i < sizeof(kTimeSizes) / sizeof(kTimeSizes[0]); i++)
    if (kTimeSizes[i][0] * kTimeSizes[i][1] == W->Pt / 2) {
      *m1 = kTimeSizes[i][0];
      *m2 = kTimeSizes[i][1];
    }
}

/* Traces per U / V tile: 16 when both transforms of the window have
 * specialised kernels and the traces form whole tiles, else the whole
 * window (untiled layout of the generic kernels). */
#define TIME_TILE 16
static int window_ug(AgdWin const *W) {
  int a, b;
  spatial_key(W, &a, &b);
  if (!a || W->ntr % TIME_TILE) return W->ntr;
  time_sizes(W, &a, &b);
  return a ? TIME_TILE : W->ntr;
}

/* the specialised time kernels serve tiled windows only */
static void time_key(AgdWin const *W, int *m1, int *m2) {
  *m1 = 0;
  *m2 = 0;
  if (W->ug == TIME_TILE && W->ntr % TIME_TILE == 0) time_sizes(W, m1, m2);
}

/* Spatial tasks (chunks of time-frequency planes) and time blocks (runs of
 * traces), each grouped into classes of windows with equal transform
 * sizes; a class is served by one kernel launch. */
/* planes per spatial task of a specialised class of padded sizes (a, b):
 * one column transform per thread of a 256-thread block (b per plane),
 * within a 48 KB LDS tile of spec_len(a) rows of b + 1 per plane;
 * AGD_SPATIAL_NK overrides */
static int default_nk(int a, int b) {
  size_t const plane = (size_t)spec_len(a) * (b + 1) * 8;
  char const *e = getenv("AGD_SPATIAL_NK");
  int nk = e ? atoi(e) : 256 / b;
  if (nk < 1) nk = 1;
  if ((size_t)nk * plane > 48 * 1024) nk = (int)(48 * 1024 / plane);
  return nk;
}

/* Spatial tasks of the classes cls[0, ncls): per window of the class, per
 * time window, runs of nk time frequencies (nk_req[c] when given and > 0,
 * else the default rule); fills each class's begin, count, wpos and LDS
 * size, and the generic kernels' LDS size in *lds_sp. */
static int spatial_tasks(AgdProblem *p, AgdClass *cls, int ncls, int const *nk_req,
                         vec *tasks, size_t *lds_sp) {
  int c, w;
  for (c = 0; c < ncls; c++) {
    AgdClass *cl = cls + c;
    int *wpos = (int *)malloc(((size_t)p->n_win + 1) * sizeof(int));
    if (wpos == NULL) return 1;
    cl->wpos = wpos;
    cl->lds = 0;
    cl->begin = (int)tasks->n;
    for (w = 0; w < p->n_win; w++) {
      AgdWin const *W = &p->win[w];
      int a, b, nk, tw, k0;
      wpos[w] = (int)tasks->n - cl->begin;
      spatial_key(W, &a, &b);
      if (a != cl->p0 || b != cl->p1 || (a && (W->ug != W->ntr) != cl->tb)) continue;
      if (a) {
        nk = (nk_req && nk_req[c] > 0) ? nk_req[c] : default_nk(a, b);
      } else {
        nk = (int)(LDS_BUDGET / (2 * (size_t)W->S * 8));
      }
      if (nk < 1) {
        fprintf(stderr, "ERROR: window of %d padded traces exceeds the LDS budget\n", W->S);
        return 1;
      }
      if (nk > W->Ft) nk = W->Ft;
      for (tw = 0; tw < W->ntw; tw++) {
        for (k0 = 0; k0 < W->Ft; k0 += nk) {
          AgdTask *tk = (AgdTask *)vec_push(tasks);
          if (tk == NULL) return 1;
          tk->win = w;
          tk->tw = tw;
          tk->k0 = k0;
          tk->nk = (k0 + nk <= W->Ft) ? nk : W->Ft - k0;
        }
      }
      if (a) {
        size_t const l = (size_t)nk * spec_len(a) * (b + 1) * 8;
        if (l > cl->lds) cl->lds = l;
      } else if ((size_t)2 * nk * W->S * 8 > *lds_sp) {
        *lds_sp = (size_t)2 * nk * W->S * 8;
      }
    }
    cl->count = (int)tasks->n - cl->begin;
    wpos[p->n_win] = cl->count;
  }
  return 0;
}

/* The adjoint time kernel's launch order: per time class, its blocks
 * sorted by (patch, patch trace index of the block's first trace, first
 * time window, window), so the blocks of overlapping windows that read the
 * same traces' record samples are neighbours (the kernel gives each XCD a
 * contiguous run). */
typedef struct {
  int64_t key;
  int tw, w, id;
} TbKey;
static int cmp_tbkey(void const *a, void const *b) {
  TbKey const *x = (TbKey const *)a, *y = (TbKey const *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  if (x->tw != y->tw) return x->tw < y->tw ? -1 : 1;
  if (x->w != y->w) return x->w < y->w ? -1 : 1;
  return (x->id > y->id) - (x->id < y->id);
}
static int *adj_time_order(AgdProblem const *p, int const *tb, int n) {
  int *ord = (int *)malloc(((size_t)n + 1) * sizeof(int));
  TbKey *k = (TbKey *)malloc(((size_t)n + 1) * sizeof(TbKey));
  int c, j;
  if (ord == NULL || k == NULL) {
    free(ord);
    free(k);
    return NULL;
  }
  for (c = 0; c < p->n_tclass; c++) {
    AgdClass const *cl = &p->tclass[c];
    for (j = 0; j < cl->count; j++) {
      int const id = cl->begin + j, w = tb[TB_INTS * id], t0 = tb[TB_INTS * id + 1];
      k[j].key = ((int64_t)p->win[w].patch << 32) | (uint32_t)p->wpatch_trace[p->win[w].trace0 + t0];
      k[j].tw = tb[TB_INTS * id + 2];
      k[j].w = w;
      k[j].id = id;
    }
    qsort(k, (size_t)cl->count, sizeof(TbKey), cmp_tbkey);
    for (j = 0; j < cl->count; j++) ord[cl->begin + j] = k[j].id;
  }
  free(k);
  return ord;
}

/* time windows per block of the specialised time kernels: all of a
 * window's (0) for the whole-volume order, where each block streams a
 * trace tile's time windows one after the other; 4 with the window-chunk
 * order (AGD_CHUNK_MB), so that a chunk of a few windows still makes
 * thousands of blocks; AGD_TIME_TW overrides */
static int time_tw_run(void) {
  char const *e = getenv("AGD_TIME_TW");
  if (e != NULL) return atoi(e) > 0 ? atoi(e) : 0;
  e = getenv("AGD_CHUNK_MB");
  return (e != NULL && atof(e) > 0.0) ? 4 : 0;
}

static int build_classes(AgdProblem *p, vec *tasks, vec *tblocks, size_t *lds_sp_out,
                         size_t *lds_t_out) {
  int w, c;
  int const tw_run = time_tw_run();
  size_t lds_sp = 0, lds_t = 0;
  vec sc, tc;
  vec_init(&sc, sizeof(AgdClass));
  vec_init(&tc, sizeof(AgdClass));
  for (w = 0; w < p->n_win; w++) {
    AgdWin const *W = &p->win[w];
    int a, b, found = 0;
    AgdClass *cl = (AgdClass *)sc.p;
    int const tiled = (W->ug != W->ntr);
    spatial_key(W, &a, &b);
    for (c = 0; c < (int)sc.n; c++)
      if (cl[c].p0 == a && cl[c].p1 == b && (!a || cl[c].tb == tiled)) found = 1;
    if (!found) {
      AgdClass *n = (AgdClass *)vec_push(&sc);
      if (n == NULL) return 1;
      memset(n, 0, sizeof(*n));
      n->p0 = a;
      n->p1 = b;
      n->tb = a ? tiled : 0; /* spatial classes: U / V tiled by TIME_TILE traces */
    }
    time_key(W, &a, &b);
    found = 0;
    cl = (AgdClass *)tc.p;
    for (c = 0; c < (int)tc.n; c++)
      if (cl[c].p0 == a && cl[c].p1 == b) found = 1;
    if (!found) {
      AgdClass *n = (AgdClass *)vec_push(&tc);
      if (n == NULL) return 1;
      memset(n, 0, sizeof(*n));
      n->p0 = a;
      n->p1 = b;
    }
  }
  if (spatial_tasks(p, (AgdClass *)sc.p, (int)sc.n, NULL, tasks, &lds_sp)) return 1;
  for (c = 0; c < (int)tc.n; c++) {
    AgdClass *cl = (AgdClass *)tc.p + c;
    int tb;
    /* traces per block: one tile for the specialised kernels, else as
     * many as fit the LDS budget, up to 16 */
    int tb_max = cl->p0 ? TIME_TILE : 16;
    int *wpos = (int *)malloc(((size_t)p->n_win + 1) * sizeof(int));
    if (wpos == NULL) return 1;
    cl->wpos = wpos;
    cl->var = 0;
    for (tb = tb_max; tb >= 1; tb /= 2) {
      size_t need = 0;
      for (w = 0; w < p->n_win; w++) {
        AgdWin const *W = &p->win[w];
        int a, b;
        size_t l;
        time_key(W, &a, &b);
        if (a != cl->p0 || b != cl->p1) continue;
        if (a) {
          /* the forward's tile, work buffer and roots (the larger of the
           * two kernels; the launcher passes each its own size) */
          size_t const M = (size_t)a * b;
          l = 8 * (2 * TIME_TILE * (M + 1) + 2 * M + 1);
        } else {
          int const M = (W->Pt % 2 == 0 && W->Pt >= 4) ? W->Pt / 2 : W->Pt;
          l = (size_t)tb * W->Ft * 8 + (size_t)2 * tb * M * 8 + (size_t)2 * tb * W->Wt * 4;
        }
        if (l > need) need = l;
      }
      cl->lds = need;
      if (cl->p0 || need <= (size_t)LDS_BUDGET || tb == 1) break;
    }
    cl->tb = tb;
    if (!cl->p0) lds_t = cl->lds > lds_t ? cl->lds : lds_t;
    cl->begin = (int)tblocks->n;
    for (w = 0; w < p->n_win; w++) {
      AgdWin const *W = &p->win[w];
      int a, b, t0, q, parts;
      wpos[w] = (int)tblocks->n - cl->begin;
      time_key(W, &a, &b);
      if (a != cl->p0 || b != cl->p1) continue;
      /* the specialised kernels' blocks split the window's time windows
       * into nearly equal runs of at most tw_run (0: one run) */
      parts = (a && tw_run > 0) ? (W->ntw + tw_run - 1) / tw_run : 1;
      if (parts < 1) parts = 1;
      for (t0 = 0; t0 < W->ntr; t0 += tb) {
        for (q = 0; q < parts; q++) {
          int *b4 = (int *)vec_push(tblocks);
          if (b4 == NULL) return 1;
          b4[0] = w;
          b4[1] = t0;
          b4[2] = (int)((int64_t)W->ntw * q / parts);
          b4[3] = (int)((int64_t)W->ntw * (q + 1) / parts);
        }
      }
    }
    cl->count = (int)tblocks->n - cl->begin;
    wpos[p->n_win] = cl->count;
  }
  p->sclass = (AgdClass *)sc.p;
  p->n_sclass = (int)sc.n;
  p->tclass = (AgdClass *)tc.p;
  p->n_tclass = (int)tc.n;
  *lds_sp_out = lds_sp;
  *lds_t_out = lds_t;
  return 0;
}

static AgdProblem *create_impl(int np, int const *volumes, int const *n_dims,
                               int const *const *wshape,
                               int const *const *coords,
                               int const *const *shape,
                               long int const *const *shot,
                               int const *const *ch,
                               enum AGDTraceType const *const *tt,
                               int const *wl, int const *const *widx,
                               float const *const *wav, int host_only) {
  AgdProblem *p = NULL;
  int *vwl = NULL;
  long **shot_taper = NULL;
  SpaceWin *all = NULL;
  int n_all = 0;
  int a, t, w, d;
  PlanSet ps;
  vec wsrc, wspec;
  vec_init(&ps.plans, sizeof(AgdFftPlan));
  vec_init(&ps.roots, 2 * sizeof(float));
  vec_init(&wsrc, sizeof(WavSrc));
  vec_init(&wspec, 2 * sizeof(float));
  if (!host_only && agdg_init()) return NULL;
  p = (AgdProblem *)calloc(1, sizeof(AgdProblem));
  if (p == NULL) return NULL;
  p->host_only = host_only;
  p->n_patches = np;
  p->nd = (int *)malloc((size_t)np * sizeof(int));
  p->ntr = (int *)malloc((size_t)np * sizeof(int));
  p->nt = (int *)malloc((size_t)np * sizeof(int));
  p->wt = (int *)malloc((size_t)np * sizeof(int));
  p->nfk = (int *)malloc((size_t)np * sizeof(int));
  p->nconv = (int *)malloc((size_t)np * sizeof(int));
  p->data_off = (int64_t *)malloc((size_t)(np + 1) * sizeof(int64_t));
  if (!p->nd || !p->ntr || !p->nt || !p->wt || !p->nfk || !p->nconv || !p->data_off)
    goto err;
  p->data_len = 0;
  for (a = 0; a < np; a++) {
    int const nd = n_dims[volumes[a]];
    p->nd[a] = nd;
    p->ntr[a] = traces_of(nd, shape[a]);
    p->nt[a] = shape[a][nd - 1];
    p->wt[a] = wshape[volumes[a]][nd - 1];
    p->data_off[a] = p->data_len;
    p->data_len += (int64_t)p->ntr[a] * p->nt[a];
  }
  p->data_off[np] = p->data_len;
  if (volume_wavelet_lengths(np, volumes, p->ntr, tt, wl, widx, &vwl)) goto err;
  /* model.c:326-357 */
  for (a = 0; a < np; a++) {
    int const half = p->wt[a] / 2;
    int const lw = vwl[volumes[a]];
    int const pre = p->nt[a] - lw + 1;
    int const ntw = (pre + half - 1) / half + 1;
    p->nfk[a] = (ntw + 1) * half;
    p->nconv[a] = p->nfk[a] + lw - 1;
  }
  /* model.c:424-446 */
  shot_taper = (long **)calloc((size_t)np, sizeof(long *));
  if (shot_taper == NULL) goto err;
  for (a = 0; a < np; a++) {
    shot_taper[a] = (long *)malloc(((size_t)p->ntr[a] + 1) * sizeof(long));
    if (shot_taper[a] == NULL) goto err;
    for (t = 0; t < p->ntr[a]; t++) shot_taper[a][t] = shot[a][t] - p->wt[a] / 2;
  }
  if (blendcfg_build(np, p->ntr, p->nconv, (long int const *const *)shot_taper, ch,
                     tt, &p->cfg))
    goto err;
  if (build_space_windows(p, volumes, n_dims, wshape, coords, shape, &all, &n_all)) {
    fprintf(stderr, "ERROR in set_space_window\n");
    goto err;
  }
  p->use_wavelets = widx != NULL;
  /* live windows (model.c:999-1006) */
  p->win = (AgdWin *)calloc((size_t)n_all + 1, sizeof(AgdWin));
  p->sw = (SpaceWin *)calloc((size_t)n_all + 1, sizeof(SpaceWin));
  if (p->win == NULL || p->sw == NULL) goto err;
  p->n_win = 0;
  p->n_wtr = 0;
  for (w = 0; w < n_all; w++) {
    SpaceWin const *s = &all[w];
    int live = 0;
    for (t = 0; t < s->ntr && !live; t++) {
      int const pt = window_patch_trace(s, shape[s->patch], t);
      if (tt[s->patch][pt] == AGDLive) live = 1;
    }
    if (!live) continue;
    p->sw[p->n_win] = *s;
    p->win[p->n_win].trace0 = p->n_wtr;
    p->n_wtr += s->ntr;
    p->n_win++;
  }
  /* per-window FK configuration and per-trace tables */
  {
    int64_t xo = 0, uo = 0, to = 0;
    vec tapers;
    float *tscale = (float *)malloc(((size_t)p->n_wtr + 1) * sizeof(float));
    float *tconst = (float *)malloc(((size_t)p->n_wtr + 1) * sizeof(float));
    int *wsrc_off = (int *)malloc(((size_t)p->n_wtr + 1) * sizeof(int));
    int *trace_win = (int *)malloc(((size_t)p->n_wtr + 1) * sizeof(int));
    AgdEntry *went = (AgdEntry *)malloc(((size_t)p->n_wtr + 1) * sizeof(AgdEntry));
    size_t lds_sp = 0, lds_t = 0, lds_w = 0;
    vec tasks, tblocks;
    p->wcoord = (TraceCoord *)malloc(((size_t)p->n_wtr + 1) * sizeof(TraceCoord));
    p->wpatch_trace = (int *)malloc(((size_t)p->n_wtr + 1) * sizeof(int));
    vec_init(&tapers, sizeof(float));
    vec_init(&tasks, sizeof(AgdTask));
    vec_init(&tblocks, TB_INTS * sizeof(int));
    if (!tscale || !tconst || !wsrc_off || !trace_win || !went || !p->wcoord ||
        !p->wpatch_trace)
      goto err;
    for (w = 0; w < p->n_win; w++) {
      SpaceWin const *s = &p->sw[w];
      AgdWin *W = &p->win[w];
      int const pa = s->patch;
      int const nsp = s->nsp;
      int Pall = 1;
      float scaler;
      float *sp_taper[MAXSP];
      int k0;
      W->nsp = nsp;
      W->patch = pa;
      W->S = 1;
      W->ntr = s->ntr;
      for (d = 0; d < nsp; d++) {
        W->W[d] = s->shape[d];
        W->P[d] = (int)ceilf((float)s->shape[d] * (1.0f + 0.25f));
        W->S *= W->P[d];
        Pall *= W->P[d];
      }
      W->Wt = p->wt[pa];
      W->half = W->Wt / 2;
      W->Pt = (int)ceilf((float)W->Wt * (1.0f + 0.25f));
      W->Ft = W->Pt / 2 + 1;
      W->nfk = p->nfk[pa];
      W->nconv = p->nconv[pa];
      W->ntw = W->nfk / W->half - 1; /* fk.c:24-26 */
      Pall *= W->Pt;
      scaler = 1.0f / sqrtf((float)Pall);
      W->scaler = scaler;
      for (d = 0; d < nsp; d++) {
        W->splan[d] = planset_get(&ps, W->P[d]);
        if (W->splan[d] < 0) goto err;
      }
      if (W->Pt % 2 == 0 && W->Pt >= 4) {
        W->tplan = planset_get(&ps, W->Pt / 2);
        W->tplan_full = planset_get(&ps, W->Pt);
      } else {
        W->tplan = planset_get(&ps, W->Pt);
        W->tplan_full = W->tplan;
      }
      if (W->tplan < 0 || W->tplan_full < 0) goto err;
      W->wplan = -1;
      W->wplan_full = -1;
      W->ug = window_ug(W);
      W->x_off = xo;
      W->u_off = uo;
      W->t_off = to;
      xo += (int64_t)W->ntw * W->Ft * W->S;
      uo += (int64_t)W->ntw * W->Ft * W->ntr;
      to += (int64_t)W->ntr * W->nconv;
      /* time taper */
      W->tt_off = (int)tapers.n;
      for (k0 = 0; k0 < W->Wt; k0++)
        if (vec_push(&tapers) == NULL) goto err;
      taper_window(W->Wt, 1, 1, (float *)tapers.p + W->tt_off);
      /* space tapers and per-trace scale (fk.c:333-347) */
      for (d = 0; d < nsp; d++) {
        sp_taper[d] = (float *)malloc((size_t)W->W[d] * sizeof(float));
        if (sp_taper[d] == NULL) goto err;
        taper_window(W->W[d], s->use_taper[2 * d], s->use_taper[2 * d + 1], sp_taper[d]);
      }
      {
        int c[MAXSP] = {0, 0, 0, 0};
        for (t = 0; t < W->ntr; t++) {
          float sc = scaler, one = 1.0f;
          for (d = nsp - 1; d >= 0; d--) {
            sc *= sp_taper[d][c[d]];
            one *= sp_taper[d][c[d]];
          }
          tscale[W->trace0 + t] = sc;
          tconst[W->trace0 + t] = one;
          for (d = nsp - 1; d >= 0; d--) {
            if (++c[d] == W->W[d]) c[d] = 0;
            else break;
          }
        }
      }
      for (d = 0; d < nsp; d++) free(sp_taper[d]);
      /* blend coordinates of each window trace (blend.c:728-760) */
      for (t = 0; t < W->ntr; t++) {
        int const g = W->trace0 + t;
        int const pt = window_patch_trace(s, shape[pa], t);
        TraceCoord *tc = &p->wcoord[g];
        p->wpatch_trace[g] = pt;
        trace_win[g] = w;
        trace_coord(&p->cfg, shot_taper[pa][pt], W->nconv, ch[pa][pt], tt[pa][pt], tc);
        went[g].rec = tc->rec;
        went[g].src = W->t_off + (int64_t)t * W->nconv;
        went[g].trace_start = tc->trace_start;
        went[g].n = tc->n;
        wsrc_off[g] = -1;
        if (p->use_wavelets && tt[pa][pt] == AGDLive) {
          /* wavelet.c:18-70: one spectrum per (wavelet, length) */
          WavSrc *ws = (WavSrc *)wsrc.p;
          size_t q;
          int const orig = widx[pa][pt];
          for (q = 0; q < wsrc.n; q++)
            if (ws[q].orig == orig && ws[q].n == W->nconv) break;
          if (q == wsrc.n) {
            WavSrc *nw = (WavSrc *)vec_push(&wsrc);
            if (nw == NULL) goto err;
            nw->orig = orig;
            nw->n = W->nconv;
            nw->off = -1;
          }
          wsrc_off[g] = (int)q; /* resolved to an offset below */
        }
      }
    }
    /* spatial tasks and time blocks, grouped by transform sizes */
    if (build_classes(p, &tasks, &tblocks, &lds_sp, &lds_t)) goto err;
    if (lds_sp > LDS_MAX || lds_t > LDS_MAX) {
      fprintf(stderr, "ERROR: window too large for the LDS-resident transforms\n");
      goto err;
    }
    /* wavelet spectra */
    if (p->use_wavelets) {
      WavSrc *ws = (WavSrc *)wsrc.p;
      size_t q;
      for (w = 0; w < p->n_win; w++) {
        AgdWin *W = &p->win[w];
        int const M = wavelet_len(W->nconv + W->nfk - 1);
        W->wM = M;
        W->wplan = planset_get(&ps, M / 2);
        W->wplan_full = planset_get(&ps, M);
        if (W->wplan < 0 || W->wplan_full < 0) goto err;
        if ((size_t)(3 * (M / 2) + 1) * 8 > lds_w) lds_w = (size_t)(3 * (M / 2) + 1) * 8;
      }
      for (q = 0; q < wsrc.n; q++) {
        /* all windows with this n share nfk (same patch family): find one */
        int nfk = 0, M = 0;
        for (w = 0; w < p->n_win; w++)
          if (p->win[w].nconv == ws[q].n) {
            nfk = p->win[w].nfk;
            M = p->win[w].wM;
            break;
          }
        ws[q].off = (int)wspec.n;
        for (t = 0; t <= M / 2; t++)
          if (vec_push(&wspec) == NULL) goto err;
        if (wavelet_conv_spectrum(wav[ws[q].orig], wl[ws[q].orig], ws[q].n, nfk, M,
                                  (float *)wspec.p + 2 * (size_t)ws[q].off))
          goto err;
      }
      for (t = 0; t < p->n_wtr; t++)
        if (wsrc_off[t] >= 0) wsrc_off[t] = ws[wsrc_off[t]].off;
      if (lds_w > LDS_MAX) {
        fprintf(stderr, "ERROR: wavelet convolution length too large\n");
        goto err;
      }
    }
    /* record gather lists */
    {
      int *tp = NULL, *te = NULL, nt = 0;
      unsigned char *mute;
      int m;
      if (build_tiles(p->cfg.rec_len, GATHER_TILE, p->wcoord, p->n_wtr, &tp, &te, &nt))
        goto err;
      p->dev.tile = GATHER_TILE;
      p->dev.n_tiles = nt;
      p->dev.tile_ptr = (int const *)dev_upload(p, tp, ((size_t)nt + 1) * sizeof(int));
      p->dev.tile_ent = (int const *)dev_upload(p, te, (size_t)tp[nt] * sizeof(int));
      free(tp);
      free(te);
      mute = (unsigned char *)calloc((size_t)p->cfg.rec_len + 1, 1);
      if (mute == NULL) goto err;
      for (m = 0; m < p->cfg.n_mute; m++)
        memset(mute + p->cfg.mute_start[m], 1, (size_t)p->cfg.mute_len[m]);
      p->dev.mute = (unsigned char const *)dev_upload(p, mute, (size_t)p->cfg.rec_len);
      free(mute);
    }
    p->dev.n_win = p->n_win;
    p->dev.win = (AgdWin const *)dev_upload(p, p->win, (size_t)p->n_win * sizeof(AgdWin));
    p->dev.plans = (AgdFftPlan const *)dev_upload(p, ps.plans.p, ps.plans.n * sizeof(AgdFftPlan));
    p->dev.roots = (float const *)dev_upload(p, ps.roots.p, ps.roots.n * 2 * sizeof(float));
    p->dev.tapers = (float const *)dev_upload(p, tapers.p, tapers.n * sizeof(float));
    p->dev.tscale = (float const *)dev_upload(p, tscale, (size_t)p->n_wtr * sizeof(float));
    p->dev.tconst = (float const *)dev_upload(p, tconst, (size_t)p->n_wtr * sizeof(float));
    p->dev.went = (AgdEntry const *)dev_upload(p, went, (size_t)p->n_wtr * sizeof(AgdEntry));
    p->dev.trace_win = (int const *)dev_upload(p, trace_win, (size_t)p->n_wtr * sizeof(int));
    p->dev.n_wtr = p->n_wtr;
    p->dev.tasks = (AgdTask const *)dev_upload(p, tasks.p, tasks.n * sizeof(AgdTask));
    p->dev.n_tasks = (int)tasks.n;
    p->dev.lds_spatial = lds_sp;
    p->dev.sclass = p->sclass;
    p->dev.n_sclass = p->n_sclass;
    p->dev.tclass = p->tclass;
    p->dev.n_tclass = p->n_tclass;
    /* window ranges: all windows by default; chunks of windows whose
     * intermediates U and V together stay below AGD_CHUNK_MB (unset or 0:
     * one chunk) for the chunked iteration order */
    p->wtrace0 = (int *)malloc(((size_t)p->n_win + 1) * sizeof(int));
    p->chunk = (int *)malloc(((size_t)p->n_win + 2) * sizeof(int));
    if (p->wtrace0 == NULL || p->chunk == NULL) goto err;
    for (w = 0; w < p->n_win; w++) p->wtrace0[w] = p->win[w].trace0;
    p->wtrace0[p->n_win] = p->n_wtr;
    p->dev.wtrace0 = p->wtrace0;
    p->dev.w0 = 0;
    p->dev.w1 = p->n_win;
    {
      char const *e = getenv("AGD_CHUNK_MB");
      double const cap = e ? atof(e) * 1e6 : 0.0;
      double acc = 0.0;
      p->n_chunk = 0;
      p->chunk[0] = 0;
      for (w = 0; w < p->n_win; w++) {
        double const b = 2.0 * 8.0 * (double)p->win[w].ntw * p->win[w].Ft * p->win[w].ntr;
        if (cap > 0.0 && acc > 0.0 && acc + b > cap) {
          p->chunk[++p->n_chunk] = w;
          acc = 0.0;
        }
        acc += b;
      }
      p->chunk[++p->n_chunk] = p->n_win;
    }
    p->dev.tblocks = (int const *)dev_upload(p, tblocks.p, tblocks.n * TB_INTS * sizeof(int));
    p->tblk = (int *)malloc((tblocks.n + 1) * TB_INTS * sizeof(int));
    if (p->tblk == NULL) goto err;
    memcpy(p->tblk, tblocks.p, tblocks.n * TB_INTS * sizeof(int));
    p->n_tblk = (int)tblocks.n;
    p->dev.n_tblocks = (int)tblocks.n;
    {
      int *ord = adj_time_order(p, (int const *)tblocks.p, (int)tblocks.n);
      p->dev.tadj = ord ? (int const *)dev_upload(p, ord, tblocks.n * sizeof(int)) : NULL;
      free(ord);
    }
    p->dev.lds_time = lds_t;
    p->dev.rec_len = p->cfg.rec_len;
    p->dev.use_wavelets = p->use_wavelets;
    p->dev.wspec = (float const *)dev_upload(p, wspec.p, wspec.n * 2 * sizeof(float));
    p->dev.wsrc_off = (int const *)dev_upload(p, wsrc_off, (size_t)p->n_wtr * sizeof(int));
    p->dev.lds_wavelet = lds_w;
    /* one transform length for every wavelet window: the specialised kernel
     * (AGD_GENERIC_WAVELET forces the generic one) */
    p->dev.wavelet_mh = 0;
    if (p->use_wavelets && getenv("AGD_GENERIC_WAVELET") == NULL) {
      int mh = -1;
      for (w = 0; w < p->n_win; w++) {
        if (p->win[w].wplan < 0) continue;
        if (mh < 0) mh = p->win[w].wM / 2;
        else if (mh != p->win[w].wM / 2) mh = 0;
      }
      p->dev.wavelet_mh = mh > 0 ? mh : 0;
    }
    p->dev.X = (float *)dev_alloc(p, (size_t)xo * 8);
    p->dev.U = (float *)dev_alloc(p, (size_t)uo * 8);
    p->dev.V = (float *)dev_alloc(p, (size_t)uo * 8);
    p->u_len = (size_t)uo;
    p->t_len = to;
    p->dev.T = (float *)dev_alloc(p, (size_t)to * 4);
    p->dev.T2 = p->use_wavelets ? (float *)dev_alloc(p, (size_t)to * 4) : p->dev.T;
    p->dev.d = (float *)dev_alloc(p, (size_t)p->cfg.rec_len * 4 + 16);
    p->dev.r = (float *)dev_alloc(p, (size_t)p->cfg.rec_len * 4 + 16);
    p->dev.scalar = (unsigned int *)dev_alloc(p, 64);
    p->sched_cap = 1 << 16;
    p->dev.lamb_sched = (float *)dev_alloc(p, (size_t)p->sched_cap * sizeof(float));
    p->dev.lamb_idx = (int *)dev_alloc(p, 64);
    p->lamb_cur = (float *)dev_alloc(p, 64);
    p->dev.lamb_dev = NULL;
    p->dev.n_partials = 1024;
    p->dev.partials = (double *)dev_alloc(p, 1024 * sizeof(double));
    p->bytes_M = (double)xo * 8.0;
    p->bytes_T = (double)to * 4.0;
    p->bytes_R = (double)p->cfg.rec_len * 4.0;
    free(tscale);
    free(tconst);
    free(wsrc_off);
    free(trace_win);
    free(went);
    vec_free(&tapers);
    vec_free(&tasks);
    vec_free(&tblocks);
    if (!p->dev.win || !p->dev.plans || !p->dev.roots || !p->dev.tapers ||
        !p->dev.tscale || !p->dev.tconst || !p->dev.went || !p->dev.trace_win ||
        !p->dev.tasks || !p->dev.tblocks || !p->dev.tadj || !p->dev.wsrc_off || !p->dev.X ||
        !p->dev.U || !p->dev.V || !p->dev.T || !p->dev.T2 || !p->dev.d || !p->dev.r ||
        !p->dev.scalar || !p->dev.lamb_sched || !p->dev.lamb_idx || !p->lamb_cur ||
        !p->dev.partials || !p->dev.tile_ptr || !p->dev.tile_ent ||
        !p->dev.mute || !p->dev.wspec)
      goto err;
    /* the wavelet adjoint writes each trace of T2 only below n_times_fk
     * (all that the adjoint time transform reads): the rest stays zero */
    if (!p->host_only && p->use_wavelets && agdg_zero(p->dev.T2, (size_t)to * 4)) goto err;
  }
  /* user-data blend coordinates (blend_input, blend.c:1294-1335) and
   * output assembly lists (set_output, model.c:1425-1515) */
  {
    int g = 0;
    AgdEntry *dent;
    int *tp = NULL, *te = NULL, nt = 0;
    int64_t *o_off;
    int *o_nt, *o_ptr;
    int64_t *o_src;
    int *cnt;
    int64_t ncontrib = 0;
    p->n_dtr = 0;
    for (a = 0; a < np; a++) p->n_dtr += p->ntr[a];
    p->dcoord = (TraceCoord *)malloc(((size_t)p->n_dtr + 1) * sizeof(TraceCoord));
    dent = (AgdEntry *)malloc(((size_t)p->n_dtr + 1) * sizeof(AgdEntry));
    if (p->dcoord == NULL || dent == NULL) goto err;
    for (a = 0; a < np; a++) {
      for (t = 0; t < p->ntr[a]; t++, g++) {
        trace_coord(&p->cfg, shot[a][t], p->nt[a], ch[a][t], tt[a][t], &p->dcoord[g]);
        dent[g].rec = p->dcoord[g].rec;
        dent[g].src = p->data_off[a] + (int64_t)t * p->nt[a];
        dent[g].trace_start = p->dcoord[g].trace_start;
        dent[g].n = p->dcoord[g].n;
      }
    }
    if (build_tiles(p->cfg.rec_len, GATHER_TILE, p->dcoord, p->n_dtr, &tp, &te, &nt)) {
      free(dent);
      goto err;
    }
    p->dgather.rec_len = p->cfg.rec_len;
    p->dgather.tile = GATHER_TILE;
    p->dgather.n_tiles = nt;
    p->dgather.tile_ptr = (int const *)dev_upload(p, tp, ((size_t)nt + 1) * sizeof(int));
    p->dgather.tile_ent = (int const *)dev_upload(p, te, (size_t)tp[nt] * sizeof(int));
    p->dgather.ent = (AgdEntry const *)dev_upload(p, dent, (size_t)p->n_dtr * sizeof(AgdEntry));
    p->dgather.mute = p->dev.mute;
    p->dgather.taper = NULL;
    p->dgather.taper_len = 0;
    free(tp);
    free(te);
    free(dent);
    /* output: per patch trace, contributions of window traces in window
     * order */
    p->n_out = p->n_dtr;
    o_off = (int64_t *)malloc(((size_t)p->n_out + 1) * sizeof(int64_t));
    o_nt = (int *)malloc(((size_t)p->n_out + 1) * sizeof(int));
    o_ptr = (int *)malloc(((size_t)p->n_out + 2) * sizeof(int));
    cnt = (int *)calloc((size_t)p->n_out + 2, sizeof(int));
    if (!o_off || !o_nt || !o_ptr || !cnt) goto err;
    {
      int64_t *pbase = (int64_t *)malloc(((size_t)np + 1) * sizeof(int64_t));
      int64_t acc = 0;
      if (pbase == NULL) goto err;
      for (a = 0; a < np; a++) {
        pbase[a] = acc;
        acc += p->ntr[a];
      }
      g = 0;
      for (a = 0; a < np; a++)
        for (t = 0; t < p->ntr[a]; t++, g++) {
          o_off[g] = p->data_off[a] + (int64_t)t * p->nt[a];
          o_nt[g] = p->nt[a];
        }
      for (w = 0; w < p->n_win; w++) {
        AgdWin const *W = &p->win[w];
        for (t = 0; t < W->ntr; t++) {
          int64_t const o = pbase[W->patch] + p->wpatch_trace[W->trace0 + t];
          cnt[o]++;
          ncontrib++;
        }
      }
      o_ptr[0] = 0;
      for (g = 0; g < p->n_out; g++) o_ptr[g + 1] = o_ptr[g] + cnt[g];
      o_src = (int64_t *)malloc(((size_t)ncontrib + 1) * sizeof(int64_t));
      if (o_src == NULL) {
        free(pbase);
        goto err;
      }
      memset(cnt, 0, ((size_t)p->n_out + 2) * sizeof(int));
      for (w = 0; w < p->n_win; w++) {
        AgdWin const *W = &p->win[w];
        for (t = 0; t < W->ntr; t++) {
          int64_t const o = pbase[W->patch] + p->wpatch_trace[W->trace0 + t];
          o_src[o_ptr[o] + cnt[o]++] = W->t_off + (int64_t)t * W->nconv + W->half;
        }
      }
      free(pbase);
    }
    p->d_o_off = (int64_t *)dev_upload(p, o_off, (size_t)p->n_out * sizeof(int64_t));
    p->d_o_nt = (int *)dev_upload(p, o_nt, (size_t)p->n_out * sizeof(int));
    p->d_o_ptr = (int *)dev_upload(p, o_ptr, ((size_t)p->n_out + 1) * sizeof(int));
    p->d_o_src = (int64_t *)dev_upload(p, o_src, (size_t)(ncontrib ? ncontrib : 1) * sizeof(int64_t));
    p->d_data = (float *)dev_alloc(p, (size_t)p->data_len * 4 + 16);
    {
      /* window traces of patch data (warm start), as entries the adjoint
       * time kernel reads like the record's: window sample t is data
       * sample t - W_t / 2 of the window trace's patch trace (set_output's
       * placement, model.c:1425-1515), zero outside the trace */
      AgdEntry *wd = (AgdEntry *)calloc((size_t)p->n_wtr + 1, sizeof(AgdEntry));
      if (wd == NULL) goto err;
      for (w = 0; w < p->n_win; w++) {
        AgdWin const *W = &p->win[w];
        int const nt = p->nt[W->patch];
        for (t = 0; t < W->ntr; t++) {
          int const gt = W->trace0 + t;
          wd[gt].rec = p->data_off[W->patch] + (int64_t)p->wpatch_trace[gt] * nt;
          wd[gt].trace_start = W->half;
          wd[gt].n = nt < W->nconv - W->half ? nt : W->nconv - W->half;
        }
      }
      p->d_went_data = (AgdEntry *)dev_upload(p, wd, (size_t)p->n_wtr * sizeof(AgdEntry));
      free(wd);
      if (!p->d_went_data) goto err;
    }
    free(o_off);
    free(o_nt);
    free(o_ptr);
    free(o_src);
    free(cnt);
    if (!p->d_o_off || !p->d_o_nt || !p->d_o_ptr || !p->d_o_src || !p->d_data ||
        !p->dgather.tile_ptr || !p->dgather.tile_ent || !p->dgather.ent)
      goto err;
  }
  p->fuse = !host_only && agdg_can_fuse(&p->dev) && getenv("AGD_NO_FUSE") == NULL;
  set_u_valid(p, 0);
  /* stepsize (model.c:1230-1312): fold of the space tapers, max with 1 */
  if (!host_only) {
    unsigned int bits = 0;
    float fold;
    if (agdg_zero(p->dev.scalar, 64)) goto err;
    if (agdg_stepsize_fold(&p->dev)) goto err;
    if (agdg_d2h(&bits, p->dev.scalar + 1, sizeof(bits))) goto err;
    memcpy(&fold, &bits, sizeof(fold));
    if (!(fold > 1.0f)) fold = 1.0f;
    p->stepsize = 0.5f / fold;
  }
  for (a = 0; a < np; a++) free(shot_taper[a]);
  free(shot_taper);
  free(vwl);
  free(all);
  vec_free(&ps.plans);
  vec_free(&ps.roots);
  vec_free(&wsrc);
  vec_free(&wspec);
  return p;
err:
  fprintf(stderr, "ERROR in set_model\n");
  if (shot_taper) {
    for (a = 0; a < np; a++) free(shot_taper[a]);
    free(shot_taper);
  }
  free(vwl);
  free(all);
  vec_free(&ps.plans);
  vec_free(&ps.roots);
  vec_free(&wsrc);
  vec_free(&wspec);
  agdx_destroy(p);
  return NULL;
}

EXPORT AgdProblem *agdx_create(int np, int const *volumes, int const *n_dims,
                               int const *const *wshape,
                               int const *const *coords,
                               int const *const *shape,
                               long int const *const *shot,
                               int const *const *ch,
                               enum AGDTraceType const *const *tt,
                               int const *wl, int const *const *widx,
                               float const *const *wav) {
  return create_impl(np, volumes, n_dims, wshape, coords, shape, shot, ch, tt, wl,
                     widx, wav, 0);
}

/* Host tables only (no device): geometry and blend tables for CPU tests */
EXPORT AgdProblem *agdx_create_host(int np, int const *volumes, int const *n_dims,
                                    int const *const *wshape,
                                    int const *const *coords,
                                    int const *const *shape,
                                    long int const *const *shot,
                                    int const *const *ch,
                                    enum AGDTraceType const *const *tt,
                                    int const *wl, int const *const *widx,
                                    float const *const *wav) {
  return create_impl(np, volumes, n_dims, wshape, coords, shape, shot, ch, tt, wl,
                     widx, wav, 1);
}

/* (channel, start, stop) per interval; returns the count */
EXPORT int agdx_intervals(AgdProblem const *p, long long *out) {
  int c, i, k = 0;
  for (c = 0; c < p->cfg.n_channels; c++)
    for (i = 0; i < p->cfg.chan[c].n; i++, k++)
      if (out) {
        out[3 * k] = c;
        out[3 * k + 1] = p->cfg.chan[c].iv[i].start;
        out[3 * k + 2] = p->cfg.chan[c].iv[i].stop;
      }
  return k;
}

/* (channel, interval, interval_start, n) per mute; returns the count */
EXPORT int agdx_mutes(AgdProblem const *p, long long *out) {
  int m;
  if (out)
    for (m = 0; m < p->cfg.n_mute; m++) {
      out[4 * m] = p->cfg.mute_iv[4 * m];
      out[4 * m + 1] = p->cfg.mute_iv[4 * m + 1];
      out[4 * m + 2] = p->cfg.mute_iv[4 * m + 2];
      out[4 * m + 3] = p->cfg.mute_iv[4 * m + 3];
    }
  return p->cfg.n_mute;
}

EXPORT int agdx_unique_channels(AgdProblem const *p, int *out) {
  if (out) memcpy(out, p->cfg.uniq, (size_t)p->cfg.n_channels * sizeof(int));
  return p->cfg.n_channels;
}

EXPORT long agdx_model_len(AgdProblem const *p) {
  int64_t n = 0;
  int w;
  for (w = 0; w < p->n_win; w++)
    n += (int64_t)p->win[w].ntw * p->win[w].Ft * p->win[w].S;
  return (long)n;
}

EXPORT long agdx_record_len(AgdProblem const *p) { return (long)p->cfg.rec_len; }
EXPORT int agdx_n_windows(AgdProblem const *p) { return p->n_win; }
EXPORT float agdx_stepsize(AgdProblem const *p) { return p->stepsize; }

EXPORT int agdx_window_info(AgdProblem const *p, int w, int *info) {
  AgdWin const *W;
  SpaceWin const *s;
  int d;
  if (w < 0 || w >= p->n_win) return 1;
  W = &p->win[w];
  s = &p->sw[w];
  memset(info, 0, 25 * sizeof(int));
  info[0] = W->patch;
  info[1] = W->nsp;
  info[2] = W->ntr;
  info[3] = W->ntw;
  info[4] = W->nfk;
  info[5] = W->nconv;
  info[6] = W->Pt;
  info[7] = W->Ft;
  info[8] = W->Wt;
  info[9] = W->S;
  info[10] = (int)(W->x_off & 0x7fffffff);
  info[11] = (int)(W->x_off >> 31);
  for (d = 0; d < W->nsp; d++) {
    info[12 + d] = W->P[d];
    info[16 + d] = W->W[d];
    info[20 + d] = s->coords[d];
  }
  info[24] = p->use_wavelets ? W->wM : 0;
  return 0;
}

/* model.c:642-653: taper flags (left, right) per space dim of window w */
EXPORT int agdx_window_taper(AgdProblem const *p, int w, int *use_taper) {
  int d;
  if (w < 0 || w >= p->n_win) return 1;
  for (d = 0; d < p->sw[w].nsp; d++) {
    use_taper[2 * d] = p->sw[w].use_taper[2 * d];
    use_taper[2 * d + 1] = p->sw[w].use_taper[2 * d + 1];
  }
  return 0;
}

/* model.c:748-805: the patch trace index of each trace of window w */
EXPORT int agdx_window_traces(AgdProblem const *p, int w, int *out) {
  int t;
  if (w < 0 || w >= p->n_win) return 1;
  for (t = 0; t < p->win[w].ntr; t++) out[t] = p->wpatch_trace[p->win[w].trace0 + t];
  return 0;
}

/* blend.c:113-342, 585-760: the blend tables agd_blend builds, on the host */
EXPORT int agdx_blend_tables(int np, int const *ntr, int const *nt,
                             long int const *const *shot, int const *const *ch,
                             enum AGDTraceType const *const *tt, int nq, int q_ntimes,
                             long int const *q_shot, int const *q_ch,
                             enum AGDTraceType const *q_tt, int *counts,
                             long long *intervals, long long *mutes, int *channels,
                             long long *coords, long long *q_coords) {
  BlendCfg cfg;
  int c, i, k, p, t, g;
  if (blendcfg_build(np, ntr, nt, shot, ch, tt, &cfg)) return 1;
  k = 0;
  for (c = 0; c < cfg.n_channels; c++)
    for (i = 0; i < cfg.chan[c].n; i++, k++)
      if (intervals) {
        intervals[3 * k] = c;
        intervals[3 * k + 1] = cfg.chan[c].iv[i].start;
        intervals[3 * k + 2] = cfg.chan[c].iv[i].stop;
      }
  counts[0] = k;
  counts[1] = cfg.n_mute;
  counts[2] = cfg.n_channels;
  if (mutes)
    for (i = 0; i < 4 * cfg.n_mute; i++) mutes[i] = cfg.mute_iv[i];
  if (channels) memcpy(channels, cfg.uniq, (size_t)cfg.n_channels * sizeof(int));
  g = 0;
  for (p = 0; p < np && coords; p++)
    for (t = 0; t < ntr[p]; t++, g++) {
      TraceCoord tc;
      trace_coord(&cfg, shot[p][t], nt[p], ch[p][t], tt[p][t], &tc);
      coords[5 * g] = tc.channel;
      coords[5 * g + 1] = tc.channel < 0 ? 0 : tc.interval;
      coords[5 * g + 2] = tc.channel < 0 ? 0 : tc.interval_start;
      coords[5 * g + 3] = tc.channel < 0 ? 0 : tc.trace_start;
      coords[5 * g + 4] = tc.channel < 0 ? 0 : tc.n;
    }
  for (t = 0; t < nq && q_coords; t++) {
    TraceCoord tc;
    trace_coord(&cfg, q_shot[t], q_ntimes, q_ch[t], q_tt[t], &tc);
    q_coords[5 * t] = tc.channel;
    q_coords[5 * t + 1] = tc.channel < 0 ? 0 : tc.interval;
    q_coords[5 * t + 2] = tc.channel < 0 ? 0 : tc.interval_start;
    q_coords[5 * t + 3] = tc.channel < 0 ? 0 : tc.trace_start;
    q_coords[5 * t + 4] = tc.channel < 0 ? 0 : tc.n;
  }
  blendcfg_free(&cfg);
  return 0;
}

EXPORT int agdx_window_blend(AgdProblem const *p, int w, long long *out) {
  AgdWin const *W;
  int t;
  if (w < 0 || w >= p->n_win) return 1;
  W = &p->win[w];
  for (t = 0; t < W->ntr; t++) {
    TraceCoord const *c = &p->wcoord[W->trace0 + t];
    out[5 * t + 0] = c->channel;
    out[5 * t + 1] = c->channel < 0 ? 0 : c->interval;
    out[5 * t + 2] = c->channel < 0 ? 0 : c->interval_start;
    out[5 * t + 3] = c->channel < 0 ? 0 : c->trace_start;
    out[5 * t + 4] = c->channel < 0 ? 0 : c->n;
  }
  return 0;
}

EXPORT double agdx_bytes(AgdProblem const *p, int which) {
  return which == 0 ? p->bytes_M : which == 1 ? p->bytes_T : p->bytes_R;
}

EXPORT int agdx_set_data(AgdProblem *p, float const *const *data, float *d_out) {
  int a;
  for (a = 0; a < p->n_patches; a++) {
    size_t const n = (size_t)p->ntr[a] * p->nt[a];
    if (agdg_h2d(p->d_data + p->data_off[a], data[a], n * 4)) return 1;
  }
  if (agdg_gather(&p->dgather, p->d_data, 1, p->dev.d)) return 1;
  /* the data as its own predictor of the record (the forward of its
   * warm-start model where the window tapers sum to one): the residual the
   * warm-start guard compares with; dev.r is recomputed before any use */
  if (agdg_gather_resid(&p->dgather, p->d_data, p->dev.d, p->dev.r) ||
      agdg_norm2(&p->dev, p->dev.r, p->cfg.rec_len, &p->input_fit2))
    return 1;
  if (d_out && agdg_d2h(d_out, p->dev.d, (size_t)p->cfg.rec_len * 4)) return 1;
  p->have_data = 1;
  return agdg_sync();
}

/* model_forward into dev.r (resid: minus d). The spatial stage is skipped
 * when the previous adjoint already produced it for the current model. */
static int forward_record(AgdProblem *p, int resid) {
  if (!p->u_valid) {
    if (agdg_fwd_spatial(&p->dev, p->dev.X)) return 1;
    set_u_valid(p, 1);
  }
  if (agdg_fwd_time(&p->dev)) return 1;
  if (p->use_wavelets && agdg_wavelet(&p->dev, 0)) return 1;
  p->t_valid = !p->use_wavelets;
  return agdg_blend_traces(&p->dev, resid);
}

/* model_adjoint_update of dev.r into dev.X (modes of agdg_adj_spatial) */
static int adjoint_update(AgdProblem *p, float alpha, float lamb, int mode) {
  int const fuse = p->fuse && mode == 0;
  if (p->use_wavelets) {
    if (agdg_wavelet(&p->dev, 1)) return 1;
    if (agdg_adj_time(&p->dev, p->dev.T2, 0)) return 1;
  } else {
    if (agdg_adj_time(&p->dev, NULL, 0)) return 1;
  }
  if (agdg_adj_spatial(&p->dev, alpha, lamb, mode, fuse)) return 1;
  set_u_valid(p, fuse);
  return 0;
}

EXPORT int agdx_forward(AgdProblem *p, float const *x, float *rec) {
  size_t const mb = (size_t)agdx_model_len(p) * 8;
  if (agdg_h2d(p->dev.X, x, mb)) return 1;
  set_u_valid(p, 0);
  if (forward_record(p, 0)) return 1;
  return agdg_d2h(rec, p->dev.r, (size_t)p->cfg.rec_len * 4);
}

EXPORT int agdx_adjoint(AgdProblem *p, float const *rec, float *g) {
  size_t const mb = (size_t)agdx_model_len(p) * 8;
  if (agdg_h2d(p->dev.r, rec, (size_t)p->cfg.rec_len * 4)) return 1;
  if (adjoint_update(p, 0.0f, 0.0f, 3)) return 1;
  return agdg_d2h(g, p->dev.X, mb);
}

EXPORT long agdx_trace_len(AgdProblem const *p) { return (long)p->t_len; }

/* Single operators of A = mute . blend . wavelet . F, each on its own, for
 * the per-operator adjoint tests (the loop composes them and fuses the
 * adjoints: blend^H into the time or wavelet kernels). Model vectors are
 * complex (model_len), trace vectors are the trace buffer (trace_len floats,
 * per window [ntr][nconv]), records record_len floats. */
static int ident_entries(AgdProblem *p) {
  AgdEntry *e;
  int *tl;
  int w, t;
  if (p->d_went_ident != NULL) return 0;
  e = (AgdEntry *)malloc(((size_t)p->n_wtr + 1) * sizeof(AgdEntry));
  tl = (int *)malloc(((size_t)p->n_wtr + 1) * sizeof(int));
  if (e == NULL || tl == NULL) {
    free(e);
    free(tl);
    return 1;
  }
  for (w = 0; w < p->n_win; w++)
    for (t = 0; t < p->win[w].ntr; t++) {
      AgdEntry *q = &e[p->win[w].trace0 + t];
      q->rec = p->win[w].t_off + (int64_t)t * p->win[w].nconv;
      q->src = q->rec;
      q->trace_start = 0;
      q->n = p->win[w].nconv;
      tl[p->win[w].trace0 + t] = p->win[w].nconv;
    }
  p->d_went_ident = (AgdEntry *)dev_upload(p, e, (size_t)p->n_wtr * sizeof(AgdEntry));
  p->d_trace_len = (int *)dev_upload(p, tl, (size_t)p->n_wtr * sizeof(int));
  free(e);
  free(tl);
  return p->d_went_ident == NULL || p->d_trace_len == NULL;
}

EXPORT int agdx_operator(AgdProblem *p, int op, int adjoint, float const *in, float *out) {
  size_t const mb = (size_t)agdx_model_len(p) * 8;
  size_t const tb = (size_t)p->t_len * 4;
  size_t const rb = (size_t)p->cfg.rec_len * 4;
  int rc = 0;
  if (p->host_only || op < 0 || op > 2) return 1;
  if ((op == 1 && !p->use_wavelets) || (op != 0 && ident_entries(p))) return 1;
  set_u_valid(p, 0);
  if (op == 0 && !adjoint) {
    /* F: model -> traces */
    if (agdg_h2d(p->dev.X, in, mb) || agdg_fwd_spatial(&p->dev, p->dev.X) ||
        agdg_fwd_time(&p->dev) || agdg_d2h(out, p->dev.T, tb))
      rc = 1;
  } else if (op == 0) {
    /* F^H: traces -> model */
    if (agdg_h2d(p->dev.T, in, tb) || agdg_adj_time(&p->dev, p->dev.T, 0) ||
        agdg_adj_spatial(&p->dev, 0.0f, 0.0f, 3, 0) || agdg_d2h(out, p->dev.X, mb))
      rc = 1;
  } else if (op == 1 && !adjoint) {
    /* wavelet convolution, in place on the traces */
    if (agdg_h2d(p->dev.T, in, tb) || agdg_wavelet(&p->dev, 0) || agdg_d2h(out, p->dev.T, tb))
      rc = 1;
  } else if (op == 1) {
    /* its adjoint reads whole traces through identity entries (in place of
     * the record's blend entries) into T2 */
    float *const r = p->dev.r;
    AgdEntry const *const went = p->dev.went;
    if (agdg_h2d(p->dev.T, in, tb)) return 1;
    p->dev.r = p->dev.T;
    p->dev.went = p->d_went_ident;
    rc = agdg_wavelet(&p->dev, 1);
    p->dev.r = r;
    p->dev.went = went;
    if (rc || agdg_d2h(out, p->dev.T2, tb)) rc = 1;
  } else if (!adjoint) {
    /* blend into the record and mute */
    if (agdg_h2d(p->dev.T, in, tb) || agdg_blend_traces(&p->dev, 0) ||
        agdg_d2h(out, p->dev.r, rb))
      rc = 1;
  } else {
    /* blend^H: record slices back into the window traces (no mute) */
    if (agdg_h2d(p->dev.r, in, rb) ||
        agdg_extract(p->dev.went, p->n_wtr, p->d_trace_len, p->dev.r, p->dev.T) ||
        agdg_sync() || agdg_d2h(out, p->dev.T, tb))
      rc = 1;
  }
  return rc || agdg_sync();
}

EXPORT int agdx_get_model(AgdProblem *p, float *x) {
  return agdg_d2h(x, p->dev.X, (size_t)agdx_model_len(p) * 8);
}

static float g_last_lambda0 = 0.0f;
static int g_last_warm_fallback = 0;
static double g_last_warm_ratio = -1.0;

EXPORT float agdx_last_lambda0(void) { return g_last_lambda0; }

EXPORT int agdx_last_warm_fallback(void) { return g_last_warm_fallback; }

EXPORT double agdx_last_warm_ratio(void) { return g_last_warm_ratio; }

/* agdeblend.c:526-629 */
/* Warm-start model s R2C(pad(y_w)) of a data estimate y (patch layout)
 * into dev.X: the plain time R2C (scaler, no tapers, no x2) of the window
 * traces of y, read in place through d_went_data (placed as set_output
 * places them), and the spatial DFTs. */
static int warm_model_flat(AgdProblem *p, float const *y_dev) {
  float *const r = p->dev.r;
  AgdEntry const *const went = p->dev.went;
  int rc;
  p->dev.r = (float *)y_dev; /* read only */
  p->dev.went = p->d_went_data;
  rc = agdg_adj_time(&p->dev, NULL, 1);
  p->dev.r = r;
  p->dev.went = went;
  if (rc) return 1;
  /* with the fused kernels the model's forward spatial stage comes along */
  if (agdg_adj_spatial(&p->dev, 0.0f, 0.0f, 3, p->fuse)) return 1;
  set_u_valid(p, p->fuse);
  return 0;
}

static int warm_model(AgdProblem *p, float const *const *y) {
  int a;
  for (a = 0; a < p->n_patches; a++) {
    size_t const n = (size_t)p->ntr[a] * p->nt[a];
    if (y[a] == NULL) {
      fprintf(stderr, "ERROR: warm-start data of patch %d is NULL\n", a);
      return 1;
    }
    if (agdg_h2d(p->d_data + p->data_off[a], y[a], n * 4)) return 1;
  }
  return warm_model_flat(p, p->d_data);
}

EXPORT float const *agdx_data_dev(AgdProblem *p) { return p->host_only ? NULL : p->d_data; }

EXPORT int64_t agdx_data_len(AgdProblem *p) { return p->data_len; }

EXPORT int agdx_warm_model_dev(AgdProblem *p, float const *y_dev) {
  if (p->host_only || y_dev == NULL) return 1;
  if (warm_model_flat(p, y_dev)) return 1;
  return agdg_sync();
}

EXPORT int agdx_warm_model(AgdProblem *p, float const *const *y) {
  if (p->host_only) return 1;
  if (warm_model(p, y)) return 1;
  return agdg_sync();
}

static int run_impl(AgdProblem *p, float f, int n_its, int print_freq, float const *x0,
                    float const *const *warm, float const *warm_dev);

static int forward_record(AgdProblem *p, int resid);
static int adjoint_update(AgdProblem *p, float alpha, float lamb, int mode);

/* One iteration in window-chunk order, starting from the residual record of
 * the current model: for each chunk of windows, the adjoint update with
 * threshold lamb and then (unless last) the chunk's forward traces for the
 * next iteration, whose intermediates are still in cache; finally the
 * residual record of the new model. The same kernels over the same
 * elements as forward_record + adjoint_update, in another launch order. */
static int chunked_iteration(AgdProblem *p, float alpha, float lamb, int last) {
  int k, rc = 0;
  for (k = 0; k < p->n_chunk && rc == 0; k++) {
    p->dev.w0 = p->chunk[k];
    p->dev.w1 = p->chunk[k + 1];
    if (adjoint_update(p, alpha, lamb, 0)) rc = 1;
    if (rc == 0 && !last) {
      if (!p->fuse && agdg_fwd_spatial(&p->dev, p->dev.X)) rc = 1;
      if (rc == 0 && agdg_fwd_time(&p->dev)) rc = 1;
      if (rc == 0 && p->use_wavelets && agdg_wavelet(&p->dev, 0)) rc = 1;
    }
  }
  p->dev.w0 = 0;
  p->dev.w1 = p->n_win;
  if (rc == 0 && !last && agdg_blend_traces(&p->dev, 1)) rc = 1;
  set_u_valid(p, p->fuse);
  return rc;
}

/* n iterations with thresholds lamb[0..n): one iteration captured into a
 * graph whose update kernels read the threshold that agdg_next_lamb copies
 * from the device schedule, launched n times (chunked: in window-chunk
 * order, none of them the last iteration) */
static int run_captured(AgdProblem *p, float alpha, float const *lamb, int n, int chunked) {
  void *exec = NULL;
  int zero = 0;
  int rc = 0;
  if (agdg_h2d(p->dev.lamb_sched, lamb, (size_t)n * sizeof(float))) return 1;
  if (agdg_h2d(p->dev.lamb_idx, &zero, sizeof(zero))) return 1;
  if (agdg_graph_begin()) return 1;
  p->dev.lamb_dev = p->lamb_cur;
  if (agdg_next_lamb(&p->dev)) rc = 1;
  else if (chunked) rc = chunked_iteration(p, alpha, 0.0f, 0);
  else if (forward_record(p, 1) || adjoint_update(p, alpha, 0.0f, 0)) rc = 1;
  p->dev.lamb_dev = NULL;
  {
    /* end the capture even after a failed launch */
    int const rc2 = agdg_graph_end(&exec);
    if (rc || rc2) {
      agdg_graph_destroy(exec);
      return 1;
    }
  }
  for (int i = 0; i < n && rc == 0; i++) rc = agdg_graph_launch(exec);
  if (rc == 0) rc = agdg_sync();
  agdg_graph_destroy(exec);
  return rc;
}

/* Iterations 1 .. n_its - 1 with thresholds sched[it], from the current
 * model; r_ready: dev.r already holds its residual A x - d (otherwise the
 * first iteration computes it). */
static int run_loop(AgdProblem *p, float alpha, float const *sched, int n_its,
                    int print_freq, int r_ready) {
  int it, rc = 0;
  /* from the second iteration on, every iteration launches the same
   * kernels with the same arguments except the threshold: capture one
   * into a graph that reads its threshold from the device schedule, and
   * replay it (not while printing residuals or timing kernels) */
  int const graph = n_its > 4 && n_its - 2 <= p->sched_cap && print_freq <= 0 &&
                    !agdg_timing_enabled() && getenv("AGD_NO_GRAPH") == NULL;
  int const chunked = p->n_chunk > 1;
  if (chunked && n_its > 1 && !r_ready && forward_record(p, 1)) rc = 1;
  for (it = 1; it < n_its && rc == 0; it++) {
    if (chunked) {
      if (graph && it == 2) {
        /* iterations 2 .. n_its - 2 replayed, the last one eager */
        rc = run_captured(p, alpha, sched + 2, n_its - 3, 1);
        it = n_its - 2;
        continue;
      }
      if (print_freq > 0 && it % print_freq == 0) {
        double r2 = 0.0;
        if (agdg_norm2(&p->dev, p->dev.r, p->cfg.rec_len, &r2)) rc = 1;
        else printf("Iteration %d Residual %f\n", it, sqrt(r2));
      }
      if (rc == 0) rc = chunked_iteration(p, alpha, sched[it], it == n_its - 1);
      continue;
    }
    if (graph && it == 2) {
      rc = run_captured(p, alpha, sched + 2, n_its - 2, 0);
      break;
    }
    if (!(it == 1 && r_ready) && forward_record(p, 1)) rc = 1;
    if (rc == 0 && print_freq > 0 && it % print_freq == 0) {
      double r2 = 0.0;
      if (agdg_norm2(&p->dev, p->dev.r, p->cfg.rec_len, &r2)) rc = 1;
      else printf("Iteration %d Residual %f\n", it, sqrt(r2));
    }
    if (rc == 0 && adjoint_update(p, alpha, sched[it], 0)) rc = 1;
  }
  return rc;
}

EXPORT int agdx_run(AgdProblem *p, float f, int n_its, int print_freq,
                    float const *x0) {
  return run_impl(p, f, n_its, print_freq, x0, NULL, NULL);
}

EXPORT int agdx_run_warm(AgdProblem *p, float f, int n_its, int print_freq,
                         float const *const *warm) {
  if (warm == NULL) {
    fprintf(stderr, "ERROR: warm-start data is NULL\n");
    return 1;
  }
  return run_impl(p, f, n_its, print_freq, NULL, warm, NULL);
}

EXPORT int agdx_run_warm_dev(AgdProblem *p, float f, int n_its, int print_freq,
                             float const *warm_dev) {
  if (warm_dev == NULL) {
    fprintf(stderr, "ERROR: warm-start data is NULL\n");
    return 1;
  }
  return run_impl(p, f, n_its, print_freq, NULL, NULL, warm_dev);
}

/* The zero start's first update x = 0 - alpha A^H d into dev.X and its
 * threshold lambda0 = max |x| (agdeblend.c:584-590, model.c:1202-1227). */
/* lambda0 = max |alpha A^H d| (the zero start's first update); keep_x: also
 * leave that update -alpha A^H d in dev.X (else dev.X is left to be
 * overwritten: a warm start or a given initial model) */
static int lambda0_update(AgdProblem *p, float alpha, float *lamb, int keep_x) {
  unsigned int bits = 0;
  float l2;
  float *tmp;
  int rc;
  if (agdg_zero(p->dev.scalar, 64)) return 1;
  /* the adjoint reads its record through dev.r: point it at d */
  tmp = p->dev.r;
  p->dev.r = p->dev.d;
  rc = adjoint_update(p, alpha, 0.0f, keep_x ? 1 : 4);
  p->dev.r = tmp;
  if (rc || agdg_d2h(&bits, p->dev.scalar, sizeof(bits))) return 1;
  memcpy(&l2, &bits, sizeof(l2));
  *lamb = sqrtf(l2);
  return 0;
}

static int run_impl(AgdProblem *p, float f, int n_its, int print_freq, float const *x0,
                    float const *const *warm, float const *warm_dev) {
  float const decay = expf(logf(0.001f / f) / (float)n_its);
  int const first_linear_it =
      (int)((logf(decay) * ((float)n_its - 1.0f) + 1.0f) / logf(decay));
  float slope = f * powf(decay, (float)first_linear_it) /
                (float)(n_its - 1 - first_linear_it);
  float const alpha = p->stepsize;
  float lamb;
  int it;
  int const warm_any = x0 == NULL && (warm != NULL || warm_dev != NULL);
  int zero_start = x0 == NULL && !warm_any;
  int r_ready = 0;
  double r2_start = -1.0;
  if (!p->have_data) {
    fprintf(stderr, "ERROR: no data set\n");
    return 1;
  }
  g_last_warm_fallback = 0;
  g_last_warm_ratio = -1.0;
  if (warm_any && n_its <= 1) {
    /* a warm start with no iteration after it uses no threshold: skip the
     * first update that only serves to compute lambda0 */
    lamb = 0.0f;
  } else {
    /* x = 0 - alpha * A^H d; lambda0 = max |x| (for a warm start or x0 the
     * schedule keeps the scale of the zero start) */
    if (lambda0_update(p, alpha, &lamb, zero_start)) return 1;
    if (x0 != NULL && agdg_h2d(p->dev.X, x0, (size_t)agdx_model_len(p) * 8)) return 1;
  }
  if (warm_any) {
    double d2 = 0.0;
    if (warm != NULL ? warm_model(p, warm) : warm_model_flat(p, warm_dev)) return 1;
    /* guard: the warm model must fit the data better than the zero model
     * and better than the data itself does as a prediction of the record,
     * ||A x_warm - d|| < min(||d||, ||mute(sum blend(data)) - d||); its
     * residual then serves the first iteration. Otherwise the plain
     * zero-start inversion runs instead. */
    if (forward_record(p, 1) || agdg_norm2(&p->dev, p->dev.r, p->cfg.rec_len, &r2_start) ||
        agdg_norm2(&p->dev, p->dev.d, p->cfg.rec_len, &d2))
      return 1;
    g_last_warm_ratio = d2 > 0.0 ? sqrt(r2_start / d2) : -1.0;
    if (r2_start < d2 && r2_start < p->input_fit2) {
      r_ready = 1;
    } else {
      printf("WARNING: the warm start fits the data no better than the zero start or the "
             "data itself (residual %g, zero start %g, data %g); starting from zero\n",
             sqrt(r2_start), sqrt(d2), sqrt(p->input_fit2));
      g_last_warm_fallback = 1;
      r2_start = -1.0;
      zero_start = 1;
      if (lambda0_update(p, alpha, &lamb, 1)) return 1;
    }
  }
  g_last_lambda0 = lamb;
  slope *= lamb;
  lamb *= f;
  if (zero_start) {
    if (agdg_shrink(&p->dev, agdx_model_len(p), lamb)) return 1;
    set_u_valid(p, 0);
  } else if (x0 != NULL) {
    set_u_valid(p, 0);
  }
  if (print_freq > 0) {
    double r2 = r2_start;
    if (r2 < 0.0 && agdg_norm2(&p->dev, p->dev.d, p->cfg.rec_len, &r2)) return 1;
    printf("Iteration 0 Residual %f\n", sqrt(r2));
  }
  {
    /* thresholds of iterations 1..n_its-1, in the reference's float
     * arithmetic (agdeblend.c:623-627) */
    float *sched = (float *)malloc((size_t)(n_its > 1 ? n_its : 1) * sizeof(float));
    int rc = 0;
    if (sched == NULL) return 1;
    for (it = 1; it < n_its; it++) {
      if (it < first_linear_it) lamb = lamb * decay;
      else lamb = (float)(n_its - 1 - it) * slope;
      sched[it] = lamb;
    }
    rc = run_loop(p, alpha, sched, n_its, print_freq, r_ready);
    free(sched);
    if (rc) return 1;
  }
  return agdg_sync();
}

/* Spatial DFT precision of the problem's operators: 0 the FP32 register
 * FFTs (default), 1 split-precision bf16 x 3 matrix-core DFTs for the
 * windows they serve (square P in {20, 40}); the others keep the FP32
 * path. Returns the number of window classes the bf16 x 3 path serves, or
 * -1 on error. The split-precision path has no fused adjoint. */
EXPORT int agdx_set_spatial_precision(AgdProblem *p, int mode) {
  if (p->host_only || (mode != 0 && mode != 1)) return -1;
  p->dev.precision = mode;
  p->fuse = agdg_can_fuse(&p->dev) && getenv("AGD_NO_FUSE") == NULL;
  set_u_valid(p, 0);
  return mode ? agdg_x3_classes(&p->dev) : 0;
}

/* Rebuild the spatial tasks with the planes per task in p->nk_req and
 * replace the device task list. */
static int rebuild_spatial(AgdProblem *p) {
  vec tasks;
  size_t lds_sp = 0;
  int c, i;
  void *q, *old = (void *)p->dev.tasks;
  vec_init(&tasks, sizeof(AgdTask));
  for (c = 0; c < p->n_sclass; c++) {
    free((void *)p->sclass[c].wpos);
    p->sclass[c].wpos = NULL;
  }
  if (spatial_tasks(p, p->sclass, p->n_sclass, p->nk_req, &tasks, &lds_sp) ||
      lds_sp > LDS_MAX) {
    vec_free(&tasks);
    return 1;
  }
  q = agdg_alloc(tasks.n ? tasks.n * sizeof(AgdTask) : 16);
  if (q == NULL || (tasks.n && agdg_h2d(q, tasks.p, tasks.n * sizeof(AgdTask)))) {
    agdg_free(q);
    vec_free(&tasks);
    return 1;
  }
  for (i = 0; i < p->n_dalloc; i++)
    if (p->dalloc[i] == old) p->dalloc[i] = q;
  agdg_free(old);
  p->dev.tasks = (AgdTask const *)q;
  p->dev.n_tasks = (int)tasks.n;
  p->dev.lds_spatial = lds_sp;
  set_u_valid(p, 0);
  vec_free(&tasks);
  return 0;
}

/* wall time of reps launches of one iteration's spatial stage (the fused
 * adjoint with alpha = 0 and threshold 0 on a zero V, which leaves X as it
 * is, or the adjoint and forward spatial kernels) */
static double time_spatial(AgdProblem *p, int reps) {
  struct timespec t0, t1;
  float *lamb_dev = p->dev.lamb_dev;
  int i, rc = 0;
  p->dev.lamb_dev = NULL;
  for (i = -1; i < reps && rc == 0; i++) {
    if (i == 0) {
      if (agdg_sync()) rc = 1;
      clock_gettime(CLOCK_MONOTONIC, &t0);
    }
    if (agdg_adj_spatial(&p->dev, 0.0f, 0.0f, p->fuse ? 0 : 2, p->fuse)) rc = 1;
    if (!p->fuse && agdg_fwd_spatial(&p->dev, p->dev.X)) rc = 1;
  }
  if (agdg_sync()) rc = 1;
  clock_gettime(CLOCK_MONOTONIC, &t1);
  p->dev.lamb_dev = lamb_dev;
  if (rc) return -1.0;
  return ((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec)) / reps;
}

/* Mean seconds of one iteration's time stage (forward and adjoint time
 * transforms of every window) over reps launches. Overwrites T and V only
 * (intermediates that every iteration recomputes). */
static double time_time(AgdProblem *p, int reps) {
  struct timespec t0, t1;
  int i, rc = 0;
  for (i = -1; i < reps && rc == 0; i++) {
    if (i == 0) {
      if (agdg_sync()) rc = 1;
      clock_gettime(CLOCK_MONOTONIC, &t0);
    }
    if (agdg_fwd_time(&p->dev) ||
        agdg_adj_time(&p->dev, p->use_wavelets ? p->dev.T2 : NULL, 0))
      rc = 1;
  }
  if (agdg_sync()) rc = 1;
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (rc) return -1.0;
  return ((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec)) / reps;
}

/* Radix choice of the time transforms: for each specialised time class,
 * every (M1, M2) factorization of its half length with a kernel is timed
 * on one iteration's time stage and the fastest kept. */
static int autotune_time(AgdProblem *p, int reps) {
  int c, k;
  p->t_valid = 0; /* the timed launches overwrite T */
  for (c = 0; c < p->n_tclass; c++) {
    AgdClass *cl = &p->tclass[c];
    int const M = cl->p0 * cl->p1;
    int b0 = cl->p0, b1 = cl->p1, bv = cl->var, v;
    double best_t = -1.0;
    if (cl->p0 == 0) continue;
    /* untimed pass first: the first candidate timed on a cold device
     * (clocks, caches, code) would read slow */
    if (time_time(p, reps) < 0.0) return 1;
    for (k = 0; k < (int)(sizeof(kTimeCands) / sizeof(kTimeCands[0])); k++) {
      if (kTimeCands[k][0] * kTimeCands[k][1] != M) continue;
      for (v = 0; v < 2; v++) {
        double t;
        cl->p0 = kTimeCands[k][0];
        cl->p1 = kTimeCands[k][1];
        cl->var = v;
        t = time_time(p, reps);
        if (t < 0.0) return 1;
        if (getenv("AGD_TUNE_VERBOSE"))
          fprintf(stderr, "autotune: time class M=%d (%d x %d, variant %d): %.4f ms\n", M,
                  cl->p0, cl->p1, v, t * 1e3);
        if (best_t < 0.0 || t < best_t) {
          best_t = t;
          b0 = cl->p0;
          b1 = cl->p1;
          bv = v;
        }
      }
    }
    cl->p0 = b0;
    cl->p1 = b1;
    cl->var = bv;
  }
  return 0;
}

EXPORT int agdx_n_time_classes(AgdProblem const *p) { return p->n_tclass; }

EXPORT int agdx_set_time_factors(AgdProblem *p, int c, int m1, int m2) {
  int k;
  if (c < 0 || c >= p->n_tclass || p->tclass[c].p0 == 0 ||
      m1 * m2 != p->tclass[c].p0 * p->tclass[c].p1) {
    fprintf(stderr, "ERROR: no time class %d of half length %d\n", c, m1 * m2);
    return 1;
  }
  for (k = 0; k < (int)(sizeof(kTimeCands) / sizeof(kTimeCands[0])); k++)
    if (kTimeCands[k][0] == m1 && kTimeCands[k][1] == m2) {
      p->tclass[c].p0 = m1;
      p->tclass[c].p1 = m2;
      return 0;
    }
  fprintf(stderr, "ERROR: no time kernel for %d x %d\n", m1, m2);
  return 1;
}

EXPORT int agdx_set_time_variant(AgdProblem *p, int c, int var) {
  if (c < 0 || c >= p->n_tclass || p->tclass[c].p0 == 0 || (var != 0 && var != 1)) {
    fprintf(stderr, "ERROR: no specialised time class %d or variant %d\n", c, var);
    return 1;
  }
  p->tclass[c].var = var;
  return 0;
}

EXPORT int agdx_time_variants(AgdProblem const *p, int *out) {
  int c;
  if (out)
    for (c = 0; c < p->n_tclass; c++) out[c] = p->tclass[c].var;
  return p->n_tclass;
}

/* the time blocks (window, first trace, time windows [tw0, tw1)), TB_INTS
 * ints each, into out when given; returns their number */
EXPORT int agdx_time_blocks(AgdProblem const *p, int *out) {
  if (out) memcpy(out, p->tblk, (size_t)p->n_tblk * TB_INTS * sizeof(int));
  return p->n_tblk;
}

EXPORT int agdx_time_factors(AgdProblem const *p, int *out) {
  int c;
  if (out)
    for (c = 0; c < p->n_tclass; c++) {
      out[2 * c] = p->tclass[c].p0;
      out[2 * c + 1] = p->tclass[c].p1;
    }
  return p->n_tclass;
}

/* Autotuner: first the radix choice of the time transforms
 * (autotune_time), then the spatial kernels' tile: for each specialised
 * spatial class in turn (the others at their current choice), time one
 * iteration's spatial stage for every candidate planes-per-task count that
 * fits the LDS, and keep the fastest. Overwrites the intermediates T and V
 * (V is left zeroed), leaves the tasks set to the choice; the model X is
 * unchanged. nk_out (n_sclass ints, or NULL)
 * receives the choice per class (0 for generic classes); ms_out (same) the
 * time of the chosen configuration. Returns 0, or 1 on error. */
EXPORT int agdx_autotune(AgdProblem *p, int reps, int *nk_out, double *ms_out) {
  static int const cand[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32};
  int c, k;
  if (p->host_only) return 1;
  if (reps < 1) reps = 3;
  if (autotune_time(p, reps)) return 1;
  if (p->nk_req == NULL) {
    p->nk_req = (int *)calloc((size_t)(p->n_sclass > 0 ? p->n_sclass : 1), sizeof(int));
    if (p->nk_req == NULL) return 1;
  }
  if (agdg_zero(p->dev.V, p->u_len * 8)) return 1;
  for (c = 0; c < p->n_sclass; c++) {
    AgdClass const *cl = &p->sclass[c];
    int best = 0, maxft = 1, w;
    double best_t = 0.0;
    if (cl->p0 == 0) {
      if (nk_out) nk_out[c] = 0;
      continue;
    }
    for (w = 0; w < p->n_win; w++) {
      int a, b;
      spatial_key(&p->win[w], &a, &b);
      if (a == cl->p0 && b == cl->p1 && p->win[w].Ft > maxft) maxft = p->win[w].Ft;
    }
    for (k = 0; k < (int)(sizeof(cand) / sizeof(cand[0])); k++) {
      size_t const plane = (size_t)spec_len(cl->p0) * (cl->p1 + 1) * 8;
      double t;
      if ((size_t)cand[k] * plane > 96 * 1024 || cand[k] > maxft) continue;
      p->nk_req[c] = cand[k];
      if (rebuild_spatial(p)) return 1;
      t = time_spatial(p, reps);
      if (t < 0.0) return 1;
      if (getenv("AGD_TUNE_VERBOSE"))
        fprintf(stderr, "autotune: spatial class %dx%d nk %2d: %.4f ms\n", cl->p0, cl->p1,
                cand[k], t * 1e3);
      if (best == 0 || t < best_t) {
        best = cand[k];
        best_t = t;
      }
    }
    p->nk_req[c] = best;
    if (rebuild_spatial(p)) return 1;
    if (nk_out) nk_out[c] = best;
    if (ms_out) ms_out[c] = best_t * 1e3;
  }
  return 0;
}

EXPORT int agdx_n_spatial_classes(AgdProblem const *p) { return p->n_sclass; }

EXPORT int agdx_iterate(AgdProblem *p, int n, float lamb) {
  float *sched;
  int it, rc;
  if (n <= 0) return agdg_sync();
  sched = (float *)malloc(((size_t)n + 1) * sizeof(float));
  if (sched == NULL) return 1;
  for (it = 0; it <= n; it++) sched[it] = lamb;
  rc = run_loop(p, p->stepsize, sched, n + 1, 0, 0);
  free(sched);
  if (rc) return 1;
  return agdg_sync();
}

/* set_output on the device: forward of the current model assembled into
 * the device copy of the patch data (no host transfer) */
EXPORT int agdx_finish(AgdProblem *p) {
  if (!p->u_valid) {
    if (agdg_fwd_spatial(&p->dev, p->dev.X)) return 1;
    set_u_valid(p, 1);
  }
  /* T is reused when the last forward was of the current model (e.g. a
   * warm start's guard with no iteration after it) */
  if (!p->t_valid) {
    if (agdg_fwd_time(&p->dev)) return 1;
    if (p->use_wavelets && agdg_wavelet(&p->dev, 0)) return 1;
    p->t_valid = !p->use_wavelets;
  }
  if (agdg_output(p->dev.T, p->n_out, p->d_o_off, p->d_o_nt, p->d_o_ptr, p->d_o_src,
                  p->d_data))
    return 1;
  return agdg_sync();
}

EXPORT int agdx_get_output(AgdProblem *p, float *const *data) {
  int a;
  if (agdx_finish(p)) return 1;
  for (a = 0; a < p->n_patches; a++) {
    size_t const n = (size_t)p->ntr[a] * p->nt[a];
    if (agdg_d2h(data[a], p->d_data + p->data_off[a], n * 4)) return 1;
  }
  return 0;
}

EXPORT int agdx_timing(int enable) { return agdg_timing(enable); }

EXPORT int agdx_timing_get(char const *kernel, double *ms, long *count) {
  return agdg_timing_get(kernel, ms, count);
}

/* ------------------------------------------------------------------------ */
/* Public entry points                                                       */
/* ------------------------------------------------------------------------ */

EXPORT int agd_deblend(int n_patches, int const *volumes, int const *n_dims,
                       int const *const *window_shapes, int const *const *coords,
                       int const *const *shapes, long int const *const *shottimes,
                       int const *const *channels,
                       enum AGDTraceType const *const *trace_types,
                       int const *wavelet_lengths, int const *const *wavelet_idxs,
                       float const *const *wavelets, float const initial_factor,
                       int n_its, int print_freq, float *const *data) {
  AgdProblem *p;
  float const step_warn = 0.3f;
  if (check_deblend_inputs(n_patches, volumes, n_dims, window_shapes, coords, shapes,
                           shottimes, channels, trace_types, wavelet_lengths,
                           wavelet_idxs, wavelets, initial_factor, n_its, data))
    goto err;
  p = agdx_create(n_patches, volumes, n_dims, window_shapes, coords, shapes, shottimes,
                  channels, trace_types, wavelet_lengths, wavelet_idxs, wavelets);
  if (p == NULL) goto err;
  if (p->stepsize > step_warn) {
    printf("WARNING: There does not appear to be any overlap between shots\n");
  }
  if (agdx_set_data(p, (float const *const *)data, NULL) ||
      agdx_run(p, initial_factor, n_its, print_freq, NULL) ||
      agdx_get_output(p, data)) {
    agdx_destroy(p);
    goto err;
  }
  agdx_destroy(p);
  return 0;
err:
  fprintf(stderr, "ERROR in agd_deblend\n");
  return 1;
}

EXPORT int agdeblend_deblend(int n_patches, int const *volumes, int const *n_dims,
                             int const *const *window_shapes,
                             int const *const *coords, int const *const *shapes,
                             long int const *const *shottimes,
                             int const *const *channels,
                             enum AGDTraceType const *const *trace_types,
                             int const *wavelet_lengths,
                             int const *const *wavelet_idxs,
                             float const *const *wavelets,
                             float const initial_factor, int n_its, int print_freq,
                             float *const *data) {
  return agd_deblend(n_patches, volumes, n_dims, window_shapes, coords, shapes,
                     shottimes, channels, trace_types, wavelet_lengths, wavelet_idxs,
                     wavelets, initial_factor, n_its, print_freq, data);
}

/* agdeblend.c:651-732 */
static int check_blend_inputs(int np, int const *ntr, int const *nt,
                              long int const *const *shot, int const *const *ch,
                              enum AGDTraceType const *const *tt,
                              float *const *data, enum AGDBlendMode mode, int taper,
                              int npo, int const *ntro, int const *nto,
                              long int const *const *shoto, int const *const *cho,
                              enum AGDTraceType const *const *tto,
                              float *const *datao) {
  int p, t;
  if (check_positive(np, "n_patches")) goto err;
  /* the reference checks n_patches a second time under this name
   * (agdeblend.c:667) */
  if (check_positive(np, "n_patches_out")) goto err;
  if (check_non_null(ntr, "n_traces_per_volume")) goto err;
  if (check_non_null(ntro, "n_traces_per_volume_out")) goto err;
  if (check_non_null(nt, "n_times_per_volume")) goto err;
  if (check_non_null(nto, "n_times_per_volume_out")) goto err;
  if (check_non_null(shot, "shottimes")) goto err;
  if (check_non_null(shoto, "shottimes_out")) goto err;
  if (check_non_null(ch, "channels")) goto err;
  if (check_non_null(cho, "channels_out")) goto err;
  if (check_non_null(tt, "trace_types")) goto err;
  if (check_non_null(tto, "trace_types_out")) goto err;
  if (check_non_null(data, "data")) goto err;
  if (check_non_null(datao, "data_out")) goto err;
  for (p = 0; p < np; p++) {
    if (check_positive(ntr[p], "n_traces_per_volume")) goto err;
    if (check_positive(nt[p], "n_times_per_volume")) goto err;
    if (check_non_null(shot[p], "shottimes")) goto err;
    if (check_non_null(ch[p], "channels")) goto err;
    if (check_non_null(tt[p], "trace_types")) goto err;
    if (check_non_null(data[p], "data")) goto err;
  }
  for (p = 0; p < npo; p++) {
    if (check_positive(ntro[p], "n_traces_per_volume_out")) goto err;
    if (check_positive(nto[p], "n_times_per_volume_out")) goto err;
    if (check_non_null(shoto[p], "shottimes_out")) goto err;
    if (check_non_null(cho[p], "channels_out")) goto err;
    if (check_non_null(tto[p], "trace_types_out")) goto err;
    if (check_non_null(datao[p], "data_out")) goto err;
    for (t = 0; t < ntro[p]; t++) {
      if (tto[p][t] == AGDBad) {
        fprintf(stderr, "ERROR: Output trace types may not be AGDBad\n");
        goto err;
      }
    }
  }
  if (mode == AGDBlendMean && taper < 0) {
    fprintf(stderr,
            "ERROR: taper_length must be >= 0 when blend_mode is AGDBlendMean\n");
    goto err;
  }
  return 0;
err:
  fprintf(stderr, "ERROR in agd_blend inputs\n");
  return 1;
}

/* Device buffers of one agd_blend call */
enum {
  BL_TILE_PTR,
  BL_TILE_ENT,
  BL_ENT_IN,
  BL_MUTE,
  BL_TAPER,
  BL_ENT_OUT,
  BL_TRACE_LEN,
  BL_IN,
  BL_OUT,
  BL_REC,
  BL_N
};

/* agdeblend.c:800-938. noalloc = 0: the device buffers are allocated for the
 * call and freed before it returns (agd_blend); noalloc = 1: they are carved
 * from the library's fixed scratch array (agd_blend_noalloc), and a call
 * whose buffers do not fit is refused before any device work. */
static int blend_impl(int np, int const *ntr, int const *nt, long int const *const *shot,
                      int const *const *ch, enum AGDTraceType const *const *tt,
                      float *const *data, enum AGDBlendMode mode, int taper, int npo,
                      int const *ntro, int const *nto, long int const *const *shoto,
                      int const *const *cho, enum AGDTraceType const *const *tto,
                      float *const *datao, int noalloc) {
  BlendCfg cfg;
  TraceCoord *ci = NULL, *co = NULL;
  AgdEntry *ei = NULL, *eo = NULL;
  int *tl = NULL, *tp = NULL, *te = NULL;
  int64_t *doff = NULL, *ooff = NULL;
  int64_t din = 0, dout = 0;
  int nin = 0, nout = 0, nt_tiles = 0;
  int p, t, g, m, b;
  float *tap = NULL;
  unsigned char *mute = NULL;
  size_t sz[BL_N];
  void *buf[BL_N];
  void const *host[BL_N];
  int owned = 0;
  float *d_in, *d_out, *d_rec;
  AgdGather gt;
  int rc = 1;
  memset(&cfg, 0, sizeof(cfg));
  memset(buf, 0, sizeof(buf));
  if (check_blend_inputs(np, ntr, nt, shot, ch, tt, data, mode, taper, npo, ntro, nto,
                         shoto, cho, tto, datao))
    goto fail;
  if (agdg_init()) goto fail;
  if (blendcfg_build(np, ntr, nt, shot, ch, tt, &cfg)) goto fail;
  for (p = 0; p < np; p++) nin += ntr[p];
  for (p = 0; p < npo; p++) nout += ntro[p];
  ci = (TraceCoord *)malloc(((size_t)nin + 1) * sizeof(TraceCoord));
  co = (TraceCoord *)malloc(((size_t)nout + 1) * sizeof(TraceCoord));
  ei = (AgdEntry *)malloc(((size_t)nin + 1) * sizeof(AgdEntry));
  eo = (AgdEntry *)malloc(((size_t)nout + 1) * sizeof(AgdEntry));
  tl = (int *)malloc(((size_t)nout + 1) * sizeof(int));
  doff = (int64_t *)malloc(((size_t)np + 1) * sizeof(int64_t));
  ooff = (int64_t *)malloc(((size_t)npo + 1) * sizeof(int64_t));
  if (!ci || !co || !ei || !eo || !tl || !doff || !ooff) goto fail;
  g = 0;
  for (p = 0; p < np; p++) {
    doff[p] = din;
    for (t = 0; t < ntr[p]; t++, g++) {
      trace_coord(&cfg, shot[p][t], nt[p], ch[p][t], tt[p][t], &ci[g]);
      ei[g].rec = ci[g].rec;
      ei[g].src = din + (int64_t)t * nt[p];
      ei[g].trace_start = ci[g].trace_start;
      ei[g].n = ci[g].n;
    }
    din += (int64_t)ntr[p] * nt[p];
  }
  g = 0;
  for (p = 0; p < npo; p++) {
    ooff[p] = dout;
    for (t = 0; t < ntro[p]; t++, g++) {
      trace_coord(&cfg, shoto[p][t], nto[p], cho[p][t], tto[p][t], &co[g]);
      eo[g].rec = co[g].rec;
      eo[g].src = dout + (int64_t)t * nto[p];
      eo[g].trace_start = co[g].trace_start;
      eo[g].n = co[g].n;
      tl[g] = nto[p];
    }
    dout += (int64_t)ntro[p] * nto[p];
  }
  if (build_tiles(cfg.rec_len, GATHER_TILE, ci, nin, &tp, &te, &nt_tiles)) goto fail;
  mute = (unsigned char *)calloc((size_t)cfg.rec_len + 1, 1);
  if (mute == NULL) goto fail;
  for (m = 0; m < cfg.n_mute; m++) memset(mute + cfg.mute_start[m], 1, (size_t)cfg.mute_len[m]);
  if (mode == AGDBlendMean && taper > 1) {
    int i;
    tap = (float *)malloc((size_t)taper * sizeof(float));
    if (tap == NULL) goto fail;
    for (i = 0; i < taper; i++)
      tap[i] = cosf(0.5f * (float)M_PI * (float)(taper - i) / (float)(taper + 1));
  }
  /* sizes and host sources of the device buffers (NULL: written on the
   * device; every element of those is written by the kernels) */
  sz[BL_TILE_PTR] = ((size_t)nt_tiles + 1) * sizeof(int);
  sz[BL_TILE_ENT] = (size_t)tp[nt_tiles] * sizeof(int);
  sz[BL_ENT_IN] = (size_t)nin * sizeof(AgdEntry);
  sz[BL_MUTE] = (size_t)cfg.rec_len;
  sz[BL_TAPER] = tap ? (size_t)taper * sizeof(float) : 0;
  sz[BL_ENT_OUT] = (size_t)nout * sizeof(AgdEntry);
  sz[BL_TRACE_LEN] = (size_t)nout * sizeof(int);
  sz[BL_IN] = (size_t)din * 4;
  sz[BL_OUT] = (size_t)dout * 4;
  sz[BL_REC] = (size_t)cfg.rec_len * 4;
  host[BL_TILE_PTR] = tp;
  host[BL_TILE_ENT] = te;
  host[BL_ENT_IN] = ei;
  host[BL_MUTE] = mute;
  host[BL_TAPER] = tap;
  host[BL_ENT_OUT] = eo;
  host[BL_TRACE_LEN] = tl;
  host[BL_IN] = NULL;
  host[BL_OUT] = NULL;
  host[BL_REC] = NULL;
  if (noalloc) {
    void *base = NULL;
    size_t cap = 0, off = 0;
    if (agdg_scratch(&base, &cap)) goto fail;
    for (b = 0; b < BL_N; b++) {
      buf[b] = (char *)base + off;
      off += (sz[b] + 255) & ~(size_t)255;
    }
    if (off > cap) {
      fprintf(stderr,
              "ERROR: agd_blend_noalloc needs %zu bytes of device scratch, more than its "
              "fixed %zu\n",
              off, cap);
      goto fail;
    }
  } else {
    owned = 1;
    for (b = 0; b < BL_N; b++)
      if ((buf[b] = agdg_alloc(sz[b] > 0 ? sz[b] : 16)) == NULL) goto fail;
  }
  for (b = 0; b < BL_N; b++)
    if (host[b] != NULL && agdg_h2d(buf[b], host[b], sz[b])) goto fail;
  gt.rec_len = cfg.rec_len;
  gt.tile = GATHER_TILE;
  gt.n_tiles = nt_tiles;
  gt.tile_ptr = (int const *)buf[BL_TILE_PTR];
  gt.tile_ent = (int const *)buf[BL_TILE_ENT];
  gt.ent = (AgdEntry const *)buf[BL_ENT_IN];
  gt.mute = (unsigned char const *)buf[BL_MUTE];
  gt.taper = (float const *)buf[BL_TAPER];
  gt.taper_len = taper;
  d_in = (float *)buf[BL_IN];
  d_out = (float *)buf[BL_OUT];
  d_rec = (float *)buf[BL_REC];
  for (p = 0; p < np; p++)
    if (agdg_h2d(d_in + doff[p], data[p], (size_t)ntr[p] * nt[p] * 4)) goto fail;
  if (agdg_gather(&gt, d_in, mode == AGDBlendSum ? 0 : (mode == AGDBlendMean ? 1 : 2), d_rec))
    goto fail;
  if (agdg_extract((AgdEntry const *)buf[BL_ENT_OUT], nout, (int const *)buf[BL_TRACE_LEN],
                   d_rec, d_out))
    goto fail;
  if (agdg_sync()) goto fail;
  for (p = 0; p < npo; p++)
    if (agdg_d2h(datao[p], d_out + ooff[p], (size_t)ntro[p] * nto[p] * 4)) goto fail;
  rc = 0;
fail:
  if (owned)
    for (b = 0; b < BL_N; b++) agdg_free(buf[b]);
  free(ci);
  free(co);
  free(ei);
  free(eo);
  free(tl);
  free(tp);
  free(te);
  free(doff);
  free(ooff);
  free(tap);
  free(mute);
  blendcfg_free(&cfg);
  if (rc) fprintf(stderr, noalloc ? "ERROR in agd_blend_noalloc\n" : "ERROR in agd_blend\n");
  return rc;
}

EXPORT int agd_blend(int np, int const *ntr, int const *nt, long int const *const *shot,
                     int const *const *ch, enum AGDTraceType const *const *tt,
                     float *const *data, enum AGDBlendMode mode, int taper, int npo,
                     int const *ntro, int const *nto, long int const *const *shoto,
                     int const *const *cho, enum AGDTraceType const *const *tto,
                     float *const *datao) {
  return blend_impl(np, ntr, nt, shot, ch, tt, data, mode, taper, npo, ntro, nto, shoto, cho,
                    tto, datao, 0);
}

EXPORT int agd_blend_noalloc(int np, int const *ntr, int const *nt,
                             long int const *const *shot, int const *const *ch,
                             enum AGDTraceType const *const *tt, float *const *data,
                             enum AGDBlendMode mode, int taper, int npo, int const *ntro,
                             int const *nto, long int const *const *shoto,
                             int const *const *cho, enum AGDTraceType const *const *tto,
                             float *const *datao) {
  return blend_impl(np, ntr, nt, shot, ch, tt, data, mode, taper, npo, ntro, nto, shoto, cho,
                    tto, datao, 1);
}

EXPORT int agdeblend_blend(int np, int const *ntr, int const *nt,
                           long int const *const *shot, int const *const *ch,
                           enum AGDTraceType const *const *tt, float *const *data,
                           enum AGDBlendMode mode, int taper, int npo,
                           int const *ntro, int const *nto,
                           long int const *const *shoto, int const *const *cho,
                           enum AGDTraceType const *const *tto, float *const *datao) {
  return agd_blend(np, ntr, nt, shot, ch, tt, data, mode, taper, npo, ntro, nto, shoto,
                   cho, tto, datao);
}
