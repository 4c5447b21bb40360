/*
 * Thin C-ABI between the C host (agd_host.c) and the HIP kernels
 * (agd_kernels.hip). Plain pointers and sizes only; device pointers are
 * opaque to the host code. All launches go to the library's own stream.
 * Every function returns 0 on success and 1 on a HIP error (message on
 * stderr).
 */
#ifndef AGD_GPU_H
#define AGD_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "agd_layout.h"

/* ints per time block: window, first trace, first and end time window */
#define TB_INTS 4

#ifdef __cplusplus
extern "C" {
#endif

/* Device-resident tables and buffers of one deblending problem. Filled by
 * the host; kernels read it by value. */
typedef struct {
  /* geometry */
  int n_win;
  AgdWin const *win;          /* [n_win] */
  AgdFftPlan const *plans;    /* FFT plans */
  float const *roots;         /* complex roots pool (interleaved) */
  float const *tapers;        /* time taper pool */
  float const *tscale;        /* per global window trace: scaler * space tapers */
  float const *tconst;        /* per global window trace: space tapers only */
  AgdEntry const *went;       /* per global window trace blend entry */
  int const *trace_win;       /* per global window trace: its window */
  int n_wtr;                  /* global window traces */
  /* spatial tasks, grouped in classes of equal sizes */
  AgdTask const *tasks;
  int n_tasks;
  size_t lds_spatial;         /* bytes (generic kernel) */
  AgdClass const *sclass;     /* host array */
  int n_sclass;
  /* time blocks: (window, first trace, time windows [tw0, tw1)) */
  int const *tblocks;         /* TB_INTS ints per block */
  int n_tblocks;
  int const *tadj;            /* [n_tblocks]: per class, its blocks sorted by
                                 their first trace (the adjoint time kernel's
                                 launch order), or NULL */
  int tb;                     /* traces per time block */
  size_t lds_time;            /* bytes (generic kernel) */
  AgdClass const *tclass;     /* host array */
  int n_tclass;
  /* record gather of window traces */
  int64_t rec_len;
  int tile;
  int n_tiles;
  int const *tile_ptr;        /* [n_tiles+1] */
  int const *tile_ent;        /* entry ids (global window traces) */
  unsigned char const *mute;  /* [rec_len] 1 = muted */
  /* wavelets */
  int use_wavelets;
  float const *wspec;         /* complex spectra pool */
  int const *wsrc_off;        /* per global window trace: complex offset of
                                 its spectrum in wspec, or -1 */
  size_t lds_wavelet;
  /* buffers */
  float *X;                   /* model, complex */
  float *U;                   /* forward intermediate, complex */
  float *V;                   /* adjoint intermediate, complex */
  float *T;                   /* traces */
  float *T2;                  /* traces after wavelet adjoint */
  float *d;                   /* blended input record */
  float *r;                   /* forward / residual record */
  unsigned int *scalar;       /* device scalars (max bits etc.) */
  /* window range [w0, w1) of the transform launches (all windows by
   * default); wtrace0 (host) holds each window's first global trace and
   * n_wtr at index n_win */
  int w0, w1;
  int const *wtrace0;
  /* device threshold of captured iterations: when lamb_dev is set, the
   * update kernels read their threshold from lamb_dev[0], which
   * agdg_next_lamb advances through lamb_sched */
  float *lamb_dev;
  float *lamb_sched;
  int *lamb_idx;
  double *partials;           /* reduction partials */
  int n_partials;
  int precision;              /* spatial DFTs: 0 FP32 register FFTs, 1 bf16 x 3 MFMA */
  int wavelet_mh;             /* half length of every window's wavelet transform when
                                 they share one (0 otherwise): a specialised kernel */
} AgdDev;

/* A gather of source traces into a flat record, or of record slices back
 * into traces (blend / pseudo-deblend). */
typedef struct {
  int64_t rec_len;
  int tile;
  int n_tiles;
  int const *tile_ptr;
  int const *tile_ent;
  AgdEntry const *ent;        /* per source trace */
  unsigned char const *mute;  /* may be NULL */
  float const *taper;         /* mean-mode weights [taper_len] */
  int taper_len;
} AgdGather;

int agdg_init(void);
/* The device address and size of the library's fixed scratch array on the
 * current device, which must be the library's device (the device current
 * at its first call). Used by the scratch-free entry point; no allocation. */
int agdg_scratch(void **ptr, size_t *bytes);
void *agdg_alloc(size_t bytes);
void agdg_free(void *p);
int agdg_h2d(void *dst, void const *src, size_t bytes);
int agdg_d2h(void *dst, void const *src, size_t bytes);
int agdg_zero(void *dst, size_t bytes);
int agdg_sync(void);
void *agdg_stream(void);

/* forward operator pieces */
int agdg_fwd_spatial(AgdDev const *g, float const *X);
int agdg_fwd_time(AgdDev const *g);
int agdg_wavelet(AgdDev const *g, int adjoint);
/* r = mute(blend(T)) - d (resid=1) or mute(blend(T)) (resid=0) */
int agdg_blend_traces(AgdDev const *g, int resid);
/* adjoint pieces */
/* V = time R2C of the windowed traces: of the residual record dev.r through
 * the blend entries dev.went (Tin = NULL; any flat buffer with entries in
 * the same terms, e.g. patch data) or of a trace buffer Tin ([ntr][nconv]
 * per window). plain = 0: the adjoint (scaler x space and time tapers, x2 on the
 * interior bins); plain = 1: scaler only (the warm-start model). */
int agdg_adj_time(AgdDev const *g, float const *Tin, int plain);
/* mode 0: X = X - alpha*G then shrink by lamb; mode 1: X = -alpha*G and
 * record max |X|^2 bits in scalar[0]; mode 2: X = X - alpha*G (no shrink);
 * mode 3: X = G; mode 4: as mode 1 without writing X where the kernels
 * allow it (X then holds the previous model or -alpha*G). fuse: also write the forward spatial transform of the new X
 * into U (the next iteration's first forward stage). */
int agdg_adj_spatial(AgdDev const *g, float alpha, float lamb, int mode, int fuse);
/* 1 when every spatial class has a specialised kernel (fusion available) */
int agdg_can_fuse(AgdDev const *g);
/* number of spatial classes the split-precision transforms can serve */
int agdg_x3_classes(AgdDev const *g);
int agdg_shrink(AgdDev const *g, int64_t n_complex, float lamb);
/* fold of spatial tapers for the stepsize; max written to scalar[1] */
int agdg_stepsize_fold(AgdDev const *g);
/* sum of squares of a float array in FP64 (deterministic order) */
int agdg_norm2(AgdDev const *g, float const *v, int64_t n, double *out);
/* output assembly: out[o_off[i] + t] = sum_c T[o_src[c] + t], t < o_nt[i] */
int agdg_output(float const *T, int n_out, int64_t const *o_off,
                int const *o_nt, int const *o_ptr, int64_t const *o_src,
                float *out);
/* gather sources into a record: mode 0 sum, 1 mean, 2 overwrite;
 * muted samples set to 0 when gt->mute is set */
int agdg_gather(AgdGather const *gt, float const *src, int mode, float *rec);
/* rec = mute(sum gather of src) - d */
int agdg_gather_resid(AgdGather const *gt, float const *src, float const *d, float *rec);
/* extract record slices into traces (blend adjoint, blend.c:1262-1292):
 * out[e.src + t] = rec[e.rec + t - e.trace_start] for t in
 * [trace_start, trace_start + n), 0 for other t < trace_len[i] */
int agdg_extract(AgdEntry const *ent, int n_ent, int const *trace_len,
                 float const *rec, float *out);
/* timing of the library's kernels (per kernel name accumulators) */
/* lamb_dev[0] = lamb_sched[*lamb_idx]; *lamb_idx += 1 */
int agdg_next_lamb(AgdDev const *g);
/* stream capture of the library stream into an executable graph */
int agdg_graph_begin(void);
int agdg_graph_end(void **exec);
int agdg_graph_launch(void *exec);
void agdg_graph_destroy(void *exec);
int agdg_timing_enabled(void);
int agdg_timing(int enable);
int agdg_timing_get(char const *name, double *ms, long *count);

#ifdef __cplusplus
}
#endif

#endif
