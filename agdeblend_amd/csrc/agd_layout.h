/*
 * Device-side tables shared by the C host (agd_host.c) and the HIP kernels
 * (agd_kernels.hip). Plain C structs only.
 *
 * HBM layout of one deblending problem (all offsets in elements):
 *   model X   complex float, per window [ntw][Ft][S]  (frequency-major: each
 *             (time window, time frequency) plane holds the S = prod(P_space)
 *             spatial coefficients contiguously, row-major, last space dim
 *             fastest). Mathematically the same coefficients as the
 *             reference's [ntw][P_space...][Ft] (fk.c:256-262), transposed.
 *   inter U   complex float, per window [ntw][ntr/ug][Ft][ug]: spatial
 *             transform of X cropped to the window's traces (forward), or
 *             time transform of the traces (adjoint); ntr = prod(W_space).
 *             ug = 16 (tiles of 16 traces, one contiguous [Ft][16] block per
 *             time window) for windows served by the specialised kernels,
 *             else ug = ntr ([ntw][Ft][ntr]).
 *   traces T  float, per window [ntr][nconv] (the reference's temp_1 trace
 *             buffer, fk.c:330-331).
 *   record    float, all (channel, interval) continuous records concatenated
 *             in channel-then-interval order (blend.c:796-829).
 */
#ifndef AGD_LAYOUT_H
#define AGD_LAYOUT_H

#include <stdint.h>

#define AGD_MAXSP 4      /* AGD_MAX_N_COORDS */
#define AGD_MAXSTAGE 24  /* FFT stages per plan */
/* half lengths Mr / 2 of the wavelet convolution with a compile-time kernel
 * (k_wavelet_t); the host prefers them to the smallest 7-smooth length when
 * one is at most 25% longer */
#define AGD_WAVELET_HALF_LENGTHS {512, 768, 1024, 1536, 2048, 3072, 4096, 6144}

/* One FFT length with its radix schedule and roots table
 * roots[m] = exp(-2 pi i m / n), m < n. */
typedef struct {
  int n;
  int nstage;
  int radix[AGD_MAXSTAGE];
  int roots_off; /* complex offset in the roots pool */
  int pad;
} AgdFftPlan;

typedef struct {
  int nsp;               /* space dims */
  int P[AGD_MAXSP];      /* padded space shape */
  int W[AGD_MAXSP];      /* window space shape */
  int S;                 /* prod P */
  int ntr;               /* prod W (traces) */
  int Pt, Ft, Wt, half, ntw, nfk, nconv;
  int tt_off;            /* time taper (Wt floats) in the taper pool */
  int tplan;             /* plan index for the time transform (Pt, or Pt/2 if even) */
  int tplan_full;        /* plan index holding roots of Pt (even Pt packing) */
  int splan[AGD_MAXSP];  /* plan index per space dim */
  int wplan;             /* wavelet convolution FFT plan (-1: none) */
  int wplan_full;        /* plan holding roots of the wavelet length M */
  int wM;                /* wavelet convolution length M */
  int trace0;            /* first global window-trace index */
  int patch;
  int ug;                /* trace group of the U / V layout (16 or ntr) */
  float scaler;          /* 1 / sqrt(prod P) (fk.c:157-163) */
  int spare;
  int64_t x_off;         /* model offset (complex) */
  int64_t u_off;         /* intermediate offset (complex) */
  int64_t t_off;         /* trace buffer offset (float) */
} AgdWin;

/* Blend coordinates of one window trace (or one user-data trace), in flat
 * record terms: samples [trace_start, trace_start + n) of the trace map to
 * record samples [rec, rec + n). n == 0: trace not blended (channel -1). */
typedef struct {
  int64_t rec;
  int64_t src;        /* flat offset of the trace's sample 0 in its buffer */
  int trace_start;
  int n;
} AgdEntry;

/* A contiguous range of spatial tasks (or time blocks) whose windows share
 * the transform sizes, so one specialised kernel instance serves them.
 * p0 == 0: no specialised kernel, the generic one runs the range. */
typedef struct {
  int p0, p1;   /* spatial: padded sizes (p0 = 1 when one space dim);
                   time: p0 = M1, p1 = M2 with M1*M2 = Pt/2 */
  int begin, count;
  int tb;       /* time classes: traces per block; specialised spatial
                   classes: 1 when U / V are tiled by 16 traces */
  int var;      /* time classes: forward kernel variant (0 register
                   prefetch, 1 lean tile); spatial classes: unused */
  size_t lds;
  int const *wpos; /* host array [n_win + 1]: first unit of the class that
                      belongs to window w or later (units are in window
                      order), relative to begin */
} AgdClass;

/* A unit of work for the spatial (per time-frequency chunk) kernels */
typedef struct {
  int win;
  int tw;
  int k0;
  int nk;
} AgdTask;

#endif
