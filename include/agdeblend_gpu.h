/*
 * Lower-level C interface of libagdeblend.so: a deblending problem that
 * stays resident on the GPU between calls.
 *
 * agd_deblend (agdeblend.h) is exactly agdx_create + agdx_set_data +
 * agdx_run + agdx_get_output + agdx_destroy. The split form exists for
 * callers that reuse one geometry (benchmarks, warm starts computed by the
 * Python training side, operator tests). It has no counterpart in the
 * reference, whose internals are all static (agdeblend.c:241-245).
 *
 * Model vectors are complex float, interleaved (re, im), in the library's
 * frequency-major window layout: window w, time window tw, time frequency
 * k, then the padded space coefficients (row-major, last space dim
 * fastest). Records are the concatenated (channel, interval) records in
 * channel-then-interval order. Return codes: 0 success, 1 failure.
 */
#ifndef AGDEBLEND_GPU_H
#define AGDEBLEND_GPU_H

#include "agdeblend.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct AgdProblem AgdProblem;

/* Validates the inputs like agd_deblend and builds the device tables.
 * Returns NULL on failure. */
AgdProblem *agdx_create(int n_patches, int const *volumes, int const *n_dims,
                        int const *const *window_shapes,
                        int const *const *coords, int const *const *shapes,
                        long int const *const *shottimes,
                        int const *const *channels,
                        enum AGDTraceType const *const *trace_types,
                        int const *wavelet_lengths,
                        int const *const *wavelet_idxs,
                        float const *const *wavelets);
void agdx_destroy(AgdProblem *p);
/* Same tables without touching the GPU (geometry and blend tables only;
 * the operators, stepsize and runs are unavailable on such a problem). */
AgdProblem *agdx_create_host(int n_patches, int const *volumes,
                             int const *n_dims,
                             int const *const *window_shapes,
                             int const *const *coords,
                             int const *const *shapes,
                             long int const *const *shottimes,
                             int const *const *channels,
                             enum AGDTraceType const *const *trace_types,
                             int const *wavelet_lengths,
                             int const *const *wavelet_idxs,
                             float const *const *wavelets);
/* record intervals (channel, start, stop), mutes (channel, interval,
 * interval_start, n) and unique channel values; each returns its count and
 * writes only when out != NULL */
int agdx_intervals(AgdProblem const *p, long long *out);
int agdx_mutes(AgdProblem const *p, long long *out);
int agdx_unique_channels(AgdProblem const *p, int *out);

long agdx_model_len(AgdProblem const *p);  /* complex elements */
long agdx_record_len(AgdProblem const *p); /* floats */
int agdx_n_windows(AgdProblem const *p);
/* info[25]: patch, nsp, ntr, ntw, nfk, nconv, Pt, Ft, Wt, S, x_off(lo),
 * x_off(hi), P[4], W[4], coords[4], wavelet convolution length (0 without
 * wavelets) */
int agdx_window_info(AgdProblem const *p, int w, int *info);
/* per window trace: channel, interval, interval_start, trace_start, n */
int agdx_window_blend(AgdProblem const *p, int w, long long *out);
/* taper flags of window w: left, right per space dim (2 * nsp ints;
 * model.c:642-653) */
int agdx_window_taper(AgdProblem const *p, int w, int *use_taper);
/* patch trace index of each trace of window w (ntr ints; model.c:748-805) */
int agdx_window_traces(AgdProblem const *p, int w, int *out);
/* The blend tables agd_blend builds for the given traces (host only, no
 * device; blend.c:113-342, 585-760): counts[3] = (intervals, mutes,
 * channels); intervals (channel, start, stop) in the reference's order,
 * mutes (channel, interval, interval_start, n), unique channel values, the
 * coordinates (channel, interval, interval_start, trace_start, n) of every
 * input trace and of nq query traces of q_ntimes samples (channel -1: not
 * blended). Arrays are filled only when not NULL. */
int agdx_blend_tables(int n_patches, int const *n_traces, int const *n_times,
                      long int const *const *shottimes, int const *const *channels,
                      enum AGDTraceType const *const *trace_types, int nq, int q_ntimes,
                      long int const *q_shottimes, int const *q_channels,
                      enum AGDTraceType const *q_trace_types, int *counts,
                      long long *intervals, long long *mutes, int *channels_out,
                      long long *coords, long long *q_coords);
/* stepsize computed on the device (0.5 / max(1, fold)) */
float agdx_stepsize(AgdProblem const *p);

/* Blends the patch data (mean of duplicates, then mute) into the device
 * record d: the pseudo-deblended input of the inversion
 * (blend.c:1294-1335). Optionally copies d to host (record_len floats). */
int agdx_set_data(AgdProblem *p, float const *const *data, float *d_out);

/* Operators on host buffers (tests): rec = A x (forward, muted record) and
 * g = A^H rec (model), without update. */
int agdx_forward(AgdProblem *p, float const *x, float *rec);
int agdx_adjoint(AgdProblem *p, float const *rec, float *g);

/* Full inversion from the zero start, or from x0 (model_len complex) when
 * x0 != NULL (warm start: the initial update is skipped and the schedule
 * starts from a threshold lambda0 computed from the first adjoint).
 * Leaves the final model on the device. */
int agdx_run(AgdProblem *p, float initial_factor, int n_its, int print_freq,
             float const *x0);
/* Warm-start model of a data estimate y (patches in the layout of
 * agd_deblend's data): per window and time window s R2C(pad(y_w)), no
 * tapers, no x2, so that the forward operator gives taper . y_w and, where
 * the window tapers sum to one, the output of the model is y. Written to
 * the device model. (The reference starts every inversion from the
 * thresholded first update, agdeblend.c:584-590: this is the hook a
 * learned or other warm start plugs into.) */
int agdx_warm_model(AgdProblem *p, float const *const *y);
/* agdx_run starting from agdx_warm_model(warm) instead of the thresholded
 * first update; lambda0 and the schedule are those of the zero start.
 * Guarded: when the warm model fits the data no better than the zero
 * model (||A x_warm - d|| >= ||d||, FP64 norms) the plain zero-start
 * inversion runs instead, with a WARNING on stdout, and
 * agdx_last_warm_fallback() returns 1. */
int agdx_run_warm(AgdProblem *p, float initial_factor, int n_its, int print_freq,
                  float const *const *warm);
/* Spatial DFT precision: 0 FP32 register FFTs (default), 1 split-precision
 * bf16 x 3 matrix-core DFTs (square spatial windows with P in {20, 40};
 * other windows keep FP32). Returns the number of window classes served by
 * the bf16 x 3 path, -1 on error. */
int agdx_set_spatial_precision(AgdProblem *p, int mode);
/* Autotuner. First the radix choice of the time transforms: per
 * specialised time class, every (M1, M2) factorization of the half length
 * Pt / 2 with a kernel (e.g. 160 = 10 x 16, 16 x 10, 20 x 8, 8 x 20) is
 * timed on one iteration's time stage and the fastest kept. Then the
 * spatial kernels' tile (planes per task): per specialised spatial class,
 * every candidate that fits the LDS is timed on one iteration's spatial
 * stage (reps launches) and the fastest kept. Overwrites the intermediates;
 * the model is unchanged. nk_out and ms_out (one entry per spatial class,
 * or NULL) receive the spatial choice (0 for generic classes) and its time
 * per iteration. */
int agdx_autotune(AgdProblem *p, int reps, int *nk_out, double *ms_out);
/* Number of spatial kernel classes of the problem. */
int agdx_n_spatial_classes(AgdProblem const *p);
/* Number of time kernel classes, and their current (M1, M2) factorization
 * (out: 2 ints per class, or NULL; 0 0 for classes of the generic
 * kernels). agdx_time_factors returns the number of classes. */
int agdx_n_time_classes(AgdProblem const *p);
int agdx_time_factors(AgdProblem const *p, int *out);
/* Forward time kernel variant per time class (out: one int per class):
 * 0 the next time window's tile prefetched into registers, 1 the lean tile
 * (one LDS tile for input and transform, the carry in registers, more
 * blocks per CU). The autotuner times both with every factorization. */
int agdx_time_variants(AgdProblem const *p, int *out);
int agdx_set_time_variant(AgdProblem *p, int c, int var);
/* Sets time class c to the factorization m1 x m2 of its half length (one
 * of the autotuner's candidates). */
int agdx_set_time_factors(AgdProblem *p, int c, int m1, int m2);
/* The time kernels' blocks, 4 ints each into out (or NULL): window, first
 * trace, and the run [tw0, tw1) of the window's time windows the block
 * transforms (all of them unless AGD_TIME_TW sets runs, or the window-chunk
 * order AGD_CHUNK_MB does: runs of 4). Returns their number. */
int agdx_time_blocks(AgdProblem const *p, int *out);
/* Device copy of the patch data set by agdx_set_data: all patches
 * concatenated in patch order, agdx_data_len floats (valid until
 * agdx_warm_model, agdx_finish or agdx_get_output overwrite it). */
float const *agdx_data_dev(AgdProblem *p);
int64_t agdx_data_len(AgdProblem *p);
/* agdx_warm_model and agdx_run_warm with the estimate already on the
 * device, in the flat layout of agdx_data_dev; the caller has finished
 * writing it (device or stream synchronised) before the call. */
int agdx_warm_model_dev(AgdProblem *p, float const *y_dev);
int agdx_run_warm_dev(AgdProblem *p, float initial_factor, int n_its, int print_freq,
                      float const *warm_dev);
/* Forward of the current model assembled into the device copy of the patch
 * data (set_output, model.c:1425-1515), without copying it to the host. */
int agdx_finish(AgdProblem *p);
/* agdx_finish, then copies the patches into data. */
int agdx_get_output(AgdProblem *p, float *const *data);
/* Floats of the trace buffer (per window [n_traces][n_times_conv]). */
long agdx_trace_len(AgdProblem const *p);
/* One operator of the forward model A = mute . blend . wavelet . F alone, on
 * host buffers (the per-operator adjoint tests): op 0 the FK operator F
 * (model -> traces; adjoint traces -> model, fk.c:307-452), 1 the wavelet
 * convolution (traces -> traces, wavelet.c:239-301; only on problems with
 * wavelets), 2 the blend with the mute (traces -> record; adjoint record ->
 * traces without the mute, blend.c:862-875, 936-960, 1262-1292). Leaves
 * the model, intermediates and traces overwritten. */
int agdx_operator(AgdProblem *p, int op, int adjoint, float const *in, float *out);
/* Copies the current model to host (model_len complex). */
int agdx_get_model(AgdProblem *p, float *x);

/* Benchmark hooks: run n ISTA iterations from the current device state with
 * a fixed threshold, synchronising at the end. */
int agdx_iterate(AgdProblem *p, int n_iterations, float lamb);
/* Threshold lambda0 of the last agdx_run (first update's max |x|). */
float agdx_last_lambda0(void);
/* 1 when the last warm-started run fell back to the zero start, else 0. */
int agdx_last_warm_fallback(void);
/* ||A x_warm - d|| / ||d|| of the last warm-started run's guard (FP64
 * norms), or -1 when the last run had no warm start. */
double agdx_last_warm_ratio(void);
/* Per-kernel device times accumulated while timing is enabled. */
int agdx_timing(int enable);
int agdx_timing_get(char const *kernel, double *ms, long *count);
/* Bytes each FK kernel is charged per launch (algorithmic, for rooflines):
 * which = 0 model bytes M, 1 trace bytes T, 2 record bytes R */
double agdx_bytes(AgdProblem const *p, int which);

#ifdef __cplusplus
}
#endif

#endif
