/*
 * Public C interface of the MI355X deblending library (libagdeblend.so).
 *
 * Drop-in replacement for the interface declared in the reference's
 * include/agdeblend.h:1-46: same enums and values, AGD_TYPE = float (the
 * reference's default), the same tuning constants and the same two entry
 * points with identical argument order, types and meaning. Existing C
 * callers built without AGD_DOUBLE / AGD_MPI, the cffi Python
 * wrapper (wrappers/python/agdeblend_build.py:7,19) and the Julia/Fortran
 * shims bind to these symbols unchanged.
 *
 *   agd_deblend  replaces agdeblend.c:511-649 (ISTA deblending)
 *   agd_blend    replaces agdeblend.c:800-938 (blend / pseudo-deblend)
 *
 * Both calls are synchronous: device memory is allocated, used and released
 * inside the call and all results are in the caller's host buffers when the
 * call returns. Return value: 0 on success, 1 on failure (diagnostics on
 * stderr, "ERROR: ..." lines as in the reference).
 *
 * Only the single-process library is provided; the MPI variants of the two
 * entry points (extra MPI_Comm argument) are not.
 */
#ifndef AGD_H
#define AGD_H

#ifdef AGD_MPI
#error "libagdeblend (MI355X) is single-process: the AGD_MPI interface is not provided"
#endif
/* The reference builds a second library, libagdeblend_double.so, for
 * callers compiled with AGD_DOUBLE (reference Makefile:13-27). This library
 * computes in float (the reference's default AGD_TYPE) and has no double
 * build: a caller compiled with AGD_DOUBLE would pass double buffers to
 * float code, so that configuration is refused here. */
#ifdef AGD_DOUBLE
#error "libagdeblend (MI355X) computes in float: the AGD_DOUBLE interface (libagdeblend_double) is not provided"
#endif

/* Planner flags the reference passes to FFTW (reference
 * include/agdeblend.h:20). This library has no FFTW: its transforms are
 * its own HIP kernels, so the value is accepted and ignored. */
#ifndef AGD_FFTW_FLAGS
#define AGD_FFTW_FLAGS 0
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum AGDTraceType { AGDLive, AGDBad, AGDMissing };
enum AGDBlendMode { AGDBlendSum, AGDBlendMean, AGDBlendOverwrite };

#define AGD_TYPE float

/* Fraction of zero padding added to each window dimension before the
 * Fourier transform (reference include/agdeblend.h:17). */
#define AGD_PAD_FRACTION ((AGD_TYPE)0.25)
/* Final threshold as a fraction of the initial one (include/agdeblend.h:18). */
#define AGD_TARGET_LAMB ((AGD_TYPE)0.001) /* 1e-3 */
/* Maximum number of space dimensions per volume (include/agdeblend.h:19). */
#define AGD_MAX_N_COORDS 4

/* Deblend: data is in/out (blended input, deblended output).
 * Reference: include/agdeblend.h:22-32, src/agdeblend.c:511-649. */
int agd_deblend(int n_patches, int const *volumes, int const *n_dims,
                int const *const *window_shapes, int const *const *coords,
                int const *const *shapes, long int const *const *shottimes,
                int const *const *channels,
                enum AGDTraceType const *const *trace_types,
                int const *wavelet_lengths, int const *const *wavelet_idxs,
                AGD_TYPE const *const *wavelets, AGD_TYPE const initial_factor,
                int n_its, int print_freq, AGD_TYPE *const *data);

/* Blend (sum / mean / overwrite) into continuous records, then extract the
 * requested output traces. data_out may alias data.
 * Reference: include/agdeblend.h:33-44, src/agdeblend.c:800-938. */
int agd_blend(int n_patches, int const *n_traces, int const *n_times,
              long int const *const *shottimes, int const *const *channels,
              enum AGDTraceType const *const *trace_types,
              AGD_TYPE *const *data, enum AGDBlendMode blend_mode,
              int taper_length, int n_patches_out, int const *n_traces_out,
              int const *n_times_out, long int const *const *shottimes_out,
              int const *const *channels_out,
              enum AGDTraceType const *const *trace_types_out,
              AGD_TYPE *const *data_out);

/* agd_blend without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory during the call
 * (code under HIP graph capture, tight serving or training-data loops,
 * harnesses that forbid hipMalloc). Same arguments, checks and results as
 * agd_blend (bit for bit), but every device buffer of the call (input copy,
 * tables, record, output staging) comes from one fixed 64 MiB array that
 * each GPU holds from library load until the process ends; a call whose
 * buffers need more is refused (return 1, "ERROR: ..." on stderr) before
 * any device work. Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (several host threads) race and must not be issued. */
int agd_blend_noalloc(int n_patches, int const *n_traces, int const *n_times,
                      long int const *const *shottimes, int const *const *channels,
                      enum AGDTraceType const *const *trace_types,
                      AGD_TYPE *const *data, enum AGDBlendMode blend_mode,
                      int taper_length, int n_patches_out, int const *n_traces_out,
                      int const *n_times_out, long int const *const *shottimes_out,
                      int const *const *channels_out,
                      enum AGDTraceType const *const *trace_types_out,
                      AGD_TYPE *const *data_out);

/* Aliases with the library-prefixed names; identical behaviour. */
int agdeblend_deblend(int n_patches, int const *volumes, int const *n_dims,
                      int const *const *window_shapes,
                      int const *const *coords, int const *const *shapes,
                      long int const *const *shottimes,
                      int const *const *channels,
                      enum AGDTraceType const *const *trace_types,
                      int const *wavelet_lengths,
                      int const *const *wavelet_idxs,
                      AGD_TYPE const *const *wavelets,
                      AGD_TYPE const initial_factor, int n_its, int print_freq,
                      AGD_TYPE *const *data);
int agdeblend_blend(int n_patches, int const *n_traces, int const *n_times,
                    long int const *const *shottimes,
                    int const *const *channels,
                    enum AGDTraceType const *const *trace_types,
                    AGD_TYPE *const *data, enum AGDBlendMode blend_mode,
                    int taper_length, int n_patches_out,
                    int const *n_traces_out, int const *n_times_out,
                    long int const *const *shottimes_out,
                    int const *const *channels_out,
                    enum AGDTraceType const *const *trace_types_out,
                    AGD_TYPE *const *data_out);

#ifdef __cplusplus
}
#endif

#endif
