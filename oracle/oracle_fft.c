/*
 * CPU FFT used by the oracle (test infrastructure only; never linked into the
 * product library).
 *
 * The reference calls FFTW3 (fftw[f]_plan_many_dft_r2c/c2r, fk.c:256-262,
 * wavelet.c:172-178,211-221); FFTW is an un-vendored external dependency with
 * no pinned version (reference Makefile:13-35) and is absent from this image.
 * This file restates the published mathematics FFTW computes, with FFTW's
 * conventions:
 *   - forward (R2C) uses e^{-2 pi i k n / N}, backward (C2R) e^{+...};
 *   - both are unnormalised;
 *   - C2R ignores the imaginary parts of the DC and (even N) Nyquist bins,
 *     as FFTW's one-dimensional hc2r does;
 *   - a multi-dimensional C2R first applies complex inverse DFTs along every
 *     dimension but the last, then a one-dimensional C2R along the last; R2C
 *     is the reverse order.
 * Algorithms: iterative mixed-radix Stockham (radices 4, 2, 3, 5, 7 and any
 * prime <= 31 through a direct butterfly), Bluestein's chirp-z for sizes with
 * a larger prime factor, and the half-length complex packing for even-length
 * real transforms. Compiled in the oracle's precision (ORACLE_DOUBLE).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  real re, im;
} ocpx;

#define OFFT_MAX_PLANS 256
#define OFFT_BLUESTEIN_PRIME 31

typedef struct offt_plan {
  int n;
  int nstage;
  int radix[64];
  ocpx *roots; /* roots[m] = exp(-2 pi i m / n) */
  /* Bluestein */
  int bluestein;
  int m;
  struct offt_plan const *sub;
  ocpx *chirp;    /* exp(-pi i k^2 / n), k < n */
  ocpx *filt_fwd; /* FFT_m of conj(chirp) wrapped (for sign -1) */
  ocpx *filt_bwd; /* FFT_m of chirp wrapped (for sign +1) */
} offt_plan;

static offt_plan g_plans[OFFT_MAX_PLANS];
static volatile int g_n_plans = 0;
static pthread_mutex_t g_plan_lock = PTHREAD_MUTEX_INITIALIZER;

static void offt_exec(offt_plan const *p, int sign, ocpx *x, ocpx *work);

static void offt_fill_roots(int n, ocpx *roots) {
  int m;
  for (m = 0; m < n; m++) {
    double const a = -2.0 * M_PI * (double)m / (double)n;
    roots[m].re = (real)cos(a);
    roots[m].im = (real)sin(a);
  }
}

static int offt_has_big_prime(int n) {
  int rem = n;
  int f;
  while (rem % 2 == 0 && rem > 1) rem /= 2;
  for (f = 3; f <= rem; f += 2) {
    while (rem % f == 0) {
      if (f > OFFT_BLUESTEIN_PRIME) return 1;
      rem /= f;
    }
  }
  return 0;
}

static int offt_bluestein_len(int n) {
  int m = 1;
  while (m < 2 * n - 1) m *= 2;
  return m;
}

static void offt_build(offt_plan *p, int n, offt_plan const *sub) {
  int rem = n;
  int f;
  memset(p, 0, sizeof(*p));
  p->n = n;
  p->roots = (ocpx *)malloc((size_t)(n > 0 ? n : 1) * sizeof(ocpx));
  offt_fill_roots(n, p->roots);
  if (sub != NULL) {
    int const m = sub->n;
    int k;
    ocpx *work;
    p->bluestein = 1;
    p->m = m;
    p->sub = sub;
    p->chirp = (ocpx *)malloc((size_t)n * sizeof(ocpx));
    for (k = 0; k < n; k++) {
      /* k^2 mod 2n keeps the angle small and exact */
      long long const k2 = ((long long)k * (long long)k) % (2LL * n);
      double const a = -M_PI * (double)k2 / (double)n;
      p->chirp[k].re = (real)cos(a);
      p->chirp[k].im = (real)sin(a);
    }
    p->filt_fwd = (ocpx *)calloc((size_t)m, sizeof(ocpx));
    p->filt_bwd = (ocpx *)calloc((size_t)m, sizeof(ocpx));
    for (k = 0; k < n; k++) {
      ocpx const c = p->chirp[k];
      p->filt_fwd[k].re = c.re;
      p->filt_fwd[k].im = -c.im;
      p->filt_bwd[k] = c;
      if (k > 0) {
        p->filt_fwd[m - k] = p->filt_fwd[k];
        p->filt_bwd[m - k] = p->filt_bwd[k];
      }
    }
    work = (ocpx *)malloc((size_t)m * sizeof(ocpx));
    offt_exec(sub, -1, p->filt_fwd, work);
    offt_exec(sub, -1, p->filt_bwd, work);
    free(work);
    return;
  }
  while (rem % 4 == 0) {
    p->radix[p->nstage++] = 4;
    rem /= 4;
  }
  while (rem % 2 == 0) {
    p->radix[p->nstage++] = 2;
    rem /= 2;
  }
  for (f = 3; f <= rem; f += 2) {
    while (rem % f == 0) {
      p->radix[p->nstage++] = f;
      rem /= f;
    }
  }
}

static offt_plan const *offt_get(int n) {
  int i;
  int count = g_n_plans;
  offt_plan const *sub = NULL;
  for (i = 0; i < count; i++) {
    if (g_plans[i].n == n) return &g_plans[i];
  }
  /* Bluestein sub-plans (powers of two) are created first, outside the
   * lock, so that a slot is never reserved while another plan is built. */
  if (offt_has_big_prime(n)) sub = offt_get(offt_bluestein_len(n));
  pthread_mutex_lock(&g_plan_lock);
  count = g_n_plans;
  for (i = 0; i < count; i++) {
    if (g_plans[i].n == n) {
      pthread_mutex_unlock(&g_plan_lock);
      return &g_plans[i];
    }
  }
  if (count >= OFFT_MAX_PLANS) {
    pthread_mutex_unlock(&g_plan_lock);
    abort();
  }
  offt_build(&g_plans[count], n, sub);
  __sync_synchronize();
  g_n_plans = count + 1;
  pthread_mutex_unlock(&g_plan_lock);
  return &g_plans[count];
}

/* root w^e for exponent e of the n-th root, direction sign */
static inline ocpx offt_root(ocpx const *roots, int n, long long e, int sign) {
  ocpx w = roots[(int)(e % n)];
  if (sign > 0) w.im = -w.im;
  return w;
}

static inline ocpx cmul(ocpx a, ocpx b) {
  ocpx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

/* One Stockham pass: radix R, Ns = product of radices already applied. */
static void offt_stage(offt_plan const *p, int R, int Ns, int sign,
                       ocpx const *x, ocpx *y) {
  int const n = p->n;
  int const nr = n / R;
  int const tw_step = n / (Ns * R); /* w_{Ns*R} = w_n^{tw_step} */
  int j;
  ocpx v[OFFT_BLUESTEIN_PRIME + 1];
  ocpx o[OFFT_BLUESTEIN_PRIME + 1];
  for (j = 0; j < nr; j++) {
    int const k = j % Ns;
    int const base = (j / Ns) * Ns * R + k;
    int r;
    for (r = 0; r < R; r++) {
      ocpx a = x[j + r * nr];
      if (r > 0 && k > 0) {
        a = cmul(a, offt_root(p->roots, n, (long long)r * k * tw_step, sign));
      }
      v[r] = a;
    }
    if (R == 2) {
      o[0].re = v[0].re + v[1].re;
      o[0].im = v[0].im + v[1].im;
      o[1].re = v[0].re - v[1].re;
      o[1].im = v[0].im - v[1].im;
    } else if (R == 4) {
      ocpx const s02 = {v[0].re + v[2].re, v[0].im + v[2].im};
      ocpx const d02 = {v[0].re - v[2].re, v[0].im - v[2].im};
      ocpx const s13 = {v[1].re + v[3].re, v[1].im + v[3].im};
      ocpx const d13 = {v[1].re - v[3].re, v[1].im - v[3].im};
      /* multiply d13 by -i (sign -1) or +i (sign +1) */
      ocpx const jd13 = {sign < 0 ? d13.im : -d13.im,
                         sign < 0 ? -d13.re : d13.re};
      o[0].re = s02.re + s13.re;
      o[0].im = s02.im + s13.im;
      o[1].re = d02.re + jd13.re;
      o[1].im = d02.im + jd13.im;
      o[2].re = s02.re - s13.re;
      o[2].im = s02.im - s13.im;
      o[3].re = d02.re - jd13.re;
      o[3].im = d02.im - jd13.im;
    } else {
      int q;
      int const rstep = n / R;
      for (q = 0; q < R; q++) {
        ocpx acc = v[0];
        for (r = 1; r < R; r++) {
          ocpx const w =
              offt_root(p->roots, n, (long long)((q * r) % R) * rstep, sign);
          ocpx const t = cmul(v[r], w);
          acc.re += t.re;
          acc.im += t.im;
        }
        o[q] = acc;
      }
    }
    for (r = 0; r < R; r++) y[base + r * Ns] = o[r];
  }
}

/* In-place unnormalised complex DFT of x (length p->n); work has p->n (or
 * p->m for Bluestein plus n) elements. sign=-1 forward, +1 backward. */
static void offt_exec(offt_plan const *p, int sign, ocpx *x, ocpx *work) {
  int const n = p->n;
  if (n <= 1) return;
  if (!p->bluestein) {
    ocpx *src = x;
    ocpx *dst = work;
    int Ns = 1;
    int s;
    for (s = 0; s < p->nstage; s++) {
      ocpx *t;
      offt_stage(p, p->radix[s], Ns, sign, src, dst);
      Ns *= p->radix[s];
      t = src;
      src = dst;
      dst = t;
    }
    if (src != x) memcpy(x, src, (size_t)n * sizeof(ocpx));
  } else {
    int const m = p->m;
    ocpx *a = work;          /* m */
    ocpx *w2 = work + m;     /* m scratch for the sub FFT */
    ocpx const *filt = sign < 0 ? p->filt_fwd : p->filt_bwd;
    real const inv_m = (real)1 / (real)m;
    int k;
    for (k = 0; k < n; k++) {
      ocpx c = p->chirp[k];
      if (sign > 0) c.im = -c.im;
      a[k] = cmul(x[k], c);
    }
    for (k = n; k < m; k++) {
      a[k].re = 0;
      a[k].im = 0;
    }
    offt_exec(p->sub, -1, a, w2);
    for (k = 0; k < m; k++) a[k] = cmul(a[k], filt[k]);
    offt_exec(p->sub, +1, a, w2);
    for (k = 0; k < n; k++) {
      ocpx c = p->chirp[k];
      if (sign > 0) c.im = -c.im;
      c.re *= inv_m;
      c.im *= inv_m;
      x[k] = cmul(a[k], c);
    }
  }
}

static size_t offt_work_elems(int n) {
  offt_plan const *p = offt_get(n);
  if (p->bluestein) return (size_t)(2 * p->m + n);
  return (size_t)n;
}

/* Real-to-half-complex, unnormalised, e^{-i}. out has n/2+1 elements.
 * work needs offt_real_work_elems(n) elements. */
static void offt_r2c(int n, real const *in, ocpx *out, ocpx *work) {
  if (n % 2 == 0 && n >= 4) {
    int const h = n / 2;
    offt_plan const *p = offt_get(h);
    offt_plan const *pn = offt_get(n);
    ocpx *z = work;
    int k;
    for (k = 0; k < h; k++) {
      z[k].re = in[2 * k];
      z[k].im = in[2 * k + 1];
    }
    offt_exec(p, -1, z, work + h);
    for (k = 0; k <= h; k++) {
      ocpx const zk = z[k % h];
      ocpx zc = z[(h - k) % h];
      ocpx e, o, w;
      zc.im = -zc.im;
      e.re = (real)0.5 * (zk.re + zc.re);
      e.im = (real)0.5 * (zk.im + zc.im);
      /* o = (zk - zc) / (2i) */
      o.re = (real)0.5 * (zk.im - zc.im);
      o.im = (real)-0.5 * (zk.re - zc.re);
      w = offt_root(pn->roots, n, k, -1);
      w = cmul(o, w);
      out[k].re = e.re + w.re;
      out[k].im = e.im + w.im;
    }
    out[0].im = 0;
    out[h].im = 0;
  } else {
    offt_plan const *p = offt_get(n);
    ocpx *z = work;
    int k;
    for (k = 0; k < n; k++) {
      z[k].re = in[k];
      z[k].im = 0;
    }
    offt_exec(p, -1, z, work + n);
    for (k = 0; k <= n / 2; k++) out[k] = z[k];
    out[0].im = 0;
  }
}

/* Half-complex-to-real, unnormalised, e^{+i}; imaginary parts of the DC and
 * (even n) Nyquist inputs are ignored. in is not modified. */
static void offt_c2r(int n, ocpx const *in, real *out, ocpx *work) {
  if (n % 2 == 0 && n >= 4) {
    int const h = n / 2;
    offt_plan const *p = offt_get(h);
    offt_plan const *pn = offt_get(n);
    ocpx *z = work;
    int k;
    for (k = 0; k < h; k++) {
      ocpx xk = in[k];
      ocpx xc = in[h - k];
      ocpx e, o, w;
      if (k == 0) {
        xk.im = 0;
        xc.im = 0;
      }
      xc.im = -xc.im; /* conj(X[h-k]) */
      e.re = xk.re + xc.re;
      e.im = xk.im + xc.im;
      o.re = xk.re - xc.re;
      o.im = xk.im - xc.im;
      w = offt_root(pn->roots, n, k, +1);
      o = cmul(o, w);
      /* z = e + i o */
      z[k].re = e.re - o.im;
      z[k].im = e.im + o.re;
    }
    offt_exec(p, +1, z, work + h);
    for (k = 0; k < h; k++) {
      out[2 * k] = z[k].re;
      out[2 * k + 1] = z[k].im;
    }
  } else {
    offt_plan const *p = offt_get(n);
    ocpx *z = work;
    int const nh = n / 2;
    int k;
    z[0].re = in[0].re;
    z[0].im = 0;
    for (k = 1; k <= nh; k++) {
      z[k] = in[k];
      if (n - k != k) {
        z[n - k].re = in[k].re;
        z[n - k].im = -in[k].im;
      }
    }
    if (n % 2 == 0) z[nh].im = 0;
    offt_exec(p, +1, z, work + n);
    for (k = 0; k < n; k++) out[k] = z[k].re;
  }
}

static size_t offt_real_work_elems(int n) {
  if (n % 2 == 0 && n >= 4) {
    offt_get(n);
    return (size_t)(n / 2) + offt_work_elems(n / 2);
  }
  return (size_t)n + offt_work_elems(n);
}

/* Complex DFT along one dimension of a row-major complex array with
 * shape dims[0..rank-1] (only the first `rank` dims are indexed). */
static void offt_along_dim(ocpx *data, int rank, int const *dims, int dim,
                           int sign, ocpx *line, ocpx *work) {
  int const n = dims[dim];
  size_t stride = 1;
  size_t outer = 1;
  size_t o, i;
  int d, k;
  offt_plan const *p;
  if (n <= 1) return;
  p = offt_get(n);
  for (d = dim + 1; d < rank; d++) stride *= (size_t)dims[d];
  for (d = 0; d < dim; d++) outer *= (size_t)dims[d];
  for (o = 0; o < outer; o++) {
    for (i = 0; i < stride; i++) {
      ocpx *const base = data + o * (size_t)n * stride + i;
      for (k = 0; k < n; k++) line[k] = base[(size_t)k * stride];
      offt_exec(p, sign, line, work);
      for (k = 0; k < n; k++) base[(size_t)k * stride] = line[k];
    }
  }
}

static size_t offt_nd_work_elems(int rank, int const *n) {
  size_t w = 0;
  int d;
  for (d = 0; d < rank - 1; d++) {
    size_t const e = (size_t)n[d] + offt_work_elems(n[d]);
    if (e > w) w = e;
  }
  {
    size_t const e = offt_real_work_elems(n[rank - 1]);
    if (e > w) w = e;
  }
  return w;
}

/* Multi-dimensional C2R of one transform: in has shape
 * [n0]..[n_{rank-2}][n_{rank-1}/2+1] (complex, destroyed), out real
 * [n0]..[n_{rank-1}]. */
static void offt_c2r_nd(int rank, int const *n, ocpx *in, real *out,
                        ocpx *work) {
  int fdims[8];
  int d;
  int const nl = n[rank - 1];
  int const fl = nl / 2 + 1;
  size_t rows = 1;
  size_t r;
  for (d = 0; d < rank - 1; d++) {
    fdims[d] = n[d];
    rows *= (size_t)n[d];
  }
  fdims[rank - 1] = fl;
  for (d = 0; d < rank - 1; d++) {
    offt_along_dim(in, rank, fdims, d, +1, work, work + n[d]);
  }
  for (r = 0; r < rows; r++) {
    offt_c2r(nl, in + r * (size_t)fl, out + r * (size_t)nl, work);
  }
}

/* Multi-dimensional R2C of one transform (inverse shapes of offt_c2r_nd). */
static void offt_r2c_nd(int rank, int const *n, real const *in, ocpx *out,
                        ocpx *work) {
  int fdims[8];
  int d;
  int const nl = n[rank - 1];
  int const fl = nl / 2 + 1;
  size_t rows = 1;
  size_t r;
  for (d = 0; d < rank - 1; d++) {
    fdims[d] = n[d];
    rows *= (size_t)n[d];
  }
  fdims[rank - 1] = fl;
  for (r = 0; r < rows; r++) {
    offt_r2c(nl, in + r * (size_t)nl, out + r * (size_t)fl, work);
  }
  for (d = 0; d < rank - 1; d++) {
    offt_along_dim(out, rank, fdims, d, -1, work, work + n[d]);
  }
}
