/*
 * A C caller of libagdeblend.so, compiled with gcc against include/agdeblend.h
 * and linked with -lagdeblend the way the reference's C examples link the
 * reference library (examples/Makefile:244-245, examples/example_4.c:58).
 *
 * 1. agd_blend known answers of the reference API tests (test.cpp:30-430):
 *    two patches on one channel (sum), missing input / output traces, one
 *    patch to two patches with a bad (muted) trace, and the refusal of a bad
 *    output trace.
 * 2. agd_deblend on a one-patch 2-D problem read from files, output written
 *    to a file (the caller compares it with the oracle).
 *
 * Usage: capi_check                                   (blend checks only)
 *        capi_check NTR NT WTR WT NITS shot.i64 data.f32 out.f32
 * Exit status 0 when every check passes.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "agdeblend.h"

static int n_fail = 0;

static void expect_vals(char const *name, float const *got, float const *want, int n) {
  int i;
  for (i = 0; i < n; i++) {
    if (got[i] != want[i]) {
      fprintf(stderr, "FAIL %s: element %d is %g, expected %g\n", name, i, got[i], want[i]);
      n_fail++;
      return;
    }
  }
  printf("ok %s\n", name);
}

/* One agd_blend call in Sum mode with at most 3 input and 3 output patches. */
struct Case {
  char const *name;
  int np, ntr[3], nt[3];
  long shot[3][8];
  int ch[3][8];
  enum AGDTraceType tt[3][8];
  float data[3][16];
  int npo, ntro[3], nto[3];
  long shoto[3][8];
  int cho[3][8];
  enum AGDTraceType tto[3][8];
  float want[3][16];
  int want_rc;
};

static void run_case(struct Case const *c) {
  long const *shot[3], *shoto[3];
  int const *ch[3], *cho[3];
  enum AGDTraceType const *tt[3], *tto[3];
  float *data[3], *out[3];
  float in[3][16], buf[3][16];
  int p, rc;
  for (p = 0; p < 3; p++) {
    shot[p] = c->shot[p];
    ch[p] = c->ch[p];
    tt[p] = c->tt[p];
    memcpy(in[p], c->data[p], sizeof in[p]);
    data[p] = in[p];
    shoto[p] = c->shoto[p];
    cho[p] = c->cho[p];
    tto[p] = c->tto[p];
    out[p] = buf[p];
    memset(buf[p], 0xff, sizeof buf[p]);
  }
  rc = agd_blend(c->np, c->ntr, c->nt, shot, ch, tt, data, AGDBlendSum, 0, c->npo, c->ntro,
                 c->nto, shoto, cho, tto, out);
  if (rc != c->want_rc) {
    fprintf(stderr, "FAIL %s: agd_blend returned %d, expected %d\n", c->name, rc, c->want_rc);
    n_fail++;
    return;
  }
  if (rc != 0) {
    printf("ok %s (refused)\n", c->name);
    return;
  }
  for (p = 0; p < c->npo; p++) {
    char name[96];
    snprintf(name, sizeof name, "%s, output patch %d", c->name, p);
    expect_vals(name, out[p], c->want[p], c->ntro[p] * c->nto[p]);
  }
}

static void blend_checks(void) {
  static struct Case const cases[] = {
      /* test.cpp:30-112: two one-trace patches on channel 123, shots 2 apart */
      {"two patches one channel", 2, {1, 1}, {3, 3}, {{25}, {27}}, {{123}, {123}},
       {{AGDLive}, {AGDLive}}, {{1, 2, 3}, {4, 5, 6}}, 2, {1, 1}, {3, 3}, {{25}, {27}},
       {{123}, {123}}, {{AGDLive}, {AGDLive}}, {{1, 2, 7}, {7, 5, 6}}, 0},
      /* test.cpp:114-...: a missing input trace blends nothing */
      {"missing in, missing out", 1, {1}, {3}, {{25}}, {{123}}, {{AGDMissing}},
       {{1, 2, 3}}, 1, {1}, {3}, {{25}}, {{123}}, {{AGDMissing}}, {{0, 0, 0}}, 0},
      {"missing in, live out", 1, {1}, {3}, {{25}}, {{123}}, {{AGDMissing}}, {{1, 2, 3}},
       1, {1}, {3}, {{25}}, {{123}}, {{AGDLive}}, {{0, 0, 0}}, 0},
      {"live in, missing out", 1, {1}, {3}, {{25}}, {{123}}, {{AGDLive}}, {{1, 2, 3}}, 1,
       {1}, {3}, {{25}}, {{123}}, {{AGDMissing}}, {{0, 0, 0}}, 0},
      /* one patch of five traces (one missing, one bad) out to two patches */
      {"one patch to two with mute", 1, {5}, {3},
       {{25, 27, 29, 31, -10}}, {{123, 123, 123, 123, -789}},
       {{AGDMissing, AGDLive, AGDLive, AGDBad, AGDLive}},
       {{1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}}, 2, {3, 2}, {2, 3},
       {{24, 26, -10}, {28, 30}}, {{123, 123, -789}, {123, 123}},
       {{AGDMissing, AGDLive, AGDLive}, {AGDLive, AGDLive}},
       {{0, 0, 0, 4, 13, 14}, {5, 13, 8, 8, 0, 0}}, 0},
      /* a bad output trace is refused */
      {"bad output trace", 1, {1}, {3}, {{25}}, {{123}}, {{AGDLive}}, {{1, 2, 3}}, 1, {1},
       {3}, {{25}}, {{123}}, {{AGDBad}}, {{0}}, 1},
  };
  size_t i;
  for (i = 0; i < sizeof cases / sizeof cases[0]; i++) run_case(&cases[i]);
}

static void *read_file(char const *path, size_t bytes) {
  FILE *f = fopen(path, "rb");
  void *buf = malloc(bytes);
  if (f == NULL || buf == NULL || fread(buf, 1, bytes, f) != bytes) {
    fprintf(stderr, "FAIL cannot read %s\n", path);
    exit(2);
  }
  fclose(f);
  return buf;
}

static void deblend_check(char **argv) {
  int const ntr = atoi(argv[1]), nt = atoi(argv[2]);
  int const wshape[2] = {atoi(argv[3]), atoi(argv[4])};
  int const n_its = atoi(argv[5]);
  long *shot = (long *)read_file(argv[6], (size_t)ntr * sizeof(long));
  float *data = (float *)read_file(argv[7], (size_t)ntr * nt * sizeof(float));
  int *ch = (int *)calloc((size_t)ntr, sizeof(int));
  enum AGDTraceType *tt = (enum AGDTraceType *)calloc((size_t)ntr, sizeof *tt);
  int const volumes[1] = {0}, n_dims[1] = {2}, coord[1] = {0}, shape[2] = {ntr, nt};
  int const *wshapes[1] = {wshape}, *coords[1] = {coord}, *shapes[1] = {shape};
  long const *shots[1] = {shot};
  int const *chans[1] = {ch};
  enum AGDTraceType const *tts[1] = {tt};
  float *datas[1] = {data};
  FILE *f;
  int rc = agd_deblend(1, volumes, n_dims, wshapes, coords, shapes, shots, chans, tts, NULL,
                       NULL, NULL, 1.0f, n_its, 0, datas);
  if (rc != 0) {
    fprintf(stderr, "FAIL agd_deblend returned %d\n", rc);
    n_fail++;
    return;
  }
  f = fopen(argv[8], "wb");
  if (f == NULL || fwrite(data, sizeof(float), (size_t)ntr * nt, f) != (size_t)ntr * nt) {
    fprintf(stderr, "FAIL cannot write %s\n", argv[8]);
    n_fail++;
  }
  if (f != NULL) fclose(f);
  printf("ok agd_deblend %d x %d, %d iterations\n", ntr, nt, n_its);
  free(shot);
  free(data);
  free(ch);
  free(tt);
}

int main(int argc, char **argv) {
  blend_checks();
  if (argc == 9) deblend_check(argv);
  else if (argc != 1) {
    fprintf(stderr, "usage: capi_check [NTR NT WTR WT NITS shot.i64 data.f32 out.f32]\n");
    return 2;
  }
  printf("%d failure(s)\n", n_fail);
  return n_fail != 0;
}
