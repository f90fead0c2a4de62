/* sanitize_driver.c -- test infrastructure: drives every entry point of the
 * oracle (nlm_oracle.c) on small media so that a build with AddressSanitizer +
 * UndefinedBehaviorSanitizer, or with ThreadSanitizer, checks the CPU
 * restatement for out-of-bounds accesses, leaks, undefined behaviour and data
 * races.  Built and run by tests/test_oracle_sanitizers.py; exits 0 when every
 * call behaved as expected, 1 otherwise (sanitizer reports abort on their own).
 *
 * Covered: every solver's constructor, pump and signal runs, Green's functions
 * and batches on several threads (the reference's thread pools,
 * src/_NonlinearMedium.cpp:263-375 and :378-444), custom pumps, Sinc pulses
 * with chirp and delay, setPump(other), the apodized intensity profile, poling,
 * the transforms at power-of-two and mixed-radix lengths, and error returns.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nlm_oracle.h"

static int failures = 0;

#define CHECK(cond, what)                                             \
  do {                                                                \
    if (!(cond)) {                                                    \
      fprintf(stderr, "FAIL %s: %s (%s)\n", what, #cond, orc_last_error()); \
      failures++;                                                     \
    }                                                                 \
  } while (0)

typedef struct {
  int model, nS, nP, poled, fnlm, nNl, nDiff;
  const char* name;
} spec_t;

/* signal / pump mode counts, poling, nonlinear lengths and phase mismatches
   per solver (the constructors of solvers/*.cpp) */
static const spec_t SPECS[] = {
    {ORC_CHI3, 1, 1, 0, 0, 1, 0, "Chi3"},
    {ORC_CHI2PDC, 1, 1, 1, 0, 1, 1, "Chi2PDC"},
    {ORC_CHI2SFG, 2, 1, 1, 0, 2, 1, "Chi2SFG"},
    {ORC_CHI2AFC, 2, 1, 0, 0, 2, 2, "Chi2AFC"},
    {ORC_CHI2SFGPDC, 2, 1, 1, 0, 2, 2, "Chi2SFGPDC"},
    {ORC_CHI2SFGII, 4, 1, 1, 0, 2, 3, "Chi2SFGII"},
    {ORC_CHI2PDCII, 2, 1, 1, 0, 2, 1, "Chi2PDCII"},
    {ORC_CHI2SFGOPA, 2, 2, 1, 0, 4, 2, "Chi2SFGOPA"},
    {ORC_CHI2SFGXPM, 2, 1, 1, 0, 4, 1, "Chi2SFGXPM"},
    {ORC_CHI2SHG, 2, 0, 1, 1, 2, 1, "Chi2SHG"},
    {ORC_CHI2SHGOPA, 4, 0, 1, 1, 4, 2, "Chi2SHGOPA"},
    {ORC_CHI2DSFG, 3, 0, 1, 1, 3, 1, "Chi2DSFG"},
    {ORC_CHI2SHGXPM, 2, 0, 1, 1, 3, 1, "Chi2SHGXPM"},
    {ORC_CHI2ASHG, 2, 0, 0, 1, 2, 2, "Chi2ASHG"},
    {ORC_CHI3GNLSE, 1, 0, 0, 1, 2, 0, "Chi3GNLSE"},
};

static const double NL[4] = {2.0, 3.0, 5.0, 7.0};
static const double B2P[2] = {0.5, -0.3}, B2S[4] = {-0.4, 0.2, 0.1, -0.25};
static const double B1P[2] = {0.0, 0.1}, B1S[4] = {0.3, -0.2, 0.15, 0.05};
static const double B3P[2] = {0.01, 0.0}, B3S[4] = {0.02, -0.01, 0.0, 0.03};
static const double DIFF[3] = {1.5, -2.0, 0.5};
static const double POLING[3] = {0.3, 0.45, 0.25};

static orc_params make_params(const spec_t* s, int N, int nZ, int profile, double zR) {
  orc_params p;
  memset(&p, 0, sizeof p);
  p.model = s->model;
  p.n_signal_modes = s->nS;
  p.n_pump_modes = s->nP;
  p.can_be_poled = s->poled;
  p.fully_nonlinear = s->fnlm;
  p.relative_length = 1.0;
  p.n_nl = s->nNl;
  p.nl_length = NL;
  p.n_beta2 = s->nP;
  p.beta2 = B2P;
  p.n_beta2s = s->nS;
  p.beta2s = B2S;
  p.n_beta1 = s->nP;
  p.beta1 = B1P;
  p.n_beta1s = s->nS;
  p.beta1s = B1S;
  p.n_beta3 = s->nP;
  p.beta3 = B3P;
  p.n_beta3s = s->nS;
  p.beta3s = B3S;
  p.n_diff = s->nDiff;
  p.diff_beta0 = DIFF;
  p.rayleigh_length = zR;
  p.t_max = 10.0;
  p.t_precision = N;
  p.z_precision = 100;
  p.nz_override = nZ;
  p.intensity_profile = profile;
  p.pulse_type = 0;
  if (s->poled) {
    p.n_poling = 3;
    p.poling = POLING;
  }
  if (s->model == ORC_CHI3GNLSE) {
    const double g[5] = {0.18, 0.21, 0.0122, 0.032, 0.096};
    memcpy(p.gnlse, g, sizeof g);
  }
  return p;
}

static double* seeded_rows(int rows, int n, unsigned seed) {
  double* x = (double*)malloc(sizeof(double) * 2 * (size_t)rows * n);
  for (size_t k = 0; k < 2 * (size_t)rows * n; k++) {
    seed = seed * 1103515245u + 12345u;
    x[k] = ((seed >> 8) & 0xffff) / 65536.0 - 0.5;
  }
  return x;
}

static void exercise(const spec_t* s, int N, int nZ, int profile, double zR) {
  orc_params p = make_params(s, N, nZ, profile, zR);
  orc_medium* m = NULL;
  CHECK(orc_create(&p, &m) == ORC_OK, s->name);
  if (!m) return;
  int32_t d[8];
  orc_dims(m, d);
  CHECK(d[0] == N && d[1] == nZ, s->name);

  if (!s->fnlm) {
    /* a Sinc pulse with chirp and delay, then a custom pump with chirp */
    CHECK(orc_set_pump_type(m, 2, 0.3, 0.2, 0) == ORC_OK, s->name);
    double* cp = seeded_rows(1, N, 7u + (unsigned)s->model);
    CHECK(orc_set_pump_custom(m, cp, N, 0.1, 0.0, 0) == ORC_OK, s->name);
    free(cp);
    CHECK(orc_run_pump(m) == ORC_OK, s->name);
  }

  /* one signal run in each domain, every input mode */
  double* in = seeded_rows(s->nS, N, 11u + (unsigned)s->model);
  for (int mode = 0; mode < s->nS; mode++) CHECK(orc_run_signal(m, in, N, mode & 1, mode) == ORC_OK, s->name);

  /* batches on three threads: every signal mode as input, all outputs */
  const int rows = 3;
  double* X = seeded_rows(rows, s->nS * N, 13u + (unsigned)s->model);
  int32_t shp[2];
  CHECK(orc_batch_shape(m, rows, s->nS * N, NULL, 0, shp) == ORC_OK, s->name);
  double* out = (double*)calloc(2 * (size_t)shp[0] * shp[1], sizeof(double));
  CHECK(orc_batch(m, X, rows, s->nS * N, 0, !s->fnlm, 3, 0, NULL, 0, out) == ORC_OK, s->name);
  CHECK(orc_batch(m, X, rows, N, 1, 0, 2, s->nS - 1, NULL, 0, out) == ORC_OK, s->name);
  /* too many threads is an error, not a crash */
  CHECK(orc_batch(m, X, rows, N, 0, 0, rows + 1, 0, NULL, 0, out) == ORC_INVALID_ARGUMENT, s->name);
  free(out);
  free(X);

  if (!s->fnlm) {
    /* Green's function on four threads, every mode in and out */
    CHECK(orc_greens_shape(m, NULL, 0, NULL, 0, shp) == ORC_OK, s->name);
    double* C = (double*)calloc(2 * (size_t)shp[0] * shp[1], sizeof(double));
    double* S = (double*)calloc(2 * (size_t)shp[0] * shp[1], sizeof(double));
    CHECK(orc_greens(m, 0, 1, 4, 0, NULL, 0, NULL, 0, C, S) == ORC_OK, s->name);
    /* normalized, where the medium has one nonlinear length per mode */
    const int norm = s->nNl == s->nS;
    CHECK(orc_greens(m, 1, 0, 3, norm, NULL, 0, NULL, 0, C, S) == ORC_OK, s->name);
    free(C);
    free(S);
  }
  free(in);
  orc_destroy(m);
}

static void pump_import(int N) {
  /* setPump(other): a Chi2PDC pump from the signal grid of a Chi3 medium */
  orc_params pa = make_params(&SPECS[0], N, 12, 0, INFINITY);
  orc_params pb = make_params(&SPECS[1], N, 9, 0, INFINITY);
  orc_medium *a = NULL, *b = NULL;
  CHECK(orc_create(&pa, &a) == ORC_OK, "import source");
  CHECK(orc_create(&pb, &b) == ORC_OK, "import target");
  if (a && b) {
    double* in = seeded_rows(1, N, 5u);
    CHECK(orc_run_pump(a) == ORC_OK, "import source pump");
    CHECK(orc_run_signal(a, in, N, 1, 0) == ORC_OK, "import source signal");
    CHECK(orc_set_pump_other(b, a, 0, 0.1, 0) == ORC_OK, "setPump(other)");
    CHECK(orc_run_pump(b) == ORC_OK, "imported pump run");
    CHECK(orc_run_signal(b, in, N, 0, 0) == ORC_OK, "imported pump signal");
    CHECK(orc_set_pump_other(b, a, 3, 0.0, 0) != ORC_OK, "mode index out of range");
    free(in);
  }
  if (a) orc_destroy(a);
  if (b) orc_destroy(b);
}

static void transforms(void) {
  const int sizes[] = {2, 6, 8, 12, 30, 64, 210, 512, 1000};
  for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; i++) {
    const int n = sizes[i], count = 3;
    double* x = seeded_rows(count, n, 17u + (unsigned)n);
    double* y = (double*)malloc(sizeof(double) * 2 * (size_t)count * n);
    double* z = (double*)malloc(sizeof(double) * 2 * (size_t)count * n);
    CHECK(orc_fft(n, count, x, y, 0) == ORC_OK, "fft");
    CHECK(orc_fft(n, count, y, z, 1) == ORC_OK, "ifft");
    double err = 0, ref = 0;
    for (size_t k = 0; k < 2 * (size_t)count * n; k++) {
      err = fmax(err, fabs(z[k] - x[k]));  /* the inverse carries 1/n */
      ref = fmax(ref, fabs(x[k]));
    }
    CHECK(err <= 1e-13 * (ref + 1), "fft round trip");
    free(x);
    free(y);
    free(z);
  }
}

static void errors(void) {
  orc_params p = make_params(&SPECS[1], 63, 10, 0, INFINITY);  /* odd nFreqs */
  orc_medium* m = NULL;
  CHECK(orc_create(&p, &m) != ORC_OK, "odd nFreqs rejected");
  if (m) orc_destroy(m);
  p = make_params(&SPECS[1], 64, 10, 0, INFINITY);
  const double badPoling[2] = {0.5, -0.1};
  p.n_poling = 2;
  p.poling = badPoling;
  m = NULL;
  CHECK(orc_create(&p, &m) != ORC_OK, "invalid domain length rejected");
  if (m) orc_destroy(m);
  p = make_params(&SPECS[1], 64, 10, 0, INFINITY);
  p.t_max = -1.0;
  m = NULL;
  CHECK(orc_create(&p, &m) != ORC_OK, "negative time span rejected");
  if (m) orc_destroy(m);
  CHECK(orc_fft(0, 1, NULL, NULL, 0) != ORC_OK, "empty transform rejected");
}

int main(void) {
  const size_t nSpecs = sizeof SPECS / sizeof SPECS[0];
  for (size_t i = 0; i < nSpecs; i++) {
    exercise(&SPECS[i], 64, 8, 0, INFINITY);
    exercise(&SPECS[i], 30, 6, 2, 0.7);  /* mixed-radix length, apodized profile */
  }
  pump_import(64);
  transforms();
  errors();
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle sanitizer driver: all checks passed\n");
  return 0;
}
