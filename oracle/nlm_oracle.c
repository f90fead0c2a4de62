/*
 * nlm_oracle.c -- CPU restatement of the split-step propagation core of
 * `nonlinearmedium`.  TEST INFRASTRUCTURE ONLY (the parity reference and the
 * CPU baseline); the product never links or loads this file's library.
 *
 * Each function names the reference lines it restates.  The arithmetic keeps
 * the reference's operation order (left-to-right evaluation of the Eigen
 * expressions) so that results agree with the reference to rounding.
 *
 * The FFT is this file's own mixed-radix decimation-in-time transform
 * (radix 4, 2, 3, 5 and generic odd factors), the same class of algorithm as
 * the reference's default Eigen/kissfft backend.  Conventions follow
 * src/CustomEigenFFT.h:54-137: forward unscaled, inverse scaled by 1/n.
 */
#define _GNU_SOURCE
#include "nlm_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------------- */
/* errors                                                                    */
/* ------------------------------------------------------------------------- */

static __thread char g_err[256];

const char* orc_last_error(void) { return g_err; }

static int fail(int kind, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return kind;
}

/* cos(t) + i sin(t): the value Eigen's .exp() gives for a complex argument
 * whose real part is (+/-)0, which is every exp(i * real) in the reference. */
static inline cplx cis(double t) { return CMPLX(cos(t), sin(t)); }

/* ------------------------------------------------------------------------- */
/* FFT                                                                       */
/* ------------------------------------------------------------------------- */

typedef struct {
  int n, nf;
  int fac[64];
  cplx* tw;          /* exp(-2 pi i k / n), k < n */
  cplx* ltw[64];     /* per recursion level: W_{n_l}^(j*k), k < m_l, j = 1..p_l-1 */
} fftplan;

static void fft_plan_init(fftplan* p, int n) {
  memset(p, 0, sizeof *p);
  p->n = n;
  int r = n;
  while (r % 4 == 0) { p->fac[p->nf++] = 4; r /= 4; }
  while (r % 2 == 0) { p->fac[p->nf++] = 2; r /= 2; }
  for (int f = 3; f * f <= r; f += 2)
    while (r % f == 0) { p->fac[p->nf++] = f; r /= f; }
  if (r > 1) p->fac[p->nf++] = r;
  p->tw = (cplx*)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n; k++) {
    long double a = -2.0L * pi * (long double)k / (long double)n;
    p->tw[k] = CMPLX((double)cosl(a), (double)sinl(a));
  }
  int nl = n;
  for (int l = 0; l < p->nf; l++) {
    const int pp = p->fac[l], m = nl / pp, ts = n / nl;
    p->ltw[l] = (cplx*)malloc(sizeof(cplx) * (size_t)(m * (pp - 1) > 0 ? m * (pp - 1) : 1));
    for (int k = 0; k < m; k++)
      for (int j = 1; j < pp; j++) p->ltw[l][(long)k * (pp - 1) + (j - 1)] = p->tw[((long)j * k * ts) % n];
    nl = m;
  }
}

static void fft_plan_free(fftplan* p) {
  free(p->tw);
  p->tw = NULL;
  for (int l = 0; l < p->nf; l++) free(p->ltw[l]);
}

/* out[0..n) = DFT_n(in[0], in[s], in[2s], ...) at recursion level `lev`. */
static void dft_rec(cplx* out, const cplx* in, long s, int lev, const fftplan* pl, int n, int inv) {
  const int p = pl->fac[lev], m = n / p;
  if (m == 1) {
    for (int j = 0; j < p; j++) out[j] = in[(long)j * s];
  } else {
    for (int j = 0; j < p; j++) dft_rec(out + (long)j * m, in + (long)j * s, s * p, lev + 1, pl, m, inv);
  }
  const cplx* tw = pl->ltw[lev];
  if (p == 2) {
    for (int k = 0; k < m; k++) {
      cplx w = inv ? conj(tw[k]) : tw[k];
      cplx t0 = out[k], t1 = out[k + m] * w;
      out[k] = t0 + t1;
      out[k + m] = t0 - t1;
    }
  } else if (p == 4) {
    for (int k = 0; k < m; k++) {
      const cplx* w = tw + 3L * k;
      cplx w1 = w[0], w2 = w[1], w3 = w[2];
      if (inv) { w1 = conj(w1); w2 = conj(w2); w3 = conj(w3); }
      cplx t0 = out[k], t1 = out[k + m] * w1, t2 = out[k + 2 * m] * w2, t3 = out[k + 3 * m] * w3;
      cplx a = t0 + t2, b = t0 - t2, c = t1 + t3, d = t1 - t3;
      /* forward: -i*d, inverse: +i*d */
      cplx jd = inv ? CMPLX(-cimag(d), creal(d)) : CMPLX(cimag(d), -creal(d));
      out[k] = a + c;
      out[k + m] = b + jd;
      out[k + 2 * m] = a - c;
      out[k + 3 * m] = b - jd;
    }
  } else {
    const int N = pl->n;
    cplx tbuf[64];
    cplx* t = p <= 64 ? tbuf : (cplx*)malloc(sizeof(cplx) * (size_t)p);
    for (int k = 0; k < m; k++) {
      t[0] = out[k];
      for (int j = 1; j < p; j++) {
        cplx w = tw[(long)k * (p - 1) + (j - 1)];
        if (inv) w = conj(w);
        t[j] = out[(long)j * m + k] * w;
      }
      for (int q = 0; q < p; q++) {
        cplx acc = t[0];
        for (int j = 1; j < p; j++) {
          cplx w = pl->tw[(((long)j * q) % p) * (N / p)];
          if (inv) w = conj(w);
          acc += t[j] * w;
        }
        out[k + (long)q * m] = acc;
      }
    }
    if (t != tbuf) free(t);
  }
}

/* Eigen::FFT::fwd (src/CustomEigenFFT.h:54-89): unscaled. out may alias in. */
static void fft_fwd(const fftplan* p, cplx* out, const cplx* in, cplx* scratch) {
  memcpy(scratch, in, sizeof(cplx) * (size_t)p->n);
  dft_rec(out, scratch, 1, 0, p, p->n, 0);
}

/* Eigen::FFT::inv (src/CustomEigenFFT.h:91-128, 132-137): inverse then *= 1/n. */
static void fft_inv(const fftplan* p, cplx* out, const cplx* in, cplx* scratch) {
  memcpy(scratch, in, sizeof(cplx) * (size_t)p->n);
  dft_rec(out, scratch, 1, 0, p, p->n, 1);
  const double s = 1. / p->n;
  for (int k = 0; k < p->n; k++) out[k] *= s;
}

int orc_fft(int n, int count, const double* in, double* out, int inverse) {
  if (n <= 0) return fail(ORC_INVALID_ARGUMENT, "invalid transform length");
  fftplan p;
  fft_plan_init(&p, n);
  cplx* scratch = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  for (int r = 0; r < count; r++) {
    const cplx* x = (const cplx*)in + (long)r * n;
    cplx* y = (cplx*)out + (long)r * n;
    if (inverse) fft_inv(&p, y, x, scratch);
    else fft_fwd(&p, y, x, scratch);
  }
  free(scratch);
  fft_plan_free(&p);
  return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* medium                                                                    */
/* ------------------------------------------------------------------------- */

#define MAXM 4

struct orc_medium {
  int model, nS, nP, canBePoled, fnlm;
  int N, nZ, nZP;
  double z, dz, dzp, tMax, zR;
  int profile;
  int nNl;
  cplx nlStep[MAXM];
  int nDb;
  double diffBeta0[MAXM];
  double beta1[MAXM], beta2[MAXM];
  double *tau, *omega;
  double *dispP[MAXM], *dispS[MAXM];
  cplx *dstepP[MAXM], *dstepS[MAXM];
  cplx* env[MAXM];
  double* poling;
  int nPol;
  cplx *pumpFreq[MAXM], *pumpTime[MAXM];
  cplx *sigFreq[MAXM], *sigTime[MAXM];
  cplx* raman;
  cplx* gnlseFilter; /* _nlStep[0] + _nlStep[1] * _omega */
  fftplan fft;
};

static double* dalloc(long n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }
static cplx* calloc_c(long n) { return (cplx*)calloc((size_t)(n > 0 ? n : 1), sizeof(cplx)); }

/* first element of largest magnitude (std::max_element with |a|<|b|) */
static double max_abs_elem(const double* v, int n) {
  if (n == 0) return 0;
  double best = v[0];
  for (int i = 1; i < n; i++)
    if (fabs(best) < fabs(v[i])) best = v[i];
  return best;
}

/* _NonlinearMedium::setLengths, src/_NonlinearMedium.cpp:43-102 */
static int set_lengths(orc_medium* m, const orc_params* p, double zR) {
  int neg = (p->relative_length <= 0 || zR <= 0);
  for (int i = 0; i < p->n_nl; i++) neg |= (p->nl_length[i] <= 0);
  if (neg) return fail(ORC_INVALID_ARGUMENT, "Non-positive length scale");

  int allNonUnit = 1;
  for (int i = 0; i < p->n_beta1; i++) allNonUnit &= (fabs(p->beta1[i]) != 1);
  for (int i = 0; i < p->n_beta2; i++) allNonUnit &= (fabs(p->beta2[i]) != 1);
  for (int i = 0; i < p->n_beta3; i++) allNonUnit &= (fabs(p->beta3[i]) != 1);
  for (int i = 0; i < p->n_beta1s; i++) allNonUnit &= (fabs(p->beta1s[i]) != 1);
  for (int i = 0; i < p->n_beta2s; i++) allNonUnit &= (fabs(p->beta2s[i]) != 1);
  for (int i = 0; i < p->n_beta3s; i++) allNonUnit &= (fabs(p->beta3s[i]) != 1);
  for (int i = 0; i < p->n_nl; i++) allNonUnit &= (p->nl_length[i] != 1);
  allNonUnit &= (p->relative_length != 1);
  allNonUnit &= (zR != 1);
  if (allNonUnit) return fail(ORC_INVALID_ARGUMENT, "No unit length scale provided: please normalize variables");

  m->z = p->relative_length;

  /* std::max({...}, absComp) over the six per-list maxima, in this order */
  double cand[6] = {max_abs_elem(p->beta2, p->n_beta2), max_abs_elem(p->beta2s, p->n_beta2s),
                    max_abs_elem(p->beta1, p->n_beta1), max_abs_elem(p->beta1s, p->n_beta1s),
                    max_abs_elem(p->beta3, p->n_beta3), max_abs_elem(p->beta3s, p->n_beta3s)};
  double mx = max_abs_elem(cand, 6);
  double minDispLength = 1 / fabs(mx);

  double minNl = 0;
  if (p->n_nl > 0) {
    minNl = p->nl_length[0];
    for (int i = 1; i < p->n_nl; i++)
      if (p->nl_length[i] < minNl) minNl = p->nl_length[i];
  }
  double den = 1.;
  if (minDispLength < den) den = minDispLength;
  if (zR < den) den = zR;
  if (minNl < den) den = minNl;

  if (p->nz_override > 0)
    m->nZ = p->nz_override;
  else
    m->nZ = (int)(unsigned int)((double)(unsigned int)p->z_precision * m->z / den);
  m->nZP = 2 * m->nZ - 1;
  m->dz = m->z / (m->nZ - 1);
  m->dzp = m->z / (m->nZP - 1);

  m->nNl = p->n_nl;
  for (int i = 0; i < p->n_nl; i++) {
    /* 1._I / nlLength * _dz  ->  (0, 1/L) * dz */
    double inv = 1. / p->nl_length[i];
    m->nlStep[i] = CMPLX(0. * m->dz, inv * m->dz);
  }
  m->zR = zR;
  return ORC_OK;
}

/* _NonlinearMedium::fftshift, src/_NonlinearMedium.cpp:447-453 */
static void fftshift_d(double* v, int n) {
  int h = n / 2;
  double* tmp = dalloc(n);
  for (int i = 0; i < h; i++) tmp[i] = v[n - h + i];
  for (int i = 0; i < h; i++) tmp[n - h + i] = v[i];
  memcpy(v, tmp, sizeof(double) * (size_t)n);
  free(tmp);
}

/* _NonlinearMedium::resetGrids, src/_NonlinearMedium.cpp:105-145 */
static int reset_grids(orc_medium* m, int nFreqs, double tMax) {
  if (nFreqs % 2 != 0 || nFreqs == 0) return fail(ORC_INVALID_ARGUMENT, "Invalid number of Frequencies");
  if (m->nZ == 0) return fail(ORC_INVALID_ARGUMENT, "Zero steps");
  if (tMax <= 0) return fail(ORC_INVALID_ARGUMENT, "Negative time span");
  m->N = nFreqs;
  m->tMax = tMax;
  const int N = nFreqs;
  m->tau = dalloc(N);
  m->omega = dalloc(N);
  const double sT = 2 * tMax / N;         /* 2 * tMax / Nt */
  const double sW = -M_PI / tMax;          /* -M_PI / _tMax */
  for (int i = 0; i < N; i++) {
    double lin = (double)(-N / 2 + i);     /* LinSpaced(Nt, -Nt/2, Nt/2-1): exact integers */
    m->tau[i] = sT * lin;
    m->omega[i] = sW * lin;
  }
  fftshift_d(m->tau, N);
  fftshift_d(m->omega, N);
  for (int p = 0; p < m->nP; p++) {
    m->pumpFreq[p] = calloc_c((long)m->nZP * N);
    m->pumpTime[p] = calloc_c((long)m->nZP * N);
  }
  for (int s = 0; s < m->nS; s++) {
    m->sigFreq[s] = calloc_c((long)m->nZ * N);
    m->sigTime[s] = calloc_c((long)m->nZ * N);
  }
  return ORC_OK;
}

/* _NonlinearMedium::setDispersion, src/_NonlinearMedium.cpp:148-177 */
static void set_dispersion(orc_medium* m, const orc_params* p) {
  const int N = m->N;
  for (int i = 0; i < MAXM; i++) { m->beta1[i] = 0; m->beta2[i] = 0; }
  for (int i = 0; i < p->n_beta1 && i < MAXM; i++) m->beta1[i] = p->beta1[i];
  for (int i = 0; i < p->n_beta2 && i < MAXM; i++) m->beta2[i] = p->beta2[i];
  m->nDb = p->n_diff;
  for (int i = 0; i < p->n_diff && i < MAXM; i++) m->diffBeta0[i] = p->diff_beta0[i];
  for (int q = 0; q < m->nP; q++) {
    m->dispP[q] = dalloc(N);
    m->dstepP[q] = calloc_c(N);
    for (int k = 0; k < N; k++) {
      double w = m->omega[k];
      m->dispP[q][k] = w * (p->beta1[q] + w * (0.5 * p->beta2[q] + w * p->beta3[q] / 6));
      m->dstepP[q][k] = cis(m->dzp * m->dispP[q][k]);
    }
  }
  for (int s = 0; s < m->nS; s++) {
    m->dispS[s] = dalloc(N);
    m->dstepS[s] = calloc_c(N);
    for (int k = 0; k < N; k++) {
      double w = m->omega[k];
      m->dispS[s][k] = w * (p->beta1s[s] + w * (0.5 * p->beta2s[s] + w * p->beta3s[s] / 6));
      m->dstepS[s][k] = cis(m->dz * m->dispS[s][k]);
    }
  }
}

/* _NonlinearMedium::setPoling, src/_NonlinearMedium.cpp:470-503 */
static int set_poling(orc_medium* m, const double* pol, int n) {
  const int nZ = m->nZ;
  free(m->poling);
  m->poling = dalloc(nZ);
  m->nPol = nZ;
  if (n <= 1) {
    for (int i = 0; i < nZ; i++) m->poling[i] = 1;
    return ORC_OK;
  }
  for (int i = 0; i < n; i++)
    if (pol[i] <= 0) return fail(ORC_INVALID_ARGUMENT, "Poling contains invalid domain length");
  double* dom = dalloc(n);
  dom[0] = pol[0];
  for (int i = 1; i < n; i++) dom[i] = pol[i] + dom[i - 1];
  const double scale = nZ / dom[n - 1];
  for (int i = 0; i < n; i++) dom[i] *= scale;
  unsigned int prevInd = 0;
  int direction = 1;
  for (int i = 0; i < n; i++) {
    const double currInd = dom[i];
    const unsigned int currIndRound = (unsigned int)currInd;
    if (currInd < prevInd) { free(dom); return fail(ORC_INVALID_ARGUMENT, "Poling period too small for simulation resolution"); }
    for (unsigned int j = prevInd; j < currIndRound; j++) m->poling[j] = direction;
    if (currIndRound < (unsigned int)nZ)
      m->poling[currIndRound] = direction * (2 * fabs(fmod(currInd, 1)) - 1);
    direction *= -1;
    prevInd = currIndRound + 1;
  }
  free(dom);
  return ORC_OK;
}

/* chirp / delay phase of setPump, src/_NonlinearMedium.cpp:198-203, 216-221 */
static void apply_chirp_delay(orc_medium* m, int q, double chirp, double delay) {
  if (chirp == 0 && delay == 0) return;
  const int N = m->N;
  cplx* tmp = calloc_c(N);
  cplx* scr = calloc_c(N);
  fft_fwd(&m->fft, tmp, m->env[q], scr);
  for (int k = 0; k < N; k++) {
    double w = m->omega[k];
    double x = m->beta1[q] * delay + 0.5 * m->beta2[q] * chirp * w;
    tmp[k] *= cis(x * w);
  }
  fft_inv(&m->fft, m->env[q], tmp, scr);
  free(tmp);
  free(scr);
}

/* _NonlinearMedium::setPump(PulseType, ...), src/_NonlinearMedium.cpp:180-204 */
static int set_pump_type(orc_medium* m, int pulseType, double chirp, double delay, int q) {
  if (q < 0 || q >= m->nP) return fail(ORC_INVALID_ARGUMENT, "Invalid pump index");
  const int N = m->N;
  for (int k = 0; k < N; k++) {
    double t = m->tau[k];
    switch (pulseType) {
      default:
      case 0: m->env[q][k] = exp(-0.5 * (t * t)); break;
      case 1: m->env[q][k] = 1 / cosh(t); break;
      case 2: m->env[q][k] = sin(t) / t; break;
    }
  }
  if (pulseType == 2) m->env[q][0] = 1;
  apply_chirp_delay(m, q, chirp, delay);
  return ORC_OK;
}

/* _NonlinearMedium::setPump(customPump, ...), src/_NonlinearMedium.cpp:207-222 */
static int set_pump_custom(orc_medium* m, const double* pump, int n, double chirp, double delay, int q) {
  if (n != m->N) return fail(ORC_INVALID_ARGUMENT, "Custom pump array length does not match number of frequency/time bins");
  if (q < 0 || q >= m->nP) return fail(ORC_INVALID_ARGUMENT, "Invalid pump index");
  memcpy(m->env[q], pump, sizeof(cplx) * (size_t)n);
  apply_chirp_delay(m, q, chirp, delay);
  return ORC_OK;
}

int orc_set_pump_type(orc_medium* m, int pulse_type, double chirp, double delay, int pump_index) {
  if (m->fnlm) return fail(ORC_RUNTIME_ERROR, "Object does not have this method.");
  return set_pump_type(m, pulse_type, chirp, delay, pump_index);
}

int orc_set_pump_custom(orc_medium* m, const double* pump, int n, double chirp, double delay, int pump_index) {
  if (m->fnlm) return fail(ORC_RUNTIME_ERROR, "Object does not have this method.");
  return set_pump_custom(m, pump, n, chirp, delay, pump_index);
}

/* _NonlinearMedium::setPump(other, ...), src/_NonlinearMedium.cpp:506-534 */
int orc_set_pump_other(orc_medium* m, const orc_medium* o, int mode, double delayLength, int q) {
  if (m->fnlm) return fail(ORC_RUNTIME_ERROR, "Object does not have this method.");
  if (o->N != m->N || o->tMax != m->tMax)
    return fail(ORC_INVALID_ARGUMENT, "Medium does not have same time and frequency axes as this one");
  if (mode < 0 || mode >= m->nS) return fail(ORC_INVALID_ARGUMENT, "Mode index larger than number of modes in medium");
  if (q < 0 || q >= m->nP) return fail(ORC_INVALID_ARGUMENT, "Invalid pump index");
  if (o->nZ < m->nZ) return fail(ORC_INVALID_ARGUMENT, "Medium does not have sufficient resolution to be used with this one");
  const int N = m->N;
  cplx* scr = calloc_c(N);
  const double bd = m->beta1[q] * delayLength;
  for (int i = 0; i < m->nZP - 1; i++) {
    double j_ = i * ((double)(o->nZ - 1) / (m->nZP - 1));
    unsigned int j = (unsigned int)j_;
    double frac = j_ - j;
    const cplx* r0 = o->sigFreq[mode] + (long)j * N;
    const cplx* r1 = o->sigFreq[mode] + (long)(j + 1) * N;
    cplx* pf = m->pumpFreq[q] + (long)i * N;
    const double a = i * m->dzp, b = (j_ + 0.5) * o->dz;
    for (int k = 0; k < N; k++) {
      double th = a * m->dispP[q][k] - b * o->dispS[mode][k] + bd * m->omega[k];
      pf[k] = ((1 - frac) * r0[k] + frac * r1[k]) * cis(th);
    }
    fft_inv(&m->fft, m->pumpTime[q] + (long)i * N, pf, scr);
  }
  {
    const cplx* rl = o->sigFreq[mode] + (long)(o->nZ - 1) * N;
    cplx* pf = m->pumpFreq[q] + (long)(m->nZP - 1) * N;
    for (int k = 0; k < N; k++) {
      double th = m->z * m->dispP[q][k] - o->z * o->dispS[mode][k] + bd * m->omega[k];
      pf[k] = rl[k] * cis(th);
    }
    /* the reference transforms row _nZSteps-1 here, not the last pump row */
    fft_inv(&m->fft, m->pumpTime[q] + (long)(m->nZ - 1) * N, m->pumpFreq[q] + (long)(m->nZ - 1) * N, scr);
  }
  free(scr);
  return ORC_OK;
}

/* Eigen LinSpaced(n, low, high) for doubles */
static double linspaced(int n, double low, double high, int i) {
  if (n == 1) return high;
  const double step = (high - low) / (double)(n - 1);
  const int flip = fabs(high) < fabs(low);
  if (flip) return (i == 0) ? low : high - (double)(n - 1 - i) * step;
  return (i == n - 1) ? high : low + (double)i * step;
}

/* _NonlinearMedium::runPumpSimulation, src/_NonlinearMedium.cpp:225-250 */
static void run_pump_linear(orc_medium* m) {
  const int N = m->N;
  cplx* scr = calloc_c(N);
  for (int q = 0; q < m->nP; q++) {
    cplx* PF = m->pumpFreq[q];
    cplx* PT = m->pumpTime[q];
    fft_fwd(&m->fft, PF, m->env[q], scr);
    memcpy(PT, m->env[q], sizeof(cplx) * (size_t)N);
    for (int i = 1; i < m->nZP; i++) {
      const double a = i * m->dzp;
      for (int k = 0; k < N; k++) PF[(long)i * N + k] = PF[k] * cis(a * m->dispP[q][k]);
      fft_inv(&m->fft, PT + (long)i * N, PF + (long)i * N, scr);
    }
    if (m->profile != 1) {
      for (int i = 0; i < m->nZP; i++) {
        double x = linspaced(m->nZP, -0.5 * m->z, 0.5 * m->z, i) / m->zR;
        double s = (m->profile == 2) ? exp(-0.5 * (x * x)) : 1 / sqrt(1 + x * x);
        for (int k = 0; k < N; k++) {
          PF[(long)i * N + k] *= s;
          PT[(long)i * N + k] *= s;
        }
      }
    }
  }
  free(scr);
}

/* Chi3::runPumpSimulation, solvers/Chi3.cpp:25-54 */
static void run_pump_chi3(orc_medium* m) {
  const int N = m->N;
  cplx* scr = calloc_c(N);
  cplx* PF = m->pumpFreq[0];
  cplx* PT = m->pumpTime[0];
  fft_fwd(&m->fft, PF, m->env[0], scr);
  for (int k = 0; k < N; k++) PF[k] *= cis((0.5 * m->dzp) * m->dispP[0][k]);
  fft_inv(&m->fft, PT, PF, scr);

  cplx* relI = calloc_c(m->nZP);
  const cplx c0 = m->dzp / m->dz * m->nlStep[0];
  for (int i = 0; i < m->nZP; i++) {
    double x = linspaced(m->nZP, -0.5 * m->z, 0.5 * m->z, i) / m->zR;
    switch (m->profile) {
      default:
      case 1: relI[i] = c0; break;
      case 0: relI[i] = c0 / (1 + x * x); break;
      case 2: relI[i] = c0 * exp(-(x * x)); break;
    }
  }
  for (int i = 1; i < m->nZP; i++) {
    const cplx* prev = PT + (long)(i - 1) * N;
    cplx* cur = PT + (long)i * N;
    for (int k = 0; k < N; k++) {
      double a2 = creal(prev[k]) * creal(prev[k]) + cimag(prev[k]) * cimag(prev[k]);
      cur[k] = prev[k] * cexp(relI[i - 1] * a2);
    }
    fft_fwd(&m->fft, PF + (long)i * N, cur, scr);
    for (int k = 0; k < N; k++) PF[(long)i * N + k] *= m->dstepP[0][k];
    fft_inv(&m->fft, cur, PF + (long)i * N, scr);
  }
  cplx* last = PF + (long)(m->nZP - 1) * N;
  for (int k = 0; k < N; k++) last[k] *= cis((-0.5 * m->dzp) * m->dispP[0][k]);
  fft_inv(&m->fft, PT + (long)(m->nZP - 1) * N, last, scr);
  free(relI);
  free(scr);
}

int orc_run_pump(orc_medium* m) {
  if (m->fnlm) return fail(ORC_RUNTIME_ERROR, "Object does not have this method.");
  if (m->model == ORC_CHI3) run_pump_chi3(m);
  else run_pump_linear(m);
  return ORC_OK;
}

/* _FullyNonlinearMedium::relativeAmplitude / relativeIntensity,
 * src/_FullyNonlinearMedium.hpp:40-63 */
static double rel_amp(const orc_medium* m, double i) {
  switch (m->profile) {
    case 0: return sqrt(1 / (1 + pow((i * m->dz - 0.5 * m->z) / m->zR, 2)));
    case 1: return 1.;
    case 2: return exp(-0.5 * pow((i * m->dz - 0.5 * m->z) / m->zR, 2));
    default: return 0;
  }
}
static double rel_int(const orc_medium* m, double i) {
  switch (m->profile) {
    case 0: return 1 / (1 + pow((i * m->dz - 0.5 * m->z) / m->zR, 2));
    case 1: return 1.;
    case 2: return exp(-pow((i * m->dz - 0.5 * m->z) / m->zR, 2));
    default: return 0;
  }
}

static inline double abs2(cplx x) { return creal(x) * creal(x) + cimag(x) * cimag(x); }

/* per-thread work buffers for one propagation */
typedef struct {
  cplx* k[4][MAXM];  /* k1..k4 per mode */
  cplx* tmp;         /* Chi3GNLSE convolution scratch */
  cplx* scr;         /* FFT scratch */
} work_t;

static void work_init(work_t* w, const orc_medium* m) {
  for (int s = 0; s < 4; s++)
    for (int q = 0; q < MAXM; q++) w->k[s][q] = calloc_c(m->N);
  w->tmp = calloc_c(m->N);
  w->scr = calloc_c(m->N);
}
static void work_free(work_t* w) {
  for (int s = 0; s < 4; s++)
    for (int q = 0; q < MAXM; q++) free(w->k[s][q]);
  free(w->tmp);
  free(w->scr);
}

#define FOR_K for (int k = 0; k < N; k++)

/* Mismatch factors exp(i * db * x) for x = (i-1)dz, (i-.5)dz, i*dz. */
static void mismatch3(const orc_medium* m, unsigned int i, double db, cplx out[3]) {
  out[0] = cis(db * ((i - 1) * m->dz));
  out[1] = cis(db * ((i - .5) * m->dz));
  out[2] = cis(db * (i * m->dz));
}

/* The model DiffEqs.  `a[q]` is the current signal row of mode q and P0, P1
 * the pump time grids.  Each block cites the reference DiffEq it restates. */
static void diffeq(const orc_medium* m, unsigned int i, cplx* const* a, work_t* w) {
  const int N = m->N;
  cplx **k1 = w->k[0], **k2 = w->k[1], **k3 = w->k[2], **k4 = w->k[3];
  const cplx* nl = m->nlStep;
  double pp = 0, pc = 0, pm = 0;
  if (m->nPol > 0) {
    pp = m->poling[i - 1];
    pc = m->poling[i];
    pm = 0.5 * (pp + pc);
  }
  const cplx *P0p = NULL, *P0i = NULL, *P0c = NULL, *P1p = NULL, *P1i = NULL, *P1c = NULL;
  if (m->nP > 0) {
    P0p = m->pumpTime[0] + (long)(2 * i - 2) * N;
    P0i = m->pumpTime[0] + (long)(2 * i - 1) * N;
    P0c = m->pumpTime[0] + (long)(2 * i) * N;
  }
  if (m->nP > 1) {
    P1p = m->pumpTime[1] + (long)(2 * i - 2) * N;
    P1i = m->pumpTime[1] + (long)(2 * i - 1) * N;
    P1c = m->pumpTime[1] + (long)(2 * i) * N;
  }

  switch (m->model) {
    case ORC_CHI3: { /* solvers/Chi3.cpp:57-69 */
      const cplx* x = a[0];
      FOR_K k1[0][k] = nl[0] * (2 * abs2(P0p[k]) * x[k] + P0p[k] * P0p[k] * conj(x[k]));
      FOR_K { cplx y = x[k] + 0.5 * k1[0][k]; k2[0][k] = nl[0] * (2 * abs2(P0i[k]) * y + P0i[k] * P0i[k] * conj(y)); }
      FOR_K { cplx y = x[k] + 0.5 * k2[0][k]; k3[0][k] = nl[0] * (2 * abs2(P0i[k]) * y + P0i[k] * P0i[k] * conj(y)); }
      FOR_K { cplx y = x[k] + k3[0][k];       k4[0][k] = nl[0] * (2 * abs2(P0c[k]) * y + P0c[k] * P0c[k] * conj(y)); }
    } break;

    case ORC_CHI2PDC: { /* solvers/Chi2PDC.cpp:27-47 */
      cplx mis[3];
      mismatch3(m, i, m->diffBeta0[0], mis);
      const cplx* x = a[0];
      const cplx c1 = pp * nl[0] * mis[0], c2 = pm * nl[0] * mis[1], c4 = pc * nl[0] * mis[2];
      FOR_K k1[0][k] = c1 * P0p[k] * conj(x[k]);
      FOR_K k2[0][k] = c2 * P0i[k] * conj(x[k] + 0.5 * k1[0][k]);
      FOR_K k3[0][k] = c2 * P0i[k] * conj(x[k] + 0.5 * k2[0][k]);
      FOR_K k4[0][k] = c4 * P0c[k] * conj(x[k] + k3[0][k]);
    } break;

    case ORC_CHI2SFG: { /* solvers/Chi2SFG.cpp:28-56 */
      cplx mis[3];
      mismatch3(m, i, m->diffBeta0[0], mis);
      const cplx inv0 = 1. / mis[0], inv1 = 1. / mis[1], inv2 = 1. / mis[2];
      const cplx *S = a[0], *O = a[1];
      const cplx a0 = pp * nl[0] * mis[0], b0 = pp * nl[1] * inv0;
      const cplx a1 = pm * nl[0] * mis[1], b1 = pm * nl[1] * inv1;
      const cplx a2 = pc * nl[0] * mis[2], b2 = pc * nl[1] * inv2;
      FOR_K { k1[0][k] = a0 * P0p[k] * O[k]; k1[1][k] = b0 * conj(P0p[k]) * S[k]; }
      FOR_K { k2[0][k] = a1 * P0i[k] * (O[k] + 0.5 * k1[1][k]); k2[1][k] = b1 * conj(P0i[k]) * (S[k] + 0.5 * k1[0][k]); }
      FOR_K { k3[0][k] = a1 * P0i[k] * (O[k] + 0.5 * k2[1][k]); k3[1][k] = b1 * conj(P0i[k]) * (S[k] + 0.5 * k2[0][k]); }
      FOR_K { k4[0][k] = a2 * P0c[k] * (O[k] + k3[1][k]); k4[1][k] = b2 * conj(P0c[k]) * (S[k] + k3[0][k]); }
    } break;

    case ORC_CHI2AFC: { /* solvers/Chi2AFC.cpp:26-51 */
      const double arg = 0.5 * (m->diffBeta0[1] - m->diffBeta0[0]) / m->z;
      const double d0 = m->diffBeta0[0];
      const double x0 = (i - 1) * m->dz, x1 = (i - .5) * m->dz, x2 = i * m->dz;
      /* exp(0.5_I * ((d0 + (i-1)*dz*arg) * (i-1)*dz)) */
      const cplx mis0 = cis(0.5 * ((d0 + (i - 1) * m->dz * arg) * (i - 1) * m->dz));
      const cplx mis1 = cis(0.5 * ((d0 + (i - .5) * m->dz * arg) * (i - .5) * m->dz));
      const cplx mis2 = cis(0.5 * ((d0 + i * m->dz * arg) * i * m->dz));
      (void)x0; (void)x1; (void)x2;
      const cplx inv0 = 1. / mis0, inv1 = 1. / mis1, inv2 = 1. / mis2;
      const cplx *S = a[0], *O = a[1];
      const cplx a0 = nl[0] * mis0, b0 = nl[1] * inv0;
      const cplx a1 = nl[0] * mis1, b1 = nl[1] * inv1;
      const cplx a2 = nl[0] * mis2, b2 = nl[1] * inv2;
      FOR_K { k1[0][k] = a0 * P0p[k] * O[k]; k1[1][k] = b0 * conj(P0p[k]) * S[k]; }
      FOR_K { k2[0][k] = a1 * P0i[k] * (O[k] + 0.5 * k1[1][k]); k2[1][k] = b1 * conj(P0i[k]) * (S[k] + 0.5 * k1[0][k]); }
      FOR_K { k3[0][k] = a1 * P0i[k] * (O[k] + 0.5 * k2[1][k]); k3[1][k] = b1 * conj(P0i[k]) * (S[k] + 0.5 * k2[0][k]); }
      FOR_K { k4[0][k] = a2 * P0c[k] * (O[k] + k3[1][k]); k4[1][k] = b2 * conj(P0c[k]) * (S[k] + k3[0][k]); }
    } break;

    case ORC_CHI2SFGPDC: { /* solvers/Chi2SFGPDC.cpp:28-59 */
      cplx mis[3], miso[3];
      mismatch3(m, i, m->diffBeta0[0], mis);
      mismatch3(m, i, m->diffBeta0[1], miso);
      const cplx inv0 = 1. / mis[0], inv1 = 1. / mis[1], inv2 = 1. / mis[2];
      const cplx *S = a[0], *O = a[1];
      const cplx a0 = pp * nl[0] * mis[0], a1 = pm * nl[0] * mis[1], a2 = pc * nl[0] * mis[2];
      const cplx b0 = pp * nl[1], b1 = pm * nl[1], b2 = pc * nl[1];
      FOR_K {
        k1[0][k] = a0 * P0p[k] * O[k];
        k1[1][k] = b0 * (inv0 * conj(P0p[k]) * S[k] + miso[0] * P0p[k] * conj(O[k]));
      }
      FOR_K {
        k2[0][k] = a1 * P0i[k] * (O[k] + 0.5 * k1[1][k]);
        k2[1][k] = b1 * (inv1 * conj(P0i[k]) * (S[k] + 0.5 * k1[0][k]) + miso[1] * P0i[k] * conj(O[k] + 0.5 * k1[1][k]));
      }
      FOR_K {
        k3[0][k] = a1 * P0i[k] * (O[k] + 0.5 * k2[1][k]);
        k3[1][k] = b1 * (inv1 * conj(P0i[k]) * (S[k] + 0.5 * k2[0][k]) + miso[1] * P0i[k] * conj(O[k] + 0.5 * k2[1][k]));
      }
      FOR_K {
        k4[0][k] = a2 * P0c[k] * (O[k] + k3[1][k]);
        k4[1][k] = b2 * (inv2 * conj(P0c[k]) * (S[k] + k3[0][k]) + miso[2] * P0c[k] * conj(O[k] + k3[1][k]));
      }
    } break;

    case ORC_CHI2SFGII: { /* solvers/Chi2SFGII.cpp:34-90 (k4 uses prevPolDir) */
      cplx cz[3], cy[3], sz[3];
      mismatch3(m, i, m->diffBeta0[0], cz);
      mismatch3(m, i, m->diffBeta0[1], cy);
      mismatch3(m, i, m->diffBeta0[2], sz);
      cplx icz[3], icy[3], isz[3];
      for (int j = 0; j < 3; j++) { icz[j] = 1. / cz[j]; icy[j] = 1. / cy[j]; isz[j] = 1. / sz[j]; }
      const cplx *Sz = a[0], *Sy = a[1], *Oz = a[2], *Oy = a[3];
      FOR_K {
        k1[0][k] = pp * P0p[k] * (nl[0] * cz[0] * Oy[k] + nl[1] * sz[0] * Oz[k]);
        k1[1][k] = (pp * nl[0] * cy[0]) * P0p[k] * Oz[k];
        k1[2][k] = pp * conj(P0p[k]) * (nl[0] * icy[0] * Sy[k] + nl[1] * isz[0] * Sz[k]);
        k1[3][k] = (pp * nl[0] * icz[0]) * conj(P0p[k]) * Sz[k];
      }
      FOR_K {
        k2[0][k] = pm * P0i[k] * (nl[0] * cz[1] * (Oy[k] + 0.5 * k1[3][k]) + nl[1] * sz[1] * (Oz[k] + 0.5 * k1[2][k]));
        k2[1][k] = (pm * nl[0] * cy[1]) * P0i[k] * (Oz[k] + 0.5 * k1[2][k]);
        k2[2][k] = pm * conj(P0i[k]) * (nl[0] * icy[1] * (Sy[k] + 0.5 * k1[1][k]) + nl[1] * isz[1] * (Sz[k] + 0.5 * k1[0][k]));
        k2[3][k] = (pm * nl[0] * icz[1]) * conj(P0i[k]) * (Sz[k] + 0.5 * k1[0][k]);
      }
      FOR_K {
        k3[0][k] = pm * P0i[k] * (nl[0] * cz[1] * (Oy[k] + 0.5 * k2[3][k]) + nl[1] * sz[1] * (Oz[k] + 0.5 * k2[2][k]));
        k3[1][k] = (pm * nl[0] * cy[1]) * P0i[k] * (Oz[k] + 0.5 * k2[2][k]);
        k3[2][k] = pm * conj(P0i[k]) * (nl[0] * icy[1] * (Sy[k] + 0.5 * k2[1][k]) + nl[1] * isz[1] * (Sz[k] + 0.5 * k2[0][k]));
        k3[3][k] = (pm * nl[0] * icz[1]) * conj(P0i[k]) * (Sz[k] + 0.5 * k2[0][k]);
      }
      FOR_K {
        k4[0][k] = pp * P0c[k] * (nl[0] * cz[2] * (Oy[k] + k3[3][k]) + nl[1] * sz[2] * (Oz[k] + k3[2][k]));
        k4[1][k] = (pp * nl[0] * cy[2]) * P0c[k] * (Oz[k] + k3[2][k]);
        k4[2][k] = pp * conj(P0c[k]) * (nl[0] * icy[2] * (Sy[k] + k3[1][k]) + nl[1] * isz[2] * (Sz[k] + k3[0][k]));
        k4[3][k] = (pp * nl[0] * icz[2]) * conj(P0c[k]) * (Sz[k] + k3[0][k]);
      }
    } break;

    case ORC_CHI2PDCII: { /* solvers/Chi2PDCII.cpp:29-55 */
      cplx mis[3];
      mismatch3(m, i, m->diffBeta0[0], mis);
      const cplx *S = a[0], *O = a[1];
      const cplx a0 = pp * nl[0] * mis[0], b0 = pp * nl[1] * mis[0];
      const cplx a1 = pm * nl[0] * mis[1], b1 = pm * nl[1] * mis[1];
      const cplx a2 = pc * nl[0] * mis[2], b2 = pc * nl[1] * mis[2];
      FOR_K { k1[0][k] = a0 * P0p[k] * conj(O[k]); k1[1][k] = b0 * P0p[k] * conj(S[k]); }
      FOR_K { k2[0][k] = a1 * P0i[k] * conj(O[k] + 0.5 * k1[1][k]); k2[1][k] = b1 * P0i[k] * conj(S[k] + 0.5 * k1[0][k]); }
      FOR_K { k3[0][k] = a1 * P0i[k] * conj(O[k] + 0.5 * k2[1][k]); k3[1][k] = b1 * P0i[k] * conj(S[k] + 0.5 * k2[0][k]); }
      FOR_K { k4[0][k] = a2 * P0c[k] * conj(O[k] + k3[1][k]); k4[1][k] = b2 * P0c[k] * conj(S[k] + k3[0][k]); }
    } break;

    case ORC_CHI2SFGOPA: { /* solvers/Chi2SFGOPA.cpp:31-77 */
      cplx ms[3], md[3];
      mismatch3(m, i, m->diffBeta0[0], ms);
      mismatch3(m, i, m->diffBeta0[1], md);
      const cplx is0 = 1. / ms[0], is1 = 1. / ms[1], is2 = 1. / ms[2];
      const cplx *H = a[0], *F = a[1];
      FOR_K {
        k1[0][k] = (pp * nl[0] * ms[0]) * P0p[k] * F[k] + (pp * nl[2] * md[0]) * P1p[k] * conj(F[k]);
        k1[1][k] = (pp * nl[1] * is0) * conj(P0p[k]) * H[k] + (pp * nl[3] * md[0]) * P1p[k] * conj(H[k]);
      }
      FOR_K {
        cplx f = F[k] + 0.5 * k1[1][k], h = H[k] + 0.5 * k1[0][k];
        k2[0][k] = (pm * nl[0] * ms[1]) * P0i[k] * f + (pm * nl[2] * md[1]) * P1i[k] * conj(f);
        k2[1][k] = (pm * nl[1] * is1) * conj(P0i[k]) * h + (pm * nl[3] * md[1]) * P1i[k] * conj(h);
      }
      FOR_K {
        cplx f = F[k] + 0.5 * k2[1][k], h = H[k] + 0.5 * k2[0][k];
        k3[0][k] = (pm * nl[0] * ms[1]) * P0i[k] * f + (pm * nl[2] * md[1]) * P1i[k] * conj(f);
        k3[1][k] = (pm * nl[1] * is1) * conj(P0i[k]) * h + (pm * nl[3] * md[1]) * P1i[k] * conj(h);
      }
      FOR_K {
        cplx f = F[k] + k3[1][k], h = H[k] + k3[0][k];
        k4[0][k] = (pc * nl[0] * ms[2]) * P0c[k] * f + (pc * nl[2] * md[2]) * P1c[k] * conj(f);
        k4[1][k] = (pc * nl[1] * is2) * conj(P0c[k]) * h + (pc * nl[3] * md[2]) * P1c[k] * conj(h);
      }
    } break;

    case ORC_CHI2SFGXPM: { /* solvers/Chi2SFGXPM.cpp:26-55 */
      cplx mis[3];
      mismatch3(m, i, m->diffBeta0[0], mis);
      const cplx inv0 = 1. / mis[0], inv1 = 1. / mis[1], inv2 = 1. / mis[2];
      const cplx *S = a[0], *O = a[1];
      FOR_K {
        k1[0][k] = ((pp * nl[0] * mis[0]) * P0p[k] + nl[2] * abs2(P0p[k])) * O[k];
        k1[1][k] = ((pp * nl[1] * inv0) * conj(P0p[k]) + nl[3] * abs2(P0p[k])) * S[k];
      }
      FOR_K {
        k2[0][k] = ((pm * nl[0] * mis[1]) * P0i[k] + nl[2] * abs2(P0i[k])) * (O[k] + 0.5 * k1[1][k]);
        k2[1][k] = ((pm * nl[1] * inv1) * conj(P0i[k]) + nl[3] * abs2(P0i[k])) * (S[k] + 0.5 * k1[0][k]);
      }
      FOR_K {
        k3[0][k] = ((pm * nl[0] * mis[1]) * P0i[k] + nl[2] * abs2(P0i[k])) * (O[k] + 0.5 * k2[1][k]);
        k3[1][k] = ((pm * nl[1] * inv1) * conj(P0i[k]) + nl[3] * abs2(P0i[k])) * (S[k] + 0.5 * k2[0][k]);
      }
      FOR_K {
        k4[0][k] = ((pc * nl[0] * mis[2]) * P0c[k] + nl[2] * abs2(P0c[k])) * (O[k] + k3[1][k]);
        k4[1][k] = ((pc * nl[1] * inv2) * conj(P0c[k]) + nl[3] * abs2(P0c[k])) * (S[k] + k3[0][k]);
      }
    } break;

    case ORC_CHI2SHG:   /* solvers/Chi2SHG.cpp:25-60 */
    case ORC_CHI2ASHG: { /* solvers/Chi2ASHG.cpp:24-56 */
      double rp, ri, rc;
      cplx mis[3];
      if (m->model == ORC_CHI2SHG) {
        rp = pp * rel_amp(m, i - 1);
        ri = pm * rel_amp(m, i - .5);
        rc = pc * rel_amp(m, i);
        mismatch3(m, i, m->diffBeta0[0], mis);
      } else {
        rp = rel_amp(m, i - 1);
        ri = rel_amp(m, i - .5);
        rc = rel_amp(m, i);
        const double arg = 0.5 * (m->diffBeta0[1] - m->diffBeta0[0]) / m->z;
        const double d0 = m->diffBeta0[0];
        mis[0] = cis(0.5 * ((d0 + (i - 1) * m->dz * arg) * (i - 1) * m->dz));
        mis[1] = cis(0.5 * ((d0 + (i - .5) * m->dz * arg) * (i - .5) * m->dz));
        mis[2] = cis(0.5 * ((d0 + i * m->dz * arg) * i * m->dz));
      }
      const cplx inv0 = 1. / mis[0], inv1 = 1. / mis[1], inv2 = 1. / mis[2];
      const cplx *A = a[0], *B = a[1];
      const cplx a0 = rp * nl[0] * mis[0], b0 = rp * nl[1] * inv0 * .5;
      const cplx a1 = ri * nl[0] * mis[1], b1 = ri * nl[1] * inv1 * .5;
      const cplx a2 = rc * nl[0] * mis[2], b2 = rc * nl[1] * inv2 * .5;
      FOR_K { k1[0][k] = a0 * conj(A[k]) * B[k]; k1[1][k] = b0 * (A[k] * A[k]); }
      FOR_K {
        cplx x = A[k] + .5 * k1[0][k];
        k2[0][k] = a1 * conj(x) * (B[k] + .5 * k1[1][k]);
        k2[1][k] = b1 * (x * x);
      }
      FOR_K {
        cplx x = A[k] + .5 * k2[0][k];
        k3[0][k] = a1 * conj(x) * (B[k] + .5 * k2[1][k]);
        k3[1][k] = b1 * (x * x);
      }
      FOR_K {
        cplx x = A[k] + k3[0][k];
        k4[0][k] = a2 * conj(x) * (B[k] + k3[1][k]);
        k4[1][k] = b2 * (x * x);
      }
    } break;

    case ORC_CHI2SHGOPA: { /* solvers/Chi2SHGOPA.cpp:32-84 */
      const double rp = pp * rel_amp(m, i - 1), ri = pm * rel_amp(m, i - .5), rc = pc * rel_amp(m, i);
      cplx ms[3], mo[3];
      mismatch3(m, i, m->diffBeta0[0], ms);
      mismatch3(m, i, m->diffBeta0[1], mo);
      cplx is[3], io[3];
      for (int j = 0; j < 3; j++) { is[j] = 1. / ms[j]; io[j] = 1. / mo[j]; }
      const double r[3] = {rp, ri, rc};
      const cplx *Pp = a[0], *SH = a[1], *A1 = a[2], *A2 = a[3];
      for (int st = 0; st < 4; st++) {
        const int j = st == 0 ? 0 : (st == 3 ? 2 : 1);
        const double h = st == 0 ? 0 : (st == 3 ? 1. : .5);
        cplx** kp = w->k[st == 0 ? 0 : st - 1];
        cplx** ko = w->k[st];
        const cplx c0 = r[j] * nl[0] * ms[j], c1 = r[j] * nl[1], c2 = r[j] * nl[2] * io[j], c3 = r[j] * nl[3] * io[j];
        FOR_K {
          cplx p_, s_, x1, x2;
          if (st == 0) { p_ = Pp[k]; s_ = SH[k]; x1 = A1[k]; x2 = A2[k]; }
          else {
            p_ = Pp[k] + h * kp[0][k]; s_ = SH[k] + h * kp[1][k];
            x1 = A1[k] + h * kp[2][k]; x2 = A2[k] + h * kp[3][k];
          }
          cplx o0 = c0 * conj(p_) * s_;
          cplx o1 = st == 0 ? c1 * (mo[j] * x2 * x1 + is[j] * .5 * (p_ * p_))
                            : c1 * (mo[j] * x1 * x2 + is[j] * .5 * (p_ * p_));
          cplx o2 = c2 * s_ * conj(x2);
          cplx o3 = c3 * s_ * conj(x1);
          ko[0][k] = o0; ko[1][k] = o1; ko[2][k] = o2; ko[3][k] = o3;
        }
      }
    } break;

    case ORC_CHI2DSFG: { /* solvers/Chi2DSFG.cpp:25-65 */
      const double rp = pp * rel_amp(m, i - 1), ri = pm * rel_amp(m, i - .5), rc = pc * rel_amp(m, i);
      cplx mis[3];
      mismatch3(m, i, m->diffBeta0[0], mis);
      const cplx inv[3] = {1. / mis[0], 1. / mis[1], 1. / mis[2]};
      const double r[3] = {rp, ri, rc};
      const cplx *Pv = a[0], *Sv = a[1], *Dv = a[2];
      for (int st = 0; st < 4; st++) {
        const int j = st == 0 ? 0 : (st == 3 ? 2 : 1);
        const double h = st == 0 ? 0 : (st == 3 ? 1. : .5);
        cplx** kp = w->k[st == 0 ? 0 : st - 1];
        cplx** ko = w->k[st];
        const cplx c0 = r[j] * nl[0] * mis[j], c1 = r[j] * nl[1] * inv[j], c2 = r[j] * nl[2] * mis[j];
        FOR_K {
          cplx p_, s_, d_;
          if (st == 0) { p_ = Pv[k]; s_ = Sv[k]; d_ = Dv[k]; }
          else { p_ = Pv[k] + h * kp[0][k]; s_ = Sv[k] + h * kp[1][k]; d_ = Dv[k] + h * kp[2][k]; }
          cplx o0 = c0 * s_ * conj(d_);
          cplx o1 = c1 * p_ * d_;
          cplx o2 = c2 * s_ * conj(p_);
          ko[0][k] = o0; ko[1][k] = o1; ko[2][k] = o2;
        }
      }
    } break;

    case ORC_CHI2SHGXPM: { /* solvers/Chi2SHGXPM.cpp:25-65 */
      const double ap = rel_amp(m, i - 1), ai = rel_amp(m, i - .5), ac = rel_amp(m, i);
      const double r[3] = {pp * ap, pm * ai, pc * ac};
      const cplx c3[3] = {pow(ap, 2) * nl[2], pow(ai, 2) * nl[2], pow(ac, 2) * nl[2]};
      cplx mis[3];
      mismatch3(m, i, m->diffBeta0[0], mis);
      const cplx inv[3] = {1. / mis[0], 1. / mis[1], 1. / mis[2]};
      const cplx *A = a[0], *B = a[1];
      for (int st = 0; st < 4; st++) {
        const int j = st == 0 ? 0 : (st == 3 ? 2 : 1);
        const double h = st == 0 ? 0 : (st == 3 ? 1. : .5);
        cplx** kp = w->k[st == 0 ? 0 : st - 1];
        cplx** ko = w->k[st];
        const cplx cA = r[j] * nl[0] * mis[j], cB = r[j] * nl[1] * inv[j] * .5;
        FOR_K {
          cplx x, y;
          if (st == 0) { x = A[k]; y = B[k]; }
          else { x = A[k] + h * kp[0][k]; y = B[k] + h * kp[1][k]; }
          cplx o0 = cA * conj(x) * y + c3[j] * (abs2(x) + 2 * abs2(y)) * x;
          cplx o1 = cB * (x * x) + 2. * c3[j] * (abs2(y) + 2 * abs2(x)) * y;
          ko[0][k] = o0; ko[1][k] = o1;
        }
      }
    } break;

    case ORC_CHI3GNLSE: { /* solvers/Chi3GNLSE.cpp:51-92 */
      const double ri[3] = {rel_int(m, i - 1), rel_int(m, i - .5), rel_int(m, i)};
      const cplx* x = a[0];
      for (int st = 0; st < 4; st++) {
        const int j = st == 0 ? 0 : (st == 3 ? 2 : 1);
        const double h = st == 0 ? 0 : (st == 3 ? 1. : .5);
        cplx* kp = w->k[st == 0 ? 0 : st - 1][0];
        cplx* ko = w->k[st][0];
        cplx* y = w->k[(st + 1) % 4][0]; /* unused stage buffer as B = a + h k */
        if (st == 3) y = w->k[0][1];
        FOR_K y[k] = (st == 0) ? x[k] : x[k] + h * kp[k];
        FOR_K ko[k] = ri[j] * abs2(y[k]);
        fft_fwd(&m->fft, w->tmp, ko, w->scr);
        FOR_K w->tmp[k] *= m->raman[k];
        fft_inv(&m->fft, ko, w->tmp, w->scr);
        FOR_K ko[k] *= y[k];
        fft_fwd(&m->fft, w->tmp, ko, w->scr);
        FOR_K w->tmp[k] *= m->gnlseFilter[k];
        fft_inv(&m->fft, ko, w->tmp, w->scr);
      }
    } break;
  }
}

/* _NonlinearMedium::signalSimulationTemplate, src/_NonlinearMedium.hpp:160-239.
 * sigF/sigT: per-mode grids with `rows` rows (1 when optimized). */
static void signal_sim(const orc_medium* m, const cplx* in, int inLen, int inTime, int inputMode,
                       cplx** sigF, cplx** sigT, int rows, int optimized, work_t* w) {
  const int N = m->N, nS = m->nS;
  const int nCh = inLen / N;
  if (nCh > 1) inputMode = 0;
  if (nS <= 1) inputMode = 0;

  if (inTime) {
    for (int q = 0; q < nS; q++) {
      if (q == inputMode || (inputMode < 1 && q < nCh)) {
        const cplx* src = in + (q == inputMode ? 0 : (long)q * N);
        memmove(sigT[q], src, sizeof(cplx) * (size_t)N);
        fft_fwd(&m->fft, sigF[q], sigT[q], w->scr);
        for (int k = 0; k < N; k++) sigF[q][k] *= cis((0.5 * m->dz) * m->dispS[q][k]);
      } else {
        memset(sigF[q], 0, sizeof(cplx) * (size_t)N);
      }
    }
  } else {
    for (int q = 0; q < nS; q++) {
      if (q == inputMode || (inputMode < 1 && q < nCh)) {
        const cplx* src = in + (q == inputMode ? 0 : (long)q * N);
        for (int k = 0; k < N; k++) sigF[q][k] = src[k] * cis((0.5 * m->dz) * m->dispS[q][k]);
      } else {
        memset(sigF[q], 0, sizeof(cplx) * (size_t)N);
      }
    }
  }
  for (int q = 0; q < nS; q++) {
    if (q == inputMode || q < nCh) fft_inv(&m->fft, sigT[q], sigF[q], w->scr);
    else memset(sigT[q], 0, sizeof(cplx) * (size_t)N);
  }

  cplx* cur[MAXM];
  for (unsigned int i = 1; i < (unsigned int)m->nZ; i++) {
    const long rp = optimized ? 0 : (long)(i - 1) * N;
    const long rc = optimized ? 0 : (long)i * N;
    for (int q = 0; q < nS; q++) cur[q] = sigT[q] + rp;
    diffeq(m, i, cur, w);
    for (int q = 0; q < nS; q++) {
      cplx* T = sigT[q] + rc;
      const cplx* Tp = sigT[q] + rp;
      const cplx *k1 = w->k[0][q], *k2 = w->k[1][q], *k3 = w->k[2][q], *k4 = w->k[3][q];
      for (int k = 0; k < N; k++) T[k] = Tp[k] + (k1[k] + 2 * k2[k] + 2 * k3[k] + k4[k]) * (1. / 6.);
      cplx* F = sigF[q] + rc;
      fft_fwd(&m->fft, F, T, w->scr);
      for (int k = 0; k < N; k++) F[k] *= m->dstepS[q][k];
      fft_inv(&m->fft, T, F, w->scr);
    }
  }
  for (int q = 0; q < nS; q++) {
    cplx* F = sigF[q] + (long)(rows - 1) * N;
    for (int k = 0; k < N; k++) F[k] *= cis((-0.5 * m->dz) * m->dispS[q][k]);
    fft_inv(&m->fft, sigT[q] + (long)(rows - 1) * N, F, w->scr);
  }
}

/* _NonlinearMedium::runSignalSimulation, src/_NonlinearMedium.cpp:253-260 */
int orc_run_signal(orc_medium* m, const double* in, int in_len, int in_time_domain, int input_mode) {
  const int N = m->N;
  if (in_len % N != 0 || in_len / N == 0 || in_len / N > m->nS)
    return fail(ORC_INVALID_ARGUMENT, "inputProf array size does not match number of frequency/time bins");
  if (input_mode < 0 || input_mode >= m->nS)
    return fail(ORC_INVALID_ARGUMENT, "inputModes does not match any mode in the system");
  work_t w;
  work_init(&w, m);
  signal_sim(m, (const cplx*)in, in_len, in_time_domain, input_mode, m->sigFreq, m->sigTime, m->nZ, 0, &w);
  work_free(&w);
  return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Green's function and batch drivers (std::thread fork/join)               */
/* ------------------------------------------------------------------------- */

typedef struct {
  const orc_medium* m;
  long start, stop;
  int inTime;
  int nIn, nOut;
  const int* inputs;
  const int* outputs;
  cplx *C, *S;          /* (nIn*N) x (nOut*N), pre-transpose */
  /* batch */
  const cplx* bin;
  int bcols, inputMode;
  cplx* bout;
  int batch;
} job_t;

static void* job_run(void* arg) {
  job_t* J = (job_t*)arg;
  const orc_medium* m = J->m;
  const int N = m->N, nS = m->nS;
  work_t w;
  work_init(&w, m);
  cplx *gF[MAXM], *gT[MAXM];
  for (int q = 0; q < nS; q++) { gF[q] = calloc_c(N); gT[q] = calloc_c(N); }
  cplx** grid = J->inTime ? gT : gF;
  const long W = (long)J->nOut * N;
  for (long i = J->start; i < J->stop; i++) {
    if (J->batch) {
      signal_sim(m, J->bin + i * J->bcols, J->bcols, J->inTime, J->inputMode, gF, gT, 1, 1, &w);
      for (int om = 0; om < J->nOut; om++)
        memcpy(J->bout + i * W + (long)om * N, grid[J->outputs[om]], sizeof(cplx) * (size_t)N);
      continue;
    }
    const int im = (int)(i / N);
    const int md = J->inputs[im];
    cplx* g = grid[md];
    /* src/_NonlinearMedium.cpp:329-345 */
    memset(g, 0, sizeof(cplx) * (size_t)N);
    g[i % N] = 1;
    signal_sim(m, g, N, J->inTime, md, gF, gT, 1, 1, &w);
    for (int om = 0; om < J->nOut; om++) {
      const cplx* o = grid[J->outputs[om]];
      cplx* c = J->C + i * W + (long)om * N;
      cplx* s = J->S + i * W + (long)om * N;
      for (int k = 0; k < N; k++) { c[k] += 0.5 * o[k]; s[k] += 0.5 * o[k]; }
    }
    memset(g, 0, sizeof(cplx) * (size_t)N);
    g[i % N] = CMPLX(0, 1);
    signal_sim(m, g, N, J->inTime, md, gF, gT, 1, 1, &w);
    for (int om = 0; om < J->nOut; om++) {
      const cplx* o = grid[J->outputs[om]];
      cplx* c = J->C + i * W + (long)om * N;
      cplx* s = J->S + i * W + (long)om * N;
      const cplx hi = CMPLX(0, 0.5);
      for (int k = 0; k < N; k++) { c[k] -= hi * o[k]; s[k] += hi * o[k]; }
    }
  }
  for (int q = 0; q < nS; q++) { free(gF[q]); free(gT[q]); }
  work_free(&w);
  return NULL;
}

static void run_jobs(job_t* proto, long total, int nThreads) {
  pthread_t* th = (pthread_t*)calloc((size_t)nThreads, sizeof(pthread_t));
  job_t* jobs = (job_t*)calloc((size_t)nThreads, sizeof(job_t));
  for (int t = 1; t < nThreads; t++) {
    jobs[t] = *proto;
    jobs[t].start = (t * total) / nThreads;
    jobs[t].stop = ((t + 1) * total) / nThreads;
    pthread_create(&th[t], NULL, job_run, &jobs[t]);
  }
  jobs[0] = *proto;
  jobs[0].start = 0;
  jobs[0].stop = total / nThreads;
  job_run(&jobs[0]);
  for (int t = 1; t < nThreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
}

/* Mode selection of computeGreensFunction, src/_NonlinearMedium.cpp:266-299.
 * The reference counts list entries below nSignalModes as the number of
 * modes but selects modes by treating entries as flags; the two agree only
 * for flag lists whose set entries are all < nSignalModes, which is the
 * case accepted here (others are undefined behaviour in the reference). */
static int select_modes(const orc_medium* m, const uint8_t* use, int n, int* list, int* count,
                        const char* tooLong, const char* none) {
  const int nS = m->nS;
  if (n > nS) return fail(ORC_INVALID_ARGUMENT, tooLong);
  int cnt = 0;
  if (n > 0) {
    for (int i = 0; i < n; i++) cnt += (use[i] < nS);
    if (cnt == 0) return fail(ORC_INVALID_ARGUMENT, none);
  } else {
    cnt = nS;
  }
  int k = 0;
  for (int q = 0; q < nS; q++)
    if (n == 0 || (q < n && use[q])) list[k++] = q;
  if (k != cnt) return fail(ORC_INVALID_ARGUMENT, "useInput/useOutput must be per-mode flags");
  *count = cnt;
  return ORC_OK;
}

int orc_greens_shape(const orc_medium* m, const uint8_t* use_in, int n_use_in, const uint8_t* use_out,
                     int n_use_out, int32_t shape[2]) {
  if (m->fnlm) return fail(ORC_RUNTIME_ERROR, "Object does not have this method.");
  int ins[MAXM], outs[MAXM], nIn, nOut, e;
  if (n_use_in > m->nS) return fail(ORC_INVALID_ARGUMENT, "List of requested inputs indices longer than number of modes!");
  if (n_use_out > m->nS) return fail(ORC_INVALID_ARGUMENT, "List of requested output indices longer than number of modes!");
  if ((e = select_modes(m, use_in, n_use_in, ins, &nIn, "List of requested inputs indices longer than number of modes!",
                        "Requested no valid inputs!")))
    return e;
  if ((e = select_modes(m, use_out, n_use_out, outs, &nOut, "List of requested output indices longer than number of modes!",
                        "Requested no valid outputs!")))
    return e;
  shape[0] = nOut * m->N;
  shape[1] = nIn * m->N;
  return ORC_OK;
}

/* fftshift2 of a block, src/_NonlinearMedium.cpp:456-467 */
static inline long shift_idx(long r, long n) { return r < n / 2 ? r + (n - n / 2) : r - n / 2; }

/* _NonlinearMedium::computeGreensFunction, src/_NonlinearMedium.cpp:263-375 */
int orc_greens(orc_medium* m, int in_time_domain, int run_pump, int n_threads, int normalize,
               const uint8_t* use_in, int n_use_in, const uint8_t* use_out, int n_use_out, double* Cout,
               double* Sout) {
  if (m->fnlm) return fail(ORC_RUNTIME_ERROR, "Object does not have this method.");
  const int N = m->N;
  int ins[MAXM], outs[MAXM], nIn, nOut, e;
  if (n_use_in > m->nS) return fail(ORC_INVALID_ARGUMENT, "List of requested inputs indices longer than number of modes!");
  if (n_use_out > m->nS) return fail(ORC_INVALID_ARGUMENT, "List of requested output indices longer than number of modes!");
  if ((e = select_modes(m, use_in, n_use_in, ins, &nIn, "List of requested inputs indices longer than number of modes!",
                        "Requested no valid inputs!")))
    return e;
  if ((e = select_modes(m, use_out, n_use_out, outs, &nOut, "List of requested output indices longer than number of modes!",
                        "Requested no valid outputs!")))
    return e;
  if (n_threads > N * nIn) return fail(ORC_INVALID_ARGUMENT, "Too many threads requested!");
  if (n_threads < 1) return fail(ORC_INVALID_ARGUMENT, "vector::reserve");
  {
    /* the reference indexes _nlStep by block index (undefined behaviour past its end) */
    const int nbm = nOut < nIn ? nOut : nIn;
    if (normalize && m->nNl < (nbm > nOut ? nbm : nOut))
      return fail(ORC_INVALID_ARGUMENT, "normalize requires one nonlinear length per mode for this medium");
  }
  if (run_pump) orc_run_pump(m);

  const long R = (long)nIn * N, Cc = (long)nOut * N;
  cplx* C = calloc_c(R * Cc);
  cplx* S = calloc_c(R * Cc);
  job_t J;
  memset(&J, 0, sizeof J);
  J.m = m; J.inTime = in_time_domain; J.nIn = nIn; J.nOut = nOut;
  J.inputs = ins; J.outputs = outs; J.C = C; J.S = S;
  run_jobs(&J, R, n_threads);

  /* transposeInPlace, then fftshift2 of each block (and normalisation):
   * out is (nOut*N) x (nIn*N); the reference loops both block indices over
   * nOutputModes, so blocks with im >= nOut are left unshifted. */
  cplx* Co = (cplx*)Cout;
  cplx* So = (cplx*)Sout;
  for (long r = 0; r < Cc; r++)
    for (long c = 0; c < R; c++) {
      Co[r * R + c] = C[c * Cc + r];
      So[r * R + c] = S[c * Cc + r];
    }
  const int nb = nOut < nIn ? nOut : nIn;
  cplx* blk = calloc_c((long)N * N);
  for (int pass = 0; pass < 2; pass++) {
    cplx* G = pass == 0 ? Co : So;
    for (int im = 0; im < nb; im++)
      for (int om = 0; om < nOut; om++) {
        cplx scale = 1;
        const int doScale = normalize && im != om;
        if (doScale) scale = csqrt(m->nlStep[im] / m->nlStep[om]);
        for (long r = 0; r < N; r++)
          for (long c = 0; c < N; c++) blk[r * N + c] = G[(om * (long)N + r) * R + im * (long)N + c];
        for (long r = 0; r < N; r++)
          for (long c = 0; c < N; c++) {
            cplx v = blk[shift_idx(r, N) * N + shift_idx(c, N)];
            G[(om * (long)N + r) * R + im * (long)N + c] = doScale ? v * scale : v;
          }
      }
  }
  free(blk);
  free(C);
  free(S);
  return ORC_OK;
}

int orc_batch_shape(const orc_medium* m, int rows, int cols, const uint8_t* use_out, int n_use_out,
                    int32_t shape[2]) {
  int nOut = 0;
  if (n_use_out > 0) {
    for (int i = 0; i < n_use_out; i++) nOut += (use_out[i] != 0);
    if (nOut == 0) return fail(ORC_INVALID_ARGUMENT, "Requested no outputs!");
  } else {
    nOut = m->nS;
  }
  (void)cols;
  shape[0] = rows;
  shape[1] = nOut * m->N;
  return ORC_OK;
}

/* _NonlinearMedium::batchSignalSimulation, src/_NonlinearMedium.cpp:378-444 */
int orc_batch(orc_medium* m, const double* in, int rows, int cols, int in_time_domain, int run_pump,
              int n_threads, int input_mode, const uint8_t* use_out, int n_use_out, double* out) {
  const int N = m->N;
  if (cols % N != 0 || cols / N == 0 || cols / N > m->nS) return fail(ORC_INVALID_ARGUMENT, "Signals not of correct length!");
  if (n_threads > rows) return fail(ORC_INVALID_ARGUMENT, "Too many threads requested!");
  if (input_mode < 0 || input_mode >= m->nS)
    return fail(ORC_INVALID_ARGUMENT, "inputModes does not match any mode in the system");
  int outs[MAXM], nOut = 0;
  if (n_use_out > 0) {
    for (int i = 0; i < n_use_out; i++) nOut += (use_out[i] != 0);
    if (nOut == 0) return fail(ORC_INVALID_ARGUMENT, "Requested no outputs!");
  } else {
    nOut = m->nS;
  }
  int k = 0;
  for (int q = 0; q < m->nS; q++)
    if (n_use_out == 0 || (q < n_use_out && use_out[q])) outs[k++] = q;
  if (k != nOut) return fail(ORC_INVALID_ARGUMENT, "useOutput must be per-mode flags");
  if (n_threads < 1) return fail(ORC_INVALID_ARGUMENT, "vector::reserve");
  if (run_pump) {
    int e = orc_run_pump(m);
    if (e) return e;
  }
  job_t J;
  memset(&J, 0, sizeof J);
  J.m = m; J.inTime = in_time_domain; J.nOut = nOut; J.outputs = outs;
  J.batch = 1; J.bin = (const cplx*)in; J.bcols = cols; J.inputMode = input_mode; J.bout = (cplx*)out;
  run_jobs(&J, rows, n_threads);
  return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* construction                                                              */
/* ------------------------------------------------------------------------- */

/* _NonlinearMedium::_NonlinearMedium, src/_NonlinearMedium.cpp:8-40, with
 * _FullyNonlinearMedium (src/_FullyNonlinearMedium.hpp:32-38) and the
 * Chi3GNLSE Raman table (solvers/Chi3GNLSE.cpp:20-36). */
int orc_create(const orc_params* p, orc_medium** out) {
  *out = NULL;
  if (p->n_signal_modes > MAXM || p->n_pump_modes > MAXM || p->n_nl > MAXM || p->n_diff > MAXM)
    return fail(ORC_INVALID_ARGUMENT, "too many modes");
  orc_medium* m = (orc_medium*)calloc(1, sizeof(orc_medium));
  m->model = p->model;
  m->nS = p->n_signal_modes;
  m->nP = p->n_pump_modes;
  m->canBePoled = p->can_be_poled;
  m->fnlm = p->fully_nonlinear;
  int e;
  double zR = p->rayleigh_length;
  if (p->intensity_profile == 1) zR = INFINITY;
  if ((e = set_lengths(m, p, zR))) { orc_destroy(m); return e; }
  if ((e = reset_grids(m, p->t_precision, p->t_max))) { orc_destroy(m); return e; }
  fft_plan_init(&m->fft, m->N);
  set_dispersion(m, p);
  if (m->canBePoled) {
    if ((e = set_poling(m, p->poling, p->n_poling))) { orc_destroy(m); return e; }
    if (p->n_poling_profile > 0) {
      if (p->n_poling_profile != m->nZ) {
        orc_destroy(m);
        return fail(ORC_INVALID_ARGUMENT, "polingProfile length does not match number of z steps");
      }
      memcpy(m->poling, p->poling_profile, sizeof(double) * (size_t)m->nZ);
    }
  }
  m->profile = (m->zR != INFINITY) ? p->intensity_profile : 1;
  for (int q = 0; q < m->nP; q++) m->env[q] = calloc_c(m->N);
  if (m->nP > 0) {
    if (p->n_custom_pump != 0) e = set_pump_custom(m, p->custom_pump, p->n_custom_pump, p->chirp, p->delay, 0);
    else e = set_pump_type(m, p->pulse_type, p->chirp, p->delay, 0);
    if (e) { orc_destroy(m); return e; }
  }
  if (m->fnlm) { m->dzp = 0; m->nZP = 0; }

  if (m->model == ORC_CHI3GNLSE) {
    const int N = m->N;
    const double fr = p->gnlse[0], fb = p->gnlse[1], t1 = p->gnlse[2], t2 = p->gnlse[3], t3 = p->gnlse[4];
    const double coeff1 = fr * (1. - fb) * (t1 / (t2 * t2) + 1. / t1);
    const double coeff2 = fr * fb / (t3 * t3);
    cplx* rt = calloc_c(N);
    for (int k = 0; k < N / 2; k++) {
      const double t = m->tau[k];
      rt[k] = coeff1 * exp(-t / t2) * sin(t / t1) + coeff2 * exp(-t / t3) * (2 * t3 - t);
    }
    rt[0] += 1. - fr;
    m->raman = calloc_c(N);
    cplx* scr = calloc_c(N);
    fft_fwd(&m->fft, m->raman, rt, scr);
    free(scr);
    free(rt);
    m->gnlseFilter = calloc_c(N);
    for (int k = 0; k < N; k++) m->gnlseFilter[k] = m->nlStep[0] + m->nlStep[1] * m->omega[k];
  }
  *out = m;
  return ORC_OK;
}

void orc_destroy(orc_medium* m) {
  if (!m) return;
  free(m->tau); free(m->omega); free(m->poling); free(m->raman); free(m->gnlseFilter);
  for (int q = 0; q < MAXM; q++) {
    free(m->dispP[q]); free(m->dispS[q]); free(m->dstepP[q]); free(m->dstepS[q]); free(m->env[q]);
    free(m->pumpFreq[q]); free(m->pumpTime[q]); free(m->sigFreq[q]); free(m->sigTime[q]);
  }
  if (m->fft.tw) fft_plan_free(&m->fft);
  free(m);
}

void orc_dims(const orc_medium* m, int32_t d[8]) {
  d[0] = m->N; d[1] = m->nZ; d[2] = m->nZP; d[3] = m->nS; d[4] = m->nP; d[5] = m->fnlm;
  d[6] = m->nPol; d[7] = m->model;
}
void orc_scalars(const orc_medium* m, double s[5]) {
  s[0] = m->z; s[1] = m->dz; s[2] = m->dzp; s[3] = m->tMax; s[4] = m->zR;
}
const double* orc_tau(const orc_medium* m) { return m->tau; }
const double* orc_omega(const orc_medium* m) { return m->omega; }
const double* orc_poling(const orc_medium* m) { return m->poling; }
const double* orc_pump_freq(const orc_medium* m, int i) { return (const double*)m->pumpFreq[i]; }
const double* orc_pump_time(const orc_medium* m, int i) { return (const double*)m->pumpTime[i]; }
const double* orc_signal_freq(const orc_medium* m, int i) { return (const double*)m->sigFreq[i]; }
const double* orc_signal_time(const orc_medium* m, int i) { return (const double*)m->sigTime[i]; }
const double* orc_envelope(const orc_medium* m, int i) { return (const double*)m->env[i]; }
const double* orc_raman(const orc_medium* m) { return (const double*)m->raman; }
