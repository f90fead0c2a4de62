/*
 * nlm_oracle.h -- CPU restatement of the split-step propagation core of
 * `nonlinearmedium` (mcmahon-lab/MultimodeNonlinearOptics).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference and the CPU
 * baseline timer.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load it; the product (multimodenonlinearoptics_amd) never
 * links, loads or calls it.
 *
 * Every function follows the reference source line by line; the citation for
 * each is given where it is defined in nlm_oracle.c.
 *
 * Complex arrays cross this interface as interleaved (re, im) doubles,
 * row-major, exactly the memory layout of the reference's Array2Dcd.
 */
#ifndef NLM_ORACLE_H
#define NLM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Solver identifiers (solvers/RegisteredSolvers.hpp:1-17). */
enum {
  ORC_CHI3 = 0, ORC_CHI2PDC, ORC_CHI2SFG, ORC_CHI2AFC, ORC_CHI2SFGPDC, ORC_CHI2SFGII,
  ORC_CHI2PDCII, ORC_CHI2SFGOPA, ORC_CHI2SFGXPM,
  ORC_CHI2SHG, ORC_CHI2SHGOPA, ORC_CHI2DSFG, ORC_CHI2SHGXPM, ORC_CHI2ASHG, ORC_CHI3GNLSE,
  ORC_NUM_MODELS
};

/* Error kinds: the C++ exception type the reference throws. */
enum { ORC_OK = 0, ORC_INVALID_ARGUMENT = 1, ORC_RUNTIME_ERROR = 2 };

/* Arguments of the base-class constructor _NonlinearMedium::_NonlinearMedium
 * (src/_NonlinearMedium.cpp:8-40) after a solver constructor has mapped its
 * own keyword arguments onto them. */
typedef struct {
  int model;
  int n_signal_modes, n_pump_modes, can_be_poled, fully_nonlinear;
  double relative_length;
  int n_nl;     const double* nl_length;
  int n_beta2;  const double* beta2;
  int n_beta2s; const double* beta2s;
  int n_beta1;  const double* beta1;
  int n_beta1s; const double* beta1s;
  int n_beta3;  const double* beta3;
  int n_beta3s; const double* beta3s;
  int n_diff;   const double* diff_beta0;
  double rayleigh_length, t_max;
  int t_precision, z_precision;
  int nz_override;                 /* >0: fix nZSteps directly (extension)        */
  int intensity_profile;           /* 0 GaussianBeam, 1 Constant, 2 Apodization   */
  double chirp, delay;
  int n_custom_pump; const double* custom_pump;   /* interleaved complex        */
  int pulse_type;                  /* 0 Gaussian, 1 Sech, 2 Sinc                  */
  int n_poling; const double* poling;              /* domain lengths            */
  int n_poling_profile; const double* poling_profile; /* per-step (extension)   */
  double gnlse[5];                 /* fr, fb, tau1, tau2, tau3 (Chi3GNLSE)        */
} orc_params;

typedef struct orc_medium orc_medium;

/* All functions return ORC_OK or an error kind; orc_last_error() holds the
 * reference's exception message for the calling thread. */
const char* orc_last_error(void);

int  orc_create(const orc_params* p, orc_medium** out);
void orc_destroy(orc_medium* m);

/* dims: N, nZ, nZP, nSignalModes, nPumpModes, fullyNonlinear, nPoling, model */
void orc_dims(const orc_medium* m, int32_t dims[8]);
/* scalars: z, dz, dzp, tMax, rayleighLength */
void orc_scalars(const orc_medium* m, double s[5]);

const double* orc_tau(const orc_medium* m);
const double* orc_omega(const orc_medium* m);
const double* orc_poling(const orc_medium* m);
const double* orc_pump_freq(const orc_medium* m, int i);
const double* orc_pump_time(const orc_medium* m, int i);
const double* orc_signal_freq(const orc_medium* m, int i);
const double* orc_signal_time(const orc_medium* m, int i);
const double* orc_envelope(const orc_medium* m, int i);
const double* orc_raman(const orc_medium* m);

int orc_set_pump_type(orc_medium* m, int pulse_type, double chirp, double delay, int pump_index);
int orc_set_pump_custom(orc_medium* m, const double* pump, int n, double chirp, double delay, int pump_index);
int orc_set_pump_other(orc_medium* m, const orc_medium* other, int mode_index, double delay, int pump_index);

int orc_run_pump(orc_medium* m);
int orc_run_signal(orc_medium* m, const double* in, int in_len, int in_time_domain, int input_mode);

/* C, S: caller-allocated, (nOut*N) x (nIn*N) complex; rows/cols reported first
 * through orc_greens_shape. */
int orc_greens_shape(const orc_medium* m, const uint8_t* use_in, int n_use_in,
                     const uint8_t* use_out, int n_use_out, int32_t shape[2]);
int orc_greens(orc_medium* m, int in_time_domain, int run_pump, int n_threads, int normalize,
               const uint8_t* use_in, int n_use_in, const uint8_t* use_out, int n_use_out,
               double* C, double* S);

int orc_batch_shape(const orc_medium* m, int rows, int cols, const uint8_t* use_out, int n_use_out,
                    int32_t shape[2]);
int orc_batch(orc_medium* m, const double* in, int rows, int cols, int in_time_domain, int run_pump,
              int n_threads, int input_mode, const uint8_t* use_out, int n_use_out, double* out);

/* Standalone transforms with the reference's conventions
 * (src/CustomEigenFFT.h:54-137): forward unscaled, inverse scaled by 1/n. */
int orc_fft(int n, int count, const double* in, double* out, int inverse);

#ifdef __cplusplus
}
#endif
#endif
