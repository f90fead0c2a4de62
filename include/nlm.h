/*
 * nlm.h -- C-ABI of the MI355X split-step propagation engine behind the
 * `nonlinearmedium` drop-in (libnlm_hip.so).
 *
 * The reference has no C interface: its hot path is the C++ template
 * _NonlinearMedium::signalSimulationTemplate (src/_NonlinearMedium.hpp:160-239)
 * driven by runPumpSimulation / runSignalSimulation / computeGreensFunction /
 * batchSignalSimulation (src/_NonlinearMedium.cpp:225-444, solvers/Chi3.cpp:25-54)
 * and reached from Python through pybind11 (src/nlmModulePy.cpp:61-98).  Each
 * entry point below replaces one of those functions; the citation is given
 * beside it.  The host classes (the pybind11 module `nonlinearmedium`) compute
 * the per-medium tables exactly as the reference's constructors do and hand
 * them to nlm_plan_create.
 *
 * Conventions
 *  - complex values are interleaved (re, im) doubles, arrays row-major, the
 *    memory layout of the reference's Array2Dcd (src/_NonlinearMedium.hpp:13-15);
 *  - every pointer argument is a HOST pointer owned by the caller, except where
 *    a name ends in _dev; output pointers may be NULL, in which case the result
 *    stays resident on the device (readable later with the *_download calls);
 *  - every function returns NLM_OK or an error kind; nlm_last_error() returns
 *    the message for the calling thread.  NLM_INVALID_ARGUMENT corresponds to
 *    the reference's std::invalid_argument, NLM_RUNTIME_ERROR to
 *    std::runtime_error;
 *  - transforms follow src/CustomEigenFFT.h:54-137: forward unscaled,
 *    inverse scaled by 1/n.  Any even n is supported: n = 2^k, 8 <= n <= 8192
 *    runs the in-register kernels, every other even n (and Chi3GNLSE batches
 *    of 16384 and 32768) the mixed-radix and staged paths; the adjoints and
 *    heterogeneous batches need powers of two.
 *  - a plan is not safe for concurrent calls; distinct plans may be used from
 *    distinct threads.
 */
#ifndef NLM_H
#define NLM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NLM_OK = 0,
  NLM_INVALID_ARGUMENT = 1,
  NLM_RUNTIME_ERROR = 2,
  NLM_DEVICE_ERROR = 3,
};

/* Solver identifiers, in the order of solvers/RegisteredSolvers.hpp:1-17. */
enum {
  NLM_CHI3 = 0, NLM_CHI2PDC, NLM_CHI2SFG, NLM_CHI2AFC, NLM_CHI2SFGPDC, NLM_CHI2SFGII,
  NLM_CHI2PDCII, NLM_CHI2SFGOPA, NLM_CHI2SFGXPM,
  NLM_CHI2SHG, NLM_CHI2SHGOPA, NLM_CHI2DSFG, NLM_CHI2SHGXPM, NLM_CHI2ASHG, NLM_CHI3GNLSE,
  NLM_NUM_MODELS
};

/* Number of complex per-position coefficients each model's nonlinear step
 * reads from step_coef (index [i][pos][c], pos = 0: z=(i-1)dz, 1: (i-1/2)dz,
 * 2: i*dz), and of constants in const_coef.  See nlm_plan_desc. */
int nlm_model_ncoef(int model, int32_t* n_step_coef, int32_t* n_const_coef, int32_t* n_signal, int32_t* n_pump);

typedef struct {
  int32_t model;
  int32_t n_freqs;          /* N                                                   */
  int32_t n_zsteps;         /* nZ (signal rows); pump grids have 2nZ-1 rows        */
  /* [nS][N] complex: exp(i dz D_s), exp(i dz/2 D_s), exp(-i dz/2 D_s)
   * (src/_NonlinearMedium.cpp:174-176, src/_NonlinearMedium.hpp:174,187,236) */
  const double* disp_step;
  const double* half_in;
  const double* half_out;
  /* [nZ][3][n_step_coef] complex: the per-step scalar factors of the model's
   * DiffEq (poling, nlStep, phase mismatch, beam profile), evaluated on the
   * host in the reference's operation order.  Row 0 is unused. */
  const double* step_coef;
  const double* const_coef; /* [n_const_coef] complex                              */
  /* Chi3GNLSE only (solvers/Chi3GNLSE.cpp:36,72): [N] complex Raman spectrum and
   * [N] complex derivative filter nlStep[0] + nlStep[1]*omega; else NULL.      */
  const double* raman;
  const double* gnlse_filter;
} nlm_plan_desc;

typedef struct nlm_plan_s nlm_plan;

/* Uploads the tables of one medium.  Replaces the device-independent setup the
 * reference does in _NonlinearMedium::_NonlinearMedium
 * (src/_NonlinearMedium.cpp:8-40); the tables themselves are built by the host. */
int nlm_plan_create(const nlm_plan_desc* desc, nlm_plan** out);
int nlm_plan_destroy(nlm_plan* plan);

/* Run this plan's work on `hip_stream` (a hipStream_t); NULL restores the
 * plan's own stream.  Lets a caller order the engine against its own work. */
int nlm_plan_set_stream(nlm_plan* plan, void* hip_stream);
void* nlm_plan_stream(nlm_plan* plan);

/* Replace the step coefficient table (same layout as at creation), e.g. after
 * a poling change; the FFT and dispersion tables are kept. */
int nlm_plan_set_coef(nlm_plan* plan, const double* step_coef);

/* Linear pump propagation, _NonlinearMedium::runPumpSimulation
 * (src/_NonlinearMedium.cpp:225-250).  envelope: [nP][N] time domain;
 * disp_pump: [nP][N] real D_p(omega); strength: [2nZ-1] beam-profile factor or
 * NULL for a constant profile. */
int nlm_pump_linear(nlm_plan* plan, int n_pump, const double* envelope, const double* disp_pump,
                    double dzp, const double* strength);

/* Nonlinear (self-phase) pump propagation, Chi3::runPumpSimulation
 * (solvers/Chi3.cpp:25-54).  disp_step_pump: [N] complex exp(i dzp D_p);
 * rel_intensity: [2nZ-1] complex (dzp/dz * nlStep * profile). */
int nlm_pump_chi3(nlm_plan* plan, const double* envelope, const double* disp_pump, double dzp,
                  const double* disp_step_pump, const double* rel_intensity);

/* Pump grid transfer ([2nZ-1][N] complex, time and frequency domain); used by
 * _NonlinearMedium::setPump(other, ...) (src/_NonlinearMedium.cpp:506-534),
 * which builds the grid on the host, and to read the grids back. */
int nlm_pump_upload(nlm_plan* plan, int pump_index, const double* pump_time, const double* pump_freq);
int nlm_pump_download(nlm_plan* plan, int pump_index, double* pump_time, double* pump_freq);

/* _NonlinearMedium::setPump(other, modeIndex, delay) (src/_NonlinearMedium.cpp:506-534)
 * on the device: pump grid `pump_index` of plan built from the signal grid of
 * mode `mode` of src (its last nlm_propagate_grid), without host copies.
 * disp_pump: [N] real D_p of this medium's pump; disp_signal_src: [N] real D_s of
 * the source mode; omega: [N]; beta1_delay = beta1 of the pump * delay; z, z_src
 * the two media's lengths.  As in the reference, the last pump time row keeps its
 * previous contents. */
int nlm_pump_import(nlm_plan* plan, int pump_index, nlm_plan* src, int mode, const double* disp_pump,
                    const double* disp_signal_src, const double* omega, double dzp, double dz_src,
                    double beta1_delay, double z, double z_src);

/* Full z-resolved propagation, runSignalSimulation -> signalSimulationTemplate
 * with optimized=false (src/_NonlinearMedium.cpp:253-260,
 * src/_NonlinearMedium.hpp:219-238).  input: in_len complex (N or k*N). */
int nlm_propagate_grid(nlm_plan* plan, const double* input, int in_len, int in_time_domain, int input_mode);
/* z-resolved batches (extension): the full-grid propagation of every one of rows
 * input rows (rows x cols, as nlm_batch) in one launch; grid_freq, grid_time:
 * [rows][nS][nZ][N] complex (host), or NULL to keep them on the device.  Row 0's
 * grids are then also what nlm_grid_download returns. */
int nlm_batch_grids(nlm_plan* plan, const double* inputs, int rows, int cols, int in_time_domain, int input_mode,
                    double* grid_freq, double* grid_time);
/* [nZ][N] complex signal grids of one mode from the last nlm_propagate_grid. */
int nlm_grid_download(nlm_plan* plan, int mode, double* signal_freq, double* signal_time);

/* Independent propagations, batchSignalSimulation
 * (src/_NonlinearMedium.cpp:378-444): rows x cols inputs, output rows x
 * (n_out*N) final rows of the selected modes. */
int nlm_batch(nlm_plan* plan, const double* inputs, int rows, int cols, int in_time_domain, int input_mode,
              int n_out, const int32_t* outputs, double* out);

/* Green's functions, computeGreensFunction (src/_NonlinearMedium.cpp:263-375):
 * 2*N*n_in impulse propagations in one launch, then the transpose / fftshift2 /
 * normalisation assembly.  block_scale: [min(n_in,n_out)][n_out] complex
 * factor per shifted block (all ones without normalisation) or NULL.
 * C, S: (n_out*N) x (n_in*N) complex, or NULL to keep them on the device. */
int nlm_greens(nlm_plan* plan, int in_time_domain, int n_in, const int32_t* inputs, int n_out,
               const int32_t* outputs, const double* block_scale, double* C, double* S);
int nlm_greens_download(nlm_plan* plan, double* C, double* S);

/* No counterpart in the reference: nlm_greens for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  It calls no hipMalloc, hipFree, memset or other allocating
 * routine: the Green's rows ([2 N n_in][N n_out] complex) and the step-propagator
 * table of one-mode undepleted media ([nZSteps][N][2] complex) live in a fixed
 * device array of NLM_NOALLOC_SCRATCH_BYTES that each GPU holds from its first
 * call until the process ends; larger requests fail with NLM_INVALID_ARGUMENT.
 * C_dev, S_dev: caller-owned device buffers of (n_out*N) x (n_in*N) complex.
 * No block normalisation.  Asynchronous on the plan's stream (nlm_plan_sync).
 * The pump grid must have been computed.  All calls on one GPU share its array:
 * calls that overlap on one GPU race. */
#define NLM_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int nlm_greens_noalloc(nlm_plan* plan, int in_time_domain, int n_in, const int32_t* inputs, int n_out,
                       const int32_t* outputs, double* C_dev, double* S_dev);

/* Device-resident variants for callers that keep data in HBM (e.g. PyTorch
 * tensors): same semantics as nlm_batch with device pointers. */
int nlm_batch_dev(nlm_plan* plan, const double* inputs_dev, int rows, int cols, int in_time_domain,
                  int input_mode, int n_out, const int32_t* outputs, double* out_dev);

/* Per-launch device timing of the last call (ms), recorded with HIP events on
 * the plan's stream when enabled: [0] pump, [1] propagation, [2] assembly. */
/* Heterogeneous batches: n independent propagations of different media (model,
 * nFreqs, nZSteps, dispersion, pump), each the batchSignalSimulation of one
 * input row (src/_NonlinearMedium.cpp:378-444, src/_FullyNonlinearMedium.hpp:8-11),
 * run as one captured graph of per-(model, nFreqs) persistent kernels fed by
 * device work queues.  Every solver; nFreqs a power of two, 32..4096.  With
 * `envelope` set, the pump rows of a single linear pump are generated in the
 * kernel from the envelope (src/_NonlinearMedium.cpp:225-250); with it NULL a
 * pumped medium's device pump grids are read (they must have been computed:
 * Chi3's nonlinear pump, Chi2SFGOPA's two pumps, or any grid set before).  The
 * plans must outlive the batch.  (Extension: no counterpart in the reference,
 * which runs instances one by one.) */
typedef struct {
  nlm_plan* plan;
  const double* input;       /* cols complex values; copied to the device at creation */
  int32_t cols;
  int32_t in_time_domain;
  int32_t input_mode;
  int32_t n_out;
  int32_t outputs[4];
  const double* envelope;    /* linear pump: [N] complex time-domain envelope, else NULL */
  const double* disp_pump;   /* [N] real D_p(omega) */
  double dzp;
  const double* strength;    /* [2nZ-1] beam-profile factor or NULL */
} nlm_multi_item;
typedef struct nlm_multi_s nlm_multi;
/* autotune (flags): bit 0 time each class alone for a few persistent grid sizes
 * and keep the fastest (once per process and class shape); bit 1 stage the
 * pump rows of nFreqs 4096 classes with a generated pump: each launch first
 * writes their time-domain rows to a device buffer (a quarter of the device
 * memory at most), which the class's kernel then reads like a pump grid. */
int nlm_multi_create(int n, const nlm_multi_item* items, int autotune, nlm_multi** out);
int nlm_multi_launch(nlm_multi* batch);             /* one graph launch, asynchronous */
int nlm_multi_sync(nlm_multi* batch, double* ms);   /* wait; ms = device time of the last launch */
int nlm_multi_output(nlm_multi* batch, int i, double* out);  /* n_out*N complex of instance i */
int nlm_multi_classes(nlm_multi* batch);
int nlm_multi_grid(nlm_multi* batch, int c, int32_t* model, int32_t* log2n, int32_t* count, int32_t* grid);
int nlm_multi_destroy(nlm_multi* batch);

/* Single-signal-mode undepleted media (Chi3, Chi2PDC): propagate with the
 * per-step RK4 propagator table (enable = 1, the default) or evaluate the
 * DiffEq in every field (enable = 0).  Both compute the RK4 step of
 * src/_NonlinearMedium.hpp:207-210; they differ by rounding only.  (Extension:
 * no counterpart in the reference.) */
int nlm_plan_set_step_table(nlm_plan* plan, int enable);
int nlm_plan_enable_timing(nlm_plan* plan, int enable);
int nlm_plan_last_timing(nlm_plan* plan, double ms[3]);

/* Differentiable Green's function of a single-mode Chi2PDC medium (linear in
 * the pump, conjugate-coupled signal; computeGreensFunction,
 * src/_NonlinearMedium.cpp:263-375, with Chi2PDC::DiffEq,
 * solvers/Chi2PDC.cpp:27-47).  nlm_greens_forward is nlm_greens with
 * time-domain field checkpoints every `ckpt_every` z steps kept on the device
 * (<= 0: the smallest interval whose 2N x ceil((nZ-1)/K) x N x 16 bytes fit in a
 * quarter of the device memory and half of the free memory, at most
 * ceil(sqrt(nZ-1))).  nlm_greens_adjoint then takes dL/dC, dL/dS (complex,
 * convention dL/dRe + i dL/dIm, the layout of C and S) and returns
 *   g_coef          [nZ][3][n_step_coef] gradient w.r.t. the step coefficients,
 *   g_pump_time     [2nZ-1][N] gradient w.r.t. the pump time grid (or NULL),
 *   g_envelope      [N] gradient w.r.t. the time-domain pump envelope through
 *                   the linear pump grid (nlm_pump_linear must have built it; or NULL).
 * The reverse sweep rebuilds the states between checkpoints; reductions are
 * done in a fixed order (no atomics). */
int nlm_greens_forward(nlm_plan* plan, int in_time_domain, int n_in, const int32_t* inputs, int n_out,
                       const int32_t* outputs, const double* block_scale, int ckpt_every, double* C, double* S);
int nlm_greens_adjoint(nlm_plan* plan, const double* gC, const double* gS, double* g_coef, double* g_pump_time,
                       double* g_envelope);

/* Differentiable Chi3GNLSE batch (batchSignalSimulation of a fully nonlinear
 * medium, src/_FullyNonlinearMedium.hpp:8-11 with Chi3GNLSE::DiffEq,
 * solvers/Chi3GNLSE.cpp:39-92, inside signalSimulationTemplate,
 * src/_NonlinearMedium.hpp:160-239).  nlm_batch_forward is nlm_batch for rows
 * input rows of length N (one mode, output in the input's domain) that also
 * keeps the time-domain field every `ckpt_every` z steps (<= 0:
 * ceil(sqrt((nZ-1)/8))) on the device.  nlm_batch_adjoint then takes dL/d(out)
 * (rows x N complex, convention dL/dRe + i dL/dIm, the layout and domain of
 * out) and returns dL/d(inputs) (rows x N, the domain of the inputs).  The
 * reverse sweep recomputes each checkpoint segment once, keeping the stage
 * inputs and Raman convolutions of its steps: device memory
 * (nSeg + 8 ckpt_every + 10) x rows x N x 16 bytes.  nFreqs 32..32768.
 * (Extension: the reference has no derivatives.) */
int nlm_batch_forward(nlm_plan* plan, const double* inputs, int rows, int in_time_domain, int ckpt_every,
                      double* out);
int nlm_batch_adjoint(nlm_plan* plan, const double* g_out, double* g_in);

/* Green's function composition of two media, Cascade::computeGreensFunction
 * (src/Cascade.cpp:128-129): C = C2 C1 + S2 conj(S1), S = C2 S1 + S2 conj(C1).
 * C2, S2: M x K; C1, S1: K x Ncol; C, S: M x Ncol (host, row-major). */
int nlm_greens_compose(int M, int K, int Ncol, const double* C2, const double* S2, const double* C1,
                       const double* S1, double* C, double* S);
/* The same on device buffers (complex, interleaved) on `stream` (a hipStream_t;
 * NULL = the null stream), asynchronous.  Runs on the FP64 matrix cores. */
int nlm_greens_compose_dev(int M, int K, int Ncol, const double* C2, const double* S2, const double* C1,
                           const double* S1, double* C, double* S, void* stream);

/* Cascade::computeGreensFunction (src/Cascade.cpp:110-135) over n media whose
 * Green's functions were computed with nlm_greens(..., C = S = NULL) and are
 * still on the device: composed in order (plans[0] first) on the device, the
 * result (rows of plans[n-1]) x (cols of plans[0]) written to C, S (host). */
int nlm_greens_cascade(int n, nlm_plan* const* plans, double* C, double* S);

/* Gaussian-state quantities of a Green's function (extension; the reference's
 * multimode.calcQuadratureGreens / calcCovarianceMtx, multimode.py:15-36, hbar = 2):
 * Z = [[Re(C+S), -Im(C-S)], [Im(C+S), Re(C-S)]] (2N x 2N real) and cov = Z Z^T.
 * C, S: N x N complex (host); Z, cov: 2N x 2N doubles (host), either may be NULL. */
int nlm_gaussian_covariance(int N, const double* C, const double* S, double* Z, double* cov);
/* out = A A^T for a real row-major M x K matrix A (host buffers), e.g. the
 * covariance of a given quadrature matrix (multimode.calcCovarianceMtx). */
int nlm_gram(int M, int K, const double* A, double* out);
/* out = A B^T for real row-major A (M x K) and B (Ncol x K), host buffers: the
 * products of the Bloch-Messiah decomposition of a quadrature matrix
 * (decompositions.py:250-353 of the reference: symplectic check S^T Omega S,
 * the polar factor and the right orthogonal factor). */
int nlm_gemm_nt(int M, int Ncol, int K, const double* A, const double* B, double* out);

/* Standalone transform with the reference's conventions (src/CustomEigenFFT.h):
 * count rows of length n, forward unscaled / inverse scaled by 1/n. */
int nlm_fft(int n, int count, const double* in, double* out, int inverse);

/* Extension: the FFT schedule variant (0..3, see csrc/nlm_fft.hpp FftShape) the
 * autotuner chose on the current device for step-table propagations of n_freqs
 * points, kind (0 Green's function, 1 batch, 2 z-resolved grid) and a field count
 * of n_fields (rounded up to a power of two); -1 if none has been tuned (always
 * for n_freqs below 1024, which run schedule 0). */
int nlm_fft_schedule(int n_freqs, int kind, int n_fields);

/* Synchronise the plan's stream. */
int nlm_plan_sync(nlm_plan* plan);

const char* nlm_last_error(void);

/* Library version string and device name of the current device. */
const char* nlm_version(void);

#ifdef __cplusplus
}
#endif
#endif
