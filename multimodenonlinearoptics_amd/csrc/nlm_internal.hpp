// nlm_internal.hpp -- plan state and host helpers shared by the engine's
// translation units (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nlm.h"
#include "nlm_prop.hpp"

int set_err(int kind, const std::string& msg);

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return set_err(NLM_DEVICE_ERROR, std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int ensure(size_t count) {
    if (count <= n && p) return NLM_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) return NLM_OK;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) return set_err(NLM_DEVICE_ERROR, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    n = count;
    return NLM_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct nlm_plan_s {
  int model = 0, N = 0, L = 0, nZ = 0, nS = 0, nP = 0, NC = 0, NCC = 0;
  int device = 0;
  hipStream_t own = nullptr, stream = nullptr;
  const cplx* tw = nullptr;
  DevBuf<cplx> dstep, halfIn, halfOut, coef, cconst, raman, filt;
  DevBuf<cplx> pumpT[2], pumpF[2];
  DevBuf<cplx> gridF, gridT;
  DevBuf<cplx> G, C, S, in, out, env, F0, dsp, relI;
  DevBuf<double> Dp, strength;
  int greensRows = 0, greensCols = 0;
  bool timing = false;
  hipEvent_t ev[6] = {};
  double last_ms[3] = {0, 0, 0};
  // Green's forward kept for the adjoint
  bool pumpLinear = false, hasStrength = false;
  double dzp = 0;
  DevBuf<cplx> ckpt, gG, gC, gS, gcoefPart, gcoef, gpumpPart, gpump, gF0Part, gF0, gtmp, gslice;
  DevBuf<cplx> scale;
  // staged transform for Chi3GNLSE with nFreqs > 8192 (nlm_big.hip)
  DevBuf<cplx> bigState, bigH;
  DevBuf<int> bigBase;
  hipGraphExec_t bigExec = nullptr;
  int bigKey[6] = {0, 0, 0, 0, 0, 0};
  const void* bigIn = nullptr;
  const void* bigOut = nullptr;
  const void* bigStatePtr = nullptr;
  DevBuf<cplx> anyScratch;  // workspaces of the any-length path (nlm_any.hip) when they exceed LDS
  DevBuf<cplx> stepTab;  // RK4 step propagators of one-mode undepleted media [nZ][N][2]
  int stepTable = 1;      // use them (nlm_plan_set_step_table)
  int ckptEvery = 0, nSeg = 0, gNIn = 0, gNOut = 0, gInTime = 0, gHasScale = 0;
  // differentiable Chi3GNLSE batch (nlm_batch_forward / nlm_batch_adjoint)
  int bRows = 0;
  DevBuf<cplx> adjWork, adjSeg;
  // state generation: bumped by every change the adjoints depend on (coefficients,
  // pump grids, the Green's rows in G); adjGen is the generation a differentiable
  // forward left, which its adjoint requires unchanged
  unsigned long long gen = 0, adjGen = 0;
};

inline void plan_touch(nlm_plan_s* p) { ++p->gen; }

void fill_common(nlm_plan_s* p, KParams& kp);
int check_pumps_ready(nlm_plan_s* p);
int launch_prop_kernel(nlm_plan_s* p, const KParams& kp);
// the same with the step-propagator table [nZ][N][2] in a caller's buffer (used when
// uses_step_table(p)); allocates nothing
int launch_prop_kernel_tab(nlm_plan_s* p, const KParams& kp, cplx* tab, bool may_tune = false);
bool uses_step_table(const nlm_plan_s* p);
int get_twiddles(int n, const cplx** out);
// input channel feeding each mode (-1: none), src/_NonlinearMedium.hpp:164-198
void input_channels(int nS, int nCh, int inputMode, int* inPos);
// Chi3GNLSE batches with nFreqs 16384 / 32768 (nlm_big.hip)
bool nlm_big_supported(int model, int L);
int nlm_big_batch(nlm_plan_s* p, const cplx* in, int rows, int cols, int inTime, int inPos, int nOut, int outPos,
                  cplx* out);
// Any even nFreqs outside the power-of-two kernels (nlm_any.hip): not a power of
// two, 2 and 4, or above 8192.  Plans of these lengths have L = 0.
bool nlm_any_size(int n);
// may_alloc = false: fail instead of allocating the device workspace that
// lengths above the LDS workspace need
int nlm_any_prop(nlm_plan_s* p, const KParams& kp, bool may_alloc = true);
int nlm_any_pump_linear(nlm_plan_s* p, const double* strength_dev, double dzp);
int nlm_any_pump_chi3(nlm_plan_s* p, double dzp);
int nlm_any_pump_import(nlm_plan_s* p, int q, const cplx* sigF, int nzs, double dzs, const double* Ds,
                        const double* Dp, const double* omega, double dzp, double bd, double z, double zs);
int nlm_any_fft(const cplx* in, cplx* out, int n, int count, bool inverse, hipStream_t s, DevBuf<cplx>* scratch);
void tic(nlm_plan_s* p, int slot);
void collect_timing(nlm_plan_s* p, int a0, int a1, int slot);
void launch_greens_assemble(const cplx* G, cplx* C, cplx* S, int N, int nIn, int nOut, int nb, const cplx* scale,
                            hipStream_t s);
