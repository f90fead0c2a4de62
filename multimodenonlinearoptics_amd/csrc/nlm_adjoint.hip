// nlm_adjoint.hip -- C-ABI of the differentiable Green's function (see
// nlm_adjoint.hpp for the method): a forward launch that keeps checkpoints,
// then the reverse sweep, the deterministic reductions and the adjoint of the
// linear pump grid.
#include <cmath>
#include <cstring>

#include "nlm_adjoint.hpp"
#include "nlm_internal.hpp"

// g_env = N * IFFT(g_F0) + s_0 * gT[row 0]  (ifft holds IFFT(g_F0), scaled)
__global__ void k_env_grad(const cplx* __restrict__ ifft, const cplx* __restrict__ gT0, double s0, int N,
                           cplx* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) out[k] = (double)N * ifft[k] + s0 * gT0[k];
}

int nlm_fft_device(const cplx* in, cplx* out, int n, int count, bool inverse, hipStream_t s);

typedef void (*adj_fn)(const AParams&, int grid, hipStream_t);

template <int L>
static void launch_adj(const AParams& ap, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_adj_rev<L>), dim3(grid), dim3(AdjLaunch<L>::THREADS), 0, s, ap);
}

#define NLM_L_SWITCH2(FN, L) \
  switch (L) {                \
    case 3: return &FN<3>;    \
    case 4: return &FN<4>;    \
    case 5: return &FN<5>;    \
    case 6: return &FN<6>;    \
    case 7: return &FN<7>;    \
    case 8: return &FN<8>;    \
    case 9: return &FN<9>;    \
    case 10: return &FN<10>;  \
    case 11: return &FN<11>;  \
    case 12: return &FN<12>;  \
    case 13: return &FN<13>;  \
  }                           \
  return nullptr;

static adj_fn pick_adj(int L) { NLM_L_SWITCH2(launch_adj, L) }

static int adj_threads(int N) {
  const int T = N / 8;
  return T >= 256 ? T : 256;
}

extern "C" int nlm_greens_forward(nlm_plan* p, int in_time, int n_in, const int32_t* inputs, int n_out,
                                  const int32_t* outputs, const double* block_scale, int ckpt_every, double* C,
                                  double* S) {
  if (!p || !inputs || !outputs) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  if (p->model != NLM_CHI2PDC)
    return set_err(NLM_INVALID_ARGUMENT, "the adjoint Green's function is implemented for Chi2PDC");
  if (n_in != 1 || n_out != 1) return set_err(NLM_INVALID_ARGUMENT, "adjoint Green's function needs one mode");
  if (int e = check_pumps_ready(p)) return e;
  const int N = p->N;
  int K = ckpt_every;
  if (K <= 0) K = (int)std::ceil(std::sqrt((double)(p->nZ - 1 > 0 ? p->nZ - 1 : 1)));
  const int nSeg = p->nZ > 1 ? (p->nZ - 1 + K - 1) / K : 1;
  const size_t nF = 2 * (size_t)N, W = N;
  if (int e = p->G.ensure(nF * W)) return e;
  if (int e = p->C.ensure(W * N)) return e;
  if (int e = p->S.ensure(W * N)) return e;
  if (int e = p->ckpt.ensure(nF * nSeg * N)) return e;
  KParams kp;
  fill_common(p, kp);
  kp.kind = KIND_GREENS;
  kp.nFields = (int)nF;
  kp.inTime = in_time;
  kp.outTime = in_time;
  kp.nIn = 1;
  kp.nOut = 1;
  kp.out = p->G.p;
  kp.inputs[0] = inputs[0];
  kp.outPos[outputs[0]] = 0;
  kp.ckpt = p->ckpt.p;
  kp.ckptEvery = K;
  kp.nSeg = nSeg;
  tic(p, 2);
  if (int e = launch_prop_kernel(p, kp)) return e;
  tic(p, 3);
  const long total = (long)W * N;
  launch_greens_assemble(p->G.p, p->C.p, p->S.p, N, 1, 1, 1, nullptr, p->stream);
  HIP_TRY(hipGetLastError());
  tic(p, 4);
  p->greensRows = (int)W;
  p->greensCols = N;
  p->ckptEvery = K;
  p->nSeg = nSeg;
  p->gNIn = 1;
  p->gNOut = 1;
  p->gInTime = in_time;
  (void)block_scale;
  if (C) HIP_TRY(hipMemcpyAsync(C, p->C.p, total * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  if (S) HIP_TRY(hipMemcpyAsync(S, p->S.p, total * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->timing) {
    collect_timing(p, 2, 3, 1);
    collect_timing(p, 3, 4, 2);
  }
  return NLM_OK;
}

extern "C" int nlm_greens_adjoint(nlm_plan* p, const double* gC, const double* gS, double* g_coef,
                                  double* g_pump_time, double* g_envelope) {
  if (!p || !gC || !gS || !g_coef) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  if (!p->nSeg || !p->ckpt.p) return set_err(NLM_RUNTIME_ERROR, "run nlm_greens_forward before the adjoint");
  if (g_envelope && !p->pumpLinear)
    return set_err(NLM_INVALID_ARGUMENT, "envelope gradient needs the linear pump simulation");
  const int N = p->N, L = p->L, nZ = p->nZ, nZP = 2 * nZ - 1;
  const size_t nF = 2 * (size_t)N, W = N, total = W * N;
  const int TH = adj_threads(N);
  const int FA = TH / (N / 8);
  const int nWG = (int)((nF + FA - 1) / FA);
  const int nWaves = nWG * (TH / 64);
  int e;
  if ((e = p->gC.ensure(total)) || (e = p->gS.ensure(total)) || (e = p->gG.ensure(nF * W)) ||
      (e = p->gcoefPart.ensure((size_t)nWaves * nZ * 3)) || (e = p->gcoef.ensure((size_t)nZ * 3)) ||
      (e = p->stepTab.ensure((size_t)nZ * N * 2)) || (e = p->stepInv.ensure((size_t)nZ * N * 2)))
    return e;
  if (g_pump_time) {
    if ((e = p->gpumpPart.ensure((size_t)nWG * nZP * N)) || (e = p->gpump.ensure((size_t)nZP * N))) return e;
  }
  if (g_envelope) {
    if ((e = p->gF0Part.ensure((size_t)nWG * 3 * N)) || (e = p->gF0.ensure(2 * (size_t)N)) ||
        (e = p->gtmp.ensure(4 * (size_t)N)))
      return e;
  }
  HIP_TRY(hipMemcpyAsync(p->gC.p, gC, total * sizeof(cplx), hipMemcpyHostToDevice, p->stream));
  HIP_TRY(hipMemcpyAsync(p->gS.p, gS, total * sizeof(cplx), hipMemcpyHostToDevice, p->stream));
  HIP_TRY(hipMemsetAsync(p->gcoefPart.p, 0, (size_t)nWaves * nZ * 3 * sizeof(cplx), p->stream));
  hipLaunchKernelGGL(k_greens_assemble_adjoint, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, p->stream,
                     p->gC.p, p->gS.p, p->gG.p, N, 1, 1, 1, (const cplx*)nullptr);
  // the states between checkpoints are rebuilt with the inverse step propagators
  nlm_launch_step_table(NLM_CHI2PDC, p->coef.p, p->cconst.p, p->pumpT[0].p, N, nZ, p->stepTab.p, p->stream,
                        p->stepInv.p);
  HIP_TRY(hipGetLastError());
  AParams ap;
  memset(&ap, 0, sizeof ap);
  ap.nZ = nZ;
  ap.nFields = (int)nF;
  ap.K = p->ckptEvery;
  ap.nSeg = p->nSeg;
  ap.inTime = p->gInTime;
  ap.outTime = p->gInTime;
  ap.nOut = 1;
  ap.outPos[0] = 0;
  ap.outPos[1] = ap.outPos[2] = ap.outPos[3] = -1;
  ap.ckpt = p->ckpt.p;
  ap.fout = p->G.p;
  ap.inv = p->stepInv.p;
  ap.gout = p->gG.p;
  ap.gin = nullptr;
  ap.gcoef = p->gcoefPart.p;
  ap.gpump = g_pump_time ? p->gpumpPart.p : nullptr;
  if (g_envelope) {
    ap.genvF = p->gF0Part.p;                   // [nWG][2][N]
    ap.genvT = p->gF0Part.p + (size_t)nWG * 2 * N;  // [nWG][N]
    ap.Dp = p->Dp.p;
    ap.ph1 = p->gtmp.p + 2 * N;
    ap.ph2 = p->gtmp.p + 3 * N;
    ap.dzp = p->dzp;
    ap.strength = p->hasStrength ? p->strength.p : nullptr;
    hipLaunchKernelGGL(k_pump_phase_steps, dim3((N + 255) / 256), dim3(256), 0, p->stream, p->Dp.p, p->dzp, N,
                       p->gtmp.p + 2 * N, p->gtmp.p + 3 * N);
  }
  ap.tw = p->tw;
  ap.dstep = p->dstep.p;
  ap.halfIn = p->halfIn.p;
  ap.halfOut = p->halfOut.p;
  ap.coef = p->coef.p;
  ap.pump0 = p->pumpT[0].p;
  ap.tab = p->stepTab.p;
  tic(p, 2);
  pick_adj(L)(ap, nWG, p->stream);
  HIP_TRY(hipGetLastError());
  tic(p, 3);
  const long nc = (long)nZ * 3;
  hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, p->stream, p->gcoefPart.p, nc,
                     nWaves, p->gcoef.p);
  if (g_pump_time) {
    const long np = (long)nZP * N;
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, p->stream, p->gpumpPart.p,
                       np, nWG, p->gpump.p);
  }
  HIP_TRY(hipGetLastError());
  if (g_envelope) {
    // g_F0 = sum of the workgroup partials; g_env = N IFFT(g_F0) + s_0 g_row0
    hipLaunchKernelGGL(k_sum_partials, dim3((N + 255) / 256), dim3(256), 0, p->stream, ap.genvF, (long)N, 2 * nWG,
                       p->gF0.p);
    hipLaunchKernelGGL(k_sum_partials, dim3((N + 255) / 256), dim3(256), 0, p->stream, ap.genvT, (long)N, nWG,
                       p->gF0.p + N);
    HIP_TRY(hipGetLastError());
    if ((e = nlm_fft_device(p->gF0.p, p->gtmp.p, N, 1, true, p->stream))) return e;
    double s0 = 1.0;
    if (p->hasStrength) HIP_TRY(hipMemcpy(&s0, p->strength.p, sizeof(double), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_env_grad, dim3((N + 255) / 256), dim3(256), 0, p->stream, p->gtmp.p, p->gF0.p + N, s0, N,
                       p->gtmp.p + N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g_envelope, p->gtmp.p + N, N * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  }
  tic(p, 4);
  HIP_TRY(hipMemcpyAsync(g_coef, p->gcoef.p, nc * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  if (g_pump_time)
    HIP_TRY(hipMemcpyAsync(g_pump_time, p->gpump.p, (size_t)nZP * N * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->timing) {
    collect_timing(p, 2, 3, 1);
    collect_timing(p, 3, 4, 2);
  }
  return NLM_OK;
}
