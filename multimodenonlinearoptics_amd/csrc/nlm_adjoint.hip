// nlm_adjoint.hip -- C-ABI of the differentiable Green's function (see
// nlm_adjoint.hpp for the method): a forward launch that keeps checkpoints,
// then the reverse sweep, the deterministic reductions and the adjoint of the
// linear pump grid.
#include <algorithm>
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

typedef void (*sweep_fn)(const AParams&, int grid, hipStream_t);
typedef void (*env_fn)(const cplx*, int, const double*, double, const double*, const cplx*, cplx*, int grid,
                       hipStream_t);

template <int L>
static void launch_sweep(const AParams& ap, int grid, hipStream_t s) {
  constexpr int FA = AdjShape<L>::FA;
  if (ap.K == 1)
    hipLaunchKernelGGL((k_adj_sweep<L, FA, false>), dim3(grid), dim3(AdjShape<L>::THREADS), 0, s, ap);
  else
    hipLaunchKernelGGL((k_adj_sweep<L, FA, true>), dim3(grid), dim3(AdjShape<L>::THREADS), 0, s, ap);
}
template <int L>
static void launch_env(const cplx* gP, int nZP, const double* Dp, double dzp, const double* strength, const cplx* tw,
                       cplx* out, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_adj_env_rows<L>), dim3(grid), dim3(FftShape<L>::T), 0, s, gP, nZP, Dp, dzp, strength, tw,
                     out);
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

static sweep_fn pick_sweep(int L) { NLM_L_SWITCH2(launch_sweep, L) }
static env_fn pick_env(int L) { NLM_L_SWITCH2(launch_env, L) }

// The sweep keeps the default FFT schedule: with the forward's schedule variants
// (see FftShape) forced on it, the backward at N = 1024, nZ = 2000 took 81.6,
// 91.4 and 85.5 ms against 77.9 ms.

static int sweep_fields_per_wg(int L) {
  switch (L) {
    case 3: return AdjShape<3>::FA;
    case 4: return AdjShape<4>::FA;
    case 5: return AdjShape<5>::FA;
    case 6: return AdjShape<6>::FA;
    case 7: return AdjShape<7>::FA;
    case 8: return AdjShape<8>::FA;
    case 9: return AdjShape<9>::FA;
    case 10: return AdjShape<10>::FA;
    case 11: return AdjShape<11>::FA;
    case 12: return AdjShape<12>::FA;
    default: return AdjShape<13>::FA;
  }
}

// Default checkpoint interval of the differentiable Green's function: the
// smallest interval whose checkpoints (2N fields x ceil((nZ-1)/K) states x N x
// 16 B) fit in a quarter of the device memory and in half of the free memory
// (the checkpoint buffer this plan already holds counts as free), at most
// ceil(sqrt(nZ - 1)).  States that are kept need not be rebuilt in the sweep: at
// N = 1024, nZ = 2000 the backward takes 71.4 ms with every state kept (67 GB),
// 75.5 ms with every second one and 78.0 ms with every 45th (1.5 GB).
static int default_ckpt_every(const nlm_plan_s* p, size_t nF) {
  const int steps = p->nZ - 1 > 0 ? p->nZ - 1 : 1;
  const int kSqrt = (int)std::ceil(std::sqrt((double)steps));
  size_t freeB = 0, totalB = 0;
  if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) {
    (void)hipGetLastError();
    return kSqrt;
  }
  const size_t held = p->ckpt.n * sizeof(cplx);
  const size_t budget = std::min(totalB / 4, (freeB + held) / 2);
  const size_t perState = nF * (size_t)p->N * sizeof(cplx);
  const size_t fit = budget / perState;
  if (fit == 0) return kSqrt;
  const size_t K = ((size_t)steps + fit - 1) / fit;
  return (int)std::max<size_t>(1, std::min<size_t>(K, (size_t)kSqrt));
}

extern "C" int nlm_greens_forward(nlm_plan* p, int in_time, int n_in, const int32_t* inputs, int n_out,
                                  const int32_t* outputs, const double* block_scale, int ckpt_every, double* C,
                                  double* S) {
  if (!p || !inputs || !outputs) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  if (p->model != NLM_CHI2PDC)
    return set_err(NLM_INVALID_ARGUMENT, "the adjoint Green's function is implemented for Chi2PDC");
  if (n_in != 1 || n_out != 1) return set_err(NLM_INVALID_ARGUMENT, "adjoint Green's function needs one mode");
  if (p->L == 0)
    return set_err(NLM_INVALID_ARGUMENT, "the adjoint Green's function needs nFreqs a power of two (8..8192)");
  if (int e = check_pumps_ready(p)) return e;
  const int N = p->N;
  const size_t nF = 2 * (size_t)N, W = N;
  const int K = ckpt_every > 0 ? ckpt_every : default_ckpt_every(p, nF);
  const int nSeg = p->nZ > 1 ? (p->nZ - 1 + K - 1) / K : 1;
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
  p->adjGen = ++p->gen;
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
  if (p->adjGen != p->gen)
    return set_err(NLM_RUNTIME_ERROR,
                   "the medium changed after nlm_greens_forward (coefficients, pump or another Green's run); run the "
                   "forward again before the adjoint");
  if (g_envelope && !p->pumpLinear)
    return set_err(NLM_INVALID_ARGUMENT, "envelope gradient needs the linear pump simulation");
  const int N = p->N, L = p->L, nZ = p->nZ, nZP = 2 * nZ - 1;
  const size_t nF = 2 * (size_t)N, W = N, total = W * N;
  const int FA = sweep_fields_per_wg(L);
  const int nWG = (int)((nF + FA - 1) / FA);
  // steps per sweep launch: the workgroup partials of one launch are kept within
  // 2 GiB where one step's partials fit (fewer, longer launches: 1 GiB chunks cost
  // 0.3 % more, 0.5 GiB chunks 10 % more at N = 1024, nZ = 2000).  One step's
  // partials are nWG * 2N * 16 B: 128 MiB at N = 2048 (16 steps per launch), 1 GiB
  // at N = 4096, 4 GiB at N = 8192, where a launch holds one step and the buffer
  // is 4 GiB
  const size_t perStep = (size_t)nWG * 2 * N;
  int chunk = (int)std::max<size_t>(1, ((size_t)1 << 31) / (perStep * sizeof(cplx)));
  if (chunk > nZ - 1) chunk = nZ - 1 > 0 ? nZ - 1 : 1;
  // The reduction's (step, tile) blocks alone leave most CUs idle where one step
  // has many workgroup partials (N = 4096: 2 steps x 16 tiles = 32 blocks reading
  // 2 GiB per launch); below 256 such blocks a first level sums slices of at
  // least 64 workgroups each (fixed order), ~256 blocks per launch, each slice a
  // multiple of the 32 workgroups of one pass of the accumulators where it can be
  // (N = 4096, nZ = 100, backward: 142 ms without; 84, 86, 91, 98 ms with 256,
  // 512, 1024, 2048 blocks; 88 ms with 384, slices of 683 workgroups).
  const int tilesR = N >= 256 ? N / 256 : (N + 63) / 64;
  int per = nWG;
  if (chunk * tilesR < 256) {
    const int want = (256 + chunk * tilesR - 1) / (chunk * tilesR);
    const int S = std::max(1, std::min(want, nWG / 64));
    per = ((nWG + S - 1) / S + RED_WAVES * RED_ACC - 1) / (RED_WAVES * RED_ACC) * (RED_WAVES * RED_ACC);
  }
  const int slices = (nWG + per - 1) / per;
  const bool rows = g_pump_time || g_envelope;
  const int nEnv = std::min(nZP - 1 > 0 ? nZP - 1 : 1, 512);
  int e;
  if ((e = p->gC.ensure(total)) || (e = p->gS.ensure(total)) || (e = p->gG.ensure(nF * W)) ||
      (e = p->gcoefPart.ensure((size_t)nZ * 3 * N)) || (e = p->gcoef.ensure((size_t)nZ * 3)) ||
      (e = p->stepTab.ensure((size_t)nZ * N * 2)) ||
      (e = p->gpumpPart.ensure(perStep * chunk)) || (e = p->gtmp.ensure(nF * 2 * N)))
    return e;
  if (slices > 1 && (e = p->gslice.ensure((size_t)slices * chunk * 2 * N))) return e;
  if (rows && (e = p->gpump.ensure((size_t)nZP * N))) return e;
  if (g_envelope && ((e = p->gF0Part.ensure((size_t)nEnv * N)) || (e = p->gF0.ensure(3 * (size_t)N)))) return e;
  HIP_TRY(hipMemcpyAsync(p->gC.p, gC, total * sizeof(cplx), hipMemcpyHostToDevice, p->stream));
  HIP_TRY(hipMemcpyAsync(p->gS.p, gS, total * sizeof(cplx), hipMemcpyHostToDevice, p->stream));
  hipLaunchKernelGGL(k_greens_assemble_adjoint, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, p->stream,
                     p->gC.p, p->gS.p, p->gG.p, N, 1, 1, 1, (const cplx*)nullptr);
  // the step propagators (the states between checkpoints are rebuilt by inverting them)
  nlm_launch_step_table(NLM_CHI2PDC, p->coef.p, p->cconst.p, p->pumpT[0].p, N, nZ, p->stepTab.p, p->stream);
  HIP_TRY(hipGetLastError());
  AParams ap;
  memset(&ap, 0, sizeof ap);
  ap.nZ = nZ;
  ap.nFields = (int)nF;
  ap.K = p->ckptEvery;
  ap.nSeg = p->nSeg;
  ap.outTime = p->gInTime;
  ap.chunk = chunk;
  ap.ckpt = p->ckpt.p;
  ap.fout = p->G.p;
  ap.gout = p->gG.p;
  ap.state = p->gtmp.p;
  ap.part = p->gpumpPart.p;
  ap.tab = p->stepTab.p;
  ap.tw = p->tw;
  ap.dstep = p->dstep.p;
  ap.halfOut = p->halfOut.p;
  tic(p, 2);
  if (nZ > 1) {  // gm of step 0 is never read (it has no RK stages)
    const sweep_fn sweep = pick_sweep(L);
    for (int iHi = nZ - 1; iHi >= 1; iHi -= chunk) {
      ap.iHi = iHi;
      ap.iLo = iHi - chunk > 0 ? iHi - chunk : 0;
      sweep(ap, nWG, p->stream);
      const int count = ap.iHi - ap.iLo;
      const unsigned blocks = (unsigned)(count * tilesR);
      const cplx* sums = ap.part;
      int nSums = nWG;
      if (slices > 1) {
        if (N >= 256)
          hipLaunchKernelGGL(k_adj_slices<4>, dim3(blocks * slices), dim3(64 * RED_WAVES), 0, p->stream, ap.part,
                             nWG, chunk, N, per, p->gslice.p);
        else
          hipLaunchKernelGGL(k_adj_slices<1>, dim3(blocks * slices), dim3(64 * RED_WAVES), 0, p->stream, ap.part,
                             nWG, chunk, N, per, p->gslice.p);
        sums = p->gslice.p;
        nSums = slices;
      }
      if (N >= 256)  // four points per lane: 1 KB per load instruction, 4 KB per wave and partial
        hipLaunchKernelGGL(k_adj_reduce<4>, dim3(blocks), dim3(64 * RED_WAVES), 0, p->stream, sums, nSums, chunk,
                           ap.iLo, N, p->coef.p, p->pumpT[0].p, p->gcoefPart.p);
      else
        hipLaunchKernelGGL(k_adj_reduce<1>, dim3(blocks), dim3(64 * RED_WAVES), 0, p->stream, sums, nSums, chunk,
                           ap.iLo, N, p->coef.p, p->pumpT[0].p, p->gcoefPart.p);
    }
  }
  HIP_TRY(hipGetLastError());
  tic(p, 3);
  hipLaunchKernelGGL(k_adj_coef, dim3((unsigned)nZ), dim3(256), 0, p->stream, p->gcoefPart.p, p->pumpT[0].p, N, nZ,
                     p->gcoef.p);
  if (rows) {
    const long np = (long)nZP * N;
    hipLaunchKernelGGL(k_adj_pump_rows, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, p->stream,
                       p->gcoefPart.p, p->coef.p, N, nZ, p->gpump.p);
  }
  HIP_TRY(hipGetLastError());
  if (g_envelope) {
    // g_F0 = sum over rows r >= 1 (workgroup partials); g_env = N IFFT(g_F0) + s_0 g_P[0]
    const double* strength = p->hasStrength ? p->strength.p : nullptr;
    if (nZP > 1) {
      pick_env(L)(p->gpump.p, nZP, p->Dp.p, p->dzp, strength, p->tw, p->gF0Part.p, nEnv, p->stream);
      hipLaunchKernelGGL(k_sum_partials, dim3((N + 255) / 256), dim3(256), 0, p->stream, p->gF0Part.p, (long)N, nEnv,
                         p->gF0.p);
    } else {
      HIP_TRY(hipMemsetAsync(p->gF0.p, 0, (size_t)N * sizeof(cplx), p->stream));
    }
    HIP_TRY(hipGetLastError());
    if ((e = nlm_fft_device(p->gF0.p, p->gF0.p + N, N, 1, true, p->stream))) return e;
    double s0 = 1.0;
    if (p->hasStrength) HIP_TRY(hipMemcpy(&s0, p->strength.p, sizeof(double), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_env_grad, dim3((N + 255) / 256), dim3(256), 0, p->stream, p->gF0.p + N, p->gpump.p, s0, N,
                       p->gF0.p + 2 * N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g_envelope, p->gF0.p + 2 * N, N * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  }
  tic(p, 4);
  HIP_TRY(hipMemcpyAsync(g_coef, p->gcoef.p, (size_t)nZ * 3 * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  if (g_pump_time)
    HIP_TRY(hipMemcpyAsync(g_pump_time, p->gpump.p, (size_t)nZP * N * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->timing) {
    collect_timing(p, 2, 3, 1);
    collect_timing(p, 3, 4, 2);
  }
  return NLM_OK;
}
