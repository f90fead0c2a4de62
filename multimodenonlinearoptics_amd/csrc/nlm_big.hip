// nlm_big.hip -- Chi3GNLSE batches with nFreqs = P * 4096 (16384, 32768): the
// field no longer fits one workgroup's LDS, so every transform is staged
// through HBM/L2 as a four-step FFT with N = P * M, M = 4096:
//
//   X[k1 + P k2] = sum_n2 W_M^(n2 k2) [ W_N^(n2 k1) sum_n1 W_P^(n1 k1) x[n1 M + n2] ]
//
// "mix" kernels (one thread per n2 of a field) apply the radix-P butterflies
// across the P blocks with the W_N twiddles, fused with the pointwise
// time-domain work of the RK4 stage that sits between two transforms; "conv"
// kernels (one workgroup per (field, k1)) run the 4096-point transform in LDS
// (nlm_fft.hpp), multiply by a frequency response stored in the same
// k1-major order, and transform back.  Between them the data sit in the
// exchange arrays X, Z ([field][k1][n2], coalesced on both sides).  One
// Chi3GNLSE step (solvers/Chi3GNLSE.cpp:39-92 inside
// src/_NonlinearMedium.hpp:205-215) is 9 conv + 9 mix launches; G steps are
// captured in one hipGraph whose kernels read the step index from a device
// counter, and the graph is replayed ceil((nZ-1)/G) times.
#include <cstring>

#include "nlm_internal.hpp"
#include "nlm_prop.hpp"

namespace {

constexpr int BIG_LM = 12;             // log2 M
constexpr int BIG_M = 1 << BIG_LM;     // local transform length
constexpr int MIX_THREADS = 256;

enum MixMode { MIX_FWD_IN = 0, MIX_RAMAN, MIX_STAGE0, MIX_STAGE1, MIX_STAGE2, MIX_LAST, MIX_DISP };
enum ConvMode { CONV_RAMAN = 0, CONV_FILTER, CONV_DISP, CONV_PRO_TIME, CONV_PRO_FREQ };

struct BigParams {
  int B, nSteps, N;
  const int* base;      // current first step of the replayed graph
  cplx* A;              // [B][N] state a (time, natural order)
  cplx* Y;              // [B][N] stage input y
  cplx* ACC;            // [B][N] k1 + 2 k2 + 2 k3
  cplx* X;              // [B][P][M] transform input exchange
  cplx* Z;              // [B][P][M] transform output exchange
  const cplx* twN;      // W_N^k
  const cplx* twM;      // W_M^k
  const cplx* Hp;       // [5][P][M] frequency responses, k1-major: raman, filter, disp, halfIn, disp*halfOut
  const cplx* coef;     // [nZ][3], .re = relative intensity of the stage
  const cplx* in;       // [B][inCols]
  int inCols, inPos;
  cplx* out;            // [B][nOut*N]
  int nOut, outPos, outTime;
};

__device__ __forceinline__ int step_of(const BigParams& bp, int di, int fixedStep) {
  return fixedStep >= 0 ? fixedStep : *bp.base + di;
}

template <int P>
__global__ void __launch_bounds__(MIX_THREADS) k_big_mix(BigParams bp, int mode, int di, int fixedStep) {
  const int i = step_of(bp, di, fixedStep);
  if (fixedStep < 0 && i > bp.nSteps) return;
  const long gid = (long)blockIdx.x * MIX_THREADS + threadIdx.x;
  if (gid >= (long)bp.B * BIG_M) return;
  const int f = (int)(gid / BIG_M), n2 = (int)(gid % BIG_M);
  const int N = bp.N;
  const size_t fo = (size_t)f * N;
  cplx v[P];
  if (mode != MIX_FWD_IN) {
    // inverse radix-P across the blocks: x[n1 M + n2] = 1/N sum_k1 W_P^(-n1 k1) W_N^(-n2 k1) z_k1[n2]
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) v[k1] = bp.Z[fo + (size_t)k1 * BIG_M + n2] * conj(bp.twN[n2 * k1]);
    dft_small<P, true>(v);
    const double s = 1.0 / N;
#pragma unroll
    for (int n1 = 0; n1 < P; n1++) v[n1] = v[n1] * s;
  }
  cplx u[P];
  switch (mode) {
    case MIX_FWD_IN:
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) u[n1] = bp.in[(size_t)f * bp.inCols + (size_t)bp.inPos * N + n1 * BIG_M + n2];
      break;
    case MIX_RAMAN:  // IFFT(FFT(r|y|^2) R) * y
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) u[n1] = v[n1] * bp.Y[fo + n1 * BIG_M + n2];
      break;
    case MIX_STAGE0:
    case MIX_STAGE1:
    case MIX_STAGE2: {
      // k_s = v; acc = ((k1 + 2 k2) + 2 k3); next stage input y = a + (1/2 | 1/2 | 1) k_s
      const int s = mode - MIX_STAGE0;
      const double h = s < 2 ? 0.5 : 1.0;
      const double r = s < 2 ? bp.coef[(size_t)i * 3 + 1].re : bp.coef[(size_t)i * 3 + 2].re;
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * BIG_M + n2;
        bp.ACC[q] = s == 0 ? v[n1] : bp.ACC[q] + 2.0 * v[n1];
        const cplx y = bp.A[q] + h * v[n1];
        bp.Y[q] = y;
        u[n1] = mk(r * abs2(y), 0.0);
      }
      break;
    }
    case MIX_LAST:  // a + (acc + k4) / 6, the input of the dispersion step
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * BIG_M + n2;
        u[n1] = bp.A[q] + (bp.ACC[q] + v[n1]) * (1. / 6.);
      }
      break;
    default: {  // MIX_DISP: new state a_i; first stage of step i+1 starts with r0 |a|^2
      if (i == bp.nSteps) {
        if (bp.outTime) {
#pragma unroll
          for (int n1 = 0; n1 < P; n1++)
            bp.out[(size_t)f * bp.nOut * N + (size_t)bp.outPos * N + n1 * BIG_M + n2] = v[n1];
        }
        return;
      }
      const double r = bp.coef[(size_t)(i + 1) * 3].re;
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * BIG_M + n2;
        bp.A[q] = v[n1];
        bp.Y[q] = v[n1];
        u[n1] = mk(r * abs2(v[n1]), 0.0);
      }
      break;
    }
  }
  // forward radix-P across the blocks, then the W_N^(n2 k1) twiddle
  dft_small<P, false>(u);
#pragma unroll
  for (int k1 = 0; k1 < P; k1++) bp.X[fo + (size_t)k1 * BIG_M + n2] = u[k1] * bp.twN[n2 * k1];
}

template <int P>
__global__ void __launch_bounds__(Launch<BIG_LM>::THREADS) k_big_conv(BigParams bp, int mode, int di, int fixedStep) {
  using F = FftShape<BIG_LM>;
  constexpr int T = F::T;
  __shared__ cplx lds[F::lds_elems];
  const int i = step_of(bp, di, fixedStep);
  if (fixedStep < 0 && i > bp.nSteps) return;
  const int f = blockIdx.x / P, k1 = blockIdx.x % P, t = threadIdx.x;
  const int N = bp.N;
  const size_t xo = (size_t)f * N + (size_t)k1 * BIG_M;
  FftTwiddles<BIG_LM> tw;
  fft_load_twiddles<BIG_LM>(tw, bp.twM, t);
  cplx v[8];
  if (mode == CONV_PRO_FREQ) {
#pragma unroll
    for (int m = 0; m < 8; m++)
      v[m] = bp.in[(size_t)f * bp.inCols + (size_t)bp.inPos * N + k1 + (size_t)P * (t + T * m)];
  } else {
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = bp.X[xo + t + T * m];
    fft_field<BIG_LM, false>(v, lds, t, tw);
  }
  int table = mode == CONV_RAMAN ? 0 : mode == CONV_FILTER ? 1 : mode == CONV_DISP ? 2 : 3;
  const bool last = mode == CONV_DISP && i == bp.nSteps;
  if (last) table = 4;  // exp(i dz D) exp(-i dz/2 D): the final half step out
  const cplx* H = bp.Hp + (size_t)table * N + (size_t)k1 * BIG_M;
#pragma unroll
  for (int m = 0; m < 8; m++) v[m] = v[m] * H[t + T * m];
  if (last && !bp.outTime) {
#pragma unroll
    for (int m = 0; m < 8; m++)
      bp.out[(size_t)f * bp.nOut * N + (size_t)bp.outPos * N + k1 + (size_t)P * (t + T * m)] = v[m];
    return;
  }
  fft_field<BIG_LM, true>(v, lds, t, tw);  // unscaled; the mix kernel applies 1/N
#pragma unroll
  for (int m = 0; m < 8; m++) bp.Z[xo + t + T * m] = v[m];
}

__global__ void k_big_set_base(int* base, int value, int add) {
  if (blockIdx.x == 0 && threadIdx.x == 0) base[0] = add ? base[0] + value : value;
}

// Hp[tab][k1 M + k2] = H[k1 + P k2]  (tab 4: disp * halfOut)
__global__ void k_big_perm(const cplx* __restrict__ ram, const cplx* __restrict__ flt, const cplx* __restrict__ ds,
                           const cplx* __restrict__ hin, const cplx* __restrict__ hout, int N, int P,
                           cplx* __restrict__ Hp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int k1 = j / BIG_M, k2 = j % BIG_M, k = k1 + P * k2;
  Hp[j] = ram[k];
  Hp[(size_t)N + j] = flt[k];
  Hp[2 * (size_t)N + j] = ds[k];
  Hp[3 * (size_t)N + j] = hin[k];
  Hp[4 * (size_t)N + j] = ds[k] * hout[k];
}

template <int P>
void launch_step(const BigParams& bp, int di, hipStream_t s) {
  const dim3 gm((unsigned)(((long)bp.B * BIG_M + MIX_THREADS - 1) / MIX_THREADS)), bm(MIX_THREADS);
  const dim3 gc((unsigned)(bp.B * P)), bc(Launch<BIG_LM>::THREADS);
  const int mixAfterFilter[4] = {MIX_STAGE0, MIX_STAGE1, MIX_STAGE2, MIX_LAST};
  for (int st = 0; st < 4; st++) {
    hipLaunchKernelGGL(k_big_conv<P>, gc, bc, 0, s, bp, (int)CONV_RAMAN, di, -1);
    hipLaunchKernelGGL(k_big_mix<P>, gm, bm, 0, s, bp, (int)MIX_RAMAN, di, -1);
    hipLaunchKernelGGL(k_big_conv<P>, gc, bc, 0, s, bp, (int)CONV_FILTER, di, -1);
    hipLaunchKernelGGL(k_big_mix<P>, gm, bm, 0, s, bp, mixAfterFilter[st], di, -1);
  }
  hipLaunchKernelGGL(k_big_conv<P>, gc, bc, 0, s, bp, (int)CONV_DISP, di, -1);
  hipLaunchKernelGGL(k_big_mix<P>, gm, bm, 0, s, bp, (int)MIX_DISP, di, -1);
}

template <int P>
void launch_prologue(const BigParams& bp, bool inTime, hipStream_t s) {
  const dim3 gm((unsigned)(((long)bp.B * BIG_M + MIX_THREADS - 1) / MIX_THREADS)), bm(MIX_THREADS);
  const dim3 gc((unsigned)(bp.B * P)), bc(Launch<BIG_LM>::THREADS);
  if (inTime) {
    hipLaunchKernelGGL(k_big_mix<P>, gm, bm, 0, s, bp, (int)MIX_FWD_IN, 0, 0);
    hipLaunchKernelGGL(k_big_conv<P>, gc, bc, 0, s, bp, (int)CONV_PRO_TIME, 0, 0);
  } else {
    hipLaunchKernelGGL(k_big_conv<P>, gc, bc, 0, s, bp, (int)CONV_PRO_FREQ, 0, 0);
  }
  hipLaunchKernelGGL(k_big_mix<P>, gm, bm, 0, s, bp, (int)MIX_DISP, 0, 0);
}

}  // namespace

// Steps per captured graph.
static constexpr int BIG_GRAPH_STEPS = 32;

bool nlm_big_supported(int model, int L) { return model == NLM_CHI3GNLSE && (L == 14 || L == 15); }

// Batch propagation of a Chi3GNLSE plan with nFreqs 16384 or 32768 (see above).
// in / out are device arrays with the layout of nlm_batch_dev.
int nlm_big_batch(nlm_plan* p, const cplx* in, int rows, int cols, int inTime, int inPos, int nOut, int outPos,
                  cplx* out) {
  const int N = p->N, P = N / BIG_M;
  if (P != 4 && P != 8) return set_err(NLM_INVALID_ARGUMENT, "unsupported size for the staged transform");
  if (p->nZ < 2) return set_err(NLM_INVALID_ARGUMENT, "the staged transform needs at least two z steps");
  if (rows == 0) return NLM_OK;
  const size_t BN = (size_t)rows * N;
  int e;
  if ((e = p->bigState.ensure(5 * BN)) || (e = p->bigH.ensure(5 * (size_t)N)) || (e = p->bigBase.ensure(1))) return e;
  const cplx* twM = nullptr;
  if ((e = get_twiddles(BIG_M, &twM))) return e;
  hipLaunchKernelGGL(k_big_perm, dim3((N + 255) / 256), dim3(256), 0, p->stream, p->raman.p, p->filt.p, p->dstep.p,
                     p->halfIn.p, p->halfOut.p, N, P, p->bigH.p);
  HIP_TRY(hipGetLastError());
  BigParams bp;
  std::memset(&bp, 0, sizeof bp);
  bp.B = rows;
  bp.nSteps = p->nZ - 1;
  bp.N = N;
  bp.base = p->bigBase.p;
  bp.A = p->bigState.p;
  bp.Y = bp.A + BN;
  bp.ACC = bp.Y + BN;
  bp.X = bp.ACC + BN;
  bp.Z = bp.X + BN;
  bp.twN = p->tw;
  bp.twM = twM;
  bp.Hp = p->bigH.p;
  bp.coef = p->coef.p;
  bp.in = in;
  bp.inCols = cols;
  bp.inPos = inPos;
  bp.out = out;
  bp.nOut = nOut;
  bp.outPos = outPos;
  bp.outTime = inTime;
  if (P == 4) launch_prologue<4>(bp, inTime != 0, p->stream);
  else launch_prologue<8>(bp, inTime != 0, p->stream);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_big_set_base, dim3(1), dim3(64), 0, p->stream, p->bigBase.p, 1, 0);
  HIP_TRY(hipGetLastError());

  // the z loop: a graph of BIG_GRAPH_STEPS steps, rebuilt when its arguments change
  const int G = BIG_GRAPH_STEPS;
  if (!p->bigExec || p->bigKey[0] != rows || p->bigKey[1] != cols || p->bigKey[2] != inTime ||
      p->bigKey[3] != inPos || p->bigKey[4] != nOut || p->bigKey[5] != outPos || p->bigIn != in ||
      p->bigOut != out || p->bigStatePtr != p->bigState.p) {
    if (p->bigExec) (void)hipGraphExecDestroy(p->bigExec);
    p->bigExec = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    for (int di = 0; di < G; di++) {
      if (P == 4) launch_step<4>(bp, di, p->stream);
      else launch_step<8>(bp, di, p->stream);
    }
    hipLaunchKernelGGL(k_big_set_base, dim3(1), dim3(64), 0, p->stream, p->bigBase.p, G, 1);
    const hipError_t ce = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(p->stream, &g);
    if (ce != hipSuccess || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return set_err(NLM_DEVICE_ERROR,
                     std::string("graph capture failed: ") + hipGetErrorString(ce != hipSuccess ? ce : ee));
    }
    const hipError_t ie = hipGraphInstantiate(&p->bigExec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      p->bigExec = nullptr;
      return set_err(NLM_DEVICE_ERROR, std::string("graph instantiation failed: ") + hipGetErrorString(ie));
    }
    const int key[6] = {rows, cols, inTime, inPos, nOut, outPos};
    std::memcpy(p->bigKey, key, sizeof key);
    p->bigIn = in;
    p->bigOut = out;
    p->bigStatePtr = p->bigState.p;
  }
  const int launches = (p->nZ - 1 + G - 1) / G;
  for (int r = 0; r < launches; r++) HIP_TRY(hipGraphLaunch(p->bigExec, p->stream));
  return NLM_OK;
}
