// nlm_big.hip -- Chi3GNLSE batches whose transforms are staged through HBM/L2,
// and the differentiable (checkpointed forward + adjoint) Chi3GNLSE batch.
//
// Staged transform.  For nFreqs = P * 4096 (16384, 32768) the field no longer
// fits one workgroup's LDS, so every transform is a four-step FFT with
// N = P * M, M = 4096:
//
//   X[k1 + P k2] = sum_n2 W_M^(n2 k2) [ W_N^(n2 k1) sum_n1 W_P^(n1 k1) x[n1 M + n2] ]
//
// "mix" kernels (one thread per n2 of a field) apply the radix-P butterflies
// across the P blocks with the W_N twiddles, fused with the pointwise
// time-domain work of the RK4 stage that sits between two transforms; "conv"
// kernels (one workgroup per (field, k1)) run the M-point transform in LDS
// (nlm_fft.hpp), multiply by a frequency response stored in the same
// k1-major order, and transform back.  Between them the data sit in the
// exchange arrays X, Z ([field][k1][n2], coalesced on both sides).  One
// Chi3GNLSE step (solvers/Chi3GNLSE.cpp:39-92 inside
// src/_NonlinearMedium.hpp:205-215) is 9 conv + 9 mix launches; G steps are
// captured in one hipGraph whose kernels read the step index from a device
// counter, and the graph is replayed ceil((nZ-1)/G) times.  With P = 1 the same
// kernels run any nFreqs up to 8192 (M = N, the mix kernels are pointwise).
//
// Adjoint.  Gradients use g_z = dL/dRe z + i dL/dIm z.  The forward keeps the
// time-domain state every K steps.  The reverse sweep takes the segments last
// to first: it recomputes the segment from its checkpoint, storing per step the
// state a, the stage inputs y2, y3, y4 and the Raman convolutions v1..v4, then
// walks the segment backwards.  With f(y) = IFFT(H FFT(y v)), v = IFFT(R FFT(r |y|^2))
// (H = nl0 + nl1 omega, R the Raman response), the adjoint of one stage is
//   g_u = IFFT(conj(H) FFT(g_k)),  g_s = Re IFFT(conj(R) FFT(conj(y) g_u)),
//   g_y = conj(v) g_u + 2 r y g_s,
// and of the dispersion step g = IFFT(conj(D) FFT(g)): 9 convolutions per step,
// as many as the forward step, each again a conv + mix pair.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "nlm_internal.hpp"
#include "nlm_prop.hpp"

namespace {

constexpr int BIG_LM = 12;  // log2 M of the staged transform
constexpr int MIX_THREADS = 256;

// mix kinds (mode & 15; mode >> 4 = RK stage 0..3)
enum MixKind {
  MIX_FWD_IN = 0,   // X = fwd(input row)
  MIX_RAMAN,        // v = raman conv; X = fwd(v y)
  MIX_STAGE0,       // k1 -> acc, y2 = a + k1/2; X = fwd(r1 |y2|^2)
  MIX_STAGE1,       // k2 -> acc, y3 = a + k2/2
  MIX_STAGE2,       // k3 -> acc, y4 = a + k3
  MIX_LAST,         // a + (acc + k4)/6 -> X
  MIX_DISP,         // new state a_i; X = fwd(r0 |a_i|^2) of step i+1
  MIX_CKPT_IN,      // MIX_DISP from the checkpoint of step i
  MIX_ADJ_IN,       // X = fwd(g_out row) (time-domain output)
  MIX_ADJ_G0,       // g of the RK4 result: GP = GA = g; X = fwd(g/6)
  MIX_ADJ_U,        // g_u: TMP = conj(v) g_u; X = fwd(conj(y) g_u)
  MIX_ADJ_S,        // g_s: g_y = TMP + 2 r y Re g_s; GA += g_y; X = fwd(next g_k or GA)
  MIX_ADJ_OUT,      // g_in = finished transform (time-domain input)
  MIX_ADJ_RESUME    // X = fwd(GA)
};
// conv mode bits: response table (0 raman, 1 filter, 2 disp, 3 halfIn, 4 disp*halfOut)
constexpr int CV_TABLE = 7;
constexpr int CV_CONJ = 8;       // multiply by conj(H)
constexpr int CV_SRC_IN = 16;    // read the frequency-domain input row instead of X
constexpr int CV_SRC_GOUT = 32;  // read N * g_out (frequency-domain output) instead of X
constexpr int CV_DST_GIN = 64;   // write g_in = conj(H) FFT(x) / N (frequency-domain input)
constexpr int CV_DISP_DYN = 128; // dispersion step: table 4 and the output on the last step
constexpr int CONV_RAMAN = 0, CONV_FILTER = 1, CONV_DISP = 2 | CV_DISP_DYN, CONV_PRO_TIME = 3,
              CONV_PRO_FREQ = 3 | CV_SRC_IN;

struct BigParams {
  int B, nSteps, N, M, LM;  // M = N / P local length, LM = log2 M
  const int* base;          // current first step of the replayed graph
  cplx* A;                  // [B][N] state a (time, natural order)
  cplx* Y;                  // [B][N] stage input y
  cplx* ACC;                // [B][N] k1 + 2 k2 + 2 k3
  cplx* X;                  // [B][P][M] transform input exchange
  cplx* Z;                  // [B][P][M] transform output exchange
  const cplx* twN;          // W_N^k
  const cplx* twM;          // W_M^k
  const cplx* Hp;           // [5][P][M] frequency responses, k1-major
  const cplx* coef;         // [nZ][3], .re = relative intensity of the stage
  const cplx* in;           // [B][inCols]
  int inCols, inPos;
  cplx* out;                // [B][nOut*N] or nullptr (recomputation)
  int nOut, outPos, outTime;
  // checkpoints [B][nSeg][N] of the states a_{sK}: written by MIX_DISP when
  // ckptWrite is set, read by MIX_CKPT_IN
  cplx* ckpt;
  int ckptEvery, nSeg, ckptWrite;
  // segment storage [K][8][B][N] for steps segFirst+1 .. segLast:
  // slot 0 = a (step start), 1..3 = y2, y3, y4, 4..7 = v1..v4
  cplx* seg;
  int segFirst, segLast;
  // adjoint work arrays [B][N] and the gradients
  cplx* GP;
  cplx* GA;
  cplx* TMP;
  const cplx* gout;  // [B][nOut*N]
  cplx* gin;         // [B][N]
};

__device__ __forceinline__ int step_of(const BigParams& bp, int di, int fixedStep) {
  return fixedStep >= 0 ? fixedStep : *bp.base + di;
}

__device__ __forceinline__ bool seg_has(const BigParams& bp, int i) {
  return bp.seg != nullptr && i > bp.segFirst && i <= bp.segLast;
}
__device__ __forceinline__ cplx* seg_slot(const BigParams& bp, int i, int which) {
  return bp.seg + ((size_t)(i - bp.segFirst - 1) * 8 + which) * ((size_t)bp.B * bp.N);
}

// relative intensity of RK stage st of step i (prev, intermediate x2, current)
__device__ __forceinline__ double stage_r(const BigParams& bp, int i, int st) {
  return bp.coef[(size_t)i * 3 + (st == 0 ? 0 : st == 3 ? 2 : 1)].re;
}

template <int P>
__global__ void __launch_bounds__(MIX_THREADS) k_big_mix(BigParams bp, int mode, int di, int fixedStep) {
  const int i = step_of(bp, di, fixedStep);
  if (fixedStep < 0 && i > bp.nSteps) return;
  const long gid = (long)blockIdx.x * MIX_THREADS + threadIdx.x;
  const int M = bp.M;
  if (gid >= ((long)bp.B << bp.LM)) return;
  const int f = (int)(gid >> bp.LM), n2 = (int)(gid & (M - 1));
  const int N = bp.N;
  const size_t fo = (size_t)f * N;
  const int kind = mode & 15, st = mode >> 4;
  cplx v[P];
  if (kind != MIX_FWD_IN && kind != MIX_CKPT_IN && kind != MIX_ADJ_IN && kind != MIX_ADJ_RESUME) {
    // inverse radix-P across the blocks: x[n1 M + n2] = 1/N sum_k1 W_P^(-n1 k1) W_N^(-n2 k1) z_k1[n2]
    if constexpr (P > 1) {
#pragma unroll
      for (int k1 = 0; k1 < P; k1++) v[k1] = bp.Z[fo + (size_t)k1 * M + n2] * conj(bp.twN[n2 * k1]);
      dft_small<P, true>(v);
    } else {
      v[0] = bp.Z[fo + n2];
    }
    const double s = 1.0 / N;
#pragma unroll
    for (int n1 = 0; n1 < P; n1++) v[n1] = v[n1] * s;
  }
  cplx u[P];
  switch (kind) {
    case MIX_FWD_IN:
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) u[n1] = bp.in[(size_t)f * bp.inCols + (size_t)bp.inPos * N + n1 * M + n2];
      break;
    case MIX_RAMAN: {  // IFFT(FFT(r|y|^2) R) * y
      const bool keep = seg_has(bp, i);
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * M + n2;
        if (keep) store_stream(seg_slot(bp, i, 4 + st) + q, v[n1]);
        u[n1] = v[n1] * bp.Y[q];
      }
      break;
    }
    case MIX_STAGE0:
    case MIX_STAGE1:
    case MIX_STAGE2: {
      // k_s = v; acc = ((k1 + 2 k2) + 2 k3); next stage input y = a + (1/2 | 1/2 | 1) k_s
      const int s = kind - MIX_STAGE0;
      const double h = s < 2 ? 0.5 : 1.0;
      const double r = stage_r(bp, i, s + 1);
      const bool keep = seg_has(bp, i);
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * M + n2;
        bp.ACC[q] = s == 0 ? v[n1] : bp.ACC[q] + 2.0 * v[n1];
        const cplx y = bp.A[q] + h * v[n1];
        bp.Y[q] = y;
        if (keep) store_stream(seg_slot(bp, i, 1 + s) + q, y);
        u[n1] = mk(r * abs2(y), 0.0);
      }
      break;
    }
    case MIX_LAST:  // a + (acc + k4) / 6, the input of the dispersion step
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * M + n2;
        u[n1] = bp.A[q] + (bp.ACC[q] + v[n1]) * (1. / 6.);
      }
      break;
    case MIX_DISP:
    case MIX_CKPT_IN: {  // new state a_i; first stage of step i+1 starts with r0 |a|^2
      if (kind == MIX_CKPT_IN) {
#pragma unroll
        for (int n1 = 0; n1 < P; n1++)
          v[n1] = bp.ckpt[((size_t)f * bp.nSeg + i / bp.ckptEvery) * N + n1 * M + n2];
      }
      if (i == bp.nSteps) {
        if (bp.out && bp.outTime) {
#pragma unroll
          for (int n1 = 0; n1 < P; n1++)
            bp.out[(size_t)f * bp.nOut * N + (size_t)bp.outPos * N + n1 * M + n2] = v[n1];
        }
        return;
      }
      if (kind == MIX_DISP && bp.ckptWrite && i % bp.ckptEvery == 0) {
#pragma unroll
        for (int n1 = 0; n1 < P; n1++) store_stream(bp.ckpt + ((size_t)f * bp.nSeg + i / bp.ckptEvery) * N + n1 * M + n2, v[n1]);
      }
      const double r = stage_r(bp, i + 1, 0);
      const bool keep = seg_has(bp, i + 1);
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * M + n2;
        bp.A[q] = v[n1];
        bp.Y[q] = v[n1];
        if (keep) store_stream(seg_slot(bp, i + 1, 0) + q, v[n1]);
        u[n1] = mk(r * abs2(v[n1]), 0.0);
      }
      break;
    }
    case MIX_ADJ_IN:
#pragma unroll
      for (int n1 = 0; n1 < P; n1++)
        u[n1] = bp.gout[(size_t)f * bp.nOut * N + (size_t)bp.outPos * N + n1 * M + n2];
      break;
    case MIX_ADJ_G0:
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * M + n2;
        bp.GP[q] = v[n1];
        bp.GA[q] = v[n1];
        u[n1] = v[n1] * (1. / 6.);
      }
      break;
    case MIX_ADJ_U: {
      const cplx* ys = seg_slot(bp, i, st);
      const cplx* vs = seg_slot(bp, i, 4 + st);
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * M + n2;
        bp.TMP[q] = conj(load_stream(vs + q)) * v[n1];  // last read of v_s
        u[n1] = conj(ys[q]) * v[n1];                     // y_s is read again by MIX_ADJ_S
      }
      break;
    }
    case MIX_ADJ_S: {
      const cplx* ys = seg_slot(bp, i, st);
      const double r2 = 2.0 * stage_r(bp, i, st);
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) {
        const size_t q = fo + n1 * M + n2;
        const cplx gy = bp.TMP[q] + (r2 * v[n1].re) * load_stream(ys + q);  // last read of y_s
        const cplx ga = bp.GA[q] + gy;
        bp.GA[q] = ga;
        if (st == 0) {
          u[n1] = ga;  // gradient of the step's start state
        } else {
          const cplx gp = bp.GP[q];
          // k3 feeds y4 = a + k3, k2 feeds y3 = a + k2/2, k1 feeds y2 = a + k1/2
          u[n1] = st == 3 ? gp * (1. / 3.) + gy : st == 2 ? gp * (1. / 3.) + 0.5 * gy : gp * (1. / 6.) + 0.5 * gy;
        }
      }
      break;
    }
    case MIX_ADJ_OUT:
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) bp.gin[fo + n1 * M + n2] = v[n1];
      return;
    default:  // MIX_ADJ_RESUME
#pragma unroll
      for (int n1 = 0; n1 < P; n1++) u[n1] = bp.GA[fo + n1 * M + n2];
      break;
  }
  // forward radix-P across the blocks, then the W_N^(n2 k1) twiddle
  if constexpr (P > 1) {
    dft_small<P, false>(u);
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) bp.X[fo + (size_t)k1 * M + n2] = u[k1] * bp.twN[n2 * k1];
  } else {
    bp.X[fo + n2] = u[0];
  }
}

template <int P, int LM>
__global__ void __launch_bounds__(FftShape<LM>::T) k_big_conv(BigParams bp, int mode, int di, int fixedStep) {
  using F = FftShape<LM>;
  constexpr int T = F::T, M = F::N;
  __shared__ cplx lds[F::lds_elems];
  const int i = step_of(bp, di, fixedStep);
  if (fixedStep < 0 && i > bp.nSteps) return;
  const int f = blockIdx.x / P, k1 = blockIdx.x % P, t = threadIdx.x;
  const int N = bp.N;
  const size_t xo = (size_t)f * N + (size_t)k1 * M;
  FftTwiddles<LM> tw;
  fft_load_twiddles<LM>(tw, bp.twM, t);
  cplx v[8];
  if (mode & CV_SRC_IN) {
#pragma unroll
    for (int m = 0; m < 8; m++)
      v[m] = bp.in[(size_t)f * bp.inCols + (size_t)bp.inPos * N + k1 + (size_t)P * (t + T * m)];
  } else if (mode & CV_SRC_GOUT) {
#pragma unroll
    for (int m = 0; m < 8; m++)
      v[m] = bp.gout[(size_t)f * bp.nOut * N + (size_t)bp.outPos * N + k1 + (size_t)P * (t + T * m)] * (double)N;
  } else {
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = bp.X[xo + t + T * m];
    fft_field<LM, false>(v, lds, t, tw);
  }
  int table = mode & CV_TABLE;
  const bool last = (mode & CV_DISP_DYN) && i == bp.nSteps;
  if (last) table = 4;  // exp(i dz D) exp(-i dz/2 D): the final half step out
  const cplx* H = bp.Hp + (size_t)table * N + (size_t)k1 * M;
  if (mode & CV_CONJ) {
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = v[m] * conj(H[t + T * m]);
  } else {
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = v[m] * H[t + T * m];
  }
  if (mode & CV_DST_GIN) {
    const double s = 1.0 / N;
#pragma unroll
    for (int m = 0; m < 8; m++) bp.gin[(size_t)f * N + k1 + (size_t)P * (t + T * m)] = v[m] * s;
    return;
  }
  if (last && !bp.outTime) {
    if (bp.out) {
#pragma unroll
      for (int m = 0; m < 8; m++)
        bp.out[(size_t)f * bp.nOut * N + (size_t)bp.outPos * N + k1 + (size_t)P * (t + T * m)] = v[m];
    }
    return;
  }
  fft_field<LM, true>(v, lds, t, tw);  // unscaled; the mix kernel applies 1/N
#pragma unroll
  for (int m = 0; m < 8; m++) bp.Z[xo + t + T * m] = v[m];
}

__global__ void k_big_set_base(int* base, int value, int add) {
  if (blockIdx.x == 0 && threadIdx.x == 0) base[0] = add ? base[0] + value : value;
}

// Hp[tab][k1 M + k2] = H[k1 + P k2]  (tab 4: disp * halfOut)
__global__ void k_big_perm(const cplx* __restrict__ ram, const cplx* __restrict__ flt, const cplx* __restrict__ ds,
                           const cplx* __restrict__ hin, const cplx* __restrict__ hout, int N, int P, int M,
                           cplx* __restrict__ Hp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int k1 = j / M, k2 = j % M, k = k1 + P * k2;
  Hp[j] = ram[k];
  Hp[(size_t)N + j] = flt[k];
  Hp[2 * (size_t)N + j] = ds[k];
  Hp[3 * (size_t)N + j] = hin[k];
  Hp[4 * (size_t)N + j] = ds[k] * hout[k];
}

// Launch shape of one (P, LM) instance.
struct Stager {
  int P = 0, LM = 0;
  void (*mix)(const BigParams&, int, int, int, hipStream_t) = nullptr;
  void (*conv)(const BigParams&, int, int, int, hipStream_t) = nullptr;
};

template <int P>
void launch_mix(const BigParams& bp, int mode, int di, int fixed, hipStream_t s) {
  const dim3 g((unsigned)((((long)bp.B << bp.LM) + MIX_THREADS - 1) / MIX_THREADS));
  hipLaunchKernelGGL(k_big_mix<P>, g, dim3(MIX_THREADS), 0, s, bp, mode, di, fixed);
}
template <int P, int LM>
void launch_conv(const BigParams& bp, int mode, int di, int fixed, hipStream_t s) {
  hipLaunchKernelGGL((k_big_conv<P, LM>), dim3((unsigned)(bp.B * P)), dim3(FftShape<LM>::T), 0, s, bp, mode, di,
                     fixed);
}

Stager pick_stager(int L) {
  Stager st;
  switch (L) {
#define NLM_STAGER1(l)          \
  case l:                       \
    st.P = 1;                   \
    st.LM = l;                  \
    st.mix = &launch_mix<1>;    \
    st.conv = &launch_conv<1, l>; \
    break;
    NLM_STAGER1(5)
    NLM_STAGER1(6)
    NLM_STAGER1(7)
    NLM_STAGER1(8)
    NLM_STAGER1(9)
    NLM_STAGER1(10)
    NLM_STAGER1(11)
    NLM_STAGER1(12)
    NLM_STAGER1(13)
#undef NLM_STAGER1
    case 14:
      st.P = 4;
      st.LM = BIG_LM;
      st.mix = &launch_mix<4>;
      st.conv = &launch_conv<4, BIG_LM>;
      break;
    case 15:
      st.P = 8;
      st.LM = BIG_LM;
      st.mix = &launch_mix<8>;
      st.conv = &launch_conv<8, BIG_LM>;
      break;
    default:
      break;
  }
  return st;
}

// One forward step: di = index within a replayed graph (fixed < 0) or the step itself.
void launch_step(const Stager& sg, const BigParams& bp, int di, int fixed, hipStream_t s) {
  const int mixAfterFilter[4] = {MIX_STAGE0, MIX_STAGE1, MIX_STAGE2, MIX_LAST};
  for (int st = 0; st < 4; st++) {
    sg.conv(bp, CONV_RAMAN, di, fixed, s);
    sg.mix(bp, MIX_RAMAN | (st << 4), di, fixed, s);
    sg.conv(bp, CONV_FILTER, di, fixed, s);
    sg.mix(bp, mixAfterFilter[st], di, fixed, s);
  }
  sg.conv(bp, CONV_DISP, di, fixed, s);
  sg.mix(bp, MIX_DISP, di, fixed, s);
}

void launch_prologue(const Stager& sg, const BigParams& bp, bool inTime, hipStream_t s) {
  if (inTime) {
    sg.mix(bp, MIX_FWD_IN, 0, 0, s);
    sg.conv(bp, CONV_PRO_TIME, 0, 0, s);
  } else {
    sg.conv(bp, CONV_PRO_FREQ, 0, 0, s);
  }
  sg.mix(bp, MIX_DISP, 0, 0, s);
}

// The adjoint of forward step i: X holds fwd(g of a_i) unless i is the last step.
void launch_adjoint_step(const Stager& sg, const BigParams& bp, int i, hipStream_t s) {
  if (i == bp.nSteps) {
    if (bp.outTime) {
      sg.mix(bp, MIX_ADJ_IN, 0, i, s);
      sg.conv(bp, 4 | CV_CONJ, 0, i, s);
    } else {
      sg.conv(bp, 4 | CV_CONJ | CV_SRC_GOUT, 0, i, s);
    }
  } else {
    sg.conv(bp, 2 | CV_CONJ, 0, i, s);
  }
  sg.mix(bp, MIX_ADJ_G0, 0, i, s);
  for (int st = 3; st >= 0; st--) {
    sg.conv(bp, 1 | CV_CONJ, 0, i, s);
    sg.mix(bp, MIX_ADJ_U | (st << 4), 0, i, s);
    sg.conv(bp, 0 | CV_CONJ, 0, i, s);
    sg.mix(bp, MIX_ADJ_S | (st << 4), 0, i, s);
  }
}

int setup_params(nlm_plan* p, const Stager& sg, int rows, BigParams& bp) {
  const int N = p->N;
  const size_t BN = (size_t)rows * N;
  int e;
  if ((e = p->bigState.ensure(5 * BN)) || (e = p->bigH.ensure(5 * (size_t)N)) || (e = p->bigBase.ensure(1))) return e;
  const cplx* twM = nullptr;
  if ((e = get_twiddles(1 << sg.LM, &twM))) return e;
  hipLaunchKernelGGL(k_big_perm, dim3((N + 255) / 256), dim3(256), 0, p->stream, p->raman.p, p->filt.p, p->dstep.p,
                     p->halfIn.p, p->halfOut.p, N, sg.P, 1 << sg.LM, p->bigH.p);
  HIP_TRY(hipGetLastError());
  std::memset(&bp, 0, sizeof bp);
  bp.B = rows;
  bp.nSteps = p->nZ - 1;
  bp.N = N;
  bp.M = 1 << sg.LM;
  bp.LM = sg.LM;
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
  bp.ckptEvery = 1;
  return NLM_OK;
}

}  // namespace

// Steps per captured graph.
static constexpr int BIG_GRAPH_STEPS = 32;

bool nlm_big_supported(int model, int L) { return model == NLM_CHI3GNLSE && (L == 14 || L == 15); }

// Batch propagation of a Chi3GNLSE plan with nFreqs 16384 or 32768 (see above).
// in / out are device arrays with the layout of nlm_batch_dev.
int nlm_big_batch(nlm_plan* p, const cplx* in, int rows, int cols, int inTime, int inPos, int nOut, int outPos,
                  cplx* out) {
  const Stager sg = pick_stager(p->L);
  if (!nlm_big_supported(p->model, p->L) || !sg.mix)
    return set_err(NLM_INVALID_ARGUMENT, "unsupported size for the staged transform");
  if (p->nZ < 2) return set_err(NLM_INVALID_ARGUMENT, "the staged transform needs at least two z steps");
  if (rows == 0) return NLM_OK;
  BigParams bp;
  if (int e = setup_params(p, sg, rows, bp)) return e;
  bp.in = in;
  bp.inCols = cols;
  bp.inPos = inPos;
  bp.out = out;
  bp.nOut = nOut;
  bp.outPos = outPos;
  bp.outTime = inTime;
  launch_prologue(sg, bp, inTime != 0, p->stream);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_big_set_base, dim3(1), dim3(64), 0, p->stream, p->bigBase.p, 1, 0);
  HIP_TRY(hipGetLastError());

  // the z loop: a graph of BIG_GRAPH_STEPS steps, rebuilt when its arguments
  // change; NLM_NO_GRAPH=1 in the environment launches the same kernels
  // directly (for tools that do not follow graph launches)
  const int G = BIG_GRAPH_STEPS;
  const int launches = (p->nZ - 1 + G - 1) / G;
  const char* noGraph = std::getenv("NLM_NO_GRAPH");
  if (noGraph && noGraph[0] == '1') {
    for (int r = 0; r < launches; r++) {
      for (int di = 0; di < G; di++) launch_step(sg, bp, di, -1, p->stream);
      hipLaunchKernelGGL(k_big_set_base, dim3(1), dim3(64), 0, p->stream, p->bigBase.p, G, 1);
    }
    HIP_TRY(hipGetLastError());
    return NLM_OK;
  }
  if (!p->bigExec || p->bigKey[0] != rows || p->bigKey[1] != cols || p->bigKey[2] != inTime ||
      p->bigKey[3] != inPos || p->bigKey[4] != nOut || p->bigKey[5] != outPos || p->bigIn != in ||
      p->bigOut != out || p->bigStatePtr != p->bigState.p) {
    if (p->bigExec) (void)hipGraphExecDestroy(p->bigExec);
    p->bigExec = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    for (int di = 0; di < G; di++) launch_step(sg, bp, di, -1, p->stream);
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
  for (int r = 0; r < launches; r++) HIP_TRY(hipGraphLaunch(p->bigExec, p->stream));
  return NLM_OK;
}

// ---------------------------------------------------------------------------
// differentiable Chi3GNLSE batch (C-ABI)
// ---------------------------------------------------------------------------

static int gnlse_diff_check(nlm_plan* p) {
  if (!p) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  if (p->model != NLM_CHI3GNLSE)
    return set_err(NLM_INVALID_ARGUMENT, "the differentiable batch is implemented for Chi3GNLSE");
  if (!pick_stager(p->L).mix) return set_err(NLM_INVALID_ARGUMENT, "the differentiable batch needs nFreqs a power of two, 32..32768");
  if (p->nZ < 2) return set_err(NLM_INVALID_ARGUMENT, "the differentiable batch needs at least two z steps");
  return NLM_OK;
}

extern "C" int nlm_batch_forward(nlm_plan* p, const double* inputs, int rows, int in_time, int ckpt_every,
                                 double* out) {
  if (int e = gnlse_diff_check(p)) return e;
  if (!inputs || !out || rows <= 0) return set_err(NLM_INVALID_ARGUMENT, "null argument or empty batch");
  const int N = p->N, nSteps = p->nZ - 1;
  int K = ckpt_every;
  if (K <= 0) K = (int)std::ceil(std::sqrt(nSteps / 8.0));
  if (K < 1) K = 1;
  if (K > nSteps) K = nSteps;
  const int nSeg = (nSteps + K - 1) / K;
  const size_t BN = (size_t)rows * N;
  const Stager sg = pick_stager(p->L);
  int e;
  if ((e = p->in.ensure(BN)) || (e = p->out.ensure(BN)) || (e = p->ckpt.ensure((size_t)nSeg * BN))) return e;
  HIP_TRY(hipMemcpyAsync(p->in.p, inputs, BN * sizeof(cplx), hipMemcpyHostToDevice, p->stream));
  BigParams bp;
  if ((e = setup_params(p, sg, rows, bp))) return e;
  bp.in = p->in.p;
  bp.inCols = N;
  bp.inPos = 0;
  bp.out = p->out.p;
  bp.nOut = 1;
  bp.outPos = 0;
  bp.outTime = in_time;
  bp.ckpt = p->ckpt.p;
  bp.ckptEvery = K;
  bp.nSeg = nSeg;
  bp.ckptWrite = 1;
  tic(p, 2);
  launch_prologue(sg, bp, in_time != 0, p->stream);
  for (int i = 1; i <= nSteps; i++) launch_step(sg, bp, 0, i, p->stream);
  HIP_TRY(hipGetLastError());
  tic(p, 3);
  HIP_TRY(hipMemcpyAsync(out, p->out.p, BN * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->timing) collect_timing(p, 2, 3, 1);
  p->ckptEvery = K;
  p->nSeg = nSeg;
  p->gInTime = in_time;
  p->bRows = rows;
  p->adjGen = ++p->gen;
  return NLM_OK;
}

extern "C" int nlm_batch_adjoint(nlm_plan* p, const double* g_out, double* g_in) {
  if (int e = gnlse_diff_check(p)) return e;
  if (!g_out || !g_in) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  if (p->bRows <= 0 || !p->nSeg || !p->ckpt.p)
    return set_err(NLM_RUNTIME_ERROR, "run nlm_batch_forward before the adjoint");
  if (p->adjGen != p->gen)
    return set_err(NLM_RUNTIME_ERROR,
                   "the medium changed after nlm_batch_forward (coefficients or another run); run the forward again "
                   "before the adjoint");
  const int N = p->N, nSteps = p->nZ - 1, rows = p->bRows, K = p->ckptEvery, nSeg = p->nSeg;
  const size_t BN = (size_t)rows * N;
  const Stager sg = pick_stager(p->L);
  int e;
  if ((e = p->gG.ensure(BN)) || (e = p->gtmp.ensure(BN)) || (e = p->adjWork.ensure(3 * BN)) ||
      (e = p->adjSeg.ensure((size_t)8 * K * BN)))
    return e;
  HIP_TRY(hipMemcpyAsync(p->gG.p, g_out, BN * sizeof(cplx), hipMemcpyHostToDevice, p->stream));
  BigParams bp;
  if ((e = setup_params(p, sg, rows, bp))) return e;
  bp.inCols = N;
  bp.nOut = 1;
  bp.outTime = p->gInTime;
  bp.ckpt = p->ckpt.p;
  bp.ckptEvery = K;
  bp.nSeg = nSeg;
  bp.seg = p->adjSeg.p;
  bp.GP = p->adjWork.p;
  bp.GA = bp.GP + BN;
  bp.TMP = bp.GA + BN;
  bp.gout = p->gG.p;
  bp.gin = p->gtmp.p;
  tic(p, 2);
  for (int s = nSeg - 1; s >= 0; s--) {
    const int i0 = s * K, i1 = i0 + K < nSteps ? i0 + K : nSteps;
    // recompute the segment from its checkpoint, keeping the stage data
    bp.segFirst = i0;
    bp.segLast = i1;
    sg.mix(bp, MIX_CKPT_IN, 0, i0, p->stream);
    for (int i = i0 + 1; i <= i1; i++) launch_step(sg, bp, 0, i, p->stream);
    // walk it backwards; GA holds the gradient of a_{i1} between segments
    if (i1 < nSteps) sg.mix(bp, MIX_ADJ_RESUME, 0, i1, p->stream);
    for (int i = i1; i > i0; i--) launch_adjoint_step(sg, bp, i, p->stream);
  }
  // the prologue a_0 = IFFT(halfIn FFT(in)) (time input) or IFFT(halfIn in)
  if (p->gInTime) {
    sg.conv(bp, 3 | CV_CONJ, 0, 0, p->stream);
    sg.mix(bp, MIX_ADJ_OUT, 0, 0, p->stream);
  } else {
    sg.conv(bp, 3 | CV_CONJ | CV_DST_GIN, 0, 0, p->stream);
  }
  HIP_TRY(hipGetLastError());
  tic(p, 3);
  HIP_TRY(hipMemcpyAsync(g_in, p->gtmp.p, BN * sizeof(cplx), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  if (p->timing) collect_timing(p, 2, 3, 1);
  return NLM_OK;
}
