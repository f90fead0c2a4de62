// nlm_prop.hpp -- the persistent split-step propagation kernels (templates).
//
// One workgroup (or one lane group of T = N/8 lanes when N < 512) owns one
// field for its whole propagation: the field stays in registers in the
// stride layout of nlm_fft.hpp, the RK4 nonlinear step runs pointwise on it,
// and the dispersion step is forward FFT -> phase multiply -> inverse FFT with
// LDS exchanges only between FFT stages.  The z loop is persistent inside the
// kernel, so a Green's function (2N independent impulse propagations) or a
// batch is a single launch.
//
// Replaces signalSimulationTemplate, src/_NonlinearMedium.hpp:160-239, with the
// model DiffEqs of nlm_models.hpp.  Each model is instantiated in its own
// translation unit (nlm_prop_inst.hip) so the library builds in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/nlm.h"
#include "nlm_fft.hpp"
#include "nlm_models.hpp"

// ---------------------------------------------------------------------------
// kernel parameters
// ---------------------------------------------------------------------------

enum { KIND_GREENS = 0, KIND_BATCH = 1, KIND_GRID = 2 };

struct KParams {
  int nZ, kind, nFields;
  int inTime, outTime;
  int inCols;          // batch/grid: complex values per input row
  int nIn;             // greens: number of input modes
  int inputs[4];       // greens: input mode list
  int inPos[4];        // batch/grid: input channel feeding mode q (-1: none)
  int outPos[4];       // output slot of mode q (-1: not written)
  int nOut;
  const cplx* in;
  cplx* out;           // [nFields][nOut*N]
  cplx* gridF;         // [nFields][nS][nZ][N] (z-resolved runs)
  cplx* gridT;
  const cplx* tw;
  const cplx* dstep;   // [nS][N]
  const cplx* halfIn;
  const cplx* halfOut;
  const cplx* coef;    // [nZ][3][NC]
  const cplx* cconst;
  const cplx* pump0;   // [2nZ-1][N] time domain
  const cplx* pump1;
  const cplx* raman;
  const cplx* filt;
  cplx* ckpt;          // optional: time-domain field every ckptEvery steps, [nFields][nSeg][NS][N]
  int ckptEvery, nSeg;
  const cplx* tab;     // k_prop_tab: RK4 step propagators [nZ][N][2] (alpha, beta) / N
  // pump rows generated in the kernel (linear pump, one pump): envelope [N]
  // (time domain), pump dispersion D_p [N], dzp and beam-profile factors
  // [2nZ-1] (or nullptr)
  const cplx* env;
  const double* Dp;
  const double* strength;
  double dzp;
};

template <int L>
struct Launch {
  static constexpr int T = FftShape<L>::T;
  static constexpr int FPB = T >= 64 ? 1 : 64 / T;  // fields per workgroup
  static constexpr int THREADS = FPB * T;
};

// Up to N = 512 every kernel that runs the steps below (k_prop, k_prop_tab,
// k_prop_gnlse: Launch; k_prop_multi: one field of T lanes) is one wave per
// workgroup, so their transforms exchange through LDS without barriers.
template <int L>
constexpr bool kWaveFft = FftShape<L>::T <= 64;

// Row i of mode q of the z-resolved grid of `field` ([nFields][nS][nZ][N]).
__device__ __forceinline__ cplx* grid_row(const KParams& kp, cplx* base, int field, int nS, int q, int i, int N) {
  return base + (((size_t)field * nS + q) * kp.nZ + i) * N;
}

// ---------------------------------------------------------------------------
// field I/O helpers (stride layout: lane t holds k = t + T*m)
// ---------------------------------------------------------------------------

template <int L>
__device__ __forceinline__ void mul_table(cplx (&v)[8], const cplx* __restrict__ tab, int t) {
  constexpr int T = FftShape<L>::T;
#pragma unroll
  for (int m = 0; m < 8; m++) v[m] = v[m] * tab[t + T * m];
}

template <int L>
__device__ __forceinline__ void scale_inv(cplx (&v)[8]) {
  constexpr double s = 1.0 / FftShape<L>::N;  // exact: N is a power of two
#pragma unroll
  for (int m = 0; m < 8; m++) v[m] = v[m] * s;
}

template <int L>
__device__ __forceinline__ void store_row(cplx* __restrict__ dst, const cplx (&v)[8], int t, bool active) {
  constexpr int T = FftShape<L>::T;
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; m++) dst[t + T * m] = v[m];
  }
}

template <int L>
__device__ __forceinline__ void load_pump(cplx (&p)[8], const cplx* __restrict__ row, int t) {
  constexpr int T = FftShape<L>::T;
#pragma unroll
  for (int m = 0; m < 8; m++) p[m] = row[t + T * m];
}

// ---------------------------------------------------------------------------
// prologue / epilogue shared by the propagation kernels
// ---------------------------------------------------------------------------

// Input of mode q for this field: reference input handling,
// src/_NonlinearMedium.hpp:164-198 and src/_NonlinearMedium.cpp:329-340.
template <int L, int NS>
__device__ __forceinline__ void load_inputs(cplx (&a)[NS][8], const KParams& kp, int field, int t, bool active) {
  constexpr int N = FftShape<L>::N, T = FftShape<L>::T;
#pragma unroll
  for (int q = 0; q < NS; q++)
#pragma unroll
    for (int m = 0; m < 8; m++) a[q][m] = mk(0.0, 0.0);
  if (!active) return;
  if (kp.kind == KIND_GREENS) {
    // field p = 2*i + w: impulse e_{i mod N} (w = 0) or i*e_{i mod N} (w = 1)
    // into mode inputs[i / N]
    const int i = field >> 1, w = field & 1;
    const int md = kp.inputs[i / N], pos = i % N;
    const cplx val = w ? mk(0.0, 1.0) : mk(1.0, 0.0);
#pragma unroll
    for (int q = 0; q < NS; q++)
      if (q == md)
#pragma unroll
        for (int m = 0; m < 8; m++)
          if (t + T * m == pos) a[q][m] = val;
  } else {
    const cplx* row = kp.in + (size_t)field * kp.inCols;
#pragma unroll
    for (int q = 0; q < NS; q++) {
      const int ch = kp.inPos[q];
      if (ch >= 0)
#pragma unroll
        for (int m = 0; m < 8; m++) a[q][m] = row[(size_t)ch * N + t + T * m];
    }
  }
}

// Half dispersion step in: a <- FFT(a) (time-domain input) * exp(i dz/2 D);
// the grid path records row 0 of the frequency grid.
template <int L, int NS, class TW>
__device__ __forceinline__ void prologue(cplx (&a)[NS][8], const KParams& kp, cplx* buf, int t,
                                         const TW& tw, int field, bool active) {
  constexpr int N = FftShape<L>::N;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    if (kp.inTime) fft_field1<L, false, kWaveFft<L>>(a[q], buf, t, tw);
    mul_table<L>(a[q], kp.halfIn + (size_t)q * N, t);
    if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridF, field, NS, q, 0, N), a[q], t, active);
  }
}

// Final half step out and output write (src/_NonlinearMedium.hpp:235-238,
// src/_NonlinearMedium.cpp:333-345, 428-431).  `a` holds the last frequency row.
template <int L, int NS, class TW>
__device__ __forceinline__ void epilogue(cplx (&a)[NS][8], const KParams& kp, cplx* buf, int t,
                                         const TW& tw, int field, bool active) {
  constexpr int N = FftShape<L>::N;
  const int last = kp.nZ - 1;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    mul_table<L>(a[q], kp.halfOut + (size_t)q * N, t);
    if (kp.kind == KIND_GRID) {
      store_row<L>(grid_row(kp, kp.gridF, field, NS, q, last, N), a[q], t, active);
      fft_field1<L, true, kWaveFft<L>>(a[q], buf, t, tw);
      scale_inv<L>(a[q]);
      store_row<L>(grid_row(kp, kp.gridT, field, NS, q, last, N), a[q], t, active);
    } else {
      const int op = kp.outPos[q];
      if (kp.outTime) {
        fft_field1<L, true, kWaveFft<L>>(a[q], buf, t, tw);
        scale_inv<L>(a[q]);
      }
      if (op >= 0) store_row<L>(kp.out + (size_t)field * kp.nOut * N + (size_t)op * N, a[q], t, active);
    }
  }
}

// ---------------------------------------------------------------------------
// the split-step propagation kernel (pointwise models)
// ---------------------------------------------------------------------------

// Linear pump generated on the fly (_NonlinearMedium::runPumpSimulation,
// src/_NonlinearMedium.cpp:225-250): row r = s_r IFFT(FFT(env) exp(i (r dzp) D_p)),
// row 0 = s_0 env.  The phase advances row by row (one complex product) and is
// re-evaluated exactly every 64 rows, which keeps it within ~1e-14 of the
// direct evaluation.  Rows are produced in order.
template <int L>
struct InlinePump {
  cplx F0[8], ph[8], e1[8];
  int row;
  template <class TW>
  NLM_DEV void init(const KParams& kp, cplx* buf, int t, const TW& tw, cplx (&row0)[8]) {
    constexpr int T = FftShape<L>::T;
    const double s0 = kp.strength ? kp.strength[0] : 1.0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      F0[m] = kp.env[t + T * m];
      row0[m] = s0 * F0[m];
      double sn, cs;
      sincos_reduced(kp.dzp * kp.Dp[t + T * m], &sn, &cs);
      e1[m] = mk(cs, sn);
      ph[m] = mk(1.0, 0.0);
    }
    fft_field1<L, false, kWaveFft<L>>(F0, buf, t, tw);
    row = 0;
  }
  template <class TW>
  NLM_DEV void next(const KParams& kp, cplx* buf, int t, const TW& tw, cplx (&out)[8]) {
    constexpr int T = FftShape<L>::T;
    row++;
    const bool exact = (row & 63) == 0;
    const double zr = (double)row * kp.dzp;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      if (exact) {
        double sn, cs;
        sincos_reduced(zr * kp.Dp[t + T * m], &sn, &cs);
        ph[m] = mk(cs, sn);
      } else {
        ph[m] = ph[m] * e1[m];
      }
      out[m] = F0[m] * ph[m];
    }
    fft_field1<L, true, kWaveFft<L>>(out, buf, t, tw);
    const double s = (kp.strength ? kp.strength[row] : 1.0) * (1.0 / FftShape<L>::N);
#pragma unroll
    for (int m = 0; m < 8; m++) out[m] = out[m] * s;
  }
};

// Read-only data that every lane of a workgroup reads at the same address, in the
// constant address space so that the compiler loads it with scalar loads.
typedef const __attribute__((address_space(4))) cplx uniform_cplx;
__device__ __forceinline__ const uniform_cplx* to_uniform(const cplx* p) { return (const uniform_cplx*)p; }

__device__ __forceinline__ int opaque_lane(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
template <int L>
__device__ __forceinline__ void retarget(TableTwiddles<L>& tw, int t) { tw.t = t; }
template <int L>
__device__ __forceinline__ void retarget(RegTwiddles<L>&, int) {}

// One propagation (the loop of signalSimulationTemplate) of field `field` on
// the T lanes that hold it; INLINE generates the pump rows instead of reading
// the pump grid.
template <class M, int L, bool INLINE = false>
__device__ __forceinline__ void prop_field(const KParams& kp, int field, bool active, int t, cplx* buf,
                                           const cplx* twl) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, NS = M::NS, NP = M::NP, NC = M::NC;
  static_assert(!INLINE || NP == 1, "the generated pump is a single linear pump");
  constexpr bool LEAN = PropTw<L>::LEAN;

  typename PropTw<L>::type tw;
  init_twiddles<L>(tw, kp.tw, twl, t);

  // exp(i dz D_s) at this lane's frequencies, reused every step (read from the
  // cached row at each use from N = 4096)
  constexpr int DS = LEAN ? 1 : 8;
  cplx ds[NS][DS];
  if constexpr (!LEAN) {
#pragma unroll
    for (int q = 0; q < NS; q++)
#pragma unroll
      for (int m = 0; m < 8; m++) ds[q][m] = kp.dstep[(size_t)q * N + t + T * m];
  }

  cplx a[NS][8];
  load_inputs<L, NS>(a, kp, field, t, active);
  prologue<L, NS>(a, kp, buf, t, tw, field, active);
  if (kp.nZ > 1) {
#pragma unroll
    for (int q = 0; q < NS; q++) {
      fft_field1<L, true, kWaveFft<L>>(a[q], buf, t, tw);
      scale_inv<L>(a[q]);
      if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridT, field, NS, q, 0, N), a[q], t, active);
      if (kp.ckpt) store_row<L>(kp.ckpt + ((size_t)field * kp.nSeg * NS + q) * N, a[q], t, active);
    }
  }

  constexpr int NPP = NP > 0 ? NP : 1;
  // pump rows 2i-2 (pp), 2i-1 (pi), 2i (pc) of step i; the next step's rows are
  // loaded before this step's transforms so their latency hides behind them
  cplx pp[NPP][8], pi[NPP][8], pc[NPP][8];
  const cplx* pumps[2] = {kp.pump0, kp.pump1};
  InlinePump<L> gen;
  if constexpr (INLINE) {
    gen.init(kp, buf, t, tw, pp[0]);
    if (kp.nZ > 1) {
      gen.next(kp, buf, t, tw, pi[0]);
      gen.next(kp, buf, t, tw, pc[0]);
    }
  } else {
#pragma unroll
    for (int r = 0; r < NP; r++) {
      load_pump<L>(pp[r], pumps[r], t);
      if (kp.nZ > 1) {
        load_pump<L>(pi[r], pumps[r] + (size_t)N, t);
        load_pump<L>(pc[r], pumps[r] + (size_t)2 * N, t);
      }
    }
  }

  // the medium's constants (Chi3: the nonlinear step), uniform like the coefficients
  constexpr int NCC = M::NCC > 0 ? M::NCC : 1;
  cplx cc[NCC];
  {
    const uniform_cplx* ccq = to_uniform(kp.cconst);
#pragma unroll
    for (int k = 0; k < M::NCC; k++) cc[k] = mk(ccq[k].re, ccq[k].im);
  }
  const int t0 = t;
  for (int i = 1; i < kp.nZ; i++) {
    // The lane index as a value the compiler cannot see through: the LDS,
    // twiddle and memory offsets derived from it are formed inside the step
    // instead of being hoisted out of the z loop, where they took up to ~120
    // registers that spilled (measured on the batch kernels: Chi2SHG at
    // N = 4096 115 -> 78 registers spilled to scratch, N = 1024 / 2048 none
    // instead of 2 / 6; Chi2PDC reading its pump rows at N = 4096 48 -> 17).
    const int t = opaque_lane(t0);
    retarget(tw, t);
    // --- RK4 nonlinear step (DiffEq + update, src/_NonlinearMedium.hpp:207-210)
    // The step's coefficients are the same for every lane: read through the
    // constant address space they are scalar loads into SGPRs, instead of vector
    // loads whose latency each RK stage waited for.
    cplx c[3 * NC > 0 ? 3 * NC : 1];
    {
      const uniform_cplx* cq = to_uniform(kp.coef + (size_t)i * 3 * NC);
#pragma unroll
      for (int k = 0; k < 3 * NC; k++) c[k] = mk(cq[k].re, cq[k].im);
    }
#pragma unroll
    for (int m = 0; m < 8; m++) {
      cplx x[NS], y[NS], k1[NS], k2[NS], k3[NS], k4[NS], P0[NPP], P1[NPP], P2[NPP];
#pragma unroll
      for (int r = 0; r < NPP; r++) {
        P0[r] = NP ? pp[r][m] : mk(0, 0);
        P1[r] = NP ? pi[r][m] : mk(0, 0);
        P2[r] = NP ? pc[r][m] : mk(0, 0);
      }
#pragma unroll
      for (int q = 0; q < NS; q++) x[q] = a[q][m];
      M::template f<0>(c, cc, x, P0, k1);
#pragma unroll
      for (int q = 0; q < NS; q++) y[q] = x[q] + 0.5 * k1[q];
      M::template f<1>(c + NC, cc, y, P1, k2);
#pragma unroll
      for (int q = 0; q < NS; q++) y[q] = x[q] + 0.5 * k2[q];
      M::template f<2>(c + NC, cc, y, P1, k3);
#pragma unroll
      for (int q = 0; q < NS; q++) y[q] = x[q] + k3[q];
      M::template f<3>(c + 2 * NC, cc, y, P2, k4);
#pragma unroll
      for (int q = 0; q < NS; q++) a[q][m] = x[q] + (((k1[q] + 2.0 * k2[q]) + 2.0 * k3[q]) + k4[q]) * (1. / 6.);
    }
#pragma unroll
    for (int r = 0; r < NP; r++) {
#pragma unroll
      for (int m = 0; m < 8; m++) pp[r][m] = pc[r][m];
      // pump grids: the next rows are loaded here so that their latency hides
      // behind the dispersion transforms; generated rows are made after them
      // (below), so that the generator's rows are not held across the transforms
      if (!INLINE && i + 1 < kp.nZ) {
        load_pump<L>(pi[r], pumps[r] + (size_t)(2 * i + 1) * N, t);
        load_pump<L>(pc[r], pumps[r] + (size_t)(2 * i + 2) * N, t);
      }
    }

    // --- dispersion step (src/_NonlinearMedium.hpp:212-215)
#pragma unroll
    for (int q = 0; q < NS; q++) {
      fft_field1<L, false, kWaveFft<L>>(a[q], buf, t, tw);
#pragma unroll
      for (int m = 0; m < 8; m++)
        a[q][m] = a[q][m] * (LEAN ? kp.dstep[(size_t)q * N + t + T * m] : ds[q][LEAN ? 0 : m]);
      if (i < kp.nZ - 1) {
        if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridF, field, NS, q, i, N), a[q], t, active);
        fft_field1<L, true, kWaveFft<L>>(a[q], buf, t, tw);
        scale_inv<L>(a[q]);
        if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridT, field, NS, q, i, N), a[q], t, active);
        if (kp.ckpt && i % kp.ckptEvery == 0)
          store_row<L>(kp.ckpt + (((size_t)field * kp.nSeg + i / kp.ckptEvery) * NS + q) * N, a[q], t, active);
      }
    }
    if constexpr (INLINE) {
      if (i + 1 < kp.nZ) {
        gen.next(kp, buf, t, tw, pi[0]);
        gen.next(kp, buf, t, tw, pc[0]);
      }
    }
  }
  epilogue<L, NS>(a, kp, buf, t, tw, field, active);
}

template <class M, int L>
__global__ void __launch_bounds__(Launch<L>::THREADS) k_prop(KParams kp) {
  using F = FftShape<L>;
  constexpr int T = F::T, FPB = Launch<L>::FPB;
  __shared__ cplx lds[FPB * F::lds_elems];
  __shared__ cplx twl[PropTw<L>::LDS_ELEMS];
  fill_twiddle_lds<L>(twl, kp.tw);
  const int lf = threadIdx.x / T, t = threadIdx.x % T;
  const int field = blockIdx.x * FPB + lf;
  prop_field<M, L>(kp, field, field < kp.nFields, t, lds + lf * F::lds_elems, twl);
}

// ---------------------------------------------------------------------------
// Chi3GNLSE: transforms inside every RK stage (solvers/Chi3GNLSE.cpp:51-92)
// ---------------------------------------------------------------------------

// k = IFFT( FFT( IFFT(FFT(r |y|^2) * R) * y ) * (nl0 + nl1*omega) )
template <int L, class TW>
__device__ __forceinline__ void gnlse_stage(const cplx (&y)[8], double r, cplx (&k)[8], const cplx (&ram)[8],
                                            const cplx (&flt)[8], cplx* buf, int t, const TW& tw) {
#pragma unroll
  for (int m = 0; m < 8; m++) k[m] = mk(r * abs2(y[m]), 0.0);
  fft_field1<L, false, kWaveFft<L>>(k, buf, t, tw);
#pragma unroll
  for (int m = 0; m < 8; m++) k[m] = k[m] * ram[m];
  fft_field1<L, true, kWaveFft<L>>(k, buf, t, tw);
  scale_inv<L>(k);
#pragma unroll
  for (int m = 0; m < 8; m++) k[m] = k[m] * y[m];
  fft_field1<L, false, kWaveFft<L>>(k, buf, t, tw);
#pragma unroll
  for (int m = 0; m < 8; m++) k[m] = k[m] * flt[m];
  fft_field1<L, true, kWaveFft<L>>(k, buf, t, tw);
  scale_inv<L>(k);
}

// One Chi3GNLSE propagation of field `field` on the T lanes that hold it.
template <int L>
__device__ __forceinline__ void gnlse_field(const KParams& kp, int field, bool active, int t, cplx* buf,
                                            const cplx* twl) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T;
  typename PropTw<L>::type tw;
  init_twiddles<L>(tw, kp.tw, twl, t);
  cplx ds[8], ram[8], flt[8];
#pragma unroll
  for (int m = 0; m < 8; m++) {
    ds[m] = kp.dstep[t + T * m];
    ram[m] = kp.raman[t + T * m];
    flt[m] = kp.filt[t + T * m];
  }

  cplx a[1][8];
  load_inputs<L, 1>(a, kp, field, t, active);
  prologue<L, 1>(a, kp, buf, t, tw, field, active);
  if (kp.nZ > 1) {
    fft_field1<L, true, kWaveFft<L>>(a[0], buf, t, tw);
    scale_inv<L>(a[0]);
    if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridT, field, 1, 0, 0, N), a[0], t, active);
  }

  for (int i = 1; i < kp.nZ; i++) {
    const cplx* c = kp.coef + (size_t)i * 3;
    const double r0 = c[0].re, r1 = c[1].re, r2 = c[2].re;
    // acc accumulates ((k1 + 2 k2) + 2 k3) + k4 in the reference's order
    cplx y[8], k[8], acc[8];
    gnlse_stage<L>(a[0], r0, k, ram, flt, buf, t, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) {
      acc[m] = k[m];
      y[m] = a[0][m] + 0.5 * k[m];
    }
    gnlse_stage<L>(y, r1, k, ram, flt, buf, t, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) {
      acc[m] = acc[m] + 2.0 * k[m];
      y[m] = a[0][m] + 0.5 * k[m];
    }
    gnlse_stage<L>(y, r1, k, ram, flt, buf, t, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) {
      acc[m] = acc[m] + 2.0 * k[m];
      y[m] = a[0][m] + k[m];
    }
    gnlse_stage<L>(y, r2, k, ram, flt, buf, t, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) a[0][m] = a[0][m] + (acc[m] + k[m]) * (1. / 6.);

    fft_field1<L, false, kWaveFft<L>>(a[0], buf, t, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) a[0][m] = a[0][m] * ds[m];
    if (i < kp.nZ - 1) {
      if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridF, field, 1, 0, i, N), a[0], t, active);
      fft_field1<L, true, kWaveFft<L>>(a[0], buf, t, tw);
      scale_inv<L>(a[0]);
      if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridT, field, 1, 0, i, N), a[0], t, active);
    }
  }
  epilogue<L, 1>(a, kp, buf, t, tw, field, active);
}

template <int L>
__global__ void __launch_bounds__(Launch<L>::THREADS) k_prop_gnlse(KParams kp) {
  using F = FftShape<L>;
  constexpr int T = F::T, FPB = Launch<L>::FPB;
  __shared__ cplx lds[FPB * F::lds_elems];
  __shared__ cplx twl[PropTw<L>::LDS_ELEMS];
  fill_twiddle_lds<L>(twl, kp.tw);
  const int lf = threadIdx.x / T, t = threadIdx.x % T;
  const int field = blockIdx.x * FPB + lf;
  gnlse_field<L>(kp, field, field < kp.nFields, t, lds + lf * F::lds_elems, twl);
}

typedef void (*prop_fn)(const KParams&, int grid, hipStream_t);

template <class M, int L>
static void launch_prop(const KParams& kp, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_prop<M, L>), dim3(grid), dim3(Launch<L>::THREADS), 0, s, kp);
}
template <int L>
static void launch_gnlse(const KParams& kp, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_prop_gnlse<L>), dim3(grid), dim3(Launch<L>::THREADS), 0, s, kp);
}

// Launcher of model MODEL for log2 N = L (3..13), or nullptr.
prop_fn nlm_pick_prop(int model, int L);
// Step-propagator kernel (one signal mode) for log2 N = L, and its table builder.
prop_fn nlm_pick_prop_tab(int L);
// schedule variant 0..3 of the table kernel (see FftShape); nullptr where a
// variant does not exist or equals the default
prop_fn nlm_pick_prop_tab_variant(int L, int variant);
bool nlm_launch_step_table(int model, const cplx* coef, const cplx* cconst, const cplx* pump, int N, int nZ,
                           cplx* tab, hipStream_t s);
