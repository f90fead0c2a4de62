// nlm_prop_tab.hip -- split step with a precomputed RK4 step propagator for the
// single-signal-mode undepleted media (Chi3, Chi2PDC).
//
// With the pump grid fixed, the DiffEq of these media is R-linear in the
// signal: Chi2PDC (solvers/Chi2PDC.cpp:27-47) has k = (c P) conj(x), Chi3
// (solvers/Chi3.cpp:57-69) has k = i nl (2|P|^2 x + P^2 conj(x)).  Every RK4
// stage (src/_NonlinearMedium.hpp:207-210) is then R-linear too, and one
// whole RK4 step at time sample k of step i is
//     x -> alpha[i][k] x + beta[i][k] conj(x),
// with alpha = (R(1) - i R(i)) / 2 and beta = (R(1) + i R(i)) / 2 for R the
// RK4 step applied to the inputs 1 and i; as a real 2 x 2 matrix acting on
// (Re x, Im x) its columns are R(1) and R(i).  k_step_table evaluates R with the
// model's own DiffEq functor once per (step, sample) for all fields and stores
// the matrix; the propagation kernel then spends 4 FP64 operations per point on
// the nonlinear step instead of four DiffEq evaluations, for every one of the 2N
// Green's impulses or batch rows.  The propagator also absorbs the inverse
// transform's 1/N (an exact power of two), so the field between steps is kept
// unscaled.
#include "nlm_prop.hpp"

template <class M>
__device__ __forceinline__ cplx rk4_step(const cplx* c, const cplx* cc, cplx x, cplx P0, cplx P1, cplx P2) {
  constexpr int NC = M::NC;
  cplx xs[1] = {x}, y[1], k1[1], k2[1], k3[1], k4[1], p0[1] = {P0}, p1[1] = {P1}, p2[1] = {P2};
  M::template f<0>(c, cc, xs, p0, k1);
  y[0] = x + 0.5 * k1[0];
  M::template f<1>(c + NC, cc, y, p1, k2);
  y[0] = x + 0.5 * k2[0];
  M::template f<2>(c + NC, cc, y, p1, k3);
  y[0] = x + k3[0];
  M::template f<3>(c + 2 * NC, cc, y, p2, k4);
  return x + (((k1[0] + 2.0 * k2[0]) + 2.0 * k3[0]) + k4[0]) * (1. / 6.);
}

// tab[(i*N + k)*2 + {0,1}] = {(Re R(1), Re R(i)), (Im R(1), Im R(i))} * scale for
// steps i = 1..nZ-1: Re x' = p.re Re x + p.im Im x, Im x' = q.re Re x + q.im Im x.
template <class M>
__global__ void k_step_table(const cplx* __restrict__ coef, const cplx* __restrict__ cconst,
                             const cplx* __restrict__ pump, int N, int nZ, double scale, cplx* __restrict__ tab) {
  static_assert(M::NS == 1 && M::NP == 1, "single signal mode, single pump");
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)(nZ - 1) * N;
  if (idx >= total) return;
  const int i = 1 + (int)(idx / N), k = (int)(idx % N);
  const cplx* c = coef + (size_t)i * 3 * M::NC;
  const cplx P0 = pump[(size_t)(2 * i - 2) * N + k];
  const cplx P1 = pump[(size_t)(2 * i - 1) * N + k];
  const cplx P2 = pump[(size_t)(2 * i) * N + k];
  const cplx r1 = rk4_step<M>(c, cconst, mk(1.0, 0.0), P0, P1, P2);
  const cplx ri = rk4_step<M>(c, cconst, mk(0.0, 1.0), P0, P1, P2);
  tab[((size_t)i * N + k) * 2] = mk(scale * r1.re, scale * ri.re);
  tab[((size_t)i * N + k) * 2 + 1] = mk(scale * r1.im, scale * ri.im);
}

template <int L>
__device__ __forceinline__ void load_prop_row(cplx (&al)[8], cplx (&be)[8], const cplx* __restrict__ row, int t) {
  constexpr int T = FftShape<L>::T;
#pragma unroll
  for (int m = 0; m < 8; m++) {
    al[m] = row[(size_t)(t + T * m) * 2];
    be[m] = row[(size_t)(t + T * m) * 2 + 1];
  }
}

template <int L>
__device__ __forceinline__ void store_scaled(cplx* __restrict__ dst, const cplx (&v)[8], int t, bool active) {
  constexpr int T = FftShape<L>::T;
  constexpr double s = 1.0 / FftShape<L>::N;
  if (active) {
#pragma unroll
    for (int m = 0; m < 8; m++) dst[t + T * m] = v[m] * s;
  }
}

// Same contract as k_prop (Green's, batch and grid kinds, checkpoints) for a
// medium with one signal mode and a step propagator table kp.tab.
template <int L>
__global__ void __launch_bounds__(Launch<L>::THREADS) k_prop_tab(KParams kp) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, FPB = Launch<L>::FPB;
  __shared__ cplx lds[FPB * F::lds_elems];
  __shared__ cplx twl[PropTw<L>::LDS_ELEMS];
  fill_twiddle_lds<L>(twl, kp.tw);

  const int lf = threadIdx.x / T, t = threadIdx.x % T;
  const int field = blockIdx.x * FPB + lf;
  const bool active = field < kp.nFields;
  cplx* buf = lds + lf * F::lds_elems;

  typename PropTw<L>::type tw;
  init_twiddles<L>(tw, kp.tw, twl, t);
  // Where the dispersion row and the next step's propagator rows live (measured
  // per size on Green's launches).  N = 512 (one wave per field and SIMD): the
  // dispersion row is read from L1/L2 where it is used and the propagator rows are
  // fetched after the forward transform, so neither is held in registers across a
  // whole step; held across the step they spilled to AGPRs (92 register moves per
  // step, 3 % of the step time).  N = 1024: the late fetch alone is 7 % faster,
  // the dispersion row read at use too is slower.  N = 2048 is fastest with both
  // held; from N = 4096 the dispersion row is read at use (PropTw::LEAN).
  constexpr int LOG = FftShape<L>::LOG;
  constexpr bool LATE = LOG == 9 || LOG == 10;
  constexpr bool DS_MEM = LOG == 9 || PropTw<L>::LEAN;
  cplx ds[DS_MEM ? 1 : 8];
  if constexpr (!DS_MEM) {
#pragma unroll
    for (int m = 0; m < 8; m++) ds[m] = kp.dstep[t + T * m];
  }

  cplx a[1][8];
  load_inputs<L, 1>(a, kp, field, t, active);
  prologue<L, 1>(a, kp, buf, t, tw, field, active);
  cplx al[8], be[8];
  if (kp.nZ > 1) {
    fft_field1<L, true, kWaveFft<L>>(a[0], buf, t, tw);  // unscaled: the table holds the 1/N
    if (kp.kind == KIND_GRID) store_scaled<L>(grid_row(kp, kp.gridT, field, 1, 0, 0, N), a[0], t, active);
    if (kp.ckpt) store_scaled<L>(kp.ckpt + (size_t)field * kp.nSeg * N, a[0], t, active);
    load_prop_row<L>(al, be, kp.tab + (size_t)N * 2, t);
  }

#ifdef NLM_TAB_SMALL
  // two steps per iteration let the scheduler overlap one step's last exchange
  // and stores with the next step's first loads (N = 512: 1.88 against 1.97 ms per
  // launch, same box; four steps 2.01 ms; from N = 1024 no change)
#pragma unroll 2
#endif
  for (int i = 1; i < kp.nZ; i++) {
    // --- RK4 nonlinear step as its R-linear propagator (a real 2 x 2 matrix)
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const cplx x = a[0][m];
      a[0][m] = mk(al[m].re * x.re + al[m].im * x.im, be[m].re * x.re + be[m].im * x.im);
    }
    if (!LATE && i + 1 < kp.nZ) load_prop_row<L>(al, be, kp.tab + (size_t)(i + 1) * N * 2, t);

    // --- dispersion step (src/_NonlinearMedium.hpp:212-215)
    fft_field1<L, false, kWaveFft<L>>(a[0], buf, t, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) a[0][m] = a[0][m] * (DS_MEM ? kp.dstep[t + T * m] : ds[DS_MEM ? 0 : m]);
    if (LATE && i + 1 < kp.nZ) load_prop_row<L>(al, be, kp.tab + (size_t)(i + 1) * N * 2, t);
    if (i < kp.nZ - 1) {
      if (kp.kind == KIND_GRID) store_row<L>(grid_row(kp, kp.gridF, field, 1, 0, i, N), a[0], t, active);
      fft_field1<L, true, kWaveFft<L>>(a[0], buf, t, tw);
      if (kp.kind == KIND_GRID) store_scaled<L>(grid_row(kp, kp.gridT, field, 1, 0, i, N), a[0], t, active);
      if (kp.ckpt && i % kp.ckptEvery == 0)
        store_scaled<L>(kp.ckpt + ((size_t)field * kp.nSeg + i / kp.ckptEvery) * N, a[0], t, active);
    }
  }
  epilogue<L, 1>(a, kp, buf, t, tw, field, active);
}

template <int L>
static void launch_prop_tab(const KParams& kp, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_prop_tab<L>), dim3(grid), dim3(Launch<L>::THREADS), 0, s, kp);
}

// This file is compiled once per schedule variant NLM_TAB_VARIANT (see FftShape);
// variants other than 0 exist for nFreqs 512..8192, where the autotuner chooses.
#ifndef NLM_TAB_VARIANT
#define NLM_TAB_VARIANT 0
#endif
#define NLM_CAT2(a, b) a##b
#define NLM_CAT(a, b) NLM_CAT2(a, b)

#ifdef NLM_TAB_SMALL
// The single-wave sizes (nFreqs 8..512, schedule 0) are compiled in a translation
// unit of their own, with LLVM's memory-clause machine scheduler: at one wave per
// SIMD nothing hides the row loads and LDS exchanges, and its order was 4 % faster
// at N = 512 (1.95 against 2.03 ms per Green's launch, same box), while from
// N = 1024 (two or more waves per SIMD) it was 1-5 % slower.
prop_fn nlm_pick_prop_tab_small(int L) {
  switch (L) {
    case 3: return &launch_prop_tab<3>;
    case 4: return &launch_prop_tab<4>;
    case 5: return &launch_prop_tab<5>;
    case 6: return &launch_prop_tab<6>;
    case 7: return &launch_prop_tab<7>;
    case 8: return &launch_prop_tab<8>;
    case 9: return &launch_prop_tab<9>;
  }
  return nullptr;
}
#else
prop_fn nlm_pick_prop_tab_small(int L);

prop_fn NLM_CAT(nlm_pick_prop_tab_var, NLM_TAB_VARIANT)(int L) {
  constexpr int V = NLM_TAB_VARIANT, C = 32 * V;
  // bit 1 (small radix last) only differs from the default when log2 N is not a multiple of 3
  if (V != 0 && (L < 9 || L > 13 || ((V & 2) && L % 3 == 0))) return nullptr;
  switch (L) {
#if NLM_TAB_VARIANT == 0
    case 3:
    case 4:
    case 5:
    case 6:
    case 7:
    case 8:
    case 9: return nlm_pick_prop_tab_small(L);
    case 12: return &launch_prop_tab<12>;
#elif (NLM_TAB_VARIANT & 2) == 0
    case 9: return &launch_prop_tab<9 + C>;
    case 12: return &launch_prop_tab<12 + C>;
#endif
    case 10: return &launch_prop_tab<10 + C>;
    case 11: return &launch_prop_tab<11 + C>;
    case 13: return &launch_prop_tab<13 + C>;
  }
  return nullptr;
}

#if NLM_TAB_VARIANT == 0
prop_fn nlm_pick_prop_tab_var1(int L);
prop_fn nlm_pick_prop_tab_var2(int L);
prop_fn nlm_pick_prop_tab_var3(int L);

prop_fn nlm_pick_prop_tab(int L) { return nlm_pick_prop_tab_var0(L); }

prop_fn nlm_pick_prop_tab_variant(int L, int variant) {
  switch (variant) {
    case 0: return nlm_pick_prop_tab_var0(L);
    case 1: return nlm_pick_prop_tab_var1(L);
    case 2: return nlm_pick_prop_tab_var2(L);
    case 3: return nlm_pick_prop_tab_var3(L);
  }
  return nullptr;
}

// Build the step propagator table of `model` (NLM_CHI3 or NLM_CHI2PDC) into tab
// ([nZ][N][2] complex; row 0 unused).  Returns false for other models.
bool nlm_launch_step_table(int model, const cplx* coef, const cplx* cconst, const cplx* pump, int N, int nZ,
                           cplx* tab, hipStream_t s) {
  const long total = (long)(nZ - 1) * N;
  if (total <= 0) return true;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  const double scale = 1.0 / N;
  if (model == NLM_CHI2PDC)
    hipLaunchKernelGGL(k_step_table<MChi2PDC>, grid, block, 0, s, coef, cconst, pump, N, nZ, scale, tab);
  else if (model == NLM_CHI3)
    hipLaunchKernelGGL(k_step_table<MChi3>, grid, block, 0, s, coef, cconst, pump, N, nZ, scale, tab);
  else
    return false;
  return true;
}
#endif  // NLM_TAB_VARIANT == 0
#endif  // NLM_TAB_SMALL
