// nlm_adjoint.hpp -- reverse-mode (adjoint) split step for the Green's function.
//
// Gradients follow one convention throughout: for a real loss L and a complex
// value z, g_z = dL/dRe(z) + i dL/dIm(z) (the convention of torch's complex
// autograd).  With it, a complex-linear map b = m a has g_a = conj(m) g_b, a
// conjugate-linear map b = m conj(a) has g_a = m conj(g_b), the forward DFT
// (b = F a) has g_a = N * IFFT(g_b) and the scaled inverse DFT has
// g_a = FFT(g_b) / N.
//
// The forward Green's launch (k_prop with kp.ckpt set) stores the time-domain
// field every K steps.  For each segment, in reverse order, k_adjoint
// recomputes the segment's states from its checkpoint (the same code path as
// the forward pass, so the states are bit-identical), then sweeps the segment
// backwards: adjoint of the dispersion step (forward FFT, conj(exp(i dz D)),
// inverse FFT), then the transposed RK4 Jacobian of the model's DiffEq.  The
// per-step coefficient gradients are reduced within each wave (deterministic
// shuffles) and written as per-wave partials; the pump-row gradients are summed
// over the workgroup's fields in LDS and written as per-workgroup partials;
// both are summed by separate passes in a fixed order (no atomics).
//
// Forward map replaced: signalSimulationTemplate (src/_NonlinearMedium.hpp:160-239)
// with Chi2PDC::DiffEq (solvers/Chi2PDC.cpp:27-47); assembly of
// computeGreensFunction (src/_NonlinearMedium.cpp:333-372); linear pump
// (src/_NonlinearMedium.cpp:225-250).
#pragma once
#include "nlm_prop.hpp"

struct AParams {
  int nZ, nFields, K, nSeg;
  int inTime, outTime, nOut;
  int outPos[4];
  const cplx* ckpt;    // [nFields][nSeg][NS][N]
  cplx* seg;           // [nFields][K][NS][N]
  const cplx* gout;    // [nFields][nOut*N], domain = outTime
  cplx* gin;           // [nFields][NS*N] gradient w.r.t. the input (input domain) or nullptr
  cplx* gcoef;         // [nWaves][nZ][3][NC]
  cplx* gpump;         // [nWG][nZP][N] or nullptr
  const cplx* tw;
  const cplx* dstep;
  const cplx* halfIn;
  const cplx* halfOut;
  const cplx* coef;
  const cplx* cconst;
  const cplx* pump0;
};

// Models with an adjoint: fadj<ST>(c, cc, y, P, gk, gy, gm) returns the
// gradient w.r.t. the stage input y (gy) and w.r.t. m = c*P (gm) of
// k = f<ST>(c, y, P) for an upstream gradient gk.
template <class M>
struct Adj {
  static constexpr bool ok = false;
};

// Chi2PDC: k = (c P) conj(y)  ->  g_y = (c P) conj(g_k),  g_m = y g_k
template <>
struct Adj<MChi2PDC> {
  static constexpr bool ok = true;
  template <int ST>
  NLM_DEV static void fadj(const cplx& c, const cplx& y, const cplx& P, const cplx& gk, cplx& gy, cplx& gm) {
    const cplx m = c * P;
    gy = m * conj(gk);
    gm = y * gk;
  }
};

template <int L>
struct AdjLaunch {
  static constexpr int T = FftShape<L>::T;
  // 256 threads (one wave per SIMD) leave each wave the full register file:
  // the adjoint holds the field, its gradient and three pump-row gradients
  static constexpr int FA = T >= 256 ? 1 : 256 / T;  // fields per workgroup
  static constexpr int THREADS = FA * T;
  static constexpr int WAVES = (THREADS + 63) / 64;
};

// wave-wide sum (xor butterfly; every lane ends with the total)
NLM_DEV double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// one full forward split step i on the time-domain field (same code as k_prop)
template <class M, int L>
NLM_DEV void fwd_step(cplx (&a)[M::NS][8], int i, const AParams& ap, cplx* buf, int t, const FftTwiddles<L>& tw,
                      const cplx (&ds)[M::NS][8]) {
  constexpr int N = FftShape<L>::N, T = FftShape<L>::T, NS = M::NS, NC = M::NC;
  const cplx* c = ap.coef + (size_t)i * 3 * NC;
#pragma unroll
  for (int m = 0; m < 8; m++) {
    cplx x[NS], y[NS], k1[NS], k2[NS], k3[NS], k4[NS], P0[1], P1[1], P2[1];
    P0[0] = ap.pump0[(size_t)(2 * i - 2) * N + t + T * m];
    P1[0] = ap.pump0[(size_t)(2 * i - 1) * N + t + T * m];
    P2[0] = ap.pump0[(size_t)(2 * i) * N + t + T * m];
#pragma unroll
    for (int q = 0; q < NS; q++) x[q] = a[q][m];
    M::template f<0>(c, ap.cconst, x, P0, k1);
#pragma unroll
    for (int q = 0; q < NS; q++) y[q] = x[q] + 0.5 * k1[q];
    M::template f<1>(c + NC, ap.cconst, y, P1, k2);
#pragma unroll
    for (int q = 0; q < NS; q++) y[q] = x[q] + 0.5 * k2[q];
    M::template f<2>(c + NC, ap.cconst, y, P1, k3);
#pragma unroll
    for (int q = 0; q < NS; q++) y[q] = x[q] + k3[q];
    M::template f<3>(c + 2 * NC, ap.cconst, y, P2, k4);
#pragma unroll
    for (int q = 0; q < NS; q++) a[q][m] = x[q] + (((k1[q] + 2.0 * k2[q]) + 2.0 * k3[q]) + k4[q]) * (1. / 6.);
  }
#pragma unroll
  for (int q = 0; q < NS; q++) {
    fft_field<L, false>(a[q], buf, t, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) a[q][m] = a[q][m] * ds[q][m];
    fft_field<L, true>(a[q], buf, t, tw);
    scale_inv<L>(a[q]);
  }
}

// Sum a pump-row gradient over the workgroup's fields (through the fields' LDS
// buffers) and write it to this workgroup's partial row.
template <int L>
NLM_DEV void emit_pump_row(const cplx (&v)[8], cplx* lds, int lf, int t, cplx* dst) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, FA = AdjLaunch<L>::FA, TH = AdjLaunch<L>::THREADS;
  lds_barrier();
#pragma unroll
  for (int m = 0; m < 8; m++) lds[lf * F::lds_elems + lpad(t + T * m)] = v[m];
  lds_barrier();
  const int tau = lf * T + t;
  for (int k = tau; k < N; k += TH) {
    cplx s = lds[lpad(k)];
#pragma unroll
    for (int f = 1; f < FA; f++) s = s + lds[f * F::lds_elems + lpad(k)];
    dst[k] = s;
  }
}

template <class M, int L>
__global__ void __launch_bounds__(AdjLaunch<L>::THREADS) k_adjoint(AParams ap) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, NS = M::NS, NC = M::NC;
  constexpr int FA = AdjLaunch<L>::FA;
  static_assert(NS == 1 && M::NP == 1, "adjoint implemented for single-mode single-pump models");
  __shared__ cplx lds[FA * F::lds_elems];
  const int lf = threadIdx.x / T, t = threadIdx.x % T;
  const int field = blockIdx.x * FA + lf;
  const bool active = field < ap.nFields;
  const int fld = active ? field : 0;
  cplx* buf = lds + lf * F::lds_elems;
  const int wave = (blockIdx.x * AdjLaunch<L>::THREADS + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int nZP = 2 * ap.nZ - 1;

  FftTwiddles<L> tw;
  fft_load_twiddles<L>(tw, ap.tw, t);
  cplx ds[NS][8];
#pragma unroll
  for (int m = 0; m < 8; m++) ds[0][m] = ap.dstep[t + T * m];

  // gradient of the last frequency row: output = F_last * halfOut (freq) or its IFFT (time)
  cplx g[8];
  const int op = ap.outPos[0];
#pragma unroll
  for (int m = 0; m < 8; m++)
    g[m] = (active && op >= 0) ? ap.gout[(size_t)fld * ap.nOut * N + (size_t)op * N + t + T * m] : mk(0, 0);
  if (ap.outTime) {  // output = IFFT(F) -> g_F = FFT(g) / N
    fft_field<L, false>(g, buf, t, tw);
    scale_inv<L>(g);
  }
#pragma unroll
  for (int m = 0; m < 8; m++) g[m] = conj(ap.halfOut[t + T * m]) * g[m];
  bool gFreq = true;  // g currently holds a frequency-domain gradient

  cplx carry[8];  // pump row 2i gradient from step i+1 (its k1 stage)
#pragma unroll
  for (int m = 0; m < 8; m++) carry[m] = mk(0, 0);
  cplx* gpumpWG = ap.gpump ? ap.gpump + (size_t)blockIdx.x * nZP * N : nullptr;

  cplx* segF = ap.seg + (size_t)fld * ap.K * N;
  for (int s = ap.nSeg - 1; s >= 0; s--) {
    const int i0 = s * ap.K;
    const int i1 = min(i0 + ap.K, ap.nZ - 1);
    cplx a[NS][8];
#pragma unroll
    for (int m = 0; m < 8; m++) a[0][m] = ap.ckpt[((size_t)fld * ap.nSeg + s) * N + t + T * m];
    for (int i = i0 + 1; i <= i1; i++) {
      if (active) {
#pragma unroll
        for (int m = 0; m < 8; m++) segF[(size_t)(i - i0 - 1) * N + t + T * m] = a[0][m];
      }
      if (i < i1) fwd_step<M, L>(a, i, ap, buf, t, tw, ds);
    }
    for (int i = i1; i > i0; i--) {
      // adjoint of the dispersion step: F = FFT(r) exp(i dz D); a_i = IFFT(F) (i < nZ-1)
      if (!gFreq) {
        fft_field<L, false>(g, buf, t, tw);
        scale_inv<L>(g);
      }
#pragma unroll
      for (int m = 0; m < 8; m++) g[m] = conj(ds[0][m]) * g[m];
      fft_field<L, true>(g, buf, t, tw);  // g_r = N * IFFT(.) = unscaled inverse
      gFreq = false;

      // adjoint of the RK4 step with the stored state a_{i-1}
      const cplx* c = ap.coef + (size_t)i * 3 * NC;
      const cplx c0 = c[0], c1 = c[NC], c2 = c[2 * NC];
      cplx gc0 = mk(0, 0), gc1 = mk(0, 0), gc2 = mk(0, 0);
      cplx gP0[8], gP1[8], gP2[8];
#pragma unroll
      for (int m = 0; m < 8; m++) {
        const int k = t + T * m;
        const cplx x = active ? segF[(size_t)(i - i0 - 1) * N + k] : mk(0, 0);
        const cplx P0 = ap.pump0[(size_t)(2 * i - 2) * N + k];
        const cplx P1 = ap.pump0[(size_t)(2 * i - 1) * N + k];
        const cplx P2 = ap.pump0[(size_t)(2 * i) * N + k];
        cplx xs[1] = {x}, ys[1], k1[1], k2[1], k3[1], Pa[1] = {P0}, Pb[1] = {P1};
        M::template f<0>(c, ap.cconst, xs, Pa, k1);
        const cplx y2 = x + 0.5 * k1[0];
        ys[0] = y2;
        M::template f<1>(c + NC, ap.cconst, ys, Pb, k2);
        const cplx y3 = x + 0.5 * k2[0];
        ys[0] = y3;
        M::template f<2>(c + NC, ap.cconst, ys, Pb, k3);
        const cplx y4 = x + k3[0];
        const cplx gr = g[m];
        cplx gx = gr, gk1 = gr * (1. / 6.), gk2 = gr * (1. / 3.), gk3 = gr * (1. / 3.), gk4 = gr * (1. / 6.);
        cplx gy, gm;
        Adj<M>::template fadj<3>(c2, y4, P2, gk4, gy, gm);
        gx = gx + gy;
        gk3 = gk3 + gy;
        gc2 = gc2 + conj(P2) * gm;
        gP2[m] = conj(c2) * gm;
        Adj<M>::template fadj<2>(c1, y3, P1, gk3, gy, gm);
        gx = gx + gy;
        gk2 = gk2 + 0.5 * gy;
        gc1 = gc1 + conj(P1) * gm;
        gP1[m] = conj(c1) * gm;
        Adj<M>::template fadj<1>(c1, y2, P1, gk2, gy, gm);
        gx = gx + gy;
        gk1 = gk1 + 0.5 * gy;
        gc1 = gc1 + conj(P1) * gm;
        gP1[m] = gP1[m] + conj(c1) * gm;
        Adj<M>::template fadj<0>(c0, x, P0, gk1, gy, gm);
        gx = gx + gy;
        gc0 = gc0 + conj(P0) * gm;
        gP0[m] = conj(c0) * gm;
        g[m] = active ? gx : mk(0, 0);
        if (!active) {
          gP0[m] = gP1[m] = gP2[m] = mk(0, 0);
        }
      }
      if (!active) gc0 = gc1 = gc2 = mk(0, 0);
      // per-wave partial of the coefficient gradients of step i
      const double r0 = wave_sum(gc0.re), i0r = wave_sum(gc0.im), r1 = wave_sum(gc1.re), i1r = wave_sum(gc1.im),
                   r2 = wave_sum(gc2.re), i2r = wave_sum(gc2.im);
      if (lane == 0) {
        cplx* dst = ap.gcoef + ((size_t)wave * ap.nZ + i) * 3 * NC;
        dst[0] = mk(r0, i0r);
        dst[NC] = mk(r1, i1r);
        dst[2 * NC] = mk(r2, i2r);
      }
      // pump rows: 2i (k4 of step i + k1 of step i+1), 2i-1 (k2, k3 of step i)
      if (gpumpWG) {
#pragma unroll
        for (int m = 0; m < 8; m++) gP2[m] = gP2[m] + carry[m];
        emit_pump_row<L>(gP2, lds, lf, t, gpumpWG + (size_t)(2 * i) * N);
        emit_pump_row<L>(gP1, lds, lf, t, gpumpWG + (size_t)(2 * i - 1) * N);
#pragma unroll
        for (int m = 0; m < 8; m++) carry[m] = gP0[m];
      }
    }
  }
  if (gpumpWG) emit_pump_row<L>(carry, lds, lf, t, gpumpWG);  // row 0

  // adjoint of the prologue: a_0 = IFFT(F0), F0 = in * halfIn (in = FFT(x) for time input)
  if (!gFreq) {
    fft_field<L, false>(g, buf, t, tw);
    scale_inv<L>(g);
  }
#pragma unroll
  for (int m = 0; m < 8; m++) g[m] = conj(ap.halfIn[t + T * m]) * g[m];
  if (ap.inTime) fft_field<L, true>(g, buf, t, tw);
  if (ap.gin && active) {
#pragma unroll
    for (int m = 0; m < 8; m++) ap.gin[(size_t)fld * N + t + T * m] = g[m];
  }
}

// Adjoint of the Green's assembly: gradients of the final rows G(e_j), G(i e_j)
// from the gradients of C and S (inverse of k_greens_assemble's index map).
__global__ void k_greens_assemble_adjoint(const cplx* __restrict__ gC, const cplx* __restrict__ gS, cplx* __restrict__ gG,
                                          int N, int nIn, int nOut, int nb, const cplx* __restrict__ scale) {
  const long rows = (long)nOut * N, cols = (long)nIn * N;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const long R = idx / cols, Cc = idx % cols;
  const int om = (int)(R / N), r = (int)(R % N), im = (int)(Cc / N), c = (int)(Cc % N);
  const int h = N / 2;
  long si, sc;
  bool doScale = false;
  if (im < nb) {
    si = (long)im * N + (c + h) % N;
    sc = (long)om * N + (r + h) % N;
    doScale = scale != nullptr && im != om;
  } else {
    si = (long)im * N + c;
    sc = (long)om * N + r;
  }
  cplx gc = gC[idx], gs = gS[idx];
  if (doScale) {
    const cplx s = conj(scale[im * nOut + om]);
    gc = s * gc;
    gs = s * gs;
  }
  const long W = (long)nOut * N;
  gG[(2 * si) * W + sc] = 0.5 * (gc + gs);
  gG[(2 * si + 1) * W + sc] = mk(0.0, 0.5) * (gc - gs);
}

// Sum `parts` partial arrays of length n (fixed order).
__global__ void k_sum_partials(const cplx* __restrict__ part, long n, int parts, cplx* __restrict__ out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  cplx s = mk(0, 0);
  for (int p = 0; p < parts; p++) s = s + part[(size_t)p * n + k];
  out[k] = s;
}

// Adjoint of the linear pump grid for one pump: from the time-row gradients
// gT[row] (already summed over the Green's fields), accumulate
//   g_F0 += s_row conj(exp(i row dzp D)) FFT(gT[row]) / N   (rows >= 1)
// per lane group over a strided set of rows; writes [groups][N] partials.
template <int L>
__global__ void __launch_bounds__(Launch<L>::THREADS)
    k_pump_linear_adjoint(const cplx* __restrict__ gT, const double* __restrict__ Dp, double dzp,
                          const double* __restrict__ strength, int nZP, cplx* __restrict__ partial,
                          const cplx* __restrict__ twt) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, FPB = Launch<L>::FPB;
  __shared__ cplx lds[FPB * F::lds_elems];
  const int lf = threadIdx.x / T, t = threadIdx.x % T;
  const int group = blockIdx.x * FPB + lf;
  const int groups = gridDim.x * FPB;
  cplx* buf = lds + lf * F::lds_elems;
  FftTwiddles<L> tw;
  fft_load_twiddles<L>(tw, twt, t);
  cplx acc[8];
#pragma unroll
  for (int m = 0; m < 8; m++) acc[m] = mk(0, 0);
  // every lane group runs the same number of iterations (uniform barriers)
  const int iters = (nZP - 1 + groups - 1) / groups;
  for (int it = 0; it < iters; it++) {
    const int row = 1 + group + it * groups;
    const bool ok = row < nZP;
    cplx v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = ok ? gT[(size_t)row * N + t + T * m] : mk(0, 0);
    fft_field<L, false>(v, buf, t, tw);
    if (ok) {
      const double s = strength ? strength[row] : 1.0;
      const double zi = (double)row * dzp;
#pragma unroll
      for (int m = 0; m < 8; m++) {
        double sn, cs;
        sincos(zi * Dp[t + T * m], &sn, &cs);
        acc[m] = acc[m] + (s / N) * (mk(cs, -sn) * v[m]);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 8; m++) partial[(size_t)group * N + t + T * m] = acc[m];
}
