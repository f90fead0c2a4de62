// nlm_adjoint.hpp -- reverse-mode (adjoint) split step for the Green's function.
//
// Gradients follow one convention throughout: for a real loss L and a complex
// value z, g_z = dL/dRe(z) + i dL/dIm(z) (the convention of torch's complex
// autograd).  With it, a complex-linear map b = m a has g_a = conj(m) g_b, a
// conjugate-linear map b = m conj(a) has g_a = m conj(g_b), the forward DFT
// (b = F a) has g_a = N * IFFT(g_b) and the scaled inverse DFT has
// g_a = FFT(g_b) / N.
//
// The forward Green's launch (k_prop_tab with kp.ckpt set) stores the
// time-domain field every K steps.  k_adj_rev sweeps all steps backwards,
// rebuilding each forward state on chip by inverting the step (the states
// between checkpoints are recomputed, backwards, instead of replayed from the
// checkpoint and stored): adjoint of the dispersion step (forward FFT, conj(exp(i dz D)),
// inverse FFT), then the transposed RK4 Jacobian of the DiffEq.  The per-step
// coefficient gradients are reduced within each wave (deterministic shuffles)
// and written as per-wave partials.  The pump enters through its linear
// propagation P_r = s_r IFFT(FFT(envelope) exp(i r dzp D_p)), so the gradient
// of each pump row, summed over the workgroup's fields in LDS, is transformed
// and folded into a per-workgroup accumulator of dL/dFFT(envelope) right
// away; nothing of size nZ x N is written per workgroup.  All partials are
// summed by separate passes in a fixed order (no atomics).
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
  const cplx* ckpt;    // [nFields][nSeg][N] scaled time-domain states a_{sK}
  const cplx* fout;    // [nFields][nOut*N] the forward's output rows, domain = outTime
  const cplx* inv;     // [nZ][N][2] inverse RK4 step propagators
  const cplx* gout;    // [nFields][nOut*N], domain = outTime
  cplx* gin;           // [nFields][N] gradient w.r.t. the input (input domain) or nullptr
  cplx* gcoef;         // [nWaves][nZ][3] per-wave partials of the step coefficients
  cplx* gpump;         // [nWG][nZP][N] per-workgroup pump-row partials, or nullptr
  cplx* genvF;         // [nWG][2][N] per-workgroup partials of dL/dFFT(envelope), or nullptr
  cplx* genvT;         // [nWG][N] per-workgroup partials of dL/d(pump row 0) (with genvF)
  const double* Dp;    // [N] pump dispersion D_p(omega)        (genvF)
  const cplx* ph1;     // [N] exp(i dzp D_p), ph2 = exp(2 i dzp D_p) (genvF)
  const cplx* ph2;
  const double* strength;  // [2nZ-1] beam-profile factor per pump row, or nullptr
  double dzp;
  const cplx* tw;
  const cplx* dstep;
  const cplx* halfIn;
  const cplx* halfOut;
  const cplx* coef;    // [nZ][3] step coefficients c = pol * nlStep * mismatch
  const cplx* pump0;   // [2nZ-1][N] pump time grid
  const cplx* tab;     // [nZ][N][2] RK4 step propagators (alpha, beta) / N
};

template <int L>
struct AdjLaunch {
  static constexpr int T = FftShape<L>::T;
  // 256 threads (one wave per SIMD) leave each wave the full register file:
  // the adjoint holds the field, its gradient and three pump-row gradients
  static constexpr int FA = T >= 256 ? 1 : 256 / T;  // fields per workgroup
  static constexpr int THREADS = FA * T;
  static constexpr int WAVES = (THREADS + 63) / 64;
  // LDS: one FFT buffer per field; with several fields also a second row
  // buffer per field (the pump-row gradients are summed over the fields through
  // the two) and the envelope-gradient accumulators and phases of groups 0, 1
  static constexpr int LDS_ELEMS = (FA == 1 ? 1 : 2 * FA + 4) * FftShape<L>::lds_elems;
  static constexpr int ACC_OFF = 2 * FA * FftShape<L>::lds_elems;
};

// wave-wide sum (xor butterfly; every lane ends with the total)
NLM_DEV double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One forward split step i from the scaled state a (time domain) with the
// step propagator table: a <- IFFT(FFT(alpha a + beta conj(a)) exp(i dz D)).
template <int L>
NLM_DEV void fwd_step_tab(cplx (&a)[8], int i, const AParams& ap, cplx* buf, int t, const FftTwiddles<L>& tw) {
  constexpr int N = FftShape<L>::N, T = FftShape<L>::T;
  const cplx* row = ap.tab + (size_t)i * N * 2;
#pragma unroll
  for (int m = 0; m < 8; m++) {
    const cplx al = row[(size_t)(t + T * m) * 2], be = row[(size_t)(t + T * m) * 2 + 1];
    const cplx x = a[m] * (double)N;  // the table holds alpha / N, beta / N (exact scaling)
    a[m] = mk(al.re * x.re - al.im * x.im + be.re * x.re + be.im * x.im,
              al.re * x.im + al.im * x.re + be.im * x.re - be.re * x.im);
  }
  fft_field<L, false>(a, buf, t, tw);
#pragma unroll
  for (int m = 0; m < 8; m++) a[m] = a[m] * ap.dstep[t + T * m];
  fft_field<L, true>(a, buf, t, tw);
  scale_inv<L>(a);
}

// Sum a pump-row gradient over the workgroup's fields (through LDS) and write
// it to this workgroup's partial row.
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
    for (int f = 1; f < FA; f++) s = s + lds[f * F::lds_elems + lpad(k)];
    dst[k] = s;
  }
}

// Reverse sweep of the Chi2PDC Green's function (k = c P conj(y), one signal
// mode, one pump): one field per T lanes, FA fields per workgroup, all steps in
// one launch.  The forward states are rebuilt backwards on chip: the
// dispersion step is unitary and the RK4 step x -> alpha x + beta conj(x) of
// this medium is invertible, so a_{i-1} = R_i^-1(IFFT(FFT(a_i) conj(exp(i dz D)))),
// starting from the forward's output row and replaced by the stored
// checkpoint at every multiple of K steps, which bounds the rounding drift.
template <int L>
__global__ void __launch_bounds__(AdjLaunch<L>::THREADS) k_adj_rev(AParams ap) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, FA = AdjLaunch<L>::FA;
  __shared__ cplx lds[AdjLaunch<L>::LDS_ELEMS];
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
  const int i1 = ap.nZ - 1;

  // gradient of the last frequency row: output = F_last * halfOut (freq) or its IFFT (time)
  cplx g[8];
  cplx carry[8];  // pump row 2i gradient from step i+1 (its k1 stage)
  const int op = ap.outPos[0];
#pragma unroll
  for (int m = 0; m < 8; m++)
    g[m] = (active && op >= 0) ? ap.gout[(size_t)fld * ap.nOut * N + (size_t)op * N + t + T * m] : mk(0, 0);
  if (ap.outTime) {  // output = IFFT(F) -> g_F = FFT(g) / N
    fft_field<L, false>(g, buf, t, tw);
    scale_inv<L>(g);
  }
#pragma unroll
  for (int m = 0; m < 8; m++) {
    g[m] = conj(ap.halfOut[t + T * m]) * g[m];
    carry[m] = mk(0, 0);
  }
  bool gFreq = true;  // g currently holds a frequency-domain gradient

  // state a_{i-1} of the step being reversed, time domain
  cplx x[8];
  if (i1 >= 1) {
    if ((i1 - 1) % ap.K == 0) {
#pragma unroll
      for (int m = 0; m < 8; m++) x[m] = ap.ckpt[((size_t)fld * ap.nSeg + (i1 - 1) / ap.K) * N + t + T * m];
    } else {
      // F_last = output * conj(halfOut), r_{nZ-1} = IFFT(F_last conj(exp(i dz D))), a = R^-1(r)
#pragma unroll
      for (int m = 0; m < 8; m++) x[m] = ap.fout[(size_t)fld * ap.nOut * N + (size_t)op * N + t + T * m];
      if (ap.outTime) fft_field<L, false>(x, buf, t, tw);
#pragma unroll
      for (int m = 0; m < 8; m++) x[m] = conj(ap.halfOut[t + T * m] * ap.dstep[t + T * m]) * x[m];
      fft_field<L, true>(x, buf, t, tw);
      scale_inv<L>(x);
      const cplx* iv = ap.inv + (size_t)i1 * N * 2;
#pragma unroll
      for (int m = 0; m < 8; m++) {
        const cplx ia = iv[(size_t)(t + T * m) * 2], ib = iv[(size_t)(t + T * m) * 2 + 1];
        x[m] = ia * x[m] + ib * conj(x[m]);
      }
    }
  }
  cplx* gpumpWG = ap.gpump ? ap.gpump + (size_t)blockIdx.x * nZP * N : nullptr;

  // envelope gradient accumulators: group 0 takes the even pump rows 2i,
  // group 1 (group 0 if FA == 1) the odd rows 2i-1; ph = exp(-i r dzp D_p)
  // for the group's next row r, advanced by exp(2 i dzp D_p) per step
  // (registers when FA == 1, LDS otherwise)
  constexpr int E = F::lds_elems;
  const bool envAcc = ap.genvF != nullptr;
  cplx acc[FA == 1 ? 8 : 1], ph[FA == 1 ? 8 : 1];
  cplx* accL = lds + AdjLaunch<L>::ACC_OFF + (lf & 1) * 2 * E;
  cplx* phL = accL + E;
  if (envAcc && lf < 2) {
    const int r0 = lf == 0 ? 2 * i1 : 2 * i1 - 1;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      double sn, cs;
      sincos_reduced((double)r0 * ap.dzp * ap.Dp[t + T * m], &sn, &cs);
      if constexpr (FA == 1) {
        acc[m] = mk(0, 0);
        ph[m] = mk(cs, -sn);
      } else {
        accL[lpad(t + T * m)] = mk(0, 0);
        phL[lpad(t + T * m)] = mk(cs, -sn);
      }
    }
  }

  {
    for (int i = i1; i >= 1; i--) {
      // adjoint of the dispersion step: F = FFT(r) exp(i dz D); a_i = IFFT(F) (i < nZ-1)
      if (!gFreq) {
        fft_field<L, false>(g, buf, t, tw);
        scale_inv<L>(g);
      }
#pragma unroll
      for (int m = 0; m < 8; m++) g[m] = conj(ap.dstep[t + T * m]) * g[m];
      fft_field<L, true>(g, buf, t, tw);  // g_r = N * IFFT(.) = unscaled inverse
      gFreq = false;

      // adjoint of the RK4 step (src/_NonlinearMedium.hpp:207-210 with
      // Chi2PDC::DiffEq, solvers/Chi2PDC.cpp:27-47) at the stored state a_{i-1}:
      // m_s = c_s P_s, k = m conj(y)  ->  g_y = m conj(g_k), g_m = y g_k
      const cplx* c = ap.coef + (size_t)i * 3;
      const cplx c0 = c[0], c1 = c[1], c2 = c[2];
      cplx gc0 = mk(0, 0), gc1 = mk(0, 0), gc2 = mk(0, 0);
      // pump-row gradients of this step: registers (FA == 1) or the field's two
      // LDS row buffers, free once the transform above has finished
      cplx rowE[FA == 1 ? 8 : 1], rowO[FA == 1 ? 8 : 1];
      const bool rows = gpumpWG || envAcc;
      if constexpr (FA > 1) {
        if (rows) lds_barrier();
      }
#pragma unroll
      for (int m = 0; m < 8; m++) {
        const int k = t + T * m;
        const cplx xs = x[m];
        const cplx P0 = ap.pump0[(size_t)(2 * i - 2) * N + k];
        const cplx P1 = ap.pump0[(size_t)(2 * i - 1) * N + k];
        const cplx P2 = ap.pump0[(size_t)(2 * i) * N + k];
        const cplx m0 = c0 * P0, m1 = c1 * P1, m2 = c2 * P2;
        const cplx y2 = xs + 0.5 * (m0 * conj(xs));
        const cplx y3 = xs + 0.5 * (m1 * conj(y2));
        const cplx y4 = xs + m1 * conj(y3);
        const cplx gr = g[m];
        const cplx gk4 = gr * (1. / 6.);
        cplx gk3 = gr * (1. / 3.), gk2 = gr * (1. / 3.), gk1 = gk4, gx = gr, gy;
        gy = m2 * conj(gk4);
        const cplx gm2 = y4 * gk4;
        gx = gx + gy;
        gk3 = gk3 + gy;
        gy = m1 * conj(gk3);
        cplx gm1 = y3 * gk3;
        gx = gx + gy;
        gk2 = gk2 + 0.5 * gy;
        gy = m1 * conj(gk2);
        gm1 = gm1 + y2 * gk2;
        gx = gx + gy;
        gk1 = gk1 + 0.5 * gy;
        gy = m0 * conj(gk1);
        const cplx gm0 = xs * gk1;
        gx = gx + gy;
        g[m] = active ? gx : mk(0, 0);
        gc0 = gc0 + conj(P0) * gm0;
        gc1 = gc1 + conj(P1) * gm1;
        gc2 = gc2 + conj(P2) * gm2;
        const cplx re = active ? conj(c2) * gm2 + carry[m] : mk(0, 0);
        const cplx ro = active ? conj(c1) * gm1 : mk(0, 0);
        carry[m] = active ? conj(c0) * gm0 : mk(0, 0);
        if constexpr (FA == 1) {
          rowE[m] = re;
          rowO[m] = ro;
        } else if (rows) {
          lds[lf * E + lpad(k)] = re;
          lds[(FA + lf) * E + lpad(k)] = ro;
        }
      }
      if constexpr (FA > 1) {
        if (rows) lds_barrier();
      }
      if (!active) gc0 = gc1 = gc2 = mk(0, 0);
      // per-wave partial of the coefficient gradients of step i
      const double r0 = wave_sum(gc0.re), i0r = wave_sum(gc0.im), r1 = wave_sum(gc1.re), i1r = wave_sum(gc1.im),
                   r2 = wave_sum(gc2.re), i2r = wave_sum(gc2.im);
      if (lane == 0) {
        cplx* dst = ap.gcoef + ((size_t)wave * ap.nZ + i) * 3;
        dst[0] = mk(r0, i0r);
        dst[1] = mk(r1, i1r);
        dst[2] = mk(r2, i2r);
      }
      // pump rows 2i (k4 of step i + k1 of step i+1) and 2i-1 (k2, k3 of step i)
      if (gpumpWG) {
        cplx* dE = gpumpWG + (size_t)(2 * i) * N;
        cplx* dO = gpumpWG + (size_t)(2 * i - 1) * N;
        if constexpr (FA == 1) {
#pragma unroll
          for (int m = 0; m < 8; m++) {
            dE[t + T * m] = rowE[m];
            dO[t + T * m] = rowO[m];
          }
        } else {
          for (int k = threadIdx.x; k < N; k += AdjLaunch<L>::THREADS) {
            cplx se = mk(0, 0), so = mk(0, 0);
            for (int f = 0; f < FA; f++) {
              se = se + lds[f * E + lpad(k)];
              so = so + lds[(FA + f) * E + lpad(k)];
            }
            dE[k] = se;
            dO[k] = so;
          }
        }
      }
      if (envAcc) {
        // g_F0 += s_r / N conj(exp(i r dzp D_p)) FFT(g_row r)   (rows r >= 1)
        const double sE = (ap.strength ? ap.strength[2 * i] : 1.0) / N;
        const double sO = (ap.strength ? ap.strength[2 * i - 1] : 1.0) / N;
        if constexpr (FA == 1) {
          fft_field<L, false>(rowE, buf, t, tw);
          fft_field<L, false>(rowO, buf, t, tw);
#pragma unroll
          for (int m = 0; m < 8; m++) {
            const cplx p1 = ap.ph1[t + T * m], p2 = ap.ph2[t + T * m];
            const cplx phO = ph[m] * p1;
            acc[m] = acc[m] + sE * (ph[m] * rowE[m]) + sO * (phO * rowO[m]);
            ph[m] = ph[m] * p2;
          }
        } else {
          // group 0 transforms the workgroup sum of row 2i, group 1 that of
          // row 2i-1; the other groups transform zeros (uniform barriers)
          cplx v[8];
          const int base = lf == 0 ? 0 : FA * E;
#pragma unroll
          for (int m = 0; m < 8; m++) {
            cplx s = mk(0, 0);
            if (lf < 2)
              for (int f = 0; f < FA; f++) s = s + lds[base + f * E + lpad(t + T * m)];
            v[m] = s;
          }
          fft_field<L, false>(v, buf, t, tw);
          if (lf < 2) {
            const double sc = lf == 1 ? sO : sE;
#pragma unroll
            for (int m = 0; m < 8; m++) {
              const int q = lpad(t + T * m);
              const cplx phm = phL[q];
              accL[q] = accL[q] + sc * (phm * v[m]);
              phL[q] = phm * ap.ph2[t + T * m];
            }
          }
        }
      }
      // state a_{i-2} for the next (earlier) step
      if (i >= 2) {
        if ((i - 2) % ap.K == 0) {
#pragma unroll
          for (int m = 0; m < 8; m++) x[m] = ap.ckpt[((size_t)fld * ap.nSeg + (i - 2) / ap.K) * N + t + T * m];
        } else {
          fft_field<L, false>(x, buf, t, tw);
#pragma unroll
          for (int m = 0; m < 8; m++) x[m] = conj(ap.dstep[t + T * m]) * x[m];
          fft_field<L, true>(x, buf, t, tw);
          scale_inv<L>(x);
          const cplx* iv = ap.inv + (size_t)(i - 1) * N * 2;
#pragma unroll
          for (int m = 0; m < 8; m++) {
            const cplx ia = iv[(size_t)(t + T * m) * 2], ib = iv[(size_t)(t + T * m) * 2 + 1];
            x[m] = ia * x[m] + ib * conj(x[m]);
          }
        }
      }
    }
  }
  if (gpumpWG) emit_pump_row<L>(carry, lds, lf, t, gpumpWG);  // row 0
  if (envAcc) {
    // row 0 enters the envelope directly: workgroup sum of its gradient
    if constexpr (FA == 1) {
#pragma unroll
      for (int m = 0; m < 8; m++) {
        ap.genvT[(size_t)blockIdx.x * N + t + T * m] = carry[m];
        ap.genvF[((size_t)blockIdx.x * 2) * N + t + T * m] = acc[m];
        ap.genvF[((size_t)blockIdx.x * 2 + 1) * N + t + T * m] = mk(0, 0);
      }
    } else {
      lds_barrier();
#pragma unroll
      for (int m = 0; m < 8; m++) lds[lf * E + lpad(t + T * m)] = carry[m];
      lds_barrier();
      for (int k = threadIdx.x; k < N; k += AdjLaunch<L>::THREADS) {
        cplx s = mk(0, 0);
        for (int f = 0; f < FA; f++) s = s + lds[f * E + lpad(k)];
        ap.genvT[(size_t)blockIdx.x * N + k] = s;
      }
      if (lf < 2) {
#pragma unroll
        for (int m = 0; m < 8; m++)
          ap.genvF[((size_t)blockIdx.x * 2 + lf) * N + t + T * m] = accL[lpad(t + T * m)];
      }
    }
  }

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

// exp(i dzp D_p), exp(2 i dzp D_p)
__global__ void k_pump_phase_steps(const double* __restrict__ Dp, double dzp, int N, cplx* __restrict__ ph1,
                                   cplx* __restrict__ ph2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  double sn, cs;
  sincos(dzp * Dp[k], &sn, &cs);
  ph1[k] = mk(cs, sn);
  sincos(2.0 * dzp * Dp[k], &sn, &cs);
  ph2[k] = mk(cs, sn);
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
