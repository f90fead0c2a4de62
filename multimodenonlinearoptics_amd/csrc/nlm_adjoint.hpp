// nlm_adjoint.hpp -- reverse-mode (adjoint) split step for the Green's function.
//
// Gradients follow one convention throughout: for a real loss L and a complex
// value z, g_z = dL/dRe(z) + i dL/dIm(z) (the convention of torch's complex
// autograd).  With it, a complex-linear map b = m a has g_a = conj(m) g_b, a
// conjugate-linear map b = m conj(a) has g_a = m conj(g_b), the forward DFT
// (b = F a) has g_a = N * IFFT(g_b) and the scaled inverse DFT has
// g_a = FFT(g_b) / N.
//
// The forward Green's launch (k_prop_tab with kp.ckpt set) runs each step as
// its R-linear propagator r = alpha a + beta conj(a) (the RK4 step of the
// medium, tabulated per step and time point) followed by the dispersion step
// a' = IFFT(D FFT(r)), and stores the time-domain field every K steps.  The
// reverse sweep (k_adj_sweep) walks the steps backwards for every field:
//   h = g_r = IFFT(conj(D) FFT(g_a')),   g_a = conj(alpha) h + beta conj(h),
// and the table gradients g_alpha = sum_fields conj(a) h, g_beta = sum_fields a h.
// The forward states are rebuilt on chip by inverting the step (dispersion is
// unitary, the R-linear step is invertible: a = (conj(alpha) r - beta conj(r)) /
// (|alpha|^2 - |beta|^2)), anchored at the checkpoints.  The table holds each
// step as the real 2 x 2 matrix with columns R(1), R(i) (k_step_table), in which
// g_a is its transpose and the inverse its adjugate over the determinant.
// Each workgroup sums g_alpha, g_beta over its fields in LDS and writes one
// partial per step; a second kernel (k_adj_reduce) sums the partials in a
// fixed order and applies the chain rule from (alpha, beta) = (R(1) -+ i R(i))/2
// through the RK4 step to m_s = c_s P_s, the products of the step
// coefficients and pump rows.  Nothing of size nZ x N is written per field.
//
// Forward map replaced: signalSimulationTemplate (src/_NonlinearMedium.hpp:160-239)
// with Chi2PDC::DiffEq (solvers/Chi2PDC.cpp:27-47); assembly of
// computeGreensFunction (src/_NonlinearMedium.cpp:333-372); linear pump
// (src/_NonlinearMedium.cpp:225-250).
#pragma once
#include "nlm_prop.hpp"

struct AParams {
  int nZ, nFields, K, nSeg;
  int outTime;
  int iHi, iLo;        // this launch: steps iHi down to iLo + 1
  int chunk;           // step capacity of `part`
  const cplx* ckpt;    // [nFields][nSeg][N] time-domain states a_{sK}
  const cplx* fout;    // [nFields][N] the forward's output rows, domain = outTime
  const cplx* gout;    // [nFields][N] dL/d(output rows), domain = outTime
  cplx* state;         // [nFields][2][N] a_{i-1} and g_{a_i} between launches
  cplx* part;          // [nWG][chunk][2][N] workgroup sums of conj(a) h, a h
  const cplx* tab;     // [nZ][N][2] RK4 step propagators as real matrices / N (k_step_table)
  const cplx* tw;
  const cplx* dstep;
  const cplx* halfOut;
};

// Fields per workgroup of the sweep: as many as keep the workgroup at 256
// lanes (one wave per SIMD, so each wave has the whole register file for the
// field, its gradient and the transform twiddles), at least 256 workgroups
// for the 2N fields, and the LDS (one transform buffer per field) within 80 KB.
// More fields per workgroup make the per-step partials smaller.
template <int L>
struct AdjShape {
  static constexpr int N = 1 << L, T = FftShape<L>::T;
  // two transform buffers per field (the gradient and the state are transformed together)
  static constexpr int byLds = (80 * 1024) / (2 * FftShape<L>::lds_elems * 16);
  static constexpr int byWG = N / 128 > 1 ? N / 128 : 1;
  static constexpr int byLanes = T < 256 ? 256 / T : 1;
  static constexpr int m1 = byLds < byWG ? byLds : byWG;
  static constexpr int m2 = m1 < byLanes ? m1 : byLanes;
  // N = 2048: two fields per workgroup (512 lanes, two waves per SIMD, lean
  // registers below), which halves the partials against one field per workgroup
  static constexpr int FA = L == 11 ? 2 : m2 >= 4 ? 4 : m2 >= 2 ? 2 : 1;
  static constexpr int THREADS = FA * T;
  // LEAN: fewer than 512 registers per lane (two or more waves per SIMD): the
  // twiddles come from the table (an LDS copy up to N = 4096, the global table
  // at N = 8192) and the dispersion and propagator rows are read where used
  static constexpr bool LEAN = L >= 12 || THREADS > 256;
  static constexpr bool TW_LDS = LEAN && L <= 12;
  static constexpr int TW_ELEMS = TW_LDS ? N : 1;
  static constexpr bool FUSE =
      !LEAN && 2 * FA * FftShape<L>::lds_elems * 16 <= 160 * 1024 && THREADS <= 256;
  static constexpr int BUFS = FUSE ? 2 * FA : FA;  // LDS transform buffers per workgroup
  static constexpr int LDS_ELEMS = BUFS * FftShape<L>::lds_elems;
};

// Sum conj(x) h (CONJ) or x h over the workgroup's FA fields and write the
// total to dst[0..N) (through the fields' transform buffers, which are free
// between transforms).

template <int L, int FA, bool CONJ>
NLM_DEV void wg_sum_product(const cplx (&x)[8], const cplx (&h)[8], bool active, cplx* lds, int lf, int t,
                            cplx* __restrict__ dst) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, E = F::lds_elems;
  if constexpr (FA == 1) {
#pragma unroll
    for (int m = 0; m < 8; m++) store_stream(dst + t + T * m, active ? (CONJ ? conj(x[m]) : x[m]) * h[m] : mk(0, 0));
  } else {
    lds_barrier();
#pragma unroll
    for (int m = 0; m < 8; m++) lds[lf * E + lpad<L>(t + T * m)] = active ? (CONJ ? conj(x[m]) : x[m]) * h[m] : mk(0, 0);
    lds_barrier();
#pragma unroll
    for (int j = 0; j < 8 / FA; j++) {
      const int k = lf * T + t + FA * T * j;
      cplx s = lds[lpad<L>(k)];
#pragma unroll
      for (int f = 1; f < FA; f++) s = s + lds[f * E + lpad<L>(k)];
      store_stream(dst + k, s);
    }
  }
}

// Both table sums at once, g_alpha partial = sum conj(x) h to dst[0..N) and
// g_beta partial = sum x h to dst[N..2N), through 2 FA transform buffers (two
// barriers instead of four).
template <int L, int FA>
NLM_DEV void wg_sum_products2(const cplx (&x)[8], const cplx (&h)[8], bool active, cplx* lds, int lf, int t,
                              cplx* __restrict__ dst) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, E = F::lds_elems;
  lds_barrier();
#pragma unroll
  for (int m = 0; m < 8; m++) {
    lds[lf * E + lpad<L>(t + T * m)] = active ? conj(x[m]) * h[m] : mk(0, 0);
    lds[(FA + lf) * E + lpad<L>(t + T * m)] = active ? x[m] * h[m] : mk(0, 0);
  }
  lds_barrier();
#pragma unroll
  for (int j = 0; j < 8 / FA; j++) {
    const int k = lf * T + t + FA * T * j;
    cplx sa = lds[lpad<L>(k)], sb = lds[FA * E + lpad<L>(k)];
#pragma unroll
    for (int f = 1; f < FA; f++) {
      sa = sa + lds[f * E + lpad<L>(k)];
      sb = sb + lds[(FA + f) * E + lpad<L>(k)];
    }
    store_stream(dst + k, sa);
    store_stream(dst + N + k, sb);
  }
}

template <int L>
NLM_DEV void prop_row(cplx (&a)[8], cplx (&b)[8], const cplx* __restrict__ row, int t) {
  constexpr int T = FftShape<L>::T;
#pragma unroll
  for (int m = 0; m < 8; m++) {
    a[m] = row[(size_t)(t + T * m) * 2];
    b[m] = row[(size_t)(t + T * m) * 2 + 1];
  }
}

// x = R^-1(r) for the step r = R x, R the real matrix [[p.re, p.im], [q.re, q.im]] / N
// acting on (Re x, Im x) (the table rows): x = adj(R) r / det(R), where det(R) N^2 =
// |alpha|^2 - |beta|^2.
template <int L>
NLM_DEV void inverse_step(cplx (&x)[8], const cplx (&r)[8], const cplx (&p)[8], const cplx (&q)[8]) {
  constexpr double N = FftShape<L>::N;
#pragma unroll
  for (int m = 0; m < 8; m++) {
    const double den = N * (p[m].re * q[m].im - p[m].im * q[m].re);
    // hardware reciprocal refined by two Newton steps (relative error ~1 ulp;
    // den is the determinant of a step close to the identity, far from 0 and
    // from the denormal range)
    double rd = __builtin_amdgcn_rcp(den);
    rd = rd * __builtin_fma(-den, rd, 2.0);
    rd = rd * __builtin_fma(-den, rd, 2.0);
    x[m] = mk((q[m].im * r[m].re - p[m].im * r[m].im) * rd, (p[m].re * r[m].im - q[m].re * r[m].re) * rd);
  }
}

// Reverse sweep over steps iHi .. iLo+1 for FA fields per workgroup (T lanes each).
// RB = false: every state of the launch's steps is a kept checkpoint (K = 1), so no
// state is rebuilt and the registers of the rebuild (the previous step's rows and
// the state's transform) are not held.
template <int L, int FA, bool RB>
__global__ void __launch_bounds__(FA * FftShape<L>::T) k_adj_sweep(AParams ap) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, E = F::lds_elems;
  constexpr bool FUSE = AdjShape<L>::FUSE;
  // With 256 or 128 registers per lane (AdjShape::LEAN) the twiddles come from
  // the table and the dispersion and propagator rows are read where they are
  // used instead of being held.
  constexpr bool LEAN = AdjShape<L>::LEAN;
  __shared__ cplx lds[AdjShape<L>::LDS_ELEMS];
  __shared__ cplx twl[AdjShape<L>::TW_ELEMS];
  if constexpr (AdjShape<L>::TW_LDS) {
    for (int k = threadIdx.x; k < N; k += blockDim.x) twl[k] = ap.tw[k];
    __syncthreads();
  }
  const int lf = threadIdx.x / T, t = threadIdx.x % T;
  const int field = blockIdx.x * FA + lf;
  const bool active = field < ap.nFields;
  const int fld = active ? field : 0;
  cplx* buf = lds + lf * E;

  typename std::conditional<LEAN, TableTwiddles<L>, RegTwiddles<L>>::type tw;
  if constexpr (LEAN) {
    tw.tab = AdjShape<L>::TW_LDS ? twl : ap.tw;
    tw.t = t;
  } else {
    fft_load_twiddles<L>(tw.w, ap.tw, t);
  }
  cplx g[8], x[8];
  bool gFreq;
  if (ap.iHi == ap.nZ - 1) {
    // the output row is halfOut * FFT(a') (frequency) or its IFFT (time)
#pragma unroll
    for (int m = 0; m < 8; m++) g[m] = active ? ap.gout[(size_t)fld * N + t + T * m] : mk(0, 0);
    if (ap.outTime) {  // output = IFFT(F) -> g_F = FFT(g) / N
      fft_field1<L, false>(g, buf, t, tw);
      scale_inv<L>(g);
    }
#pragma unroll
    for (int m = 0; m < 8; m++) g[m] = conj(ap.halfOut[t + T * m]) * g[m];
    gFreq = true;
    const int i1 = ap.nZ - 1;
    if ((i1 - 1) % ap.K == 0) {
#pragma unroll
      for (int m = 0; m < 8; m++) x[m] = ap.ckpt[((size_t)fld * ap.nSeg + (i1 - 1) / ap.K) * N + t + T * m];
    } else {
      // F_last = output * conj(halfOut), r = IFFT(F_last conj(exp(i dz D))), a = R^-1(r)
#pragma unroll
      for (int m = 0; m < 8; m++) x[m] = ap.fout[(size_t)fld * N + t + T * m];
      if (ap.outTime) fft_field1<L, false>(x, buf, t, tw);
#pragma unroll
      for (int m = 0; m < 8; m++) x[m] = conj(ap.halfOut[t + T * m] * ap.dstep[t + T * m]) * x[m];
      fft_field1<L, true>(x, buf, t, tw);
      scale_inv<L>(x);
      cplx ra[8], rb[8], r[8];
      prop_row<L>(ra, rb, ap.tab + (size_t)i1 * N * 2, t);
#pragma unroll
      for (int m = 0; m < 8; m++) r[m] = x[m];
      inverse_step<L>(x, r, ra, rb);
    }
  } else {
    const cplx* st = ap.state + (size_t)fld * 2 * N;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      x[m] = st[t + T * m];
      g[m] = st[N + t + T * m];
    }
    gFreq = false;
  }

  cplx dsr[LEAN ? 1 : 8];  // conj(exp(i dz D)) / N
  if constexpr (!LEAN) {
#pragma unroll
    for (int m = 0; m < 8; m++) dsr[m] = conj(ap.dstep[t + T * m]) * (1.0 / N);
  }
  // propagator rows (matrix rows p, q) / N of this step (al, be) and of the step
  // before it (pa, pb, which also inverts it), loaded a step ahead so their
  // latency hides behind the transforms (at the point of use from N = 4096)
  // (with RB false this step's rows are loaded at the top of the step instead)
  constexpr int PR = LEAN || !RB ? 1 : 8;
  // LATE (every state kept, full registers): this step's rows and the next
  // state are fetched after the forward transform of g, so neither is held in
  // registers across it (their latency hides behind the inverse transform)
  constexpr bool LATE = !LEAN && !RB;
  cplx al[8], be[8], pa[PR], pb[PR];
  if constexpr (!LEAN && RB) {
    prop_row<L>(al, be, ap.tab + (size_t)ap.iHi * N * 2, t);
    if (ap.iHi >= 2) prop_row<L>(pa, pb, ap.tab + (size_t)(ap.iHi - 1) * N * 2, t);
  }

  // N = 2048 and 4096 (two waves per SIMD): the lane index is opaque to the
  // compiler inside each step, so the offsets derived from it are not hoisted out
  // of the loop into registers (the rebuilding sweep spilled 38 registers to
  // scratch at N = 2048 and 16 at N = 4096 with them hoisted, 2 without; at
  // N = 8192, four waves per SIMD, it spills more, so it keeps them hoisted)
  constexpr bool OPQ = LEAN && L <= 12;
  const int t0 = t;
  for (int i = ap.iHi; i > ap.iLo; i--) {
    const int t = OPQ ? opaque_lane(t0) : t0;
    if constexpr (OPQ) retarget(tw, t);
    // adjoint of the dispersion step: h = g_r = IFFT(conj(exp(i dz D)) FFT(g));
    // with FUSE the state's dispersion step is undone in the same transforms:
    // xr = IFFT(conj(exp(i dz D)) FFT(a_{i-1})), from which a_{i-2} = R_{i-1}^-1(xr)
    const bool rebuild = RB && i >= 2 && (i - 2) % ap.K != 0;
    const bool fused = FUSE && rebuild && !gFreq;
    cplx xr[8];
    if (fused) {
#pragma unroll
      for (int m = 0; m < 8; m++) xr[m] = x[m];
      fft_field2<L, false>(g, xr, buf, lds + (FA + lf) * E, t, tw);
#pragma unroll
      for (int m = 0; m < 8; m++) {
        g[m] = (LEAN ? conj(ap.dstep[t + T * m]) * (1.0 / N) : dsr[LEAN ? 0 : m]) * g[m];
        xr[m] = (LEAN ? conj(ap.dstep[t + T * m]) * (1.0 / N) : dsr[LEAN ? 0 : m]) * xr[m];
      }
      fft_field2<L, true>(g, xr, buf, lds + (FA + lf) * E, t, tw);
    } else {
      if (!gFreq) {
        fft_field1<L, false>(g, buf, t, tw);
      } else {
#pragma unroll
        for (int m = 0; m < 8; m++) g[m] = g[m] * (double)N;  // ds carries 1/N
      }
#pragma unroll
      for (int m = 0; m < 8; m++) g[m] = (LEAN ? conj(ap.dstep[t + T * m]) * (1.0 / N) : dsr[LEAN ? 0 : m]) * g[m];
      if constexpr (LATE) {
        prop_row<L>(al, be, ap.tab + (size_t)i * N * 2, t);
        // state a_{i-1} of this step (kept checkpoint; the first step's is set above)
        if (i != ap.iHi) {
#pragma unroll
          for (int m = 0; m < 8; m++) x[m] = ap.ckpt[((size_t)fld * ap.nSeg + (i - 1) / ap.K) * N + t + T * m];
        }
      }
      fft_field1<L, true>(g, buf, t, tw);  // unscaled inverse
    }
    gFreq = false;

    // adjoint of r = alpha x + beta conj(x), h = g: g_alpha += conj(x) h,
    // g_beta += x h (workgroup sums), g_x = conj(alpha) h + beta conj(h)
    cplx* dst = ap.part + (((size_t)blockIdx.x * ap.chunk + (i - ap.iLo - 1)) * 2) * N;
    if constexpr (FA > 1 && FUSE) {
      wg_sum_products2<L, FA>(x, g, active, lds, lf, t, dst);
    } else {
      wg_sum_product<L, FA, true>(x, g, active, lds, lf, t, dst);
      wg_sum_product<L, FA, false>(x, g, active, lds, lf, t, dst + N);
    }
    if constexpr (LEAN) prop_row<L>(al, be, ap.tab + (size_t)i * N * 2, t);
#pragma unroll
    for (int m = 0; m < 8; m++) {
      // the transpose of the step's real matrix: g = conj(alpha) h + beta conj(h)
      const cplx h = g[m];
      g[m] = mk((al[m].re * h.re + be[m].re * h.im) * (double)N, (al[m].im * h.re + be[m].im * h.im) * (double)N);
    }

    // state a_{i-2} for the next (earlier) step
    if (!LATE && i >= 2) {
      if (!rebuild) {
#pragma unroll
        for (int m = 0; m < 8; m++) x[m] = ap.ckpt[((size_t)fld * ap.nSeg + (i - 2) / ap.K) * N + t + T * m];
      } else {
        if (!fused) {
#pragma unroll
          for (int m = 0; m < 8; m++) xr[m] = x[m];
          fft_field1<L, false>(xr, buf, t, tw);
#pragma unroll
          for (int m = 0; m < 8; m++) xr[m] = (LEAN ? conj(ap.dstep[t + T * m]) * (1.0 / N) : dsr[LEAN ? 0 : m]) * xr[m];
          fft_field1<L, true>(xr, buf, t, tw);
        }
        if constexpr (LEAN || !RB) {
          prop_row<L>(al, be, ap.tab + (size_t)(i - 1) * N * 2, t);
          inverse_step<L>(x, xr, al, be);
        } else {
          inverse_step<L>(x, xr, pa, pb);
        }
      }
      if constexpr (!LEAN && RB) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
          al[m] = pa[m];
          be[m] = pb[m];
        }
        if (i >= 3) prop_row<L>(pa, pb, ap.tab + (size_t)(i - 2) * N * 2, t);
      }
    }
  }
  if (LATE && ap.iLo > 0 && ap.iLo < ap.iHi) {
#pragma unroll
    for (int m = 0; m < 8; m++) x[m] = ap.ckpt[((size_t)fld * ap.nSeg + (ap.iLo - 1) / ap.K) * N + t + T * m];
  }
  if (ap.iLo > 0 && active) {
    cplx* st = ap.state + (size_t)fld * 2 * N;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      st[t + T * m] = x[m];
      st[N + t + T * m] = g[m];
    }
  }
}

constexpr int RED_WAVES = 8;
constexpr int RED_ACC = 4;

// Wave q of a block sums the partials of workgroups wLo + q, wLo + q + RED_WAVES,
// ... < wHi (fixed order, RED_ACC independent accumulators so that many loads are
// in flight per point) for step jj of the launch and the block's 64 TPL points
// from k0, into red[0 / 1][q].
template <int TPL>
__device__ __forceinline__ void sum_partials(const cplx* __restrict__ part, int wLo, int wHi, int chunk, int jj,
                                             int N, int k0, int q, int lane, bool valid,
                                             cplx (&red)[2][RED_WAVES][64 * TPL]) {
  cplx a4[TPL][RED_ACC], b4[TPL][RED_ACC];
#pragma unroll
  for (int u = 0; u < TPL; u++)
#pragma unroll
    for (int v = 0; v < RED_ACC; v++) a4[u][v] = b4[u][v] = mk(0, 0);
  if (valid) {
    const size_t stride = (size_t)RED_WAVES * chunk * 2 * N;
    int w = wLo + q;
    const cplx* p = part + (((size_t)w * chunk + jj) * 2) * N + k0 + lane;
    for (; w + (RED_ACC - 1) * RED_WAVES < wHi; w += RED_ACC * RED_WAVES, p += RED_ACC * stride) {
#pragma unroll
      for (int v = 0; v < RED_ACC; v++)
#pragma unroll
        for (int u = 0; u < TPL; u++) {
          a4[u][v] = a4[u][v] + load_stream(p + v * stride + 64 * u);
          b4[u][v] = b4[u][v] + load_stream(p + v * stride + N + 64 * u);
        }
    }
    for (int v = 0; w < wHi; w += RED_WAVES, p += stride, v++)
#pragma unroll
      for (int u = 0; u < TPL; u++) {
        a4[u][v] = a4[u][v] + load_stream(p + 64 * u);
        b4[u][v] = b4[u][v] + load_stream(p + N + 64 * u);
      }
  }
#pragma unroll
  for (int u = 0; u < TPL; u++) {
#pragma unroll
    for (int h = RED_ACC / 2; h >= 1; h /= 2) {
#pragma unroll
      for (int v = 0; v < h; v++) {
        a4[u][v] = a4[u][v] + a4[u][v + h];
        b4[u][v] = b4[u][v] + b4[u][v + h];
      }
    }
    red[0][q][lane + 64 * u] = a4[u][0];
    red[1][q][lane + 64 * u] = b4[u][0];
  }
}

// First level of the reduction where one step's workgroups are too many for the
// (step, tile) blocks alone to fill the device (N = 2048 and above): block
// (jj, tile, s) sums the workgroups of slice s, [s per, (s + 1) per), in the
// fixed order above and writes the slice total to out[s][jj][2][N], the layout
// of the partials, so that k_adj_reduce then sums the slices (fixed order).
template <int TPL>
__global__ void __launch_bounds__(64 * RED_WAVES) k_adj_slices(const cplx* __restrict__ part, int nWG, int chunk,
                                                               int N, int per, cplx* __restrict__ out) {
  constexpr int TILE = 64 * TPL;
  __shared__ cplx red[2][RED_WAVES][TILE];
  const int tiles = (N + TILE - 1) / TILE, slices = (nWG + per - 1) / per;
  const int s = blockIdx.x % slices, rest = blockIdx.x / slices;
  const int jj = rest / tiles, k0 = (rest % tiles) * TILE;
  const int q = threadIdx.x / 64, lane = threadIdx.x % 64;
  const bool valid = TPL > 1 || k0 + lane < N;
  const int wLo = s * per, wHi = wLo + per < nWG ? wLo + per : nWG;
  sum_partials<TPL>(part, wLo, wHi, chunk, jj, N, k0, q, lane, valid, red);
  __syncthreads();
  if (q != 0 || !valid) return;
  cplx* o = out + (((size_t)s * chunk + jj) * 2) * N + k0 + lane;
#pragma unroll
  for (int u = 0; u < TPL; u++) {
    cplx ga = red[0][0][lane + 64 * u], gb = red[1][0][lane + 64 * u];
#pragma unroll
    for (int r = 1; r < RED_WAVES; r++) {
      ga = ga + red[0][r][lane + 64 * u];
      gb = gb + red[1][r][lane + 64 * u];
    }
    o[64 * u] = ga;
    o[N + 64 * u] = gb;
  }
}

// Sum the partials of steps iLo+1 .. iLo+count and apply the chain rule from
// (alpha, beta) to m_s = c_s P_s: R(x) = x + (k1 + 2 k2 + 2 k3 + k4) / 6 with
// k_s = m_s conj(y_s) (Chi2PDC::DiffEq, solvers/Chi2PDC.cpp:27-47),
// alpha = (R(1) - i R(i)) / 2, beta = (R(1) + i R(i)) / 2.  Writes gm[i][s][tau].
// One block per (step, 64 TPL time points): wave q sums workgroups
// q, q + RED_WAVES, ... (fixed order), lane l taking the points k0 + l + 64 u
// (u < TPL, so each load instruction of a wave reads 1 KB contiguous and a wave
// streams 64 TPL points of each partial); then wave 0 adds the wave sums in order.
template <int TPL>
__global__ void __launch_bounds__(64 * RED_WAVES) k_adj_reduce(const cplx* __restrict__ part, int nWG, int chunk,
                                                               int iLo, int N, const cplx* __restrict__ coef,
                                                               const cplx* __restrict__ pump, cplx* __restrict__ gm) {
  constexpr int TILE = 64 * TPL;
  __shared__ cplx red[2][RED_WAVES][TILE];
  const int tiles = (N + TILE - 1) / TILE;
  const int jj = blockIdx.x / tiles, k0 = (blockIdx.x % tiles) * TILE;
  const int q = threadIdx.x / 64, lane = threadIdx.x % 64;
  // TPL > 1 is launched only when TILE divides N
  const bool valid = TPL > 1 || k0 + lane < N;
  sum_partials<TPL>(part, 0, nWG, chunk, jj, N, k0, q, lane, valid, red);
  __syncthreads();
  if (q != 0 || !valid) return;
#pragma unroll 1
  for (int u = 0; u < TPL; u++) {
    const int k = k0 + lane + 64 * u;
    cplx ga = red[0][0][lane + 64 * u], gb = red[1][0][lane + 64 * u];
#pragma unroll
    for (int r = 1; r < RED_WAVES; r++) {
      ga = ga + red[0][r][lane + 64 * u];
      gb = gb + red[1][r][lane + 64 * u];
    }
    const int i = iLo + 1 + jj;
    const cplx* c = coef + (size_t)i * 3;
    const cplx m0 = c[0] * pump[(size_t)(2 * i - 2) * N + k];
    const cplx m1 = c[1] * pump[(size_t)(2 * i - 1) * N + k];
    const cplx m2 = c[2] * pump[(size_t)(2 * i) * N + k];
    cplx gm0 = mk(0, 0), gm1 = mk(0, 0), gm2 = mk(0, 0);
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
      // x = 1 carries g_R(1) = (g_alpha + g_beta) / 2, x = i carries g_R(i) = i (g_alpha - g_beta) / 2
      const cplx x = pass == 0 ? mk(1, 0) : mk(0, 1);
      const cplx gr = pass == 0 ? 0.5 * (ga + gb) : mk(0, 0.5) * (ga - gb);
      const cplx y2 = x + 0.5 * (m0 * conj(x));
      const cplx y3 = x + 0.5 * (m1 * conj(y2));
      const cplx y4 = x + m1 * conj(y3);
      const cplx gk4 = gr * (1. / 6.);
      cplx gk3 = gr * (1. / 3.), gk2 = gr * (1. / 3.), gk1 = gk4;
      gm2 = gm2 + y4 * gk4;
      gk3 = gk3 + m2 * conj(gk4);
      gm1 = gm1 + y3 * gk3;
      gk2 = gk2 + 0.5 * (m1 * conj(gk3));
      gm1 = gm1 + y2 * gk2;
      gk1 = gk1 + 0.5 * (m1 * conj(gk2));
      gm0 = gm0 + x * gk1;
    }
    cplx* o = gm + (size_t)i * 3 * N;
    o[k] = gm0;
    o[N + k] = gm1;
    o[2 * N + k] = gm2;
  }
}

// g_coef[i][s] = sum_tau conj(P_s) gm[i][s] (fixed-order tree; one block per step).
__global__ void __launch_bounds__(256) k_adj_coef(const cplx* __restrict__ gm, const cplx* __restrict__ pump, int N,
                                                  int nZ, cplx* __restrict__ gcoef) {
  __shared__ cplx red[3][256];
  const int i = blockIdx.x, t = threadIdx.x;
  cplx s0 = mk(0, 0), s1 = mk(0, 0), s2 = mk(0, 0);
  if (i >= 1) {
    const cplx* g = gm + (size_t)i * 3 * N;
    for (int k = t; k < N; k += 256) {
      s0 = s0 + conj(pump[(size_t)(2 * i - 2) * N + k]) * g[k];
      s1 = s1 + conj(pump[(size_t)(2 * i - 1) * N + k]) * g[N + k];
      s2 = s2 + conj(pump[(size_t)(2 * i) * N + k]) * g[2 * N + k];
    }
  }
  red[0][t] = s0;
  red[1][t] = s1;
  red[2][t] = s2;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (t < h) {
      red[0][t] = red[0][t] + red[0][t + h];
      red[1][t] = red[1][t] + red[1][t + h];
      red[2][t] = red[2][t] + red[2][t + h];
    }
    __syncthreads();
  }
  if (t < 3) gcoef[(size_t)i * 3 + t] = red[t][0];
}

// Pump-row gradients g_P[r] = conj(c) gm: row 2j from k4 of step j and k1 of
// step j+1, row 2j-1 from k2, k3 of step j.
__global__ void k_adj_pump_rows(const cplx* __restrict__ gm, const cplx* __restrict__ coef, int N, int nZ,
                                cplx* __restrict__ gP) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nZP = 2 * nZ - 1;
  if (idx >= (long)nZP * N) return;
  const int r = (int)(idx / N), k = (int)(idx % N);
  cplx v = mk(0, 0);
  if (r % 2 == 0) {
    const int j = r / 2;
    if (j >= 1) v = v + conj(coef[(size_t)j * 3 + 2]) * gm[((size_t)j * 3 + 2) * N + k];
    if (j + 1 <= nZ - 1) v = v + conj(coef[(size_t)(j + 1) * 3]) * gm[((size_t)(j + 1) * 3) * N + k];
  } else {
    const int j = (r + 1) / 2;
    v = conj(coef[(size_t)j * 3 + 1]) * gm[((size_t)j * 3 + 1) * N + k];
  }
  gP[idx] = v;
}

// Per-workgroup partial of g_F0 = sum_{r >= 1} s_r / N conj(exp(i r dzp D_p)) FFT(g_P[r])
// over rows r = 1 + blockIdx.x, 1 + blockIdx.x + gridDim.x, ... (one field-transform per workgroup).
template <int L>
__global__ void __launch_bounds__(FftShape<L>::T) k_adj_env_rows(const cplx* __restrict__ gP, int nZP,
                                                                 const double* __restrict__ Dp, double dzp,
                                                                 const double* __restrict__ strength,
                                                                 const cplx* __restrict__ twt, cplx* __restrict__ out) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T;
  __shared__ cplx lds[F::lds_elems];
  const int t = threadIdx.x;
  FftTwiddles<L> tw;
  fft_load_twiddles<L>(tw, twt, t);
  cplx acc[8];
#pragma unroll
  for (int m = 0; m < 8; m++) acc[m] = mk(0, 0);
  for (int r = 1 + blockIdx.x; r < nZP; r += gridDim.x) {
    cplx v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = gP[(size_t)r * N + t + T * m];
    fft_field<L, false>(v, lds, t, tw);
    const double s = (strength ? strength[r] : 1.0) / N;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      double sn, cs;
      sincos_reduced((double)r * dzp * Dp[t + T * m], &sn, &cs);
      acc[m] = acc[m] + s * (mk(cs, -sn) * v[m]);
    }
  }
#pragma unroll
  for (int m = 0; m < 8; m++) out[(size_t)blockIdx.x * N + t + T * m] = acc[m];
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
