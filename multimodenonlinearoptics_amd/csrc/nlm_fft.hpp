// nlm_fft.hpp -- complex FP64 FFT for one field held by a group of lanes.
//
// A field of N = 2^L points is held by T = N/8 lanes, eight points per lane in
// the "stride layout": lane t holds x[t + T*m] for m = 0..7.  The forward and
// inverse transforms take the field in this layout and return it in the same
// layout (frequency index k = t + T*m), so the pointwise work of a split step
// (RK4 stages, dispersion phase, Green's impulse injection) never moves data.
//
// Algorithm: Cooley-Tukey decimation in frequency over the radix sequence
// 8, (2|4), 8, 8, ...  Stage 0 works on the eight points a lane already holds;
// every later stage exchanges points through LDS.  After stage s the data are
// stored sub-array-major, sub-array g (digits c_1..c_s of the output index,
// least significant first) at LDS index g*M_s + b.  The last stage's outputs
// land in natural order, which is the stride layout again.  LDS indices are
// padded by one element per eight so that the 16-byte accesses of eight
// consecutive lanes fall in distinct banks.
//
// Conventions follow the reference's Eigen wrapper (src/CustomEigenFFT.h:54-137):
// forward X_k = sum_j x_j exp(-2 pi i jk/N) unscaled; the inverse uses
// exp(+2 pi i jk/N) and its 1/N factor is applied by the caller.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>

struct __align__(16) cplx {
  double re, im;
};

__host__ __device__ __forceinline__ cplx mk(double re, double im) { return cplx{re, im}; }
__host__ __device__ __forceinline__ cplx operator+(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ cplx operator-(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__host__ __device__ __forceinline__ cplx operator*(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ __forceinline__ cplx operator*(double s, cplx a) { return cplx{s * a.re, s * a.im}; }
__host__ __device__ __forceinline__ cplx operator*(cplx a, double s) { return cplx{a.re * s, a.im * s}; }
__host__ __device__ __forceinline__ cplx conj(cplx a) { return cplx{a.re, -a.im}; }
__host__ __device__ __forceinline__ double abs2(cplx a) { return a.re * a.re + a.im * a.im; }

// Loads and stores of data that is touched once (workgroup partials, stored
// stage data of a checkpoint segment): non-temporal, so that it streams past
// L2 and the Infinity Cache instead of evicting the working set.
__device__ __forceinline__ cplx load_stream(const cplx* src) {
  return mk(__builtin_nontemporal_load(&src->re), __builtin_nontemporal_load(&src->im));
}
__device__ __forceinline__ void store_stream(cplx* dst, cplx v) {
  __builtin_nontemporal_store(v.re, &dst->re);
  __builtin_nontemporal_store(v.im, &dst->im);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ cplx mul_mi(cplx a) {
  return INV ? cplx{-a.im, a.re} : cplx{a.im, -a.re};
}

// A transform is described by a shape code C = log2(N) + 32 * V.  V selects a
// schedule variant, picked by the autotuner of the propagation launch (V = 0 is
// the default schedule):
//   bit 0: the other LDS padding (one slot per 8 elements where the default is
//          one per 32, and the reverse);
//   bit 1: the stage of radix 2 or 4 (log2 N not a multiple of 3) goes last
//          instead of second.
template <int C>
struct FftShape {
  static constexpr int LOG = C & 31;
  static constexpr int V = C >> 5;
  static constexpr int N = 1 << LOG;
  static constexpr int T = N / 8;                             // lanes per field
  static constexpr int S = LOG / 3 + ((LOG % 3) ? 1 : 0);     // number of stages
  static constexpr int SMALL = (V & 2) ? S - 1 : 1;           // stage of the small radix
  static constexpr int radix(int s) { return (s == SMALL && (LOG % 3)) ? (1 << (LOG % 3)) : 8; }
  static constexpr int before(int s) {                  // product of radices of stages < s
    int p = 1;
    for (int i = 0; i < s; i++) p *= radix(i);
    return p;
  }
  // LDS padding: one slot after every 2^PAD elements.  Measured on the step
  // kernels (N = 512: 5.6 % faster with PAD 5 than 3; N = 1024: the adjoint
  // sweep is 4 % slower with 5, PAD 2 and 4 slower everywhere).
  static constexpr int PAD_DEFAULT = LOG <= 9 ? 5 : 3;
  static constexpr int PAD = (V & 1) ? (PAD_DEFAULT == 5 ? 3 : 5) : PAD_DEFAULT;
  static constexpr int lds_elems = N + (N >> PAD);       // padded LDS footprint (cplx)
};

template <int L>
__device__ __forceinline__ int lpad(int x) { return x + (x >> FftShape<L>::PAD); }

// sin and cos of a large argument (dispersion phases reach 1e7 rad) to about
// one ulp: Cody-Waite reduction by pi/2 split into two doubles, each
// subtraction done with a fused multiply-add, then the library kernel on
// |r| <= pi/4.  Past 2^30 quarter turns the library routine is used as is.
__device__ __forceinline__ void sincos_reduced(double a, double* s, double* c) {
  constexpr double P1 = 1.5707963267948966;     // nearest double to pi/2
  constexpr double P2 = 6.123233995736766e-17;  // pi/2 - P1
  constexpr double TWO_OVER_PI = 0.6366197723675814;
  const double k = rint(a * TWO_OVER_PI);
  if (fabs(k) > 1073741824.0) {
    sincos(a, s, c);
    return;
  }
  const double r = fma(-k, P2, fma(-k, P1, a));
  double sr, cr;
  sincos(r, &sr, &cr);
  switch (((long long)k) & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

// Workgroup barrier ordering LDS only: unlike __syncthreads() it does not wait
// for outstanding global loads, so loads issued before a transform (the next
// z step's pump rows) stay in flight across its LDS exchanges.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Ordering of a transform's LDS exchanges.  WAVE: the field's lanes and the whole
// workgroup are one wave, whose LDS operations the CU performs in the order the
// wave issues them, so the exchange needs neither a barrier nor a wait for its
// writes before its reads; a wavefront-scope fence (no instruction) keeps the
// compiler from moving the accesses across it.  Otherwise a workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void exchange_sync() {
  if constexpr (WAVE)
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront", "local");
  else
    lds_barrier();
}

// In-place R-point DFT of v[o], v[o+1], ..., v[o+R-1]; outputs in natural order.
template <int R, bool INV>
__device__ __forceinline__ void dft_small(cplx* v) {
  if constexpr (R == 2) {
    cplx a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (R == 4) {
    cplx s0 = v[0] + v[2], d0 = v[0] - v[2], s1 = v[1] + v[3], d1 = mul_mi<INV>(v[1] - v[3]);
    v[0] = s0 + s1;
    v[2] = s0 - s1;
    v[1] = d0 + d1;
    v[3] = d0 - d1;
  } else {
    constexpr double r = 0.70710678118654752440;
    cplx u0 = v[0] + v[4], w0 = v[0] - v[4];
    cplx u1 = v[1] + v[5], w1 = v[1] - v[5];
    cplx u2 = v[2] + v[6], w2 = v[2] - v[6];
    cplx u3 = v[3] + v[7], w3 = v[3] - v[7];
    // w_a *= W_8^a (forward W_8 = exp(-i pi/4))
    if (!INV) {
      w1 = cplx{(w1.re + w1.im) * r, (w1.im - w1.re) * r};
      w3 = cplx{(w3.im - w3.re) * r, -(w3.re + w3.im) * r};
    } else {
      w1 = cplx{(w1.re - w1.im) * r, (w1.re + w1.im) * r};
      w3 = cplx{-(w3.re + w3.im) * r, (w3.re - w3.im) * r};
    }
    w2 = mul_mi<INV>(w2);
    cplx a0 = u0 + u2, a1 = u0 - u2, a2 = u1 + u3, a3 = mul_mi<INV>(u1 - u3);
    cplx b0 = w0 + w2, b1 = w0 - w2, b2 = w1 + w3, b3 = mul_mi<INV>(w1 - w3);
    v[0] = a0 + a2;
    v[4] = a0 - a2;
    v[2] = a1 + a3;
    v[6] = a1 - a3;
    v[1] = b0 + b2;
    v[5] = b0 - b2;
    v[3] = b1 + b3;
    v[7] = b1 - b3;
  }
}

template <bool INV>
__device__ __forceinline__ cplx twmul(cplx x, cplx w) {
  return INV ? x * conj(w) : x * w;
}

// Per-lane twiddles, fixed for the whole propagation: tw[s][j*R + c] for the
// twiddled stages s = 0..S-2 (the last stage needs none).
template <int L>
struct FftTwiddles {
  cplx w[FftShape<L>::S > 1 ? FftShape<L>::S - 1 : 1][8];
};

template <int L>
__device__ __forceinline__ void fft_load_twiddles(FftTwiddles<L>& tw, const cplx* __restrict__ table, int t) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, S = F::S;
#pragma unroll
  for (int s = 0; s < S - 1; s++) {
    const int R = F::radix(s), P = F::before(s), M = N / P / R, NB = 8 / R;
#pragma unroll
    for (int j = 0; j < 8; j++) tw.w[s][j] = mk(1.0, 0.0);
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int b = (t + T * j) % M;
#pragma unroll
      for (int c = 1; c < R; c++) tw.w[s][j * R + c] = table[(b * c * P) & (N - 1)];
    }
  }
}

// Twiddle sources for the accessor-based transforms (fft_field1 / fft_field2):
// tw(stage, entry) returns FftTwiddles<L>::w[stage][entry] of this lane.
//  RegTwiddles:   the lane's twiddles held in registers (24 complex at most);
//  TableTwiddles: looked up in the N-entry table on each use, in LDS or in global
//                 memory, for fields whose lanes cannot spare the registers.
template <int L>
struct RegTwiddles {
  FftTwiddles<L> w;
  __device__ __forceinline__ cplx operator()(int s, int idx) const { return w.w[s][idx]; }
};

template <int L>
struct TableTwiddles {
  const cplx* tab;
  int t;
  __device__ __forceinline__ cplx operator()(int s, int idx) const {
    using F = FftShape<L>;
    const int R = F::radix(s), P = F::before(s), M = F::N / P / R;
    const int j = idx / R, c = idx % R;
    const int b = (t + F::T * j) % M;
    return tab[(b * c * P) & (F::N - 1)];
  }
};

// Twiddles of the propagation kernels.  Up to N = 2048 a field's lanes have 512
// registers each (one wave per SIMD) and keep their twiddles in registers.  From
// N = 4096 a field needs 512 or 1024 lanes, i.e. two or four waves per SIMD with
// 256 or 128 registers: the twiddles are looked up in the table instead, an LDS
// copy beside the field's buffer at N = 4096, the global table (through L1/L2)
// at N = 8192, whose buffer fills most of the LDS.  (Measured on a batch of 2000
// SHG / PDC media: the N = 4096 Chi2SHG instances 13 % and the longest Chi2PDC
// instance 10 % faster together with the dispersion row read where it is used;
// the same from N = 2048 was 1 % slower.)
template <int L>
struct PropTw {
  static constexpr bool LEAN = FftShape<L>::LOG >= 12;
  static constexpr bool IN_LDS = FftShape<L>::LOG == 12;
  static constexpr int LDS_ELEMS = IN_LDS ? FftShape<L>::N : 1;
  using type = typename std::conditional<LEAN, TableTwiddles<L>, RegTwiddles<L>>::type;
};

// Fill the workgroup's LDS copy of the twiddle table (all threads; ends with a barrier).
template <int L>
__device__ __forceinline__ void fill_twiddle_lds(cplx* lds_tab, const cplx* table) {
  if constexpr (PropTw<L>::IN_LDS) {
    for (int k = threadIdx.x; k < FftShape<L>::N; k += blockDim.x) lds_tab[k] = table[k];
    __syncthreads();
  }
}

template <int L>
__device__ __forceinline__ void init_twiddles(RegTwiddles<L>& tw, const cplx* table, const cplx*, int t) {
  fft_load_twiddles<L>(tw.w, table, t);
}
template <int L>
__device__ __forceinline__ void init_twiddles(TableTwiddles<L>& tw, const cplx* table, const cplx* lds_tab, int t) {
  tw.tab = PropTw<L>::IN_LDS ? lds_tab : table;
  tw.t = t;
}

// Stages s >= 1 (recursive template so every index is a compile-time shape).
template <int L, bool INV, int s>
__device__ __forceinline__ void fft_stages(cplx (&v)[8], cplx* __restrict__ buf, int t, const FftTwiddles<L>& tw) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, S = F::S;
  constexpr int R = F::radix(s), P = F::before(s), Mp = N / P, M = Mp / R, NB = 8 / R;
  lds_barrier();
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const int beta = t + T * j;
    const int b = beta % M, g = beta / M;
#pragma unroll
    for (int a = 0; a < R; a++) v[j * R + a] = buf[lpad<L>(g * Mp + a * M + b)];
  }
#pragma unroll
  for (int j = 0; j < NB; j++) dft_small<R, INV>(v + j * R);
  if constexpr (s == S - 1) {
    if constexpr (NB > 1) {
      cplx o[8];
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int c = 0; c < R; c++) o[j + NB * c] = v[j * R + c];
#pragma unroll
      for (int m = 0; m < 8; m++) v[m] = o[m];
    }
  } else {
    lds_barrier();
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int beta = t + T * j;
      const int b = beta % M, g = beta / M;
#pragma unroll
      for (int c = 0; c < R; c++) {
        cplx x = c == 0 ? v[j * R] : twmul<INV>(v[j * R + c], tw.w[s][j * R + c]);
        buf[lpad<L>((g + P * c) * M + b)] = x;
      }
    }
    fft_stages<L, INV, s + 1>(v, buf, t, tw);
  }
}

// Transform the field (stride layout in, stride layout out).  `buf` is this
// field's LDS region of FftShape<L>::lds_elems elements.  All lanes of the
// workgroup must call this together (it contains workgroup barriers).
template <int L, bool INV>
__device__ __forceinline__ void fft_field(cplx (&v)[8], cplx* __restrict__ buf, int t, const FftTwiddles<L>& tw) {
  using F = FftShape<L>;
  constexpr int T = F::T, S = F::S;
  dft_small<8, INV>(v);
  if constexpr (S > 1) {
#pragma unroll
    for (int c = 1; c < 8; c++) v[c] = twmul<INV>(v[c], tw.w[0][c]);
    lds_barrier();  // previous readers of buf are done
#pragma unroll
    for (int c = 0; c < 8; c++) buf[lpad<L>(c * T + t)] = v[c];
    fft_stages<L, INV, 1>(v, buf, t, tw);
  }
}

// Two fields transformed together (same direction, separate LDS buffers): the
// exchanges of both share each barrier, so every lane has twice the independent
// arithmetic between two waits.  Same layouts and results as fft_field; the
// twiddles come from an accessor tw(s, idx) returning FftTwiddles::w[s][idx].
template <int L, bool INV, int s, class TW>
__device__ __forceinline__ void fft_stages2(cplx (&v)[8], cplx (&u)[8], cplx* __restrict__ bv, cplx* __restrict__ bu,
                                            int t, const TW& tw) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, S = F::S;
  constexpr int R = F::radix(s), P = F::before(s), Mp = N / P, M = Mp / R, NB = 8 / R;
  lds_barrier();
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const int beta = t + T * j;
    const int b = beta % M, g = beta / M;
#pragma unroll
    for (int a = 0; a < R; a++) {
      v[j * R + a] = bv[lpad<L>(g * Mp + a * M + b)];
      u[j * R + a] = bu[lpad<L>(g * Mp + a * M + b)];
    }
  }
#pragma unroll
  for (int j = 0; j < NB; j++) {
    dft_small<R, INV>(v + j * R);
    dft_small<R, INV>(u + j * R);
  }
  if constexpr (s == S - 1) {
    if constexpr (NB > 1) {
      cplx o[8], q[8];
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int c = 0; c < R; c++) {
          o[j + NB * c] = v[j * R + c];
          q[j + NB * c] = u[j * R + c];
        }
#pragma unroll
      for (int m = 0; m < 8; m++) {
        v[m] = o[m];
        u[m] = q[m];
      }
    }
  } else {
    lds_barrier();
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int beta = t + T * j;
      const int b = beta % M, g = beta / M;
#pragma unroll
      for (int c = 0; c < R; c++) {
        if (c == 0) {
          bv[lpad<L>(g * M + b)] = v[j * R];
          bu[lpad<L>(g * M + b)] = u[j * R];
        } else {
          const cplx w = tw(s, j * R + c);
          bv[lpad<L>((g + P * c) * M + b)] = twmul<INV>(v[j * R + c], w);
          bu[lpad<L>((g + P * c) * M + b)] = twmul<INV>(u[j * R + c], w);
        }
      }
    }
    fft_stages2<L, INV, s + 1>(v, u, bv, bu, t, tw);
  }
}

template <int L, bool INV, class TW>
__device__ __forceinline__ void fft_field2(cplx (&v)[8], cplx (&u)[8], cplx* __restrict__ bv, cplx* __restrict__ bu,
                                           int t, const TW& tw) {
  using F = FftShape<L>;
  constexpr int T = F::T, S = F::S;
  dft_small<8, INV>(v);
  dft_small<8, INV>(u);
  if constexpr (S > 1) {
#pragma unroll
    for (int c = 1; c < 8; c++) {
      const cplx w = tw(0, c);
      v[c] = twmul<INV>(v[c], w);
      u[c] = twmul<INV>(u[c], w);
    }
    lds_barrier();
#pragma unroll
    for (int c = 0; c < 8; c++) {
      bv[lpad<L>(c * T + t)] = v[c];
      bu[lpad<L>(c * T + t)] = u[c];
    }
    fft_stages2<L, INV, 1>(v, u, bv, bu, t, tw);
  }
}

// One field with twiddles from tw(s, idx) (see fft_field).
template <int L, bool INV, int s, bool WAVE = false, class TW>
__device__ __forceinline__ void fft_stages1(cplx (&v)[8], cplx* __restrict__ buf, int t, const TW& tw) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T, S = F::S;
  constexpr int R = F::radix(s), P = F::before(s), Mp = N / P, M = Mp / R, NB = 8 / R;
  exchange_sync<WAVE>();
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const int beta = t + T * j;
    const int b = beta % M, g = beta / M;
#pragma unroll
    for (int a = 0; a < R; a++) v[j * R + a] = buf[lpad<L>(g * Mp + a * M + b)];
  }
#pragma unroll
  for (int j = 0; j < NB; j++) dft_small<R, INV>(v + j * R);
  if constexpr (s == S - 1) {
    if constexpr (NB > 1) {
      cplx o[8];
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int c = 0; c < R; c++) o[j + NB * c] = v[j * R + c];
#pragma unroll
      for (int m = 0; m < 8; m++) v[m] = o[m];
    }
  } else {
    exchange_sync<WAVE>();
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int beta = t + T * j;
      const int b = beta % M, g = beta / M;
#pragma unroll
      for (int c = 0; c < R; c++)
        buf[lpad<L>((g + P * c) * M + b)] = c == 0 ? v[j * R] : twmul<INV>(v[j * R + c], tw(s, j * R + c));
    }
    fft_stages1<L, INV, s + 1, WAVE>(v, buf, t, tw);
  }
}

template <int L, bool INV, bool WAVE = false, class TW>
__device__ __forceinline__ void fft_field1(cplx (&v)[8], cplx* __restrict__ buf, int t, const TW& tw) {
  using F = FftShape<L>;
  constexpr int T = F::T, S = F::S;
  dft_small<8, INV>(v);
  if constexpr (S > 1) {
#pragma unroll
    for (int c = 1; c < 8; c++) v[c] = twmul<INV>(v[c], tw(0, c));
    exchange_sync<WAVE>();
#pragma unroll
    for (int c = 0; c < 8; c++) buf[lpad<L>(c * T + t)] = v[c];
    fft_stages1<L, INV, 1, WAVE>(v, buf, t, tw);
  }
}
