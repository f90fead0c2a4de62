// nlm_any.hip -- the split-step engine for every even nFreqs that the in-register
// power-of-two kernels do not take: lengths that are not powers of two (the
// reference accepts any even nFreqs, src/_NonlinearMedium.cpp:108-109, with
// kissfft's mixed-radix transform, src/CustomEigenFFT.h:17-40), lengths 2 and
// 4, and lengths above 8192.
//
// One workgroup of 256 lanes owns one field at a time (a persistent grid walks
// the fields).  The field's modes and a second buffer per mode live in LDS when
// they fit in 64 KB, otherwise in a per-workgroup slice of a device scratch
// buffer (served from L2).  Transforms are self-sorting mixed-radix Stockham
// passes, radices 8, 4, 2 and then the odd prime factors, one output element
// per lane and pass:
//   out[(j / Ns) Ns R + (j mod Ns) + k Ns] = sum_r in[j + r N/R] W_N^(r ((j mod Ns) N/(Ns R) + k N/R))
// with every twiddle read from the table W_N^m, m < N (each product uses one
// correctly rounded table entry).  Conventions follow src/CustomEigenFFT.h:54-137:
// forward unscaled, inverse scaled by 1/N.
//
// Replaced here for these lengths: signalSimulationTemplate
// (src/_NonlinearMedium.hpp:160-239) with every solver DiffEq (Chi3GNLSE's
// transforms inside each RK stage included, solvers/Chi3GNLSE.cpp:51-92), the
// linear pump (src/_NonlinearMedium.cpp:225-250), Chi3::runPumpSimulation
// (solvers/Chi3.cpp:25-54), setPump(other) (src/_NonlinearMedium.cpp:506-534)
// and the transforms of nlm_fft.
#include <algorithm>
#include <cstring>
#include <string>

#include "nlm_internal.hpp"

extern __shared__ cplx any_lds[];

namespace {

constexpr int ANY_THREADS = 256;
constexpr int ANY_MAX_PASSES = 40;
constexpr size_t ANY_LDS_BYTES = 64 * 1024;

struct AnyFft {
  int N, nPass;
  int radix[ANY_MAX_PASSES];
  const cplx* tw;  // W_N^m = exp(-2 pi i m / N)
  double invN;
};

// Transform the N points in `src` through Stockham passes alternating between
// `src` and `dst`; returns the buffer that holds the result.  The workgroup
// synchronizes before the first pass and after every pass.
template <bool INV>
__device__ cplx* wg_fft(const AnyFft& f, cplx* src, cplx* dst) {
  const int N = f.N;
  int Ns = 1;
  __syncthreads();
  for (int p = 0; p < f.nPass; p++) {
    const int R = f.radix[p], M = N / R, twStride = N / (Ns * R);
    for (int o = threadIdx.x; o < N; o += blockDim.x) {
      const int jm = o % Ns, k = (o / Ns) % R, jq = o / (Ns * R);
      const int j = jq * Ns + jm;
      const int step = jm * twStride + k * M;  // < N / R + N: reduced once below
      const int st = step >= N ? step - N : step;
      cplx acc = src[j];
      int e = 0;
      for (int r = 1; r < R; r++) {
        e += st;
        if (e >= N) e -= N;
        const cplx w = f.tw[e];
        acc = acc + src[j + r * M] * (INV ? conj(w) : w);
      }
      dst[o] = acc;
    }
    __syncthreads();
    cplx* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

// Forward or scaled inverse transform of one mode's buffer pair (cur holds the
// data on entry and on exit).
template <bool INV>
__device__ void mode_fft(const AnyFft& f, cplx*& cur, cplx*& oth) {
  cplx* res = wg_fft<INV>(f, cur, oth);
  if (res != cur) {
    oth = cur;
    cur = res;
  }
  if (INV)
    for (int k = threadIdx.x; k < f.N; k += blockDim.x) cur[k] = cur[k] * f.invN;
}

__device__ void store_any(cplx* __restrict__ dst, const cplx* src, int N) {
  for (int k = threadIdx.x; k < N; k += blockDim.x) dst[k] = src[k];
}

__device__ void mul_any(cplx* v, const cplx* __restrict__ tab, int N) {
  for (int k = threadIdx.x; k < N; k += blockDim.x) v[k] = v[k] * tab[k];
}

__device__ cplx* any_row(const KParams& kp, cplx* base, int field, int nS, int q, int i, int N) {
  return base + (((size_t)field * nS + q) * kp.nZ + i) * N;
}

// Inputs and the half dispersion step in (src/_NonlinearMedium.hpp:164-198,
// src/_NonlinearMedium.cpp:329-340): leaves the time-domain field in cur[q]
// (frequency domain when nZ == 1) and writes the grids' row 0.  Pointwise phases
// use the lane-to-point mapping k = threadIdx.x + blockDim.x * j throughout, so
// they need no barrier between them; every transform synchronizes on entry and exit.
template <int NS>
__device__ void any_prologue(const KParams& kp, const AnyFft& f, int field, cplx* (&cur)[NS], cplx* (&oth)[NS]) {
  const int N = f.N, tid = threadIdx.x, nt = blockDim.x;
  __syncthreads();  // the previous field's readers are done
  for (int q = 0; q < NS; q++) {
    if (kp.kind == KIND_GREENS) {
      // field 2i + w: impulse e_(i mod N) (w = 0) or i e_(i mod N) (w = 1) into mode inputs[i / N]
      const int i = field >> 1, w = field & 1, pos = q == kp.inputs[i / N] ? i % N : -1;
      const cplx val = w ? mk(0.0, 1.0) : mk(1.0, 0.0);
      for (int k = tid; k < N; k += nt) cur[q][k] = k == pos ? val : mk(0.0, 0.0);
    } else if (kp.inPos[q] >= 0) {
      const cplx* row = kp.in + (size_t)field * kp.inCols + (size_t)kp.inPos[q] * N;
      for (int k = tid; k < N; k += nt) cur[q][k] = row[k];
    } else {
      for (int k = tid; k < N; k += nt) cur[q][k] = mk(0.0, 0.0);
    }
  }
  for (int q = 0; q < NS; q++) {
    if (kp.inTime) mode_fft<false>(f, cur[q], oth[q]);
    mul_any(cur[q], kp.halfIn + (size_t)q * N, N);
    if (kp.kind == KIND_GRID) store_any(any_row(kp, kp.gridF, field, NS, q, 0, N), cur[q], N);
  }
  if (kp.nZ > 1)
    for (int q = 0; q < NS; q++) {
      mode_fft<true>(f, cur[q], oth[q]);
      if (kp.kind == KIND_GRID) store_any(any_row(kp, kp.gridT, field, NS, q, 0, N), cur[q], N);
    }
}

// Dispersion step i of mode q (src/_NonlinearMedium.hpp:212-215) with the grids' row i.
__device__ void any_dispersion(const KParams& kp, const AnyFft& f, int field, int NS, int q, int i, cplx*& cur,
                               cplx*& oth) {
  const int N = f.N;
  mode_fft<false>(f, cur, oth);
  mul_any(cur, kp.dstep + (size_t)q * N, N);
  if (i < kp.nZ - 1) {
    if (kp.kind == KIND_GRID) store_any(any_row(kp, kp.gridF, field, NS, q, i, N), cur, N);
    mode_fft<true>(f, cur, oth);
    if (kp.kind == KIND_GRID) store_any(any_row(kp, kp.gridT, field, NS, q, i, N), cur, N);
  }
}

// Half step out and outputs (src/_NonlinearMedium.hpp:235-238, src/_NonlinearMedium.cpp:333-345, 428-431).
template <int NS>
__device__ void any_epilogue(const KParams& kp, const AnyFft& f, int field, cplx* (&cur)[NS], cplx* (&oth)[NS]) {
  const int N = f.N, last = kp.nZ - 1;
  for (int q = 0; q < NS; q++) {
    mul_any(cur[q], kp.halfOut + (size_t)q * N, N);
    if (kp.kind == KIND_GRID) {
      store_any(any_row(kp, kp.gridF, field, NS, q, last, N), cur[q], N);
      mode_fft<true>(f, cur[q], oth[q]);
      store_any(any_row(kp, kp.gridT, field, NS, q, last, N), cur[q], N);
    } else {
      if (kp.outTime) mode_fft<true>(f, cur[q], oth[q]);
      if (kp.outPos[q] >= 0)
        store_any(kp.out + (size_t)field * kp.nOut * N + (size_t)kp.outPos[q] * N, cur[q], N);
    }
  }
}

// One propagation of field `field` (signalSimulationTemplate) for a pointwise DiffEq.
template <class M>
__device__ void any_field(const KParams& kp, const AnyFft& f, int field, cplx* (&cur)[M::NS], cplx* (&oth)[M::NS]) {
  constexpr int NS = M::NS, NP = M::NP, NC = M::NC;
  constexpr int NPP = NP > 0 ? NP : 1;
  const int N = f.N, tid = threadIdx.x, nt = blockDim.x;
  any_prologue<NS>(kp, f, field, cur, oth);
  const cplx* pumps[2] = {kp.pump0, kp.pump1};
  for (int i = 1; i < kp.nZ; i++) {
    // RK4 nonlinear step (src/_NonlinearMedium.hpp:207-210), pointwise
    const cplx* c = kp.coef + (size_t)i * 3 * NC;
    for (int k = tid; k < N; k += nt) {
      cplx x[NS], y[NS], k1[NS], k2[NS], k3[NS], k4[NS], P0[NPP], P1[NPP], P2[NPP];
      for (int r = 0; r < NPP; r++) {
        P0[r] = NP ? pumps[r][(size_t)(2 * i - 2) * N + k] : mk(0, 0);
        P1[r] = NP ? pumps[r][(size_t)(2 * i - 1) * N + k] : mk(0, 0);
        P2[r] = NP ? pumps[r][(size_t)(2 * i) * N + k] : mk(0, 0);
      }
      for (int q = 0; q < NS; q++) x[q] = cur[q][k];
      M::template f<0>(c, kp.cconst, x, P0, k1);
      for (int q = 0; q < NS; q++) y[q] = x[q] + 0.5 * k1[q];
      M::template f<1>(c + NC, kp.cconst, y, P1, k2);
      for (int q = 0; q < NS; q++) y[q] = x[q] + 0.5 * k2[q];
      M::template f<2>(c + NC, kp.cconst, y, P1, k3);
      for (int q = 0; q < NS; q++) y[q] = x[q] + k3[q];
      M::template f<3>(c + 2 * NC, kp.cconst, y, P2, k4);
      for (int q = 0; q < NS; q++) cur[q][k] = x[q] + (((k1[q] + 2.0 * k2[q]) + 2.0 * k3[q]) + k4[q]) * (1. / 6.);
    }
    for (int q = 0; q < NS; q++) any_dispersion(kp, f, field, NS, q, i, cur[q], oth[q]);
  }
  any_epilogue<NS>(kp, f, field, cur, oth);
}

// Chi3GNLSE stage increment (solvers/Chi3GNLSE.cpp:51-92):
// k = IFFT( FFT( IFFT(FFT(r |y|^2) * R) * y ) * (nl0 + nl1*omega) ), into K (KO its
// ping-pong partner); y is not modified.
__device__ void any_gnlse_stage(const KParams& kp, const AnyFft& f, const cplx* y, double r, cplx*& K, cplx*& KO) {
  const int N = f.N;
  for (int k = threadIdx.x; k < N; k += blockDim.x) K[k] = mk(r * abs2(y[k]), 0.0);
  mode_fft<false>(f, K, KO);
  mul_any(K, kp.raman, N);
  mode_fft<true>(f, K, KO);
  for (int k = threadIdx.x; k < N; k += blockDim.x) K[k] = K[k] * y[k];
  mode_fft<false>(f, K, KO);
  mul_any(K, kp.filt, N);
  mode_fft<true>(f, K, KO);
}

// One Chi3GNLSE propagation; workspace a, its transform partner, y, acc, k, k's partner.
__device__ void any_gnlse_field(const KParams& kp, const AnyFft& f, int field, cplx* ws) {
  const int N = f.N, tid = threadIdx.x, nt = blockDim.x;
  cplx* cur[1] = {ws};
  cplx* oth[1] = {ws + N};
  cplx* Y = ws + 2 * (size_t)N;
  cplx* acc = ws + 3 * (size_t)N;
  cplx* K = ws + 4 * (size_t)N;
  cplx* KO = ws + 5 * (size_t)N;
  any_prologue<1>(kp, f, field, cur, oth);
  for (int i = 1; i < kp.nZ; i++) {
    const cplx* c = kp.coef + (size_t)i * 3;
    const double r0 = c[0].re, r1 = c[1].re, r2 = c[2].re;
    cplx* A = cur[0];
    // acc accumulates ((k1 + 2 k2) + 2 k3) + k4 in the reference's order
    any_gnlse_stage(kp, f, A, r0, K, KO);
    for (int k = tid; k < N; k += nt) {
      acc[k] = K[k];
      Y[k] = A[k] + 0.5 * K[k];
    }
    any_gnlse_stage(kp, f, Y, r1, K, KO);
    __syncthreads();  // every lane has read Y before it is rewritten
    for (int k = tid; k < N; k += nt) {
      acc[k] = acc[k] + 2.0 * K[k];
      Y[k] = A[k] + 0.5 * K[k];
    }
    any_gnlse_stage(kp, f, Y, r1, K, KO);
    __syncthreads();
    for (int k = tid; k < N; k += nt) {
      acc[k] = acc[k] + 2.0 * K[k];
      Y[k] = A[k] + K[k];
    }
    any_gnlse_stage(kp, f, Y, r2, K, KO);
    for (int k = tid; k < N; k += nt) A[k] = A[k] + (acc[k] + K[k]) * (1. / 6.);
    any_dispersion(kp, f, field, 1, 0, i, cur[0], oth[0]);
  }
  any_epilogue<1>(kp, f, field, cur, oth);
}

// Workspace of one workgroup: 2 NS buffers of N points, in LDS (dynamic shared
// memory) or at gws + blockIdx.x * 2 NS N.
__device__ cplx* any_workspace(cplx* gws, int NS, int N) {
  return gws ? gws + (size_t)blockIdx.x * 2 * NS * N : any_lds;
}

template <class M>
__global__ void __launch_bounds__(ANY_THREADS) k_any_prop(KParams kp, AnyFft f, cplx* gws) {
  constexpr int NS = M::NS;
  cplx* ws = any_workspace(gws, NS, f.N);
  for (int field = blockIdx.x; field < kp.nFields; field += gridDim.x) {
    cplx* cur[NS];
    cplx* oth[NS];
    for (int q = 0; q < NS; q++) {
      cur[q] = ws + (size_t)(2 * q) * f.N;
      oth[q] = ws + (size_t)(2 * q + 1) * f.N;
    }
    any_field<M>(kp, f, field, cur, oth);
  }
}

__global__ void __launch_bounds__(ANY_THREADS) k_any_gnlse(KParams kp, AnyFft f, cplx* gws) {
  cplx* ws = any_workspace(gws, 3, f.N);
  for (int field = blockIdx.x; field < kp.nFields; field += gridDim.x) any_gnlse_field(kp, f, field, ws);
}

// Rows of `count` transforms (nlm_fft; forward unscaled, inverse scaled by 1/N).
__global__ void __launch_bounds__(ANY_THREADS)
    k_any_fft(const cplx* __restrict__ in, cplx* __restrict__ out, int count, int inverse, AnyFft f, cplx* gws) {
  cplx* ws = any_workspace(gws, 1, f.N);
  for (int row = blockIdx.x; row < count; row += gridDim.x) {
    __syncthreads();
    cplx *cur = ws, *oth = ws + f.N;
    for (int k = threadIdx.x; k < f.N; k += blockDim.x) cur[k] = in[(size_t)row * f.N + k];
    if (inverse) mode_fft<true>(f, cur, oth);
    else mode_fft<false>(f, cur, oth);
    store_any(out + (size_t)row * f.N, cur, f.N);
  }
}

// Linear pump rows (src/_NonlinearMedium.cpp:225-250): row i: freq = F0 exp(i (i dzp) D_p),
// time = IFFT(freq), row 0's time row the envelope itself; both times strength[i].
__global__ void __launch_bounds__(ANY_THREADS)
    k_any_pump_linear(const cplx* __restrict__ F0, const cplx* __restrict__ env, const double* __restrict__ Dp,
                      double dzp, const double* __restrict__ strength, int nZP, cplx* __restrict__ pumpF,
                      cplx* __restrict__ pumpT, AnyFft f, cplx* gws) {
  cplx* ws = any_workspace(gws, 1, f.N);
  const int N = f.N;
  for (int i = blockIdx.x; i < nZP; i += gridDim.x) {
    __syncthreads();
    cplx *cur = ws, *oth = ws + N;
    const double s = strength ? strength[i] : 1.0;
    const double zi = (double)i * dzp;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      cplx v = F0[k];
      if (i > 0) {
        double sn, cs;
        sincos_reduced(zi * Dp[k], &sn, &cs);
        v = v * mk(cs, sn);
      }
      cur[k] = v;
      pumpF[(size_t)i * N + k] = strength ? v * s : v;
    }
    mode_fft<true>(f, cur, oth);
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      const cplx v = i == 0 ? env[k] : cur[k];
      pumpT[(size_t)i * N + k] = strength ? v * s : v;
    }
  }
}

// Chi3 pump (solvers/Chi3.cpp:25-54), serial over the 2nZ-1 rows in one workgroup.
__global__ void __launch_bounds__(ANY_THREADS)
    k_any_pump_chi3(const cplx* __restrict__ env, const double* __restrict__ Dp, double dzp,
                    const cplx* __restrict__ dstepP, const cplx* __restrict__ relI, int nZP, cplx* __restrict__ pumpF,
                    cplx* __restrict__ pumpT, AnyFft f, cplx* gws) {
  cplx* ws = any_workspace(gws, 1, f.N);
  const int N = f.N;
  cplx *cur = ws, *oth = ws + N;
  for (int k = threadIdx.x; k < N; k += blockDim.x) cur[k] = env[k];
  mode_fft<false>(f, cur, oth);
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    double s, c;
    sincos((0.5 * dzp) * Dp[k], &s, &c);
    cur[k] = cur[k] * mk(c, s);
    pumpF[k] = cur[k];
  }
  mode_fft<true>(f, cur, oth);
  store_any(pumpT, cur, N);
  for (int i = 1; i < nZP; i++) {
    const cplx r = relI[i - 1];
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      const double a2 = abs2(cur[k]);
      const double e = exp(r.re * a2);
      double s, c;
      sincos(r.im * a2, &s, &c);
      cur[k] = cur[k] * mk(e * c, e * s);
    }
    mode_fft<false>(f, cur, oth);
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      cplx v = cur[k] * dstepP[k];
      if (i == nZP - 1) {
        double s, c;
        sincos((-0.5 * dzp) * Dp[k], &s, &c);
        v = v * mk(c, s);
      }
      cur[k] = v;
      pumpF[(size_t)i * N + k] = v;
    }
    mode_fft<true>(f, cur, oth);
    store_any(pumpT + (size_t)i * N, cur, N);
  }
}

// Pump rows from another medium's frequency grid, setPump(other, modeIndex, delay)
// (src/_NonlinearMedium.cpp:506-534); see k_pump_import for the row formula.
__global__ void __launch_bounds__(ANY_THREADS)
    k_any_pump_import(const cplx* __restrict__ sigF, int nzs, double dzs, const double* __restrict__ Ds,
                      const double* __restrict__ Dp, const double* __restrict__ omega, double dzp, double bd, double z,
                      double zs, int nZP, cplx* __restrict__ pumpF, cplx* __restrict__ pumpT, AnyFft f, cplx* gws) {
  cplx* ws = any_workspace(gws, 1, f.N);
  const int N = f.N;
  for (int i = blockIdx.x; i < nZP; i += gridDim.x) {
    __syncthreads();
    cplx *cur = ws, *oth = ws + N;
    const bool lastRow = i == nZP - 1;
    const double j_ = i * ((double)(nzs - 1) / (nZP - 1));
    const int j = (int)j_;
    const double frac = j_ - j;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
      cplx v;
      double ph;
      if (!lastRow) {
        ph = ((double)i * dzp) * Dp[k] - ((j_ + 0.5) * dzs) * Ds[k] + bd * omega[k];
        v = (1 - frac) * sigF[(size_t)j * N + k] + frac * sigF[(size_t)(j + 1) * N + k];
      } else {
        ph = z * Dp[k] - zs * Ds[k] + bd * omega[k];
        v = sigF[(size_t)(nzs - 1) * N + k];
      }
      double sn, cs;
      sincos_reduced(ph, &sn, &cs);
      v = v * mk(cs, sn);
      cur[k] = v;
      pumpF[(size_t)i * N + k] = v;
    }
    mode_fft<true>(f, cur, oth);
    if (!lastRow) store_any(pumpT + (size_t)i * N, cur, N);
  }
}

// Radices of N: 8, 4 and 2 first, then the odd primes in increasing order.
int any_plan(int N, AnyFft* f) {
  memset(f, 0, sizeof *f);
  if (N < 2 || N % 2) return set_err(NLM_INVALID_ARGUMENT, "Invalid number of Frequencies");
  f->N = N;
  f->invN = 1.0 / N;
  int n = N;
  auto add = [&](int r) {
    if (f->nPass >= ANY_MAX_PASSES) return false;
    f->radix[f->nPass++] = r;
    n /= r;
    return true;
  };
  while (n % 8 == 0) add(8);
  if (n % 4 == 0) add(4);
  if (n % 2 == 0) add(2);
  for (int p = 3; n > 1 && (long)p * p <= n; p += 2)
    while (n % p == 0) add(p);
  if (n > 1) add(n);
  if (int e = get_twiddles(N, &f->tw)) return e;
  return NLM_OK;
}

// Launch shape: a persistent grid over `items`, LDS workspaces when 2 NS N
// points fit in 64 KB, else a device scratch of grid x 2 NS N points held by the
// plan (or by `own` when there is no plan).
struct AnyLaunch {
  int grid;
  size_t lds;
  cplx* gws;
};

int any_launch(nlm_plan_s* p, DevBuf<cplx>* own, int NS, int N, long items, AnyLaunch* out, bool may_alloc = true) {
  const size_t bytes = (size_t)2 * NS * N * sizeof(cplx);
  int cus = 256, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus <= 0) cus = 256;
  if (bytes <= ANY_LDS_BYTES) {
    const long per = std::max<long>(1, (long)(160 * 1024 / bytes));
    out->grid = (int)std::max<long>(1, std::min<long>(items, cus * std::min<long>(per, 8)));
    out->lds = bytes;
    out->gws = nullptr;
    return NLM_OK;
  }
  out->grid = (int)std::max<long>(1, std::min<long>(items, 4L * cus));
  out->lds = 0;
  DevBuf<cplx>& buf = p ? p->anyScratch : *own;
  if (!may_alloc && (!buf.p || buf.n < (size_t)out->grid * 2 * NS * N))
    return set_err(NLM_INVALID_ARGUMENT, "nFreqs " + std::to_string(N) +
                                             " needs a device workspace beyond LDS; use the allocating entry point");
  if (int e = buf.ensure((size_t)out->grid * 2 * NS * N)) return e;
  out->gws = buf.p;
  return NLM_OK;
}

typedef void (*any_prop_fn)(const KParams&, const AnyFft&, const AnyLaunch&, hipStream_t);

template <class M>
void launch_any_prop(const KParams& kp, const AnyFft& f, const AnyLaunch& al, hipStream_t s) {
  hipLaunchKernelGGL(k_any_prop<M>, dim3(al.grid), dim3(ANY_THREADS), al.lds, s, kp, f, al.gws);
}

void launch_any_gnlse(const KParams& kp, const AnyFft& f, const AnyLaunch& al, hipStream_t s) {
  hipLaunchKernelGGL(k_any_gnlse, dim3(al.grid), dim3(ANY_THREADS), al.lds, s, kp, f, al.gws);
}

template <int MODEL>
any_prop_fn any_prop_of() {
  if constexpr (MODEL == NLM_CHI3GNLSE) return &launch_any_gnlse;
  else return &launch_any_prop<typename ModelOf<MODEL>::type>;
}

any_prop_fn pick_any_prop(int model) {
  switch (model) {
    case NLM_CHI3: return any_prop_of<NLM_CHI3>();
    case NLM_CHI2PDC: return any_prop_of<NLM_CHI2PDC>();
    case NLM_CHI2SFG: return any_prop_of<NLM_CHI2SFG>();
    case NLM_CHI2AFC: return any_prop_of<NLM_CHI2AFC>();
    case NLM_CHI2SFGPDC: return any_prop_of<NLM_CHI2SFGPDC>();
    case NLM_CHI2SFGII: return any_prop_of<NLM_CHI2SFGII>();
    case NLM_CHI2PDCII: return any_prop_of<NLM_CHI2PDCII>();
    case NLM_CHI2SFGOPA: return any_prop_of<NLM_CHI2SFGOPA>();
    case NLM_CHI2SFGXPM: return any_prop_of<NLM_CHI2SFGXPM>();
    case NLM_CHI2SHG: return any_prop_of<NLM_CHI2SHG>();
    case NLM_CHI2SHGOPA: return any_prop_of<NLM_CHI2SHGOPA>();
    case NLM_CHI2DSFG: return any_prop_of<NLM_CHI2DSFG>();
    case NLM_CHI2SHGXPM: return any_prop_of<NLM_CHI2SHGXPM>();
    case NLM_CHI2ASHG: return any_prop_of<NLM_CHI2ASHG>();
    case NLM_CHI3GNLSE: return any_prop_of<NLM_CHI3GNLSE>();
  }
  return nullptr;
}

}  // namespace

bool nlm_any_size(int n) { return n >= 2 && n % 2 == 0 && (n < 8 || n > 8192 || (n & (n - 1)) != 0); }

int nlm_any_prop(nlm_plan_s* p, const KParams& kp, bool may_alloc) {
  const any_prop_fn fn = pick_any_prop(p->model);
  if (!fn) return set_err(NLM_INVALID_ARGUMENT, "unsupported model");
  if (kp.nFields <= 0) return NLM_OK;
  AnyFft f;
  if (int e = any_plan(p->N, &f)) return e;
  AnyLaunch al;
  // a Chi3GNLSE field needs six buffers (the state, y, the increment sum, the
  // increment and two transform partners), a pointwise model two per mode
  const int bufPairs = p->model == NLM_CHI3GNLSE ? 3 : p->nS;
  if (int e = any_launch(p, nullptr, bufPairs, p->N, kp.nFields, &al, may_alloc)) return e;
  fn(kp, f, al, p->stream);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}

int nlm_any_pump_linear(nlm_plan_s* p, const double* strength_dev, double dzp) {
  const int N = p->N, nZP = 2 * p->nZ - 1;
  AnyFft f;
  if (int e = any_plan(N, &f)) return e;
  AnyLaunch al;
  if (int e = any_launch(p, nullptr, 1, N, std::max(nZP, p->nP), &al)) return e;
  hipLaunchKernelGGL(k_any_fft, dim3(std::min(al.grid, p->nP)), dim3(ANY_THREADS), al.lds, p->stream, p->env.p,
                     p->F0.p, p->nP, 0, f, al.gws);
  for (int q = 0; q < p->nP; q++)
    hipLaunchKernelGGL(k_any_pump_linear, dim3(std::min(al.grid, nZP)), dim3(ANY_THREADS), al.lds, p->stream,
                       p->F0.p + (size_t)q * N, p->env.p + (size_t)q * N, p->Dp.p + (size_t)q * N, dzp, strength_dev,
                       nZP, p->pumpF[q].p, p->pumpT[q].p, f, al.gws);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}

int nlm_any_pump_chi3(nlm_plan_s* p, double dzp) {
  const int N = p->N, nZP = 2 * p->nZ - 1;
  AnyFft f;
  if (int e = any_plan(N, &f)) return e;
  AnyLaunch al;
  if (int e = any_launch(p, nullptr, 1, N, 1, &al)) return e;
  hipLaunchKernelGGL(k_any_pump_chi3, dim3(1), dim3(ANY_THREADS), al.lds, p->stream, p->env.p, p->Dp.p, dzp, p->dsp.p,
                     p->relI.p, nZP, p->pumpF[0].p, p->pumpT[0].p, f, al.gws);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}

int nlm_any_pump_import(nlm_plan_s* p, int q, const cplx* sigF, int nzs, double dzs, const double* Ds,
                        const double* Dp, const double* omega, double dzp, double bd, double z, double zs) {
  const int N = p->N, nZP = 2 * p->nZ - 1;
  AnyFft f;
  if (int e = any_plan(N, &f)) return e;
  AnyLaunch al;
  if (int e = any_launch(p, nullptr, 1, N, nZP, &al)) return e;
  hipLaunchKernelGGL(k_any_pump_import, dim3(al.grid), dim3(ANY_THREADS), al.lds, p->stream, sigF, nzs, dzs, Ds, Dp,
                     omega, dzp, bd, z, zs, nZP, p->pumpF[q].p, p->pumpT[q].p, f, al.gws);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}

int nlm_any_fft(const cplx* in, cplx* out, int n, int count, bool inverse, hipStream_t s, DevBuf<cplx>* scratch) {
  AnyFft f;
  if (int e = any_plan(n, &f)) return e;
  if (count <= 0) return NLM_OK;
  AnyLaunch al;
  if (int e = any_launch(nullptr, scratch, 1, n, count, &al)) return e;
  hipLaunchKernelGGL(k_any_fft, dim3(al.grid), dim3(ANY_THREADS), al.lds, s, in, out, count, inverse ? 1 : 0, f,
                     al.gws);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}
