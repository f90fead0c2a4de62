// nlm_compose.hip -- Green's function composition of two media on the FP64
// matrix cores (Cascade::computeGreensFunction, src/Cascade.cpp:128-129):
//
//   C = C2 C1 + S2 conj(S1),   S = C2 S1 + S2 conj(C1)
//
// with C2, S2: M x K and C1, S1: K x Nc complex (row-major, interleaved).  In
// real arithmetic each of Re C, Im C, Re S, Im S is a sum of four real
// products of the real / imaginary parts, e.g.
//   Re C = Re C2 Re C1 - Im C2 Im C1 + Re S2 Re S1 + Im S2 Im S1,
// so one pass over K feeds 16 real 16x16x4 FP64 MFMAs per wave and k-step from
// four A fragments (Re/Im of C2, S2, two of them negated) and four B fragments.
// A workgroup of four waves computes a 32 x 32 output tile; the K dimension is
// staged through LDS in blocks of KB, de-interleaved into real and imaginary
// planes.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/nlm.h"
#include "nlm_internal.hpp"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int TILE = 32;  // output tile per workgroup (2 x 2 waves of 16 x 16)
constexpr int KB = 32;    // K block staged in LDS
constexpr int APAD = KB + 1, BPAD = TILE + 1;

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One output tile.  MFMA operand layout (16x16x4, FP64): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register r of lane l
// holds D[l / 16 + 4 r][l % 16].
__global__ void __launch_bounds__(256) k_compose_mfma(const cplx* __restrict__ C2, const cplx* __restrict__ S2,
                                                      const cplx* __restrict__ C1, const cplx* __restrict__ S1,
                                                      cplx* __restrict__ C, cplx* __restrict__ S, int M, int K,
                                                      int Nc) {
  // A planes: 0 Re C2, 1 Im C2, 2 Re S2, 3 Im S2; B planes: 0 Re C1, 1 Im C1, 2 Re S1, 3 Im S1
  __shared__ double As[4][TILE][APAD];
  __shared__ double Bs[4][KB][BPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row0 = blockIdx.y * TILE, col0 = blockIdx.x * TILE;
  const int wr = (wave >> 1) * 16, wc = (wave & 1) * 16;
  v4d cr = {0, 0, 0, 0}, ci = cr, sr = cr, si = cr;

  for (int k0 = 0; k0 < K; k0 += KB) {
    // stage A: TILE rows x KB of C2, S2; B: KB rows x TILE of C1, S1 (zeros outside)
    for (int e = tid; e < TILE * KB; e += 256) {
      const int r = e / KB, k = e % KB;
      const int gr = row0 + r, gk = k0 + k;
      cplx c2 = mk(0, 0), s2 = mk(0, 0);
      if (gr < M && gk < K) {
        c2 = C2[(size_t)gr * K + gk];
        s2 = S2[(size_t)gr * K + gk];
      }
      As[0][r][k] = c2.re;
      As[1][r][k] = c2.im;
      As[2][r][k] = s2.re;
      As[3][r][k] = s2.im;
    }
    for (int e = tid; e < KB * TILE; e += 256) {
      const int k = e / TILE, c = e % TILE;
      const int gk = k0 + k, gc = col0 + c;
      cplx c1 = mk(0, 0), s1 = mk(0, 0);
      if (gk < K && gc < Nc) {
        c1 = C1[(size_t)gk * Nc + gc];
        s1 = S1[(size_t)gk * Nc + gc];
      }
      Bs[0][k][c] = c1.re;
      Bs[1][k][c] = c1.im;
      Bs[2][k][c] = s1.re;
      Bs[3][k][c] = s1.im;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      const int ar = wr + (lane & 15), ak = kk + (lane >> 4);
      const int bk = kk + (lane >> 4), bc = wc + (lane & 15);
      const double c2r = As[0][ar][ak], c2i = As[1][ar][ak], s2r = As[2][ar][ak], s2i = As[3][ar][ak];
      const double c1r = Bs[0][bk][bc], c1i = Bs[1][bk][bc], s1r = Bs[2][bk][bc], s1i = Bs[3][bk][bc];
      // Re C = C2r C1r - C2i C1i + S2r S1r + S2i S1i
      cr = mfma(c2r, c1r, cr);
      cr = mfma(-c2i, c1i, cr);
      cr = mfma(s2r, s1r, cr);
      cr = mfma(s2i, s1i, cr);
      // Im C = C2r C1i + C2i C1r - S2r S1i + S2i S1r
      ci = mfma(c2r, c1i, ci);
      ci = mfma(c2i, c1r, ci);
      ci = mfma(-s2r, s1i, ci);
      ci = mfma(s2i, s1r, ci);
      // Re S = C2r S1r - C2i S1i + S2r C1r + S2i C1i
      sr = mfma(c2r, s1r, sr);
      sr = mfma(-c2i, s1i, sr);
      sr = mfma(s2r, c1r, sr);
      sr = mfma(s2i, c1i, sr);
      // Im S = C2r S1i + C2i S1r - S2r C1i + S2i C1r
      si = mfma(c2r, s1i, si);
      si = mfma(c2i, s1r, si);
      si = mfma(-s2r, c1i, si);
      si = mfma(s2i, c1r, si);
    }
    __syncthreads();
  }
  const int oc = col0 + wc + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int orow = row0 + wr + (lane >> 4) + 4 * r;
    if (orow < M && oc < Nc) {
      C[(size_t)orow * Nc + oc] = mk(cr[r], ci[r]);
      S[(size_t)orow * Nc + oc] = mk(sr[r], si[r]);
    }
  }
}

void launch_compose(const cplx* C2, const cplx* S2, const cplx* C1, const cplx* S1, cplx* C, cplx* S, int M, int K,
                    int Nc, hipStream_t s) {
  const dim3 grid((unsigned)((Nc + TILE - 1) / TILE), (unsigned)((M + TILE - 1) / TILE));
  hipLaunchKernelGGL(k_compose_mfma, grid, dim3(256), 0, s, C2, S2, C1, S1, C, S, M, K, Nc);
}

}  // namespace

extern "C" int nlm_greens_compose_dev(int M, int K, int Nc, const double* C2, const double* S2, const double* C1,
                                      const double* S1, double* C, double* S, void* stream) {
  if (M < 0 || K < 0 || Nc < 0) return set_err(NLM_INVALID_ARGUMENT, "invalid arguments");
  if (M == 0 || Nc == 0) return NLM_OK;
  if (!C || !S || (K > 0 && (!C2 || !S2 || !C1 || !S1))) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  launch_compose((const cplx*)C2, (const cplx*)S2, (const cplx*)C1, (const cplx*)S1, (cplx*)C, (cplx*)S, M, K, Nc,
                 (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}

extern "C" int nlm_greens_compose(int M, int K, int Nc, const double* C2, const double* S2, const double* C1,
                                  const double* S1, double* C, double* S) {
  if (M < 0 || K < 0 || Nc < 0 || !C2 || !S2 || !C1 || !S1 || !C || !S)
    return set_err(NLM_INVALID_ARGUMENT, "invalid arguments");
  if (M == 0 || Nc == 0) return NLM_OK;
  DevBuf<cplx> dC2, dS2, dC1, dS1, dC, dS;
  const size_t a = (size_t)M * K, b = (size_t)K * Nc, c = (size_t)M * Nc;
  int e = NLM_OK;
  if ((e = dC2.ensure(a)) || (e = dS2.ensure(a)) || (e = dC1.ensure(b)) || (e = dS1.ensure(b)) || (e = dC.ensure(c)) ||
      (e = dS.ensure(c)))
    return e;
  hipError_t err = hipSuccess;
  auto h2d = [&](cplx* d, const double* h, size_t n) {
    if (err == hipSuccess && n) err = hipMemcpy(d, h, n * sizeof(cplx), hipMemcpyHostToDevice);
  };
  h2d(dC2.p, C2, a);
  h2d(dS2.p, S2, a);
  h2d(dC1.p, C1, b);
  h2d(dS1.p, S1, b);
  if (err == hipSuccess) {
    launch_compose(dC2.p, dS2.p, dC1.p, dS1.p, dC.p, dS.p, M, K, Nc, 0);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipMemcpy(C, dC.p, c * sizeof(cplx), hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(S, dS.p, c * sizeof(cplx), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return set_err(NLM_DEVICE_ERROR, std::string("nlm_greens_compose: ") + hipGetErrorString(err));
  return NLM_OK;
}

// Cascade::computeGreensFunction (src/Cascade.cpp:110-135) on device-resident
// Green's functions: G = G_{n-1} o ... o G_1 o G_0 with the plans' last
// nlm_greens results (kept on the device), composed on the matrix cores without
// host round trips; only the final C, S are copied out.
extern "C" int nlm_greens_cascade(int n, nlm_plan* const* plans, double* C, double* S) {
  if (n < 1 || !plans || !C || !S) return set_err(NLM_INVALID_ARGUMENT, "invalid arguments");
  for (int i = 0; i < n; i++) {
    if (!plans[i]) return set_err(NLM_INVALID_ARGUMENT, "null plan");
    if (!plans[i]->greensRows) return set_err(NLM_RUNTIME_ERROR, "no Green's function has been computed");
    HIP_TRY(hipStreamSynchronize(plans[i]->stream));
  }
  hipStream_t s = plans[n - 1]->stream;
  int rows = plans[0]->greensRows;
  const int cols = plans[0]->greensCols;
  const cplx* aC = plans[0]->C.p;
  const cplx* aS = plans[0]->S.p;
  DevBuf<cplx> buf[2][2];
  for (int i = 1; i < n; i++) {
    const int M = plans[i]->greensRows, K = plans[i]->greensCols;
    if (K != rows) return set_err(NLM_INVALID_ARGUMENT, "Green's function shapes of adjacent media do not match");
    DevBuf<cplx>* o = buf[i % 2];
    int e;
    if ((e = o[0].ensure((size_t)M * cols)) || (e = o[1].ensure((size_t)M * cols))) return e;
    launch_compose(plans[i]->C.p, plans[i]->S.p, aC, aS, o[0].p, o[1].p, M, K, cols, s);
    HIP_TRY(hipGetLastError());
    aC = o[0].p;
    aS = o[1].p;
    rows = M;
  }
  const size_t total = (size_t)rows * cols;
  HIP_TRY(hipMemcpyAsync(C, aC, total * sizeof(cplx), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(S, aS, total * sizeof(cplx), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return NLM_OK;
}

// ---------------------------------------------------------------------------
// Gaussian-state quantities of a Green's function (multimode.py:15-36 of the
// reference, downstream of computeGreensFunction): the quadrature transfer
// matrix Z = [[Re(C+S), -Im(C-S)], [Im(C+S), Re(C-S)]] and the covariance Z Z^T
// (real, 2N x 2N), the product on the FP64 matrix cores.
// ---------------------------------------------------------------------------

namespace {

__global__ void k_quadrature(const cplx* __restrict__ C, const cplx* __restrict__ S, int N, double* __restrict__ Z) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)N * N) return;
  const int r = (int)(idx / N), c = (int)(idx % N);
  const cplx a = C[idx], b = S[idx];
  const size_t W = 2 * (size_t)N;
  Z[(size_t)r * W + c] = a.re + b.re;            // Re(C + S)
  Z[(size_t)r * W + N + c] = -(a.im - b.im);      // -Im(C - S)
  Z[(size_t)(N + r) * W + c] = a.im + b.im;      // Im(C + S)
  Z[(size_t)(N + r) * W + N + c] = a.re - b.re;  // Re(C - S)
}

// out = A B^T for real row-major A (M x K) and B (Ncol x K): 32 x 32 tiles, four
// waves of 16 x 16 FP64 MFMAs (operand layout as in k_compose_mfma).  B = A
// gives the Gram matrix A A^T.
__global__ void __launch_bounds__(256) k_gemm_nt_mfma(const double* __restrict__ A, const double* __restrict__ B,
                                                      int M, int Ncol, int K, double* __restrict__ out) {
  __shared__ double Ar[TILE][APAD], Br[TILE][APAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row0 = blockIdx.y * TILE, col0 = blockIdx.x * TILE;
  const int wr = (wave >> 1) * 16, wc = (wave & 1) * 16;
  v4d acc = {0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += KB) {
    for (int e = tid; e < TILE * KB; e += 256) {
      const int r = e / KB, k = e % KB, gk = k0 + k;
      Ar[r][k] = (row0 + r < M && gk < K) ? A[(size_t)(row0 + r) * K + gk] : 0.0;
      Br[r][k] = (col0 + r < Ncol && gk < K) ? B[(size_t)(col0 + r) * K + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      const int k = kk + (lane >> 4);
      acc = mfma(Ar[wr + (lane & 15)][k], Br[wc + (lane & 15)][k], acc);
    }
    __syncthreads();
  }
  const int oc = col0 + wc + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int orow = row0 + wr + (lane >> 4) + 4 * r;
    if (orow < M && oc < Ncol) out[(size_t)orow * Ncol + oc] = acc[r];
  }
}

}  // namespace

extern "C" int nlm_gaussian_covariance(int N, const double* C, const double* S, double* Z, double* cov) {
  if (N <= 0 || !C || !S || (!Z && !cov)) return set_err(NLM_INVALID_ARGUMENT, "invalid arguments");
  const size_t n2 = (size_t)N * N, W = 2 * (size_t)N;
  DevBuf<cplx> dC, dS;
  DevBuf<double> dZ, dV;
  int e;
  if ((e = dC.ensure(n2)) || (e = dS.ensure(n2)) || (e = dZ.ensure(W * W)) || (cov && (e = dV.ensure(W * W)))) return e;
  hipError_t err = hipMemcpy(dC.p, C, n2 * sizeof(cplx), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(dS.p, S, n2 * sizeof(cplx), hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(k_quadrature, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, 0, dC.p, dS.p, N, dZ.p);
    err = hipGetLastError();
  }
  if (err == hipSuccess && cov) {
    const dim3 grid((unsigned)((W + TILE - 1) / TILE), (unsigned)((W + TILE - 1) / TILE));
    hipLaunchKernelGGL(k_gemm_nt_mfma, grid, dim3(256), 0, 0, dZ.p, dZ.p, (int)W, (int)W, (int)W, dV.p);
    err = hipGetLastError();
  }
  if (err == hipSuccess && Z) err = hipMemcpy(Z, dZ.p, W * W * sizeof(double), hipMemcpyDeviceToHost);
  if (err == hipSuccess && cov) err = hipMemcpy(cov, dV.p, W * W * sizeof(double), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return set_err(NLM_DEVICE_ERROR, std::string("nlm_gaussian_covariance: ") + hipGetErrorString(err));
  return NLM_OK;
}

extern "C" int nlm_gemm_nt(int M, int Ncol, int K, const double* A, const double* B, double* out) {
  if (M < 0 || Ncol < 0 || K < 0 || (M > 0 && Ncol > 0 && (!A || !B || !out)))
    return set_err(NLM_INVALID_ARGUMENT, "invalid arguments");
  if (M == 0 || Ncol == 0) return NLM_OK;
  const bool same = A == B && M == Ncol;
  DevBuf<double> dA, dB, dO;
  int e;
  if ((e = dA.ensure((size_t)M * (K > 0 ? K : 1))) || (!same && (e = dB.ensure((size_t)Ncol * (K > 0 ? K : 1)))) ||
      (e = dO.ensure((size_t)M * Ncol)))
    return e;
  hipError_t err = K > 0 ? hipMemcpy(dA.p, A, (size_t)M * K * sizeof(double), hipMemcpyHostToDevice) : hipSuccess;
  if (err == hipSuccess && !same && K > 0)
    err = hipMemcpy(dB.p, B, (size_t)Ncol * K * sizeof(double), hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    const dim3 grid((unsigned)((Ncol + TILE - 1) / TILE), (unsigned)((M + TILE - 1) / TILE));
    hipLaunchKernelGGL(k_gemm_nt_mfma, grid, dim3(256), 0, 0, dA.p, same ? dA.p : dB.p, M, Ncol, K, dO.p);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipMemcpy(out, dO.p, (size_t)M * Ncol * sizeof(double), hipMemcpyDeviceToHost);
  if (err != hipSuccess) return set_err(NLM_DEVICE_ERROR, std::string("nlm_gemm_nt: ") + hipGetErrorString(err));
  return NLM_OK;
}

extern "C" int nlm_gram(int M, int K, const double* A, double* out) {
  if (M < 0 || K < 0 || (M > 0 && (!A || !out))) return set_err(NLM_INVALID_ARGUMENT, "invalid arguments");
  return nlm_gemm_nt(M, M, K, A, A, out);
}
