// nlm_multi.hpp -- the persistent queue kernel of heterogeneous batches
// (nlm_multi.hip), instantiated per model in nlm_multi_inst.hip.
#pragma once
#include <type_traits>

#include "nlm_prop.hpp"

struct MultiInst {
  const cplx* dstep;
  const cplx* halfIn;
  const cplx* halfOut;
  const cplx* coef;
  const cplx* cconst;
  const cplx* in;
  cplx* out;
  const cplx* env;     // generated linear pump: envelope, D_p, beam-profile factors, dzp
  const double* Dp;
  const double* strength;
  double dzp;
  const cplx* pump0;   // otherwise the medium's device pump grids
  const cplx* pump1;
  const cplx* raman;   // Chi3GNLSE
  const cplx* filt;
  int nZ, inTime, inCols, nOut;
  int inPos[4], outPos[4];
};

template <int L>
struct MultiLaunch {
  static constexpr int THREADS = FftShape<L>::T;  // one instance per workgroup
};

// Workgroups take instances from the class's queue (one atomic counter) until it is
// empty; INLINE generates the pump rows of a single linear pump in the kernel.
template <class M, int L, bool INLINE>
__global__ void __launch_bounds__(MultiLaunch<L>::THREADS)
    k_prop_multi(const MultiInst* __restrict__ inst, const int* __restrict__ order, int n, int* counter,
                 const cplx* tw) {
  using F = FftShape<L>;
  __shared__ cplx lds[F::lds_elems];
  __shared__ cplx twl[PropTw<L>::LDS_ELEMS];
  __shared__ int item;
  const int t = threadIdx.x;
  fill_twiddle_lds<L>(twl, tw);
  for (;;) {
    __syncthreads();  // the previous instance's readers of `item` are done
    if (t == 0) item = atomicAdd(counter, 1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(item);
    if (w >= n) break;  // uniform: every workgroup leaves once the queue is empty
    const MultiInst& I = inst[order[w]];
    KParams kp{};
    kp.nZ = I.nZ;
    kp.kind = KIND_BATCH;
    kp.nFields = 1;
    kp.inTime = I.inTime;
    kp.outTime = I.inTime;
    kp.inCols = I.inCols;
    kp.nOut = I.nOut;
    for (int q = 0; q < 4; q++) {
      kp.inPos[q] = I.inPos[q];
      kp.outPos[q] = I.outPos[q];
    }
    kp.in = I.in;
    kp.out = I.out;
    kp.tw = tw;
    kp.dstep = I.dstep;
    kp.halfIn = I.halfIn;
    kp.halfOut = I.halfOut;
    kp.coef = I.coef;
    kp.cconst = I.cconst;
    kp.env = I.env;
    kp.Dp = I.Dp;
    kp.strength = I.strength;
    kp.dzp = I.dzp;
    kp.pump0 = I.pump0;
    kp.pump1 = I.pump1;
    kp.raman = I.raman;
    kp.filt = I.filt;
    if constexpr (std::is_same<M, MChi3GNLSE>::value)
      gnlse_field<L>(kp, 0, true, t, lds, twl);
    else
      prop_field<M, L, INLINE>(kp, 0, true, t, lds, twl);
  }
}

typedef void (*multi_fn)(const MultiInst*, const int*, int, int*, const cplx*, int grid, hipStream_t);

template <class M, int L, bool INLINE>
void launch_multi(const MultiInst* inst, const int* order, int n, int* counter, const cplx* tw, int grid,
                  hipStream_t s) {
  hipLaunchKernelGGL((k_prop_multi<M, L, INLINE>), dim3(grid), dim3(MultiLaunch<L>::THREADS), 0, s, inst, order, n,
                     counter, tw);
}

// Per-model launchers (nlm_multi_inst.hip), log2 N = 5..12; nullptr where the
// model has no such kernel (a generated pump needs a single linear pump).
multi_fn nlm_pick_multi_0(int L, bool inl), nlm_pick_multi_1(int L, bool inl), nlm_pick_multi_2(int L, bool inl),
    nlm_pick_multi_4(int L, bool inl), nlm_pick_multi_5(int L, bool inl), nlm_pick_multi_6(int L, bool inl),
    nlm_pick_multi_7(int L, bool inl), nlm_pick_multi_8(int L, bool inl), nlm_pick_multi_9(int L, bool inl),
    nlm_pick_multi_10(int L, bool inl), nlm_pick_multi_11(int L, bool inl), nlm_pick_multi_12(int L, bool inl),
    nlm_pick_multi_14(int L, bool inl);
