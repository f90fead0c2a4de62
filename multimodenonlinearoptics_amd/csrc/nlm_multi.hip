// nlm_multi.hip -- heterogeneous batches: many independent propagations of
// different media (model, nFreqs, nZSteps, dispersion, pump) in one graph launch.
//
// Each instance is one batchSignalSimulation row of its own medium
// (src/_NonlinearMedium.cpp:378-444; the FNLM overload
// src/_FullyNonlinearMedium.hpp:8-11).  Instances are grouped into classes by
// (model, log2 N); every class runs as one persistent kernel whose workgroups
// take instances from a device work queue (one atomic counter per class) in
// order of decreasing nZ, so long propagations start first and short ones fill
// the tail.  For media with a single linear pump the pump rows are generated
// inside the kernel (InlinePump), so no (2nZ-1) x N pump grid is written or
// read per instance; Chi3 (nonlinear pump) and Chi2SFGOPA (two pumps) read their
// plan's device pump grids.  Every solver has a kernel (Chi3GNLSE with its
// transforms inside the RK stages).  All classes are captured into one hipGraph (one branch
// per class on its own stream, joined at the end); nlm_multi_launch replays it.
//
// The number of persistent workgroups of a class is chosen by a small
// autotuner at creation: each class is timed alone for a few grid sizes and
// the fastest kept (per process, keyed by model, log2 N and instance count).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "nlm_internal.hpp"
#include "nlm_multi.hpp"

__global__ void k_zero_ints(int* p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

// Staged pump rows.  From N = 4096 a field's 512 lanes have 256 registers each,
// and a propagation that generates its pump rows (InlinePump: the envelope's
// spectrum and the running phases held per lane, two inverse transforms per
// step) spills to scratch.  Such a class instead starts with this kernel, which
// writes every instance's time-domain pump rows (src/_NonlinearMedium.cpp:
// 225-250: row 0 the envelope, row r > 0 IFFT(FFT(envelope) exp(i r dzp D_p)),
// each times the beam-profile factor) into a buffer that the class's
// propagation kernel then reads like a medium's pump grid.  One workgroup per
// chunk of ROWS_CHUNK rows of one instance (chunks[2b] = instance, chunks[2b+1]
// = first row); every workgroup transforms the envelope once.
constexpr int ROWS_L = 12;        // sizes that stage their pump rows
constexpr int ROWS_CHUNK = 32;    // rows per workgroup

template <int L>
__global__ void __launch_bounds__(FftShape<L>::T)
    k_multi_rows(const MultiInst* __restrict__ inst, const int* __restrict__ chunks, const cplx* __restrict__ twt) {
  using F = FftShape<L>;
  constexpr int N = F::N, T = F::T;
  __shared__ cplx lds[F::lds_elems];
  __shared__ cplx twl[PropTw<L>::LDS_ELEMS];
  fill_twiddle_lds<L>(twl, twt);
  const int t = threadIdx.x;
  const MultiInst& I = inst[chunks[2 * blockIdx.x]];
  const int r0 = chunks[2 * blockIdx.x + 1];
  const int nZP = 2 * I.nZ - 1;
  const int r1 = r0 + ROWS_CHUNK < nZP ? r0 + ROWS_CHUNK : nZP;
  cplx* rows = const_cast<cplx*>(I.pump0);
  typename PropTw<L>::type tw;
  init_twiddles<L>(tw, twt, twl, t);
  cplx f0[8];
#pragma unroll
  for (int m = 0; m < 8; m++) f0[m] = I.env[t + T * m];
  if (r0 == 0) {
    const double s0 = I.strength ? I.strength[0] : 1.0;
#pragma unroll
    for (int m = 0; m < 8; m++) rows[t + T * m] = s0 * f0[m];
  }
  fft_field1<L, false>(f0, lds, t, tw);
  // the phase exp(i r dzp D_p): exact at the chunk's first row, then advanced by
  // exp(i dzp D_p) per row (as the in-kernel generator does between its exact
  // re-evaluations)
  const int rb = r0 > 0 ? r0 : 1;
  cplx ph[8], e1[8];
#pragma unroll
  for (int m = 0; m < 8; m++) {
    const double d = I.Dp[t + T * m];
    double sn, cs;
    sincos_reduced((double)rb * I.dzp * d, &sn, &cs);
    ph[m] = mk(cs, sn);
    sincos_reduced(I.dzp * d, &sn, &cs);
    e1[m] = mk(cs, sn);
  }
  for (int r = rb; r < r1; r++) {
    cplx f[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
      if (r > rb) ph[m] = ph[m] * e1[m];
      f[m] = f0[m] * ph[m];
    }
    fft_field1<L, true>(f, lds, t, tw);
    const double sc = (I.strength ? I.strength[r] : 1.0) * (1.0 / N);
#pragma unroll
    for (int m = 0; m < 8; m++) rows[(size_t)r * N + t + T * m] = f[m] * sc;
  }
}

// Kernel of a class: models sharing a right-hand side share kernels (Chi2AFC
// runs Chi2SFG's, Chi2ASHG Chi2SHG's).
static int class_model(int model) {
  return model == NLM_CHI2AFC ? NLM_CHI2SFG : model == NLM_CHI2ASHG ? NLM_CHI2SHG : model;
}

static multi_fn pick_multi(int model, int L, bool inl) {
  switch (class_model(model)) {
    case NLM_CHI3: return nlm_pick_multi_0(L, inl);
    case NLM_CHI2PDC: return nlm_pick_multi_1(L, inl);
    case NLM_CHI2SFG: return nlm_pick_multi_2(L, inl);
    case NLM_CHI2SFGPDC: return nlm_pick_multi_4(L, inl);
    case NLM_CHI2SFGII: return nlm_pick_multi_5(L, inl);
    case NLM_CHI2PDCII: return nlm_pick_multi_6(L, inl);
    case NLM_CHI2SFGOPA: return nlm_pick_multi_7(L, inl);
    case NLM_CHI2SFGXPM: return nlm_pick_multi_8(L, inl);
    case NLM_CHI2SHG: return nlm_pick_multi_9(L, inl);
    case NLM_CHI2SHGOPA: return nlm_pick_multi_10(L, inl);
    case NLM_CHI2DSFG: return nlm_pick_multi_11(L, inl);
    case NLM_CHI2SHGXPM: return nlm_pick_multi_12(L, inl);
    case NLM_CHI3GNLSE: return nlm_pick_multi_14(L, inl);
  }
  return nullptr;
}

struct MultiClass {
  int model, L, count, grid;
  bool inl;   // pump rows generated in the kernel
  bool rows;  // pump rows staged by k_multi_rows before the propagation kernel
  int chunkOff = 0, nChunks = 0;
  multi_fn fn;
  size_t orderOff;  // offset of this class's instance order in the order array
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
};

struct nlm_multi_s {
  int n = 0;
  std::vector<int> N, nOut;
  std::vector<size_t> outOff;
  DevBuf<cplx> inBuf, outBuf, envBuf;
  DevBuf<double> realBuf;
  DevBuf<MultiInst> inst;
  DevBuf<int> order, counters;
  DevBuf<cplx> rowsBuf;  // staged pump rows of the classes with MultiClass::rows
  DevBuf<int> chunks;    // their generation chunks (instance, first row)
  std::vector<MultiClass> classes;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  float lastMs = 0.f;
};

static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, int> g_tuned;  // (model, L, generated pump, count) -> grid

static int device_cus() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return cus > 0 ? cus : 256;
}

static void launch_rows(nlm_multi_s* m, const MultiClass& c, const cplx* tw, hipStream_t s) {
  if (c.rows && c.nChunks > 0)
    hipLaunchKernelGGL((k_multi_rows<ROWS_L>), dim3(c.nChunks), dim3(FftShape<ROWS_L>::T), 0, s, m->inst.p,
                       m->chunks.p + 2 * (size_t)c.chunkOff, tw);
}

static int run_class(nlm_multi_s* m, const MultiClass& c, int grid, const cplx* tw, hipStream_t s) {
  launch_rows(m, c, tw, s);
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, m->counters.p + (&c - m->classes.data()), 1);
  c.fn(m->inst.p, m->order.p + c.orderOff, c.count, m->counters.p + (&c - m->classes.data()), tw, grid, s);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}

extern "C" int nlm_multi_destroy(nlm_multi* m) {
  if (!m) return NLM_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->exec) (void)hipGraphExecDestroy(m->exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (auto& c : m->classes) {
    if (c.stream) (void)hipStreamDestroy(c.stream);
    if (c.done) (void)hipEventDestroy(c.done);
  }
  if (m->fork) (void)hipEventDestroy(m->fork);
  if (m->t0) (void)hipEventDestroy(m->t0);
  if (m->t1) (void)hipEventDestroy(m->t1);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return NLM_OK;
}

extern "C" int nlm_multi_create(int n, const nlm_multi_item* items, int autotune, nlm_multi** out) {
  if (!out || n < 0 || (n > 0 && !items)) return set_err(NLM_INVALID_ARGUMENT, "invalid arguments");
  *out = nullptr;
  const bool stage = (autotune & 2) != 0;
  autotune &= 1;
  nlm_multi_s* m = new nlm_multi_s();
  auto fail = [&](int code) {
    nlm_multi_destroy(m);
    return code;
  };
  m->n = n;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&m->fork) != hipSuccess || hipEventCreate(&m->t0) != hipSuccess ||
      hipEventCreate(&m->t1) != hipSuccess)
    return fail(set_err(NLM_DEVICE_ERROR, "stream/event creation failed"));

  // sizes of the packed per-instance buffers
  size_t inTot = 0, outTot = 0, envTot = 0, realTot = 0;
  std::map<std::tuple<int, int, int>, std::vector<int>> byClass;  // (class model, L, generated pump)
  m->N.resize(n);
  m->nOut.resize(n);
  m->outOff.resize(n);
  for (int i = 0; i < n; i++) {
    const nlm_multi_item& it = items[i];
    nlm_plan* p = it.plan;
    if (!p) return fail(set_err(NLM_INVALID_ARGUMENT, "null plan in item"));
    const bool pumped = p->nP > 0;
    const bool inl = pumped && it.envelope != nullptr;
    if (inl && (p->nP != 1 || p->model == NLM_CHI3))
      return fail(set_err(NLM_INVALID_ARGUMENT,
                          "a pump generated in the batch needs a single linear pump; pass no envelope and run the "
                          "medium's pump simulation instead"));
    if (!pick_multi(p->model, p->L, inl))
      return fail(set_err(NLM_INVALID_ARGUMENT, "heterogeneous batches take nFreqs a power of two, 32..4096"));
    if (!it.input || it.cols <= 0 || it.cols % p->N != 0 || it.cols / p->N > p->nS)
      return fail(set_err(NLM_INVALID_ARGUMENT, "Signals not of correct length!"));
    if (it.input_mode < 0 || it.input_mode >= p->nS)
      return fail(set_err(NLM_INVALID_ARGUMENT, "inputModes does not match any mode in the system"));
    if (it.n_out < 1 || it.n_out > p->nS) return fail(set_err(NLM_INVALID_ARGUMENT, "Requested no outputs!"));
    for (int q = 0; q < it.n_out; q++)
      if (it.outputs[q] < 0 || it.outputs[q] >= p->nS)
        return fail(set_err(NLM_INVALID_ARGUMENT, "output mode out of range"));
    if (inl && !it.disp_pump)
      return fail(set_err(NLM_INVALID_ARGUMENT, "a generated pump needs the pump dispersion"));
    if (pumped && !inl)
      if (int e = check_pumps_ready(p)) return fail(e);
    m->N[i] = p->N;
    m->nOut[i] = it.n_out;
    m->outOff[i] = outTot;
    inTot += it.cols;
    outTot += (size_t)it.n_out * p->N;
    if (inl) {
      envTot += p->N;
      realTot += p->N + (it.strength ? 2 * (size_t)p->nZ - 1 : 0);
    }
    byClass[std::make_tuple(class_model(p->model), p->L, inl ? 1 : 0)].push_back(i);
  }
  int e;
  if ((e = m->inBuf.ensure(inTot)) || (e = m->outBuf.ensure(outTot)) || (e = m->envBuf.ensure(envTot)) ||
      (e = m->realBuf.ensure(realTot)) || (e = m->inst.ensure(n)) || (e = m->order.ensure(n)))
    return fail(e);

  // descriptors and packed uploads
  std::vector<MultiInst> hi(n);
  std::vector<cplx> hin(inTot), henv(envTot);
  std::vector<double> hreal(realTot);
  size_t io = 0, eo = 0, ro = 0;
  for (int i = 0; i < n; i++) {
    const nlm_multi_item& it = items[i];
    nlm_plan* p = it.plan;
    MultiInst& d = hi[i];
    memset(&d, 0, sizeof d);
    d.dstep = p->dstep.p;
    d.halfIn = p->halfIn.p;
    d.halfOut = p->halfOut.p;
    d.coef = p->coef.p;
    d.cconst = p->cconst.p;
    std::memcpy(hin.data() + io, it.input, sizeof(cplx) * it.cols);
    d.in = m->inBuf.p + io;
    io += it.cols;
    d.out = m->outBuf.p + m->outOff[i];
    d.nZ = p->nZ;
    d.inTime = it.in_time_domain ? 1 : 0;
    d.inCols = it.cols;
    d.nOut = it.n_out;
    for (int q = 0; q < 4; q++) d.inPos[q] = d.outPos[q] = -1;
    input_channels(p->nS, it.cols / p->N, it.input_mode, d.inPos);
    for (int q = 0; q < it.n_out; q++) d.outPos[it.outputs[q]] = q;
    d.raman = p->raman.p;
    d.filt = p->filt.p;
    if (p->nP > 0 && !it.envelope) {
      d.pump0 = p->pumpT[0].p;
      d.pump1 = p->nP > 1 ? p->pumpT[1].p : nullptr;
    }
    if (p->nP > 0 && it.envelope) {
      std::memcpy(henv.data() + eo, it.envelope, sizeof(cplx) * p->N);
      d.env = m->envBuf.p + eo;
      eo += p->N;
      std::memcpy(hreal.data() + ro, it.disp_pump, sizeof(double) * p->N);
      d.Dp = m->realBuf.p + ro;
      ro += p->N;
      if (it.strength) {
        std::memcpy(hreal.data() + ro, it.strength, sizeof(double) * (2 * (size_t)p->nZ - 1));
        d.strength = m->realBuf.p + ro;
        ro += 2 * (size_t)p->nZ - 1;
      }
      d.dzp = it.dzp;
    }
  }
  // classes that stage their pump rows: ROWS_L with a generated pump (key 2),
  // when all their rows fit a quarter of the device memory and half of what is
  // free.  (Staging only the longer half of such a class, the shorter half
  // generating its rows in the kernel from the start, was slower: 211 against
  // 197 ms on 2000 SHG / PDC media.)
  if (stage) {
    std::vector<std::tuple<int, int, int>> keys;
    for (auto& kv : byClass)
      if (std::get<2>(kv.first) == 1 && std::get<1>(kv.first) == ROWS_L) keys.push_back(kv.first);
    for (const auto& k : keys) {
      byClass[std::make_tuple(std::get<0>(k), ROWS_L, 2)] = byClass[k];
      byClass.erase(k);
    }
  }
  if ((e = m->counters.ensure(byClass.size() + 1))) return fail(e);
  size_t rowsTot = 0;
  for (auto& kv : byClass)
    if (std::get<2>(kv.first) == 2)
      for (int i : kv.second) rowsTot += (2 * (size_t)hi[i].nZ - 1) * m->N[i];
  bool useRows = rowsTot > 0;
  if (useRows) {
    size_t freeB = 0, totalB = 0;
    useRows = hipMemGetInfo(&freeB, &totalB) == hipSuccess && rowsTot * sizeof(cplx) <= totalB / 4 &&
              rowsTot * sizeof(cplx) <= freeB / 2;
    if (useRows && (e = m->rowsBuf.ensure(rowsTot))) return fail(e);
  }
  std::vector<int> hchunks;
  size_t rowsOff = 0;
  // classes and their longest-first orders
  std::vector<int> hord;
  hord.reserve(n);
  for (auto& kv : byClass) {
    MultiClass c;
    c.model = std::get<0>(kv.first);
    c.L = std::get<1>(kv.first);
    c.inl = std::get<2>(kv.first) != 0;
    c.rows = useRows && std::get<2>(kv.first) == 2;
    c.fn = pick_multi(c.model, c.L, c.inl && !c.rows);
    if (c.rows) {
      c.chunkOff = (int)(hchunks.size() / 2);
      for (int i : kv.second) {
        hi[i].pump0 = m->rowsBuf.p + rowsOff;
        const int nZP = 2 * hi[i].nZ - 1;
        rowsOff += (size_t)nZP * m->N[i];
        for (int r = 0; r < nZP; r += ROWS_CHUNK) {
          hchunks.push_back(i);
          hchunks.push_back(r);
        }
      }
      c.nChunks = (int)(hchunks.size() / 2) - c.chunkOff;
    }
    std::vector<int> idx = kv.second;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return hi[a].nZ > hi[b].nZ; });
    c.orderOff = hord.size();
    hord.insert(hord.end(), idx.begin(), idx.end());
    c.count = (int)idx.size();
    c.grid = 1;
    m->classes.push_back(c);
  }
  if (!hchunks.empty()) {
    if ((e = m->chunks.ensure(hchunks.size()))) return fail(e);
    HIP_TRY(hipMemcpy(m->chunks.p, hchunks.data(), sizeof(int) * hchunks.size(), hipMemcpyHostToDevice));
  }
  if (n > 0) {
    HIP_TRY(hipMemcpy(m->inst.p, hi.data(), sizeof(MultiInst) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->order.p, hord.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->inBuf.p, hin.data(), sizeof(cplx) * inTot, hipMemcpyHostToDevice));
    if (envTot) HIP_TRY(hipMemcpy(m->envBuf.p, henv.data(), sizeof(cplx) * envTot, hipMemcpyHostToDevice));
    if (realTot) HIP_TRY(hipMemcpy(m->realBuf.p, hreal.data(), sizeof(double) * realTot, hipMemcpyHostToDevice));
  }

  // grid sizes: autotuned per class (timed alone) or the default of four
  // workgroups per CU, never more workgroups than instances
  const int cus = device_cus();
  for (auto& c : m->classes) {
    const cplx* tw = nullptr;
    if ((e = get_twiddles(1 << c.L, &tw))) return fail(e);
    const int cand[4] = {cus, 2 * cus, 4 * cus, 8 * cus};
    int best = std::min(c.count, 4 * cus);
    const auto key = std::make_tuple(c.model, c.L, c.rows ? 2 : c.inl ? 1 : 0, c.count);
    bool cached = false;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto f = g_tuned.find(key);
      if (f != g_tuned.end()) {
        best = f->second;
        cached = true;
      }
    }
    if (autotune && !cached && c.count > cus) {
      float bestMs = 1e30f;
      for (int g : cand) {
        const int grid = std::min(c.count, g);
        if ((e = run_class(m, c, grid, tw, m->stream))) return fail(e);  // warm
        HIP_TRY(hipEventRecord(m->t0, m->stream));
        if ((e = run_class(m, c, grid, tw, m->stream))) return fail(e);
        HIP_TRY(hipEventRecord(m->t1, m->stream));
        HIP_TRY(hipEventSynchronize(m->t1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, m->t0, m->t1));
        if (ms < bestMs) {
          bestMs = ms;
          best = grid;
        }
        if (grid == c.count) break;
      }
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tuned[key] = best;
    }
    c.grid = std::max(1, best);
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c.done, hipEventDisableTiming) != hipSuccess)
      return fail(set_err(NLM_DEVICE_ERROR, "stream/event creation failed"));
  }

  // capture: fork one branch per class, join on the main stream
  if (!m->classes.empty()) {
    HIP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipEventRecord(m->fork, m->stream);
    for (auto& c : m->classes) {
      const cplx* tw = nullptr;
      (void)get_twiddles(1 << c.L, &tw);
      if (ce == hipSuccess) ce = hipStreamWaitEvent(c.stream, m->fork, 0);
      if (ce == hipSuccess) {
        launch_rows(m, c, tw, c.stream);
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, c.stream, m->counters.p + (&c - m->classes.data()), 1);
        c.fn(m->inst.p, m->order.p + c.orderOff, c.count, m->counters.p + (&c - m->classes.data()), tw, c.grid,
             c.stream);
        ce = hipGetLastError();
      }
      if (ce == hipSuccess) ce = hipEventRecord(c.done, c.stream);
      if (ce == hipSuccess) ce = hipStreamWaitEvent(m->stream, c.done, 0);
    }
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(m->stream, &g);
    if (ce != hipSuccess || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(set_err(NLM_DEVICE_ERROR, std::string("graph capture failed: ") +
                                                hipGetErrorString(ce != hipSuccess ? ce : ee)));
    }
    m->graph = g;
    HIP_TRY(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
  }
  *out = m;
  return NLM_OK;
}

extern "C" int nlm_multi_launch(nlm_multi* m) {
  if (!m) return set_err(NLM_INVALID_ARGUMENT, "null batch");
  if (!m->exec) return NLM_OK;
  HIP_TRY(hipEventRecord(m->t0, m->stream));
  HIP_TRY(hipGraphLaunch(m->exec, m->stream));
  HIP_TRY(hipEventRecord(m->t1, m->stream));
  return NLM_OK;
}

extern "C" int nlm_multi_sync(nlm_multi* m, double* ms) {
  if (!m) return set_err(NLM_INVALID_ARGUMENT, "null batch");
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (ms) {
    float f = 0.f;
    if (m->exec) HIP_TRY(hipEventElapsedTime(&f, m->t0, m->t1));
    *ms = f;
  }
  return NLM_OK;
}

extern "C" int nlm_multi_output(nlm_multi* m, int i, double* out) {
  if (!m || !out) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  if (i < 0 || i >= m->n) return set_err(NLM_INVALID_ARGUMENT, "instance index out of range");
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(out, m->outBuf.p + m->outOff[i], sizeof(cplx) * m->nOut[i] * m->N[i], hipMemcpyDeviceToHost));
  return NLM_OK;
}

extern "C" int nlm_multi_grid(nlm_multi* m, int c, int32_t* model, int32_t* log2n, int32_t* count, int32_t* grid) {
  if (!m) return set_err(NLM_INVALID_ARGUMENT, "null batch");
  if (c < 0 || c >= (int)m->classes.size()) return set_err(NLM_INVALID_ARGUMENT, "class index out of range");
  if (model) *model = m->classes[c].model;
  if (log2n) *log2n = m->classes[c].L;
  if (count) *count = m->classes[c].count;
  if (grid) *grid = m->classes[c].grid;
  return NLM_OK;
}

extern "C" int nlm_multi_classes(nlm_multi* m) { return m ? (int)m->classes.size() : 0; }
