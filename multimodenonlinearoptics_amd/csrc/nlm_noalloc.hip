// nlm_noalloc.hip -- the Green's-function solve for callers that cannot allocate
// device memory (code under graph capture, serving loops, harnesses that forbid
// hipMalloc): nlm_greens_noalloc.
//
// nlm_greens keeps two scratch buffers on the plan, allocated on first use: the
// final rows G of the 2 N nIn impulse propagations ([2 N nIn][N nOut]) and, for
// the one-mode undepleted media, the RK4 step-propagator table ([nZ][N][2]).
// Here both live in one file-scope device array of NLM_NOALLOC_SCRATCH_BYTES,
// and C and S are written to the caller's device buffers.  Each GPU holds its
// own copy of the array from the first call on it until the process ends; every
// call on a GPU shares that GPU's copy, so calls that overlap on one GPU (other
// host threads, other streams) race.  Nothing in the array needs clearing: every
// element a call reads it has written first.
#include <mutex>

#include "nlm_internal.hpp"

namespace {

constexpr size_t kScratchElems = NLM_NOALLOC_SCRATCH_BYTES / sizeof(cplx);
static_assert(NLM_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must keep 256-byte alignment");

// In host code this name is a host-side stand-in: kernels get the per-device
// address obtained from hipGetSymbolAddress below, never the name itself.
static __device__ __align__(256) cplx scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
cplx* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on `device`, resolved once per device with that
// device current.
int scratch_on(int device, cplx** out) {
  if (device < 0 || device >= kMaxDevices) return set_err(NLM_DEVICE_ERROR, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess || !ptr)
      return set_err(NLM_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<cplx*>(ptr);
  }
  *out = g_scratch_ptr[device];
  return NLM_OK;
}

}  // namespace

extern "C" int nlm_greens_noalloc(nlm_plan* p, int in_time, int n_in, const int32_t* inputs, int n_out,
                                  const int32_t* outputs, double* C_dev, double* S_dev) {
  if (!p || !inputs || !outputs || !C_dev || !S_dev) return set_err(NLM_INVALID_ARGUMENT, "null argument");
  if (p->model >= NLM_CHI2SHG) return set_err(NLM_RUNTIME_ERROR, "Object does not have this method.");
  if (n_in < 1 || n_in > p->nS || n_out < 1 || n_out > p->nS)
    return set_err(NLM_INVALID_ARGUMENT, "invalid number of input / output modes");
  if (int e = check_pumps_ready(p)) return e;
  const int N = p->N;
  const size_t W = (size_t)n_out * N, nF = 2 * (size_t)n_in * N;
  const size_t gElems = nF * W;
  const size_t tElems = uses_step_table(p) ? (size_t)p->nZ * N * 2 : 0;
  // the table starts on a 256-byte boundary after G
  const size_t tOff = (gElems + 15) / 16 * 16;
  if (tOff + tElems > kScratchElems)
    return set_err(NLM_INVALID_ARGUMENT,
                   "nlm_greens_noalloc: the Green's rows (" + std::to_string(gElems * sizeof(cplx) >> 20) +
                       " MiB) and step table (" + std::to_string(tElems * sizeof(cplx) >> 20) +
                       " MiB) exceed the fixed scratch of " + std::to_string(NLM_NOALLOC_SCRATCH_BYTES >> 20) +
                       " MiB; use nlm_greens");
  cplx* base = nullptr;
  if (int e = scratch_on(p->device, &base)) return e;
  KParams kp;
  fill_common(p, kp);
  kp.kind = KIND_GREENS;
  kp.nFields = (int)nF;
  kp.inTime = in_time;
  kp.outTime = in_time;
  kp.nIn = n_in;
  kp.nOut = n_out;
  kp.out = base;
  for (int i = 0; i < n_in; i++) {
    if (inputs[i] < 0 || inputs[i] >= p->nS) return set_err(NLM_INVALID_ARGUMENT, "input mode out of range");
    kp.inputs[i] = inputs[i];
  }
  for (int om = 0; om < n_out; om++) {
    if (outputs[om] < 0 || outputs[om] >= p->nS) return set_err(NLM_INVALID_ARGUMENT, "output mode out of range");
    kp.outPos[outputs[om]] = om;
  }
  if (int e = launch_prop_kernel_tab(p, kp, tElems ? base + tOff : nullptr)) return e;
  launch_greens_assemble(base, reinterpret_cast<cplx*>(C_dev), reinterpret_cast<cplx*>(S_dev), N, n_in, n_out,
                         n_in < n_out ? n_in : n_out, nullptr, p->stream);
  HIP_TRY(hipGetLastError());
  return NLM_OK;
}
