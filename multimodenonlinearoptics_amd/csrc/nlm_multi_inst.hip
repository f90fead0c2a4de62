// nlm_multi_inst.hip -- heterogeneous-batch kernels of one model for nFreqs
// 32..4096.  Compiled once per model with -DNLM_INST_MODEL=<id>.
#include "nlm_multi.hpp"

#ifndef NLM_INST_MODEL
#error "NLM_INST_MODEL must be defined"
#endif

#define NLM_CAT2(a, b) a##b
#define NLM_CAT(a, b) NLM_CAT2(a, b)

namespace {
using MM = typename ModelOf<NLM_INST_MODEL>::type;
// the pump rows can be generated in the kernel for a single linear pump
// (_NonlinearMedium::runPumpSimulation); Chi3's pump is nonlinear (solvers/Chi3.cpp:25-54)
constexpr bool CAN_INLINE = MM::NP == 1 && NLM_INST_MODEL != NLM_CHI3;

template <bool INL>
multi_fn pick(int L) {
  if constexpr (INL && !CAN_INLINE) {
    return nullptr;
  } else {
    switch (L) {
      case 5: return &launch_multi<MM, 5, INL>;
      case 6: return &launch_multi<MM, 6, INL>;
      case 7: return &launch_multi<MM, 7, INL>;
      case 8: return &launch_multi<MM, 8, INL>;
      case 9: return &launch_multi<MM, 9, INL>;
      case 10: return &launch_multi<MM, 10, INL>;
      case 11: return &launch_multi<MM, 11, INL>;
      case 12: return &launch_multi<MM, 12, INL>;
    }
    return nullptr;
  }
}
}  // namespace

multi_fn NLM_CAT(nlm_pick_multi_, NLM_INST_MODEL)(int L, bool inl) { return inl ? pick<true>(L) : pick<false>(L); }
