// nlm_prop_inst.hip -- instantiates the propagation kernel of one model for
// every supported size.  Compiled once per model with -DNLM_INST_MODEL=<id>.
#include "nlm_prop.hpp"

#ifndef NLM_INST_MODEL
#error "NLM_INST_MODEL must be defined"
#endif

template <int MODEL, int L>
static prop_fn pick_L() {
  if constexpr (MODEL == NLM_CHI3GNLSE) return &launch_gnlse<L>;
  else return &launch_prop<typename ModelOf<MODEL>::type, L>;
}

#define NLM_CAT2(a, b) a##b
#define NLM_CAT(a, b) NLM_CAT2(a, b)

prop_fn NLM_CAT(nlm_pick_prop_, NLM_INST_MODEL)(int L) {
  constexpr int M = NLM_INST_MODEL;
  switch (L) {
    case 3: return pick_L<M, 3>();
    case 4: return pick_L<M, 4>();
    case 5: return pick_L<M, 5>();
    case 6: return pick_L<M, 6>();
    case 7: return pick_L<M, 7>();
    case 8: return pick_L<M, 8>();
    case 9: return pick_L<M, 9>();
    case 10: return pick_L<M, 10>();
    case 11: return pick_L<M, 11>();
    case 12: return pick_L<M, 12>();
    case 13: return pick_L<M, 13>();
  }
  return nullptr;
}
