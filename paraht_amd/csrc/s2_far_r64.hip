// Far-update kernels of stage 2 for r = 64 (see s2_far.hpp); one translation
// unit per r so that the instantiations compile in parallel.
#include "s2_far.hpp"

namespace pht {
PHT_S2_FAR_INST(64, 8)
PHT_S2_FAR_INST(64, 16)
PHT_S2_FAR_INST(64, 32)
}  // namespace pht
