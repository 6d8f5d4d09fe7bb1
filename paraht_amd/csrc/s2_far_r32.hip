// Far-update kernels of stage 2 for r = 32 (see s2_far.hpp); one translation
// unit per r so that the instantiations compile in parallel.
#include "s2_far.hpp"

namespace pht {
PHT_S2_FAR_INST(32, 8)
PHT_S2_FAR_INST(32, 16)
PHT_S2_FAR_INST(32, 32)
}  // namespace pht
