// Far-update kernels of stage 2 for r = 16 (see s2_far.hpp); one translation
// unit per r so that the instantiations compile in parallel.
#include "s2_far.hpp"

namespace pht {
PHT_S2_FAR_INST(16, 4)
PHT_S2_FAR_INST(16, 8)
PHT_S2_FAR_INST(16, 16)
}  // namespace pht
