// Far-update kernels of stage 2 for r = 8 (see s2_far.hpp); one translation
// unit per r so that the instantiations compile in parallel.
#include "s2_far.hpp"

namespace pht {
PHT_S2_FAR_INST(8, 4)
PHT_S2_FAR_INST(8, 8)
}  // namespace pht
