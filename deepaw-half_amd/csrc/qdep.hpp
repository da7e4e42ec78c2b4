// Potential-dependent correction of the per-atom non-local strengths D
// ("QDEP", MIMIC_US) on the device (see qdep.hip).
#pragma once

#include "paw.hpp"

namespace half {

// Replaces the D block of every atom by DION + the QDEP correction of the
// current Veff at the current atom positions (sys.positions_cart), and
// fills dD/dR (sys.ddblk / sys.d_ddblk, [alpha][a][b] per atom).  Requires
// the datasets' PAW partial waves; all-zero partial waves give D = DION.
void build_qdep(System &sys, hipStream_t s);

}  // namespace half
