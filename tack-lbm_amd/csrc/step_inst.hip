// Instantiates every step kernel of one (storage, collision) pair, single-lattice and batched:
// 3 obstacles x 3 sides x {quiet, report}.  Compiled once per pair with -DTLBM_STORAGE=<0..2>
// -DTLBM_COLL=<0..2> so the 9 units build in parallel.
#include "registry.hpp"

#if !defined(TLBM_STORAGE) || !defined(TLBM_COLL)
#error "compile with -DTLBM_STORAGE=<storage> -DTLBM_COLL=<collision>"
#endif

#define TLBM_UNIT_NAME_(s, c) tlbm_register_##s##_##c
#define TLBM_UNIT_NAME(s, c) TLBM_UNIT_NAME_(s, c)

void TLBM_UNIT_NAME(TLBM_STORAGE, TLBM_COLL)(tlbm::StepTable& t, tlbm::BatchTable& bt) {
  tlbm::fill_table<TLBM_COLL, TLBM_STORAGE>(t, bt);
}
