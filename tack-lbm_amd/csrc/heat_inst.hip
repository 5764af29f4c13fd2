// Instantiates the temperature variant of the step kernel for one collision (FP32 populations):
// 5 obstacles x 3 sides x {quiet, report}.  Compiled once per collision with -DTLBM_COLL=<0..2>.
#include "registry.hpp"

#if !defined(TLBM_COLL)
#error "compile with -DTLBM_COLL=<collision>"
#endif

#define TLBM_HEAT_UNIT_NAME_(c) tlbm_register_heat_##c
#define TLBM_HEAT_UNIT_NAME(c) TLBM_HEAT_UNIT_NAME_(c)

void TLBM_HEAT_UNIT_NAME(TLBM_COLL)(tlbm::HeatTable& t) { tlbm::fill_heat_table<TLBM_COLL>(t); }
