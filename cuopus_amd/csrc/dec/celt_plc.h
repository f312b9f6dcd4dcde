// celt_plc.h -- pitch-based CELT loss concealment (celt_decoder.c:700-969).
#pragma once
#include "celt.h"

namespace cuo {

// Not yet restated: a chain that needs periodic CELT concealment is reported
// as a decode error (never silently decoded differently).
CUO_FN int celt_plc_periodic(CeltState *, int, CeltScratch *) { return -1; }

}  // namespace cuo
