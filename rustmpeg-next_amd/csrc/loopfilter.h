// loopfilter.h - host-side table builders of the LMCS and ALF plans
// (loopfilter.cpp). Internal to librmh.
#pragma once
#include <stdint.h>

#include "../../include/rmh.h"

namespace rmh {

// The tables of LmcsReshaper16::from_aps (vvc/filter.rs:2628-2705) at
// bit_depth 8..12 (false otherwise); any output may be null.
bool lmcs_lut(const rmh_lmcs_aps& aps, int bit_depth, uint16_t* fwd, uint16_t* inv, int32_t* chroma_adj,
              int32_t* pivot);

// The luma filter bank of a picture (kAlfBankEntries int16): the 16 fixed
// sets and the APS (zero when aps is null) for every class and transposition.
void alf_luma_bank(const rmh_alf_aps* aps, int bit_depth, int16_t* out);

// [alternative 0..7][6 coefficients, 6 clip values] of a chroma APS.
void alf_chroma_bank(const rmh_alf_aps& aps, int bit_depth, int16_t* out);

}  // namespace rmh
