// Setup ordering for domains of more than 2^26 blocks (setup_wide.hip).
#pragma once

#include <algorithm>

#include "kernels.h"

namespace ckmpm {

// The order of SetupOrder (particles sorted by block key, then by their rank among the
// earlier input particles of their binning cell, then by cell) without packing the block
// key and the cell into one 32-bit word: for key_bits + 6 > 32.  Same outputs (sorted block
// keys, the input index of each processing position, the occupied-block bitmap) and the
// same temporaries.
void SetupOrderWide(const DomainDev& d, const float* xyz, uint32_t n, int key_bits, uint32_t* keys, uint32_t* order,
                    const SetupTemps& t, uint32_t* bitmap, Status* st, hipStream_t s);

}  // namespace ckmpm
