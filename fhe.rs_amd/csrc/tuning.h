// Launch shapes of the fused row kernels that the multiplication autotuner
// chooses, per degree and process-wide.
#pragma once

#include "fhehip_internal.h"

namespace fhehip {

struct FusedTuning {
    int xcd = 1;          // XCD-aware workgroup order in the 1-D fused grids
    int ks_mac_inst = 4;  // instances sharing one load of the key rows in ks_mac (2, 4 or 8)
    int re_tensor = 4;    // log2 values per thread of tensor_intt (4 or 5)
    int re_lift = 4;      // of lift_ntt
    int re_finish = 4;    // of ks_finish
};

// The current choice for degree 2^logn (the defaults until set).
FusedTuning fused_tuning(int logn);
// Validates and stores a choice for degree 2^logn.
void set_fused_tuning(int logn, const FusedTuning& t);

}  // namespace fhehip
