// Launch shapes of the fused row kernels that the multiplication autotuner
// chooses, per degree and process-wide.
#pragma once

#include "fhehip_internal.h"

namespace fhehip {

enum : int { HALVES_FINISH = 1, HALVES_INPUT = 2, HALVES_TENSOR = 4 };

struct FusedTuning {
    int xcd = 1;          // XCD-aware workgroup order in the 1-D fused grids
    int ks_mac_inst = 4;  // instances sharing one load of the key rows in ks_mac (2, 4 or 8)
    int re_tensor = 4;    // log2 values per thread of tensor_intt (4 or 5)
    int re_lift = 4;      // of lift_ntt
    int re_finish = 4;    // of ks_finish
    // Stages that run as half rows (two workgroups per CU) in launches of any
    // size at N = 4096 .. 16384 (without the bit, only launches of at most
    // split_row_limit() rows do): HALVES_FINISH the key-switch finish,
    // HALVES_INPUT the multiplication's input inverse transforms, extension
    // and forward transform (the stages pairing the halves in a paired
    // scaler), HALVES_TENSOR the tensor product + inverse transform and the
    // down-scaling (likewise).
    int halves = HALVES_FINISH;
    // The column sums of the extension and down-scaling on the matrix cores
    // (scale_mfma.hip) where the scaler's shape allows it (1), or the
    // per-column VALU kernels (0).
    int scaler_mfma = 1;
};

// The current choice for degree 2^logn (the defaults until set).
FusedTuning fused_tuning(int logn);
// Validates and stores a choice for degree 2^logn.
void set_fused_tuning(int logn, const FusedTuning& t);

}  // namespace fhehip
