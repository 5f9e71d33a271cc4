// Allocation-free evaluator (scratch in a file-scope device array).
#pragma once
#include "ctx.h"

namespace pggc {

// resolves the scratch array's address on `device` (which must be current)
void noalloc_init(int device);
i64 noalloc_max_nnz();
i64 noalloc_scratch_bytes();
// pggc_evaluate without any device allocation: vol[L] (zeroed here), uncut
// and the 128-bit sum of squared volumes (high, low) on the host
void evaluate_noalloc(Ctx& c, const DGraph& g, const i32* labels, i64 L, i64* vol, i64* uncut_out, u64* sq_out);

}  // namespace pggc
