// sk_fast.hip — register-resident row and column passes for the lengths that have them (filled in below).
#include <hip/hip_runtime.h>

#include "sk_internal.h"

namespace sk {

bool fast_row_supported(int, int, int) { return false; }
bool fast_col_supported(int, int, int) { return false; }
void launch_row_fast(const RowArgs&, int, hipStream_t) { fail(SK_ERR_DEVICE, "no fast row pass"); }
void launch_col_fast(const ColArgs&, int, hipStream_t) { fail(SK_ERR_DEVICE, "no fast column pass"); }
int fast_col_blocks_per_wave(int, int) { return 0; }

}  // namespace sk
