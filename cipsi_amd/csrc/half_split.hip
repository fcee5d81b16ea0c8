// Split-FP16 (Ozaki) half-transform.  Placeholder until the slice kernels land:
// every request for CIPSI_PREC_SPLIT_FP16 reports CIPSI_ERR_UNSUPPORTED.
#include <cmath>

#include "../../include/cipsi_ijkl.h"
#include "ctx.h"

bool split_supported(int) { return false; }
int64_t split_workspace_bytes(int, int) { return 0; }
double split_unit_roundoff(int, int) { return std::ldexp(1.0, -53); }
int run_half_split(cipsi_ijkl_ctx*, const cipsi::HalfArgs&, int, int) { return CIPSI_ERR_UNSUPPORTED; }
