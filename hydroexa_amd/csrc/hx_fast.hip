// hx_fast.hip -- fast arithmetic build of the fused advance: divisions that
// share a denominator become one Newton-refined reciprocal, and a product
// feeding a sum in the same expression becomes one FMA (compiled with
// -ffp-contract=on: contraction within an expression only, so that every
// inlined copy of a face rounds identically).  Selected at run
// time with hx_set_arith(ctx, HX_ARITH_FAST); the parity tests hold it to
// 1e-12 relative L-infinity against the FP64 oracle.
#include <hip/hip_runtime.h>

#include "hx_types.hpp"
#include "hx_adjoint.hpp"

namespace hx {
#define HX_NS fast
#define HX_FAST 1
#include "hx_device.hpp"
#include "hx_advance.inc"
#undef HX_NS
#undef HX_FAST
}  // namespace hx
