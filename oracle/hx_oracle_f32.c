/* hx_oracle_f32.c -- the oracle's restatement (hx_oracle.c) compiled in
 * single precision.  TEST INFRASTRUCTURE ONLY: the checker of the FP32
 * build of the fused advance, which computes in float throughout; the tests
 * load it, nothing on the product path does.
 *
 * Every double of hx_oracle.c becomes float, every floating literal a float
 * (built with -fsingle-precision-constant) and every libm call its float
 * form, so each operation rounds to single precision in the order
 * hx_oracle.c states it (include/solvers/kernels/Roe.H:50-309,
 * src/solvers/HydrostaticRoe.cpp:173-606, as cited there).  The reference's
 * own SINGLE build (cmake/presets/01-precision.json:21) keeps double
 * literals, so C++ promotes some of its subexpressions to double: this file
 * restates the algorithm in pure single precision, the arithmetic the FP32
 * kernels use.  Only hxo_run is bound (oracle.run32).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define double float
#define sqrt sqrtf
#define cbrt cbrtf
#define fabs fabsf

#include "hx_oracle.c"
