// Pin for the oracle's directed-rounding square root: the x86-64 sqrtsd
// instruction is correctly rounded in the current MXCSR rounding mode, so
// sqrt() under FE_DOWNWARD is the IEEE-754 round-toward-minus-infinity root.
// TEST INFRASTRUCTURE ONLY.  Compiled with -frounding-math so that the
// compiler does not fold or reorder the operation across the mode switch.
#include <cfenv>
#include <cmath>
#include "oracle.h"

extern "C" double oracle_sqrt_rd_fenv(double x) {
    const int old = std::fegetround();
    std::fesetround(FE_DOWNWARD);
    volatile double v = x;
    volatile double s = std::sqrt(v);
    std::fesetround(old);
    return s;
}
