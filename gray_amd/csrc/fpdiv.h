/*
 * fpdiv.h -- correctly rounded FP64 division by 6 without the division
 * sequence (the RK4 combination of the bit-exact FP64 forms).
 */
#pragma once

#include <hip/hip_runtime.h>

namespace gray {

/* x / 6, correctly rounded.  In the normal range: q = x RN(1/6) is within
 * one ulp of x/6, r = x - 6 q is exact with an FMA, and q + r RN(1/6) rounds
 * to the correctly rounded quotient (Markstein's correction theorem); zero
 * returns itself (keeping its sign), and tiny, huge and non-finite x take
 * the division. */
__device__ __forceinline__ double div6(double x)
{
	const double y = 1.0 / 6.0;
	const double ax = fabs(x);
	if (ax >= 0x1p-960 && ax <= 0x1p1000) {
		const double q = x * y;
		return __builtin_fma(__builtin_fma(-6.0, q, x), y, q);
	}
	if (x == 0.0)
		return x;
	return x / 6.0;
}

} // namespace gray
