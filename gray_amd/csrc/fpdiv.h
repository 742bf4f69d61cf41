/*
 * fpdiv.h -- correctly rounded FP64 division by 6 without the division
 * sequence (the RK4 combination of the bit-exact FP64 forms), and the
 * compiler's correctly rounded FP64 square root and division without their
 * range steps, for operands where those steps change nothing.
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

/* For gfx9 the compiler lowers the correctly rounded sqrt(x) to: scale x by
 * 2^256 when x < 2^-767; y = v_rsq(x); g = x y, h = y / 2; r = 1/2 - h g,
 * g += g r, h += h r; twice d = x - g^2, g += d h; scale g back; and x
 * itself where x is +-0 or +inf.  For a finite x >= 2^-767 the scalings are
 * by 2^0 and the last select keeps g, so sqrt_core, the same instructions
 * without them, returns the same bits. */
__device__ __forceinline__ double sqrt_core(double x)
{
	const double y = __builtin_amdgcn_rsq(x);
	double g = x * y, h = y * 0.5;
	const double r = __builtin_fma(-h, g, 0.5);
	g = __builtin_fma(g, r, g);
	h = __builtin_fma(h, r, h);
	g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
	return __builtin_fma(__builtin_fma(-g, g, x), h, g);
}

/* The correctly rounded a / b is lowered to: b' = v_div_scale(b), a' =
 * v_div_scale(a); y = v_rcp(b'), twice y += y (1 - b' y); q = a' y;
 * q = v_div_fmas(a' - b' q, y, q); v_div_fixup(q, b, a).  v_div_scale leaves
 * its operand unchanged and clears the flag that v_div_fmas scales by unless
 * b is subnormal, 1/b or a/b is subnormal, exp(a) - exp(b) >= 768, or
 * |a| < 2^-969 with a != 0; v_div_fixup returns q with the quotient's sign
 * for finite normal operands and quotient, and the zero of the quotient's
 * sign for a = +0 and a finite normal b, as does div_core, the same
 * instructions without those steps (not so for a = -0).  Callers guarantee
 * operands outside all of those cases. */
__device__ __forceinline__ double div_core(double a, double b)
{
	double y = __builtin_amdgcn_rcp(b);
	y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
	y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
	const double q = a * y;
	return __builtin_fma(__builtin_fma(-b, q, a), y, q);
}

/* lo <= x <= hi (false for NaN). */
__device__ __forceinline__ bool in_range(double x, double lo, double hi)
{
	return x >= lo && x <= hi;
}

} // namespace gray
