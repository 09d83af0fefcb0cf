// detmath.h - deterministic elementary functions for fp32 shading code.
//
// The wavefront tracer and its CPU checker must take identical path decisions
// (Russian roulette, lobe and light selection all compare a float against a
// uniform sample), so every transcendental on the shading path has to return
// the same bits on gfx950 and on x86-64.  Neither libm nor the HIP device
// library promises that, so the functions below are written out from their
// textbook definitions and evaluated in IEEE binary64 with only +, -, *, /
// (compiled with contraction off), then rounded once to binary32:
//
//  * sin/cos: Cody-Waite argument reduction by pi/2 with a two-part constant
//    (W. J. Cody and W. Waite, "Software Manual for the Elementary Functions",
//    1980) and the Taylor series to degree 15/16 on |r| <= pi/4
//    (truncation error < 5e-17);
//  * log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//    s = (m-1)/(m+1), series to s^21 (|s| < 0.172, error < 1e-19);
//  * exp: exp(x) = 2^k exp(r), r = x - k ln2 (Cody-Waite two-part ln2),
//    Taylor series of exp(r) to degree 13 on |r| <= ln2/2;
//  * atan: reduction atan(t) = pi/2 - atan(1/t) and atan(t) = pi/6 +
//    atan((t sqrt3 - 1)/(t + sqrt3)) to |u| <= tan(pi/12), series to u^25.
//
// Every binary64 result is within about one binary64 ulp of the exact value,
// so the final rounding to binary32 is correct except in ties closer than
// ~2^-52 relative; the tests check the float results against libm double.
// sqrt and division are IEEE-exact on both sides already (correctly rounded
// f32 division/sqrt is the HIP default) and are not wrapped.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define DM_HD __host__ __device__
#else
#define DM_HD
#endif

namespace detmath {

DM_HD inline uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
DM_HD inline double double_of(uint64_t u) { return __builtin_bit_cast(double, u); }
DM_HD inline uint32_t bits_of(float x) { return __builtin_bit_cast(uint32_t, x); }
DM_HD inline float float_of(uint32_t u) { return __builtin_bit_cast(float, u); }

DM_HD inline bool is_nan(double x) { return x != x; }
DM_HD inline bool is_inf(double x) { return x == __builtin_inf() || x == -__builtin_inf(); }

// Nearest integer of a finite double of moderate magnitude (|x| < 2^52).
DM_HD inline double round_nearest(double x) {
	double f = (double) (int64_t) x;   // truncation toward zero
	double d = x - f;
	if (d >= 0.5) f += 1.0;
	else if (d <= -0.5) f -= 1.0;
	return f;
}

// sin/cos of r with |r| <= pi/4 (binary64).
DM_HD inline double sin_kernel(double r) {
	const double r2 = r * r;
	double p		= -1.0 / 1307674368000.0;   // -1/15!
	p				= p * r2 + 1.0 / 6227020800.0;  // 1/13!
	p				= p * r2 - 1.0 / 39916800.0;    // -1/11!
	p				= p * r2 + 1.0 / 362880.0;      // 1/9!
	p				= p * r2 - 1.0 / 5040.0;        // -1/7!
	p				= p * r2 + 1.0 / 120.0;         // 1/5!
	p				= p * r2 - 1.0 / 6.0;           // -1/3!
	return r + r * r2 * p;
}

DM_HD inline double cos_kernel(double r) {
	const double r2 = r * r;
	double p		= 1.0 / 20922789888000.0;      // 1/16!
	p				= p * r2 - 1.0 / 87178291200.0; // -1/14!
	p				= p * r2 + 1.0 / 479001600.0;   // 1/12!
	p				= p * r2 - 1.0 / 3628800.0;     // -1/10!
	p				= p * r2 + 1.0 / 40320.0;       // 1/8!
	p				= p * r2 - 1.0 / 720.0;         // -1/6!
	p				= p * r2 + 1.0 / 24.0;          // 1/4!
	p				= p * r2 - 0.5;                 // -1/2!
	return 1.0 + r2 * p;
}

// Reduce x to r in [-pi/4, pi/4] and quadrant q: x = q*pi/2 + r.
// Valid for |x| < 2^20 (k * PIO2_HI exact); shading arguments are < 100.
DM_HD inline double reduce_pio2(double x, int *q) {
	const double TWO_OVER_PI = 0.636619772367581343075535;
	const double PIO2_HI	 = 1.57079632673412561417e+00; // first 33 bits of pi/2
	const double PIO2_LO	 = 6.07710050650619224932e-11; // pi/2 - PIO2_HI
	double k				 = round_nearest(x * TWO_OVER_PI);
	*q						 = (int) ((int64_t) k & 3);
	return (x - k * PIO2_HI) - k * PIO2_LO;
}

DM_HD inline void sincos_d(double x, double *s, double *c) {
	int q;
	double r  = reduce_pio2(x, &q);
	double sr = sin_kernel(r), cr = cos_kernel(r);
	switch (q) {
		case 0: *s = sr; *c = cr; break;
		case 1: *s = cr; *c = -sr; break;
		case 2: *s = -sr; *c = -cr; break;
		default: *s = -cr; *c = sr; break;
	}
}

DM_HD inline float sinf(float x) {
	double s, c;
	sincos_d((double) x, &s, &c);
	return (float) s;
}

DM_HD inline float cosf(float x) {
	double s, c;
	sincos_d((double) x, &s, &c);
	return (float) c;
}

DM_HD inline void sincosf(float x, float *s, float *c) {
	double sd, cd;
	sincos_d((double) x, &sd, &cd);
	*s = (float) sd;
	*c = (float) cd;
}

DM_HD inline float tanf(float x) {
	double s, c;
	sincos_d((double) x, &s, &c);
	return (float) (s / c);
}

DM_HD inline double log_d(double x) {
	if (is_nan(x) || x < 0.0) return __builtin_nan("");
	if (x == 0.0) return -__builtin_inf();
	if (is_inf(x)) return x;
	const double LN2_HI = 6.93147180369123816490e-01;
	const double LN2_LO = 1.90821492927058770002e-10;
	uint64_t u			= bits_of(x);
	int e				= (int) ((u >> 52) & 0x7ff);
	if (e == 0) { // subnormal binary64: scale by 2^54 first
		x *= 18014398509481984.0;
		u = bits_of(x);
		e = (int) ((u >> 52) & 0x7ff) - 54;
	}
	e -= 1023;
	double m = double_of((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull); // [1,2)
	if (m > 1.41421356237309504880) {
		m *= 0.5;
		e += 1;
	}
	const double s	= (m - 1.0) / (m + 1.0);
	const double s2 = s * s;
	double p		= 1.0 / 21.0;
	p				= p * s2 + 1.0 / 19.0;
	p				= p * s2 + 1.0 / 17.0;
	p				= p * s2 + 1.0 / 15.0;
	p				= p * s2 + 1.0 / 13.0;
	p				= p * s2 + 1.0 / 11.0;
	p				= p * s2 + 1.0 / 9.0;
	p				= p * s2 + 1.0 / 7.0;
	p				= p * s2 + 1.0 / 5.0;
	p				= p * s2 + 1.0 / 3.0;
	const double logm = 2.0 * s + 2.0 * s * s2 * p;
	const double de	  = (double) e;
	return de * LN2_HI + (de * LN2_LO + logm);
}

DM_HD inline float logf(float x) { return (float) log_d((double) x); }

DM_HD inline double exp_d(double x) {
	if (is_nan(x)) return x;
	if (x > 709.0) return __builtin_inf();
	if (x < -745.0) return 0.0;
	const double INV_LN2 = 1.44269504088896338700;
	const double LN2_HI	 = 6.93147180369123816490e-01;
	const double LN2_LO	 = 1.90821492927058770002e-10;
	double k			 = round_nearest(x * INV_LN2);
	double r			 = (x - k * LN2_HI) - k * LN2_LO;
	double p			 = 1.0 / 6227020800.0; // 1/13!
	p					 = p * r + 1.0 / 479001600.0;
	p					 = p * r + 1.0 / 39916800.0;
	p					 = p * r + 1.0 / 3628800.0;
	p					 = p * r + 1.0 / 362880.0;
	p					 = p * r + 1.0 / 40320.0;
	p					 = p * r + 1.0 / 5040.0;
	p					 = p * r + 1.0 / 720.0;
	p					 = p * r + 1.0 / 120.0;
	p					 = p * r + 1.0 / 24.0;
	p					 = p * r + 1.0 / 6.0;
	p					 = p * r + 0.5;
	p					 = p * r + 1.0;
	p					 = p * r + 1.0;
	int ki				 = (int) k;
	// 2^k as two factors so that k down to -1074 and up to 1023 stays exact.
	int k1 = ki / 2, k2 = ki - k1;
	double s1 = double_of((uint64_t) (k1 + 1023) << 52);
	double s2 = double_of((uint64_t) (k2 + 1023) << 52);
	return p * s1 * s2;
}

DM_HD inline float expf(float x) {
	if (x > 89.0f) return __builtin_inff();
	if (x < -104.0f) return 0.0f;
	return (float) exp_d((double) x);
}

// atan of t >= 0 (binary64).
DM_HD inline double atan_pos(double t) {
	const double PIO2			= 1.57079632679489661923;
	const double PIO6			= 0.52359877559829887308;
	const double SQRT3			= 1.73205080756887729353;
	const double TAN_PI_OVER_12 = 0.26794919243112270647;
	bool inv					= false;
	if (t > 1.0) {
		t	= 1.0 / t;
		inv = true;
	}
	double add = 0.0;
	if (t > TAN_PI_OVER_12) {
		t	= (t * SQRT3 - 1.0) / (t + SQRT3);
		add = PIO6;
	}
	const double t2 = t * t;
	double p		= 1.0 / 25.0;
	p				= p * t2 - 1.0 / 23.0;
	p				= p * t2 + 1.0 / 21.0;
	p				= p * t2 - 1.0 / 19.0;
	p				= p * t2 + 1.0 / 17.0;
	p				= p * t2 - 1.0 / 15.0;
	p				= p * t2 + 1.0 / 13.0;
	p				= p * t2 - 1.0 / 11.0;
	p				= p * t2 + 1.0 / 9.0;
	p				= p * t2 - 1.0 / 7.0;
	p				= p * t2 + 1.0 / 5.0;
	p				= p * t2 - 1.0 / 3.0;
	double a		= add + (t + t * t2 * p);
	return inv ? PIO2 - a : a;
}

DM_HD inline float atan2f(float yf, float xf) {
	const double PI = 3.14159265358979323846;
	double y = yf, x = xf;
	if (is_nan(x) || is_nan(y)) return __builtin_nanf("");
	if (x == 0.0 && y == 0.0) return (bits_of(xf) >> 31) ? (float) __builtin_copysign(PI, y) : yf;
	double ay = y < 0 ? -y : y, ax = x < 0 ? -x : x;
	double a;
	if (ax == 0.0) a = 1.57079632679489661923;
	else if (is_inf(ay) && is_inf(ax)) a = 0.78539816339744830962;
	else a = atan_pos(ay / ax);
	if (x < 0) a = PI - a;
	return (float) (y < 0 ? -a : a);
}

// acos(x) for x in [-1, 1] (clamped), from the binary64 arctangent:
// acos(x) = atan(sqrt(1 - x^2) / x), reflected for x < 0.
DM_HD inline float acosf(float xf) {
	const double PI = 3.14159265358979323846;
	if (is_nan((double) xf)) return __builtin_nanf("");
	double x = xf;
	if (x > 1.0) x = 1.0;
	if (x < -1.0) x = -1.0;
	const double ax = x < 0 ? -x : x;
	const double s	= __builtin_sqrt(1.0 - x * x);
	const double a	= ax == 0.0 ? 1.57079632679489661923 : atan_pos(s / ax);
	return (float) (x < 0 ? PI - a : a);
}

} // namespace detmath
