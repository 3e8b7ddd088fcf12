// Deterministic single-precision transcendentals for the catalog hot path.
//
// The reference calls the platform's expf/logf/tanf/powf/exp10f/log10f inside
// the Poisson and Gaussian samplers (common/cuda_rng.h:165-376), the density
// models (skygen/model_expDisk.h:113, model_powerLawEllipsoid.h:93) and the
// per-cell distance (skygen/skygen.cu.h:222, 408).  Rejection loops compare
// those results against uniforms, so integer star counts depend on the last
// bit of every transcendental.  To make the GPU path and a CPU restatement
// agree bit for bit, every such function here is built only from +, *, /,
// sqrt, fma and rint, which are correctly rounded IEEE operations on gfx950
// and on x86-64, so the results are identical on both:
//
//   expf, logf (the samplers, the density models): float arithmetic, see
//         gf_expf / gf_logf; at most 1 ulp from the correctly rounded value.
//   exp10f (and the distance modulus's 10^(0.2 DM + 1)): float arithmetic,
//         see gf_pow10f; at most 1 ulp.
//   log10f, powf, sinf, cosf, tanf: evaluated in IEEE double and rounded
//         once to float (the argument scaling of pow needs the extra
//         precision):
//     exp:  Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, Taylor series
//           of e^r to degree 10 (truncation < 3e-13 relative), scaling by 2^k.
//     log:  x = 2^e * m, m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//           log m = 2*atanh(s) = 2s * sum z^n/(2n+1), z = s^2 <= 0.0295,
//           summed to n = 7 (truncation < 4e-14 relative).
//     sin/cos: reduction by pi/2 (two-part constant, fma), Taylor series to
//           degree 17/18 on |r| <= pi/4 (truncation < 1e-18).
//     The float result is the correctly rounded one except within 2^-40 of
//     a rounding boundary.
// The double-precision sin/cos/atan/atan2/asin here (a few 1e-16 relative)
// replace the platform's in the deprojection, direction cosines and sky
// rotations for the same reason: the GPU and a CPU evaluation agree to the
// bit.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GF_HD __host__ __device__ inline
#else
#include <math.h>
#include <string.h>
#define GF_HD inline
#endif

namespace gfm {

GF_HD double as_double(uint64_t u) { double d; __builtin_memcpy(&d, &u, 8); return d; }
GF_HD uint64_t as_u64(double d) { uint64_t u; __builtin_memcpy(&u, &d, 8); return u; }
GF_HD float as_float(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
GF_HD uint32_t as_u32(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }

GF_HD double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
GF_HD double drint(double a) { return __builtin_rint(a); }

static constexpr double LN2_HI = 6.93147180369123816490e-01;   // ln2 rounded to 32 significant bits
static constexpr double LN2_LO = 1.90821492927058770002e-10;   // ln2 - LN2_HI
static constexpr double INV_LN2 = 1.44269504088896338700e+00;
static constexpr double LN10 = 2.30258509299404590109e+00;
static constexpr double INV_LN10 = 4.34294481903251816668e-01;
static constexpr double PIO2_HI = 1.57079632679489655800e+00;   // double nearest pi/2
static constexpr double PIO2_LO = 6.12323399573676603587e-17;   // pi/2 - PIO2_HI
static constexpr double INV_PIO2 = 6.36619772367581382433e-01;

// 2^k for integer k in [-1022, 1023]
GF_HD double pow2i(int k) { return as_double((uint64_t)(k + 1023) << 52); }

// e^x in double
GF_HD double exp(double x)
{
	if (x != x) return x;
	if (x > 709.0) return as_double(0x7ff0000000000000ull);
	if (x < -745.0) return 0.0;
	double kd = drint(x * INV_LN2);
	double r = dfma(-kd, LN2_HI, x);
	r = dfma(-kd, LN2_LO, r);
	double p = 1.0 / 3628800.0;               // 1/10!
	p = dfma(p, r, 1.0 / 362880.0);           // 1/9!
	p = dfma(p, r, 1.0 / 40320.0);            // 1/8!
	p = dfma(p, r, 1.0 / 5040.0);             // 1/7!
	p = dfma(p, r, 1.0 / 720.0);              // 1/6!
	p = dfma(p, r, 1.0 / 120.0);              // 1/5!
	p = dfma(p, r, 1.0 / 24.0);               // 1/4!
	p = dfma(p, r, 1.0 / 6.0);                // 1/3!
	p = dfma(p, r, 0.5);                      // 1/2!
	p = dfma(p, r, 1.0);
	p = dfma(p, r, 1.0);
	int k = (int)kd;
	if (k > 1023) { p *= pow2i(1023); k -= 1023; }
	if (k < -1022) { p *= pow2i(-1022); k += 1022; }
	return p * pow2i(k);
}

// natural log in double
GF_HD double log(double x)
{
	if (x != x || x < 0.0) return as_double(0x7ff8000000000000ull);
	if (x == 0.0) return -as_double(0x7ff0000000000000ull);
	uint64_t u = as_u64(x);
	int e = (int)((u >> 52) & 0x7ff);
	if (e == 0x7ff) return x;   // +inf
	if (e == 0) { x *= pow2i(54); u = as_u64(x); e = (int)((u >> 52) & 0x7ff) - 54; }
	e -= 1023;
	double m = as_double((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);   // [1,2)
	if (m > 1.41421356237309504880) { m *= 0.5; e += 1; }
	double s = (m - 1.0) / (m + 1.0);
	double z = s * s;
	double p = 1.0 / 15.0;
	p = dfma(p, z, 1.0 / 13.0);
	p = dfma(p, z, 1.0 / 11.0);
	p = dfma(p, z, 1.0 / 9.0);
	p = dfma(p, z, 1.0 / 7.0);
	p = dfma(p, z, 1.0 / 5.0);
	p = dfma(p, z, 1.0 / 3.0);
	p = dfma(p, z, 1.0);
	double lm = 2.0 * s * p;
	double ed = (double)e;
	return dfma(ed, LN2_HI, dfma(ed, LN2_LO, lm));
}

// sin and cos of x in double (|x| < 1e6)
GF_HD void sincos(double x, double *sn, double *cs)
{
	double kd = drint(x * INV_PIO2);
	double r = dfma(-kd, PIO2_HI, x);
	r = dfma(-kd, PIO2_LO, r);
	double z = r * r;
	// sin r = r * (1 - z/3! + z^2/5! - ... ), to r^17/17!
	double ps = -1.0 / 355687428096000.0;     // -1/17!
	ps = dfma(ps, z, 1.0 / 1307674368000.0);  // 1/15!
	ps = dfma(ps, z, -1.0 / 6227020800.0);    // -1/13!
	ps = dfma(ps, z, 1.0 / 39916800.0);       // 1/11!
	ps = dfma(ps, z, -1.0 / 362880.0);        // -1/9!
	ps = dfma(ps, z, 1.0 / 5040.0);           // 1/7!
	ps = dfma(ps, z, -1.0 / 120.0);           // -1/5!
	ps = dfma(ps, z, 1.0 / 6.0);              // 1/3!  (subtracted below)
	double s = dfma(-r * z, ps, r);
	// cos r = 1 - z/2! + z^2/4! - ..., to r^18/18!
	double pc = 1.0 / 6402373705728000.0;     // 1/18!
	pc = dfma(pc, z, -1.0 / 20922789888000.0);// -1/16!
	pc = dfma(pc, z, 1.0 / 87178291200.0);    // 1/14!
	pc = dfma(pc, z, -1.0 / 479001600.0);     // -1/12!
	pc = dfma(pc, z, 1.0 / 3628800.0);        // 1/10!
	pc = dfma(pc, z, -1.0 / 40320.0);         // -1/8!
	pc = dfma(pc, z, 1.0 / 720.0);            // 1/6!
	pc = dfma(pc, z, -1.0 / 24.0);            // -1/4!
	pc = dfma(pc, z, 0.5);                    // 1/2!  (subtracted below)
	double c = dfma(-z, pc, 1.0);
	int q = ((int)(int64_t)kd) & 3;
	switch (q) {
	case 0: *sn = s; *cs = c; break;
	case 1: *sn = c; *cs = -s; break;
	case 2: *sn = -s; *cs = -c; break;
	default: *sn = -c; *cs = s; break;
	}
}

GF_HD double sin(double x) { double s, c; sincos(x, &s, &c); return s; }
GF_HD double cos(double x) { double s, c; sincos(x, &s, &c); return c; }

// atan(k/8), k = 0..8, each rounded to the nearest double (computed with
// 200-bit arithmetic).
static constexpr double ATAN_K8[9] = {
	0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2, 0x1.dac670561bb4fp-2,
	0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1, 0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1};

// atan t for 0 <= t <= 1: c = k/8 nearest t, atan t = atan c + atan u with
// u = (t - c) / (1 + t c) (t - c is exact, |u| <= 1/16), atan u by its
// Taylor series to u^13/13 (truncation < 6e-20).  One division.
GF_HD double atan_unit(double t)
{
	double kd = drint(t * 8.0);
	double c = kd * 0.125;
	double u = (t - c) / dfma(t, c, 1.0);
	double z = u * u;
	double p = 1.0 / 13.0;
	p = dfma(p, z, -1.0 / 11.0);
	p = dfma(p, z, 1.0 / 9.0);
	p = dfma(p, z, -1.0 / 7.0);
	p = dfma(p, z, 1.0 / 5.0);
	p = dfma(p, z, -1.0 / 3.0);
	return ATAN_K8[(int)kd] + dfma(u * z, p, u);
}

// atan(x) for double x; |x| > 1 uses pi/2 - atan(1/x).
GF_HD double atan(double x)
{
	if (x != x) return x;
	double ax = x < 0 ? -x : x;
	bool inv = ax > 1.0;
	double a = atan_unit(inv ? 1.0 / ax : ax);
	if (inv) a = (PIO2_HI - a) + PIO2_LO;
	return x < 0 ? -a : a;
}

// atan2(y, x) with the usual quadrant conventions for finite arguments.
GF_HD double atan2(double y, double x)
{
	static constexpr double PI_HI = 3.14159265358979311600e+00, PI_LO = 1.22464679914735317720e-16;
	if (x != x || y != y) return x + y;
	if (x == 0.0 && y == 0.0) {
		bool xneg = __builtin_signbit(x);
		double r = xneg ? PI_HI : 0.0;
		return __builtin_signbit(y) ? -r : r;
	}
	double ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
	double a;
	if (ay <= ax) {
		a = atan_unit(ay / ax);
	} else {
		a = (PIO2_HI - atan_unit(ax / ay)) + PIO2_LO;
	}
	if (x < 0) a = (PI_HI - a) + PI_LO;
	return y < 0 || (y == 0.0 && __builtin_signbit(y)) ? -a : a;
}

// asin(x) = atan2(x, sqrt((1 - x)(1 + x))) for |x| <= 1.
GF_HD double asin(double x)
{
	if (!(x >= -1.0 && x <= 1.0)) return as_double(0x7ff8000000000000ull);
	return atan2(x, __builtin_sqrt((1.0 - x) * (1.0 + x)));
}

} // namespace gfm

// Single-precision entry points used by the kernels.
// e^x in float arithmetic: k = rint(x / ln2), r = x - k ln2 with ln2 split
// into a 16-bit head (k * head is exact) and a tail, e^r by its Taylor
// series to r^7/7! (|r| <= ln2/2), scaled by 2^k.  At most 1 ulp from the
// correctly rounded value.
GF_HD float gf_expf(float x)
{
	const float LN2_HI = 0.693145751953125f, LN2_LO = 1.428606765330187e-06f;
	if (x != x) return x;
	if (x > 88.72283935546875f) return gfm::as_float(0x7f800000u);
	if (x < -103.97208404541016f) return 0.f;
	float kf = __builtin_rintf(x * 1.44269502162933349609375f);
	float r = __builtin_fmaf(-kf, LN2_HI, x);
	r = __builtin_fmaf(-kf, LN2_LO, r);
	float p = 1.f / 5040.f;
	p = __builtin_fmaf(p, r, 1.f / 720.f);
	p = __builtin_fmaf(p, r, 1.f / 120.f);
	p = __builtin_fmaf(p, r, 1.f / 24.f);
	p = __builtin_fmaf(p, r, 1.f / 6.f);
	p = __builtin_fmaf(p, r, 0.5f);
	p = __builtin_fmaf(p, r, 1.f);
	p = __builtin_fmaf(p, r, 1.f);
	int k = (int)kf;
	if (k > 127) { p *= gfm::as_float(0x7f000000u); k -= 127; }
	if (k < -126) { p *= gfm::as_float(0x00800000u); k += 126; }
	return p * gfm::as_float((uint32_t)(k + 127) << 23);
}

// log x in float arithmetic: x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m-1
// (exact), s = f/(2+f), log(1+f) = f - (f^2/2 - s (f^2/2 + R)) with
// R = 2 atanh(s)/s - 2 = z (2/3 + 2z/5 + ... + 2z^4/11), z = s^2 <= 0.0295,
// plus e ln2 (head exact, tail folded into the correction).  At most 1 ulp.
GF_HD float gf_logf(float x)
{
	const float LN2_HI = 0.693145751953125f, LN2_LO = 1.428606765330187e-06f;
	if (x != x || x < 0.f) return gfm::as_float(0x7fc00000u);
	if (x == 0.f) return gfm::as_float(0xff800000u);
	uint32_t ix = gfm::as_u32(x);
	if ((ix >> 23) == 0xffu) return x;
	int e = 0;
	if ((ix >> 23) == 0u) { x *= gfm::as_float(0x4c000000u); ix = gfm::as_u32(x); e = -25; }
	e += (int)(ix >> 23) - 127;
	float m = gfm::as_float((ix & 0x7fffffu) | 0x3f800000u);
	if (m > 1.41421353816986083984375f) { m *= 0.5f; e += 1; }
	float f = m - 1.f;
	float hfsq = 0.5f * f * f;
	float s = f / (2.f + f);
	float z = s * s;
	float R = 2.f / 11.f;
	R = __builtin_fmaf(R, z, 2.f / 9.f);
	R = __builtin_fmaf(R, z, 2.f / 7.f);
	R = __builtin_fmaf(R, z, 2.f / 5.f);
	R = __builtin_fmaf(R, z, 2.f / 3.f);
	R = R * z;
	float ef = (float)e;
	return __builtin_fmaf(ef, LN2_HI, -((hfsq - __builtin_fmaf(s, hfsq + R, ef * LN2_LO)) - f));
}
// 10^y in float arithmetic (Cody and Waite's argument reduction): y log2(10)
// = k + r with k = rint(y Lh), r = (y Lh - k) + y Ll, log2(10) split into a
// float head Lh and tail Ll (the fma forms y Lh - k with one rounding);
// 2^r = e^(r ln 2) by its Taylor series to degree 7 (|r ln 2| <= 0.35:
// truncation below 1e-8 relative); times 2^k.  Within 1 ulp of the
// correctly rounded value.
GF_HD float gf_pow10f(float y)
{
	const float Lh = 0x1.a934fp+1f, Ll = 0x1.2f346ep-24f;
	if (y != y) return y;
	if (y > 38.5318413f) return gfm::as_float(0x7f800000u);
	if (y < -45.3534698f) return 0.f;
	const float k = __builtin_rintf(y * Lh);
	float r = __builtin_fmaf(y, Lh, -k);
	r = __builtin_fmaf(y, Ll, r);
	float p = 0x1.ffcbfcp-17f;                 // ln2^7 / 7!
	p = __builtin_fmaf(p, r, 0x1.430912p-13f); // ln2^6 / 6!
	p = __builtin_fmaf(p, r, 0x1.5d87fep-10f); // ln2^5 / 5!
	p = __builtin_fmaf(p, r, 0x1.3b2ab6p-7f);  // ln2^4 / 4!
	p = __builtin_fmaf(p, r, 0x1.c6b08ep-5f);  // ln2^3 / 3!
	p = __builtin_fmaf(p, r, 0x1.ebfbe0p-3f);  // ln2^2 / 2!
	p = __builtin_fmaf(p, r, 0x1.62e430p-1f);  // ln2
	p = __builtin_fmaf(p, r, 1.f);
	int ki = (int)k;
	if (ki > 127) { p *= gfm::as_float(0x7f000000u); ki -= 127; }
	if (ki < -126) { p *= gfm::as_float(0x00800000u); ki += 126; }
	return p * gfm::as_float((uint32_t)(ki + 127) << 23);
}
GF_HD float gf_exp10f(float x) { return gf_pow10f(x); }
GF_HD float gf_log10f(float x) { return (float)(gfm::log((double)x) * gfm::INV_LN10); }
GF_HD float gf_sinf(float x) { double s, c; gfm::sincos((double)x, &s, &c); return (float)s; }
GF_HD float gf_cosf(float x) { double s, c; gfm::sincos((double)x, &s, &c); return (float)c; }
GF_HD float gf_tanf(float x) { double s, c; gfm::sincos((double)x, &s, &c); return (float)(s / c); }
GF_HD float gf_sqrtf(float x) { return __builtin_sqrtf(x); }
GF_HD float gf_fabsf(float x) { return __builtin_fabsf(x); }

// x^y for float x >= 0 (and NaN), with log(x) supplied by logx() only
// when it is needed: callers that raise one x to several powers pass a
// cached log and get exactly gf_powf's results.
template <typename LogX>
GF_HD float gf_powf_nonneg(float x, float y, LogX logx)
{
	// y first and alone: a kernel's exponents are the same for every lane,
	// so this test is a scalar branch there
	if (y == 0.0f) return 1.0f;
	if (x == 1.0f) return 1.0f;
	if (x != x || y != y) return x + y;
	if (x == 0.0f) return y > 0.0f ? 0.0f : gfm::as_float(0x7f800000u);
	return (float)gfm::exp((double)y * logx());
}

// x^y for float x >= 0 (and the integer-y cases of x < 0)
GF_HD float gf_powf(float x, float y)
{
	if (x < 0.0f && y != 0.0f && y == y) {
		float yi = __builtin_truncf(y);
		if (yi != y) return gfm::as_float(0x7fc00000u);
		double r = gfm::exp((double)y * gfm::log(-(double)x));
		int64_t iy = (int64_t)yi;
		return (float)((iy & 1) ? -r : r);
	}
	return gf_powf_nonneg(x, y, [x]() { return gfm::log((double)x); });
}
