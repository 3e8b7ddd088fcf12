/*
 * ORACLE -- test infrastructure only.  Never linked into the product.
 *
 * Single-precision transcendentals of the oracle.  The reference calls the
 * platform expf/logf/tanf/powf/exp10f/log10f/sinf/cosf (common/cuda_rng.h,
 * skygen/model_expDisk.h and the module kernels).  The oracle evaluates each one in
 * IEEE double and rounds once to float:
 *   expf, logf: in float arithmetic (see gfo_expf / gfo_logf), <= 1 ulp;
 *   the others in IEEE double, rounded once to float:
 *   exp: k = rint(x/ln2), r = x - k*ln2 (two-part ln2, fma), e^r by its
 *        Taylor series to r^10/10!, times 2^k;
 *   log: x = 2^e m, m in [1/sqrt2, sqrt2), log m = 2 atanh((m-1)/(m+1)) by
 *        its series to z^7/15 (both ~40 bits: they only feed float results);
 *   sin/cos: k = rint(x/(pi/2)), r = x - k*pi/2 (two-part, fma), Taylor
 *        series to degree 17/18, quadrant by k mod 4.
 * Pinned in tests/test_oracle_math.py against glibc (at most one float ulp,
 * almost always exact) and by identities (exp(log x) = x, sin^2+cos^2 = 1).
 */
#ifndef GFO_MATH_H
#define GFO_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline double gfo_bits2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static inline uint64_t gfo_d2bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

#define GFO_LN2_HI 6.93147180369123816490e-01
#define GFO_LN2_LO 1.90821492927058770002e-10
#define GFO_PIO2_HI 1.57079632679489655800e+00
#define GFO_PIO2_LO 6.12323399573676603587e-17

static inline double gfo_pow2i(int k) { return gfo_bits2d((uint64_t)(k + 1023) << 52); }

static inline double gfo_exp(double x)
{
	if (x != x) return x;
	if (x > 709.0) return INFINITY;
	if (x < -745.0) return 0.0;
	double kd = rint(x * 1.44269504088896338700e+00);
	double r = fma(-kd, GFO_LN2_HI, x);
	r = fma(-kd, GFO_LN2_LO, r);
	static const double inv_fact[11] = {
		1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0,
		1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0};
	double p = inv_fact[10];
	for (int i = 9; i >= 0; i--) p = fma(p, r, inv_fact[i]);
	int k = (int)kd;
	if (k > 1023) { p *= gfo_pow2i(1023); k -= 1023; }
	if (k < -1022) { p *= gfo_pow2i(-1022); k += 1022; }
	return p * gfo_pow2i(k);
}

static inline double gfo_log(double x)
{
	if (x != x || x < 0.0) return NAN;
	if (x == 0.0) return -INFINITY;
	uint64_t u = gfo_d2bits(x);
	int e = (int)((u >> 52) & 0x7ff);
	if (e == 0x7ff) return x;
	if (e == 0) { x *= gfo_pow2i(54); u = gfo_d2bits(x); e = (int)((u >> 52) & 0x7ff) - 54; }
	e -= 1023;
	double m = gfo_bits2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
	if (m > 1.41421356237309504880) { m *= 0.5; e += 1; }
	double s = (m - 1.0) / (m + 1.0);
	double z = s * s;
	double p = 1.0 / 15.0;
	for (int n = 6; n >= 0; n--) p = fma(p, z, 1.0 / (double)(2 * n + 1));
	double lm = 2.0 * s * p;
	double ed = (double)e;
	return fma(ed, GFO_LN2_HI, fma(ed, GFO_LN2_LO, lm));
}

static inline void gfo_sincos(double x, double *sn, double *cs)
{
	double kd = rint(x * 6.36619772367581382433e-01);
	double r = fma(-kd, GFO_PIO2_HI, x);
	r = fma(-kd, GFO_PIO2_LO, r);
	double z = r * r;
	/* sin: r - r^3/3! + ... ; coefficients of z^n in (1/3! - z/5! + ...) */
	static const double sc[8] = {1.0 / 6.0, -1.0 / 120.0, 1.0 / 5040.0, -1.0 / 362880.0,
	                             1.0 / 39916800.0, -1.0 / 6227020800.0, 1.0 / 1307674368000.0,
	                             -1.0 / 355687428096000.0};
	double ps = sc[7];
	for (int i = 6; i >= 0; i--) ps = fma(ps, z, sc[i]);
	double s = fma(-r * z, ps, r);
	static const double cc[9] = {0.5, -1.0 / 24.0, 1.0 / 720.0, -1.0 / 40320.0, 1.0 / 3628800.0,
	                             -1.0 / 479001600.0, 1.0 / 87178291200.0, -1.0 / 20922789888000.0,
	                             1.0 / 6402373705728000.0};
	double pc = cc[8];
	for (int i = 7; i >= 0; i--) pc = fma(pc, z, cc[i]);
	double c = fma(-z, pc, 1.0);
	int q = ((int)(int64_t)kd) & 3;
	switch (q) {
	case 0: *sn = s; *cs = c; break;
	case 1: *sn = c; *cs = -s; break;
	case 2: *sn = -s; *cs = -c; break;
	default: *sn = -c; *cs = s; break;
	}
}

/* atan(k/8), k = 0..8, each rounded to the nearest double (computed with
 * 200-bit arithmetic) */
static const double GFO_ATAN_K8[9] = {
	0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2, 0x1.dac670561bb4fp-2,
	0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1, 0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1};

/* atan t for 0 <= t <= 1: c = k/8 nearest t, atan t = atan c + atan u with
 * u = (t - c) / (1 + t c) (t - c exact, |u| <= 1/16), atan u by its Taylor
 * series to u^13/13 (truncation < 6e-20) */
static inline double gfo_atan_unit(double t)
{
	double k = rint(t * 8.0);
	double c = k * 0.125;
	double u = (t - c) / fma(t, c, 1.0);
	double z = u * u;
	double p = 1.0 / 13.0;
	p = fma(p, z, -1.0 / 11.0);
	p = fma(p, z, 1.0 / 9.0);
	p = fma(p, z, -1.0 / 7.0);
	p = fma(p, z, 1.0 / 5.0);
	p = fma(p, z, -1.0 / 3.0);
	return GFO_ATAN_K8[(int)k] + fma(u * z, p, u);
}

static inline double gfo_atan(double x)
{
	if (x != x) return x;
	double ax = fabs(x);
	int inv = ax > 1.0;
	double a = gfo_atan_unit(inv ? 1.0 / ax : ax);
	if (inv) a = (GFO_PIO2_HI - a) + GFO_PIO2_LO;
	return x < 0 ? -a : a;
}

static inline double gfo_atan2(double y, double x)
{
	const double PI_HI = 3.14159265358979311600e+00, PI_LO = 1.22464679914735317720e-16;
	if (x != x || y != y) return x + y;
	if (x == 0.0 && y == 0.0) {
		double r = signbit(x) ? PI_HI : 0.0;
		return signbit(y) ? -r : r;
	}
	double ax = fabs(x), ay = fabs(y), a;
	if (ay <= ax) a = gfo_atan_unit(ay / ax);
	else a = (GFO_PIO2_HI - gfo_atan_unit(ax / ay)) + GFO_PIO2_LO;
	if (x < 0) a = (PI_HI - a) + PI_LO;
	return (y < 0 || (y == 0.0 && signbit(y))) ? -a : a;
}

static inline double gfo_asin(double x)
{
	if (!(x >= -1.0 && x <= 1.0)) return NAN;
	return gfo_atan2(x, sqrt((1.0 - x) * (1.0 + x)));
}

/* e^x in float arithmetic: k = rint(x/ln2), r = x - k*ln2 with ln2 in a
 * 16-bit head (k*head exact) and a tail, Taylor series to r^7/7!, times 2^k */
static inline float gfo_expf(float x)
{
	const float hi = 0.693145751953125f, lo = 1.428606765330187e-06f;
	if (x != x) return x;
	if (x > 88.72283935546875f) return INFINITY;
	if (x < -103.97208404541016f) return 0.f;
	float k = rintf(x * 1.44269502162933349609375f);
	float r = fmaf(-k, hi, x);
	r = fmaf(-k, lo, r);
	static const float c[8] = {1.f / 5040.f, 1.f / 720.f, 1.f / 120.f, 1.f / 24.f, 1.f / 6.f, 0.5f, 1.f, 1.f};
	float p = c[0];
	for (int i = 1; i < 8; i++) p = fmaf(p, r, c[i]);
	int ki = (int)k;
	float s;
	uint32_t u;
	if (ki > 127) { u = 0x7f000000u; memcpy(&s, &u, 4); p *= s; ki -= 127; }
	if (ki < -126) { u = 0x00800000u; memcpy(&s, &u, 4); p *= s; ki += 126; }
	u = (uint32_t)(ki + 127) << 23;
	memcpy(&s, &u, 4);
	return p * s;
}
/* log x in float arithmetic: x = 2^e m, m in [1/sqrt2, sqrt2), f = m - 1,
 * s = f/(2+f), log(1+f) = f - (f^2/2 - s (f^2/2 + R)),
 * R = z (2/3 + 2z/5 + 2z^2/7 + 2z^3/9 + 2z^4/11), z = s^2; plus e*ln2 */
static inline float gfo_logf(float x)
{
	const float hi = 0.693145751953125f, lo = 1.428606765330187e-06f;
	if (x != x || x < 0.f) return NAN;
	if (x == 0.f) return -INFINITY;
	if (isinf(x)) return x;
	int e;
	float m = frexpf(x, &e) * 2.f;   /* x = m 2^(e-1), m in [1, 2) */
	e -= 1;
	if (m > 1.41421353816986083984375f) { m *= 0.5f; e += 1; }
	float f = m - 1.f;
	float h = 0.5f * f * f;
	float s = f / (2.f + f);
	float z = s * s;
	float R = 2.f / 11.f;
	R = fmaf(R, z, 2.f / 9.f);
	R = fmaf(R, z, 2.f / 7.f);
	R = fmaf(R, z, 2.f / 5.f);
	R = fmaf(R, z, 2.f / 3.f);
	R = R * z;
	float ef = (float)e;
	return fmaf(ef, hi, -((h - fmaf(s, h + R, ef * lo)) - f));
}
/* 10^y in float arithmetic (Cody and Waite): y log2(10) = k + r, k =
 * rint(y*Lh), r = fma(y, Lh, -k) + y*Ll with log2(10) = Lh + Ll (float head
 * and tail); 2^r by the Taylor series of e^(r ln2) to degree 7; times 2^k */
static inline float gfo_pow10f(float y)
{
	const float Lh = 0x1.a934fp+1f, Ll = 0x1.2f346ep-24f;
	if (y != y) return y;
	if (y > 38.5318413f) return INFINITY;
	if (y < -45.3534698f) return 0.f;
	float k = rintf(y * Lh);
	float r = fmaf(y, Lh, -k);
	r = fmaf(y, Ll, r);
	static const float c[8] = {0x1.ffcbfcp-17f, 0x1.430912p-13f, 0x1.5d87fep-10f, 0x1.3b2ab6p-7f,
	                           0x1.c6b08ep-5f,  0x1.ebfbe0p-3f,  0x1.62e430p-1f,  1.f};
	float p = c[0];
	for (int i = 1; i < 8; i++) p = fmaf(p, r, c[i]);
	int ki = (int)k;
	float s;
	uint32_t u;
	if (ki > 127) { u = 0x7f000000u; memcpy(&s, &u, 4); p *= s; ki -= 127; }
	if (ki < -126) { u = 0x00800000u; memcpy(&s, &u, 4); p *= s; ki += 126; }
	u = (uint32_t)(ki + 127) << 23;
	memcpy(&s, &u, 4);
	return p * s;
}
static inline float gfo_exp10f(float x) { return gfo_pow10f(x); }
static inline float gfo_log10f(float x) { return (float)(gfo_log((double)x) * 4.34294481903251816668e-01); }
static inline float gfo_sinf(float x) { double s, c; gfo_sincos((double)x, &s, &c); return (float)s; }
static inline float gfo_cosf(float x) { double s, c; gfo_sincos((double)x, &s, &c); return (float)c; }
static inline float gfo_tanf(float x) { double s, c; gfo_sincos((double)x, &s, &c); return (float)(s / c); }

static inline float gfo_powf(float x, float y)
{
	if (y == 0.0f || x == 1.0f) return 1.0f;
	if (x != x || y != y) return x + y;
	if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
	if (x < 0.0f) {
		float yi = truncf(y);
		if (yi != y) return NAN;
		double r = gfo_exp((double)y * gfo_log(-(double)x));
		return (float)((((int64_t)yi) & 1) ? -r : r);
	}
	return (float)gfo_exp((double)y * gfo_log((double)x));
}

#endif
