// Multiply-with-carry random streams and the samplers built on them.
//
// Follows prngs::mwc_impl::uniform (common/cuda_rng.h:439-463) and the
// samplers of prngs::rng<> (common/cuda_rng.h:158-376): polar Box-Muller
// gaussian, GSL-derived gamma_large/gamma_int/gamma_frac/gamma/beta,
// binomial (Knuth) and Poisson (mu > 10 split through gamma_int, product of
// uniforms below).  Each lane keeps its stream (a, c, x) in registers; the
// reference kept it in shared memory, which has no effect on the sequence.
// All float arithmetic is written operation by operation (the library is
// built with -ffp-contract=off) and transcendentals come from gf_math.h, so
// a CPU evaluation of the same steps yields the same bits.
#pragma once

#include "gf_math.h"

struct gf_mwc
{
	uint32_t a, c, x;

	GF_HD float uniform()
	{
		uint64_t t = (uint64_t)a * (uint64_t)x + (uint64_t)c;
		c = (uint32_t)(t >> 32);
		x = (uint32_t)t;
		return 2.32830643708e-10f * (float)x;
	}

	GF_HD float uniform_pos()
	{
		float u;
		do { u = uniform(); } while (u == 0.f);
		return u;
	}

	// One coordinate of a polar try, -1 + 2*uniform(): uniform() is
	// 2^-32 * (float)x exactly (the constant is 2^-32), so the product by 2
	// is exact and one fma with 2^-31 rounds the same single sum.
	GF_HD float polar_coord()
	{
		uint64_t t = (uint64_t)a * (uint64_t)x + (uint64_t)c;
		c = (uint32_t)(t >> 32);
		x = (uint32_t)t;
		return __builtin_fmaf((float)x, 0x1p-31f, -1.f);
	}

	// The acceptance test of a try, !(r2 > 1 || r2 == 0 || r2 == 4) of the
	// reference's loop, for the r2 >= 0 a try produces.
	GF_HD static bool polar_accept(float r2) { return r2 <= 1.0f && r2 != 0.f; }

	// The draws of one polar Box-Muller deviate: uniform pairs until the
	// point falls inside the unit disc.  gaussian() finishes it; modules
	// that only need to move the stream past a deviate stop here.
	GF_HD void polar(float &yy, float &r2)
	{
		float xx;
		do {
			xx = polar_coord();
			yy = polar_coord();
			r2 = xx * xx + yy * yy;
		} while (!polar_accept(r2));
	}

	// One try of a polar deviate; true when it is accepted.
	GF_HD bool polar_try()
	{
		float xx = polar_coord();
		float yy = polar_coord();
		return polar_accept(xx * xx + yy * yy);
	}

	// Move the stream past n polar deviates: the same draws as n calls of
	// polar(), in one loop, so that lanes of a wave that need different
	// numbers of tries for each deviate wait once instead of n times.
	GF_HD void polar_skip(int n)
	{
		while (n > 0) n -= polar_try() ? 1 : 0;
	}

	GF_HD float gaussian(const float sigma)
	{
		float yy, r2;
		polar(yy, r2);
		float t = -2.0f * gf_logf(r2);
		t = t / r2;
		return sigma * yy * gf_sqrtf(t);
	}

	GF_HD float gamma_large(const float aa)
	{
		float sqa, xv, yv, v;
		sqa = gf_sqrtf(2 * aa - 1);
		for (;;) {
			do {
				// tanf(M_PI * u): the product is formed in double and narrowed
				yv = gf_tanf((float)(3.14159265358979323846 * (double)uniform()));
				xv = sqa * yv + aa - 1;
			} while (xv <= 0);
			v = uniform();
			float e = (aa - 1) * gf_logf(xv / (aa - 1)) - sqa * yv;
			if (!(v > (1 + yv * yv) * gf_expf(e))) break;
		}
		return xv;
	}

	GF_HD float gamma_int(const unsigned int aa)
	{
		if (aa < 3) {
			float prod = 1;
			for (unsigned int i = 0; i < aa; i++) prod *= uniform_pos();
			return -gf_logf(prod);
		}
		return gamma_large((float)aa);
	}

	GF_HD float gamma_frac(const float aa)
	{
		// p = M_E / (a + M_E), evaluated in double as written in the reference
		float p, q, xv, u, v;
		p = (float)(2.71828182845904523536 / ((double)aa + 2.71828182845904523536));
		do {
			u = uniform();
			v = uniform_pos();
			if (u < p) {
				xv = gf_expf((1 / aa) * gf_logf(v));
				q = gf_expf(-xv);
			} else {
				xv = 1 - gf_logf(v);
				q = gf_expf((aa - 1) * gf_logf(xv));
			}
		} while (uniform() >= q);
		return xv;
	}

	GF_HD float gamma(const float aa, const float b)
	{
		unsigned int na = (unsigned int)__builtin_floorf(aa);
		if (aa == (float)na) return b * gamma_int(na);
		if (na == 0) return b * gamma_frac(aa);
		// the reference leaves the order of these two draws unspecified;
		// here the integer part is drawn first
		float gi = gamma_int(na);
		float gf = gamma_frac(aa - (float)na);
		return b * (gi + gf);
	}

	GF_HD float beta(const float aa, const float b)
	{
		float x1 = gamma(aa, 1.0f);
		float x2 = gamma(b, 1.0f);
		return x1 / (x1 + x2);
	}

	GF_HD unsigned int binomial(float p, unsigned int n)
	{
		unsigned int i, aa, b, k = 0;
		while (n > 10) {
			aa = 1 + (n / 2);
			b = 1 + n - aa;
			float X = beta((float)aa, (float)b);
			if (X >= p) {
				n = aa - 1;
				p /= X;
			} else {
				k += aa;
				n = b - 1;
				p = (p - X) / (1 - X);
			}
		}
		for (i = 0; i < n; i++) {
			float u = uniform();
			if (u < p) k++;
		}
		return k;
	}

	GF_HD unsigned int poisson(float mu)
	{
		float emu;
		float prod = 1.0f;
		unsigned int k = 0;
		while (mu > 10) {
			// m = mu * (7.0/8.0): double product truncated to unsigned
			unsigned int m = (unsigned int)((double)mu * 0.875);
			float X = gamma_int(m);
			if (X >= mu) {
				return k + binomial(mu / X, m - 1);
			} else {
				k += m;
				mu -= X;
			}
		}
		emu = gf_expf(-mu);
		do {
			prod *= uniform();
			k++;
		} while (prod > emu);
		return k - 1;
	}
};

// Jump-ahead.  With base 2^32 a stream's state (c, x), read as
// z = c*2^32 + x, lives in Z/mZ for m = a*2^32 - 1: one step maps z to
// a*(z mod 2^32) + (z div 2^32) = a*z mod m (a*2^32 = 1 mod m), so n steps
// multiply z by a^n.  c < a holds for every seeded state and is kept by the
// recurrence, so z < m; the two fixed points z = 0 and z = m are never
// reached from a seeded state.
//
// Products mod m are Montgomery products with R = 2^64, made of two
// reductions by 2^32: t*2^-32 = (t div 2^32) + a*(t mod 2^32) (mod m),
// because 2^-32 = a (mod m).  A jump table holds a^(2^b) * R mod m; its
// first entry a*R mod m is 2^32 exactly (a*2^64 = 2^32 mod m), and each next
// entry is the Montgomery square of the previous one.  mul(z, table[b]) is
// z * a^(2^b) mod m for z in plain form.
constexpr int GF_JUMP_BITS = 64;   // jump tables cover n < 2^64

// One stream table to seed on the device: nstreams streams from `seed`, at
// word `offset` of the output ([a | c | x], each plane nstreams words).
struct gf_seed_seg {
	uint32_t seed;
	uint32_t nstreams;
	uint64_t offset;
};

struct gf_mwc_mod
{
	uint64_t m;
	uint32_t a;

	GF_HD explicit gf_mwc_mod(uint32_t a_) : m(((uint64_t)a_ << 32) - 1u), a(a_) {}

	// x * Y * 2^-64 mod m for x, Y < m
	GF_HD uint64_t mul(uint64_t x, uint64_t Y) const
	{
		const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
		const uint32_t y0 = (uint32_t)Y, y1 = (uint32_t)(Y >> 32);
		// t = x*y0 < m*2^32; r = t*2^-32 mod m
		const uint64_t p0 = (uint64_t)x0 * y0;
		const uint64_t p1 = (uint64_t)x1 * y0 + (p0 >> 32);             // t div 2^32 < m
		uint64_t r = p1 + (uint64_t)a * (uint32_t)p0;                    // < 2m - a
		if (r < p1 || r >= m) r -= m;
		// t = r + x*y1 < m*2^32; result = t*2^-32 mod m
		const uint64_t q0 = (uint64_t)x0 * y1 + (uint32_t)r;
		const uint64_t q1 = (uint64_t)x1 * y1 + (q0 >> 32) + (r >> 32);  // t div 2^32 < m
		uint64_t s = q1 + (uint64_t)a * (uint32_t)q0;                    // < 2m - a
		if (s < q1 || s >= m) s -= m;
		return s;
	}
};

GF_HD uint64_t gf_mwc_z(const gf_mwc &r) { return ((uint64_t)r.c << 32) | r.x; }
GF_HD void gf_mwc_set_z(gf_mwc &r, uint64_t z) { r.c = (uint32_t)(z >> 32); r.x = (uint32_t)z; }

// z * a^n mod m with the jump table of the stream (entry b at tab[b * stride]
// for every set bit b of n).
GF_HD uint64_t gf_mwc_jump_z(const gf_mwc_mod &md, uint64_t z, uint64_t n, const uint64_t *tab, size_t stride)
{
	for (int b = 0; n; b++, n >>= 1)
		if (n & 1u) z = md.mul(z, tab[(size_t)b * stride]);
	return z;
}

// Plane layout of a stream table: [a | c | x], each plane nstreams words
// (common/cuda_rng.h:45-69).
GF_HD gf_mwc gf_mwc_load(const uint32_t *st, uint32_t nstreams, uint32_t s)
{
	gf_mwc r;
	r.a = st[s];
	r.c = st[nstreams + s];
	r.x = st[2 * nstreams + s];
	return r;
}

GF_HD void gf_mwc_store(uint32_t *st, uint32_t nstreams, uint32_t s, const gf_mwc &r)
{
	st[nstreams + s] = r.c;
	st[2 * nstreams + s] = r.x;
}
