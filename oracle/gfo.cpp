/*
 * ORACLE -- CPU restatement of the galfast catalog hot path.
 *
 * This file is test infrastructure: only tests/, __graft_entry__.smoke() and
 * bench.py's CPU baseline load it, and only as the checker.  It restates the
 * reference's algorithms from the reference's source text (paths relative to
 * the galfast tree are cited on every function), single-threaded, in plain
 * C++ with FP contraction disabled.  It shares the parameter-block layouts of
 * include/galfast_hip.h (host pointers in place of device pointers) and no
 * code with the product.
 *
 * Pinning (see the tests named in each section):
 *   MT19937, drand48/lrand48 ........ library routines and the C++ standard's
 *                                      known answer (tests/test_oracle_rng.py)
 *   MWC uniform ...................... brute-force 64-bit recurrence
 *   samplers .......................... moments / KS vs scipy distributions
 *   float transcendentals ............ glibc within 1 ulp (test_oracle_math.py)
 *   LF texture, spline ............... closed forms (piecewise-linear inputs)
 *   count integral ................... crawl-order per-cell sum vs separable
 *                                      FP64 form vs closed-form 1-D integrals
 *   gal2other ........................ radec golden of tests/os_vel2pm/result
 *   vel2pm ........................... commented known answer in
 *                                      modules/vel2pm_gpu.cu.h:223-231 and the
 *                                      pmlb golden column
 *   FeH, GaussianFeH, Bond2010, ...... KS / chi-square / covariance against the
 *   unresolvedMultiples, photometric   distributions their formulas fix, each
 *   errors                             test also rejecting a planted mistake
 *                                      (tests/test_oracle_modules.py)
 *   photometry ....................... closed form on polynomial colour tables
 *   footprint regions, coverage ...... textbook Galactic pole / centre in
 *                                      equatorial coordinates, cap, ring and
 *                                      band areas in closed form, union
 *                                      additivity (tests/test_oracle_footprint.py)
 *   countsMap binning ................ known bins of bin centres and edges,
 *                                      closed-form polar Lambert radius and
 *                                      azimuth, coadd monotonicity, numpy
 *                                      histogram (tests/test_oracle_countsmap.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <vector>

#include "../include/galfast_hip.h"
#include "gfo_math.h"

#define GFO_API extern "C" __attribute__((visibility("default")))

/* ======================================================================== */
/* MT19937 (the GSL default generator behind rng_gsl_t, common/gpu.h:362-382) */
/* Published algorithm of Matsumoto & Nishimura (1998), init_genrand seeding. */
/* ======================================================================== */
struct gfo_mt {
	uint32_t mt[624];
	int mti;
	void seed(uint32_t s)
	{
		mt[0] = s;
		for (int i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
		mti = 624;
	}
	uint32_t next()
	{
		if (mti >= 624) {
			for (int k = 0; k < 624; k++) {
				uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % 624] & 0x7fffffffu);
				mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
			}
			mti = 0;
		}
		uint32_t y = mt[mti++];
		y ^= (y >> 11);
		y ^= (y << 7) & 0x9d2c5680u;
		y ^= (y << 15) & 0xefc60000u;
		y ^= (y >> 18);
		return y;
	}
};

GFO_API void gfo_mt19937(uint32_t seed, uint32_t n, uint32_t *out)
{
	gfo_mt m;
	m.seed(seed);
	for (uint32_t i = 0; i < n; i++) out[i] = m.next();
}

/* persistent_rng::get (common/cux.cpp:404-410): GSL seeds 0 as 4357 */
GFO_API uint32_t gfo_seed_from_pipeline_seed(uint32_t pipeline_seed)
{
	gfo_mt m;
	m.seed(pipeline_seed == 0 ? 4357u : pipeline_seed);
	float u = (float)((double)m.next() / 4294967296.0);
	return (uint32_t)(u * (float)(1 << 24));
}

/* ======================================================================== */
/* drand48 family (POSIX): X <- a X + c mod 2^48, a = 0x5DEECE66D, c = 0xB   */
/* ======================================================================== */
struct gfo_r48 {
	uint64_t X;
	void seed(uint32_t s) { X = ((uint64_t)s << 16) | 0x330Eu; }
	void step() { X = (X * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1); }
	double drand() { step(); return ldexp((double)X, -48); }
	uint32_t lrand() { step(); return (uint32_t)(X >> 17); }
};

GFO_API void gfo_rand48(uint32_t seed, uint32_t n, double *d_out, uint32_t *l_out)
{
	gfo_r48 r;
	r.seed(seed);
	for (uint32_t i = 0; i < n; i++) {
		if (d_out) d_out[i] = r.drand();
		if (l_out) l_out[i] = r.lrand();
	}
}

/* mwc_impl::srand (common/cuda_rng.h:466-497).  (int)(drand48()*prime)
 * overflows for prime > 2^31; x86-64 cvttsd2si yields INT_MIN then. */
GFO_API int gfo_mwc_seed(uint32_t seed, const uint32_t *primes, uint32_t nprimes, uint32_t nstreams, uint32_t *out)
{
	if (nstreams % 64) nstreams += 64 - nstreams % 64;
	if (nstreams > nprimes) return -1;
	gfo_r48 r;
	r.seed(seed);
	for (uint32_t i = 0; i < nstreams; i++) {
		uint32_t a = primes[i];
		double v = r.drand() * (double)a;
		int32_t c = (v >= 2147483648.0 || v <= -2147483649.0) ? INT32_MIN : (int32_t)v;
		out[i] = a;
		out[nstreams + i] = (uint32_t)c;
		out[2 * nstreams + i] = r.lrand();
	}
	return (int)nstreams;
}

/* ======================================================================== */
/* MWC uniform and samplers (common/cuda_rng.h:158-376, 439-463)             */
/* ======================================================================== */
struct gfo_rng {
	uint32_t a, c, x;
	float uniform()
	{
		uint64_t t = (uint64_t)a * x + c;
		c = (uint32_t)(t >> 32);
		x = (uint32_t)(t & 0xffffffffu);
		return 2.32830643708e-10f * (float)x;
	}
	float uniform_pos()
	{
		float v;
		do { v = uniform(); } while (v == 0.f);
		return v;
	}
	/* polar Box-Muller, rejecting r2 > 1, r2 == 0, r2 == 4 */
	float gaussian(float sigma)
	{
		float xx, yy, r2;
		do {
			xx = -1.f + 2.f * uniform();
			yy = -1.f + 2.f * uniform();
			r2 = xx * xx + yy * yy;
		} while (r2 > 1.0f || r2 == 0.f || r2 == 4.f);
		float q = -2.0f * gfo_logf(r2);
		q = q / r2;
		return sigma * yy * sqrtf(q);
	}
	float gamma_large(float aa)
	{
		float sqa = sqrtf(2 * aa - 1), xv, yv, v;
		for (;;) {
			do {
				float arg = (float)(3.14159265358979323846 * (double)uniform());
				yv = gfo_tanf(arg);
				xv = sqa * yv + aa - 1;
			} while (xv <= 0);
			v = uniform();
			float e = (aa - 1) * gfo_logf(xv / (aa - 1)) - sqa * yv;
			if (!(v > (1 + yv * yv) * gfo_expf(e))) return xv;
		}
	}
	float gamma_int(unsigned int aa)
	{
		if (aa < 3) {
			float prod = 1;
			for (unsigned int i = 0; i < aa; i++) prod *= uniform_pos();
			return -gfo_logf(prod);
		}
		return gamma_large((float)aa);
	}
	float gamma_frac(float aa)
	{
		float p = (float)(2.71828182845904523536 / ((double)aa + 2.71828182845904523536));
		float q, xv, u, v;
		do {
			u = uniform();
			v = uniform_pos();
			if (u < p) { xv = gfo_expf((1 / aa) * gfo_logf(v)); q = gfo_expf(-xv); }
			else { xv = 1 - gfo_logf(v); q = gfo_expf((aa - 1) * gfo_logf(xv)); }
		} while (uniform() >= q);
		return xv;
	}
	float gamma(float aa, float b)
	{
		unsigned int na = (unsigned int)floorf(aa);
		if (aa == (float)na) return b * gamma_int(na);
		if (na == 0) return b * gamma_frac(aa);
		float gi = gamma_int(na);
		float gf = gamma_frac(aa - (float)na);
		return b * (gi + gf);
	}
	float beta(float aa, float b)
	{
		float x1 = gamma(aa, 1.0f);
		float x2 = gamma(b, 1.0f);
		return x1 / (x1 + x2);
	}
	unsigned int binomial(float p, unsigned int n)
	{
		unsigned int k = 0;
		while (n > 10) {
			unsigned int aa = 1 + (n / 2);
			unsigned int b = 1 + n - aa;
			float X = beta((float)aa, (float)b);
			if (X >= p) { n = aa - 1; p /= X; }
			else { k += aa; n = b - 1; p = (p - X) / (1 - X); }
		}
		for (unsigned int i = 0; i < n; i++) {
			float u = uniform();
			if (u < p) k++;
		}
		return k;
	}
	unsigned int poisson(float mu)
	{
		float prod = 1.0f;
		unsigned int k = 0;
		while (mu > 10) {
			unsigned int m = (unsigned int)((double)mu * (7.0 / 8.0));
			float X = gamma_int(m);
			if (X >= mu) return k + binomial(mu / X, m - 1);
			k += m;
			mu -= X;
		}
		float emu = gfo_expf(-mu);
		do {
			prod *= uniform();
			k++;
		} while (prod > emu);
		return k - 1;
	}
};

static gfo_rng rng_load(const uint32_t *st, uint32_t ns, uint32_t s)
{
	gfo_rng r;
	r.a = st[s]; r.c = st[ns + s]; r.x = st[2 * ns + s];
	return r;
}
static void rng_store(uint32_t *st, uint32_t ns, uint32_t s, const gfo_rng &r)
{
	st[ns + s] = r.c; st[2 * ns + s] = r.x;
}

/* Jump-ahead of a multiply-with-carry stream (the recurrence of
 * common/cuda_rng.h:439-463).  With base b = 2^32, read the state as
 * z = c*2^32 + x.  One step maps z to a*(z mod 2^32) + (z div 2^32), which is
 * a*z mod m for m = a*2^32 - 1 (because a*2^32 = 1 mod m), so n steps
 * multiply z by a^n mod m.  Evaluated here with exact 128-bit products
 * (square-and-multiply); pinned against stepping the recurrence
 * (tests/test_oracle_rng.py). */
static uint64_t gfo_mulmod(uint64_t x, uint64_t y, uint64_t m)
{
	return (uint64_t)(((unsigned __int128)x * y) % m);
}
static uint64_t gfo_powmod(uint64_t a, uint64_t n, uint64_t m)
{
	uint64_t r = 1 % m, b = a % m;
	while (n) {
		if (n & 1) r = gfo_mulmod(r, b, m);
		b = gfo_mulmod(b, b, m);
		n >>= 1;
	}
	return r;
}
static void gfo_jump(gfo_rng &r, uint64_t n)
{
	const uint64_t m = ((uint64_t)r.a << 32) - 1;
	uint64_t z = ((uint64_t)r.c << 32) | r.x;
	z = gfo_mulmod(z, gfo_powmod(r.a, n, m), m);
	r.c = (uint32_t)(z >> 32);
	r.x = (uint32_t)z;
}
/* every stream of st moved n steps ahead */
GFO_API void gfo_mwc_jump(uint32_t *st, uint32_t ns, uint64_t n)
{
	for (uint32_t s = 0; s < ns; s++) {
		gfo_rng r = rng_load(st, ns, s);
		gfo_jump(r, n);
		rng_store(st, ns, s, r);
	}
}
/* every stream of st moved n steps ahead by stepping */
GFO_API void gfo_mwc_step(uint32_t *st, uint32_t ns, uint64_t n)
{
	for (uint32_t s = 0; s < ns; s++) {
		gfo_rng r = rng_load(st, ns, s);
		for (uint64_t i = 0; i < n; i++) (void)r.uniform();
		rng_store(st, ns, s, r);
	}
}

GFO_API void gfo_mwc_sample(uint32_t *st, uint32_t ns, int kind, float param, uint32_t n, void *out)
{
	for (uint32_t s = 0; s < ns; s++) {
		gfo_rng r = rng_load(st, ns, s);
		for (uint32_t i = 0; i < n; i++) {
			size_t o = (size_t)i * ns + s;
			if (kind == 0) ((float *)out)[o] = r.uniform();
			else if (kind == 1) ((float *)out)[o] = r.gaussian(param);
			else if (kind == 2) ((uint32_t *)out)[o] = r.poisson(param);
			else ((uint32_t *)out)[o] = r.poisson(param * (float)(1 + s));
		}
		rng_store(st, ns, s, r);
	}
}

GFO_API void gfo_math_f(int fn, const float *x, const float *y, float *out, size_t n)
{
	for (size_t i = 0; i < n; i++) {
		switch (fn) {
		case 0: out[i] = gfo_expf(x[i]); break;
		case 1: out[i] = gfo_logf(x[i]); break;
		case 2: out[i] = gfo_sinf(x[i]); break;
		case 3: out[i] = gfo_cosf(x[i]); break;
		case 4: out[i] = gfo_tanf(x[i]); break;
		case 5: out[i] = gfo_powf(x[i], y[i]); break;
		case 6: out[i] = gfo_exp10f(x[i]); break;
		case 7: out[i] = gfo_log10f(x[i]); break;
		}
	}
}

/* ======================================================================== */
/* Texture sampling (common/cux.h:849-874) with exact linear filtering      */
/* ======================================================================== */
static float coord(float x, float tc0, float tc1, int n, int *i)
{
	float xi = (x - tc0) * tc1 + 0.5f;
	float t = xi - 0.5f;
	if (!(t > 0.f)) { *i = 0; return 0.f; }
	if (t >= (float)(n - 1)) { *i = n > 1 ? n - 2 : 0; return n > 1 ? 1.f : 0.f; }
	float fl = floorf(t);
	*i = (int)fl;
	return t - fl;
}

static float tex1(const float *d, int n, const float tc[2], float x)
{
	int i;
	float a = coord(x, tc[0], tc[1], n, &i);
	if (n == 1) return d[0];
	return d[i] + a * (d[i + 1] - d[i]);
}

static float tex3(const gf_tex3d &t, float x, float y, float z)
{
	if (!t.data) return 0.f;
	int i, j, k;
	float a = coord(x, t.tc[0][0], t.tc[0][1], t.nx, &i);
	float b = coord(y, t.tc[1][0], t.tc[1][1], t.ny, &j);
	float c = coord(z, t.tc[2][0], t.tc[2][1], t.nz, &k);
	int i1 = t.nx > 1 ? i + 1 : i, j1 = t.ny > 1 ? j + 1 : j, k1 = t.nz > 1 ? k + 1 : k;
	auto at = [&](int ii, int jj, int kk) { return t.data[(size_t)ii + (size_t)jj * t.nx + (size_t)kk * t.nx * t.ny]; };
	float c00 = at(i, j, k) + a * (at(i1, j, k) - at(i, j, k));
	float c10 = at(i, j1, k) + a * (at(i1, j1, k) - at(i, j1, k));
	float c01 = at(i, j, k1) + a * (at(i1, j, k1) - at(i, j, k1));
	float c11 = at(i, j1, k1) + a * (at(i1, j1, k1) - at(i, j1, k1));
	float c0 = c00 + b * (c10 - c00);
	float c1 = c01 + b * (c11 - c01);
	return c0 + c * (c1 - c0);
}

GFO_API float gfo_tex3d(const gf_tex3d *t, float x, float y, float z) { return tex3(*t, x, y, z); }

/* ======================================================================== */
/* Models (skygen/model_expDisk.h:71-121, model_powerLawEllipsoid.h:60-105)  */
/* ======================================================================== */
static void xform(float &x, float &y, float &z, const float T[3], const float M[3][3])
{
	float vx = x + T[0], vy = y + T[1], vz = z + T[2];
	x = M[0][0] * vx + M[0][1] * vy + M[0][2] * vz;
	y = M[1][0] * vx + M[1][1] * vy + M[1][2] * vz;
	z = M[2][0] * vx + M[2][1] * vy + M[2][2] * vz;
}

static float rho_xyz(const gf_model &m, float x, float y, float z)
{
	xform(x, y, z, m.T, m.M);
	if (m.kind == GF_MODEL_EXPDISK) {
		float r2 = x * x + y * y;
		if (m.r_cut2 != 0.f && r2 + z * z > m.r_cut2) return 0.f;
		float r = sqrtf(r2);
		return m.f * gfo_expf(-r / m.l - fabsf(z) / m.h);
	}
	float DSq = x * x + y * y + z * z;
	if (m.rmaxSq != 0.f && (DSq < m.rminSq || DSq >= m.rmaxSq)) return 0.f;
	float yb = y / m.ba, zc = z / m.ca;
	float rSq = x * x + yb * yb + zc * zc;
	return m.f * gfo_powf(rSq, 0.5f * m.n);
}

GFO_API float gfo_model_rho(const gf_model *m, float x, float y, float z) { return rho_xyz(*m, x, y, z); }

/* LF normalisation of load_lf (skygen/model_lib.h:68-95): f = f0 / rho at
 * the LF site with f = 1 and no cuts. */
GFO_API float gfo_lf_norm_f(gf_model m, float f0, float lfx, float lfy, float lfz)
{
	m.f = 1.f;
	m.r_cut2 = 0.f;
	m.rminSq = m.rmaxSq = 0.f;
	float norm = rho_xyz(m, lfx, lfy, lfz);
	return f0 / norm;
}

/* ======================================================================== */
/* Linear spline (common/spline.cpp:59-65: gsl_interp_linear), extrapolating */
/* linearly past the end knots                                               */
/* ======================================================================== */
static double spline_at(const double *x, const double *y, int n, double v)
{
	if (n == 1) return y[0];
	int lo = 0, hi = n - 1;
	while (hi > lo + 1) {
		int mid = (lo + hi) / 2;
		if (x[mid] > v) hi = mid; else lo = mid;
	}
	double dx = x[lo + 1] - x[lo];
	double D = (y[lo + 1] - y[lo]) / dx;
	return y[lo] + D * (v - x[lo]);
}

GFO_API double gfo_spline(const double *x, const double *y, int n, double v) { return spline_at(x, y, n, v); }

/* load_resample_and_clip_texture_1D + construct_texture_by_resampling_1D
 * (common/cux.cpp:336-384): zero points are added 1e-5 of the range outside
 * both ends, then the linear spline is resampled to nsamp texels;
 * tc = (X0, 1/dX) in float. */
GFO_API void gfo_resample_clip(const double *X, const double *Y, int n, float clip, int nsamp, float *tex, float tc[2])
{
	std::vector<double> x(n + 2), y(n + 2);
	for (int i = 0; i < n; i++) { x[i + 1] = X[i]; y[i + 1] = Y[i]; }
	double xrange = x[n] - x[1];
	x[0] = x[1] - 1e-5 * xrange; y[0] = clip;
	x[n + 1] = x[n] + 1e-5 * xrange; y[n + 1] = clip;
	double X0 = x[0], X1 = x[n + 1], dX = (X1 - X0) / (nsamp - 1);
	for (int i = 0; i < nsamp; i++) tex[i] = (float)spline_at(x.data(), y.data(), n + 2, X0 + i * dX);
	tc[0] = (float)X0;
	tc[1] = (float)(1. / dX);
}

/* construct_texture_by_resampling_1D without clipping (common/cux.cpp:336-355) */
GFO_API void gfo_resample(const double *X, const double *Y, int n, int nsamp, float *tex, float tc[2])
{
	double X0 = X[0], X1 = X[n - 1], dX = (X1 - X0) / (nsamp - 1);
	for (int i = 0; i < nsamp; i++) tex[i] = (float)spline_at(X, Y, n, X0 + i * dX);
	tc[0] = (float)X0;
	tc[1] = (float)(1. / dX);
}

GFO_API float gfo_tex1d(const float *d, int n, const float tc[2], float x) { return tex1(d, n, tc, x); }

/* ======================================================================== */
/* Skygen geometry                                                          */
/* ======================================================================== */
static const float POGSON = 0.4605170185988091f;
static const double R2D = 57.2957795130823208767981548;

/* lambert::deproject (skygen/skygen.h:131-146) */
static void deproject(const gf_lambert &p, double &l, double &b, float x, float y)
{
	double r2 = (double)(x * x + y * y);
	if (r2 == 0.) { b = p.b0; l = 0.; return; }
	double ir = 1. / sqrt(r2);
	double sc = 1. / ir * sqrt(1. - 0.25 * r2);
	double cc = 1. - 0.5 * r2;
	b = gfo_asin(cc * (double)p.sb + (double)y * sc * (double)p.cb * ir);
	float ysb = y * p.sb;
	l = p.l0 + gfo_atan2((double)x * sc, (double)p.cb * cc / ir - (double)ysb * sc);
}

/* deproject with the direction cosines (cos b cos l, cos b sin l, sin b)
 * of the point from its own terms: cos b sin(l - l0) = x sc ir and
 * cos b cos(l - l0) = (cb cc / ir - y sb sc) ir (the atan2 arguments times
 * ir = 1/rho), sin b the asin argument, rotated by l0 (cl, sl of the
 * projection). */
static void deproject_dir(const gf_lambert &p, double &l, double &b, double &cbcl, double &cbsl, double &sbo,
                          float x, float y)
{
	double r2 = (double)(x * x + y * y);
	if (r2 == 0.) { b = p.b0; l = 0.; cbcl = (double)p.cb; cbsl = 0.; sbo = (double)p.sb; return; }
	double ir = 1. / sqrt(r2);
	double sc = 1. / ir * sqrt(1. - 0.25 * r2);
	double cc = 1. - 0.5 * r2;
	double S = cc * (double)p.sb + (double)y * sc * (double)p.cb * ir;
	b = gfo_asin(S);
	float ysb = y * p.sb;
	double Pn = (double)x * sc, Qd = (double)p.cb * cc / ir - (double)ysb * sc;
	l = p.l0 + gfo_atan2(Pn, Qd);
	double Pr = Pn * ir, Qr = Qd * ir;
	cbcl = (double)p.cl * Qr - (double)p.sl * Pr;
	cbsl = (double)p.sl * Qr + (double)p.cl * Pr;
	sbo = S;
}

GFO_API void gfo_deproject(const gf_lambert *p, float x, float y, double *l, double *b) { deproject(*p, *l, *b, x, y); }

GFO_API void gfo_galequ(double l, double b, double *ra, double *dec);

/* One footprint region as footprint.cpp builds it from one configuration
 * file: rect in Galactic or equatorial lon/lat (load_footprint_rect,
 * footprint.cpp:486-560), beam with optional hole (load_footprint_beam,
 * :675-700), minus the zone of avoidance |b| < bmin (:704-727).  The
 * reference clips gpc polygons sampled along these boundaries; the oracle
 * tests the boundaries themselves.  1 if (l, b) [radians] is outside. */
static int outside_region(const gf_foot_region &f, double l, double b)
{
	if (f.bmin > 0. && fabs(b * R2D) < f.bmin) return 1;
	if (f.kind == 1) {
		double lon = l, lat = b;
		if (f.equ) gfo_galequ(l, b, &lon, &lat);
		double ld = lon * R2D, bd = lat * R2D;
		while (ld < 0.) ld += 360.;
		while (ld >= 360.) ld -= 360.;
		if (bd < f.b0 || bd > f.b1) return 1;
		if (f.l1 - f.l0 >= 360.) return 0;
		double t = ld - f.l0;
		while (t < 0.) t += 360.;
		while (t >= 360.) t -= 360.;
		return t > (f.l1 - f.l0) ? 1 : 0;
	}
	if (f.kind == 2) {
		double cb, sb, cl, sl;
		gfo_sincos(b, &sb, &cb);
		gfo_sincos(l, &sl, &cl);
		double d = cb * cl * f.cb * f.cl + cb * sl * f.cb * f.sl + sb * f.sb;
		if (d < f.cosr) return 1;
		return (f.hole > 0. && 1. - d < f.hole) ? 1 : 0;
	}
	return 1;
}

/* The union of the footprint files (load_footprints, footprint.cpp:746-759). */
static int outside_foot(const gf_footprint &f, double l, double b)
{
	if (f.nregion <= 0) return 0;
	for (int k = 0; k < f.nregion && k < GF_FOOT_MAX; k++)
		if (!outside_region(f.r[k], l, b)) return 0;
	return 1;
}

GFO_API int gfo_outside_footprint(const gf_footprint *f, double l, double b) { return outside_foot(*f, l, b); }

/* Pixel-grid coverage (os_skygen.cpp:235-283, 940-995 pixelizes the clipped
 * polygons; here each cell of the pole-centred Lambert grid counts its 8 x 8
 * cell-centred sample points inside the footprint).  cover[(h n + i) n + j]. */
GFO_API void gfo_footprint_cover(const gf_footprint *foot, double dx, int n, uint8_t *cover)
{
	const double PI = 3.14159265358979323846;
	for (int h = 0; h < 2; h++)
		for (int i = 0; i < n; i++)
			for (int j = 0; j < n; j++) {
				const double x0 = -n * dx / 2, xa = x0 + i * dx, ya = x0 + j * dx;
				const double fx = (xa > 0 || xa + dx < 0) ? std::min(fabs(xa), fabs(xa + dx)) : 0.;
				const double fy = (ya > 0 || ya + dx < 0) ? std::min(fabs(ya), fabs(ya + dx)) : 0.;
				int inside = 0;
				if (fx * fx + fy * fy <= 2.0)
					for (int a = 0; a < 8; a++)
						for (int c = 0; c < 8; c++) {
							double x = xa + dx * (a + 0.5) / 8, y = ya + dx * (c + 0.5) / 8;
							double r2 = x * x + y * y;
							if (r2 > 2.0) continue;
							/* hemisphere of the sample: z = 1 - r^2/2 north, -z south */
							double z = 1.0 - 0.5 * r2;
							if ((h == 0 && z < 0) || (h == 1 && -z >= 0)) continue;
							double sb = std::min(1.0, std::max(-1.0, z)), l, b;
							if (h == 0) { b = gfo_asin(sb); l = 0.5 * PI + gfo_atan2(x, -y); }
							else { b = -gfo_asin(sb); l = 1.5 * PI + gfo_atan2(x, y); }
							while (l < 0.) l += 2 * PI;
							while (l >= 2 * PI) l -= 2 * PI;
							if (foot->nregion <= 0) { inside++; continue; }
							bool in = false;
							for (int k = 0; k < foot->nregion && k < GF_FOOT_MAX && !in; k++) {
								const gf_foot_region &f = foot->r[k];
								if (f.kind == 2) {
									/* unit vector of the sample point straight from the plane */
									double s = sqrt(std::max(0.0, 1.0 - 0.25 * r2));
									double v0 = h == 0 ? -x * s : x * s, v1 = -y * s, v2 = h == 0 ? z : -z;
									if (f.bmin > 0. && fabs(b * R2D) < f.bmin) continue;
									double d = v0 * (f.cb * f.cl) + v1 * (f.cb * f.sl) + v2 * f.sb;
									in = d >= f.cosr && !(f.hole > 0. && 1. - d < f.hole);
								} else {
									in = !outside_region(f, l, b);
								}
							}
							inside += in ? 1 : 0;
						}
				cover[((size_t)h * n + i) * n + j] = (uint8_t)inside;
			}
}

/* compute_pos (skygen/skygen.cu.h:216-229) */
static void cell_pos(const gf_skygen_params &sp, const gf_extinction &ext, const gf_pencil_beam &pix, float M, int im,
                     float &D, float &Am, float &x, float &y, float &z)
{
	float m = sp.m0 + im * sp.dm;
	float DM = m - M;
	D = gfo_pow10f(0.2f * DM + 1.f);
	int xidx = pix.extIdx / 2048, yidx = pix.extIdx % 2048;
	Am = ext.beam.data ? tex3(ext.beam, (float)xidx, (float)yidx, DM) : 0.f;
	x = D * pix.cl * pix.cb;
	y = D * pix.sl * pix.cb;
	z = D * pix.sb;
}

/* diagIndexToIJ and advance (skygen/skygen.cu.h:257-358) */
static void diag_ij(int &ilb, int &i, int &j, int &k, int x, int y)
{
	int kmax = x * y;
	if (k >= kmax) { ilb += k / kmax; k %= kmax; }
	int l = x < y ? x : y, m = x < y ? y : x, d;
	if (2 * k < l * (l - 1)) {
		d = (int)(0.5 * (double)(sqrtf((float)(1 + 8 * k)) - 1));
		i = k - d * (d + 1) / 2;
	} else if (2 * k < l * (2 * m + 1 - l)) {
		if (x >= y) { int ka = k - y * (y - 1) / 2; d = ka / y + (y - 1); i = (ka % y) + d - (y - 1); }
		else { int ka = k - x * (x - 1) / 2; d = ka / x + (x - 1); i = ka % x; }
	} else {
		int ka = x * y - k - 1;
		int dd = (int)(0.5 * (double)(sqrtf((float)(1 + 8 * ka)) - 1));
		d = (x + y - 2) - dd;
		i = (x - 1) - (ka - dd * (dd + 1) / 2);
	}
	j = d - i;
}

static bool advance(int &ilb, int &i, int &j, gf_pencil_beam &pix, const gf_pencil_beam *pixels, int npix, int x, int y)
{
	i++; j--;
	if (j < 0) {
		if (i < y) { j += i + 1; i = 0; }
		else { i -= (y - 1); j += y; }
	} else if (i >= x) {
		if (x > y) { i -= y - j - 2; j = y - 1; }
		else if (j + 1 < y - x) { i -= x; j += x + 1; }
		else { i = x - y + j + 2; j = y - 1; }
		if (i == x) { ilb++; if (ilb != npix) pix = pixels[ilb]; }
	} else return false;
	return true;
}

struct Sink {
	gf_skygen_cols c;
	bool write;
	uint64_t row;
};

template <typename T> static T &C(const gf_col &c, uint64_t row, int plane = 0)
{
	return *((T *)((char *)c.ptr + (size_t)plane * c.pitch) + row);
}

/* draw_stars body (skygen/skygen.cu.h:373-456) for one star + clipper */
static void one_star(const gf_skygen_params &sp, const gf_model &m, const gf_extinction &ext, const gf_footprint &foot,
                     const gf_pencil_beam &pix, float M, int im, float AmMin, gfo_rng &rng, Sink &s)
{
	float x = pix.X, y = pix.Y;
	x += pix.dx * (rng.uniform() - 0.5f);
	y += pix.dx * (rng.uniform() - 0.5f);
	uint64_t r = s.row;
	if (x * x + y * y > 2.f) {
		if (s.write) {
			C<int>(s.c.hidden, r) = 1; C<int>(s.c.projIdx, r) = pix.projIdx;
			C<float>(s.c.projXY, r, 0) = x; C<float>(s.c.projXY, r, 1) = y;
			C<double>(s.c.lb, r, 0) = 0.; C<double>(s.c.lb, r, 1) = 0.;
			C<float>(s.c.M, r) = 0.f; C<float>(s.c.DM, r) = 0.f;
			for (int q = 0; q < 3; q++) C<float>(s.c.XYZ, r, q) = 0.f;
			C<int>(s.c.comp, r) = m.comp; C<float>(s.c.Am, r) = 0.f; C<float>(s.c.AmInf, r) = 0.f;
		}
		return;
	}
	double l, b, cbcl, cbsl, sbd;
	deproject_dir(sp.proj[pix.projIdx], l, b, cbcl, cbsl, sbd, x, y);
	int outside = outside_foot(foot, l, b);
	l *= R2D;
	if (l < 0.) l += 360.;
	if (l > 360.) l -= 360.;
	b *= R2D;
	float Mt = M + sp.dM * (rng.uniform() - 0.5f);
	float mt = sp.m0 + sp.dm * ((float)im + rng.uniform() - 0.5f);
	float DM = mt - Mt;
	float D = gfo_pow10f(0.2f * DM + 1.f);
	const gf_tex3d &em = pix.projIdx == 0 ? ext.north : ext.south;
	float Am0 = em.data ? tex3(em, x, y, DM) : 0.f;
	float Am1 = em.data ? tex3(em, x, y, 100.f) : 0.f;
	float AmInf = Am1 > Am0 ? Am1 : Am0;
	if (AmMin > Am0) AmInf = -AmMin;
	int hidden = (mt + Am0 > sp.m1) ? 1 : 0;
	hidden |= outside;
	if (s.write) {
		C<int>(s.c.projIdx, r) = pix.projIdx;
		C<float>(s.c.projXY, r, 0) = x; C<float>(s.c.projXY, r, 1) = y;
		C<double>(s.c.lb, r, 0) = l; C<double>(s.c.lb, r, 1) = b;
		C<float>(s.c.M, r) = Mt; C<float>(s.c.DM, r) = DM;
		C<float>(s.c.XYZ, r, 0) = D * (float)cbcl; C<float>(s.c.XYZ, r, 1) = D * (float)cbsl;
		C<float>(s.c.XYZ, r, 2) = D * (float)sbd;
		C<int>(s.c.comp, r) = m.comp;
		C<float>(s.c.Am, r) = Am0; C<float>(s.c.AmInf, r) = AmInf;
		C<int>(s.c.hidden, r) = hidden;
	}
}

/* The reference's count pass (skygen/skygen.cu.h:469-618, kernel<0>) with
 * sp.nthreads threads crawling in blocks of 10: per-thread FP64 sums of the
 * float per-cell expectations, summed on the host in thread order
 * (skygen/skyconfig_impl.h:137-156).  One launch over all pixels (the
 * reference restarts the crawl every PIXBLOCK pixels; the sum is the same
 * set of cells). */
GFO_API void gfo_count_crawl(const gf_skygen_params *spp, const gf_model *mp, const gf_pencil_beam *pixels,
                             const gf_extinction *extp, gf_count_result *out, double *per_pixel)
{
	const gf_skygen_params &sp = *spp;
	const gf_model &m = *mp;
	gf_extinction ext{};
	if (extp) ext = *extp;
	const int block = 10;
	double total = 0, covered = 0, mx = 0, ncell = 0;
	if (per_pixel) for (int p = 0; p < sp.npixels; p++) per_pixel[p] = 0;
	for (int tid = 0; tid < sp.nthreads; tid++) {
		int ilb = 0, im = 0, iM = 0, k = block * (tid - sp.nthreads), bc = 0;
		float D = 0, Am = 0, R = 0;
		gf_pencil_beam pix{};
		double count = 0, countCovered = 0;
		for (;;) {
			bool moved;
			if (bc == 0) {
				bc = block;
				k += block * sp.nthreads;
				diag_ij(ilb, im, iM, k, sp.nm, sp.nM);
				if (ilb >= sp.npixels) break;
				pix = pixels[ilb];
				moved = true;
			} else {
				moved = advance(ilb, im, iM, pix, pixels, sp.npixels, sp.nm, sp.nM);
			}
			bc--;
			float M = sp.M1 - iM * sp.dM;
			if (moved) {
				if (ilb >= sp.npixels) break;
				float x, y, z;
				cell_pos(sp, ext, pix, M, im, D, Am, x, y, z);
				R = rho_xyz(m, x, y, z);
			}
			if ((sp.dmin != 0.f || sp.dmax != 0.f) && (sp.dmin > D || sp.dmax <= D)) continue;
			float mm = sp.m0 + sp.dm * im + Am;
			if (mm > sp.m1) continue;
			float rho = tex1(m.lf.data, m.lf.n, m.lf.tc, M) * R;
			rho *= m.norm;
			rho *= D * D * D;
			rho *= pix.dA * POGSON * sp.dm * sp.dM;
			count += rho;
			countCovered += rho * pix.coveredFraction;
			if (per_pixel) per_pixel[ilb] += (double)(rho * pix.coveredFraction);
			if (rho > mx) mx = rho;
			ncell += 1;
		}
		total += count;
		covered += countCovered;
	}
	out->nstarsExpectedToGenerate = total;
	out->nstarsExpected = covered;
	out->maxCount = mx;
	out->ncells = ncell;
}

/* Reference-order draw (kernel<1>, skygen/skygen.cu.h:469-618) with
 * unlimited capacity; stars in (thread, crawl) order. */
GFO_API int64_t gfo_draw_compat(const gf_skygen_params *spp, const gf_model *mp, const gf_pencil_beam *pixels,
                                const gf_extinction *extp, const gf_footprint *footp, uint32_t *st, uint32_t ns,
                                const gf_skygen_cols *out, uint64_t capacity)
{
	const gf_skygen_params &sp = *spp;
	const gf_model &m = *mp;
	gf_extinction ext{};
	if (extp) ext = *extp;
	gf_footprint foot{};
	if (footp) foot = *footp;
	if ((uint32_t)sp.nthreads > ns) return -1;
	const int block = 10;
	Sink s{out ? *out : gf_skygen_cols{}, out != nullptr, 0};
	for (int tid = 0; tid < sp.nthreads; tid++) {
		gfo_rng rng = rng_load(st, ns, tid);
		int ilb = 0, im = 0, iM = 0, k = block * (tid - sp.nthreads), bc = 0;
		float D = 0, Am = 0, R = 0;
		gf_pencil_beam pix{};
		for (;;) {
			bool moved;
			if (bc == 0) {
				bc = block;
				k += block * sp.nthreads;
				diag_ij(ilb, im, iM, k, sp.nm, sp.nM);
				if (ilb >= sp.npixels) break;
				pix = pixels[ilb];
				moved = true;
			} else {
				moved = advance(ilb, im, iM, pix, pixels, sp.npixels, sp.nm, sp.nM);
			}
			bc--;
			float M = sp.M1 - iM * sp.dM;
			if (moved) {
				if (ilb >= sp.npixels) break;
				float x, y, z;
				cell_pos(sp, ext, pix, M, im, D, Am, x, y, z);
				R = rho_xyz(m, x, y, z);
			}
			if ((sp.dmin != 0.f || sp.dmax != 0.f) && (sp.dmin > D || sp.dmax <= D)) continue;
			float mm = sp.m0 + sp.dm * im + Am;
			if (mm > sp.m1) continue;
			float rho = tex1(m.lf.data, m.lf.n, m.lf.tc, M) * R;
			rho *= m.norm;
			rho *= D * D * D;
			rho *= pix.dA * POGSON * sp.dm * sp.dM;
			unsigned int nd = rng.poisson(rho);
			for (unsigned int q = 0; q < nd; q++) {
				if (s.write && s.row >= capacity) s.write = false;
				one_star(sp, m, ext, foot, pix, M, im, Am, rng, s);
				s.row++;
			}
		}
		rng_store(st, ns, tid, rng);
	}
	return (int64_t)s.row;
}

/* Rows [r0, r1) of the reference-order draw from the streams' states at the
 * start of the draw (st0, only read): every thread's crawl replayed in
 * thread order, the stars before r0 drawn without being written (their
 * draws move the stream all the same).  The capacity-limited draw of
 * skygenHost::drawSources (skygen/skyconfig_impl.h:434-534) writes these
 * rows in its batch [r0, r1).  Returns the number of rows written. */
GFO_API uint64_t gfo_draw_window(const gf_skygen_params *spp, const gf_model *mp, const gf_pencil_beam *pixels,
                                 const gf_extinction *extp, const gf_footprint *footp, const uint32_t *st0,
                                 uint32_t ns, uint64_t r0, uint64_t r1, const gf_skygen_cols *out)
{
	const gf_skygen_params &sp = *spp;
	const gf_model &m = *mp;
	gf_extinction ext{};
	if (extp) ext = *extp;
	gf_footprint foot{};
	if (footp) foot = *footp;
	const int block = 10;
	uint64_t row = 0;
	for (int tid = 0; tid < sp.nthreads && row < r1; tid++) {
		gfo_rng rng = rng_load(st0, ns, tid);
		int ilb = 0, im = 0, iM = 0, k = block * (tid - sp.nthreads), bc = 0;
		float D = 0, Am = 0, R = 0;
		gf_pencil_beam pix{};
		for (;;) {
			bool moved;
			if (bc == 0) {
				bc = block;
				k += block * sp.nthreads;
				diag_ij(ilb, im, iM, k, sp.nm, sp.nM);
				if (ilb >= sp.npixels) break;
				pix = pixels[ilb];
				moved = true;
			} else {
				moved = advance(ilb, im, iM, pix, pixels, sp.npixels, sp.nm, sp.nM);
			}
			bc--;
			float M = sp.M1 - iM * sp.dM;
			if (moved) {
				if (ilb >= sp.npixels) break;
				float x, y, z;
				cell_pos(sp, ext, pix, M, im, D, Am, x, y, z);
				R = rho_xyz(m, x, y, z);
			}
			if ((sp.dmin != 0.f || sp.dmax != 0.f) && (sp.dmin > D || sp.dmax <= D)) continue;
			float mm = sp.m0 + sp.dm * im + Am;
			if (mm > sp.m1) continue;
			float rho = tex1(m.lf.data, m.lf.n, m.lf.tc, M) * R;
			rho *= m.norm;
			rho *= D * D * D;
			rho *= pix.dA * POGSON * sp.dm * sp.dM;
			unsigned int nd = rng.poisson(rho);
			for (unsigned int q = 0; q < nd && row < r1; q++, row++) {
				Sink s{*out, row >= r0, row >= r0 ? row - r0 : 0};
				one_star(sp, m, ext, foot, pix, M, im, Am, rng, s);
			}
			if (row >= r1) break;
		}
	}
	return row > r0 ? row - r0 : 0;
}

/* ---- cumulative-table draw: the build's (pixel, DM-bin) restatement ---- */

/* Luminosity function at the absolute-magnitude bin centres, and its FP64
 * exclusive prefix sums. */
static void lf_prefix(const gf_skygen_params &sp, const gf_model &m, std::vector<float> &lfc, std::vector<double> &P)
{
	lfc.resize(sp.nM);
	P.assign(sp.nM + 1, 0.0);
	for (int iM = 0; iM < sp.nM; iM++) {
		float M = sp.M1 - iM * sp.dM;
		lfc[iM] = tex1(m.lf.data, m.lf.n, m.lf.tc, M);
		P[iM + 1] = P[iM] + (double)lfc[iM];
	}
}

/* DM bin k: distance of the diagonal's first cell in crawl order, window of
 * valid absolute-magnitude bins under the flux and distance cuts. */
static bool dm_bin(const gf_skygen_params &sp, const gf_extinction &ext, const gf_pencil_beam &pix, int k,
                   float &D, float &Am, float &x, float &y, float &z, int &lo, int &hi)
{
	int iM0 = k < sp.nM - 1 ? k : sp.nM - 1;
	float M = sp.M1 - iM0 * sp.dM;
	cell_pos(sp, ext, pix, M, k - iM0, D, Am, x, y, z);
	if ((sp.dmin != 0.f || sp.dmax != 0.f) && (sp.dmin > D || sp.dmax <= D)) return false;
	lo = std::max(0, k - (sp.nm - 1));
	hi = std::min(sp.nM - 1, k);
	if (lo > hi) return false;
	/* the cells of the window that pass m0 + dm*im + Am <= m1, by direct test */
	int first = -1, last = -1;
	for (int iM = hi; iM >= lo; iM--) {
		int im = k - iM;
		if (sp.m0 + sp.dm * im + Am > sp.m1) continue;
		if (first < 0) first = iM;
		last = iM;
	}
	if (first < 0) return false;
	hi = first;
	lo = last;
	return true;
}

/* drawSources' batch order (skygen/skyconfig_impl.h:467-515): std::sort of
 * the rows with star_comp (:34-68) -- hidden, then l, b, DM, M, Am, each with
 * < -- and the table truncated at the first hidden row.  std::stable_sort
 * here, so rows equal on every key keep their input order (std::sort leaves
 * it unspecified).  Gathers the visible rows into out; returns their
 * number. */
GFO_API uint64_t gfo_order(const gf_skygen_cols *inp, uint64_t n, const gf_skygen_cols *outp)
{
	const gf_skygen_cols &in = *inp, &out = *outp;
	std::vector<uint64_t> s(n);
	for (uint64_t i = 0; i < n; i++) s[i] = i;
	auto less = [&](uint64_t a, uint64_t b) {
		int ha = C<int>(in.hidden, a), hb = C<int>(in.hidden, b);
		if (ha != hb) return ha < hb;
		double la = C<double>(in.lb, a, 0), lb = C<double>(in.lb, b, 0);
		if (la < lb) return true;
		if (!(la == lb)) return false;
		double ba = C<double>(in.lb, a, 1), bb = C<double>(in.lb, b, 1);
		if (ba < bb) return true;
		if (!(ba == bb)) return false;
		float da = C<float>(in.DM, a), db = C<float>(in.DM, b);
		if (da < db) return true;
		if (!(da == db)) return false;
		float ma = C<float>(in.M, a), mb = C<float>(in.M, b);
		if (ma < mb) return true;
		if (!(ma == mb)) return false;
		return C<float>(in.Am, a) < C<float>(in.Am, b);
	};
	std::stable_sort(s.begin(), s.end(), less);
	uint64_t nv = 0;
	while (nv < n && C<int>(in.hidden, s[nv]) == 0) nv++;
	const gf_col *src[] = {&in.lb, &in.projIdx, &in.projXY, &in.DM, &in.M, &in.XYZ, &in.Am, &in.AmInf, &in.comp,
	                       &in.hidden};
	const gf_col *dst[] = {&out.lb, &out.projIdx, &out.projXY, &out.DM, &out.M, &out.XYZ, &out.Am, &out.AmInf,
	                       &out.comp, &out.hidden};
	for (int k = 0; k < 10; k++)
		for (uint32_t p = 0; p < src[k]->width; p++)
			for (uint64_t j = 0; j < nv; j++) {
				if (src[k]->elem == 8) C<uint64_t>(*dst[k], j, p) = C<uint64_t>(*src[k], s[j], p);
				else C<uint32_t>(*dst[k], j, p) = C<uint32_t>(*src[k], s[j], p);
			}
	return nv;
}

/* The clipper stage for the analytic footprints (os_clipper::process,
 * skygen/os_skygen.cpp:997-1044): a visible row whose projection-plane
 * position deprojects outside the footprint is hidden. */
GFO_API void gfo_clip(const gf_skygen_params *sp, const gf_footprint *foot, const gf_skygen_cols *c, uint64_t n)
{
	if (foot->nregion <= 0) return;
	for (uint64_t r = 0; r < n; r++) {
		if (C<int>(c->hidden, r)) continue;
		double l, b;
		deproject(sp->proj[C<int>(c->projIdx, r)], l, b, C<float>(c->projXY, r, 0), C<float>(c->projXY, r, 1));
		if (outside_foot(*foot, l, b)) C<int>(c->hidden, r) = 1;
	}
}

/* The countsMap stage's binning (os_countsMap::threaded_process,
 * pipeline.cpp:418-494, find_bin :361-367, lambert::project
 * common/projections.h:46-55), the reference's way: a std::map of beams
 * (X, Y, map), each a vector of nz x zwidth int counts created on the beam's
 * first row, coadds added to every bin from the depth bin up.  The result is
 * added to counts / beams in the product's dense layout
 * ([beam][Z][zwidth], beam = (X ny + Y) nmaps + map) and the number of
 * z-element counts is returned.  coadd_offsets is a host pointer. */
static int gfo_find_bin(double x, double x0, double dx, int n)
{
	const double v = (x - x0) / dx + 0.5;
	/* the x86-64 conversion of an out-of-range or NaN value: INT_MIN */
	int b = (v > -2147483649.0 && v < 2147483648.0) ? (int)v : (int)0x80000000u;
	b = std::max(0, b);
	b = std::min(n - 1, b);
	return b;
}

GFO_API uint64_t gfo_countsmap_bin(const gf_countsmap_params *pp, gf_col xy, gf_col z, gf_col hidden, uint64_t n,
                                   uint32_t *counts, uint8_t *beams)
{
	const gf_countsmap_params &p = *pp;
	const int zm = p.zmag_width, zw = zm + p.ncoadd, nmaps = p.equalarea ? 2 : 1;
	const double d2r = 0.01745329251994329576923691;
	struct key {
		int X, Y, map;
		bool operator<(const key &a) const
		{
			return X < a.X || (X == a.X && Y < a.Y) || (X == a.X && Y == a.Y && map < a.map);
		}
	};
	std::map<key, std::vector<int>> m;
	uint64_t total = 0;
	for (uint64_t row = 0; row < n; row++) {
		if (C<int>(hidden, row)) continue;
		key b;
		double x = C<double>(xy, row, 0), y = C<double>(xy, row, 1);
		b.map = 0;
		if (p.equalarea) {
			b.map = y > 0 ? 0 : 1;
			const double *pr = p.proj[b.map];
			const double l = x * d2r, phi = y * d2r;
			double sp, cp, sdl, cdl;
			gfo_sincos(phi, &sp, &cp);
			gfo_sincos(l - pr[0], &sdl, &cdl);
			const double denom = 1 + pr[1] * sp + pr[2] * cp * cdl;
			if (!(denom < 1e-10)) {
				const double kp = sqrt(2. / denom);
				x = kp * cp * sdl;
				y = kp * (pr[2] * sp - pr[1] * cp * cdl);
			}
		}
		b.X = gfo_find_bin(x, p.x0, p.dx, p.nx);
		b.Y = gfo_find_bin(y, p.y0, p.dy, p.ny);
		std::vector<int> &c = m[b];
		if (c.empty()) c.assign((size_t)p.nz * zw, 0);
		for (int i = 0; i != zm; i++) {
			int Z = gfo_find_bin(C<float>(z, row, i), p.z0, p.dz, p.nz);
			c[(size_t)Z * zw + i]++;
			total++;
		}
		for (int k = 0; k < p.ncoadd; k++) {
			const float *coeffs = p.coadd_offsets + (size_t)k * zm;
			float magCut = INFINITY;
			for (int i = 0; i != zm; i++) magCut = std::min(magCut, C<float>(z, row, i) - coeffs[i]);
			for (int Z = gfo_find_bin(magCut, p.z0, p.dz, p.nz); Z < p.nz; Z++) c[(size_t)Z * zw + zm + k]++;
		}
	}
	for (auto &kv : m) {
		const size_t beam = ((size_t)kv.first.X * p.ny + kv.first.Y) * nmaps + kv.first.map;
		beams[beam] = 1;
		for (size_t j = 0; j < kv.second.size(); j++) counts[beam * kv.second.size() + j] += (uint32_t)kv.second[j];
	}
	return total;
}

GFO_API void gfo_count_bins(const gf_skygen_params *spp, const gf_model *mp, const gf_pencil_beam *pixels,
                            const gf_extinction *extp, gf_count_result *out, double *per_pixel, double *wpk)
{
	const gf_skygen_params &sp = *spp;
	const gf_model &m = *mp;
	gf_extinction ext{};
	if (extp) ext = *extp;
	std::vector<float> lfc;
	std::vector<double> P;
	lf_prefix(sp, m, lfc, P);
	const int K = sp.nm + sp.nM - 1;
	double total = 0, cov = 0, mx = 0, nc = 0;
	for (int p = 0; p < sp.npixels; p++) {
		const gf_pencil_beam &pix = pixels[p];
		double pc = 0;
		for (int k = 0; k < K; k++) {
			float D, Am, x, y, z;
			int lo, hi;
			double w = -1.0;
			if (dm_bin(sp, ext, pix, k, D, Am, x, y, z, lo, hi)) {
				float R = rho_xyz(m, x, y, z);
				double vol = (double)pix.dA * (double)POGSON * (double)sp.dm * (double)sp.dM;
				w = (double)R * (double)m.norm * ((double)D * D * D) * vol * (P[hi + 1] - P[lo]);
				total += w;
				cov += w * (double)pix.coveredFraction;
				pc += w * (double)pix.coveredFraction;
				if (w > mx) mx = w;
				nc += hi - lo + 1;
			}
			if (wpk) wpk[(size_t)p * K + k] = w;
		}
		if (per_pixel) per_pixel[p] = pc;
	}
	out->nstarsExpectedToGenerate = total;
	out->nstarsExpected = cov;
	out->maxCount = mx;
	out->ncells = nc;
}

/* Plan: unit u = p*K + k Poisson-drawn from stream u mod ns. */
/* The cumulative-table draw's use of the streams (this build's own draw,
 * statistically equivalent to the reference's per-cell draw of
 * skygen/skygen.cu.h:366-457 by Poisson superposition).  Unit u = p*K + k
 * of a model belongs to stream s = u mod ns with index i = u div ns; with
 * the streams' states Z at the plan:
 *   the plan's Poisson draw of u starts at Z_s advanced i*2^32 draws;
 *   star j of u (0-based within the unit) starts at Z_s advanced
 *   i*2^32 + 2^31 + 8 j draws (a star takes 3 or 5);
 *   once the model is planned, the streams move on past every unit's draws,
 *   ceil(U / ns) * 2^32 draws (gfo_plan_advance). */
static const uint64_t GFO_UNIT_SPAN = 1ull << 32;
static const uint64_t GFO_EMIT_AT = 1ull << 31;
static const uint64_t GFO_STAR_SPAN = 8;

/* Poisson counts of the U units; st (the streams at the plan) is only read */
GFO_API void gfo_plan(const double *wpk, uint64_t U, const uint32_t *st, uint32_t ns, uint32_t *counts)
{
	for (uint32_t s = 0; s < ns && s < U; s++) {
		gfo_rng start = rng_load(st, ns, s);
		const uint64_t m = ((uint64_t)start.a << 32) - 1;
		const uint64_t A = gfo_powmod(start.a, GFO_UNIT_SPAN, m);   /* one unit's draws */
		for (uint64_t u = s; u < U; u += ns) {
			gfo_rng r = start;
			counts[u] = wpk[u] > 0.0 ? r.poisson((float)wpk[u]) : 0;
			/* the next unit of the stream */
			const uint64_t z = gfo_mulmod(((uint64_t)start.c << 32) | start.x, A, m);
			start.c = (uint32_t)(z >> 32);
			start.x = (uint32_t)z;
		}
	}
}

/* the streams once a model of U units is planned */
GFO_API void gfo_plan_advance(uint32_t *st, uint32_t ns, uint64_t U)
{
	const uint64_t nI = (U + ns - 1) / ns;
	for (uint32_t s = 0; s < ns; s++) {
		gfo_rng r = rng_load(st, ns, s);
		gfo_jump(r, nI * GFO_UNIT_SPAN);
		rng_store(st, ns, s, r);
	}
}

/* Emit rows [row_begin, row_end) of a plan with offsets offs (U+1); st
 * holds the streams as they were at the plan (only read). */
GFO_API void gfo_emit(const gf_skygen_params *spp, const gf_model *mp, const gf_pencil_beam *pixels,
                      const gf_extinction *extp, const gf_footprint *footp, const uint64_t *offs, uint64_t U,
                      uint64_t row_begin, uint64_t row_end, const uint32_t *st, uint32_t ns,
                      const gf_skygen_cols *out)
{
	const gf_skygen_params &sp = *spp;
	const gf_model &m = *mp;
	gf_extinction ext{};
	if (extp) ext = *extp;
	gf_footprint foot{};
	if (footp) foot = *footp;
	std::vector<float> lfc;
	std::vector<double> P;
	lf_prefix(sp, m, lfc, P);
	const int K = sp.nm + sp.nM - 1;
	if (row_begin >= row_end) return;
	/* the unit of row_begin: the largest u with offs[u] <= row_begin */
	uint64_t u = (uint64_t)(std::upper_bound(offs, offs + U + 1, row_begin) - offs) - 1;
	for (; u < U && offs[u] < row_end; u++) {
		const uint64_t o0 = offs[u], o1 = offs[u + 1];
		if (o1 == o0) continue;
		int p = (int)(u / K), k = (int)(u % K);
		const gf_pencil_beam &pix = pixels[p];
		float D, Am, x, y, z;
		int lo, hi;
		dm_bin(sp, ext, pix, k, D, Am, x, y, z, lo, hi);
		double base = P[lo], span = P[hi + 1] - P[lo];
		const uint64_t j0 = std::max(row_begin, o0), j1 = std::min(row_end, o1);
		gfo_rng star = rng_load(st, ns, (uint32_t)(u % ns));
		gfo_jump(star, (u / ns) * GFO_UNIT_SPAN + GFO_EMIT_AT + GFO_STAR_SPAN * (j0 - o0));
		Sink snk{*out, true, j0 - row_begin};
		for (uint64_t j = j0; j < j1; j++) {
			gfo_rng r = star;
			double tgt = base + (double)r.uniform() * span;
			int iM = lo;
			while (iM < hi && !(P[iM + 1] > tgt)) iM++;
			float M = sp.M1 - iM * sp.dM;
			one_star(sp, m, ext, foot, pix, M, k - iM, Am, r, snk);
			snk.row++;
			for (uint64_t q = 0; q < GFO_STAR_SPAN; q++) (void)star.uniform();   /* the next star */
		}
	}
}

/* ======================================================================== */
/* Per-star modules                                                          */
/* ======================================================================== */
static bool isset(const gf_bitmap &m, int bit)
{
	if (bit < 0 || bit >= 64) return false;   /* no component outside [0, 64) */
	return bit < 32 ? ((1u << bit) & m.bits[0]) != 0 : ((1u << (bit - 32)) & m.bits[1]) != 0;
}

struct RowIt {
	uint32_t begin, step, end;
	uint32_t nthreads() const { uint32_t n = end - begin; return n / step + (n % step ? 1 : 0); }
	void range(uint32_t t, uint32_t &r0, uint32_t &r1) const
	{
		uint64_t a = (uint64_t)begin + (uint64_t)step * t, b = a + step;
		r0 = (uint32_t)a; r1 = (uint32_t)std::min<uint64_t>(b, end);
	}
};

/* os_FeH_kernel (modules/FeH_gpu.cu.h:43-87) */
static void feh_rows(const gf_FeH_data &par, gfo_rng &rng, uint32_t r0, uint32_t r1, const gf_table &t)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		int cmp = C<int>(t.comp, row);
		if (isset(par.comp_thin, cmp) || isset(par.comp_thick, cmp)) {
			float p = rng.uniform() * (par.A[0] + par.A[1]);
			int i = p < par.A[0] ? 0 : 1;
			float x = C<float>(t.XYZ, row, 0), y = C<float>(t.XYZ, row, 1), z = C<float>(t.XYZ, row, 2);
			xform(x, y, z, par.T, par.M);
			float muD = par.muInf + par.DeltaMu * gfo_expf(-fabsf(z) / par.Hmu);
			float aZ = muD - 0.067f;
			C<float>(t.FeH, row) = rng.gaussian(par.sigma[i]) + aZ + par.offs[i];
		} else if (isset(par.comp_halo, cmp)) {
			C<float>(t.FeH, row) = par.offs[2] + rng.gaussian(par.sigma[2]);
		}
	}
}

/* os_GaussianFeH_kernel (modules/GaussianFeH_gpu.cu.h:42-73) */
static void gfeh_rows(const gf_GaussianFeH_data &par, gf_bitmap apply, gfo_rng &rng, uint32_t r0, uint32_t r1,
                      const gf_table &t)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		int cmp = C<int>(t.comp, row);
		if (!isset(apply, cmp)) continue;
		int at = 0;
		if (par.ncomp != 1) {
			float u = rng.uniform();
			while (at < par.ncomp - 1 && par.pcum[at] < u) ++at;
		}
		C<float>(t.FeH, row) = par.mean[at] + rng.gaussian(par.sigma[at]);
		if (t.FeH_comp.ptr) C<int>(t.FeH_comp, row) = at;
	}
}

/* modfun, trivar_gaussK_draw, disk/halo kinematics
 * (modules/Bond2010_gpu.cu.h:81-218) */
static float modfun(float R2, float Z, const float v[5])
{
	float p1 = gfo_powf(fabsf(Z), v[2]);
	float p2 = gfo_powf(R2, (float)(0.5 * (double)v[4]));
	return v[0] + v[1] * p1 + v[3] * p2;
}

static void trivar(float y[3], const float s[6], gfo_rng &rng)
{
	float b10 = s[1], b11 = s[3] * s[3], b20 = s[2], b21 = s[4], b22 = s[5] * s[5];
	float a00 = s[0], inv = 1.0f / a00;
	float a10 = b10 * inv;
	float a11 = sqrtf(b11 - a10 * a10);
	float a20 = b20 * inv;
	float a21 = (b21 - a20 * a10) / a11;
	float a22 = sqrtf(b22 - a20 * a20 - a21 * a21);
	float z0 = rng.gaussian(1.0f), z1 = rng.gaussian(1.0f), z2 = rng.gaussian(1.0f);
	y[0] += a00 * z0;
	y[1] += a10 * z0 + a11 * z1;
	y[2] += a20 * z0 + a21 * z1 + a22 * z2;
}

static void bond_rows(const gf_Bond2010_data &par, gfo_rng &rng, uint32_t r0, uint32_t r1, const gf_table &t)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		int cmp = C<int>(t.comp, row);
		float x = C<float>(t.XYZ, row, 0), y = C<float>(t.XYZ, row, 1), z = C<float>(t.XYZ, row, 2);
		xform(x, y, z, par.T, par.M);
		float R2 = (float)(1e-6 * (double)(x * x + y * y));
		float Z = (float)(1e-3 * (double)z);
		float v[3], s[6];
		if (isset(par.comp_thin, cmp) || isset(par.comp_thick, cmp)) {
			float p = rng.uniform();
			const gf_vmoments &d = p < par.fk ? par.disk_1 : par.disk_2;
			for (int i = 0; i < 3; i++) v[i] = modfun(R2, Z, d.m[i]);
			if (v[1] > 0.f) v[1] = 0.f;
			for (int i = 0; i < 6; i++) s[i] = modfun(R2, Z, d.ss[i]);
			trivar(v, s, rng);
		} else if (isset(par.comp_halo, cmp)) {
			for (int i = 0; i < 3; i++) v[i] = modfun(R2, Z, par.halo.m[i]);
			for (int i = 0; i < 6; i++) s[i] = modfun(R2, Z, par.halo.ss[i]);
			trivar(v, s, rng);
			float invR = 1.f / sqrtf(R2 + Z * Z);
			float ca = sqrtf(R2) * invR, sa = Z * invR;
			float vR = ca * v[0] - sa * v[1], vZ = sa * v[0] + ca * v[1];
			v[0] = vR; v[1] = v[2]; v[2] = vZ;
		} else continue;
		for (int i = 0; i < 3; i++) C<float>(t.vcyl, row, i) = v[i];
	}
}

static const double GE_ANGP = 0.01745329251994329576923691 * 192.859508333;
static const double GE_L0 = 0.01745329251994329576923691 * 32.932;
static const double GE_CE = 0.88998740217659689;
static const double GE_SE = 0.45598511375586859;
static const double D2R = 0.01745329251994329576923691;

/* transform_pm (modules/vel2pm_gpu.cu.h:134-164) */
static void transform_pm(float &vlout, float &vbout, float l, float b, float vl, float vb, double ce, double se, double l0)
{
	double cb, sb, cl, sl;
	gfo_sincos((double)b, &sb, &cb);
	gfo_sincos((double)l - l0, &sl, &cl);
	double tmp1 = cb * se - ce * sb * sl;
	double tmp2 = ce * sb - cb * se * sl;
	double tmp4 = cb * cl;
	double tmp3 = sb * se + cb * ce * sl;
	if (fabs(tmp3) > 0.9999999) { vlout = vbout = 999.99f; return; }
	double denom1 = 1. / (tmp4 * tmp4 + tmp2 * tmp2);
	double cd = sqrt(1. - tmp3 * tmp3);
	vl = (float)((double)vl / cb);
	vlout = (float)((-ce * cl * (double)vb + cb * tmp1 * (double)vl) * denom1);
	vbout = (float)((tmp1 * (double)vb + cb * ce * cl * (double)vl) / cd);
	vlout = (float)((double)vlout * cd);
}

/* vcyl2pm (modules/vel2pm_gpu.cu.h:221-251) */
static void vcyl2pm(const gf_vel2pm_data &par, float l, float b, float vr, float vphi, float vz0, float X, float Y, float Z,
                    float pm[3])
{
	float rho = sqrtf(X * X + Y * Y);
	float cphi = X / rho, sphi = Y / rho;
	float px = -sphi * vphi + cphi * vr, py = cphi * vphi + sphi * vr, pz = vz0;
	px -= par.u0;
	py -= par.v0 + par.vLSR;
	pz -= par.w0;
	float vx = -px, vy = -py, vz = pz;
	float cl = gfo_cosf(l), sl = gfo_sinf(l), cb = gfo_cosf(b), sb = gfo_sinf(b);
	float vl = -vx * sl + vy * cl;
	float tmp = vx * cl + vy * sl;
	float vrad = cb * tmp + vz * sb;
	float vb = -sb * tmp + vz * cb;
	float dx = 8000.f - X;
	float D = sqrtf(dx * dx + Y * Y + Z * Z);
	double den = 4.74 * (double)D * 1e-3;
	pm[0] = (float)((double)vl / den);
	pm[1] = (float)((double)vb / den);
	pm[2] = vrad;
}

GFO_API void gfo_vcyl2pm(const gf_vel2pm_data *par, float l, float b, float vr, float vphi, float vz, float X, float Y,
                         float Z, float *pm)
{
	vcyl2pm(*par, l, b, vr, vphi, vz, X, Y, Z, pm);
}

static void vel2pm_rows(const gf_vel2pm_data &par, uint32_t r0, uint32_t r1, const gf_table &t)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		float l = 0.017453292519943295769f * (float)C<double>(t.lb, row, 0);
		float b = 0.017453292519943295769f * (float)C<double>(t.lb, row, 1);
		float X = par.Rg - C<float>(t.XYZ, row, 0), Y = -C<float>(t.XYZ, row, 1), Z = C<float>(t.XYZ, row, 2);
		float pm[3];
		vcyl2pm(par, l, b, C<float>(t.vcyl, row, 0), C<float>(t.vcyl, row, 1), C<float>(t.vcyl, row, 2), X, Y, Z, pm);
		if (par.coordsys == 1) transform_pm(pm[0], pm[1], l, b, pm[0], pm[1], GE_CE, GE_SE, GE_L0);
		else if (par.coordsys != 0) pm[0] = pm[1] = pm[2] = -9.99f;
		for (int i = 0; i < 3; i++) C<float>(t.pm, row, i) = pm[i];
	}
}

/* galequ (modules/gal2other_gpu.cu.h:39-60), radians in and out */
GFO_API void gfo_galequ(double l, double b, double *ra, double *dec)
{
	double cb, sb, cl, sl;
	gfo_sincos(b, &sb, &cb);
	gfo_sincos(l - GE_L0, &sl, &cl);
	double rx = gfo_atan2(cb * cl, sb * GE_CE - cb * GE_SE * sl) + GE_ANGP;
	double ry = gfo_asin(cb * GE_CE * sl + sb * GE_SE);
	while (rx < 0.) rx += 6.28318530717958647692528677;
	while (rx >= 6.28318530717958647692528677) rx -= 6.28318530717958647692528677;
	*ra = rx;
	*dec = ry;
}

static void gal2other_rows(int cs, uint32_t r0, uint32_t r1, const gf_table &t)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		double ra = -9999., dec = -9999.;
		if (cs == 1) {
			gfo_galequ(C<double>(t.lb, row, 0) * D2R, C<double>(t.lb, row, 1) * D2R, &ra, &dec);
		}
		C<double>(t.radec, row, 0) = ra / D2R;
		C<double>(t.radec, row, 1) = dec / D2R;
	}
}

/* bilinear sample of float4 texels [j*nx + i] */
static void tex2_4(const float *tab, int nx, int ny, const float tcx[2], const float tcy[2], float x, float y, float out[4])
{
	int i, j;
	float a = coord(x, tcx[0], tcx[1], nx, &i);
	float b = coord(y, tcy[0], tcy[1], ny, &j);
	int i1 = nx > 1 ? i + 1 : i, j1 = ny > 1 ? j + 1 : j;
	for (int c = 0; c < 4; c++) {
		float t00 = tab[((size_t)j * nx + i) * 4 + c], t10 = tab[((size_t)j * nx + i1) * 4 + c];
		float t01 = tab[((size_t)j1 * nx + i) * 4 + c], t11 = tab[((size_t)j1 * nx + i1) * 4 + c];
		float q0 = t00 + a * (t10 - t00);
		float q1 = t01 + a * (t11 - t01);
		out[c] = q0 + b * (q1 - q0);
	}
}

static uint32_t brev32(uint32_t a)
{
	uint32_t r = 0;
	for (int i = 0; i < 32; i++) if (a & (1u << i)) r |= 1u << (31 - i);
	return r;
}

/* os_photometry_kernel (modules/photometry_gpu.cu.h:87-174) */
static void phot_rows(const gf_photometry_data &lt, gf_bitmap apply, uint32_t r0, uint32_t r1, const gf_table &t,
                      const gf_col &Mr)
{
	int nb = (int)lt.ncolors, bidx = (int)lt.bidx;
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		int cmp = C<int>(t.comp, row);
		if (!isset(apply, cmp)) continue;
		float FeH = C<float>(t.FeH, row);
		uint32_t flag = 0;
		float L[17] = {0}, Mfirst[17] = {0};
		float c[16];
		int ncomp = 0;
		for (int sc = 0; sc < (int)Mr.width; sc++) {
			float fMr = C<float>(Mr, row, sc);
			if (fMr >= 99.999f) break;
			uint32_t cidx = 0;
			for (int tt = 0; tt < 4 && (int)cidx < nb; tt++) {
				float cv[4], fv[4];
				tex2_4(lt.color[tt], lt.nFeH, lt.nMr, lt.tcFeH, lt.tcMr, FeH, fMr, cv);
				tex2_4(lt.cflags[tt], lt.nFeH, lt.nMr, lt.tcFeH, lt.tcMr, FeH, fMr, fv);
				for (int q = 0; q < 4 && (int)cidx < nb; q++) {
					c[cidx] = cv[q];
					flag |= (uint32_t)(fv[q] != 0.f) << cidx;
					cidx++;
				}
			}
			for (int b = 0; b <= nb; b++) {
				float M = fMr;
				if (b < bidx) for (int i = b; i != bidx; i++) M += c[i];
				if (b > bidx) for (int i = bidx; i != b; i++) M -= c[i];
				/* a lone star keeps its magnitude; a system sums luminosities,
				 * newest component first as in the reference's running sum */
				if (sc == 0) Mfirst[b] = M;
				else L[b] = gfo_exp10f(-0.4f * M) + (sc == 1 ? gfo_exp10f(-0.4f * Mfirst[b]) : L[b]);
			}
			ncomp++;
		}
		bool fl = false;
		for (int i = bidx - 1; i >= 0; i--) { fl = fl || (flag & (1u << i)); flag |= (uint32_t)fl << i; }
		fl = false;
		for (int i = bidx; i < nb; i++) { fl = fl || (flag & (1u << i)); flag |= (uint32_t)fl << i << 1; }
		flag &= ~(1u << bidx);
		flag = brev32(flag);
		flag >>= (32 - nb - 1);
		C<int>(t.photoFlags, row) = (int)flag;
		float dm = C<float>(t.DM, row), Am0 = C<float>(t.Am, row);
		for (int b = 0; b <= nb; b++) {
			float Msys = ncomp == 1 ? Mfirst[b] : -2.5f * gfo_log10f(L[b]);
			float msys = dm + Msys;
			C<float>(t.mags, row, b) = msys + Am0 * lt.reddening[b];
		}
	}
}

/* os_unresolvedMultiples_kernel (modules/unresolvedMultiples_gpu.cu.h:48-122) */
static void mult_rows(const gf_multiples_data &par, gf_bitmap apply, gfo_rng &rng, uint32_t r0, uint32_t r1,
                      const gf_table &t)
{
	int nab = (int)t.absmagSys.width;
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		int cmp = C<int>(t.comp, row);
		if (!isset(apply, cmp)) continue;
		float M1 = C<float>(t.absmag, row);
		C<float>(t.absmagSys, row, 0) = M1;
		float Ltot = gfo_exp10f(-0.4f * M1);
		int nc = 1;
		for (int i = 1; i < nab; i++) {
			float psec = tex1(par.secProb.data, par.secProb.n, par.secProb.tc, M1);
			float u = rng.uniform();
			if (u > psec) { C<float>(t.absmagSys, row, i) = 99.999f; continue; }
			float M2;
			if (par.algo == 3) M2 = M1;
			else {
				float pprim = tex1(par.cumLF.data, par.cumLF.n, par.cumLF.tc, M1);
				u = rng.uniform();
				if (par.algo == 1) u = (float)((double)pprim + (double)u * (1. - (double)pprim));
				M2 = tex1(par.invCumLF.data, par.invCumLF.n, par.invCumLF.tc, u);
				if (par.algo == 1 && M2 < M1) M2 = M1;
			}
			C<float>(t.absmagSys, row, i) = M2;
			Ltot += gfo_exp10f(-0.4f * M2);
			nc++;
		}
		C<float>(t.absmag, row) = (float)(-2.5 * (double)gfo_log10f(Ltot));
		C<int>(t.absmagSysNcomp, row) = nc;
	}
}

/* photometric errors drawn from the chunk stream, rows outer, bands inner
 * (modules/photometricErrors_host.cpp:112-131 draws from one serial stream
 * instead) */
static void perr_rows(const gf_photoerr_data &par, gfo_rng &rng, uint32_t r0, uint32_t r1, const gf_table &t)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (C<int>(t.hidden, row)) continue;
		for (int b = 0; b < par.nbands; b++) {
			int k0 = par.koff[b], k1 = par.koff[b + 1];
			float mag = C<float>(t.mags, row, b);
			if (k1 > k0) {
				float sg = (float)spline_at(par.kx + k0, par.ky + k0, k1 - k0, (double)mag);
				C<float>(t.obsMags, row, b) = mag + rng.gaussian(sg);
			} else {
				C<float>(t.obsMags, row, b) = mag;
			}
		}
	}
}

/* The module chain over rows [begin, end) in step-row chunks, one stream per
 * chunk, modules in pipeline order.  Running it with one module enabled is
 * the stand-alone kernel of that module. */
GFO_API int gfo_module_chain(const gf_rows *ks, const gf_chain *ch, uint32_t *st, uint32_t ns, const gf_table *tp)
{
	RowIt it{ks->begin, ks->step, ks->end};
	uint32_t nthr = it.nthreads();
	if (nthr > ns) return -1;
	const gf_table &t = *tp;
	for (uint32_t th = 0; th < nthr; th++) {
		uint32_t r0, r1;
		it.range(th, r0, r1);
		gfo_rng rng = rng_load(st, ns, th);
		if (ch->feh) feh_rows(*ch->feh, rng, r0, r1, t);
		if (ch->gfeh) gfeh_rows(*ch->gfeh, ch->gfeh_comps, rng, r0, r1, t);
		if (ch->mult) mult_rows(*ch->mult, ch->mult_comps, rng, r0, r1, t);
		if (ch->bond) bond_rows(*ch->bond, rng, r0, r1, t);
		if (ch->phot) phot_rows(*ch->phot, ch->phot_comps, r0, r1, t, ch->mult ? t.absmagSys : t.absmag);
		if (ch->perr) perr_rows(*ch->perr, rng, r0, r1, t);
		if (ch->v2pm) vel2pm_rows(*ch->v2pm, r0, r1, t);
		if (ch->gal2other_coordsys) gal2other_rows(ch->gal2other_coordsys, r0, r1, t);
		rng_store(st, ns, th, rng);
	}
	return 0;
}

/* ======================================================================== */
/* Host-side tables                                                          */
/* ======================================================================== */

/* photometry colour-table construction (modules/photometry_host.cpp:175-349):
 * per FeH knot a col(Mr) linear spline per colour, then per Mr a col(FeH)
 * spline through the knots, resampled on FeH0 + f*dFeH, Mr0 + m*dMr.
 * Input rows (Mr, FeH, c0..c{nc-1}); out: ntex float4 tables [nMr][nFeH]
 * for colours and flags; returns nFeH, nMr and texture coordinates. */
GFO_API int gfo_photometry_tables(const double *rows, int nrows, int ncolors, double dFeH, double dMr, float *colors,
                                  float *flags, int cap_texels, int *nFeH_out, int *nMr_out, float *tc)
{
	std::vector<double> feh_keys;
	for (int r = 0; r < nrows; r++) feh_keys.push_back(rows[r * (2 + ncolors) + 1]);
	std::sort(feh_keys.begin(), feh_keys.end());
	feh_keys.erase(std::unique(feh_keys.begin(), feh_keys.end()), feh_keys.end());
	int nk = (int)feh_keys.size();
	double FeH0 = feh_keys.front(), FeH1 = feh_keys.back();
	double Mr0 = rows[0], Mr1 = rows[0];
	for (int r = 0; r < nrows; r++) { Mr0 = std::min(Mr0, rows[r * (2 + ncolors)]); Mr1 = std::max(Mr1, rows[r * (2 + ncolors)]); }
	int nFeH = (int)((FeH1 - FeH0) / dFeH + 1);
	int nMr = (int)((Mr1 - Mr0) / dMr + 1);
	int ntex = (ncolors + 3) / 4;
	if ((size_t)nFeH * nMr * 4 * ntex > (size_t)cap_texels) return -1;
	/* per FeH knot: sorted (Mr, colours) */
	std::vector<std::vector<std::pair<double, std::vector<double>>>> knot(nk);
	for (int r = 0; r < nrows; r++) {
		const double *row = rows + (size_t)r * (2 + ncolors);
		int ki = (int)(std::lower_bound(feh_keys.begin(), feh_keys.end(), row[1]) - feh_keys.begin());
		knot[ki].push_back({row[0], std::vector<double>(row + 2, row + 2 + ncolors)});
	}
	std::vector<std::vector<double>> kMr(nk);
	std::vector<std::vector<std::vector<double>>> kC(nk);
	for (int ki = 0; ki < nk; ki++) {
		std::stable_sort(knot[ki].begin(), knot[ki].end(),
		                 [](const std::pair<double, std::vector<double>> &a, const std::pair<double, std::vector<double>> &b) { return a.first < b.first; });
		for (auto &e : knot[ki]) kMr[ki].push_back(e.first);
		kC[ki].resize(ncolors);
		for (int c = 0; c < ncolors; c++) for (auto &e : knot[ki]) kC[ki][c].push_back(e.second[c]);
	}
	std::vector<double> vcol(nk), vecol(nk);
	for (int ic = 0; ic < ncolors; ic++) {
		int tex = ic / 4, comp = ic % 4;
		for (int m = 0; m < nMr; m++) {
			double Mr = Mr0 + m * dMr;
			for (int ki = 0; ki < nk; ki++) {
				vcol[ki] = spline_at(kMr[ki].data(), kC[ki][ic].data(), (int)kMr[ki].size(), Mr);
				vecol[ki] = (kMr[ki].front() > Mr || Mr > kMr[ki].back()) ? 1.0 : 0.0;
			}
			for (int f = 0; f < nFeH; f++) {
				double FeH = FeH0 + f * dFeH;
				float color = (float)spline_at(feh_keys.data(), vcol.data(), nk, FeH);
				bool ef = feh_keys.front() > FeH || FeH > feh_keys.back() ||
				          spline_at(feh_keys.data(), vecol.data(), nk, FeH) != 0.;
				size_t o = (((size_t)tex * nMr + m) * nFeH + f) * 4 + comp;
				colors[o] = color;
				flags[o] = ef ? 1.f : 0.f;
			}
		}
	}
	/* texcoord_from_range(0, n, x0, x0 + n*d) (common/cux.cpp:389-397) */
	auto tcr = [](float imgx, float imgy, float x, float y, float *o) {
		o[0] = (-imgy * x + imgx * y) / (imgx - imgy);
		o[1] = (imgx - imgy) / (x - y);
	};
	tcr(0.f, (float)nFeH, (float)FeH0, (float)(FeH0 + nFeH * dFeH), tc);
	tcr(0.f, (float)nMr, (float)Mr0, (float)(Mr0 + nMr * dMr), tc + 2);
	*nFeH_out = nFeH;
	*nMr_out = nMr;
	return ntex;
}

/* the oracle's double functions: 0 sin, 1 cos, 2 atan2(x, y), 3 asin, 4 atan */
GFO_API void gfo_math_d(int fn, const double *x, const double *y, double *out, size_t n)
{
	for (size_t i = 0; i < n; i++) {
		double s, c;
		switch (fn) {
		case 0: gfo_sincos(x[i], &s, &c); out[i] = s; break;
		case 1: gfo_sincos(x[i], &s, &c); out[i] = c; break;
		case 2: out[i] = gfo_atan2(x[i], y[i]); break;
		case 3: out[i] = gfo_asin(x[i]); break;
		case 4: out[i] = gfo_atan(x[i]); break;
		}
	}
}

/* glibc's float functions, for pinning the oracle's own */
GFO_API void gfo_libm_f(int fn, const float *x, const float *y, float *out, size_t n)
{
	for (size_t i = 0; i < n; i++) {
		switch (fn) {
		case 0: out[i] = expf(x[i]); break;
		case 1: out[i] = logf(x[i]); break;
		case 2: out[i] = sinf(x[i]); break;
		case 3: out[i] = cosf(x[i]); break;
		case 4: out[i] = tanf(x[i]); break;
		case 5: out[i] = powf(x[i], y[i]); break;
		case 6: out[i] = exp10f(x[i]); break;
		case 7: out[i] = log10f(x[i]); break;
		}
	}
}
