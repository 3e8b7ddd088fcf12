// Device helpers shared by the skygen and module kernels: column access,
// texture sampling, coordinate transforms and the density models.
#pragma once

#include <hip/hip_runtime.h>
#include "../../include/galfast_hip.h"
#include "gf_math.h"
#include "gf_rng.h"
#include "gf_sky.h"

#define GF_D __device__ __forceinline__

// Element (row, plane) of a structure-of-arrays column (common/column.h:131-136).
template <typename T>
GF_D T &col(const gf_col &c, uint32_t row, uint32_t plane = 0)
{
	return *((T *)((char *)c.ptr + (size_t)plane * c.pitch) + row);
}

GF_D bool bit_isset(const gf_bitmap &m, int bit)
{
	if ((unsigned)bit >= 64u) return false;   // no component outside [0, 64)
	if (bit < 32) return ((1u << bit) & m.bits[0]) != 0;
	return ((1u << (bit - 32)) & m.bits[1]) != 0;
}

// Linear filtering of a 1D table with clamping, in the texture coordinate
// convention of sample_impl (common/cux.h:849-855): xi = (x - tc0)*tc1 + 0.5,
// texel centres at integer + 0.5.  The reference's CUDA path filters in
// hardware and its CPU path takes the nearest texel (common/cux.h:778-810);
// here the interpolation is evaluated exactly in FP32.
GF_D float lerp_coord(float x, float tc0, float tc1, int n, int &i)
{
	float xi = (x - tc0) * tc1 + 0.5f;
	float t = xi - 0.5f;
	if (!(t > 0.f)) { i = 0; return 0.f; }
	float tmax = (float)(n - 1);
	if (t >= tmax) { i = n > 1 ? n - 2 : 0; return n > 1 ? 1.f : 0.f; }
	float fl = __builtin_floorf(t);
	i = (int)fl;
	return t - fl;
}

GF_D float tex1d(const float *data, int n, float tc0, float tc1, float x)
{
	int i;
	float a = lerp_coord(x, tc0, tc1, n, i);
	if (n == 1) return data[0];
	float v0 = data[i], v1 = data[i + 1];
	return v0 + a * (v1 - v0);
}

GF_D float tex1d(const gf_tex1d &t, float x) { return tex1d(t.data, t.n, t.tc[0], t.tc[1], x); }

GF_D float tex3d(const gf_tex3d &t, float x, float y, float z)
{
	if (!t.data) return 0.f;
	int i, j, k;
	float a = lerp_coord(x, t.tc[0][0], t.tc[0][1], t.nx, i);
	float b = lerp_coord(y, t.tc[1][0], t.tc[1][1], t.ny, j);
	float c = lerp_coord(z, t.tc[2][0], t.tc[2][1], t.nz, k);
	int i1 = t.nx > 1 ? i + 1 : i, j1 = t.ny > 1 ? j + 1 : j, k1 = t.nz > 1 ? k + 1 : k;
	const size_t sx = 1, sy = (size_t)t.nx, sz = (size_t)t.nx * t.ny;
	auto at = [&](int ii, int jj, int kk) { return t.data[ii * sx + jj * sy + kk * sz]; };
	float c00 = at(i, j, k) + a * (at(i1, j, k) - at(i, j, k));
	float c10 = at(i, j1, k) + a * (at(i1, j1, k) - at(i, j1, k));
	float c01 = at(i, j, k1) + a * (at(i1, j, k1) - at(i, j, k1));
	float c11 = at(i, j1, k1) + a * (at(i1, j1, k1) - at(i, j1, k1));
	float c0 = c00 + b * (c10 - c00);
	float c1 = c01 + b * (c11 - c01);
	return c0 + c * (c1 - c0);
}

// transform(): translation followed by a 3x3 rotation (modules/module_lib.h:122-148).
GF_D void transform3(float &x, float &y, float &z, const float T[3], const float M[3][3])
{
	float vx = x + T[0], vy = y + T[1], vz = z + T[2];
	x = M[0][0] * vx + M[0][1] * vy + M[0][2] * vz;
	y = M[1][0] * vx + M[1][1] * vy + M[1][2] * vz;
	z = M[2][0] * vx + M[2][1] * vy + M[2][2] * vz;
}

// Density of a model at heliocentric Galactic (x, y, z), before the LF factor:
// expDisk::rho (skygen/model_expDisk.h:71-121) and
// powerLawEllipsoid::rho (skygen/model_powerLawEllipsoid.h:60-105).
GF_D float model_rho_xyz(const gf_model &m, float x, float y, float z)
{
	transform3(x, y, z, m.T, m.M);
	if (m.kind == GF_MODEL_EXPDISK) {
		float r2 = x * x + y * y;
		if (m.r_cut2 != 0.f && r2 + z * z > m.r_cut2) return 0.f;
		float r = gf_sqrtf(r2);
		float e = -r / m.l - gf_fabsf(z) / m.h;
		return m.f * gf_expf(e);
	} else {
		float DSq = x * x + y * y + z * z;
		if (m.rmaxSq != 0.f && (DSq < m.rminSq || DSq >= m.rmaxSq)) return 0.f;
		float yb = y / m.ba, zc = z / m.ca;
		float rSq = x * x + yb * yb + zc * zc;
		return m.f * gf_powf(rSq, 0.5f * m.n);
	}
}

// lambert::deproject (skygen/skygen.h:131-146), with the reference's mixed
// float/double promotions kept.
GF_D void lambert_deproject(const gf_lambert &p, double &l, double &b, const float x, const float y)
{
	double r2 = (double)(x * x + y * y);
	if (r2 == 0.) { b = p.b0; l = 0.; return; }
	double ir = 1. / sqrt(r2);
	double sc = 1. / ir * sqrt(1. - 0.25 * r2);
	double cc = 1. - 0.5 * r2;
	b = gfm::asin(cc * (double)p.sb + (double)y * sc * (double)p.cb * ir);
	float ysb = y * p.sb;
	l = p.l0 + gfm::atan2((double)x * sc, (double)p.cb * cc / ir - (double)ysb * sc);
}

// lambert_deproject that also gives the direction cosines (cos b cos l,
// cos b sin l, sin b) of the point from the deprojection's own terms: with
// rho = 1/ir, cos b sin(l - l0) = x sc / rho and cos b cos(l - l0) =
// (cb cc rho - y sb sc) / rho (the atan2 arguments over rho), sin b the asin
// argument; rotated by l0 with the projection's cos l0, sin l0.
GF_D void lambert_deproject_dir(const gf_lambert &p, double &l, double &b, double &cbcl, double &cbsl, double &sb,
                                const float x, const float y)
{
	double r2 = (double)(x * x + y * y);
	if (r2 == 0.) {
		b = p.b0; l = 0.;
		cbcl = (double)p.cb; cbsl = 0.; sb = (double)p.sb;
		return;
	}
	double ir = 1. / sqrt(r2);
	double sc = 1. / ir * sqrt(1. - 0.25 * r2);
	double cc = 1. - 0.5 * r2;
	const double S = cc * (double)p.sb + (double)y * sc * (double)p.cb * ir;
	b = gfm::asin(S);
	float ysb = y * p.sb;
	const double P = (double)x * sc, Q = (double)p.cb * cc / ir - (double)ysb * sc;
	l = p.l0 + gfm::atan2(P, Q);
	const double Pr = P * ir, Qr = Q * ir;
	cbcl = (double)p.cl * Qr - (double)p.sl * Pr;
	cbsl = (double)p.sl * Qr + (double)p.cl * Pr;
	sb = S;
}

static constexpr float POGSON = 0.4605170185988091f;   // 0.2 ln 10 (skygen/skygen.cu.h:466)
// Valid absolute-magnitude-bin window of DM bin k of one pixel
// (the cell skips of skygen/skygen.cu.h:551-559): cells (im = k - iM, iM)
// with 0 <= im < nm, 0 <= iM < nM, not cut by the distance limits and with
// m0 + dm*im + Am <= m1.  Returns false if the window is empty.
GF_D bool dm_window(const gf_skygen_params &sp, int k, float D, float Am, int &iMlo, int &iMhi)
{
	if ((sp.dmin != 0.f || sp.dmax != 0.f) && (sp.dmin > D || sp.dmax <= D)) return false;
	int lo = k - (sp.nm - 1); if (lo < 0) lo = 0;
	int hi = k; if (hi > sp.nM - 1) hi = sp.nM - 1;
	if (lo > hi) return false;
	// flux limit: the largest im passing m0 + dm*im + Am <= m1, found by the
	// same float expression the reference evaluates per cell
	int imlo = k - hi, imhi = k - lo;
	if (sp.m0 + sp.dm * imlo + Am > sp.m1) return false;
	if (!(sp.m0 + sp.dm * imhi + Am > sp.m1)) { iMlo = lo; iMhi = hi; return true; }
	int a = imlo, b = imhi;   // pass at a, fail at b
	while (b - a > 1) {
		int mid = (a + b) >> 1;
		if (sp.m0 + sp.dm * mid + Am > sp.m1) b = mid; else a = mid;
	}
	iMlo = k - a; iMhi = hi;
	return true;
}

// compute_pos (skygen/skygen.cu.h:216-229): distance and heliocentric XYZ of
// DM bin (im, M) of a pixel, and the per-beam minimum extinction.
GF_D void cell_pos(const gf_skygen_params &sp, const gf_extinction &ext, const gf_pencil_beam &pix,
                   float M, int im, float &D, float &Am, float &x, float &y, float &z)
{
	float m = sp.m0 + im * sp.dm;
	float DM = m - M;
	D = gf_pow10f(0.2f * DM + 1.f);
	int xidx = pix.extIdx / 2048;
	int yidx = pix.extIdx % 2048;
	Am = ext.beam.data ? tex3d(ext.beam, (float)xidx, (float)yidx, DM) : 0.f;
	x = D * pix.cl * pix.cb;
	y = D * pix.sl * pix.cb;
	z = D * pix.sb;
}

// Absolute-magnitude bin of a draw: the smallest a in [lo, hi] with
// P[a + 1] > tgt (hi if there is none) for the nondecreasing FP64 LF prefix
// P, as a binary search over [lo, hi] finds it.  A guess table G of
// LF_GUESS entries (G[q] = that bin over the whole range for the q-th
// equally spaced value of P) gives the start; walking down while P[a] > tgt
// and up while P[a + 1] <= tgt then ends on the same bin as the binary
// search from any start, in a step or two instead of log2(nM).
constexpr int LF_GUESS = 1024;

GF_D void lf_guess_build(const double *P, int nM, uint16_t *G, int tid, int nthr)
{
	const double P0 = P[0], Pn = P[nM];
	for (int q = tid; q < LF_GUESS; q += nthr) {
		double t = P0 + (Pn - P0) * ((double)q / (double)LF_GUESS);
		int a = 0, b = nM - 1;
		while (a < b) {
			int mid = (a + b) >> 1;
			if (P[mid + 1] > t) b = mid; else a = mid + 1;
		}
		G[q] = (uint16_t)a;
	}
}

GF_D double lf_guess_scale(const double *P, int nM)
{
	const double r = P[nM] - P[0];
	return r > 0.0 ? (double)LF_GUESS / r : 0.0;
}

// The guess table's thresholds: GP[q] = (P[G[q]], P[G[q] + 1]), so that a
// draw reads its guess and both the guess's neighbours at once.
GF_D void lf_guess_pairs(const double *P, const uint16_t *G, double2 *GP, int tid, int nthr)
{
	for (int q = tid; q < LF_GUESS; q += nthr) {
		const int g = G[q];
		GP[q] = make_double2(P[g], P[g + 1]);
	}
}

// lf_bin with the thresholds of the guess from GP (the same bin).
GF_D int lf_bin_pairs(const double *P, const uint16_t *G, const double2 *GP, double scale, int lo, int hi,
                      double tgt)
{
	double qd = (tgt - P[0]) * scale;
	int q = qd > 0.0 ? (qd < (double)(LF_GUESS - 1) ? (int)qd : LF_GUESS - 1) : 0;
	const int g = G[q];
	const double2 pp = GP[q];
	int a = g < lo ? lo : (g > hi ? hi : g);
	double pa = pp.x, pa1 = pp.y;
	if (a != g) {
		pa = P[a];
		pa1 = P[a + 1];
	}
	if (a > lo && pa > tgt) {
		a--;
		while (a > lo && P[a] > tgt) a--;
	} else if (a < hi && pa1 <= tgt) {
		a++;
		while (a < hi && P[a + 1] <= tgt) a++;
	}
	return a;
}

GF_D int lf_bin(const double *P, const uint16_t *G, double scale, int lo, int hi, double tgt)
{
	double qd = (tgt - P[0]) * scale;
	int q = qd > 0.0 ? (qd < (double)(LF_GUESS - 1) ? (int)qd : LF_GUESS - 1) : 0;
	int a = G[q];
	a = a < lo ? lo : (a > hi ? hi : a);
	// both neighbours of the guess are read at once; a walk down ends
	// where the walk up would not start (P is nondecreasing)
	const double pa = P[a], pa1 = P[a + 1];
	if (a > lo && pa > tgt) {
		a--;
		while (a > lo && P[a] > tgt) a--;
	} else if (a < hi && pa1 <= tgt) {
		a++;
		while (a < hi && P[a + 1] <= tgt) a++;
	}
	return a;
}

// The random part of one star of draw_stars (skygen/skygen.cu.h:376-385,
// 404-405): position inside the pixel (2 uniforms) and, when that stays
// inside the projection, the absolute and apparent magnitude (2 more).
GF_D bool star_draws(const gf_skygen_params &sp, const gf_pencil_beam &pix, float M, int im, gf_mwc &rng,
                     float &x, float &y, float &Mtmp, float &mtmp)
{
	x = pix.X;
	y = pix.Y;
	x += pix.dx * (rng.uniform() - 0.5f);
	y += pix.dx * (rng.uniform() - 0.5f);
	if (x * x + y * y > 2.f) return false;
	Mtmp = M + sp.dM * (rng.uniform() - 0.5f);
	mtmp = sp.m0 + sp.dm * ((float)im + rng.uniform() - 0.5f);
	return true;
}

// Columns of a star outside the projection: only `hidden` is meaningful
// (the reference writes nothing else); the rest are set to zero.
GF_D void write_star_outside(const gf_skygen_cols &c, uint32_t row, const gf_pencil_beam &pix, int comp, float x,
                             float y)
{
	col<int>(c.hidden, row) = 1;
	col<int>(c.projIdx, row) = pix.projIdx;
	col<float>(c.projXY, row, 0) = x; col<float>(c.projXY, row, 1) = y;
	col<double>(c.lb, row, 0) = 0.; col<double>(c.lb, row, 1) = 0.;
	col<float>(c.M, row) = 0.f; col<float>(c.DM, row) = 0.f;
	col<float>(c.XYZ, row, 0) = 0.f; col<float>(c.XYZ, row, 1) = 0.f; col<float>(c.XYZ, row, 2) = 0.f;
	col<int>(c.comp, row) = comp;
	col<float>(c.Am, row) = 0.f; col<float>(c.AmInf, row) = 0.f;
}

// The deterministic rest of draw_stars (skygen/skygen.cu.h:387-456):
// deprojection, distance, XYZ, extinction and the flux limit, followed by
// the clipper's footprint test.
template <bool FULL = true>
GF_D void write_star(const gf_skygen_params &sp, int comp, const gf_extinction &ext, const gf_footprint &foot,
                     const gf_pencil_beam &pix, float x, float y, float Mtmp, float mtmp, float AmMin,
                     const gf_skygen_cols &c, uint32_t row)
{
	double l, b, cbcl, cbsl, sbd;
	lambert_deproject_dir(sp.proj[pix.projIdx], l, b, cbcl, cbsl, sbd, x, y);
	int outside = outside_footprint<FULL>(foot, l, b);
	l *= DBL_R2D;
	if (l < 0.) l += 360.;
	if (l > 360.) l -= 360.;
	b *= DBL_R2D;
	float DM = mtmp - Mtmp;
	float D = gf_pow10f(0.2f * DM + 1.f);
	float X = D * (float)cbcl, Y = D * (float)cbsl, Z = D * (float)sbd;
	float Am0 = 0.f, Am1 = 0.f;
	const gf_tex3d &em = pix.projIdx == 0 ? ext.north : ext.south;
	if (em.data) {
		Am0 = tex3d(em, x, y, DM);
		Am1 = tex3d(em, x, y, 100.f);
	}
	float AmInf = Am1 > Am0 ? Am1 : Am0;
	if (AmMin > Am0) AmInf = -AmMin;
	int hidden = (mtmp + Am0 > sp.m1) ? 1 : 0;
	hidden |= outside;
	col<int>(c.projIdx, row) = pix.projIdx;
	col<float>(c.projXY, row, 0) = x;
	col<float>(c.projXY, row, 1) = y;
	col<double>(c.lb, row, 0) = l;
	col<double>(c.lb, row, 1) = b;
	col<float>(c.M, row) = Mtmp;
	col<float>(c.DM, row) = DM;
	col<float>(c.XYZ, row, 0) = X;
	col<float>(c.XYZ, row, 1) = Y;
	col<float>(c.XYZ, row, 2) = Z;
	col<int>(c.comp, row) = comp;
	col<float>(c.Am, row) = Am0;
	col<float>(c.AmInf, row) = AmInf;
	col<int>(c.hidden, row) = hidden;
}

// The cumulative-table draw's use of the streams.  Unit u = p*K + k of a
// model belongs to stream s = u mod ns with index i = u div ns; from the
// streams' states at the plan, the unit's Poisson draw starts i*2^32 draws
// on, its star j (0-based within the unit) i*2^32 + 2^31 + 8 j draws on (a
// star takes 3 or 5), and once the model is planned the streams move
// ceil(U / ns) * 2^32 draws on, past every unit.  Positions are reached by
// jump-ahead (gf_mwc_mod) with the streams' jump tables.
constexpr int GF_UNIT_SPAN_BITS = 32;
constexpr int GF_EMIT_AT_BIT = 31;
constexpr int GF_STAR_SPAN_BITS = 3;

// The absolute-magnitude window and minimum extinction of a unit that holds
// stars (its distance cut passed): DM bin k's first cell in crawl order as
// cell_pos reads it, then dm_window; the distance itself only when a
// distance cut is configured.
GF_D bool unit_window(const gf_skygen_params &sp, const gf_extinction &ext, const gf_pencil_beam &pix, int k,
                      float &Am, int &lo, int &hi)
{
	const int iM0 = k < sp.nM - 1 ? k : sp.nM - 1;
	const int im0 = k - iM0;
	const float M = sp.M1 - iM0 * sp.dM;
	const float m = sp.m0 + im0 * sp.dm;
	const float DM = m - M;
	const int xidx = pix.extIdx / 2048, yidx = pix.extIdx % 2048;
	Am = ext.beam.data ? tex3d(ext.beam, (float)xidx, (float)yidx, DM) : 0.f;
	const float D = (sp.dmin != 0.f || sp.dmax != 0.f) ? gf_pow10f(0.2f * DM + 1.f) : 0.f;
	return dm_window(sp, k, D, Am, lo, hi);
}

// The absolute-magnitude bin of a draw by binary search over the FP64 LF
// prefix P: the smallest a in [lo, hi] with P[a + 1] > tgt, else hi (the
// bin lf_bin_pairs finds with its guess table).
GF_D int lf_bin_search(const double *P, int lo, int hi, double tgt)
{
	int a = lo, b = hi;
	while (a < b) {
		const int mid = (a + b) >> 1;
		if (P[mid + 1] > tgt) b = mid; else a = mid + 1;
	}
	return a;
}

// One star of DM bin k: its absolute-magnitude bin from the FP64 cumulative
// LF window [lo, hi] (one uniform; bin(lo, hi, tgt) finds it in P), the
// draws of draw_stars and the deterministic rest, written to row `row` of
// out.
template <bool FULL, typename Bin>
GF_D void emit_star(const gf_skygen_params &sp, int comp, const gf_extinction &ext, const gf_footprint &foot,
                    const gf_pencil_beam &pix, int k, int lo, int hi, float Am, const double *P, Bin bin,
                    gf_mwc &rng, const gf_skygen_cols &out, uint32_t row)
{
	const double tgt = P[lo] + (double)rng.uniform() * (P[hi + 1] - P[lo]);
	const int iM = bin(lo, hi, tgt);
	const float M = sp.M1 - iM * sp.dM;
	float x, y, Mt = 0.f, mt = 0.f;
	if (!star_draws(sp, pix, M, k - iM, rng, x, y, Mt, mt)) {
		write_star_outside(out, row, pix, comp, x, y);
		return;
	}
	write_star<FULL>(sp, comp, ext, foot, pix, x, y, Mt, mt, Am, out, row);
}

