// Sky geometry shared by the kernels and the host: the Galactic ->
// equatorial rotation, the footprint regions' point test, the pole-centred
// Lambert deprojection and the pixel-grid sample test.  One definition for
// both sides, so the host's pixelization of a request and the device's
// coverage of a catalog's sky count the same samples.
#pragma once

#include "../../include/galfast_hip.h"
#include "gf_math.h"

static constexpr double DBL_R2D = 57.2957795130823208767981548;
static constexpr double DBL_D2R = 0.01745329251994329576923691;

// J2000 Galactic <-> equatorial constants (modules/module_lib.h:62-71)
namespace galequ_c {
static constexpr double angp = 0.01745329251994329576923691 * 192.859508333;
static constexpr double l0 = 0.01745329251994329576923691 * 32.932;
static constexpr double ce = 0.88998740217659689;
static constexpr double se = 0.45598511375586859;
static constexpr double pi2 = 6.28318530717958647692528677;
static constexpr double d2r = 0.01745329251994329576923691;
}

// Galactic (l, b) -> equatorial (ra, dec), radians, ra in [0, 2 pi)
// (galequ, modules/gal2other_gpu.cu.h:39-60)
GF_HD void gal_to_equ(double l, double b, double &ra, double &dec)
{
	using namespace galequ_c;
	double cb, sb, cl, sl;
	gfm::sincos(b, &sb, &cb);
	gfm::sincos(l - l0, &sl, &cl);
	ra = gfm::atan2(cb * cl, sb * ce - cb * se * sl) + angp;
	dec = gfm::asin(cb * ce * sl + sb * se);
	while (ra < 0.) ra += pi2;
	while (ra >= pi2) ra -= pi2;
}

// One footprint region (footprint.cpp:486-727): 1 if the Galactic direction
// (l, b) [radians] lies outside it.  EQU = false compiles rects in Galactic
// coordinates only (the equatorial transform's registers stay out of the
// kernels that never need it).
template <bool EQU = true>
GF_HD int outside_region(const gf_foot_region &f, double l, double b)
{
	const double R2D = DBL_R2D;
	if (f.bmin > 0. && fabs(b * R2D) < f.bmin) return 1;
	if (f.kind == 1) {
		double lon = l, lat = b;
		if (EQU && f.equ) gal_to_equ(l, b, lon, lat);
		double ld = lon * R2D, bd = lat * R2D;
		while (ld < 0.) ld += 360.;
		while (ld >= 360.) ld -= 360.;
		if (bd < f.b0 || bd > f.b1) return 1;
		double lo = f.l0, hi = f.l1;
		if (hi - lo >= 360.) return 0;
		double t = ld - lo;
		while (t < 0.) t += 360.;
		while (t >= 360.) t -= 360.;
		return t > (hi - lo) ? 1 : 0;
	}
	if (f.kind == 2) {
		double cb, sb, cl, sl;
		gfm::sincos(b, &sb, &cb);
		gfm::sincos(l, &sl, &cl);
		double d = cb * cl * f.cb * f.cl + cb * sl * f.cb * f.sl + sb * f.sb;
		if (d < f.cosr) return 1;
		return (f.hole > 0. && 1. - d < f.hole) ? 1 : 0;
	}
	return 1;
}

// The footprint (the union of its regions; no regions = the whole sky).
// FULL = false: the footprint is one region, not an equatorial rect
// (footprint_is_simple), and only that region is read.
template <bool FULL = true>
GF_HD int outside_footprint(const gf_footprint &f, double l, double b)
{
	if (f.nregion <= 0) return 0;
	if (!FULL) return outside_region<false>(f.r[0], l, b);
	for (int k = 0; k < f.nregion && k < GF_FOOT_MAX; k++)
		if (!outside_region(f.r[k], l, b)) return 0;
	return 1;
}

inline bool footprint_is_simple(const gf_footprint &f)
{
	return f.nregion == 0 || (f.nregion == 1 && !(f.r[0].kind == 1 && f.r[0].equ));
}

// Deprojection of a pole-centred Lambert plane (hemi 0: pole b = +90 deg,
// l0 = 90 deg; hemi 1: b = -90 deg, l0 = 270 deg), radians, l in [0, 2 pi)
// (peyton::math::lambert::deproject with b0 = +-pi/2).
GF_HD void pole_deproject(int hemi, double x, double y, double &l, double &b)
{
	const double PI = 3.14159265358979323846;
	double r2 = x * x + y * y;
	double sb = fmin(1.0, fmax(-1.0, 1. - 0.5 * r2));
	if (hemi == 0) { b = gfm::asin(sb); l = 0.5 * PI + gfm::atan2(x, -y); }
	else { b = -gfm::asin(sb); l = 1.5 * PI + gfm::atan2(x, y); }
	while (l < 0.) l += 2 * PI;
	while (l >= 2 * PI) l -= 2 * PI;
}

// Sample point (x, y) of hemisphere `hemi`'s Lambert plane inside the
// footprint?  Beams are tested on the point's unit vector, which the
// equal-area projection gives without transcendentals:
// (-x s, -y s, 1 - r^2/2) in the north, (x s, -y s, r^2/2 - 1) in the south,
// s = sqrt(1 - r^2/4); rects and the zone of avoidance on the deprojected
// (l, b).
GF_HD bool sample_inside(const gf_footprint &foot, int hemi, double x, double y)
{
	double r2 = x * x + y * y;
	if (r2 > 2.0) return false;
	double s = sqrt(fmax(0.0, 1.0 - 0.25 * r2)), z = 1.0 - 0.5 * r2;
	double v0 = hemi == 0 ? -x * s : x * s, v1 = -y * s, v2 = hemi == 0 ? z : -z;
	if ((hemi == 0 && v2 < 0) || (hemi == 1 && v2 >= 0)) return false;
	if (foot.nregion <= 0) return true;
	double l = 0., b = 0.;
	bool have_lb = false;
	for (int k = 0; k < foot.nregion && k < GF_FOOT_MAX; k++) {
		const gf_foot_region &f = foot.r[k];
		if (f.kind == 2 && f.bmin <= 0.) {
			double d = v0 * (f.cb * f.cl) + v1 * (f.cb * f.sl) + v2 * f.sb;
			if (d >= f.cosr && !(f.hole > 0. && 1. - d < f.hole)) return true;
			continue;
		}
		if (!have_lb) { pole_deproject(hemi, x, y, l, b); have_lb = true; }
		if (f.kind == 2) {
			if (fabs(b * DBL_R2D) < f.bmin) continue;
			double d = v0 * (f.cb * f.cl) + v1 * (f.cb * f.sl) + v2 * f.sb;
			if (d >= f.cosr && !(f.hole > 0. && 1. - d < f.hole)) return true;
			continue;
		}
		if (!outside_region(f, l, b)) return true;
	}
	return false;
}

