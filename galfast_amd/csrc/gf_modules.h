// Per-star column modules as device functions over one thread's row range.
// Each follows the reference kernel named in its comment; the global kernels
// (modules.hip) and the fused chain (chain.hip) call them.
#pragma once

#include "gf_device.h"

static constexpr float ABSMAG_NOT_PRESENT = 99.999f;   // modules/module_lib.h:57

// One row of os_FeH_kernel (modules/FeH_gpu.cu.h:60-84); false when the row's
// component is not handled by the module.  Disk rows draw their component
// choice first; the Gaussian deviate, the last draw of both branches, is
// drawn once after them (offs + g and g + offs are the same sum).
GF_D bool feh_row(const gf_FeH_data &par, gf_mwc &rng, int cmp, float x, float y, float z, float &feh)
{
	const bool disk = bit_isset(par.comp_thin, cmp) || bit_isset(par.comp_thick, cmp);
	if (!disk && !bit_isset(par.comp_halo, cmp)) return false;
	int i = 2;
	float aZ = 0.f;
	if (disk) {
		float p = rng.uniform() * (par.A[0] + par.A[1]);
		i = p < par.A[0] ? 0 : 1;
		transform3(x, y, z, par.T, par.M);
		float muD = par.muInf + par.DeltaMu * gf_expf(-gf_fabsf(z) / par.Hmu);
		aZ = muD - 0.067f;
	}
	const float g = rng.gaussian(par.sigma[i]);
	feh = disk ? g + aZ + par.offs[i] : g + par.offs[2];
	return true;
}

// The draws feh_row makes for a row of component cmp, without the
// arithmetic: the stream ends where feh_row leaves it.
GF_D void feh_skip(const gf_FeH_data &par, gf_mwc &rng, int cmp)
{
	float yy, r2;
	if (bit_isset(par.comp_thin, cmp) || bit_isset(par.comp_thick, cmp)) {
		(void)rng.uniform();
		rng.polar(yy, r2);
	} else if (bit_isset(par.comp_halo, cmp)) {
		rng.polar(yy, r2);
	}
}

// os_FeH_kernel body (modules/FeH_gpu.cu.h:55-86)
GF_D void mod_FeH(const gf_FeH_data &par, gf_mwc &rng, uint32_t r0, uint32_t r1,
                  const gf_col &comp, const gf_col &hidden, const gf_col &XYZ, const gf_col &FeH)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		float feh;
		if (feh_row(par, rng, col<int>(comp, row), col<float>(XYZ, row, 0), col<float>(XYZ, row, 1),
		            col<float>(XYZ, row, 2), feh))
			col<float>(FeH, row) = feh;
	}
}

// One row of os_GaussianFeH_kernel (modules/GaussianFeH_gpu.cu.h:55-72) for
// a visible row of component cmp.
GF_D void gfeh_row(const gf_GaussianFeH_data &par, gf_bitmap apply, gf_mwc &rng, int cmp, uint32_t row,
                   const gf_col &FeH_comp, const gf_col &FeH)
{
	if (!bit_isset(apply, cmp)) return;
	int at = 0;
	if (par.ncomp != 1) {
		float u = rng.uniform();
		while (at < par.ncomp - 1 && par.pcum[at] < u) ++at;
	}
	col<float>(FeH, row) = par.mean[at] + rng.gaussian(par.sigma[at]);
	if (FeH_comp.ptr) col<int>(FeH_comp, row) = at;
}

GF_D void gfeh_skip(const gf_GaussianFeH_data &par, gf_bitmap apply, gf_mwc &rng, int cmp)
{
	if (!bit_isset(apply, cmp)) return;
	float yy, r2;
	if (par.ncomp != 1) (void)rng.uniform();
	rng.polar(yy, r2);
}

// os_GaussianFeH_kernel body (modules/GaussianFeH_gpu.cu.h:55-72)
GF_D void mod_GaussianFeH(const gf_GaussianFeH_data &par, gf_bitmap apply, gf_mwc &rng,
                          uint32_t r0, uint32_t r1, const gf_col &comp, const gf_col &hidden,
                          const gf_col &FeH_comp, const gf_col &FeH)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		gfeh_row(par, apply, rng, col<int>(comp, row), row, FeH_comp, FeH);
	}
}

// Powers of one row's |Z| and R^2 (both >= 0): each log is taken once, on
// first use, and every power equals gf_powf's.
struct BondPow {
	float aZ, R2;
	double lZ = 0.0, lR = 0.0;
	bool hZ = false, hR = false;
	GF_D BondPow(float Rsquared, float Z) : aZ(gf_fabsf(Z)), R2(Rsquared) {}
	GF_D float z(float y)
	{
		return gf_powf_nonneg(aZ, y, [this]() {
			if (!hZ) { lZ = gfm::log((double)aZ); hZ = true; }
			return lZ;
		});
	}
	GF_D float r(float y)
	{
		return gf_powf_nonneg(R2, y, [this]() {
			if (!hR) { lR = gfm::log((double)R2); hR = true; }
			return lR;
		});
	}
};

// modfun (modules/Bond2010_gpu.cu.h:106-110): a + b|Z|^c + d R^e, with the
// exponent 0.5*e formed in double and narrowed as in the reference.
GF_D float bond_modfun(BondPow &bp, const float val[5])
{
	float p1 = bp.z(val[2]);
	float p2 = bp.r((float)(0.5 * (double)val[4]));
	return val[0] + val[1] * p1 + val[3] * p2;
}

// trivar_gaussK_draw (modules/Bond2010_gpu.cu.h:81-103): s11, s22, s33 are
// standard deviations, s12, s13, s23 covariances.
GF_D void bond_trivar(float y[3], const float s[6], gf_mwc &rng)
{
	float b10 = s[1];
	float b11 = s[3] * s[3];
	float b20 = s[2];
	float b21 = s[4];
	float b22 = s[5] * s[5];
	float a00 = s[0];
	float oneDivA00 = 1.0f / a00;
	float a10 = b10 * oneDivA00;
	float a11 = gf_sqrtf(b11 - a10 * a10);
	float a20 = b20 * oneDivA00;
	float a21 = (b21 - a20 * a10) / a11;
	float a22 = gf_sqrtf(b22 - a20 * a20 - a21 * a21);
	float z0 = rng.gaussian(1.0f);
	float z1 = rng.gaussian(1.0f);
	float z2 = rng.gaussian(1.0f);
	y[0] += a00 * z0;
	y[1] += a10 * z0 + a11 * z1;
	y[2] += a20 * z0 + a21 * z1 + a22 * z2;
}

// One row of os_Bond2010_kernel (modules/Bond2010_gpu.cu.h:185-216); false
// when the row's component is not handled by the module.  The moments are
// evaluated in the component's own branch (get_disk_kinematics, :130-140;
// get_halo_kinematics, :142-159), whose parameters are the same for every
// lane; the three Gaussian deviates of trivar_gaussK_draw, which both
// branches draw last, are drawn after the branches, once per wave however
// its rows mix the components (every row's draws stay in their order: a
// disk row's uniform, then its deviates).  ONE_DISK: disk_1 and disk_2 hold
// the same bits, so the disk moments are read from disk_1 whatever the draw
// chooses (the uniform is drawn as always).
template <bool ONE_DISK = false>
GF_D bool bond_row(const gf_Bond2010_data &par, gf_mwc &rng, int cmp, float x, float y, float z, float v[3])
{
	transform3(x, y, z, par.T, par.M);
	const float Rsquared = (float)(1e-6 * (double)(x * x + y * y));
	const float Z = (float)(1e-3 * (double)z);
	BondPow bp(Rsquared, Z);
	const bool disk = bit_isset(par.comp_thin, cmp) || bit_isset(par.comp_thick, cmp);
	if (!disk && !bit_isset(par.comp_halo, cmp)) return false;
	float s[6];
	if (disk) {
		float p = rng.uniform();
		const bool first = ONE_DISK || p < par.fk;
		// the chosen set's coefficients by value (with ONE_DISK, disk_1's,
		// the same for every lane)
		auto pick = [first](const float (&a)[5], const float (&b)[5], float c[5]) {
#pragma unroll
			for (int e = 0; e < 5; e++) c[e] = first ? a[e] : b[e];
		};
#pragma unroll
		for (int i = 0; i < 3; i++) {
			float c[5];
			pick(par.disk_1.m[i], par.disk_2.m[i], c);
			v[i] = bond_modfun(bp, c);
		}
		if (v[1] > 0.f) v[1] = 0.f;
#pragma unroll
		for (int i = 0; i < 6; i++) {
			float c[5];
			pick(par.disk_1.ss[i], par.disk_2.ss[i], c);
			s[i] = bond_modfun(bp, c);
		}
	} else {
#pragma unroll
		for (int i = 0; i < 3; i++) v[i] = bond_modfun(bp, par.halo.m[i]);
#pragma unroll
		for (int i = 0; i < 6; i++) s[i] = bond_modfun(bp, par.halo.ss[i]);
	}
	bond_trivar(v, s, rng);
	if (!disk) {
		float invR = 1.f / gf_sqrtf(Rsquared + Z * Z);
		float cosa = gf_sqrtf(Rsquared) * invR;
		float sina = Z * invR;
		float vR = cosa * v[0] - sina * v[1];
		float vZ = sina * v[0] + cosa * v[1];
		v[0] = vR;
		v[1] = v[2];
		v[2] = vZ;
	}
	return true;
}

// The draws bond_row makes for a row of component cmp.
GF_D void bond_skip(const gf_Bond2010_data &par, gf_mwc &rng, int cmp)
{
	if (bit_isset(par.comp_thin, cmp) || bit_isset(par.comp_thick, cmp)) {
		(void)rng.uniform();
	} else if (!bit_isset(par.comp_halo, cmp)) {
		return;
	}
	rng.polar_skip(3);
}

// os_Bond2010_kernel body (modules/Bond2010_gpu.cu.h:181-217)
GF_D void mod_Bond2010(const gf_Bond2010_data &par, gf_mwc &rng, uint32_t r0, uint32_t r1,
                       const gf_col &comp, const gf_col &hidden, const gf_col &XYZ, const gf_col &vcyl)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		float v[3];
		if (!bond_row(par, rng, col<int>(comp, row), col<float>(XYZ, row, 0), col<float>(XYZ, row, 1),
		              col<float>(XYZ, row, 2), v))
			continue;
		col<float>(vcyl, row, 0) = v[0];
		col<float>(vcyl, row, 1) = v[1];
		col<float>(vcyl, row, 2) = v[2];
	}
}

// transform_pm in double (modules/vel2pm_gpu.cu.h:134-164), with float
// in/out parameters narrowed where the reference narrows them.
GF_D void transform_pm(float &vlout, float &vbout, float l, float b, float vl, float vb,
                       const double ce, const double se, const double l0)
{
	double cb, sb, cl, sl;
	gfm::sincos((double)b, &sb, &cb);
	gfm::sincos((double)l - l0, &sl, &cl);
	double tmp1 = cb * se - ce * sb * sl;
	double tmp2 = ce * sb - cb * se * sl;
	double tmp4 = cb * cl;
	double tmp3 = sb * se + cb * ce * sl;
	if (fabs(tmp3) > 0.9999999) {
		vlout = vbout = 999.99f;
	} else {
		double denom1 = 1. / (tmp4 * tmp4 + tmp2 * tmp2);
		double cd = sqrt(1. - tmp3 * tmp3);
		vl = (float)((double)vl / cb);
		vlout = (float)((-ce * cl * (double)vb + cb * tmp1 * (double)vl) * denom1);
		vbout = (float)((tmp1 * (double)vb + cb * ce * cl * (double)vl) / cd);
		vlout = (float)((double)vlout * cd);
	}
}

// One row of os_vel2pm_kernel (modules/vel2pm_gpu.cu.h:268-300) with
// vcyl2pm (:221-251), vel_cyl2xyz (:112-121) and vel_xyz2lbr (:81-94), from
// the row's (l, b) in degrees, XYZ and cylindrical velocity.
GF_D void vel2pm_row(const gf_vel2pm_data &par, uint32_t row, double l_deg, double b_deg, float x, float y, float z,
                     float vr, float vphi, float vz0, const gf_col &pmout)
{
	{
		float l = 0.017453292519943295769f * (float)l_deg;
		float b = 0.017453292519943295769f * (float)b_deg;
		float X = par.Rg - x;
		float Y = -y;
		float Z = z;

		float rho = gf_sqrtf(X * X + Y * Y);
		float cphi = X / rho;
		float sphi = Y / rho;
		float px = -sphi * vphi + cphi * vr;
		float py = cphi * vphi + sphi * vr;
		float pz = vz0;
		px -= par.u0;
		py -= par.v0 + par.vLSR;
		pz -= par.w0;

		float vx = -px, vy = -py, vz = pz;
		float cl = gf_cosf(l), sl = gf_sinf(l), cb = gf_cosf(b), sb = gf_sinf(b);
		float vl = -vx * sl + vy * cl;
		float tmp = vx * cl + vy * sl;
		float vrad = cb * tmp + vz * sb;
		float vb = -sb * tmp + vz * cb;

		float dx = 8000.f - X;
		float D = gf_sqrtf(dx * dx + Y * Y + Z * Z);
		double den = 4.74 * (double)D * 1e-3;
		float pmx = (float)((double)vl / den);
		float pmy = (float)((double)vb / den);
		float pmz = vrad;

		if (par.coordsys == 0) {
		} else if (par.coordsys == 1) {
			transform_pm(pmx, pmy, l, b, pmx, pmy, galequ_c::ce, galequ_c::se, galequ_c::l0);
		} else {
			pmx = pmy = pmz = -9.99f;
		}
		col<float>(pmout, row, 0) = pmx;
		col<float>(pmout, row, 1) = pmy;
		col<float>(pmout, row, 2) = pmz;
	}
}

// os_vel2pm_kernel body over rows [r0, r1).
GF_D void mod_vel2pm(const gf_vel2pm_data &par, uint32_t r0, uint32_t r1, const gf_col &lb0,
                     const gf_col &XYZ, const gf_col &vcyl, const gf_col &pmout, const gf_col &hidden)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		vel2pm_row(par, row, col<double>(lb0, row, 0), col<double>(lb0, row, 1), col<float>(XYZ, row, 0),
		           col<float>(XYZ, row, 1), col<float>(XYZ, row, 2), col<float>(vcyl, row, 0),
		           col<float>(vcyl, row, 1), col<float>(vcyl, row, 2), pmout);
	}
}

// One row of os_gal2other_kernel with galequ (modules/gal2other_gpu.cu.h:39-95)
// from the row's (l, b) in degrees.
GF_D void gal2other_row(int coordsys, uint32_t row, double l_deg, double b_deg, const gf_col &out)
{
	using namespace galequ_c;
	double lx = l_deg * d2r;
	double ly = b_deg * d2r;
	double rx, ry;
	if (coordsys == 1) {
		gal_to_equ(lx, ly, rx, ry);
	} else {
		rx = ry = -9999.;
	}
	col<double>(out, row, 0) = rx / d2r;
	col<double>(out, row, 1) = ry / d2r;
}

// os_gal2other_kernel body over rows [r0, r1).
GF_D void mod_gal2other(int coordsys, uint32_t r0, uint32_t r1, const gf_col &hidden,
                        const gf_col &lb0, const gf_col &out)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		gal2other_row(coordsys, row, col<double>(lb0, row, 0), col<double>(lb0, row, 1), out);
	}
}

// Bilinear sample of a float4 colour table (the TEX2D lookups of
// sampleColors, modules/photometry_gpu.cu.h:56-71).
GF_D float4 bilerp4(float4 t00, float4 t10, float4 t01, float4 t11, float a, float b)
{
	float4 r;
#define GF_BL(c)                                             \
	{                                                        \
		float r0 = t00.c + a * (t10.c - t00.c);              \
		float r1 = t01.c + a * (t11.c - t01.c);              \
		r.c = r0 + b * (r1 - r0);                            \
	}
	GF_BL(x) GF_BL(y) GF_BL(z) GF_BL(w)
#undef GF_BL
	return r;
}

GF_D float4 tex2d4(const float *tab, int nx, int ny, const float tcx[2], const float tcy[2], float x, float y)
{
	int i, j;
	float a = lerp_coord(x, tcx[0], tcx[1], nx, i);
	float b = lerp_coord(y, tcy[0], tcy[1], ny, j);
	int i1 = nx > 1 ? i + 1 : i, j1 = ny > 1 ? j + 1 : j;
	const float4 *t = (const float4 *)tab;
	return bilerp4(t[(size_t)j * nx + i], t[(size_t)j * nx + i1], t[(size_t)j1 * nx + i], t[(size_t)j1 * nx + i1],
	               a, b);
}

// sampleColors (modules/photometry_gpu.cu.h:56-71).  NT: the most colour
// tables (float4 each) the caller's arrays hold; the loops are unrolled to
// it, so a small NT keeps the caller's register footprint small.
template <int NT = 4>
GF_D uint32_t sample_colors(const gf_photometry_data &lt, float *c, float FeH, float Mr)
{
	uint32_t flags = 0, cidx = 0;
	// texel coordinates, shared by every table
	int i, j;
	const float a = lerp_coord(FeH, lt.tcFeH[0], lt.tcFeH[1], lt.nFeH, i);
	const float b = lerp_coord(Mr, lt.tcMr[0], lt.tcMr[1], lt.nMr, j);
	const int ntex = ((int)lt.ncolors + 3) / 4;
	// the lookup quad's record (quad layout): the four corners' float4s per
	// table and their flag bits, one contiguous fetch each
	const float4 *tx = lt.flagbits ? (const float4 *)lt.texels : nullptr;
	const size_t k = (size_t)j * lt.nFeH + i;
	uint32_t b00 = 0, b10 = 0, b01 = 0, b11 = 0;
	if (tx) {
		const uint2 fw = *(const uint2 *)(lt.flagbits + 4 * k);
		b00 = fw.x & 0xffffu; b10 = fw.x >> 16;
		b01 = fw.y & 0xffffu; b11 = fw.y >> 16;
	}
	for (int t = 0; t < NT && cidx < lt.ncolors; t++) {
		float4 clr, f;
		if (tx) {
			const float4 *qd = tx + 4 * ((size_t)ntex * k + t);
			clr = bilerp4(qd[0], qd[1], qd[2], qd[3], a, b);
			// the flag texels, 0 or 1, filtered as the colours are
			auto unpack = [t](uint32_t bits) {
				bits >>= 4 * t;
				return make_float4((float)(bits & 1u), (float)((bits >> 1) & 1u), (float)((bits >> 2) & 1u),
				                   (float)((bits >> 3) & 1u));
			};
			f = bilerp4(unpack(b00), unpack(b10), unpack(b01), unpack(b11), a, b);
		} else {
			clr = tex2d4(lt.color[t], lt.nFeH, lt.nMr, lt.tcFeH, lt.tcMr, FeH, Mr);
			f = tex2d4(lt.cflags[t], lt.nFeH, lt.nMr, lt.tcFeH, lt.tcMr, FeH, Mr);
		}
		float cv[4] = {clr.x, clr.y, clr.z, clr.w};
		float fv[4] = {f.x, f.y, f.z, f.w};
		for (int q = 0; q < 4 && cidx < lt.ncolors; q++) {
			c[cidx] = cv[q];
			flags |= (uint32_t)(fv[q] != 0.f) << cidx;
			cidx++;
		}
	}
	return flags;
}

// One row of os_photometry_kernel (modules/photometry_gpu.cu.h:100-173) for
// a visible row of a selected component, from the row's [Fe/H], distance
// modulus, extinction and first two absolute magnitudes (fMr1 is read only
// when nabsmag > 1), for at most 4*NT colours (the caller checks ncolors <=
// 4*NT).  LONE: every row is a lone star (one absolute-magnitude column), so
// the luminosity sum over a system's components is compiled out.
template <int NT, bool LONE>
GF_D void phot_row(const gf_photometry_data &lt, uint32_t row, float fFeH, float dm, float Am0, float fMr0,
                   float fMr1, const gf_col &flags, const gf_col &Mr, int nabsmag, const gf_col &mags)
{
	constexpr int NC = 4 * NT;
	float c[NC];
	const int nb = (int)lt.ncolors < NC ? (int)lt.ncolors : NC;
	{
		uint32_t flag = 0;
		const bool system = !LONE && fMr0 < ABSMAG_NOT_PRESENT && nabsmag > 1 && fMr1 < ABSMAG_NOT_PRESENT;
		if (!system) {
			// a lone star: its magnitudes are the colour sums themselves
			// (-2.5 log10 10^(-0.4 M) = M, without the round trip)
			if (fMr0 < ABSMAG_NOT_PRESENT) flag = sample_colors<NT>(lt, c, fFeH, fMr0);
			for (int b = 0; b <= nb; b++) {
				float M = fMr0;
				if (fMr0 < ABSMAG_NOT_PRESENT) {
					if (b < (int)lt.bidx) { for (int i = b; i != (int)lt.bidx; i++) M += c[i]; }
					if (b > (int)lt.bidx) { for (int i = lt.bidx; i != b; i++) M -= c[i]; }
				} else {
					M = gfm::as_float(0x7f800000u);   // no component: zero luminosity
				}
				col<float>(mags, row, b) = (dm + M) + Am0 * lt.reddening[b];
			}
		} else {
			// a system: luminosities summed over its components, newest first
			float L[NC + 1];
			for (int b = 0; b <= nb; b++) L[b] = 0.f;
			for (int syscomp = 0; syscomp < nabsmag; syscomp++) {
				float fMr = col<float>(Mr, row, syscomp);
				if (fMr >= ABSMAG_NOT_PRESENT) break;
				flag |= sample_colors<NT>(lt, c, fFeH, fMr);
				for (int b = 0; b <= nb; b++) {
					float M = fMr;
					if (b < (int)lt.bidx) { for (int i = b; i != (int)lt.bidx; i++) M += c[i]; }
					if (b > (int)lt.bidx) { for (int i = lt.bidx; i != b; i++) M -= c[i]; }
					float Lb = gf_exp10f(-0.4f * M);
					if (syscomp) Lb += L[b];
					L[b] = Lb;
				}
			}
			for (int b = 0; b <= nb; b++)
				col<float>(mags, row, b) = (dm + -2.5f * gf_log10f(L[b])) + Am0 * lt.reddening[b];
		}
		bool fl = false;
		for (int i = (int)lt.bidx - 1; i >= 0; i--) {
			uint32_t mask = 1u << i;
			fl = fl || (flag & mask);
			flag |= (uint32_t)fl << i;
		}
		fl = false;
		for (int i = lt.bidx; i < nb; i++) {
			uint32_t mask = 1u << i;
			fl = fl || (flag & mask);
			flag |= (uint32_t)fl << i << 1;
		}
		flag &= ~(1u << lt.bidx);
		flag = __brev(flag);
		flag >>= (32 - nb - 1);
		col<int>(flags, row) = (int)flag;
	}
}

// os_photometry_kernel body (modules/photometry_gpu.cu.h:100-173) over rows
// [r0, r1), for at most 4*NT colours.
template <int NT = 4, bool LONE = false>
GF_D void mod_photometry(const gf_photometry_data &lt, gf_bitmap apply, uint32_t r0, uint32_t r1,
                         const gf_col &Am, const gf_col &flags, const gf_col &DM, const gf_col &Mr,
                         int nabsmag, const gf_col &mags, const gf_col &FeH, const gf_col &comp,
                         const gf_col &hidden)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		if (!bit_isset(apply, col<int>(comp, row))) continue;
		const float fMr1 = !LONE && nabsmag > 1 ? col<float>(Mr, row, 1) : ABSMAG_NOT_PRESENT;
		phot_row<NT, LONE>(lt, row, col<float>(FeH, row), col<float>(DM, row), col<float>(Am, row),
		                   col<float>(Mr, row, 0), fMr1, flags, Mr, nabsmag, mags);
	}
}

// The companion draws of one row of os_unresolvedMultiples_kernel
// (draw_companion, modules/unresolvedMultiples_gpu.cu.h:48-78, called from
// :104-117) whose primary has magnitude M1: put(i, M, has) receives the
// magnitude of companion i in [1, nabsmag), ABSMAG_NOT_PRESENT when it has
// none.
template <typename Put>
GF_D void multiples_draw(const gf_multiples_data &par, gf_mwc &rng, float M1, int nabsmag, Put put)
{
	for (int i = 1; i < nabsmag; i++) {
		float M2 = 0.f;
		bool has = false;
		float psec = tex1d(par.secProb, M1);
		float u = rng.uniform();
		if (!(u > psec)) {
			has = true;
			if (par.algo == 3) {
				M2 = M1;
			} else {
				float pprim = tex1d(par.cumLF, M1);
				u = rng.uniform();
				if (par.algo == 1) u = (float)((double)pprim + (double)u * (1. - (double)pprim));
				M2 = tex1d(par.invCumLF, u);
				if (par.algo == 1 && M2 < M1) M2 = M1;
			}
		}
		put(i, has ? M2 : ABSMAG_NOT_PRESENT, has);
	}
}

// The system's luminosity sum (modules/unresolvedMultiples_gpu.cu.h:104-118):
// each present component adds 10^(-0.4 M) in component order, and the total
// magnitude is -2.5 log10 of the sum.
GF_D float multiples_lum(float M) { return gf_exp10f(-0.4f * M); }
GF_D float multiples_mag(float Ltot) { return (float)(-2.5 * (double)gf_log10f(Ltot)); }

// os_unresolvedMultiples_kernel (modules/unresolvedMultiples_gpu.cu.h:80-123)
// over rows [r0, r1)
GF_D void mod_multiples(const gf_multiples_data &par, gf_bitmap apply, gf_mwc &rng, uint32_t r0, uint32_t r1,
                        int nabsmag, const gf_col &M, const gf_col &Msys, const gf_col &ncomp,
                        const gf_col &comp, const gf_col &hidden)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		int cmp = col<int>(comp, row);
		if (!bit_isset(apply, cmp)) continue;
		float M1 = col<float>(M, row);
		col<float>(Msys, row, 0) = M1;
		float Ltot = multiples_lum(M1);
		int n = 1;
		multiples_draw(par, rng, M1, nabsmag, [&](int i, float v, bool has) {
			col<float>(Msys, row, i) = v;
			if (has) {
				Ltot += multiples_lum(v);
				n++;
			}
		});
		col<float>(M, row) = multiples_mag(Ltot);
		col<int>(ncomp, row) = n;
	}
}

// Linear spline with linear extrapolation past the end knots (the GSL
// linear interpolation behind spline::operator(), common/spline.cpp:59-65).
// The interval is the largest lo in [0, n-2] with kx[lo] <= x (0 below the
// first knot), as a binary search finds it; it is guessed from the mean knot
// spacing first and the search runs only when the guess is not that interval.
GF_D double spline_eval(const double *kx, const double *ky, int n, double x)
{
	if (n == 1) return ky[0];
	int lo;
	double x0 = kx[0], xn = kx[n - 1];
	double g = (x - x0) * ((double)(n - 1) / (xn - x0));
	int gi = g < 0. ? 0 : (g > (double)(n - 2) ? n - 2 : (int)g);
	bool ok = (gi == 0 || kx[gi] <= x) && (gi == n - 2 || kx[gi + 1] > x);
	if (ok) {
		lo = gi;
	} else {
		lo = 0;
		int hi = n - 1;
		while (hi > lo + 1) {
			int mid = (lo + hi) >> 1;
			if (kx[mid] > x) hi = mid; else lo = mid;
		}
	}
	double dx = kx[lo + 1] - kx[lo];
	double D = (ky[lo + 1] - ky[lo]) / dx;
	return ky[lo] + D * (x - kx[lo]);
}

// spline_eval over the interval records q[0 .. n-1] of one band (the layout
// of gf_photoerr_data::kq) when the guessed interval gi was wrong: the
// interval by binary search, with the same arithmetic.
GF_D double spline_eval_rec_search(const double4 *q, int n, double x)
{
	int lo = 0, hi = n - 1;
	while (hi > lo + 1) {
		int mid = (lo + hi) >> 1;
		if (q[mid].x > x) hi = mid; else lo = mid;
	}
	const double4 r = q[lo];
	return r.y + r.z * (x - r.x);
}

// The guessed interval of x among n knots from the band's first knot x0
// and scale (n-1)/(x_last-x0), clamped to [0, n-2].
GF_D int spline_guess(int n, double x, double x0, double scale)
{
	const double g = (x - x0) * scale;
	return g < 0. ? 0 : (g > (double)(n - 2) ? n - 2 : (int)g);
}

// Whether record r of interval gi (of n knots) holds x: spline_eval's interval.
GF_D bool spline_holds(const double4 &r, int gi, int n, double x)
{
	return (gi == 0 || r.x <= x) && (gi == n - 2 || r.w > x);
}

// Bands whose error lookups are all issued before the draws (interval
// records present): perr_row_fast.
constexpr int PERR_FAST_BANDS = 6;

GF_D bool perr_fast(const gf_photoerr_data &par) { return par.kq && par.nbands <= PERR_FAST_BANDS; }

// Photometric errors of one visible row from its true magnitudes mag[0 ..
// nbands-1] (perr_fast(par)): every band's guessed interval record (one
// 32-byte fetch each, all in flight together), then the values (a wrong
// guess, from irregular knots, falls back to the binary search), then the
// deviates in band order.
GF_D void perr_row_fast(const gf_photoerr_data &par, gf_mwc &rng, uint32_t row, const float mag[PERR_FAST_BANDS],
                        const gf_col &magObs)
{
	constexpr int NB = PERR_FAST_BANDS;
	const double4 *q = (const double4 *)par.kq;
	float sigma[NB];
	double4 r[NB];
	int gi[NB];
#pragma unroll
	for (int b = 0; b < NB; b++) {
		const int n = b < par.nbands ? par.koff[b + 1] - par.koff[b] : 0;
		gi[b] = n > 1 ? spline_guess(n, (double)mag[b], par.gx0[b], par.gscale[b]) : 0;
		if (n > 0) r[b] = q[par.koff[b] + gi[b]];
	}
#pragma unroll
	for (int b = 0; b < NB; b++) {
		const int n = b < par.nbands ? par.koff[b + 1] - par.koff[b] : 0;
		const double x = (double)mag[b];
		double v = n == 1 ? r[b].y : r[b].y + r[b].z * (x - r[b].x);
		if (n > 1 && !spline_holds(r[b], gi[b], n, x)) v = spline_eval_rec_search(q + par.koff[b], n, x);
		sigma[b] = (float)v;
	}
#pragma unroll
	for (int b = 0; b < NB; b++) {
		if (b >= par.nbands) break;
		const bool curve = par.koff[b + 1] > par.koff[b];
		col<float>(magObs, row, b) = curve ? mag[b] + rng.gaussian(sigma[b]) : mag[b];
	}
}

// Photometric errors (modules/photometricErrors_host.cpp:112-131) drawn
// from the chunk's MWC stream, rows in order and the bands of a row in
// order.  One visible row: a deviate for each band that has an error curve.
GF_D void perr_row(const gf_photoerr_data &par, gf_mwc &rng, uint32_t row, const gf_col &magTrue,
                   const gf_col &magObs)
{
	if (perr_fast(par)) {
		float mag[PERR_FAST_BANDS];
#pragma unroll
		for (int b = 0; b < PERR_FAST_BANDS; b++)
			if (b < par.nbands) mag[b] = col<float>(magTrue, row, b);
		perr_row_fast(par, rng, row, mag, magObs);
		return;
	}
	for (int b = 0; b < par.nbands; b++) {
		int k0 = par.koff[b], k1 = par.koff[b + 1];
		float mag = col<float>(magTrue, row, b);
		if (k1 > k0) {
			float sigma = (float)spline_eval(par.kx + k0, par.ky + k0, k1 - k0, (double)mag);
			col<float>(magObs, row, b) = mag + rng.gaussian(sigma);
		} else {
			col<float>(magObs, row, b) = mag;
		}
	}
}

// ncurves: the number of bands with an error curve
GF_D void perr_skip(int ncurves, gf_mwc &rng)
{
	rng.polar_skip(ncurves);
}

GF_D int perr_ncurves(const gf_photoerr_data &par)
{
	int n = 0;
	for (int b = 0; b < par.nbands; b++) n += par.koff[b + 1] > par.koff[b] ? 1 : 0;
	return n;
}

GF_D void mod_photoerr(const gf_photoerr_data &par, gf_mwc &rng, uint32_t r0, uint32_t r1,
                       const gf_col &magTrue, const gf_col &magObs, const gf_col &hidden)
{
	for (uint32_t row = r0; row < r1; row++) {
		if (col<int>(hidden, row)) continue;
		perr_row(par, rng, row, magTrue, magObs);
	}
}
