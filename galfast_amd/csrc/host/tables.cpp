// Configuration parsing and host-side table construction.
#include "host.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <map>
#include <stdlib.h>
#include <sstream>

namespace gfh {

void fail(const std::string &msg) { throw ConfigError{msg}; }

static std::string trim(const std::string &s)
{
	size_t a = s.find_first_not_of(" \t\r\n");
	if (a == std::string::npos) return "";
	size_t b = s.find_last_not_of(" \t\r\n");
	return s.substr(a, b - a + 1);
}

void Config::parse(const std::string &text)
{
	std::istringstream in(text);
	std::string line;
	while (std::getline(in, line)) {
		size_t h = line.find('#');
		if (h != std::string::npos) line = line.substr(0, h);
		size_t eq = line.find('=');
		if (eq == std::string::npos) continue;
		std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
		if (!k.empty()) kv[k] = v;
	}
}

bool Config::load(const std::string &file)
{
	std::ifstream f(file);
	if (!f) return false;
	std::stringstream ss;
	ss << f.rdbuf();
	path = file;
	parse(ss.str());
	return true;
}

bool Config::count(const std::string &key) const
{
	return kv.count(key) || (globals && globals->count(key));
}

std::string Config::get(const std::string &key) const
{
	auto it = kv.find(key);
	if (it != kv.end()) return it->second;
	if (globals && globals->count(key)) return globals->get(key);
	fail("configuration key '" + key + "' missing" + (path.empty() ? "" : " in " + path));
}

std::string Config::get(const std::string &key, const std::string &dflt) const
{
	return count(key) ? get(key) : dflt;
}

double Config::getd(const std::string &key) const
{
	std::string v = get(key);
	char *end = nullptr;
	double d = strtod(v.c_str(), &end);
	if (end == v.c_str()) fail("configuration key '" + key + "' is not a number: " + v);
	return d;
}

double Config::getd(const std::string &key, double dflt) const { return count(key) ? getd(key) : dflt; }

std::vector<double> Config::getv(const std::string &key, const std::string &dflt) const
{
	std::istringstream ss(get(key, dflt));
	std::vector<double> out;
	double d;
	while (ss >> d) out.push_back(d);
	return out;
}

std::string Config::resolve(const std::string &rel) const
{
	if (rel.empty() || rel[0] == '/' || path.empty()) return rel;
	size_t s = path.find_last_of('/');
	if (s == std::string::npos) return rel;
	return path.substr(0, s + 1) + rel;
}

std::vector<std::vector<double>> load_columns(const std::string &file, int ncols)
{
	std::ifstream f(file);
	if (!f) fail("cannot open " + file);
	std::vector<std::vector<double>> rows;
	std::string line;
	while (std::getline(f, line)) {
		size_t h = line.find('#');
		if (h != std::string::npos) line = line.substr(0, h);
		std::istringstream ss(line);
		std::vector<double> r;
		double d;
		while ((int)r.size() < ncols && ss >> d) r.push_back(d);
		if (r.empty()) continue;
		if ((int)r.size() < ncols) fail("short row in " + file);
		rows.push_back(r);
	}
	return rows;
}

double Spline::operator()(double v) const
{
	int n = (int)x.size();
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

Tex1D resample_1d(const std::vector<double> &X, const std::vector<double> &Y, int nsamp)
{
	Spline s(X, Y);
	Tex1D t;
	t.v.resize(nsamp);
	double X0 = X.front(), X1 = X.back(), dX = (X1 - X0) / (nsamp - 1);
	for (int i = 0; i < nsamp; i++) t.v[i] = (float)s(X0 + i * dX);
	t.tc[0] = (float)X0;
	t.tc[1] = (float)(1. / dX);
	return t;
}

Tex1D resample_clip_1d(std::vector<double> X, std::vector<double> Y, float clip, int nsamp)
{
	X.insert(X.begin(), -123.0);
	Y.insert(Y.begin(), (double)clip);
	double xrange = X.back() - X[1];
	X[0] = X[1] - 1e-5 * xrange;
	X.push_back(X.back() + 1e-5 * xrange);
	Y.push_back((double)clip);
	return resample_1d(X, Y, nsamp);
}

Tex1D constant_1d(float val, float X0, float X1)
{
	Tex1D t;
	t.v = {val, val};
	t.tc[0] = X0;
	t.tc[1] = (float)(1. / ((double)X1 - (double)X0));
	return t;
}

PhotTables build_photometry_tables(const std::vector<std::vector<double>> &rows, int ncolors, double dFeH,
                                   double dMr)
{
	// group rows by FeH knot; sort each group by Mr
	std::map<double, std::vector<std::vector<double>>> byFeH;
	for (const auto &r : rows) byFeH[r[1]].push_back(r);
	if (byFeH.empty()) fail("empty photometric system table");
	double FeH0 = byFeH.begin()->first, FeH1 = byFeH.rbegin()->first;
	double Mr0 = rows[0][0], Mr1 = rows[0][0];
	for (const auto &r : rows) { Mr0 = std::min(Mr0, r[0]); Mr1 = std::max(Mr1, r[0]); }
	PhotTables pt;
	pt.ncolors = ncolors;
	pt.nFeH = (int)((FeH1 - FeH0) / dFeH + 1);
	pt.nMr = (int)((Mr1 - Mr0) / dMr + 1);
	pt.ntex = (ncolors + 3) / 4;
	size_t per = (size_t)pt.nFeH * pt.nMr * 4;
	pt.color.assign(per * pt.ntex, 0.f);
	pt.flags.assign(per * pt.ntex, 0.f);

	std::vector<double> keys;
	std::vector<double> mrmin, mrmax;
	std::vector<std::vector<Spline>> spl;   // [knot][colour] col(Mr)
	for (auto &kvp : byFeH) {
		auto &g = kvp.second;
		std::stable_sort(g.begin(), g.end(), [](const std::vector<double> &a, const std::vector<double> &b) { return a[0] < b[0]; });
		keys.push_back(kvp.first);
		std::vector<double> mr;
		for (auto &r : g) mr.push_back(r[0]);
		std::vector<Spline> s(ncolors);
		for (int c = 0; c < ncolors; c++) {
			std::vector<double> cv;
			for (auto &r : g) cv.push_back(r[2 + c]);
			s[c] = Spline(mr, cv);
		}
		spl.push_back(std::move(s));
		mrmin.push_back(mr.front());
		mrmax.push_back(mr.back());
	}
	const int nk = (int)keys.size();
	std::vector<double> vcol(nk), vecol(nk);
	for (int ic = 0; ic < ncolors; ic++) {
		int tex = ic / 4, comp = ic % 4;
		for (int m = 0; m < pt.nMr; m++) {
			double Mr = Mr0 + m * dMr;
			for (int k = 0; k < nk; k++) {
				vcol[k] = spl[k][ic](Mr);
				vecol[k] = (mrmin[k] > Mr || Mr > mrmax[k]) ? 1.0 : 0.0;
			}
			Spline fc(keys, vcol), fe(keys, vecol);
			for (int f = 0; f < pt.nFeH; f++) {
				double FeH = FeH0 + f * dFeH;
				float color = (float)fc(FeH);
				bool ef = keys.front() > FeH || FeH > keys.back() || fe(FeH) != 0.;
				size_t o = (((size_t)tex * pt.nMr + m) * pt.nFeH + f) * 4 + comp;
				pt.color[o] = color;
				pt.flags[o] = ef ? 1.f : 0.f;
			}
		}
	}
	// texcoord_from_range(0, n, x0, x0 + n*d), evaluated in float (common/cux.cpp:389-397)
	auto tcr = [](float imgx, float imgy, float x, float y, float *o) {
		o[0] = (-imgy * x + imgx * y) / (imgx - imgy);
		o[1] = (imgx - imgy) / (x - y);
	};
	tcr(0.f, (float)pt.nFeH, (float)FeH0, (float)(FeH0 + pt.nFeH * dFeH), pt.tcFeH);
	tcr(0.f, (float)pt.nMr, (float)Mr0, (float)(Mr0 + pt.nMr * dMr), pt.tcMr);
	return pt;
}

// J2000 Galactic <-> equatorial (the constants of galequ_constants,
// modules/module_lib.h:62-71): (l, b) -> (ra, dec) as galequ
// (modules/gal2other_gpu.cu.h:39-60) and its inverse rotation, radians.
namespace {
const double GE_ANGP = M_PI / 180. * 192.859508333, GE_L0 = M_PI / 180. * 32.932;
const double GE_CE = 0.88998740217659689, GE_SE = 0.45598511375586859;
}

void gal_to_equ_host(double l, double b, double &ra, double &dec)
{
	const double cb = cos(b), sb = sin(b), cl = cos(l - GE_L0), sl = sin(l - GE_L0);
	ra = atan2(cb * cl, sb * GE_CE - cb * GE_SE * sl) + GE_ANGP;
	dec = asin(cb * GE_CE * sl + sb * GE_SE);
	while (ra < 0.) ra += 2 * M_PI;
	while (ra >= 2 * M_PI) ra -= 2 * M_PI;
}

void equ_to_gal(double ra, double dec, double &l, double &b)
{
	double A = ra - GE_ANGP;
	double X = cos(dec) * cos(A), Y = cos(dec) * sin(A), Z = sin(dec);
	double u0 = Y, u1 = -GE_SE * X + GE_CE * Z, u2 = GE_CE * X + GE_SE * Z;
	b = asin(std::max(-1.0, std::min(1.0, u2)));
	l = atan2(u1, u0) + GE_L0;
	l = fmod(l, 2 * M_PI);
	if (l < 0) l += 2 * M_PI;
}

namespace {

// (l, b, D) -> Galactic cartesian (lbd2xyz, common/transform.cpp:58-68)
void lbd2xyz(double v[3])
{
	const double cl = cos(v[0]), cb = cos(v[1]), sl = sin(v[0]), sb = sin(v[1]), d = v[2];
	v[0] = d * cl * cb;
	v[1] = d * sl * cb;
	v[2] = d * sb;
}

// get_gal2xxx_matrix (common/transform.cpp:70-148): Galactic cartesian to
// coordinate system cs.
void gal2xxx(double go[3][3], const std::string &cs, const Config &cfg)
{
	if (cs == "gal" || cs == "galcentric") {
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++) go[i][j] = i == j ? 1. : 0.;
	} else if (cs == "equ") {
		// the Galactic axes in equatorial cartesian form the columns
		double x[3], y[3], z[3];
		gal_to_equ_host(0., 0., x[0], x[1]); x[2] = 1;
		gal_to_equ_host(M_PI / 2, 0., y[0], y[1]); y[2] = 1;
		gal_to_equ_host(0., M_PI / 2, z[0], z[1]); z[2] = 1;
		lbd2xyz(x); lbd2xyz(y); lbd2xyz(z);
		for (int i = 0; i < 3; i++) { go[i][0] = x[i]; go[i][1] = y[i]; go[i][2] = z[i]; }
	} else if (cs == "galplane") {
		double Rg = cfg.getd("Rg"), z0 = cfg.getd("z0");
		double sina = -z0 / Rg, cosa = sqrt(1 - sina * sina);
		double G[3][3] = {{-cosa, 0, sina}, {0, -1, 0}, {sina, 0, cosa}};
		memcpy(go, G, sizeof(G));
	} else {
		fail("Unknown coordinate system " + cs + ".");
	}
}

// xxx2galxyz (common/transform.cpp:150-224): a point given in coordinate
// system cs (spherical: degrees; an empty type is the origin) to Galactic
// cartesian.
void xxx2galxyz(const Config &cfg, const std::string &cs, const std::string &type, double v[3])
{
	bool unkt = false;
	if (type.empty()) v[0] = v[1] = v[2] = 0.;
	if (cs == "gal") {
		if (type == "sph") {
			v[0] *= M_PI / 180.; v[1] *= M_PI / 180.;
			lbd2xyz(v);
		} else if (!(type == "xyz" || type.empty())) unkt = true;
	} else if (cs == "equ") {
		if (type == "sph" || type.empty()) {
			v[0] *= M_PI / 180.; v[1] *= M_PI / 180.;
			equ_to_gal(v[0], v[1], v[0], v[1]);
			lbd2xyz(v);
		} else unkt = true;
	} else if (cs == "galcentric") {
		if (type == "xyz" || type.empty()) {
			double Rg = cfg.getd("Rg");
			v[0] = Rg - v[0];
			v[1] = -v[1];
		} else unkt = true;
	} else if (cs == "galplane") {
		if (type == "xyz" || type.empty()) {
			double Rg = cfg.getd("Rg"), z0 = cfg.getd("z0");
			double sina = -z0 / Rg, cosa = sqrt(1 - sina * sina);
			double x = cosa * v[0] - sina * v[2];
			double z = sina * v[0] + cosa * v[2];
			v[0] = Rg - x;
			v[1] = -v[1];
			v[2] = z;
		} else unkt = true;
	} else {
		fail("Unknown coordinate system " + cs + ".");
	}
	if (unkt) fail("Unknown rotation type " + type + ".");
}

// get_rotation_matrix (common/transform.cpp:226-343): '<coordsys> [<kind>
// params]' with kind axisangle (x y z angle[deg]) or euler (phi theta psi
// [deg]), composed as rot * gal2xxx.
void rotation_matrix(double M[3][3], const std::string &spec, const Config &cfg)
{
	std::istringstream in(spec);
	std::string cs;
	if (!(in >> cs)) {
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++) M[i][j] = i == j ? 1. : 0.;
		return;
	}
	double R[3][3];
	gal2xxx(R, cs, cfg);
	std::string kind;
	if (!(in >> kind)) {
		memcpy(M, R, sizeof(R));
		return;
	}
	double rot[3][3];
	if (kind == "axisangle") {
		double x, y, z, alpha;
		if (!(in >> x >> y >> z >> alpha)) fail("Four numbers (x, y, z, angle) required for 'axisangle' rotation.");
		double norm = sqrt(x * x + y * y + z * z);
		x /= norm; y /= norm; z /= norm;
		alpha = -alpha * M_PI / 180.;
		double c = cos(alpha), sn = sin(alpha), Cc = 1. - c;
		double xs = x * sn, ys = y * sn, zs = z * sn, xC = x * Cc, yC = y * Cc, zC = z * Cc;
		double xyC = x * yC, yzC = y * zC, zxC = z * xC;
		rot[0][0] = x * xC + c; rot[0][1] = xyC - zs;   rot[0][2] = zxC + ys;
		rot[1][0] = xyC + zs;   rot[1][1] = y * yC + c; rot[1][2] = yzC - xs;
		rot[2][0] = zxC - ys;   rot[2][1] = yzC + xs;   rot[2][2] = z * zC + c;
	} else if (kind == "euler") {
		double phi, theta, psi;
		if (!(in >> phi >> theta >> psi)) fail("Three angles (phi, theta, psi) required for 'euler' rotation.");
		phi *= M_PI / 180.; theta *= M_PI / 180.; psi *= M_PI / 180.;
		rot[0][0] = cos(psi) * cos(phi) - cos(theta) * sin(phi) * sin(psi);
		rot[0][1] = cos(psi) * sin(phi) + cos(theta) * cos(phi) * sin(psi);
		rot[0][2] = sin(psi) * sin(theta);
		rot[1][0] = -sin(psi) * cos(phi) - cos(theta) * sin(phi) * cos(psi);
		rot[1][1] = -sin(psi) * sin(phi) + cos(theta) * cos(phi) * cos(psi);
		rot[1][2] = cos(psi) * sin(theta);
		rot[2][0] = sin(theta) * sin(phi);
		rot[2][1] = -sin(theta) * cos(phi);
		rot[2][2] = cos(theta);
	} else {
		fail("Unknown rotation type '" + kind + "'");
	}
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) {
			double s = 0;
			for (int k = 0; k < 3; k++) s += rot[i][k] * R[k][j];
			M[i][j] = s;
		}
}

} // namespace

// load_transform (common/transform.cpp:345-388): xyz' = M (xyz + T) for a
// centre '<coordsys> <type> v0 v1 v2' and an orientation as above.
void load_transform(float T[3], float M[3][3], const std::string &center, const std::string &orientation,
                    const Config &cfg)
{
	double ctr[3] = {0, 0, 0};
	std::string cs, type;
	std::istringstream ss(center);
	ss >> cs >> type >> ctr[0] >> ctr[1] >> ctr[2];
	xxx2galxyz(cfg, cs, type, ctr);
	for (int i = 0; i < 3; i++) T[i] = (float)-ctr[i];
	double R[3][3];
	rotation_matrix(R, orientation, cfg);
	for (int i = 0; i < 3; i++)
		for (int j = 0; j < 3; j++) M[i][j] = (float)R[i][j];
}

void texcoord_from_range(float imgx, float imgy, float x, float y, float tc[2])
{
	tc[0] = (-imgy * x + imgx * y) / (imgx - imgy);
	tc[1] = (imgx - imgy) / (x - y);
}

namespace {

// FITS header card value: the text after "= " up to a '/' comment
bool card_value(const std::string &card, const std::string &key, std::string &val)
{
	if (card.compare(0, key.size(), key) != 0) return false;
	size_t k = key.size();
	while (k < 8 && card[k] == ' ') k++;
	if (k != 8 || card.size() < 10 || card[8] != '=' ) return false;
	std::string v = card.substr(10);
	size_t q = v.find('/');
	if (q != std::string::npos && v.find('\'') == std::string::npos) v = v.substr(0, q);
	size_t a = v.find_first_not_of(' '), b = v.find_last_not_of(' ');
	val = a == std::string::npos ? std::string() : v.substr(a, b - a + 1);
	return true;
}

} // namespace

Cube load_fits_cube(const std::string &file)
{
	std::ifstream in(file, std::ios::binary);
	if (!in) fail("cannot open extinction map " + file);
	std::map<std::string, std::string> kv;
	bool end = false;
	char block[2880];
	while (!end) {
		in.read(block, 2880);
		if (in.gcount() != 2880) fail("truncated FITS header in " + file);
		for (int i = 0; i < 36 && !end; i++) {
			std::string card(block + 80 * i, 80);
			if (card.compare(0, 3, "END") == 0 && card.find_first_not_of(' ', 3) == std::string::npos) {
				end = true;
				break;
			}
			std::string key = card.substr(0, 8);
			key = key.substr(0, key.find_last_not_of(' ') + 1);
			std::string val;
			if (!key.empty() && card_value(card, key, val)) kv[key] = val;
		}
	}
	auto num = [&](const std::string &k, bool &ok) {
		auto it = kv.find(k);
		ok = it != kv.end();
		return ok ? strtod(it->second.c_str(), nullptr) : 0.0;
	};
	bool ok;
	if ((int)num("NAXIS", ok) != 3 || !ok)
		fail("Supplied extinction map file " + file + " is not 3-dimensional.");
	if ((int)num("BITPIX", ok) != -32 || !ok)
		fail("Supplied extinction map file " + file + " is not a data cube of 32-bit floats");
	Cube c;
	c.nx = (int)num("NAXIS1", ok);
	c.ny = (int)num("NAXIS2", ok);
	c.nz = (int)num("NAXIS3", ok);
	if (c.nx < 1 || c.ny < 1 || c.nz < 1) fail("bad extinction map dimensions in " + file);
	size_t n = (size_t)c.nx * c.ny * c.nz;
	std::vector<unsigned char> raw(n * 4);
	in.read((char *)raw.data(), (std::streamsize)raw.size());
	if ((size_t)in.gcount() != raw.size()) fail("truncated FITS data in " + file);
	c.v.resize(n);
	for (size_t i = 0; i < n; i++) {
		uint32_t u = (uint32_t)raw[4 * i] << 24 | (uint32_t)raw[4 * i + 1] << 16 | (uint32_t)raw[4 * i + 2] << 8 |
		             (uint32_t)raw[4 * i + 3];
		memcpy(&c.v[i], &u, 4);
	}
	// texture coordinates from the WCS keys (all three axes) or the defaults
	bool have = kv.count("CRPIX1") != 0;
	const int dims[3] = {c.nx, c.ny, c.nz};
	for (int a = 0; a < 3; a++) {
		std::string s = std::to_string(a + 1);
		if (kv.count("CROTA" + s)) fail("Axis rotations not supported (CROTA" + s + " keyword present in " + file + ")");
		if (have) {
			bool o1, o2, o3;
			float imgx = (float)(num("CRPIX" + s, o1) - 1);
			float x = (float)num("CRVAL" + s, o2), dx = (float)num("CDELT" + s, o3);
			if (!(o1 && o2 && o3)) fail("incomplete coordinate keys for axis " + s + " in " + file);
			c.tc[a][0] = x - imgx * dx;
			c.tc[a][1] = 1.f / dx;
		} else if (a < 2) {
			texcoord_from_range(0, (float)(dims[a] - 1), -2, 2, c.tc[a]);
		} else {
			texcoord_from_range(0, (float)(dims[a] - 1), 10, 15, c.tc[a]);
		}
	}
	bool o1, o2;
	double l0 = num("LAMBDA0", o1), phi1 = num("PHI1", o2);
	if (o1 && o2) { c.l0 = l0; c.phi1 = phi1; }
	return c;
}

} // namespace gfh
