// Host-side layer: configuration files, table construction and the catalog
// pipeline behind the C ABI.  Mirrors the reference's host modules
// (pipeline.cpp, skygen/os_skygen.cpp, modules/*_host.cpp).
#pragma once

#include <map>
#include <strings.h>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/galfast_hip.h"

namespace gfh {

// key = value configuration file (the libpeyton Config format the reference
// reads everywhere): '#' starts a comment, keys are trimmed, the last
// occurrence of a key wins; lookups fall back to `globals` (definitions).
struct Config {
	std::map<std::string, std::string> kv;
	const Config *globals = nullptr;
	std::string path;   // file the config came from (for relative paths)

	bool load(const std::string &file);
	void parse(const std::string &text);
	bool count(const std::string &key) const;
	std::string get(const std::string &key) const;   // throws if missing
	std::string get(const std::string &key, const std::string &dflt) const;
	double getd(const std::string &key, double dflt) const;
	double getd(const std::string &key) const;
	std::vector<double> getv(const std::string &key, const std::string &dflt) const;
	std::string resolve(const std::string &relpath) const;   // relative to the config's directory
};

struct ConfigError {
	std::string msg;
};

[[noreturn]] void fail(const std::string &msg);

// Whitespace-separated numeric columns of a text table ('#' comments).
std::vector<std::vector<double>> load_columns(const std::string &file, int ncols);

// Linear spline with linear extrapolation (common/spline.cpp:59-65).
struct Spline {
	std::vector<double> x, y;
	Spline() {}
	Spline(std::vector<double> xx, std::vector<double> yy) : x(std::move(xx)), y(std::move(yy)) {}
	double operator()(double v) const;
};

// A 1D texture on the host: texels plus (X0, 1/dX).
struct Tex1D {
	std::vector<float> v;
	float tc[2] = {0, 1};
};

// construct_texture_by_resampling_1D (common/cux.cpp:336-355)
Tex1D resample_1d(const std::vector<double> &X, const std::vector<double> &Y, int nsamp);
// load_resample_and_clip_texture_1D (common/cux.cpp:368-384)
Tex1D resample_clip_1d(std::vector<double> X, std::vector<double> Y, float clip, int nsamp);
// load_constant_texture_1D (common/cux.cpp:310-318)
Tex1D constant_1d(float val, float X0, float X1);

// Photometric system colour tables (modules/photometry_host.cpp:175-349).
struct PhotTables {
	int ncolors = 0, ntex = 0, nFeH = 0, nMr = 0;
	float tcFeH[2], tcMr[2];
	std::vector<float> color, flags;   // [tex][nMr][nFeH][4]
};
PhotTables build_photometry_tables(const std::vector<std::vector<double>> &rows, int ncolors, double dFeH,
                                   double dMr);

// A 3D float cube (x fastest) with the texture coordinates of each axis
// (xi = (x - tc[0]) * tc[1] + 0.5 at texel centres).
struct Cube {
	int nx = 0, ny = 0, nz = 0;
	std::vector<float> v;
	float tc[3][2] = {{0, 0}, {0, 0}, {0, 0}};
	double l0 = -100., phi1 = -100.;   // LAMBDA0 / PHI1 (degrees) when present
	float &at(int x, int y, int z) { return v[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)]; }
	float at(int x, int y, int z) const { return v[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)]; }
};
// load_extinction_map (skygen/os_skygen.cpp:518-568): a FITS primary image
// of 32-bit floats, NAXIS = 3, axes X, Y (Lambert plane) and DM, with
// CRPIXn / CRVALn / CDELTn coordinates (texcoord_from_wcs, :465-488) or the
// defaults X, Y in [-2, 2] and DM in [10, 15].
Cube load_fits_cube(const std::string &file);
// texcoord_from_range (common/cux.cpp:389-397)
void texcoord_from_range(float imgx, float imgy, float x, float y, float tc[2]);

// Coordinate transform of a model or module (common/transform.cpp:345-388).
// J2000 Galactic <-> equatorial, radians (galequ, modules/gal2other_gpu.cu.h:39-60,
// and its inverse rotation).
void gal_to_equ_host(double l, double b, double &ra, double &dec);
void equ_to_gal(double ra, double dec, double &l, double &b);
void load_transform(float T[3], float M[3][3], const std::string &center, const std::string &orientation,
                    const Config &cfg);

// One column of the output stages: its name, the catalog table column, its
// FITS form (width, 'E' / 'D' / 'J'), its text format and its otable
// definition (the text header and the FITS METADATA extension).
struct OutColumn {
	std::string name;
	gf_col gf_table::*member = nullptr;
	uint32_t width = 1;
	char type = 'E';
	std::string fmt, def;
	bool map_comp = false;   // write the user's component id
};
void output_schema(const gf_catalog &c, std::vector<OutColumn> &cols);
const std::vector<uint32_t> &component_ids(const gf_catalog &c);
// The job's countsMap output stage, or NULL.
const Config *countsmap_config(const gf_catalog &c);
// The message gf_last_error returns on this thread.
void set_last_error(const std::string &msg);

} // namespace gfh
