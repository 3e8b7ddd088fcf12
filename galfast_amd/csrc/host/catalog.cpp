// The catalog pipeline behind gf_catalog_*: reads a galfast `catalog`
// command configuration (pipeline.cpp:1313-1445), constructs the sky
// generator (skygen/os_skygen.cpp:816-860) and the per-star modules
// (modules/*_host.cpp construct()), keeps every table resident on the
// device, and runs count -> draw -> module chain in output chunks.
#include "host.h"
#include "../gf_runtime.h"
#include "../gf_chain_types.h"
#include "../gf_rng.h"
#include "../gf_sky.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <cmath>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <limits>
#include <mutex>
#include <sstream>

using namespace gfh;

namespace {

thread_local std::string g_last_error;

std::vector<std::string> split_ws(const std::string &s)
{
	std::istringstream ss(s);
	std::vector<std::string> out;
	std::string w;
	while (ss >> w) out.push_back(w);
	return out;
}

// Device buffer owned by the catalog.
struct DevBuf {
	void *p = nullptr;
	size_t bytes = 0;
	void alloc(size_t b)
	{
		free();
		if (hipMalloc(&p, b ? b : 1) != hipSuccess) fail("device allocation failed");
		bytes = b;
	}
	template <typename T> void upload(const std::vector<T> &v)
	{
		alloc(v.size() * sizeof(T));
		if (hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) fail("upload failed");
	}
	void free()
	{
		if (p) (void)hipFree(p);
		p = nullptr;
		bytes = 0;
	}
	~DevBuf() { free(); }
};

// componentMap (common/compmgr.h:35-49): user component IDs -> sequential indices.
struct ComponentMap {
	std::vector<uint32_t> seq2comp;
	uint32_t seqIdx(uint32_t id)
	{
		for (size_t i = 0; i < seq2comp.size(); i++)
			if (seq2comp[i] == id) return (uint32_t)i;
		seq2comp.push_back(id);
		return (uint32_t)seq2comp.size() - 1;
	}
};

// read_component_map (pipeline.cpp:133-191): "a, b-c, ..." -> bitmap over
// sequential indices of components that exist.
gf_bitmap read_component_map(const Config &cfg, const std::string &key, ComponentMap &cm)
{
	gf_bitmap b{{0, 0}};
	std::string line = cfg.get(key);
	std::stringstream ss(line);
	std::string item;
	while (std::getline(ss, item, ',')) {
		size_t dash = item.find('-', item.find_first_not_of(" \t") + 1);
		uint32_t a, c;
		if (dash != std::string::npos) { a = (uint32_t)atof(item.substr(0, dash).c_str()); c = (uint32_t)atof(item.substr(dash + 1).c_str()); }
		else { a = c = (uint32_t)atof(item.c_str()); }
		for (size_t i = 0; i < cm.seq2comp.size(); i++) {
			uint32_t id = cm.seq2comp[i];
			if (id >= a && id <= c && i < 64) b.bits[i / 32] |= 1u << (i % 32);
		}
	}
	return b;
}

struct Model {
	gf_model m{};
	gf_skygen_params sp{};   // the job's binning with this model's absolute-magnitude range
	Tex1D lf;
	DevBuf lf_dev;
	std::string name;
	int user_comp = 0;
	gf_count_result counts{};
	DevBuf wpk;
	gf_draw_plan *plan = nullptr;   // the cumulative-table draw's plan, reused across runs
	~Model() { gf_skygen_plan_free(plan); }
};

struct Footprint {
	gf_footprint f{};
	double area_sr = 0;
};

} // namespace

struct gf_catalog {
	Config globals, skygen;
	std::vector<Config> model_cfgs, foot_cfgs, module_cfgs;
	int countsmap = -1;   // module_cfgs index of the countsMap output stage
	ComponentMap cm;
	gf_skygen_params sp{};
	std::vector<gf_pencil_beam> pixels;
	DevBuf pix_dev;
	std::vector<std::unique_ptr<Model>> models;
	Footprint foot;
	uint32_t seed = 42;
	std::string datadir;
	std::string band = "LSSTr";
	uint64_t chunk_rows = 30000000;
	int acc_mode = GF_ACC_FP64;
	double nstars_forced = 0;
	double maxstars = 1e8;        // skygen's maxstars (0: unlimited)
	double counted = -1;          // expected stars of the last count (after nstars), -1 before a count
	DevBuf per_pixel;             // covered-fraction weighted counts per pixel, for the self-check

	// modules
	bool has_feh = false, has_gfeh = false, has_mult = false, has_bond = false, has_phot = false,
	     has_perr = false, has_v2pm = false;
	int gal2other = 0;
	gf_FeH_data feh{};
	gf_GaussianFeH_data gfeh{};
	gf_bitmap gfeh_comps{}, mult_comps{}, phot_comps{};
	gf_multiples_data mult{};
	Tex1D mult_sec, mult_cum, mult_inv;
	DevBuf mult_dev[3];
	int nabsmag = 1;
	gf_Bond2010_data bond{};
	gf_photometry_data phot{};
	PhotTables phot_tables;
	DevBuf phot_dev[8], phot_il, phot_fb;
	int nbands = 0;
	std::vector<std::string> band_names;
	std::string bandset;   // the photometric system's name (photometry_host.cpp:94)
	gf_photoerr_data perr{};
	DevBuf perr_dev[3];
	gf_vel2pm_data v2pm{};

	// extinction (os_skygen::load_extinction_maps): north/south maps and the
	// per-pixel minimum along each DM slice; ext holds device pointers,
	// ext_h host pointers
	bool has_ext = false;
	Cube ext_north, ext_south, ext_beam;
	DevBuf ext_dev[3];
	gf_extinction ext{}, ext_h{};

	// host copies for gf_catalog_host_view
	std::vector<gf_model> models_h;
	std::vector<gf_skygen_params> models_sp_h;
	gf_photometry_data phot_h{};
	gf_multiples_data mult_h{};
	gf_photoerr_data perr_h{};
	std::vector<double> perr_kx, perr_ky;
	std::vector<uint32_t> rng_state0;

	// device state
	gf_rng rng{};
	hipStream_t stream = nullptr;    // stream-order work: plans, emission, draw passes
	hipStream_t rows = nullptr;      // the module chain's row passes
	hipStream_t side = nullptr;      // kinematics and astrometry row passes, beside photometry
	hipStream_t emit = nullptr;      // the cumulative-table draw's emission (independent of the draw passes)
	hipEvent_t planned = nullptr, emitted[2] = {};
	hipEvent_t ev[2] = {nullptr, nullptr};
	// Two output chunks are in flight: while the row passes of chunk k run
	// on `rows`, chunk k+1 is emitted and drawn on `stream` into the other
	// slot.
	struct Slot {
		DevBuf table_dev;
		uint64_t cap = 0;
		gf_table table{};
		hipEvent_t done = nullptr;        // the chunk's last launch (on `rows`)
		std::vector<hipEvent_t> ev;       // timing events of the chunk's launches
		bool pending = false;
		uint64_t n = 0;
		int comp = 0, nsub = 0;
	} slot[2];
	// Position planes of the draw passes, alternating per chain call: the
	// draw pass into buffer b waits for the row passes that read it before.
	DevBuf pos[2];
	DevBuf jtab;          // jump tables of the 65,536 streams (GF_JUMP_BITS x 65536)
	hipEvent_t drawn[2] = {}, pos_free[2] = {}, pos_free_side[2] = {}, side_done[2] = {};
	DevBuf vis_dev;                            // visible-row counter per slot
	unsigned long long *vis_host = nullptr;    // pinned copies of the counters
	int overlap = 1;
	int draw_mode = 0;   // 0: cumulative-table draw; 1: the reference's crawl, batch order and clipper

	~gf_catalog()
	{
		if (rng.state) gf_rng_free(&rng);
		for (auto e : ev) if (e) (void)hipEventDestroy(e);
		for (auto &sl : slot) {
			if (sl.done) (void)hipEventDestroy(sl.done);
			for (auto e : sl.ev) (void)hipEventDestroy(e);
		}
		for (auto e : drawn) if (e) (void)hipEventDestroy(e);
		for (auto e : pos_free) if (e) (void)hipEventDestroy(e);
		for (auto e : side_done) if (e) (void)hipEventDestroy(e);
		for (auto e : pos_free_side) if (e) (void)hipEventDestroy(e);
		if (vis_host) (void)hipHostFree(vis_host);
		if (planned) (void)hipEventDestroy(planned);
		for (auto e : emitted) if (e) (void)hipEventDestroy(e);
		models.clear();   // plans hold device buffers: freed before the streams go
		if (rows) (void)hipStreamDestroy(rows);
		if (side) (void)hipStreamDestroy(side);
		if (emit) (void)hipStreamDestroy(emit);
	}
};

namespace {

// ---------------------------------------------------------------------------
// Sky pixelization: square pixels of side dx in the north (pole b = +90, l0 =
// 90 deg) and south (b = -90, l0 = 270 deg) Lambert projections
// (skygen/os_skygen.cpp:239-283, 940-995), covered fractions by 8x8 point
// sampling of the footprint (gf_footprint_cover on the device).
// ---------------------------------------------------------------------------
void lambert_init(gf_lambert &p, double l0, double b0)
{
	p.l0 = l0; p.b0 = b0;
	p.cl = (float)cos(l0); p.cb = (float)cos(b0); p.sl = (float)sin(l0); p.sb = (float)sin(b0);
}

// deprojection of a pole projection (b0 = +-90 deg) in double
void deproj_pole(int hemi, double x, double y, double &l, double &b)
{
	double r2 = x * x + y * y;
	double sb = 1. - 0.5 * r2;
	if (hemi == 0) { b = asin(std::max(-1.0, std::min(1.0, sb))); l = M_PI / 2 + atan2(x, -y); }
	else { b = -asin(std::max(-1.0, std::min(1.0, sb))); l = 3 * M_PI / 2 + atan2(x, y); }
	l = fmod(l, 2 * M_PI); if (l < 0) l += 2 * M_PI;
}

// Project (l, b) [radians] into hemisphere `hemi`'s Lambert plane.
void proj_pole(int hemi, double l, double b, double &x, double &y)
{
	if (hemi == 0) {
		double kp = sqrt(2.0 / (1.0 + sin(b)));
		x = kp * cos(b) * sin(l - M_PI / 2);
		y = -kp * cos(b) * cos(l - M_PI / 2);
	} else {
		double kp = sqrt(2.0 / (1.0 - sin(b)));
		x = kp * cos(b) * sin(l - 3 * M_PI / 2);
		y = kp * cos(b) * cos(l - 3 * M_PI / 2);
	}
}

int pixel_grid_n(double dx) { return (int)ceil(2 * sqrt(2.0) / dx); }

// The pixel of grid cell (hemi, i, j) with `inside` of its 64 samples in the
// footprint.
gf_pencil_beam cell_pixel(int hemi, int n, int i, int j, double dx, int inside)
{
	double x0 = -n * dx / 2, xc = x0 + i * dx + 0.5 * dx, yc = x0 + j * dx + 0.5 * dx;
	double l, b;
	deproj_pole(hemi, xc, yc, l, b);
	gf_pencil_beam pb{};
	pb.cl = (float)cos(l); pb.cb = (float)cos(b); pb.sl = (float)sin(l); pb.sb = (float)sin(b);
	pb.coveredFraction = (float)inside / 64.f;
	pb.dx = (float)dx;
	pb.dA = pb.dx * pb.dx;
	pb.projIdx = hemi;
	pb.X = (float)xc; pb.Y = (float)yc;
	pb.extIdx = -1;
	return pb;
}

// The pixels of a one-region footprint -- a beam without a hole, or a rect
// in Galactic coordinates -- on the catalog's pole grids: the cells around
// the region's projection, each with the number of its 8 x 8 samples that
// the device's coverage test (sample_inside) finds inside, in (hemisphere,
// i, j) order: the pixels a catalog job on that footprint gets.
std::vector<gf_pencil_beam> pixelize_region(const gf_footprint &fp, double dx)
{
	std::vector<gf_pencil_beam> out;
	if (fp.nregion != 1 || !((fp.r[0].kind == 2 && fp.r[0].hole <= 0.) || (fp.r[0].kind == 1 && !fp.r[0].equ)))
		fail("pixelize_region: not a single beam or Galactic rect");
	const gf_foot_region &f = fp.r[0];
	const int n = pixel_grid_n(dx);
	const double x0 = -n * dx / 2, y0 = -n * dx / 2, d2r = M_PI / 180.;
	for (int hemi = 0; hemi < 2; hemi++) {
		// points of the region in this hemisphere whose projections bound it
		std::vector<std::pair<double, double>> lb;
		if (f.kind == 2) {
			// the cap's centre and boundary circle, and the arc of the equator
			// inside it: the projection maps the hemisphere's part of the cap
			// onto a region bounded by the images of those two curves, so
			// their samples bound it (to within the sampling step, which the
			// two-cell margin below covers); a cap that crosses the equator by
			// less than the step still reaches both hemispheres along the arc
			const double bc[3] = {f.cb * f.cl, f.cb * f.sl, f.sb};
			const double r = acos(std::max(-1.0, std::min(1.0, f.cosr)));
			const double e1[3] = {-f.sl, f.cl, 0}, e2[3] = {-f.sb * f.cl, -f.sb * f.sl, f.cb};
			for (int ir = 0; ir <= 1; ir++) {
				const double rr = ir ? r : 0.0;
				const int na = ir == 0 ? 1 : 192;
				for (int ia = 0; ia < na; ia++) {
					double a = 2 * M_PI * ia / na, v[3];
					for (int q = 0; q < 3; q++) v[q] = cos(rr) * bc[q] + sin(rr) * (cos(a) * e1[q] + sin(a) * e2[q]);
					lb.emplace_back(atan2(v[1], v[0]), asin(std::max(-1.0, std::min(1.0, v[2]))));
				}
			}
			const double bcen = asin(std::max(-1.0, std::min(1.0, f.sb)));
			if (bcen - r <= 0 && bcen + r >= 0) {
				const double lc = atan2(f.sl, f.cl);
				const double cr = f.cb > 1e-12 ? f.cosr / f.cb : (f.cosr <= 0 ? -1.0 : 2.0);
				const double half = cr >= 1.0 ? 0.0 : (cr <= -1.0 ? M_PI : acos(cr));
				for (int t = 0; t <= 96; t++) lb.emplace_back(lc + half * (2.0 * t / 96.0 - 1.0), 0.0);
			}
		} else {
			// the rect's edges (its projection is bounded by their images),
			// cut at the equator
			double b0 = f.b0, b1 = f.b1;
			if (hemi == 0) b0 = std::max(b0, 0.0); else b1 = std::min(b1, 0.0);
			if (b0 > b1) continue;
			const double l0 = f.l0, l1 = f.l1 - f.l0 >= 360. ? f.l0 + 360. : f.l1;
			for (int t = 0; t <= 256; t++) {
				const double l = l0 + (l1 - l0) * t / 256.0, b = b0 + (b1 - b0) * t / 256.0;
				lb.emplace_back(l * d2r, b0 * d2r);
				lb.emplace_back(l * d2r, b1 * d2r);
				lb.emplace_back(l0 * d2r, b * d2r);
				lb.emplace_back(l1 * d2r, b * d2r);
			}
		}
		double xmin = 1e9, xmax = -1e9, ymin = 1e9, ymax = -1e9;
		bool any = false;
		for (auto &p : lb) {
			if ((hemi == 0 && p.second < 0) || (hemi == 1 && p.second > 0)) continue;
			double x, y;
			proj_pole(hemi, p.first, p.second, x, y);
			xmin = std::min(xmin, x); xmax = std::max(xmax, x);
			ymin = std::min(ymin, y); ymax = std::max(ymax, y);
			any = true;
		}
		if (!any) continue;
		const int i0 = std::max(0, (int)floor((xmin - x0) / dx) - 2), i1 = std::min(n, (int)floor((xmax - x0) / dx) + 3);
		const int j0 = std::max(0, (int)floor((ymin - y0) / dx) - 2), j1 = std::min(n, (int)floor((ymax - y0) / dx) + 3);
		// a beam's cells well inside or well outside the cap need no samples:
		// within a cell of the pole grids whose corners are all at plane
		// radius < sqrt 2 the unit vector moves by at most sqrt 2 times the
		// plane distance, and no sample is farther than dx / sqrt 2 from the
		// centre, so every sample's cap product is within dx of the centre's
		const bool beam = f.kind == 2 && f.bmin <= 0. && f.hole <= 0.;
		const double margin = 1.01 * dx + 1e-9;
		for (int i = i0; i < i1; i++)
			for (int j = j0; j < j1; j++) {
				double xa = x0 + i * dx, ya = y0 + j * dx;
				double fx = (xa > 0 || xa + dx < 0) ? std::min(fabs(xa), fabs(xa + dx)) : 0.;
				double fy = (ya > 0 || ya + dx < 0) ? std::min(fabs(ya), fabs(ya + dx)) : 0.;
				if (fx * fx + fy * fy > 2.0) continue;
				if (beam) {
					const double gx = std::max(fabs(xa), fabs(xa + dx)), gy = std::max(fabs(ya), fabs(ya + dx));
					if (gx * gx + gy * gy < 2.0 - 1e-9) {
						const double xc = xa + 0.5 * dx, yc = ya + 0.5 * dx, r2 = xc * xc + yc * yc;
						const double sc = sqrt(1.0 - 0.25 * r2), z = 1.0 - 0.5 * r2;
						const double v0 = hemi == 0 ? -xc * sc : xc * sc, v1 = -yc * sc, v2 = hemi == 0 ? z : -z;
						const double dc = v0 * (f.cb * f.cl) + v1 * (f.cb * f.sl) + v2 * f.sb;
						if (dc - f.cosr > margin) { out.push_back(cell_pixel(hemi, n, i, j, dx, 64)); continue; }
						if (f.cosr - dc > margin) continue;
					}
				}
				int inside = 0;
				for (int a = 0; a < 8; a++)
					for (int b2 = 0; b2 < 8; b2++)
						inside += sample_inside(fp, hemi, xa + dx * (a + 0.5) / 8, ya + dx * (b2 + 0.5) / 8) ? 1 : 0;
				if (inside) out.push_back(cell_pixel(hemi, n, i, j, dx, inside));
			}
	}
	return out;
}

// The catalog's pixels: cell coverage on the device, the covered cells in
// (hemisphere, i, j) order.
void pixelize(gf_catalog &c, double dx)
{
	const int n = pixel_grid_n(dx);
	const size_t cells = 2 * (size_t)n * n;
	DevBuf d;
	d.alloc(cells);
	if (gf_footprint_cover(&c.foot.f, dx, n, (uint8_t *)d.p, nullptr)) fail("gf_footprint_cover failed");
	std::vector<uint8_t> cover(cells);
	if (hipMemcpy(cover.data(), d.p, cells, hipMemcpyDeviceToHost) != hipSuccess) fail("footprint cover copy failed");
	c.pixels.clear();
	for (int hemi = 0; hemi < 2; hemi++)
		for (int i = 0; i < n; i++)
			for (int j = 0; j < n; j++) {
				int inside = cover[((size_t)hemi * n + i) * n + j];
				if (inside) c.pixels.push_back(cell_pixel(hemi, n, i, j, dx, inside));
			}
	c.sp.npixels = (int)c.pixels.size();
}

// One footprint configuration file -> one region (load_footprint,
// footprint.cpp:729-744).
// A rect region, longitudes as footprint.cpp:486-515 reduces them: to
// [0, 360), l0 >= l1 wrapping through 0 (l0 = l1 = 0 is the whole ring),
// and within 1e-5 of a full turn the whole ring (:515-521).
gf_foot_region rect_region(double l0, double l1, double b0, double b1)
{
	gf_foot_region f{};
	l0 = fmod(l0, 360.);
	l1 = fmod(l1, 360.);
	if (l0 < 0) l0 += 360.;
	if (l1 < 0) l1 += 360.;
	if (l0 >= l1) { if (l0 == 0) l1 = 360.; else l0 -= 360.; }
	if (fabs((l1 - l0) / 360. - 1.) < 1e-5) { l0 = 0.; l1 = 360.; }
	f.kind = 1;
	f.l0 = l0; f.l1 = l1; f.b0 = b0; f.b1 = b1;
	return f;
}

gf_foot_region load_region(const Config &fc)
{
	std::string type = fc.get("type", "");
	if (type.empty()) fail("The footprint configuration file must contain the 'type' keyword.");
	std::string cs = fc.get("coordsys", "gal");
	if (cs != "gal" && cs != "equ")
		fail("Don't know how to convert from coordinate system '" + cs + "' to galactic coordinates.");
	gf_foot_region f{};
	f.equ = cs == "equ";
	if (type == "rect") {
		// footprint.cpp:486-515: longitudes reduced to [0, 360), l0 >= l1
		// wraps (l0 = l1 = 0 is the whole ring)
		std::vector<double> r = fc.getv("rect", "0 1 0 1");
		if (r.size() < 4) fail("rect needs l0 l1 b0 b1");
		const bool equ = f.equ;
		f = rect_region(r[0], r[1], r[2], r[3]);
		f.equ = equ;
	} else if (type == "beam") {
		// footprint.cpp:675-700: centre l b, radius r, optional hole radius
		if (!fc.count("footprint_beam")) fail("Configuration key footprint_beam must be set");
		std::vector<double> v = fc.getv("footprint_beam", "");
		if (v.size() < 3) fail("footprint_beam needs l b r");
		double l = v[0] * M_PI / 180., b = v[1] * M_PI / 180.;
		if (f.equ) equ_to_gal(l, b, l, b);
		f.kind = 2;
		f.equ = 0;
		f.cl = cos(l); f.cb = cos(b); f.sl = sin(l); f.sb = sin(b);
		f.cosr = cos(v[2] * M_PI / 180.);
		double rhole = v.size() > 3 ? v[3] : 0.;
		f.hole = rhole > 0. ? 1. - cos(rhole * M_PI / 180.) : 0.;
	} else fail("The requested footprint type " + type + " is unknown.");
	f.bmin = fc.getd("bmin", 0.);
	return f;
}

void load_footprints(gf_catalog &c)
{
	gf_footprint &f = c.foot.f;
	f = gf_footprint{};
	if (c.foot_cfgs.size() > GF_FOOT_MAX)
		fail("at most " + std::to_string(GF_FOOT_MAX) + " footprint configurations are supported");
	for (auto &fc : c.foot_cfgs) f.r[f.nregion++] = load_region(fc);
}

// ---------------------------------------------------------------------------
// Extinction maps (skygen/os_skygen.cpp:285-342, 699-814)
// ---------------------------------------------------------------------------

// find_extinction_minima (skygen/os_skygen.cpp:285-341): for every DM slice
// z, the smallest extinction one slice closer over the map texels that can
// overlap the pixel (texels outside the hemisphere with zero extinction
// skipped), folded into out[z]; proj_x, proj_y: the pixel centre in the
// map's projection.
void extinction_minima(float *out, double proj_x, double proj_y, double dx, const Cube &tex)
{
	const float *tcx = tex.tc[0], *tcy = tex.tc[1];
	double dr = 1.0001 / 2. * sqrt(2.) * (dx + std::max(1. / tcx[1], 1. / tcy[1]));
	int xi = (int)((proj_x - dr - tcx[0]) * tcx[1]);  xi = std::max(0, std::min(xi - 1, tex.nx));
	int yi = (int)((proj_y - dr - tcy[0]) * tcy[1]);  yi = std::max(0, std::min(yi - 1, tex.ny));
	int xie = (int)((proj_x + dr - tcx[0]) * tcx[1]); xie = std::max(0, std::min(xie + 2, tex.nx));
	int yie = (int)((proj_y + dr - tcy[0]) * tcy[1]); yie = std::max(0, std::min(yie + 2, tex.ny));
	for (int x = xi; x < xie; x++) {
		for (int y = yi; y < yie; y++) {
			if (tex.at(x, y, tex.nz - 1) == 0.f) {
				double xx = (double)x / tcx[1] + tcx[0], yy = (double)y / tcy[1] + tcy[0];
				if (xx * xx + yy * yy > 2.) continue;
			}
			for (int z = 0; z < tex.nz; z++) {
				float Ar = z == 0 ? 0.f : tex.at(x, y, z - 1);
				out[z] = std::min(out[z], Ar);
			}
		}
	}
}

// The minima of every pixel, nz per pixel ([pixel][z]), over the north and
// south maps that are loaded (os_skygen.cpp:773-784).
std::vector<float> pixel_extinction_minima(const gf_catalog &c, const std::vector<gf_pencil_beam> &pix)
{
	const Cube &ref = c.ext_north.v.empty() ? c.ext_south : c.ext_north;
	const int nz = ref.nz;
	std::vector<float> out((size_t)pix.size() * nz, std::numeric_limits<float>::infinity());
	for (size_t i = 0; i < pix.size(); i++) {
		const gf_pencil_beam &p = pix[i];
		double l = atan2((double)p.sl, (double)p.cl), b = atan2((double)p.sb, (double)p.cb);
		double x, y;
		if (!c.ext_north.v.empty()) { proj_pole(0, l, b, x, y); extinction_minima(&out[i * nz], x, y, p.dx, c.ext_north); }
		if (!c.ext_south.v.empty()) { proj_pole(1, l, b, x, y); extinction_minima(&out[i * nz], x, y, p.dx, c.ext_south); }
	}
	return out;
}

// The minima as the draw samples them: a 3D texture with pixel g at texel
// (g div 2048, g mod 2048) and the DM slices along z
// (skygen/os_skygen.cpp:767-775).  Texels past the last pixel hold 0, not
// the reference's infinity: a pixel's own texel is read with weight 1 and its
// neighbour in x with weight 0, and 0 * infinity would make the sample NaN.
void beam_texture(Cube &beam, const std::vector<float> &minima, size_t npix, int nz, const float tcz[2])
{
	beam.nx = std::max<int>(1, (int)(npix / 2048 + (npix % 2048 ? 1 : 0)));
	beam.ny = std::max<int>(1, (int)std::min<size_t>(npix, 2048));
	beam.nz = nz;
	beam.v.assign((size_t)beam.nx * beam.ny * beam.nz, 0.f);
	texcoord_from_range(0, (float)beam.nx, 0, (float)beam.nx, beam.tc[0]);
	texcoord_from_range(0, (float)beam.ny, 0, (float)beam.ny, beam.tc[1]);
	beam.tc[2][0] = tcz[0];
	beam.tc[2][1] = tcz[1];
	for (size_t g = 0; g < npix; g++)
		for (int z = 0; z < nz; z++) beam.at((int)(g / 2048), (int)(g % 2048), z) = minima[g * nz + z];
}

gf_tex3d tex3d_of(const Cube &c, const float *data)
{
	gf_tex3d t{};
	t.data = data;
	t.nx = c.nx; t.ny = c.ny; t.nz = c.nz;
	for (int a = 0; a < 3; a++) { t.tc[a][0] = c.tc[a][0]; t.tc[a][1] = c.tc[a][1]; }
	return t;
}

// os_skygen::load_extinction_maps: "north", "south" map files, "nscale",
// "sscale" factors; pixels get extIdx = their index.
void load_extinction(gf_catalog &c, const std::string &econf)
{
	if (econf.empty()) return;
	Config cfg;
	if (!cfg.load(c.skygen.resolve(econf))) fail("cannot read " + c.skygen.resolve(econf));
	std::string nf = cfg.get("north", ""), sf = cfg.get("south", "");
	float nscale = (float)cfg.getd("nscale", 1.), sscale = (float)cfg.getd("sscale", 1.);
	if (nf.empty() && sf.empty()) return;
	if (!nf.empty()) {
		c.ext_north = load_fits_cube(cfg.resolve(nf));
		for (auto &v : c.ext_north.v) v *= nscale;
	}
	if (!sf.empty()) {
		c.ext_south = load_fits_cube(cfg.resolve(sf));
		for (auto &v : c.ext_south.v) v *= sscale;
	}
	if (!nf.empty() && !sf.empty() && c.ext_north.nz != c.ext_south.nz)
		fail("north and south extinction maps differ in depth");
	const Cube &ref = nf.empty() ? c.ext_south : c.ext_north;
	const int np = (int)c.pixels.size();
	for (int i = 0; i < np; i++) c.pixels[i].extIdx = i;
	beam_texture(c.ext_beam, pixel_extinction_minima(c, c.pixels), c.pixels.size(), ref.nz, ref.tc[2]);
	Cube &beam = c.ext_beam;
	c.ext_dev[0].upload(c.ext_north.v);
	c.ext_dev[1].upload(c.ext_south.v);
	c.ext_dev[2].upload(beam.v);
	if (!nf.empty()) {
		c.ext.north = tex3d_of(c.ext_north, (const float *)c.ext_dev[0].p);
		c.ext_h.north = tex3d_of(c.ext_north, c.ext_north.v.data());
	}
	if (!sf.empty()) {
		c.ext.south = tex3d_of(c.ext_south, (const float *)c.ext_dev[1].p);
		c.ext_h.south = tex3d_of(c.ext_south, c.ext_south.v.data());
	}
	c.ext.beam = tex3d_of(beam, (const float *)c.ext_dev[2].p);
	c.ext_h.beam = tex3d_of(beam, beam.v.data());
	c.has_ext = true;
}

// os_skygen::load_skyPixelizationConfig + reset_absmag (skygen/os_skygen.cpp:401-430)
void load_sky_pixelization(gf_catalog &c, double &dx)
{
	const Config &cfg = c.skygen;
	dx = cfg.getd("dx", 1.0) * M_PI / 180.;
	auto any = [&](const char *a, const char *b) { return cfg.count(a) ? cfg.getd(a) : cfg.getd(b); };
	gf_skygen_params &sp = c.sp;
	sp.M0 = (float)any("M0", "ri0");
	sp.M1 = (float)any("M1", "ri1");
	sp.m0 = (float)any("m0", "r0");
	sp.m1 = (float)any("m1", "r1");
	sp.dm = (float)cfg.getd("dm");
	sp.dM = sp.dm;
	sp.dmin = (float)cfg.getd("dmin", 0.);
	sp.dmax = (float)cfg.getd("dmax", 0.);
	sp.nM = (int)round((sp.M1 - sp.M0) / sp.dM);
	sp.M1 = (float)(sp.M1 - 0.5 * sp.dM);
	sp.nm = (int)round((sp.m1 - sp.m0) / sp.dm);
	sp.m0 = (float)(sp.m0 + 0.5 * sp.dm);
	sp.nthreads = 7680;   // the reference's 120 x 64 crawl (skygen/skyconfig_impl.h:276-289)
	lambert_init(sp.proj[0], M_PI / 2, M_PI / 2);
	lambert_init(sp.proj[1], 3 * M_PI / 2, -M_PI / 2);
	c.band = cfg.get("band", "LSSTr");
}

// A model (skygen/model_expDisk.cpp, model_powerLawEllipsoid.cpp) with load_lf
// (skygen/model_lib.h:37-95).
void load_model(gf_catalog &c, const Config &mc)
{
	auto md = std::make_unique<Model>();
	gf_model &m = md->m;
	std::string kind = mc.get("model");
	md->name = kind;
	std::string ctr = mc.get("center", "galplane"), orient = mc.get("orientation", "galplane");
	load_transform(m.T, m.M, ctr, orient, mc);
	if (kind == "exponential disk") {
		m.kind = GF_MODEL_EXPDISK;
		m.l = (float)mc.getd("l");
		m.h = (float)mc.getd("h");
	} else if (kind == "power law ellipsoid") {
		m.kind = GF_MODEL_POWERLAW;
		m.ca = (float)mc.getd("c_a", 1.0);
		m.ba = (float)mc.getd("b_a", 1.0);
		m.n = (float)mc.getd("n", -3.0);
	} else fail("unsupported density model '" + kind + "'");
	md->user_comp = (int)mc.getd("comp");
	m.comp = (int)c.cm.seqIdx(md->user_comp);
	// luminosity function
	if (mc.count("lumfunc")) {
		auto rows = load_columns(mc.resolve(mc.get("lumfunc")), 2);
		std::vector<double> X, Y;
		for (auto &r : rows) { X.push_back(r[0]); Y.push_back(r[1]); }
		md->lf = resample_clip_1d(X, Y, 0.f, 1024);
	} else {
		md->lf = constant_1d(1.f, -100.f, 100.f);
	}
	// renormalise f so rho(LF site) = f (the cuts are applied after)
	float lfl = (float)mc.getd("lf_l", 0.), lfb = (float)mc.getd("lf_b", 0.), lfD = (float)mc.getd("lf_D", 0.);
	double lr = (double)(0.01745329251994329576923691f * lfl), br = (double)(0.01745329251994329576923691f * lfb);
	float dcl = (float)cos(lr), dcb = (float)cos(br), dsl = (float)sin(lr), dsb = (float)sin(br);
	float px = lfD * dcl * dcb, py = lfD * dsl * dcb, pz = lfD * dsb;
	m.f = 1.f;
	float vx = px + m.T[0], vy = py + m.T[1], vz = pz + m.T[2];
	float x = m.M[0][0] * vx + m.M[0][1] * vy + m.M[0][2] * vz;
	float y = m.M[1][0] * vx + m.M[1][1] * vy + m.M[1][2] * vz;
	float z = m.M[2][0] * vx + m.M[2][1] * vy + m.M[2][2] * vz;
	float norm;
	if (m.kind == GF_MODEL_EXPDISK) {
		float r = sqrtf(x * x + y * y);
		float e = -r / m.l - fabsf(z) / m.h;
		norm = (float)exp((double)e);
	} else {
		float yb = y / m.ba, zc = z / m.ca;
		float rSq = x * x + yb * yb + zc * zc;
		norm = (float)exp((double)(0.5f * m.n) * log((double)rSq));
	}
	if (!(norm != 0.f) || !std::isfinite(norm)) fail("luminosity-function site has zero density");
	float f0 = (float)mc.getd("f", 1.0);
	m.f = f0 / norm;
	if (m.kind == GF_MODEL_EXPDISK) {
		float rc = (float)mc.getd("rcut", 0.);
		m.r_cut2 = rc * rc;
	} else {
		float rmin = (float)mc.getd("rmin", 0.), rmax = (float)mc.getd("rmax", 0.);
		m.rminSq = rmin * rmin;
		m.rmaxSq = rmax * rmax;
	}
	m.norm = 1.f;
	md->lf_dev.upload(md->lf.v);
	m.lf.data = (const float *)md->lf_dev.p;
	m.lf.n = (int)md->lf.v.size();
	m.lf.tc[0] = md->lf.tc[0];
	m.lf.tc[1] = md->lf.tc[1];
	c.models.push_back(std::move(md));
}

void construct_FeH(gf_catalog &c, const Config &cfg)
{
	gf_FeH_data &p = c.feh;
	load_transform(p.T, p.M, cfg.get("center", "galplane"), cfg.get("orientation", "galplane"), cfg);
	p.A[0] = (float)cfg.getd("A0", 0.63);
	p.sigma[0] = (float)cfg.getd("sigma0", 0.20);
	p.offs[0] = (float)cfg.getd("offs0", 0.00);
	p.A[1] = (float)cfg.getd("A1", 0.37);
	p.sigma[1] = (float)cfg.getd("sigma1", 0.20);
	p.offs[1] = (float)cfg.getd("offs1", 0.14);
	p.Hmu = (float)cfg.getd("Hmu", 500.);
	p.muInf = (float)cfg.getd("muInf", -0.82);
	p.DeltaMu = (float)cfg.getd("deltaMu", 0.55);
	double sumA = (double)(p.A[0] + p.A[1]);
	p.A[0] = (float)(p.A[0] / sumA);
	p.A[1] = (float)(p.A[1] / sumA);
	p.sigma[2] = (float)cfg.getd("sigmaH", 0.30);
	p.offs[2] = (float)cfg.getd("offsH", -1.46);
	p.comp_thin = read_component_map(cfg, "comp_thin", c.cm);
	p.comp_thick = read_component_map(cfg, "comp_thick", c.cm);
	p.comp_halo = read_component_map(cfg, "comp_halo", c.cm);
	c.has_feh = true;
}

void construct_GaussianFeH(gf_catalog &c, const Config &cfg)
{
	c.gfeh_comps = read_component_map(cfg, "comp", c.cm);
	std::vector<double> mean = cfg.getv("mean", ""), sigma = cfg.getv("sigma", "");
	if (mean.empty() || mean.size() >= 10 || sigma.size() != mean.size()) fail("GaussianFeH: bad mean/sigma");
	gf_GaussianFeH_data &p = c.gfeh;
	p.ncomp = (int)mean.size();
	for (int i = 0; i < p.ncomp; i++) { p.mean[i] = (float)mean[i]; p.sigma[i] = (float)sigma[i]; }
	std::vector<double> wt = p.ncomp > 1 ? cfg.getv("wt", "") : std::vector<double>(1, 1.);
	if ((int)wt.size() != p.ncomp) fail("GaussianFeH: wt size");
	double s = 0;
	for (int i = 0; i < p.ncomp; i++) { s += wt[i]; p.pcum[i] = (float)s; }
	for (int i = 0; i < p.ncomp; i++) p.pcum[i] /= p.pcum[p.ncomp - 1];
	c.has_gfeh = true;
}

void farr5(float dst[5], const std::vector<double> &v)
{
	for (int i = 0; i < 5; i++) dst[i] = i < (int)v.size() ? (float)v[i] : 0.f;
}

// os_Bond2010::construct (modules/Bond2010_host.cpp:136-227)
void construct_Bond2010(gf_catalog &c, const Config &cfg)
{
	gf_Bond2010_data &p = c.bond;
	p.comp_thin = read_component_map(cfg, "comp_thin", c.cm);
	p.comp_thick = read_component_map(cfg, "comp_thick", c.cm);
	p.comp_halo = read_component_map(cfg, "comp_halo", c.cm);
	load_transform(p.T, p.M, cfg.get("center", "galplane"), cfg.get("orientation", "galplane"), cfg);
	auto g = [&](const char *k, const char *d) {
		std::vector<double> v = cfg.getv(k, d);
		std::vector<double> f;
		for (double x : v) f.push_back((double)(float)x);
		return f;
	};
	farr5(p.disk_1.m[0], g("vR", "0 0 0 0 0"));
	farr5(p.disk_1.m[1], g("vPhi", "-194 19.2 1.25 0 0"));
	farr5(p.disk_1.m[2], g("vZ", "0 0 0 0 0"));
	farr5(p.disk_1.ss[0], g("sigmaR", "40 5 1.5 0 0"));
	farr5(p.disk_1.ss[1], g("sigmaRPhi", "0 0 0 0 0"));
	farr5(p.disk_1.ss[2], g("sigmaRZ", "0 0 0 0 0"));
	farr5(p.disk_1.ss[3], g("sigmaPhi1", "12 1.8 2 0 0"));
	farr5(p.disk_1.ss[4], g("sigmaPhiZ", "0 0 0 0 0"));
	farr5(p.disk_1.ss[5], g("sigmaZ", "25 4 1.5 0 0"));
	p.disk_2 = p.disk_1;
	farr5(p.disk_2.ss[3], g("sigmaPhi2", "34 1.2 2 0 0"));
	float DeltavPhi = (float)cfg.getd("DeltavPhi", 34.0);
	p.disk_2.m[1][0] += DeltavPhi;
	float fk = (float)cfg.getd("fk", 3.0);
	p.fk = (float)(fk / (1. + fk));
	farr5(p.halo.m[0], g("HvR", "0 0 0 0 0"));
	farr5(p.halo.m[1], g("HvTh", "0 0 0 0 0"));
	farr5(p.halo.m[2], g("HvPhi", "0 0 0 0 0"));
	farr5(p.halo.ss[0], g("HsigmaR", "141 0 0 0 0"));
	farr5(p.halo.ss[1], g("HsigmaRTh", "0 0 0 0 0"));
	farr5(p.halo.ss[2], g("HsigmaRPhi", "0 0 0 0 0"));
	farr5(p.halo.ss[3], g("HsigmaTh", "75 0 0 0 0"));
	farr5(p.halo.ss[4], g("HsigmaThPhi", "0 0 0 0 0"));
	farr5(p.halo.ss[5], g("HsigmaPhi", "85 0 0 0 0"));
	c.has_bond = true;
}

// os_vel2pm::construct (modules/vel2pm_host.cpp:79-97)
void construct_vel2pm(gf_catalog &c, const Config &cfg)
{
	std::string cs = cfg.get("coordsys", "gal");
	gf_vel2pm_data &p = c.v2pm;
	if (cs == "gal") p.coordsys = 0;
	else if (cs == "equ") p.coordsys = 1;
	else fail("vel2pm: unknown coordinate system " + cs);
	p.Rg = (float)cfg.getd("Rg");
	p.vLSR = (float)cfg.getd("vLSR", -220.0);
	p.u0 = (float)cfg.getd("u0", -10.0);
	p.v0 = (float)cfg.getd("v0", -5.3);
	p.w0 = (float)cfg.getd("w0", 7.2);
	c.has_v2pm = true;
}

void construct_gal2other(gf_catalog &c, const Config &cfg)
{
	std::string cs = cfg.get("coordsys", "gal");
	if (cs == "equ") c.gal2other = 1;
	else if (cs != "gal") fail("gal2other: unknown coordinate system " + cs);
}

// os_photometry::construct (modules/photometry_host.cpp:87-349)
void construct_photometry(gf_catalog &c, const Config &cfg)
{
	c.phot_comps = read_component_map(cfg, "comp", c.cm);
	std::string bandset = cfg.get("bandset", "LSSTugrizy");
	c.bandset = bandset;
	c.band_names = split_ws(cfg.get("bands", "LSSTu LSSTg LSSTr LSSTi LSSTz LSSTy"));
	c.nbands = (int)c.band_names.size();
	int ncolors = c.nbands - 1;
	if (ncolors < 1 || ncolors >= 16) fail("photometry: 2..16 bands supported");
	std::string bband = cfg.get("bootstrap_band", "LSSTr");
	auto it = std::find(c.band_names.begin(), c.band_names.end(), bband);
	if (it == c.band_names.end()) fail("Bootstrap band must be listed in the 'bands' keyword.");
	gf_photometry_data &p = c.phot;
	p.ncolors = (uint32_t)ncolors;
	p.bidx = (uint32_t)(it - c.band_names.begin());
	for (int i = 0; i < 17; i++) p.reddening[i] = 1.f;
	std::vector<double> red = cfg.getv("reddening", "");
	for (int i = 0; i < c.nbands && i < (int)red.size(); i++) p.reddening[i] = (float)red[i];
	std::string file = cfg.get("file", "");
	if (file.empty()) file = c.datadir + "/" + bandset + ".photosys.txt";
	else file = cfg.resolve(file);
	auto rows = load_columns(file, 2 + ncolors);
	double dFeH = cfg.getd("dFeH", 0.), dMr = cfg.getd("dM", 0.);
	if (dFeH == 0.) { auto v = cfg.getv("FeH_grid", "-3 0.5 0.01"); dFeH = v.size() > 2 ? v[2] : 0.01; }
	if (dMr == 0.) { auto v = cfg.getv("absmag_grid", "3 15 0.01"); dMr = v.size() > 2 ? v[2] : 0.01; }
	c.phot_tables = build_photometry_tables(rows, ncolors, dFeH, dMr);
	PhotTables &pt = c.phot_tables;
	size_t per = (size_t)pt.nFeH * pt.nMr * 4;
	for (int t = 0; t < pt.ntex; t++) {
		std::vector<float> col(pt.color.begin() + t * per, pt.color.begin() + (t + 1) * per);
		std::vector<float> fl(pt.flags.begin() + t * per, pt.flags.begin() + (t + 1) * per);
		c.phot_dev[t].upload(col);
		c.phot_dev[4 + t].upload(fl);
		p.color[t] = (const float *)c.phot_dev[t].p;
		p.cflags[t] = (const float *)c.phot_dev[4 + t].p;
	}
	// the copy by lookup quad: per lower corner (i, j) the four corner
	// texels' ntex colour float4s, and each corner's flags (0 or 1) as one
	// bit per colour
	{
		const int nF = pt.nFeH, nM = pt.nMr;
		const size_t ntexel = (size_t)nF * nM, nt = (size_t)pt.ntex;
		std::vector<uint16_t> bits(ntexel, 0);
		for (size_t k = 0; k < ntexel; k++)
			for (size_t t = 0; t < nt; t++)
				for (int q = 0; q < 4; q++) {
					const float f = pt.flags[t * per + k * 4 + q];
					if (f != 0.f && f != 1.f) fail("photometry flags must be 0 or 1");
					if (f != 0.f) bits[k] |= (uint16_t)(1u << (4 * t + q));
				}
		std::vector<float> qd(ntexel * nt * 16);
		std::vector<uint16_t> fb(ntexel * 4);
		for (int j = 0; j < nM; j++)
			for (int i = 0; i < nF; i++) {
				const size_t k = (size_t)j * nF + i;
				const int i1 = std::min(i + 1, nF - 1), j1 = std::min(j + 1, nM - 1);
				const size_t corner[4] = {k, (size_t)j * nF + i1, (size_t)j1 * nF + i, (size_t)j1 * nF + i1};
				for (int cn = 0; cn < 4; cn++) {
					fb[4 * k + cn] = bits[corner[cn]];
					for (size_t t = 0; t < nt; t++)
						for (int q = 0; q < 4; q++)
							qd[((k * nt + t) * 4 + cn) * 4 + q] = pt.color[t * per + corner[cn] * 4 + q];
				}
			}
		c.phot_il.upload(qd);
		c.phot_fb.upload(fb);
		p.texels = (const float *)c.phot_il.p;
		p.flagbits = (const uint16_t *)c.phot_fb.p;
	}
	p.nFeH = pt.nFeH;
	p.nMr = pt.nMr;
	p.tcFeH[0] = pt.tcFeH[0]; p.tcFeH[1] = pt.tcFeH[1];
	p.tcMr[0] = pt.tcMr[0]; p.tcMr[1] = pt.tcMr[1];
	c.has_phot = true;
}

// os_unresolvedMultiples::construct (modules/unresolvedMultiples_host.cpp:126-189)
void construct_multiples(gf_catalog &c, const Config &cfg)
{
	c.mult_comps = read_component_map(cfg, "comp", c.cm);
	std::string algo = cfg.get("algorithm");
	if (algo == "LF_M2_gt_M1") c.mult.algo = 1;
	else if (algo == "LF") c.mult.algo = 2;
	else if (algo == "equal_mass") c.mult.algo = 3;
	else fail("unknown secondary magnitude algorithm " + algo);
	std::string ff = cfg.get("fraction_file", "");
	if (!ff.empty()) {
		auto rows = load_columns(cfg.resolve(ff), 2);
		std::vector<double> X, Y;
		for (auto &r : rows) { X.push_back(r[0]); Y.push_back(r[1]); }
		c.mult_sec = resample_1d(X, Y, 64);
	} else {
		float frac = (float)cfg.getd("fraction", 1.0);
		c.mult_sec = constant_1d(frac, -100.f, 100.f);
	}
	std::vector<double> x, y;
	std::string lff = cfg.get("lumfunc", "");
	if (!lff.empty()) {
		auto rows = load_columns(cfg.resolve(lff), 2);
		for (auto &r : rows) { x.push_back(r[0]); y.push_back(r[1]); }
	} else {
		x = {-100, 100};
		y = {1, 1};
	}
	const int NPIX = 256;
	Spline lf(x, y);
	double dx = (x.back() - x.front()) / (NPIX - 1);
	std::vector<double> xc(NPIX), yc(NPIX);
	xc[0] = x.front();
	yc[0] = 0;
	double yprev = lf(x.front());
	for (int i = 1; i < NPIX; i++) {
		double xx = x.front() + i * dx, yy = lf(xx);
		double dy = yy - yprev;
		double dA = (yprev + 0.5 * dy) * dx;
		xc[i] = xx;
		yc[i] = yc[i - 1] + dA;
		yprev = yy;
	}
	double norm = yc.back();
	for (auto &v : yc) v /= norm;
	c.mult_cum = resample_1d(xc, yc, NPIX);
	c.mult_inv = resample_1d(yc, xc, NPIX);
	Tex1D *ts[3] = {&c.mult_sec, &c.mult_cum, &c.mult_inv};
	gf_tex1d *dst[3] = {&c.mult.secProb, &c.mult.cumLF, &c.mult.invCumLF};
	for (int i = 0; i < 3; i++) {
		c.mult_dev[i].upload(ts[i]->v);
		dst[i]->data = (const float *)c.mult_dev[i].p;
		dst[i]->n = (int)ts[i]->v.size();
		dst[i]->tc[0] = ts[i]->tc[0];
		dst[i]->tc[1] = ts[i]->tc[1];
	}
	c.nabsmag = (int)cfg.getd("ncomp", 2);
	c.has_mult = true;
}

// os_photometricErrors::construct (modules/photometricErrors_host.cpp:232-283):
// <bandset>.<band>.file = path ('internal' = <datadir>/<bandset>.<band>.photoerr.txt)
void construct_photoerr(gf_catalog &c, const Config &cfg)
{
	if (!c.has_phot) fail("photometricErrors needs the photometry module");
	std::vector<double> kx, ky;
	c.perr.nbands = c.nbands;
	c.perr.koff[0] = 0;
	for (int b = 0; b < c.nbands; b++) {
		std::string key;
		for (auto &kvp : cfg.kv) {
			const std::string &k = kvp.first;
			std::string suffix = "." + c.band_names[b] + ".file";
			if (k.size() > suffix.size() && k.compare(k.size() - suffix.size(), suffix.size(), suffix) == 0) key = k;
		}
		if (!key.empty()) {
			std::string file = cfg.get(key);
			if (file == "internal") {
				std::string bandset = key.substr(0, key.find('.'));
				file = c.datadir + "/" + bandset + "." + c.band_names[b] + ".photoerr.txt";
			} else file = cfg.resolve(file);
			auto rows = load_columns(file, 2);
			for (auto &r : rows) { kx.push_back(r[0]); ky.push_back(r[1]); }
		}
		c.perr.koff[b + 1] = (int)kx.size();
	}
	c.perr_kx = kx;
	c.perr_ky = ky;
	c.perr_dev[0].upload(kx);
	c.perr_dev[1].upload(ky);
	c.perr.kx = (const double *)c.perr_dev[0].p;
	c.perr.ky = (const double *)c.perr_dev[1].p;
	// interval records: one fetch and no division per band and star
	std::vector<double> kq(4 * kx.size());
	if (gf_photoerr_records(c.perr.nbands, c.perr.koff, kx.data(), ky.data(), kq.data(), c.perr.gx0,
	                        c.perr.gscale))
		fail("photometricErrors: bad error-curve knots");
	c.perr_dev[2].upload(kq);
	c.perr.kq = kq.empty() ? nullptr : (const double *)c.perr_dev[2].p;
	c.has_perr = true;
}

std::string normalize_keyword(std::string s)
{
	std::string o;
	for (char ch : s) if (ch != ' ' && ch != '\t' && ch != '_') o += (char)tolower((unsigned char)ch);
	return o;
}

void open_catalog(gf_catalog &c, const std::string &cmdfile)
{
	Config cmd;
	if (!cmd.load(cmdfile)) fail("cannot read " + cmdfile);
	if (cmd.count("definitions")) {
		for (auto &f : split_ws(cmd.get("definitions"))) {
			Config d;
			if (!d.load(cmd.resolve(f))) fail("cannot read " + f);
			for (auto &kv : d.kv) if (kv.first != "module") c.globals.kv[kv.first] = kv.second;
		}
	}
	auto loadc = [&](const std::string &f) {
		Config x;
		x.globals = &c.globals;
		if (!x.load(cmd.resolve(f))) fail("cannot read " + cmd.resolve(f));
		return x;
	};
	c.skygen = loadc(cmd.get("input"));
	for (auto &f : split_ws(cmd.get("models"))) c.model_cfgs.push_back(loadc(f));
	for (auto &f : split_ws(cmd.get("footprints"))) c.foot_cfgs.push_back(loadc(f));
	for (auto &f : split_ws(cmd.get("modules", ""))) c.module_cfgs.push_back(loadc(f));
	// the output stages are module configurations too (galfast.cpp:227-233)
	for (auto &f : split_ws(cmd.get("output", ""))) c.module_cfgs.push_back(loadc(f));
	// nstars / maxstars belong to the skygen stage (skygen/os_skygen.cpp:816-820);
	// a command-file value is kept as a fallback
	c.nstars_forced = c.skygen.count("nstars") ? c.skygen.getd("nstars") : cmd.getd("nstars", 0.);
	c.maxstars = c.skygen.count("maxstars") ? c.skygen.getd("maxstars") : cmd.getd("maxstars", 1e8);
	if (c.foot_cfgs.empty()) fail("no footprint configuration given");
	for (auto &mc : c.model_cfgs) mc.globals = &c.globals;

	double dx;
	load_sky_pixelization(c, dx);
	load_footprints(c);
	pixelize(c, dx);
	load_extinction(c, c.skygen.get("extmaps", ""));
	for (auto &mc : c.model_cfgs) load_model(c, mc);
	// hint_absmag: each model's absolute-magnitude range restricted to its
	// LF's support (skygen/skyconfig_impl.h:258-263, skygen/model_lib.h:
	// 97-112), then reset_absmag (skygen/os_skygen.cpp:401-407)
	for (auto &m : c.models) {
		m->sp = c.sp;
		float X0 = m->lf.tc[0];
		float X1 = (float)((double)(m->lf.v.size() - 1) / m->lf.tc[1] + X0);
		float M0 = c.sp.M0, M1 = (float)(c.sp.M1 + 0.5 * c.sp.dM);
		bool narrowed = false;
		if (X0 > M0) { narrowed = true; M0 = X0; }
		if (X1 < M1) { narrowed = true; M1 = X1; }
		if (narrowed) {
			m->sp.M0 = M0;
			m->sp.nM = (int)round((M1 - M0) / c.sp.dM);
			m->sp.M1 = (float)(M1 - 0.5 * c.sp.dM);
			if (m->sp.nM < 1) fail("luminosity function of model " + m->name + " has no absolute magnitude in range");
		}
	}
	// modules in pipeline order (pipeline.h:75-89)
	for (auto &mc : c.module_cfgs) {
		std::string mod = normalize_keyword(mc.get("module"));
		if (mod == "feh") construct_FeH(c, mc);
		else if (mod == "gaussianfeh") construct_GaussianFeH(c, mc);
		else if (mod == "unresolvedmultiples") construct_multiples(c, mc);
		else if (mod == "bond2010") construct_Bond2010(c, mc);
		else if (mod == "photometry") construct_photometry(c, mc);
		else if (mod == "vel2pm") construct_vel2pm(c, mc);
		else if (mod == "gal2other") construct_gal2other(c, mc);
		else if (mod == "photometricerrors") { /* needs photometry; constructed below */ }
		else if (mod == "countsmap") c.countsmap = (int)(&mc - c.module_cfgs.data());
		else if (mod == "fitsout" || mod == "textout") { /* the writer's kind and path choose the output */ }
		else fail("unsupported module '" + mod + "'");
	}
	for (auto &mc : c.module_cfgs)
		if (normalize_keyword(mc.get("module")) == "photometricerrors") construct_photoerr(c, mc);
	if (c.has_feh && c.has_gfeh) fail("FeH and GaussianFeH are both configured");

	c.pix_dev.upload(c.pixels);
	if (hipStreamCreate(&c.stream) != hipSuccess || hipStreamCreate(&c.rows) != hipSuccess ||
	    hipStreamCreate(&c.side) != hipSuccess || hipStreamCreate(&c.emit) != hipSuccess)
		fail("stream creation failed");
	if (hipEventCreateWithFlags(&c.planned, hipEventDisableTiming) != hipSuccess ||
	    hipEventCreateWithFlags(&c.emitted[0], hipEventDisableTiming) != hipSuccess ||
	    hipEventCreateWithFlags(&c.emitted[1], hipEventDisableTiming) != hipSuccess)
		fail("event creation failed");
	(void)hipEventCreate(&c.ev[0]);
	(void)hipEventCreate(&c.ev[1]);
	for (auto &sl : c.slot)
		if (hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) fail("event creation failed");
	for (int b = 0; b < 2; b++)
		if (hipEventCreateWithFlags(&c.drawn[b], hipEventDisableTiming) != hipSuccess ||
		    hipEventCreateWithFlags(&c.pos_free[b], hipEventDisableTiming) != hipSuccess ||
		    hipEventCreateWithFlags(&c.side_done[b], hipEventDisableTiming) != hipSuccess ||
		    hipEventCreateWithFlags(&c.pos_free_side[b], hipEventDisableTiming) != hipSuccess)
			fail("event creation failed");
	for (int b = 0; b < 2; b++)   // the position buffers start free
		if (hipEventRecord(c.pos_free[b], c.rows) != hipSuccess ||
		    hipEventRecord(c.pos_free_side[b], c.side) != hipSuccess)
			fail("event record failed");
	c.vis_dev.alloc(2 * sizeof(unsigned long long));
	if (hipHostMalloc((void **)&c.vis_host, 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
		fail("pinned allocation failed");
	// persistent streams seeded from the pipeline seed (common/cux.cpp:404-435)
	std::string pf = c.datadir + "/safeprimes32.u32";
	std::ifstream f(pf, std::ios::binary);
	if (!f) fail("cannot read " + pf);
	std::vector<uint32_t> primes(70763);
	f.read((char *)primes.data(), primes.size() * 4);
	primes.resize(f.gcount() / 4);
	uint32_t s32 = gf_seed_from_pipeline_seed(c.seed);
	std::vector<uint32_t> st(3 * 65536);
	if (gf_mwc_seed_host(s32, primes.data(), (uint32_t)primes.size(), 65536, st.data()) != 65536)
		fail("seeding failed");
	if (gf_rng_alloc(&c.rng, 65536) || gf_rng_upload(&c.rng, st.data(), c.stream)) fail("stream upload failed");
	c.jtab.alloc((size_t)GF_JUMP_BITS * 65536 * sizeof(uint64_t));
	if (gf::mwc_jump_table(&c.rng, (uint64_t *)c.jtab.p, c.stream) || hipStreamSynchronize(c.stream) != hipSuccess)
		fail("jump tables failed");
	c.rng_state0 = st;

	// host mirrors of the device parameter blocks
	c.models_h.clear();
	c.models_sp_h.clear();
	for (auto &m : c.models) {
		gf_model h = m->m;
		h.lf.data = m->lf.v.data();
		c.models_h.push_back(h);
		c.models_sp_h.push_back(m->sp);
	}
	if (c.has_phot) {
		c.phot_h = c.phot;
		size_t per = (size_t)c.phot_tables.nFeH * c.phot_tables.nMr * 4;
		for (int t = 0; t < 4; t++) {
			c.phot_h.color[t] = t < c.phot_tables.ntex ? c.phot_tables.color.data() + t * per : nullptr;
			c.phot_h.cflags[t] = t < c.phot_tables.ntex ? c.phot_tables.flags.data() + t * per : nullptr;
		}
		c.phot_h.texels = nullptr;
		c.phot_h.flagbits = nullptr;
	}
	if (c.has_mult) {
		c.mult_h = c.mult;
		c.mult_h.secProb.data = c.mult_sec.v.data();
		c.mult_h.cumLF.data = c.mult_cum.v.data();
		c.mult_h.invCumLF.data = c.mult_inv.v.data();
	}
	if (c.has_perr) {
		c.perr_h = c.perr;
		c.perr_h.kx = c.perr_kx.data();
		c.perr_h.ky = c.perr_ky.data();
		c.perr_h.kq = nullptr;   // the host view holds the knots only
	}
}

// Allocate the device catalog table for `cap` rows.
void ensure_table(gf_catalog &c, gf_catalog::Slot &sl, uint64_t cap)
{
	if (sl.cap >= cap && sl.table_dev.p) return;
	struct ColDef { gf_col gf_table::*m; uint32_t elem, width; };
	int nb = c.nbands > 0 ? c.nbands : 1;
	ColDef defs[] = {
		{&gf_table::lb, 8, 2}, {&gf_table::projIdx, 4, 1}, {&gf_table::projXY, 4, 2}, {&gf_table::DM, 4, 1},
		{&gf_table::absmag, 4, 1}, {&gf_table::XYZ, 4, 3}, {&gf_table::Am, 4, 1}, {&gf_table::AmInf, 4, 1},
		{&gf_table::comp, 4, 1}, {&gf_table::hidden, 4, 1}, {&gf_table::FeH, 4, 1}, {&gf_table::FeH_comp, 4, 1},
		{&gf_table::absmagSys, 4, (uint32_t)std::max(c.nabsmag, 1)}, {&gf_table::absmagSysNcomp, 4, 1},
		{&gf_table::vcyl, 4, 3}, {&gf_table::mags, 4, (uint32_t)nb}, {&gf_table::photoFlags, 4, 1},
		{&gf_table::obsMags, 4, (uint32_t)nb}, {&gf_table::pm, 4, 3}, {&gf_table::radec, 8, 2},
	};
	size_t total = 0;
	for (auto &d : defs) {
		size_t pitch = ((cap * d.elem + 127) / 128) * 128;
		total += pitch * d.width;
	}
	sl.table_dev.alloc(total);
	char *p = (char *)sl.table_dev.p;
	for (auto &d : defs) {
		uint32_t pitch = (uint32_t)(((cap * d.elem + 127) / 128) * 128);
		gf_col col{p, pitch, d.width, d.elem, 0};
		sl.table.*(d.m) = col;
		p += (size_t)pitch * d.width;
	}
	sl.cap = cap;
}

double now_s()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The catalog's module chain (device parameter blocks).
gf_chain chain_of(gf_catalog &c)
{
	gf_chain ch{};
	if (c.has_feh) ch.feh = &c.feh;
	if (c.has_gfeh) { ch.gfeh = &c.gfeh; ch.gfeh_comps = c.gfeh_comps; }
	if (c.has_mult) { ch.mult = &c.mult; ch.mult_comps = c.mult_comps; }
	if (c.has_bond) ch.bond = &c.bond;
	if (c.has_phot) { ch.phot = &c.phot; ch.phot_comps = c.phot_comps; }
	if (c.has_perr) ch.perr = &c.perr;
	if (c.has_v2pm) ch.v2pm = &c.v2pm;
	ch.gal2other_coordsys = c.gal2other;
	return ch;
}

gf_skygen_cols skygen_cols_of(const gf_table &t)
{
	return gf_skygen_cols{t.lb, t.projIdx, t.projXY, t.DM, t.absmag, t.XYZ, t.Am, t.AmInf, t.comp, t.hidden};
}

// The reference's draw, model by model (skygenHost<T>::drawSources,
// skygen/skyconfig_impl.h:434-545, then os_clipper and the modules): the
// diagonal crawl of kernel<1> with sp.nthreads threads on streams 0 .. nthreads-1,
// the batch sorted with star_comp and truncated at its first hidden row, the
// clipper, then the module chain over the visible rows, the stream of row r
// being r / 256.  One batch per model: the table grows to hold the model's
// stars.  (The reference suspends the crawl when its table fills and
// resumes it for the next batch; which stars make a batch then depends on
// the order of its threads' atomics.)
int run_catalog_reference(gf_catalog &c, gf_chunk_fn fn, void *user, gf_run_stats *stats)
{
	gf_run_stats s{};
	double t0 = now_s();
	hipStream_t A = c.stream;
	const gf_chain ch = chain_of(c);
	const uint64_t sub = 65536ull * 256ull;   // rows the 65,536 streams cover per chain call
	gf_catalog::Slot &drawn = c.slot[0], &sorted = c.slot[1];
	for (auto &mp : c.models) {
		Model &m = *mp;
		if (!m.wpk.p) return GF_E_STATE;   // the count pass comes first, as in os_skygen::run
		const gf_extinction *ext = c.has_ext ? &c.ext : nullptr;
		const gf_pencil_beam *pix = (const gf_pencil_beam *)c.pix_dev.p;
		// batches of at most chunk_rows stars, the output table's capacity
		// (drawSources, skyconfig_impl.h:434-534): each is sorted, truncated
		// at its first hidden row, clipped and run through the modules
		gf_draw_state *ds = nullptr;
		uint64_t total = 0;
		int rc = gf_skygen_draw_begin(&m.sp, &m.m, pix, ext, nullptr, &c.rng, &ds, &total, A);
		if (rc) return rc;
		std::unique_ptr<gf_draw_state, void (*)(gf_draw_state *)> hold(ds, gf_skygen_draw_end);
		const uint64_t cap = std::max<uint64_t>(1, std::min<uint64_t>(c.chunk_rows, total));
		ensure_table(c, drawn, cap);
		ensure_table(c, sorted, cap);
		gf_skygen_cols dc = skygen_cols_of(drawn.table);
		gf_skygen_cols sc = skygen_cols_of(sorted.table);
		s.generated += total;
		s.expected += m.counts.nstarsExpectedToGenerate;
		for (;;) {
			uint64_t n = 0;
			rc = gf_skygen_draw_next(ds, &dc, cap, &n, A);
			if (rc) return rc;
			if (n == 0) break;
			s.n_emit++;
			uint64_t nv = 0;
			rc = gf_skygen_order(&dc, n, &sc, &nv, A);
			if (rc) return rc;
			rc = gf_clip_footprint(&c.sp, &c.foot.f, sorted.table.projIdx, sorted.table.projXY, sorted.table.hidden,
			                       nv, A);
			if (rc) return rc;
			for (uint64_t b0 = 0; b0 < nv; b0 += sub) {
				gf_rows ks{(uint32_t)b0, 256, (uint32_t)std::min(nv, b0 + sub), 0};
				rc = gf::module_chain(&ks, &ch, &c.rng, &sorted.table, A, nullptr, nullptr);
				if (rc) return rc;
				s.n_chain++;
			}
			s.rows_chain += nv;
			uint64_t vis = 0;
			if (nv) {
				rc = gf_count_visible(sorted.table.hidden, nv, &vis, A);
				if (rc) return rc;
			}
			s.visible += vis;
			if (fn && nv) {
				rc = fn(user, &sorted.table, nv, m.m.comp);
				if (rc) return rc;
			}
		}
	}
	GF_HIP(hipStreamSynchronize(A));
	s.t_total = now_s() - t0;
	if (stats) *stats = s;
	return 0;
}

// Events of one chunk: 3 around emission (on the catalog stream), then per
// chain call 2 around each launch of the module chain (on the stream that
// runs it).
constexpr int EV_EMIT = 3, EV_CALL = 2 * GF_CHAIN_LAUNCHES;

void ensure_events(gf_catalog::Slot &sl, int ncalls)
{
	size_t want = EV_EMIT + (size_t)EV_CALL * ncalls;
	while (sl.ev.size() < want) {
		hipEvent_t e;
		if (hipEventCreate(&e) != hipSuccess) fail("event creation failed");
		sl.ev.push_back(e);
	}
}

double elapsed_s(hipEvent_t a, hipEvent_t b)
{
	float ms = 0.f;
	(void)hipEventElapsedTime(&ms, a, b);
	return ms * 1e-3;
}

// Waits for the chunk in slot si, adds its visible rows and launch times to
// the statistics and hands its columns to the callback.
int deliver(gf_catalog &c, int si, gf_chunk_fn fn, void *user, gf_run_stats &s)
{
	gf_catalog::Slot &sl = c.slot[si];
	GF_HIP(hipEventSynchronize(sl.done));
	sl.pending = false;
	s.visible += c.vis_host[si];
	const hipEvent_t *e = sl.ev.data();
	double a = elapsed_s(e[0], e[1]), b = elapsed_s(e[1], e[2]);   // the tiles' units, then the stars
	s.t_emit_units += a;
	s.t_emit_stars += b;
	s.t_emit += a + b;
	for (int j = 0; j < sl.nsub; j++) {
		const hipEvent_t *q = e + EV_EMIT + EV_CALL * j;
		for (int g = 0; g < GF_CHAIN_LAUNCHES; g++) {
			double r = elapsed_s(q[2 * g], q[2 * g + 1]);
			s.t_stage[g] += r;
			s.t_chain += r;
		}
	}
	if (fn) return fn(user, &sl.table, sl.n, sl.comp);
	return 0;
}

// The catalog in output chunks.  Stream order: emission and the draw passes
// advance the RNG streams and run in order on c.stream; the row passes only
// read the streams' multipliers and run on c.rows as soon as the draw pass
// of their rows is done, so that they overlap the next chunk's emission and
// draws (the draw-type kernels run one lane per stream and leave most of the
// device idle).  The kinematics and astrometry passes depend on neither
// metallicity nor photometry and run on c.side beside them (memory-latency
// bound texture lookups beside double-precision arithmetic).  Two chunk slots: chunk k is handed to the callback after
// chunk k+1 has been queued, and before chunk k+2 overwrites its slot.
int run_catalog(gf_catalog &c, gf_chunk_fn fn, void *user, gf_run_stats *stats)
{
	gf_run_stats s{};
	double t0 = now_s();
	hipStream_t A = c.stream, B = c.overlap ? c.rows : c.stream, C = c.overlap ? c.side : c.stream;
	hipStream_t E = c.overlap ? c.emit : c.stream;
	const gf_chain ch = chain_of(c);
	const uint64_t sub = 65536ull * 256ull;   // rows the 65,536 streams cover per chain call
	const size_t pos_bytes = gf::chain_positions_bytes(&ch, (uint32_t)sub);
	for (auto &pb : c.pos)
		if (pos_bytes && pb.bytes < pos_bytes) pb.alloc(pos_bytes);
	unsigned long long *vis_dev = (unsigned long long *)c.vis_dev.p;
	uint32_t pb = 0;      // position buffer of the next chain call
	uint32_t last = 0;    // position buffer of the last chain call queued
	uint64_t k = 0;       // chunks queued
	int rc;
	for (auto &mp : c.models) {
		Model &m = *mp;
		if (!m.wpk.p) return GF_E_STATE;
		uint64_t total = 0;
		// a model's plan is rewritten only by its next run, which starts
		// after every stream of this one has finished
		GF_HIP(hipEventRecord(c.ev[0], A));
		rc = gf::skygen_plan(&m.sp, &m.m, (const gf_pencil_beam *)c.pix_dev.p, c.has_ext ? &c.ext : nullptr,
		                     (const double *)m.wpk.p, &c.rng, &m.plan, &total, A);
		if (rc) return rc;
		gf_draw_plan *plan = m.plan;
		// the emission reads the plan and the streams as they were at the
		// plan (kept in the plan), never the streams the draw passes advance
		GF_HIP(hipEventRecord(c.planned, A));
		GF_HIP(hipStreamWaitEvent(E, c.planned, 0));
		GF_HIP(hipEventRecord(c.ev[1], A));
		GF_HIP(hipEventSynchronize(c.ev[1]));
		s.t_plan += elapsed_s(c.ev[0], c.ev[1]);
		s.generated += total;
		s.expected += m.counts.nstarsExpectedToGenerate;
		if (!total) continue;
		const uint64_t chunk = std::min<uint64_t>(c.chunk_rows, total);
		GF_CHECK(chunk < (1ull << 32), GF_E_ARG);
		const uint64_t nchunks = (total + chunk - 1) / chunk;
		GF_CHECK(nchunks < (1ull << 32), GF_E_ARG);
		std::vector<uint64_t> units(nchunks);
		rc = gf::skygen_chunk_units(plan, chunk, (uint32_t)nchunks, units.data(), A);
		if (rc) return rc;
		for (uint64_t q = 0; q < nchunks; q++, k++) {
			const int si = (int)(k & 1);
			gf_catalog::Slot &sl = c.slot[si];
			if (sl.pending) {
				rc = deliver(c, si, fn, user, s);
				if (rc) return rc;
			}
			const uint64_t r0 = q * chunk, r1 = std::min(total, r0 + chunk), n = r1 - r0;
			const int nsub = (int)((n + sub - 1) / sub);
			ensure_table(c, sl, chunk);
			ensure_events(sl, nsub);
			const hipEvent_t *e = sl.ev.data();
			gf_skygen_cols sc{sl.table.lb, sl.table.projIdx, sl.table.projXY, sl.table.DM, sl.table.absmag,
			                  sl.table.XYZ, sl.table.Am, sl.table.AmInf, sl.table.comp, sl.table.hidden};
			rc = gf::skygen_emit(plan, r0, r1, &c.foot.f, &c.rng, &sc, E, sl.ev.data(), nullptr, units[q]);
			if (rc) return rc;
			GF_HIP(hipEventRecord(c.emitted[si], E));
			GF_HIP(hipStreamWaitEvent(A, c.emitted[si], 0));
			s.n_emit++;
			bool counted = false;   // the draw passes count the chunk's visible rows
			for (int j = 0; j < nsub; j++, pb ^= 1) {
				const uint64_t b0 = (uint64_t)j * sub;
				gf_rows ks{(uint32_t)b0, 256, (uint32_t)std::min(n, b0 + sub), 0};
				gf_chain_launch L;
				rc = gf::chain_setup(&ks, &ch, &c.rng, &sl.table, pos_bytes ? (uint2 *)c.pos[pb].p : nullptr, L);
				if (rc) return rc;
				if (j == 0 && gf::chain_counts_visible(L)) {
					counted = true;
					GF_HIP(hipMemsetAsync(vis_dev + si, 0, sizeof(unsigned long long), A));
				}
				if (counted) L.vis = vis_dev + si;
				const hipEvent_t *qe = e + EV_EMIT + EV_CALL * j;
				// the draw pass overwrites the planes the row passes two calls back read
				GF_HIP(hipStreamWaitEvent(A, c.pos_free[pb], 0));
				GF_HIP(hipStreamWaitEvent(A, c.pos_free_side[pb], 0));
				auto launch = [&](int g, hipStream_t st) -> int {
					GF_HIP(hipEventRecord(qe[2 * g], st));
					int r = gf::chain_launch(L, &c.rng, g, st);
					if (r) return r;
					GF_HIP(hipEventRecord(qe[2 * g + 1], st));
					return 0;
				};
				if ((rc = launch(0, A))) return rc;
				GF_HIP(hipEventRecord(c.drawn[pb], A));
				GF_HIP(hipStreamWaitEvent(B, c.drawn[pb], 0));
				GF_HIP(hipStreamWaitEvent(C, c.drawn[pb], 0));
				if ((rc = launch(1 + GF_ROWS_BOND, C))) return rc;
				GF_HIP(hipEventRecord(c.pos_free_side[pb], C));
				if ((rc = launch(1 + GF_ROWS_ASTRO, C))) return rc;
				GF_HIP(hipEventRecord(c.side_done[pb], C));
				if ((rc = launch(1 + GF_ROWS_FEH, B))) return rc;
				if ((rc = launch(1 + GF_ROWS_PHOT, B))) return rc;
				if ((rc = launch(1 + GF_ROWS_PERR, B))) return rc;
				GF_HIP(hipEventRecord(c.pos_free[pb], B));
				last = pb;
				s.n_chain++;
			}
			// the chunk is complete when the astrometry passes are
			GF_HIP(hipStreamWaitEvent(B, c.side_done[last], 0));
			s.rows_chain += n;
			if (!counted) {
				GF_HIP(hipMemsetAsync(vis_dev + si, 0, sizeof(unsigned long long), B));
				rc = gf::count_visible(sl.table.hidden, n, vis_dev + si, B);
				if (rc) return rc;
			}
			GF_HIP(hipMemcpyAsync(c.vis_host + si, vis_dev + si, sizeof(unsigned long long), hipMemcpyDeviceToHost,
			                      B));
			GF_HIP(hipEventRecord(sl.done, B));
			// the next chunk's emission rewrites the rows' inputs only in the
			// other slot; the streams are not touched by the row passes
			sl.pending = true;
			sl.n = n;
			sl.comp = m.m.comp;
			sl.nsub = nsub;
		}
	}
	for (uint64_t q = 0; q < 2; q++, k++) {
		const int si = (int)(k & 1);
		if (c.slot[si].pending) {
			rc = deliver(c, si, fn, user, s);
			if (rc) return rc;
		}
	}
	GF_HIP(hipStreamSynchronize(A));
	GF_HIP(hipStreamSynchronize(B));
	GF_HIP(hipStreamSynchronize(C));
	GF_HIP(hipStreamSynchronize(E));
	s.t_total = now_s() - t0;
	if (stats) *stats = s;
	return 0;
}

} // namespace

extern "C" {

const char *gf_last_error(void) { return g_last_error.c_str(); }

int gf_catalog_open(const char *cmd_conf, uint32_t seed, const char *datadir, gf_catalog **out)
{
	GF_CHECK(cmd_conf && datadir && out, GF_E_ARG);
	auto c = std::make_unique<gf_catalog>();
	c->seed = seed;
	c->datadir = datadir;
	try {
		open_catalog(*c, cmd_conf);
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
	*out = c.release();
	return 0;
}

int gf_catalog_set(gf_catalog *c, const char *key, double value)
{
	GF_CHECK(c && key, GF_E_ARG);
	std::string k = key;
	if (k == "chunk_rows") {
		if (!(value >= 1.0 && value <= (double)GF_MAX_TABLE_ROWS)) return GF_E_ARG;
		c->chunk_rows = (uint64_t)value;
	}
	else if (k == "overlap") c->overlap = value != 0.;
	else if (k == "draw") {
		if (value != 0. && value != 1.) return GF_E_ARG;
		c->draw_mode = (int)value;
	}
	else if (k == "acc_mode") c->acc_mode = (int)value;
	else if (k == "nstars") c->nstars_forced = value;
	else if (k == "maxstars") c->maxstars = value;
	else return GF_E_ARG;
	return 0;
}

int gf_catalog_component_ids(const gf_catalog *c, uint32_t *ids, int cap)
{
	GF_CHECK(c && (ids || cap == 0), GF_E_ARG);
	const int n = (int)c->cm.seq2comp.size();
	for (int i = 0; i < n && i < cap; i++) ids[i] = c->cm.seq2comp[i];
	return n;
}

int gf_catalog_info(const gf_catalog *c, gf_catalog_info_t *info)
{
	GF_CHECK(c && info, GF_E_ARG);
	info->npixels = (uint64_t)c->sp.npixels;
	info->nm = c->sp.nm;
	info->nM = c->sp.nM;
	info->nmodels = (int)c->models.size();
	info->nbands = c->nbands;
	info->nabsmag = c->nabsmag;
	info->footprint_covered_sr = 0;
	for (auto &p : c->pixels) info->footprint_covered_sr += (double)p.coveredFraction * (double)p.dA;
	return 0;
}

int gf_catalog_count(gf_catalog *c, gf_count_result *per_model, double *expected)
{
	GF_CHECK(c, GF_E_ARG);
	double tot = 0;
	const size_t np = (size_t)c->sp.npixels;
	c->counted = -1;
	if (c->per_pixel.bytes < np * sizeof(double)) c->per_pixel.alloc(np * sizeof(double));
	std::vector<double> pp(np), den(np, 0.0);
	// every count starts from the model's own density (norm 1), so that a
	// count is idempotent and a changed nstars scales the raw total
	for (auto &m : c->models) m->m.norm = 1.f;
	auto count_all = [&](bool check) -> int {
		tot = 0;
		for (size_t i = 0; i < c->models.size(); i++) {
			Model &m = *c->models[i];
			const uint64_t K = (uint64_t)(m.sp.nm + m.sp.nM - 1);
			// (allocated also for a footprint without pixels: a run checks
			// for it to know that the count came first)
			if (!m.wpk.p || m.wpk.bytes < np * K * sizeof(double)) m.wpk.alloc(np * K * sizeof(double));
			int rc = gf_skygen_count(&m.sp, &m.m, (const gf_pencil_beam *)c->pix_dev.p,
			                         c->has_ext ? &c->ext : nullptr, c->acc_mode, &m.counts,
			                         check ? (double *)c->per_pixel.p : nullptr, (double *)m.wpk.p, c->stream);
			if (rc) return rc;
			if (check && np) {
				GF_HIP(hipMemcpy(pp.data(), c->per_pixel.p, np * sizeof(double), hipMemcpyDeviceToHost));
				// the second summation of skygenHost::integrateCounts: the
				// per-pixel densities added in pixel order must agree with
				// the total to 1e-5 (skygen/skyconfig_impl.h:387-402)
				double total = 0;
				for (size_t p = 0; p < np; p++) { total += pp[p]; den[p] += pp[p]; }
				if (m.counts.nstarsExpected > 0 && !(fabs(total / m.counts.nstarsExpected - 1.) < 1e-5)) {
					g_last_error = "The expected number of stars, computed in two different ways is inconsistent "
					               "(nstarsExpected=" + std::to_string(m.counts.nstarsExpected) +
					               ", total=" + std::to_string(total) + ")";
					return GF_E_STATE;
				}
			}
			if (per_model) per_model[i] = m.counts;
			tot += m.counts.nstarsExpected;
		}
		return 0;
	};
	int rc = count_all(true);
	if (rc) return rc;
	// nstars renormalisation (skygen/os_skygen.cpp:883-901): the densities
	// are scaled by nstars / (the model total at norm 1)
	if (c->nstars_forced > 0 && tot > 0) {
		double norm = c->nstars_forced / tot;
		for (auto &m : c->models) m->m.norm = (float)norm;
		rc = count_all(false);
		if (rc) return rc;
	}
	c->counted = tot;
	if (expected) *expected = tot;
	return 0;
}

int gf_catalog_run(gf_catalog *c, gf_chunk_fn fn, void *user, gf_run_stats *stats)
{
	GF_CHECK(c, GF_E_ARG);
	int rc;
	// the expected number of stars may not exceed maxstars (skygen/os_skygen.cpp:908-917)
	if (c->counted >= 0 && c->maxstars != 0 && c->counted > c->maxstars) {
		g_last_error = "The expected number of generated stars (" + std::to_string(c->counted) + ") "
		               "exceeds the configuration limit (" + std::to_string(c->maxstars) + "). Increase "
		               "maxstars or set to zero for unlimited.";
		return GF_E_CONFIG;
	}
	try {
		rc = c->draw_mode == 1 ? run_catalog_reference(*c, fn, user, stats) : run_catalog(*c, fn, user, stats);
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		rc = GF_E_CONFIG;
	}
	if (rc) {
		// drop the chunks in flight
		(void)hipStreamSynchronize(c->stream);
		(void)hipStreamSynchronize(c->rows);
		(void)hipStreamSynchronize(c->side);
		for (auto &sl : c->slot) sl.pending = false;
	}
	return rc;
}

} // extern "C"

namespace gfh {

// The columns the reference's output stages write for this job, in its
// order (otable::getSortedColumnsForOutput, otable.cpp:502-540: ranked by
// the first of lb, radec, XYZ, DM, absmag, comp, FeH, vcyl, pmlb, pmradec,
// Am, AmInf that the column or one of its aliases is named, the others after
// them by case-insensitive name), with the built-in and module column
// definitions (otable.cpp:715-745; os_skygen.cpp:841-858;
// photometry_host.cpp:108-122; photometricErrors_host.cpp:156-177;
// unresolvedMultiples_host.cpp:62-77; GaussianFeH_host.cpp:46-47;
// vel2pm_host.cpp:80-83; gal2other_host.cpp:74).  Hidden columns (hidden,
// projIdx, projXY) are not written.
void output_schema(const gf_catalog &c, std::vector<OutColumn> &out)
{
	struct Col {
		OutColumn oc;
		std::vector<std::string> aliases;
	};
	std::vector<Col> cols;
	auto add = [&](std::string name, gf_col gf_table::*m, uint32_t w, char type, std::string fmt, std::string def,
	               std::vector<std::string> aliases = {}, bool comp = false) {
		Col k;
		k.oc.name = name;
		k.oc.member = m;
		k.oc.width = w;
		k.oc.type = type;
		k.oc.fmt = fmt;
		k.oc.def = def;
		k.oc.map_comp = comp;
		k.aliases = aliases;
		cols.push_back(k);
	};
	auto fields = [](const std::vector<std::string> &names, const std::string &prefix) {
		std::string f = "fieldNames=";
		for (size_t i = 0; i < names.size(); i++) f += (i ? "," : "") + std::to_string(i) + ":" + prefix + names[i];
		return f + ";";
	};
	const std::string absmag = "abs" + c.band;
	add("lb", &gf_table::lb, 2, 'D', "% 13.8f", "lb[2]");
	add("XYZ", &gf_table::XYZ, 3, 'E', "% 10.2f", "XYZ[3]");
	add("DM", &gf_table::DM, 1, 'E', "% 7.3f", "DM");
	if (c.has_mult)
		add(absmag, &gf_table::absmag, 1, 'E', "% 7.3f", absmag + "{type=float;alias=absmag;band=" + c.band + ";}",
		    {"absmag"});
	else
		add(absmag, &gf_table::absmag, 1, 'E', "% 7.3f",
		    absmag + "{type=float;alias=M1;alias=absmag;band=" + c.band + ";}", {"M1", "absmag"});
	add("comp", &gf_table::comp, 1, 'J', "%3d", "comp", {}, true);
	add("Am", &gf_table::Am, 1, 'E', "%6.3f", "Am");
	add("AmInf", &gf_table::AmInf, 1, 'E', "%6.3f", "AmInf");
	if (c.has_feh || c.has_gfeh) add("FeH", &gf_table::FeH, 1, 'E', "% 6.3f", "FeH");
	if (c.has_gfeh) add("FeH_comp", &gf_table::FeH_comp, 1, 'J', "%3d", "FeH_comp{type=int;fmt=%3d;}");
	if (c.has_bond) add("vcyl", &gf_table::vcyl, 3, 'E', "% 7.1f", "vcyl[3]");
	if (c.has_v2pm) {
		const char *pm = c.v2pm.coordsys == 1 ? "pmradec" : "pmlb";
		add(pm, &gf_table::pm, 3, 'E', "% 8.2f", std::string(pm) + "[3]");
	}
	if (c.gal2other == 1) add("radec", &gf_table::radec, 2, 'D', "% 13.8f", "radec[2]");
	if (c.has_phot) {
		const std::string n = std::to_string(c.nbands);
		add(c.bandset, &gf_table::mags, (uint32_t)c.nbands, 'E', "% 7.3f",
		    c.bandset + "[" + n + "]{type=float;class=magnitude;" + fields(c.band_names, "") + "}");
		add(c.bandset + "PhotoFlags", &gf_table::photoFlags, 1, 'J', "% 4d",
		    c.bandset + "PhotoFlags{type=int;class=flags;}");
		if (c.has_perr)
			add("obs" + c.bandset, &gf_table::obsMags, (uint32_t)c.nbands, 'E', "% 7.3f",
			    "obs" + c.bandset + "[" + n + "]{type=float;class=magnitude;" + fields(c.band_names, "obs") + "}");
	}
	if (c.has_mult) {
		add(absmag + "Sys", &gf_table::absmagSys, 2, 'E', "% 7.3f",
		    absmag + "Sys[2]{type=float;class=magnitude;fieldNames=0:M1,1:M2;alias=absmagSys;band=" + c.band + ";}",
		    {"absmagSys"});
		add(absmag + "SysNcomp", &gf_table::absmagSysNcomp, 1, 'J', "%1d",
		    absmag + "SysNcomp{type=int;fmt=%1d;}");
	}
	static const char *ranked[] = {"lb", "radec", "XYZ", "DM", "absmag", "comp", "FeH", "vcyl", "pmlb", "pmradec",
	                               "Am", "AmInf"};
	auto rank = [&](const Col &k) {
		int r = 1000;
		for (int i = 0; i < 12; i++) {
			if (k.oc.name == ranked[i]) r = std::min(r, i + 1);
			for (auto &a : k.aliases)
				if (a == ranked[i]) r = std::min(r, i + 1);
		}
		return r;
	};
	std::stable_sort(cols.begin(), cols.end(), [&](const Col &a, const Col &b) {
		int ra = rank(a), rb = rank(b);
		return ra == rb ? strcasecmp(a.oc.name.c_str(), b.oc.name.c_str()) < 0 : ra < rb;
	});
	out.clear();
	for (auto &k : cols) out.push_back(k.oc);
}

const std::vector<uint32_t> &component_ids(const gf_catalog &c) { return c.cm.seq2comp; }

const Config *countsmap_config(const gf_catalog &c)
{
	return c.countsmap >= 0 ? &c.module_cfgs[c.countsmap] : nullptr;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

} // namespace gfh

extern "C" {

int gf_catalog_host_view_get(const gf_catalog *c, gf_catalog_host_view *v)
{
	GF_CHECK(c && v, GF_E_ARG);
	*v = gf_catalog_host_view{};
	v->sp = c->sp;
	v->pixels = c->pixels.data();
	v->nmodels = (int)c->models_h.size();
	v->nbands = c->nbands;
	v->models = c->models_h.data();
	v->model_sp = c->models_sp_h.data();
	v->foot = c->foot.f;
	gf_chain &ch = v->chain;
	if (c->has_feh) ch.feh = &c->feh;
	if (c->has_gfeh) { ch.gfeh = &c->gfeh; ch.gfeh_comps = c->gfeh_comps; }
	if (c->has_mult) { ch.mult = &c->mult_h; ch.mult_comps = c->mult_comps; }
	if (c->has_bond) ch.bond = &c->bond;
	if (c->has_phot) { ch.phot = &c->phot_h; ch.phot_comps = c->phot_comps; }
	if (c->has_perr) ch.perr = &c->perr_h;
	if (c->has_v2pm) ch.v2pm = &c->v2pm;
	ch.gal2other_coordsys = c->gal2other;
	v->nabsmag = c->nabsmag;
	v->nstreams = 65536;
	v->rng_state0 = c->rng_state0.data();
	if (c->has_ext) v->ext = c->ext_h;
	return 0;
}

int gf_beam_pixels_host(double l, double b, double radius, double dx, gf_pencil_beam *out, int cap)
{
	GF_CHECK(radius > 0 && dx > 0 && cap >= 0 && (out || cap == 0), GF_E_ARG);
	try {
		gf_footprint fp{};
		fp.nregion = 1;
		gf_foot_region &f = fp.r[0];
		f.kind = 2;
		const double d2r = M_PI / 180.0;
		f.cl = cos(l * d2r); f.sl = sin(l * d2r); f.cb = cos(b * d2r); f.sb = sin(b * d2r);
		f.cosr = cos(radius * d2r);
		std::vector<gf_pencil_beam> v = pixelize_region(fp, dx * d2r);
		for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = v[i];
		return (int)v.size();
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
}

int gf_rect_pixels_host(double l0, double l1, double b0, double b1, double dx, gf_pencil_beam *out, int cap)
{
	GF_CHECK(l1 > l0 && l1 - l0 <= 360. && b0 >= -90. && b0 < b1 && b1 <= 90. && dx > 0 && cap >= 0 &&
	         (out || cap == 0), GF_E_ARG);
	try {
		gf_footprint fp{};
		fp.nregion = 1;
		fp.r[0] = rect_region(l0, l1, b0, b1);
		std::vector<gf_pencil_beam> v = pixelize_region(fp, dx * M_PI / 180.0);
		for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = v[i];
		return (int)v.size();
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
}

int gf_transform_host(const char *center, const char *orientation, double Rg, double z0, float T[3], float M[9])
{
	GF_CHECK(center && orientation && T && M, GF_E_ARG);
	try {
		Config cfg;
		char buf[64];
		snprintf(buf, sizeof buf, "%.17g", Rg);
		cfg.kv["Rg"] = buf;
		snprintf(buf, sizeof buf, "%.17g", z0);
		cfg.kv["z0"] = buf;
		float MM[3][3];
		load_transform(T, MM, center, orientation, cfg);
		for (int i = 0; i < 3; i++)
			for (int j = 0; j < 3; j++) M[3 * i + j] = MM[i][j];
		return 0;
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
}

int gf_lf_texture_host(const double *X, const double *Y, int n, int nsamp, float *tex, float tc[2])
{
	GF_CHECK(X && Y && tex && tc && n >= 1 && nsamp >= 2, GF_E_ARG);
	Tex1D t = resample_clip_1d(std::vector<double>(X, X + n), std::vector<double>(Y, Y + n), 0.f, nsamp);
	std::copy(t.v.begin(), t.v.end(), tex);
	tc[0] = t.tc[0];
	tc[1] = t.tc[1];
	return 0;
}

int gf_photometry_tables_host(const double *rows, int nrows, int ncolors, double dFeH, double dMr,
                              float *colors, float *flags, size_t cap_floats, int *nFeH, int *nMr, float tc[4])
{
	GF_CHECK(rows && nrows > 0 && ncolors >= 1 && ncolors <= 16 && colors && flags && nFeH && nMr && tc, GF_E_ARG);
	std::vector<std::vector<double>> rr;
	for (int r = 0; r < nrows; r++) rr.emplace_back(rows + (size_t)r * (2 + ncolors), rows + (size_t)(r + 1) * (2 + ncolors));
	PhotTables pt;
	try {
		pt = build_photometry_tables(rr, ncolors, dFeH, dMr);
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
	if (pt.color.size() > cap_floats) return GF_E_CAPACITY;
	std::copy(pt.color.begin(), pt.color.end(), colors);
	std::copy(pt.flags.begin(), pt.flags.end(), flags);
	*nFeH = pt.nFeH;
	*nMr = pt.nMr;
	tc[0] = pt.tcFeH[0]; tc[1] = pt.tcFeH[1]; tc[2] = pt.tcMr[0]; tc[3] = pt.tcMr[1];
	return pt.ntex;
}

void gf_catalog_close(gf_catalog *c)
{
	if (!c) return;
	if (c->stream) (void)hipStreamDestroy(c->stream);
	delete c;
}

} // extern "C"

// ===========================================================================
// Request server: continuous batching of footprint requests.
//
// Submitted requests are prepared on a host thread (pixelization of the
// beam, seeding of the request's own stream table) and queued.  Each worker
// owns a HIP stream and device buffers; it takes every ready request up to
// max_batch, runs count -> plan -> emit -> module chain for the whole batch
// with one launch per stage (serve.hip), hands each request its rows and
// takes the next batch.  With two workers one batch's host work overlaps the
// other's kernels.  A request's streams, pixels and binning are its own, so
// its rows do not depend on the batch it travels in.
// ===========================================================================
#include "../serve.h"
#include <condition_variable>
#include <deque>
#include <thread>

namespace {

double server_clock() { return now_s(); }

struct SReq {
	gf_request r;
	double t_submit = 0, t_start = 0;
	gf_footprint foot{};
	std::vector<gf_pencil_beam> pix;
	gf_skygen_params sp{};
	uint32_t S = 0;
	uint32_t seed32 = 0;           // the seed of its S streams (seeded on the device with its batch)
	uint64_t nunits_model = 0;     // npix * K of the job's binning: the size of the stream table
	std::vector<gf_skygen_params> msp;   // per model: the binning with its absolute-magnitude range
	std::vector<uint64_t> munits;        // per model: npix * K of its binning
	uint64_t nunits = 0;                 // all models' units
	uint64_t nI = 0;                     // stream indices all models' units take
	double expected = 0;
	double counted = -1;           // expected rows from a count of an earlier batch, -1 before
	std::vector<float> extmin;     // extinction minima per pixel and DM slice, with extinction maps
};

struct Worker {
	hipStream_t st = nullptr;
	DevBuf W, counts, offs, pixsum, segexp, pix, state, seeds, segs, segall, reqs, table, pos, zu, vis, extbeam, roff;
	uint64_t table_cap = 0;
	gf_table T{};
	std::thread th;
	// the module chain captured once (serve_chain_graph) and its batch arguments
	hipGraphExec_t chain_exec = nullptr;
	DevBuf chain_args;
	gf_serve_chain_args chain_args_h{};
	// device bytes held by the batch buffers (the graph's argument block aside)
	uint64_t held() const
	{
		return W.bytes + counts.bytes + offs.bytes + pixsum.bytes + segexp.bytes + pix.bytes + state.bytes +
		       seeds.bytes + segs.bytes + segall.bytes + reqs.bytes + table.bytes + pos.bytes + zu.bytes + vis.bytes +
		       extbeam.bytes + roff.bytes;
	}
	void release()
	{
		for (DevBuf *b : {&W, &counts, &offs, &pixsum, &segexp, &pix, &state, &seeds, &segs, &segall, &reqs, &table,
		                  &pos, &zu, &vis, &extbeam, &roff})
			b->free();
		table_cap = 0;
		T = gf_table{};
	}
};

} // namespace

struct gf_server {
	gf_catalog base;                 // models, binning, module parameters
	double dx = 0;
	float m0raw = 0;
	std::vector<uint32_t> primes;
	gf_server_config cfg{};
	gf_result_fn fn = nullptr;
	void *user = nullptr;
	DevBuf dmodels, dlfP, dchain;
	DevBuf dprimes;                  // the stream multipliers, for seeding requests' streams on the device
	std::vector<gf_model> hmodels;
	gf_serve_chain hchain{};
	std::mutex mu;
	std::condition_variable cv_sub, cv_ready, cv_idle;
	std::deque<std::unique_ptr<SReq>> submitted, ready;
	uint64_t n_submitted = 0, n_completed = 0, n_batches = 0, n_rows = 0, n_visible = 0, batch_seq = 0;
	double t_prep = 0, t_batch = 0, t_wait = 0, t_deliver = 0;
	double t_first = 0, t_last = 0;
	bool stop = false;
	std::vector<std::thread> prep;
	std::vector<std::unique_ptr<Worker>> workers;
	std::mutex cb_mu;
};

namespace {

void prepare(gf_server &s, SReq &q)
{
	gf_footprint &fp = q.foot;
	fp = gf_footprint{};
	fp.nregion = 1;
	gf_foot_region &f = fp.r[0];
	if (q.r.radius > 0) {
		double l = q.r.l * M_PI / 180., b = q.r.b * M_PI / 180.;
		f.kind = 2;
		f.cl = cos(l); f.cb = cos(b); f.sl = sin(l); f.sb = sin(b);
		f.cosr = cos(q.r.radius * M_PI / 180.);
	} else {
		f = rect_region(q.r.l, q.r.l1, q.r.b, q.r.b1);
	}
	q.pix = pixelize_region(fp, s.dx);
	if (s.base.has_ext) q.extmin = pixel_extinction_minima(s.base, q.pix);
	q.sp = s.base.sp;
	q.sp.m1 = q.r.m1;
	q.sp.nm = (int)round((q.sp.m1 - s.m0raw) / q.sp.dm);
	q.sp.npixels = (int)q.pix.size();
	const uint64_t K = (uint64_t)(q.sp.nm + q.sp.nM - 1);
	q.nunits_model = (uint64_t)q.pix.size() * K;
	q.msp.clear();
	q.munits.clear();
	q.nunits = 0;
	for (auto &m : s.base.models) {
		gf_skygen_params sp = q.sp;
		sp.M0 = m->sp.M0; sp.M1 = m->sp.M1; sp.nM = m->sp.nM;
		q.msp.push_back(sp);
		q.munits.push_back((uint64_t)q.pix.size() * (uint64_t)(sp.nm + sp.nM - 1));
		q.nunits += q.munits.back();
	}
	uint64_t want = std::max<uint64_t>(64, q.nunits_model);
	want = (want + 63) / 64 * 64;
	q.S = (uint32_t)std::min<uint64_t>(65536, want);
	q.nI = 0;
	for (uint64_t u : q.munits) q.nI += (u + q.S - 1) / q.S;
	q.seed32 = gf_seed_from_pipeline_seed(q.r.seed);
}

// The server table's columns (the catalog table of gf_table)
struct ColDef { gf_col gf_table::*m; uint32_t elem, width; };

std::vector<ColDef> table_defs(const gf_server &s)
{
	const uint32_t nb = (uint32_t)std::max(1, s.base.nbands), na = (uint32_t)std::max(1, s.base.nabsmag);
	return {
		{&gf_table::lb, 8, 2}, {&gf_table::projIdx, 4, 1}, {&gf_table::projXY, 4, 2}, {&gf_table::DM, 4, 1},
		{&gf_table::absmag, 4, 1}, {&gf_table::XYZ, 4, 3}, {&gf_table::Am, 4, 1}, {&gf_table::AmInf, 4, 1},
		{&gf_table::comp, 4, 1}, {&gf_table::hidden, 4, 1}, {&gf_table::FeH, 4, 1}, {&gf_table::FeH_comp, 4, 1},
		{&gf_table::absmagSys, 4, na}, {&gf_table::absmagSysNcomp, 4, 1},
		{&gf_table::vcyl, 4, 3}, {&gf_table::mags, 4, nb}, {&gf_table::photoFlags, 4, 1},
		{&gf_table::obsMags, 4, nb}, {&gf_table::pm, 4, 3}, {&gf_table::radec, 8, 2},
	};
}

uint32_t col_pitch(uint64_t cap, uint32_t elem) { return (uint32_t)(((cap * elem + 127) / 128) * 128); }

size_t table_bytes(const gf_server &s, uint64_t cap)
{
	size_t total = 0;
	for (auto &d : table_defs(s)) total += (size_t)col_pitch(cap, d.elem) * d.width;
	return total;
}

void worker_table(gf_server &s, Worker &w, uint64_t cap, uint64_t share)
{
	if (w.table_cap >= cap && w.table.p) return;
	// a quarter of headroom (at least 4,096 rows, at most GF_MAX_TABLE_ROWS)
	// when it stays within the share
	const uint64_t others = w.held() - w.table.bytes;
	uint64_t want = std::min<uint64_t>(std::max<uint64_t>(cap + cap / 4, 4096), GF_MAX_TABLE_ROWS);
	if (others + table_bytes(s, want) > share) want = cap;
	cap = want;
	w.table.alloc(table_bytes(s, cap));
	char *p = (char *)w.table.p;
	for (auto &d : table_defs(s)) {
		const uint32_t pitch = col_pitch(cap, d.elem);
		w.T.*(d.m) = gf_col{p, pitch, d.width, d.elem, 0};
		p += (size_t)pitch * d.width;
	}
	w.table_cap = cap;
}

// Grow a worker buffer to n elements: with a quarter of headroom when the
// worker's buffers stay within its share of the budget, else exactly.
template <typename T> void grow(Worker &w, uint64_t share, DevBuf &b, size_t n)
{
	const size_t need = n * sizeof(T);
	if (b.bytes >= need) return;
	const uint64_t others = w.held() - b.bytes;
	size_t want = need + need / 4 + 256;
	if (others + want > share) want = need;
	b.alloc(want);
}

// Device bytes of one request in a batch apart from its rows: per unit of
// every model its weight, count, offset and emission base (8 + 4 + 8 + 8 B),
// per pixel the pencil beam and its partial sum, per stream its state, its
// seeding record, segments and request record, and its pixels' extinction
// minima.
uint64_t request_fixed_bytes(const gf_server &s, const SReq &q)
{
	const uint64_t nm = s.hmodels.size();
	return q.nunits * 28 + (uint64_t)q.pix.size() * (sizeof(gf_pencil_beam) + 8) + (uint64_t)q.S * 12 +
	       sizeof(gf_seed_seg) + nm * (2 * sizeof(gf_serve_seg) + 8) + sizeof(gf_serve_req) + 16 +
	       q.extmin.size() * 4;
}

// Device bytes of a batch beyond its requests' own: each buffer's spare end
// element and the table's column pitches (rounded up to 128 bytes).
uint64_t batch_slack(const gf_server &s)
{
	uint64_t planes = 0;
	for (auto &d : table_defs(s)) planes += d.width;
	return planes * 128 + 4096;
}

// Device bytes per output row: the table's columns and the module chain's
// stream positions.
uint64_t row_device_bytes(const gf_server &s)
{
	uint64_t b = 0;
	for (auto &d : table_defs(s)) b += (uint64_t)d.elem * d.width;
	return b + 8 * (uint64_t)s.hchain.nplanes;
}

// Rows a request of expectation N stays below except with probability
// ~1e-9: six Poisson standard deviations and a floor.
uint64_t rows_bound(double N)
{
	return (uint64_t)ceil(N + 6.0 * sqrt(N) + 32.0);
}

// The rows a request is admitted with: its counted expectation's bound, or
// before any count a rough bound from its area (3e8 stars per steradian).
uint64_t admission_rows(const SReq &q)
{
	if (q.counted >= 0) return rows_bound(q.counted);
	double area = 0;
	for (auto &p : q.pix) area += (double)p.dA;
	return rows_bound(area * 3.0e8);
}

uint64_t worker_share(const gf_server &s)
{
	return s.cfg.hbm_budget / (uint64_t)std::max(1, s.cfg.workers);
}

// Run one batch.  The count runs for every candidate of the batch; the
// requests whose counted rows no longer fit the worker's share after the
// first are moved to `deferred` (to go back to the queue, their count kept),
// and a first request that does not fit the share alone to `rejected`.
// A worker's wait for its stream, added to the server's device-wait time.
int wait_stream(gf_server &s, hipStream_t st)
{
	const double t0 = server_clock();
	GF_HIP(hipStreamSynchronize(st));
	const double dt = server_clock() - t0;
	std::lock_guard<std::mutex> lk(s.mu);
	s.t_wait += dt;
	return 0;
}

// Device bytes a batch of `units` units, `pixels` pixels, `words` stream
// table words, nr requests of nm models, `extf` extinction-minima floats and
// `rows` table rows needs beyond what the worker's buffers already hold: for
// each buffer the part of its need above its size (the sizes run_batch grows
// them to).
uint64_t beyond_held(const gf_server &s, const Worker &w, uint64_t units, uint64_t pixels, uint64_t words, int nr,
                     int nm, uint64_t extf, uint64_t rows)
{
	auto over = [](const DevBuf &b, uint64_t need) -> uint64_t { return need > b.bytes ? need - b.bytes : 0; };
	const uint64_t nseg = (uint64_t)nr * nm + 1, nq = (uint64_t)nr + 1;
	uint64_t e = over(w.pix, std::max<uint64_t>(pixels, 1) * sizeof(gf_pencil_beam)) +
	             over(w.state, std::max<uint64_t>(words, 1) * 4) +
	             over(w.seeds, std::max<uint64_t>(nr, 1) * sizeof(gf_seed_seg)) + over(w.W, (units + 1) * 8) +
	             over(w.counts, (units + 1) * 4) + over(w.offs, (units + 2) * 8) + over(w.zu, (units + 1) * 8) +
	             over(w.pixsum, (pixels + 1) * 8) + over(w.segexp, nseg * 8) +
	             over(w.segs, nseg * sizeof(gf_serve_seg)) + over(w.segall, nseg * sizeof(gf_serve_seg)) +
	             over(w.reqs, nq * sizeof(gf_serve_req)) + over(w.vis, nq * 8) + over(w.roff, nq * 8) +
	             over(w.extbeam, extf * 4) + over(w.pos, (uint64_t)s.hchain.nplanes * rows * sizeof(uint2));
	if (!(w.table.p && w.table_cap >= rows)) {
		const uint64_t tb = table_bytes(s, std::min<uint64_t>(rows, GF_MAX_TABLE_ROWS));
		e += tb > w.table.bytes ? tb - w.table.bytes : 0;
	}
	return e + batch_slack(s);
}

int run_batch(gf_server &s, Worker &w, std::vector<std::unique_ptr<SReq>> &batch,
              std::vector<std::unique_ptr<SReq>> &deferred, std::vector<std::unique_ptr<SReq>> &rejected)
{
	hipStream_t st = w.st;
	const int nm = (int)s.hmodels.size();
	int nr = (int)batch.size();
	const uint64_t share = worker_share(s);
	const uint64_t rowb = row_device_bytes(s);
	// the batch's admitted size; the buffers of earlier batches are released
	// first when growing them for it would not stay within the share (they
	// are reused as they are otherwise: an allocation and a free can cost
	// far more than the batch).  The buffers other than the table grow with
	// headroom only while that leaves the share room for the table of the
	// admitted rows.
	uint64_t reserve = batch_slack(s);
	{
		uint64_t units = 0, pixels = 0, words = 0, extf = 0, rows = 0;
		for (auto &q : batch) {
			reserve += admission_rows(*q) * rowb;
			units += q->nunits;
			pixels += q->pix.size();
			words += 3 * (uint64_t)q->S;
			extf += q->extmin.size();
			rows += admission_rows(*q);
		}
		// the batch's minima cube pads its pixels to rows of 2,048 (beam_texture)
		if (extf && pixels) extf = (pixels + 2047) * (extf / pixels);
		if (w.held() + beyond_held(s, w, units, pixels, words, nr, nm, extf, rows) > share) w.release();
	}
	const uint64_t fshare = share > reserve ? share - reserve : 0;
	// unit layout: request-major, model-minor
	std::vector<uint64_t> uoff(nr + 1, 0);
	std::vector<uint32_t> poff(nr + 1, 0), loff(nr + 1, 0);
	for (int r = 0; r < nr; r++) {
		uoff[r + 1] = uoff[r] + batch[r]->nunits;
		poff[r + 1] = poff[r] + (uint32_t)batch[r]->pix.size();
		loff[r + 1] = loff[r] + batch[r]->S;
	}
	const uint64_t Ucand = uoff[nr];
	std::vector<gf_pencil_beam> pix;
	pix.reserve(poff[nr]);
	// the requests' stream tables, seeded on the device (gf_mwc_seed_host's
	// tables, 3 S words each)
	std::vector<size_t> soff(nr + 1, 0);
	std::vector<gf_seed_seg> seeds(nr);
	uint32_t maxS = 0;
	for (int r = 0; r < nr; r++) {
		pix.insert(pix.end(), batch[r]->pix.begin(), batch[r]->pix.end());
		soff[r + 1] = soff[r] + 3 * (size_t)batch[r]->S;
		seeds[r] = gf_seed_seg{batch[r]->seed32, batch[r]->S, (uint64_t)soff[r]};
		maxS = std::max(maxS, batch[r]->S);
	}
	if (s.base.has_ext)
		for (size_t g = 0; g < pix.size(); g++) pix[g].extIdx = (int32_t)g;   // its texel of the batch's minima
	grow<gf_pencil_beam>(w, fshare, w.pix, std::max<size_t>(pix.size(), 1));
	grow<uint32_t>(w, fshare, w.state, std::max<size_t>(soff[nr], 1));
	grow<gf_seed_seg>(w, fshare, w.seeds, std::max<size_t>(seeds.size(), 1));
	grow<double>(w, fshare, w.W, Ucand + 1);
	grow<uint32_t>(w, fshare, w.counts, Ucand + 1);
	grow<uint64_t>(w, fshare, w.offs, Ucand + 2);
	grow<double>(w, fshare, w.pixsum, (size_t)poff[nr] + 1);
	grow<double>(w, fshare, w.segexp, (size_t)nr * nm + 1);
	grow<gf_serve_seg>(w, fshare, w.segs, (size_t)nr * nm + 1);
	grow<gf_serve_seg>(w, fshare, w.segall, (size_t)nr * nm + 1);
	grow<gf_serve_req>(w, fshare, w.reqs, (size_t)nr + 1);
	GF_HIP(hipMemcpyAsync(w.pix.p, pix.data(), pix.size() * sizeof(gf_pencil_beam), hipMemcpyHostToDevice, st));
	GF_HIP(hipMemcpyAsync(w.seeds.p, seeds.data(), seeds.size() * sizeof(gf_seed_seg), hipMemcpyHostToDevice, st));
	uint32_t *dstate = (uint32_t *)w.state.p;
	{
		int rc = gf::mwc_seed_streams((const gf_seed_seg *)w.seeds.p, (uint32_t)nr, maxS,
		                              (const uint32_t *)s.dprimes.p, dstate, st);
		if (rc) return rc;
	}
	// segments per model (launch order: requests), and all segments in row order
	std::vector<gf_serve_seg> segs((size_t)nm * nr), segall((size_t)nr * nm);
	for (int r = 0; r < nr; r++) {
		SReq &q = *batch[r];
		uint64_t u0 = uoff[r], i0 = 0;
		for (int m = 0; m < nm; m++) {
			gf_serve_seg g{};
			g.sp = q.msp[m];
			g.foot = q.foot;
			g.lfP = (const double *)s.dlfP.p + (size_t)m * (s.base.sp.nM + 1);
			g.state = dstate + soff[r];
			g.unit0 = u0;
			g.nunits = q.munits[m];
			g.pix0 = poff[r];
			g.npix = (uint32_t)q.pix.size();
			g.slot0 = poff[r];
			g.lane0 = loff[r];
			g.S = q.S;
			g.model = m;
			g.req = r;
			// the models of a request use its streams in turn
			g.i0 = (uint32_t)i0;
			u0 += q.munits[m];
			i0 += (q.munits[m] + q.S - 1) / q.S;
			segs[(size_t)m * nr + r] = g;
			segall[(size_t)r * nm + m] = g;
		}
	}
	GF_HIP(hipMemcpyAsync(w.segs.p, segs.data(), segs.size() * sizeof(gf_serve_seg), hipMemcpyHostToDevice, st));
	GF_HIP(hipMemcpyAsync(w.segall.p, segall.data(), segall.size() * sizeof(gf_serve_seg), hipMemcpyHostToDevice, st));
	gf_serve_seg *dsegs = (gf_serve_seg *)w.segs.p;
	gf_serve_seg *dsegall = (gf_serve_seg *)w.segall.p;
	const uint64_t *jtab = (const uint64_t *)s.base.jtab.p;
	grow<uint64_t>(w, fshare, w.zu, Ucand + 1);
	const gf_model *dmodels = (const gf_model *)s.dmodels.p;
	const gf_pencil_beam *dpix = (const gf_pencil_beam *)w.pix.p;
	// the batch's extinction minima: pixel g of the batch at texel
	// (g div 2048, g mod 2048), the maps themselves resident
	gf_extinction ext{};
	if (s.base.has_ext) {
		const Cube &ref = s.base.ext_north.v.empty() ? s.base.ext_south : s.base.ext_north;
		std::vector<float> minima;
		minima.reserve((size_t)poff[nr] * ref.nz);
		for (int r = 0; r < nr; r++) minima.insert(minima.end(), batch[r]->extmin.begin(), batch[r]->extmin.end());
		Cube beam;
		beam_texture(beam, minima, poff[nr], ref.nz, ref.tc[2]);
		grow<float>(w, fshare, w.extbeam, beam.v.size());
		GF_HIP(hipMemcpyAsync(w.extbeam.p, beam.v.data(), beam.v.size() * 4, hipMemcpyHostToDevice, st));
		ext = s.base.ext;
		ext.beam = tex3d_of(beam, (const float *)w.extbeam.p);
		// the copy reads the host vector: wait before it goes out of scope
		if (int wrc = wait_stream(s, st)) return wrc;
	}
	// count
	for (int m = 0; m < nm; m++) {
		int rc = gf::serve_count(dsegs + (size_t)m * nr, nr, poff[nr], dmodels, ext, dpix, (double *)w.W.p,
		                         (double *)w.pixsum.p, (double *)w.segexp.p + (size_t)m * nr, st);
		if (rc) return rc;
	}
	std::vector<double> segexp((size_t)nm * nr);
	GF_HIP(hipMemcpyAsync(segexp.data(), w.segexp.p, segexp.size() * 8, hipMemcpyDeviceToHost, st));
	if (int wrc = wait_stream(s, st)) return wrc;
	// admission by the counted expectation: the longest prefix of the
	// candidates whose fixed bytes and bounded rows fit the share (units,
	// pixels and streams are request-major, so a prefix keeps its layout)
	{
		uint64_t used = batch_slack(s), rows = 0;
		int k = 0;
		for (int r = 0; r < nr; r++) {
			SReq &q = *batch[r];
			q.counted = 0;
			for (int m = 0; m < nm; m++) q.counted += segexp[(size_t)m * nr + r];
			const uint64_t need = request_fixed_bytes(s, q) + rows_bound(q.counted) * rowb;
			// and the prefix's rows fit one table
			if (k == r && used + need <= share && rows + rows_bound(q.counted) <= GF_MAX_TABLE_ROWS) {
				used += need;
				rows += rows_bound(q.counted);
				k++;
			}
		}
		if (k == 0) {
			rejected.push_back(std::move(batch[0]));
			for (int r = 1; r < nr; r++) deferred.push_back(std::move(batch[r]));
			batch.clear();
			return 0;
		}
		if (k < nr) {
			// the buffers were sized for every candidate, the deferred ones
			// included: run the prefix again from released buffers, so that
			// what the worker holds is the prefix's own (its counts are kept,
			// so the second count admits all of it)
			for (int r = k; r < nr; r++) deferred.push_back(std::move(batch[r]));
			batch.resize(k);
			w.release();
			return run_batch(s, w, batch, deferred, rejected);
		}
	}
	const uint64_t U = uoff[nr];
	uint64_t bseq;
	{
		std::lock_guard<std::mutex> lk(s.mu);
		bseq = s.batch_seq++;
	}
	// plan: every unit of the batch at once (a segment's units start at its
	// models' offset on its request's streams)
	int rc = gf::serve_plan(dsegall, nr * nm, U, (const double *)w.W.p, jtab, (uint32_t *)w.counts.p,
	                        (uint64_t *)w.zu.p, st);
	if (rc) return rc;
	rc = gf::exclusive_scan_u32_u64((const uint32_t *)w.counts.p, (uint64_t *)w.offs.p, U, st);
	if (rc) return rc;
	std::vector<uint64_t> roff(nr + 1);
	grow<uint64_t>(w, fshare, w.roff, (size_t)nr + 1);
	rc = gf::serve_req_rows(dsegall, nr, nm, (const uint64_t *)w.offs.p, U, (uint64_t *)w.roff.p, st);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(roff.data(), w.roff.p, roff.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	if (int wrc = wait_stream(s, st)) return wrc;
	for (int r = 0; r < nr; r++) batch[r]->expected = batch[r]->counted;
	const uint64_t nrows = roff[nr];
	if (nrows > GF_MAX_TABLE_ROWS) return GF_E_BUDGET;
	// the rows drawn past their bound (a ~1e-9 event) may not fit the share
	if (w.held() - w.table.bytes - w.pos.bytes + table_bytes(s, nrows) +
	    (uint64_t)s.hchain.nplanes * nrows * sizeof(uint2) > share)
		return GF_E_BUDGET;
	{
		// the table's headroom leaves room for the stream positions
		const uint64_t posb = (uint64_t)s.hchain.nplanes * nrows * sizeof(uint2);
		const uint64_t need_pos = w.pos.bytes >= posb ? 0 : posb - w.pos.bytes;
		worker_table(s, w, nrows, share > need_pos ? share - need_pos : 0);
	}
	// row0/row1 of the row-ordered table from the scanned offsets
	rc = gf::serve_seg_rows(dsegall, nr * nm, (const uint64_t *)w.offs.p, st);
	if (rc) return rc;
	gf_skygen_cols sc{w.T.lb, w.T.projIdx, w.T.projXY, w.T.DM, w.T.absmag, w.T.XYZ, w.T.Am, w.T.AmInf, w.T.comp, w.T.hidden};
	rc = gf::serve_emit(dsegall, nr * nm, dmodels, ext, dpix, (const uint64_t *)w.offs.p, (const uint64_t *)w.zu.p, jtab,
	                    nrows, sc, st);
	if (rc) return rc;
	// the module chain's lanes: per request the streams that own a 256-row
	// chunk of its rows (stream s owns chunks s, s + S, ...; streams past
	// the request's chunk count have nothing to draw and are left out)
	std::vector<gf_serve_req> rq(nr);
	std::vector<uint32_t> aoff(nr + 1, 0);
	for (int r = 0; r < nr; r++) {
		const SReq &q = *batch[r];
		const uint64_t chunks = (roff[r + 1] - roff[r] + 255) / 256;
		aoff[r + 1] = aoff[r] + (uint32_t)std::min<uint64_t>(q.S, chunks);
		rq[r].state = dstate + soff[r];
		rq[r].row0 = roff[r];
		rq[r].row1 = roff[r + 1];
		rq[r].lane0 = aoff[r];
		rq[r].S = q.S;
		rq[r].advance = q.nI << GF_UNIT_SPAN_BITS;
	}
	const uint32_t nlanes = aoff[nr];
	GF_HIP(hipMemcpyAsync(w.reqs.p, rq.data(), rq.size() * sizeof(gf_serve_req), hipMemcpyHostToDevice, st));
	// the streams move past every unit before the module chain draws from them
	rc = gf::serve_advance((const gf_serve_req *)w.reqs.p, nr, nlanes, jtab, st);
	if (rc) return rc;
	grow<uint2>(w, fshare, w.pos, std::max<size_t>(1, (size_t)s.hchain.nplanes * nrows));
	if (s.cfg.eager_chain) {
		rc = gf::serve_chain((const gf_serve_req *)w.reqs.p, nr, nlanes, (const gf_serve_chain *)s.dchain.p,
		                     s.hchain, nrows, w.T, (uint2 *)w.pos.p, st);
		if (rc) return rc;
	} else if (nrows) {
		if (!w.chain_exec) {
			w.chain_args.alloc(sizeof(gf_serve_chain_args));
			rc = gf::serve_chain_graph((const gf_serve_chain_args *)w.chain_args.p, (const gf_serve_chain *)s.dchain.p,
			                           s.hchain, st, &w.chain_exec);
			if (rc) return rc;
		}
		// the previous batch's replay has finished (its results were read back)
		w.chain_args_h = gf_serve_chain_args{(const gf_serve_req *)w.reqs.p, nr, nlanes, nrows, w.T,
		                                     gf_positions{(uint2 *)w.pos.p, nrows}};
		GF_HIP(hipMemcpyAsync(w.chain_args.p, &w.chain_args_h, sizeof(gf_serve_chain_args), hipMemcpyHostToDevice,
		                      st));
		GF_HIP(hipGraphLaunch(w.chain_exec, st));
	}
	uint64_t maxn = 0;
	for (int r = 0; r < nr; r++) maxn = std::max(maxn, roff[r + 1] - roff[r]);
	grow<unsigned long long>(w, fshare, w.vis, (size_t)nr + 1);
	rc = gf::serve_visible((const gf_serve_req *)w.reqs.p, nr, maxn, w.T.hidden, (unsigned long long *)w.vis.p, st);
	if (rc) return rc;
	std::vector<unsigned long long> visv(nr);
	GF_HIP(hipMemcpyAsync(visv.data(), w.vis.p, (size_t)nr * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
	if (int wrc = wait_stream(s, st)) return wrc;
	// results
	const double td0 = server_clock();
	uint64_t vis_total = 0;
	for (int r = 0; r < nr; r++) {
		SReq &q = *batch[r];
		uint64_t n = roff[r + 1] - roff[r], vis = visv[r];
		vis_total += vis;
		gf_result res{};
		res.id = q.r.id;
		res.status = 0;
		res.batch_size = nr;
		res.npixels = q.pix.size();
		res.generated = n;
		res.visible = vis;
		res.expected = q.expected;
		res.t_submit = q.t_submit;
		res.t_start = q.t_start;
		res.t_done = server_clock();
		res.device_bytes = request_fixed_bytes(s, q) + rows_bound(q.counted) * rowb;
		res.batch = bseq;
		if (s.fn) {
			std::lock_guard<std::mutex> lk(s.cb_mu);
			s.fn(s.user, &res, &w.T, roff[r], n);
		}
		std::lock_guard<std::mutex> lk(s.mu);
		s.n_completed++;
		s.t_last = res.t_done;
	}
	{
		std::lock_guard<std::mutex> lk(s.mu);
		s.t_deliver += server_clock() - td0;
		s.n_batches++;
		s.n_rows += nrows;
		s.n_visible += vis_total;
	}
	s.cv_idle.notify_all();
	return 0;
}

void fail_batch(gf_server &s, std::vector<std::unique_ptr<SReq>> &batch, int code)
{
	for (auto &q : batch) {
		gf_result res{};
		res.id = q->r.id;
		res.status = code;
		res.batch_size = (int)batch.size();
		res.t_submit = q->t_submit;
		res.t_start = q->t_start;
		res.t_done = server_clock();
		if (q->counted >= 0)
			res.device_bytes = request_fixed_bytes(s, *q) + rows_bound(q->counted) * row_device_bytes(s);
		if (s.fn) {
			std::lock_guard<std::mutex> lk(s.cb_mu);
			s.fn(s.user, &res, nullptr, 0, 0);
		}
		std::lock_guard<std::mutex> lk(s.mu);
		s.n_completed++;
	}
	s.cv_idle.notify_all();
}

void prep_loop(gf_server &s)
{
	for (;;) {
		std::unique_ptr<SReq> q;
		{
			std::unique_lock<std::mutex> lk(s.mu);
			s.cv_sub.wait(lk, [&] { return s.stop || !s.submitted.empty(); });
			if (s.stop && s.submitted.empty()) return;
			q = std::move(s.submitted.front());
			s.submitted.pop_front();
		}
		double t0 = server_clock();
		prepare(s, *q);
		double t1 = server_clock();
		{
			std::lock_guard<std::mutex> lk(s.mu);
			s.t_prep += t1 - t0;
			s.ready.push_back(std::move(q));
		}
		s.cv_ready.notify_all();
	}
}

void worker_loop(gf_server &s, Worker &w)
{
	(void)hipSetDevice(0);
	const uint64_t share = worker_share(s);
	const uint64_t rowb = row_device_bytes(s);
	for (;;) {
		std::vector<std::unique_ptr<SReq>> batch, deferred, rejected;
		{
			std::unique_lock<std::mutex> lk(s.mu);
			auto &src = s.cfg.plain ? s.submitted : s.ready;
			s.cv_ready.wait(lk, [&] { return s.stop || !src.empty(); });
			if (s.stop && src.empty()) return;
			if (!s.cfg.plain && s.cfg.linger_us > 0 && (int)src.size() < s.cfg.max_batch)
				s.cv_ready.wait_for(lk, std::chrono::microseconds(s.cfg.linger_us),
				                    [&] { return s.stop || (int)src.size() >= s.cfg.max_batch; });
			// admission: up to max_batch requests whose rows (counted when an
			// earlier batch counted them, else bounded from their area) fit the
			// worker's share of the device budget; a counted request that does
			// not fit the share alone is rejected
			uint64_t used = batch_slack(s);
			while (!src.empty() && (int)batch.size() < std::max(1, s.cfg.max_batch)) {
				SReq &q = *src.front();
				const uint64_t need = request_fixed_bytes(s, q) + admission_rows(q) * rowb;
				if (q.counted >= 0 && need + batch_slack(s) > share) {
					rejected.push_back(std::move(src.front()));
					src.pop_front();
					continue;
				}
				if (!batch.empty() && used + need > share) break;
				used += need;
				batch.push_back(std::move(src.front()));
				src.pop_front();
			}
		}
		if (!rejected.empty()) {
			fail_batch(s, rejected, GF_E_BUDGET);
			rejected.clear();
		}
		if (batch.empty()) continue;
		double t0 = server_clock();
		for (auto &q : batch) {
			if (s.cfg.plain) prepare(s, *q);
			q->t_start = t0;
		}
		double t1 = server_clock();
		int rc;
		try {
			rc = run_batch(s, w, batch, deferred, rejected);
		} catch (const ConfigError &e) {   // a device allocation failed
			g_last_error = e.msg;
			w.release();
			rc = GF_E_BUDGET;
		}
		if (rc) fail_batch(s, batch, rc);
		if (!rejected.empty()) fail_batch(s, rejected, GF_E_BUDGET);
		if (!deferred.empty()) {
			{
				std::lock_guard<std::mutex> lk(s.mu);
				auto &src = s.cfg.plain ? s.submitted : s.ready;
				for (auto it = deferred.rbegin(); it != deferred.rend(); ++it) src.push_front(std::move(*it));
			}
			s.cv_ready.notify_all();
		}
		{
			std::lock_guard<std::mutex> lk(s.mu);
			s.t_batch += server_clock() - t1;
			if (s.cfg.plain) s.t_prep += t1 - t0;
		}
	}
}

} // namespace

extern "C" {

double gf_server_clock(void) { return server_clock(); }

int gf_server_create(const char *cmd_conf, const char *datadir, const gf_server_config *cfg, gf_result_fn fn,
                     void *user, gf_server **out)
{
	GF_CHECK(cmd_conf && datadir && cfg && out, GF_E_ARG);
	auto s = std::make_unique<gf_server>();
	s->base.datadir = datadir;
	s->cfg = *cfg;
	if (s->cfg.workers < 1) s->cfg.workers = 1;
	if (s->cfg.max_batch < 1) s->cfg.max_batch = 1;
	if (s->cfg.plain) { s->cfg.workers = 1; s->cfg.max_batch = 1; }
	s->fn = fn;
	s->user = user;
	try {
		open_catalog(s->base, cmd_conf);
		s->dx = s->base.skygen.getd("dx", 1.0) * M_PI / 180.;
		s->m0raw = (float)(s->base.skygen.count("m0") ? s->base.skygen.getd("m0") : s->base.skygen.getd("r0"));
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
	gf_catalog &c = s->base;
	s->primes.resize(70763);
	{
		std::ifstream f(c.datadir + "/safeprimes32.u32", std::ios::binary);
		if (!f) { g_last_error = "cannot read safeprimes32.u32"; return GF_E_CONFIG; }
		f.read((char *)s->primes.data(), s->primes.size() * 4);
		s->primes.resize(f.gcount() / 4);
	}
	try {
		s->dprimes.upload(s->primes);
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
	// models and their LF prefix tables (device)
	for (auto &m : c.models) s->hmodels.push_back(m->m);
	try {
		s->dmodels.upload(s->hmodels);
		s->dlfP.alloc(s->hmodels.size() * ((size_t)c.sp.nM + 1) * sizeof(double));
	} catch (const ConfigError &e) {
		g_last_error = e.msg;
		return GF_E_CONFIG;
	}
	for (size_t i = 0; i < s->hmodels.size(); i++) {
		// each model's prefix over its own absolute-magnitude bins (at most the job's)
		int rc = gf::lf_prefix(c.models[i]->sp, s->hmodels[i], (double *)s->dlfP.p + i * ((size_t)c.sp.nM + 1),
		                       c.stream);
		if (rc) return rc;
	}
	// module chain
	gf_serve_chain &h = s->hchain;
	if (c.has_feh) { h.feh = c.feh; h.has_feh = 1; }
	if (c.has_gfeh) { h.gfeh = c.gfeh; h.has_gfeh = 1; h.gfeh_comps = c.gfeh_comps; }
	if (c.has_mult) { h.mult = c.mult; h.has_mult = 1; h.mult_comps = c.mult_comps; }
	if (c.has_bond) { h.bond = c.bond; h.has_bond = 1; }
	if (c.has_phot) { h.phot = c.phot; h.has_phot = 1; h.phot_comps = c.phot_comps; }
	if (c.has_perr) { h.perr = c.perr; h.has_perr = 1; }
	if (c.has_v2pm) { h.v2pm = c.v2pm; h.has_v2pm = 1; }
	h.gal2other = c.gal2other;
	gf_chain_planes(h);
	try {
		s->dchain.alloc(sizeof(gf_serve_chain));
	} catch (const ConfigError &e) {
		return GF_E_CONFIG;
	}
	GF_HIP(hipMemcpy(s->dchain.p, &h, sizeof(h), hipMemcpyHostToDevice));
	gf_server *sp = s.get();
	for (int i = 0; i < s->cfg.workers; i++) {
		auto w = std::make_unique<Worker>();
		GF_HIP(hipStreamCreate(&w->st));
		s->workers.push_back(std::move(w));
	}
	if (!s->cfg.plain)
		for (int i = 0; i < 8; i++) s->prep.emplace_back(prep_loop, std::ref(*sp));
	for (auto &w : s->workers) {
		Worker *wp = w.get();
		w->th = std::thread([sp, wp] { worker_loop(*sp, *wp); });
	}
	*out = s.release();
	return 0;
}

int gf_server_submit(gf_server *s, const gf_request *req)
{
	GF_CHECK(s && req, GF_E_ARG);
	GF_CHECK((req->radius > 0 && req->radius < 90) ||
	         (req->radius == 0 && req->l1 > req->l && req->l1 - req->l <= 360. && req->b >= -90. &&
	          req->b < req->b1 && req->b1 <= 90.), GF_E_ARG);
	GF_CHECK(std::isfinite(req->l) && std::isfinite(req->b) && req->b >= -90. && req->b <= 90., GF_E_ARG);
	// at least one apparent-magnitude bin between the server's m0 and m1
	GF_CHECK(std::isfinite(req->m1) && std::round((req->m1 - s->m0raw) / s->base.sp.dm) >= 1.0, GF_E_ARG);
	auto q = std::make_unique<SReq>();
	q->r = *req;
	q->t_submit = server_clock();
	{
		std::lock_guard<std::mutex> lk(s->mu);
		if (s->n_submitted == 0) s->t_first = q->t_submit;
		s->n_submitted++;
		s->submitted.push_back(std::move(q));
	}
	if (s->cfg.plain) s->cv_ready.notify_one();
	else s->cv_sub.notify_one();
	return 0;
}

int gf_server_wait_idle(gf_server *s)
{
	GF_CHECK(s, GF_E_ARG);
	std::unique_lock<std::mutex> lk(s->mu);
	s->cv_idle.wait(lk, [&] { return s->n_completed >= s->n_submitted; });
	return 0;
}

int gf_server_stats_get(gf_server *s, gf_server_stats *st)
{
	GF_CHECK(s && st, GF_E_ARG);
	std::lock_guard<std::mutex> lk(s->mu);
	st->submitted = s->n_submitted;
	st->completed = s->n_completed;
	st->batches = s->n_batches;
	st->rows = s->n_rows;
	st->visible = s->n_visible;
	st->t_first_submit = s->t_first;
	st->t_last_done = s->t_last;
	st->t_prep = s->t_prep;
	st->t_batch = s->t_batch;
	st->t_wait = s->t_wait;
	st->t_deliver = s->t_deliver;
	return 0;
}

void gf_server_destroy(gf_server *s)
{
	if (!s) return;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		s->stop = true;
	}
	s->cv_sub.notify_all();
	s->cv_ready.notify_all();
	for (auto &t : s->prep)
		if (t.joinable()) t.join();
	for (auto &w : s->workers) {
		if (w->th.joinable()) w->th.join();
		if (w->chain_exec) (void)hipGraphExecDestroy(w->chain_exec);
		if (w->st) (void)hipStreamDestroy(w->st);
	}
	if (s->base.stream) (void)hipStreamDestroy(s->base.stream);
	s->base.stream = nullptr;
	delete s;
}

} // extern "C"
