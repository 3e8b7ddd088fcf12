// Output stages behind the chunk callback: os_textout (pipeline.cpp:199-288)
// and os_fitsout (pipeline.cpp:767-1035).  The visible rows of every chunk
// are formatted or packed on the device (gf_text_format_rows,
// gf_fits_pack_rows), copied to pinned host memory and appended to the file;
// the file's structure is written on the host.
//
// Text: a header line "# " followed by every column's definition and a space
// (otable::serialize_header, otable.cpp:542-559), then one line per visible
// row.  FITS: an empty primary HDU, the binary table CATALOG with one column
// per output column (TTYPE = its name, TFORM = width and E / D / J), and the
// binary table METADATA holding the text header's definitions in one string
// column HEADER (os_fitsout::~os_fitsout, :1006-1034).
//
// countsMap (os_countsMap, :293-762): the job's countsMap configuration
// read as os_countsMap::construct reads it (:590-682), every chunk binned
// on the device into one dense map (gf_countsmap_bin), and the map written
// when the writer closes, as ~os_countsMap writes it (:684-762): a header,
// then for every beam that received a row (every map-0 beam with "output
// all bins") one line per z bin: map, the beam's x and y (Lambert x, y
// with equalarea), z, the counts of every z element and coadd, then the
// beam centre's (lon, lat) in degrees with equalarea or its area in square
// degrees without.  Numbers are written as an ostream writes them: "%g"
// (six significant digits) for floating point values, decimal integers.
#include "host.h"
#include "../gf_runtime.h"
#include <hip/hip_runtime.h>
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <memory>
#include <string>
#include <vector>

using namespace gfh;

namespace {

// os_countsMap's state (pipeline.cpp:293-355) and its device map.
struct CountsMap {
	gf_countsmap_params p{};
	std::string xy, z;                                  // column names
	gf_col gf_table::*xy_member = nullptr, gf_table::*z_member = nullptr;
	std::vector<std::string> coadd_names;
	std::vector<float> coadd;                           // ncoadd x zmag_width
	int dense = 0;
	double l0[2] = {}, phi1[2] = {};                    // the two projections' poles (radians)
	size_t nbeams = 0, nvals = 0;                       // beams; uint32 counts in the map
	float *dcoadd = nullptr;
	uint32_t *dcounts = nullptr;
	uint8_t *dbeams = nullptr;
	unsigned long long *dtotal = nullptr;

	~CountsMap()
	{
		if (dcoadd) (void)hipFree(dcoadd);
		if (dcounts) (void)hipFree(dcounts);
		if (dbeams) (void)hipFree(dbeams);
		if (dtotal) (void)hipFree(dtotal);
	}
};

} // namespace

struct gf_writer {
	int kind = 0;
	FILE *f = nullptr;
	std::vector<OutColumn> cols;
	std::vector<gf_outcol> oc;       // per chunk: the table's columns
	uint32_t *dmap = nullptr;        // device copy of the component ids
	uint32_t nmap = 0;
	hipStream_t st = nullptr;        // packing / formatting
	hipStream_t cs = nullptr;        // copies to the host
	// two buffers of packed or formatted rows: a piece is copied to the host
	// while the next one is produced and the one before is written
	void *dbuf[2] = {};              // device
	size_t dcap[2] = {};
	char *hbuf[2] = {};              // pinned host copies
	size_t hcap[2] = {};
	hipEvent_t copied[2] = {};
	int pending = -1;                // buffer whose copy is queued and not yet written
	size_t pending_bytes = 0;
	int next = 0;                    // buffer of the next piece
	uint64_t rows = 0;
	uint32_t rec_bytes = 0;          // FITS record length
	long naxis2_at = -1;             // file offset of the CATALOG header's NAXIS2 card
	std::string header_def;
	std::unique_ptr<CountsMap> cm;
	int error = 0;

	~gf_writer()
	{
		if (cs) (void)hipStreamSynchronize(cs);
		for (int b = 0; b < 2; b++) {
			if (dbuf[b]) (void)hipFree(dbuf[b]);
			if (hbuf[b]) (void)hipHostFree(hbuf[b]);
			if (copied[b]) (void)hipEventDestroy(copied[b]);
		}
		if (dmap) (void)hipFree(dmap);
		if (st) (void)hipStreamDestroy(st);
		if (cs) (void)hipStreamDestroy(cs);
		if (f) fclose(f);
	}
};

namespace {

// "%[ ]W.Pf" / "%[ ]Wd" -> (flags, width, precision)
bool parse_fmt(const std::string &fmt, gf_outcol &o)
{
	size_t i = 0;
	if (i >= fmt.size() || fmt[i++] != '%') return false;
	o.flags = 0;
	if (i < fmt.size() && fmt[i] == ' ') { o.flags = 1; i++; }
	int w = 0;
	while (i < fmt.size() && isdigit((unsigned char)fmt[i])) w = 10 * w + (fmt[i++] - '0');
	o.width = (int16_t)w;
	o.prec = -1;
	if (i < fmt.size() && fmt[i] == '.') {
		i++;
		int p = 0;
		while (i < fmt.size() && isdigit((unsigned char)fmt[i])) p = 10 * p + (fmt[i++] - '0');
		o.prec = p;
	}
	if (i + 1 != fmt.size()) return false;
	if (fmt[i] == 'f') return o.prec >= 0 || (o.prec = 6, true);
	if (fmt[i] == 'd') return o.prec < 0;
	return false;
}

// One 80-character FITS header card.
std::string card(const std::string &key, const std::string &value, const std::string &comment = "")
{
	char b[81];
	std::string v = value;
	if (!comment.empty()) v += " / " + comment;
	snprintf(b, sizeof b, "%-8.8s= %-70.70s", key.c_str(), v.c_str());
	return std::string(b, 80);
}

std::string str_value(const std::string &s)
{
	std::string q = "'";
	for (char ch : s) q += ch == '\'' ? std::string("''") : std::string(1, ch);
	while (q.size() < 9) q += ' ';   // fixed-format strings are at least 8 characters
	return q + "'";
}

std::string int_value(long v)
{
	char b[32];
	snprintf(b, sizeof b, "%20ld", v);
	return b;
}

std::string end_block(std::string h)
{
	char b[81];
	snprintf(b, sizeof b, "%-80s", "END");
	h += std::string(b, 80);
	while (h.size() % 2880) h += ' ';
	return h;
}

int write_all(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) == n ? 0 : GF_E_IO; }

int pad_data(FILE *f)
{
	long at = ftell(f);
	if (at < 0) return GF_E_IO;
	size_t pad = (2880 - (size_t)at % 2880) % 2880;
	std::vector<char> z(pad, 0);
	return pad ? write_all(f, z.data(), pad) : 0;
}

// Buffer b of at least `bytes` on the device and the host (its previous
// contents are written out before it is reused).
int ensure(gf_writer &w, int b, size_t bytes)
{
	if (w.dcap[b] < bytes) {
		if (w.dbuf[b]) GF_HIP(hipFree(w.dbuf[b]));
		w.dbuf[b] = nullptr;
		w.dcap[b] = 0;
		GF_HIP(hipMalloc(&w.dbuf[b], bytes));
		w.dcap[b] = bytes;
	}
	if (w.hcap[b] < bytes) {
		if (w.hbuf[b]) GF_HIP(hipHostFree(w.hbuf[b]));
		w.hbuf[b] = nullptr;
		w.hcap[b] = 0;
		GF_HIP(hipHostMalloc((void **)&w.hbuf[b], bytes, hipHostMallocDefault));
		w.hcap[b] = bytes;
	}
	return 0;
}

// Writes the piece whose copy was queued last, once the copy is done.
int flush_pending(gf_writer &w)
{
	if (w.pending < 0) return 0;
	const int b = w.pending;
	w.pending = -1;
	GF_HIP(hipEventSynchronize(w.copied[b]));
	return w.pending_bytes ? write_all(w.f, w.hbuf[b], w.pending_bytes) : 0;
}

// The rows of a column from row r0 on.
gf_col from_row(gf_col c, uint64_t r0)
{
	c.ptr = (char *)c.ptr + r0 * c.elem;
	return c;
}

constexpr uint64_t PIECE_ROWS = 1ull << 22;   // rows packed or formatted per piece

// One piece of rows [r0, r0 + n) of the chunk: packed or formatted into
// buffer b on w.st, its copy to the host queued on w.cs, then the previous
// piece written to the file while this one copies.
int write_piece(gf_writer &w, const gf_table *t, uint64_t r0, uint64_t n)
{
	const int b = w.next;
	std::vector<gf_outcol> oc = w.oc;
	for (auto &o : oc) o.col = from_row(o.col, r0);
	const gf_col hidden = from_row(t->hidden, r0);
	int rc;
	uint64_t nv = 0;
	size_t bytes = 0;
	if (w.kind == GF_OUT_FITS) {
		if ((rc = ensure(w, b, std::max<size_t>(1, (size_t)n * w.rec_bytes)))) return rc;
		rc = gf_fits_pack_rows(oc.data(), (int)oc.size(), hidden, n, w.dbuf[b], n, &nv, w.st);
		if (rc) return rc;
		bytes = (size_t)nv * w.rec_bytes;
	} else {
		if (!w.dbuf[b] && (rc = ensure(w, b, 1 << 20))) return rc;
		uint64_t nb = 0;
		rc = gf_text_format_rows(oc.data(), (int)oc.size(), hidden, n, (char *)w.dbuf[b], w.dcap[b], &nb, &nv, w.st);
		if (rc == GF_E_CAPACITY) {
			if ((rc = ensure(w, b, nb + nb / 8))) return rc;
			rc = gf_text_format_rows(oc.data(), (int)oc.size(), hidden, n, (char *)w.dbuf[b], w.dcap[b], &nb, &nv,
			                         w.st);
		}
		if (rc) return rc;
		bytes = nb;
	}
	// the copy waits for the packing kernel, which may still run
	GF_HIP(hipEventRecord(w.copied[b], w.st));
	GF_HIP(hipStreamWaitEvent(w.cs, w.copied[b], 0));
	if (bytes) GF_HIP(hipMemcpyAsync(w.hbuf[b], w.dbuf[b], bytes, hipMemcpyDeviceToHost, w.cs));
	GF_HIP(hipEventRecord(w.copied[b], w.cs));
	if ((rc = flush_pending(w))) return rc;
	w.pending = b;
	w.pending_bytes = bytes;
	w.next = b ^ 1;
	w.rows += nv;
	return 0;
}

constexpr double D2R = 0.017453292519943295;   // pi / 180
constexpr double R2D = 57.29577951308232;      // 180 / pi

double rad(double d) { return d * D2R; }
double deg(double r) { return r * R2D; }

// A float key as an istream reads it (strtof), or dflt.
float cfg_float(const Config &cfg, const std::string &key, float dflt)
{
	if (!cfg.count(key)) return dflt;
	const std::string v = cfg.get(key);
	char *end = nullptr;
	const float f = strtof(v.c_str(), &end);
	if (end == v.c_str()) fail(key + " = '" + v + "' is not a number");
	return f;
}

// Bins of an axis from x0 to x1: (x1 - x0) / dx rounded when it is within
// 1e-4 of an integer, else its ceiling, plus one so that x1 is covered
// (bins are centred on x0 + k dx; os_countsMap::construct :612-637).
int axis_bins(float x0, double x1, float dx)
{
	const float tmp = (float)((x1 - (double)x0) / (double)dx);
	const double r = round((double)tmp);
	return (int)(fabs((double)tmp - r) < 1e-4 ? r : ceil((double)tmp)) + 1;
}

std::string fmt_g(double v)
{
	char b[40];
	snprintf(b, sizeof b, "%g", v);
	return b;
}

void append_u32(std::string &o, uint32_t v)
{
	char b[12];
	int n = 0;
	do { b[n++] = (char)('0' + v % 10); v /= 10; } while (v);
	while (n) o += b[--n];
}

// os_countsMap::construct (pipeline.cpp:590-682) and runtime_init (:577-588)
// on the job's countsMap configuration; allocates the device map.
int countsmap_open(const gf_catalog &c, gf_writer &w)
{
	const Config *cfgp = countsmap_config(c);
	if (!cfgp) {
		set_last_error("the job has no countsMap output stage");
		return GF_E_CONFIG;
	}
	const Config &cfg = *cfgp;
	auto m = std::make_unique<CountsMap>();
	gf_countsmap_params &p = m->p;
	try {
		p.equalarea = cfg.getd("equalarea", 0) != 0;
		float x0, dx, y0, dy, z0 = 0, dz = 1;
		double x1, y1, z1 = 40;
		if (p.equalarea) {   // Lambert x, y: [-2, 2] in 1-degree steps
			x0 = y0 = -2;
			x1 = y1 = 2;
			dx = dy = (float)rad(1.);
		} else {             // longitude, latitude in degrees
			x0 = 0; x1 = 360; dx = 1;
			y0 = -90; y1 = 90; dy = 1;
		}
		m->xy = cfg.get("xy");
		x0 = cfg_float(cfg, "x0", x0);
		x1 = cfg.getd("x1", x1);
		dx = cfg_float(cfg, "dx", dx);
		y0 = cfg_float(cfg, "y0", y0);
		y1 = cfg.getd("y1", y1);
		dy = cfg_float(cfg, "dy", dy);
		m->z = cfg.get("z");
		z0 = cfg_float(cfg, "z0", z0);
		z1 = cfg.getd("z1", z1);
		dz = cfg_float(cfg, "dz", dz);
		if (!(dx > 0 && dy > 0 && dz > 0)) fail("bin widths must be positive");
		p.x0 = x0; p.dx = dx; p.nx = axis_bins(x0, x1, dx);
		p.y0 = y0; p.dy = dy; p.ny = axis_bins(y0, y1, dy);
		p.z0 = z0; p.dz = dz; p.nz = axis_bins(z0, z1, dz);
		if (p.nx < 1 || p.ny < 1 || p.nz < 1) fail("empty bin range");
		// coadds: keys coadd.*, in key order, each a list of per-band offsets
		std::vector<std::vector<float>> offs;
		for (auto &kv : cfg.kv) {
			if (kv.first.compare(0, 6, "coadd.") != 0) continue;
			std::vector<float> o;
			const char *q = kv.second.c_str();
			for (char *end;; q = end) {
				const float f = strtof(q, &end);
				if (end == q) break;
				o.push_back(f);
			}
			m->coadd_names.push_back(kv.first);
			offs.push_back(o);
		}
		m->dense = (int)cfg.getd("output all bins", 0);
		// the columns
		std::vector<OutColumn> cols;
		output_schema(c, cols);
		const OutColumn *cx = nullptr, *cz = nullptr;
		for (auto &k : cols) {
			if (k.name == m->xy) cx = &k;
			if (k.name == m->z) cz = &k;
		}
		if (!cx || cx->type != 'D' || cx->width < 2)
			fail("xy = " + m->xy + " is not a two-element double column of the job");
		if (!cz || cz->type != 'E') fail("z = " + m->z + " is not a float column of the job");
		m->xy_member = cx->member;
		m->z_member = cz->member;
		p.zmag_width = (int32_t)cz->width;
		p.ncoadd = (int32_t)offs.size();
		for (size_t k = 0; k < offs.size(); k++) {
			if ((int)offs[k].size() < p.zmag_width)
				fail(m->coadd_names[k] + " needs " + std::to_string(p.zmag_width) + " offsets");
			m->coadd.insert(m->coadd.end(), offs[k].begin(), offs[k].begin() + p.zmag_width);
		}
	} catch (const ConfigError &e) {
		set_last_error("countsMap: " + e.msg);
		return GF_E_CONFIG;
	}
	// the north and south polar projections (os_countsMap() :349-354)
	m->l0[0] = m->phi1[0] = rad(90);
	m->l0[1] = m->phi1[1] = rad(-90);
	for (int i = 0; i < 2; i++) {
		p.proj[i][0] = m->l0[i];
		p.proj[i][1] = sin(m->phi1[i]);
		p.proj[i][2] = cos(m->phi1[i]);
		p.proj[i][3] = 0;
	}
	const size_t nmaps = p.equalarea ? 2 : 1;
	const size_t zw = (size_t)p.zmag_width + p.ncoadd;
	m->nbeams = (size_t)p.nx * p.ny * nmaps;
	m->nvals = m->nbeams * p.nz * zw;
	if (m->nvals > (16ull << 30)) {   // 64 GiB of counts
		set_last_error("countsMap: the map's " + std::to_string(m->nvals) + " bins exceed 2^34");
		return GF_E_BUDGET;
	}
	GF_HIP(hipMalloc((void **)&m->dcounts, m->nvals * sizeof(uint32_t)));
	GF_HIP(hipMalloc((void **)&m->dbeams, m->nbeams));
	GF_HIP(hipMalloc((void **)&m->dtotal, sizeof(unsigned long long)));
	GF_HIP(hipMemset(m->dcounts, 0, m->nvals * sizeof(uint32_t)));
	GF_HIP(hipMemset(m->dbeams, 0, m->nbeams));
	GF_HIP(hipMemset(m->dtotal, 0, sizeof(unsigned long long)));
	if (!m->coadd.empty()) {
		GF_HIP(hipMalloc((void **)&m->dcoadd, m->coadd.size() * sizeof(float)));
		GF_HIP(hipMemcpy(m->dcoadd, m->coadd.data(), m->coadd.size() * sizeof(float), hipMemcpyHostToDevice));
	}
	p.coadd_offsets = m->dcoadd;
	w.cm = std::move(m);
	return 0;
}

// ~os_countsMap (pipeline.cpp:684-762): the header, then the map.
int countsmap_write(gf_writer &w)
{
	CountsMap &m = *w.cm;
	const gf_countsmap_params &p = m.p;
	int rc = gf_countsmap_finish(&p, m.dcounts, w.st);
	if (rc) return rc;
	std::vector<uint32_t> cnt(m.nvals);
	std::vector<uint8_t> touched(m.nbeams);
	unsigned long long total = 0;
	GF_HIP(hipMemcpyAsync(cnt.data(), m.dcounts, m.nvals * sizeof(uint32_t), hipMemcpyDeviceToHost, w.st));
	GF_HIP(hipMemcpyAsync(touched.data(), m.dbeams, m.nbeams, hipMemcpyDeviceToHost, w.st));
	GF_HIP(hipMemcpyAsync(&total, m.dtotal, sizeof total, hipMemcpyDeviceToHost, w.st));
	GF_HIP(hipStreamSynchronize(w.st));

	const int zm = p.zmag_width, zw = zm + p.ncoadd, nmaps = p.equalarea ? 2 : 1;
	std::string o;
	if (p.equalarea) {
		o += "# Binned in lambert equal area projection, using the following poles:\n";
		for (int i = 0; i < 2; i++)
			o += "#\tmap = " + std::to_string(i) + "   :  l0 = " + fmt_g(deg(m.l0[i])) + ",  phi1 = " +
			     fmt_g(deg(m.phi1[i])) + "\n";
		o += "#\n";
		o += "# map\tlam_X\tlam_Y\t" + m.z;
	} else {
		o += "# map\t" + m.xy + "[0]\t" + m.xy + "[1]\t" + m.z;
	}
	for (int k = 0; k < zm; k++) o += "\tN(" + m.z + "[" + std::to_string(k) + "])";
	for (auto &n : m.coadd_names) o += "\t" + n;
	if (p.equalarea) o += "\t" + m.xy + "[0]\t" + m.xy + "[1]";
	else o += "\tdA";
	o += "\n#\n";

	// z of each z bin, as the reference computes it (in float)
	std::vector<std::string> zs(p.nz);
	for (int Z = 0; Z < p.nz; Z++) {
		const int i = Z * zw;
		const float zf = p.z0 + p.dz * ((float)i / (float)zw);
		zs[Z] = "\t" + fmt_g(zf);
	}
	unsigned long long sum = 0;
	for (int X = 0; X < p.nx; X++) {
		for (int Y = 0; Y < p.ny; Y++) {
			for (int map = 0; map < nmaps; map++) {
				const size_t beam = ((size_t)X * p.ny + Y) * nmaps + map;
				if (!touched[beam] && !(m.dense && map == 0)) continue;
				const float xf = p.x0 + p.dx * (float)X, yf = p.y0 + p.dy * (float)Y;
				const double x = xf, y = yf;
				const std::string head = std::to_string(map) + "\t" + fmt_g(x) + "\t" + fmt_g(y);
				std::string tail;
				if (p.equalarea) {
					// peyton::math::lambert::deproject (common/projections.h:58-66)
					const double sinphi1 = p.proj[map][1], cosphi1 = p.proj[map][2];
					const double r = sqrt(x * x + y * y);
					const double cc = 2 * asin(0.5 * r);
					const double lat = r != 0. ? asin(cos(cc) * sinphi1 + y * sin(cc) * cosphi1 / r) : m.phi1[map];
					const double lon = m.l0[map] +
					                   (r != 0. ? atan2(x * sin(cc), r * cosphi1 * cos(cc) - y * sinphi1 * sin(cc)) : 0.);
					tail = "\t" + fmt_g(deg(lon)) + "\t" + fmt_g(deg(lat));
				} else {
					// the pixel's area for (lon, lat) bins in degrees:
					// dx (sin(lat2) - sin(lat1)) in square degrees
					const double area = (double)p.dx * deg(sin(rad(std::min(y + 0.5 * p.dy, 90.))) -
					                                       sin(rad(std::max(y - 0.5 * p.dy, -90.))));
					tail = "\t" + fmt_g(area);
				}
				const uint32_t *cb = cnt.data() + beam * (size_t)p.nz * zw;
				for (int Z = 0; Z < p.nz; Z++) {
					o += head;
					o += zs[Z];
					for (int k = 0; k < zw; k++) {
						o += '\t';
						append_u32(o, cb[(size_t)Z * zw + k]);
						if (k < zm) sum += cb[(size_t)Z * zw + k];
					}
					o += tail;
					o += '\n';
				}
				if (o.size() > (64u << 20)) {
					if ((rc = write_all(w.f, o.data(), o.size()))) return rc;
					o.clear();
				}
			}
		}
	}
	if ((rc = write_all(w.f, o.data(), o.size()))) return rc;
	// the reference's own check: every z-element count was written once
	if (sum != total) {
		set_last_error("countsMap: map total " + std::to_string(sum) + " != " + std::to_string(total));
		return GF_E_STATE;
	}
	w.rows = total / (unsigned long long)zm;
	return 0;
}

} // namespace

extern "C" {

int gf_writer_open(const gf_catalog *c, const char *path, int kind, gf_writer **out)
{
	GF_CHECK(c && path && out && (kind == GF_OUT_TEXT || kind == GF_OUT_FITS || kind == GF_OUT_COUNTSMAP), GF_E_ARG);
	auto w = std::make_unique<gf_writer>();
	w->kind = kind;
	if (kind == GF_OUT_COUNTSMAP) {
		int rc = countsmap_open(*c, *w);
		if (rc) return rc;
		GF_HIP(hipStreamCreate(&w->st));
		w->f = fopen(path, "wb");
		if (!w->f) return GF_E_IO;
		*out = w.release();
		return 0;
	}
	output_schema(*c, w->cols);
	for (auto &k : w->cols) w->header_def += k.def + " ";
	const std::vector<uint32_t> &ids = component_ids(*c);
	w->nmap = (uint32_t)ids.size();
	if (w->nmap) {
		GF_HIP(hipMalloc((void **)&w->dmap, w->nmap * sizeof(uint32_t)));
		GF_HIP(hipMemcpy(w->dmap, ids.data(), w->nmap * sizeof(uint32_t), hipMemcpyHostToDevice));
	}
	w->oc.resize(w->cols.size());
	for (size_t i = 0; i < w->cols.size(); i++) {
		const OutColumn &k = w->cols[i];
		gf_outcol &o = w->oc[i];
		o = gf_outcol{};
		o.type = k.type;
		if (!parse_fmt(k.fmt, o)) return GF_E_ARG;
		if (k.map_comp) { o.map = w->dmap; o.nmap = w->nmap; }
		w->rec_bytes += k.width * (k.type == 'D' ? 8 : 4);
	}
	GF_HIP(hipStreamCreate(&w->st));
	GF_HIP(hipStreamCreate(&w->cs));
	for (auto &e : w->copied) GF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
	w->f = fopen(path, "wb");
	if (!w->f) return GF_E_IO;
	if (kind == GF_OUT_TEXT) {
		std::string h = "# " + w->header_def + "\n";
		if (write_all(w->f, h.data(), h.size())) return GF_E_IO;
	} else {
		std::string p = card("SIMPLE", std::string(19, ' ') + "T", "file does conform to FITS standard");
		p += card("BITPIX", int_value(8), "number of bits per data pixel");
		p += card("NAXIS", int_value(0), "number of data axes");
		p += card("EXTEND", std::string(19, ' ') + "T", "FITS dataset may contain extensions");
		p = end_block(p);
		std::string h = card("XTENSION", str_value("BINTABLE"), "binary table extension");
		h += card("BITPIX", int_value(8), "8-bit bytes");
		h += card("NAXIS", int_value(2), "2-dimensional binary table");
		h += card("NAXIS1", int_value(w->rec_bytes), "width of table in bytes");
		const long naxis2_card = (long)(p.size() + h.size());
		h += card("NAXIS2", int_value(0), "number of rows in table");
		h += card("PCOUNT", int_value(0), "size of special data area");
		h += card("GCOUNT", int_value(1), "one data group (required keyword)");
		h += card("TFIELDS", int_value((long)w->cols.size()), "number of fields in each row");
		for (size_t i = 0; i < w->cols.size(); i++) {
			const OutColumn &k = w->cols[i];
			h += card("TTYPE" + std::to_string(i + 1), str_value(k.name), "label for field");
			h += card("TFORM" + std::to_string(i + 1), str_value(std::to_string(k.width) + k.type),
			          "data format of field");
		}
		h += card("EXTNAME", str_value("CATALOG"), "name of this binary table extension");
		h = end_block(h);
		if (write_all(w->f, p.data(), p.size()) || write_all(w->f, h.data(), h.size())) return GF_E_IO;
		w->naxis2_at = naxis2_card;
	}
	*out = w.release();
	return 0;
}

int gf_writer_chunk(void *user, const gf_table *t, uint64_t nrows, int comp)
{
	(void)comp;
	gf_writer *w = (gf_writer *)user;
	GF_CHECK(w && t && w->f, GF_E_ARG);
	if (w->error) return w->error;
	int rc;
	if (w->kind == GF_OUT_COUNTSMAP) {
		CountsMap &m = *w->cm;
		rc = gf_countsmap_bin(&m.p, t->*(m.xy_member), t->*(m.z_member), t->hidden, nrows, m.dcounts, m.dbeams,
		                      m.dtotal, w->st);
		if (rc) return w->error = rc;
		// the chunk's columns are the run's again once the callback returns
		GF_HIP(hipStreamSynchronize(w->st));
		return 0;
	}
	for (size_t i = 0; i < w->cols.size(); i++) {
		w->oc[i].col = t->*(w->cols[i].member);
		if (!w->oc[i].col.ptr) return w->error = GF_E_ARG;
	}
	for (uint64_t r0 = 0; r0 < nrows; r0 += PIECE_ROWS)
		if ((rc = write_piece(*w, t, r0, std::min(PIECE_ROWS, nrows - r0)))) return w->error = rc;
	// the chunk's columns are the run's again once the callback returns:
	// every piece has been packed or formatted (the copies may still run)
	GF_HIP(hipStreamSynchronize(w->st));
	return 0;
}

int gf_writer_close(gf_writer *w, uint64_t *rows)
{
	if (!w) return GF_E_ARG;
	int rc = w->error;
	if (!rc && w->f) rc = flush_pending(*w);
	if (!rc && w->kind == GF_OUT_COUNTSMAP && w->f) rc = countsmap_write(*w);
	if (!rc && w->kind == GF_OUT_FITS && w->f) {
		// the CATALOG table's row count, its data padded to a 2880-byte block,
		// then the METADATA table
		rc = pad_data(w->f);
		const long end = ftell(w->f);
		if (!rc && fseek(w->f, w->naxis2_at, SEEK_SET) == 0) {
			std::string c2 = card("NAXIS2", int_value((long)w->rows), "number of rows in table");
			rc = write_all(w->f, c2.data(), c2.size());
			if (!rc && fseek(w->f, end, SEEK_SET) != 0) rc = GF_E_IO;
		} else if (!rc) {
			rc = GF_E_IO;
		}
		if (!rc) {
			const std::string &def = w->header_def;
			std::string h = card("XTENSION", str_value("BINTABLE"), "binary table extension");
			h += card("BITPIX", int_value(8), "8-bit bytes");
			h += card("NAXIS", int_value(2), "2-dimensional binary table");
			h += card("NAXIS1", int_value((long)def.size()), "width of table in bytes");
			h += card("NAXIS2", int_value(1), "number of rows in table");
			h += card("PCOUNT", int_value(0), "size of special data area");
			h += card("GCOUNT", int_value(1), "one data group (required keyword)");
			h += card("TFIELDS", int_value(1), "number of fields in each row");
			h += card("TTYPE1", str_value("HEADER"), "label for field");
			h += card("TFORM1", str_value(std::to_string(def.size()) + "A"), "data format of field");
			h += card("EXTNAME", str_value("METADATA"), "name of this binary table extension");
			h = end_block(h);
			rc = write_all(w->f, h.data(), h.size());
			if (!rc) rc = write_all(w->f, def.data(), def.size());
			if (!rc) rc = pad_data(w->f);
		}
	}
	if (w->f && fclose(w->f) != 0 && !rc) rc = GF_E_IO;
	w->f = nullptr;
	if (rows) *rows = w->rows;
	delete w;
	return rc;
}

} // extern "C"
