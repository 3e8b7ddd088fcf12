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
#include "host.h"
#include "../gf_runtime.h"
#include <hip/hip_runtime.h>
#include <ctype.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <memory>
#include <string>
#include <vector>

using namespace gfh;

struct gf_writer {
	int kind = 0;
	FILE *f = nullptr;
	std::vector<OutColumn> cols;
	std::vector<gf_outcol> oc;       // per chunk: the table's columns
	uint32_t *dmap = nullptr;        // device copy of the component ids
	uint32_t nmap = 0;
	hipStream_t st = nullptr;
	void *dbuf = nullptr;            // device output of one chunk
	size_t dcap = 0;
	char *hbuf = nullptr;            // pinned copy
	size_t hcap = 0;
	uint64_t rows = 0;
	uint32_t rec_bytes = 0;          // FITS record length
	long naxis2_at = -1;             // file offset of the CATALOG header's NAXIS2 card
	std::string header_def;
	int error = 0;

	~gf_writer()
	{
		if (dbuf) (void)hipFree(dbuf);
		if (hbuf) (void)hipHostFree(hbuf);
		if (dmap) (void)hipFree(dmap);
		if (st) (void)hipStreamDestroy(st);
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

int ensure(gf_writer &w, size_t bytes)
{
	if (w.dcap < bytes) {
		if (w.dbuf) GF_HIP(hipFree(w.dbuf));
		w.dbuf = nullptr;
		w.dcap = 0;
		GF_HIP(hipMalloc(&w.dbuf, bytes));
		w.dcap = bytes;
	}
	if (w.hcap < bytes) {
		if (w.hbuf) GF_HIP(hipHostFree(w.hbuf));
		w.hbuf = nullptr;
		w.hcap = 0;
		GF_HIP(hipHostMalloc((void **)&w.hbuf, bytes, hipHostMallocDefault));
		w.hcap = bytes;
	}
	return 0;
}

} // namespace

extern "C" {

int gf_writer_open(const gf_catalog *c, const char *path, int kind, gf_writer **out)
{
	GF_CHECK(c && path && out && (kind == GF_OUT_TEXT || kind == GF_OUT_FITS), GF_E_ARG);
	auto w = std::make_unique<gf_writer>();
	w->kind = kind;
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
	for (size_t i = 0; i < w->cols.size(); i++) {
		w->oc[i].col = t->*(w->cols[i].member);
		if (!w->oc[i].col.ptr) return w->error = GF_E_ARG;
	}
	int rc;
	uint64_t nv = 0;
	if (w->kind == GF_OUT_FITS) {
		if ((rc = ensure(*w, std::max<size_t>(1, (size_t)nrows * w->rec_bytes)))) return w->error = rc;
		rc = gf_fits_pack_rows(w->oc.data(), (int)w->oc.size(), t->hidden, nrows, w->dbuf, nrows, &nv, w->st);
		if (rc) return w->error = rc;
		const size_t bytes = (size_t)nv * w->rec_bytes;
		if (bytes) {
			GF_HIP(hipMemcpyAsync(w->hbuf, w->dbuf, bytes, hipMemcpyDeviceToHost, w->st));
			GF_HIP(hipStreamSynchronize(w->st));
			if ((rc = write_all(w->f, w->hbuf, bytes))) return w->error = rc;
		}
	} else {
		uint64_t nb = 0;
		if (!w->dbuf && (rc = ensure(*w, 1 << 20))) return w->error = rc;
		rc = gf_text_format_rows(w->oc.data(), (int)w->oc.size(), t->hidden, nrows, (char *)w->dbuf, w->dcap, &nb,
		                         &nv, w->st);
		if (rc == GF_E_CAPACITY) {
			if ((rc = ensure(*w, nb))) return w->error = rc;
			rc = gf_text_format_rows(w->oc.data(), (int)w->oc.size(), t->hidden, nrows, (char *)w->dbuf, w->dcap,
			                         &nb, &nv, w->st);
		}
		if (rc) return w->error = rc;
		if (nb) {
			GF_HIP(hipMemcpyAsync(w->hbuf, w->dbuf, nb, hipMemcpyDeviceToHost, w->st));
			GF_HIP(hipStreamSynchronize(w->st));
			if ((rc = write_all(w->f, w->hbuf, nb))) return w->error = rc;
		}
	}
	w->rows += nv;
	return 0;
}

int gf_writer_close(gf_writer *w, uint64_t *rows)
{
	if (!w) return GF_E_ARG;
	int rc = w->error;
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
