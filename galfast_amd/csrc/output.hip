// Output stages on the device: the visible rows of a chunk packed as FITS
// binary-table records (os_fitsout, pipeline.cpp:767-1035) or formatted as
// text lines (os_textout, pipeline.cpp:199-288, with otable's row format,
// otable.cpp:562-582 and otable.h:39-94).  Hidden rows are dropped, rows keep
// their order, `comp` values can be remapped (the sequential component index
// to the user's component id, osink::transformComponentIds).
//
// Text: every vector element is one field, fields are separated by one
// space and a line ends with '\n'; each field is what snprintf(fmt, v) makes
// of it for the formats the catalog's columns use: "%[ ]W.Pf" (a float is
// converted to double first, as a variadic argument is) and "%[ ]Wd".  The
// decimal digits come from exact integer arithmetic: v = m 2^e, the field's
// digits are round-half-even(|v| 10^P) (glibc's rounding of the exact
// binary value), for |v| < 1e10; a row with a larger value, a NaN or an
// infinity fails the call.
#include "gf_runtime.h"
#include "gf_device.h"

namespace {

constexpr int OB = 256;
constexpr int MAX_FIELD = 40;   // the longest field a supported format makes

__global__ void __launch_bounds__(OB) k_flags(gf_col hidden, uint64_t n, uint32_t *flag)
{
	uint64_t i = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (i < n) flag[i] = col<int>(hidden, (uint32_t)i) == 0 ? 1u : 0u;
}

GF_D uint32_t value_bits32(const gf_outcol &c, uint32_t row, uint32_t plane)
{
	uint32_t v = col<uint32_t>(c.col, row, plane);
	if (c.map && c.type == 'J') v = v < c.nmap ? c.map[v] : v;
	return v;
}

GF_D uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// FITS records of the block's visible rows, assembled in LDS and written as
// one contiguous run (the block's visible rows are consecutive in the
// output).
__global__ void __launch_bounds__(OB) k_fits_pack(const gf_outcol *cols, int ncols, uint32_t rec_words,
                                                  const uint64_t *pos, uint64_t n, uint32_t *out)
{
	extern __shared__ uint32_t rec[];
	const uint64_t r0 = (uint64_t)blockIdx.x * OB;
	const uint64_t r = r0 + threadIdx.x;
	const uint64_t p0 = pos[r0];
	const uint64_t p1 = pos[r0 + OB < n ? r0 + OB : n];
	if (r < n && pos[r + 1] > pos[r]) {
		uint32_t *dst = rec + (size_t)(pos[r] - p0) * rec_words;
		uint32_t w = 0;
		for (int k = 0; k < ncols; k++) {
			const gf_outcol &c = cols[k];
			for (uint32_t e = 0; e < c.col.width; e++) {
				if (c.col.elem == 8) {
					uint64_t v = col<uint64_t>(c.col, (uint32_t)r, e);
					dst[w++] = bswap32((uint32_t)(v >> 32));
					dst[w++] = bswap32((uint32_t)v);
				} else {
					dst[w++] = bswap32(value_bits32(c, (uint32_t)r, e));
				}
			}
		}
	}
	__syncthreads();
	const size_t words = (size_t)(p1 - p0) * rec_words;
	uint32_t *o = out + (size_t)p0 * rec_words;
	for (size_t i = threadIdx.x; i < words; i += OB) o[i] = rec[i];
}

// ---- text ---------------------------------------------------------------

struct Field {
	char s[MAX_FIELD];
	int len;
};

GF_D void put_uint(Field &f, uint64_t q, int mindigits)
{
	char d[24];
	int nd = 0;
	do {
		d[nd++] = (char)('0' + q % 10);
		q /= 10;
	} while (q || nd < mindigits);
	while (nd) f.s[f.len++] = d[--nd];
}

// The field of value v formatted as "%[ ]W.Pf" (prec >= 0) or "%[ ]Wd"
// (prec < 0); false when v is outside the supported range.
GF_D bool format_field(const gf_outcol &c, uint32_t row, uint32_t plane, Field &f)
{
	char body[MAX_FIELD];
	int blen = 0;
	bool neg;
	if (c.prec < 0) {
		int32_t v = (int32_t)value_bits32(c, row, plane);
		neg = v < 0;
		uint64_t a = neg ? (uint64_t)(-(int64_t)v) : (uint64_t)v;
		Field t;
		t.len = 0;
		put_uint(t, a, 1);
		for (int i = 0; i < t.len; i++) body[blen++] = t.s[i];
	} else {
		double v = c.type == 'D' ? col<double>(c.col, row, plane) : (double)col<float>(c.col, row, plane);
		uint64_t u = gfm::as_u64(v);
		neg = (u >> 63) != 0;
		const int bexp = (int)((u >> 52) & 0x7ff);
		uint64_t m = u & 0xfffffffffffffull;
		if (bexp == 0x7ff) return false;           // inf, NaN
		if (!(__builtin_fabs(v) < 1e10)) return false;
		int e;
		if (bexp == 0) { e = -1074; } else { m |= 1ull << 52; e = bexp - 1075; }
		// |v| 10^P = m 10^P 2^e, e < 0 here; N = m 10^P in 128 bits
		uint64_t p10 = 1;
		for (int i = 0; i < c.prec; i++) p10 *= 10;
		const uint64_t lo = m * p10, hi = __umul64hi(m, p10);
		const int s = -e;
		uint64_t q, rem_hi_bit, rem_lo_nonzero;   // q = N >> s; the bits shifted out
		if (s >= 128) {
			q = 0;
			rem_hi_bit = 0;   // N < 2^117 < 2^(s-1): below one half
			rem_lo_nonzero = 1;
		} else if (s >= 64) {
			q = s == 64 ? hi : hi >> (s - 64);
			// bit s-1 and the bits below it
			const int b = s - 1;
			rem_hi_bit = b >= 64 ? (hi >> (b - 64)) & 1 : (lo >> b) & 1;
			uint64_t below_hi = b >= 64 ? (b - 64 == 0 ? 0 : hi & ((1ull << (b - 64)) - 1)) : 0;
			uint64_t below_lo = b >= 64 ? lo : (b == 0 ? 0 : lo & ((1ull << b) - 1));
			rem_lo_nonzero = (below_hi | below_lo) != 0;
		} else if (s > 0) {
			q = (lo >> s) | (hi << (64 - s));
			rem_hi_bit = (lo >> (s - 1)) & 1;
			rem_lo_nonzero = s > 1 ? (lo & ((1ull << (s - 1)) - 1)) != 0 : 0;
		} else {
			q = lo;   // not reached for |v| < 1e10 with 53-bit m
			rem_hi_bit = 0;
			rem_lo_nonzero = 0;
		}
		if (rem_hi_bit && (rem_lo_nonzero || (q & 1))) q++;   // round half to even
		Field t;
		t.len = 0;
		put_uint(t, q / p10, 1);
		for (int i = 0; i < t.len; i++) body[blen++] = t.s[i];
		if (c.prec > 0) {
			body[blen++] = '.';
			t.len = 0;
			put_uint(t, q % p10, c.prec);
			for (int i = 0; i < t.len; i++) body[blen++] = t.s[i];
		}
	}
	const bool sign = neg || (c.flags & 1);
	const int len = blen + (sign ? 1 : 0);
	f.len = 0;
	for (int i = len; i < c.width; i++) f.s[f.len++] = ' ';
	if (sign) f.s[f.len++] = neg ? '-' : ' ';
	for (int i = 0; i < blen; i++) f.s[f.len++] = body[i];
	return true;
}

// Line length of every row (0 for a hidden row); bad[0] set when a value
// is outside the formatter's range.
__global__ void __launch_bounds__(OB) k_text_len(const gf_outcol *cols, int ncols, gf_col hidden, uint64_t n,
                                                 uint32_t *len, uint32_t *bad)
{
	uint64_t r = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (r >= n) return;
	if (col<int>(hidden, (uint32_t)r)) { len[r] = 0; return; }
	uint32_t l = 0;
	for (int k = 0; k < ncols; k++) {
		const gf_outcol &c = cols[k];
		for (uint32_t e = 0; e < c.col.width; e++) {
			Field f;
			if (!format_field(c, (uint32_t)r, e, f)) { atomicOr(bad, 1u); f.len = 0; }
			l += (uint32_t)f.len + 1;   // the field and a space or the newline
		}
	}
	len[r] = l;
}

__global__ void __launch_bounds__(OB) k_text_write(const gf_outcol *cols, int ncols, uint64_t n, const uint64_t *off,
                                                   char *out)
{
	uint64_t r = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (r >= n || off[r + 1] == off[r]) return;
	char *o = out + off[r];
	bool first = true;
	for (int k = 0; k < ncols; k++) {
		const gf_outcol &c = cols[k];
		for (uint32_t e = 0; e < c.col.width; e++) {
			Field f;
			if (!format_field(c, (uint32_t)r, e, f)) f.len = 0;
			if (!first) *o++ = ' ';
			first = false;
			for (int i = 0; i < f.len; i++) *o++ = f.s[i];
		}
	}
	*o = '\n';
}

int upload_cols(const gf_outcol *cols, int ncols, hipStream_t st, gf_outcol **dcols)
{
	void *p;
	int rc = gf::scratch(GF_SLOT_OUT_COLS, (size_t)ncols * sizeof(gf_outcol), &p);
	if (rc) return rc;
	// the caller's array may go away when the call returns
	GF_HIP(hipMemcpyAsync(p, cols, (size_t)ncols * sizeof(gf_outcol), hipMemcpyHostToDevice, st));
	GF_HIP(hipStreamSynchronize(st));
	*dcols = (gf_outcol *)p;
	return 0;
}

} // namespace

extern "C" {

int gf_fits_pack_rows(const gf_outcol *cols, int ncols, gf_col hidden, uint64_t nrows, void *out, uint64_t cap_rows,
                      uint64_t *nvisible, void *stream)
{
	GF_CHECK(cols && ncols > 0 && out && nvisible && nrows < (1ull << 32), GF_E_ARG);
	*nvisible = 0;
	if (nrows == 0) return 0;
	uint32_t rec = 0;
	for (int k = 0; k < ncols; k++) {
		GF_CHECK(cols[k].col.ptr && (cols[k].col.elem == 4 || cols[k].col.elem == 8), GF_E_ARG);
		rec += cols[k].col.width * cols[k].col.elem;
	}
	GF_CHECK(rec % 4 == 0 && rec <= 256, GF_E_ARG);   // 256 records of LDS within 64 KiB
	hipStream_t st = gf::as_stream(stream);
	void *p;
	int rc = gf::scratch(GF_SLOT_OUT, (size_t)nrows * 4 + ((size_t)nrows + 1) * 8 + 16, &p);
	if (rc) return rc;
	uint32_t *flag = (uint32_t *)p;
	uint64_t *pos = (uint64_t *)(((uintptr_t)(flag + nrows) + 7) & ~(uintptr_t)7);
	const unsigned g = gf::blocks_for(nrows, OB);
	k_flags<<<g, OB, 0, st>>>(hidden, nrows, flag);
	GF_LAUNCHED();
	rc = gf::exclusive_scan_u32_u64(flag, pos, nrows, st);
	if (rc) return rc;
	uint64_t nv = 0;
	GF_HIP(hipMemcpyAsync(&nv, pos + nrows, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	if (nv > cap_rows) return GF_E_CAPACITY;
	gf_outcol *dcols;
	if ((rc = upload_cols(cols, ncols, st, &dcols))) return rc;
	k_fits_pack<<<g, OB, (size_t)OB * rec, st>>>(dcols, ncols, rec / 4, pos, nrows, (uint32_t *)out);
	GF_LAUNCHED();
	*nvisible = nv;
	return 0;
}

int gf_text_format_rows(const gf_outcol *cols, int ncols, gf_col hidden, uint64_t nrows, char *out, uint64_t cap,
                        uint64_t *nbytes, uint64_t *nvisible, void *stream)
{
	GF_CHECK(cols && ncols > 0 && out && nbytes && nvisible && nrows < (1ull << 32), GF_E_ARG);
	*nbytes = *nvisible = 0;
	if (nrows == 0) return 0;
	for (int k = 0; k < ncols; k++) {
		const gf_outcol &c = cols[k];
		GF_CHECK(c.col.ptr && (c.type == 'E' || c.type == 'D' || c.type == 'J'), GF_E_ARG);
		GF_CHECK(c.prec <= 8 && c.width >= 0 && c.width <= 24, GF_E_ARG);
		GF_CHECK(c.prec >= 0 || c.type == 'J', GF_E_ARG);
		GF_CHECK(c.prec < 0 || c.type != 'J', GF_E_ARG);
	}
	hipStream_t st = gf::as_stream(stream);
	void *p;
	int rc = gf::scratch(GF_SLOT_OUT, (size_t)nrows * 4 + ((size_t)nrows + 1) * 8 + 32, &p);
	if (rc) return rc;
	uint32_t *len = (uint32_t *)p;
	uint64_t *off = (uint64_t *)(((uintptr_t)(len + nrows) + 7) & ~(uintptr_t)7);
	uint32_t *bad = (uint32_t *)(off + nrows + 1);
	gf_outcol *dcols;
	if ((rc = upload_cols(cols, ncols, st, &dcols))) return rc;
	GF_HIP(hipMemsetAsync(bad, 0, sizeof(uint32_t), st));
	const unsigned g = gf::blocks_for(nrows, OB);
	k_text_len<<<g, OB, 0, st>>>(dcols, ncols, hidden, nrows, len, bad);
	GF_LAUNCHED();
	rc = gf::exclusive_scan_u32_u64(len, off, nrows, st);
	if (rc) return rc;
	uint64_t total = 0;
	uint32_t hbad = 0;
	GF_HIP(hipMemcpyAsync(&total, off + nrows, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipMemcpyAsync(&hbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	if (hbad) return GF_E_ARG;
	if (total > cap) { *nbytes = total; return GF_E_CAPACITY; }
	k_text_write<<<g, OB, 0, st>>>(dcols, ncols, nrows, off, out);
	GF_LAUNCHED();
	*nbytes = total;
	// visible rows: lines written
	uint64_t nv = 0;
	{
		// count lines: rows with a non-zero length
		void *q;
		if ((rc = gf::scratch(GF_SLOT_OUT_COUNT, sizeof(unsigned long long), &q))) return rc;
		GF_HIP(hipMemsetAsync(q, 0, sizeof(unsigned long long), st));
		if ((rc = gf::count_visible(hidden, nrows, (unsigned long long *)q, st))) return rc;
		GF_HIP(hipMemcpyAsync(&nv, q, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
		GF_HIP(hipStreamSynchronize(st));
	}
	*nvisible = nv;
	return 0;
}

} // extern "C"
