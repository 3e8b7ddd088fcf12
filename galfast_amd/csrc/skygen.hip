// Sky generator kernels: expected-count integral, reference-order draw,
// and the cumulative-table draw (plan + emit).
//
// Cell geometry (skygen/skygen.cu.h:216-229, 541-565): pixel p, apparent bin
// im, absolute bin iM; M = M1 - iM*dM, m = m0 + im*dm, and because dm == dM
// the distance modulus depends only on k = im + iM.  A (pixel, k) pair is a
// "DM bin"; its density rho(xyz(D_k)) is shared by all its cells, and a cell
// contributes LF(M) * rho * norm * D^3 * dA * POGSON * dm * dM.
#include "gf_runtime.h"
#include "gf_device.h"
#include <mutex>

namespace {

constexpr unsigned SK_BS = 256;

// ---------------------------------------------------------------------------
// Luminosity-function tables of one model: LFc[iM] = LF(M1 - iM*dM) and its
// exclusive prefix sums in the three accumulation arithmetics.
// ---------------------------------------------------------------------------
struct LFTables {
	float *lfc;     // nM
	double *P;      // nM + 1, FP64 prefix
	float *Pk;      // nM + 1, Kahan-compensated FP32 prefix
	float *Pf;      // nM + 1, naive FP32 prefix
};

__global__ void k_lf_tables(gf_skygen_params sp, gf_model m, LFTables t)
{
	for (int iM = threadIdx.x; iM < sp.nM; iM += blockDim.x) {
		float M = sp.M1 - iM * sp.dM;
		t.lfc[iM] = tex1d(m.lf, M);
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double s = 0.0;
		float sk = 0.f, ck = 0.f, sf = 0.f;
		t.P[0] = 0.0; t.Pk[0] = 0.f; t.Pf[0] = 0.f;
		for (int iM = 0; iM < sp.nM; iM++) {
			float v = t.lfc[iM];
			s += (double)v;
			float y = v - ck;
			float tt = sk + y;
			ck = (tt - sk) - y;
			sk = tt;
			sf += v;
			t.P[iM + 1] = s; t.Pk[iM + 1] = sk; t.Pf[iM + 1] = sf;
		}
	}
}

// Bytes of the LF tables of nM absolute-magnitude bins.
size_t lf_tables_bytes(int nM)
{
	size_t n = (size_t)nM + 1;
	return n * 4 + n * 8 + n * 4 + n * 4 + 64;
}

// The LF tables in the buffer p (lf_tables_bytes(sp.nM) bytes).
int lf_tables_at(const gf_skygen_params &sp, const gf_model &m, void *p, LFTables &t, hipStream_t st)
{
	size_t n = (size_t)sp.nM + 1;
	char *c = (char *)p;
	t.P = (double *)c; c += n * 8;
	t.lfc = (float *)c; c += n * 4;
	t.Pk = (float *)c; c += n * 4;
	t.Pf = (float *)c;
	k_lf_tables<<<1, SK_BS, 0, st>>>(sp, m, t);
	GF_LAUNCHED();
	return 0;
}

int lf_tables(const gf_skygen_params &sp, const gf_model &m, LFTables &t, hipStream_t st)
{
	void *p;
	int rc = gf::scratch(GF_SLOT_LFTAB, lf_tables_bytes(sp.nM), &p);
	if (rc) return rc;
	return lf_tables_at(sp, m, p, t, st);
}

// Distance of DM bin k: compute_pos evaluated at the diagonal's first cell in
// crawl order (largest iM), which is where the reference's advance() enters a
// diagonal.
GF_D void dm_bin(const gf_skygen_params &sp, const gf_extinction &ext, const gf_pencil_beam &pix, int k,
                 float &D, float &Am, float &x, float &y, float &z)
{
	int iM0 = k < sp.nM - 1 ? k : sp.nM - 1;
	int im0 = k - iM0;
	float M = sp.M1 - iM0 * sp.dM;
	cell_pos(sp, ext, pix, M, im0, D, Am, x, y, z);
}

// ---------------------------------------------------------------------------
// Count integral: one workgroup per pixel, lanes over DM bins.
// ---------------------------------------------------------------------------
struct CountAcc {
	double c, cov, mx, ncell;
};

template <int MODE>
struct Acc;   // per-mode running sums

template <> struct Acc<GF_ACC_FP64> {
	double c = 0, cov = 0;
	GF_D void add(double w, float cf) { c += w; cov += w * (double)cf; }
	GF_D double total() const { return c; }
	GF_D double covered() const { return cov; }
};
template <> struct Acc<GF_ACC_KAHAN32> {
	float c = 0, cc = 0, cov = 0, ccov = 0;
	GF_D static void kadd(float &s, float &comp, float v) { float y = v - comp; float t = s + y; comp = (t - s) - y; s = t; }
	GF_D void add(double w, float cf) { float wf = (float)w; kadd(c, cc, wf); kadd(cov, ccov, wf * cf); }
	GF_D double total() const { return (double)c; }
	GF_D double covered() const { return (double)cov; }
};
template <> struct Acc<GF_ACC_FP32> {
	float c = 0, cov = 0;
	GF_D void add(double w, float cf) { float wf = (float)w; c += wf; cov += wf * cf; }
	GF_D double total() const { return (double)c; }
	GF_D double covered() const { return (double)cov; }
};

// Expectation of DM bin (pixel, k); returns -1 for an empty window.
template <int MODE>
GF_D double bin_expectation(const gf_skygen_params &sp, const gf_model &m, const gf_extinction &ext,
                            const gf_pencil_beam &pix, const LFTables &t, int k, int &ncell)
{
	float D, Am, x, y, z;
	dm_bin(sp, ext, pix, k, D, Am, x, y, z);
	int lo, hi;
	ncell = 0;
	if (!dm_window(sp, k, D, Am, lo, hi)) return -1.0;
	ncell = hi - lo + 1;
	float R = model_rho_xyz(m, x, y, z);
	if (MODE == GF_ACC_FP64) {
		double vol = (double)pix.dA * (double)POGSON * (double)sp.dm * (double)sp.dM;
		double fac = (double)R * (double)m.norm * ((double)D * (double)D * (double)D) * vol;
		return fac * (t.P[hi + 1] - t.P[lo]);
	} else {
		float rho = R * m.norm;
		rho *= D * D * D;
		rho *= pix.dA * POGSON * sp.dm * sp.dM;
		float S = MODE == GF_ACC_KAHAN32 ? (t.Pk[hi + 1] - t.Pk[lo]) : (t.Pf[hi + 1] - t.Pf[lo]);
		return (double)(rho * S);
	}
}

template <int MODE>
__global__ void __launch_bounds__(SK_BS) k_count(gf_skygen_params sp, gf_model m, const gf_pencil_beam *pixels,
                                                 gf_extinction ext, LFTables t, CountAcc *pix_out,
                                                 double *per_pixel, double *wpk)
{
	const int p = blockIdx.x;
	const int K = sp.nm + sp.nM - 1;
	const gf_pencil_beam pix = pixels[p];
	Acc<MODE> acc;
	double mx = 0.0, nc = 0.0;
	for (int k = threadIdx.x; k < K; k += blockDim.x) {
		int ncell;
		double w = bin_expectation<MODE>(sp, m, ext, pix, t, k, ncell);
		if (wpk) wpk[(size_t)p * K + k] = w;
		if (w >= 0.0) {
			acc.add(w, pix.coveredFraction);
			if (w > mx) mx = w;
			nc += (double)ncell;
		}
	}
	__shared__ double sc[SK_BS], scov[SK_BS], smx[SK_BS], snc[SK_BS];
	sc[threadIdx.x] = acc.total();
	scov[threadIdx.x] = acc.covered();
	smx[threadIdx.x] = mx;
	snc[threadIdx.x] = nc;
	__syncthreads();
	for (int o = SK_BS / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) {
			if (MODE == GF_ACC_FP64) {
				sc[threadIdx.x] += sc[threadIdx.x + o];
				scov[threadIdx.x] += scov[threadIdx.x + o];
			} else {
				sc[threadIdx.x] = (double)((float)sc[threadIdx.x] + (float)sc[threadIdx.x + o]);
				scov[threadIdx.x] = (double)((float)scov[threadIdx.x] + (float)scov[threadIdx.x + o]);
			}
			smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + o]);
			snc[threadIdx.x] += snc[threadIdx.x + o];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		pix_out[p] = CountAcc{sc[0], scov[0], smx[0], snc[0]};
		if (per_pixel) per_pixel[p] = scov[0];
	}
}

template <int MODE>
__global__ void __launch_bounds__(SK_BS) k_count_final(const CountAcc *pix_out, int np, gf_count_result *res)
{
	// the uncovered total is reduced in the mode's arithmetic (the precision
	// study compares it); covered totals, maxima and cell counts in FP64
	Acc<MODE> acc;
	double cov = 0.0, mx = 0.0, nc = 0.0;
	for (int p = threadIdx.x; p < np; p += blockDim.x) {
		CountAcc a = pix_out[p];
		acc.add(a.c, 1.0f);
		cov += a.cov;
		mx = fmax(mx, a.mx);
		nc += a.ncell;
	}
	__shared__ double sc[SK_BS], scov[SK_BS], smx[SK_BS], snc[SK_BS];
	sc[threadIdx.x] = acc.total();
	scov[threadIdx.x] = cov;
	smx[threadIdx.x] = mx;
	snc[threadIdx.x] = nc;
	__syncthreads();
	for (int o = SK_BS / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) {
			if (MODE == GF_ACC_FP64) sc[threadIdx.x] += sc[threadIdx.x + o];
			else sc[threadIdx.x] = (double)((float)sc[threadIdx.x] + (float)sc[threadIdx.x + o]);
			scov[threadIdx.x] += scov[threadIdx.x + o];
			smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + o]);
			snc[threadIdx.x] += snc[threadIdx.x + o];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		res->nstarsExpectedToGenerate = sc[0];
		res->nstarsExpected = scov[0];
		res->maxCount = smx[0];
		res->ncells = snc[0];
	}
}

// Count result of an empty pixel list.
__global__ void k_count_empty(gf_count_result *res)
{
	*res = gf_count_result{0, 0, 0, 0};
}

// The count kernels in the accumulation arithmetic acc_mode: per-pixel sums
// into pix_out (npixels entries), the job's totals into *dres (device).
int launch_count(const gf_skygen_params &sp, const gf_model &m, const gf_pencil_beam *pixels,
                 const gf_extinction &ex, int acc_mode, const LFTables &t, CountAcc *pix_out,
                 gf_count_result *dres, double *per_pixel, double *wpk, hipStream_t st)
{
	switch (acc_mode) {
	case GF_ACC_FP64:
		k_count<GF_ACC_FP64><<<sp.npixels, SK_BS, 0, st>>>(sp, m, pixels, ex, t, pix_out, per_pixel, wpk);
		GF_LAUNCHED();
		k_count_final<GF_ACC_FP64><<<1, SK_BS, 0, st>>>(pix_out, sp.npixels, dres);
		break;
	case GF_ACC_KAHAN32:
		k_count<GF_ACC_KAHAN32><<<sp.npixels, SK_BS, 0, st>>>(sp, m, pixels, ex, t, pix_out, per_pixel, wpk);
		GF_LAUNCHED();
		k_count_final<GF_ACC_KAHAN32><<<1, SK_BS, 0, st>>>(pix_out, sp.npixels, dres);
		break;
	default:
		k_count<GF_ACC_FP32><<<sp.npixels, SK_BS, 0, st>>>(sp, m, pixels, ex, t, pix_out, per_pixel, wpk);
		GF_LAUNCHED();
		k_count_final<GF_ACC_FP32><<<1, SK_BS, 0, st>>>(pix_out, sp.npixels, dres);
		break;
	}
	GF_LAUNCHED();
	return 0;
}

// Layout of gf_skygen_count_noalloc's scratch: the LF tables of up to
// GF_NOALLOC_MAX_NM bins, then the per-pixel partial sums of up to
// GF_NOALLOC_MAX_PIXELS pixels.
constexpr size_t NOALLOC_LF_BYTES = ((((size_t)GF_NOALLOC_MAX_NM + 1) * 20 + 64) + 255) & ~(size_t)255;
constexpr size_t NOALLOC_BYTES = NOALLOC_LF_BYTES + (size_t)GF_NOALLOC_MAX_PIXELS * sizeof(CountAcc);

// ---------------------------------------------------------------------------
// Diagonal crawl of the reference (skygen/skygen.cu.h:231-358).
// ---------------------------------------------------------------------------
GF_D void diag_index_to_ij(int &ilb, int &i, int &j, int &k, const int x, const int y)
{
	int kmax = x * y;
	if (k >= kmax) {
		ilb += k / kmax;
		k %= kmax;
	}
	int l = x < y ? x : y;
	int m = x < y ? y : x;
	int d;
	if (2 * k < l * (l - 1)) {
		d = int(0.5 * (double)(gf_sqrtf((float)(1 + 8 * k)) - 1));
		i = k - d * (d + 1) / 2;
	} else if (2 * k < l * (2 * m + 1 - l)) {
		if (x >= y) {
			int ka = k - y * (y - 1) / 2;
			d = int(ka / y) + (y - 1);
			i = (ka % y) + d - (y - 1);
		} else {
			int ka = k - x * (x - 1) / 2;
			d = int(ka / x) + (x - 1);
			i = ka % x;
		}
	} else {
		int ka = x * y - k - 1;
		int dd = int(0.5 * (double)(gf_sqrtf((float)(1 + 8 * ka)) - 1));
		d = (x + y - 2) - dd;
		i = (x - 1) - (ka - dd * (dd + 1) / 2);
	}
	j = d - i;
}

GF_D bool crawl_advance(int &ilb, int &i, int &j, gf_pencil_beam &pix, const gf_pencil_beam *pixels,
                        int npixels, const int x, const int y)
{
	i++; j--;
	if (j < 0) {
		if (i < y) { j += i + 1; i = 0; }
		else { i -= (y - 1); j += y; }
	} else if (i >= x) {
		if (x > y) { i -= y - j - 2; j = y - 1; }
		else {
			if (j + 1 < y - x) { i -= x; j += x + 1; }
			else { i = x - y + j + 2; j = y - 1; }
		}
		if (i == x) {
			ilb++;
			if (ilb != npixels) pix = pixels[ilb];
		}
	} else {
		return false;
	}
	return true;
}

// Destination of one generated star.
struct StarSink {
	gf_skygen_cols c;
	bool write;
	uint64_t row;
};

// One star of draw_stars with the clipper's footprint test.
GF_D void draw_one_star(const gf_skygen_params &sp, const gf_model &m, const gf_extinction &ext,
                        const gf_footprint &foot, const gf_pencil_beam &pix, float M, int im, float AmMin,
                        gf_mwc &rng, StarSink &s)
{
	float x, y, Mtmp, mtmp;
	const uint32_t row = (uint32_t)s.row;
	if (!star_draws(sp, pix, M, im, rng, x, y, Mtmp, mtmp)) {
		if (s.write) write_star_outside(s.c, row, pix, m.comp, x, y);
		return;
	}
	if (s.write) write_star(sp, m.comp, ext, foot, pix, x, y, Mtmp, mtmp, AmMin, s.c, row);
}

// kernel<1> of the reference with unlimited output capacity.  Pass 0 counts
// the stars of each thread, pass 1 regenerates them from the same stream
// state and writes them at the thread's offset.
__global__ void __launch_bounds__(64) k_draw_compat(gf_skygen_params sp, gf_model m, const gf_pencil_beam *pixels,
                                                    gf_extinction ext, gf_footprint foot, uint32_t *st, uint32_t ns,
                                                    int pass, uint32_t *counts, const uint64_t *offs,
                                                    gf_skygen_cols out)
{
	const int tid = blockIdx.x * blockDim.x + threadIdx.x;
	if (tid >= sp.nthreads) return;
	const int block = 10;
	const int nthreads = sp.nthreads, nm = sp.nm, nM = sp.nM;
	int ilb = 0, im = 0, iM = 0, k = block * (tid - nthreads), bc = 0;
	float D = 0.f, Am = 0.f, R = 0.f;
	gf_pencil_beam pix;
	gf_mwc rng = gf_mwc_load(st, ns, tid);
	StarSink s{out, pass == 1, pass == 1 ? offs[tid] : 0};
	uint32_t n = 0;
	for (;;) {
		bool moved;
		if (bc == 0) {
			bc = block;
			k += block * nthreads;
			diag_index_to_ij(ilb, im, iM, k, nm, nM);
			if (ilb >= sp.npixels) break;
			pix = pixels[ilb];
			moved = true;
		} else {
			moved = crawl_advance(ilb, im, iM, pix, pixels, sp.npixels, nm, nM);
		}
		bc--;
		float M = sp.M1 - iM * sp.dM;
		if (moved) {
			if (ilb >= sp.npixels) break;
			float x, y, z;
			cell_pos(sp, ext, pix, M, im, D, Am, x, y, z);
			R = model_rho_xyz(m, x, y, z);
		}
		if ((sp.dmin != 0.f || sp.dmax != 0.f) && (sp.dmin > D || sp.dmax <= D)) continue;
		float mm = sp.m0 + sp.dm * im + Am;
		if (mm > sp.m1) continue;
		float phi = tex1d(m.lf, M);
		float rho = phi * R;
		rho *= m.norm;
		rho *= D * D * D;
		rho *= pix.dA * POGSON * sp.dm * sp.dM;
		unsigned int ndraw = rng.poisson(rho);
		for (unsigned int q = 0; q < ndraw; q++) {
			draw_one_star(sp, m, ext, foot, pix, M, im, Am, rng, s);
			s.row++;
			n++;
		}
	}
	if (pass == 0) counts[tid] = n;
	else gf_mwc_store(st, ns, tid, rng);
}

// ---------------------------------------------------------------------------
// Cumulative-table draw.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(SK_BS) k_plan(const double *wpk, uint64_t U, uint32_t *st, uint32_t ns,
                                                uint32_t *counts)
{
	uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= ns) return;
	__builtin_amdgcn_s_setprio(3);   // the catalog's serial stream first among resident waves
	gf_mwc rng = gf_mwc_load(st, ns, s);
	// two units per trip, each unit's weight loaded a trip ahead (one wave
	// per SIMD has nothing else to hide a load's latency behind); the two
	// weights keep their registers, so no copy waits on a load in flight,
	// and every trip issues the same loads (indices clamped to the last
	// unit), so the waits count them exactly
	const uint64_t step = 2ull * ns, last = U - 1;
	double wa = wpk[s < last ? s : last];
	double wb = wpk[s + (uint64_t)ns < last ? s + ns : last];
	for (uint64_t u = s; u < U; u += step) {
		uint32_t n = 0;
		if (wa > 0.0) n = rng.poisson((float)wa);
		counts[u] = n;
		wa = wpk[u + step < last ? u + step : last];
		const uint64_t v = u + ns;
		if (v >= U) break;
		n = 0;
		if (wb > 0.0) n = rng.poisson((float)wb);
		counts[v] = n;
		wb = wpk[v + step < last ? v + step : last];
	}
	gf_mwc_store(st, ns, s, rng);
}

constexpr int EMIT_MAX_NM = 4096;

// LDS of the emission: the LF prefix (nM + 1 doubles), the guess table
// (LF_GUESS uint16) and, 16-byte aligned, the guess thresholds.
GF_HD size_t lf_pairs_offset(int nM)
{
	const size_t b = ((size_t)nM + 1) * sizeof(double) + LF_GUESS * sizeof(uint16_t);
	return (b + 15) & ~(size_t)15;
}

__global__ void __launch_bounds__(SK_BS) k_emit(gf_skygen_params sp, gf_model m, const gf_pencil_beam *pixels,
                                                gf_extinction ext, gf_footprint foot, const double *Pg,
                                                const uint64_t *offs, uint64_t U, uint64_t row_begin,
                                                uint64_t row_end, uint64_t u_lo, uint32_t *st, uint32_t ns,
                                                gf_skygen_cols out)
{
	extern __shared__ double P[];
	for (int i = threadIdx.x; i <= sp.nM; i += blockDim.x) P[i] = Pg[i];
	__syncthreads();
	uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= ns) return;
	const int K = sp.nm + sp.nM - 1;
	uint64_t first = u_lo + ((uint64_t)s + ns - (u_lo % ns)) % ns;
	gf_mwc rng = gf_mwc_load(st, ns, s);
	bool touched = false;
	for (uint64_t u = first; u < U; u += ns) {
		uint64_t o0 = offs[u], o1 = offs[u + 1];
		if (o0 >= row_end) break;
		if (o1 <= row_begin || o1 == o0) continue;
		touched = true;
		int p = (int)(u / K), k = (int)(u % K);
		const gf_pencil_beam pix = pixels[p];
		float D, Am, x, y, z;
		dm_bin(sp, ext, pix, k, D, Am, x, y, z);
		int lo, hi;
		dm_window(sp, k, D, Am, lo, hi);
		double base = P[lo], span = P[hi + 1] - P[lo];
		uint64_t j0 = row_begin > o0 ? row_begin : o0;
		uint64_t j1 = row_end < o1 ? row_end : o1;
		StarSink snk{out, true, j0 - row_begin};
		for (uint64_t j = j0; j < j1; j++) {
			// absolute-magnitude bin from the cumulative LF window
			double tgt = base + (double)rng.uniform() * span;
			int a = lo, b = hi;
			while (a < b) {
				int mid = (a + b) >> 1;
				if (P[mid + 1] > tgt) b = mid; else a = mid + 1;
			}
			int iM = a;
			int im = k - iM;
			float M = sp.M1 - iM * sp.dM;
			draw_one_star(sp, m, ext, foot, pix, M, im, Am, rng, snk);
			snk.row++;
		}
	}
	if (touched) gf_mwc_store(st, ns, s, rng);
}

// Random half of the emit for maps without extinction: per star the
// absolute-magnitude bin and the star_draws values, parked in the output
// columns (projXY = x, y; M = Mtmp; DM = mtmp; comp = pixel index;
// hidden = 1 outside the projection).  k_emit_rows finishes every row.
__global__ void __launch_bounds__(SK_BS) k_emit_rng(gf_skygen_params sp, gf_model m, const gf_pencil_beam *pixels,
                                                    gf_extinction ext, const double *Pg, const uint64_t *offs,
                                                    uint64_t U, uint64_t row_begin, uint64_t row_end, uint64_t u_lo,
                                                    uint32_t *st, uint32_t ns, gf_parked park)
{
	extern __shared__ double P[];
	uint16_t *G = (uint16_t *)(P + sp.nM + 1);
	double2 *GP = (double2 *)((char *)P + lf_pairs_offset(sp.nM));
	for (int i = threadIdx.x; i <= sp.nM; i += blockDim.x) P[i] = Pg[i];
	__syncthreads();
	lf_guess_build(P, sp.nM, G, threadIdx.x, blockDim.x);
	__syncthreads();
	lf_guess_pairs(P, G, GP, threadIdx.x, blockDim.x);
	__syncthreads();
	const double gscale = lf_guess_scale(P, sp.nM);
	uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= ns) return;
	const int K = sp.nm + sp.nM - 1;
	uint64_t first = u_lo + ((uint64_t)s + ns - (u_lo % ns)) % ns;
	__builtin_amdgcn_s_setprio(3);   // one wave per SIMD: first among the resident waves
	gf_mwc rng = gf_mwc_load(st, ns, s);
	bool touched = false;
	for (uint64_t u = first; u < U; u += ns) {
		uint64_t o0 = offs[u], o1 = offs[u + 1];
		if (o0 >= row_end) break;
		if (o1 <= row_begin || o1 == o0) continue;
		touched = true;
		int p = (int)(u / K), k = (int)(u % K);
		const gf_pencil_beam pix = pixels[p];
		float D, Am, x, y, z;
		dm_bin(sp, ext, pix, k, D, Am, x, y, z);
		int lo, hi;
		dm_window(sp, k, D, Am, lo, hi);
		double base = P[lo], span = P[hi + 1] - P[lo];
		uint64_t j0 = row_begin > o0 ? row_begin : o0;
		uint64_t j1 = row_end < o1 ? row_end : o1;
		for (uint64_t j = j0; j < j1; j++) {
			double tgt = base + (double)rng.uniform() * span;
			int iM = lf_bin_pairs(P, G, GP, gscale, lo, hi, tgt);
			float M = sp.M1 - iM * sp.dM;
			float sx, sy, Mtmp = 0.f, mtmp = 0.f;
			bool in = star_draws(sp, pix, M, k - iM, rng, sx, sy, Mtmp, mtmp);
			uint64_t row = j - row_begin;
			park.xyMm[row] = make_float4(sx, sy, Mtmp, mtmp);
			park.pix[row] = (uint32_t)p | (in ? 0u : 0x80000000u);
		}
	}
	if (touched) gf_mwc_store(st, ns, s, rng);
}

// Deterministic half of the emit, one lane per row.  FULL: the footprint
// test handles unions and equatorial rects.
template <bool FULL>
__global__ void __launch_bounds__(SK_BS) k_emit_rows(gf_skygen_params sp, int comp, const gf_pencil_beam *pixels,
                                                     gf_extinction ext, gf_footprint foot, uint32_t nrows,
                                                     gf_parked park, gf_skygen_cols out)
{
	uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= nrows) return;
	const float4 v = park.xyMm[row];
	const uint32_t ph = park.pix[row];
	const gf_pencil_beam pix = pixels[ph & 0x7fffffffu];
	if (ph & 0x80000000u) {
		write_star_outside(out, row, pix, comp, v.x, v.y);
		return;
	}
	write_star<FULL>(sp, comp, ext, foot, pix, v.x, v.y, v.z, v.w, 0.f, out, row);
}

__global__ void k_find_unit(const uint64_t *offs, uint64_t U, uint64_t row, uint64_t *res)
{
	// largest u with offs[u] <= row (u < U)
	uint64_t a = 0, b = U;   // offs[a] <= row; search in [a, b)
	while (b - a > 1) {
		uint64_t mid = (a + b) / 2;
		if (offs[mid] <= row) a = mid; else b = mid;
	}
	*res = a;
}

// The unit of row q * chunk for q < nq (the first unit of every output chunk).
__global__ void k_find_units(const uint64_t *offs, uint64_t U, uint64_t chunk, uint32_t nq, uint64_t *res)
{
	uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= nq) return;
	uint64_t row = (uint64_t)q * chunk;
	uint64_t a = 0, b = U;
	while (b - a > 1) {
		uint64_t mid = (a + b) / 2;
		if (offs[mid] <= row) a = mid; else b = mid;
	}
	res[q] = a;
}

} // namespace

struct gf_draw_plan {
	gf_skygen_params sp;
	gf_model model;
	const gf_pencil_beam *pixels;
	gf_extinction ext;
	uint64_t U;
	uint64_t total;
	uint32_t *counts;
	uint64_t *offs;
	double *P;
	uint64_t *tmp;
};

extern "C" {

int gf_skygen_count(const gf_skygen_params *sp, const gf_model *model,
                    const gf_pencil_beam *pixels, const gf_extinction *ext,
                    int acc_mode, gf_count_result *out, double *per_pixel,
                    double *wpk, void *stream)
{
	GF_CHECK(sp && model && pixels && out && sp->npixels >= 0 && sp->nm > 0 && sp->nM > 0, GF_E_ARG);
	GF_CHECK(acc_mode >= 0 && acc_mode <= 2, GF_E_ARG);
	hipStream_t st = gf::as_stream(stream);
	gf_extinction ex{};
	if (ext) ex = *ext;
	if (sp->npixels == 0) { *out = gf_count_result{0, 0, 0, 0}; return 0; }
	LFTables t;
	int rc = lf_tables(*sp, *model, t, st);
	if (rc) return rc;
	void *p;
	rc = gf::scratch(GF_SLOT_PARTIALS, (size_t)sp->npixels * sizeof(CountAcc) + sizeof(gf_count_result) + 64, &p);
	if (rc) return rc;
	CountAcc *pix_out = (CountAcc *)p;
	gf_count_result *dres = (gf_count_result *)((char *)p + (size_t)sp->npixels * sizeof(CountAcc));
	rc = launch_count(*sp, *model, pixels, ex, acc_mode, t, pix_out, dres, per_pixel, wpk, st);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(out, dres, sizeof(gf_count_result), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	return 0;
}

int gf_skygen_count_noalloc(const gf_skygen_params *sp, const gf_model *model,
                            const gf_pencil_beam *pixels, const gf_extinction *ext,
                            int acc_mode, gf_count_result *out_dev, double *per_pixel,
                            double *wpk, void *stream)
{
	GF_CHECK(sp && model && pixels && out_dev && sp->npixels >= 0 && sp->nm > 0 && sp->nM > 0, GF_E_ARG);
	GF_CHECK(acc_mode >= 0 && acc_mode <= 2, GF_E_ARG);
	GF_CHECK(sp->npixels <= GF_NOALLOC_MAX_PIXELS && sp->nM <= GF_NOALLOC_MAX_NM, GF_E_SCRATCH);
	void *base;
	int rc = gf::noalloc_scratch(&base);
	if (rc) return rc;
	hipStream_t st = gf::as_stream(stream);
	gf_extinction ex{};
	if (ext) ex = *ext;
	// the LF tables first, then the per-pixel partial sums (256-byte aligned)
	LFTables t;
	rc = lf_tables_at(*sp, *model, base, t, st);
	if (rc) return rc;
	CountAcc *pix_out = (CountAcc *)((char *)base + NOALLOC_LF_BYTES);
	if (sp->npixels == 0) {
		k_count_empty<<<1, 1, 0, st>>>(out_dev);
		GF_LAUNCHED();
		return 0;
	}
	return launch_count(*sp, *model, pixels, ex, acc_mode, t, pix_out, out_dev, per_pixel, wpk, st);
}

int gf_skygen_draw_compat(const gf_skygen_params *sp, const gf_model *model,
                          const gf_pencil_beam *pixels, const gf_extinction *ext,
                          const gf_footprint *foot, gf_rng *rng,
                          const gf_skygen_cols *out, uint64_t capacity,
                          uint64_t *nstars, void *stream)
{
	GF_CHECK(sp && model && pixels && rng && rng->state && out && nstars, GF_E_ARG);
	GF_CHECK(sp->nthreads > 0 && (uint32_t)sp->nthreads <= rng->nstreams, GF_E_STREAMS);
	hipStream_t st = gf::as_stream(stream);
	gf_extinction ex{};
	if (ext) ex = *ext;
	gf_footprint fp{};
	if (foot) fp = *foot;
	const int nt = sp->nthreads;
	void *pc, *po;
	int rc = gf::scratch(GF_SLOT_COMPAT_COUNTS, (size_t)nt * sizeof(uint32_t), &pc);
	if (rc) return rc;
	rc = gf::scratch(GF_SLOT_COMPAT_STATE, ((size_t)nt + 1) * sizeof(uint64_t), &po);
	if (rc) return rc;
	uint32_t *counts = (uint32_t *)pc;
	uint64_t *offs = (uint64_t *)po;
	unsigned grid = gf::blocks_for(nt, 64);
	k_draw_compat<<<grid, 64, 0, st>>>(*sp, *model, pixels, ex, fp, rng->state, rng->nstreams, 0, counts, offs, *out);
	GF_LAUNCHED();
	rc = gf::exclusive_scan_u32_u64(counts, offs, nt, st);
	if (rc) return rc;
	uint64_t total = 0;
	GF_HIP(hipMemcpyAsync(&total, offs + nt, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	*nstars = total;
	if (total > capacity) return GF_E_CAPACITY;
	k_draw_compat<<<grid, 64, 0, st>>>(*sp, *model, pixels, ex, fp, rng->state, rng->nstreams, 1, counts, offs, *out);
	GF_LAUNCHED();
	return 0;
}

int gf_skygen_plan(const gf_skygen_params *sp, const gf_model *model,
                   const gf_pencil_beam *pixels, const gf_extinction *ext,
                   const double *wpk, gf_rng *rng, gf_draw_plan **plan,
                   uint64_t *nstars, void *stream)
{
	GF_CHECK(sp && model && pixels && wpk && rng && rng->state && plan && nstars, GF_E_ARG);
	GF_CHECK(sp->nM <= EMIT_MAX_NM, GF_E_ARG);
	hipStream_t st = gf::as_stream(stream);
	gf_draw_plan *pl = new gf_draw_plan{};
	pl->sp = *sp;
	pl->model = *model;
	pl->pixels = pixels;
	if (ext) pl->ext = *ext;
	const uint64_t K = (uint64_t)(sp->nm + sp->nM - 1);
	pl->U = (uint64_t)sp->npixels * K;
	hipError_t e = hipMalloc((void **)&pl->counts, (pl->U + 1) * sizeof(uint32_t));
	if (e == hipSuccess) e = hipMalloc((void **)&pl->offs, (pl->U + 1) * sizeof(uint64_t));
	if (e == hipSuccess) e = hipMalloc((void **)&pl->P, ((size_t)sp->nM + 1) * sizeof(double));
	if (e == hipSuccess) e = hipMalloc((void **)&pl->tmp, 2 * sizeof(uint64_t));
	if (e != hipSuccess) { gf_skygen_plan_free(pl); return (int)e; }
	LFTables t;
	int rc = lf_tables(*sp, *model, t, st);
	if (rc) { gf_skygen_plan_free(pl); return rc; }
	e = hipMemcpyAsync(pl->P, t.P, ((size_t)sp->nM + 1) * sizeof(double), hipMemcpyDeviceToDevice, st);
	if (e != hipSuccess) { gf_skygen_plan_free(pl); return (int)e; }
	if (pl->U > 0) {
		k_plan<<<gf::blocks_for(rng->nstreams, SK_BS), SK_BS, 0, st>>>(wpk, pl->U, rng->state, rng->nstreams, pl->counts);
		e = hipGetLastError();
		if (e != hipSuccess) { gf_skygen_plan_free(pl); return (int)e; }
	}
	rc = gf::exclusive_scan_u32_u64(pl->counts, pl->offs, pl->U, st);
	if (rc) { gf_skygen_plan_free(pl); return rc; }
	e = hipMemcpyAsync(&pl->total, pl->offs + pl->U, sizeof(uint64_t), hipMemcpyDeviceToHost, st);
	if (e == hipSuccess) e = hipStreamSynchronize(st);
	if (e != hipSuccess) { gf_skygen_plan_free(pl); return (int)e; }
	*nstars = pl->total;
	*plan = pl;
	return 0;
}

int gf_skygen_emit(const gf_draw_plan *plan, uint64_t row_begin, uint64_t row_end,
                   const gf_footprint *foot, gf_rng *rng,
                   const gf_skygen_cols *out, void *stream)
{
	return gf::skygen_emit(plan, row_begin, row_end, foot, rng, out, gf::as_stream(stream), nullptr, nullptr,
	                       UINT64_MAX);
}

} // extern "C"

namespace gf {

int skygen_chunk_units(const gf_draw_plan *plan, uint64_t chunk, uint32_t nq, uint64_t *units, hipStream_t st)
{
	GF_CHECK(plan && units && chunk > 0, GF_E_ARG);
	if (!nq) return 0;
	void *p;
	int rc = gf::scratch(GF_SLOT_UNITS, (size_t)nq * sizeof(uint64_t), &p);
	if (rc) return rc;
	k_find_units<<<gf::blocks_for(nq, 256), 256, 0, st>>>(plan->offs, plan->U, chunk, nq, (uint64_t *)p);
	GF_LAUNCHED();
	GF_HIP(hipMemcpyAsync(units, p, (size_t)nq * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	return 0;
}

int skygen_emit(const gf_draw_plan *plan, uint64_t row_begin, uint64_t row_end, const gf_footprint *foot,
                gf_rng *rng, const gf_skygen_cols *out, hipStream_t st, hipEvent_t *ev, float *ms, uint64_t unit)
{
	GF_CHECK(plan && rng && rng->state && out && row_begin <= row_end && row_end <= plan->total, GF_E_ARG);
	if (ms) ms[0] = ms[1] = 0.f;
	if (row_begin == row_end) return 0;
	gf_footprint fp{};
	if (foot) fp = *foot;
	uint64_t u_lo = unit;
	if (u_lo == UINT64_MAX) {
		k_find_unit<<<1, 1, 0, st>>>(plan->offs, plan->U, row_begin, plan->tmp);
		GF_LAUNCHED();
		GF_HIP(hipMemcpyAsync(&u_lo, plan->tmp, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
		GF_HIP(hipStreamSynchronize(st));
	}
	GF_CHECK(u_lo < plan->U, GF_E_ARG);
	size_t shb = ((size_t)plan->sp.nM + 1) * sizeof(double);
	const size_t shb_guess = lf_pairs_offset(plan->sp.nM) + LF_GUESS * sizeof(double2);
	if (plan->ext.beam.data || plan->ext.north.data || plan->ext.south.data) {
		// with extinction the star needs its DM bin's minimum extinction: one pass
		if (ev) GF_HIP(hipEventRecord(ev[0], st));
		k_emit<<<gf::blocks_for(rng->nstreams, SK_BS), SK_BS, shb, st>>>(
		    plan->sp, plan->model, plan->pixels, plan->ext, fp, plan->P, plan->offs, plan->U, row_begin, row_end,
		    u_lo, rng->state, rng->nstreams, *out);
		GF_LAUNCHED();
		if (ev) {
			GF_HIP(hipEventRecord(ev[1], st));
			GF_HIP(hipEventRecord(ev[2], st));
			if (ms) {
				GF_HIP(hipEventSynchronize(ev[2]));
				GF_HIP(hipEventElapsedTime(&ms[0], ev[0], ev[1]));
			}
		}
		return 0;
	}
	uint64_t n = row_end - row_begin;
	GF_CHECK(n < (1ull << 32), GF_E_ARG);
	void *pk;
	int rc = gf::scratch(GF_SLOT_PARK, (size_t)n * (sizeof(float4) + sizeof(uint32_t)), &pk);
	if (rc) return rc;
	gf_parked park{(float4 *)pk, (uint32_t *)((char *)pk + (size_t)n * sizeof(float4))};
	if (ev) GF_HIP(hipEventRecord(ev[0], st));
	k_emit_rng<<<gf::blocks_for(rng->nstreams, SK_BS), SK_BS, shb_guess, st>>>(
	    plan->sp, plan->model, plan->pixels, plan->ext, plan->P, plan->offs, plan->U, row_begin, row_end, u_lo,
	    rng->state, rng->nstreams, park);
	GF_LAUNCHED();
	if (ev) GF_HIP(hipEventRecord(ev[1], st));
	if (footprint_is_simple(fp))
		k_emit_rows<false><<<gf::blocks_for(n, SK_BS), SK_BS, 0, st>>>(plan->sp, plan->model.comp, plan->pixels,
		                                                                 plan->ext, fp, (uint32_t)n, park, *out);
	else
		k_emit_rows<true><<<gf::blocks_for(n, SK_BS), SK_BS, 0, st>>>(plan->sp, plan->model.comp, plan->pixels,
		                                                                plan->ext, fp, (uint32_t)n, park, *out);
	GF_LAUNCHED();
	if (ev) {
		GF_HIP(hipEventRecord(ev[2], st));
		if (ms) {
			GF_HIP(hipEventSynchronize(ev[2]));
			GF_HIP(hipEventElapsedTime(&ms[0], ev[0], ev[1]));
			GF_HIP(hipEventElapsedTime(&ms[1], ev[1], ev[2]));
		}
	}
	return 0;
}

} // namespace gf

extern "C" {

int gf_skygen_plan_free(gf_draw_plan *plan)
{
	if (!plan) return 0;
	if (plan->counts) (void)hipFree(plan->counts);
	if (plan->offs) (void)hipFree(plan->offs);
	if (plan->P) (void)hipFree(plan->P);
	if (plan->tmp) (void)hipFree(plan->tmp);
	delete plan;
	return 0;
}

} // extern "C"

namespace gf {
int lf_prefix(const gf_skygen_params &sp, const gf_model &m, double *dst, hipStream_t st)
{
	LFTables t;
	int rc = lf_tables(sp, m, t, st);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(dst, t.P, ((size_t)sp.nM + 1) * sizeof(double), hipMemcpyDeviceToDevice, st));
	GF_HIP(hipStreamSynchronize(st));
	return 0;
}
} // namespace gf

// The scratch of gf_skygen_count_noalloc (8.1 MiB, held by every device the
// library runs on until the process ends).  At global scope, so that
// ::scratch never names gf::scratch (the allocating cache).  Host code never
// passes the array's name to a kernel: each device's address comes from
// hipGetSymbolAddress (gf::noalloc_setup).
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace gf {

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES];
}

int noalloc_setup(int device)
{
	GF_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, GF_E_ARG);
	std::lock_guard<std::mutex> lk(g_noalloc_mu);
	if (g_noalloc_ptr[device]) return 0;
	int cur = 0;
	GF_HIP(hipGetDevice(&cur));
	if (cur != device) GF_HIP(hipSetDevice(device));
	void *p = nullptr;
	hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
	if (cur != device) (void)hipSetDevice(cur);
	if (e != hipSuccess) return (int)e;
	if (!p) return GF_E_STATE;
	g_noalloc_ptr[device] = p;
	return 0;
}

int noalloc_scratch(void **ptr)
{
	int dev = 0;
	GF_HIP(hipGetDevice(&dev));
	GF_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, GF_E_ARG);
	{
		std::lock_guard<std::mutex> lk(g_noalloc_mu);
		if (g_noalloc_ptr[dev]) { *ptr = g_noalloc_ptr[dev]; return 0; }
	}
	int rc = noalloc_setup(dev);
	if (rc) return rc;
	std::lock_guard<std::mutex> lk(g_noalloc_mu);
	*ptr = g_noalloc_ptr[dev];
	return 0;
}

} // namespace gf
