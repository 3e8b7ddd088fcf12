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
#include <algorithm>
#include <memory>
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

// The count's prefix table in LDS (LDS = true, launched with nM + 1 entries
// of dynamic shared memory) when it has at most COUNT_LDS_NM + 1 entries,
// else read from global memory.
constexpr int COUNT_LDS_NM = 4095;

template <int MODE, bool LDS>
__global__ void __launch_bounds__(SK_BS) k_count(gf_skygen_params sp, gf_model m, const gf_pencil_beam *pixels,
                                                 gf_extinction ext, LFTables t, CountAcc *pix_out,
                                                 double *per_pixel, double *wpk)
{
	const int p = blockIdx.x;
	const int K = sp.nm + sp.nM - 1;
	if (LDS) {
		// the mode's prefix table, staged once per workgroup
		extern __shared__ double count_prefix[];
		if (MODE == GF_ACC_FP64) {
			for (int i = threadIdx.x; i <= sp.nM; i += blockDim.x) count_prefix[i] = t.P[i];
			t.P = count_prefix;
		} else {
			float *f = (float *)count_prefix;
			const float *src = MODE == GF_ACC_KAHAN32 ? t.Pk : t.Pf;
			for (int i = threadIdx.x; i <= sp.nM; i += blockDim.x) f[i] = src[i];
			if (MODE == GF_ACC_KAHAN32) t.Pk = f; else t.Pf = f;
		}
		__syncthreads();
	}
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
// k_count in accumulation MODE, its prefix table (entries of type T) staged
// in LDS when it fits.
template <int MODE, typename T>
void count_kernel(const gf_skygen_params &sp, const gf_model &m, const gf_pencil_beam *pixels,
                  const gf_extinction &ex, const LFTables &t, CountAcc *pix_out, double *per_pixel, double *wpk,
                  hipStream_t st)
{
	if (sp.nM <= COUNT_LDS_NM)
		k_count<MODE, true><<<sp.npixels, SK_BS, ((size_t)sp.nM + 1) * sizeof(T), st>>>(sp, m, pixels, ex, t, pix_out,
		                                                                                per_pixel, wpk);
	else
		k_count<MODE, false><<<sp.npixels, SK_BS, 0, st>>>(sp, m, pixels, ex, t, pix_out, per_pixel, wpk);
}

int launch_count(const gf_skygen_params &sp, const gf_model &m, const gf_pencil_beam *pixels,
                 const gf_extinction &ex, int acc_mode, const LFTables &t, CountAcc *pix_out,
                 gf_count_result *dres, double *per_pixel, double *wpk, hipStream_t st)
{
	switch (acc_mode) {
	case GF_ACC_FP64:
		count_kernel<GF_ACC_FP64, double>(sp, m, pixels, ex, t, pix_out, per_pixel, wpk, st);
		GF_LAUNCHED();
		k_count_final<GF_ACC_FP64><<<1, SK_BS, 0, st>>>(pix_out, sp.npixels, dres);
		break;
	case GF_ACC_KAHAN32:
		count_kernel<GF_ACC_KAHAN32, float>(sp, m, pixels, ex, t, pix_out, per_pixel, wpk, st);
		GF_LAUNCHED();
		k_count_final<GF_ACC_KAHAN32><<<1, SK_BS, 0, st>>>(pix_out, sp.npixels, dres);
		break;
	default:
		count_kernel<GF_ACC_FP32, float>(sp, m, pixels, ex, t, pix_out, per_pixel, wpk, st);
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
                                                    gf_skygen_cols out, int store = 0)
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
	if (pass == 1 || store) gf_mwc_store(st, ns, tid, rng);
}

// The crawl state of the thread a batch of the capacity-limited draw stopped
// inside (runtime_state, skygen/skygen.h:184-257): its place in the crawl,
// the current cell's values, the stars of that cell still to draw and its
// stream.
struct gf_draw_resume {
	int32_t tid, ilb, im, iM, k, bc;
	float D, Am, R;
	uint32_t ndraw;
	uint32_t a, c, x;
};

// Rows [r0, r1) of the reference-order draw: thread t writes its stars
// offs[t] .. offs[t+1] - 1 (the count pass's exclusive scan), from its
// stream's state at the start of the draw (st0), or, for the thread whose
// stars began in the previous batch, from the state that batch left in
// res_in; the thread that reaches row r1 with stars still to draw leaves
// its state in res_out.  The batches concatenated are the unlimited draw.
__global__ void __launch_bounds__(64) k_draw_window(gf_skygen_params sp, gf_model m, const gf_pencil_beam *pixels,
                                                    gf_extinction ext, gf_footprint foot, const uint32_t *st0,
                                                    uint32_t ns, const uint64_t *offs, uint64_t r0, uint64_t r1,
                                                    const gf_draw_resume *res_in, gf_draw_resume *res_out,
                                                    gf_skygen_cols out)
{
	const int tid = blockIdx.x * blockDim.x + threadIdx.x;
	if (tid >= sp.nthreads) return;
	const uint64_t o0 = offs[tid], o1 = offs[tid + 1];
	if (o1 <= r0 || o0 >= r1) return;
	const int block = 10;
	const int nthreads = sp.nthreads, nm = sp.nm, nM = sp.nM;
	int ilb = 0, im = 0, iM = 0, k = block * (tid - nthreads), bc = 0;
	float D = 0.f, Am = 0.f, R = 0.f;
	gf_pencil_beam pix;
	gf_mwc rng;
	unsigned int ndraw = 0;
	uint64_t row = o0;
	if (o0 < r0) {
		const gf_draw_resume q = *res_in;
		ilb = q.ilb; im = q.im; iM = q.iM; k = q.k; bc = q.bc;
		D = q.D; Am = q.Am; R = q.R;
		ndraw = q.ndraw;
		rng.a = q.a; rng.c = q.c; rng.x = q.x;
		pix = pixels[ilb];
		row = r0;
	} else {
		rng = gf_mwc_load(st0, ns, tid);
	}
	StarSink s{out, true, row - r0};
	for (;;) {
		while (ndraw > 0) {
			if (row == r1) {
				*res_out = gf_draw_resume{tid, ilb, im, iM, k, bc, D, Am, R, ndraw, rng.a, rng.c, rng.x};
				return;
			}
			draw_one_star(sp, m, ext, foot, pix, sp.M1 - iM * sp.dM, im, Am, rng, s);
			s.row++;
			row++;
			ndraw--;
		}
		if (row >= o1) return;
		bool moved;
		if (bc == 0) {
			bc = block;
			k += block * nthreads;
			diag_index_to_ij(ilb, im, iM, k, nm, nM);
			if (ilb >= sp.npixels) return;
			pix = pixels[ilb];
			moved = true;
		} else {
			moved = crawl_advance(ilb, im, iM, pix, pixels, sp.npixels, nm, nM);
		}
		bc--;
		float M = sp.M1 - iM * sp.dM;
		if (moved) {
			if (ilb >= sp.npixels) return;
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
		ndraw = rng.poisson(rho);
	}
}

// ---------------------------------------------------------------------------
// Cumulative-table draw on sub-streams (GF_UNIT_SPAN_BITS in gf_device.h).
// ---------------------------------------------------------------------------
// Plan: lane per unit.  The unit's Poisson count from its own position of
// its stream, and the state where its stars' draws begin.
__global__ void __launch_bounds__(SK_BS) k_plan_units(const double *wpk, uint64_t U, const uint32_t *base,
                                                      uint32_t ns, const uint64_t *jtab, uint32_t *counts,
                                                      uint64_t *zu)
{
	const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (u >= U) return;
	const double w = wpk[u];
	uint32_t n = 0;
	if (w > 0.0) {
		const uint32_t s = (uint32_t)(u % ns);
		const uint64_t i = u / ns;
		gf_mwc r = gf_mwc_load(base, ns, s);
		const gf_mwc_mod md(r.a);
		// i << GF_UNIT_SPAN_BITS draws: the table from that bit on (the same products)
		const uint64_t z = gf_mwc_jump_z(md, gf_mwc_z(r), i, jtab + (size_t)GF_UNIT_SPAN_BITS * ns + s, ns);
		gf_mwc_set_z(r, z);
		n = r.poisson((float)w);
		zu[u] = md.mul(z, jtab[(size_t)GF_EMIT_AT_BIT * ns + s]);
	}
	counts[u] = n;
}

constexpr int EMIT_MAX_NM = 4096;

// LDS of the emission: the LF prefix (nM + 1 doubles), the guess table
// (LF_GUESS uint16) and, 16-byte aligned, the guess thresholds.
GF_HD size_t lf_pairs_offset(int nM)
{
	const size_t b = ((size_t)nM + 1) * sizeof(double) + LF_GUESS * sizeof(uint16_t);
	return (b + 15) & ~(size_t)15;
}

constexpr unsigned EMIT_TILE = SK_BS;   // rows per tile of the emission

// tu[q] = the unit of the first row of tile q (q < ntiles) and tu[ntiles]
// = the unit of row row_end - 1: the largest u in [u_lo, U) with
// offs[u] <= row (offs[u_lo] <= row_begin).
__global__ void k_tile_units(const uint64_t *offs, uint64_t U, uint64_t row_begin, uint64_t row_end, uint64_t u_lo,
                             uint32_t ntiles, uint64_t *tu)
{
	const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q > ntiles) return;
	const uint64_t row = q < ntiles ? row_begin + (uint64_t)q * EMIT_TILE : row_end - 1;
	uint64_t a = u_lo, b = U;
	while (b - a > 1) {
		const uint64_t mid = a + (b - a) / 2;
		if (offs[mid] <= row) a = mid; else b = mid;
	}
	tu[q] = a;
}

// The emission's LF tables of a plan, built once: the prefix P, the guess
// table G and its thresholds GP, laid out as in the emission's LDS
// (lf_pairs_offset).
__global__ void __launch_bounds__(SK_BS) k_lf_guess(const double *Pg, int nM, char *dst)
{
	double *P = (double *)dst;
	uint16_t *G = (uint16_t *)(P + nM + 1);
	double2 *GP = (double2 *)(dst + lf_pairs_offset(nM));
	lf_guess_build(Pg, nM, G, threadIdx.x, blockDim.x);
	for (int i = threadIdx.x; i <= nM; i += blockDim.x) P[i] = Pg[i];
	__syncthreads();
	lf_guess_pairs(Pg, G, GP, threadIdx.x, blockDim.x);
}

constexpr unsigned EMIT_OFFS_LDS = 512;   // tile unit offsets staged in LDS (wider ranges read global memory)

// Emission: lane per star, one tile of EMIT_TILE rows per workgroup.  A row
// finds its unit between its tile's first and last units (their offsets
// staged in LDS when the range is at most EMIT_OFFS_LDS units), jumps from
// the unit's star state 8 draws per star before it, and draws the star.
template <bool FULL>
__global__ void __launch_bounds__(SK_BS) k_emit_stars(gf_skygen_params sp, int comp, const gf_pencil_beam *pixels,
                                                      gf_extinction ext, gf_footprint foot, const char *lfg,
                                                      uint32_t lfg_words, const uint64_t *offs, uint64_t row_begin,
                                                      uint64_t row_end, const uint64_t *tu, const uint32_t *aplane,
                                                      uint32_t ns, const uint64_t *zu, const uint64_t *jtab,
                                                      gf_skygen_cols out)
{
	extern __shared__ double P[];
	__shared__ uint64_t so[EMIT_OFFS_LDS];
	const uint32_t q = blockIdx.x;
	const uint64_t ua = tu[q], ub = tu[q + 1] + 1;   // offs[ua] <= the tile's rows < offs[ub], ub <= U
	const bool staged = ub - ua < EMIT_OFFS_LDS;     // so[k] = offs[ua + k], k <= ub - ua
	for (uint32_t i = threadIdx.x; i < lfg_words; i += blockDim.x) ((uint4 *)P)[i] = ((const uint4 *)lfg)[i];
	if (staged)
		for (uint32_t i = threadIdx.x; i <= (uint32_t)(ub - ua); i += blockDim.x) so[i] = offs[ua + i];
	__syncthreads();
	const uint16_t *G = (const uint16_t *)(P + sp.nM + 1);
	const double2 *GP = (const double2 *)((const char *)P + lf_pairs_offset(sp.nM));
	const double gscale = lf_guess_scale(P, sp.nM);
	const uint64_t j = row_begin + (uint64_t)q * EMIT_TILE + threadIdx.x;
	if (j >= row_end) return;
	uint64_t u, o;   // the row's unit and its first row
	if (staged) {
		uint32_t a = 0, b = (uint32_t)(ub - ua);
		while (b - a > 1) {
			const uint32_t mid = a + (b - a) / 2;
			if (so[mid] <= j) a = mid; else b = mid;
		}
		u = ua + a;
		o = so[a];
	} else {
		uint64_t a = ua, b = ub;
		while (b - a > 1) {
			const uint64_t mid = a + (b - a) / 2;
			if (offs[mid] <= j) a = mid; else b = mid;
		}
		u = a;
		o = offs[u];
	}
	const uint64_t jj = j - o;
	const uint32_t s = (uint32_t)(u % ns);
	const int K = sp.nm + sp.nM - 1;
	const int p = (int)(u / K), k = (int)(u % K);
	gf_mwc r;
	r.a = aplane[s];
	const gf_mwc_mod md(r.a);
	gf_mwc_set_z(r, gf_mwc_jump_z(md, zu[u], jj, jtab + (size_t)GF_STAR_SPAN_BITS * ns + s, ns));
	const gf_pencil_beam pix = pixels[p];
	float Am;
	int lo, hi;
	unit_window(sp, ext, pix, k, Am, lo, hi);
	emit_star<FULL>(sp, comp, ext, foot, pix, k, lo, hi, Am, P,
	                [&](int l, int h, double tg) { return lf_bin_pairs(P, G, GP, gscale, l, h, tg); }, r, out,
	                (uint32_t)(j - row_begin));
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
	uint32_t ns;          // streams
	uint32_t *counts;
	uint64_t *offs;
	double *P;
	uint64_t *tmp;
	uint32_t *base;       // the streams [a | c | x] at the plan
	uint64_t *jtab;       // their jump tables (GF_JUMP_BITS x ns)
	uint64_t *zu;         // per unit: the state its stars' draws begin from
	char *lfg;            // the emission's LF tables (k_lf_guess)
	uint32_t lfg_words;   // their size in 16-byte words
	// allocated element counts (a plan is reused across runs)
	size_t cap_counts, cap_offs, cap_P, cap_tmp, cap_base, cap_jtab, cap_zu, cap_lfg;
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

// The capacity-limited reference draw (skygenHost::drawSources,
// skygen/skyconfig_impl.h:434-534): the streams' starting states, each
// thread's star count and their exclusive scan, the two resume records and
// the next row to write.
struct gf_draw_state {
	gf_skygen_params sp;
	gf_model m;
	const gf_pencil_beam *pixels;
	gf_extinction ext;
	gf_footprint foot;
	uint32_t ns;
	uint32_t *st0 = nullptr;          // 3 * ns words
	uint32_t *counts = nullptr;       // nthreads
	uint64_t *offs = nullptr;         // nthreads + 1
	gf_draw_resume *res = nullptr;    // 2
	uint64_t total = 0, next = 0;
	int batch = 0;
	~gf_draw_state()
	{
		if (st0) (void)hipFree(st0);
		if (counts) (void)hipFree(counts);
		if (offs) (void)hipFree(offs);
		if (res) (void)hipFree(res);
	}
};

int gf_skygen_draw_begin(const gf_skygen_params *sp, const gf_model *model, const gf_pencil_beam *pixels,
                         const gf_extinction *ext, const gf_footprint *foot, gf_rng *rng, gf_draw_state **state,
                         uint64_t *nstars, void *stream)
{
	GF_CHECK(sp && model && pixels && rng && rng->state && state && nstars, GF_E_ARG);
	GF_CHECK(sp->nthreads > 0 && (uint32_t)sp->nthreads <= rng->nstreams, GF_E_STREAMS);
	*state = nullptr;
	hipStream_t st = gf::as_stream(stream);
	auto d = std::make_unique<gf_draw_state>();
	d->sp = *sp;
	d->m = *model;
	d->pixels = pixels;
	d->ext = ext ? *ext : gf_extinction{};
	d->foot = foot ? *foot : gf_footprint{};
	d->ns = rng->nstreams;
	const int nt = sp->nthreads;
	GF_HIP(hipMalloc(&d->st0, (size_t)3 * d->ns * sizeof(uint32_t)));
	GF_HIP(hipMalloc(&d->counts, (size_t)nt * sizeof(uint32_t)));
	GF_HIP(hipMalloc(&d->offs, ((size_t)nt + 1) * sizeof(uint64_t)));
	GF_HIP(hipMalloc(&d->res, 2 * sizeof(gf_draw_resume)));
	GF_HIP(hipMemcpyAsync(d->st0, rng->state, (size_t)3 * d->ns * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
	// the count pass leaves the streams where the whole draw leaves them
	const unsigned grid = gf::blocks_for(nt, 64);
	k_draw_compat<<<grid, 64, 0, st>>>(d->sp, d->m, pixels, d->ext, d->foot, rng->state, rng->nstreams, 0,
	                                   d->counts, d->offs, gf_skygen_cols{}, 1);
	GF_LAUNCHED();
	int rc = gf::exclusive_scan_u32_u64(d->counts, d->offs, nt, st);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(&d->total, d->offs + nt, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	*nstars = d->total;
	*state = d.release();
	return 0;
}

int gf_skygen_draw_next(gf_draw_state *d, const gf_skygen_cols *out, uint64_t capacity, uint64_t *nwritten,
                        void *stream)
{
	GF_CHECK(d && out && nwritten && capacity > 0, GF_E_ARG);
	hipStream_t st = gf::as_stream(stream);
	const uint64_t r0 = d->next, r1 = std::min(d->total, r0 + capacity);
	*nwritten = r1 - r0;
	if (r1 == r0) return 0;
	const gf_draw_resume *in = d->res + (d->batch & 1);
	gf_draw_resume *outr = d->res + ((d->batch + 1) & 1);
	k_draw_window<<<gf::blocks_for(d->sp.nthreads, 64), 64, 0, st>>>(d->sp, d->m, d->pixels, d->ext, d->foot, d->st0,
	                                                                  d->ns, d->offs, r0, r1, in, outr, *out);
	GF_LAUNCHED();
	d->next = r1;
	d->batch++;
	return 0;
}

void gf_skygen_draw_end(gf_draw_state *d)
{
	delete d;
}

int gf_skygen_plan(const gf_skygen_params *sp, const gf_model *model,
                   const gf_pencil_beam *pixels, const gf_extinction *ext,
                   const double *wpk, gf_rng *rng, gf_draw_plan **plan,
                   uint64_t *nstars, void *stream)
{
	GF_CHECK(plan, GF_E_ARG);
	*plan = nullptr;
	gf_draw_plan *pl = nullptr;
	int rc = gf::skygen_plan(sp, model, pixels, ext, wpk, rng, &pl, nstars, gf::as_stream(stream));
	if (rc) { gf_skygen_plan_free(pl); return rc; }
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

namespace {
template <typename T> hipError_t grow_dev(T *&p, size_t &cap, size_t n)
{
	if (p && cap >= n) return hipSuccess;
	if (p) (void)hipFree(p);
	p = nullptr;
	cap = 0;
	hipError_t e = hipMalloc((void **)&p, n * sizeof(T));
	if (e == hipSuccess) cap = n;
	return e;
}
} // namespace

int skygen_plan(const gf_skygen_params *sp, const gf_model *model, const gf_pencil_beam *pixels,
                const gf_extinction *ext, const double *wpk, gf_rng *rng, gf_draw_plan **plan, uint64_t *nstars,
                hipStream_t st)
{
	GF_CHECK(sp && model && pixels && wpk && rng && rng->state && plan && nstars, GF_E_ARG);
	GF_CHECK(sp->nM <= EMIT_MAX_NM && rng->nstreams > 0, GF_E_ARG);
	gf_draw_plan *pl = *plan;
	if (!pl) *plan = pl = new gf_draw_plan{};
	pl->sp = *sp;
	pl->model = *model;
	pl->pixels = pixels;
	pl->ext = ext ? *ext : gf_extinction{};
	pl->total = 0;
	const uint64_t K = (uint64_t)(sp->nm + sp->nM - 1);
	const uint32_t ns = rng->nstreams;
	pl->U = (uint64_t)sp->npixels * K;
	pl->ns = ns;
	const size_t lfg_bytes = lf_pairs_offset(sp->nM) + LF_GUESS * sizeof(double2);
	pl->lfg_words = (uint32_t)((lfg_bytes + 15) / 16);
	hipError_t e = grow_dev(pl->counts, pl->cap_counts, pl->U + 1);
	if (e == hipSuccess) e = grow_dev(pl->offs, pl->cap_offs, pl->U + 1);
	if (e == hipSuccess) e = grow_dev(pl->P, pl->cap_P, (size_t)sp->nM + 1);
	if (e == hipSuccess) e = grow_dev(pl->tmp, pl->cap_tmp, 2);
	if (e == hipSuccess) e = grow_dev(pl->base, pl->cap_base, (size_t)3 * ns);
	if (e == hipSuccess) e = grow_dev(pl->jtab, pl->cap_jtab, (size_t)GF_JUMP_BITS * ns);
	if (e == hipSuccess) e = grow_dev(pl->zu, pl->cap_zu, pl->U + 1);
	if (e == hipSuccess) e = grow_dev(pl->lfg, pl->cap_lfg, (size_t)pl->lfg_words * 16);
	if (e != hipSuccess) return (int)e;
	LFTables t;
	int rc = lf_tables(*sp, *model, t, st);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(pl->P, t.P, ((size_t)sp->nM + 1) * sizeof(double), hipMemcpyDeviceToDevice, st));
	GF_HIP(hipMemcpyAsync(pl->base, rng->state, (size_t)3 * ns * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
	rc = gf::mwc_jump_table(rng, pl->jtab, st);
	if (rc) return rc;
	k_lf_guess<<<1, SK_BS, 0, st>>>(pl->P, sp->nM, pl->lfg);
	GF_LAUNCHED();
	if (pl->U > 0) {
		k_plan_units<<<gf::blocks_for(pl->U, SK_BS), SK_BS, 0, st>>>(wpk, pl->U, pl->base, ns, pl->jtab, pl->counts,
		                                                             pl->zu);
		GF_LAUNCHED();
		// the streams move past every unit's draws
		rc = gf::mwc_jump(rng, ((pl->U + ns - 1) / ns) << GF_UNIT_SPAN_BITS, pl->jtab, st);
		if (rc) return rc;
	}
	rc = gf::exclusive_scan_u32_u64(pl->counts, pl->offs, pl->U, st);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(&pl->total, pl->offs + pl->U, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	*nstars = pl->total;
	return 0;
}

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
	(void)rng;   // the stars' draws come from the streams as they were at the plan
	GF_CHECK(plan && out && row_begin <= row_end && row_end <= plan->total, GF_E_ARG);
	if (ms) ms[0] = ms[1] = 0.f;
	if (row_begin == row_end) return 0;
	gf_footprint fp{};
	if (foot) fp = *foot;
	const uint64_t n = row_end - row_begin;
	GF_CHECK(n < (1ull << 32), GF_E_ARG);
	const uint32_t ntiles = (uint32_t)((n + EMIT_TILE - 1) / EMIT_TILE);
	void *p;
	int rc = gf::scratch(GF_SLOT_PARK, ((size_t)ntiles + 1) * sizeof(uint64_t), &p);
	if (rc) return rc;
	uint64_t *tu = (uint64_t *)p;
	const uint64_t u_lo = unit == UINT64_MAX ? 0 : unit;
	GF_CHECK(u_lo < plan->U, GF_E_ARG);
	if (ev) GF_HIP(hipEventRecord(ev[0], st));
	k_tile_units<<<gf::blocks_for((uint64_t)ntiles + 1, 256), 256, 0, st>>>(plan->offs, plan->U, row_begin, row_end,
	                                                                        u_lo, ntiles, tu);
	GF_LAUNCHED();
	if (ev) GF_HIP(hipEventRecord(ev[1], st));
	const size_t shb = (size_t)plan->lfg_words * 16;
	if (footprint_is_simple(fp))
		k_emit_stars<false><<<ntiles, SK_BS, shb, st>>>(plan->sp, plan->model.comp, plan->pixels, plan->ext, fp,
		                                                plan->lfg, plan->lfg_words, plan->offs, row_begin, row_end, tu,
		                                                plan->base, plan->ns, plan->zu, plan->jtab, *out);
	else
		k_emit_stars<true><<<ntiles, SK_BS, shb, st>>>(plan->sp, plan->model.comp, plan->pixels, plan->ext, fp,
		                                               plan->lfg, plan->lfg_words, plan->offs, row_begin, row_end, tu,
		                                               plan->base, plan->ns, plan->zu, plan->jtab, *out);
	GF_LAUNCHED();
	if (ev) {
		GF_HIP(hipEventRecord(ev[2], st));
		if (ms) {
			GF_HIP(hipEventSynchronize(ev[2]));
			GF_HIP(hipEventElapsedTime(&ms[0], ev[1], ev[2]));
			GF_HIP(hipEventElapsedTime(&ms[1], ev[0], ev[1]));
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
	if (plan->base) (void)hipFree(plan->base);
	if (plan->jtab) (void)hipFree(plan->jtab);
	if (plan->zu) (void)hipFree(plan->zu);
	if (plan->lfg) (void)hipFree(plan->lfg);
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
