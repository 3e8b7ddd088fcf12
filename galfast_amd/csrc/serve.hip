// Batched kernels of the request server: many footprint requests, each with
// its own sky binning, pixels and random-stream table, processed by one
// launch per pipeline stage.  A "segment" is one (request, model) pair.
// Every request's arithmetic and draws depend only on that request (its
// pixels, binning, seed and stream table), never on which other requests
// share the launch, so a request's rows are identical alone and in a batch.
#include "gf_runtime.h"
#include "gf_modules.h"
#include "serve.h"

#include <algorithm>

namespace {

constexpr unsigned SB = 256;

template <typename T, typename F>
GF_D int upper_index(const T *arr, int n, F key, uint64_t v)
{
	// largest i in [0, n) with key(arr[i]) <= v
	int lo = 0, hi = n;
	while (hi - lo > 1) {
		int mid = (lo + hi) >> 1;
		if (key(arr[mid]) <= v) lo = mid; else hi = mid;
	}
	return lo;
}

// ---------------------------------------------------------------------------
// Count: one workgroup per (segment pixel), lanes over the segment's DM bins.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(SB) k_b_count(const gf_serve_seg *segs, int nseg, const gf_model *models,
                                                const gf_extinction ext, const gf_pencil_beam *pixels, double *W,
                                                double *pix_sum)
{
	const uint32_t gp = blockIdx.x;   // batch pixel-slot index over this launch's segments
	int si = upper_index(segs, nseg, [](const gf_serve_seg &s) { return (uint64_t)s.slot0; }, gp);
	const gf_serve_seg &sg = segs[si];
	const uint32_t p = gp - sg.slot0;
	if (p >= sg.npix) return;
	const gf_skygen_params &sp = sg.sp;
	const gf_model &m = models[sg.model];
	const int K = sp.nm + sp.nM - 1;
	const gf_pencil_beam pix = pixels[sg.pix0 + p];
	const double *P = sg.lfP;
	double acc = 0.0;
	for (int k = threadIdx.x; k < K; k += blockDim.x) {
		float D, Am, x, y, z;
		int iM0 = k < sp.nM - 1 ? k : sp.nM - 1;
		float M = sp.M1 - iM0 * sp.dM;
		cell_pos(sp, ext, pix, M, k - iM0, D, Am, x, y, z);
		int lo, hi;
		double w = -1.0;
		if (dm_window(sp, k, D, Am, lo, hi)) {
			float R = model_rho_xyz(m, x, y, z);
			double vol = (double)pix.dA * (double)POGSON * (double)sp.dm * (double)sp.dM;
			double fac = (double)R * (double)m.norm * ((double)D * (double)D * (double)D) * vol;
			w = fac * (P[hi + 1] - P[lo]);
			acc += w;
		}
		W[sg.unit0 + (uint64_t)p * K + k] = w;
	}
	__shared__ double sh[SB];
	sh[threadIdx.x] = acc;
	__syncthreads();
	for (int o = SB / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) pix_sum[gp] = sh[0];
}

// Per-segment expected totals, summed over the segment's pixels in order.
__global__ void k_b_seg_expected(const gf_serve_seg *segs, int nseg, const double *pix_sum, double *expected)
{
	int si = blockIdx.x * blockDim.x + threadIdx.x;
	if (si >= nseg) return;
	double s = 0.0;
	for (uint32_t p = 0; p < segs[si].npix; p++) s += pix_sum[segs[si].slot0 + p];
	expected[si] = s;
}

// ---------------------------------------------------------------------------
// Plan: lane per batch unit.  Unit u of a segment belongs to stream u mod S
// of its request with index i0 + u div S (the cumulative-table draw's
// sub-streams, gf_device.h): its Poisson count and the state its stars
// start from, from the request's streams as they were at the batch start.
// ---------------------------------------------------------------------------
constexpr uint32_t JSTRIDE = 65536;   // stride of the server's jump tables

__global__ void __launch_bounds__(SB) k_b_plan_units(const gf_serve_seg *segs, int nseg, uint64_t nunits,
                                                     const double *W, const uint64_t *jtab, uint32_t *counts,
                                                     uint64_t *zu)
{
	const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nunits) return;
	const double w = W[g];
	uint32_t n = 0;
	if (w > 0.0) {
		const int si = upper_index(segs, nseg, [](const gf_serve_seg &q) { return q.unit0; }, g);
		const gf_serve_seg &sg = segs[si];
		const uint64_t u = g - sg.unit0;
		const uint32_t s = (uint32_t)(u % sg.S);
		const uint64_t i = sg.i0 + u / sg.S;
		gf_mwc r = gf_mwc_load(sg.state, sg.S, s);
		const gf_mwc_mod md(r.a);
		const uint64_t z = gf_mwc_jump_z(md, gf_mwc_z(r), i, jtab + (size_t)GF_UNIT_SPAN_BITS * JSTRIDE + s, JSTRIDE);
		gf_mwc_set_z(r, z);
		n = r.poisson((float)w);
		zu[g] = md.mul(z, jtab[(size_t)GF_EMIT_AT_BIT * JSTRIDE + s]);
	}
	counts[g] = n;
}

// ---------------------------------------------------------------------------
// Emission: lane per batch row.  The row's segment, its unit (binary search
// over the segment's offsets), the star's state 8 draws per earlier star of
// the unit on, and the star; the LF bin by binary search over the model's
// prefix.  A request's footprint is one Galactic beam; no extinction.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(SB) k_b_emit_stars(const gf_serve_seg *segs, int nseg, const gf_model *models,
                                                     const gf_extinction ext, const gf_pencil_beam *pixels,
                                                     const uint64_t *offs,
                                                     const uint64_t *zu, const uint64_t *jtab, uint64_t nrows,
                                                     gf_skygen_cols out)
{
	const uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= nrows) return;
	const int si = upper_index(segs, nseg, [](const gf_serve_seg &q) { return q.row0; }, row);
	const gf_serve_seg &sg = segs[si];
	const gf_skygen_params &sp = sg.sp;
	uint64_t a = sg.unit0, b = sg.unit0 + sg.nunits;   // offs[a] <= row < offs[b]
	while (b - a > 1) {
		const uint64_t mid = a + (b - a) / 2;
		if (offs[mid] <= row) a = mid; else b = mid;
	}
	const uint64_t u = a - sg.unit0;
	const uint64_t jj = row - offs[a];
	const uint32_t s = (uint32_t)(u % sg.S);
	const int K = sp.nm + sp.nM - 1;
	const int p = (int)(u / K), k = (int)(u % K);
	gf_mwc r;
	r.a = sg.state[s];
	const gf_mwc_mod md(r.a);
	gf_mwc_set_z(r, gf_mwc_jump_z(md, zu[a], jj, jtab + (size_t)GF_STAR_SPAN_BITS * JSTRIDE + s, JSTRIDE));
	const gf_pencil_beam pix = pixels[sg.pix0 + p];
	float Am;
	int lo, hi;
	unit_window(sp, ext, pix, k, Am, lo, hi);
	const double *P = sg.lfP;
	emit_star<false>(sp, models[sg.model].comp, ext, sg.foot, pix, k, lo, hi, Am, P,
	                 [&](int l, int h, double tg) { return lf_bin_search(P, l, h, tg); }, r, out, (uint32_t)row);
}

// Every request's streams that own a chunk of its rows move on past all its
// units: lane per (request, stream); the others draw nothing more.
__global__ void __launch_bounds__(SB) k_b_advance(const gf_serve_req *reqs, int nreq, uint32_t nlanes,
                                                  const uint64_t *jtab)
{
	const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nlanes) return;
	const int ri = upper_index(reqs, nreq, [](const gf_serve_req &q) { return (uint64_t)q.lane0; }, g);
	const gf_serve_req &rq = reqs[ri];
	const uint32_t s = g - rq.lane0;
	if (s >= rq.S || rq.advance == 0) return;
	gf_mwc r = gf_mwc_load(rq.state, rq.S, s);
	const gf_mwc_mod md(r.a);
	gf_mwc_set_z(r, gf_mwc_jump_z(md, gf_mwc_z(r), rq.advance, jtab + s, JSTRIDE));
	gf_mwc_store(rq.state, rq.S, s, r);
}

// Row starts of the segments from the scanned offsets.
__global__ void k_b_seg_rows(gf_serve_seg *segs, int nseg, const uint64_t *offs)
{
	int si = blockIdx.x * blockDim.x + threadIdx.x;
	if (si >= nseg) return;
	segs[si].row0 = offs[segs[si].unit0];
	segs[si].row1 = offs[segs[si].unit0 + segs[si].nunits];
}

// roff[r] = the first row of request r (its first segment's first unit's
// offset), roff[nreq] = offs[nunits], the batch's rows.  segall holds nm
// segments per request, request-major.
__global__ void k_b_req_rows(const gf_serve_seg *segall, int nreq, int nm, const uint64_t *offs, uint64_t nunits,
                             uint64_t *roff)
{
	int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r > nreq) return;
	roff[r] = offs[r < nreq ? segall[(size_t)r * nm].unit0 : nunits];
}

// ---------------------------------------------------------------------------
// Module chain over requests: lane (request, s) owns the request's 256-row
// chunks s, s + S, s + 2S, ... and runs the draw pass on them in that order
// (gf_chain.h); the row pass runs one lane per batch row.
// ---------------------------------------------------------------------------
// One lane g of the batch's draw pass.
GF_D void serve_draw_lane(const gf_serve_req *reqs, int nreq, uint32_t g, const gf_serve_chain &ch,
                          const gf_table &T, const gf_positions &P, gfc::DrawLdsTile &L)
{
	int ri = upper_index(reqs, nreq, [](const gf_serve_req &q) { return (uint64_t)q.lane0; }, g);
	const gf_serve_req &rq = reqs[ri];
	uint32_t s = g - rq.lane0;
	if (s >= rq.S) return;
	gf_mwc rng = gf_mwc_load(rq.state, rq.S, s);
	const uint64_t nrows = rq.row1 - rq.row0;
	for (uint64_t c = s; c * 256 < nrows; c += rq.S) {
		uint32_t r0 = (uint32_t)(rq.row0 + c * 256);
		uint32_t r1 = (uint32_t)(rq.row0 + std::min<uint64_t>(nrows, (c + 1) * 256));
		gfc::chunk_draws<false>(ch, rng, r0, r1, r0, T, P, L, threadIdx.x);
	}
	gf_mwc_store(rq.state, rq.S, s, rng);
}

// One row of the batch's row pass group GROUP.
template <int GROUP, int NT>
GF_D void serve_row(const gf_serve_req *reqs, int nreq, const gf_serve_chain &ch, uint64_t row, const gf_table &T,
                    const gf_positions &P, int need_stream)
{
	uint32_t a = 0;
	if (need_stream) {
		int ri = upper_index(reqs, nreq, [](const gf_serve_req &q) { return q.row0; }, row);
		const gf_serve_req &rq = reqs[ri];
		uint64_t c = (row - rq.row0) / 256;
		a = rq.state[c % rq.S];
	}
	gfc::chain_rows<GROUP, NT>(ch, a, (uint32_t)row, row, T, P);
}

__global__ void __launch_bounds__(gfc::BS) k_b_draws(const gf_serve_req *reqs, int nreq, uint32_t nlanes,
                                                     const gf_serve_chain *ch, gf_table T, gf_positions P)
{
	__shared__ gfc::DrawLdsTile L;
	uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nlanes) return;
	serve_draw_lane(reqs, nreq, g, *ch, T, P, L);
}

template <int GROUP, int NT>
__global__ void __launch_bounds__(gfc::BS) k_b_rows(const gf_serve_req *reqs, int nreq, const gf_serve_chain *ch,
                                                    uint64_t nrows, gf_table T, gf_positions P, int need_stream)
{
	uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= nrows) return;
	serve_row<GROUP, NT>(reqs, nreq, *ch, row, T, P, need_stream);
}

// The same passes for the captured chain: fixed grids that stride over the
// batch, which they read from the batch's argument block.
constexpr unsigned GRAPH_DRAW_BLOCKS = 1024;   // 262,144 lanes a sweep
constexpr unsigned GRAPH_ROW_BLOCKS = 8192;    // 2,097,152 rows a sweep

__global__ void __launch_bounds__(gfc::BS) k_g_draws(const gf_serve_chain_args *args, const gf_serve_chain *ch)
{
	__shared__ gfc::DrawLdsTile L;
	const gf_serve_chain_args a = *args;
	const uint32_t stride = gridDim.x * blockDim.x;
	for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.nlanes; g += stride)
		serve_draw_lane(a.reqs, a.nreq, g, *ch, a.T, a.P, L);
}

template <int GROUP, int NT>
__global__ void __launch_bounds__(gfc::BS) k_g_rows(const gf_serve_chain_args *args, const gf_serve_chain *ch,
                                                    int need_stream)
{
	const gf_serve_chain_args a = *args;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t row = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; row < a.nrows; row += stride)
		serve_row<GROUP, NT>(a.reqs, a.nreq, *ch, row, a.T, a.P, need_stream);
}

// Visible rows of each request: a workgroup per (request, 64K-row slice).
__global__ void __launch_bounds__(SB) k_b_visible(const gf_serve_req *reqs, gf_col hidden,
                                                  unsigned long long *vis)
{
	const gf_serve_req &rq = reqs[blockIdx.y];
	uint64_t b0 = rq.row0 + (uint64_t)blockIdx.x * 65536;
	uint64_t b1 = std::min<uint64_t>(rq.row1, b0 + 65536);
	unsigned long long c = 0;
	for (uint64_t r = b0 + threadIdx.x; r < b1; r += SB) c += col<int>(hidden, (uint32_t)r) == 0 ? 1ull : 0ull;
	__shared__ unsigned long long sh[SB];
	sh[threadIdx.x] = c;
	__syncthreads();
	for (int o = SB / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0 && b0 < b1) atomicAdd(&vis[blockIdx.y], sh[0]);
}

__global__ void k_b_zero(unsigned long long *v, int n)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) v[i] = 0;
}

} // namespace

namespace gf {

int serve_visible(const gf_serve_req *dreqs, int nreq, uint64_t max_rows, const gf_col &hidden,
                  unsigned long long *vis, hipStream_t st)
{
	if (nreq == 0) return 0;
	k_b_zero<<<gf::blocks_for(nreq, 64), 64, 0, st>>>(vis, nreq);
	GF_LAUNCHED();
	if (max_rows == 0) return 0;
	dim3 grid((unsigned)((max_rows + 65535) / 65536), (unsigned)nreq);
	k_b_visible<<<grid, SB, 0, st>>>(dreqs, hidden, vis);
	GF_LAUNCHED();
	return 0;
}


int serve_count(const gf_serve_seg *dsegs, int nseg, uint32_t nslots, const gf_model *dmodels, const gf_extinction &ext,
                const gf_pencil_beam *dpix, double *W, double *pix_sum, double *seg_expected, hipStream_t st)
{
	if (nslots) {
		k_b_count<<<nslots, SB, 0, st>>>(dsegs, nseg, dmodels, ext, dpix, W, pix_sum);
		GF_LAUNCHED();
	}
	k_b_seg_expected<<<gf::blocks_for(nseg, 64), 64, 0, st>>>(dsegs, nseg, pix_sum, seg_expected);
	GF_LAUNCHED();
	return 0;
}

int serve_plan(const gf_serve_seg *dsegs, int nseg, uint64_t nunits, const double *W, const uint64_t *jtab,
               uint32_t *counts, uint64_t *zu, hipStream_t st)
{
	if (!nunits) return 0;
	k_b_plan_units<<<gf::blocks_for(nunits, SB), SB, 0, st>>>(dsegs, nseg, nunits, W, jtab, counts, zu);
	GF_LAUNCHED();
	return 0;
}

int serve_req_rows(const gf_serve_seg *dsegall, int nreq, int nm, const uint64_t *offs, uint64_t nunits,
                   uint64_t *roff, hipStream_t st)
{
	k_b_req_rows<<<gf::blocks_for((uint64_t)nreq + 1, 64), 64, 0, st>>>(dsegall, nreq, nm, offs, nunits, roff);
	GF_LAUNCHED();
	return 0;
}

int serve_seg_rows(gf_serve_seg *dsegs, int nseg, const uint64_t *offs, hipStream_t st)
{
	k_b_seg_rows<<<gf::blocks_for(nseg, 64), 64, 0, st>>>(dsegs, nseg, offs);
	GF_LAUNCHED();
	return 0;
}

int serve_emit(const gf_serve_seg *dsegs, int nseg, const gf_model *dmodels, const gf_extinction &ext,
               const gf_pencil_beam *dpix,
               const uint64_t *offs, const uint64_t *zu, const uint64_t *jtab, uint64_t nrows,
               const gf_skygen_cols &out, hipStream_t st)
{
	if (!nrows) return 0;
	k_b_emit_stars<<<gf::blocks_for(nrows, SB), SB, 0, st>>>(dsegs, nseg, dmodels, ext, dpix, offs, zu, jtab, nrows,
	                                                         out);
	GF_LAUNCHED();
	return 0;
}

int serve_advance(const gf_serve_req *dreqs, int nreq, uint32_t nlanes, const uint64_t *jtab, hipStream_t st)
{
	if (!nlanes) return 0;
	k_b_advance<<<gf::blocks_for(nlanes, SB), SB, 0, st>>>(dreqs, nreq, nlanes, jtab);
	GF_LAUNCHED();
	return 0;
}

int serve_chain(const gf_serve_req *dreqs, int nreq, uint32_t nlanes, const gf_serve_chain *dch,
                const gf_serve_chain &hch, uint64_t nrows, const gf_table &T, uint2 *pos, hipStream_t st)
{
	if (nrows == 0) return 0;
	gf_positions P{pos, nrows};
	if (hch.nplanes || hch.has_mult) {
		k_b_draws<<<gf::blocks_for(nlanes, gfc::BS), gfc::BS, 0, st>>>(dreqs, nreq, nlanes, dch, T, P);
		GF_LAUNCHED();
	}
	const unsigned g = gf::blocks_for(nrows, gfc::BS);
	if (gfc::rows_on(hch, gfc::ROWS_FEH)) {
		k_b_rows<gfc::ROWS_FEH, 1><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 1);
		GF_LAUNCHED();
	}
	if (gfc::rows_on(hch, gfc::ROWS_BOND)) {
		k_b_rows<gfc::ROWS_BOND, 1><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 1);
		GF_LAUNCHED();
	}
	if (gfc::rows_on(hch, gfc::ROWS_PHOT)) {
		switch (gfc::phot_tables(hch)) {
		case 1: k_b_rows<gfc::ROWS_PHOT, 1><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 0); break;
		case 2: k_b_rows<gfc::ROWS_PHOT, 2><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 0); break;
		case 3: k_b_rows<gfc::ROWS_PHOT, 3><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 0); break;
		default: k_b_rows<gfc::ROWS_PHOT, 4><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 0); break;
		}
		GF_LAUNCHED();
	}
	if (gfc::rows_on(hch, gfc::ROWS_PERR)) {
		k_b_rows<gfc::ROWS_PERR, 1><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 1);
		GF_LAUNCHED();
	}
	if (gfc::rows_on(hch, gfc::ROWS_ASTRO)) {
		k_b_rows<gfc::ROWS_ASTRO, 1><<<g, gfc::BS, 0, st>>>(dreqs, nreq, dch, nrows, T, P, 0);
		GF_LAUNCHED();
	}
	return 0;
}

int serve_chain_graph(const gf_serve_chain_args *dargs, const gf_serve_chain *dch, const gf_serve_chain &hch,
                      hipStream_t st, hipGraphExec_t *exec)
{
	GF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
	// launches only: nothing that a capture would record as a memset or copy
	const unsigned gr = GRAPH_ROW_BLOCKS, gd = GRAPH_DRAW_BLOCKS;
	if (hch.nplanes || hch.has_mult) k_g_draws<<<gd, gfc::BS, 0, st>>>(dargs, dch);
	if (gfc::rows_on(hch, gfc::ROWS_FEH)) k_g_rows<gfc::ROWS_FEH, 1><<<gr, gfc::BS, 0, st>>>(dargs, dch, 1);
	if (gfc::rows_on(hch, gfc::ROWS_BOND)) k_g_rows<gfc::ROWS_BOND, 1><<<gr, gfc::BS, 0, st>>>(dargs, dch, 1);
	if (gfc::rows_on(hch, gfc::ROWS_PHOT)) {
		switch (gfc::phot_tables(hch)) {
		case 1: k_g_rows<gfc::ROWS_PHOT, 1><<<gr, gfc::BS, 0, st>>>(dargs, dch, 0); break;
		case 2: k_g_rows<gfc::ROWS_PHOT, 2><<<gr, gfc::BS, 0, st>>>(dargs, dch, 0); break;
		case 3: k_g_rows<gfc::ROWS_PHOT, 3><<<gr, gfc::BS, 0, st>>>(dargs, dch, 0); break;
		default: k_g_rows<gfc::ROWS_PHOT, 4><<<gr, gfc::BS, 0, st>>>(dargs, dch, 0); break;
		}
	}
	if (gfc::rows_on(hch, gfc::ROWS_PERR)) k_g_rows<gfc::ROWS_PERR, 1><<<gr, gfc::BS, 0, st>>>(dargs, dch, 1);
	if (gfc::rows_on(hch, gfc::ROWS_ASTRO)) k_g_rows<gfc::ROWS_ASTRO, 1><<<gr, gfc::BS, 0, st>>>(dargs, dch, 0);
	hipError_t le = hipGetLastError();
	hipGraph_t g = nullptr;
	hipError_t ce = hipStreamEndCapture(st, &g);
	if (le != hipSuccess) { if (g) (void)hipGraphDestroy(g); return (int)le; }
	if (ce != hipSuccess) return (int)ce;
	hipError_t ie = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
	(void)hipGraphDestroy(g);
	return (int)ie;
}

} // namespace gf
