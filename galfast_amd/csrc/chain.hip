// Per-star module chain in pipeline order (pipeline.h:75-89: FeH,
// multiples, kinematics, photometry, detector errors, database
// conversions) in two launches: the draw pass, one lane per 256-row chunk
// with the chunk's stream, and the row pass, one lane per row (gf_chain.h).
// Each module consumes its draws in the same order as its stand-alone
// kernel, so the output equals the sequence of gf_os_*_kernel calls.
#include "gf_runtime.h"
#include "gf_chain.h"
#include <algorithm>

namespace {

// Draw pass: lane t owns chunk t of the rows and stream t.  The launch
// descriptor travels by value in the kernel arguments.
template <bool CACHED>
__global__ void __launch_bounds__(gfc::BS) k_chain_draws(const gf_chain_launch c, uint32_t nthr, uint32_t *st,
                                                         uint32_t ns)
{
	__shared__ gfc::DrawLds L;
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nthr) return;
	__builtin_amdgcn_s_setprio(3);
	uint64_t a = (uint64_t)c.begin + (uint64_t)c.step * t;
	uint32_t r0 = (uint32_t)a, r1 = (uint32_t)std::min<uint64_t>(a + c.step, c.end);
	gf_mwc r = gf_mwc_load(st, ns, t);
	gfc::chunk_draws<CACHED>(c.m, r, r0, r1, (uint64_t)(r0 - c.begin), c.t, c.P, L, threadIdx.x, c.vis);
	gf_mwc_store(st, ns, t, r);
}

// Row pass: one lane per row; the row's stream is its chunk's.  It reads
// only the multiplier plane of the streams, which no launch changes.
template <int GROUP, int NT>
__global__ void __launch_bounds__(gfc::BS) k_chain_rows(const gf_chain_launch c, const uint32_t *st, int need_stream)
{
	uint32_t row = c.begin + blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= c.end) return;
	uint32_t i = row - c.begin;
	gfc::chain_rows<GROUP, NT>(c.m, need_stream ? st[i / c.step] : 0u, row, i, c.t, c.P);
}

} // namespace

namespace gf {

size_t chain_positions_bytes(const gf_chain *chain, uint32_t nrows)
{
	int n = (chain->feh ? 1 : 0) + (chain->gfeh ? 1 : 0) + (chain->bond ? 1 : 0) + (chain->perr ? 1 : 0);
	return (size_t)n * nrows * sizeof(uint2);
}

int chain_setup(const gf_rows *ks, const gf_chain *chain, const gf_rng *rng, const gf_table *t, uint2 *pos,
                gf_chain_launch &L)
{
	GF_CHECK(ks && chain && t && rng && rng->state && ks->step > 0 && ks->end >= ks->begin, GF_E_ARG);
	GF_CHECK(!(chain->feh && chain->gfeh), GF_E_ARG);
	uint32_t n = ks->end - ks->begin;
	uint32_t nthr = n / ks->step + (n % ks->step ? 1 : 0);
	GF_CHECK(nthr <= rng->nstreams, GF_E_STREAMS);
	if (chain->phot) GF_CHECK(chain->phot->ncolors >= 1 && chain->phot->ncolors <= 16, GF_E_ARG);
	L = gf_chain_launch{};
	gf_chain_mods &m = L.m;
	if (chain->feh) { m.feh = *chain->feh; m.has_feh = 1; }
	if (chain->gfeh) { m.gfeh = *chain->gfeh; m.has_gfeh = 1; m.gfeh_comps = chain->gfeh_comps; }
	if (chain->mult) { m.mult = *chain->mult; m.has_mult = 1; m.mult_comps = chain->mult_comps; }
	if (chain->bond) { m.bond = *chain->bond; m.has_bond = 1; }
	if (chain->phot) { m.phot = *chain->phot; m.has_phot = 1; m.phot_comps = chain->phot_comps; }
	if (chain->perr) { m.perr = *chain->perr; m.has_perr = 1; }
	if (chain->v2pm) { m.v2pm = *chain->v2pm; m.has_v2pm = 1; }
	m.gal2other = chain->gal2other_coordsys;
	gf_chain_planes(m);
	L.t = *t;
	L.begin = ks->begin; L.step = ks->step; L.end = ks->end;
	if (m.nplanes) {
		if (!pos) {
			void *p;
			int rc = gf::scratch(GF_SLOT_NSLOTS + 4, (size_t)m.nplanes * n * sizeof(uint2), &p);
			if (rc) return rc;
			pos = (uint2 *)p;
		}
		L.P.pos = pos;
		L.P.pitch = n;
	}
	return 0;
}

bool chain_counts_visible(const gf_chain_launch &L)
{
	return chain_on(L, 0) && L.step <= (uint32_t)gfc::MAXC;
}

bool chain_on(const gf_chain_launch &L, int launch)
{
	if (L.end <= L.begin) return false;
	if (launch == 0) return L.m.nplanes || L.m.has_mult;
	return gfc::rows_on(L.m, launch - 1);
}

int chain_launch(const gf_chain_launch &L, const gf_rng *rng, int launch, hipStream_t st)
{
	if (!chain_on(L, launch)) return 0;
	const uint32_t n = L.end - L.begin;
	const unsigned grows = gf::blocks_for(n, gfc::BS);
	switch (launch) {
	case 0: {
		uint32_t nthr = n / L.step + (n % L.step ? 1 : 0);
		unsigned g = gf::blocks_for(nthr, gfc::BS);
		if (L.step <= (uint32_t)gfc::MAXC)
			k_chain_draws<true><<<g, gfc::BS, 0, st>>>(L, nthr, rng->state, rng->nstreams);
		else
			k_chain_draws<false><<<g, gfc::BS, 0, st>>>(L, nthr, rng->state, rng->nstreams);
		break;
	}
	case 1 + gfc::ROWS_FEH:
		k_chain_rows<gfc::ROWS_FEH, 1><<<grows, gfc::BS, 0, st>>>(L, rng->state, 1);
		break;
	case 1 + gfc::ROWS_BOND:
		k_chain_rows<gfc::ROWS_BOND, 1><<<grows, gfc::BS, 0, st>>>(L, rng->state, 1);
		break;
	case 1 + gfc::ROWS_PHOT:
		switch (gfc::phot_variant(L.m, L.t)) {
		case 1: k_chain_rows<gfc::ROWS_PHOT, 1><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		case 2: k_chain_rows<gfc::ROWS_PHOT, 2><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		case 3: k_chain_rows<gfc::ROWS_PHOT, 3><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		case 4: k_chain_rows<gfc::ROWS_PHOT, 4><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		case 1 | gfc::PHOT_LONE: k_chain_rows<gfc::ROWS_PHOT, 1 | gfc::PHOT_LONE><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		case 2 | gfc::PHOT_LONE: k_chain_rows<gfc::ROWS_PHOT, 2 | gfc::PHOT_LONE><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		case 3 | gfc::PHOT_LONE: k_chain_rows<gfc::ROWS_PHOT, 3 | gfc::PHOT_LONE><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		default: k_chain_rows<gfc::ROWS_PHOT, 4 | gfc::PHOT_LONE><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0); break;
		}
		break;
	case 1 + gfc::ROWS_PERR:
		k_chain_rows<gfc::ROWS_PERR, 1><<<grows, gfc::BS, 0, st>>>(L, rng->state, 1);
		break;
	default:
		k_chain_rows<gfc::ROWS_ASTRO, 1><<<grows, gfc::BS, 0, st>>>(L, rng->state, 0);
		break;
	}
	GF_LAUNCHED();
	return 0;
}

// The chain on one stream, with optional per-launch timing: when ev is
// non-NULL it holds GF_CHAIN_LAUNCHES + 1 events; ms[0] receives the draw
// pass's duration, ms[1..5] the row pass's groups (metallicity, kinematics,
// photometry, photometric errors, astrometry); 0 for a launch with nothing
// to do.
int module_chain(const gf_rows *ks, const gf_chain *chain, gf_rng *rng, const gf_table *t, hipStream_t st,
                 hipEvent_t *ev, float *ms)
{
	if (ms) for (int i = 0; i < GF_CHAIN_LAUNCHES; i++) ms[i] = 0.f;
	GF_CHECK(ks && ks->end >= ks->begin, GF_E_ARG);
	if (ks->end == ks->begin) return 0;
	gf_chain_launch L;
	int rc = chain_setup(ks, chain, rng, t, nullptr, L);
	if (rc) return rc;
	for (int i = 0; i < GF_CHAIN_LAUNCHES; i++) {
		if (ev) GF_HIP(hipEventRecord(ev[i], st));
		rc = chain_launch(L, rng, i, st);
		if (rc) return rc;
	}
	if (ev) {
		GF_HIP(hipEventRecord(ev[GF_CHAIN_LAUNCHES], st));
		if (ms) {
			GF_HIP(hipEventSynchronize(ev[GF_CHAIN_LAUNCHES]));
			for (int i = 0; i < GF_CHAIN_LAUNCHES; i++)
				if (chain_on(L, i)) GF_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
		}
	}
	return 0;
}

} // namespace gf

extern "C" int gf_module_chain(const gf_rows *ks, const gf_chain *chain, gf_rng *rng,
                               const gf_table *t, void *stream)
{
	return gf::module_chain(ks, chain, rng, t, gf::as_stream(stream), nullptr, nullptr);
}

namespace {
__global__ void __launch_bounds__(256) k_count_visible(gf_col hidden, uint64_t n, unsigned long long *out)
{
	__shared__ unsigned long long sh[256];
	unsigned long long c = 0;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
		c += col<int>(hidden, (uint32_t)i) == 0 ? 1ull : 0ull;
	sh[threadIdx.x] = c;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}
} // namespace

namespace gf {
int count_visible(gf_col hidden, uint64_t n, unsigned long long *counter, hipStream_t st)
{
	GF_CHECK(counter && hidden.ptr && n < (1ull << 32), GF_E_ARG);
	if (!n) return 0;
	unsigned grid = (unsigned)std::min<uint64_t>(2048, (n + 255) / 256);
	k_count_visible<<<grid, 256, 0, st>>>(hidden, n, counter);
	GF_LAUNCHED();
	return 0;
}
} // namespace gf

extern "C" int gf_count_visible(gf_col hidden, uint64_t n, uint64_t *visible, void *stream)
{
	GF_CHECK(visible && hidden.ptr && n < (1ull << 32), GF_E_ARG);
	hipStream_t st = gf::as_stream(stream);
	void *p;
	int rc = gf::scratch(GF_SLOT_NSLOTS + 3, sizeof(unsigned long long), &p);
	if (rc) return rc;
	GF_HIP(hipMemsetAsync(p, 0, sizeof(unsigned long long), st));
	rc = gf::count_visible(hidden, n, (unsigned long long *)p, st);
	if (rc) return rc;
	unsigned long long h = 0;
	GF_HIP(hipMemcpyAsync(&h, p, sizeof(h), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	*visible = h;
	return 0;
}
