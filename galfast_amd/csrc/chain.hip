// Per-star module chain in pipeline order (pipeline.h:75-89: FeH,
// multiples, kinematics, photometry, detector errors, database
// conversions), in at most four launches: the modules that draw random
// numbers run one lane per 256-row chunk with the chunk's stream in
// registers (stage A: FeH/GaussianFeH, multiples, Bond2010; stage C:
// photometric errors), the deterministic ones one lane per row (stage B:
// photometry; stage D: vel2pm, gal2other).  Each module consumes its draws
// in the same order as its stand-alone kernel, so the output equals the
// sequence of gf_os_*_kernel calls.
#include "gf_runtime.h"
#include "gf_modules.h"
#include <algorithm>

namespace {

constexpr unsigned CHAIN_BS = 256;

struct ChainDev {
	gf_FeH_data feh;
	gf_GaussianFeH_data gfeh;
	gf_multiples_data mult;
	gf_Bond2010_data bond;
	gf_photometry_data phot;
	gf_photoerr_data perr;
	gf_vel2pm_data v2pm;
	gf_bitmap gfeh_comps, mult_comps, phot_comps;
	int has_feh, has_gfeh, has_mult, has_bond, has_phot, has_perr, has_v2pm;
	int gal2other;
	gf_table t;
	uint32_t begin, step, end;
};

// Row range of chunk t.
GF_D bool chunk_rows(const ChainDev &c, uint32_t t, uint32_t &r0, uint32_t &r1)
{
	uint64_t a = (uint64_t)c.begin + (uint64_t)c.step * t;
	if (a >= c.end) return false;
	uint64_t b = a + c.step;
	r0 = (uint32_t)a;
	r1 = (uint32_t)(b < c.end ? b : c.end);
	return true;
}

// Stage A: the modules that draw random numbers before photometry
// (FeH or GaussianFeH, unresolvedMultiples, Bond2010), one lane per chunk.
__global__ void __launch_bounds__(CHAIN_BS) k_chain_rng_a(const ChainDev *cd, uint32_t nthr, uint32_t *st, uint32_t ns)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr) return;
	const ChainDev &c = *cd;
	if (!chunk_rows(c, t, r0, r1)) return;
	const gf_table &T = c.t;
	gf_mwc r = gf_mwc_load(st, ns, t);
	if (c.has_feh) mod_FeH(c.feh, r, r0, r1, T.comp, T.hidden, T.XYZ, T.FeH);
	if (c.has_gfeh) mod_GaussianFeH(c.gfeh, c.gfeh_comps, r, r0, r1, T.comp, T.hidden, T.FeH_comp, T.FeH);
	if (c.has_mult)
		mod_multiples(c.mult, c.mult_comps, r, r0, r1, (int)T.absmagSys.width, T.absmag, T.absmagSys,
		              T.absmagSysNcomp, T.comp, T.hidden);
	if (c.has_bond) mod_Bond2010(c.bond, r, r0, r1, T.comp, T.hidden, T.XYZ, T.vcyl);
	gf_mwc_store(st, ns, t, r);
}

// Stage B: photometry draws nothing, so it runs one lane per row.
__global__ void __launch_bounds__(CHAIN_BS) k_rows_phot(const ChainDev *cd)
{
	const ChainDev &c = *cd;
	uint32_t row = c.begin + blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= c.end) return;
	const gf_table &T = c.t;
	const gf_col &Mr = c.has_mult ? T.absmagSys : T.absmag;
	mod_photometry(c.phot, c.phot_comps, row, row + 1, T.Am, T.photoFlags, T.DM, Mr, (int)Mr.width, T.mags,
	               T.FeH, T.comp, T.hidden);
}

// Stage C: photometric errors, one lane per chunk (draws random numbers).
__global__ void __launch_bounds__(CHAIN_BS) k_chain_rng_c(const ChainDev *cd, uint32_t nthr, uint32_t *st, uint32_t ns)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr) return;
	const ChainDev &c = *cd;
	if (!chunk_rows(c, t, r0, r1)) return;
	gf_mwc r = gf_mwc_load(st, ns, t);
	mod_photoerr(c.perr, r, r0, r1, c.t.mags, c.t.obsMags, c.t.hidden);
	gf_mwc_store(st, ns, t, r);
}

// Stage D: vel2pm and gal2other, one lane per row.
__global__ void __launch_bounds__(CHAIN_BS) k_rows_astro(const ChainDev *cd)
{
	const ChainDev &c = *cd;
	uint32_t row = c.begin + blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= c.end) return;
	const gf_table &T = c.t;
	if (c.has_v2pm) mod_vel2pm(c.v2pm, row, row + 1, T.lb, T.XYZ, T.vcyl, T.pm, T.hidden);
	if (c.gal2other) mod_gal2other(c.gal2other, row, row + 1, T.hidden, T.lb, T.radec);
}

} // namespace

namespace gf {

// The chain with optional per-stage timing: when ev is non-NULL it holds 5
// events recorded around the four stages and ms[4] receives their durations
// (stages that are not enabled get 0).
int module_chain(const gf_rows *ks, const gf_chain *chain, gf_rng *rng, const gf_table *t, hipStream_t st,
                 hipEvent_t *ev, float *ms)
{
	GF_CHECK(ks && chain && t && rng && rng->state && ks->step > 0 && ks->end >= ks->begin, GF_E_ARG);
	GF_CHECK(!(chain->feh && chain->gfeh), GF_E_ARG);
	uint32_t n = ks->end - ks->begin;
	uint32_t nthr = n / ks->step + (n % ks->step ? 1 : 0);
	if (ms) for (int i = 0; i < 4; i++) ms[i] = 0.f;
	if (nthr == 0) return 0;
	GF_CHECK(nthr <= rng->nstreams, GF_E_STREAMS);
	if (chain->phot) GF_CHECK(chain->phot->ncolors >= 1 && chain->phot->ncolors <= 16, GF_E_ARG);

	static thread_local ChainDev h;   // host staging copy; lives past the async copy
	h = ChainDev{};
	if (chain->feh) { h.feh = *chain->feh; h.has_feh = 1; }
	if (chain->gfeh) { h.gfeh = *chain->gfeh; h.has_gfeh = 1; h.gfeh_comps = chain->gfeh_comps; }
	if (chain->mult) { h.mult = *chain->mult; h.has_mult = 1; h.mult_comps = chain->mult_comps; }
	if (chain->bond) { h.bond = *chain->bond; h.has_bond = 1; }
	if (chain->phot) { h.phot = *chain->phot; h.has_phot = 1; h.phot_comps = chain->phot_comps; }
	if (chain->perr) { h.perr = *chain->perr; h.has_perr = 1; }
	if (chain->v2pm) { h.v2pm = *chain->v2pm; h.has_v2pm = 1; }
	h.gal2other = chain->gal2other_coordsys;
	h.t = *t;
	h.begin = ks->begin; h.step = ks->step; h.end = ks->end;

	void *p;
	int rc = gf::scratch(GF_SLOT_NSLOTS + 2, sizeof(ChainDev), &p);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(p, &h, sizeof(ChainDev), hipMemcpyHostToDevice, st));
	const ChainDev *dp = (const ChainDev *)p;
	unsigned gchunks = gf::blocks_for(nthr, CHAIN_BS), grows = gf::blocks_for(n, CHAIN_BS);
	bool on[4] = {h.has_feh || h.has_gfeh || h.has_mult || h.has_bond, (bool)h.has_phot, (bool)h.has_perr,
	              h.has_v2pm || h.gal2other != 0};
	if (ev) GF_HIP(hipEventRecord(ev[0], st));
	if (on[0]) {
		k_chain_rng_a<<<gchunks, CHAIN_BS, 0, st>>>(dp, nthr, rng->state, rng->nstreams);
		GF_LAUNCHED();
	}
	if (ev) GF_HIP(hipEventRecord(ev[1], st));
	if (on[1]) {
		k_rows_phot<<<grows, CHAIN_BS, 0, st>>>(dp);
		GF_LAUNCHED();
	}
	if (ev) GF_HIP(hipEventRecord(ev[2], st));
	if (on[2]) {
		k_chain_rng_c<<<gchunks, CHAIN_BS, 0, st>>>(dp, nthr, rng->state, rng->nstreams);
		GF_LAUNCHED();
	}
	if (ev) GF_HIP(hipEventRecord(ev[3], st));
	if (on[3]) {
		k_rows_astro<<<grows, CHAIN_BS, 0, st>>>(dp);
		GF_LAUNCHED();
	}
	if (ev) GF_HIP(hipEventRecord(ev[4], st));
	// the staging copy must complete before the host buffer is reused
	GF_HIP(hipStreamSynchronize(st));
	if (ev && ms)
		for (int i = 0; i < 4; i++)
			if (on[i]) GF_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
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

extern "C" int gf_count_visible(gf_col hidden, uint64_t n, uint64_t *visible, void *stream)
{
	GF_CHECK(visible && hidden.ptr && n < (1ull << 32), GF_E_ARG);
	hipStream_t st = gf::as_stream(stream);
	void *p;
	int rc = gf::scratch(GF_SLOT_NSLOTS + 3, sizeof(unsigned long long), &p);
	if (rc) return rc;
	GF_HIP(hipMemsetAsync(p, 0, sizeof(unsigned long long), st));
	if (n) {
		unsigned grid = (unsigned)std::min<uint64_t>(2048, (n + 255) / 256);
		k_count_visible<<<grid, 256, 0, st>>>(hidden, n, (unsigned long long *)p);
		GF_LAUNCHED();
	}
	unsigned long long h = 0;
	GF_HIP(hipMemcpyAsync(&h, p, sizeof(h), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	*visible = h;
	return 0;
}
