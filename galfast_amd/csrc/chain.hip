// Fused per-star module chain: one launch runs every enabled module over
// each thread's 256-row chunk, in pipeline order (pipeline.h:75-89:
// FeH, multiples, kinematics, photometry, detector errors, database
// conversions).  The chunk's RNG stream is loaded once and stored once; each
// module consumes draws in the same order as its stand-alone kernel, so the
// output equals the unfused sequence of gf_os_*_kernel calls.
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

__global__ void __launch_bounds__(CHAIN_BS) k_chain(const ChainDev *cd, uint32_t nthr, uint32_t *st, uint32_t ns)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nthr) return;
	const ChainDev &c = *cd;
	uint64_t a = (uint64_t)c.begin + (uint64_t)c.step * t;
	if (a >= c.end) return;
	uint64_t b = a + c.step;
	uint32_t r0 = (uint32_t)a, r1 = (uint32_t)(b < c.end ? b : c.end);
	const gf_table &T = c.t;
	gf_mwc r = gf_mwc_load(st, ns, t);
	if (c.has_feh) mod_FeH(c.feh, r, r0, r1, T.comp, T.hidden, T.XYZ, T.FeH);
	if (c.has_gfeh) mod_GaussianFeH(c.gfeh, c.gfeh_comps, r, r0, r1, T.comp, T.hidden, T.FeH_comp, T.FeH);
	if (c.has_mult)
		mod_multiples(c.mult, c.mult_comps, r, r0, r1, (int)T.absmagSys.width, T.absmag, T.absmagSys,
		              T.absmagSysNcomp, T.comp, T.hidden);
	if (c.has_bond) mod_Bond2010(c.bond, r, r0, r1, T.comp, T.hidden, T.XYZ, T.vcyl);
	if (c.has_phot) {
		const gf_col &Mr = c.has_mult ? T.absmagSys : T.absmag;
		mod_photometry(c.phot, c.phot_comps, r0, r1, T.Am, T.photoFlags, T.DM, Mr, (int)Mr.width, T.mags,
		               T.FeH, T.comp, T.hidden);
	}
	if (c.has_perr) mod_photoerr(c.perr, r, r0, r1, T.mags, T.obsMags, T.hidden);
	if (c.has_v2pm) mod_vel2pm(c.v2pm, r0, r1, T.lb, T.XYZ, T.vcyl, T.pm, T.hidden);
	if (c.gal2other) mod_gal2other(c.gal2other, r0, r1, T.hidden, T.lb, T.radec);
	gf_mwc_store(st, ns, t, r);
}

} // namespace

extern "C" int gf_module_chain(const gf_rows *ks, const gf_chain *chain, gf_rng *rng,
                               const gf_table *t, void *stream)
{
	GF_CHECK(ks && chain && t && rng && rng->state && ks->step > 0 && ks->end >= ks->begin, GF_E_ARG);
	GF_CHECK(!(chain->feh && chain->gfeh), GF_E_ARG);
	uint32_t n = ks->end - ks->begin;
	uint32_t nthr = n / ks->step + (n % ks->step ? 1 : 0);
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

	hipStream_t st = gf::as_stream(stream);
	void *p;
	int rc = gf::scratch(GF_SLOT_NSLOTS + 2, sizeof(ChainDev), &p);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(p, &h, sizeof(ChainDev), hipMemcpyHostToDevice, st));
	k_chain<<<gf::blocks_for(nthr, CHAIN_BS), CHAIN_BS, 0, st>>>((const ChainDev *)p, nthr, rng->state, rng->nstreams);
	GF_LAUNCHED();
	// the staging copy must complete before the host buffer is reused
	GF_HIP(hipStreamSynchronize(st));
	return 0;
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
