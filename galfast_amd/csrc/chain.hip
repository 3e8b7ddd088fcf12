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

// Rows per tile of the RNG stages.  A lane walks its 256-row chunk in
// tiles: the tile's inputs are fetched with independent vector loads into
// the lane's own LDS slots (so the loads of 16 rows are in flight together),
// then the rows are processed in order from LDS.  Only the load schedule
// changes; each row's arithmetic and draws are those of the module.
constexpr int TILE = 16;

template <typename T>
GF_D void tile_load(const gf_col &c, uint32_t plane, uint32_t row0, int n, T (*dst)[CHAIN_BS], int lane, T pad)
{
	const T *src = &col<T>(c, row0, plane);
	if (n == TILE && ((uintptr_t)src & 15) == 0) {
		const uint4 *v = (const uint4 *)src;
		uint4 q[TILE / 4];
#pragma unroll
		for (int i = 0; i < TILE / 4; i++) q[i] = v[i];
#pragma unroll
		for (int i = 0; i < TILE / 4; i++) {
			uint32_t w[4] = {q[i].x, q[i].y, q[i].z, q[i].w};
#pragma unroll
			for (int k = 0; k < 4; k++) dst[4 * i + k][lane] = __builtin_bit_cast(T, w[k]);
		}
	} else {
		for (int i = 0; i < TILE; i++) dst[i][lane] = i < n ? src[i] : pad;
	}
}

// Stage A: the modules that draw random numbers before photometry
// (FeH or GaussianFeH, unresolvedMultiples, Bond2010), one lane per chunk.
__global__ void __launch_bounds__(CHAIN_BS) k_chain_rng_a(const ChainDev *cd, uint32_t nthr, uint32_t *st, uint32_t ns)
{
	__shared__ int s_hid[TILE][CHAIN_BS], s_cmp[TILE][CHAIN_BS];
	__shared__ float s_x[TILE][CHAIN_BS], s_y[TILE][CHAIN_BS], s_z[TILE][CHAIN_BS];
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr) return;
	const ChainDev &c = *cd;
	if (!chunk_rows(c, t, r0, r1)) return;
	const gf_table &T = c.t;
	const int lane = threadIdx.x;
	gf_mwc r = gf_mwc_load(st, ns, t);
	if (c.has_feh) {
		for (uint32_t base = r0; base < r1; base += TILE) {
			int n = (int)(r1 - base < (uint32_t)TILE ? r1 - base : TILE);
			tile_load<int>(T.hidden, 0, base, n, s_hid, lane, 1);
			tile_load<int>(T.comp, 0, base, n, s_cmp, lane, 0);
			tile_load<float>(T.XYZ, 0, base, n, s_x, lane, 0.f);
			tile_load<float>(T.XYZ, 1, base, n, s_y, lane, 0.f);
			tile_load<float>(T.XYZ, 2, base, n, s_z, lane, 0.f);
			for (int i = 0; i < n; i++) {
				if (s_hid[i][lane]) continue;
				float feh;
				if (feh_row(c.feh, r, s_cmp[i][lane], s_x[i][lane], s_y[i][lane], s_z[i][lane], feh))
					col<float>(T.FeH, base + i) = feh;
			}
		}
	}
	if (c.has_gfeh) mod_GaussianFeH(c.gfeh, c.gfeh_comps, r, r0, r1, T.comp, T.hidden, T.FeH_comp, T.FeH);
	if (c.has_mult)
		mod_multiples(c.mult, c.mult_comps, r, r0, r1, (int)T.absmagSys.width, T.absmag, T.absmagSys,
		              T.absmagSysNcomp, T.comp, T.hidden);
	if (c.has_bond) {
		for (uint32_t base = r0; base < r1; base += TILE) {
			int n = (int)(r1 - base < (uint32_t)TILE ? r1 - base : TILE);
			tile_load<int>(T.hidden, 0, base, n, s_hid, lane, 1);
			tile_load<int>(T.comp, 0, base, n, s_cmp, lane, 0);
			tile_load<float>(T.XYZ, 0, base, n, s_x, lane, 0.f);
			tile_load<float>(T.XYZ, 1, base, n, s_y, lane, 0.f);
			tile_load<float>(T.XYZ, 2, base, n, s_z, lane, 0.f);
			for (int i = 0; i < n; i++) {
				if (s_hid[i][lane]) continue;
				float v[3];
				if (!bond_row(c.bond, r, s_cmp[i][lane], s_x[i][lane], s_y[i][lane], s_z[i][lane], v)) continue;
				col<float>(T.vcyl, base + i, 0) = v[0];
				col<float>(T.vcyl, base + i, 1) = v[1];
				col<float>(T.vcyl, base + i, 2) = v[2];
			}
		}
	}
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

// Stage C: photometric errors, one lane per chunk (draws random numbers),
// bands outer and rows inner as mod_photoerr, inputs fetched per tile.
__global__ void __launch_bounds__(CHAIN_BS) k_chain_rng_c(const ChainDev *cd, uint32_t nthr, uint32_t *st, uint32_t ns)
{
	__shared__ int s_hid[TILE][CHAIN_BS];
	__shared__ float s_mag[TILE][CHAIN_BS];
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr) return;
	const ChainDev &c = *cd;
	if (!chunk_rows(c, t, r0, r1)) return;
	const gf_table &T = c.t;
	const gf_photoerr_data &par = c.perr;
	const int lane = threadIdx.x;
	gf_mwc r = gf_mwc_load(st, ns, t);
	for (int b = 0; b < par.nbands; b++) {
		int k0 = par.koff[b], k1 = par.koff[b + 1];
		for (uint32_t base = r0; base < r1; base += TILE) {
			int n = (int)(r1 - base < (uint32_t)TILE ? r1 - base : TILE);
			tile_load<int>(T.hidden, 0, base, n, s_hid, lane, 1);
			tile_load<float>(T.mags, b, base, n, s_mag, lane, 0.f);
			for (int i = 0; i < n; i++) {
				if (s_hid[i][lane]) continue;
				float mag = s_mag[i][lane];
				float obs = mag;
				if (k1 > k0) {
					float sigma = (float)spline_eval(par.kx + k0, par.ky + k0, k1 - k0, (double)mag);
					obs = mag + r.gaussian(sigma);
				}
				col<float>(T.obsMags, base + i, b) = obs;
			}
		}
	}
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
