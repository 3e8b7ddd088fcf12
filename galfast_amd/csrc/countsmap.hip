// Sky-count maps on the device: the countsMap output stage (os_countsMap,
// pipeline.cpp:293-762).  The reference bins rows on the host, one thread
// per core into its own std::map of beams, then merges the maps
// (os_countsMap::process :531-575).  Here the map is one dense array in HBM,
// [beam][Z][zwidth] uint32 with beam = (X ny + Y) nmaps + map, that every
// row of a chunk adds to with device atomics, plus a byte per beam for "a
// row landed here" (the reference creates a beam's vector on its first row,
// and only such beams are written unless every bin is asked for).
//
// A coadd adds the row to every z bin from its depth bin up
// (:463-476); here it adds one count at the depth bin and
// gf_countsmap_finish turns the coadd entries into running sums over Z
// once, after the last chunk: the same counts with one atomic per row
// instead of up to nz.
#include "gf_runtime.h"
#include "gf_device.h"
#include <climits>

namespace {

constexpr int CB = 256;
constexpr double D2R = 0.017453292519943295;   // pi / 180, the double nearest

// find_bin (pipeline.cpp:361-367): b = (int)((x - x0) / dx + 0.5), clamped
// to [0, n - 1].  The conversion of a value outside int's range, or of a
// NaN, gives INT_MIN on the x86-64 hosts the reference runs on: bin 0.
GF_D int find_bin(double x, double x0, double dx, int n)
{
	const double v = (x - x0) / dx + 0.5;
	int b = (v > -2147483649.0 && v < 2147483648.0) ? (int)v : INT_MIN;
	b = b < 0 ? 0 : b;
	return b > n - 1 ? n - 1 : b;
}

// peyton::math::lambert::project (common/projections.h:46-55) of (l, phi)
// = rad(x, y); x and y keep their values when the point is the projection's
// antipode (denominator < 1e-10), as the reference's do.
GF_D void lambert_project(const double *pr, double &x, double &y)
{
	const double l = x * D2R, phi = y * D2R;
	double sp, cp, sdl, cdl;
	gfm::sincos(phi, &sp, &cp);
	gfm::sincos(l - pr[0], &sdl, &cdl);
	const double denom = 1 + pr[1] * sp + pr[2] * cp * cdl;
	if (denom < 1e-10) return;
	const double kp = __builtin_sqrt(2. / denom);
	x = kp * cp * sdl;
	y = kp * (pr[2] * sp - pr[1] * cp * cdl);
}

// Adds one to counts[key] for every active lane of the wave.  Lanes with
// equal keys are combined into one atomic (rows arrive in sky order, so a
// wave's rows mostly share a beam and a few z bins); after AGG_ROUNDS
// distinct keys the remaining lanes add their own.
constexpr int AGG_ROUNDS = 12;

GF_D void wave_add(uint32_t *counts, uint64_t key, bool on)
{
	const int lane = (int)__lane_id();
	uint64_t active = __ballot(on);
	for (int round = 0; active; round++) {
		if (round == AGG_ROUNDS) {
			if ((active >> lane) & 1) atomicAdd(counts + key, 1u);
			return;
		}
		const int leader = __ffsll((unsigned long long)active) - 1;
		const uint64_t k0 = __shfl(key, leader);
		const uint64_t same = __ballot(on && key == k0) & active;
		if (lane == leader) atomicAdd(counts + k0, (uint32_t)__popcll(same));
		active &= ~same;
	}
}

__global__ void __launch_bounds__(CB) k_countsmap(const gf_countsmap_params p, gf_col xy, gf_col z, gf_col hidden,
                                                  uint32_t n, uint32_t *counts, uint8_t *beams,
                                                  unsigned long long *total)
{
	const int zm = p.zmag_width, zw = zm + p.ncoadd, nmaps = p.equalarea ? 2 : 1;
	unsigned long long mine = 0;
	// every lane of a wave takes the same trips, so that the wave's lanes
	// combine their counts
	const uint64_t wave0 = (uint64_t)blockIdx.x * CB + (threadIdx.x & ~63u);
	for (uint64_t w = wave0; w < n; w += (uint64_t)gridDim.x * CB) {
		const uint64_t i = w + __lane_id();
		const uint32_t r = (uint32_t)i;
		const bool on = i < n && col<int>(hidden, r) == 0;
		uint64_t base = 0;
		if (on) {
			double x = col<double>(xy, r, 0), y = col<double>(xy, r, 1);
			int map = 0;
			if (p.equalarea) {
				map = y > 0 ? 0 : 1;
				lambert_project(p.proj[map], x, y);
			}
			const int X = find_bin(x, p.x0, p.dx, p.nx), Y = find_bin(y, p.y0, p.dy, p.ny);
			const uint64_t beam = ((uint64_t)X * p.ny + Y) * nmaps + map;
			beams[beam] = 1;
			base = beam * (uint64_t)(p.nz * zw);
			mine += zm;
		}
		for (int e = 0; e < zm; e++) {
			const uint64_t key = on ? base + (uint64_t)find_bin(col<float>(z, r, e), p.z0, p.dz, p.nz) * zw + e : 0;
			wave_add(counts, key, on);
		}
		// the coadd depth bin: the smallest z_e - offset_e over the bands
		// (std::min as the reference writes it: a NaN term is skipped)
		for (int k = 0; k < p.ncoadd; k++) {
			uint64_t key = 0;
			if (on) {
				float cut = __builtin_inff();
				for (int e = 0; e < zm; e++) {
					const float m = col<float>(z, r, e) - p.coadd_offsets[k * zm + e];
					cut = m < cut ? m : cut;
				}
				key = base + (uint64_t)find_bin(cut, p.z0, p.dz, p.nz) * zw + zm + k;
			}
			wave_add(counts, key, on);
		}
	}
	for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o);
	if ((threadIdx.x & 63) == 0 && mine) atomicAdd(total, mine);
}

// The same binning with each block's counts gathered in LDS first: a block
// takes tiles of TILE consecutive rows (a few sky pixels of a chunk in sky
// order, so a few beams), gives each beam it meets one of LDS_BEAMS slots of
// nz x zwidth counts, and adds the slots to the map once per tile.  Rows
// whose beam finds no free slot go to the map directly.
constexpr int LDS_BEAMS = 8;
constexpr int TILE = CB * 16;

__global__ void __launch_bounds__(CB) k_countsmap_lds(const gf_countsmap_params p, gf_col xy, gf_col z,
                                                      gf_col hidden, uint32_t n, uint32_t *counts, uint8_t *beams,
                                                      unsigned long long *total)
{
	extern __shared__ uint32_t h[];            // LDS_BEAMS x nz x zwidth
	__shared__ uint32_t tag[LDS_BEAMS];        // beam + 1 of each slot, 0 = free
	const int zm = p.zmag_width, zw = zm + p.ncoadd, nmaps = p.equalarea ? 2 : 1;
	const int per = p.nz * zw;
	unsigned long long mine = 0;
	for (uint64_t t0 = (uint64_t)blockIdx.x * TILE; t0 < n; t0 += (uint64_t)gridDim.x * TILE) {
		for (int j = threadIdx.x; j < LDS_BEAMS * per; j += CB) h[j] = 0;
		if (threadIdx.x < LDS_BEAMS) tag[threadIdx.x] = 0;
		__syncthreads();
		for (int j = 0; j < TILE / CB; j++) {
			const uint64_t i = t0 + (uint64_t)j * CB + threadIdx.x;
			if (i >= n) break;
			const uint32_t r = (uint32_t)i;
			if (col<int>(hidden, r)) continue;
			double x = col<double>(xy, r, 0), y = col<double>(xy, r, 1);
			int map = 0;
			if (p.equalarea) {
				map = y > 0 ? 0 : 1;
				lambert_project(p.proj[map], x, y);
			}
			const int X = find_bin(x, p.x0, p.dx, p.nx), Y = find_bin(y, p.y0, p.dy, p.ny);
			const uint64_t beam = ((uint64_t)X * p.ny + Y) * nmaps + map;
			// the beam's slot: the one tagged with it, else the first free one
			const uint32_t want = (uint32_t)beam + 1u;
			int slot = -1;
			for (int q = 0; q < LDS_BEAMS; q++) {
				const uint32_t old = atomicCAS(&tag[q], 0u, want);
				if (old == 0u || old == want) { slot = q; break; }
			}
			uint32_t *c;
			if (slot >= 0) {
				c = h + slot * per;
			} else {
				beams[beam] = 1;
				c = counts + beam * (uint64_t)per;
			}
			for (int e = 0; e < zm; e++) atomicAdd(c + find_bin(col<float>(z, r, e), p.z0, p.dz, p.nz) * zw + e, 1u);
			mine += zm;
			// the coadd depth bin: the smallest z_e - offset_e over the
			// bands (std::min as the reference writes it: a NaN is skipped)
			for (int k = 0; k < p.ncoadd; k++) {
				float cut = __builtin_inff();
				for (int e = 0; e < zm; e++) {
					const float m = col<float>(z, r, e) - p.coadd_offsets[k * zm + e];
					cut = m < cut ? m : cut;
				}
				atomicAdd(c + find_bin(cut, p.z0, p.dz, p.nz) * zw + zm + k, 1u);
			}
		}
		__syncthreads();
		for (int q = 0; q < LDS_BEAMS; q++) {
			const uint32_t tg = tag[q];
			if (tg == 0u) break;                   // slots fill in order
			const uint64_t beam = tg - 1u;
			if (threadIdx.x == 0) beams[beam] = 1;
			uint32_t *dst = counts + beam * (uint64_t)per;
			for (int j = threadIdx.x; j < per; j += CB) {
				const uint32_t v = h[q * per + j];
				if (v) atomicAdd(dst + j, v);
			}
		}
		__syncthreads();
	}
	for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o);
	if ((threadIdx.x & 63) == 0 && mine) atomicAdd(total, mine);
}

__global__ void __launch_bounds__(CB) k_countsmap_finish(uint64_t npairs, int ncoadd, int zm, int zw, int nz,
                                                         uint32_t *counts)
{
	const uint64_t q = (uint64_t)blockIdx.x * CB + threadIdx.x;
	if (q >= npairs) return;
	uint32_t *c = counts + (q / ncoadd) * (uint64_t)(nz * zw) + zm + (int)(q % ncoadd);
	uint32_t s = 0;
	for (int Z = 0; Z < nz; Z++) {
		s += c[(size_t)Z * zw];
		c[(size_t)Z * zw] = s;
	}
}

int check_params(const gf_countsmap_params *p)
{
	GF_CHECK(p && p->nx > 0 && p->ny > 0 && p->nz > 0 && p->zmag_width > 0 && p->ncoadd >= 0, GF_E_ARG);
	GF_CHECK(p->ncoadd == 0 || p->coadd_offsets, GF_E_ARG);
	const uint64_t nmaps = p->equalarea ? 2 : 1;
	GF_CHECK((uint64_t)p->nx * p->ny * nmaps * p->nz * (p->zmag_width + p->ncoadd) < (1ull << 40), GF_E_ARG);
	GF_CHECK((uint64_t)p->nz * (p->zmag_width + p->ncoadd) < (1ull << 31), GF_E_ARG);
	return 0;
}

} // namespace

extern "C" {

int gf_countsmap_bin(const gf_countsmap_params *p, gf_col xy, gf_col z, gf_col hidden, uint64_t nrows,
                     uint32_t *counts, uint8_t *beams, unsigned long long *total, void *stream)
{
	int rc = check_params(p);
	if (rc) return rc;
	GF_CHECK(counts && beams && total && nrows < (1ull << 32), GF_E_ARG);
	if (nrows == 0) return 0;
	GF_CHECK(xy.ptr && xy.elem == 8 && xy.width >= 2, GF_E_ARG);
	GF_CHECK(z.ptr && z.elem == 4 && z.width >= (uint32_t)p->zmag_width && hidden.ptr, GF_E_ARG);
	const size_t lds = (size_t)LDS_BEAMS * p->nz * (p->zmag_width + p->ncoadd) * sizeof(uint32_t);
	const uint64_t nbeams = (uint64_t)p->nx * p->ny * (p->equalarea ? 2 : 1);
	if (lds <= 48 * 1024 && nbeams < 0xffffffffull) {   // slots tag beams as beam + 1 in 32 bits
		const unsigned g = (unsigned)std::min<uint64_t>((nrows + TILE - 1) / TILE, 4096);
		k_countsmap_lds<<<g, CB, lds, gf::as_stream(stream)>>>(*p, xy, z, hidden, (uint32_t)nrows, counts, beams,
		                                                      total);
	} else {
		const unsigned g = (unsigned)std::min<uint64_t>(gf::blocks_for(nrows, CB), 8192);
		k_countsmap<<<g, CB, 0, gf::as_stream(stream)>>>(*p, xy, z, hidden, (uint32_t)nrows, counts, beams, total);
	}
	GF_LAUNCHED();
	return 0;
}

int gf_countsmap_finish(const gf_countsmap_params *p, uint32_t *counts, void *stream)
{
	int rc = check_params(p);
	if (rc) return rc;
	GF_CHECK(counts, GF_E_ARG);
	if (p->ncoadd == 0) return 0;
	const uint64_t npairs = (uint64_t)p->nx * p->ny * (p->equalarea ? 2 : 1) * p->ncoadd;
	GF_CHECK(npairs < (1ull << 31) * (uint64_t)CB, GF_E_ARG);
	k_countsmap_finish<<<gf::blocks_for(npairs, CB), CB, 0, gf::as_stream(stream)>>>(
		npairs, p->ncoadd, p->zmag_width, p->zmag_width + p->ncoadd, p->nz, counts);
	GF_LAUNCHED();
	return 0;
}

} // extern "C"
