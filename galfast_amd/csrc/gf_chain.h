// The per-star module chain as two passes over a table.
//
// Draw pass (one lane per 256-row chunk, the chunk's MWC stream in
// registers): for every module that draws random numbers the lane walks the
// chunk's visible rows in module order, as the module's own kernel would,
// records the stream position (c, x) at the start of each row's draws and
// moves the stream past them (feh_skip, bond_skip, ...: the uniform draws
// and rejection loops only, none of the arithmetic that turns them into
// values).  unresolvedMultiples, whose number of draws depends on the row's
// data, draws its companions here; with two absolute-magnitude planes the
// FEH row pass sums the system's magnitude, with more the draw pass does.
//
// Row pass (one lane per row, five launches): every module in pipeline
// order (FeH, GaussianFeH, Bond2010 | photometry | photometric errors |
// vel2pm, gal2other) for the row, the random-number modules restarting the
// stream from the recorded position.  A row's draws are the same numbers in the same order
// as in the sequential module kernels, so the output is identical; the
// transcendental arithmetic runs at full occupancy instead of on one lane
// per stream.
#pragma once

#include "gf_modules.h"
#include "gf_chain_types.h"


namespace gfc {

constexpr int BS = 256;      // lanes per workgroup of both passes
constexpr int TILE = 16;     // rows per tile of the draw pass
constexpr int MAXC = 256;    // rows of a chunk kept in LDS
constexpr uint8_t HIDDEN = 0xFF, NOCOMP = 0xFE;
constexpr int PHOT_LONE = 8;   // photometry launch variant flag (rows_phot)

// LDS of one draw-pass workgroup: a tile of two 32-bit values per lane, and
// the (hidden, comp) code of ROWS rows of each lane's chunk: all MAXC rows
// when the chunk's codes are read once for every module (CACHED walks,
// 96 KiB, one workgroup per CU), or the current tile's (16 KiB less per
// lane: four workgroups per CU, for launches with more chunks than lanes).
template <int ROWS>
struct DrawLdsN {
	uint32_t a[TILE][BS];
	uint32_t b[TILE][BS];
	uint8_t hc[ROWS][BS];
};
using DrawLds = DrawLdsN<MAXC>;
using DrawLdsTile = DrawLdsN<TILE>;

// Fetch n <= TILE consecutive 32-bit values starting at src into the lane's
// LDS column (vector loads when the run is whole and aligned).
GF_D void tile_fetch(const uint32_t *src, int n, uint32_t (*dst)[BS], int lane, uint32_t pad)
{
	if (n == TILE && ((uintptr_t)src & 15) == 0) {
		const uint4 *v = (const uint4 *)src;
		uint4 q[TILE / 4];
#pragma unroll
		for (int i = 0; i < TILE / 4; i++) q[i] = v[i];
#pragma unroll
		for (int i = 0; i < TILE / 4; i++) {
			dst[4 * i + 0][lane] = q[i].x;
			dst[4 * i + 1][lane] = q[i].y;
			dst[4 * i + 2][lane] = q[i].z;
			dst[4 * i + 3][lane] = q[i].w;
		}
	} else {
		for (int i = 0; i < TILE; i++) dst[i][lane] = i < n ? src[i] : pad;
	}
}

GF_D uint8_t hc_code(uint32_t hidden, uint32_t comp)
{
	if (hidden) return HIDDEN;
	return comp < 64u ? (uint8_t)comp : NOCOMP;
}

// Write the lane's n recorded positions (LDS columns a = c, b = x) to dst.
template <class Lds>
GF_D void tile_store(uint2 *dst, int n, const Lds &L, int lane)
{
	if (n == TILE && ((uintptr_t)dst & 15) == 0) {
		uint4 *v = (uint4 *)dst;
#pragma unroll
		for (int i = 0; i < TILE / 2; i++)
			v[i] = make_uint4(L.a[2 * i][lane], L.b[2 * i][lane], L.a[2 * i + 1][lane], L.b[2 * i + 1][lane]);
	} else {
		for (int i = 0; i < n; i++) dst[i] = make_uint2(L.a[i][lane], L.b[i][lane]);
	}
}

// Write the lane's n values of LDS column src to dst.
GF_D void tile_store1(uint32_t *dst, int n, const uint32_t (*src)[BS], int lane)
{
	if (n == TILE && ((uintptr_t)dst & 15) == 0) {
		uint4 *v = (uint4 *)dst;
#pragma unroll
		for (int i = 0; i < TILE / 4; i++)
			v[i] = make_uint4(src[4 * i][lane], src[4 * i + 1][lane], src[4 * i + 2][lane], src[4 * i + 3][lane]);
	} else {
		for (int i = 0; i < n; i++) dst[i] = src[i][lane];
	}
}

GF_D bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// The companion draws of unresolvedMultiples over the chunk with two
// absolute-magnitude planes, a tile at a time: the tile's absmag comes in
// through vector loads; the primary's and the companion's magnitudes and
// the component count leave through vector stores when every row of a whole
// tile takes part, otherwise row by row for the rows that take part.  The
// system magnitudes follow in the FEH row pass (rows_mult_total).
template <bool CACHED, class Lds>
GF_D void mult_walk(const gf_chain_mods &m, gf_mwc &rng, uint32_t r0, uint32_t r1, const gf_table &T, Lds &L,
                    int lane)
{
	for (uint32_t base = r0; base < r1; base += TILE) {
		const int n = (int)(r1 - base < (uint32_t)TILE ? r1 - base : TILE);
		if (!CACHED) {
			tile_fetch(&col<uint32_t>(T.hidden, base), n, L.a, lane, 1u);
			tile_fetch(&col<uint32_t>(T.comp, base), n, L.b, lane, 0u);
			for (int i = 0; i < n; i++) L.hc[i][lane] = hc_code(L.a[i][lane], L.b[i][lane]);
		}
		const uint32_t h0 = CACHED ? base - r0 : 0;
		bool all = n == TILE;
		for (int i = 0; i < n; i++) {
			const uint8_t code = L.hc[h0 + i][lane];
			all = all && code != HIDDEN && bit_isset(m.mult_comps, code == NOCOMP ? -1 : (int)code);
		}
		uint32_t *S0 = &col<uint32_t>(T.absmagSys, base, 0), *S1 = &col<uint32_t>(T.absmagSys, base, 1);
		uint32_t *Nc = &col<uint32_t>(T.absmagSysNcomp, base);
		tile_fetch(&col<uint32_t>(T.absmag, base), n, L.a, lane, 0u);
		const bool vec = all && aligned16(S0) && aligned16(S1) && aligned16(Nc);
		uint32_t two = 0;   // bit i: row base+i has a companion (vector path)
		for (int i = 0; i < n; i++) {
			if (!vec) {
				const uint8_t code = L.hc[h0 + i][lane];
				if (code == HIDDEN || !bit_isset(m.mult_comps, code == NOCOMP ? -1 : (int)code)) continue;
			}
			const float M1 = __uint_as_float(L.a[i][lane]);
			float M2 = ABSMAG_NOT_PRESENT;
			bool has = false;
			multiples_draw(m.mult, rng, M1, 2, [&](int, float v, bool h) { M2 = v; has = h; });
			if (vec) {
				L.b[i][lane] = __float_as_uint(M2);
				two |= (uint32_t)has << i;
			} else {
				S0[i] = __float_as_uint(M1);
				S1[i] = __float_as_uint(M2);
				Nc[i] = has ? 2u : 1u;
			}
		}
		if (vec) {
			tile_store1(S0, TILE, L.a, lane);
			tile_store1(S1, TILE, L.b, lane);
			uint4 *v = (uint4 *)Nc;
#pragma unroll
			for (int i = 0; i < TILE / 4; i++)
				v[i] = make_uint4(1u + ((two >> (4 * i)) & 1u), 1u + ((two >> (4 * i + 1)) & 1u),
				                  1u + ((two >> (4 * i + 2)) & 1u), 1u + ((two >> (4 * i + 3)) & 1u));
		}
	}
}

// Whether unresolvedMultiples is split between the passes (its draws in the
// draw pass, the system magnitude in the FEH row pass): with two
// absolute-magnitude planes.
GF_D bool mult_split(const gf_chain_mods &m, const gf_table &T) { return m.has_mult && T.absmagSys.width == 2; }

// One module's walk over the chunk: record each visible row's position in
// plane pos, then step past the row's draws with skip(rng, comp).
template <bool CACHED, class Lds, typename Skip>
GF_D void draw_walk(gf_mwc &rng, uint32_t r0, uint32_t r1, uint64_t prow0, uint2 *plane, const gf_table &T,
                    Lds &L, int lane, Skip skip)
{
	for (uint32_t base = r0; base < r1; base += TILE) {
		int n = (int)(r1 - base < (uint32_t)TILE ? r1 - base : TILE);
		if (!CACHED) {
			tile_fetch(&col<uint32_t>(T.hidden, base), n, L.a, lane, 1u);
			tile_fetch(&col<uint32_t>(T.comp, base), n, L.b, lane, 0u);
			for (int i = 0; i < n; i++) L.hc[i][lane] = hc_code(L.a[i][lane], L.b[i][lane]);
		}
		for (int i = 0; i < n; i++) {
			uint8_t code = CACHED ? L.hc[base - r0 + i][lane] : L.hc[i][lane];
			if (code == HIDDEN) continue;
			L.a[i][lane] = rng.c;
			L.b[i][lane] = rng.x;
			skip(rng, code == NOCOMP ? -1 : (int)code);
		}
		tile_store(plane + prow0 + (base - r0), n, L, lane);
	}
}

// The draw pass over one chunk [r0, r1) whose rows have positions at
// prow0 + (row - r0) of every plane.
// Adds every active lane's count (at most 511) to *out with one atomic per
// wave: the sum bit by bit from ballots, which see only the lanes that are
// active (lanes past the launch's end have returned).
GF_D void wave_add(unsigned long long *out, uint32_t count)
{
	unsigned long long sum = 0;
#pragma unroll
	for (int b = 0; b < 9; b++) sum += (unsigned long long)__popcll(__ballot((count >> b) & 1u)) << b;
	const unsigned long long active = __ballot(1);
	if ((int)__lane_id() == __ffsll((long long)active) - 1) atomicAdd(out, sum);
}

// vis (CACHED only): when set, the chunk's visible rows are added to it.
template <bool CACHED, class Lds>
GF_D void chunk_draws(const gf_chain_mods &m, gf_mwc &rng, uint32_t r0, uint32_t r1, uint64_t prow0,
                      const gf_table &T, const gf_positions &P, Lds &L, int lane, unsigned long long *vis = nullptr)
{
	static_assert(!CACHED || sizeof(L.hc) >= sizeof(uint8_t[MAXC][BS]), "the cached walk keeps every row's code");
	if constexpr (CACHED) {
		uint32_t nvis = 0;
		for (uint32_t base = r0; base < r1; base += TILE) {
			int n = (int)(r1 - base < (uint32_t)TILE ? r1 - base : TILE);
			tile_fetch(&col<uint32_t>(T.hidden, base), n, L.a, lane, 1u);
			tile_fetch(&col<uint32_t>(T.comp, base), n, L.b, lane, 0u);
			for (int i = 0; i < n; i++) {
				L.hc[base - r0 + i][lane] = hc_code(L.a[i][lane], L.b[i][lane]);
				nvis += L.a[i][lane] == 0u ? 1u : 0u;
			}
		}
		if (vis) wave_add(vis, nvis);
	}
	if (m.has_feh)
		draw_walk<CACHED>(rng, r0, r1, prow0, P.pos + m.pl_feh * P.pitch, T, L, lane,
		                  [&](gf_mwc &r, int cmp) { feh_skip(m.feh, r, cmp); });
	if (m.has_gfeh)
		draw_walk<CACHED>(rng, r0, r1, prow0, P.pos + m.pl_gfeh * P.pitch, T, L, lane,
		                  [&](gf_mwc &r, int cmp) { gfeh_skip(m.gfeh, m.gfeh_comps, r, cmp); });
	if (m.has_mult) {
		if (mult_split(m, T))
			mult_walk<CACHED>(m, rng, r0, r1, T, L, lane);
		else
			mod_multiples(m.mult, m.mult_comps, rng, r0, r1, (int)T.absmagSys.width, T.absmag, T.absmagSys,
			              T.absmagSysNcomp, T.comp, T.hidden);
	}
	if (m.has_bond)
		draw_walk<CACHED>(rng, r0, r1, prow0, P.pos + m.pl_bond * P.pitch, T, L, lane,
		                  [&](gf_mwc &r, int cmp) { bond_skip(m.bond, r, cmp); });
	if (m.has_perr) {
		const int nc = perr_ncurves(m.perr);
		draw_walk<CACHED>(rng, r0, r1, prow0, P.pos + m.pl_perr * P.pitch, T, L, lane,
		                  [&](gf_mwc &r, int) { perr_skip(nc, r); });
	}
}

GF_D gf_mwc restart_at(uint32_t a, uint2 q)
{
	gf_mwc r;
	r.a = a;
	r.c = q.x;
	r.x = q.y;
	return r;
}

GF_D uint2 position(const gf_positions &P, int plane, uint64_t prow) { return P.pos[plane * P.pitch + prow]; }

// The row pass, in five groups so that each launch stays small enough in
// registers to keep several waves per SIMD, and so that independent groups
// can run side by side: FEH (FeH, GaussianFeH, the system magnitude of
// unresolvedMultiples), BOND (Bond2010), PHOT
// (photometry), PERR (photometric errors), ASTRO (vel2pm, gal2other).  a is
// the multiplier of the row's stream and prow its index in the position
// planes.  Each group reads all of a row's inputs before it tests the row's
// hidden flag, so that the loads travel together (the columns exist for
// every row; a hidden row's values are read and left unused).
enum {
	ROWS_FEH = GF_ROWS_FEH,
	ROWS_BOND = GF_ROWS_BOND,
	ROWS_PHOT = GF_ROWS_PHOT,
	ROWS_PERR = GF_ROWS_PERR,
	ROWS_ASTRO = GF_ROWS_ASTRO,
	ROWS_GROUPS = GF_ROWS_GROUPS
};

inline bool rows_on(const gf_chain_mods &m, int group)
{
	if (group == ROWS_FEH) return m.has_feh || m.has_gfeh || m.has_mult;
	if (group == ROWS_BOND) return m.has_bond;
	if (group == ROWS_PHOT) return m.has_phot;
	if (group == ROWS_PERR) return m.has_perr;
	return m.has_v2pm || m.gal2other;
}

// whether the group restarts streams (needs the row's multiplier)
inline bool rows_stream(const gf_chain_mods &m, int group)
{
	return group == ROWS_FEH || group == ROWS_BOND || group == ROWS_PERR;
}

// The system magnitude of a row the draw pass gave a companion draw
// (mult_split): the luminosity sum of its one or two components.
GF_D void rows_mult_total(const gf_chain_mods &m, uint32_t row, int hidden, int cmp, const gf_table &T)
{
	const float M1 = col<float>(T.absmagSys, row, 0), M2 = col<float>(T.absmagSys, row, 1);
	const int n = col<int>(T.absmagSysNcomp, row);
	if (hidden || !bit_isset(m.mult_comps, cmp)) return;
	float Ltot = multiples_lum(M1);
	if (n == 2) Ltot += multiples_lum(M2);
	col<float>(T.absmag, row) = multiples_mag(Ltot);
}

GF_D void rows_feh(const gf_chain_mods &m, uint32_t a, uint32_t row, uint64_t prow, const gf_table &T,
                   const gf_positions &P)
{
	const int hidden = col<int>(T.hidden, row);
	const int cmp = col<int>(T.comp, row);
	if (mult_split(m, T)) rows_mult_total(m, row, hidden, cmp, T);
	if (m.has_feh) {
		const float x = col<float>(T.XYZ, row, 0), y = col<float>(T.XYZ, row, 1), z = col<float>(T.XYZ, row, 2);
		const uint2 q = position(P, m.pl_feh, prow);
		if (hidden) return;
		gf_mwc r = restart_at(a, q);
		float feh;
		if (feh_row(m.feh, r, cmp, x, y, z, feh)) col<float>(T.FeH, row) = feh;
	}
	if (m.has_gfeh) {
		const uint2 q = position(P, m.pl_gfeh, prow);
		if (hidden) return;
		gf_mwc r = restart_at(a, q);
		gfeh_row(m.gfeh, m.gfeh_comps, r, cmp, row, T.FeH_comp, T.FeH);
	}
}

GF_D void rows_bond(const gf_chain_mods &m, uint32_t a, uint32_t row, uint64_t prow, const gf_table &T,
                    const gf_positions &P)
{
	const int hidden = col<int>(T.hidden, row);
	const int cmp = col<int>(T.comp, row);
	const float x = col<float>(T.XYZ, row, 0), y = col<float>(T.XYZ, row, 1), z = col<float>(T.XYZ, row, 2);
	const uint2 q = position(P, m.pl_bond, prow);
	if (hidden) return;
	gf_mwc r = restart_at(a, q);
	float v[3];
	if (m.bond_one_disk ? bond_row<true>(m.bond, r, cmp, x, y, z, v) : bond_row<false>(m.bond, r, cmp, x, y, z, v)) {
		col<float>(T.vcyl, row, 0) = v[0];
		col<float>(T.vcyl, row, 1) = v[1];
		col<float>(T.vcyl, row, 2) = v[2];
	}
}

// NT: the colour tables (1..4) plus PHOT_LONE when every row is a lone star
template <int NT>
GF_D void rows_phot(const gf_chain_mods &m, uint32_t row, const gf_table &T)
{
	constexpr bool LONE = (NT & PHOT_LONE) != 0;
	const gf_col &Mr = m.has_mult ? T.absmagSys : T.absmag;
	const int nabsmag = (int)Mr.width;
	const int hidden = col<int>(T.hidden, row);
	const int cmp = col<int>(T.comp, row);
	const float feh = col<float>(T.FeH, row), dm = col<float>(T.DM, row), am = col<float>(T.Am, row);
	const float mr0 = col<float>(Mr, row, 0);
	const float mr1 = !LONE && nabsmag > 1 ? col<float>(Mr, row, 1) : ABSMAG_NOT_PRESENT;
	if (hidden || !bit_isset(m.phot_comps, cmp)) return;
	phot_row<(NT & 7), LONE>(m.phot, row, feh, dm, am, mr0, mr1, T.photoFlags, Mr, nabsmag, T.mags);
}

GF_D void rows_perr(const gf_chain_mods &m, uint32_t a, uint32_t row, uint64_t prow, const gf_table &T,
                    const gf_positions &P)
{
	const int hidden = col<int>(T.hidden, row);
	const uint2 q = position(P, m.pl_perr, prow);
	if (perr_fast(m.perr)) {
		float mag[PERR_FAST_BANDS];
#pragma unroll
		for (int b = 0; b < PERR_FAST_BANDS; b++)
			if (b < m.perr.nbands) mag[b] = col<float>(T.mags, row, b);
		if (hidden) return;
		gf_mwc r = restart_at(a, q);
		perr_row_fast(m.perr, r, row, mag, T.obsMags);
		return;
	}
	if (hidden) return;
	gf_mwc r = restart_at(a, q);
	perr_row(m.perr, r, row, T.mags, T.obsMags);
}

GF_D void rows_astro(const gf_chain_mods &m, uint32_t row, const gf_table &T)
{
	const int hidden = col<int>(T.hidden, row);
	const double l = col<double>(T.lb, row, 0), b = col<double>(T.lb, row, 1);
	if (m.has_v2pm) {
		const float x = col<float>(T.XYZ, row, 0), y = col<float>(T.XYZ, row, 1), z = col<float>(T.XYZ, row, 2);
		const float vr = col<float>(T.vcyl, row, 0), vp = col<float>(T.vcyl, row, 1), vz = col<float>(T.vcyl, row, 2);
		if (hidden) return;
		vel2pm_row(m.v2pm, row, l, b, x, y, z, vr, vp, vz, T.pm);
	}
	if (hidden) return;
	if (m.gal2other) gal2other_row(m.gal2other, row, l, b, T.radec);
}

// GROUP: one of ROWS_*; NT: colour tables of the photometry (1..4, at
// least ceil(ncolors / 4); see phot_tables())
template <int GROUP, int NT>
GF_D void chain_rows(const gf_chain_mods &m, uint32_t a, uint32_t row, uint64_t prow, const gf_table &T,
                     const gf_positions &P)
{
	if (GROUP == ROWS_FEH) rows_feh(m, a, row, prow, T, P);
	else if (GROUP == ROWS_BOND) rows_bond(m, a, row, prow, T, P);
	else if (GROUP == ROWS_PHOT) rows_phot<NT>(m, row, T);
	else if (GROUP == ROWS_PERR) rows_perr(m, a, row, prow, T, P);
	else rows_astro(m, row, T);
}

inline int phot_tables(const gf_chain_mods &m)
{
	return m.has_phot ? ((int)m.phot.ncolors + 3) / 4 : 1;
}

// The photometry launch variant: phot_tables(), plus PHOT_LONE when the
// absolute-magnitude input has one column (no unresolved multiples), so that
// no row is a system of several stars.
inline int phot_variant(const gf_chain_mods &m, const gf_table &T)
{
	const gf_col &Mr = m.has_mult ? T.absmagSys : T.absmag;
	return phot_tables(m) | (Mr.width == 1 ? PHOT_LONE : 0);
}

} // namespace gfc
