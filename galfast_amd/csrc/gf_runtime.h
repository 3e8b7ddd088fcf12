// Host-side runtime helpers: error propagation, launch checks and a small
// per-device cache of scratch allocations reused across calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "../../include/galfast_hip.h"

#define GF_HIP(expr)                                   \
	do {                                               \
		hipError_t e__ = (expr);                       \
		if (e__ != hipSuccess) return (int)e__;        \
	} while (0)

#define GF_CHECK(cond, code)                           \
	do {                                               \
		if (!(cond)) return (code);                    \
	} while (0)

// Launch-status check after <<<>>>.
#define GF_LAUNCHED() GF_HIP(hipGetLastError())

namespace gf {

// Grow-only scratch buffer cached per (host thread, device, slot).  Slots
// are owned by the entry point that uses them; calls from one host thread
// are ordered, calls from different threads use different buffers.
int scratch(int slot, size_t bytes, void **ptr);

inline hipStream_t as_stream(void *s) { return (hipStream_t)s; }

inline unsigned blocks_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Exclusive prefix sum of n 32-bit counts into 64-bit offsets; offs has n+1
// entries, offs[n] = total.  Deterministic (fixed block decomposition).
int exclusive_scan_u32_u64(const uint32_t *in, uint64_t *offs, uint64_t n, hipStream_t st);

struct gf_chain_fwd;
// gf_skygen_emit with optional timing: ev holds 3 events, ms[0] and ms[1]
// receive the durations of the emission (k_emit_stars) and of the lookup of
// its tiles' units (k_tile_units) (and make the call wait for them).  unit:
// a unit at or below the one holding row_begin (from skygen_chunk_units),
// or UINT64_MAX to search from unit 0.
int skygen_emit(const gf_draw_plan *plan, uint64_t row_begin, uint64_t row_end, const gf_footprint *foot,
                gf_rng *rng, const gf_skygen_cols *out, hipStream_t st, hipEvent_t *ev, float *ms, uint64_t unit);
// gf_skygen_plan into *plan, reusing its buffers when *plan is non-NULL
// (creating it otherwise).
int skygen_plan(const gf_skygen_params *sp, const gf_model *model, const gf_pencil_beam *pixels,
                const gf_extinction *ext, const double *wpk, gf_rng *rng, gf_draw_plan **plan, uint64_t *nstars,
                hipStream_t st);
// units[q] = the plan unit holding row q * chunk, q < nq (waits for the stream).
int skygen_chunk_units(const gf_draw_plan *plan, uint64_t chunk, uint32_t nq, uint64_t *units, hipStream_t st);

// ev: GF_CHAIN_LAUNCHES + 1 events (or NULL); ms: GF_CHAIN_LAUNCHES
// durations (draw pass, then the row pass's groups)
int module_chain(const gf_rows *ks, const gf_chain *chain, gf_rng *rng, const gf_table *t, hipStream_t st,
                 hipEvent_t *ev, float *ms);
} // namespace gf

// The module chain in separate launches (launch 0: the draw pass, which
// advances the streams; launches 1..4: the row pass's groups, which only
// read the streams' multipliers and may run on another HIP stream once the
// draw pass of their rows is done).  pos: the position planes
// (chain_positions_bytes), or NULL for a per-thread scratch buffer.
struct gf_chain_launch;
struct gf_seed_seg;
namespace gf {
size_t chain_positions_bytes(const gf_chain *chain, uint32_t nrows);
int chain_setup(const gf_rows *ks, const gf_chain *chain, const gf_rng *rng, const gf_table *t, uint2 *pos,
                gf_chain_launch &L);
bool chain_on(const gf_chain_launch &L, int launch);
// Whether the launch's draw pass adds its visible rows to L.vis when set.
bool chain_counts_visible(const gf_chain_launch &L);
int chain_launch(const gf_chain_launch &L, const gf_rng *rng, int launch, hipStream_t st);
// The device address of the no-allocation scratch array on `device`
// (noalloc_setup, once per device; gf_init calls it) and on the current
// device (noalloc_scratch, setting it up on first use).
int noalloc_setup(int device);
int noalloc_scratch(void **ptr);
// Jump tables of the streams of rng (GF_JUMP_BITS * nstreams entries,
// entry b of stream s at tab[b * nstreams + s]) and every stream n steps
// ahead with such a table.
int mwc_jump_table(const gf_rng *rng, uint64_t *tab, hipStream_t st);
int mwc_jump(gf_rng *rng, uint64_t n, const uint64_t *tab, hipStream_t st);
// gf_mwc_seed_host on the device for nsegs stream tables at once (segs_dev
// in device memory, max_streams the largest nstreams among them): table g
// gets segs_dev[g].nstreams streams seeded from segs_dev[g].seed at word
// segs_dev[g].offset of out, multipliers primes[0 ..] (device memory).
int mwc_seed_streams(const gf_seed_seg *segs_dev, uint32_t nsegs, uint32_t max_streams, const uint32_t *primes,
                     uint32_t *out, hipStream_t st);
// Adds the number of rows [0, n) with hidden == 0 to *counter (device memory).
int count_visible(gf_col hidden, uint64_t n, unsigned long long *counter, hipStream_t st);

} // namespace gf

constexpr int GF_CHAIN_LAUNCHES = 6;   // the draw pass and the row pass's five groups

// Rows of the largest table a catalog chunk or a server batch may use: a
// column's pitch (gf_col.pitch, 32 bits) must hold its 8-byte planes, rounded
// up to 128 bytes.
constexpr uint64_t GF_MAX_TABLE_ROWS = ((1ull << 32) - 128) / 8;

enum gf_scratch_slot {
	GF_SLOT_LFTAB = 0,
	GF_SLOT_PARTIALS,
	GF_SLOT_SCAN,
	GF_SLOT_COMPAT_COUNTS,
	GF_SLOT_COMPAT_STATE,
	GF_SLOT_SAMPLE,
	GF_SLOT_COUNT_W,
	GF_SLOT_PARK,
	GF_SLOT_UNITS,
	GF_SLOT_ORDER,
	GF_SLOT_ORDER_TMP,
	GF_SLOT_OUT,
	GF_SLOT_OUT_COLS,
	GF_SLOT_OUT_COUNT,
	GF_SLOT_JUMP,
	GF_SLOT_SEED,
	GF_SLOT_NSLOTS
};
