// Device-side descriptors of the request server (host layer: host/server.cpp,
// kernels: serve.hip).
#pragma once

#include <hip/hip_runtime.h>
#include "../../include/galfast_hip.h"
#include "gf_chain.h"
#include "gf_device.h"

// One (request, model) segment of a batch.
struct gf_serve_seg {
	gf_skygen_params sp;    // the request's binning (its own m1)
	gf_footprint foot;      // the request's footprint
	const double *lfP;      // FP64 prefix of the model's LF at the bin centres (nM + 1)
	uint32_t *state;        // the request's stream table [a | c | x], S streams
	uint64_t unit0;         // first unit (pixel x DM bin) of the segment in the batch
	uint64_t nunits;
	uint64_t row0, row1;    // output rows (after the plan scan)
	uint32_t pix0, npix;    // the request's pixels in the batch pixel array
	uint32_t slot0;         // first count-kernel workgroup of the segment in its launch
	uint32_t lane0;         // first lane of the segment's request in per-stream launches
	uint32_t S;             // streams of the request
	int32_t model;          // index into the model array
	int32_t req;            // request index in the batch
	uint32_t i0;            // unit index of the segment's first unit on its streams (models before it)
};

// One request of a batch, for the module chain.
struct gf_serve_req {
	uint32_t *state;
	uint64_t row0, row1;
	uint32_t lane0;         // first lane of the request: it has one per stream that owns a 256-row chunk
	uint32_t S;
	uint64_t advance;       // draws its streams move on once all its models are planned
};

// Module parameter blocks of the server (device copy).
typedef gf_chain_mods gf_serve_chain;

// What the captured module chain of a batch reads from device memory: the
// batch's requests, lanes and rows, its table and position planes.
struct gf_serve_chain_args {
	const gf_serve_req *reqs;
	int32_t nreq;
	uint32_t nlanes;
	uint64_t nrows;
	gf_table T;
	gf_positions P;
};

namespace gf {
int serve_count(const gf_serve_seg *dsegs, int nseg, uint32_t nslots, const gf_model *dmodels, const gf_extinction &ext,
                const gf_pencil_beam *dpix, double *W, double *pix_sum, double *seg_expected, hipStream_t st);
// Plan of a batch: lane per unit over all nseg segments (in unit order);
// counts and zu (the states the units' stars start from) per batch unit.
// jtab: jump tables of the first 65,536 streams (stride 65,536).
int serve_plan(const gf_serve_seg *dsegs, int nseg, uint64_t nunits, const double *W, const uint64_t *jtab,
               uint32_t *counts, uint64_t *zu, hipStream_t st);
int serve_seg_rows(gf_serve_seg *dsegs, int nseg, const uint64_t *offs, hipStream_t st);
// roff[r] (r <= nreq): request r's first row, roff[nreq] the batch's rows,
// from the plan's scanned offsets (segall: nm segments per request,
// request-major; nunits: the batch's units)
int serve_req_rows(const gf_serve_seg *dsegall, int nreq, int nm, const uint64_t *offs, uint64_t nunits,
                   uint64_t *roff, hipStream_t st);
// Emission: lane per batch row over all segments (in row order).
int serve_emit(const gf_serve_seg *dsegs, int nseg, const gf_model *dmodels, const gf_extinction &ext,
               const gf_pencil_beam *dpix,
               const uint64_t *offs, const uint64_t *zu, const uint64_t *jtab, uint64_t nrows,
               const gf_skygen_cols &out, hipStream_t st);
// Every request's streams move on by its `advance` (nlanes = its streams in all).
int serve_advance(const gf_serve_req *dreqs, int nreq, uint32_t nlanes, const uint64_t *jtab, hipStream_t st);
// pos: hch.nplanes * nrows stream positions (device scratch of the caller)
int serve_chain(const gf_serve_req *dreqs, int nreq, uint32_t nlanes, const gf_serve_chain *dch,
                const gf_serve_chain &hch, uint64_t nrows, const gf_table &T, uint2 *pos, hipStream_t st);
// The module chain of serve_chain captured as a HIP graph whose launches
// read the batch from *dargs (fixed grids that stride over it), so that one
// instantiation serves every batch: write the batch's gf_serve_chain_args to
// dargs, then launch *exec on the same stream.
int serve_chain_graph(const gf_serve_chain_args *dargs, const gf_serve_chain *dch, const gf_serve_chain &hch,
                      hipStream_t st, hipGraphExec_t *exec);
// vis[r] = rows of request r with hidden == 0 (max_rows: the largest request)
int serve_visible(const gf_serve_req *dreqs, int nreq, uint64_t max_rows, const gf_col &hidden,
                  unsigned long long *vis, hipStream_t st);
// FP64 exclusive prefix of a model's LF at the absolute-magnitude bin centres
// (nM + 1 doubles) into dst (device).
int lf_prefix(const gf_skygen_params &sp, const gf_model &m, double *dst, hipStream_t st);
} // namespace gf
