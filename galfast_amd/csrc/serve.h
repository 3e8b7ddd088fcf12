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
	uint32_t lane0;         // first plan/emit lane of the segment in its launch
	uint32_t S;             // streams of the request
	int32_t model;          // index into the model array
	int32_t req;            // request index in the batch
	int32_t pad;
};

// One request of a batch, for the module chain.
struct gf_serve_req {
	uint32_t *state;
	uint64_t row0, row1;
	uint32_t lane0;
	uint32_t S;
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
int serve_count(const gf_serve_seg *dsegs, int nseg, uint32_t nslots, const gf_model *dmodels,
                const gf_pencil_beam *dpix, double *W, double *pix_sum, double *seg_expected, hipStream_t st);
int serve_plan(const gf_serve_seg *dsegs, int nseg, uint32_t nlanes, const double *W, uint32_t *counts,
               hipStream_t st);
int serve_seg_rows(gf_serve_seg *dsegs, int nseg, const uint64_t *offs, hipStream_t st);
// park: n rows of parked stars (device scratch of the caller)
int serve_emit(const gf_serve_seg *dsegs, int nseg, uint32_t nlanes, int nM, const gf_pencil_beam *dpix,
               const uint64_t *offs, const gf_parked &park, hipStream_t st);
int serve_emit_rows(const gf_serve_seg *dsegs, int nseg, const gf_model *dmodels, const gf_pencil_beam *dpix,
                    uint64_t row0, uint64_t row1, const gf_parked &park, const gf_skygen_cols &out, hipStream_t st);
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
