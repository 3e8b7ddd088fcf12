// Launch descriptor of the per-star module chain: module parameter blocks,
// position planes and the row range, shared by the kernels and the host code
// that queues them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include "../../include/galfast_hip.h"

// Module parameter blocks of a chain (device copy), with the index of each
// random-number module's position plane (-1 when the module is off).
struct gf_chain_mods {
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
	int pl_feh, pl_gfeh, pl_bond, pl_perr;
	int nplanes;
	int bond_one_disk;   // Bond2010's two disk moment sets are the same bits
};

// Assign the position planes of the enabled modules, and note whether
// Bond2010's disk_1 and disk_2 are identical (then the per-star choice
// between them, drawn as always, selects the same numbers, and the kernel
// may read disk_1 as a value common to every lane).
inline void gf_chain_planes(gf_chain_mods &m)
{
	int n = 0;
	m.pl_feh = m.has_feh ? n++ : -1;
	m.pl_gfeh = m.has_gfeh ? n++ : -1;
	m.pl_bond = m.has_bond ? n++ : -1;
	m.pl_perr = m.has_perr ? n++ : -1;
	m.nplanes = n;
	m.bond_one_disk = m.has_bond && memcmp(&m.bond.disk_1, &m.bond.disk_2, sizeof(gf_vmoments)) == 0;
}

// Stream positions: plane p holds the (c, x) of row i at pos[p * pitch + i].
struct gf_positions {
	uint2 *pos;
	uint64_t pitch;
};

// One chain over rows [begin, end) in step-row chunks: the module
// parameters, the table and the position planes, passed to every launch by
// value.
struct gf_chain_launch {
	gf_chain_mods m;
	gf_table t;
	gf_positions P;
	uint32_t begin, step, end;
	unsigned long long *vis;   // when set, the draw pass adds the call's visible rows to *vis
};

// The row pass's launch groups, in pipeline order: metallicity (FeH,
// GaussianFeH), kinematics (Bond2010), photometry, photometric errors,
// astrometry (vel2pm, gal2other).  Photometry needs the metallicity,
// photometric errors the photometry, astrometry the kinematics.
enum gf_row_group {
	GF_ROWS_FEH = 0,
	GF_ROWS_BOND = 1,
	GF_ROWS_PHOT = 2,
	GF_ROWS_PERR = 3,
	GF_ROWS_ASTRO = 4,
	GF_ROWS_GROUPS = 5
};
