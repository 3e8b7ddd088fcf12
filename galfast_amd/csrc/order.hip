// Reference row order of a drawn batch and the footprint clip.
//
// skygenHost<T>::drawSources (skygen/skyconfig_impl.h:467-515) sorts every
// drawn batch with star_comp (:34-68) -- hidden first, then l, b, DM, M, Am
// -- and truncates the table at the first hidden row, so the module stages
// see only the visible rows, in that order, and their stream of row r is
// r / 256.  gf_skygen_order does the same on the device: it compacts the
// visible rows (their emission order kept), sorts their indices by
// (l, b, DM, M, Am) with five stable radix sorts from the least significant
// key up, and gathers the skygen columns into the output table.  Rows that
// compare equal on every key keep their emission order (std::sort leaves
// that order unspecified).  gf_clip_footprint is the clipper stage
// (os_clipper::process, skygen/os_skygen.cpp:997-1044) for the analytic
// footprints (rect, beam): a visible row outside the footprint becomes
// hidden, judged from its projection-plane position exactly as the draw
// judges it.
#include "gf_runtime.h"
#include "gf_device.h"
#include <rocprim/device/device_radix_sort.hpp>

namespace {

constexpr int OB = 256;

__global__ void __launch_bounds__(OB) k_visible(gf_col hidden, uint64_t n, uint32_t *flag)
{
	uint64_t i = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (i < n) flag[i] = col<int>(hidden, (uint32_t)i) == 0 ? 1u : 0u;
}

__global__ void __launch_bounds__(OB) k_compact(const uint32_t *flag, const uint64_t *pos, uint64_t n, uint32_t *idx)
{
	uint64_t i = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (i < n && flag[i]) idx[pos[i]] = (uint32_t)i;
}

// Order-preserving unsigned images of floats and doubles, with -0 taken as
// +0 (star_comp compares with <, under which they are equal).
GF_D uint32_t ukey(float f)
{
	uint32_t u = gfm::as_u32(f == 0.f ? 0.f : f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

GF_D uint64_t ukey(double d)
{
	uint64_t u = gfm::as_u64(d == 0.0 ? 0.0 : d);
	return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

template <typename T, typename K>
__global__ void __launch_bounds__(OB) k_keys(gf_col c, uint32_t plane, const uint32_t *idx, uint64_t n, K *key)
{
	uint64_t j = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (j < n) key[j] = ukey(col<T>(c, idx[j], plane));
}

template <typename T>
__global__ void __launch_bounds__(OB) k_gather(gf_col src, gf_col dst, const uint32_t *idx, uint64_t n)
{
	uint64_t j = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (j >= n) return;
	const uint32_t i = idx[j];
	for (uint32_t p = 0; p < src.width; p++) col<T>(dst, (uint32_t)j, p) = col<T>(src, i, p);
}

__global__ void __launch_bounds__(OB) k_clip(gf_skygen_params sp, gf_footprint foot, gf_col projIdx, gf_col projXY,
                                             gf_col hidden, uint64_t n)
{
	uint64_t i = (uint64_t)blockIdx.x * OB + threadIdx.x;
	if (i >= n) return;
	const uint32_t r = (uint32_t)i;
	if (col<int>(hidden, r)) return;
	double l, b;
	lambert_deproject(sp.proj[col<int>(projIdx, r)], l, b, col<float>(projXY, r, 0), col<float>(projXY, r, 1));
	if (outside_footprint(foot, l, b)) col<int>(hidden, r) = 1;
}

// One thread per pixel-grid cell: the number of its 8 x 8 sample points
// inside the footprint (os_skygen.cpp:940-995 clips polygons instead).
__global__ void __launch_bounds__(OB) k_footprint_cover(gf_footprint foot, double dx, int n, uint8_t *cover)
{
	const uint64_t t = (uint64_t)blockIdx.x * OB + threadIdx.x;
	const uint64_t nn = (uint64_t)n * n;
	if (t >= 2 * nn) return;
	const int hemi = (int)(t / nn);
	const int i = (int)((t % nn) / n), j = (int)(t % n);
	const double x0 = -n * dx / 2, xa = x0 + i * dx, ya = x0 + j * dx;
	const double fx = (xa > 0 || xa + dx < 0) ? fmin(fabs(xa), fabs(xa + dx)) : 0.;
	const double fy = (ya > 0 || ya + dx < 0) ? fmin(fabs(ya), fabs(ya + dx)) : 0.;
	int inside = 0;
	if (fx * fx + fy * fy <= 2.0) {
		for (int a = 0; a < 8; a++)
			for (int c = 0; c < 8; c++)
				inside += sample_inside(foot, hemi, xa + dx * (a + 0.5) / 8, ya + dx * (c + 0.5) / 8) ? 1 : 0;
	}
	cover[t] = (uint8_t)inside;
}

int gather(const gf_col &src, const gf_col &dst, const uint32_t *idx, uint64_t n, hipStream_t st)
{
	GF_CHECK(src.width == dst.width && src.elem == dst.elem, GF_E_ARG);
	const unsigned g = gf::blocks_for(n, OB);
	if (src.elem == 8) k_gather<uint64_t><<<g, OB, 0, st>>>(src, dst, idx, n);
	else k_gather<uint32_t><<<g, OB, 0, st>>>(src, dst, idx, n);
	GF_LAUNCHED();
	return 0;
}

// One stable sort of (key, index) pairs; the sorted indices end in idx.
template <typename K>
int sort_pass(K *k0, K *k1, uint32_t *&idx, uint32_t *&idx_alt, uint64_t n, hipStream_t st)
{
	size_t tb = 0;
	GF_HIP(rocprim::radix_sort_pairs(nullptr, tb, k0, k1, idx, idx_alt, (unsigned)n, 0, 8 * (int)sizeof(K), st));
	void *tmp;
	int rc = gf::scratch(GF_SLOT_ORDER_TMP, tb, &tmp);
	if (rc) return rc;
	GF_HIP(rocprim::radix_sort_pairs(tmp, tb, k0, k1, idx, idx_alt, (unsigned)n, 0, 8 * (int)sizeof(K), st));
	uint32_t *t = idx;
	idx = idx_alt;
	idx_alt = t;
	return 0;
}

} // namespace

extern "C" {

int gf_skygen_order(const gf_skygen_cols *in, uint64_t n, const gf_skygen_cols *out, uint64_t *nvisible, void *stream)
{
	GF_CHECK(in && out && nvisible && n < (1ull << 32), GF_E_ARG);
	*nvisible = 0;
	if (n == 0) return 0;
	hipStream_t st = gf::as_stream(stream);
	// scratch: flags (4 B), positions (8 B), two index arrays (4 B each) and
	// two key arrays (8 B each) per row
	void *p;
	int rc = gf::scratch(GF_SLOT_ORDER, (size_t)n * 36 + 64, &p);
	if (rc) return rc;
	uint32_t *flag = (uint32_t *)p;
	uint64_t *pos = (uint64_t *)(((uintptr_t)(flag + n) + 7) & ~(uintptr_t)7);
	uint32_t *idx = (uint32_t *)(pos + n + 1);
	uint32_t *idx_alt = idx + n;
	uint64_t *key0 = (uint64_t *)(((uintptr_t)(idx_alt + n) + 7) & ~(uintptr_t)7);
	uint64_t *key1 = key0 + n;
	const unsigned g = gf::blocks_for(n, OB);
	k_visible<<<g, OB, 0, st>>>(in->hidden, n, flag);
	GF_LAUNCHED();
	rc = gf::exclusive_scan_u32_u64(flag, pos, n, st);
	if (rc) return rc;
	uint64_t nv = 0;
	GF_HIP(hipMemcpyAsync(&nv, pos + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
	GF_HIP(hipStreamSynchronize(st));
	k_compact<<<g, OB, 0, st>>>(flag, pos, n, idx);
	GF_LAUNCHED();
	if (nv) {
		const unsigned gv = gf::blocks_for(nv, OB);
		uint32_t *k32a = (uint32_t *)key0, *k32b = (uint32_t *)key1;
		// least significant key first: Am, M, DM, b, l
		const gf_col *fk[3] = {&in->Am, &in->M, &in->DM};
		for (const gf_col *c : fk) {
			k_keys<float, uint32_t><<<gv, OB, 0, st>>>(*c, 0, idx, nv, k32a);
			GF_LAUNCHED();
			if ((rc = sort_pass(k32a, k32b, idx, idx_alt, nv, st))) return rc;
		}
		for (uint32_t plane : {1u, 0u}) {
			k_keys<double, uint64_t><<<gv, OB, 0, st>>>(in->lb, plane, idx, nv, key0);
			GF_LAUNCHED();
			if ((rc = sort_pass(key0, key1, idx, idx_alt, nv, st))) return rc;
		}
		const gf_col *src[] = {&in->lb, &in->projIdx, &in->projXY, &in->DM, &in->M,
		                       &in->XYZ, &in->Am, &in->AmInf, &in->comp, &in->hidden};
		const gf_col *dst[] = {&out->lb, &out->projIdx, &out->projXY, &out->DM, &out->M,
		                       &out->XYZ, &out->Am, &out->AmInf, &out->comp, &out->hidden};
		for (int k = 0; k < 10; k++)
			if ((rc = gather(*src[k], *dst[k], idx, nv, st))) return rc;
	}
	*nvisible = nv;
	return 0;
}

int gf_clip_footprint(const gf_skygen_params *sp, const gf_footprint *foot, gf_col projIdx, gf_col projXY,
                      gf_col hidden, uint64_t n, void *stream)
{
	GF_CHECK(sp && foot && n < (1ull << 32), GF_E_ARG);
	GF_CHECK(foot->nregion >= 0 && foot->nregion <= GF_FOOT_MAX, GF_E_ARG);
	if (n == 0 || foot->nregion == 0) return 0;
	k_clip<<<gf::blocks_for(n, OB), OB, 0, gf::as_stream(stream)>>>(*sp, *foot, projIdx, projXY, hidden, n);
	GF_LAUNCHED();
	return 0;
}

int gf_footprint_cover(const gf_footprint *foot, double dx, int32_t n, uint8_t *cover, void *stream)
{
	GF_CHECK(foot && cover && dx > 0. && n > 0 && n <= 65536, GF_E_ARG);
	GF_CHECK(foot->nregion >= 0 && foot->nregion <= GF_FOOT_MAX, GF_E_ARG);
	const uint64_t cells = 2ull * (uint64_t)n * (uint64_t)n;
	k_footprint_cover<<<gf::blocks_for(cells, OB), OB, 0, gf::as_stream(stream)>>>(*foot, dx, n, cover);
	GF_LAUNCHED();
	return 0;
}

} // extern "C"
