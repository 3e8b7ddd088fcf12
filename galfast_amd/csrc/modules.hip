// Global kernels and C-ABI entry points of the per-star modules.  Thread t
// of a launch owns rows [begin + step*t, min(begin + step*(t+1), end)) and
// RNG stream t, as KERNEL/kernel_state do in the reference (common/gpu.h:124-156,
// 240-257); the workgroups are 256 lanes.
#include "gf_runtime.h"
#include "gf_modules.h"

namespace {

constexpr unsigned MOD_BS = 256;

struct RowMap {
	uint32_t begin, step, end;
	GF_D bool range(uint32_t t, uint32_t &r0, uint32_t &r1) const
	{
		uint64_t a = (uint64_t)begin + (uint64_t)step * t;
		if (a >= end) return false;
		uint64_t b = a + step;
		r0 = (uint32_t)a;
		r1 = (uint32_t)(b < end ? b : end);
		return true;
	}
};

inline uint32_t nthreads_of(const gf_rows *ks)
{
	uint32_t n = ks->end - ks->begin;
	return n / ks->step + (n % ks->step ? 1 : 0);
}

int check_rows(const gf_rows *ks)
{
	GF_CHECK(ks && ks->step > 0 && ks->end >= ks->begin, GF_E_ARG);
	return 0;
}

__global__ void __launch_bounds__(MOD_BS) k_FeH(RowMap rm, uint32_t nthr, gf_FeH_data par, uint32_t *st, uint32_t ns,
                                                gf_col comp, gf_col hidden, gf_col XYZ, gf_col FeH)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	gf_mwc r = gf_mwc_load(st, ns, t);
	mod_FeH(par, r, r0, r1, comp, hidden, XYZ, FeH);
	gf_mwc_store(st, ns, t, r);
}

__global__ void __launch_bounds__(MOD_BS) k_GaussianFeH(RowMap rm, uint32_t nthr, gf_GaussianFeH_data par, gf_bitmap apply,
                                                        uint32_t *st, uint32_t ns, gf_col comp, gf_col hidden,
                                                        gf_col FeH_comp, gf_col FeH)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	gf_mwc r = gf_mwc_load(st, ns, t);
	mod_GaussianFeH(par, apply, r, r0, r1, comp, hidden, FeH_comp, FeH);
	gf_mwc_store(st, ns, t, r);
}

__global__ void __launch_bounds__(MOD_BS) k_Bond2010(RowMap rm, uint32_t nthr, const gf_Bond2010_data *par, uint32_t *st,
                                                     uint32_t ns, gf_col comp, gf_col hidden, gf_col XYZ, gf_col vcyl)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	gf_mwc r = gf_mwc_load(st, ns, t);
	mod_Bond2010(*par, r, r0, r1, comp, hidden, XYZ, vcyl);
	gf_mwc_store(st, ns, t, r);
}

__global__ void __launch_bounds__(MOD_BS) k_vel2pm(RowMap rm, uint32_t nthr, gf_vel2pm_data par, gf_col lb, gf_col XYZ,
                                                   gf_col vcyl, gf_col pm, gf_col hidden)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	mod_vel2pm(par, r0, r1, lb, XYZ, vcyl, pm, hidden);
}

__global__ void __launch_bounds__(MOD_BS) k_gal2other(RowMap rm, uint32_t nthr, int coordsys, gf_col hidden, gf_col lb, gf_col out)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	mod_gal2other(coordsys, r0, r1, hidden, lb, out);
}

__global__ void __launch_bounds__(MOD_BS) k_photometry(RowMap rm, uint32_t nthr, const gf_photometry_data *lt, gf_bitmap apply,
                                                       gf_col Am, gf_col flags, gf_col DM, gf_col Mr, int nabsmag,
                                                       gf_col mags, gf_col FeH, gf_col comp, gf_col hidden)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	mod_photometry(*lt, apply, r0, r1, Am, flags, DM, Mr, nabsmag, mags, FeH, comp, hidden);
}

__global__ void __launch_bounds__(MOD_BS) k_multiples(RowMap rm, uint32_t nthr, gf_multiples_data par, gf_bitmap apply,
                                                      uint32_t *st, uint32_t ns, int nabsmag, gf_col M, gf_col Msys,
                                                      gf_col ncomp, gf_col comp, gf_col hidden)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	gf_mwc r = gf_mwc_load(st, ns, t);
	mod_multiples(par, apply, r, r0, r1, nabsmag, M, Msys, ncomp, comp, hidden);
	gf_mwc_store(st, ns, t, r);
}

__global__ void __launch_bounds__(MOD_BS) k_photoerr(RowMap rm, uint32_t nthr, gf_photoerr_data par, uint32_t *st, uint32_t ns,
                                                     gf_col magTrue, gf_col magObs, gf_col hidden)
{
	uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, r0, r1;
	if (t >= nthr || !rm.range(t, r0, r1)) return;
	gf_mwc r = gf_mwc_load(st, ns, t);
	mod_photoerr(par, r, r0, r1, magTrue, magObs, hidden);
	gf_mwc_store(st, ns, t, r);
}

// Parameter blocks too large for a comfortable kernel argument are staged
// in a device copy (the reference uploads them as __constant__).
template <typename T>
int stage_param(int slot, const T &host, const T **dev, hipStream_t st)
{
	void *p;
	int rc = gf::scratch(slot, sizeof(T), &p);
	if (rc) return rc;
	GF_HIP(hipMemcpyAsync(p, &host, sizeof(T), hipMemcpyHostToDevice, st));
	*dev = (const T *)p;
	return 0;
}

} // namespace

#define GF_ROWS_PROLOGUE(needs_rng)                                            \
	int rc__ = check_rows(ks);                                                 \
	if (rc__) return rc__;                                                     \
	uint32_t nthr = nthreads_of(ks);                                           \
	if (nthr == 0) return 0;                                                   \
	if (needs_rng) {                                                           \
		GF_CHECK(rng && rng->state, GF_E_ARG);                                 \
		GF_CHECK(nthr <= rng->nstreams, GF_E_STREAMS);                         \
	}                                                                          \
	RowMap rm{ks->begin, ks->step, ks->end};                                   \
	hipStream_t st = gf::as_stream(stream);                                    \
	unsigned grid = gf::blocks_for(nthr, MOD_BS);

extern "C" {

int gf_os_FeH_kernel(const gf_rows *ks, const gf_FeH_data *par, gf_rng *rng,
                     gf_col comp, gf_col hidden, gf_col XYZ, gf_col FeH, void *stream)
{
	GF_CHECK(par, GF_E_ARG);
	GF_ROWS_PROLOGUE(true)
	k_FeH<<<grid, MOD_BS, 0, st>>>(rm, nthr, *par, rng->state, rng->nstreams, comp, hidden, XYZ, FeH);
	GF_LAUNCHED();
	return 0;
}

int gf_os_GaussianFeH_kernel(const gf_rows *ks, gf_bitmap applyToComponents,
                             const gf_GaussianFeH_data *par, gf_rng *rng,
                             gf_col comp, gf_col hidden, gf_col XYZ,
                             gf_col FeH_comp, gf_col FeH, void *stream)
{
	(void)XYZ;
	GF_CHECK(par && par->ncomp >= 1 && par->ncomp <= 10, GF_E_ARG);
	GF_ROWS_PROLOGUE(true)
	k_GaussianFeH<<<grid, MOD_BS, 0, st>>>(rm, nthr, *par, applyToComponents, rng->state, rng->nstreams,
	                                       comp, hidden, FeH_comp, FeH);
	GF_LAUNCHED();
	return 0;
}

int gf_os_Bond2010_kernel(const gf_rows *ks, const gf_Bond2010_data *par, gf_rng *rng,
                          gf_col comp, gf_col hidden, gf_col XYZ, gf_col vcyl, void *stream)
{
	GF_CHECK(par, GF_E_ARG);
	GF_ROWS_PROLOGUE(true)
	const gf_Bond2010_data *dpar;
	int rc = stage_param(GF_SLOT_NSLOTS + 0, *par, &dpar, st);
	if (rc) return rc;
	k_Bond2010<<<grid, MOD_BS, 0, st>>>(rm, nthr, dpar, rng->state, rng->nstreams, comp, hidden, XYZ, vcyl);
	GF_LAUNCHED();
	return 0;
}

int gf_os_vel2pm_kernel(const gf_rows *ks, const gf_vel2pm_data *par,
                        gf_col lb, gf_col XYZ, gf_col vcyl, gf_col pmout,
                        gf_col hidden, void *stream)
{
	GF_CHECK(par, GF_E_ARG);
	gf_rng *rng = nullptr;
	(void)rng;
	GF_ROWS_PROLOGUE(false)
	k_vel2pm<<<grid, MOD_BS, 0, st>>>(rm, nthr, *par, lb, XYZ, vcyl, pmout, hidden);
	GF_LAUNCHED();
	return 0;
}

int gf_os_gal2other_kernel(const gf_rows *ks, int coordsys, gf_col hidden,
                           gf_col lb, gf_col out, void *stream)
{
	gf_rng *rng = nullptr;
	(void)rng;
	GF_ROWS_PROLOGUE(false)
	k_gal2other<<<grid, MOD_BS, 0, st>>>(rm, nthr, coordsys, hidden, lb, out);
	GF_LAUNCHED();
	return 0;
}

int gf_os_photometry_kernel(const gf_rows *ks, gf_bitmap applyToComponents,
                            const gf_photometry_data *par,
                            gf_col Am, gf_col flags, gf_col DM, gf_col Mr,
                            int nabsmag, gf_col mags, gf_col FeH, gf_col comp,
                            gf_col hidden, void *stream)
{
	GF_CHECK(par && par->ncolors >= 1 && par->ncolors <= 16 && par->bidx <= par->ncolors, GF_E_ARG);
	gf_rng *rng = nullptr;
	(void)rng;
	GF_ROWS_PROLOGUE(false)
	const gf_photometry_data *dpar;
	int rc = stage_param(GF_SLOT_NSLOTS + 1, *par, &dpar, st);
	if (rc) return rc;
	k_photometry<<<grid, MOD_BS, 0, st>>>(rm, nthr, dpar, applyToComponents, Am, flags, DM, Mr, nabsmag,
	                                      mags, FeH, comp, hidden);
	GF_LAUNCHED();
	return 0;
}

int gf_os_unresolvedMultiples_kernel(const gf_rows *ks, gf_bitmap applyToComponents,
                                     const gf_multiples_data *par, gf_rng *rng,
                                     int nabsmag, gf_col M, gf_col Msys,
                                     gf_col ncomp, gf_col comp, gf_col hidden,
                                     void *stream)
{
	GF_CHECK(par && par->algo >= 1 && par->algo <= 3, GF_E_ARG);
	GF_ROWS_PROLOGUE(true)
	k_multiples<<<grid, MOD_BS, 0, st>>>(rm, nthr, *par, applyToComponents, rng->state, rng->nstreams,
	                                     nabsmag, M, Msys, ncomp, comp, hidden);
	GF_LAUNCHED();
	return 0;
}

int gf_photometric_errors_kernel(const gf_rows *ks, const gf_photoerr_data *par,
                                 gf_rng *rng, gf_col magTrue, gf_col magObs,
                                 gf_col hidden, void *stream)
{
	GF_CHECK(par && par->nbands >= 0 && par->nbands <= 17, GF_E_ARG);
	GF_ROWS_PROLOGUE(true)
	k_photoerr<<<grid, MOD_BS, 0, st>>>(rm, nthr, *par, rng->state, rng->nstreams, magTrue, magObs, hidden);
	GF_LAUNCHED();
	return 0;
}

int gf_photoerr_records(int32_t nbands, const int32_t *koff, const double *kx, const double *ky, double *kq,
                        double *gx0, double *gscale)
{
	GF_CHECK(nbands >= 0 && nbands <= 17 && koff && gx0 && gscale, GF_E_ARG);
	GF_CHECK(koff[0] == 0 && (koff[nbands] == 0 || (kx && ky && kq)), GF_E_ARG);
	for (int b = 0; b < nbands; b++) {
		const int k0 = koff[b], k1 = koff[b + 1];
		GF_CHECK(k1 >= k0, GF_E_ARG);
		gx0[b] = k1 > k0 ? kx[k0] : 0.0;
		gscale[b] = k1 - k0 > 1 ? (double)(k1 - k0 - 1) / (kx[k1 - 1] - kx[k0]) : 0.0;
		for (int j = k0; j < k1; j++) {
			double *r = kq + 4 * (size_t)j;
			const bool last = j + 1 == k1;
			r[0] = kx[j];
			r[1] = ky[j];
			r[2] = last ? 0.0 : (ky[j + 1] - ky[j]) / (kx[j + 1] - kx[j]);
			r[3] = last ? kx[j] : kx[j + 1];
		}
	}
	return 0;
}

} // extern "C"
