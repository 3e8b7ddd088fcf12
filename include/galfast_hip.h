/*
 * galfast_hip.h -- C ABI of the MI355X (gfx950) catalog hot path.
 *
 * Every entry point takes plain pointers, sizes and POD parameter blocks.
 * Device pointers are HIP device addresses; `stream` is a hipStream_t (NULL
 * selects the null stream).  Functions return 0 on success, a positive
 * hipError_t value when a HIP call failed, or a negative GF_E* code for an
 * argument error.  The parameter blocks keep the field order and types of
 * the reference's own structures so that a reference-side module fills them
 * exactly as it fills its kernel arguments today.
 *
 * Reference paths are relative to the galfast source tree.
 */
#ifndef GALFAST_HIP_H
#define GALFAST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Error codes                                                             */
/* ---------------------------------------------------------------------- */
#define GF_OK              0
#define GF_E_ARG         (-1)   /* invalid argument (NULL pointer, bad size) */
#define GF_E_CAPACITY    (-2)   /* output table too small */
#define GF_E_STREAMS     (-3)   /* more 256-row chunks than RNG streams */
#define GF_E_CONFIG      (-4)   /* configuration file missing a key or malformed */
#define GF_E_IO          (-5)   /* file could not be read */
#define GF_E_NODEVICE    (-6)   /* no HIP device */
#define GF_E_BUDGET      (-7)   /* request exceeds the device-memory budget */
#define GF_E_STATE       (-8)   /* call out of order */
#define GF_E_SCRATCH     (-9)   /* input exceeds a no-allocation entry point's fixed scratch */

/* Human-readable text for an error code (static storage). */
const char *gf_strerror(int code);

/* ---------------------------------------------------------------------- */
/* Data model                                                              */
/* ---------------------------------------------------------------------- */

/* Structure-of-arrays column view, the device side of a column<T>.
 * Element (row r, plane j) lives at (char*)ptr + j*pitch + r*elem.
 * Replaces gptr<T,2> = { T *ptr; uint32_t extent[1]; }
 * (common/cux.h:183-216, common/column.h:131-136).  pitch is a multiple
 * of 128 bytes for multi-plane columns (common/otable.h:275-288). */
typedef struct gf_col {
    void    *ptr;
    uint32_t pitch;   /* bytes between planes */
    uint32_t width;   /* number of planes */
    uint32_t elem;    /* bytes per element (4 or 8) */
    uint32_t reserved;
} gf_col;

/* Row range of one kernel call: replaces kernel_state / otable_ks
 * (common/gpu.h:124-156).  Thread t of the reference owns rows
 * [begin + step*t, min(begin + step*(t+1), end)) and RNG stream t. */
typedef struct gf_rows {
    uint32_t begin;
    uint32_t step;    /* rows per stream; 256 in the reference */
    uint32_t end;
    uint32_t shmem;   /* unused, kept for layout parity */
} gf_rows;

/* Component mask: bit_map (common/column.h:51-113). */
typedef struct gf_bitmap {
    uint32_t bits[2];
} gf_bitmap;

/* Device-resident multiply-with-carry streams: three planes [a | c | x] of
 * nstreams words each (common/cuda_rng.h:45-69, 414-500). */
typedef struct gf_rng {
    uint32_t *state;     /* device pointer, 3*nstreams words */
    uint32_t  nstreams;
    uint32_t  reserved;
} gf_rng;

/* ---------------------------------------------------------------------- */
/* Runtime                                                                 */
/* ---------------------------------------------------------------------- */

/* Select the HIP device used by later calls (cux_init, common/cux.cpp:713). */
int gf_init(int device);
/* Wait for all work on `stream` (cudaThreadSynchronize in gpu.h:253). */
int gf_sync(void *stream);
/* Device memory helpers for callers without their own allocator. */
int gf_malloc(void **ptr, size_t bytes);
int gf_free(void *ptr);
int gf_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int gf_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int gf_memset(void *dst, int value, size_t bytes, void *stream);

/* ---------------------------------------------------------------------- */
/* Random streams                                                          */
/* ---------------------------------------------------------------------- */

/* Seed derivation of gpu_rng_t::persistent_rng::get (common/cux.cpp:404-410)
 * with rng_gsl_t (common/gpu.h:362-382): the first MT19937 output of
 * `pipeline_seed`, divided by 2^32, rounded to float, times 2^24, truncated. */
uint32_t gf_seed_from_pipeline_seed(uint32_t pipeline_seed);

/* mwc_impl::srand (common/cuda_rng.h:466-497), on the host.
 * primes: multipliers (first column of safeprimes32.txt), at least nstreams.
 * nstreams is rounded up to a multiple of 64.  out_state receives
 * 3*nstreams words [a | c | x].  Returns the rounded stream count, or a
 * negative error code. */
int gf_mwc_seed_host(uint32_t seed, const uint32_t *primes, uint32_t nprimes,
                     uint32_t nstreams, uint32_t *out_state);

/* mwc_impl::srand (common/cuda_rng.h:466-497) on the device: the same
 * tables as gf_mwc_seed_host, each stream's two drand48/lrand48 draws
 * reached by jump-ahead, so that every stream is seeded by its own lane.
 * primes and state are device memory; nstreams must be a multiple of 64
 * and at most nprimes; state receives 3*nstreams words [a | c | x]. */
int gf_mwc_seed_device(uint32_t seed, const uint32_t *primes, uint32_t nprimes,
                       uint32_t nstreams, uint32_t *state, void *stream);

/* Allocate / copy device stream state (rng_base::upload/download,
 * common/cuda_rng.h:74-103). */
int gf_rng_alloc(gf_rng *rng, uint32_t nstreams);
int gf_rng_upload(gf_rng *rng, const uint32_t *host_state, void *stream);
int gf_rng_download(const gf_rng *rng, uint32_t *host_state, void *stream);
int gf_rng_free(gf_rng *rng);

/* Move every stream n steps ahead in O(log n) (multiply-with-carry
 * jump-ahead: with base 2^32 the state c*2^32 + x is multiplied by a^n
 * modulo a*2^32 - 1).  Same states as n calls of uniform() per stream
 * (common/cuda_rng.h:439-463); no counterpart in the reference, which only
 * steps.  gf_mwc_jump_host does the same on a host stream table. */
int gf_mwc_jump(gf_rng *rng, uint64_t n, void *stream);
int gf_mwc_jump_host(uint32_t *state, uint32_t nstreams, uint64_t n);

/* Sampler test kernel: stream s produces n values of the chosen sampler,
 * written to out[i*nstreams + s] (float for uniform/gaussian, uint32 for
 * poisson).  kind: 0 uniform, 1 gaussian(param), 2 poisson(param),
 * 3 poisson(param * (1 + s)) -- a sweep of means across streams. */
int gf_mwc_sample(gf_rng *rng, int kind, float param, uint32_t n, void *out, void *stream);

/* ---------------------------------------------------------------------- */
/* Per-star module kernels (the GPU twin of cpulaunch_/gpulaunch_ pairs)   */
/* ---------------------------------------------------------------------- */

/* os_FeH_data (modules/FeH_gpu.cu.h:27-36). */
typedef struct gf_FeH_data {
    float A[2], sigma[3], offs[3];
    float Hmu, muInf, DeltaMu;
    gf_bitmap comp_thin, comp_thick, comp_halo;
    float M[3][3];
    float T[3];
} gf_FeH_data;

/* os_GaussianFeH_data (modules/GaussianFeH_gpu.cu.h:27-36). */
typedef struct gf_GaussianFeH_data {
    int   ncomp;
    float mean[10];
    float sigma[10];
    float pcum[10];
} gf_GaussianFeH_data;

/* Bond2010::vmoments and os_Bond2010_data (modules/Bond2010_gpu.cu.h:29-52):
 * each quantity is a + b|Z|^c + d R^e, stored as 5 floats. */
typedef struct gf_vmoments {
    float m[3][5];
    float ss[6][5];
} gf_vmoments;

typedef struct gf_Bond2010_data {
    gf_bitmap comp_thin, comp_thick, comp_halo;
    float fk;
    gf_vmoments disk_1, disk_2;
    gf_vmoments halo;
    float M[3][3];
    float T[3];
} gf_Bond2010_data;

/* os_vel2pm_data (modules/vel2pm_gpu.cu.h:27-35). coordsys: 0 GAL, 1 EQU. */
typedef struct gf_vel2pm_data {
    int   coordsys;
    float vLSR, u0, v0, w0;
    float Rg;
} gf_vel2pm_data;

/* os_photometry_data (modules/photometry.h:27-33) plus the colour tables the
 * reference binds as textures color0..3 / cflags0..3.  Table i holds colours
 * 4i..4i+3 as float4 texels at [m*nFeH + f]; FeH of texel f is
 * (f - 0.5 + 0.5)/tcFeH[1] + tcFeH[0] (texcoord_from_range convention). */
typedef struct gf_photometry_data {
    uint32_t ncolors, bidx;
    float    reddening[17];
    int32_t  nFeH, nMr;
    float    tcFeH[2], tcMr[2];
    const float *color[4];   /* device, float4 texels */
    const float *cflags[4];  /* device, float4 texels */
    /* optional (device): the colour texels by lookup quad, so that a lookup
     * is one contiguous fetch.  Quad k = j*nFeH + i holds the four texels a
     * lookup with lower corner (i, j) interpolates, (i, j), (i1, j),
     * (i, j1), (i1, j1) with i1 = min(i+1, nFeH-1), j1 = min(j+1, nMr-1):
     * for each of the ntex colour tables t the four corners' float4s at
     * texels + 16*(ntex*k + t); flagbits[4k + corner] holds that corner's
     * flags as bits (bit 4t+c set when colour 4t+c's flag texel is
     * nonzero).  Used instead of color/cflags when both are set. */
    const float *texels;
    const uint16_t *flagbits;
} gf_photometry_data;

/* Tabulated 1D function on a uniform grid, sampled with the reference's
 * texture convention xi = (x - tc[0])*tc[1] + 0.5 and linear filtering with
 * clamping (common/cux.h:849-855). */
typedef struct gf_tex1d {
    const float *data;   /* device */
    int32_t n;
    float   tc[2];
} gf_tex1d;

/* unresolvedMultiples tables (modules/unresolvedMultiples_host.cpp:140-180).
 * algo: 1 LF_M2_gt_M1, 2 LF, 3 equal_mass. */
typedef struct gf_multiples_data {
    gf_tex1d secProb, cumLF, invCumLF;
    int32_t algo;
} gf_multiples_data;

/* Photometric-error curves sigma(m) as linear splines over knots
 * (modules/photometricErrors_host.cpp:112-131).  Band b uses knots
 * [koff[b], koff[b+1]); nbands <= 17; a band with no knots gets no error. */
typedef struct gf_photoerr_data {
    int32_t nbands;
    int32_t koff[18];
    const double *kx;   /* device knots: magnitude */
    const double *ky;   /* device knots: sigma */
    /* Optional (NULL: evaluate from kx, ky): device records of 4 doubles, one
     * per knot j of band b (koff[b] <= j < koff[b+1]): {kx[j], ky[j],
     * (ky[j+1]-ky[j]) / (kx[j+1]-kx[j]), kx[j+1]}, the band's last knot
     * {kx[j], ky[j], 0, kx[j]}; and per band the first knot gx0[b] and
     * gscale[b] = (n-1) / (kx[last]-kx[first]) (n knots, 0 when n = 1), from
     * which the interval of a magnitude is guessed.  The slopes are the
     * quotients the evaluation would form, so the results are the same;
     * gf_photoerr_records fills all three. */
    const double *kq;
    double gx0[18];
    double gscale[18];
} gf_photoerr_data;

/* Fill the interval records `kq` (4 * koff[nbands] doubles, host memory),
 * gx0 and gscale (nbands doubles each) of gf_photoerr_data from host copies
 * of its knots. */
int gf_photoerr_records(int32_t nbands, const int32_t *koff, const double *kx, const double *ky, double *kq,
                        double *gx0, double *gscale);

/* os_FeH_kernel (modules/FeH_gpu.cu.h:43-87; called at FeH_host.cpp:76). */
int gf_os_FeH_kernel(const gf_rows *ks, const gf_FeH_data *par, gf_rng *rng,
                     gf_col comp, gf_col hidden, gf_col XYZ, gf_col FeH, void *stream);

/* os_GaussianFeH_kernel (modules/GaussianFeH_gpu.cu.h:42-73;
 * called at GaussianFeH_host.cpp:74). */
int gf_os_GaussianFeH_kernel(const gf_rows *ks, gf_bitmap applyToComponents,
                             const gf_GaussianFeH_data *par, gf_rng *rng,
                             gf_col comp, gf_col hidden, gf_col XYZ,
                             gf_col FeH_comp, gf_col FeH, void *stream);

/* os_Bond2010_kernel (modules/Bond2010_gpu.cu.h:164-218; called at
 * Bond2010_host.cpp:88; the parameters travel as a __constant__ there). */
int gf_os_Bond2010_kernel(const gf_rows *ks, const gf_Bond2010_data *par, gf_rng *rng,
                          gf_col comp, gf_col hidden, gf_col XYZ, gf_col vcyl, void *stream);

/* os_vel2pm_kernel (modules/vel2pm_gpu.cu.h:253-301; called at
 * vel2pm_host.cpp:72).  Draws no random numbers. */
int gf_os_vel2pm_kernel(const gf_rows *ks, const gf_vel2pm_data *par,
                        gf_col lb, gf_col XYZ, gf_col vcyl, gf_col pmout,
                        gf_col hidden, void *stream);

/* os_gal2other_kernel (modules/gal2other_gpu.cu.h:62-96; called at
 * gal2other_host.cpp:62).  coordsys: 1 EQU. */
int gf_os_gal2other_kernel(const gf_rows *ks, int coordsys, gf_col hidden,
                           gf_col lb, gf_col out, void *stream);

/* os_photometry_kernel (modules/photometry_gpu.cu.h:87-174; called at
 * photometry_host.cpp:391). */
int gf_os_photometry_kernel(const gf_rows *ks, gf_bitmap applyToComponents,
                            const gf_photometry_data *par,
                            gf_col Am, gf_col flags, gf_col DM, gf_col Mr,
                            int nabsmag, gf_col mags, gf_col FeH, gf_col comp,
                            gf_col hidden, void *stream);

/* os_unresolvedMultiples_kernel (modules/unresolvedMultiples_gpu.cu.h:80-123;
 * called at unresolvedMultiples_host.cpp:103). */
int gf_os_unresolvedMultiples_kernel(const gf_rows *ks, gf_bitmap applyToComponents,
                                     const gf_multiples_data *par, gf_rng *rng,
                                     int nabsmag, gf_col M, gf_col Msys,
                                     gf_col ncomp, gf_col comp, gf_col hidden,
                                     void *stream);

/* os_photometricErrors::process (modules/photometricErrors_host.cpp:112-131).
 * The reference runs this on the host with the GSL seeder stream; this
 * kernel draws the Gaussian errors from the per-chunk MWC streams instead
 * (one stream per `ks->step` rows, bands in order within each row chunk). */
int gf_photometric_errors_kernel(const gf_rows *ks, const gf_photoerr_data *par,
                                 gf_rng *rng, gf_col magTrue, gf_col magObs,
                                 gf_col hidden, void *stream);

/* ---------------------------------------------------------------------- */
/* Fused column chain                                                      */
/* ---------------------------------------------------------------------- */

/* One pass over a row range running, for each 256-row chunk, the enabled
 * modules in pipeline order (pipeline.h:75-89): FeH or GaussianFeH,
 * unresolvedMultiples, Bond2010, photometry, photometric errors, vel2pm,
 * gal2other.  Each module consumes its random numbers from the chunk's
 * stream in the same order as the unfused kernels, so the outputs are
 * identical to calling the per-module entry points one after another.
 * Disabled modules have a NULL parameter pointer. */
typedef struct gf_chain {
    const gf_FeH_data         *feh;
    const gf_GaussianFeH_data *gfeh;
    gf_bitmap                  gfeh_comps;
    const gf_multiples_data   *mult;
    gf_bitmap                  mult_comps;
    const gf_Bond2010_data    *bond;
    const gf_photometry_data  *phot;
    gf_bitmap                  phot_comps;
    const gf_photoerr_data    *perr;
    const gf_vel2pm_data      *v2pm;
    int32_t                    gal2other_coordsys;  /* 0 disables */
    int32_t                    pad;
} gf_chain;

/* Column set of the catalog table (names as in otable.cpp:714-756). */
typedef struct gf_table {
    gf_col lb, projIdx, projXY, DM, absmag, XYZ, Am, AmInf, comp, hidden;
    gf_col FeH, FeH_comp, absmagSys, absmagSysNcomp, vcyl, mags, photoFlags,
           obsMags, pm, radec;
} gf_table;

int gf_module_chain(const gf_rows *ks, const gf_chain *chain, gf_rng *rng,
                    const gf_table *t, void *stream);

/* ---------------------------------------------------------------------- */
/* Sky generator (skygen)                                                  */
/* ---------------------------------------------------------------------- */

/* lambert (skygen/skygen.h:108-147): pole (l0, b0) in radians plus the
 * float cosines/sines of the pole. */
typedef struct gf_lambert {
    double l0, b0;
    float  cl, cb, sl, sb;
} gf_lambert;

/* pencilBeam (skygen/skygen.h:89-103): direction cosines of the pixel
 * centre, covered fraction, linear size (rad), area (sr), projection index,
 * Lambert coordinates and the per-beam extinction index. */
typedef struct gf_pencil_beam {
    float   cl, cb, sl, sb;
    float   coveredFraction;
    float   dx, dA;
    int32_t projIdx;
    float   X, Y;
    int32_t extIdx;
    int32_t pad;
} gf_pencil_beam;

/* skygenParams (skygen/skygen.h:152-165). */
typedef struct gf_skygen_params {
    float   m0, m1, dm;
    float   M0, M1, dM;
    float   dmin, dmax;
    int32_t npixels, nm, nM;
    int32_t nthreads;
    int32_t stopstars;
    int32_t pad;
    gf_lambert proj[2];
} gf_skygen_params;

#define GF_MODEL_EXPDISK   1
#define GF_MODEL_POWERLAW  2

/* One density model (skygen/model_expDisk.h:39-124,
 * skygen/model_powerLawEllipsoid.h:30-108) with its luminosity function
 * texture (1024 texels, skygen/model_lib.h:37-95). */
typedef struct gf_model {
    int32_t kind;
    int32_t comp;            /* sequential component index written to `comp` */
    float   f, l, h, r_cut2; /* exponential disk */
    float   n, ba, ca;       /* power-law ellipsoid */
    float   rminSq, rmaxSq;
    float   T[3];
    float   M[3][3];
    float   norm;            /* overall density normalisation (setDensityNorm) */
    gf_tex1d lf;
} gf_model;

/* Extinction: ext_beam minimum-extinction cube and north/south maps
 * (skygen/skygen.cu.h:199-226).  A NULL data pointer means zero extinction. */
typedef struct gf_tex3d {
    const float *data;   /* device, [k][j][i] */
    int32_t nx, ny, nz;
    float   tc[3][2];
} gf_tex3d;

typedef struct gf_extinction {
    gf_tex3d north, south, beam;
} gf_extinction;

/* Footprint predicate used to hide stars drawn outside the requested area
 * (the clipper, skygen/os_skygen.cpp:999-1044) and to pixelize the sky.
 * A footprint is the union of `nregion` regions, one per footprint
 * configuration file (load_footprints, footprint.cpp:746-759); nregion 0
 * means the whole sky.  Each region is what footprint.cpp builds for one
 * file:
 *   kind 1 rect (load_footprint_rect, footprint.cpp:486-560): lon l0..l1,
 *          lat b0..b1 in degrees, in Galactic (equ = 0) or J2000
 *          equatorial (equ = 1) coordinates; l0 < l1, l1 - l0 <= 360;
 *   kind 2 beam (load_footprint_beam, footprint.cpp:675-700): Galactic
 *          direction cosines of the centre, cos(radius), and
 *          hole = 1 - cos(hole radius) (0 = no hole);
 * minus the zone of avoidance |b| < bmin degrees, Galactic
 * (clip_zone_of_avoidance, footprint.cpp:704-727). */
#define GF_FOOT_MAX 8
typedef struct gf_foot_region {
    int32_t kind;
    int32_t equ;
    double  l0, l1, b0, b1;         /* rect, degrees */
    double  cl, cb, sl, sb;         /* beam centre direction */
    double  cosr, hole;             /* beam cos(radius), 1 - cos(hole radius) */
    double  bmin;                   /* |b| < bmin removed, degrees */
} gf_foot_region;

typedef struct gf_footprint {
    int32_t        nregion;
    int32_t        pad;
    gf_foot_region r[GF_FOOT_MAX];
} gf_footprint;

/* Sky pixelization (os_skygen::load_footprints, skygen/os_skygen.cpp:
 * 235-283, 940-995): square pixels of side dx [radians of the Lambert
 * plane] on a grid of n x n cells centred on each pole (north: pole
 * b = +90 deg, l0 = 90 deg; south: b = -90 deg, l0 = 270 deg),
 * n = ceil(2 sqrt(2) / dx).  Cell (h, i, j) spans x in [-n dx/2 + i dx, +dx),
 * y likewise with j; its coverage is the number of 8 x 8 sample points
 * (cell-centred) that lie in its own hemisphere and inside `foot`.
 * Writes cover[(h n + i) n + j] (0..64) for the 2 n^2 cells. */
int gf_footprint_cover(const gf_footprint *foot, double dx, int32_t n, uint8_t *cover, void *stream);

/* Count-pass results (skygen/skyconfig_impl.h:112-184). */
typedef struct gf_count_result {
    double nstarsExpectedToGenerate;   /* sum over pixels at full pixel area */
    double nstarsExpected;             /* covered-fraction weighted */
    double maxCount;                   /* largest single-cell expectation */
    double ncells;                     /* cells inside flux and distance limits */
} gf_count_result;

/* Accumulation variants for the count integral. */
#define GF_ACC_FP64    0
#define GF_ACC_KAHAN32 1
#define GF_ACC_FP32    2

/* Expected star counts for one model, in one pass over all (pixel, DM-bin)
 * pairs.  pixels: device array of npixels beams.  per_pixel (optional,
 * device, npixels doubles) receives covered-fraction weighted counts per
 * pixel.  wpk (optional, device, npixels*(nm+nM-1) doubles) receives the
 * expectation of every (pixel, DM-bin) pair for the draw.
 * Replaces skygenHost<T>::integrateCounts (skygen/skyconfig_impl.h:322-429)
 * and the count branch of skygenGPU<T>::kernel<0> (skygen/skygen.cu.h:469-618). */
int gf_skygen_count(const gf_skygen_params *sp, const gf_model *model,
                    const gf_pencil_beam *pixels, const gf_extinction *ext,
                    int acc_mode, gf_count_result *out, double *per_pixel,
                    double *wpk, void *stream);

/* gf_skygen_count without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as gf_skygen_count except that out_dev is DEVICE memory and the
 * call neither copies to the host nor waits: the totals are in *out_dev
 * once `stream` reaches that point.  Its scratch (LF tables and per-pixel
 * partial sums) is a fixed device array of the library, one per GPU,
 * limited to GF_NOALLOC_MAX_PIXELS pixels and GF_NOALLOC_MAX_NM
 * absolute-magnitude bins (GF_E_SCRATCH beyond).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (different
 * streams, or host threads) race; calls ordered on one stream do not. */
#define GF_NOALLOC_MAX_PIXELS 262144
#define GF_NOALLOC_MAX_NM     4096
int gf_skygen_count_noalloc(const gf_skygen_params *sp, const gf_model *model,
                            const gf_pencil_beam *pixels, const gf_extinction *ext,
                            int acc_mode, gf_count_result *out_dev, double *per_pixel,
                            double *wpk, void *stream);

/* Skygen output columns (skygen/skygen.h:170-177). */
typedef struct gf_skygen_cols {
    gf_col lb, projIdx, projXY, DM, M, XYZ, Am, AmInf, comp, hidden;
} gf_skygen_cols;

/* Reference-order draw: the diagonal crawl of skygenGPU<T>::kernel<1>
 * (skygen/skygen.cu.h:366-457, 469-618) with `sp->nthreads` threads, each
 * bound to RNG stream tid, Poisson draw per cell, 2 or 4 uniforms per star.
 * Stars are written in (thread, crawl) order instead of atomicAdd order.
 * Needs sp->nthreads * 12 bytes of scratch plus the output capacity; the
 * number of stars is returned in *nstars; GF_E_CAPACITY if it exceeds
 * `capacity` (nothing is written then, the streams are left unchanged). */
int gf_skygen_draw_compat(const gf_skygen_params *sp, const gf_model *model,
                          const gf_pencil_beam *pixels, const gf_extinction *ext,
                          const gf_footprint *foot, gf_rng *rng,
                          const gf_skygen_cols *out, uint64_t capacity,
                          uint64_t *nstars, void *stream);

/* Reference row order of a drawn batch (skygenHost<T>::drawSources,
 * skygen/skyconfig_impl.h:467-515, with star_comp :34-68): the visible rows
 * of `in` (hidden == 0) sorted by (l, b, DM, M, Am) and gathered into `out`,
 * whose capacity must hold them; *nvisible receives their number (the
 * reference truncates the table there).  Rows equal on every key keep
 * their input order. */
int gf_skygen_order(const gf_skygen_cols *in, uint64_t n, const gf_skygen_cols *out,
                    uint64_t *nvisible, void *stream);

/* The clipper stage (os_clipper::process, skygen/os_skygen.cpp:997-1044)
 * for the analytic footprints (unions of rects and beams): a visible row of [0, n) whose
 * projection-plane position deprojects outside `foot` gets hidden = 1. */
int gf_clip_footprint(const gf_skygen_params *sp, const gf_footprint *foot, gf_col projIdx,
                      gf_col projXY, gf_col hidden, uint64_t n, void *stream);

/* Draw plan for the cumulative-table draw below: per (pixel, DM-bin) star
 * counts and their exclusive prefix sums, device resident. */
/* Capacity-limited, resumable reference-order draw: skygenHost::drawSources
 * with runtime_state (skygen/skyconfig_impl.h:434-534, skygen/skygen.h:
 * 184-257), where a draw kernel stops when the output table is full and the
 * next call resumes every thread where it stopped.  gf_skygen_draw_begin
 * keeps the streams' starting states, counts every thread's stars (moving
 * the streams past the whole draw, as the unlimited draw leaves them) and
 * returns their number; each gf_skygen_draw_next writes the next
 * min(capacity, remaining) stars in (thread, crawl) order into `out`,
 * resuming the thread the previous call stopped inside; *nwritten is 0 once
 * all are written.  The calls' rows concatenated equal one
 * gf_skygen_draw_compat of unlimited capacity.  The state holds device
 * memory for the stream copy and 12 bytes per thread until
 * gf_skygen_draw_end. */
typedef struct gf_draw_state gf_draw_state;
int  gf_skygen_draw_begin(const gf_skygen_params *sp, const gf_model *model, const gf_pencil_beam *pixels,
                          const gf_extinction *ext, const gf_footprint *foot, gf_rng *rng, gf_draw_state **state,
                          uint64_t *nstars, void *stream);
int  gf_skygen_draw_next(gf_draw_state *state, const gf_skygen_cols *out, uint64_t capacity, uint64_t *nwritten,
                         void *stream);
void gf_skygen_draw_end(gf_draw_state *state);

typedef struct gf_draw_plan gf_draw_plan;

/* Poisson-draw the star count of every (pixel, DM-bin) pair from the
 * expectations `wpk` written by gf_skygen_count.  Unit u = p*K + k is drawn
 * from stream u mod nstreams, units of one stream in increasing u.
 * Returns the total in *nstars. */
int gf_skygen_plan(const gf_skygen_params *sp, const gf_model *model,
                   const gf_pencil_beam *pixels, const gf_extinction *ext,
                   const double *wpk, gf_rng *rng, gf_draw_plan **plan,
                   uint64_t *nstars, void *stream);

/* Emit the stars of rows [row_begin, row_end) of a plan (a chunk of the
 * catalog) into `out` starting at row 0: each star's absolute-magnitude bin
 * is drawn from the cumulative luminosity-function window of its (pixel,
 * DM-bin) pair (FP64 prefix table), then position, M, m, DM, D, XYZ,
 * extinction and the footprint flag are drawn as in draw_stars.  Chunks
 * must be emitted in increasing row order (each stream continues where the
 * previous chunk left it). */
int gf_skygen_emit(const gf_draw_plan *plan, uint64_t row_begin, uint64_t row_end,
                   const gf_footprint *foot, gf_rng *rng,
                   const gf_skygen_cols *out, void *stream);

int gf_skygen_plan_free(gf_draw_plan *plan);

/* ---------------------------------------------------------------------- */
/* Output stages (os_fitsout, pipeline.cpp:767-1035; os_textout, :199-288)  */
/* ---------------------------------------------------------------------- */

/* One output column: a device column of col.width fields per row. */
typedef struct gf_outcol {
    gf_col          col;
    const uint32_t *map;     /* int columns: a value v < nmap is written as map[v]
                              * (device memory; the sequential component index to the
                              * user's component id), or NULL */
    uint32_t        nmap;
    char            type;    /* 'E' float, 'D' double, 'J' int32 (the FITS TFORM codes) */
    char            flags;   /* text: 1 = the ' ' flag (a space before a non-negative value) */
    int16_t         width;   /* text: minimum field width */
    int32_t         prec;    /* text: "%.Pf" digits after the point; -1 for "%d" */
} gf_outcol;

/* The visible rows (hidden == 0) of rows [0, nrows), in order, as FITS
 * binary-table records: the columns' fields in order, big-endian, written
 * to `out` (device memory, cap_rows records).  *nvisible receives the
 * number of records (the rows os_fitsout writes, write_fits_rows :804-866).
 * GF_E_CAPACITY if they do not fit. */
int gf_fits_pack_rows(const gf_outcol *cols, int ncols, gf_col hidden, uint64_t nrows, void *out,
                      uint64_t cap_rows, uint64_t *nvisible, void *stream);

/* The visible rows as text lines (otable::serialize_body, otable.cpp:562-582,
 * with fmtout, otable.h:39-94): each field as snprintf formats it with
 * "%[ ]W.Pf" (floats promoted to double) or "%[ ]Wd", fields separated by
 * one space, '\n' at the end of a line.  Values must be finite with
 * |v| < 1e10 (GF_E_ARG otherwise).  *nbytes receives the text's length (also
 * when it exceeds cap: GF_E_CAPACITY), *nvisible the number of lines. */
int gf_text_format_rows(const gf_outcol *cols, int ncols, gf_col hidden, uint64_t nrows, char *out,
                        uint64_t cap, uint64_t *nbytes, uint64_t *nvisible, void *stream);

/* Sky-count maps (os_countsMap, pipeline.cpp:293-762): every visible row
 * lands in the beam (X, Y) of its position column xy (degrees; with
 * equalarea, first Lambert-projected onto the north (map 0, b > 0) or south
 * (map 1) polar map, os_countsMap::threaded_process :418-494) and, for each
 * element i < zmag_width of its z column, in z bin Z; a coadd c counts the
 * row in every bin from the bin of min_i(z_i - coadd_offsets[c][i]) up.
 * Bin b of an axis with origin a0 and width d holds values whose
 * (int)((v - a0) / d + 0.5) is b, clamped to [0, n - 1] (find_bin :361-367).
 * Counts are uint32 in [beam][Z][zwidth] order, beam = (X ny + Y) nmaps + map,
 * zwidth = zmag_width + ncoadd, nmaps = equalarea ? 2 : 1. */
typedef struct gf_countsmap_params {
    float        x0, dx, y0, dy, z0, dz;
    int32_t      nx, ny, nz;
    int32_t      zmag_width, ncoadd, equalarea;
    const float *coadd_offsets;   /* device, ncoadd x zmag_width, or NULL */
    double       proj[2][4];      /* per map: l0, sin(phi1), cos(phi1), unused (radians) */
} gf_countsmap_params;

/* Adds the visible rows of [0, nrows) to `counts` (device, nmaps nx ny nz
 * zwidth uint32); the coadd entries receive one count at the row's coadd
 * bin only (gf_countsmap_finish makes them cumulative).  beams[beam] (device
 * bytes) is set to 1 for every beam a row lands in; *total (device) grows by
 * the number of z-bin counts added (m_total). */
int gf_countsmap_bin(const gf_countsmap_params *p, gf_col xy, gf_col z, gf_col hidden, uint64_t nrows,
                     uint32_t *counts, uint8_t *beams, unsigned long long *total, void *stream);
/* The coadd entries of every (beam, coadd) made cumulative over Z. */
int gf_countsmap_finish(const gf_countsmap_params *p, uint32_t *counts, void *stream);

/* ---------------------------------------------------------------------- */
/* Catalog pipeline (generate_catalog, pipeline.cpp:1313-1445)             */
/* ---------------------------------------------------------------------- */

/* A loaded `galfast catalog cmd.conf` job: definitions, skygen input,
 * models, footprint and per-star modules, with every table resident on the
 * device and the persistent RNG streams seeded from `seed`
 * (common/cux.cpp:404-435).  Output stages (textout, fitsout, countsMap)
 * are accepted in the module list and ignored: rows are handed to the
 * caller's chunk callback instead. */
typedef struct gf_catalog gf_catalog;

typedef struct gf_catalog_info_t {
    uint64_t npixels;
    int32_t  nm, nM, nmodels, nbands, nabsmag, pad;
    double   footprint_covered_sr;
} gf_catalog_info_t;

/* Per-run statistics; times in seconds, measured with HIP events on the
 * catalog's stream. */
typedef struct gf_run_stats {
    double   expected;      /* sum of nstarsExpectedToGenerate over models */
    uint64_t generated;     /* rows drawn (hidden included) */
    uint64_t rows_chain;    /* rows passed through the module chain */
    uint64_t n_emit, n_chain;
    double   t_plan, t_emit, t_chain, t_total;
    uint64_t visible;       /* rows with hidden == 0 */
    double   t_stage[6];    /* module chain launches: draw pass; row pass for
                               FeH/GaussianFeH, Bond2010, photometry,
                               photometric errors, vel2pm + gal2other */
    double   t_emit_stars, t_emit_units;   /* the two launches inside t_emit: the stars and their units */
} gf_run_stats;

/* Called once per output chunk, in row order, with device columns holding
 * `nrows` rows (hidden rows included; hidden != 0 marks rows the reference
 * drops).  The work that wrote them has completed; the columns stay valid
 * until the callback returns (the run may already be computing the next
 * chunk into other buffers). */
typedef int (*gf_chunk_fn)(void *user, const gf_table *cols, uint64_t nrows, int comp);

/* Message of the last GF_E_CONFIG failure on this thread. */
const char *gf_last_error(void);
int  gf_catalog_open(const char *cmd_conf, uint32_t seed, const char *datadir, gf_catalog **out);
/* Keys: "chunk_rows" (rows per output chunk, default 3e7), "draw" (0, the
 * default: the cumulative-table draw, chunks in (pixel, DM bin) order with
 * hidden rows kept; 1: the reference's crawl, each model one batch in the
 * order of drawSources -- hidden rows dropped, the rest sorted by l, b, DM,
 * M, Am -- then the clipper and the modules), "overlap" (1,
 * the default: the module chain's row passes run on a second HIP stream
 * while the next rows are emitted and drawn; 0: one stream), "acc_mode"
 * (GF_ACC_*), "nstars" (force the expected total, os_skygen.cpp:886-901). */
int  gf_catalog_set(gf_catalog *c, const char *key, double value);
int  gf_catalog_info(const gf_catalog *c, gf_catalog_info_t *info);
/* The user's component id of each sequential component index (the `comp`
 * column holds the index; componentMap, common/compmgr.h:35-49): writes up
 * to cap ids, returns their number. */
int  gf_catalog_component_ids(const gf_catalog *c, uint32_t *ids, int cap);
/* Count pass for every model (os_skygen::run, skygen/os_skygen.cpp:876-884);
 * per_model (optional) receives nmodels results. */
int  gf_catalog_count(gf_catalog *c, gf_count_result *per_model, double *expected);
/* Draw every model and run the module chain chunk by chunk
 * (skygenHost<T>::drawSources, skygen/skyconfig_impl.h:434-545).  The RNG
 * streams persist: a second run on the same catalog continues them and
 * draws another realization, a catalog opened again with the same seed
 * repeats the first. */
int  gf_catalog_run(gf_catalog *c, gf_chunk_fn fn, void *user, gf_run_stats *stats);

/* Output stages for a catalog run (os_textout, pipeline.cpp:199-288;
 * os_fitsout, :767-1035): the job's output columns in the reference's order
 * and names, the visible rows of every chunk formatted (text) or packed
 * (FITS binary table CATALOG, then a METADATA table with the column
 * definitions) on the device and appended to `path`.  Pass
 * gf_writer_chunk with the writer as the chunk callback of gf_catalog_run;
 * gf_writer_close finishes the file, returns the rows written and frees the
 * writer. */
#define GF_OUT_TEXT 0
#define GF_OUT_FITS 1
/* GF_OUT_COUNTSMAP: the job's countsMap output stage (its configuration from
 * the cmd.conf `output` or `modules` list; GF_E_CONFIG without one) binned
 * on the device chunk by chunk and written to `path` as the reference's
 * text table when the writer closes (os_countsMap::~os_countsMap :684-762);
 * the rows it returns are the visible rows binned. */
#define GF_OUT_COUNTSMAP 2
typedef struct gf_writer gf_writer;
int  gf_writer_open(const gf_catalog *c, const char *path, int kind, gf_writer **out);
int  gf_writer_chunk(void *writer, const gf_table *cols, uint64_t nrows, int comp);
int  gf_writer_close(gf_writer *w, uint64_t *rows);
void gf_catalog_close(gf_catalog *c);

/* Host-side copies of everything a catalog job was built from: the sky
 * parameters, pixels, models (luminosity-function texels in host memory),
 * footprint, module parameter blocks (tables in host memory) and the
 * initial stream table.  Valid until gf_catalog_close.  Lets a CPU
 * implementation replay the same job. */
typedef struct gf_catalog_host_view {
    gf_skygen_params      sp;
    const gf_pencil_beam *pixels;
    int32_t               nmodels;
    int32_t               nbands;
    const gf_model       *models;
    gf_footprint          foot;
    gf_chain              chain;
    int32_t               nabsmag;
    uint32_t              nstreams;
    const uint32_t       *rng_state0;
    gf_extinction         ext;      /* host tables; all data NULL without extinction */
    const gf_skygen_params *model_sp;   /* per model: sp with the model's absolute-magnitude
                                         * range (its LF's support, hint_absmag) */
} gf_catalog_host_view;

int gf_catalog_host_view_get(const gf_catalog *c, gf_catalog_host_view *view);

/* ---------------------------------------------------------------------- */
/* Request server (the build's own; the reference has no serving mode)     */
/* ---------------------------------------------------------------------- */

/* A footprint-generation request: a circular beam (Galactic l, b, radius in
 * degrees) or, with radius = 0, the Galactic rect l in [l, l1], b in [b, b1]
 * (degrees; 0 < l1 - l <= 360, -90 <= b < b1 <= 90), its own faint limit and
 * its own seed.  Everything else (models, absolute-magnitude binning, bright
 * limit, pixel size, extinction maps, modules) comes from the server's base
 * configuration. */
typedef struct gf_request {
    uint64_t id;
    double   l, b, radius;
    float    m1;
    uint32_t seed;
    double   l1, b1;
} gf_request;

typedef struct gf_result {
    uint64_t id;
    int32_t  status;        /* 0, or a GF_E_* code */
    int32_t  batch_size;    /* requests in the batch that carried it */
    uint64_t npixels;
    uint64_t generated;     /* rows drawn */
    uint64_t visible;       /* rows with hidden == 0 */
    double   expected;      /* expected stars (pixelized footprint) */
    double   t_submit, t_start, t_done;   /* seconds on the server's clock */
    uint64_t device_bytes;  /* device bytes the request was admitted with (its units,
                             * pixels and streams, and its rows bounded by the counted
                             * expectation plus six Poisson deviations); 0 before a count */
    uint64_t batch;         /* sequence number of the batch that carried it */
} gf_result;

/* Called on a server thread when a request completes.  cols are the batch's
 * device columns; the request owns rows [row0, row0 + nrows).  They are
 * valid only during the call. */
typedef int (*gf_result_fn)(void *user, const gf_result *res, const gf_table *cols, uint64_t row0,
                            uint64_t nrows);

typedef struct gf_server_config {
    uint64_t hbm_budget;    /* device bytes the server may use for batches */
    int32_t  max_batch;     /* requests per batch */
    int32_t  workers;       /* batches in flight (one HIP stream each) */
    int32_t  plain;         /* 1: no batching and no overlap (one request at a time) */
    int32_t  linger_us;     /* a worker with fewer than max_batch ready requests waits up
                             * to this long for more before it launches */
    int32_t  eager_chain;   /* 1: launch the module chain of each batch directly instead of
                             * replaying the worker's captured HIP graph of it */
    int32_t  pad;
} gf_server_config;

typedef struct gf_server_stats {
    uint64_t submitted, completed, batches, rows, visible;
    double   t_first_submit, t_last_done;
    double   t_prep;    /* host seconds spent preparing requests (all threads) */
    double   t_batch;   /* seconds workers spent running batches (all workers) */
    double   t_wait;    /* of which waiting for their device streams */
    double   t_deliver; /* of which handing results to the callback */
} gf_server_stats;

typedef struct gf_server gf_server;

/* cmd_conf: a catalog command configuration (its models, skygen binning and
 * modules are used; its footprint is ignored). */
int    gf_server_create(const char *cmd_conf, const char *datadir, const gf_server_config *cfg,
                        gf_result_fn fn, void *user, gf_server **out);
/* Thread-safe; the submission time is taken from the server's clock. */
int    gf_server_submit(gf_server *s, const gf_request *req);
/* Block until every submitted request has completed. */
int    gf_server_wait_idle(gf_server *s);
int    gf_server_stats_get(gf_server *s, gf_server_stats *st);
double gf_server_clock(void);
void   gf_server_destroy(gf_server *s);

/* Number of rows with hidden == 0 among the first n (synchronous). */
int gf_count_visible(gf_col hidden, uint64_t n, uint64_t *visible, void *stream);

/* Host-side table construction, exposed for hosts that build their own
 * parameter blocks:
 * - the 1024-texel luminosity-function texture with zero clipping
 *   (load_resample_and_clip_texture_1D, common/cux.cpp:368-384);
 * - the photometric-system colour/flag tables
 *   (modules/photometry_host.cpp:175-349), [tex][nMr][nFeH][4] floats,
 *   returning the number of float4 tables or a negative error. */
int gf_lf_texture_host(const double *X, const double *Y, int n, int nsamp, float *tex, float tc[2]);
/* The translation T and rotation M (row-major) of a model or module's
 * `center` and `orientation` keys, xyz' = M (xyz + T), given the
 * definitions' Rg and z0 (load_transform, common/transform.cpp:345-388):
 * centre systems gal (sph | xyz), equ (sph), galcentric, galplane;
 * orientations gal, equ, galcentric, galplane, each optionally followed by
 * `axisangle x y z angle` or `euler phi theta psi` (degrees).  GF_E_CONFIG
 * (with gf_last_error) for anything else. */
int gf_transform_host(const char *center, const char *orientation, double Rg, double z0, float T[3], float M[9]);
/* The pixels of a server request's circular beam (centre l, b and radius
 * in degrees, dx-degree cells), as gf_footprint_cover and the catalog's
 * pixelization would give them, computed on the host: writes up to cap
 * beams, returns the number of pixels.  No reference counterpart (the
 * reference pixelizes footprints through its polygon clipper,
 * skygen/os_skygen.cpp:940-995). */
int gf_beam_pixels_host(double l, double b, double radius, double dx, gf_pencil_beam *out, int cap);
/* The same for a server request's Galactic rect [l0, l1] x [b0, b1]
 * (degrees). */
int gf_rect_pixels_host(double l0, double l1, double b0, double b1, double dx, gf_pencil_beam *out, int cap);
int gf_photometry_tables_host(const double *rows, int nrows, int ncolors, double dFeH, double dMr,
                              float *colors, float *flags, size_t cap_floats, int *nFeH, int *nMr,
                              float tc[4]);

#ifdef __cplusplus
}
#endif

#endif /* GALFAST_HIP_H */
