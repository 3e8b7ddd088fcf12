// Seeding and device management of the multiply-with-carry streams
// (common/cuda_rng.h:74-103, 466-497; common/cux.cpp:404-435).
#include "gf_runtime.h"
#include "gf_rng.h"
#include <random>

namespace {

// POSIX drand48 family: X <- (0x5DEECE66D * X + 0xB) mod 2^48.
struct Rand48 {
	uint64_t X;
	void seed(uint32_t s) { X = ((uint64_t)s << 16) | 0x330Eull; }
	void step() { X = (0x5DEECE66Dull * X + 0xBull) & 0xFFFFFFFFFFFFull; }
	double drand() { step(); return (double)X * (1.0 / 281474976710656.0); }
	uint32_t lrand() { step(); return (uint32_t)(X >> 17); }
};

// (int)v with the x86-64 cvttsd2si result for out-of-range values.
__host__ __device__ inline int32_t trunc_to_int_x86(double v)
{
	if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
	return (int32_t)v;
}

// The drand48 state n steps after X: the n-th power of the step
// X <- (0x5DEECE66D X + 0xB) mod 2^48, formed by squaring (the step k times
// is X <- (A X + C) mod 2^48; twice that is A^2 and C (A + 1)).  Products
// are taken mod 2^64, whose low 48 bits are the products mod 2^48.
__device__ inline uint64_t rand48_skip(uint64_t X, uint64_t n)
{
	const uint64_t M48 = 0xFFFFFFFFFFFFull;
	uint64_t A = 0x5DEECE66Dull, C = 0xBull;
	for (; n; n >>= 1) {
		if (n & 1u) X = (A * X + C) & M48;
		C = (C * (A + 1u)) & M48;
		A = (A * A) & M48;
	}
	return X;
}

// gf_mwc_seed_host on the device: stream i of table g takes the generator's
// draws 2i+1 (its carry, drand48 times its multiplier) and 2i+2 (its value,
// lrand48), so every lane jumps to its own pair.  Block row y seeds table
// segs[y].
__global__ void k_seed_streams(const gf_seed_seg *segs, const uint32_t *primes, uint32_t *out)
{
	const gf_seed_seg g = segs[blockIdx.y];
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.nstreams) return;
	const uint64_t M48 = 0xFFFFFFFFFFFFull;
	uint64_t X = rand48_skip(((uint64_t)g.seed << 16) | 0x330Eull, 2 * (uint64_t)i);
	X = (0x5DEECE66Dull * X + 0xBull) & M48;
	const double u = (double)X * (1.0 / 281474976710656.0);
	X = (0x5DEECE66Dull * X + 0xBull) & M48;
	const uint32_t a = primes[i];
	uint32_t *t = out + g.offset;
	t[i] = a;
	t[g.nstreams + i] = (uint32_t)trunc_to_int_x86(u * (double)a);
	t[2 * (size_t)g.nstreams + i] = (uint32_t)(X >> 17);
}

__global__ void mwc_sample_kernel(uint32_t *state, uint32_t ns, int kind, float param, uint32_t n, void *out)
{
	uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= ns) return;
	gf_mwc r = gf_mwc_load(state, ns, s);
	for (uint32_t i = 0; i < n; i++) {
		size_t o = (size_t)i * ns + s;
		if (kind == 0) ((float *)out)[o] = r.uniform();
		else if (kind == 1) ((float *)out)[o] = r.gaussian(param);
		else if (kind == 2) ((uint32_t *)out)[o] = r.poisson(param);
		else ((uint32_t *)out)[o] = r.poisson(param * (float)(1 + s));
	}
	gf_mwc_store(state, ns, s, r);
}

// Jump tables of the streams: tab[b * ns + s] = a_s^(2^b) * 2^64 mod m_s for
// b < GF_JUMP_BITS (gf_mwc_mod).
__global__ void k_jump_table(const uint32_t *state, uint32_t ns, uint64_t *tab)
{
	uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= ns) return;
	const gf_mwc_mod md(state[s]);
	uint64_t y = 1ull << 32;
	for (int b = 0; b < GF_JUMP_BITS; b++) {
		tab[(size_t)b * ns + s] = y;
		y = md.mul(y, y);
	}
}

// Every stream n steps ahead.
__global__ void k_jump(uint32_t *state, uint32_t ns, uint64_t n, const uint64_t *tab)
{
	uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= ns) return;
	gf_mwc r = gf_mwc_load(state, ns, s);
	const gf_mwc_mod md(r.a);
	gf_mwc_set_z(r, gf_mwc_jump_z(md, gf_mwc_z(r), n, tab + s, ns));
	gf_mwc_store(state, ns, s, r);
}

} // namespace

namespace gf {

int mwc_jump_table(const gf_rng *rng, uint64_t *tab, hipStream_t st)
{
	GF_CHECK(rng && rng->state && tab, GF_E_ARG);
	k_jump_table<<<gf::blocks_for(rng->nstreams, 256), 256, 0, st>>>(rng->state, rng->nstreams, tab);
	GF_LAUNCHED();
	return 0;
}

int mwc_seed_streams(const gf_seed_seg *segs_dev, uint32_t nsegs, uint32_t max_streams, const uint32_t *primes,
                     uint32_t *out, hipStream_t st)
{
	GF_CHECK(segs_dev && primes && out && nsegs <= 65535u, GF_E_ARG);
	if (!nsegs || !max_streams) return 0;
	k_seed_streams<<<dim3((max_streams + 255) / 256, nsegs), 256, 0, st>>>(segs_dev, primes, out);
	GF_LAUNCHED();
	return 0;
}

int mwc_jump(gf_rng *rng, uint64_t n, const uint64_t *tab, hipStream_t st)
{
	GF_CHECK(rng && rng->state && tab, GF_E_ARG);
	if (!n) return 0;
	k_jump<<<gf::blocks_for(rng->nstreams, 256), 256, 0, st>>>(rng->state, rng->nstreams, n, tab);
	GF_LAUNCHED();
	return 0;
}

} // namespace gf

extern "C" {

uint32_t gf_seed_from_pipeline_seed(uint32_t pipeline_seed)
{
	// gsl_rng_set(mt19937, 0) seeds with 4357; the first output divided by
	// 2^32 is gsl_rng_uniform, narrowed to float by rng_gsl_t::uniform
	std::mt19937 mt(pipeline_seed == 0 ? 4357u : pipeline_seed);
	float u = (float)((double)mt() / 4294967296.0);
	return (uint32_t)(u * (float)(1 << 24));
}

int gf_mwc_seed_host(uint32_t seed, const uint32_t *primes, uint32_t nprimes,
                     uint32_t nstreams, uint32_t *out_state)
{
	GF_CHECK(primes && out_state, GF_E_ARG);
	if (nstreams % 64) nstreams += 64 - nstreams % 64;
	GF_CHECK(nstreams <= nprimes, GF_E_ARG);
	Rand48 r;
	r.seed(seed);
	for (uint32_t i = 0; i < nstreams; i++) {
		uint32_t a = primes[i];
		out_state[i] = a;
		out_state[nstreams + i] = (uint32_t)trunc_to_int_x86(r.drand() * (double)a);
		out_state[2 * nstreams + i] = r.lrand();
	}
	return (int)nstreams;
}

int gf_mwc_seed_device(uint32_t seed, const uint32_t *primes, uint32_t nprimes, uint32_t nstreams, uint32_t *state,
                       void *stream)
{
	GF_CHECK(primes && state && nstreams > 0 && nstreams % 64 == 0 && nstreams <= nprimes, GF_E_ARG);
	hipStream_t st = gf::as_stream(stream);
	void *p;
	int rc = gf::scratch(GF_SLOT_SEED, sizeof(gf_seed_seg), &p);
	if (rc) return rc;
	const gf_seed_seg g{seed, nstreams, 0};
	GF_HIP(hipMemcpyAsync(p, &g, sizeof g, hipMemcpyHostToDevice, st));
	rc = gf::mwc_seed_streams((const gf_seed_seg *)p, 1, nstreams, primes, state, st);
	if (rc) return rc;
	// the copy reads g on this stack frame
	GF_HIP(hipStreamSynchronize(st));
	return 0;
}

int gf_rng_alloc(gf_rng *rng, uint32_t nstreams)
{
	GF_CHECK(rng && nstreams > 0, GF_E_ARG);
	rng->nstreams = nstreams;
	rng->reserved = 0;
	GF_HIP(hipMalloc((void **)&rng->state, (size_t)3 * nstreams * sizeof(uint32_t)));
	return 0;
}

int gf_rng_upload(gf_rng *rng, const uint32_t *host_state, void *stream)
{
	GF_CHECK(rng && rng->state && host_state, GF_E_ARG);
	GF_HIP(hipMemcpyAsync(rng->state, host_state, (size_t)3 * rng->nstreams * sizeof(uint32_t),
	                      hipMemcpyHostToDevice, gf::as_stream(stream)));
	GF_HIP(hipStreamSynchronize(gf::as_stream(stream)));
	return 0;
}

int gf_rng_download(const gf_rng *rng, uint32_t *host_state, void *stream)
{
	GF_CHECK(rng && rng->state && host_state, GF_E_ARG);
	GF_HIP(hipMemcpyAsync(host_state, rng->state, (size_t)3 * rng->nstreams * sizeof(uint32_t),
	                      hipMemcpyDeviceToHost, gf::as_stream(stream)));
	GF_HIP(hipStreamSynchronize(gf::as_stream(stream)));
	return 0;
}

int gf_rng_free(gf_rng *rng)
{
	GF_CHECK(rng, GF_E_ARG);
	if (rng->state) GF_HIP(hipFree(rng->state));
	rng->state = nullptr;
	rng->nstreams = 0;
	return 0;
}

int gf_mwc_jump(gf_rng *rng, uint64_t n, void *stream)
{
	GF_CHECK(rng && rng->state, GF_E_ARG);
	if (!n) return 0;
	hipStream_t st = gf::as_stream(stream);
	void *p;
	int rc = gf::scratch(GF_SLOT_JUMP, (size_t)GF_JUMP_BITS * rng->nstreams * sizeof(uint64_t), &p);
	if (rc) return rc;
	rc = gf::mwc_jump_table(rng, (uint64_t *)p, st);
	if (rc) return rc;
	return gf::mwc_jump(rng, n, (const uint64_t *)p, st);
}

int gf_mwc_jump_host(uint32_t *state, uint32_t nstreams, uint64_t n)
{
	GF_CHECK(state, GF_E_ARG);
	uint64_t tab[GF_JUMP_BITS];
	for (uint32_t s = 0; s < nstreams; s++) {
		gf_mwc r = gf_mwc_load(state, nstreams, s);
		const gf_mwc_mod md(r.a);
		uint64_t y = 1ull << 32;
		for (int b = 0; b < GF_JUMP_BITS; b++) { tab[b] = y; y = md.mul(y, y); }
		gf_mwc_set_z(r, gf_mwc_jump_z(md, gf_mwc_z(r), n, tab, 1));
		gf_mwc_store(state, nstreams, s, r);
	}
	return 0;
}

int gf_mwc_sample(gf_rng *rng, int kind, float param, uint32_t n, void *out, void *stream)
{
	GF_CHECK(rng && rng->state && out && kind >= 0 && kind <= 3, GF_E_ARG);
	unsigned bs = 256;
	mwc_sample_kernel<<<gf::blocks_for(rng->nstreams, bs), bs, 0, gf::as_stream(stream)>>>(
	    rng->state, rng->nstreams, kind, param, n, out);
	GF_LAUNCHED();
	return 0;
}

} // extern "C"
