// Runtime entry points: device selection, memory helpers, scratch cache,
// a deterministic prefix scan and the error strings.
#include "gf_runtime.h"
#include <mutex>
#include <vector>

namespace gf {

namespace {
struct ScratchEntry { void *ptr = nullptr; size_t bytes = 0; };
// per host thread, so that threads driving different streams never share a
// buffer; [device][slot]
thread_local std::vector<std::vector<ScratchEntry>> g_scratch;
}

int scratch(int slot, size_t bytes, void **ptr)
{
	int dev = 0;
	GF_HIP(hipGetDevice(&dev));
	if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
	auto &v = g_scratch[dev];
	if ((int)v.size() <= slot) v.resize(slot + 1);
	ScratchEntry &e = v[slot];
	if (e.bytes < bytes) {
		if (e.ptr) GF_HIP(hipFree(e.ptr));
		e.ptr = nullptr;
		e.bytes = 0;
		size_t want = bytes < 256 ? 256 : bytes;
		GF_HIP(hipMalloc(&e.ptr, want));
		e.bytes = want;
	}
	*ptr = e.ptr;
	return 0;
}

// Three-phase scan: per-block sums, a single-block scan of the block sums,
// then per-block scans offset by the block prefix.
static constexpr int SCAN_BS = 256;
static constexpr int SCAN_ITEMS = 16;
static constexpr int SCAN_TILE = SCAN_BS * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_BS) scan_reduce(const uint32_t *in, uint64_t n, uint64_t *bsum)
{
	__shared__ uint64_t sh[SCAN_BS];
	uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
	uint64_t s = 0;
	for (int i = 0; i < SCAN_ITEMS; i++) {
		uint64_t idx = base + (uint64_t)i * SCAN_BS + threadIdx.x;
		if (idx < n) s += in[idx];
	}
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int o = SCAN_BS / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(SCAN_BS) scan_blocksums(uint64_t *bsum, uint64_t nb)
{
	// single block: sequential chunks of SCAN_BS, Hillis-Steele within a chunk
	__shared__ uint64_t sh[SCAN_BS];
	uint64_t carry = 0;
	for (uint64_t base = 0; base < nb; base += SCAN_BS) {
		uint64_t idx = base + threadIdx.x;
		uint64_t v = idx < nb ? bsum[idx] : 0;
		sh[threadIdx.x] = v;
		__syncthreads();
		for (int o = 1; o < SCAN_BS; o <<= 1) {
			uint64_t t = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
			__syncthreads();
			sh[threadIdx.x] += t;
			__syncthreads();
		}
		uint64_t incl = sh[threadIdx.x];
		if (idx < nb) bsum[idx] = carry + incl - v;   // exclusive
		__syncthreads();
		carry += sh[SCAN_BS - 1];
		__syncthreads();
	}
	if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ void __launch_bounds__(SCAN_BS) scan_final(const uint32_t *in, uint64_t n, const uint64_t *bsum, uint64_t *offs)
{
	__shared__ uint64_t sh[SCAN_BS];
	uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
	// each thread owns SCAN_ITEMS consecutive elements
	uint64_t my = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
	uint32_t v[SCAN_ITEMS];
	uint64_t s = 0;
	for (int i = 0; i < SCAN_ITEMS; i++) {
		uint64_t idx = my + i;
		v[i] = idx < n ? in[idx] : 0;
		s += v[i];
	}
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int o = 1; o < SCAN_BS; o <<= 1) {
		uint64_t t = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
		__syncthreads();
		sh[threadIdx.x] += t;
		__syncthreads();
	}
	uint64_t run = bsum[blockIdx.x] + sh[threadIdx.x] - s;
	for (int i = 0; i < SCAN_ITEMS; i++) {
		uint64_t idx = my + i;
		if (idx < n) offs[idx] = run;
		run += v[i];
	}
}

__global__ void scan_total(const uint64_t *bsum, uint64_t nb, uint64_t *offs, uint64_t n)
{
	offs[n] = bsum[nb];
}

int exclusive_scan_u32_u64(const uint32_t *in, uint64_t *offs, uint64_t n, hipStream_t st)
{
	uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
	if (nb == 0) nb = 1;
	void *p;
	int rc = scratch(GF_SLOT_SCAN, (nb + 1) * sizeof(uint64_t), &p);
	if (rc) return rc;
	uint64_t *bsum = (uint64_t *)p;
	if (n > 0) {
		scan_reduce<<<(unsigned)nb, SCAN_BS, 0, st>>>(in, n, bsum);
		GF_LAUNCHED();
	} else {
		GF_HIP(hipMemsetAsync(bsum, 0, sizeof(uint64_t), st));
	}
	scan_blocksums<<<1, SCAN_BS, 0, st>>>(bsum, n > 0 ? nb : 1);
	GF_LAUNCHED();
	if (n > 0) {
		scan_final<<<(unsigned)nb, SCAN_BS, 0, st>>>(in, n, bsum, offs);
		GF_LAUNCHED();
	}
	scan_total<<<1, 1, 0, st>>>(bsum, n > 0 ? nb : 1, offs, n);
	GF_LAUNCHED();
	return 0;
}

} // namespace gf

extern "C" {

const char *gf_strerror(int code)
{
	switch (code) {
	case GF_OK: return "success";
	case GF_E_ARG: return "invalid argument";
	case GF_E_CAPACITY: return "output table too small";
	case GF_E_STREAMS: return "more row chunks than random streams";
	case GF_E_CONFIG: return "configuration error";
	case GF_E_IO: return "file could not be read";
	case GF_E_NODEVICE: return "no HIP device";
	case GF_E_BUDGET: return "request exceeds the device-memory budget";
	case GF_E_STATE: return "call out of order";
	case GF_E_SCRATCH: return "input exceeds the fixed scratch of a no-allocation entry point";
	default:
		if (code > 0) return hipGetErrorString((hipError_t)code);
		return "unknown error";
	}
}

int gf_init(int device)
{
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n == 0) return GF_E_NODEVICE;
	GF_CHECK(device >= 0 && device < n, GF_E_ARG);
	GF_HIP(hipSetDevice(device));
	// the address of the no-allocation scratch on this device; a failure
	// here is reported by gf_skygen_count_noalloc, which repeats the set-up
	if (gf::noalloc_setup(device) != 0) (void)hipGetLastError();
	return 0;
}

int gf_sync(void *stream)
{
	GF_HIP(hipStreamSynchronize(gf::as_stream(stream)));
	return 0;
}

int gf_malloc(void **ptr, size_t bytes)
{
	GF_CHECK(ptr, GF_E_ARG);
	GF_HIP(hipMalloc(ptr, bytes ? bytes : 1));
	return 0;
}

int gf_free(void *ptr)
{
	if (ptr) GF_HIP(hipFree(ptr));
	return 0;
}

int gf_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
	GF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, gf::as_stream(stream)));
	GF_HIP(hipStreamSynchronize(gf::as_stream(stream)));
	return 0;
}

int gf_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
	GF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, gf::as_stream(stream)));
	GF_HIP(hipStreamSynchronize(gf::as_stream(stream)));
	return 0;
}

int gf_memset(void *dst, int value, size_t bytes, void *stream)
{
	GF_HIP(hipMemsetAsync(dst, value, bytes, gf::as_stream(stream)));
	return 0;
}

} // extern "C"
