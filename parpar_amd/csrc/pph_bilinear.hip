// Bilinear dense form of the recovery product.
//
// Multiplication in GF(2^16) is a bilinear map over GF(2).  Four levels of
// Karatsuba on the 16-term polynomials give a rank-81 formula:
//   c * x = Red( sum_t E(c)_t E(x)_t e_t ),   t = 0..80,
// where E: GF(2)^16 -> GF(2)^81 is the Karatsuba expansion (at each level a
// block of n bits becomes its low half, high half and their XOR), e_t is the
// recombination of product t into the 31 coefficients of the polynomial product
// and Red reduces modulo x^16 + x^12 + x^3 + x + 1.  Summed over the inputs:
//   out_r[w] = Red( sum_t parity( sum_i E(c_ri)_t E(x_i[w])_t ) e_t ).
// For each t the inner sum is an integer GEMM over 0/1 operands:
//   G_t[r][w] = sum_i Ac_t[r][i] Xw_t[w][i],
// 81 MACs per GF(2^16) mul-add instead of the 256 of the 16x16 bit-matrix.
// The 81 GEMMs run in the hand-written fp4 MFMA kernel over bit-packed planes
// (pph_bl_mfma.hip); this file holds the host formula, the shared workspace
// and the pph_muladd_bilinear entry point.
#include "pph_internal.h"
#include "pph_karatsuba.h"
#include "../../include/parpar_hip.h"

#include <atomic>
#include <map>
#include <mutex>
#include <vector>

namespace pph {

// E(v) in recursive order: E(lo) ++ E(hi) ++ E(lo ^ hi).
static void bl_expand_host(uint32_t v, unsigned n, std::vector<uint8_t>& out) {
	if(n == 1) {
		out.push_back((uint8_t)(v & 1));
		return;
	}
	const unsigned h = n / 2;
	const uint32_t lo = v & ((1u << h) - 1), hi = v >> h;
	bl_expand_host(lo, h, out);
	bl_expand_host(hi, h, out);
	bl_expand_host(lo ^ hi, h, out);
}

// Recombination words W_t (the compile-time table of pph_karatsuba.h): the
// product is the XOR of W_t over the components whose parity is 1.
static const uint16_t* bl_recombine_table() {
	static constexpr BlTable kW = bl_table();
	return kW.w;
}

namespace {

struct BlDevice {
	std::mutex mu;
	uint8_t* buf = nullptr;  // workspace of pph_muladd_bilinear: coefficient planes | word planes | parity bits
	size_t buf_bytes = 0;
	hipEvent_t tail = nullptr;  // end of the last call on the shared workspace
	bool tail_valid = false;
	std::mutex call_mu;         // held across one call's wait, enqueue and record
};

BlDevice* bl_device(int dev) {
	static std::mutex mu;
	static std::map<int, BlDevice*> m;
	std::lock_guard<std::mutex> lk(mu);
	auto& p = m[dev];
	if(!p) p = new BlDevice;
	return p;
}

}  // namespace

}  // namespace pph

using namespace pph;

extern "C" uint16_t pph_bilinear_mul(uint16_t c, uint16_t x) {
	std::vector<uint8_t> ec, ex;
	bl_expand_host(c, 16, ec);
	bl_expand_host(x, 16, ex);
	const uint16_t* w = bl_recombine_table();
	uint16_t r = 0;
	for(unsigned t = 0; t < kBl; t++)
		if(ec[t] & ex[t]) r ^= w[t];
	return r;
}

extern "C" int pph_bilinear_workspace(size_t slot_bytes, unsigned n_in, unsigned n_out, size_t* bytes) {
	if(!bytes || !n_in || !n_out || slot_bytes % kSlotAlign) return PPH_E_INVALID;
	*bytes = blm_workspace_bytes(slot_bytes, n_in, n_out);
	return PPH_OK;
}

extern "C" int pph_bilinear_release(void) {
	int dev = 0;
	if(hipGetDevice(&dev) != hipSuccess) return PPH_E_DEVICE;
	BlDevice* d = bl_device(dev);
	std::lock_guard<std::mutex> lk(d->mu);
	if(d->buf) {
		hipError_t e = hipDeviceSynchronize();
		if(e != hipSuccess) return set_error(e);
		(void)hipFree(d->buf);
	}
	d->buf = nullptr;
	d->buf_bytes = 0;
	return PPH_OK;
}

namespace pph {

size_t bilinear_workspace_bytes(size_t slot_bytes, unsigned n_in, unsigned n_out) {
	return blm_workspace_bytes(slot_bytes, n_in, n_out);
}

// Whether the bilinear form takes this shape (a host check: no device work).
int bilinear_supported(size_t slot_bytes, unsigned n_in, unsigned n_out) {
	if(!n_in || !n_out || n_in > 32768 || slot_bytes % kSlotAlign) return PPH_E_INVALID;
	return PPH_OK;
}

// Enqueue one bilinear mul-add on s with a caller-owned workspace of at least
// bilinear_workspace_bytes.
int bilinear_enqueue(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                     const uint16_t* exps, void* out_slots, unsigned n_out, bool accumulate, void* ws, size_t ws_bytes,
                     const DeviceTables& tab, hipStream_t s, const uint32_t* coef, unsigned coef_stride) {
	if(!n_out) return PPH_OK;
	return blm_enqueue(in_slots, slot_bytes, n_in, input_logs, exps, out_slots, n_out, accumulate, ws, ws_bytes, tab,
	                   s, 0, coef, coef_stride);
}

}  // namespace pph

namespace pph {
namespace {

std::atomic<size_t> g_ws_limit{0};  // test hook: workspace allocations above this fail (0: no limit)

// Grow the shared workspace to `need` bytes; d->mu held.
int reserve_locked(BlDevice* d, size_t need) {
	if(d->buf_bytes >= need) return PPH_OK;
	const size_t lim = g_ws_limit.load(std::memory_order_relaxed);
	if(lim && need > lim) return PPH_E_NOMEM;
	if(d->buf) {
		hipError_t e = hipDeviceSynchronize();
		if(e != hipSuccess) return set_error(e);
		(void)hipFree(d->buf);
		d->buf = nullptr;
		d->buf_bytes = 0;
	}
	hipError_t e = hipMalloc((void**)&d->buf, need);
	if(e != hipSuccess) {
		d->buf = nullptr;
		(void)hipGetLastError();
		return e == hipErrorOutOfMemory ? PPH_E_NOMEM : set_error(e);
	}
	d->buf_bytes = need;
	return PPH_OK;
}

}  // namespace

// A shared workspace large enough for the shape on the current device: after
// PPH_OK, pph_muladd_bilinear of that shape allocates nothing.  PPH_E_NOMEM
// when the workspace does not fit.
int bilinear_reserve(size_t slot_bytes, unsigned n_in, unsigned n_out) {
	int rc = bilinear_supported(slot_bytes, n_in, n_out);
	if(rc) return rc;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return set_error(e);
	BlDevice* d = bl_device(dev);
	std::lock_guard<std::mutex> call(d->call_mu);
	std::lock_guard<std::mutex> lk(d->mu);
	return reserve_locked(d, bilinear_workspace_bytes(slot_bytes, n_in, n_out));
}

}  // namespace pph

// ---------------------------------------------------------------- without allocations
//
// pph_muladd_bilinear_noalloc keeps every buffer its kernels use beyond the
// caller's inputs and outputs in static device memory of this file: the
// GEMM's planes and parity bits in `scratch` (chunked to fit) and the exp
// table in `scratch_exp` (filled by the first call on each device).  Each GPU
// holds its own copy of these
// arrays for the life of the process; in host code their names are
// stand-ins, so their device addresses come from hipGetSymbolAddress, once
// per device, and only those addresses are used there.

static __device__ __align__(256) uint8_t scratch[PPH_NOALLOC_SCRATCH_BYTES];
static __device__ __align__(256) uint16_t scratch_exp[65536];

namespace pph {
namespace {

struct StaticScratch {
	bool ready = false;
	uint8_t* ws = nullptr;
	uint16_t* exp = nullptr;
	int cus = 0;
};
std::mutex g_scratch_mu;
StaticScratch g_scratch[64];

// The static exp table: scratch_exp[k] = 2^k in GF(2^16) (k < 65535),
// scratch_exp[65535] = 0 (the tables' zero entry).
__global__ __launch_bounds__(256) void k_fill_scratch_tables() {
	const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
	if(k > 65535) return;
	if(k == 65535) {
		scratch_exp[k] = 0;
		return;
	}
	auto mul = [](uint32_t a, uint32_t b) {
		uint32_t r = 0;
		for(int i = 0; i < 16; i++)
			if((b >> i) & 1) r ^= a << i;
		for(int i = 30; i >= 16; i--)
			if((r >> i) & 1) r ^= kPoly << (i - 16);
		return r;
	};
	uint32_t r = 1, base = 2;
	for(uint32_t e = k; e; e >>= 1) {
		if(e & 1) r = mul(r, base);
		base = mul(base, base);
	}
	scratch_exp[k] = (uint16_t)r;
}

// Set-up on the first call on a device: the arrays' device addresses and the
// tables (filled on the caller's stream and waited for).
int scratch_setup(int dev, hipStream_t s, StaticScratch** out) {
	if(dev < 0 || dev >= 64) return PPH_E_UNSUPPORTED;
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	StaticScratch& st = g_scratch[dev];
	if(!st.ready) {
		void* ws = nullptr;
		void* ex = nullptr;
		hipError_t e = hipGetSymbolAddress(&ws, HIP_SYMBOL(scratch));
		if(e == hipSuccess) e = hipGetSymbolAddress(&ex, HIP_SYMBOL(scratch_exp));
		int cus = 0;
		if(e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
		if(e != hipSuccess) return set_error(e);
		hipLaunchKernelGGL(k_fill_scratch_tables, dim3(256), dim3(256), 0, s);
		e = hipGetLastError();
		if(e == hipSuccess) e = hipStreamSynchronize(s);
		if(e != hipSuccess) return set_error(e);
		st.ws = (uint8_t*)ws;
		st.exp = (uint16_t*)ex;
		st.cus = cus;
		st.ready = true;
	}
	*out = &st;
	return PPH_OK;
}

}  // namespace
}  // namespace pph

extern "C" int pph_muladd_bilinear_noalloc(const void* in_slots, size_t slot_bytes, unsigned n_in,
                                           const uint16_t* input_logs, const uint16_t* exps, void* out_slots,
                                           unsigned n_out, int accumulate, void* stream) {
	if(!n_out) return PPH_OK;
	if(!in_slots || !out_slots || !input_logs || !exps || !n_in || n_in > 32768 || slot_bytes % kSlotAlign)
		return PPH_E_INVALID;
	// the planes must fit the static scratch however finely the work is chunked
	if(blm_workspace_bytes(slot_bytes, n_in, n_out, PPH_NOALLOC_SCRATCH_BYTES) > PPH_NOALLOC_SCRATCH_BYTES)
		return PPH_E_INVALID;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return set_error(e);
	StaticScratch* st = nullptr;
	int rc = scratch_setup(dev, (hipStream_t)stream, &st);
	if(rc) return rc;
	DeviceTables t;
	t.exp = st->exp;
	t.cus = st->cus;
	return blm_enqueue(in_slots, slot_bytes, n_in, input_logs, exps, out_slots, n_out, accumulate != 0, st->ws,
	                   PPH_NOALLOC_SCRATCH_BYTES, t, (hipStream_t)stream, PPH_NOALLOC_SCRATCH_BYTES);
}

extern "C" int pph_debug_bilinear_workspace_limit(size_t bytes) {
	g_ws_limit.store(bytes, std::memory_order_relaxed);
	return PPH_OK;
}

namespace pph {
int bilinear_call(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                  const uint16_t* exps, const uint32_t* coef, unsigned coef_stride, void* out_slots, unsigned n_out,
                  int accumulate, hipStream_t stream) {
	if(!n_out) return PPH_OK;
	if(!in_slots || !out_slots || !n_in || n_in > 32768 || slot_bytes % kSlotAlign) return PPH_E_INVALID;
	if(coef ? coef_stride < n_out : (!input_logs || !exps)) return PPH_E_INVALID;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return set_error(e);
	DeviceTables* tab = nullptr;
	int rc = device_tables(&tab);
	if(rc) return rc;
	rc = bilinear_supported(slot_bytes, n_in, n_out);
	if(rc) return rc;
	BlDevice* d = bl_device(dev);
	std::lock_guard<std::mutex> call(d->call_mu);
	void* ws = nullptr;
	size_t ws_bytes = 0;
	{
		std::lock_guard<std::mutex> lk(d->mu);
		rc = reserve_locked(d, bilinear_workspace_bytes(slot_bytes, n_in, n_out));
		if(rc) return rc;
		ws = d->buf;
		ws_bytes = d->buf_bytes;
		if(!d->tail) {
			e = hipEventCreateWithFlags(&d->tail, hipEventDisableTiming);
			if(e != hipSuccess) return set_error(e);
		}
		// calls from different streams share the workspace: each waits for the last
		if(d->tail_valid) {
			e = hipStreamWaitEvent((hipStream_t)stream, d->tail, 0);
			if(e != hipSuccess) return set_error(e);
		}
	}
	rc = bilinear_enqueue(in_slots, slot_bytes, n_in, input_logs, exps, out_slots, n_out, accumulate != 0, ws, ws_bytes,
	                      *tab, (hipStream_t)stream, coef, coef_stride);
	std::lock_guard<std::mutex> lk(d->mu);
	e = hipEventRecord(d->tail, (hipStream_t)stream);
	if(e != hipSuccess) return set_error(e);
	d->tail_valid = true;
	return rc;
}
}  // namespace pph

extern "C" int pph_muladd_bilinear(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                                   const uint16_t* exps, void* out_slots, unsigned n_out, int accumulate,
                                   void* stream) {
	if(n_out && (!input_logs || !exps)) return PPH_E_INVALID;
	return bilinear_call(in_slots, slot_bytes, n_in, input_logs, exps, nullptr, 0, out_slots, n_out, accumulate,
	                     (hipStream_t)stream);
}

extern "C" int pph_muladd_bilinear_coef(const void* in_slots, size_t slot_bytes, unsigned n_in,
                                        const uint32_t* coef_log2, unsigned coef_stride, void* out_slots,
                                        unsigned n_out, int accumulate, void* stream) {
	if(n_out && !coef_log2) return PPH_E_INVALID;
	return bilinear_call(in_slots, slot_bytes, n_in, nullptr, nullptr, coef_log2, coef_stride, out_slots, n_out,
	                     accumulate, (hipStream_t)stream);
}
