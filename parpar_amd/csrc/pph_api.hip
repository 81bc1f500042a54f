// C ABI of libparpar_hip.so: device-layer entry points and the pph_proc
// recovery backend (the IPAR2ProcBackend contract, gf16/controller.h:77-217).
#include "pph_internal.h"
#include "pph_copy_pool.h"
#include "pph_task_queue.h"
#include "../../include/parpar_hip.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <thread>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

namespace pph {

static thread_local int g_last_hip = 0;

int set_error(hipError_t e) {
	g_last_hip = (int)e;
	return PPH_E_DEVICE;
}

#define PPH_HIP(expr)                                  \
	do {                                               \
		hipError_t _e = (expr);                        \
		if(_e != hipSuccess) return pph::set_error(_e); \
	} while(0)

// ---------------------------------------------------------------- host tables
struct HostTables {
	uint16_t exp[65536];
	uint16_t log[65536];
	uint16_t in_log[32768];
	uint16_t rowmask[16 * 64];
	uint32_t crc[16 * 256];  // slice-by-16: table t maps a byte followed by t zero bytes
	HostTables() {
		for(uint32_t i = 0; i < 256; i++) {
			uint32_t c = i;
			for(int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : (c >> 1);
			crc[i] = c;
		}
		for(int tab = 1; tab < 16; tab++)
			for(uint32_t i = 0; i < 256; i++)
				crc[tab * 256 + i] = (crc[(tab - 1) * 256 + i] >> 8) ^ crc[crc[(tab - 1) * 256 + i] & 255];
		uint32_t n = 1;
		for(uint32_t e = 0; e < kOrder; e++) {
			exp[e] = (uint16_t)n;
			log[n] = (uint16_t)e;
			n <<= 1;
			if(n > 0xFFFF) n ^= kPoly;
		}
		exp[65535] = 0;        // zero entry of the lookup table
		log[0] = kLogZero;     // log-domain encoding of 0
		// PAR2 input constants: exponents n >= 1 not divisible by 3, 5, 17 or 257,
		// in increasing order (gf16/gfmat_coeff.c:12-22)
		uint32_t e = 0;
		for(int i = 0; i < 32768; i++) {
			do { e++; } while(!(e % 3) || !(e % 5) || !(e % 17) || !(e % 257));
			in_log[i] = (uint16_t)e;
		}
		// bit-matrix rows: M_c[j][b] = bit j of (c * z^b) = XOR_k c_k bit j of z^(k+b);
		// Q_j[k] = sum_b 2^b bit j of z^(k+b); T[j][q][n] = XOR of Q_j[4q+t] over bits t of n
		uint32_t pw[31];
		pw[0] = 1;
		for(int m = 1; m < 31; m++) {
			pw[m] = pw[m - 1] << 1;
			if(pw[m] & 0x10000) pw[m] ^= kPoly;
		}
		for(int jj = 0; jj < 16; jj++) {
			uint16_t Q[16];
			for(int k = 0; k < 16; k++) {
				uint32_t m = 0;
				for(int b = 0; b < 16; b++) m |= ((pw[k + b] >> jj) & 1u) << b;
				Q[k] = (uint16_t)m;
			}
			for(int qq = 0; qq < 4; qq++)
				for(int nn = 0; nn < 16; nn++) {
					uint16_t v = 0;
					for(int t = 0; t < 4; t++)
						if(nn & (1 << t)) v ^= Q[4 * qq + t];
					rowmask[jj * 64 + qq * 16 + nn] = v;
				}
		}
	}
};
static const HostTables& tables() {
	static HostTables t;
	return t;
}
const uint16_t* host_exp() { return tables().exp; }
const uint16_t* host_log() { return tables().log; }
const uint16_t* host_rowmask() { return tables().rowmask; }
const uint32_t* host_crc() { return tables().crc; }
uint16_t host_input_log(uint16_t input) { return tables().in_log[input & 32767]; }
uint16_t host_coeff_log(uint16_t input_log, uint16_t exponent) {
	uint32_t r = (uint32_t)input_log * exponent;
	r = (r >> 16) + (r & 0xFFFF);
	r += r >> 16;
	r &= 0xFFFF;
	return (uint16_t)(r == kOrder ? 0 : r);
}

// ---------------------------------------------------------------- device tables
static std::mutex g_dev_mu;
static DeviceTables g_dev[64];

int device_tables(DeviceTables** out) {
	int dev = 0;
	PPH_HIP(hipGetDevice(&dev));
	if(dev < 0 || dev >= 64) return PPH_E_UNSUPPORTED;
	std::lock_guard<std::mutex> lk(g_dev_mu);
	DeviceTables& t = g_dev[dev];
	if(!t.exp) {
		const HostTables& h = tables();
		uint16_t *e = nullptr, *l = nullptr, *rm = nullptr;
		uint32_t* cr = nullptr;
		PPH_HIP(hipMalloc(&e, sizeof(h.exp)));
		if(hipMalloc(&l, sizeof(h.log)) != hipSuccess || hipMalloc(&rm, sizeof(h.rowmask)) != hipSuccess ||
		   hipMalloc(&cr, sizeof(h.crc)) != hipSuccess) {
			(void)hipFree(e);
			if(l) (void)hipFree(l);
			if(rm) (void)hipFree(rm);
			return PPH_E_NOMEM;
		}
		PPH_HIP(hipMemcpy(cr, h.crc, sizeof(h.crc), hipMemcpyHostToDevice));
		t.crc = cr;
		PPH_HIP(hipMemcpy(e, h.exp, sizeof(h.exp), hipMemcpyHostToDevice));
		PPH_HIP(hipMemcpy(l, h.log, sizeof(h.log), hipMemcpyHostToDevice));
		PPH_HIP(hipMemcpy(rm, h.rowmask, sizeof(h.rowmask), hipMemcpyHostToDevice));
		t.rowmask = rm;
		hipDeviceProp_t prop;
		PPH_HIP(hipGetDeviceProperties(&prop, dev));
		t.cus = prop.multiProcessorCount;
		t.log = l;
		t.exp = e;
	}
	*out = &t;
	return PPH_OK;
}

}  // namespace pph

using namespace pph;

// ====================================================================== misc ABI

extern "C" const char* pph_strerror(int status) {
	switch(status) {
		case PPH_OK: return "ok";
		case PPH_E_INVALID: return "invalid argument";
		case PPH_E_NOMEM: return "out of memory";
		case PPH_E_STATE: return "call out of sequence";
		case PPH_E_DEVICE: return "HIP runtime error";
		case PPH_E_UNSUPPORTED: return "unsupported (no HIP device or configuration)";
		default: return "unknown status";
	}
}
extern "C" int pph_last_hip_error(void) { return g_last_hip; }

extern "C" int pph_device_count(void) {
	int n = 0;
	if(hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
}
extern "C" int pph_device_info(int device, char* name, size_t name_len, int* cu_count) {
	hipDeviceProp_t prop;
	PPH_HIP(hipGetDeviceProperties(&prop, device));
	if(name && name_len) {
		snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
	}
	if(cu_count) *cu_count = prop.multiProcessorCount;
	return PPH_OK;
}

extern "C" int pph_device_memory(int device, size_t* total_bytes) {
	if(!total_bytes) return PPH_E_INVALID;
	hipDeviceProp_t prop;
	PPH_HIP(hipGetDeviceProperties(&prop, device));
	*total_bytes = prop.totalGlobalMem;
	return PPH_OK;
}

extern "C" uint16_t pph_input_log(uint16_t input_num) { return host_input_log(input_num); }
extern "C" uint16_t pph_coeff_log(uint16_t input_log, uint16_t exponent) {
	return host_coeff_log(input_log, exponent);
}
extern "C" uint16_t pph_coeff(uint16_t input_num, uint16_t exponent) {
	return host_exp()[host_coeff_log(host_input_log(input_num), exponent)];
}
extern "C" uint16_t pph_gf_mul(uint16_t a, uint16_t b) {
	if(!a || !b) return 0;
	uint32_t s = (uint32_t)host_log()[a] + host_log()[b];
	if(s >= kOrder) s -= kOrder;
	return host_exp()[s];
}

extern "C" int pph_slot_layout(size_t slice_size, unsigned cksum_size, size_t* ck_off, size_t* slot_bytes) {
	if(!valid_ck(cksum_size)) return PPH_E_INVALID;
	if(ck_off) *ck_off = ck_offset(slice_size);
	if(slot_bytes) *slot_bytes = slot_size(slice_size, cksum_size);
	return PPH_OK;
}

// ====================================================================== device layer

extern "C" int pph_set_device(int device) {
	PPH_HIP(hipSetDevice(device));
	return PPH_OK;
}
extern "C" int pph_mem_get_info(size_t* free_bytes, size_t* total_bytes) {
	if(!free_bytes || !total_bytes) return PPH_E_INVALID;
	PPH_HIP(hipMemGetInfo(free_bytes, total_bytes));
	return PPH_OK;
}
extern "C" int pph_malloc(void** ptr, size_t bytes) {
	if(!ptr) return PPH_E_INVALID;
	hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
	if(e == hipErrorOutOfMemory) return PPH_E_NOMEM;
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}
extern "C" int pph_free(void* ptr) {
	if(ptr) PPH_HIP(hipFree(ptr));
	return PPH_OK;
}
extern "C" int pph_memcpy_h2d(void* dst, const void* src, size_t bytes) {
	if(bytes) PPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
	return PPH_OK;
}
extern "C" int pph_memcpy_d2h(void* dst, const void* src, size_t bytes) {
	if(bytes) PPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
	return PPH_OK;
}
extern "C" int pph_copy(void* dst, const void* src, size_t bytes, void* stream) {
	if(!dst || !src || bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return PPH_E_INVALID;
	PPH_HIP(launch_copy(dst, src, bytes, (hipStream_t)stream));
	return PPH_OK;
}
extern "C" int pph_lds_gather_ceiling(unsigned rounds, double* lookups_per_s, float* ms) {
	if(!rounds || !lookups_per_s) return PPH_E_INVALID;
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	const unsigned blocks = (unsigned)(t->cus > 0 ? t->cus : 256);
	uint32_t* sink = nullptr;
	PPH_HIP(hipMalloc((void**)&sink, (size_t)blocks * kPlainThreads * sizeof(uint32_t)));
	hipEvent_t a = nullptr, b = nullptr;
	hipError_t e = hipEventCreate(&a);
	if(e == hipSuccess) e = hipEventCreate(&b);
	// a first launch warms the code object and the table; the second is timed
	if(e == hipSuccess) e = launch_lds_gather_ceiling(t->exp, blocks, rounds, 12345u, sink, nullptr);
	if(e == hipSuccess) e = hipEventRecord(a, nullptr);
	if(e == hipSuccess) e = launch_lds_gather_ceiling(t->exp, blocks, rounds, 777u, sink, nullptr);
	if(e == hipSuccess) e = hipEventRecord(b, nullptr);
	if(e == hipSuccess) e = hipEventSynchronize(b);
	float t_ms = 0;
	if(e == hipSuccess) e = hipEventElapsedTime(&t_ms, a, b);
	if(a) (void)hipEventDestroy(a);
	if(b) (void)hipEventDestroy(b);
	(void)hipFree(sink);
	if(e != hipSuccess) return set_error(e);
	*lookups_per_s = (double)blocks * kPlainThreads * 16.0 * rounds / (t_ms * 1e-3);
	if(ms) *ms = t_ms;
	return PPH_OK;
}
extern "C" int pph_memset(void* dst, int value, size_t bytes) {
	if(bytes) PPH_HIP(hipMemset(dst, value, bytes));
	return PPH_OK;
}
extern "C" int pph_synchronize(void) {
	PPH_HIP(hipDeviceSynchronize());
	return PPH_OK;
}
extern "C" int pph_stream_create(void** stream) {
	if(!stream) return PPH_E_INVALID;
	hipStream_t s;
	PPH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
	*stream = (void*)s;
	return PPH_OK;
}
extern "C" int pph_stream_destroy(void* stream) {
	if(stream) PPH_HIP(hipStreamDestroy((hipStream_t)stream));
	return PPH_OK;
}
extern "C" int pph_stream_synchronize(void* stream) {
	PPH_HIP(hipStreamSynchronize((hipStream_t)stream));
	return PPH_OK;
}

struct pph_timer_t {
	hipEvent_t a, b;
};
extern "C" int pph_timer_create(void** timer) {
	if(!timer) return PPH_E_INVALID;
	auto* t = new(std::nothrow) pph_timer_t;
	if(!t) return PPH_E_NOMEM;
	if(hipEventCreate(&t->a) != hipSuccess || hipEventCreate(&t->b) != hipSuccess) {
		delete t;
		return set_error(hipErrorUnknown);
	}
	*timer = t;
	return PPH_OK;
}
extern "C" int pph_timer_destroy(void* timer) {
	auto* t = (pph_timer_t*)timer;
	if(!t) return PPH_OK;
	(void)hipEventDestroy(t->a);
	(void)hipEventDestroy(t->b);
	delete t;
	return PPH_OK;
}
extern "C" int pph_timer_start(void* timer, void* stream) {
	PPH_HIP(hipEventRecord(((pph_timer_t*)timer)->a, (hipStream_t)stream));
	return PPH_OK;
}
extern "C" int pph_timer_stop(void* timer, void* stream) {
	PPH_HIP(hipEventRecord(((pph_timer_t*)timer)->b, (hipStream_t)stream));
	return PPH_OK;
}
extern "C" int pph_timer_elapsed_ms(void* timer, float* ms) {
	auto* t = (pph_timer_t*)timer;
	PPH_HIP(hipEventSynchronize(t->b));
	PPH_HIP(hipEventElapsedTime(ms, t->a, t->b));
	return PPH_OK;
}

extern "C" int pph_fill_random(void* slots, size_t slot_bytes, unsigned count, size_t len, uint64_t seed,
                               uint32_t job, uint32_t first_slice, void* stream) {
	if(!slots || slot_bytes % 16 || len > slot_bytes) return PPH_E_INVALID;
	PPH_HIP(launch_fill_random((uint8_t*)slots, slot_bytes, count, len, seed, job, first_slice, (hipStream_t)stream));
	return PPH_OK;
}

extern "C" int pph_prep_slots_from(void* dst, const void* src, size_t slot_bytes, unsigned count,
                                   const uint32_t* lens, size_t slice_size, unsigned cksum_size, int to_log,
                                   void* stream) {
	if(!dst || !valid_ck(cksum_size) || slot_bytes < slot_size(slice_size, cksum_size) || slot_bytes % kSlotAlign)
		return PPH_E_INVALID;
	if(src && src != dst) {
		// distinct source and destination slot arrays must not overlap
		const uint8_t *a = (const uint8_t*)src, *b = (const uint8_t*)dst;
		const size_t n = slot_bytes * count;
		if(a < b + n && b < a + n) return PPH_E_INVALID;
	}
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	PPH_HIP(launch_prep((uint8_t*)dst, (const uint8_t*)src, slot_bytes, count, lens, nullptr, slice_size,
	                    ck_offset(slice_size), cksum_size, to_log != 0, *t, (hipStream_t)stream));
	return PPH_OK;
}

extern "C" int pph_prep_slots(void* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                              size_t slice_size, unsigned cksum_size, int to_log, void* stream) {
	return pph_prep_slots_from(slots, nullptr, slot_bytes, count, lens, slice_size, cksum_size, to_log, stream);
}

extern "C" int pph_muladd_plain(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                                const uint16_t* exps, const uint32_t* coef_log2, unsigned coef_stride,
                                void* out_slots, unsigned n_out, int accumulate, unsigned group, void* stream) {
	if((!in_slots && n_in) || !out_slots || slot_bytes % kSlotAlign) return PPH_E_INVALID;
	if(group && group != 1 && group != 2 && group != 4 && group != 8 && group != 16) return PPH_E_INVALID;
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	hipError_t e = launch_muladd_plain((const uint16_t*)in_slots, slot_bytes / 2, n_in, input_logs, exps, coef_log2,
	                                   coef_stride, (uint16_t*)out_slots, n_out, accumulate != 0, group, *t,
	                                   (hipStream_t)stream);
	if(e == hipErrorInvalidValue) return PPH_E_INVALID;
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}

extern "C" int pph_muladd_dense(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                                const uint16_t* exps, void* out_slots, unsigned n_out, int accumulate, void* stream) {
	if((!in_slots && n_in) || !out_slots || slot_bytes % kSlotAlign) return PPH_E_INVALID;
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	hipError_t e = launch_muladd_dense((const uint16_t*)in_slots, slot_bytes / 2, n_in, input_logs, exps,
	                                   (uint16_t*)out_slots, n_out, accumulate != 0, *t, (hipStream_t)stream);
	if(e == hipErrorInvalidValue) return PPH_E_INVALID;
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}

extern "C" int pph_slice_hash(const void* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                              size_t slice_size, uint8_t* md5_out, uint32_t* crc_out, void* stream) {
	if((!slots && count) || (count && (!md5_out || !crc_out))) return PPH_E_INVALID;
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	hipError_t e = launch_slice_hash((const uint8_t*)slots, slot_bytes, count, lens, slice_size, md5_out, crc_out, *t,
	                                 (hipStream_t)stream);
	if(e == hipErrorInvalidValue) return PPH_E_INVALID;
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}

extern "C" int pph_slice_hash_ifsc(const void* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                                   size_t slice_size, uint8_t* ifsc_out, void* stream) {
	if((!slots || !ifsc_out) && count) return PPH_E_INVALID;
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	hipError_t e = launch_slice_hash_to((const uint8_t*)slots, slot_bytes, count, lens, slice_size, ifsc_out, 20,
	                                    ifsc_out + 16, 20, *t, (hipStream_t)stream);
	if(e == hipErrorInvalidValue) return PPH_E_INVALID;
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}

// ---------------------------------------------------------------- recovery-packet MD5

struct pph_packet_md5 {
	int device = 0;
	unsigned regions = 0;
	uint32_t* d_state = nullptr;  // 20 words per region
	uint16_t* d_exps = nullptr;
	uint32_t head[8] = {};
	uint64_t done = 0;  // message bytes of each region so far
	bool begun = false, first = false;
};

extern "C" int pph_packet_md5_create(unsigned regions, pph_packet_md5** out) {
	if(!out || !regions || regions > 65535) return PPH_E_INVALID;
	*out = nullptr;
	auto* h = new(std::nothrow) pph_packet_md5;
	if(!h) return PPH_E_NOMEM;
	h->regions = regions;
	if(hipGetDevice(&h->device) != hipSuccess || hipMalloc((void**)&h->d_state, (size_t)regions * 80) != hipSuccess ||
	   hipMalloc((void**)&h->d_exps, (size_t)regions * 2) != hipSuccess) {
		(void)hipGetLastError();
		pph_packet_md5_destroy(h);
		return PPH_E_NOMEM;
	}
	*out = h;
	return PPH_OK;
}

extern "C" void pph_packet_md5_destroy(pph_packet_md5* h) {
	if(!h) return;
	if(h->d_state) (void)hipFree(h->d_state);
	if(h->d_exps) (void)hipFree(h->d_exps);
	delete h;
}

extern "C" int pph_packet_md5_begin(pph_packet_md5* h, const uint8_t* set_id, const uint16_t* exps, void* stream) {
	if(!h || !set_id || !exps) return PPH_E_INVALID;
	for(unsigned r = 0; r < h->regions; r++)
		if(exps[r] == 65535) return PPH_E_INVALID;
	// the host array is read before the call returns
	PPH_HIP(hipMemcpyAsync(h->d_exps, exps, (size_t)h->regions * 2, hipMemcpyHostToDevice, (hipStream_t)stream));
	PPH_HIP(hipStreamSynchronize((hipStream_t)stream));
	static const char kType[16] = {'P', 'A', 'R', ' ', '2', '.', '0', '\0', 'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
	memcpy(h->head, set_id, 16);
	memcpy(h->head + 4, kType, 16);
	h->done = 36;
	h->begun = h->first = true;
	return PPH_OK;
}

namespace {
int packet_md5_step(pph_packet_md5* h, const uint8_t* slots, size_t slot_bytes, size_t len, uint8_t* digest,
                    hipStream_t s) {
	if(!h->begun) return PPH_E_STATE;
	hipError_t e = launch_packet_md5(h->d_state, slots, slot_bytes, h->regions, h->d_exps, h->head,
	                                 h->first ? 0 : h->done, len, h->first, digest != nullptr, digest, s);
	if(e == hipErrorInvalidValue) return PPH_E_INVALID;
	if(e != hipSuccess) return set_error(e);
	h->done += len;
	h->first = false;
	if(digest) h->begun = false;
	return PPH_OK;
}
}  // namespace

extern "C" int pph_packet_md5_update(pph_packet_md5* h, const void* slots, size_t slot_bytes, size_t len,
                                     void* stream) {
	if(!h) return PPH_E_INVALID;
	return packet_md5_step(h, (const uint8_t*)slots, slot_bytes, len, nullptr, (hipStream_t)stream);
}

extern "C" int pph_packet_md5_final(pph_packet_md5* h, uint8_t* md5_out, void* stream) {
	if(!h || !md5_out) return PPH_E_INVALID;
	return packet_md5_step(h, nullptr, 0, 0, md5_out, (hipStream_t)stream);
}

extern "C" int pph_file_md5(const void* slots, size_t slot_bytes, unsigned count, size_t slice_size,
                            uint64_t file_size, uint8_t* md5_out, void* stream) {
	if(!slots || !md5_out) return PPH_E_INVALID;
	hipError_t e = launch_file_md5((const uint8_t*)slots, slot_bytes, count, slice_size, file_size, md5_out,
	                               (hipStream_t)stream);
	if(e == hipErrorInvalidValue) return PPH_E_INVALID;
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}

extern "C" int pph_verify_slots(const void* slots, size_t slot_bytes, unsigned count, size_t slice_size,
                                unsigned cksum_size, int* valid, void* stream) {
	if(!slots || !valid || !valid_ck(cksum_size) || slot_bytes < slot_size(slice_size, cksum_size) ||
	   slot_bytes % kSlotAlign)
		return PPH_E_INVALID;
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	PPH_HIP(launch_verify((const uint8_t*)slots, slot_bytes, count, slice_size, ck_offset(slice_size), cksum_size,
	                      valid, *t, (hipStream_t)stream));
	return PPH_OK;
}

// ====================================================================== backend

// One staging area: `batch` device input slots, their pinned host bounce
// buffers and per-slot coefficient data.  Active from submit until the
// mul-add that consumes it has completed (the role of IPAR2ProcStaging's
// isActive, gf16/controller.h:45-75).
struct Area {
	uint8_t* d_in = nullptr;        // batch * slot_bytes
	uint8_t* h_stage = nullptr;     // pinned, batch * slot_bytes (the device slots' layout)
	uint32_t* h_meta = nullptr;     // pinned: lens[batch] | slot idx[batch] | in_log[batch] (u16 packed)
	uint32_t* d_meta = nullptr;
	uint32_t* h_coef = nullptr;     // pinned explicit table batch * r_pad (or null)
	uint32_t* d_coef = nullptr;
	hipEvent_t done = nullptr;
	bool active = false;
	unsigned h2d_first = 0, h2d_count = 0;  // staged slots whose copy to the device is not issued yet
	bool explicit_coef = false;
	unsigned prep_count = 0;        // slots added since last submit needing preparation
	int filled_form = -1;           // fillInput slots: -1 none, else 1 prepared for the bilinear form, 0 plain
};

struct pph_proc {
	int device = 0;
	hipStream_t s_copy = nullptr, s_comp = nullptr;
	hipEvent_t ev_copied = nullptr;
	DeviceTables* tab = nullptr;
	unsigned n_areas = 2, batch = 0, min_batch = 0;
	unsigned ck = 64;
	int form = PPH_FORM_AUTO;
	bool inited = false;
	size_t slice_size = 0;       // allocated (setSliceSize)
	size_t cur_size = 0;         // current (setCurrentSliceSize)
	size_t slot_bytes = 0, alloc_slot_bytes = 0, alloc_slice = 0;
	std::vector<Area> areas;
	std::vector<uint16_t> exps;
	unsigned r_pad = 0, r_alloc = 0;
	uint8_t* d_out = nullptr;
	uint16_t* d_exps = nullptr;
	int failed = PPH_OK;   // first error of a batch submission: its inputs are lost, so the
	                       // recovery slices are incomplete until the next one starts afresh
	unsigned cur_area = 0, cur_inputs = 0;
	unsigned active_count = 0;
	unsigned batches_started = 0;
	bool processing_add = false;
	std::unique_ptr<pph::CopyPool> copier;  // host copies into and out of pinned memory
	uint8_t* h_bounce[2] = {nullptr, nullptr};  // pinned, one slot each: getOutput's device-to-host copies
	size_t bounce_bytes = 0;
	std::vector<uint8_t> h_pageable;  // getOutput's bounce where pinned memory is not available
	int pf_index = -1;     // recovery slice already copied (or in flight on s_comp) into h_bounce[pf_buf]
	unsigned pf_buf = 0;
	size_t dbg_flip_off = 0;  // test hook: one byte of the next bounced slice flipped on the host
	uint8_t dbg_flip_mask = 0;
	int dbg_fail = PPH_OK;    // test hook: the next batch submission fails with this status

	// Asynchronous entry points: a worker thread waits for submitted batches
	// and runs queued getOutputs, in the order they were posted, and calls the
	// caller's completion functions from there.
	struct Task {
		hipEvent_t ev = nullptr;        // batch: recorded after its mul-add (null: failed submission)
		int status = PPH_OK;
		pph_done_fn done = nullptr;
		pph_output_fn out_done = nullptr;  // set for a getOutput task
		void* user = nullptr;
		unsigned index = 0;
		void* output = nullptr;
	};
	pph::TaskQueue<Task> tasks;
	std::mutex wmu;                       // the worker's start and the event pool
	std::vector<hipEvent_t> free_events;
	std::thread worker;
};

namespace {

// Default input batch: 256 inputs, one bilinear k-block, while a staging
// area stays within kDefaultAreaBytes (fewer inputs for large slices).  From
// host buffers, 2000 x 1 MiB with 200 recovery slices took 51.6 ms per pass
// in 256-input batches (the bilinear form) against 59.4 ms in 64-input ones;
// with 20 or 64 recovery slices both took 41-44 ms.
constexpr unsigned kDefaultBatch = 256;
constexpr size_t kDefaultAreaBytes = 512ull << 20;
// PPH_FORM_AUTO takes the bilinear form for numbered batches of at least this
// many inputs (the bilinear GEMM pads inputs to a multiple of 256) with at
// least this many recovery slices.  2000 inputs x 1 MiB with 200 recovery
// slices run in 11.7 ms that way against 51.6 ms in the plain kernel, and a
// pass from host buffers in 256-input batches in 48 ms against 60 ms in
// 64-input plain batches; below about 32 recovery slices the plane generation
// and the row padding cost more than they save.
constexpr unsigned kAutoBilinearBatch = 256, kAutoBilinearRecovery = 64;

bool bilinear_batches(const pph_proc* p) {
	if(p->exps.empty()) return false;
	if(p->form == PPH_FORM_BILINEAR) return true;
	return p->form == PPH_FORM_AUTO && p->batch >= kAutoBilinearBatch && p->exps.size() >= kAutoBilinearRecovery;
}

void meta_ptrs(pph_proc* p, Area& a, uint32_t** lens, uint32_t** slots, uint16_t** logs) {
	*lens = a.h_meta;
	*slots = a.h_meta + p->batch;
	*logs = reinterpret_cast<uint16_t*>(a.h_meta + 2 * p->batch);
}

void free_areas(pph_proc* p) {
	for(auto& a : p->areas) {
		if(a.done) (void)hipEventSynchronize(a.done);
		if(a.d_in) (void)hipFree(a.d_in);
		if(a.h_stage) (void)hipHostFree(a.h_stage);
		if(a.h_meta) (void)hipHostFree(a.h_meta);
		if(a.d_meta) (void)hipFree(a.d_meta);
		if(a.h_coef) (void)hipHostFree(a.h_coef);
		if(a.d_coef) (void)hipFree(a.d_coef);
		if(a.done) (void)hipEventDestroy(a.done);
		a = Area();
	}
}

void free_outputs(pph_proc* p) {
	if(p->d_out) (void)hipFree(p->d_out);
	if(p->d_exps) (void)hipFree(p->d_exps);
	p->d_out = nullptr;
	p->d_exps = nullptr;
	p->r_alloc = 0;
	p->pf_index = -1;
}

// refresh isActive flags from the device
void poll_areas(const pph_proc* cp) {
	pph_proc* p = const_cast<pph_proc*>(cp);
	for(auto& a : p->areas) {
		if(a.active && hipEventQuery(a.done) == hipSuccess) {
			a.active = false;
			p->active_count--;
		}
	}
}

int wait_area(pph_proc* p, Area& a) {
	if(a.active) {
		PPH_HIP(hipEventSynchronize(a.done));
		a.active = false;
		p->active_count--;
	}
	return PPH_OK;
}

int alloc_areas(pph_proc* p) {
	free_areas(p);
	p->areas.resize(p->n_areas);
	for(auto& a : p->areas) {
		if(hipMalloc(&a.d_in, (size_t)p->batch * p->slot_bytes) != hipSuccess) return PPH_E_NOMEM;
		if(hipHostMalloc(&a.h_stage, (size_t)p->batch * p->slot_bytes, hipHostMallocDefault) != hipSuccess)
			return PPH_E_NOMEM;
		const size_t meta = (size_t)p->batch * 3 * sizeof(uint32_t);
		if(hipHostMalloc(&a.h_meta, meta, hipHostMallocDefault) != hipSuccess) return PPH_E_NOMEM;
		if(hipMalloc(&a.d_meta, meta) != hipSuccess) return PPH_E_NOMEM;
		PPH_HIP(hipEventCreateWithFlags(&a.done, hipEventDisableTiming));
	}
	p->alloc_slot_bytes = p->slot_bytes;
	p->alloc_slice = p->slice_size;
	return PPH_OK;
}

int ensure_coef_tables(pph_proc* p) {
	for(auto& a : p->areas) {
		if(a.h_coef) (void)hipHostFree(a.h_coef);
		if(a.d_coef) (void)hipFree(a.d_coef);
		a.h_coef = nullptr;
		a.d_coef = nullptr;
		if(!p->r_pad) continue;
		const size_t bytes = (size_t)p->batch * p->r_pad * sizeof(uint32_t);
		if(hipHostMalloc(&a.h_coef, bytes, hipHostMallocDefault) != hipSuccess) return PPH_E_NOMEM;
		if(hipMalloc(&a.d_coef, bytes) != hipSuccess) return PPH_E_NOMEM;
	}
	return PPH_OK;
}

// submit the current area's `n` inputs (PAR2ProcOCL::run_kernel, controller_ocl.cpp:496-527)
// Staged slots go to the device in runs of up to kH2DRun slots, one contiguous
// copy each (the staging has the device slots' layout; bytes past a slice's
// length are zeroed by the preparation): one DMA per 1 MiB slice runs well
// below the link rate.
constexpr unsigned kH2DRun = 16;

int flush_h2d(pph_proc* p, Area& a) {
	if(!a.h2d_count) return PPH_OK;
	const size_t off = (size_t)a.h2d_first * p->slot_bytes;
	PPH_HIP(hipMemcpyAsync(a.d_in + off, a.h_stage + off, (size_t)a.h2d_count * p->slot_bytes, hipMemcpyHostToDevice,
	                       p->s_copy));
	a.h2d_count = 0;
	return PPH_OK;
}

int submit_area(pph_proc* p, unsigned area_idx, unsigned n) {
	Area& a = p->areas[area_idx];
	if(p->dbg_fail) {
		const int rc = p->dbg_fail;
		p->dbg_fail = PPH_OK;
		return rc;
	}
	if(int rc = flush_h2d(p, a)) return rc;
	uint32_t *lens, *slots;
	uint16_t* logs;
	meta_ptrs(p, a, &lens, &slots, &logs);
	const size_t meta = (size_t)p->batch * 3 * sizeof(uint32_t);
	// the bilinear form takes data-form inputs (numbered or explicit
	// coefficients alike); a batch whose bilinear workspace cannot be
	// allocated runs the plain form (the same bytes, more time)
	bool bl = bilinear_batches(p);
	if(bl && a.filled_form != 1) {
		const int rc = bilinear_reserve(p->slot_bytes, n, (unsigned)p->exps.size());
		if(rc == PPH_E_NOMEM)
			bl = false;
		else if(rc)
			return rc;
	}
	if(a.filled_form >= 0 && !p->exps.empty() && a.filled_form != (bl ? 1 : 0))
		return PPH_E_STATE;  // fillInput prepared this area's slots for the other form
	PPH_HIP(hipMemcpyAsync(a.d_meta, a.h_meta, meta, hipMemcpyHostToDevice, p->s_copy));
	if(a.prep_count) {
		PPH_HIP(launch_prep(a.d_in, nullptr, p->slot_bytes, a.prep_count, a.d_meta, a.d_meta + p->batch, p->cur_size,
		                    ck_offset(p->cur_size), p->ck, !bl, *p->tab, p->s_copy));
		a.prep_count = 0;
	}
	if(a.explicit_coef) {
		PPH_HIP(hipMemcpyAsync(a.d_coef, a.h_coef, (size_t)p->batch * p->r_pad * sizeof(uint32_t),
		                       hipMemcpyHostToDevice, p->s_copy));
	}
	PPH_HIP(hipEventRecord(p->ev_copied, p->s_copy));
	PPH_HIP(hipStreamWaitEvent(p->s_comp, p->ev_copied, 0));
	if(bl) {
		const uint16_t* d_logs = reinterpret_cast<const uint16_t*>(a.d_meta + 2 * p->batch);
		const int rc = bilinear_call(a.d_in, p->slot_bytes, n, d_logs, p->d_exps, a.explicit_coef ? a.d_coef : nullptr,
		                             p->r_pad, p->d_out, (unsigned)p->exps.size(), p->processing_add ? 1 : 0, p->s_comp);
		if(rc) return rc;
		p->processing_add = true;
	} else if(!p->exps.empty()) {
		const uint16_t* d_logs = reinterpret_cast<const uint16_t*>(a.d_meta + 2 * p->batch);
		hipError_t e = launch_muladd_plain(reinterpret_cast<const uint16_t*>(a.d_in), p->slot_bytes / 2, n, d_logs,
		                                   p->d_exps, a.explicit_coef ? a.d_coef : nullptr, p->r_pad, (uint16_t*)p->d_out,
		                                   (unsigned)p->exps.size(), p->processing_add, 0, *p->tab, p->s_comp);
		if(e != hipSuccess) return set_error(e);
		p->processing_add = true;
	}
	PPH_HIP(hipEventRecord(a.done, p->s_comp));
	a.active = true;
	a.explicit_coef = false;
	p->active_count++;
	p->batches_started++;
	p->pf_index = -1;
	return PPH_OK;
}

// Submit a batch; a failed submission loses its inputs, so the backend keeps
// the first such error and reports it from endInput and every getOutput until
// a new recovery computation starts (discardOutput, init, a grown slice).
int run_kernel(pph_proc* p, unsigned area_idx, unsigned n) {
	const int rc = submit_area(p, area_idx, n);
	if(rc == PPH_OK) return PPH_OK;
	Area& a = p->areas[area_idx];
	// whatever was enqueued for the area finishes before it is refilled
	(void)hipStreamSynchronize(p->s_copy);
	(void)hipStreamSynchronize(p->s_comp);
	(void)hipGetLastError();
	a.active = false;
	a.explicit_coef = false;
	a.prep_count = 0;
	a.h2d_count = 0;
	p->batches_started++;
	p->pf_index = -1;
	if(p->failed == PPH_OK) p->failed = rc;
	return rc;
}

void advance_area(pph_proc* p) {
	p->cur_inputs = 0;
	if(++p->cur_area == p->areas.size()) p->cur_area = 0;
}

bool should_submit(const pph_proc* p, bool flush) {
	return flush || p->cur_inputs == p->batch ||
	       (p->active_count == 0 && p->areas.size() > 1 && p->cur_inputs >= p->min_batch);
}

// write one input's coefficient data into the area (set_coeffs, controller_ocl.cpp:387-403)
void set_coeffs_num(pph_proc* p, Area& a, unsigned idx, uint16_t input_num) {
	uint32_t *lens, *slots;
	uint16_t* logs;
	meta_ptrs(p, a, &lens, &slots, &logs);
	const uint16_t lg = host_input_log(input_num);
	logs[idx] = lg;
	if(a.explicit_coef) {
		for(unsigned r = 0; r < p->r_pad; r++)
			a.h_coef[(size_t)idx * p->r_pad + r] =
			    r < p->exps.size() ? 2u * host_coeff_log(lg, p->exps[r]) : kSkipCoef;
	}
}

// switch an area to an explicit table, materialising the coefficients of the
// inputs already staged in it
void to_explicit(pph_proc* p, Area& a, unsigned upto) {
	if(a.explicit_coef) return;
	a.explicit_coef = true;
	uint32_t *lens, *slots;
	uint16_t* logs;
	meta_ptrs(p, a, &lens, &slots, &logs);
	for(unsigned k = 0; k < upto; k++)
		for(unsigned r = 0; r < p->r_pad; r++)
			a.h_coef[(size_t)k * p->r_pad + r] =
			    r < p->exps.size() ? 2u * host_coeff_log(logs[k], p->exps[r]) : kSkipCoef;
}

// Stage one input into slot idx of an area that is not active: its
// coefficient data, and (unless dummy) its bytes into the pinned staging,
// queued for preparation and sent to the device in contiguous runs.
int stage_slot(pph_proc* p, Area& a, unsigned idx, const void* buffer, size_t size, const uint16_t* coeffs,
               uint16_t input_num, bool dummy) {
	uint32_t *lens, *slots;
	uint16_t* logs;
	meta_ptrs(p, a, &lens, &slots, &logs);
	if(coeffs) {
		if(!a.h_coef) return PPH_E_STATE;
		to_explicit(p, a, idx);
		for(unsigned r = 0; r < p->r_pad; r++) {
			uint32_t v = kSkipCoef;
			if(r < p->exps.size() && coeffs[r]) v = 2u * host_log()[coeffs[r]];
			a.h_coef[(size_t)idx * p->r_pad + r] = v;
		}
		logs[idx] = 0;
	} else {
		set_coeffs_num(p, a, idx, input_num);
	}
	if(!dummy) {
		uint8_t* stage = a.h_stage + (size_t)idx * p->slot_bytes;
		if(size) p->copier->copy(stage, buffer, size);
		if(a.h2d_count && a.h2d_first + a.h2d_count != idx) {
			const int rc = flush_h2d(p, a);
			if(rc) return rc;
		}
		if(!a.h2d_count) a.h2d_first = idx;
		if(++a.h2d_count == kH2DRun) {
			const int rc = flush_h2d(p, a);
			if(rc) return rc;
		}
		lens[a.prep_count] = (uint32_t)size;
		slots[a.prep_count] = idx;
		a.prep_count++;
	}
	return PPH_OK;
}

int add_common(pph_proc* p, const void* buffer, size_t size, const uint16_t* coeffs, uint16_t input_num, int flush,
               bool dummy) {
	if(!p || !p->inited) return PPH_E_STATE;
	if(!dummy && size && !buffer) return PPH_E_INVALID;
	if(size > p->cur_size) return PPH_E_INVALID;
	Area& a = p->areas[p->cur_area];
	int rc = wait_area(p, a);
	if(rc) return rc;
	rc = stage_slot(p, a, p->cur_inputs, buffer, size, coeffs, input_num, dummy);
	if(rc) return rc;
	p->cur_inputs++;
	poll_areas(p);
	if(should_submit(p, flush != 0)) {
		const unsigned area = p->cur_area;
		rc = run_kernel(p, area, p->cur_inputs);
		advance_area(p);
		if(rc) return rc;
	}
	// the caller's buffer has been copied: it may be reused now
	return PPH_OK;
}

}  // namespace

extern "C" int pph_proc_create(int device, unsigned staging_areas, pph_proc** out) {
	if(!out) return PPH_E_INVALID;
	*out = nullptr;
	int n = 0;
	if(hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PPH_E_UNSUPPORTED;
	if(device < 0 || device >= n) return PPH_E_INVALID;
	auto* p = new(std::nothrow) pph_proc;
	if(!p) return PPH_E_NOMEM;
	p->device = device;
	p->n_areas = staging_areas ? staging_areas : 2;
	hipError_t e = hipSetDevice(device);
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&p->s_copy, hipStreamNonBlocking);
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&p->s_comp, hipStreamNonBlocking);
	if(e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_copied, hipEventDisableTiming);
	if(e != hipSuccess) {
		pph_proc_destroy(p);
		return set_error(e);
	}
	int rc = device_tables(&p->tab);
	if(rc) {
		pph_proc_destroy(p);
		return rc;
	}
	// helper threads for the staging copies (PPH_COPY_THREADS, default 3: with
	// the caller's thread, four streaming copies at once fill the host's copy
	// bandwidth into pinned memory; more threads contend and slow every copy)
	unsigned helpers = 3;
	if(const char* v = getenv("PPH_COPY_THREADS")) helpers = (unsigned)std::min(15l, std::max(0l, strtol(v, nullptr, 10)));
	p->copier.reset(new(std::nothrow) CopyPool(helpers));
	if(!p->copier) {
		pph_proc_destroy(p);
		return PPH_E_NOMEM;
	}
	*out = p;
	return PPH_OK;
}

extern "C" void pph_proc_destroy(pph_proc* p) {
	if(!p) return;
	// the worker runs every queued task (its completion functions included) first
	p->tasks.close();
	if(p->worker.joinable()) p->worker.join();
	(void)hipSetDevice(p->device);
	for(hipEvent_t e : p->free_events) (void)hipEventDestroy(e);
	if(p->s_comp) (void)hipStreamSynchronize(p->s_comp);
	if(p->s_copy) (void)hipStreamSynchronize(p->s_copy);
	free_areas(p);
	free_outputs(p);
	if(p->ev_copied) (void)hipEventDestroy(p->ev_copied);
	if(p->s_copy) (void)hipStreamDestroy(p->s_copy);
	if(p->s_comp) (void)hipStreamDestroy(p->s_comp);
	for(uint8_t* b : p->h_bounce)
		if(b) (void)hipHostFree(b);
	delete p;
}

extern "C" int pph_proc_set_slice_size(pph_proc* p, size_t slice_size) {
	if(!p) return PPH_E_INVALID;
	if(slice_size > 0xFFFFFFFFull) return PPH_E_INVALID;
	p->slice_size = slice_size;
	if(!p->inited) p->cur_size = slice_size;
	return PPH_OK;
}

extern "C" int pph_proc_init(pph_proc* p, unsigned cksum_size, unsigned input_batch, int form) {
	if(!p || !valid_ck(cksum_size)) return PPH_E_INVALID;
	if(form != PPH_FORM_AUTO && form != PPH_FORM_PLAIN && form != PPH_FORM_BILINEAR) return PPH_E_UNSUPPORTED;
	PPH_HIP(hipSetDevice(p->device));
	if(p->s_comp) PPH_HIP(hipStreamSynchronize(p->s_comp));
	p->ck = cksum_size;
	p->form = form;
	p->batch = input_batch ? input_batch : kDefaultBatch;
	p->min_batch = p->batch;
	p->cur_size = p->slice_size;
	p->slot_bytes = slot_size(p->cur_size, p->ck);
	if(!input_batch)
		p->batch = (unsigned)std::max<size_t>(1, std::min<size_t>(kDefaultBatch, kDefaultAreaBytes / p->slot_bytes));
	p->min_batch = p->batch;
	p->exps.clear();
	free_outputs(p);
	int rc = alloc_areas(p);
	if(rc) return rc;
	p->cur_area = p->cur_inputs = 0;
	p->active_count = 0;
	p->batches_started = 0;
	p->processing_add = false;
	p->failed = PPH_OK;
	p->pf_index = -1;
	p->r_pad = 0;
	p->inited = true;
	return PPH_OK;
}

extern "C" int pph_proc_set_current_slice_size(pph_proc* p, size_t slice_size) {
	if(!p || !p->inited) return PPH_E_STATE;
	if(slice_size > 0xFFFFFFFFull) return PPH_E_INVALID;
	// only between passes: every staged batch must have finished
	if(p->cur_inputs) return PPH_E_STATE;
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	poll_areas(p);
	const size_t slot = slot_size(slice_size, p->ck);
	p->cur_size = slice_size;
	p->slot_bytes = slot;
	if(slot > p->alloc_slot_bytes || slice_size > p->alloc_slice) {
		// grow: PAR2ProcOCL re-inits (controller_ocl.cpp:211-222), and init
		// starts the recovery afresh (processingAdd = false, :147)
		p->slice_size = std::max(p->slice_size, slice_size);
		p->processing_add = false;
		p->failed = PPH_OK;
		p->pf_index = -1;
		p->cur_area = p->cur_inputs = 0;
		p->active_count = 0;
		int rc = alloc_areas(p);
		if(rc) return rc;
		rc = ensure_coef_tables(p);
		if(rc) return rc;
		if(p->r_alloc) {
			auto ex = p->exps;
			free_outputs(p);
			p->exps.clear();
			return pph_proc_set_recovery_slices(p, (unsigned)ex.size(), ex.data());
		}
	}
	return PPH_OK;
}

extern "C" int pph_proc_set_recovery_slices(pph_proc* p, unsigned n, const uint16_t* exponents) {
	if(!p || !p->inited) return PPH_E_STATE;
	if(n > 65535) return PPH_E_INVALID;
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	poll_areas(p);
	const std::vector<uint16_t> old_exps = p->exps;
	p->exps.assign(n, 1);  // default exponent 1 (controller_cpu.cpp:176)
	if(exponents) {
		for(unsigned r = 0; r < n; r++) {
			if(exponents[r] == 65535) return PPH_E_INVALID;  // gfmat_coeff_log's precondition
			p->exps[r] = exponents[r];
		}
	}
	if(!n) return PPH_OK;
	// slots prepared by fillInput keep their form: a count that moves PPH_FORM_AUTO
	// to the other form would run them through the wrong kernel
	const int now_form = bilinear_batches(p) ? 1 : 0;
	for(const Area& a : p->areas)
		if(a.filled_form >= 0 && a.filled_form != now_form) {
			p->exps = old_exps;
			return PPH_E_STATE;
		}
	const unsigned group = 16;
	const unsigned r_pad = (n + group - 1) / group * group;
	if(!p->d_out || r_pad > p->r_alloc || p->slot_bytes > p->alloc_slot_bytes) {
		free_outputs(p);
		if(hipMalloc(&p->d_out, (size_t)r_pad * p->alloc_slot_bytes) != hipSuccess) return PPH_E_NOMEM;
		if(hipMalloc(&p->d_exps, (size_t)r_pad * sizeof(uint16_t)) != hipSuccess) return PPH_E_NOMEM;
		p->r_alloc = r_pad;
	}
	if(r_pad != p->r_pad) {
		p->r_pad = r_pad;
		int rc = ensure_coef_tables(p);
		if(rc) return rc;
	}
	std::vector<uint16_t> ex(r_pad, 0);
	std::copy(p->exps.begin(), p->exps.end(), ex.begin());
	PPH_HIP(hipMemcpy(p->d_exps, ex.data(), r_pad * sizeof(uint16_t), hipMemcpyHostToDevice));
	p->pf_index = -1;
	return PPH_OK;
}

extern "C" int pph_proc_can_add(const pph_proc* p) {
	if(!p || !p->inited) return PPH_ADD_FULL;
	poll_areas(p);
	if(p->areas[p->cur_area].active) return PPH_ADD_FULL;
	return p->active_count < p->areas.size() - 1 ? PPH_ADD_OK : PPH_ADD_OK_BUSY;
}

extern "C" int pph_proc_wait_for_add(pph_proc* p) {
	if(!p || !p->inited) return PPH_E_STATE;
	return wait_area(p, p->areas[p->cur_area]);
}

extern "C" int pph_proc_add_input(pph_proc* p, const void* buffer, size_t size, uint16_t input_num, int flush) {
	return add_common(p, buffer, size, nullptr, input_num, flush, false);
}
extern "C" int pph_proc_add_input_coeffs(pph_proc* p, const void* buffer, size_t size, const uint16_t* coeffs,
                                         int flush) {
	if(!coeffs) return PPH_E_INVALID;
	return add_common(p, buffer, size, coeffs, 0, flush, false);
}
extern "C" int pph_proc_dummy_input(pph_proc* p, uint16_t input_num, int flush) {
	return add_common(p, nullptr, 0, nullptr, input_num, flush, true);
}

extern "C" int pph_proc_fill_input(pph_proc* p, const void* buffer) {
	// PAR2ProcOCL::fillInput (controller_ocl.cpp:366-380): prepare the next slot of
	// the current area in place, without scheduling any processing
	if(!p || !p->inited || !buffer) return PPH_E_INVALID;
	Area& a = p->areas[p->cur_area];
	int rc = wait_area(p, a);
	if(rc) return rc;
	if((rc = flush_h2d(p, a))) return rc;
	uint8_t* stage = a.h_stage + (size_t)p->cur_inputs * p->slot_bytes;
	p->copier->copy(stage, buffer, p->cur_size);
	uint8_t* dst = a.d_in + (size_t)p->cur_inputs * p->slot_bytes;
	PPH_HIP(hipMemcpyAsync(dst, stage, p->cur_size, hipMemcpyHostToDevice, p->s_copy));
	const bool bl = bilinear_batches(p);
	if(a.filled_form >= 0 && a.filled_form != (bl ? 1 : 0)) return PPH_E_STATE;
	PPH_HIP(launch_prep(dst, nullptr, p->slot_bytes, 1, nullptr, nullptr, p->cur_size, ck_offset(p->cur_size), p->ck,
	                    !bl, *p->tab, p->s_copy));
	a.filled_form = bl ? 1 : 0;
	PPH_HIP(hipStreamSynchronize(p->s_copy));
	if(++p->cur_inputs == p->batch) {
		p->cur_inputs = 0;
		if(++p->cur_area == p->areas.size()) {
			p->cur_area = 0;
			return 1;
		}
	}
	return 0;
}

extern "C" int pph_proc_flush(pph_proc* p) {
	if(!p || !p->inited) return PPH_E_STATE;
	if(!p->cur_inputs) return PPH_OK;
	const unsigned area = p->cur_area;
	int rc = run_kernel(p, area, p->cur_inputs);
	advance_area(p);
	return rc;
}

extern "C" int pph_proc_end_input(pph_proc* p) {
	if(!p || !p->inited) return PPH_E_STATE;
	int rc = pph_proc_flush(p);
	if(rc) return rc;
	PPH_HIP(hipStreamSynchronize(p->s_copy));
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	poll_areas(p);
	return p->failed;
}

namespace {

// Pinned bounce slots for getOutput (the slice and its checksum block).
void ensure_bounce(pph_proc* p) {
	if(p->bounce_bytes >= p->alloc_slot_bytes) return;
	for(uint8_t*& b : p->h_bounce) {
		if(b) (void)hipHostFree(b);
		b = nullptr;
	}
	p->bounce_bytes = 0;
	p->pf_index = -1;
	bool ok = true;
	for(uint8_t*& b : p->h_bounce)
		ok = ok && hipHostMalloc((void**)&b, p->alloc_slot_bytes, hipHostMallocDefault) == hipSuccess;
	if(ok) p->bounce_bytes = p->alloc_slot_bytes;
	(void)hipGetLastError();
}

}  // namespace

extern "C" int pph_proc_get_output(pph_proc* p, unsigned index, void* output, int* valid) {
	if(valid) *valid = 0;
	if(!p || !p->inited) return PPH_E_STATE;
	if(index >= p->exps.size() || (!output && p->cur_size)) return PPH_E_INVALID;
	if(p->failed) return p->failed;
	if(!p->processing_add) {
		// nothing computed on this backend: zero fill (controller.cpp:376-382)
		if(p->cur_size) memset(output, 0, p->cur_size);
		if(valid) *valid = 1;
		return PPH_OK;
	}
	// The slice and its checksum block come to host memory, and the checksum is
	// recomputed over the bytes as they are copied to the caller (the check of
	// copy_cksum_check, gf16/gf16_cksum_base.h:48-59, on the delivered bytes).
	const size_t ck_off = ck_offset(p->cur_size), span = ck_off + p->ck;
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	ensure_bounce(p);
	const uint8_t* src = p->d_out + (size_t)index * p->slot_bytes;
	const uint8_t* host;
	if(p->bounce_bytes >= span) {
		// pinned bounce slots (full-rate DMA); the next slice's device-to-host
		// copy runs during this slice's host copy
		const unsigned b = p->pf_buf;
		if(p->pf_index != (int)index) {  // (a prefetched slice was waited for above)
			PPH_HIP(hipMemcpyAsync(p->h_bounce[b], src, span, hipMemcpyDeviceToHost, p->s_comp));
			PPH_HIP(hipStreamSynchronize(p->s_comp));
		}
		p->pf_index = -1;
		if(index + 1 < p->exps.size()) {  // callers read the slices in order
			PPH_HIP(hipMemcpyAsync(p->h_bounce[b ^ 1u], src + p->slot_bytes, span, hipMemcpyDeviceToHost, p->s_comp));
			p->pf_index = (int)index + 1;
			p->pf_buf = b ^ 1u;
		}
		host = p->h_bounce[b];
	} else {
		p->h_pageable.resize(span);
		PPH_HIP(hipMemcpy(p->h_pageable.data(), src, span, hipMemcpyDeviceToHost));
		host = p->h_pageable.data();
	}
	if(p->dbg_flip_mask && p->dbg_flip_off < p->cur_size) {
		const_cast<uint8_t*>(host)[p->dbg_flip_off] ^= p->dbg_flip_mask;
		p->dbg_flip_mask = 0;
	}
	alignas(16) uint8_t h[64];
	p->copier->copy_cksum(output, host, p->cur_size, p->ck, h, host_exp(), host_log());
	if(valid) *valid = memcmp(h, host + ck_off, p->ck) == 0 ? 1 : 0;
	return PPH_OK;
}

extern "C" int pph_proc_packet_md5_update(pph_proc* p, pph_packet_md5* h) {
	if(!p || !h) return PPH_E_INVALID;
	if(!p->inited) return PPH_E_STATE;
	if(h->regions != p->exps.size() || h->device != p->device) return PPH_E_INVALID;
	// after every submitted batch (same stream); nothing computed reads as zeros.  The
	// backend's stream does not order with the caller's, so the call returns once done.
	const int rc =
	    packet_md5_step(h, p->processing_add ? p->d_out : nullptr, p->slot_bytes, p->cur_size, nullptr, p->s_comp);
	if(rc) return rc;
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	return PPH_OK;
}

extern "C" void pph_proc_discard_output(pph_proc* p) {
	if(!p) return;
	p->processing_add = false;
	p->failed = PPH_OK;
}
extern "C" int pph_proc_has_added(const pph_proc* p) { return p && p->processing_add ? 1 : 0; }
extern "C" int pph_proc_is_empty(const pph_proc* p) {
	if(!p) return 1;
	poll_areas(p);
	return p->active_count == 0 ? 1 : 0;
}
extern "C" int pph_proc_free_processing_mem(pph_proc* p) {
	if(!p) return PPH_E_INVALID;
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	poll_areas(p);
	free_outputs(p);
	p->exps.clear();
	p->processing_add = false;
	p->failed = PPH_OK;
	return PPH_OK;
}
extern "C" unsigned pph_proc_input_batch_size(const pph_proc* p) { return p ? p->batch : 0; }
extern "C" unsigned pph_proc_batches_started(const pph_proc* p) { return p ? p->batches_started : 0; }
extern "C" void pph_proc_set_min_input_batch_size(pph_proc* p, unsigned n) {
	if(p) p->min_batch = n ? std::min(n, p->batch) : p->batch;
}
extern "C" int pph_proc_num_recovery_slices(const pph_proc* p) { return p ? (int)p->exps.size() : 0; }

// ====================================================================== asynchronous backend

namespace {

void worker_main(pph_proc* p) {
	(void)hipSetDevice(p->device);
	pph_proc::Task t;
	while(p->tasks.pop(&t)) {
		if(t.out_done) {
			int valid = 0;
			const int rc = pph_proc_get_output(p, t.index, t.output, &valid);
			t.out_done(t.user, rc, valid);
		} else {
			int status = t.status;
			if(t.ev) {
				const hipError_t e = hipEventSynchronize(t.ev);
				if(e != hipSuccess && status == PPH_OK) status = set_error(e);
			}
			if(t.done) t.done(t.user, status);
			if(t.ev) {
				std::lock_guard<std::mutex> lk(p->wmu);
				p->free_events.push_back(t.ev);
			}
		}
		p->tasks.finished();
	}
}

int post_task(pph_proc* p, const pph_proc::Task& t) {
	{
		std::lock_guard<std::mutex> lk(p->wmu);
		if(!p->worker.joinable() && !p->tasks.closed()) {
			try {
				p->worker = std::thread(worker_main, p);
			} catch(...) {
				return PPH_E_NOMEM;
			}
		}
	}
	return p->tasks.push(t) ? PPH_OK : PPH_E_STATE;
}

hipEvent_t take_event(pph_proc* p) {
	{
		std::lock_guard<std::mutex> lk(p->wmu);
		if(!p->free_events.empty()) {
			hipEvent_t e = p->free_events.back();
			p->free_events.pop_back();
			return e;
		}
	}
	hipEvent_t e = nullptr;
	if(hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
	return e;
}

}  // namespace

extern "C" unsigned pph_proc_staging_areas(const pph_proc* p) { return p ? (unsigned)p->areas.size() : 0; }

extern "C" int pph_proc_stage_input(pph_proc* p, unsigned area, unsigned slot, const void* buffer, size_t size,
                                    uint16_t input_num, const uint16_t* coeffs) {
	if(!p || !p->inited) return PPH_E_STATE;
	if(area >= p->areas.size() || slot >= p->batch || size > p->cur_size || (size && !buffer)) return PPH_E_INVALID;
	Area& a = p->areas[area];
	const int rc = wait_area(p, a);
	if(rc) return rc;
	return stage_slot(p, a, slot, buffer, size, coeffs, input_num, false);
}

extern "C" int pph_proc_stage_dummy(pph_proc* p, unsigned area, unsigned slot, uint16_t input_num) {
	if(!p || !p->inited) return PPH_E_STATE;
	if(area >= p->areas.size() || slot >= p->batch) return PPH_E_INVALID;
	Area& a = p->areas[area];
	const int rc = wait_area(p, a);
	if(rc) return rc;
	return stage_slot(p, a, slot, nullptr, 0, nullptr, input_num, true);
}

extern "C" int pph_proc_fill_slot(pph_proc* p, unsigned area, unsigned slot, const void* buffer) {
	if(!p || !p->inited || !buffer) return PPH_E_INVALID;
	if(area >= p->areas.size() || slot >= p->batch) return PPH_E_INVALID;
	const unsigned ca = p->cur_area, ci = p->cur_inputs;
	p->cur_area = area;
	p->cur_inputs = slot;
	const int rc = pph_proc_fill_input(p, buffer);
	p->cur_area = ca;
	p->cur_inputs = ci;
	return rc < 0 ? rc : PPH_OK;
}

extern "C" int pph_proc_submit(pph_proc* p, unsigned area, unsigned n_inputs, int first, pph_done_fn done,
                               void* user) {
	pph_proc::Task t;
	t.done = done;
	t.user = user;
	if(!p || !p->inited) {
		t.status = PPH_E_STATE;
	} else if(area >= p->areas.size() || !n_inputs || n_inputs > p->batch) {
		t.status = PPH_E_INVALID;
	} else {
		if(first) pph_proc_discard_output(p);
		t.status = run_kernel(p, area, n_inputs);
		if(t.status == PPH_OK) {
			t.ev = take_event(p);
			if(!t.ev || hipEventRecord(t.ev, p->s_comp) != hipSuccess) {
				// completion cannot be observed through an event: wait here instead
				(void)hipStreamSynchronize(p->s_comp);
				if(t.ev) {
					std::lock_guard<std::mutex> lk(p->wmu);
					p->free_events.push_back(t.ev);
				}
				t.ev = nullptr;
			}
		}
	}
	// done runs exactly once: from the worker, or here when nothing can be posted
	const int rc = p ? post_task(p, t) : t.status;
	if(!p || rc != PPH_OK) {
		if(t.ev) (void)hipEventSynchronize(t.ev);
		if(done) done(user, rc != PPH_OK ? rc : t.status);
		return rc != PPH_OK ? rc : t.status;
	}
	return t.status;
}

extern "C" int pph_proc_get_output_async(pph_proc* p, unsigned index, void* output, pph_output_fn done, void* user) {
	if(!p || !p->inited || !done) return PPH_E_INVALID;
	pph_proc::Task t;
	t.out_done = done;
	t.user = user;
	t.index = index;
	t.output = output;
	return post_task(p, t);
}

extern "C" int pph_proc_drain(pph_proc* p) {
	if(!p) return PPH_E_INVALID;
	p->tasks.wait_idle();
	return PPH_OK;
}

extern "C" int pph_proc_debug_corrupt_output(pph_proc* p, unsigned index, size_t offset, uint8_t mask) {
	if(!p || !p->inited || index >= p->exps.size() || offset >= p->slot_bytes) return PPH_E_INVALID;
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	PPH_HIP(launch_xor_byte(p->d_out + (size_t)index * p->slot_bytes + offset, mask, p->s_comp));
	PPH_HIP(hipStreamSynchronize(p->s_comp));
	p->pf_index = -1;
	return PPH_OK;
}

extern "C" int pph_proc_debug_corrupt_host_copy(pph_proc* p, size_t offset, uint8_t mask) {
	if(!p || !p->inited || offset >= p->cur_size) return PPH_E_INVALID;
	p->dbg_flip_off = offset;
	p->dbg_flip_mask = mask;
	return PPH_OK;
}

extern "C" int pph_proc_debug_fail_next_batch(pph_proc* p, int status) {
	if(!p || !p->inited || status >= 0) return PPH_E_INVALID;
	p->dbg_fail = status;
	return PPH_OK;
}
