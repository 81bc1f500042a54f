// Internal declarations shared by the kernel and API translation units of
// libparpar_hip.so (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

struct pph_tuning;

namespace pph {

// GF(2^16) with the PAR2 polynomial x^16 + x^12 + x^3 + x + 1 (0x1100B),
// generator 2 (gf16/gf16_global.h:9-10, gf16/gfmat_coeff.c:12-19).
constexpr uint32_t kPoly = 0x1100B;
constexpr uint32_t kOrder = 65535;
// Byte address of the zero entry of the LDS exp table and the "no contribution"
// marker of explicit coefficient tables.
constexpr uint32_t kZeroIdx2 = 2u * 65535u;
constexpr uint32_t kSkipCoef = 0xFFFFFFFFu;
// Log-domain encoding of the zero word.
constexpr uint16_t kLogZero = 0xFFFF;
// Decoded log of zero: large enough that both candidates of the modular
// reduction exceed kZeroIdx2 for every coefficient.
constexpr uint32_t kLogZeroDecoded = 0x40000000u;

constexpr unsigned kSlotAlign = 1024;   // slot size granularity (bytes)
constexpr unsigned kPlainThreads = 1024;
constexpr unsigned kPlainWordsPerLane = 8;

struct DeviceTables {
	uint16_t* exp = nullptr;   // exp[k] = 2^k for k < 65535; exp[65535] = 0
	uint16_t* log = nullptr;   // log[x] = log2(x) for x != 0; log[0] = 0xFFFF
	uint16_t* rowmask = nullptr;  // [j][q][n]: row j of the bit-matrix of n * z^(4q), as a mask
	uint32_t* crc = nullptr;      // CRC-32 slice-by-4 tables [4][256]
	int cus = 0;
};

// Host-side tables (built once).
const uint16_t* host_exp();
const uint16_t* host_log();
const uint16_t* host_rowmask();  // 16 x 4 x 16 nibble tables of the coefficient bit-matrices
const uint32_t* host_crc();      // CRC-32 (reflected 0xEDB88320) slice-by-16 tables, 16 x 256
uint16_t host_input_log(uint16_t input);
uint16_t host_coeff_log(uint16_t input_log, uint16_t exponent);

// Per-device tables, uploaded on first use of a device (current device).
int device_tables(DeviceTables** out);

int set_error(hipError_t e);

// Kernel launchers (pph_kernels.hip); all return hipError_t.
hipError_t launch_fill_random(uint8_t* slots, size_t slot_bytes, unsigned count, size_t len,
                              uint64_t seed, uint32_t job, uint32_t first_slice, hipStream_t s);
hipError_t launch_prep(uint8_t* slots, const uint8_t* src, size_t slot_bytes, unsigned count, const uint32_t* lens,
                       const uint32_t* slot_index, size_t slice_size, size_t ck_off, unsigned ck,
                       bool to_log, const DeviceTables& t, hipStream_t s);
hipError_t launch_verify(const uint8_t* slots, size_t slot_bytes, unsigned count, size_t slice_size,
                         size_t ck_off, unsigned ck, int* valid, const DeviceTables& t, hipStream_t s);
// Slots [0, count) of one slot array in a launch over many (device-resident
// tables): data-form preparation (lens, optional: bytes of each input; first:
// number of slot 0 among the table's slots) or verification (valid: one flag
// per slot).
struct CkJob {
	uint8_t* slots;
	const uint32_t* lens;
	int* valid;
	uint64_t slot_bytes, slice_size, ck_off;
	uint32_t count, first;
};
// Host-side check of one entry against the launchers' shape rules.
bool ck_job_valid(const CkJob& j, unsigned ck);
hipError_t launch_prep_data_jobs(const CkJob* table, unsigned n_jobs, unsigned total, unsigned ck, unsigned per_cu,
                                 const DeviceTables& t, hipStream_t s);
hipError_t launch_verify_jobs(const CkJob* table, unsigned n_jobs, unsigned max_count, unsigned ck,
                              const DeviceTables& t, hipStream_t s);
// split: persistent grid with the work split evenly over the CUs (items
// shared by two workgroups combined with atomic XORs); false: one workgroup
// per (column tile, output group) item, for launches that share the CUs
hipError_t launch_muladd_plain(const uint16_t* in, size_t slot_words, unsigned n_in,
                               const uint16_t* in_logs, const uint16_t* exps,
                               const uint32_t* coef, unsigned coef_stride,
                               uint16_t* out, unsigned n_out, bool accumulate, unsigned group,
                               const DeviceTables& t, hipStream_t s, bool split = true);
hipError_t launch_muladd_dense(const uint16_t* in, size_t slot_words, unsigned n_in, const uint16_t* in_logs,
                               const uint16_t* exps, uint16_t* out, unsigned n_out, bool accumulate,
                               const DeviceTables& t, hipStream_t s);
// md5 / crc strides in bytes (16 and 4: two arrays; 20 and 20 with crc = md5 + 16: IFSC entries)
hipError_t launch_slice_hash_to(const uint8_t* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                                size_t slice_size, uint8_t* md5, size_t md5_stride, uint8_t* crc, size_t crc_stride,
                                const DeviceTables& t, hipStream_t s);
hipError_t launch_slice_hash(const uint8_t* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                             size_t slice_size, uint8_t* md5_out, uint32_t* crc_out, const DeviceTables& t,
                             hipStream_t s);
// Recovery-packet MD5 over `count` regions (state: 20 words each, device).
// first: start with head (set ID, packet type) and u32 exps[r]; done: message
// bytes before this call; n: data bytes of each slot appended (even; slots
// NULL appends zeros); final: pad and write 16-byte digests.
hipError_t launch_packet_md5(uint32_t* state, const uint8_t* slots, size_t slot_bytes, unsigned count,
                             const uint16_t* exps, const uint32_t head[8], uint64_t done, uint64_t n, bool first,
                             bool final, uint8_t* digest, hipStream_t s);
hipError_t launch_file_md5(const uint8_t* slots, size_t slot_bytes, unsigned count, size_t slice_size,
                           uint64_t file_size, uint8_t* md5_out, hipStream_t s);
// Spans of file MD5 chains (n Md5Span records, device-resident), one wave each.
hipError_t launch_file_md5_round(const void* spans, unsigned n, hipStream_t s);
// Rate model of the autotuner's table (pph_batch.hip): mul-adds per second.
// 0: unavailable; batched: the padding of the batch's mixed row tiles
double bilinear_rate(const struct ::pph_tuning& tu, unsigned n_in, unsigned n_out, bool batched = false);
double plain_rate(const struct ::pph_tuning& tu, unsigned n_out);
int autotune(struct ::pph_tuning* out);

// Bilinear form (pph_bilinear.hip); int status codes.
size_t bilinear_workspace_bytes(size_t slot_bytes, unsigned n_in, unsigned n_out);
int bilinear_supported(size_t slot_bytes, unsigned n_in, unsigned n_out);
int bilinear_reserve(size_t slot_bytes, unsigned n_in, unsigned n_out);
// coef (optional): explicit coefficients as in pph_muladd_plain's table
int bilinear_enqueue(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                     const uint16_t* exps, void* out_slots, unsigned n_out, bool accumulate, void* ws, size_t ws_bytes,
                     const DeviceTables& tab, hipStream_t s, const uint32_t* coef = nullptr, unsigned coef_stride = 0);
// pph_muladd_bilinear with an optional explicit coefficient table
int bilinear_call(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                  const uint16_t* exps, const uint32_t* coef, unsigned coef_stride, void* out_slots, unsigned n_out,
                  int accumulate, hipStream_t stream);
// fp4 MFMA kernel of the bilinear form (pph_bl_mfma.hip).  cap > 0 chunks the
// work so that the workspace stays within cap bytes (SIZE_MAX: it cannot).
size_t blm_workspace_bytes(size_t slot_bytes, unsigned n_in, unsigned n_out, size_t cap = 0);
int blm_enqueue(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                const uint16_t* exps, void* out_slots, unsigned n_out, bool accumulate, void* ws, size_t ws_bytes,
                const DeviceTables& tab, hipStream_t s, size_t cap = 0, const uint32_t* coef = nullptr,
                unsigned coef_stride = 0);
// Row tile of the bilinear GEMM for n_out outputs: the one that pads the
// outputs least, a 64-row tile's rows weighed 1.42 (measured: its waves expand
// ten masks per step where the others expand eight, and issue twice the
// loads), ties to the larger tile.  So 64 only up to 64 outputs.
inline size_t blm_row_tile(size_t n_out) {
	const auto ru = [](size_t v, size_t a) { return (v + a - 1) / a * a; };
	const double c256 = (double)ru(n_out, 256), c128 = (double)ru(n_out, 128), c64 = 1.42 * (double)ru(n_out, 64);
	if(c256 <= c128 && c256 <= c64) return 256;
	return c128 <= c64 ? 128 : 64;
}

// Rows of one job in the batched form, covered by tiles of several heights:
// n[0] tiles of 256 rows, then n[1] of 128, then n[2] of 64, the split that
// costs least with a row of the 128-row tile weighed 1.05 and of the 64-row
// tile 1.35 against the 256-row tile's (measured per call on 768..3072
// inputs: 1.00-1.08 and 1.22-1.41), fewer tiles on ties.  Returns the padded
// row count.
struct BlmRowSplit {
	unsigned n[3];
	size_t rows;
	double cost;
};
inline BlmRowSplit blm_row_split(size_t n_out) {
	BlmRowSplit best{{0, 0, 0}, 0, 1e300};
	const size_t top = (n_out + 255) / 256;
	for(size_t a = 0; a <= top; a++)
		for(unsigned b = 0; b < 4; b++)
			for(unsigned c = 0; c < 4; c++) {
				const size_t rows = 256 * a + 128 * b + 64 * c;
				if(rows < n_out) continue;
				const double cost = 256.0 * a + 1.05 * 128 * b + 1.35 * 64 * c + 1e-6 * (a + b + c);
				if(cost < best.cost) best = BlmRowSplit{{(unsigned)a, b, c}, rows, cost};
			}
	return best;
}

// Many bilinear jobs as one sequence of launches (pph_bl_mfma.hip): waves of
// word chunks whose planes share the workspace, cross-job GEMM item tables.
struct BlmBatchJob {
	const void* in_slots;
	size_t slot_bytes;
	unsigned n_in;
	const uint16_t* in_logs;
	const uint16_t* exps;
	void* out_slots;  // overwritten
	unsigned n_out;
	const uint32_t* in_lens;  // preparation: bytes of each input (null: slice_size)
	size_t slice_size;
	int* valid;  // verification: n_out checksum flags
};
struct BlmBatch;
bool blm_batch_eligible(const BlmBatchJob& j, size_t ws_budget);
size_t blm_batch_workspace(const BlmBatchJob* jobs, unsigned n, size_t ws_budget);
// ck > 0: the batch also prepares the jobs' inputs (data form, checksum width
// ck), each job before the word planes of its first wave, and verifies their
// outputs' checksums after the recombination of its last wave
int blm_batch_create(const BlmBatchJob* jobs, unsigned n, void* ws, size_t ws_bytes, unsigned ck, BlmBatch** out);
int blm_batch_enqueue(const BlmBatch* b, const DeviceTables& tab, hipStream_t s);
void blm_batch_destroy(BlmBatch* b);
hipError_t launch_xor_byte(uint8_t* p, uint8_t mask, hipStream_t s);
hipError_t launch_lds_gather_ceiling(const uint16_t* exp_tab, unsigned blocks, unsigned rounds, uint32_t seed,
                                     uint32_t* sink, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

// A span [begin, end) of one file's MD5 chain, the file laid out over slots
// (slot k holds file bytes [k*slice_size, ...)).  begin is a multiple of 64;
// end too unless the span is final (then end = file_size and the MD5 padding
// follows).  The chain state lives in `state` (4 words, device) between
// spans; a final span writes the digest to `digest` (16 bytes, device).
struct Md5Span {
	const uint8_t* slots;
	uint64_t slot_bytes, slice_size, file_size, begin, end;
	uint32_t* state;
	uint8_t* digest;
	uint32_t flags;  // kMd5SpanFirst | kMd5SpanFinal
	uint32_t pad;
};
constexpr uint32_t kMd5SpanFirst = 1, kMd5SpanFinal = 2;

// A descriptor read from a table, made wave-uniform (scalar registers): the
// tables are read with vector loads (the kernels also store to global memory),
// which the compiler would otherwise treat as per-lane values.
// (Word-wise bit casts: a byte-wise copy through memcpy made the compiler keep
// the fields as separate bytes in scalar registers and reassemble them on
// every use, ~100 instructions per GEMM item.)
template <class T>
__device__ __forceinline__ T wave_uniform(const T& x) {
	static_assert(sizeof(T) % 4 == 0, "descriptor words");
	struct Words {
		uint32_t w[sizeof(T) / 4];
	};
	Words w = __builtin_bit_cast(Words, x);
#pragma unroll
	for(unsigned k = 0; k < sizeof(T) / 4; k++) w.w[k] = __builtin_amdgcn_readfirstlane(w.w[k]);
	return __builtin_bit_cast(T, w);
}

inline size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
inline size_t ck_offset(size_t slice_size) { return round_up(slice_size, 16); }
inline size_t slot_size(size_t slice_size, unsigned ck) { return round_up(ck_offset(slice_size) + ck, kSlotAlign); }
inline bool valid_ck(unsigned ck) { return ck == 8 || ck == 16 || ck == 32 || ck == 64; }

}  // namespace pph
