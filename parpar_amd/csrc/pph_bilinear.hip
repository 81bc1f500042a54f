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
//
// Two implementations of the 81 GEMMs: the default hand-written fp4 MFMA
// kernel over bit-packed planes (pph_bl_mfma.hip), and 81 batched int8 GEMMs in
// hipBLASLt (PPH_BILINEAR_GEMM=library), whose pipeline is:
//   k_bl_coef_planes   Ac_t[r][i] = E(c_ri)_t as int8 0/1 (coefficients from the
//                      PAR2 input logs and recovery exponents)
//   for each chunk of slot words:
//     k_bl_word_planes Xw_t[w][i] = E(x_i[w])_t as int8 0/1
//     hipBLASLt        81 batched int8 GEMMs, int32 results G_t
//     k_bl_epilogue    parity of G_t, recombination to 16-bit words, store
//                      (or XOR) into the recovery slots
// The expansion is computed bit-sliced: 32 values' bit b packed in one word,
// so each Karatsuba level is a handful of XORs on whole words.
#include "pph_internal.h"
#include "pph_plain_tile.h"
#include "pph_karatsuba.h"
#include "../../include/parpar_hip.h"

#include <hipblaslt/hipblaslt.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace pph {

// ------------------------------------------------------------------ host maps

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

// ------------------------------------------------------------------ device

// four mask bits -> four 0/1 bytes
__device__ __forceinline__ uint32_t bytes4(uint32_t nib) { return ((nib & 0xFu) * 0x00204081u) & 0x01010101u; }

// 32 component masks -> 32 bytes at dst (16-byte aligned)
__device__ __forceinline__ void store_bytes32(uint8_t* dst, uint32_t m) {
	uint4 a, b;
	a.x = bytes4(m);
	a.y = bytes4(m >> 4);
	a.z = bytes4(m >> 8);
	a.w = bytes4(m >> 12);
	b.x = bytes4(m >> 16);
	b.y = bytes4(m >> 20);
	b.z = bytes4(m >> 24);
	b.w = bytes4(m >> 28);
	reinterpret_cast<uint4*>(dst)[0] = a;
	reinterpret_cast<uint4*>(dst)[1] = b;
}

// Coefficient planes: Ac[t][r][i] (row stride k_pad, plane stride n_rows * k_pad)
// for outputs r0 .. r0 + n_rows.  Thread: one output, 32 consecutive inputs.
__global__ __launch_bounds__(256) void k_bl_coef_planes(uint8_t* __restrict__ ac, size_t k_pad, unsigned n_in,
                                                        unsigned r0, unsigned n_rows,
                                                        const uint16_t* __restrict__ in_logs,
                                                        const uint16_t* __restrict__ exps,
                                                        const uint16_t* __restrict__ exp_tab) {
	const size_t groups = k_pad / 32;
	const size_t total = groups * n_rows;
	const size_t plane = (size_t)n_rows * k_pad;
	for(size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
		const unsigned r = (unsigned)(g / groups);
		const unsigned i0 = (unsigned)(g - (size_t)r * groups) * 32;
		const uint32_t er = exps[r0 + r];
		uint32_t s[16];
#pragma unroll
		for(int b = 0; b < 16; b++) s[b] = 0;
		for(int k = 0; k < 32; k++) {
			const unsigned i = i0 + k;
			uint32_t c = 0;
			if(i < n_in) {
				uint32_t l = mod65535((uint32_t)in_logs[i] * er);
				c = exp_tab[l == kOrder ? 0u : l];
			}
#pragma unroll
			for(int b = 0; b < 16; b++) s[b] |= ((c >> b) & 1u) << k;
		}
		uint32_t e[kBl];
		bl_expand_slices(s, e);
		uint8_t* dst = ac + (size_t)r * k_pad + i0;
#pragma unroll
		for(unsigned t = 0; t < kBl; t++) store_bytes32(dst + t * plane, e[t]);
	}
}

constexpr unsigned kBlTileI = 256, kBlTileW = 64, kBlLdsStride = kBlTileI + 8;

// Input-word planes for words [w0, w0 + wc) of every input slot:
// Xw[t][w - w0][i] (row stride k_pad, plane stride wc_stride * k_pad).
// Tile: 256 inputs x 64 words staged through LDS (coalesced slot reads), then
// one thread per (word, 32 inputs) expands and writes 32 bytes per plane.
__global__ __launch_bounds__(256) void k_bl_word_planes(uint8_t* __restrict__ xw, size_t k_pad, size_t wc,
                                                        size_t wc_stride, const uint8_t* __restrict__ in_slots,
                                                        size_t slot_bytes, unsigned n_in, size_t w0,
                                                        size_t slot_words) {
	__shared__ __align__(16) uint16_t s_x[kBlTileW * kBlLdsStride];
	const size_t ti_count = k_pad / kBlTileI;
	const size_t tw_count = (wc + kBlTileW - 1) / kBlTileW;
	const size_t plane = wc_stride * k_pad;
	for(size_t tile = blockIdx.x; tile < ti_count * tw_count; tile += gridDim.x) {
		const size_t ti = tile % ti_count, tw = tile / ti_count;
		const size_t i_base = ti * kBlTileI, wl_base = tw * kBlTileW;
		__syncthreads();  // the previous tile's reads are done
#pragma unroll
		for(int k = 0; k < 8; k++) {
			const unsigned idx = threadIdx.x + 256 * k;
			const unsigned il = idx >> 3, q = idx & 7;
			const size_t i = i_base + il;
			const size_t wl = wl_base + 8 * q;
			const size_t w = w0 + wl;
			uint4 v = make_uint4(0, 0, 0, 0);
			if(i < n_in && wl < wc && w < slot_words)
				v = *reinterpret_cast<const uint4*>(in_slots + i * slot_bytes + 2 * w);
			const uint32_t d[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
			for(int j = 0; j < 4; j++) {
				s_x[(8 * q + 2 * j) * kBlLdsStride + il] = (uint16_t)d[j];
				s_x[(8 * q + 2 * j + 1) * kBlLdsStride + il] = (uint16_t)(d[j] >> 16);
			}
		}
		__syncthreads();
#pragma unroll 1
		for(int h = 0; h < 2; h++) {
			const unsigned item = threadIdx.x + 256 * h;
			const unsigned ig = item & 7, wl = item >> 3;
			if(wl_base + wl >= wc) continue;
			const uint4* src = reinterpret_cast<const uint4*>(s_x + wl * kBlLdsStride + ig * 32);
			uint32_t s[16];
#pragma unroll
			for(int b = 0; b < 16; b++) s[b] = 0;
#pragma unroll
			for(int q = 0; q < 4; q++) {
				const uint4 v = src[q];
				const uint32_t d[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
				for(int j = 0; j < 4; j++) {
					const uint32_t lo = d[j] & 0xFFFFu, hi = d[j] >> 16;
					const int k = 8 * q + 2 * j;
#pragma unroll
					for(int b = 0; b < 16; b++) s[b] |= (((lo >> b) & 1u) << k) | (((hi >> b) & 1u) << (k + 1));
				}
			}
			uint32_t e[kBl];
			bl_expand_slices(s, e);
			uint8_t* dst = xw + (wl_base + wl) * k_pad + i_base + ig * 32;
#pragma unroll
			for(unsigned t = 0; t < kBl; t++) store_bytes32(dst + t * plane, e[t]);
		}
	}
}

// Parity and recombination: word w of recovery slot r0 + r from G[t][r][w]
// (row stride wc, plane stride wc * n_rows).  Two words per thread.
__global__ __launch_bounds__(256) void k_bl_epilogue(const int32_t* __restrict__ g, size_t wc, unsigned n_rows,
                                                     unsigned r0, uint8_t* __restrict__ out_slots, size_t slot_bytes,
                                                     size_t w0, int accumulate, const uint16_t* __restrict__ wtab) {
	const size_t pairs = wc / 2;
	const size_t total = pairs * n_rows;
	const size_t plane = wc * n_rows;
	for(size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
		const size_t r = idx / pairs;
		const size_t w = 2 * (idx - r * pairs);
		const int32_t* src = g + r * wc + w;
		uint32_t o0 = 0, o1 = 0;
#pragma unroll 9
		for(unsigned t = 0; t < kBl; t++) {
			const int2 v = *reinterpret_cast<const int2*>(src + t * plane);
			const uint32_t wt = wtab[t];
			o0 ^= wt & (0u - ((uint32_t)v.x & 1u));
			o1 ^= wt & (0u - ((uint32_t)v.y & 1u));
		}
		uint32_t* dst = reinterpret_cast<uint32_t*>(out_slots + (r0 + r) * slot_bytes + 2 * (w0 + w));
		const uint32_t o = o0 | (o1 << 16);
		*dst = accumulate ? (*dst ^ o) : o;
	}
}

// ------------------------------------------------------------------ driver

namespace {

struct GemmPlan {
	hipblasLtMatmulDesc_t desc = nullptr;
	hipblasLtMatrixLayout_t a = nullptr, b = nullptr, c = nullptr;
	hipblasLtMatmulAlgo_t algo{};
	size_t ws = 0;
	bool ok = false;
};

struct BlDevice {
	std::mutex mu;
	hipblasLtHandle_t lt = nullptr;
	uint8_t* buf = nullptr;  // workspace: coefficient planes | word planes | G | library workspace
	size_t buf_bytes = 0;
	uint16_t* d_wtab = nullptr;
	hipEvent_t tail = nullptr;  // end of the last call on the shared workspace
	bool tail_valid = false;
	std::mutex call_mu;         // held across one call's wait, enqueue and record
	std::map<std::tuple<size_t, size_t, size_t, size_t, size_t>, GemmPlan> plans;  // (M, N, K, strides)
};

constexpr size_t kLtWorkspace = 128ull << 20;
constexpr size_t kCoefBudget = 24ull << 30;   // coefficient planes per output block
constexpr size_t kChunkBudget = 20ull << 30;  // word planes + G per word chunk

BlDevice* bl_device(int dev) {
	static std::mutex mu;
	static std::map<int, BlDevice*> m;
	std::lock_guard<std::mutex> lk(mu);
	auto& p = m[dev];
	if(!p) p = new BlDevice;
	return p;
}

// C(M x N, int32, column-major ld M) = A^T B, A: K x M int8 (ld k_pad, batch
// stride sa), B: K x N int8 (ld k_pad, batch stride sb), 81 batches.
GemmPlan* gemm_plan(BlDevice* d, size_t M, size_t N, size_t K, size_t sa, size_t sb) {
	auto key = std::make_tuple(M, N, K, sa, sb);
	auto it = d->plans.find(key);
	if(it != d->plans.end()) return &it->second;
	GemmPlan& p = d->plans[key];
	const int32_t batch = (int32_t)kBl;
	const int64_t strA = (int64_t)sa, strB = (int64_t)sb, strC = (int64_t)(M * N);
	hipblasOperation_t ta = HIPBLAS_OP_T, tb = HIPBLAS_OP_N;
	bool ok = hipblasLtMatmulDescCreate(&p.desc, HIPBLAS_COMPUTE_32I, HIP_R_32I) == HIPBLAS_STATUS_SUCCESS;
	ok = ok && hipblasLtMatmulDescSetAttribute(p.desc, HIPBLASLT_MATMUL_DESC_TRANSA, &ta, sizeof(ta)) == HIPBLAS_STATUS_SUCCESS;
	ok = ok && hipblasLtMatmulDescSetAttribute(p.desc, HIPBLASLT_MATMUL_DESC_TRANSB, &tb, sizeof(tb)) == HIPBLAS_STATUS_SUCCESS;
	ok = ok && hipblasLtMatrixLayoutCreate(&p.a, HIP_R_8I, K, M, K) == HIPBLAS_STATUS_SUCCESS;
	ok = ok && hipblasLtMatrixLayoutCreate(&p.b, HIP_R_8I, K, N, K) == HIPBLAS_STATUS_SUCCESS;
	ok = ok && hipblasLtMatrixLayoutCreate(&p.c, HIP_R_32I, M, N, M) == HIPBLAS_STATUS_SUCCESS;
	struct {
		hipblasLtMatrixLayout_t l;
		int64_t stride;
	} ls[3] = {{p.a, strA}, {p.b, strB}, {p.c, strC}};
	for(auto& x : ls) {
		ok = ok && hipblasLtMatrixLayoutSetAttribute(x.l, HIPBLASLT_MATRIX_LAYOUT_BATCH_COUNT, &batch, sizeof(batch)) ==
		               HIPBLAS_STATUS_SUCCESS;
		ok = ok && hipblasLtMatrixLayoutSetAttribute(x.l, HIPBLASLT_MATRIX_LAYOUT_STRIDED_BATCH_OFFSET, &x.stride,
		                                             sizeof(x.stride)) == HIPBLAS_STATUS_SUCCESS;
	}
	hipblasLtMatmulPreference_t pref = nullptr;
	ok = ok && hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS;
	uint64_t wsmax = kLtWorkspace;
	ok = ok && hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsmax,
	                                                 sizeof(wsmax)) == HIPBLAS_STATUS_SUCCESS;
	hipblasLtMatmulHeuristicResult_t res[4];
	int n = 0;
	ok = ok && hipblasLtMatmulAlgoGetHeuristic(d->lt, p.desc, p.a, p.b, p.c, p.c, pref, 4, res, &n) ==
	               HIPBLAS_STATUS_SUCCESS;
	if(pref) hipblasLtMatmulPreferenceDestroy(pref);
	ok = ok && n > 0 && res[0].state == HIPBLAS_STATUS_SUCCESS && res[0].workspaceSize <= kLtWorkspace;
	if(ok) {
		p.algo = res[0].algo;
		p.ws = res[0].workspaceSize;
	}
	p.ok = ok;
	return &p;
}

struct Shape {
	size_t k_pad, rows, wc, chunks, slot_words;
	size_t coef_bytes, word_bytes, g_bytes, total;
};

// Tuning overrides (positive integers): PPH_BILINEAR_ROWS caps the outputs per
// coefficient block, PPH_BILINEAR_CHUNK_WORDS the words per chunk.
size_t env_size(const char* name) {
	const char* v = getenv(name);
	return v ? (size_t)strtoull(v, nullptr, 10) : 0;
}

Shape bl_shape(size_t slot_bytes, unsigned n_in, unsigned n_out) {
	Shape s{};
	s.k_pad = round_up(n_in, kBlTileI);
	s.slot_words = slot_bytes / 2;
	s.rows = std::min<size_t>(n_out, std::max<size_t>(32, kCoefBudget / (kBl * s.k_pad) / 32 * 32));
	if(size_t cap = env_size("PPH_BILINEAR_ROWS")) s.rows = std::min(s.rows, cap);
	size_t wc = kChunkBudget / (kBl * (s.k_pad + 4 * s.rows));
	if(size_t cap = env_size("PPH_BILINEAR_CHUNK_WORDS")) wc = std::min(wc, cap);
	wc = std::max<size_t>(64, wc / 64 * 64);
	wc = std::min(wc, round_up(s.slot_words, 64));
	s.chunks = (s.slot_words + wc - 1) / wc;
	// even chunks (multiples of 64 words)
	s.wc = round_up((s.slot_words + s.chunks - 1) / s.chunks, 64);
	s.coef_bytes = round_up(kBl * s.rows * s.k_pad, 256);
	s.word_bytes = round_up(kBl * s.wc * s.k_pad, 256);
	s.g_bytes = round_up(kBl * s.wc * s.rows * 4, 256);
	s.total = s.coef_bytes + s.word_bytes + s.g_bytes + kLtWorkspace;
	return s;
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
	*bytes = bl_shape(slot_bytes, n_in, n_out).total;
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

// PPH_BILINEAR_GEMM=library selects the hipBLASLt implementation.
bool bl_library() {
	const char* v = getenv("PPH_BILINEAR_GEMM");
	return v && !strcmp(v, "library");
}

size_t bilinear_workspace_bytes(size_t slot_bytes, unsigned n_in, unsigned n_out) {
	return bl_library() ? bl_shape(slot_bytes, n_in, n_out).total : blm_workspace_bytes(slot_bytes, n_in, n_out);
}

// Library handle, recombination table and the GEMM plans of a call's shapes.
// Host work only: run it before a stream capture that encloses the enqueue.
int bilinear_prepare(size_t slot_bytes, unsigned n_in, unsigned n_out) {
	if(!n_in || !n_out || n_in > 32768 || slot_bytes % kSlotAlign) return PPH_E_INVALID;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return set_error(e);
	BlDevice* d = bl_device(dev);
	std::lock_guard<std::mutex> lk(d->mu);
	if(!d->lt && hipblasLtCreate(&d->lt) != HIPBLAS_STATUS_SUCCESS) return PPH_E_DEVICE;
	if(!d->d_wtab) {
		e = hipMalloc((void**)&d->d_wtab, kBl * sizeof(uint16_t));
		if(e == hipSuccess) e = hipMemcpy(d->d_wtab, bl_recombine_table(), kBl * sizeof(uint16_t), hipMemcpyHostToDevice);
		if(e != hipSuccess) return set_error(e);
	}
	if(!bl_library()) return PPH_OK;
	const Shape sh = bl_shape(slot_bytes, n_in, n_out);
	for(size_t r0 = 0; r0 < n_out; r0 += sh.rows) {
		const size_t rows = std::min<size_t>(sh.rows, n_out - r0);
		for(size_t w0 = 0; w0 < sh.slot_words; w0 += sh.wc) {
			const size_t wc = std::min(sh.wc, sh.slot_words - w0);
			if(!gemm_plan(d, wc, rows, sh.k_pad, sh.wc * sh.k_pad, rows * sh.k_pad)->ok) return PPH_E_UNSUPPORTED;
		}
	}
	return PPH_OK;
}

// Enqueue one bilinear mul-add on s with a caller-owned workspace of at least
// bilinear_workspace_bytes.  bilinear_prepare must have succeeded for the shape.
int bilinear_enqueue(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                     const uint16_t* exps, void* out_slots, unsigned n_out, bool accumulate, void* ws, size_t ws_bytes,
                     const DeviceTables& tab, hipStream_t s) {
	if(!n_out) return PPH_OK;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return set_error(e);
	BlDevice* d = bl_device(dev);
	if(!bl_library()) {
		if(!d->d_wtab) return PPH_E_STATE;
		return blm_enqueue(in_slots, slot_bytes, n_in, input_logs, exps, out_slots, n_out, accumulate, ws, ws_bytes,
		                   d->d_wtab, tab, s);
	}
	const Shape sh = bl_shape(slot_bytes, n_in, n_out);
	if(!ws || ws_bytes < sh.total) return PPH_E_INVALID;
	std::lock_guard<std::mutex> lk(d->mu);
	if(!d->lt || !d->d_wtab) return PPH_E_STATE;
	uint8_t* coef = static_cast<uint8_t*>(ws);
	uint8_t* words = coef + sh.coef_bytes;
	int32_t* g = reinterpret_cast<int32_t*>(words + sh.word_bytes);
	void* ltws = words + sh.word_bytes + sh.g_bytes;
	const int cus = tab.cus > 0 ? tab.cus : 256;
	const int32_t alpha = 1, beta = 0;
	for(size_t r0 = 0; r0 < n_out; r0 += sh.rows) {
		const unsigned rows = (unsigned)std::min<size_t>(sh.rows, n_out - r0);
		const size_t coef_plane = (size_t)rows * sh.k_pad;
		{
			const size_t items = sh.k_pad / 32 * rows;
			const unsigned grid = (unsigned)std::min<size_t>((items + 255) / 256, (size_t)cus * 16);
			hipLaunchKernelGGL(k_bl_coef_planes, dim3(grid), dim3(256), 0, s, coef, sh.k_pad, n_in, (unsigned)r0, rows,
			                   input_logs, exps, tab.exp);
			e = hipGetLastError();
			if(e != hipSuccess) return set_error(e);
		}
		for(size_t w0 = 0; w0 < sh.slot_words; w0 += sh.wc) {
			const size_t wc = std::min(sh.wc, sh.slot_words - w0);  // multiple of 64
			const size_t tiles = sh.k_pad / kBlTileI * ((wc + kBlTileW - 1) / kBlTileW);
			hipLaunchKernelGGL(k_bl_word_planes, dim3((unsigned)std::min<size_t>(tiles, (size_t)cus * 8)), dim3(256), 0,
			                   s, words, sh.k_pad, wc, sh.wc, (const uint8_t*)in_slots, slot_bytes, n_in, w0,
			                   sh.slot_words);
			e = hipGetLastError();
			if(e != hipSuccess) return set_error(e);
			auto it = d->plans.find(std::make_tuple(wc, (size_t)rows, sh.k_pad, sh.wc * sh.k_pad, coef_plane));
			if(it == d->plans.end() || !it->second.ok) return PPH_E_STATE;
			const GemmPlan* p = &it->second;
			if(hipblasLtMatmul(d->lt, p->desc, &alpha, words, p->a, coef, p->b, &beta, g, p->c, g, p->c, &p->algo, ltws,
			                   p->ws, s) != HIPBLAS_STATUS_SUCCESS)
				return PPH_E_DEVICE;
			const size_t items = wc / 2 * rows;
			hipLaunchKernelGGL(k_bl_epilogue, dim3((unsigned)std::min<size_t>((items + 255) / 256, (size_t)cus * 16)),
			                   dim3(256), 0, s, g, wc, rows, (unsigned)r0, (uint8_t*)out_slots, slot_bytes, w0,
			                   accumulate ? 1 : 0, d->d_wtab);
			e = hipGetLastError();
			if(e != hipSuccess) return set_error(e);
		}
	}
	return PPH_OK;
}

}  // namespace pph

extern "C" int pph_muladd_bilinear(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                                   const uint16_t* exps, void* out_slots, unsigned n_out, int accumulate,
                                   void* stream) {
	if(!n_out) return PPH_OK;
	if(!in_slots || !out_slots || !input_logs || !exps || !n_in || n_in > 32768 || slot_bytes % kSlotAlign)
		return PPH_E_INVALID;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess) return set_error(e);
	DeviceTables* tab = nullptr;
	int rc = device_tables(&tab);
	if(rc) return rc;
	rc = bilinear_prepare(slot_bytes, n_in, n_out);
	if(rc) return rc;
	BlDevice* d = bl_device(dev);
	std::lock_guard<std::mutex> call(d->call_mu);
	const size_t need = bilinear_workspace_bytes(slot_bytes, n_in, n_out);
	void* ws = nullptr;
	{
		std::lock_guard<std::mutex> lk(d->mu);
		if(d->buf_bytes < need) {
			if(d->buf) {
				e = hipDeviceSynchronize();
				if(e != hipSuccess) return set_error(e);
				(void)hipFree(d->buf);
				d->buf = nullptr;
				d->buf_bytes = 0;
			}
			e = hipMalloc((void**)&d->buf, need);
			if(e != hipSuccess) {
				(void)hipGetLastError();
				return e == hipErrorOutOfMemory ? PPH_E_NOMEM : set_error(e);
			}
			d->buf_bytes = need;
		}
		ws = d->buf;
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
	rc = bilinear_enqueue(in_slots, slot_bytes, n_in, input_logs, exps, out_slots, n_out, accumulate != 0, ws,
	                      d->buf_bytes, *tab, (hipStream_t)stream);
	std::lock_guard<std::mutex> lk(d->mu);
	e = hipEventRecord(d->tail, (hipStream_t)stream);
	if(e != hipSuccess) return set_error(e);
	d->tail_valid = true;
	return rc;
}
