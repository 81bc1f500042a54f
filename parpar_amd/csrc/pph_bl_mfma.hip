// Bilinear form on fp4 MFMA: the 81 integer GEMMs of the rank-81 formula
// (see pph_bilinear.hip) computed by a hand-written kernel on
// v_mfma_scale_f32_32x32x64_f8f6f4 with e2m1 operands (0 or a power of two,
// paired so that every product of set elements is the same power of two; the
// fp32 sums are exact), twice the int8 rate.
//
// Operands are bit-packed planes, one bit per element:
//   Cb[t][kw][r]  bit of E(c(r, 32kw + k))_t     (coefficients, outputs r)
//   Wb[t][kw][w]  bit of E(x_{32kw + k}[w])_t     (input words w)
// with element k of a 32-input group at bit 4(k % 8) + q(k / 8) (blm_pos,
// blm_pos_c), so that most operand dwords are one AND of the mask.  The GEMM
// kernel (persistent; items of 256 words x 256 outputs, or 512 words x 128
// outputs, and component t) stages them in LDS (one 64-input step per
// iteration), runs 16 MFMAs per wave per step and writes the parity of each
// sum as a bit:
//   P[t][w / 32][r]  bit w % 32 = parity of sum_i E(c_ri)_t E(x_i[w])_t
// k_blm_combine recombines the 81 parity bits of each word into the 16-bit
// recovery word (bit-sliced over 32 words) and stores or XORs it.
#include "pph_internal.h"
#include "pph_plain_tile.h"
#include "pph_karatsuba.h"
#include "../../include/parpar_hip.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <type_traits>
#include <vector>

namespace pph {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Packed masks: element k (0..31) of a 32-input group sits at bit 4 (k % 8)
// + q of its mask, q a per-operand function of k / 8, so that the MFMA
// operand (nibble k % 8 of dword k / 8) is mostly one AND.  A set nibble bit
// is an e2m1 value (bit 0: 0.5, bit 1: 1.0, bit 2: 2.0); the word operand
// puts groups 0, 1, 2, 3 on bits 0, 1, 2, 2 and the coefficient operand on
// bits 2, 1, 0, 0, so every product of two set elements is exactly 1 and the
// fp32 sums are the integer counts.
//   word masks:        q = k / 8           (dword 3 shifted down one bit)
//   coefficient masks: q = (2, 1, 0, 3)[k / 8]  (dword 3 shifted down three)
__device__ __forceinline__ unsigned blm_pos(unsigned k) { return 4 * (k & 7) + (k >> 3); }
__device__ __forceinline__ unsigned blm_pos_c(unsigned k) { return 4 * (k & 7) + ((0x3012u >> (4 * (k >> 3))) & 15u); }

// 32 packed elements -> 16 bytes of e2m1 nibbles
__device__ __forceinline__ uint4 blm_nibbles(uint32_t m) {
	return make_uint4(m & 0x11111111u, m & 0x22222222u, m & 0x44444444u, (m >> 1) & 0x44444444u);
}
__device__ __forceinline__ uint4 blm_nibbles_c(uint32_t m) {
	return make_uint4(m & 0x44444444u, m & 0x22222222u, m & 0x11111111u, (m >> 3) & 0x11111111u);
}

// Coefficient planes of one job's output block: outputs r0 .. r0 + n_rows
// (exps, or the explicit table, already offset by r0), rows padded to r_pad
// with zeros.  Coefficients are PAR2's, 2^(in_logs[i] * exps[r] mod 65535),
// or, with `coef`, explicit: coef[i * coef_stride + r] = 2 log2(c), 0xFFFFFFFF
// for c = 0 (the encoding of pph_muladd_plain's table).
struct BlmCoefJob {
	uint32_t* cb;  // [t][kw_count][r_pad]
	const uint16_t* in_logs;
	const uint16_t* exps;
	const uint32_t* coef;
	uint32_t kw_count, r_pad, n_in, n_rows, coef_stride, pad;
};

// Coefficient planes: job `one`, or job blockIdx.y of `table`.  Thread: one
// output row r, one 32-input group kw; r varies fastest.
__global__ __launch_bounds__(256) void k_blm_coef(BlmCoefJob one, const BlmCoefJob* __restrict__ table,
                                                  const uint16_t* __restrict__ exp_tab) {
	const BlmCoefJob J = table ? wave_uniform(table[blockIdx.y]) : one;
	const size_t total = (size_t)J.kw_count * J.r_pad;
	for(size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
		const size_t kw = g / J.r_pad;
		const unsigned r = (unsigned)(g - kw * J.r_pad);
		uint32_t s[16];
#pragma unroll
		for(int b = 0; b < 16; b++) s[b] = 0;
		if(r < J.n_rows) {
			const uint32_t er = J.coef ? 0u : J.exps[r];
			for(unsigned k = 0; k < 32; k++) {
				const size_t i = kw * 32 + k;
				if(i >= J.n_in) break;
				uint32_t c;
				if(J.coef) {
					const uint32_t l2 = J.coef[i * J.coef_stride + r];
					c = l2 == kSkipCoef ? 0u : exp_tab[l2 >> 1];
				} else {
					const uint32_t l = mod65535((uint32_t)J.in_logs[i] * er);
					c = exp_tab[l == kOrder ? 0u : l];
				}
				const unsigned p = blm_pos_c(k);
#pragma unroll
				for(int b = 0; b < 16; b++) s[b] |= ((c >> b) & 1u) << p;
			}
		}
		uint32_t e[kBl];
		bl_expand_slices(s, e);
#pragma unroll
		for(unsigned t = 0; t < kBl; t++) J.cb[((size_t)t * J.kw_count + kw) * J.r_pad + r] = e[t];
	}
}

// 32 x 32 bit-matrix transpose (rows a[j], bit b of row j -> bit j of a[b]),
// five butterfly stages of masked swaps.
__device__ __forceinline__ void blm_transpose32(uint32_t* a) {
#pragma unroll
	for(unsigned s = 16, m = 0x0000FFFFu; s; s >>= 1, m ^= m << s) {
#pragma unroll
		for(unsigned j = 0; j < 32; j++) {
			if(j & s) continue;
			const uint32_t t = ((a[j] >> s) ^ a[j + s]) & m;
			a[j + s] ^= t;
			a[j] ^= t << s;
		}
	}
}

// Word planes of one word chunk of one job: words w0 .. w0 + wv of every input
// slot (`in` already offset to word w0), in plane rows of wc words (wc a
// multiple of the item's word tile, wv of 512; the rest of a row is zero).
struct BlmWordsPiece {
	uint32_t* wb;  // [t][kw_count][wc]
	const uint8_t* in;
	uint64_t slot_bytes;
	uint32_t kw_count, wc, wv, n_in;
};

// Word planes: piece `one`, or piece blockIdx.y of `table`.  A workgroup
// covers 512 words of one 32-input group (chunks are multiples of 512 words,
// so the group, the unit's first word and whether it holds data are uniform:
// loads and stores take a scalar base and a 32-bit lane offset).  Thread: two
// adjacent words; each of the 32 slot reads (4 bytes) is coalesced.  Row j of
// the 32 x 32 bit matrix is input 8 (j % 4) + j / 4 (the packed position
// order), so after the transpose row b < 16 is bit slice b of the first word
// and row 16 + b that of the second.
__global__ __launch_bounds__(256) void k_blm_words(BlmWordsPiece one, const BlmWordsPiece* __restrict__ table) {
	const BlmWordsPiece P = table ? wave_uniform(table[blockIdx.y]) : one;
	const unsigned per_kw = P.wc / 512;
	const size_t units = (size_t)P.kw_count * per_kw;
	const size_t plane = (size_t)P.kw_count * P.wc;
	const unsigned tid = threadIdx.x;
	for(size_t u = blockIdx.x; u < units; u += gridDim.x) {
		const unsigned kw = (unsigned)(u / per_kw);
		const unsigned w0 = 512 * (unsigned)(u - (size_t)kw * per_kw);
		uint32_t a[32];
		if(w0 < P.wv) {
			// The slot pointer and the group's input count are made wave-uniform
			// explicitly: a descriptor the compiler takes for per-lane is issued in
			// a loop over lanes (one per load).
			const uint8_t* in = wave_uniform(P.in + 2 * (size_t)w0 + (size_t)kw * 32 * P.slot_bytes);
			const unsigned n = __builtin_amdgcn_readfirstlane(P.n_in > kw * 32 ? P.n_in - kw * 32 : 0);
#pragma unroll
			for(unsigned k = 0; k < 32; k++) {
				// past the last input: an empty resource, whose loads return 0
				const uint8_t* src = wave_uniform(in + k * P.slot_bytes);
				a[4 * (k % 8) + k / 8] = __builtin_amdgcn_raw_buffer_load_b32(
				    __builtin_amdgcn_make_buffer_rsrc((void*)src, 0, k < n ? 2048 : 0, 0x00020000), (int)(4 * tid), 0, 0);
			}
		} else {
#pragma unroll
			for(unsigned j = 0; j < 32; j++) a[j] = 0;
		}
		blm_transpose32(a);
		uint32_t e0[kBl], e1[kBl];
		bl_expand_slices(a, e0);
		bl_expand_slices(a + 16, e1);
		uint32_t* dst = P.wb + (size_t)kw * P.wc + w0;  // uniform
#pragma unroll
		for(unsigned t = 0; t < kBl; t++) {
			typedef unsigned v2u __attribute__((ext_vector_type(2)));
			__builtin_amdgcn_raw_buffer_store_b64(v2u{e0[t], e1[t]},
			                                      __builtin_amdgcn_make_buffer_rsrc(dst, 0, 2048, 0x00020000),
			                                      (int)(8 * tid), 0, 0);
			dst += plane;
		}
	}
}

constexpr unsigned kBlmWordTile = 512;  // word chunks are multiples of this
constexpr unsigned kBlmKStep = 256;     // inputs are padded to multiples of this
constexpr unsigned kBlmMinKw = 16;      // and to at least 512 (eight steps)

// One GEMM piece: the planes of one word chunk of one job.  Its items are
// (component t, word tile, row tile), numbered from item0, row tiles varying
// fastest (word tiles with row_fast 0).
struct BlmPiece {
	const uint32_t* wb;  // word planes [t][kw][wc_pad]
	const uint32_t* cb;  // coefficient planes [t][kw][r_pad]
	uint32_t* par;       // parity bits [t][wc_pad / 32][r_pad]
	uint32_t steps;      // 64-input steps, a multiple of 8
	uint32_t wc_pad;     // words, a multiple of 512
	uint32_t r_pad;      // outputs (the planes' row stride)
	uint32_t r_base;     // first output of the piece's row tiles (a multiple of 64)
	uint32_t w_tiles, r_tiles;
	uint32_t item0, items;
	uint32_t row_fast;
	// division by w_tiles * r_tiles and by the fastest tile count (r_tiles or
	// w_tiles) as a multiply-high by floor(2^32 / d) and one correction
	// (blm_magic)
	uint32_t m_tiles, m_inner;
};

// floor(2^32 / d) for d >= 2, 2^32 - 1 for d = 1: umulhi(n, m) is floor(n / d)
// or one less for every 32-bit n
inline uint32_t blm_magic(uint32_t d) { return d <= 1 ? 0xFFFFFFFFu : (uint32_t)((1ull << 32) / d); }
__device__ __forceinline__ unsigned blm_divmod(unsigned n, unsigned d, unsigned m, unsigned& r) {
	unsigned q = __umulhi(n, m);
	r = n - q * d;
	if(r >= d) {
		q++;
		r -= d;
	}
	return q;
}

// Work item: component t, RT outputs r0 .. r0+RT-1 and 65536/RT words w0 ..
// (256 x 256, waves 2x2; or 128 outputs x 512 words, waves 4x1 along the
// words).  Each wave covers 128 words x 128 outputs = 4 x 4 blocks of 32 x 32.
// The kernel is persistent: workgroup b runs items b and b + G (G = grid
// size), then b + 2G, b + 3G, ... or, with `counter`, the items a counter
// hands out, of one piece (`one`, table null) or of the pieces of `table`
// (consecutive item ranges), one workgroup per CU, so the launch is paid once
// per workgroup rather than once per item.
// One step = 64 inputs (two 32-input groups, one MFMA k-step); every item has
// a multiple of four steps and at least eight.  LDS holds the bit masks
// ([group][row], 8 bytes per row and step, four buffers); each wave expands
// its operands in registers, one step after reading their masks, so that no
// expansion waits on LDS.  One barrier per two steps.  The load pipeline runs
// on across items: the last eight steps of an item fetch the first eight of
// the next, so an item starts with its masks in LDS and registers.
// Sums: the first step of an item adds its products to 2.0 (an inline
// constant, so no accumulator is reset), with every product scaled by 2^-22
// (both operand scales 2^-11), so that an accumulator holds 2 + n 2^-22
// exactly (n < 2^22 inputs), in [2, 4) where the encoding's unit is 2^-22: the
// parity of n is bit 0 of the encoding.
template <unsigned RT>
__global__ __launch_bounds__(256) void k_blm_gemm(BlmPiece one, const BlmPiece* __restrict__ table, unsigned items,
                                                  unsigned* __restrict__ counter) {
	static_assert(RT == 256 || RT == 128 || RT == 64, "row tile");
	constexpr unsigned WT = 65536 / RT;          // words per item
	constexpr unsigned kA = 2 * WT / 256;        // word masks per thread and step
	constexpr unsigned kB = RT == 64 ? 1 : 2 * RT / 256;  // coefficient masks per thread and step (64: half the threads)
	constexpr unsigned NA = RT == 64 ? 8 : 4;    // 32-word blocks per wave
	constexpr unsigned NB = RT == 64 ? 2 : 4;    // 32-output blocks per wave
	constexpr unsigned kBufs = 4, kDepth = 4;
	__shared__ __align__(16) uint32_t sA[kBufs][2][WT];
	__shared__ __align__(16) uint32_t sB[kBufs][2][RT];
	const unsigned G = gridDim.x, b = blockIdx.x;
	if(b >= items) return;
	const unsigned tid = threadIdx.x, lane = tid & 63;
	const unsigned wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const unsigned wr = RT == 256 ? wave >> 1 : wave, wcol = RT == 256 ? wave & 1 : 0;
	const unsigned wa0 = 32 * NA * wr, wb0 = 32 * NB * wcol;  // this wave's first word / output of the item
	const unsigned half = lane >> 5, row32 = lane & 31;

	// Item descriptors.  A workgroup's items only move forward through the
	// table, so the piece is found by stepping from the last one.
	BlmPiece P = table ? wave_uniform(table[0]) : one;
	unsigned pc = 0;
	struct Src {
		__amdgpu_buffer_rsrc_t a, b;
		uint32_t* pbase;
		uint32_t pbytes;
		unsigned w0, r0, steps, lda, ldb;
	};
	constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit access, gfx9 family
	auto locate = [&](unsigned it) {
		if(table)
			while(it >= P.item0 + P.items) P = wave_uniform(table[++pc]);
		const unsigned li = it - P.item0;
		unsigned tile, inner;
		const unsigned t = blm_divmod(li, P.w_tiles * P.r_tiles, P.m_tiles, tile);
		const unsigned outer = blm_divmod(tile, P.row_fast ? P.r_tiles : P.w_tiles, P.m_inner, inner);
		const unsigned wt = P.row_fast ? outer : inner;
		const unsigned rt = P.row_fast ? inner : outer;
		const size_t pa = (size_t)P.steps * P.wc_pad * 2, pb = (size_t)P.steps * P.r_pad * 2;  // 2 groups a step
		const size_t pp = (size_t)(P.wc_pad / 32) * P.r_pad;
		return Src{__builtin_amdgcn_make_buffer_rsrc((void*)(P.wb + t * pa), 0, (int)(pa * 4), kRsrcWord3),
		           __builtin_amdgcn_make_buffer_rsrc((void*)(P.cb + t * pb), 0, (int)(pb * 4), kRsrcWord3),
		           P.par + t * pp,
		           (uint32_t)(pp * 4),
		           wt * WT,
		           P.r_base + rt * RT,
		           P.steps,
		           P.wc_pad,
		           P.r_pad};
	};

	v16f acc[NA][NB];
	// LDS layout [group][row]: a half-wave reads 32 consecutive masks.
	// m[0 .. kA) = word masks (group g = index / (kA / 2), row tid + 256 (index % (kA / 2))),
	// m[kA .. kA + kB) = coefficient masks (RT = 256: group = index; RT = 128:
	// group tid / 128, row tid % 128; RT = 64: threads below 128, group tid /
	// 64, row tid % 64)
	auto stage = [&](unsigned buf, const uint32_t* m) {
#pragma unroll
		for(unsigned q = 0; q < kA; q++) sA[buf][q / (kA / 2)][tid + 256 * (q % (kA / 2))] = m[q];
		if constexpr(RT == 256) {
			sB[buf][0][tid] = m[kA];
			sB[buf][1][tid] = m[kA + 1];
		} else if constexpr(RT == 128) {
			sB[buf][tid / 128][tid % 128] = m[kA];
		} else {
			if(wave < 2) sB[buf][tid / 64][tid % 64] = m[kA];
		}
	};
	// this lane's masks (group `half`) of its NA A rows and NB B rows, then
	// their e2m1 expansion into MFMA operands
	auto read_masks = [&](unsigned buf, uint32_t* rm) {
#pragma unroll
		for(unsigned i = 0; i < NA; i++) rm[i] = sA[buf][half][wa0 + 32 * i + row32];
#pragma unroll
		for(unsigned j = 0; j < NB; j++) rm[NA + j] = sB[buf][half][wb0 + 32 * j + row32];
	};
	auto expand = [&](const uint32_t* rm, v8i* fa, v8i* fb) {
#pragma unroll
		for(unsigned i = 0; i < NA; i++) {
			const uint4 qa = blm_nibbles(rm[i]);
			fa[i] = v8i{(int)qa.x, (int)qa.y, (int)qa.z, (int)qa.w, 0, 0, 0, 0};
		}
#pragma unroll
		for(unsigned j = 0; j < NB; j++) {
			const uint4 qb = blm_nibbles_c(rm[NA + j]);
			fb[j] = v8i{(int)qb.x, (int)qb.y, (int)qb.z, (int)qb.w, 0, 0, 0, 0};
		}
	};
	// Loads are buffer loads: a resource per item and plane (scalar), a byte
	// offset per thread and load that changes only with the piece (the second
	// group of a step is a plane row further), and the step's uniform byte
	// offset in a scalar register, so the address costs no VALU work.  A plane
	// is < 2^31 bytes (the workspace chunking keeps it so).
	struct Voff {
		int v[2 + kB];  // word masks of group 0, 1 (further masks of a group at +1 KiB each), then coefficients
	};
	auto voffsets = [&](const Src& f) {
		Voff o;
		o.v[0] = (int)(tid * 4);
		o.v[1] = (int)((f.lda + tid) * 4);
		if constexpr(RT == 256) {
			o.v[2] = (int)(tid * 4);
			o.v[3] = (int)((f.ldb + tid) * 4);
		} else if constexpr(RT == 128) {
			o.v[2] = (int)(((tid / 128) * f.ldb + tid % 128) * 4);
		} else {  // threads 128 and up load what threads 0 .. 127 do (in range, not staged)
			o.v[2] = (int)((((tid / 64) & 1) * f.ldb + tid % 64) * 4);
		}
		return o;
	};
	auto fetch = [&](uint32_t* m, const Src& f, const Voff& o, unsigned st) {
		const int sa = (int)((2 * st * f.lda + f.w0) * 4), sb = (int)((2 * st * f.ldb + f.r0) * 4);
#pragma unroll
		for(unsigned q = 0; q < kA; q++)
			m[q] = __builtin_amdgcn_raw_buffer_load_b32(f.a, o.v[q / (kA / 2)] + (int)(1024 * (q % (kA / 2))), sa, 0);
#pragma unroll
		for(unsigned q = 0; q < kB; q++) m[kA + q] = __builtin_amdgcn_raw_buffer_load_b32(f.b, o.v[2 + q], sb, 0);
	};
	// Items of this workgroup: b and b + G, then b + 2G, b + 3G, ... (counter
	// null) or 2G plus the values `counter` hands out, one per item, so that a
	// workgroup that starts late or shares its CU with other work takes fewer
	// items.  The next index is taken by one thread at the start of an item and
	// shared through LDS at its end (the atomic's latency lies under the item).
	unsigned i_cur = b, i_nxt = b + G;
	__shared__ unsigned s_next;
	Src cur = locate(i_cur);
	Src nxt = i_nxt < items ? locate(i_nxt) : cur;
	Voff vc = voffsets(cur), vn = voffsets(nxt);

	// Pipeline, step s (buffer and register-set indices static in the unrolled
	// pass): MFMAs on the operands of step s; expansion of the masks of step
	// s + 1 (read from LDS during step s - 1); LDS reads of the masks of step
	// s + 2; the global masks of step s + 4 into LDS; the global fetch of step
	// s + 8.  A barrier after every second step: a buffer is written two steps
	// after its last read and read two steps after its write.
	uint32_t pf[kDepth][kA + kB];
	v8i fa[2][NA], fb[2][NB];
	uint32_t rm[2][NA + NB];
	{
		uint32_t m0[4][kA + kB];
#pragma unroll
		for(unsigned k = 0; k < 4; k++) fetch(m0[k], cur, vc, k);
#pragma unroll
		for(unsigned k = 0; k < kDepth; k++) fetch(pf[k], cur, vc, 4 + k);
#pragma unroll
		for(unsigned k = 0; k < 4; k++) stage(k, m0[k]);
	}
	__syncthreads();
	read_masks(0, rm[0]);
	expand(rm[0], fa[0], fb[0]);
	read_masks(1, rm[1]);
	__syncthreads();  // buffers 0 and 1 are written again in the first two steps

	// The parity words of an item are stored during the fourth step of the next
	// item, after that step's loads: vector memory operations complete in
	// order, so a staging wait for a load issued behind the stores waits for
	// them too, and the first such wait now comes four steps after them (0.2 us
	// less per item than stores in the first step).
	// Before the first item the resource is empty, so its stores are dropped.
	// The destination is made wave-uniform where it is used (a loop-carried
	// value the compiler would otherwise treat as per-lane, looping over lanes).
	// po[i * NB / 2 + p]: lanes 0..31 hold the mask of block (i, 2p), lanes
	// 32..63 that of block (i, 2p + 1) (adjacent 32-output blocks: one 256-byte
	// store per pair)
	uint32_t po[NA * NB / 2];
#pragma unroll
	for(unsigned q = 0; q < NA * NB / 2; q++) po[q] = 0;
	struct Dst {
		uint32_t* base;  // parity plane of the item's component
		uint32_t bytes;  // its size (0: no item)
		int sp;          // byte offset of this wave's first block
		uint32_t ldb;
	};
	Dst pd{P.par, 0u, 0, 0u};
	auto store_prev = [&]() {
		const Dst d = wave_uniform(pd);
		const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)d.base, 0, (int)d.bytes, kRsrcWord3);
#pragma unroll
		for(unsigned i = 0; i < NA; i++)
#pragma unroll
			for(unsigned p = 0; p < NB / 2; p++)
				__builtin_amdgcn_raw_buffer_store_b32(po[i * NB / 2 + p], rs, (int)(lane * 4),
				                                      d.sp + (int)((i * d.ldb + 64 * p) * 4), 0);
	};

	constexpr int kScale = 116;  // e8m0 2^-11 on both operands
	const v16f init = {2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f,
	                   2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f, 2.0f};
	// Four steps s0 .. s0 + 3: a common multiple of the four LDS buffers, the two
	// operand and mask sets and the four prefetch sets; FIRST: s0 = 0, whose
	// first step starts the sums.  The pass's four fetches (steps s0 + 8 ..
	// s0 + 11) are all of this item or all of the next (step counts are
	// multiples of four): steps st0 .. st0 + 3 of item f.
	auto pass = [&](unsigned s0, auto first, const Src& f, const Voff& o, unsigned st0) __attribute__((always_inline)) {
#pragma unroll
		for(unsigned k = 0; k < 4; k++) {
			const unsigned c = k & 1, n = c ^ 1;
			read_masks((k + 2) % kBufs, rm[c]);
			expand(rm[n], fa[n], fb[n]);
#pragma unroll
			for(unsigned i = 0; i < NA; i++)
#pragma unroll
				for(unsigned j = 0; j < NB; j++)
					acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
					    fa[c][i], fb[c][j], (decltype(first)::value && k == 0) ? init : acc[i][j], 4, 4, 0, kScale, 0,
					    kScale);
			stage((k + 4) % kBufs, pf[k]);
			fetch(pf[k], f, o, st0 + k);
			if constexpr(decltype(first)::value)
				if(k == 3) store_prev();
			// Every MFMA followed by a few instructions of the other kinds, spread
			// evenly over the step, in the counts the step has: 40 VALU (the
			// expansion of masks that arrived a step ago), 4 LDS reads (two masks
			// each) in the first quarter, then the LDS writes (2, or 3 for the
			// 128-row tile) and the global loads (4, or 5).
			if constexpr(RT == 64) {
				// 16 MFMA, 50 VALU (10 masks), 5 LDS reads, 5 LDS writes, 9 loads
#pragma unroll
				for(int q = 0; q < 16; q++) {
					__builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
					if(q < 5) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
					if(q == 0 || q == 8)
						__builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
					else
						__builtin_amdgcn_sched_group_barrier(0x002, 3, 0);
					if(q >= 5 && q <= 13 && (q & 1)) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
					if(q >= 6 && q <= 14) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
					if(decltype(first)::value && k == 3 && q >= 14) __builtin_amdgcn_sched_group_barrier(0x040, 8, 0);
				}
			} else {
#pragma unroll
				for(int q = 0; q < 16; q++) {
					__builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // 1 MFMA
					if(q < 4) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // 1 LDS read
					if(q & 1)
						__builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // 2 VALU
					else
						__builtin_amdgcn_sched_group_barrier(0x002, 3, 0);  // 3 VALU
					if(q == 4 || q == 6 || (RT == 128 && q == 8))
						__builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // 1 LDS write
					if(q == 7 || q == 9 || q == 11 || q == 13 || (RT == 128 && q == 15))
						__builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // 1 global load
					if(decltype(first)::value && k == 3 && q >= 14)
						__builtin_amdgcn_sched_group_barrier(0x040, 8, 0);  // 8 stores (previous item)
				}
			}
			if(k & 1) __syncthreads();
		}
	};
	for(;;) {
		unsigned ticket = 0;
		if(counter && i_nxt < items && tid == 0)
			ticket = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		// the first pass fetches steps 8 .. 11: of the next item when this one
		// has eight steps
		{
			const bool more = cur.steps > 8;
			Src f0 = nxt;
			Voff o0 = vn;
			if(more) {
				f0 = cur;
				o0 = vc;
			}
			pass(0u, std::true_type{}, f0, o0, more ? 8u : 0u);
		}
		unsigned s0 = 4;
		for(; s0 + 4 + kDepth < cur.steps; s0 += 4) pass(s0, std::false_type{}, cur, vc, s0 + 4 + kDepth);
		for(; s0 < cur.steps; s0 += 4) pass(s0, std::false_type{}, nxt, vn, s0 + 4 + kDepth - cur.steps);
		// item done: parity bits, lane (half h, row32 = output column) holding
		// words m = 8(v / 4) + 4h + v % 4 of each block; gathering the low bytes
		// of sums c, 4 + c, 8 + c, 12 + c into one dword and shifting their bit 0
		// up by 4h + c puts word m at bit m of the block's mask.  The two
		// half-waves hold disjoint bits of the same mask; one lane swap per pair
		// of blocks joins them.
		// The sums are read with an AGPR-constrained move, which keeps the
		// accumulators in AGPRs across the item loop (a plain VALU read inside
		// the loop makes the allocator move them to VGPRs and spill).  A VALU
		// read of an MFMA result needs up to 19 wait states after the MFMA, and
		// asm reads are outside the compiler's hazard tracking: the padding gives
		// 32, and every read is volatile so that none is scheduled above it
		// (volatile asm statements keep their order).
		asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7\n\ts_nop 7");
		const int sp = (int)((((cur.w0 + wa0) / 32) * cur.ldb + cur.r0 + wb0) * 4);
#pragma unroll
		for(unsigned i = 0; i < NA; i++)
#pragma unroll
			for(unsigned p = 0; p < NB / 2; p++) {
				uint32_t m2[2];
#pragma unroll
				for(unsigned q = 0; q < 2; q++) {
					const unsigned j = 2 * p + q;
					uint32_t u[16];
#pragma unroll
					for(int v = 0; v < 16; v++) asm volatile("v_accvgpr_read_b32 %0, %1" : "=v"(u[v]) : "a"(acc[i][j][v]));
					uint32_t m = 0;
#pragma unroll
					for(int c = 0; c < 4; c++) {
						const uint32_t w = __builtin_amdgcn_perm(u[4 + c], u[c], 0x0c0c0400u) |
						                   __builtin_amdgcn_perm(u[12 + c], u[8 + c], 0x04000c0cu);
						m |= (w & 0x01010101u) << (4 * half + c);
					}
					m2[q] = m;
				}
				// lanes 32..63 of the first with lanes 0..31 of the second: the OR
				// joins each block's two half-waves, block 2p in lanes 0..31 and
				// block 2p + 1 in lanes 32..63
				const auto sw = __builtin_amdgcn_permlane32_swap(m2[0], m2[1], false, false);
				po[i * NB / 2 + p] = sw[0] | sw[1];
			}
		pd = Dst{cur.pbase, cur.pbytes, sp, cur.ldb};
		if(i_nxt >= items) break;
		unsigned i_after = i_nxt + G;
		if(counter) {
			if(tid == 0) s_next = 2 * G + ticket;
			__syncthreads();
			i_after = __builtin_amdgcn_readfirstlane(s_next);
		}
		cur = nxt;
		vc = vn;
		i_cur = i_nxt;
		i_nxt = i_after;
		// past the last item: the current one again (its planes are fetched
		// and never used)
		if(i_nxt < items) nxt = locate(i_nxt);
		vn = voffsets(nxt);
	}
	store_prev();
}

// Recombination of one piece: outputs r0 .. r0 + n_rows, words w0 .. w0 +
// 32 wbk (`out` already offset to row r0 and word w0) from parity planes of
// plane_wbk 32-word blocks (wbk of them real).
struct BlmCombinePiece {
	const uint32_t* par;  // [t][plane_wbk][r_pad]
	uint8_t* out;
	uint64_t slot_bytes;
	uint32_t r_pad, wbk, plane_wbk, n_rows, accumulate, pad;
};

// Recovery words from the parity bits of piece `one`, or of piece blockIdx.z
// of `table`: a block of 32 output rows x 8 32-word blocks, rows varying
// fastest within a wave (coalesced parity reads); the 16 result bit-slices come
// from the 81 parity masks by the Karatsuba recombination and the reduction
// (bl_recombine, bl_reduce_slices: XORs fixed at compile time).  Words 2i and
// 2i + 1 go to one dword: a 32 x 32 bit transpose of the slices' even bits
// (rows 0 .. 15) and odd bits shifted down (rows 16 .. 31) leaves that dword
// in row 2i.  Each thread stores its 64 contiguous bytes as four 16-byte
// stores (slots are 1 KiB multiples, chunks 512-word multiples).
__global__ __launch_bounds__(256) void k_blm_combine(BlmCombinePiece one, const BlmCombinePiece* __restrict__ table) {
	const BlmCombinePiece P = table ? wave_uniform(table[blockIdx.z]) : one;
	const unsigned r = blockIdx.x * 32 + (threadIdx.x & 31);
	if(r >= P.n_rows) return;
	const size_t plane = (size_t)P.plane_wbk * P.r_pad;
	for(size_t wb = (size_t)blockIdx.y * 8 + (threadIdx.x >> 5); wb < P.wbk; wb += (size_t)gridDim.y * 8) {
		const __attribute__((address_space(1))) uint32_t* src =
		    (const __attribute__((address_space(1))) uint32_t*)(P.par + wb * P.r_pad + r);
		uint32_t pm[kBl];
#pragma unroll
		for(unsigned t = 0; t < kBl; t++) {
			pm[t] = *src;
			src += plane;
		}
		uint32_t acc[31];
		bl_recombine<16>(pm, acc);
		bl_reduce_slices(acc);
		uint32_t a[32];
#pragma unroll
		for(int j = 0; j < 16; j++) {
			a[j] = acc[j] & 0x55555555u;
			a[16 + j] = (acc[j] >> 1) & 0x55555555u;
		}
		blm_transpose32(a);
		uint4* dst = reinterpret_cast<uint4*>(P.out + r * P.slot_bytes + 64 * wb);
#pragma unroll
		for(int q = 0; q < 4; q++) {
			uint4 o = make_uint4(a[8 * q], a[8 * q + 2], a[8 * q + 4], a[8 * q + 6]);
			if(P.accumulate) {
				const uint4 d = dst[q];
				o = make_uint4(o.x ^ d.x, o.y ^ d.y, o.z ^ d.z, o.w ^ d.w);
			}
			dst[q] = o;
		}
	}
}

}  // namespace pph

// ------------------------------------------------------------------ driver

namespace pph {
namespace {

constexpr size_t kBlmCoefBudget = 16ull << 30;   // coefficient planes per output block
constexpr size_t kBlmChunkBudget = 16ull << 30;  // word planes + parity bits per word chunk

size_t blm_env(const char* name) {
	const char* v = getenv(name);
	return v ? (size_t)strtoull(v, nullptr, 10) : 0;
}

struct BlmShape {
	size_t kw_count, rt, r_blk, wc, slot_words;
	unsigned row_fast;  // item order: row tiles (1) or word tiles (0) vary fastest
	size_t cb_bytes, wb_bytes, p_bytes, total;  // total: + kBlmCounterBytes (the GEMM's item counter)
};
constexpr size_t kBlmCounterBytes = 256;

// Row tile: 128 where it pads the outputs less than 256 does.  With cap > 0
// the whole workspace stays within cap bytes: the coefficient block takes at
// most a quarter of it (at least one row tile), the word chunk what is left
// (at least one word tile); total is SIZE_MAX when even that does not fit.
BlmShape blm_shape(size_t slot_bytes, unsigned n_in, unsigned n_out, size_t cap) {
	BlmShape s{};
	// whole groups of four 64-input steps, at least eight steps
	s.kw_count = std::max<size_t>(kBlmMinKw, round_up(n_in, kBlmKStep) / 32);
	s.slot_words = slot_bytes / 2;                // a multiple of 512
	s.rt = blm_row_tile(n_out);
	if(size_t v = blm_env("PPH_BILINEAR_ROW_TILE")) s.rt = v == 64 ? 64 : v == 128 ? 128 : 256;
	// Items run rows-fastest by default: the 32 items one XCD holds at a time
	// then share 4 coefficient tiles and 8 word tiles in its L2 (words-fastest:
	// 2 and 16); 32768 inputs x 64 KiB, 8192 recovery slices ran 2 % faster.
	// PPH_BILINEAR_ORDER=2 runs word tiles fastest.
	s.row_fast = blm_env("PPH_BILINEAR_ORDER") == 2 ? 0u : 1u;
	const size_t per_row = kBl * s.kw_count * 4;
	size_t rb = std::max<size_t>(s.rt, kBlmCoefBudget / per_row / s.rt * s.rt);
	if(size_t c = blm_env("PPH_BILINEAR_ROWS")) rb = std::min(rb, round_up(c, s.rt));
	if(cap) rb = std::min(rb, std::max<size_t>(s.rt, cap / 4 / per_row / s.rt * s.rt));
	s.r_blk = std::min(rb, round_up(n_out, s.rt));
	const size_t per_word = kBl * s.kw_count * 4 + kBl * s.r_blk / 8;
	const size_t wt = std::max<size_t>(kBlmWordTile, 65536 / s.rt);  // chunks are multiples of the word tile
	size_t wc = std::max<size_t>(wt, kBlmChunkBudget / per_word / wt * wt);
	if(size_t c = blm_env("PPH_BILINEAR_CHUNK_WORDS")) wc = std::min(wc, round_up(c, wt));
	s.cb_bytes = round_up(kBl * s.kw_count * s.r_blk * 4, 256);
	if(cap) {
		const size_t used = s.cb_bytes + 512 + kBlmCounterBytes;  // (two more regions rounded to 256, the counter)
		wc = std::min(wc, cap > used ? (cap - used) / per_word / wt * wt : 0);
		if(wc < wt) {
			s.total = SIZE_MAX;
			return s;
		}
	}
	wc = std::min(wc, round_up(s.slot_words, wt));
	const size_t chunks = (s.slot_words + wc - 1) / wc;
	s.wc = round_up((s.slot_words + chunks - 1) / chunks, wt);
	s.wb_bytes = round_up(kBl * s.kw_count * s.wc * 4, 256);
	s.p_bytes = round_up(kBl * s.r_blk * (s.wc / 32) * 4, 256);
	s.total = s.cb_bytes + s.wb_bytes + s.p_bytes + kBlmCounterBytes;
	return s;
}

}  // namespace

size_t blm_workspace_bytes(size_t slot_bytes, unsigned n_in, unsigned n_out, size_t cap) {
	return blm_shape(slot_bytes, n_in, n_out, cap).total;
}

namespace {

unsigned blm_blocks(size_t threads, const DeviceTables& tab) {
	const size_t cap = (size_t)(tab.cus > 0 ? tab.cus : 256) * 16;
	return (unsigned)std::max<size_t>(1, std::min(cap, (threads + 255) / 256));
}

// a plane of one component is addressed through a buffer resource
bool blm_planes_fit(size_t kw_count, size_t wc, size_t r_pad) {
	const size_t lim = 1ull << 31;
	return kw_count * wc * 4 < lim && kw_count * r_pad * 4 < lim && (wc / 32) * r_pad * 4 < lim;
}

// counter (optional): zeroed before the launch, items handed out dynamically
hipError_t blm_launch_gemm(unsigned rt, const BlmPiece& one, const BlmPiece* table, unsigned items,
                           const DeviceTables& tab, hipStream_t s, unsigned* counter = nullptr) {
	if(!items) return hipSuccess;
	const unsigned grid = std::min<unsigned>(items, (unsigned)(tab.cus > 0 ? tab.cus : 256));  // one per CU
	if(rt == 64)
		hipLaunchKernelGGL(k_blm_gemm<64>, dim3(grid), dim3(256), 0, s, one, table, items, counter);
	else if(rt == 128)
		hipLaunchKernelGGL(k_blm_gemm<128>, dim3(grid), dim3(256), 0, s, one, table, items, counter);
	else
		hipLaunchKernelGGL(k_blm_gemm<256>, dim3(grid), dim3(256), 0, s, one, table, items, counter);
	return hipGetLastError();
}

// n counters to zero (a kernel: no memset node in a captured graph)
__global__ void k_blm_zero(unsigned* p, unsigned n) {
	for(unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

dim3 blm_combine_grid(size_t rows, size_t wbk, unsigned pieces) {
	return dim3((unsigned)((rows + 31) / 32), (unsigned)std::min<size_t>((wbk + 7) / 8, 4096), pieces);
}

}  // namespace

int blm_enqueue(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                const uint16_t* exps, void* out_slots, unsigned n_out, bool accumulate, void* ws, size_t ws_bytes,
                const DeviceTables& tab, hipStream_t s, size_t cap, const uint32_t* coef, unsigned coef_stride) {
	if(!n_out) return PPH_OK;
	const BlmShape sh = blm_shape(slot_bytes, n_in, n_out, cap);
	if(!ws || sh.total == SIZE_MAX || ws_bytes < sh.total) return PPH_E_INVALID;
	if(!blm_planes_fit(sh.kw_count, sh.wc, round_up(sh.r_blk, sh.rt))) return PPH_E_INVALID;
	uint32_t* cb = static_cast<uint32_t*>(ws);
	uint32_t* wb = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(ws) + sh.cb_bytes);
	uint32_t* par = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(ws) + sh.cb_bytes + sh.wb_bytes);
	// the GEMM hands out its items from a counter (zeroed before each launch):
	// workgroups on CUs that run slower take fewer items (32768 x 64 KiB,
	// 8192 recovery slices: 192.6 ms against 203.4 with a fixed split)
	unsigned* counter = reinterpret_cast<unsigned*>(static_cast<uint8_t*>(ws) + sh.cb_bytes + sh.wb_bytes + sh.p_bytes);
	for(size_t r0 = 0; r0 < n_out; r0 += sh.r_blk) {
		const unsigned rows = (unsigned)std::min<size_t>(sh.r_blk, n_out - r0);
		const size_t r_pad = round_up(rows, sh.rt);
		const BlmCoefJob cj{cb,
		                    input_logs,
		                    coef ? nullptr : exps + r0,
		                    coef ? coef + r0 : nullptr,
		                    (uint32_t)sh.kw_count,
		                    (uint32_t)r_pad,
		                    n_in,
		                    rows,
		                    coef_stride,
		                    0u};
		hipLaunchKernelGGL(k_blm_coef, dim3(blm_blocks(sh.kw_count * r_pad, tab)), dim3(256), 0, s, cj,
		                   (const BlmCoefJob*)nullptr, tab.exp);
		hipError_t e = hipGetLastError();
		if(e != hipSuccess) return set_error(e);
		for(size_t w0 = 0; w0 < sh.slot_words; w0 += sh.wc) {
			const size_t wv = std::min(sh.wc, sh.slot_words - w0);          // a multiple of 512
			const size_t wc = round_up(wv, std::max<size_t>(kBlmWordTile, 65536 / sh.rt));  // <= sh.wc
			const BlmWordsPiece wp{wb, (const uint8_t*)in_slots + 2 * w0, slot_bytes, (uint32_t)sh.kw_count,
			                       (uint32_t)wc, (uint32_t)wv, n_in};
			hipLaunchKernelGGL(k_blm_words, dim3(blm_blocks(sh.kw_count * wc / 2, tab)), dim3(256), 0, s, wp,
			                   (const BlmWordsPiece*)nullptr);
			const unsigned w_tiles = (unsigned)(wc / (65536 / sh.rt)), r_tiles = (unsigned)(r_pad / sh.rt);
			const unsigned items = kBl * w_tiles * r_tiles;
			const BlmPiece piece{wb,
			                     cb,
			                     par,
			                     (uint32_t)(sh.kw_count / 2),
			                     (uint32_t)wc,
			                     (uint32_t)r_pad,
			                     0u,
			                     w_tiles,
			                     r_tiles,
			                     0u,
			                     items,
			                     sh.row_fast,
			                     blm_magic(w_tiles * r_tiles),
			                     blm_magic(sh.row_fast ? r_tiles : w_tiles)};
			e = hipGetLastError();
			if(e == hipSuccess) {
				hipLaunchKernelGGL(k_blm_zero, dim3(1), dim3(64), 0, s, counter, 1u);
				e = hipGetLastError();
			}
			if(e == hipSuccess) e = blm_launch_gemm((unsigned)sh.rt, piece, nullptr, items, tab, s, counter);
			const BlmCombinePiece cp{par,           (uint8_t*)out_slots + r0 * slot_bytes + 2 * w0,
			                         slot_bytes,    (uint32_t)r_pad,
			                         (uint32_t)(wv / 32), (uint32_t)(wc / 32),
			                         rows,          accumulate ? 1u : 0u,
			                         0u};
			if(e == hipSuccess)
				hipLaunchKernelGGL(k_blm_combine, blm_combine_grid(rows, wv / 32, 1), dim3(256), 0, s, cp,
				                   (const BlmCombinePiece*)nullptr);
			if(e == hipSuccess) e = hipGetLastError();
			if(e != hipSuccess) return set_error(e);
		}
	}
	return PPH_OK;
}

// ------------------------------------------------------------ many jobs

// The bilinear jobs of a batch as one sequence of launches: every job's
// coefficient planes in one launch (one output block per job), then waves of
// pieces (one word chunk of one job each) whose word planes and parity bits
// fit the workspace together: per wave one word-plane launch, one GEMM launch
// per row tile over a cross-job item table, and one recombination launch.
// Pieces run longest first (input count descending), so that a wave's last
// items are its shortest.
//
// Overlapped form (two halves of the workspace, waves alternating between
// them): the GEMMs run back to back on the caller's stream while a second
// stream writes the next wave's word planes and recombines the previous
// wave's parity bits.  Those kernels are memory-bound and the GEMM is
// MFMA-bound, and they fit on a CU beside a GEMM workgroup (which holds one
// wave of ~380 registers per SIMD): the word planes run as about one
// workgroup per two CUs for that reason.
struct BlmBatch {
	std::vector<BlmCoefJob> coef;
	struct Wave {
		unsigned words0, words_n;     // slice of the word-piece table
		size_t words_threads;         // largest piece's thread count
		unsigned gemm0[3], gemm_n[3]; // slices of the GEMM table (row tile 256, 128, 64)
		unsigned items[3];
		unsigned comb0, comb_n;
		size_t comb_rows, comb_wbk;   // largest piece's
	};
	std::vector<Wave> waves;
	size_t coef_threads = 0;
	BlmCoefJob* d_coef = nullptr;
	BlmWordsPiece* d_words = nullptr;
	BlmPiece* d_gemm = nullptr;
	BlmCombinePiece* d_comb = nullptr;
	unsigned* d_counters = nullptr;  // item counters of the GEMM launches (3 per wave)
	// input preparation per wave (ck > 0): the jobs whose first piece is in
	// the wave, entries prep0[k] .. prep0[k] + prep_n[k] of d_prep, slots in
	// all; output verification per wave: the jobs whose last piece is in it,
	// entries ver0[k] .. of d_prep, the largest output count
	unsigned ck = 0;
	CkJob* d_prep = nullptr;
	std::vector<unsigned> prep0, prep_n, prep_slots, ver0, ver_n, ver_max;
	bool overlap = false;
	hipStream_t side = nullptr;
	hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_words[2] = {}, ev_gemm[2] = {};
};

namespace {

// A job's rows are covered by row tiles of several heights (blm_row_split:
// 256-row tiles first, then 128 and 64), each height in its own GEMM launch;
// plane rows are multiples of the largest word tile used.
struct BlmJobShape {
	size_t kw_count, wt, r_pad, slot_words, cb_bytes;  // wt: word tile, plane rows are multiples
	unsigned tiles[3];                                 // row tiles of 256, 128, 64 rows
};

BlmJobShape blm_job_shape(const BlmBatchJob& j) {
	BlmJobShape s{};
	s.kw_count = std::max<size_t>(kBlmMinKw, round_up(j.n_in, kBlmKStep) / 32);
	const BlmRowSplit rs = blm_row_split(j.n_out);
	s.wt = kBlmWordTile;
	for(unsigned k = 0; k < 3; k++) {
		s.tiles[k] = rs.n[k];
		if(rs.n[k]) s.wt = std::max<size_t>(s.wt, 65536 / (256u >> k));
	}
	s.r_pad = rs.rows;
	s.slot_words = j.slot_bytes / 2;
	s.cb_bytes = round_up(kBl * s.kw_count * s.r_pad * 4, 256);
	return s;
}

// Word chunk of a piece (at most): 65536 words, or PPH_BATCH_PIECE_WORDS
// rounded up to 512 (a test hook: many pieces from small jobs).
size_t blm_piece_words() {
	const size_t v = blm_env("PPH_BATCH_PIECE_WORDS");
	return v ? round_up(v, kBlmWordTile) : 65536;
}

// planes of a piece of wv words (its rows padded to the word tile)
size_t blm_piece_bytes(const BlmJobShape& s, size_t wv) {
	const size_t wc = round_up(wv, s.wt);
	return round_up(kBl * s.kw_count * wc * 4, 256) + round_up(kBl * (wc / 32) * s.r_pad * 4, 256);
}

}  // namespace

bool blm_batch_eligible(const BlmBatchJob& j, size_t ws_budget) {
	if(!j.n_in || !j.n_out || j.n_in > 32768 || j.slot_bytes % kSlotAlign) return false;
	const BlmJobShape s = blm_job_shape(j);
	const size_t wc = std::min(blm_piece_words(), s.slot_words);
	// coefficient planes at most a quarter of the workspace, a piece at most a quarter
	return s.cb_bytes <= ws_budget / 4 && blm_piece_bytes(s, wc) <= ws_budget / 4 &&
	       blm_planes_fit(s.kw_count, round_up(wc, s.wt), s.r_pad);
}

size_t blm_batch_workspace(const BlmBatchJob* jobs, unsigned n, size_t ws_budget) {
	size_t cb = 0, big = 0, all = 0;
	const size_t pw = blm_piece_words();
	for(unsigned k = 0; k < n; k++) {
		const BlmJobShape s = blm_job_shape(jobs[k]);
		cb += s.cb_bytes;
		for(size_t w0 = 0; w0 < s.slot_words; w0 += pw) {
			const size_t pb = blm_piece_bytes(s, std::min(pw, s.slot_words - w0));
			big = std::max(big, pb);
			all += pb;
		}
	}
	if(!n) return 0;
	// every piece at once when that fits, else the budget (at least the
	// coefficients and one piece, two for the overlapped form)
	return std::max(cb + (all > big ? 2 : 1) * big, std::min(cb + all, ws_budget));
}

void blm_batch_destroy(BlmBatch* b) {
	if(!b) return;
	if(b->side) (void)hipStreamSynchronize(b->side);
	if(b->d_coef) (void)hipFree(b->d_coef);
	if(b->d_words) (void)hipFree(b->d_words);
	if(b->d_gemm) (void)hipFree(b->d_gemm);
	if(b->d_comb) (void)hipFree(b->d_comb);
	if(b->d_counters) (void)hipFree(b->d_counters);
	if(b->d_prep) (void)hipFree(b->d_prep);
	for(hipEvent_t e : {b->ev_start, b->ev_end, b->ev_words[0], b->ev_words[1], b->ev_gemm[0], b->ev_gemm[1]})
		if(e) (void)hipEventDestroy(e);
	if(b->side) (void)hipStreamDestroy(b->side);
	delete b;
}

int blm_batch_create(const BlmBatchJob* jobs, unsigned n, void* ws, size_t ws_bytes, unsigned ck, BlmBatch** out) {
	*out = nullptr;
	std::unique_ptr<BlmBatch> b(new(std::nothrow) BlmBatch);
	if(!b) return PPH_E_NOMEM;
	b->ck = ck;
	uint8_t* base = static_cast<uint8_t*>(ws);
	size_t off = 0;
	std::vector<BlmJobShape> sh(n);
	std::vector<uint32_t*> cb(n);
	for(unsigned k = 0; k < n; k++) {
		sh[k] = blm_job_shape(jobs[k]);
		cb[k] = reinterpret_cast<uint32_t*>(base + off);
		b->coef.push_back(BlmCoefJob{cb[k], jobs[k].in_logs, jobs[k].exps, nullptr, (uint32_t)sh[k].kw_count,
		                             (uint32_t)sh[k].r_pad, jobs[k].n_in, jobs[k].n_out, 0u, 0u});
		b->coef_threads = std::max(b->coef_threads, sh[k].kw_count * sh[k].r_pad);
		off += sh[k].cb_bytes;
	}
	if(off > ws_bytes) return PPH_E_INVALID;
	struct Piece {
		unsigned job;
		size_t w0, wc;
	};
	std::vector<Piece> pieces;
	const size_t pw = blm_piece_words();
	for(unsigned k = 0; k < n; k++)
		for(size_t w0 = 0; w0 < sh[k].slot_words; w0 += pw)
			pieces.push_back(Piece{k, w0, std::min(pw, sh[k].slot_words - w0)});
	std::stable_sort(pieces.begin(), pieces.end(),
	                 [&](const Piece& x, const Piece& y) { return sh[x.job].kw_count > sh[y.job].kw_count; });
	// Overlapped form: waves alternate between two halves of the piece area
	// (when each half holds the largest piece and there are two pieces or more).
	size_t area = ws_bytes - off, big = 0;
	for(const Piece& pc : pieces) big = std::max(big, blm_piece_bytes(sh[pc.job], pc.wc));
	const char* ov = getenv("PPH_BATCH_OVERLAP");
	b->overlap = pieces.size() >= 2 && area / 2 / 256 * 256 >= big && !(ov && ov[0] == '0');
	if(b->overlap) area = area / 2 / 256 * 256;
	std::vector<unsigned> first_wave(n, ~0u), last_wave(n, 0u);
	std::vector<BlmWordsPiece> words;
	std::vector<BlmPiece> gemm[3];
	std::vector<BlmCombinePiece> comb;
	std::vector<BlmPiece> gemm_all;
	size_t p = 0;
	while(p < pieces.size()) {
		BlmBatch::Wave w{};
		w.words0 = (unsigned)words.size();
		w.comb0 = (unsigned)comb.size();
		for(auto& g : gemm) g.clear();
		const size_t half = b->overlap ? (b->waves.size() % 2) * area : 0;
		// Overlapped form: the first two waves fill a quarter and a half of their
		// area (at least one piece each), so that the word planes of the first
		// wave, which nothing runs beside, are short, and those of the second
		// still fit under the first wave's GEMMs.
		const size_t fill = !b->overlap || b->waves.size() > 1 ? area : b->waves.empty() ? area / 4 : area / 2;
		size_t used = 0;
		for(; p < pieces.size(); p++) {
			const Piece& pc = pieces[p];
			const BlmJobShape& s = sh[pc.job];
			const size_t wc = round_up(pc.wc, s.wt);  // plane row: pc.wc words, padded to the word tile
			const size_t wb_bytes = round_up(kBl * s.kw_count * wc * 4, 256);
			const size_t need = blm_piece_bytes(s, pc.wc);
			if(used + need > (used ? fill : area)) {
				if(!used) return PPH_E_INVALID;  // a piece larger than the area
				break;
			}
			uint32_t* wbp = reinterpret_cast<uint32_t*>(base + off + half + used);
			uint32_t* par = reinterpret_cast<uint32_t*>(base + off + half + used + wb_bytes);
			used += need;
			const BlmBatchJob& j = jobs[pc.job];
			if(first_wave[pc.job] == ~0u) first_wave[pc.job] = (unsigned)b->waves.size();
			last_wave[pc.job] = (unsigned)b->waves.size();
			words.push_back(BlmWordsPiece{wbp, (const uint8_t*)j.in_slots + 2 * pc.w0, j.slot_bytes,
			                              (uint32_t)s.kw_count, (uint32_t)wc, (uint32_t)pc.wc, j.n_in});
			w.words_threads = std::max(w.words_threads, s.kw_count * wc / 2);
			unsigned r_base = 0;
			for(unsigned ti = 0; ti < 3; ti++) {
				const unsigned rt = 256u >> ti, r_tiles = s.tiles[ti];
				if(!r_tiles) continue;
				const unsigned w_tiles = (unsigned)(wc / (65536 / rt));
				const unsigned items = kBl * w_tiles * r_tiles;
				gemm[ti].push_back(BlmPiece{wbp, cb[pc.job], par, (uint32_t)(s.kw_count / 2), (uint32_t)wc,
				                            (uint32_t)s.r_pad, r_base, w_tiles, r_tiles, w.items[ti], items, 1u,
				                            blm_magic(w_tiles * r_tiles), blm_magic(r_tiles)});
				w.items[ti] += items;
				r_base += rt * r_tiles;
			}
			comb.push_back(BlmCombinePiece{par, (uint8_t*)j.out_slots + 2 * pc.w0, j.slot_bytes, (uint32_t)s.r_pad,
			                               (uint32_t)(pc.wc / 32), (uint32_t)(wc / 32), j.n_out, 0u, 0u});
			w.comb_rows = std::max<size_t>(w.comb_rows, j.n_out);
			w.comb_wbk = std::max<size_t>(w.comb_wbk, pc.wc / 32);
		}
		w.words_n = (unsigned)words.size() - w.words0;
		w.comb_n = (unsigned)comb.size() - w.comb0;
		for(unsigned ti = 0; ti < 3; ti++) {
			w.gemm0[ti] = (unsigned)gemm_all.size();
			w.gemm_n[ti] = (unsigned)gemm[ti].size();
			gemm_all.insert(gemm_all.end(), gemm[ti].begin(), gemm[ti].end());
		}
		b->waves.push_back(w);
	}
	// preparation tables, one per wave (largest slots first), slots numbered
	// within each table
	std::vector<CkJob> prep;
	if(ck) {
		for(size_t k = 0; k < b->waves.size(); k++) {
			std::vector<CkJob> t;
			for(unsigned q = 0; q < n; q++) {
				if(first_wave[q] != k) continue;
				const BlmBatchJob& j = jobs[q];
				const CkJob c{(uint8_t*)j.in_slots, j.in_lens, nullptr, j.slot_bytes, j.slice_size,
				              ck_offset(j.slice_size), j.n_in, 0u};
				if(!ck_job_valid(c, ck)) return PPH_E_INVALID;
				t.push_back(c);
			}
			std::stable_sort(t.begin(), t.end(),
			                 [](const CkJob& x, const CkJob& y) { return x.slot_bytes > y.slot_bytes; });
			unsigned slots = 0;
			for(CkJob& c : t) {
				c.first = slots;
				slots += c.count;
			}
			b->prep0.push_back((unsigned)prep.size());
			b->prep_n.push_back((unsigned)t.size());
			b->prep_slots.push_back(slots);
			prep.insert(prep.end(), t.begin(), t.end());
		}
		for(size_t k = 0; k < b->waves.size(); k++) {
			unsigned cnt = 0, mx = 0;
			b->ver0.push_back((unsigned)prep.size());
			for(unsigned q = 0; q < n; q++) {
				if(last_wave[q] != k || !jobs[q].n_out) continue;
				const BlmBatchJob& j = jobs[q];
				const CkJob c{(uint8_t*)j.out_slots, nullptr, j.valid, j.slot_bytes, j.slice_size, ck_offset(j.slice_size),
				              j.n_out, 0u};
				if(!j.valid || !ck_job_valid(c, ck)) return PPH_E_INVALID;
				prep.push_back(c);
				cnt++;
				mx = std::max(mx, j.n_out);
			}
			b->ver_n.push_back(cnt);
			b->ver_max.push_back(mx);
		}
	}
	auto upload = [](auto** d, const auto& v) -> hipError_t {
		if(v.empty()) return hipSuccess;
		const size_t bytes = v.size() * sizeof(v[0]);
		hipError_t e = hipMalloc((void**)d, bytes);
		if(e == hipSuccess) e = hipMemcpy(*d, v.data(), bytes, hipMemcpyHostToDevice);
		return e;
	};
	hipError_t e = upload(&b->d_coef, b->coef);
	if(e == hipSuccess) e = upload(&b->d_words, words);
	if(e == hipSuccess) e = upload(&b->d_gemm, gemm_all);
	if(e == hipSuccess) e = upload(&b->d_comb, comb);
	if(e == hipSuccess) e = upload(&b->d_prep, prep);
	if(e == hipSuccess) e = hipMalloc((void**)&b->d_counters, 3 * b->waves.size() * sizeof(unsigned));
	if(b->overlap) {
		if(e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
		for(hipEvent_t* ev : {&b->ev_start, &b->ev_end, &b->ev_words[0], &b->ev_words[1], &b->ev_gemm[0], &b->ev_gemm[1]})
			if(e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
	}
	if(e != hipSuccess) {
		blm_batch_destroy(b.release());
		return e == hipErrorOutOfMemory ? PPH_E_NOMEM : set_error(e);
	}
	*out = b.release();
	return PPH_OK;
}

int blm_batch_enqueue(const BlmBatch* b, const DeviceTables& tab, hipStream_t s) {
	if(!b) return PPH_E_INVALID;
	if(b->coef.empty()) return PPH_OK;
	const BlmCoefJob none_c{};
	const BlmWordsPiece none_w{};
	const BlmPiece none_g{};
	const BlmCombinePiece none_m{};
	hipLaunchKernelGGL(k_blm_coef, dim3(blm_blocks(b->coef_threads, tab), (unsigned)b->coef.size()), dim3(256), 0, s,
	                   none_c, b->d_coef, tab.exp);
	// verification of the jobs that end in wave k (after its recombination)
	auto verify = [&](size_t k, hipStream_t st) -> hipError_t {
		if(!b->ck || !b->ver_n[k]) return hipSuccess;
		return launch_verify_jobs(b->d_prep + b->ver0[k], b->ver_n[k], b->ver_max[k], b->ck, tab, st);
	};
	const unsigned n_counters = 3 * (unsigned)b->waves.size();
	hipLaunchKernelGGL(k_blm_zero, dim3((n_counters + 255) / 256), dim3(256), 0, s, b->d_counters, n_counters);
	hipError_t e = hipGetLastError();
	if(b->overlap) {
		// Caller's stream: GEMM(k) after the word planes of wave k.  Side stream:
		// words(0); then per wave k: words(k + 1) (its half was last read by
		// GEMM(k - 1) and recombination(k - 1), both before it in this order),
		// recombination(k) after GEMM(k).  The side stream joins back at the end.
		const unsigned cus = (unsigned)(tab.cus > 0 ? tab.cus : 256);
		// wave 0 runs alone (every CU); later waves run beside a GEMM, as a
		// workgroup per two CUs in all (the 64-job workload: 465-468 ms per
		// replay, against 473-483 with one per CU, 478-481 with two and
		// 473-478 with one per four; the GEMM hands out its items
		// dynamically, so the CUs that host them take fewer items)
		// preparation of the jobs that start in wave k: on every CU before wave
		// 0, beside a GEMM (one workgroup per CU) later
		auto prep = [&](size_t k, unsigned per_cu) -> hipError_t {
			if(!b->ck || !b->prep_n[k]) return hipSuccess;
			return launch_prep_data_jobs(b->d_prep + b->prep0[k], b->prep_n[k], b->prep_slots[k], b->ck, per_cu, tab,
			                             b->side);
		};
		auto words = [&](const BlmBatch::Wave& w) {
			const unsigned full = blm_blocks(w.words_threads, tab);
			const unsigned gx = &w == &b->waves[0]
			                        ? full
			                        : std::min(full, std::max(1u, (cus + 2 * w.words_n - 1) / (2 * w.words_n)));
			hipLaunchKernelGGL(k_blm_words, dim3(gx, w.words_n), dim3(256), 0, b->side, none_w, b->d_words + w.words0);
			return hipGetLastError();
		};
		if(e == hipSuccess) e = hipEventRecord(b->ev_start, s);
		if(e == hipSuccess) e = hipStreamWaitEvent(b->side, b->ev_start, 0);
		if(e == hipSuccess) e = prep(0, 8);
		if(e == hipSuccess) e = words(b->waves[0]);
		if(e == hipSuccess) e = hipEventRecord(b->ev_words[0], b->side);
		for(size_t k = 0; k < b->waves.size() && e == hipSuccess; k++) {
			const BlmBatch::Wave& w = b->waves[k];
			e = hipStreamWaitEvent(s, b->ev_words[k % 2], 0);
			for(unsigned ti = 0; ti < 3 && e == hipSuccess; ti++)
				if(w.gemm_n[ti])
					e = blm_launch_gemm(256u >> ti, none_g, b->d_gemm + w.gemm0[ti], w.items[ti], tab, s,
					                    b->d_counters + 3 * k + ti);
			if(e == hipSuccess) e = hipEventRecord(b->ev_gemm[k % 2], s);
			if(e == hipSuccess && k + 1 < b->waves.size()) {
				e = prep(k + 1, 1);
				if(e == hipSuccess) e = words(b->waves[k + 1]);
				if(e == hipSuccess) e = hipEventRecord(b->ev_words[(k + 1) % 2], b->side);
			}
			if(e == hipSuccess) e = hipStreamWaitEvent(b->side, b->ev_gemm[k % 2], 0);
			if(e == hipSuccess) {
				hipLaunchKernelGGL(k_blm_combine, blm_combine_grid(w.comb_rows, w.comb_wbk, w.comb_n), dim3(256), 0,
				                   b->side, none_m, b->d_comb + w.comb0);
				e = hipGetLastError();
			}
			if(e == hipSuccess) e = verify(k, b->side);
		}
		// join even after an error, so that a capture on `s` can end
		const hipError_t j = hipEventRecord(b->ev_end, b->side);
		const hipError_t j2 = j == hipSuccess ? hipStreamWaitEvent(s, b->ev_end, 0) : j;
		if(e == hipSuccess) e = j2;
		return e == hipSuccess ? PPH_OK : set_error(e);
	}
	for(size_t k = 0; k < b->waves.size(); k++) {
		const BlmBatch::Wave& w = b->waves[k];
		if(e != hipSuccess) break;
		if(b->ck && b->prep_n[k]) {
			e = launch_prep_data_jobs(b->d_prep + b->prep0[k], b->prep_n[k], b->prep_slots[k], b->ck, 8, tab, s);
			if(e != hipSuccess) break;
		}
		hipLaunchKernelGGL(k_blm_words, dim3(blm_blocks(w.words_threads, tab), w.words_n), dim3(256), 0, s, none_w,
		                   b->d_words + w.words0);
		e = hipGetLastError();
		for(unsigned ti = 0; ti < 3 && e == hipSuccess; ti++)
			if(w.gemm_n[ti])
				e = blm_launch_gemm(256u >> ti, none_g, b->d_gemm + w.gemm0[ti], w.items[ti], tab, s,
				                    b->d_counters + 3 * k + ti);
		if(e == hipSuccess) {
			hipLaunchKernelGGL(k_blm_combine, blm_combine_grid(w.comb_rows, w.comb_wbk, w.comb_n), dim3(256), 0, s,
			                   none_m, b->d_comb + w.comb0);
			e = hipGetLastError();
		}
		if(e == hipSuccess) e = verify(k, s);
	}
	return e == hipSuccess ? PPH_OK : set_error(e);
}

}  // namespace pph
