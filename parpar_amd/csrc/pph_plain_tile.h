// The plain-form mul-add tile, shared by the standalone kernel and the
// persistent batched kernel.
#pragma once
#include "pph_internal.h"

namespace pph {

// out[r] (^)= sum_i c(r,i) * in[i] in GF(2^16).  Inputs are in log domain.
// Each lane owns 8 consecutive words of a 8192-word column tile and G recovery
// slices; the full 64 Ki-entry exp table sits in LDS (entry 65535 = 0), so one
// product is   exp[min(lx + lc, lx + lc - 65535, 65535)]   with doubled
// (byte-address) logs: two adds and a min3, no branch for zero words.
// Coefficients are uniform across the wavefront: either computed from the
// input log and the recovery exponent on the scalar unit (PAR2 coefficients,
// gf16/gfmat_coeff.c:60-77) or read from an explicit table.
__device__ __forceinline__ uint32_t mod65535(uint32_t x) {
	x = (x & 0xFFFFu) + (x >> 16);
	x = (x & 0xFFFFu) + (x >> 16);
	return x;  // <= 65535, congruent; 65535 acts as 0 in the lookup below
}

__device__ __forceinline__ uint32_t exp_pair(const uint16_t* s_exp, uint32_t a, uint32_t b) {
	uint32_t lo = *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(s_exp) + a);
	uint32_t hi = *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(s_exp) + b);
	return lo | (hi << 16);
}

// Load the exp table into LDS (kPlainThreads threads).
__device__ __forceinline__ void load_exp_table(uint16_t* s_exp, const uint16_t* __restrict__ exp_tab) {
	const uint4* src = reinterpret_cast<const uint4*>(exp_tab);
	uint4* dst = reinterpret_cast<uint4*>(s_exp);
	for(unsigned k = threadIdx.x; k < 65536 * 2 / 16; k += kPlainThreads) dst[k] = src[k];
}

// One work item: column tile `col` (8192 words) x recovery slices [G*grp, G*grp+G),
// over inputs [i0, i1).  partial: the item's other inputs are summed by other
// workgroups, so the sums are XORed into `out` with atomics (which then holds
// zeros or, accumulating, the previous output); otherwise they are stored
// (accumulate: XORed with the previous output).  s_lin: LDS for
// kPlainThreads input logs.  Every thread of the workgroup calls it (it
// synchronises the workgroup).
template <int G, bool TABLE>
__device__ __forceinline__ void plain_tile(const uint16_t* s_exp, uint16_t* s_lin, const uint16_t* __restrict__ in,
                                           size_t slot_words,
                                           unsigned n_in, const uint16_t* __restrict__ in_logs,
                                           const uint16_t* __restrict__ exps, const uint32_t* __restrict__ coef,
                                           unsigned coef_stride, uint16_t* __restrict__ out, unsigned n_out,
                                           int accumulate, unsigned col, unsigned grp, unsigned i0 = 0,
                                           unsigned i1 = ~0u, bool partial = false) {
	if(i1 > n_in) i1 = n_in;
	const size_t w0 = (size_t)col * (kPlainThreads * kPlainWordsPerLane) + (size_t)threadIdx.x * kPlainWordsPerLane;
	const bool active = w0 < slot_words;
	const unsigned r0 = grp * G;

	uint32_t e[G];
	if constexpr(!TABLE) {
#pragma unroll
		for(int r = 0; r < G; r++) e[r] = (r0 + r < n_out) ? exps[r0 + r] : 0u;
	}

	uint32_t acc[G][4];
#pragma unroll
	for(int r = 0; r < G; r++) {
		acc[r][0] = acc[r][1] = acc[r][2] = acc[r][3] = 0;
		if(accumulate && !partial && active && r0 + r < n_out) {
			uint4 v = *reinterpret_cast<const uint4*>(out + (size_t)(r0 + r) * slot_words + w0);
			acc[r][0] = v.x; acc[r][1] = v.y; acc[r][2] = v.z; acc[r][3] = v.w;
		}
	}

	// The next input's words are loaded one input ahead, and the input logs
	// come from LDS (staged kPlainThreads at a time), so that no wave waits on
	// a global load at the top of its input loop (a log read from global memory
	// into a scalar register waited for every load in flight).
	const uint4 kNone = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
	uint4 v_next = kNone;
	if(i0 < i1 && active) v_next = *reinterpret_cast<const uint4*>(in + (size_t)i0 * slot_words + w0);
	for(unsigned i = i0; i < i1; i++) {
		const uint4 v = v_next;
		if(i + 1 < i1 && active) v_next = *reinterpret_cast<const uint4*>(in + (size_t)(i + 1) * slot_words + w0);
		if constexpr(!TABLE) {
			if((i - i0) % kPlainThreads == 0) {
				__syncthreads();  // the previous chunk has been read by every wave
				const unsigned k = i + threadIdx.x;
				s_lin[threadIdx.x] = k < i1 ? in_logs[k] : 0;
				__syncthreads();
			}
		}
		const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
		uint32_t lx[8], lxm[8];
#pragma unroll
		for(int j = 0; j < 4; j++) {
			const uint32_t L0 = wv[j] & 0xFFFFu, L1 = wv[j] >> 16;
			lx[2 * j] = L0 == kLogZero ? kLogZeroDecoded : 2 * L0;
			lx[2 * j + 1] = L1 == kLogZero ? kLogZeroDecoded : 2 * L1;
			lxm[2 * j] = lx[2 * j] - kZeroIdx2;
			lxm[2 * j + 1] = lx[2 * j + 1] - kZeroIdx2;
		}
		uint32_t lin = 0;
		if constexpr(!TABLE) lin = __builtin_amdgcn_readfirstlane(s_lin[(i - i0) % kPlainThreads]);
#pragma unroll
		for(int r = 0; r < G; r++) {
			uint32_t lc;
			bool use;
			if constexpr(TABLE) {
				lc = coef[(size_t)i * coef_stride + r0 + r];
				use = lc != kSkipCoef;
			} else {
				// rows past n_out have exponent 0: their sums are never stored, and
				// without a branch per row the lookups of neighbouring rows overlap
				lc = 2 * mod65535(lin * e[r]);
				use = true;
			}
			if(use) {
#pragma unroll
				for(int j = 0; j < 4; j++) {
					const uint32_t a = min(min(lx[2 * j] + lc, lxm[2 * j] + lc), kZeroIdx2);
					const uint32_t b = min(min(lx[2 * j + 1] + lc, lxm[2 * j + 1] + lc), kZeroIdx2);
					acc[r][j] ^= exp_pair(s_exp, a, b);
				}
			}
		}
	}

	if(active) {
#pragma unroll
		for(int r = 0; r < G; r++) {
			if(r0 + r >= n_out) continue;
			uint32_t* o = reinterpret_cast<uint32_t*>(out + (size_t)(r0 + r) * slot_words + w0);
			if(partial) {
#pragma unroll
				for(int j = 0; j < 4; j++)
					if(acc[r][j]) atomicXor(o + j, acc[r][j]);
			} else {
				*reinterpret_cast<uint4*>(o) = make_uint4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
			}
		}
	}
}

}  // namespace pph
