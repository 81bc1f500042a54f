// Karatsuba expansion of GF(2^16) elements for the bilinear recovery form,
// computed bit-sliced on the device (shared by pph_bilinear.hip and
// pph_bl_mfma.hip).
#pragma once
#include <stdint.h>

namespace pph {

constexpr unsigned kBl = 81;  // rank of the four-level Karatsuba formula

// Bit-sliced expansion: s[b] holds bit b of 32 values (value k at a bit position
// of the caller's choosing); e[t] receives component t of the same 32 values at
// the same positions.  Components in recursive order: E(v) = E(lo) ++ E(hi) ++
// E(lo ^ hi) over the 8-bit halves, then the 4-, 2- and 1-bit ones.
__device__ __forceinline__ void bl_expand_slices(const uint32_t s[16], uint32_t e[kBl]) {
	uint32_t l1[24], l2[36], l3[54];
#pragma unroll
	for(int b = 0; b < 8; b++) {
		l1[b] = s[b];
		l1[8 + b] = s[8 + b];
		l1[16 + b] = s[b] ^ s[8 + b];
	}
#pragma unroll
	for(int k = 0; k < 3; k++)
#pragma unroll
		for(int b = 0; b < 4; b++) {
			l2[12 * k + b] = l1[8 * k + b];
			l2[12 * k + 4 + b] = l1[8 * k + 4 + b];
			l2[12 * k + 8 + b] = l1[8 * k + b] ^ l1[8 * k + 4 + b];
		}
#pragma unroll
	for(int k = 0; k < 9; k++)
#pragma unroll
		for(int b = 0; b < 2; b++) {
			l3[6 * k + b] = l2[4 * k + b];
			l3[6 * k + 2 + b] = l2[4 * k + 2 + b];
			l3[6 * k + 4 + b] = l2[4 * k + b] ^ l2[4 * k + 2 + b];
		}
#pragma unroll
	for(int k = 0; k < 27; k++) {
		e[3 * k] = l3[2 * k];
		e[3 * k + 1] = l3[2 * k + 1];
		e[3 * k + 2] = l3[2 * k] ^ l3[2 * k + 1];
	}
}

}  // namespace pph
