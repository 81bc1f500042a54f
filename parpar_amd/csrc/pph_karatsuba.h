// Karatsuba expansion of GF(2^16) elements for the bilinear recovery form,
// computed bit-sliced on the device (shared by pph_bilinear.hip and
// pph_bl_mfma.hip).
#pragma once
#include <stdint.h>

namespace pph {

constexpr unsigned kBl = 81;  // rank of the four-level Karatsuba formula

// Recombination of component t: the product polynomial of the unit vector e_t
// (Karatsuba recombination L + (D + L + H) x^h + H x^n per level, components
// in the order of bl_expand_slices), reduced modulo the PAR2 polynomial
// 0x1100B.  W_t is a 16-bit mask: bit j set where component t's parity enters
// bit j of the product.  Evaluated at compile time.
constexpr uint64_t bl_combine_unit(unsigned t, unsigned count, unsigned n) {
	if(n == 1) return 1;
	const unsigned h = n / 2, k = count / 3;
	const uint64_t L = t < k ? bl_combine_unit(t, k, h) : 0;
	const uint64_t H = t >= k && t < 2 * k ? bl_combine_unit(t - k, k, h) : 0;
	const uint64_t D = t >= 2 * k ? bl_combine_unit(t - 2 * k, k, h) : 0;
	return L ^ ((D ^ L ^ H) << h) ^ (H << n);
}
constexpr uint16_t bl_reduce(uint64_t v) {
	for(int b = 30; b >= 16; b--)
		if((v >> b) & 1) v ^= (uint64_t)0x1100B << (b - 16);
	return (uint16_t)v;
}
struct BlTable {
	uint16_t w[kBl];
};
constexpr BlTable bl_table() {
	BlTable r{};
	for(unsigned t = 0; t < kBl; t++) r.w[t] = bl_reduce(bl_combine_unit(t, kBl, 16));
	return r;
}

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

// Karatsuba recombination, bit-sliced: products p of the components of two
// N-term polynomials (the order of bl_expand_slices) -> the 2N - 1 coefficients
// of their product, L + (D + L + H) x^(N/2) + H x^N per level.
template <unsigned N, unsigned C>
__device__ __forceinline__ void bl_recombine_n(const uint32_t* p, uint32_t* r) {
	if constexpr(N == 1) {
		r[0] = p[0];
	} else {
		constexpr unsigned H = N / 2, K = C / 3;
		uint32_t l[2 * H - 1], h[2 * H - 1], d[2 * H - 1];
		bl_recombine_n<H, K>(p, l);
		bl_recombine_n<H, K>(p + K, h);
		bl_recombine_n<H, K>(p + 2 * K, d);
#pragma unroll
		for(unsigned i = 0; i < 2 * N - 1; i++) r[i] = 0;
#pragma unroll
		for(unsigned i = 0; i < 2 * H - 1; i++) {
			r[i] ^= l[i];
			r[i + H] ^= d[i] ^ l[i] ^ h[i];
			r[i + N] ^= h[i];
		}
	}
}
template <unsigned N>
__device__ __forceinline__ void bl_recombine(const uint32_t* p, uint32_t* r) {
	bl_recombine_n<N, kBl>(p, r);
}
// coefficients 16 .. 30 folded down modulo x^16 + x^12 + x^3 + x + 1
__device__ __forceinline__ void bl_reduce_slices(uint32_t* r) {
#pragma unroll
	for(unsigned b = 30; b >= 16; b--) {
		r[b - 4] ^= r[b];
		r[b - 13] ^= r[b];
		r[b - 15] ^= r[b];
		r[b - 16] ^= r[b];
	}
}

}  // namespace pph
