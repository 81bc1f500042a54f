// gf16_cksum on the host, computed while recovery bytes are copied out to the
// caller: the check of copy_cksum_check (gf16/gf16_cksum_base.h:48-59) over
// the bytes the caller receives, not over the device copy they came from.
//
// The checksum of a slice of n bytes is a vector of ck/2 little-endian 16-bit
// lanes, H = sum_k B_k * 2^(m-1-k) over its m = ceil(n/ck) blocks of ck bytes
// (the last one zero-padded), lane-wise in GF(2^16) with polynomial 0x1100B,
// evaluated by Horner: acc = acc * 2 ^ block.  copy_cksum_check rewinds the
// stored block by 2^-m and runs the same Horner steps from it, so the slice is
// valid exactly when H equals the stored block.
#pragma once
#include <emmintrin.h>

#include <cstddef>
#include <cstdint>
#include <cstring>

namespace pph {

// Lane-wise accumulator for ck = 8, 16, 32 or 64 bytes (the low ck bytes of v).
struct CkAcc {
	__m128i v[4];
};

inline void ck_zero(CkAcc& a) {
	for(auto& x : a.v) x = _mm_setzero_si128();
}

// acc = acc * 2 in every 16-bit lane
static inline __m128i ck_mul2(__m128i x) {
	const __m128i top = _mm_srai_epi16(x, 15);
	return _mm_xor_si128(_mm_slli_epi16(x, 1), _mm_and_si128(top, _mm_set1_epi16(0x100B)));
}

// Horner over the whole blocks of [src, src + blocks * ck), copying them to
// dst when dst is not null: non-temporal stores (16-byte aligned dst) when
// NT, ordinary unaligned stores otherwise.
template<unsigned CK, bool NT>
static inline void ck_blocks(CkAcc& a, uint8_t* dst, const uint8_t* src, size_t blocks) {
	constexpr unsigned V = CK >= 16 ? CK / 16 : 1;
	__m128i acc[V];
	for(unsigned k = 0; k < V; k++) acc[k] = a.v[k];
	for(size_t b = 0; b < blocks; b++, src += CK) {
		if(CK >= 16) {
			for(unsigned k = 0; k < V; k++) {
				const __m128i d = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + 16 * k));
				acc[k] = _mm_xor_si128(ck_mul2(acc[k]), d);
				if(dst) {
					if(NT)
						_mm_stream_si128(reinterpret_cast<__m128i*>(dst + 16 * k), d);
					else
						_mm_storeu_si128(reinterpret_cast<__m128i*>(dst + 16 * k), d);
				}
			}
		} else {
			const __m128i d = _mm_loadl_epi64(reinterpret_cast<const __m128i*>(src));
			acc[0] = _mm_xor_si128(ck_mul2(acc[0]), d);
			if(dst) _mm_storel_epi64(reinterpret_cast<__m128i*>(dst), d);
		}
		if(dst) dst += CK;
	}
	for(unsigned k = 0; k < V; k++) a.v[k] = acc[k];
}

template<bool NT>
static inline void ck_blocks_any(CkAcc& a, uint8_t* dst, const uint8_t* src, size_t blocks, unsigned ck) {
	switch(ck) {
		case 8: ck_blocks<8, false>(a, dst, src, blocks); break;
		case 16: ck_blocks<16, NT>(a, dst, src, blocks); break;
		case 32: ck_blocks<32, NT>(a, dst, src, blocks); break;
		default: ck_blocks<64, NT>(a, dst, src, blocks); break;
	}
}

// Horner over n bytes (the last block zero-padded) while copying them to dst
// (null: no copy).  A 16-byte aligned dst gets non-temporal stores, fenced
// before return.
inline void ck_copy_accumulate(CkAcc& a, void* dst_, const void* src_, size_t n, unsigned ck) {
	uint8_t* dst = static_cast<uint8_t*>(dst_);
	const uint8_t* src = static_cast<const uint8_t*>(src_);
	const size_t blocks = n / ck;
	const bool nt = dst && ((uintptr_t)dst & 15) == 0;
	if(nt)
		ck_blocks_any<true>(a, dst, src, blocks, ck);
	else
		ck_blocks_any<false>(a, dst, src, blocks, ck);
	const size_t rem = n - blocks * ck;
	if(rem) {
		alignas(16) uint8_t last[64] = {0};
		memcpy(last, src + blocks * ck, rem);
		ck_blocks_any<false>(a, nullptr, last, 1, ck);
		if(dst) memcpy(dst + blocks * ck, src + blocks * ck, rem);
	}
	if(nt) _mm_sfence();
}

inline void ck_store(const CkAcc& a, uint8_t* out, unsigned ck) {
	alignas(16) uint8_t buf[64];
	for(unsigned k = 0; k < 4; k++) _mm_store_si128(reinterpret_cast<__m128i*>(buf + 16 * k), a.v[k]);
	memcpy(out, buf, ck);
}

// out ^= in * 2^e in every lane (exp/log: the field's tables, log[0] unused,
// exp over 0..65534).
inline void ck_mul_exp_xor(uint16_t* out, const uint16_t* in, unsigned lanes, uint32_t e, const uint16_t* exp_tab,
                           const uint16_t* log_tab) {
	e %= 65535;
	for(unsigned l = 0; l < lanes; l++) {
		const uint16_t v = in[l];
		if(!v) continue;
		uint32_t s = (uint32_t)log_tab[v] + e;
		if(s >= 65535) s -= 65535;
		out[l] ^= exp_tab[s];
	}
}

}  // namespace pph
