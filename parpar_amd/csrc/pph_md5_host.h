// MD5 (RFC 1321) on the host, for files whose bytes are in host memory: one
// serial chain runs several times faster on a CPU core than on one wavefront.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

namespace pph {

class HostMd5 {
	uint32_t st_[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
	uint8_t buf_[64];
	size_t fill_ = 0;
	uint64_t total_ = 0;

	template <unsigned C>
	static uint32_t rotl(uint32_t x) {
		return (x << C) | (x >> (32 - C));
	}
	// One step: a = b + rotl(a + f(b, c, d) + K[i] + m[g], s), with the four
	// round functions of RFC 1321 section 3.4.  Only b is produced by the
	// previous step, so everything that does not read b is summed into t
	// first: a + K + m, and for G = (b & d) | (c & ~d) the disjoint half
	// c & ~d (the halves share no bits, so | is +).  The empty asm pins that
	// sum; without it the compiler reassociates K after f and puts an extra
	// add on the dependent chain.  The chain per step is then the ops of f
	// that read b (2, 1, 1, 2 for F, G, H, I), one add, the rotate and the
	// final add.
	template <unsigned R, unsigned C>
	static void step(uint32_t& a, uint32_t b, uint32_t c, uint32_t d, uint32_t km) {
		uint32_t t = a + km;
		if(R == 1) t += c & ~d;
		asm("" : "+r"(t));
		if(R == 0) t += d ^ (b & (c ^ d));
		else if(R == 1) t += b & d;
		else if(R == 2) t += b ^ (c ^ d);
		else t += c ^ (b | ~d);
		a = b + rotl<C>(t);
	}

	void block(const uint8_t* p) {
		uint32_t m[16];
		memcpy(m, p, 64);  // little-endian hosts
		uint32_t a = st_[0], b = st_[1], c = st_[2], d = st_[3];
#define PPH_MD5_Q(R, g0, g1, g2, g3, s0, s1, s2, s3, k0, k1, k2, k3) \
	step<R, s0>(a, b, c, d, k0 + m[g0]);                                             \
	step<R, s1>(d, a, b, c, k1 + m[g1]);                                             \
	step<R, s2>(c, d, a, b, k2 + m[g2]);                                             \
	step<R, s3>(b, c, d, a, k3 + m[g3]);
		PPH_MD5_Q(0, 0, 1, 2, 3, 7, 12, 17, 22, 0xd76aa478, 0xe8c7b756, 0x242070db, 0xc1bdceee)
		PPH_MD5_Q(0, 4, 5, 6, 7, 7, 12, 17, 22, 0xf57c0faf, 0x4787c62a, 0xa8304613, 0xfd469501)
		PPH_MD5_Q(0, 8, 9, 10, 11, 7, 12, 17, 22, 0x698098d8, 0x8b44f7af, 0xffff5bb1, 0x895cd7be)
		PPH_MD5_Q(0, 12, 13, 14, 15, 7, 12, 17, 22, 0x6b901122, 0xfd987193, 0xa679438e, 0x49b40821)
		PPH_MD5_Q(1, 1, 6, 11, 0, 5, 9, 14, 20, 0xf61e2562, 0xc040b340, 0x265e5a51, 0xe9b6c7aa)
		PPH_MD5_Q(1, 5, 10, 15, 4, 5, 9, 14, 20, 0xd62f105d, 0x02441453, 0xd8a1e681, 0xe7d3fbc8)
		PPH_MD5_Q(1, 9, 14, 3, 8, 5, 9, 14, 20, 0x21e1cde6, 0xc33707d6, 0xf4d50d87, 0x455a14ed)
		PPH_MD5_Q(1, 13, 2, 7, 12, 5, 9, 14, 20, 0xa9e3e905, 0xfcefa3f8, 0x676f02d9, 0x8d2a4c8a)
		PPH_MD5_Q(2, 5, 8, 11, 14, 4, 11, 16, 23, 0xfffa3942, 0x8771f681, 0x6d9d6122, 0xfde5380c)
		PPH_MD5_Q(2, 1, 4, 7, 10, 4, 11, 16, 23, 0xa4beea44, 0x4bdecfa9, 0xf6bb4b60, 0xbebfbc70)
		PPH_MD5_Q(2, 13, 0, 3, 6, 4, 11, 16, 23, 0x289b7ec6, 0xeaa127fa, 0xd4ef3085, 0x04881d05)
		PPH_MD5_Q(2, 9, 12, 15, 2, 4, 11, 16, 23, 0xd9d4d039, 0xe6db99e5, 0x1fa27cf8, 0xc4ac5665)
		PPH_MD5_Q(3, 0, 7, 14, 5, 6, 10, 15, 21, 0xf4292244, 0x432aff97, 0xab9423a7, 0xfc93a039)
		PPH_MD5_Q(3, 12, 3, 10, 1, 6, 10, 15, 21, 0x655b59c3, 0x8f0ccc92, 0xffeff47d, 0x85845dd1)
		PPH_MD5_Q(3, 8, 15, 6, 13, 6, 10, 15, 21, 0x6fa87e4f, 0xfe2ce6e0, 0xa3014314, 0x4e0811a1)
		PPH_MD5_Q(3, 4, 11, 2, 9, 6, 10, 15, 21, 0xf7537e82, 0xbd3af235, 0x2ad7d2bb, 0xeb86d391)
#undef PPH_MD5_Q
		st_[0] += a;
		st_[1] += b;
		st_[2] += c;
		st_[3] += d;
	}

public:
	void update(const uint8_t* p, size_t n) {
		total_ += n;
		if(fill_) {
			const size_t k = n < 64 - fill_ ? n : 64 - fill_;
			memcpy(buf_ + fill_, p, k);
			fill_ += k;
			p += k;
			n -= k;
			if(fill_ < 64) return;
			block(buf_);
			fill_ = 0;
		}
		for(; n >= 64; p += 64, n -= 64) block(p);
		memcpy(buf_, p, n);
		fill_ = n;
	}
	void final(uint8_t out[16]) {
		const uint64_t bits = total_ * 8;
		const uint8_t pad = 0x80;
		update(&pad, 1);
		const uint8_t zero[64] = {};
		update(zero, fill_ <= 56 ? 56 - fill_ : 120 - fill_);
		uint8_t len[8];
		for(int i = 0; i < 8; i++) len[i] = (uint8_t)(bits >> (8 * i));
		update(len, 8);
		memcpy(out, st_, 16);
	}
};

}  // namespace pph
