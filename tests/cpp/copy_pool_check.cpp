// The backend's parallel host copy: every size around the split threshold,
// any number of helpers, many calls in a row; and the checksummed copy of
// getOutput against a scalar lane-by-lane Horner sum (checksum widths 8, 16,
// 32, 64; odd sizes; aligned and unaligned destinations).  Prints "ok" when
// every copy equals its source and every checksum the scalar one.
#include "../../parpar_amd/csrc/pph_copy_pool.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

static uint16_t g_exp[65536], g_log[65536];

static void tables() {
	uint32_t n = 1;
	for(uint32_t e = 0; e < 65535; e++) {
		g_exp[e] = (uint16_t)n;
		g_log[n] = (uint16_t)e;
		n <<= 1;
		if(n > 0xFFFF) n ^= 0x1100B;
	}
}

// H over n bytes: per 16-bit lane, acc = acc * x ^ word over blocks of ck
// bytes, the last block zero-padded
static void scalar_cksum(const uint8_t* src, size_t n, unsigned ck, uint16_t* out) {
	const unsigned lanes = ck / 2;
	for(unsigned l = 0; l < lanes; l++) out[l] = 0;
	for(size_t pos = 0; pos < n; pos += ck) {
		for(unsigned l = 0; l < lanes; l++) {
			uint32_t v = (uint32_t)out[l] << 1;
			if(v & 0x10000) v ^= 0x1100B;
			const size_t b = pos + 2 * l;
			uint16_t w = 0;
			if(b < n) w |= src[b];
			if(b + 1 < n) w |= (uint16_t)src[b + 1] << 8;
			out[l] = (uint16_t)(v ^ w);
		}
	}
}

int main() {
	tables();
	const size_t P = pph::CopyPool::kPart;
	const size_t sizes[] = {0, 1, 63, P - 1, P, 2 * P - 1, 2 * P, 2 * P + 1, 3 * P + 17, 7 * P + 5, 16 * P, 33 * P + 1};
	for(unsigned helpers = 0; helpers < 8; helpers++) {
		pph::CopyPool pool(helpers);
		for(int rep = 0; rep < 3; rep++)
			for(size_t n : sizes) {
				std::vector<uint8_t> src(n + 64), dst(n + 64, 0xEE);
				for(size_t k = 0; k < src.size(); k++) src[k] = (uint8_t)(k * 131 + n + rep);
				pool.copy(dst.data() + 1, src.data() + 3, n);
				for(size_t k = 0; k < n; k++)
					if(dst[1 + k] != src[3 + k]) {
						printf("mismatch helpers %u size %zu at %zu\n", helpers, n, k);
						return 1;
					}
				if(dst[0] != 0xEE || dst[1 + n] != 0xEE) {
					printf("overrun helpers %u size %zu\n", helpers, n);
					return 1;
				}
				for(unsigned ck : {8u, 16u, 32u, 64u})
					for(size_t dst_off : {0, 16, 1}) {
						std::vector<uint8_t> d2(n + 64, 0xEE);
						uint8_t* out = d2.data() + dst_off;
						alignas(16) uint8_t got[64];
						uint16_t want[32];
						pool.copy_cksum(out, src.data() + 3, n, ck, got, g_exp, g_log);
						scalar_cksum(src.data() + 3, n, ck, want);
						for(size_t k = 0; k < n; k++)
							if(out[k] != src[3 + k]) {
								printf("cksum copy mismatch helpers %u size %zu ck %u at %zu\n", helpers, n, ck, k);
								return 1;
							}
						if(out[n] != 0xEE || (dst_off && out[-1] != 0xEE)) {
							printf("cksum copy overrun helpers %u size %zu ck %u\n", helpers, n, ck);
							return 1;
						}
						if(memcmp(got, want, ck) != 0) {
							printf("cksum mismatch helpers %u size %zu ck %u off %zu\n", helpers, n, ck, dst_off);
							return 1;
						}
					}
			}
	}
	printf("ok\n");
	return 0;
}
