/*
 * ORACLE — test infrastructure, not product code.  Only tests/, the smoke
 * check and the benchmark's CPU-baseline leg may load this library, and only as
 * the checker; the HIP product path never calls it.
 *
 * Restates ParPar's CPU recovery path:
 *   po_mul / po_exp / po_log    gf16/gf16_global.h:9-10 (polynomial 0x1100B, generator 2)
 *   po_input_log                gf16/gfmat_coeff.c:12-22,56-58 (PAR2 sieve)
 *   po_coeff_log / po_coeff     gf16/gfmat_coeff.c:60-77
 *   po_cksum / po_cksum_verify  gf16/gf16_cksum_base.h:13-59 with the lane-wise
 *                               Horner step of gf16/gf16_checksum_generic.h:25-124
 *                               and gf16/gf16_checksum_x86.h:6-150
 *   po_muladd_region            gf16/gf16_lookup.c:124-249 (split 8-bit tables)
 *   po_recover                  gf16/controller_cpu.cpp:463-537 (+ zero-pad of short
 *                               inputs, gf16/gf16_cksum_base.h:38-44)
 *   po_md5 / po_crc32           hasher/hasher_base.h:83-198 semantics, RFC 1321 and
 *                               IEEE CRC-32 (reflected 0xEDB88320)
 *
 * Pinning (tests/test_oracle.py):
 *   - po_mul against po_mul_slow (shift-and-add) on sampled and exhaustive-row pairs,
 *     and field axioms (a*inv(a) = 1 via log/exp);
 *   - po_input_log / po_coeff against the reference's gfmat_* (oracle/_ref, every
 *     input 0..32767 x sampled exponents) and the committed spot table;
 *   - po_cksum / po_cksum_verify against the reference's copy_cksum at widths
 *     8/16/32/64 (fixtures) and checksum linearity;
 *   - po_recover against full PAR2ProcCPU outputs (fixtures, oracle/_ref) and
 *     invariants (exp 0 = XOR, single input = c*in);
 *   - po_md5 / po_crc32 against Python hashlib / zlib.
 */
#include "gf16_oracle.h"
#include "po_prng.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PO_POLY 0x1100B

static uint16_t po_exp_tab[65536 * 2];
static uint16_t po_log_tab[65536];
static int po_ready = 0;
static uint16_t po_in_log[32768];

void po_init(void) {
	if(po_ready) return;
	uint32_t n = 1;
	for(uint32_t e = 0; e < 65535; e++) {
		po_exp_tab[e] = (uint16_t)n;
		po_exp_tab[e + 65535] = (uint16_t)n;
		po_log_tab[n] = (uint16_t)e;
		n <<= 1;
		if(n > 65535) n ^= PO_POLY;
	}
	po_exp_tab[131070] = po_exp_tab[0];
	po_log_tab[0] = 0; /* unused: zero has no logarithm */
	/* PAR2 input constants: the i-th exponent n >= 1 with n coprime to 65535,
	 * i.e. n not divisible by 3, 5, 17 or 257 (gfmat_coeff.c:12-22). */
	uint32_t e = 0;
	for(int i = 0; i < 32768; i++) {
		do { e++; } while(!(e % 3) || !(e % 5) || !(e % 17) || !(e % 257));
		po_in_log[i] = (uint16_t)e;
	}
	po_ready = 1;
}

uint16_t po_mul_slow(uint16_t a, uint16_t b) {
	uint32_t r = 0, x = a;
	for(int bit = 0; bit < 16; bit++) {
		if(b & (1u << bit)) r ^= x;
		x <<= 1;
		if(x & 0x10000) x ^= PO_POLY;
	}
	return (uint16_t)r;
}

uint16_t po_exp(uint32_t v) {
	return po_exp_tab[v % 65535];
}
uint16_t po_log(uint16_t x) {
	return po_log_tab[x];
}
uint16_t po_mul(uint16_t a, uint16_t b) {
	if(!a || !b) return 0;
	return po_exp_tab[(uint32_t)po_log_tab[a] + po_log_tab[b]];
}

uint16_t po_input_log(uint16_t input) {
	po_init();
	return po_in_log[input & 32767];
}
uint16_t po_coeff_log(uint16_t inputLog, uint16_t exp) {
	return (uint16_t)(((uint32_t)inputLog * exp) % 65535);
}
uint16_t po_coeff(uint16_t input, uint16_t exp) {
	po_init();
	return po_exp(po_coeff_log(po_input_log(input), exp));
}

/* ---- gf16_cksum ---- */
static inline uint16_t po_mul2(uint16_t v) {
	return (uint16_t)((v << 1) ^ ((v & 0x8000) ? (PO_POLY & 0xffff) : 0));
}
static void po_ck_block(uint16_t* ck, const uint8_t* src, size_t amount, unsigned lanes) {
	for(unsigned l = 0; l < lanes; l++) {
		uint16_t d = 0;
		size_t p = (size_t)l * 2;
		if(p < amount) d |= src[p];
		if(p + 1 < amount) d |= (uint16_t)src[p + 1] << 8;
		ck[l] = po_mul2(ck[l]) ^ d;
	}
}
static void po_ck_mul(uint16_t* ck, uint16_t c, unsigned lanes) {
	for(unsigned l = 0; l < lanes; l++) ck[l] = po_mul(ck[l], c);
}
static void po_ck_store(uint8_t* out, const uint16_t* ck, unsigned lanes) {
	for(unsigned l = 0; l < lanes; l++) {
		out[2 * l] = (uint8_t)ck[l];
		out[2 * l + 1] = (uint8_t)(ck[l] >> 8);
	}
}

void po_cksum(const uint8_t* src, size_t srcLen, size_t sliceLen, unsigned width, uint8_t* out) {
	po_init();
	unsigned lanes = width / 2;
	uint16_t ck[512];
	memset(ck, 0, sizeof(ck));
	for(size_t pos = 0; pos < srcLen; pos += width) {
		size_t amount = srcLen - pos < width ? srcLen - pos : width;
		po_ck_block(ck, src + pos, amount, lanes);
	}
	if(srcLen < sliceLen) {
		size_t done = (srcLen + width - 1) / width * width;
		if(done < sliceLen)
			po_ck_mul(ck, po_exp((uint32_t)(((sliceLen - done + width - 1) / width) % 65535)), lanes);
	}
	po_ck_store(out, ck, lanes);
}

int po_cksum_verify(const uint8_t* data, size_t len, const uint8_t* stored, unsigned width) {
	po_init();
	unsigned lanes = width / 2;
	uint16_t ck[512];
	for(unsigned l = 0; l < lanes; l++) ck[l] = (uint16_t)(stored[2 * l] | (stored[2 * l + 1] << 8));
	po_ck_mul(ck, po_exp((uint32_t)(65535 - (((len + width - 1) / width) % 65535))), lanes);
	for(size_t pos = 0; pos < len; pos += width) {
		size_t amount = len - pos < width ? len - pos : width;
		po_ck_block(ck, data + pos, amount, lanes);
	}
	for(unsigned l = 0; l < lanes; l++)
		if(ck[l]) return 0;
	return 1;
}

/* ---- mul-add ---- */
static void po_build_tables(uint16_t coeff, uint16_t* lo, uint16_t* hi) {
	lo[0] = hi[0] = 0;
	for(int b = 0; b < 8; b++) {
		uint16_t vl = po_mul(coeff, (uint16_t)(1u << b));
		uint16_t vh = po_mul(coeff, (uint16_t)(1u << (b + 8)));
		int step = 1 << b;
		for(int j = 0; j < step; j++) {
			lo[step + j] = lo[j] ^ vl;
			hi[step + j] = hi[j] ^ vh;
		}
	}
}

/* dst[0..dstLen) ^= coeff * src over ceil(srcLen/2) words; source bytes at or
 * beyond srcLen read as zero (the zero padding of a short input slice). */
static void po_muladd_tab(uint8_t* dst, size_t dstLen, const uint8_t* src, size_t srcLen, const uint16_t* lo, const uint16_t* hi) {
	size_t words = srcLen / 2;
	for(size_t w = 0; w < words; w++) {
		uint16_t x = (uint16_t)(src[2 * w] | (src[2 * w + 1] << 8));
		uint16_t y = lo[x & 255] ^ hi[x >> 8];
		dst[2 * w] ^= (uint8_t)y;
		dst[2 * w + 1] ^= (uint8_t)(y >> 8);
	}
	if(srcLen & 1) {
		uint16_t y = lo[src[srcLen - 1]];
		dst[srcLen - 1] ^= (uint8_t)y;
		if(srcLen < dstLen) dst[srcLen] ^= (uint8_t)(y >> 8);
	}
}

void po_muladd_region(uint8_t* dst, const uint8_t* src, size_t len, uint16_t coeff) {
	uint16_t lo[256], hi[256];
	po_init();
	po_build_tables(coeff, lo, hi);
	po_muladd_tab(dst, len, src, len, lo, hi);
}

int po_recover(unsigned nIn, const uint8_t* inputs, size_t inStride, const size_t* inLens,
               const uint16_t* inputNums, const uint16_t* coeffs, unsigned nOut,
               const uint16_t* exps, size_t sliceSize, uint8_t* outputs, int threads) {
	po_init();
	memset(outputs, 0, (size_t)nOut * sliceSize);
	if(!nIn || !nOut || !sliceSize) return 1;
	const size_t tile = 32768; /* bytes of every slice per work item */
	const unsigned outBlock = 8; /* recovery slices per work item */
	size_t nTiles = (sliceSize + tile - 1) / tile;
	size_t nBlocks = (nOut + outBlock - 1) / outBlock;
	long nItems = (long)(nTiles * nBlocks);
	int used = 1;
#ifdef _OPENMP
	if(threads <= 0) threads = omp_get_max_threads();
	#pragma omp parallel num_threads(threads)
	{
		#pragma omp single
		used = omp_get_num_threads();
	}
	#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#else
	(void)threads;
#endif
	for(long item = 0; item < nItems; item++) {
		uint16_t lo[256], hi[256];
		size_t t = (size_t)item / nBlocks;
		unsigned r0 = (unsigned)((size_t)item % nBlocks) * outBlock;
		unsigned r1 = r0 + outBlock < nOut ? r0 + outBlock : nOut;
		size_t off = t * tile;
		size_t len = sliceSize - off < tile ? sliceSize - off : tile;
		for(unsigned i = 0; i < nIn; i++) {
			size_t inLen = inLens ? inLens[i] : sliceSize;
			if(inLen <= off) continue; /* zero padding contributes nothing */
			size_t avail = inLen - off < len ? inLen - off : len;
			const uint8_t* src = inputs + (size_t)i * inStride + off;
			uint16_t logI = coeffs ? 0 : po_input_log(inputNums[i]);
			for(unsigned r = r0; r < r1; r++) {
				uint16_t c = coeffs ? coeffs[(size_t)i * nOut + r] : po_exp(po_coeff_log(logI, exps[r]));
				if(!c) continue;
				po_build_tables(c, lo, hi);
				po_muladd_tab(outputs + (size_t)r * sliceSize + off, len, src, avail, lo, hi);
			}
		}
	}
	return used;
}

void po_prng_fill_c(uint64_t seed, uint32_t job, uint32_t slice, uint64_t offset, uint8_t* dst, size_t len) {
	po_prng_fill(seed, job, slice, offset, dst, len);
}

/* gf16_cksum of synthetic slices without materialising them: slice k (of
 * `count`, numbered first + k) holds lens[k] bytes of the generator's stream
 * (seed, job, first + k), zero-padded to sliceLen (lens NULL: sliceLen).
 * The composition of po_prng_fill and po_cksum, computed 64-byte block by
 * block, slices in parallel; out[k * width .. ) receives slice k's checksum. */
void po_prng_cksums(uint64_t seed, uint32_t job, uint32_t first, unsigned count, const uint32_t* lens,
                    size_t sliceLen, unsigned width, uint8_t* out, int threads) {
	po_init();
	if(threads <= 0) threads = omp_get_max_threads();
	#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
	for(long k = 0; k < (long)count; k++) {
		const size_t len = lens ? lens[k] : sliceLen;
		const unsigned lanes = width / 2;
		uint16_t ck[32];
		uint8_t blk[4096];
		memset(ck, 0, sizeof(ck));
		for(size_t pos = 0; pos < len; pos += sizeof(blk)) {
			const size_t n = len - pos < sizeof(blk) ? len - pos : sizeof(blk);
			po_prng_fill(seed, job, first + (uint32_t)k, pos, blk, n);
			for(size_t b = 0; b < n; b += width) po_ck_block(ck, blk + b, n - b < width ? n - b : width, lanes);
		}
		if(len < sliceLen) {
			const size_t done = (len + width - 1) / width * width;
			if(done < sliceLen)
				po_ck_mul(ck, po_exp((uint32_t)(((sliceLen - done + width - 1) / width) % 65535)), lanes);
		}
		po_ck_store(out + (size_t)k * width, ck, lanes);
	}
}

/* ---- MD5 (RFC 1321) ---- */
static const uint32_t po_md5_k[64] = {
	0xd76aa478, 0xe8c7b756, 0x242070db, 0xc1bdceee, 0xf57c0faf, 0x4787c62a, 0xa8304613, 0xfd469501,
	0x698098d8, 0x8b44f7af, 0xffff5bb1, 0x895cd7be, 0x6b901122, 0xfd987193, 0xa679438e, 0x49b40821,
	0xf61e2562, 0xc040b340, 0x265e5a51, 0xe9b6c7aa, 0xd62f105d, 0x02441453, 0xd8a1e681, 0xe7d3fbc8,
	0x21e1cde6, 0xc33707d6, 0xf4d50d87, 0x455a14ed, 0xa9e3e905, 0xfcefa3f8, 0x676f02d9, 0x8d2a4c8a,
	0xfffa3942, 0x8771f681, 0x6d9d6122, 0xfde5380c, 0xa4beea44, 0x4bdecfa9, 0xf6bb4b60, 0xbebfbc70,
	0x289b7ec6, 0xeaa127fa, 0xd4ef3085, 0x04881d05, 0xd9d4d039, 0xe6db99e5, 0x1fa27cf8, 0xc4ac5665,
	0xf4292244, 0x432aff97, 0xab9423a7, 0xfc93a039, 0x655b59c3, 0x8f0ccc92, 0xffeff47d, 0x85845dd1,
	0x6fa87e4f, 0xfe2ce6e0, 0xa3014314, 0x4e0811a1, 0xf7537e82, 0xbd3af235, 0x2ad7d2bb, 0xeb86d391};
static const uint8_t po_md5_r[64] = {
	7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
	5, 9, 14, 20, 5, 9, 14, 20, 5, 9, 14, 20, 5, 9, 14, 20,
	4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
	6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};

static void po_md5_block(uint32_t st[4], const uint8_t* b) {
	uint32_t m[16];
	for(int i = 0; i < 16; i++)
		m[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) | ((uint32_t)b[4 * i + 3] << 24);
	uint32_t a = st[0], bb = st[1], c = st[2], d = st[3];
	for(int i = 0; i < 64; i++) {
		uint32_t f;
		int g;
		if(i < 16) { f = (bb & c) | (~bb & d); g = i; }
		else if(i < 32) { f = (d & bb) | (~d & c); g = (5 * i + 1) & 15; }
		else if(i < 48) { f = bb ^ c ^ d; g = (3 * i + 5) & 15; }
		else { f = c ^ (bb | ~d); g = (7 * i) & 15; }
		uint32_t tmp = d;
		d = c;
		c = bb;
		uint32_t x = a + f + po_md5_k[i] + m[g];
		bb = bb + ((x << po_md5_r[i]) | (x >> (32 - po_md5_r[i])));
		a = tmp;
	}
	st[0] += a; st[1] += bb; st[2] += c; st[3] += d;
}

void po_md5(const uint8_t* data, size_t len, size_t padLen, uint8_t out[16]) {
	uint32_t st[4] = {0x67452301, 0xefcdab89, 0x98badcfe, 0x10325476};
	if(padLen < len) padLen = len;
	uint8_t blk[64];
	size_t pos = 0;
	while(pos + 64 <= padLen) {
		if(pos + 64 <= len) po_md5_block(st, data + pos);
		else {
			memset(blk, 0, 64);
			if(pos < len) memcpy(blk, data + pos, len - pos);
			po_md5_block(st, blk);
		}
		pos += 64;
	}
	size_t rem = padLen - pos;
	memset(blk, 0, 64);
	if(pos < len) memcpy(blk, data + pos, len - pos);
	blk[rem] = 0x80;
	if(rem >= 56) {
		po_md5_block(st, blk);
		memset(blk, 0, 64);
	}
	uint64_t bits = (uint64_t)padLen * 8;
	for(int i = 0; i < 8; i++) blk[56 + i] = (uint8_t)(bits >> (8 * i));
	po_md5_block(st, blk);
	for(int i = 0; i < 4; i++)
		for(int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(st[i] >> (8 * j));
}

/* ---- CRC-32 ---- */
static uint32_t po_crc_tab[256];
static int po_crc_ready = 0;
uint32_t po_crc32(const uint8_t* data, size_t len, size_t padLen) {
	if(!po_crc_ready) {
		for(uint32_t i = 0; i < 256; i++) {
			uint32_t c = i;
			for(int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : (c >> 1);
			po_crc_tab[i] = c;
		}
		po_crc_ready = 1;
	}
	if(padLen < len) padLen = len;
	uint32_t c = 0xFFFFFFFFu;
	for(size_t i = 0; i < padLen; i++) {
		uint8_t b = i < len ? data[i] : 0;
		c = po_crc_tab[(c ^ b) & 255] ^ (c >> 8);
	}
	return c ^ 0xFFFFFFFFu;
}
