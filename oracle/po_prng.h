/*
 * Test infrastructure (oracle side). Counter-based synthetic-input generator.
 *
 * Every synthetic input slice is a pure function of (seed, job, slice): its
 * byte stream is the little-endian concatenation of 64-bit words
 *     word[k] = mix64(key + (k + 1) * GOLDEN),   key = slice_key(seed, job, slice)
 * so any byte range of any slice can be produced independently (the device
 * generator in the HIP library produces the identical stream in parallel).
 * mix64 is the SplitMix64 output finaliser (Steele, Lea, Flood, "Fast
 * splittable pseudorandom number generators", OOPSLA 2014).
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark use
 * this file; the product library carries its own copy of the generator.
 */
#ifndef PO_PRNG_H
#define PO_PRNG_H

#include <stdint.h>
#include <stddef.h>
#include <string.h>

#define PO_GOLDEN 0x9E3779B97F4A7C15ULL

static inline uint64_t po_mix64(uint64_t z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

static inline uint64_t po_slice_key(uint64_t seed, uint32_t job, uint32_t slice) {
	uint64_t k = po_mix64(seed * PO_GOLDEN + 0x632BE59BD9B4E019ULL);
	k = po_mix64(k ^ (((uint64_t)job << 32) | (uint64_t)slice));
	return k;
}

/* Fill dst[0..len) with bytes [offset, offset+len) of the stream of (seed, job, slice). */
static inline void po_prng_fill(uint64_t seed, uint32_t job, uint32_t slice, uint64_t offset, uint8_t* dst, size_t len) {
	uint64_t key = po_slice_key(seed, job, slice);
	size_t done = 0;
	while(done < len) {
		uint64_t pos = offset + done;
		uint64_t k = pos >> 3;
		unsigned b = (unsigned)(pos & 7);
		uint64_t w = po_mix64(key + (k + 1) * PO_GOLDEN);
		uint8_t tmp[8];
		for(int i = 0; i < 8; i++) tmp[i] = (uint8_t)(w >> (8 * i));
		size_t take = 8 - b;
		if(take > len - done) take = len - done;
		memcpy(dst + done, tmp + b, take);
		done += take;
	}
}

#endif
