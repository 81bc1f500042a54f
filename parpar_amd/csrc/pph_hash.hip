// Hashes of the PAR2 input path on the device (the hasher/ path ParPar runs on
// CPU threads, hasher/hasher_base.h:83-198):
//   - per slice: MD5 and CRC-32 of the slice zero-padded to the slice size
//     (IFSC entries; getBlock(..., sliceSize - pos), hasher_base.h:154-176);
//   - per file: MD5 of the file bytes (hasher_base.h:178-186).
// MD5 follows RFC 1321; CRC-32 is the IEEE 802.3 reflected CRC (0xEDB88320).
// MD5 is one serial chain per message: slices get one lane each; a file gets
// one wavefront whose lanes all run the same chain on wave-uniform values, so
// the compiler keeps it on the scalar unit.
#include "pph_internal.h"

namespace pph {

// RFC 1321 sine-derived constants; indexed with compile-time constants only, so
// they become instruction literals.
constexpr uint32_t kMd5K[64] = {
    0xd76aa478, 0xe8c7b756, 0x242070db, 0xc1bdceee, 0xf57c0faf, 0x4787c62a, 0xa8304613, 0xfd469501,
    0x698098d8, 0x8b44f7af, 0xffff5bb1, 0x895cd7be, 0x6b901122, 0xfd987193, 0xa679438e, 0x49b40821,
    0xf61e2562, 0xc040b340, 0x265e5a51, 0xe9b6c7aa, 0xd62f105d, 0x02441453, 0xd8a1e681, 0xe7d3fbc8,
    0x21e1cde6, 0xc33707d6, 0xf4d50d87, 0x455a14ed, 0xa9e3e905, 0xfcefa3f8, 0x676f02d9, 0x8d2a4c8a,
    0xfffa3942, 0x8771f681, 0x6d9d6122, 0xfde5380c, 0xa4beea44, 0x4bdecfa9, 0xf6bb4b60, 0xbebfbc70,
    0x289b7ec6, 0xeaa127fa, 0xd4ef3085, 0x04881d05, 0xd9d4d039, 0xe6db99e5, 0x1fa27cf8, 0xc4ac5665,
    0xf4292244, 0x432aff97, 0xab9423a7, 0xfc93a039, 0x655b59c3, 0x8f0ccc92, 0xffeff47d, 0x85845dd1,
    0x6fa87e4f, 0xfe2ce6e0, 0xa3014314, 0x4e0811a1, 0xf7537e82, 0xbd3af235, 0x2ad7d2bb, 0xeb86d391};

// rotate left by c (1..31) as one v_alignbit_b32
__device__ __forceinline__ uint32_t rotl(uint32_t x, unsigned c) { return __builtin_amdgcn_alignbit(x, x, 32 - c); }

// One 64-byte block (16 little-endian words) into the state.
__device__ __forceinline__ void md5_block(uint32_t st[4], const uint32_t m[16]) {
	uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
#define PPH_MD5_STEP(f, g, r, i)                                         \
	{                                                                    \
		const uint32_t x = a + (f) + kMd5K[i] + m[g];                    \
		a = d;                                                           \
		d = c;                                                           \
		c = b;                                                           \
		b = b + rotl(x, r);                                              \
	}
#pragma unroll
	for(int i = 0; i < 16; i++) {
		constexpr unsigned R0[4] = {7, 12, 17, 22};
		PPH_MD5_STEP((b & c) | (~b & d), i, R0[i & 3], i);
	}
#pragma unroll
	for(int i = 16; i < 32; i++) {
		constexpr unsigned R1[4] = {5, 9, 14, 20};
		PPH_MD5_STEP((d & b) | (~d & c), (5 * i + 1) & 15, R1[i & 3], i);
	}
#pragma unroll
	for(int i = 32; i < 48; i++) {
		constexpr unsigned R2[4] = {4, 11, 16, 23};
		PPH_MD5_STEP(b ^ c ^ d, (3 * i + 5) & 15, R2[i & 3], i);
	}
#pragma unroll
	for(int i = 48; i < 64; i++) {
		constexpr unsigned R3[4] = {6, 10, 15, 21};
		PPH_MD5_STEP(c ^ (b | ~d), (7 * i) & 15, R3[i & 3], i);
	}
#undef PPH_MD5_STEP
	st[0] += a;
	st[1] += b;
	st[2] += c;
	st[3] += d;
}

// One block with the per-step sums km[i] = K[i] + m[g(i)] precomputed: each
// step is then bit-select, three-way add, rotate, add.
__device__ __forceinline__ void md5_block_km(uint32_t st[4], const uint32_t km[64]) {
	uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
#define PPH_MD5_STEP_KM(f, r, i)                                         \
	{                                                                    \
		const uint32_t x = a + (f) + km[i];                              \
		a = d;                                                           \
		d = c;                                                           \
		c = b;                                                           \
		b = b + rotl(x, r);                                              \
	}
#pragma unroll
	for(int i = 0; i < 16; i++) {
		constexpr unsigned R0[4] = {7, 12, 17, 22};
		PPH_MD5_STEP_KM((b & c) | (~b & d), R0[i & 3], i);
	}
#pragma unroll
	for(int i = 16; i < 32; i++) {
		constexpr unsigned R1[4] = {5, 9, 14, 20};
		PPH_MD5_STEP_KM((d & b) | (~d & c), R1[i & 3], i);
	}
#pragma unroll
	for(int i = 32; i < 48; i++) {
		constexpr unsigned R2[4] = {4, 11, 16, 23};
		PPH_MD5_STEP_KM(b ^ c ^ d, R2[i & 3], i);
	}
#pragma unroll
	for(int i = 48; i < 64; i++) {
		constexpr unsigned R3[4] = {6, 10, 15, 21};
		PPH_MD5_STEP_KM(c ^ (b | ~d), R3[i & 3], i);
	}
#undef PPH_MD5_STEP_KM
	st[0] += a;
	st[1] += b;
	st[2] += c;
	st[3] += d;
}

// message word used by MD5 step i (RFC 1321 round schedules)
__device__ __forceinline__ unsigned md5_word_index(unsigned i) {
	if(i < 16) return i;
	if(i < 32) return (5 * i + 1) & 15;
	if(i < 48) return (3 * i + 5) & 15;
	return (7 * i) & 15;
}

// byte mask of the dword at byte position pos for a message of len bytes
__device__ __forceinline__ uint32_t keep4(size_t pos, size_t len) {
	if(pos + 4 <= len) return 0xFFFFFFFFu;
	if(pos >= len) return 0u;
	return (1u << (8 * (unsigned)(len - pos))) - 1u;
}

// ------------------------------------------------------------------ per slice

// The 64-byte block at byte `base` of a message of len bytes, zero past len.
__device__ __forceinline__ void load_block(const uint32_t* __restrict__ src, size_t base, size_t len, uint32_t m[16]) {
	if(base + 64 <= len) {
#pragma unroll
		for(int w = 0; w < 16; w += 4) {
			const uint4 v = *reinterpret_cast<const uint4*>(src + base / 4 + w);
			m[w] = v.x;
			m[w + 1] = v.y;
			m[w + 2] = v.z;
			m[w + 3] = v.w;
		}
	} else if(base >= len) {
#pragma unroll
		for(int w = 0; w < 16; w++) m[w] = 0;
	} else {
#pragma unroll
		for(int w = 0; w < 16; w++) m[w] = src[base / 4 + w] & keep4(base + 4 * w, len);
	}
}

// CRC-32 over 16 bytes (words w[0..3]) with the slice-by-16 tables: byte j of
// the 16 goes through table 15 - j, so only the first word waits for the
// running value and the other 12 lookups are independent of it.
__device__ __forceinline__ uint32_t crc_16bytes(uint32_t crc, const uint32_t* w, const uint32_t (*t)[256]) {
	const uint32_t x = crc ^ w[0];
	return t[15][x & 255] ^ t[14][(x >> 8) & 255] ^ t[13][(x >> 16) & 255] ^ t[12][x >> 24] ^ t[11][w[1] & 255] ^
	       t[10][(w[1] >> 8) & 255] ^ t[9][(w[1] >> 16) & 255] ^ t[8][w[1] >> 24] ^ t[7][w[2] & 255] ^
	       t[6][(w[2] >> 8) & 255] ^ t[5][(w[2] >> 16) & 255] ^ t[4][w[2] >> 24] ^ t[3][w[3] & 255] ^
	       t[2][(w[3] >> 8) & 255] ^ t[1][(w[3] >> 16) & 255] ^ t[0][w[3] >> 24];
}

__device__ __forceinline__ uint32_t crc_4bytes(uint32_t crc, uint32_t w, const uint32_t (*t)[256]) {
	const uint32_t x = crc ^ w;
	return t[3][x & 255] ^ t[2][(x >> 8) & 255] ^ t[1][(x >> 16) & 255] ^ t[0][x >> 24];
}

// a * b mod P for CRC-32 register values (reflected: bit 31 is the x^0
// coefficient, so multiplying by x is a right shift with the reflected
// polynomial folded back in).  Multiplying a register by x^(8n) mod P is
// feeding it n zero bytes.
__host__ __device__ inline uint32_t crc_mulmod(uint32_t a, uint32_t b) {
	uint32_t p = 0;
	for(int i = 31; i >= 0; i--) {
		if((a >> i) & 1u) p ^= b;
		b = (b >> 1) ^ ((b & 1u) ? 0xEDB88320u : 0u);
	}
	return p;
}

// Where slice k's hashes go: MD5 (16 bytes) at md5 + k * md5_stride, CRC-32
// (4 bytes, little endian) at crc + k * crc_stride.  Strides 16 and 4 give two
// arrays; 20 and 20 with crc = md5 + 16 give IFSC entries.
struct SliceHashOut {
	uint8_t* md5;
	size_t md5_stride;
	uint8_t* crc;
	size_t crc_stride;
};

// MD5 of slot k's bytes [0, lens[k]) followed by zeros up to slice_size, on
// one lane.  The chain is serial, so a file with few large slices is bound by
// it, not by bandwidth; the next block's loads are issued before the current
// block is hashed so that their latency hides behind it.
__device__ __forceinline__ void slice_md5_lane(const uint8_t* __restrict__ slots, size_t slot_bytes, unsigned k,
                                               const uint32_t* __restrict__ lens, size_t slice_size,
                                               const SliceHashOut& out) {
	const uint32_t* src = reinterpret_cast<const uint32_t*>(slots + (size_t)k * slot_bytes);
	size_t len = lens ? lens[k] : slice_size;
	if(len > slice_size) len = slice_size;
	uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
	const size_t full = slice_size / 64;
	uint32_t m[16], nx[16];
	if(full) load_block(src, 0, len, nx);
	for(size_t blk = 0; blk < full; blk++) {
#pragma unroll
		for(int w = 0; w < 16; w++) m[w] = nx[w];
		if(blk + 1 < full) load_block(src, (blk + 1) * 64, len, nx);
		md5_block(st, m);
	}
	// tail (slice_size % 64 bytes, a multiple of 4) + MD5 padding
	const size_t rem = slice_size - full * 64;
	const size_t base = full * 64;
#pragma unroll
	for(int w = 0; w < 16; w++) {
		const size_t pos = base + 4 * w;
		m[w] = pos < slice_size && pos < len ? src[pos / 4] & keep4(pos, len) : 0u;
	}
	// append 0x80 after the message
	m[rem / 4] |= 0x80u << (8 * (rem & 3));
	const uint64_t bits = (uint64_t)slice_size * 8;
	if(rem >= 56) {
		md5_block(st, m);
#pragma unroll
		for(int w = 0; w < 16; w++) m[w] = 0;
	}
	m[14] = (uint32_t)bits;
	m[15] = (uint32_t)(bits >> 32);
	md5_block(st, m);
	uint32_t* o = reinterpret_cast<uint32_t*>(out.md5 + out.md5_stride * k);
	o[0] = st[0];
	o[1] = st[1];
	o[2] = st[2];
	o[3] = st[3];
}

// How the CRC lanes split a slice of S bytes over L = 2^lanes_log2 lanes.
// The slice is read as lead = L*C - S zero bytes followed by the slice, in L
// chunks of C bytes: leading zeros leave a zero register at zero, so the
// register (started at zero) of that message is the slice's.  Lane j takes
// chunk j; pairs of neighbouring results combine as left * x^(8 * right
// length) + right, and every right-hand part at tree level l is 2^l whole
// chunks, so the factors are per launch.  The CRC's initial value
// 0xFFFFFFFF, carried through the S bytes, is added at the end.
struct CrcSplit {
	uint64_t chunk;       // C, a multiple of 16
	uint64_t lead;        // L * C - S
	uint32_t lanes_log2;  // L <= 64: a slice's lanes are in one wavefront
	uint32_t init;        // 0xFFFFFFFF * x^(8 S) mod P
	uint32_t x[6];        // x^(8 C 2^l) mod P for tree level l
};

// CRC-32 of slot k's bytes [0, lens[k]) followed by zeros up to slice_size,
// on L lanes per slice (slice-by-16 tables in LDS), combined in registers;
// `lane` numbers the (slice, chunk) pairs.  Every lane of the wavefront
// calls it: the combine shuffles read neighbouring lanes.
__device__ __forceinline__ void slice_crc_lanes(const uint8_t* __restrict__ slots, size_t slot_bytes, unsigned count,
                                                const uint32_t* __restrict__ lens, size_t slice_size,
                                                const SliceHashOut& out, const uint32_t (*s_crc)[256],
                                                const CrcSplit& sp, size_t lane) {
	const size_t k = lane >> sp.lanes_log2;
	const size_t j = lane & ((1u << sp.lanes_log2) - 1);
	uint32_t r = 0;
	if(k < count) {
		const uint32_t* w = reinterpret_cast<const uint32_t*>(slots + k * slot_bytes);
		size_t len = lens ? lens[k] : slice_size;
		if(len > slice_size) len = slice_size;
		const size_t c0 = j * sp.chunk, c1 = c0 + sp.chunk;  // chunk j of the lead-padded message
		size_t p = c0 > sp.lead ? c0 - sp.lead : 0;
		const size_t end = c1 > sp.lead ? (c1 - sp.lead < slice_size ? c1 - sp.lead : slice_size) : 0;
		if(p < len) {  // a chunk that starts past the data is all zeros: its register stays zero
			// words up to 64-byte alignment, 64-byte blocks with the next one's loads
			// in flight, then the remaining words
			for(; p < end && (p & 63); p += 4) r = crc_4bytes(r, w[p / 4] & keep4(p, len), s_crc);
			if(p + 64 <= end) {
				uint32_t m[16], nx[16];
				load_block(w, p, len, nx);
				for(; p + 64 <= end; p += 64) {
#pragma unroll
					for(int q = 0; q < 16; q++) m[q] = nx[q];
					if(p + 128 <= end) load_block(w, p + 64, len, nx);
#pragma unroll
					for(int q = 0; q < 16; q += 4) r = crc_16bytes(r, m + q, s_crc);
				}
			}
			for(; p < end; p += 4) r = crc_4bytes(r, w[p / 4] & keep4(p, len), s_crc);
		}
	}
	// every lane takes part in the shuffles; slice groups never straddle a wavefront
	for(unsigned l = 0; l < sp.lanes_log2; l++) {
		const uint32_t right = __shfl_down(r, 1u << l, 64);
		if(!(j & ((2u << l) - 1))) r = crc_mulmod(r, sp.x[l]) ^ right;
	}
	if(k < count && j == 0) *reinterpret_cast<uint32_t*>(out.crc + out.crc_stride * k) = ~(r ^ sp.init);
}

// Per-slice MD5 and CRC-32 in one launch: the first md5_blocks blocks run the
// MD5 chains (64 slices each, on the block's first wavefront, so that the
// chains spread over the CUs), the rest the CRC lanes, which fill the SIMDs
// the chains leave idle.
__global__ __launch_bounds__(256) void k_slice_hash(const uint8_t* __restrict__ slots, size_t slot_bytes,
                                                    unsigned count, const uint32_t* __restrict__ lens,
                                                    size_t slice_size, SliceHashOut out,
                                                    const uint32_t* __restrict__ crc_tab, CrcSplit sp,
                                                    unsigned md5_blocks) {
	__shared__ uint32_t s_crc[16][256];
	if(blockIdx.x < md5_blocks) {
		const unsigned k = blockIdx.x * 64 + threadIdx.x;
		if(threadIdx.x < 64 && k < count) slice_md5_lane(slots, slot_bytes, k, lens, slice_size, out);
		return;
	}
	for(unsigned i = threadIdx.x; i < 16 * 256; i += 256) s_crc[i >> 8][i & 255] = crc_tab[i];
	__syncthreads();
	slice_crc_lanes(slots, slot_bytes, count, lens, slice_size, out, s_crc, sp,
	                (size_t)(blockIdx.x - md5_blocks) * 256 + threadIdx.x);
}

// ------------------------------------------------------------------ recovery packets

// One call of the recovery-packet hash.  Region r's message is the packet from
// its recovery set ID on (lib/par2.js:505-513): set ID (16 bytes) ||
// "PAR 2.0\0RecvSlic" || u32le(exps[r]) || the recovery slice, which may come
// in pieces of even length (chunked mode, lib/par2.js:300-327).  Every region
// has the same length, so the fill of the partial block is uniform.
struct PacketMd5Call {
	uint32_t head[8];  // set ID, then the packet type, as little-endian words
	uint64_t done;     // message bytes of every region before this call (ignored with kPacketFirst)
	uint64_t n;        // data bytes this call appends to every region (even)
	unsigned flags;
};
constexpr unsigned kPacketFirst = 1, kPacketFinal = 2;

// Data bytes [o, o + 4) as a little-endian word (o even), zero past n (n even).
__device__ __forceinline__ uint32_t packet_word(const uint32_t* __restrict__ d, size_t o, size_t n) {
	if(!d || o >= n) return 0u;
	const uint32_t v = (o & 3) ? __builtin_amdgcn_alignbyte(d[(o + 2) / 4], d[(o - 2) / 4], 2) : d[o / 4];
	return o + 2 == n ? v & 0xFFFFu : v;
}

// `blocks` whole 64-byte blocks from data offset o (SHIFT: o = 2 mod 4, each
// word is the high half of one data word and the low half of the next), the
// next block's loads in flight while the current one is hashed.
template <bool SHIFT>
__device__ __forceinline__ void packet_blocks(uint32_t st[4], const uint32_t* __restrict__ d, size_t o, size_t blocks) {
	const uint32_t* p = d + (SHIFT ? (o + 2) / 4 : o / 4);
	uint32_t carry = SHIFT ? d[(o - 2) / 4] : 0u;
	uint32_t nx[16], m[16];
#pragma unroll
	for(int i = 0; i < 16; i++) nx[i] = p[i];
	for(size_t b = 0; b < blocks; b++) {
#pragma unroll
		for(int i = 0; i < 16; i++) m[i] = SHIFT ? __builtin_amdgcn_alignbyte(nx[i], i ? nx[i - 1] : carry, 2) : nx[i];
		if(SHIFT) carry = nx[15];
		if(b + 1 < blocks) {
#pragma unroll
			for(int i = 0; i < 16; i++) nx[i] = p[16 * (b + 1) + i];
		}
		md5_block(st, m);
	}
}

// Lane r: region r's MD5 state (4 words) and partial block (16 words, zero
// past the fill) live in state[20 r ...] between calls.  The partial block is
// completed from the data, whole blocks are hashed straight from the slot,
// and the rest becomes the new partial block; with kPacketFinal the message
// is padded (RFC 1321) and the digest written.  Slots NULL: the data is zeros.
__global__ __launch_bounds__(64) void k_packet_md5(uint32_t* __restrict__ state, const uint8_t* __restrict__ slots,
                                                   size_t slot_bytes, unsigned count,
                                                   const uint16_t* __restrict__ exps, PacketMd5Call c,
                                                   uint8_t* __restrict__ digest) {
	const unsigned r = blockIdx.x * 64 + threadIdx.x;
	if(r >= count) return;
	uint32_t* sv = state + (size_t)r * 20;
	uint32_t st[4], buf[16];
	uint64_t done = c.done;
	if(c.flags & kPacketFirst) {
		st[0] = 0x67452301u;
		st[1] = 0xefcdab89u;
		st[2] = 0x98badcfeu;
		st[3] = 0x10325476u;
#pragma unroll
		for(int k = 0; k < 16; k++) buf[k] = k < 8 ? c.head[k] : 0u;
		buf[8] = exps[r];
		done = 36;
	} else {
#pragma unroll
		for(int k = 0; k < 4; k++) st[k] = sv[k];
#pragma unroll
		for(int k = 0; k < 16; k++) buf[k] = sv[4 + k];
	}
	const uint32_t* d = slots ? reinterpret_cast<const uint32_t*>(slots + (size_t)r * slot_bytes) : nullptr;
	const size_t n = c.n;
	const unsigned f = (unsigned)(done & 63);  // bytes in the partial block
#pragma unroll
	for(int k = 0; k < 16; k++) {
		const int delta = 4 * k - (int)f;  // data offset of word k
		if(delta == -2) {
			if(d && n >= 2) buf[k] |= d[0] << 16;
		} else if(delta >= 0) {
			buf[k] = packet_word(d, (size_t)delta, n);
		}
	}
	if(f + n >= 64) {
		md5_block(st, buf);
		size_t o = 64 - f;
		const size_t blocks = (n - o) / 64;
		if(blocks) {
			if(!d) {
				uint32_t z[16];
#pragma unroll
				for(int k = 0; k < 16; k++) z[k] = 0u;
				for(size_t b = 0; b < blocks; b++) md5_block(st, z);
			} else if(o & 3) {
				packet_blocks<true>(st, d, o, blocks);
			} else {
				packet_blocks<false>(st, d, o, blocks);
			}
		}
		o += 64 * blocks;
#pragma unroll
		for(int k = 0; k < 16; k++) buf[k] = packet_word(d, o + 4 * k, n);
	}
	done += n;
	if(c.flags & kPacketFinal) {
		const unsigned fill = (unsigned)(done & 63);
#pragma unroll
		for(int k = 0; k < 16; k++)
			if(k == (int)(fill / 4)) buf[k] |= 0x80u << (8 * (fill & 3));
		if(fill >= 56) {
			md5_block(st, buf);
#pragma unroll
			for(int k = 0; k < 16; k++) buf[k] = 0u;
		}
		const uint64_t bits = done * 8;
		buf[14] = (uint32_t)bits;
		buf[15] = (uint32_t)(bits >> 32);
		md5_block(st, buf);
		uint32_t* o = reinterpret_cast<uint32_t*>(digest + 16 * (size_t)r);
#pragma unroll
		for(int k = 0; k < 4; k++) o[k] = st[k];
	} else {
#pragma unroll
		for(int k = 0; k < 4; k++) sv[k] = st[k];
#pragma unroll
		for(int k = 0; k < 16; k++) sv[4 + k] = buf[k];
	}
}

// ------------------------------------------------------------------ per file

// One wavefront runs one chain.  The state is kept in vector registers (made
// non-uniform through an opaque zero) so that each step is a handful of VALU
// operations (bit-select, three-way add, align-bit rotate, add) with the
// message words as scalar operands: lane l loads word l of four consecutive
// 64-byte blocks, and v_readlane hands each word to the chain.
__device__ void md5_span(const Md5Span& e) {
	uint32_t z;
	asm volatile("v_mov_b32 %0, 0" : "=v"(z));
	const unsigned lane = threadIdx.x & 63;
	uint32_t st[4];
	if(e.flags & kMd5SpanFirst) {
		st[0] = 0x67452301u ^ z;
		st[1] = 0xefcdab89u ^ z;
		st[2] = 0x98badcfeu ^ z;
		st[3] = 0x10325476u ^ z;
	} else {
		for(int i = 0; i < 4; i++) st[i] = e.state[i] ^ z;
	}
	const uint64_t wps = e.slice_size / 4;  // words per slot
	auto word_at = [&](uint64_t wi) -> uint32_t {
		const uint64_t sl = wi / wps;
		const uint64_t off = wi - sl * wps;
		return reinterpret_cast<const uint32_t*>(e.slots + sl * e.slot_bytes)[off];
	};
	const bool last = e.flags & kMd5SpanFinal;
	uint64_t blk = e.begin / 64;
	const uint64_t blk_end = last ? e.file_size / 64 : e.end / 64;
	// Groups of four blocks: lane l holds word l of the group; for step i = l
	// of block b it forms K[i] + m[b][g(i)] (a lane shuffle and an add), the
	// 256 sums go to LDS, and the chain reads them back four at a time at a
	// uniform address (one broadcast ds_read_b128 per four steps).
	__shared__ uint4 s_km[64];
	uint32_t k_lane = 0;
#pragma unroll
	for(int i = 0; i < 64; i++)
		if((unsigned)i == lane) k_lane = kMd5K[i];
	const unsigned g_lane = md5_word_index(lane);
	uint32_t m[16];
	uint32_t v = blk + 4 <= blk_end ? word_at(blk * 16 + lane) : 0u;
	for(; blk + 4 <= blk_end; blk += 4) {
		uint32_t vn = 0;
		if(blk + 8 <= blk_end) vn = word_at((blk + 4) * 16 + lane);
		__syncthreads();  // the previous group's sums have been read
		uint32_t* skm = reinterpret_cast<uint32_t*>(s_km);
#pragma unroll
		for(int b = 0; b < 4; b++) skm[64 * b + lane] = k_lane + __shfl(v, 16 * b + g_lane, 64);
		__syncthreads();
#pragma unroll
		for(int b = 0; b < 4; b++) {
			uint32_t km[64];
#pragma unroll
			for(int q = 0; q < 16; q++) {
				const uint4 t = s_km[16 * b + q];
				km[4 * q] = t.x;
				km[4 * q + 1] = t.y;
				km[4 * q + 2] = t.z;
				km[4 * q + 3] = t.w;
			}
			md5_block_km(st, km);
		}
		v = vn;
	}
	for(; blk < blk_end; blk++) {
		const uint32_t u = word_at(blk * 16 + (lane & 15));
#pragma unroll
		for(int w = 0; w < 16; w++) m[w] = __builtin_amdgcn_readlane(u, w);
		md5_block(st, m);
	}
	if(last) {
		const uint64_t fs = e.file_size;
		const uint64_t rem = fs - blk_end * 64;
#pragma unroll
		for(int w = 0; w < 16; w++) {
			const uint64_t pos = blk_end * 64 + 4 * (uint64_t)w;
			uint32_t x = 0;
			if(pos < fs) x = __builtin_amdgcn_readfirstlane(word_at(blk_end * 16 + w)) & keep4(pos, fs);
			m[w] = x;
		}
		m[rem / 4] |= 0x80u << (8 * (rem & 3));
		const uint64_t bits = fs * 8;
		if(rem >= 56) {
			md5_block(st, m);
#pragma unroll
			for(int w = 0; w < 16; w++) m[w] = 0;
		}
		m[14] = (uint32_t)bits;
		m[15] = (uint32_t)(bits >> 32);
		md5_block(st, m);
	}
	if(lane == 0) {
		uint32_t* o = last ? reinterpret_cast<uint32_t*>(e.digest) : e.state;
		for(int i = 0; i < 4; i++) o[i] = st[i];
	}
}

__global__ __launch_bounds__(64) void k_file_md5(Md5Span e) { md5_span(e); }

// One wavefront per span: the chains of many files advance together.
__global__ __launch_bounds__(64) void k_file_md5_round(const Md5Span* __restrict__ spans) {
	const Md5Span e = spans[blockIdx.x];
	md5_span(e);
}

namespace {
// x^(8n) mod P by square and multiply (x^1 is 0x40000000 in the reflected form)
uint32_t crc_xpow8n(uint64_t n) {
	uint32_t r = 0x80000000u, sq = 0x40000000u;
	for(uint64_t e = 8 * n; e; e >>= 1) {
		if(e & 1) r = crc_mulmod(r, sq);
		sq = crc_mulmod(sq, sq);
	}
	return r;
}

CrcSplit crc_split(size_t slice_size) {
	// about 4 KiB or more per lane, at most one wavefront per slice
	unsigned lg = 0;
	while(lg < 6 && (slice_size >> (lg + 1)) >= 4096) lg++;
	CrcSplit sp{};
	const size_t L = (size_t)1 << lg;
	sp.lanes_log2 = lg;
	sp.chunk = ((slice_size + L - 1) / L + 15) / 16 * 16;
	sp.lead = L * sp.chunk - slice_size;
	sp.init = crc_mulmod(0xFFFFFFFFu, crc_xpow8n(slice_size));
	for(unsigned l = 0; l < lg; l++) sp.x[l] = crc_xpow8n(sp.chunk << l);
	return sp;
}
}  // namespace

hipError_t launch_slice_hash_to(const uint8_t* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                                size_t slice_size, uint8_t* md5, size_t md5_stride, uint8_t* crc, size_t crc_stride,
                                const DeviceTables& t, hipStream_t s) {
	if(!count) return hipSuccess;
	if(slice_size % 4 || slice_size > slot_bytes || slot_bytes % 16 || !t.crc) return hipErrorInvalidValue;
	if(md5_stride % 4 || crc_stride % 4 || (size_t)md5 % 4 || (size_t)crc % 4) return hipErrorInvalidValue;
	const CrcSplit sp = crc_split(slice_size);
	const SliceHashOut out{md5, md5_stride, crc, crc_stride};
	const size_t md5_blocks = (count + 63) / 64, crc_blocks = (((size_t)count << sp.lanes_log2) + 255) / 256;
	hipLaunchKernelGGL(k_slice_hash, dim3((unsigned)(md5_blocks + crc_blocks)), dim3(256), 0, s, slots, slot_bytes,
	                   count, lens, slice_size, out, t.crc, sp, (unsigned)md5_blocks);
	return hipGetLastError();
}

hipError_t launch_slice_hash(const uint8_t* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                             size_t slice_size, uint8_t* md5_out, uint32_t* crc_out, const DeviceTables& t,
                             hipStream_t s) {
	return launch_slice_hash_to(slots, slot_bytes, count, lens, slice_size, md5_out, 16,
	                            reinterpret_cast<uint8_t*>(crc_out), 4, t, s);
}

hipError_t launch_file_md5(const uint8_t* slots, size_t slot_bytes, unsigned count, size_t slice_size,
                           uint64_t file_size, uint8_t* md5_out, hipStream_t s) {
	if(slice_size % 4 || slice_size > slot_bytes || !slice_size) return hipErrorInvalidValue;
	if(file_size > (uint64_t)count * slice_size) return hipErrorInvalidValue;
	Md5Span e{};
	e.slots = slots;
	e.slot_bytes = slot_bytes;
	e.slice_size = slice_size;
	e.file_size = file_size;
	e.begin = 0;
	e.end = file_size;
	e.state = nullptr;
	e.digest = md5_out;
	e.flags = kMd5SpanFirst | kMd5SpanFinal;
	hipLaunchKernelGGL(k_file_md5, dim3(1), dim3(64), 0, s, e);
	return hipGetLastError();
}

hipError_t launch_file_md5_round(const void* spans, unsigned n, hipStream_t s) {
	if(!n) return hipSuccess;
	hipLaunchKernelGGL(k_file_md5_round, dim3(n), dim3(64), 0, s, reinterpret_cast<const Md5Span*>(spans));
	return hipGetLastError();
}

hipError_t launch_packet_md5(uint32_t* state, const uint8_t* slots, size_t slot_bytes, unsigned count,
                             const uint16_t* exps, const uint32_t head[8], uint64_t done, uint64_t n, bool first,
                             bool final, uint8_t* digest, hipStream_t s) {
	if(!count) return hipSuccess;
	if(n % 2 || (slots && (slot_bytes % 16 || n + 4 > slot_bytes))) return hipErrorInvalidValue;
	PacketMd5Call c{};
	for(int k = 0; k < 8; k++) c.head[k] = head[k];
	c.done = done;
	c.n = n;
	c.flags = (first ? kPacketFirst : 0u) | (final ? kPacketFinal : 0u);
	hipLaunchKernelGGL(k_packet_md5, dim3((count + 63) / 64), dim3(64), 0, s, state, slots, slot_bytes, count, exps, c,
	                   digest);
	return hipGetLastError();
}

}  // namespace pph
