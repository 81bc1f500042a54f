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

// Lane k hashes slot k: bytes [0, lens[k]) followed by zeros up to slice_size.
// CRC-32 by 4-byte steps through the slice-by-4 tables in LDS.
__global__ __launch_bounds__(64) void k_slice_hash(const uint8_t* __restrict__ slots, size_t slot_bytes, unsigned count,
                                                   const uint32_t* __restrict__ lens, size_t slice_size,
                                                   uint8_t* __restrict__ md5_out, uint32_t* __restrict__ crc_out,
                                                   const uint32_t* __restrict__ crc_tab) {
	__shared__ uint32_t s_crc[4][256];
	for(unsigned k = threadIdx.x; k < 1024; k += 64) s_crc[k >> 8][k & 255] = crc_tab[k];
	__syncthreads();
	const unsigned k = blockIdx.x * 64 + threadIdx.x;
	if(k >= count) return;
	const uint32_t* src = reinterpret_cast<const uint32_t*>(slots + (size_t)k * slot_bytes);
	size_t len = lens ? lens[k] : slice_size;
	if(len > slice_size) len = slice_size;
	uint32_t st[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
	uint32_t crc = 0xFFFFFFFFu;
	const size_t full = slice_size / 64;
	uint32_t m[16];
	for(size_t blk = 0; blk < full; blk++) {
		const size_t base = blk * 64;
		if(base + 64 <= len) {
#pragma unroll
			for(int w = 0; w < 16; w += 4) {
				const uint4 v = *reinterpret_cast<const uint4*>(src + base / 4 + w);
				m[w] = v.x; m[w + 1] = v.y; m[w + 2] = v.z; m[w + 3] = v.w;
			}
		} else if(base >= len) {
#pragma unroll
			for(int w = 0; w < 16; w++) m[w] = 0;
		} else {
#pragma unroll
			for(int w = 0; w < 16; w++) m[w] = src[base / 4 + w] & keep4(base + 4 * w, len);
		}
#pragma unroll
		for(int w = 0; w < 16; w++) {
			const uint32_t x = crc ^ m[w];
			crc = s_crc[3][x & 255] ^ s_crc[2][(x >> 8) & 255] ^ s_crc[1][(x >> 16) & 255] ^ s_crc[0][x >> 24];
		}
		md5_block(st, m);
	}
	// tail (slice_size % 64 bytes, a multiple of 4 for PAR2 slices) + MD5 padding
	const size_t rem = slice_size - full * 64;
	const size_t base = full * 64;
#pragma unroll
	for(int w = 0; w < 16; w++) {
		const size_t pos = base + 4 * w;
		uint32_t v = 0;
		if(pos < slice_size) {
			v = pos < len ? src[pos / 4] & keep4(pos, len) : 0u;
			const unsigned nb = (unsigned)(slice_size - pos < 4 ? slice_size - pos : 4);
			v &= nb == 4 ? 0xFFFFFFFFu : ((1u << (8 * nb)) - 1u);
			// CRC over the remaining bytes of the slice, byte by byte
			for(unsigned bb = 0; bb < nb; bb++) crc = s_crc[0][(crc ^ (v >> (8 * bb))) & 255] ^ (crc >> 8);
		}
		m[w] = v;
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
	uint32_t* o = reinterpret_cast<uint32_t*>(md5_out + 16 * (size_t)k);
	o[0] = st[0];
	o[1] = st[1];
	o[2] = st[2];
	o[3] = st[3];
	crc_out[k] = crc ^ 0xFFFFFFFFu;
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

hipError_t launch_slice_hash(const uint8_t* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                             size_t slice_size, uint8_t* md5_out, uint32_t* crc_out, const DeviceTables& t,
                             hipStream_t s) {
	if(!count) return hipSuccess;
	if(slice_size % 4 || slice_size > slot_bytes || slot_bytes % 16 || !t.crc) return hipErrorInvalidValue;
	hipLaunchKernelGGL(k_slice_hash, dim3((count + 63) / 64), dim3(64), 0, s, slots, slot_bytes, count, lens, slice_size,
	                   md5_out, crc_out, t.crc);
	return hipGetLastError();
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

}  // namespace pph
