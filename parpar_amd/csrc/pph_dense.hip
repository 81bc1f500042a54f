// Dense form: GF(2^16) mul-add as a GF(2) contraction on int8 MFMA (gfx950).
//
// Multiplication by a coefficient c is linear over GF(2): bit j of c*x is
// XOR_b M_c[j][b] x_b with M_c[j][b] = bit j of (c * z^b mod 0x1100B).  So
//     out_r bit j (word w) = parity( sum_i sum_b M_{c(r,i)}[j][b] * bit_b(in_i[w]) )
// which is an integer GEMM with rows (r, j), reduction index (i, b) and columns
// w, on 0/1 int8 operands: v_mfma_i32_32x32x32_i8 accumulates exact int32 sums
// (at most 16 * inputs <= 2^19) and the low bit is the GF(2) result.
//
// Workgroup: 8 waves (two per SIMD), 8 recovery slices x 512 words.  Wave q
// owns words [64q, 64q+64) of the block as 4 x 2 MFMA tiles of 32 x 32 (128
// accumulators).  A K-step covers 2 inputs (K = 32).  The A fragments (rows of
// the coefficient bit-matrices) are the same for all waves: waves q and q+4
// build fragment q for the next chunk into LDS (double buffered, one barrier per
// chunk of steps); every wave expands its own B fragments (16 bits of one
// input word per lane -> 16 bytes).
//
// Row j of M_c, as a 16-bit mask over b, is XOR_k c_k Q_j[k] with
// Q_j[k] = sum_b 2^b * bit j of z^(k+b) mod p; it is read from four
// 16-entry nibble tables T[j][q][n] (2 KiB in LDS).
#include "pph_internal.h"

namespace pph {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr unsigned kDenseWaves = 8;    // two waves per SIMD: one's VALU overlaps the other's MFMAs
constexpr unsigned kDenseNT = 2;       // 32-word MFMA column tiles per wave
constexpr unsigned kDenseThreads = 64 * kDenseWaves;
constexpr unsigned kDenseRows = 8;     // recovery slices per workgroup
constexpr unsigned kDenseCols = kDenseWaves * 32 * kDenseNT;  // words per workgroup (512)
constexpr unsigned kRmStride = 66;     // LDS row stride of the rowmask tables (odd dword count: no bank aliasing)

// bit b of x (b < 16) -> byte b, as 0 or 1
__device__ __forceinline__ v4i expand16(uint32_t x) {
	v4i r;
	r[0] = (int)((__umul24(x & 0xFu, 0x00204081u)) & 0x01010101u);
	r[1] = (int)((__umul24((x >> 4) & 0xFu, 0x00204081u)) & 0x01010101u);
	r[2] = (int)((__umul24((x >> 8) & 0xFu, 0x00204081u)) & 0x01010101u);
	r[3] = (int)((__umul24((x >> 12) & 0xFu, 0x00204081u)) & 0x01010101u);
	return r;
}

__device__ __forceinline__ uint32_t mod65535d(uint32_t x) {
	x = (x & 0xFFFFu) + (x >> 16);
	x = (x & 0xFFFFu) + (x >> 16);
	return x == 65535u ? 0u : x;
}

__device__ __forceinline__ uint32_t rowmask(const uint16_t* t_j, uint32_t c) {
	return (uint32_t)t_j[c & 15u] ^ (uint32_t)t_j[16 + ((c >> 4) & 15u)] ^ (uint32_t)t_j[32 + ((c >> 8) & 15u)] ^
	       (uint32_t)t_j[48 + (c >> 12)];
}

constexpr unsigned kDenseChunk = 4;        // K-steps (2 inputs each) per pipeline chunk
constexpr unsigned kDenseMaxInputs = 32768;  // PAR2 input limit; logs resident in LDS

// The K loop runs in chunks of kDenseChunk steps.  While the MFMAs of chunk k
// run, the wave has already issued the word loads of chunk k+1 and the
// coefficient gathers of chunk k+2, and builds the A fragments of chunk k+1 from
// coefficients gathered one iteration earlier: global latency stays hidden
// behind 128 MFMAs, and there is one barrier per chunk.
__global__ __launch_bounds__(kDenseThreads) void k_muladd_dense(
    const uint16_t* __restrict__ in, size_t slot_words, unsigned n_in, const uint16_t* __restrict__ in_logs,
    const uint16_t* __restrict__ exps, uint16_t* __restrict__ out, unsigned n_out, int accumulate,
    const uint16_t* __restrict__ exp_tab, const uint16_t* __restrict__ rm_tab, unsigned row_blocks,
    unsigned col_blocks) {
	__shared__ uint16_t s_rm[16 * kRmStride];
	__shared__ uint16_t s_logs[kDenseMaxInputs];
	__shared__ v4i s_a[2][kDenseChunk][4][64];
	for(unsigned k = threadIdx.x; k < 16 * 64; k += kDenseThreads) s_rm[(k >> 6) * kRmStride + (k & 63)] = rm_tab[k];
	for(unsigned k = threadIdx.x; k < n_in; k += kDenseThreads) s_logs[k] = in_logs[k];

	// blocks of one column block run on the same XCD (blockIdx round-robins over 8 XCDs)
	const unsigned total = row_blocks * col_blocks;
	unsigned L = blockIdx.x;
	if((total & 7u) == 0) L = (blockIdx.x & 7u) * (total >> 3) + (blockIdx.x >> 3);
	const unsigned cb = L / row_blocks;
	const unsigned rb = L - cb * row_blocks;

	const unsigned q = threadIdx.x >> 6;     // wave
	const unsigned lane = threadIdx.x & 63u;
	const unsigned n = lane & 31u;           // MFMA column / A row within the tile
	const unsigned g = lane >> 5;            // K group: which of the step's two inputs
	const unsigned r_base = rb * kDenseRows;
	const size_t w_base = (size_t)cb * kDenseCols + q * (32 * kDenseNT);

	// the A row this lane builds: fragment q%4, slice r_base + 2(q%4) + n/16, bit
	// j = n%16; waves q and q+4 build the even / odd steps of each chunk
	const unsigned fq = q & 3u;
	const unsigned t_par = q >> 2;
	const unsigned r_a = r_base + 2 * fq + (n >> 4);
	const unsigned j = n & 15u;
	const bool row_ok = r_a < n_out;
	const uint32_t e_a = row_ok ? exps[r_a] : 0u;
	const uint16_t* t_j = s_rm + j * kRmStride;
	__syncthreads();

	v16i acc[4][kDenseNT];
#pragma unroll
	for(int a = 0; a < 4; a++)
#pragma unroll
		for(int b = 0; b < (int)kDenseNT; b++)
#pragma unroll
			for(int v = 0; v < 16; v++) acc[a][b][v] = 0;

	const unsigned steps = (n_in + 1) / 2;
	const unsigned chunks = (steps + kDenseChunk - 1) / kDenseChunk;
	const uint16_t* in_lane = in + w_base + kDenseNT * n;
	constexpr unsigned kBuild = kDenseChunk / 2;  // steps of a chunk this wave builds A for

	uint32_t cv_next[kBuild], cv_next2[kBuild];
	uint32_t w_cur[kDenseChunk], w_next[kDenseChunk];
	auto gather = [&](unsigned chunk, uint32_t* cv) {
#pragma unroll
		for(unsigned u = 0; u < kBuild; u++) {
			const unsigned i = 2 * (chunk * kDenseChunk + 2 * u + t_par) + g;
			cv[u] = (i < n_in && row_ok) ? (uint32_t)exp_tab[mod65535d((uint32_t)s_logs[i] * e_a)] : 0u;
		}
	};
	auto load_words = [&](unsigned chunk, uint32_t* w) {
#pragma unroll
		for(unsigned t = 0; t < kDenseChunk; t++) {
			const unsigned i = 2 * (chunk * kDenseChunk + t) + g;
			w[t] = i < n_in ? *reinterpret_cast<const uint32_t*>(in_lane + (size_t)i * slot_words) : 0u;
		}
	};
	auto build_a = [&](unsigned chunk, const uint32_t* cv) {
#pragma unroll
		for(unsigned u = 0; u < kBuild; u++)
			s_a[chunk & 1][2 * u + t_par][fq][lane] = expand16(rowmask(t_j, cv[u]));
	};

	gather(0, cv_next);
	build_a(0, cv_next);
	gather(1, cv_next);
	load_words(0, w_cur);
	__syncthreads();
	for(unsigned k = 0; k < chunks; k++) {
		gather(k + 2, cv_next2);
		load_words(k + 1, w_next);
		if(k + 1 < chunks) build_a(k + 1, cv_next);
#pragma unroll
		for(unsigned t = 0; t < kDenseChunk; t++) {
			v4i a[4], b[kDenseNT];
#pragma unroll
			for(int f = 0; f < 4; f++) a[f] = s_a[k & 1][t][f][lane];
			b[0] = expand16(w_cur[t] & 0xFFFFu);
			b[1] = expand16(w_cur[t] >> 16);
#pragma unroll
			for(int fa = 0; fa < 4; fa++)
#pragma unroll
				for(int fb = 0; fb < (int)kDenseNT; fb++)
					acc[fa][fb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[fa], b[fb], acc[fa][fb], 0, 0, 0);
		}
		__syncthreads();
#pragma unroll
		for(unsigned t = 0; t < kDenseChunk; t++) w_cur[t] = w_next[t];
#pragma unroll
		for(unsigned u = 0; u < kBuild; u++) cv_next[u] = cv_next2[u];
	}

	// parity -> words.  32x32 i32 result: lane l, element v holds row
	// 8*(v/4) + 4*(l/32) + v%4, column l%32.  Rows 16h..16h+15 are the 16 bits of
	// slice 2*fa + h; this lane holds 8 of them, its partner lane l^32 the rest.
	// Column n of tile fb is word w_base + NT*n + fb.
#pragma unroll
	for(int fa = 0; fa < 4; fa++) {
		uint32_t word[kDenseNT][2];
#pragma unroll
		for(int fb = 0; fb < (int)kDenseNT; fb++) {
#pragma unroll
			for(int h = 0; h < 2; h++) {
				uint32_t bits = 0;
#pragma unroll
				for(int v = 8 * h; v < 8 * h + 8; v++) {
					const unsigned jj = 8 * ((v >> 2) & 1) + 4 * g + (v & 3);
					bits |= ((uint32_t)acc[fa][fb][v] & 1u) << jj;
				}
				word[fb][h] = bits | (uint32_t)__shfl_xor((int)bits, 32);
			}
		}
		// lane group g writes slice 2*fa + g: words w_base + NT*n .. +NT-1
		const unsigned r = r_base + 2 * fa + g;
		if(r < n_out) {
			const uint32_t o = (g ? word[0][1] : word[0][0]) | ((g ? word[1][1] : word[1][0]) << 16);
			uint32_t* dst = reinterpret_cast<uint32_t*>(out + (size_t)r * slot_words + w_base + kDenseNT * n);
			*dst = accumulate ? (o ^ *dst) : o;
		}
	}
}

hipError_t launch_muladd_dense(const uint16_t* in, size_t slot_words, unsigned n_in, const uint16_t* in_logs,
                               const uint16_t* exps, uint16_t* out, unsigned n_out, bool accumulate,
                               const DeviceTables& t, hipStream_t s) {
	if(!n_out || !slot_words) return hipSuccess;
	if(slot_words % kDenseCols) return hipErrorInvalidValue;
	if(n_in && (!in_logs || !exps || !in)) return hipErrorInvalidValue;
	if(n_in > kDenseMaxInputs) return hipErrorInvalidValue;
	if(!t.rowmask) return hipErrorInvalidValue;
	const unsigned row_blocks = (n_out + kDenseRows - 1) / kDenseRows;
	const unsigned col_blocks = (unsigned)(slot_words / kDenseCols);
	hipLaunchKernelGGL(k_muladd_dense, dim3(row_blocks * col_blocks), dim3(kDenseThreads), 0, s, in, slot_words, n_in,
	                   in_logs, exps, out, n_out, accumulate ? 1 : 0, t.exp, t.rowmask, row_blocks, col_blocks);
	return hipGetLastError();
}

}  // namespace pph
