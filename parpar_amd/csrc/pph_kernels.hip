// HIP kernels for gfx950: synthetic fill, input preparation (zero padding,
// gf16_cksum, log-domain conversion), the plain mul-add and checksum
// verification.  Launchers validate shapes on the host before launching.
#include "pph_internal.h"
#include "pph_plain_tile.h"

namespace pph {

// ------------------------------------------------------------------ helpers

// Two GF(2^16) lanes packed in a dword, each multiplied by 2 (x) modulo 0x1100B:
// the lane-wise Horner step of gf16_cksum (gf16/gf16_checksum_generic.h:25-39).
__device__ __forceinline__ uint32_t mul2x2(uint32_t v) {
	return ((v << 1) & 0xFFFEFFFEu) ^ (((v >> 15) & 0x00010001u) * (kPoly & 0xFFFFu));
}

// v * 2^e for one lane, e < 65535, through the global log/exp tables.
__device__ __forceinline__ uint32_t mul_pow2(uint32_t v, uint32_t e, const uint16_t* __restrict__ lg,
                                             const uint16_t* __restrict__ ex) {
	if(!v) return 0;
	uint32_t s = (uint32_t)lg[v] + e;
	if(s >= kOrder) s -= kOrder;
	return ex[s];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t slice_key(uint64_t seed, uint32_t job, uint32_t slice) {
	auto m = [](uint64_t z) {
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	};
	uint64_t k = m(seed * kGolden + 0x632BE59BD9B4E019ull);
	return m(k ^ (((uint64_t)job << 32) | (uint64_t)slice));
}

// byte mask keeping positions < len of the dword starting at byte `pos`
__device__ __forceinline__ uint32_t keep_mask(size_t pos, size_t len) {
	if(pos + 4 <= len) return 0xFFFFFFFFu;
	if(pos >= len) return 0u;
	return (1u << (8 * (unsigned)(len - pos))) - 1u;
}

// ------------------------------------------------------------------ fill

__global__ void k_fill_random(uint8_t* __restrict__ slots, size_t slot_bytes, unsigned count, size_t len,
                              uint64_t seed, uint32_t job, uint32_t first_slice) {
	const size_t pieces_per_slot = slot_bytes / 16;
	const size_t total = pieces_per_slot * count;
	for(size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (size_t)gridDim.x * blockDim.x) {
		const unsigned k = (unsigned)(p / pieces_per_slot);
		const size_t off = (p - (size_t)k * pieces_per_slot) * 16;
		uint4 v = make_uint4(0, 0, 0, 0);
		if(off < len) {
			const uint64_t key = slice_key(seed, job, first_slice + k);
			const uint64_t w = off >> 3;
			const uint64_t a = mix64(key + (w + 1) * kGolden);
			const uint64_t b = mix64(key + (w + 2) * kGolden);
			v.x = (uint32_t)a & keep_mask(off, len);
			v.y = (uint32_t)(a >> 32) & keep_mask(off + 4, len);
			v.z = (uint32_t)b & keep_mask(off + 8, len);
			v.w = (uint32_t)(b >> 32) & keep_mask(off + 12, len);
		}
		*reinterpret_cast<uint4*>(slots + (size_t)k * slot_bytes + off) = v;
	}
}

hipError_t launch_fill_random(uint8_t* slots, size_t slot_bytes, unsigned count, size_t len,
                              uint64_t seed, uint32_t job, uint32_t first_slice, hipStream_t s) {
	if(!count) return hipSuccess;
	const size_t total = slot_bytes / 16 * count;
	size_t blocks = (total + 255) / 256;
	if(blocks > 65536) blocks = 65536;
	hipLaunchKernelGGL(k_fill_random, dim3((unsigned)blocks), dim3(256), 0, s, slots, slot_bytes, count, len, seed,
	                   job, first_slice);
	return hipGetLastError();
}

// ------------------------------------------------------------------ checksum

// gf16_cksum of a slot on the device.  The slice is a sequence of CK-byte rows
// (zero padded); the checksum is the lane-wise Horner sum
//     ck = sum_k 2^(nb-1-k) * row_k,   nb = ceil(slice_size / CK)
// which is what gf16_cksum_copy stores (gf16/gf16_cksum_base.h:32-46,
// including the 2^n step for zero rows after a short input).  One workgroup per
// slot: each group of TPR threads runs Horner over a contiguous run of rows, the
// partial sums are scaled by 2^(rows after the run) and XOR-reduced in LDS.
constexpr unsigned kCkThreads = 256;     // verification: one workgroup per slot
constexpr unsigned kPrepThreads = 1024;  // preparation: persistent, log table in LDS

template <unsigned CK, unsigned T>
struct CkGeom {
	static constexpr unsigned P = CK < 16 ? CK : 16;  // bytes per thread per row
	static constexpr unsigned TPR = CK / P;           // threads per row
	static constexpr unsigned NG = T / TPR;           // row groups
	static constexpr unsigned DW = P / 4;             // dwords per piece
};

template <unsigned DW>
struct Piece {
	uint32_t d[DW];
};

template <unsigned DW>
__device__ __forceinline__ Piece<DW> load_piece(const uint8_t* p) {
	Piece<DW> r;
	if constexpr(DW == 4) {
		uint4 v = *reinterpret_cast<const uint4*>(p);
		r.d[0] = v.x; r.d[1] = v.y; r.d[2] = v.z; r.d[3] = v.w;
	} else {
		uint2 v = *reinterpret_cast<const uint2*>(p);
		r.d[0] = v.x; r.d[1] = v.y;
	}
	return r;
}
template <unsigned DW>
__device__ __forceinline__ void store_piece(uint8_t* p, const Piece<DW>& r) {
	if constexpr(DW == 4) {
		*reinterpret_cast<uint4*>(p) = make_uint4(r.d[0], r.d[1], r.d[2], r.d[3]);
	} else {
		*reinterpret_cast<uint2*>(p) = make_uint2(r.d[0], r.d[1]);
	}
}

// Horner partials of one workgroup over slot rows [0, nrows) of `src`; rows >= nb
// do not enter the checksum.  MODE 0: verify (read only), 1: prep to data form
// (write the zero-padded slice to `slot`), 2: prep to log domain (write every
// word of `slot` in log domain through the table `lg`).  src may equal slot.
// Rows are read four at a time so that several loads are in flight per thread.
template <unsigned CK, unsigned T, int MODE>
__device__ void ck_rows(const uint8_t* src, uint8_t* slot, size_t nrows, size_t nb, size_t len, uint32_t* h,
                        const uint16_t* lg, size_t& run_end) {
	using G = CkGeom<CK, T>;
	constexpr unsigned U = 4;
	const bool inplace = src == slot;
	const unsigned t = threadIdx.x % G::TPR;
	const unsigned g = threadIdx.x / G::TPR;
	const size_t rpg = (nrows + G::NG - 1) / G::NG;
	const size_t r0 = (size_t)g * rpg;
	size_t r1 = r0 + rpg;
	if(r1 > nrows) r1 = nrows;
	for(unsigned j = 0; j < G::DW; j++) h[j] = 0;
	for(size_t rb = r0; rb < r1; rb += U) {
		Piece<G::DW> v[U];
#pragma unroll
		for(unsigned u = 0; u < U; u++) {
			if(rb + u < r1) v[u] = load_piece<G::DW>(src + (rb + u) * CK + (size_t)t * G::P);
		}
#pragma unroll
		for(unsigned u = 0; u < U; u++) {
			const size_t row = rb + u;
			if(row >= r1) break;
			const size_t off = row * CK + (size_t)t * G::P;
			const bool partial = off + G::P > len;
			if(partial) {
				for(unsigned j = 0; j < G::DW; j++) v[u].d[j] &= keep_mask(off + 4 * j, len);
			}
			if(row < nb) {
				for(unsigned j = 0; j < G::DW; j++) h[j] = mul2x2(h[j]) ^ v[u].d[j];
			}
			if constexpr(MODE == 1) {
				if(partial || !inplace) store_piece<G::DW>(slot + off, v[u]);
			} else if constexpr(MODE == 2) {
				Piece<G::DW> o;
				for(unsigned j = 0; j < G::DW; j++) {
					uint32_t lo = lg[v[u].d[j] & 0xFFFFu];
					uint32_t hi = lg[v[u].d[j] >> 16];
					o.d[j] = lo | (hi << 16);
				}
				store_piece<G::DW>(slot + off, o);
			}
		}
	}
	run_end = r1 < nb ? r1 : nb;
	if(r0 >= nb) run_end = (size_t)-1;  // run contributes nothing
}

// The rows of ck_rows for in-place data-form preparation (MODE 1, src ==
// slot), in three runs per thread: rows wholly inside the input (read eight
// at a time, no mask and no store: the common case, with few registers live,
// so that few workgroups per CU still keep enough loads in flight), the row
// holding the input's end (masked and stored), and the rows past it (zeros
// stored without a read; Horner steps of zero while in the checksum).
template <unsigned CK, unsigned T>
__device__ void ck_rows_inplace(uint8_t* slot, size_t nrows, size_t nb, size_t len, uint32_t* h, size_t& run_end) {
	using G = CkGeom<CK, T>;
	constexpr unsigned U = 8;
	const unsigned t = threadIdx.x % G::TPR;
	const unsigned g = threadIdx.x / G::TPR;
	const size_t rpg = (nrows + G::NG - 1) / G::NG;
	const size_t r0 = (size_t)g * rpg;
	const size_t r1 = r0 + rpg < nrows ? r0 + rpg : nrows;
	const auto clamp_row = [&](size_t x) { return x < r0 ? r0 : (x > r1 ? r1 : x); };
	const size_t rf = clamp_row(len / CK), rz = clamp_row((len + CK - 1) / CK);  // rows < rz < nb + 1
	for(unsigned j = 0; j < G::DW; j++) h[j] = 0;
	uint8_t* p = slot + (size_t)t * G::P;
	size_t r = r0;
	for(; r + U <= rf; r += U) {
		Piece<G::DW> v[U];
#pragma unroll
		for(unsigned u = 0; u < U; u++) v[u] = load_piece<G::DW>(p + (r + u) * CK);
#pragma unroll
		for(unsigned u = 0; u < U; u++)
			for(unsigned j = 0; j < G::DW; j++) h[j] = mul2x2(h[j]) ^ v[u].d[j];
	}
	for(; r < rz; r++) {
		Piece<G::DW> v = load_piece<G::DW>(p + r * CK);
		const size_t off = r * CK + (size_t)t * G::P;
		if(off + G::P > len) {
			for(unsigned j = 0; j < G::DW; j++) v.d[j] &= keep_mask(off + 4 * j, len);
			store_piece<G::DW>(p + r * CK, v);
		}
		for(unsigned j = 0; j < G::DW; j++) h[j] = mul2x2(h[j]) ^ v.d[j];
	}
	Piece<G::DW> zero;
	for(unsigned j = 0; j < G::DW; j++) zero.d[j] = 0;
	for(; r < r1; r++) {
		store_piece<G::DW>(p + r * CK, zero);
		if(r < nb)
			for(unsigned j = 0; j < G::DW; j++) h[j] = mul2x2(h[j]);
	}
	run_end = r1 < nb ? r1 : nb;
	if(r0 >= nb) run_end = (size_t)-1;  // run contributes nothing
}

// Scale this thread's partial to its final position and reduce across groups.
// Returns (in threads < CK/2) the checksum lane `threadIdx.x`.
template <unsigned CK, unsigned T>
__device__ uint32_t ck_reduce(const uint32_t* h, size_t nb, size_t run_end, const uint16_t* __restrict__ lg,
                              const uint16_t* __restrict__ ex) {
	using G = CkGeom<CK, T>;
	__shared__ uint16_t red[G::NG][CK / 2];
	const unsigned t = threadIdx.x % G::TPR;
	const unsigned g = threadIdx.x / G::TPR;
	const uint32_t e = run_end == (size_t)-1 ? 0 : (uint32_t)((nb - run_end) % kOrder);
	for(unsigned j = 0; j < G::DW; j++) {
		uint32_t lo = 0, hi = 0;
		if(run_end != (size_t)-1) {
			lo = mul_pow2(h[j] & 0xFFFFu, e, lg, ex);
			hi = mul_pow2(h[j] >> 16, e, lg, ex);
		}
		red[g][t * (G::P / 2) + 2 * j] = (uint16_t)lo;
		red[g][t * (G::P / 2) + 2 * j + 1] = (uint16_t)hi;
	}
	__syncthreads();
	uint32_t acc = 0;
	if(threadIdx.x < CK / 2) {
		for(unsigned k = 0; k < G::NG; k++) acc ^= red[k][threadIdx.x];
	}
	__syncthreads();  // red is reused by the next slot
	return acc;
}

// Persistent over slots: workgroup b prepares slots b, b + gridDim.x, ...
template <unsigned CK>
__global__ __launch_bounds__(kPrepThreads) void k_prep(uint8_t* slots, const uint8_t* src_slots, size_t slot_bytes,
                                                       unsigned count, const uint32_t* __restrict__ lens,
                                                       const uint32_t* __restrict__ slot_index, size_t slice_size,
                                                       size_t ck_off, int to_log, const uint16_t* __restrict__ lg,
                                                       const uint16_t* __restrict__ ex) {
	__shared__ uint16_t s_log[65536];
	if(to_log) {
		const uint4* a = reinterpret_cast<const uint4*>(lg);
		uint4* b = reinterpret_cast<uint4*>(s_log);
		for(unsigned k = threadIdx.x; k < 65536 * 2 / 16; k += kPrepThreads) b[k] = a[k];
		__syncthreads();
	}
	const size_t nrows = slot_bytes / CK;
	const size_t nb = (slice_size + CK - 1) / CK;
	for(unsigned k = blockIdx.x; k < count; k += gridDim.x) {
		const size_t sidx = slot_index ? slot_index[k] : k;
		uint8_t* slot = slots + sidx * slot_bytes;
		const uint8_t* src = src_slots ? src_slots + sidx * slot_bytes : slot;
		size_t len = lens ? lens[k] : slice_size;
		if(len > slice_size) len = slice_size;
		uint32_t h[CkGeom<CK, kPrepThreads>::DW];
		size_t run_end;
		if(to_log)
			ck_rows<CK, kPrepThreads, 2>(src, slot, nrows, nb, len, h, s_log, run_end);
		else
			ck_rows<CK, kPrepThreads, 1>(src, slot, nrows, nb, len, h, s_log, run_end);
		// row rewrites complete before the checksum block is stored over them: all
		// of them come from this workgroup (one CU, one L2), so a workgroup barrier
		// orders them; a device-scope fence would write back the whole L2 per slot
		__syncthreads();
		uint32_t lane = ck_reduce<CK, kPrepThreads>(h, nb, run_end, lg, ex);
		if(threadIdx.x < CK / 2) {
			uint16_t v = (uint16_t)lane;
			if(to_log) v = s_log[v];
			*reinterpret_cast<uint16_t*>(slot + ck_off + 2 * threadIdx.x) = v;
		}
	}
}

// Data-form preparation (no log table) of one slot: zero padding past `len`,
// gf16_cksum into the checksum block.
// INPLACE: src == slot, known at compile time.
template <unsigned CK, bool INPLACE>
__device__ __forceinline__ void prep_data_slot(uint8_t* slot, const uint8_t* src, size_t slot_bytes, size_t len,
                                               size_t slice_size, size_t ck_off, const uint16_t* __restrict__ lg,
                                               const uint16_t* __restrict__ ex) {
	const size_t nrows = slot_bytes / CK;
	const size_t nb = (slice_size + CK - 1) / CK;
	if(len > slice_size) len = slice_size;
	uint32_t h[CkGeom<CK, kCkThreads>::DW];
	size_t run_end;
	if constexpr(INPLACE)
		ck_rows_inplace<CK, kCkThreads>(slot, nrows, nb, len, h, run_end);
	else
		ck_rows<CK, kCkThreads, 1>(src, slot, nrows, nb, len, h, nullptr, run_end);
	// row rewrites complete before the checksum block is stored over them: all
	// of them come from this workgroup (one CU, one L2), so a workgroup barrier
	// orders them; a device-scope fence would write back the whole L2 per slot
	__syncthreads();
	uint32_t lane = ck_reduce<CK, kCkThreads>(h, nb, run_end, lg, ex);
	if(threadIdx.x < CK / 2) *reinterpret_cast<uint16_t*>(slot + ck_off + 2 * threadIdx.x) = (uint16_t)lane;
}

// 256-thread workgroups, several per CU, persistent over slots — the form the
// dense kernels read.
template <unsigned CK>
__global__ __launch_bounds__(kCkThreads) void k_prep_data(uint8_t* slots, const uint8_t* src_slots, size_t slot_bytes,
                                                          unsigned count, const uint32_t* __restrict__ lens,
                                                          const uint32_t* __restrict__ slot_index, size_t slice_size,
                                                          size_t ck_off, const uint16_t* __restrict__ lg,
                                                          const uint16_t* __restrict__ ex) {
	for(unsigned k = blockIdx.x; k < count; k += gridDim.x) {
		const size_t sidx = slot_index ? slot_index[k] : k;
		uint8_t* slot = slots + sidx * slot_bytes;
		const uint8_t* src = src_slots ? src_slots + sidx * slot_bytes : slot;
		prep_data_slot<CK, false>(slot, src, slot_bytes, lens ? lens[k] : slice_size, slice_size, ck_off, lg, ex);
	}
}

// Many jobs: workgroup b prepares slots b, b + G, ... of the jobs' slots
// numbered in table order (entry j's slot k is number table[j].first + k), so
// a small grid can take its share of the CUs and leave the rest.
template <unsigned CK>
__global__ __launch_bounds__(kCkThreads) void k_prep_data_jobs(const CkJob* __restrict__ table, unsigned n_jobs,
                                                               unsigned total, const uint16_t* __restrict__ lg,
                                                               const uint16_t* __restrict__ ex) {
	unsigned j = 0;
	CkJob J = wave_uniform(table[0]);
	for(unsigned n = blockIdx.x; n < total; n += gridDim.x) {
		while(n >= J.first + J.count && j + 1 < n_jobs) J = wave_uniform(table[++j]);
		const unsigned k = n - J.first;
		uint8_t* slot = J.slots + (size_t)k * J.slot_bytes;
		prep_data_slot<CK, true>(slot, slot, J.slot_bytes, J.lens ? J.lens[k] : J.slice_size, J.slice_size, J.ck_off, lg,
		                         ex);
	}
}

// One workgroup per slot; with `table`, slot blockIdx.x of job blockIdx.y
// (workgroups past the job's slot count do nothing).
template <unsigned CK>
__global__ __launch_bounds__(kCkThreads) void k_verify(const uint8_t* __restrict__ slots, size_t slot_bytes,
                                                       size_t slice_size, size_t ck_off, int* __restrict__ valid,
                                                       const uint16_t* __restrict__ lg,
                                                       const uint16_t* __restrict__ ex,
                                                       const CkJob* __restrict__ table) {
	const unsigned k = blockIdx.x;
	if(table) {
		const CkJob J = wave_uniform(table[blockIdx.y]);
		if(k >= J.count) return;
		slots = J.slots;
		slot_bytes = J.slot_bytes;
		slice_size = J.slice_size;
		ck_off = J.ck_off;
		valid = J.valid;
	}
	uint8_t* slot = const_cast<uint8_t*>(slots) + (size_t)k * slot_bytes;
	const size_t nb = (slice_size + CK - 1) / CK;
	uint32_t h[CkGeom<CK, kCkThreads>::DW];
	size_t run_end;
	ck_rows<CK, kCkThreads, 0>(slot, slot, nb, nb, slice_size, h, lg, run_end);
	uint32_t lane = ck_reduce<CK, kCkThreads>(h, nb, run_end, lg, ex);
	__shared__ int bad;
	if(threadIdx.x == 0) bad = 0;
	__syncthreads();
	if(threadIdx.x < CK / 2) {
		uint16_t stored = *reinterpret_cast<const uint16_t*>(slot + ck_off + 2 * threadIdx.x);
		if(stored != (uint16_t)lane) atomicOr(&bad, 1);
	}
	__syncthreads();
	if(threadIdx.x == 0) valid[k] = bad ? 0 : 1;
}

hipError_t launch_prep(uint8_t* slots, const uint8_t* src, size_t slot_bytes, unsigned count, const uint32_t* lens,
                       const uint32_t* slot_index, size_t slice_size, size_t ck_off, unsigned ck, bool to_log,
                       const DeviceTables& t, hipStream_t s) {
	if(!count) return hipSuccess;
	if(slot_bytes % kSlotAlign || ck_off + ck > slot_bytes || slice_size > ck_off || !valid_ck(ck))
		return hipErrorInvalidValue;
	const unsigned cus = (unsigned)(t.cus > 0 ? t.cus : 256);
	if(!to_log) {
		const unsigned blocks = count < cus * 8 ? count : cus * 8;
		dim3 grid(blocks), block(kCkThreads);
		switch(ck) {
			case 8: hipLaunchKernelGGL(k_prep_data<8>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, t.log, t.exp); break;
			case 16: hipLaunchKernelGGL(k_prep_data<16>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, t.log, t.exp); break;
			case 32: hipLaunchKernelGGL(k_prep_data<32>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, t.log, t.exp); break;
			default: hipLaunchKernelGGL(k_prep_data<64>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, t.log, t.exp); break;
		}
		return hipGetLastError();
	}
	const unsigned blocks = count < cus ? count : cus;
	dim3 grid(blocks), block(kPrepThreads);
	const int tl = 1;
	switch(ck) {
		case 8: hipLaunchKernelGGL(k_prep<8>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, tl, t.log, t.exp); break;
		case 16: hipLaunchKernelGGL(k_prep<16>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, tl, t.log, t.exp); break;
		case 32: hipLaunchKernelGGL(k_prep<32>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, tl, t.log, t.exp); break;
		default: hipLaunchKernelGGL(k_prep<64>, grid, block, 0, s, slots, src, slot_bytes, count, lens, slot_index, slice_size, ck_off, tl, t.log, t.exp); break;
	}
	return hipGetLastError();
}

hipError_t launch_verify(const uint8_t* slots, size_t slot_bytes, unsigned count, size_t slice_size, size_t ck_off,
                         unsigned ck, int* valid, const DeviceTables& t, hipStream_t s) {
	if(!count) return hipSuccess;
	if(slot_bytes % kSlotAlign || ck_off + ck > slot_bytes || slice_size > ck_off || !valid_ck(ck))
		return hipErrorInvalidValue;
	dim3 grid(count), block(kCkThreads);
	switch(ck) {
		case 8: hipLaunchKernelGGL(k_verify<8>, grid, block, 0, s, slots, slot_bytes, slice_size, ck_off, valid, t.log, t.exp, (const CkJob*)nullptr); break;
		case 16: hipLaunchKernelGGL(k_verify<16>, grid, block, 0, s, slots, slot_bytes, slice_size, ck_off, valid, t.log, t.exp, (const CkJob*)nullptr); break;
		case 32: hipLaunchKernelGGL(k_verify<32>, grid, block, 0, s, slots, slot_bytes, slice_size, ck_off, valid, t.log, t.exp, (const CkJob*)nullptr); break;
		default: hipLaunchKernelGGL(k_verify<64>, grid, block, 0, s, slots, slot_bytes, slice_size, ck_off, valid, t.log, t.exp, (const CkJob*)nullptr); break;
	}
	return hipGetLastError();
}

bool ck_job_valid(const CkJob& j, unsigned ck) {
	return j.slots && j.slot_bytes % kSlotAlign == 0 && j.ck_off + ck <= j.slot_bytes && j.slice_size <= j.ck_off &&
	       valid_ck(ck);
}

// Many jobs in one launch.  The entries were checked with ck_job_valid and
// numbered (first) when the table was built; `total` slots in all.
// Preparation: persistent, at most per_cu workgroups per CU, so that other
// work can share the CUs.
hipError_t launch_prep_data_jobs(const CkJob* table, unsigned n_jobs, unsigned total, unsigned ck, unsigned per_cu,
                                 const DeviceTables& t, hipStream_t s) {
	if(!n_jobs || !total) return hipSuccess;
	if(!table || !valid_ck(ck) || !per_cu) return hipErrorInvalidValue;
	const unsigned cus = (unsigned)(t.cus > 0 ? t.cus : 256);
	dim3 grid(total < cus * per_cu ? total : cus * per_cu), block(kCkThreads);
	switch(ck) {
		case 8: hipLaunchKernelGGL(k_prep_data_jobs<8>, grid, block, 0, s, table, n_jobs, total, t.log, t.exp); break;
		case 16: hipLaunchKernelGGL(k_prep_data_jobs<16>, grid, block, 0, s, table, n_jobs, total, t.log, t.exp); break;
		case 32: hipLaunchKernelGGL(k_prep_data_jobs<32>, grid, block, 0, s, table, n_jobs, total, t.log, t.exp); break;
		default: hipLaunchKernelGGL(k_prep_data_jobs<64>, grid, block, 0, s, table, n_jobs, total, t.log, t.exp); break;
	}
	return hipGetLastError();
}

// Verification: grid (largest slot count, jobs), one workgroup per slot.
hipError_t launch_verify_jobs(const CkJob* table, unsigned n_jobs, unsigned max_count, unsigned ck,
                              const DeviceTables& t, hipStream_t s) {
	if(!n_jobs || !max_count) return hipSuccess;
	if(!table || !valid_ck(ck) || n_jobs > 65535) return hipErrorInvalidValue;
	dim3 grid(max_count, n_jobs), block(kCkThreads);
	switch(ck) {
		case 8: hipLaunchKernelGGL(k_verify<8>, grid, block, 0, s, nullptr, 0, 0, 0, nullptr, t.log, t.exp, table); break;
		case 16: hipLaunchKernelGGL(k_verify<16>, grid, block, 0, s, nullptr, 0, 0, 0, nullptr, t.log, t.exp, table); break;
		case 32: hipLaunchKernelGGL(k_verify<32>, grid, block, 0, s, nullptr, 0, 0, 0, nullptr, t.log, t.exp, table); break;
		default: hipLaunchKernelGGL(k_verify<64>, grid, block, 0, s, nullptr, 0, 0, 0, nullptr, t.log, t.exp, table); break;
	}
	return hipGetLastError();
}

// ------------------------------------------------------------------ plain mul-add

// Persistent, one workgroup per CU, over the units (item, input), item =
// (column tile, output group) numbered column-major, split evenly between the
// workgroups (a "stream-K" split): a workgroup takes a run of whole items
// and, at its ends, parts of the items it shares with its neighbours, whose
// sums it XORs into the output with atomics.  With one workgroup per item
// (845 items for 1 MiB slices and 200 outputs in groups of 16) the last of
// ~3.3 rounds of workgroups ran a quarter of the CUs: 77 % of the LDS
// ceiling against ~95 % per workgroup.  The shares go to the workgroups in
// XCD-major order (blockIdx round-robins over the 8 XCDs), so that the
// items of one column tile, which read the same inputs, meet in one L2.
template <int G, bool TABLE>
__global__ __launch_bounds__(kPlainThreads) void k_muladd_plain(
    const uint16_t* __restrict__ in, size_t slot_words, unsigned n_in, const uint16_t* __restrict__ in_logs,
    const uint16_t* __restrict__ exps, const uint32_t* __restrict__ coef, unsigned coef_stride,
    uint16_t* __restrict__ out, unsigned n_out, int accumulate, const uint16_t* __restrict__ exp_tab,
    unsigned col_tiles, unsigned groups) {
	__shared__ uint16_t s_exp[65536];
	__shared__ uint16_t s_lin[kPlainThreads];
	load_exp_table(s_exp, exp_tab);
	__syncthreads();
	const unsigned P = gridDim.x;
	const unsigned sh = (P & 7u) == 0 ? (blockIdx.x & 7u) * (P >> 3) + (blockIdx.x >> 3) : blockIdx.x;
	const uint64_t total = (uint64_t)col_tiles * groups * n_in;
	uint64_t u = total * sh / P;
	const uint64_t u1 = total * (sh + 1) / P;
	while(u < u1) {
		const uint64_t item = u / n_in;
		const unsigned i0 = (unsigned)(u - item * n_in);
		const uint64_t left = u1 - u;
		const unsigned i1 = left < (uint64_t)(n_in - i0) ? i0 + (unsigned)left : n_in;
		const unsigned col = (unsigned)(item / groups);
		const unsigned grp = (unsigned)(item - (uint64_t)col * groups);
		plain_tile<G, TABLE>(s_exp, s_lin, in, slot_words, n_in, in_logs, exps, coef, coef_stride, out, n_out, accumulate,
		                     col, grp, i0, i1, i0 != 0 || i1 != n_in);
		u += i1 - i0;
	}
}

// One workgroup per item (the form concurrent launches share the CUs in, as
// the serving engine runs them): blocks of one column tile run on the same
// XCD (blockIdx round-robins over 8 XCDs).
template <int G, bool TABLE>
__global__ __launch_bounds__(kPlainThreads) void k_muladd_plain_items(
    const uint16_t* __restrict__ in, size_t slot_words, unsigned n_in, const uint16_t* __restrict__ in_logs,
    const uint16_t* __restrict__ exps, const uint32_t* __restrict__ coef, unsigned coef_stride,
    uint16_t* __restrict__ out, unsigned n_out, int accumulate, const uint16_t* __restrict__ exp_tab,
    unsigned col_tiles, unsigned groups) {
	__shared__ uint16_t s_exp[65536];
	__shared__ uint16_t s_lin[kPlainThreads];
	load_exp_table(s_exp, exp_tab);
	__syncthreads();
	const unsigned total = col_tiles * groups;
	unsigned L = blockIdx.x;
	if((total & 7u) == 0) L = (blockIdx.x & 7u) * (total >> 3) + (blockIdx.x >> 3);
	const unsigned col = L / groups;
	const unsigned grp = L - col * groups;
	plain_tile<G, TABLE>(s_exp, s_lin, in, slot_words, n_in, in_logs, exps, coef, coef_stride, out, n_out, accumulate, col,
	                     grp);
}

// Before an overwriting launch: the output tiles of the items that the split
// shares between workgroups (those a share starts inside), zeroed for the
// atomic XORs.  Block (s, r): the start of share s, row r of its group.
template <int G>
__global__ __launch_bounds__(kPlainThreads) void k_plain_zero_split(uint16_t* __restrict__ out, size_t slot_words,
                                                                    unsigned n_in, unsigned n_out, unsigned col_tiles,
                                                                    unsigned groups) {
	const uint64_t total = (uint64_t)col_tiles * groups * n_in;
	const uint64_t u = total * blockIdx.x / gridDim.x;
	if(u >= total || u % n_in == 0) return;
	const uint64_t item = u / n_in;
	const unsigned col = (unsigned)(item / groups);
	const unsigned row = (unsigned)(item - (uint64_t)col * groups) * G + blockIdx.y;
	const size_t w0 = (size_t)col * (kPlainThreads * kPlainWordsPerLane) + (size_t)threadIdx.x * kPlainWordsPerLane;
	if(row < n_out && w0 < slot_words)
		*reinterpret_cast<uint4*>(out + (size_t)row * slot_words + w0) = make_uint4(0, 0, 0, 0);
}

// rows [0, n_out) of the output slots zeroed (an overwriting launch without inputs)
__global__ void k_plain_zero_rows(uint16_t* __restrict__ out, size_t slot_words, unsigned n_out) {
	const size_t n = (size_t)n_out * slot_words / 8;
	for(size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
		reinterpret_cast<uint4*>(out)[k] = make_uint4(0, 0, 0, 0);
}

template <int G>
static hipError_t launch_plain_g(const uint16_t* in, size_t slot_words, unsigned n_in, const uint16_t* in_logs,
                                 const uint16_t* exps, const uint32_t* coef, unsigned coef_stride, uint16_t* out,
                                 unsigned n_out, bool accumulate, bool split, const DeviceTables& t, hipStream_t s) {
	const unsigned tile_words = kPlainThreads * kPlainWordsPerLane;
	const unsigned col_tiles = (unsigned)((slot_words + tile_words - 1) / tile_words);
	const unsigned groups = (n_out + G - 1) / G;
	if(!split) {
		dim3 grid(col_tiles * groups), block(kPlainThreads);
		if(coef)
			hipLaunchKernelGGL((k_muladd_plain_items<G, true>), grid, block, 0, s, in, slot_words, n_in, in_logs, exps,
			                   coef, coef_stride, out, n_out, accumulate ? 1 : 0, t.exp, col_tiles, groups);
		else
			hipLaunchKernelGGL((k_muladd_plain_items<G, false>), grid, block, 0, s, in, slot_words, n_in, in_logs, exps,
			                   coef, coef_stride, out, n_out, accumulate ? 1 : 0, t.exp, col_tiles, groups);
		return hipGetLastError();
	}
	if(!n_in) {
		if(!accumulate) hipLaunchKernelGGL(k_plain_zero_rows, dim3(1024), dim3(256), 0, s, out, slot_words, n_out);
		return hipGetLastError();
	}
	const unsigned P = (unsigned)(t.cus > 0 ? t.cus : 256);  // one workgroup per CU
	if(!accumulate)
		hipLaunchKernelGGL((k_plain_zero_split<G>), dim3(P, G), dim3(kPlainThreads), 0, s, out, slot_words, n_in, n_out,
		                   col_tiles, groups);
	dim3 grid(P), block(kPlainThreads);
	if(coef)
		hipLaunchKernelGGL((k_muladd_plain<G, true>), grid, block, 0, s, in, slot_words, n_in, in_logs, exps, coef,
		                   coef_stride, out, n_out, accumulate ? 1 : 0, t.exp, col_tiles, groups);
	else
		hipLaunchKernelGGL((k_muladd_plain<G, false>), grid, block, 0, s, in, slot_words, n_in, in_logs, exps, coef,
		                   coef_stride, out, n_out, accumulate ? 1 : 0, t.exp, col_tiles, groups);
	return hipGetLastError();
}

hipError_t launch_muladd_plain(const uint16_t* in, size_t slot_words, unsigned n_in, const uint16_t* in_logs,
                               const uint16_t* exps, const uint32_t* coef, unsigned coef_stride, uint16_t* out,
                               unsigned n_out, bool accumulate, unsigned group, const DeviceTables& t,
                               hipStream_t s, bool split) {
	if(!n_out || !slot_words) return hipSuccess;
	if(slot_words % (kSlotAlign / 2)) return hipErrorInvalidValue;
	if(!coef && (!in_logs || !exps) && n_in) return hipErrorInvalidValue;
	if(group == 0) {
		group = 16;
		while(group > 1 && group / 2 >= n_out) group /= 2;
	}
	// an explicit table must cover every output group it is read for
	if(coef) {
		const unsigned padded = (n_out + group - 1) / group * group;
		if(coef_stride < padded) return hipErrorInvalidValue;
	}
	switch(group) {
		case 1: return launch_plain_g<1>(in, slot_words, n_in, in_logs, exps, coef, coef_stride, out, n_out, accumulate, split, t, s);
		case 2: return launch_plain_g<2>(in, slot_words, n_in, in_logs, exps, coef, coef_stride, out, n_out, accumulate, split, t, s);
		case 4: return launch_plain_g<4>(in, slot_words, n_in, in_logs, exps, coef, coef_stride, out, n_out, accumulate, split, t, s);
		case 8: return launch_plain_g<8>(in, slot_words, n_in, in_logs, exps, coef, coef_stride, out, n_out, accumulate, split, t, s);
		case 16: return launch_plain_g<16>(in, slot_words, n_in, in_logs, exps, coef, coef_stride, out, n_out, accumulate, split, t, s);
		default: return hipErrorInvalidValue;
	}
}

// ------------------------------------------------------------------ misc

// Plain streaming copy (16-byte vectors): the reference point for HBM traffic counters.
__global__ void k_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n) {
	for(size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
		dst[k] = src[k];
}
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
	const size_t n = bytes / 16;
	if(!n) return hipSuccess;
	size_t blocks = (n + 255) / 256;
	if(blocks > 8192) blocks = 8192;
	hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, n);
	return hipGetLastError();
}

// Ceiling of random 16-bit lookups into the 128 KiB exp table in LDS, the
// plain form's inner operation: one 1024-thread workgroup per CU holding the
// table, as k_muladd_plain does; each lane gathers from 16 random byte
// offsets, XORed each round by a wave-uniform random even offset (so every
// round is a fresh random set of addresses with the statistics of random
// 16-bit indices), one VALU XOR to form each address and one to keep the
// value.  `sink` (one word per thread) keeps the results live.
__global__ __launch_bounds__(kPlainThreads) void k_lds_gather_ceiling(const uint16_t* __restrict__ exp_tab,
                                                                      unsigned rounds, uint32_t seed,
                                                                      uint32_t* __restrict__ sink) {
	__shared__ uint16_t s_exp[65536];
	load_exp_table(s_exp, exp_tab);
	__syncthreads();
	uint32_t x = seed ^ (blockIdx.x * 0x9E3779B1u) ^ (threadIdx.x * 0x85EBCA77u);
	uint32_t base[16];
#pragma unroll
	for(int k = 0; k < 16; k++) {
		x ^= x << 13;
		x ^= x >> 17;
		x ^= x << 5;
		base[k] = x & 0x1FFFEu;
	}
	uint32_t acc = 0;
	for(unsigned i = 0; i < rounds; i++) {
		const uint32_t u = ((i + 1) * 0x9E3779B9u ^ seed) & 0x1FFFEu;
#pragma unroll
		for(int k = 0; k < 16; k++)
			acc ^= *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(s_exp) + (base[k] ^ u));
	}
	sink[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}
hipError_t launch_lds_gather_ceiling(const uint16_t* exp_tab, unsigned blocks, unsigned rounds, uint32_t seed,
                                     uint32_t* sink, hipStream_t s) {
	hipLaunchKernelGGL(k_lds_gather_ceiling, dim3(blocks), dim3(kPlainThreads), 0, s, exp_tab, rounds, seed, sink);
	return hipGetLastError();
}

__global__ void k_xor_byte(uint8_t* p, uint8_t mask) {
	if(threadIdx.x == 0 && blockIdx.x == 0) *p ^= mask;
}
hipError_t launch_xor_byte(uint8_t* p, uint8_t mask, hipStream_t s) {
	hipLaunchKernelGGL(k_xor_byte, dim3(1), dim3(64), 0, s, p, mask);
	return hipGetLastError();
}

}  // namespace pph
