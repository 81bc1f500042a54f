// Parallel host copy: a copy of many bytes is split into parts that a few
// helper threads and the calling thread take from a shared list; the call
// returns once every part is done.  Staging copies into pinned memory run at
// one core's copy rate otherwise, which is below the PCIe rate they feed.
#pragma once
#include "pph_cksum_host.h"

#include <emmintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace pph {

// Copy with non-temporal stores.  The destination is staging memory that the
// DMA engine reads next, never this core: its lines are written without first
// being read for ownership, and the copy does not evict the caller's data from
// the caches.  The fence makes the stores visible before the copy is reported
// done (they are not ordered by ordinary stores).
inline void copy_stream(void* dst_, const void* src_, size_t n) {
	uint8_t* dst = static_cast<uint8_t*>(dst_);
	const uint8_t* src = static_cast<const uint8_t*>(src_);
	const size_t head = std::min<size_t>(n, (16 - ((uintptr_t)dst & 15)) & 15);
	memcpy(dst, src, head);
	dst += head;
	src += head;
	n -= head;
	size_t i = 0;
	for(; i + 64 <= n; i += 64) {
		const __m128i a = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + i));
		const __m128i b = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + i + 16));
		const __m128i c = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + i + 32));
		const __m128i d = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + i + 48));
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + i), a);
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + i + 16), b);
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + i + 32), c);
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + i + 48), d);
	}
	memcpy(dst + i, src + i, n - i);
	_mm_sfence();
}


class CopyPool {
	struct Part {
		uint8_t* dst;
		const uint8_t* src;
		size_t n;
		CkAcc* acc;   // with a checksum: the part's Horner sum over its blocks (from zero)
		unsigned ck;
	};
	std::mutex mu_;
	std::condition_variable cv_;
	std::vector<Part> parts_;      // the current copy's parts (mu_)
	size_t next_ = 0;              // next part to take (mu_)
	std::atomic<size_t> left_{0};  // parts not yet copied
	std::atomic<unsigned> gen_{0};  // bumped each time parts are posted
	std::atomic<bool> stop_{false};
	static constexpr std::chrono::microseconds kSpin{200};
	std::vector<std::thread> threads_;

	bool take(Part* p) {  // with mu_ held
		if(next_ >= parts_.size()) return false;
		*p = parts_[next_++];
		return true;
	}
	static void run(const Part& p) {
		if(p.acc)
			ck_copy_accumulate(*p.acc, p.dst, p.src, p.n, p.ck);
		else
			copy_stream(p.dst, p.src, p.n);
	}
	// the part's stores are fenced (or ordinary): the release publishes them to
	// the caller's acquire
	void finish() { left_.fetch_sub(1, std::memory_order_release); }
	void worker() {
		unsigned seen = 0;
		std::unique_lock<std::mutex> lk(mu_);
		for(;;) {
			if(!stop_ && next_ >= parts_.size()) {
				// Idle: watch for the next parts a while before sleeping.  Adds come
				// back to back, and a wake-up from sleep takes longer than a part.
				lk.unlock();
				const auto t0 = std::chrono::steady_clock::now();
				for(unsigned k = 1; gen_.load(std::memory_order_acquire) == seen && !stop_; k++) {
					_mm_pause();
					if(k % 64 == 0 && std::chrono::steady_clock::now() - t0 > kSpin) break;
				}
				lk.lock();
				cv_.wait(lk, [&] { return stop_ || next_ < parts_.size(); });
			}
			if(stop_) return;
			seen = gen_.load(std::memory_order_relaxed);
			Part p;
			while(take(&p)) {
				lk.unlock();
				run(p);
				finish();
				lk.lock();
			}
		}
	}

public:
	static constexpr size_t kPart = 256 << 10;  // below this a copy stays on the caller's thread

	explicit CopyPool(unsigned helpers) {
		for(unsigned k = 0; k < helpers; k++) threads_.emplace_back(&CopyPool::worker, this);
	}
	~CopyPool() {
		{
			std::lock_guard<std::mutex> lk(mu_);
			stop_ = true;
		}
		cv_.notify_all();
		for(auto& t : threads_) t.join();
	}
	CopyPool(const CopyPool&) = delete;
	CopyPool& operator=(const CopyPool&) = delete;

	// One caller at a time (the backend's controlling thread).
	void copy(void* dst, const void* src, size_t n) { split(dst, src, n, 0); }

	// Copy n bytes and return in ck_out (ck bytes) their gf16_cksum Horner sum
	// H = sum_k B_k * 2^(m-1-k) over the m = ceil(n/ck) blocks of ck bytes, the
	// last one zero-padded (ck: 8, 16, 32 or 64).  Parts start on block
	// boundaries; each sums its own blocks from zero, and the parts' sums are
	// combined by H = sum_j H_j * 2^(blocks after part j), with the field's
	// exp/log tables.
	void copy_cksum(void* dst, const void* src, size_t n, unsigned ck, uint8_t* ck_out, const uint16_t* exp_tab,
	                const uint16_t* log_tab) {
		split(dst, src, n, ck);
		alignas(16) uint16_t total[32] = {0};
		const size_t m = (n + ck - 1) / ck;
		size_t blocks_done = 0;
		for(size_t j = 0; j < plens_.size(); j++) {
			blocks_done += (plens_[j] + ck - 1) / ck;
			alignas(16) uint16_t h[32];
			ck_store(accs_[j], reinterpret_cast<uint8_t*>(h), ck);
			ck_mul_exp_xor(total, h, ck / 2, (uint32_t)((m - blocks_done) % 65535), exp_tab, log_tab);
		}
		memcpy(ck_out, total, ck);
	}

private:
	// the current copy's part lengths and, for copy_cksum, their sums (the caller's thread)
	std::vector<size_t> plens_;
	std::vector<CkAcc> accs_;

	// Post the parts of one copy (checksummed when ck != 0), take parts on the
	// caller's thread too, return once all are done.  Parts are multiples of 64
	// bytes but the last, so they start on checksum block boundaries.
	void split(void* dst, const void* src, size_t n, unsigned ck) {
		const size_t ways = std::max<size_t>(1, std::min<size_t>(threads_.size() + 1, n / kPart));
		const size_t step = ways < 2 ? std::max<size_t>(n, 1) : (n / ways + 63) & ~size_t(63);
		plens_.clear();
		for(size_t off = 0; off < n; off += step) plens_.push_back(std::min(step, n - off));
		if(ck) {
			accs_.resize(plens_.size());
			for(auto& a : accs_) ck_zero(a);
		}
		if(plens_.size() < 2) {
			if(n) run(Part{(uint8_t*)dst, (const uint8_t*)src, n, ck ? &accs_[0] : nullptr, ck});
			return;
		}
		{
			std::lock_guard<std::mutex> lk(mu_);
			parts_.clear();
			next_ = 0;
			for(size_t j = 0; j < plens_.size(); j++)
				parts_.push_back({(uint8_t*)dst + j * step, (const uint8_t*)src + j * step, plens_[j],
				                  ck ? &accs_[j] : nullptr, ck});
			left_.store(parts_.size(), std::memory_order_relaxed);
			gen_.fetch_add(1, std::memory_order_release);
		}
		cv_.notify_all();
		for(Part p;;) {
			{
				std::lock_guard<std::mutex> lk(mu_);
				if(!take(&p)) break;
			}
			run(p);
			finish();
		}
		// the helpers' last parts are in progress: at most one part's copy time
		while(left_.load(std::memory_order_acquire) != 0) _mm_pause();
	}
};

}  // namespace pph
