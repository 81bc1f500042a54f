// Parallel host copy: a copy of many bytes is split into parts that a few
// helper threads and the calling thread take from a shared list; the call
// returns once every part is done.  Staging copies into pinned memory run at
// one core's copy rate otherwise, which is below the PCIe rate they feed.
#pragma once
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
	// copy_stream has fenced the part's stores: the release publishes them to
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
				copy_stream(p.dst, p.src, p.n);
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
	void copy(void* dst, const void* src, size_t n) {
		const size_t ways = std::min<size_t>(threads_.size() + 1, n / kPart);
		if(ways < 2) {
			copy_stream(dst, src, n);
			return;
		}
		const size_t step = (n / ways + 63) & ~size_t(63);
		{
			std::lock_guard<std::mutex> lk(mu_);
			parts_.clear();
			next_ = 0;
			for(size_t off = 0; off < n; off += step)
				parts_.push_back({(uint8_t*)dst + off, (const uint8_t*)src + off, std::min(step, n - off)});
			left_.store(parts_.size(), std::memory_order_relaxed);
			gen_.fetch_add(1, std::memory_order_release);
		}
		cv_.notify_all();
		for(Part p;;) {
			{
				std::lock_guard<std::mutex> lk(mu_);
				if(!take(&p)) break;
			}
			copy_stream(p.dst, p.src, p.n);
			finish();
		}
		// the helpers' last parts are in progress: at most one part's copy time
		while(left_.load(std::memory_order_acquire) != 0) _mm_pause();
	}
};

}  // namespace pph
