// A blocking task queue for the library's worker threads: any thread pushes;
// workers pop in push order, blocking while it is empty; a worker reports
// each popped task finished, so that callers can wait until every task pushed
// so far has run; close() lets the workers drain what is queued and stop.
// Host C++ only (no HIP), so it is also built and run under ThreadSanitizer.
#pragma once
#include <condition_variable>
#include <deque>
#include <mutex>

namespace pph {

template<class T>
class TaskQueue {
	std::mutex mu_;
	std::condition_variable cv_;
	std::deque<T> q_;
	unsigned running_ = 0;  // popped, not yet reported finished
	bool closed_ = false;

public:
	// false once the queue is closed (the task is not queued)
	bool push(T v) {
		{
			std::lock_guard<std::mutex> lk(mu_);
			if(closed_) return false;
			q_.push_back(std::move(v));
		}
		cv_.notify_all();
		return true;
	}
	// Blocks for the next task; false when the queue is closed and empty.
	// A true return must be followed by finished() once the task has run.
	bool pop(T* out) {
		std::unique_lock<std::mutex> lk(mu_);
		cv_.wait(lk, [&] { return closed_ || !q_.empty(); });
		if(q_.empty()) return false;
		*out = std::move(q_.front());
		q_.pop_front();
		running_++;
		return true;
	}
	void finished() {
		{
			std::lock_guard<std::mutex> lk(mu_);
			running_--;
		}
		cv_.notify_all();
	}
	// Until every task pushed before the call has run.
	void wait_idle() {
		std::unique_lock<std::mutex> lk(mu_);
		cv_.wait(lk, [&] { return q_.empty() && running_ == 0; });
	}
	void close() {
		{
			std::lock_guard<std::mutex> lk(mu_);
			closed_ = true;
		}
		cv_.notify_all();
	}
	bool closed() {
		std::lock_guard<std::mutex> lk(mu_);
		return closed_;
	}
};

}  // namespace pph
