// Host threading of the library under ThreadSanitizer (built by
// tests/test_tsan.py with -fsanitize=thread; no HIP):
//  - TaskQueue, the queue of the backend's completion worker and of the Node
//    binding's worker: several producers and consumers, wait_idle while
//    tasks run, close with tasks still queued (every pushed task runs once);
//  - CopyPool, the backend's staging and getOutput copies: three pools, each
//    driven by its own thread with its own helpers, plain and checksummed
//    copies checked against a single-threaded copy and a scalar checksum;
//  - the serving engine's host MD5 pool: threads taking files from a shared
//    queue by the highest response ratio while others submit, every file
//    hashed once; and the order itself on fixed queues.
// Prints "ok" when every check holds (ThreadSanitizer reports races itself).
#include "../../parpar_amd/csrc/pph_copy_pool.h"
#include "../../parpar_amd/csrc/pph_md5_order.h"
#include "../../parpar_amd/csrc/pph_task_queue.h"

#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <thread>
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

static void scalar_cksum(const uint8_t* src, size_t n, unsigned ck, uint16_t* out) {
	for(unsigned l = 0; l < ck / 2; l++) out[l] = 0;
	for(size_t pos = 0; pos < n; pos += ck)
		for(unsigned l = 0; l < ck / 2; l++) {
			uint32_t v = (uint32_t)out[l] << 1;
			if(v & 0x10000) v ^= 0x1100B;
			const size_t b = pos + 2 * l;
			uint16_t w = 0;
			if(b < n) w |= src[b];
			if(b + 1 < n) w |= (uint16_t)src[b + 1] << 8;
			out[l] = (uint16_t)(v ^ w);
		}
}

static int check_queue() {
	struct Item {
		unsigned producer, seq;
	};
	pph::TaskQueue<Item> q;
	const unsigned P = 4, C = 3, N = 5000;
	std::vector<std::atomic<unsigned>> seen(P * N);
	for(auto& s : seen) s = 0;
	std::atomic<unsigned> done{0};
	std::vector<std::thread> cons;
	for(unsigned c = 0; c < C; c++)
		cons.emplace_back([&] {
			Item it;
			while(q.pop(&it)) {
				seen[it.producer * N + it.seq]++;
				done++;
				q.finished();
			}
		});
	std::vector<std::thread> prod;
	for(unsigned p = 0; p < P; p++)
		prod.emplace_back([&, p] {
			for(unsigned k = 0; k < N; k++) {
				q.push(Item{p, k});
				if(k % 1000 == 999) q.wait_idle();  // every task pushed so far has run
			}
		});
	for(auto& t : prod) t.join();
	q.wait_idle();
	if(done != P * N) return printf("queue: %u of %u tasks ran before wait_idle returned\n", done.load(), P * N), 1;
	// tasks queued at close still run, then the consumers stop
	for(unsigned k = 0; k < 100; k++) q.push(Item{0, k});
	q.close();
	if(q.push(Item{0, 0})) return printf("queue: push after close\n"), 1;
	for(auto& t : cons) t.join();
	for(unsigned k = 0; k < P * N; k++)
		if(seen[k] != (k < 100 ? 2u : 1u)) return printf("queue: task %u ran %u times\n", k, seen[k].load()), 1;
	return 0;
}

static int check_pools() {
	std::atomic<int> bad{0};
	std::vector<std::thread> drivers;
	for(unsigned d = 0; d < 3; d++)
		drivers.emplace_back([&, d] {
			pph::CopyPool pool(2 + d);
			const size_t P = pph::CopyPool::kPart;
			const size_t sizes[] = {P - 1, 3 * P + 17, 7 * P + 5, 16 * P};
			for(int rep = 0; rep < 4; rep++)
				for(size_t n : sizes) {
					std::vector<uint8_t> src(n), dst(n + 16, 0xEE);
					for(size_t k = 0; k < n; k++) src[k] = (uint8_t)(k * 131 + n + rep + d);
					pool.copy(dst.data(), src.data(), n);
					if(memcmp(dst.data(), src.data(), n)) bad++;
					alignas(16) uint8_t got[64];
					uint16_t want[32];
					const unsigned ck = 8u << (rep % 4);
					pool.copy_cksum(dst.data() + 16 * (rep & 1), src.data(), n, ck, got, g_exp, g_log);
					scalar_cksum(src.data(), n, ck, want);
					if(memcmp(got, want, ck) || memcmp(dst.data() + 16 * (rep & 1), src.data(), n)) bad++;
				}
		});
	for(auto& t : drivers) t.join();
	if(bad) return printf("copy pools: %d mismatches\n", bad.load()), 1;
	return 0;
}

struct File {
	uint64_t size;
	double arrival;
	unsigned id;
};

static int check_md5_order() {
	// fixed queues: a short file that waited beats a long one that waited as long;
	// a long file that waited long enough beats a fresh short one; ties: the first
	std::vector<File> q = {{8ull << 30, 0.0, 0}, {1 << 20, 0.0, 1}, {1 << 20, 0.0, 2}};
	auto size = [](const File& f) { return f.size; };
	auto arr = [](const File& f) { return f.arrival; };
	if(pph::pick_highest_response_ratio(q.begin(), q.end(), 1.0, size, arr)->id != 1)
		return printf("md5 order: the short file should go first\n"), 1;
	q = {{8ull << 30, 0.0, 0}, {1 << 20, 99.99, 1}};
	if(pph::pick_highest_response_ratio(q.begin(), q.end(), 100.0, size, arr)->id != 0)
		return printf("md5 order: the long file that waited 100 s should go first\n"), 1;
	q.clear();
	if(pph::pick_highest_response_ratio(q.begin(), q.end(), 0.0, size, arr) != q.end())
		return printf("md5 order: empty queue\n"), 1;
	// a pool: 4 threads take files by response ratio while 2 threads submit
	std::mutex mu;
	std::condition_variable cv;
	std::deque<File> queue;
	bool closed = false;
	std::vector<std::atomic<unsigned>> hashed(2000);
	for(auto& h : hashed) h = 0;
	std::vector<std::thread> pool;
	for(int t = 0; t < 4; t++)
		pool.emplace_back([&] {
			for(;;) {
				File f;
				{
					std::unique_lock<std::mutex> lk(mu);
					cv.wait(lk, [&] { return closed || !queue.empty(); });
					if(queue.empty()) return;
					auto it = pph::pick_highest_response_ratio(queue.begin(), queue.end(), 1.0, size, arr);
					f = *it;
					queue.erase(it);
				}
				hashed[f.id]++;
			}
		});
	std::vector<std::thread> submitters;
	for(int t = 0; t < 2; t++)
		submitters.emplace_back([&, t] {
			for(unsigned k = 0; k < 1000; k++) {
				{
					std::lock_guard<std::mutex> lk(mu);
					queue.push_back(File{(uint64_t)(k * 7919 % 4096 + 1) << 20, k * 1e-3, t * 1000 + k});
				}
				cv.notify_one();
			}
		});
	for(auto& t : submitters) t.join();
	{
		std::lock_guard<std::mutex> lk(mu);
		closed = true;
	}
	cv.notify_all();
	for(auto& t : pool) t.join();
	for(unsigned k = 0; k < 2000; k++)
		if(hashed[k] != 1) return printf("md5 pool: file %u hashed %u times\n", k, hashed[k].load()), 1;
	return 0;
}

int main() {
	tables();
	if(check_queue() || check_pools() || check_md5_order()) return 1;
	printf("ok\n");
	return 0;
}
