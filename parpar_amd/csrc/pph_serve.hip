// Serving: a request loop with continuous batching.  Jobs (one file each: its
// slices are the PAR2 inputs) are submitted at any time; a scheduler thread
// admits them while device memory (a fixed arena, sized by the caller) and a
// compute stream are free, so new jobs start while earlier ones still run and
// their kernels overlap on the device.  An admitted job runs on its own stream:
//   inputs -> per-slice MD5/CRC-32 -> preparation (gf16_cksum, out of place)
//   -> mul-add -> gf16_cksum check of the recovery slices
// The mul-add is the plain form (log-domain preparation) or, for jobs where
// the autotuned rates predict it faster, the bilinear form (data-form
// preparation, the fp4 MFMA GEMM in an engine-owned workspace; the bilinear
// phases of different jobs are ordered through one event so that they take
// the workspace one at a time).
// Inputs come from the caller's host buffer through pinned staging chunks
// (double-buffered hipMemcpyAsync; one or more upload threads, each with its
// own pair of chunks), or from the counter-based generator on the device.
//
// The whole-file MD5 is one serial chain per file.  Running it as a kernel per
// job would hold a hardware queue for seconds on large files, so every file
// whose inputs are on the device joins a shared hash round instead: one
// kernel on the hash stream, one wavefront per file, each advancing its chain
// by at most `md5_quantum` bytes.  Rounds repeat while chains are unfinished;
// files join and leave at round boundaries.  Optionally, a pool of host
// threads hashes every file instead (a core runs the chain several times
// faster than a wavefront): bytes in host memory from submit on, bytes
// generated in HBM once they are there, read back through pinned chunks.
//
// A job's memory is two arena ranges: the inputs and hash results (held until
// its MD5 chain ends) and the log-domain inputs and recovery slices (released
// as soon as the recovery and its checks are copied back).
#include "pph_internal.h"
#include "pph_md5_order.h"
#include "pph_copy_pool.h"
#include "pph_md5_host.h"
#include "../../include/parpar_hip.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

namespace pph {
namespace {

double now_s() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over offsets of a device arena, 256-byte granules.
class Arena {
	std::map<size_t, size_t> free_;  // offset -> length
public:
	size_t capacity = 0, used = 0, peak = 0;
	void init(size_t cap) {
		capacity = cap;
		free_.clear();
		if(cap) free_[0] = cap;
	}
	bool alloc(size_t len, size_t* off) {
		len = round_up(len, 256);
		for(auto it = free_.begin(); it != free_.end(); ++it) {
			if(it->second < len) continue;
			*off = it->first;
			const size_t rest = it->second - len, noff = it->first + len;
			free_.erase(it);
			if(rest) free_[noff] = rest;
			used += len;
			peak = std::max(peak, used);
			return true;
		}
		return false;
	}
	void release(size_t off, size_t len) {
		len = round_up(len, 256);
		used -= len;
		auto it = free_.emplace(off, len).first;
		auto nx = std::next(it);
		if(nx != free_.end() && it->first + it->second == nx->first) {
			it->second += nx->second;
			free_.erase(nx);
		}
		if(it != free_.begin()) {
			auto pv = std::prev(it);
			if(pv->first + pv->second == it->first) {
				pv->second += it->second;
				free_.erase(it);
			}
		}
	}
};

constexpr unsigned kMaxBypass = 16;   // admissions that may pass a waiting head job
constexpr unsigned kLookBehind = 8;   // queued jobs considered behind the head
constexpr unsigned kMaxSpans = 1024;  // files per hash round
constexpr unsigned kMaxUploaders = 16;
constexpr size_t kMd5ReadChunk = 4u << 20;  // pinned read-back chunk of a host MD5 thread

struct Job {
	pph_serve_job spec;
	std::vector<uint16_t> exps;
	uint64_t ticket = 0;
	double t_arrival = 0, t_admit = 0, t_computed = 0;
	size_t n_in = 0, slot = 0, last_len = 0;
	unsigned n_out = 0, bypassed = 0;
	// arena ranges: A = inputs + hash results, B = compute buffers
	size_t off_a = 0, bytes_a = 0, off_b = 0, bytes_b = 0;
	bool have_a = false, have_b = false;
	uint8_t *data = nullptr, *ifsc = nullptr, *digest = nullptr;  // ifsc: slice MD5 + CRC-32, 20 bytes each
	uint32_t *lens = nullptr, *md5_state = nullptr;
	uint8_t *logd = nullptr, *out = nullptr;
	uint32_t* pk_state = nullptr;  // recovery-packet MD5 state (20 words per recovery slice)
	uint8_t* pk_md5 = nullptr;
	bool packet = false;           // recovery-packet MD5 requested
	uint32_t packet_head[8] = {};  // set ID and packet type words
	uint16_t *logs = nullptr, *d_exps = nullptr;
	int* valid = nullptr;
	int stream = -1;
	hipEvent_t ev_data = nullptr, ev_done = nullptr;
	int form = PPH_FORM_PLAIN;
	bool launched = false;   // inputs and compute enqueued (guarded by the engine mutex)
	bool delivering = false; // handed to the delivery thread
	bool computed = false;   // compute results copied back, B released (guarded by the engine mutex)
	bool hashing = false;    // in the MD5 rounds
	uint64_t md5_pos = 0;    // bytes of the chain dispatched
	bool md5_done = false;   // for host_md5 jobs: set by the hashing thread, guarded by the engine mutex
	bool host_md5 = false;   // file MD5 on a host thread (fixed at submit)
	bool host_queued = false;  // handed to the host MD5 threads (guarded by the engine mutex)
	uint8_t host_digest[16] = {};
	int status = PPH_OK;
};

struct Round {
	hipEvent_t ev = nullptr;
	std::vector<Job*> finals;  // jobs whose last span is in this round
};

// Two pinned chunks of one upload thread: the host fills one while the
// other's copy is in flight.
struct Stager {
	uint8_t* buf[2] = {nullptr, nullptr};
	hipEvent_t ev[2] = {nullptr, nullptr};  // last copy out of each chunk
};

}  // namespace
}  // namespace pph

using namespace pph;

struct pph_serve {
	int device = 0;
	pph_serve_config cfg{};
	DeviceTables* tab = nullptr;
	unsigned ck = 64;
	uint8_t* arena_mem = nullptr;
	Arena arena;
	std::vector<hipStream_t> streams;  // compute streams, one per running job
	std::vector<uint8_t> stream_busy;  // guarded by mu
	hipStream_t hash_stream = nullptr;
	Md5Span* h_spans[2] = {nullptr, nullptr};  // pinned
	Md5Span* d_spans[2] = {nullptr, nullptr};
	hipEvent_t round_ev[2] = {nullptr, nullptr};
	std::deque<Round> rounds;  // in flight, oldest first
	unsigned next_round_buf = 0;
	std::vector<Stager> stagers;  // one per upload thread
	std::vector<Stager> md5_stagers;  // one per host MD5 thread, allocated at its first device-resident file
	void* bl_ws = nullptr;               // bilinear workspace
	hipEvent_t bl_tail = nullptr;        // end of the last enqueued bilinear phase
	bool bl_any = false;
	std::mutex bl_mu;                    // orders the bilinear phases' enqueue
	pph_tuning tuning{};

	std::mutex mu;
	std::condition_variable cv;
	std::deque<Job*> pending;
	std::vector<Job*> active;  // admitted, not finished
	std::deque<Job*> uploads;
	std::deque<Job*> deliveries;
	std::deque<Job*> host_md5;  // file MD5 not yet started on a host thread: host bytes from submit,
	                            // device-resident bytes once they are in HBM
	std::vector<pph_serve_result> done;
	pph_serve_stats stats{};
	uint64_t next_ticket = 1;
	unsigned outstanding = 0;
	bool stop = false, sched_done = false;
	std::thread sched, deliverer;
	std::vector<std::thread> uploaders, md5_threads;
};

namespace {

size_t bytes_a(size_t n_in, size_t slot) {
	return round_up(n_in * slot, 256) + round_up(n_in * 4, 256) + round_up(n_in * 20, 256) + 2 * 256;
}
size_t bytes_b(size_t n_in, unsigned n_out, size_t slot, bool packet) {
	return round_up(n_in * slot, 256) + round_up((size_t)n_out * slot, 256) + round_up(n_in * 2, 256) +
	       round_up((size_t)n_out * 2, 256) + round_up((size_t)n_out * 4, 256) +
	       (packet ? round_up((size_t)n_out * 80, 256) + round_up((size_t)n_out * 16, 256) : 0);
}

void carve(pph_serve* s, Job* j) {
	uint8_t* p = s->arena_mem + j->off_a;
	auto take = [&](size_t n) {
		uint8_t* r = p;
		p += round_up(n, 256);
		return r;
	};
	j->data = take(j->n_in * j->slot);
	j->lens = (uint32_t*)take(j->n_in * 4);
	j->ifsc = take(j->n_in * 20);
	j->md5_state = (uint32_t*)take(16);
	j->digest = take(16);
	p = s->arena_mem + j->off_b;
	j->logd = take(j->n_in * j->slot);
	j->out = take((size_t)j->n_out * j->slot);
	j->logs = (uint16_t*)take(j->n_in * 2);
	j->d_exps = (uint16_t*)take((size_t)j->n_out * 2);
	j->valid = (int*)take((size_t)j->n_out * 4);
	if(j->packet) {
		j->pk_state = (uint32_t*)take((size_t)j->n_out * 80);
		j->pk_md5 = take((size_t)j->n_out * 16);
	}
}

// Hashes and recovery of one job on its stream, after its inputs.
hipError_t enqueue_compute(pph_serve* s, Job* j) {
	hipStream_t a = s->streams[j->stream];
	const size_t S = j->spec.slice_size;
	hipError_t e = hipEventRecord(j->ev_data, a);
	if(e == hipSuccess)
		e = launch_slice_hash_to(j->data, j->slot, (unsigned)j->n_in, j->lens, S, j->ifsc, 20, j->ifsc + 16, 20,
		                         *s->tab, a);
	if(e == hipSuccess && j->n_out && j->form == PPH_FORM_BILINEAR &&
	   bilinear_supported(j->slot, (unsigned)j->n_in, j->n_out) != PPH_OK)
		j->form = PPH_FORM_PLAIN;
	if(e == hipSuccess && j->n_out) {
		const bool bl = j->form == PPH_FORM_BILINEAR;
		e = launch_prep(j->logd, j->data, j->slot, (unsigned)j->n_in, j->lens, nullptr, S, ck_offset(S), s->ck, !bl,
		                *s->tab, a);
		if(e == hipSuccess && bl) {
			std::lock_guard<std::mutex> lk(s->bl_mu);
			if(s->bl_any) e = hipStreamWaitEvent(a, s->bl_tail, 0);
			if(e == hipSuccess &&
			   bilinear_enqueue(j->logd, j->slot, (unsigned)j->n_in, j->logs, j->d_exps, j->out, j->n_out, false, s->bl_ws,
			                    s->cfg.bilinear_workspace, *s->tab, a) != PPH_OK)
				e = hipErrorLaunchFailure;
			if(e == hipSuccess) e = hipEventRecord(s->bl_tail, a);
			if(e == hipSuccess) s->bl_any = true;
		} else if(e == hipSuccess) {
			// one workgroup per item: the engine's compute streams share the CUs
			e = launch_muladd_plain((const uint16_t*)j->logd, j->slot / 2, (unsigned)j->n_in, j->logs, j->d_exps,
			                        nullptr, 0, (uint16_t*)j->out, j->n_out, false, 0, *s->tab, a, false);
		}
		if(e == hipSuccess) e = launch_verify(j->out, j->slot, j->n_out, S, ck_offset(S), s->ck, j->valid, *s->tab, a);
		// recovery packets hashed while the slices are still in HBM (lib/par2.js:505-513)
		if(e == hipSuccess && j->packet)
			e = launch_packet_md5(j->pk_state, j->out, j->slot, j->n_out, j->d_exps, j->packet_head, 0, S, true, true,
			                      j->pk_md5, a);
	}
	if(e == hipSuccess) e = hipEventRecord(j->ev_done, a);
	return e;
}

// Reserve memory and a stream, upload the metadata, start the inputs.
// Runs on the scheduler thread.
bool admit(pph_serve* s, Job* j) {
	int st = -1;
	{
		// the delivery thread frees streams and memory under the same lock
		std::lock_guard<std::mutex> lk(s->mu);
		for(unsigned k = 0; k < s->streams.size(); k++)
			if(!s->stream_busy[k]) {
				st = (int)k;
				break;
			}
		if(st < 0) return false;
		if(!s->arena.alloc(j->bytes_a, &j->off_a)) return false;
		if(!s->arena.alloc(j->bytes_b, &j->off_b)) {
			s->arena.release(j->off_a, j->bytes_a);
			return false;
		}
		s->stream_busy[st] = 1;
	}
	j->have_a = j->have_b = true;
	j->stream = st;
	j->t_admit = now_s();
	carve(s, j);
	hipStream_t a = s->streams[st];
	const size_t S = j->spec.slice_size;
	std::vector<uint32_t> lens(j->n_in, (uint32_t)S);
	lens[j->n_in - 1] = (uint32_t)j->last_len;
	std::vector<uint16_t> logs(j->n_in);
	for(size_t i = 0; i < j->n_in; i++) logs[i] = host_input_log((uint16_t)(j->spec.first_input + i));
	hipError_t e = hipEventCreateWithFlags(&j->ev_data, hipEventDisableTiming);
	if(e == hipSuccess) e = hipEventCreateWithFlags(&j->ev_done, hipEventDisableTiming);
	if(e == hipSuccess) e = hipMemcpyAsync(j->lens, lens.data(), lens.size() * 4, hipMemcpyHostToDevice, a);
	if(e == hipSuccess) e = hipMemcpyAsync(j->logs, logs.data(), logs.size() * 2, hipMemcpyHostToDevice, a);
	if(e == hipSuccess && j->n_out)
		e = hipMemcpyAsync(j->d_exps, j->exps.data(), j->exps.size() * 2, hipMemcpyHostToDevice, a);
	if(e == hipSuccess) e = hipStreamSynchronize(a);  // the host vectors end here
	bool to_upload = false;
	if(e == hipSuccess && !j->spec.host_data) {
		// generated inputs: slice k is the generator's (seed, job_id, k), the last one truncated
		if(j->n_in > 1)
			e = launch_fill_random(j->data, j->slot, (unsigned)(j->n_in - 1), S, j->spec.seed, j->spec.job_id, 0, a);
		if(e == hipSuccess)
			e = launch_fill_random(j->data + (j->n_in - 1) * j->slot, j->slot, 1, j->last_len, j->spec.seed,
			                       j->spec.job_id, (uint32_t)(j->n_in - 1), a);
		if(e == hipSuccess) e = enqueue_compute(s, j);
	} else if(e == hipSuccess) {
		to_upload = true;
	}
	std::lock_guard<std::mutex> lk(s->mu);
	if(e != hipSuccess) j->status = set_error(e);
	if(to_upload) {
		s->uploads.push_back(j);
		s->cv.notify_all();
	} else {
		j->launched = true;
	}
	return true;
}

// Host-sourced inputs: chunked copies through two pinned buffers (the host
// fills one while the other is in flight), then the compute chain.
void upload_job(pph_serve* s, Job* j, Stager& st) {
	hipStream_t a = s->streams[j->stream];
	const uint8_t* src = (const uint8_t*)j->spec.host_data;
	const size_t S = j->spec.slice_size, chunk = s->cfg.staging_bytes;
	hipError_t e = hipSuccess;
	unsigned which = 0;
	for(size_t k = 0; k < j->n_in && e == hipSuccess; k++) {
		const size_t len = k + 1 < j->n_in ? S : j->last_len;
		for(size_t done = 0; done < len && e == hipSuccess;) {
			const size_t n = std::min(chunk, len - done);
			e = hipEventSynchronize(st.ev[which]);  // the last copy out of this chunk has finished
			if(e != hipSuccess) break;
			copy_stream(st.buf[which], src + k * S + done, n);  // streaming stores: the DMA reads it next
			e = hipMemcpyAsync(j->data + k * j->slot + done, st.buf[which], n, hipMemcpyHostToDevice, a);
			if(e == hipSuccess) e = hipEventRecord(st.ev[which], a);
			which ^= 1u;
			done += n;
		}
	}
	if(e == hipSuccess) e = enqueue_compute(s, j);
	std::lock_guard<std::mutex> lk(s->mu);
	if(e != hipSuccess) j->status = set_error(e);
	j->launched = true;
	s->cv.notify_all();
}

// Recovery, flags and slice hashes to the caller (synchronous copies into
// caller memory, on the delivery thread so that admission and the hash rounds
// go on meanwhile); compute memory and stream back.
void finish_compute(pph_serve* s, Job* j) {
	const pph_serve_job& sp = j->spec;
	hipError_t e = hipSuccess;
	int status;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		status = j->status;
	}
	if(status == PPH_OK) {
		if(sp.slice_md5) e = hipMemcpy2D(sp.slice_md5, 16, j->ifsc, 20, 16, j->n_in, hipMemcpyDeviceToHost);
		if(e == hipSuccess && sp.slice_crc)
			e = hipMemcpy2D(sp.slice_crc, 4, j->ifsc + 16, 20, 4, j->n_in, hipMemcpyDeviceToHost);
		if(e == hipSuccess && sp.ifsc) e = hipMemcpy(sp.ifsc, j->ifsc, j->n_in * 20, hipMemcpyDeviceToHost);
		if(e == hipSuccess && j->packet && sp.packet_md5)
			e = hipMemcpy(sp.packet_md5, j->pk_md5, (size_t)j->n_out * 16, hipMemcpyDeviceToHost);
		if(e == hipSuccess && sp.valid && j->n_out)
			e = hipMemcpy(sp.valid, j->valid, (size_t)j->n_out * 4, hipMemcpyDeviceToHost);
		if(e == hipSuccess && sp.recovery_out && j->n_out)
			e = hipMemcpy2D(sp.recovery_out, sp.slice_size, j->out, j->slot, sp.slice_size, j->n_out,
			                hipMemcpyDeviceToHost);
	} else {
		(void)hipStreamSynchronize(s->streams[j->stream]);
	}
	std::lock_guard<std::mutex> lk(s->mu);
	if(e != hipSuccess && j->status == PPH_OK) j->status = set_error(e);
	j->t_computed = now_s();
	j->computed = true;
	if(j->have_b) s->arena.release(j->off_b, j->bytes_b);
	j->have_b = false;
	s->stream_busy[j->stream] = 0;
	s->cv.notify_all();
}

void finish_job(pph_serve* s, Job* j) {
	const pph_serve_job& sp = j->spec;
	if(!j->md5_done) {
		// failed job still named by a queued round: let the rounds drain
		(void)hipStreamSynchronize(s->hash_stream);
	}
	for(Round& r : s->rounds) r.finals.erase(std::remove(r.finals.begin(), r.finals.end(), j), r.finals.end());
	int status;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		status = j->status;
	}
	if(status == PPH_OK && sp.file_md5) {
		if(j->host_md5) {
			memcpy(sp.file_md5, j->host_digest, 16);
		} else {
			hipError_t e = hipMemcpy(sp.file_md5, j->digest, 16, hipMemcpyDeviceToHost);
			if(e != hipSuccess) status = set_error(e);
		}
	}
	const double t = now_s();
	pph_serve_result r{};
	r.ticket = j->ticket;
	r.status = status;
	r.latency_s = t - j->t_arrival;
	r.queue_s = j->t_admit - j->t_arrival;
	r.compute_s = j->t_computed - j->t_admit;
	r.run_s = t - j->t_admit;
	r.mul_adds = (double)j->n_in * j->n_out * (double)(sp.slice_size / 2);
	r.bytes = (double)sp.file_size;
	if(j->ev_data) (void)hipEventDestroy(j->ev_data);
	if(j->ev_done) (void)hipEventDestroy(j->ev_done);
	std::lock_guard<std::mutex> lk(s->mu);
	if(j->have_a) s->arena.release(j->off_a, j->bytes_a);
	if(j->have_b) s->arena.release(j->off_b, j->bytes_b);
	s->active.erase(std::find(s->active.begin(), s->active.end(), j));
	s->done.push_back(r);
	s->stats.completed++;
	if(j->form == PPH_FORM_BILINEAR) s->stats.bilinear_jobs++;
	if(j->host_md5) s->stats.host_md5_jobs++;
	if(r.status == PPH_OK) {
		s->stats.mul_adds += r.mul_adds;
		s->stats.bytes += r.bytes;
	}
	s->outstanding--;
	s->cv.notify_all();
	delete j;
}

// Launch one hash round over the files whose chains are not fully dispatched.
bool launch_round(pph_serve* s, const std::vector<Job*>& hashing) {
	const unsigned buf = s->next_round_buf;
	Md5Span* h = s->h_spans[buf];
	Round r;
	unsigned n = 0;
	const uint64_t q = s->cfg.md5_quantum;
	std::vector<Job*> members;
	for(Job* j : hashing) {
		if(n == kMaxSpans) break;
		const uint64_t fs = j->spec.file_size;
		if(j->md5_pos >= fs) continue;  // fully dispatched
		Md5Span& e = h[n++];
		e.slots = j->data;
		e.slot_bytes = j->slot;
		e.slice_size = j->spec.slice_size;
		e.file_size = fs;
		e.begin = j->md5_pos;
		e.state = j->md5_state;
		e.digest = j->digest;
		e.flags = j->md5_pos == 0 ? kMd5SpanFirst : 0;
		e.pad = 0;
		if(fs - j->md5_pos <= q) {
			e.end = fs;
			e.flags |= kMd5SpanFinal;
			r.finals.push_back(j);
		} else {
			e.end = j->md5_pos + q;
		}
		members.push_back(j);
	}
	if(!n) return false;
	hipError_t err = hipMemcpyAsync(s->d_spans[buf], h, n * sizeof(Md5Span), hipMemcpyHostToDevice, s->hash_stream);
	if(err == hipSuccess) err = launch_file_md5_round(s->d_spans[buf], n, s->hash_stream);
	if(err == hipSuccess) err = hipEventRecord(s->round_ev[buf], s->hash_stream);
	if(err != hipSuccess) {
		(void)hipStreamSynchronize(s->hash_stream);
		std::lock_guard<std::mutex> lk(s->mu);
		for(Job* j : members) {
			j->status = set_error(err);
			j->md5_pos = j->spec.file_size;
			j->md5_done = true;
		}
		return false;
	}
	for(Job* j : members) j->md5_pos = std::min<uint64_t>(j->spec.file_size, j->md5_pos + q);
	r.ev = s->round_ev[buf];
	s->rounds.push_back(std::move(r));
	s->next_round_buf ^= 1u;
	s->stats.md5_rounds++;
	return true;
}

void deliver(pph_serve* s, Job* j) {
	j->delivering = true;
	std::lock_guard<std::mutex> lk(s->mu);
	s->deliveries.push_back(j);
	s->cv.notify_all();
}

void scheduler(pph_serve* s) {
	(void)hipSetDevice(s->device);
	for(;;) {
		bool progress = false;
		std::vector<Job*> act;
		{
			std::unique_lock<std::mutex> lk(s->mu);
			if(s->stop && s->pending.empty() && s->active.empty()) break;
			// admission: FIFO; up to kLookBehind queued jobs may pass a head
			// that does not fit yet, at most kMaxBypass times per head
			while(!s->pending.empty()) {
				Job* head = s->pending.front();
				lk.unlock();
				bool ok = admit(s, head);
				lk.lock();
				if(ok) {
					s->pending.pop_front();
					s->active.push_back(head);
					progress = true;
					continue;
				}
				unsigned looked = 0;
				for(auto it = s->pending.begin() + 1;
				    it != s->pending.end() && looked < kLookBehind && head->bypassed < kMaxBypass;) {
					Job* j = *it;
					lk.unlock();
					ok = admit(s, j);
					lk.lock();
					if(ok) {
						it = s->pending.erase(it);
						s->active.push_back(j);
						head->bypassed++;
						progress = true;
					} else {
						++it;
						looked++;
					}
				}
				break;
			}
			s->stats.peak_jobs = std::max(s->stats.peak_jobs, (unsigned)s->active.size());
			s->stats.peak_bytes = std::max<uint64_t>(s->stats.peak_bytes, s->arena.peak);
			act = s->active;
		}
		// compute completions and inputs that have arrived
		std::vector<Job*> hashing;
		for(Job* j : act) {
			bool launched;
			int status;
			{
				std::lock_guard<std::mutex> lk(s->mu);
				launched = j->launched;
				status = j->status;
			}
			if(!launched) continue;
			if(status != PPH_OK) {
				if(!j->delivering) deliver(s, j);
				continue;
			}
			if(!j->hashing && hipEventQuery(j->ev_data) == hipSuccess) {
				j->hashing = true;
				progress = true;
				if(j->host_md5 && !j->spec.host_data) {
					std::lock_guard<std::mutex> lk(s->mu);
					j->host_queued = true;
					s->host_md5.push_back(j);
					s->cv.notify_all();
				}
			}
			if(!j->delivering && hipEventQuery(j->ev_done) == hipSuccess) {
				deliver(s, j);
				progress = true;
			}
			if(j->hashing && !j->host_md5 && !j->md5_done) hashing.push_back(j);
		}
		// finished rounds, oldest first
		while(!s->rounds.empty() && hipEventQuery(s->rounds.front().ev) == hipSuccess) {
			for(Job* j : s->rounds.front().finals) j->md5_done = true;
			s->rounds.pop_front();
			progress = true;
		}
		// keep two rounds queued on the hash stream
		while(s->rounds.size() < 2 && launch_round(s, hashing)) progress = true;
		for(Job* j : act) {
			bool computed, md5_done, host_queued;
			{
				std::lock_guard<std::mutex> lk(s->mu);
				computed = j->computed;
				md5_done = j->md5_done;
				host_queued = j->host_queued;
			}
			// a host thread may still read a failed job's bytes: it ends only after its hash
			if(computed && (md5_done || (j->status != PPH_OK && !host_queued))) {
				finish_job(s, j);
				progress = true;
			}
		}
		if(!progress) {
			std::unique_lock<std::mutex> lk(s->mu);
			s->cv.wait_for(lk, std::chrono::microseconds(200));
		}
	}
}

void deliverer(pph_serve* s) {
	(void)hipSetDevice(s->device);
	for(;;) {
		Job* j = nullptr;
		{
			std::unique_lock<std::mutex> lk(s->mu);
			s->cv.wait(lk, [&] { return s->sched_done || !s->deliveries.empty(); });
			if(s->deliveries.empty()) break;
			j = s->deliveries.front();
			s->deliveries.pop_front();
		}
		finish_compute(s, j);
	}
}

void uploader(pph_serve* s, unsigned k) {
	(void)hipSetDevice(s->device);
	for(;;) {
		Job* j = nullptr;
		{
			std::unique_lock<std::mutex> lk(s->mu);
			s->cv.wait(lk, [&] { return s->sched_done || !s->uploads.empty(); });
			if(s->uploads.empty()) break;  // the scheduler has ended: nothing more can come
			j = s->uploads.front();
			s->uploads.pop_front();
		}
		upload_job(s, j, s->stagers[k]);
	}
}

// File MD5 of a job whose inputs are in HBM: its slices are read back in
// file order through the thread's two pinned chunks on the hash stream, the
// copy of the next piece in flight while the current one is hashed.  The
// job's input memory stays allocated until its MD5 is done.
int hash_device_file(pph_serve* s, Job* j, Stager& st, HostMd5& h) {
	const size_t chunk = std::min<size_t>(s->cfg.staging_bytes, kMd5ReadChunk);
	hipError_t e = hipSuccess;
	for(int b = 0; b < 2 && e == hipSuccess; b++) {
		if(!st.buf[b]) e = hipHostMalloc((void**)&st.buf[b], chunk, hipHostMallocDefault);
		if(e == hipSuccess && !st.ev[b]) e = hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming);
	}
	const size_t S = j->spec.slice_size;
	size_t k = 0, off = 0, n[2] = {0, 0};
	auto issue = [&](unsigned b) {  // the next piece (within one slice) into chunk b
		const size_t len = k + 1 < j->n_in ? S : j->last_len;
		n[b] = std::min(chunk, len - off);
		hipError_t r =
		    hipMemcpyAsync(st.buf[b], j->data + k * j->slot + off, n[b], hipMemcpyDeviceToHost, s->hash_stream);
		if(r == hipSuccess) r = hipEventRecord(st.ev[b], s->hash_stream);
		off += n[b];
		if(off == len) {
			k++;
			off = 0;
		}
		return r;
	};
	unsigned b = 0;
	if(e == hipSuccess) e = issue(0);
	while(e == hipSuccess) {
		const bool more = k < j->n_in;
		if(more) e = issue(b ^ 1u);  // chunk b^1 was hashed in the previous pass
		if(e == hipSuccess) e = hipEventSynchronize(st.ev[b]);
		if(e != hipSuccess) break;
		h.update(st.buf[b], n[b]);
		if(!more) break;
		b ^= 1u;
	}
	return e == hipSuccess ? PPH_OK : set_error(e);
}

// The queued file a free host MD5 thread takes (engine mutex held): the
// highest response ratio, (wait + hash time) / hash time with the wait counted
// from submit, or the oldest with md5_order = 1.  In a window that offers more
// file bytes than the threads hash, submit order leaves small files behind
// multi-second chains of large ones.
Job* next_host_md5(pph_serve* s) {
	auto pick = s->host_md5.begin();
	if(s->cfg.md5_order == 0)
		pick = pick_highest_response_ratio(s->host_md5.begin(), s->host_md5.end(), now_s(),
		                                   [](const Job* j) { return (uint64_t)j->spec.file_size; },
		                                   [](const Job* j) { return j->t_arrival; });
	Job* j = *pick;
	s->host_md5.erase(pick);
	return j;
}

// File MD5 on host thread k: host-sourced files from submit on, files
// generated in HBM once their inputs are there.
void md5_worker(pph_serve* s, unsigned k) {
	bool device_set = false;
	for(;;) {
		Job* j = nullptr;
		{
			std::unique_lock<std::mutex> lk(s->mu);
			s->cv.wait(lk, [&] { return s->sched_done || !s->host_md5.empty(); });
			if(s->host_md5.empty()) break;
			j = next_host_md5(s);
		}
		HostMd5 h;
		int rc = PPH_OK;
		if(j->spec.host_data) {
			h.update((const uint8_t*)j->spec.host_data, (size_t)j->spec.file_size);
		} else {
			if(!device_set) device_set = hipSetDevice(s->device) == hipSuccess;
			rc = device_set ? hash_device_file(s, j, s->md5_stagers[k], h) : PPH_E_DEVICE;
		}
		h.final(j->host_digest);
		std::lock_guard<std::mutex> lk(s->mu);
		if(rc != PPH_OK && j->status == PPH_OK) j->status = rc;
		j->md5_done = true;
		s->cv.notify_all();
	}
}

}  // namespace

extern "C" void pph_serve_destroy(pph_serve* s) {
	if(!s) return;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		s->stop = true;
		s->cv.notify_all();
	}
	if(s->sched.joinable()) s->sched.join();
	{
		std::lock_guard<std::mutex> lk(s->mu);
		s->sched_done = true;
		s->cv.notify_all();
	}
	for(auto& t : s->uploaders)
		if(t.joinable()) t.join();
	for(auto& t : s->md5_threads)
		if(t.joinable()) t.join();
	if(s->deliverer.joinable()) s->deliverer.join();
	(void)hipSetDevice(s->device);
	for(auto st : s->streams)
		if(st) {
			(void)hipStreamSynchronize(st);
			(void)hipStreamDestroy(st);
		}
	if(s->hash_stream) {
		(void)hipStreamSynchronize(s->hash_stream);
		(void)hipStreamDestroy(s->hash_stream);
	}
	for(auto* group : {&s->stagers, &s->md5_stagers})
		for(Stager& st : *group)
			for(int k = 0; k < 2; k++) {
				if(st.ev[k]) (void)hipEventDestroy(st.ev[k]);
				if(st.buf[k]) (void)hipHostFree(st.buf[k]);
			}
	for(int k = 0; k < 2; k++) {
		if(s->round_ev[k]) (void)hipEventDestroy(s->round_ev[k]);
		if(s->h_spans[k]) (void)hipHostFree(s->h_spans[k]);
		if(s->d_spans[k]) (void)hipFree(s->d_spans[k]);
	}
	if(s->arena_mem) (void)hipFree(s->arena_mem);
	if(s->bl_tail) (void)hipEventDestroy(s->bl_tail);
	if(s->bl_ws) (void)hipFree(s->bl_ws);
	delete s;
}

extern "C" int pph_serve_create(int device, const pph_serve_config* cfg, pph_serve** out) {
	if(!out || !cfg) return PPH_E_INVALID;
	*out = nullptr;
	pph_serve_config c = *cfg;
	if(!c.md5_quantum) c.md5_quantum = 8u << 20;
	if(!c.staging_bytes) c.staging_bytes = 64u << 20;
	if(!c.upload_threads) c.upload_threads = 1;
	if(!c.hbm_budget || !c.max_compute || c.max_compute > 64 || c.md5_quantum % 64 ||
	   c.upload_threads > kMaxUploaders || c.md5_host_threads > 256 || c.md5_order > 1)
		return PPH_E_INVALID;
	int n = 0;
	if(hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PPH_E_UNSUPPORTED;
	if(device < 0 || device >= n) return PPH_E_INVALID;
	auto* s = new(std::nothrow) pph_serve;
	if(!s) return PPH_E_NOMEM;
	s->device = device;
	s->cfg = c;
	hipError_t e = hipSetDevice(device);
	int rc = e == hipSuccess ? device_tables(&s->tab) : set_error(e);
	if(rc) {
		delete s;
		return rc;
	}
	e = hipMalloc((void**)&s->arena_mem, c.hbm_budget);
	if(e != hipSuccess) {
		s->arena_mem = nullptr;
		delete s;
		(void)hipGetLastError();
		return e == hipErrorOutOfMemory ? PPH_E_NOMEM : set_error(e);
	}
	s->arena.init(c.hbm_budget);
	s->streams.assign(c.max_compute, nullptr);
	s->stream_busy.assign(c.max_compute, 0);
	for(auto& st : s->streams)
		if(e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&s->hash_stream, hipStreamNonBlocking);
	s->stagers.resize(c.upload_threads);
	s->md5_stagers.resize(c.md5_host_threads);
	for(Stager& st : s->stagers)
		for(int k = 0; k < 2 && e == hipSuccess; k++) {
			e = hipHostMalloc((void**)&st.buf[k], c.staging_bytes, hipHostMallocDefault);
			if(e == hipSuccess) e = hipEventCreateWithFlags(&st.ev[k], hipEventDisableTiming);
			if(e == hipSuccess) e = hipEventRecord(st.ev[k], s->hash_stream);  // "chunk free"
		}
	for(int k = 0; k < 2 && e == hipSuccess; k++) {
		e = hipHostMalloc((void**)&s->h_spans[k], kMaxSpans * sizeof(Md5Span), hipHostMallocDefault);
		if(e == hipSuccess) e = hipMalloc((void**)&s->d_spans[k], kMaxSpans * sizeof(Md5Span));
		if(e == hipSuccess) e = hipEventCreateWithFlags(&s->round_ev[k], hipEventDisableTiming);
	}
	if(e == hipSuccess) e = hipStreamSynchronize(s->hash_stream);
	if(e == hipSuccess && c.bilinear_workspace) {
		e = hipMalloc(&s->bl_ws, c.bilinear_workspace);
		if(e == hipSuccess) e = hipEventCreateWithFlags(&s->bl_tail, hipEventDisableTiming);
		if(e == hipSuccess) {
			if(c.tuning) {
				s->tuning = *c.tuning;
			} else {
				rc = autotune(&s->tuning);
				if(rc) {
					pph_serve_destroy(s);
					return rc;
				}
			}
		}
	}
	if(e != hipSuccess) {
		pph_serve_destroy(s);
		return e == hipErrorOutOfMemory ? PPH_E_NOMEM : set_error(e);
	}
	s->stats.arena_bytes = c.hbm_budget;
	s->sched = std::thread(scheduler, s);
	for(unsigned k = 0; k < c.upload_threads; k++) s->uploaders.emplace_back(uploader, s, k);
	for(unsigned k = 0; k < c.md5_host_threads; k++) s->md5_threads.emplace_back(md5_worker, s, k);
	s->deliverer = std::thread(deliverer, s);
	*out = s;
	return PPH_OK;
}

extern "C" int pph_serve_submit(pph_serve* s, const pph_serve_job* job, uint64_t* ticket) {
	if(!s || !job) return PPH_E_INVALID;
	const uint64_t S = job->slice_size;
	if(!S || S % 4 || S > 0xFFFFFFF0ull || !job->file_size) return PPH_E_INVALID;
	const uint64_t n_in = (job->file_size + S - 1) / S;
	if(n_in + job->first_input > 32768 || job->n_out > 65535) return PPH_E_INVALID;
	auto* j = new(std::nothrow) Job;
	if(!j) return PPH_E_NOMEM;
	j->spec = *job;
	j->n_in = (size_t)n_in;
	j->n_out = job->n_out;
	j->last_len = (size_t)(job->file_size - (n_in - 1) * S);
	j->slot = slot_size((size_t)S, s->ck);
	j->exps.resize(job->n_out);
	for(unsigned r = 0; r < job->n_out; r++) {
		j->exps[r] = job->exps ? job->exps[r] : (uint16_t)r;
		if(j->exps[r] == 65535) {
			delete j;
			return PPH_E_INVALID;
		}
	}
	if(s->bl_ws && j->n_out) {
		// form by predicted time, the bilinear form only where its workspace fits
		const double ma = (double)j->n_in * j->n_out * (double)(j->slot / 2);
		const double rb = bilinear_rate(s->tuning, (unsigned)j->n_in, j->n_out);
		if(rb > 0 && ma / rb < ma / plain_rate(s->tuning, j->n_out) &&
		   bilinear_workspace_bytes(j->slot, (unsigned)j->n_in, j->n_out) <= s->cfg.bilinear_workspace)
			j->form = PPH_FORM_BILINEAR;
	}
	if(job->packet_set_id && j->n_out) {
		static const char kType[16] = {'P', 'A', 'R', ' ', '2', '.', '0', '\0', 'R', 'e', 'c', 'v', 'S', 'l', 'i', 'c'};
		j->packet = true;
		memcpy(j->packet_head, job->packet_set_id, 16);
		memcpy(j->packet_head + 4, kType, 16);
	}
	j->spec.packet_set_id = nullptr;  // copied: the caller's buffer may go
	j->bytes_a = bytes_a(j->n_in, j->slot);
	j->bytes_b = bytes_b(j->n_in, j->n_out, j->slot, j->packet);
	if(j->bytes_a + j->bytes_b > s->cfg.hbm_budget) {
		delete j;
		return PPH_E_NOMEM;  // could never be admitted
	}
	j->host_md5 = s->cfg.md5_host_threads != 0;
	j->host_queued = j->host_md5 && job->host_data;
	j->t_arrival = now_s();
	std::lock_guard<std::mutex> lk(s->mu);
	if(s->stop) {
		delete j;
		return PPH_E_STATE;
	}
	j->ticket = s->next_ticket++;
	if(ticket) *ticket = j->ticket;
	s->pending.push_back(j);
	if(j->host_queued) s->host_md5.push_back(j);
	s->outstanding++;
	s->stats.submitted++;
	s->cv.notify_all();
	return PPH_OK;
}

extern "C" int pph_serve_poll(pph_serve* s, pph_serve_result* out, unsigned max, unsigned* n, int wait_ms) {
	if(!s || !n || (max && !out)) return PPH_E_INVALID;
	std::unique_lock<std::mutex> lk(s->mu);
	if(s->done.empty() && wait_ms > 0)
		s->cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
	const unsigned k = (unsigned)std::min<size_t>(max, s->done.size());
	std::copy(s->done.begin(), s->done.begin() + k, out);
	s->done.erase(s->done.begin(), s->done.begin() + k);
	*n = k;
	return PPH_OK;
}

extern "C" int pph_serve_drain(pph_serve* s, int wait_ms) {
	if(!s) return PPH_E_INVALID;
	std::unique_lock<std::mutex> lk(s->mu);
	s->cv.wait_for(lk, std::chrono::milliseconds(wait_ms < 0 ? 0 : wait_ms), [&] { return s->outstanding == 0; });
	return (int)s->outstanding;
}

extern "C" int pph_serve_get_stats(pph_serve* s, pph_serve_stats* out) {
	if(!s || !out) return PPH_E_INVALID;
	std::lock_guard<std::mutex> lk(s->mu);
	*out = s->stats;
	return PPH_OK;
}
