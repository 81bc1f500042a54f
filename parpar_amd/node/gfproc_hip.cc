// Node binding: a GfProc whose backends are AMD GPUs (PAR2ProcHIP through the
// C ABI of libparpar_hip.so).  Same JavaScript interface as ParPar's GfProc
// (src/gf.cc:236-733): constructor (sliceSize, cpuOpts, oclOpts, stagingCount)
// plus a fifth argument, the `proc_hip` specs (the fields of GfOclSpec,
// src/gf.cc:217-224,301-330: device, slice_offset, slice_size,
// input_batchsize, input_minbatchsize, method; and cksum_size), and the
// methods add / end / get / setRecoverySlices / setCurrentSliceSize /
// setProgressCb / info / freeMem / close, with the callbacks lib/par2.js
// expects (lib/par2.js:97-230).  The byte-range split over several backends
// follows PAR2Proc (gf16/controller.cpp:42-130).
//
// Threads: add() copies the caller's buffer into pinned staging on the
// JavaScript thread and returns false when every staging area is full, as
// PAR2Proc::addInput does; end() and get() run on one worker thread per
// object.  Every pph_proc call holds the object's mutex (a pph_proc is driven
// by one thread at a time).  Completions reach JavaScript through a
// thread-safe function.  After a refused add the worker waits for a staging
// area and then calls the progress callback, where lib/par2.js retries.
#include <node_api.h>

#include "parpar_hip.h"
#include "../csrc/pph_task_queue.h"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

#define NAPI_OK(call)                                                         \
	do {                                                                      \
		if((call) != napi_ok) {                                               \
			napi_throw_error(env, nullptr, "N-API call failed: " #call);      \
			return nullptr;                                                   \
		}                                                                     \
	} while(0)

struct Backend {
	pph_proc* p = nullptr;
	int device = 0, form = PPH_FORM_AUTO;
	unsigned batch = 0, min_batch = 0, cksum = 64;
	size_t alloc_offset = 0, alloc_size = 0;  // initial allocation (PAR2ProcBackendAlloc)
	size_t offset = 0, size = 0;              // current range
};

struct Task {
	std::function<void()> work;          // worker thread
	std::function<void(napi_env)> done;  // JavaScript thread
	napi_ref owner = nullptr;            // keeps the JavaScript object alive until done has run
};

struct GfProc {
	std::vector<Backend> be;
	size_t slice_size = 0, cur_size = 0;
	unsigned n_out = 0;
	bool running = false, closed = false, pending_discard = true, has_output = false, has_added = false;
	bool waiting_for_add = false;
	napi_ref progress_cb = nullptr;

	std::mutex proc_mu;  // every pph_proc call
	pph::TaskQueue<Task*> queue;
	std::thread worker;
	napi_threadsafe_function tsfn = nullptr;
	unsigned outstanding = 0;  // posted, completion not yet run (JavaScript thread only)

	// From the JavaScript thread.  While anything is outstanding the completion
	// channel keeps the event loop alive, as a pending libuv request would.
	void post(napi_env env, Task* t) {
		if(outstanding++ == 0) napi_ref_threadsafe_function(env, tsfn);
		queue.push(t);
	}
	void run() {
		Task* t;
		while(queue.pop(&t)) {
			if(t->work) t->work();
			napi_call_threadsafe_function(tsfn, t, napi_tsfn_blocking);
			queue.finished();
		}
	}
	void destroy_procs() {
		std::lock_guard<std::mutex> lk(proc_mu);
		for(auto& b : be)
			if(b.p) pph_proc_destroy(b.p);
		be.clear();
	}
};

// completions on the JavaScript thread
void call_js(napi_env env, napi_value, void* context, void* data) {
	Task* t = static_cast<Task*>(data);
	GfProc* g = static_cast<GfProc*>(context);
	if(env) {
		if(--g->outstanding == 0) napi_unref_threadsafe_function(env, g->tsfn);
		if(t->done) t->done(env);
		if(t->owner) napi_delete_reference(env, t->owner);
	}
	delete t;
}

std::string status_text(int rc) { return std::string(pph_strerror(rc)); }

bool get_num(napi_env env, napi_value obj, const char* key, double* out) {
	bool has = false;
	if(napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
	napi_value v;
	napi_valuetype t;
	if(napi_get_named_property(env, obj, key, &v) != napi_ok || napi_typeof(env, v, &t) != napi_ok) return false;
	if(t != napi_number) return false;
	return napi_get_value_double(env, v, out) == napi_ok;
}

bool is_nullish(napi_env env, napi_value v) {
	napi_valuetype t;
	return napi_typeof(env, v, &t) != napi_ok || t == napi_null || t == napi_undefined;
}

GfProc* unwrap(napi_env env, napi_callback_info info, size_t* argc, napi_value* argv, napi_value* this_out = nullptr) {
	napi_value self;
	if(napi_get_cb_info(env, info, argc, argv, &self, nullptr) != napi_ok) return nullptr;
	if(this_out) *this_out = self;
	GfProc* g = nullptr;
	if(napi_unwrap(env, self, reinterpret_cast<void**>(&g)) != napi_ok || !g) {
		napi_throw_error(env, nullptr, "Not a GfProc");
		return nullptr;
	}
	if(g->closed) {
		napi_throw_error(env, nullptr, "Already closed");
		return nullptr;
	}
	return g;
}

// Call a persistent callback with args and drop the references.
void call_ref(napi_env env, napi_ref cb, std::vector<napi_value> args) {
	napi_value fn, global, ret;
	if(napi_get_reference_value(env, cb, &fn) == napi_ok && fn && napi_get_global(env, &global) == napi_ok)
		napi_call_function(env, global, fn, args.size(), args.data(), &ret);
}

void finalize(napi_env env, void* data, void*) {
	GfProc* g = static_cast<GfProc*>(data);
	g->queue.close();
	if(g->worker.joinable()) g->worker.join();
	if(g->tsfn) napi_release_threadsafe_function(g->tsfn, napi_tsfn_abort);
	g->destroy_procs();
	if(g->progress_cb) napi_delete_reference(env, g->progress_cb);
	delete g;
}

// PAR2Proc::init's check of the allocation (gf16/controller.cpp:20-40): the
// ranges, sorted by offset, tile [0, slice_size); only the last may be odd.
bool check_allocation(std::vector<Backend>& be, size_t slice_size) {
	std::sort(be.begin(), be.end(), [](const Backend& a, const Backend& b) { return a.alloc_offset < b.alloc_offset; });
	size_t pos = 0;
	for(size_t i = 0; i < be.size(); i++) {
		const Backend& b = be[i];
		if(b.alloc_offset != pos || !b.alloc_size || (b.alloc_offset & 1)) return false;
		if((b.alloc_size & 1) && i + 1 != be.size()) return false;
		pos += b.alloc_size;
	}
	return pos == slice_size;
}

napi_value New(napi_env env, napi_callback_info info) {
	size_t argc = 5;
	napi_value argv[5];
	napi_value self;
	NAPI_OK(napi_get_cb_info(env, info, &argc, argv, &self, nullptr));
	if(argc < 1) {
		napi_throw_error(env, nullptr, "Slice size required");
		return nullptr;
	}
	double slice = 0;
	NAPI_OK(napi_get_value_double(env, argv[0], &slice));
	const size_t slice_size = (size_t)slice;
	if(slice_size < 2 || (slice_size & 1)) {
		napi_throw_error(env, nullptr, "Slice size is invalid");
		return nullptr;
	}
	if(argc >= 2 && !is_nullish(env, argv[1])) {
		napi_throw_error(env, nullptr, "CPU processing is ParPar's own binding; pass null for the CPU options");
		return nullptr;
	}
	if(argc >= 3 && !is_nullish(env, argv[2])) {
		uint32_t n = 0;
		bool arr = false;
		napi_is_array(env, argv[2], &arr);
		if(arr) napi_get_array_length(env, argv[2], &n);
		if(!arr || n) {
			napi_throw_error(env, nullptr, "OpenCL devices are ParPar's own binding; pass [] or null");
			return nullptr;
		}
	}
	int staging = 2;
	if(argc >= 4 && !is_nullish(env, argv[3])) {
		NAPI_OK(napi_get_value_int32(env, argv[3], &staging));
		if(staging < 1 || staging > 32768) {
			napi_throw_error(env, nullptr, "Staging area count is invalid");
			return nullptr;
		}
	}
	auto* g = new GfProc;
	g->slice_size = g->cur_size = slice_size;
	if(argc >= 5 && !is_nullish(env, argv[4])) {
		bool arr = false;
		napi_is_array(env, argv[4], &arr);
		uint32_t n = 0;
		if(arr) napi_get_array_length(env, argv[4], &n);
		if(!arr || !n) {
			delete g;
			napi_throw_error(env, nullptr, "proc_hip must be a non-empty array of device specs");
			return nullptr;
		}
		for(uint32_t i = 0; i < n; i++) {
			napi_value o;
			NAPI_OK(napi_get_element(env, argv[4], i, &o));
			Backend b;
			double v;
			if(get_num(env, o, "device", &v)) b.device = (int)v;
			b.alloc_offset = get_num(env, o, "slice_offset", &v) ? (size_t)v : 0;
			b.alloc_size = get_num(env, o, "slice_size", &v) ? (size_t)v : slice_size - std::min(b.alloc_offset, slice_size);
			if(get_num(env, o, "input_batchsize", &v)) b.batch = (unsigned)v;
			if(get_num(env, o, "input_minbatchsize", &v)) b.min_batch = (unsigned)v;
			if(get_num(env, o, "method", &v)) b.form = (int)v;
			if(get_num(env, o, "cksum_size", &v)) b.cksum = (unsigned)v;
			if(b.batch > 32768 || b.alloc_offset > slice_size) {
				delete g;
				napi_throw_error(env, nullptr, "Invalid proc_hip spec");
				return nullptr;
			}
			g->be.push_back(b);
		}
	} else {
		Backend b;
		b.alloc_size = slice_size;
		g->be.push_back(b);
	}
	if(!check_allocation(g->be, slice_size)) {
		delete g;
		napi_throw_error(env, nullptr, "Slice portions allocated to the GPUs are invalid");
		return nullptr;
	}
	for(auto& b : g->be) {
		int rc = pph_proc_create(b.device, (unsigned)staging, &b.p);
		if(!rc) rc = pph_proc_set_slice_size(b.p, b.alloc_size);
		if(!rc) rc = pph_proc_init(b.p, b.cksum, b.batch, b.form);
		if(rc) {
			g->destroy_procs();
			delete g;
			napi_throw_error(env, nullptr, ("Failed to initialise the GPU: " + status_text(rc)).c_str());
			return nullptr;
		}
		pph_proc_set_min_input_batch_size(b.p, b.min_batch);
		b.offset = b.alloc_offset;
		b.size = b.alloc_size;
	}
	napi_value name;
	NAPI_OK(napi_create_string_utf8(env, "gfproc_hip", NAPI_AUTO_LENGTH, &name));
	if(napi_create_threadsafe_function(env, nullptr, nullptr, name, 0, 1, nullptr, nullptr, g, call_js,
	                                   &g->tsfn) != napi_ok) {
		g->destroy_procs();
		delete g;
		napi_throw_error(env, nullptr, "Could not create the completion channel");
		return nullptr;
	}
	// idle, the completion channel does not keep the process alive
	napi_unref_threadsafe_function(env, g->tsfn);
	g->worker = std::thread([g] { g->run(); });
	if(napi_wrap(env, self, g, finalize, nullptr, nullptr) != napi_ok) {
		finalize(env, g, nullptr);
		napi_throw_error(env, nullptr, "Could not wrap the object");
		return nullptr;
	}
	return nullptr;
}

napi_value SetRecoverySlices(napi_env env, napi_callback_info info) {
	size_t argc = 1;
	napi_value argv[1];
	GfProc* g = unwrap(env, info, &argc, argv);
	if(!g) return nullptr;
	if(g->running) {
		napi_throw_error(env, nullptr, "Cannot change params whilst running");
		return nullptr;
	}
	bool arr = false;
	if(argc < 1 || napi_is_array(env, argv[0], &arr) != napi_ok || !arr) {
		napi_throw_error(env, nullptr, "List of recovery indicies required");
		return nullptr;
	}
	uint32_t n = 0;
	NAPI_OK(napi_get_array_length(env, argv[0], &n));
	if(n < 1 || n > 65534) {
		napi_throw_error(env, nullptr, n ? "Too many recovery indicies specified"
		                                 : "At least one recovery index must be supplied");
		return nullptr;
	}
	std::vector<uint16_t> exps(n);
	for(uint32_t i = 0; i < n; i++) {
		napi_value v;
		uint32_t e = 0;
		NAPI_OK(napi_get_element(env, argv[0], i, &v));
		NAPI_OK(napi_get_value_uint32(env, v, &e));
		if(e > 65534) {
			napi_throw_error(env, nullptr, "Invalid recovery index supplied");
			return nullptr;
		}
		exps[i] = (uint16_t)e;
	}
	std::lock_guard<std::mutex> lk(g->proc_mu);
	for(auto& b : g->be) {
		const int rc = pph_proc_set_recovery_slices(b.p, n, exps.data());
		if(rc) {
			napi_throw_error(env, nullptr, ("Failed to allocate memory: " + status_text(rc)).c_str());
			return nullptr;
		}
	}
	g->n_out = n;
	g->has_output = false;
	return nullptr;
}

// PAR2Proc::setCurrentSliceSize (gf16/controller.cpp:87-115)
napi_value SetCurrentSliceSize(napi_env env, napi_callback_info info) {
	size_t argc = 1;
	napi_value argv[1];
	GfProc* g = unwrap(env, info, &argc, argv);
	if(!g) return nullptr;
	if(g->running) {
		napi_throw_error(env, nullptr, "Cannot change params whilst running");
		return nullptr;
	}
	double v = 0;
	if(argc < 1 || napi_get_value_double(env, argv[0], &v) != napi_ok) {
		napi_throw_error(env, nullptr, "Argument required");
		return nullptr;
	}
	const size_t size = (size_t)v;
	if(size < 2 || (size & 1)) {
		napi_throw_error(env, nullptr, "Slice size is invalid");
		return nullptr;
	}
	g->has_output = false;
	std::lock_guard<std::mutex> lk(g->proc_mu);
	bool ok = true;
	if(g->be.size() == 1) {
		g->be[0].offset = 0;
		g->be[0].size = size;
		ok = !pph_proc_set_current_slice_size(g->be[0].p, size);
	} else {
		size_t total = 0;
		for(auto& b : g->be) total += b.alloc_size;
		ok = size <= total;
		size_t pos = 0;
		for(auto& b : g->be) {
			if(!ok) break;
			b.offset = pos;
			b.size = std::min(size - pos, b.alloc_size);
			if(b.size) ok = !pph_proc_set_current_slice_size(b.p, b.size);
			pos += b.size;
		}
	}
	if(!ok) {
		napi_throw_error(env, nullptr, "Failed to allocate memory");
		return nullptr;
	}
	g->cur_size = size;
	return nullptr;
}

napi_value SetNumThreads(napi_env env, napi_callback_info info) {
	size_t argc = 0;
	if(!unwrap(env, info, &argc, nullptr)) return nullptr;
	napi_throw_error(env, nullptr, "CPU processing not enabled");
	return nullptr;
}

napi_value SetProgressCb(napi_env env, napi_callback_info info) {
	size_t argc = 1;
	napi_value argv[1];
	GfProc* g = unwrap(env, info, &argc, argv);
	if(!g) return nullptr;
	if(g->progress_cb) napi_delete_reference(env, g->progress_cb);
	g->progress_cb = nullptr;
	if(argc >= 1 && !is_nullish(env, argv[0])) {
		napi_valuetype t;
		NAPI_OK(napi_typeof(env, argv[0], &t));
		if(t != napi_function) {
			napi_throw_error(env, nullptr, "Callback required");
			return nullptr;
		}
		NAPI_OK(napi_create_reference(env, argv[0], 1, &g->progress_cb));
	}
	return nullptr;
}

// After a refused add: wait for a staging area, then tell JavaScript (the
// progress callback is where lib/par2.js retries its queued adds).
void schedule_wait_for_add(napi_env env, GfProc* g, napi_value self) {
	if(g->waiting_for_add) return;
	g->waiting_for_add = true;
	auto* t = new Task;
	napi_create_reference(env, self, 1, &t->owner);
	t->work = [g] {
		std::lock_guard<std::mutex> lk(g->proc_mu);
		for(auto& b : g->be)
			if(b.size) pph_proc_wait_for_add(b.p);
	};
	t->done = [g](napi_env env) {
		g->waiting_for_add = false;
		if(g->progress_cb) {
			unsigned started = 0;
			for(auto& b : g->be)
				started = std::max(started, pph_proc_batches_started(b.p) * pph_proc_input_batch_size(b.p));
			napi_value n;
			napi_create_uint32(env, started, &n);
			call_ref(env, g->progress_cb, {n});
		}
	};
	g->post(env, t);
}

// add(inputIndex, buffer, cb(idx, buffer)) -> added (src/gf.cc:570-622)
napi_value AddSlice(napi_env env, napi_callback_info info) {
	size_t argc = 3;
	napi_value argv[3], self;
	GfProc* g = unwrap(env, info, &argc, argv, &self);
	if(!g) return nullptr;
	if(!g->n_out) {
		napi_throw_error(env, nullptr, "setRecoverySlices not yet called");
		return nullptr;
	}
	bool is_buf = false;
	napi_valuetype t;
	if(argc < 3 || napi_is_buffer(env, argv[1], &is_buf) != napi_ok || !is_buf ||
	   napi_typeof(env, argv[2], &t) != napi_ok || t != napi_function) {
		napi_throw_error(env, nullptr, "Requires an input index, an input buffer and a callback");
		return nullptr;
	}
	int32_t idx = -1;
	NAPI_OK(napi_get_value_int32(env, argv[0], &idx));
	if(idx < 0 || idx > 32767) {
		napi_throw_error(env, nullptr, "Input index not valid");
		return nullptr;
	}
	void* data = nullptr;
	size_t len = 0;
	NAPI_OK(napi_get_buffer_info(env, argv[1], &data, &len));
	if(len > g->cur_size) {
		napi_throw_error(env, nullptr, "Input buffer too large");
		return nullptr;
	}
	napi_value result;
	{
		std::lock_guard<std::mutex> lk(g->proc_mu);
		// PAR2Proc::addInput refuses when a backend is full (gf16/controller.cpp:142-170)
		bool full = false;
		for(auto& b : g->be)
			if(b.size && b.offset < len && pph_proc_can_add(b.p) == PPH_ADD_FULL) full = true;
		if(full) {
			NAPI_OK(napi_get_boolean(env, false, &result));
			schedule_wait_for_add(env, g, self);
			return result;
		}
		g->running = true;
		g->has_output = false;
		if(g->pending_discard) {
			g->pending_discard = false;
			for(auto& b : g->be) pph_proc_discard_output(b.p);
		}
		for(auto& b : g->be) {
			if(!b.size || b.offset >= len) continue;
			const size_t amount = std::min(len - b.offset, b.size);
			const int rc =
			    pph_proc_add_input(b.p, static_cast<const uint8_t*>(data) + b.offset, amount, (uint16_t)idx, 0);
			if(rc < 0) {
				napi_throw_error(env, nullptr, ("addInput failed: " + status_text(rc)).c_str());
				return nullptr;
			}
		}
		g->has_added = true;
	}
	// the buffer has been copied: hand it back, asynchronously as the reference does
	auto* task = new Task;
	NAPI_OK(napi_create_reference(env, self, 1, &task->owner));
	napi_ref cb_ref, buf_ref;
	NAPI_OK(napi_create_reference(env, argv[2], 1, &cb_ref));
	NAPI_OK(napi_create_reference(env, argv[1], 1, &buf_ref));
	task->done = [cb_ref, buf_ref, idx](napi_env env) {
		napi_value i, buf;
		napi_create_int32(env, idx, &i);
		napi_get_reference_value(env, buf_ref, &buf);
		call_ref(env, cb_ref, {i, buf});
		napi_delete_reference(env, cb_ref);
		napi_delete_reference(env, buf_ref);
	};
	g->post(env, task);
	NAPI_OK(napi_get_boolean(env, true, &result));
	return result;
}

// end(cb) (src/gf.cc:624-648)
napi_value EndInput(napi_env env, napi_callback_info info) {
	size_t argc = 1;
	napi_value argv[1], self;
	GfProc* g = unwrap(env, info, &argc, argv, &self);
	if(!g) return nullptr;
	napi_valuetype t;
	if(argc < 1 || napi_typeof(env, argv[0], &t) != napi_ok || t != napi_function) {
		napi_throw_error(env, nullptr, "Callback required");
		return nullptr;
	}
	if(!g->n_out) {
		napi_throw_error(env, nullptr, "setRecoverySlices not yet called");
		return nullptr;
	}
	if(g->pending_discard) {
		std::lock_guard<std::mutex> lk(g->proc_mu);
		for(auto& b : g->be) pph_proc_discard_output(b.p);
		g->has_added = false;
	}
	napi_ref cb_ref;
	NAPI_OK(napi_create_reference(env, argv[0], 1, &cb_ref));
	auto* task = new Task;
	NAPI_OK(napi_create_reference(env, self, 1, &task->owner));
	auto* rc = new int(0);
	task->work = [g, rc] {
		std::lock_guard<std::mutex> lk(g->proc_mu);
		for(auto& b : g->be)
			if(b.size) {
				const int r = pph_proc_end_input(b.p);
				if(r < 0 && !*rc) *rc = r;
			}
	};
	task->done = [g, cb_ref, rc](napi_env env) {
		g->running = false;
		g->has_output = true;
		g->pending_discard = true;
		std::vector<napi_value> args;
		if(*rc) {
			napi_value msg, err;
			napi_create_string_utf8(env, ("endInput failed: " + status_text(*rc)).c_str(), NAPI_AUTO_LENGTH, &msg);
			napi_create_error(env, nullptr, msg, &err);
			args.push_back(err);
		}
		delete rc;
		call_ref(env, cb_ref, args);
		napi_delete_reference(env, cb_ref);
	};
	g->post(env, task);
	return nullptr;
}

// get(index, buffer, cb(idx, cksumValid, buffer)) (src/gf.cc:650-700)
napi_value GetOutputSlice(napi_env env, napi_callback_info info) {
	size_t argc = 3;
	napi_value argv[3], self;
	GfProc* g = unwrap(env, info, &argc, argv, &self);
	if(!g) return nullptr;
	if(g->running) {
		napi_throw_error(env, nullptr, "Cannot get output whilst running");
		return nullptr;
	}
	if(!g->has_output) {
		napi_throw_error(env, nullptr, "No finalized output to retrieve");
		return nullptr;
	}
	bool is_buf = false;
	napi_valuetype t;
	if(argc < 3 || napi_is_buffer(env, argv[1], &is_buf) != napi_ok || !is_buf ||
	   napi_typeof(env, argv[2], &t) != napi_ok || t != napi_function) {
		napi_throw_error(env, nullptr, "Requires an index, an output buffer and a callback");
		return nullptr;
	}
	int32_t idx = -1;
	NAPI_OK(napi_get_value_int32(env, argv[0], &idx));
	if(idx < 0 || (unsigned)idx >= g->n_out) {
		napi_throw_error(env, nullptr, "Recovery index is not valid");
		return nullptr;
	}
	void* data = nullptr;
	size_t len = 0;
	NAPI_OK(napi_get_buffer_info(env, argv[1], &data, &len));
	if(len < g->cur_size) {
		napi_throw_error(env, nullptr, "Output buffer too small");
		return nullptr;
	}
	napi_ref cb_ref, buf_ref;
	NAPI_OK(napi_create_reference(env, argv[2], 1, &cb_ref));
	NAPI_OK(napi_create_reference(env, argv[1], 1, &buf_ref));
	auto* task = new Task;
	NAPI_OK(napi_create_reference(env, self, 1, &task->owner));
	auto* valid = new int(1);
	uint8_t* out = static_cast<uint8_t*>(data);
	const bool added = g->has_added;
	task->work = [g, idx, out, valid, added] {
		std::lock_guard<std::mutex> lk(g->proc_mu);
		// PAR2Proc::getOutput (gf16/controller.cpp:360-400): nothing added reads zeros
		for(auto& b : g->be) {
			if(!b.size) continue;
			int v = 1;
			if(!added || pph_proc_get_output(b.p, (unsigned)idx, out + b.offset, &v) < 0) {
				if(!added) memset(out + b.offset, 0, b.size);
				else v = 0;
			}
			if(!v) *valid = 0;
		}
	};
	task->done = [cb_ref, buf_ref, idx, valid](napi_env env) {
		napi_value i, ok, buf;
		napi_create_int32(env, idx, &i);
		napi_get_boolean(env, *valid != 0, &ok);
		napi_get_reference_value(env, buf_ref, &buf);
		delete valid;
		call_ref(env, cb_ref, {i, ok, buf});
		napi_delete_reference(env, cb_ref);
		napi_delete_reference(env, buf_ref);
	};
	g->post(env, task);
	return nullptr;
}

napi_value FreeMem(napi_env env, napi_callback_info info) {
	size_t argc = 0;
	GfProc* g = unwrap(env, info, &argc, nullptr);
	if(!g) return nullptr;
	if(g->running) {
		napi_throw_error(env, nullptr, "Cannot free memory whilst running");
		return nullptr;
	}
	std::lock_guard<std::mutex> lk(g->proc_mu);
	for(auto& b : g->be) pph_proc_free_processing_mem(b.p);
	g->has_output = false;
	return nullptr;
}

napi_value GetInfo(napi_env env, napi_callback_info info) {
	size_t argc = 0;
	GfProc* g = unwrap(env, info, &argc, nullptr);
	if(!g) return nullptr;
	if(!g->n_out) {
		napi_throw_error(env, nullptr, "setRecoverySlices not yet called");
		return nullptr;
	}
	napi_value ret, devs;
	NAPI_OK(napi_create_object(env, &ret));
	NAPI_OK(napi_create_array_with_length(env, g->be.size(), &devs));
	std::lock_guard<std::mutex> lk(g->proc_mu);
	for(size_t i = 0; i < g->be.size(); i++) {
		const Backend& b = g->be[i];
		char name[256] = {};
		int cus = 0;
		pph_device_info(b.device, name, sizeof(name), &cus);
		napi_value o, v;
		NAPI_OK(napi_create_object(env, &o));
		NAPI_OK(napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &v));
		NAPI_OK(napi_set_named_property(env, o, "device_name", v));
		NAPI_OK(napi_create_uint32(env, (uint32_t)cus, &v));
		NAPI_OK(napi_set_named_property(env, o, "compute_units", v));
		NAPI_OK(napi_create_uint32(env, pph_proc_input_batch_size(b.p), &v));
		NAPI_OK(napi_set_named_property(env, o, "staging_size", v));
		NAPI_OK(napi_create_double(env, (double)b.offset, &v));
		NAPI_OK(napi_set_named_property(env, o, "slice_offset", v));
		NAPI_OK(napi_create_double(env, (double)b.size, &v));
		NAPI_OK(napi_set_named_property(env, o, "slice_size", v));
		NAPI_OK(napi_create_uint32(env, (uint32_t)pph_proc_num_recovery_slices(b.p), &v));
		NAPI_OK(napi_set_named_property(env, o, "num_output_slices", v));
		NAPI_OK(napi_set_element(env, devs, (uint32_t)i, o));
	}
	NAPI_OK(napi_set_named_property(env, ret, "hip_devices", devs));
	return ret;
}

// close([cb]) (src/gf.cc:402-424)
napi_value Close(napi_env env, napi_callback_info info) {
	size_t argc = 1;
	napi_value argv[1], self;
	GfProc* g = unwrap(env, info, &argc, argv, &self);
	if(!g) return nullptr;
	if(g->running) {
		napi_throw_error(env, nullptr, "Cannot close whilst running");
		return nullptr;
	}
	g->closed = true;
	auto* task = new Task;
	NAPI_OK(napi_create_reference(env, self, 1, &task->owner));
	task->work = [g] { g->destroy_procs(); };
	if(argc >= 1 && !is_nullish(env, argv[0])) {
		napi_ref cb_ref;
		NAPI_OK(napi_create_reference(env, argv[0], 1, &cb_ref));
		task->done = [cb_ref](napi_env env) {
			call_ref(env, cb_ref, {});
			napi_delete_reference(env, cb_ref);
		};
	}
	g->post(env, task);
	return nullptr;
}

// hip_device_info(device) and hip_devices(): the fields of ParPar's OpenCL
// device info (src/gf.cc:759-787) that its planner reads (lib/par2gen.js:
// 498-516), for HIP devices.
napi_value device_object(napi_env env, int device) {
	char name[256] = {};
	int cus = 0;
	size_t mem = 0;
	int rc = pph_device_info(device, name, sizeof(name), &cus);
	if(!rc) rc = pph_device_memory(device, &mem);
	if(rc) {
		napi_throw_error(env, nullptr, ("Unable to obtain HIP device info: " + status_text(rc)).c_str());
		return nullptr;
	}
	napi_value o, v;
	NAPI_OK(napi_create_object(env, &o));
	NAPI_OK(napi_create_int32(env, device, &v));
	NAPI_OK(napi_set_named_property(env, o, "id", v));
	NAPI_OK(napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &v));
	NAPI_OK(napi_set_named_property(env, o, "name", v));
	NAPI_OK(napi_create_double(env, (double)mem, &v));
	NAPI_OK(napi_set_named_property(env, o, "memory_global", v));
	NAPI_OK(napi_create_uint32(env, (uint32_t)cus, &v));
	NAPI_OK(napi_set_named_property(env, o, "compute_units", v));
	return o;
}

napi_value HipDeviceInfo(napi_env env, napi_callback_info info) {
	size_t argc = 1;
	napi_value argv[1];
	NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
	int32_t device = 0;
	if(argc >= 1 && !is_nullish(env, argv[0])) NAPI_OK(napi_get_value_int32(env, argv[0], &device));
	const int n = pph_device_count();
	if(device < 0 || device >= n) {
		napi_throw_error(env, nullptr, "Invalid HIP device");
		return nullptr;
	}
	return device_object(env, device);
}

napi_value HipDevices(napi_env env, napi_callback_info) {
	const int n = std::max(0, pph_device_count());
	napi_value arr;
	NAPI_OK(napi_create_array_with_length(env, (size_t)n, &arr));
	for(int d = 0; d < n; d++) {
		napi_value o = device_object(env, d);
		if(!o) return nullptr;
		NAPI_OK(napi_set_element(env, arr, (uint32_t)d, o));
	}
	return arr;
}

napi_value Init(napi_env env, napi_value exports) {
	napi_property_descriptor methods[] = {
	    {"close", nullptr, Close, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"freeMem", nullptr, FreeMem, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"setRecoverySlices", nullptr, SetRecoverySlices, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"setCurrentSliceSize", nullptr, SetCurrentSliceSize, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"setNumThreads", nullptr, SetNumThreads, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"setProgressCb", nullptr, SetProgressCb, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"info", nullptr, GetInfo, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"add", nullptr, AddSlice, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"end", nullptr, EndInput, nullptr, nullptr, nullptr, napi_default, nullptr},
	    {"get", nullptr, GetOutputSlice, nullptr, nullptr, nullptr, napi_default, nullptr},
	};
	napi_value cls;
	NAPI_OK(napi_define_class(env, "GfProc", NAPI_AUTO_LENGTH, New, nullptr, sizeof(methods) / sizeof(methods[0]),
	                          methods, &cls));
	NAPI_OK(napi_set_named_property(env, exports, "GfProc", cls));
	napi_value fn;
	NAPI_OK(napi_create_function(env, "hip_device_info", NAPI_AUTO_LENGTH, HipDeviceInfo, nullptr, &fn));
	NAPI_OK(napi_set_named_property(env, exports, "hip_device_info", fn));
	NAPI_OK(napi_create_function(env, "hip_devices", NAPI_AUTO_LENGTH, HipDevices, nullptr, &fn));
	NAPI_OK(napi_set_named_property(env, exports, "hip_devices", fn));
	return exports;
}

}  // namespace

NAPI_MODULE(gfproc_hip, Init)
