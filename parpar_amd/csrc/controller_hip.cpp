// PAR2ProcHIP: IPAR2ProcBackend over the libparpar_hip C ABI.
//
// Semantics follow PAR2ProcOCL (gf16/controller_ocl.cpp): inputs are staged
// into device slots with their gf16_cksum, batches of inputBatchSize inputs
// are submitted as one mul-add launch (the first overwrites the recovery
// slices, later ones accumulate: processingAdd), and getOutput checks the
// checksum of each recovery slice as it is copied out.
//
// std::future form: an add's future is ready as soon as the caller's buffer
// has been copied (controller_ocl.cpp:296 resolves it after the map/copy);
// getOutput's future carries the checksum result.
//
// libuv form: the loop thread decides batch boundaries and keeps the areas'
// isActive flags (controller_ocl.cpp:303-350,405-426); a transfer thread
// copies inputs into the staging and submits batches; completions come from
// the library's worker thread and are posted to the loop through the notify
// queues.  Callbacks run on the loop thread: an add's once its buffer has
// been copied (_notifySent), the progress callback once a batch's mul-add has
// completed (_notifyProc), a getOutput's with the checksum result
// (_notifyRecv).
#include "controller_hip.h"

#include <cassert>
#include <cstring>

PAR2ProcHIP::PAR2ProcHIP(IF_LIBUV(uv_loop_t* _loop, ) int _deviceId, int _stagingAreas)
: IPAR2ProcBackend(IF_LIBUV(_loop)), deviceId(_deviceId), stagingAreas(_stagingAreas > 0 ? (unsigned)_stagingAreas : 2) {
	IF_NOT_LIBUV(processingAdd = false);
	initSuccess = check(pph_proc_create(deviceId, stagingAreas, &proc));
#ifdef USE_LIBUV
	processingAdd = false;
	currentStagingArea = currentStagingInputs = 0;
	inputBatchSize = minInBatchSize = 0;
	statBatchesStarted = 0;
	staging.resize(stagingAreas);
	for(auto& area : staging) area.setIsActive(false);
#endif
}

PAR2ProcHIP::~PAR2ProcHIP() {
	deinit();
}

void PAR2ProcHIP::_deinit() {
#ifdef USE_LIBUV
	// staged work is submitted first, then every completion is posted
	transferThread.reset();
	if(proc) pph_proc_drain(proc);
#endif
	if(proc) {
		pph_proc_destroy(proc);
		proc = nullptr;
	}
}

bool PAR2ProcHIP::check(int status) {
	lastStatus = status;
	return status == PPH_OK;
}

bool PAR2ProcHIP::init(unsigned _cksumSize, unsigned inputBatch, int _form) {
	if(!proc) return false;
	cksumSize = _cksumSize;
	form = _form;
	outputExponents.clear();
	if(!check(pph_proc_set_slice_size(proc, sliceSize))) return false;
	if(!check(pph_proc_init(proc, cksumSize, inputBatch, form))) return false;
	inputBatchSize = pph_proc_input_batch_size(proc);
	minInBatchSize = inputBatchSize;
	statBatchesStarted = 0;
	processingAdd = false;
#ifdef USE_LIBUV
	currentStagingArea = currentStagingInputs = 0;
#endif
	return true;
}

void PAR2ProcHIP::setSliceSize(size_t size) {
	sliceSize = size;
	if(proc) check(pph_proc_set_slice_size(proc, size));
}

bool PAR2ProcHIP::setCurrentSliceSize(size_t size) {
	if(!proc) return false;
	const bool ok = check(pph_proc_set_current_slice_size(proc, size));
	// a grown slice re-initialises the backend and starts the recovery afresh
	// (PAR2ProcOCL::setCurrentSliceSize, controller_ocl.cpp:211-222)
	if(ok && !pph_proc_has_added(proc)) processingAdd = false;
	return ok;
}

bool PAR2ProcHIP::setRecoverySlices(unsigned numSlices, const uint16_t* exponents) {
	if(!proc) return false;
	if(!check(pph_proc_set_recovery_slices(proc, numSlices, exponents))) return false;
	outputExponents.assign(numSlices, 1);
	if(exponents) outputExponents.assign(exponents, exponents + numSlices);
	return true;
}

std::string PAR2ProcHIP::getDeviceName() const {
	char name[256] = {0};
	if(pph_device_info(deviceId, name, sizeof(name), nullptr) != PPH_OK) return std::string();
	return std::string(name);
}

void PAR2ProcHIP::freeProcessingMem() {
	if(proc) check(pph_proc_free_processing_mem(proc));
	outputExponents.clear();
	processingAdd = false;
}

#ifndef USE_LIBUV
// ============================================================ std::future form

static std::future<void> ready_future() {
	std::promise<void> p;
	p.set_value();
	return p.get_future();
}
static std::future<bool> ready_future(bool v) {
	std::promise<bool> p;
	p.set_value(v);
	return p.get_future();
}

void PAR2ProcHIP::sync_in() {
	if(proc && !processingAdd && pph_proc_has_added(proc)) pph_proc_discard_output(proc);
}
void PAR2ProcHIP::sync_out() {
	if(!proc) return;
	processingAdd = pph_proc_has_added(proc) != 0;
	statBatchesStarted = pph_proc_batches_started(proc);
	inputBatchSize = pph_proc_input_batch_size(proc);
	const unsigned active = pph_proc_is_empty(proc) ? 0u : 1u;
	stagingActiveCount.store(active, std::memory_order_relaxed);
}

PAR2ProcBackendAddResult PAR2ProcHIP::canAdd() const {
	switch(pph_proc_can_add(proc)) {
		case PPH_ADD_OK: return PROC_ADD_OK;
		case PPH_ADD_OK_BUSY: return PROC_ADD_OK_BUSY;
		default: return PROC_ADD_FULL;
	}
}

void PAR2ProcHIP::waitForAdd() {
	if(proc) check(pph_proc_wait_for_add(proc));
}

FUTURE_RETURN_T PAR2ProcHIP::addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush) {
	sync_in();
	pph_proc_set_min_input_batch_size(proc, minInBatchSize);
	check(pph_proc_add_input(proc, buffer, size, inputNum, flush ? 1 : 0));
	sync_out();
	return ready_future();
}

FUTURE_RETURN_T PAR2ProcHIP::addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush) {
	sync_in();
	pph_proc_set_min_input_batch_size(proc, minInBatchSize);
	check(pph_proc_add_input_coeffs(proc, buffer, size, coeffs, flush ? 1 : 0));
	sync_out();
	return ready_future();
}

void PAR2ProcHIP::dummyInput(uint16_t inputNum, bool flush) {
	sync_in();
	check(pph_proc_dummy_input(proc, inputNum, flush ? 1 : 0));
	sync_out();
}

bool PAR2ProcHIP::fillInput(const void* buffer) {
	int rc = pph_proc_fill_input(proc, buffer);
	if(rc < 0) {
		check(rc);
		return false;
	}
	return rc == 1;
}

void PAR2ProcHIP::flush() {
	sync_in();
	check(pph_proc_flush(proc));
	sync_out();
}

void PAR2ProcHIP::run_kernel(unsigned, unsigned) {
	// batches are submitted by the device layer as they fill; an explicit
	// request flushes the current one
	flush();
}

FUTURE_RETURN_T PAR2ProcHIP::endInput() {
	sync_in();
	check(pph_proc_end_input(proc));
	sync_out();
	processing_finished();
	return ready_future();
}

FUTURE_RETURN_BOOL_T PAR2ProcHIP::getOutput(unsigned index, void* output) {
	// a failed batch submission makes every getOutput report the slice invalid
	int valid = 0;
	if(!check(pph_proc_get_output(proc, index, output, &valid))) return ready_future(false);
	return ready_future(valid != 0);
}

#else
// ============================================================ libuv form

enum { HIP_REQ_STAGE, HIP_REQ_DUMMY, HIP_REQ_SUBMIT };

// One request to the transfer thread, posted back to the loop through
// _queueSent once done.
struct hip_transfer_req {
	int kind;
	PAR2ProcHIP* parent;
	pph_proc* proc;
	const void* buffer;
	size_t size;
	uint16_t inputNum;
	bool hasCoeffs;
	std::vector<uint16_t> coeffs;
	unsigned area, slot;
	unsigned submitInputs;  // > 0: submit the area's first submitInputs slots after staging
	bool first;             // the batch starts a new recovery computation
	PAR2ProcPlainCb cb;
	int status;
};
// A submitted batch, posted through _queueProc once its mul-add completed.
struct hip_compute_req {
	PAR2ProcHIP* parent;
	unsigned area, numInputs;
	int status;
};
// A getOutput, posted through _queueRecv with its checksum result.
struct hip_output_req {
	PAR2ProcHIP* parent;
	PAR2ProcOutputCb cb;
	int status, valid;
};

static void set_req_coeffs(hip_transfer_req* req, uint16_t inputNum, size_t) {
	req->inputNum = inputNum;
	req->hasCoeffs = false;
}
static void set_req_coeffs(hip_transfer_req* req, const uint16_t* coeffs, size_t n) {
	// copied now: the caller may reuse its coefficient list once addInput returns
	req->inputNum = 0;
	req->hasCoeffs = true;
	req->coeffs.assign(coeffs, coeffs + n);
}

void PAR2ProcHIP::batch_done(void* _req, int status) {
	// the library's worker thread: only post to the loop
	auto* req = static_cast<hip_compute_req*>(_req);
	req->status = status;
	req->parent->_queueProc.notify(req);
}

void PAR2ProcHIP::output_done(void* _req, int status, int valid) {
	auto* req = static_cast<hip_output_req*>(_req);
	req->status = status;
	req->valid = valid;
	req->parent->_queueRecv.notify(req);
}

void PAR2ProcHIP::transfer_main(ThreadMessageQueue<void*>& q) {
	hip_transfer_req* req;
	while((req = static_cast<hip_transfer_req*>(q.pop())) != nullptr) {
		int rc = PPH_OK;
		if(req->kind == HIP_REQ_STAGE)
			rc = pph_proc_stage_input(req->proc, req->area, req->slot, req->buffer, req->size, req->inputNum,
			                          req->hasCoeffs ? req->coeffs.data() : nullptr);
		else if(req->kind == HIP_REQ_DUMMY)
			rc = pph_proc_stage_dummy(req->proc, req->area, req->slot, req->inputNum);
		if(req->submitInputs) {
			auto* creq = new hip_compute_req{req->parent, req->area, req->submitInputs, PPH_OK};
			// batch_done runs exactly once, also for a failed submission
			const int src = pph_proc_submit(req->proc, req->area, req->submitInputs, req->first ? 1 : 0,
			                                &PAR2ProcHIP::batch_done, creq);
			if(rc == PPH_OK) rc = src;
		}
		req->status = rc;
		req->parent->_queueSent.notify(req);
	}
}

void PAR2ProcHIP::_notifySent(void* _req) {
	auto* req = static_cast<hip_transfer_req*>(_req);
	pendingInCallbacks--;
	if(req->status != PPH_OK) check(req->status);
	if(req->cb && req->buffer) req->cb();
	delete req;
	// the last add's callback may come after the last batch completed
	if(endSignalled && isEmpty() && progressCb) progressCb(0);
}

void PAR2ProcHIP::_notifyProc(void* _req) {
	auto* req = static_cast<hip_compute_req*>(_req);
	if(req->status != PPH_OK) check(req->status);
	stagingActiveCount_dec();
	staging[req->area].setIsActive(false);
	// an add refused while the area was busy can now be retried
	if(progressCb) progressCb(req->numInputs);
	delete req;
}

void PAR2ProcHIP::_notifyRecv(void* _req) {
	auto* req = static_cast<hip_output_req*>(_req);
	pendingOutCallbacks--;
	if(req->status != PPH_OK) check(req->status);
	req->cb(req->status == PPH_OK && req->valid != 0);
	delete req;
	if(pendingOutCallbacks < 1 && deinitCallback) deinit(deinitCallback);
}

PAR2ProcBackendAddResult PAR2ProcHIP::canAdd() const {
	if(staging[currentStagingArea].getIsActive()) return PROC_ADD_FULL;
	return stagingActiveCount_get() < staging.size() - 1 ? PROC_ADD_OK : PROC_ADD_OK_BUSY;
}

// The area's batch starts: lock the area until its completion is posted.
void PAR2ProcHIP::begin_batch(hip_transfer_req* req) {
	req->first = !processingAdd;
	processingAdd = true;
	stagingActiveCount_inc();
	staging[currentStagingArea].setIsActive(true);
	statBatchesStarted++;
	currentStagingInputs = 0;
	if(++currentStagingArea == staging.size()) currentStagingArea = 0;
}

template<typename T>
void PAR2ProcHIP::queue_input(int kind, const void* buffer, size_t size, T inputNumOrCoeffs, bool flush,
                              const PAR2ProcPlainCb& cb) {
	assert(!staging[currentStagingArea].getIsActive());
	auto* req = new hip_transfer_req();
	req->kind = kind;
	req->parent = this;
	req->proc = proc;
	req->buffer = buffer;
	req->size = size;
	set_req_coeffs(req, inputNumOrCoeffs, outputExponents.size());
	req->area = currentStagingArea;
	req->slot = currentStagingInputs;
	req->cb = cb;
	req->status = PPH_OK;
	currentStagingInputs++;
	req->submitInputs = (flush || currentStagingInputs == inputBatchSize ||
	                     // submit early while nothing is being processed
	                     (stagingActiveCount_get() == 0 && staging.size() > 1 && currentStagingInputs >= minInBatchSize))
	                        ? currentStagingInputs
	                        : 0;
	if(req->submitInputs) begin_batch(req);
	pendingInCallbacks++;
	if(!transferThread) {
		transferThread.reset(new MessageThread(&PAR2ProcHIP::transfer_main));
		transferThread->name = "hip_transfer";
	}
	transferThread->send(req);
}

void PAR2ProcHIP::addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush, const PAR2ProcPlainCb& cb) {
	queue_input(HIP_REQ_STAGE, buffer, size, inputNum, flush, cb);
}

void PAR2ProcHIP::addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush,
                           const PAR2ProcPlainCb& cb) {
	queue_input(HIP_REQ_STAGE, buffer, size, coeffs, flush, cb);
}

void PAR2ProcHIP::dummyInput(uint16_t inputNum, bool flush) {
	queue_input(HIP_REQ_DUMMY, nullptr, 0, inputNum, flush, nullptr);
}

bool PAR2ProcHIP::fillInput(const void* buffer) {
	// benchmark set-up before any add (PAR2ProcOCL::fillInput, controller_ocl.cpp:372-385)
	if(!check(pph_proc_fill_slot(proc, currentStagingArea, currentStagingInputs, buffer))) return false;
	if(++currentStagingInputs == inputBatchSize) {
		currentStagingInputs = 0;
		if(++currentStagingArea == staging.size()) {
			currentStagingArea = 0;
			return true;
		}
	}
	return false;
}

void PAR2ProcHIP::flush() {
	if(!currentStagingInputs) return;
	auto* req = new hip_transfer_req();
	req->kind = HIP_REQ_SUBMIT;
	req->parent = this;
	req->proc = proc;
	req->buffer = nullptr;
	req->area = currentStagingArea;
	req->submitInputs = currentStagingInputs;
	req->status = PPH_OK;
	begin_batch(req);
	pendingInCallbacks++;
	if(!transferThread) {
		transferThread.reset(new MessageThread(&PAR2ProcHIP::transfer_main));
		transferThread->name = "hip_transfer";
	}
	transferThread->send(req);
}

void PAR2ProcHIP::run_kernel(unsigned, unsigned) {
	// batches are submitted by the transfer thread as the loop thread decides
	// (queue_input, flush); an explicit request flushes the current one
	flush();
}

void PAR2ProcHIP::processing_finished() {
	endSignalled = false;
}

void PAR2ProcHIP::getOutput(unsigned index, void* output, const PAR2ProcOutputCb& cb) {
	auto* req = new hip_output_req{this, cb, PPH_OK, 0};
	pendingOutCallbacks++;
	const int rc = pph_proc_get_output_async(proc, index, output, &PAR2ProcHIP::output_done, req);
	if(rc != PPH_OK) output_done(req, rc, 0);
}
#endif
