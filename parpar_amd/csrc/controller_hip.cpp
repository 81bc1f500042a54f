// PAR2ProcHIP: IPAR2ProcBackend over the libparpar_hip C ABI.
//
// Semantics follow PAR2ProcOCL (gf16/controller_ocl.cpp): inputs are staged
// into device slots with their gf16_cksum, batches of inputBatchSize inputs
// are submitted as one mul-add launch (the first overwrites the recovery
// slices, later ones accumulate: processingAdd), and getOutput checks the
// checksum of each recovery slice.  An add's future is ready as soon as the
// caller's buffer has been copied (controller_ocl.cpp:296 resolves it after
// the map/copy); getOutput's future carries the checksum result.
#include "controller_hip.h"

#include <cstring>

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

PAR2ProcHIP::PAR2ProcHIP(int _deviceId, int _stagingAreas)
: IPAR2ProcBackend(), deviceId(_deviceId), stagingAreas(_stagingAreas > 0 ? (unsigned)_stagingAreas : 2) {
	initSuccess = check(pph_proc_create(deviceId, stagingAreas, &proc));
}

PAR2ProcHIP::~PAR2ProcHIP() {
	deinit();
}

void PAR2ProcHIP::_deinit() {
	if(proc) {
		pph_proc_destroy(proc);
		proc = nullptr;
	}
}

bool PAR2ProcHIP::check(int status) {
	lastStatus = status;
	return status == PPH_OK;
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
	return true;
}

void PAR2ProcHIP::setSliceSize(size_t size) {
	sliceSize = size;
	if(proc) check(pph_proc_set_slice_size(proc, size));
}

bool PAR2ProcHIP::setCurrentSliceSize(size_t size) {
	if(!proc) return false;
	return check(pph_proc_set_current_slice_size(proc, size));
}

bool PAR2ProcHIP::setRecoverySlices(unsigned numSlices, const uint16_t* exponents) {
	if(!proc) return false;
	if(!check(pph_proc_set_recovery_slices(proc, numSlices, exponents))) return false;
	outputExponents.assign(numSlices, 1);
	if(exponents) outputExponents.assign(exponents, exponents + numSlices);
	return true;
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
	int valid = 0;
	if(!check(pph_proc_get_output(proc, index, output, &valid))) return ready_future(false);
	return ready_future(valid != 0);
}

void PAR2ProcHIP::freeProcessingMem() {
	if(proc) check(pph_proc_free_processing_mem(proc));
	outputExponents.clear();
	processingAdd = false;
}

std::string PAR2ProcHIP::getDeviceName() const {
	char name[256] = {0};
	if(pph_device_info(deviceId, name, sizeof(name), nullptr) != PPH_OK) return std::string();
	return std::string(name);
}
