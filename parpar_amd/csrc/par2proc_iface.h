// The backend plugin interface of ParPar's PAR2 processing controller, in its
// non-libuv (std::future) form: the abstract class IPAR2ProcBackend and the
// enums and staging base it relies on, with the names and semantics of
// gf16/controller.h:37-217.  It exists so PAR2ProcHIP builds on its own; inside
// a ParPar tree define PPH_USE_PARPAR_CONTROLLER and put gf16/ on the include
// path, and controller_hip.h uses ParPar's controller.h instead.
#ifndef PPH_PAR2PROC_IFACE_H
#define PPH_PAR2PROC_IFACE_H

#include <atomic>
#include <chrono>
#include <cstdint>
#include <future>
#include <vector>

#define FUTURE_RETURN_T std::future<void>
#define FUTURE_RETURN_BOOL_T std::future<bool>
#define IF_LIBUV(...)
#define IF_NOT_LIBUV(...) __VA_ARGS__

enum PAR2ProcBackendAddResult { PROC_ADD_OK, PROC_ADD_OK_BUSY, PROC_ADD_FULL, PROC_ADD_ALL_FULL };

// a staging area is "active" while the batch it holds is being processed
class IPAR2ProcStaging {
	std::promise<void> prom;
public:
	std::shared_future<void> promFuture;
	std::vector<uint16_t> procCoeffs;
	IPAR2ProcStaging() { promFuture = prom.get_future(); }
	bool getIsActive() const {
		return promFuture.wait_for(std::chrono::seconds::zero()) == std::future_status::timeout;
	}
	void setIsActive(bool v) {
		if(v) {
			promFuture.get();
			prom = std::promise<void>();
			promFuture = prom.get_future();
		} else if(getIsActive())
			prom.set_value();
	}
};

class IPAR2ProcBackend {
protected:
	std::vector<uint16_t> outputExponents;
	bool processingAdd = false;
	virtual void run_kernel(unsigned stagingArea, unsigned numInputs) = 0;
	unsigned currentStagingArea = 0, currentStagingInputs = 0;
	unsigned inputBatchSize = 0, minInBatchSize = 0;
	unsigned statBatchesStarted = 0;
	std::atomic<unsigned> stagingActiveCount;
	static void _waitForAdd(IPAR2ProcStaging& area) { area.promFuture.get(); }
	void stagingActiveCount_inc() { stagingActiveCount.fetch_add(1, std::memory_order_relaxed); }
	unsigned stagingActiveCount_get() const { return stagingActiveCount.load(std::memory_order_relaxed); }
	virtual void _deinit() = 0;

public:
	IPAR2ProcBackend() : stagingActiveCount(0) {}
	void stagingActiveCount_dec() { stagingActiveCount.fetch_sub(1, std::memory_order_relaxed); }
	int getNumRecoverySlices() const { return (int)outputExponents.size(); }
	virtual void setSliceSize(size_t size) = 0;
	virtual bool setCurrentSliceSize(size_t size) = 0;
	virtual bool setRecoverySlices(unsigned numSlices, const uint16_t* exponents = nullptr) = 0;
	virtual PAR2ProcBackendAddResult canAdd() const = 0;
	virtual FUTURE_RETURN_T addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush) = 0;
	virtual FUTURE_RETURN_T addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush) = 0;
	virtual void dummyInput(uint16_t inputNum, bool flush = false) = 0;
	virtual bool fillInput(const void* buffer) = 0;
	virtual void flush() = 0;
	virtual FUTURE_RETURN_T endInput() = 0;
	bool isEmpty() const { return stagingActiveCount_get() == 0; }
	virtual FUTURE_RETURN_BOOL_T getOutput(unsigned index, void* output) = 0;
	void discardOutput() { processingAdd = false; }
	bool _hasAdded() const { return processingAdd; }
	virtual void processing_finished() {}
	virtual void waitForAdd() = 0;
	void deinit() { _deinit(); }
	virtual void freeProcessingMem() = 0;
	unsigned getInputBatchSize() const { return inputBatchSize; }
	unsigned getBatchesStarted() const { return statBatchesStarted; }
	void setMinInputBatchSize(unsigned size) { minInBatchSize = size ? size : inputBatchSize; }
	virtual ~IPAR2ProcBackend() {}
};

#endif
