// PAR2ProcHIP: a ParPar PAR2 processing backend for one AMD Instinct GPU,
// beside PAR2ProcCPU (gf16/controller_cpu.h) and PAR2ProcOCL
// (gf16/controller_ocl.h).  Host-side C++ only: every device operation goes
// through the C ABI of libparpar_hip.so (include/parpar_hip.h).
//
// Both forms of ParPar's backend interface are implemented:
//   - the std::future form (no USE_LIBUV), which builds on its own against a
//     restatement of the interface (par2proc_iface.h) or, with
//     PPH_USE_PARPAR_CONTROLLER, against ParPar's gf16/controller.h;
//   - the libuv form (USE_LIBUV, the form ParPar's Node binding is built in,
//     binding.gyp:47), which needs ParPar's controller.h (define
//     PPH_USE_PARPAR_CONTROLLER): callbacks are posted to the loop through the
//     interface's _queueSent / _queueRecv / _queueProc notify queues, as
//     PAR2ProcOCL posts them (gf16/controller_ocl.cpp:253-268,434-494).
#ifndef PPH_CONTROLLER_HIP_H
#define PPH_CONTROLLER_HIP_H

#ifdef PPH_USE_PARPAR_CONTROLLER
#include "controller.h"
#else
#ifdef USE_LIBUV
#error "the libuv form of PAR2ProcHIP builds against ParPar's gf16/controller.h (define PPH_USE_PARPAR_CONTROLLER)"
#endif
#include "par2proc_iface.h"
#endif
#include "parpar_hip.h"

#include <memory>
#include <string>
#include <vector>

class PAR2ProcHIP : public IPAR2ProcBackend {
	pph_proc* proc = nullptr;
	int deviceId;
	unsigned stagingAreas;
	size_t sliceSize = 0;
	unsigned cksumSize = 64;
	int form = PPH_FORM_AUTO;
	bool initSuccess = false;
	int lastStatus = PPH_OK;

	bool check(int status);

	void run_kernel(unsigned stagingArea, unsigned numInputs) override;
	void _deinit() override;

#ifdef USE_LIBUV
	// Staging state lives on the loop thread, as in PAR2ProcOCL
	// (controller_ocl.cpp:303-350): the area an add goes to, whether it starts
	// a batch, which areas have a batch in flight.  A transfer thread stages
	// the inputs and submits the batches; the backend library's worker thread
	// reports completed batches and outputs.
	std::vector<IPAR2ProcStaging> staging;
	std::unique_ptr<MessageThread> transferThread;
	static void transfer_main(ThreadMessageQueue<void*>& q);
	static void batch_done(void* req, int status);
	static void output_done(void* req, int status, int valid);
	template<typename T>
	void queue_input(int kind, const void* buffer, size_t size, T inputNumOrCoeffs, bool flush,
	                 const PAR2ProcPlainCb& cb);
	void begin_batch(struct hip_transfer_req* req);
	void _notifySent(void* req) override;
	void _notifyRecv(void* req) override;
	void _notifyProc(void* req) override;
#else
	void sync_in();   // propagate discardOutput() to the device side
	void sync_out();  // mirror the device side's state into the base members
#endif

	PAR2ProcHIP(const PAR2ProcHIP&) = delete;
	PAR2ProcHIP& operator=(const PAR2ProcHIP&) = delete;

public:
	// PAR2ProcOCL(platformId, deviceId, stagingAreas) (gf16/controller_ocl.cpp:27)
	explicit PAR2ProcHIP(IF_LIBUV(uv_loop_t* loop, ) int deviceId = 0, int stagingAreas = 2);
	~PAR2ProcHIP();
	// PAR2ProcOCL::init (gf16/controller_ocl.cpp:67): cksumSize is the checksum
	// width of the host CPU method (Galois16Mul::info().cksumSize).
	bool init(unsigned cksumSize = 64, unsigned inputBatch = 0, int form = PPH_FORM_AUTO);

	void setSliceSize(size_t size) override;
	bool setCurrentSliceSize(size_t size) override;
	bool setRecoverySlices(unsigned numSlices, const uint16_t* exponents = nullptr) override;
	PAR2ProcBackendAddResult canAdd() const override;
	FUTURE_RETURN_T addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush
	                         IF_LIBUV(, const PAR2ProcPlainCb& cb)) override;
	FUTURE_RETURN_T addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush
	                         IF_LIBUV(, const PAR2ProcPlainCb& cb)) override;
	void dummyInput(uint16_t inputNum, bool flush = false) override;
	bool fillInput(const void* buffer) override;
	void flush() override;
	FUTURE_RETURN_BOOL_T getOutput(unsigned index, void* output IF_LIBUV(, const PAR2ProcOutputCb& cb)) override;
	void freeProcessingMem() override;
#ifdef USE_LIBUV
	void processing_finished() override;
#else
	FUTURE_RETURN_T endInput() override;
	void waitForAdd() override;
#endif

	// Test hook: the next batch submission fails with `status` (pph_proc_debug_fail_next_batch).
	int debugFailNextBatch(int status) { return proc ? pph_proc_debug_fail_next_batch(proc, status) : PPH_E_STATE; }
	bool isInitSuccess() const { return initSuccess; }
	int getLastStatus() const { return lastStatus; }
	std::string getDeviceName() const;
	static int getNumDevices() { return pph_device_count(); }
};

#endif
