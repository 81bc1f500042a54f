// PAR2ProcHIP: a ParPar PAR2 processing backend for one AMD Instinct GPU,
// beside PAR2ProcCPU (gf16/controller_cpu.h) and PAR2ProcOCL
// (gf16/controller_ocl.h).  Host-side C++ only: every device operation goes
// through the C ABI of libparpar_hip.so (include/parpar_hip.h).
#ifndef PPH_CONTROLLER_HIP_H
#define PPH_CONTROLLER_HIP_H

#ifdef PPH_USE_PARPAR_CONTROLLER
#include "controller.h"
#else
#include "par2proc_iface.h"
#endif
#include "parpar_hip.h"

#include <string>

class PAR2ProcHIP : public IPAR2ProcBackend {
	pph_proc* proc = nullptr;
	int deviceId;
	unsigned stagingAreas;
	size_t sliceSize = 0;
	unsigned cksumSize = 64;
	int form = PPH_FORM_AUTO;
	bool initSuccess = false;
	int lastStatus = PPH_OK;

	void sync_in();   // propagate discardOutput() to the device side
	void sync_out();  // mirror the device side's state into the base members
	bool check(int status);

	void run_kernel(unsigned stagingArea, unsigned numInputs) override;
	void _deinit() override;

	PAR2ProcHIP(const PAR2ProcHIP&) = delete;
	PAR2ProcHIP& operator=(const PAR2ProcHIP&) = delete;

public:
	// PAR2ProcOCL(platformId, deviceId, stagingAreas) (gf16/controller_ocl.cpp:27)
	explicit PAR2ProcHIP(int deviceId = 0, int stagingAreas = 2);
	~PAR2ProcHIP();
	// PAR2ProcOCL::init (gf16/controller_ocl.cpp:67): cksumSize is the checksum
	// width of the host CPU method (Galois16Mul::info().cksumSize).
	bool init(unsigned cksumSize = 64, unsigned inputBatch = 0, int form = PPH_FORM_AUTO);

	void setSliceSize(size_t size) override;
	bool setCurrentSliceSize(size_t size) override;
	bool setRecoverySlices(unsigned numSlices, const uint16_t* exponents = nullptr) override;
	PAR2ProcBackendAddResult canAdd() const override;
	FUTURE_RETURN_T addInput(const void* buffer, size_t size, uint16_t inputNum, bool flush) override;
	FUTURE_RETURN_T addInput(const void* buffer, size_t size, const uint16_t* coeffs, bool flush) override;
	void dummyInput(uint16_t inputNum, bool flush = false) override;
	bool fillInput(const void* buffer) override;
	void flush() override;
	FUTURE_RETURN_T endInput() override;
	FUTURE_RETURN_BOOL_T getOutput(unsigned index, void* output) override;
	void waitForAdd() override;
	void freeProcessingMem() override;

	bool isInitSuccess() const { return initSuccess; }
	int getLastStatus() const { return lastStatus; }
	std::string getDeviceName() const;
	static int getNumDevices() { return pph_device_count(); }
};

#endif
