/*
 * Test infrastructure: a C-ABI harness around the reference's own CPU path.
 *
 * This file is compiled together with the reference's gf16/ sources, where
 * they lie (see oracle/Makefile), into oracle/_ref/libparpar_ref.so.  It drives
 * PAR2Proc + PAR2ProcCPU exactly as the non-libuv C++ API is used
 * (gf16/controller.h:248-326, gf16/controller_cpu.h, src/gf.cc:697-727 for the
 * init order) and exposes the result to Python so the golden fixtures under
 * tests/golden/ can be written and the C restatement in oracle/ can be pinned
 * against the reference.  Nothing here is shipped or measured as the product.
 */
#include "controller.h"
#include "controller_cpu.h"
#include "gfmat_coeff.h"
#include "gf16mul.h"

#include <chrono>
#include <cstring>
#include <vector>
#include <memory>

extern "C" {

/* Name and checksum width of the method GF16_AUTO selects on this host
 * (gf16/gf16mul.cpp:1221-1275, width from :215-409). */
int ref_info(char* name, size_t nameLen, unsigned* cksumSize, unsigned* threads) {
	Galois16Mul gf(Galois16Mul::default_method());
	const Galois16MethodInfo& info = gf.info();
	if(name && nameLen) {
		strncpy(name, info.name, nameLen - 1);
		name[nameLen - 1] = 0;
	}
	if(cksumSize) *cksumSize = info.cksumSize;
	if(threads) *threads = hardware_concurrency();
	return (int)info.id;
}

uint16_t ref_gfmat_input_log(uint16_t input) {
	gfmat_init();
	return gfmat_input_log(input);
}
uint16_t ref_gfmat_coeff(uint16_t input, uint16_t exp) {
	gfmat_init();
	return gfmat_coeff(input, exp);
}
uint16_t ref_gf16_exp(uint16_t v) {
	gfmat_init();
	return gf16_exp(v);
}

/* Galois16Mul::copy_cksum / copy_cksum_check of a given method (0 = AUTO),
 * i.e. gf16_cksum_copy / gf16_cksum_copy_check (gf16/gf16_cksum_base.h:13-59).
 * dst must hold sliceLen + cksumSize bytes.  Returns the checksum width used. */
int ref_cksum_copy(int method, void* dst, const void* src, size_t srcLen, size_t sliceLen) {
	gfmat_init();
	Galois16Mul gf((Galois16Methods)method);
	gf.copy_cksum(dst, src, srcLen, sliceLen);
	return (int)gf.info().cksumSize;
}
int ref_cksum_check(int method, void* dst, const void* src, size_t len) {
	gfmat_init();
	Galois16Mul gf((Galois16Methods)method);
	return gf.copy_cksum_check(dst, src, len);
}
int ref_method_cksum_size(int method) {
	Galois16Mul gf((Galois16Methods)method);
	return (int)gf.info().cksumSize;
}

/*
 * Full recovery through PAR2Proc + PAR2ProcCPU (GF16_AUTO or a given method).
 *   inputs   : nIn slices, input i at inputs + i*inStride, inLens[i] bytes (<= sliceSize)
 *   inputNums: PAR2 input numbers (used when coeffs == NULL)
 *   coeffs   : optional explicit coefficients, coeffs[i*nOut + r] (controller.h:170)
 *   exps     : nOut recovery exponents
 *   split    : 0 = one backend covering the slice; otherwise two CPU backends
 *              covering [0, split) and [split, sliceSize) (controller.cpp:13-84)
 *   outputs  : nOut * sliceSize bytes; valid[r] = getOutput()'s checksum flag
 *   seconds  : time from the first addInput to the last getOutput resolving
 */
int ref_recover(int method, unsigned nIn, const uint8_t* inputs, size_t inStride, const size_t* inLens,
                const uint16_t* inputNums, const uint16_t* coeffs, unsigned nOut, const uint16_t* exps,
                size_t sliceSize, size_t split, uint8_t* outputs, int* valid, int threads, double* seconds) {
	PAR2Proc proc;
	std::vector<std::unique_ptr<PAR2ProcCPU>> cpus;
	std::vector<PAR2ProcBackendAlloc> allocs;
	if(split && split < sliceSize) {
		cpus.emplace_back(new PAR2ProcCPU());
		cpus.emplace_back(new PAR2ProcCPU());
		allocs.push_back({cpus[0].get(), 0, split});
		allocs.push_back({cpus[1].get(), split, sliceSize - split});
	} else {
		cpus.emplace_back(new PAR2ProcCPU());
		allocs.push_back({cpus[0].get(), 0, sliceSize});
	}
	if(!proc.init(sliceSize, allocs)) return -1;
	for(auto& cpu : cpus) {
		if(threads > 0) cpu->setNumThreads(threads);
		if(!cpu->init((Galois16Methods)method)) return -2;
	}
	if(!proc.setRecoverySlices(nOut, exps)) return -3;

	std::vector<uint16_t> colCoeffs(nOut);
	auto t0 = std::chrono::steady_clock::now();
	std::vector<std::future<void>> pending;
	for(unsigned i = 0; i < nIn; i++) {
		PAR2ProcBackendAddResult st = proc.canAdd();
		if(st == PROC_ADD_FULL || st == PROC_ADD_ALL_FULL)
			proc.waitForAdd();
		const uint8_t* src = inputs + (size_t)i * inStride;
		if(coeffs) {
			memcpy(colCoeffs.data(), coeffs + (size_t)i * nOut, nOut * sizeof(uint16_t));
			pending.push_back(proc.addInput(src, inLens[i], colCoeffs.data(), false));
			// the coefficient array is read at add time (controller_cpu.cpp:380-385)
		} else {
			pending.push_back(proc.addInput(src, inLens[i], inputNums[i], false));
		}
	}
	for(auto& f : pending) f.get();
	proc.endInput().get();
	for(unsigned r = 0; r < nOut; r++) {
		bool ok = proc.getOutput(r, outputs + (size_t)r * sliceSize).get();
		if(valid) valid[r] = ok ? 1 : 0;
	}
	auto t1 = std::chrono::steady_clock::now();
	if(seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
	proc.deinit();
	return 0;
}

} // extern "C"
