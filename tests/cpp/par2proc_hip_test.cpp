// Drives the C++ PAR2ProcHIP backend the way ParPar's controller drives a
// backend in its non-libuv build (gf16/controller.cpp:243-399): futures from
// addInput / endInput / getOutput, checked against the oracle.  Prints
// "ALL OK" and exits 0 on success.
#include "controller_hip.h"
#include "gf16_oracle.h"
#include "po_prng.h"

#include <cstdio>
#include <cstring>
#include <vector>

static int fail(const char* what) {
	printf("FAIL: %s\n", what);
	fflush(stdout);
	return 1;
}

static int run(size_t S, unsigned n, const std::vector<uint16_t>& exps, unsigned batch, unsigned areas, unsigned ck,
               bool explicitHalf) {
	printf("case S=%zu n=%u R=%zu batch=%u areas=%u ck=%u explicit=%d\n", S, n, exps.size(), batch, areas, ck,
	       (int)explicitHalf);
	fflush(stdout);
	const unsigned R = (unsigned)exps.size();
	std::vector<uint8_t> in((size_t)n * S);
	std::vector<size_t> lens(n, S);
	lens[n - 1] = S / 3 + 1;
	for(unsigned i = 0; i < n; i++) po_prng_fill(77, 1, i, 0, &in[(size_t)i * S], lens[i]);
	for(unsigned i = 0; i < n; i++) memset(&in[(size_t)i * S + lens[i]], 0, S - lens[i]);
	std::vector<uint16_t> nums(n), coeffs((size_t)n * R);
	for(unsigned i = 0; i < n; i++) {
		nums[i] = (uint16_t)(i * 7 % 32768);
		for(unsigned r = 0; r < R; r++) coeffs[(size_t)i * R + r] = po_coeff(nums[i], exps[r]);
	}

	PAR2ProcHIP be(0, (int)areas);
	if(!be.isInitSuccess()) return fail("create");
	be.setSliceSize(S);
	if(!be.init(ck, batch)) return fail("init");
	if(!be.setRecoverySlices(R, exps.data())) return fail("setRecoverySlices");
	for(unsigned i = 0; i < n; i++) {
		if(be.canAdd() == PROC_ADD_FULL) be.waitForAdd();
		std::future<void> f;
		if(explicitHalf && (i & 1))
			f = be.addInput(&in[(size_t)i * S], lens[i], &coeffs[(size_t)i * R], false);
		else
			f = be.addInput(&in[(size_t)i * S], lens[i], nums[i], false);
		f.get();
		if(be.getLastStatus() != PPH_OK) return fail("addInput");
	}
	be.endInput().get();
	if(!be._hasAdded()) return fail("processingAdd");

	std::vector<uint8_t> ref((size_t)R * S), out(S);
	po_recover(n, in.data(), S, lens.data(), nums.data(), nullptr, R, exps.data(), S, ref.data(), 0);
	for(unsigned r = 0; r < R; r++) {
		bool ok = be.getOutput(r, out.data()).get();
		if(!ok) return fail("checksum");
		if(memcmp(out.data(), &ref[(size_t)r * S], S)) return fail("output bytes");
	}
	return 0;
}

// A batch whose submission fails loses its inputs: every getOutput of that
// recovery computation must then report invalid, until a new one starts.
static int run_failed_batch() {
	printf("case failed batch\n");
	fflush(stdout);
	const size_t S = 4096;
	const unsigned n = 6;
	std::vector<uint8_t> in((size_t)n * S);
	for(unsigned i = 0; i < n; i++) po_prng_fill(78, 2, i, 0, &in[(size_t)i * S], S);
	std::vector<uint16_t> exps = {0, 5};
	PAR2ProcHIP be(0, 2);
	be.setSliceSize(S);
	if(!be.init(64, 2)) return fail("init");
	if(!be.setRecoverySlices(2, exps.data())) return fail("setRecoverySlices");
	std::vector<uint8_t> out(S);
	for(int pass = 0; pass < 2; pass++) {
		if(pass == 0 && be.debugFailNextBatch(PPH_E_NOMEM) != PPH_OK) return fail("fault hook");
		for(unsigned i = 0; i < n; i++) {
			if(be.canAdd() == PROC_ADD_FULL) be.waitForAdd();
			be.addInput(&in[(size_t)i * S], S, (uint16_t)i, false).get();
		}
		be.endInput().get();
		bool any = false;
		for(unsigned r = 0; r < 2; r++) any = be.getOutput(r, out.data()).get() || any;
		if(pass == 0 && any) return fail("an output of a computation with a failed batch was reported valid");
		if(pass == 1) {
			std::vector<uint8_t> ref(2 * S);
			std::vector<uint16_t> nums = {0, 1, 2, 3, 4, 5};
			std::vector<size_t> lens(n, S);
			po_recover(n, in.data(), S, lens.data(), nums.data(), nullptr, 2, exps.data(), S, ref.data(), 0);
			for(unsigned r = 0; r < 2; r++) {
				if(!be.getOutput(r, out.data()).get()) return fail("output after a fresh computation");
				if(memcmp(out.data(), &ref[(size_t)r * S], S)) return fail("output bytes after a fresh computation");
			}
		}
		be.discardOutput();
	}
	return 0;
}

int main() {
	po_init();
	if(PAR2ProcHIP::getNumDevices() <= 0) return fail("no device");
	{
		PAR2ProcHIP probe(0, 2);
		printf("device: %s\n", probe.getDeviceName().c_str());
	}
	std::vector<uint16_t> e1 = {0, 1, 2, 3, 100, 65534};
	std::vector<uint16_t> e2;
	for(unsigned r = 0; r < 37; r++) e2.push_back((uint16_t)(r * 1771 % 65535));
	int rc = 0;
	rc |= run(65536, 37, e1, 8, 2, 64, false);
	rc |= run(65536 + 6, 37, e2, 5, 3, 64, true);
	rc |= run(1000, 9, e1, 64, 2, 16, true);
	rc |= run(2, 3, e1, 1, 2, 8, false);
	rc |= run_failed_batch();
	if(rc) return 1;
	printf("ALL OK\n");
	return 0;
}
