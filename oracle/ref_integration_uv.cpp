/*
 * Test infrastructure: ParPar's own controller in its libuv build (PAR2Proc,
 * gf16/controller.cpp compiled with USE_LIBUV, as binding.gyp:47 builds it)
 * driving PAR2ProcHIP's libuv form on a uv loop, alone or beside ParPar's own
 * PAR2ProcCPU (gf16/controller_cpu.cpp, also USE_LIBUV) with the slice split
 * between them by byte range.  The backends are assembled as GfProc does
 * (src/gf.cc:697-721: construct, PAR2Proc::init, then each backend's init),
 * and inputs are fed the way lib/par2.js:130-160 feeds GfProc: adds refused
 * while a backend is busy are retried from the progress callback.  Every
 * recovery slice is compared byte for byte with the oracle (po_recover) and
 * its checksum flag must be valid.  Prints "ALL OK" and exits 0 on success.
 *
 * Built by oracle/Makefile (target `uv`) from the reference sources where they
 * lie into oracle/_ref/ (git-ignored).
 *
 *   ref_integration_uv --construct-only
 *   ref_integration_uv hip|split|cpu <slice> <inputs> <recovery> <batch> [cpu_share]
 */
#include "controller.h"
#include "controller_cpu.h"
#include "gf16mul.h"
#include "controller_hip.h"
#include "gf16_oracle.h"
#include "po_prng.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

struct Run {
	uv_loop_t* loop;
	PAR2Proc proc;
	size_t S = 0;
	unsigned n = 0, R = 0;
	std::vector<uint8_t> in, out;
	std::vector<size_t> lens;
	std::vector<uint16_t> exps;
	unsigned next = 0, addCbs = 0, outCbs = 0, invalid = 0;
	bool ended = false, finished = false, closed = false;
	unsigned refusals = 0;
};

static void pump(Run* r);

static void fetch(Run* r) {
	for(unsigned k = 0; k < r->R; k++) {
		r->proc.getOutput(k, &r->out[(size_t)k * r->S], [r](bool valid) {
			if(!valid) r->invalid++;
			if(++r->outCbs == r->R) {
				r->proc.deinit([r]() { r->closed = true; });
			}
		});
	}
}

static void pump(Run* r) {
	while(r->next < r->n) {
		const unsigned i = r->next;
		const bool ok = r->proc.addInput(&r->in[(size_t)i * r->S], r->lens[i], (uint16_t)i, false, [r]() { r->addCbs++; });
		if(!ok) {
			r->refusals++;
			return;  // retried from the progress callback
		}
		r->next++;
	}
	if(!r->ended) {
		r->ended = true;
		r->proc.endInput([r]() {
			r->finished = true;
			fetch(r);
		});
	}
}

int main(int argc, char** argv) {
	uv_loop_t* loop = uv_default_loop();
	if(argc > 1 && !strcmp(argv[1], "--construct-only")) {
		PAR2ProcHIP hip(loop, 0, 2);
		printf("constructed (device available: %d)\n", hip.isInitSuccess() ? 1 : 0);
		hip.deinit();
		uv_run(loop, UV_RUN_DEFAULT);
		return 0;
	}
	if(argc < 6) {
		fprintf(stderr, "usage: %s hip|split S n R batch [cpu_share]\n", argv[0]);
		return 2;
	}
	const bool split = !strcmp(argv[1], "split");
	const bool cpuOnly = !strcmp(argv[1], "cpu");  // PAR2ProcCPU alone: the harness itself, without a GPU
	Run* r = new Run;
	r->loop = loop;
	r->S = strtoull(argv[2], nullptr, 10);
	r->n = (unsigned)atoi(argv[3]);
	r->R = (unsigned)atoi(argv[4]);
	const unsigned batch = (unsigned)atoi(argv[5]);
	size_t cpuShare = argc > 6 ? strtoull(argv[6], nullptr, 10) : r->S / 3;
	cpuShare &= ~(size_t)1;  // byte ranges start at even offsets (controller.cpp:42-84)
	if(!split) cpuShare = 0;
	if(cpuOnly) cpuShare = r->S;
	printf("case %s S=%zu n=%u R=%u batch=%u cpu_share=%zu\n", argv[1], r->S, r->n, r->R, batch, cpuShare);
	fflush(stdout);

	r->in.assign((size_t)r->n * r->S, 0);
	r->out.assign((size_t)r->R * r->S, 0xAA);
	r->lens.assign(r->n, r->S);
	r->lens[r->n - 1] = r->S / 3 + 1;  // a short last input, zero padded
	for(unsigned i = 0; i < r->n; i++) po_prng_fill(91, 7, i, 0, &r->in[(size_t)i * r->S], r->lens[i]);
	r->exps.resize(r->R);
	for(unsigned k = 0; k < r->R; k++) r->exps[k] = (uint16_t)(k * 97 % 65535);

	// construction, as GfProc does (src/gf.cc:697-708): devices first, then the CPU
	const size_t hipOff = cpuShare, hipSize = r->S - cpuShare;
	PAR2ProcHIP* hip = cpuOnly ? nullptr : new PAR2ProcHIP(loop, 0, 2);
	PAR2ProcCPU* cpu = cpuShare ? new PAR2ProcCPU(loop, 2) : nullptr;
	std::vector<PAR2ProcBackendAlloc> allocs;
	if(hip) allocs.push_back({hip, hipOff, hipSize});
	if(cpu) allocs.push_back({cpu, 0, cpuShare});
	if(!r->proc.init(r->S, allocs, [r](unsigned) {
		   if(r->next < r->n) pump(r);
	   })) {
		printf("FAIL: PAR2Proc::init\n");
		return 3;
	}
	Galois16Mul cksumGf(Galois16Mul::default_method());
	const unsigned ck = cksumGf.info().cksumSize;
	if(hip && (!hip->isInitSuccess() || !hip->init(ck, batch))) {
		printf("FAIL: PAR2ProcHIP init (status %d)\n", hip->getLastStatus());
		return 4;
	}
	if(cpu && !cpu->init(GF16_AUTO)) {
		printf("FAIL: PAR2ProcCPU init\n");
		return 5;
	}
	// (the CPU backend frees its method object at deinit, inside the loop)
	const std::string method = cpu ? cpu->getMethodName() : "-";
	if(!r->proc.setRecoverySlices(r->exps)) {
		printf("FAIL: setRecoverySlices\n");
		return 6;
	}
	pump(r);
	uv_run(loop, UV_RUN_DEFAULT);

	if(!r->finished || !r->closed || r->outCbs != r->R || r->addCbs != r->n) {
		printf("FAIL: loop ended early (finished %d closed %d outputs %u/%u adds %u/%u)\n", (int)r->finished,
		       (int)r->closed, r->outCbs, r->R, r->addCbs, r->n);
		return 7;
	}
	std::vector<uint16_t> nums(r->n);
	for(unsigned i = 0; i < r->n; i++) nums[i] = (uint16_t)i;
	std::vector<uint8_t> ref((size_t)r->R * r->S);
	po_recover(r->n, r->in.data(), r->S, r->lens.data(), nums.data(), nullptr, r->R, r->exps.data(), r->S, ref.data(), 0);
	for(unsigned k = 0; k < r->R; k++)
		if(memcmp(&ref[(size_t)k * r->S], &r->out[(size_t)k * r->S], r->S) != 0) {
			printf("FAIL: recovery slice %u differs from the oracle\n", k);
			return 8;
		}
	if(r->invalid) {
		printf("FAIL: %u recovery slices reported invalid\n", r->invalid);
		return 9;
	}
	printf("%s: %u inputs x %zu bytes -> %u recovery slices byte-exact (backend method %s, %u refused adds, "
	       "%u batches on the GPU)\n",
	       split ? "PAR2ProcCPU + PAR2ProcHIP" : cpuOnly ? "PAR2ProcCPU" : "PAR2ProcHIP", r->n, r->S, r->R,
	       method.c_str(), r->refusals, hip ? hip->getBatchesStarted() : 0);
	delete cpu;
	delete hip;
	delete r;
	printf("ALL OK\n");
	return 0;
}
