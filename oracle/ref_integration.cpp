/*
 * Test infrastructure: link check of the drop-in.  Builds ParPar's own
 * controller (PAR2Proc, gf16/controller.cpp, compiled from the reference
 * sources into oracle/_ref/) with PAR2ProcHIP as its backend, exactly as
 * src/gf.cc:697-721 assembles backends.  Compiled and linked in tests; it only
 * runs its "--construct-only" path where no GPU is present.
 */
#include "controller.h"
#include "controller_hip.h"

#include <cstdio>
#include <cstring>

int main(int argc, char** argv) {
	const bool constructOnly = argc > 1 && !strcmp(argv[1], "--construct-only");
	PAR2ProcHIP hip(0, 2);
	if(constructOnly) {
		printf("constructed (device available: %d)\n", hip.isInitSuccess() ? 1 : 0);
		return 0;
	}
	if(!hip.isInitSuccess()) return 2;
	const size_t S = 4096;
	PAR2Proc proc;
	if(!proc.init(S, {{&hip, 0, S}})) return 3;
	if(!hip.init(64, 4)) return 4;
	uint16_t exps[2] = {0, 7};
	if(!proc.setRecoverySlices(2, exps)) return 5;
	uint8_t buf[S];
	for(unsigned i = 0; i < 6; i++) {
		memset(buf, (int)(i * 37 + 1), S);
		proc.addInput(buf, S, (uint16_t)i).get();
	}
	proc.endInput().get();
	uint8_t out[S];
	bool ok = proc.getOutput(0, out).get();
	uint8_t x = 0;
	for(unsigned i = 0; i < 6; i++) x ^= (uint8_t)(i * 37 + 1);
	for(size_t b = 0; b < S; b++)
		if(out[b] != x) return 6;
	printf("PAR2Proc + PAR2ProcHIP ok=%d\n", ok ? 1 : 0);
	return ok ? 0 : 7;
}
