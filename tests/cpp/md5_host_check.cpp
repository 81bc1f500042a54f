// Reads stdin, writes the MD5 digest of it (hex) computed by the serving
// engine's host routine, fed in pieces of the sizes given as arguments.
#include "../../parpar_amd/csrc/pph_md5_host.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv) {
	std::vector<uint8_t> data;
	uint8_t buf[65536];
	size_t n;
	while((n = fread(buf, 1, sizeof(buf), stdin)) > 0) data.insert(data.end(), buf, buf + n);
	pph::HostMd5 h;
	size_t pos = 0;
	for(int a = 1; a < argc && pos < data.size(); a++) {
		size_t k = strtoull(argv[a], nullptr, 10);
		if(k > data.size() - pos) k = data.size() - pos;
		h.update(data.data() + pos, k);
		pos += k;
	}
	h.update(data.data() + pos, data.size() - pos);
	uint8_t d[16];
	h.final(d);
	for(int i = 0; i < 16; i++) printf("%02x", d[i]);
	printf("\n");
	return 0;
}
