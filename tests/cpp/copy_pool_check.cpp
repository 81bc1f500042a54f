// The backend's parallel host copy: every size around the split threshold,
// any number of helpers, many calls in a row; prints "ok" when every copy
// equals its source.
#include "../../parpar_amd/csrc/pph_copy_pool.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main() {
	const size_t P = pph::CopyPool::kPart;
	const size_t sizes[] = {0, 1, 63, P - 1, P, 2 * P - 1, 2 * P, 2 * P + 1, 3 * P + 17, 7 * P + 5, 16 * P, 33 * P + 1};
	for(unsigned helpers = 0; helpers < 8; helpers++) {
		pph::CopyPool pool(helpers);
		for(int rep = 0; rep < 3; rep++)
			for(size_t n : sizes) {
				std::vector<uint8_t> src(n + 64), dst(n + 64, 0xEE);
				for(size_t k = 0; k < src.size(); k++) src[k] = (uint8_t)(k * 131 + n + rep);
				pool.copy(dst.data() + 1, src.data() + 3, n);
				for(size_t k = 0; k < n; k++)
					if(dst[1 + k] != src[3 + k]) {
						printf("mismatch helpers %u size %zu at %zu\n", helpers, n, k);
						return 1;
					}
				if(dst[0] != 0xEE || dst[1 + n] != 0xEE) {
					printf("overrun helpers %u size %zu\n", helpers, n);
					return 1;
				}
			}
	}
	printf("ok\n");
	return 0;
}
