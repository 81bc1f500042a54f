// The order in which free host MD5 threads take queued files: the highest
// response ratio, (wait + expected hash time) / expected hash time, the wait
// counted from submit.  Short files then do not queue behind multi-second
// chains of large ones, and long ones still age to the front.  Host C++ only.
#pragma once
#include <cstdint>

namespace pph {

constexpr double kHostMd5Rate = 0.8e9;  // bytes per second of one host MD5 chain (the expected hash time)

inline double md5_response_ratio(uint64_t file_size, double waited_s) {
	const double t = (double)file_size / kHostMd5Rate + 1e-3;
	return (waited_s + t) / t;
}

// The element of [first, last) with the highest response ratio at time now;
// size(x) and arrival(x) give a file's bytes and submit time.  The first of
// equal ratios wins.  last when the range is empty.
template<class It, class Size, class Arrival>
It pick_highest_response_ratio(It first, It last, double now, Size size, Arrival arrival) {
	It pick = last;
	double best = -1;
	for(It it = first; it != last; ++it) {
		const double r = md5_response_ratio(size(*it), now - arrival(*it));
		if(r > best) {
			best = r;
			pick = it;
		}
	}
	return pick;
}

}  // namespace pph
