// See host_math.hpp.
#include "host_math.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <thread>
#include <vector>

namespace tbtkx {

void direct_factor_table(int kind, int M, int E, double e0, double de, double s, double *fac,
                         double *den) {
	std::vector<double> acosE(E);
	for (int e = 0; e < E; e++) {
		const double En = (e0 + e * de) / s;
		acosE[e] = std::acos(En);
		den[e] = std::sqrt(1 - En * En);
	}
	const size_t per = kind == 0 ? 2 : 1;
	auto fill = [&](int first, int stride) {
		const std::complex<double> i(0, 1);
		for (int n = 1 + first; n < M; n += stride) {
			double *row = fac + (size_t)(n - 1) * E * per;
			for (int e = 0; e < E; e++) {
				if (kind == 0) {
					const std::complex<double> v = std::exp(-i * ((double)n) * acosE[e]);
					row[2 * e] = v.real();
					row[2 * e + 1] = v.imag();
				} else if (kind == 1) {
					row[e] = std::sin(((double)n) * acosE[e]);
				} else {
					row[e] = std::cos(((double)n) * acosE[e]);
				}
			}
		}
	};
	const size_t rows = M > 1 ? (size_t)(M - 1) : 0;
	const int threads = (int)std::min<size_t>(
	    std::max(1u, std::min(16u, std::thread::hardware_concurrency())), rows / 64 + 1);
	std::vector<std::thread> pool;
	for (int t = 1; t < threads; t++) pool.emplace_back(fill, t, threads);
	fill(0, threads);
	for (auto &t : pool) t.join();
}

}  // namespace tbtkx
