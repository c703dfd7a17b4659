// RANSAC (placeholder).
#include "msift_internal.hpp"
namespace msift {
size_t ransac_workspace(int N, int B) { (void)N; (void)B; return 256; }
int ransac(const double*, const double*, int, uint64_t, int, double, double, int, int, int, void*, size_t, double*,
           uint8_t*, int*, int*, hipStream_t) { set_error("ransac: not available"); return MSIFT_ESTATE; }
}
