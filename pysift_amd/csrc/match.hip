// Exact brute-force matcher (placeholder until the MFMA kernels land).
#include "msift_internal.hpp"
namespace msift {
size_t match_workspace(int N, int M) { (void)N; (void)M; return 256; }
int match_u8(const uint8_t*, int, const uint8_t*, int, double, int, void*, size_t, int32_t*, int32_t*, int32_t*,
             int*, hipStream_t) { set_error("match_u8: not available"); return MSIFT_ESTATE; }
int match_f16(const float*, int, const float*, int, double, void*, size_t, int32_t*, int32_t*, int32_t*, int*,
              hipStream_t) { set_error("match_f16: not available"); return MSIFT_ESTATE; }
}
