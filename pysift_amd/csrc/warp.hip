// Warp + feather (placeholder).
#include "msift_internal.hpp"
namespace msift {
int warp_feather(const uint8_t*, int, int, const uint8_t*, int, int, const double*, int, int, double, double,
                 uint8_t*, uint8_t*, hipStream_t) { set_error("warp: not available"); return MSIFT_ESTATE; }
int bgr_to_gray(const uint8_t*, int, int, uint8_t*, hipStream_t) { set_error("gray: not available"); return MSIFT_ESTATE; }
}
