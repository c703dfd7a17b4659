// Internal host-side definitions shared by the library's translation units.
#pragma once
#include "msift_common.hpp"
#include "pysift_amd.h"
#include <string>

struct KpTables {
    float sig_ori[32];     // orientation sigma per scale index (float32)
    int ph[32];            // descriptor patch half-size per scale index
    double base_size[32];  // sigma0 * k^si (double)
    float dsp[5];          // DSP-SIFT patch multipliers (core.py:1198-1202)
    int n_dsp;             // 0: plain descriptor
    int q16_oct;           // first octave read as binary16 (fp16 mode: 2, else past the last)
};

struct msift_handle {
    msift_params p;
    int S;
    double sd0[2];                   // base blur: [0] single-size input, [1] doubled
    double inc[64];                  // incremental sigmas, levels 1..S+2
    msift::Taps taps_base[2];
    msift::Taps taps_inc[64];
    msift::UpsTaps ups;
    KpTables tab;
    float edge32;
    hipEvent_t ev[10];
    bool ev_init;
    int blur_th;                     // blur tile height: 0 = by level size, else 8/16/32/64
    // second stream for the pyramid levels off the octave-to-octave chain,
    // with its fork/join events; created on first use on a device
    hipStream_t side;
    hipEvent_t fork[msift::kMaxOct + 2];   // per octave, the join, the threshold's fork
    int side_dev;
    bool side_init;
};

namespace msift {

void set_error(const std::string& msg);

#define MSIFT_HIP(call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            msift::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return MSIFT_EHIP;                                                 \
        }                                                                      \
    } while (0)

#define MSIFT_LAUNCH_CHECK()                                                   \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) {                                                \
            msift::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return MSIFT_EHIP;                                                 \
        }                                                                      \
    } while (0)

void make_taps(double sigma, Taps* t);

// detect.hip
// the handle's side stream on the device of `s` (created on first use, or
// re-created when the device changes); release frees it
int side_stream(msift_handle* h, hipStream_t s, hipStream_t* out);
void release_side_stream(msift_handle* h);
int detect_workspace(msift_handle* h, int H, int W, size_t* bytes);
int detect_describe(msift_handle* h, const uint8_t* img, int H, int W, int pitch, void* ws,
                    size_t ws_bytes, int cap, float* kp, float* df, uint8_t* du, int* count,
                    float* phase_ms, hipStream_t s);
// one pyramid level: gout = blur(in) (and dec = gout[::2, ::2] when
// non-NULL); sigma < 0.01 copies.  nb images of one size, pointers
// advancing by zin / zout floats per image; th = tile height (0: chosen from
// the level size; results do not depend on it)
int blur(const float* in, float* gout, float* dec, int H, int W, double sigma, const Taps& t,
         hipStream_t s, int nb = 1, size_t zin = 0, size_t zout = 0, int th = 0);
int pyramid_view(msift_handle* h, int H, int W, void* ws, int o, int* Ho, int* Wo, float** g,
                 float** d);

// match.hip
size_t match_workspace(int N, int M);
int match_u8(const uint8_t* a, int N, const uint8_t* b, int M, double ratio, int method, void* ws,
             size_t ws_bytes, int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs,
             hipStream_t s);
int match_set_tile(int shape, int cpw);
int match_u8_noalloc(const uint8_t* a, int N, const uint8_t* b, int M, double ratio, int method, int32_t* fwd,
                     int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s);
int match_f16(const float* a, int N, const float* b, int M, double ratio, void* ws,
              size_t ws_bytes, int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs,
              hipStream_t s);

// ransac.hip
size_t ransac_workspace(int N, int B);
int ransac(const double* p1, const double* p2, int N, uint64_t seed, int B, double thr,
           double sigma_max, int method, int lo, int min_inl, void* ws, size_t ws_bytes,
           double* H, uint8_t* mask, int* ninl, int* best, hipStream_t s);

// warp.hip
int warp_feather(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2,
                 const double* H, int cw, int ch, double ox, double oy, uint8_t* pano,
                 uint8_t* valid, hipStream_t s);
int bgr_to_gray(const uint8_t* bgr, int H, int W, uint8_t* gray, hipStream_t s);
size_t clahe_workspace(int tiles_x, int tiles_y);
size_t pca_workspace(int N, int M);
size_t composite_workspace(int cw, int ch, int n_levels);
int composite(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2, const double* H, int cw,
              int ch, double ox, double oy, int n_levels, int color_match, void* ws, size_t ws_bytes, uint8_t* pano,
              uint8_t* valid, float* seam_out, void* stats_out, hipStream_t s);
int l2_normalize_rows(const float* x, int n, float* y, hipStream_t s);
int pca_fit(const float* a, int N, const float* b, int M, int dims, void* ws, size_t ws_bytes, float* mean,
            float* proj, double* evals, int* sweeps, hipStream_t s);
int pca_project(const float* x, int n, const float* mean, const float* proj, int dims, float* y, hipStream_t s);
int match_sim16(const float* a, int N, const float* b, int M, int D, double ratio, void* ws, size_t ws_bytes,
                int32_t* fwd, int32_t* bwd, int32_t* pairs, int* npairs, hipStream_t s);
int clahe(const uint8_t* src, int H, int W, int pitch, double clip, int tiles_x, int tiles_y, void* ws,
          size_t ws_bytes, uint8_t* dst, hipStream_t s);

// scan.hip helpers (device-wide exclusive scan of int32, single workgroup)
// with tmp (scan_tmp_ints(n) ints) large inputs use a multi-workgroup scan
void launch_scan(const int* in, int* out, int n, int* total_a, int* total_b, const int* gate,
                 int gate_below, hipStream_t s, int* tmp = nullptr);
int scan_tmp_ints(int n);

}  // namespace msift
