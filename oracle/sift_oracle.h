/*
 * CPU oracle for the PySIFT detect -> describe -> match -> RANSAC -> warp/feather
 * path.  TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library to check results.  The product
 * (pysift_amd) never links, imports or calls it.
 *
 * Every function restates the reference's algorithm (pysift/core.py, cited per
 * function) with the numerics pysift_amd defines where the reference leaves
 * them open (summation order, transcendental functions, FP64 solves).  The GPU
 * library is required to match these outputs bit for bit.
 */
#ifndef SIFT_ORACLE_H
#define SIFT_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int    n_octaves;            /* core.py:1191 (5 on a >=12 GB device) */
    int    n_scales;             /* S, core.py:1184 */
    double sigma0;               /* core.py:1184 */
    double contrast_thresh;      /* core.py:1185 */
    double edge_thresh;          /* core.py:1185 */
    int    double_image;         /* core.py:1188 */
    int    rootsift;             /* core.py:1188 */
    int    double_px_cap;        /* core.py:1270: 16e6 high-VRAM */
    int    max_cands_per_octave; /* core.py:1395 */
    int    max_keypoints;        /* core.py:2078 */
    int    border;               /* core.py:1381 */
    int    retry_below;          /* core.py:2064: retry at ct/2 when fewer */
    int    adaptive_thresh;      /* core.py:2045 */
    int    dsp_n_scales;         /* core.py:1196-1206: 0 = plain SIFT, 3/4/5 = DSP-SIFT preset */
    int    fp16_pyramid;         /* core.py:1207, 1244, 1308: binary16 base and octaves >= 2 */
} or_params;

/* ---- elementary functions (double, IEEE basic operations only) ---- */
double or_exp(double x);
double or_log2(double x);
double or_atan2(double y, double x);
float or_exp_f(float x);
float or_f16_round(float x);
void or_f16_round_arr(const float* x, float* y, int n);
float or_atan2_f(float y, float x);
void or_exp_f_arr(const float* x, float* y, int n);
void or_atan2_f_arr(const float* y, const float* x, float* a, int n);
void   or_sincos(double x, double* s, double* c);

/* ---- pyramid pieces ---- */
int    or_radius(double sigma);                                   /* int(4*sigma+0.5) */
void   or_taps(double sigma, int r, float* taps);                 /* 2r+1 normalised taps */
void   or_blur(const float* in, float* out, int H, int W, double sigma);
void   or_upsample2x(const float* in, int H, int W, float* out);  /* zoom(x,2,order=3) */
void   or_u8_to_unit(const uint8_t* in, int H, int W, int pitch, float* out);
double or_adaptive_thresh(const uint8_t* img, int H, int W, int pitch, double base);

/* Full plain detectAndCompute.  Outputs are allocated by the caller with room
 * for `cap` keypoints: kp[cap*6] = x,y,size,angle,response,octave;
 * desc[cap*128] (f32), desc_u8[cap*128].  Returns the keypoint count
 * (clipped to cap) or <0 on error. */
int or_detect_describe(const uint8_t* img, int H, int W, int pitch, const or_params* p,
                       int cap, float* kp, float* desc, uint8_t* desc_u8);

/* Pyramid only: writes octave shapes and, when non-NULL, the Gaussian
 * (S+3 levels) and DoG (S+2 levels) stacks of octave `which`. */
int or_pyramid_octave(const uint8_t* img, int H, int W, int pitch, const or_params* p,
                      int which, int* Ho, int* Wo, float* gauss, float* dog);

/* ---- matching (exact, integer) ---- */
void or_match_u8(const uint8_t* a, int N, const uint8_t* b, int M, double ratio,
                 int32_t* fwd /*N*3: j,d1,d2*/, int32_t* bwd /*M*3: i,d1,d2*/,
                 int32_t* pairs /*min(N,M)*2*/, int* npairs);

/* ---- the reference's float matcher: L2-normalise, PCA, fp16 similarity ---- */
#define OR_PCA_CHUNK 128
#define OR_PCA_MAX_SWEEPS 40
void or_l2n_rows(const float* x, int n, float* y /*n*128*/);
int  or_pca_fit(const float* xa, int N, const float* xb, int M, int dims, float* mean /*128*/,
                float* proj /*128*dims*/, double* evals /*128, descending*/, int* sweeps);
void or_pca_project(const float* x, int n, const float* mean, const float* proj, int dims, float* y /*n*dims*/);
void or_sim16(const float* a, int N, const float* b, int M, int D, float* s16 /*N*M*/);
void or_match_sim16(const float* a, int N, const float* b, int M, int D, double ratio,
                    int32_t* fwd /*N*3: j, d1 bits, d2 bits*/, int32_t* bwd, int32_t* pairs, int* npairs);

/* ---- F3: gain + DP seam + multiband compositing ---- */
void or_grid_warp(const uint8_t* img, int h, int w, const double* H, int cw, int ch, double ox, double oy,
                  uint8_t* out /*ch*cw*3*/, uint8_t* mask /*ch*cw*/);
void or_gauss_weights_sp(float* w9);                  /* sigma 1, radius 4, float32 */
void or_gauss_weights_cv(float* w25);                 /* cv2 getGaussianKernel(25, 3), float32 */
void or_gauss_down(const float* in, int H, int W, const float* w9, float* out /*((H+1)/2)*((W+1)/2)*/);
void or_zoom_up(const float* in, int h, int w, int th, int tw, float* out);
void or_sobel(const float* in, int H, int W, int axis, float* out);
void or_cv_blur25(const float* in, int H, int W, const float* w25, float* out);
void or_dp_seam(const float* cost, const uint8_t* ov, int H, int W, int vertical, int32_t* seam);
void or_hist_lut(const uint64_t* hs, const uint64_t* hr, float* lut);
int  or_composite(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2, const double* H,
                  int cw, int ch, double ox, double oy, int n_levels, int color_match,
                  uint8_t* pano /*ch*cw*3*/, uint8_t* valid, float* seam_mask /*ch*cw, may be NULL*/,
                  float* gain /*3*/, double* rmse, int* vertical);

/* ---- Philox4x32-10 ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_ransac_sample(int B, int N, uint64_t seed, int32_t* idx /*B*4*/);

/* ---- homography ---- */
int  or_dlt_ls(const double* p1, const double* p2, int N, double* H);
int  or_ransac(const double* p1, const double* p2, int N, uint64_t seed, int B,
               double thr, double sigma_max, int method, int lo, int min_inliers,
               double* H, uint8_t* mask, double* scores /*B or NULL*/,
               double* Hs /*B*9 or NULL*/);

/* ---- warp + feather ---- */
void or_canvas(int h1, int w1, int h2, int w2, const double* H,
               int* cw, int* ch, double* ox, double* oy);
void or_warp_feather(const uint8_t* img1, int h1, int w1, const uint8_t* img2, int h2, int w2,
                     const double* H, int cw, int ch, double ox, double oy,
                     uint8_t* pano /*ch*cw*3*/, uint8_t* valid /*ch*cw*/);
void or_bgr_to_gray(const uint8_t* bgr, int H, int W, uint8_t* gray);
int  or_set_threads(int n);

/* ---- single-stage hooks (one octave / one keypoint) ---- */
int  or_extrema_octave(const float* D, int S, int H, int W, float thr, int border, int cap, int* out);
int  or_refine_octave(const float* D, int S, int H, int W, int border, float thr, float edge,
                      const int* cands, int n, int* valid, float* out);
int  or_orientation(const float* G, int H, int W, float ri, float ci, float sigma, float* angles);
/* CLAHE with OpenCV semantics on a u8 image (core.py:2562-2565) */
int  or_clahe(const uint8_t* src, int H, int W, int pitch, double clip, int tiles_x, int tiles_y, uint8_t* dst);
void or_descriptor_dsp(const float* G, int H, int W, float ri, float ci, float ang_rad, int ph, int rootsift,
                       int dsp_n_scales, float* out, uint8_t* out_u8);
void or_descriptor(const float* G, int H, int W, float ri, float ci, float ang_rad, int ph, int rootsift,
                   float* out, uint8_t* out_u8);

#ifdef __cplusplus
}
#endif
#endif
