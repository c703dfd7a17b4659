/*
 * pysift_amd C-ABI (libpysift_amd.so): MI355X-native SIFT detect -> describe
 * -> match -> RANSAC -> warp/feather.
 *
 * Conventions
 *   - Every entry point returns int status: 0 on success, <0 on error
 *     (MSIFT_E*); msift_last_error() returns the thread's last message.
 *     Nothing throws across the ABI.
 *   - Device buffers are caller-owned (torch tensors in the Python facade),
 *     passed as raw pointers plus sizes.  `stream` is a hipStream_t; all work
 *     is enqueued asynchronously on it.
 *   - Counts are written to device memory, so a caller may capture the calls
 *     in a HIP graph.  Workspaces are sized by the *_workspace_bytes queries.
 *   - Not thread-safe per handle: use one handle per thread or stream.
 *
 * Each declaration names the reference interface it replaces
 * (SivaIITM/PySIFT pysift/core.py:<line>).
 */
#ifndef PYSIFT_AMD_H
#define PYSIFT_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSIFT_ABI_VERSION 2

#define MSIFT_OK          0
#define MSIFT_EINVAL     -1   /* bad argument (Python: ValueError) */
#define MSIFT_EHIP       -2   /* HIP runtime failure */
#define MSIFT_EWORKSPACE -3   /* workspace too small */
#define MSIFT_ESTATE     -4   /* operation not valid in this state (RuntimeError) */

/* Parameters of PySIFT.__init__ (core.py:1184-1232) plus the caps that the
 * reference hard-codes (core.py:1395, 2078-2080, 1270). */
typedef struct msift_params {
    int    n_octaves;            /* core.py:1191 (5 on a >=12 GB device) */
    int    n_scales;             /* S */
    double sigma0;
    double contrast_thresh;
    double edge_thresh;
    int    double_image;
    int    rootsift;
    int    double_px_cap;        /* core.py:1270: 16 000 000 on a >=12 GB device */
    int    max_cands_per_octave; /* core.py:1395: 25 000 */
    int    max_keypoints;        /* core.py:2078: 8000 */
    int    border;               /* core.py:1381: 5 */
    int    retry_below;          /* core.py:2064: 100 */
    int    adaptive_thresh;      /* core.py:2045: 1 */
    int    dsp_n_scales;         /* core.py:1196-1206: 0 = plain SIFT descriptor, 3/4/5 = DSP-SIFT preset */
    int    fp16_pyramid;         /* core.py:1207: 1 = binary16 base and octaves >= 2 (the reference's
                                    low-VRAM numerics; storage stays float32) */
} msift_params;

typedef struct msift_handle msift_handle;

int msift_abi_version(void);
int msift_last_error(char* buf, size_t n);

/* PySIFT.__init__ (core.py:1184): validates parameters, precomputes taps. */
int msift_create(const msift_params* p, msift_handle** h);
int msift_destroy(msift_handle* h);
/* No counterpart (the reference's Gaussian kernels have one launch shape):
 * tile height of the pyramid blur kernels, 0 = chosen per level size (the
 * default), or 8, 16, 32, 64.  Results do not depend on it. */
int msift_set_blur_tile(msift_handle* h, int tile_h);

/* Workspace for detect_describe on an H x W image (pyramid, candidates,
 * keypoint lists).  Replaces the CuPy memory-pool allocations of
 * core.py:1234-1460. */
int msift_workspace_bytes(msift_handle* h, int H, int W, size_t* bytes);

/* PySIFT.detectAndCompute (core.py:2021-2172) on a device-resident uint8
 * image.  Writes up to `cap` keypoints sorted by (x, y):
 *   d_kp[cap*6]      = x, y, size, angle(deg), response, octave
 *   d_desc_f32       = [cap,128] RootSIFT float descriptors (or NULL)
 *   d_desc_u8        = [cap,128] uint8 = min(255, trunc(512 v + 0.5)) (or NULL)
 *   d_count          = number of keypoints (before clipping to cap)
 * Set phase_ms (8 floats) non-NULL to time the stages with HIP events
 * (synchronises the stream; keys of core.py:2152-2161). */
int msift_detect_describe(msift_handle* h, const uint8_t* d_img, int H, int W, int pitch,
                          void* d_ws, size_t ws_bytes, int cap,
                          float* d_kp, float* d_desc_f32, uint8_t* d_desc_u8, int* d_count,
                          float* phase_ms, void* stream);

/* Batched detectAndCompute over n images of uneven size (no reference
 * counterpart: the reference processes one image per call, core.py:2021).
 * A plan fixes the image sizes, builds the per-image / per-octave tile tables
 * in device memory (the only allocation, done here) and the workspace layout.
 * Inputs are packed back to back (row pitch = width) at the offsets that
 * msift_batch_sizes reports.  Every stage runs as one launch over all images
 * and the call never synchronises, so it can be captured in a HIP graph.
 * Outputs: images concatenated in order, each sorted by (x, y) exactly as
 * msift_detect_describe would return it; d_offsets[n+1], d_counts[n]. */
typedef struct msift_batch msift_batch;
int msift_batch_create(msift_handle* h, int n, const int* H, const int* W, msift_batch** plan);
int msift_batch_destroy(msift_batch* plan);
int msift_batch_sizes(msift_batch* plan, size_t* ws_bytes, size_t* input_bytes, size_t* input_offsets);
/* tile height of the blur kernels (0 = default, 16, 32, 64); results do not
 * depend on it (set by the autotuner). */
int msift_batch_set_blur_tile(msift_batch* plan, int tile_h);
int msift_detect_describe_batched(msift_batch* plan, const uint8_t* d_packed, void* d_ws, size_t ws_bytes, int cap,
                                  float* d_kp, float* d_desc_f32, uint8_t* d_desc_u8, int* d_counts, int* d_offsets,
                                  void* stream);

/* Differentiable descriptor (no reference counterpart: the reference has no
 * gradient path; the forward is its pyramid, core.py:1234-1352, and
 * descriptor, core.py:693-835 + RootSIFT core.py:1870-1873).
 * B float32 images H x W in [0, 1] (row-major, back to back).  With the
 * handle's double_image set (and H * W below its double_px_cap) the base is
 * the x2 cubic B-spline upsample (core.py:1270-1274) and octave 0 has
 * 2H x 2W pixels; otherwise the single-size base.  The handle must describe
 * the plain descriptor (dsp_n_scales = 0, fp16_pyramid = 0).  d_kps[B*K*5] = (octave,
 * scale index 1..S, row, col in octave pixels, angle in degrees) per slot;
 * an octave < 0 marks an empty slot.  Forward: d_desc[B*K*128].  Backward:
 * given d_gdesc = dL/d(desc) it recomputes the pyramid and writes
 * d_gimg[B*H*W] = dL/d(image); bitwise reproducible (fixed-point integer
 * accumulation of the per-keypoint scatters).  Workspace from
 * msift_grad_workspace_bytes (both calls). */
int msift_grad_workspace_bytes(msift_handle* h, int B, int H, int W, size_t* bytes);
int msift_desc_forward(msift_handle* h, const float* d_imgs, int B, int H, int W, const float* d_kps, int K,
                       void* d_ws, size_t ws_bytes, float* d_desc, void* stream);
int msift_desc_backward(msift_handle* h, const float* d_imgs, int B, int H, int W, const float* d_kps, int K,
                        const float* d_gdesc, void* d_ws, size_t ws_bytes, float* d_gimg, void* stream);
/* Level access for tests, inside a workspace filled by msift_desc_backward:
 * image b, octave o: Gaussian levels 0..S (float, Ho x Wo each) and the
 * descriptor stage's gradient of levels 1..S before the blur adjoints (int64,
 * fixed point: value * 2^32). */
int msift_grad_level_view(msift_handle* h, int B, int H, int W, void* d_ws, int b, int o, int* Ho, int* Wo,
                          float** d_gauss, long long** d_grad_fixed);

/* Octave-level access for tests: the Gaussian stack (S+3 levels) of octave
 * `o` inside a workspace filled by msift_detect_describe.  The DoG is not
 * stored (its readers form D[s] = G[s+1] - G[s]): *d_dog is set to NULL. */
int msift_pyramid_view(msift_handle* h, int H, int W, void* d_ws, int o,
                       int* Ho, int* Wo, float** d_gauss, float** d_dog);

/* Brute-force ratio-test matcher, exact on uint8 descriptors: replaces
 * _match_features / _match_symmetric (core.py:2840-2940, 2652-2701).
 * d2 = sum (a-b)^2 is computed on INT8 MFMA with INT32 accumulation.
 *   method 0: high/low byte split (a = 128 h + l)  1: signed shift (a-128)
 *   2: the shift arithmetic as two one-sided passes (top-2 over b of every
 *      a, then over a of every b): twice the MFMA work, about half the
 *      integer work per distance of 0 and 1; the fastest on MI355X
 * Outputs: d_fwd[N*3] = best j, d1, d2 over b;  d_bwd[M*3] = best i, d1, d2
 * over a (ties keep the smallest index);  d_pairs[2*min(N,M)] mutual pairs
 * (i, j) passing d1 < ratio^2 d2 in both directions, in increasing i;
 * d_npairs = count. */
int msift_match_workspace_bytes(int N, int M, size_t* bytes);
int msift_match_u8(const uint8_t* d_a, int N, const uint8_t* d_b, int M, double ratio, int method,
                   void* d_ws, size_t ws_bytes, int32_t* d_fwd, int32_t* d_bwd,
                   int32_t* d_pairs, int* d_npairs, void* stream);

/* Tile shape of method 2 for the whole process (no reference counterpart;
 * the reference has no tile choice): shape 0 = 64 a-rows per wave with the
 * next tile's MFMAs issued ahead, 1 = 64 rows, 2 = 32 rows, 3 = 96 rows,
 * 4 = 128 rows (each staged y fragment feeds more MFMAs); chunks_per_block
 * = 512-column chunks per workgroup, 1..64.  Outputs do not depend on it; the
 * autotuner (pysift_amd.matcher.autotune) sets it.  Default (0, 4). */
int msift_match_set_tile(int shape, int chunks_per_block);

/* Allocation-free variant of msift_match_u8 (no counterpart in the
 * reference: it is for callers that cannot allocate device memory, such as
 * code under HIP graph capture or a serving loop that forbids hipMalloc).
 * Same arguments and outputs as msift_match_u8 without the workspace: the
 * scratch (norms, the two-pass method's y' copies and key terms, partial
 * top-2 keys, block counts) lives in a static device
 * array of 17.0 MB per GPU, held from the first call on that GPU until the
 * process ends, which limits N and M to 16384 each (larger sizes return
 * MSIFT_EINVAL).  The stream must belong to the current device.  Every call
 * on one GPU shares that GPU's scratch, so calls that overlap on one GPU
 * (concurrent streams or threads) race; serialise them.  Nothing is
 * allocated and no scratch needs clearing: every call writes all of the
 * scratch it reads. */
int msift_match_u8_noalloc(const uint8_t* d_a, int N, const uint8_t* d_b, int M, double ratio, int method,
                           int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs, int* d_npairs, void* stream);

/* Batched matcher (no reference counterpart: the reference matches one pair
 * per call, core.py:2840-2940): n_pairs descriptor-set pairs (a_p: N[p] rows,
 * b_p: M[p] rows of 128 B), packed back to back at the row offsets that
 * msift_match_batch_sizes reports.  One launch computes every tile of every
 * pair; per pair the outputs are exactly msift_match_u8's:
 *   d_fwd[3 * (a_off + i)], d_bwd[3 * (b_off + j)],
 *   d_pairs[2 * (a_off + k)] for k < d_npairs[p] (room for N[p] pairs).
 * The plan (device pair table) is the only allocation; run is capturable. */
typedef struct msift_match_batch msift_match_batch;
int msift_match_batch_create(int n_pairs, const int* N, const int* M, msift_match_batch** plan);
int msift_match_batch_destroy(msift_match_batch* plan);
int msift_match_batch_sizes(msift_match_batch* plan, size_t* ws_bytes, int* a_offsets, int* b_offsets);
int msift_match_u8_batched(msift_match_batch* plan, const uint8_t* d_a, const uint8_t* d_b, double ratio, int method,
                           void* d_ws, size_t ws_bytes, int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs,
                           int* d_npairs, void* stream);
/* Matched point coordinates for batched RANSAC: keypoint rows (x, y, ...) of
 * stride_a / stride_b floats in the same packed order as the descriptors;
 * d_p1/d_p2[2 * (a_off + k)] = (x, y) of match k of pair p (doubles). */
int msift_match_batch_gather_points(msift_match_batch* plan, const float* d_kpa, int stride_a, const float* d_kpb,
                                    int stride_b, const int32_t* d_pairs, const int* d_npairs, double* d_p1,
                                    double* d_p2, void* stream);

/* Float descriptor path (the reference's approximate float matcher, core.py:
 * 2574-2743, 2840-2940, as an FP16 GEMM): inputs rounded to fp16, a.b on FP16
 * MFMA with FP32 accumulation, norms in FP32 from the rounded values,
 * d2 = |a|^2 + |b|^2 - 2 a.b clamped to [0, 4] (unit-norm descriptors) and
 * quantised to 2^-21 for the top-2 keys.  Same outputs as msift_match_u8 with
 * the distance slots holding float bits (+inf = none).  Workspace from
 * msift_match_workspace_bytes. */
int msift_match_f16(const float* d_a, int N, const float* d_b, int M, double ratio,
                    void* d_ws, size_t ws_bytes, int32_t* d_fwd, int32_t* d_bwd,
                    int32_t* d_pairs, int* d_npairs, void* stream);

/* GPU-batched RANSAC + LO refit: replaces _ransac / _ransac_batch_eval /
 * _compute_homography (core.py:3011-3097, 2272-2313, 2959-2981).
 * Hypotheses: Philox4x32-10, counter (b,0,0,0), key = seed; FP64 DLT.
 * method 0 = MAGSAC soft score, 1 = inlier count.  d_H[9] row-major,
 * d_mask[N] inliers, d_ninl = inlier count, d_best = winning hypothesis. */
int msift_ransac_workspace_bytes(int N, int n_hyp, size_t* bytes);
int msift_ransac_homography(const double* d_p1, const double* d_p2, int N, uint64_t seed, int n_hyp,
                            double thr_px, double sigma_max, int method, int lo, int min_inliers,
                            void* d_ws, size_t ws_bytes, double* d_H, uint8_t* d_mask,
                            int* d_ninl, int* d_best, void* stream);

/* Batched RANSAC (no reference counterpart: _ransac runs per pair,
 * core.py:3011-3097): n_pairs point sets packed back to back, pair p at the
 * point offset msift_ransac_batch_sizes reports with room for cap[p] points;
 * the counts d_npts[p] are read on the device (pairs below 4 points get
 * H = 0, no inliers, best = -1).  Pair p uses key seed + p, so it equals
 * msift_ransac_homography with seed + p.  Outputs d_H[9 * p], d_mask at the
 * pair's point offset, d_ninl[p], d_best[p]. */
typedef struct msift_ransac_batch msift_ransac_batch;
int msift_ransac_batch_create(int n_pairs, const int* cap, int n_hyp, msift_ransac_batch** plan);
int msift_ransac_batch_destroy(msift_ransac_batch* plan);
int msift_ransac_batch_sizes(msift_ransac_batch* plan, size_t* ws_bytes, long long* point_offsets);
int msift_ransac_homography_batched(msift_ransac_batch* plan, const double* d_p1, const double* d_p2,
                                    const int* d_npts, uint64_t seed, double thr_px, double sigma_max, int method,
                                    int lo, int min_inliers, void* d_ws, size_t ws_bytes, double* d_H,
                                    uint8_t* d_mask, int* d_ninl, int* d_best, void* stream);

/* _warp_image + compositing (core.py:3210-3239, 3794-3802) with a feather
 * blend: img1 (BGR u8, h1 x w1) warped by H into a cw x ch canvas,
 * img2 pasted at (round(ox), round(oy)); out pano [ch*cw*3], valid [ch*cw]. */
int msift_warp_feather(const uint8_t* d_img1, int h1, int w1, const uint8_t* d_img2, int h2, int w2,
                       const double* H, int cw, int ch, double ox, double oy,
                       uint8_t* d_pano, uint8_t* d_valid, void* stream);

/* cv2.cvtColor(BGR2GRAY) fixed-point BT.601 (core.py:2548). */
int msift_bgr_to_gray(const uint8_t* d_bgr, int H, int W, uint8_t* d_gray, void* stream);

/* The reference's float matcher (core.py:2840-2940 with int8_matching off,
 * _match_symmetric core.py:2652-2701, _pca_compress core.py:2703-2743):
 *   msift_l2_normalize   rows of 128 floats -> x / (||x|| + 1e-7)
 *                        (core.py:2879-2880);
 *   msift_pca_fit        PCA of the normalised rows of a and b stacked
 *                        (core.py:2728-2736): d_mean[128], d_proj[128 * dims]
 *                        (columns = eigenvectors of the dims largest
 *                        eigenvalues, ascending, as eigh(cov)[:, -dims:]),
 *                        d_evals[128] descending (double), *d_sweeps =
 *                        Jacobi sweeps; FP64, deterministic; workspace from
 *                        msift_pca_workspace_bytes, 8-byte aligned;
 *   msift_pca_project    unit rows -> ((x - mean) @ proj) / (||.|| + 1e-7)
 *                        (core.py:2737-2742), n x dims;
 *   msift_match_sim_f16  unit rows a (N x D), b (M x D), 1 <= D <= 128:
 *                        binary16 similarity on FP16 MFMA (fp32 accumulation,
 *                        result rounded to binary16), top-2 by similarity in
 *                        both directions, ratio test on d = max(2 - 2 s, 0),
 *                        mutual pairs; fwd/bwd triples (index, d1, d2) hold d
 *                        as float32 bits; workspace from
 *                        msift_match_workspace_bytes(N, M). */
int msift_l2_normalize(const float* d_x, int n, float* d_y, void* stream);
int msift_pca_workspace_bytes(int N, int M, size_t* bytes);
int msift_pca_fit(const float* d_a, int N, const float* d_b, int M, int dims, void* d_ws, size_t ws_bytes,
                  float* d_mean, float* d_proj, double* d_evals, int* d_sweeps, void* stream);
int msift_pca_project(const float* d_x, int n, const float* d_mean, const float* d_proj, int dims, float* d_y,
                      void* stream);
int msift_match_sim_f16(const float* d_a, int N, const float* d_b, int M, int D, double ratio, void* d_ws,
                        size_t ws_bytes, int32_t* d_fwd, int32_t* d_bwd, int32_t* d_pairs, int* d_npairs,
                        void* stream);

/* The reference's composition of a pair (_blend_pair, core.py:3686-3714;
 * color_match 0 = 'gain', 1 = 'hist' (_hist_match, core.py:3591-3626); plus
 * the warp and paste of _stitch_pair,
 * core.py:3783-3802): image 1 (BGR u8, h1 x w1, at least 2 x 2) warped by H
 * with grid_sample semantics, image 2 pasted at (round(ox), round(oy)),
 * colour matching over the overlap (gain, or histogram matching), a DP seam through the colour + gradient
 * cost, the seam mask softened by a sigma-3 Gaussian, and an n_levels
 * Laplacian-pyramid blend (1 <= n_levels <= 12).  Outputs: d_pano
 * [ch*cw*3] (0 outside both images), d_valid [ch*cw], optional d_seam_mask
 * [ch*cw] float (1 = image 1), and *d_stats (device memory) below.
 * Workspace from msift_composite_workspace_bytes, 256-byte aligned. */
typedef struct msift_composite_stats {
    unsigned long long sum1[3], sum2[3];  /* overlap channel sums of the warped image 1 and of image 2 */
    unsigned long long n_overlap;         /* overlap pixels */
    unsigned long long sq_diff;           /* sum of squared u8 differences over the overlap (seam RMSE) */
    unsigned long long reserved;
    float gain[3];                        /* per-channel gain applied to image 1 */
    int vertical;                         /* 1: vertical seam, 0: horizontal */
    int img1_first;                       /* image 1 left of (vertical) / above (horizontal) the seam */
    int has_overlap;
    double rmse;                          /* _seam_rmse (core.py:3556-3562); NaN without overlap */
} msift_composite_stats;
int msift_composite_workspace_bytes(int cw, int ch, int n_levels, size_t* bytes);
int msift_composite_multiband(const uint8_t* d_img1, int h1, int w1, const uint8_t* d_img2, int h2, int w2,
                              const double* H, int cw, int ch, double ox, double oy, int n_levels,
                              int color_match, void* d_ws, size_t ws_bytes, uint8_t* d_pano, uint8_t* d_valid,
                              float* d_seam_mask, msift_composite_stats* d_stats, void* stream);

/* cv2.createCLAHE(clipLimit, (tiles_x, tiles_y)).apply(gray) (core.py:2562-2565,
 * applied to every image ahead of detectAndCompute) with OpenCV's 8-bit
 * semantics: BORDER_REFLECT_101 tiling of ragged images, clip limit
 * max(1, (int)(clip * tile_area / 256)) (clip_limit <= 0: no clipping),
 * redistribution of the clipped count, bilinear blend of the 4 nearest tile
 * LUTs.  d_src is H x W with row pitch `pitch` bytes; d_dst is dense H x W
 * (d_dst == d_src with pitch == W runs in place).  1 <= tiles <= 128; d_ws is
 * 4-byte aligned. */
int msift_clahe_workspace_bytes(int tiles_x, int tiles_y, size_t* bytes);
int msift_clahe(const uint8_t* d_src, int H, int W, int pitch, double clip_limit, int tiles_x, int tiles_y,
                void* d_ws, size_t ws_bytes, uint8_t* d_dst, void* stream);

#ifdef __cplusplus
}
#endif
#endif
