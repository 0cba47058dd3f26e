/*
 * nli_oracle.h -- CPU oracle for the inner inpainting loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a clean-room CPU restatement of
 * the reference algorithm (Tom-Zheng/non-local-inpainting) used as the parity
 * checker of the HIP implementation and as the timed CPU baseline.  Nothing
 * in the product path (the non-local-inpainting_amd package and libnli.so)
 * links, loads or calls it; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do.
 *
 * Layout conventions (identical to the reference containers):
 *   images  float32, interleaved HWC, index = C*(W*y + x) + c   (image.hpp:481-485)
 *   masks   uint8 0/1, index = W*y + x                          (mask.h, bool = 1 byte)
 *   NNF     int32 pairs (x, y), index = 2*(W*y + x); (-1,-1) = none (patch_match.cpp:89-90)
 *   point lists are row-major (mask.cpp:459-486)
 */
#ifndef NLI_ORACLE_H_
#define NLI_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* patch distance kinds (main.cpp:245-263 pairing) */
enum { NLO_DIST_L2 = 0, NLO_DIST_L1 = 1, NLO_DIST_L2COMBINED = 2, NLO_DIST_L2RGBD = 3 };
/* image updating kinds */
enum { NLO_UPD_MEANS = 0, NLO_UPD_MEDIANS = 1, NLO_UPD_POISSON = 2, NLO_UPD_DEPTH = 3 };
/* coarsest-level initialisation, same order as ImageInpainting::InitType (image_inpainting.h) */
enum { NLO_INIT_BLACK = 0, NLO_INIT_AVG = 1, NLO_INIT_NONE = 2, NLO_INIT_POISSON = 3 };

typedef struct {
    int kind;          /* NLO_DIST_* */
    int patch;         /* square patch side P */
    float sigma;       /* Gaussian sigma of the intra-patch weights */
    float lambda;      /* L2Combined lambda, or L2RGBD lambda_rgb */
    float lambda_d;    /* L2RGBD lambda_d */
} nlo_dist_params;

typedef struct {
    int iterations;    /* sweeps (PatchMatch::_iteration_count) */
    int shots;         /* random shots limit */
    int window;        /* search window size, -1 = max(W, H) */
} nlo_pm_params;

typedef struct {
    int kind;          /* NLO_UPD_* */
    int patch;
    float sigma;
    float lambda;      /* Poisson lambda, or Depth lambda_rgb */
    float lambda_d;    /* Depth lambda_d */
    float cg_tol;      /* Poisson only; Depth hard-codes 1e-6 (patch_non_local_depth.cpp:31,33) */
    int cg_limit;
} nlo_upd_params;

typedef struct {
    int channels_solved;
    int cg_iterations[4];   /* per channel (depth: rgb channels then depth) */
    double cg_nr0[4];
    double cg_nr[4];
} nlo_cg_stats;

typedef struct {
    int method;          /* NLO_UPD_*: selects the (distance, updater) pair */
    int patch;
    float sigma;
    float lambda;
    float lambda_rgb;
    float lambda_d;
    int iters;           /* inpainting iterations per scale */
    float tol;           /* stop when total difference <= tol */
    int scales;
    float coarse;        /* coarsest size ratio */
    float conf_t;
    float conf_a;
    int init;            /* NLO_INIT_* */
    int pm_iters;
    int pm_shots;
    int pm_window;
    float cg_tol;
    int cg_limit;
    unsigned seed;       /* srand() seed for the serial PatchMatch */
    int pm_threads;      /* 0: serial PatchMatch; >0: OpenMP odd/even-buffer PatchMatch */
    int max_iters_per_scale; /* <=0: unlimited; >0: cap (bounded CPU-baseline samples) */
} nlo_process_params;

typedef struct {
    int iterations[32];      /* inpainting iterations run per scale, finest = index 0 */
    int width[32], height[32];
    double seconds_pyramid, seconds_patchmatch, seconds_update, seconds_other;
    double seconds_iterations[32];   /* per scale: the inpainting iterations */
    double seconds_propagate[32];    /* per scale: propagate_weights into it */
} nlo_process_stats;

/* ---- numerics ---- */
void  nlo_gaussian_weights(int size_x, int size_y, float sigma_x, float sigma_y, float *out);
void  nlo_gradient(const float *img, int W, int H, int C, float *grad);
void  nlo_distance_transform(const uint8_t *mask, int W, int H, float *out);
void  nlo_confidence(const uint8_t *mask, int W, int H, float decay, float asym, float *out);
float nlo_subsampling_rate(float ratio, int scales);
int   nlo_sample_size(int size, float factor);
int   nlo_downsample_with_mask(const float *img, const uint8_t *mask, int W, int H, int C,
                               float factor, float threshold, int mask_val,
                               float *out_img, uint8_t *out_mask);
int   nlo_downsample(const float *img, int W, int H, int C, float factor, float *out);
int   nlo_downsample_mask(const uint8_t *mask, int W, int H, float factor, float threshold,
                          uint8_t *out);
void  nlo_upsample(const float *in, int W, int H, int C, int W2, int H2, float *out);
/* sRGB (0-255) <-> CIE L*a*b*, io_utility.cpp:268-318 and :380-450.  Images of other
 * channel counts are copied through (rgb_to_lab: C != 3; lab_to_rgb: C not 3 or 4);
 * lab_to_rgb of a 4-channel image converts channels 0-2.  out holds W*H*3 floats when
 * converting, W*H*C otherwise. */
void  nlo_rgb_to_lab(const float *rgb, int W, int H, int C, float *lab);
void  nlo_lab_to_rgb(const float *lab, int W, int H, int C, float *rgb);

/* ---- per-scale masks ---- */
void  nlo_extended_domain(const uint8_t *hole, int W, int H, int P, uint8_t *out);
int   nlo_scale_masks(const uint8_t *hole, int W, int H, int P, int erode,
                      uint8_t *ext, uint8_t *target, uint8_t *source);
void  nlo_add_margin(uint8_t *a, uint8_t *b, int W, int H, int margin);
void  nlo_scale_nnf(const int32_t *lower_nnf, const uint8_t *lower_hole, int lW, int lH,
                    int uW, int uH, int32_t *upper_nnf);

/* ---- patch distances ---- */
float nlo_patch_distance(const nlo_dist_params *d, const float *weights,
                         const float *src, const float *src_grad, int sW,
                         const float *tgt, const float *tgt_grad, int tW, int C,
                         int sx, int sy, int tx, int ty);
int   nlo_patch_distance_field(const nlo_dist_params *d, const float *src, int sW, int sH,
                               const float *tgt, int tW, int tH, int C,
                               const int32_t *nnf, const uint8_t *tmask, float *out);

/* ---- PatchMatch ---- */
int   nlo_patchmatch(const nlo_dist_params *d, const nlo_pm_params *pm,
                     const float *src, const uint8_t *smask, int sW, int sH,
                     const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                     const int32_t *init_nnf, unsigned seed,
                     int32_t *out_nnf, float *out_dist, double *energy_per_iter);
int   nlo_patchmatch_omp(const nlo_dist_params *d, const nlo_pm_params *pm,
                         const float *src, const uint8_t *smask, int sW, int sH,
                         const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                         const int32_t *init_nnf, unsigned seed, int threads,
                         int32_t *out_nnf);

/* ---- image updating ---- */
double nlo_update(const nlo_upd_params *u, float *image, const float *orig,
                  const uint8_t *domain, const uint8_t *ext, const int32_t *nnf,
                  const float *conf, int W, int H, int C, nlo_cg_stats *stats);
void   nlo_poisson_coefficients(int patch, float sigma, float lambda,
                                const float *image, const uint8_t *ext, const int32_t *nnf,
                                const float *conf, int W, int H, int C,
                                double *a1, double *a2, double *f1, double *f2);
void   nlo_anisotropic_laplacian(double *u, const double *a, const uint8_t *mask, int W, int H);
int    nlo_conjugate_gradient(double *x, const double *a1, const double *a2, const double *b,
                              const uint8_t *mask, int W, int H, float tol, int limit,
                              double *nr0_out, double *nr_out);

/* ---- coarsest-level init (3rdparty/simpois) ---- */
void   nlo_simpois(float *out, const float *in, int w, int h, int pd);

/* ---- driver ---- */
int    nlo_process(const nlo_process_params *p, const float *in, const uint8_t *mask,
                   int W, int H, int C, float *out, nlo_process_stats *stats);

#ifdef __cplusplus
}
#endif

#endif
