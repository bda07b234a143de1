/*
 * gpb_capi.h — C ABI of the MI355X gPb library (libgpb.so).
 *
 * The library computes the gPb contour signal that feeds OWT-UCM: Schmid
 * texton assignment, half-disc chi-squared cues, mPb and its oriented nonmax,
 * the intervening-contour affinity W, the generalised eigensolve
 * (D - W) v = lambda D v, sPb and gPb.  All functions return GPB_OK (0) or a
 * negative status; gpb_ctx_last_error() gives the message of the last failure
 * on a context.  A context owns one HIP device, one stream and grow-only
 * workspace; calls on one context must not overlap (use one context per host
 * thread).  Images are row-major, width = columns, height = rows.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the yanyp/UCM-MPI-GPU tree).
 */
#ifndef GPB_CAPI_H
#define GPB_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPB_OK 0
#define GPB_ERR_ARG (-1)     /* invalid argument (sizes, null pointers) */
#define GPB_ERR_HIP (-2)     /* HIP runtime error (allocation, launch) */
#define GPB_ERR_NOCONV (-3)  /* eigensolver stopped before the residual target */
#define GPB_ERR_INTERNAL (-4)

typedef struct gpb_ctx gpb_ctx;

/* Context lifetime.  Replaces the per-call device selection and
 * cudaDeviceReset of computeGPb (damascene/src/damascene.cu:258-268, :660). */
int gpb_ctx_create(int device, gpb_ctx** out);
void gpb_ctx_destroy(gpb_ctx* ctx);
const char* gpb_ctx_last_error(const gpb_ctx* ctx);

/* Number of visible GPUs.  Replaces getCudaDeviceCount
 * (damascene/include/globalPb.h:44, damascene/src/damascene.cu:242). */
int gpb_device_count(void);

/* ---- inputs -------------------------------------------------------------- */

/* Pack H*W*3 RGB bytes into 0x00BBGGRR words, the pixel format computeGPb
 * reads (src/globalPb.cpp:264-276; decode damascene/src/convert/convert.cu:19-22). */
void gpb_pack_rgb(const uint8_t* rgb, size_t npixels, uint32_t* packed);

/* Seeded Voronoi mosaic (regions, oriented sinusoid textures, Gaussian
 * noise), written as H*W*3 RGB bytes; region ids optional (may be NULL).
 * No reference counterpart: the reference reads photographs (src/main.cpp:119). */
int gpb_synth_voronoi(int width, int height, int regions, double noise_sigma, uint64_t seed, int classes,
                      uint8_t* rgb, int32_t* region);

/* 13x32 texton codebook (row-major, filter x codeword) by seeded k-means++
 * and Lloyd iterations on the image's own Schmid responses.  Replaces the
 * codebook file read by SPFeatures::genTextonMap (src/SPFeatures.cpp:317-323). */
int gpb_make_codebook(const uint32_t* packed, int width, int height, uint64_t seed, int iters, int max_samples,
                      float* codebook);

/* ---- stage (a): textons -------------------------------------------------- */

/* Texton label per pixel (0..31): Schmid 13x49x49 correlation with REFLECT
 * border and nearest codebook column.  Replaces SPFeatures::genTextonMap
 * (src/SPFeatures.cpp:277-360).  Host pointers. */
int gpb_texton(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const float* codebook, uint8_t* labels);

/* ---- stage (b): local cues and mPb ------------------------------------------ */

/* Gamma 2.5, RGB->Lab (D65), normalise, quantise to 25 bins; bins[3][N]
 * (L, a, b planes) and nbins[3] = max(bin)+1 per plane.  Replaces
 * rgb_gamma_correct / rgb_to_lab / lab_normalize / quantize_values
 * (globalPb/src/math/libraries/lib_image.cc:123-245, 2315-2344) as called at
 * src/mex_pb_parts_final_selected.cpp:115-134. */
int gpb_lab_quantize(gpb_ctx* ctx, int width, int height, const uint32_t* packed, uint8_t* bins, int* nbins);

/* The 12 half-disc chi-squared cues x 8 orientations and their mPb weighted
 * sum.  cues[12][8][N] (bg r3,5,10; cga, cgb, tg r5,10,20; may be NULL) and
 * mpb_all[8][N].  Replaces det_mPb (src/det_mPb.cpp:9-22) ->
 * mex_pb_parts_final_selected (src/mex_pb_parts_final_selected.cpp:76-227) ->
 * lib_image::hist_gradient_2D (lib_image.cc:3575-3686), and the weighted sum
 * of multiscalePb (src/multiscalePb.cpp:216-245). */
int gpb_local_cues(gpb_ctx* ctx, int width, int height, const uint32_t* packed, const uint8_t* textons, float* cues,
                   float* mpb_all);

/* Max over orientations, oriented nonmax suppression (tolerance pi/8) and the
 * clip to [0,1]: mpb_all[8][N] -> nmax[N].  Replaces nonmax_channels
 * (src/nonmax_channels.cpp:8-80), nonmax_oriented_2D (lib_image.cc:4083-4200)
 * and the clip of src/multiscalePb.cpp:251-261. */
int gpb_nonmax(gpb_ctx* ctx, int width, int height, const float* mpb_all, float* nmax);

#ifdef __cplusplus
}
#endif

#endif /* GPB_CAPI_H */
