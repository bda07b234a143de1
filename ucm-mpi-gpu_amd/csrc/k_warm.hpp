// Learned warm start for block Lanczos: host API of the convolutional
// encoder-decoder in k_warm.hip (no counterpart in the reference, whose
// eigensolvers start from a fixed or random vector: spectralPb.cpp:221-244,
// damascene/src/noReorthog/lanczos.cu:393-752).
#pragma once
#include <cstdint>

#include "gpb_pipeline.hpp"

namespace gpb {

// Network input: 32 cue channels (4 cues x 8 orientations, each the
// scale-weighted mPb contribution averaged over 4x4 pixel cells) plus two
// cell-coordinate channels, zero-padded to 64; output: 8 vectors per cell.
constexpr int kWarmIn = 64;
constexpr int kWarmFeat = 32;
constexpr int kWarmOut = 8;
constexpr int kWarmLayers = 12;

size_t warm_param_count();
// Offsets of layer l's weights W[9][cout][cin] and bias[cout] in the
// parameter vector, and its channel counts.
void warm_layer_info(int l, int* cin, int* cout, int* relu, size_t* w_off, size_t* b_off);

void warm_init(Context& c, uint64_t seed);
void warm_set_params(Context& c, const float* host);
void warm_get_params(Context& c, float* host);
bool warm_ready(const Context& c);

// Cell grid of a W x H image: ceil(H/4) x ceil(W/4) cells, padded to
// multiples of 8 for the three 2x poolings.
void warm_cell_grid(int W, int H, int* hq, int* wq, int* h0, int* w0);

// Predicted starting block from the cue features d_feat ([32][hq][wq] FP32,
// see launch_local_cues): d_block [8][Np] in the padded eigensolver frame
// (generalised vectors, nearest-cell upsampling, zero halo).
void warm_predict_block(Context& c, const float* d_feat, int W, int H, float* d_block, size_t Np);

// Training set held on the device: all images W x H.
void warm_dataset_reset(Context& c, int W, int H, int capacity);
// Appends one image: features and the D-weighted cell sums of eigenvectors
// 1..8 of d_evecs ([K][Np], K >= 9, padded frame) with the degrees st.D.
void warm_dataset_add(Context& c, const float* d_feat, const Stencil& st, const float* d_evecs);
int warm_dataset_size(const Context& c);
// Adam over dataset entries [first, first + count) for `epochs` passes of
// shuffled mini-batches, each image under a random symmetry of the square
// cell grid when augment != 0; epoch_loss[e] (may be null) receives the mean
// training loss of epoch e.
void warm_train(Context& c, int first, int count, int epochs, int batch, float lr, uint64_t seed, int augment,
                float* epoch_loss);
// Mean loss over dataset entries [first, first + count).
double warm_eval_loss(Context& c, int first, int count);
// Loss and parameter gradient for host inputs at cell resolution (hq, wq
// multiples of 8): x [n][hq][wq][34] (32 cues + 2 coordinates), T [n][hq][wq][8],
// Db [n][hq][wq]; y (optional) receives the network output [n][hq][wq][8].
double warm_loss_grad_host(Context& c, int n, int hq, int wq, const float* x, const float* T, const float* Db,
                           float* grad, float* y);

}  // namespace gpb
