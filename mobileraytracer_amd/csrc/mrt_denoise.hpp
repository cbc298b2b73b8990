// Denoiser kernels (mrt_denoise.hip): BF16 NHWC tensors, 3x3 convolutions on
// MFMA with FP32 accumulation and 2x2 max pooling.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrt {

// One 3x3 convolution (stride 1, zero padding 1) over an H x W frame.
//
// Input, cin channels (a multiple of 32), in 32-channel chunks:
//   * channels [0, cu) (cu a multiple of 32, possibly 0): xu, an H/2 x W/2
//     tensor of cu channels per pixel, upsampled 2x (nearest);
//   * channels [cu, cin): channels [x_off, x_off + cin - cu) of x (x_stride
//     channels per pixel);
//   * or, when rad is set (cin = 32, cu = 0): the network input packed from a
//     frame of fw x fh pixels at the top left of the H x W frame, zero
//     elsewhere: log(1 + max(rad * rad_scale, 0)) (3 channels), aov[0..5],
//     aov[6] * depth_scale, then 22 zero channels.
// Weights W[tap][cout][cin] BF16 (tap = 3 * dy + dx for offsets dy - 1,
// dx - 1) in fragment order, wf (frag_index; launch_frag_weights reorders
// [9][cout][cin] weights): each 16-output-channel x 32-input-channel block of
// a tap is one contiguous 1 KB, which a wave loads in one coalesced access
// (rows cin apart, the loads took up to 45 % longer on the coarse layers);
// FP32 bias (or null); ReLU when relu is set.
//
// Output (cout a multiple of 32):
//   * BF16 channels [y_off, y_off + cout) of y (y_stride per pixel);
//   * or, when rad_out is set: FP32 radiance max(exp(v) - 1, 0) of output
//     channels 0..2 for the pixels of the fw x fh frame (3 floats per pixel,
//     row length fw); the other channels are dropped.
// Offsets and strides are multiples of 8.
//
// Batches (training): batch > 1 frames of H x W pixels, each tensor's frames
// one after the other (x: H W x_stride elements apart, xu: H W cu / 4, y and
// mask: H W times their strides); only for tensor inputs and outputs.
//
// Masked output (training, the gradient through a ReLU): when mask is set,
// output channel k of a pixel is zeroed unless channel mask_off + k of the
// same pixel of mask (mask_stride channels per pixel, BF16) is above zero.
// Position of weight (tap, co, ci) of a cin x cout layer in fragment order:
// tap, then 16-channel output block, then 32-channel input chunk, each block
// 64 lanes x 8 channels, lane 16 lg + li holding output channel 16 n + li and
// input channels 32 k + 8 lg .. + 7 (the B operand of
// v_mfma_f32_16x16x32_bf16).
__host__ __device__ inline size_t frag_index(uint32_t tap, uint32_t co, uint32_t ci, uint32_t cin, uint32_t cout) {
    return ((static_cast<size_t>(tap) * (cout >> 4) + (co >> 4)) * (cin >> 5) + (ci >> 5)) * 512u +
           (((ci >> 3) & 3u) * 16u + (co & 15u)) * 8u + (ci & 7u);
}

struct ConvArgs {
    const uint16_t* x;
    uint32_t x_stride, x_off, cin;
    const uint16_t* xu;
    uint32_t cu;
    const float* rad;
    const float* aov;
    float rad_scale, depth_scale;
    const uint16_t* wf;  // weights in fragment order
    const float* bias;
    uint32_t cout;
    uint16_t* y;
    uint32_t y_stride, y_off;
    float* rad_out;
    uint32_t fw, fh;
    uint32_t H, W;
    uint32_t relu;
    uint32_t batch;  // frames (0 or 1: one)
    // Tensor outputs only: also the 2x2 max pooling of the output (H/2 x W/2
    // pixels of cout channels per frame; H and W even), what launch_pool2
    // would write from y.
    uint16_t* pool;
    const uint16_t* mask;
    uint32_t mask_stride, mask_off;
};

// min_blocks: the channel block per workgroup (32 or 64 output channels) is
// the widest that still gives this many workgroups.
hipError_t launch_conv3x3(const ConvArgs& a, uint32_t min_blocks, hipStream_t stream);
// w ([9][cout][cin] BF16) into fragment order (frag_index) in wf.
hipError_t launch_frag_weights(const uint16_t* w, uint32_t cin, uint32_t cout, uint16_t* wf, hipStream_t stream);
// 2x2 max pooling: x (H x W, c channels per pixel, H and W even, c a
// multiple of 8) into y (H/2 x W/2, c channels per pixel).  A batch of frames
// of even height pools as one frame of the summed height.
hipError_t launch_pool2(const uint16_t* x, uint32_t c, uint32_t H, uint32_t W, uint16_t* y, hipStream_t stream);

// ---- training (mrt_denoise_train.hip) --------------------------------------
//
// Octet layout, the operand layout of the weight-gradient kernel: a tensor of
// B frames, R rows, W pixels per row (W a multiple of 8) and C channels, with
// element (b, r, x, c) at (((b R + r) (W / 8) + x / 8) C + c) 8 + x % 8: eight
// consecutive pixels of one channel are 16 contiguous bytes.

// What a convolution reads, for the octet transpose: channels [0, cu) of xu
// upsampled 2x (nearest), then channels [x_off, x_off + c - cu) of x.
struct TensorIn {
    const uint16_t* x;
    uint32_t x_stride, x_off;
    const uint16_t* xu;
    uint32_t cu;
    uint32_t c;  // channels (multiple of 8)
};

// Octet copy of a tensor (rowpad = 0: R = H rows) or of a convolution input
// with a zero row above and below each frame (rowpad = 1: R = H + 2 rows, row
// r holding input row r - 1).
hipError_t launch_octet(const TensorIn& in, uint32_t batch, uint32_t H, uint32_t W, int rowpad, uint16_t* out,
                        hipStream_t stream);

// Weight gradients of one 3x3 convolution: partial sums per K split of
// dW[tap][co][ci] = sum over pixels of dZ(pixel, co) X(pixel + tap offset, ci)
// and db[co] = sum of dZ(pixel, co), from dZ in octet layout (B x H x W, cout
// channels) and X as the row-padded octet copy (cin channels); FP32
// accumulation on MFMA.  part_w: [splits][9][cout][cin], part_b:
// [splits][cout]; `splits` is chosen here (at most max_splits) and returned.
// K steps of 4 segments per split and the number of splits launch_wgrad
// uses: about target_groups workgroups (of 4 waves), at least 8 K steps
// (256 pixels) per wave, at most max_splits splits.
inline uint32_t wgrad_splits(uint32_t batch, uint32_t H, uint32_t W, uint32_t cin, uint32_t cout, uint32_t max_splits,
                             uint32_t target_groups, uint32_t* kps) {
    const uint32_t nseg = batch * H * (W / 8u), ksteps = (nseg + 3u) / 4u;
    const uint32_t per_split = 3u * (cin / 32u) * (cout / 32u);
    uint32_t s = (target_groups + per_split - 1) / per_split;
    if (s > max_splits) s = max_splits;
    if (s > (ksteps + 31u) / 32u) s = (ksteps + 31u) / 32u;
    if (s < 1u) s = 1u;
    const uint32_t k = (ksteps + s - 1) / s;
    if (kps) *kps = k;
    return (ksteps + k - 1) / k;
}
hipError_t launch_wgrad(const uint16_t* dzt, const uint16_t* xt, uint32_t batch, uint32_t H, uint32_t W, uint32_t cin,
                        uint32_t cout, uint32_t max_splits, uint32_t target_groups, float* part_w, float* part_b,
                        uint32_t* splits, hipStream_t stream);
// out[i] = sum over s < splits of part[s * n + i] (in order of s).
hipError_t launch_sum_splits(const float* part, uint32_t splits, size_t n, float* out, hipStream_t stream);

// Loss of a batch: y (n pixels x 32 BF16 channels, 0..2 used), target t (n x
// 3 FP32): loss_sum += sum of (y - t)^2 (FP64 accumulation), dy = BF16((y -
// t) scale) in channels 0..2 of 32 (the others zero).
hipError_t launch_mse(const uint16_t* y, const float* t, size_t n, float scale, uint16_t* dy, double* loss_sum,
                      hipStream_t stream);

// Gradient of an encoder output e (n_rows x W pixels, c channels; rows in
// pairs) that feeds both a 2x2 max pooling and a skip connection:
// dz = BF16((dskip + routed dp) * (e > 0)), dp (pooled grid, c channels)
// going to the first position of its window, in row-major order, that holds
// the window's maximum; dskip channels at dskip_off of dskip_stride per pixel.
hipError_t launch_pool2_back(const uint16_t* e, const uint16_t* dp, const uint16_t* dskip, uint32_t dskip_stride,
                             uint32_t dskip_off, uint32_t c, uint32_t n_rows, uint32_t W, uint16_t* dz,
                             hipStream_t stream);
// Gradient of the coarse tensor of a decoder's 2x upsampling:
// dz(y, x) = BF16((g(2y, 2x) + g(2y, 2x + 1) + g(2y + 1, 2x) + g(2y + 1, 2x + 1)) * (coarse(y, x) > 0)),
// g = channels [0, c) of dcat (dcat_stride per pixel, n_rows x W fine grid).
hipError_t launch_up2_back(const uint16_t* dcat, uint32_t dcat_stride, const uint16_t* coarse, uint32_t c,
                           uint32_t n_rows, uint32_t W, uint16_t* dz, hipStream_t stream);

// One crop of a training batch: frame n of the data set, top-left (oy, ox),
// mirrored left-right (fx) and/or top-bottom (fy).
struct Crop {
    uint32_t n, oy, ox, fx, fy;
};
constexpr uint32_t kMaxCrops = 64;
struct CropTable {
    Crop c[kMaxCrops];
};
// Gathers `batch` crops of side S from a data set of frames (fw x fh pixels):
// packed inputs xin (10 FP32 per pixel) to x (S x S x 32 BF16 per crop,
// channels 10..31 zero) and targets tin (3 FP32 per pixel) to t.
hipError_t launch_gather(const float* xin, const float* tin, uint32_t fw, uint32_t fh, const CropTable& crops,
                         uint32_t batch, uint32_t S, uint16_t* x, float* t, hipStream_t stream);

// Adam over a flat FP32 parameter vector (n_w convolution weights in the
// BF16 weight layout [tap][cout][cin] per layer, then the biases): the update
// of torch.optim.Adam (m += omb1 (g - m); v = v b2 + (omb2 g) g;
// p += -step_size (m / (sqrt(v) / bc2_sqrt + eps)), omb = 1 - beta and the
// others rounded to FP32 from double as torch rounds them), then the BF16 forward
// weights wq and the transposed, tap-flipped weights wt of the
// backward-data convolutions ([8 - tap][cin][cout] per layer), both in
// fragment order (wqf, wtf).
struct AdamLayer {
    uint32_t w_off, cin, cout;
};
constexpr uint32_t kMaxLayers = 19;
struct AdamArgs {
    float* p;
    const float* g;
    float* m;
    float* v;
    uint16_t* wqf;
    uint16_t* wtf;
    size_t n, n_w;
    float omb1, b2, omb2, eps, step_size, bc2_sqrt;
    uint32_t update;  // 0: only write wqf and wtf from p
    uint32_t n_layers;
    AdamLayer layer[kMaxLayers];
};
hipError_t launch_adam(const AdamArgs& a, hipStream_t stream);

}  // namespace mrt
