// windhip_net.hpp -- learned initial guess for the wind solve: a 3-D
// convolutional encoder-decoder (4 scales, 16/32/64/64 channels, 3x3x3
// kernels) on bf16 MFMA with FP32 accumulation. Internal interface shared by
// windhip_net.hip (network) and windhip.hip (solver hook).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace wnet {

constexpr int kInCh = 8;      // 6 input fields + 2 zero channels (8-channel granularity of the loads)
constexpr int kNumConv = 14;  // 3x3x3 convolutions; plus the 1x1x1 head

// Batch of volumes, channels-last: voxel v = ((b*D + z)*H + y)*W + x.
struct Shape {
    int B, D, H, W;
    __host__ __device__ long voxels() const { return (long)B * D * H * W; }
    __host__ __device__ Shape half() const { return Shape{B, D / 2, H / 2, W / 2}; }
};

struct ConvSpec {
    int ci, co;  // channels
    int level;   // resolution level 0..3
};

// The fixed architecture: encoder c0a c0b | c1a c1b | c2a c2b | c3a c3b,
// decoder d2a d2b | d1a d1b | d0a d0b, head 16 -> 1.
extern const ConvSpec kConvs[kNumConv];

// Parameter layout (flat FP32): per convolution W[co][t][ci] (t = 9 (dz+1) +
// 3 (dy+1) + (dx+1)) then bias[co]; then the head w[16], b.
long param_count();

}  // namespace wnet

struct windhip_net;

// Forward pass of a batch already assembled on the device: in [B][D][H][W][8]
// FP32 (channels-last, padded dims multiples of 8), out [B][D][H][W] FP32.
int wnet_forward_dev(windhip_net *net, hipStream_t stream, const float *in, wnet::Shape s, float *out);
// The same with the input already in bf16 ([B][D][H][W][8], the rounding the
// forward pass applies to an FP32 input): no conversion pass.
int wnet_forward_dev_bf16(windhip_net *net, hipStream_t stream, const __bf16 *in, wnet::Shape s, float *out);
