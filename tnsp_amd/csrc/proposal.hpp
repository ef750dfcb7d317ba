// Hop-proposal network kernels (proposal.hip): 3x3 convolutions with bf16
// weights and fp32 activations / accumulation, masked hop softmax, loss
// gradient and Adam.  Buffers are caller-owned device pointers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tns {

constexpr int CONV_CHUNK = 64;  // samples per workgroup of the weight gradient

void launch_conv_fwd(const float* x, const uint16_t* w, const float* bias, float* y, int B, int cin, int cout, int H,
                     int W, int relu, hipStream_t st);
void launch_conv_bwd_data(const float* dy, const float* relu_y, const uint16_t* w, float* dx, int B, int cin, int cout,
                          int H, int W, hipStream_t st);
int64_t conv_bwd_weight_workspace(int B, int cin, int cout);
void launch_conv_bwd_weight(const float* dy, const float* relu_y, const float* x, float* dw, float* db, int B, int cin,
                            int cout, int H, int W, float* workspace, hipStream_t st);
void launch_hop(const float* logits, const int8_t* occ, const float* u, int32_t* hop, float* logq, int B, int H, int W,
                hipStream_t st);
void launch_hop_regression(const float* logits, const int32_t* hop, const float* target, float* dlogits, float* loss,
                           int B, int n, hipStream_t st);
void launch_adam(float* w, float* m, float* v, const float* g, uint16_t* wbf, int n, float lr, float b1, float b2,
                 float eps, int step, hipStream_t st);
void launch_to_bf16(const float* w, uint16_t* wbf, int n, hipStream_t st);
void launch_occ_to_input(const int8_t* occ, float* x, int64_t n, hipStream_t st);

}  // namespace tns
