// Hop-proposal network for Metropolis sampling of fermion occupations on an
// H x W lattice: 3x3 convolutions (zero padding, NCHW, fp32 activations and
// accumulation, bf16 weights), ReLU, and the masked softmax over the 4 H W
// nearest-neighbour hops (direction d at site s: 0 up, 1 down, 2 left,
// 3 right; valid when s is filled and its neighbour exists and is empty).
// Forward, both backward passes, the loss gradient, Adam and the proposal
// sampling are written here; the caller owns every buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "proposal.hpp"

namespace tns {
namespace {

constexpr int NT = 256;

__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ inline uint16_t f_to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// y[b, co] = bias[co] + sum_ci w[co, ci] (*) x[b, ci]; one workgroup per
// sample; the padded input plane stack and the weights sit in LDS.
__global__ __launch_bounds__(NT) void conv_fwd_kernel(const float* __restrict__ x, const uint16_t* __restrict__ w,
                                                     const float* __restrict__ bias, float* __restrict__ y, int cin,
                                                     int cout, int H, int W, int relu) {
    extern __shared__ float sm[];
    const int HP = H + 2, WP = W + 2;
    float* xs = sm;                       // cin x HP x WP
    float* ws = sm + cin * HP * WP;       // cout x cin x 9
    const int b = blockIdx.x, tid = threadIdx.x;
    const float* xb = x + int64_t(b) * cin * H * W;
    for (int e = tid; e < cin * HP * WP; e += NT) {
        const int c = e / (HP * WP), r = (e / WP) % HP, q = e % WP;
        const int i = r - 1, j = q - 1;
        xs[e] = (i >= 0 && i < H && j >= 0 && j < W) ? xb[(c * H + i) * W + j] : 0.f;
    }
    for (int e = tid; e < cout * cin * 9; e += NT) ws[e] = bf16_to_f(w[e]);
    __syncthreads();
    float* yb = y + int64_t(b) * cout * H * W;
    for (int o = tid; o < cout * H * W; o += NT) {
        const int co = o / (H * W), i = (o / W) % H, j = o % W;
        float acc = bias[co];
        for (int ci = 0; ci < cin; ci++) {
            const float* xp = xs + (ci * HP + i) * WP + j;
            const float* wp = ws + (co * cin + ci) * 9;
#pragma unroll
            for (int ky = 0; ky < 3; ky++)
#pragma unroll
                for (int kx = 0; kx < 3; kx++) acc += wp[ky * 3 + kx] * xp[ky * WP + kx];
        }
        yb[o] = relu ? fmaxf(acc, 0.f) : acc;
    }
}

// dx[b, ci] = sum_co w[co, ci] (*)^T dy[b, co]; dy is masked by the ReLU of
// the layer output y when relu_y is non-null.
__global__ __launch_bounds__(NT) void conv_bwd_data_kernel(const float* __restrict__ dy, const float* __restrict__ relu_y,
                                                           const uint16_t* __restrict__ w, float* __restrict__ dx,
                                                           int cin, int cout, int H, int W) {
    extern __shared__ float sm[];
    const int HP = H + 2, WP = W + 2;
    float* gs = sm;                   // cout x HP x WP (padded, masked dy)
    float* ws = sm + cout * HP * WP;  // cout x cin x 9
    const int b = blockIdx.x, tid = threadIdx.x;
    const float* gb = dy + int64_t(b) * cout * H * W;
    const float* yb = relu_y ? relu_y + int64_t(b) * cout * H * W : nullptr;
    for (int e = tid; e < cout * HP * WP; e += NT) {
        const int c = e / (HP * WP), r = (e / WP) % HP, q = e % WP;
        const int i = r - 1, j = q - 1;
        float v = 0.f;
        if (i >= 0 && i < H && j >= 0 && j < W) {
            const int k = (c * H + i) * W + j;
            v = gb[k];
            if (yb && !(yb[k] > 0.f)) v = 0.f;
        }
        gs[e] = v;
    }
    for (int e = tid; e < cout * cin * 9; e += NT) ws[e] = bf16_to_f(w[e]);
    __syncthreads();
    float* xb = dx + int64_t(b) * cin * H * W;
    for (int o = tid; o < cin * H * W; o += NT) {
        const int ci = o / (H * W), i = (o / W) % H, j = o % W;
        float acc = 0.f;
        for (int co = 0; co < cout; co++) {
            // y(i', j') uses x(i' + ky - 1, j' + kx - 1): x(i, j) feeds
            // y(i - ky + 1, j - kx + 1), i.e. padded row i - ky + 2
            const float* gp = gs + (co * HP + i) * WP + j;
            const float* wp = ws + (co * cin + ci) * 9;
#pragma unroll
            for (int ky = 0; ky < 3; ky++)
#pragma unroll
                for (int kx = 0; kx < 3; kx++) acc += wp[ky * 3 + kx] * gp[(2 - ky) * WP + (2 - kx)];
        }
        xb[o] = acc;
    }
}

// Partial weight / bias gradients over a chunk of samples per workgroup:
// part[chunk][co][ci][9] and partb[chunk][co].
__global__ __launch_bounds__(NT) void conv_bwd_weight_kernel(const float* __restrict__ dy,
                                                             const float* __restrict__ relu_y,
                                                             const float* __restrict__ x, float* __restrict__ part,
                                                             float* __restrict__ partb, int B, int per_chunk, int cin,
                                                             int cout, int H, int W) {
    extern __shared__ float sm[];
    const int HP = H + 2, WP = W + 2;
    float* xs = sm;                   // cin x HP x WP
    float* gs = sm + cin * HP * WP;   // cout x H x W
    const int tid = threadIdx.x, chunk = blockIdx.x;
    const int nw = cout * cin * 9;
    constexpr int MAXPT = 64;  // accumulators per thread: nw <= NT * MAXPT
    float acc[MAXPT];
    float accb = 0.f;
#pragma unroll
    for (int q = 0; q < MAXPT; q++) acc[q] = 0.f;
    const int b0 = chunk * per_chunk, b1 = min(B, b0 + per_chunk);
    for (int b = b0; b < b1; b++) {
        const float* xb = x + int64_t(b) * cin * H * W;
        const float* gb = dy + int64_t(b) * cout * H * W;
        const float* yb = relu_y ? relu_y + int64_t(b) * cout * H * W : nullptr;
        __syncthreads();
        for (int e = tid; e < cin * HP * WP; e += NT) {
            const int c = e / (HP * WP), r = (e / WP) % HP, q = e % WP;
            const int i = r - 1, j = q - 1;
            xs[e] = (i >= 0 && i < H && j >= 0 && j < W) ? xb[(c * H + i) * W + j] : 0.f;
        }
        for (int e = tid; e < cout * H * W; e += NT) {
            float v = gb[e];
            if (yb && !(yb[e] > 0.f)) v = 0.f;
            gs[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXPT; q++) {
            const int k = tid + q * NT;
            if (k < nw) {
                const int co = k / (cin * 9), ci = (k / 9) % cin, kk = k % 9, ky = kk / 3, kx = kk % 3;
                const float* gp = gs + co * H * W;
                const float* xp = xs + (ci * HP + ky) * WP + kx;
                float s = 0.f;
                for (int i = 0; i < H; i++)
                    for (int j = 0; j < W; j++) s += gp[i * W + j] * xp[i * WP + j];
                acc[q] += s;
            }
        }
        if (tid < cout) {
            const float* gp = gs + tid * H * W;
            float s = 0.f;
            for (int e = 0; e < H * W; e++) s += gp[e];
            accb += s;
        }
    }
#pragma unroll
    for (int q = 0; q < MAXPT; q++) {
        const int k = tid + q * NT;
        if (k < nw) part[int64_t(chunk) * nw + k] = acc[q];
    }
    if (tid < cout) partb[int64_t(chunk) * cout + tid] = accb;
}

__global__ void reduce_chunks_kernel(const float* __restrict__ part, int chunks, int n, float* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float s = 0.f;
    for (int c = 0; c < chunks; c++) s += part[int64_t(c) * n + k];
    out[k] = s;
}

// hop h = d * H W + s: valid when site s is filled and its neighbour in
// direction d exists and is empty; returns the neighbour or -1
__device__ inline int hop_target(const int8_t* occ, int h, int H, int W) {
    const int HW = H * W, d = h / HW, s = h % HW, i = s / W, j = s % W;
    int ti = i, tj = j;
    if (d == 0) ti--;
    else if (d == 1) ti++;
    else if (d == 2) tj--;
    else tj++;
    if (ti < 0 || ti >= H || tj < 0 || tj >= W) return -1;
    const int t = ti * W + tj;
    return (occ[s] == 1 && occ[t] == 0) ? t : -1;
}

__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline float wave_sumf(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// one wave per chain: masked log-softmax over the hops; with u != null draw
// a hop by inverse CDF (u in [0, 1)), else score the given hop.  logq gets
// log q(hop); hop gets -1 if no hop is valid.
__global__ __launch_bounds__(64) void hop_kernel(const float* __restrict__ logits, const int8_t* __restrict__ occ,
                                                 const float* __restrict__ u, int32_t* __restrict__ hop,
                                                 float* __restrict__ logq, int H, int W) {
    const int b = blockIdx.x, lane = threadIdx.x;
    const int n = 4 * H * W;
    const float* lg = logits + int64_t(b) * n;
    const int8_t* oc = occ + int64_t(b) * H * W;
    float mx = -INFINITY;
    for (int h = lane; h < n; h += 64)
        if (hop_target(oc, h, H, W) >= 0) mx = fmaxf(mx, lg[h]);
    mx = wave_max(mx);
    if (mx == -INFINITY) {
        if (lane == 0) {
            if (u) hop[b] = -1;
            logq[b] = 0.f;
        }
        return;
    }
    float z = 0.f;
    for (int h = lane; h < n; h += 64)
        if (hop_target(oc, h, H, W) >= 0) z += expf(lg[h] - mx);
    z = wave_sumf(z);
    if (!u) {
        if (lane == 0) {
            const int h = hop[b];
            logq[b] = (h >= 0 && h < n && hop_target(oc, h, H, W) >= 0) ? lg[h] - mx - logf(z) : -INFINITY;
        }
        return;
    }
    // inverse CDF in hop order, done by lane 0 over the (small) hop list
    if (lane == 0) {
        const float target = u[b] * z;
        float c = 0.f;
        int pick = -1, last = -1;
        for (int h = 0; h < n; h++) {
            if (hop_target(oc, h, H, W) < 0) continue;
            last = h;
            c += expf(lg[h] - mx);
            if (c > target) {
                pick = h;
                break;
            }
        }
        if (pick < 0) pick = last;
        hop[b] = pick;
        logq[b] = lg[pick] - mx - logf(z);
    }
}

// regression of the hop logit onto its target: loss_b = (l_h - t_b)^2,
// dlogits = 2 (l_h - t_b) / B at h, zero elsewhere
__global__ void hop_regression_kernel(const float* __restrict__ logits, const int32_t* __restrict__ hop,
                                      const float* __restrict__ target, float* __restrict__ dlogits,
                                      float* __restrict__ loss, int B, int n) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= int64_t(B) * n) return;
    const int b = int(e / n), h = int(e % n);
    float g = 0.f;
    if (h == hop[b]) {
        const float d = logits[e] - target[b];
        g = 2.f * d / float(B);
        loss[b] = d * d;
    }
    dlogits[e] = g;
}

__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, uint16_t* __restrict__ wbf, int n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    w[k] = wk;
    if (wbf) wbf[k] = f_to_bf16(wk);
}

__global__ void to_bf16_kernel(const float* __restrict__ w, uint16_t* __restrict__ wbf, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) wbf[k] = f_to_bf16(w[k]);
}

__global__ void occ_to_input_kernel(const int8_t* __restrict__ occ, float* __restrict__ x, int64_t n) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) x[k] = float(occ[k]);
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void allow(const void* k, size_t bytes) {
    if (bytes > 64 * 1024) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
}

}  // namespace

void conv_check(int cin, int cout, int H, int W) {
    if (cin < 1 || cout < 1 || H < 1 || W < 1 || cin > 64 || cout > 64 || H > 32 || W > 32)
        throw std::runtime_error("conv3x3: channels must be in [1, 64] and the lattice at most 32 x 32");
    if (cout * cin * 9 > NT * 64) throw std::runtime_error("conv3x3: too many weights for the gradient kernel");
}

void launch_conv_fwd(const float* x, const uint16_t* w, const float* bias, float* y, int B, int cin, int cout, int H,
                     int W, int relu, hipStream_t st) {
    conv_check(cin, cout, H, W);
    if (B <= 0) return;
    const size_t lds = size_t(cin * (H + 2) * (W + 2) + cout * cin * 9) * sizeof(float);
    if (lds > 160 * 1024) throw std::runtime_error("conv3x3: working set exceeds LDS");
    allow(reinterpret_cast<const void*>(conv_fwd_kernel), lds);
    hipLaunchKernelGGL(conv_fwd_kernel, dim3(unsigned(B)), dim3(NT), lds, st, x, w, bias, y, cin, cout, H, W, relu);
    check(hipGetLastError(), "conv forward launch");
}

void launch_conv_bwd_data(const float* dy, const float* relu_y, const uint16_t* w, float* dx, int B, int cin, int cout,
                          int H, int W, hipStream_t st) {
    conv_check(cin, cout, H, W);
    if (B <= 0) return;
    const size_t lds = size_t(cout * (H + 2) * (W + 2) + cout * cin * 9) * sizeof(float);
    if (lds > 160 * 1024) throw std::runtime_error("conv3x3: working set exceeds LDS");
    allow(reinterpret_cast<const void*>(conv_bwd_data_kernel), lds);
    hipLaunchKernelGGL(conv_bwd_data_kernel, dim3(unsigned(B)), dim3(NT), lds, st, dy, relu_y, w, dx, cin, cout, H, W);
    check(hipGetLastError(), "conv backward-data launch");
}

int64_t conv_bwd_weight_workspace(int B, int cin, int cout) {
    const int chunks = (B + CONV_CHUNK - 1) / CONV_CHUNK;
    return int64_t(chunks) * (int64_t(cout) * cin * 9 + cout) * int64_t(sizeof(float));
}

void launch_conv_bwd_weight(const float* dy, const float* relu_y, const float* x, float* dw, float* db, int B, int cin,
                            int cout, int H, int W, float* workspace, hipStream_t st) {
    conv_check(cin, cout, H, W);
    if (B <= 0) return;
    const int chunks = (B + CONV_CHUNK - 1) / CONV_CHUNK;
    const int nw = cout * cin * 9;
    float* part = workspace;
    float* partb = workspace + int64_t(chunks) * nw;
    const size_t lds = size_t(cin * (H + 2) * (W + 2) + cout * H * W) * sizeof(float);
    if (lds > 160 * 1024) throw std::runtime_error("conv3x3: working set exceeds LDS");
    allow(reinterpret_cast<const void*>(conv_bwd_weight_kernel), lds);
    hipLaunchKernelGGL(conv_bwd_weight_kernel, dim3(unsigned(chunks)), dim3(NT), lds, st, dy, relu_y, x, part, partb, B,
                       CONV_CHUNK, cin, cout, H, W);
    check(hipGetLastError(), "conv backward-weight launch");
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3(unsigned((nw + 255) / 256)), dim3(256), 0, st, part, chunks, nw, dw);
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3(unsigned((cout + 255) / 256)), dim3(256), 0, st, partb, chunks, cout,
                       db);
    check(hipGetLastError(), "conv gradient reduction launch");
}

void launch_hop(const float* logits, const int8_t* occ, const float* u, int32_t* hop, float* logq, int B, int H, int W,
                hipStream_t st) {
    if (B <= 0) return;
    if (H < 1 || W < 1 || H > 32 || W > 32) throw std::runtime_error("hop: lattice at most 32 x 32");
    hipLaunchKernelGGL(hop_kernel, dim3(unsigned(B)), dim3(64), 0, st, logits, occ, u, hop, logq, H, W);
    check(hipGetLastError(), "hop launch");
}

void launch_hop_regression(const float* logits, const int32_t* hop, const float* target, float* dlogits, float* loss,
                           int B, int n, hipStream_t st) {
    if (B <= 0) return;
    const int64_t total = int64_t(B) * n;
    hipLaunchKernelGGL(hop_regression_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, logits, hop, target,
                       dlogits, loss, B, n);
    check(hipGetLastError(), "hop regression launch");
}

void launch_adam(float* w, float* m, float* v, const float* g, uint16_t* wbf, int n, float lr, float b1, float b2,
                 float eps, int step, hipStream_t st) {
    if (n <= 0) return;
    const float c1 = 1.f - std::pow(b1, float(step)), c2 = 1.f - std::pow(b2, float(step));
    hipLaunchKernelGGL(adam_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, w, m, v, g, wbf, n, lr, b1, b2,
                       eps, c1, c2);
    check(hipGetLastError(), "adam launch");
}

void launch_to_bf16(const float* w, uint16_t* wbf, int n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, w, wbf, n);
    check(hipGetLastError(), "bf16 launch");
}

void launch_occ_to_input(const int8_t* occ, float* x, int64_t n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(occ_to_input_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, occ, x, n);
    check(hipGetLastError(), "input launch");
}

}  // namespace tns
