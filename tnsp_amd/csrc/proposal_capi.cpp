// C ABI of the hop-proposal network kernels (declared in include/tnsgpu.h).
#include <stdexcept>

#include "handles.hpp"
#include "proposal.hpp"

extern "C" {

int tns_conv3x3_forward(const float* x, const uint16_t* w, const float* bias, float* y, int batch, int cin, int cout,
                        int height, int width, int relu, void* stream) {
    TNS_TRY(tns::launch_conv_fwd(x, w, bias, y, batch, cin, cout, height, width, relu,
                                 static_cast<hipStream_t>(stream)))
}

int tns_conv3x3_backward_data(const float* dy, const float* relu_y, const uint16_t* w, float* dx, int batch, int cin,
                              int cout, int height, int width, void* stream) {
    TNS_TRY(tns::launch_conv_bwd_data(dy, relu_y, w, dx, batch, cin, cout, height, width,
                                      static_cast<hipStream_t>(stream)))
}

int tns_conv3x3_backward_weight_workspace(int batch, int cin, int cout, int64_t* bytes) {
    TNS_TRY({
        if (!bytes) throw std::runtime_error("null output");
        *bytes = tns::conv_bwd_weight_workspace(batch, cin, cout);
    })
}

int tns_conv3x3_backward_weight(const float* dy, const float* relu_y, const float* x, float* dw, float* dbias,
                                int batch, int cin, int cout, int height, int width, void* workspace, void* stream) {
    TNS_TRY(tns::launch_conv_bwd_weight(dy, relu_y, x, dw, dbias, batch, cin, cout, height, width,
                                        static_cast<float*>(workspace), static_cast<hipStream_t>(stream)))
}

int tns_hop_sample(const float* logits, const int8_t* occupations, const float* uniforms, int32_t* hops,
                   float* log_q, int batch, int height, int width, void* stream) {
    TNS_TRY({
        if (!uniforms) throw std::runtime_error("tns_hop_sample needs uniforms");
        tns::launch_hop(logits, occupations, uniforms, hops, log_q, batch, height, width,
                        static_cast<hipStream_t>(stream));
    })
}

int tns_hop_log_prob(const float* logits, const int8_t* occupations, int32_t* hops, float* log_q, int batch,
                     int height, int width, void* stream) {
    TNS_TRY(tns::launch_hop(logits, occupations, nullptr, hops, log_q, batch, height, width,
                            static_cast<hipStream_t>(stream)))
}

int tns_hop_regression(const float* logits, const int32_t* hops, const float* targets, float* dlogits, float* loss,
                       int batch, int n, void* stream) {
    TNS_TRY(tns::launch_hop_regression(logits, hops, targets, dlogits, loss, batch, n,
                                       static_cast<hipStream_t>(stream)))
}

int tns_adam_step(float* w, float* m, float* v, const float* grad, uint16_t* w_bf16, int n, float lr, float beta1,
                  float beta2, float eps, int step, void* stream) {
    TNS_TRY(tns::launch_adam(w, m, v, grad, w_bf16, n, lr, beta1, beta2, eps, step, static_cast<hipStream_t>(stream)))
}

int tns_to_bf16(const float* w, uint16_t* w_bf16, int n, void* stream) {
    TNS_TRY(tns::launch_to_bf16(w, w_bf16, n, static_cast<hipStream_t>(stream)))
}

int tns_occupations_to_input(const int8_t* occupations, float* x, int64_t n, void* stream) {
    TNS_TRY(tns::launch_occ_to_input(occupations, x, n, static_cast<hipStream_t>(stream)))
}

}  // extern "C"
