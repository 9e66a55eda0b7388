// unet.hpp — internal interface of the denoiser: kernel launchers
// (unet_kernels.hip) and the U-Net that strings them together (unet.hip).
#pragma once

#include "internal.hpp"

namespace aur {

// Convolution input: full-resolution A (Ca channels) concatenated with a
// half-resolution B (Cb channels) read through nearest upsampling.
struct ConvIn {
    const uint16_t* a;
    const uint16_t* b;
    int ca, cb;
};

Status conv3x3_forward(const ConvIn& in, const uint16_t* wk, int kpad, const float* bias, void* out, bool out_f32,
                       bool accum, bool relu, int N, int H, int W, int cout, hipStream_t s);
Status conv3x3_weight_grad(const ConvIn& in, const uint16_t* g, int cout, int N, int H, int W, float* dw, float* db,
                           DevBuf& scratch, hipStream_t s);
Status mask_to_bf16(const float* grad, const uint16_t* act, size_t n, uint16_t* out, hipStream_t s);
Status avgpool2(const uint16_t* in, int N, int H, int W, int C, uint16_t* out, hipStream_t s);
Status avgpool2_backward(const float* gout, int N, int H, int W, int C, float* gin, hipStream_t s);
Status split_concat_grad(const float* d, int N, int H, int W, int ca, int cb, float* ga, float* gb, bool gb_accum,
                         hipStream_t s);
Status prep_input(const float* rgb, const float* alb, const float* nrm, size_t npix, float inv_spp, uint16_t* x,
                  float* noisy_log, hipStream_t s);
Status output_image(const float* out16, const float* noisy_log, size_t npix, float* img, hipStream_t s);
Status loss_grad(const float* out16, const float* noisy_log, const float* target, size_t npix, int kind,
                 float* g16f, float* loss_partial, hipStream_t s);
Status adam_step(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                 int t, hipStream_t s);
Status pack_weights(const float* w, int cout, int cin, int kpad, int kpad_t, uint16_t* wk, uint16_t* wt,
                    hipStream_t s);
Status f32_to_bf16(const float* in, size_t n, uint16_t* out, hipStream_t s);
Status mfma_selftest(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s);

// 4-level U-Net (channels 32/48/64/64), 15 3x3 convolutions.
class Unet {
   public:
    struct Layer {
        int a, b;          // activation slots of the A / B inputs (b < 0: none)
        int out;           // activation slot written
        int ca, cb, cout;  // channels
        int level;         // resolution level 0..3
        bool relu;
        int kpad, kpad_t;
        size_t w_off, b_off;
        DevBuf wk, wt;     // BF16 forward / data-gradient weight copies
    };
    explicit Unet(hipStream_t s) : s_(s) {}
    Status init(uint32_t seed);
    size_t param_count() const { return n_params_; }
    Status get_params(float* host);
    Status set_params(const float* host);
    // forward on N images of H x W (H, W multiples of 8); out16 kept
    Status forward(const float* rgb, const float* alb, const float* nrm, float inv_spp, int N, int H, int W);
    Status write_output(float* img);
    // loss against targets (mean radiance) and gradients into grads_
    Status backward(const float* target, float* loss);
    Status adam(float lr);
    void set_loss(int kind) { loss_kind_ = kind; }  // 0 log-space L2, 1 relative L2
    Status grads_to_host(float* host);
    void release();

   private:
    Status alloc(int N, int H, int W);
    Status repack();
    hipStream_t s_;
    std::vector<Layer> layers_;
    size_t n_params_ = 0;
    DevBuf params_, grads_, m_, v_;
    int N_ = 0, H_ = 0, W_ = 0;
    int t_ = 0;
    int loss_kind_ = 0;
    static constexpr int kSlots = 19;  // x, 14 hidden, 3 pools, out
    DevBuf act_[kSlots];               // BF16 activations (slot 18: FP32 out16)
    DevBuf grad_[kSlots];              // FP32 gradients of activations
    DevBuf noisy_log_, gbf_, dcat_, scratch_, loss_part_;
    int slot_c_[kSlots] = {0};
    int slot_lv_[kSlots] = {0};
};

}  // namespace aur
