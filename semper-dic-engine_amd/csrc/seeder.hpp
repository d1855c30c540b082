// Learned seeding (no reference counterpart): a small CNN that maps a
// reference / deformed patch pair to the 6-parameter first-order warp of the
// deformed patch, trained on the GPU by the engine's own forward and backward
// kernels on pairs the engine renders itself (its synthetic speckle, warped by
// random affine maps through a Keys bicubic interpolator). IC-GN always refines
// its output; the plain anchor seeding is the fallback wherever the refined
// solve fails.
//
// Network (NCHW, FP32): input 2 x 64 x 64 (the two patches at pyramid level 2,
// each normalised to zero mean and unit variance); four 3x3 convolutions,
// stride 2, padding 1, 32 channels, ReLU (64 -> 32 -> 16 -> 8 -> 4); a linear
// layer 512 -> 6. Output: (u, v) / kUvScale in full-resolution pixels and
// (ux, uy, vx, vy) / kGradScale, at the patch centre.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace sdic {

struct SeederShape {
    static constexpr int kPatch = 64;        // level-2 pixels per side
    static constexpr int kLevel = 2;         // pyramid level of the patches
    static constexpr int kScale = 1 << kLevel;
    static constexpr int kCh = 32;
    static constexpr int kInCh = 2;
    static constexpr int kFeat = kCh * 4 * 4;
    static constexpr int kOut = 6;
    static constexpr float kUvScale = 16.0f;
    static constexpr float kGradScale = 0.1f;
    // flat parameter layout: conv l weights [32][cin][3][3], biases [32];
    // linear weights [6][512], biases [6]
    static constexpr int conv_w(int l) { return kCh * (l == 0 ? kInCh : kCh) * 9; }
    static constexpr int off_w(int l) { return l == 0 ? 0 : off_b(l - 1) + kCh; }
    static constexpr int off_b(int l) { return off_w(l) + conv_w(l); }
    static constexpr int off_fw() { return off_b(3) + kCh; }
    static constexpr int off_fb() { return off_fw() + kOut * kFeat; }
    static constexpr int total() { return off_fb() + kOut; }
};

struct SeederTrainConfig {
    int steps = 8000;
    int batch = 256;
    float lr = 1e-3f, lr_end = 1e-5f;
    float max_shift = 28.0f;      // full-resolution px, per component
    float max_grad = 0.12f;
    float noise = 4.0f;           // grey levels on the deformed samples
    // smooth heterogeneous fields on top of the affine: the share of pairs
    // with them, the bump amplitude (px, per component) and width range (px)
    float het_frac = 0.5f;
    float bump_amp = 8.0f;
    float bump_min = 100.0f, bump_max = 400.0f;
    uint64_t seed = 1;
};

class Seeder {
public:
    Seeder();
    ~Seeder();
    Seeder(const Seeder&) = delete;
    Seeder& operator=(const Seeder&) = delete;

    bool ready() const { return ready_; }
    static int param_count() { return SeederShape::total(); }
    // Uniform +-1 / sqrt(fan-in) initialisation of weights and biases from a seed.
    int init(uint64_t seed);
    int load(const std::string& path);
    int save(const std::string& path) const;
    // Trains on pairs cut from `ref` (device u8, w x h, full resolution).
    // losses receives the mean loss of every 100 steps. 0 or -3.
    int train(const uint8_t* ref, int w, int h, const SeederTrainConfig& cfg, std::vector<float>* losses);
    // One forward and backward pass on a given batch (x: B x 2 x 64 x 64, t:
    // B x 6 scaled targets, host) with the given parameters: the loss and the
    // gradient (host). For the gradient check against the CPU restatement.
    int grad(const float* params, const float* x, const float* t, int batch, float* grads, float* loss);
    // The forward pass alone on a given batch: y receives B x 6 scaled outputs.
    int forward(const float* params, const float* x, int batch, float* y);
    // The same through the inference kernel: fp16 activations and weights on
    // MFMA with FP32 accumulation, one workgroup per pair.
    int forward_mfma(const float* params, const float* x, int batch, float* y);
    // infer runs the MFMA kernel (true) or the FP32 forward pass of training.
    bool use_mfma = true;
    // Warps of n patch pairs cut from level-2 float planes: ref patch centred
    // on level-2 pixel c[2k], c[2k+1] (covering c - 32 .. c + 31), deformed
    // patch on c + off. out: n x 6, full-resolution (u, v, ux, uy, vx, vy) of
    // the patch centre (4 c - 0.5 in full-resolution pixels), relative to the
    // deformed patch's offset. 0 or -3.
    int infer(const float* r2, const float* d2, int w2, int h2, const int* c, const int* off, int n, float* out,
              hipStream_t st);
    std::vector<float> params() const;

private:
    struct Impl;
    Impl* d_;
    bool ready_ = false;
};

}  // namespace sdic
