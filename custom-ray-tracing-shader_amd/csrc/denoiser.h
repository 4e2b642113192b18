// denoiser.h - the learned denoiser: a 5-level U-Net of 3x3 convolutions whose forward and
// backward passes run on gfx950 BF16 MFMA with FP32 accumulation (denoiser.hip).
#pragma once
#include <string>
#include <vector>
#include "rt_core.h"

namespace rt {

constexpr int kDenoiseInputChannels = 10;   // log radiance 3, albedo 3, normal 3, log depth 1
constexpr unsigned kWeightFileMagic = 0x4E445452u;   // "RTDN"
constexpr unsigned kWeightFileVersion = 2;   // 2: albedo-demodulated radiance features

struct DenoiserNet;

DenoiserNet* denoiserCreate();
void denoiserDestroy(DenoiserNet* net);
long long denoiserParamCount(const DenoiserNet* net);
hipError_t denoiserInit(DenoiserNet* net, unsigned seed);
hipError_t denoiserGetParams(DenoiserNet* net, float* host);
hipError_t denoiserSetParams(DenoiserNet* net, const float* host);
bool denoiserSave(DenoiserNet* net, const char* path, std::string& error);
bool denoiserLoad(DenoiserNet* net, const char* path, std::string& error);

// Inference on device-resident planes: input f32 [n][h][w][10] (already transformed as
// described in denoiser.hip), output f32 radiance [n][h][w][4] (channel 3 unused).
hipError_t denoiserInferDevice(DenoiserNet* net, const float* dInput, int n, int h, int w, float* dOutput,
                               hipStream_t s);
// Host-buffer wrappers for tests and tools.
hipError_t denoiserInferHost(DenoiserNet* net, const float* input, int n, int h, int w, float* output);
hipError_t denoiserGradients(DenoiserNet* net, const float* input, const float* target, int n, int h, int w,
                             float* loss, float* grads);
hipError_t denoiserSetDataset(DenoiserNet* net, const float* inputs, const float* targets, int count, int h, int w);
hipError_t denoiserTrain(DenoiserNet* net, int steps, int batch, int crop, float lr, unsigned seed, float* losses);
// Analytic multiply-add FLOPs of one forward pass over n x h x w pixels (h, w multiples of 16).
double denoiserForwardFlops(int n, int h, int w);
// Weight of the display-space L1 term of the training loss (0 = log-space L1 only).
bool denoiserSetDisplayWeight(DenoiserNet* net, float weight);
// Device time of the last forward and of the last training step's forward+backward.
float denoiserLastForwardMs(const DenoiserNet* net);
float denoiserLastTrainStepMs(const DenoiserNet* net);

// Single-layer checks: 3x3 convolution, stride 1, zero padding, NHWC, weights [9][cout][cin]
// (tap = 3*(dy+1) + (dx+1)).  Inputs are rounded to BF16 first.
hipError_t convForwardHost(const float* in, int n, int h, int w, int cin, const float* weight, const float* bias,
                           int cout, int relu, float* out);
hipError_t convBackwardHost(const float* in, int n, int h, int w, int cin, const float* weight, const float* gradOut,
                            int cout, float* gradIn, float* gradWeight, float* gradBias);

}  // namespace rt
