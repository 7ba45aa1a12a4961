// Learned spike-presence detector: launchers shared by tcn.hip (network
// kernels) and detect.hip (the gated scan that turns window classes into
// event times).
#pragma once

#include "common.h"

namespace msamd {
namespace tcn {

constexpr int kH = 32;     // hidden channels
constexpr int kCI = 7;     // input channels: central -3..+3
constexpr int kK = 9;      // taps
constexpr int kWin = 32;   // samples per scored window
constexpr int kS1 = 8;     // layer-1 stride
constexpr int kS2 = 4;     // layer-2 stride

constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kCI * kK;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH * kK;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH * kK;
constexpr int kNumParams = kOffB3 + 1;

inline int64_t num_windows(int64_t N) { return ceil_div(N, kWin); }

// Bytes of the MFMA weight fragments built from a parameter vector.
int64_t frag_bytes();

// Builds the inference fragments (bf16 weights in MFMA lane order, fp32
// biases and layer-3 weights) from the fp32 parameter vector.
int prepare(const float* params, void* frag, hipStream_t s);

// Logits (optional, [nW][M] float) and classes (optional, [nW][M] uint8:
// 0 when z < lo, 2 when z > hi, else 1) of every window of every channel.
int classify(const float* X, int64_t M, int64_t N, const void* frag, float lo_logit,
             float hi_logit, float* logits, uint8_t* classes, hipStream_t s);

}  // namespace tcn
}  // namespace msamd
