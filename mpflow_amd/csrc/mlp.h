// Multilayer perceptron (dense + GELU layers, MSE loss, Adam) on MFMA; see mlp.hip.
#pragma once

#include <vector>

#include "mpf_internal.h"

namespace mpf {

struct Mlp {
    hipStream_t st;
    std::vector<int> dims;          // widths: input, hidden..., output
    int L = 0;                      // layers
    std::vector<DBuf<float>> W, b, gW, gb, mW, vW, mb, vb;
    std::vector<DBuf<float>> Z, H;  // pre-activations / activations of the last forward
    DBuf<float> dA, dB;             // backward work (batch x widest layer)
    DBuf<double> part, loss;
    int cap = 0, batch = 0;
    const float* x_in = nullptr;    // input of the last forward (caller-owned)
    int64_t t = 0;                  // Adam steps taken

    Mlp(hipStream_t s, const std::vector<int>& d, uint64_t seed);
    void ensure_batch(int B);
    // Y (nullable) = network(X); X (B x in) stays caller-owned until backward
    void forward(int B, const float* X, float* Y);
    // MSE against T (B x out) of the last forward; fills the gradients; returns the loss
    double backward(const float* T);
    void adam(float lr, float b1, float b2, float eps);
    int64_t count() const;
    void params(int which, bool set, float* host);
};

}  // namespace mpf
