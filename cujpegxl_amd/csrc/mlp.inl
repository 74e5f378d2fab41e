// The learned transform-selection / adaptive-quantization network: its
// training kernels (included into device.hip).
//
// Network (one row of 96 features per 32x32 region):
//   h = ReLU(W1 x + b1)  (64),  z = W2 h + b2  (7)
//   z[0], z[1]  root logits (split, DCT32) under a softmax
//   z[2..5]     quad q: DCT16 (> 0) or four DCT8s (< 0), sigmoid
//   z[6]        the region's mean log2 quant field
// Parameters (FP32 masters, laid out W1[64][96], b1[64], W2[7][64], b2[7]);
// the encoder runs FP16 copies of W1 and W2 (region kernel, FP32
// accumulation).  Labels per region (12 floats): root, quad[4], quant-field
// target, root mask, quad mask[4], unused; a mask of 0 drops the term.
// Loss: mean over regions of m_r CE(root) + sum_q m_q BCE(quad q) +
// lambda (z6 - target)^2.
//
// mlp_grad_kernel: one workgroup per 32 regions of a batch: forward into LDS,
// loss gradient of the outputs, backward, and the workgroup's gradient sums
// into its row of a partial-gradient buffer.  mlp_reduce_kernel sums the rows
// (fixed order) into the batch's mean gradient and, with a learning rate,
// takes one Adam step.

namespace cjx {

constexpr int kMlpIn = 96;
constexpr int kMlpHid = 64;
constexpr int kMlpOut = 7;
constexpr int kMlpLabels = 12;
constexpr int kMlpParams = kMlpHid * kMlpIn + kMlpHid + kMlpOut * kMlpHid + kMlpOut;
constexpr int kMlpB1 = kMlpHid * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpOut * kMlpHid;
constexpr int kMlpTiles = 32;      // regions per workgroup of the gradient kernel
constexpr int kMlpW1Stride = 97;   // LDS row stride of W1 (odd: rows in distinct banks)

struct MlpShared {
    float w1[kMlpHid * kMlpW1Stride];
    float w2[kMlpOut * kMlpHid];
    float b1[kMlpHid];
    float b2[kMlpOut + 1];
    float x[kMlpTiles][kMlpIn];
    float h[kMlpTiles][kMlpHid];   // activations, then their gradients
    float dz[kMlpTiles][8];
    float loss[kMlpTiles];
};

__global__ __launch_bounds__(256) void mlp_grad_kernel(const float* __restrict__ params, const float* __restrict__ x,
                                                       const float* __restrict__ labels, std::uint32_t n,
                                                       float lambda, float* __restrict__ partial) {
    __shared__ MlpShared S;
    const int t{static_cast<int>(threadIdx.x)};
    const std::uint32_t first{blockIdx.x * static_cast<std::uint32_t>(kMlpTiles)};
    const int count{static_cast<int>(n - first < static_cast<std::uint32_t>(kMlpTiles) ? n - first : kMlpTiles)};
    for (int i{t}; i < kMlpHid * kMlpIn; i += 256) {
        S.w1[(i / kMlpIn) * kMlpW1Stride + i % kMlpIn] = params[i];
    }
    for (int i{t}; i < kMlpOut * kMlpHid; i += 256) {
        S.w2[i] = params[kMlpW2 + i];
    }
    if (t < kMlpHid) {
        S.b1[t] = params[kMlpB1 + t];
    }
    if (t < kMlpOut) {
        S.b2[t] = params[kMlpB2 + t];
    }
    for (int i{t}; i < kMlpTiles * kMlpIn; i += 256) {
        const int r{i / kMlpIn};
        S.x[r][i % kMlpIn] = r < count ? x[(static_cast<std::size_t>(first) + r) * kMlpIn + i % kMlpIn] : 0.0f;
    }
    __syncthreads();
    // Hidden layer: thread t takes unit j = t % 64 of regions t / 64 + 4k.
    const int j{t & 63};
    for (int k{0}; k < kMlpTiles / 4; ++k) {
        const int r{(t >> 6) + 4 * k};
        float s{S.b1[j]};
        for (int i{0}; i < kMlpIn; ++i) {
            s = __builtin_fmaf(S.w1[j * kMlpW1Stride + i], S.x[r][i], s);
        }
        S.h[r][j] = fmaxf(s, 0.0f);
    }
    __syncthreads();
    // Outputs and their loss gradient: thread r < count takes region r.
    if (t < kMlpTiles) {
        const int r{t};
        float z[kMlpOut];
        for (int o{0}; o < kMlpOut; ++o) {
            float s{S.b2[o]};
            for (int q{0}; q < kMlpHid; ++q) {
                s = __builtin_fmaf(S.w2[o * kMlpHid + q], S.h[r][q], s);
            }
            z[o] = s;
        }
        float dz[kMlpOut]{};
        float loss{0.0f};
        if (r < count) {
            const float* l{labels + (static_cast<std::size_t>(first) + r) * kMlpLabels};
            const float mr{l[6]};
            const float mx{fmaxf(z[0], z[1])};
            const float e0{expf(z[0] - mx)}, e1{expf(z[1] - mx)};
            const float p1{e1 / (e0 + e1)};
            const bool y32{l[0] > 0.5f};
            loss += mr * (mx + logf(e0 + e1) - (y32 ? z[1] : z[0]));
            dz[0] = mr * ((1.0f - p1) - (y32 ? 0.0f : 1.0f));
            dz[1] = mr * (p1 - (y32 ? 1.0f : 0.0f));
            for (int q{0}; q < 4; ++q) {
                const float m{l[7 + q]}, y{l[1 + q]}, v{z[2 + q]};
                const float sp{v > 30.0f ? v : log1pf(expf(v))};
                loss += m * (sp - y * v);
                dz[2 + q] = m * (1.0f / (1.0f + expf(-v)) - y);
            }
            const float d{z[6] - l[5]};
            loss += lambda * d * d;
            dz[6] = 2.0f * lambda * d;
        }
        for (int o{0}; o < kMlpOut; ++o) {
            S.dz[r][o] = dz[o];
        }
        S.loss[r] = loss;
    }
    __syncthreads();
    float* const out{partial + static_cast<std::size_t>(blockIdx.x) * (kMlpParams + 1)};
    // Second layer's gradient: W2 (448) and b2 (7) entries, sums over regions.
    for (int e{t}; e < kMlpOut * kMlpHid + kMlpOut; e += 256) {
        float s{0.0f};
        if (e < kMlpOut * kMlpHid) {
            const int o{e / kMlpHid}, q{e % kMlpHid};
            for (int r{0}; r < kMlpTiles; ++r) {
                s = __builtin_fmaf(S.dz[r][o], S.h[r][q], s);
            }
            out[kMlpW2 + e] = s;
        } else {
            const int o{e - kMlpOut * kMlpHid};
            for (int r{0}; r < kMlpTiles; ++r) {
                s += S.dz[r][o];
            }
            out[kMlpB2 + o] = s;
        }
    }
    if (t == 0) {
        float s{0.0f};
        for (int r{0}; r < kMlpTiles; ++r) {
            s += S.loss[r];
        }
        out[kMlpParams] = s;
    }
    __syncthreads();
    // Hidden gradients (the ReLU mask from the stored activations), in place.
    for (int k{0}; k < kMlpTiles / 4; ++k) {
        const int r{(t >> 6) + 4 * k};
        float s{0.0f};
        for (int o{0}; o < kMlpOut; ++o) {
            s = __builtin_fmaf(S.w2[o * kMlpHid + j], S.dz[r][o], s);
        }
        S.h[r][j] = S.h[r][j] > 0.0f ? s : 0.0f;
    }
    __syncthreads();
    // First layer's gradient: W1 (6,144) and b1 (64) entries.
    for (int e{t}; e < kMlpHid * kMlpIn + kMlpHid; e += 256) {
        float s{0.0f};
        if (e < kMlpHid * kMlpIn) {
            const int q{e / kMlpIn}, i{e % kMlpIn};
            for (int r{0}; r < kMlpTiles; ++r) {
                s = __builtin_fmaf(S.h[r][q], S.x[r][i], s);
            }
            out[e] = s;
        } else {
            const int q{e - kMlpHid * kMlpIn};
            for (int r{0}; r < kMlpTiles; ++r) {
                s += S.h[r][q];
            }
            out[kMlpB1 + q] = s;
        }
    }
}

// grad[p] = (sum over the rows of partial[p]) / n; loss[0] likewise.  With
// lr > 0, one Adam step (beta1 0.9, beta2 0.999, eps 1e-8; `step` counts
// from 1) on params with moments m, v.
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ partial, std::uint32_t rows,
                                                         std::uint32_t n, float* __restrict__ grad,
                                                         float* __restrict__ loss, float* __restrict__ params,
                                                         float* __restrict__ m, float* __restrict__ v,
                                                         std::uint32_t step, float lr) {
    const std::uint32_t p{blockIdx.x * 256u + threadIdx.x};
    if (p > static_cast<std::uint32_t>(kMlpParams)) {
        return;
    }
    float s{0.0f};
    for (std::uint32_t r{0}; r < rows; ++r) {
        s += partial[static_cast<std::size_t>(r) * (kMlpParams + 1) + p];
    }
    const float g{s / static_cast<float>(n)};
    if (p == static_cast<std::uint32_t>(kMlpParams)) {
        loss[0] = g;
        return;
    }
    grad[p] = g;
    if (lr > 0.0f) {
        const float b1{0.9f}, b2{0.999f};
        const float mp{b1 * m[p] + (1.0f - b1) * g};
        const float vp{b2 * v[p] + (1.0f - b2) * g * g};
        m[p] = mp;
        v[p] = vp;
        const float c1{1.0f - powf(b1, static_cast<float>(step))};
        const float c2{1.0f - powf(b2, static_cast<float>(step))};
        params[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + 1e-8f);
    }
}

// Outputs z[n][7] of the FP32 network (evaluation of training runs).
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* __restrict__ params,
                                                          const float* __restrict__ x, std::uint32_t n,
                                                          float* __restrict__ z) {
    const std::uint32_t r{blockIdx.x * 256u + threadIdx.x};
    if (r >= n) {
        return;
    }
    const float* xr{x + static_cast<std::size_t>(r) * kMlpIn};
    float h[kMlpHid];
    for (int q{0}; q < kMlpHid; ++q) {
        float s{params[kMlpB1 + q]};
        for (int i{0}; i < kMlpIn; ++i) {
            s = __builtin_fmaf(params[q * kMlpIn + i], xr[i], s);
        }
        h[q] = fmaxf(s, 0.0f);
    }
    for (int o{0}; o < kMlpOut; ++o) {
        float s{params[kMlpB2 + o]};
        for (int q{0}; q < kMlpHid; ++q) {
            s = __builtin_fmaf(params[kMlpW2 + o * kMlpHid + q], h[q], s);
        }
        z[static_cast<std::size_t>(r) * kMlpOut + o] = s;
    }
}

}  // namespace cjx
