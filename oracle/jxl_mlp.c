/* Test infrastructure only: the CPU oracle of the learned transform-selection
 * / adaptive-quantization network (no counterpart in the reference; its
 * labels are the reference heuristic's decisions, transform_select_impl.cuh
 * :130-201 for the transform map and frontend_fused.cu:177-201 for the quant
 * field).  Everything here is float64.
 *
 * Network: x (96 features) -> h = ReLU(W1 x + b1) (64) -> z = W2 h + b2 (7).
 *   z[0], z[1]  root logits (split, DCT32), softmax; label 1 = DCT32
 *   z[2..5]     quad q logits (4 DCT8 vs DCT16), sigmoid; label 1 = DCT16
 *   z[6]        the region's mean log2(quant field), squared error
 * Parameters (floats): W1[64][96], b1[64], W2[7][64], b2[7] (6,663).
 * Labels per tile (12 floats): root, quad[4], aq target, root mask,
 * quad mask[4], unused.  A mask of 0 drops that term (the size does not fit
 * the frame).  Loss = mean over tiles of
 *   m_r CE(root) + sum_q m_q BCE(quad q) + lambda (z6 - aq)^2. */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "jxl_oracle.h"

#define IN JO_MLP_IN
#define HID JO_MLP_HID
#define OUT JO_MLP_OUT

static void forward_one(const float* p, const float* x, double* h, double* z) {
    const float* w1 = p;
    const float* b1 = w1 + HID * IN;
    const float* w2 = b1 + HID;
    const float* b2 = w2 + OUT * HID;
    for (int j = 0; j < HID; ++j) {
        double s = b1[j];
        for (int i = 0; i < IN; ++i) {
            s += (double)w1[j * IN + i] * x[i];
        }
        h[j] = s > 0.0 ? s : 0.0;
    }
    for (int o = 0; o < OUT; ++o) {
        double s = b2[o];
        for (int j = 0; j < HID; ++j) {
            s += (double)w2[o * HID + j] * h[j];
        }
        z[o] = s;
    }
}

void jo_mlp_forward(const float* params, const float* x, uint32_t n, double* z) {
    double h[HID];
    for (uint32_t t = 0; t < n; ++t) {
        forward_one(params, x + (size_t)t * IN, h, z + (size_t)t * OUT);
    }
}

static double softplus(double v) {
    return v > 30.0 ? v : log1p(exp(v));
}

double jo_mlp_loss_grad(const float* params, const float* x, const float* lab, uint32_t n, float lambda,
                        double* grad) {
    const float* w2 = params + HID * IN + HID;
    double* g1 = grad;
    double* gb1 = g1 + HID * IN;
    double* g2 = gb1 + HID;
    double* gb2 = g2 + OUT * HID;
    if (grad) {
        memset(grad, 0, sizeof(double) * JO_MLP_PARAMS);
    }
    double loss = 0.0;
    for (uint32_t t = 0; t < n; ++t) {
        const float* xt = x + (size_t)t * IN;
        const float* l = lab + (size_t)t * JO_MLP_LABELS;
        double h[HID], z[OUT], dz[OUT];
        forward_one(params, xt, h, z);
        /* root: softmax cross-entropy over (z0, z1) */
        const double mr = l[6];
        const double mx = z[0] > z[1] ? z[0] : z[1];
        const double e0 = exp(z[0] - mx), e1 = exp(z[1] - mx);
        const double p1 = e1 / (e0 + e1);
        const double lse = mx + log(e0 + e1);
        loss += mr * (lse - (l[0] > 0.5f ? z[1] : z[0]));
        dz[0] = mr * ((1.0 - p1) - (l[0] > 0.5f ? 0.0 : 1.0));
        dz[1] = mr * (p1 - (l[0] > 0.5f ? 1.0 : 0.0));
        /* quads: binary cross-entropy with logits */
        for (int q = 0; q < 4; ++q) {
            const double m = l[7 + q], y = l[1 + q], v = z[2 + q];
            loss += m * (softplus(v) - y * v);
            dz[2 + q] = m * (1.0 / (1.0 + exp(-v)) - y);
        }
        /* quant field scale: squared error */
        const double r = z[6] - l[5];
        loss += lambda * r * r;
        dz[6] = 2.0 * lambda * r;
        if (!grad) {
            continue;
        }
        double dh[HID];
        for (int j = 0; j < HID; ++j) {
            double s = 0.0;
            for (int o = 0; o < OUT; ++o) {
                s += (double)w2[o * HID + j] * dz[o];
                g2[o * HID + j] += dz[o] * h[j];
            }
            dh[j] = h[j] > 0.0 ? s : 0.0;
        }
        for (int o = 0; o < OUT; ++o) {
            gb2[o] += dz[o];
        }
        for (int j = 0; j < HID; ++j) {
            gb1[j] += dh[j];
            for (int i = 0; i < IN; ++i) {
                g1[j * IN + i] += dh[j] * xt[i];
            }
        }
    }
    if (grad) {
        for (int k = 0; k < JO_MLP_PARAMS; ++k) {
            grad[k] /= n;
        }
    }
    return loss / n;
}
