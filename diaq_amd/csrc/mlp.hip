// The learned (m, s) selector's network on gfx950: a small FP32 MLP (n_in -> h1 -> h2 -> n_out, tanh hidden layers,
// softmax output) with hand-written forward, softmax cross-entropy backward, Adam update and argmax inference
// kernels. No reference counterpart.
//
// Parameters live in one device array in the order W1 [h1 x n_in], b1 [h1], W2 [h2 x h1], b2 [h2], W3 [n_out x h2],
// b3 [n_out] (row-major, out x in). The layers are tiny (at most 64 x 64 and 195 x 64 weights), so every kernel
// gives one lane one output value and loops over the reduction in a fixed order: results do not depend on the
// launch geometry, and there are no atomics, so training is deterministic.
#include <algorithm>
#include <cmath>
#include <random>

#include "dq_common.hpp"

namespace dq {

// y[b, o] = act(sum_i W[o, i] x[b, i] + bias[o]); act = tanh when tanh_act, identity otherwise.
__global__ void dense_fwd_kernel(int B, int n_in, int n_out, const float *__restrict__ W, const float *__restrict__ bias,
                                 const float *__restrict__ x, float *__restrict__ y, int tanh_act) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)B * n_out) return;
    const int b = static_cast<int>(e / n_out), o = static_cast<int>(e % n_out);
    float acc = bias[o];
    for (int i = 0; i < n_in; ++i) acc += W[o * n_in + i] * x[b * n_in + i];
    y[e] = tanh_act ? tanhf(acc) : acc;
}

// One wave per sample: softmax over n_out logits, loss_b = -log p[label], dlogits = (p - onehot) / B.
__global__ __launch_bounds__(64) void softmax_xent_kernel(int B, int n_out, const float *__restrict__ logits,
                                                          const int32_t *__restrict__ labels, float *__restrict__ dlog,
                                                          float *__restrict__ loss) {
    const int b = blockIdx.x;
    const float *z = logits + static_cast<int64_t>(b) * n_out;
    float mx = -INFINITY;
    for (int o = threadIdx.x; o < n_out; o += 64) mx = fmaxf(mx, z[o]);
    for (int w = 32; w > 0; w >>= 1) mx = fmaxf(mx, __shfl_xor(mx, w, 64));
    float se = 0.f;
    for (int o = threadIdx.x; o < n_out; o += 64) se += expf(z[o] - mx);
    for (int w = 32; w > 0; w >>= 1) se += __shfl_xor(se, w, 64);
    const int lab = labels[b];
    for (int o = threadIdx.x; o < n_out; o += 64) {
        const float p = expf(z[o] - mx) / se;
        dlog[static_cast<int64_t>(b) * n_out + o] = (p - (o == lab ? 1.f : 0.f)) / static_cast<float>(B);
    }
    if (threadIdx.x == 0) loss[b] = -(z[lab] - mx - logf(se));
}

// dW[o, i] = sum_b dy[b, o] x[b, i]; db[o] = sum_b dy[b, o] (one lane per weight; lanes i == n_in do the bias).
__global__ void dense_wgrad_kernel(int B, int n_in, int n_out, const float *__restrict__ dy, const float *__restrict__ x,
                                   float *__restrict__ dW, float *__restrict__ db) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)n_out * (n_in + 1)) return;
    const int o = static_cast<int>(e / (n_in + 1)), i = static_cast<int>(e % (n_in + 1));
    float acc = 0.f;
    if (i < n_in) {
        for (int b = 0; b < B; ++b) acc += dy[static_cast<int64_t>(b) * n_out + o] * x[static_cast<int64_t>(b) * n_in + i];
        dW[o * n_in + i] = acc;
    } else {
        for (int b = 0; b < B; ++b) acc += dy[static_cast<int64_t>(b) * n_out + o];
        db[o] = acc;
    }
}

// dx[b, i] = (sum_o W[o, i] dy[b, o]) * (1 - h[b, i]^2): back through the weights and the tanh of the layer below.
__global__ void dense_xgrad_tanh_kernel(int B, int n_in, int n_out, const float *__restrict__ W,
                                        const float *__restrict__ dy, const float *__restrict__ h,
                                        float *__restrict__ dx) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (int64_t)B * n_in) return;
    const int b = static_cast<int>(e / n_in), i = static_cast<int>(e % n_in);
    float acc = 0.f;
    for (int o = 0; o < n_out; ++o) acc += W[o * n_in + i] * dy[static_cast<int64_t>(b) * n_out + o];
    const float hv = h[e];
    dx[e] = acc * (1.f - hv * hv);
}

// Adam (Kingma & Ba 2015) with bias correction: m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
__global__ void adam_kernel(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// One wave per sample: the index of the largest logit among the allowed classes (mask[o] != 0), lowest index on ties.
__global__ __launch_bounds__(64) void argmax_kernel(int B, int n_out, const float *__restrict__ logits,
                                                    const int32_t *__restrict__ mask, int32_t *__restrict__ cls) {
    const int b = blockIdx.x;
    float best = -INFINITY;
    int bi = 0x7fffffff;
    for (int o = threadIdx.x; o < n_out; o += 64) {
        if (mask && !mask[o]) continue;
        const float z = logits[static_cast<int64_t>(b) * n_out + o];
        if (z > best || (z == best && o < bi)) { best = z; bi = o; }
    }
    for (int w = 32; w > 0; w >>= 1) {
        const float ob = __shfl_xor(best, w, 64);
        const int oi = __shfl_xor(bi, w, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (threadIdx.x == 0) cls[b] = bi == 0x7fffffff ? 0 : bi;
}

}  // namespace dq

struct dq_mlp {
    int n_in = 0, h1 = 0, h2 = 0, n_out = 0;
    int64_t n_params = 0;
    dq::DevBuf params, grads, adam_m, adam_v;
    int64_t step = 0;
    // activations / gradients for up to `cap` samples
    int cap = 0;
    dq::DevBuf a1, a2, logits, d3, d2, d1, loss;
    int64_t off_W1() const { return 0; }
    int64_t off_b1() const { return off_W1() + (int64_t)h1 * n_in; }
    int64_t off_W2() const { return off_b1() + h1; }
    int64_t off_b2() const { return off_W2() + (int64_t)h2 * h1; }
    int64_t off_W3() const { return off_b2() + h2; }
    int64_t off_b3() const { return off_W3() + (int64_t)n_out * h2; }
};

namespace dq {

static void ensure_cap(dq_mlp &M, int B) {
    if (B <= M.cap) return;
    M.a1 = DevBuf(sizeof(float) * B * M.h1);
    M.a2 = DevBuf(sizeof(float) * B * M.h2);
    M.logits = DevBuf(sizeof(float) * B * M.n_out);
    M.d3 = DevBuf(sizeof(float) * B * M.n_out);
    M.d2 = DevBuf(sizeof(float) * B * M.h2);
    M.d1 = DevBuf(sizeof(float) * B * M.h1);
    M.loss = DevBuf(sizeof(float) * B);
    M.cap = B;
}

static inline unsigned g256(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

static void forward(dq_mlp &M, int B, const float *x, hipStream_t s) {
    ensure_cap(M, B);
    const float *P = M.params.as<float>();
    hipLaunchKernelGGL(dense_fwd_kernel, dim3(g256((int64_t)B * M.h1)), dim3(256), 0, s, B, M.n_in, M.h1,
                       P + M.off_W1(), P + M.off_b1(), x, M.a1.as<float>(), 1);
    hipLaunchKernelGGL(dense_fwd_kernel, dim3(g256((int64_t)B * M.h2)), dim3(256), 0, s, B, M.h1, M.h2,
                       P + M.off_W2(), P + M.off_b2(), M.a1.as<float>(), M.a2.as<float>(), 1);
    hipLaunchKernelGGL(dense_fwd_kernel, dim3(g256((int64_t)B * M.n_out)), dim3(256), 0, s, B, M.h2, M.n_out,
                       P + M.off_W3(), P + M.off_b3(), M.a2.as<float>(), M.logits.as<float>(), 0);
    DQ_LAUNCH_CHECK();
}

static void backward(dq_mlp &M, int B, const float *x, const int32_t *y, hipStream_t s) {
    const float *P = M.params.as<float>();
    float *G = M.grads.as<float>();
    hipLaunchKernelGGL(softmax_xent_kernel, dim3(B), dim3(64), 0, s, B, M.n_out, M.logits.as<float>(), y,
                       M.d3.as<float>(), M.loss.as<float>());
    hipLaunchKernelGGL(dense_wgrad_kernel, dim3(g256((int64_t)M.n_out * (M.h2 + 1))), dim3(256), 0, s, B, M.h2,
                       M.n_out, M.d3.as<float>(), M.a2.as<float>(), G + M.off_W3(), G + M.off_b3());
    hipLaunchKernelGGL(dense_xgrad_tanh_kernel, dim3(g256((int64_t)B * M.h2)), dim3(256), 0, s, B, M.h2, M.n_out,
                       P + M.off_W3(), M.d3.as<float>(), M.a2.as<float>(), M.d2.as<float>());
    hipLaunchKernelGGL(dense_wgrad_kernel, dim3(g256((int64_t)M.h2 * (M.h1 + 1))), dim3(256), 0, s, B, M.h1, M.h2,
                       M.d2.as<float>(), M.a1.as<float>(), G + M.off_W2(), G + M.off_b2());
    hipLaunchKernelGGL(dense_xgrad_tanh_kernel, dim3(g256((int64_t)B * M.h1)), dim3(256), 0, s, B, M.h1, M.h2,
                       P + M.off_W2(), M.d2.as<float>(), M.a1.as<float>(), M.d1.as<float>());
    hipLaunchKernelGGL(dense_wgrad_kernel, dim3(g256((int64_t)M.h1 * (M.n_in + 1))), dim3(256), 0, s, B, M.n_in,
                       M.h1, M.d1.as<float>(), x, G + M.off_W1(), G + M.off_b1());
    DQ_LAUNCH_CHECK();
}

static float mean_loss(dq_mlp &M, int B, hipStream_t s) {
    std::vector<float> h(B);
    DQ_HIP(hipMemcpyAsync(h.data(), M.loss.p, sizeof(float) * B, hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for (float v : h) acc += v;
    return static_cast<float>(acc / B);
}

}  // namespace dq

using namespace dq;

extern "C" int dq_mlp_create(int n_in, int h1, int h2, int n_out, uint64_t seed, dq_mlp **out) {
    return guard([&] {
        if (!out || n_in <= 0 || h1 <= 0 || h2 <= 0 || n_out <= 0 || n_in > 4096 || h1 > 4096 || h2 > 4096 ||
            n_out > 4096)
            throw ArgError("dq_mlp_create: bad sizes");
        auto M = std::make_unique<dq_mlp>();
        M->n_in = n_in; M->h1 = h1; M->h2 = h2; M->n_out = n_out;
        M->n_params = M->off_b3() + n_out;
        // Glorot-uniform weights, zero biases (host PRNG, seeded)
        std::vector<float> p(M->n_params, 0.f);
        std::mt19937_64 rng(seed);
        auto fill = [&](int64_t off, int fan_out, int fan_in) {
            const float a = std::sqrt(6.f / static_cast<float>(fan_in + fan_out));
            std::uniform_real_distribution<float> u(-a, a);
            for (int64_t i = 0; i < (int64_t)fan_out * fan_in; ++i) p[off + i] = u(rng);
        };
        fill(M->off_W1(), h1, n_in);
        fill(M->off_W2(), h2, h1);
        fill(M->off_W3(), n_out, h2);
        const size_t pb = sizeof(float) * M->n_params;
        M->params = DevBuf(pb); M->grads = DevBuf(pb); M->adam_m = DevBuf(pb); M->adam_v = DevBuf(pb);
        DQ_HIP(hipMemcpy(M->params.p, p.data(), pb, hipMemcpyHostToDevice));
        DQ_HIP(hipMemset(M->grads.p, 0, pb));
        DQ_HIP(hipMemset(M->adam_m.p, 0, pb));
        DQ_HIP(hipMemset(M->adam_v.p, 0, pb));
        *out = M.release();
    });
}

extern "C" void dq_mlp_destroy(dq_mlp *M) { delete M; }

extern "C" int64_t dq_mlp_num_params(const dq_mlp *M) { return M ? M->n_params : 0; }

extern "C" int dq_mlp_get_params(const dq_mlp *M, float *host) {
    return guard([&] {
        if (!M || !host) throw ArgError("NULL argument");
        DQ_HIP(hipMemcpy(host, M->params.p, sizeof(float) * M->n_params, hipMemcpyDeviceToHost));
    });
}

extern "C" int dq_mlp_set_params(dq_mlp *M, const float *host) {
    return guard([&] {
        if (!M || !host) throw ArgError("NULL argument");
        DQ_HIP(hipMemcpy(M->params.p, host, sizeof(float) * M->n_params, hipMemcpyHostToDevice));
        M->step = 0;
        DQ_HIP(hipMemset(M->adam_m.p, 0, sizeof(float) * M->n_params));
        DQ_HIP(hipMemset(M->adam_v.p, 0, sizeof(float) * M->n_params));
    });
}

// logits_host (may be NULL): B x n_out logits copied back.
extern "C" int dq_mlp_forward(dq_mlp *M, int B, const float *x_dev, float *logits_host) {
    return guard([&] {
        if (!M || B <= 0 || !x_dev) throw ArgError("dq_mlp_forward: bad arguments");
        forward(*M, B, x_dev, 0);
        if (logits_host)
            DQ_HIP(hipMemcpy(logits_host, M->logits.p, sizeof(float) * B * M->n_out, hipMemcpyDeviceToHost));
        DQ_HIP(hipStreamSynchronize(0));
    });
}

// Mean loss and the gradient of the mean loss (grad_host: n_params floats, may be NULL) for one batch.
extern "C" int dq_mlp_loss_grad(dq_mlp *M, int B, const float *x_dev, const int32_t *y_dev, float *loss,
                                float *grad_host) {
    return guard([&] {
        if (!M || B <= 0 || !x_dev || !y_dev) throw ArgError("dq_mlp_loss_grad: bad arguments");
        forward(*M, B, x_dev, 0);
        backward(*M, B, x_dev, y_dev, 0);
        if (loss) *loss = mean_loss(*M, B, 0);
        if (grad_host)
            DQ_HIP(hipMemcpy(grad_host, M->grads.p, sizeof(float) * M->n_params, hipMemcpyDeviceToHost));
    });
}

// One Adam step on one batch; loss (may be NULL) receives the batch's mean loss before the step.
extern "C" int dq_mlp_train_step(dq_mlp *M, int B, const float *x_dev, const int32_t *y_dev, float lr, float *loss) {
    return guard([&] {
        if (!M || B <= 0 || !x_dev || !y_dev) throw ArgError("dq_mlp_train_step: bad arguments");
        forward(*M, B, x_dev, 0);
        backward(*M, B, x_dev, y_dev, 0);
        M->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, static_cast<float>(M->step));
        const float c2 = 1.f - std::pow(b2, static_cast<float>(M->step));
        hipLaunchKernelGGL(adam_kernel, dim3(g256(M->n_params)), dim3(256), 0, 0, (int)M->n_params,
                           M->params.as<float>(), M->grads.as<float>(), M->adam_m.as<float>(), M->adam_v.as<float>(),
                           lr, b1, b2, 1e-8f, c1, c2);
        DQ_LAUNCH_CHECK();
        if (loss) *loss = mean_loss(*M, B, 0);
        else DQ_HIP(hipStreamSynchronize(0));
    });
}

// Class of every sample: argmax of the logits over the allowed classes (mask_dev: n_out int32 flags, may be NULL).
extern "C" int dq_mlp_predict(dq_mlp *M, int B, const float *x_dev, const int32_t *mask_dev, int32_t *cls_dev,
                              uintptr_t stream) {
    return guard([&] {
        if (!M || B <= 0 || !x_dev || !cls_dev) throw ArgError("dq_mlp_predict: bad arguments");
        hipStream_t s = as_stream(stream);
        forward(*M, B, x_dev, s);
        hipLaunchKernelGGL(argmax_kernel, dim3(B), dim3(64), 0, s, B, M->n_out, M->logits.as<float>(), mask_dev,
                           cls_dev);
        DQ_LAUNCH_CHECK();
    });
}
