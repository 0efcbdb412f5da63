// Learned schedule selector: a 16-64-64-K MLP that predicts the log time of
// each fused-BFS step mode from device-side frontier statistics; inference
// runs inside the BFS decide kernel (BF16 weights, FP32 accumulation), and
// training runs here with the project's own forward / backward / Adam
// kernels in FP32. The reference has no learned component; GraphIt's fixed
// hybrid rule (operators.cpp:526) is the fallback and the baseline.
#include <cmath>
#include <cstring>
#include <vector>

#include "gi_bfs_fused.h"

namespace gi {

// ------------------------------------------------------------ model storage
struct SelectorModel {
    uint16_t *w = nullptr;  // BF16 inference blob: w1 [hid][in], w2 [hid][hid], w3 [hid][kSelMaxOut], b1, b2, b3
    float norm[2 * kSelInputs] = {};  // feat_mean, feat_scale
    int32_t n_out = 0;
    int32_t allowed = 0;
    bool loaded = false;
};
static SelectorModel g_model;
static unsigned long long g_model_version = 0;  // advanced by every load and clear

static uint16_t f32_to_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    uint32_t rounding = 0x7fffu + ((u >> 16) & 1u);     // round to nearest even
    return (uint16_t)((u + rounding) >> 16);
}

static int64_t param_count(int32_t n_out) {
    return (int64_t)kSelInputs * kSelHidden + kSelHidden + (int64_t)kSelHidden * kSelHidden +
           kSelHidden + (int64_t)kSelHidden * n_out + n_out;
}

SelectorView selector_view() {
    SelectorView v{};
    if (!g_model.loaded) return v;
    const uint16_t *p = g_model.w;  // 16-byte rows: w1 at 0, w2 at 2048, w3 at 10240 bytes
    v.w1 = p + kSelOffW1;
    v.w2 = p + kSelOffW2;
    v.w3 = p + kSelOffW3;
    v.b1 = p + kSelOffB1;
    v.b2 = p + kSelOffB2;
    v.b3 = p + kSelOffB3;
    v.version = g_model_version;
    v.n_out = g_model.n_out;
    v.allowed = g_model.allowed;
    v.mode = GI_SELECT_LEARNED;
    for (int k = 0; k < kSelInputs; ++k) {
        v.feat_mean[k] = g_model.norm[k];
        v.feat_scale[k] = g_model.norm[kSelInputs + k];
    }
    return v;
}

// ------------------------------------------------------------ training kernels
// Y[b][j] = act(sum_k X[b][k] W[k][j] + bias[j]); thread per (b, j).
__global__ void mlp_fwd_kernel(const float *X, const float *W, const float *bias, float *Y,
                               int32_t B, int32_t nin, int32_t nout, int32_t relu) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * nout) return;
    int32_t b = (int32_t)(t / nout), j = (int32_t)(t % nout);
    float acc = bias[j];
    for (int32_t k = 0; k < nin; ++k) acc += X[(int64_t)b * nin + k] * W[(int64_t)k * nout + j];
    Y[t] = (relu && acc < 0.f) ? 0.f : acc;
}

// Weighted mean squared error L = sum m (y - t)^2 / sum m (m = 0 masks an
// entry out): G = 2 m (y - t) / sum m; loss partials per block.
__global__ void mlp_loss_kernel(const float *Y, const float *T, const float *M, float *G,
                                int64_t n, float inv_msum, float *loss) {
    float part = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float d = Y[i] - T[i];
        part += M[i] * d * d;
        G[i] = 2.f * M[i] * d * inv_msum;
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if (lane_id() == 0) atomicAdd(loss, part * inv_msum);
}

// GW[k][j] = sum_b X[b][k] GY[b][j]; GB[j] = sum_b GY[b][j]; thread per (k, j)
// plus nout threads for the bias.
__global__ void mlp_wgrad_kernel(const float *X, const float *GY, float *GW, float *GB,
                                 int32_t B, int32_t nin, int32_t nout) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nw = (int64_t)nin * nout;
    if (t < nw) {
        int32_t k = (int32_t)(t / nout), j = (int32_t)(t % nout);
        float acc = 0.f;
        for (int32_t b = 0; b < B; ++b) acc += X[(int64_t)b * nin + k] * GY[(int64_t)b * nout + j];
        GW[t] = acc;
    } else if (t < nw + nout) {
        int32_t j = (int32_t)(t - nw);
        float acc = 0.f;
        for (int32_t b = 0; b < B; ++b) acc += GY[(int64_t)b * nout + j];
        GB[j] = acc;
    }
}

// GX[b][k] = (sum_j GY[b][j] W[k][j]) * (H[b][k] > 0); thread per (b, k).
__global__ void mlp_xgrad_kernel(const float *GY, const float *W, const float *H, float *GX,
                                 int32_t B, int32_t nin, int32_t nout) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * nin) return;
    int32_t b = (int32_t)(t / nin), k = (int32_t)(t % nin);
    float acc = 0.f;
    for (int32_t j = 0; j < nout; ++j) acc += GY[(int64_t)b * nout + j] * W[(int64_t)k * nout + j];
    GX[t] = H[t] > 0.f ? acc : 0.f;
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, int64_t n, float lr,
                            float b1, float b2, float c1, float c2, float eps, float wd) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi_ = g[i] + wd * p[i];
    m[i] = b1 * m[i] + (1.f - b1) * gi_;
    v[i] = b2 * v[i] + (1.f - b2) * gi_ * gi_;
    float mh = m[i] / c1, vh = v[i] / c2;
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

static int blocks_for(int64_t n) { return (int)((n + 255) / 256); }

// Device MLP workspace for one batch.
struct MlpWork {
    int32_t B = 0, nout = 0;
    float *X = nullptr, *T = nullptr, *M = nullptr;
    float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;
    float *h1 = nullptr, *h2 = nullptr, *y = nullptr, *gy = nullptr, *g2 = nullptr, *g1 = nullptr;
    float *loss = nullptr;
    float inv_msum = 1.f;
    ~MlpWork() {
        for (float *p : {X, T, M, P, G, m, v, h1, h2, y, gy, g2, g1, loss}) dfree(p);
    }
    int init(int32_t b, int32_t no, const float *hX, const float *hT, const float *hM,
             const float *hP) {
        B = b;
        nout = no;
        int64_t np = param_count(no);
        GI_TRY(dalloc(&X, (size_t)B * kSelInputs));
        GI_TRY(dalloc(&T, (size_t)B * nout));
        GI_TRY(dalloc(&M, (size_t)B * nout));
        GI_TRY(dalloc(&P, (size_t)np));
        GI_TRY(dalloc(&G, (size_t)np));
        GI_TRY(dalloc(&m, (size_t)np));
        GI_TRY(dalloc(&v, (size_t)np));
        GI_TRY(dalloc(&h1, (size_t)B * kSelHidden));
        GI_TRY(dalloc(&h2, (size_t)B * kSelHidden));
        GI_TRY(dalloc(&y, (size_t)B * nout));
        GI_TRY(dalloc(&gy, (size_t)B * nout));
        GI_TRY(dalloc(&g2, (size_t)B * kSelHidden));
        GI_TRY(dalloc(&g1, (size_t)B * kSelHidden));
        GI_TRY(dalloc(&loss, 1));
        GI_HIP(hipMemcpy(X, hX, sizeof(float) * B * kSelInputs, hipMemcpyHostToDevice));
        GI_HIP(hipMemcpy(T, hT, sizeof(float) * B * nout, hipMemcpyHostToDevice));
        GI_HIP(hipMemcpy(M, hM, sizeof(float) * B * nout, hipMemcpyHostToDevice));
        GI_HIP(hipMemcpy(P, hP, sizeof(float) * np, hipMemcpyHostToDevice));
        GI_TRY(fill_f32(m, np, 0.f, 0));
        GI_TRY(fill_f32(v, np, 0.f, 0));
        double msum = 0.0;
        for (int64_t i = 0; i < (int64_t)B * nout; ++i) msum += hM[i];
        inv_msum = (float)(1.0 / (msum > 0 ? msum : 1.0));
        return GI_OK;
    }
    // forward + loss + backward into G
    int step(hipStream_t s) {
        float *w1 = P, *b1 = w1 + kSelInputs * kSelHidden, *w2 = b1 + kSelHidden;
        float *b2 = w2 + kSelHidden * kSelHidden, *w3 = b2 + kSelHidden, *b3 = w3 + kSelHidden * nout;
        float *gw1 = G, *gb1 = gw1 + kSelInputs * kSelHidden, *gw2 = gb1 + kSelHidden;
        float *gb2 = gw2 + kSelHidden * kSelHidden, *gw3 = gb2 + kSelHidden, *gb3 = gw3 + kSelHidden * nout;
        mlp_fwd_kernel<<<blocks_for((int64_t)B * kSelHidden), 256, 0, s>>>(X, w1, b1, h1, B, kSelInputs, kSelHidden, 1);
        mlp_fwd_kernel<<<blocks_for((int64_t)B * kSelHidden), 256, 0, s>>>(h1, w2, b2, h2, B, kSelHidden, kSelHidden, 1);
        mlp_fwd_kernel<<<blocks_for((int64_t)B * nout), 256, 0, s>>>(h2, w3, b3, y, B, kSelHidden, nout, 0);
        GI_TRY(fill_f32(loss, 1, 0.f, s));
        mlp_loss_kernel<<<blocks_for((int64_t)B * nout), 256, 0, s>>>(y, T, M, gy, (int64_t)B * nout, inv_msum, loss);
        mlp_wgrad_kernel<<<blocks_for((int64_t)kSelHidden * nout + nout), 256, 0, s>>>(h2, gy, gw3, gb3, B, kSelHidden, nout);
        mlp_xgrad_kernel<<<blocks_for((int64_t)B * kSelHidden), 256, 0, s>>>(gy, w3, h2, g2, B, kSelHidden, nout);
        mlp_wgrad_kernel<<<blocks_for((int64_t)kSelHidden * kSelHidden + kSelHidden), 256, 0, s>>>(h1, g2, gw2, gb2, B, kSelHidden, kSelHidden);
        mlp_xgrad_kernel<<<blocks_for((int64_t)B * kSelHidden), 256, 0, s>>>(g2, w2, h1, g1, B, kSelHidden, kSelHidden);
        mlp_wgrad_kernel<<<blocks_for((int64_t)kSelInputs * kSelHidden + kSelHidden), 256, 0, s>>>(X, g1, gw1, gb1, B, kSelInputs, kSelHidden);
        GI_LAUNCH_CHECK();
        return GI_OK;
    }
};

}  // namespace gi

using namespace gi;

extern "C" {

int gi_selector_load(const float *params, int32_t n_out, int32_t allowed_mask,
                     const float *feat_mean, const float *feat_scale) {
    if (!params || !feat_mean || !feat_scale) return fail(GI_ERR_INVALID, "null argument");
    if (n_out < 1 || n_out > kSelMaxOut) return fail(GI_ERR_RANGE, "n_out must be in [1,%d]", kSelMaxOut);
    if ((allowed_mask & ((1 << n_out) - 1)) == 0) return fail(GI_ERR_INVALID, "no allowed mode");
    int64_t np = param_count(n_out);
    // params in layer order: w1 [in][hid], b1, w2 [hid][hid], b2, w3 [hid][out], b3;
    // the blob keeps one row per unit of the layer each weight feeds
    const float *w1 = params, *b1 = w1 + kSelInputs * kSelHidden, *w2 = b1 + kSelHidden;
    const float *b2 = w2 + kSelHidden * kSelHidden, *w3 = b2 + kSelHidden, *b3 = w3 + kSelHidden * n_out;
    // w3 rows padded to kSelMaxOut outputs: one 16-byte load per hidden unit
    std::vector<uint16_t> h((size_t)kSelBlob, 0);
    uint16_t *o = h.data();
    for (int j = 0; j < kSelHidden; ++j)
        for (int k = 0; k < kSelInputs; ++k) *o++ = f32_to_bf16(w1[k * kSelHidden + j]);
    for (int j = 0; j < kSelHidden; ++j)
        for (int k = 0; k < kSelHidden; ++k) *o++ = f32_to_bf16(w2[k * kSelHidden + j]);
    for (int j = 0; j < kSelHidden; ++j)
        for (int k = 0; k < kSelMaxOut; ++k) *o++ = k < n_out ? f32_to_bf16(w3[j * n_out + k]) : 0;
    for (int i = 0; i < kSelHidden; ++i) *o++ = f32_to_bf16(b1[i]);
    for (int i = 0; i < kSelHidden; ++i) *o++ = f32_to_bf16(b2[i]);
    for (int i = 0; i < n_out; ++i) *o++ = f32_to_bf16(b3[i]);
    if (g_model.w) dfree(g_model.w);
    g_model = SelectorModel();
    GI_TRY(dalloc(&g_model.w, h.size()));
    GI_HIP(hipMemcpy(g_model.w, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice));
    for (int k = 0; k < kSelInputs; ++k) {
        g_model.norm[k] = feat_mean[k];
        g_model.norm[kSelInputs + k] = feat_scale[k];
    }
    g_model.n_out = n_out;
    g_model.allowed = allowed_mask & ((1 << n_out) - 1);
    g_model.loaded = true;
    ++g_model_version;
    return GI_OK;
}

int gi_selector_clear(void) {
    if (g_model.w) dfree(g_model.w);
    g_model = SelectorModel();
    ++g_model_version;
    return GI_OK;
}

int32_t gi_selector_loaded(void) { return g_model.loaded ? 1 : 0; }

int32_t gi_selector_param_count(int32_t n_out) {
    return n_out < 1 || n_out > kSelMaxOut ? -1 : (int32_t)param_count(n_out);
}

int gi_selector_loss_grad(const float *X, const float *T, const float *M, int32_t batch,
                          int32_t n_out, const float *params, float *grads, float *loss) {
    if (!X || !T || !M || !params || !grads || !loss) return fail(GI_ERR_INVALID, "null argument");
    if (batch < 1 || n_out < 1 || n_out > kSelMaxOut) return fail(GI_ERR_RANGE, "bad sizes");
    MlpWork w;
    GI_TRY(w.init(batch, n_out, X, T, M, params));
    GI_TRY(w.step(0));
    GI_HIP(hipMemcpy(grads, w.G, sizeof(float) * param_count(n_out), hipMemcpyDeviceToHost));
    GI_HIP(hipMemcpy(loss, w.loss, sizeof(float), hipMemcpyDeviceToHost));
    return GI_OK;
}

int gi_selector_train(const float *X, const float *T, const float *M, int32_t batch,
                      int32_t n_out, int32_t epochs, float lr, float weight_decay,
                      float *params, float *loss_history) {
    if (!X || !T || !M || !params) return fail(GI_ERR_INVALID, "null argument");
    if (batch < 1 || n_out < 1 || n_out > kSelMaxOut || epochs < 0) return fail(GI_ERR_RANGE, "bad sizes");
    MlpWork w;
    GI_TRY(w.init(batch, n_out, X, T, M, params));
    int64_t np = param_count(n_out);
    const float b1 = 0.9f, b2 = 0.999f;
    double c1 = 1.0, c2 = 1.0;
    for (int32_t ep = 0; ep < epochs; ++ep) {
        GI_TRY(w.step(0));
        c1 *= b1;
        c2 *= b2;
        adam_kernel<<<blocks_for(np), 256>>>(w.P, w.G, w.m, w.v, np, lr, b1, b2, (float)(1.0 - c1),
                                             (float)(1.0 - c2), 1e-8f, weight_decay);
        GI_LAUNCH_CHECK();
        if (loss_history) GI_HIP(hipMemcpy(&loss_history[ep], w.loss, sizeof(float), hipMemcpyDeviceToHost));
    }
    GI_HIP(hipMemcpy(params, w.P, sizeof(float) * np, hipMemcpyDeviceToHost));
    return GI_OK;
}

}  // extern "C"
