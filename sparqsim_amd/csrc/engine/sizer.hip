// Learned output sizer: a 4-layer FP32 MLP (16 -> 64 -> 64 -> 64 -> 1, ReLU)
// that predicts log2 of the group count G of a branching operator from
// features available before the count pass (input size, target width, bits
// that vary in and outside the target, sortedness, operator kind and a
// 6-bin popcount histogram of the target field, all from the key-statistics
// pass).  The engine sizes the expansion buffers with the prediction and
// skips the host readback of the exact count; the exact count, computed on
// the device anyway, checks the prediction and an overflow reruns the exact
// path.  No key or amplitude is ever approximated.
//
// Training runs on the GPU with the kernels below: a fused forward/backward
// over tiles of 16 samples (per-tile gradient partials, reduced in a fixed
// order, so a step is deterministic) and an Adam update.  Loss is an
// asymmetric squared error on the prediction (under-prediction, which costs a
// rerun, weighs `alpha` times more than over-prediction, which costs memory).
#include <cmath>
#include <cstring>
#include <mutex>

#include "engine.hpp"
#include "keyops.hpp"

namespace sparq {

constexpr int kF = SPARQ_SIZER_FEATURES, kH = 64, kS = 16;
constexpr int oW1 = 0, oB1 = oW1 + kH * kF, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH * kH,
              oW4 = oB3 + kH, oB4 = oW4 + kH, kP = oB4 + 1;
static_assert(kP == SPARQ_SIZER_PARAMS, "parameter count");

struct MlpLds {
    float w2[kH * kH];
    float w3[kH * kH];
    float x[kS][kF];
    float h1[kS][kH];
    float h2[kS][kH];
    float h3[kS][kH];
    float y[kS];
};

// Forward pass of the kS samples in L.x; weights W2/W3 are staged in LDS.
__device__ void mlp_forward_lds(const float* __restrict__ P, MlpLds& L) {
    for (int t = threadIdx.x; t < kH * kH; t += blockDim.x) {
        L.w2[t] = P[oW2 + t];
        L.w3[t] = P[oW3 + t];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kS * kH; t += blockDim.x) {
        const int s = t / kH, j = t % kH;
        float acc = P[oB1 + j];
        for (int f = 0; f < kF; ++f) acc += P[oW1 + j * kF + f] * L.x[s][f];
        L.h1[s][j] = fmaxf(acc, 0.f);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kS * kH; t += blockDim.x) {
        const int s = t / kH, j = t % kH;
        float acc = P[oB2 + j];
        for (int k = 0; k < kH; ++k) acc += L.w2[j * kH + k] * L.h1[s][k];
        L.h2[s][j] = fmaxf(acc, 0.f);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kS * kH; t += blockDim.x) {
        const int s = t / kH, j = t % kH;
        float acc = P[oB3 + j];
        for (int k = 0; k < kH; ++k) acc += L.w3[j * kH + k] * L.h2[s][k];
        L.h3[s][j] = fmaxf(acc, 0.f);
    }
    __syncthreads();
    if (threadIdx.x < kS) {
        const int s = threadIdx.x;
        float acc = P[oB4];
        for (int k = 0; k < kH; ++k) acc += P[oW4 + k] * L.h3[s][k];
        L.y[s] = acc;
    }
    __syncthreads();
}

__device__ void load_samples(MlpLds& L, const float* __restrict__ X, const uint32_t* __restrict__ idx, uint32_t n, uint32_t s0) {
    for (int t = threadIdx.x; t < kS * kF; t += blockDim.x) {
        const int s = t / kF, f = t % kF;
        const uint32_t i = s0 + s;
        L.x[s][f] = (i < n) ? X[(uint64_t)(idx ? idx[i] : i) * kF + f] : 0.f;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kBlock) mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X, uint32_t n,
                                                        float* __restrict__ Y) {
    __shared__ MlpLds L;
    const uint32_t s0 = blockIdx.x * kS;
    load_samples(L, X, nullptr, n, s0);
    mlp_forward_lds(P, L);
    if (threadIdx.x < kS && s0 + threadIdx.x < n) Y[s0 + threadIdx.x] = L.y[threadIdx.x];
}

// Forward + backward of one tile of samples idx[s0 .. s0+kS); writes the tile's
// gradient partial (kP values) and its loss sum at partial[blockIdx.x*(kP+1)].
// `scale` = 1/batch (mean loss).
__global__ void __launch_bounds__(kBlock) mlp_grad_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                     const float* __restrict__ Yt, const uint32_t* __restrict__ idx, uint32_t n,
                                                     float alpha, float scale, float* __restrict__ partial) {
    __shared__ MlpLds L;
    __shared__ float d3[kS][kH], d2[kS][kH], d1[kS][kH];
    __shared__ float dy[kS], ls[kS];
    const uint32_t s0 = blockIdx.x * kS;
    load_samples(L, X, idx, n, s0);
    mlp_forward_lds(P, L);
    if (threadIdx.x < kS) {
        const int s = threadIdx.x;
        const uint32_t i = s0 + s;
        float g = 0.f, l = 0.f;
        if (i < n) {
            const float r = L.y[s] - Yt[idx ? idx[i] : i];
            const float w = r >= 0.f ? 1.f : alpha;
            l = w * r * r;
            g = 2.f * w * r * scale;
        }
        dy[s] = g;
        ls[s] = l;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kS * kH; t += blockDim.x) {
        const int s = t / kH, k = t % kH;
        d3[s][k] = L.h3[s][k] > 0.f ? dy[s] * P[oW4 + k] : 0.f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kS * kH; t += blockDim.x) {
        const int s = t / kH, j = t % kH;
        float acc = 0.f;
        for (int k = 0; k < kH; ++k) acc += d3[s][k] * L.w3[k * kH + j];
        d2[s][j] = L.h2[s][j] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kS * kH; t += blockDim.x) {
        const int s = t / kH, j = t % kH;
        float acc = 0.f;
        for (int k = 0; k < kH; ++k) acc += d2[s][k] * L.w2[k * kH + j];
        d1[s][j] = L.h1[s][j] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    float* out = partial + (uint64_t)blockIdx.x * (kP + 1);
    for (int p = threadIdx.x; p < kP + 1; p += blockDim.x) {
        float acc = 0.f;
        if (p < oB1) {                       // W1[k][f]
            const int k = p / kF, f = p % kF;
            for (int s = 0; s < kS; ++s) acc += d1[s][k] * L.x[s][f];
        } else if (p < oW2) {                // b1
            for (int s = 0; s < kS; ++s) acc += d1[s][p - oB1];
        } else if (p < oB2) {                // W2[k][j]
            const int q = p - oW2, k = q / kH, j = q % kH;
            for (int s = 0; s < kS; ++s) acc += d2[s][k] * L.h1[s][j];
        } else if (p < oW3) {
            for (int s = 0; s < kS; ++s) acc += d2[s][p - oB2];
        } else if (p < oB3) {                // W3[k][j]
            const int q = p - oW3, k = q / kH, j = q % kH;
            for (int s = 0; s < kS; ++s) acc += d3[s][k] * L.h2[s][j];
        } else if (p < oW4) {
            for (int s = 0; s < kS; ++s) acc += d3[s][p - oB3];
        } else if (p < oB4) {                // W4[k]
            for (int s = 0; s < kS; ++s) acc += dy[s] * L.h3[s][p - oW4];
        } else if (p == oB4) {
            for (int s = 0; s < kS; ++s) acc += dy[s];
        } else {                             // loss sum of the tile
            for (int s = 0; s < kS; ++s) acc += ls[s];
        }
        out[p] = acc;
    }
}

// grad[p] = sum over tiles, in tile order.
__global__ void __launch_bounds__(kBlock) mlp_reduce_kernel(const float* __restrict__ partial, uint32_t tiles, float* __restrict__ grad) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= kP + 1) return;
    float acc = 0.f;
    for (uint32_t b = 0; b < tiles; ++b) acc += partial[(uint64_t)b * (kP + 1) + p];
    grad[p] = acc;
}

__global__ void __launch_bounds__(kBlock) adam_kernel(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
                                                 float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= kP) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// ---------------------------------------------------------------- features
// Built on the device from the key-statistics record, one sample, and (when
// P != nullptr) run through the network: out[0..kF) features, out[kF] y.
// One wave: lane j owns hidden unit j of every layer (weights read from L2).
__global__ void __launch_bounds__(64) sizer_feature_kernel(const uint64_t* __restrict__ stats, SizerQuery q, const float* __restrict__ P,
                                                      float* __restrict__ out) {
    __shared__ float x[kF], h1[kH], h2[kH];
    const int j = threadIdx.x;
    if (j == 0) {
        const uint32_t W = q.words;
        const uint64_t tmask = fmask(q.nd) << q.target_shift;
        int vt = 0, vo = 0;
        for (uint32_t w = 0; w < W; ++w) {
            const uint64_t vary = stats[w] ^ stats[W + w];
            const uint64_t in_t = (w == q.target_word) ? (vary & tmask) : 0;
            vt += __popcll(in_t);
            vo += __popcll(vary & ~in_t);
        }
        const bool sorted = stats[2 * W] == 0;
        const float lg = log2f((float)q.n);
        const float lower = fmaxf(0.f, lg - (float)vt), upper = fminf(lg, (float)vo);
        x[0] = lg / 32.f;
        x[1] = q.nd / 32.f;
        x[2] = vt / 32.f;
        x[3] = fminf((float)vo, 40.f) / 32.f;
        x[4] = lower / 32.f;
        x[5] = upper / 32.f;
        x[6] = sorted ? 1.f : 0.f;
        x[7] = q.kind == 0 ? 1.f : 0.f;
        x[8] = q.kind == 1 ? 1.f : 0.f;
        x[9] = q.kind == 2 ? 1.f : 0.f;
        for (int b = 0; b < kStatsHistBins; ++b) x[10 + b] = (float)stats[3 * W + 2 + b] / (float)q.n;
    }
    __syncthreads();
    if (j < kF) out[j] = x[j];
    if (!P) {
        if (j == 0) out[kF] = 0.f;
        return;
    }
    float acc = P[oB1 + j];
    for (int f = 0; f < kF; ++f) acc += P[oW1 + j * kF + f] * x[f];
    h1[j] = fmaxf(acc, 0.f);
    __syncthreads();
    acc = P[oB2 + j];
    for (int k = 0; k < kH; ++k) acc += P[oW2 + j * kH + k] * h1[k];
    h2[j] = fmaxf(acc, 0.f);
    __syncthreads();
    acc = P[oB3 + j];
    for (int k = 0; k < kH; ++k) acc += P[oW3 + j * kH + k] * h2[k];
    float y = P[oW4 + j] * fmaxf(acc, 0.f);
    for (int off = 32; off > 0; off >>= 1) y += __shfl_xor(y, off, 64);
    if (j == 0) out[kF] = y + P[oB4];
}

// ---------------------------------------------------------------- engine state
namespace {
struct SizerGlobal {
    std::mutex mu;
    float* params = nullptr;    // device copy of the active network
    float margin = 0.5f;        // log2 units added to the prediction
    bool active = false, logging = false;
    std::vector<float> log_x, log_y;
    uint64_t predicted = 0, overflowed = 0, exact = 0;
} g_sizer;
}  // namespace

bool sizer_wanted(bool& predict) {
    std::lock_guard<std::mutex> lock(g_sizer.mu);
    predict = g_sizer.active && g_sizer.params;
    return predict || g_sizer.logging;
}

void launch_sizer_features(State& st, const uint64_t* stats, const SizerQuery& q, bool predict, float* out) {
    const float* P = nullptr;
    if (predict) {
        std::lock_guard<std::mutex> lock(g_sizer.mu);
        P = g_sizer.params;
    }
    {
        SPQ_PROF(st.stream, "sizer_feature_kernel", 0.0);
        sizer_feature_kernel<<<1, 64, 0, st.stream>>>(stats, q, P, out);
    }
    SPQ_HIP(hipGetLastError());
}

float sizer_margin() {
    std::lock_guard<std::mutex> lock(g_sizer.mu);
    return g_sizer.margin;
}

void sizer_record(const float* features, float target) {
    std::lock_guard<std::mutex> lock(g_sizer.mu);
    if (!g_sizer.logging) return;
    g_sizer.log_x.insert(g_sizer.log_x.end(), features, features + kF);
    g_sizer.log_y.push_back(target);
}

void sizer_count(int what) {
    std::lock_guard<std::mutex> lock(g_sizer.mu);
    if (what == 0) ++g_sizer.exact;
    else if (what == 1) ++g_sizer.predicted;
    else ++g_sizer.overflowed;
}

// ---------------------------------------------------------------- training helpers
namespace {
struct Scratch {
    hipStream_t s = nullptr;
    Scratch() { SPQ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~Scratch() {
        if (s) {
            (void)hipStreamSynchronize(s);
            forget_stream(s);
            (void)hipStreamDestroy(s);
        }
    }
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    hipStream_t s;
    DevBuf(size_t n, hipStream_t st) : s(st) { p = dalloc<T>(n ? n : 1, st); }
    ~DevBuf() { dfree(p, s); }
};

void upload(hipStream_t s, void* d, const void* h, size_t bytes) {
    if (bytes) SPQ_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
}

// One gradient evaluation over samples idx[0..n) (idx == nullptr: all in order).
void grad_step(hipStream_t s, const float* P, const float* X, const float* Y, const uint32_t* idx, uint32_t n, float alpha,
               float* partial, float* grad) {
    const uint32_t tiles = (n + kS - 1) / kS;
    mlp_grad_kernel<<<tiles, kBlock, 0, s>>>(P, X, Y, idx, n, alpha, 1.f / (float)n, partial);
    SPQ_HIP(hipGetLastError());
    mlp_reduce_kernel<<<(kP + 1 + kBlock - 1) / kBlock, kBlock, 0, s>>>(partial, tiles, grad);
    SPQ_HIP(hipGetLastError());
}

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

}  // namespace sparq

using namespace sparq;

extern "C" {

int sparq_mlp_forward(const float* params, const float* x, uint32_t n, float* y) {
    return guarded_call([&] {
        SPQ_REQUIRE(params && (x || !n) && (y || !n), "null argument");
        ensure_device();
        if (n == 0) return;
        Scratch sc;
        DevBuf<float> P(kP, sc.s), X((size_t)n * kF, sc.s), Y(n, sc.s);
        upload(sc.s, P.p, params, kP * 4);
        upload(sc.s, X.p, x, (size_t)n * kF * 4);
        mlp_forward_kernel<<<(n + kS - 1) / kS, kBlock, 0, sc.s>>>(P.p, X.p, n, Y.p);
        SPQ_HIP(hipGetLastError());
        SPQ_HIP(hipMemcpyAsync(y, Y.p, (size_t)n * 4, hipMemcpyDeviceToHost, sc.s));
        SPQ_HIP(hipStreamSynchronize(sc.s));
    });
}

int sparq_mlp_grad(const float* params, const float* x, const float* y, uint32_t n, float alpha, float* grad, float* loss) {
    return guarded_call([&] {
        SPQ_REQUIRE(params && x && y && grad && loss && n > 0, "bad argument");
        ensure_device();
        Scratch sc;
        const uint32_t tiles = (n + kS - 1) / kS;
        DevBuf<float> P(kP, sc.s), X((size_t)n * kF, sc.s), Y(n, sc.s), part((size_t)tiles * (kP + 1), sc.s), G(kP + 1, sc.s);
        upload(sc.s, P.p, params, kP * 4);
        upload(sc.s, X.p, x, (size_t)n * kF * 4);
        upload(sc.s, Y.p, y, (size_t)n * 4);
        grad_step(sc.s, P.p, X.p, Y.p, nullptr, n, alpha, part.p, G.p);
        std::vector<float> h(kP + 1);
        SPQ_HIP(hipMemcpyAsync(h.data(), G.p, (kP + 1) * 4, hipMemcpyDeviceToHost, sc.s));
        SPQ_HIP(hipStreamSynchronize(sc.s));
        std::memcpy(grad, h.data(), kP * 4);
        *loss = h[kP] / (float)n;
    });
}

int sparq_mlp_train(float* params, const float* x, const float* y, uint32_t n, uint32_t steps, uint32_t batch, float lr,
                    float alpha, uint64_t seed, float* loss_history) {
    return guarded_call([&] {
        SPQ_REQUIRE(params && x && y && n > 0 && batch > 0 && steps > 0, "bad argument");
        ensure_device();
        batch = std::min(batch, n);
        Scratch sc;
        const uint32_t tiles = (batch + kS - 1) / kS;
        DevBuf<float> P(kP, sc.s), X((size_t)n * kF, sc.s), Y(n, sc.s), part((size_t)tiles * (kP + 1), sc.s), G(kP + 1, sc.s),
            M(kP, sc.s), V(kP, sc.s), hist(steps, sc.s);
        DevBuf<uint32_t> order((size_t)n, sc.s);
        upload(sc.s, P.p, params, kP * 4);
        upload(sc.s, X.p, x, (size_t)n * kF * 4);
        upload(sc.s, Y.p, y, (size_t)n * 4);
        SPQ_HIP(hipMemsetAsync(M.p, 0, kP * 4, sc.s));
        SPQ_HIP(hipMemsetAsync(V.p, 0, kP * 4, sc.s));
        // minibatches walk a seeded permutation, reshuffled every epoch
        std::vector<uint32_t> perm(n);
        for (uint32_t i = 0; i < n; ++i) perm[i] = i;
        uint64_t rs = seed;
        uint32_t at = n;
        const float b1 = 0.9f, b2 = 0.999f;
        float c1 = 1.f, c2 = 1.f;
        std::vector<std::vector<uint32_t>> epochs;
        for (uint32_t t = 0; t < steps; ++t) {
            if (at + batch > n) {
                for (uint32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(rs) % (i + 1)]);
                SPQ_HIP(hipStreamSynchronize(sc.s));   // the previous permutation upload has been consumed
                upload(sc.s, order.p, perm.data(), (size_t)n * 4);
                at = 0;
            }
            grad_step(sc.s, P.p, X.p, Y.p, order.p + at, batch, alpha, part.p, G.p);
            SPQ_HIP(hipMemcpyAsync(hist.p + t, G.p + kP, 4, hipMemcpyDeviceToDevice, sc.s));
            c1 *= b1;
            c2 *= b2;
            adam_kernel<<<(kP + kBlock - 1) / kBlock, kBlock, 0, sc.s>>>(P.p, G.p, M.p, V.p, lr, b1, b2, 1e-8f, 1.f - c1, 1.f - c2);
            SPQ_HIP(hipGetLastError());
            at += batch;
        }
        SPQ_HIP(hipMemcpyAsync(params, P.p, kP * 4, hipMemcpyDeviceToHost, sc.s));
        std::vector<float> h(steps);
        SPQ_HIP(hipMemcpyAsync(h.data(), hist.p, steps * 4, hipMemcpyDeviceToHost, sc.s));
        SPQ_HIP(hipStreamSynchronize(sc.s));
        if (loss_history)
            for (uint32_t t = 0; t < steps; ++t) loss_history[t] = h[t] / (float)batch;
    });
}

int sparq_sizer_load(const float* params, uint32_t count, float margin) {
    return guarded_call([&] {
        SPQ_REQUIRE(params && count == (uint32_t)kP, "the sizer takes SPARQ_SIZER_PARAMS floats");
        SPQ_REQUIRE(margin >= 0.f && margin < 20.f, "margin out of range");
        ensure_device();
        float* d = nullptr;
        SPQ_HIP(hipMalloc(&d, kP * 4));
        SPQ_HIP(hipMemcpy(d, params, kP * 4, hipMemcpyHostToDevice));
        std::lock_guard<std::mutex> lock(g_sizer.mu);
        if (g_sizer.params) {
            (void)hipDeviceSynchronize();   // no launch still reads the old copy
            (void)hipFree(g_sizer.params);
        }
        g_sizer.params = d;
        g_sizer.margin = margin;
    });
}

int sparq_sizer_enable(int on) {
    return guarded_call([&] {
        std::lock_guard<std::mutex> lock(g_sizer.mu);
        SPQ_REQUIRE(!on || g_sizer.params, "load the sizer's parameters first");
        g_sizer.active = on != 0;
    });
}

int sparq_sizer_log(int on) {
    return guarded_call([&] {
        std::lock_guard<std::mutex> lock(g_sizer.mu);
        g_sizer.logging = on != 0;
    });
}

int sparq_sizer_log_fetch(float* x, float* y, uint32_t capacity, uint32_t* count, int clear) {
    return guarded_call([&] {
        SPQ_REQUIRE(count != nullptr, "null count");
        std::lock_guard<std::mutex> lock(g_sizer.mu);
        const uint32_t n = (uint32_t)g_sizer.log_y.size();
        *count = n;
        const uint32_t c = std::min(n, capacity);
        if (c && x) std::memcpy(x, g_sizer.log_x.data(), (size_t)c * kF * 4);
        if (c && y) std::memcpy(y, g_sizer.log_y.data(), (size_t)c * 4);
        if (clear) {
            g_sizer.log_x.clear();
            g_sizer.log_y.clear();
        }
    });
}

int sparq_sizer_stats(uint64_t out[3]) {
    return guarded_call([&] {
        SPQ_REQUIRE(out != nullptr, "null output");
        std::lock_guard<std::mutex> lock(g_sizer.mu);
        out[0] = g_sizer.exact;
        out[1] = g_sizer.predicted;
        out[2] = g_sizer.overflowed;
    });
}

}  // extern "C"
