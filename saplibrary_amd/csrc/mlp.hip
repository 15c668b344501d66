// Small fully connected network for the learned option selector: layers
// in -> h -> h -> h -> out, ReLU between them, FP32.  Forward, backward
// (masked mean squared error on the outputs), and Adam, as hand-written
// kernels; every reduction is a fixed-order loop (no atomics), so training
// is deterministic.  Sizes are small (inputs tens, hidden 64, batches up
// to a few thousand): one thread per output element is the right shape.
#include "sap_internal.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <random>

namespace sap_amd {

namespace {

constexpr int TB = 256;

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + TB - 1) / TB); }

// Z[b, j] = sum_i X[b, i] W[i, j] + bias[j]; A = relu(Z) (or Z for the
// last layer).  X: B x I, W: I x J (row-major), Z/A: B x J.
__global__ void dense_forward_kernel(int B, int I, int J, const float* __restrict__ X, const float* __restrict__ W,
                                     const float* __restrict__ bias, float* __restrict__ Z, float* __restrict__ A,
                                     int relu)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)B * J) return;
    const int b = (int)(e / J), j = (int)(e % J);
    float s = bias[j];
    for (int i = 0; i < I; ++i) s = fmaf(X[(int64_t)b * I + i], W[(int64_t)i * J + j], s);
    Z[e] = s;
    A[e] = relu ? fmaxf(s, 0.f) : s;
}

// dY = 2 M (Y - T) / count(M), loss partials per row: L_b = sum_j M (Y - T)^2
__global__ void mse_grad_kernel(int B, int J, const float* __restrict__ Y, const float* __restrict__ T,
                                const float* __restrict__ M, float inv_count, float* __restrict__ dY,
                                float* __restrict__ row_loss)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float l = 0.f;
    for (int j = 0; j < J; ++j) {
        const int64_t e = (int64_t)b * J + j;
        const float d = M[e] * (Y[e] - T[e]);
        l = fmaf(d, Y[e] - T[e], l);
        dY[e] = 2.f * d * inv_count;
    }
    row_loss[b] = l;
}

// loss = sum_b row_loss[b] * inv_count (one thread, fixed order)
__global__ void loss_sum_kernel(int B, const float* __restrict__ row_loss, float inv_count, float* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += row_loss[b];
    *out = s * inv_count;
}

// dZ = dA * relu'(Z)
__global__ void relu_back_kernel(int64_t n, const float* __restrict__ Z, float* __restrict__ dA)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n && Z[e] <= 0.f) dA[e] = 0.f;
}

// dW[i, j] = sum_b X[b, i] dZ[b, j]; db[j] = sum_b dZ[b, j]
__global__ void dense_wgrad_kernel(int B, int I, int J, const float* __restrict__ X, const float* __restrict__ dZ,
                                   float* __restrict__ dW, float* __restrict__ db)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)(I + 1) * J) return;
    const int i = (int)(e / J), j = (int)(e % J);
    float s = 0.f;
    if (i < I) {
        for (int b = 0; b < B; ++b) s = fmaf(X[(int64_t)b * I + i], dZ[(int64_t)b * J + j], s);
        dW[(int64_t)i * J + j] = s;
    } else {
        for (int b = 0; b < B; ++b) s += dZ[(int64_t)b * J + j];
        db[j] = s;
    }
}

// dX[b, i] = sum_j dZ[b, j] W[i, j]
__global__ void dense_xgrad_kernel(int B, int I, int J, const float* __restrict__ dZ, const float* __restrict__ W,
                                   float* __restrict__ dX)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)B * I) return;
    const int b = (int)(e / I), i = (int)(e % I);
    float s = 0.f;
    for (int j = 0; j < J; ++j) s = fmaf(dZ[(int64_t)b * J + j], W[(int64_t)i * J + j], s);
    dX[e] = s;
}

// Adam (Kingma & Ba 2015) on all parameters: m, v moments, bias-corrected
__global__ void adam_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

// Parameters of layer l (l = 0..3): W_l (I_l x J_l, row-major) then b_l
// (J_l), packed in one array in that order.
class Mlp {
public:
    int in, hid, out;
    static constexpr int L = 4;
    int I[L], J[L];
    size_t woff[L], boff[L], nparam = 0;
    DevBuf<float> p, g, m, v;
    // activations of the last forward (batch cap_)
    DevBuf<float> Z[L], A[L], dA[L], dX0, rowloss, loss;
    int cap_ = 0;
    int64_t step_ = 0;
    hipStream_t stream = nullptr;

    Mlp(int in_, int hid_, int out_, uint64_t seed) : in(in_), hid(hid_), out(out_)
    {
        const int dims[L + 1] = {in, hid, hid, hid, out};
        for (int l = 0; l < L; ++l) {
            I[l] = dims[l];
            J[l] = dims[l + 1];
            woff[l] = nparam;
            nparam += (size_t)I[l] * J[l];
            boff[l] = nparam;
            nparam += J[l];
        }
        p.resize(nparam);
        g.resize(nparam);
        m.resize(nparam);
        v.resize(nparam);
        loss.resize(1);
        // He initialisation (normal, std sqrt(2 / fan_in)), zero biases
        std::vector<float> h(nparam, 0.f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < L; ++l) {
            std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)I[l]));
            for (size_t e = 0; e < (size_t)I[l] * J[l]; ++e) h[woff[l] + e] = nd(rng);
        }
        SAP_HIP_CHECK(hipMemcpy(p.get(), h.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
        zero_moments();
    }
    void zero_moments()
    {
        std::vector<float> z(nparam, 0.f);
        SAP_HIP_CHECK(hipMemcpy(m.get(), z.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
        SAP_HIP_CHECK(hipMemcpy(v.get(), z.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
        step_ = 0;
    }
    void ensure(int B)
    {
        if (B <= cap_) return;
        for (int l = 0; l < L; ++l) {
            Z[l].resize((size_t)B * J[l]);
            A[l].resize((size_t)B * J[l]);
            dA[l].resize((size_t)B * J[l]);
        }
        dX0.resize((size_t)B * in);
        rowloss.resize(B);
        cap_ = B;
    }
    void forward(const float* X, int B)
    {
        ensure(B);
        const float* x = X;
        for (int l = 0; l < L; ++l) {
            hipLaunchKernelGGL(dense_forward_kernel, dim3(blocks_for((int64_t)B * J[l])), dim3(TB), 0, stream, B, I[l],
                               J[l], x, p.get() + woff[l], p.get() + boff[l], Z[l].get(), A[l].get(), l < L - 1);
            SAP_LAUNCH_CHECK();
            x = A[l].get();
        }
    }
    // gradient of the masked mean squared error into g; the loss into loss[0]
    void backward(const float* X, const float* T, const float* M, int B, float inv_count)
    {
        hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks_for(B)), dim3(TB), 0, stream, B, out, A[L - 1].get(), T, M,
                           inv_count, dA[L - 1].get(), rowloss.get());
        SAP_LAUNCH_CHECK();
        hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, stream, B, rowloss.get(), inv_count, loss.get());
        SAP_LAUNCH_CHECK();
        for (int l = L - 1; l >= 0; --l) {
            if (l < L - 1) {
                hipLaunchKernelGGL(relu_back_kernel, dim3(blocks_for((int64_t)B * J[l])), dim3(TB), 0, stream,
                                   (int64_t)B * J[l], Z[l].get(), dA[l].get());
                SAP_LAUNCH_CHECK();
            }
            const float* xin = l == 0 ? X : A[l - 1].get();
            hipLaunchKernelGGL(dense_wgrad_kernel, dim3(blocks_for((int64_t)(I[l] + 1) * J[l])), dim3(TB), 0, stream,
                               B, I[l], J[l], xin, dA[l].get(), g.get() + woff[l], g.get() + boff[l]);
            SAP_LAUNCH_CHECK();
            float* dx = l == 0 ? dX0.get() : dA[l - 1].get();
            hipLaunchKernelGGL(dense_xgrad_kernel, dim3(blocks_for((int64_t)B * I[l])), dim3(TB), 0, stream, B, I[l],
                               J[l], dA[l].get(), p.get() + woff[l], dx);
            SAP_LAUNCH_CHECK();
        }
    }
    void adam(float lr)
    {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        ++step_;
        const float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for((int64_t)nparam)), dim3(TB), 0, stream, (int64_t)nparam,
                           p.get(), g.get(), m.get(), v.get(), lr, b1, b2, eps, c1, c2);
        SAP_LAUNCH_CHECK();
    }
};

}  // namespace sap_amd

struct sap_mlp {
    std::unique_ptr<sap_amd::Mlp> m;
    std::string err;
};

namespace {
template <class F>
int mlp_guard(sap_mlp* h, F&& f)
{
    try {
        return f();
    } catch (const sap_amd::Error& e) {
        if (h) h->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (h) h->err = e.what();
        return SAP_ERR_HIP;
    }
}

float mask_count(const float* d_M, int64_t n, hipStream_t st)
{
    std::vector<float> h(n);
    SAP_HIP_CHECK(hipMemcpyAsync(h.data(), d_M, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    SAP_HIP_CHECK(hipStreamSynchronize(st));
    double c = 0;
    for (float x : h) c += x;
    return (float)c;
}
}  // namespace

extern "C" {

sap_mlp* sap_mlp_create(int n_in, int n_hidden, int n_out, uint64_t seed)
{
    if (n_in < 1 || n_hidden < 1 || n_out < 1) return nullptr;
    auto* h = new (std::nothrow) sap_mlp();
    if (!h) return nullptr;
    if (mlp_guard(h, [&] {
            h->m = std::make_unique<sap_amd::Mlp>(n_in, n_hidden, n_out, seed);
            return 0;
        }) != 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void sap_mlp_destroy(sap_mlp* h) { delete h; }

int64_t sap_mlp_num_params(const sap_mlp* h) { return h ? (int64_t)h->m->nparam : SAP_ERR_BADARG; }

int sap_mlp_get_params(const sap_mlp* h, float* out)
{
    if (!h || !out) return SAP_ERR_BADARG;
    return mlp_guard(const_cast<sap_mlp*>(h), [&] {
        SAP_HIP_CHECK(hipMemcpy(out, h->m->p.get(), sizeof(float) * h->m->nparam, hipMemcpyDeviceToHost));
        return 0;
    });
}

int sap_mlp_set_params(sap_mlp* h, const float* in)
{
    if (!h || !in) return SAP_ERR_BADARG;
    return mlp_guard(h, [&] {
        SAP_HIP_CHECK(hipMemcpy(h->m->p.get(), in, sizeof(float) * h->m->nparam, hipMemcpyHostToDevice));
        h->m->zero_moments();
        return 0;
    });
}

int sap_mlp_forward(sap_mlp* h, const float* d_X, int batch, float* d_Y)
{
    if (!h || !d_X || !d_Y || batch < 1) return SAP_ERR_BADARG;
    return mlp_guard(h, [&] {
        auto& m = *h->m;
        m.forward(d_X, batch);
        SAP_HIP_CHECK(hipMemcpyAsync(d_Y, m.A[sap_amd::Mlp::L - 1].get(), sizeof(float) * (size_t)batch * m.out,
                                     hipMemcpyDeviceToDevice, m.stream));
        SAP_HIP_CHECK(hipStreamSynchronize(m.stream));
        return 0;
    });
}

int sap_mlp_gradient(sap_mlp* h, const float* d_X, const float* d_T, const float* d_M, int batch, float* h_grad,
                     float* loss)
{
    if (!h || !d_X || !d_T || !d_M || batch < 1) return SAP_ERR_BADARG;
    return mlp_guard(h, [&] {
        auto& m = *h->m;
        const float cnt = mask_count(d_M, (int64_t)batch * m.out, m.stream);
        m.forward(d_X, batch);
        m.backward(d_X, d_T, d_M, batch, cnt > 0 ? 1.f / cnt : 0.f);
        if (h_grad)
            SAP_HIP_CHECK(hipMemcpyAsync(h_grad, m.g.get(), sizeof(float) * m.nparam, hipMemcpyDeviceToHost, m.stream));
        if (loss) SAP_HIP_CHECK(hipMemcpyAsync(loss, m.loss.get(), sizeof(float), hipMemcpyDeviceToHost, m.stream));
        SAP_HIP_CHECK(hipStreamSynchronize(m.stream));
        return 0;
    });
}

int sap_mlp_train(sap_mlp* h, const float* d_X, const float* d_T, const float* d_M, int batch, int epochs, float lr,
                  float* h_losses)
{
    if (!h || !d_X || !d_T || !d_M || batch < 1 || epochs < 0) return SAP_ERR_BADARG;
    return mlp_guard(h, [&] {
        auto& m = *h->m;
        const float cnt = mask_count(d_M, (int64_t)batch * m.out, m.stream);
        const float inv = cnt > 0 ? 1.f / cnt : 0.f;
        for (int e = 0; e < epochs; ++e) {
            m.forward(d_X, batch);
            m.backward(d_X, d_T, d_M, batch, inv);
            m.adam(lr);
            if (h_losses)
                SAP_HIP_CHECK(hipMemcpyAsync(h_losses + e, m.loss.get(), sizeof(float), hipMemcpyDeviceToHost,
                                             m.stream));
        }
        SAP_HIP_CHECK(hipStreamSynchronize(m.stream));
        return 0;
    });
}

const char* sap_mlp_last_error(const sap_mlp* h) { return h ? h->err.c_str() : ""; }

}  // extern "C"
