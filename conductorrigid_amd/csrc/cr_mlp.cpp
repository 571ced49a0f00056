// cr_mlp.cpp -- C ABI of the learned warm start's network (include/conductorrigid.h):
// FP32 master parameters, Adam moments, bf16 compute copies and training buffers on
// the device.  No reference counterpart.
#include "../../include/conductorrigid.h"
#include "cr_host.h"
#include "cr_kernels.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <vector>

using namespace cr;

static_assert(CR_MLP_IN == kMlpIn && CR_MLP_HIDDEN == kMlpHidden && CR_MLP_OUT == kMlpOut &&
                  CR_MLP_PARAMS == kMlpParams,
              "ABI constants out of step with the kernels");

struct cr_mlp {
    int device = 0;
    int step = 0;
    DevBuf theta, grad, m1, m2;              // FP32 [kMlpParams]
    DevBuf W1t, W2t, W2n, W3;                // bf16 compute copies
    DevBuf z1, z2, du, part, h1, h2, dz1, dz2, ybuf;
    int cap = 0;                             // rows the training buffers hold
    bool packed = false;

    MlpDevWeights weights() const
    {
        MlpDevWeights w;
        w.W1t = W1t.as<unsigned short>();
        w.W2t = W2t.as<unsigned short>();
        w.W2n = W2n.as<unsigned short>();
        w.W3 = W3.as<unsigned short>();
        w.b1 = theta.as<float>() + kMlpOffB1;
        w.b2 = theta.as<float>() + kMlpOffB2;
        w.b3 = theta.as<float>() + kMlpOffB3;
        return w;
    }
};

namespace {

constexpr int kSlices = 16;

int pack(cr_mlp* m, hipStream_t st)
{
    CR_HIP(launch_mlp_pack(m->theta.as<float>(), m->weights(), st));
    m->packed = true;
    return CR_OK;
}

int ensure_train(cr_mlp* m, int M)
{
    if (M <= m->cap) return CR_OK;
    const size_t rows = (size_t)M, H = kMlpHidden;
    CR_HIP(m->z1.ensure(sizeof(float) * rows * H));
    CR_HIP(m->z2.ensure(sizeof(float) * rows * H));
    CR_HIP(m->du.ensure(sizeof(float) * rows * kMlpOut));
    CR_HIP(m->ybuf.ensure(sizeof(float) * rows));
    CR_HIP(m->h1.ensure(2 * rows * H));
    CR_HIP(m->h2.ensure(2 * rows * H));
    CR_HIP(m->dz1.ensure(2 * rows * H));
    CR_HIP(m->dz2.ensure(2 * rows * H));
    CR_HIP(m->part.ensure(sizeof(float) * kSlices * (size_t)kMlpIn * H));
    m->cap = M;
    return CR_OK;
}

MlpTrainBufs bufs(cr_mlp* m)
{
    MlpTrainBufs b;
    b.z1 = m->z1.as<float>();
    b.z2 = m->z2.as<float>();
    b.du = m->du.as<float>();
    b.part = m->part.as<float>();
    b.h1 = m->h1.as<unsigned short>();
    b.h2 = m->h2.as<unsigned short>();
    b.dz1 = m->dz1.as<unsigned short>();
    b.dz2 = m->dz2.as<unsigned short>();
    b.slices = kSlices;
    return b;
}

}  // namespace

extern "C" {

int cr_mlp_create(cr_mlp** out, int device, unsigned long long seed)
{
    if (!out) return set_error(CR_ERR_INVALID, "null output pointer");
    *out = nullptr;
    CR_HIP(hipSetDevice(device));
    cr_mlp* m = new cr_mlp();
    m->device = device;
    const size_t P = kMlpParams;
    if (m->theta.ensure(4 * P) != hipSuccess || m->grad.ensure(4 * P) != hipSuccess ||
        m->m1.ensure(4 * P) != hipSuccess || m->m2.ensure(4 * P) != hipSuccess ||
        m->W1t.ensure(2 * (size_t)kMlpIn * kMlpHidden) != hipSuccess ||
        m->W2t.ensure(2 * (size_t)kMlpHidden * kMlpHidden) != hipSuccess ||
        m->W2n.ensure(2 * (size_t)kMlpHidden * kMlpHidden) != hipSuccess ||
        m->W3.ensure(2 * (size_t)kMlpHidden * kMlpOut) != hipSuccess) {
        delete m;
        return set_error(CR_ERR_DEVICE, "allocation of network parameters failed");
    }
    // He-uniform initialisation for the SiLU layers, small output layer, zero biases
    std::vector<float> h(P, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int count, float gain) {
        const float a = gain * std::sqrt(6.0f / (float)fan_in);
        std::uniform_real_distribution<float> U(-a, a);
        for (int i = 0; i < count; ++i) h[off + i] = U(rng);
    };
    fill(kMlpOffW1, kMlpIn, kMlpIn * kMlpHidden, 1.0f);
    fill(kMlpOffW2, kMlpHidden, kMlpHidden * kMlpHidden, 1.0f);
    fill(kMlpOffW3, kMlpHidden, kMlpHidden * kMlpOut, 0.1f);
    int rc = cr_mlp_set_params(m, h.data());
    if (rc != CR_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return CR_OK;
}

void cr_mlp_destroy(cr_mlp* m) { delete m; }

int cr_mlp_set_params(cr_mlp* m, const float* theta)
{
    if (!m || !theta) return set_error(CR_ERR_INVALID, "null argument");
    const size_t bytes = 4 * (size_t)kMlpParams;
    CR_HIP(hipMemcpy(m->theta.p, theta, bytes, hipMemcpyHostToDevice));
    CR_HIP(hipMemset(m->m1.p, 0, bytes));
    CR_HIP(hipMemset(m->m2.p, 0, bytes));
    m->step = 0;
    CR_RET(pack(m, nullptr));
    CR_HIP(hipStreamSynchronize(nullptr));
    return CR_OK;
}

int cr_mlp_get_params(const cr_mlp* m, float* theta)
{
    if (!m || !theta) return set_error(CR_ERR_INVALID, "null argument");
    CR_HIP(hipMemcpy(theta, m->theta.p, 4 * (size_t)kMlpParams, hipMemcpyDeviceToHost));
    return CR_OK;
}

int cr_mlp_predict(cr_mlp* m, int M, const unsigned short* X, float* y, void* stream)
{
    if (!m || M < 0 || (M > 0 && (!X || !y))) return set_error(CR_ERR_INVALID, "bad argument");
    CR_HIP(launch_mlp_forward(M, X, m->weights(), y, nullptr, (hipStream_t)stream));
    return CR_OK;
}

int cr_mlp_gradient(cr_mlp* m, int M, const unsigned short* X, const float* t, float* y, float* grad, void* stream)
{
    if (!m || M <= 0 || !X || !t || !grad) return set_error(CR_ERR_INVALID, "bad argument");
    CR_RET(ensure_train(m, M));
    float* yy = y ? y : m->ybuf.as<float>();
    CR_HIP(launch_mlp_gradient(M, X, t, m->weights(), bufs(m), yy, grad, (hipStream_t)stream));
    return CR_OK;
}

int cr_mlp_train_step(cr_mlp* m, int M, const unsigned short* X, const float* t, float lr, float* y, void* stream)
{
    if (!m || M <= 0 || !X || !t || !(lr > 0.0f)) return set_error(CR_ERR_INVALID, "bad argument");
    hipStream_t st = (hipStream_t)stream;
    CR_RET(ensure_train(m, M));
    float* yy = y ? y : m->ybuf.as<float>();
    CR_HIP(launch_mlp_gradient(M, X, t, m->weights(), bufs(m), yy, m->grad.as<float>(), st));
    ++m->step;
    CR_HIP(launch_mlp_adam(m->theta.as<float>(), m->grad.as<float>(), m->m1.as<float>(), m->m2.as<float>(), lr,
                           m->step, st));
    CR_RET(pack(m, st));
    return CR_OK;
}

}  // extern "C"
