// Bracket-position network entry points of include/rfd.h: creation, weights
// I/O, forward, gradients and Adam training, all on the device kernels of
// mlp_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "mlp_internal.hpp"
#include "mlp_kernels.hpp"

using namespace rfd;

namespace {

using rfd_host::set_err;

constexpr char kMagic[8] = {'R', 'F', 'D', 'M', 'L', 'P', '0', '2'};

unsigned blocks(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

GemmArgs gemm_args(const uint16_t *A, long long sai, long long sak, const uint16_t *B,
                   long long sbk, long long sbj, int M, int N, int K) {
    GemmArgs g{};
    g.A = A;
    g.B = B;
    g.M = M;
    g.N = N;
    g.K = K;
    g.sai = sai;
    g.sak = sak;
    g.sbk = sbk;
    g.sbj = sbj;
    return g;
}

// The layers' products (K = 400 or 256) stage their whole K at once; the
// weight gradients' (K = the batch's rows) stream it.
template <int EPI>
void gemm(hipStream_t st, const GemmArgs &g) {
    dim3 grid((unsigned)((g.N + 63) / 64), (unsigned)((g.M + 63) / 64));
    if (g.K <= kGemmStepsAll * kGemmK) {
        hipLaunchKernelGGL((mlp_gemm<EPI, kGemmStepsAll>), grid, dim3(256), 0, st, g);
    } else {
        hipLaunchKernelGGL((mlp_gemm<EPI, 1>), grid, dim3(256), 0, st, g);
    }
}

// Backward + gradient buffer fill from the activations of the last forward.
void backward_dev(rfd_mlp *n, const uint16_t *x0, int rows, const float *labels) {
    hipStream_t st = n->ctx->stream;
    hipLaunchKernelGGL(mlp_head, dim3(blocks(rows)), dim3(256), 0, st, n->act[3], n->params, rows,
                       labels, rows, n->p, n->dzh, n->loss);
    hipLaunchKernelGGL(mlp_head_back, dim3(blocks((long long)rows * MLP_H)), dim3(256), 0, st,
                       n->act[3], n->params, n->dzh, rows, n->dz[1], n->grads);
    int cur = 1;
    for (int l = MLP_LAYERS - 1; l >= 0; --l) {
        const int K = l == 0 ? MLP_IN : MLP_H;
        const uint16_t *X = l == 0 ? x0 : n->act[l - 1];
        const uint16_t *dZ = n->dz[cur];
        // dW_l = X^T dZ : [K x 256]
        GemmArgs g = gemm_args(X, 1, K, dZ, MLP_H, 1, K, MLP_H, rows);
        g.Cf = n->grads + mlp_w_off(l);
        gemm<EPI_F32>(st, g);
        hipLaunchKernelGGL(mlp_colsum, dim3(1), dim3(256), 0, st, dZ, rows, MLP_H,
                           n->grads + mlp_b_off(l));
        if (l > 0) {
            // dZ_{l-1} = (dZ W_l^T) masked by a_{l-1} > 0
            GemmArgs h = gemm_args(dZ, MLP_H, 1, n->wb + (l == 0 ? 0 : (long long)MLP_IN * MLP_H +
                                                      (long long)(l - 1) * MLP_H * MLP_H),
                                   1, MLP_H, rows, MLP_H, MLP_H);
            h.Cb = n->dz[1 - cur];
            h.mask = n->act[l - 1];
            gemm<EPI_MASK_BF16>(st, h);
            cur = 1 - cur;
        }
    }
}

rfd_status upload_rows(rfd_mlp *n, const float *features, int rows, std::vector<float> &staging,
                       float **dtmp) {
    rfd_ctx *ctx = n->ctx;
    RFD_HIP(ctx, hipMalloc(dtmp, sizeof(float) * (size_t)rows * MLP_FEAT));
    RFD_HIP(ctx, hipMemcpyAsync(*dtmp, features, sizeof(float) * (size_t)rows * MLP_FEAT,
                                hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(mlp_rows_to_bf16, dim3(blocks((long long)rows * MLP_IN)), dim3(256), 0,
                       ctx->stream, *dtmp, rows, n->x0);
    (void)staging;
    return RFD_OK;
}

}  // namespace

namespace rfd_host {

rfd_status mlp_reserve(rfd_mlp *n, int rows) {
    rfd_ctx *ctx = n->ctx;
    if (rows <= n->cap) return RFD_OK;
    int cap = std::max(rows, 256);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(n->x0);
    for (auto &a : n->act) (void)hipFree(a);
    for (auto &d : n->dz) (void)hipFree(d);
    (void)hipFree(n->p);
    (void)hipFree(n->dzh);
    (void)hipFree(n->loss);
    (void)hipFree(n->labels);
    RFD_HIP(ctx, hipMalloc(&n->x0, sizeof(uint16_t) * (size_t)cap * MLP_IN));
    for (auto &a : n->act) RFD_HIP(ctx, hipMalloc(&a, sizeof(uint16_t) * (size_t)cap * MLP_H));
    for (auto &d : n->dz) RFD_HIP(ctx, hipMalloc(&d, sizeof(uint16_t) * (size_t)cap * MLP_H));
    RFD_HIP(ctx, hipMalloc(&n->p, sizeof(float) * cap));
    RFD_HIP(ctx, hipMalloc(&n->dzh, sizeof(float) * cap));
    RFD_HIP(ctx, hipMalloc(&n->loss, sizeof(float) * cap));
    RFD_HIP(ctx, hipMalloc(&n->labels, sizeof(float) * cap));
    n->cap = cap;
    return RFD_OK;
}

void mlp_forward_dev(rfd_mlp *n, const uint16_t *x0, int rows, float *p_out) {
    hipStream_t st = n->ctx->stream;
    const uint16_t *X = x0;
    for (int l = 0; l < MLP_LAYERS; ++l) {
        const int K = l == 0 ? MLP_IN : MLP_H;
        const uint16_t *W = n->wb + (l == 0 ? 0 : (long long)MLP_IN * MLP_H + (long long)(l - 1) * MLP_H * MLP_H);
        GemmArgs g = gemm_args(X, K, 1, W, MLP_H, 1, rows, MLP_H, K);
        g.Cb = n->act[l];
        g.bias = n->params + mlp_b_off(l);
        gemm<EPI_BIAS_RELU_BF16>(st, g);
        X = n->act[l];
    }
    hipLaunchKernelGGL(mlp_head, dim3(blocks(rows)), dim3(256), 0, st, n->act[3], n->params, rows,
                       (const float *)nullptr, rows, p_out, (float *)nullptr, (float *)nullptr);
}

}  // namespace rfd_host

extern "C" {

int64_t rfd_mlp_param_count(void) { return MLP_NPARAMS; }

rfd_status rfd_mlp_destroy(rfd_mlp *n) {
    if (!n) return RFD_ERR_ARG;
    (void)hipStreamSynchronize(n->ctx->stream);
    for (void *p : {(void *)n->params, (void *)n->grads, (void *)n->m1, (void *)n->m2, (void *)n->wb,
                    (void *)n->x0, (void *)n->act[0], (void *)n->act[1], (void *)n->act[2],
                    (void *)n->act[3], (void *)n->dz[0], (void *)n->dz[1], (void *)n->p,
                    (void *)n->dzh, (void *)n->loss, (void *)n->labels})
        (void)hipFree(p);
    delete n;
    return RFD_OK;
}

rfd_status rfd_mlp_create(rfd_ctx *ctx, uint64_t seed, rfd_mlp **out) {
    if (!ctx || !out) return RFD_ERR_ARG;
    *out = nullptr;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    rfd_mlp *n = new rfd_mlp();
    n->ctx = ctx;
    // He initialisation for the ReLU layers, 1/sqrt(fan-in) for the head.
    std::vector<float> h((size_t)MLP_NPARAMS, 0.f);
    std::mt19937_64 g(seed);
    for (int l = 0; l < MLP_LAYERS; ++l) {
        const int K = l == 0 ? MLP_FEAT : MLP_H;  // padded inputs stay zero
        std::normal_distribution<float> d(0.f, std::sqrt(2.f / (float)K));
        const int rows = l == 0 ? MLP_IN : MLP_H;
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < MLP_H; ++c)
                h[(size_t)(mlp_w_off(l) + (long long)r * MLP_H + c)] = (l == 0 && r >= MLP_FEAT) ? 0.f : d(g);
    }
    std::normal_distribution<float> dh(0.f, 1.f / std::sqrt((float)MLP_H));
    for (int c = 0; c < MLP_H; ++c) h[(size_t)(MLP_W5 + c)] = dh(g);
    auto fail = [&](hipError_t e) {
        set_err(ctx, "MLP allocation failed: %s", hipGetErrorString(e));
        rfd_mlp_destroy(n);
        return RFD_ERR_NOMEM;
    };
    hipError_t e;
    if ((e = hipMalloc(&n->params, sizeof(float) * MLP_NPARAMS)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&n->grads, sizeof(float) * MLP_NPARAMS)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&n->m1, sizeof(float) * MLP_NPARAMS)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&n->m2, sizeof(float) * MLP_NPARAMS)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&n->wb, sizeof(uint16_t) * ((size_t)MLP_IN * MLP_H + 3 * (size_t)MLP_H * MLP_H))) != hipSuccess)
        return fail(e);
    *out = n;
    return rfd_mlp_set_params(n, h.data());
}

rfd_status rfd_mlp_get_params(rfd_mlp *n, float *out) {
    if (!n || !out) return RFD_ERR_ARG;
    RFD_HIP(n->ctx, hipMemcpyAsync(out, n->params, sizeof(float) * MLP_NPARAMS, hipMemcpyDeviceToHost,
                                   n->ctx->stream));
    RFD_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return RFD_OK;
}

rfd_status rfd_mlp_set_params(rfd_mlp *n, const float *in) {
    if (!n || !in) return RFD_ERR_ARG;
    rfd_ctx *ctx = n->ctx;
    hipStream_t st = ctx->stream;
    RFD_HIP(ctx, hipMemcpyAsync(n->params, in, sizeof(float) * MLP_NPARAMS, hipMemcpyHostToDevice, st));
    RFD_HIP(ctx, hipMemsetAsync(n->m1, 0, sizeof(float) * MLP_NPARAMS, st));
    RFD_HIP(ctx, hipMemsetAsync(n->m2, 0, sizeof(float) * MLP_NPARAMS, st));
    hipLaunchKernelGGL(mlp_cast_weights, dim3(blocks((long long)MLP_IN * MLP_H + 3LL * MLP_H * MLP_H)),
                       dim3(256), 0, st, n->params, n->wb);
    RFD_HIP(ctx, hipGetLastError());
    RFD_HIP(ctx, hipStreamSynchronize(st));
    n->step = 0;
    return RFD_OK;
}

rfd_status rfd_mlp_save(rfd_mlp *n, const char *path) {
    if (!n || !path) return RFD_ERR_ARG;
    std::vector<float> h((size_t)MLP_NPARAMS);
    rfd_status s = rfd_mlp_get_params(n, h.data());
    if (s != RFD_OK) return s;
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        set_err(n->ctx, "cannot write %s", path);
        return RFD_ERR_ARG;
    }
    const long long cnt = MLP_NPARAMS;
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&cnt, sizeof(cnt), 1, f) == 1 &&
              std::fwrite(&n->tau, sizeof(double), 1, f) == 1 &&
              std::fwrite(&n->delta, sizeof(double), 1, f) == 1 &&
              std::fwrite(h.data(), sizeof(float), h.size(), f) == h.size();
    std::fclose(f);
    if (!ok) {
        set_err(n->ctx, "short write to %s", path);
        return RFD_ERR_ARG;
    }
    return RFD_OK;
}

rfd_status rfd_mlp_load(rfd_mlp *n, const char *path) {
    if (!n || !path) return RFD_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_err(n->ctx, "cannot read %s", path);
        return RFD_ERR_ARG;
    }
    char magic[8];
    long long cnt = 0;
    double delta = 0, tau = 0;
    std::vector<float> h((size_t)MLP_NPARAMS);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(&cnt, sizeof(cnt), 1, f) == 1 && cnt == MLP_NPARAMS &&
              std::fread(&tau, sizeof(double), 1, f) == 1 &&
              std::fread(&delta, sizeof(double), 1, f) == 1 &&
              std::fread(h.data(), sizeof(float), h.size(), f) == h.size();
    std::fclose(f);
    if (!ok) {
        set_err(n->ctx, "%s is not a network file of this layout", path);
        return RFD_ERR_ARG;
    }
    rfd_status s = rfd_mlp_set_params(n, h.data());
    if (s == RFD_OK) {
        n->tau = tau;
        n->delta = delta;
    }
    return s;
}

rfd_status rfd_mlp_set_policy(rfd_mlp *n, double tau, double delta) {
    if (!n || !(tau >= 0) || tau > 0.5 || !(delta >= 0) || delta > 0.5) return RFD_ERR_ARG;
    n->tau = tau;
    n->delta = delta;
    return RFD_OK;
}

rfd_status rfd_mlp_get_policy(rfd_mlp *n, double *tau, double *delta) {
    if (!n || !tau || !delta) return RFD_ERR_ARG;
    *tau = n->tau;
    *delta = n->delta;
    return RFD_OK;
}

rfd_status rfd_mlp_forward(rfd_mlp *n, const float *features, int64_t rows, float *out) {
    if (!n || !features || !out || rows <= 0 || rows > (1 << 24)) return RFD_ERR_ARG;
    rfd_ctx *ctx = n->ctx;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    rfd_status s = rfd_host::mlp_reserve(n, (int)rows);
    if (s != RFD_OK) return s;
    std::vector<float> staging;
    float *tmp = nullptr;
    if ((s = upload_rows(n, features, (int)rows, staging, &tmp)) != RFD_OK) return s;
    rfd_host::mlp_forward_dev(n, n->x0, (int)rows, n->p);
    RFD_HIP(ctx, hipGetLastError());
    RFD_HIP(ctx, hipMemcpyAsync(out, n->p, sizeof(float) * rows, hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RFD_HIP(ctx, hipFree(tmp));
    return RFD_OK;
}

rfd_status rfd_mlp_gradients(rfd_mlp *n, const float *features, const float *labels, int64_t rows,
                             float *grads, float *loss) {
    if (!n || !features || !labels || rows <= 0 || rows > (1 << 24)) return RFD_ERR_ARG;
    rfd_ctx *ctx = n->ctx;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    rfd_status s = rfd_host::mlp_reserve(n, (int)rows);
    if (s != RFD_OK) return s;
    std::vector<float> staging;
    float *tmp = nullptr;
    if ((s = upload_rows(n, features, (int)rows, staging, &tmp)) != RFD_OK) return s;
    RFD_HIP(ctx, hipMemcpyAsync(n->labels, labels, sizeof(float) * rows, hipMemcpyHostToDevice, ctx->stream));
    rfd_host::mlp_forward_dev(n, n->x0, (int)rows, n->p);
    backward_dev(n, n->x0, (int)rows, n->labels);
    RFD_HIP(ctx, hipGetLastError());
    if (grads)
        RFD_HIP(ctx, hipMemcpyAsync(grads, n->grads, sizeof(float) * MLP_NPARAMS, hipMemcpyDeviceToHost,
                                    ctx->stream));
    std::vector<float> lr((size_t)rows);
    RFD_HIP(ctx, hipMemcpyAsync(lr.data(), n->loss, sizeof(float) * rows, hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RFD_HIP(ctx, hipFree(tmp));
    if (loss) *loss = std::accumulate(lr.begin(), lr.end(), 0.f);
    return RFD_OK;
}

rfd_status rfd_mlp_train(rfd_mlp *n, const float *features, const float *labels, int64_t rows,
                         int32_t epochs, int32_t batch, float lr, uint64_t seed, float *loss_per_epoch) {
    if (!n || !features || !labels || rows <= 0 || epochs <= 0 || batch <= 0 || rows > (1 << 24))
        return RFD_ERR_ARG;
    rfd_ctx *ctx = n->ctx;
    hipStream_t st = ctx->stream;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    const int B = (int)std::min<int64_t>(batch, rows);
    rfd_status s = rfd_host::mlp_reserve(n, B);
    if (s != RFD_OK) return s;
    // All records in BF16 on the device, reshuffled each epoch on the host.
    uint16_t *xall = nullptr;
    float *fall = nullptr, *yall = nullptr;
    RFD_HIP(ctx, hipMalloc(&xall, sizeof(uint16_t) * (size_t)rows * MLP_IN));
    RFD_HIP(ctx, hipMalloc(&fall, sizeof(float) * (size_t)rows * MLP_FEAT));
    RFD_HIP(ctx, hipMalloc(&yall, sizeof(float) * (size_t)rows));
    std::vector<int64_t> perm((size_t)rows);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 g(seed);
    std::vector<float> fx((size_t)rows * MLP_FEAT), fy((size_t)rows), lrows((size_t)B);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), g);
        for (int64_t i = 0; i < rows; ++i) {
            std::memcpy(&fx[(size_t)i * MLP_FEAT], features + perm[(size_t)i] * MLP_FEAT, sizeof(float) * MLP_FEAT);
            fy[(size_t)i] = labels[perm[(size_t)i]];
        }
        RFD_HIP(ctx, hipMemcpyAsync(fall, fx.data(), sizeof(float) * fx.size(), hipMemcpyHostToDevice, st));
        RFD_HIP(ctx, hipMemcpyAsync(yall, fy.data(), sizeof(float) * fy.size(), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(mlp_rows_to_bf16, dim3(blocks((long long)rows * MLP_IN)), dim3(256), 0, st, fall,
                           (int)rows, xall);
        double ep_loss = 0;
        int nb = 0;
        for (int64_t off = 0; off + B <= rows || (off == 0); off += B) {
            const int cur = (int)std::min<int64_t>(B, rows - off);
            const uint16_t *x0 = xall + off * MLP_IN;
            rfd_host::mlp_forward_dev(n, x0, cur, n->p);
            backward_dev(n, x0, cur, yall + off);
            n->step += 1;
            const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
            hipLaunchKernelGGL(mlp_adam, dim3(blocks(MLP_NPARAMS)), dim3(256), 0, st, n->params, n->grads,
                               n->m1, n->m2, (long long)MLP_NPARAMS, lr, b1, b2, eps, c1, c2);
            hipLaunchKernelGGL(mlp_cast_weights, dim3(blocks((long long)MLP_IN * MLP_H + 3LL * MLP_H * MLP_H)),
                               dim3(256), 0, st, n->params, n->wb);
            if (loss_per_epoch) {
                RFD_HIP(ctx, hipMemcpyAsync(lrows.data(), n->loss, sizeof(float) * cur, hipMemcpyDeviceToHost, st));
                RFD_HIP(ctx, hipStreamSynchronize(st));
                for (int i = 0; i < cur; ++i) ep_loss += lrows[(size_t)i];
                ++nb;
            }
            if (off + B >= rows) break;
        }
        RFD_HIP(ctx, hipGetLastError());
        if (loss_per_epoch) loss_per_epoch[ep] = nb ? (float)(ep_loss / nb) : 0.f;
    }
    RFD_HIP(ctx, hipStreamSynchronize(st));
    RFD_HIP(ctx, hipFree(xall));
    RFD_HIP(ctx, hipFree(fall));
    RFD_HIP(ctx, hipFree(yall));
    return RFD_OK;
}

}  // extern "C"
