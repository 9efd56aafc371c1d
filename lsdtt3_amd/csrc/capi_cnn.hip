// capi_cnn.hip -- C ABI of the learned hillslope warm start (lsdtt3_gpu.h):
// the U-Net, its training set captured from LEM steps, training, and the
// LEM hook.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <chrono>
#include <memory>
#include <random>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
#include "capi_lem.hpp"
#include "capi_util.hpp"
#include "cnn.hpp"
#include "warmstart.hpp"

using namespace lsdtt_gpu;

struct lsdtt_gpu_cnn {
    std::unique_ptr<UNet> net;
    FeatureSpec fs;
    std::unique_ptr<CnnTrainer> trainer;
    std::unique_ptr<WarmStartPredictor> predictor;
    dbuf<float> x, y, t;
    dbuf<uint8_t> m;
};

struct lsdtt_gpu_cnn_dataset {
    std::unique_ptr<CnnDataset> ds;
};

extern "C" {

lsdtt_gpu_cnn* lsdtt_gpu_cnn_create(int32_t in_channels, int32_t base_channels, int32_t levels, uint64_t seed,
                                    int32_t tile, int32_t halo)
{
    return guarded_ptr<lsdtt_gpu_cnn>([&] {
        if (tile < 1 || halo < 0) throw std::invalid_argument("tile must be >= 1 and halo >= 0");
        auto c = std::make_unique<lsdtt_gpu_cnn>();
        c->net = std::make_unique<UNet>(in_channels, base_channels, levels, seed);
        c->fs.tile = tile;
        c->fs.halo = halo;
        if (!c->net->shape_ok(c->fs.window(), c->fs.window()))
            throw std::invalid_argument("tile + 2 halo must be divisible by 2^(levels-1)");
        c->trainer = std::make_unique<CnnTrainer>(*c->net, c->fs);
        c->predictor = std::make_unique<WarmStartPredictor>(*c->net, c->fs);
        return c.release();
    });
}

void lsdtt_gpu_cnn_destroy(lsdtt_gpu_cnn* net)
{
    if (!net) return;
    (void)hipDeviceSynchronize();
    delete net;
}

int64_t lsdtt_gpu_cnn_num_params(const lsdtt_gpu_cnn* net) { return net ? (int64_t)net->net->num_params() : 0; }

int lsdtt_gpu_cnn_get_params(const lsdtt_gpu_cnn* net, float* params)
{
    return guarded([&] {
        if (!net || !params) return fail_arg("null argument");
        net->net->get_params(params);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_set_params(lsdtt_gpu_cnn* net, const float* params)
{
    return guarded([&] {
        if (!net || !params) return fail_arg("null argument");
        net->net->set_params(params);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_get_grads(const lsdtt_gpu_cnn* net, float* grads)
{
    return guarded([&] {
        if (!net || !grads) return fail_arg("null argument");
        LSDTT_HIP(hipDeviceSynchronize());
        net->net->get_grads(grads);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_forward(lsdtt_gpu_cnn* net, const float* x, int32_t n, int32_t h, int32_t w, float* y)
{
    return guarded([&] {
        if (!net || !x || !y) return fail_arg("null argument");
        if (n < 1 || !net->net->shape_ok(h, w)) return fail_arg("shape not divisible by the network's depth");
        const size_t P = (size_t)n * h * w, C = (size_t)net->net->in_channels();
        net->x.reserve(P * C);
        LSDTT_HIP(hipMemcpy(net->x.p, x, sizeof(float) * P * C, hipMemcpyHostToDevice));
        net->net->forward(net->x.p, n, h, w, nullptr, nullptr);
        LSDTT_HIP(hipMemcpy(y, net->net->out_device(), sizeof(float) * P, hipMemcpyDeviceToHost));
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_backward(lsdtt_gpu_cnn* net, const float* target, const uint8_t* mask, double* loss)
{
    return guarded([&] {
        if (!net || !target || !loss) return fail_arg("null argument");
        (void)0;
        const size_t P = net->x.n / (size_t)net->net->in_channels();
        if (P == 0) return fail_arg("backward before forward");
        net->t.reserve(P);
        LSDTT_HIP(hipMemcpy(net->t.p, target, sizeof(float) * P, hipMemcpyHostToDevice));
        const uint8_t* dm = nullptr;
        if (mask) {
            net->m.reserve(P);
            LSDTT_HIP(hipMemcpy(net->m.p, mask, P, hipMemcpyHostToDevice));
            dm = net->m.p;
        }
        *loss = net->net->backward(net->t.p, dm, nullptr);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_zero_grad(lsdtt_gpu_cnn* net)
{
    return guarded([&] {
        if (!net) return fail_arg("null argument");
        net->net->zero_grad(nullptr);
        LSDTT_HIP(hipDeviceSynchronize());
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_adam_step(lsdtt_gpu_cnn* net, double lr, double beta1, double beta2, double eps)
{
    return guarded([&] {
        if (!net) return fail_arg("null argument");
        if (!(lr > 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) || !(eps > 0.0))
            return fail_arg("Adam needs lr > 0, 0 <= beta < 1, eps > 0");
        net->net->adam_step(lr, beta1, beta2, eps, nullptr);
        LSDTT_HIP(hipDeviceSynchronize());
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_benchmark(lsdtt_gpu_cnn* net, int32_t n, int32_t h, int32_t w, int32_t reps, double* forward_ms,
                            double* conv_ms, double* train_ms, double* flops)
{
    return guarded([&] {
        if (!net || !forward_ms || !conv_ms || !train_ms || !flops) return fail_arg("null argument");
        if (n < 1 || reps < 1 || !net->net->shape_ok(h, w)) return fail_arg("bad shape or repetition count");
        UNet& u = *net->net;
        const size_t P = (size_t)n * h * w, C = (size_t)u.in_channels();
        std::vector<float> hx(P * C), ht(P);
        std::mt19937_64 rng(12345);
        std::normal_distribution<float> nd(0.f, 1.f);
        for (auto& v : hx) v = nd(rng);
        for (auto& v : ht) v = nd(rng);
        dbuf<float> dx, dt;
        dx.reserve(P * C);
        dt.reserve(P);
        LSDTT_HIP(hipMemcpy(dx.p, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice));
        LSDTT_HIP(hipMemcpy(dt.p, ht.data(), sizeof(float) * ht.size(), hipMemcpyHostToDevice));
        std::vector<float> saved(u.num_params());
        u.get_params(saved.data());
        hipEvent_t e0, e1;
        LSDTT_HIP(hipEventCreate(&e0));
        LSDTT_HIP(hipEventCreate(&e1));
        u.forward(dx.p, n, h, w, nullptr, nullptr);  // warm-up (allocations)
        LSDTT_HIP(hipDeviceSynchronize());
        double cms = 0.0;
        u.time_convs = true;
        for (int r = 0; r < reps; ++r) {
            u.forward(dx.p, n, h, w, nullptr, nullptr);
            cms += u.conv_ms;
        }
        u.time_convs = false;
        LSDTT_HIP(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; ++r) u.forward(dx.p, n, h, w, nullptr, nullptr);
        LSDTT_HIP(hipEventRecord(e1, nullptr));
        LSDTT_HIP(hipEventSynchronize(e1));
        float fms = 0.f;
        LSDTT_HIP(hipEventElapsedTime(&fms, e0, e1));
        u.zero_grad(nullptr);
        (void)u.backward(dt.p, nullptr, nullptr);  // warm-up
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r) {
            u.forward(dx.p, n, h, w, nullptr, nullptr);
            u.zero_grad(nullptr);
            (void)u.backward(dt.p, nullptr, nullptr);
            u.adam_step(1e-6, 0.9, 0.999, 1e-8, nullptr);
        }
        LSDTT_HIP(hipDeviceSynchronize());
        const double tms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        // restore the parameters (the Adam moments keep the tiny steps)
        u.set_params(saved.data());
        u.zero_grad(nullptr);
        LSDTT_HIP(hipDeviceSynchronize());
        *forward_ms = fms / reps;
        *conv_ms = cms / reps;
        *train_ms = tms / reps;
        *flops = u.conv_flops(n, h, w);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_conv3x3(const float* x, const float* w, const float* b, int32_t n, int32_t h, int32_t w_px,
                          int32_t cin, int32_t cout, int32_t relu, float* y, const float* gy, float* gx, float* gw,
                          float* gb)
{
    return guarded([&] {
        if (!x || !w || !b || !y || (gy && (!gw || !gb))) return fail_arg("null argument");
        conv3x3_layer(x, w, b, n, h, w_px, cin, cout, relu, y, gy, gx, gw, gb);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_capture_hillslope(lsdtt_gpu_lem* lem, int32_t enable)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        lem->eng->capture = enable != 0;
        lem->eng->cap_valid = false;
        return LSDTT_GPU_OK;
    });
}

lsdtt_gpu_cnn_dataset* lsdtt_gpu_cnn_dataset_create(void)
{
    return guarded_ptr<lsdtt_gpu_cnn_dataset>([&] { return new lsdtt_gpu_cnn_dataset(); });
}

void lsdtt_gpu_cnn_dataset_destroy(lsdtt_gpu_cnn_dataset* ds)
{
    if (!ds) return;
    (void)hipDeviceSynchronize();
    delete ds;
}

int lsdtt_gpu_cnn_dataset_add(lsdtt_gpu_cnn_dataset* ds, lsdtt_gpu_lem* lem)
{
    return guarded([&] {
        if (!ds || !lem) return fail_arg("null argument");
        LemEngine& e = *lem->eng;
        if (!e.capture || !e.cap_valid) return fail_arg("no captured hillslope step (enable capture, then step)");
        if (!ds->ds) ds->ds = std::make_unique<CnnDataset>(e.nr, e.nc, e.dx, e.mask.p, e.dof.p);
        if (ds->ds->nr != e.nr || ds->ds->nc != e.nc) return fail_arg("dataset grid differs from the LEM's");
        ds->ds->add(e.z_hs0.p, e.cap_dz.p, e.cap_kappa, e.p.Sc, nullptr);
        LSDTT_HIP(hipDeviceSynchronize());
        return LSDTT_GPU_OK;
    });
}

int32_t lsdtt_gpu_cnn_dataset_size(const lsdtt_gpu_cnn_dataset* ds)
{
    return (ds && ds->ds) ? ds->ds->size() : 0;
}

int lsdtt_gpu_cnn_train(lsdtt_gpu_cnn* net, lsdtt_gpu_cnn_dataset* ds, int32_t n_iterations, int32_t batch, double lr,
                        uint64_t seed, float* losses)
{
    return guarded([&] {
        if (!net || !ds) return fail_arg("null argument");
        if (!ds->ds || ds->ds->size() == 0) return fail_arg("empty dataset");
        if (n_iterations < 0 || batch < 1 || !(lr > 0.0)) return fail_arg("need n_iterations >= 0, batch >= 1, lr > 0");
        net->trainer->train(*ds->ds, n_iterations, batch, lr, seed, losses, nullptr);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_cnn_evaluate(lsdtt_gpu_cnn* net, lsdtt_gpu_cnn_dataset* ds, int32_t n_batches, int32_t batch,
                           uint64_t seed, double* loss, double* zero_loss)
{
    return guarded([&] {
        if (!net || !ds || !loss || !zero_loss) return fail_arg("null argument");
        if (!ds->ds || ds->ds->size() == 0) return fail_arg("empty dataset");
        if (n_batches < 1 || batch < 1) return fail_arg("need n_batches >= 1, batch >= 1");
        net->trainer->evaluate(*ds->ds, n_batches, batch, seed, loss, zero_loss, nullptr);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_lem_set_warm_start(lsdtt_gpu_lem* lem, lsdtt_gpu_cnn* net, int32_t mode)
{
    return guarded([&] {
        if (!lem) return fail_arg("null argument");
        if (mode < 0 || mode > 3) return fail_arg("warm start mode must be 0, 1, 2 or 3");
        if (mode == 2 && !net) return fail_arg("the learned warm start needs a network");
        if (mode == 2 && net->net->in_channels() != 4) return fail_arg("the warm-start network takes 4 input channels");
        lem->eng->ws_mode = mode;
        lem->eng->ws_dz_valid = false;
        lem->eng->ws = mode == 2 ? net->predictor.get() : nullptr;
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
