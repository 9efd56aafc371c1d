// warmstart.hip -- features, targets, training loop and predictor of the
// learned hillslope warm start.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>

#include "warmstart.hpp"

namespace lsdtt_gpu {

namespace {

constexpr int kB = 256;

inline int nblk(long long n) { return static_cast<int>((n + kB - 1) / kB); }

// mean of the valid z in each window (one workgroup per window, FP64)
__global__ void __launch_bounds__(kB) k_win_mean(const WinDesc* __restrict__ desc, int win, int halo,
                                                 double* __restrict__ means)
{
    __shared__ double ss[kB], sc[kB];
    const WinDesc d = desc[blockIdx.x];
    double s = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < win * win; i += kB) {
        const int r = d.r0 - halo + i / win, q = d.c0 - halo + i % win;
        if (r < 0 || r >= d.nr || q < 0 || q >= d.nc) continue;
        const long long f = (long long)r * d.nc + q;
        if (!d.mask[f]) continue;
        s += d.z[f];
        c += 1.0;
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            ss[threadIdx.x] += ss[threadIdx.x + o];
            sc[threadIdx.x] += sc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) means[blockIdx.x] = sc[0] > 0.0 ? ss[0] / sc[0] : 0.0;
}

__device__ __forceinline__ bool valid_at(const WinDesc& d, int r, int c)
{
    return r >= 0 && r < d.nr && c >= 0 && c < d.nc && d.mask[(long long)r * d.nc + c];
}

// features [N][win][win][4]
__global__ void k_features(const WinDesc* __restrict__ desc, int N, int win, int halo, const double* __restrict__ means, float z_scale, float lap_scale, float kappa_scale,
                           float* __restrict__ x)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long per = (long long)win * win;
    if (i >= N * per) return;
    const int n = static_cast<int>(i / per);
    const int wi = static_cast<int>(i - n * per);
    const WinDesc d = desc[n];
    const int r = d.r0 - halo + wi / win, c = d.c0 - halo + wi % win;
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid_at(d, r, c)) {
        const long long k = (long long)r * d.nc + c;
        const double z = d.z[k];
        const bool vn = valid_at(d, r - 1, c), vs = valid_at(d, r + 1, c);
        const bool vw = valid_at(d, r, c - 1), ve = valid_at(d, r, c + 1);
        const double zn = vn ? d.z[k - d.nc] : z, zs = vs ? d.z[k + d.nc] : z;
        const double zw = vw ? d.z[k - 1] : z, ze = ve ? d.z[k + 1] : z;
        const double gx = (ze - zw) / (((ve && vw) ? 2.0 : 1.0) * d.h);
        const double gy = (zs - zn) / (((vn && vs) ? 2.0 : 1.0) * d.h);
        const double lap = (zn - z) + (zs - z) + (zw - z) + (ze - z);
        f.x = static_cast<float>((z - means[n]) / z_scale);
        f.y = static_cast<float>(std::sqrt(gx * gx + gy * gy) / d.Sc);
        f.z = static_cast<float>(lap / lap_scale);
        f.w = d.kappa * kappa_scale;
    }
    reinterpret_cast<float4*>(x)[i] = f;
}

__global__ void k_targets(const WinDesc* __restrict__ desc, int N, int win, int halo, int tile, float lap_scale,
                          float* __restrict__ target, uint8_t* __restrict__ lmask)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long per = (long long)win * win;
    if (i >= N * per) return;
    const int n = static_cast<int>(i / per);
    const int wi = static_cast<int>(i - n * per);
    const WinDesc d = desc[n];
    const int wy = wi / win, wx = wi % win;
    const int r = d.r0 - halo + wy, c = d.c0 - halo + wx;
    const bool interior = wy >= halo && wy < halo + tile && wx >= halo && wx < halo + tile;
    float t = 0.f;
    uint8_t m = 0;
    if (interior && r >= 0 && r < d.nr && c >= 0 && c < d.nc) {
        const long long k = (long long)r * d.nc + c;
        if (d.dof[k]) {
            m = 1;
            t = d.dz[k] / (d.kappa * lap_scale);
        }
    }
    target[i] = t;
    lmask[i] = m;
}

__global__ void k_scatter(const WinDesc* __restrict__ desc, int N, int win, int halo, int tile, float lap_scale,
                          const float* __restrict__ y, float* __restrict__ dz)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long per = (long long)tile * tile;
    if (i >= N * per) return;
    const int n = static_cast<int>(i / per);
    const int ti = static_cast<int>(i - n * per);
    const WinDesc d = desc[n];
    const int ty = ti / tile, tx = ti % tile;
    const int r = d.r0 + ty, c = d.c0 + tx;
    if (r >= d.nr || c >= d.nc) return;
    const long long k = (long long)r * d.nc + c;
    dz[k] = d.dof[k] ? y[(long long)n * win * win + (long long)(ty + halo) * win + (tx + halo)] * d.kappa * lap_scale
                     : 0.f;
}

__global__ void k_fill0(float* __restrict__ a, long long n)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.f;
}

__global__ void k_d2f_diff(const double* __restrict__ a, const double* __restrict__ b, long long n,
                           float* __restrict__ out)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(a[i] - b[i]);
}

}  // namespace

void ws_features(const WinDesc* d_desc, int N, const FeatureSpec& fs, double* d_means, float* x, hipStream_t s)
{
    const int win = fs.window();
    k_win_mean<<<N, kB, 0, s>>>(d_desc, win, fs.halo, d_means);
    k_features<<<nblk((long long)N * win * win), kB, 0, s>>>(d_desc, N, win, fs.halo, d_means, (float)fs.z_scale,
                                                            (float)fs.lap_scale, (float)fs.kappa_scale, x);
    LSDTT_LAUNCH_CHECK();
}

void ws_targets(const WinDesc* d_desc, int N, const FeatureSpec& fs, float* target, uint8_t* lmask, hipStream_t s)
{
    const int win = fs.window();
    k_targets<<<nblk((long long)N * win * win), kB, 0, s>>>(d_desc, N, win, fs.halo, fs.tile, (float)fs.lap_scale,
                                                           target, lmask);
    LSDTT_LAUNCH_CHECK();
}

void ws_scatter(const WinDesc* d_desc, int N, const FeatureSpec& fs, const float* y, float* dz, hipStream_t s)
{
    k_scatter<<<nblk((long long)N * fs.tile * fs.tile), kB, 0, s>>>(d_desc, N, fs.window(), fs.halo, fs.tile,
                                                                    (float)fs.lap_scale, y, dz);
    LSDTT_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
CnnDataset::CnnDataset(int nr_, int nc_, double dx_, const uint8_t* d_mask, const uint8_t* d_dof)
    : nr(nr_), nc(nc_), dx(dx_)
{
    const size_t n = (size_t)nr * nc;
    mask.reserve(n);
    dof.reserve(n);
    LSDTT_HIP(hipMemcpy(mask.p, d_mask, n, hipMemcpyDeviceToDevice));
    LSDTT_HIP(hipMemcpy(dof.p, d_dof, n, hipMemcpyDeviceToDevice));
}

void CnnDataset::add(const double* d_z, const float* d_dz, double kappa, double Sc, hipStream_t s)
{
    const size_t n = (size_t)nr * nc;
    auto sn = std::make_unique<Snap>();
    sn->z.reserve(n);
    sn->dz.reserve(n);
    LSDTT_HIP(hipMemcpyAsync(sn->z.p, d_z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    LSDTT_HIP(hipMemcpyAsync(sn->dz.p, d_dz, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    sn->kappa = kappa;
    sn->Sc = Sc;
    snaps.push_back(std::move(sn));
}

void CnnTrainer::make_batch(CnnDataset& ds, int batch, std::mt19937_64& rng, hipStream_t s)
{
    const int win = fs.window();
    const size_t per = (size_t)win * win;
    desc.reserve(batch);
    hdesc.reserve(batch);
    means.reserve(batch);
    x.reserve(per * batch * 4);
    target.reserve(per * batch);
    lmask.reserve(per * batch);
    // the previous batch's descriptors may still be in flight
    LSDTT_HIP(hipStreamSynchronize(s));
    std::uniform_int_distribution<int> us(0, ds.size() - 1);
    std::uniform_int_distribution<int> ur(0, std::max(0, ds.nr - fs.tile)), uc(0, std::max(0, ds.nc - fs.tile));
    for (int i = 0; i < batch; ++i) {
        const auto& sn = *ds.snaps[us(rng)];
        WinDesc d;
        d.z = sn.z.p;
        d.dz = sn.dz.p;
        d.mask = ds.mask.p;
        d.dof = ds.dof.p;
        d.h = ds.dx;
        d.nr = ds.nr;
        d.nc = ds.nc;
        d.r0 = ur(rng);
        d.c0 = uc(rng);
        d.kappa = static_cast<float>(sn.kappa);
        d.Sc = static_cast<float>(sn.Sc);
        hdesc.p[i] = d;
    }
    LSDTT_HIP(hipMemcpyAsync(desc.p, hdesc.p, sizeof(WinDesc) * batch, hipMemcpyHostToDevice, s));
    ws_features(desc.p, batch, fs, means.p, x.p, s);
    ws_targets(desc.p, batch, fs, target.p, lmask.p, s);
}

double CnnTrainer::train(CnnDataset& ds, int iters, int batch, double lr, uint64_t seed, float* losses, hipStream_t s)
{
    if (ds.size() == 0) throw std::invalid_argument("CnnTrainer: empty dataset");
    const int win = fs.window();
    if (!net.shape_ok(win, win)) throw std::invalid_argument("CnnTrainer: window size does not fit the network depth");
    std::mt19937_64 rng(seed);
    double last = 0.0;
    for (int it = 0; it < iters; ++it) {
        make_batch(ds, batch, rng, s);
        net.forward(x.p, batch, win, win, nullptr, s);
        net.zero_grad(s);
        last = net.backward(target.p, lmask.p, s);
        net.adam_step(lr, 0.9, 0.999, 1e-8, s);
        if (losses) losses[it] = static_cast<float>(last);
    }
    LSDTT_HIP(hipStreamSynchronize(s));
    return last;
}

void CnnTrainer::evaluate(CnnDataset& ds, int n_batches, int batch, uint64_t seed, double* loss, double* zero_loss,
                          hipStream_t s)
{
    if (ds.size() == 0) throw std::invalid_argument("CnnTrainer: empty dataset");
    const int win = fs.window();
    std::mt19937_64 rng(seed);
    double a = 0.0, b = 0.0;
    for (int it = 0; it < n_batches; ++it) {
        make_batch(ds, batch, rng, s);
        net.forward(x.p, batch, win, win, nullptr, s);
        // loss of the network; the parameter gradient it also forms is
        // discarded by the next zero_grad
        a += net.backward(target.p, lmask.p, s);
        // zero prediction: MSE = mean of target^2 over the mask
        std::vector<float> ht((size_t)batch * win * win);
        std::vector<uint8_t> hm(ht.size());
        LSDTT_HIP(hipMemcpy(ht.data(), target.p, sizeof(float) * ht.size(), hipMemcpyDeviceToHost));
        LSDTT_HIP(hipMemcpy(hm.data(), lmask.p, hm.size(), hipMemcpyDeviceToHost));
        double s2 = 0.0, c = 0.0;
        for (size_t i = 0; i < ht.size(); ++i)
            if (hm[i]) {
                s2 += (double)ht[i] * ht[i];
                c += 1.0;
            }
        b += c > 0.0 ? s2 / c : 0.0;
    }
    net.zero_grad(s);
    LSDTT_HIP(hipStreamSynchronize(s));
    *loss = a / n_batches;
    *zero_loss = b / n_batches;
}

// ---------------------------------------------------------------------------
void WarmStartPredictor::predict(const double* z, const uint8_t* mask, const uint8_t* dof, int nr, int nc, double h,
                                 double kappa, double Sc, float* dz, hipStream_t s)
{
    const int win = fs.window();
    const int tr = (nr + fs.tile - 1) / fs.tile, tc = (nc + fs.tile - 1) / fs.tile;
    const int N = tr * tc;
    if (hdesc.size() != (size_t)N) {
        hdesc.resize(N);
        desc.reserve(N);
    }
    for (int i = 0; i < tr; ++i)
        for (int j = 0; j < tc; ++j) {
            WinDesc d;
            d.z = z;
            d.dz = nullptr;
            d.mask = mask;
            d.dof = dof;
            d.h = h;
            d.nr = nr;
            d.nc = nc;
            d.r0 = i * fs.tile;
            d.c0 = j * fs.tile;
            d.kappa = static_cast<float>(kappa);
            d.Sc = static_cast<float>(Sc);
            hdesc[(size_t)i * tc + j] = d;
        }
    means.reserve(N);
    x.reserve((size_t)N * win * win * 4);
    // descriptors are plain data: a synchronous upload keeps the host vector
    // free to change on the next call
    LSDTT_HIP(hipMemcpyAsync(desc.p, hdesc.data(), sizeof(WinDesc) * N, hipMemcpyHostToDevice, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    ws_features(desc.p, N, fs, means.p, x.p, s);
    net.forward(x.p, N, win, win, nullptr, s);
    ws_scatter(desc.p, N, fs, net.out_device(), dz, s);
}

}  // namespace lsdtt_gpu
