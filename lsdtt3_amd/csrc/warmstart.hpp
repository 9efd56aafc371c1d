// warmstart.hpp -- inputs, targets and training for the learned hillslope
// warm start, and the predictor the LEM calls before its Newton solve.
//
// No counterpart in the reference.  The network sees square windows of the
// post-incision surface: a tile of side `tile` plus `halo` cells on each
// side.  Per pixel it gets four features:
//   0: (z - mean of the window's valid z) / z_scale
//   1: |grad z| / Sc            (central differences, one-sided at edges)
//   2: (sum of valid 4-neighbours' z - z) / lap_scale   (h^2 Laplacian, m)
//   3: kappa * kappa_scale, kappa = dt D / h^2          (per window)
// and predicts dz / (kappa lap_scale): the hillslope step's elevation change
// in the units of the discrete Laplacian (for linear diffusion the explicit
// estimate of that quantity is feature 2 itself).  Pixels outside the grid
// or invalid get zero features.  Only the tile's interior DOF pixels carry
// loss and are scattered back into the solver's start.
#pragma once

#include <cstdint>
#include <memory>
#include <random>
#include <vector>

#include "cnn.hpp"
#include "common.hpp"
#include "dbuf.hpp"

namespace lsdtt_gpu {

struct FeatureSpec {
    int tile = 256, halo = 16;
    double z_scale = 100.0, lap_scale = 1.0, kappa_scale = 100.0;
    int window() const { return tile + 2 * halo; }
};

// one window: the raster it reads and the interior tile's origin
struct WinDesc {
    const double* z;   // FP64 surface (n cells)
    const float* dz;   // target elevation change (training) or null
    const uint8_t* mask;
    const uint8_t* dof;
    double h;          // cell size
    int nr, nc, r0, c0;
    float kappa, Sc;
};

// device helpers (stream-ordered)
// x: FP32 [N][window][window][4]
void ws_features(const WinDesc* d_desc, int N, const FeatureSpec& fs, double* d_means, float* x, hipStream_t s);
void ws_targets(const WinDesc* d_desc, int N, const FeatureSpec& fs, float* target, uint8_t* lmask, hipStream_t s);
// dz[cell] = y * kappa * lap_scale on interior DOF pixels (others untouched)
void ws_scatter(const WinDesc* d_desc, int N, const FeatureSpec& fs, const float* y, float* dz, hipStream_t s);

// training pairs captured from LEM steps (post-incision surface, hillslope
// elevation change), all on one grid
class CnnDataset {
public:
    CnnDataset(int nr, int nc, double dx, const uint8_t* d_mask, const uint8_t* d_dof);
    void add(const double* d_z, const float* d_dz, double kappa, double Sc, hipStream_t s);
    int size() const { return static_cast<int>(snaps.size()); }
    int nr, nc;
    double dx;
    struct Snap {
        dbuf<double> z;
        dbuf<float> dz;
        double kappa, Sc;
    };
    std::vector<std::unique_ptr<Snap>> snaps;
    dbuf<uint8_t> mask, dof;
};

class CnnTrainer {
public:
    CnnTrainer(UNet& net, const FeatureSpec& fs) : net(net), fs(fs) {}
    // `iters` Adam steps on random windows (uniform snapshot, uniform tile
    // origin inside the grid); losses[i] (host, may be null) the batch MSE
    double train(CnnDataset& ds, int iters, int batch, double lr, uint64_t seed, float* losses, hipStream_t s);
    // mean MSE of the network and of the zero prediction over n_batches
    // random batches (no parameter change)
    void evaluate(CnnDataset& ds, int n_batches, int batch, uint64_t seed, double* loss, double* zero_loss,
                  hipStream_t s);

private:
    UNet& net;
    FeatureSpec fs;
    dbuf<WinDesc> desc;
    hbuf<WinDesc> hdesc;
    dbuf<double> means;
    dbuf<float> x, target;
    dbuf<uint8_t> lmask;
    dbuf<float> zero_out;
    void make_batch(CnnDataset& ds, int batch, std::mt19937_64& rng, hipStream_t s);
};

// the LEM's predictor: tiles the grid, runs the network, scatters dz
class WarmStartPredictor {
public:
    WarmStartPredictor(UNet& net, const FeatureSpec& fs) : net(net), fs(fs) {}
    // dz (FP32, n cells) = the predicted hillslope change of z over a step
    // of dt (kappa = dt D / h^2) on the DOFs, 0 elsewhere
    void predict(const double* z, const uint8_t* mask, const uint8_t* dof, int nr, int nc, double h, double kappa,
                 double Sc, float* dz, hipStream_t s);
    UNet& net;
    FeatureSpec fs;

private:
    dbuf<WinDesc> desc;
    std::vector<WinDesc> hdesc;
    dbuf<double> means;
    dbuf<float> x, y;
};

}  // namespace lsdtt_gpu
