// capi_flow.hip -- C ABI for the flow object and the fill (lsdtt3_gpu.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
#include "capi_util.hpp"
#include "chi.hpp"
#include "capi_steady.hpp"
#include "flow.hpp"

namespace lsdtt_gpu {
size_t fill_scratch_ints(int nr, int nc);
int fill_device(const float* z, const uint8_t* mask, int nr, int nc, double nodata, BC bc, float* out,
                int* scratch, hipStream_t s);
}

using namespace lsdtt_gpu;

struct lsdtt_gpu_flow {
    lsdtt_gpu_flow_options o;
    std::unique_ptr<FlowEngine> eng;
    dbuf<float> z32;
    dbuf<double> z64;
    dbuf<uint8_t> mask, forced;
    dbuf<double> wbuf, abuf;
    dbuf<float> chi;
    ChiEngine chi_eng;
    bool built = false;
    bool has_topo = false, has_sv = false;
};

namespace {

bool parse_bc(const char* s, BC& bc)
{
    for (int i = 0; i < 4; ++i)
        if (s[i] != 'o' && s[i] != 'n' && s[i] != 'p' && s[i] != 'b') return false;
    bc = BC{s[0], s[1], s[2], s[3]};
    return true;
}

__global__ void k_widen(const int* __restrict__ a, long long n, long long* __restrict__ out)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i];
}

// download an int32 device array widened to int64 on the host
void get_i64(const int* d, long long n, int64_t* h, hipStream_t s)
{
    if (!h || n <= 0) return;
    std::vector<int> tmp(static_cast<size_t>(n));
    LSDTT_HIP(hipMemcpyAsync(tmp.data(), d, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    LSDTT_HIP(hipStreamSynchronize(s));
    for (long long i = 0; i < n; ++i) h[i] = tmp[static_cast<size_t>(i)];
}

int finish_build(lsdtt_gpu_flow* f, long long rc)
{
    if (rc < 0) {
        const long long flat = -rc - 1;
        set_error("lsdtt_flowinfo: could not assign a receiver at row " + std::to_string(flat / f->o.n_cols) +
                  ", col " + std::to_string(flat % f->o.n_cols) +
                  ". The DEM likely contains an unresolved sink or flat.");
        return LSDTT_GPU_ERR_UNRESOLVED;
    }
    hipStream_t s = nullptr;
    const bool topo = f->o.build_topological_order != 0;
    const bool sv = f->o.build_s_vector != 0;
    f->eng->upstream(true, nullptr, nullptr, s, sv);
    if (topo) f->eng->topological_order(s);
    if (sv) f->eng->s_vector(s);
    LSDTT_HIP(hipStreamSynchronize(s));
    f->built = true;
    f->has_topo = topo;
    f->has_sv = sv;
    return LSDTT_GPU_OK;
}

template <class T>
int build_host(lsdtt_gpu_flow* f, const T* z, const uint8_t* mask, const int64_t* forced, int64_t n_forced,
               dbuf<T>& zb)
{
    if (!f || !z || !mask) return fail_arg("null argument");
    const long long n = (long long)f->o.n_rows * f->o.n_cols;
    std::vector<uint8_t> plane;
    if (forced && n_forced > 0) {
        plane.assign(static_cast<size_t>(n), 0);
        for (int64_t i = 0; i < n_forced; ++i) {
            const int64_t k = forced[i];
            if (k < 0 || k >= n)
                return fail_arg("lsdtt_flowinfo: forced outlet flat index " + std::to_string(k) +
                                " is out of range");
            if (!mask[k])
                return fail_arg("lsdtt_flowinfo: forced outlet flat index " + std::to_string(k) +
                                " lands on a nodata or masked cell");
            plane[static_cast<size_t>(k)] = 1;
        }
    }
    hipStream_t s = nullptr;
    zb.reserve(n);
    f->mask.reserve(n);
    LSDTT_HIP(hipMemcpyAsync(zb.p, z, sizeof(T) * n, hipMemcpyHostToDevice, s));
    LSDTT_HIP(hipMemcpyAsync(f->mask.p, mask, n, hipMemcpyHostToDevice, s));
    const uint8_t* fp = nullptr;
    if (!plane.empty()) {
        f->forced.reserve(n);
        LSDTT_HIP(hipMemcpyAsync(f->forced.p, plane.data(), n, hipMemcpyHostToDevice, s));
        fp = f->forced.p;
    }
    f->built = false;
    const long long rc = f->eng->route<T>(zb.p, f->mask.p, fp, s);
    return finish_build(f, rc);
}

}  // namespace

extern "C" {

int lsdtt_gpu_available(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

const char* lsdtt_gpu_last_error(void) { return last_error_cstr(); }

void* lsdtt_gpu_device_alloc(int64_t bytes)
{
    return guarded_ptr<void>([&]() -> void* {
        if (bytes <= 0) throw std::invalid_argument("allocation size must be positive");
        void* p = nullptr;
        LSDTT_HIP(hipMalloc(&p, static_cast<size_t>(bytes)));
        return p;
    });
}

void lsdtt_gpu_device_free(void* p)
{
    if (p) (void)hipFree(p);
}

int lsdtt_gpu_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind)
{
    return guarded([&] {
        if ((!dst || !src) && bytes > 0) return fail_arg("null argument");
        if (bytes < 0 || kind < 1 || kind > 3) return fail_arg("bytes >= 0 and kind 1, 2 or 3");
        const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                          : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
        if (bytes) LSDTT_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), k));
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_synchronize(void)
{
    return guarded([&] {
        LSDTT_HIP(hipDeviceSynchronize());
        return LSDTT_GPU_OK;
    });
}

void lsdtt_gpu_flow_options_init(lsdtt_gpu_flow_options* o, int32_t n_rows, int32_t n_cols, double cell_size)
{
    if (!o) return;
    o->n_rows = n_rows;
    o->n_cols = n_cols;
    o->cell_size = cell_size;
    std::memcpy(o->boundary_conditions, "oooo", 4);
    o->interior_nodata_is_outlet = 1;
    o->allow_pits_as_outlets = 0;
    o->build_topological_order = 1;
    o->build_s_vector = 1;
}

lsdtt_gpu_flow* lsdtt_gpu_flow_create(const lsdtt_gpu_flow_options* o)
{
    return guarded_ptr<lsdtt_gpu_flow>([&]() -> lsdtt_gpu_flow* {
        if (!o) throw std::invalid_argument("null options");
        BC bc;
        if (!parse_bc(o->boundary_conditions, bc))
            throw std::invalid_argument("boundary conditions must be four of o/n/p/b");
        if (!(o->cell_size > 0.0) || !std::isfinite(o->cell_size))
            throw std::invalid_argument("cell_size must be finite and positive");
        FlowOptions fo;
        fo.nr = o->n_rows;
        fo.nc = o->n_cols;
        fo.dx = o->cell_size;
        fo.bc = bc;
        fo.interior_nodata_is_outlet = o->interior_nodata_is_outlet != 0;
        fo.allow_pits_as_outlets = o->allow_pits_as_outlets != 0;
        auto f = std::make_unique<lsdtt_gpu_flow>();
        f->o = *o;
        f->eng = std::make_unique<FlowEngine>(fo);
        return f.release();
    });
}

void lsdtt_gpu_flow_destroy(lsdtt_gpu_flow* f) { delete f; }

int lsdtt_gpu_flow_build_f32(lsdtt_gpu_flow* f, const float* z, const uint8_t* mask, const int64_t* forced,
                             int64_t n_forced)
{
    return guarded([&] { return build_host<float>(f, z, mask, forced, n_forced, f->z32); });
}

int lsdtt_gpu_flow_build_f64(lsdtt_gpu_flow* f, const double* z, const uint8_t* mask, const int64_t* forced,
                             int64_t n_forced)
{
    return guarded([&] { return build_host<double>(f, z, mask, forced, n_forced, f->z64); });
}

int lsdtt_gpu_flow_build_device(lsdtt_gpu_flow* f, const void* d_z, int32_t z_is_f64, const uint8_t* d_mask,
                                void* stream)
{
    return guarded([&] {
        if (!f || !d_z || !d_mask) return fail_arg("null argument");
        hipStream_t s = static_cast<hipStream_t>(stream);
        f->built = false;
        const long long rc = z_is_f64 ? f->eng->route<double>(static_cast<const double*>(d_z), d_mask, nullptr, s)
                                      : f->eng->route<float>(static_cast<const float*>(d_z), d_mask, nullptr, s);
        if (rc < 0) return finish_build(f, rc);
        const bool topo = f->o.build_topological_order != 0;
        const bool sv = f->o.build_s_vector != 0;
        f->eng->upstream(true, nullptr, nullptr, s, sv);
        if (topo) f->eng->topological_order(s);
        if (sv) f->eng->s_vector(s);
        f->built = true;
        f->has_topo = topo;
        f->has_sv = sv;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_flow_sizes(const lsdtt_gpu_flow* f, int64_t* n_valid, int64_t* n_outlets, int64_t* max_depth)
{
    return guarded([&] {
        if (!f || !f->built) return fail_arg("flow object not built");
        if (n_valid) *n_valid = f->eng->n_valid;
        if (n_outlets) *n_outlets = f->eng->n_outlets;
        if (max_depth) *max_depth = f->eng->max_depth;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_flow_get(lsdtt_gpu_flow* f, int64_t* receiver, uint8_t* dir, uint8_t* donors, int64_t* outlets,
                       int64_t* topo, int64_t* s_vector, int64_t* s_start, uint64_t* s_size)
{
    return guarded([&] {
        if (!f || !f->built) return fail_arg("flow object not built");
        if ((topo && !f->has_topo) || ((s_vector || s_start || s_size) && !f->has_sv))
            return fail_arg("requested a component that was not built");
        FlowEngine& e = *f->eng;
        hipStream_t s = nullptr;
        LSDTT_HIP(hipStreamSynchronize(s));
        get_i64(e.rcv.p, e.n, receiver, s);
        if (dir) LSDTT_HIP(hipMemcpy(dir, e.dir.p, e.n, hipMemcpyDeviceToHost));
        if (donors) LSDTT_HIP(hipMemcpy(donors, e.donors.p, e.n, hipMemcpyDeviceToHost));
        get_i64(e.outlets.p, e.n_outlets, outlets, s);
        if (topo) get_i64(e.topo.p, e.n_valid, topo, s);
        if (s_vector) get_i64(e.svec.p, e.n_valid, s_vector, s);
        if (s_start || s_size) {
            std::vector<uint8_t> m(static_cast<size_t>(e.n));
            std::vector<int> st(static_cast<size_t>(e.n)), c(static_cast<size_t>(e.n));
            LSDTT_HIP(hipMemcpy(m.data(), e.mask, e.n, hipMemcpyDeviceToHost));
            LSDTT_HIP(hipMemcpy(st.data(), e.start.p, sizeof(int) * e.n, hipMemcpyDeviceToHost));
            LSDTT_HIP(hipMemcpy(c.data(), e.cnt.p, sizeof(int) * e.n, hipMemcpyDeviceToHost));
            for (long long i = 0; i < e.n; ++i) {
                const bool v = m[static_cast<size_t>(i)] != 0;
                if (s_start) s_start[i] = v ? st[static_cast<size_t>(i)] : -1;
                if (s_size) s_size[i] = v ? static_cast<uint64_t>(c[static_cast<size_t>(i)]) : 0u;
            }
        }
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_flow_area(lsdtt_gpu_flow* f, double nodata, float* area, uint8_t* area_mask, int64_t* counts)
{
    return guarded([&] {
        if (!f || !f->built) return fail_arg("flow object not built");
        FlowEngine& e = *f->eng;
        std::vector<uint8_t> m(static_cast<size_t>(e.n));
        std::vector<int> c(static_cast<size_t>(e.n));
        LSDTT_HIP(hipMemcpy(m.data(), e.mask, e.n, hipMemcpyDeviceToHost));
        LSDTT_HIP(hipMemcpy(c.data(), e.cnt.p, sizeof(int) * e.n, hipMemcpyDeviceToHost));
        const double pa = f->o.cell_size * f->o.cell_size;
        for (long long i = 0; i < e.n; ++i) {
            const bool v = m[static_cast<size_t>(i)] != 0;
            if (area) area[i] = v ? static_cast<float>(static_cast<double>(c[static_cast<size_t>(i)]) * pa)
                                  : static_cast<float>(nodata);
            if (area_mask) area_mask[i] = v ? 1 : 0;
            if (counts) counts[i] = v ? c[static_cast<size_t>(i)] : 0;
        }
        return LSDTT_GPU_OK;
    });
}

void lsdtt_gpu_steady_state_options_init(lsdtt_gpu_steady_state_options* o)
{
    if (!o) return;
    const SteadyOptions d;
    o->U = d.U;
    o->K = d.K;
    o->m = d.m;
    o->n = d.n;
    o->relief_mode = d.relief_mode ? 1 : 0;
    o->target_relief = d.target_relief;
    o->baselevel_elevation = d.baselevel_elevation;
    o->divide_area_by_dx = d.divide_area_by_dx ? 1 : 0;
    o->channel_threshold_area = d.channel_threshold_area;
}

int lsdtt_gpu_flow_steady_state(lsdtt_gpu_flow* f, const lsdtt_gpu_steady_state_options* o,
                                const float* erodibility_field, double nodata, float* z_out,
                                lsdtt_gpu_steady_state_result* res)
{
    return guarded([&] {
        if (!f || !f->built || !o || !z_out) return fail_arg("flow object not built or null argument");
        FlowEngine& e = *f->eng;
        const long long n = e.n;
        const SteadyOptions so = steady_options(o);
        hipStream_t s = nullptr;
        dbuf<float> w, Kd, zd;
        if (erodibility_field) {
            std::vector<uint8_t> m(static_cast<size_t>(n));
            LSDTT_HIP(hipMemcpy(m.data(), e.mask, n, hipMemcpyDeviceToHost));
            const std::string err = erodibility_field_error(erodibility_field, m.data(), n, f->o.n_cols, so.K);
            if (!err.empty()) return fail_arg(err);
            Kd.reserve(n);
            w.reserve(n);
            LSDTT_HIP(hipMemcpyAsync(Kd.p, erodibility_field, sizeof(float) * n, hipMemcpyHostToDevice, s));
            chi_weights(e.mask, Kd.p, n, so.K, so.n, w.p, s);
        }
        zd.reserve(n);
        const SteadyResult r =
            steady_state_d8(e, so, erodibility_field ? w.p : nullptr, static_cast<float>(nodata), zd.p, s);
        LSDTT_HIP(hipMemcpy(z_out, zd.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        copy_steady_result(r, res);
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_flow_accumulate(lsdtt_gpu_flow* f, const double* weights, double* accum)
{
    return guarded([&] {
        if (!f || !f->built || !accum) return fail_arg("flow object not built or null output");
        FlowEngine& e = *f->eng;
        const long long n = e.n;
        std::vector<uint8_t> m(static_cast<size_t>(n));
        LSDTT_HIP(hipMemcpy(m.data(), e.mask, n, hipMemcpyDeviceToHost));
        std::vector<double> w(static_cast<size_t>(n), 0.0);
        const double w0 = f->o.cell_size * f->o.cell_size;
        for (long long i = 0; i < n; ++i) {
            if (!m[static_cast<size_t>(i)]) continue;
            const double v = weights ? weights[i] : w0;
            if (!std::isfinite(v))
                return fail_arg("lsdtt_flowinfo::accumulate: cell_weights must be finite for all valid cells");
            w[static_cast<size_t>(i)] = v;
        }
        f->wbuf.reserve(n);
        f->abuf.reserve(n);
        hipStream_t s = nullptr;
        LSDTT_HIP(hipMemcpyAsync(f->wbuf.p, w.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        LSDTT_HIP(hipMemcpyAsync(f->abuf.p, w.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        e.upstream(true, f->wbuf.p, f->abuf.p, s);
        LSDTT_HIP(hipMemcpyAsync(accum, f->abuf.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        return LSDTT_GPU_OK;
    });
}

namespace {

// validate_flow_for_chi / validate_chi_options
// (lsdtt_chi_extraction_cpu.cpp:57-92), then the chi raster on the device
int chi_prepare(lsdtt_gpu_flow* f, double m_over_n, double A_0, float nodata, float* d_out, hipStream_t s)
{
    if (!f || !f->built) return fail_arg("flow object not built");
    if (!f->has_sv) {
        set_error("lsdtt_chi_extraction_cpu: chi extraction requires a flow object with s_vector built.");
        return LSDTT_GPU_ERR_RUNTIME;
    }
    if (!std::isfinite(m_over_n)) return fail_arg("lsdtt_chi_extraction_cpu: m_over_n must be finite.");
    if (!std::isfinite(A_0) || A_0 <= 0.0) return fail_arg("lsdtt_chi_extraction_cpu: A_0 must be finite and > 0.");
    if (!d_out) {
        f->chi.reserve(f->eng->n);
        d_out = f->chi.p;
    }
    f->chi_eng.raster(*f->eng, m_over_n, A_0, nodata, d_out, s);
    return LSDTT_GPU_OK;
}

}  // namespace

int lsdtt_gpu_flow_chi(lsdtt_gpu_flow* f, double m_over_n, double A_0, double nodata, float* chi, uint8_t* chi_mask)
{
    return guarded([&] {
        hipStream_t s = nullptr;
        const int rc = chi_prepare(f, m_over_n, A_0, static_cast<float>(nodata), nullptr, s);
        if (rc != LSDTT_GPU_OK) return rc;
        FlowEngine& e = *f->eng;
        if (chi) LSDTT_HIP(hipMemcpyAsync(chi, f->chi.p, sizeof(float) * e.n, hipMemcpyDeviceToHost, s));
        if (chi_mask) LSDTT_HIP(hipMemcpyAsync(chi_mask, e.mask, e.n, hipMemcpyDeviceToHost, s));
        LSDTT_HIP(hipStreamSynchronize(s));
        if (chi_mask)
            for (long long i = 0; i < e.n; ++i) chi_mask[i] = chi_mask[i] ? 1 : 0;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_flow_chi_device(lsdtt_gpu_flow* f, double m_over_n, double A_0, double nodata, float* d_chi,
                              void* stream)
{
    return guarded([&] {
        if (!d_chi) return fail_arg("null argument");
        return chi_prepare(f, m_over_n, A_0, static_cast<float>(nodata), d_chi, static_cast<hipStream_t>(stream));
    });
}

int lsdtt_gpu_flow_chi_threshold(lsdtt_gpu_flow* f, double m_over_n, double A_0, int64_t min_pixels,
                                 int64_t* n_pixels, int64_t* n_basins, int64_t* flats, float* chi,
                                 int64_t* receivers, int64_t* basin_start)
{
    return guarded([&] {
        if (min_pixels <= 0)
            return fail_arg("lsdtt_chi_extraction_cpu: minimum_contributing_pixels must be > 0.");
        hipStream_t s = nullptr;
        const int rc = chi_prepare(f, m_over_n, A_0, 0.0f, nullptr, s);
        if (rc != LSDTT_GPU_OK) return rc;
        FlowEngine& e = *f->eng;
        ChiEngine& c = f->chi_eng;
        const long long m = c.threshold(e, f->chi.p, min_pixels, s);
        std::vector<int> fl(static_cast<size_t>(m)), rv(static_cast<size_t>(m));
        if (m > 0) {
            LSDTT_HIP(hipMemcpyAsync(fl.data(), c.sel_flat.p, sizeof(int) * m, hipMemcpyDeviceToHost, s));
            LSDTT_HIP(hipMemcpyAsync(rv.data(), c.sel_rcv.p, sizeof(int) * m, hipMemcpyDeviceToHost, s));
            if (chi) LSDTT_HIP(hipMemcpyAsync(chi, c.sel_chi.p, sizeof(float) * m, hipMemcpyDeviceToHost, s));
        }
        LSDTT_HIP(hipStreamSynchronize(s));
        // a selected pixel whose receiver is an outlet code starts a basin
        // (outlets meeting the threshold, row-major: the s-vector's order)
        long long nb = 0;
        for (long long i = 0; i < m; ++i) {
            const int r = rv[static_cast<size_t>(i)];
            if (r < 0) {
                if (basin_start) basin_start[nb] = i;
                ++nb;
            }
            if (flats) flats[i] = fl[static_cast<size_t>(i)];
            if (receivers) receivers[i] = r;
        }
        if (n_pixels) *n_pixels = m;
        if (n_basins) *n_basins = nb;
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_fill_device(const float* d_z, const uint8_t* d_mask, int32_t n_rows, int32_t n_cols, double nodata,
                          const char* boundary_conditions, float* d_filled, void* stream)
{
    return guarded([&] {
        if (!d_z || !d_mask || !d_filled || !boundary_conditions) return fail_arg("null argument");
        if (n_rows <= 0 || n_cols <= 0) return fail_arg("grid must have at least one row and column");
        BC bc;
        if (!parse_bc(boundary_conditions, bc)) return fail_arg("boundary conditions must be four of o/n/p/b");
        const bool pns = bc.n == 'p', pew = bc.e == 'p';
        if (pns != (bc.s == 'p')) return fail_arg("lsdtt::fill_pits: north and south periodic flags must match");
        if (pew != (bc.w == 'p')) return fail_arg("lsdtt::fill_pits: east and west periodic flags must match");
        if (pns && pew)
            return fail_arg("lsdtt::fill_pits: periodic in BOTH directions, so the grid has no drainage exit");
        const long long n = (long long)n_rows * n_cols;
        if (n >= (1ll << 31) - 1) return fail_arg("grids of 2^31 cells or more are not supported");
        // working keys; kept for the life of the process (never released at
        // exit, when the HIP runtime may already be torn down).  One
        // workspace serves every caller: calls hold the lock until their
        // fill has finished on their stream, so calls from other threads or
        // on other streams never share it while it is in use.
        static std::mutex* work_lock = new std::mutex();
        static dbuf<int>* work = new dbuf<int>();
        std::lock_guard<std::mutex> hold(*work_lock);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        work->reserve(fill_scratch_ints(n_rows, n_cols));
        fill_device(d_z, d_mask, n_rows, n_cols, nodata, bc, d_filled, work->p, s);
        LSDTT_HIP(hipStreamSynchronize(s));
        return LSDTT_GPU_OK;
    });
}

int lsdtt_gpu_fill(const float* z, const uint8_t* mask, int32_t n_rows, int32_t n_cols, double nodata,
                   const char* boundary_conditions, float* filled)
{
    return guarded([&] {
        if (!z || !mask || !filled) return fail_arg("null argument");
        if (n_rows <= 0 || n_cols <= 0) return fail_arg("grid must have at least one row and column");
        const long long n = (long long)n_rows * n_cols;
        dbuf<float> dz, dout;
        dbuf<uint8_t> dm;
        dz.reserve(n);
        dout.reserve(n);
        dm.reserve(n);
        LSDTT_HIP(hipMemcpy(dz.p, z, sizeof(float) * n, hipMemcpyHostToDevice));
        LSDTT_HIP(hipMemcpy(dm.p, mask, n, hipMemcpyHostToDevice));
        const int rc = lsdtt_gpu_fill_device(dz.p, dm.p, n_rows, n_cols, nodata, boundary_conditions, dout.p, nullptr);
        if (rc != LSDTT_GPU_OK) return rc;
        LSDTT_HIP(hipMemcpy(filled, dout.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        return LSDTT_GPU_OK;
    });
}

}  // extern "C"
