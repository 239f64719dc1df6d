// cltem_amd engine: HIP runtime state and the C ABI of include/cltem_amd.h.
//
// One engine = one GPU, one HIP stream (default priority), device buffers owned by
// the engine and reused across calls, as clTEM's SimulationGeneral owns its clMemory
// members (microscope/simulationgeneral.h:74-95).  Errors are C++ exceptions inside
// and become an int status plus a thread-local message at the ABI.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cltem_amd.h"
#include "atom_kernels.hpp"
#include "engine_internal.hpp"
#include "potential_table.hpp"

namespace mse_geom {
void kgrid(int res, double real_scale, double* kx, double* ky, double* bandwidthkmax);
}

// Device workspace of mse_stem_detect_noalloc: detector bitmasks followed by the FP64
// per-tile partial sums.  Each GPU holds its own copy for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress returns for the current device
// (noalloc_scratch below), never through this name.
static constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];


namespace mse_rt {


void destroy_stem_graph(mse_engine* e);

// Per-device address of `scratch`, looked up once per device with that device current.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

void init_noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)device < g_scratch_ptr.size() && g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p)
        throw Error(MSE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    g_scratch_ptr[device] = p;
}

void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    void* p = (size_t)device < g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
    REQUIRE(p, MSE_ERR_STATE, "no workspace address for device " + std::to_string(device));
    return p;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        HIPX(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

void drain_profile(mse_engine* e) {
    if (e->pending.empty()) return;
    HIPX(hipStreamSynchronize(e->stream));
    for (auto& t : e->pending) {
        float ms = 0.f;
        HIPX(hipEventElapsedTime(&ms, std::get<1>(t), std::get<2>(t)));
        e->prof_ms[std::get<0>(t)] += ms;
        e->prof_n[std::get<0>(t)] += 1;
        e->ev_pool.push_back(std::get<1>(t));
        e->ev_pool.push_back(std::get<2>(t));
    }
    e->pending.clear();
}

template <class F>
void dispatch_res(int res, F&& f) {
    switch (res) {
        case 256: f(std::integral_constant<int, 256>{}); break;
        case 384: f(std::integral_constant<int, 384>{}); break;
        case 512: f(std::integral_constant<int, 512>{}); break;
        case 768: f(std::integral_constant<int, 768>{}); break;
        case 1024: f(std::integral_constant<int, 1024>{}); break;
        default:
            throw Error(MSE_ERR_UNSUPPORTED,
                        "resolution " + std::to_string(res) +
                            " has no FFT plan (supported: 256, 384, 512, 768, 1024)");
    }
}

// ---- sweeps: dispatch on the resolution to the per-size launchers ---------------------------
// the variant of a sweep over nsets waves: the one tuned for that wave count, else the default
int tuned_variant(const mse_engine* e, int nsets, bool column) {
    for (const auto& t : e->tuned)
        if (t.first == nsets) return column ? t.second.second : t.second.first;
    return column ? e->col_var : e->row_var;
}

void row_pass(mse_engine* e, int mode, const float2* in, float2* out, int nlines_sets,
              size_t wave_stride, const mse::RowArgs& args, int variant) {
    if (variant < 0) variant = tuned_variant(e, nlines_sets, false);
    switch (e->res) {
        case 256: row_pass_n<256>(e, mode, in, out, nlines_sets, wave_stride, args, variant); break;
        case 384: row_pass_n<384>(e, mode, in, out, nlines_sets, wave_stride, args, variant); break;
        case 512: row_pass_n<512>(e, mode, in, out, nlines_sets, wave_stride, args, variant); break;
        case 768: row_pass_n<768>(e, mode, in, out, nlines_sets, wave_stride, args, variant); break;
        case 1024: row_pass_n<1024>(e, mode, in, out, nlines_sets, wave_stride, args, variant); break;
        default: dispatch_res(e->res, [](auto) {});
    }
}

// returns the number of column tiles of the launch (detector partials per wave)
int col_pass(mse_engine* e, int mode, const float2* in, float2* out, int nsets, size_t wave_stride,
             const mse::ColArgs& args, int kind, int variant) {
    if (variant < 0) variant = tuned_variant(e, nsets, true);
    switch (e->res) {
        case 256: return col_pass_n<256>(e, mode, in, out, nsets, wave_stride, args, kind, variant);
        case 384: return col_pass_n<384>(e, mode, in, out, nsets, wave_stride, args, kind, variant);
        case 512: return col_pass_n<512>(e, mode, in, out, nsets, wave_stride, args, kind, variant);
        case 768: return col_pass_n<768>(e, mode, in, out, nsets, wave_stride, args, kind, variant);
        case 1024: return col_pass_n<1024>(e, mode, in, out, nsets, wave_stride, args, kind, variant);
        default: dispatch_res(e->res, [](auto) {});
    }
    return 0;
}

mse::ColArgs base_col_args(mse_engine* e) {
    mse::ColArgs a{};
    a.px = e->px.p;
    a.py = e->py.p;
    a.wave_prop = e->any_beam ? e->wave_prop.p : nullptr;
    a.lim = e->prop_lim.p;
    a.kx = e->kx.p;
    a.ky = e->ky.p;
    a.scale = 1.0f;
    a.det_scale = 1.0;
    a.ndet = 0;
    return a;
}

mse::RowArgs base_row_args(mse_engine* e) {
    mse::RowArgs a{};
    a.cfg_trans = e->cfg_ptrs.p;
    a.cfg_nslices = e->cfg_nslices.p;
    a.dout = nullptr;
    a.wave_cfg = e->wave_cfg.p;
    a.slice = 0;
    a.plane_value = 0.f;
    a.scale = 1.0f;
    a.mix_key = e->mix_key.p;
    a.mix_seed = make_uint2((unsigned)(e->mix_seed & 0xffffffffu), (unsigned)(e->mix_seed >> 32));
    a.mix_first = e->mix_first;
    a.mix_sets = e->mix_sets >= 2 ? e->mix_sets : 0;
    return a;
}

void require_geom(mse_engine* e) { REQUIRE(e->have_geom, MSE_ERR_STATE, "mse_set_geometry has not been called"); }

const mse_geometry& cfg_geom(mse_engine* e, const ConfigData& c) { return c.own_geom ? c.g : e->g; }

// device tables of every configuration's transmission base pointer and slice count
void upload_cfg_ptrs(mse_engine* e) {
    std::vector<const float2*> h(e->configs.size());
    std::vector<int> ns(e->configs.size());
    for (size_t i = 0; i < h.size(); ++i) {
        h[i] = e->configs[i].trans.p;
        ns[i] = cfg_geom(e, e->configs[i]).n_slices;
    }
    const size_t cap = std::max<size_t>(h.size(), 1);
    if (e->cfg_ptrs.n < cap || e->cfg_nslices.n < cap) {
        HIPX(hipStreamSynchronize(e->stream));
        e->cfg_ptrs.ensure(cap);
        e->cfg_nslices.ensure(cap);
    }
    if (!h.empty()) {
        HIPX(hipMemcpyAsync(e->cfg_ptrs.p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->cfg_nslices.p, ns.data(), ns.size() * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    }
}

mse::PotArgs pot_args(mse_engine* e, const ConfigData& c) {
    const mse_geometry& g = cfg_geom(e, c);
    mse::PotArgs a{};
    a.res = e->res;
    a.pixelscale = g.real_scale;
    a.startx = g.sim_x[0] + c.ref_off[0];  // reference_perturb (simulationgeneral.cpp:559-560)
    a.starty = g.sim_y[0] + c.ref_off[1];
    a.min_x = g.full_x[0];
    a.max_x = g.full_x[1];
    a.min_y = g.full_y[0];
    a.max_y = g.full_y[1];
    a.blocks_x = g.blocks_x;
    a.blocks_y = g.blocks_y;
    a.rrx = 1.0 / (a.max_x - a.min_x);
    a.rry = 1.0 / (a.max_y - a.min_y);
    a.inv_ps = (float)(1.0 / a.pixelscale);
    a.reach_px2 = (float)(64.01 / (a.pixelscale * a.pixelscale));
    a.load_x = g.load_x;
    a.load_y = g.load_y;
    a.load_z = g.load_z;
    a.n_slices = g.n_slices;
    const double beam_theta = M_PI_2 - g.beam_tilt_mrad * 0.001;  // ..._projected_d.cl:221
    a.cos_phi = std::cos(g.beam_azimuth);
    a.sin_phi = std::sin(g.beam_azimuth);
    a.sin_2theta = std::sin(2.0 * beam_theta);
    a.tan_theta = std::tan(beam_theta);
    a.untilted = g.beam_tilt_mrad == 0.0;
    a.sigma = g.sigma_eff;
    a.dz = g.slice_thickness;
    a.z_min = g.z_lims[0];
    a.shift_x = (g.k_vec[0] / g.k_vec[2]) * g.slice_thickness / g.full3d_integrals;  // simulationgeneral.cpp:562
    a.shift_y = (g.k_vec[1] / g.k_vec[2]) * g.slice_thickness / g.full3d_integrals;
    if (c.own_beam) {
        // modifyBeamTilt (simulationgeneral.cpp:649-679): tilt and azimuth of the configuration's
        // beam vector; the tilt is passed in mrad as the kernel expects it, and the full-3D y
        // shift uses ky
        const double kx = c.k_vec[0], ky = c.k_vec[1], kz = c.k_vec[2];
        const double theta = M_PI_2 - std::atan(std::sqrt(kx * kx + ky * ky) / kz);
        const double phi = std::atan2(ky, kx);
        a.cos_phi = std::cos(phi);
        a.sin_phi = std::sin(phi);
        a.sin_2theta = std::sin(2.0 * theta);
        a.tan_theta = std::tan(theta);
        a.untilted = kx == 0.0 && ky == 0.0;
        a.shift_x = (kx / kz) * g.slice_thickness / g.full3d_integrals;
        a.shift_y = (ky / kz) * g.slice_thickness / g.full3d_integrals;
    }
    a.integrals = g.full3d_integrals;
    a.per_atom = e->per_atom;
    a.i_count = e->i_count;
    return a;
}

// Propagator (kernels/propagator_d.cl:49-83) for beam vector kv, split as
// f e^{i pi dz (k^2-kz^2)/kz} e^{-i pi dz (bx+kx)^2/kz} e^{-i pi dz (by+ky)^2/kz}
// (separable for any tilt), with the two unitary FFT scales 1/res^2 folded into px.
// Needs the engine's k-grid (h_kx, h_ky) and geometry.
void propagator_tables(mse_engine* e, const double* kv, std::vector<float2>& hpx, std::vector<float2>& hpy) {
    const int res = e->res;
    const double dz = e->g.slice_thickness;
    const double k = e->g.wavenumber, bkx = kv[0], bky = kv[1], bkz = kv[2];
    const double f = bkz / k;
    const double c0 = M_PI * dz * (k * k - bkz * bkz) / bkz;
    const double inv_n = 1.0 / ((double)res * (double)res);
    hpx.resize(res);
    hpy.resize(res);
    for (int i = 0; i < res; ++i) {
        double ax = std::remainder(c0 - M_PI * dz * (bkx + e->h_kx[i]) * (bkx + e->h_kx[i]) / bkz, 2.0 * M_PI);
        hpx[i] = make_float2((float)(f * inv_n * std::cos(ax)), (float)(f * inv_n * std::sin(ax)));
        double ay = std::remainder(-M_PI * dz * (bky + e->h_ky[i]) * (bky + e->h_ky[i]) / bkz, 2.0 * M_PI);
        hpy[i] = make_float2((float)std::cos(ay), (float)std::sin(ay));
    }
}

// Per-wave buffers for n waves: the waves, their configuration table, probe positions and
// mixed-potential keys (identity keys when the key buffer is new).  Kernels that may read the
// old buffers must have finished (callers synchronise first).
void size_wave_buffers(mse_engine* e, int n) {
    e->waves.ensure((size_t)n * e->N);
    e->wave_cfg.ensure(n);
    e->probe_pos.ensure(2 * (size_t)n);
    if (e->mix_key.n < (size_t)n || !e->mix_key.p) {
        e->mix_key.ensure(n);
        std::vector<unsigned> keys(n);
        for (int i = 0; i < n; ++i) keys[i] = (unsigned)i;
        HIPX(hipMemcpyAsync(e->mix_key.p, keys.data(), n * sizeof(unsigned), hipMemcpyHostToDevice, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    }
}

void sync_wave_prop(mse_engine* e) {
    const int nw = std::min<int>(e->n_waves, (int)e->h_wave_cfg.size());
    bool any = false;
    for (int w = 0; w < nw; ++w) {
        const int c = e->h_wave_cfg[w];
        if (c >= 0 && c < (int)e->configs.size() && e->configs[c].own_beam) any = true;
    }
    e->any_beam = any;
    if (!any) return;
    e->h_wave_prop.assign(2 * (size_t)nw, nullptr);
    for (int w = 0; w < nw; ++w) {
        const int ci = e->h_wave_cfg[w];
        const bool own = ci >= 0 && ci < (int)e->configs.size() && e->configs[ci].own_beam;
        e->h_wave_prop[2 * w] = own ? e->configs[ci].px.p : e->px.p;
        e->h_wave_prop[2 * w + 1] = own ? e->configs[ci].py.p : e->py.p;
    }
    HIPX(hipStreamSynchronize(e->stream));  // kernels in flight may read the table
    e->wave_prop.ensure(2 * (size_t)nw);
    HIPX(hipMemcpyAsync(e->wave_prop.p, e->h_wave_prop.data(), 2 * (size_t)nw * sizeof(const float2*),
                        hipMemcpyHostToDevice, e->stream));
    HIPX(hipStreamSynchronize(e->stream));
}

ConfigData& config(mse_engine* e, int c) {
    REQUIRE(c >= 0 && c < 4096, MSE_ERR_ARG, "configuration index out of range");
    if ((size_t)c >= e->configs.size()) e->configs.resize(c + 1);
    return e->configs[c];
}

// Detector bitmasks over the fft-shifted image, decided in double with band_pass's test
// (kernels/band_pass_d.cl:98-121) on the radii of getStemPixel (simulationstem.cpp:105-110).
void ensure_detectors(mse_engine* e, const mse_detector* dets, int n_det) {
    REQUIRE(n_det >= 0 && n_det <= mse::kMaxDet, MSE_ERR_ARG, "at most 8 detectors");
    std::vector<double> key;
    for (int d = 0; d < n_det; ++d) {
        key.push_back(dets[d].inner_mrad);
        key.push_back(dets[d].outer_mrad);
        key.push_back(dets[d].xc_mrad);
        key.push_back(dets[d].yc_mrad);
    }
    key.push_back((double)e->res);
    key.push_back(e->g.inverse_scale_angle);
    if (key == e->det_key && e->det_bits.p) return;
    const int W = e->res, H = e->res;
    const size_t words = (size_t)W * H / 32;
    std::vector<unsigned> bits(words * std::max(n_det, 1), 0u);
    const double angle_scale = e->g.inverse_scale_angle;
    for (int d = 0; d < n_det; ++d) {
        const double inner = dets[d].inner_mrad / angle_scale;
        const double outer = dets[d].outer_mrad / angle_scale;
        const double centX = W / 2.0 + dets[d].xc_mrad / angle_scale;
        const double centY = H / 2.0 + dets[d].yc_mrad / angle_scale;
        for (int yid = 0; yid < H; ++yid)
            for (int xid = 0; xid < W; ++xid) {
                const double radius = std::sqrt((xid - centX) * (xid - centX) + (yid - centY) * (yid - centY));
                if (radius <= outer && radius >= inner) {
                    const size_t b = (size_t)yid * W + xid;
                    bits[d * words + (b >> 5)] |= 1u << (b & 31);
                }
            }
    }
    e->det_bits.ensure(bits.size());
    HIPX(hipMemcpyAsync(e->det_bits.p, bits.data(), bits.size() * sizeof(unsigned), hipMemcpyHostToDevice, e->stream));
    HIPX(hipStreamSynchronize(e->stream));
    e->det_key = key;
}

// Detector radii in pixels of the fft-shifted image, as ensure_detectors derives them.
struct DetectorRadii {
    double inner[mse::kMaxDet], outer[mse::kMaxDet], cx[mse::kMaxDet], cy[mse::kMaxDet];
};

DetectorRadii detector_radii(const mse_engine* e, const mse_detector* dets, int n_det) {
    DetectorRadii r{};
    const double angle_scale = e->g.inverse_scale_angle;
    for (int d = 0; d < n_det; ++d) {
        r.inner[d] = dets[d].inner_mrad / angle_scale;
        r.outer[d] = dets[d].outer_mrad / angle_scale;
        r.cx[d] = e->res / 2.0 + dets[d].xc_mrad / angle_scale;
        r.cy[d] = e->res / 2.0 + dets[d].yc_mrad / angle_scale;
    }
    return r;
}

// The bitmasks of ensure_detectors built on the device: one 32-pixel word per thread,
// blockIdx.y = detector.  The same double expression without contraction and a correctly
// rounded square root give the host's decisions bit for bit.  Every word is written.
__global__ void k_detector_bits(unsigned* __restrict__ bits, int res, DetectorRadii r) {
#pragma clang fp contract(off)
    const size_t words = (size_t)res * res / 32;
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int d = blockIdx.y;
    if (w >= words) return;
    unsigned m = 0u;
    for (int k = 0; k < 32; ++k) {
        const size_t b = w * 32 + k;
        const int yid = (int)(b / res), xid = (int)(b % res);
        const double dx = xid - r.cx[d], dy = yid - r.cy[d];
        const double radius = sqrt(dx * dx + dy * dy);
        if (radius <= r.outer[d] && radius >= r.inner[d]) m |= 1u << k;
    }
    bits[(size_t)d * words + w] = m;
}

int run_guarded(const char* fn, mse_engine* e, const std::function<void()>& body) {
    try {
        if (e) {
            DeviceGuard dg(e->device);
            body();
        } else {
            body();
        }
        return MSE_OK;
    } catch (const Error& err) {
        g_err = std::string(fn) + ": " + err.what();
        return err.code;
    } catch (const std::exception& ex) {
        g_err = std::string(fn) + ": " + ex.what();
        return MSE_ERR_STATE;
    }
}

}  // namespace mse_rt

extern "C" {

const char* mse_last_error(void) { return g_err.c_str(); }

int mse_create(int device_ordinal, mse_engine** out) {
    return run_guarded("mse_create", nullptr, [&]() {
        REQUIRE(out != nullptr, MSE_ERR_ARG, "out is NULL");
        int count = 0;
        HIPX(hipGetDeviceCount(&count));
        REQUIRE(device_ordinal >= 0 && device_ordinal < count, MSE_ERR_ARG,
                "device ordinal " + std::to_string(device_ordinal) + " not present (" + std::to_string(count) + " devices)");
        DeviceGuard dg(device_ordinal);
        init_noalloc_scratch(device_ordinal);
        auto* e = new mse_engine();
        e->device = device_ordinal;
        {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess && prop.multiProcessorCount > 0)
                e->num_cus = prop.multiProcessorCount;
        }
        hipError_t st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
        if (st != hipSuccess) {
            delete e;
            throw Error(MSE_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(st));
        }
        *out = e;
    });
}

int mse_destroy(mse_engine* e) {
    if (!e) return MSE_OK;
    int rc = run_guarded("mse_destroy", e, [&]() {
        (void)hipStreamSynchronize(e->stream);
        for (auto& t : e->pending) {
            (void)hipEventDestroy(std::get<1>(t));
            (void)hipEventDestroy(std::get<2>(t));
        }
        for (auto ev : e->ev_pool) (void)hipEventDestroy(ev);
        destroy_stem_graph(e);
        e->configs.clear();
        hipStream_t s = e->stream;
        delete e;
        (void)hipStreamDestroy(s);
    });
    return rc;
}

int mse_synchronize(mse_engine* e) {
    return run_guarded("mse_synchronize", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_device_info(mse_engine* e, char* name, int name_len, int* cu_count, double* hbm_gbs) {
    return run_guarded("mse_device_info", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        hipDeviceProp_t p;
        HIPX(hipGetDeviceProperties(&p, e->device));
        if (name && name_len > 0) {
            std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
        }
        if (cu_count) *cu_count = p.multiProcessorCount;
        if (hbm_gbs) *hbm_gbs = 2.0 * p.memoryClockRate * 1e3 * (p.memoryBusWidth / 8) / 1e9;
    });
}

int mse_set_geometry(mse_engine* e, const mse_geometry* g) {
    return run_guarded("mse_set_geometry", e, [&]() {
        REQUIRE(e && g, MSE_ERR_ARG, "NULL argument");
        const int res = g->resolution;
        dispatch_res(res, [](auto) {});
        REQUIRE(g->n_slices > 0 && g->blocks_x > 0 && g->blocks_y > 0, MSE_ERR_ARG, "geometry not planned");
        HIPX(hipStreamSynchronize(e->stream));
        const bool res_changed = res != e->res;
        e->g = *g;
        e->res = res;
        e->N = (size_t)res * res;
        if (res_changed) {
            for (auto& c : e->configs) c.trans.release();
            e->waves.release();
            e->n_waves = 0;
            e->det_key.clear();
        }
        e->adj_ready = false;  // checkpoints belong to the previous geometry
        for (auto& c : e->configs) {  // bins and transmission functions depend on the geometry
            c.sorted = false;
            c.built = false;
            ++c.version;
            if (c.own_geom) c.trans.release();
            c.own_geom = false;
            c.own_beam = false;  // its propagator tables belong to the previous k-grid
        }
        // the device table of transmission pointers and slice counts must not keep pointers
        // released above or slice counts of the previous geometry
        if (!e->configs.empty()) upload_cfg_ptrs(e);
        // k-grid exactly as the reference builds it (double)
        e->h_kx.assign(res, 0.0);
        e->h_ky.assign(res, 0.0);
        double bw = 0.0;
        mse_geom::kgrid(res, g->real_scale, e->h_kx.data(), e->h_ky.data(), &bw);
        // twiddles tw[q] = exp(-2 pi i q / N), evaluated in double
        std::vector<float2> tw(res);
        for (int q = 0; q < res; ++q) {
            const double ang = -2.0 * M_PI * (double)q / (double)res;
            tw[q] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        std::vector<float2> hpx, hpy;
        propagator_tables(e, g->k_vec, hpx, hpy);
        // masks, decided in double with the reference formulas, stored as per-row |kx index| limits
        const double k_max = bw * g->limit_factor;           // simulationgeneral.cpp:633
        const double k_max_2 = k_max * k_max;                 // propagator_d.cl:69-71 (strict <)
        const double band_lim = bw * g->limit_factor;         // band_limit_d.cl:24-30 (<=)
        std::vector<int> plim(res), blim(res);
        for (int y = 0; y < res; ++y) {
            int pl = -1, bl = -1;
            for (int x = 0; x < res; ++x) {
                const int xs = (x < res / 2) ? x : res - x;
                const double k0x = e->h_kx[x] * e->h_kx[x];
                const double k0y = e->h_ky[y] * e->h_ky[y];
                if (k0x + k0y < k_max_2) pl = std::max(pl, xs);
                if (std::sqrt(e->h_kx[x] * e->h_kx[x] + e->h_ky[y] * e->h_ky[y]) <= band_lim) bl = std::max(bl, xs);
            }
            // the kept sets are {x : |x| <= limit}; verify (monotone in |kx|)
            for (int x = 0; x < res; ++x) {
                const int xs = (x < res / 2) ? x : res - x;
                const double k0x = e->h_kx[x] * e->h_kx[x];
                const double k0y = e->h_ky[y] * e->h_ky[y];
                const bool pk = k0x + k0y < k_max_2;
                const bool bk = std::sqrt(e->h_kx[x] * e->h_kx[x] + e->h_ky[y] * e->h_ky[y]) <= band_lim;
                REQUIRE(pk == (xs <= pl) && bk == (xs <= bl), MSE_ERR_STATE, "non-radial aperture mask");
            }
            plim[y] = pl;
            blim[y] = bl;
        }
        e->tw.ensure(res);
        e->px.ensure(res);
        e->py.ensure(res);
        e->prop_lim.ensure(res);
        e->band_lim.ensure(res);
        // the band limit's column pass runs on the propagation kernels with these factors:
        // x the unitary scale of the 2D transform pair, y one
        const float band_scale = (float)(1.0 / ((double)res * (double)res));
        std::vector<float2> bxy(2 * (size_t)res);
        for (int i = 0; i < res; ++i) {
            bxy[i] = make_float2(band_scale, 0.f);
            bxy[res + i] = make_float2(1.f, 0.f);
        }
        e->band_xy.ensure(2 * (size_t)res);
        HIPX(hipMemcpyAsync(e->band_xy.p, bxy.data(), 2 * (size_t)res * sizeof(float2), hipMemcpyHostToDevice,
                            e->stream));
        e->kx.ensure(res);
        e->ky.ensure(res);
        HIPX(hipMemcpyAsync(e->tw.p, tw.data(), res * sizeof(float2), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->px.p, hpx.data(), res * sizeof(float2), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->py.p, hpy.data(), res * sizeof(float2), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->prop_lim.p, plim.data(), res * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->band_lim.p, blim.data(), res * sizeof(int), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->kx.p, e->h_kx.data(), res * sizeof(double), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->ky.p, e->h_ky.data(), res * sizeof(double), hipMemcpyHostToDevice, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
        e->have_geom = true;
        upload_cfg_ptrs(e);
        sync_wave_prop(e);
    });
}

int mse_set_parameterisation(mse_engine* e, int form, int i_per_atom, const double* params, int n_params) {
    return run_guarded("mse_set_parameterisation", e, [&]() {
        REQUIRE(e && params, MSE_ERR_ARG, "NULL argument");
        REQUIRE(form >= 0 && form <= 2, MSE_ERR_ARG, "form must be 0 (Kirkland), 1 (Peng) or 2 (Lobato)");
        const int per = i_per_atom * (form == 0 ? 4 : 2);
        REQUIRE(i_per_atom > 0 && per > 0 && n_params % per == 0, MSE_ERR_ARG,
                "parameterisation does not contain correct parameters per atom");
        REQUIRE(form == 0 ? i_per_atom == 3 : i_per_atom == 5, MSE_ERR_UNSUPPORTED,
                "parameter layout expects 3 Kirkland or 5 Peng/Lobato terms");
        std::vector<float> f(params, params + n_params);
        e->params.ensure(n_params);
        HIPX(hipMemcpyAsync(e->params.p, f.data(), n_params * sizeof(float), hipMemcpyHostToDevice, e->stream));
        // full-3D constants per element, in double then rounded (v3d_tab's table layout)
        const int nz = n_params / per;
        const double log2e = 1.4426950408889634, pi = 3.141592653589793;
        std::vector<float> t3((size_t)nz * mse::kPot3dStride, 0.f);
        for (int z = 0; z < nz; ++z) {
            const double* p = params + (size_t)z * per;
            float* t = t3.data() + (size_t)z * mse::kPot3dStride;
            for (int i = 0; i < i_per_atom; ++i) {
                if (form == 0) {  // a_i, b_i pairs then c_i, d_i pairs
                    const double a = p[2 * i], b = p[2 * i + 1], c = p[2 * i_per_atom + 2 * i],
                                 d = p[2 * i_per_atom + 2 * i + 1];
                    t[i] = (float)a;
                    t[3 + i] = (float)(2.0 * pi * log2e * std::sqrt(b));
                    t[6 + i] = (float)(c / (d * std::sqrt(d)));
                    t[9 + i] = (float)(pi * pi * log2e / d);
                } else {  // a_i then b_i
                    const double a = p[i], b = p[i_per_atom + i];
                    t[i] = (float)(a / (b * std::sqrt(b)));
                    if (form == 1) {
                        t[5 + i] = (float)(pi * pi * log2e / b);
                    } else {
                        t[5 + i] = (float)(2.0 * pi * log2e / std::sqrt(b));
                        t[10 + i] = (float)(std::sqrt(b) / pi);
                    }
                }
            }
        }
        // projected constants per element (vproj_tab's table layout)
        std::vector<float> tp((size_t)nz * mse::kPotProjStride, 0.f);
        for (int z = 0; z < nz; ++z) {
            const double* p = params + (size_t)z * per;
            float* t = tp.data() + (size_t)z * mse::kPotProjStride;
            for (int i = 0; i < i_per_atom; ++i) {
                if (form == 0) {
                    const double a = p[2 * i], b = p[2 * i + 1], c = p[2 * i_per_atom + 2 * i],
                                 d = p[2 * i_per_atom + 2 * i + 1];
                    t[i] = (float)a;
                    t[3 + i] = (float)(2.0 * pi * std::sqrt(b));
                    t[6 + i] = (float)(c / d);
                    t[9 + i] = (float)(pi * pi * log2e / d);
                } else {
                    const double a = p[i], b = p[i_per_atom + i];
                    if (form == 1) {
                        t[i] = (float)(a / b);
                        t[5 + i] = (float)(pi * pi * log2e / b);
                    } else {
                        t[i] = (float)(a / (b * std::sqrt(b)));
                        t[5 + i] = (float)(2.0 * pi / std::sqrt(b));
                    }
                }
            }
        }
        e->params_proj.ensure(tp.size());
        HIPX(hipMemcpyAsync(e->params_proj.p, tp.data(), tp.size() * sizeof(float), hipMemcpyHostToDevice,
                            e->stream));
        e->params3d.ensure(t3.size());
        HIPX(hipMemcpyAsync(e->params3d.p, t3.data(), t3.size() * sizeof(float), hipMemcpyHostToDevice, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
        e->form = form;
        e->i_count = i_per_atom;
        e->per_atom = per;
        e->max_Z = n_params / per;
        // tabulated potentials: built per element on first use (potential_tables)
        e->h_params.assign(params, params + n_params);
        e->h_vtab.assign((size_t)2 * e->max_Z * mse_tab::kBuckets * 4, 0.f);
        e->vtab_on.assign((size_t)2 * e->max_Z, 0);
        e->vtab.ensure((size_t)2 * e->max_Z * mse_tab::kBuckets);
    });
}

int mse_set_potential_tables(mse_engine* e, int enable) {
    return run_guarded("mse_set_potential_tables", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        REQUIRE(enable >= 0 && enable <= 2, MSE_ERR_ARG, "enable must be 0, 1 or 2");
        e->use_vtab = enable != 0;
        e->vtab_lds = enable == 1;
    });
}

int mse_potential_table(int form, int full3d, const double* params, int n_params, int i_per_atom, int Z,
                        float* out) {
    return run_guarded("mse_potential_table", nullptr, [&]() {
        REQUIRE(params && out, MSE_ERR_ARG, "NULL argument");
        REQUIRE(form >= 0 && form <= 2, MSE_ERR_ARG, "form must be 0 (Kirkland), 1 (Peng) or 2 (Lobato)");
        REQUIRE(form == 0 ? i_per_atom == 3 : i_per_atom == 5, MSE_ERR_UNSUPPORTED,
                "parameter layout expects 3 Kirkland or 5 Peng/Lobato terms");
        const int per = i_per_atom * (form == 0 ? 4 : 2);
        REQUIRE(n_params > 0 && n_params % per == 0 && Z >= 1 && Z <= n_params / per, MSE_ERR_ARG,
                "atomic number outside the parameterisation");
        mse_tab::build(form, full3d != 0, params, i_per_atom, Z, out);
    });
}

int mse_sort_atoms(mse_engine* e, int cfg, const double* x, const double* y, const double* z, const int* Z,
                   int n, int* block_ids, int* z_ids, int* block_start, int* n_in_range) {
    return run_guarded("mse_sort_atoms", e, [&]() {
        REQUIRE(e && x && y && z && Z, MSE_ERR_ARG, "NULL argument");
        REQUIRE(n >= 0, MSE_ERR_ARG, "negative atom count");
        require_geom(e);
        if (e->max_Z > 0)
            for (int i = 0; i < n; ++i)
                REQUIRE(Z[i] >= 1 && Z[i] <= e->max_Z, MSE_ERR_ARG,
                        "atomic number " + std::to_string(Z[i]) + " outside the parameterisation");
        ConfigData& c = config(e, cfg);
        const mse_geometry& g = cfg_geom(e, c);
        const int nblk = g.blocks_x * g.blocks_y;
        const long long nkeys = (long long)g.n_slices * nblk;
        REQUIRE(nkeys < (1LL << 30), MSE_ERR_UNSUPPORTED, "too many bins");
        const size_t na = std::max(n, 1);
        e->tmp_xyz.ensure(3 * na);
        e->tmp_Z.ensure(na);
        e->tmp_bid.ensure(na);
        e->tmp_zid.ensure(na);
        e->keys.ensure(na);
        e->keys_sorted.ensure(na);
        e->idx.ensure(na);
        e->idx_sorted.ensure(na);
        c.block_start.ensure(nkeys + 1);
        c.x.ensure(na);
        c.y.ensure(na);
        c.z.ensure(na);
        c.Z.ensure(na);
        double* dx = e->tmp_xyz.p;
        double* dy = dx + na;
        double* dzp = dy + na;
        if (n > 0) {
            HIPX(hipMemcpyAsync(dx, x, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
            HIPX(hipMemcpyAsync(dy, y, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
            HIPX(hipMemcpyAsync(dzp, z, n * sizeof(double), hipMemcpyHostToDevice, e->stream));
            HIPX(hipMemcpyAsync(e->tmp_Z.p, Z, n * sizeof(int), hipMemcpyHostToDevice, e->stream));
            mse::BinGeom bg{g.full_x[0], g.full_x[1], g.full_y[0], g.full_y[1], g.z_lims[0], g.z_lims[1],
                            g.slice_thickness, g.blocks_x, g.blocks_y, g.n_slices};
            launch(e, PK_OTHER, [&]() {
                mse::k_bin_atoms<<<(n + 255) / 256, 256, 0, e->stream>>>(dx, dy, dzp, n, bg, e->keys.p, e->idx.p,
                                                                         e->tmp_bid.p, e->tmp_zid.p);
            });
            int end_bit = 1;
            while ((1LL << end_bit) <= nkeys) ++end_bit;
            size_t tmp_bytes = 0;
            HIPX(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, e->keys.p, e->keys_sorted.p, e->idx.p,
                                                    e->idx_sorted.p, n, 0, end_bit, e->stream));
            e->cub_tmp.ensure(tmp_bytes + 1);
            HIPX(hipcub::DeviceRadixSort::SortPairs(e->cub_tmp.p, tmp_bytes, e->keys.p, e->keys_sorted.p, e->idx.p,
                                                    e->idx_sorted.p, n, 0, end_bit, e->stream));
            launch(e, PK_OTHER, [&]() {
                mse::k_block_starts<<<(int)((nkeys + 1 + 255) / 256), 256, 0, e->stream>>>(e->keys_sorted.p, n,
                                                                                       (int)nkeys, c.block_start.p);
            });
            int nr = 0;
            HIPX(hipMemcpyAsync(&nr, c.block_start.p + nkeys, sizeof(int), hipMemcpyDeviceToHost, e->stream));
            HIPX(hipStreamSynchronize(e->stream));
            c.n_in_range = nr;
            if (nr > 0) {
                launch(e, PK_OTHER, [&]() {
                    mse::k_gather_atoms<<<(nr + 255) / 256, 256, 0, e->stream>>>(e->idx_sorted.p, nr, dx, dy, dzp,
                                                                               e->tmp_Z.p, c.x.p, c.y.p, c.z.p, c.Z.p);
                });
                c.orig.ensure(nr);
                HIPX(hipMemcpyAsync(c.orig.p, e->idx_sorted.p, nr * sizeof(int), hipMemcpyDeviceToDevice, e->stream));
            }
        } else {
            std::vector<int> zeros(nkeys + 1, 0);
            HIPX(hipMemcpyAsync(c.block_start.p, zeros.data(), zeros.size() * sizeof(int), hipMemcpyHostToDevice,
                                e->stream));
            c.n_in_range = 0;
        }
        if (block_ids && n > 0)
            HIPX(hipMemcpyAsync(block_ids, e->tmp_bid.p, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
        if (z_ids && n > 0)
            HIPX(hipMemcpyAsync(z_ids, e->tmp_zid.p, n * sizeof(int), hipMemcpyDeviceToHost, e->stream));
        if (block_start)
            HIPX(hipMemcpyAsync(block_start, c.block_start.p, (nkeys + 1) * sizeof(int), hipMemcpyDeviceToHost,
                                e->stream));
        HIPX(hipStreamSynchronize(e->stream));
        if (n_in_range) *n_in_range = c.n_in_range;
        c.n_input = n;
        c.has_Z.assign((size_t)std::max(e->max_Z, 0) + 1, 0);
        for (int i = 0; i < n; ++i)
            if (Z[i] >= 1 && Z[i] < (int)c.has_Z.size()) c.has_Z[Z[i]] = 1;
        ++c.version;
        c.sorted = true;
    });
}

int mse_build_transmission(mse_engine* e, int cfg, int first, int count) {
    return run_guarded("mse_build_transmission", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        require_geom(e);
        REQUIRE(e->form >= 0, MSE_ERR_STATE, "mse_set_parameterisation has not been called");
        ConfigData& c = config(e, cfg);
        REQUIRE(c.sorted, MSE_ERR_STATE, "atoms of this configuration are not sorted for this geometry");
        const mse_geometry& g = cfg_geom(e, c);
        REQUIRE(first >= 0 && count >= 0 && first + count <= g.n_slices, MSE_ERR_ARG, "slice range out of bounds");
        if (count == 0) return;
        ++c.version;
        // ensure() may move the buffer (a larger geometry than before): the device pointer
        // table the sweeps read must then be refreshed, or they would read released memory
        const float2* before = c.trans.p;
        c.trans.ensure((size_t)g.n_slices * e->N);
        c.built = true;
        if (c.trans.p != before) upload_cfg_ptrs(e);
        const mse::PotArgs a = pot_args(e, c);
        float2* outp = c.trans.p + (size_t)first * e->N;
        // the full-3D kernels build up to 32 slices per launch (about 9 ms for 20,000 SiO2
        // atoms at 1024^2), the projected ones up to 256 (under a millisecond): few launches,
        // each with a tail that is small against its work
        const int chunk = g.full3d ? 32 : 256;
        for (int s0 = 0; s0 < count; s0 += chunk) {
            const int cs = std::min(chunk, count - s0);
            potential_planes(e, c, g, a, outp + (size_t)s0 * e->N, cs, first + s0);
        }
        band_limit_planes(e, outp, count);
    });
}

int mse_read_transmission(mse_engine* e, int cfg, int slice, float* out) {
    return run_guarded("mse_read_transmission", e, [&]() {
        REQUIRE(e && out, MSE_ERR_ARG, "NULL argument");
        require_geom(e);
        ConfigData& c = config(e, cfg);
        REQUIRE(c.trans.p && c.built, MSE_ERR_STATE, "no transmission functions for this configuration");
        REQUIRE(slice >= 0 && slice < cfg_geom(e, c).n_slices, MSE_ERR_ARG, "slice out of range");
        HIPX(hipMemcpyAsync(out, c.trans.p + (size_t)slice * e->N, e->N * sizeof(float2), hipMemcpyDeviceToHost,
                            e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_set_wave_count(mse_engine* e, int n_waves) {
    return run_guarded("mse_set_wave_count", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        require_geom(e);
        REQUIRE(n_waves >= 1 && n_waves <= 65535, MSE_ERR_ARG, "wave count out of range");
        HIPX(hipStreamSynchronize(e->stream));
        size_wave_buffers(e, n_waves);
        {  // a new wave set starts with identity keys
            std::vector<unsigned> keys(n_waves);
            for (int i = 0; i < n_waves; ++i) keys[i] = (unsigned)i;
            HIPX(hipMemcpyAsync(e->mix_key.p, keys.data(), n_waves * sizeof(unsigned), hipMemcpyHostToDevice,
                                e->stream));
        }
        if ((int)e->h_wave_cfg.size() != n_waves) {
            e->h_wave_cfg.assign(n_waves, 0);
            HIPX(hipMemcpyAsync(e->wave_cfg.p, e->h_wave_cfg.data(), n_waves * sizeof(int), hipMemcpyHostToDevice,
                                e->stream));
        }
        e->n_waves = n_waves;
        HIPX(hipStreamSynchronize(e->stream));
        sync_wave_prop(e);
    });
}

int mse_set_wave_configs(mse_engine* e, const int* configs, int n_waves) {
    return run_guarded("mse_set_wave_configs", e, [&]() {
        REQUIRE(e && configs, MSE_ERR_ARG, "NULL argument");
        REQUIRE(n_waves == e->n_waves, MSE_ERR_ARG, "n_waves must equal the wave count");
        for (int i = 0; i < n_waves; ++i)
            REQUIRE(configs[i] >= 0 && configs[i] < (int)e->configs.size() && e->configs[configs[i]].trans.p,
                    MSE_ERR_STATE, "wave refers to a configuration without transmission functions");
        e->h_wave_cfg.assign(configs, configs + n_waves);
        HIPX(hipMemcpyAsync(e->wave_cfg.p, e->h_wave_cfg.data(), n_waves * sizeof(int), hipMemcpyHostToDevice,
                            e->stream));
        HIPX(hipStreamSynchronize(e->stream));
        sync_wave_prop(e);
    });
}

}  // extern "C"

namespace mse_rt {
__global__ void k_fill_plane(float2* __restrict__ w, int res, size_t N, int n_waves, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * (size_t)n_waves) return;
    const size_t p = i % N;
    // S = F_x psi of a constant wave: value * res at kx = 0, zero elsewhere
    w[i] = make_float2((p % res) == 0 ? v : 0.f, 0.f);
}

// bilinear_translate (kernels/bilinear_translate_d.cl:21-72) of each wave's fft-shifted
// diffraction image by (ix + sx, iy + sy) pixels: the pixel moved from (x - ix, y - iy),
// wrapped once (an image moved further is zero there), blended with its left, upper and
// upper-left neighbours (wrapped).  sh holds on, ix, iy, sx, sy per wave; on = 0 copies.
__global__ void k_translate_images(const double* __restrict__ in, double* __restrict__ out, int res, size_t N,
                                   const double* __restrict__ sh) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int w = blockIdx.y;
    const double* img = in + (size_t)w * N;
    const double* s = sh + 5 * (size_t)w;
    double v;
    if (s[0] == 0.0) {
        v = img[p];
    } else {
        const int W = res, H = res;
        int nx = (int)(p % W) - (int)s[1];
        int ny = (int)(p / W) - (int)s[2];
        if (nx < 0) nx += W; else if (nx >= W) nx -= W;
        if (ny < 0) ny += H; else if (ny >= H) ny -= H;
        if (nx < 0 || nx >= W || ny < 0 || ny >= H) {
            v = 0.0;
        } else {
            const int xl = nx == 0 ? W - 1 : nx - 1;
            const int yu = ny == 0 ? H - 1 : ny - 1;
            const double sx = s[3], sy = s[4];
            v = (1.0 - sx) * (1.0 - sy) * img[(size_t)ny * W + nx] + sx * (1.0 - sy) * img[(size_t)ny * W + xl] +
                (1.0 - sx) * sy * img[(size_t)yu * W + nx] + sx * sy * img[(size_t)yu * W + xl];
        }
    }
    out[(size_t)w * N + p] = v;
}

// Each wave's image summed over each detector's bitmask in double (band_pass + sum_reduction
// of getStemPixel, simulationstem.cpp:100-131): out[w * ndet + d] = (or +=) scale * sum.
__global__ __launch_bounds__(256) void k_masked_sums(const double* __restrict__ img, size_t N,
                                                     const unsigned* __restrict__ bits, int ndet, double scale,
                                                     int accumulate, double* __restrict__ out) {
    const int d = blockIdx.x, w = blockIdx.y;
    const size_t words = N / 32;
    const unsigned* b = bits + (size_t)d * words;
    const double* im = img + (size_t)w * N;
    double acc = 0.0;
    for (size_t i = threadIdx.x; i < words; i += 256) {
        unsigned m = b[i];
        while (m) {
            acc += im[i * 32 + __builtin_ctz(m)];
            m &= m - 1;
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t o = (size_t)w * ndet + d;
        const double r = scale * red[0];
        out[o] = accumulate ? out[o] + r : r;
    }
}

// translateDiffImage's shift (simulationgeneral.cpp:832-854) of waves [w0, w0 + nw): the
// difference of the wave's beam vector from the geometry's over inverseScale, split into
// whole pixels and a fraction; uploaded to wave_shift.  Returns whether any wave moves.
bool upload_wave_shifts(mse_engine* e, int w0, int nw) {
    std::vector<double> sh(5 * (size_t)nw, 0.0);
    bool any = false;
    for (int i = 0; i < nw; ++i) {
        const int c = e->h_wave_cfg[w0 + i];
        if (c < 0 || c >= (int)e->configs.size() || !e->configs[c].own_beam) continue;
        const double dkx = e->configs[c].k_vec[0] - e->g.k_vec[0];
        const double dky = e->configs[c].k_vec[1] - e->g.k_vec[1];
        if (dkx == 0.0 && dky == 0.0) continue;
        const double shx = dkx / e->g.inverse_scale, shy = dky / e->g.inverse_scale;
        const double ix = std::floor(shx), iy = std::floor(shy);
        REQUIRE(std::fabs(ix) < 1e9 && std::fabs(iy) < 1e9, MSE_ERR_ARG, "diffraction translation out of range");
        double* o = sh.data() + 5 * (size_t)i;
        o[0] = 1.0;
        o[1] = ix;
        o[2] = iy;
        o[3] = shx - ix;
        o[4] = shy - iy;
        any = true;
    }
    if (!any) return false;
    e->wave_shift.ensure(sh.size());
    HIPX(hipMemcpyAsync(e->wave_shift.p, sh.data(), sh.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HIPX(hipStreamSynchronize(e->stream));
    return true;
}

// Diffraction images of waves [w0, w0 + nw) (COL_FFT_DIFF into diff_img), translated into
// diff_moved where their beam has moved; returns the buffer that holds the final images.
const double* diffraction_images(mse_engine* e, int w0, int nw) {
    e->diff_img.ensure((size_t)nw * e->N);
    mse::ColArgs ca = base_col_args(e);
    ca.scale = (float)(1.0 / e->res);
    ca.diff = e->diff_img.p;
    ca.diff_accumulate = 0;
    col_pass(e, mse::COL_FFT_DIFF, e->waves.p + (size_t)w0 * e->N, nullptr, nw, e->N, ca);
    if (!upload_wave_shifts(e, w0, nw)) return e->diff_img.p;
    e->diff_moved.ensure((size_t)nw * e->N);
    launch(e, PK_OTHER, [&]() {
        k_translate_images<<<dim3((unsigned)((e->N + 255) / 256), nw), 256, 0, e->stream>>>(
            e->diff_img.p, e->diff_moved.p, e->res, e->N, e->wave_shift.p);
    });
    return e->diff_moved.p;
}

void check_waves(mse_engine* e) {
    require_geom(e);
    REQUIRE(e->n_waves > 0 && e->waves.p, MSE_ERR_STATE, "mse_set_wave_count has not been called");
}

void check_mix(mse_engine* e) {
    if (e->mix_sets < 2) return;
    for (int c = e->mix_first; c < e->mix_first + e->mix_sets; ++c) {
        REQUIRE(c < (int)e->configs.size(), MSE_ERR_STATE, "mixed potential set " + std::to_string(c) + " does not exist");
        const ConfigData& cd = e->configs[c];
        REQUIRE(!cd.own_geom && !cd.own_beam, MSE_ERR_STATE,
                "mixed potential sets must use the engine's geometry and beam");
        REQUIRE(cd.trans.p && cd.built && cd.trans.n >= (size_t)e->g.n_slices * e->N, MSE_ERR_STATE,
                "transmission functions of mixed potential set " + std::to_string(c) + " are not built");
    }
}

void check_trans(mse_engine* e) {
    check_mix(e);
    for (int i = 0; i < e->n_waves; ++i) {
        const int c = e->h_wave_cfg[i];
        REQUIRE(c >= 0 && c < (int)e->configs.size(), MSE_ERR_STATE, "wave refers to an unknown configuration");
        const ConfigData& cd = e->configs[c];
        REQUIRE(cd.trans.p && cd.built && cd.trans.n >= (size_t)cfg_geom(e, cd).n_slices * e->N, MSE_ERR_STATE,
                "transmission functions of configuration " + std::to_string(c) +
                    " are not built for the current geometry");
    }
}

mse::ProbeArgs probe_args(mse_engine* e, const mse_probe* p) {
    mse::ProbeArgs a{};
    a.pos = e->probe_pos.p;
    a.wavelength = e->g.wavelength;
    a.cond_ap2 = (p->cond_ap_mrad * 0.001) / e->g.wavelength;
    a.smooth_r = (p->smooth_mrad * 0.001) / e->g.wavelength;
    a.C10 = p->C10;
    a.C30 = p->C30;
    a.C50 = p->C50;
    a.C12 = make_double2(p->C12[0], p->C12[1]);
    a.C21 = make_double2(p->C21[0], p->C21[1]);
    a.C23 = make_double2(p->C23[0], p->C23[1]);
    a.C32 = make_double2(p->C32[0], p->C32[1]);
    a.C34 = make_double2(p->C34[0], p->C34[1]);
    a.C41 = make_double2(p->C41[0], p->C41[1]);
    a.C43 = make_double2(p->C43[0], p->C43[1]);
    a.C45 = make_double2(p->C45[0], p->C45[1]);
    a.C52 = make_double2(p->C52[0], p->C52[1]);
    a.C54 = make_double2(p->C54[0], p->C54[1]);
    a.C56 = make_double2(p->C56[0], p->C56[1]);
    a.taper = p->edge_taper;
    a.dC10 = e->h_focus.empty() ? nullptr : e->probe_df.p;
    return a;
}

// positions of n probes (relative to the padded start) into waves first_wave.., and, when
// focus changes are set, those of probes first_probe.. of the caller's list
void upload_positions(mse_engine* e, const double* pos_xy, int n, int first_wave = 0, int first_probe = 0) {
    std::vector<double> rel(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        rel[2 * i] = pos_xy[2 * i] - e->g.sim_x[0];  // simulationcbed.cpp:54-59
        rel[2 * i + 1] = pos_xy[2 * i + 1] - e->g.sim_y[0];
    }
    HIPX(hipMemcpyAsync(e->probe_pos.p + 2 * (size_t)first_wave, rel.data(), rel.size() * sizeof(double),
                        hipMemcpyHostToDevice, e->stream));
    // keys of the mixed-potential draws: the probe indices
    REQUIRE(e->mix_key.p && e->mix_key.n >= (size_t)first_wave + n, MSE_ERR_STATE, "wave buffers are not sized");
    std::vector<unsigned> keys(n);
    for (int i = 0; i < n; ++i) keys[i] = (unsigned)(first_probe + i);
    HIPX(hipMemcpyAsync(e->mix_key.p + first_wave, keys.data(), n * sizeof(unsigned), hipMemcpyHostToDevice,
                        e->stream));
    if (!e->h_focus.empty()) {
        REQUIRE((size_t)first_probe + n <= e->h_focus.size(), MSE_ERR_ARG,
                "fewer focus changes than probes (mse_set_focus_spread)");
        if (e->probe_df.n < (size_t)first_wave + n) {
            HIPX(hipStreamSynchronize(e->stream));
            e->probe_df.ensure(std::max<size_t>((size_t)first_wave + n, e->n_waves));
        }
        HIPX(hipMemcpyAsync(e->probe_df.p + first_wave, e->h_focus.data() + first_probe, n * sizeof(double),
                            hipMemcpyHostToDevice, e->stream));
    }
}

void probes_into_waves(mse_engine* e, int n, const mse_probe* p) {
    mse::ColArgs ca = base_col_args(e);
    ca.probe = probe_args(e, p);
    col_pass(e, mse::COL_PROBE, e->waves.p, e->waves.p, n, e->N, ca, PK_PROBE);
}

// Waves whose configuration has fewer slices than the engine's geometry (thinner
// specimens batched with thicker ones) stop at their own exit surface, as separate clTEM
// runs would.  For each slice, the active waves are those whose configuration still has
// that slice; when they are the first k waves (waves ordered by thickness, or all waves)
// the passes run on k sets directly, otherwise on a device list of their indices.
void active_waves(mse_engine* e, int nw) {
    std::vector<int> key(e->h_wave_cfg.begin(), e->h_wave_cfg.begin() + nw);
    std::vector<int> ns(e->configs.size());
    // under mixed static potentials every wave reads sets of the engine's thickness
    for (size_t c = 0; c < e->configs.size(); ++c)
        ns[c] = e->mix_sets >= 2 ? e->g.n_slices : cfg_geom(e, e->configs[c]).n_slices;
    key.push_back(-1);
    key.insert(key.end(), ns.begin(), ns.end());
    key.push_back(e->g.n_slices);
    if (key == e->map_key && !e->map_count.empty()) return;
    const int S = e->g.n_slices;
    e->map_count.assign(S, 0);
    e->map_offset.assign(S, -1);
    std::vector<int> table;
    for (int s = 0; s < S; ++s) {
        std::vector<int> act;
        for (int w = 0; w < nw; ++w)
            if (ns[e->h_wave_cfg[w]] > s) act.push_back(w);
        e->map_count[s] = (int)act.size();
        bool prefix = true;
        for (int i = 0; i < (int)act.size(); ++i) prefix = prefix && act[i] == i;
        if (!prefix) {
            e->map_offset[s] = (int)table.size();
            table.insert(table.end(), act.begin(), act.end());
        }
    }
    if (!table.empty()) {
        if (e->wave_map.n < table.size()) {
            HIPX(hipStreamSynchronize(e->stream));
            e->wave_map.ensure(table.size());
        }
        HIPX(hipMemcpyAsync(e->wave_map.p, table.data(), table.size() * sizeof(int), hipMemcpyHostToDevice,
                            e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    }
    e->map_key = key;
}

void require_elements(const mse_engine* e, const ConfigData& c) {
    // the potential kernels read per-element rows of the parameterisation (and its tables)
    // by atomic number: a configuration sorted under a larger parameterisation must not reach
    // past the current one
    for (size_t Z = (size_t)std::max(e->max_Z, 0) + 1; Z < c.has_Z.size(); ++Z)
        REQUIRE(!c.has_Z[Z], MSE_ERR_STATE,
                "the configuration holds atomic numbers beyond the current parameterisation "
                "(sort its atoms again after mse_set_parameterisation)");
}

const float4* potential_tables(mse_engine* e, const ConfigData& c, const mse_geometry& g) {
    if (!e->use_vtab || e->max_Z <= 0 || e->h_params.empty()) return nullptr;
    // the distances the kernels evaluate must lie inside the tables: r >= r_min = px / 4, and
    // r^2 <= 64 + (depth reach)^2 with the full-3D depth below (load_z + 1) slices
    const double r_min = 0.25 * g.real_scale;
    // (with a margin, so that r_min^2 formed in float is above the floor too)
    if (r_min * r_min < std::ldexp(1.0, mse_tab::kOctaveLo) * (1.0 + 1e-6)) return nullptr;
    const double reach = (g.load_z + 2) * g.slice_thickness;
    if (g.full3d && 64.0 + reach * reach >= std::ldexp(1.0, mse_tab::kOctaveHi)) return nullptr;
    const int t = g.full3d ? 1 : 0;
    const size_t per = (size_t)mse_tab::kBuckets * 4;
    for (int Z = 1; Z <= e->max_Z && Z < (int)c.has_Z.size(); ++Z) {
        if (!c.has_Z[Z] || e->vtab_on[(size_t)t * e->max_Z + Z - 1]) continue;
        float* h = e->h_vtab.data() + ((size_t)t * e->max_Z + Z - 1) * per;
        mse_tab::build(e->form, g.full3d != 0, e->h_params.data(), e->i_count, Z, h);
        // the kernels' form of the cubics: in the integer m = t 2^kMantBits (exact rescaling)
        for (int b = 0; b < mse_tab::kBuckets; ++b)
            for (int q = 1; q < 4; ++q) h[4 * b + q] = std::ldexp(h[4 * b + q], -q * mse_tab::kMantBits);
        HIPX(hipMemcpyAsync(e->vtab.p + ((size_t)t * e->max_Z + Z - 1) * mse_tab::kBuckets, h,
                            per * sizeof(float), hipMemcpyHostToDevice, e->stream));
        e->vtab_on[(size_t)t * e->max_Z + Z - 1] = 1;
    }
    return e->vtab.p + (size_t)t * e->max_Z * mse_tab::kBuckets;
}

// The full-3D kernels' LDS copy of the tables (PotArgs::vslot .. vsel_at): the elements of
// the configuration and the buckets an untilted slice reaches, from r_min^2 as the kernels
// round it (at least the tables' floor) to 64 + ((load_z + 2) dz)^2.  Returns the bytes of
// LDS it takes, or 0 where the global tables are used instead: a tilted beam, or tables over
// 32 KB.  (The projected kernels, one evaluation per atom and pixel, read the global tables:
// an LDS copy per block costs them more than it saves.)
static size_t lds_tables(mse_engine* e, const ConfigData& c, const mse_geometry& g, mse::PotArgs& a) {
    if (!g.full3d || a.shift_x != 0.0 || a.shift_y != 0.0) return 0;
    const float r_min = (float)(0.25 * a.pixelscale);
    const float r_lo2 = std::max(r_min * r_min, mse::kVtabFloor);
    const double reach = (g.load_z + 2) * g.slice_thickness;
    const float r_hi2 = (float)((64.0 + reach * reach) * 1.001);
    const auto bucket = [](float r2) {
        unsigned u;
        std::memcpy(&u, &r2, sizeof(u));
        return (int)(u >> mse_tab::kMantBits);
    };
    const int top = ((127 + mse_tab::kOctaveHi) << mse_tab::kSubBits) - 1;
    const int lo = bucket(r_lo2), hi = std::min(bucket(r_hi2), top);
    if (hi < lo) return 0;
    const int n_z = (int)c.has_Z.size();
    std::vector<int> hv((size_t)n_z, 0);
    int n_el = 0;
    for (int Z = 1; Z < n_z; ++Z) {
        if (!c.has_Z[Z]) continue;
        if (Z > e->max_Z) return 0;  // no table for it
        hv[Z] = n_el++;
    }
    for (int Z = 1; Z < n_z; ++Z)
        if (c.has_Z[Z]) hv.push_back(Z);
    const size_t bytes = (size_t)n_el * (hi - lo + 1) * sizeof(float4);
    if (n_el == 0 || bytes > 32768) return 0;
    if (hv != e->h_vslot) {
        // the previous contents may still be read or uploaded
        HIPX(hipStreamSynchronize(e->stream));
        e->vslot.ensure(hv.size());
        e->h_vslot = hv;
        HIPX(hipMemcpyAsync(e->vslot.p, e->h_vslot.data(), hv.size() * sizeof(int), hipMemcpyHostToDevice,
                            e->stream));
    }
    a.vslot = e->vslot.p;
    a.v_nel = n_el;
    a.vb_lo = lo;
    a.vb_n = hi - lo + 1;
    a.vsel_at = n_z;
    return bytes;
}

void potential_planes(mse_engine* e, const ConfigData& c, const mse_geometry& g, const mse::PotArgs& args,
                      float2* out, int planes, int first) {
    require_elements(e, c);
    const dim3 grid(e->res / 16, e->res / 16, planes);
    const dim3 grid_proj(e->res / (16 * mse::kProjPX), e->res / (16 * mse::kProjPY), planes);
    const dim3 grid_lds(e->res / (16 * mse::kFull3dNX), e->res / 16, planes);
    mse::PotArgs a = args;
    a.vtab = potential_tables(e, c, g);
    const bool tab = a.vtab != nullptr;
    const size_t lds = tab && e->vtab_lds ? lds_tables(e, c, g, a) : 0;
    launch(e, PK_POT, [&]() {
        if (g.full3d && lds) {
            if (e->form == 0)
                mse::k_potential_full3d<0, true, true><<<grid_lds, 256, lds, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first);
            else if (e->form == 1)
                mse::k_potential_full3d<1, true, true><<<grid_lds, 256, lds, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first);
            else
                mse::k_potential_full3d<2, true, true><<<grid_lds, 256, lds, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first);
        } else if (g.full3d) {
            if (e->form == 0)
                tab ? mse::k_potential_full3d<0, true><<<grid, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first)
                    : mse::k_potential_full3d<0><<<grid, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first);
            else if (e->form == 1)
                tab ? mse::k_potential_full3d<1, true><<<grid, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first)
                    : mse::k_potential_full3d<1><<<grid, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first);
            else
                tab ? mse::k_potential_full3d<2, true><<<grid, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first)
                    : mse::k_potential_full3d<2><<<grid, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, a, first);
        } else {
            if (e->form == 0)
                tab ? mse::k_potential_projected<0, true><<<grid_proj, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, a, first)
                    : mse::k_potential_projected<0><<<grid_proj, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, a, first);
            else if (e->form == 1)
                tab ? mse::k_potential_projected<1, true><<<grid_proj, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, a, first)
                    : mse::k_potential_projected<1><<<grid_proj, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, a, first);
            else
                tab ? mse::k_potential_projected<2, true><<<grid_proj, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, a, first)
                    : mse::k_potential_projected<2><<<grid_proj, 256, 0, e->stream>>>(out, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, a, first);
        }
    });
}

// band limit: FFT -> keep |k| <= kmax*2/3 -> IFFT (simulationgeneral.cpp:596-602).  The
// column pass is a propagation step whose factors are the constant scale (x) and one (y),
// with the band's per-row limits as the mask: cmul(v, cmul((s, 0), (1, 0))) = cmul(v, (s, 0))
// bit for bit, on the persistent column kernels.  Their variant is fixed per grid size (not
// tuned), so that transmission functions do not depend on the tuning state.
void band_limit_planes(mse_engine* e, float2* p, int planes) {
    mse::RowArgs ra = base_row_args(e);
    ra.scale = 1.0f;
    row_pass(e, mse::ROW_FFT, p, p, planes, e->N, ra);
    mse::ColArgs ca = base_col_args(e);
    ca.wave_map = nullptr;
    ca.wave_prop = nullptr;
    ca.px = e->band_xy.p;
    ca.py = e->band_xy.p + e->res;
    ca.lim = e->band_lim.p;
    const int variant = e->res == 1024 ? 12 : e->res == 768 ? 11 : 7;
    col_pass(e, mse::COL_PROP, p, p, planes, e->N, ca, PK_COL, variant);
    row_pass(e, mse::ROW_IFFT, p, p, planes, e->N, ra);
}

void slice_step(mse_engine* e, float2* waves, int nw, int slice) {
    mse::RowArgs ra = base_row_args(e);
    ra.slice = slice;
    row_pass(e, mse::ROW_IFFT_T_FFT, waves, waves, nw, e->N, ra);
    col_pass(e, mse::COL_PROP, waves, waves, nw, e->N, base_col_args(e));
}

void multislice(mse_engine* e, int nw, int first, int count, bool last_detect, const mse::ColArgs* det_args) {
    active_waves(e, nw);
    mse::RowArgs ra = base_row_args(e);
    mse::ColArgs ca = base_col_args(e);
    for (int s = first; s < first + count; ++s) {
        const int na = e->map_count[s];
        const int* map = e->map_offset[s] >= 0 ? e->wave_map.p + e->map_offset[s] : nullptr;
        if (last_detect && s == first + count - 1)
            REQUIRE(na == nw && !map, MSE_ERR_STATE, "detector slice reached by finished waves");
        if (na == 0) continue;
        ra.slice = s;
        ra.wave_map = map;
        ca.wave_map = map;
        row_pass(e, mse::ROW_IFFT_T_FFT, e->waves.p, e->waves.p, na, e->N, ra);
        if (last_detect && s == first + count - 1)
            e->det_tiles = col_pass(e, mse::COL_PROP_DETECT, e->waves.p, e->waves.p, nw, e->N, *det_args);
        else
            col_pass(e, mse::COL_PROP, e->waves.p, e->waves.p, na, e->N, ca);
    }
}

void reduce_detectors(mse_engine* e, int nw, int n_det, int sqrt2, double* dev_out, int accumulate,
                      const long long* dev_offset = nullptr) {
    const int nwd = nw * n_det;
    const int ntiles = e->det_tiles;
    launch(e, PK_OTHER, [&]() {
        mse::k_reduce_partials<<<(nwd + 255) / 256, 256, 0, e->stream>>>(e->partials.p, ntiles, nwd, dev_out,
                                                                          accumulate, sqrt2 ? M_SQRT2 : 1.0,
                                                                          dev_offset);
    });
}

// Captured STEM batch: probe init, every slice and the fused detector reduction of one batch
// of `nb` probes, replayed per batch.  Positions and the output offset live in device
// memory and are updated (outside the graph) before each replay.
struct StemGraphKey {
    int res, cfg, nb, n_det, sqrt2, row_var, col_var, n_slices, mix_first, mix_sets;
    unsigned tune_version;
    unsigned long long mix_seed;
    const void *waves, *partials, *det_out, *det_bits, *trans, *probe_pos, *probe_df, *mix_key;
    const void *cfg_ptrs, *cfg_nslices, *wave_cfg, *wave_map, *tw, *px, *py, *wave_prop, *lim, *kx, *ky;
    mse_probe probe;
};
}  // namespace mse_rt

extern "C" {

int mse_init_plane_wave(mse_engine* e, double value) {
    return run_guarded("mse_init_plane_wave", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        check_waves(e);
        const size_t total = e->N * e->n_waves;
        const float v = (float)(value * e->res);
        launch(e, PK_OTHER, [&]() {
            k_fill_plane<<<(unsigned)((total + 255) / 256), 256, 0, e->stream>>>(e->waves.p, e->res, e->N, e->n_waves, v);
        });
    });
}

int mse_init_probes(mse_engine* e, const double* pos_xy, int n, const mse_probe* probe) {
    return run_guarded("mse_init_probes", e, [&]() {
        REQUIRE(e && pos_xy && probe, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        REQUIRE(n >= 1 && n <= e->n_waves, MSE_ERR_ARG, "probe count must be 1..wave count");
        upload_positions(e, pos_xy, n);
        probes_into_waves(e, n, probe);
    });
}

int mse_multislice(mse_engine* e, int first, int count) {
    return run_guarded("mse_multislice", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        check_waves(e);
        check_trans(e);
        REQUIRE(first >= 0 && count >= 0 && first + count <= e->g.n_slices, MSE_ERR_ARG, "slice range out of bounds");
        multislice(e, e->n_waves, first, count, false, nullptr);
    });
}

// Everything a captured slice loop depends on: launch shapes, kernel variants and every
// device pointer baked into its kernel arguments.
struct LoopGraphKey {
    int res, nw, first, count, row_var, col_var, num_cus, mix_first, mix_sets;
    unsigned tune_version;
    unsigned long long mix_seed;
    const void *waves, *cfg_ptrs, *cfg_nslices, *wave_cfg, *tw, *px, *py, *wave_prop, *lim, *mix_key;
};

}  // extern "C"

namespace mse_rt {
// The key of a captured slice loop over [first, first + count) for every wave; plans the
// per-slice wave sets (host work and uploads that must happen before a capture).
std::vector<unsigned char> loop_graph_key(mse_engine* e, int first, int count) {
    LoopGraphKey k;
    std::memset(&k, 0, sizeof(k));
    k.res = e->res;
    k.nw = e->n_waves;
    k.first = first;
    k.count = count;
    k.row_var = e->row_var;
    k.col_var = e->col_var;
    k.tune_version = e->tune_version;
    k.num_cus = e->num_cus;
    k.waves = e->waves.p;
    k.cfg_ptrs = e->cfg_ptrs.p;
    k.cfg_nslices = e->cfg_nslices.p;
    k.wave_cfg = e->wave_cfg.p;
    k.tw = e->tw.p;
    k.px = e->px.p;
    k.py = e->py.p;
    k.wave_prop = e->any_beam ? e->wave_prop.p : nullptr;
    k.mix_first = e->mix_first;
    k.mix_sets = e->mix_sets;
    k.mix_seed = e->mix_seed;
    k.mix_key = e->mix_key.p;
    k.lim = e->prop_lim.p;
    std::vector<unsigned char> key((unsigned char*)&k, (unsigned char*)&k + sizeof(k));
    active_waves(e, e->n_waves);  // the per-slice wave sets are baked into the graph
    const void* mp = e->wave_map.p;
    key.insert(key.end(), (const unsigned char*)&mp, (const unsigned char*)&mp + sizeof(mp));
    key.insert(key.end(), (const unsigned char*)e->map_key.data(),
               (const unsigned char*)(e->map_key.data() + e->map_key.size()));
    return key;
}
}  // namespace mse_rt

extern "C" {

int mse_multislice_graph(mse_engine* e, int first, int count) {
    return run_guarded("mse_multislice_graph", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        check_waves(e);
        check_trans(e);
        REQUIRE(first >= 0 && count >= 0 && first + count <= e->g.n_slices, MSE_ERR_ARG, "slice range out of bounds");
        if (count == 0) return;
        if (e->prof) {  // per-kernel events need eager launches
            multislice(e, e->n_waves, first, count, false, nullptr);
            return;
        }
        const std::vector<unsigned char> key = loop_graph_key(e, first, count);
        if (!e->loop_exec || key != e->loop_key) {
            if (e->loop_exec) (void)hipGraphExecDestroy(e->loop_exec);
            e->loop_exec = nullptr;
            e->loop_key.clear();
            hipGraph_t gr = nullptr;
            HIPX(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            try {
                multislice(e, e->n_waves, first, count, false, nullptr);
            } catch (...) {
                (void)hipStreamEndCapture(e->stream, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            HIPX(hipStreamEndCapture(e->stream, &gr));
            hipError_t st = hipGraphInstantiate(&e->loop_exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (st != hipSuccess) {
                e->loop_exec = nullptr;
                throw Error(MSE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(st));
            }
            e->loop_key = key;
        }
        HIPX(hipGraphLaunch(e->loop_exec, e->stream));
    });
}

}  // extern "C"

// Engines of one device whose slice loops run as one graph (mixed grid sizes).
struct mse_group {
    std::vector<mse_engine*> engines;
    int device = 0;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> join, pre;  // per engine: capture join, launch-time dependency
};

namespace mse_rt {
void destroy_group_events(mse_group* g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    g->exec = nullptr;
    for (hipEvent_t ev : g->join) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : g->pre) (void)hipEventDestroy(ev);
    if (g->fork) (void)hipEventDestroy(g->fork);
    if (g->done) (void)hipEventDestroy(g->done);
    g->join.clear();
    g->pre.clear();
    g->fork = g->done = nullptr;
}
}  // namespace mse_rt

extern "C" {

int mse_group_create(mse_engine* const* engines, int n, mse_group** out) {
    return run_guarded("mse_group_create", nullptr, [&]() {
        REQUIRE(engines && out && n >= 1, MSE_ERR_ARG, "NULL argument or empty group");
        for (int i = 0; i < n; ++i) {
            REQUIRE(engines[i], MSE_ERR_ARG, "NULL engine in the group");
            REQUIRE(engines[i]->device == engines[0]->device, MSE_ERR_ARG, "a group's engines share one device");
            for (int j = 0; j < i; ++j) REQUIRE(engines[j] != engines[i], MSE_ERR_ARG, "an engine appears twice");
        }
        auto* g = new mse_group();
        g->engines.assign(engines, engines + n);
        g->device = engines[0]->device;
        try {
            DeviceGuard dg(g->device);
            HIPX(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
            HIPX(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
            for (int i = 0; i < n; ++i) {
                hipEvent_t a = nullptr, b = nullptr;
                HIPX(hipEventCreateWithFlags(&a, hipEventDisableTiming));
                g->join.push_back(a);
                HIPX(hipEventCreateWithFlags(&b, hipEventDisableTiming));
                g->pre.push_back(b);
            }
        } catch (...) {
            destroy_group_events(g);
            delete g;
            throw;
        }
        *out = g;
    });
}

int mse_group_destroy(mse_group* g) {
    if (!g) return MSE_OK;
    return run_guarded("mse_group_destroy", nullptr, [&]() {
        DeviceGuard dg(g->device);
        (void)hipStreamSynchronize(g->engines[0]->stream);
        destroy_group_events(g);
        delete g;
    });
}

int mse_group_multislice(mse_group* g) {
    return run_guarded("mse_group_multislice", nullptr, [&]() {
        REQUIRE(g, MSE_ERR_ARG, "group is NULL");
        DeviceGuard dg(g->device);
        const int n = (int)g->engines.size();
        std::vector<unsigned char> key;
        bool eager = false;
        for (mse_engine* e : g->engines) {
            check_waves(e);
            check_trans(e);
            eager = eager || e->prof;
            const std::vector<unsigned char> k = loop_graph_key(e, 0, e->g.n_slices);
            key.insert(key.end(), (const unsigned char*)&e, (const unsigned char*)&e + sizeof(e));
            key.insert(key.end(), k.begin(), k.end());
        }
        if (eager) {  // per-kernel profiling events need eager launches on each engine's stream
            for (mse_engine* e : g->engines) multislice(e, e->n_waves, 0, e->g.n_slices, false, nullptr);
            return;
        }
        hipStream_t s0 = g->engines[0]->stream;
        if (!g->exec || key != g->key) {
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
            g->exec = nullptr;
            g->key.clear();
            hipGraph_t gr = nullptr;
            HIPX(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
            try {
                // every other engine's stream joins the capture, runs its own slice loop as a
                // branch of the graph, and joins back into the first stream
                HIPX(hipEventRecord(g->fork, s0));
                for (int i = 1; i < n; ++i) HIPX(hipStreamWaitEvent(g->engines[i]->stream, g->fork, 0));
                for (mse_engine* e : g->engines) multislice(e, e->n_waves, 0, e->g.n_slices, false, nullptr);
                for (int i = 1; i < n; ++i) {
                    HIPX(hipEventRecord(g->join[i], g->engines[i]->stream));
                    HIPX(hipStreamWaitEvent(s0, g->join[i], 0));
                }
            } catch (...) {  // leave the streams out of capture mode for later calls
                (void)hipStreamEndCapture(s0, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            HIPX(hipStreamEndCapture(s0, &gr));
            hipError_t st = hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (st != hipSuccess) {
                g->exec = nullptr;
                throw Error(MSE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(st));
            }
            g->key = key;
        }
        // the graph starts after the work queued on every engine (probe set-ups) and every
        // engine's later work waits for it
        for (int i = 1; i < n; ++i) {
            HIPX(hipEventRecord(g->pre[i], g->engines[i]->stream));
            HIPX(hipStreamWaitEvent(s0, g->pre[i], 0));
        }
        HIPX(hipGraphLaunch(g->exec, s0));
        HIPX(hipEventRecord(g->done, s0));
        for (int i = 1; i < n; ++i) HIPX(hipStreamWaitEvent(g->engines[i]->stream, g->done, 0));
    });
}

int mse_stem_detect(mse_engine* e, const mse_detector* dets, int n_det, int sqrt2, double* out) {
    return run_guarded("mse_stem_detect", e, [&]() {
        REQUIRE(e && dets && out, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        ensure_detectors(e, dets, n_det);
        const int nw = e->n_waves;
        e->det_out.ensure((size_t)nw * n_det);
        if (e->any_beam) {
            // plasmon-tilted waves: detectors over the translated diffraction images
            // (getStemPixel with d_kx, d_ky: simulationstem.cpp:80-131)
            const double* im = diffraction_images(e, 0, nw);
            if (n_det > 0)
                launch(e, PK_OTHER, [&]() {
                    k_masked_sums<<<dim3(n_det, nw), 256, 0, e->stream>>>(im, e->N, e->det_bits.p, n_det,
                                                                         sqrt2 ? std::sqrt(2.0) : 1.0, 0,
                                                                         e->det_out.p);
                });
        } else {
            e->partials.ensure((size_t)nw * n_det * (e->res / 8));
            mse::ColArgs ca = base_col_args(e);
            ca.scale = (float)(1.0 / e->res);
            ca.det_scale = 1.0;
            ca.det_bits = e->det_bits.p;
            ca.ndet = n_det;
            ca.partials = e->partials.p;
            e->det_tiles = col_pass(e, mse::COL_FFT_DETECT, e->waves.p, e->waves.p, nw, e->N, ca);
            reduce_detectors(e, nw, n_det, sqrt2, e->det_out.p, 0);
        }
        HIPX(hipMemcpyAsync(out, e->det_out.p, (size_t)nw * n_det * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_stem_detect_noalloc(mse_engine* e, const mse_detector* dets, int n_det, int sqrt2, double* out_dev) {
    return run_guarded("mse_stem_detect_noalloc", e, [&]() {
        REQUIRE(e && dets && out_dev, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        REQUIRE(n_det >= 1 && n_det <= mse::kMaxDet, MSE_ERR_ARG, "1 to 8 detectors");
        REQUIRE(!e->any_beam, MSE_ERR_UNSUPPORTED,
                "waves of a configuration with its own beam need translated diffraction images; use mse_stem_detect");
        const int nw = e->n_waves;
        const size_t words = e->N / 32;
        const size_t mask_bytes = (n_det * words * sizeof(unsigned) + 255) / 256 * 256;
        // at most res/8 column tiles per wave and detector (8-column tiles)
        const size_t partial_bytes = (size_t)nw * n_det * (e->res / 8) * sizeof(double);
        REQUIRE(mask_bytes + partial_bytes <= kNoAllocScratchBytes, MSE_ERR_ARG,
                "workspace of " + std::to_string(mask_bytes + partial_bytes) + " bytes (" + std::to_string(n_det) +
                    " detectors at " + std::to_string(e->res) + "^2, " + std::to_string(nw) +
                    " waves) exceeds the fixed " + std::to_string(kNoAllocScratchBytes) +
                    " bytes: n_det*res^2/8 + n_waves*n_det*res must fit");
        auto* base = static_cast<unsigned char*>(noalloc_scratch(e->device));
        auto* bits = reinterpret_cast<unsigned*>(base);
        auto* partials = reinterpret_cast<double*>(base + mask_bytes);
        const DetectorRadii radii = detector_radii(e, dets, n_det);
        launch(e, PK_OTHER, [&]() {
            k_detector_bits<<<dim3((unsigned)((words + 255) / 256), n_det), 256, 0, e->stream>>>(bits, e->res, radii);
        });
        mse::ColArgs ca = base_col_args(e);
        ca.scale = (float)(1.0 / e->res);
        ca.det_scale = 1.0;
        ca.det_bits = bits;
        ca.ndet = n_det;
        ca.partials = partials;
        // the detector pass runs the default plan on 16- or 8-column tiles: at most res/8 tiles
        const int ntiles = col_pass(e, mse::COL_FFT_DETECT, e->waves.p, e->waves.p, nw, e->N, ca);
        const int nwd = nw * n_det;
        launch(e, PK_OTHER, [&]() {
            mse::k_reduce_partials<<<(nwd + 255) / 256, 256, 0, e->stream>>>(partials, ntiles, nwd, out_dev, 0,
                                                                              sqrt2 ? M_SQRT2 : 1.0, nullptr);
        });
    });
}

int mse_diffraction(mse_engine* e, int wave, double* out) {
    return run_guarded("mse_diffraction", e, [&]() {
        REQUIRE(e && out, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        REQUIRE(wave >= 0 && wave < e->n_waves, MSE_ERR_ARG, "wave index out of range");
        const double* im = diffraction_images(e, wave, 1);  // translated when the beam moved
        HIPX(hipMemcpyAsync(out, im, e->N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_exit_wave(mse_engine* e, int wave, double* out) {
    return run_guarded("mse_exit_wave", e, [&]() {
        REQUIRE(e && out, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        REQUIRE(wave >= 0 && wave < e->n_waves, MSE_ERR_ARG, "wave index out of range");
        e->scratch.ensure(e->N);
        mse::RowArgs ra = base_row_args(e);
        ra.scale = (float)(1.0 / e->res);
        row_pass(e, mse::ROW_IFFT, e->waves.p + (size_t)wave * e->N, e->scratch.p, 1, e->N, ra);
        std::vector<float2> h(e->N);
        HIPX(hipMemcpyAsync(h.data(), e->scratch.p, e->N * sizeof(float2), hipMemcpyDeviceToHost, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
        for (size_t i = 0; i < e->N; ++i) {
            out[2 * i] = h[i].x;
            out[2 * i + 1] = h[i].y;
        }
    });
}

int mse_stem_batch(mse_engine* e, int cfg, const double* pos_xy, int n_probes, int batch, const mse_probe* probe,
                   const mse_detector* dets, int n_det, int sqrt2, int use_graph, double* accum) {
    return run_guarded("mse_stem_batch", e, [&]() {
        REQUIRE(e && pos_xy && probe && dets && accum, MSE_ERR_ARG, "NULL argument");
        require_geom(e);
        REQUIRE(n_probes >= 1 && batch >= 1 && n_det >= 1, MSE_ERR_ARG, "counts must be positive");
        ConfigData& c = config(e, cfg);
        REQUIRE(c.trans.p && c.built && c.trans.n >= (size_t)cfg_geom(e, c).n_slices * e->N, MSE_ERR_STATE,
                "transmission functions of this configuration are not built for the current geometry");
        check_mix(e);
        REQUIRE(!c.own_beam, MSE_ERR_UNSUPPORTED,
                "a configuration with its own beam vector runs through mse_multislice and mse_stem_detect "
                "(its detector images are translated)");
        if (e->n_waves < batch || !e->waves.p || e->mix_key.n < (size_t)e->n_waves) {
            HIPX(hipStreamSynchronize(e->stream));
            size_wave_buffers(e, std::max(batch, e->n_waves));
            e->n_waves = std::max(batch, e->n_waves);
        }
        // every wave (also those past the batch) refers to the configuration, so the host and
        // device tables keep one entry per wave
        std::vector<int> wc(e->n_waves, cfg);
        if (e->h_wave_cfg != wc) {
            e->h_wave_cfg = wc;
            HIPX(hipMemcpyAsync(e->wave_cfg.p, wc.data(), wc.size() * sizeof(int), hipMemcpyHostToDevice,
                                e->stream));
        }
        sync_wave_prop(e);
        // the probes leave the specimen at the configuration's own exit surface (a thinner
        // specimen than the engine geometry's; mixed static potentials use the engine's slicing)
        const int ns = e->mix_sets >= 2 ? e->g.n_slices : cfg_geom(e, c).n_slices;
        ensure_detectors(e, dets, n_det);
        e->partials.ensure((size_t)batch * n_det * (e->res / 8));
        e->det_out.ensure((size_t)n_probes * n_det);
        HIPX(hipMemcpyAsync(e->det_out.p, accum, (size_t)n_probes * n_det * sizeof(double), hipMemcpyHostToDevice,
                            e->stream));
        mse::ColArgs da = base_col_args(e);
        da.det_scale = 1.0;  // V = P' F_y A is Psi(k) itself (unitary scales folded into P')
        da.det_bits = e->det_bits.p;
        da.ndet = n_det;
        da.partials = e->partials.p;
        const bool graph = use_graph && !e->prof;
        if (graph) {
            e->dev_offset.ensure(1);
            StemGraphKey k;
            std::memset(&k, 0, sizeof(k));
            k.res = e->res;
            k.cfg = cfg;
            k.nb = batch;
            k.n_det = n_det;
            k.sqrt2 = sqrt2;
            k.row_var = e->row_var;
            k.col_var = e->col_var;
            k.tune_version = e->tune_version;
            k.n_slices = ns;
            k.waves = e->waves.p;
            k.partials = e->partials.p;
            k.det_out = e->det_out.p;
            k.det_bits = e->det_bits.p;
            k.trans = c.trans.p;
            k.probe_pos = e->probe_pos.p;
            k.cfg_ptrs = e->cfg_ptrs.p;  // every device pointer the captured kernels read
            k.cfg_nslices = e->cfg_nslices.p;
            k.wave_cfg = e->wave_cfg.p;
            k.wave_map = e->wave_map.p;
            k.tw = e->tw.p;
            k.px = e->px.p;
            k.py = e->py.p;
            k.wave_prop = e->any_beam ? e->wave_prop.p : nullptr;
            k.mix_first = e->mix_first;
            k.mix_sets = e->mix_sets;
            k.mix_seed = e->mix_seed;
            k.mix_key = e->mix_key.p;
            k.lim = e->prop_lim.p;
            k.kx = e->kx.p;
            k.ky = e->ky.p;
            if (!e->h_focus.empty()) {  // the device array the captured probe kernel reads
                if (e->probe_df.n < (size_t)batch) {
                    HIPX(hipStreamSynchronize(e->stream));
                    e->probe_df.ensure(batch);
                }
                k.probe_df = e->probe_df.p;
            }
            k.probe = *probe;
            std::vector<unsigned char> key((unsigned char*)&k, (unsigned char*)&k + sizeof(k));
            active_waves(e, batch);  // host-side plan (no stream work may happen inside the capture)
            if (!e->stem_exec || key != e->stem_key) {
                destroy_stem_graph(e);
                hipGraph_t gr = nullptr;
                HIPX(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                try {
                    probes_into_waves(e, batch, probe);
                    multislice(e, batch, 0, ns, true, &da);
                    reduce_detectors(e, batch, n_det, sqrt2, e->det_out.p, 1, e->dev_offset.p);
                } catch (...) {  // leave the stream out of capture mode for later calls
                    (void)hipStreamEndCapture(e->stream, &gr);
                    if (gr) (void)hipGraphDestroy(gr);
                    throw;
                }
                HIPX(hipStreamEndCapture(e->stream, &gr));
                hipError_t st = hipGraphInstantiate(&e->stem_exec, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (st != hipSuccess) {
                    e->stem_exec = nullptr;
                    throw Error(MSE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(st));
                }
                e->stem_key = key;
            }
        }
        std::vector<long long> offsets;
        offsets.reserve((n_probes + batch - 1) / batch);
        for (int p0 = 0; p0 < n_probes; p0 += batch) offsets.push_back((long long)p0 * n_det);
        for (int p0 = 0, bi = 0; p0 < n_probes; p0 += batch, ++bi) {
            const int nb = std::min(batch, n_probes - p0);
            upload_positions(e, pos_xy + 2 * (size_t)p0, nb, 0, p0);
            if (graph && nb == batch) {
                HIPX(hipMemcpyAsync(e->dev_offset.p, &offsets[bi], sizeof(long long), hipMemcpyHostToDevice,
                                    e->stream));
                HIPX(hipGraphLaunch(e->stem_exec, e->stream));
            } else {
                probes_into_waves(e, nb, probe);
                multislice(e, nb, 0, ns, true, &da);
                reduce_detectors(e, nb, n_det, sqrt2, e->det_out.p + (size_t)p0 * n_det, 1);
            }
        }
        HIPX(hipMemcpyAsync(accum, e->det_out.p, (size_t)n_probes * n_det * sizeof(double), hipMemcpyDeviceToHost,
                            e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_stem_batch_frames(mse_engine* e, int cfg, const double* pos_xy, const double* frame_xy, int n_probes,
                          int batch, const mse_probe* probe, const mse_detector* dets, int n_det, int sqrt2,
                          double* accum) {
    return run_guarded("mse_stem_batch_frames", e, [&]() {
        REQUIRE(e && pos_xy && frame_xy && probe && dets && accum, MSE_ERR_ARG, "NULL argument");
        require_geom(e);
        REQUIRE(e->form >= 0, MSE_ERR_STATE, "mse_set_parameterisation has not been called");
        REQUIRE(n_probes >= 1 && batch >= 1 && n_det >= 1, MSE_ERR_ARG, "counts must be positive");
        REQUIRE(batch <= 65535, MSE_ERR_ARG, "at most 65535 probes per batch");
        for (int i = 0; i < 2 * n_probes; ++i)
            REQUIRE(std::isfinite(frame_xy[i]) && std::isfinite(pos_xy[i]), MSE_ERR_ARG, "position not finite");
        ConfigData& c = config(e, cfg);
        REQUIRE(c.sorted, MSE_ERR_STATE, "atoms of this configuration are not sorted for this geometry");
        REQUIRE(!c.own_beam, MSE_ERR_UNSUPPORTED, "moving frames run with the engine's beam");
        REQUIRE(e->mix_sets < 2, MSE_ERR_UNSUPPORTED, "moving frames build their own transmission functions");
        const mse_geometry& g = cfg_geom(e, c);
        const int ns = g.n_slices;
        if (e->n_waves < batch || !e->waves.p || e->mix_key.n < (size_t)e->n_waves) {
            HIPX(hipStreamSynchronize(e->stream));
            size_wave_buffers(e, std::max(batch, e->n_waves));
            e->n_waves = std::max(batch, e->n_waves);
        }
        std::vector<int> wc(e->n_waves, cfg);
        if (e->h_wave_cfg != wc) {
            e->h_wave_cfg = wc;
            HIPX(hipMemcpyAsync(e->wave_cfg.p, wc.data(), wc.size() * sizeof(int), hipMemcpyHostToDevice,
                                e->stream));
        }
        sync_wave_prop(e);
        ensure_detectors(e, dets, n_det);
        HIPX(hipStreamSynchronize(e->stream));  // buffers below may be reallocated
        e->partials.ensure((size_t)batch * n_det * (e->res / 8));
        e->det_out.ensure((size_t)n_probes * n_det);
        e->frame_xy.ensure(2 * (size_t)batch);
        e->frame_trans.ensure((size_t)batch * e->N);
        // the row pass finds wave w's transmission function as plane w of frame_trans: a table
        // of one "configuration" per wave, each one slice deep, read at slice 0
        e->frame_ptrs.ensure(batch);
        e->frame_ns.ensure(batch);
        e->frame_wcfg.ensure(batch);
        if (!e->h_focus.empty()) e->probe_df.ensure(std::max<size_t>(batch, e->n_waves));
        {
            std::vector<const float2*> ptrs(batch);
            std::vector<int> ones(batch, 1), ids(batch);
            for (int w = 0; w < batch; ++w) {
                ptrs[w] = e->frame_trans.p + (size_t)w * e->N;
                ids[w] = w;
            }
            HIPX(hipMemcpyAsync(e->frame_ptrs.p, ptrs.data(), batch * sizeof(const float2*), hipMemcpyHostToDevice,
                                e->stream));
            HIPX(hipMemcpyAsync(e->frame_ns.p, ones.data(), batch * sizeof(int), hipMemcpyHostToDevice, e->stream));
            HIPX(hipMemcpyAsync(e->frame_wcfg.p, ids.data(), batch * sizeof(int), hipMemcpyHostToDevice, e->stream));
            HIPX(hipMemcpyAsync(e->det_out.p, accum, (size_t)n_probes * n_det * sizeof(double),
                                hipMemcpyHostToDevice, e->stream));
            HIPX(hipStreamSynchronize(e->stream));  // the host vectors go out of scope
        }
        mse::PotArgs pa = pot_args(e, c);
        pa.frame_xy = e->frame_xy.p;
        mse::RowArgs ra = base_row_args(e);
        ra.cfg_trans = e->frame_ptrs.p;
        ra.cfg_nslices = e->frame_ns.p;
        ra.wave_cfg = e->frame_wcfg.p;
        ra.slice = 0;
        ra.mix_sets = 0;
        const mse::ColArgs ca = base_col_args(e);
        mse::ColArgs da = base_col_args(e);
        da.det_scale = 1.0;
        da.det_bits = e->det_bits.p;
        da.ndet = n_det;
        da.partials = e->partials.p;
        std::vector<double> rel(2 * (size_t)batch), org(2 * (size_t)batch);
        for (int p0 = 0; p0 < n_probes; p0 += batch) {
            const int nb = std::min(batch, n_probes - p0);
            for (int i = 0; i < 2 * nb; ++i) {
                const double f = frame_xy[2 * (size_t)p0 + i];
                // each probe relative to its own field's origin (simulationcbed.cpp:54-59); the
                // potential grid at that origin plus the source-size offset (simulationgeneral.cpp:559-560)
                rel[i] = pos_xy[2 * (size_t)p0 + i] - f;
                org[i] = f + c.ref_off[i & 1];
            }
            HIPX(hipMemcpyAsync(e->probe_pos.p, rel.data(), 2 * (size_t)nb * sizeof(double), hipMemcpyHostToDevice,
                                e->stream));
            HIPX(hipMemcpyAsync(e->frame_xy.p, org.data(), 2 * (size_t)nb * sizeof(double), hipMemcpyHostToDevice,
                                e->stream));
            if (!e->h_focus.empty()) {
                REQUIRE((size_t)p0 + nb <= e->h_focus.size(), MSE_ERR_ARG,
                        "fewer focus changes than probes (mse_set_focus_spread)");
                HIPX(hipMemcpyAsync(e->probe_df.p, e->h_focus.data() + p0, nb * sizeof(double),
                                    hipMemcpyHostToDevice, e->stream));
            }
            probes_into_waves(e, nb, probe);
            for (int s = 0; s < ns; ++s) {
                potential_planes(e, c, g, pa, e->frame_trans.p, nb, s);
                band_limit_planes(e, e->frame_trans.p, nb);
                row_pass(e, mse::ROW_IFFT_T_FFT, e->waves.p, e->waves.p, nb, e->N, ra);
                if (s == ns - 1)
                    e->det_tiles = col_pass(e, mse::COL_PROP_DETECT, e->waves.p, e->waves.p, nb, e->N, da);
                else
                    col_pass(e, mse::COL_PROP, e->waves.p, e->waves.p, nb, e->N, ca);
            }
            reduce_detectors(e, nb, n_det, sqrt2, e->det_out.p + (size_t)p0 * n_det, 1);
            // rel and the frames are reused by the next batch: its uploads must not overtake
            HIPX(hipStreamSynchronize(e->stream));
        }
        HIPX(hipMemcpyAsync(accum, e->det_out.p, (size_t)n_probes * n_det * sizeof(double), hipMemcpyDeviceToHost,
                            e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

}  // extern "C"

namespace mse_rt {
void destroy_stem_graph(mse_engine* e) {
    if (e->stem_exec) (void)hipGraphExecDestroy(e->stem_exec);
    e->stem_exec = nullptr;
    e->stem_key.clear();
    if (e->loop_exec) (void)hipGraphExecDestroy(e->loop_exec);
    e->loop_exec = nullptr;
    e->loop_key.clear();
}
}  // namespace mse_rt

extern "C" {

int mse_set_mixed_potentials(mse_engine* e, int first_config, int n_sets, unsigned long long seed) {
    return run_guarded("mse_set_mixed_potentials", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        REQUIRE(n_sets >= 0 && first_config >= 0 && (size_t)first_config + n_sets <= 4096, MSE_ERR_ARG,
                "mixed potential sets out of range");
        HIPX(hipStreamSynchronize(e->stream));
        e->mix_first = n_sets >= 2 ? first_config : 0;
        e->mix_sets = n_sets >= 2 ? n_sets : 0;
        e->mix_seed = seed;
        e->map_key.clear();  // the active wave sets depend on the mixing
    });
}

int mse_set_config_beam(mse_engine* e, int cfg, const double* k_vec) {
    return run_guarded("mse_set_config_beam", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        require_geom(e);
        ConfigData& c = config(e, cfg);
        HIPX(hipStreamSynchronize(e->stream));  // kernels in flight may read its tables
        if (!k_vec) {
            c.own_beam = false;
        } else {
            const double kx = k_vec[0], ky = k_vec[1], kz = k_vec[2];
            REQUIRE(std::isfinite(kx) && std::isfinite(ky) && std::isfinite(kz) && kz > 0.0, MSE_ERR_ARG,
                    "beam vector must be finite with kz > 0");
            const double kk = std::sqrt(kx * kx + ky * ky + kz * kz);
            REQUIRE(std::fabs(kk - e->g.wavenumber) <= 1e-9 * e->g.wavenumber, MSE_ERR_ARG,
                    "beam vector length must equal the wavenumber");
            c.k_vec[0] = kx;
            c.k_vec[1] = ky;
            c.k_vec[2] = kz;
            std::vector<float2> hpx, hpy;
            propagator_tables(e, c.k_vec, hpx, hpy);
            c.px.ensure(e->res);
            c.py.ensure(e->res);
            HIPX(hipMemcpyAsync(c.px.p, hpx.data(), e->res * sizeof(float2), hipMemcpyHostToDevice, e->stream));
            HIPX(hipMemcpyAsync(c.py.p, hpy.data(), e->res * sizeof(float2), hipMemcpyHostToDevice, e->stream));
            HIPX(hipStreamSynchronize(e->stream));
            c.own_beam = true;
        }
        ++c.version;  // transmission functions built before no longer match its potentials
        sync_wave_prop(e);
    });
}

int mse_set_config_geometry(mse_engine* e, int cfg, const mse_geometry* g) {
    return run_guarded("mse_set_config_geometry", e, [&]() {
        REQUIRE(e && g, MSE_ERR_ARG, "NULL argument");
        require_geom(e);
        const mse_geometry& m = e->g;
        // only the slicing of the specimen may differ: same grid, field, blocks and optics
        const bool same_xy = g->resolution == m.resolution && g->real_scale == m.real_scale &&
                             g->full_x[0] == m.full_x[0] && g->full_x[1] == m.full_x[1] &&
                             g->full_y[0] == m.full_y[0] && g->full_y[1] == m.full_y[1] &&
                             g->sim_x[0] == m.sim_x[0] && g->sim_y[0] == m.sim_y[0] &&
                             g->blocks_x == m.blocks_x && g->blocks_y == m.blocks_y && g->load_x == m.load_x &&
                             g->load_y == m.load_y && g->load_z == m.load_z &&
                             g->slice_thickness == m.slice_thickness && g->wavelength == m.wavelength &&
                             g->sigma_eff == m.sigma_eff && g->k_vec[0] == m.k_vec[0] &&
                             g->k_vec[1] == m.k_vec[1] && g->k_vec[2] == m.k_vec[2] &&
                             g->full3d == m.full3d && g->full3d_integrals == m.full3d_integrals &&
                             g->kmax_band == m.kmax_band && g->limit_factor == m.limit_factor;
        REQUIRE(same_xy, MSE_ERR_ARG,
                "configuration geometry must share the engine's grid, field, blocks, slice thickness and optics");
        REQUIRE(g->n_slices >= 1 && g->n_slices <= m.n_slices, MSE_ERR_ARG,
                "configuration has more slices than the engine geometry (set the thickest specimen's geometry "
                "on the engine)");
        ConfigData& c = config(e, cfg);
        HIPX(hipStreamSynchronize(e->stream));
        c.trans.release();
        c.built = false;
        c.sorted = false;
        ++c.version;
        c.g = *g;
        c.own_geom = true;
        upload_cfg_ptrs(e);
    });
}

int mse_set_focus_spread(mse_engine* e, const double* delta_c10, int n) {
    return run_guarded("mse_set_focus_spread", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        REQUIRE(n >= 0 && (n == 0 || delta_c10), MSE_ERR_ARG, "NULL focus changes");
        for (int i = 0; i < n; ++i) REQUIRE(std::isfinite(delta_c10[i]), MSE_ERR_ARG, "focus change not finite");
        HIPX(hipStreamSynchronize(e->stream));  // earlier uploads read the previous list
        e->h_focus.assign(delta_c10, delta_c10 + n);
    });
}

int mse_set_reference_offset(mse_engine* e, int cfg, double dx, double dy) {
    return run_guarded("mse_set_reference_offset", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        REQUIRE(std::isfinite(dx) && std::isfinite(dy), MSE_ERR_ARG, "offset not finite");
        ConfigData& c = config(e, cfg);
        c.ref_off[0] = dx;
        c.ref_off[1] = dy;
    });
}

int mse_init_probes_range(mse_engine* e, int first_wave, const double* pos_xy, int n, const mse_probe* probe) {
    return run_guarded("mse_init_probes_range", e, [&]() {
        REQUIRE(e && pos_xy && probe, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        REQUIRE(first_wave >= 0 && n >= 1 && first_wave + n <= e->n_waves, MSE_ERR_ARG, "wave range out of bounds");
        upload_positions(e, pos_xy, n, first_wave, 0);
        mse::ColArgs ca = base_col_args(e);
        ca.probe = probe_args(e, probe);
        ca.probe.pos = e->probe_pos.p + 2 * (size_t)first_wave;
        if (ca.probe.dC10) ca.probe.dC10 += first_wave;
        float2* w = e->waves.p + (size_t)first_wave * e->N;
        col_pass(e, mse::COL_PROBE, w, w, n, e->N, ca, PK_PROBE);
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_ctem_image(mse_engine* e, const mse_imaging* im, int accumulate, double* out) {
    return run_guarded("mse_ctem_image", e, [&]() {
        REQUIRE(e && im && out, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        REQUIRE(im->obj_ap_mrad >= 0 && im->obj_smooth_mrad >= 0 && im->beta_mrad >= 0, MSE_ERR_ARG,
                "apertures and beta must be non-negative");
        const int nw = e->n_waves;
        const double lam = e->g.wavelength;
        mse::ColArgs ca = base_col_args(e);
        mse::ProbeArgs& p = ca.probe;
        p.wavelength = lam;
        p.cond_ap2 = (im->obj_ap_mrad * 0.001) / lam;  // ctem_image_d.cl: obj_ap2, ap_smooth_radius
        p.smooth_r = (im->obj_smooth_mrad * 0.001) / lam;
        p.beta2 = (im->beta_mrad * 0.001) / lam;
        p.delta = im->delta;
        p.C10 = im->C10;
        p.C30 = im->C30;
        p.C50 = im->C50;
        p.C12 = make_double2(im->C12[0], im->C12[1]);
        p.C21 = make_double2(im->C21[0], im->C21[1]);
        p.C23 = make_double2(im->C23[0], im->C23[1]);
        p.C32 = make_double2(im->C32[0], im->C32[1]);
        p.C34 = make_double2(im->C34[0], im->C34[1]);
        p.C41 = make_double2(im->C41[0], im->C41[1]);
        p.C43 = make_double2(im->C43[0], im->C43[1]);
        p.C45 = make_double2(im->C45[0], im->C45[1]);
        p.C52 = make_double2(im->C52[0], im->C52[1]);
        p.C54 = make_double2(im->C54[0], im->C54[1]);
        p.C56 = make_double2(im->C56[0], im->C56[1]);
        // column pass: F_y, x transfer(kx, ky) / res^2, F_y^-1 into a work copy of the waves;
        // row pass: |F_x^-1|^2 = |unitary IFFT2(transfer x Psi)|^2 in double
        ca.scale = (float)(1.0 / ((double)e->res * (double)e->res));
        e->img_tmp.ensure((size_t)nw * e->N);
        col_pass(e, mse::COL_CTEM_IMAGE, e->waves.p, e->img_tmp.p, nw, e->N, ca, PK_OTHER);
        e->diff_img.ensure((size_t)nw * e->N);
        e->det_out.ensure(e->N);
        if (accumulate)
            HIPX(hipMemcpyAsync(e->det_out.p, out, e->N * sizeof(double), hipMemcpyHostToDevice, e->stream));
        mse::RowArgs ra = base_row_args(e);
        ra.scale = 1.0f;
        ra.dout = e->diff_img.p;
        row_pass(e, mse::ROW_IFFT_INTENSITY, e->img_tmp.p, e->img_tmp.p, nw, e->N, ra);
        launch(e, PK_OTHER, [&]() {
            mse::k_sum_waves<<<(unsigned)((e->N + 255) / 256), 256, 0, e->stream>>>(e->diff_img.p, nw, e->N,
                                                                                   e->det_out.p, accumulate);
        });
        HIPX(hipMemcpyAsync(out, e->det_out.p, e->N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_exit_intensity_sum(mse_engine* e, int accumulate, double* out) {
    return run_guarded("mse_exit_intensity_sum", e, [&]() {
        REQUIRE(e && out, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        const int nw = e->n_waves;
        e->diff_img.ensure((size_t)nw * e->N);
        e->det_out.ensure(e->N);
        if (accumulate)
            HIPX(hipMemcpyAsync(e->det_out.p, out, e->N * sizeof(double), hipMemcpyHostToDevice, e->stream));
        mse::RowArgs ra = base_row_args(e);
        ra.scale = (float)(1.0 / e->res);
        ra.dout = e->diff_img.p;
        row_pass(e, mse::ROW_IFFT_INTENSITY, e->waves.p, e->scratch.p, nw, e->N, ra);
        launch(e, PK_OTHER, [&]() {
            mse::k_sum_waves<<<(unsigned)((e->N + 255) / 256), 256, 0, e->stream>>>(e->diff_img.p, nw, e->N,
                                                                                   e->det_out.p, accumulate);
        });
        HIPX(hipMemcpyAsync(out, e->det_out.p, e->N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_diffraction_sum(mse_engine* e, const int* wave_group, int n_groups, int accumulate, double* out) {
    return run_guarded("mse_diffraction_sum", e, [&]() {
        REQUIRE(e && wave_group && out, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        REQUIRE(n_groups >= 1, MSE_ERR_ARG, "n_groups must be positive");
        const int nw = e->n_waves;
        for (int w = 0; w < nw; ++w)
            REQUIRE(wave_group[w] >= 0 && wave_group[w] < n_groups, MSE_ERR_ARG, "wave group out of range");
        e->diff_img.ensure((size_t)nw * e->N);
        e->det_out.ensure((size_t)n_groups * e->N);
        e->tmp_bid.ensure(nw);
        HIPX(hipMemcpyAsync(e->tmp_bid.p, wave_group, nw * sizeof(int), hipMemcpyHostToDevice, e->stream));
        if (accumulate)
            HIPX(hipMemcpyAsync(e->det_out.p, out, (size_t)n_groups * e->N * sizeof(double), hipMemcpyHostToDevice,
                                e->stream));
        const double* im = diffraction_images(e, 0, nw);  // translated where a beam moved
        launch(e, PK_OTHER, [&]() {
            mse::k_sum_groups<<<(unsigned)((e->N + 255) / 256), 256, 0, e->stream>>>(
                im, nw, e->N, e->tmp_bid.p, n_groups, e->det_out.p, accumulate);
        });
        HIPX(hipMemcpyAsync(out, e->det_out.p, (size_t)n_groups * e->N * sizeof(double), hipMemcpyDeviceToHost,
                            e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

int mse_set_kernel_variant(mse_engine* e, int kernel, int variant) {
    return run_guarded("mse_set_kernel_variant", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        REQUIRE(kernel == 0 || kernel == 1, MSE_ERR_ARG, "kernel must be 0 (row) or 1 (column)");
        REQUIRE(variant >= 0 && variant < (kernel == 0 ? kRowVariants : kColVariants), MSE_ERR_ARG,
                "variant out of range");
        if (kernel == 0)
            e->row_var = variant;
        else
            e->col_var = variant;
        e->tuned.clear();  // an explicit choice holds for every wave count
        ++e->tune_version;
    });
}

int mse_autotune(mse_engine* e, int cfg, int n_waves, int reps, double* row_us, double* col_us,
                 int* best_row, int* best_col) {
    return run_guarded("mse_autotune", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        require_geom(e);
        REQUIRE(n_waves >= 1 && reps >= 1, MSE_ERR_ARG, "n_waves and reps must be positive");
        ConfigData& c = config(e, cfg);
        REQUIRE(c.trans.p && c.built && c.trans.n >= (size_t)cfg_geom(e, c).n_slices * e->N, MSE_ERR_STATE,
                "transmission functions of this configuration are not built for the current geometry");
        if (e->n_waves < n_waves || !e->waves.p) {
            HIPX(hipStreamSynchronize(e->stream));
            size_wave_buffers(e, n_waves);
            e->n_waves = n_waves;
        }
        e->h_wave_cfg.assign(e->n_waves, cfg);
        HIPX(hipMemcpyAsync(e->wave_cfg.p, e->h_wave_cfg.data(), e->n_waves * sizeof(int), hipMemcpyHostToDevice,
                            e->stream));
        sync_wave_prop(e);  // the per-wave propagator table must cover the new wave set
        const size_t total = e->N * n_waves;
        k_fill_plane<<<(unsigned)((total + 255) / 256), 256, 0, e->stream>>>(e->waves.p, e->res, e->N, n_waves, 1.0f);
        HIPX(hipGetLastError());
        hipEvent_t a, b;
        HIPX(hipEventCreate(&a));
        HIPX(hipEventCreate(&b));
        const bool prof = e->prof;
        e->prof = false;
        mse::RowArgs ra = base_row_args(e);
        mse::ColArgs ca = base_col_args(e);
        double best = 1e30;
        int bi = 0;
        for (int v = 0; v < kRowVariants; ++v) {
            row_pass(e, mse::ROW_IFFT_T_FFT, e->waves.p, e->waves.p, n_waves, e->N, ra, v);
            HIPX(hipEventRecord(a, e->stream));
            for (int r = 0; r < reps; ++r) row_pass(e, mse::ROW_IFFT_T_FFT, e->waves.p, e->waves.p, n_waves, e->N, ra, v);
            HIPX(hipEventRecord(b, e->stream));
            HIPX(hipEventSynchronize(b));
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, a, b));
            const double us = 1e3 * ms / reps;
            if (row_us) row_us[v] = us;
            if (us < best) { best = us; bi = v; }
        }
        e->row_var = bi;
        best = 1e30;
        bi = 0;
        for (int v = 0; v < kColVariants; ++v) {
            col_pass(e, mse::COL_PROP, e->waves.p, e->waves.p, n_waves, e->N, ca, PK_COL, v);
            HIPX(hipEventRecord(a, e->stream));
            for (int r = 0; r < reps; ++r) col_pass(e, mse::COL_PROP, e->waves.p, e->waves.p, n_waves, e->N, ca, PK_COL, v);
            HIPX(hipEventRecord(b, e->stream));
            HIPX(hipEventSynchronize(b));
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, a, b));
            const double us = 1e3 * ms / reps;
            if (col_us) col_us[v] = us;
            if (us < best) { best = us; bi = v; }
        }
        e->col_var = bi;
        // remembered for this wave count; other wave counts use the latest choice
        bool found = false;
        for (auto& t : e->tuned)
            if (t.first == n_waves) {
                t.second = {e->row_var, e->col_var};
                found = true;
            }
        if (!found) e->tuned.push_back({n_waves, {e->row_var, e->col_var}});
        ++e->tune_version;
        e->prof = prof;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (best_row) *best_row = e->row_var;
        if (best_col) *best_col = e->col_var;
    });
}

int mse_time_sweeps(mse_engine* e, int slice, int reps, double* row_us, double* col_us) {
    return run_guarded("mse_time_sweeps", e, [&]() {
        REQUIRE(e && row_us && col_us, MSE_ERR_ARG, "NULL argument");
        check_waves(e);
        check_trans(e);
        REQUIRE(reps >= 1 && slice >= 0 && slice < e->g.n_slices, MSE_ERR_ARG, "bad slice or repetition count");
        const bool prof = e->prof;
        e->prof = false;
        hipEvent_t a, b, c;
        HIPX(hipEventCreate(&a));
        HIPX(hipEventCreate(&b));
        HIPX(hipEventCreate(&c));
        mse::RowArgs ra = base_row_args(e);
        ra.slice = slice;
        const mse::ColArgs ca = base_col_args(e);
        const int nw = e->n_waves;
        row_pass(e, mse::ROW_IFFT_T_FFT, e->waves.p, e->waves.p, nw, e->N, ra);  // warm
        col_pass(e, mse::COL_PROP, e->waves.p, e->waves.p, nw, e->N, ca);
        HIPX(hipEventRecord(a, e->stream));
        for (int r = 0; r < reps; ++r) row_pass(e, mse::ROW_IFFT_T_FFT, e->waves.p, e->waves.p, nw, e->N, ra);
        HIPX(hipEventRecord(b, e->stream));
        for (int r = 0; r < reps; ++r) col_pass(e, mse::COL_PROP, e->waves.p, e->waves.p, nw, e->N, ca);
        HIPX(hipEventRecord(c, e->stream));
        HIPX(hipEventSynchronize(c));
        float ms_r = 0, ms_c = 0;
        HIPX(hipEventElapsedTime(&ms_r, a, b));
        HIPX(hipEventElapsedTime(&ms_c, b, c));
        *row_us = 1e3 * ms_r / reps;
        *col_us = 1e3 * ms_c / reps;
        e->prof = prof;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        (void)hipEventDestroy(c);
    });
}

int mse_bandwidth_probe(mse_engine* e, int pattern, int variant, int n_waves, int reps, double* us_per_launch) {
    return run_guarded("mse_bandwidth_probe", e, [&]() {
        REQUIRE(e && us_per_launch, MSE_ERR_ARG, "NULL argument");
        require_geom(e);
        REQUIRE(pattern == 0 || pattern == 1, MSE_ERR_ARG, "pattern must be 0 (rows) or 1 (column tiles)");
        REQUIRE(n_waves >= 1 && reps >= 1, MSE_ERR_ARG, "n_waves and reps must be positive");
        REQUIRE(variant >= 0 && variant < kColVariants, MSE_ERR_ARG, "variant out of range");
        if (e->n_waves < n_waves || !e->waves.p) {
            HIPX(hipStreamSynchronize(e->stream));
            size_wave_buffers(e, n_waves);
            e->h_wave_cfg.assign(n_waves, 0);
            HIPX(hipMemcpyAsync(e->wave_cfg.p, e->h_wave_cfg.data(), n_waves * sizeof(int), hipMemcpyHostToDevice,
                                e->stream));
            e->n_waves = n_waves;
            sync_wave_prop(e);
        }
        hipEvent_t a, b;
        HIPX(hipEventCreate(&a));
        HIPX(hipEventCreate(&b));
        const bool prof = e->prof;
        e->prof = false;
        mse::RowArgs ra = base_row_args(e);
        mse::ColArgs ca = base_col_args(e);
        auto one = [&]() {
            if (pattern == 0)
                row_pass(e, mse::ROW_COPY, e->waves.p, e->waves.p, n_waves, e->N, ra, 0);
            else
                col_pass(e, mse::COL_COPY, e->waves.p, e->waves.p, n_waves, e->N, ca, PK_COL, variant);
        };
        one();
        HIPX(hipEventRecord(a, e->stream));
        for (int r = 0; r < reps; ++r) one();
        HIPX(hipEventRecord(b, e->stream));
        HIPX(hipEventSynchronize(b));
        float ms = 0;
        HIPX(hipEventElapsedTime(&ms, a, b));
        *us_per_launch = 1e3 * ms / reps;
        e->prof = prof;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    });
}

int mse_profile_enable(mse_engine* e, int enable) {
    return run_guarded("mse_profile_enable", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        drain_profile(e);
        e->prof = enable != 0;
    });
}

int mse_profile_read(mse_engine* e, int kernel_id, int* launches, double* total_ms) {
    return run_guarded("mse_profile_read", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        REQUIRE(kernel_id >= 0 && kernel_id < PK_COUNT, MSE_ERR_ARG, "kernel id out of range");
        drain_profile(e);
        if (launches) *launches = e->prof_n[kernel_id];
        if (total_ms) *total_ms = e->prof_ms[kernel_id];
    });
}

int mse_profile_reset(mse_engine* e) {
    return run_guarded("mse_profile_reset", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        drain_profile(e);
        for (int i = 0; i < PK_COUNT; ++i) {
            e->prof_ms[i] = 0.0;
            e->prof_n[i] = 0;
        }
    });
}

}  // extern "C"

