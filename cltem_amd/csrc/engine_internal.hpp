// Engine state shared by the translation units of libcltem_amd.so (not part of the ABI).
#pragma once
// cltem_amd engine: HIP runtime state and the C ABI of include/cltem_amd.h.
//
// One engine = one GPU, one HIP stream (default priority), device buffers owned by
// the engine and reused across calls, as clTEM's SimulationGeneral owns its clMemory
// members (microscope/simulationgeneral.h:74-95).  Errors are C++ exceptions inside
// and become an int status plus a thread-local message at the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cltem_amd.h"
#include "fft_wave.hpp"
#include "slice_kernels.hpp"



namespace mse {
struct PotArgs;
}

namespace mse_rt {

inline thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPX(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw Error(MSE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

#define REQUIRE(cond, code, msg)              \
    do {                                      \
        if (!(cond)) throw Error(code, msg);  \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        const hipError_t rc = hipMalloc((void**)&p, count * sizeof(T));
        if (rc != hipSuccess) {
            // a failed allocation also leaves its code as the runtime's last error, which the
            // launch check of the next kernel would report: clear it, the error is thrown here
            (void)hipGetLastError();
            p = nullptr;
            throw Error(MSE_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(rc));
        }
        n = count;
    }
};

struct ConfigData {
    DevBuf<double> x, y, z;
    DevBuf<int> Z;
    DevBuf<int> block_start;
    DevBuf<int> orig;       // input index of each sorted atom
    int n_input = 0;        // atoms passed to the last sort (in range or not)
    unsigned version = 0;   // bumped whenever the atoms or transmission functions change
    double ref_off[2] = {0.0, 0.0};  // potential grid origin shift (source-size incoherence)
    int n_in_range = 0;
    std::vector<unsigned char> has_Z;  // atomic numbers present in the last sort (index Z)
    bool sorted = false;
    DevBuf<float2> trans;  // n_slices * N
    bool built = false;     // trans holds the transmission functions of the current geometry
    bool own_geom = false;  // z-geometry of its own (a specimen of another thickness)
    mse_geometry g{};
    // beam wave vector of its own (plasmon tilt): tilted potentials and propagator tables
    bool own_beam = false;
    double k_vec[3] = {0.0, 0.0, 0.0};
    DevBuf<float2> px, py;
};

enum ProfKind : int { PK_ROW = 0, PK_COL = 1, PK_POT = 2, PK_PROBE = 3, PK_OTHER = 4, PK_COUNT = 5 };



}  // namespace mse_rt

using namespace mse_rt;

struct mse_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    bool have_geom = false;
    mse_geometry g{};
    int res = 0;
    size_t N = 0;
    DevBuf<float2> tw, px, py;
    DevBuf<int> prop_lim, band_lim;
    DevBuf<float2> band_xy;  // x then y factors of the band limit's column pass
    DevBuf<double> kx, ky;
    std::vector<double> h_kx, h_ky;
    DevBuf<float> params;
    DevBuf<float> params3d;  // per-element constants of the full-3D potentials (kPot3dStride per Z)
    DevBuf<float> params_proj;  // per-element constants of the projected potentials (kPotProjStride per Z)
    int form = -1, i_count = 0, per_atom = 0, max_Z = 0;
    // tabulated potentials (potential_table.hpp): the parameterisation's values, and per
    // table (0 projected, 1 full 3D) and element the host copy and whether it is on the device
    std::vector<double> h_params;
    std::vector<float> h_vtab;           // [2][max_Z][kBuckets][4], as uploaded (cubics in m)
    std::vector<unsigned char> vtab_on;  // [2][max_Z]
    DevBuf<float4> vtab;                 // [2][max_Z][kBuckets]
    std::vector<int> h_vslot;            // element slots of the LDS tables (PotArgs::vslot)
    DevBuf<int> vslot;
    bool use_vtab = true;                // mse_set_potential_tables
    bool vtab_lds = true;                // full-3D tables copied into LDS where they fit
    std::vector<ConfigData> configs;
    DevBuf<const float2*> cfg_ptrs;
    DevBuf<int> cfg_nslices;
    int n_waves = 0;
    DevBuf<float2> waves;
    DevBuf<int> wave_cfg;
    std::vector<int> h_wave_cfg;
    DevBuf<double> probe_pos;
    DevBuf<double> probe_df;            // per wave: focus change of the probe (chromatic spread)
    std::vector<double> h_focus;        // per probe of the next probe set-ups (mse_set_focus_spread)
    DevBuf<unsigned> det_bits;
    std::vector<double> det_key;  // detectors the bitmasks were built for
    DevBuf<double> partials;
    DevBuf<double> det_out;
    DevBuf<double> diff_img;
    DevBuf<double> diff_moved;   // bilinearly translated diffraction images (plasmon tilt)
    DevBuf<double> wave_shift;   // per wave: on, ix, iy, sx, sy of the translation
    // per wave: the px, py tables of its configuration's propagator, uploaded when some wave
    // uses a configuration with a beam of its own (then any_beam; otherwise px, py for all)
    DevBuf<const float2*> wave_prop;
    std::vector<const float2*> h_wave_prop;
    bool any_beam = false;
    // mixed static potentials (mse_set_mixed_potentials): per wave, the key of its set draws
    DevBuf<unsigned> mix_key;
    int mix_first = 0, mix_sets = 0;
    unsigned long long mix_seed = 0;
    DevBuf<float2> scratch;
    DevBuf<float2> img_tmp;  // CTEM image formation work copy of the waves
    DevBuf<float> ccd_filter;  // sqrt(DQE) and NTF factors per pixel (2 res*res)
    // sort scratch
    DevBuf<double> tmp_xyz;
    DevBuf<int> tmp_Z, tmp_bid, tmp_zid;
    DevBuf<unsigned> keys, keys_sorted, idx, idx_sorted;
    DevBuf<unsigned char> cub_tmp;
    int row_var = 4, col_var = 7;  // kernel variants (see row_pass / col_pass); pipelined by default
    // per wave count: the (row, column) variants mse_autotune found fastest for that many waves;
    // sweeps over a tuned wave count use them, any other count the defaults above
    std::vector<std::pair<int, std::pair<int, int>>> tuned;
    unsigned tune_version = 0;     // bumped by every change of variants (captured graphs depend on them)
    hipGraphExec_t stem_exec = nullptr;
    std::vector<unsigned char> stem_key;
    hipGraphExec_t loop_exec = nullptr;   // captured slice loop of mse_multislice_graph
    std::vector<unsigned char> loop_key;
    // waves still inside their specimen, per slice (see active_waves in engine.hip)
    DevBuf<int> wave_map;
    std::vector<int> map_key;                 // wave configurations and slice counts it was built for
    std::vector<int> map_count, map_offset;   // per slice: active waves, offset into wave_map (-1: prefix)
    DevBuf<long long> dev_offset;
    // moving STEM frame (mse_stem_batch_frames): per wave of a batch its frame origin, the
    // transmission plane of the current slice, and the row pass's lookup tables into them
    DevBuf<double> frame_xy;
    DevBuf<float2> frame_trans;
    DevBuf<const float2*> frame_ptrs;
    DevBuf<int> frame_ns, frame_wcfg;
    int num_cus = 256;
    int det_tiles = 0;             // column tiles of the last detector launch
    // profiling
    bool prof = false;
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> ev_pool;
    double prof_ms[PK_COUNT] = {0};
    int prof_n[PK_COUNT] = {0};
    // adjoint slice loop: checkpoints of the last checkpointed forward run
    DevBuf<float2> adj_ckpt, adj_exit, adj_seg, adj_g, adj_gt, adj_gtau;
    DevBuf<float2> adj_pconj;  // conj(px), conj(py): the backward column pass's tables
    // forward states of all slices are kept (no recompute) when they take at most this many bytes
    size_t adj_budget = (size_t)1 << 30;
    DevBuf<float> adj_phase;
    DevBuf<double> adj_w, adj_grad;
    int adj_cfg = -1, adj_k = 0, adj_nw = 0, adj_nslices = 0;
    unsigned adj_version = 0;
    bool adj_ready = false;
};

namespace mse_rt {

inline hipEvent_t take_event(mse_engine* e) {
    if (!e->ev_pool.empty()) {
        hipEvent_t ev = e->ev_pool.back();
        e->ev_pool.pop_back();
        return ev;
    }
    hipEvent_t ev;
    HIPX(hipEventCreate(&ev));
    return ev;
}

template <class F>
inline void launch(mse_engine* e, int kind, F&& f) {
    hipEvent_t a = nullptr, b = nullptr;
    if (e->prof) {
        a = take_event(e);
        b = take_event(e);
        HIPX(hipEventRecord(a, e->stream));
    }
    f();
    HIPX(hipGetLastError());
    if (e->prof) {
        HIPX(hipEventRecord(b, e->stream));
        e->pending.emplace_back(kind, a, b);
    }
}

// ---- helpers defined in engine.hip ---------------------------------------------------------
// potential-kernel arguments for a configuration's geometry
mse::PotArgs pot_args(mse_engine* e, const ConfigData& c);
const mse_geometry& cfg_geom(mse_engine* e, const ConfigData& c);
ConfigData& config(mse_engine* e, int c);
void require_geom(mse_engine* e);
void check_waves(mse_engine* e);
void check_trans(mse_engine* e);
int run_guarded(const char* fn, mse_engine* e, const std::function<void()>& body);
mse::ColArgs base_col_args(mse_engine* e);
mse::RowArgs base_row_args(mse_engine* e);
void row_pass(mse_engine* e, int mode, const float2* in, float2* out, int nlines_sets, size_t wave_stride,
              const mse::RowArgs& args, int variant = -1);
int col_pass(mse_engine* e, int mode, const float2* in, float2* out, int nsets, size_t wave_stride,
             const mse::ColArgs& args, int kind = PK_COL, int variant = -1);
// one forward slice (row pass with t, column pass with the propagator) on nw waves in place
void slice_step(mse_engine* e, float2* waves, int nw, int slice);
// exp(i sigma V) of `planes` planes into out from the sorted atoms of c: slices
// first..first+planes-1, or with a.frame_xy set one plane per frame of slice `first`
void potential_planes(mse_engine* e, const ConfigData& c, const mse_geometry& g, const mse::PotArgs& a,
                      float2* out, int planes, int first);
// transmission band limit of `planes` planes in place: FFT, |k| <= 2/3 kmax, IFFT
void band_limit_planes(mse_engine* e, float2* planes_p, int planes);
// the tables of configuration c's elements for the geometry g (uploaded on first use), or
// NULL when tabulation is off or the geometry's distances fall outside the tables' range
const float4* potential_tables(mse_engine* e, const ConfigData& c, const mse_geometry& g);
// throws MSE_ERR_STATE if the configuration holds elements beyond the parameterisation
void require_elements(const mse_engine* e, const ConfigData& c);
// refresh wave_prop / any_beam after the waves' configurations or a configuration's beam changed
void sync_wave_prop(mse_engine* e);

// ---- sweeps ------------------------------------------------------------------------------
// Kernel variants (chosen by mse_autotune or mse_set_kernel_variant):
//   row    0: default plan, wave-level LDS sync   1: half lanes per line   2: block barriers
//          3: twiddles read at use, >= 8 waves per SIMD
//   column 0: default plan, 16 columns per block  1: half lanes per column
//          2: 8 columns per block                 3: block barriers
//          4: 16 columns, twiddles read at use, >= 8 waves per SIMD (2 blocks per CU)
//          5: 8 columns, twiddles read at use, >= 8 waves per SIMD (4 blocks per CU)
//          6: 32 columns (256-byte row runs), half lanes per column
//   row    4: persistent lines with the next row prefetched into registers
//          5: as 4 with half lanes per line (res <= 512; otherwise as 4)
//          6: as 4 with twiddles read at use
//   column 7: persistent blocks with the next tile prefetched into registers
//          8: as 7 with half lanes per column, two blocks per CU (res <= 512; otherwise as 7)
//          9: as 7 with twiddles read at use
//         10: as 7 with 8-column tiles (512 threads: no register spills at 1024 points)
constexpr int kRowVariants = 7;
//         12: as 10 (8-column tiles) with XCD-aware item order, up to four blocks per CU
//         13: as 7 (16-column tiles) with XCD-aware item order
constexpr int kColVariants = 14;

// Sweep launchers, one translation unit per resolution (sweeps_<N>.hip).
template <int N>
void row_pass_n(mse_engine* e, int mode, const float2* in, float2* out, int nlines_sets, size_t wave_stride,
                const mse::RowArgs& args, int variant);
template <int N>
int col_pass_n(mse_engine* e, int mode, const float2* in, float2* out, int nsets, size_t wave_stride,
               const mse::ColArgs& args, int kind, int variant);

}  // namespace mse_rt
