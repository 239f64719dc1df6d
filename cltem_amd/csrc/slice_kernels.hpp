// Multislice sweep kernels (gfx950, wave64).
//
// The wave of every probe lives in HBM in the mixed (k_x, y) domain: S = F_x psi,
// the unnormalised 1D DFT along x of the real-space wave.  One slice of
// doMultiSliceStep (microscope/simulationgeneral.cpp:744-764),
//     psi' = IFFT2( P . FFT2( t . psi ) )     (unitary 2D FFTs),
// is two sweeps over the wave:
//   row pass    (one line per row):    A = F_x (t . F_x^-1 S)   (= res F_x(t psi))
//   column pass (16 columns per group): V = P' . F_y A ; S' = F_y^-1 V
// with P' = P / res^2 folding both unitary scales, so that V = P . FFT2_unitary(t psi); each sweep reads and writes the
// wave once (2 x 8N bytes) and the row pass reads t once (8N bytes).  The propagator
// is separable, P(kx,ky) = px[kx] py[ky] inside the aperture mask, so it is applied
// from two res-length tables and a per-row mask limit instead of an N-sized array.
// V = Psi(k) is the propagated k-space wave (clWaveFunctionRecip), so detector
// sums are taken from it in the last slice's column pass.
#pragma once
#include "fft_wave.hpp"
#include "philox.hpp"

namespace mse {

enum RowMode : int {
    ROW_IFFT_T_FFT = 0,   // slice step, row half
    ROW_PLANE_T_FFT = 1,  // first slice of a plane wave: psi = value (no load)
    ROW_FFT = 2,          // forward x-DFT (transmission band limit, step 1)
    ROW_IFFT = 3,         // inverse x-DFT (band limit step 3, exit-wave readback)
    ROW_COPY = 4,         // read and write the rows only (bandwidth reference)
    ROW_IFFT_INTENSITY = 5,  // |inverse x-DFT|^2 into a double image per wave (exit intensity)
    ROW_ADJ_SEED = 6,     // adjoint seed: g = 2 w psi for L = sum w |psi|^2, stored as res F_x g
    ROW_ADJ = 7,          // adjoint row half: g_t = conj(psi_s) u, then res F_x(conj(t) u)
};

enum ColMode : int {
    COL_PROP = 0,         // slice step, column half
    COL_PROP_DETECT = 1,  // last slice: propagate and reduce detectors, no store
    COL_PROBE = 3,        // generate the k-space probe and inverse y-DFT
    COL_FFT_DETECT = 4,   // detectors from a stored wave (F_y S / res)
    COL_FFT_DIFF = 5,     // |fftshift Psi|^2 of a stored wave into a double image
    COL_PROP_STORE_DETECT = 6,  // last slice: propagate, reduce detectors and store
    COL_COPY = 7,         // move the tile through LDS only (bandwidth reference)
    COL_PROP_CONJ = 8,    // adjoint column half: propagation by conj(P')
    COL_CTEM_IMAGE = 9,   // objective-lens transfer of Psi(k) (ctem_image), back to the mixed domain
    COL_FILTER = 10,      // multiply the spectrum by a real per-pixel filter (CCD DQE / NTF)
};

constexpr int kMaxDet = 8;

// A pointer known to address global memory.  Pointers read from memory are generic, and
// generic (flat) loads are counted by both the vector-memory and the LDS wait counters, so
// every LDS wait would also wait for them; global loads are counted only by the former.
__device__ __forceinline__ float2 gload(const float2* p) {
    const unsigned long long u = *(const __attribute__((address_space(1))) unsigned long long*)p;
    float2 r;
    __builtin_memcpy(&r, &u, sizeof(r));
    return r;
}
__device__ __forceinline__ int gload(const int* p) { return *(const __attribute__((address_space(1))) int*)p; }
// Load of a table that no kernel writes while it runs (wave maps, configurations, table
// pointers), through the constant address space: at a wave-uniform address this is a scalar
// load, which waits neither for the wave's vector loads nor for its stores.
template <class T>
__device__ __forceinline__ T cload(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}
__device__ __forceinline__ float2 cload(const float2* p) {
    const unsigned long long u = *(const __attribute__((address_space(4))) unsigned long long*)p;
    float2 r;
    __builtin_memcpy(&r, &u, sizeof(r));
    return r;
}

struct RowArgs {
    const float2* const* cfg_trans;  // per configuration: base of its slices
    const int* cfg_nslices;          // per configuration: its number of slices (t = 1 beyond)
    const int* wave_cfg;             // per wave: configuration
    double* dout;                    // ROW_IFFT_INTENSITY: res*res doubles per wave
    const double* weights;           // ROW_ADJ_SEED: dL/dI, res*res doubles per wave
    const int* wave_map;             // set -> wave index (NULL: identity); skips finished waves
    const float2* aux_in;            // ROW_ADJ: forward state S_s = F_x psi_s (wave_stride per wave)
    float2* aux_out;                 // ROW_ADJ: g_t = conj(psi_s) u in real space (wave_stride per wave)
    int slice;
    float plane_value;               // ROW_PLANE_T_FFT: psi amplitude times res
    float scale;                     // ROW_FFT / ROW_IFFT output scale
    // mixed static potentials: with mix_sets >= 2 a wave reads configuration mix_first + a
    // set drawn per (mix_key[wave], slice) instead of its own (row_config)
    const unsigned* mix_key;
    uint2 mix_seed;
    int mix_first, mix_sets;
};

// The configuration whose transmission function `wave` multiplies at slice a.slice: its own,
// or under mixed static potentials floor(n r / 2^32) past mix_first with r the first word of
// Philox4x32-10 (key mix_seed) at counter (mix_key[wave], slice, 0x3A, 6).
__device__ __forceinline__ int row_config(const RowArgs& a, int wave) {
    if (a.mix_sets > 1) {
        const uint4 r = philox4x32_10(make_uint4(cload(a.mix_key + wave), (uint32_t)a.slice, 0x3Au, 6u), a.mix_seed);
        return a.mix_first + (int)(((uint64_t)r.x * (uint64_t)a.mix_sets) >> 32);
    }
    return cload(a.wave_cfg + wave);
}

struct ProbeArgs {
    const double* pos;   // per wave: (x, y) relative to the padded start, Angstrom
    double wavelength;
    double cond_ap2, smooth_r;  // k-space radii (1/Angstrom)
    double C10, C30, C50;
    double2 C12, C21, C23, C32, C34, C41, C43, C45, C52, C54, C56;
    int taper;
    double beta2, delta;        // COL_CTEM_IMAGE: illumination semi-angle (1/Angstrom), defocus spread
    const double* dC10;         // per wave: focus change added to C10 (chromatic spread), or NULL
};

struct ColArgs {
    const int* wave_map;     // set -> wave index (NULL: identity); skips finished waves
    const float2* px;        // propagator factors (COL_PROP*)
    const float2* py;
    const float2* const* wave_prop;  // per wave: its own px, py (beam tilt); NULL: px, py for all
    const int* lim;          // per ky row: largest |kx index| kept (mask), -1 for none
    const float* filter;     // COL_FILTER: res*res real factors, row-major [ky][kx] (unshifted)
    const double* kx;        // k-grid (COL_PROBE)
    const double* ky;
    float scale;             // COL_FFT_*: multiply after the forward DFT
    double det_scale;        // intensity factor applied to |V|^2 in detector sums
    const unsigned* det_bits;  // ndet bitmasks over the fft-shifted image, res*res bits each
    int ndet;
    double* partials;        // [wave][det][tile] partial sums
    double* diff;            // COL_FFT_DIFF: res*res doubles per wave, fft-shifted
    int diff_accumulate;
    ProbeArgs probe;
};

// --------------------------------------------------------------------------------------
// Row pass: 256 threads, 256/L lines per block; grid (res / lines, n_waves)
// --------------------------------------------------------------------------------------
template <class PL, int MODE, class SYNC = WaveSync, bool TWREG = true, int WPE = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_row(
    const float2* __restrict__ in, float2* __restrict__ out, size_t wave_stride,
    const float2* __restrict__ tw, RowArgs a) {
    constexpr int N = PL::N, L = PL::L, E = PL::E;
    constexpr int LINES = 256 / L;
    __shared__ float2 lds[LINES][LdsLine<N>::size];
    const int tid = threadIdx.x;
    const int line = tid / L;
    const int t = tid % L;
    const int row = blockIdx.x * LINES + line;
    const int wave = a.wave_map ? a.wave_map[blockIdx.y] : (int)blockIdx.y;
    float2* ln = lds[line];

    PlanTw<PL, TWREG> tws;
    tws.load(tw, t);
    const SYNC sync{};

    const size_t base = (size_t)wave * wave_stride + (size_t)row * N;
    float2 v[E];
    if constexpr (MODE == ROW_PLANE_T_FFT) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = make_float2(a.plane_value, 0.f);
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = in[base + t + L * m];
    }
    if constexpr (MODE == ROW_COPY) {
#pragma unroll
        for (int m = 0; m < E; ++m) out[base + t + L * m] = v[m];
        return;
    }
    if constexpr (MODE == ROW_ADJ_SEED || MODE == ROW_ADJ) {
        // u = F_x^-1 v / res (real space)
        fft_line<PL, true>(v, ln, t, tws, sync);
        sync();
        const float inv = 1.0f / N;
        if constexpr (MODE == ROW_ADJ_SEED) {
            const double* w = a.weights + (size_t)wave * N * N + (size_t)row * N;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const float f = 2.0f * (float)w[t + L * m] * inv;
                v[m] = make_float2(v[m].x * f, v[m].y * f);  // g = 2 w psi
            }
        } else {
            float2 ps[E];
            const float2* sp = a.aux_in + (size_t)wave * wave_stride + (size_t)row * N;
#pragma unroll
            for (int m = 0; m < E; ++m) ps[m] = sp[t + L * m];
            fft_line<PL, true>(ps, ln, t, tws, sync);
            sync();
            float2* gt = a.aux_out + (size_t)wave * wave_stride + (size_t)row * N;
            const int cfg = row_config(a, wave);
            const float2* tr = a.cfg_trans[cfg] + (size_t)a.slice * N * N + (size_t)row * N;
            const bool has_t = a.slice < a.cfg_nslices[cfg];
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const float2 u = make_float2(v[m].x * inv, v[m].y * inv);
                const float2 psi = make_float2(ps[m].x * inv, ps[m].y * inv);
                gt[t + L * m] = cmulc(u, psi);  // conj(psi) u
                const float2 tv = has_t ? tr[t + L * m] : make_float2(1.f, 0.f);
                v[m] = cmulc(u, tv);  // conj(t) u
            }
        }
        // stored adjoint state res F_x g
        fft_line<PL, false>(v, ln, t, tws, sync);
        const float rs = (float)N;
#pragma unroll
        for (int m = 0; m < E; ++m) out[base + t + L * m] = make_float2(v[m].x * rs, v[m].y * rs);
        return;
    }
    if constexpr (MODE == ROW_IFFT_T_FFT || MODE == ROW_IFFT || MODE == ROW_IFFT_INTENSITY) {
        fft_line<PL, true>(v, ln, t, tws, sync);
        sync();
    }
    if constexpr (MODE == ROW_IFFT_INTENSITY) {
        double* d = a.dout + (size_t)wave * N * N + (size_t)row * N;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const double re = (double)v[m].x * a.scale, im = (double)v[m].y * a.scale;
            d[t + L * m] = re * re + im * im;
        }
        return;
    }
    if constexpr (MODE == ROW_IFFT_T_FFT || MODE == ROW_PLANE_T_FFT) {
        const int cfg = row_config(a, wave);
        if (a.slice < a.cfg_nslices[cfg]) {  // below a thinner specimen the slice is vacuum
            const float2* tr = a.cfg_trans[cfg] + (size_t)a.slice * N * N + (size_t)row * N;
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = cmul(v[m], tr[t + L * m]);
        }
    }
    if constexpr (MODE != ROW_IFFT) {
        fft_line<PL, false>(v, ln, t, tws, sync);
    }
    if constexpr (MODE == ROW_FFT || MODE == ROW_IFFT) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) out[base + t + L * m] = v[m];
}

// complex helpers in double for the probe
__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 dpow(double2 a, int n) {
    double2 r = a;
    for (int j = 1; j < n; ++j) r = dmul(r, a);
    return r;
}

// ctem_image (kernels/ctem_image_d.cl) for one k-space pixel, in double: the factor that
// multiplies Psi(k): smoothed objective aperture x temporal and spatial coherence envelopes x
// exp(-i chi), chi = 2 pi / lambda * sum of the aberration terms in w = lambda (kx + i ky).
__device__ __forceinline__ float2 ctem_transfer(const ProbeArgs& p, double kx, double ky) {
    const double lam = p.wavelength;
    const double k = sqrt(kx * kx + ky * ky);
    if (!(k < p.cond_ap2 + p.smooth_r)) return make_float2(0.f, 0.f);
    const double2 w = make_double2(lam * kx, lam * ky);
    const double2 wc = make_double2(w.x, -w.y);
    const double w2 = w.x * w.x + w.y * w.y;
    const double temporal = exp(-0.25 * M_PI * M_PI * p.delta * p.delta * w2 * w2 / (lam * lam));
    const double defs = p.C10 + p.C30 * w2 + p.C50 * w2 * w2;
    const double spatial = exp(-1.0 * M_PI * M_PI * p.beta2 * p.beta2 * w2 * (defs * defs));
    const double2 wc2 = dpow(wc, 2), wc3 = dpow(wc, 3), wc4 = dpow(wc, 4), wc5 = dpow(wc, 5), wc6 = dpow(wc, 6);
    const double2 w_2 = dpow(w, 2);
    double cchi = 0.5 * p.C10 * w2;
    cchi += 0.5 * dmul(p.C12, wc2).x;
    cchi += dmul(p.C21, dmul(wc2, w)).x / 3.0;
    cchi += dmul(p.C23, wc3).x / 3.0;
    cchi += 0.25 * p.C30 * w2 * w2;
    cchi += 0.25 * dmul(p.C32, dmul(wc3, w)).x;
    cchi += 0.25 * dmul(p.C34, wc4).x;
    cchi += 0.2 * dmul(p.C41, dmul(wc3, w_2)).x;
    cchi += 0.2 * dmul(p.C43, dmul(wc4, w)).x;
    cchi += 0.2 * dmul(p.C45, wc5).x;
    cchi += p.C50 * w2 * w2 * w2 / 6.0;
    cchi += dmul(p.C52, dmul(wc4, w_2)).x / 6.0;
    cchi += dmul(p.C54, dmul(wc5, w)).x / 6.0;
    cchi += dmul(p.C56, wc6).x / 6.0;
    double chi = 2.0 * M_PI * cchi / lam;
    chi = remainder(chi, 2.0 * M_PI);
    double edge = 1.0;
    if (fabs(k - p.cond_ap2) < p.smooth_r) {  // 1 - smoothstep(ap - r, ap + r, k)
        const double t = fmin(fmax((k - (p.cond_ap2 - p.smooth_r)) / (2.0 * p.smooth_r), 0.0), 1.0);
        edge = 1.0 - t * t * (3.0 - 2.0 * t);
    }
    const double amp = edge * temporal * spatial;
    double sc, cc;
    sincos(chi, &sc, &cc);
    return make_float2((float)(amp * cc), (float)(-amp * sc));
}

// init_probe_wave (kernels/init_probe_wave_f.cl:49-118) for one k-space pixel, in double.
__device__ __forceinline__ float2 probe_value(const ProbeArgs& p, double kx, double ky, double px,
                                              double py, double dC10) {
    const double k = sqrt(kx * kx + ky * ky);
    if (!(k < p.cond_ap2 + p.smooth_r)) return make_float2(0.f, 0.f);
    const double lam = p.wavelength;
    const double posTerm = 2.0 * M_PI * (kx * px + ky * py);
    const double2 w = make_double2(lam * kx, lam * ky);
    const double2 wc = make_double2(w.x, -w.y);
    const double w2 = w.x * w.x + w.y * w.y;
    double cchi = 0.5 * (p.C10 + dC10) * w2;
    cchi += 0.5 * dmul(p.C12, dpow(wc, 2)).x;
    cchi += dmul(p.C21, dmul(dpow(wc, 2), w)).x / 3.0;
    cchi += dmul(p.C23, dpow(wc, 3)).x / 3.0;
    cchi += 0.25 * p.C30 * w2 * w2;
    cchi += 0.25 * dmul(p.C32, dmul(dpow(wc, 3), w)).x;
    cchi += 0.25 * dmul(p.C34, dpow(wc, 4)).x;
    cchi += 0.2 * dmul(p.C41, dmul(dpow(wc, 3), dpow(w, 2))).x;
    cchi += 0.2 * dmul(p.C43, dmul(dpow(wc, 4), w)).x;
    cchi += 0.2 * dmul(p.C45, dpow(wc, 5)).x;
    cchi += p.C50 * w2 * w2 * w2 / 6.0;
    cchi += dmul(p.C52, dmul(dpow(wc, 4), dpow(w, 2))).x / 6.0;
    cchi += dmul(p.C54, dmul(dpow(wc, 5), w)).x / 6.0;
    cchi += dmul(p.C56, dpow(wc, 6)).x / 6.0;
    const double chi = 2.0 * M_PI * cchi / lam;
    double edge = 1.0;
    if (p.taper && fabs(k - p.cond_ap2) < p.smooth_r) {
        const double e0 = p.cond_ap2 - p.smooth_r, e1 = p.cond_ap2 + p.smooth_r;
        double tt = (k - e0) / (e1 - e0);
        tt = tt < 0.0 ? 0.0 : (tt > 1.0 ? 1.0 : tt);
        edge = 1.0 - tt * tt * (3.0 - 2.0 * tt);
    }
    double ph = -posTerm - chi;
    ph = remainder(ph, 2.0 * M_PI);
    float s, c;
    sincosf((float)ph, &s, &c);
    return make_float2((float)(c * edge), (float)(s * edge));
}

// wave64 double reduction
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// --------------------------------------------------------------------------------------
// Column pass: TC columns per block, L lanes per column (TC*L threads); grid
// (res / TC, n_waves).  The TC x res tile is staged through LDS column-major with a
// column stride of 1 mod 16 complex, so the coalesced 128-byte row loads and the
// per-column FFT exchanges are both free of bank conflicts; each column's LDS line
// doubles as its FFT exchange buffer.
// --------------------------------------------------------------------------------------
template <class PL, int TC, int MODE, class SYNC = WaveSync, bool TWREG = true, int WPE = 1>
__global__ __launch_bounds__(TC * PL::L) __attribute__((amdgpu_waves_per_eu(WPE))) void k_col(
    const float2* __restrict__ in, float2* __restrict__ out, size_t wave_stride,
    const float2* __restrict__ tw, ColArgs a) {
    constexpr int N = PL::N, L = PL::L, E = PL::E;
    constexpr int NT = TC * L;
    constexpr int CS = LdsLine<N>::col_stride;
    static_assert(TC % 2 == 0, "tiles are moved two complex at a time");
    __shared__ float2 tile[TC * CS];
    __shared__ double red[(NT + 63) / 64][kMaxDet];
    const int tid = threadIdx.x;
    const int c = tid / L;
    const int t = tid % L;
    const int x0 = blockIdx.x * TC;
    const int x = x0 + c;
    const int wave = a.wave_map ? a.wave_map[blockIdx.y] : (int)blockIdx.y;
    float2* ln = tile + c * CS;
    const size_t wbase = (size_t)wave * wave_stride;

    PlanTw<PL, TWREG> tws;
    tws.load(tw, t);
    const SYNC sync{};

    float2 v[E];
    if constexpr (MODE == COL_PROBE) {
        const double kxv = a.kx[x];
        const double ppx = a.probe.pos[2 * wave], ppy = a.probe.pos[2 * wave + 1];
        const double df = a.probe.dC10 ? a.probe.dC10[wave] : 0.0;
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = probe_value(a.probe, kxv, a.ky[t + L * m], ppx, ppy, df);
    } else {
        // coalesced load of the TC x N tile: TC consecutive complex per row, 16 bytes a lane
        const float4* in4 = reinterpret_cast<const float4*>(in + wbase + x0);
#pragma unroll 4
        for (int idx = tid; idx < TC * N / 2; idx += NT) {
            const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
            const float4 q = in4[((size_t)r * N) / 2 + cc / 2];
            tile[cc * CS + lpad(r)] = make_float2(q.x, q.y);
            tile[(cc + 1) * CS + lpad(r)] = make_float2(q.z, q.w);
        }
        __syncthreads();
        if constexpr (MODE == COL_COPY) {
            float4* out4 = reinterpret_cast<float4*>(out + wbase + x0);
#pragma unroll 4
            for (int idx = tid; idx < TC * N / 2; idx += NT) {
                const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
                const float2 a0 = tile[cc * CS + lpad(r)], a1 = tile[(cc + 1) * CS + lpad(r)];
                out4[((size_t)r * N) / 2 + cc / 2] = make_float4(a0.x, a0.y, a1.x, a1.y);
            }
            return;
        }
        line_gather<N, L>(v, ln, t);
        sync();
        fft_line<PL, false>(v, ln, t, tws, sync);
    }

    if constexpr (MODE == COL_PROP || MODE == COL_PROP_DETECT || MODE == COL_PROP_STORE_DETECT ||
                  MODE == COL_PROP_CONJ) {
        const int xs = (x < N / 2) ? x : N - x;  // |signed kx index|
        const float2* pxt = a.px;
        const float2* pyt = a.py;
        if (a.wave_prop) {
            pxt = a.wave_prop[2 * wave];
            pyt = a.wave_prop[2 * wave + 1];
        }
        const float2 pxv = pxt[x];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int y = t + L * m;
            const bool keep = xs <= a.lim[y];
            float2 f = cmul(pxv, pyt[y]);
            if constexpr (MODE == COL_PROP_CONJ) f.y = -f.y;
            v[m] = keep ? cmul(v[m], f) : make_float2(0.f, 0.f);
        }
    } else if constexpr (MODE == COL_FFT_DETECT || MODE == COL_FFT_DIFF) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale);
    } else if constexpr (MODE == COL_FILTER) {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale * a.filter[(size_t)(t + L * m) * N + x]);
    } else if constexpr (MODE == COL_CTEM_IMAGE) {
        const double kxv = a.kx[x];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cscale(cmul(v[m], ctem_transfer(a.probe, kxv, a.ky[t + L * m])), a.scale);
    }

    if constexpr (MODE == COL_PROP_DETECT || MODE == COL_PROP_STORE_DETECT ||
                  MODE == COL_FFT_DETECT) {
        const int xs = (x + N / 2) % N;  // fft-shifted column
        const int nw = NT / 64;
        const int wid = tid / 64, lane = tid % 64;
        for (int d = 0; d < a.ndet; ++d) {
            const unsigned* bits = a.det_bits + (size_t)d * ((size_t)N * N / 32);
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int ys = (t + L * m + N / 2) % N;
                const size_t bit = (size_t)ys * N + xs;
                if ((bits[bit >> 5] >> (bit & 31)) & 1u) {
                    const float i2 = v[m].x * v[m].x + v[m].y * v[m].y;
                    acc += (double)i2;
                }
            }
            acc = wave_sum(acc);
            if (lane == 0) red[wid][d] = acc;
        }
        __syncthreads();
        if (tid < a.ndet) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += red[w][tid];
            a.partials[((size_t)wave * a.ndet + tid) * gridDim.x + blockIdx.x] = s * a.det_scale;
        }
    }
    if constexpr (MODE == COL_FFT_DIFF) {
        double* img = a.diff + (size_t)wave * N * N;
        const int xs = (x + N / 2) % N;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int ys = (t + L * m + N / 2) % N;
            const double i2 = (double)v[m].x * v[m].x + (double)v[m].y * v[m].y;
            const size_t o = (size_t)ys * N + xs;
            img[o] = a.diff_accumulate ? img[o] + i2 : i2;
        }
    }

    if constexpr (MODE == COL_PROP || MODE == COL_PROP_STORE_DETECT ||
                  MODE == COL_PROBE || MODE == COL_PROP_CONJ || MODE == COL_CTEM_IMAGE || MODE == COL_FILTER) {
        if constexpr (MODE != COL_PROBE) sync();  // the forward pass's last reads precede new writes
        fft_line<PL, true>(v, ln, t, tws, sync);
        sync();
#pragma unroll
        for (int m = 0; m < E; ++m) ln[lpad(t + L * m)] = v[m];
        __syncthreads();
        float4* out4 = reinterpret_cast<float4*>(out + wbase + x0);
#pragma unroll 4
        for (int idx = tid; idx < TC * N / 2; idx += NT) {
            const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
            const float2 a0 = tile[cc * CS + lpad(r)], a1 = tile[(cc + 1) * CS + lpad(r)];
            out4[((size_t)r * N) / 2 + cc / 2] = make_float4(a0.x, a0.y, a1.x, a1.y);
        }
    }
}

// --------------------------------------------------------------------------------------
// Persistent, software-pipelined column propagation (COL_PROP).  One block per CU walks
// the (wave, column tile) items; while the columns of item i are transformed, the global
// loads of item i+1 are already in flight into registers, and the stores of item i drain
// during the next item's transforms.  Same arithmetic as k_col<..., COL_PROP>.
// --------------------------------------------------------------------------------------
// XCD: blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8), each with
// its own L2.  With XCD set (and gridDim.x a multiple of 8) the blocks of one XCD take
// consecutive items in every round, so tiles that share 128-byte lines (8-column tiles: two
// per line) are read and written through the same L2 at about the same time.
template <class PL, int TC, class SYNC = WaveSync, bool TWREG = true, int WPE = 1, bool XCD = false>
__global__ __launch_bounds__(TC * PL::L) __attribute__((amdgpu_waves_per_eu(WPE))) void k_col_prop_pipe(float2* __restrict__ data, size_t wave_stride,
                                                              const float2* __restrict__ tw, ColArgs a,
                                                              int nsets) {
    constexpr int N = PL::N, L = PL::L, E = PL::E;
    constexpr int NT = TC * L;
    constexpr int CS = LdsLine<N>::col_stride;
    constexpr int NQ = (TC * N / 2) / NT;  // float4 per thread per tile
    static_assert((TC * N / 2) % NT == 0, "tile must split evenly over the block");
    constexpr int ntiles = N / TC;
    __shared__ float2 tile[TC * CS];
    const int tid = threadIdx.x;
    const int c = tid / L;
    const int t = tid % L;
    const int nitems = nsets * ntiles;

    PlanTw<PL, TWREG> tws;
    tws.load(tw, t);
    const SYNC sync{};

    float4 q[NQ];
    auto fetch = [&](int item) {
        const int wv = a.wave_map ? cload(a.wave_map + item / ntiles) : item / ntiles;
        const int xx = (item % ntiles) * TC;
        const float4* in4 = reinterpret_cast<const float4*>(data + (size_t)wv * wave_stride + xx);
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int idx = tid + k * NT;
            const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
            q[k] = in4[((size_t)r * N) / 2 + cc / 2];
        }
    };
    // the previous tile's output, held in registers and stored once this tile's input is in
    // LDS: its stores then drain during this tile's transforms instead of delaying the LDS
    // writes (the vector-memory counter orders loads and stores together)
    float4 o[NQ];
    auto store = [&](int item) {
        const int wv = a.wave_map ? cload(a.wave_map + item / ntiles) : item / ntiles;
        const int xx = (item % ntiles) * TC;
        float4* out4 = reinterpret_cast<float4*>(data + (size_t)wv * wave_stride + xx);
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int idx = tid + k * NT;
            const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
            out4[((size_t)r * N) / 2 + cc / 2] = o[k];
        }
    };
    // With one propagator for every wave the y factors are the same in every tile: a lane
    // keeps its E of them in registers when E <= 8, larger plans read them from an LDS copy
    // (registers are short there).  Per-wave propagators are loaded per tile.
    constexpr bool PY_REGS = E <= 8;
    const bool per_wave = a.wave_prop != nullptr;
    float2 pyv[PY_REGS ? E : 1];
    __shared__ float2 spy[PY_REGS ? 1 : N];
    if constexpr (PY_REGS) {
        if (!per_wave) {
#pragma unroll
            for (int m = 0; m < E; ++m) pyv[m] = gload(a.py + t + L * m);
        }
    } else {
        if (!per_wave)
            for (int i = tid; i < N; i += NT) spy[i] = a.py[i];  // ordered by the first tile's barrier
    }
    const float2* pyt = a.py;
    int item = XCD ? (int)((blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8) : (int)blockIdx.x;
    int prev = -1;
    if (item < nitems) fetch(item);
    for (; item < nitems; item += gridDim.x) {
        const int wv = a.wave_map ? cload(a.wave_map + item / ntiles) : item / ntiles;
        const int x0 = (item % ntiles) * TC;
        const int x = x0 + c;
        // propagator factors of this column, loaded before the next tile's prefetch so that
        // waiting for them does not wait for the prefetch
        float2 pxv;
        int lim;
        {
            const float2* pxt = a.px;
            if (per_wave) {
                pxt = cload(a.wave_prop + 2 * wv);
                pyt = cload(a.wave_prop + 2 * wv + 1);
                if constexpr (PY_REGS) {
#pragma unroll
                    for (int m = 0; m < E; ++m) pyv[m] = gload(pyt + t + L * m);
                }
            }
            // the 64 / L columns of this wave: scalar loads, which the stores issued before the
            // multiply do not hold up, then each lane takes its own column's values (the
            // aperture mask is symmetric in kx and ky)
            constexpr int CPW = 64 / L;
            static_assert(L <= 64 && 64 % L == 0 && TC % CPW == 0, "whole columns per wave");
            const int xw = __builtin_amdgcn_readfirstlane(x - c % CPW);
            const int j = c % CPW;
            pxv = cload(pxt + xw);
            lim = cload(a.lim + ((xw < N / 2) ? xw : N - xw));
#pragma unroll
            for (int k = 1; k < CPW; ++k) {
                const int xk = xw + k;
                const float2 pk = cload(pxt + xk);
                const int lk = cload(a.lim + ((xk < N / 2) ? xk : N - xk));
                if (j == k) {
                    pxv = pk;
                    lim = lk;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int idx = tid + k * NT;
            const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
            tile[cc * CS + lpad(r)] = make_float2(q[k].x, q[k].y);
            tile[(cc + 1) * CS + lpad(r)] = make_float2(q[k].z, q[k].w);
        }
        if (prev >= 0) store(prev);
        __syncthreads();
        if (item + (int)gridDim.x < nitems) fetch(item + gridDim.x);
        float2* ln = tile + c * CS;
        float2 v[E];
        line_gather<N, L>(v, ln, t);
        sync();
        fft_line<PL, false>(v, ln, t, tws, sync);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int y = t + L * m;
            const int ys = (y < N / 2) ? y : N - y;
            float2 pyy;
            if constexpr (PY_REGS)
                pyy = pyv[m];
            else
                pyy = per_wave ? gload(pyt + y) : spy[y];
            const float2 f = cmul(v[m], cmul(pxv, pyy));
            v[m] = ys <= lim ? f : make_float2(0.f, 0.f);
        }
        sync();
        fft_line<PL, true>(v, ln, t, tws, sync);
        sync();
#pragma unroll
        for (int m = 0; m < E; ++m) ln[lpad(t + L * m)] = v[m];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const int idx = tid + k * NT;
            const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
            const float2 a0 = tile[cc * CS + lpad(r)], a1 = tile[(cc + 1) * CS + lpad(r)];
            o[k] = make_float4(a0.x, a0.y, a1.x, a1.y);
        }
        prev = item;
        __syncthreads();
    }
    if (prev >= 0) store(prev);
}

// --------------------------------------------------------------------------------------
// Column slice step, one 16-column tile per block with no prefetch, sized for two blocks
// per CU (at most 64 VGPRs): twiddles and y factors are read from per-block copies in LDS,
// so that the second block's loads and stores overlap the first one's transforms.  Same
// arithmetic as k_col<..., COL_PROP>.  Grid (N / TC, nsets).
// --------------------------------------------------------------------------------------
// Twiddle table entries a plan's passes read: pass with radix R after p points reads
// r k N / (p R) for r < R, k < p.
template <class PL>
constexpr int plan_twiddles() {
    constexpr int N = PL::N;
    const int R[4] = {PL::R0, PL::R1, PL::R2, PL::R3};
    int p = R[0], most = 0;
    for (int i = 1; i < 4 && R[i] > 1; ++i) {
        const int top = (R[i] - 1) * (p - 1) * (N / (p * R[i]));
        most = top > most ? top : most;
        p *= R[i];
    }
    return most + 1;
}

template <class PL, int TC>
__device__ __forceinline__ void col_prop_lean(float2* __restrict__ data, size_t wave_stride,
                                              const float2* __restrict__ tw, const ColArgs& a) {
    constexpr int N = PL::N, L = PL::L, E = PL::E;
    constexpr int NT = TC * L;
    constexpr int CS = LdsLine<N>::col_stride;
    constexpr int NQ = (TC * N / 2) / NT;
    static_assert(L == 64 && (TC * N / 2) % NT == 0, "one column per wave");
    constexpr int NTW = plan_twiddles<PL>();
    __shared__ float2 tile[TC * CS];
    __shared__ float2 stw[NTW];
    __shared__ float2 spy[N];
    const int tid = threadIdx.x;
    const int c = tid / L;
    const int t = tid % L;
    const int wv = a.wave_map ? cload(a.wave_map + blockIdx.y) : (int)blockIdx.y;
    const int x0 = blockIdx.x * TC;
    const int x = x0 + c;
    const float4* in4 = reinterpret_cast<const float4*>(data + (size_t)wv * wave_stride + x0);
    float4 q[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const int idx = tid + k * NT;
        const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
        q[k] = in4[((size_t)r * N) / 2 + cc / 2];
    }
    const float2* pxt = a.px;
    const float2* pyt = a.py;
    if (a.wave_prop) {
        pxt = cload(a.wave_prop + 2 * wv);
        pyt = cload(a.wave_prop + 2 * wv + 1);
    }
    for (int i = tid; i < N; i += NT) {
        if (i < NTW) stw[i] = tw[i];
        spy[i] = pyt[i];
    }
    const int xs = (x < N / 2) ? x : N - x;
    const float2 pxv = cload(pxt + __builtin_amdgcn_readfirstlane(x));
    const int lim = cload(a.lim + __builtin_amdgcn_readfirstlane(xs));  // the mask is symmetric in kx and ky
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const int idx = tid + k * NT;
        const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
        tile[cc * CS + lpad(r)] = make_float2(q[k].x, q[k].y);
        tile[(cc + 1) * CS + lpad(r)] = make_float2(q[k].z, q[k].w);
    }
    __syncthreads();
    PlanTw<PL, false> tws;
    tws.load(stw, t);
    const WaveSync sync{};
    float2* ln = tile + c * CS;
    float2 v[E];
    line_gather<N, L>(v, ln, t);
    sync();
    fft_line<PL, false>(v, ln, t, tws, sync);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int y = t + L * m;
        const int ys = (y < N / 2) ? y : N - y;
        const float2 f = cmul(v[m], cmul(pxv, spy[y]));
        v[m] = ys <= lim ? f : make_float2(0.f, 0.f);
    }
    sync();
    fft_line<PL, true>(v, ln, t, tws, sync);
    sync();
#pragma unroll
    for (int m = 0; m < E; ++m) ln[lpad(t + L * m)] = v[m];
    __syncthreads();
    float4* out4 = reinterpret_cast<float4*>(data + (size_t)wv * wave_stride + x0);
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const int idx = tid + k * NT;
        const int r = idx / (TC / 2), cc = 2 * (idx % (TC / 2));
        const float2 a0 = tile[cc * CS + lpad(r)], a1 = tile[(cc + 1) * CS + lpad(r)];
        out4[((size_t)r * N) / 2 + cc / 2] = make_float4(a0.x, a0.y, a1.x, a1.y);
    }
}

// 16-column tiles (512 points): 64 VGPRs, two 1024-thread blocks per CU
template <class PL>
__global__ __launch_bounds__(16 * PL::L) __attribute__((amdgpu_waves_per_eu(8), amdgpu_num_vgpr(64))) void k_col_prop_lean(
    float2* __restrict__ data, size_t wave_stride, const float2* __restrict__ tw, ColArgs a) {
    col_prop_lean<PL, 16>(data, wave_stride, tw, a);
}
// 8-column tiles (four-pass plans, 12 elements per lane): two 512-thread blocks per CU
template <class PL>
__global__ __launch_bounds__(8 * PL::L) __attribute__((amdgpu_waves_per_eu(4))) void k_col_prop_lean8(
    float2* __restrict__ data, size_t wave_stride, const float2* __restrict__ tw, ColArgs a) {
    col_prop_lean<PL, 8>(data, wave_stride, tw, a);
}

// --------------------------------------------------------------------------------------
// Persistent, software-pipelined row slice step (ROW_IFFT_T_FFT).  Every line walks the
// (wave, row) items independently; the next item's wave row and transmission row are
// loaded into registers while the current row is transformed.  MODE = ROW_FFT or ROW_IFFT:
// the one transform and the scale of k_row's same modes, in place (no transmission row).
// --------------------------------------------------------------------------------------
template <class PL, class SYNC = WaveSync, bool TWREG = true, int MODE = ROW_IFFT_T_FFT>
__global__ __launch_bounds__(256) void k_row_step_pipe(float2* __restrict__ data, size_t wave_stride,
                                                       const float2* __restrict__ tw, RowArgs a, int nsets) {
    static_assert(MODE == ROW_IFFT_T_FFT || MODE == ROW_FFT || MODE == ROW_IFFT, "row step modes");
    constexpr bool STEP = MODE == ROW_IFFT_T_FFT;
    constexpr int N = PL::N, L = PL::L, E = PL::E;
    constexpr int LINES = 256 / L;
    __shared__ float2 lds[LINES][LdsLine<N>::size];
    const int tid = threadIdx.x;
    const int line = tid / L;
    const int t = tid % L;
    float2* ln = lds[line];
    const int nitems = nsets * N;
    const int stride = gridDim.x * LINES;

    PlanTw<PL, TWREG> tws;
    tws.load(tw, t);
    const SYNC sync{};

    float2 qv[E], qt[E];
    auto fetch = [&](int item) {
        // with one line per wave the item is wave-uniform: its wave, configuration and
        // transmission row then come from scalar loads
        if constexpr (L == 64) item = __builtin_amdgcn_readfirstlane(item);
        const int wv = a.wave_map ? cload(a.wave_map + item / N) : item / N, row = item % N;
        const float2* src = data + (size_t)wv * wave_stride + (size_t)row * N;
#pragma unroll
        for (int m = 0; m < E; ++m) qv[m] = src[t + L * m];
        if constexpr (!STEP) return;
        const int cfg = row_config(a, wv);
        if (a.slice < cload(a.cfg_nslices + cfg)) {
            const float2* tr = cload(a.cfg_trans + cfg) + (size_t)a.slice * N * N + (size_t)row * N;
#pragma unroll
            for (int m = 0; m < E; ++m) qt[m] = gload(tr + t + L * m);
        } else {  // vacuum below a thinner specimen
#pragma unroll
            for (int m = 0; m < E; ++m) qt[m] = make_float2(1.f, 0.f);
        }
    };
    int item = blockIdx.x * LINES + line;
    if (item < nitems) fetch(item);
    for (; item < nitems; item += stride) {
        float2 v[E], tr[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            v[m] = qv[m];
            tr[m] = qt[m];
        }
        if (item + stride < nitems) fetch(item + stride);
        if constexpr (STEP) {
            fft_line<PL, true>(v, ln, t, tws, sync);
            sync();
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = cmul(v[m], tr[m]);
            fft_line<PL, false>(v, ln, t, tws, sync);
        } else {
            fft_line<PL, MODE == ROW_IFFT>(v, ln, t, tws, sync);
#pragma unroll
            for (int m = 0; m < E; ++m) v[m] = cscale(v[m], a.scale);
        }
        sync();
        int it = item;
        if constexpr (L == 64) it = __builtin_amdgcn_readfirstlane(it);
        const int wv = a.wave_map ? cload(a.wave_map + it / N) : it / N, row = it % N;
        float2* dst = data + (size_t)wv * wave_stride + (size_t)row * N;
#pragma unroll
        for (int m = 0; m < E; ++m) dst[t + L * m] = v[m];
    }
}

// sum per-tile partials in fixed order: out[w*ndet + d] (+)= sum_tiles; when dev_offset is
// given the output starts dev_offset[0] entries further (read on the device, so a captured
// graph can be replayed for successive probe batches)
static __global__ void k_reduce_partials(const double* __restrict__ partials, int ntiles, int nwd,
                                  double* __restrict__ out, int accumulate, double factor,
                                  const long long* __restrict__ dev_offset) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nwd) return;
    double s = 0.0;
    for (int k = 0; k < ntiles; ++k) s += partials[(size_t)i * ntiles + k];
    s *= factor;
    double* o = out + (dev_offset ? dev_offset[0] : 0LL);
    o[i] = accumulate ? o[i] + s : s;
}

// out[p] (+)= sum over waves of img[w][p], in wave order
static __global__ void k_sum_waves(const double* __restrict__ img, int nw, size_t npix, double* __restrict__ out,
                                   int accumulate) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    double s = accumulate ? out[p] : 0.0;
    for (int w = 0; w < nw; ++w) s += img[(size_t)w * npix + p];
    out[p] = s;
}

// out[g][p] (+)= sum of img[w][p] over the waves w of group g, in wave order
static __global__ void k_sum_groups(const double* __restrict__ img, int nw, size_t npix,
                                    const int* __restrict__ group, int ngroups, double* __restrict__ out,
                                    int accumulate) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    if (!accumulate)
        for (int g = 0; g < ngroups; ++g) out[(size_t)g * npix + p] = 0.0;
    for (int w = 0; w < nw; ++w) out[(size_t)group[w] * npix + p] += img[(size_t)w * npix + p];
}

}  // namespace mse
