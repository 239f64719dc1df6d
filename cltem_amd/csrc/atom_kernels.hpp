// Atom binning and transmission-function kernels (gfx950).
#pragma once
#include "potential_table.hpp"
#include <hip/hip_runtime.h>
#include <type_traits>

namespace mse {

// ---------------------------------------------------------------------------------------
// Binning.  SimulationGeneral::sortAtoms (microscope/simulationgeneral.cpp:188-328) filters
// atoms with strict limits on the host, bins them with atom_sort_d
// (kernels/atom_sort_d.cl:28-69) and fills the bins stably in input order.  Here the
// filter and the bin formula run per atom in IEEE double with the reference's operation
// order (no contraction is possible in these expressions and it is disabled anyway), and
// the stable fill is a stable radix sort of (slice*blocks + block, index).
// ---------------------------------------------------------------------------------------
struct BinGeom {
    double full_x0, full_x1, full_y0, full_y1;  // paddedFullLimitsX/Y
    double z0, z1;                              // paddedSimLimitsZ
    double dz;
    int blocks_x, blocks_y, n_slices;
};

static __global__ void k_bin_atoms(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, int n, BinGeom g, unsigned* __restrict__ key,
                            unsigned* __restrict__ idx, int* __restrict__ bid_out,
                            int* __restrict__ zid_out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned out_key = (unsigned)(g.n_slices * g.blocks_x * g.blocks_y);
    const double xi = x[i], yi = y[i], zi = z[i];
    // host filter (simulationgeneral.cpp:206-210): strict against padded limits
    const bool in_x = xi > g.full_x0 && xi < g.full_x1;
    const bool in_y = yi > g.full_y0 && yi < g.full_y1;
    const bool in_z = zi > g.z0 && zi < g.z1;
    int bid = -1, zid = -1;
    unsigned k = out_key;
    if (in_x && in_y && in_z) {
        const double min_x = g.full_x0, max_x = g.full_x1, min_y = g.full_y0, max_y = g.full_y1;
        const double max_z = g.z1;
        if (xi >= min_x && xi <= max_x && yi >= min_y && yi <= max_y) {
            int bidx = (int)floor((xi - min_x) * (1.0 / (max_x - min_x)) * g.blocks_x);
            int bidy = (int)floor((yi - min_y) * (1.0 / (max_y - min_y)) * g.blocks_y);
            int z_id = (int)floor((max_z - zi) * (1.0 / (g.dz)));
            z_id -= (z_id == g.n_slices);
            bidx -= (bidx == g.blocks_x);
            bidy -= (bidy == g.blocks_y);
            bid = bidx + g.blocks_x * bidy;
            zid = z_id;
            if (zid >= 0 && bid >= 0)
                k = (unsigned)(zid * g.blocks_x * g.blocks_y + bid);
        }
    }
    key[i] = k;
    idx[i] = (unsigned)i;
    bid_out[i] = bid;
    zid_out[i] = zid;
}

// block_start[k] = first sorted position with key >= k, for k in [0, nkeys]
static __global__ void k_block_starts(const unsigned* __restrict__ sorted_keys, int n, int nkeys,
                               int* __restrict__ block_start) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nkeys) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (sorted_keys[mid] < (unsigned)k)
            lo = mid + 1;
        else
            hi = mid;
    }
    block_start[k] = lo;
}

static __global__ void k_gather_atoms(const unsigned* __restrict__ order, int n, const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               const int* __restrict__ Z, double* __restrict__ sx,
                               double* __restrict__ sy, double* __restrict__ sz, int* __restrict__ sZ) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned j = order[i];
    sx[i] = x[j];
    sy[i] = y[j];
    sz[i] = z[j];
    sZ[i] = Z[j];
}

// ---------------------------------------------------------------------------------------
// Potentials.  Distances and the selection tests (r_min clamp, r <= 8, full-3D window)
// are evaluated in double exactly as the `_d` kernels do; the parameterised potential
// itself is evaluated in float (VALU transcendental cost) and summed in double.
// ---------------------------------------------------------------------------------------

// Bessel K0/K1, Numerical Recipes in C 2nd ed. section 6.6 (the rational forms with the
// coefficient tables of kernels/transmission_potentials_projected_f.cl:57-69), in float.
__device__ __forceinline__ float poly4(float x, float c0, float c1, float c2, float c3, float c4) {
    return (((c4 * x + c3) * x + c2) * x + c1) * x + c0;
}
__device__ __forceinline__ float poly2(float x, float c0, float c1, float c2) {
    return (c2 * x + c1) * x + c0;
}
__device__ __forceinline__ float poly7(float x, const float* c) {
    float a = c[7];
#pragma unroll
    for (int i = 6; i >= 0; --i) a = a * x + c[i];
    return a;
}

__device__ __forceinline__ float bessk0f(float x) {
    if (x <= 1.0f) {
        const float z = x * x;
        const float term = poly4(z, 1.0f, 2.346487949187396e-1f, 1.187082088663404e-2f,
                                 2.150707366040937e-4f, 1.425433617130587e-6f) *
                           logf(x) /
                           poly2(1.0f - z, 9.847324170755358e-1f, 1.518396076767770e-2f,
                                 8.362215678646257e-5f);
        return poly4(z, 1.159315156584126e-1f, 2.770731240515333e-1f, 2.066458134619875e-2f,
                     4.574734709978264e-4f, 3.454715527986737e-6f) /
                   poly2(1.0f - z, 9.836249671709183e-1f, 1.627693622304549e-2f,
                         9.809660603621949e-5f) -
               term;
    }
    const float pp[8] = {1.253314137315499f, 1.475731032429900e1f, 6.123767403223466e1f,
                         1.121012633939949e2f, 9.285288485892228e1f, 3.198289277679660e1f,
                         3.595376024148513f, 6.160228690102976e-2f};
    const float qq[8] = {1.0f, 1.189963006673403e1f, 5.027773590829784e1f, 9.496513373427093e1f,
                         8.318077493230258e1f, 3.181399777449301e1f, 4.443672926432041f,
                         1.408295601966600e-1f};
    const float z = 1.0f / x;
    return expf(-x) * poly7(z, pp) / poly7(z, qq) * rsqrtf(x);
}

__device__ __forceinline__ float bessk1f(float x) {
    if (x <= 1.0f) {
        const float z = x * x;
        const float term = poly4(z, 0.5f, 5.598072040178741e-2f, 1.818666382168295e-3f,
                                 2.397509908859959e-5f, 1.239567816344855e-7f) *
                           logf(x) /
                           poly2(1.0f - z, 9.870202601341150e-1f, 1.292092053534579e-2f,
                                 5.881933053917096e-5f);
        return x * (poly4(z, -3.079657578292062e-1f, -8.109417631822442e-2f, -3.477550948593604e-3f,
                          -5.385594871975406e-5f, -3.110372465429008e-7f) /
                        poly2(1.0f - z, 9.861813171751389e-1f, 1.375094061153160e-2f,
                              6.774221332947002e-5f) +
                    term) +
               1.0f / x;
    }
    const float pp[8] = {1.253314137315502f, 1.457171340220454e1f, 6.063161173098803e1f,
                         1.147386690867892e2f, 1.040442011439181e2f, 4.356596656837691e1f,
                         7.265230396353690f, 3.144418558991021e-1f};
    const float qq[8] = {1.0f, 1.125154514806458e1f, 4.427488496597630e1f, 7.616113213117645e1f,
                         5.863377227890893e1f, 1.850303673841586e1f, 1.857244676566022f,
                         2.538540887654872e-2f};
    const float z = 1.0f / x;
    return expf(-x) * poly7(z, pp) / poly7(z, qq) * rsqrtf(x);
}

// Projected potentials of one atom at distance r (kernels/transmission_potentials_projected_f.cl:104-181)
// The same K0 and K1 with the hardware reciprocal, log2 and exp2 (v_rcp_f32, v_log_f32,
// v_exp_f32, about 1 ulp each) in place of IEEE division and the library log and exp.
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float ln_hw(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

__device__ __forceinline__ float bessk0_hw(float x) {
    if (x <= 1.0f) {
        const float z = x * x;
        const float term = poly4(z, 1.0f, 2.346487949187396e-1f, 1.187082088663404e-2f,
                                 2.150707366040937e-4f, 1.425433617130587e-6f) *
                           ln_hw(x) *
                           rcp_hw(poly2(1.0f - z, 9.847324170755358e-1f, 1.518396076767770e-2f,
                                        8.362215678646257e-5f));
        return poly4(z, 1.159315156584126e-1f, 2.770731240515333e-1f, 2.066458134619875e-2f,
                     4.574734709978264e-4f, 3.454715527986737e-6f) *
                   rcp_hw(poly2(1.0f - z, 9.836249671709183e-1f, 1.627693622304549e-2f,
                                9.809660603621949e-5f)) -
               term;
    }
    const float pp[8] = {1.253314137315499f, 1.475731032429900e1f, 6.123767403223466e1f,
                         1.121012633939949e2f, 9.285288485892228e1f, 3.198289277679660e1f,
                         3.595376024148513f, 6.160228690102976e-2f};
    const float qq[8] = {1.0f, 1.189963006673403e1f, 5.027773590829784e1f, 9.496513373427093e1f,
                         8.318077493230258e1f, 3.181399777449301e1f, 4.443672926432041f,
                         1.408295601966600e-1f};
    const float z = rcp_hw(x);
    return exp_hw(-x) * poly7(z, pp) * rcp_hw(poly7(z, qq)) * rsqrtf(x);
}

__device__ __forceinline__ float bessk1_hw(float x) {
    if (x <= 1.0f) {
        const float z = x * x;
        const float term = poly4(z, 0.5f, 5.598072040178741e-2f, 1.818666382168295e-3f,
                                 2.397509908859959e-5f, 1.239567816344855e-7f) *
                           ln_hw(x) *
                           rcp_hw(poly2(1.0f - z, 9.870202601341150e-1f, 1.292092053534579e-2f,
                                        5.881933053917096e-5f));
        return x * (poly4(z, -3.079657578292062e-1f, -8.109417631822442e-2f, -3.477550948593604e-3f,
                          -5.385594871975406e-5f, -3.110372465429008e-7f) *
                        rcp_hw(poly2(1.0f - z, 9.861813171751389e-1f, 1.375094061153160e-2f,
                                     6.774221332947002e-5f)) +
                    term) +
               rcp_hw(x);
    }
    const float pp[8] = {1.253314137315502f, 1.457171340220454e1f, 6.063161173098803e1f,
                         1.147386690867892e2f, 1.040442011439181e2f, 4.356596656837691e1f,
                         7.265230396353690f, 3.144418558991021e-1f};
    const float qq[8] = {1.0f, 1.125154514806458e1f, 4.427488496597630e1f, 7.616113213117645e1f,
                         5.863377227890893e1f, 1.850303673841586e1f, 1.857244676566022f,
                         2.538540887654872e-2f};
    const float z = rcp_hw(x);
    return exp_hw(-x) * poly7(z, pp) * rcp_hw(poly7(z, qq)) * rsqrtf(x);
}

// Terms per parameterisation (mse_set_parameterisation accepts no other count): three
// Lorentzian + three Gaussian pairs for Kirkland, five terms for Peng and Lobato.  As a
// compile-time count the term loops unroll and an atom's constants are loaded once.
template <int FORM>
constexpr int kTerms = FORM == 0 ? 3 : 5;

// Projected potentials from per-element constants precomputed on the host in double (the
// square roots and reciprocals of the parameters hoisted out of the per-point sums; the
// Gaussian exponents carry log2(e)).  Table layout per Z (kPotProjStride floats):
//   Kirkland  a_i | 2 pi sqrt(b_i) | c_i / d_i | pi^2 log2e / d_i
//   Peng      a_i / b_i | pi^2 log2e / b_i
//   Lobato    a_i b_i^-3/2 | 2 pi / sqrt(b_i)
constexpr int kPotProjStride = 16;

template <int FORM>
__device__ __forceinline__ float vproj_tab(const float* __restrict__ t, int, float r) {
    constexpr int IC = kTerms<FORM>;
    if constexpr (FORM == 0) {  // kernels/transmission_potentials_projected_f.cl:71-112
        float suml = 0.f, sumg = 0.f;
#pragma unroll
        for (int i = 0; i < IC; ++i) suml += t[i] * bessk0_hw(r * t[3 + i]);
        const float r2 = r * r;
#pragma unroll
        for (int i = 0; i < IC; ++i) sumg += t[6 + i] * __builtin_amdgcn_exp2f(-r2 * t[9 + i]);
        return 300.8242834f * suml + 150.4121417f * sumg;
    } else if constexpr (FORM == 1) {
        const float r2 = r * r;
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < IC; ++i) sum += t[i] * __builtin_amdgcn_exp2f(-r2 * t[5 + i]);
        return 150.4121417f * sum;
    } else {
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < IC; ++i) {
            const float x = r * t[5 + i];
            sum += t[i] * (bessk0_hw(x) + r * bessk1_hw(x));
        }
        return 945.090144399935f * sum;
    }
}

// Full-3D potentials (kernels/transmission_potentials_full_3d_f.cl:57-134), evaluated from
// per-element constants precomputed on the host in double: the square roots, reciprocals and
// b^-3/2 factors of the parameters are hoisted out of the per-point sums, and the exponent
// scales carry log2(e) so that every term is one hardware exp2 (v_exp_f32, about 1 ulp).
// r2 is the squared distance, already clamped to r_min^2.  Table layout per Z
// (kPot3dStride floats):
//   Kirkland  a_i | 2 pi log2e sqrt(b_i) | c_i d_i^-3/2 | pi^2 log2e / d_i
//   Peng      a_i b_i^-3/2 | pi^2 log2e / b_i
//   Lobato    a_i b_i^-3/2 | 2 pi log2e / sqrt(b_i) | sqrt(b_i) / pi
constexpr int kPot3dStride = 16;

template <int FORM>
__device__ __forceinline__ float v3d_tab(const float* __restrict__ t, int, float r2) {
    constexpr int IC = kTerms<FORM>;
    if constexpr (FORM == 0) {
        const float r = __builtin_amdgcn_sqrtf(r2);
        float suml = 0.f, sumg = 0.f;
#pragma unroll
        for (int i = 0; i < IC; ++i) suml += t[i] * __builtin_amdgcn_exp2f(-r * t[3 + i]);
#pragma unroll
        for (int i = 0; i < IC; ++i) sumg += t[6 + i] * __builtin_amdgcn_exp2f(-r2 * t[9 + i]);
        return 150.4121417f * __builtin_amdgcn_rcpf(r) * suml + 266.5157269f * sumg;
    } else if constexpr (FORM == 1) {
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < IC; ++i) sum += t[i] * __builtin_amdgcn_exp2f(-r2 * t[5 + i]);
        return 266.5157269f * sum;
    } else {
        const float r = __builtin_amdgcn_sqrtf(r2);
        const float inv_r = __builtin_amdgcn_rcpf(r);
        float sum = 0.f;
#pragma unroll
        for (int i = 0; i < IC; ++i)
            sum += t[i] * __builtin_amdgcn_exp2f(-r * t[5 + i]) * (t[10 + i] * inv_r + 1.0f);
        return 472.545072199968f * sum;
    }
}

struct PotArgs {
    int res;
    double pixelscale, startx, starty;
    double min_x, max_x, min_y, max_y;  // paddedFullLimits
    int blocks_x, blocks_y, load_x, load_y, load_z, n_slices;
    double cos_phi, sin_phi, sin_2theta, tan_theta;  // beam tilt geometry (projected)
    int untilted;        // no tilt: the tilted metric reduces to rad_x^2 + rad_y^2 bit for bit
    double sigma;        // sigma_eff
    double dz, z_min;    // full 3D
    double shift_x, shift_y;
    int integrals;
    int per_atom;        // parameters per Z (12 or 10)
    int i_count;
    // moving STEM frames: when set, blockIdx.z selects a frame whose potential-grid origin is
    // (frame_xy[2z], frame_xy[2z+1]) and every frame builds the same slice `first`
    const double* frame_xy;
    // tabulated potentials (potential_table.hpp), kBuckets cubic pieces per element from
    // element 1; read by the kernels instantiated with TAB = true
    const float4* vtab;
    // the full-3D kernels with LDS = true first copy into LDS, for each of the v_nel elements
    // present, the buckets whose index bits(r^2) >> kMantBits lies in [vb_lo, vb_lo + vb_n):
    // all that the slice's distances reach.  vslot[Z] is element Z's slot there and
    // vslot[vsel_at + s] the element of slot s.
    const int* vslot;
    int v_nel, vb_lo, vb_n, vsel_at;
    // per launch, from the host: 1 / (max - min) of the padded limits (the binning's own
    // factors), 1 / pixel size and (8 A + a margin of 0.01 A^2)^2 in pixels^2
    double rrx, rry;
    float inv_ps, reach_px2;
};

// An element's device table as an address biased so that the bucket of r2 sits at byte
// offset (bits(r2) >> (kMantBits - 4)) with the low four bits cleared: the bucket index plus
// (127 + kOctaveLo) << kSubBits, times 16.
// The element must be the same in every lane of the wave (the callers take it with
// readfirstlane from the staged atom): the base is then held in a scalar register pair that
// the compiler treats as opaque (an empty asm statement), so that it cannot split the bias
// off again, and the table loads take it as their scalar base with the bucket offset as the
// 32-bit vector offset.
__device__ __forceinline__ unsigned long long vtab_base(const float4* tz) {
    unsigned long long b = reinterpret_cast<unsigned long long>(tz) -
                           ((unsigned long long)((127 + mse_tab::kOctaveLo) << mse_tab::kSubBits) << 4);
    asm volatile("" : "+s"(b));
    return b;
}

// Squared distances are clamped below by at least this before a table lookup, so that the
// bucket offset never falls below the element's table.
constexpr float kVtabFloor = 1.0f / (float)(1u << -mse_tab::kOctaveLo);

// V(r2) of one element from its device table (base from vtab_base), for r2 >= kVtabFloor.
// Above the table's range the offset is clamped to the last bucket, so that no r2 (NaN
// included) reads past the element's table.  The device table holds each bucket's cubic in
// the integer m = t 2^kMantBits of the low mantissa bits, as (c0, c1 2^-kMantBits,
// c2 2^-2kMantBits, c3 2^-3kMantBits): scaling by powers of two is exact in every rounding
// step, so the Horner sum in m equals ((c3 t + c2) t + c1) t + c0 bit for bit while saving
// the multiply that forms t.
__device__ __forceinline__ float vtab_eval(unsigned long long base, float r2) {
    constexpr unsigned top = (unsigned)(((127 + mse_tab::kOctaveHi) << mse_tab::kSubBits) - 1) << 4;
    const unsigned bits = __float_as_uint(r2);
    const unsigned off = min((bits >> (mse_tab::kMantBits - 4)) & ~15u, top);
    // a global-memory load at base + off (the address space 1 pointer tells the compiler so)
    using gfloat4 = const __attribute__((address_space(1))) float4;
    const float4 c = *(const float4*)reinterpret_cast<gfloat4*>(base + off);
    const float m = (float)(bits & ((1u << mse_tab::kMantBits) - 1u));
    return fmaf(fmaf(fmaf(c.w, m, c.z), m, c.y), m, c.x);
}

// V(r2) as vtab_eval computes it, for kVtabFloor <= r2 < 2^kOctaveHi, where the offset needs
// no clamp.
__device__ __forceinline__ float vtab_eval_in(unsigned long long base, float r2) {
    const unsigned bits = __float_as_uint(r2);
    const unsigned off = (bits >> (mse_tab::kMantBits - 4)) & ~15u;
    using gfloat4 = const __attribute__((address_space(1))) float4;
    const float4 c = *(const float4*)reinterpret_cast<gfloat4*>(base + off);
    const float m = (float)(bits & ((1u << mse_tab::kMantBits) - 1u));
    return fmaf(fmaf(fmaf(c.w, m, c.z), m, c.y), m, c.x);
}

// V(r2) as vtab_eval computes it, from the LDS copy of an element's buckets
// [vb_lo, vb_lo + vb_n), which starts at lt[off + vb_lo]; r2 must lie in those buckets
// (the caller clamps it to them).
__device__ __forceinline__ float vtab_eval_lds(const float4* __restrict__ lt, int off, float r2) {
    const unsigned bits = __float_as_uint(r2);
    // the bucket index, shifted in an asm statement of its own so that the compiler folds the
    // byte scaling and the base into one shift-and-add rather than shift, mask and add
    unsigned q;
    asm("v_lshrrev_b32 %0, %1, %2" : "=v"(q) : "i"(mse_tab::kMantBits), "v"(bits));
    const float4 c = lt[(int)q + off];
    const float m = (float)(bits & ((1u << mse_tab::kMantBits) - 1u));
    return fmaf(fmaf(fmaf(c.w, m, c.z), m, c.y), m, c.x);
}

// The arguments of the plane blockIdx.z of a launch over slices [first, first + gridDim.z),
// or over frames of slice `first` (frame_xy set); returns the slice.
__device__ __forceinline__ int plane_args(const PotArgs& a, int first, PotArgs& b) {
    b = a;
    if (a.frame_xy) {
        b.startx = a.frame_xy[2 * blockIdx.z];
        b.starty = a.frame_xy[2 * blockIdx.z + 1];
        return first;
    }
    return first + (int)blockIdx.z;
}

// The blocks of the 16x16 pixel groups gx..gx + nx - 1, gy..gy + ny - 1 (the union of their
// ranges: the range's ends are monotone in the group)
__device__ __forceinline__ void tile_block_range(const PotArgs& a, int gx, int gy, int& starti,
                                                 int& endi, int& startj, int& endj, int nx = 1, int ny = 1) {
    // kernels/transmission_potentials_projected_d.cl:227-246 with a 16x16 group
    const double gsx = 16 * a.pixelscale, gsy = 16 * a.pixelscale;
    const double group_start_x = a.startx + gx * gsx;
    const double group_end_x = (a.startx + (gx + nx - 1) * gsx) + gsx;
    const double group_start_y = a.starty + gy * gsy;
    const double group_end_y = (a.starty + (gy + ny - 1) * gsy) + gsy;
    const double rrx = a.rrx, rry = a.rry;
    starti = (int)fmax(floor(a.blocks_x * (group_start_x - a.min_x) * rrx) - a.load_x, 0.0);
    endi = (int)fmin(ceil(a.blocks_x * (group_end_x - a.min_x) * rrx) + a.load_x, (double)(a.blocks_x - 1));
    startj = (int)fmax(floor(a.blocks_y * (group_start_y - a.min_y) * rry) - a.load_y, 0.0);
    endj = (int)fmin(ceil(a.blocks_y * (group_end_y - a.min_y) * rry) + a.load_y, (double)(a.blocks_y - 1));
}

// The same range cut down to the blocks that can hold an atom within `reach` A of the tile's
// pixels: an atom at x >= x_lo is binned (k_bin_atoms: floor((x - min_x) * (1 / (max_x -
// min_x)) * blocks_x)) into a block >= floor((x_lo - min_x) * (1 / (max_x - min_x)) *
// blocks_x), each rounding step being monotone, and likewise at the other ends.  Atoms of the
// blocks left out lie beyond the reach of every pixel of the tile: they add nothing and, in the
// full-3D trapezoid, carry nothing to the next atom.
__device__ __forceinline__ void tile_block_range_within(const PotArgs& a, int gx, int gy, double reach,
                                                        int& starti, int& endi, int& startj, int& endj,
                                                        int nx = 1, int ny = 1) {
    tile_block_range(a, gx, gy, starti, endi, startj, endj, nx, ny);
    const double rrx = a.rrx, rry = a.rry;
    const double x_lo = a.startx + (gx * 16) * a.pixelscale - reach;
    const double x_hi = a.startx + ((gx + nx) * 16 - 1) * a.pixelscale + reach;
    const double y_lo = a.starty + (gy * 16) * a.pixelscale - reach;
    const double y_hi = a.starty + ((gy + ny) * 16 - 1) * a.pixelscale + reach;
    starti = max(starti, (int)fmax(floor((x_lo - a.min_x) * rrx * a.blocks_x), -1.0));
    endi = min(endi, (int)fmin(floor((x_hi - a.min_x) * rrx * a.blocks_x), (double)a.blocks_x));
    startj = max(startj, (int)fmax(floor((y_lo - a.min_y) * rry * a.blocks_y), -1.0));
    endj = min(endj, (int)fmin(floor((y_hi - a.min_y) * rry * a.blocks_y), (double)a.blocks_y));
    if (starti > endi) {  // no block in reach: no rows (and no row reads)
        startj = 1;
        endj = 0;
    }
}

// ph - n 2pi with n the integer nearest ph / 2pi (2pi as the double 2.0 * M_PI).  The fma
// takes n 2pi exactly and rounds the difference once; remainder(ph, 2pi) is exactly
// representable, so the two agree whenever n is its quotient, which holds except where
// ph / 2pi lies within rounding of a half-integer (both results are then +-pi).  The phases
// here are small (a few hundred radians at most), so ph / 2pi is formed to far below 1/2.
__device__ __forceinline__ double reduce_2pi(double ph) {
    const double n = rint(ph * (1.0 / (2.0 * M_PI)));
    return fma(-n, 2.0 * M_PI, ph);
}

__device__ __forceinline__ float2 phase_to_t(double ph) {
    ph = reduce_2pi(ph);
    float s, c;
    sincosf((float)ph, &s, &c);
    return make_float2(c, s);
}

// The pixels of a thread in the builders' (16 PX)x(16 PY) pixel tiles (block (blockIdx.x,
// blockIdx.y), the 16x16 groups of the reference's tiling blockIdx.x * PX .. + PX - 1 by
// blockIdx.y * PY .. + PY - 1): wave w takes the (8 PX)x(8 PY) quadrant (w & 1, w >> 1) of
// the tile, a compact footprint for the waves' cull, and lane (lx, ly) = (l & 7, l >> 3) of
// it the pixels (lx + 8 i, ly + 8 j).
template <int PX, int PY>
__device__ __forceinline__ int tile_pixel_x(int i) {
    const int w = threadIdx.x >> 6;
    return (int)blockIdx.x * 16 * PX + (w & 1) * 8 * PX + (threadIdx.x & 7) + 8 * i;
}
template <int PX, int PY>
__device__ __forceinline__ int tile_pixel_y(int j) {
    const int w = threadIdx.x >> 6;
    return (int)blockIdx.y * 16 * PY + (w >> 1) * 8 * PY + ((threadIdx.x >> 3) & 7) + 8 * j;
}

// Projected potential sums (sumz) of slice `slice` at PX x PY pixels per thread, sumz[j PX + i]
// for pixel (tile_pixel_x(i), tile_pixel_y(j)).  The atoms of the tile's block rows are staged
// through LDS as one sequence, in chunks of 256, with no limit on how many a row holds.  (The
// tables stay in global memory here: one evaluation per atom and pixel is too little work to
// repay an LDS copy per block.)
template <int FORM, bool TAB, int PX = 1, int PY = 1>
__device__ __forceinline__ void projected_sumz(const double* __restrict__ ax, const double* __restrict__ ay,
                                               const int* __restrict__ aZ, const int* __restrict__ bsp,
                                               const float* __restrict__ ptab, const PotArgs& a, int slice,
                                               double* sx, double* sy, int* sZ, double (&sumz)[PX * PY]) {
#pragma clang fp contract(off)
    const int lid = threadIdx.x;
    int starti, endi, startj, endj;
    // untilted, a pixel reaches r <= 8 A (r2 <= 64.0000000001) laterally
    if (a.untilted)
        tile_block_range_within(a, blockIdx.x * PX, blockIdx.y * PY, 8.01, starti, endi, startj, endj, PX, PY);
    else
        tile_block_range(a, blockIdx.x * PX, blockIdx.y * PY, starti, endi, startj, endj, PX, PY);
    int k = slice < 0 ? 0 : (slice >= a.n_slices ? a.n_slices - 1 : slice);
    double im_pos_x[PX], im_pos_y[PY];
#pragma unroll
    for (int i = 0; i < PX; ++i) im_pos_x[i] = a.startx + tile_pixel_x<PX, PY>(i) * a.pixelscale;
#pragma unroll
    for (int j = 0; j < PY; ++j) im_pos_y[j] = a.starty + tile_pixel_y<PX, PY>(j) * a.pixelscale;
    const double r_min = 0.25 * a.pixelscale;
    const double r_min2 = r_min * r_min;
    const float r_minf = (float)r_min;
    const int wave = __builtin_amdgcn_readfirstlane(lid >> 6);
    // the tile's first pixel, 1 / pixel size and (8 A + margin)^2 in pixels^2, for the waves' reach
    const double tile_x0 = a.startx + (int)(blockIdx.x * PX * 16) * a.pixelscale;
    const double tile_y0 = a.starty + (int)(blockIdx.y * PY * 16) * a.pixelscale;
    const float inv_ps = a.inv_ps;
    const float reach_px2 = a.reach_px2;
#pragma unroll
    for (int i = 0; i < PX * PY; ++i) sumz[i] = 0.0;
    const int nb = a.blocks_x * a.blocks_y;
    // The atoms of block rows startj..endj (row j: sorted positions [bsp(j, starti),
    // bsp(j, endi + 1))) are staged as one sequence in row order, 256 at a time, so that rows
    // holding a few atoms each share a staging round.  Nothing carries from one atom to the
    // next, and every pixel sums its atoms in the reference's order.
    const int* brow = bsp + (size_t)k * nb;
    int j = startj, p = 0, re = 0;
    if (j <= endj) {
        p = brow[a.blocks_x * j + starti];
        re = brow[a.blocks_x * j + endi + 1];
    }
    while (j <= endj) {
        // the next (up to) 256 atoms: lane lid takes the one at position lid of the chunk
        int mine = -1, cnt = 0;
        while (cnt < 256 && j <= endj) {
            const int take = min(256 - cnt, re - p);
            if (lid >= cnt && lid < cnt + take) mine = p + (lid - cnt);
            cnt += take;
            p += take;
            if (p == re && ++j <= endj) {
                p = brow[a.blocks_x * j + starti];
                re = brow[a.blocks_x * j + endi + 1];
            }
        }
        __syncthreads();
        if (mine >= 0) {
            const int c = mine;
            const double xa = ax[c], ya = ay[c];
            sx[lid] = xa;
            sy[lid] = ya;
            // bits 16..19: the waves (pixel quadrants of the tile) that may have a pixel
            // within 8 A; the distance to each wave's pixel rectangle, in pixels and in
            // float, is tested against 64.01 A^2, a margin far above the float error.
            // Untilted only (a tilted metric is no lateral distance): all waves otherwise.
            unsigned waves = 15u;
            if (a.untilted) {
                waves = 0;
                const float u = (float)(xa - tile_x0) * inv_ps;
                const float v = (float)(ya - tile_y0) * inv_ps;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const float u0 = (float)((w & 1) * 8 * PX), v0 = (float)((w >> 1) * 8 * PY);
                    const float du = fmaxf(fmaxf(u0 - u, u - (u0 + (8.f * PX - 1.f))), 0.f);
                    const float dv = fmaxf(fmaxf(v0 - v, v - (v0 + (8.f * PY - 1.f))), 0.f);
                    if (du * du + dv * dv <= reach_px2) waves |= 1u << w;
                }
            }
            sZ[lid] = aZ[c] | (int)(waves << 16);
        }
        __syncthreads();
        for (int l = 0; l < cnt; ++l) {
            const int info = __builtin_amdgcn_readfirstlane(sZ[l]);
            if (!((info >> (16 + wave)) & 1)) continue;
            const int zs = info & 0xffff;
            // the element's constants (or table), loaded ahead of the distance arithmetic
            float pt[TAB ? 1 : kPotProjStride];
            unsigned long long tz = 0;
            if constexpr (TAB) {
                tz = vtab_base(a.vtab + (size_t)(zs - 1) * mse_tab::kBuckets);
            } else {
                const float* row = ptab + (size_t)(zs - 1) * kPotProjStride;
#pragma unroll
                for (int q = 0; q < kPotProjStride; ++q) pt[q] = row[q];
            }
            const double xl = sx[l], yl = sy[l];
#pragma unroll
            for (int jy = 0; jy < PY; ++jy) {
                const double rad_y = im_pos_y[jy] - yl;
#pragma unroll
                for (int i = 0; i < PX; ++i) {
                    const double rad_x = im_pos_x[i] - xl;
                    double r2;
                    if (a.untilted) {
                        // theta = pi/2: z' = O(1e-16) rad and z'/tan(theta) = O(1e-32) rad
                        // vanish below the rounding of x'^2 + y'^2, which equals this exactly
                        r2 = rad_x * rad_x + rad_y * rad_y;
                    } else {
                        const double z_prime = -0.5 * (rad_x * a.cos_phi + rad_y * a.sin_phi) * a.sin_2theta;
                        const double zt = z_prime / a.tan_theta;
                        const double x_prime = rad_x + zt * a.cos_phi;
                        const double y_prime = rad_y + zt * a.sin_phi;
                        r2 = z_prime * z_prime + x_prime * x_prime + y_prime * y_prime;
                    }
                    // rad = max(sqrt(r2), r_min) <= 8 with r2 <= 64.0000000001 (the
                    // reference's test, correctly rounded double sqrt): sqrt(64 + 2^-46) =
                    // 8 + 2^-50 - 2^-103... rounds to 8 while sqrt of the next double,
                    // 64 + 2^-45, rounds to 8 + 2^-49, so the test is exactly
                    // r2 <= 64 + 2^-46.  The potential takes rad in float.
                    if (r2 <= 0x1.0000000000001p+6) {
                        if constexpr (TAB) {
                            // r_min^2 is above the tables' floor (the host enables them only
                            // then) and r2 <= 64 + 2^-46 is 64 or less in float, far inside
                            // their top: the lookup needs no clamp
                            const float f = r2 < r_min2 ? r_minf * r_minf : (float)r2;
                            sumz[jy * PX + i] += (double)vtab_eval_in(tz, f);
                        } else {
                            const float rf = r2 < r_min2 ? r_minf : __builtin_amdgcn_sqrtf((float)r2);
                            sumz[jy * PX + i] += (double)vproj_tab<FORM>(pt, a.i_count, rf);
                        }
                    }
                }
            }
        }
    }
}

// The LDS copy of the tables a kernel instantiated with LDS = true reads: buckets
// [vb_lo, vb_lo + vb_n) of each element present, slot after slot (see PotArgs::vslot).  The
// first staging round of the sums synchronises the block before any table read.
__device__ __forceinline__ void fill_lds_tables(const PotArgs& a, float4* lt) {
    const int total = a.v_nel * a.vb_n;
    const int base = a.vb_lo - ((127 + mse_tab::kOctaveLo) << mse_tab::kSubBits);
    for (int i = threadIdx.x; i < total; i += 256) {
        const int s = i / a.vb_n;
        const int Z = a.vslot[a.vsel_at + s];
        lt[i] = a.vtab[(size_t)(Z - 1) * mse_tab::kBuckets + base + (i - s * a.vb_n)];
    }
}

// Transmission function exp(i sigma V) of slices [first, first + gridDim.z)
// (kernels/transmission_potentials_projected_f.cl:183-317); out has one res*res plane per
// slice (relative to first).  Grid (res / (16 kProjPX), res / (16 kProjPY), planes):
// (16 kProjPX)x(16 kProjPY) pixel tiles, 16x16 pixels per wave.
constexpr int kProjPX = 2, kProjPY = 2;
template <int FORM, bool TAB = false>
__global__ __launch_bounds__(256) void k_potential_projected(
    float2* __restrict__ out, const double* __restrict__ ax, const double* __restrict__ ay,
    const int* __restrict__ aZ, const int* __restrict__ bsp, const float* __restrict__ ptab,
    PotArgs a, int first) {
    __shared__ double sx[256], sy[256];
    __shared__ int sZ[256];
    PotArgs b;
    const int slice = plane_args(a, first, b);
    // kProjPX x kProjPY pixels per thread: the atom's staging, loads and wave test serve them all
    double sumz[kProjPX * kProjPY];
    projected_sumz<FORM, TAB, kProjPX, kProjPY>(ax, ay, aZ, bsp, ptab, b, slice, sx, sy, sZ, sumz);
#pragma unroll
    for (int j = 0; j < kProjPY; ++j) {
        const int yid = tile_pixel_y<kProjPX, kProjPY>(j);
#pragma unroll
        for (int i = 0; i < kProjPX; ++i) {
            const int xid = tile_pixel_x<kProjPX, kProjPY>(i);
            out[(size_t)blockIdx.z * a.res * a.res + (size_t)yid * a.res + xid] =
                phase_to_t(a.sigma * sumz[j * kProjPX + i]);
        }
    }
}

// The phase sigma V of the same slices, reduced to [-pi, pi] as phase_to_t reduces it, so
// that sincosf of the stored float reproduces the transmission function's exp(i sigma V)
// before the band limit (used by the adjoint).  Grid (res / 16, res / 16, planes).
template <int FORM, bool TAB = false>
__global__ __launch_bounds__(256) void k_phase_projected(
    float* __restrict__ out, const double* __restrict__ ax, const double* __restrict__ ay,
    const int* __restrict__ aZ, const int* __restrict__ bsp, const float* __restrict__ ptab,
    PotArgs a, int first) {
    __shared__ double sx[256], sy[256];
    __shared__ int sZ[256];
    double sumz[1];
    projected_sumz<FORM, TAB, 1, 1>(ax, ay, aZ, bsp, ptab, a, first + (int)blockIdx.z, sx, sy, sZ, sumz);
    const int xid = tile_pixel_x<1, 1>(0);
    const int yid = tile_pixel_y<1, 1>(0);
    out[(size_t)blockIdx.z * a.res * a.res + (size_t)yid * a.res + xid] =
        (float)reduce_2pi(a.sigma * sumz[0]);
}

// transmission_potentials_full_3d: trapezoid over integrals+1 sub-slice points for the
// slices k +- load_z; p2 carries between atoms of one block row exactly as the reference
// kernel does (kernels/transmission_potentials_full_3d_f.cl:229-278).  Sums sumz of slice
// `slice` at NX pixels per thread, sumz[i] for pixel (tile_pixel_x<NX, 1>(i),
// tile_pixel_y<NX, 1>(0)) of the block's (16 NX)x16 pixel tile; each pixel keeps its own
// carried p2.
template <int FORM, bool TAB = false, bool LDS = false, int NX = 1>
__device__ __forceinline__ void full3d_sumz(const double* __restrict__ ax, const double* __restrict__ ay,
                                            const double* __restrict__ az, const int* __restrict__ aZ,
                                            const int* __restrict__ bsp, const float* __restrict__ tab3d,
                                            const PotArgs& a, int slice, double* sx, double* sy, double* sz,
                                            int* sZ, int* sh0, float* srz, const float4* lt, double (&sumz)[NX]) {
#pragma clang fp contract(off)
    const int lid = threadIdx.x;
    const int yid = tile_pixel_y<NX, 1>(0);
    int starti, endi, startj, endj;
    // a pixel reaches 8 A laterally, plus the beam-tilt shift over the integration points
    const double drift = (a.integrals + 1) * fmax(fabs(a.shift_x), fabs(a.shift_y));
    tile_block_range_within(a, blockIdx.x * NX, blockIdx.y, 8.01 + drift, starti, endi, startj, endj, NX);
    int topz = slice - a.load_z, bottomz = slice + a.load_z;
    if (topz < 0) topz = 0;
    if (bottomz >= a.n_slices) bottomz = a.n_slices - 1;
    const double int_r = 1.0 / a.integrals;
    const double current_z = a.z_min + (a.n_slices - slice) * a.dz;
    double im_pos_x[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) im_pos_x[i] = a.startx + tile_pixel_x<NX, 1>(i) * a.pixelscale;
    const double im_pos_y = a.starty + yid * a.pixelscale;
    const double r_min = 0.25 * a.pixelscale;
    // distances below r_min are evaluated at r_min (squared, as the potentials take r^2)
    const double r_min2 = r_min * r_min;
    const float r_min2f = (float)r_min * (float)r_min;
    const bool no_shift = (a.shift_x == 0.0 && a.shift_y == 0.0);
    // table lookups: r^2 at least r_min^2 (and the tables' floor, which r_min^2 meets on every
    // geometry the tables accept); with the tables in LDS also at most the top of the last
    // bucket there, above every distance an untilted slice reaches
    const float r_lo2 = fmaxf(r_min2f, kVtabFloor);
    const float r_hi2 = __uint_as_float(((unsigned)(a.vb_lo + a.vb_n) << mse_tab::kMantBits) - 1u);
    const int wave = __builtin_amdgcn_readfirstlane(lid >> 6);
    // the tile's first pixel, 1 / pixel size and (8 A + margin)^2 in pixels^2, for the waves' reach
    const double tile_x0 = a.startx + (int)(blockIdx.x * NX * 16) * a.pixelscale;
    const double tile_y0 = a.starty + (int)(blockIdx.y * 16) * a.pixelscale;
    const float inv_ps = a.inv_ps;
    const float reach_px2 = a.reach_px2;
#pragma unroll
    for (int i = 0; i < NX; ++i) sumz[i] = 0.0;
    const int nb = a.blocks_x * a.blocks_y;
    // The atoms of the rows (k, j), k = topz..bottomz, j = startj..endj in that order (row
    // (k, j): sorted positions [bsp(k, j, starti), bsp(k, j, endi + 1))) are staged as one
    // sequence, 256 at a time, so that rows holding a few atoms each share a staging round.
    // Bit 24 of sh0 marks the first atom of a row, where p2 starts again from 0; bit 25 an
    // atom whose points may reach past the LDS tables' top bucket (see the clamp below).
    constexpr int kRowStart = 1 << 24, kDeep = 1 << 25;
    double p2[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) p2[i] = 0.0;
    bool row_fresh = false;
    int k = topz, j = startj, p = 0, re = 0;
    bool more = topz <= bottomz && startj <= endj, fresh = true;
    if (more) {
        p = bsp[k * nb + a.blocks_x * j + starti];
        re = bsp[k * nb + a.blocks_x * j + endi + 1];
    }
    while (more) {
        // the next (up to) 256 atoms: lane lid takes the one at position lid of the chunk
        int mine = -1, cnt = 0;
        bool mine_first = false;
        while (cnt < 256 && more) {
            const int take = min(256 - cnt, re - p);
            if (lid >= cnt && lid < cnt + take) {
                mine = p + (lid - cnt);
                mine_first = fresh && lid == cnt;
            }
            if (take > 0) fresh = false;
            cnt += take;
            p += take;
            if (p == re) {
                if (++j > endj) {
                    j = startj;
                    more = ++k <= bottomz;
                }
                if (more) {
                    p = bsp[k * nb + a.blocks_x * j + starti];
                    re = bsp[k * nb + a.blocks_x * j + endi + 1];
                    fresh = true;
                }
            }
        }
        __syncthreads();
        if (mine >= 0) {
            const int c = mine;
            sx[lid] = ax[c];
            sy[lid] = ay[c];
            const double zl = az[c];
            sz[lid] = zl;
            // with the tables in LDS, the element's slot there
            sZ[lid] = LDS ? a.vslot[aZ[c]] : aZ[c];
            int info_w = mine_first ? kRowStart : 0;
            if (no_shift) {
                // rad_z(h) = (current_z - h dz / integrals) - z falls with h (each
                // rounding step is monotone), so the points with rad_z <= 3 are
                // h0..integrals; h0 is found with the very expression of the per-point
                // test, once per atom for the whole tile, with the depth rad_z(h0)
                auto within = [&](int h) { return (current_z - h * a.dz * int_r) - zl <= 3.0; };
                const double h_est = ceil((current_z - zl - 3.0) / (a.dz * int_r));
                int h0 = (int)fmin(fmax(h_est, 0.0), (double)(a.integrals + 1));
#pragma clang loop unroll(disable) vectorize(disable) interleave(disable)
                while (h0 > 0 && within(h0 - 1)) --h0;
#pragma clang loop unroll(disable) vectorize(disable) interleave(disable)
                while (h0 <= a.integrals && !within(h0)) ++h0;
                // bit 16 + w: wave w (its (8 NX)x8 pixel quadrant of the tile) may have
                // a pixel within 8 A.  The atom's distance to the wave's pixel
                // rectangle, in pixels and in float, is tested against 64.01 A^2: the
                // float error (~1e-6 A) is far below that margin, so a clear bit means
                // every lane of the wave fails its own exact test xyrad2 <= 64.
                unsigned waves = 0;
                if (h0 <= a.integrals) {
                    const float u = (float)(sx[lid] - tile_x0) * inv_ps;
                    const float v = (float)(sy[lid] - tile_y0) * inv_ps;
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        const float u0 = (float)((w & 1) * 8 * NX), v0 = (float)((w >> 1) * 8);
                        const float du = fmaxf(fmaxf(u0 - u, u - (u0 + (8.f * NX - 1.f))), 0.f);
                        const float dv = fmaxf(fmaxf(v0 - v, v - (v0 + 7.f)), 0.f);
                        if (du * du + dv * dv <= reach_px2) waves |= 1u << w;
                    }
                }
                info_w |= h0 | (int)(waves << 16);
                const float rz0 = (float)((current_z - h0 * a.dz * int_r) - zl);
                srz[lid] = rz0;
                if constexpr (LDS) {
                    const float stepf = (float)(a.dz * int_r);
                    const float zmax = (fabsf(rz0) + (float)(a.integrals - h0) * stepf) * 1.001f;
                    if (!(fmaf(zmax, zmax, 64.f) <= r_hi2)) info_w |= kDeep;
                }
            }
            sh0[lid] = info_w;
        }
        __syncthreads();
        for (int l = 0; l < cnt; ++l) {
            // the atom's first point (the same in every lane); atoms out of this
            // wave's reach, in depth or laterally, are passed over by the whole wave
            int h0 = 0;
            const int info = __builtin_amdgcn_readfirstlane(sh0[l]);
            // a new row restarts p2; atoms the wave passes over leave p2 as it is, so the
            // restart is applied at the row's first atom the wave takes
            row_fresh = row_fresh || (info & kRowStart);
            if (no_shift) {
                if (!((info >> (16 + wave)) & 1)) continue;
                h0 = info & 0xffff;
            }
            if (row_fresh) {
#pragma unroll
                for (int i = 0; i < NX; ++i) p2[i] = 0.0;
                row_fresh = false;
            }
            const double xl = sx[l];
            const double rad_y0 = im_pos_y - sy[l];
            // the element's constants (scalar registers) or table, loaded once per atom
            float pz[TAB ? 1 : kPot3dStride];
            unsigned long long tz = 0;
            int loff = 0;
            if constexpr (LDS) {
                loff = __builtin_amdgcn_readfirstlane(sZ[l]) * a.vb_n - a.vb_lo;
            } else if constexpr (TAB) {
                tz = vtab_base(a.vtab + (size_t)(__builtin_amdgcn_readfirstlane(sZ[l]) - 1) * mse_tab::kBuckets);
            } else {
                const float* row = tab3d + (size_t)(sZ[l] - 1) * kPot3dStride;
#pragma unroll
                for (int q = 0; q < kPot3dStride; ++q) pz[q] = row[q];
            }
            if (no_shift) {
                // the xy distance of each pixel is the same at every integration point;
                // a pixel beyond 8 A takes no part (no sum, no carried p2)
                bool in[NX];
                float xy2f[NX];
                bool any = false, low = false;
#pragma unroll
                for (int i = 0; i < NX; ++i) {
                    const double rad_x = im_pos_x[i] - xl;
                    const double xyrad2 = rad_x * rad_x + rad_y0 * rad_y0;
                    in[i] = xyrad2 <= 64.0;
                    // a pixel that takes no part is evaluated at the smallest table
                    // distance, so its discarded lookups stay within the tables
                    xy2f[i] = in[i] ? (float)xyrad2 : r_lo2;
                    any = any || in[i];
                    low = low || xy2f[i] < r_lo2;
                }
                if (!any) continue;
                // from there the depth steps in float and the atom's trapezoid terms are
                // summed in float, then added to the pixel's double sum once
                const float stepf = (float)(a.dz * int_r);
                const float rz0 = srz[l];
                if constexpr (TAB) {
                    // points h0..integrals at depths rz0 - (h - h0) step; the trapezoid
                    //   sum_{h > 0} (p(h) + p(h - 1)) = 2 sum_h p(h) - p(integrals) + c,
                    // c = p2 (carried from the row's previous atom) when h0 > 0 and
                    // -p(0) when h0 = 0, summed as 2 s - p(last) + (2 p(h0) + c)
                    // with s the sum over h0 + 1..integrals.  The pixels share the depths;
                    // one beyond 8 A is evaluated (within the table) and discarded.
                    //
                    // With the tables in LDS the clamp to the staged range is left out
                    // when no lane of the wave needs it: every r^2 = fma(rz, rz, xy2) is
                    // at least xy2 >= r_lo2, and at most 64 + zmax^2, zmax bounding |rz|
                    // over the atom's points (|rz0| + its depth span; the 1e-3 margin
                    // covers the float steps' rounding, < 4096 * 2^-24 relative).  The
                    // bound is the same in every lane, tested once per atom at staging
                    // (kDeep).
                    bool clamp = true;
                    if constexpr (LDS) clamp = (info & kDeep) || __builtin_amdgcn_ballot_w64(low) != 0;
                    auto sweep = [&](auto clamped) {
                        auto eval = [&](float r2) {
                            if constexpr (LDS)
                                return vtab_eval_lds(lt, loff, decltype(clamped)::value ? __builtin_amdgcn_fmed3f(r2, r_lo2, r_hi2) : r2);
                            else
                                return vtab_eval(tz, fmaxf(r2, r_lo2));
                        };
                        float rz = rz0;
                        float pf[NX], s[NX], pl[NX];
#pragma unroll
                        for (int i = 0; i < NX; ++i) {
                            pf[i] = eval(fmaf(rz, rz, xy2f[i]));
                            s[i] = 0.f;
                            pl[i] = pf[i];
                        }
                        // four points per trip (the lookups' asm keeps the compiler from
                        // unrolling), so that their table reads are in flight together;
                        // added in order
                        int h = h0;
                        for (; h + 3 < a.integrals; h += 4) {
                            float rzk[4];
                            rzk[0] = rz - stepf;
                            rzk[1] = rzk[0] - stepf;
                            rzk[2] = rzk[1] - stepf;
                            rzk[3] = rzk[2] - stepf;
                            rz = rzk[3];
                            if constexpr (LDS) {
                                // the same lookups as eval, in phases: every bucket index,
                                // then every table read, then the cubics
                                float r2k[4][NX];
                                unsigned qk[4][NX];
                                float4 ck[4][NX];
#pragma unroll
                                for (int k = 0; k < 4; ++k)
#pragma unroll
                                    for (int i = 0; i < NX; ++i) {
                                        const float r2 = fmaf(rzk[k], rzk[k], xy2f[i]);
                                        r2k[k][i] = decltype(clamped)::value ? __builtin_amdgcn_fmed3f(r2, r_lo2, r_hi2) : r2;
                                        asm("v_lshrrev_b32 %0, %1, %2" : "=v"(qk[k][i]) : "i"(mse_tab::kMantBits), "v"(__float_as_uint(r2k[k][i])));
                                    }
#pragma unroll
                                for (int k = 0; k < 4; ++k)
#pragma unroll
                                    for (int i = 0; i < NX; ++i) ck[k][i] = lt[(int)qk[k][i] + loff];
#pragma unroll
                                for (int i = 0; i < NX; ++i)
#pragma unroll
                                    for (int k = 0; k < 4; ++k) {
                                        const float4 c = ck[k][i];
                                        const float m = (float)(__float_as_uint(r2k[k][i]) & ((1u << mse_tab::kMantBits) - 1u));
                                        const float p = fmaf(fmaf(fmaf(c.w, m, c.z), m, c.y), m, c.x);
                                        s[i] += p;
                                        if (k == 3) pl[i] = p;
                                    }
                            } else {
#pragma unroll
                                for (int i = 0; i < NX; ++i) {
                                    const float pa = eval(fmaf(rzk[0], rzk[0], xy2f[i]));
                                    const float pb = eval(fmaf(rzk[1], rzk[1], xy2f[i]));
                                    const float pc = eval(fmaf(rzk[2], rzk[2], xy2f[i]));
                                    pl[i] = eval(fmaf(rz, rz, xy2f[i]));
                                    s[i] += pa;
                                    s[i] += pb;
                                    s[i] += pc;
                                    s[i] += pl[i];
                                }
                            }
                        }
#pragma clang loop unroll(disable)
                        for (; h < a.integrals; ++h) {
                            rz -= stepf;
#pragma unroll
                            for (int i = 0; i < NX; ++i) {
                                pl[i] = eval(fmaf(rz, rz, xy2f[i]));
                                s[i] += pl[i];
                            }
                        }
#pragma unroll
                        for (int i = 0; i < NX; ++i) {
                            if (!in[i]) continue;
                            const float acc = (2.f * s[i] - pl[i]) + (h0 > 0 ? 2.f * pf[i] + (float)p2[i] : pf[i]);
                            sumz[i] += 0.5 * (double)acc;
                            p2[i] = (double)pl[i];
                        }
                    };
                    if (!LDS || clamp)
                        sweep(std::true_type{});
                    else
                        sweep(std::false_type{});
                    continue;
                }
#pragma unroll
                for (int i = 0; i < NX; ++i) {
                    if (!in[i]) continue;
                    float p2f = (float)p2[i], acc = 0.f;
                    for (int h = h0; h <= a.integrals; ++h) {
                        const float rz = fmaf(-(float)(h - h0), stepf, rz0);
                        const float r2 = fmaf(rz, rz, xy2f[i]);
                        const float p1 = v3d_tab<FORM>(pz, a.i_count, r2 < r_min2f ? r_min2f : r2);
                        if (h != 0) acc += p1 + p2f;
                        p2f = p1;
                    }
                    sumz[i] += 0.5 * (double)acc;
                    p2[i] = (double)p2f;
                }
                continue;
            }
#pragma unroll
            for (int i = 0; i < NX; ++i) {
                double rad_x = im_pos_x[i] - xl;
                double rad_y = rad_y0;
                for (int h = 0; h <= a.integrals; ++h) {
                    rad_x -= a.shift_x;
                    rad_y -= a.shift_y;
                    const double xyrad2 = rad_x * rad_x + rad_y * rad_y;
                    const double rad_z = (current_z - h * a.dz * int_r) - sz[l];
                    if (xyrad2 <= 64.0 && rad_z <= 3.0) {
                        const double r2 = xyrad2 + rad_z * rad_z;
                        float p1f;
                        if constexpr (LDS)  // not launched with a tilt; kept within the copy
                            p1f = vtab_eval_lds(lt, loff, __builtin_amdgcn_fmed3f(r2 < r_min2 ? r_min2f : (float)r2, r_lo2, r_hi2));
                        else if constexpr (TAB)
                            p1f = vtab_eval(tz, fmaxf(r2 < r_min2 ? r_min2f : (float)r2, kVtabFloor));
                        else
                            p1f = v3d_tab<FORM>(pz, a.i_count, r2 < r_min2 ? r_min2f : (float)r2);
                        const double p1 = (double)p1f;
                        sumz[i] += (h != 0) * (p1 + p2[i]) * 0.5;
                        p2[i] = p1;
                    }
                }
            }
        }
    }
}

// LDS = true (untilted beams only): the elements' tables are first copied into dynamic LDS
// (v_nel * vb_n * 16 bytes), where the integration points read them, and each thread sums
// kFull3dNX pixels (tile_pixel_x<kFull3dNX, 1>): grid (res / (16 kFull3dNX), res / 16, planes)
// then, (res / 16, res / 16, planes) otherwise.
constexpr int kFull3dNX = 2;
template <int FORM, bool TAB = false, bool LDS = false>
__global__ __launch_bounds__(256) void k_potential_full3d(
    float2* __restrict__ out, const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, const int* __restrict__ aZ, const int* __restrict__ bsp,
    const float* __restrict__ tab3d, PotArgs a, int first) {
    __shared__ double sx[256], sy[256], sz[256];
    __shared__ int sZ[256], sh0[256];
    __shared__ float srz[256];
    extern __shared__ float4 lt[];
    if constexpr (LDS) fill_lds_tables(a, lt);
    PotArgs b;
    const int slice = plane_args(a, first, b);
    constexpr int NX = LDS ? kFull3dNX : 1;
    double sumz[NX];
    full3d_sumz<FORM, TAB, LDS, NX>(ax, ay, az, aZ, bsp, tab3d, b, slice, sx, sy, sz, sZ, sh0, srz, lt, sumz);
    const int yid = tile_pixel_y<NX, 1>(0);
    const double sst = a.dz * (1.0 / a.integrals);
#pragma unroll
    for (int i = 0; i < NX; ++i) {
        const int xid = tile_pixel_x<NX, 1>(i);
        out[(size_t)blockIdx.z * a.res * a.res + (size_t)yid * a.res + xid] = phase_to_t(sst * a.sigma * sumz[i]);
    }
}

// The full-3D phase sst sigma sumz of the same slices, reduced to [-pi, pi] (used by the adjoint).
template <int FORM, bool TAB = false>
__global__ __launch_bounds__(256) void k_phase_full3d(
    float* __restrict__ out, const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, const int* __restrict__ aZ, const int* __restrict__ bsp,
    const float* __restrict__ tab3d, PotArgs a, int first) {
    __shared__ double sx[256], sy[256], sz[256];
    __shared__ int sZ[256], sh0[256];
    __shared__ float srz[256];
    double sumz1[1];
    full3d_sumz<FORM, TAB>(ax, ay, az, aZ, bsp, tab3d, a, first + blockIdx.z, sx, sy, sz, sZ, sh0, srz, nullptr, sumz1);
    const double sumz = sumz1[0];
    const int xid = tile_pixel_x<1, 1>(0);
    const int yid = tile_pixel_y<1, 1>(0);
    const double sst = a.dz * (1.0 / a.integrals);
    out[(size_t)blockIdx.z * a.res * a.res + (size_t)yid * a.res + xid] =
        (float)reduce_2pi(sst * a.sigma * sumz);
}

}  // namespace mse
