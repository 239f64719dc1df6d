// Adjoint kernels: the derivative of the projected potentials and the reduction of
// dL/d(phase) over pixels into per-atom position gradients (gfx950).
//
// The forward transmission of slice s is t = BL(exp(i phi)), phi(p) = sigma * sum_a v_a(r_ap)
// (projected potential sum of k_potential_projected).  With g_tau = BL(g_t) the gradient of a
// real loss with respect to exp(i phi), dL/dphi(p) = -Im(conj(g_tau(p)) exp(i phi(p))), and
//   dL/dx_a = sum_p dL/dphi(p) * sigma * v_a'(r_ap) * dr_ap/dx_a
// over exactly the pixels whose potential sum includes atom a (same tile block ranges, the
// same r2 <= 64 and r <= 8 tests; no derivative where the r_min clamp holds).
#pragma once
#include <hip/hip_runtime.h>

#include "atom_kernels.hpp"
#include "slice_kernels.hpp"

namespace mse {

// d/dr of the projected potentials of vproj_tab<FORM> (kernels/transmission_potentials_projected_f.cl:71-181),
// from the raw parameters (p = a1,b1,..,c1,d1,.. for Kirkland; a1..a5,b1..b5 otherwise):
//   Kirkland  300.82 sum a K0(2 pi r sqrt b) + 150.41 sum (c/d) exp(-(pi r)^2/d)
//             -> 300.82 sum a (-2 pi sqrt b) K1(2 pi r sqrt b) + 150.41 sum (c/d)(-2 pi^2 r/d) exp(..)
//   Peng      150.41 sum (a/b) exp(-(pi r)^2/b) -> 150.41 sum (a/b)(-2 pi^2 r/b) exp(..)
//   Lobato    945.09 sum a b^-3/2 (K0(c r) + r K1(c r)), c = 2 pi / sqrt b
//             -> 945.09 sum a b^-3/2 (-c)(K1(c r) + r K0(c r))   (K0' = -K1, (z K1)' = -z K0)
template <int FORM>
__device__ __forceinline__ float dvproj(const float* __restrict__ p, int, float r) {
    constexpr int ic = kTerms<FORM>;
    const float pi = (float)M_PI;
    if constexpr (FORM == 0) {
        const float x = 2.0f * pi * r;
        float suml = 0.f, sumg = 0.f;
        for (int i = 0; i < ic * 2; i += 2) {
            const float sb = sqrtf(p[i + 1]);
            suml += p[i] * (-2.0f * pi * sb) * bessk1f(x * sb);
        }
        const float xg = (pi * r) * (pi * r);
        for (int i = ic * 2; i < ic * 4; i += 2) {
            const float d_inv = 1.0f / p[i + 1];
            sumg += (p[i] * d_inv) * expf(-xg * d_inv) * (-2.0f * pi * pi * r * d_inv);
        }
        return 300.8242834f * suml + 150.4121417f * sumg;
    } else if constexpr (FORM == 1) {
        const float x = (pi * r) * (pi * r);
        float sum = 0.f;
        for (int i = 0; i < ic; ++i) {
            const float b_inv = 1.0f / p[i + 5];
            sum += p[i] * b_inv * expf(-x * b_inv) * (-2.0f * pi * pi * r * b_inv);
        }
        return 150.4121417f * sum;
    } else {
        const float x = 2.0f * pi * r;
        float sum = 0.f;
        for (int i = 0; i < ic; ++i) {
            const float bir = rsqrtf(p[i + 5]);
            const float c = 2.0f * pi * bir;
            sum += p[i] * (bir * bir * bir) * (-c) * (bessk1f(x * bir) + r * bessk0f(x * bir));
        }
        return 945.090144399935f * sum;
    }
}

// Sum of nw complex planes of N points (the per-wave g_t of one configuration).
static __global__ void k_sum_planes(const float2* __restrict__ in, int nw, size_t stride, size_t N,
                             float2* __restrict__ out) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    float2 s = in[p];
    for (int w = 1; w < nw; ++w) {
        const float2 v = in[(size_t)w * stride + p];
        s.x += v.x;
        s.y += v.y;
    }
    out[p] = s;
}

// out[0, n) = conj(px), out[n, 2n) = conj(py): conj(P') = conj(px) conj(py) is the
// propagator of the backward column pass (negation is exact, so the products equal those of
// the conjugated forward factor bit for bit).
static __global__ void k_conj_tables(const float2* __restrict__ px, const float2* __restrict__ py, int n,
                                     float2* __restrict__ out) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= 2 * n) return;
    const float2 v = i < n ? px[i] : py[i - n];
    out[i] = make_float2(v.x, -v.y);
}

static __global__ void k_fill_zero(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// dV/d(r2) of the cubic that vtab_eval_in evaluates at r2, for kVtabFloor <= r2 < 2^kOctaveHi:
// the derivative of what the tabulated forward sums.  With r2 = 2^e (1 + (i + m 2^-kMantBits)
// 2^-kSubBits), m the low kMantBits mantissa bits, dm/d(r2) = 2^(23 - e), the float of biased
// exponent 127 + 23 - e = 277 - (bits(r2) >> 23) (142..170 over the tables' range).
__device__ __forceinline__ float vtab_slope_in(unsigned long long base, float r2) {
    const unsigned bits = __float_as_uint(r2);
    const unsigned off = (bits >> (mse_tab::kMantBits - 4)) & ~15u;
    using gfloat4 = const __attribute__((address_space(1))) float4;
    const float4 c = *(const float4*)reinterpret_cast<gfloat4*>(base + off);
    const float m = (float)(bits & ((1u << mse_tab::kMantBits) - 1u));
    const float dv_dm = fmaf(fmaf(3.0f * c.w, m, 2.0f * c.z), m, c.y);
    return dv_dm * __uint_as_float((277u - (bits >> 23)) << 23);
}

// One block of 256 threads per sorted atom of slices 0..n_slices-1 (sorted range
// [first, first + gridDim.x)); its slice s is found from the slice starts bs[k * nblk].
// The block visits the pixels of the atom's 8 A neighbourhood, keeps those whose tile block
// range contains the atom's block (as the forward sum does), and reduces
//   -sigma * dL/dphi(p) * v'(r) * dr/d(rad) in double, with dL/dphi from slice s's planes
// g_tau + s N and phase + s N.  TAB: v'(r) / r = 2 dV/d(r2) of the element's table (the
// derivative of the tabulated forward, with its clamp r2 < r_min^2 and its cutoff), else the
// derivative of the formulas.  grad has 3 doubles (x, y, z) per input atom, indexed through
// orig.
template <int FORM, bool TAB = false>
__global__ __launch_bounds__(256) void k_atom_grad_projected(
    const double* __restrict__ ax, const double* __restrict__ ay, const int* __restrict__ aZ,
    const int* __restrict__ orig, int first, const int* __restrict__ bs, int nblk, int n_slices,
    const float* __restrict__ params, PotArgs a, const float2* __restrict__ g_tau_all,
    const float* __restrict__ phase_all, double* __restrict__ grad) {
#pragma clang fp contract(off)
    __shared__ double red[2][4];
    const int i = first + blockIdx.x;
    // the slice holding sorted atom i: the last k with bs[k * nblk] <= i
    int lo = 0, hi = n_slices - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bs[(size_t)mid * nblk] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    const size_t plane = (size_t)a.res * a.res;
    const float2* __restrict__ g_tau = g_tau_all + (size_t)lo * plane;
    const float* __restrict__ phase = phase_all + (size_t)lo * plane;
    const double xa = ax[i], ya = ay[i];
    const float* pz = params + (size_t)(aZ[i] - 1) * a.per_atom;
    // the atom's xy block, as k_bin_atoms computes it
    int bidx = (int)floor((xa - a.min_x) * (1.0 / (a.max_x - a.min_x)) * a.blocks_x);
    int bidy = (int)floor((ya - a.min_y) * (1.0 / (a.max_y - a.min_y)) * a.blocks_y);
    bidx -= (bidx == a.blocks_x);
    bidy -= (bidy == a.blocks_y);
    // pixel window: every pixel within 8 A (plus a margin for the tilted metric)
    const double reach = 8.5;
    const int x0 = max(0, (int)floor((xa - reach - a.startx) / a.pixelscale));
    const int x1 = min(a.res - 1, (int)ceil((xa + reach - a.startx) / a.pixelscale));
    const int y0 = max(0, (int)floor((ya - reach - a.starty) / a.pixelscale));
    const int y1 = min(a.res - 1, (int)ceil((ya + reach - a.starty) / a.pixelscale));
    const int wx = x1 - x0 + 1, wy = y1 - y0 + 1;
    const int npix = (wx > 0 && wy > 0) ? wx * wy : 0;
    const double r_min = 0.25 * a.pixelscale;
    const double r_min2 = r_min * r_min;
    unsigned long long tz = 0;
    if constexpr (TAB) tz = vtab_base(a.vtab + (size_t)(aZ[i] - 1) * mse_tab::kBuckets);
    // d(z', x', y')/d(rad_x, rad_y) of the tilted metric
    const double dzx = -0.5 * a.cos_phi * a.sin_2theta, dzy = -0.5 * a.sin_phi * a.sin_2theta;
    const double dxx = 1.0 + dzx / a.tan_theta * a.cos_phi, dxy = dzy / a.tan_theta * a.cos_phi;
    const double dyx = dzx / a.tan_theta * a.sin_phi, dyy = 1.0 + dzy / a.tan_theta * a.sin_phi;
    double gx = 0.0, gy = 0.0;
    for (int q = threadIdx.x; q < npix; q += 256) {
        const int xid = x0 + q % wx;
        const int yid = y0 + q / wx;
        const double rad_x = (a.startx + xid * a.pixelscale) - xa;
        const double rad_y = (a.starty + yid * a.pixelscale) - ya;
        const double z_prime = -0.5 * (rad_x * a.cos_phi + rad_y * a.sin_phi) * a.sin_2theta;
        const double zt = z_prime / a.tan_theta;
        const double x_prime = rad_x + zt * a.cos_phi;
        const double y_prime = rad_y + zt * a.sin_phi;
        const double r2 = z_prime * z_prime + x_prime * x_prime + y_prime * y_prime;
        double rad = 0.0;
        if constexpr (TAB) {
            // clamped to V(r_min) (constant) or outside the cutoff r2 <= 64 + 2^-46
            if (!(r2 <= 0x1.0000000000001p+6) || r2 < r_min2) continue;
        } else {
            if (!(r2 <= 64.0000000001)) continue;
            rad = sqrt(r2);
            if (rad < r_min || rad > 8.0) continue;  // clamped (constant) or outside the cutoff
        }
        int starti, endi, startj, endj;
        tile_block_range(a, xid >> 4, yid >> 4, starti, endi, startj, endj);
        if (bidx < starti || bidx > endi || bidy < startj || bidy > endj) continue;
        const size_t p = (size_t)yid * a.res + xid;
        float s, c;
        sincosf(phase[p], &s, &c);
        const float2 g = g_tau[p];
        // dL/dphi = -Im(conj(g) tau)
        const double dl_dphi = -((double)g.x * s - (double)g.y * c);
        double dv_r;  // v'(r) / r
        if constexpr (TAB)
            dv_r = 2.0 * (double)vtab_slope_in(tz, (float)r2);
        else
            dv_r = (double)dvproj<FORM>(pz, a.i_count, (float)rad) / rad;
        const double f = a.sigma * dl_dphi * dv_r;
        const double dr_drx = z_prime * dzx + x_prime * dxx + y_prime * dyx;  // times rad
        const double dr_dry = z_prime * dzy + x_prime * dxy + y_prime * dyy;
        gx -= f * dr_drx;  // d rad_x / d x_a = -1
        gy -= f * dr_dry;
    }
    // block reduction: 64-wide wave shuffles, then the 4 waves through LDS
    for (int off = 32; off > 0; off >>= 1) {
        gx += __shfl_down(gx, off, 64);
        gy += __shfl_down(gy, off, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = gx;
        red[1][wv] = gy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t o = 3 * (size_t)orig[i];
        grad[o + 0] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        grad[o + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
        grad[o + 2] = 0.0;  // projected: the potential does not depend on z within a slice
    }
}

// d/dr of the full-3D potentials of v3d_tab<FORM>, from the same per-element table
// (exponent scales carry log2 e, so d/dr 2^(-u r) = -ln2 u 2^(-u r)):
//   Kirkland  V = 150.41/r sum a 2^(-r u) + 266.52 sum c d^-3/2 2^(-r^2 w)
//   Peng      V = 266.52 sum a b^-3/2 2^(-r^2 w)
//   Lobato    V = 472.55 sum a b^-3/2 2^(-r u) (s/r + 1),  s = sqrt(b)/pi
template <int FORM>
__device__ __forceinline__ float dv3d_tab(const float* __restrict__ t, int, float r) {
    constexpr int ic = kTerms<FORM>;
    const float ln2 = 0.6931471805599453f;
    const float r2 = r * r;
    const float inv_r = 1.0f / r;
    if constexpr (FORM == 0) {
        float s0 = 0.f, s1 = 0.f, sg = 0.f;
        for (int i = 0; i < ic; ++i) {
            const float e = exp2f(-r * t[3 + i]);
            s0 += t[i] * e;
            s1 += t[i] * t[3 + i] * e;
        }
        for (int i = 0; i < ic; ++i) sg += t[6 + i] * t[9 + i] * exp2f(-r2 * t[9 + i]);
        return 150.4121417f * (-s0 * inv_r * inv_r - ln2 * s1 * inv_r) + 266.5157269f * (-2.0f * ln2 * r * sg);
    } else if constexpr (FORM == 1) {
        float sg = 0.f;
        for (int i = 0; i < ic; ++i) sg += t[i] * t[5 + i] * exp2f(-r2 * t[5 + i]);
        return 266.5157269f * (-2.0f * ln2 * r * sg);
    } else {
        float sum = 0.f;
        for (int i = 0; i < ic; ++i) {
            const float e = exp2f(-r * t[5 + i]);
            sum += t[i] * e * (-ln2 * t[5 + i] * (t[10 + i] * inv_r + 1.0f) - t[10 + i] * inv_r * inv_r);
        }
        return 472.545072199968f * sum;
    }
}

__device__ __forceinline__ void add_grad(double* __restrict__ grad, const int* __restrict__ orig, int i,
                                         double gx, double gy, double gz) {
    double* o = grad + 3 * (size_t)orig[i];
    unsafeAtomicAdd(o + 0, gx);
    unsafeAtomicAdd(o + 1, gy);
    unsafeAtomicAdd(o + 2, gz);
}

// Position gradients of the full-3D phase of slices first + blockIdx.z: per pixel the same
// (slice, block row) sequence, atoms and integration points as full3d_sumz (k_phase_full3d), in
// the same order, over 16x16 pixel tiles of its own.  With p1_e the
// potential at evaluation e of a (k, j) row sequence, the trapezoid gives
//   sumz = sum_e 0.5 [h_e != 0] (p1_e + p1_prev(e)),
// so p1_e carries 0.5 [h_e != 0] from its own step and 0.5 from the next evaluation when
// that one has h != 0; prev(e) may belong to the atom before (the value carried between
// atoms of a row).  Per pixel f = sigma sst dL/dphi, and dp1/d(x, y, z)_atom =
// -v'(r) rad / r (no derivative where the r_min clamp holds).  Each atom's own terms are
// wave-reduced and added with one atomic per wave; the carried term of an atom's first
// evaluation goes to the atom before, reduced when the whole wave agrees on that atom.
template <int FORM>
__global__ __launch_bounds__(256) void k_atom_grad_full3d(
    const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
    const int* __restrict__ aZ, const int* __restrict__ bsp, const int* __restrict__ orig,
    const float* __restrict__ tab3d, PotArgs a, int first, const float2* __restrict__ g_tau_all,
    const float* __restrict__ phase_all, double* __restrict__ grad) {
#pragma clang fp contract(off)
    // slice first + blockIdx.z and its planes
    const int slice = first + blockIdx.z;
    const float2* __restrict__ g_tau = g_tau_all + (size_t)slice * a.res * a.res;
    const float* __restrict__ phase = phase_all + (size_t)slice * a.res * a.res;
    __shared__ double sx[256], sy[256], sz[256];
    __shared__ int sZ[256];
    const int lid = threadIdx.x;
    const int lane = lid & 63;
    const int xid = blockIdx.x * 16 + (lid & 15);
    const int yid = blockIdx.y * 16 + (lid >> 4);
    int starti, endi, startj, endj;
    tile_block_range(a, blockIdx.x, blockIdx.y, starti, endi, startj, endj);
    int topz = slice - a.load_z, bottomz = slice + a.load_z;
    if (topz < 0) topz = 0;
    if (bottomz >= a.n_slices) bottomz = a.n_slices - 1;
    const double int_r = 1.0 / a.integrals;
    const double sst = a.dz * int_r;
    const double current_z = a.z_min + (a.n_slices - slice) * a.dz;
    const double im_pos_x = a.startx + xid * a.pixelscale;
    const double im_pos_y = a.starty + yid * a.pixelscale;
    const double r_min = 0.25 * a.pixelscale;
    const double r_min2 = r_min * r_min;
    // f = sigma sst dL/dphi, dL/dphi = -Im(conj(g) exp(i phi))
    double f;
    {
        const size_t p = (size_t)yid * a.res + xid;
        float s, c;
        sincosf(phase[p], &s, &c);
        const float2 g = g_tau[p];
        f = 0.5 * a.sigma * sst * (-((double)g.x * s - (double)g.y * c));  // with the trapezoid's 1/2
    }
    const int nb = a.blocks_x * a.blocks_y;
    for (int k = topz; k <= bottomz; ++k) {
        for (int j = startj; j <= endj; ++j) {
            const int start = bsp[k * nb + a.blocks_x * j + starti];
            const int end = bsp[k * nb + a.blocks_x * j + endi + 1];
            // the previous evaluation of this row sequence: its atom and f dp1/dpos
            int prev = -1;
            double px = 0.0, py = 0.0, pz = 0.0;
            for (int c0 = start; c0 < end; c0 += 256) {
                const int cnt = min(256, end - c0);
                __syncthreads();
                if (lid < cnt) {
                    sx[lid] = ax[c0 + lid];
                    sy[lid] = ay[c0 + lid];
                    sz[lid] = az[c0 + lid];
                    sZ[lid] = aZ[c0 + lid];
                }
                __syncthreads();
                for (int l = 0; l < cnt; ++l) {
                    const int atom = c0 + l;
                    double rad_x = im_pos_x - sx[l];
                    double rad_y = im_pos_y - sy[l];
                    const float* tz = tab3d + (size_t)(sZ[l] - 1) * kPot3dStride;
                    double gx = 0.0, gy = 0.0, gz = 0.0;      // this atom's own terms
                    double cx = 0.0, cy = 0.0, cz = 0.0;      // carried term for `carry_to`
                    int carry_to = -1;
                    bool any = false;
                    for (int h = 0; h <= a.integrals; ++h) {
                        rad_x -= a.shift_x;
                        rad_y -= a.shift_y;
                        const double xyrad2 = rad_x * rad_x + rad_y * rad_y;
                        const double rad_z = (current_z - h * a.dz * int_r) - sz[l];
                        if (!(xyrad2 <= 64.0 && rad_z <= 3.0)) continue;
                        any = true;
                        const double r2 = xyrad2 + rad_z * rad_z;
                        double dx = 0.0, dy = 0.0, dzz = 0.0;
                        if (!(r2 < r_min2)) {
                            const double r = sqrt(r2);
                            const double q = -f * (double)dv3d_tab<FORM>(tz, a.i_count, (float)r) / r;
                            dx = q * rad_x;
                            dy = q * rad_y;
                            dzz = q * rad_z;
                        }
                        if (h != 0) {
                            gx += dx;
                            gy += dy;
                            gz += dzz;
                            if (prev == atom) {
                                gx += px;
                                gy += py;
                                gz += pz;
                            } else if (prev >= 0) {
                                carry_to = prev;
                                cx = px;
                                cy = py;
                                cz = pz;
                            }
                        }
                        prev = atom;
                        px = dx;
                        py = dy;
                        pz = dzz;
                    }
                    // own terms: one atomic triple per wave
                    if (__ballot(any)) {
                        gx = wave_sum(gx);
                        gy = wave_sum(gy);
                        gz = wave_sum(gz);
                        if (lane == 0 && (gx != 0.0 || gy != 0.0 || gz != 0.0))
                            add_grad(grad, orig, atom, gx, gy, gz);
                    }
                    // carried terms
                    const unsigned long long m = __ballot(carry_to >= 0);
                    if (m) {
                        const int t0 = __shfl(carry_to, __ffsll((long long)m) - 1, 64);
                        if (__ballot(carry_to >= 0 && carry_to != t0) == 0) {
                            cx = wave_sum(cx);
                            cy = wave_sum(cy);
                            cz = wave_sum(cz);
                            if (lane == 0) add_grad(grad, orig, t0, cx, cy, cz);
                        } else if (carry_to >= 0) {
                            add_grad(grad, orig, carry_to, cx, cy, cz);
                        }
                    }
                }
            }
        }
    }
}

}  // namespace mse
