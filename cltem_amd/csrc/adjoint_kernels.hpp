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

namespace mse {

// d/dr of the projected potentials of vproj_tab<FORM> (kernels/transmission_potentials_projected_f.cl:71-181),
// from the raw parameters (p = a1,b1,..,c1,d1,.. for Kirkland; a1..a5,b1..b5 otherwise):
//   Kirkland  300.82 sum a K0(2 pi r sqrt b) + 150.41 sum (c/d) exp(-(pi r)^2/d)
//             -> 300.82 sum a (-2 pi sqrt b) K1(2 pi r sqrt b) + 150.41 sum (c/d)(-2 pi^2 r/d) exp(..)
//   Peng      150.41 sum (a/b) exp(-(pi r)^2/b) -> 150.41 sum (a/b)(-2 pi^2 r/b) exp(..)
//   Lobato    945.09 sum a b^-3/2 (K0(c r) + r K1(c r)), c = 2 pi / sqrt b
//             -> 945.09 sum a b^-3/2 (-c)(K1(c r) + r K0(c r))   (K0' = -K1, (z K1)' = -z K0)
template <int FORM>
__device__ __forceinline__ float dvproj(const float* __restrict__ p, int ic, float r) {
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

static __global__ void k_fill_zero(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// One block of 256 threads per atom of slice s (sorted range [first, first + gridDim.x)).
// The block visits the pixels of the atom's 8 A neighbourhood, keeps those whose tile block
// range contains the atom's block (as the forward sum does), and reduces
//   -sigma * dL/dphi(p) * v'(r) * dr/d(rad) in double.
// grad has 3 doubles (x, y, z) per input atom, indexed through orig.
template <int FORM>
__global__ __launch_bounds__(256) void k_atom_grad_projected(
    const double* __restrict__ ax, const double* __restrict__ ay, const int* __restrict__ aZ,
    const int* __restrict__ orig, int first, const float* __restrict__ params, PotArgs a,
    const float2* __restrict__ g_tau, const float* __restrict__ phase, double* __restrict__ grad) {
#pragma clang fp contract(off)
    __shared__ double red[2][4];
    const int i = first + blockIdx.x;
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
    // d(z', x', y')/d(rad_x, rad_y) of the tilted metric
    const double dzx = -0.5 * a.cos_phi * a.sin_2theta, dzy = -0.5 * a.sin_phi * a.sin_2theta;
    const double dxx = 1.0 + dzx / a.tan_theta * a.cos_phi, dxy = dzy / a.tan_theta * a.cos_phi;
    const double dyx = dzx / a.tan_theta * a.sin_phi, dyy = 1.0 + dzy / a.tan_theta * a.sin_phi;
    double gx = 0.0, gy = 0.0;
    for (int q = threadIdx.x; q < npix; q += 256) {
        const int xid = x0 + q % wx;
        const int yid = y0 + q / wx;
        int starti, endi, startj, endj;
        tile_block_range(a, xid >> 4, yid >> 4, starti, endi, startj, endj);
        if (bidx < starti || bidx > endi || bidy < startj || bidy > endj) continue;
        const double rad_x = (a.startx + xid * a.pixelscale) - xa;
        const double rad_y = (a.starty + yid * a.pixelscale) - ya;
        const double z_prime = -0.5 * (rad_x * a.cos_phi + rad_y * a.sin_phi) * a.sin_2theta;
        const double zt = z_prime / a.tan_theta;
        const double x_prime = rad_x + zt * a.cos_phi;
        const double y_prime = rad_y + zt * a.sin_phi;
        const double r2 = z_prime * z_prime + x_prime * x_prime + y_prime * y_prime;
        if (!(r2 <= 64.0000000001)) continue;
        const double rad = sqrt(r2);
        if (rad < r_min || rad > 8.0) continue;  // clamped (constant) or outside the cutoff
        const size_t p = (size_t)yid * a.res + xid;
        float s, c;
        sincosf(phase[p], &s, &c);
        const float2 g = g_tau[p];
        // dL/dphi = -Im(conj(g) tau)
        const double dl_dphi = -((double)g.x * s - (double)g.y * c);
        const double dv = (double)dvproj<FORM>(pz, a.i_count, (float)rad);
        const double f = a.sigma * dl_dphi * dv / rad;
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

}  // namespace mse
