// CTEM detector kernels (gfx950): shot noise with a counter-based generator and the
// real <-> complex moves around the CCD response filters.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"

namespace mse {

__device__ __forceinline__ double unit_open(uint32_t u) { return ((double)u + 0.5) * (1.0 / 4294967296.0); }

// Poisson deviate of mean lam for stream `id`: inversion by sequential search below a mean
// of 10 (one uniform, counter (id, 0, 0, 0)); above, the transformed rejection method PTRS
// (W. Hormann, Insurance: Mathematics and Economics 12 (1993) 39-45) with the uniforms of
// counter (id, attempt + 1, 0, 0).  Both loops are bounded.
__device__ __forceinline__ double poisson_deviate(double lam, uint32_t id, uint2 key) {
    if (!(lam > 0.0)) return 0.0;
    if (lam < 10.0) {
        const uint4 r = philox4x32_10(make_uint4(id, 0u, 0u, 0u), key);
        const double u = unit_open(r.x);
        double p = exp(-lam), F = p;
        int k = 0;
        while (u > F && k < 200) {
            ++k;
            p *= lam / k;
            F += p;
        }
        return (double)k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (uint32_t it = 0; it < 64; ++it) {
        const uint4 r = philox4x32_10(make_uint4(id, it + 1u, 0u, 0u), key);
        const double U = unit_open(r.x) - 0.5;
        const double V = unit_open(r.y);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
    return floor(lam + 0.5);  // not reached in practice (acceptance > 0.9 per attempt)
}

// z[p] = (x[p], 0) from a double image
static __global__ void k_real_to_complex(const double* __restrict__ x, float2* __restrict__ z, size_t n) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) z[p] = make_float2((float)x[p], 0.f);
}

// z[p] = (conversion * Poisson(dose * Re z[p]), 0)
static __global__ void k_shot_noise(float2* __restrict__ z, size_t n, double dose, double conversion, uint2 key) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double lam = dose * (double)z[p].x;
    z[p] = make_float2((float)(conversion * poisson_deviate(lam, (uint32_t)p, key)), 0.f);
}

// out[p] = Re z[p]
static __global__ void k_real_part(const float2* __restrict__ z, double* __restrict__ out, size_t n) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = (double)z[p].x;
}

}  // namespace mse
