// Register-resident Stockham FFT for one line of length N held by L lanes of a
// 64-wide wavefront (gfx950).
//
// Layout: lane t (0 <= t < L) of a line holds x[t + L*m] in v[m], m = 0..E-1,
// E = N / L.  A radix-R pass (R | E) gives lane t the butterflies i = t + L*b,
// b = 0..E/R-1, whose inputs x[i + r*N/R] are exactly v[b + r*E/R]: a pass never
// needs data from another lane on its input side.  Outputs go to j + r*p with
// j = (i - k)*R + k, k = i mod p (Stockham autosort), which is exchanged through a
// padded LDS line; the last pass (p*R == N) writes x[i + r*N/R] and therefore
// lands back in the register layout without an exchange.  A plan of P passes costs
// P-1 LDS round trips, and the first pass (p == 1) has no twiddles.
//
// Twiddles come from a per-size table tw[q] = exp(-2 pi i q / N) evaluated in
// double on the host (accurate to float rounding); inverse transforms use the
// conjugate.  All transforms are unnormalised; scaling is folded by the callers.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace mse {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// ---- compile-time trigonometry for the small-DFT internal twiddles -----------------
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double ct_cos_series(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
constexpr double ct_sin_series(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 30; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
// cos/sin of 2*pi*J/R reduced to [-pi/4, pi/4] with exact symmetries
constexpr double ct_cos2pi(int J, int R) {
    int j = ((J % R) + R) % R;
    // octant handling using multiples of R/8 where exact; else plain series on reduced angle
    double a = kTwoPi * (double)j / (double)R;
    if (a > kTwoPi / 2) a -= kTwoPi;
    return ct_cos_series(a);
}
constexpr double ct_sin2pi(int J, int R) {
    int j = ((J % R) + R) % R;
    double a = kTwoPi * (double)j / (double)R;
    if (a > kTwoPi / 2) a -= kTwoPi;
    return ct_sin_series(a);
}

// multiply by W_R^J = exp(-+ 2 pi i J / R) (forward: minus, inverse: plus)
template <int R, int J, bool INV>
__device__ __forceinline__ float2 rot(float2 a) {
    constexpr int j = ((J % R) + R) % R;
    if constexpr (j == 0) {
        return a;
    } else if constexpr (4 * j == R) {  // -i (fwd) / +i (inv)
        return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
    } else if constexpr (2 * j == R) {
        return make_float2(-a.x, -a.y);
    } else if constexpr (4 * j == 3 * R) {  // +i (fwd) / -i (inv)
        return INV ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
    } else if constexpr (8 * j == R || 8 * j == 3 * R || 8 * j == 5 * R || 8 * j == 7 * R) {
        constexpr float h = 0.70710678118654752440f;
        constexpr float c = (float)ct_cos2pi(j, R);
        constexpr float s = (float)(INV ? ct_sin2pi(j, R) : -ct_sin2pi(j, R));
        // (a.x + i a.y)(c + i s) with |c| = |s| = 1/sqrt2
        constexpr float sc = c > 0 ? h : -h;
        constexpr float ss = s > 0 ? h : -h;
        return make_float2(a.x * sc - a.y * ss, a.x * ss + a.y * sc);
    } else {
        constexpr float c = (float)ct_cos2pi(j, R);
        constexpr float s = (float)(INV ? ct_sin2pi(j, R) : -ct_sin2pi(j, R));
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// ---- small DFTs in registers --------------------------------------------------------
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
    static __device__ __forceinline__ void run(float2* a) {
        float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    }
};

template <bool INV>
struct Dft<3, INV> {
    static __device__ __forceinline__ void run(float2* a) {
        constexpr float s3 = 0.86602540378443864676f;
        float2 t1 = cadd(a[1], a[2]);
        float2 t2 = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
        float2 d = csub(a[1], a[2]);
        // forward: y1 = t2 - i s3 d, y2 = t2 + i s3 d
        float2 isd = INV ? make_float2(-s3 * d.y, s3 * d.x) : make_float2(s3 * d.y, -s3 * d.x);
        a[0] = cadd(a[0], t1);
        a[1] = cadd(t2, isd);
        a[2] = csub(t2, isd);
    }
};

template <bool INV>
struct Dft<4, INV> {
    static __device__ __forceinline__ void run(float2* a) {
        float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        float2 t2 = cadd(a[1], a[3]), t3 = csub(a[1], a[3]);
        float2 t3r = INV ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3r);
        a[3] = csub(t1, t3r);
    }
};

// R = R1 * R2: r = r1 + R1*r2, s = R2*s1 + s2,
// y_s = sum_r1 W_R1^{r1 s1} W_R^{r1 s2} [sum_r2 W_R2^{r2 s2} a_{r1 + R1 r2}]
template <int R1, int R2, bool INV>
struct DftComposite {
    static constexpr int R = R1 * R2;
    template <int r1>
    static __device__ __forceinline__ void inner(const float2* a, float2* b) {
        float2 tmp[R2];
#pragma unroll
        for (int r2 = 0; r2 < R2; ++r2) tmp[r2] = a[r1 + R1 * r2];
        Dft<R2, INV>::run(tmp);
        twid<r1, 0>(tmp, b);
    }
    template <int r1, int s2>
    static __device__ __forceinline__ void twid(const float2* tmp, float2* b) {
        if constexpr (s2 < R2) {
            b[r1 * R2 + s2] = rot<R, r1 * s2, INV>(tmp[s2]);
            twid<r1, s2 + 1>(tmp, b);
        }
    }
    template <int r1>
    static __device__ __forceinline__ void all_inner(const float2* a, float2* b) {
        if constexpr (r1 < R1) {
            inner<r1>(a, b);
            all_inner<r1 + 1>(a, b);
        }
    }
    static __device__ __forceinline__ void run(float2* a) {
        float2 b[R];
        all_inner<0>(a, b);
#pragma unroll
        for (int s2 = 0; s2 < R2; ++s2) {
            float2 tmp[R1];
#pragma unroll
            for (int r1 = 0; r1 < R1; ++r1) tmp[r1] = b[r1 * R2 + s2];
            Dft<R1, INV>::run(tmp);
#pragma unroll
            for (int s1 = 0; s1 < R1; ++s1) a[R2 * s1 + s2] = tmp[s1];
        }
    }
};

template <bool INV> struct Dft<6, INV> : DftComposite<3, 2, INV> {};
template <bool INV> struct Dft<8, INV> : DftComposite<4, 2, INV> {};
template <bool INV> struct Dft<12, INV> : DftComposite<4, 3, INV> {};
template <bool INV> struct Dft<16, INV> : DftComposite<4, 4, INV> {};

// ---- radix plans ---------------------------------------------------------------------
// A plan: N points over L lanes per line, radices R0, R1, R2 (R | E), with an LDS
// padding of one complex every 2^PS entries.
template <int N_, int L_, int R0_, int R1_, int R2_>
struct Plan3 {
    static constexpr int N = N_, L = L_, E = N_ / L_;
    static constexpr int R0 = R0_, R1 = R1_, R2 = R2_, R3 = 1;
    static_assert(R0_ * R1_ * R2_ == N_, "radices must multiply to N");
    static_assert(E % R0_ == 0 && E % R1_ == 0 && E % R2_ == 0, "radix must divide E");
};
// Four passes, for sizes whose elements per lane admit no three-radix split (768 points on
// 64 lanes: 12 per lane, 4 * 4 * 4 * 12).  With the radices in this order every pass's p
// divides L, so each pass needs one twiddle set per lane.
template <int N_, int L_, int R0_, int R1_, int R2_, int R3_>
struct Plan4 {
    static constexpr int N = N_, L = L_, E = N_ / L_;
    static constexpr int R0 = R0_, R1 = R1_, R2 = R2_, R3 = R3_;
    static_assert(R0_ * R1_ * R2_ * R3_ == N_, "radices must multiply to N");
    static_assert(E % R0_ == 0 && E % R1_ == 0 && E % R2_ == 0 && E % R3_ == 0, "radix must divide E");
};

template <int N>
struct PlanFor;
template <> struct PlanFor<256> { using P = Plan3<256, 32, 8, 8, 4>; };
template <> struct PlanFor<384> { using P = Plan4<384, 32, 4, 4, 2, 12>; };
template <> struct PlanFor<512> { using P = Plan3<512, 64, 8, 8, 8>; };
template <> struct PlanFor<768> { using P = Plan4<768, 64, 4, 4, 4, 12>; };
template <> struct PlanFor<1024> { using P = Plan3<1024, 64, 16, 8, 8>; };

// Alternative plans with half the lanes per line (twice the elements per lane): more
// independent work per lane and fewer threads per column tile.
template <int N>
struct PlanHalfLanes { using P = typename PlanFor<N>::P; };
template <> struct PlanHalfLanes<256> { using P = Plan3<256, 16, 8, 8, 4>; };
template <> struct PlanHalfLanes<384> { using P = Plan3<384, 16, 8, 8, 6>; };
template <> struct PlanHalfLanes<768> { using P = Plan3<768, 32, 8, 8, 12>; };
template <> struct PlanHalfLanes<512> { using P = Plan3<512, 32, 8, 8, 8>; };
template <> struct PlanHalfLanes<1024> { using P = Plan3<1024, 32, 16, 8, 8>; };

// Orders LDS traffic among the lanes of one wavefront: waits for this wave's LDS
// operations and stops the compiler moving memory accesses across the point.  Valid
// when every lane sharing an LDS line belongs to the same wavefront (L <= 64).
struct WaveSync {
    __device__ __forceinline__ void operator()() const {
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_wave_barrier();
    }
};
struct BlockSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// padded LDS index: one extra complex every 16 entries.  The LDS serves 16 lanes of an
// 8-byte access per cycle over 16 bank pairs, so 16 lanes at consecutive indices, or at
// indices 8 apart (the first pass's radix-8 writes), hit 16 different pairs.
__device__ __forceinline__ int lpad(int q) { return q + (q >> 4); }
// Index of the second pass's exchange: lpad plus 4 more per block of 64 (monotone, so
// one-to-one).  With a first radix of 8 the second pass's 16-lane groups write two runs of 8
// entries 64 apart, which lpad puts 4 pairs apart (a 2-way conflict); the extra 4 moves the
// second run onto the other 8 pairs.  Consecutive reads stay conflict-free (the shift is the
// same within an aligned block of 16).
__device__ __forceinline__ int lpad2(int q) { return q + (q >> 4) + ((q >> 6) << 2); }
template <int N>
struct LdsLine {
    static constexpr int size = N + N / 8;  // lpad2(N-1) + 1 <= size (and lpad(N-1) + 1)
    // column tiles use a stride that is 1 mod 16 complex so 16 columns hit 16 bank pairs
    static constexpr int col_stride = size + ((17 - size % 16) % 16);
};

// Twiddles of one pass, preloaded into registers.  k = i mod p with i = t + L*b; when
// p divides L, k = t mod p for every b and one set of R-1 values serves all butterflies.
template <int N, int L, int R, int P_>
struct PassTw {
    static constexpr int NB = (N / L) / R;
    static constexpr bool shared = (L % P_) == 0;
    static constexpr int SETS = shared ? 1 : NB;
    float2 w[SETS][R - 1];
    __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
        constexpr int step = N / (P_ * R);
#pragma unroll
        for (int s = 0; s < SETS; ++s) {
            const int k = (t + L * s) % P_;
#pragma unroll
            for (int r = 1; r < R; ++r) w[s][r - 1] = tw[r * k * step];
        }
    }
    // twiddle W_{pR}^{r k} of butterfly b
    __device__ __forceinline__ float2 get(int b, int r, int) const { return w[shared ? 0 : b][r - 1]; }
};

// Same twiddles read from the (cache-resident) table at the point of use: fewer
// registers held across the transform, one extra load per twiddle.
template <int N, int L, int R, int P_>
struct PassTwMem {
    static constexpr int step = N / (P_ * R);
    const float2* __restrict__ tw;
    __device__ __forceinline__ void load(const float2* __restrict__ table, int) { tw = table; }
    __device__ __forceinline__ float2 get(int, int r, int k) const { return tw[r * k * step]; }
};

struct NoPassTw {
    __device__ __forceinline__ void load(const float2* __restrict__, int) {}
};

template <class PL, bool IN_REGS = true>
struct PlanTw {
    using T1 = typename std::conditional<IN_REGS, PassTw<PL::N, PL::L, PL::R1, PL::R0>,
                                         PassTwMem<PL::N, PL::L, PL::R1, PL::R0>>::type;
    using T2 = typename std::conditional<IN_REGS, PassTw<PL::N, PL::L, PL::R2, PL::R0 * PL::R1>,
                                         PassTwMem<PL::N, PL::L, PL::R2, PL::R0 * PL::R1>>::type;
    // fourth pass (four-pass plans only)
    static constexpr int R3 = PL::R3 > 1 ? PL::R3 : 2;
    using T3 = typename std::conditional<
        (PL::R3 > 1),
        typename std::conditional<IN_REGS, PassTw<PL::N, PL::L, R3, PL::R0 * PL::R1 * PL::R2>,
                                  PassTwMem<PL::N, PL::L, R3, PL::R0 * PL::R1 * PL::R2>>::type,
        NoPassTw>::type;
    T1 p1;
    T2 p2;
    T3 p3;
    __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
        p1.load(tw, t);
        p2.load(tw, t);
        p3.load(tw, t);
    }
};

// One Stockham pass.  v holds the line in register layout; line is this line's LDS.
// P_ = product of radices already done (p); LAST = this pass ends the transform.
template <int N, int L, int R, int P_, bool INV, bool LAST, class TW, bool PAD2 = false>
__device__ __forceinline__ void stockham_pass(float2 (&v)[N / L], float2* line, int t,
                                              const TW* tw) {
    constexpr int E = N / L;
    constexpr int NB = E / R;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        float2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[b + r * NB];
        const int i = t + L * b;
        const int k = (P_ == 1) ? 0 : (i % P_);
        if constexpr (P_ > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const float2 w = tw->get(b, r, k);
                a[r] = INV ? cmulc(a[r], w) : cmul(a[r], w);
            }
        }
        Dft<R, INV>::run(a);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * NB] = a[r];
        } else {
            const int j = (i - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) line[PAD2 ? lpad2(j + r * P_) : lpad(j + r * P_)] = a[r];
        }
    }
}

template <int N, int L, bool PAD2 = false>
__device__ __forceinline__ void line_gather(float2 (&v)[N / L], const float2* line, int t) {
    constexpr int E = N / L;
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = line[PAD2 ? lpad2(t + L * m) : lpad(t + L * m)];
}

struct NoTw {
    __device__ __forceinline__ float2 get(int, int, int) const { return make_float2(1.f, 0.f); }
};

// Full transform of one line in register layout.  sync() must order LDS writes before
// reads and reads before the next pass's writes for every lane sharing the LDS line.
template <class PL, bool INV, class TWS, class Sync>
__device__ __forceinline__ void fft_line(float2 (&v)[PL::E], float2* line, int t, const TWS& tw,
                                         Sync sync) {
    constexpr int N = PL::N, L = PL::L;
    stockham_pass<N, L, PL::R0, 1, INV, false, NoTw>(v, line, t, nullptr);
    sync();
    line_gather<N, L>(v, line, t);
    sync();
    stockham_pass<N, L, PL::R1, PL::R0, INV, false, typename TWS::T1, true>(v, line, t, &tw.p1);
    sync();
    line_gather<N, L, true>(v, line, t);
    sync();
    if constexpr (PL::R3 == 1) {
        stockham_pass<N, L, PL::R2, PL::R0 * PL::R1, INV, true>(v, line, t, &tw.p2);
    } else {
        stockham_pass<N, L, PL::R2, PL::R0 * PL::R1, INV, false, typename TWS::T2>(v, line, t, &tw.p2);
        sync();
        line_gather<N, L>(v, line, t);
        sync();
        stockham_pass<N, L, PL::R3, PL::R0 * PL::R1 * PL::R2, INV, true>(v, line, t, &tw.p3);
    }
}

}  // namespace mse
