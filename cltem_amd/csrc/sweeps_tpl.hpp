// Sweep launchers (row and column passes and their variants) for one resolution N; each
// sweeps_<N>.hip instantiates them so the resolutions compile in parallel.
#pragma once
#include "engine_internal.hpp"

namespace mse_rt {

// Blocks of a persistent kernel that stay resident per CU (HIP's occupancy query for that
// kernel, asked once), at most `cap`: the persistent sweeps size their grids to what runs
// at once, so no block waits for another to finish before it starts.
template <auto KERNEL>
int resident_per_cu(int block_threads, int cap) {
    static const int n = [&]() {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, KERNEL, block_threads, 0) != hipSuccess) {
            (void)hipGetLastError();
            b = cap;
        }
        return std::max(1, std::min(b, cap));
    }();
    return n;
}

template <int N>
void row_pass_n(mse_engine* e, int mode, const float2* in, float2* out, int nlines_sets, size_t wave_stride,
                const mse::RowArgs& args, int variant) {
    {
        auto go = [&](auto plan, auto sync, auto twreg) {
            using PL = decltype(plan);
            using SY = decltype(sync);
            constexpr bool TR = decltype(twreg)::value;
            constexpr int WPE = TR ? 1 : 8;
            constexpr int LINES = 256 / PL::L;
            dim3 grid(N / LINES, nlines_sets);
            launch(e, PK_ROW, [&]() {
                switch (mode) {
                    case mse::ROW_IFFT_T_FFT:
                        mse::k_row<PL, mse::ROW_IFFT_T_FFT, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                    case mse::ROW_PLANE_T_FFT:
                        mse::k_row<PL, mse::ROW_PLANE_T_FFT, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                    case mse::ROW_FFT:
                        mse::k_row<PL, mse::ROW_FFT, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                    case mse::ROW_IFFT:
                        mse::k_row<PL, mse::ROW_IFFT, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                    case mse::ROW_COPY:
                        mse::k_row<PL, mse::ROW_COPY, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                    case mse::ROW_IFFT_INTENSITY:
                        mse::k_row<PL, mse::ROW_IFFT_INTENSITY, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                    case mse::ROW_ADJ_SEED:
                        mse::k_row<PL, mse::ROW_ADJ_SEED, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                    case mse::ROW_ADJ:
                        mse::k_row<PL, mse::ROW_ADJ, SY, TR, WPE><<<grid, 256, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args);
                        break;
                }
            });
        };
        using P0 = typename mse::PlanFor<N>::P;
        using P1 = typename mse::PlanHalfLanes<N>::P;
        using REG = std::true_type;
        using MEM = std::false_type;
        if ((mode == mse::ROW_FFT || mode == mse::ROW_IFFT) && in == out) {
            // one transform in place (the band limit of transmission functions): the persistent
            // pipelined kernel, whatever the tuned variant
            auto go_pipe = [&](auto m) {
                constexpr int MD = decltype(m)::value;
                constexpr int LINES = 256 / P0::L;
                const int per_cu = resident_per_cu<mse::k_row_step_pipe<P0, mse::WaveSync, true, MD>>(256, 4);
                const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N * nlines_sets + LINES - 1) / LINES));
                launch(e, PK_ROW, [&]() {
                    mse::k_row_step_pipe<P0, mse::WaveSync, true, MD><<<blocks, 256, 0, e->stream>>>(
                        out, wave_stride, e->tw.p, args, nlines_sets);
                });
            };
            if (mode == mse::ROW_FFT)
                go_pipe(std::integral_constant<int, mse::ROW_FFT>{});
            else
                go_pipe(std::integral_constant<int, mse::ROW_IFFT>{});
            return;
        }
        if constexpr (N <= 512) {  // half-lane lines keep the prefetch registers within budget up to 512
            if (mode == mse::ROW_IFFT_T_FFT && variant == 5 && in == out) {
                constexpr int LINES = 256 / P1::L;
                const int per_cu = resident_per_cu<mse::k_row_step_pipe<P1, mse::WaveSync>>(256, 4);
                const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N * nlines_sets + LINES - 1) / LINES));
                launch(e, PK_ROW, [&]() {
                    mse::k_row_step_pipe<P1, mse::WaveSync><<<blocks, 256, 0, e->stream>>>(out, wave_stride, e->tw.p,
                                                                                         args, nlines_sets);
                });
                return;
            }
        }
        if (mode == mse::ROW_IFFT_T_FFT && variant == 6 && in == out) {  // twiddles read at use
            constexpr int LINES = 256 / P0::L;
            const int per_cu = resident_per_cu<mse::k_row_step_pipe<P0, mse::WaveSync, false>>(256, 4);
            const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N * nlines_sets + LINES - 1) / LINES));
            launch(e, PK_ROW, [&]() {
                mse::k_row_step_pipe<P0, mse::WaveSync, false><<<blocks, 256, 0, e->stream>>>(out, wave_stride, e->tw.p,
                                                                                            args, nlines_sets);
            });
            return;
        }
        if (mode == mse::ROW_IFFT_T_FFT && variant >= 4 && in == out) {
            constexpr int LINES = 256 / P0::L;
            const int per_cu = resident_per_cu<mse::k_row_step_pipe<P0, mse::WaveSync>>(256, 4);
            const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N * nlines_sets + LINES - 1) / LINES));
            launch(e, PK_ROW, [&]() {
                mse::k_row_step_pipe<P0, mse::WaveSync><<<blocks, 256, 0, e->stream>>>(out, wave_stride, e->tw.p, args,
                                                                                     nlines_sets);
            });
            return;
        }
        if (mode == mse::ROW_IFFT_T_FFT && variant == 1)
            go(P1{}, mse::WaveSync{}, REG{});
        else if (mode == mse::ROW_IFFT_T_FFT && variant == 2)
            go(P0{}, mse::BlockSync{}, REG{});
        else if (mode == mse::ROW_IFFT_T_FFT && variant == 3)
            go(P0{}, mse::WaveSync{}, MEM{});
        else
            go(P0{}, mse::WaveSync{}, REG{});
    }
}

// returns the number of column tiles of the launch (detector partials per wave)
template <int N>
int col_pass_n(mse_engine* e, int mode, const float2* in, float2* out, int nsets, size_t wave_stride,
               const mse::ColArgs& args, int kind, int variant) {
    const bool hot = mode == mse::COL_PROP || mode == mse::COL_PROP_DETECT || mode == mse::COL_COPY;
    if (!hot) variant = 0;
    int tiles = 0;
    {
        auto go = [&](auto plan, auto tc, auto sync, auto twreg) {
            using PL = decltype(plan);
            using SY = decltype(sync);
            constexpr bool TR = decltype(twreg)::value;
            constexpr int WPE = TR ? 1 : 8;
            constexpr int TC = decltype(tc)::value;
            constexpr int NT = TC * PL::L;
            dim3 grid(N / TC, nsets);
            tiles = N / TC;
            launch(e, kind, [&]() {
                switch (mode) {
#define MSE_COL_CASE(M) \
    case mse::M: mse::k_col<PL, TC, mse::M, SY, TR, WPE><<<grid, NT, 0, e->stream>>>(in, out, wave_stride, e->tw.p, args); break;
                    MSE_COL_CASE(COL_PROP)
                    MSE_COL_CASE(COL_PROP_DETECT)
                    MSE_COL_CASE(COL_PROBE)
                    MSE_COL_CASE(COL_FFT_DETECT)
                    MSE_COL_CASE(COL_FFT_DIFF)
                    MSE_COL_CASE(COL_PROP_STORE_DETECT)
                    MSE_COL_CASE(COL_COPY)
                    MSE_COL_CASE(COL_PROP_CONJ)
                    MSE_COL_CASE(COL_CTEM_IMAGE)
                    MSE_COL_CASE(COL_FILTER)
#undef MSE_COL_CASE
                }
            });
        };
        using P0 = typename mse::PlanFor<N>::P;
        using P1 = typename mse::PlanHalfLanes<N>::P;
        using C16 = std::integral_constant<int, 16>;
        using C8 = std::integral_constant<int, 8>;
        using C32 = std::integral_constant<int, 32>;
        using REG = std::true_type;
        using MEM = std::false_type;
        if constexpr (N <= 512) {
            if (variant == 8 && mode == mse::COL_PROP && in == out) {  // half lanes: two blocks per CU
                constexpr int NT = 16 * P1::L;
                const int per_cu = resident_per_cu<mse::k_col_prop_pipe<P1, 16, mse::WaveSync>>(NT, 2048 / NT);
                const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N / 16) * nsets));
                tiles = N / 16;
                launch(e, kind, [&]() {
                    mse::k_col_prop_pipe<P1, 16, mse::WaveSync><<<blocks, NT, 0, e->stream>>>(out, wave_stride,
                                                                                            e->tw.p, args, nsets);
                });
                return tiles;
            }
        }
        if ((variant == 12 || variant == 13) && mode == mse::COL_PROP && in == out) {
            // persistent tiles of 8 (12) or 16 (13) columns with XCD-aware item order; the grid
            // is a multiple of the 8 XCDs
            auto go_xcd = [&](auto tc) {
                constexpr int TCX = decltype(tc)::value;
                constexpr int NT = TCX * P0::L;
                const int per_cu = resident_per_cu<mse::k_col_prop_pipe<P0, TCX, mse::WaveSync, true, 1, true>>(
                    NT, 2048 / NT);
                int blocks = std::min(e->num_cus * per_cu, (N / TCX) * nsets);
                blocks = blocks >= 8 ? blocks / 8 * 8 : 8;
                tiles = N / TCX;
                launch(e, kind, [&]() {
                    mse::k_col_prop_pipe<P0, TCX, mse::WaveSync, true, 1, true><<<blocks, NT, 0, e->stream>>>(
                        out, wave_stride, e->tw.p, args, nsets);
                });
            };
            if (variant == 12)
                go_xcd(C8{});
            else
                go_xcd(C16{});
            return tiles;
        }
        if (variant == 10 && mode == mse::COL_PROP && in == out) {  // 8-column tiles (no spills at 1024)
            constexpr int NT = 8 * P0::L;
            const int per_cu = resident_per_cu<mse::k_col_prop_pipe<P0, 8, mse::WaveSync>>(NT, 1024 / NT);
            const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N / 8) * nsets));
            tiles = N / 8;
            launch(e, kind, [&]() {
                mse::k_col_prop_pipe<P0, 8, mse::WaveSync><<<blocks, NT, 0, e->stream>>>(out, wave_stride, e->tw.p,
                                                                                        args, nsets);
            });
            return tiles;
        }
        if (variant == 9 && mode == mse::COL_PROP && in == out) {  // twiddles read at use
            constexpr int NT = 16 * P0::L;
            const int per_cu = resident_per_cu<mse::k_col_prop_pipe<P0, 16, mse::WaveSync, false>>(NT, 1024 / NT);
            const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N / 16) * nsets));
            tiles = N / 16;
            launch(e, kind, [&]() {
                mse::k_col_prop_pipe<P0, 16, mse::WaveSync, false><<<blocks, NT, 0, e->stream>>>(out, wave_stride,
                                                                                                e->tw.p, args, nsets);
            });
            return tiles;
        }
        if (variant == 11 && mode == mse::COL_PROP && in == out) {  // lean: two blocks per CU
            if constexpr (P0::L == 64 && N <= 512) {  // 16-column tiles, 64 VGPRs
                dim3 grid(N / 16, nsets);
                tiles = N / 16;
                launch(e, kind, [&]() {
                    mse::k_col_prop_lean<P0><<<grid, 16 * P0::L, 0, e->stream>>>(out, wave_stride, e->tw.p, args);
                });
                return tiles;
            } else if constexpr (P0::L == 64 && P0::R3 > 1) {  // four-pass plans: 8-column tiles, 128 VGPRs
                dim3 grid(N / 8, nsets);
                tiles = N / 8;
                launch(e, kind, [&]() {
                    mse::k_col_prop_lean8<P0><<<grid, 8 * P0::L, 0, e->stream>>>(out, wave_stride, e->tw.p, args);
                });
                return tiles;
            }
        }
        if (variant >= 7 && mode == mse::COL_PROP && in == out) {
            constexpr int NT = 16 * P0::L;
            const int per_cu = resident_per_cu<mse::k_col_prop_pipe<P0, 16, mse::WaveSync>>(NT, 1024 / NT);
            const int blocks = std::max(1, std::min(e->num_cus * per_cu, (N / 16) * nsets));
            tiles = N / 16;
            launch(e, kind, [&]() {
                mse::k_col_prop_pipe<P0, 16, mse::WaveSync><<<blocks, NT, 0, e->stream>>>(out, wave_stride, e->tw.p, args,
                                                                                          nsets);
            });
            return tiles;
        }
        if (variant >= 7) variant = 0;
        if (variant == 1)
            go(P1{}, C16{}, mse::WaveSync{}, REG{});
        else if (variant == 2)
            go(P0{}, C8{}, mse::WaveSync{}, REG{});
        else if (variant == 3)
            go(P0{}, C16{}, mse::BlockSync{}, REG{});
        else if (variant == 4)
            go(P0{}, C16{}, mse::WaveSync{}, MEM{});
        else if (variant == 5)
            go(P0{}, C8{}, mse::WaveSync{}, MEM{});
        else if (variant == 6) {
            if constexpr (N <= 512)  // 32 columns exceed the 160 KiB LDS above 512 points
                go(P1{}, C32{}, mse::WaveSync{}, REG{});
            else
                go(P0{}, C16{}, mse::WaveSync{}, REG{});
        }
        else if (!hot && P0::L == 64 && P0::E > 8)  // 16-column tiles of 1024 threads spill here
            go(P0{}, C8{}, mse::WaveSync{}, REG{});
        else
            go(P0{}, C16{}, mse::WaveSync{}, REG{});
    }
    return tiles;
}

}  // namespace mse_rt

#define MSE_INSTANTIATE_SWEEPS(NN)                                                                          \
    template void mse_rt::row_pass_n<NN>(mse_engine*, int, const float2*, float2*, int, size_t,             \
                                         const mse::RowArgs&, int);                                        \
    template int mse_rt::col_pass_n<NN>(mse_engine*, int, const float2*, float2*, int, size_t,              \
                                        const mse::ColArgs&, int, int);
