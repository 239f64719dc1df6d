// CCD response and shot noise of a CTEM image (SimulationCtem::simulateImageDose,
// microscope/simulationctem.cpp:174-263, with ccd_dqe / ccd_ntf, kernels/ccd_dqe_d.cl and
// kernels/ccd_ntf_d.cl).  The two filters are radial tables interpolated per pixel on the
// host in double; each filter is one row pass, one column pass with the filter fused and
// one row pass.  The Poisson noise uses a seeded counter-based generator (clTEM seeds
// std::mt19937_64 from the clock, so its noise cannot be reproduced).
#include "engine_internal.hpp"
#include "imaging_kernels.hpp"

namespace mse_rt {
namespace {

// ccd_dqe_d / ccd_ntf_d: radius in pixels of the unshifted FFT, linear interpolation of the
// table at radius / binning, indices clamped to the last entry.
void radial_filter(int res, const double* table, int n_table, int binning, bool take_sqrt, float* out) {
    for (int yid = 0; yid < res; ++yid)
        for (int xid = 0; xid < res; ++xid) {
            const int midx = xid < res / 2 ? 0 : res;
            const int midy = yid < res / 2 ? 0 : res;
            const double xp = xid - midx, yp = yid - midy;
            const double rad = std::sqrt(xp * xp + yp * yp);
            const int i0 = (int)std::floor(rad / binning);
            const int i1 = i0 + 1;
            const double v0 = table[std::min(i0, n_table - 1)];
            const double v1 = table[std::min(i1, n_table - 1)];
            const double interp = rad / binning - std::floor(rad / binning);
            const double f = interp * v1 + (1 - interp) * v0;
            out[(size_t)yid * res + xid] = (float)(take_sqrt ? std::sqrt(f) : f);
        }
}

void filter_pass(mse_engine* e, float2* z, const float* filter) {
    mse::RowArgs ra = base_row_args(e);
    ra.scale = 1.0f;
    row_pass(e, mse::ROW_FFT, z, z, 1, e->N, ra);
    mse::ColArgs ca = base_col_args(e);
    ca.filter = filter;
    ca.scale = (float)(1.0 / ((double)e->res * (double)e->res));
    col_pass(e, mse::COL_FILTER, z, z, 1, e->N, ca, PK_OTHER);
    row_pass(e, mse::ROW_IFFT, z, z, 1, e->N, ra);
}

}  // namespace
}  // namespace mse_rt

extern "C" int mse_ccd_image(mse_engine* e, const double* image, const double* dqe, const double* ntf, int n_table,
                             int binning, double dose_per_pixel, double conversion, unsigned long long seed,
                             double* out) {
    return run_guarded("mse_ccd_image", e, [&]() {
        REQUIRE(e && image && dqe && ntf && out, MSE_ERR_ARG, "NULL argument");
        require_geom(e);
        REQUIRE(n_table >= 1 && binning >= 1, MSE_ERR_ARG, "table length and binning must be positive");
        REQUIRE(dose_per_pixel >= 0.0 && std::isfinite(dose_per_pixel), MSE_ERR_ARG, "dose must be finite and >= 0");
        for (int i = 0; i < n_table; ++i)
            REQUIRE(dqe[i] >= 0.0 && std::isfinite(dqe[i]) && std::isfinite(ntf[i]), MSE_ERR_ARG,
                    "DQE must be finite and >= 0, NTF finite");
        const size_t N = e->N;
        std::vector<float> f(2 * N);
        radial_filter(e->res, dqe, n_table, binning, true, f.data());
        radial_filter(e->res, ntf, n_table, binning, false, f.data() + N);
        e->ccd_filter.ensure(2 * N);
        e->img_tmp.ensure(N);
        e->diff_img.ensure(N);
        HIPX(hipMemcpyAsync(e->ccd_filter.p, f.data(), 2 * N * sizeof(float), hipMemcpyHostToDevice, e->stream));
        HIPX(hipMemcpyAsync(e->diff_img.p, image, N * sizeof(double), hipMemcpyHostToDevice, e->stream));
        float2* z = e->img_tmp.p;
        const unsigned blocks = (unsigned)((N + 255) / 256);
        launch(e, PK_OTHER, [&]() { mse::k_real_to_complex<<<blocks, 256, 0, e->stream>>>(e->diff_img.p, z, N); });
        filter_pass(e, z, e->ccd_filter.p);  // DQE
        const double n_tot = dose_per_pixel * binning * binning;  // simulationctem.cpp:229
        const uint2 key = make_uint2((uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32));
        launch(e, PK_OTHER, [&]() { mse::k_shot_noise<<<blocks, 256, 0, e->stream>>>(z, N, n_tot, conversion, key); });
        filter_pass(e, z, e->ccd_filter.p + N);  // NTF
        launch(e, PK_OTHER, [&]() { mse::k_real_part<<<blocks, 256, 0, e->stream>>>(z, e->diff_img.p, N); });
        HIPX(hipMemcpyAsync(out, e->diff_img.p, N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}
