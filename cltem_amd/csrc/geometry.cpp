// Scenario geometry: the SimulationManager / SimulationCell / MicroscopeParameters
// quantities the multislice path consumes, restated from the reference's host code
// (the formulas and their evaluation order must match, because slice counts, block
// counts and the padded limits decide the atom binning bit for bit).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/cltem_amd.h"

namespace mse_geom {

// utilities/commonstructs.cpp:5-12
static const double kPi = M_PI;
static const double kEMass = 9.10938291e-031;
static const double kECharge = 1.6021773e-019;
static const double kH = 6.6262e-034;
static const double kC = 299792458.0;

// MicroscopeParameters::Wavelength (utilities/commonstructs.h:269-271)
double wavelength(double voltage) {
    return kH * kC /
           std::sqrt(kECharge * (voltage * 1000) *
                     (2 * kEMass * kC * kC + kECharge * (voltage * 1000))) *
           1e+10;
}

// MicroscopeParameters::Sigma (utilities/commonstructs.h:278-280)
double sigma(double voltage) {
    return (2 * kPi / (wavelength(voltage) * (voltage * 1000))) *
           (kEMass * kC * kC + kECharge * (voltage * 1000)) /
           (2 * kEMass * kC * kC + kECharge * (voltage * 1000));
}

// k-grid in unshifted FFT order (microscope/simulationgeneral.cpp:427-455)
void kgrid(int res, double real_scale, double* kx, double* ky, double* bandwidthkmax) {
    const unsigned resolution = (unsigned)res;
    const double sim_size = real_scale * resolution;
    const unsigned imid = (unsigned)std::floor(static_cast<double>(resolution) / 2.0 + 0.5);
    for (unsigned i = 0; i < resolution; ++i) {
        double v = (i >= imid) ? (int)(i - resolution) / sim_size : (int)i / sim_size;
        kx[i] = v;
        ky[i] = v;
    }
    *bandwidthkmax = std::min(std::abs(ky[imid]), std::abs(kx[imid]));
}

// StemArea::getPixelSimArea (utilities/commonstructs.h:379-396): the pixel's scan position,
// snapped down to a multiple of the full field's pixel size, from the STEM area's start
void frame_position(const mse_geometry* g, int pixel, double* xp, double* yp) {
    const unsigned nx = (unsigned)g->stem_pixels[0];
    const unsigned x_px = (unsigned)pixel % nx;
    const unsigned y_px = (unsigned)pixel / nx;
    const double pos_x = g->stem_scale[0] * x_px;
    const double pos_y = g->stem_scale[1] * y_px;
    *xp = g->stem_start[0] + g->frame_step * std::floor(pos_x / g->frame_step);
    *yp = g->stem_start[1] + g->frame_step * std::floor(pos_y / g->frame_step);
}

}  // namespace mse_geom

extern "C" int mse_plan_geometry(const mse_scenario* s, mse_geometry* g) {
    if (!s || !g || s->resolution < 16 || !(s->slice_thickness > 0) || !(s->voltage_kv > 0))
        return MSE_ERR_ARG;
    if (s->full3d && (s->full3d_integrals < 1 || s->full3d_integrals > 4096)) return MSE_ERR_ARG;
    std::memset(g, 0, sizeof(*g));
    g->resolution = s->resolution;
    g->slice_thickness = s->slice_thickness;
    g->limit_factor = s->limit_factor;
    g->beam_tilt_mrad = s->beam_tilt_mrad;
    g->beam_azimuth = s->beam_azimuth;
    g->full3d = s->full3d;
    g->full3d_integrals = s->full3d_integrals;
    g->stem_pixels[0] = s->stem_pixels[0];
    g->stem_pixels[1] = s->stem_pixels[1];

    // SimulationArea::getCorrectedLimitsX/Y (utilities/commonstructs.cpp:33-59)
    const double* a = s->area;
    double range_x = a[1] - a[0], range_y = a[3] - a[2];
    double range_total = std::max(range_x, range_y) + s->area_padding;
    double x_diff = (range_total - range_x) / 2;
    double y_diff = (range_total - range_y) / 2;
    // paddedFullLimits = corrected + SimulationCell xy padding (simulationmanager.cpp:446-452)
    g->full_x[0] = (a[0] - x_diff) + s->pad_xy[0];
    g->full_x[1] = (a[1] + x_diff) + s->pad_xy[1];
    g->full_y[0] = (a[2] - y_diff) + s->pad_xy[0];
    g->full_y[1] = (a[3] + y_diff) + s->pad_xy[1];
    // parallel STEM, CTEM and CBED simulate the full area (simulationmanager.cpp:485-505)
    for (int i = 0; i < 2; ++i) {
        g->sim_x[i] = g->full_x[i];
        g->sim_y[i] = g->full_y[i];
        g->pad_xy[i] = s->pad_xy[i];
    }
    if (s->stem_pixels[0] > 0 && s->stem_pixels[1] > 0) {
        // StemArea::getScaleX/Y (utilities/commonstructs.cpp:99-107)
        g->stem_scale[0] = (a[1] - a[0]) / s->stem_pixels[0];
        g->stem_scale[1] = (a[3] - a[2]) / s->stem_pixels[1];
    }
    g->stem_start[0] = a[0];
    g->stem_start[1] = a[2];
    if (s->mode == 2 && s->moving_stem_frame) {
        if (!(s->stem_pixels[0] > 0 && s->stem_pixels[1] > 0)) return MSE_ERR_ARG;
        // moving frame: each pixel's field is a zero-width area at the pixel's snapped
        // position, padded by the STEM area padding (getPixelSimArea, commonstructs.h:379-396);
        // the snapping step is the full field's pixel size (simulationmanager.cpp:485-501)
        g->moving_frame = 1;
        const double full_range = (s->pad_xy[1] - s->pad_xy[0]) + ((a[1] + x_diff) - (a[0] - x_diff));
        g->frame_step = full_range / (double)s->resolution;
        // a zero-width area: range_total = padding, half of it each side
        g->frame_half[0] = g->frame_half[1] = (s->area_padding - 0.0) / 2;
        double xp, yp;
        mse_geom::frame_position(g, 0, &xp, &yp);  // pixel 0's field sets realScale
        g->sim_x[0] = (xp - g->frame_half[0]) + s->pad_xy[0];
        g->sim_x[1] = (xp + g->frame_half[0]) + s->pad_xy[1];
        g->sim_y[0] = (yp - g->frame_half[1]) + s->pad_xy[0];
        g->sim_y[1] = (yp + g->frame_half[1]) + s->pad_xy[1];
    }

    // SimulationCell::roundPaddingZ (structure/simulationcell.cpp:43-61)
    const double dz = s->slice_thickness, off = s->slice_offset;
    const int pre_n = (int)std::ceil((std::fabs(s->pad_z[1]) - off) / dz);
    const double pre_pad = off + pre_n * dz;
    const double zw = s->struct_max[2] - s->struct_min[2];
    const int struct_n = (int)std::ceil((zw + off) / dz);
    const double left_over = struct_n * dz - off - zw;
    const int post_n = (int)std::ceil((std::fabs(s->pad_z[0]) - left_over) / dz);
    const double post_pad = post_n * dz + left_over;
    g->pad_z[0] = -post_pad;
    g->pad_z[1] = pre_pad;
    g->z_lims[0] = s->struct_min[2] + g->pad_z[0];
    g->z_lims[1] = s->struct_max[2] + g->pad_z[1];

    // sliceCount / preSliceCount (structure/simulationcell.cpp:75-87)
    unsigned n_slices = (unsigned)std::ceil(((g->z_lims[1] - g->z_lims[0]) / dz) - 0.000001);
    n_slices += (n_slices == 0);
    g->n_slices = (int)n_slices;
    g->pre_slices = (int)(unsigned)(g->pad_z[1] / dz);

    // realScale (simulationmanager.cpp:211-222)
    g->real_scale = std::max(g->sim_x[1] - g->sim_x[0], g->sim_y[1] - g->sim_y[0]) / s->resolution;

    // calculateBlocks and blockScaleX/Y (simulationmanager.cpp:371-408)
    const double xr = g->full_x[1] - g->full_x[0], yr = g->full_y[1] - g->full_y[0];
    const int nb = (int)std::ceil(std::max(xr, yr) / 8.0);
    g->blocks_x = g->blocks_y = nb;
    g->block_scale_x = xr / nb;
    g->block_scale_y = yr / nb;
    // simulationgeneral.cpp:517-521
    g->load_x = (int)std::ceil(8.0 / g->block_scale_x);
    g->load_y = (int)std::ceil(8.0 / g->block_scale_y);
    g->load_z = (int)std::ceil(3.0 / dz);

    g->wavelength = mse_geom::wavelength(s->voltage_kv);
    g->wavenumber = 1.0 / g->wavelength;
    g->sigma = mse_geom::sigma(s->voltage_kv);
    // MicroscopeParameters::Wavevector (utilities/commonstructs.h:282-289)
    const double k = g->wavenumber;
    g->k_vec[0] = k * std::sin(s->beam_tilt_mrad / 1000.0) * std::cos(s->beam_azimuth);
    g->k_vec[1] = k * std::sin(s->beam_tilt_mrad / 1000.0) * std::sin(s->beam_azimuth);
    g->k_vec[2] = k * std::cos(s->beam_tilt_mrad / 1000.0);
    g->sigma_eff = g->sigma * g->wavenumber / g->k_vec[2];  // simulationgeneral.cpp:418
    g->inverse_scale = 1.0 / (g->real_scale * s->resolution);
    g->inverse_scale_angle = 1000.0 * g->inverse_scale * g->wavelength;

    {
        double* kx = new double[s->resolution];
        double* ky = new double[s->resolution];
        mse_geom::kgrid(s->resolution, g->real_scale, kx, ky, &g->kmax_band);
        delete[] kx;
        delete[] ky;
    }
    return MSE_OK;
}

extern "C" int mse_stem_pixel_frame(const mse_geometry* g, int pixel, double* start_xy) {
    if (!g || !start_xy || pixel < 0) return MSE_ERR_ARG;
    if (!g->moving_frame) {
        start_xy[0] = g->sim_x[0];
        start_xy[1] = g->sim_y[0];
        return MSE_OK;
    }
    if (pixel >= g->stem_pixels[0] * g->stem_pixels[1]) return MSE_ERR_ARG;
    double xp, yp;
    mse_geom::frame_position(g, pixel, &xp, &yp);
    // the padded lower limits of the zero-width area at (xp, yp)
    start_xy[0] = (xp - g->frame_half[0]) + g->pad_xy[0];
    start_xy[1] = (yp - g->frame_half[1]) + g->pad_xy[0];
    return MSE_OK;
}
