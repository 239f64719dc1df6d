// CPU restatement of clTEM's multislice path -- TEST INFRASTRUCTURE ONLY.
//
// The HIP product never links or calls this file.  It is the oracle the tests
// compare the product with, and (through orc_run_stem_f32) the CPU baseline that
// bench.py times.  Built with -ffp-contract=off so that every double expression is
// evaluated op for op as written in the reference's `_d` kernels.
//
// Parity status of each function is listed in mse_oracle.h.
#include "mse_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------
// Physical constants and microscope quantities (utilities/commonstructs.cpp:5-12,
// utilities/commonstructs.h:269-289).  Expressions keep the reference's evaluation order.
// ---------------------------------------------------------------------------------------
const double kPi = M_PI;
const double kEMass = 9.10938291e-031;
const double kECharge = 1.6021773e-019;
const double kH = 6.6262e-034;
const double kC = 299792458.0;

double wavelength(double voltage) {
    return kH * kC /
           std::sqrt(kECharge * (voltage * 1000) *
                     (2 * kEMass * kC * kC + kECharge * (voltage * 1000))) *
           1e+10;
}

double sigma(double voltage) {
    return (2 * kPi / (wavelength(voltage) * (voltage * 1000))) *
           (kEMass * kC * kC + kECharge * (voltage * 1000)) /
           (2 * kEMass * kC * kC + kECharge * (voltage * 1000));
}

// SimulationArea::getCorrectedLimitsX/Y (utilities/commonstructs.cpp:33-59)
void corrected_limits(const double* a, double padding, double* lx, double* ly) {
    double range_x = a[1] - a[0];
    double range_y = a[3] - a[2];
    double range_max = std::max(range_x, range_y);
    double range_total = range_max + padding;
    double x_diff = (range_total - range_x) / 2;
    double y_diff = (range_total - range_y) / 2;
    lx[0] = a[0] - x_diff;
    lx[1] = a[1] + x_diff;
    ly[0] = a[2] - y_diff;
    ly[1] = a[3] + y_diff;
}

// SimulationManager::currentAreaBase for STEM without parallel pixels
// (simulationmanager.cpp:485-501): StemArea::getPixelSimArea(pixel, rs)
// (utilities/commonstructs.h:379-396), whose corrected limits are returned.  rs is the pixel
// size of the full field (cell padding range + corrected STEM area range over the resolution).
void moving_frame_area(const OrcScenario* s, int pixel, double* lx, double* ly) {
    const double pad_range = s->pad_xy[1] - s->pad_xy[0];
    double area_x[2], area_y[2];
    corrected_limits(s->area, s->area_padding, area_x, area_y);
    const double area_range = area_x[1] - area_x[0];
    const double rs = (pad_range + area_range) / static_cast<double>(s->resolution);
    const unsigned int xPixels = (unsigned int)s->stem_pixels[0];
    const unsigned int yPixels = (unsigned int)s->stem_pixels[1];
    const unsigned int x_px = pixel % xPixels;
    const unsigned int y_px = (unsigned int)std::floor(pixel / xPixels);
    // StemArea::getScaleX/Y: (finish - start) / pixels
    const double pos_x = ((s->area[1] - s->area[0]) / xPixels) * x_px;
    const double pos_y = ((s->area[3] - s->area[2]) / yPixels) * y_px;
    const double real_pos_x = rs * std::floor(pos_x / rs);
    const double real_pos_y = rs * std::floor(pos_y / rs);
    const double xPos = s->area[0] + real_pos_x;
    const double yPos = s->area[2] + real_pos_y;
    const double frame[4] = {xPos, xPos, yPos, yPos};
    corrected_limits(frame, s->area_padding, lx, ly);
}

// ---------------------------------------------------------------------------------------
// Bessel K0/K1: Numerical Recipes in C 2nd ed. section 6.6 rational forms, with the
// coefficient tables the reference kernel uses
// (kernels/transmission_potentials_projected_d.cl:57-102).  Published method; checked
// against scipy.special.k0/k1 in tests/test_oracle_potentials.py.
// ---------------------------------------------------------------------------------------
const double k0pi[5] = {1.0, 2.346487949187396e-1, 1.187082088663404e-2, 2.150707366040937e-4,
                        1.425433617130587e-6};
const double k0qi[3] = {9.847324170755358e-1, 1.518396076767770e-2, 8.362215678646257e-5};
const double k0p[5] = {1.159315156584126e-1, 2.770731240515333e-1, 2.066458134619875e-2,
                       4.574734709978264e-4, 3.454715527986737e-6};
const double k0q[3] = {9.836249671709183e-1, 1.627693622304549e-2, 9.809660603621949e-5};
const double k0pp[8] = {1.253314137315499, 1.475731032429900e1, 6.123767403223466e1,
                        1.121012633939949e2, 9.285288485892228e1, 3.198289277679660e1,
                        3.595376024148513, 6.160228690102976e-2};
const double k0qq[8] = {1.0, 1.189963006673403e1, 5.027773590829784e1, 9.496513373427093e1,
                        8.318077493230258e1, 3.181399777449301e1, 4.443672926432041,
                        1.408295601966600e-1};
const double k1pi[5] = {0.5, 5.598072040178741e-2, 1.818666382168295e-3, 2.397509908859959e-5,
                        1.239567816344855e-7};
const double k1qi[3] = {9.870202601341150e-1, 1.292092053534579e-2, 5.881933053917096e-5};
const double k1p[5] = {-3.079657578292062e-1, -8.109417631822442e-2, -3.477550948593604e-3,
                       -5.385594871975406e-5, -3.110372465429008e-7};
const double k1q[3] = {9.861813171751389e-1, 1.375094061153160e-2, 6.774221332947002e-5};
const double k1pp[8] = {1.253314137315502, 1.457171340220454e1, 6.063161173098803e1,
                        1.147386690867892e2, 1.040442011439181e2, 4.356596656837691e1,
                        7.265230396353690, 3.144418558991021e-1};
const double k1qq[8] = {1.0, 1.125154514806458e1, 4.427488496597630e1, 7.616113213117645e1,
                        5.863377227890893e1, 1.850303673841586e1, 1.857244676566022,
                        2.538540887654872e-2};

template <typename T>
T horner(const double* cof, int n, T x) {  // poly() in the kernel
    T ans = (T)cof[n];
    for (int i = n - 1; i >= 0; --i) ans = ans * x + (T)cof[i];
    return ans;
}

template <typename T>
T bessk0(T x) {
    T ax = std::fabs(x);
    if (ax > (T)0 && ax <= (T)1) {
        T z = x * x;
        T term = horner(k0pi, 4, z) * std::log(x) * ((T)1 / horner(k0qi, 2, (T)1 - z));
        return horner(k0p, 4, z) * ((T)1 / horner(k0q, 2, (T)1 - z)) - term;
    } else if (ax > (T)1) {
        T z = (T)1 / x;
        return std::exp(-x) * horner(k0pp, 7, z) * ((T)1 / horner(k0qq, 7, z)) *
               ((T)1 / std::sqrt(x));
    }
    return sizeof(T) == 8 ? (T)DBL_MAX : (T)FLT_MAX;
}

template <typename T>
T bessk1(T x) {
    T ax = std::fabs(x);
    if (ax > (T)0 && ax <= (T)1) {
        T z = x * x;
        T term = horner(k1pi, 4, z) * std::log(x) * ((T)1 / horner(k1qi, 2, (T)1 - z));
        return x * (horner(k1p, 4, z) * ((T)1 / horner(k1q, 2, (T)1 - z)) + term) + (T)1 / x;
    } else if (ax > (T)1) {
        T z = (T)1 / x;
        return std::exp(-x) * horner(k1pp, 7, z) * ((T)1 / horner(k1qq, 7, z)) *
               ((T)1 / std::sqrt(x));
    }
    return sizeof(T) == 8 ? (T)DBL_MAX : (T)FLT_MAX;
}

// Projected potentials (kernels/transmission_potentials_projected_d.cl:104-181).
double v_kirkland(const double* params, int i_lim, int Z, double rad) {
    double suml = 0.0, sumg = 0.0;
    int z_ofst = (Z - 1) * 12;
    double x = 2.0 * M_PI * rad;
    for (int i = 0; i < i_lim * 2; i += 2) {
        double a = params[z_ofst + i];
        double b = params[z_ofst + i + 1];
        suml += a * bessk0(x * std::sqrt(b));
    }
    x = M_PI * rad;
    x = x * x;
    for (int i = i_lim * 2; i < i_lim * 4; i += 2) {
        double c = params[z_ofst + i];
        double d = params[z_ofst + i + 1];
        double d_inv = 1.0 / d;
        sumg += (c * d_inv) * std::exp(-x * d_inv);
    }
    return 300.8242834 * suml + 150.4121417 * sumg;
}

double v_lobato(const double* params, int i_lim, int Z, double rad) {
    double sum = 0.0;
    int z_ofst = (Z - 1) * 10;
    double x = 2.0 * M_PI * rad;
    for (int i = 0; i < i_lim; ++i) {
        double a = params[z_ofst + i];
        double b = params[z_ofst + i + 5];
        double b_inv_root = 1.0 / std::sqrt(b);
        sum += a * (b_inv_root * b_inv_root * b_inv_root) *
               (bessk0(x * b_inv_root) + rad * bessk1(x * b_inv_root));
    }
    return 945.090144399935 * sum;
}

double v_peng(const double* params, int i_lim, int Z, double rad) {
    double sum = 0.0;
    int z_ofst = (Z - 1) * 10;
    double x = M_PI * rad;
    x = x * x;
    for (int i = 0; i < i_lim; ++i) {
        double a = params[z_ofst + i];
        double b = params[z_ofst + i + 5];
        double b_inv = 1.0 / b;
        sum += a * b_inv * std::exp(-x * b_inv);
    }
    return 150.4121417 * sum;
}

double vproj(int form, const double* params, int i_lim, int Z, double rad) {
    if (form == 0) return v_kirkland(params, i_lim, Z, rad);
    if (form == 1) return v_peng(params, i_lim, Z, rad);
    if (form == 2) return v_lobato(params, i_lim, Z, rad);
    return 0.0;
}

// Full 3D potentials (kernels/transmission_potentials_full_3d_d.cl:57-134).
double v3_kirkland(const double* params, int i_lim, int Z, double rad) {
    double suml = 0.0, sumg = 0.0;
    int z_ofst = (Z - 1) * 12;
    double x = 2.0 * M_PI * rad;
    for (int i = 0; i < i_lim * 2; i += 2) {
        double a = params[z_ofst + i];
        double b = params[z_ofst + i + 1];
        suml += a * std::exp(-x * std::sqrt(b));
    }
    x = M_PI * rad;
    x = x * x;
    for (int i = i_lim * 2; i < i_lim * 4; i += 2) {
        double c = params[z_ofst + i];
        double d = params[z_ofst + i + 1];
        double d_inv_root = 1.0 / std::sqrt(d);
        sumg += c * (d_inv_root * d_inv_root * d_inv_root) * std::exp(-x * (1.0 / d));
    }
    return 150.4121417 * (1.0 / rad) * suml + 266.5157269 * sumg;
}

double v3_lobato(const double* params, int i_lim, int Z, double rad) {
    double sum = 0.0;
    int z_ofst = (Z - 1) * 10;
    double x = M_PI * rad;
    for (int i = 0; i < i_lim; ++i) {
        double a = params[z_ofst + i];
        double b = params[z_ofst + i + 5];
        double b_inv_root = 1.0 / std::sqrt(b);
        sum += a * (b_inv_root * b_inv_root * b_inv_root) * std::exp(-2.0 * x * b_inv_root) *
               (std::sqrt(b) * (1.0 / x) + 1.0);
    }
    return 472.545072199968 * sum;
}

double v3_peng(const double* params, int i_lim, int Z, double rad) {
    double sum = 0.0;
    int z_ofst = (Z - 1) * 10;
    double x = M_PI * rad;
    x = x * x;
    for (int i = 0; i < i_lim; ++i) {
        double a = params[z_ofst + i];
        double b = params[z_ofst + i + 5];
        double b_inv_root = 1.0 / std::sqrt(b);
        sum += a * (b_inv_root * b_inv_root * b_inv_root) * std::exp(-x * (1.0 / b));
    }
    return 266.5157269 * sum;
}

double v3d(int form, const double* params, int i_lim, int Z, double rad) {
    if (form == 0) return v3_kirkland(params, i_lim, Z, rad);
    if (form == 1) return v3_peng(params, i_lim, Z, rad);
    if (form == 2) return v3_lobato(params, i_lim, Z, rad);
    return 0.0;
}

// Work-group block range of the potential kernels
// (kernels/transmission_potentials_projected_d.cl:227-246, full_3d_d.cl:191-210).
struct BlockRange {
    int starti, endi, startj, endj;
};

BlockRange block_range(const OrcGeometry* g, int gx, int gy) {
    const double pixelscale = g->real_scale;
    double startx = g->sim_x[0], starty = g->sim_y[0];
    double group_size_x = 16 * pixelscale;
    double group_size_y = 16 * pixelscale;
    double group_start_x = startx + gx * group_size_x;
    double group_end_x = group_start_x + group_size_x;
    double group_start_y = starty + gy * group_size_y;
    double group_end_y = group_start_y + group_size_y;
    double min_x = g->full_x[0], max_x = g->full_x[1];
    double min_y = g->full_y[0], max_y = g->full_y[1];
    double recip_range_x = 1.0 / (max_x - min_x);
    double recip_range_y = 1.0 / (max_y - min_y);
    BlockRange b;
    b.starti = (int)std::fmax(
        std::floor(g->blocks_x * (group_start_x - min_x) * recip_range_x) - g->load_x, 0);
    b.endi = (int)std::fmin(
        std::ceil(g->blocks_x * (group_end_x - min_x) * recip_range_x) + g->load_x,
        g->blocks_x - 1);
    b.startj = (int)std::fmax(
        std::floor(g->blocks_y * (group_start_y - min_y) * recip_range_y) - g->load_y, 0);
    b.endj = (int)std::fmin(
        std::ceil(g->blocks_y * (group_end_y - min_y) * recip_range_y) + g->load_y,
        g->blocks_y - 1);
    return b;
}

// transmission_potentials_projected_d: raw sum over the slice's atoms for one pixel.
double pixel_projected(const OrcGeometry* g, const double* sx, const double* sy, const int* sZ,
                       const int* bsp, const double* params, int form, int icount, int slice,
                       int xid, int yid, const BlockRange& br) {
    const double pixelscale = g->real_scale;
    double startx = g->sim_x[0], starty = g->sim_y[0];
    double beam_theta = M_PI_2 - g->beam_tilt_mrad * 0.001;
    double beam_phi = g->beam_azimuth;
    int k = slice;
    if (k < 0) k = 0;
    if (k >= g->n_slices) k = g->n_slices - 1;
    const int bx = g->blocks_x, by = g->blocks_y;
    double sumz = 0.0;
    for (int j = br.startj; j <= br.endj; j++) {
        int start = bsp[k * bx * by + bx * j + br.starti];
        int end = bsp[k * bx * by + bx * j + br.endi + 1];
        for (int l = start; l < end; l++) {
            double im_pos_x = startx + xid * pixelscale;
            double rad_x = im_pos_x - sx[l];
            double im_pos_y = starty + yid * pixelscale;
            double rad_y = im_pos_y - sy[l];
            double cos_beam_phi = std::cos(beam_phi);
            double sin_beam_phi = std::sin(beam_phi);
            double sin_beam_2theta = std::sin(2.0 * beam_theta);
            double z_prime = -0.5 * (rad_x * cos_beam_phi + rad_y * sin_beam_phi) * sin_beam_2theta;
            double z_by_tan_beam_theta = z_prime / std::tan(beam_theta);
            double x_prime = rad_x + z_by_tan_beam_theta * cos_beam_phi;
            double y_prime = rad_y + z_by_tan_beam_theta * sin_beam_phi;
            double rad = std::sqrt(z_prime * z_prime + x_prime * x_prime + y_prime * y_prime);
            double r_min = 0.25 * pixelscale;
            if (rad < r_min) rad = r_min;
            if (rad <= 8.0) sumz += vproj(form, params, icount, sZ[l], rad);
        }
    }
    return sumz;
}

// transmission_potentials_full_3d_d: raw trapezoid sum for one pixel (reproduces the
// cumulative rad_x shift and the stale p2 carried between atoms of a block row).
double pixel_full3d(const OrcGeometry* g, const double* sx, const double* sy, const double* sz,
                    const int* sZ, const int* bsp, const double* params, int form, int icount,
                    int slice, int xid, int yid, const BlockRange& br) {
    const double pixel_scale = g->real_scale;
    double startx = g->sim_x[0], starty = g->sim_y[0];
    const int integrals = g->full3d_integrals;
    const double dz = g->slice_thickness;
    const double current_z = g->z_lims[0] + (g->n_slices - slice) * dz;
    const double slice_shift_x = (g->k_vec[0] / g->k_vec[2]) * dz / integrals;
    const double slice_shift_y = (g->k_vec[1] / g->k_vec[2]) * dz / integrals;
    int topz = slice - g->load_z;
    int bottomz = slice + g->load_z;
    double int_r = 1.0 / integrals;
    if (topz < 0) topz = 0;
    if (bottomz >= g->n_slices) bottomz = g->n_slices - 1;
    const int bx = g->blocks_x, by = g->blocks_y;
    double sumz = 0.0;
    for (int k = topz; k <= bottomz; k++) {
        for (int j = br.startj; j <= br.endj; j++) {
            int start = bsp[k * bx * by + bx * j + br.starti];
            int end = bsp[k * bx * by + bx * j + br.endi + 1];
            double p2 = 0.0;
            for (int l = start; l < end; l++) {
                double im_pos_x = startx + xid * pixel_scale;
                double rad_x = im_pos_x - sx[l];
                double im_pos_y = starty + yid * pixel_scale;
                double rad_y = im_pos_y - sy[l];
                for (int h = 0; h <= integrals; h++) {
                    rad_x -= slice_shift_x;
                    rad_y -= slice_shift_y;
                    double xyrad2 = rad_x * rad_x + rad_y * rad_y;
                    double im_pos_z = current_z - h * dz * int_r;
                    double rad_z = im_pos_z - sz[l];
                    double rad = std::sqrt(xyrad2 + rad_z * rad_z);
                    double r_min = 0.25 * pixel_scale;
                    if (rad < r_min) rad = r_min;
                    if (xyrad2 <= 64.0 && rad_z <= 3.0) {
                        double p1 = v3d(form, params, icount, sZ[l], rad);
                        sumz += (h != 0) * (p1 + p2) * 0.5;
                        p2 = p1;
                    }
                }
            }
        }
    }
    return sumz;
}

template <typename F>
void parallel_for(int n, int nthreads, F f) {
    if (nthreads <= 1 || n <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    nthreads = std::min(nthreads, n);
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&, t]() {
            for (int i = t; i < n; i += nthreads) f(i);
        });
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------------------
// FFT.  clFFT is a third-party library (unpinned, not vendored); clTEM calls it as a 2D
// out-of-place interleaved transform with scale 1/sqrt(W*H) in both directions and the
// exp(-2 pi i jk/N) sign for forwards (clwrapper/clfourier.cpp:30-73).  This is an
// independent recursive mixed-radix (2,3,4,5) decimation-in-time transform with twiddles
// evaluated in double; it is pinned against numpy.fft.fft2(norm="ortho").
// ---------------------------------------------------------------------------------------
template <typename T>
struct Fft1 {
    typedef std::complex<T> C;
    int n = 0;
    std::vector<int> fac;
    std::vector<C> tw;  // tw[j] = exp(-2 pi i j / n)

    explicit Fft1(int n_) : n(n_) {
        int m = n;
        while (m % 4 == 0) { fac.push_back(4); m /= 4; }
        while (m % 2 == 0) { fac.push_back(2); m /= 2; }
        while (m % 3 == 0) { fac.push_back(3); m /= 3; }
        while (m % 5 == 0) { fac.push_back(5); m /= 5; }
        if (m != 1) fac.push_back(m);  // generic DFT for any other factor
        tw.resize(n);
        for (int j = 0; j < n; ++j) {
            // reduce the angle exactly before evaluating
            long double a = -2.0L * (long double)M_PI * (long double)j / (long double)n;
            tw[j] = C((T)std::cos((double)a), (T)std::sin((double)a));
        }
    }

    C w(long j, int tws, bool inv) const {
        C v = tw[(size_t)((j * tws) % n)];
        return inv ? std::conj(v) : v;
    }

    void rec(const C* in, int is, C* out, int m, int fi, int tws, bool inv) const {
        if (m == 1) {
            out[0] = in[0];
            return;
        }
        int R = fac[fi];
        int q = m / R;
        for (int r = 0; r < R; ++r) rec(in + r * is, is * R, out + r * q, q, fi + 1, tws * R, inv);
        std::vector<C> a(R), y(R);
        for (int k = 0; k < q; ++k) {
            for (int r = 0; r < R; ++r) a[r] = (r == 0) ? out[k] : out[r * q + k] * w((long)r * k, tws, inv);
            if (R == 2) {
                y[0] = a[0] + a[1];
                y[1] = a[0] - a[1];
            } else if (R == 4) {
                C t0 = a[0] + a[2], t1 = a[0] - a[2];
                C t2 = a[1] + a[3], t3 = a[1] - a[3];
                // multiply t3 by -i (forward) or +i (inverse)
                C t3r = inv ? C(-t3.imag(), t3.real()) : C(t3.imag(), -t3.real());
                y[0] = t0 + t2;
                y[2] = t0 - t2;
                y[1] = t1 + t3r;
                y[3] = t1 - t3r;
            } else {
                for (int s = 0; s < R; ++s) {
                    C acc = a[0];
                    for (int r = 1; r < R; ++r) acc += a[r] * w((long)((r * s) % R) * q, tws, inv);
                    y[s] = acc;
                }
            }
            for (int s = 0; s < R; ++s) out[k + s * q] = y[s];
        }
    }

    void run(C* data, int stride, bool inv, C* scratch_in, C* scratch_out) const {
        for (int i = 0; i < n; ++i) scratch_in[i] = data[(size_t)i * stride];
        rec(scratch_in, 1, scratch_out, n, 0, 1, inv);
        for (int i = 0; i < n; ++i) data[(size_t)i * stride] = scratch_out[i];
    }
};

template <typename T>
struct Fft2 {
    typedef std::complex<T> C;
    int w, h;
    Fft1<T> fx, fy;
    T scale;
    Fft2(int w_, int h_) : w(w_), h(h_), fx(w_), fy(h_) {
        if (sizeof(T) == 4)
            scale = (T)(1.0f / std::sqrt((float)(w_ * h_)));  // clfftSetPlanScale(1.0f/sqrtf(W*H))
        else
            scale = (T)(1.0 / std::sqrt((double)w_ * (double)h_));
    }
    void run(C* d, bool inv) const {
        std::vector<C> s1(std::max(w, h)), s2(std::max(w, h));
        for (int y = 0; y < h; ++y) fx.run(d + (size_t)y * w, 1, inv, s1.data(), s2.data());
        for (int x = 0; x < w; ++x) fy.run(d + x, w, inv, s1.data(), s2.data());
        for (size_t i = 0; i < (size_t)w * h; ++i) d[i] *= scale;
    }
};

// The CPU baseline's FFT for power-of-two sizes: an iterative Stockham transform (radix-4
// stages and one radix-2 stage when log2 n is odd), twiddles per stage computed once, no
// allocation per call; columns go through it eight at a time from a contiguous copy.  The
// same sign convention and 1/sqrt(W*H) scale as Fft2; checked against Fft2 (and so against
// numpy) by orc_fft2_fast.  The checker itself keeps Fft2.
template <typename T>
struct StockhamFft1 {
    typedef std::complex<T> C;
    int n = 0;
    std::vector<int> radix;            // stage radices, 4s first, then one 2 if needed
    std::vector<std::vector<C>> tw;    // stage of length L: w^(r p), w = exp(-2 pi i / L),
                                       // p < L / R, r = 1..R-1 (at [(r - 1) (L / R) + p])

    explicit StockhamFft1(int n_) : n(n_) {
        int m = n;
        while (m % 4 == 0) { radix.push_back(4); m /= 4; }
        if (m == 2) radix.push_back(2);
        int len = n;
        for (int R : radix) {
            const int q = len / R;
            std::vector<C> w((size_t)(R - 1) * q);
            for (int r = 1; r < R; ++r)
                for (int p = 0; p < q; ++p) {
                    long double a = -2.0L * (long double)M_PI * (long double)(r * p) / (long double)len;
                    w[(size_t)(r - 1) * q + p] = C((T)std::cos((double)a), (T)std::sin((double)a));
                }
            tw.push_back(std::move(w));
            len = q;
        }
    }
    static bool fits(int n) {
        return n >= 2 && (n & (n - 1)) == 0;
    }
    // the plain complex product (std::complex's operator* also handles infinities and NaNs
    // out of line)
    static C mul(const C& u, const C& v) {
        return C(u.real() * v.real() - u.imag() * v.imag(), u.real() * v.imag() + u.imag() * v.real());
    }
    // Stockham decimation in frequency: the stage of length L and stride s (L s = n) takes
    // a[t + s (p + r L/R)], r < R, through the R-point DFT, multiplies output r by w^(r p)
    // and writes it to b[t + s (R p + r)].  x (n points) is transformed in place, y is scratch.
    void run(C* x, C* y, bool inv) const {
        int len = n, s = 1;
        C* a = x;
        C* b = y;
        for (size_t st = 0; st < radix.size(); ++st) {
            const int R = radix[st];
            const int q = len / R;
            const C* w = tw[st].data();
            for (int p = 0; p < q; ++p) {
                C w1 = w[p], w2(1, 0), w3(1, 0);
                if (R == 4) {
                    w2 = w[(size_t)q + p];
                    w3 = w[(size_t)2 * q + p];
                }
                if (inv) {
                    w1 = std::conj(w1);
                    w2 = std::conj(w2);
                    w3 = std::conj(w3);
                }
                const C* in = a + (size_t)s * p;
                C* out = b + (size_t)s * R * p;
                const size_t sq = (size_t)s * q;
                if (R == 2) {
                    for (int t = 0; t < s; ++t) {
                        const C c0 = in[t], c1 = in[t + sq];
                        out[t] = c0 + c1;
                        out[t + s] = mul(c0 - c1, w1);
                    }
                } else {
                    for (int t = 0; t < s; ++t) {
                        const C c0 = in[t], c1 = in[t + sq], c2 = in[t + 2 * sq], c3 = in[t + 3 * sq];
                        const C apc = c0 + c2, amc = c0 - c2, bpd = c1 + c3, bmd = c1 - c3;
                        // (b - d) times -i (forward) or +i (inverse)
                        const C jb = inv ? C(-bmd.imag(), bmd.real()) : C(bmd.imag(), -bmd.real());
                        out[t] = apc + bpd;
                        out[t + s] = mul(amc + jb, w1);
                        out[t + 2 * s] = mul(apc - bpd, w2);
                        out[t + 3 * s] = mul(amc - jb, w3);
                    }
                }
            }
            std::swap(a, b);
            len = q;
            s *= R;
        }
        if (a != x) std::memcpy((void*)x, (const void*)a, (size_t)n * sizeof(C));
    }
};

template <typename T>
struct FastFft2 {
    typedef std::complex<T> C;
    int w, h;
    StockhamFft1<T> fx, fy;
    T scale;
    FastFft2(int w_, int h_) : w(w_), h(h_), fx(w_), fy(h_) {
        if (sizeof(T) == 4)
            scale = (T)(1.0f / std::sqrt((float)(w_ * h_)));  // clfftSetPlanScale(1.0f/sqrtf(W*H))
        else
            scale = (T)(1.0 / std::sqrt((double)w_ * (double)h_));
    }
    void run(C* d, bool inv) const {
        constexpr int G = 8;  // columns per gather
        thread_local std::vector<C> cols, scratch;
        cols.resize((size_t)G * h);
        scratch.resize((size_t)std::max(w, h));
        for (int y = 0; y < h; ++y) fx.run(d + (size_t)y * w, scratch.data(), inv);
        for (int x0 = 0; x0 < w; x0 += G) {
            const int g = std::min(G, w - x0);
            for (int y = 0; y < h; ++y)
                for (int c = 0; c < g; ++c) cols[(size_t)c * h + y] = d[(size_t)y * w + x0 + c];
            for (int c = 0; c < g; ++c) fy.run(cols.data() + (size_t)c * h, scratch.data(), inv);
            for (int y = 0; y < h; ++y)
                for (int c = 0; c < g; ++c) d[(size_t)y * w + x0 + c] = cols[(size_t)c * h + y] * scale;
        }
    }
};

// k-grid of SimulationGeneral::initialiseSimulation (microscope/simulationgeneral.cpp:427-455)
template <typename T>
void kgrid(const OrcGeometry* g, std::vector<T>& kx, std::vector<T>& ky, double* bandwidthkmax) {
    unsigned int resolution = (unsigned int)g->resolution;
    double SimSizeX = g->real_scale * resolution;
    double SimSizeY = SimSizeX;
    kx.assign(resolution, 0);
    ky.assign(resolution, 0);
    unsigned int imidx = (unsigned int)std::floor(static_cast<double>(resolution) / 2.0 + 0.5);
    unsigned int imidy = imidx;
    double temp;
    for (int i = 0; i < (int)resolution; i++) {
        if ((unsigned)i >= imidx)
            temp = (int)(i - resolution) / SimSizeX;
        else
            temp = i / SimSizeX;
        kx[i] = (T)temp;
    }
    for (int i = 0; i < (int)resolution; i++) {
        if ((unsigned)i >= imidy)
            temp = (int)(i - resolution) / SimSizeY;
        else
            temp = i / SimSizeY;
        ky[i] = (T)temp;
    }
    T kmaxx = std::abs(kx[imidx]);
    T kmaxy = std::abs(ky[imidy]);
    *bandwidthkmax = std::min(kmaxy, kmaxx);
}

// band_limit_{d,f} (kernels/band_limit_d.cl:13-32): keep |k| <= k_max * limit_factor
template <typename T>
void band_limit_mask(const OrcGeometry* g, std::complex<T>* d) {
    std::vector<T> kx, ky;
    double kmax;
    kgrid<T>(g, kx, ky, &kmax);
    const int res = g->resolution;
    T k_max = (T)kmax;
    T limit_factor = (T)g->limit_factor;
    T lim = k_max * limit_factor;
    for (int y = 0; y < res; ++y)
        for (int x = 0; x < res; ++x) {
            T k = std::sqrt(kx[x] * kx[x] + ky[y] * ky[y]);
            T keep = (T)(k <= lim);
            d[(size_t)y * res + x] = std::complex<T>(d[(size_t)y * res + x].real() * keep,
                                                     d[(size_t)y * res + x].imag() * keep);
        }
}

// propagator_{d,f} (kernels/propagator_d.cl:18-51), args from simulationgeneral.cpp:623-633
template <typename T>
void propagator(const OrcGeometry* g, std::complex<T>* P) {
    std::vector<T> kx, ky;
    double bw;
    kgrid<T>(g, kx, ky, &bw);
    const int res = g->resolution;
    T dz = (T)g->slice_thickness;
    T beam_k = (T)g->wavenumber;
    T beam_k_x = (T)g->k_vec[0], beam_k_y = (T)g->k_vec[1], beam_k_z = (T)g->k_vec[2];
    T k_max = (T)(bw * g->limit_factor);
    const T pi = (T)M_PI;
    for (int y = 0; y < res; ++y)
        for (int x = 0; x < res; ++x) {
            T k0x = kx[x] * kx[x];
            T k0y = ky[y] * ky[y];
            T k_max_2 = k_max * k_max;
            std::complex<T> v(0, 0);
            if (k0x + k0y < k_max_2) {
                T f = beam_k_z / beam_k;
                T s_u = beam_k * beam_k - beam_k_z * beam_k_z -
                        (beam_k_x + kx[x]) * (beam_k_x + kx[x]) -
                        (beam_k_y + ky[y]) * (beam_k_y + ky[y]);
                s_u = s_u / (beam_k_z);
                v = std::complex<T>(f * std::cos(pi * s_u * dz), f * std::sin(pi * s_u * dz));
            }
            P[(size_t)y * res + x] = v;
        }
}

// init_probe_wave_{d,f} (kernels/init_probe_wave_d.cl:28-118).  taper=1 gives the `_f`
// kernel's smoothstep edge (init_probe_wave_f.cl:103-110); the `_d` kernel omits it.
template <typename T>
void probe_kspace(const OrcGeometry* g, T pos_x, T pos_y, const double* ab, T cond_ap,
                  T ap_smooth, int taper, std::complex<T>* out) {
    typedef std::complex<T> C;
    std::vector<T> kx, ky;
    double bw;
    kgrid<T>(g, kx, ky, &bw);
    const int res = g->resolution;
    const T wavelength = (T)g->wavelength;
    const T pi = (T)M_PI;
    auto cpow = [](C a, int n) {
        C temp = a;
        for (int j = 1; j < n; j++) temp = temp * a;
        return temp;
    };
    auto modsq = [](C a) { return a.real() * a.real() + a.imag() * a.imag(); };
    const T C10 = (T)ab[0];
    const C C12((T)ab[1], (T)ab[2]), C21((T)ab[3], (T)ab[4]), C23((T)ab[5], (T)ab[6]);
    const T C30 = (T)ab[7];
    const C C32((T)ab[8], (T)ab[9]), C34((T)ab[10], (T)ab[11]);
    const C C41((T)ab[12], (T)ab[13]), C43((T)ab[14], (T)ab[15]), C45((T)ab[16], (T)ab[17]);
    const T C50 = (T)ab[18];
    const C C52((T)ab[19], (T)ab[20]), C54((T)ab[21], (T)ab[22]), C56((T)ab[23], (T)ab[24]);
    for (int yid = 0; yid < res; ++yid)
        for (int xid = 0; xid < res; ++xid) {
            size_t id = (size_t)xid + (size_t)yid * res;
            T cond_ap2 = (cond_ap * (T)0.001) / wavelength;
            T k = std::sqrt((kx[xid] * kx[xid]) + (ky[yid] * ky[yid]));
            T ap_smooth_radius = (ap_smooth * (T)0.001) / wavelength;
            if (k < cond_ap2 + ap_smooth_radius) {
                T posTerm = (T)2.0 * pi * (kx[xid] * pos_x + ky[yid] * pos_y);
                C w(wavelength * kx[xid], wavelength * ky[yid]);
                C wc = std::conj(w);
                T tC10 = (T)0.5 * C10 * modsq(w);
                C tC12 = (T)0.5 * (C12 * cpow(wc, 2));
                C tC21 = (C21 * (cpow(wc, 2) * w)) / (T)3.0;
                C tC23 = (C23 * cpow(wc, 3)) / (T)3.0;
                T tC30 = (T)0.25 * C30 * modsq(w) * modsq(w);
                C tC32 = (T)0.25 * (C32 * (cpow(wc, 3) * w));
                C tC34 = (T)0.25 * (C34 * cpow(wc, 4));
                C tC41 = (T)0.2 * (C41 * (cpow(wc, 3) * cpow(w, 2)));
                C tC43 = (T)0.2 * (C43 * (cpow(wc, 4) * w));
                C tC45 = (T)0.2 * (C45 * cpow(wc, 5));
                T tC50 = C50 * modsq(w) * modsq(w) * modsq(w) / (T)6.0;
                C tC52 = (C52 * (cpow(wc, 4) * cpow(w, 2))) / (T)6.0;
                C tC54 = (C54 * (cpow(wc, 5) * w)) / (T)6.0;
                C tC56 = (C56 * cpow(wc, 6)) / (T)6.0;
                T cchi = tC10 + tC12.real() + tC21.real() + tC23.real() + tC30 + tC32.real() +
                         tC34.real() + tC41.real() + tC43.real() + tC45.real() + tC50 +
                         tC52.real() + tC54.real() + tC56.real();
                T chi = (T)2.0 * pi * cchi / wavelength;
                T edge_factor = 1;
                if (taper && std::fabs(k - cond_ap2) < ap_smooth_radius) {
                    // OpenCL smoothstep(e0, e1, x) = t*t*(3-2t), t = clamp((x-e0)/(e1-e0), 0, 1)
                    T e0 = cond_ap2 - ap_smooth_radius, e1 = cond_ap2 + ap_smooth_radius;
                    T t = std::min(std::max((k - e0) / (e1 - e0), (T)0), (T)1);
                    edge_factor = (T)1 - t * t * ((T)3 - (T)2 * t);
                }
                out[id] = C(std::cos(-posTerm - chi) * edge_factor,
                            std::sin(-posTerm - chi) * edge_factor);
            } else {
                out[id] = C(0, 0);
            }
        }
}

// doMultiSliceStep for one wave (microscope/simulationgeneral.cpp:744-764)
template <typename T, typename FFT>
void slice_step(const FFT& fft, size_t N, std::complex<T>* psi, std::complex<T>* recip,
                const std::complex<T>* t, const std::complex<T>* P) {
    for (size_t i = 0; i < N; ++i) {  // complex_multiply_d.cl:19-24
        T ax = t[i].real(), ay = t[i].imag(), bx = psi[i].real(), by = psi[i].imag();
        recip[i] = std::complex<T>(ax * bx - ay * by, ax * by + ay * bx);
    }
    fft.run(recip, false);
    for (size_t i = 0; i < N; ++i) {
        T ax = recip[i].real(), ay = recip[i].imag(), bx = P[i].real(), by = P[i].imag();
        recip[i] = std::complex<T>(ax * bx - ay * by, ax * by + ay * bx);
    }
    std::memcpy((void*)psi, (const void*)recip, N * sizeof(std::complex<T>));
    fft.run(psi, true);
}

// getStemPixel + doSumReduction (microscope/simulationstem.cpp:42-131) with fft_shift,
// complex_to_real(4), band_pass and sum_reduction (kernels/*_d.cl).
template <typename T>
double stem_pixel(int res, const std::complex<T>* recip, double inner, double outer, double xc,
                  double yc, double angle_scale, int sqrt2) {
    const int W = res, H = res;
    T innerPx = (T)(inner / angle_scale);
    T outerPx = (T)(outer / angle_scale);
    T xcPx = (T)(xc / angle_scale);
    T ycPx = (T)(yc / angle_scale);
    T centX = W / (T)2.0 + xcPx;
    T centY = H / (T)2.0 + ycPx;
    const size_t total = (size_t)W * H;
    std::vector<T> masked(total);
    for (int yid = 0; yid < H; ++yid)
        for (int xid = 0; xid < W; ++xid) {
            // shifted pixel (xid, yid) holds the unshifted pixel ((xid+W/2)%W, (yid+H/2)%H)
            int ux = (xid + W / 2) % W, uy = (yid + H / 2) % H;
            std::complex<T> v = recip[(size_t)uy * W + ux];
            T val = v.real() * v.real() + v.imag() * v.imag();
            T radius = std::sqrt((xid - centX) * (xid - centX) + (yid - centY) * (yid - centY));
            masked[(size_t)yid * W + xid] = (radius <= outerPx && radius >= innerPx) ? val : (T)0;
        }
    double sum = 0;
    for (size_t g0 = 0; g0 < total; g0 += 256) {
        T s = 0;
        for (size_t i = 0; i < 256; ++i) s += std::fabs(masked[g0 + i]);
        if (sqrt2) s = s * (T)M_SQRT2;
        sum += s;
    }
    return sum;
}

template <typename T, typename FFT = Fft2<T>>
int run_stem(const OrcGeometry* g, const double* trans, const double* Pd, const double* pos,
             int n_probes, const double* aberr, double cond_ap, double smooth, int taper,
             const double* det, int ndet, int sqrt2, double* out, int nthreads) {
    typedef std::complex<T> C;
    const int res = g->resolution;
    const size_t N = (size_t)res * res;
    const int ns = g->n_slices;
    std::vector<C> t((size_t)ns * N), P(N);
    for (size_t i = 0; i < (size_t)ns * N; ++i) t[i] = C((T)trans[2 * i], (T)trans[2 * i + 1]);
    for (size_t i = 0; i < N; ++i) P[i] = C((T)Pd[2 * i], (T)Pd[2 * i + 1]);
    FFT fft(res, res);
    parallel_for(n_probes, nthreads, [&](int p) {
        std::vector<C> psi(N), recip(N);
        // simulationstem.cpp:178-188 -> simulationcbed.cpp:54-59: position relative to padded start
        T px = (T)(pos[2 * p] - g->sim_x[0]);
        T py = (T)(pos[2 * p + 1] - g->sim_y[0]);
        probe_kspace<T>(g, px, py, aberr, (T)cond_ap, (T)smooth, taper, recip.data());
        std::memcpy((void*)psi.data(), (const void*)recip.data(), N * sizeof(C));
        fft.run(psi.data(), true);
        for (int s = 0; s < ns; ++s) slice_step<T, FFT>(fft, N, psi.data(), recip.data(), &t[(size_t)s * N], P.data());
        for (int d = 0; d < ndet; ++d)
            out[(size_t)p * ndet + d] = stem_pixel<T>(res, recip.data(), det[4 * d], det[4 * d + 1],
                                                      det[4 * d + 2], det[4 * d + 3],
                                                      g->inverse_scale_angle, sqrt2);
    });
    return 0;
}

}  // namespace

// =========================================================================================
// C ABI
// =========================================================================================
extern "C" {

double orc_wavelength(double kv) { return wavelength(kv); }
double orc_sigma(double kv) { return sigma(kv); }

int orc_geometry(const OrcScenario* s, OrcGeometry* g) {
    std::memset(g, 0, sizeof(*g));
    if (!s || s->resolution <= 0 || s->slice_thickness <= 0) return -1;
    g->resolution = s->resolution;
    g->slice_thickness = s->slice_thickness;
    g->limit_factor = s->limit_factor;
    g->beam_tilt_mrad = s->beam_tilt_mrad;
    g->beam_azimuth = s->beam_azimuth;
    g->full3d = s->full3d;
    g->full3d_integrals = s->full3d_integrals;
    g->stem_pixels[0] = s->stem_pixels[0];
    g->stem_pixels[1] = s->stem_pixels[1];

    // fullAreaBase (simulationmanager.cpp:470-483): STEM area, CBED point, CTEM area
    double lx[2], ly[2];
    corrected_limits(s->area, s->area_padding, lx, ly);
    // paddedFullLimitsX/Y = corrected + cell xy padding (simulationmanager.cpp:446-452)
    g->full_x[0] = lx[0] + s->pad_xy[0];
    g->full_x[1] = lx[1] + s->pad_xy[1];
    g->full_y[0] = ly[0] + s->pad_xy[0];
    g->full_y[1] = ly[1] + s->pad_xy[1];
    // parallel STEM / CTEM / CBED: currentAreaBase == fullAreaBase (simulationmanager.cpp:485-505)
    for (int i = 0; i < 2; ++i) {
        g->sim_x[i] = g->full_x[i];
        g->sim_y[i] = g->full_y[i];
    }
    if (s->mode == 2 && s->moving_stem_frame) {
        // moving frame: realScale and every scale after it follow paddedSimLimitsX/Y(0)
        if (s->stem_pixels[0] <= 0 || s->stem_pixels[1] <= 0) return -1;
        double fx[2], fy[2];
        moving_frame_area(s, 0, fx, fy);
        g->sim_x[0] = fx[0] + s->pad_xy[0];
        g->sim_x[1] = fx[1] + s->pad_xy[1];
        g->sim_y[0] = fy[0] + s->pad_xy[0];
        g->sim_y[1] = fy[1] + s->pad_xy[1];
        g->moving_frame = 1;
    }
    // SimulationCell::roundPaddingZ (structure/simulationcell.cpp:43-61)
    {
        const double slice_thickness = s->slice_thickness, slice_offset = s->slice_offset;
        int pad_slices_pre = (int)std::ceil((std::fabs(s->pad_z[1]) - slice_offset) / slice_thickness);
        double pre_pad = slice_offset + pad_slices_pre * slice_thickness;
        double zw = s->struct_max[2] - s->struct_min[2];
        int struct_slices = (int)std::ceil((zw + slice_offset) / slice_thickness);
        double struct_slice_thick = struct_slices * slice_thickness;
        double left_over = struct_slice_thick - slice_offset - zw;
        int pad_slices_post = (int)std::ceil((std::fabs(s->pad_z[0]) - left_over) / slice_thickness);
        double post_pad = pad_slices_post * slice_thickness + left_over;
        g->pad_z[0] = -post_pad;
        g->pad_z[1] = pre_pad;
        g->z_lims[0] = s->struct_min[2] + g->pad_z[0];
        g->z_lims[1] = s->struct_max[2] + g->pad_z[1];
    }
    // SimulationCell::sliceCount / preSliceCount (structure/simulationcell.cpp:75-87)
    {
        double z_range = g->z_lims[1] - g->z_lims[0];
        unsigned int n_slices = (unsigned int)std::ceil((z_range / s->slice_thickness) - 0.000001);
        n_slices += (n_slices == 0);
        g->n_slices = (int)n_slices;
        g->pre_slices = (int)(unsigned int)(g->pad_z[1] / s->slice_thickness);
    }
    // realScale (simulationmanager.cpp:211-222)
    {
        double x_r = g->sim_x[1] - g->sim_x[0];
        double y_r = g->sim_y[1] - g->sim_y[0];
        g->real_scale = std::max(x_r, y_r) / s->resolution;
    }
    // calculateBlocks / blockScale (simulationmanager.cpp:371-408)
    {
        double xr = g->full_x[1] - g->full_x[0];
        double yr = g->full_y[1] - g->full_y[0];
        int n_blocks = (int)std::ceil(std::max(xr, yr) / 8.0);
        g->blocks_x = n_blocks;
        g->blocks_y = n_blocks;
        g->block_scale_x = xr / g->blocks_x;
        g->block_scale_y = yr / g->blocks_y;
        // simulationgeneral.cpp:517-521
        g->load_x = (int)std::ceil(8.0 / g->block_scale_x);
        g->load_y = (int)std::ceil(8.0 / g->block_scale_y);
        g->load_z = (int)std::ceil(3.0 / s->slice_thickness);
    }
    g->wavelength = wavelength(s->voltage_kv);
    g->wavenumber = 1.0 / g->wavelength;
    g->sigma = sigma(s->voltage_kv);
    {
        double k = g->wavenumber;
        g->k_vec[0] = k * std::sin(s->beam_tilt_mrad / 1000.0) * std::cos(s->beam_azimuth);
        g->k_vec[1] = k * std::sin(s->beam_tilt_mrad / 1000.0) * std::sin(s->beam_azimuth);
        g->k_vec[2] = k * std::cos(s->beam_tilt_mrad / 1000.0);
    }
    // simulationgeneral.cpp:418
    g->sigma_eff = g->sigma * g->wavenumber / g->k_vec[2];
    // simulationmanager.cpp:224-247
    g->inverse_scale = 1.0 / (g->real_scale * s->resolution);
    g->inverse_scale_angle = 1000.0 * g->inverse_scale * g->wavelength;
    {
        std::vector<double> kx, ky;
        kgrid<double>(g, kx, ky, &g->kmax_band);
    }
    // StemArea::getScaleX/Y (utilities/commonstructs.cpp:99-107)
    if (s->stem_pixels[0] > 0 && s->stem_pixels[1] > 0) {
        g->stem_scale[0] = (s->area[1] - s->area[0]) / s->stem_pixels[0];
        g->stem_scale[1] = (s->area[3] - s->area[2]) / s->stem_pixels[1];
    }
    g->stem_start[0] = s->area[0];
    g->stem_start[1] = s->area[2];
    return 0;
}

int orc_geometry_pixel(const OrcScenario* s, int pixel, OrcGeometry* g) {
    if (orc_geometry(s, g) != 0) return -1;
    if (!g->moving_frame) return 0;
    if (pixel < 0 || pixel >= s->stem_pixels[0] * s->stem_pixels[1]) return -1;
    // paddedSimLimitsX/Y(pixel) = currentAreaBase(pixel) corrected + cell padding
    // (simulationmanager.cpp:434-440); the scales stay those of pixel 0 (realScale)
    double fx[2], fy[2];
    moving_frame_area(s, pixel, fx, fy);
    g->sim_x[0] = fx[0] + s->pad_xy[0];
    g->sim_x[1] = fx[1] + s->pad_xy[1];
    g->sim_y[0] = fy[0] + s->pad_xy[0];
    g->sim_y[1] = fy[1] + s->pad_xy[1];
    return 0;
}

int orc_sort_atoms(const OrcGeometry* g, const double* x, const double* y, const double* z,
                   const int* Z, int n, double* sx, double* sy, double* sz, int* sZ,
                   int* block_ids, int* z_ids, int* block_start, int* n_filtered,
                   int* n_in_range) {
    // host filter with strict limits (microscope/simulationgeneral.cpp:188-217)
    std::vector<double> fx, fy, fz;
    std::vector<int> fa;
    for (int i = 0; i < n; ++i) {
        bool in_x = x[i] > g->full_x[0] && x[i] < g->full_x[1];
        bool in_y = y[i] > g->full_y[0] && y[i] < g->full_y[1];
        bool in_z = z[i] > g->z_lims[0] && z[i] < g->z_lims[1];
        if (in_x && in_y && in_z) {
            fa.push_back(Z[i]);
            fx.push_back(x[i]);
            fy.push_back(y[i]);
            fz.push_back(z[i]);
        }
    }
    const int count = (int)fa.size();
    const int blocks_x = g->blocks_x, blocks_y = g->blocks_y, n_slices = g->n_slices;
    const double min_x = g->full_x[0], max_x = g->full_x[1];
    const double min_y = g->full_y[0], max_y = g->full_y[1];
    const double max_z = g->z_lims[1];
    const double dz = g->slice_thickness;
    // atom_sort_d (kernels/atom_sort_d.cl:28-69), native_recip -> (1.0 / (x))
    for (int xid = 0; xid < count; ++xid) {
        if (fx[xid] >= min_x && fx[xid] <= max_x && fy[xid] >= min_y && fy[xid] <= max_y) {
            int bidx = (int)std::floor((fx[xid] - min_x) * (1.0 / (max_x - min_x)) * blocks_x);
            int bidy = (int)std::floor((fy[xid] - min_y) * (1.0 / (max_y - min_y)) * blocks_y);
            int zid = (int)std::floor((max_z - fz[xid]) * (1.0 / (dz)));
            zid -= (zid == n_slices);
            bidx -= (bidx == blocks_x);
            bidy -= (bidy == blocks_y);
            block_ids[xid] = bidx + blocks_x * bidy;
            z_ids[xid] = zid;
        } else {
            block_ids[xid] = -1;
            z_ids[xid] = -1;
        }
    }
    // host bucket fill, stable (microscope/simulationgeneral.cpp:271-328)
    const int nblk = blocks_x * blocks_y;
    std::vector<std::vector<int>> bins((size_t)nblk * n_slices);
    int count_in_range = 0;
    for (int i = 0; i < count; ++i) {
        if (z_ids[i] >= 0 && block_ids[i] >= 0) {
            bins[(size_t)block_ids[i] * n_slices + z_ids[i]].push_back(i);
            ++count_in_range;
        }
    }
    for (int i = 0; i < count; ++i) {  // unsorted tail keeps the filtered values
        sx[i] = fx[i];
        sy[i] = fy[i];
        sz[i] = fz[i];
        sZ[i] = fa[i];
    }
    int it = 0;
    for (int s = 0; s < n_slices; ++s)
        for (int j = 0; j < blocks_y; ++j)
            for (int k = 0; k < blocks_x; ++k) {
                block_start[s * nblk + j * blocks_x + k] = it;
                for (int idx : bins[(size_t)(j * blocks_x + k) * n_slices + s]) {
                    sx[it] = fx[idx];
                    sy[it] = fy[idx];
                    sz[it] = fz[idx];
                    sZ[it] = fa[idx];
                    ++it;
                }
            }
    block_start[n_slices * nblk] = count_in_range;
    *n_filtered = count;
    *n_in_range = count_in_range;
    return 0;
}

double orc_bessk0(double x) { return bessk0<double>(x); }
double orc_bessk1(double x) { return bessk1<double>(x); }
double orc_vproj(int form, const double* params, int i_count, int Z, double r) {
    return vproj(form, params, i_count, Z, r);
}
double orc_v3d(int form, const double* params, int i_count, int Z, double r) {
    return v3d(form, params, i_count, Z, r);
}

int orc_kgrid(const OrcGeometry* g, double* kx, double* ky) {
    std::vector<double> a, b;
    double km;
    kgrid<double>(g, a, b, &km);
    std::copy(a.begin(), a.end(), kx);
    std::copy(b.begin(), b.end(), ky);
    return 0;
}

int orc_potential_slice(const OrcGeometry* g, const double* sx, const double* sy, const double* sz,
                        const int* sZ, const int* block_start, const double* params, int form,
                        int i_count, int slice, double* out, int nthreads) {
    const int res = g->resolution;
    if (res % 16) return -1;
    const int ng = res / 16;
    parallel_for(ng * ng, nthreads, [&](int gidx) {
        int gx = gidx % ng, gy = gidx / ng;
        BlockRange br = block_range(g, gx, gy);
        for (int ly = 0; ly < 16; ++ly)
            for (int lx = 0; lx < 16; ++lx) {
                int xid = gx * 16 + lx, yid = gy * 16 + ly;
                double v = g->full3d ? pixel_full3d(g, sx, sy, sz, sZ, block_start, params, form,
                                                    i_count, slice, xid, yid, br)
                                     : pixel_projected(g, sx, sy, sZ, block_start, params, form,
                                                       i_count, slice, xid, yid, br);
                out[(size_t)yid * res + xid] = v;
            }
    });
    return 0;
}

int orc_transmission(const OrcGeometry* g, const double* sx, const double* sy, const double* sz,
                     const int* sZ, const int* block_start, const double* params, int form,
                     int i_count, int first, int count, int band_limit, double* out, int nthreads) {
    const int res = g->resolution;
    const size_t N = (size_t)res * res;
    std::vector<double> V(N);
    Fft2<double> fft(res, res);
    for (int s = first; s < first + count; ++s) {
        orc_potential_slice(g, sx, sy, sz, sZ, block_start, params, form, i_count, s, V.data(),
                            nthreads);
        std::complex<double>* t = reinterpret_cast<std::complex<double>*>(out + (size_t)(s - first) * 2 * N);
        const double sig = g->sigma_eff;
        const double sst = g->slice_thickness * (1.0 / g->full3d_integrals);
        for (size_t i = 0; i < N; ++i) {
            double ph = g->full3d ? sst * sig * V[i] : sig * V[i];
            t[i] = std::complex<double>(std::cos(ph), std::sin(ph));
        }
        if (band_limit) {  // simulationgeneral.cpp:596-602: FFT, band_limit, IFFT
            fft.run(t, false);
            band_limit_mask<double>(g, t);
            fft.run(t, true);
        }
    }
    return 0;
}

int orc_fft2(double* data, int w, int h, int inverse) {
    Fft2<double> fft(w, h);
    fft.run(reinterpret_cast<std::complex<double>*>(data), inverse != 0);
    return 0;
}

int orc_band_limit(const OrcGeometry* g, double* t) {
    band_limit_mask<double>(g, reinterpret_cast<std::complex<double>*>(t));
    return 0;
}

int orc_propagator(const OrcGeometry* g, double* P) {
    propagator<double>(g, reinterpret_cast<std::complex<double>*>(P));
    return 0;
}

int orc_probe_kspace(const OrcGeometry* g, double px, double py, const double* aberr,
                     double cond_ap_mrad, double smooth_mrad, int taper, double* out) {
    probe_kspace<double>(g, px, py, aberr, cond_ap_mrad, smooth_mrad, taper,
                         reinterpret_cast<std::complex<double>*>(out));
    return 0;
}

int orc_slice_step(int res, double* psi, double* recip, const double* t, const double* P) {
    Fft2<double> fft(res, res);
    typedef std::complex<double> C;
    slice_step<double>(fft, (size_t)res * res, reinterpret_cast<C*>(psi), reinterpret_cast<C*>(recip),
                       reinterpret_cast<const C*>(t), reinterpret_cast<const C*>(P));
    return 0;
}

double orc_stem_pixel(int res, const double* recip, double inner_mrad, double outer_mrad,
                      double xc_mrad, double yc_mrad, double angle_scale, int sqrt2) {
    return stem_pixel<double>(res, reinterpret_cast<const std::complex<double>*>(recip), inner_mrad,
                              outer_mrad, xc_mrad, yc_mrad, angle_scale, sqrt2);
}

// getDiffractionImage (simulationgeneral.cpp:772-805): |fftshift(recip)|^2
int orc_diffraction(int res, const double* recip, double* out) {
    for (int yid = 0; yid < res; ++yid)
        for (int xid = 0; xid < res; ++xid) {
            int ux = (xid + res / 2) % res, uy = (yid + res / 2) % res;
            size_t u = (size_t)uy * res + ux;
            out[(size_t)yid * res + xid] = recip[2 * u] * recip[2 * u] + recip[2 * u + 1] * recip[2 * u + 1];
        }
    return 0;
}

// translateDiffImage (simulationgeneral.cpp:832-854): shift = d_k / inverseScale split into a
// floor and a fraction; bilinear_translate_d (kernels/bilinear_translate_d.cl:21-72): the pixel
// taken from (x - ix, y - iy) wrapped once (further: 0), blended with its left, upper and
// upper-left neighbours with the factors of the opposite quadrants.
int orc_translate_image(int res, const double* in, double* out, double dkx, double dky, double inverse_scale) {
    const double shift_x = dkx / inverse_scale, shift_y = dky / inverse_scale;
    const int ix = (int)std::floor(shift_x), iy = (int)std::floor(shift_y);
    const double sx = shift_x - ix, sy = shift_y - iy;
    const int W = res, H = res;
    for (int yid = 0; yid < H; ++yid)
        for (int xid = 0; xid < W; ++xid) {
            const int id = xid + W * yid;
            int nx = xid - ix, ny = yid - iy;
            if (nx < 0)
                nx = W + nx;
            else if (nx >= W)
                nx = nx - W;
            if (ny < 0)
                ny = H + ny;
            else if (ny >= H)
                ny = ny - H;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) {
                out[id] = 0.0;
                continue;
            }
            const int nid = nx + W * ny;
            int br = nid, bl = nid - 1, tr = nid - W, tl = nid - W - 1;
            if (nx == 0) {
                bl += W;
                tl += W;
            }
            if (ny == 0) {
                tr += W * H;
                tl += W * H;
            }
            const double f_br = sx * sy, f_bl = (1.0 - sx) * sy, f_tr = sx * (1.0 - sy),
                         f_tl = (1.0 - sx) * (1.0 - sy);
            out[id] = f_tl * in[br] + f_tr * in[bl] + f_bl * in[tr] + f_br * in[tl];
        }
    return 0;
}

// generateRotationAroundVector (utilities/vectorutils.cpp:8-18): I + sin(a) C + (1 - cos(a)) C^2
// with C the cross-product matrix of the normalised axis
static void rotation_about(const double* ax_in, double angle, double R[3][3]) {
    const double n = std::sqrt(ax_in[0] * ax_in[0] + ax_in[1] * ax_in[1] + ax_in[2] * ax_in[2]);
    const double a0 = ax_in[0] / n, a1 = ax_in[1] / n, a2 = ax_in[2] / n;
    const double C[3][3] = {{0.0, -a2, a1}, {a2, 0.0, -a0}, {-a1, a0, 0.0}};
    const double s = std::sin(angle), c1 = 1.0 - std::cos(angle);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double cc = 0.0;
            for (int k = 0; k < 3; ++k) cc += C[i][k] * C[k][j];
            R[i][j] = (i == j ? 1.0 : 0.0) + C[i][j] * s + cc * c1;
        }
}

static void apply(const double R[3][3], double* v) {
    const double r0 = R[0][0] * v[0] + R[0][1] * v[1] + R[0][2] * v[2];
    const double r1 = R[1][0] * v[0] + R[1][1] * v[1] + R[1][2] * v[2];
    const double r2 = R[2][0] * v[0] + R[2][1] * v[1] + R[2][2] * v[2];
    v[0] = r0;
    v[1] = r1;
    v[2] = r2;
}

void orc_rotate_vector_spherical(double* z, double* y, double theta, double phi) {
    double r_theta[3][3], r_phi[3][3];
    rotation_about(y, theta, r_theta);
    rotation_about(z, phi, r_phi);
    apply(r_theta, z);
    apply(r_phi, z);
    apply(r_phi, y);
}

// CTEM: init_plane_wave (value 1) then the slice loop (simulationctem.cpp:77-82, 340-341)
int orc_run_plane(int res, int n_slices, const double* trans, const double* P, double* psi_out,
                  double* recip_out) {
    typedef std::complex<double> C;
    const size_t N = (size_t)res * res;
    Fft2<double> fft(res, res);
    std::vector<C> psi(N, C(1.0, 0.0)), recip(N, C(0.0, 0.0));
    for (int s = 0; s < n_slices; ++s)
        slice_step<double>(fft, N, psi.data(), recip.data(),
                           reinterpret_cast<const C*>(trans + (size_t)s * 2 * N),
                           reinterpret_cast<const C*>(P));
    std::memcpy(psi_out, psi.data(), N * sizeof(C));
    if (recip_out) std::memcpy(recip_out, recip.data(), N * sizeof(C));
    return 0;
}

int orc_run_stem(const OrcGeometry* g, const double* trans, const double* P, const double* pos,
                 int n_probes, const double* aberr, double cond_ap_mrad, double smooth_mrad,
                 int taper, const double* det, int ndet, int sqrt2, double* out, int nthreads) {
    return run_stem<double>(g, trans, P, pos, n_probes, aberr, cond_ap_mrad, smooth_mrad, taper,
                            det, ndet, sqrt2, out, nthreads);
}

int orc_run_stem_f32(const OrcGeometry* g, const double* trans, const double* P, const double* pos,
                     int n_probes, const double* aberr, double cond_ap_mrad, double smooth_mrad,
                     int taper, const double* det, int ndet, int sqrt2, double* out, int nthreads) {
    if (StockhamFft1<float>::fits(g->resolution))
        return run_stem<float, FastFft2<float>>(g, trans, P, pos, n_probes, aberr, cond_ap_mrad, smooth_mrad,
                                                taper, det, ndet, sqrt2, out, nthreads);
    return run_stem<float>(g, trans, P, pos, n_probes, aberr, cond_ap_mrad, smooth_mrad, taper,
                           det, ndet, sqrt2, out, nthreads);
}

int orc_fft2_fast(double* data, int w, int h, int inverse) {
    if (!StockhamFft1<double>::fits(w) || !StockhamFft1<double>::fits(h)) return -1;
    FastFft2<double> fft(w, h);
    fft.run(reinterpret_cast<std::complex<double>*>(data), inverse != 0);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Loss of a CTEM plane-wave exit intensity and its central finite differences with respect
// to atom coordinates: the checker of the adjoint slice loop.  The forward is the path
// above (sortAtoms, transmission with band limit, doMultiSliceStep over every slice from
// init_plane_wave value 1).  A perturbed coordinate only changes the transmission of the
// slices whose potential sums see the atom (its z bin, +-load_z in full 3D), so each
// evaluation rebuilds those slices and reruns the slice loop.
// ---------------------------------------------------------------------------------------
namespace {

struct CtemProblem {
    const OrcGeometry* g;
    const double* params;
    int form, i_count;
    size_t N;
    std::vector<std::complex<double>> P;
};

struct SortedAtoms {
    std::vector<double> x, y, z;
    std::vector<int> Z, bid, zid, bs;
};

SortedAtoms sort_all(const OrcGeometry* g, const double* x, const double* y, const double* z, const int* Z, int n) {
    SortedAtoms s;
    const size_t na = std::max(n, 1);
    s.x.resize(na); s.y.resize(na); s.z.resize(na); s.Z.resize(na); s.bid.resize(na); s.zid.resize(na);
    s.bs.resize((size_t)g->n_slices * g->blocks_x * g->blocks_y + 1);
    int nf = 0, nr = 0;
    orc_sort_atoms(g, x, y, z, Z, n, s.x.data(), s.y.data(), s.z.data(), s.Z.data(), s.bid.data(),
                   s.zid.data(), s.bs.data(), &nf, &nr);
    return s;
}

void slice_transmission(const CtemProblem& pb, const SortedAtoms& s, int slice, std::complex<double>* t) {
    orc_transmission(pb.g, s.x.data(), s.y.data(), s.z.data(), s.Z.data(), s.bs.data(), pb.params, pb.form,
                     pb.i_count, slice, 1, 1, reinterpret_cast<double*>(t), 1);
}

// exit intensity through the slice stack; slice s uses over[s] when present
void exit_intensity(const CtemProblem& pb, const std::vector<const std::complex<double>*>& stack,
                    std::vector<double>& I) {
    typedef std::complex<double> C;
    const int res = pb.g->resolution;
    Fft2<double> fft(res, res);
    std::vector<C> psi(pb.N, C(1.0, 0.0)), recip(pb.N);
    for (size_t s = 0; s < stack.size(); ++s) slice_step<double>(fft, pb.N, psi.data(), recip.data(), stack[s], pb.P.data());
    I.resize(pb.N);
    for (size_t i = 0; i < pb.N; ++i) I[i] = std::norm(psi[i]);
}

double loss_of(const std::vector<double>& I, const double* target) {
    double L = 0.0;
    for (size_t i = 0; i < I.size(); ++i) {
        const double d = I[i] - target[i];
        L += d * d;
    }
    return L;
}

int z_bin(const OrcGeometry* g, double x, double y, double z) {
    if (!(x > g->full_x[0] && x < g->full_x[1] && y > g->full_y[0] && y < g->full_y[1] && z > g->z_lims[0] &&
          z < g->z_lims[1]))
        return -1;
    int zid = (int)std::floor((g->z_lims[1] - z) * (1.0 / (g->slice_thickness)));
    zid -= (zid == g->n_slices);
    return zid;
}

}  // namespace

extern "C" {

int orc_ctem_intensity(const OrcGeometry* g, const double* x, const double* y, const double* z, const int* Z,
                       int n, const double* params, int form, int i_count, double* intensity, int nthreads) {
    typedef std::complex<double> C;
    CtemProblem pb{g, params, form, i_count, (size_t)g->resolution * g->resolution, {}};
    pb.P.resize(pb.N);
    propagator<double>(g, pb.P.data());
    SortedAtoms s = sort_all(g, x, y, z, Z, n);
    std::vector<C> T((size_t)g->n_slices * pb.N);
    orc_transmission(g, s.x.data(), s.y.data(), s.z.data(), s.Z.data(), s.bs.data(), params, form, i_count, 0,
                     g->n_slices, 1, reinterpret_cast<double*>(T.data()), nthreads);
    std::vector<const C*> stack(g->n_slices);
    for (int k = 0; k < g->n_slices; ++k) stack[k] = T.data() + (size_t)k * pb.N;
    std::vector<double> I;
    exit_intensity(pb, stack, I);
    std::memcpy(intensity, I.data(), pb.N * sizeof(double));
    return 0;
}

int orc_ctem_loss_fd(const OrcGeometry* g, const double* x, const double* y, const double* z, const int* Z, int n,
                     const double* params, int form, int i_count, const double* target, const int* atom,
                     const int* axis, int m, double h, double* grad, double* loss0, int nthreads) {
    typedef std::complex<double> C;
    for (int i = 0; i < m; ++i)
        if (atom[i] < 0 || atom[i] >= n || axis[i] < 0 || axis[i] > 2) return -1;
    CtemProblem pb{g, params, form, i_count, (size_t)g->resolution * g->resolution, {}};
    pb.P.resize(pb.N);
    propagator<double>(g, pb.P.data());
    const SortedAtoms s0 = sort_all(g, x, y, z, Z, n);
    std::vector<C> T((size_t)g->n_slices * pb.N);
    orc_transmission(g, s0.x.data(), s0.y.data(), s0.z.data(), s0.Z.data(), s0.bs.data(), params, form, i_count,
                     0, g->n_slices, 1, reinterpret_cast<double*>(T.data()), nthreads);
    std::vector<const C*> base(g->n_slices);
    for (int k = 0; k < g->n_slices; ++k) base[k] = T.data() + (size_t)k * pb.N;
    if (loss0) {
        std::vector<double> I;
        exit_intensity(pb, base, I);
        *loss0 = loss_of(I, target);
    }
    const int lz = g->full3d ? g->load_z : 0;
    std::vector<double> Lv(2 * (size_t)m, 0.0);
    parallel_for(2 * m, nthreads, [&](int e) {
        const int i = e / 2;
        const double sgn = (e % 2) ? -1.0 : 1.0;
        std::vector<double> px(x, x + n), py(y, y + n), pz(z, z + n);
        double* c = axis[i] == 0 ? px.data() : axis[i] == 1 ? py.data() : pz.data();
        const int a = atom[i];
        const int k_old = z_bin(g, x[a], y[a], z[a]);
        c[a] += sgn * h;
        const int k_new = z_bin(g, px[a], py[a], pz[a]);
        const SortedAtoms sp = sort_all(g, px.data(), py.data(), pz.data(), Z, n);
        std::vector<int> redo;
        for (int k0 : {k_old, k_new}) {
            if (k0 < 0) continue;
            for (int k = std::max(0, k0 - lz); k <= std::min(g->n_slices - 1, k0 + lz); ++k)
                if (std::find(redo.begin(), redo.end(), k) == redo.end()) redo.push_back(k);
        }
        std::vector<C> mod(redo.size() * pb.N);
        std::vector<const C*> stack = base;
        for (size_t r = 0; r < redo.size(); ++r) {
            slice_transmission(pb, sp, redo[r], mod.data() + r * pb.N);
            stack[redo[r]] = mod.data() + r * pb.N;
        }
        std::vector<double> I;
        exit_intensity(pb, stack, I);
        Lv[e] = loss_of(I, target);
    });
    for (int i = 0; i < m; ++i) grad[i] = (Lv[2 * i] - Lv[2 * i + 1]) / (2.0 * h);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// CTEM image formation: ctem_image_d (kernels/ctem_image_d.cl) on Psi(k), the inverse FFT
// and sqabs_d (SimulationCtem::simulateImagePerfect, microscope/simulationctem.cpp:113-171).
// lens = 25 aberration doubles in the probe order, then obj_ap_mrad, ap_smooth_mrad,
// beta_mrad, delta.  recip is the unshifted Psi(k) (2N doubles); out has N doubles.
// ---------------------------------------------------------------------------------------
extern "C" int orc_ctem_image(const OrcGeometry* g, const double* recip, const double* lens, double* out) {
    typedef std::complex<double> C;
    const int res = g->resolution;
    const size_t N = (size_t)res * res;
    std::vector<double> kx, ky;
    double bw;
    kgrid<double>(g, kx, ky, &bw);
    const double wavelength = g->wavelength;
    auto cpow = [](C a, int n) {
        C temp = a;
        for (int j = 1; j < n; j++) temp = temp * a;
        return temp;
    };
    auto modsq = [](C a) { return a.real() * a.real() + a.imag() * a.imag(); };
    const double* ab = lens;
    const double C10 = ab[0];
    const C C12(ab[1], ab[2]), C21(ab[3], ab[4]), C23(ab[5], ab[6]);
    const double C30 = ab[7];
    const C C32(ab[8], ab[9]), C34(ab[10], ab[11]);
    const C C41(ab[12], ab[13]), C43(ab[14], ab[15]), C45(ab[16], ab[17]);
    const double C50 = ab[18];
    const C C52(ab[19], ab[20]), C54(ab[21], ab[22]), C56(ab[23], ab[24]);
    const double obj_ap = lens[25], ap_smooth = lens[26], beta = lens[27], delta = lens[28];
    const C* in = reinterpret_cast<const C*>(recip);
    std::vector<C> img(N);
    for (int yid = 0; yid < res; ++yid)
        for (int xid = 0; xid < res; ++xid) {
            const size_t id = (size_t)xid + (size_t)yid * res;
            const double obj_ap2 = (obj_ap * 0.001) / wavelength;
            const double beta2 = (beta * 0.001) / wavelength;
            const double k = std::sqrt((kx[xid] * kx[xid]) + (ky[yid] * ky[yid]));
            const double ap_smooth_radius = (ap_smooth * 0.001) / wavelength;
            if (k < obj_ap2 + ap_smooth_radius) {
                const C w(wavelength * kx[xid], wavelength * ky[yid]);
                const C wc = std::conj(w);
                const double temporalCoh =
                    std::exp(-0.25 * M_PI * M_PI * delta * delta * modsq(w) * modsq(w) / (wavelength * wavelength));
                const double spatialCoh = std::exp(-1.0 * M_PI * M_PI * beta2 * beta2 * modsq(w) *
                                                   std::pow((C10 + C30 * modsq(w) + C50 * modsq(w) * modsq(w)), 2));
                const double tC10 = 0.5 * C10 * modsq(w);
                const C tC12 = 0.5 * (C12 * cpow(wc, 2));
                const C tC21 = (C21 * (cpow(wc, 2) * w)) / 3.0;
                const C tC23 = (C23 * cpow(wc, 3)) / 3.0;
                const double tC30 = 0.25 * C30 * modsq(w) * modsq(w);
                const C tC32 = 0.25 * (C32 * (cpow(wc, 3) * w));
                const C tC34 = 0.25 * (C34 * cpow(wc, 4));
                const C tC41 = 0.2 * (C41 * (cpow(wc, 3) * cpow(w, 2)));
                const C tC43 = 0.2 * (C43 * (cpow(wc, 4) * w));
                const C tC45 = 0.2 * (C45 * cpow(wc, 5));
                const double tC50 = C50 * modsq(w) * modsq(w) * modsq(w) / 6.0;
                const C tC52 = (C52 * (cpow(wc, 4) * cpow(w, 2))) / 6.0;
                const C tC54 = (C54 * (cpow(wc, 5) * w)) / 6.0;
                const C tC56 = (C56 * cpow(wc, 6)) / 6.0;
                const double cchi = tC10 + tC12.real() + tC21.real() + tC23.real() + tC30 + tC32.real() +
                                    tC34.real() + tC41.real() + tC43.real() + tC45.real() + tC50 + tC52.real() +
                                    tC54.real() + tC56.real();
                const double chi = 2.0 * M_PI * cchi / wavelength;
                double edge_factor = 1.0;
                if (std::fabs(k - obj_ap2) < ap_smooth_radius) {
                    const double e0 = obj_ap2 - ap_smooth_radius, e1 = obj_ap2 + ap_smooth_radius;
                    const double t = std::min(std::max((k - e0) / (e1 - e0), 0.0), 1.0);
                    edge_factor = 1.0 - t * t * (3.0 - 2.0 * t);
                }
                const double f = edge_factor * temporalCoh * spatialCoh;
                img[id] = C(f * (in[id].real() * std::cos(chi) + in[id].imag() * std::sin(chi)),
                            f * (in[id].imag() * std::cos(chi) - in[id].real() * std::sin(chi)));
            } else {
                img[id] = C(0.0, 0.0);
            }
        }
    Fft2<double> fft(res, res);
    fft.run(img.data(), true);
    for (size_t i = 0; i < N; ++i) out[i] = std::norm(img[i]);  // sqabs_d
    return 0;
}
