// Host construction of the tabulated potentials (see potential_table.hpp).  The potentials
// are the reference's `_d` kernel formulas evaluated in double; K0 and K1 are the Numerical
// Recipes in C (2nd ed., section 6.6) rational approximations the reference cites, with
// the same published coefficients the device kernels carry.
#include "potential_table.hpp"

#include <cfloat>
#include <cmath>

namespace mse_tab {
namespace {

double horner(const double* c, int n, double x) {
    double a = c[n];
    for (int i = n - 1; i >= 0; --i) a = a * x + c[i];
    return a;
}

// Numerical Recipes section 6.6 coefficients (small argument: numerator of the regular part,
// its denominator in 1 - x^2, the log part's numerator and denominator; large argument:
// numerator and denominator in 1/x).
const double kK0Small[5] = {1.159315156584126e-1, 2.770731240515333e-1, 2.066458134619875e-2,
                            4.574734709978264e-4, 3.454715527986737e-6};
const double kK0SmallQ[3] = {9.836249671709183e-1, 1.627693622304549e-2, 9.809660603621949e-5};
const double kK0Log[5] = {1.0, 2.346487949187396e-1, 1.187082088663404e-2, 2.150707366040937e-4,
                          1.425433617130587e-6};
const double kK0LogQ[3] = {9.847324170755358e-1, 1.518396076767770e-2, 8.362215678646257e-5};
const double kK0Large[8] = {1.253314137315499, 1.475731032429900e1, 6.123767403223466e1,
                            1.121012633939949e2, 9.285288485892228e1, 3.198289277679660e1,
                            3.595376024148513, 6.160228690102976e-2};
const double kK0LargeQ[8] = {1.0, 1.189963006673403e1, 5.027773590829784e1, 9.496513373427093e1,
                             8.318077493230258e1, 3.181399777449301e1, 4.443672926432041,
                             1.408295601966600e-1};
const double kK1Small[5] = {-3.079657578292062e-1, -8.109417631822442e-2, -3.477550948593604e-3,
                            -5.385594871975406e-5, -3.110372465429008e-7};
const double kK1SmallQ[3] = {9.861813171751389e-1, 1.375094061153160e-2, 6.774221332947002e-5};
const double kK1Log[5] = {0.5, 5.598072040178741e-2, 1.818666382168295e-3, 2.397509908859959e-5,
                          1.239567816344855e-7};
const double kK1LogQ[3] = {9.870202601341150e-1, 1.292092053534579e-2, 5.881933053917096e-5};
const double kK1Large[8] = {1.253314137315502, 1.457171340220454e1, 6.063161173098803e1,
                            1.147386690867892e2, 1.040442011439181e2, 4.356596656837691e1,
                            7.265230396353690, 3.144418558991021e-1};
const double kK1LargeQ[8] = {1.0, 1.125154514806458e1, 4.427488496597630e1, 7.616113213117645e1,
                             5.863377227890893e1, 1.850303673841586e1, 1.857244676566022,
                             2.538540887654872e-2};

double k0(double x) {
    if (x > 0.0 && x <= 1.0) {
        const double z = x * x;
        return horner(kK0Small, 4, z) / horner(kK0SmallQ, 2, 1.0 - z) -
               horner(kK0Log, 4, z) * std::log(x) / horner(kK0LogQ, 2, 1.0 - z);
    }
    if (x > 1.0) {
        const double z = 1.0 / x;
        return std::exp(-x) * horner(kK0Large, 7, z) / horner(kK0LargeQ, 7, z) / std::sqrt(x);
    }
    return DBL_MAX;
}

double k1(double x) {
    if (x > 0.0 && x <= 1.0) {
        const double z = x * x;
        return x * (horner(kK1Small, 4, z) / horner(kK1SmallQ, 2, 1.0 - z) +
                    horner(kK1Log, 4, z) * std::log(x) / horner(kK1LogQ, 2, 1.0 - z)) +
               1.0 / x;
    }
    if (x > 1.0) {
        const double z = 1.0 / x;
        return std::exp(-x) * horner(kK1Large, 7, z) / horner(kK1LargeQ, 7, z) / std::sqrt(x);
    }
    return DBL_MAX;
}

}  // namespace

double vproj_ref(int form, const double* p, int ic, int Z, double r) {
    if (form == 0) {  // Kirkland C.20: 12 parameters per Z, (a, b) Lorentzians then (c, d) Gaussians
        const double* q = p + (Z - 1) * 12;
        double suml = 0.0, sumg = 0.0;
        for (int i = 0; i < ic; ++i) suml += q[2 * i] * k0(2.0 * M_PI * r * std::sqrt(q[2 * i + 1]));
        const double x = (M_PI * r) * (M_PI * r);
        for (int i = ic; i < 2 * ic; ++i) {
            const double d_inv = 1.0 / q[2 * i + 1];
            sumg += (q[2 * i] * d_inv) * std::exp(-x * d_inv);
        }
        return 300.8242834 * suml + 150.4121417 * sumg;
    }
    const double* q = p + (Z - 1) * 10;  // a_1..a_5, b_1..b_5
    double sum = 0.0;
    if (form == 1) {  // Peng
        const double x = (M_PI * r) * (M_PI * r);
        for (int i = 0; i < ic; ++i) {
            const double b_inv = 1.0 / q[i + 5];
            sum += q[i] * b_inv * std::exp(-x * b_inv);
        }
        return 150.4121417 * sum;
    }
    for (int i = 0; i < ic; ++i) {  // Lobato
        const double bir = 1.0 / std::sqrt(q[i + 5]);
        const double x = 2.0 * M_PI * r * bir;
        sum += q[i] * (bir * bir * bir) * (k0(x) + r * k1(x));
    }
    return 945.090144399935 * sum;
}

double v3d_ref(int form, const double* p, int ic, int Z, double r) {
    if (form == 0) {
        const double* q = p + (Z - 1) * 12;
        double suml = 0.0, sumg = 0.0;
        for (int i = 0; i < ic; ++i) suml += q[2 * i] * std::exp(-2.0 * M_PI * r * std::sqrt(q[2 * i + 1]));
        const double x = (M_PI * r) * (M_PI * r);
        for (int i = ic; i < 2 * ic; ++i) {
            const double dir = 1.0 / std::sqrt(q[2 * i + 1]);
            sumg += q[2 * i] * (dir * dir * dir) * std::exp(-x / q[2 * i + 1]);
        }
        return 150.4121417 / r * suml + 266.5157269 * sumg;
    }
    const double* q = p + (Z - 1) * 10;
    double sum = 0.0;
    if (form == 1) {
        const double x = (M_PI * r) * (M_PI * r);
        for (int i = 0; i < ic; ++i) {
            const double bir = 1.0 / std::sqrt(q[i + 5]);
            sum += q[i] * (bir * bir * bir) * std::exp(-x / q[i + 5]);
        }
        return 266.5157269 * sum;
    }
    const double x = M_PI * r;
    for (int i = 0; i < ic; ++i) {
        const double bir = 1.0 / std::sqrt(q[i + 5]);
        sum += q[i] * (bir * bir * bir) * std::exp(-2.0 * x * bir) * (std::sqrt(q[i + 5]) / x + 1.0);
    }
    return 472.545072199968 * sum;
}

void build(int form, bool full3d, const double* params, int i_count, int Z, float* out) {
    // cubic through the values at the Chebyshev nodes t_k of [0, 1]: c = M^-1 v with
    // M the Vandermonde matrix of the nodes, inverted once by Gauss-Jordan elimination
    static double minv[4][4];
    static double nodes[4];
    static bool ready = false;
    if (!ready) {
        double a[4][8];
        for (int k = 0; k < 4; ++k) {
            nodes[k] = 0.5 - 0.5 * std::cos((2 * k + 1) * M_PI / 8.0);
            for (int j = 0; j < 4; ++j) {
                a[k][j] = std::pow(nodes[k], j);
                a[k][4 + j] = (k == j) ? 1.0 : 0.0;
            }
        }
        for (int c = 0; c < 4; ++c) {
            int piv = c;
            for (int k = c + 1; k < 4; ++k)
                if (std::fabs(a[k][c]) > std::fabs(a[piv][c])) piv = k;
            for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[piv][j]);
            const double d = a[c][c];
            for (int j = 0; j < 8; ++j) a[c][j] /= d;
            for (int k = 0; k < 4; ++k) {
                if (k == c) continue;
                const double f = a[k][c];
                for (int j = 0; j < 8; ++j) a[k][j] -= f * a[c][j];
            }
        }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) minv[i][j] = a[i][4 + j];
        ready = true;
    }
    for (int b = 0; b < kBuckets; ++b) {
        const int e = kOctaveLo + (b >> kSubBits);
        const int m = b & ((1 << kSubBits) - 1);
        const double lo = std::ldexp(1.0 + (double)m / (1 << kSubBits), e);
        const double w = std::ldexp(1.0 / (1 << kSubBits), e);
        double v[4];
        for (int k = 0; k < 4; ++k) {
            const double r = std::sqrt(lo + w * nodes[k]);
            v[k] = full3d ? v3d_ref(form, params, i_count, Z, r) : vproj_ref(form, params, i_count, Z, r);
        }
        for (int i = 0; i < 4; ++i) {
            double c = 0.0;
            for (int j = 0; j < 4; ++j) c += minv[i][j] * v[j];
            out[4 * b + i] = (float)c;
        }
    }
}

}  // namespace mse_tab
