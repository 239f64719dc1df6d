// Host side of the engine in the shape of clTEM's simulation classes (see cltem_host.hpp).
#include "cltem_host.hpp"

#include <complex>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

namespace cltem_amd {

void check(int rc) {
    if (rc != MSE_OK) throw std::runtime_error(mse_last_error());
}

void AtomList::limits(double* lo, double* hi) const {
    // CrystalStructure::updateLimits (structure/crystalstructure.cpp:238-255)
    lo[0] = lo[1] = lo[2] = 1e300;
    hi[0] = hi[1] = hi[2] = -1e300;
    for (size_t i = 0; i < x.size(); ++i) {
        lo[0] = std::min(lo[0], x[i]);
        lo[1] = std::min(lo[1], y[i]);
        lo[2] = std::min(lo[2], z[i]);
        hi[0] = std::max(hi[0], x[i]);
        hi[1] = std::max(hi[1], y[i]);
        hi[2] = std::max(hi[2], z[i]);
    }
}

static const char* kSymbols[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl",
    "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se",
    "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb",
    "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er",
    "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At",
    "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No",
    "Lr"};

int element_number(const std::string& symbol) {
    for (int i = 0; i < 103; ++i)
        if (symbol == kSymbols[i]) return i + 1;
    throw std::runtime_error("unknown element symbol " + symbol);
}

AtomList read_xyz(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Error opening .xyz file.");
    std::string line;
    std::getline(in, line);
    size_t count = std::stoul(line);
    std::getline(in, line);
    double scale = 1.0;
    {
        std::istringstream hs(line);
        std::string tok;
        while (hs >> tok)
            if (tok == "nm") scale = 10.0;
    }
    AtomList a;
    while (std::getline(in, line)) {
        if (line.empty() || line == "\r") break;
        std::istringstream ls(line);
        std::string sym, sx, sy, sz;
        if (!(ls >> sym >> sx >> sy >> sz))
            throw std::runtime_error(".xyz file columns are fewer than header entries. line: " +
                                     std::to_string(3 + a.size()));
        a.Z.push_back(element_number(sym));
        a.x.push_back((double)std::stof(sx) * scale);
        a.y.push_back((double)std::stof(sy) * scale);
        a.z.push_back((double)std::stof(sz) * scale);
    }
    if (a.size() != count)
        throw std::runtime_error("Number of atoms does not match .xyz first line: " + std::to_string(a.size()) +
                                 " instead of " + std::to_string(count));
    return a;
}

AtomList srtio3(int nx, int ny, int nz, double a) {
    const struct { int Z; double x, y, z; } basis[5] = {
        {38, 0.0, 0.0, 0.0}, {22, 0.5, 0.5, 0.5}, {8, 0.5, 0.5, 0.0}, {8, 0.5, 0.0, 0.5}, {8, 0.0, 0.5, 0.5}};
    AtomList out;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (const auto& b : basis) {
                    out.Z.push_back(b.Z);
                    out.x.push_back((double)(float)((i + b.x) * a));
                    out.y.push_back((double)(float)((j + b.y) * a));
                    out.z.push_back((double)(float)((k + b.z) * a));
                }
    return out;
}

// Philox4x32-10 (Salmon et al., SC11)
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

AtomList displaced(const AtomList& a, double u_rms, uint64_t seed, int config) {
    AtomList o = a;
    const uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    for (size_t i = 0; i < a.size(); ++i) {
        const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)config, 0u, 0u};
        uint32_t r[4];
        philox4x32_10(ctr, key, r);
        double u[4];
        for (int k = 0; k < 4; ++k) u[k] = ((double)r[k] + 0.5) / 4294967296.0;
        const double rad0 = std::sqrt(-2.0 * std::log(u[0]));
        const double rad1 = std::sqrt(-2.0 * std::log(u[2]));
        const double n0 = rad0 * std::cos(2.0 * M_PI * u[1]);
        const double n1 = rad0 * std::sin(2.0 * M_PI * u[1]);
        const double n2 = rad1 * std::cos(2.0 * M_PI * u[3]);
        o.x[i] = (double)(float)(a.x[i] + u_rms * n0);
        o.y[i] = (double)(float)(a.y[i] + u_rms * n1);
        o.z[i] = (double)(float)(a.z[i] + u_rms * n2);
    }
    return o;
}

// ---- incoherent effects -----------------------------------------------------------------
static void uniforms4(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double u[4]) {
    const uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    for (int k = 0; k < 4; ++k) u[k] = ((double)r[k] + 0.5) / 4294967296.0;
}

void ChromaticSpread::setHalfWidthHalfMaxs(double hwhm_neg, double hwhm_pos) {
    const double f = 2.0 * std::sqrt(2.0 * std::log(2.0));
    sigma_neg = std::fabs(2.0 * hwhm_neg / f);
    sigma_pos = std::fabs(2.0 * hwhm_pos / f);
}

void SourceSize::setFullWidthHalfMax(double fwhm) { sigma = fwhm / (2.0 * std::sqrt(2.0 * std::log(2.0))); }

std::vector<double> PlasmonScattering::depths(uint64_t seed, int job, double thickness) const {
    const int attempts = individual > 0 ? (1 << 20) : 1;
    for (int a = 0; a < attempts; ++a) {
        std::vector<double> d;
        double t = 0.0;
        for (uint32_t k = 0;; ++k) {
            double u[4];
            uniforms4(seed, (uint32_t)job, 0x9Au + 256u * (uint32_t)a, k, 4u, u);
            t += -mean_free_path * std::log(u[0]);
            if (t > thickness) break;
            d.push_back(t);
            if (individual > 0 && (int)d.size() > individual) break;
        }
        if (individual <= 0 || (int)d.size() == individual) return d;
    }
    throw std::runtime_error("Could not generate valid plasmon configuration.");
}

void PlasmonScattering::angles(uint64_t seed, int job, int n, double* polar_mrad, double* azimuth) const {
    double u[4];
    uniforms4(seed, (uint32_t)job, 0xA5u, (uint32_t)n, 5u, u);
    const double ce2 = characteristic_angle * characteristic_angle;
    const double cc2 = critical_angle * critical_angle;
    *polar_mrad = std::sqrt(ce2 * (std::pow((cc2 + ce2) / ce2, u[0]) - 1.0));
    *azimuth = 2.0 * M_PI * u[1];
}

static void rotation_about(const double* ax, double angle, double R[3][3]) {
    const double n = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    const double a[3] = {ax[0] / n, ax[1] / n, ax[2] / n};
    const double C[3][3] = {{0.0, -a[2], a[1]}, {a[2], 0.0, -a[0]}, {-a[1], a[0], 0.0}};
    const double s = std::sin(angle), c1 = 1.0 - std::cos(angle);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double cc = 0.0;
            for (int k = 0; k < 3; ++k) cc += C[i][k] * C[k][j];
            R[i][j] = (i == j ? 1.0 : 0.0) + C[i][j] * s + cc * c1;
        }
}

void rotate_vector_spherical(double z[3], double y[3], double theta, double phi) {
    double rt[3][3], rp[3][3];
    rotation_about(y, theta, rt);
    rotation_about(z, phi, rp);
    auto apply = [](const double R[3][3], double* v) {
        const double r0 = R[0][0] * v[0] + R[0][1] * v[1] + R[0][2] * v[2];
        const double r1 = R[1][0] * v[0] + R[1][1] * v[1] + R[1][2] * v[2];
        const double r2 = R[2][0] * v[0] + R[2][1] * v[1] + R[2][2] * v[2];
        v[0] = r0;
        v[1] = r1;
        v[2] = r2;
    };
    apply(rt, z);
    apply(rp, z);
    apply(rp, y);
}

int SimulationManager::mixedSets() const {
    // clTEM needs frozen phonons, precalculated transmission functions (no plasmons) and a
    // count below the number of iterations (simulationmanager.h:218-231)
    const bool phonons = phonon_configs > 1 && u_rms > 0.0;
    if (!phonons || plasmon.enabled || parallel_potentials < 2) return 0;
    // STEM shares sets only with the static area (parallelStem())
    if (scenario.mode == 2 && !parallel_stem) return 0;
    if (scenario.mode != 2 && parallel_potentials >= phonon_configs) return 0;
    return parallel_potentials;
}

uint64_t SimulationManager::mixedSeed(int job, int batch) const {
    const uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)job, 0x3Bu, (uint32_t)batch, 7u};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    return (uint64_t)r[0] | ((uint64_t)r[1] << 32);
}

double SimulationManager::focusChange(int job, int probe) const {
    if (!chromatic.enabled) return 0.0;
    double u[4];
    uniforms4(seed, (uint32_t)job, 0xC0u, (uint32_t)probe, 1u, u);
    // generateDist: a selector picks the negative or positive half-Gaussian in proportion
    // sigma_neg : sigma_pos; ratio = (n/p) / (n/p + 1) is written n / (n + p)
    const double tot = chromatic.sigma_neg + chromatic.sigma_pos;
    const double ratio = tot > 0.0 ? chromatic.sigma_neg / tot : 0.0;
    const double mag = std::fabs(std::sqrt(-2.0 * std::log(u[1])) * std::cos(2.0 * M_PI * u[2]));
    const double de = u[0] >= ratio ? mag * chromatic.sigma_pos : -mag * chromatic.sigma_neg;
    return 10000.0 * chromatic.cc_mm * de / scenario.voltage_kv;
}

void SimulationManager::sourceOffset(int job, double* dx, double* dy) const {
    *dx = *dy = 0.0;
    if (!source.enabled) return;
    double u[4];
    uniforms4(seed, (uint32_t)job, 0x50u, 0u, 2u, u);
    const double r = std::sqrt(-2.0 * std::log(u[0]));
    *dx = r * std::cos(2.0 * M_PI * u[1]) * source.sigma;
    *dy = r * std::sin(2.0 * M_PI * u[1]) * source.sigma;
}

// ---- SimulationManager ------------------------------------------------------------------
void SimulationManager::updateImages(std::map<std::string, Image>& ims, int jobCount) {
    std::lock_guard<std::mutex> lck(image_update_mutex);
    for (auto& kv : ims) {
        auto it = images.find(kv.first);
        if (it == images.end()) {
            images[kv.first] = kv.second;
            continue;
        }
        Image& cur = it->second;
        const Image& im = kv.second;
        if (im.slices.size() != cur.slices.size() ||
            (!im.slices.empty() && im.slices[0].size() != cur.slices[0].size()))
            throw std::runtime_error("Tried to merge simulation jobs with different output size");
        if (im.weighting.size() != cur.weighting.size())
            throw std::runtime_error("Tried to merge simulation jobs with different weighting sizes");
        for (size_t j = 0; j < cur.slices.size(); ++j)
            for (size_t k = 0; k < cur.slices[j].size(); ++k) cur.slices[j][k] += im.slices[j][k];
        for (size_t k = 0; k < cur.weighting.size(); ++k) cur.weighting[k] += im.weighting[k];
    }
    complete_jobs += jobCount;
}

std::map<std::string, Image> SimulationManager::averaged() const {
    std::map<std::string, Image> out = images;
    for (auto& kv : out) {
        Image& im = kv.second;
        for (auto& sl : im.slices) {
            const size_t per = im.weighting.empty() ? 1 : sl.size() / im.weighting.size();
            for (size_t k = 0; k < sl.size(); ++k) {
                const double w = im.weighting.empty() ? 1.0 : im.weighting[k / per];
                sl[k] = w > 0 ? sl[k] / w : 0.0;
            }
        }
    }
    return out;
}

mse_geometry SimulationManager::geometry() const {
    mse_scenario s = scenario;
    s.moving_stem_frame = (s.mode == 2 && !parallel_stem) ? 1 : 0;
    mse_geometry g;
    check(mse_plan_geometry(&s, &g));
    return g;
}

// ---- SimulationGeneral --------------------------------------------------------------------
SimulationGeneral::SimulationGeneral(int device) { check(mse_create(device, &eng)); }

SimulationGeneral::~SimulationGeneral() {
    if (eng) mse_destroy(eng);
}

void SimulationGeneral::setJob(std::shared_ptr<SimulationManager> m, int c, std::vector<int> px) {
    sm = std::move(m);
    config = c;
    pixels = std::move(px);
}

static void load_params(const std::string& path, int* form, int* per, std::vector<double>* values) {
    // utilities/stringutils.cpp:149-184 file format
    std::ifstream in(path);
    if (in.fail()) throw std::runtime_error("Error opening resource file: " + path);
    std::string name, form_name, tmp;
    std::getline(in, name);
    std::getline(in, form_name);
    std::getline(in, tmp);
    *per = std::stoi(tmp);
    double p;
    values->clear();
    while (in >> p) values->push_back(p);
    auto trim = [](std::string s) {
        while (!s.empty() && (s.back() == '\r' || s.back() == ' ')) s.pop_back();
        return s;
    };
    form_name = trim(form_name);
    if (form_name == "kirkland")
        *form = 0;
    else if (form_name == "peng")
        *form = 1;
    else if (form_name == "lobato")
        *form = 2;
    else
        throw std::runtime_error("Unknown parameterisation form " + form_name + " in file " + path);
}

void SimulationGeneral::sortAtoms() {
    const bool phonon = sm->phonon_configs > 1 || sm->u_rms > 0.0;
    const AtomList a = phonon ? displaced(sm->atoms, sm->u_rms, sm->seed, config) : sm->atoms;
    check(mse_sort_atoms(eng, 0, a.x.data(), a.y.data(), a.z.data(), a.Z.data(), (int)a.size(), nullptr, nullptr,
                         nullptr, nullptr));
}

bool SimulationGeneral::initialiseSimulation() {
    const int sets = sm->mixedSets();
    if (sets >= 2 && sets_for == sm.get() && !sm->source.enabled) {
        // the shared sets of this manager are built: only the job's draws change
        // (simulationgeneral.cpp:358-362 reuses them the same way)
        check(mse_set_mixed_potentials(eng, 0, sets, sm->mixedSeed(config)));
        return true;
    }
    geom = sm->geometry();
    check(mse_set_geometry(eng, &geom));  // every configuration back to the geometry's beam
    beam_moved = false;
    sets_for = nullptr;
    int form = 0, per = 0;
    std::vector<double> values;
    load_params(sm->params_file, &form, &per, &values);
    check(mse_set_parameterisation(eng, form, per, values.data(), (int)values.size()));
    if (sets >= 2) {
        // one phonon configuration per set (simulationgeneral.cpp:576-612), in configurations 0..sets-1
        for (int j = 0; j < sets; ++j) {
            const AtomList a = displaced(sm->atoms, sm->u_rms, sm->seed, j);
            check(mse_sort_atoms(eng, j, a.x.data(), a.y.data(), a.z.data(), a.Z.data(), (int)a.size(), nullptr,
                                 nullptr, nullptr, nullptr));
            check(mse_set_reference_offset(eng, j, reference_perturb_x, reference_perturb_y));
            check(mse_build_transmission(eng, j, 0, geom.n_slices));
        }
        check(mse_set_mixed_potentials(eng, 0, sets, sm->mixedSeed(config)));
        sets_for = sm.get();
        return true;
    }
    check(mse_set_mixed_potentials(eng, 0, 0, 0));
    sortAtoms();
    // simulationgeneral.cpp:559-560: the potential grid origin carries the source-size offset
    check(mse_set_reference_offset(eng, 0, reference_perturb_x, reference_perturb_y));
    // a moving STEM frame builds every pixel's transmission functions in its own field
    // (mse_stem_batch_frames), none in pixel 0's
    if (!geom.moving_frame) check(mse_build_transmission(eng, 0, 0, geom.n_slices));
    return true;
}

void SimulationGeneral::doMultiSliceStep(int slice) { check(mse_multislice(eng, slice, 1)); }

void SimulationGeneral::resetBeam() {
    if (!beam_moved) return;
    check(mse_set_config_beam(eng, 0, nullptr));
    check(mse_build_transmission(eng, 0, 0, geom.n_slices));
    beam_moved = false;
}

int SimulationGeneral::runSlices(int job, const std::function<void(int)>& output) {
    const int n = geom.n_slices;
    const int step = sm->slice_step;
    const int n_out = sm->outputCount(n);
    const bool plasmons = sm->plasmon.enabled;
    std::vector<double> depths;
    // beam vector and the azimuth axis at the top (simulationcbed.cpp:152-162)
    double k[3] = {0.0, 0.0, geom.wavenumber}, yax[3] = {0.0, 1.0, 0.0};
    if (plasmons) {
        // the structure's own thickness (console/main.cpp:711-717)
        double zmin = 0.0, zmax = 0.0;
        if (sm->atoms.size() > 0) {
            zmin = *std::min_element(sm->atoms.z.begin(), sm->atoms.z.end());
            zmax = *std::max_element(sm->atoms.z.begin(), sm->atoms.z.end());
        }
        depths = sm->plasmon.depths(sm->seed, job, zmax - zmin);
        rotate_vector_spherical(k, yax, geom.beam_tilt_mrad / 1000.0, geom.beam_azimuth);
    }
    size_t next = 0;
    int first = 0, events = 0, outs = 0;
    for (int i = 0; i < n; ++i) {
        const double depth = (i + 1 - geom.pre_slices) * geom.slice_thickness;
        const bool scatter = plasmons && next < depths.size() && depth >= depths[next];
        const bool out = step > 0 && (i + 1) % step == 0;
        if (!scatter && !out) continue;
        check(mse_multislice(eng, first, i + 1 - first));
        first = i + 1;
        if (scatter) {
            double polar = 0.0, azimuth = 0.0;
            sm->plasmon.angles(sm->seed, job, (int)next, &polar, &azimuth);
            rotate_vector_spherical(k, yax, polar / 1000.0, azimuth);
            check(mse_set_config_beam(eng, 0, k));
            beam_moved = true;
            if (i + 1 < n) check(mse_build_transmission(eng, 0, i + 1, n - i - 1));
            ++next;
            ++events;
        }
        if (out) output(outs++);
    }
    if (first < n) check(mse_multislice(eng, first, n - first));
    if (outs < n_out) output(outs);
    return events;
}

std::vector<double> SimulationGeneral::getDiffractionImage(int parallel_ind) {
    std::vector<double> out((size_t)geom.resolution * geom.resolution);
    check(mse_diffraction(eng, parallel_ind, out.data()));
    return out;
}

std::vector<double> SimulationGeneral::getExitWaveImage(unsigned t, unsigned l, unsigned b, unsigned r) {
    const unsigned res = (unsigned)geom.resolution;
    std::vector<double> full(2 * (size_t)res * res);
    check(mse_exit_wave(eng, 0, full.data()));
    // cropping of simulationgeneral.cpp:816-826
    std::vector<double> out(2 * (size_t)(res - t - b) * (res - l - r));
    size_t cnt = 0;
    for (unsigned j = 0; j < res; ++j)
        if (j >= b && j < res - t)
            for (unsigned i = 0; i < res; ++i)
                if (i >= l && i < res - r) {
                    const size_t k = i + (size_t)j * res;
                    out[cnt] = full[2 * k];
                    out[cnt + 1] = full[2 * k + 1];
                    cnt += 2;
                }
    return out;
}

// ---- CTEM -------------------------------------------------------------------------------
bool SimulationCtem::initialiseSimulation() {
    reference_perturb_x = reference_perturb_y = 0.0;  // simulationctem.cpp:59-60
    SimulationGeneral::initialiseSimulation();
    check(mse_set_wave_count(eng, 1));
    const int c0 = 0;
    check(mse_set_wave_configs(eng, &c0, 1));
    check(mse_init_plane_wave(eng, 1.0));  // simulationctem.cpp:77-82
    n_waves = 1;
    return true;
}

void SimulationCtem::simulate() {
    initialiseSimulation();
    const int res = geom.resolution;
    const int n_out = sm->outputCount(geom.n_slices);
    Image ew(res, res, n_out), diff(res, res, n_out);
    for (auto& sl : ew.slices) sl.resize(2 * (size_t)res * res);  // interleaved re/im (averaged as complex)
    runSlices(config, [&](int k) {
        ew.slices[k] = getExitWaveImage();
        diff.slices[k] = getDiffractionImage(0);
    });
    std::fill(ew.weighting.begin(), ew.weighting.end(), 1.0);
    std::fill(diff.weighting.begin(), diff.weighting.end(), 1.0);
    std::map<std::string, Image> ims;
    ims["EW"] = ew;
    ims["Diff"] = diff;
    sm->updateImages(ims, 1);
}

// ---- CBED -------------------------------------------------------------------------------
bool SimulationCbed::initialiseSimulation() {
    sm->sourceOffset(config, &reference_perturb_x, &reference_perturb_y);
    return SimulationGeneral::initialiseSimulation();
}

void SimulationCbed::initialiseProbeWave(double posx, double posy, int n_parallel) {
    if (n_parallel != 0) throw std::runtime_error("this adapter initialises probes from wave 0");
    const double pos[2] = {posx, posy};
    const double df = sm->focusChange(config, 0);  // simulationcbed.cpp:62-71
    check(mse_set_focus_spread(eng, &df, sm->chromatic.enabled ? 1 : 0));
    check(mse_init_probes(eng, pos, 1, &sm->probe));
}

void SimulationCbed::simulate() {
    initialiseSimulation();
    check(mse_set_wave_count(eng, 1));
    const int c0 = 0;
    check(mse_set_wave_configs(eng, &c0, 1));
    n_waves = 1;
    // CbedPosition: the mode's area is the point (x, x, y, y)
    initialiseProbeWave(sm->scenario.area[0], sm->scenario.area[2]);
    const int res = geom.resolution;
    Image diff(res, res, sm->outputCount(geom.n_slices));
    runSlices(config, [&](int k) { diff.slices[k] = getDiffractionImage(0); });
    std::map<std::string, Image> ims;
    std::fill(diff.weighting.begin(), diff.weighting.end(), 1.0);
    ims["Diff"] = diff;
    sm->updateImages(ims, 1);
}

// ---- STEM -------------------------------------------------------------------------------
double SimulationStem::getStemPixel(double inner, double outer, double xc, double yc, int parallel_ind) {
    if (parallel_ind < 0 || parallel_ind >= n_waves) throw std::runtime_error("parallel index out of range");
    mse_detector d{inner, outer, xc, yc};
    std::vector<double> out(n_waves);
    check(mse_stem_detect(eng, &d, 1, sm->sqrt2 ? 1 : 0, out.data()));
    return out[parallel_ind];
}

void SimulationStem::simulate() {
    SimulationCbed::initialiseSimulation();
    const int nx = geom.stem_pixels[0], ny = geom.stem_pixels[1];
    if (pixels.empty())
        for (int p = 0; p < nx * ny; ++p) pixels.push_back(p);
    // probe positions of simulationstem.cpp:156-187
    std::vector<double> pos(2 * pixels.size());
    for (size_t i = 0; i < pixels.size(); ++i) {
        const int p = pixels[i];
        pos[2 * i] = geom.stem_start[0] + (p % nx) * geom.stem_scale[0];
        pos[2 * i + 1] = geom.stem_start[1] + (p / nx) * geom.stem_scale[1];
    }
    const int nd = (int)sm->detectors.size();
    std::vector<mse_detector> dets(nd);
    for (int d = 0; d < nd; ++d) dets[d] = sm->detectors[d].d;
    const int n_out = sm->outputCount(geom.n_slices);
    std::vector<double> acc(pixels.size() * nd * n_out, 0.0);  // [output][pixel][detector]
    const int batch = std::max(1, std::min(sm->parallel_pixels, (int)pixels.size()));
    // every probe draws its own focus change (initialiseProbeWave is called per probe)
    std::vector<double> df(sm->chromatic.enabled ? pixels.size() : 0);
    for (size_t i = 0; i < df.size(); ++i) df[i] = sm->focusChange(config, pixels[i]);
    check(mse_set_focus_spread(eng, df.data(), (int)df.size()));
    if (geom.moving_frame) {
        // parallel_stem off: one field per pixel (currentAreaBase(pixel), simulationmanager.cpp:485-501)
        if (sm->plasmon.enabled || sm->slice_step > 0)
            throw std::runtime_error("the moving STEM frame runs without plasmons or intermediate outputs here");
        std::vector<double> frames(2 * pixels.size());
        for (size_t i = 0; i < pixels.size(); ++i) check(mse_stem_pixel_frame(&geom, pixels[i], &frames[2 * i]));
        check(mse_stem_batch_frames(eng, 0, pos.data(), frames.data(), (int)pixels.size(), batch, &sm->probe,
                                    dets.data(), nd, sm->sqrt2 ? 1 : 0, acc.data()));
        n_waves = batch;
    } else if (!sm->plasmon.enabled && sm->slice_step <= 0) {
        check(mse_stem_batch(eng, 0, pos.data(), (int)pixels.size(), batch, &sm->probe, dets.data(), nd,
                             sm->sqrt2 ? 1 : 0, 1, acc.data()));
        n_waves = batch;
    } else {
        // intermediate outputs or plasmons: the batches go through the slice loop of
        // runSlices with getStemPixel at every output (simulationstem.cpp:210-270); with
        // plasmons every batch of pixels is one job with its own scattering history, as each
        // of clTEM's STEM jobs is, and the detectors read the diffraction pattern moved back by
        // the beam change (getStemPixel with d_kx, d_ky)
        const int n_batches = ((int)pixels.size() + batch - 1) / batch;
        std::vector<double> out;
        for (int b = 0; b < n_batches; ++b) {
            const int p0 = b * batch, nb = std::min(batch, (int)pixels.size() - p0);
            resetBeam();
            if (n_waves != nb) {
                check(mse_set_wave_count(eng, nb));
                std::vector<int> wc(nb, 0);
                check(mse_set_wave_configs(eng, wc.data(), nb));
                n_waves = nb;
            }
            if (!df.empty()) check(mse_set_focus_spread(eng, df.data() + p0, nb));
            // the probe set-up keys the mixed-set draws 0..nb-1: each batch draws with its own seed
            if (sm->mixedSets() >= 2) check(mse_set_mixed_potentials(eng, 0, sm->mixedSets(), sm->mixedSeed(config, b)));
            check(mse_init_probes(eng, pos.data() + 2 * (size_t)p0, nb, &sm->probe));
            runSlices(config * n_batches + b, [&](int k) {
                out.assign((size_t)nb * nd, 0.0);
                check(mse_stem_detect(eng, dets.data(), nd, sm->sqrt2 ? 1 : 0, out.data()));
                std::copy(out.begin(), out.end(), acc.begin() + ((size_t)k * pixels.size() + p0) * nd);
            });
        }
    }
    std::map<std::string, Image> ims;
    for (int d = 0; d < nd; ++d) {
        Image im(nx, ny, n_out);
        for (int k = 0; k < n_out; ++k)
            for (size_t i = 0; i < pixels.size(); ++i) im.slices[k][pixels[i]] = acc[((size_t)k * pixels.size() + i) * nd + d];
        for (size_t i = 0; i < pixels.size(); ++i) im.weighting[pixels[i]] = 1.0;
        ims[sm->detectors[d].name] = im;
    }
    sm->updateImages(ims, 1);
}

// ---- output writers -----------------------------------------------------------------------
namespace {
void put16(std::vector<unsigned char>& b, uint16_t v) {
    b.push_back(v & 0xff);
    b.push_back(v >> 8);
}
void put32(std::vector<unsigned char>& b, uint32_t v) {
    for (int k = 0; k < 4; ++k) b.push_back((v >> (8 * k)) & 0xff);
}
}  // namespace

void save_tiff(const std::string& path, const std::vector<double>& data, int width, int height) {
    if ((size_t)width * height != data.size())
        throw std::runtime_error("Attempting to save image with incommensurate data size and image dimensions");
    // little-endian baseline TIFF: header, pixel strip, then the IFD
    std::vector<unsigned char> b;
    b.push_back('I');
    b.push_back('I');
    put16(b, 42);
    const uint32_t strip_off = 8, strip_bytes = (uint32_t)(data.size() * sizeof(float));
    const uint32_t ifd_off = strip_off + strip_bytes + (strip_bytes & 1);
    put32(b, ifd_off);
    for (double v : data) {
        const float f = (float)v;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        put32(b, u);
    }
    if (strip_bytes & 1) b.push_back(0);
    struct Tag { uint16_t id, type; uint32_t count, value; };
    const Tag tags[] = {
        {256, 4, 1, (uint32_t)width},   // ImageWidth
        {257, 4, 1, (uint32_t)height},  // ImageLength
        {258, 3, 1, 32},                // BitsPerSample
        {259, 3, 1, 1},                 // Compression: none
        {262, 3, 1, 1},                 // Photometric: min-is-black
        {273, 4, 1, strip_off},         // StripOffsets
        {274, 3, 1, 4},                 // Orientation: bottom-left
        {277, 3, 1, 1},                 // SamplesPerPixel
        {278, 4, 1, (uint32_t)height},  // RowsPerStrip
        {279, 4, 1, strip_bytes},       // StripByteCounts
        {284, 3, 1, 1},                 // PlanarConfiguration: contiguous
        {339, 3, 1, 3},                 // SampleFormat: IEEE float
    };
    put16(b, (uint16_t)(sizeof(tags) / sizeof(tags[0])));
    for (const Tag& t : tags) {
        put16(b, t.id);
        put16(b, t.type);
        put32(b, t.count);
        if (t.type == 3) {  // SHORT values sit left-justified in the value field
            put16(b, (uint16_t)t.value);
            put16(b, 0);
        } else {
            put32(b, t.value);
        }
    }
    put32(b, 0);  // no further IFD
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Could not open .tif file for saving: " + path);
    f.write((const char*)b.data(), (std::streamsize)b.size());
    if (!f) throw std::runtime_error("Unable to write data to .tif file");
}

namespace {
// A JSON value with the object ordering of nlohmann::json's default type (keys sorted) and
// its setw(4) text, enough for the settings files.
struct Json {
    enum Kind { Null, Bool, Int, Num, Str, Obj } kind = Null;
    bool b = false;
    long long i = 0;
    double d = 0.0;
    std::string s;
    std::map<std::string, Json> o;
    Json& operator[](const std::string& k) {
        kind = Obj;
        return o[k];
    }
    Json& operator=(bool v) { kind = Bool; b = v; return *this; }
    Json& operator=(int v) { kind = Int; i = v; return *this; }
    Json& operator=(unsigned v) { kind = Int; i = v; return *this; }
    Json& operator=(long long v) { kind = Int; i = v; return *this; }
    Json& operator=(double v) { kind = Num; d = v; return *this; }
    Json& operator=(const char* v) { kind = Str; s = v; return *this; }
    Json& operator=(const std::string& v) { kind = Str; s = v; return *this; }
    void erase(const std::string& k) { o.erase(k); }

    static std::string number(double v) {
        if (!std::isfinite(v)) return "null";
        char buf[40];
        for (int p = 1; p <= 17; ++p) {  // shortest text that reads back as the same double
            std::snprintf(buf, sizeof(buf), "%.*g", p, v);
            if (std::strtod(buf, nullptr) == v) break;
        }
        std::string t(buf);
        if (t.find_first_of(".e") == std::string::npos) t += ".0";
        return t;
    }
    static std::string quoted(const std::string& t) {
        std::string r = "\"";
        for (unsigned char c : t) {
            if (c == '"' || c == '\\') {
                r += '\\';
                r += (char)c;
            } else if (c < 0x20) {
                char u[8];
                std::snprintf(u, sizeof(u), "\\u%04x", c);
                r += u;
            } else {
                r += (char)c;
            }
        }
        return r + "\"";
    }
    void dump(std::ostream& os, int depth) const {
        switch (kind) {
            case Null: os << "null"; break;
            case Bool: os << (b ? "true" : "false"); break;
            case Int: os << i; break;
            case Num: os << number(d); break;
            case Str: os << quoted(s); break;
            case Obj: {
                if (o.empty()) {
                    os << "{}";
                    break;
                }
                os << "{\n";
                size_t n = 0;
                for (const auto& kv : o) {
                    os << std::string(4 * (depth + 1), ' ') << quoted(kv.first) << ": ";
                    kv.second.dump(os, depth + 1);
                    os << (++n < o.size() ? ",\n" : "\n");
                }
                os << std::string(4 * depth, ' ') << "}";
                break;
            }
        }
    }
};

const double kDeg = 180.0 / M_PI;

// (magnitude, angle in degrees) of a complex aberration held as (re, im)
void polar_aberration(Json& j, const double* c, double scale, const char* units) {
    j["mag"] = std::hypot(c[0], c[1]) / scale;
    j["ang"] = std::atan2(c[1], c[0]) * kDeg;
    j["units"] = units;
}

std::string base_name(const std::string& path) {
    std::string b = path.substr(path.find_last_of("/\\") + 1);
    const size_t dot = b.find_last_of('.');
    return dot == std::string::npos ? b : b.substr(0, dot);
}
}  // namespace

std::string settings_json(const SimulationManager& sm, const std::string& image) {
    const mse_scenario& sc = sm.scenario;
    const mse_geometry g = sm.geometry();
    const int mode = sc.mode;
    const bool probe_mode = mode == 2 || mode == 3;
    static const char* kModes[] = {"None", "CTEM", "STEM", "CBED"};
    Json j;
    j["double precision"] = false;
    j["mode"]["id"] = mode;
    j["mode"]["name"] = kModes[std::min(std::max(mode, 0), 3)];
    j["potentials"] = base_name(sm.params_file);
    j["resolution"] = sc.resolution;
    j["slice thickness"]["val"] = sc.slice_thickness;
    j["slice thickness"]["units"] = "Å";
    j["slice offset"]["val"] = sc.slice_offset;
    j["slice offset"]["units"] = "Å";
    j["slice count"] = g.n_slices;
    j["intermediate output"]["slice interval"] = sm.slice_step;
    j["intermediate output"]["enabled"] = sm.slice_step > 0;
    j["maintain areas"] = false;
    const char* axes[] = {"x", "y", "z"};
    const double starts[] = {g.sim_x[0], g.sim_y[0], g.z_lims[0]}, ends[] = {g.sim_x[1], g.sim_y[1], g.z_lims[1]};
    const double pads[] = {sc.pad_xy[1], sc.pad_xy[1], g.pad_z[1]};
    for (int k = 0; k < 3; ++k) {
        Json& ax = j["simulation area"][axes[k]];
        ax["start"] = starts[k];
        ax["finish"] = ends[k];
        ax["padding"] = pads[k];
        ax["units"] = "Å";
    }
    j["default padding"]["z"]["val"] = sc.pad_z[1];
    j["default padding"]["z"]["units"] = "Å";
    j["default padding"]["xy"]["val"] = sc.pad_xy[1];
    j["default padding"]["xy"]["units"] = "Å";
    j["full 3d"]["state"] = sc.full3d != 0;
    if (sc.full3d) j["full 3d"]["integrals"] = sc.full3d_integrals;
    j["precalculate transmission"] = !sm.plasmon.enabled;

    Json& mic = j["microscope"];
    mic["voltage"]["val"] = sc.voltage_kv;
    mic["voltage"]["units"] = "kV";
    if (mode == 1) {  // the objective lens of CTEM image formation (not run by this driver)
        mic["objective aperture"]["semi-angle"] = 20.0;
        mic["objective aperture"]["smoothing"] = 0.0;
        mic["objective aperture"]["units"] = "mrad";
        mic["alpha"]["val"] = 0.0;
        mic["alpha"]["units"] = "mrad";
        mic["delta"]["val"] = 0.0;
        mic["delta"]["units"] = "nm";
    } else {
        mic["condenser aperture"]["semi-angle"] = sm.probe.cond_ap_mrad;
        mic["condenser aperture"]["smoothing"] = sm.probe.smooth_mrad;
        mic["condenser aperture"]["units"] = "mrad";
    }
    mic["beam tilt"]["inclination"]["val"] = sc.beam_tilt_mrad;
    mic["beam tilt"]["inclination"]["units"] = "mrad";
    mic["beam tilt"]["azimuth"]["val"] = sc.beam_azimuth * kDeg;
    mic["beam tilt"]["azimuth"]["units"] = "°";
    const mse_probe& p = sm.probe;
    Json& ab = mic["aberrations"];
    ab["C10"]["val"] = p.C10 / 10.0;
    ab["C10"]["units"] = "nm";
    polar_aberration(ab["C12"], p.C12, 10.0, "nm, °");
    polar_aberration(ab["C21"], p.C21, 10.0, "nm, °");
    polar_aberration(ab["C23"], p.C23, 10.0, "nm, °");
    ab["C30"]["val"] = p.C30 / 10000.0;
    ab["C30"]["units"] = "μm";
    polar_aberration(ab["C32"], p.C32, 10000.0, "μm, °");
    polar_aberration(ab["C34"], p.C34, 10000.0, "μm, °");
    polar_aberration(ab["C41"], p.C41, 10000.0, "μm, °");
    polar_aberration(ab["C43"], p.C43, 10000.0, "μm, °");
    polar_aberration(ab["C45"], p.C45, 10000.0, "μm, °");
    ab["C50"]["val"] = p.C50 / 10000.0;
    ab["C50"]["units"] = "μm";
    polar_aberration(ab["C52"], p.C52, 10000.0, "μm, °");
    polar_aberration(ab["C54"], p.C54, 10000.0, "μm, °");
    polar_aberration(ab["C56"], p.C56, 10000.0, "μm, °");

    if (mode == 1) {
        j["ctem"]["simulate image"] = false;
        j["ctem"]["area"]["x"]["start"] = sc.area[0];
        j["ctem"]["area"]["x"]["finish"] = sc.area[1];
        j["ctem"]["area"]["y"]["start"] = sc.area[2];
        j["ctem"]["area"]["y"]["finish"] = sc.area[3];
    }
    if (mode == 2) {
        Json& st = j["stem"];
        st["area"]["x"]["start"] = sc.area[0];
        st["area"]["x"]["finish"] = sc.area[1];
        st["area"]["x"]["units"] = "Å";
        st["area"]["y"]["start"] = sc.area[2];
        st["area"]["y"]["finish"] = sc.area[3];
        st["area"]["y"]["units"] = "Å";
        st["scan"]["x"]["pixels"] = sc.stem_pixels[0];
        st["scan"]["y"]["pixels"] = sc.stem_pixels[1];
        st["area"]["padding"]["val"] = sc.area_padding;
        st["area"]["padding"]["units"] = "Å";
        st["static area"]["concurrent pixels"] = sm.parallel_pixels;
        st["static area"]["enabled"] = sm.parallel_stem;
    }
    if (mode == 3) {
        j["cbed"]["position"]["x"] = sc.area[0];
        j["cbed"]["position"]["y"] = sc.area[2];
        j["cbed"]["position"]["padding"] = sc.area_padding;
        j["cbed"]["position"]["units"] = "Å";
    }

    // incoherent effects (jsonutils.cpp:657-778)
    const bool phonons = sm.phonon_configs > 1 && sm.u_rms > 0.0;
    const bool probe_effects = probe_mode && (sm.chromatic.enabled || sm.source.enabled);
    if (phonons || probe_effects || sm.plasmon.enabled)
        j["incoherence"]["iterations"] = std::max(1, sm.phonon_configs);
    if (probe_mode && (phonons || probe_effects || sm.plasmon.enabled)) {
        Json& pr = j["incoherence"]["probe"];
        const double fwhm = 2.0 * std::sqrt(2.0 * std::log(2.0));
        pr["chromatic"]["enabled"] = sm.chromatic.enabled;
        if (sm.chromatic.enabled) {
            pr["chromatic"]["Cc"]["val"] = sm.chromatic.cc_mm;
            pr["chromatic"]["Cc"]["units"] = "mm";
            pr["chromatic"]["dE"]["HWHM +"] = sm.chromatic.sigma_pos * fwhm / 2.0;
            pr["chromatic"]["dE"]["HWHM -"] = sm.chromatic.sigma_neg * fwhm / 2.0;
            pr["chromatic"]["dE"]["units"] = "eV";
        }
        pr["source size"]["enabled"] = sm.source.enabled;
        if (sm.source.enabled) {
            pr["source size"]["FWHM"]["val"] = sm.source.sigma * fwhm;
            pr["source size"]["FWHM"]["units"] = "Å";
        }
    }
    if (phonons) {
        Json& ph = j["incoherence"]["inelastic scattering"]["phonon"];
        if (sm.mixedSets() >= 2) {
            ph["mixed static potentials"]["count"] = sm.mixedSets();
            ph["mixed static potentials"]["enabled"] = true;
        }
        ph["force atom resort"] = false;
    }
    if (sm.plasmon.enabled) {
        Json& pl = j["incoherence"]["inelastic scattering"]["plasmon"];
        pl["enabled"] = true;
        pl["type"] = sm.plasmon.individual > 0 ? "individual" : "full";
        if (sm.plasmon.individual > 0) pl["individual"] = sm.plasmon.individual;
        pl["mean free path"]["value"] = sm.plasmon.mean_free_path / 10.0;
        pl["mean free path"]["units"] = "nm";
        pl["characteristic angle"]["value"] = sm.plasmon.characteristic_angle;
        pl["characteristic angle"]["units"] = "mrad";
        pl["critical angle"]["value"] = sm.plasmon.critical_angle;
        pl["critical angle"]["units"] = "mrad";
    }

    // imageReturned (console/main.cpp:189-228)
    j["filename"] = sm.structure_file;
    if (image == "EW" || image == "Diff") {
        mic.erase("aberrations");
        mic.erase("alpha");
        mic.erase("delta");
        mic.erase("objective aperture");
    } else if (image != "Image") {
        for (const auto& d : sm.detectors)
            if (d.name == image) {
                Json& dj = j["stem"]["detectors"][d.name];
                dj["radius"]["inner"] = d.d.inner_mrad;
                dj["radius"]["outer"] = d.d.outer_mrad;
                dj["radius"]["units"] = "mrad";
                dj["centre"]["x"] = d.d.xc_mrad;
                dj["centre"]["y"] = d.d.yc_mrad;
                dj["centre"]["units"] = "mrad";
            }
        mic.erase("alpha");
        mic.erase("delta");
    }
    std::ostringstream os;
    j.dump(os, 0);
    return os.str();
}

void save_outputs(const SimulationManager& sm, const std::string& prefix) {
    for (const auto& kv : sm.averaged()) {
        const std::string& name = kv.first;
        const Image& im = kv.second;
        const size_t npix = (size_t)im.width * im.height;
        auto write = [&](const std::string& base, auto value) {
            std::ofstream(base + ".json") << settings_json(sm, name);
            for (size_t j = 0; j < im.slices.size(); ++j) {
                std::vector<double> d(npix);
                for (size_t k = 0; k < npix; ++k) d[k] = value(im.slices[j], k);
                std::string fn = base + ".tif";
                if (im.slices.size() > 1) {
                    // saveTiffOutput (console/main.cpp:158-186): the slice that ends output j,
                    // zero-padded to the width of the last slice index
                    const int sc = sm.geometry().n_slices;
                    const int slice_id = std::min((int)(j + 1) * sm.slice_step - 1, sc - 1);
                    std::string id = std::to_string(slice_id);
                    const size_t width = std::to_string(sc - 1).size();
                    id = std::string(width > id.size() ? width - id.size() : 0, '0') + id;
                    fn = base + "_" + id + ".tif";
                }
                save_tiff(fn, d, im.width, im.height);
            }
        };
        if (name == "EW") {  // complex: amplitude and phase
            write(prefix + "_" + name + "_amplitude",
                  [](const std::vector<double>& s, size_t k) { return std::abs(std::complex<double>(s[2 * k], s[2 * k + 1])); });
            write(prefix + "_" + name + "_phase",
                  [](const std::vector<double>& s, size_t k) { return std::arg(std::complex<double>(s[2 * k], s[2 * k + 1])); });
        } else {
            write(prefix + "_" + name, [](const std::vector<double>& s, size_t k) { return s[k]; });
        }
    }
}

void run_simulation(std::shared_ptr<SimulationManager> sm) {
    const int configs = std::max(1, sm->phonon_configs);
    std::unique_ptr<SimulationGeneral> worker;
    if (sm->scenario.mode == 1)
        worker.reset(new SimulationCtem(sm->device));
    else if (sm->scenario.mode == 2)
        worker.reset(new SimulationStem(sm->device));
    else if (sm->scenario.mode == 3)
        worker.reset(new SimulationCbed(sm->device));
    else
        throw std::runtime_error("unknown simulation mode");
    for (int c = 0; c < configs; ++c) {
        worker->setJob(sm, c);
        worker->simulate();
    }
}

}  // namespace cltem_amd
