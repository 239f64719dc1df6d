// mse_run: command-line driver of the engine through the clTEM-shaped host classes.
//
//   mse_run --mode ctem|stem|cbed --params cltem_amd/params/kirkland.dat
//           (--xyz file.xyz | --srtio3 NX,NY,NZ | --cif file.cif --widths X,Y,Z [--zone u,v,w]
//            [--normal a,b,c] [--tilts a,b,g]) --res 512 --dz 1.9525 --kv 200
//           [--area x0,x1,y0,y1] [--stem-pixels N] [--det name,inner,outer[,xc,yc]]...
//           [--alpha mrad] [--batch B] [--moving-frame] [--phonons K --urms A --seed S] [--tilt mrad,azimuth]
//           [--full3d integrals] [--chromatic Cc_mm,hwhm_neg_eV,hwhm_pos_eV] [--source-size fwhm_A]
//           [--plasmon mfp_nm,characteristic_mrad,critical_mrad[,n_individual]] [--parallel-potentials N]
//           [--intermediate SLICES]
//           [--device D] [--tiff] --out PREFIX
//   mse_run ... --dump-atoms FILE [--dump-config C]   (writes the displaced atoms of phonon
//           configuration C as x,y,z doubles and Z ints, and exits without using the GPU)
//   mse_run ... --dump-incoherence N [--dump-config C]   (prints the focus changes of probes
//           0..N-1 and the source offset of job C, one number per line, and exits)
//   mse_run ... --dump-plasmons THICKNESS [--dump-config C]   (prints the scattering count,
//           depths, then polar (mrad) and azimuth of each scattering of job C, and exits)
//   mse_run ... --dump-settings IMAGE   (prints the settings .json written beside image IMAGE
//           -- EW, Diff or a detector name -- and exits without using the GPU)
//   mse_run ... --dump-mixed-seed [--dump-config C]   (prints the seed of job C's mixed
//           potential draws and exits)
//   mse_run --write-test-tiff FILE W H   (writes the ramp x + 1000 y, W x H, and exits)
//
// Writes PREFIX_<image>.bin (doubles, averaged over the phonon configurations as
// SimulationManager::updateImages's weighting does) and PREFIX.json describing them; with
// --tiff also clTEM's outputs (32-bit float .tif + settings .json per image, EW as
// amplitude and phase).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "cltem_host.hpp"

using namespace cltem_amd;

static std::vector<double> split_doubles(const std::string& s) {
    std::vector<double> v;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) v.push_back(std::stod(tok));
    return v;
}

int main(int argc, char** argv) {
    try {
        auto sm = std::make_shared<SimulationManager>();
        mse_scenario& s = sm->scenario;
        s.mode = 1;
        s.resolution = 256;
        s.voltage_kv = 200.0;
        s.slice_thickness = 1.0;
        s.pad_xy[0] = -8.0;
        s.pad_xy[1] = 8.0;
        s.pad_z[0] = -3.0;
        s.pad_z[1] = 3.0;
        s.limit_factor = 2.0 / 3.0;
        s.full3d_integrals = 20;
        sm->probe.cond_ap_mrad = 20.0;
        std::string out = "mse_out";
        std::string dump;
        int dump_config = 0;
        int dump_incoherence = -1;
        double dump_plasmons = -1.0;
        bool dump_mixed = false;
        std::string dump_settings;
        bool tiff = false;
        std::string cif_path;
        double zone[3] = {0, 0, 1}, normal[3] = {0, 0, 0}, widths[3] = {0, 0, 0}, tilts[3] = {0, 0, 0};
        bool have_area = false;
        int stem_px = 64;
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            auto next = [&]() -> std::string {
                if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
                return argv[++i];
            };
            if (a == "--mode") {
                std::string m = next();
                s.mode = m == "ctem" ? 1 : m == "stem" ? 2 : m == "cbed" ? 3 : 0;
            } else if (a == "--params") {
                sm->params_file = next();
            } else if (a == "--xyz") {
                sm->structure_file = next();
                sm->atoms = read_xyz(sm->structure_file);
            } else if (a == "--srtio3") {
                auto v = split_doubles(next());
                sm->atoms = srtio3((int)v.at(0), (int)v.at(1), (int)v.at(2));
            } else if (a == "--res") {
                s.resolution = std::stoi(next());
            } else if (a == "--dz") {
                s.slice_thickness = std::stod(next());
            } else if (a == "--kv") {
                s.voltage_kv = std::stod(next());
            } else if (a == "--area") {
                auto v = split_doubles(next());
                for (int k = 0; k < 4; ++k) s.area[k] = v.at(k);
                have_area = true;
            } else if (a == "--stem-pixels") {
                stem_px = std::stoi(next());
            } else if (a == "--det") {
                std::string d = next();
                auto c = d.find(',');
                Detector det;
                det.name = d.substr(0, c);
                auto v = split_doubles(d.substr(c + 1));
                det.d = mse_detector{v.at(0), v.at(1), v.size() > 2 ? v[2] : 0.0, v.size() > 3 ? v[3] : 0.0};
                sm->detectors.push_back(det);
            } else if (a == "--alpha") {
                sm->probe.cond_ap_mrad = std::stod(next());
            } else if (a == "--batch") {
                sm->parallel_pixels = std::stoi(next());
            } else if (a == "--moving-frame") {  // the settings' "static area enabled": false
                sm->parallel_stem = false;
            } else if (a == "--phonons") {
                sm->phonon_configs = std::stoi(next());
            } else if (a == "--urms") {
                sm->u_rms = std::stod(next());
            } else if (a == "--seed") {
                sm->seed = std::stoull(next());
            } else if (a == "--tilt") {
                auto v = split_doubles(next());
                s.beam_tilt_mrad = v.at(0);
                s.beam_azimuth = v.size() > 1 ? v[1] : 0.0;
            } else if (a == "--full3d") {
                s.full3d = 1;
                s.full3d_integrals = std::stoi(next());
            } else if (a == "--chromatic") {
                auto v = split_doubles(next());
                sm->chromatic.enabled = true;
                sm->chromatic.cc_mm = v.at(0);
                sm->chromatic.setHalfWidthHalfMaxs(v.at(1), v.at(2));
            } else if (a == "--plasmon") {
                auto v = split_doubles(next());
                sm->plasmon.enabled = true;
                sm->plasmon.mean_free_path = v.at(0) * 10.0;  // nm -> Angstrom (jsonutils.cpp:350)
                sm->plasmon.characteristic_angle = v.at(1);
                sm->plasmon.critical_angle = v.at(2);
                sm->plasmon.individual = v.size() > 3 ? (int)v[3] : 0;
            } else if (a == "--intermediate") {
                sm->slice_step = std::stoi(next());
            } else if (a == "--parallel-potentials") {
                sm->parallel_potentials = std::stoi(next());
            } else if (a == "--dump-settings") {
                dump_settings = next();
            } else if (a == "--dump-mixed-seed") {
                dump_mixed = true;
            } else if (a == "--dump-plasmons") {
                dump_plasmons = std::stod(next());
            } else if (a == "--source-size") {
                sm->source.enabled = true;
                sm->source.setFullWidthHalfMax(std::stod(next()));
            } else if (a == "--device") {
                sm->device = std::stoi(next());
            } else if (a == "--out") {
                out = next();
            } else if (a == "--dump-atoms") {
                dump = next();
            } else if (a == "--write-test-tiff") {
                const std::string fn = next();
                const int w = std::stoi(next()), h = std::stoi(next());
                std::vector<double> d((size_t)w * h);
                for (int y = 0; y < h; ++y)
                    for (int x = 0; x < w; ++x) d[(size_t)y * w + x] = x + 1000.0 * y;
                save_tiff(fn, d, w, h);
                return 0;
            } else if (a == "--cif") {
                cif_path = next();
                sm->structure_file = cif_path;
            } else if (a == "--zone" || a == "--normal" || a == "--widths" || a == "--tilts") {
                auto v = split_doubles(next());
                double* dst = a == "--zone" ? zone : a == "--normal" ? normal : a == "--widths" ? widths : tilts;
                for (int k = 0; k < 3; ++k) dst[k] = v.at(k);
            } else if (a == "--tiff") {
                tiff = true;
            } else if (a == "--dump-incoherence") {
                dump_incoherence = std::stoi(next());
            } else if (a == "--dump-config") {
                dump_config = std::stoi(next());
            } else {
                throw std::runtime_error("unknown option " + a);
            }
        }
        if (!cif_path.empty()) {
            if (widths[0] <= 0 || widths[1] <= 0 || widths[2] <= 0)
                throw std::runtime_error("--cif needs --widths x,y,z (Angstrom)");
            sm->atoms = cif_supercell(read_cif(cif_path), zone, normal, widths, tilts, sm->seed);
        }
        if (dump_mixed) {
            std::cout << sm->mixedSeed(dump_config) << std::endl;
            return 0;
        }
        if (dump_plasmons >= 0.0) {
            std::cout.precision(17);
            const std::vector<double> d = sm->plasmon.depths(sm->seed, dump_config, dump_plasmons);
            std::cout << d.size() << "\n";
            for (double v : d) std::cout << v << "\n";
            for (size_t i = 0; i < d.size(); ++i) {
                double polar, azimuth;
                sm->plasmon.angles(sm->seed, dump_config, (int)i, &polar, &azimuth);
                std::cout << polar << "\n" << azimuth << "\n";
            }
            std::cout << std::flush;
            return 0;
        }
        if (dump_incoherence >= 0) {
            std::cout.precision(17);
            for (int i = 0; i < dump_incoherence; ++i) std::cout << sm->focusChange(dump_config, i) << "\n";
            double dx, dy;
            sm->sourceOffset(dump_config, &dx, &dy);
            std::cout << dx << "\n" << dy << std::endl;
            return 0;
        }
        if (sm->atoms.size() == 0) throw std::runtime_error("no structure: use --xyz or --srtio3");
        if (!dump.empty()) {
            const AtomList a = displaced(sm->atoms, sm->u_rms, sm->seed, dump_config);
            std::ofstream f(dump, std::ios::binary);
            for (const auto* v : {&a.x, &a.y, &a.z}) f.write((const char*)v->data(), v->size() * sizeof(double));
            f.write((const char*)a.Z.data(), a.Z.size() * sizeof(int));
            return 0;
        }
        if (sm->params_file.empty()) throw std::runtime_error("no parameterisation: use --params");
        double lo[3], hi[3];
        sm->atoms.limits(lo, hi);
        for (int k = 0; k < 3; ++k) {
            s.struct_min[k] = lo[k];
            s.struct_max[k] = hi[k];
        }
        if (!have_area) {
            s.area[0] = lo[0];
            s.area[1] = hi[0];
            s.area[2] = lo[1];
            s.area[3] = hi[1];
        }
        if (s.mode == 2) {
            s.stem_pixels[0] = s.stem_pixels[1] = stem_px;
            if (sm->detectors.empty()) sm->detectors.push_back({"ADF", mse_detector{60.0, 200.0, 0.0, 0.0}});
        }
        if (!dump_settings.empty()) {
            std::cout << settings_json(*sm, dump_settings) << std::endl;
            return 0;
        }
        run_simulation(sm);
        std::ofstream js(out + ".json");
        js << "{";
        bool first = true;
        for (auto& kv : sm->averaged()) {
            const Image& im = kv.second;
            const std::string fn = out + "_" + kv.first + ".bin";
            std::ofstream f(fn, std::ios::binary);
            for (auto& sl : im.slices) f.write((const char*)sl.data(), sl.size() * sizeof(double));
            const size_t per = im.weighting.empty() ? 1 : im.slices[0].size() / im.weighting.size();
            js << (first ? "" : ", ") << "\"" << kv.first << "\": {\"file\": \"" << fn << "\", \"width\": "
               << im.width << ", \"height\": " << im.height << ", \"depth\": " << im.slices.size()
               << ", \"values_per_pixel\": " << per << "}";
            first = false;
        }
        js << ", \"jobs\": " << sm->complete_jobs << "}\n";
        if (tiff) save_outputs(*sm, out);
        std::cout << "mse_run: " << sm->complete_jobs << " jobs done, results in " << out << ".json" << std::endl;
        return 0;
    } catch (const std::exception& ex) {
        std::cerr << "mse_run: " << ex.what() << std::endl;
        return 1;
    }
}
