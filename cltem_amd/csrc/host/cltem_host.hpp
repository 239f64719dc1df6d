// Host side of the engine in the shape of clTEM's simulation classes.
//
// clTEM's CTEM, CBED and STEM modes are written against SimulationGeneral<T>'s protected
// methods (microscope/simulationgeneral.h:53-71), SimulationCbed::initialiseProbeWave
// (microscope/simulationcbed.h:33) and SimulationStem::getStemPixel
// (microscope/simulationstem.h:53-56), and merge results with
// SimulationManager::updateImages (simulationmanager.cpp:274-350).  These classes keep those
// names, argument meanings and the exception-on-failure behaviour of clError::Throw
// (clwrapper/clerror.h:17-24), and implement them on the C ABI of include/cltem_amd.h.
#pragma once

#include <array>
#include <climits>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/cltem_amd.h"

namespace cltem_amd {

// A failing engine call throws std::runtime_error with the engine's message.
void check(int rc);

struct AtomList {
    std::vector<double> x, y, z;
    std::vector<int> Z;
    size_t size() const { return x.size(); }
    void limits(double* lo, double* hi) const;
};

// Element symbol <-> atomic number (1..103), as Utils::ElementSymbolToNumber.
int element_number(const std::string& symbol);

// .xyz reader in clTEM's default-header form (structure/crystalstructure.cpp:44-188):
// line 1 atom count, line 2 header/comment, then "A x y z" per atom; coordinates pass
// through float like std::stof; a header containing "nm" scales by 10.
AtomList read_xyz(const std::string& path);

// A crystal read from a CIF file: cell geometry (Angstrom, degrees), symmetry operations as
// rows (x, y, z coefficients, translation), and the sites with their symmetry copies in the
// unit cell (fractional).
struct CifStructure {
    double a = 0, b = 0, c = 0, alpha = 0, beta = 0, gamma = 0;
    struct Op { std::array<double, 4> row[3]; };
    std::vector<Op> ops;
    struct Site {
        std::string label;
        std::vector<int> Z;        // several elements may share a site
        std::vector<double> occ;
        double fx, fy, fz;
        std::vector<std::array<double, 3>> positions;
    };
    std::vector<Site> sites;
};
CifStructure read_cif(const std::string& path);
// Supercell of widths[3] (Angstrom) with zone axis [uvw] along z, optional [abc] along x and
// tilts (degrees) about x, y, z; partial occupancies resolved per site with a seeded draw.
// `occupancy` (optional) receives the occupancy of every atom before that selection.
AtomList cif_supercell(const CifStructure& cif, const double uvw[3], const double abc[3], const double widths[3],
                       const double tilts_deg[3], uint64_t seed, std::vector<double>* occupancy = nullptr);

// SrTiO3 perovskite tiled nx*ny*nz along [001] (a = 3.905 A): Sr(0,0,0), Ti(1/2,1/2,1/2),
// O(1/2,1/2,0), O(1/2,0,1/2), O(0,1/2,1/2); coordinates float32-exact.
AtomList srtio3(int nx, int ny, int nz, double a = 3.905);

// Philox4x32-10 frozen-phonon displacement: each axis + u_rms * N(0,1), counter
// (atom, config, 0, 0), key (seed lo, seed hi), Box-Muller; results rounded to float32.
AtomList displaced(const AtomList& a, double u_rms, uint64_t seed, int config);
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

// Image<double> with per-pixel weighting (utilities/commonstructs.h:31-199).
struct Image {
    int width = 0, height = 0;
    std::vector<std::vector<double>> slices;
    std::vector<double> weighting;
    Image() = default;
    Image(int w, int h, int depth) : width(w), height(h), slices(depth, std::vector<double>((size_t)w * h, 0.0)),
                                     weighting((size_t)w * h, 0.0) {}
};

struct Detector {
    std::string name;
    mse_detector d;
};

// Incoherent effects drawn once per job (incoherence/chromaticaberration.cpp,
// incoherence/probesourcesize.cpp), from Philox4x32-10 keyed by the manager's seed instead of
// clTEM's clock-seeded generators, so a run can be repeated.
struct ChromaticSpread {
    bool enabled = false;
    double cc_mm = 0.0;                      // chromatic aberration coefficient
    double sigma_neg = 0.0, sigma_pos = 0.0; // energy spread (eV) below / above the mean
    // setHalfWidthHalfMaxs (chromaticaberration.cpp:108-120): sigma = 2 hwhm / (2 sqrt(2 ln 2))
    void setHalfWidthHalfMaxs(double hwhm_neg, double hwhm_pos);
};
struct SourceSize {
    bool enabled = false;
    double sigma = 0.0;                      // Angstrom
    void setFullWidthHalfMax(double fwhm);   // probesourcesize.cpp:17-19
};

// PlasmonScattering (incoherence/inelastic/plasmon.h:25-100) with Philox draws keyed by the
// manager's seed in place of its clock-seeded mt19937: counter (job, 0x9A + 256 attempt, k, 4)
// for the k-th free path of an attempt, (job, 0xA5, n, 5) for the angles of scattering n.
struct PlasmonScattering {
    bool enabled = false;
    int individual = 0;              // 0: PlasmonType::Full; n > 0: exactly n scatterings
    double mean_free_path = 100.0;   // Angstrom (clTEM's settings hold nm, times 10)
    double characteristic_angle = 1.0, critical_angle = 0.1;  // mrad
    // generateScatteringDepths (plasmon.cpp:37-127): depths below the specimen top
    std::vector<double> depths(uint64_t seed, int job, double thickness) const;
    // getScatteringPolar (mrad) and getScatteringAzimuth of scattering n (plasmon.cpp:14-29)
    void angles(uint64_t seed, int job, int n, double* polar_mrad, double* azimuth) const;
};

// Utils::rotateVectorSpherical (utilities/vectorutils.cpp:8-31)
void rotate_vector_spherical(double z[3], double y[3], double theta, double phi);

// The inputs clTEM's SimulationManager holds for this path, and its result merging.
struct SimulationManager {
    mse_scenario scenario{};   // mode, resolution, voltage, slicing, padding, areas, tilt
    std::string params_file;   // params/*.dat
    std::string structure_file;  // CrystalStructure::fileName (the settings' "filename")
    mse_probe probe{};         // STEM/CBED condenser and aberrations
    std::vector<Detector> detectors;
    AtomList atoms;            // undisplaced structure
    int parallel_pixels = 64;  // probes per batch
    // STEM "static area" (parallelStem, simulationmanager.h:194-200): off, every pixel runs in
    // a field of its own (getPixelSimArea) with its transmission functions rebuilt for it
    bool parallel_stem = true;
    int phonon_configs = 1;    // frozen-phonon iterations (0 or 1 = no phonons)
    double u_rms = 0.0;
    uint64_t seed = 20261202;
    bool sqrt2 = true;         // sum_reduction's sqrt(2) factor
    int device = 0;
    ChromaticSpread chromatic;
    SourceSize source;
    PlasmonScattering plasmon;
    // mixed static potentials (setParallelPotentialsCount): with phonons and >= 2, jobs share
    // this many phonon configurations and draw one per probe and slice
    int parallel_potentials = 0;
    // intermediateSliceStep: outputs after every slice_step slices and at the exit (0: exit only)
    int slice_step = 0;
    // intermediate outputs of a run: ceil(slices / slice_step), or 1
    int outputCount(int n_slices) const { return slice_step > 0 ? (n_slices + slice_step - 1) / slice_step : 1; }

    // ChromaticAberration::getFocusChange for probe `probe` of job `job` (Angstrom): an
    // asymmetric Gaussian in energy, 10000 * Cc[mm] * dE / kV (chromaticaberration.cpp:43-58, 99-106)
    double focusChange(int job, int probe) const;
    // ProbeSourceSize::getOffset for x and y of job `job` (Angstrom)
    void sourceOffset(int job, double* dx, double* dy) const;
    // parallelPotentialsCount (simulationmanager.h:218-231): the number of shared sets, or 0
    int mixedSets() const;
    // seed of the set draws of job `job` (batch `batch` of its probes when they are set up
    // one batch at a time): Philox4x32-10 words 0, 1 at (job, 0x3B, batch, 7)
    uint64_t mixedSeed(int job, int batch = 0) const;

    // SimulationManager::updateImages: FP64 sum of every job's images and weights
    void updateImages(std::map<std::string, Image>& ims, int jobCount);
    // result of a finished simulation: each image divided by its weighting
    std::map<std::string, Image> averaged() const;
    std::map<std::string, Image> images;
    int complete_jobs = 0;

    mse_geometry geometry() const;

   private:
    std::mutex image_update_mutex;
};

class SimulationGeneral {
   public:
    explicit SimulationGeneral(int device);
    virtual ~SimulationGeneral();
    SimulationGeneral(const SimulationGeneral&) = delete;
    SimulationGeneral& operator=(const SimulationGeneral&) = delete;

    // one job: a frozen-phonon configuration and (STEM) the raster pixels to simulate
    void setJob(std::shared_ptr<SimulationManager> sm, int config, std::vector<int> pixels = {});
    virtual void simulate() = 0;
    mse_engine* engine() { return eng; }

   protected:
    void sortAtoms();
    bool initialiseSimulation();
    void doMultiSliceStep(int slice);
    // the slice loop of simulate() (simulationcbed.cpp:142-199): with plasmons, after the slice
    // whose end reaches the next scattering depth the beam vector is rotated and the job's
    // propagator and remaining transmission functions follow it (modifyBeamTilt).  `job` keys
    // the scattering draws.  `output(k)` is called for intermediate output k after every
    // slice_step slices (after that slice's scattering) and, if the last one was not at the
    // exit surface, once more at the end.  Returns the number of scatterings applied.
    int runSlices(int job, const std::function<void(int)>& output);
    void resetBeam();  // the geometry's beam again, with its transmission functions rebuilt
    std::vector<double> getDiffractionImage(int parallel_ind);
    std::vector<double> getExitWaveImage(unsigned t = 0, unsigned l = 0, unsigned b = 0, unsigned r = 0);

    mse_engine* eng = nullptr;
    std::shared_ptr<SimulationManager> sm;
    int config = 0;
    double reference_perturb_x = 0.0, reference_perturb_y = 0.0;  // source-size offset of the job
    std::vector<int> pixels;
    mse_geometry geom{};
    int n_waves = 0;
    bool beam_moved = false;  // a plasmon scattering changed configuration 0's beam
    const SimulationManager* sets_for = nullptr;  // manager whose mixed potential sets are built
};

class SimulationCtem : public SimulationGeneral {
   public:
    using SimulationGeneral::SimulationGeneral;
    void simulate() override;  // simulationctem.cpp:285-400 without image formation
   protected:
    bool initialiseSimulation();
};

class SimulationCbed : public SimulationGeneral {
   public:
    using SimulationGeneral::SimulationGeneral;
    void simulate() override;  // simulationcbed.cpp:122-210
   protected:
    bool initialiseSimulation();  // draws the source-size offset (simulationcbed.cpp:102-118)
    void initialiseProbeWave(double posx, double posy, int n_parallel = 0);
};

class SimulationStem : public SimulationCbed {
   public:
    using SimulationCbed::SimulationCbed;
    void simulate() override;  // simulationstem.cpp:143-278, batched through mse_stem_batch
   protected:
    double getStemPixel(double inner, double outer, double xc, double yc, int parallel_ind);
};

// fileio::SaveTiff<float> (utilities/fileio.h:44-82): one uncompressed strip of 32-bit IEEE
// floats, one sample per pixel, min-is-black, orientation bottom-left, rows in data order.
void save_tiff(const std::string& path, const std::vector<double>& data, int width, int height);

// JSONUtils::BasicManagerToJson(man, false, true) (utilities/jsonutils.cpp:432-781) with the
// structure file name and the per-image edits of imageReturned (console/main.cpp:189-228),
// written as SaveSettingsJson does (keys sorted, 4-space indents).
std::string settings_json(const SimulationManager& sm, const std::string& image);

// The outputs of imageReturned (console/main.cpp:189-266): "EW" as <name>_amplitude and
// <name>_phase, every other image as <name>; each a .tif (one per slice of a stack, suffixed
// _<slice>) and a .json with the run's settings.
void save_outputs(const SimulationManager& sm, const std::string& prefix);

// Runs every job of a configured manager on one GPU (the role of SimulationRunner's
// worker, threading/simulationrunner.cpp:52-131, for one device).
void run_simulation(std::shared_ptr<SimulationManager> sm);

}  // namespace cltem_amd
