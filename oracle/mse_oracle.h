/*
 * CPU restatement of clTEM's multislice path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the oracle that the HIP product is checked against.  It is
 * linked by nothing in the product: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker / CPU baseline.
 *
 * Every function restates one piece of the reference in double precision
 * (the `_d` kernels and the host code that drives them), op for op where the
 * result is decided by rounding (atom binning, masks), built with
 * -ffp-contract=off.  Each function cites the reference file:line it follows.
 *
 * Parity pins (what checks each function, see tests/test_oracle_*.py):
 *   orc_wavelength/orc_sigma ........ published closed forms (lambda(200kV), sigma)
 *   orc_geometry .................... hand-derived SrTiO3 numbers (L, px, slices, blocks)
 *   orc_sort_atoms .................. brute-force tiny cases, stability, edge clamps
 *   orc_bessk0/orc_bessk1 ........... scipy.special.k0/k1
 *   orc_potential_* ................. scipy radial tables, superposition, cutoffs
 *   orc_potential_slice (full 3D) ... sum over slices = scipy quad of V3D over the signed
 *                                     3 A window + Euler-Maclaurin term (1e-9); the value
 *                                     carried between atoms of a block row and the
 *                                     cumulative tilt shift by explicit point-value sums
 *   orc_fft2 ........................ numpy.fft.fft2(norm="ortho") golden vectors
 *   orc_fft2_fast ................... orc_fft2 and numpy on power-of-two sizes
 *   orc_propagator / band limit ..... exact masks and exp(-i pi lambda k^2 dz)
 *   orc_probe ....................... numpy chi(), aperture-disc pixel count
 *   orc_run_* ....................... empty specimen identities, numpy multislice
 *   orc_stem_pixel .................. centred and off-centre annuli on an empty specimen =
 *                                     sqrt(2) x a pixel set counted in numpy
 *   orc_ctem_image .................. aberration-free wide aperture = |psi|^2; two-beam
 *                                     closed form with defocus, C30, both envelopes, edge
 *   orc_ctem_intensity .............. orc_run_plane over orc_transmission of the sorted atoms
 *   orc_ctem_loss_fd ................ brute-force central differences (every slice rebuilt,
 *                                     perturbed atoms moved in numpy), O(h^2) convergence
 *   orc_translate_image ............. whole-pixel moves = numpy.roll; fractional moves = the
 *                                     linear blend of the two neighbouring rolls; sum kept
 *   orc_rotate_vector_spherical ..... MicroscopeParameters::Wavevector from (0, 0, k); lengths
 *                                     and the y . z = 0 frame kept; scipy Rotation composition
 */
#ifndef MSE_ORACLE_H
#define MSE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Inputs that clTEM's SimulationManager / SimulationCell hold
 * (simulationmanager.h, structure/simulationcell.h, utilities/commonstructs.h). */
typedef struct OrcScenario {
    int mode;                 /* 1 CTEM, 2 STEM, 3 CBED (utilities/enums.h:8-14) */
    int resolution;
    double voltage_kv;
    double slice_thickness;
    double slice_offset;
    double pad_xy[2];         /* SimulationCell default_xy_padding {-8, 8} */
    double pad_z[2];          /* SimulationCell default_z_padding {-3, 3} */
    double struct_min[3];     /* CrystalStructure limits */
    double struct_max[3];
    double area[4];           /* xStart, xFinish, yStart, yFinish of the mode's area */
    double area_padding;      /* SimulationArea::padding */
    int stem_pixels[2];
    double beam_tilt_mrad;
    double beam_azimuth;
    double limit_factor;      /* max_inverse_factor = 2/3 */
    int full3d;
    int full3d_integrals;
    int moving_stem_frame;    /* STEM: 1 = !SimulationManager::parallel_stem */
} OrcScenario;

typedef struct OrcGeometry {
    int resolution;
    double full_x[2], full_y[2];   /* paddedFullLimitsX/Y */
    double sim_x[2], sim_y[2];     /* paddedSimLimitsX/Y(0) */
    double z_lims[2];              /* paddedSimLimitsZ */
    double pad_z[2];               /* rounded z padding */
    double real_scale;
    double inverse_scale;
    double inverse_scale_angle;
    int n_slices;
    int pre_slices;
    int blocks_x, blocks_y;
    double block_scale_x, block_scale_y;
    int load_x, load_y, load_z;
    double slice_thickness;
    double wavelength, wavenumber, sigma, sigma_eff;
    double k_vec[3];
    double kmax_band;              /* bandwidthkmax */
    double limit_factor;
    double beam_tilt_mrad, beam_azimuth;
    double stem_scale[2];
    double stem_start[2];
    int stem_pixels[2];
    int full3d, full3d_integrals;
    int moving_frame;              /* sim_x/sim_y are the field of one STEM pixel */
} OrcGeometry;

double orc_wavelength(double kv);
double orc_sigma(double kv);
int orc_geometry(const OrcScenario* s, OrcGeometry* g);
/* the geometry of STEM pixel `pixel`'s job: with the moving frame, sim_x/sim_y are
 * paddedSimLimitsX/Y(pixel) (simulationmanager.cpp:434-440,485-501); otherwise orc_geometry */
int orc_geometry_pixel(const OrcScenario* s, int pixel, OrcGeometry* g);

/* sortAtoms: host filter, atom_sort_d, stable bucket fill.
 * out arrays need capacity n; block_start needs n_slices*bx*by+1. */
int orc_sort_atoms(const OrcGeometry* g, const double* x, const double* y, const double* z,
                   const int* Z, int n, double* sx, double* sy, double* sz, int* sZ,
                   int* block_ids, int* z_ids, int* block_start, int* n_filtered, int* n_in_range);

double orc_bessk0(double x);
double orc_bessk1(double x);
double orc_vproj(int form, const double* params, int i_count, int Z, double r);
double orc_v3d(int form, const double* params, int i_count, int Z, double r);

int orc_kgrid(const OrcGeometry* g, double* kx, double* ky);

/* raw potential sum (sumz) for one slice; out has N doubles */
int orc_potential_slice(const OrcGeometry* g, const double* sx, const double* sy, const double* sz,
                        const int* sZ, const int* block_start, const double* params, int form,
                        int i_count, int slice, double* out, int nthreads);
/* transmission for slices [first, first+count): out is count * 2N doubles (interleaved) */
int orc_transmission(const OrcGeometry* g, const double* sx, const double* sy, const double* sz,
                     const int* sZ, const int* block_start, const double* params, int form,
                     int i_count, int first, int count, int band_limit, double* out, int nthreads);

int orc_fft2(double* data, int w, int h, int inverse);
/* the CPU baseline's FFT (iterative Stockham, power-of-two w and h; -1 otherwise), the same
 * transform as orc_fft2 */
int orc_fft2_fast(double* data, int w, int h, int inverse);
int orc_band_limit(const OrcGeometry* g, double* t);
int orc_propagator(const OrcGeometry* g, double* P);
/* k-space probe (before the IFFT) for position (px, py) in Angstrom relative to
 * the padded simulation start; aberr = 25 doubles:
 *   C10, C12(re,im), C21(re,im), C23(re,im), C30, C32(re,im), C34(re,im),
 *   C41(re,im), C43(re,im), C45(re,im), C50, C52(re,im), C54(re,im), C56(re,im) */
int orc_probe_kspace(const OrcGeometry* g, double px, double py, const double* aberr,
                     double cond_ap_mrad, double smooth_mrad, int taper, double* out);

/* One doMultiSliceStep for one wave: recip = FFT(t psi) P ; psi = IFFT(recip) */
int orc_slice_step(int res, double* psi, double* recip, const double* t, const double* P);
double orc_stem_pixel(int res, const double* recip, double inner_mrad, double outer_mrad,
                      double xc_mrad, double yc_mrad, double angle_scale, int sqrt2);
int orc_diffraction(int res, const double* recip, double* out);
/* translateDiffImage (simulationgeneral.cpp:832-854) + bilinear_translate_d
 * (kernels/bilinear_translate_d.cl:21-72): an fft-shifted image moved by (dkx, dky) / inverse_scale
 * pixels (plasmon tilt), with one wrap-around */
int orc_translate_image(int res, const double* in, double* out, double dkx, double dky, double inverse_scale);
/* Utils::rotateVectorSpherical (utilities/vectorutils.cpp:8-31): z rotated by theta about y, then
 * z and y rotated by phi about the original z (Rodrigues matrices of the normalised axes) */
void orc_rotate_vector_spherical(double* z, double* y, double theta, double phi);

/* Whole-path drivers.  trans = n_slices * 2N (already band limited). */
int orc_run_plane(int res, int n_slices, const double* trans, const double* P,
                  double* psi_out, double* recip_out);
/* STEM: probes at pos[2*i] (Angstrom, absolute), detectors det[4*d] = inner, outer, xc, yc (mrad),
 * out[i*ndet + d] */
int orc_run_stem(const OrcGeometry* g, const double* trans, const double* P, const double* pos,
                 int n_probes, const double* aberr, double cond_ap_mrad, double smooth_mrad,
                 int taper, const double* det, int ndet, int sqrt2, double* out, int nthreads);
/* FP32 restatement of the same loop (the `_f` path), used as the CPU baseline. */
int orc_run_stem_f32(const OrcGeometry* g, const double* trans, const double* P, const double* pos,
                     int n_probes, const double* aberr, double cond_ap_mrad, double smooth_mrad,
                     int taper, const double* det, int ndet, int sqrt2, double* out, int nthreads);

/* CTEM image (ctem_image + IFFT + sqabs) of the unshifted Psi(k) recip (2N doubles);
 * lens = 25 aberrations (probe order), obj_ap_mrad, ap_smooth_mrad, beta_mrad, delta. */
int orc_ctem_image(const OrcGeometry* g, const double* recip, const double* lens, double* out);
/* CTEM plane wave (value 1) through the full path from unsorted atoms: exit intensity
 * |psi|^2, N doubles. */
int orc_ctem_intensity(const OrcGeometry* g, const double* x, const double* y, const double* z, const int* Z,
                       int n, const double* params, int form, int i_count, double* intensity, int nthreads);
/* L = sum_p (|psi_exit(p)|^2 - target(p))^2 of the same forward; grad[i] = central difference
 * (L(+h) - L(-h)) / 2h in coordinate axis[i] (0 x, 1 y, 2 z) of atom atom[i] (input order).
 * loss0 (optional) receives L at the given positions. */
int orc_ctem_loss_fd(const OrcGeometry* g, const double* x, const double* y, const double* z, const int* Z, int n,
                     const double* params, int form, int i_count, const double* target, const int* atom,
                     const int* axis, int m, double h, double* grad, double* loss0, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
