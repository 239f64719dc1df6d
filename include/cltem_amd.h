/*
 * cltem_amd -- MI355X-native multislice engine behind clTEM's simulation interfaces.
 *
 * Plain C ABI: opaque engine handle, int status returns (0 = success), a
 * thread-local error string, caller-owned host buffers passed as pointer + size,
 * inputs in double.  One engine owns one HIP stream on one GPU; engines are not
 * shared between threads (clTEM runs one worker thread per device,
 * threading/threadpool.cpp:9-18, and each worker owns its own context).
 *
 * Every entry point names the reference method or kernel whose contract it
 * replaces (paths relative to the clTEM source tree).
 */
#ifndef CLTEM_AMD_H
#define CLTEM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define MSE_OK 0
#define MSE_ERR_ARG -1
#define MSE_ERR_STATE -2
#define MSE_ERR_HIP -3
#define MSE_ERR_UNSUPPORTED -4

typedef struct mse_engine mse_engine;

/* ---- scenario geometry: SimulationManager / SimulationCell / MicroscopeParameters ---- */

/* Inputs held by SimulationManager and SimulationCell (simulationmanager.h,
 * structure/simulationcell.h, utilities/commonstructs.h:210-401). */
typedef struct mse_scenario {
    int mode;                /* 1 CTEM, 2 STEM, 3 CBED (utilities/enums.h:8-14) */
    int resolution;          /* 256, 384, 512, 768, 1024, 1536, 2048 */
    double voltage_kv;
    double slice_thickness;  /* SimulationCell::slice_thickness */
    double slice_offset;     /* SimulationCell::slice_offset */
    double pad_xy[2];        /* SimulationCell default_xy_padding, {-8, 8} */
    double pad_z[2];         /* SimulationCell default_z_padding, {-3, 3} */
    double struct_min[3];    /* CrystalStructure::limitsX/Y/Z lower */
    double struct_max[3];    /* CrystalStructure::limitsX/Y/Z upper */
    double area[4];          /* xStart, xFinish, yStart, yFinish of the mode's area */
    double area_padding;     /* SimulationArea::padding */
    int stem_pixels[2];      /* StemArea xPixels, yPixels (STEM only) */
    double beam_tilt_mrad;   /* MicroscopeParameters::BeamTilt */
    double beam_azimuth;     /* MicroscopeParameters::BeamAzimuth (rad) */
    double limit_factor;     /* SimulationManager::max_inverse_factor, 2/3 */
    int full3d;              /* SimulationManager::use_full_3d */
    int full3d_integrals;    /* SimulationManager::full_3d_integrals, 20 (1..4096 with full3d) */
    int moving_stem_frame;   /* STEM only: 1 = !SimulationManager::parallel_stem, every pixel in a
                                field of its own (StemArea::getPixelSimArea); 0 = clTEM's default */
} mse_scenario;

/* Derived quantities the reference computes on demand from the manager. */
typedef struct mse_geometry {
    int resolution;
    double full_x[2], full_y[2]; /* paddedFullLimitsX/Y (simulationmanager.cpp:446-452) */
    double sim_x[2], sim_y[2];   /* paddedSimLimitsX/Y(0) (simulationmanager.cpp:434-440) */
    double z_lims[2];            /* paddedSimLimitsZ (simulationmanager.cpp:442-444) */
    double pad_z[2];             /* SimulationCell::roundPaddingZ (simulationcell.cpp:43-61) */
    double real_scale;           /* realScale (simulationmanager.cpp:211-222) */
    double inverse_scale;        /* inverseScale (simulationmanager.cpp:224-230) */
    double inverse_scale_angle;  /* inverseScaleAngle, mrad/px (simulationmanager.cpp:241-247) */
    int n_slices;                /* SimulationCell::sliceCount (simulationcell.cpp:75-83) */
    int pre_slices;              /* SimulationCell::preSliceCount (simulationcell.cpp:85-87) */
    int blocks_x, blocks_y;      /* calculateBlocks (simulationmanager.cpp:393-408) */
    double block_scale_x, block_scale_y;
    int load_x, load_y, load_z;  /* simulationgeneral.cpp:517-521 */
    double slice_thickness;
    double wavelength, wavenumber, sigma, sigma_eff; /* commonstructs.h:269-289, simulationgeneral.cpp:418 */
    double k_vec[3];             /* MicroscopeParameters::Wavevector */
    double kmax_band;            /* bandwidthkmax (simulationgeneral.cpp:451-455) */
    double limit_factor;
    double beam_tilt_mrad, beam_azimuth;
    double stem_scale[2];        /* StemArea::getScaleX/Y (commonstructs.cpp:99-107) */
    double stem_start[2];        /* StemArea raw start */
    int stem_pixels[2];
    int full3d, full3d_integrals;
    /* moving STEM frame (simulationmanager.cpp:485-501): sim_x/sim_y above are pixel 0's field
     * and fix real_scale; pixel p's field starts at mse_stem_pixel_frame(g, p) */
    int moving_frame;
    double frame_step;           /* the full field's pixel size the frame origins snap to */
    double frame_half[2];        /* half the STEM area padding, each side of a frame */
    double pad_xy[2];            /* SimulationCell xy padding */
} mse_geometry;

/* Probe (condenser) parameters: SimulationCbed::initialiseProbeWave
 * (microscope/simulationcbed.cpp:35-100) -> init_probe_wave (kernels/init_probe_wave_f.cl:49-118).
 * Complex aberrations are ComplexAberration::getComplex() = polar(mag, ang) as (re, im). */
typedef struct mse_probe {
    double C10;
    double C12[2], C21[2], C23[2];
    double C30;
    double C32[2], C34[2];
    double C41[2], C43[2], C45[2];
    double C50;
    double C52[2], C54[2], C56[2];
    double cond_ap_mrad;     /* CondenserAperture */
    double smooth_mrad;      /* CondenserApertureSmoothing */
    int edge_taper;          /* 1: smoothstep edge of init_probe_wave_f; 0: init_probe_wave_d */
} mse_probe;

/* Objective-lens imaging of a CTEM exit wave: the MicroscopeParameters fields that
 * SimulationCtem::simulateImagePerfect passes to ctem_image (microscope/simulationctem.cpp:113-171).
 * Aberrations as in mse_probe; apertures and the illumination semi-angle in mrad. */
typedef struct mse_imaging {
    double C10;
    double C12[2], C21[2], C23[2];
    double C30;
    double C32[2], C34[2];
    double C41[2], C43[2], C45[2];
    double C50;
    double C52[2], C54[2], C56[2];
    double obj_ap_mrad;      /* ObjectiveAperture */
    double obj_smooth_mrad;  /* ObjectiveApertureSmoothing */
    double beta_mrad;        /* Alpha: illumination semi-angle (spatial coherence envelope) */
    double delta;            /* Delta: defocus spread, Angstrom (temporal coherence envelope) */
} mse_imaging;

/* StemDetector (utilities/commonstructs.h:292-300): radii and centre in mrad. */
typedef struct mse_detector {
    double inner_mrad, outer_mrad, xc_mrad, yc_mrad;
} mse_detector;

/* ---- lifecycle (replaces clContext creation, SimulationGeneral ctor/dtor,
 *      microscope/simulationgeneral.h:20-34) ---- */
int mse_create(int device_ordinal, mse_engine** out);
int mse_destroy(mse_engine* e);
/* thread-local message for the last failing call on this thread (clError::Throw text,
 * clwrapper/clerror.h:17-24) */
const char* mse_last_error(void);
/* ctx->WaitForQueueFinish() (clwrapper/clcontext.h) */
int mse_synchronize(mse_engine* e);
/* device name and compute-unit count, for reports */
int mse_device_info(mse_engine* e, char* name, int name_len, int* cu_count, double* hbm_gbs);

/* ---- geometry (host only; SimulationManager geometry methods, see mse_geometry) ---- */
int mse_plan_geometry(const mse_scenario* s, mse_geometry* g);
/* paddedSimLimitsX/Y(pixel)[0] of a moving-frame STEM geometry: the potential-grid origin of
 * STEM pixel `pixel` (SimulationManager::currentAreaBase, simulationmanager.cpp:485-501, and
 * StemArea::getPixelSimArea, utilities/commonstructs.h:379-396).  For any other geometry
 * every pixel shares sim_x[0], sim_y[0]. */
int mse_stem_pixel_frame(const mse_geometry* g, int pixel, double* start_xy);

/* ---- SimulationGeneral::initialiseSimulation (microscope/simulationgeneral.cpp:345-646):
 *      k-grid, band-limit and propagator tables (propagator kernel, kernels/propagator_f.cl:18-51) ---- */
int mse_set_geometry(mse_engine* e, const mse_geometry* g);
/* ClParameterisation upload (simulationgeneral.cpp:375-378); form 0 Kirkland, 1 Peng, 2 Lobato */
int mse_set_parameterisation(mse_engine* e, int form, int i_per_atom, const double* params, int n_params);

/* ---- SimulationGeneral::sortAtoms (microscope/simulationgeneral.cpp:147-342) with
 *      atom_sort_d (kernels/atom_sort_d.cl:28-69), binned in IEEE double on the GPU.
 *      x/y/z are already-displaced coordinates of one frozen-phonon configuration.
 *      Optional outputs (may be NULL): block_ids/z_ids per input atom (-1 when the host
 *      filter drops it), block_start (n_slices*bx*by + 1 entries), n_in_range. ---- */
int mse_sort_atoms(mse_engine* e, int config, const double* x, const double* y, const double* z,
                   const int* Z, int n, int* block_ids, int* z_ids, int* block_start, int* n_in_range);

/* ---- transmission functions for slices [first, first+count) of a configuration:
 *      transmission_potentials_projected / _full_3d (kernels/transmission_potentials_*_f.cl)
 *      plus the FFT -> band_limit -> IFFT of simulationgeneral.cpp:576-613 ---- */
int mse_build_transmission(mse_engine* e, int config, int first, int count);
/* read one transmission function back as interleaved float (2N) */
int mse_read_transmission(mse_engine* e, int config, int slice, float* out);

/* ---- a configuration with its own slicing: a specimen of another thickness sharing the
 *      engine's grid, field, blocks, slice thickness and optics (only z_lims, pad_z, n_slices,
 *      pre_slices may differ; n_slices may not exceed the engine geometry's).  Its slices
 *      beyond n_slices are vacuum (t = 1) in the slice loop.  Replaces re-initialising
 *      SimulationGeneral for every specimen (simulationgeneral.cpp:345-391). ---- */
int mse_set_config_geometry(mse_engine* e, int config, const mse_geometry* g);

/* ---- plasmon beam tilt: SimulationGeneral::modifyBeamTilt (microscope/simulationgeneral.cpp:649-679)
 *      as a per-configuration state.  k_vec (1/Angstrom, |k_vec| = the wavenumber, kz > 0) becomes
 *      the configuration's beam vector; NULL restores the geometry's.  Transmission functions
 *      built afterwards use its tilt and azimuth (projected: beam_theta/beam_phi of
 *      transmission_potentials_projected; full 3D: the per-integral shifts (kx, ky)/kz * dz/integrals),
 *      and every wave of the configuration propagates with propagator_d at k_vec.  Diffraction
 *      images, their sums and STEM detector values of such waves are bilinearly translated by
 *      (k_vec - geometry k_vec) / inverse_scale pixels (translateDiffImage, simulationgeneral.cpp:832-854,
 *      kernels/bilinear_translate_d.cl).  Rebuild the configuration's transmission functions
 *      for the slices still to come (mse_build_transmission).  Not reproduced: clTEM passes the
 *      projected tilt as a 30th argument of a 29-argument kernel, and its full-3D y shift uses kx.
 *      mse_stem_batch refuses such a configuration (use mse_multislice + mse_stem_detect). ---- */
int mse_set_config_beam(mse_engine* e, int config, const double* k_vec);

/* ---- mixed static potentials: doMultiSliceStep with useParallelPotentials
 *      (microscope/simulationgeneral.cpp:731-745; the sets are built per phonon configuration
 *      at simulationgeneral.cpp:576-612).  With n_sets >= 2, at every slice every wave
 *      multiplies the transmission function of configuration first_config + floor(n_sets r / 2^32),
 *      r the first word of Philox4x32-10 keyed by seed at counter (wave key, slice, 0x3A, 6);
 *      a wave's key is the probe index of its last probe set-up (mse_init_probes*: the index
 *      into pos_xy; mse_stem_batch: the probe's index over all batches), else its wave index.
 *      clTEM draws the set from a clock-seeded generator; this draw is reproducible.  The sets
 *      must use the engine's geometry and beam.  n_sets 0 or 1: each wave uses its own
 *      configuration again.  Position gradients refuse mixed sets. ---- */
int mse_set_mixed_potentials(mse_engine* e, int first_config, int n_sets, unsigned long long seed);

/* ---- waves: clWaveFunctionReal/Recip[parallel_pixels] (simulationgeneral.h:85-86) ---- */
int mse_set_wave_count(mse_engine* e, int n_waves);
/* wave -> frozen-phonon configuration whose transmission functions it uses (default 0) */
int mse_set_wave_configs(mse_engine* e, const int* configs, int n_waves);
/* init_plane_wave (kernels/init_plane_wave_f.cl:12-25; simulationctem.cpp:77-82) */
int mse_init_plane_wave(mse_engine* e, double value);
/* SimulationCbed::initialiseProbeWave for waves 0..n-1 at absolute positions pos_xy[2*i]
 * (Angstrom), made relative to the padded simulation start (simulationcbed.cpp:54-59) */
int mse_init_probes(mse_engine* e, const double* pos_xy, int n, const mse_probe* probe);
/* Chromatic focus spread (ChromaticAberration::getFocusChange, incoherence/chromaticaberration.cpp,
 * added to C10 in SimulationCbed::initialiseProbeWave, microscope/simulationcbed.cpp:62-71):
 * delta_c10[i] (Angstrom) is added to C10 for the i-th probe position of each later
 * mse_init_probes / mse_init_probes_range / mse_stem_batch call, so every probe of a batch
 * can carry its own focus.  n = 0 clears the list. */
int mse_set_focus_spread(mse_engine* e, const double* delta_c10, int n);
/* Source-size offset (SimulationCbed::initialiseSimulation drawing reference_perturb_x/y,
 * microscope/simulationcbed.cpp:104-118, applied to the potential grid origin at
 * simulationgeneral.cpp:559-560): the transmission functions of configuration cfg built after
 * this call sample the potential on a grid shifted by (dx, dy) Angstrom. */
int mse_set_reference_offset(mse_engine* e, int cfg, double dx, double dy);
/* the same for waves first_wave..first_wave+n-1 (probes of different convergence per wave) */
int mse_init_probes_range(mse_engine* e, int first_wave, const double* pos_xy, int n, const mse_probe* probe);
/* SimulationGeneral::doMultiSliceStep for slices [first, first+count) on every wave
 * (microscope/simulationgeneral.cpp:682-769), batched over the waves */
int mse_multislice(mse_engine* e, int first, int count);
/* mse_multislice replayed from a captured hipGraph: the whole slice range for every wave
 * (probes, phonon configurations, specimens of other thicknesses) is one graph launch.
 * Captured on first use and again whenever the wave set, slice range, kernel variants or
 * engine buffers change; results are identical to mse_multislice. */
int mse_multislice_graph(mse_engine* e, int first, int count);
/* ---- mixed grid sizes in one launch.  No clTEM counterpart (clTEM runs one simulation
 *      per grid).  A group holds engines of one device, typically one per grid size;
 *      mse_group_multislice runs every engine's whole slice loop (all its slices and waves,
 *      as mse_multislice) from one captured hipGraph launched once on the first engine's
 *      stream: each engine's kernels form a branch of the graph and the branches run
 *      concurrently.  The launch waits for the work already queued on every engine, and
 *      every engine's later calls wait for it.  Captured again when any engine's wave set,
 *      slice plan, kernel variants or buffers change.  Destroy a group before its engines. */
typedef struct mse_group mse_group;
int mse_group_create(mse_engine* const* engines, int n, mse_group** out);
int mse_group_multislice(mse_group* g);
int mse_group_destroy(mse_group* g);
/* SimulationStem::getStemPixel for every wave and detector (microscope/simulationstem.cpp:42-131):
 * out[wave*n_det + d] = (sqrt2 ? sqrt(2) : 1) * sum over the annulus of |Psi_shifted|^2, FP64 */
int mse_stem_detect(mse_engine* e, const mse_detector* dets, int n_det, int sqrt2, double* out);
/* mse_stem_detect without any allocation.  No clTEM counterpart: it is for callers that
 * cannot allocate device memory (code under stream capture, serving loops, harnesses that
 * forbid hipMalloc).  out_dev is caller-owned DEVICE memory of n_waves*n_det doubles, written
 * with the same values as mse_stem_detect; the call only queues work on the engine stream
 * (mse_synchronize before reading out_dev on another stream).  The detector bitmasks and the
 * per-tile partial sums live in a fixed 4 MiB device workspace per GPU, shared by every engine
 * on that GPU: n_det*res^2/8 + n_waves*n_det*res bytes must fit (MSE_ERR_ARG otherwise),
 * calls that overlap on one GPU race, and each call rewrites what it reads.
 * MSE_ERR_UNSUPPORTED when a wave's configuration has a beam of its own (plasmon tilt). */
int mse_stem_detect_noalloc(mse_engine* e, const mse_detector* dets, int n_det, int sqrt2, double* out_dev);
/* getDiffractionImage (simulationgeneral.cpp:772-805): |fftshift Psi(k)|^2, N doubles */
int mse_diffraction(mse_engine* e, int wave, double* out);
/* getExitWaveImage (simulationgeneral.cpp:808-829) without cropping: interleaved re/im, 2N doubles */
int mse_exit_wave(mse_engine* e, int wave, double* out);
/* CTEM image of every wave, summed over the waves into N doubles (out += when accumulate):
 * ctem_image (kernels/ctem_image_d.cl: objective aperture with smoothed edge, aberration
 * phase exp(-i chi), temporal and spatial coherence envelopes) applied to Psi(k), the inverse
 * FFT and sqabs (SimulationCtem::simulateImagePerfect, microscope/simulationctem.cpp:113-171),
 * fused into one column pass and one row pass; the exit waves are left unchanged.  With one
 * wave per frozen-phonon configuration this is the FP64 image sum of updateImages. */
int mse_ctem_image(mse_engine* e, const mse_imaging* im, int accumulate, double* out);
/* CCD response and dose of a CTEM image (SimulationCtem::simulateImageDose,
 * microscope/simulationctem.cpp:174-263): image (N doubles, e.g. from mse_ctem_image) ->
 * FFT, x sqrt(DQE), IFFT (ccd_dqe, kernels/ccd_dqe_d.cl) -> conversion * Poisson(dose_per_pixel
 * * binning^2 * real part) -> FFT, x NTF, IFFT (ccd_ntf, kernels/ccd_ntf_d.cl) -> real part
 * into out.  dqe/ntf are n_table radial samples (the files of clTEM's ccds folder hold 725) at radius /
 * binning in pixels of the unshifted FFT, linearly interpolated, clamped to the last entry.
 * The noise comes from Philox4x32-10 keyed by seed (counter = pixel), so a seed reproduces
 * an image exactly; clTEM's generator is clock-seeded. */
int mse_ccd_image(mse_engine* e, const double* image, const double* dqe, const double* ntf, int n_table,
                  int binning, double dose_per_pixel, double conversion, unsigned long long seed, double* out);
/* sum over all waves of |psi_exit|^2 into N doubles (out += when accumulate): the phonon
 * average of the exit-wave intensity as SimulationManager::updateImages sums jobs in FP64
 * (simulationmanager.cpp:274-350) */
int mse_exit_intensity_sum(mse_engine* e, int accumulate, double* out);
/* out[g*N + p] (+)= sum over waves w with wave_group[w] == g of |fftshift Psi_w(k)|^2 (FP64):
 * CBED diffraction patterns accumulated over frozen-phonon configurations */
int mse_diffraction_sum(mse_engine* e, const int* wave_group, int n_groups, int accumulate, double* out);

/* ---- batched STEM hot path: SimulationStem::simulate (microscope/simulationstem.cpp:143-278)
 *      for n_probes positions of one configuration, in batches of `batch` waves, with the
 *      detector reduction fused into the last slice.  accum[p*n_det + d] += intensity
 *      (FP64, so repeated calls over configurations accumulate like updateImages,
 *      simulationmanager.cpp:274-350).  use_graph=1 replays a captured hipGraph per batch.
 *      Afterwards every wave refers to `config` and waves 0..batch-1 hold the last batch. ---- */
int mse_stem_batch(mse_engine* e, int config, const double* pos_xy, int n_probes, int batch,
                   const mse_probe* probe, const mse_detector* dets, int n_det, int sqrt2,
                   int use_graph, double* accum);
/* The same with the moving STEM frame (parallel_stem off): probe p runs in a field of its own
 * whose origin is frame_xy[2p], frame_xy[2p+1] (mse_stem_pixel_frame of its pixel); its
 * position is taken relative to that origin (simulationcbed.cpp:54-59) and its potential grid
 * starts there plus the configuration's reference offset (mse_set_reference_offset).  clTEM
 * regenerates every transmission function for each pixel job (initialiseSimulation,
 * microscope/simulationgeneral.cpp:345-358); here, slice by slice, the transmission functions
 * of the batch's frames are built from the sorted atoms of `config` (its transmission
 * functions need not be built), band limited, and each multiplies its own wave.  The engine
 * geometry (grid, field size, optics) is the moving-frame geometry.  Device memory: one
 * transmission plane per wave of the batch.  Eager launches only. */
int mse_stem_batch_frames(mse_engine* e, int config, const double* pos_xy, const double* frame_xy,
                          int n_probes, int batch, const mse_probe* probe, const mse_detector* dets,
                          int n_det, int sqrt2, double* accum);

/* ---- potential evaluation.  The potential kernels read each element's potential from a
 *      table of cubic pieces in r^2 (32 per octave of r^2 from 2^-20 to 2^8 A^2, indexed by the
 *      float bits of r^2), filled on the host from the reference formulas in double
 *      (kernels/transmission_potentials_projected_d.cl:71-181,
 *      kernels/transmission_potentials_full_3d_d.cl:57-134).  Geometries whose distances leave
 *      that range (pixels below 0.004 A, full-3D depth reach above 13.8 A) evaluate the
 *      formulas directly.  The full-3D kernels of an untilted beam copy the buckets a slice
 *      reaches into LDS when they take at most 32 KB; enable = 2 keeps every read in global
 *      memory (same values; for tests), enable = 0 always evaluates the formulas directly.
 *      The next mse_build_transmission uses the choice.  No clTEM counterpart. ---- */
int mse_set_potential_tables(mse_engine* e, int enable);
/* host only: the table of element Z (1-based) as 4 floats (c0..c3 of V = ((c3 t + c2) t + c1) t
 * + c0) per bucket into out[4 * 896], for form 0/1/2 (Kirkland, Peng, Lobato), full3d 0/1 */
int mse_potential_table(int form, int full3d, const double* params, int n_params, int i_per_atom, int Z,
                        float* out);

/* ---- tuning: kernel variants of the slice sweeps.  kernel 0 = row pass (variants 0..6):
 *      0 wave-level LDS sync, 1 half lanes per line, 2 block barriers, 3 twiddles read at
 *      use, 4 persistent with the next row prefetched (default), 5 as 4 with half lanes,
 *      6 as 4 with twiddles read at use.
 *      kernel 1 = column pass (variants 0..13): 0 16-column tiles, 1 half lanes, 2 8-column
 *      tiles, 3 block barriers, 4/5 twiddles read at use (16/8 columns), 6 32-column tiles,
 *      7 persistent with the next tile prefetched (default), 8 as 7 with half lanes, 9 as 7
 *      with twiddles read at use, 10 as 7 with 8-column tiles, 11 one tile per block at two
 *      blocks per CU, twiddles and propagator rows from LDS (16-column tiles at 512 points,
 *      8-column tiles for the four-pass 768-point plan; every other size, 384 included,
 *      runs variant 7), 12 / 13 as 10 / 7 with the persistent blocks of each XCD taking
 *      consecutive tiles, so that 8-column tiles sharing 128-byte lines meet in one L2.
 *      mse_autotune times every variant of both sweeps on n_waves waves of `config`
 *      (reps launches each; times in microseconds per launch) and selects the fastest: sweeps
 *      over that number of waves use it from then on, sweeps over any other number of waves the
 *      latest choice.  It leaves the waves' contents undefined.  mse_set_kernel_variant
 *      forces a variant for every number of waves. ---- */
int mse_set_kernel_variant(mse_engine* e, int kernel, int variant);
int mse_autotune(mse_engine* e, int config, int n_waves, int reps, double* row_us, double* col_us,
                 int* best_row, int* best_col);

/* ---- adjoint (position gradients).  No clTEM counterpart: clTEM only runs forward.  The
 *      forward differentiated is doMultiSliceStep over all slices
 *      (microscope/simulationgeneral.cpp:682-769) from the entrance waves, with the
 *      projected or full-3D transmission functions of initialiseSimulation
 *      (simulationgeneral.cpp:345-646).  ---- */
/* From the current entrance waves (all referring to configuration cfg), run every slice of
 * cfg, keeping ceil(n/K) checkpoints of the waves: K = 1 (every state, so that the adjoint
 * recomputes nothing) when the n_slices states of all waves take at most the checkpoint
 * budget, else K = ceil(sqrt(n_slices)).  Leaves the exit waves in the engine like
 * mse_multislice; when intensity is not NULL writes |psi_exit|^2 as n_waves * N doubles
 * (wave-major, rows of res). */
int mse_forward_checkpointed(mse_engine* e, int cfg, double* intensity);
/* Device bytes that mse_forward_checkpointed may spend on keeping every forward state
 * (n_slices * n_waves * res^2 * 8); default 1 GiB, 0 always keeps sqrt(n) checkpoints. */
int mse_set_checkpoint_budget(mse_engine* e, unsigned long long bytes);
/* Gradient of a loss L(I) of the exit intensities of the last mse_forward_checkpointed run:
 * dl_di = dL/dI (n_waves * N doubles, same layout as the intensities; for
 * L = sum (I - I_target)^2 it is 2 (I - I_target)).  Recomputes each segment of K slices
 * from its checkpoint (nothing when every state was kept) and runs the adjoint slice loop
 * backwards.  grad_xyz receives
 * dL/dx, dL/dy, dL/dz for every atom of the last mse_sort_atoms call on cfg, in input order
 * (3 doubles per atom; atoms outside the simulated volume get 0).  With projected
 * potentials dL/dz = 0 because an atom's potential does not depend on its depth within its
 * slice; with full-3D potentials the z derivative follows the sub-slice integration
 * (kernels/transmission_potentials_full_3d_f.cl:229-278), including the value carried
 * between atoms of one block row.  Gradients are undefined where a coordinate sits exactly
 * on a cutoff of the forward sum (slice and block edges, the 8 A lateral and 3 A depth
 * limits); MSE_ERR_UNSUPPORTED under mixed static potentials.  Device memory: the
 * checkpoints (ceil(n / K) * n_waves planes), one segment ((K - 1) * n_waves planes), and the
 * wave-summed g_t and the phase of every slice (n complex and n float planes). */
int mse_adjoint_positions(mse_engine* e, const double* dl_di, double* grad_xyz);
/* ---- sweep timing: `reps` back-to-back row passes (slice `slice`, every wave with its own
 *      configuration, current variant) and then `reps` column passes, timed with hipEvents on
 *      the engine stream; average microseconds per launch.  Leaves the waves' contents
 *      undefined. ---- */
int mse_time_sweeps(mse_engine* e, int slice, int reps, double* row_us, double* col_us);

/* ---- bandwidth reference: copy n_waves waves in place with the access pattern of the row
 *      sweep (pattern 0) or of the column sweep's LDS tiles (pattern 1, column variant
 *      `variant`), no arithmetic; average microseconds per launch.  Used to calibrate HBM
 *      traffic counters and as the achievable-bandwidth reference of each sweep. ---- */
int mse_bandwidth_probe(mse_engine* e, int pattern, int variant, int n_waves, int reps, double* us_per_launch);

/* ---- profiling: per-kernel device time measured with hipEvents on the engine stream ---- */
int mse_profile_enable(mse_engine* e, int enable);
/* kernel ids: 0 row pass, 1 column pass, 2 potential, 3 probe init, 4 other */
int mse_profile_read(mse_engine* e, int kernel_id, int* launches, double* total_ms);
int mse_profile_reset(mse_engine* e);

#ifdef __cplusplus
}
#endif
#endif
