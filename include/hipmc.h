/* C ABI of the MI355X transport-and-scoring engine (libhipmc.so).
 *
 * The boundary the Python HipEngine binds through ctypes. Plain pointers and sizes;
 * no torch or HIP types cross it. Every entry returns HMC_OK (0) or a negative status;
 * hmc_last_error() holds the message. Host arrays passed in are copied; the caller
 * keeps ownership. Device memory is owned by the context and released by
 * hmc_destroy(). One context per (device, host thread).
 *
 * Each entry names the reference interface it replaces (paths in the pyRadMC tree).
 */
#ifndef HIPMC_H
#define HIPMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HMC_OK = 0,
  HMC_EINVAL = -1,    /* invalid argument: ValueError on the Python side */
  HMC_EHIP = -2,      /* HIP runtime failure: RuntimeError */
  HMC_EOVERFLOW = -3, /* secondary queue overflow: RuntimeError */
  HMC_EMAJORANT = -4, /* Woodcock majorant violated in-kernel: RuntimeError */
  HMC_ENOMEM = -5,    /* device allocation failed: MemoryError */
  HMC_ESTATE = -6     /* call order (e.g. transport before tables): RuntimeError */
};

typedef struct hmc_ctx hmc_ctx;

/* Flattened cross-section tables, float32, as the reference's CrossSectionTables
 * fields (pyradmc/data/tables.py:139-219). Per-material arrays are [n_mat][n_points]. */
typedef struct hmc_tables {
  int n_points, n_mat, n_coherent;
  const float *majorant;                 /* [n_points] */
  const float *mu_compton, *mu_photo, *mu_pair, *mu_rayleigh;
  const float *csda_range, *moller, *restricted_range, *stopping_radiative;
  const float *scattering_power, *log_eta, *stopping_restricted;
  const float *energy_of_restricted_range;
  const float *coherent_x;               /* [n_coherent] */
  const float *coherent_cumulative;      /* [n_mat][n_coherent] */
  double p_log_min, p_inv_dlog, e_log_min, e_inv_dlog, r_log_min, r_inv_dlog;
} hmc_tables;

/* Source kinds generated in-kernel (pyradmc/backends/warp/kernels.py:1888-2401). */
enum {
  HMC_SRC_PENCIL = 0,          /* PencilBeamSource */
  HMC_SRC_PARALLEL = 1,        /* ParallelBeamSource */
  HMC_SRC_PARALLEL_SPECTRAL = 2, /* parallel field with spectral energies (extension) */
  HMC_SRC_SPECTRAL_FAN = 3,    /* SpectralBeamSource / SpectralBeamletSource */
  HMC_SRC_LATTICE = 4,         /* BeamletGridSource */
  HMC_SRC_UPLOADED = 5,        /* host-sampled primaries */
  HMC_SRC_RESIDENT = 6,        /* records drawn in-kernel from a device-resident population */
  HMC_SRC_SPOT_FAN = 7         /* GaussianSpotBeamSource / GaussianSpotBeamletSource */
};

typedef struct hmc_source {
  int kind;
  int particle;        /* 0 photon, 1 electron, 2 positron (mono sources) */
  double energy;       /* mono energy, MeV */
  double position[3];  /* pencil start */
  double direction[3]; /* pencil direction (unit) */
  double z0, x0, x_extent, y0, y_extent; /* parallel field */
  double focal[3], u_axis[3], v_axis[3], width_u, width_v; /* divergent fan */
  int n_bins;          /* spectrum bins */
  const double *edges; /* [n_bins + 1] */
  const double *cdf;   /* [n_bins] */
  int n_beamlets;      /* per-column parameters below */
  const double *beamlet_params; /* lattice: [n][4] x_lo, x_ext, y_lo, y_ext; fans: [n][3] centre */
  double sigma_u, sigma_v;     /* Gaussian focal spot (spot fan) */
} hmc_source;

/* A deterministic per-primary weight factor applied by the generator to in-kernel
 * sources (and resident phase spaces): the transmission of a beam-limiting stack,
 * exp(-sum_i rho_i mu_i(E) t_i), or a 0..1 mask read where the ray crosses a plane
 * rectangle. It replaces the host-side attenuation of sampled batches
 * (pyradmc/geometry/collimation.py:1071-1151) for sources generated on the device;
 * geometry, tables and mask conventions as pyradmc/geometry/collimation.py:280-489,
 * 927-1068. */
typedef struct hmc_ray_weights {
  int kind;                     /* 0 none, 1 beam-limiting stack, 2 plane mask */
  int n_devices;                /* stack: devices in order */
  double origin[3];             /* focal spot (engine frame) */
  double basis[9];              /* rows u, v, w of the beam frame (engine frame) */
  const int32_t *device_kind;   /* [n_devices] 0 jaw pair, 1 MLC */
  const int32_t *flags;         /* jaw: axis (0 u, 1 v) | focused << 1; MLC: focused sides << 1 */
  const int32_t *n_pairs;       /* MLC leaf pairs */
  const int32_t *edge_offset;   /* MLC: first strip edge in leaf_edges */
  const int32_t *pair_offset;   /* MLC: first pair in tips_neg / tips_pos */
  const double *z_top, *z_bottom;
  const double *edge_neg, *edge_pos; /* jaw edges at the mid-plane; MLC tip radius in edge_neg */
  const double *density;        /* g/cm^3 */
  int n_leaf_edges, n_tips;
  const double *leaf_edges, *tips_neg, *tips_pos;
  int n_table;                  /* log-log total mu/rho per device */
  const double *log_energies;   /* [n_table] increasing */
  const double *log_mu;         /* [n_devices][n_table] log(mu/rho), cm^2/g */
  double center[3], u_axis[3], v_axis[3], width_u, width_v; /* mask plane */
  int mask_nu, mask_nv;
  const double *mask;           /* [mask_nu][mask_nv], values in 0..1 */
} hmc_ray_weights;

typedef struct hmc_physics {
  double pcut, ecut, step_energy_fraction, deposit_resolution_cm;
  int transport_electrons, msc_gs, dose_to_water, reserved;
} hmc_physics;

/* One contiguous run of histories feeding one tally column: the history key of the
 * i-th primary is key_base + i. column = batch_slot * n_columns + group-local column;
 * beamlet indexes the source's per-beamlet parameters (0 for open-field sources). */
typedef struct hmc_work_item {
  int64_t key_base;
  int64_t count;
  int32_t column;
  int32_t beamlet;
} hmc_work_item;

/* Integer books of one hmc_transport call (quanta of 1e-9 MeV), per-kernel event
 * counts (the algorithmic-bytes denominators), and per-kernel HIP-event times. */
typedef struct hmc_ledger {
  int64_t escaped, unscored, deposited, emitted;
  int64_t violations, trapped, overflow, rounds;
  /* photon round, then electron round: voxel fetches, tally atomics, queue pushes,
   * particles popped, steps (Woodcock flights / condensed substeps) */
  int64_t photon_voxels, photon_deposits, photon_pushes, photon_pops, photon_steps;
  int64_t electron_voxels, electron_deposits, electron_pushes, electron_pops, electron_steps;
  int64_t photon_launches, electron_launches;
  double photon_kernel_ms, electron_kernel_ms; /* only with hmc_set_profiling(ctx, 1) */
  int64_t graph_launches; /* groups of rounds replayed from the captured graph */
} hmc_ledger;

/* Lifetime ------------------------------------------------------------------- */
/* Replaces WarpEngine construction (pyradmc/backends/warp/engine.py:369-407). */
int hmc_create(int device, hmc_ctx **out);
void hmc_destroy(hmc_ctx *ctx);
const char *hmc_last_error(const hmc_ctx *ctx);
/* Device name, CU count and memory, for auto-sizing (engine.py:175-255). Each output
   may be NULL; with only free_bytes asked for, the device properties are not read. */
int hmc_device_info(int device, char *name, int name_len, int *cu_count, int64_t *total_bytes,
                    int64_t *free_bytes);

/* Uploads -------------------------------------------------------------------- */
/* WarpEngine._upload_grid (engine.py:3336-3366). */
int hmc_set_grid(hmc_ctx *ctx, int nx, int ny, int nz, const double *origin, const double *spacing,
                 const float *density, const uint8_t *material);
/* _scoring_info + dense per-column maps (engine.py:287-330, 1413-1426). The int64
 * quanta maps hold n_slots batches side by side ([n_slots][n_columns][voxels]) so a
 * block of batches transports in one call; the float64 sums are [n_columns][voxels]. */
int hmc_set_scoring(hmc_ctx *ctx, int nx, int ny, int nz, const double *origin,
                    const double *spacing, const double *voxel_mass, int64_t n_columns,
                    int n_slots);
/* WarpEngine._upload_tables (engine.py:3368-3429). */
int hmc_set_tables(hmc_ctx *ctx, const hmc_tables *tables);
/* GS grid upload (engine.py:3418-3425, goudsmit_saunderson.py:789-823). */
int hmc_set_gs(hmc_ctx *ctx, const float *values, int n_eta, int n_theta2, int n_u, int ix0,
               int iy0);
/* Generator parameters (engine.py:2833-3006). */
int hmc_set_source(hmc_ctx *ctx, const hmc_source *source);
/* Host-sampled primaries for HMC_SRC_UPLOADED (kernels.py:1922-1960): columns of n. */
int hmc_set_uploaded(hmc_ctx *ctx, int64_t n, const int32_t *particle, const float *energy,
                     const float *x, const float *y, const float *z, const float *ux,
                     const float *uy, const float *uz, const float *weight);
/* A phase-space population kept in device memory for HMC_SRC_RESIDENT sources (an
 * extension: the reference samples a file's records on the host for every chunk,
 * pyradmc/backends/warp/engine.py:2506-2597, and keeps only its own head pre-solve
 * output on the device, pyradmc/backends/warp/presolve.py:710-742). Writes rows
 * [offset, offset + count) of a population of n records, in order: offset 0 starts a
 * new population of n (and unsets a resident source until hmc_set_source is called
 * again), each later call continues where the previous one ended, and a resident
 * source can be set once all n rows are written. A history draws row floor(U * n), U
 * from 64 bits of its own stream. Columns as hmc_set_uploaded (kinds 0 photon,
 * 1 electron, 2 positron). */
int hmc_set_population(hmc_ctx *ctx, int64_t n, int64_t offset, int64_t count,
                       const int32_t *particle, const float *energy, const float *x,
                       const float *y, const float *z, const float *ux, const float *uy,
                       const float *uz, const float *weight);
int hmc_set_physics(hmc_ctx *ctx, const hmc_physics *physics);

/* Treatment-head pre-solve (pyradmc/backends/warp/presolve.py:572-687, physics of
 * pyradmc/geometry/head.py:377-583). Per photon primary: Beer-Lambert transmission
 * through the stack, optionally a forced first Compton interaction at the mid-plane of
 * a device chosen by its share of the optical depth (the scattered photon attenuated
 * on its way out), an air column (attenuation, and a forced Compton photon and recoil
 * electron at an exponentially sampled depth), fair roulette of forced secondaries
 * below weight_floor, then every forward particle at or above its cutoff moved to the
 * exit plane w = exit_z. The population (weights scaled by kept / primaries) becomes
 * the context's resident population in primary order (stack photon, air photon, air
 * electron, transmitted photon of each primary), ready for a HMC_SRC_RESIDENT source. */
typedef struct hmc_presolve {
  int first_compton;            /* 0 attenuation only, 1 forced first Compton */
  int has_air;
  double exit_z;                /* local w of the exit plane */
  double air_z_start, air_z_end;
  double pcut, ecut, weight_floor;
  double origin[3], basis[9];   /* the beam frame (the stack's, when there is a stack) */
  int n_table;                  /* shared energy grid of the attenuation tables */
  int n_rows;                   /* stack devices in order, then the air column */
  const double *log_energies;   /* [n_table] increasing */
  const double *log_mu_total;   /* [n_rows][n_table] log of the macroscopic mu (1/cm) */
  const double *log_mu_compton; /* [n_rows][n_table] */
} hmc_presolve;
/* n photon primaries: drawn in-kernel from generator (an open-field photon source:
 * pencil, parallel, spectral parallel, or a one-aperture spectral or spot fan; it
 * becomes the context's source) on history h's stream keyed (seed, h), or, with
 * generator NULL, given as columns (sample_batch); stack: a kind-1 model (its tables
 * unused) or NULL; *n_out receives the population size. At most 16 devices. */
int hmc_presolve_head(hmc_ctx *ctx, uint64_t seed, int64_t n, const hmc_source *generator,
                      const float *energy, const float *x, const float *y, const float *z,
                      const float *ux, const float *uy, const float *uz, const float *weight,
                      const hmc_ray_weights *stack, const hmc_presolve *presolve, int64_t *n_out);
/* Rows [offset, offset + count) of the resident population (kinds 0 photon, 1 electron,
 * 2 positron) into host columns. */
int hmc_get_population(hmc_ctx *ctx, int64_t offset, int64_t count, int32_t *particle, float *energy,
                       float *x, float *y, float *z, float *ux, float *uy, float *uz, float *weight);
/* The ray weight model of the following transports (NULL or kind 0: none). */
int hmc_set_ray_weights(hmc_ctx *ctx, const hmc_ray_weights *weights);
/* The current model's factor for each of n rays (energy, position [n][3], direction
 * [n][3], float32 as the generator holds a primary) into out[n] (a probe of the
 * device evaluation; no reference counterpart). */
int hmc_ray_weight_factors(hmc_ctx *ctx, int64_t n, const float *energy, const float *position,
                           const float *direction, double *out);
/* Time every transport launch with HIP events on the context's stream (no
 * reference counterpart: the measurement hook for the per-kernel roofline). */
int hmc_set_profiling(hmc_ctx *ctx, int enabled);
/* Launch shape of the transport rounds: lanes per CU of each round's grid (multiple
 * of 256 in [256, 8192]; default 2048), queue indices a wave claims at a time and
 * output slots it reserves at a time (each a multiple of 64 in [64, 1024]; default
 * 64), and whether each group of four rounds is replayed from a captured HIP graph
 * (default 1; profiled transports launch one by one, so that every kernel is
 * timed). None of them changes a result. No reference counterpart beyond the
 * chunk_size / queue_factor knobs of WarpEngine (engine.py:369-407): these are the
 * autotuner's knobs. */
int hmc_set_launch(hmc_ctx *ctx, int round_lanes_per_cu, int refill_block, int push_chunk,
                   int use_graphs);

/* Work ----------------------------------------------------------------------- */
/* Generate and transport every primary of the work items into the dense tallies:
 * replaces _transport_chunk* + _drain_queues (engine.py:2097-2831). chunk bounds
 * the histories resident at once; queue_factor sizes the secondary queues. */
int hmc_transport(hmc_ctx *ctx, uint64_t seed, const hmc_work_item *items, int n_items,
                  int64_t chunk, int queue_factor, hmc_ledger *ledger);
/* Fold batch slots 0..n_slots_used-1, in slot order, into the running f64 sums and
 * consume the maps: accumulate_run_batch / accumulate_dij_batch (kernels.py:804-878).
 * nhist[slot * n_columns + column] = histories of that column in that batch.
 * total_quanta receives the folded quanta after the fold has run on the device. With
 * total_quanta NULL the call returns once the fold is queued, and the quanta of such
 * folds add up until hmc_finalize reads them back (hmc_folded_quanta). */
int hmc_fold(hmc_ctx *ctx, int n_slots_used, const double *nhist, int64_t *total_quanta);
/* The quanta folded by hmc_fold calls without total_quanta since the previous call of
 * this function: read back by the hmc_finalize after them (or here, waiting for the
 * device, if none came). It counts the energy the reference's accumulate_*_batch add
 * to the deposited total (kernels.py:804-878); resets the count to zero. */
int hmc_folded_quanta(hmc_ctx *ctx, int64_t *total_quanta);
/* Mean and sigma per (column, voxel): finalize_run (kernels.py:881-909). With
 * mean/sigma NULL the result stays on the device for hmc_truncate. */
int hmc_finalize(hmc_ctx *ctx, int n_batches, double *mean, double *sigma);
/* Per-column truncation and ordered CSC compaction on the device: column_max,
 * count_kept_per_chunk, compact_column_chunked (kernels.py:1032-1168) and the
 * host prefix (engine.py:1496-1557). counts[n_columns] out; entries fetched next. */
int hmc_truncate(hmc_ctx *ctx, double truncation, int64_t *counts, int64_t *total);
int hmc_fetch_csc(hmc_ctx *ctx, int64_t *indices, double *dose, double *sigma);
/* The same entries with float32 dose and sigma (rounded to nearest on the device), for a
 * float32 CSC at half the bytes over PCIe; the reference's DijAssembler.finalize
 * (pyradmc/scoring/dij.py:397-434) converts its float64 columns on the host. */
int hmc_fetch_csc_f32(hmc_ctx *ctx, int64_t *indices, float *dose, float *sigma);

/* Page-locked host memory for the fetched CSC arrays (no counterpart in the reference,
 * whose DijAssembler.finalize, pyradmc/scoring/dij.py:397-434, returns ordinary host
 * arrays). A device-to-host copy into it runs at the link's rate; into freshly allocated
 * pageable memory it is bound by first-touch page faults at a fraction of that.
 * hmc_host_alloc returns HMC_ENOMEM when the allocation fails; hmc_host_free(NULL) is a
 * no-op. */
int hmc_host_alloc(size_t bytes, void **out);
void hmc_host_free(void *ptr);

/* Beam-aligned windows of the finalized columns, for the learned denoiser (no
 * counterpart in the reference; the denoised Dij builder in pyradmc_amd/denoise.py).
 * origins: host int32 [n_cols][2] (x0, y0), n_cols equal to the scoring column count;
 * a window is wx x wy scoring voxels laterally and the full depth.
 * hmc_gather_windows writes float32 out[c][z][x][y] (device pointer) from the means
 * kept on the device by hmc_finalize(ctx, n, NULL, NULL).
 * hmc_scatter_windows replaces every column's mean and sigma by the window values of
 * mean / sigma (device float32 [c][z][x][y]; sigma may be NULL for zero) and zero
 * elsewhere, ready for hmc_truncate. Both return after the device work is done. */
int hmc_gather_windows(hmc_ctx *ctx, const int32_t *origins, int64_t n_cols, int wx, int wy,
                       float *out);
int hmc_scatter_windows(hmc_ctx *ctx, const int32_t *origins, int64_t n_cols, int wx, int wy,
                        const float *mean, const float *sigma);
/* Overwrite mean and sigma of one column at n voxels (flat scoring indices) after
 * hmc_finalize or hmc_scatter_windows, before hmc_truncate: the denoised builder puts
 * the plain Monte Carlo entries of its audited columns there, so one device truncation
 * and compaction yields the whole Dij in column order. Host arrays; indices outside the
 * grid are refused. */
int hmc_set_column_entries(hmc_ctx *ctx, int64_t column, int64_t n, const int64_t *indices,
                           const double *dose, const double *sigma);
/* Clear the dense maps and sums for a new group. */
int hmc_reset_maps(hmc_ctx *ctx);
/* Replace / read the dense int64 quanta maps [n_slots][n_columns][n_voxels]: the synthetic
 * fixed-point maps the reference drives its device reduction with
 * (tests/integration/test_warp_device_reduction.py:38-62). */
int hmc_set_quanta(hmc_ctx *ctx, const int64_t *quanta);
int hmc_get_quanta(hmc_ctx *ctx, int64_t *quanta);

/* Sampler draws for the device-level physics tests (no counterpart in the reference's
 * engine API; the probe its own device sampler tests launch, tests/physics/
 * test_warp_physics.py:54-120 and test_warp_gs_sampler.py:43-128). n independent
 * streams (keys 0..n-1 of seed) each draw every scalar sampler the transport rounds
 * use, with the rounds' own device functions; out is host float [n][HMC_SAMPLE_COLUMNS]:
 *   0 free path at path_mu             1 channel (0 Compton .. 3 Rayleigh) from channel_mu
 *   2 Compton ratio at compton_energy[0]  3 its photon cosine  4 its electron cosine
 *   5 Compton ratio at compton_energy[1]  6 |isotropic|^2   7 isotropic z
 *   8 |isotropic turned by (rotate_cos, rotate_phi)|^2
 *   9 Moller delta energy at (moller_energy, moller_cut)  10 delta cosine  11 primary cosine
 *  12 Gaussian hinge cosine at hinge_ms
 *  13 bremsstrahlung emission probability and 14 local part for (brems_energy,
 *     brems_radiative_loss, brems_pcut)   15 bremsstrahlung photon energy
 *  16 coherent cosine at (coherent_energy, coherent_material) from the uploaded tables
 *  17, 18 GS cosine at (gs_log_eta, gs_ms[0]) and (gs_log_eta, gs_ms[1]) from the uploaded grid
 *  19 the uniform drawn after a GS sample, 20 after a Gaussian hinge, from the same state
 *  21 z of (0, 0, 1) turned by (rotate_cos, rotate_phi) (the rotation's pole branch)
 * Columns 16-19 are NaN without hmc_set_tables (and hmc_set_gs for 17-19). */
#define HMC_SAMPLE_COLUMNS 22
typedef struct hmc_sampler_args {
  double path_mu;
  double channel_mu[4];
  double compton_energy[2];
  double moller_energy, moller_cut;
  double hinge_ms;
  double brems_energy, brems_radiative_loss, brems_pcut;
  double coherent_energy;
  int coherent_material, reserved;
  double gs_log_eta, gs_ms[2];
  double rotate_cos, rotate_phi;
} hmc_sampler_args;
int hmc_sample_physics(hmc_ctx *ctx, uint64_t seed, int64_t n, const hmc_sampler_args *args,
                       float *out);

/* Host-side Goudsmit-Saunderson node grid (build_gs_grid, goudsmit_saunderson.py:
 * 1108-1219): fills out[(ix1-ix0+1)][(iy1-iy0+1)][n_u] with scaled-deflection tables
 * at eta = exp(i/bins), <theta^2> = exp(j/bins). n_threads <= 0 uses all cores. */
int hmc_build_gs_grid(int ix0, int ix1, int iy0, int iy1, int n_u, double bins_per_log,
                      int n_threads, double *out);

#ifdef __cplusplus
}
#endif

#endif /* HIPMC_H */
