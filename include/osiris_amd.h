/*
 * osiris_amd.h -- C ABI of the MI355X (gfx950) PIC hot path.
 *
 * Plain C types only (pointers, sizes, status codes).  Every entry point returns an int
 * status (OSR_OK = 0) unless it mirrors a reference entry point that returns nothing;
 * the message of the last failure of the calling thread is osr_last_error().
 *
 * Three groups of entry points:
 *
 *  1. Stateless drop-ins for the reference's native fused kernels.  Same argument list,
 *     meaning and layout as FNAME(vadvance_deposit_{2d,3d}_s{1,2,3}) in
 *     simd/os-spec-push-avx2.c:471-475 (2D) and :761-765 (3D), called from
 *     spec/os-spec-vpush.f03:144-169 / :294-318 with every argument by reference.
 *     They copy the caller's host arrays to the GPU, run gather + Boris push + position
 *     advance + charge-conserving deposit, and copy the results back.
 *
 *  2. A stateless Yee advance on the caller's guarded field arrays, replacing
 *     t_emf_solver%advance for periodic boundaries
 *     (emf/solver/os-emf-solver-define.f03:26-34, os-emf-solver-yee.f03:542-568).
 *
 *  3. The resident fast path: a simulation handle whose fields and particles live in
 *     HBM; osr_sim_run advances whole time steps (gather, push, deposit, current fold,
 *     Yee, energy diagnostics) without host round trips.  This replaces, per time
 *     step, t_particles%advance_deposit (particles/os-particles-define.f03:256-269)
 *     followed by jay%update_boundary and t_emf%advance (os-simulation.f03:588-661).
 *
 * Array conventions of the reference (spec/os-spec-define.f03:391-412,
 * vdf/os-vdf-define.f03:35-116): particle ix(D,N) are 1-based cell indices, x(D,N) the
 * position in [-0.5,0.5) relative to the cell, u(3,N) proper velocities, q(N) charges;
 * grids are f(3, 1-gcl:nx+gcu, 1-gcl:ny+gcu[, ...]) stored component fastest.
 */
#ifndef OSIRIS_AMD_H
#define OSIRIS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OSR_OK = 0,
  OSR_ERR_ARG = 1,      /* invalid argument */
  OSR_ERR_HIP = 2,      /* HIP runtime error (no device, launch failure, ...) */
  OSR_ERR_NOMEM = 3,    /* device allocation failed */
  OSR_ERR_STATE = 4,    /* call not valid in the current state */
  OSR_ERR_LAYOUT = 5    /* a particle was found outside its tile (internal invariant) */
};

enum { OSR_BORIS = 0, OSR_VAY = 1 };

/* Message of the last error on this thread ("" when none). */
const char *osr_last_error(void);
/* ABI version: major*10000 + minor*100 + patch. */
int osr_version(void);
/* Number of visible GPUs (0 when none); never initialises a context on failure. */
int osr_device_count(void);

/* ---------------------------------------------------------------------------------
 * 1. Stateless drop-ins (reference: simd/os-spec-push-avx2.c:471-475, 761-765).
 *
 * osr_advance_deposit is the status-returning core; pusher = OSR_BORIS or OSR_VAY
 * (the reference's SIMD kernels are Boris only).  The particles i0..i1 (1-based,
 * inclusive) are pushed; ix, x, u are updated in place; the current of their motion is
 * ADDED to j (the caller zeroes it, particles/os-particles.f03:1426,1453); *ene is
 * incremented by the time-centred Sum q*u^2/(gamma+1) (not normalised), as the
 * reference does.  emf_size/j_size are the array extents in cells (guards included),
 * emf_offset/j_offset the number of lower guard cells per dimension.
 * Differences from the reference: no dummy particles are written past i1, and the
 * current is accumulated exactly (int64 fixed point) so the result does not depend on
 * thread scheduling.
 * ------------------------------------------------------------------------------- */
int osr_advance_deposit(int dims, int order, int pusher, int *ix, float *x, float *u,
                        const float *q, int i0, int i1, float rqm, const float *efield,
                        const float *bfield, const int *emf_size, const int *emf_offset,
                        float *j, const int *j_size, const int *j_offset, const double *dx,
                        double dt, double *ene);

/* osr_advance_deposit on arrays already in device memory, with no device allocation.
 * It has no counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Arguments and results are osr_advance_deposit's, except that ix, x, u, q,
 * efield, bfield and j are DEVICE pointers (same layouts); emf_size, emf_offset, j_size,
 * j_offset, dx and ene stay host pointers.  The working memory (int64 fixed-point
 * current, per-block energy partials, range-check flag) is a fixed 64 MiB device array of
 * the library, one per GPU for the life of the process: 8192 + 8 * ceil(np / 256) bytes
 * (rounded up to 256) + 8 * 3 * prod(j_size) must fit in it, else OSR_ERR_NOMEM and
 * nothing is touched.  Every call on one GPU shares that array, so calls that overlap on
 * one GPU race and must be serialised.  Kernels run on the null stream; the call returns
 * after they finished. */
int osr_advance_deposit_noalloc(int dims, int order, int pusher, int *ix, float *x, float *u,
                                const float *q, int i0, int i1, float rqm,
                                const float *efield, const float *bfield,
                                const int *emf_size, const int *emf_offset, float *j,
                                const int *j_size, const int *j_offset, const double *dx,
                                double dt, double *ene);

/* Fortran-callable twins of FNAME(vadvance_deposit_*_sN): every argument by reference,
 * no return value.  On failure they print the error and exit(1), as the reference's
 * kernels do on a violated precondition (simd/os-spec-push-avx2.c:537-547). */
void hip_vadvance_deposit_2d_s1_(int *ix, float *x, float *u, float *q, int *i0, int *i1,
                                 float *rqm, float *efield, float *bfield, int *emf_size,
                                 int *emf_offset, float *j, int *j_size, int *j_offset,
                                 double *dx, double *dt, double *ene);
void hip_vadvance_deposit_2d_s2_(int *ix, float *x, float *u, float *q, int *i0, int *i1,
                                 float *rqm, float *efield, float *bfield, int *emf_size,
                                 int *emf_offset, float *j, int *j_size, int *j_offset,
                                 double *dx, double *dt, double *ene);
void hip_vadvance_deposit_2d_s3_(int *ix, float *x, float *u, float *q, int *i0, int *i1,
                                 float *rqm, float *efield, float *bfield, int *emf_size,
                                 int *emf_offset, float *j, int *j_size, int *j_offset,
                                 double *dx, double *dt, double *ene);
void hip_vadvance_deposit_3d_s1_(int *ix, float *x, float *u, float *q, int *i0, int *i1,
                                 float *rqm, float *efield, float *bfield, int *emf_size,
                                 int *emf_offset, float *j, int *j_size, int *j_offset,
                                 double *dx, double *dt, double *ene);
void hip_vadvance_deposit_3d_s2_(int *ix, float *x, float *u, float *q, int *i0, int *i1,
                                 float *rqm, float *efield, float *bfield, int *emf_size,
                                 int *emf_offset, float *j, int *j_size, int *j_offset,
                                 double *dx, double *dt, double *ene);
void hip_vadvance_deposit_3d_s3_(int *ix, float *x, float *u, float *q, int *i0, int *i1,
                                 float *rqm, float *efield, float *bfield, int *emf_size,
                                 int *emf_offset, float *j, int *j_size, int *j_offset,
                                 double *dx, double *dt, double *ene);

/* ---------------------------------------------------------------------------------
 * 2. Stateless Yee advance, periodic boundaries (os-emf-solver-yee.f03:542-568):
 * B += -dt/2 curl E, E += dt curl B - dt J, B += -dt/2 curl E, guard cells refreshed
 * with the periodic images after each sub-step (os-emf-boundary.f90:785-786).
 * e, b, jay share the extents ld[] (cells incl. guards) and lower guards gcl[]; nx[]
 * are the interior cells.  jay must already be folded (jay%update_boundary).
 * ------------------------------------------------------------------------------- */
int osr_yee_advance(int dims, const int *nx, const int *ld, const int *gcl, const double *dx,
                    double dt, float *e, float *b, const float *jay);

/* ---------------------------------------------------------------------------------
 * 3. Resident simulation
 * ------------------------------------------------------------------------------- */
typedef struct osr_sim osr_sim;

typedef struct {
  int dims;       /* 2 or 3 */
  int nx[3];      /* interior cells per dimension (unused entries ignored) */
  double dx[3];   /* cell size */
  double dt;      /* time step (Courant condition of the Yee scheme is checked) */
  int order;      /* interpolation / deposit order 1..3 */
  int tile[3];    /* tile size in cells; 0 = default (2D 16x16, 3D 8x8x4) */
  int path;       /* OSR_PATH_FUSED (default) or OSR_PATH_PLAIN (2D only) */
  int kernel;     /* fused push kernel: 0 auto, 1 generic (runtime tile shape), 2
                     compile-time tile shape (2D tiles 8x8, 16x16, 32x16; 3D tile 8x8x4);
                     all accumulate the current in int64 fixed point in LDS and give
                     bit-identical results */
  int xbnd;       /* x walls: OSR_BND_PERIODIC (0), OSR_BND_LINDMAN (1, also named
                     OSR_BND_OPEN) or OSR_BND_VPML (2) */
  int reorder;    /* fused path: restore the (rank in cell, cell) particle order inside the
                     tiles every this many steps; 0 = default (4 in 3D, off in 2D),
                     negative = off.  Changes no result beyond FP64 round-off of the
                     energy sums (the counterpart of the reference's periodic sort,
                     sort_species, spec/os-species.f03:269-332) */
  int vpml_size;  /* OSR_BND_VPML: wall thickness in cells, 4..64; 0 = 8 (vpml_bnd_size,
                     p_vpml_size_default, emf/boundary/os-emf-boundary.f90:30, 230) */
  int vpml_diffuse[2]; /* OSR_BND_VPML in 2D, orders 2 and 3: nonzero mixes the wall's
                     split B3 pair before each wall E update, lower / upper x wall
                     (vpml_diffuse(:, 1), default .false., os-emf-boundary.f90:231;
                     os-emf-vpml.f03:977-1006) */
  int push_waves; /* fused path: waves per push workgroup, 0 = default (8 for 2D
                     compile-time tiles of 512 cells or more, else 4); 4 or 8 for the 2D
                     compile-time tiles 16x16 and 32x16.  Changes no result */
} osr_sim_desc;

/* x boundaries of a resident simulation.  Both open types have absorbing particles
 * (spec_bound type "absorbing" / "open", spec/os-spec-boundary.f03:150-151, 476-513:
 * particles leaving through x are removed after the push); y (and z) stay periodic.
 * OSR_BND_LINDMAN (= OSR_BND_OPEN): Lindman absorbing field walls at both x ends
 * (emf_bound type "lindman", emf/boundary/os-emf-lindman.f90:190-941, applied by
 * os-emf-boundary.f90:723-1112), 2D and 3D.  OSR_BND_VPML: the split-field perfectly
 * matched layer of vpml_size cells at both x ends (emf_bound type "vpml", which the
 * reference's deck also names "open", os-emf-boundary.f90:262-263; emf/boundary/
 * os-emf-vpml.f03: setup_vpml :270-474, setup_att_coeff :3586-3689, update_e_vpml_2d/3d
 * :598-856, update_b_vpml_2d/3d :1742-...), 2D and 3D, with the optional 2D diffusion
 * (osr_sim_desc.vpml_diffuse).  Fused path, nx a multiple of the tile width, no
 * current smoothing; osr_sim_gauss is not available. */
enum { OSR_BND_PERIODIC = 0, OSR_BND_OPEN = 1, OSR_BND_LINDMAN = 1, OSR_BND_VPML = 2 };

/* OSR_PATH_FUSED: tiled fused push+deposit, LDS-staged fields and current, particles
 * sorted into tiles every step, reproducible current.  OSR_PATH_PLAIN: one kernel per
 * phase (gather, push, move/split, deposit with global FP32 atomics), intermediate
 * arrays in HBM, no sorting -- the comparison baseline; not reproducible run to run. */
enum { OSR_PATH_FUSED = 0, OSR_PATH_PLAIN = 1 };

int osr_sim_create(const osr_sim_desc *desc, osr_sim **out);
int osr_sim_destroy(osr_sim *sim);

/* Adds a particle species (rqm = charge/mass ratio, pusher OSR_BORIS/OSR_VAY);
 * returns its id in *sid.  At most 8 species. */
int osr_sim_add_species(osr_sim *sim, float rqm, int pusher, int *sid);

/* Replaces the particles of a species with np particles in the reference layout.
 * Positions must satisfy -0.5 <= x < 0.5 and 1 <= ix <= nx. */
int osr_species_set(osr_sim *sim, int sid, int64_t np, const int *ix, const float *x,
                    const float *u, const float *q);
/* Number of particles of a species. */
int osr_species_count(osr_sim *sim, int sid, int64_t *np);
/* Copies the particles back in the order they were set (or generated). */
int osr_species_get(osr_sim *sim, int sid, int *ix, float *x, float *u, float *q);

/* Generates particles on the device: ppc[d] particles per cell along each dimension
 * (regular lattice as spec/psource/os-psource-std.f03:822-869 when random_pos == 0,
 * uniform random in the cell otherwise), density n (q = sign(rqm) * n / ppc), momenta
 * u = ufl + uth * N(0,1) per component, counter-based RNG seeded with seed. */
int osr_species_generate(osr_sim *sim, int sid, const int *ppc, double density,
                         const double *ufl, const double *uth, uint64_t seed, int random_pos);

/* Fields in the reference layout with ld[] extents and gcl[] lower guard cells; only the
 * interior cells are read (set) or written (get; guards filled with periodic images). */
int osr_fields_set(osr_sim *sim, const float *e, const float *b, const int *ld,
                   const int *gcl);
int osr_fields_get(osr_sim *sim, float *e, float *b, const int *ld, const int *gcl);
/* Uniform initial fields (e.g. type_init_b = "uniform", os-emf-gridval.f03). */
int osr_fields_fill(osr_sim *sim, const double *e0, const double *b0);

/* Advances nsteps full time steps (blocking). */
int osr_sim_run(osr_sim *sim, int nsteps);
/* Completed time steps. */
int osr_sim_steps(osr_sim *sim, int64_t *steps);
/* Energy history rows first..first+count-1: fld[6*count] = (B1,B2,B3,E1,E2,E3) energies
 * of the fields the particles were pushed with, par[nspecies*count] = per-species
 * time-centred kinetic energy, normalised as HIST/fld_ene and HIST/par_ene
 * (os-emf-diagnostics.f03:44-55, os-spec-define.f03:1030-1086). */
int osr_sim_history(osr_sim *sim, int64_t first, int64_t count, double *fld, double *par);
/* Gauss-law residual div E - rho (FP64) on the interior cell corners, with rho deposited
 * by DEPOSIT_RHO (spec/os-spec-charge.f90:500-580) and div E by div_vdf
 * (vdf/os-vdf-math.f90:1013-1073).  res and rho (ncell doubles each) may be NULL;
 * *maxabs receives max |res|. */
int osr_sim_gauss(osr_sim *sim, double *res, double *rho, double *maxabs);
/* Particle re-layouts done because a tile ran out of capacity. */
int osr_sim_relayouts(osr_sim *sim, int64_t *count);

/* Per-kernel timing of nsteps steps, measured with HIP events on the simulation's
 * stream (the steps are really advanced).  ms[k] = mean duration per step of kernel
 * class k: 0 push+deposit (all species), 1 tile fill and re-order, 2 field advance
 * (smoothing, Yee sub-steps, wall updates), 3 step end (energy reduction; with Lindman
 * walls it runs in the last wall kernel and this is 0); ms[4] = mean time per step from
 * the first to the last event; ms[5] = 0. */
int osr_sim_time_kernels(osr_sim *sim, int nsteps, double *ms);
/* Bytes of device memory held by a simulation. */
int osr_sim_footprint(osr_sim *sim, int64_t *bytes);
/* Launch configuration the fused push resolved to (no counterpart in the reference):
 * tile[3] in cells, waves per workgroup, and *compiled_tile = 1 when the compile-time
 * tile variant runs, 0 for the generic kernel (or the plain path). */
int osr_sim_launch_config(osr_sim *sim, int *tile, int *waves, int *compiled_tile);

/* ---------------------------------------------------------------------------------
 * 4. Ensemble serving (no counterpart in the reference, which runs one simulation per
 * process).  Independent jobs -- each one simulation as osr_sim_create would run it, set
 * up by osr_species_generate -- are submitted at any time; jobs are admitted in
 * submission order while their device footprint fits the budget, and all admitted jobs
 * that share dims, order and tile shape are advanced together: one launch per kernel
 * and time step covers the tiles of every job (descriptor arrays + work lists), and
 * each step is a captured hipGraph.  A job parks itself on the device when it reaches
 * its step count; finished jobs free their memory for waiting ones.  A job's results
 * are bit-identical to running it alone.
 * ------------------------------------------------------------------------------- */
typedef struct osr_ensemble osr_ensemble;

typedef struct {
  int64_t hbm_budget;  /* device bytes the admitted jobs may hold (0: 80% of free memory) */
  int max_active;      /* max admitted jobs (0: 256) */
  int chunk_steps;     /* graph replays between scheduling points (0: 16) */
  int reserved[5];
} osr_ens_desc;

typedef struct {
  float rqm;          /* charge/mass ratio */
  int pusher;         /* OSR_BORIS / OSR_VAY */
  int ppc[3];         /* particles per cell per dimension */
  int random_pos;     /* 1: uniform random in the cell, 0: lattice */
  double density;
  double ufl[3];      /* fluid proper velocity */
  double uth[3];      /* thermal spread */
  uint64_t seed;
} osr_job_species;

typedef struct {
  osr_sim_desc sim;   /* grid, order, tile (path must be OSR_PATH_FUSED) */
  int64_t nsteps;     /* time steps to run */
  double e0[3], b0[3];/* uniform initial fields */
  int nspecies;       /* 1..8 */
  int reserved;
  osr_job_species species[8];
} osr_job_desc;

enum { OSR_JOB_PENDING = 0, OSR_JOB_ACTIVE = 1, OSR_JOB_DONE = 2, OSR_JOB_FAILED = 3 };

typedef struct {
  int state;          /* OSR_JOB_* */
  int nspecies;
  int64_t steps;      /* completed steps */
  int64_t particles;  /* total particles (0 before admission) */
  int64_t footprint;  /* device bytes held while active */
  double t_submit, t_admit, t_done;  /* seconds since osr_ens_create */
} osr_job_info;

int osr_ens_create(const osr_ens_desc *desc, osr_ensemble **out);
/* Frees every job, finished or not. */
int osr_ens_destroy(osr_ensemble *ens);
/* Queues a job; *job_id identifies it from now on.  Fails if the job alone can never
 * fit the budget. */
int osr_ens_submit(osr_ensemble *ens, const osr_job_desc *job, int64_t *job_id);
/* One scheduling pass, never blocks on the GPU: collects finished step chunks (re-laying
 * out overflowed jobs, retiring finished ones), admits waiting jobs, launches the next
 * chunk of every idle batch.  *active / *pending (may be NULL) receive the job counts. */
int osr_ens_progress(osr_ensemble *ens, int *active, int *pending);
/* Runs the scheduler until every submitted job has finished. */
int osr_ens_wait(osr_ensemble *ens);
int osr_ens_job_info(osr_ensemble *ens, int64_t job_id, osr_job_info *info);
/* Energy history of a finished job, as osr_sim_history. */
int osr_ens_job_history(osr_ensemble *ens, int64_t job_id, int64_t first, int64_t count,
                        double *fld, double *par);
/* Forgets a finished job's results. */
int osr_ens_job_release(osr_ensemble *ens, int64_t job_id);
/* Device bytes currently held by admitted jobs. */
int osr_ens_usage(osr_ensemble *ens, int64_t *bytes);

/* ---------------------------------------------------------------------------------
 * 5. Current smoothing (vdf/os-vdf-smooth.f90), the per-step option of the current's
 * nl_smooth section (current/os-current-define.f03:298-312, applied after the current's
 * boundary update, os-simulation.f03:650-657).  Filter types use the reference's codes
 * (os-vdf-smooth.f90:22-31).
 * ------------------------------------------------------------------------------- */
enum {
  OSR_SMOOTH_NONE = 0,
  OSR_SMOOTH_CUSTOM = 1,
  OSR_SMOOTH_BINOMIAL = 2,
  OSR_SMOOTH_COMPENSATED = 3,
  OSR_SMOOTH_5PASS = 4,
  OSR_SMOOTH_TRI6 = 5,
  OSR_SMOOTH_TRI8 = 6,
  OSR_SMOOTH_LASERKDX = 10,
  OSR_SMOOTH_DIGITAL = 11
};
#define OSR_SMOOTH_ORDER_MAX 15 /* p_order_max */
#define OSR_SMOOTH_NCOEF 31     /* 2 * OSR_SMOOTH_ORDER_MAX + 1 */

/* The nl_smooth namelist (os-vdf-smooth.f90:214-216), one entry per direction. */
typedef struct {
  int type[3];                 /* type(:) as an OSR_SMOOTH_* code */
  int order[3];                /* order(:) */
  double swfj[3][15][3];       /* swfj(1:3, pass, dir) of the custom filter */
  double laserkdx_omega0;      /* laserkdx_omega0 */
  int laserkdx_dir;            /* laserkdx_dir (1-based) */
  int reserved;
  double digital_A[3];         /* digital_A(:) */
  double digital_fc[3];        /* digital_fc(:) */
} osr_smooth_params;

/* read_nml_smooth + setup_smooth (os-vdf-smooth.f90:183-470) for the single-precision
 * build: order[3] receives the filter order per direction (0 = none), coef[3 * 31] the
 * normalised FP32 coefficients, coef[31 d + 15 + k] = s(k) for k = -order[d]..order[d].
 * gamma is the simulation's boost factor (1 when not boosted), used by laserkdx only.
 * Host only; returns OSR_ERR_ARG where the reference aborts. */
int osr_smooth_setup(int dims, const osr_smooth_params *p, const double *dx, double gamma,
                     int *order, float *coef);
/* smooth_vdf (os-vdf-smooth.f90:515-540) on a host array f(fdim, ld[0], ld[1][, ld[2]])
 * with guard cells: direction by direction, every line of the array, cells at least
 * order[d] away from both ends replaced by the filtered value; in place. */
int osr_smooth_vdf(int dims, const int *ld, int fdim, const int *order, const float *coef,
                   float *f);
/* Resident simulation: smooth the current every step before the field advance
 * (t_current%smooth); all orders 0 switches it off. */
int osr_sim_set_smooth(osr_sim *sim, const int *order, const float *coef);

/* ---------------------------------------------------------------------------------
 * 6. Output in the reference's formats (host code; the resident forms download first)
 * ------------------------------------------------------------------------------- */
/* HIST energy files: appends nrows rows to <dir>/fld_ene (report_energy_emf,
 * os-emf-diagnostics.f03:44-85), <dir>/par_ene (report_energy_particles,
 * os-particles.f03:1710-1780) and <dir>/parNN_ene, NN = 01.. (report_energy_species,
 * os-spec-diagnostics.f03:1840-1885), creating each file with its header when a row's time
 * is 0, with the reference's Fortran formats ('(I6,1X,G15.8,6(1X,ES23.16))' etc.).
 * fld[6 r + c]: B1 B2 B3 E1 E2 E3 of row r; par[nspec r + s]; npar[s] particle counts;
 * names[s] species names (NULL: "species <s+1>").  Creates dir if needed. */
int osr_hist_append(const char *dir, int nspec, const char *const *names, int64_t nrows,
                    const int64_t *iter, const double *time, const double *fld,
                    const double *par, const int64_t *npar);
/* The same for history rows first .. first+count-1 of a resident simulation (iteration n,
 * time n dt). */
int osr_sim_write_hist(osr_sim *sim, const char *dir, const char *const *names, int64_t first,
                       int64_t count);

/* ZDF (zdf/zdf.h): data types and metadata as t_zdf_grid_info / t_zdf_grid_axis /
 * t_zdf_iteration / t_zdf_part_info. */
enum { OSR_ZDF_INT8 = 1, OSR_ZDF_UINT8, OSR_ZDF_INT16, OSR_ZDF_UINT16, OSR_ZDF_INT32,
       OSR_ZDF_UINT32, OSR_ZDF_INT64, OSR_ZDF_UINT64, OSR_ZDF_FLOAT32, OSR_ZDF_FLOAT64 };
typedef struct {
  const char *name;
  int type; /* 0 linear, 1 log10, 2 log2 */
  double min, max;
  const char *label, *units;
} osr_zdf_grid_axis;
typedef struct {
  const char *name;
  int ndims;
  uint64_t count[3];
  const char *label, *units;
  const osr_zdf_grid_axis *axis; /* ndims entries or NULL */
} osr_zdf_grid_info;
typedef struct {
  const char *name; /* "ITERATION" in OSIRIS */
  int32_t n;
  double t;
  const char *time_units;
} osr_zdf_iteration;
typedef struct {
  const char *name, *label;
  uint64_t np;
  int nquants;
  const char *const *quants, *const *qlabels, *const *qunits;
} osr_zdf_part_info;

/* zdf_save_grid (zdf.c:1585-1610): <path>/<label>-<n %06u>.zdf with one dataset. */
int osr_zdf_save_grid(const void *data, int data_type, const osr_zdf_grid_info *info,
                      const osr_zdf_iteration *iter, const char *path);
/* A grid file as OSIRIS's diagnostic files write it from one node (open_file_zdf +
 * start_cdset / write_cdset / end_cdset, os-diagfile-zdf.f03:297-347, 481-595): the
 * dataset as a chunked dataset of one chunk (start 0, stride 1); file is the full name. */
int osr_zdf_write_grid_file(const char *file, const void *data, int data_type,
                            const osr_zdf_grid_info *info, const osr_zdf_iteration *iter);
/* zdf_open_part_file + zdf_add_quant_part_file (zdf.c:1612-1660): <path>/particles-<name>-
 * <n %06u>.zdf, one float32 dataset of np values per quantity. */
int osr_zdf_save_particles(const osr_zdf_part_info *info, const osr_zdf_iteration *iter,
                           const char *path, const float *const *quants);
/* Field dump of a resident simulation as OSIRIS writes MS/FLD/<q>/<q>-<n %06u>.zdf
 * (labels and units of os-emf-diag-define.f03:164-190): quant 0..5 = e1 e2 e3 b1 b2 b3,
 * interior cells, axes x1..x3 over [xmin, xmin + nx dx), iteration = steps taken,
 * time = steps * dt. */
int osr_sim_save_field(osr_sim *sim, int quant, const char *ms_dir, const double *xmin);

/* Initial fields of species sid (init_fields, spec/os-spec-initfields.f90:38-100), 2D.
 * The species' charge density (folded like the reference's update_boundary(rho, add)) is
 * turned into a field and added.  ufl1 is the species' fluid momentum along x
 * (udist%ufl(1)): 0 selects es_solver_2d (emf/os-emf-es-solver.f03:125-236), the
 * free-space direct sum (charge * r / r^2, FP64 terms, FP32 accumulation in source order,
 * O(cells^2) work) added to E; otherwise es_solver_2d_beam (:457-616), the sheet field of
 * a beam drifting along x, added to E2 and B3.  Call it before the first step, as the
 * reference does at species set-up. */
int osr_sim_init_fields(osr_sim *sim, int sid, float ufl1);

/* ---------------------------------------------------------------------------------
 * 7. Laser injection from an x wall (zpulse_wall, zpulse/os-zpulse-wall.f03), 2D and 3D
 *
 * Replaces the zpulse_wall section of an input deck (read_input_zpulse_wall,
 * os-zpulse-wall.f03:90-628) for direction 1.  Before every push (os-simulation.f03:
 * 528-531, t = n dt, os-time.f90:258-276) each active pulse adds a column of B2, B3 at
 * x = 1 (forward) or x = nx (backward) (launch_em_wall_2d / 3d, os-zpulse-wall.f03:
 * 865-960, 965-1082; in 3D a plane of values at (y, z+1/2) and (y+1/2, z));
 * a circularly polarised pulse (pol_type = +-1) adds a second column with the phase
 * shifted by +-90 degrees and the polarisation by 90 degrees (:633-700).  Temporal envelopes "polynomial",
 * "sin2" and "gaussian" (t_env_zpulse_wall :705-746, os-math.f90:83-165); transverse
 * profiles "plane" and "gaussian"/"hermite" TEM modes 0..7 (per_envelope_2d_zpulse /
 * 3d, zpulse/os-zpulse-std.f03:2599-2753, 2760-3049, hermite_double os-math.f90:
 * 478-509).  Chirps, tilts, math-function envelopes, astigmatic and Laguerre beams and
 * other directions are not available.
 * The launch wall must be open (osr_sim_desc.xbnd = OSR_BND_LINDMAN or OSR_BND_VPML), as the reference
 * requires (:600-608).  The FP64 amplitudes are evaluated on the host, a block of steps
 * at a time, and added in FP32 on the device.
 * ------------------------------------------------------------------------------- */
enum { OSR_TENV_POLYNOMIAL = 0, OSR_TENV_SIN2 = 1, OSR_TENV_GAUSSIAN = 2 };
enum { OSR_PER_PLANE = 0, OSR_PER_HERMITE = 1 };
enum { OSR_PROP_FORWARD = 0, OSR_PROP_BACKWARD = 1 };
/* "unset" marker of per_center, per_focus and tenv_range (the reference's -huge(1.0)) */
#define OSR_ZP_UNSET (-1.7976931348623157e308)

typedef struct {
  double a0, omega0;      /* normalised vector potential, frequency (1.0, 10.0) */
  double phase;           /* degrees (0) */
  double pol;             /* polarisation angle, degrees (90) */
  int pol_type;           /* 0 linear, +1 / -1 circular (0) */
  int propagation;        /* OSR_PROP_* (forward: from x-min) */
  int tenv_type;          /* OSR_TENV_* (polynomial) */
  int per_type;           /* OSR_PER_* (plane) */
  double tenv_fwhm;       /* > 0 overrides rise/flat/fall or duration (-1) */
  double tenv_rise, tenv_flat, tenv_fall;  /* (0, 0, 0) */
  double tenv_duration, tenv_range;        /* gaussian envelope (0, unset) */
  double per_center[2];   /* transverse centre, y (and z) (unset: middle of the box) */
  double per_w0, per_fwhm, per_focus;      /* gaussian beam waist, fwhm, focus (0,0,unset) */
  int per_tem_mode[2];    /* Hermite modes 0..7 along y (and z) (0, 0) */
  double launch_time;     /* (0) */
  double xmin[3];         /* lower corner of the box (g_space xmin) */
} osr_zpulse_wall;

/* The reference's defaults (os-zpulse-wall.f03:196-239), box corner at the origin. */
void osr_zpulse_wall_defaults(osr_zpulse_wall *zp);
/* Adds a pulse to the simulation (at most 8; they launch in the order added). */
int osr_sim_add_zpulse_wall(osr_sim *sim, const osr_zpulse_wall *zp);
/* The FP32 values pulse k adds at time t (n dt) to B2 and B3 at its column, rows
 * y = 1..ny (y fastest, then z = 1..nz in 3D), for its first (second, circular only)
 * launch: b12[ny nz], b3[ny nz]; *active = 0 when it adds nothing at t.  Host code; used
 * to check the injection. */
int osr_sim_zpulse_column(osr_sim *sim, int k, int second, double t, float *b12, float *b3,
                          int *active);

/* In-box laser pulse: the zpulse input section (zpulse/os-zpulse-std.f03,
 * read_input_zpulse :157-805) for a pulse along x1 (direction 1), without boost, tilt,
 * chirp or math envelopes.  The pulse is launched at once (launch_zpulse :2350-2469,
 * launch_time 0): its field (launch_zpulse_2d / 3d :1072-1387) is divergence-corrected
 * unless plane or no_div_corr (div_corr_zpulse_2d / 3d :1681-2205) and added to E and B;
 * pol_type != 0 adds a second pulse shifted by 90 degrees in phase and polarisation.
 * The reference launches it in the loop of the first step, before the push
 * (os-simulation.f03:528-531, t >= launch_time = 0).  Defaults in (), as the
 * reference's. */
typedef struct {
  double a0, omega0;      /* (1.0, 10.0) */
  double phase;           /* degrees (0) */
  double pol;             /* polarisation angle, degrees (90) */
  int pol_type;           /* 0 linear, +1 / -1 circular (0) */
  int propagation;        /* OSR_PROP_* (forward: towards +x) */
  int lon_type;           /* OSR_TENV_* longitudinal envelope (polynomial) */
  int per_type;           /* OSR_PER_* (plane) */
  double lon_start;       /* front of a polynomial / sin2 envelope (unset) */
  double lon_fwhm;        /* > 0 overrides rise/flat/fall or duration (-1) */
  double lon_rise, lon_flat, lon_fall;      /* (0, 0, 0) */
  double lon_duration, lon_x0, lon_range;   /* gaussian envelope (0, unset, unset) */
  double per_center[2];   /* transverse centre, y (and z) (unset: middle of the box) */
  double per_w0, per_fwhm, per_focus;       /* gaussian beam waist, fwhm, focus (0,0,unset) */
  int per_tem_mode[2];    /* Hermite modes 0..7 along y (and z) (0, 0) */
  int no_div_corr;        /* 1: no longitudinal divergence correction (0) */
  double xmin[3];         /* lower corner of the box (g_space xmin) */
} osr_zpulse;

/* The reference's defaults (os-zpulse-std.f03:246-302), box corner at the origin. */
void osr_zpulse_defaults(osr_zpulse *zp);
/* Launches the pulse into the simulation's fields now (before the first step).  With open
 * x walls the pulse's values on the x guard cells go to the guard band as well. */
int osr_sim_add_zpulse(osr_sim *sim, const osr_zpulse *zp);

/* ---------------------------------------------------------------------------------
 * Measurement helper (no counterpart in the reference): streams `bytes` bytes through a
 * plain device copy kernel `reps` times and returns the mean kernel time in *ms; used
 * to calibrate hardware traffic counters (FETCH_SIZE / WRITE_SIZE) and as a bandwidth
 * probe.  Allocates and frees its own buffers.
 * ------------------------------------------------------------------------------- */
int osr_copy_probe(int64_t bytes, int reps, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* OSIRIS_AMD_H */
