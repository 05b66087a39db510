/*
 * m3d.h -- C-ABI of the MI355X FFT homogenization solver (libmat3d.so).
 *
 * Drop-in boundary for the reference's two solver classes.  Each entry point
 * below replaces one public method; the reference interface it stands for is
 * cited as file:line relative to the reference tree.  Plain pointers and
 * sizes only; arrays use the reference's memory order so Eigen callers pass
 * .data() unchanged:
 *   - real fields: C x N column-major (C = 6 elastic, 3 thermal), i.e. the
 *     C components of voxel i are contiguous, voxel i = z + Nz*(y + Ny*x);
 *   - material tables: nph x 7 column-major [E, nu, sigma_y, H, n, P, Y0]
 *     (example/3d/main.cpp:104-111);
 *   - 6x6 / 3x3 results: column-major.
 *
 * Errors: every function returns M3D_OK (0) or a negative M3D_ERR_* code and
 * never exits, throws or aborts across the ABI; m3d_*_last_error() returns
 * a message.  Handles are independent (no global cleanup; compare the
 * reference destructor's fftw_cleanup, src/3d/FFT_analysis3D.cpp:1141-1142).
 * One handle must not be used from two threads at once.
 */
#ifndef M3D_H
#define M3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M3D_OK 0
#define M3D_ERR_ARG -1         /* invalid argument (sizes, phase ids, behaviour) */
#define M3D_ERR_UNSUPPORTED -2 /* grid size with a prime factor > 13, > 256 phases */
#define M3D_ERR_OOM -3         /* device allocation failed */
#define M3D_ERR_HIP -4         /* HIP runtime error */
#define M3D_ERR_STATE -5       /* call out of order (e.g. run before material) */
#define M3D_ERR_NUMERIC -6     /* singular 6x6 / 3x3 or non-finite state */

/* precision policy */
#define M3D_FP64 0            /* fields, FFTs and reductions in FP64 */
#define M3D_FP32_POLISH 1     /* FP32 fields/FFTs, FP64 reductions, FP64 polish to the stop rule */

typedef struct m3d_elastic m3d_elastic;
typedef struct m3d_thermal m3d_thermal;

/* per-solve outcome of a run (no counterpart in the reference, which prints
 * Itter/eqVar/diff_sum_field to stdout every iteration: FFT_analysis3D.cpp:394-412) */
typedef struct {
    int iterations;     /* itNo at exit */
    int status;         /* 1 stop rule met, 2 iteration cap reached */
    int iters_fp32;     /* iterations executed in FP32 (0 under M3D_FP64) */
    int iters_fp64;     /* iterations executed in FP64 */
    double eqVar;       /* last evaluated avr_eqVar */
    double diff_sum;    /* last diff_sum_field */
    double diff_eq;     /* last diff_avr_eqVar (only evaluated when diff_sum < tol_field) */
    double time_ms;     /* device time of the solve loop */
} m3d_run_info;

/* ---------------------------------------------------------------- elastic */

/* FFT_Analysis3D(int Nx, int Ny, int Nz, string mater_behav)
 * -- include/FFT_analysis3D.hpp:196, src/3d/FFT_analysis3D.cpp:25-152.
 * behav: "e" (accelerated scheme), "ep" (J2 elastoplastic) or "ed" (damage);
 * "ep"/"ed" runAnalysis runs the basic scheme in FP64 whatever `precision`. */
int m3d_elastic_create(int nx, int ny, int nz, const char *behav, int precision, m3d_elastic **out);

/* ~FFT_Analysis3D() -- src/3d/FFT_analysis3D.cpp:1107-1147 */
int m3d_elastic_destroy(m3d_elastic *h);

/* void material(MatrixDdd& mat, MatrixIdd& RVE, int noPhases)
 * -- include/FFT_analysis3D.hpp:200, src/3d/FFT_analysis3D.cpp:420-549.
 * mat: nph x 7 column-major; rve: N phase ids in [0, nph). */
int m3d_elastic_material(m3d_elastic *h, const double *mat, const int32_t *rve, int nph);

/* void setLoad(MatrixDdd& E, MatrixDdd& X0)
 * -- include/FFT_analysis3D.hpp:202, src/3d/FFT_analysis3D.cpp:551-564. */
int m3d_elastic_set_load(m3d_elastic *h, const double *E, const double *X0);

/* void runAnalysis(double tol_field, double tol_eq, int nInc)
 * -- include/FFT_analysis3D.hpp:206, src/3d/FFT_analysis3D.cpp:749-854.
 * "e": accelerated scheme to the stop rule (nInc ignored, as in the
 * reference).  "ep"/"ed": nInc increments of the basic scheme with mixed
 * control along X0 (updateE); info (may be NULL) describes the last increment,
 * iters_fp64 the total. */
int m3d_elastic_run(m3d_elastic *h, double tol_field, double tol_eq, int nInc, m3d_run_info *info);

/* public member prop_homogenized -- include/FFT_analysis3D.hpp:156, filled by
 * export_homogen_var (src/3d/FFT_analysis3D.cpp:722-746) during an "ep"/"ed"
 * runAnalysis: rows x 9 column-major [eps_t, sigma_1..6, mean sigma_eff (ep)
 * or d (ed), mean eps_p (ep) or 0].  prop may be NULL to query *rows. */
int m3d_elastic_get_prop(m3d_elastic *h, double *prop, int *rows);

/* public members epsilon_p_eff and sigma_eff ("ep"; sig_eff is zero for
 * "ed") or d ("ed") after runAnalysis, N values each (either pointer may be
 * NULL) -- FFT_analysis3D.hpp:56-98, read by printer / vtiWriter
 * (src/3d/FFT_analysis3D.cpp:1001-1104) */
int m3d_elastic_get_state(m3d_elastic *h, double *ep_eff_or_d, double *sig_eff);

/* double homogenization(double tol_field, double tol_eq, string out)
 * -- include/FFT_analysis3D.hpp:209, src/3d/FFT_analysis3D.cpp:856-999.
 * The six load cases run as one batched launch.  Outputs (column-major
 * 6x6, any may be NULL): CHomo, SHomo, epHomo, sigHomo; *ret =
 * 1/S11 + 1/S22 + 1/S33 (the reference's return value); info6: 6 entries. */
int m3d_elastic_homogenization(m3d_elastic *h, double tol_field, double tol_eq, double *C, double *S,
                               double *epHomo, double *sigHomo, double *ret, m3d_run_info *info6);

/* public members epsilon / sigma (6 x N) -- include/FFT_analysis3D.hpp:98-99.
 * Copies the fields of the last run (for homogenization: load case `lc`,
 * 0..5; for run: lc = 0) in the reference's memory order. */
int m3d_elastic_get_fields(m3d_elastic *h, int lc, double *eps, double *sig);

/* public member vol_fractions and the chosen algorithm (strain/stress base),
 * include/FFT_analysis3D.hpp:58-59,141.  vol: nph entries (may be NULL);
 * strain_base: 1 strain-based, 0 stress-based. */
int m3d_elastic_info(m3d_elastic *h, double *vol, int *strain_base, double *la0, double *mu0);

/* Tuning knobs (no reference counterpart): key "polish_factor" (the FP32
 * phase stops at max(tol*factor, fp32_floor), default 10), "fp32_floor"
 * (default 1e-6), "rollback" (FP32 iterations redone in FP64 when the switch
 * iterate already satisfies the stop rule within 2 x tol, default 5), "cap"
 * (iteration cap, default 2000 elastic / 500 thermal), "chunk" (iterations
 * per graph replay, 0 = automatic), "refine" (1, default: the polish iterates
 * an FP32 correction of one FP64 anchor with FP64 stop-rule sums and checks;
 * 0: plain FP64 polish iterations), "scheme" (0, default: the reference's
 * accelerated "e" loop; 1: the basic Moulinec-Suquet iteration with the load
 * held fixed, for homogenization and the "e" runAnalysis), "basic_fp32" (0,
 * default: "ep"/"ed" runs iterate in FP64; 1 on an FP32+polish "ep" handle:
 * each increment iterates in FP32 -- fields, FFTs, return map -- to
 * polish_factor x tol, then in FP64 under the reference rule; the history
 * stays FP64.  Plasticity carries the FP32 rounding of the early iterations
 * into the final fields: about 5e-5 of the largest stress at polish_factor
 * 10, 1e-5 at 1000). */
int m3d_elastic_set_option(m3d_elastic *h, const char *key, double value);

/* Kernel timing for measurement tools (no reference counterpart).  Times
 * `reps` standalone launches of each per-iteration kernel of the given
 * precision over the batch of the last run/homogenization, with HIP events
 * on the handle's stream.  ms6 = average ms per launch of {z pass with
 * update, y pass forward, x pass with Green step, y pass inverse, z pass of
 * the check FFT, x pass of the check FFT}; geom9 (may be NULL) =
 * {Nx, Ny, Nz, Nz/2+1, C, solves, TY, TKy, TKx}.  Overwrites the fields. */
int m3d_elastic_profile(m3d_elastic *h, int precision, int reps, double *ms6, int *geom9);

/* Variants of the FP32 x/y passes the autotuner chose for the last solve's
 * shape (no reference counterpart; it stands where the reference plans its
 * FFTs, FFT_analysis3D.cpp:101-150): threads per workgroup of the x pass with
 * the Green step, its wide kz tile (1) or not (0), threads of the y pass, the
 * radix plan of each (0: radix 16 first, 1: radix 8 first) and whether the
 * choice came from the tuning file (1) or was timed in this process (0).
 * Option "autotune" (default 1) times the variants at the first FP32 solve
 * of a (device, grid, solves) shape; every pointer may be NULL. */
int m3d_elastic_get_tuning(m3d_elastic *h, int *xg_threads, int *xg_wide, int *y_threads, int *xg_radix, int *y_radix,
                           int *from_file);

/* Every tuning choice this process timed or read from the tuning file, one
 * line each, into buf (len bytes, NUL-terminated, truncated to fit; *needed =
 * bytes of the whole report).  No reference counterpart. */
int m3d_tuning_report(char *buf, int len, int *needed);

/* Path of the tuning file: $M3D_TUNE_FILE, else $XDG_CACHE_HOME or
 * $HOME/.cache, then /mat3d_amd/tune.txt; "" when M3D_TUNE_FILE is "none"
 * (choices kept in the process only).  No reference counterpart. */
const char *m3d_tuning_file(void);

const char *m3d_elastic_last_error(m3d_elastic *h);

/* Device-to-device copy rate for measurement tools (no reference
 * counterpart): one fill and one untimed copy of gib GiB, then `reps` timed
 * copies with the runtime's copy kernel (HIP events on the null stream);
 * ms_per_copy = average ms per copy (each reads and writes gib GiB).  Also
 * the known-size traffic that calibrates the HBM byte counters. */
int m3d_device_copy(double gib, int reps, double *ms_per_copy);

/* ---------------------------------------------------------------- thermal */

/* FFT_analysis3d_thermal(int Nx, int Ny, int Nz)
 * -- include/FFT_thermal3D.hpp:133, src/3dTherm/FFT_thermal3D.cpp:14-91 */
int m3d_thermal_create(int nx, int ny, int nz, int precision, m3d_thermal **out);

int m3d_thermal_destroy(m3d_thermal *h);

/* void material(double km, double ki, MatrixIdd& RVE)
 * -- include/FFT_thermal3D.hpp:138, src/3dTherm/FFT_thermal3D.cpp:373-422.
 * RVE != 0 -> matrix (km), RVE == 0 -> inclusion (ki). */
int m3d_thermal_material(m3d_thermal *h, double km, double ki, const int32_t *rve);

/* N-phase extension of material(): conductivity k[p] for phase id p.  No
 * reference counterpart (the reference supports two phases only,
 * FFT_thermal3D.cpp:186-189); rule: k0 = sqrt(kmin*kmax), gradient base iff
 * the first max-k phase has the largest volume fraction. */
int m3d_thermal_material_n(m3d_thermal *h, const double *k, int nph, const int32_t *rve);

/* void setLoad(MatrixDdd& E) -- include/FFT_thermal3D.hpp:140, FFT_thermal3D.cpp:426-441 */
int m3d_thermal_set_load(m3d_thermal *h, const double *E);

/* void runAnalysis(double tol_field, double tol_eq)
 * -- include/FFT_thermal3D.hpp:145, FFT_thermal3D.cpp:446-485 */
int m3d_thermal_run(m3d_thermal *h, double tol_field, double tol_eq, m3d_run_info *info);

/* void homogenization(double, double, string) -- include/FFT_thermal3D.hpp:148,
 * FFT_thermal3D.cpp:489-622.  Outputs 3x3 column-major: KHomo, RHomo,
 * tempGradHomo, heatFluxHomo (any may be NULL); info3: 3 entries. */
int m3d_thermal_homogenization(m3d_thermal *h, double tol_field, double tol_eq, double *K, double *R,
                               double *tempGradHomo, double *heatFluxHomo, m3d_run_info *info3);

/* public members tempGradient / heatFlux (3 x N) -- include/FFT_thermal3D.hpp:82,92 */
int m3d_thermal_get_fields(m3d_thermal *h, int lc, double *grad, double *flux);

int m3d_thermal_set_option(m3d_thermal *h, const char *key, double value);

const char *m3d_thermal_last_error(m3d_thermal *h);

/* ------------------------------------------------------------------ batch */

/* Many independent elastic homogenizations (RVEs of any mix of grid sizes).
 * The reference has no batch entry point: its callers loop over RVEs,
 * constructing one FFT_Analysis3D per RVE and calling material() and
 * homogenization() on each (example/3d/main.cpp:117-128,
 * src/3d/FFT_analysis3D.cpp:420-549,856-999); results here are those of
 * that loop.  The library groups the problems by grid shape and runs each
 * group's 6 x RVEs load cases as one batch of solves on the device. */
typedef struct m3d_batch m3d_batch;

int m3d_batch_create(int precision, m3d_batch **out);

int m3d_batch_destroy(m3d_batch *b);

/* Adds one RVE: material() of a behaviour-"e" FFT_Analysis3D(nx, ny, nz)
 * (FFT_analysis3D.cpp:420-549).  mat: nph x 7 column-major as in
 * m3d_elastic_material; rve: N phase ids, z fastest.  *index receives the
 * problem's position (0, 1, 2, ...). */
int m3d_batch_add(m3d_batch *b, int nx, int ny, int nz, const double *mat, const int32_t *rve, int nph,
                  int *index);

/* homogenization(tol_field, tol_eq, out) of every added RVE
 * (FFT_analysis3D.cpp:856-999). */
int m3d_batch_homogenization(m3d_batch *b, double tol_field, double tol_eq);

/* Results of problem `index` after m3d_batch_homogenization, as
 * m3d_elastic_homogenization returns them (any pointer may be NULL). */
int m3d_batch_results(m3d_batch *b, int index, double *C, double *S, double *epHomo, double *sigHomo, double *ret,
                      m3d_run_info *info6);

/* Problem count, device time (ms) of the last m3d_batch_homogenization and
 * its voxel-iterations (sum over solves of N * iterations). */
int m3d_batch_stats(m3d_batch *b, int *count, double *ms, double *voxel_iterations);

/* options as m3d_elastic_set_option, plus "max_solves": upper bound on the
 * solves of one device batch (0 = as many as fit in 70% of free memory). */
int m3d_batch_set_option(m3d_batch *b, const char *key, double value);

const char *m3d_batch_last_error(m3d_batch *b);

/* library version string */
const char *m3d_version(void);

/* ---------------------------------------------------------------------
 * Phase-retrieval reconstruction (src/rec/reconstruction.cpp), FP64.
 * ------------------------------------------------------------------- */
typedef struct m3d_rec m3d_rec;

/* recPlan(double *TPCF, int Nx1, int Ny1, int Nz1) -- include/reconstruction.hpp:55,
 * src/rec/reconstruction.cpp:211-243.  tpcf: Nx*Ny*Nz values, index
 * (x*Ny + y)*Nz + z (the reference's k + j*Nz + i*Nz*Ny); normC =
 * sqrt(|tpcf| * Nx*Ny*Nz).  Each length must be 32, 48, 64, 96, 128, 192 or
 * 256 (M3D_ERR_UNSUPPORTED otherwise). */
int m3d_rec_create(const double *tpcf, int nx, int ny, int nz, m3d_rec **out);
int m3d_rec_destroy(m3d_rec *r);
const char *m3d_rec_last_error(m3d_rec *r);

/* Start field in place of the random one (no reference counterpart; for
 * tests and warm starts): Nx*Ny*Nz real values. */
int m3d_rec_set_initial(m3d_rec *r, const double *m0);

/* runRec(int n, double b) -- include/reconstruction.hpp:66,
 * src/rec/reconstruction.cpp:285-406.  The reference draws its random
 * phases with Eigen's setRandom; here value i of stream k is
 * splitmix64(seed + (k << 40) + i) mapped to [-1, 1): stream 0 starts the
 * field ((u+1)/2), stream k is the k-th re-randomisation (every 150
 * iterations).  iterations = entries of trackrealerrorA written; status 1 =
 * error below 0.1, 2 = iteration number exceeded; search_steps = steps of
 * the final volume-fraction search. */
int m3d_rec_run(m3d_rec *r, int n, double beta, uint64_t seed, int *iterations, int *status, int *search_steps);

/* public mTilda (interleaved re, im; Nx*Ny*Nz complex) and trackrealerrorA
 * (up to maxn values; nerr = how many the last run wrote) --
 * include/reconstruction.hpp:73-74. */
int m3d_rec_result(m3d_rec *r, double *mtilda_ri, double *err, int maxn, int *nerr);

/* ---------------------------------------------------------------------
 * Voxel work of the seeded structure generators (mat3d_amd/gen.py; no
 * reference counterpart: the reference reads RVE files).  The random draws
 * and placement decisions stay with the caller; these rasterise on the GPU
 * with the generators' own tests and operation order, so the phase ids are
 * theirs bit for bit.  ids: Nx*Ny*Nz int32, index (x*Ny + y)*Nz + z.
 * ------------------------------------------------------------------- */
const char *m3d_gen_last_error(void);

/* ids = fill, then every ball {cx, cy, cz, r, phase} (5 int32 each) stamps
 * its phase on the voxels c + o with |o|^2 <= r^2 (periodic).  Balls must not
 * share voxels (random sequential addition guarantees it). */
int m3d_gen_balls(int nx, int ny, int nz, const int32_t *balls, int nballs, int32_t fill, int32_t *ids);

/* Straight finite cylinders stamped onto ids (in/out) in batches of `batch`
 * until the count of `incl` voxels reaches `target`: cylinder k = {axis v
 * (unit, 3), centre c (3), box lo (3), box hi (3)} (12 doubles); voxel a of the
 * box with X = a - c, t = X.v: |t| <= half and |X|^2 - t^2 <= radius^2.
 * used = cylinders stamped, count = final count (stops early when the list
 * runs out: the caller draws more and calls again). */
int m3d_gen_cylinders(int nx, int ny, int nz, const double *cyl, int ncyl, double half, double radius, int32_t incl,
                      int batch, long long target, int32_t *ids, int *used, long long *count);

/* No counterpart in the reference.  m3d_gen_cylinders for callers that cannot
 * allocate device memory (code under HIP graph capture, tight loops, harnesses
 * that forbid hipMalloc): it calls no hipMalloc/hipFree.  ids_dev is the
 * caller's DEVICE array of nx*ny*nz int32 (in/out, voxel i = z + nz*(y + ny*x));
 * cyl stays a host array, staged into a fixed device buffer the library holds
 * per GPU for the life of the process (8192 cylinders at most per call, more
 * returns M3D_ERR_ARG; the per-workgroup counts live in the same buffer).
 * Work runs on the null stream and the call returns when it is done.  Calls
 * on one GPU share that buffer: calls that overlap on one GPU (other host
 * threads) race. */
int m3d_gen_cylinders_noalloc(int nx, int ny, int nz, const double *cyl, int ncyl, double half, double radius,
                              int32_t incl, int batch, long long target, int32_t *ids_dev, int *used, long long *count);

/* Periodic Voronoi cells: ids = index of the nearest of the g seeds (3
 * doubles each, inside the box) under the minimum-image distance, the first
 * seed on exact ties. */
int m3d_gen_voronoi(int nx, int ny, int nz, const double *pts, int g, int32_t *ids);

/* ---------------------------------------------------------------------
 * FFT_Analysis2D (include/FFT_analysis2D.hpp, src/2d/FFT_analysis2D.cpp),
 * FP64: basic scheme on complex 2D fields [E11, E22, E12] (tensor shear),
 * 25 fixed iterations per run.  Fields: Nx*Ny points, index y + x*Ny, three
 * interleaved complex components (re, im).  Nx and Ny: 32, 48, 64, 96, 128,
 * 192, 256 or 512.
 * ------------------------------------------------------------------- */
typedef struct m3d_2d m3d_2d;

/* FFT_Analysis2D(int Nx1, int Ny1) -- FFT_analysis2D.hpp:82, .cpp:3-51.  The
 * fields start at zero (the reference leaves them uninitialised). */
int m3d_2d_create(int nx, int ny, m3d_2d **out);
int m3d_2d_destroy(m3d_2d *h);
const char *m3d_2d_last_error(m3d_2d *h);

/* material(double *properties, MatrixDdd &RVE1) -- .hpp:89, .cpp:189-221:
 * properties {E matrix, nu matrix, E inclusion, nu inclusion}; rve Nx*Ny
 * values, 1 = matrix, anything else = inclusion. */
int m3d_2d_material(m3d_2d *h, const double *properties, const double *rve);

/* setLoad(MatrixCdd &E1) -- .hpp:92, .cpp:223-230: E (3 complex, re/im). */
int m3d_2d_set_load(m3d_2d *h, const double *E_ri);

/* runAnalysis() -- .hpp:95, .cpp:232-246: 25 iterations from the current
 * fields; itNo (may be NULL) = iterations run by this handle so far. */
int m3d_2d_run(m3d_2d *h, int *itNo);

/* homogenization() -- .hpp:97, .cpp:248-272: three runs with E(w) = 0.005
 * from the current fields (the reference does not reset them), then CHomo
 * = epHomo^-1 sigHomo; all 3x3 column-major (may be NULL).  The reference's
 * report file "Homogenized_Properties" is not written. */
int m3d_2d_homogenization(m3d_2d *h, double *CHomo, double *epHomo, double *sigHomo);

/* public epsilon / sigma (3 x Nx*Ny complex) -- .hpp:28-29 */
int m3d_2d_get_fields(m3d_2d *h, double *eps_ri, double *sig_ri);

/* ---------------------------------------------------------------------
 * percolation (include/percolation.hpp, src/percolation/percolation.cpp):
 * clusters of one phase and the face contacts of two phases.  rve: Nx*Ny*Nz
 * int32 phase ids, index (x*Ny + y)*Nz + z (the reference's C1[k + j*NZ +
 * i*NY*NZ]).  Clusters: the interest voxels joined through every 2x2x2 cube
 * (largest corner (x, y, z)) for which one of the six index permutations of
 * (x, y, z) is an interest voxel -- the fixed point of the reference's
 * min-label sweeps (:117-286), which its 5 x 8 x 6 sweeps reach on all but
 * long winding clusters.
 * ------------------------------------------------------------------- */
typedef struct m3d_perc m3d_perc;

/* percolation(MatrixDdd &RVE, Nx1, Ny1, Nz1, noOfPhases, intrestPhase) --
 * percolation.hpp:39, .cpp:7-30.  The field is copied to the GPU. */
int m3d_perc_create(const int32_t *rve, int nx, int ny, int nz, int noOfPhases, int intrestPhase, m3d_perc **out);
int m3d_perc_destroy(m3d_perc *h);
const char *m3d_perc_last_error(m3d_perc *h);

/* runPercolationAnalysis(int total_iter) -- .hpp:42, .cpp:117-382
 * (total_iter is unused there and here).  Needs Nx = Ny = Nz, and every
 * phase id below noOfPhases + 1 (M3D_ERR_ARG otherwise: such ids read as
 * cluster labels in the reference).  nclusters = all_clusters.size();
 * vol_intrstd = vol_intrstd_phs (interest voxels; the reference adds them up
 * again on every call, this returns one call's count). */
int m3d_perc_run(m3d_perc *h, int total_iter, int *nclusters, double *vol_intrstd);

/* C1 after the analysis (interest voxels = noOfPhases + 1 + rank of their
 * cluster by size, equal sizes equal rank; other voxels their id) or after
 * SurfaceOfTwoPhase (the field given there); all_clusters -- .hpp:59: the
 * clusters' labels (noOfPhases + 1 + raster rank of their first voxel among
 * the interest voxels) in first-appearance order; clusters_vol_frac -- the
 * clusters' voxel counts / vol_intrstd_phs, descending.  Each may be NULL;
 * at most maxn clusters are written. */
int m3d_perc_result(m3d_perc *h, int32_t *c1, int32_t *all_clusters, double *clusters_vol_frac, int maxn);

/* SurfaceOfTwoPhase(MatrixDdd &RVE, phase1, phase2) -- .hpp:47, .cpp:384-414:
 * voxels of phase1 times their -x, -y and -z face neighbours of phase2 (-1
 * outside the box); the field replaces the handle's, as it replaces C1.
 * Counted in 64 bits (the reference's int). */
int m3d_perc_surface(m3d_perc *h, const int32_t *rve, int phase1, int phase2, long long *commonSurface);

/* No reference counterpart: stage durations (ms) of the next runs, from
 * events on the handle's stream: init, link, enumerate, sizes, sort,
 * relabel (6 values). */
int m3d_perc_set_timing(m3d_perc *h, int on);
int m3d_perc_timings(m3d_perc *h, double *ms, int n);

#ifdef __cplusplus
}
#endif
#endif /* M3D_H */
