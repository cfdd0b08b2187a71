/*
 * specsim3d.h -- C ABI of the MI355X spectral-simulation engine (libspecsim3d.so).
 *
 * The reference program (src/specsim3d.f90) has no in-process API: its hot path is the inline
 * block of the main program f90:1234-1345 calling module sim3d_mod (f90:89-304).  Each entry
 * point below replaces one piece of that block; the Fortran driver (specsim3d_amd/fortran)
 * binds them with ISO_C_BINDING and keeps the reference's parameter file, stdout and output
 * files.
 *
 * Conventions
 *  - Plain C types only.  Every function returns SS3D_OK (0) or a positive SS3D_E* code; the
 *    message of the last failure on the calling thread is available from ss3d_last_error().
 *  - Arrays follow the reference's Fortran layout: x fastest, then y, then z, then realization.
 *  - The library owns device memory (spectra, work buffers); callers own host arrays.  Entry
 *    points that take device pointers say so; those pointers must come from hipMalloc on the
 *    context's device.
 *  - A context is used by one host thread at a time; work is ordered on the context's stream.
 */
#ifndef SPECSIM3D_H
#define SPECSIM3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS3D_OK 0
#define SS3D_EINVAL 1   /* invalid argument (message says which) */
#define SS3D_ETOOBIG 2  /* grid or axis beyond what the engine supports */
#define SS3D_ENOMEM 3   /* device or host allocation failed */
#define SS3D_EHIP 4     /* HIP runtime error */
#define SS3D_ENODEV 5   /* no usable GPU */
#define SS3D_EIO 6      /* an output file could not be opened or written (message names it) */

/* arithmetic of the per-realization path */
#define SS3D_FP64 0 /* drop-in: C2C x->y->z in FP64, reference butterfly order (f90:1335) */
#define SS3D_FP32 1 /* fast: FP32 half-spectrum z->y passes + C2R x pass, FP64 statistics */

/* One nested structure: the scalar members of `type structure` (f90:95-100).  The rotation
 * (setrot, f90:107-139) is recomputed inside the library.  Angles in degrees. */
typedef struct ss3d_structure {
    int32_t itype; /* 1 spherical, 2 exponential, 3 gaussian */
    int32_t pad_;
    double sill, ahmax, ahmin, avert, az, dip, plg;
} ss3d_structure;

/* Spectral-density diagnostics: the values the reference prints at f90:1289-1292. */
typedef struct ss3d_spec_diag {
    double max_imag; /* max |Im S| */
    double max_real; /* max |Re S| */
    double c0;       /* C(0), the covariance at lag 0 (f90:1270) */
    double spos;     /* sum of the non-negative Re S */
    double sneg;     /* sum of the negative Re S (clipped to zero amplitude) */
} ss3d_spec_diag;

typedef struct ss3d_context ss3d_context;
typedef struct ss3d_spectrum ss3d_spectrum;

/* ABI version, major * 10000 + minor * 100: 1.1 adds the output writer, spectrum files and
 * conditioning to 1.0 (grid sizing, spectra, realizations, variograms, request loop, batch graph);
 * 1.2 adds the allocation-free realization entry point and the tile entries of the tuning table;
 * 1.3 adds the options SS3D_OPT_ZRUN and SS3D_OPT_SPEC_FP32. */
#define SS3D_VERSION 10300
int ss3d_version(void);

/* Copies the last error message of the calling thread (NUL-terminated, truncated to len). */
int ss3d_last_error(char* buf, int64_t len);

/* Opens a context on HIP device `device` (stream, twiddle tables, work-buffer pool). */
int ss3d_context_create(int device, ss3d_context** out);
int ss3d_context_destroy(ss3d_context* ctx);
int ss3d_context_synchronize(ss3d_context* ctx);

/* Periodic grid size (replaces f90:1234-1252 with cutoff f90:142 and nextpow2 f90:181).
 * ipad=0: next power of two >= n; ipad=1: nextpow2(max(n+r, 2r)) with r the covariance reach
 * in cells; nz=1 keeps npz=1.  Unlike the reference (f90:1253) no 2^27 cap is applied here:
 * the caller decides (the Fortran driver keeps the reference's cap unless told otherwise). */
int ss3d_periodic_grid(const ss3d_structure* st, int nst, double dx, double dy, double dz, int nx, int ny,
                       int nz, int ipad, int* npx, int* npy, int* npz);

/* Spectral density on the device (replaces f90:1256-1288 with covariance f90:153 and the
 * forward fft3d f90:254): covariance sampled on the periodic grid with minimum-image lags,
 * forward FP64 transform, S = Re/N, amplitude sqrt(max(S,0)).  Keeps the FP64 amplitude (for
 * SS3D_FP64 realizations) and the FP32 half-spectrum amplitude (for SS3D_FP32).  The handle
 * caches the random-stream tables of the grid.  diag may be NULL. */
int ss3d_spectrum_create(ss3d_context* ctx, const ss3d_structure* st, int nst, double nugget, double dx,
                         double dy, double dz, int npx, int npy, int npz, ss3d_spectrum** out,
                         ss3d_spec_diag* diag);
int ss3d_spectrum_destroy(ss3d_spectrum* spec);
/* Device bytes held by a spectrum handle. */
int64_t ss3d_spectrum_bytes(const ss3d_spectrum* spec);
/* Copies the FP64 amplitude sqrt(max(S,0)) (npx*npy*npz doubles, x fastest) to host memory. */
int ss3d_spectrum_amplitude(const ss3d_spectrum* spec, double* host_amp);

/* Random stream bookkeeping (f90:191-195, 1302-1303, fill order f90:1305-1334).
 * draws per realization: (N + n_sc)/2 points draw once (igauss=0) or twice (igauss=1). */
int64_t ss3d_draws_per_realization(int npx, int npy, int npz, int igauss);
/* Park-Miller state that realization r (0-based) starts from, seeded as f90:1302-1303. */
int64_t ss3d_realization_state(int iseed, int npx, int npy, int npz, int igauss, int64_t r);

/* The spectral density as a file (no counterpart in the reference, which recomputes it every run):
 * the FP64 amplitudes, the FP32 half spectrum, the diagnostics and the key (structures, nugget,
 * spacing, periodic grid) of a spectrum made by ss3d_spectrum_create.  ss3d_spectrum_load makes a
 * spectrum of ctx from such a file; with st != NULL the file must have been made for exactly that
 * model, spacing and grid (SS3D_EINVAL otherwise); st = NULL takes the file's key.  The payload is
 * checked against the diagnostics (sum of amp^2 = positive spectral mass).  Realizations of a loaded
 * spectrum are identical to those of the spectrum that was saved.  Spectra without FP64 amplitudes
 * (SS3D_OPT_SPEC_FP32) are not saved (SS3D_EINVAL). */
int ss3d_spectrum_save(const ss3d_spectrum* spec, const char* path);
int ss3d_spectrum_load(ss3d_context* ctx, const char* path, const ss3d_structure* st, int nst, double nugget,
                       double dx, double dy, double dz, int npx, int npy, int npz, ss3d_spectrum** out,
                       ss3d_spec_diag* diag);

/* Realizations r0 .. r0+nreal-1 of the stream seeded by iseed (replaces f90:1305-1344):
 * Hermitian coefficient fill, inverse transform, crop of the first nx*ny*nz corner.
 * host_out: nx*ny*nz*nreal values, double for SS3D_FP64 and float for SS3D_FP32.
 * mean, var, maximag: nreal doubles each (may be NULL) -- the per-realization statistics the
 * reference prints at f90:1343 (population variance, max |Im| of the cropped corner). */
int ss3d_realize(ss3d_context* ctx, const ss3d_spectrum* spec, int iseed, int igauss, int64_t r0, int nreal,
                 int nx, int ny, int nz, int precision, void* host_out, double* mean, double* var,
                 double* maximag);

/* Same work with the output left in device memory and no host synchronisation: dev_out holds
 * nx*ny*nz*nreal values (double or float by precision), dev_stats 3*nreal doubles
 * (mean, variance, max|Im| per realization) or NULL.  Ordered on the context stream. */
int ss3d_realize_device(ss3d_context* ctx, const ss3d_spectrum* spec, int iseed, int igauss, int64_t r0,
                        int nreal, int nx, int ny, int nz, int precision, void* dev_out, double* dev_stats);

/* The same work as ss3d_realize_device, with no device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or other allocating routine.
 * Every work buffer of its passes is a window of one static device array of SS3D_SCRATCH_BYTES
 * (256 MiB) that each GPU holds from the library's first use on it until the process ends; the
 * caller still passes the output (and optionally the statistics) buffer.  Results and statistics
 * are those of ss3d_realize_device bit for bit.
 *  - SS3D_ETOOBIG when one realization of the grid needs more scratch than the array holds (the
 *    message gives both sizes); e.g. 128^3 outputs in a 256^3 periodic grid fit (FP32: 3
 *    realizations per launch, FP64: 1), 256^3 in 512^3 does not.
 *  - SS3D_EINVAL when the spectrum's random-stream tables for igauss are not built yet: any one
 *    ss3d_realize / ss3d_realize_device call with that igauss builds them.
 *  - Every call on a GPU uses that GPU's one array (which keeps the last call's contents; no
 *    buffer needs clearing, every word read is written earlier in the same call).  Calls ordered
 *    on one context's stream run one after the other; calls that overlap on one GPU (two contexts,
 *    or two host threads) race.  Ordered on the context stream; no profiling records. */
#define SS3D_SCRATCH_BYTES (256ll << 20)
int ss3d_realize_device_noalloc(ss3d_context* ctx, const ss3d_spectrum* spec, int iseed, int igauss, int64_t r0,
                                int nreal, int nx, int ny, int nz, int precision, void* dev_out, double* dev_stats);

/* Per-pass timing with HIP events recorded on the context stream around every pass launch.
 * Pass kinds: 0..2 spectrum x/y/z (FP64), 3..5 FP64 realization x/y/z, 6..8 FP32 realization
 * z/y/x, 9 FP32 z fill beside the y pass of the previous batch (SS3D_OPT_FUSE).  enable=1 clears and starts recording, enable=0 stops.  profile_read synchronises the
 * stream and returns, per kind, total milliseconds, algorithmic HBM bytes (bytes each pass must
 * read and write) and launch count.  No counterpart in the reference (its timers are the wall
 * clocks printed at f90:1299-1300, 1350-1351). */
#define SS3D_NPASS_KINDS 10
int ss3d_context_profile(ss3d_context* ctx, int enable);
int ss3d_context_profile_read(ss3d_context* ctx, double* ms, double* bytes, int64_t* launches, int nkinds);

/* Autotuner of the FP32 passes (no counterpart in the reference): times every radix plan of
 * each FP32 pass (largest radix 8, 16 or 32) and, for the strided passes, every tile variant
 * (below) on this spectrum's grid and crop with `reps` timed batches per candidate, and keeps the
 * fastest for later calls on the same context (the fill and y passes prefer a pair that has a
 * fused launch).
 * tuning_get returns the plan (log2 of the largest radix) of pass kind 6..8 and length
 * 2^logn; tuning_set pins one (e.g. a plan stored from an earlier run). */
int ss3d_autotune(ss3d_context* ctx, const ss3d_spectrum* spec, int nx, int ny, int nz, int igauss, int reps);
int ss3d_tuning_get(const ss3d_context* ctx, int kind, int logn);
int ss3d_tuning_set(ss3d_context* ctx, int kind, int logn, int rm);
/* The tile of a tuning entry (no counterpart in the reference).  The autotuner also picks a tile
 * variant of the strided passes (kinds 6 and 7): variant 0 is the default tile, variant 1 halves
 * the tile width for lengths up to 2^8 and drops the LDS position padding at length 2^10 (plans 3
 * and 4 only).  The values computed do not depend on the tile.  get_tile returns the plan, the
 * variant, the columns per tile (rows per block for kind 8) and whether the LDS exchange is padded;
 * set_tile pins a plan and variant (SS3D_EINVAL where the variant does not exist).  Fused launches
 * use variant 0. */
int ss3d_tuning_get_tile(const ss3d_context* ctx, int kind, int logn, int* rm, int* variant, int* tile, int* lds_pad);
int ss3d_tuning_set_tile(ss3d_context* ctx, int kind, int logn, int rm, int variant);

/* Context options (no counterpart in the reference).
 * SS3D_OPT_PIPELINE (default 0): FP32 realizations run the fill pass of batch k+1 on a second
 * stream while the remaining passes of batch k run, double-buffering the fill output. */
#define SS3D_OPT_PIPELINE 1
/* SS3D_OPT_FP64_GENERIC (default 0): FP64 realizations through the generic (runtime-length) pass
 * kernel instead of the compile-time kernels; both compute the same values bit for bit. */
#define SS3D_OPT_FP64_GENERIC 2
/* SS3D_OPT_FUSE (default 1): FP32 realizations whose z and y axes have one length (or a y axis
 * twice as long) run the fill pass of batch k+1 and the y pass of batch k in one launch whose
 * blocks interleave (arithmetic-bound and memory-bound tiles share each CU), where the passes' radix
 * plans have such a launch; 0 runs them as separate launches.  Same values. */
#define SS3D_OPT_FUSE 3
/* SS3D_OPT_BATCH_AUTOTUNE (default 1): ss3d_batch_create autotunes every distinct grid of its
 * jobs; 0 keeps the context's tuning table as it is (for comparisons with untuned plans). */
#define SS3D_OPT_BATCH_AUTOTUNE 4
/* SS3D_OPT_ZRUN (default 1): FP32 realizations whose z and y axes have length 2^9 keep the fill
 * and y outputs in runs of 8 z per pair of kx columns (cropped nz a multiple of 8), written by a
 * z fill whose waves exchange without block barriers and read by a C2R pass that takes whole runs;
 * 0 keeps the row layouts.  Same field values bit for bit (the statistics' block sums group other
 * rows, so they may differ in the last bits). */
#define SS3D_OPT_ZRUN 5
/* SS3D_OPT_SPEC_FP32 (default 0): ss3d_spectrum_create builds spectra for FP32 realizations only.
 * No FP64 amplitudes are kept (FP64 realizations from such a spectrum fail with SS3D_EINVAL), and
 * the transforms after the x pass run over the columns kx <= npx/2 alone: the spectral density of
 * the real covariance grid is Hermitian, S(-k) = conj S(k), and the FP32 amplitudes use no other
 * column.
 * The FP32 amplitudes equal those of a full build bit for bit; the diagnostics agree to rounding
 * (their sums count the columns 0 < kx < npx/2 twice).  The request loop and ss3d_batch_create
 * build the spectra of their FP32-only work this way whatever the option. */
#define SS3D_OPT_SPEC_FP32 6
int ss3d_context_set_option(ss3d_context* ctx, int option, int value);

/* ---------------------------------------------------------------- request loop
 * No counterpart in the reference (one process per parameter file).  A server owns `nslots`
 * lanes of a context and a worker thread.  Requests are admitted in arrival order as soon as a
 * lane is free and their device footprint fits `hbm_budget` bytes (<= 0: 80% of free memory);
 * requests admitted while others run overlap on the device.  Spectral densities are cached by
 * (model, spacing, periodic grid), shared, and evicted least recently used when unreferenced.
 * While a server runs, its context must not be used from other threads. */
typedef struct ss3d_server ss3d_server;

typedef struct ss3d_request {
    int64_t id;
    int32_t nst, iseed;
    ss3d_structure st[10];
    double nugget, dx, dy, dz;
    int32_t nx, ny, nz, ipad;
    int32_t igauss, nreal, precision;
    int32_t spot_check; /* FP32 requests: also compute realization 0 in FP64 and report the RMS gap */
    void* host_out;     /* optional: nreal*nx*ny*nz values copied here before completion */
} ss3d_request;

typedef struct ss3d_completion {
    int64_t id;
    int32_t status;    /* SS3D_OK or the error code */
    int32_t cache_hit; /* spectral density found in the cache */
    double t_submit, t_admit, t_done; /* seconds on the server clock */
    double spot_rel_rms;              /* relative RMS FP32 vs FP64 of realization 0, or -1 */
    double mean0, var0;               /* statistics of realization 0 */
    int64_t out_bytes;
} ss3d_completion;

int ss3d_server_create(ss3d_context* ctx, int nslots, int64_t hbm_budget, ss3d_server** out);
int ss3d_server_submit(ss3d_server* s, const ss3d_request* r);
/* Returns up to `max` finished requests; waits up to timeout_s seconds when none is ready. */
int ss3d_server_poll(ss3d_server* s, ss3d_completion* out, int max, int* n, double timeout_s);
int ss3d_server_info(ss3d_server* s, int64_t* queued, int64_t* running, int64_t* cached, int64_t* cache_bytes);
double ss3d_server_clock(const ss3d_server* s);
/* Waits for every submitted request, then releases the lanes and the cache. */
int ss3d_server_destroy(ss3d_server* s);

/* ---------------------------------------------------------------- output files
 * The reference's writers (src/specsim3d.f90:307-391 write_vti / write_vtk_bin, and the output
 * block f90:1357-1414), as a writer that streams files while realizations are computed.
 * Files (NULL or vtk_count = 0: not written), each byte-identical to the reference's:
 *   bin_path        raw float32 little-endian, realizations one after the other, x fastest
 *                   (f90:1362-1376);
 *   vti_path        VTK XML ImageData with every realization as appended raw float32 blocks,
 *                   each preceded by its UInt64 byte count (f90:307-359, 1377-1380);
 *   vtk_base        binary legacy VTK (big-endian float32) of realizations 1..vtk_count, named
 *                   <vtk_base>_<r>.vtk (f90:362-391, 1370-1373);
 *   gslib_path      GSLIB grid file: title, nreal, sim1..simN, then one line per node with every
 *                   realization in es15.7 (f90:1381-1396);
 *   vtk_ascii_path  ASCII legacy VTK of realization 1, es13.5 (f90:1397-1413).
 * Values written as float32 are real(x, 4) of the FP64 field (round to nearest even). */
typedef struct ss3d_outputs {
    const char* bin_path;
    const char* vti_path;
    const char* vtk_base;
    int32_t vtk_count;
    int32_t reserved; /* 0 */
    const char* gslib_path;
    const char* vtk_ascii_path;
    double origin[3];  /* xmin, ymin, zmin */
    double spacing[3]; /* dx, dy, dz */
} ss3d_outputs;

typedef struct ss3d_writer ss3d_writer;
/* Creates (replaces) the files and writes every header; all file offsets follow from the grid and
 * nreal.  ctx may be NULL for a writer fed only by ss3d_writer_put.  precision: of the fields the
 * writer receives (SS3D_FP64 as the reference's `res`, or SS3D_FP32). */
int ss3d_writer_open(ss3d_context* ctx, const ss3d_outputs* o, int nx, int ny, int nz, int nreal, int precision,
                     ss3d_writer** out);
/* Largest number of realizations per ss3d_writer_realize call (about 256 MiB of output). */
int ss3d_writer_slot(const ss3d_writer* w);
/* Computes the next nreal (1..slot) realizations of the stream (iseed, igauss), as ss3d_realize
 * does, and returns their statistics (mean, var, maximag: nreal doubles each, may be NULL).  Their
 * output goes to the files in the background: a copy stream moves them to pinned host memory and a
 * writer thread writes them while the following calls compute.  *dev_field (may be NULL) receives
 * the device address of these realizations (nx*ny*nz*nreal values of the writer's precision),
 * valid for work ordered on the context's stream before the next call.  Must be called in order
 * until nreal realizations have been given. */
int ss3d_writer_realize(ss3d_writer* w, const ss3d_spectrum* spec, int iseed, int igauss, int nreal, double* mean,
                        double* var, double* maximag, const void** dev_field);
/* Writes the next nreal realizations from host memory (the writer's precision); returns when the
 * buffer may be reused. */
int ss3d_writer_put(ss3d_writer* w, const void* host_field, int nreal);
/* Waits for every pending write, writes the GSLIB and ASCII VTK files (several threads, fixed-width
 * lines at known offsets) and the .vti trailer, closes the files and frees the writer, also on
 * failure.  Returns the first error of any write. */
int ss3d_writer_close(ss3d_writer* w);
/* Fortran ESw.d of n values into out (n*w characters, no terminator), as the writers print them:
 * correctly rounded digits (ties to even), E+dd exponents, +ddd beyond 99, '*' fill on overflow. */
int ss3d_format_es(const double* v, int64_t n, int w, int d, char* out);

/* ---------------------------------------------------------------- conditioning
 * Conditioning of realizations on data by simple kriging of their residual (module cond_mod,
 * src/specsim3d.f90:548-1083): z_c = z*_d + z_s - z*_s at every node, from the ndmax nearest data
 * in the anisotropic search ellipsoid, then the normal-score back transform (itrans = 1) or the
 * simple-kriging mean added (f90:910-925).  The caller prepares the data as the reference does
 * (f90:760-846: snapped to nodes, merged, normal scores or value - skmean) and passes them as 1-based
 * node indices (i, j, k) and Gaussian values.  ss3d_cond_create computes the kriging weights of
 * every node once (neighbour search, Cholesky solve with the reference's ridge rule);
 * ss3d_cond_apply conditions nreal realizations in device memory on the context's stream. */
typedef struct ss3d_cond_params {
    double rmax, rmed, rvert; /* search radii (f90:741) */
    double sa1, sa2, sa3;     /* search angles (f90:742) */
    int32_t ndmax;            /* nearest data per node, 1..100 (f90:743) */
    int32_t itrans;           /* 1: back transform with the table below; 0: add skmean */
    double skmean;            /* simple-kriging mean (itrans = 0) */
    int32_t nt;               /* normal-score table: vr (values) and gt (scores), ascending */
    int32_t ltail, utail;     /* tail options of the back transform (f90:736-737) */
    int32_t reserved;         /* 0 */
    const double* vr;
    const double* gt;
    double zlo, zhi, ltpar, utpar;
} ss3d_cond_params;

/* What the reference prints about the kriging (f90:890, 927-931), and the time spent. */
typedef struct ss3d_cond_info {
    int32_t nsbx, nsby, nsbz; /* super blocks */
    int32_t nnone;            /* nodes without data in the search ellipsoid */
    int32_t nexact;           /* nodes holding a datum */
    int32_t nfail;            /* systems that needed a larger ridge, or were not solved */
    int64_t nsum;             /* neighbours summed over nodes */
    double weights_s;         /* ss3d_cond_create, host clock */
    double apply_s;           /* kriging of realizations: device time of every ss3d_cond_apply */
    double backtr_s;          /* back transform / mean: device time */
} ss3d_cond_info;

/* The data-file part of the conditioning block (f90:732-734). */
typedef struct ss3d_cond_columns {
    int32_t ixc, iyc, izc, ivc, iwc; /* 1-based columns of x, y, z, value, weight; izc, iwc may be 0 */
    int32_t itrans;                  /* 1: normal-score transform of the data */
    double tmin, tmax;               /* trimming limits: values outside [tmin, tmax) are ignored */
    double skmean;                   /* simple-kriging mean (itrans = 0) */
} ss3d_cond_columns;

/* The data preparation of the reference (f90:760-846): rows (nrow records of ncol values, as read
 * from the GSLIB data file) trimmed, snapped to the nearest node, merged into the weighted mean of
 * each node (in the reference's heap-sort order), and transformed: normal scores (itrans = 1, with
 * the table vr/gt, nd entries, ascending) or value - skmean.  Outputs sized nrow: ijk (3 per datum,
 * 1-based), dval (merged values), dg (Gaussian values); nd, the number merged away and the number
 * outside the grid.  SS3D_EINVAL when no datum is inside the grid. */
int ss3d_cond_prepare(const double* rows, int nrow, int ncol, const ss3d_cond_columns* cols, int nx, int ny, int nz,
                      double xmin, double ymin, double zmin, double dx, double dy, double dz, int32_t* nd,
                      int32_t* nmerged, int32_t* nout, int32_t* ijk, double* dval, double* dg, double* vr, double* gt);

typedef struct ss3d_cond ss3d_cond;
int ss3d_cond_create(ss3d_context* ctx, const ss3d_structure* st, int nst, double nugget, int nx, int ny, int nz,
                     double xmin, double ymin, double zmin, double dx, double dy, double dz, int nd,
                     const int32_t* ijk, const double* g, const ss3d_cond_params* p, ss3d_cond** out,
                     ss3d_cond_info* info);
/* dev_field: nreal realizations (nx*ny*nz values each, x fastest) of the given precision. */
int ss3d_cond_apply(ss3d_cond* cond, void* dev_field, int precision, int nreal);
/* Counters and accumulated times (waits for the context's stream). */
int ss3d_cond_info_get(ss3d_cond* cond, ss3d_cond_info* info);
int ss3d_cond_destroy(ss3d_cond* cond);
/* Conditions every later ss3d_writer_realize chunk with cond (same context) before it is written,
 * as the reference conditions res before its writers (f90:1347). */
int ss3d_writer_condition(ss3d_writer* w, ss3d_cond* cond);

/* ---------------------------------------------------------------- variograms
 * Experimental semivariograms, replacing vario_lag (src/specsim3d.f90:405-431) as vario_report
 * calls it (f90:462-483).  For realization r, direction d with step (sx, sy, sz) and lag
 * l = 1..nlag_d:
 *   gam[(r*ndir + d)*nlagmax + l-1] = 0.5 * sum over all node pairs (f(x + l s) - f(x))^2 / npairs,
 * 0 where no pair exists (and for l > nlag_d); nlagmax = max(1, max_d nlag_d).  field: nreal fields
 * of nx*ny*nz values, x fastest, FP64 (SS3D_FP64) or float32 (SS3D_FP32), in host memory
 * (on_device = 0, copied in) or device memory.  dirs: ndir x {sx, sy, sz, nlag}.  Sums are FP64 in
 * a fixed order, so repeated calls give identical results. */
int ss3d_variogram(ss3d_context* ctx, const void* field, int on_device, int precision, int nx, int ny, int nz,
                   int nreal, int ndir, const int32_t* dirs, double* gam);
/* Number of node pairs at lag (lx, ly, lz) on an nx x ny x nz grid (npair_count, f90:394-402). */
int64_t ss3d_npairs(int nx, int ny, int nz, int lx, int ly, int lz);
/* Model covariance C(h) of nested structures plus nugget at lag (hx, hy, hz) (covariance, f90:153-178),
 * FP64 on the host; the model variogram is C(0) - C(h). */
double ss3d_covariance(const ss3d_structure* st, int nst, double nugget, double hx, double hy, double hz);

/* ---------------------------------------------------------------- batch graph
 * No counterpart in the reference.  Many FP32 jobs of uneven grids (ss3d_request, host_out
 * ignored), simulated together: creation builds their spectral densities (shared between
 * identical models and grids), autotunes every distinct grid, runs every job once eagerly to
 * size device buffers, and captures all jobs' passes into one HIP graph, jobs spread over
 * `nlanes` (1..16) concurrent branches.  ss3d_batch_run replays it on the context's stream
 * (use_graph = 1) or enqueues every pass again (use_graph = 0), with the same results; it
 * returns before the work ends (ss3d_context_synchronize waits).  The results of job i stay on
 * the device until ss3d_batch_output copies them (nreal*nx*ny*nz floats, x fastest; per
 * realization mean, variance and 0).  The context must not be used by other threads meanwhile. */
typedef struct ss3d_batch ss3d_batch;
int ss3d_batch_create(ss3d_context* ctx, const ss3d_request* jobs, int njobs, int nlanes, ss3d_batch** out);
int ss3d_batch_run(ss3d_batch* b, int use_graph);
int ss3d_batch_output(ss3d_batch* b, int job, void* host_out, double* host_stats);
int64_t ss3d_batch_cells(const ss3d_batch* b);
int ss3d_batch_destroy(ss3d_batch* b);

#ifdef __cplusplus
}
#endif
#endif
