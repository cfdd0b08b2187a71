/*
 * oracle.h -- CPU restatement of SPECSIM3D's unconditional spectral simulation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.  The product path
 * (specsim3d_amd, libspecsim3d.so, the Fortran driver) never calls it.
 *
 * Every function follows a block of the reference program src/specsim3d.f90 (cited per
 * function in oracle.c).  Arithmetic is IEEE binary64 compiled with -ffp-contract=off and the
 * same glibc libm routines the reference calls (cos, sin, exp, log, sqrt), so on x86-64 the
 * fields agree with the reference build bit for bit (checked against the golden fixtures).
 */
#ifndef SS3D_ORACLE_H
#define SS3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One nested structure: the scalar members of `type structure` (f90:95-100). */
typedef struct or_structure {
    int32_t itype;   /* 1 spherical, 2 exponential, 3 gaussian */
    int32_t pad_;
    double sill, ahmax, ahmin, avert, az, dip, plg;
} or_structure;

/* Spectral-density diagnostics printed by the reference (f90:1289-1292). */
typedef struct or_spec_diag {
    double max_imag;   /* max |Im S| */
    double max_real;   /* max |Re S| */
    double c0;         /* C(0) = Re c(1,1,1) before the transform */
    double spos;       /* sum of Re S >= 0 */
    double sneg;       /* sum of Re S < 0 */
} or_spec_diag;

void    or_setrot(const or_structure* s, double rot[9], double axes[9]);
double  or_cutoff(int itype);
int     or_nextpow2(int n);
int     or_periodic_grid(const or_structure* st, int nst, double dx, double dy, double dz,
                         int nx, int ny, int nz, int ipad, int* npx, int* npy, int* npz);
double  or_covariance(const or_structure* st, const double* rot, int nst, double nugget,
                      double hx, double hy, double hz);
void    or_twiddles(int n, int isign, double* tw /* 2*max(n/2,1) */);
void    or_fft1d(double* x /* 2n interleaved */, int n, const double* tw);
void    or_fft3d(double* a /* 2*nx*ny*nz */, int nx, int ny, int nz, int isign);

/* Covariance grid + forward transform + normalise/clip/sqrt.  amp has npx*npy*npz entries. */
void    or_spectrum(const or_structure* st, int nst, double nugget, double dx, double dy, double dz,
                    int npx, int npy, int npz, double* amp, double* spec_re, double* spec_im,
                    or_spec_diag* diag);

/* Park-Miller stream (f90:191-205, 1302-1303). */
int64_t or_seed_state(int iseed);
double  or_urand(int64_t* state);
void    or_gauss_pair(int64_t* state, double* g1, double* g2);
/* Closed forms (derived from the fill loop f90:1305-1334), pinned against or_fill's serial walk. */
int64_t or_pm_pow(int64_t a, uint64_t e);
int64_t or_draws_per_realization(int npx, int npy, int npz, int igauss);
int64_t or_point_draw_index(int i, int j, int k, int npx, int npy, int npz, int igauss);

/* One realization's coefficient fill, literal serial walk (f90:1305-1334).  c is 2*N doubles.
 * draw_log, when non-NULL, receives for every point the 0-based index (within the realization)
 * of the first draw that produced its coefficient, or -1 for points the walk never writes. */
void    or_fill(const double* amp, double* c, int npx, int npy, int npz, int igauss,
                int64_t* state, int64_t* draw_log);

/* Whole unconditional run: spectrum, then nreal realizations (f90:1234-1345).
 * res: nx*ny*nz*nreal doubles (x fastest); mean/var/maximag: nreal each (may be NULL). */
int     or_simulate(const or_structure* st, int nst, double nugget, double dx, double dy, double dz,
                    int nx, int ny, int nz, int ipad, int iseed, int nreal, int igauss,
                    double* res, double* mean, double* var, double* maximag, or_spec_diag* diag,
                    int* np_out /* 3 ints, may be NULL */);

/* Experimental semivariogram of one field (x fastest) for the integer lag vector (lx, ly, lz):
 * 0.5 * mean over all node pairs of (f(x + l) - f(x))^2, 0 when no pair exists
 * (npair_count f90:394-402, vario_lag f90:405-431). */
int64_t or_npair_count(int nx, int ny, int nz, int lx, int ly, int lz);
double  or_vario_lag(const double* f, int nx, int ny, int nz, int lx, int ly, int lz);

/* ---- conditioning (cond_mod, f90:548-1083), oracle_cond.c ---- */
/* The conditioning block without its data file name (f90:731-747). */
typedef struct or_cond_params {
    int32_t ixc, iyc, izc, ivc, iwc;   /* GSLIB columns, 1-based; izc, iwc may be 0 */
    int32_t itrans, ltail, utail, ndmax, pad_;
    double tmin, tmax, skmean, zlo, zhi, ltpar, utpar;
    double rmax, rmed, rvert, sa1, sa2, sa3;
} or_cond_params;

/* What the reference prints about the conditioning (f90:818-819, 890, 927-931). */
typedef struct or_cond_info {
    int32_t nd, nmerged, nout, nsbx, nsby, nsbz, nnone, nexact, nfail, nt;
    int64_t nsum;
} or_cond_info;

double  or_gcum(double x);
double  or_gauinv(double p);
void    or_heapsort_idx(const double* key, int* idx /* 1-based entries */, int n);
void    or_nscore(int nd, const double* v, const double* w, double* g, double* vr, double* gt);
double  or_backtr(double y, int nt, const double* vr, const double* gt, double zmin, double zmax, int lt,
                  double ltp, int ut, double utp);
/* condition_field without its file I/O: returns 0, or 1 (ndmax), 2 (radii), 3 (no data in grid). */
int     or_condition(double* res, int nx, int ny, int nz, int nreal, double xmin, double ymin, double zmin,
                     double dx, double dy, double dz, const or_structure* st, int nst, double nugget,
                     const or_cond_params* p, const double* rows, int nrow, int ncol, or_cond_info* info,
                     int* ix, int* iy, int* iz, double* dval, double* dg, double* vrt, double* gtt);

#ifdef __cplusplus
}
#endif
#endif
