/*
 * oracle.c -- CPU restatement of SPECSIM3D's unconditional spectral simulation path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle.so.
 *
 * Citations "f90:N" are lines of the reference program src/specsim3d.f90.  Each routine
 * restates the reference's arithmetic in the same operation order so that, compiled with
 * -O2 -ffp-contract=off against glibc libm (as the reference build is), results agree bit for
 * bit with the reference executable.  Pinned by tests/test_oracle.py against:
 *   - golden fields written by the reference build (tests/golden/, FP64 dumps),
 *   - the Park-Miller known answer (state 1 -> 10,000th state 1043618065),
 *   - brute-force DFTs, Parseval and closed forms of the covariance models,
 *   - a brute-force walk of the fill loop for the draw-index closed form.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979323846 /* f90:91 */
#define PM_A 16807LL                 /* f90:193 */
#define PM_M 2147483647LL

/* f90:107-139 setrot: GSLIB scaled rotation (rows 2,3 scaled by ahmax/ahmin, ahmax/avert) and
 * the principal axes scaled by their ranges.  rot and axes are row-major [row*3+col]. */
void or_setrot(const or_structure* s, double rot[9], double axes[9]) {
    double deg = OR_PI / 180.0;
    double alpha;
    if (s->az >= 0.0 && s->az < 270.0)
        alpha = (90.0 - s->az) * deg;
    else
        alpha = (450.0 - s->az) * deg;
    double beta = -s->dip * deg;
    double theta = s->plg * deg;
    double sina = sin(alpha), cosa = cos(alpha);
    double sinb = sin(beta), cosb = cos(beta);
    double sint = sin(theta), cost = cos(theta);
    double anis1 = s->ahmin / s->ahmax;
    double anis2 = s->avert / s->ahmax;
    double afac1 = 1.0 / (anis1 > 1.0e-20 ? anis1 : 1.0e-20);
    double afac2 = 1.0 / (anis2 > 1.0e-20 ? anis2 : 1.0e-20);
    rot[0] = cosb * cosa;
    rot[1] = cosb * sina;
    rot[2] = -sinb;
    rot[3] = afac1 * (-(cost * sina) + sint * sinb * cosa);
    rot[4] = afac1 * (cost * cosa + sint * sinb * sina);
    rot[5] = afac1 * (sint * cosb);
    rot[6] = afac2 * (sint * sina + cost * sinb * cosa);
    rot[7] = afac2 * (-(sint * cosa) + cost * sinb * sina);
    rot[8] = afac2 * (cost * cosb);
    for (int c = 0; c < 3; ++c) {
        axes[0 + c] = rot[0 + c] * s->ahmax;
        axes[3 + c] = rot[3 + c] / afac1 * s->ahmin;
        axes[6 + c] = rot[6 + c] / afac2 * s->avert;
    }
}

/* f90:142-150 */
double or_cutoff(int itype) {
    if (itype == 1) return 1.0;
    if (itype == 2) return 2.3;
    return 1.6;
}

/* f90:181-188 */
int or_nextpow2(int n) {
    int m = 1;
    while (m < n) m *= 2;
    return m;
}

/* f90:1234-1253 periodic grid size.  Returns 0, or 1 when the grid exceeds the reference's
 * 2^27-node limit (the reference stops there). */
int or_periodic_grid(const or_structure* st, int nst, double dx, double dy, double dz, int nx, int ny,
                     int nz, int ipad, int* npx, int* npy, int* npz) {
    int px = or_nextpow2(nx), py = or_nextpow2(ny), pz = or_nextpow2(nz);
    if (ipad == 1) {
        double ex = 0, ey = 0, ez = 0;
        for (int s = 0; s < nst; ++s) {
            double rot[9], axes[9];
            or_setrot(&st[s], rot, axes);
            double co = or_cutoff(st[s].itype);
            /* sqrt(sum(axes(:,d)**2)): column d of axes */
            double vx = co * sqrt(axes[0] * axes[0] + axes[3] * axes[3] + axes[6] * axes[6]) / dx;
            double vy = co * sqrt(axes[1] * axes[1] + axes[4] * axes[4] + axes[7] * axes[7]) / dy;
            double vz = co * sqrt(axes[2] * axes[2] + axes[5] * axes[5] + axes[8] * axes[8]) / dz;
            if (vx > ex) ex = vx;
            if (vy > ey) ey = vy;
            if (vz > ez) ez = vz;
        }
        int rx = (int)ceil(ex), ry = (int)ceil(ey), rz = (int)ceil(ez);
        px = or_nextpow2(nx + rx > 2 * rx ? nx + rx : 2 * rx);
        py = or_nextpow2(ny + ry > 2 * ry ? ny + ry : 2 * ry);
        if (nz > 1)
            pz = or_nextpow2(nz + rz > 2 * rz ? nz + rz : 2 * rz);
        else
            pz = 1;
    }
    *npx = px;
    *npy = py;
    *npz = pz;
    return ((double)px * (double)py * (double)pz > 134217728.0) ? 1 : 0;
}

/* f90:153-178 covariance at lag (hx,hy,hz).  rot holds nst row-major 3x3 matrices. */
double or_covariance(const or_structure* st, const double* rot, int nst, double nugget, double hx,
                     double hy, double hz) {
    double c = 0.0;
    if (hx == 0.0 && hy == 0.0 && hz == 0.0) c = nugget;
    for (int i = 0; i < nst; ++i) {
        const double* r = rot + 9 * i;
        double rx = r[0] * hx + r[1] * hy + r[2] * hz;
        double ry = r[3] * hx + r[4] * hy + r[5] * hz;
        double rz = r[6] * hx + r[7] * hy + r[8] * hz;
        double h = sqrt(rx * rx + ry * ry + rz * rz);
        double hr = h / st[i].ahmax;
        if (st[i].itype == 1) {
            if (hr < 1.0) c = c + st[i].sill * (1.0 - 1.5 * hr + 0.5 * (hr * hr * hr));
        } else if (st[i].itype == 2) {
            c = c + st[i].sill * exp(-3.0 * hr);
        } else {
            c = c + st[i].sill * exp(-3.0 * hr * hr);
        }
    }
    return c;
}

/* f90:244-251 twiddles: tw(k) = cos(2 pi k/n) + i*isign*sin(2 pi k/n), k < max(n/2,1). */
void or_twiddles(int n, int isign, double* tw) {
    int m = n / 2 > 1 ? n / 2 : 1;
    for (int k = 0; k < m; ++k) {
        double ang = 2.0 * OR_PI * (double)k / (double)n;
        tw[2 * k] = cos(ang);
        tw[2 * k + 1] = (double)isign * sin(ang);
    }
}

/* f90:208-241 fft1d: in-place radix-2 DIT, bit-reversal then len = 2, 4, ..., n. */
void or_fft1d(double* x, int n, const double* tw) {
    if (n < 2) return;
    /* bit reversal, 1-based as in f90:215-226 */
    int j = 1;
    for (int i = 1; i <= n - 1; ++i) {
        if (i < j) {
            double tr = x[2 * (i - 1)], ti = x[2 * (i - 1) + 1];
            x[2 * (i - 1)] = x[2 * (j - 1)];
            x[2 * (i - 1) + 1] = x[2 * (j - 1) + 1];
            x[2 * (j - 1)] = tr;
            x[2 * (j - 1) + 1] = ti;
        }
        int m = n / 2;
        while (m >= 1 && j > m) {
            j = j - m;
            m = m / 2;
        }
        j = j + m;
    }
    for (int len = 2; len <= n; len *= 2) {
        int half = len / 2, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                double ur = x[2 * (i + k)], ui = x[2 * (i + k) + 1];
                double wr = tw[2 * (k * step)], wi = tw[2 * (k * step) + 1];
                double vr = x[2 * (i + k + half)], vi = x[2 * (i + k + half) + 1];
                double trr = wr * vr - wi * vi;
                double tri = wr * vi + wi * vr;
                x[2 * (i + k)] = ur + trr;
                x[2 * (i + k) + 1] = ui + tri;
                x[2 * (i + k + half)] = ur - trr;
                x[2 * (i + k + half) + 1] = ui - tri;
            }
        }
    }
}

/* f90:254-304 fft3d: x pencils, then y, then z; unscaled; a is (nx,ny,nz) complex, x fastest. */
void or_fft3d(double* a, int nx, int ny, int nz, int isign) {
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    double* buf = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
    double* tw = (double*)malloc(sizeof(double) * 2 * (size_t)(nmax / 2 + 1));
    if (nx > 1) {
        or_twiddles(nx, isign, tw);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                size_t base = (size_t)nx * ((size_t)j + (size_t)ny * k);
                memcpy(buf, a + 2 * base, sizeof(double) * 2 * nx);
                or_fft1d(buf, nx, tw);
                memcpy(a + 2 * base, buf, sizeof(double) * 2 * nx);
            }
    }
    if (ny > 1) {
        or_twiddles(ny, isign, tw);
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                for (int j = 0; j < ny; ++j) {
                    size_t p = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    buf[2 * j] = a[2 * p];
                    buf[2 * j + 1] = a[2 * p + 1];
                }
                or_fft1d(buf, ny, tw);
                for (int j = 0; j < ny; ++j) {
                    size_t p = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    a[2 * p] = buf[2 * j];
                    a[2 * p + 1] = buf[2 * j + 1];
                }
            }
    }
    if (nz > 1) {
        or_twiddles(nz, isign, tw);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                for (int k = 0; k < nz; ++k) {
                    size_t p = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    buf[2 * k] = a[2 * p];
                    buf[2 * k + 1] = a[2 * p + 1];
                }
                or_fft1d(buf, nz, tw);
                for (int k = 0; k < nz; ++k) {
                    size_t p = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                    a[2 * p] = buf[2 * k];
                    a[2 * p + 1] = buf[2 * k + 1];
                }
            }
    }
    free(buf);
    free(tw);
}

/* f90:1256-1288: covariance on the periodic grid with minimum-image lags (lag n/2 stays +n/2),
 * forward transform, S = c/N, amp = sqrt(Re S) for Re S >= 0 else 0, and the diagnostics. */
void or_spectrum(const or_structure* st, int nst, double nugget, double dx, double dy, double dz, int npx,
                 int npy, int npz, double* amp, double* spec_re, double* spec_im, or_spec_diag* diag) {
    size_t ntot = (size_t)npx * npy * npz;
    double* rot = (double*)malloc(sizeof(double) * 9 * (size_t)(nst > 0 ? nst : 1));
    for (int s = 0; s < nst; ++s) {
        double axes[9];
        or_setrot(&st[s], rot + 9 * s, axes);
    }
    double* c = (double*)malloc(sizeof(double) * 2 * ntot);
    for (int k = 0; k < npz; ++k)
        for (int j = 0; j < npy; ++j)
            for (int i = 0; i < npx; ++i) {
                int kp = k; if (kp > npz / 2) kp -= npz;
                int jp = j; if (jp > npy / 2) jp -= npy;
                int ip = i; if (ip > npx / 2) ip -= npx;
                size_t p = (size_t)i + (size_t)npx * ((size_t)j + (size_t)npy * k);
                c[2 * p] = or_covariance(st, rot, nst, nugget, (double)ip * dx, (double)jp * dy, (double)kp * dz);
                c[2 * p + 1] = 0.0;
            }
    double c0 = c[0];
    or_fft3d(c, npx, npy, npz, -1);
    double rn = (double)ntot;
    double maximag = 0, maxreal = 0, sneg = 0, spos = 0;
    for (size_t p = 0; p < ntot; ++p) {
        double a = c[2 * p] / rn;
        double b = c[2 * p + 1] / rn;
        if (spec_re) spec_re[p] = a;
        if (spec_im) spec_im[p] = b;
        if (fabs(b) > maximag) maximag = fabs(b);
        if (fabs(a) > maxreal) maxreal = fabs(a);
        if (a < 0.0) {
            sneg = sneg + a;
            amp[p] = 0.0;
        } else {
            spos = spos + a;
            amp[p] = sqrt(a);
        }
    }
    if (diag) {
        diag->max_imag = maximag;
        diag->max_real = maxreal;
        diag->c0 = c0;
        diag->spos = spos;
        diag->sneg = sneg;
    }
    free(c);
    free(rot);
}

/* f90:1302-1303 seeding: state = mod(|iseed|, 2147483646) + 1, then 10 discarded draws. */
int64_t or_seed_state(int iseed) {
    int64_t a = iseed < 0 ? -(int64_t)iseed : (int64_t)iseed;
    int64_t s = a % 2147483646LL + 1;
    for (int i = 0; i < 10; ++i) s = (PM_A * s) % PM_M;
    return s;
}

/* f90:191-195 urand */
double or_urand(int64_t* state) {
    *state = (PM_A * *state) % PM_M;
    return (double)*state / 2147483647.0;
}

/* f90:198-205 gauss_pair (Box-Muller) */
void or_gauss_pair(int64_t* state, double* g1, double* g2) {
    double u1 = or_urand(state);
    double u2 = or_urand(state);
    double rr = sqrt(-2.0 * log(u1));
    *g1 = rr * cos(2.0 * OR_PI * u2);
    *g2 = rr * sin(2.0 * OR_PI * u2);
}

/* a^e mod (2^31-1), square and multiply (jump-ahead of the f90:193 recurrence). */
int64_t or_pm_pow(int64_t a, uint64_t e) {
    int64_t r = 1, b = a % PM_M;
    while (e) {
        if (e & 1) r = (r * b) % PM_M;
        b = (b * b) % PM_M;
        e >>= 1;
    }
    return r;
}

/* Draws consumed by one realization of the fill loop f90:1305-1334: every point with
 * lin <= linp draws once (igauss=0) or twice (igauss=1); (N + n_sc)/2 such points, n_sc = number
 * of self-conjugate points = 2^(axes with n >= 2). */
int64_t or_draws_per_realization(int npx, int npy, int npz, int igauss) {
    int64_t n = (int64_t)npx * npy * npz;
    int64_t nsc = (npx >= 2 ? 2 : 1) * (npy >= 2 ? 2 : 1) * (npz >= 2 ? 2 : 1);
    return (igauss ? 2 : 1) * ((n + nsc) / 2);
}

/* Closed form of the fill order (derived from f90:1305-1334): index, within one realization,
 * of the first draw that produces the coefficient at (i,j,k) (the partner's draw when
 * lin > linp).  Planes k = 0 and k = npz/2 and rows j = 0 and j = npy/2 inside them are
 * self-conjugate and draw only for i = 0..npx/2; other planes/rows draw for every i if their
 * partner comes later in the walk and never otherwise. */
int64_t or_point_draw_index(int i, int j, int k, int npx, int npy, int npz, int igauss) {
    int ip = (npx - i) % npx, jp = (npy - j) % npy, kp = (npz - k) % npz;
    int64_t lin = i + (int64_t)npx * (j + (int64_t)npy * k);
    int64_t linp = ip + (int64_t)npx * (jp + (int64_t)npy * kp);
    if (lin > linp) {
        i = ip; j = jp; k = kp;
    }
    int64_t hxc = npx >= 2 ? npx / 2 + 1 : 1;
    int64_t hp = npy >= 2 ? (int64_t)(npy / 2 - 1) * npx + 2 * hxc : hxc;
    int64_t fp = (int64_t)npy * npx;
    int64_t q;
    int self_plane = (k == 0) || (npz >= 2 && k == npz / 2);
    int64_t plane_start = (k == 0) ? 0 : hp + (int64_t)(k - 1) * fp;
    if (!self_plane) {
        q = plane_start + (int64_t)j * npx + i;
    } else {
        int64_t row_start = (j == 0) ? 0 : hxc + (int64_t)(j - 1) * npx;
        q = plane_start + row_start + i;
    }
    return (igauss ? 2 : 1) * q;
}

/* f90:1305-1334, literal serial walk (k, j, i with x fastest). */
void or_fill(const double* amp, double* c, int npx, int npy, int npz, int igauss, int64_t* state,
             int64_t* draw_log) {
    const double s05 = sqrt(0.5);
    int64_t first_state = 0;
    int64_t ndraw = 0; /* draws taken so far in this realization */
    if (draw_log) {
        size_t n = (size_t)npx * npy * npz;
        for (size_t p = 0; p < n; ++p) draw_log[p] = -1;
    }
    (void)first_state;
    for (int k = 0; k < npz; ++k)
        for (int j = 0; j < npy; ++j)
            for (int i = 0; i < npx; ++i) {
                int ip = (npx - i) % npx, jp = (npy - j) % npy, kp = (npz - k) % npz;
                int64_t lin = i + (int64_t)npx * (j + (int64_t)npy * k);
                int64_t linp = ip + (int64_t)npx * (jp + (int64_t)npy * kp);
                size_t p = (size_t)lin, pp = (size_t)linp;
                if (lin < linp) {
                    if (draw_log) draw_log[p] = draw_log[pp] = ndraw;
                    if (igauss == 1) {
                        double g1, g2;
                        or_gauss_pair(state, &g1, &g2);
                        ndraw += 2;
                        c[2 * p] = amp[p] * g1 * s05;
                        c[2 * p + 1] = amp[p] * g2 * s05;
                        c[2 * pp] = amp[pp] * g1 * s05;
                        c[2 * pp + 1] = amp[pp] * (-g2) * s05;
                    } else {
                        double phi = 2.0 * OR_PI * or_urand(state);
                        ndraw += 1;
                        double cp = cos(phi), sp = sin(phi);
                        c[2 * p] = amp[p] * cp;
                        c[2 * p + 1] = amp[p] * sp;
                        c[2 * pp] = amp[pp] * cp;
                        c[2 * pp + 1] = amp[pp] * (-sp);
                    }
                } else if (lin == linp) {
                    if (draw_log) draw_log[p] = ndraw;
                    if (igauss == 1) {
                        double g1, g2;
                        or_gauss_pair(state, &g1, &g2);
                        ndraw += 2;
                        c[2 * p] = amp[p] * g1;
                        c[2 * p + 1] = 0.0;
                    } else {
                        ndraw += 1;
                        if (or_urand(state) < 0.5) {
                            c[2 * p] = amp[p];
                        } else {
                            c[2 * p] = -amp[p];
                        }
                        c[2 * p + 1] = 0.0;
                    }
                }
            }
}

/* f90:1234-1345: sizing, spectrum, seeding, then per realization fill, inverse transform,
 * crop of the first nx*ny*nz corner and the mean/variance/max|Im| statistics. */
int or_simulate(const or_structure* st, int nst, double nugget, double dx, double dy, double dz, int nx,
                int ny, int nz, int ipad, int iseed, int nreal, int igauss, double* res, double* mean,
                double* var, double* maximag, or_spec_diag* diag, int* np_out) {
    int npx, npy, npz;
    if (or_periodic_grid(st, nst, dx, dy, dz, nx, ny, nz, ipad, &npx, &npy, &npz)) return 1;
    if (np_out) {
        np_out[0] = npx;
        np_out[1] = npy;
        np_out[2] = npz;
    }
    size_t ntot = (size_t)npx * npy * npz;
    double* amp = (double*)malloc(sizeof(double) * ntot);
    double* c = (double*)malloc(sizeof(double) * 2 * ntot);
    or_spectrum(st, nst, nugget, dx, dy, dz, npx, npy, npz, amp, NULL, NULL, diag);
    int64_t state = or_seed_state(iseed);
    size_t nout = (size_t)nx * ny * nz;
    /* c keeps its contents across realizations in the reference (it is never cleared); the
     * walk rewrites every point each realization, so this does not matter, but keep it. */
    for (size_t p = 0; p < 2 * ntot; ++p) c[p] = 0.0;
    for (int r = 0; r < nreal; ++r) {
        or_fill(amp, c, npx, npy, npz, igauss, &state, NULL);
        or_fft3d(c, npx, npy, npz, +1);
        double mi = 0.0;
        double* out = res + (size_t)r * nout;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    size_t p = (size_t)i + (size_t)npx * ((size_t)j + (size_t)npy * k);
                    double im = fabs(c[2 * p + 1]);
                    if (im > mi) mi = im;
                    out[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = c[2 * p];
                }
        double s = 0.0;
        for (size_t p = 0; p < nout; ++p) s += out[p];
        double zm = s / (double)nout;
        double v = 0.0;
        for (size_t p = 0; p < nout; ++p) v += (out[p] - zm) * (out[p] - zm);
        if (mean) mean[r] = zm;
        if (var) var[r] = v / (double)nout;
        if (maximag) maximag[r] = mi;
    }
    free(c);
    free(amp);
    return 0;
}

/* npair_count, f90:394-402: pairs at lag (lx, ly, lz) on an nx x ny x nz grid. */
int64_t or_npair_count(int nx, int ny, int nz, int lx, int ly, int lz) {
    if (abs(lx) >= nx || abs(ly) >= ny || abs(lz) >= nz) return 0;
    return (int64_t)(nx - abs(lx)) * (int64_t)(ny - abs(ly)) * (int64_t)(nz - abs(lz));
}

/* vario_lag, f90:405-431: the loop runs over the nodes whose partner x + l is inside the grid,
 * k outermost, accumulating (f(x + l) - f(x))^2 in that order. */
double or_vario_lag(const double* f, int nx, int ny, int nz, int lx, int ly, int lz) {
    const int64_t np = or_npair_count(nx, ny, nz, lx, ly, lz);
    if (np == 0) return 0.0;
    const int i0 = lx < 0 ? -lx : 0, i1 = lx > 0 ? nx - lx : nx;
    const int j0 = ly < 0 ? -ly : 0, j1 = ly > 0 ? ny - ly : ny;
    const int k0 = lz < 0 ? -lz : 0, k1 = lz > 0 ? nz - lz : nz;
    double acc = 0.0;
    for (int k = k0; k < k1; ++k)
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i) {
                const double d = f[(i + lx) + (int64_t)nx * ((j + ly) + (int64_t)ny * (k + lz))] -
                                 f[i + (int64_t)nx * (j + (int64_t)ny * k)];
                acc = acc + d * d;
            }
    return 0.5 * acc / (double)np;
}
