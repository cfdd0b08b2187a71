/*
 * oracle_cond.c -- CPU restatement of SPECSIM3D's conditioning (module cond_mod,
 * src/specsim3d.f90:548-1083): simple kriging of the residual of each unconditional
 * realization, with GSLIB-style super-block neighbour search, Cholesky solves, normal-score
 * transform and back transform.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): only tests/ load liboracle.so.
 *
 * Every routine restates the reference's arithmetic in its operation order (compiled with
 * -ffp-contract=off against glibc libm, as the reference build is), so that conditioned fields
 * agree with the reference executable's.  Pinned by tests/test_oracle_cond.py against the
 * reference's own FP64 conditioned fields and output files (tests/golden/cond), exact simple
 * kriging solved by numpy, and the standard normal quantiles.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQRT2 1.4142135623730951     /* f90:551 */
#define SQRT2PI 2.5066282746310002   /* f90:552 */

/* f90:557-562: standard normal cdf. */
double or_gcum(double x) { return 0.5 * erfc(-x / SQRT2); }

/* f90:564-590: inverse standard normal cdf, Acklam's rational approximation (coefficients as
 * published by P. J. Acklam) followed by one Halley step on erfc. */
double or_gauinv(double p0) {
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                1.383577518672690e+02,  -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                6.680131188771972e+01,  -1.328068155288572e+01};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                -2.549732539343734e+00, 4.374664141464968e+00,  2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                3.754408661907416e+00};
    double p = p0 < 1.0e-15 ? 1.0e-15 : p0;
    if (p > 1.0 - 1.0e-15) p = 1.0 - 1.0e-15;
    double x, q, r;
    if (p < 0.02425) {
        q = sqrt(-2.0 * log(p));
        x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
            ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    } else if (p <= 1.0 - 0.02425) {
        q = p - 0.5;
        r = q * q;
        x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
            (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
    } else {
        q = sqrt(-2.0 * log(1.0 - p));
        x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
            ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    }
    double e = 0.5 * erfc(-x / SQRT2) - p;
    double u = e * SQRT2PI * exp(0.5 * x * x);
    return x - u / (1.0 + 0.5 * x * u);
}

/* f90:593-626: heapsort of an index array by key, ascending.  Positions are 1-based as in the
 * reference (children of position p are 2p and 2p+1), stored in idx[pos-1]. */
static void sift(const double* key, int* idx, int root, int last) {
    int parent = root, v = idx[parent - 1];
    for (;;) {
        int child = 2 * parent;
        if (child > last) break;
        if (child < last && key[idx[child] - 1] > key[idx[child - 1] - 1]) child = child + 1;
        if (key[idx[child - 1] - 1] > key[v - 1]) {
            idx[parent - 1] = idx[child - 1];
            parent = child;
        } else {
            break;
        }
    }
    idx[parent - 1] = v;
}

/* idx holds 1-based entries into key. */
void or_heapsort_idx(const double* key, int* idx, int n) {
    for (int i = n / 2; i >= 1; --i) sift(key, idx, i, n);
    for (int i = n; i >= 2; --i) {
        int t = idx[0];
        idx[0] = idx[i - 1];
        idx[i - 1] = t;
        sift(key, idx, 1, i - 1);
    }
}

/* f90:628-636. */
static double powint(double xlow, double xhigh, double ylow, double yhigh, double xval, double pw) {
    if (xhigh - xlow < 1.0e-20) return 0.5 * (ylow + yhigh);
    return ylow + (yhigh - ylow) * pow((xval - xlow) / (xhigh - xlow), pw);
}

/* f90:638-659: weighted normal-score transform.  g: score of each datum; vr, gt: the table. */
void or_nscore(int nd, const double* v, const double* w, double* g, double* vr, double* gt) {
    int* idx = malloc(sizeof(int) * (size_t)(nd > 0 ? nd : 1));
    for (int i = 0; i < nd; ++i) idx[i] = i + 1;
    or_heapsort_idx(v, idx, nd);
    double sw = 0.0;
    for (int i = 0; i < nd; ++i) sw = sw + w[i];
    double cp = 0.0, oldcp = 0.0;
    for (int i = 0; i < nd; ++i) {
        int k = idx[i] - 1;
        cp = cp + w[k] / sw;
        gt[i] = or_gauinv(0.5 * (cp + oldcp));
        vr[i] = v[k];
        g[k] = gt[i];
        oldcp = cp;
    }
    free(idx);
}

/* f90:662-698: GSLIB back transform with lower/upper tail options. */
double or_backtr(double y, int nt, const double* vr, const double* gt, double zmin, double zmax, int lt, double ltp,
                 int ut, double utp) {
    double z;
    if (y <= gt[0]) {
        double cdflo = or_gcum(gt[0]), cdfbt = or_gcum(y);
        if (lt == 2)
            z = powint(0.0, cdflo, zmin, vr[0], cdfbt, ltp);
        else
            z = powint(0.0, cdflo, zmin, vr[0], cdfbt, 1.0);
        if (nt == 1) z = vr[0];
    } else if (y >= gt[nt - 1]) {
        double cdfhi = or_gcum(gt[nt - 1]), cdfbt = or_gcum(y);
        if (ut == 2) {
            z = powint(cdfhi, 1.0, vr[nt - 1], zmax, cdfbt, utp);
        } else if (ut == 4) {
            double lam = pow(vr[nt - 1], utp) * (1.0 - cdfhi);
            double den = 1.0 - cdfbt;
            if (den < 1.0e-300) den = 1.0e-300;
            z = pow(lam / den, 1.0 / utp);
        } else {
            z = powint(cdfhi, 1.0, vr[nt - 1], zmax, cdfbt, 1.0);
        }
    } else {
        int lo = 1, hi = nt;
        while (hi - lo > 1) {
            int mid = (lo + hi) / 2;
            if (gt[mid - 1] <= y)
                lo = mid;
            else
                hi = mid;
        }
        z = powint(gt[lo - 1], gt[hi - 1], vr[lo - 1], vr[hi - 1], y, 1.0);
    }
    return z;
}

/* Fortran nint: nearest integer, halves away from zero. */
static int nint(double x) { return (int)lround(x); }

/* Fortran ceiling of a double. */
static int ceil_i(double x) { return (int)ceil(x); }

/* f90:700-1004 condition_field, minus its file reading and writing: rows are the data file's
 * records (nrow x ncol, row-major, GSLIB columns p->ixc.. 1-based).  res (nx*ny*nz*nreal, x
 * fastest) is conditioned in place.  Outputs (arrays sized nrow, may be NULL): the data nodes used
 * (1-based), their merged values and Gaussian values, and the normal-score table (nt entries). */
int or_condition(double* res, int nx, int ny, int nz, int nreal, double xmin, double ymin, double zmin, double dx,
                 double dy, double dz, const or_structure* st, int nst, double nugget, const or_cond_params* p,
                 const double* rows, int nrow, int ncol, or_cond_info* info, int* oix, int* oiy, int* oiz,
                 double* odval, double* odg, double* ovrt, double* ogtt) {
    memset(info, 0, sizeof *info);
    if (p->ndmax < 1 || p->ndmax > 100) return 1;
    if (p->rmax <= 0 || p->rmed <= 0 || p->rvert <= 0) return 2;
    const int ndmax = p->ndmax;
    /* search ellipsoid: an isotropic-sill spherical structure of ranges rmax/rmed/rvert (f90:750-754) */
    or_structure srch = {1, 0, 1.0, p->rmax, p->rmed, p->rvert, p->sa1, p->sa2, p->sa3};
    double srot[9], saxes[9];
    or_setrot(&srch, srot, saxes);
    double ex = sqrt(saxes[0] * saxes[0] + saxes[3] * saxes[3] + saxes[6] * saxes[6]);
    double ey = sqrt(saxes[1] * saxes[1] + saxes[4] * saxes[4] + saxes[7] * saxes[7]);
    double ez = sqrt(saxes[2] * saxes[2] + saxes[5] * saxes[5] + saxes[8] * saxes[8]);

    /* data inside the trimming limits and the grid, snapped to the nearest node (f90:780-797) */
    double* raw = malloc(sizeof(double) * 5 * (size_t)(nrow > 0 ? nrow : 1));
    double* key = malloc(sizeof(double) * (size_t)(nrow > 0 ? nrow : 1));
    int n = 0, nout = 0;
    for (int i = 0; i < nrow; ++i) {
        const double* row = rows + (size_t)i * ncol;
        double v = row[p->ivc - 1];
        if (v < p->tmin || v >= p->tmax) continue;
        int j = nint((row[p->ixc - 1] - xmin) / dx) + 1;
        int k = nint((row[p->iyc - 1] - ymin) / dy) + 1;
        int r = 1;
        if (p->izc > 0 && nz > 1) r = nint((row[p->izc - 1] - zmin) / dz) + 1;
        if (j < 1 || j > nx || k < 1 || k > ny || r < 1 || r > nz) {
            nout = nout + 1;
            continue;
        }
        double* q = raw + 5 * (size_t)n;
        q[0] = (double)j;
        q[1] = (double)k;
        q[2] = (double)r;
        q[3] = v;
        q[4] = 1.0;
        if (p->iwc > 0) q[4] = row[p->iwc - 1];
        key[n] = (double)j + (double)nx * ((double)(k - 1) + (double)ny * (double)(r - 1));
        n = n + 1;
    }
    info->nout = nout;
    if (n == 0) {
        free(raw);
        free(key);
        return 3;
    }
    /* merge data of the same node: weighted mean (f90:799-817) */
    int* ord = malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) ord[i] = i + 1;
    or_heapsort_idx(key, ord, n);
    int *ix = malloc(sizeof(int) * (size_t)n), *iy = malloc(sizeof(int) * (size_t)n), *iz = malloc(sizeof(int) * (size_t)n);
    double *dval = malloc(sizeof(double) * (size_t)n), *dwt = malloc(sizeof(double) * (size_t)n);
    int nd = 0;
    int i = 1;
    while (i <= n) {
        int j = i;
        while (j < n) {
            if (key[ord[j] - 1] != key[ord[i - 1] - 1]) break;
            j = j + 1;
        }
        const double* q0 = raw + 5 * (size_t)(ord[i - 1] - 1);
        ix[nd] = nint(q0[0]);
        iy[nd] = nint(q0[1]);
        iz[nd] = nint(q0[2]);
        double sw = 0.0, svw = 0.0;
        for (int m = i; m <= j; ++m) sw = sw + raw[5 * (size_t)(ord[m - 1] - 1) + 4];
        for (int m = i; m <= j; ++m) {
            const double* q = raw + 5 * (size_t)(ord[m - 1] - 1);
            svw = svw + q[3] * q[4];
        }
        dwt[nd] = sw;
        dval[nd] = svw / dwt[nd];
        nd = nd + 1;
        i = j + 1;
    }
    info->nd = nd;
    info->nmerged = n - nd;
    double *dxs = malloc(sizeof(double) * (size_t)nd), *dys = malloc(sizeof(double) * (size_t)nd),
           *dzs = malloc(sizeof(double) * (size_t)nd), *dg = malloc(sizeof(double) * (size_t)nd);
    for (int d = 0; d < nd; ++d) {
        dxs[d] = xmin + (double)(ix[d] - 1) * dx;
        dys[d] = ymin + (double)(iy[d] - 1) * dy;
        dzs[d] = zmin + (double)(iz[d] - 1) * dz;
    }
    /* normal scores or the simple-kriging mean (f90:821-846) */
    double *vrt = malloc(sizeof(double) * (size_t)nd), *gtt = malloc(sizeof(double) * (size_t)nd);
    int nt = nd;
    if (p->itrans == 1) {
        or_nscore(nd, dval, dwt, dg, vrt, gtt);
    } else {
        for (int d = 0; d < nd; ++d) dg[d] = dval[d] - p->skmean;
    }
    info->nt = p->itrans == 1 ? nt : 0;
    /* simulated values at the data nodes (f90:857-863) */
    const size_t ncell = (size_t)nx * ny * nz;
    double* simd = malloc(sizeof(double) * (size_t)nd * (size_t)(nreal > 0 ? nreal : 1));
    for (int r = 0; r < nreal; ++r)
        for (int d = 0; d < nd; ++d)
            simd[(size_t)r * nd + d] = res[(size_t)r * ncell + (size_t)(ix[d] - 1) + (size_t)nx * ((iy[d] - 1) + (size_t)ny * (iz[d] - 1))];
    /* super blocks (f90:865-891) */
    int nsbx = ceil_i((double)nx * dx / (0.5 * ex)), nsby = ceil_i((double)ny * dy / (0.5 * ey)),
        nsbz = ceil_i((double)nz * dz / (0.5 * ez));
    nsbx = nsbx > 100 ? 100 : nsbx;
    nsbx = nsbx < 1 ? 1 : nsbx;
    nsby = nsby > 100 ? 100 : nsby;
    nsby = nsby < 1 ? 1 : nsby;
    nsbz = nsbz > 100 ? 100 : nsbz;
    nsbz = nsbz < 1 ? 1 : nsbz;
    if (nz == 1) nsbz = 1;
    while ((double)nsbx * nsby * nsbz > 2.0e6) {
        nsbx = nsbx * 3 / 4 > 1 ? nsbx * 3 / 4 : 1;
        nsby = nsby * 3 / 4 > 1 ? nsby * 3 / 4 : 1;
        nsbz = nsbz * 3 / 4 > 1 ? nsbz * 3 / 4 : 1;
    }
    info->nsbx = nsbx;
    info->nsby = nsby;
    info->nsbz = nsbz;
    const int nsb = nsbx * nsby * nsbz;
    const double bsx = (double)nx * dx / nsbx, bsy = (double)ny * dy / nsby, bsz = (double)nz * dz / nsbz;
    const int mbx = ceil_i(ex / bsx), mby = ceil_i(ey / bsy), mbz = ceil_i(ez / bsz);
#define BLOCKOF(X, Y, Z, BX, BY, BZ)                                      \
    do {                                                                  \
        BX = (int)(((X) - xmin + 0.5 * dx) / bsx) + 1;                    \
        BX = BX < 1 ? 1 : (BX > nsbx ? nsbx : BX);                        \
        BY = (int)(((Y) - ymin + 0.5 * dy) / bsy) + 1;                    \
        BY = BY < 1 ? 1 : (BY > nsby ? nsby : BY);                        \
        BZ = (int)(((Z) - zmin + 0.5 * dz) / bsz) + 1;                    \
        BZ = BZ < 1 ? 1 : (BZ > nsbz ? nsbz : BZ);                        \
    } while (0)
    int *sbstart = calloc((size_t)nsb + 1, sizeof(int)), *sbcount = calloc((size_t)nsb, sizeof(int));
    int *sbidx = malloc(sizeof(int) * (size_t)nd), *sbof = malloc(sizeof(int) * (size_t)nd);
    for (int d = 0; d < nd; ++d) {
        int bx, by, bz;
        BLOCKOF(dxs[d], dys[d], dzs[d], bx, by, bz);
        sbof[d] = bx + nsbx * (by - 1 + nsby * (bz - 1));
        sbcount[sbof[d] - 1] += 1;
    }
    sbstart[0] = 1;
    for (int b = 0; b < nsb; ++b) sbstart[b + 1] = sbstart[b] + sbcount[b];
    memset(sbcount, 0, sizeof(int) * (size_t)nsb);
    for (int d = 0; d < nd; ++d) {
        int b = sbof[d];
        sbidx[sbstart[b - 1] + sbcount[b - 1] - 1] = d + 1;
        sbcount[b - 1] += 1;
    }
    double* rot = malloc(sizeof(double) * 9 * (size_t)nst);
    for (int s = 0; s < nst; ++s) {
        double ax[9];
        or_setrot(&st[s], rot + 9 * s, ax);
    }
    const double c0 = or_covariance(st, rot, nst, nugget, 0.0, 0.0, 0.0);

    /* kriging of the residual at every node (f90:895-907, krige_node 950-1004) */
    double *cd = malloc(sizeof(double) * ndmax), *a = malloc(sizeof(double) * ndmax * ndmax),
           *bvec = malloc(sizeof(double) * ndmax), *wgt = malloc(sizeof(double) * ndmax),
           *lmat = malloc(sizeof(double) * ndmax * ndmax);
    int* cid = malloc(sizeof(int) * ndmax);
    int64_t nsum = 0;
    int nfail = 0, nnone = 0, nexact = 0;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int ii = 1; ii <= nx; ++ii) {
                const double x = xmin + (ii - 1) * dx, y = ymin + (j - 1) * dy, z = zmin + (k - 1) * dz;
                int bx, by, bz;
                BLOCKOF(x, y, z, bx, by, bz);
                int nn = 0;
                for (int kb = (bz - mbz > 1 ? bz - mbz : 1); kb <= (bz + mbz < nsbz ? bz + mbz : nsbz); ++kb)
                    for (int jb = (by - mby > 1 ? by - mby : 1); jb <= (by + mby < nsby ? by + mby : nsby); ++jb)
                        for (int ib = (bx - mbx > 1 ? bx - mbx : 1); ib <= (bx + mbx < nsbx ? bx + mbx : nsbx); ++ib) {
                            const int blk = ib + nsbx * (jb - 1 + nsby * (kb - 1));
                            for (int pp = sbstart[blk - 1]; pp <= sbstart[blk] - 1; ++pp) {
                                const int q = sbidx[pp - 1];
                                const double hx = dxs[q - 1] - x, hy = dys[q - 1] - y, hz = dzs[q - 1] - z;
                                const double rx = srot[0] * hx + srot[1] * hy + srot[2] * hz;
                                const double ry = srot[3] * hx + srot[4] * hy + srot[5] * hz;
                                const double rz = srot[6] * hx + srot[7] * hy + srot[8] * hz;
                                const double d2 = (rx * rx + ry * ry + rz * rz) / (p->rmax * p->rmax);
                                if (d2 > 1.0) continue;
                                int m;
                                if (nn < ndmax) {
                                    nn = nn + 1;
                                    m = nn;
                                } else if (d2 < cd[nn - 1]) {
                                    m = nn;
                                } else {
                                    continue;
                                }
                                while (m > 1) {
                                    if (cd[m - 2] <= d2) break;
                                    cd[m - 1] = cd[m - 2];
                                    cid[m - 1] = cid[m - 2];
                                    m = m - 1;
                                }
                                cd[m - 1] = d2;
                                cid[m - 1] = q;
                            }
                        }
                nsum = nsum + nn;
                if (nn == 0) {
                    nnone = nnone + 1;
                    continue;
                }
                const size_t node = (size_t)(ii - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1));
                if (cd[0] == 0.0) {
                    const int q = cid[0] - 1;
                    if (ix[q] == ii && iy[q] == j && iz[q] == k) {
                        for (int l = 0; l < nreal; ++l) res[(size_t)l * ncell + node] = dg[q];
                        nexact = nexact + 1;
                        continue;
                    }
                }
                for (int pp = 0; pp < nn; ++pp) {
                    const int cp = cid[pp] - 1;
                    for (int q = 0; q <= pp; ++q) {
                        const int cq = cid[q] - 1;
                        a[pp * ndmax + q] = or_covariance(st, rot, nst, nugget, dxs[cp] - dxs[cq], dys[cp] - dys[cq],
                                                          dzs[cp] - dzs[cq]);
                        a[q * ndmax + pp] = a[pp * ndmax + q];
                    }
                    bvec[pp] = or_covariance(st, rot, nst, nugget, dxs[cp] - x, dys[cp] - y, dzs[cp] - z);
                }
                double ridge = 1.0e-10 * c0;
                int ok = 0, jj;
                for (jj = 1; jj <= 3; ++jj) {
                    /* chol_solve (f90:1006-1040): L L^T = A + ridge I, then two triangular solves */
                    ok = 1;
                    for (int pp = 0; pp < nn && ok; ++pp)
                        for (int q = 0; q <= pp; ++q) {
                            double s = a[pp * ndmax + q];
                            if (pp == q) s = s + ridge;
                            for (int m = 0; m < q; ++m) s = s - lmat[pp * ndmax + m] * lmat[q * ndmax + m];
                            if (pp == q) {
                                if (s <= 0.0) {
                                    ok = 0;
                                    break;
                                }
                                lmat[pp * ndmax + pp] = sqrt(s);
                            } else {
                                lmat[pp * ndmax + q] = s / lmat[q * ndmax + q];
                            }
                        }
                    if (ok) {
                        for (int pp = 0; pp < nn; ++pp) {
                            double s = bvec[pp];
                            for (int m = 0; m < pp; ++m) s = s - lmat[pp * ndmax + m] * wgt[m];
                            wgt[pp] = s / lmat[pp * ndmax + pp];
                        }
                        for (int pp = nn - 1; pp >= 0; --pp) {
                            double s = wgt[pp];
                            for (int m = pp + 1; m < nn; ++m) s = s - lmat[m * ndmax + pp] * wgt[m];
                            wgt[pp] = s / lmat[pp * ndmax + pp];
                        }
                        break;
                    }
                    ridge = ridge * 1.0e3;
                }
                if (!ok) {
                    nfail = nfail + 1;
                    continue;
                }
                if (jj > 1) nfail = nfail + 1;
                double zk = 0.0;
                for (int pp = 0; pp < nn; ++pp) zk = zk + wgt[pp] * dg[cid[pp] - 1];
                for (int l = 0; l < nreal; ++l) {
                    double sks = 0.0;
                    for (int pp = 0; pp < nn; ++pp) sks = sks + wgt[pp] * simd[(size_t)l * nd + cid[pp] - 1];
                    double* rv = &res[(size_t)l * ncell + node];
                    *rv = zk + *rv - sks;
                }
            }
#undef BLOCKOF
    info->nsum = nsum;
    info->nnone = nnone;
    info->nexact = nexact;
    info->nfail = nfail;
    /* back transform, or the mean restored (f90:910-925) */
    if (p->itrans == 1) {
        for (size_t c = 0; c < ncell * (size_t)nreal; ++c)
            res[c] = or_backtr(res[c], nt, vrt, gtt, p->zlo, p->zhi, p->ltail, p->ltpar, p->utail, p->utpar);
    } else if (p->skmean != 0.0) {
        for (size_t c = 0; c < ncell * (size_t)nreal; ++c) res[c] = res[c] + p->skmean;
    }
    for (int d = 0; d < nd; ++d) {
        if (oix) oix[d] = ix[d];
        if (oiy) oiy[d] = iy[d];
        if (oiz) oiz[d] = iz[d];
        if (odval) odval[d] = dval[d];
        if (odg) odg[d] = dg[d];
        if (ovrt && p->itrans == 1) ovrt[d] = vrt[d];
        if (ogtt && p->itrans == 1) ogtt[d] = gtt[d];
    }
    free(raw); free(key); free(ord); free(ix); free(iy); free(iz); free(dval); free(dwt);
    free(dxs); free(dys); free(dzs); free(dg); free(vrt); free(gtt); free(simd);
    free(sbstart); free(sbcount); free(sbidx); free(sbof); free(rot);
    free(cd); free(a); free(bvec); free(wgt); free(lmat); free(cid);
    return 0;
}
