/*
 * CPU oracle for CMAQ's SMVGEAR gas-chemistry solver.
 *
 * TEST INFRASTRUCTURE ONLY (see ros3_oracle.h).  A plain-C restatement of
 *   CCTM/src/gas/smvgear/grsubfun.F   (dc/dt)                    -> om_subfun
 *   CCTM/src/gas/smvgear/grpderiv.F   (predictor matrix I - a h J) -> om_pderiv
 *   CCTM/src/gas/smvgear/grdecomp.F   (sparse LU)                 -> om_decomp
 *   CCTM/src/gas/smvgear/grbacksub.F  (substitutions)             -> om_solve
 *   CCTM/src/gas/smvgear/GRVARS.F:550-707 (OPTIMAL_ATOL_PPM)      -> om_optimal_atol
 *   CCTM/src/gas/smvgear/grsmvgear.F  (Gear BDF + step/order control) -> om_smvgear
 *   CCTM/src/gas/smvgear/grdriver.F:400-560 (block driver)        -> om_gear_run
 * The symbolic analysis (grsprse.F, JSPARSE) is the same algorithm as
 * ros3's RBSPARSE (species order, fill, storage order, decomposition and
 * back-substitution lists, PD scatter lists), so om_create serves both; its
 * coefficient tables (grsprse.F:675-713, GRVARS.F:251-254) are below.
 * grdecomp.F / grbacksub.F perform the operations of rbdecomp.F /
 * rbsolve.F in the same order, so om_decomp / om_solve are reused.
 *
 * REAL(4) literals and runtime variables of the Fortran (0.2, 1.1, 1.4,
 * 0.1, GEAR_RTOL, GEAR_*_ATOL, ENQQ1..3) keep their single-precision value
 * here, promoted to double where the Fortran promotes them.
 */
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

#define A2(a, i, cc) (a)[(size_t)(i) * nc + (cc)]
#define MXORDER 7
#define MAXORD 5                 /* grsprse.F:677 */
#define MSTEP 3                  /* grsprse.F:676 */
#define MBETWEEN 50              /* grsprse.F:678 */
#define MXLOWRED 20              /* grsmvgear.F:91  */
#define HMIN 1.0e-10             /* grinit.F:186    */
#define HMAXDAY 15.0             /* grinit.F:184    */
#define HMAXNIT 5.0              /* grinit.F:185    */
#define FRACDEC 0.25             /* grinit.F:189    */
#define GCONCMIN 1.0e-30         /* grinit.F:190    */
#define ZBOUND 1.0e-30           /* grinit.F:191    */
#define OMIC 0.000001
#define OPT2MIC 0.0000012
#define OPT3MIC 0.0000013
#define OPT4MIC 0.0000014
#define PT2 0.2
#define PT003 0.003
#define OM_GEAR_MAX_STEPS 10000000L

/* Test knobs of the pow() of the step/order selection (grsmvgear.F:
 * 880-905).  By default it is libm's pow, as in the reference's build.
 * om_gear_set_pow_exact(1) takes x^y rounded once from quad precision
 * (libquadmath's powq, 113-bit), the double nearest the exact value up to
 * about 2^-60 of an ulp: the result of any correctly rounded pow, which is
 * what the device's step control computes.  om_gear_set_pow_nudge(n) then
 * moves every result by n ulps.  tests/test_oracle_gear.py uses both to
 * measure how far a last-bit difference of pow -- as between two libms --
 * moves SMVGEAR's results. */
static int g_pow_nudge = 0;
static int g_pow_exact = 0;
void om_gear_set_pow_nudge(int ulps) { g_pow_nudge = ulps; }
void om_gear_set_pow_exact(int on) { g_pow_exact = on; }
static double gpow(double a, double b)
{
    double r = g_pow_exact ? (double)powq((__float128)a, (__float128)b) : pow(a, b);
    for (int k = 0; k < g_pow_nudge; ++k) r = nextafter(r, INFINITY);
    for (int k = 0; k > g_pow_nudge; --k) r = nextafter(r, -INFINITY);
    return r;
}

/* Gear coefficient tables, as JSPARSE leaves them (grsprse.F:675-713,
 * GRVARS.F:251-254): PERTST squared, ENQQ1..3 from REAL(4) divisions,
 * CONPST / CONP15, ASET(order, j). */
typedef struct {
    double pertst[MXORDER + 1][4];
    double enqq1[MXORDER + 1], enqq2[MXORDER + 1], enqq3[MXORDER + 1];
    double conpst[MXORDER + 1], conp15[MXORDER + 1];
    double aset[11][9];
} gear_coef;

static void gear_coefficients(gear_coef *g)
{
    static const double pertst0[3][MXORDER] = {
        {2.0, 4.5, 7.333, 10.42, 13.7, 17.15, 1.0},
        {3.0, 6.0, 9.167, 12.5, 15.98, 1.0, 1.0},
        {1.0, 1.0, 0.5, 0.1667, 0.04133, 0.008267, 1.0}};
    memset(g, 0, sizeof *g);
    for (int q = 1; q <= MXORDER; ++q) {
        for (int k = 1; k <= 3; ++k) g->pertst[q][k] = pertst0[k - 1][q - 1];
        g->enqq1[q] = (double)(0.5f / (float)q);
        g->enqq2[q] = (double)(0.5f / (float)(q + 1));
        g->enqq3[q] = (double)(0.5f / (float)(q + 2));
        g->conpst[q] = 1.0 / (g->pertst[q][1] * g->enqq3[q]);
        g->conp15[q] = 1.5 * g->conpst[q];
        for (int k = 1; k <= 3; ++k) g->pertst[q][k] = g->pertst[q][k] * g->pertst[q][k];
    }
    for (int i = 1; i <= 6; ++i) { g->aset[i][2] = 1.0; g->aset[i][8] = 0.0; }
    g->aset[1][1] = 1.0;
    g->aset[2][1] = 2.0 / 3.0;     g->aset[2][3] = 1.0 / 3.0;
    g->aset[3][1] = 6.0 / 11.0;    g->aset[3][3] = 6.0 / 11.0;    g->aset[3][4] = 1.0 / 11.0;
    g->aset[4][1] = 12.0 / 25.0;   g->aset[4][3] = 0.70;          g->aset[4][4] = 0.20;
    g->aset[4][5] = 0.02;
    g->aset[5][1] = 60.0 / 137.0;  g->aset[5][3] = 225.0 / 274.0; g->aset[5][4] = 85.0 / 274.0;
    g->aset[5][5] = 15.0 / 274.0;  g->aset[5][6] = 1.0 / 274.0;
    g->aset[6][1] = 180.0 / 441.0; g->aset[6][3] = 406.0 / 441.0; g->aset[6][4] = 735.0 / 1764.0;
    g->aset[6][5] = 175.0 / 1764.0; g->aset[6][6] = 21.0 / 1764.0; g->aset[6][7] = 1.0 / 1764.0;
}

void om_gear_coefficients(double *aset /* [6][7]: ASET(q, j), q = 1..6, j = 1..7 */,
                          double *pertst /* [7][3] */, double *enqq /* [7][3] */,
                          double *conpst /* [7][2]: CONPST, CONP15 */)
{
    gear_coef g;
    gear_coefficients(&g);
    for (int q = 1; q <= 6; ++q)
        for (int j = 1; j <= 7; ++j) aset[(q - 1) * 7 + j - 1] = g.aset[q][j];
    for (int q = 1; q <= MXORDER; ++q) {
        for (int k = 1; k <= 3; ++k) pertst[(q - 1) * 3 + k - 1] = g.pertst[q][k];
        enqq[(q - 1) * 3 + 0] = g.enqq1[q];
        enqq[(q - 1) * 3 + 1] = g.enqq2[q];
        enqq[(q - 1) * 3 + 2] = g.enqq3[q];
        conpst[(q - 1) * 2 + 0] = g.conpst[q];
        conpst[(q - 1) * 2 + 1] = g.conp15[q];
    }
}

/* SUBFUN (grsubfun.F:83-171) with no emissions in the chemistry (EMBLK = 0):
 * GLOSS = sum over active reactions of stoichiometry x rate; species at or
 * below ZBOUND do not lose mass. */
void om_subfun(const om_ctx *c, int ncsp, int nc, const double *rki, const double *cnew, double *gloss)
{
    const om_tables *t = &c->t;
    for (int isp = 1; isp <= c->ischan; ++isp)
        for (int cc = 0; cc < nc; ++cc) A2(gloss, isp - 1, cc) = 0.0;
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int nrk = c->nkuserat[ncsp][nrx], nre = t->nreact[nrk - 1];
        for (int cc = 0; cc < nc; ++cc) {
            double rx = A2(rki, nrk - 1, cc);
            if (nre >= 1) rx = rx * A2(cnew, IRM2(c, nrk, 1) - 1, cc);
            if (nre >= 2) rx = rx * A2(cnew, IRM2(c, nrk, 2) - 1, cc);
            if (nre >= 3) rx = rx * A2(cnew, IRM2(c, nrk, 3) - 1, cc);
            for (int nr = 1; nr <= nre; ++nr) {
                int s = IRM2(c, nrk, nr) - 1;
                A2(gloss, s, cc) = A2(gloss, s, cc) - rx;
            }
            for (int np = 1; np <= t->nprdct[nrk - 1]; ++np) {
                int s = IRM2(c, nrk, np + 3) - 1;
                A2(gloss, s, cc) = A2(gloss, s, cc) + SC(c, nrk, np) * rx;
            }
        }
    }
    for (int isp = 1; isp <= c->ischan; ++isp)
        for (int cc = 0; cc < nc; ++cc)
            if (A2(cnew, isp - 1, cc) <= ZBOUND && A2(gloss, isp - 1, cc) < 0.0) A2(gloss, isp - 1, cc) = 0.0;
}

/* PDERIV (grpderiv.F:88-224): CC2 = I + R1DELT * J(CEFF), CEFF = CNEW with
 * values at or below ZBOUND taken as zero; each partial derivative is scaled
 * by R1DELT (or SC * R1DELT) as it is added. */
void om_pderiv(const om_ctx *c, int ncsp, int nc, const double *rki, const double *cnew, double r1delt,
               double *cc2)
{
    const om_tables *t = &c->t;
    const int iarry = c->iarray[ncsp], nondiag = iarry - c->ischan;
    for (int iar = 0; iar <= nondiag; ++iar)
        for (int cc = 0; cc < nc; ++cc) A2(cc2, iar, cc) = 0.0;
    for (int iar = nondiag + 1; iar <= iarry; ++iar)
        for (int cc = 0; cc < nc; ++cc) A2(cc2, iar, cc) = 1.0;
    double ex[4];
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int nrk = c->nkuserat[ncsp][nrx], nre = t->nreact[nrk - 1];
        for (int cc = 0; cc < nc; ++cc) {
#define CEFF(s) (A2(cnew, (s) - 1, cc) <= ZBOUND ? 0.0 : A2(cnew, (s) - 1, cc))
            double rk = A2(rki, nrk - 1, cc);
            if (nre == 1) {
                ex[1] = rk;
            } else if (nre == 2) {
                int jr1 = IRM2(c, nrk, 1), jr2 = IRM2(c, nrk, 2);
                ex[1] = rk * CEFF(jr2);
                ex[2] = rk * CEFF(jr1);
            } else if (nre == 3) {
                int jr1 = IRM2(c, nrk, 1), jr2 = IRM2(c, nrk, 2), jr3 = IRM2(c, nrk, 3);
                double cr2 = rk * CEFF(jr2);
                ex[1] = cr2 * CEFF(jr3);
                ex[2] = rk * CEFF(jr1) * CEFF(jr3);
                ex[3] = cr2 * CEFF(jr1);
            }
#undef CEFF
            for (int nl = 1; nl <= c->nderivl[ncsp][nrk]; ++nl) {
                int iarp = c->jarrl[ncsp][nrk * (c->mxrr + 1) + nl];
                int ialp = c->jlial[ncsp][nrk * (c->mxrr + 1) + nl];
                A2(cc2, iarp, cc) = A2(cc2, iarp, cc) - r1delt * ex[ialp];
            }
            for (int np = 1; np <= c->nderivp[ncsp][nrk]; ++np) {
                int q = nrk * (c->mxrp + 1) + np;
                int iarp = c->jarrp[ncsp][q], ialp = c->jpial[ncsp][q], isc = c->icoeff[ncsp][q];
                if (isc == 0) {
                    A2(cc2, iarp, cc) = A2(cc2, iarp, cc) + r1delt * ex[ialp];
                } else {
                    double fracn = SC(c, nrk, isc) * r1delt;
                    A2(cc2, iarp, cc) = A2(cc2, iarp, cc) + fracn * ex[ialp];
                }
            }
        }
    }
}

/* OPTIMAL_ATOL_PPM (GRVARS.F:550-707): per cell, the absolute tolerance is
 * the highest of six levels (log-spaced from GEAR_MAX_ATOL down to
 * GEAR_MIN_ATOL) above which more than 40 % of the species that exceed
 * GEAR_CONC_FLOOR lie.  cnew [ns][nc] in solver order. */
void om_optimal_atol(const om_ctx *c, int nc, const double *cnew, const om_gear_tol *tol, double *ylowa)
{
    enum { NBINS = 5 };
    double lev[NBINS + 2];
    const double abhi = log10((double)tol->max_atol), ablo = log10((double)tol->min_atol);
    lev[1] = pow(10.0, abhi);
    lev[NBINS + 1] = pow(10.0, ablo);
    const double floor_ = (double)tol->conc_floor;
    for (int i = 2; i <= NBINS; ++i)
        lev[i] = pow(10.0, ablo + (abhi - ablo) * (double)(NBINS - i + 1) / (double)NBINS);
    for (int cc = 0; cc < nc; ++cc) {
        int bins[NBINS + 2] = {0}, cum[NBINS + 2] = {0};
        for (int j = 1; j <= c->ns; ++j) {
            double v = A2(cnew, j - 1, cc);
            if (v >= lev[1]) bins[1]++;
            else if (v >= lev[2]) bins[2]++;
            else if (v >= lev[3]) bins[3]++;
            else if (v >= lev[4]) bins[4]++;
            else if (v >= lev[5]) bins[5]++;
            else if (v >= floor_) bins[6]++;
        }
        cum[1] = bins[1];
        for (int b = 2; b <= NBINS + 1; ++b) cum[b] = bins[b] + cum[b - 1];
        int cutoff = (int)(0.4f * (float)cum[6]);
        double a = lev[6];
        for (int b = 1; b <= NBINS; ++b)
            if (cum[b] > cutoff) { a = lev[b]; break; }
        ylowa[cc] = a;
    }
}

/* SMVGEAR (grsmvgear.F:232-872) on one block of nc cells sharing the step
 * size and order (nc = 1: per-cell control).  y [ns][nc] in solver order,
 * in/out.  The counters are per block. */
int om_smvgear(const om_ctx *c, int ncsp, int nc, double chemstep, const double *rki, double *y,
               const om_gear_tol *tol, int sunlit, om_gear_stats *st)
{
    const int ns = c->ns, ischan = c->ischan, ischan1 = ischan - 1;
    gear_coef g;
    gear_coefficients(&g);
    const size_t sz = (size_t)ns * nc;
    double *cnew = y;
    double *gloss = calloc(sz, sizeof(double));
    double *conc = calloc(sz * MXORDER, sizeof(double));
    double *dtlos = calloc(sz, sizeof(double)), *cest = calloc(sz, sizeof(double));
    double *chold = calloc(sz, sizeof(double)), *blkconc = malloc(sz * sizeof(double));
    double *dely = calloc(nc, sizeof(double)), *ylowa = malloc(nc * sizeof(double));
    double *yloweps = malloc(nc * sizeof(double)), *yloweps1 = malloc(nc * sizeof(double));
    double *cc2 = malloc((size_t)(c->iarray[ncsp] + 1) * nc * sizeof(double));
    double *vdiag = malloc((size_t)ischan * nc * sizeof(double));
#define CONC(cc, i) conc[(size_t)((i) - 1) * nc + (cc)]     /* CONC(NCELL, I), I = 1..ISCHAN*MXORDER */
    memcpy(blkconc, cnew, sz * sizeof(double));
    memset(st, 0, sizeof *st);
    const double hmax = sunlit ? HMAXDAY : HMAXNIT;
    int rc = 0;

    om_optimal_atol(c, nc, cnew, tol, ylowa);
    const double eps = (double)tol->rtol, eps1 = 1.0 / eps;
    const double epslow = eps < PT003 ? eps : PT003;
    const double order = (double)ischan;
    int jfail = 0;
    const double h0fac = sunlit ? 0.0625 : 1.0;
    const double abst2 = 1.0 / (chemstep * chemstep);
    int nylowred = 0, nsteps = 0;
    long guard = 0;

    /* LOOP_BEGIN */
    for (;;) {
        int ifsuccess = 1, idoub = 2, nslp = MBETWEEN, jrestar = 0;
        double xelaps = 0.0, told = xelaps, rdelmax = 10000.0, hrmax = 0.3, hratio = 0.0, asn1 = 1.0;
        for (int cc = 0; cc < nc; ++cc) yloweps1[cc] = ylowa[cc] * eps1;
        double timremain = chemstep;
        int restart_begin = 0;
        /* LOOP_INITIALIZE */
        for (;;) {
            om_subfun(c, ncsp, nc, rki, cnew, gloss);
            st->nsubfun++;
            for (int cc = 0; cc < nc; ++cc) yloweps[cc] = ylowa[cc] / epslow;
            for (int cc = 0; cc < nc; ++cc) dely[cc] = 0.0;
            for (int j = 1; j <= ischan; ++j)
                for (int cc = 0; cc < nc; ++cc) {
                    double e = A2(gloss, j - 1, cc) / (A2(cnew, j - 1, cc) + yloweps[cc]);
                    dely[cc] = dely[cc] + e * e;
                }
            double rmstop = 0.0;
            for (int cc = 0; cc < nc; ++cc)
                if (dely[cc] > rmstop) rmstop = dely[cc];
            double delt = sqrt(epslow / (abst2 + rmstop / order)) * h0fac;
            if (delt > chemstep) delt = chemstep;
            if (delt > hmax) delt = hmax;
            if (delt < HMIN) delt = HMIN;
            int nqq = 1;
            for (int j = 1; j <= ischan; ++j) {
                int js1 = ischan + j;
                for (int cc = 0; cc < nc; ++cc) {
                    CONC(cc, j) = A2(cnew, j - 1, cc);
                    CONC(cc, js1) = delt * A2(gloss, j - 1, cc);
                }
            }
            int nqqold = 0, jeval = 1, kstep = 0, nqqisc = 0;
            double rdelt = 1.0, enqq = 0, eup = 0, edwn = 0, drate = 0, rmserr = 0, der2max = 0;
            double rdeltup = 0;
            int restart_init = 0;
            /* LOOP_ORDER */
            for (;;) {
                if (++guard > OM_GEAR_MAX_STEPS) { rc = -3; goto done; }
                int higher_order = 1;
                timremain = chemstep - xelaps;
                if (timremain <= OMIC) goto done;
                if (nqq != nqqold) {
                    nqqold = nqq;
                    kstep = nqq + 1;
                    hratio = hratio * g.aset[nqq][1] / asn1;
                    asn1 = g.aset[nqq][1];
                    enqq = g.pertst[nqq][1] * order;
                    eup = g.pertst[nqq][2] * order;
                    edwn = g.pertst[nqq][3] * order;
                    nqqisc = nqq * ischan;
                }
                {
                    double m = rdelt;
                    if (rdelmax < m) m = rdelmax;
                    if (hmax / delt < m) m = hmax / delt;
                    if (timremain / delt < m) m = timremain / delt;
                    rdelt = m;
                }
                delt = delt * rdelt;
                hratio = hratio * rdelt;
                if (delt < HMIN) {
                    /* lower YLOWA and restart the sync step (grsmvgear.F:425-466) */
                    for (int cc = 0; cc < nc; ++cc) ylowa[cc] = ylowa[cc] * (double)0.1f;
                    nylowred++;
                    st->nylowred = nylowred;
                    if (nylowred < MXLOWRED) {
                        memcpy(cnew, blkconc, sz * sizeof(double));
                        restart_begin = 1;
                        break;
                    }
                    rc = -1;
                    goto done;
                }
                if (rdelt != 1.0) {
                    double rdelta = 1.0;
                    int i1 = 1;
                    for (int j = 2; j <= kstep; ++j) {
                        rdelta = rdelta * rdelt;
                        i1 = i1 + ischan;
                        for (int i = i1; i <= i1 + ischan1; ++i)
                            for (int cc = 0; cc < nc; ++cc) CONC(cc, i) = CONC(cc, i) * rdelta;
                    }
                }
                if (rdelt < 1.0) st->numbkups++;
                if (ifsuccess == 1) {
                    rdelmax = 10.0;
                    if (nsteps % 3 == 2) om_optimal_atol(c, nc, cnew, tol, ylowa);
                    for (int cc = 0; cc < nc; ++cc) yloweps1[cc] = ylowa[cc] * eps1;
                    for (int j = 1; j <= ischan; ++j)
                        for (int cc = 0; cc < nc; ++cc)
                            A2(chold, j - 1, cc) = eps1 / (fabs(A2(cnew, j - 1, cc)) + yloweps1[cc]);
                }
                /* prediction */
                if (fabs(hratio - 1.0) > hrmax || nsteps >= nslp) jeval = 1;
                xelaps = xelaps + delt;
                {
                    int i1 = nqqisc + 1;
                    for (int jb = 1; jb <= nqq; ++jb) {
                        i1 = i1 - ischan;
                        for (int i = i1; i <= nqqisc; ++i) {
                            int j = i + ischan;
                            for (int cc = 0; cc < nc; ++cc) CONC(cc, i) = CONC(cc, i) + CONC(cc, j);
                        }
                    }
                }
                /* corrector (label 440) */
                int l3;
                int converged;
            newton:
                l3 = 0;
                st->numnewt++;
                for (int i = 1; i <= ischan; ++i)
                    for (int cc = 0; cc < nc; ++cc) {
                        A2(cnew, i - 1, cc) = CONC(cc, i);
                        A2(dtlos, i - 1, cc) = 0.0;
                    }
                if (jeval == 1) {
                    double r1delt = -asn1 * delt;
                    om_pderiv(c, ncsp, nc, rki, cnew, r1delt, cc2);
                    om_decomp(c, ncsp, nc, cc2, vdiag);
                    st->npderiv++;
                    jeval = -1;
                    hratio = 1.0;
                    nslp = nsteps + MBETWEEN;
                    drate = 0.70;
                }
                converged = 0;
                for (;;) {   /* label 520 */
                    om_subfun(c, ncsp, nc, rki, cnew, gloss);
                    st->nsubfun++;
                    for (int i = 1; i <= ischan; ++i) {
                        int j = i + ischan;
                        for (int cc = 0; cc < nc; ++cc)
                            A2(gloss, i - 1, cc) = delt * A2(gloss, i - 1, cc) -
                                                   (CONC(cc, j) + A2(dtlos, i - 1, cc));
                    }
                    om_solve(c, ncsp, nc, cc2, vdiag, gloss);
                    for (int cc = 0; cc < nc; ++cc) dely[cc] = 0.0;
                    for (int i = 1; i <= ischan; ++i)
                        for (int cc = 0; cc < nc; ++cc) {
                            A2(dtlos, i - 1, cc) = A2(dtlos, i - 1, cc) + A2(gloss, i - 1, cc);
                            A2(cnew, i - 1, cc) = CONC(cc, i) + asn1 * A2(dtlos, i - 1, cc);
                            double e = A2(gloss, i - 1, cc) * A2(chold, i - 1, cc);
                            dely[cc] = dely[cc] + e * e;
                        }
                    double rmserrp = rmserr;
                    der2max = 0.0;
                    for (int cc = 0; cc < nc; ++cc)
                        if (dely[cc] > der2max) der2max = dely[cc];
                    rmserr = sqrt(der2max / order);
                    l3 = l3 + 1;
                    if (l3 > 1) {
                        double rmsrat = rmserr / rmserrp;
                        double d = (double)0.2f * drate;
                        drate = d > rmsrat ? d : rmsrat;
                    }
                    double m = g.conp15[nqq] * drate;
                    if (g.conpst[nqq] < m) m = g.conpst[nqq];
                    double dcon = rmserr * m;
                    if (dcon <= 1.0) { converged = 1; break; }
                    if (l3 >= MSTEP) break;
                }
                if (!converged) {
                    /* convergence failure (grsmvgear.F:658-683) */
                    st->ncfail++;
                    if (jeval == 0) { jeval = 1; goto newton; }
                    rdelmax = 2.0;
                    xelaps = told;
                    rdelt = FRACDEC;
                    jeval = 1;
                    ifsuccess = 0;
                    int i1 = nqqisc + 1;
                    for (int jb = 1; jb <= nqq; ++jb) {
                        i1 = i1 - ischan;
                        for (int i = i1; i <= nqqisc; ++i) {
                            int j = i + ischan;
                            for (int cc = 0; cc < nc; ++cc) CONC(cc, i) = CONC(cc, i) - CONC(cc, j);
                        }
                    }
                    continue;   /* CYCLE LOOP_ORDER */
                }
                /* convergence achieved (label 780) */
                jeval = 0;
                if (l3 > 1) {
                    for (int cc = 0; cc < nc; ++cc) dely[cc] = 0.0;
                    for (int j = 1; j <= ischan; ++j)
                        for (int cc = 0; cc < nc; ++cc) {
                            double e = A2(dtlos, j - 1, cc) * A2(chold, j - 1, cc);
                            dely[cc] = dely[cc] + e * e;
                        }
                    der2max = 0.0;
                    for (int cc = 0; cc < nc; ++cc)
                        if (dely[cc] > der2max) der2max = dely[cc];
                }
                if (der2max > enqq) {
                    /* error test failure (grsmvgear.F:728-777) */
                    st->nefail++;
                    xelaps = told;
                    jfail = jfail + 1;
                    int i1 = nqqisc + 1;
                    for (int jb = 1; jb <= nqq; ++jb) {
                        i1 = i1 - ischan;
                        for (int i = i1; i <= nqqisc; ++i) {
                            int j = i + ischan;
                            for (int cc = 0; cc < nc; ++cc) CONC(cc, i) = CONC(cc, i) - CONC(cc, j);
                        }
                    }
                    rdelmax = 2.0;
                    if (jfail <= 6) {
                        ifsuccess = 0;
                        rdeltup = 0.0;
                        higher_order = 0;
                    } else if (jfail <= 20) {
                        ifsuccess = 0;
                        rdelt = FRACDEC;
                        continue;   /* CYCLE LOOP_ORDER */
                    } else {
                        ifsuccess = 1;
                        delt = delt * 0.1;
                        rdelt = 1.0;
                        jfail = 0;
                        jrestar = jrestar + 1;
                        idoub = 5;
                        for (int i = 1; i <= ischan; ++i)
                            for (int cc = 0; cc < nc; ++cc) A2(cnew, i - 1, cc) = CONC(cc, i);
                        if (jrestar == 100) { rc = -2; goto done; }
                        restart_init = 1;
                        break;      /* CYCLE LOOP_INITIALIZE */
                    }
                } else {
                    /* error test passed (grsmvgear.F:789-865) */
                    jfail = 0;
                    ifsuccess = 1;
                    told = xelaps;
                    nsteps = nsteps + 1;
                    st->nsteps = nsteps;
                    if (nqq > st->mxordused) st->mxordused = nqq;
                    int i1 = 1 - ischan;
                    for (int j = 1; j <= kstep; ++j) {
                        i1 = i1 + ischan;
                        for (int i = i1; i <= i1 + ischan1; ++i) {
                            int jspc = i - i1 + 1;
                            for (int cc = 0; cc < nc; ++cc)
                                CONC(cc, i) = CONC(cc, i) + g.aset[nqq][j] * A2(dtlos, jspc - 1, cc);
                        }
                    }
                    for (int i = 1; i <= ischan; ++i)
                        for (int cc = 0; cc < nc; ++cc)
                            if (CONC(cc, i) <= GCONCMIN) {
                                CONC(cc, i) = ZBOUND;
                                A2(cnew, i - 1, cc) = ZBOUND;
                                for (int k = 1; k <= 5; ++k) CONC(cc, i + k * ischan) = 0.0;
                            }
                    if (idoub > 1) {
                        idoub = idoub - 1;
                        if (idoub == 1)
                            for (int j = 1; j <= ischan; ++j)
                                for (int cc = 0; cc < nc; ++cc) A2(cest, j - 1, cc) = A2(dtlos, j - 1, cc);
                        rdelt = 1.0;
                        continue;   /* CYCLE LOOP_ORDER */
                    }
                }
                if (higher_order) {
                    if (nqq < MAXORD) {
                        for (int cc = 0; cc < nc; ++cc) dely[cc] = 0.0;
                        for (int j = 1; j <= ischan; ++j)
                            for (int cc = 0; cc < nc; ++cc) {
                                double e = (A2(dtlos, j - 1, cc) - A2(cest, j - 1, cc)) * A2(chold, j - 1, cc);
                                dely[cc] = dely[cc] + e * e;
                            }
                        double der3max = 0.0;
                        for (int cc = 0; cc < nc; ++cc)
                            if (dely[cc] > der3max) der3max = dely[cc];
                        double pr3 = (double)1.4f * gpow(der3max / eup, g.enqq3[nqq]);
                        rdeltup = 1.0 / (pr3 + OPT4MIC);
                    } else {
                        rdeltup = 0.0;
                    }
                }
                double pr2 = 1.2 * gpow(der2max / enqq, g.enqq2[nqq]);
                double rdeltsm = 1.0 / (pr2 + OPT2MIC);
                double rdeltdn;
                if (nqq > 1) {
                    for (int cc = 0; cc < nc; ++cc) dely[cc] = 0.0;
                    for (int j = 1; j <= ischan; ++j) {
                        int i = j + (kstep - 1) * ischan;
                        for (int cc = 0; cc < nc; ++cc) {
                            double e = CONC(cc, i) * A2(chold, j - 1, cc);
                            dely[cc] = dely[cc] + e * e;
                        }
                    }
                    double der1max = 0.0;
                    for (int cc = 0; cc < nc; ++cc)
                        if (dely[cc] > der1max) der1max = dely[cc];
                    double pr1 = 1.3 * gpow(der1max / edwn, g.enqq1[nqq]);
                    rdeltdn = 1.0 / (pr1 + OPT3MIC);
                } else {
                    rdeltdn = 0.0;
                }
                rdelt = rdeltsm;
                if (rdeltup > rdelt) rdelt = rdeltup;
                if (rdeltdn > rdelt) rdelt = rdeltdn;
                if (rdelt < (double)1.1f && ifsuccess == 1) {
                    idoub = 3;
                    continue;   /* CYCLE LOOP_ORDER */
                }
                if (rdelt == rdeltdn) {
                    nqq = nqq - 1;
                    if (ifsuccess == 0 && rdelt > 1.0) rdelt = 1.0;
                } else if (rdelt == rdeltup) {
                    double consmult = g.aset[nqq][kstep] / (double)kstep;
                    nqq = kstep;
                    int nqisc = nqq * ischan;
                    for (int j = 1; j <= ischan; ++j) {
                        int i1 = j + nqisc;
                        for (int cc = 0; cc < nc; ++cc) CONC(cc, i1) = A2(dtlos, j - 1, cc) * consmult;
                    }
                }
                if (jfail >= 2 && rdelt > PT2) rdelt = PT2;
                idoub = nqq + 1;
            }   /* LOOP_ORDER */
            if (restart_begin) break;
            if (!restart_init) break;
        }   /* LOOP_INITIALIZE */
        if (!restart_begin) break;
    }   /* LOOP_BEGIN */
done:
    if (rc == 0)
        for (size_t q = 0; q < sz; ++q)
            if (cnew[q] <= ZBOUND) cnew[q] = GCONCMIN;
#undef CONC
    free(gloss); free(conc); free(dtlos); free(cest); free(chold); free(blkconc); free(dely);
    free(ylowa); free(yloweps); free(yloweps1); free(cc2); free(vdiag);
    return rc;
}

/* The SMVGEAR block loop of CHEM (grdriver.F:400-560): blocks of blksize
 * cells (1 = per-cell control), CALC_RCONST per block, the block's
 * LSUNLIGHT choosing the day or night pattern and HMAX, SMVGEAR, and the
 * results back.  y [ns][ncells] in MECHANISM order, in/out, clamped to
 * CONCMIN on entry as the driver's gather does (grdriver.F:419-422).
 * nsteps / nfail per cell (the block's successful steps / error-test plus
 * convergence failures), -1 in nsteps marks a failed block.  Returns 0 or
 * the first failing block's code. */
int om_gear_run(const om_ctx *c, int64_t ncells, int blksize, double chemstep, const double *temp,
                const double *pres, const double *h2o, const double *rj, const double *het,
                const double *seawater, double *y, const om_gear_tol *tol, int32_t *nsteps,
                int32_t *nfail, const double *rki_in)
{
    const om_tables *t = &c->t;
    const int ns = c->ns, nr = c->nr;
    if (ncells <= 0) return 0;
    if (blksize < 1) return -4;
    int64_t nblks = 1 + (ncells - 1) / blksize;
    int64_t iavg = 1 + (ncells - 1) / nblks;
    if (iavg > blksize) iavg = blksize;
    int bmax = (int)(ncells - (nblks - 1) * iavg);
    if (bmax < iavg) bmax = (int)iavg;
    double *bt = malloc(bmax * sizeof(double)), *bp = malloc(bmax * sizeof(double));
    double *bh = malloc(bmax * sizeof(double)), *bs = malloc(bmax * sizeof(double));
    double *brj = malloc((size_t)(t->nphotab ? t->nphotab : 1) * bmax * sizeof(double));
    double *bhet = malloc((size_t)(t->nhetero ? t->nhetero : 1) * bmax * sizeof(double));
    double *brk = malloc((size_t)nr * bmax * sizeof(double));
    double *by = malloc((size_t)ns * bmax * sizeof(double));
    int rc = 0;
    for (int64_t b = 0; b < nblks; ++b) {
        int64_t off = b * iavg;
        int nc = (int)((b == nblks - 1) ? ncells - off : iavg);
        for (int cc = 0; cc < nc; ++cc) {
            bt[cc] = temp[off + cc]; bp[cc] = pres[off + cc]; bh[cc] = h2o[off + cc];
            bs[cc] = seawater ? seawater[off + cc] : 0.0;
            for (int i = 0; i < t->nphotab; ++i) brj[(size_t)i * nc + cc] = rj[(size_t)i * ncells + off + cc];
            for (int i = 0; i < t->nhetero; ++i) bhet[(size_t)i * nc + cc] = het ? het[(size_t)i * ncells + off + cc] : 0.0;
            for (int iold = 1; iold <= ns; ++iold) {
                double v = y[(size_t)(iold - 1) * ncells + off + cc];
                by[(size_t)(c->iold2new[iold] - 1) * nc + cc] = v > GCONCMIN ? v : GCONCMIN;
            }
        }
        int sunlit = om_rates_block(c, nc, bt, bp, bh, brj, bhet, bs, brk);
        if (rki_in) {
            /* given rate constants (the solver alone); LSUNLIGHT from J as above */
            for (int r = 0; r < nr; ++r)
                for (int cc = 0; cc < nc; ++cc) brk[(size_t)r * nc + cc] = rki_in[(size_t)r * ncells + off + cc];
        }
        om_gear_stats s;
        int brc = om_smvgear(c, sunlit ? 1 : 2, nc, chemstep, brk, by, tol, sunlit, &s);
        if (brc != 0 && rc == 0) rc = brc;
        for (int cc = 0; cc < nc; ++cc) {
            if (nsteps) nsteps[off + cc] = brc != 0 ? -1 : (int32_t)s.nsteps;
            if (nfail) nfail[off + cc] = (int32_t)(s.nefail + s.ncfail);
            for (int inew = 1; inew <= ns; ++inew)
                y[(size_t)(c->inew2old[inew] - 1) * ncells + off + cc] = by[(size_t)(inew - 1) * nc + cc];
        }
    }
    free(bt); free(bp); free(bh); free(bs); free(brj); free(bhet); free(brk); free(by);
    return rc;
}
