/*
 * CPU oracle for CMAQ's Rosenbrock ROS3 gas-chemistry path.
 *
 * TEST INFRASTRUCTURE ONLY (see ros3_oracle.h).  A plain-C restatement of
 *   CCTM/src/gas/ros3/rbsparse.F  (symbolic analysis)      -> om_create
 *   CCTM/src/gas/ros3/rbfeval.F   (f(y))                   -> om_feval
 *   CCTM/src/gas/ros3/rbjacob.F   (Jacobian)               -> om_jacob
 *   CCTM/src/gas/ros3/rbdecomp.F  (sparse LU)              -> om_decomp
 *   CCTM/src/gas/ros3/rbsolve.F   (triangular solves)      -> om_solve
 *   CCTM/src/gas/ros3/rbsolver.F  (ROS3 + step control)    -> om_rbsolver
 *   CCTM/src/gas/ros3/rbdriver.F:440-595 (block driver)    -> om_run
 *   MECHS/<mech>/RXNS_FUNC_MODULE.F90 CALC_RCONST, as written by
 *   UTIL/chemmech/src/wrt_reactions_module.f:580-800       -> om_rates_block
 * Arithmetic follows the Fortran expression order term by term; compile
 * with -ffp-contract=off (the reference's gcc flags, config_cmaq.csh:164,
 * do not enable FMA contraction on x86-64).
 *
 * Arrays here keep the Fortran 1-based indexing (slot 0 unused) so that each
 * loop can be compared line by line with the reference.
 */
#include "ros3_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"


static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); if (!p) abort(); return p; }

void om_destroy(om_ctx *c)
{
    if (!c) return;
    free(c->inew2old); free(c->iold2new); free(c->irm2);
    for (int n = 1; n <= 2; ++n) {
        free(c->nkuserat[n]); free(c->jarraypt[n]); free(c->kzero[n]);
        free(c->kzilch[n]); free(c->mzilch[n]); free(c->opbeg[n]);
        free(c->op_ij[n]); free(c->op_ik[n]); free(c->op_kj[n]);
        free(c->lbeg[n]); free(c->lsc[n]);
        free(c->nderivl[n]); free(c->nderivp[n]); free(c->jarrl[n]);
        free(c->jlial[n]); free(c->jarrp[n]); free(c->jpial[n]); free(c->icoeff[n]);
    }
    free(c->fall_idx); free(c->het_idx);
    free(c);
}

om_ctx *om_create(const om_tables *t, char *err, int errlen)
{
    om_ctx *c = xcalloc(1, sizeof *c);
    c->t = *t;
    const int ns = c->ns = t->ns, nr = c->nr = t->nr, mxprd = c->mxprd = t->mxprd;
    const int n1 = ns + 1;
    c->mxrr = 3 * MXRCT;             /* rbinit.F:101 */
    c->mxrp = 3 * mxprd;             /* rbinit.F:102 */

    /* ---- rate-table lookups and unsupported features ---- */
    c->fall_idx = xcalloc(nr + 1, sizeof(int));
    c->het_idx = xcalloc(nr + 1, sizeof(int));
    for (int f = t->nfalloff; f >= 1; --f) c->fall_idx[t->irrfall[f - 1]] = f;   /* first match */
    for (int h = t->mhetero; h >= 1; --h) c->het_idx[t->ihetero[(h - 1) * 2]] = t->ihetero[(h - 1) * 2 + 1];
    for (int r = 1; r <= nr; ++r) {
        int k = t->ktype[r - 1];
        if (k == 11 || k == 13 || k < -1 || k > 13) {
            snprintf(err, errlen, "reaction %d: rate type %d not supported", r, k);
            om_destroy(c); return NULL;
        }
        if ((k >= 8 && k <= 10) || k == 12) {
            if (!c->fall_idx[r]) { snprintf(err, errlen, "reaction %d: no RFDAT row", r); om_destroy(c); return NULL; }
        }
        if (k == 12 && t->iorder[r - 1] != 1) {
            snprintf(err, errlen, "reaction %d: halogen rate of order %d not supported", r, t->iorder[r - 1]);
            om_destroy(c); return NULL;
        }
        if (k == -1 && !c->het_idx[r]) { snprintf(err, errlen, "reaction %d: no IHETERO row", r); om_destroy(c); return NULL; }
        if (t->nreact[r - 1] > MXRCT) { snprintf(err, errlen, "reaction %d: >3 reactants", r); om_destroy(c); return NULL; }
    }
    for (int n = 0; n < t->nmphot; ++n) {
        if (t->iph[n * 3 + 2] == 0) {
            snprintf(err, errlen, "IPH(%d,3)=0 (photolysis relative to a reaction) not supported", n + 1);
            om_destroy(c); return NULL;
        }
    }

    /* ---- rbsparse.F:185-202: PD terms per species ---- */
    int *isaporl = xcalloc(n1, sizeof(int));
    char *isparder = xcalloc((size_t)n1 * n1, 1);      /* (iporr, ireact) */
    for (int nrx = 1; nrx <= nr; ++nrx)
        for (int nrr = 1; nrr <= 3; ++nrr) {
            int ireact = IRR(c, nrx, nrr);
            if (ireact != 0)
                for (int nrpp = 1; nrpp <= 3 + mxprd; ++nrpp) {
                    int iporr = IRR(c, nrx, nrpp);
                    if (iporr != 0) isparder[iporr * n1 + ireact] = 1;
                }
        }
    for (int ireact = 1; ireact <= ns; ++ireact)
        for (int iporr = 1; iporr <= ns; ++iporr)
            if (isparder[iporr * n1 + ireact]) isaporl[iporr]++;

    /* ---- rbsparse.F:209-252: species order ---- */
    c->inew2old = xcalloc(n1, sizeof(int));
    c->iold2new = xcalloc(n1, sizeof(int));
    int ischan = 0, nochang = ns;
    for (int jold = 1; jold <= ns; ++jold) {
        if (isaporl[jold] > 0) {
            ++ischan; c->inew2old[ischan] = jold; c->iold2new[jold] = ischan;
        } else {
            c->inew2old[nochang] = jold; c->iold2new[jold] = nochang; --nochang;
        }
    }
    for (int jnew = 1; jnew <= ischan; ++jnew) {
        int jold = c->inew2old[jnew], minvalu = isaporl[jold], iminold = jold, iminnew = jnew;
        for (int inew = jnew + 1; inew <= ischan; ++inew) {
            int iold = c->inew2old[inew];
            if (isaporl[iold] < minvalu) { minvalu = isaporl[iold]; iminold = iold; iminnew = inew; }
        }
        c->inew2old[iminnew] = jold; c->inew2old[jnew] = iminold;
        c->iold2new[jold] = iminnew; c->iold2new[iminold] = jnew;
    }
    c->ischan = ischan;
    free(isaporl); free(isparder);

    /* ---- rbsparse.F:259-281: IRM2 and active reaction lists ---- */
    c->irm2 = xcalloc((size_t)(nr + 1) * (mxprd + 4), sizeof(int));
    for (int n = 1; n <= 2; ++n) c->nkuserat[n] = xcalloc(nr + 1, sizeof(int));
    for (int nrx = 1; nrx <= nr; ++nrx) {
        for (int nrr = 1; nrr <= t->nreact[nrx - 1]; ++nrr) IRM2(c, nrx, nrr) = c->iold2new[IRR(c, nrx, nrr)];
        for (int np = 1; np <= t->nprdct[nrx - 1]; ++np) IRM2(c, nrx, np + 3) = c->iold2new[IRR(c, nrx, np + 3)];
        if (t->nreact[nrx - 1] > 0) {
            c->nkuserat[1][++c->nuserat[1]] = nrx;
            if (!(t->irxbits[nrx - 1] & 2)) c->nkuserat[2][++c->nuserat[2]] = nrx;
        }
    }

    /* ---- rbsparse.F:360-655, once for day (1) and night (2) ---- */
    char *lzero = xcalloc((size_t)n1 * n1, 1);           /* (row, col) */
#define LZ(i, j) lzero[(i) * n1 + (j)]
    for (int ncsp = 1; ncsp <= 2; ++ncsp) {
        memset(lzero, 0, (size_t)n1 * n1);
        for (int i = 1; i <= ns; ++i) LZ(i, i) = 1;
        for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
            int nk = c->nkuserat[ncsp][nrx];
            for (int nrr = 1; nrr <= t->nreact[nk - 1]; ++nrr) {
                int ire = IRM2(c, nk, nrr);
                for (int jal = 1; jal <= t->nreact[nk - 1]; ++jal) LZ(IRM2(c, nk, jal), ire) = 1;
                for (int iap = 1; iap <= t->nprdct[nk - 1]; ++iap) LZ(IRM2(c, nk, 3 + iap), ire) = 1;
            }
        }
        /* symbolic decomposition (:423-455) */
        int cap = 16, nop = 0, nls = 0, lcap = 16;
        int *oi = xcalloc(cap, sizeof(int)), *ok = xcalloc(cap, sizeof(int)), *oj = xcalloc(cap, sizeof(int));
        int *opcol = xcalloc(ischan + 2, sizeof(int));
        int *li = xcalloc(lcap, sizeof(int)), *lj = xcalloc(lcap, sizeof(int));
        int *lcol = xcalloc(ischan + 2, sizeof(int));
        for (int j = 1; j <= ischan; ++j) {
            int j1 = j - 1;
            opcol[j] = nop;
            for (int i = 2; i <= ischan; ++i) {
                int i1 = j1;
                if (i <= j1) i1 = i - 1;
                for (int k = 1; k <= i1; ++k) {
                    if (LZ(i, k) == 1 && LZ(k, j) == 1) {
                        if (nop == cap) {
                            cap *= 2;
                            oi = realloc(oi, cap * sizeof(int)); ok = realloc(ok, cap * sizeof(int)); oj = realloc(oj, cap * sizeof(int));
                        }
                        oi[nop] = i; ok[nop] = k; oj[nop] = j; ++nop;
                        LZ(i, j) = 1;
                    }
                }
            }
            lcol[j] = nls;
            for (int i = j + 1; i <= ischan; ++i) {
                if (LZ(i, j) == 1) {
                    if (nls == lcap) { lcap *= 2; li = realloc(li, lcap * sizeof(int)); lj = realloc(lj, lcap * sizeof(int)); }
                    li[nls] = i; lj[nls] = j; ++nls;
                }
            }
        }
        opcol[ischan + 1] = nop; lcol[ischan + 1] = nls;

        /* storage layout (:461-497) */
        c->jarraypt[ncsp] = xcalloc((size_t)n1 * n1, sizeof(int));
        c->kzilch[ncsp] = xcalloc(n1, sizeof(int));
        c->mzilch[ncsp] = xcalloc(n1, sizeof(int));
        int nnz = 0;
        for (int i = 1; i <= ischan; ++i) for (int j = 1; j <= ischan; ++j) nnz += LZ(i, j);
        c->kzero[ncsp] = xcalloc(nnz + 1, sizeof(int));
        int ia = 0;
        for (int i = 2; i <= ischan; ++i)
            for (int j = 1; j <= i - 1; ++j)
                if (LZ(i, j) == 1) { c->kzilch[ncsp][i]++; ++ia; c->kzero[ncsp][ia] = j; JAP(c, ncsp, i, j) = ia; }
        for (int i = ischan - 1; i >= 1; --i)
            for (int j = i + 1; j <= ischan; ++j)
                if (LZ(i, j) == 1) { c->mzilch[ncsp][i]++; ++ia; c->kzero[ncsp][ia] = j; JAP(c, ncsp, i, j) = ia; }
        for (int i = 1; i <= ischan; ++i) { ++ia; JAP(c, ncsp, i, i) = ia; c->kzero[ncsp][ia] = i; }
        c->iarray[ncsp] = ia;

        /* decomposition op lists in storage pointers (:503-554) */
        c->nops[ncsp] = nop;
        c->opbeg[ncsp] = xcalloc(ischan + 2, sizeof(int));
        c->op_ij[ncsp] = xcalloc(nop + 1, sizeof(int));
        c->op_ik[ncsp] = xcalloc(nop + 1, sizeof(int));
        c->op_kj[ncsp] = xcalloc(nop + 1, sizeof(int));
        for (int j = 1; j <= ischan + 1; ++j) c->opbeg[ncsp][j] = opcol[j];
        for (int q = 0; q < nop; ++q) {
            c->op_ij[ncsp][q] = JAP(c, ncsp, oi[q], oj[q]);
            c->op_ik[ncsp][q] = JAP(c, ncsp, oi[q], ok[q]);
            c->op_kj[ncsp][q] = JAP(c, ncsp, ok[q], oj[q]);
        }
        c->lbeg[ncsp] = xcalloc(ischan + 2, sizeof(int));
        c->lsc[ncsp] = xcalloc(nls + 1, sizeof(int));
        for (int j = 1; j <= ischan + 1; ++j) c->lbeg[ncsp][j] = lcol[j];
        for (int q = 0; q < nls; ++q) c->lsc[ncsp][q] = JAP(c, ncsp, li[q], lj[q]);
        free(oi); free(ok); free(oj); free(opcol); free(li); free(lj); free(lcol);

        /* PD pointers (:626-655) */
        c->nderivl[ncsp] = xcalloc(nr + 1, sizeof(int));
        c->nderivp[ncsp] = xcalloc(nr + 1, sizeof(int));
        c->jarrl[ncsp] = xcalloc((size_t)(nr + 1) * (c->mxrr + 1), sizeof(int));
        c->jlial[ncsp] = xcalloc((size_t)(nr + 1) * (c->mxrr + 1), sizeof(int));
        c->jarrp[ncsp] = xcalloc((size_t)(nr + 1) * (c->mxrp + 1), sizeof(int));
        c->jpial[ncsp] = xcalloc((size_t)(nr + 1) * (c->mxrp + 1), sizeof(int));
        c->icoeff[ncsp] = xcalloc((size_t)(nr + 1) * (c->mxrp + 1), sizeof(int));
        for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
            int nk = c->nkuserat[ncsp][nrx];
            for (int ial = 1; ial <= t->nreact[nk - 1]; ++ial) {
                int ir = IRM2(c, nk, ial);
                for (int jal = 1; jal <= t->nreact[nk - 1]; ++jal) {
                    int jr = IRM2(c, nk, jal);
                    int nl = ++c->nderivl[ncsp][nk];
                    c->jarrl[ncsp][nk * (c->mxrr + 1) + nl] = JAP(c, ncsp, jr, ir);
                    c->jlial[ncsp][nk * (c->mxrr + 1) + nl] = ial;
                }
                for (int iap = 1; iap <= t->nprdct[nk - 1]; ++iap) {
                    int jp = IRM2(c, nk, iap + 3);
                    int npr = ++c->nderivp[ncsp][nk];
                    c->jarrp[ncsp][nk * (c->mxrp + 1) + npr] = JAP(c, ncsp, jp, ir);
                    c->jpial[ncsp][nk * (c->mxrp + 1) + npr] = ial;
                    c->icoeff[ncsp][nk * (c->mxrp + 1) + npr] = (fabs(SC(c, nk, iap) - 1.0) > 1.0e-06) ? iap : 0;
                }
            }
        }
    }
#undef LZ
    free(lzero);
    return c;
}

int om_ischan(const om_ctx *c) { return c->ischan; }
int om_iarray(const om_ctx *c, int ncsp) { return c->iarray[ncsp]; }
void om_inew2old(const om_ctx *c, int32_t *out) { for (int i = 1; i <= c->ns; ++i) out[i - 1] = c->inew2old[i]; }
int om_nuserat(const om_ctx *c, int ncsp) { return c->nuserat[ncsp]; }
void om_nkuserat(const om_ctx *c, int ncsp, int32_t *out) { for (int i = 1; i <= c->nuserat[ncsp]; ++i) out[i - 1] = c->nkuserat[ncsp][i]; }
int om_nlu_ops(const om_ctx *c, int ncsp) { return c->nops[ncsp]; }
void om_lu_ops(const om_ctx *c, int ncsp, int32_t *ij, int32_t *ik, int32_t *kj)
{
    for (int q = 0; q < c->nops[ncsp]; ++q) { ij[q] = c->op_ij[ncsp][q]; ik[q] = c->op_ik[ncsp][q]; kj[q] = c->op_kj[ncsp][q]; }
}
void om_pattern(const om_ctx *c, int ncsp, int32_t *rows, int32_t *cols)
{
    for (int i = 1; i <= c->ischan; ++i)
        for (int j = 1; j <= c->ischan; ++j) {
            int s = JAP(c, ncsp, i, j);
            if (s) { rows[s - 1] = i; cols[s - 1] = j; }
        }
}

/* ------------------------------------------------------------------ */
/* CALC_RCONST (RXNS_FUNC_MODULE.F90:204-1116 for cb6r5_ae7_aq)        */
/* ------------------------------------------------------------------ */
static double power_t02(double t300, double a0, double b0) { return a0 * pow(t300, b0); }
static double arr_t03(double inv_t, double a0, double b0) { return a0 * exp(b0 * inv_t); }
static double arr_t04(double inv_t, double t300, double a0, double b0, double c0)
{ return a0 * exp(b0 * inv_t) * pow(t300, c0); }
static double fall_t08(double inv_t, double cair, double a0, double c0, double a2, double c2, double a3, double c3)
{
    double k0 = a0 * exp(c0 * inv_t), k2 = a2 * exp(c2 * inv_t), k3 = a3 * exp(c3 * inv_t);
    k3 = k3 * cair;
    return k0 + k3 / (1.0 + k3 / k2);
}
static double fall_t09(double inv_t, double cair, double a1, double c1, double a2, double c2)
{
    double k1 = a1 * exp(c1 * inv_t), k2 = a2 * exp(c2 * inv_t);
    return k1 + k2 * cair;
}
static double fall_t10(double inv_t, double t300, double cair, double a0, double b0, double c0,
                       double a1, double b1, double c1, double ce, double cf)
{
    double k0 = a0 * cair * exp(b0 * inv_t) * pow(t300, c0);
    double k1 = a1 * exp(b1 * inv_t) * pow(t300, c1);
    double g = (1.0 / ce) * log10(k0 / k1);
    double kend = 1.0 + g * g;
    kend = 1.0 / kend;
    return (k0 / (1.0 + k0 / k1)) * pow(cf, kend);
}
static double fall_t11(double inv_t, double t300, double cair, double a1, double b1, double c1,
                       double a2, double b2, double c2, double d1, double d2)
{
    double k1 = a1 * exp(c1 * inv_t) * pow(t300, b1);
    double k2 = a2 * exp(c2 * inv_t) * pow(t300, b2);
    double k3 = d1 * exp(d2 * inv_t);
    return k1 + k2 * cair + k3;
}
static double halogen_falloff(double p, double a1, double b1, double a2, double b2, double a3)
{
    double h = a1 * exp(b1 * p) + a2 * exp(b2 * p);
    return a3 < h ? a3 : h;
}
/* chemmech writes computed constants with FORMAT 1PD12.4 */
static double d12_4(double v) { char b[64]; snprintf(b, sizeof b, "%.4e", v); return strtod(b, NULL); }

typedef struct { double cfact, cfact_squ, inv_cfact, sfact, rfact, rfact_squ; } om_conv;
/* WRITE_RATE_CONVERT_BEFORE/AFTER (UTIL/chemmech/src/BASIC_WRITE_ROUTINES.f:1102-1152);
 * returns 0 when chemmech writes no factor for this order. */
static int conv_factor(const om_conv *u, int order, double *f)
{
    switch (order) {
    case 0: *f = u->inv_cfact; return 1;
    case 1: *f = u->sfact; return 1;
    case 2: *f = u->cfact; return 1;
    case 3: *f = u->cfact_squ; return 1;
    case -1: *f = u->rfact; return 1;
    case -2: *f = u->rfact_squ; return 1;
    default: return 0;
    }
}

int om_rates_block(const om_ctx *c, int nc, const double *temp, const double *pres,
                   const double *h2o, const double *rj, const double *het,
                   const double *seawater, double *rki)
{
    const om_tables *t = &c->t;
    const double COEF1 = 7.33981e+15, TI300 = 1.0 / 300.0, SFACT = 60.0;
    int sunlit = 0;
    for (int i = 0; i < t->nphotab; ++i)
        for (int cc = 0; cc < nc; ++cc)
            if (rj[(size_t)i * nc + cc] > 0.0) sunlit = 1;
    for (size_t q = 0; q < (size_t)t->nr * nc; ++q) rki[q] = 0.0;
#define RK(r) rki[(size_t)((r) - 1) * nc + cc]
    if (sunlit) {
        for (int cc = 0; cc < nc; ++cc) {
            for (int n = 0; n < t->nmphot; ++n) {
                int nxx = t->iph[n * 3], idx = t->iph[n * 3 + 1];
                if (nxx <= 0) continue;
                double j = rj[(size_t)(idx - 1) * nc + cc];
                RK(nxx) = (RTDAT(c, 1, nxx) != 1.0) ? RTDAT(c, 1, nxx) * j : j;
            }
            for (int nxx = 1; nxx <= t->nr; ++nxx) {
                if (t->ktype[nxx - 1] != 12) continue;
                if (seawater[cc] > 0.001) {
                    int f = c->fall_idx[nxx];
                    double h = halogen_falloff(pres[cc], RTDAT(c, 1, nxx), RFDAT(c, 1, f),
                                               RTDAT(c, 2, nxx), RFDAT(c, 2, f), RTDAT(c, 3, nxx));
                    RK(nxx) = seawater[cc] * SFACT * h;     /* first order only */
                } else {
                    RK(nxx) = 0.0;
                }
            }
        }
    }
    for (int cc = 0; cc < nc; ++cc) {
        double inv_temp = 1.0 / temp[cc];
        double cair = 1.0e+06 * COEF1 * pres[cc] * inv_temp;
        om_conv u;
        u.cfact = 6.0e-05 * cair;
        u.cfact_squ = 6.0e-11 * cair * cair;
        double inv_cair = 1.0 / cair;
        u.inv_cfact = 6.0e+07 * inv_cair;
        double t300 = temp[cc] * TI300;
        u.rfact = 1.0e+06 * inv_cair;
        u.rfact_squ = 1.0e+12 * inv_cair * inv_cair;
        u.sfact = SFACT;
        for (int nxx = 1; nxx <= t->nr; ++nxx) {
            int kt = t->ktype[nxx - 1], ord = t->iorder[nxx - 1];
            double a = RTDAT(c, 1, nxx), b = RTDAT(c, 2, nxx), cc3 = RTDAT(c, 3, nxx);
            double f, v;
            int f_ok;
            int fi = c->fall_idx[nxx];
            switch (kt) {
            case -1: {
                double h = het[(size_t)(c->het_idx[nxx] - 1) * nc + cc];
                RK(nxx) = (a != 1.0) ? a * h : h;
                break;
            }
            case 1:
                RK(nxx) = conv_factor(&u, ord, &f) ? a * f : a;
                break;
            case 2:
                v = power_t02(t300, a, b);
                RK(nxx) = (t->kunits == 2 && conv_factor(&u, ord, &f)) ? f * v : v;
                break;
            case 3:
                v = arr_t03(inv_temp, a, cc3);
                RK(nxx) = (t->kunits == 2 && conv_factor(&u, ord, &f)) ? f * v : v;
                break;
            case 4:
                v = arr_t04(inv_temp, t300, a, cc3, b);
                RK(nxx) = (t->kunits == 2 && conv_factor(&u, ord, &f)) ? f * v : v;
                break;
            case 5: {
                int irx = (int)cc3;
                if (irx > nxx) break;                      /* never written by chemmech */
                int idiff = ord - t->iorder[irx - 1];
                int kti = t->ktype[irx - 1];
                int falloff = (kti > 7 && kti < 11);
                double eq = d12_4(1.0 / a) * exp(-b * inv_temp);
                if (idiff != 0 && (t->kunits == 2 || falloff) && conv_factor(&u, idiff, &f))
                    RK(nxx) = f * RK(irx) * eq;
                else
                    RK(nxx) = RK(irx) * eq;
                break;
            }
            case 6: {
                int irx = (int)b;
                if (irx > nxx) break;                      /* never written by chemmech */
                int idiff = ord - t->iorder[irx - 1];
                int kti = t->ktype[irx - 1];
                int falloff = (kti > 7 && kti < 11);
                f_ok = idiff != 0 && (t->kunits == 2 || falloff) && conv_factor(&u, idiff, &f);
                if (a != 1.0) RK(nxx) = f_ok ? f * a * RK(irx) : a * RK(irx);
                else RK(nxx) = f_ok ? f * RK(irx) : RK(irx);
                break;
            }
            case 7:   /* wrt_reactions_module.f:731-737, formats 5007 and 5114: no unit factor */
                if (b != 0.0) RK(nxx) = arr_t04(inv_temp, t300, d12_4(a), 0.0, d12_4(b)) * pres[cc];
                else RK(nxx) = d12_4(a) * (1.0 + 0.6 * pres[cc]);
                break;
            case 8:
                v = fall_t08(inv_temp, cair, a, b, cc3, RFDAT(c, 1, fi), RFDAT(c, 2, fi), RFDAT(c, 3, fi));
                RK(nxx) = conv_factor(&u, ord, &f) ? f * v : v;
                break;
            case 9:
                if (RFDAT(c, 2, fi) == 0.0 && RFDAT(c, 3, fi) == 0.0)
                    v = fall_t09(inv_temp, cair, a, b, cc3, RFDAT(c, 1, fi));
                else
                    v = fall_t11(inv_temp, t300, cair, a, RFDAT(c, 2, fi), b, cc3, RFDAT(c, 3, fi),
                                 RFDAT(c, 1, fi), RFDAT(c, 4, fi), RFDAT(c, 5, fi));
                RK(nxx) = conv_factor(&u, ord, &f) ? f * v : v;
                break;
            case 10:
                v = fall_t10(inv_temp, t300, cair, a, cc3, b, RFDAT(c, 1, fi), RFDAT(c, 3, fi),
                             RFDAT(c, 2, fi), RFDAT(c, 5, fi), RFDAT(c, 4, fi));
                RK(nxx) = conv_factor(&u, ord, &f) ? f * v : v;
                break;
            default:
                break;                                     /* 0 and 12: sunlit block */
            }
        }
    }
    /* third-body and fixed-species multipliers (:1065-1115) */
    struct { int n; const int32_t *l; int h2o; double v; } mult[6] = {
        {t->nwm, t->nrxwm, 0, t->atm_air}, {t->nwo2, t->nrxwo2, 0, t->atm_o2},
        {t->nwn2, t->nrxwn2, 0, t->atm_n2}, {t->nww, t->nrxww, 1, 0.0},
        {t->nwh2, t->nrxwh2, 0, t->atm_h2}, {t->nwch4, t->nrxwch4, 0, t->atm_ch4}};
    for (int m = 0; m < 6; ++m)
        for (int q = 0; q < mult[m].n; ++q) {
            int r = mult[m].l[q];
            for (int cc = 0; cc < nc; ++cc) RK(r) = RK(r) * (mult[m].h2o ? h2o[cc] : mult[m].v);
        }
#undef RK
    return sunlit;
}

/* ------------------------------------------------------------------ */
/* Numerical kernels on one block                                     */
/* ------------------------------------------------------------------ */
#define A2(a, i, cc) (a)[(size_t)(i) * nc + (cc)]

void om_feval(const om_ctx *c, int ncsp, int nc, const double *rki, const double *y, double *ydot)
{
    const om_tables *t = &c->t;
    /* y / ydot are indexed by new species id - 1 */
    for (int isp = 1; isp <= c->ns; ++isp)
        for (int cc = 0; cc < nc; ++cc) A2(ydot, isp - 1, cc) = 0.0;
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int nrk = c->nkuserat[ncsp][nrx], nre = t->nreact[nrk - 1];
        for (int cc = 0; cc < nc; ++cc) {
            double rx = A2(rki, nrk - 1, cc);
            if (nre >= 1) rx = rx * A2(y, IRM2(c, nrk, 1) - 1, cc);
            if (nre >= 2) rx = rx * A2(y, IRM2(c, nrk, 2) - 1, cc);
            if (nre >= 3) rx = rx * A2(y, IRM2(c, nrk, 3) - 1, cc);
            for (int nrr = 1; nrr <= nre; ++nrr) {
                int s = IRM2(c, nrk, nrr) - 1;
                A2(ydot, s, cc) = A2(ydot, s, cc) - rx;
            }
            for (int np = 1; np <= t->nprdct[nrk - 1]; ++np) {
                int s = IRM2(c, nrk, np + 3) - 1;
                A2(ydot, s, cc) = A2(ydot, s, cc) + SC(c, nrk, np) * rx;
            }
        }
    }
}

void om_jacob(const om_ctx *c, int ncsp, int nc, const double *rki, const double *y, double *cc2)
{
    const om_tables *t = &c->t;
    for (int iar = 0; iar <= c->iarray[ncsp]; ++iar)
        for (int cc = 0; cc < nc; ++cc) A2(cc2, iar, cc) = 0.0;
    double ex[4];
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int nrk = c->nkuserat[ncsp][nrx], nre = t->nreact[nrk - 1];
        for (int cc = 0; cc < nc; ++cc) {
            double rk = A2(rki, nrk - 1, cc);
            if (nre == 1) {
                ex[1] = rk;
            } else if (nre == 2) {
                int jr1 = IRM2(c, nrk, 1) - 1, jr2 = IRM2(c, nrk, 2) - 1;
                ex[1] = rk * A2(y, jr2, cc);
                ex[2] = rk * A2(y, jr1, cc);
            } else if (nre == 3) {
                int jr1 = IRM2(c, nrk, 1) - 1, jr2 = IRM2(c, nrk, 2) - 1, jr3 = IRM2(c, nrk, 3) - 1;
                double cr2 = rk * A2(y, jr2, cc);
                ex[1] = cr2 * A2(y, jr3, cc);
                ex[2] = rk * A2(y, jr1, cc) * A2(y, jr3, cc);
                ex[3] = cr2 * A2(y, jr1, cc);
            }
            for (int nl = 1; nl <= c->nderivl[ncsp][nrk]; ++nl) {
                int iarp = c->jarrl[ncsp][nrk * (c->mxrr + 1) + nl];
                int ialp = c->jlial[ncsp][nrk * (c->mxrr + 1) + nl];
                A2(cc2, iarp, cc) = A2(cc2, iarp, cc) - ex[ialp];
            }
            for (int np = 1; np <= c->nderivp[ncsp][nrk]; ++np) {
                int q = nrk * (c->mxrp + 1) + np;
                int iarp = c->jarrp[ncsp][q], ialp = c->jpial[ncsp][q], isc = c->icoeff[ncsp][q];
                if (isc == 0) A2(cc2, iarp, cc) = A2(cc2, iarp, cc) + ex[ialp];
                else A2(cc2, iarp, cc) = A2(cc2, iarp, cc) + SC(c, nrk, isc) * ex[ialp];
            }
        }
    }
}

void om_decomp(const om_ctx *c, int ncsp, int nc, double *cc2, double *vdiag)
{
    for (int j = 1; j <= c->ischan; ++j) {
        for (int q = c->opbeg[ncsp][j]; q < c->opbeg[ncsp][j + 1]; ++q) {
            int ij = c->op_ij[ncsp][q], ik = c->op_ik[ncsp][q], kj = c->op_kj[ncsp][q];
            for (int cc = 0; cc < nc; ++cc)
                A2(cc2, ij, cc) = A2(cc2, ij, cc) - A2(cc2, ik, cc) * A2(cc2, kj, cc);
        }
        int iar = JAP(c, ncsp, j, j);
        for (int cc = 0; cc < nc; ++cc) A2(vdiag, j - 1, cc) = 1.0 / A2(cc2, iar, cc);
        for (int q = c->lbeg[ncsp][j]; q < c->lbeg[ncsp][j + 1]; ++q) {
            int ija = c->lsc[ncsp][q];
            for (int cc = 0; cc < nc; ++cc) A2(cc2, ija, cc) = A2(cc2, ija, cc) * A2(vdiag, j - 1, cc);
        }
    }
}

void om_solve(const om_ctx *c, int ncsp, int nc, const double *cc2, const double *vdiag, double *rhs)
{
    const int *kz = c->kzero[ncsp];
    int ij = 1;
    for (int i = 1; i <= c->ischan; ++i)
        for (int q = 0; q < c->kzilch[ncsp][i]; ++q, ++ij)
            for (int cc = 0; cc < nc; ++cc)
                A2(rhs, i - 1, cc) = A2(rhs, i - 1, cc) - A2(cc2, ij, cc) * A2(rhs, kz[ij] - 1, cc);
    for (int i = c->ischan; i >= 1; --i) {
        for (int q = 0; q < c->mzilch[ncsp][i]; ++q, ++ij)
            for (int cc = 0; cc < nc; ++cc)
                A2(rhs, i - 1, cc) = A2(rhs, i - 1, cc) - A2(cc2, ij, cc) * A2(rhs, kz[ij] - 1, cc);
        for (int cc = 0; cc < nc; ++cc) A2(rhs, i - 1, cc) = A2(rhs, i - 1, cc) * A2(vdiag, i - 1, cc);
    }
}

/* ROS3 constants (rbsolver.F:85-113) */
static const double GAM = 0.43586652150845899941601945119356e+00;
static const double C21 = -0.10156171083877702091975600115545e+01;
static const double C31 = 0.40759956452537699824805835358067e+01;
static const double C32 = 0.92076794298330791242156818474003e+01;
static const double B1 = 0.10000000000000000000000000000000e+01;
static const double B2 = 0.61697947043828245592553615689730e+01;
static const double B3 = -0.42772256543218573326238373806514e+00;
static const double D1 = 0.50000000000000000000000000000000e+00;
static const double D2 = -0.29079558716805469821718236208017e+01;
static const double D3 = 0.22354069897811569627360909276199e+00;
#define GROW (1.0 / 3.00)
#define DTSTART 0.5e-01
#define DTMIN 1.0e-08
#define DTMAX 1.0e+01
#define UROUND 1.0e-18
#define FACMAX 1.0e+01
#define FACMIN 1.0e-01
#define FACONE 1.0
#define CONMIN 1.0e-30
#define OM_MAX_ATTEMPTS 10000000L

int om_rbsolver(const om_ctx *c, int ncsp, int nc, double chemstep, const double *rki,
                double *y, const double *rtol, const double *atol,
                int32_t *nacc, int32_t *nrej, om_fail *fail)
{
    const int ns = c->ns, ischan = c->ischan;
    const double RGAM = 1.0 / GAM;
    const double rnspec = (double)(1.0f / (float)ns);   /* REAL(4) division, rbsolver.F:246 */
    const int idiagbeg = c->iarray[ncsp] - ischan + 1;
    size_t sz = (size_t)ns * nc;
    double *ydot = malloc(sz * sizeof(double)), *k1 = malloc(sz * sizeof(double));
    double *k2 = malloc(sz * sizeof(double)), *k3 = malloc(sz * sizeof(double));
    double *yp = malloc(sz * sizeof(double)), *err = malloc(nc * sizeof(double));
    double *cc2 = malloc((size_t)(c->iarray[ncsp] + 1) * nc * sizeof(double));
    double *vdiag = malloc((size_t)ischan * nc * sizeof(double));
    double dt = DTMIN > DTSTART ? DTMIN : DTSTART;       /* MAX(DTMIN, DTSTART) */
    double tnow = 0.0;
    int lpass = 0, rc = 0;
    long attempts = 0;
    *nacc = 0; *nrej = 0;
    while (tnow < chemstep) {
        if (++attempts > OM_MAX_ATTEMPTS) { rc = -2; break; }
        om_feval(c, ncsp, nc, rki, y, ydot);
        double tend = tnow + dt;
        if (tend > chemstep) { dt = chemstep - tnow; tend = chemstep; }
        double dtinv = 1.0 / dt;
        double gdtinv = dtinv * RGAM;
        om_jacob(c, ncsp, nc, rki, y, cc2);
        for (int j = idiagbeg; j <= c->iarray[ncsp]; ++j)
            for (int cc = 0; cc < nc; ++cc) A2(cc2, j, cc) = A2(cc2, j, cc) - gdtinv;
        om_decomp(c, ncsp, nc, cc2, vdiag);
        for (size_t q = 0; q < sz; ++q) k1[q] = -ydot[q];
        om_solve(c, ncsp, nc, cc2, vdiag, k1);
        for (size_t q = 0; q < sz; ++q) yp[q] = y[q] + k1[q];
        om_feval(c, ncsp, nc, rki, yp, ydot);
        double x1 = C21 * dtinv;
        for (size_t q = 0; q < sz; ++q) k2[q] = -ydot[q] - x1 * k1[q];
        om_solve(c, ncsp, nc, cc2, vdiag, k2);
        x1 = C31 * dtinv;
        double x2 = C32 * dtinv;
        for (size_t q = 0; q < sz; ++q) k3[q] = -ydot[q] - x1 * k1[q] - x2 * k2[q];
        om_solve(c, ncsp, nc, cc2, vdiag, k3);
        for (size_t q = 0; q < sz; ++q) yp[q] = y[q] + B1 * k1[q] + B2 * k2[q] + B3 * k3[q];
        for (int cc = 0; cc < nc; ++cc) err[cc] = 0.0;
        for (int n = 0; n < ns; ++n)
            for (int cc = 0; cc < nc; ++cc) {
                double ytol = atol[n] + rtol[n] * fabs(A2(yp, n, cc));
                double e = (D1 * A2(k1, n, cc) + D2 * A2(k2, n, cc) + D3 * A2(k3, n, cc)) / ytol;
                err[cc] = err[cc] + e * e;
            }
        double maxerr = 0.0;
        for (int cc = 0; cc < nc; ++cc) {
            double s = sqrt(err[cc] * rnspec);
            if (UROUND > maxerr) maxerr = UROUND;
            if (!(s <= maxerr)) maxerr = s;          /* a NaN error rejects the step */
        }
        double dtfac = 0.9 / pow(maxerr, GROW);
        if (maxerr <= 1.0) {
            (*nacc)++;
            for (size_t q = 0; q < sz; ++q) y[q] = yp[q] > CONMIN ? yp[q] : CONMIN;
            tnow = tend;
            double hi = lpass ? FACMAX : FACONE;
            if (dtfac > hi) dtfac = hi;
            if (dtfac < FACMIN) dtfac = FACMIN;
            double dn = dtfac * dt;
            if (dn < DTMIN) dn = DTMIN;
            dt = dn < DTMAX ? dn : DTMAX;
            lpass = 1;
        } else {
            (*nrej)++;
            dt = FACMIN * dt;
            lpass = 0;
            if (dt < DTMIN) {
                if (fail) { fail->maxerr = maxerr; fail->dt = dt; fail->tnow = tnow; }
                if (fail && fail->dump) {
                    /* CELL_MAXERR: the first cell holding the block maximum
                       (rbsolver.F:446-455), then its ERR, Y, YP, K1..K3 */
                    int cm = 0;
                    double best = UROUND;
                    for (int cc = 0; cc < nc; ++cc) {
                        double s = sqrt(err[cc] * rnspec);
                        if (s > best) { best = s; cm = cc; }
                    }
                    double *d = fail->dump;
                    d[0] = err[cm];
                    for (int n = 0; n < ns; ++n) {
                        d[1 + n] = A2(y, n, cm);
                        d[1 + ns + n] = A2(yp, n, cm);
                        d[1 + 2 * ns + n] = A2(k1, n, cm);
                        d[1 + 3 * ns + n] = A2(k2, n, cm);
                        d[1 + 4 * ns + n] = A2(k3, n, cm);
                    }
                }
                rc = -1;
                break;
            }
        }
    }
    free(ydot); free(k1); free(k2); free(k3); free(yp); free(err); free(cc2); free(vdiag);
    return rc;
}

int om_run(const om_ctx *c, int64_t ncells, int blksize, double chemstep,
           const double *temp, const double *pres, const double *h2o,
           const double *rj, const double *het, const double *seawater,
           double *y, const double *rtol, const double *atol,
           int32_t *nacc, int32_t *nrej, om_fail *fail)
{
    const om_tables *t = &c->t;
    const int ns = c->ns, nr = c->nr;
    if (fail) { fail->cell = -1; fail->maxerr = 0; fail->dt = 0; fail->tnow = 0; }
    if (ncells <= 0) return 0;
    if (fail && fail->dump && blksize != 1) return -3;   /* the dump names one cell */
    if (blksize < 1) return -3;
    /* block partition (rbinit.F:286-303) */
    int64_t nblks = 1 + (ncells - 1) / blksize;
    int64_t iavg = 1 + (ncells - 1) / nblks;
    if (iavg > blksize) iavg = blksize;
    double *rtn = malloc(ns * sizeof(double)), *atn = malloc(ns * sizeof(double));
    for (int i = 1; i <= ns; ++i) { rtn[i - 1] = rtol[c->inew2old[i] - 1]; atn[i - 1] = atol[c->inew2old[i] - 1]; }
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
                by[(size_t)(c->iold2new[iold] - 1) * nc + cc] = y[(size_t)(iold - 1) * ncells + off + cc];
            }
        }
        int sunlit = om_rates_block(c, nc, bt, bp, bh, brj, bhet, bs, brk);
        int ncsp = sunlit ? 1 : 2;
        int32_t na = 0, nj = 0;
        om_fail f1;
        f1.dump = (rc == 0 && fail) ? fail->dump : NULL;
        int brc = om_rbsolver(c, ncsp, nc, chemstep, brk, by, rtn, atn, &na, &nj, &f1);
        if (brc != 0) {
            if (rc == 0 && fail) { *fail = f1; fail->cell = off; }
            if (rc == 0) rc = brc;
            na = -1;                          /* marks the cells of a failed block */
        }
        for (int cc = 0; cc < nc; ++cc) {
            if (nacc) nacc[off + cc] = na;
            if (nrej) nrej[off + cc] = nj;
            for (int inew = 1; inew <= ns; ++inew)
                y[(size_t)(c->inew2old[inew] - 1) * ncells + off + cc] = by[(size_t)(inew - 1) * nc + cc];
        }
    }
    free(rtn); free(atn); free(bt); free(bp); free(bh); free(bs); free(brj); free(bhet); free(brk); free(by);
    return rc;
}

/* The LORDERING pre-pass of the ros3 CHEM driver (rbdriver.F:430-434,
 * 440-521; rbsolver.F:346-360): for every cell, in blocks of blksize
 * consecutive cells with the block's LSUNLIGHT deciding the pattern, one
 * CALC_RCONST and one RBFEVAL at the initial state and the stiffness
 * estimate ERRMX2 = sum over the ISCHAN species of
 * (YDOT / (Y + ATOL / MIN(RTOL, 0.003)))^2.  The driver then heap-sorts the
 * cells by ERRMX2 (rbdriver.F:683-724) and integrates them in that order.
 * y is in MECHANISM order [ns][ncells]; rtol/atol per species in mechanism
 * order; errmx2 [ncells] out. */
void om_stiffness(const om_ctx *c, int64_t ncells, int blksize, const double *temp, const double *pres,
                  const double *h2o, const double *rj, const double *het, const double *seawater,
                  const double *y, const double *rtol, const double *atol, double *errmx2)
{
    const om_tables *t = &c->t;
    const int ns = c->ns, nr = c->nr;
    if (ncells <= 0 || blksize < 1) return;
    double *bt = malloc(blksize * sizeof(double)), *bp = malloc(blksize * sizeof(double));
    double *bh = malloc(blksize * sizeof(double)), *bs = malloc(blksize * sizeof(double));
    double *brj = malloc((size_t)(t->nphotab ? t->nphotab : 1) * blksize * sizeof(double));
    double *bhet = malloc((size_t)(t->nhetero ? t->nhetero : 1) * blksize * sizeof(double));
    double *brk = malloc((size_t)nr * blksize * sizeof(double));
    double *by = malloc((size_t)ns * blksize * sizeof(double));
    double *bd = malloc((size_t)ns * blksize * sizeof(double));
    for (int64_t off = 0; off < ncells; off += blksize) {
        const int nc = (int)(ncells - off < blksize ? ncells - off : blksize);
        for (int cc = 0; cc < nc; ++cc) {
            bt[cc] = temp[off + cc]; bp[cc] = pres[off + cc]; bh[cc] = h2o[off + cc];
            bs[cc] = seawater ? seawater[off + cc] : 0.0;
            for (int i = 0; i < t->nphotab; ++i) brj[(size_t)i * nc + cc] = rj[(size_t)i * ncells + off + cc];
            for (int i = 0; i < t->nhetero; ++i) bhet[(size_t)i * nc + cc] = het ? het[(size_t)i * ncells + off + cc] : 0.0;
            for (int iold = 1; iold <= ns; ++iold)
                by[(size_t)(c->iold2new[iold] - 1) * nc + cc] = y[(size_t)(iold - 1) * ncells + off + cc];
        }
        const int ncsp = om_rates_block(c, nc, bt, bp, bh, brj, bhet, bs, brk) ? 1 : 2;
        om_feval(c, ncsp, nc, brk, by, bd);
        for (int cc = 0; cc < nc; ++cc) errmx2[off + cc] = 0.0;
        for (int j = 1; j <= c->ischan; ++j) {
            const int old = c->inew2old[j] - 1;
            const double rt = rtol[old] < 0.003 ? rtol[old] : 0.003;
            const double ylow = atol[old] / rt;
            for (int cc = 0; cc < nc; ++cc) {
                const double e = A2(bd, j - 1, cc) / (A2(by, j - 1, cc) + ylow);
                errmx2[off + cc] = errmx2[off + cc] + e * e;
            }
        }
    }
    free(bt); free(bp); free(bh); free(bs); free(brj); free(bhet); free(brk); free(by); free(bd);
}

/* ------------------------------------------------------------------ */
/* Per-cell step tape, fixed-step replay and discrete adjoint          */
/* ------------------------------------------------------------------ */
/* The reference path has no adjoint (its sensitivity option is DDM-3D,
 * ddm3d/DDM3D_CHEM.F).  These functions hold the accepted-step sequence of
 * om_rbsolver fixed and differentiate the resulting map y0 -> y_end (and
 * rki -> y_end) exactly; the step-size controller is not differentiated.
 * Pinned by central finite differences of om_replay_cell and by
 * om_replay_cell reproducing om_rbsolver bit for bit.  One cell (nc = 1),
 * arrays in the solver's (new) species order. */

/* One ROS3 step of size dt from y, as in om_rbsolver (rbsolver.F:357-430):
 * cc2/vdiag = LU of J(y) - I/(GAM dt); k1..k3 the stages; yp = y + K1;
 * yn = y + B1 K1 + B2 K2 + B3 K3 (before the CONMIN clamp). */
static void ros3_stages1(const om_ctx *c, int ncsp, const double *rki, const double *y, double dt,
                         double *ydot, double *cc2, double *vdiag, double *k1, double *k2,
                         double *k3, double *yp, double *yn)
{
    const int ns = c->ns, nc = 1;
    const double RGAM = 1.0 / GAM;
    const int idiagbeg = c->iarray[ncsp] - c->ischan + 1;
    om_feval(c, ncsp, nc, rki, y, ydot);
    double dtinv = 1.0 / dt;
    double gdtinv = dtinv * RGAM;
    om_jacob(c, ncsp, nc, rki, y, cc2);
    for (int j = idiagbeg; j <= c->iarray[ncsp]; ++j) cc2[j] = cc2[j] - gdtinv;
    om_decomp(c, ncsp, nc, cc2, vdiag);
    for (int q = 0; q < ns; ++q) k1[q] = -ydot[q];
    om_solve(c, ncsp, nc, cc2, vdiag, k1);
    for (int q = 0; q < ns; ++q) yp[q] = y[q] + k1[q];
    om_feval(c, ncsp, nc, rki, yp, ydot);
    double x1 = C21 * dtinv;
    for (int q = 0; q < ns; ++q) k2[q] = -ydot[q] - x1 * k1[q];
    om_solve(c, ncsp, nc, cc2, vdiag, k2);
    x1 = C31 * dtinv;
    double x2 = C32 * dtinv;
    for (int q = 0; q < ns; ++q) k3[q] = -ydot[q] - x1 * k1[q] - x2 * k2[q];
    om_solve(c, ncsp, nc, cc2, vdiag, k3);
    for (int q = 0; q < ns; ++q) yn[q] = y[q] + B1 * k1[q] + B2 * k2[q] + B3 * k3[q];
}

typedef struct { double *ydot, *cc2, *vdiag, *k1, *k2, *k3, *yp, *yn, *u1, *u2, *u3, *yb, *s2, *s3; } om_work;

static void work_alloc(const om_ctx *c, int ncsp, om_work *w)
{
    size_t ns = c->ns;
    w->ydot = xcalloc(ns, 8); w->k1 = xcalloc(ns, 8); w->k2 = xcalloc(ns, 8); w->k3 = xcalloc(ns, 8);
    w->yp = xcalloc(ns, 8); w->yn = xcalloc(ns, 8); w->u1 = xcalloc(ns, 8); w->u2 = xcalloc(ns, 8);
    w->u3 = xcalloc(ns, 8); w->yb = xcalloc(ns, 8); w->vdiag = xcalloc(ns, 8);
    w->cc2 = xcalloc((size_t)c->iarray[ncsp] + 1, 8);
    w->s2 = xcalloc((size_t)c->nr + 1, 8); w->s3 = xcalloc((size_t)c->nr + 1, 8);
}

static void work_free(om_work *w)
{
    free(w->ydot); free(w->k1); free(w->k2); free(w->k3); free(w->yp); free(w->yn); free(w->u1);
    free(w->u2); free(w->u3); free(w->yb); free(w->vdiag); free(w->cc2); free(w->s2); free(w->s3);
}

int om_steps_cell(const om_ctx *c, int ncsp, double chemstep, const double *rki, double *y,
                  const double *rtol, const double *atol, double *h, int hcap, int *nh)
{
    /* om_rbsolver's controller on one cell, recording each accepted DT */
    const int ns = c->ns;
    const double rnspec = (double)(1.0f / (float)ns);
    om_work w;
    work_alloc(c, ncsp, &w);
    double dt = DTMIN > DTSTART ? DTMIN : DTSTART, tnow = 0.0;
    int lpass = 0, rc = 0;
    long attempts = 0;
    *nh = 0;
    while (tnow < chemstep) {
        if (++attempts > OM_MAX_ATTEMPTS) { rc = -2; break; }
        double tend = tnow + dt;
        if (tend > chemstep) { dt = chemstep - tnow; tend = chemstep; }
        ros3_stages1(c, ncsp, rki, y, dt, w.ydot, w.cc2, w.vdiag, w.k1, w.k2, w.k3, w.yp, w.yn);
        double err = 0.0;
        for (int n = 0; n < ns; ++n) {
            double ytol = atol[n] + rtol[n] * fabs(w.yn[n]);
            double e = (D1 * w.k1[n] + D2 * w.k2[n] + D3 * w.k3[n]) / ytol;
            err = err + e * e;
        }
        double s = sqrt(err * rnspec), maxerr = UROUND;
        if (!(s <= maxerr)) maxerr = s;
        double dtfac = 0.9 / pow(maxerr, GROW);
        if (maxerr <= 1.0) {
            if (*nh >= hcap) { rc = -3; break; }
            h[(*nh)++] = dt;
            for (int q = 0; q < ns; ++q) y[q] = w.yn[q] > CONMIN ? w.yn[q] : CONMIN;
            tnow = tend;
            double hi = lpass ? FACMAX : FACONE;
            if (dtfac > hi) dtfac = hi;
            if (dtfac < FACMIN) dtfac = FACMIN;
            double dn = dtfac * dt;
            if (dn < DTMIN) dn = DTMIN;
            dt = dn < DTMAX ? dn : DTMAX;
            lpass = 1;
        } else {
            dt = FACMIN * dt;
            lpass = 0;
            if (dt < DTMIN) { rc = -1; break; }
        }
    }
    work_free(&w);
    return rc;
}

void om_replay_cell(const om_ctx *c, int ncsp, const double *rki, double *y, const double *h, int nh)
{
    om_work w;
    work_alloc(c, ncsp, &w);
    for (int n = 0; n < nh; ++n) {
        ros3_stages1(c, ncsp, rki, y, h[n], w.ydot, w.cc2, w.vdiag, w.k1, w.k2, w.k3, w.yp, w.yn);
        for (int q = 0; q < c->ns; ++q) y[q] = w.yn[q] > CONMIN ? w.yn[q] : CONMIN;
    }
    work_free(&w);
}

/* Solves (L U)^T u = v in place with om_decomp's factors: U^T z = v by rows
 * ascending, then L^T u = z by rows descending (scatter form). */
static void tsolve1(const om_ctx *c, int ncsp, const double *cc2, const double *vdiag, double *v)
{
    const int N = c->ischan;
    const int *kz = c->kzero[ncsp];
    int ltot = 0;
    for (int i = 1; i <= N; ++i) ltot += c->kzilch[ncsp][i];
    /* U rows are stored N..1 after the L block */
    int *ubeg = xcalloc((size_t)N + 2, sizeof(int)), *lbeg = xcalloc((size_t)N + 2, sizeof(int));
    int q = ltot + 1;
    for (int i = N; i >= 1; --i) { ubeg[i] = q; q += c->mzilch[ncsp][i]; }
    q = 1;
    for (int i = 1; i <= N; ++i) { lbeg[i] = q; q += c->kzilch[ncsp][i]; }
    for (int i = 1; i <= N; ++i) {
        v[i - 1] = v[i - 1] * vdiag[i - 1];
        for (int e = ubeg[i]; e < ubeg[i] + c->mzilch[ncsp][i]; ++e)
            v[kz[e] - 1] = v[kz[e] - 1] - cc2[e] * v[i - 1];
    }
    for (int i = N; i >= 1; --i)
        for (int e = lbeg[i]; e < lbeg[i] + c->kzilch[ncsp][i]; ++e)
            v[kz[e] - 1] = v[kz[e] - 1] - cc2[e] * v[i - 1];
    free(ubeg); free(lbeg);
}

/* (S^T w)_r: -w over reactant slots, +SC w over product slots (rbfeval.F:120-147) */
static double st_dot(const om_ctx *c, int r, const double *w)
{
    double s = 0.0;
    for (int q = 1; q <= c->t.nreact[r - 1]; ++q) s -= w[IRM2(c, r, q) - 1];
    for (int p = 1; p <= c->t.nprdct[r - 1]; ++p) s += SC(c, r, p) * w[IRM2(c, r, p + 3) - 1];
    return s;
}

/* product of reactant concentrations except slots a and b (1-based; 0 = none) */
static double pexcept(const om_ctx *c, int r, const double *y, int a, int b)
{
    double p = 1.0;
    for (int q = 1; q <= c->t.nreact[r - 1]; ++q)
        if (q != a && q != b) p *= y[IRM2(c, r, q) - 1];
    return p;
}

/* Reverse of one step: lam holds dL/dy_{n+1} on entry, dL/dy_n on exit;
 * gk[r-1] += dL/dk_r.  y is the step's initial state, dt its size. */
static void adj_step1(const om_ctx *c, int ncsp, const double *rki, const double *y, double dt,
                      double *lam, double *gk, om_work *w)
{
    const int ns = c->ns;
    ros3_stages1(c, ncsp, rki, y, dt, w->ydot, w->cc2, w->vdiag, w->k1, w->k2, w->k3, w->yp, w->yn);
    const double dtinv = 1.0 / dt;
    const double x21 = C21 * dtinv, x31 = C31 * dtinv, x32 = C32 * dtinv;
    for (int i = 0; i < ns; ++i) if (!(w->yn[i] > CONMIN)) lam[i] = 0.0;  /* clamp (rbsolver.F:473) */
    for (int i = 0; i < ns; ++i) w->u3[i] = B3 * lam[i];
    tsolve1(c, ncsp, w->cc2, w->vdiag, w->u3);
    for (int i = 0; i < ns; ++i) w->u2[i] = B2 * lam[i] - x32 * w->u3[i];
    tsolve1(c, ncsp, w->cc2, w->vdiag, w->u2);
    /* F2 = f(YP) with F2bar = -(u2 + u3): YPbar = J(YP)^T F2bar */
    for (int i = 0; i < ns; ++i) w->yb[i] = 0.0;
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int r = c->nkuserat[ncsp][nrx];
        w->s2[r] = st_dot(c, r, w->u2);
        w->s3[r] = st_dot(c, r, w->u3);
        double g = -rki[r - 1] * (w->s2[r] + w->s3[r]);
        for (int q = 1; q <= c->t.nreact[r - 1]; ++q)
            w->yb[IRM2(c, r, q) - 1] += g * pexcept(c, r, w->yp, q, 0);
    }
    for (int i = 0; i < ns; ++i) w->u1[i] = B1 * lam[i] - x31 * w->u3[i] - x21 * w->u2[i] + w->yb[i];
    tsolve1(c, ncsp, w->cc2, w->vdiag, w->u1);
    /* ybar = lam + YPbar + J(y)^T F1bar - sum_s H(u_s, K_s), F1bar = -u1 */
    for (int i = 0; i < ns; ++i) w->yb[i] += lam[i];
    const double *K[3] = {w->k1, w->k2, w->k3};
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int r = c->nkuserat[ncsp][nrx], nre = c->t.nreact[r - 1];
        double s1 = st_dot(c, r, w->u1), s2 = w->s2[r], s3 = w->s3[r];
        double sig[3] = {s1, s2, s3};
        double k = rki[r - 1];
        double w1 = -s1, w2 = -(s2 + s3);
        double dk = w1 * pexcept(c, r, y, 0, 0) + w2 * pexcept(c, r, w->yp, 0, 0);
        for (int q = 1; q <= nre; ++q) {
            int a = IRM2(c, r, q) - 1;
            double gy = pexcept(c, r, y, q, 0);
            double hs = 0.0;
            for (int s = 0; s < 3; ++s) {
                double hk = 0.0;   /* sum over the other reactant slots of K * remaining product */
                for (int q2 = 1; q2 <= nre; ++q2)
                    if (q2 != q) hk += K[s][IRM2(c, r, q2) - 1] * pexcept(c, r, y, q, q2);
                hs += sig[s] * hk;
                dk -= sig[s] * gy * K[s][a];
            }
            w->yb[a] += k * (w1 * gy - hs);
        }
        gk[r - 1] += dk;
    }
    for (int i = 0; i < ns; ++i) lam[i] = w->yb[i];
}

void om_adjoint_cell(const om_ctx *c, int ncsp, const double *rki, const double *y0,
                     const double *h, int nh, double *lam, double *gk)
{
    const int ns = c->ns;
    double *tape = xcalloc((size_t)(nh ? nh : 1) * ns, 8);
    om_work w;
    work_alloc(c, ncsp, &w);
    double *y = xcalloc(ns, 8);
    memcpy(y, y0, (size_t)ns * 8);
    for (int n = 0; n < nh; ++n) {
        memcpy(tape + (size_t)n * ns, y, (size_t)ns * 8);
        ros3_stages1(c, ncsp, rki, y, h[n], w.ydot, w.cc2, w.vdiag, w.k1, w.k2, w.k3, w.yp, w.yn);
        for (int q = 0; q < ns; ++q) y[q] = w.yn[q] > CONMIN ? w.yn[q] : CONMIN;
    }
    for (int n = nh - 1; n >= 0; --n) adj_step1(c, ncsp, rki, tape + (size_t)n * ns, h[n], lam, gk, &w);
    work_free(&w);
    free(y); free(tape);
}

/* ------------------------------------------------------------------ */
/* Tangent-linear (forward-mode) sensitivity of the same fixed-step map
 * (no reference counterpart; an exact alternative to DDM-3D's CN step,
 * ddm3d/DDM3D_CHEM.F:155-511).  Along a direction (dy0, dk), one step with
 * M = J(y) - I/(GAM h):
 *   dK1 = M^-1 (-S w1),  w1_r = q_r(y; dy) + dk_r p_r(y) + dq_r(y; K1)
 *   dYP = dy + dK1,      dF1 = S (q_r(yp; dYP) + dk_r p_r(yp))
 *   dK2 = M^-1 (-dF1 - x21 dK1 - S dq_r(y; K2))
 *   dK3 = M^-1 (-dF1 - x31 dK1 - x32 dK2 - S dq_r(y; K3))
 *   dy_new = [yn > CONMIN] (dy + B1 dK1 + B2 dK2 + B3 dK3)
 * with p_r the reactant product, q_r(y; v) = k_r grad p_r(y) . v its
 * directional derivative and dq_r(y; K) = d/de [k_r grad p_r(y) . K] along
 * (dy, dk): the Jacobian's change applied to a stage. */

/* q_r(y; v) / k_r */
static double om_qdir(const om_ctx *c, int r, const double *y, const double *v)
{
    double s = 0.0;
    for (int q = 1; q <= c->t.nreact[r - 1]; ++q) s += v[IRM2(c, r, q) - 1] * pexcept(c, r, y, q, 0);
    return s;
}

/* dq_r(y; K) along (dy, dk_r) */
static double om_dq(const om_ctx *c, int r, double k, double dkr, const double *y, const double *dy,
                    const double *K)
{
    const int nre = c->t.nreact[r - 1];
    double s = dkr * om_qdir(c, r, y, K);
    for (int q = 1; q <= nre; ++q)
        for (int q2 = 1; q2 <= nre; ++q2)
            if (q2 != q) s += k * K[IRM2(c, r, q) - 1] * dy[IRM2(c, r, q2) - 1] * pexcept(c, r, y, q, q2);
    return s;
}

/* out += sgn * S w (w per reaction, 1-based) */
static void om_smul(const om_ctx *c, int ncsp, const double *w, double sgn, double *out)
{
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int r = c->nkuserat[ncsp][nrx];
        for (int q = 1; q <= c->t.nreact[r - 1]; ++q) out[IRM2(c, r, q) - 1] -= sgn * w[r];
        for (int p = 1; p <= c->t.nprdct[r - 1]; ++p) out[IRM2(c, r, p + 3) - 1] += sgn * SC(c, r, p) * w[r];
    }
}

static void tlm_step1(const om_ctx *c, int ncsp, const double *rki, const double *y, double dt,
                      double *dy, const double *dk, om_work *w, double *wr, double *df1)
{
    const int ns = c->ns;
    ros3_stages1(c, ncsp, rki, y, dt, w->ydot, w->cc2, w->vdiag, w->k1, w->k2, w->k3, w->yp, w->yn);
    const double dtinv = 1.0 / dt;
    const double x21 = C21 * dtinv, x31 = C31 * dtinv, x32 = C32 * dtinv;
    /* stage 1 */
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int r = c->nkuserat[ncsp][nrx];
        double k = rki[r - 1], dkr = dk ? dk[r - 1] : 0.0;
        wr[r] = k * om_qdir(c, r, y, dy) + dkr * pexcept(c, r, y, 0, 0) + om_dq(c, r, k, dkr, y, dy, w->k1);
    }
    for (int i = 0; i < ns; ++i) w->u1[i] = 0.0;
    om_smul(c, ncsp, wr, -1.0, w->u1);
    om_solve(c, ncsp, 1, w->cc2, w->vdiag, w->u1);
    /* dF1 at the stage point */
    for (int i = 0; i < ns; ++i) w->yb[i] = dy[i] + w->u1[i];
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int r = c->nkuserat[ncsp][nrx];
        double k = rki[r - 1], dkr = dk ? dk[r - 1] : 0.0;
        wr[r] = k * om_qdir(c, r, w->yp, w->yb) + dkr * pexcept(c, r, w->yp, 0, 0);
    }
    for (int i = 0; i < ns; ++i) df1[i] = 0.0;
    om_smul(c, ncsp, wr, 1.0, df1);
    /* stages 2 and 3 */
    const double *K[2] = {w->k2, w->k3};
    double *U[2] = {w->u2, w->u3};
    for (int s = 0; s < 2; ++s) {
        for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
            int r = c->nkuserat[ncsp][nrx];
            wr[r] = om_dq(c, r, rki[r - 1], dk ? dk[r - 1] : 0.0, y, dy, K[s]);
        }
        for (int i = 0; i < ns; ++i)
            U[s][i] = s == 0 ? -df1[i] - x21 * w->u1[i] : -df1[i] - x31 * w->u1[i] - x32 * w->u2[i];
        om_smul(c, ncsp, wr, -1.0, U[s]);
        om_solve(c, ncsp, 1, w->cc2, w->vdiag, U[s]);
    }
    for (int i = 0; i < ns; ++i)
        dy[i] = w->yn[i] > CONMIN ? dy[i] + B1 * w->u1[i] + B2 * w->u2[i] + B3 * w->u3[i] : 0.0;
}

void om_tlm_cell(const om_ctx *c, int ncsp, const double *rki, const double *y0, const double *h, int nh,
                 double *dy, const double *dk)
{
    const int ns = c->ns;
    om_work w;
    work_alloc(c, ncsp, &w);
    double *y = xcalloc(ns, 8), *wr = xcalloc((size_t)c->nr + 1, 8), *df1 = xcalloc(ns, 8);
    memcpy(y, y0, (size_t)ns * 8);
    for (int n = 0; n < nh; ++n) {
        tlm_step1(c, ncsp, rki, y, h[n], dy, dk, &w, wr, df1);
        for (int q = 0; q < ns; ++q) y[q] = w.yn[q] > CONMIN ? w.yn[q] : CONMIN;
    }
    work_free(&w);
    free(y); free(wr); free(df1);
}

int om_adjoint_run(const om_ctx *c, int64_t ncells, double chemstep, const double *rki,
                   const double *y0, const double *rtol, const double *atol,
                   double *lam, double *gk, double *yend, int32_t *nacc)
{
    const int ns = c->ns, nr = c->nr, hcap = 1 << 20;
    double *h = xcalloc(hcap, 8), *y = xcalloc(ns, 8), *ys = xcalloc(ns, 8), *l = xcalloc(ns, 8);
    double *rk = xcalloc(nr, 8), *g = xcalloc(nr, 8), *rt = xcalloc(ns, 8), *at = xcalloc(ns, 8);
    for (int i = 1; i <= ns; ++i) { rt[i - 1] = rtol[c->inew2old[i] - 1]; at[i - 1] = atol[c->inew2old[i] - 1]; }
    int rc = 0;
    for (int64_t cc = 0; cc < ncells; ++cc) {
        for (int r = 0; r < nr; ++r) { rk[r] = rki[(size_t)r * ncells + cc]; g[r] = 0.0; }
        for (int i = 1; i <= ns; ++i) {
            size_t o = (size_t)(c->inew2old[i] - 1) * ncells + cc;
            y[i - 1] = ys[i - 1] = y0[o];
            l[i - 1] = lam[o];
        }
        int nh = 0;
        int brc = om_steps_cell(c, 1, chemstep, rk, y, rt, at, h, hcap, &nh);
        if (nacc) nacc[cc] = brc ? -1 : nh;
        if (brc) { if (!rc) rc = brc; continue; }
        om_adjoint_cell(c, 1, rk, ys, h, nh, l, g);
        for (int i = 1; i <= ns; ++i) {
            size_t o = (size_t)(c->inew2old[i] - 1) * ncells + cc;
            lam[o] = l[i - 1];
            if (yend) yend[o] = y[i - 1];
        }
        if (gk) for (int r = 0; r < nr; ++r) gk[(size_t)r * ncells + cc] += g[r];
    }
    free(h); free(y); free(ys); free(l); free(rk); free(g); free(rt); free(at);
    return rc;
}

/* ------------------------------------------------------------------ */
/* CMAQ-DDM-3D first-order chemistry sensitivities                     */
/* ------------------------------------------------------------------ */
/* Restates the DDM-3D chemistry step as the ros3 build with -Dsens runs it:
 * the forward step accumulates the time-averaged concentration
 * YAVE = sum over accepted steps of (Y + max(YNEW, CONMIN)) / 2 * DT, divided
 * by CHEMSTEP (rbsolver.F:341-343, 468-472, 592-596); then, per parameter,
 * one Crank-Nicolson step over the whole sync step with the Jacobian at YAVE
 * (ddm3d/DDM3D_CHEM.F:296-395): A = I - h/2 J, A1 = I + h/2 J in REAL(4),
 * A S_new = A1 S_old + P_rate, LINPACK SGEFA/SGESL (ddm3d/s_lu.F).  Rate
 * parameters (IPT = 5) add P_rate = h k p(YAVE) S_r (DDM3D_CHEM.F:340-384),
 * i.e. sensitivities to ln k; INIT parameters are seeded with S0 = C0
 * (init/initscen.F:306-314), i.e. sensitivities to ln C0. */

/* LINPACK SGEFA: Gaussian elimination with partial pivoting, column major. */
static int sgefa(float *a, int n, int *ipvt)
{
    int info = 0;
    for (int k = 0; k < n - 1; ++k) {
        int l = k;
        float m = fabsf(a[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabsf(a[k * n + i]) > m) { m = fabsf(a[k * n + i]); l = i; }
        ipvt[k] = l;
        if (a[k * n + l] == 0.0f) { info = k + 1; continue; }
        if (l != k) { float t = a[k * n + l]; a[k * n + l] = a[k * n + k]; a[k * n + k] = t; }
        float t = -1.0f / a[k * n + k];
        for (int i = k + 1; i < n; ++i) a[k * n + i] *= t;
        for (int j = k + 1; j < n; ++j) {
            float tj = a[j * n + l];
            if (l != k) { a[j * n + l] = a[j * n + k]; a[j * n + k] = tj; }
            for (int i = k + 1; i < n; ++i) a[j * n + i] += tj * a[k * n + i];
        }
    }
    ipvt[n - 1] = n - 1;
    if (a[(n - 1) * n + n - 1] == 0.0f) info = n;
    return info;
}

/* LINPACK SGESL with job = 0: solves A x = b in place. */
static void sgesl(const float *a, int n, const int *ipvt, float *b)
{
    for (int k = 0; k < n - 1; ++k) {
        int l = ipvt[k];
        float t = b[l];
        if (l != k) { b[l] = b[k]; b[k] = t; }
        for (int i = k + 1; i < n; ++i) b[i] += t * a[k * n + i];
    }
    for (int k = n - 1; k >= 0; --k) {
        b[k] /= a[k * n + k];
        float t = -b[k];
        for (int i = 0; i < k; ++i) b[i] += t * a[k * n + i];
    }
}

/* The Crank-Nicolson part alone, from a given YAVE (NEW order). */
void om_ddm_step(const om_ctx *c, double chemstep, const double *rki, const double *yave, int np,
                 const int32_t *prate, float *sens)
{
    const int ns = c->ns, ncsp = 1;
    /* Jacobian at YAVE accumulated in REAL(4) (MECHANISM_FUNCTIONS.F:831-869) */
    float *jac = xcalloc((size_t)ns * ns, 4), *a = xcalloc((size_t)ns * ns, 4), *a1 = xcalloc((size_t)ns * ns, 4);
    for (int nrx = 1; nrx <= c->nuserat[ncsp]; ++nrx) {
        int r = c->nkuserat[ncsp][nrx], nre = c->t.nreact[r - 1];
        for (int q = 1; q <= nre; ++q) {
            int col = IRM2(c, r, q) - 1;
            double ex = rki[r - 1] * pexcept(c, r, yave, q, 0);
            for (int q2 = 1; q2 <= nre; ++q2) {
                int row = IRM2(c, r, q2) - 1;
                jac[col * ns + row] += (float)(-1.0 * ex);
            }
            for (int p = 1; p <= c->t.nprdct[r - 1]; ++p) {
                int row = IRM2(c, r, p + 3) - 1;
                jac[col * ns + row] += (float)(SC(c, r, p) * ex);
            }
        }
    }
    for (int j = 0; j < ns; ++j)
        for (int i = 0; i < ns; ++i) {
            a[j * ns + i] = (float)(-0.5 * chemstep * (double)jac[j * ns + i]);
            a1[j * ns + i] = (float)(0.5 * chemstep * (double)jac[j * ns + i]);
            if (i == j) { a[j * ns + i] = 1.0f + a[j * ns + i]; a1[j * ns + i] = 1.0f + a1[j * ns + i]; }
        }
    int *ipvt = xcalloc(ns, sizeof(int));
    sgefa(a, ns, ipvt);
    float *sold = xcalloc(ns, 4), *prd = xcalloc(ns, 4), *prate_v = xcalloc(ns, 4);
    for (int p = 0; p < np; ++p) {
        float *sp = sens + (size_t)p * ns;
        for (int q = 0; q < ns; ++q) {
            sold[q] = fabsf(sp[q]) < 1.0e-25f ? 0.0f : sp[q];
            prate_v[q] = 0.0f;
        }
        int r = prate ? prate[p] : 0;
        if (r > 0) {
            float kstep = (float)(rki[r - 1] * chemstep);
            float rxm = (float)((double)kstep * pexcept(c, r, yave, 0, 0));
            for (int q = 1; q <= c->t.nreact[r - 1]; ++q) {
                int s = IRM2(c, r, q) - 1;
                prate_v[s] = prate_v[s] - rxm * 1.0f;
            }
            for (int q = 1; q <= c->t.nprdct[r - 1]; ++q) {
                int s = IRM2(c, r, q + 3) - 1;
                prate_v[s] = (float)((double)prate_v[s] + (double)rxm * SC(c, r, q) * 1.0);
            }
        }
        for (int s = 0; s < ns; ++s) {
            double total = 0.0;
            for (int j = 0; j < ns; ++j) total = total + (double)(a1[j * ns + s] * sold[j]);
            prd[s] = (float)(total + (double)prate_v[s]);
        }
        sgesl(a, ns, ipvt, prd);
        for (int s = 0; s < ns; ++s) sp[s] = fabsf(prd[s]) < 1.0e-25f ? 0.0f : prd[s];
    }
    free(jac); free(a); free(a1); free(ipvt); free(sold); free(prd); free(prate_v);
}

int om_ddm_cell(const om_ctx *c, double chemstep, const double *rki, double *y,
                const double *rtol, const double *atol, int np, const int32_t *prate, float *sens)
{
    const int ns = c->ns, ncsp = 1;
    const double rnspec = (double)(1.0f / (float)ns);
    om_work w;
    work_alloc(c, ncsp, &w);
    double *yave = xcalloc(ns, 8);
    double dt = DTMIN > DTSTART ? DTMIN : DTSTART, tnow = 0.0;
    int lpass = 0, rc = 0;
    long attempts = 0;
    while (tnow < chemstep) {
        if (++attempts > OM_MAX_ATTEMPTS) { rc = -2; break; }
        double tend = tnow + dt;
        if (tend > chemstep) { dt = chemstep - tnow; tend = chemstep; }
        ros3_stages1(c, ncsp, rki, y, dt, w.ydot, w.cc2, w.vdiag, w.k1, w.k2, w.k3, w.yp, w.yn);
        double err = 0.0;
        for (int n = 0; n < ns; ++n) {
            double ytol = atol[n] + rtol[n] * fabs(w.yn[n]);
            double e = (D1 * w.k1[n] + D2 * w.k2[n] + D3 * w.k3[n]) / ytol;
            err = err + e * e;
        }
        double s = sqrt(err * rnspec), maxerr = UROUND;
        if (!(s <= maxerr)) maxerr = s;
        double dtfac = 0.9 / pow(maxerr, GROW);
        if (maxerr <= 1.0) {
            for (int q = 0; q < ns; ++q) {
                double yn = w.yn[q] > CONMIN ? w.yn[q] : CONMIN;
                yave[q] = yave[q] + 0.5 * (y[q] + yn) * dt;
                y[q] = yn;
            }
            tnow = tend;
            double hi = lpass ? FACMAX : FACONE;
            if (dtfac > hi) dtfac = hi;
            if (dtfac < FACMIN) dtfac = FACMIN;
            double dn = dtfac * dt;
            if (dn < DTMIN) dn = DTMIN;
            dt = dn < DTMAX ? dn : DTMAX;
            lpass = 1;
        } else {
            dt = FACMIN * dt;
            lpass = 0;
            if (dt < DTMIN) { rc = -1; break; }
        }
    }
    work_free(&w);
    if (rc) { free(yave); return rc; }
    for (int q = 0; q < ns; ++q) yave[q] = yave[q] / chemstep;
    om_ddm_step(c, chemstep, rki, yave, np, prate, sens);
    free(yave);
    return 0;
}


/* ---------------------------------------------------------------------------
 * CHEM's CGRID gather and scatter (rbdriver.F:460-509, 577-595).  Every
 * operation is REAL(4) as in the driver, REAL(8) where the driver widens.
 * ------------------------------------------------------------------------- */
static const float OM_CONCMIN = 1.0e-30f;                 /* rbdriver.F:110 */
static const float OM_CONCOFM = 1.0e+06f;                 /* rbdriver.F:111 */
static const float OM_STDATMPA = 101325.0f;               /* CONST.EXT:95   */
static const float OM_MWAIR = 28.9628f, OM_MWWAT = 18.0153f;   /* CONST.EXT:113, 119 */

static float om_forward_conv(double molwt)               /* rbinit.F:260   */
{
    float a = 1.0e-3f * OM_MWAIR;
    return a / (float)molwt;
}

void om_gather(int ns, int nphotab, int nhetero, int ncols, int nrows, int nlays,
               const int32_t *cgrid_index, const int32_t *convert_conc,
               const double *species_molwt, const float *cgrid, const float *ta,
               const float *dens, const float *pres, const float *qv,
               const float *seawater_zone, const float *seaice, const float *rj,
               const double *khetero, int ld_het, double *temp, double *pres_atm,
               double *h2o, double *sea, double *rjd, double *het, double *y)
{
    const float pa2atm = 1.0f / OM_STDATMPA;             /* rbdriver.F:112 */
    const float maomv = OM_MWAIR / OM_MWWAT;              /* rbdriver.F:113 */
    const float qv_to_ppm = OM_CONCOFM * maomv;           /* rbdriver.F:114 */
    const size_t n = (size_t)ncols * nrows * nlays;
    float *svol = (float *)malloc(n * sizeof(float));
    for (int lev = 0; lev < nlays; ++lev)
        for (int row = 0; row < nrows; ++row)
            for (int col = 0; col < ncols; ++col) {
                size_t c = (size_t)col + (size_t)ncols * (row + (size_t)nrows * lev);
                size_t c2 = (size_t)col + (size_t)ncols * row;
                float zone = seawater_zone ? seawater_zone[c2] : 0.0f;
                float ice = seaice ? seaice[c2] : 0.0f;
                float s, h;
                temp[c] = (double)ta[c];                                 /* :465 */
                svol[c] = 1.0f / dens[c];                                /* :467 */
                pres_atm[c] = (double)(pa2atm * pres[c]);                /* :468 */
                h = qv_to_ppm * qv[c];                                   /* :469 */
                h2o[c] = (double)(h > 0.0f ? h : 0.0f);
                s = zone - ice;                                          /* :470 */
                sea[c] = (double)(0.0f > s ? 0.0f : s);
            }
    for (int isp = 0; isp < ns; ++isp) {                                 /* :476-492 */
        const size_t spc = (size_t)(cgrid_index[isp] - 1);
        const float fwd = om_forward_conv(species_molwt[isp]);
        for (size_t c = 0; c < n; ++c) {
            float v = cgrid[spc * n + c];
            if (convert_conc[isp]) v = fwd * svol[c] * v;
            y[(size_t)isp * n + c] = (double)(v > OM_CONCMIN ? v : OM_CONCMIN);
        }
    }
    for (int k = 0; k < nphotab; ++k)                                     /* :507-508 */
        for (size_t c = 0; c < n; ++c) rjd[(size_t)k * n + c] = (double)rj[(size_t)k * n + c];
    if (het)
        for (int k = 0; k < nhetero; ++k)                                 /* :504-505 */
            for (size_t c = 0; c < n; ++c)
                het[(size_t)k * n + c] = khetero ? khetero[c * (size_t)ld_het + k] : 0.0;
    free(svol);
}

void om_scatter(int ns, int ncols, int nrows, int nlays, const int32_t *cgrid_index,
                const int32_t *convert_conc, const double *species_molwt, const float *dens,
                const double *y, float *cgrid)
{
    const size_t n = (size_t)ncols * nrows * nlays;
    for (int isp = 0; isp < ns; ++isp) {                                 /* :580-595 */
        const size_t spc = (size_t)(cgrid_index[isp] - 1);
        const double rev = (double)(1.0f / om_forward_conv(species_molwt[isp]));   /* rbinit.F:261 */
        for (size_t c = 0; c < n; ++c) {
            const double yc = y[(size_t)isp * n + c];
            cgrid[spc * n + c] = convert_conc[isp] ? (float)(rev * (double)dens[c] * yc) : (float)yc;
        }
    }
}
