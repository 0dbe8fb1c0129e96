/*
 * Internal layout of the oracle's mechanism context, shared by the ROS3
 * restatement (ros3_oracle.c) and the SMVGEAR restatement (gear_oracle.c).
 * TEST INFRASTRUCTURE ONLY (see ros3_oracle.h).
 */
#ifndef CMAQ_OM_INTERNAL_H
#define CMAQ_OM_INTERNAL_H

#include "ros3_oracle.h"

#define MXRCT 3

struct om_ctx {
    om_tables t;
    int ns, nr, mxprd, ischan;
    int *inew2old, *iold2new;          /* [ns+1] */
    int *irm2;                         /* [(nr+1)*(mxprd+4)] */
    int nuserat[3];
    int *nkuserat[3];                  /* [nr+1] */
    int iarray[3];
    int *jarraypt[3];                  /* [(ns+1)*(ns+1)], (row, col) */
    int *kzero[3];                     /* [iarray+1] column of slot */
    int *kzilch[3], *mzilch[3];        /* [ns+1] */
    /* decomposition: loop-1 ops of column j are ops[opbeg[j] .. opbeg[j+1]) */
    int nops[3];
    int *opbeg[3], *op_ij[3], *op_ik[3], *op_kj[3];
    /* loop-2 (L scaling): slots lsc[lbeg[j] .. lbeg[j+1]) */
    int *lbeg[3], *lsc[3];
    /* PD lists */
    int mxrr, mxrp;
    int *nderivl[3], *nderivp[3];
    int *jarrl[3], *jlial[3], *jarrp[3], *jpial[3], *icoeff[3];
    /* rate evaluation: falloff row and het table per reaction */
    int *fall_idx;                     /* [nr+1] 1-based RFDAT column or 0 */
    int *het_idx;                      /* [nr+1] */
};

#define IRR(c, r, s)  ((c)->t.irr[((r) - 1) * ((c)->mxprd + 3) + (s) - 1])
#define SC(c, r, p)   ((c)->t.sc[((r) - 1) * (c)->mxprd + (p) - 1])
#define RTDAT(c, k, r) ((c)->t.rtdat[((r) - 1) * 3 + (k) - 1])
#define RFDAT(c, k, f) ((c)->t.rfdat[((f) - 1) * 5 + (k) - 1])
#define IRM2(c, r, s) ((c)->irm2[(r) * ((c)->mxprd + 4) + (s)])
#define JAP(c, n, i, j) ((c)->jarraypt[n][(i) * ((c)->ns + 1) + (j)])

#endif
