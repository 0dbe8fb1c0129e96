// SMVGEAR (CCTM/src/gas/smvgear) on the GPU: the launch record of the
// persistent Gear kernel and the Gear coefficient tables.
//
// The symbolic analysis of SMVGEAR (grsprse.F, JSPARSE) is the algorithm of
// ros3's RBSPARSE, so the Gear kernels share the mechanism's species order,
// sparse pattern and factorization schedule with the ROS3 kernels; only the
// matrix (P = I - ASET(q,1) h J, grpderiv.F) and the integrator differ.
#pragma once

#include <stdint.h>

#include "ros3_pow.h"

namespace ros3 {

// Constants of the Gear solver (grinit.F:184-191, grsmvgear.F:91-103,
// grsprse.F:675-678).
constexpr int kGearMaxOrd = 5;            // MAXORD
constexpr int kGearMstep = 3;             // corrector iterations per step
constexpr int kGearMbetween = 50;         // steps between Jacobian updates
constexpr int kGearMxLowRed = 20;         // times YLOWA may be lowered
constexpr int kGearCols = 7;             // Nordsieck columns held (MXORDER, GRVARS.F:67)
constexpr double kGearHmin = 1.0e-10;
constexpr double kGearHmaxDay = 15.0;
constexpr double kGearHmaxNight = 5.0;
constexpr double kGearFracdec = 0.25;
constexpr double kGearConcmin = 1.0e-30;
constexpr double kGearZbound = 1.0e-30;
constexpr double kGearOmic = 0.000001;
constexpr double kGearOpt2mic = 0.0000012;
constexpr double kGearOpt3mic = 0.0000013;
constexpr double kGearOpt4mic = 0.0000014;
// passes a lane's PDERIV + DECOMP may wait for more lanes of its wavefront
constexpr int kGearPdWait = 2;

// ASET, PERTST (squared), ENQQ1..3, CONPST, CONP15 as JSPARSE leaves them
// (grsprse.F:675-713; PERTST's DATA statement GRVARS.F:251-254).  ENQQ1..3
// are REAL(4) quotients stored in REAL(8).
struct GearCoef {
    double aset[7][8] = {};
    double pertst[8][4] = {};
    double enqq1[8] = {}, enqq2[8] = {}, enqq3[8] = {}, conpst[8] = {}, conp15[8] = {};
    constexpr GearCoef()
    {
        const double p0[3][7] = {{2.0, 4.5, 7.333, 10.42, 13.7, 17.15, 1.0},
                                 {3.0, 6.0, 9.167, 12.5, 15.98, 1.0, 1.0},
                                 {1.0, 1.0, 0.5, 0.1667, 0.04133, 0.008267, 1.0}};
        for (int q = 1; q <= 7; ++q) {
            enqq1[q] = (double)(0.5f / (float)q);
            enqq2[q] = (double)(0.5f / (float)(q + 1));
            enqq3[q] = (double)(0.5f / (float)(q + 2));
            conpst[q] = 1.0 / (p0[0][q - 1] * enqq3[q]);
            conp15[q] = 1.5 * conpst[q];
            for (int k = 1; k <= 3; ++k) pertst[q][k] = p0[k - 1][q - 1] * p0[k - 1][q - 1];
        }
        for (int i = 1; i <= 6; ++i) aset[i][2] = 1.0;
        aset[1][1] = 1.0;
        aset[2][1] = 2.0 / 3.0;     aset[2][3] = 1.0 / 3.0;
        aset[3][1] = 6.0 / 11.0;    aset[3][3] = 6.0 / 11.0;    aset[3][4] = 1.0 / 11.0;
        aset[4][1] = 12.0 / 25.0;   aset[4][3] = 0.70;          aset[4][4] = 0.20;   aset[4][5] = 0.02;
        aset[5][1] = 60.0 / 137.0;  aset[5][3] = 225.0 / 274.0; aset[5][4] = 85.0 / 274.0;
        aset[5][5] = 15.0 / 274.0;  aset[5][6] = 1.0 / 274.0;
        aset[6][1] = 180.0 / 441.0; aset[6][3] = 406.0 / 441.0; aset[6][4] = 735.0 / 1764.0;
        aset[6][5] = 175.0 / 1764.0; aset[6][6] = 21.0 / 1764.0; aset[6][7] = 1.0 / 1764.0;
    }
};

// A cell that gear_kernel hands to gear_tail_kernel: the lane's whole control
// state at the top of LOOP_ORDER (its workspace rows go to tail_ws).  The
// X-macro lists name every variable that lives across kernel-loop passes.
#define GEAR_LANE_INTS(X) X(pc) X(nqq) X(nqqold) X(kstep) X(idoub) X(ifsuccess) X(jfail) X(jrestar) X(jeval) \
    X(nslp) X(nsteps) X(nylowred) X(l3) X(nfails) X(attempts) X(fcode)
#define GEAR_LANE_BOOLS(X) X(init_phase) X(pd_pending) X(higher_order)
#define GEAR_LANE_DBLS(X) X(delt) X(xelaps) X(told) X(rdelt) X(rdelmax) X(hratio) X(asn1) X(enqq) X(eup) \
    X(edwn) X(drate) X(rmserr) X(der2max) X(rdeltup) X(r1delt) X(ylowa) X(hmax) X(h0fac)
struct GearLaneState {
    int64_t cell;
#define GEAR_DECL_INT(v) int32_t v;
#define GEAR_DECL_DBL(v) double v;
    GEAR_LANE_INTS(GEAR_DECL_INT)
    GEAR_LANE_BOOLS(GEAR_DECL_INT)
    int32_t pad;
    GEAR_LANE_DBLS(GEAR_DECL_DBL)
#undef GEAR_DECL_INT
#undef GEAR_DECL_DBL
};

struct GearParams {
    int64_t ncells;            // cells 0 .. ncells - 1 of the arrays
    int64_t stride;            // leading dimension of the cell arrays
    double chemstep;           // minutes
    const double *rki;         // [nr][stride]
    const uint8_t *sunlit;     // [stride] LSUNLIGHT per cell, or nullptr (all sunlit)
    double *y;                 // [ns][stride] mechanism order, in/out (ppm)
    int32_t *nsteps;           // [stride] successful steps (-1: failed cell), or nullptr
    int32_t *nfail;            // [stride] error-test + convergence failures, or nullptr
    double *ws;                // workspace: wave w's rows start at ws + w * ws_stride
    int64_t ws_stride;
    unsigned long long *queue; // next unassigned cell
    int32_t *fail_flag;
    int64_t *fail_cell;
    double *fail_info;         // DELT, XELAPS, YLOWA of the first failing cell
    int32_t *fail_code;
    double eps;                // ERRMAX: GEAR_RTOL (REAL(4)) in REAL(8)
    double atol_level[7];      // OPTIMAL_ATOL_PPM levels 1..6 ([0] unused)
    double conc_floor;         // GEAR_CONC_FLOOR in REAL(8)
    int32_t max_steps;         // attempted steps per cell before code -5
    int32_t refill_min;
    // long cells: after tail_after attempted steps (0 = never) a cell moves to
    // gear_tail_kernel, one cell per wavefront with its workspace in LDS;
    // at most tail_cap cells per call (the others stay where they are)
    int32_t tail_after;
    int32_t tail_cap;
    GearLaneState *tail_state;   // [tail_cap]
    double *tail_ws;             // [tail_cap][NWS]
    unsigned int *tail_count;    // cells handed over (may exceed tail_cap)
    unsigned int *tail_queue;    // next handed-over cell for gear_tail_kernel
};

}  // namespace ros3
