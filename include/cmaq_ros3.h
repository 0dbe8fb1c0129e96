/*
 * cmaq_ros3.h -- C ABI of the MI355X (gfx950) ROS3 gas-chemistry solver.
 *
 * Drop-in for the per-block loop of CMAQ's ros3 CHEM driver
 * (CCTM/src/gas/ros3/rbdriver.F:440-677), i.e. for
 *     CALL CALC_RCONST(...)   (RXNS_FUNC_MODULE.F90:204, rbdriver.F:513)
 *     CALL RBSOLVER(...)      (rbsolver.F:26, rbdriver.F:570)
 * over every cell of a process's sub-domain in one call.  Symbolic set-up
 * (RBINIT rbinit.F:26 + RBSPARSE rbsparse.F:26) happens in cmaq_ros3_init.
 *
 * Units and layouts follow the reference exactly:
 *   concentrations ppm, time minutes, J-values and heterogeneous rates 1/min,
 *   pressure atm, water vapour ppm, temperature K;
 *   every per-cell array is "cell fastest" (Fortran (cell, index) order):
 *   a[k * ncells + cell];  species in mechanism (CHEMISTRY_SPC) order.
 * Step control is per cell: the result for a cell equals RBSOLVER run on a
 * block holding that cell alone (NUMCELLS = 1), which decides day/night per
 * cell (any J > 0) instead of per 50-cell block.
 *
 * All functions return 0 on success or a negative CMAQ_ROS3_E_* code; they
 * never exit the process (the reference calls M3EXIT, rbsolver.F:583).
 */
#ifndef CMAQ_ROS3_H
#define CMAQ_ROS3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CMAQ_ROS3_OK              0
#define CMAQ_ROS3_E_ARG          -1  /* bad argument / size                         */
#define CMAQ_ROS3_E_MECH         -2  /* unsupported table feature or no kernel built */
#define CMAQ_ROS3_E_HIP          -3  /* HIP runtime error                           */
#define CMAQ_ROS3_E_CONVERGENCE  -4  /* DT fell below DTMIN (rbsolver.F:528-583)     */
#define CMAQ_ROS3_E_STEPS        -5  /* attempted-step limit per cell exceeded       */
#define CMAQ_ROS3_E_NOMEM        -6  /* device allocation failed                    */
#define CMAQ_ROS3_E_TAPE         -7  /* adjoint replay differs from the forward run  */

/*
 * Mechanism tables exactly as MODULE RXNS_DATA holds them
 * (MECHS/<mech>/RXNS_DATA_MODULE.F90).  Pointers refer to Fortran arrays
 * passed with C_LOC; 2-D Fortran arrays are passed as they lie in memory
 * (column-major), so IRR(NRXNS, MXPRD+3) arrives as irr[slot*nrxns + r].
 * Ids inside the arrays stay 1-based.
 */
typedef struct cmaq_ros3_mech {
    int32_t numb_mech_spc;      /* NUMB_MECH_SPC                                  */
    int32_t nrxns;              /* NRXNS                                          */
    int32_t mxprd;              /* MXPRD                                          */
    int32_t kunits;             /* KUNITS                                         */
    const int32_t *irr;         /* IRR(NRXNS, MXPRD+3)                            */
    const double  *sc;          /* SC(NRXNS, MXPRD)                               */
    const int32_t *nreact;      /* NREACT(NRXNS)                                  */
    const int32_t *nprdct;      /* NPRDCT(NRXNS)                                  */
    const int32_t *ktype;       /* KTYPE(NRXNS)                                   */
    const int32_t *irxbits;     /* IRXBITS(NRXNS)                                 */
    const int32_t *iorder;      /* IORDER(NRXNS)                                  */
    const double  *rtdat;       /* RTDAT(3, NRXNS)                                */
    int32_t nfalloff;           /* NFALLOFF                                       */
    const int32_t *irrfall;     /* IRRFALL(NFALLOFF)                              */
    const double  *rfdat;       /* RFDAT(5, NFALLOFF)                             */
    int32_t nmphot;             /* NMPHOT                                         */
    int32_t nphotab;            /* NPHOTAB                                        */
    const int32_t *iph;         /* IPH(NMPHOT, 3)                                 */
    int32_t mhetero;            /* MHETERO                                        */
    int32_t nhetero;            /* NHETERO                                        */
    const int32_t *ihetero;     /* IHETERO(MHETERO, 2)                            */
    int32_t nwm, nww, nwo2, nwn2, nwch4, nwh2;
    const int32_t *nrxwm, *nrxww, *nrxwo2, *nrxwn2, *nrxwch4, *nrxwh2;
    double atm_air, atm_o2, atm_n2, atm_ch4, atm_h2;
    int32_t nspecial_rxn;       /* must be 0 (SPECIAL_RATES not supported)        */
} cmaq_ros3_mech;

typedef struct cmaq_ros3_opts {
    int32_t device;             /* HIP device ordinal; -1 = current device        */
    int32_t max_attempts;       /* per cell and call; 0 = 1000000                 */
    int32_t waves_per_cu;       /* cap on the solver grid: at most waves_per_cu x
                                   CUs waves, in workgroups of 4; 0 = as many as
                                   the kernel's occupancy allows                  */
    int32_t refill_min;         /* lanes that must be idle before a wave refills;
                                   0 = autotuned                                  */
} cmaq_ros3_opts;

typedef struct cmaq_ros3_fail {
    int32_t code;               /* CMAQ_ROS3_E_CONVERGENCE / _STEPS, or 0          */
    int32_t pad;
    int64_t cell;               /* failing cell (0-based), -1 if none              */
    double maxerr;              /* MAXERR of the last rejected step                */
    double dt;                  /* DT after the last reduction, minutes            */
    double tnow;                /* elapsed minutes in the sync step                */
} cmaq_ros3_fail;

/*
 * The first failure of the last call in full: what RBSOLVER writes to LOGDEV
 * before M3EXIT (rbsolver.F:532-583).  temp_k .. seawater are the failing
 * cell's BLKTEMP, BLKPRES, BLKCH2O, BLKSEAWATER and dens its DENS (kg/m3,
 * BLKDENS); each is NaN where the call did not take that input (dens is set
 * by the CHEM-level calls only; the met fields are not known to
 * cmaq_ros3_solve_device and the adjoint calls).
 */
typedef struct cmaq_ros3_fail_detail {
    int32_t code;               /* as cmaq_ros3_fail; 0 = the last call had no failure */
    int32_t numb_mech_spc;      /* lengths of the arrays cmaq_ros3_fail_detail fills */
    int32_t nrxns;
    int32_t nphotab;
    int64_t cell;               /* failing cell (0-based), -1 if none              */
    double maxerr;              /* MAXERR of the last step                         */
    double err;                 /* ERR(cell): sum of the squared scaled errors     */
    double dt, tnow;            /* minutes                                         */
    double temp_k, pres_atm, h2o_ppm, seawater, dens;
} cmaq_ros3_fail_detail;

typedef struct cmaq_ros3_ctx cmaq_ros3_ctx;

/*
 * RBINIT + RBSPARSE replacement.  Validates the tables, derives the solver
 * species order and sparse pattern, and selects the generated gfx950 kernels
 * built for this mechanism (matched by a hash of the structural tables).
 * rtol/atol: per species, mechanism order (rbinit.F:257-262 sets them all to
 * RB_RTOL / RB_ATOL); NULL selects the defaults 1e-3 and 1e-7 ppm
 * (util/util/RUNTIME_VARS.F:177-178).  opts may be NULL.
 */
int cmaq_ros3_init(cmaq_ros3_ctx **out, const cmaq_ros3_mech *mech,
                   const double *rtol, const double *atol,
                   const cmaq_ros3_opts *opts);

/*
 * CALC_RCONST + RBSOLVER over ncells independent cells for one chemistry
 * step of chemstep_min minutes (rbdriver.F:385 CHEMSTEP).  Host pointers;
 * synchronous.  rj [nphotab][ncells], het [nhetero][ncells] (may be NULL =
 * 0), seawater [ncells] (may be NULL = 0), y [numb_mech_spc][ncells] in/out.
 * n_accepted / n_rejected (may be NULL) receive the per-cell step counts that
 * the reference's rbstats build reports (rbsolver.F:462-464, 521-524).
 * Where the reference would stop the run (DT below DTMIN, rbsolver.F:528-583)
 * this call integrates every other cell, sets n_accepted = -1 for each
 * failing cell, leaves that cell's y at its last accepted state and returns
 * CMAQ_ROS3_E_CONVERGENCE with the first failure in *fail.
 */
int cmaq_ros3_run(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                  const double *temp_k, const double *pres_atm,
                  const double *h2o_ppm, const double *rj_per_min,
                  const double *het_per_min, const double *seawater,
                  double *y_ppm, int32_t *n_accepted, int32_t *n_rejected,
                  cmaq_ros3_fail *fail);

/*
 * Same as cmaq_ros3_run with every array already in device memory (HBM),
 * enqueued on `stream` (a hipStream_t; NULL = default stream).  Returns after
 * the stream has finished so that failures are reported.
 */
int cmaq_ros3_run_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                         const double *temp_k, const double *pres_atm,
                         const double *h2o_ppm, const double *rj_per_min,
                         const double *het_per_min, const double *seawater,
                         double *y_ppm, int32_t *n_accepted, int32_t *n_rejected,
                         cmaq_ros3_fail *fail, void *stream);

/*
 * cmaq_ros3_run_device with the reference's stock block semantics: the
 * ncells cells, in array order, form NBLKS = 1 + (ncells-1)/blksize blocks of
 * IAVG = 1 + (ncells-1)/NBLKS consecutive cells, the last block holding the
 * rest (rbinit.F:286-303, blksize = BLKSIZE of rbdata_mod.F, 50 in the
 * reference; 1..64 here).  Each block decides LSUNLIGHT as any J > 0 among
 * its cells (rbdriver.F:497-519) and integrates with one step size, from
 * MAXERR = the largest error over its cells (rbsolver.F:443-456), so its
 * cells share their accepted and rejected step counts.  A block that fails
 * marks all its cells with n_accepted = -1 and *fail names its first cell;
 * cmaq_ros3_get_fail_detail then returns the cell's inputs and NaN for the
 * per-cell state.  A caller that wants the LORDERING order (rbdriver.F:
 * 683-724) permutes the cell arrays before the call.
 */
int cmaq_ros3_run_blocks_device(cmaq_ros3_ctx *ctx, int64_t ncells, int32_t blksize,
                                double chemstep_min, const double *temp_k,
                                const double *pres_atm, const double *h2o_ppm,
                                const double *rj_per_min, const double *het_per_min,
                                const double *seawater, double *y_ppm,
                                int32_t *n_accepted, int32_t *n_rejected,
                                cmaq_ros3_fail *fail, void *stream);

/*
 * nsteps consecutive chemistry sync steps (CHEM calls, rbdriver.F:20) with
 * the same met inputs, on device arrays.  Each sync step -- reset of the cell
 * queue, CALC_RCONST, RBSOLVER over all cells, checkpoint -- is captured once
 * in a hipGraph and replayed per step; the graph is rebuilt only when the
 * sizes, pointers or kernel variant change.  y is advanced in place;
 * y_ckpt [nsteps][numb_mech_spc][ncells] (may be NULL) receives y at the end
 * of every sync step, n_accepted / n_rejected [nsteps][ncells] (may be NULL)
 * the per-step counts.  The first failure of any step is reported as in
 * cmaq_ros3_run_device; the other cells and steps still run.  Ordered after
 * prior work on `stream`; returns when all steps have finished.
 */
int cmaq_ros3_run_steps_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                               int32_t nsteps, const double *temp_k, const double *pres_atm,
                               const double *h2o_ppm, const double *rj_per_min,
                               const double *het_per_min, const double *seawater,
                               double *y_ppm, double *y_ckpt, int32_t *n_accepted,
                               int32_t *n_rejected, cmaq_ros3_fail *fail, void *stream);

/* Device time of all sync steps of the last cmaq_ros3_run_steps_device call. */
int cmaq_ros3_last_steps_timing(const cmaq_ros3_ctx *ctx, double *ms);

/*
 * RBSOLVER alone over ncells cells with the rate constants supplied
 * (the RKI array of rbdata_mod.F, filled by CALC_RCONST at rbdriver.F:521):
 * rki [nrxns][ncells] in ppm-min units, device memory.  Otherwise as
 * cmaq_ros3_run_device.  Every cell uses the day pattern; for a cell with
 * J = 0 the result is identical to the reference's night pattern.
 */
int cmaq_ros3_solve_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                           const double *rki, double *y_ppm, int32_t *n_accepted,
                           int32_t *n_rejected, cmaq_ros3_fail *fail, void *stream);

/*
 * Allocation-free variant of cmaq_ros3_solve_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results, bit for bit; the call makes no hipMalloc, hipFree
 * or other allocating call.  Its solver workspace, cell queue and failure
 * record live in one static device array of 256 MiB per GPU, held until the
 * process ends, instead of the context's buffers: the persistent kernel runs
 * with as many waves as that array holds (the forward's workspace rows x 512
 * B per wave in 1 MiB less than 256 MiB: 152 waves for CB6r5), so it is
 * slower than cmaq_ros3_solve_device.  A mechanism whose one workgroup (4
 * waves) does not fit is refused with CMAQ_ROS3_E_NOMEM.  The array keeps the
 * last call's contents and is shared by every context on the GPU: calls that
 * overlap on one GPU race.  The context (cmaq_ros3_init, which allocates) is
 * created beforehand.
 */
int cmaq_ros3_solve_device_noalloc(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep,
                                   const double *rki, double *y_ppm, int32_t *n_accepted,
                                   int32_t *n_rejected, cmaq_ros3_fail *fail, void *stream);

/*
 * Discrete adjoint of one cmaq_ros3_solve_device call (no counterpart in the
 * reference, whose sensitivity option is DDM-3D, ddm3d/DDM3D_CHEM.F:155-511).
 * Given the call's inputs -- rki, the initial state y0 [ns][ncells] and the
 * n_accepted it returned -- and lambda = dL/dy_end [ns][ncells] (mechanism
 * order), overwrites lambda with dL/dy0 and adds dL/drki to grad_rki
 * [nrxns][ncells] (NULL skips it).  The accepted step sizes are held fixed
 * (the step-size controller is not differentiated).  The forward is replayed
 * cell by cell into per-step tapes, longest cells first, in chunks of cells
 * whose tapes fit 80 % of the free device memory (CMAQ_ROS3_TAPE_BYTES lowers
 * the budget; tapes above 8 GiB are released when the call returns); a
 * replay that does not reproduce n_accepted returns CMAQ_ROS3_E_TAPE.  Cells with
 * n_accepted < 0 (failed forward) get NaN in lambda.  Device memory, `stream`
 * as in cmaq_ros3_run_device; returns after the stream has finished.
 */
int cmaq_ros3_adjoint_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                             const double *rki, const double *y0_ppm,
                             const int32_t *n_accepted, double *lambda,
                             double *grad_rki, cmaq_ros3_fail *fail, void *stream);

/*
 * Stored-tape adjoint, in two calls per sync step (no counterpart in the
 * reference, as for cmaq_ros3_adjoint_device).
 *
 * cmaq_ros3_solve_tape_device integrates exactly as cmaq_ros3_solve_device
 * (same y, n_accepted, n_rejected, failures) and also stores every accepted
 * step (Y, K1, K2, K3, h) as one record of
 * cmaq_ros3_query(ctx, "tape_record_doubles") doubles in the caller's arena
 * `tape` of tape_cap records.  *tape_count (device, uint64) is the number of
 * records taken so far: zero it before the first sync step; later calls
 * append, so one arena holds a whole integration.  tape_head [ncells]
 * (device, int64) receives each cell's last record of this sync step (-1 for
 * a failed cell).  If the arena overflows, the integration is still complete
 * but the tape is not: the call returns CMAQ_ROS3_E_NOMEM.
 *
 * cmaq_ros3_reverse_device is the reverse sweep of cmaq_ros3_adjoint_device
 * over one such sync step, with no replay: lambda and grad_rki as there,
 * n_accepted and tape_head from the matching cmaq_ros3_solve_tape_device
 * call.  It returns CMAQ_ROS3_E_NOMEM without launching if *tape_count
 * exceeds tape_cap.  Device memory, `stream` as in cmaq_ros3_run_device;
 * both return after the stream has finished.
 */
int cmaq_ros3_solve_tape_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                                const double *rki, double *y_ppm, int32_t *n_accepted,
                                int32_t *n_rejected, double *tape, int64_t tape_cap,
                                uint64_t *tape_count, int64_t *tape_head,
                                cmaq_ros3_fail *fail, void *stream);
int cmaq_ros3_reverse_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                             const double *rki, const int32_t *n_accepted,
                             const double *tape, int64_t tape_cap,
                             const uint64_t *tape_count, const int64_t *tape_head,
                             double *lambda, double *grad_rki, cmaq_ros3_fail *fail,
                             void *stream);

/*
 * Tangent-linear (forward-mode) sensitivities along ndir directions (1..8),
 * carried through a cmaq_ros3_solve_device call (no counterpart in the
 * reference; an exact alternative to DDM-3D's Crank-Nicolson sensitivity
 * step, ddm3d/DDM3D_CHEM.F:155-511, whose parameters are the directions).
 * Integrates exactly as cmaq_ros3_solve_device (same steps, same y,
 * n_accepted, n_rejected) and, with each accepted step's own stages and
 * factorization, advances dy [ndir][ns][ncells] (mechanism order, in/out:
 * dy0 in, d y_end out) along the rate-constant directions dk
 * [ndir][nrxns][ncells] (NULL = 0).  The accepted step sizes are held fixed,
 * as in cmaq_ros3_adjoint_device, so for each direction
 * w . dy_end = lambda0 . dy0 + grad_rki . dk.  Cells that fail get NaN in dy.
 * More than 8 directions: further calls from the same y0.  Device memory,
 * `stream` as in cmaq_ros3_run_device.
 */
int cmaq_ros3_tlm_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                         const double *rki, double *y_ppm, int32_t *n_accepted,
                         int32_t *n_rejected, int32_t ndir, double *dy, const double *dk,
                         cmaq_ros3_fail *fail, void *stream);

/*
 * cmaq_ros3_solve_device that also returns the time-averaged state of the
 * ros3 -Dsens build, YAVE = sum over accepted steps of 0.5 * (Y + MAX(YP,
 * CONMIN)) * DT, divided by CHEMSTEP (rbsolver.F:341-343, 468-472,
 * 592-596): yave [numb_mech_spc][ncells] out, mechanism order, NaN for a
 * failing cell.  Same steps, y and counts as cmaq_ros3_solve_device.
 */
int cmaq_ros3_solve_yave_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep,
                                const double *rki, double *y_ppm, double *yave,
                                int32_t *n_accepted, int32_t *n_rejected, cmaq_ros3_fail *fail,
                                void *stream);

/*
 * CMAQ-DDM-3D's chemistry step (SOLVE_DDM3D_CHEM, ddm3d/DDM3D_CHEM.F:155-511)
 * after one sync step: cmaq_ros3_solve_yave_device (y advanced, yave out --
 * NULL keeps it internal), then for each of np <= 16 sensitivity parameters
 * one Crank-Nicolson step over the sync step with the REAL(4) Jacobian at
 * YAVE, A S_new = (I + h/2 J) S_old + P, solved with LINPACK SGEFA / SGESL
 * (ddm3d/s_lu.F) in the solver's species order as the reference does.
 * prate [np] (host): 0 for a parameter whose chemistry step is propagation
 * only (initial conditions, emissions, boundaries), r > 0 for the rate
 * constant of reaction r (1-based; the sensitivity to ln k, IPT = 5).
 * sens [np][numb_mech_spc][ncells] REAL(4), mechanism order, in/out (the
 * cell-fastest slices of SENGRID); a failing cell gets NaN.  n_accepted is
 * required.  Device memory, `stream` as in cmaq_ros3_run_device.
 */
int cmaq_ros3_ddm_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep, const double *rki,
                         double *y_ppm, double *yave, int32_t np, const int32_t *prate, float *sens,
                         int32_t *n_accepted, int32_t *n_rejected, cmaq_ros3_fail *fail, void *stream);

/*
 * Scheduling hint for the forward calls (cmaq_ros3_run_device,
 * cmaq_ros3_solve_device and cmaq_ros3_run): cells are handed to lanes in
 * the order order[0..n-1] (a device array holding a permutation of
 * 0..n-1, e.g. the cells sorted by their attempted-step counts of the
 * previous sync step, longest first) whenever a call has exactly n cells.
 * Results do not depend on the order.  NULL clears the hint.  The array must
 * stay valid while the hint is set.  No counterpart in the reference (whose
 * LORDERING pass, rbdriver.F:430-434, 682-724, reorders cells into blocks).
 */
int cmaq_ros3_set_queue_order(cmaq_ros3_ctx *ctx, const int32_t *order, int64_t n);

/* Device time of the last cmaq_ros3_adjoint_device call: replay (tape
 * writing) and reverse kernels, and the number of cell chunks. */
int cmaq_ros3_last_adjoint_timing(const cmaq_ros3_ctx *ctx, double *replay_ms,
                                  double *reverse_ms, int64_t *chunks);

/*
 * CALC_RCONST alone (RXNS_FUNC_MODULE.F90:204-1116) on device arrays:
 * rki [nrxns][ncells] in ppm-min units.  Day/night is decided per cell.
 */
int cmaq_ros3_rates_device(cmaq_ros3_ctx *ctx, int64_t ncells,
                           const double *temp_k, const double *pres_atm,
                           const double *h2o_ppm, const double *rj_per_min,
                           const double *het_per_min, const double *seawater,
                           double *rki, void *stream);

/*
 * Device time of the kernels of the last cmaq_ros3_run / _run_device call on
 * this context, measured with HIP events on the call's stream:
 * rates_ms = CALC_RCONST kernel, solve_ms = persistent ROS3 kernel.
 */
int cmaq_ros3_last_timing(const cmaq_ros3_ctx *ctx, double *rates_ms, double *solve_ms);

/* ---- CHEM level: CGRID in, CGRID out ----------------------------------- */

/*
 * The arrays of one CHEM call (rbdriver.F:20) as CCTM holds them, Fortran
 * column-major REAL(4) unless noted; "grid" = (NCOLS,NROWS,NLAYS):
 *   cgrid          CONC(NCOLS,NROWS,NLAYS,NSPCSD), in/out (rbdriver.F:100):
 *                  gas ppm, aerosol species ug/m3 (converted as CONVERT_CONC)
 *   ta, dens,      TA (K), DENS (kg/m3), PRES (Pa), QV (kg/kg) on the grid
 *   pres, qv       (rbdriver.F:169-172, interpolated by CHEM at :390-403)
 *   seawater_zone  SEAWATER_ZONE(NCOLS,NROWS) = OCEAN + SZONE (rbdriver.F:253)
 *   seaice         SEAICE(NCOLS,NROWS) (rbdriver.F:390); either may be NULL = 0
 *   rj             RJ(NCOLS,NROWS,NLAYS,NPHOTAB) 1/min (PHOT_MOD.F:109)
 *   khetero        REAL(8) KHETERO(ld_het,NCOLS,NROWS,NLAYS) 1/min
 *                  (AEROSOL_CHEMISTRY.F:46, 712), ld_het >= NHETERO; NULL = 0
 * Cells are integrated in CGRID's own order, cell = col + ncols*(row +
 * nrows*lay); results do not depend on the order (per-cell step control).
 */
typedef struct cmaq_ros3_grid {
    int32_t ncols, nrows, nlays, ld_het;
    float *cgrid;
    const float *ta, *dens, *pres, *qv;
    const float *seawater_zone, *seaice;
    const float *rj;
    const double *khetero;
} cmaq_ros3_grid;

/*
 * Species map of MAP_CHEMISTRY_SPECIES / RESET_SPECIES_POINTERS
 * (RXNS_FUNC_MODULE.F90:1117-1305) and RBINIT's conversion factors
 * (rbinit.F:257-262), per mechanism species in CHEMISTRY_SPC order:
 * cgrid_index (1-based CGRID species, <= nspcsd, all distinct),
 * convert_conc (nonzero = CGRID holds ug/m3: FORWARD_CONV = 1e-3*MWAIR/MOLWT
 * in REAL(4)) and species_molwt (RXNS_DATA SPECIES_MOLWT, g/mol).  Required
 * before the calls below.
 */
int cmaq_ros3_set_cgrid_map(cmaq_ros3_ctx *ctx, int32_t nspcsd, const int32_t *cgrid_index,
                            const int32_t *convert_conc, const double *species_molwt);

/*
 * The whole gas-chemistry part of CHEM for every cell of the grid: the
 * gather with unit conversion (rbdriver.F:460-509), CALC_RCONST + RBSOLVER
 * (as cmaq_ros3_run_device) and the scatter back into CGRID
 * (rbdriver.F:577-595).  Host arrays; the CGRID planes of the mechanism
 * species, the met fields, RJ and KHETERO are copied to the device and the
 * planes back.  n_accepted / n_rejected [ncells] (may be NULL), failures as
 * in cmaq_ros3_run (a failing cell's CGRID holds its last accepted state).
 */
int cmaq_ros3_chem(cmaq_ros3_ctx *ctx, const cmaq_ros3_grid *grid, double chemstep_min,
                   int32_t *n_accepted, int32_t *n_rejected, cmaq_ros3_fail *fail);

/* Same as cmaq_ros3_chem with every array of *grid and the counts in device
 * memory, enqueued on `stream`; returns after the stream has finished. */
int cmaq_ros3_chem_device(cmaq_ros3_ctx *ctx, const cmaq_ros3_grid *grid, double chemstep_min,
                          int32_t *n_accepted, int32_t *n_rejected, cmaq_ros3_fail *fail,
                          void *stream);

/*
 * The two halves of cmaq_ros3_chem_device on device arrays.  gather fills the
 * block arrays of rbdriver.F:460-509 for every cell -- temp_k, pres_atm,
 * h2o_ppm, seawater [ncells]; rj_per_min [nphotab][ncells]; het_per_min
 * [nhetero][ncells] (may be NULL); y_ppm [numb_mech_spc][ncells] in
 * mechanism order -- in the reference's REAL(4) arithmetic, bit for bit.
 * scatter writes y_ppm back into grid->cgrid (rbdriver.F:577-595), reading
 * grid->dens.  Enqueued on `stream`, asynchronous.
 */
int cmaq_ros3_gather_device(cmaq_ros3_ctx *ctx, const cmaq_ros3_grid *grid, double *temp_k,
                            double *pres_atm, double *h2o_ppm, double *rj_per_min,
                            double *het_per_min, double *seawater, double *y_ppm, void *stream);
int cmaq_ros3_scatter_device(cmaq_ros3_ctx *ctx, const cmaq_ros3_grid *grid, const double *y_ppm,
                             void *stream);

/*
 * Kernel variants: every mechanism is generated in several register-budget /
 * load-lookahead configurations (chosen per mechanism by the autotuner and
 * overridable with the CMAQ_ROS3_VARIANT environment variable at init).
 * cmaq_ros3_variant(ctx, -1) names the active one, index >= 0 lists them
 * (NULL past the end); cmaq_ros3_set_variant switches.
 */
int cmaq_ros3_set_variant(cmaq_ros3_ctx *ctx, const char *variant);
const char *cmaq_ros3_variant(const cmaq_ros3_ctx *ctx, int index);

/* Sizes of the selected pattern: key in {"ischan","nnz_day","nnz_night",
 * "lu_ops_day","lu_ops_night","nrxns","numb_mech_spc","grid_waves"
 * (workspace waves),"solve_waves" (waves of the forward kernel's grid)}. */
int cmaq_ros3_query(const cmaq_ros3_ctx *ctx, const char *key, int64_t *value);

/* Solver species order: inew2old[new] = old (both 0-based), numb_mech_spc entries. */
int cmaq_ros3_species_order(const cmaq_ros3_ctx *ctx, int32_t *inew2old);

/*
 * Details of the first failing cell of the last run/solve/CHEM call on this
 * context (rbsolver.F:532-583 writes the same for CELL_MAXERR).  Per species,
 * in mechanism (CHEMISTRY_SPC) order: spc_err = MAX_SPC_ERR, the species'
 * share of ERR, ((D1*K1 + D2*K2 + D3*K3) / (ATOL + RTOL*|YP|))**2 / ERR
 * (rbsolver.F:551-555); y = the last accepted state Y; yp = the rejected
 * step's new state YP.  rj [nphotab] = RJBLK and rki [nrxns] = RKI of the
 * cell (NaN where the call had no such input).  Any array may be NULL.  With
 * no failure d->code = 0 and the arrays are NaN.
 */
int cmaq_ros3_get_fail_detail(const cmaq_ros3_ctx *ctx, cmaq_ros3_fail_detail *d, double *spc_err,
                              double *y, double *yp, double *rj, double *rki);

/* ---- SMVGEAR (CCTM/src/gas/smvgear) --------------------------------------
 * The Gear solver of CMAQ's smvgear CHEM driver on the same context: the
 * mechanism's tables, species order and sparse pattern are shared with ROS3
 * (grsprse.F's JSPARSE is RBSPARSE's algorithm).  Drop-in for the block loop
 * of grdriver.F:400-560 -- CALC_RCONST (grdriver.F:450) + SMVGEAR
 * (grsmvgear.F:27, grdriver.F:520) -- over every cell in one call, with
 * per-cell step and order control (NUMCELLS = 1), LSUNLIGHT per cell (any
 * J > 0) choosing HMAX (15 / 5 min) and the first-step factor H0FAC.
 * Concentrations enter as MAX(y, CONCMIN) (grdriver.F:419-422).
 */

/* Gear tolerances, REAL(4) as RUNTIME_VARS.F:173-176 holds them: GEAR_RTOL,
 * GEAR_MAX_ATOL, GEAR_MIN_ATOL, GEAR_CONC_FLOOR (defaults 1e-3, 1e-7, 1e-9,
 * 1e-12 ppm); the absolute tolerance of each cell is chosen among six levels
 * between the two bounds (OPTIMAL_ATOL_PPM, GRVARS.F:550-707). */
typedef struct cmaq_gear_tol {
    float rtol, max_atol, min_atol, conc_floor;
} cmaq_gear_tol;

/* Sets the context's Gear tolerances (NULL = the defaults).  Requires
 * max_atol > min_atol > conc_floor > 0 and rtol > 0, as OPTIMAL_ATOL_PPM
 * does (GRVARS.F:588-605, where the reference stops the run). */
int cmaq_gear_set_tolerances(cmaq_ros3_ctx *ctx, const cmaq_gear_tol *tol);

/* CALC_RCONST + SMVGEAR over ncells cells, host arrays, synchronous; the
 * arguments as cmaq_ros3_run.  n_steps receives the cell's successful Gear
 * steps (NSTEPS), n_fail its error-test plus convergence failures (NEFAIL +
 * NCFAIL).  Where the reference would stop the run (YLOWA lowered
 * MXLOWRED = 20 times, grsmvgear.F:425-466, or 100 restarts,
 * grsmvgear.F:768-772) the cell gets n_steps = -1 and the call returns
 * CMAQ_ROS3_E_CONVERGENCE with *fail holding the cell, DELT (dt), XELAPS
 * (tnow) and YLOWA (maxerr); every other cell is integrated. */
int cmaq_gear_run(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                  const double *temp_k, const double *pres_atm, const double *h2o_ppm,
                  const double *rj_per_min, const double *het_per_min, const double *seawater,
                  double *y_ppm, int32_t *n_steps, int32_t *n_fail, cmaq_ros3_fail *fail);

/* Same as cmaq_gear_run on device arrays, enqueued on `stream`; returns
 * after the stream has finished. */
int cmaq_gear_run_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min,
                         const double *temp_k, const double *pres_atm, const double *h2o_ppm,
                         const double *rj_per_min, const double *het_per_min, const double *seawater,
                         double *y_ppm, int32_t *n_steps, int32_t *n_fail, cmaq_ros3_fail *fail,
                         void *stream);

/* SMVGEAR alone with the rate constants supplied (RK of GRVARS, filled by
 * CALC_RCONST at grdriver.F:450): rki [nrxns][ncells] and sunlit [ncells]
 * (uint8 LSUNLIGHT per cell; NULL = every cell sunlit), device memory;
 * otherwise as cmaq_gear_run_device. */
int cmaq_gear_solve_device(cmaq_ros3_ctx *ctx, int64_t ncells, double chemstep_min, const double *rki,
                           const uint8_t *sunlit, double *y_ppm, int32_t *n_steps, int32_t *n_fail,
                           cmaq_ros3_fail *fail, void *stream);

/* Scheduling of long cells; no counterpart in the reference, whose block
 * loop runs each block to its end on one core.  A cell still integrating
 * after after_attempts attempted Gear steps (default 2048; 0 = never) leaves
 * the cell-per-lane kernel at the top of its next step and is finished by a
 * second kernel that runs one cell per wavefront with its workspace in LDS,
 * at a lower latency per step; at most max_cells cells per call do so
 * (0 = automatic: ncells / 32, at least 1024).  Results do not depend on
 * either value: both kernels execute the same operations in the same order.
 * cmaq_ros3_query(ctx, "gear_tail_cells_last") gives the number of cells the
 * last Gear call handed over. */
int cmaq_gear_set_tail(cmaq_ros3_ctx *ctx, int32_t after_attempts, int32_t max_cells);

/* Device time (ms) of the last SMVGEAR call's rates and Gear kernels. */
int cmaq_gear_last_timing(const cmaq_ros3_ctx *ctx, double *rates_ms, double *solve_ms);

/* Message for the last failure on this context (static storage, never NULL). */
const char *cmaq_ros3_last_error(const cmaq_ros3_ctx *ctx);

void cmaq_ros3_free(cmaq_ros3_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* CMAQ_ROS3_H */
