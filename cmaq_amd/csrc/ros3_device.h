// Device-side pieces shared by every generated mechanism: the ROS3
// constants and step controller (CCTM/src/gas/ros3/rbsolver.F:85-113,
// 433-530), the table-driven rate-constant evaluator (replacing the
// chemmech-generated CALC_RCONST, RXNS_FUNC_MODULE.F90:204-1116) and the
// persistent, queue-refilled integration kernel.
//
// Memory layout.  Every per-cell vector lives in a lane-private workspace
// laid out [wave][entry][64 lanes]: entry k of lane l of wave w is at
// ws[(w * NWS + k) * 64 + l].  A wave's 64 lanes therefore read one entry of
// 64 different cells as one contiguous 512-byte segment, and the generated
// code addresses every entry with a compile-time offset from one per-lane
// base pointer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ros3_registry.h"

namespace ros3 {

constexpr int kWave = 64;
typedef __attribute__((ext_vector_type(2))) unsigned int u32x2;

// ---- ROS3 (Sandu et al. 1997, KPP) in the reference's G-form ----------
constexpr double GAM = 0.43586652150845899941601945119356e+00;
constexpr double C21 = -0.10156171083877702091975600115545e+01;
constexpr double C31 = 0.40759956452537699824805835358067e+01;
constexpr double C32 = 0.92076794298330791242156818474003e+01;
constexpr double B1 = 0.10000000000000000000000000000000e+01;
constexpr double B2 = 0.61697947043828245592553615689730e+01;
constexpr double B3 = -0.42772256543218573326238373806514e+00;
constexpr double D1 = 0.50000000000000000000000000000000e+00;
constexpr double D2 = -0.29079558716805469821718236208017e+01;
constexpr double D3 = 0.22354069897811569627360909276199e+00;
constexpr double GROW = 1.0 / 3.00;
constexpr double RGAM = 1.0 / GAM;
constexpr double DTSTART = 0.5e-01;   // minutes
constexpr double DTMIN = 1.0e-08;
constexpr double DTMAX = 1.0e+01;
constexpr double UROUND = 1.0e-18;
constexpr double FACMAX = 1.0e+01;
constexpr double FACMIN = 1.0e-01;
constexpr double FACONE = 1.0;
constexpr double CONMIN = 1.0e-30;

// ---- rate-constant program --------------------------------------------
// One record per reaction, decoded on the host from KTYPE/RTDAT/RFDAT/
// IORDER/IPH/IHETERO with chemmech's formula and unit-factor rules
// (UTIL/chemmech/src/wrt_reactions_module.f:580-800,
//  BASIC_WRITE_ROUTINES.f:1102-1152).  Parameters are stored in the
// argument order of the reference rate functions.
enum RateKind : int32_t {
    RK_NONE = 0,      // stays 0 (e.g. relative rate to a later reaction)
    RK_PHOT,          // [a *] J(ref)                 (sunlit only)
    RK_HALOGEN,       // SEAWATER * SFACT * min(..)   (sunlit, seawater > 0.001)
    RK_HET,           // [a *] HET(ref)
    RK_CONST,         // a [* conv]
    RK_T02,           // conv * a*(T/300)^b
    RK_T03,           // conv * a*exp(b/T)
    RK_T04,           // conv * a*exp(b/T)*(T/300)^c
    RK_T05,           // [conv *] K(ref) * (a*exp(b/T))
    RK_T06,           // [conv *] [a *] K(ref)
    RK_T08,
    RK_T09,
    RK_T10,
    RK_T11,
    RK_T07,           // a*(1 + 0.6 P), or a*(T/300)^b * P (P in atm)
};
enum ConvKind : int32_t {   // chemmech unit factor; CV_NONE writes none
    CV_NONE = -1, CV_INV_CFACT = 0, CV_SFACT, CV_CFACT, CV_CFACT_SQU, CV_RFACT, CV_RFACT_SQU
};
struct RateOp {
    int32_t kind;
    int32_t conv;
    int32_t ref;      // 0-based photolysis table, het table or reaction
    int32_t scaled;   // 1 if the leading coefficient `a` multiplies
    double p[8];
};
static_assert(sizeof(RateOp) == 80, "RateOp layout");
// multipliers, applied in the reference order M, O2, N2, H2O, H2, CH4
struct RateMul {
    int32_t rxn;
    int32_t use_h2o; // 1: multiply by the cell's H2O ppm, else by `v`
    double v;
};

struct CellConv {
    double inv_temp, t300, cair, c[6];
};

__device__ inline void cell_conv(double temp, double pres, CellConv &u)
{
    const double COEF1 = 7.33981e+15, TI300 = 1.0 / 300.0;
    u.inv_temp = 1.0 / temp;
    u.cair = 1.0e+06 * COEF1 * pres * u.inv_temp;
    u.c[CV_CFACT] = 6.0e-05 * u.cair;
    u.c[CV_CFACT_SQU] = 6.0e-11 * u.cair * u.cair;
    double inv_cair = 1.0 / u.cair;
    u.c[CV_INV_CFACT] = 6.0e+07 * inv_cair;
    u.t300 = temp * TI300;
    u.c[CV_RFACT] = 1.0e+06 * inv_cair;
    u.c[CV_RFACT_SQU] = 1.0e+12 * inv_cair * inv_cair;
    u.c[CV_SFACT] = 60.0;
}

// Evaluates reaction `op` for one cell.  `kref` is K(ref) for RK_T05/06.
__device__ inline double rate_eval(const RateOp &op, const CellConv &u, double kref,
                                   double jval, double hval, double sea, double pres,
                                   bool sunlit, bool &skip)
{
    const double *p = op.p;
    double v;
    skip = false;
    switch (op.kind) {
    case RK_PHOT:
        if (!sunlit) return 0.0;
        return op.scaled ? p[0] * jval : jval;
    case RK_HALOGEN: {
        if (!sunlit) return 0.0;
        if (!(sea > 0.001)) return 0.0;
        double h = p[0] * exp(p[1] * pres) + p[2] * exp(p[3] * pres);
        h = p[4] < h ? p[4] : h;
        return sea * u.c[CV_SFACT] * h;
    }
    case RK_HET:
        return op.scaled ? p[0] * hval : hval;
    case RK_CONST:
        return op.conv >= 0 ? p[0] * u.c[op.conv] : p[0];
    case RK_T02:
        v = p[0] * pow(u.t300, p[1]);
        break;
    case RK_T03:
        v = p[0] * exp(p[1] * u.inv_temp);
        break;
    case RK_T04:
        v = p[0] * exp(p[1] * u.inv_temp) * pow(u.t300, p[2]);
        break;
    case RK_T05: {
        double eq = p[0] * exp(p[1] * u.inv_temp);
        return op.conv >= 0 ? u.c[op.conv] * kref * eq : kref * eq;
    }
    case RK_T06:
        if (op.scaled) return op.conv >= 0 ? u.c[op.conv] * p[0] * kref : p[0] * kref;
        return op.conv >= 0 ? u.c[op.conv] * kref : kref;
    case RK_T08: {
        double k0 = p[0] * exp(p[1] * u.inv_temp);
        double k2 = p[2] * exp(p[3] * u.inv_temp);
        double k3 = p[4] * exp(p[5] * u.inv_temp);
        k3 = k3 * u.cair;
        v = k0 + k3 / (1.0 + k3 / k2);
        break;
    }
    case RK_T09: {
        double k1 = p[0] * exp(p[1] * u.inv_temp);
        double k2 = p[2] * exp(p[3] * u.inv_temp);
        v = k1 + k2 * u.cair;
        break;
    }
    case RK_T10: {
        double k0 = p[0] * u.cair * exp(p[1] * u.inv_temp) * pow(u.t300, p[2]);
        double k1 = p[3] * exp(p[4] * u.inv_temp) * pow(u.t300, p[5]);
        double g = (1.0 / p[6]) * log10(k0 / k1);
        double kend = 1.0 + g * g;
        kend = 1.0 / kend;
        v = (k0 / (1.0 + k0 / k1)) * pow(p[7], kend);
        break;
    }
    case RK_T11: {
        double k1 = p[0] * exp(p[2] * u.inv_temp) * pow(u.t300, p[1]);
        double k2 = p[3] * exp(p[5] * u.inv_temp) * pow(u.t300, p[4]);
        double k3 = p[6] * exp(p[7] * u.inv_temp);
        v = k1 + k2 * u.cair + k3;
        break;
    }
    case RK_T07:
        // chemmech writes no unit factor for this type
        if (p[1] != 0.0) return p[0] * exp(0.0 * u.inv_temp) * pow(u.t300, p[1]) * pres;
        return p[0] * (1.0 + 0.6 * pres);
    default:
        skip = true;
        return 0.0;
    }
    return op.conv >= 0 ? u.c[op.conv] * v : v;
}

struct RatesParams {
    int64_t ncells;
    int32_t nr, nphotab, nhetero, nmul;
    const RateOp *ops;
    const RateMul *mul;
    const double *temp, *pres, *h2o, *rj, *het, *sea;
    double *rki;      // [nr][ncells]
    uint8_t *sunlit;  // [ncells] or nullptr
    // blocks of cells sharing LSUNLIGHT (rbdriver.F:497-519): blk_nblks
    // blocks of blk_iavg consecutive cells, the last one holding the rest;
    // blk_iavg = 0 decides it per cell
    int64_t blk_iavg, blk_nblks;
};

// CALC_RCONST for every cell; the rate program is staged in LDS once per
// workgroup and read as wave-uniform broadcasts.  With MAXR = 0 (programs
// longer than the LDS stage, e.g. SAPRC07's 929 reactions) each wave reads
// the program from global memory at the same wave-uniform addresses.
template <int MAXR, int MAXMUL>
__global__ __launch_bounds__(256) void rates_kernel(RatesParams P)
{
    __shared__ RateOp s_ops[MAXR > 0 ? MAXR : 1];
    __shared__ RateMul s_mul[MAXMUL];
    if (MAXR > 0)
        for (int i = threadIdx.x; i < P.nr; i += blockDim.x) s_ops[i] = P.ops[i];
    for (int i = threadIdx.x; i < P.nmul; i += blockDim.x) s_mul[i] = P.mul[i];
    __syncthreads();
    const RateOp *const ops = MAXR > 0 ? s_ops : P.ops;
    const int64_t n = P.ncells;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        bool sunlit = false;
        if (P.blk_iavg > 0) {   // any J > 0 in the cell's block
            int64_t b = c / P.blk_iavg;
            if (b > P.blk_nblks - 1) b = P.blk_nblks - 1;
            const int64_t c0 = b * P.blk_iavg, c1 = b == P.blk_nblks - 1 ? n : c0 + P.blk_iavg;
            for (int k = 0; k < P.nphotab; ++k)
                for (int64_t q = c0; q < c1; ++q) sunlit |= P.rj[k * n + q] > 0.0;
        } else {
            for (int k = 0; k < P.nphotab; ++k) sunlit |= P.rj[k * n + c] > 0.0;
        }
        const double temp = P.temp[c], pres = P.pres[c];
        const double sea = P.sea ? P.sea[c] : 0.0;
        CellConv u;
        cell_conv(temp, pres, u);
        double *rk = P.rki + c;
        for (int r = 0; r < P.nr; ++r) {
            const RateOp &op = ops[r];
            double jv = 0.0, hv = 0.0, kref = 0.0;
            if (op.kind == RK_PHOT) jv = P.rj[op.ref * n + c];
            else if (op.kind == RK_HET) hv = P.het ? P.het[op.ref * n + c] : 0.0;
            else if (op.kind == RK_T05 || op.kind == RK_T06) kref = rk[op.ref * n];
            bool skip;
            double v = rate_eval(op, u, kref, jv, hv, sea, pres, sunlit, skip);
            rk[r * n] = v;
        }
        const double h2o = P.h2o[c];
        for (int q = 0; q < P.nmul; ++q) {
            const RateMul &m = s_mul[q];
            rk[m.rxn * n] = rk[m.rxn * n] * (m.use_h2o ? h2o : m.v);
        }
        if (P.sunlit) P.sunlit[c] = sunlit;
    }
}

// ---- persistent integration kernel ------------------------------------
constexpr int kTlmMaxDirections = 8;   // tangent-linear directions per call

struct SolveParams {
    int64_t ncells;           // cells of this launch: cell0 .. cell0 + ncells - 1,
    int64_t cell0;            // or order[0 .. ncells - 1] when order is set
    const int32_t *order;     // optional queue order (cell indices), else nullptr
    int64_t stride;           // leading dimension of the cell arrays
    double chemstep;
    const double *rki;        // [nr][stride]
    double *y;                // [ns][stride] mechanism order, in/out
    int32_t *nacc, *nrej;     // [stride] or nullptr
    const double *rtol, *atol;// [ns] solver (new) order
    double *ws;               // workspace: wave w's rows start at ws + w * ws_stride
    int64_t ws_stride;        // doubles per wave (>= NWS * 64)
    unsigned long long *queue;// next unassigned cell
    int32_t *fail_flag;       // 0 until the first failure
    int64_t *fail_cell;
    double *fail_info;        // maxerr, dt, tnow
    int32_t *fail_code;
    double *fail_detail;      // first failure: ERR, then Y, YP, K1, K2, K3 [ns] (solver order), or nullptr
    int32_t max_attempts;
    int32_t refill_min;       // idle lanes needed before a wave refills
    // adjoint: per-cell step tapes (records of Y, K1, K2, K3, h per accepted step)
    double *tape;             // record storage of this launch
    const int64_t *tape_off;  // [stride] first record of each cell
    const int32_t *nacc_fwd;  // [stride] accepted steps of the forward run
    double *lam;              // [ns][stride] mechanism order: dL/dy_end in, dL/dy0 out;
                              // tangent-linear run: [ndir][ns][stride] dy0 in, dy_end out
    double *grad_rki;         // [nr][stride] accumulated dL/drki, or nullptr
    const double *dk;         // tangent-linear run: [ndir][nr][stride] rate directions, or nullptr
    int32_t ndir;             // tangent-linear run: directions (1..kTlmMaxDirections)
    // stored tape (no replay): records appended to an arena of tape_cap
    // records, each linked to its cell's previous record of the sync step
    int64_t *tape_head;       // [stride] last record of each cell (-1: none), or nullptr
    unsigned long long *tape_count;   // records taken from the arena so far
    int64_t tape_cap;         // arena capacity in records
    double *yave;             // YAVE run: [ns][stride] time-averaged state, mechanism order
    // BLOCK run: blk_nblks blocks of blk_iavg consecutive cells (the last
    // one holds the rest, at most kWave cells each) sharing one step size
    int64_t blk_iavg, blk_nblks;
};

__device__ inline int lane_id() { return __lane_id(); }

}  // namespace ros3
