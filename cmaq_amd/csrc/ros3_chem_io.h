// CHEM-level gather and scatter between CCTM's CGRID and the solver's
// per-cell arrays (CCTM/src/gas/ros3/rbdriver.F:460-509 and :577-595), in
// the reference's REAL(4) arithmetic so that the block arrays are bit for bit
// those the Fortran driver builds.  One thread per cell; every array is read
// and written one species (or field) plane at a time, so a wave touches 64
// consecutive cells of one plane per access.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ros3 {

// REAL(4) constants of rbdriver.F:110-114 and ICL/fixed/const/CONST.EXT:95-119
constexpr float kConcMin = 1.0e-30f;                          // CONCMIN
constexpr float kMwAir = 28.9628f, kMwWat = 18.0153f;         // MWAIR, MWWAT
constexpr float kPa2Atm = 1.0f / 101325.0f;                   // PA2ATM = 1 / STDATMPA
constexpr float kQvToPpm = 1.0e6f * (kMwAir / kMwWat);        // CONCOFM * MAOMV

// one mechanism species, CHEMISTRY_SPC order
struct ChemMap {
    int32_t spc;     // 0-based CGRID species (CGRID_INDEX - 1)
    int32_t conv;    // CONVERT_CONC
    float fwd;       // FORWARD_CONV = 1e-3 * MWAIR / MOLWT, REAL(4) (rbinit.F:260)
    float pad;
    double rev;      // REVERSE_CONV = REAL(1.0 / FORWARD_CONV, 8) (rbinit.F:261)
};
static_assert(sizeof(ChemMap) == 24, "ChemMap layout");

constexpr int kMaxChemSpc = 512;   // species staged in LDS (12 KB)
constexpr int kBatch = 16;         // planes loaded together per thread

struct ChemParams {
    int64_t n;                 // cells = ncols * nrows * nlays
    int64_t plane2;            // ncols * nrows (2-D fields)
    int32_t ns, nphotab, nhetero, ld_het;
    const ChemMap *map;
    float *cgrid;              // [nspcsd][n]
    const float *ta, *dens, *pres, *qv;   // [n]
    const float *zone, *ice;   // [plane2] or nullptr
    const float *rj;           // [nphotab][n]
    const double *khet;        // [n][ld_het] or nullptr
    double *temp, *pres_atm, *h2o, *sea, *rjd, *het, *y;   // block arrays (cell fastest)
};

__global__ __launch_bounds__(256) void gather_kernel(ChemParams P)
{
    __shared__ ChemMap s_map[kMaxChemSpc];
    for (int i = threadIdx.x; i < P.ns; i += blockDim.x) s_map[i] = P.map[i];
    __syncthreads();
    const int64_t n = P.n;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        // rbdriver.F:465-470
        const float dens = P.dens[c];
        const float svol = 1.0f / dens;
        P.temp[c] = (double)P.ta[c];
        P.pres_atm[c] = (double)(kPa2Atm * P.pres[c]);
        P.h2o[c] = (double)fmaxf(kQvToPpm * P.qv[c], 0.0f);
        const int64_t c2 = c % P.plane2;
        const float zone = P.zone ? P.zone[c2] : 0.0f, ice = P.ice ? P.ice[c2] : 0.0f;
        P.sea[c] = (double)fmaxf(0.0f, zone - ice);
        // rbdriver.F:476-492, kBatch planes loaded before any is stored so
        // that each wave keeps several loads in flight
        for (int i0 = 0; i0 < P.ns; i0 += kBatch) {
            float conc[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                conc[u] = i0 + u < P.ns ? P.cgrid[(int64_t)s_map[i0 + u].spc * n + c] : 0.0f;
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                if (i0 + u >= P.ns) break;
                const ChemMap m = s_map[i0 + u];
                const float v = m.conv ? (m.fwd * svol) * conc[u] : conc[u];
                P.y[(int64_t)(i0 + u) * n + c] = (double)fmaxf(v, kConcMin);
            }
        }
        // rbdriver.F:499-510
        for (int k0 = 0; k0 < P.nphotab; k0 += kBatch) {
            float j[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                j[u] = k0 + u < P.nphotab ? P.rj[(int64_t)(k0 + u) * n + c] : 0.0f;
#pragma unroll
            for (int u = 0; u < kBatch; ++u)
                if (k0 + u < P.nphotab) P.rjd[(int64_t)(k0 + u) * n + c] = (double)j[u];
        }
        if (P.het)
            for (int k = 0; k < P.nhetero; ++k)
                P.het[(int64_t)k * n + c] = P.khet ? P.khet[c * P.ld_het + k] : 0.0;
    }
}

__global__ __launch_bounds__(256) void scatter_kernel(ChemParams P)
{
    __shared__ ChemMap s_map[kMaxChemSpc];
    for (int i = threadIdx.x; i < P.ns; i += blockDim.x) s_map[i] = P.map[i];
    __syncthreads();
    const int64_t n = P.n;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        const double dens = (double)P.dens[c];          // BLKDENS, rbdriver.F:466
        // rbdriver.F:580-595, batched as in the gather
        for (int i0 = 0; i0 < P.ns; i0 += kBatch) {
            double yv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) yv[u] = i0 + u < P.ns ? P.y[(int64_t)(i0 + u) * n + c] : 0.0;
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                if (i0 + u >= P.ns) break;
                const ChemMap m = s_map[i0 + u];
                P.cgrid[(int64_t)m.spc * n + c] = m.conv ? (float)((m.rev * dens) * yv[u]) : (float)yv[u];
            }
        }
    }
}

}  // namespace ros3
