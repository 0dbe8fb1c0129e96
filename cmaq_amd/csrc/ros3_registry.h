// Record that each generated mechanism translation unit exports so that
// cmaq_ros3_init can select the kernels matching the tables it is given.
#pragma once

#include <stdint.h>

namespace ros3 {

struct MechKernels {
    const char *name;
    unsigned long long struct_hash;   // FNV-1a of the structural tables
    int ns, nr, nnz_day, lu_ops_day, nnz_night, lu_ops_night, nws;
    const void *solve_kernel;         // host stub of solve_kernel
    const int32_t *inew2old;          // solver order -> mechanism order (0-based)
    const char *variant;              // code-generation variant, e.g. "b48_l8"
    int reg_budget, lookahead, waves_per_eu;
    int is_default;                   // chosen when no variant is requested
    int refill_min;                   // tuned refill threshold of this variant
    const void *replay_kernel;        // solve_kernel writing per-step tapes
    const void *reverse_kernel;       // discrete adjoint over the tapes
    int tape_rs;                      // doubles per tape record
    const void *tlm_kernel;           // solve_kernel carrying a tangent-linear direction
    int nws_fwd;                      // workspace rows (entries) the forward kernel touches
    const void *yave_kernel;          // solve_kernel accumulating the time-averaged state
    const void *gear_kernel;          // SMVGEAR kernel (default variants), or nullptr
    int gear_nws;                     // workspace rows (entries) the SMVGEAR kernel uses
    const void *gear_tail_kernel;     // SMVGEAR for long cells, one per wavefront in LDS, or nullptr
    const void *block_kernel;         // solve_kernel with a step size shared by each block of cells
};

}  // namespace ros3

extern "C" const ros3::MechKernels *const ros3_mech_registry[];
extern "C" const int ros3_mech_registry_size;
