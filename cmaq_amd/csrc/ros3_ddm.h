// CMAQ-DDM-3D's chemistry step on the GPU (ddm3d/DDM3D_CHEM.F:155-511):
// for each cell, the dense Jacobian at the time-averaged state YAVE
// accumulated in REAL(4) (MECHANISM_FUNCTIONS.F:831-869), A = I - h/2 J and
// A1 = I + h/2 J in REAL(4), and per sensitivity parameter one
// Crank-Nicolson step A S_new = A1 S_old + P with LINPACK's SGEFA / SGESL
// (ddm3d/s_lu.F:271-572; column-major, partial pivoting).  Rate parameters
// (IPT = 5) add P = h k p(YAVE) S_r, the sensitivity to ln k
// (DDM3D_CHEM.F:340-384).  Species are in the solver (RBSPARSE) order inside
// the kernel, as in the reference's dense matrices, so the pivot choices and
// every REAL(4) rounding are the reference's.
//
// One workgroup integrates one cell at a time (a persistent loop over the
// cells): the N x N matrix lives in LDS when it fits (CB6r5: 89 KB, CRACMM2:
// 143 KB), else in a per-workgroup slice of global memory.  Each pivot step
// of SGEFA and each substitution step of SGESL is one or two barrier-
// separated parallel phases over the rows (and columns / parameters).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ros3 {

constexpr int kDdmMaxParams = 16;   // sensitivity parameters per call
constexpr int kDdmBlock = 256;

// One reaction in the solver's species order (reactant order as IRR holds it).
struct DdmRxn {
    int32_t nre, npr, poff, pad;
    int32_t re[3];
    int32_t pad2;
};

struct DdmParams {
    int64_t n;                     // cells
    int32_t ns, nact, np, a_in_lds;
    const DdmRxn *rx;              // [nr], by reaction
    const int32_t *act;            // [nact] active (NREACT > 0) reactions, reference order
    const int32_t *pid;            // product species (solver order), by rx[r].poff
    const double *psc;             // product coefficients SC, same indexing
    const int32_t *inew2old;       // [ns] solver order -> mechanism order
    int32_t prate[kDdmMaxParams];  // 0: initial-condition-like parameter; r > 0: ln k of reaction r (1-based)
    double chemstep;
    const double *rki;             // [nr][n]
    const double *yave;            // [ns][n] mechanism order
    const int32_t *nacc;           // [n] accepted steps of the sync step, < 0: failed cell
    float *sens;                   // [np][ns][n] mechanism order, in/out
    float *ga;                     // [gridDim.x][ns * ns] when the matrix is not in LDS
};

__device__ inline double ddm_pexcept(const DdmRxn &x, const double *y, int skip)
{
    double p = 1.0;
    for (int q = 0; q < x.nre; ++q)
        if (q != skip) p *= y[x.re[q]];
    return p;
}

__global__ __launch_bounds__(kDdmBlock) void ddm_kernel(DdmParams P)
{
    extern __shared__ __align__(16) unsigned char ddm_smem[];
    const int ns = P.ns, np = P.np, t = threadIdx.x;
    const int64_t n = P.n;
    // LDS: [A ns*ns floats if a_in_lds] yav[ns] doubles, rhs[np*ns], sold[ns], tvec[ns], ipvt[ns],
    // tp[np], reduction scratch
    unsigned char *sp = ddm_smem;
    float *A = P.a_in_lds ? (float *)sp : P.ga + (int64_t)blockIdx.x * ns * ns;
    if (P.a_in_lds) sp += (size_t)ns * ns * sizeof(float);
    double *yav = (double *)sp; sp += (size_t)ns * sizeof(double);
    float *rhs = (float *)sp; sp += (size_t)np * ns * sizeof(float);
    float *sold = (float *)sp; sp += (size_t)ns * sizeof(float);
    float *tvec = (float *)sp; sp += (size_t)ns * sizeof(float);
    int *ipvt = (int *)sp; sp += (size_t)ns * sizeof(int);
    float *tp = (float *)sp; sp += kDdmMaxParams * sizeof(float);
    float *rv = (float *)sp; sp += (kDdmBlock / 64) * sizeof(float);
    int *ri = (int *)sp; sp += (kDdmBlock / 64) * sizeof(int);
    __shared__ float s_scal;
    __shared__ int s_piv, s_zero;
    const double h = P.chemstep;

    for (int64_t cell = blockIdx.x; cell < n; cell += gridDim.x) {
        if (P.nacc[cell] < 0) {          // the forward failed: no sensitivities
            for (int idx = t; idx < np * ns; idx += kDdmBlock)
                P.sens[((int64_t)(idx / ns) * ns + idx % ns) * n + cell] = __builtin_nanf("");
            continue;
        }
        for (int i = t; i < ns; i += kDdmBlock) yav[i] = P.yave[(int64_t)P.inew2old[i] * n + cell];
        for (int idx = t; idx < ns * ns; idx += kDdmBlock) A[idx] = 0.0f;
        __syncthreads();
        // ---- Jacobian at YAVE in REAL(4): column col by one thread, its
        // entries accumulated reaction by reaction in the reference order
        for (int col = t; col < ns; col += kDdmBlock) {
            float *acol = A + (int64_t)col * ns;
            for (int a = 0; a < P.nact; ++a) {
                const int r = P.act[a];
                const DdmRxn x = P.rx[r];
                for (int q = 0; q < x.nre; ++q) {
                    if (x.re[q] != col) continue;
                    const double ex = P.rki[(int64_t)r * n + cell] * ddm_pexcept(x, yav, q);
                    for (int q2 = 0; q2 < x.nre; ++q2) acol[x.re[q2]] += (float)(-1.0 * ex);
                    for (int p = 0; p < x.npr; ++p) acol[P.pid[x.poff + p]] += (float)(P.psc[x.poff + p] * ex);
                }
            }
        }
        __syncthreads();
        // ---- right-hand sides A1 S_old + P_rate, parameter by parameter
        for (int p = 0; p < np; ++p) {
            float *b = rhs + p * ns;
            for (int i = t; i < ns; i += kDdmBlock) {
                const float s = P.sens[((int64_t)p * ns + P.inew2old[i]) * n + cell];
                sold[i] = fabsf(s) < 1.0e-25f ? 0.0f : s;
                b[i] = 0.0f;
            }
            __syncthreads();
            const int r1 = P.prate[p];
            if (r1 > 0 && t == 0) {          // rate parameter: h k p(YAVE) on reactants and products
                const DdmRxn x = P.rx[r1 - 1];
                const float kstep = (float)(P.rki[(int64_t)(r1 - 1) * n + cell] * h);
                const float rxm = (float)((double)kstep * ddm_pexcept(x, yav, -1));
                for (int q = 0; q < x.nre; ++q) b[x.re[q]] = b[x.re[q]] - rxm * 1.0f;
                for (int q = 0; q < x.npr; ++q) {
                    const int s = P.pid[x.poff + q];
                    b[s] = (float)((double)b[s] + (double)rxm * P.psc[x.poff + q] * 1.0);
                }
            }
            __syncthreads();
            for (int s = t; s < ns; s += kDdmBlock) {
                double total = 0.0;
                for (int j = 0; j < ns; ++j) {
                    float a1 = (float)(0.5 * h * (double)A[(int64_t)j * ns + s]);
                    if (j == s) a1 = 1.0f + a1;
                    total = total + (double)(a1 * sold[j]);
                }
                b[s] = (float)(total + (double)b[s]);
            }
            __syncthreads();
        }
        // ---- A = I - h/2 J in place
        for (int idx = t; idx < ns * ns; idx += kDdmBlock) {
            float a = (float)(-0.5 * h * (double)A[idx]);
            if (idx / ns == idx % ns) a = 1.0f + a;
            A[idx] = a;
        }
        __syncthreads();
        // ---- SGEFA: column k's pivot (first index of the largest magnitude),
        // swap, scaling by -1/pivot, then the column updates
        for (int k = 0; k < ns - 1; ++k) {
            const float *ak = A + (int64_t)k * ns;
            float bv = -1.0f;
            int bi = ns;
            for (int i = k + t; i < ns; i += kDdmBlock) {
                const float v = fabsf(ak[i]);
                if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const float ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bi, off);
                if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
            }
            if ((t & 63) == 0) { rv[t >> 6] = bv; ri[t >> 6] = bi; }
            __syncthreads();
            if (t == 0) {
                for (int w = 1; w < kDdmBlock / 64; ++w)
                    if (rv[w] > bv || (rv[w] == bv && ri[w] < bi)) { bv = rv[w]; bi = ri[w]; }
                // the sequential scan keeps k unless a later entry is strictly
                // larger than |a(k,k)|: same result, NaN aside
                const int l = (bi < ns && bv > fabsf(A[(int64_t)k * ns + k])) ? bi : k;
                ipvt[k] = l;
                s_piv = l;
                s_zero = A[(int64_t)k * ns + l] == 0.0f;
                if (!s_zero) {
                    float *akk = A + (int64_t)k * ns;
                    if (l != k) { const float tt = akk[l]; akk[l] = akk[k]; akk[k] = tt; }
                    s_scal = -1.0f / akk[k];
                }
            }
            __syncthreads();
            if (s_zero) continue;            // SGEFA records INFO and goes on (uniform branch)
            const int l = s_piv;
            const float sc = s_scal;
            float *akw = A + (int64_t)k * ns;
            for (int i = k + 1 + t; i < ns; i += kDdmBlock) akw[i] *= sc;
            for (int j = k + 1 + t; j < ns; j += kDdmBlock) {
                float *aj = A + (int64_t)j * ns;
                const float tj = aj[l];
                if (l != k) { aj[l] = aj[k]; aj[k] = tj; }
                tvec[j] = tj;
            }
            __syncthreads();
            const int m = ns - k - 1;
            for (int idx = t; idx < m * m; idx += kDdmBlock) {
                const int j = k + 1 + idx / m, i = k + 1 + idx % m;
                A[(int64_t)j * ns + i] += tvec[j] * akw[i];
            }
            __syncthreads();
        }
        if (t == 0) ipvt[ns - 1] = ns - 1;
        __syncthreads();
        // ---- SGESL (job = 0) for every parameter at once
        for (int k = 0; k < ns - 1; ++k) {
            const int l = ipvt[k];
            if (t < np) {
                float *b = rhs + t * ns;
                const float tt = b[l];
                if (l != k) { b[l] = b[k]; b[k] = tt; }
                tp[t] = tt;
            }
            __syncthreads();
            const float *ak = A + (int64_t)k * ns;
            const int m = ns - k - 1;
            for (int idx = t; idx < np * m; idx += kDdmBlock) {
                const int p = idx / m, i = k + 1 + idx % m;
                rhs[p * ns + i] += tp[p] * ak[i];
            }
            __syncthreads();
        }
        for (int k = ns - 1; k >= 0; --k) {
            const float *ak = A + (int64_t)k * ns;
            if (t < np) {
                float *b = rhs + t * ns;
                b[k] /= ak[k];
                tp[t] = -b[k];
            }
            __syncthreads();
            for (int idx = t; idx < np * k; idx += kDdmBlock) {
                const int p = idx / k, i = idx % k;
                rhs[p * ns + i] += tp[p] * ak[i];
            }
            __syncthreads();
        }
        for (int idx = t; idx < np * ns; idx += kDdmBlock) {
            const int p = idx / ns, s = idx % ns;
            const float v = rhs[idx];
            P.sens[((int64_t)p * ns + P.inew2old[s]) * n + cell] = fabsf(v) < 1.0e-25f ? 0.0f : v;
        }
        __syncthreads();
    }
}

// LDS bytes the kernel needs besides the matrix
inline size_t ddm_lds_extra(int ns, int np)
{
    return (size_t)ns * sizeof(double) + (size_t)np * ns * 4 + 3 * (size_t)ns * 4 + kDdmMaxParams * 4 +
           (kDdmBlock / 64) * 8 + 64;
}

}  // namespace ros3
