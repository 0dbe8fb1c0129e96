// hbm_rows: read throughput of the solver's memory access stream, without
// the solver.  Every resident wave owns a private region like its solver
// workspace (2.75 MB) and reads whole lane-rows at pseudo-random row indices:
// 64 lanes x 8 B = 512 B per load instruction (the workspace rows), or
// 64 x 16 B = 1 KB.  Sixteen independent loads are kept in flight per wave.
// The result is the practical HBM ceiling for that stream, against which the
// solve kernel's measured traffic is compared.
//
//     hbm_rows [iterations]      prints one line per (waves per CU, width)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kWave = 64;
constexpr int kUnroll = 16;
constexpr long kRowsPerWave = 2688;   // rows of 1 KB: 2.75 MB per wave

__device__ inline unsigned mix32(unsigned x)
{
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

template <int W>   // doubles per lane per load
__global__ __launch_bounds__(256) void row_reads(const double *__restrict__ buf, double *__restrict__ out, int iters)
{
    const int lane = threadIdx.x & (kWave - 1);
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const double *base = buf + wave * kRowsPerWave * kWave * 2;
    const unsigned rows = (unsigned)(kRowsPerWave * 2 / W);
    double acc = 0.0;
    for (int it = 0; it < iters; ++it) {
        double v[kUnroll * W];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const unsigned r = mix32((unsigned)(it * kUnroll + u) * 2654435761U ^ (unsigned)wave) % rows;
            const double *p = base + ((long)r * kWave + lane) * W;
            if (W == 1) {
                v[u] = p[0];
            } else {
                const double2 t = *(const double2 *)p;
                v[2 * u] = t.x;
                v[2 * u + 1] = t.y;
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll * W; ++u) acc += v[u];
    }
    out[wave * kWave + lane] = acc;
}

#define CHECK(e) do { hipError_t _e = (e); if (_e != hipSuccess) { std::fprintf(stderr, "%s: %s\n", #e, hipGetErrorString(_e)); return 1; } } while (0)

}  // namespace

int main(int argc, char **argv)
{
    const int iters = argc > 1 ? std::atoi(argv[1]) : 4000;
    CHECK(hipSetDevice(0));
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    for (int wpc : {1, 2, 4, 8, 16}) {
        const long waves = (long)cus * wpc;
        const size_t bytes = (size_t)waves * kRowsPerWave * kWave * 2 * sizeof(double);
        double *buf = nullptr, *out = nullptr;
        CHECK(hipMalloc(&buf, bytes));
        CHECK(hipMalloc(&out, (size_t)waves * kWave * sizeof(double)));
        CHECK(hipMemset(buf, 0, bytes));
        hipEvent_t a, b;
        CHECK(hipEventCreate(&a));
        CHECK(hipEventCreate(&b));
        for (int w : {1, 2}) {
            const int it = w == 1 ? iters : iters / 2;      // same bytes for both widths
            float ms = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                CHECK(hipEventRecord(a));
                if (w == 1) hipLaunchKernelGGL(row_reads<1>, dim3(waves / 4), dim3(256), 0, 0, buf, out, it);
                else hipLaunchKernelGGL(row_reads<2>, dim3(waves / 4), dim3(256), 0, 0, buf, out, it);
                CHECK(hipGetLastError());
                CHECK(hipEventRecord(b));
                CHECK(hipEventSynchronize(b));
                CHECK(hipEventElapsedTime(&ms, a, b));
            }
            const double moved = (double)waves * it * kUnroll * kWave * 8.0 * w;
            std::printf("waves_per_cu %d bytes_per_lane %d ms %.2f GBs %.0f\n", wpc, 8 * w, ms, moved / ms / 1e6);
        }
        CHECK(hipEventDestroy(a));
        CHECK(hipEventDestroy(b));
        CHECK(hipFree(buf));
        CHECK(hipFree(out));
    }
    return 0;
}
