// copy_probe.hip -- calibration for rocprofv3 memory counters: a plain copy
// kernel over a buffer of known size (default 1 GiB read + 1 GiB written per
// launch), so FETCH_SIZE / WRITE_SIZE readings of the solver's kernels can be
// scaled to bytes.  `copy_probe ceiling` instead measures the best HBM
// bandwidth a copy reaches on this part: one 16-byte element per thread over
// 4 GiB, and a stream of 15 arrays in and 15 out shaped like the Poisson
// iteration (each thread reads one double of each input array and writes one
// double of each output array).  Not part of the solver library.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ in,
                                                   float4* __restrict__ out, size_t n)
{
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        out[i] = in[i];
}

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e = (x);                                                        \
        if (e != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));            \
            return 1;                                                              \
        }                                                                          \
    } while (0)

__global__ __launch_bounds__(256) void copy_once(const float4* __restrict__ in, float4* __restrict__ out,
                                                 size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// 15 SoA arrays of n doubles in, 15 out (+1 when shift: the Poisson push's
// stores, each array's element moved by its own small offset, wrapped).
template <int SHIFT>
__global__ __launch_bounds__(256) void copy_soa15(const double* __restrict__ in, double* __restrict__ out,
                                                  size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) v[q] = in[q * n + i];
#pragma unroll
    for (int q = 0; q < 15; ++q) {
        size_t j = i;
        if (SHIFT) {
            const size_t off = q == 0 ? 0 : (q & 1 ? 1 : n - 1);  // +1 or -1 cell
            j = (i + off) % n;
        }
        out[q * n + j] = v[q];
    }
}

// The Poisson push's store pattern on an lx x ly x lz grid: array q's element
// moves by c_qx + c_qy lx + c_qz lx ly cells (D3Q15), wrapped linearly.
__constant__ int D3Q15_C[15][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},   {0, 0, 1},   {-1, 0, 0},
                                   {0, -1, 0}, {0, 0, -1}, {1, 1, 1},   {-1, 1, 1},  {1, -1, 1},
                                   {1, 1, -1}, {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};

template <typename T>
__global__ __launch_bounds__(256) void copy_d3q15(const T* __restrict__ in, T* __restrict__ out,
                                                  long long lx, long long ly, long long n)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) v[q] = in[q * n + i];
#pragma unroll
    for (int q = 0; q < 15; ++q) {
        const long long off = D3Q15_C[q][0] + D3Q15_C[q][1] * lx + D3Q15_C[q][2] * lx * ly;
        long long j = i + off;
        j = j < 0 ? j + n : (j >= n ? j - n : j);
        out[q * n + j] = v[q];
    }
}

int ceiling()
{
    const size_t bytes = size_t(4) << 30;
    const size_t n4 = bytes / sizeof(float4);
    const size_t nd = (size_t(256) * 256 * 256);  // 256^3 cells, 15 arrays in and out
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    double *h0 = nullptr, *h1 = nullptr;
    CHECK(hipMalloc(&h0, 15 * nd * sizeof(double)));
    CHECK(hipMalloc(&h1, 15 * nd * sizeof(double)));
    CHECK(hipMemset(h0, 0, 15 * nd * sizeof(double)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    auto timed = [&](const char* name, double moved, auto launch) -> int {
        float best = 1e30f;
        for (int it = 0; it < 6; ++it) {
            CHECK(hipEventRecord(e0));
            launch();
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (it > 0 && ms < best) best = ms;
        }
        std::printf("%s: %.3f ms, %.1f GB/s\n", name, best, moved / (best * 1e-3) / 1e9);
        return 0;
    };
    const unsigned g4 = static_cast<unsigned>((n4 + 255) / 256), gd = static_cast<unsigned>((nd + 255) / 256);
    if (timed("copy 4 GiB, one float4 per thread", 2.0 * bytes,
              [&] { copy_once<<<g4, 256>>>(a, b, n4); })) return 1;
    if (timed("copy 4 GiB, grid-stride 8192 blocks", 2.0 * bytes,
              [&] { copy_kernel<<<8192, 256>>>(a, b, n4); })) return 1;
    if (timed("15 SoA arrays in, 15 out, 256^3 doubles", 2.0 * 15 * nd * sizeof(double),
              [&] { copy_soa15<0><<<gd, 256>>>(h0, h1, nd); })) return 1;
    if (timed("15 SoA arrays in, 15 out shifted by +-1 cell", 2.0 * 15 * nd * sizeof(double),
              [&] { copy_soa15<1><<<gd, 256>>>(h0, h1, nd); })) return 1;
    const long long shapes[5][3] = {{256, 256, 256}, {256, 256, 216}, {192, 192, 384}, {384, 192, 192},
                                    {256, 128, 128}};
    for (const auto& sh : shapes) {
        const long long n = sh[0] * sh[1] * sh[2];
        char name[96];
        std::snprintf(name, sizeof(name), "D3Q15 push pattern %lldx%lldx%lld", sh[0], sh[1], sh[2]);
        if (timed(name, 2.0 * 15 * n * sizeof(double), [&] {
                copy_d3q15<double><<<static_cast<unsigned>((n + 255) / 256), 256>>>(h0, h1, sh[0], sh[1], n);
            }))
            return 1;
    }
    // The same pattern on 4-byte values (the mixed-precision iteration's h).
    for (const auto& sh : {shapes[0], shapes[2], shapes[4]}) {
        const long long n = sh[0] * sh[1] * sh[2];
        char name[96];
        std::snprintf(name, sizeof(name), "D3Q15 push pattern FP32 %lldx%lldx%lld", sh[0], sh[1], sh[2]);
        if (timed(name, 2.0 * 15 * n * sizeof(float), [&] {
                copy_d3q15<float><<<static_cast<unsigned>((n + 255) / 256), 256>>>(
                    reinterpret_cast<const float*>(h0), reinterpret_cast<float*>(h1), sh[0], sh[1], n);
            }))
            return 1;
    }
    CHECK(hipGetLastError());
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(h0));
    CHECK(hipFree(h1));
    return 0;
}

int main(int argc, char** argv)
{
    if (argc > 1 && std::string(argv[1]) == "ceiling") return ceiling();
    const size_t bytes = (argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024ull) << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 1, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int it = 0; it < 4; ++it) {
        CHECK(hipEventRecord(e0));
        copy_kernel<<<8192, 256>>>(a, b, n);
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("copy %zu bytes: %.3f ms, %.1f GB/s (read+write)\n", bytes, ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
