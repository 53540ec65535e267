// gather_probe.hip - the cost of random bound gathers on this GPU.
//
// A propagation sweep reads, per nonzero, a 4-byte column index and an
// 8-byte coefficient (a sequential stream) and gathers the bounds of that
// column from a table of records (random access).  This program times those
// two access patterns alone, without any propagation arithmetic: N entries,
// each with a streamed index and value, and one gather per entry from a table
// of `table_mb` megabytes with uniformly random indices.  K entries per lane
// are loaded before they are consumed (the sweep's lanes take 4), and the
// gather is 16, 8 or 4 bytes wide.  The result is the floor a sweep of N
// nonzeros cannot go below with its gathers served by the cache hierarchy.
//
// usage: gather_probe [entries] [table_mb ...]     (defaults: 19300000, 0.0625 1 3.8 8 32)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

namespace {

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

typedef unsigned u4 __attribute__((ext_vector_type(4)));
typedef unsigned u2 __attribute__((ext_vector_type(2)));

template <int K, int W>
__global__ __launch_bounds__(256) void probe_kernel(long long n, const int *__restrict__ idx,
                                                    const double *__restrict__ val, const char *__restrict__ table,
                                                    double *__restrict__ out) {
    const long long tid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    double acc = 0.0;
    for (long long base = tid * K; base < n; base += stride * K) {
        int ix[K];
        double v[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const long long e = base + k;
            ix[k] = e < n ? idx[e] : 0;
            v[k] = e < n ? val[e] : 0.0;
        }
        if constexpr (W == 0) {
#pragma unroll
            for (int k = 0; k < K; k++) acc += v[k] * static_cast<double>(ix[k]);
        } else {
            double g[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const char *p = table + static_cast<size_t>(ix[k]) * 16;
                if constexpr (W == 16) {
                    const u4 c = *reinterpret_cast<const u4 *>(p);
                    g[k] = static_cast<double>(c.x ^ c.y ^ c.z ^ c.w);
                } else if constexpr (W == 8) {
                    const u2 c = *reinterpret_cast<const u2 *>(p);
                    g[k] = static_cast<double>(c.x ^ c.y);
                } else {
                    g[k] = static_cast<double>(*reinterpret_cast<const unsigned *>(p));
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) acc += v[k] * g[k];
        }
    }
    out[tid] = acc;
}

}  // namespace

int main(int argc, char **argv) {
    const long long n = argc > 1 ? std::atoll(argv[1]) : 19300000LL;
    std::vector<double> sizes;
    for (int a = 2; a < argc; a++) sizes.push_back(std::atof(argv[a]));
    if (sizes.empty()) sizes = {0.0625, 1.0, 3.8, 8.0, 32.0};
    if (n < 1) return 1;
    const int grid = 256 * 32, threads = 256;
    int *d_idx = nullptr;
    double *d_val = nullptr, *d_out = nullptr;
    char *d_table = nullptr;
    double max_mb = 0.0;
    for (double s : sizes) max_mb = s > max_mb ? s : max_mb;
    const long long max_records = static_cast<long long>(max_mb * 1048576.0 / 16.0) + 1;
    check(hipMalloc(&d_idx, sizeof(int) * n), "hipMalloc");
    check(hipMalloc(&d_val, sizeof(double) * n), "hipMalloc");
    check(hipMalloc(&d_table, 16 * max_records), "hipMalloc");
    check(hipMalloc(&d_out, sizeof(double) * grid * threads), "hipMalloc");
    check(hipMemset(d_val, 0, sizeof(double) * n), "hipMemset");
    check(hipMemset(d_table, 0, 16 * max_records), "hipMemset");
    hipEvent_t a, b;
    check(hipEventCreate(&a), "event");
    check(hipEventCreate(&b), "event");
    std::vector<int> h_idx(n);
    std::mt19937 rng(1);
    std::printf("entries %lld; time per pass (best of 5 after one warmup) and gathers per second\n", n);
    for (double mb : sizes) {
        const long long records = static_cast<long long>(mb * 1048576.0 / 16.0);
        if (records < 1) continue;
        std::uniform_int_distribution<int> dist(0, static_cast<int>(records) - 1);
        for (long long e = 0; e < n; e++) h_idx[e] = dist(rng);
        check(hipMemcpy(d_idx, h_idx.data(), sizeof(int) * n, hipMemcpyHostToDevice), "H2D");
        auto run = [&](const char *name, auto kernel) {
            float best = 1e30f;
            for (int it = 0; it < 6; it++) {
                check(hipEventRecord(a), "record");
                hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, 0, n, d_idx, d_val, d_table, d_out);
                check(hipEventRecord(b), "record");
                check(hipEventSynchronize(b), "sync");
                float ms = 0.0f;
                check(hipEventElapsedTime(&ms, a, b), "elapsed");
                if (it > 0 && ms < best) best = ms;
            }
            std::printf("table %8.4f MB  %-16s %9.1f us  %7.1f G/s\n", mb, name, best * 1000.0, n / (best * 1e6));
            std::fflush(stdout);
        };
        run("stream only", probe_kernel<4, 0>);
        run("16 B, K=1", probe_kernel<1, 16>);
        run("16 B, K=2", probe_kernel<2, 16>);
        run("16 B, K=4", probe_kernel<4, 16>);
        run("8 B, K=4", probe_kernel<4, 8>);
        run("4 B, K=4", probe_kernel<4, 4>);
        run("4 B, K=1", probe_kernel<1, 4>);
    }
    check(hipFree(d_idx), "free");
    check(hipFree(d_val), "free");
    check(hipFree(d_table), "free");
    check(hipFree(d_out), "free");
    return 0;
}
