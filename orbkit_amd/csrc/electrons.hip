// Integrated electron count sum_p rho[p] * d3r in FP64 with compensation
// (orbkit prints numpy.sum(rho) * d3r, orbkit/core.py:590-592).  Each thread
// keeps a Neumaier (sum, compensation) pair over a grid-stride range; pairs
// are merged with Knuth's TwoSum in a fixed tree order within a workgroup,
// per-workgroup pairs go to the workspace and one workgroup merges them in
// order.  The result is deterministic for a given launch shape.
#include "okmi_internal.h"

namespace {

constexpr int EC_BLOCKS = 512;
constexpr int EC_THREADS = 256;

struct Pair {
    double s, c;
};

__device__ __forceinline__ Pair two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return Pair{s, e};
}

__device__ __forceinline__ Pair merge(Pair x, Pair y)
{
    const Pair t = two_sum(x.s, y.s);
    return Pair{t.s, t.c + x.c + y.c};
}

__device__ Pair block_merge(Pair v)
{
    __shared__ double ss[EC_THREADS], cc[EC_THREADS];
    ss[threadIdx.x] = v.s;
    cc[threadIdx.x] = v.c;
    __syncthreads();
    for (int o = EC_THREADS / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const Pair m = merge(Pair{ss[threadIdx.x], cc[threadIdx.x]}, Pair{ss[threadIdx.x + o], cc[threadIdx.x + o]});
            ss[threadIdx.x] = m.s;
            cc[threadIdx.x] = m.c;
        }
        __syncthreads();
    }
    return Pair{ss[0], cc[0]};
}

__global__ __launch_bounds__(EC_THREADS) void k_electrons_part(const double* rho, int64_t n, double* part)
{
    Pair acc{0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * EC_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * EC_THREADS) {
        const Pair t = two_sum(acc.s, rho[i]);
        acc = Pair{t.s, acc.c + t.c};
    }
    const Pair b = block_merge(acc);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = b.s;
        part[2 * blockIdx.x + 1] = b.c;
    }
}

__global__ __launch_bounds__(EC_THREADS) void k_electrons_final(const double* part, int nb, double d3r, double* out)
{
    Pair acc{0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += EC_THREADS) acc = merge(acc, Pair{part[2 * i], part[2 * i + 1]});
    const Pair b = block_merge(acc);
    if (threadIdx.x == 0) out[0] = (b.s + b.c) * d3r;
}

}  // namespace

size_t okmi_electrons_ws_bytes() { return sizeof(double) * 2 * EC_BLOCKS; }

int okmi_launch_electrons(const double* d_rho, int64_t n, double d3r, void* ws, size_t ws_bytes, double* d_out,
                          hipStream_t st)
{
    if (!ws || ws_bytes < okmi_electrons_ws_bytes()) {
        okmi_set_error("electron_count: workspace of %zu bytes is too small (need %zu)", ws_bytes,
                       okmi_electrons_ws_bytes());
        return 1;
    }
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(EC_BLOCKS, (n + EC_THREADS - 1) / EC_THREADS));
    hipLaunchKernelGGL(k_electrons_part, dim3(nb), dim3(EC_THREADS), 0, st, d_rho, n, (double*)ws);
    hipLaunchKernelGGL(k_electrons_final, dim3(1), dim3(EC_THREADS), 0, st, (const double*)ws, nb, d3r, d_out);
    return okmi_check_launch("electron_count");
}
