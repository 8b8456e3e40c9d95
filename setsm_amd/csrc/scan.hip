// scan.hip — exclusive prefix sums over per-cell counts (CSR offsets of the
// voxel cube, of the height-run work list and of the per-tile run lists).
//
// Three kernels and no library call: a chunk-sum pass, one workgroup scanning
// the chunk sums, and a pass that rescans each chunk from its offset. Nothing
// in it allocates, clears memory with a runtime call or reads back to the
// host, so the sequence can be captured in a hipGraph and driven from a
// caller-provided partials buffer.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smx {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr long kChunk = (long)kThreads * kPerThread;   // elements per workgroup
constexpr int kTopThreads = 1024;
constexpr int kTopPerThread = 64;
constexpr long kMaxChunks = (long)kTopThreads * kTopPerThread;

// Exclusive scan of one value per thread across a workgroup of kT threads;
// returns the workgroup total in `total`.
template <int kT>
__device__ __forceinline__ int64_t block_exclusive(int64_t v, int64_t *lds_waves, int64_t &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0;
        for (int w = 0; w < kT / 64; w++) {
            const int64_t t = lds_waves[w];
            lds_waves[w] = s;
            s += t;
        }
        lds_waves[kT / 64] = s;
    }
    __syncthreads();
    total = lds_waves[kT / 64];
    const int64_t r = lds_waves[wave] + x - v;
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kThreads) k_chunk_sums(const int64_t *__restrict__ in, long n,
                                                         int64_t *__restrict__ part)
{
    __shared__ int64_t w[kThreads / 64 + 1];
    const long base = (long)blockIdx.x * kChunk;
    int64_t s = 0;
    for (int k = 0; k < kPerThread; k++) {
        const long i = base + (long)k * kThreads + threadIdx.x;   // coalesced
        if (i < n) s += in[i];
    }
    int64_t total;
    (void)block_exclusive<kThreads>(s, w, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kTopThreads) k_scan_chunks(int64_t *__restrict__ part, long nchunks)
{
    __shared__ int64_t w[kTopThreads / 64 + 1];
    const long b = (long)threadIdx.x * kTopPerThread;
    int64_t s = 0;
    for (int k = 0; k < kTopPerThread; k++)
        if (b + k < nchunks) s += part[b + k];
    int64_t total;
    int64_t run = block_exclusive<kTopThreads>(s, w, total);
    for (int k = 0; k < kTopPerThread; k++)
        if (b + k < nchunks) {
            const int64_t t = part[b + k];
            part[b + k] = run;
            run += t;
        }
}

// LDS slot of chunk element e: one pad word per 16 elements, so the threads
// of a wave walking their 16 consecutive elements hit different banks.
__device__ __forceinline__ int padded(int e) { return e + (e >> 4); }

__global__ void __launch_bounds__(kThreads) k_chunk_scan(const int64_t *__restrict__ in, long n,
                                                         const int64_t *__restrict__ part, int64_t *__restrict__ out)
{
    __shared__ int64_t tile[kChunk + kChunk / 16];
    __shared__ int64_t w[kThreads / 64 + 1];
    const long base = (long)blockIdx.x * kChunk;
    for (int k = 0; k < kPerThread; k++) {
        const int e = k * kThreads + threadIdx.x;
        const long i = base + e;
        tile[padded(e)] = i < n ? in[i] : 0;
    }
    __syncthreads();
    // each thread scans its kPerThread consecutive elements
    const int e0 = threadIdx.x * kPerThread;
    int64_t s = 0;
    for (int k = 0; k < kPerThread; k++) s += tile[padded(e0 + k)];
    int64_t total;
    int64_t run = part[blockIdx.x] + block_exclusive<kThreads>(s, w, total);
    for (int k = 0; k < kPerThread; k++) {
        const int64_t t = tile[padded(e0 + k)];
        tile[padded(e0 + k)] = run;
        run += t;
    }
    __syncthreads();
    for (int k = 0; k < kPerThread; k++) {
        const int e = k * kThreads + threadIdx.x;
        const long i = base + e;
        if (i < n) out[i] = tile[padded(e)];
    }
}

}  // namespace

// int64 slots of the partials buffer scan_i64 needs for n values.
long scan_partials(long n) { return (n + kChunk - 1) / kChunk + 1; }

// Largest n scan_i64 accepts.
long scan_max_n() { return kMaxChunks * kChunk; }

// out[i] = in[0] + ... + in[i-1] for i < n (out may not alias in). `part`
// holds scan_partials(n) int64 values of scratch.
hipError_t scan_i64(const int64_t *in, int64_t *out, long n, int64_t *part, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const long nchunks = (n + kChunk - 1) / kChunk;
    if (nchunks > kMaxChunks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_chunk_sums, dim3((unsigned)nchunks), dim3(kThreads), 0, s, in, n, part);
    hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(kTopThreads), 0, s, part, nchunks);
    hipLaunchKernelGGL(k_chunk_scan, dim3((unsigned)nchunks), dim3(kThreads), 0, s, in, n, part, out);
    return hipGetLastError();
}

}  // namespace smx
