// ereb_scan.hip -- device-wide exclusive scan of int32 counts and fill
// kernels (fills replace hipMemsetAsync so that a captured step graph holds
// only kernel nodes).
#include "ereb_internal.h"

namespace ereb {

constexpr int SCAN_T = 1024;        // threads per block
constexpr int SCAN_PER = 4;         // items per thread
constexpr int SCAN_B = SCAN_T * SCAN_PER;

__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* sh, int32_t* total) {
    // sh: SCAN_T/64 + 1 ints
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int nw = blockDim.x >> 6;
        int32_t s = lane < nw ? sh[lane] : 0;
        int32_t t = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int32_t y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < nw) sh[lane] = t - s;  // exclusive wave offsets
        if (lane == nw - 1) sh[SCAN_T / 64] = t;
    }
    __syncthreads();
    int32_t r = sh[wid] + x - v;
    *total = sh[SCAN_T / 64];
    __syncthreads();
    return r;
}

// sum of each block's SCAN_B inputs
// run (optional): the scan runs only when run[0] != 0 (a device-side
// condition, so that a captured graph can skip it)
__global__ __launch_bounds__(SCAN_T) void k_scan_blocksum(const int32_t* __restrict__ in,
                                                          int64_t n, int32_t* __restrict__ bs,
                                                          const int32_t* __restrict__ run) {
    __shared__ int32_t sh[SCAN_T / 64 + 1];
    if (run && run[0] == 0) return;
    int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_PER;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
        if (base + k < n) s += in[base + k];
    int32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) bs[blockIdx.x] = tot;
}

// out[i] = exclusive prefix of in, the block's offset being the sum of the
// block sums before it (each block adds them up itself); the last block also
// writes the total to out[n].  Optionally also out2[i] = out[i], and in[i] = 0
// once read (for inputs that must start from zero next time).
__global__ __launch_bounds__(SCAN_T) void k_scan_apply(int32_t* __restrict__ in, int64_t n,
                                                       const int32_t* __restrict__ bs,
                                                       int32_t* __restrict__ out,
                                                       int32_t* __restrict__ out2, int zero_in,
                                                       const int32_t* __restrict__ run) {
    __shared__ int32_t sh[SCAN_T / 64 + 1];
    if (run && run[0] == 0) return;
    // the block's inputs are loaded first, so that their latency overlaps the
    // reduction of the block sums before this block
    int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_PER;
    int32_t v[SCAN_PER];
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) v[k] = base + k < n ? in[base + k] : 0;
    int32_t off = 0, tot;
    for (int64_t b0 = 0; b0 < blockIdx.x; b0 += SCAN_T) {
        const int64_t b = b0 + threadIdx.x;
        block_exclusive_scan(b < blockIdx.x ? bs[b] : 0, sh, &tot);
        off += tot;
    }
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) s += v[k];
    int32_t e = block_exclusive_scan(s, sh, &tot) + off;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + tot;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        if (base + k < n) {
            out[base + k] = e;
            if (out2) out2[base + k] = e;
            if (zero_in) in[base + k] = 0;
        }
        e += v[k];
    }
}

int64_t scan_blocks(int64_t n) { return std::max<int64_t>((n + SCAN_B - 1) / SCAN_B, 1); }

int scan_exclusive_buf(ereb_ctx* c, const int32_t* in, int32_t* out, int64_t n,
                       int32_t* blocksums, int32_t* out2, bool zero_in, const int32_t* run) {
    const int64_t nb = scan_blocks(n);
    hipLaunchKernelGGL(k_scan_blocksum, dim3((unsigned)nb), dim3(SCAN_T), 0, c->stream, in, n,
                       blocksums, run);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_T), 0, c->stream,
                       const_cast<int32_t*>(in), n, (const int32_t*)blocksums, out, out2,
                       zero_in ? 1 : 0, run);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

int scan_exclusive(ereb_ctx* c, const int32_t* in, int32_t* out, int64_t n, int32_t* out2,
                   bool zero_in, const int32_t* run) {
    int rc = ensure_scan(c, n);
    if (rc) return rc;
    return scan_exclusive_buf(c, in, out, n, c->d_scan, out2, zero_in, run);
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int fill_i32(ereb_ctx* c, int32_t* p, int32_t v, int64_t n) {
    if (n <= 0) return EREB_OK;
    int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<int32_t>, dim3((unsigned)nb), dim3(256), 0, c->stream, p, v, n);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

int fill_f64(ereb_ctx* c, double* p, double v, int64_t n) {
    if (n <= 0) return EREB_OK;
    int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)nb), dim3(256), 0, c->stream, p, v, n);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

}  // namespace ereb
