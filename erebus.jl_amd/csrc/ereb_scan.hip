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

// Block sums; the last block to finish (counted on done[0], which it resets
// to zero for the next scan) then scans the nb block sums in place and
// writes the total to out[n].  No block waits for another.
__global__ __launch_bounds__(SCAN_T) void k_scan_blocksum(const int32_t* __restrict__ in,
                                                          int64_t n, int32_t* __restrict__ bs,
                                                          int64_t nb, int32_t* __restrict__ out,
                                                          int32_t* __restrict__ done) {
    __shared__ int32_t sh[SCAN_T / 64 + 1];
    __shared__ int32_t last;
    int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_PER;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
        if (base + k < n) s += in[base + k];
    int32_t tot;
    block_exclusive_scan(s, sh, &tot);
    if (threadIdx.x == 0) {
        bs[blockIdx.x] = tot;
        __threadfence();  // the sum is visible before this block is counted
        last = atomicAdd(done, 1) == (int32_t)nb - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // every block's sum is visible to the last block
    int32_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += SCAN_T) {
        const int64_t i = b0 + threadIdx.x;
        const int32_t v = i < nb ? __atomic_load_n(&bs[i], __ATOMIC_RELAXED) : 0;
        int32_t t2;
        const int32_t e = block_exclusive_scan(v, sh, &t2);
        if (i < nb) bs[i] = carry + e;
        carry += t2;
    }
    if (threadIdx.x == 0) {
        out[n] = carry;
        *done = 0;
    }
}

// out[i] = exclusive prefix of in; optionally also out2[i] = out[i], and
// in[i] = 0 once read (for inputs that must start from zero next time)
__global__ __launch_bounds__(SCAN_T) void k_scan_apply(int32_t* __restrict__ in, int64_t n,
                                                       const int32_t* __restrict__ bs,
                                                       int32_t* __restrict__ out,
                                                       int32_t* __restrict__ out2, int zero_in) {
    __shared__ int32_t sh[SCAN_T / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_B + (int64_t)threadIdx.x * SCAN_PER;
    int32_t v[SCAN_PER];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int32_t tot;
    int32_t e = block_exclusive_scan(s, sh, &tot) + bs[blockIdx.x];
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

int scan_exclusive(ereb_ctx* c, const int32_t* in, int32_t* out, int64_t n, int32_t* out2,
                   bool zero_in) {
    int rc = ensure_scan(c, n);
    if (rc) return rc;
    int64_t nb = (n + SCAN_B - 1) / SCAN_B;
    if (nb == 0) nb = 1;
    int32_t* done = c->d_scan + c->scan_cap;  // last-block counter, zero between scans
    hipLaunchKernelGGL(k_scan_blocksum, dim3((unsigned)nb), dim3(SCAN_T), 0, c->stream, in, n,
                       c->d_scan, nb, out, done);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_T), 0, c->stream,
                       const_cast<int32_t*>(in), n, (const int32_t*)c->d_scan, out, out2,
                       zero_in ? 1 : 0);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
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
