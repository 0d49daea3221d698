// ereb_sort.hip -- stable counting sort of marker indices by scatter tile.
//
// key(m) = tile of the basic cell (fix with the basic clamps, E:1456) of
// marker m; tiles are TILE x TILE basic cells numbered column-major.  The
// result perm[] lists marker indices grouped by tile in ascending tile
// order and, inside a tile, in ascending marker index (a stable sort).
//
//   1. k_tile_count : key per marker, per-tile counts (wavefront-aggregated
//                     integer atomics: counts do not depend on order)
//   2. scan         : tile_start = exclusive scan of the counts
//   3. k_tile_place : wavefront-aggregated slot reservation per tile
//   4. k_segsort_*  : each tile segment is sorted ascending in LDS (bitonic),
//                     which removes the arrival order of step 3 and makes the
//                     result deterministic.  Segments longer than 4096
//                     markers go to a 1024-thread variant (LDS up to 16384,
//                     global-memory bitonic beyond).
#include "ereb_internal.h"

namespace ereb {

// For every active lane: reserve one slot in counters[key]; returns the
// slot index.  One atomic per distinct key in the wavefront.
__device__ __forceinline__ int32_t wave_reserve(int32_t* counters, uint32_t key, bool active) {
    uint64_t todo = __ballot(active);
    int32_t result = 0;
    const int lane = __lane_id();
    while (todo) {
        int leader = __ffsll((unsigned long long)todo) - 1;
        uint32_t lk = __shfl(key, leader);
        bool mine = active && key == lk;
        uint64_t m = __ballot(mine);
        int32_t base = 0;
        if (lane == leader) base = atomicAdd(&counters[lk], (int32_t)__popcll(m));
        base = __shfl(base, leader);
        if (mine) result = base + (int32_t)__popcll(m & lanemask_lt());
        todo &= ~m;
    }
    return result;
}

__global__ __launch_bounds__(256) void k_tile_count(GeomK g, int nty, const double* __restrict__ xm,
                                                    const double* __restrict__ ym, int64_t M,
                                                    uint32_t* __restrict__ key,
                                                    int32_t* __restrict__ tcount,
                                                    int32_t* __restrict__ flags) {
    int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool active = m < M;
    uint32_t k = 0;
    if (active) {
        Cells c = marker_cells(g, xm[m], ym[m]);
        int di = c.ip - c.ib, dj = c.jp - c.jb;
        if (di < 0 || di > 1 || dj < 0 || dj > 1) flags[0] = 1;
        k = (uint32_t)(((c.jb - 1) / TILE) * nty + (c.ib - 1) / TILE);
        key[m] = k;
    }
    wave_reserve(tcount, k, active);
}

__global__ __launch_bounds__(256) void k_copy_i32(const int32_t* __restrict__ a,
                                                  int32_t* __restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_tile_place(const uint32_t* __restrict__ key, int64_t M,
                                                    int32_t* __restrict__ cursor,
                                                    uint32_t* __restrict__ perm) {
    int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool active = m < M;
    uint32_t k = active ? key[m] : 0u;
    int32_t pos = wave_reserve(cursor, k, active);
    if (active) perm[pos] = (uint32_t)m;
}

// bitonic sort (ascending) of s[0..n), n a power of two, by nthreads threads
__device__ __forceinline__ void bitonic_lds(uint32_t* s, int n) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    uint32_t a = s[i], b = s[ixj];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

constexpr int SEG_SMALL = 4096;
constexpr int SEG_BIG = 16384;  // 64 KB of LDS

__global__ __launch_bounds__(256) void k_segsort_small(const int32_t* __restrict__ tstart,
                                                       uint32_t* __restrict__ perm,
                                                       int32_t* __restrict__ big) {
    __shared__ uint32_t s[SEG_SMALL];
    const int t = blockIdx.x;
    const int32_t a = tstart[t], L = tstart[t + 1] - a;
    if (L <= 1) return;
    if (L > SEG_SMALL) {
        if (threadIdx.x == 0) {
            int slot = atomicAdd(&big[0], 1);
            big[1 + slot] = t;
        }
        return;
    }
    int n = 1;
    while (n < L) n <<= 1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = i < L ? perm[a + i] : 0xFFFFFFFFu;
    __syncthreads();
    bitonic_lds(s, n);
    for (int i = threadIdx.x; i < L; i += blockDim.x) perm[a + i] = s[i];
}

__global__ __launch_bounds__(1024) void k_segsort_big(const int32_t* __restrict__ tstart,
                                                      uint32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ big,
                                                      uint32_t* __restrict__ tmp) {
    extern __shared__ uint32_t sdyn[];
    const int nbig = big[0];
    for (int e = blockIdx.x; e < nbig; e += gridDim.x) {
        const int t = big[1 + e];
        const int32_t a = tstart[t], L = tstart[t + 1] - a;
        int n = 1;
        while (n < L) n <<= 1;
        // LDS when it fits, else a scratch region [2a, 2a + n) of tmp, which
        // is disjoint between tiles because n < 2L
        uint32_t* s = n <= SEG_BIG ? sdyn : tmp + 2 * (int64_t)a;
        for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = i < L ? perm[a + i] : 0xFFFFFFFFu;
        __syncthreads();
        bitonic_lds(s, n);
        for (int i = threadIdx.x; i < L; i += blockDim.x) perm[a + i] = s[i];
        __syncthreads();
    }
}

int launch_tile_sort(ereb_ctx* c) {
    const int64_t M = c->marknum;
    int rc;
    hipEvent_t ev;
    prof_begin(c, EREB_K_SORT, &ev);
    if ((rc = fill_i32(c, c->d_tcount, 0, c->ntiles))) return rc;
    if ((rc = fill_i32(c, c->d_flags, 0, 4))) return rc;
    if ((rc = fill_i32(c, c->d_big, 0, 1))) return rc;
    GeomK g = geomk(c);
    if (M > 0) {
        unsigned nb = (unsigned)((M + 255) / 256);
        hipLaunchKernelGGL(k_tile_count, dim3(nb), dim3(256), 0, c->stream, g, c->nty,
                           mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), M, c->d_key, c->d_tcount,
                           c->d_flags);
    }
    if ((rc = scan_exclusive(c, c->d_tcount, c->d_tstart, c->ntiles))) return rc;
    if (M > 0) {
        unsigned nt = (unsigned)((c->ntiles + 255) / 256);
        hipLaunchKernelGGL(k_copy_i32, dim3(nt), dim3(256), 0, c->stream, c->d_tstart,
                           c->d_tcursor, (int64_t)c->ntiles);
        unsigned nb = (unsigned)((M + 255) / 256);
        hipLaunchKernelGGL(k_tile_place, dim3(nb), dim3(256), 0, c->stream, c->d_key, M,
                           c->d_tcursor, c->d_perm);
        hipLaunchKernelGGL(k_segsort_small, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream,
                           c->d_tstart, c->d_perm, c->d_big);
        hipLaunchKernelGGL(k_segsort_big, dim3(64), dim3(1024), SEG_BIG * sizeof(uint32_t),
                           c->stream, c->d_tstart, c->d_perm, c->d_big, c->d_tmp);
    }
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_SORT, ev);
    c->have_sort = true;
    return EREB_OK;
}

}  // namespace ereb
