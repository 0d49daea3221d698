// ereb_sort.hip -- stable counting sort of marker indices by fine key.
//
// The fine key of a marker is its "fine cell": the basic cell (ib, jb)
// (fix with the basic clamps, E:1456) refined by the P cell offsets
// (ip-ib, jp-jb) in {0,1}^2, numbered tile-major: tiles of TILE x TILE
// basic cells in column-major order, then FINE x FINE fine cells inside a
// tile, column-major.  All markers of one fine cell share the same four
// stencils on every grid.  The result perm[] lists marker indices in
// ascending key and, inside a key, ascending marker index -- exactly a
// stable counting sort, which makes every later summation order canonical.
//
//   1. k_fine_count : key per marker and its rank among the markers of its
//                     key in arrival order (one returning integer atomic per
//                     run of equal keys in a wavefront: the counts do not
//                     depend on order); before a scatter also the marker
//                     properties
//   2. scan         : key starts
//   3. k_fine_place : perm[start of the key + rank] = marker, no atomics
//   4. k_fine_order : each key segment (about 9 markers) is sorted by marker
//                     index in place (in registers up to 16 markers),
//                     removing the arrival order of step 3;
//                     segments longer than SEG_THREAD go to k_fine_order_big
//                     (bitonic in LDS up to 16384, global memory beyond).
#include "ereb_internal.h"

namespace ereb {

// Lanes of a wavefront grouped into runs: maximal stretches of consecutive
// active lanes with equal keys (markers in storage order come in runs of a
// few per fine cell).  The first lane of a run is its leader; every lane
// learns its leader and its rank in the run, a leader the run length.  A key
// split over several runs gets one count per run, which the counts and the
// slot reservations below allow (k_fine_order fixes the order within a key
// afterwards).  Two ballots and a shuffle, no loop.
struct KeyGroup {
    int leader, rank, count;  // count is non-zero on leaders only
};
__device__ __forceinline__ KeyGroup wave_group(uint32_t key, bool active) {
    const int lane = __lane_id();
    const uint32_t prev = __shfl_up(key, 1, 64);
    const uint64_t act = __ballot(active);
    const bool prev_active = lane > 0 && ((act >> (lane - 1)) & 1ull);
    const bool head = active && (!prev_active || prev != key);
    const uint64_t heads = __ballot(head);
    const uint64_t le = lanemask_lt() | (1ull << lane);
    // my run's leader: the highest head at or below this lane
    const uint64_t below = heads & le;
    KeyGroup g{lane, 0, 0};
    if (active) {
        g.leader = 63 - __clzll((long long)below);
        g.rank = lane - g.leader;
    }
    if (head) {
        // the run ends at the next head or at the next inactive lane
        const uint64_t stop = (heads | ~act) & ~le;
        const int next = stop ? __ffsll((unsigned long long)stop) - 1 : 64;
        g.count = next - lane;
    }
    return g;
}

// For every active lane: reserve one slot in counters[key]; returns the slot
// index.  All leaders issue their atomic together (one round trip per
// wavefront); within a run the slots follow lane order.
__device__ __forceinline__ int32_t wave_reserve(int32_t* counters, uint32_t key, bool active) {
    const KeyGroup g = wave_group(key, active);
    int32_t base = 0;
    if (g.count > 0) base = atomicAdd(&counters[key], g.count);
    base = __shfl(base, g.leader);
    return base + g.rank;
}

// fine key of a marker (see the file header); sets *bad when the P cell is
// not within one cell of the basic cell (coordinates beyond ~1e15 m)
__device__ __forceinline__ uint32_t fine_key(const Cells& c, int nty, bool* bad) {
    int di = c.ip - c.ib, dj = c.jp - c.jb;
    *bad = di < 0 || di > 1 || dj < 0 || dj > 1;
    di = min(max(di, 0), 1);
    dj = min(max(dj, 0), 1);
    const int ty = (c.ib - 1) / TILE, tx = (c.jb - 1) / TILE;
    const int fi = 2 * (c.ib - 1 - ty * TILE) + di;
    const int fj = 2 * (c.jb - 1 - tx * TILE) + dj;
    return (uint32_t)((tx * nty + ty) * NFINE + fj * FINE + fi);
}

// fine key, rank within the key (arrival order) and count of every marker;
// with PropArgs also the marker properties, in the same pass over the markers
// in storage order
template <class P>
__global__ __launch_bounds__(256) void k_fine_count(GeomK g, int nty, const double* __restrict__ xm,
                                                    const double* __restrict__ ym, int64_t M,
                                                    uint32_t* __restrict__ key,
                                                    int32_t* __restrict__ rank,
                                                    int32_t* __restrict__ count,
                                                    int32_t* __restrict__ flags,
                                                    int32_t* __restrict__ big, P props) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = m < M;
    if (m == 0) big[0] = 0;  // k_fine_order, which fills the list, runs after
    uint32_t k = 0;
    if (active) {
        bool bad;
        k = fine_key(marker_cells(g, xm[m], ym[m]), nty, &bad);
        if (bad) flags[0] = 1;
        key[m] = k;
    }
    const int32_t r = wave_reserve(count, k, active);
    if (active) rank[m] = r;
    if (active) props.apply(m);
    // last, where nothing else is live: the step scalars, when they come with
    // this launch, to the device block for the kernels after it
    if (m == 0) props.publish();
}

template <class R>
static PropArgs<R> prop_args(ereb_ctx* c, const PropsReq& q) {
    PropArgs<R> P;
    P.M = c->d_mat;
    P.S = c->d_scal;  // hrs, hrf written there by set_heating
    // the step's scalars with this launch when they ride with the key pass
    P.Sv = c->scal;
    P.Sw = c->scal_ride && c->marknum > 0 ? c->d_scal : nullptr;
    P.mode = q.mode;
    P.tm = tm_ptr(c);
    P.tkm = mptr(c, EREB_M_TKM);
    P.phim = mptr(c, EREB_M_PHIM);
    auto f = [&](int field) { return (R*)c->d_m[field]; };
    P.XW = f(EREB_M_XWSOLIDM0);
    P.rhototal = f(EREB_M_RHOTOTALM);
    P.rhocptotal = f(EREB_M_RHOCPTOTALM);
    P.etatotal = f(EREB_M_ETATOTALM);
    P.hrtotal = f(EREB_M_HRTOTALM);
    P.ktotal = f(EREB_M_KTOTALM);
    P.tkrhocp = f(EREB_M_TKM_RHOCPTOTALM);
    P.etafkphi = f(EREB_M_ETAFLUIDCUR_INV_KPHIM);
    return P;
}

// each marker to its sorted slot: the key's start plus its rank.  The
// workgroups are numbered XCD-contiguously, so that neighbouring marker
// columns, whose markers fill the same key segments, write through one L2.
__global__ __launch_bounds__(256) void k_fine_place(const uint32_t* __restrict__ key,
                                                    const int32_t* __restrict__ rank, int64_t M,
                                                    const int32_t* __restrict__ start,
                                                    uint32_t* __restrict__ perm) {
    const int64_t m = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (m < M) perm[start[key[m]] + rank[m]] = (uint32_t)m;
}

constexpr int SEG_REG = 16;     // segments ordered in registers by rank
constexpr int SEG_THREAD = 48;  // longest segment ordered by one thread
constexpr int SEG_BIG = 16384;  // longest segment ordered in LDS (64 KB)

__global__ __launch_bounds__(256) void k_fine_order(const int32_t* __restrict__ start, int64_t nkeys,
                                                    uint32_t* __restrict__ perm,
                                                    int32_t* __restrict__ big) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const int32_t a = start[k], e = start[k + 1];
    if (e - a <= 1) return;
    if (e - a > SEG_THREAD) {
        const int slot = atomicAdd(&big[0], 1);
        big[1 + slot] = (int32_t)k;
        return;
    }
    if (e - a <= SEG_REG) {
        // registers: each value's rank among the (distinct) segment values is
        // its position
        uint32_t v[SEG_REG];
        const int n = e - a;
#pragma unroll
        for (int i = 0; i < SEG_REG; ++i) v[i] = i < n ? perm[a + i] : 0xffffffffu;
#pragma unroll
        for (int i = 0; i < SEG_REG; ++i) {
            int r = 0;
#pragma unroll
            for (int j = 0; j < SEG_REG; ++j) r += (v[j] < v[i]) ? 1 : 0;
            if (i < n) perm[a + r] = v[i];
        }
        return;
    }
    for (int32_t p = a + 1; p < e; ++p) {  // insertion sort, ascending
        const uint32_t v = perm[p];
        int32_t q = p - 1;
        while (q >= a && perm[q] > v) {
            perm[q + 1] = perm[q];
            --q;
        }
        perm[q + 1] = v;
    }
}

__global__ __launch_bounds__(1024) void k_fine_order_big(const int32_t* __restrict__ start,
                                                         uint32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ big,
                                                         uint32_t* __restrict__ tmp) {
    extern __shared__ uint32_t sdyn[];
    const int nbig = big[0];
    for (int e = blockIdx.x; e < nbig; e += gridDim.x) {
        const int32_t k = big[1 + e];
        const int32_t a = start[k], L = start[k + 1] - a;
        int n = 1;
        while (n < L) n <<= 1;
        // LDS when it fits, else the scratch region [2a, 2a + n) of tmp,
        // disjoint between segments because n < 2L
        uint32_t* s = n <= SEG_BIG ? sdyn : tmp + 2 * (int64_t)a;
        for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = i < L ? perm[a + i] : 0xFFFFFFFFu;
        __syncthreads();
        bitonic(s, n);
        for (int i = threadIdx.x; i < L; i += blockDim.x) perm[a + i] = s[i];
        __syncthreads();
    }
}

int launch_fine_sort(ereb_ctx* c, const PropsReq* props) {
    const int64_t M = c->marknum;
    const int64_t nk = c->nkeys;
    int rc;
    hipEvent_t ev;
    if (props && (rc = set_heating(c, props->hrs, props->hrf))) return rc;
    // perm is rewritten below: it orders no marker set until the sort completes
    c->have_sort = false;
    prof_begin(c, EREB_K_SORT, &ev);
    // the key counts are zero here (left so by the previous sort's scan)
    // unless a sort stopped half-way; a domain error flag stays set until
    // the host reads it (check_flags)
    if (c->fcount_dirty && (rc = fill_i32(c, c->d_fcount, 0, nk))) return rc;
    // ranks live in the first M words of the oversize-segment scratch, which
    // k_fine_order_big uses only after the placement
    int32_t* rank = (int32_t*)c->d_tmp;
    if (M > 0) {
        c->fcount_dirty = true;
        const unsigned nb = (unsigned)((M + 255) / 256);
        const GeomK g = geomk(c);
        const double *xm = mptr(c, EREB_M_XM), *ym = mptr(c, EREB_M_YM);
        // the property vectors are float exactly when the context is in the mixed form
        if (!props)
            hipLaunchKernelGGL(k_fine_count<NoProps>, dim3(nb), dim3(256), 0, c->stream, g,
                               c->nty, xm, ym, M, c->d_key, rank, c->d_fcount, c->d_flags,
                               c->d_big, NoProps{});
        else {
            // the key pass's completion as ev_keys when asked (rec_keys)
            hipEvent_t done = c->rec_keys ? c->ev_keys : nullptr;
            c->rec_keys = false;
            if (marker_is_f32(c->precision, EREB_M_ETATOTALM)) {
                const auto P = prop_args<float>(c, *props);
                if (done)
                    hipExtLaunchKernelGGL(k_fine_count<PropArgs<float>>, dim3(nb), dim3(256), 0,
                                          c->stream, nullptr, done, 0, g, c->nty, xm, ym, M,
                                          c->d_key, rank, c->d_fcount, c->d_flags, c->d_big, P);
                else
                    hipLaunchKernelGGL(k_fine_count<PropArgs<float>>, dim3(nb), dim3(256), 0,
                                       c->stream, g, c->nty, xm, ym, M, c->d_key, rank,
                                       c->d_fcount, c->d_flags, c->d_big, P);
            } else {
                const auto P = prop_args<double>(c, *props);
                if (done)
                    hipExtLaunchKernelGGL(k_fine_count<PropArgs<double>>, dim3(nb), dim3(256),
                                          0, c->stream, nullptr, done, 0, g, c->nty, xm, ym, M,
                                          c->d_key, rank, c->d_fcount, c->d_flags, c->d_big, P);
                else
                    hipLaunchKernelGGL(k_fine_count<PropArgs<double>>, dim3(nb), dim3(256), 0,
                                       c->stream, g, c->nty, xm, ym, M, c->d_key, rank,
                                       c->d_fcount, c->d_flags, c->d_big, P);
            }
        }
        if (props && c->scal_ride) {  // the device block now follows this launch
            c->scal_ride = false;
            c->scal_dev = c->scal;
            c->scal_written = true;
        }
    }
    // key starts; the counts are cleared for the next sort
    if ((rc = scan_exclusive(c, c->d_fcount, c->d_fstart, nk, nullptr, true))) return rc;
    c->fcount_dirty = false;
    if (M > 0) {
        const unsigned nkb = (unsigned)((nk + 255) / 256);
        const unsigned nb = (unsigned)((M + 255) / 256);
        hipLaunchKernelGGL(k_fine_place, dim3(nb), dim3(256), 0, c->stream, c->d_key,
                           (const int32_t*)rank, M, c->d_fstart, c->d_perm);
        hipLaunchKernelGGL(k_fine_order, dim3(nkb), dim3(256), 0, c->stream, c->d_fstart, nk,
                           c->d_perm, c->d_big);
        hipLaunchKernelGGL(k_fine_order_big, dim3(64), dim3(1024), SEG_BIG * sizeof(uint32_t),
                           c->stream, c->d_fstart, c->d_perm, c->d_big, c->d_tmp);
    }
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_SORT, ev);
    c->have_sort = true;
    c->sort_gen = c->pos_gen;
    c->sort_m = M;
    return EREB_OK;
}

int ensure_sort(ereb_ctx* c) {
    if (c->have_sort && c->sort_gen == c->pos_gen) return EREB_OK;
    return launch_fine_sort(c);
}

}  // namespace ereb
