// ereb_replenish.hip -- replenish_markers! (E:5586-5698): add markers where
// the marker launch lattice has gone empty.
//
// The reference runs three sequential loops; here each is a data-parallel
// pass with the same result:
//   1. nearest marker of every anchor among the markers whose marker cell
//      touches it (update_marker_population_geometry!, E:5518-5544): a
//      64-bit atomicMin of the distance bits (non-negative doubles order as
//      their bit patterns), then an atomicMin of the marker index among the
//      markers at that distance -- the lowest index, as the strict `<` of the
//      sequential loop keeps the first;
//   2. every anchor left without a marker scans its (dii, djj) window in the
//      reference's order for the markers found in pass 1 (E:5641-5654), one
//      wavefront per such anchor; it reads only pass-1 results, as the
//      reference does (it tests mnum > 0);
//   3. an exclusive scan of the "new marker" flags in column-major anchor
//      order gives each new marker its index, so markers are appended in
//      the reference's order, copying every resident marker vector.
#include <algorithm>
#include <cstring>

#include "ereb_internal.h"

namespace ereb {

struct MGridK {
    int Nxm, Nym, dii, djj;
    double dxm, dym, idxm, idym, mdis_init;
    const double* xxm;
    const double* yym;
};

// distance E:1097-1099
__device__ __forceinline__ double mdist(double x1, double y1, double x2, double y2) {
    const double a = x1 - x2, b = y1 - y2;
    return sqrt(a * a + b * b);
}

// fix on the marker lattice (clamps 1..Nxm-1, 1..Nym-1), 1-based (i, j)
__device__ __forceinline__ void mcell(const MGridK& G, double x, double y, int& i, int& j) {
    j = jl_cell((x - G.xxm[0]) * G.idxm, 1, G.Nxm - 1);
    i = jl_cell((y - G.yym[0]) * G.idym, 1, G.Nym - 1);
}

__global__ __launch_bounds__(256) void k_anchor_init(uint64_t* __restrict__ dbits,
                                                     int32_t* __restrict__ pick, uint64_t init,
                                                     int64_t n) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    dbits[a] = init;
    pick[a] = INT32_MAX;
}

constexpr int AWIN = 2048;  // anchors in an LDS window (16 KB of minima)

// Anchors touched by a workgroup's markers: rows i0 .. i0+rows-1 and
// columns j0 .. j0+cols-1 (1-based; cells i..i+1 x j..j+1 of every active
// marker).  `local` (uniform): the box holds at most AWIN anchors.  Every
// thread of the workgroup calls it (it synchronises).
struct AnchorBox {
    int i0, j0, rows, cols;
    bool local;
};
__device__ AnchorBox anchor_box(bool active, int i, int j, int* box) {
    const int t = threadIdx.x;
    if (t == 0) {
        box[0] = INT32_MAX;
        box[1] = INT32_MIN;
        box[2] = INT32_MAX;
        box[3] = INT32_MIN;
    }
    int imin = active ? i : INT32_MAX, imax = active ? i : INT32_MIN;
    int jmin = active ? j : INT32_MAX, jmax = active ? j : INT32_MIN;
    for (int o = 32; o > 0; o >>= 1) {
        imin = min(imin, __shfl_xor(imin, o));
        imax = max(imax, __shfl_xor(imax, o));
        jmin = min(jmin, __shfl_xor(jmin, o));
        jmax = max(jmax, __shfl_xor(jmax, o));
    }
    __syncthreads();
    if ((t & 63) == 0) {
        atomicMin(&box[0], imin);
        atomicMax(&box[1], imax);
        atomicMin(&box[2], jmin);
        atomicMax(&box[3], jmax);
    }
    __syncthreads();
    AnchorBox b;
    b.i0 = box[0];
    b.j0 = box[2];
    b.rows = box[1] - b.i0 + 2;
    b.cols = box[3] - b.j0 + 2;
    b.local = box[1] >= b.i0 && (int64_t)b.rows * b.cols <= AWIN;
    return b;
}

// Both passes visit the markers in `order` when it is given (a permutation
// of the markers that keeps neighbours together: the fine-key sort of the
// timestep's scatter, made before the advection moved the markers by less
// than a cell), else in storage order.  Minima and picks do not depend on
// the order; the window of pass 1a does: advected markers drift away from
// their storage-order neighbours, and a workgroup's anchors then span more
// than the window.
__device__ __forceinline__ int64_t visit(const uint32_t* order, int64_t k) {
    return order ? (int64_t)order[k] : k;
}

// pass 1a: smallest distance per anchor.  A workgroup's markers (256
// consecutive ones in the visiting order: part of a tile of the sort, or in
// the lattice order of the reference's markers a few lattice columns) first
// take their minima in an LDS window over the anchors they touch, then one
// global atomic per touched anchor; a workgroup whose anchors span more than
// the window goes straight to the global minima.
__global__ __launch_bounds__(256) void k_anchor_min(MGridK G, const double* __restrict__ xm,
                                                    const double* __restrict__ ym, int64_t M,
                                                    const uint32_t* __restrict__ order,
                                                    uint64_t* __restrict__ dbits) {
    __shared__ unsigned long long win[AWIN];
    __shared__ int box[4];  // min cell row, max cell row, min cell column, max cell column
    const int t = threadIdx.x;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + t;
    const bool active = k < M;
    double x = 0.0, y = 0.0;
    int i = 1, j = 1;
    if (active) {
        const int64_t m = visit(order, k);
        x = xm[m];
        y = ym[m];
        mcell(G, x, y, i, j);
    }
    const AnchorBox B = anchor_box(active, i, j, box);
    const int i0 = B.i0, j0 = B.j0, rows = B.rows, cols = B.cols;
    const bool local = B.local;
    if (local) {
        for (int k = t; k < rows * cols; k += blockDim.x) win[k] = ~0ull;
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int ii = i + (q & 1), jj = j + (q >> 1);
        const double d = active ? mdist(x, y, G.xxm[jj - 1], G.yym[ii - 1]) : G.mdis_init;
        const uint64_t b = (uint64_t)__double_as_longlong(d);
        if (!(d < G.mdis_init)) continue;
        if (local) {
            atomicMin(&win[(jj - j0) * rows + (ii - i0)], (unsigned long long)b);
        } else {
            uint64_t* a = &dbits[(size_t)(jj - 1) * G.Nym + (ii - 1)];
            // the stored value only decreases: skip atomics that cannot win
            if (b < __atomic_load_n(a, __ATOMIC_RELAXED))
                atomicMin((unsigned long long*)a, (unsigned long long)b);
        }
    }
    if (!local) return;
    __syncthreads();
    for (int k = t; k < rows * cols; k += blockDim.x) {
        const unsigned long long b = win[k];
        if (b == ~0ull) continue;
        const int ii = i0 + k % rows, jj = j0 + k / rows;
        uint64_t* a = &dbits[(size_t)(jj - 1) * G.Nym + (ii - 1)];
        if (b < __atomic_load_n(a, __ATOMIC_RELAXED))
            atomicMin((unsigned long long*)a, b);
    }
}

// pass 1b: lowest marker index at that distance (about one atomic per
// anchor: an LDS window as in pass 1a measured slower here)
__global__ __launch_bounds__(256) void k_anchor_pick(MGridK G, const double* __restrict__ xm,
                                                     const double* __restrict__ ym, int64_t M,
                                                     const uint32_t* __restrict__ order,
                                                     const uint64_t* __restrict__ dbits,
                                                     int32_t* __restrict__ pick) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const int64_t m = visit(order, k);
    const double x = xm[m], y = ym[m];
    int i, j;
    mcell(G, x, y, i, j);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int ii = i + (q & 1), jj = j + (q >> 1);
        const double d = mdist(x, y, G.xxm[jj - 1], G.yym[ii - 1]);
        const size_t a = (size_t)(jj - 1) * G.Nym + (ii - 1);
        if (d < G.mdis_init && (uint64_t)__double_as_longlong(d) == dbits[a])
            atomicMin(&pick[a], (int32_t)(m + 1));
    }
}

// pass 2a: mdis / mnum of the anchors that have a pass-1 marker; the others
// get (mdis_init, 0, no flag) and go to the list of empty anchors
__global__ __launch_bounds__(256) void k_anchor_fill(MGridK G,
                                                     const uint64_t* __restrict__ dbits,
                                                     const int32_t* __restrict__ pick,
                                                     double* __restrict__ mdis,
                                                     int32_t* __restrict__ mnum,
                                                     int32_t* __restrict__ flag,
                                                     int32_t* __restrict__ empty,
                                                     int32_t* __restrict__ nempty) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= (int64_t)G.Nxm * G.Nym) return;
    const int p = pick[a];
    flag[a] = 0;
    if (p != INT32_MAX) {
        mdis[a] = __longlong_as_double((long long)dbits[a]);
        mnum[a] = p;
        return;
    }
    mdis[a] = G.mdis_init;
    mnum[a] = 0;
    empty[atomicAdd(nempty, 1)] = (int32_t)a;  // list order does not affect any result
}

// pass 2b: one workgroup per empty anchor scans its window in the
// reference's order (jj outer, ii inner), reading only pass-1 picks (the
// reference tests mnum > 0, so anchors filled here never feed another).
// Thread t takes window positions t, t + 256, ..., loading the picks of
// four positions, then their markers, before comparing; the workgroup keeps
// the smallest distance and, among equal ones, the first position -- the
// pick of the sequential strict-< scan.
constexpr int SEARCH_THREADS = 256;

struct Best {
    double d;
    int w, q;
};
// the better of two candidates: smaller distance, then earlier position
__device__ __forceinline__ void take_better(Best& a, const Best& b) {
    if (b.d < a.d || (b.d == a.d && b.w < a.w)) a = b;
}

__global__ __launch_bounds__(SEARCH_THREADS) void k_anchor_search(
    MGridK G, const double* __restrict__ xm, const double* __restrict__ ym,
    const int32_t* __restrict__ pick, const int32_t* __restrict__ empty,
    const int32_t* __restrict__ nempty, double* __restrict__ mdis, int32_t* __restrict__ mnum,
    int32_t* __restrict__ flag) {
    constexpr int U = 4;
    __shared__ Best wbest[SEARCH_THREADS / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int ne = *nempty;
    for (int e = blockIdx.x; e < ne; e += gridDim.x) {  // block-uniform bound: every block exits
        const int64_t a = empty[e];
        const int i = (int)(a % G.Nym) + 1, j = (int)(a / G.Nym) + 1;
        const double xa = G.xxm[j - 1], ya = G.yym[i - 1];
        const int jlo = max(j - G.djj, 1), jhi = min(j + G.djj, G.Nxm);
        const int ilo = max(i - G.dii, 1), ihi = min(i + G.dii, G.Nym);
        const int h = ihi - ilo + 1, nw = h * (jhi - jlo + 1);
        Best best{G.mdis_init, INT32_MAX, 0};
        for (int w0 = t; w0 < nw; w0 += U * SEARCH_THREADS) {
            int q[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int w = w0 + u * SEARCH_THREADS;
                q[u] = INT32_MAX;
                if (w < nw) {
                    const int jj = jlo + w / h, ii = ilo + w % h;
                    q[u] = pick[(size_t)(jj - 1) * G.Nym + (ii - 1)];
                }
            }
            double px[U], py[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                px[u] = py[u] = 0.0;
                if (q[u] != INT32_MAX) {
                    px[u] = xm[q[u] - 1];
                    py[u] = ym[q[u] - 1];
                }
            }
            // positions rise with u and w0: strict < keeps a thread's first
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (q[u] == INT32_MAX) continue;
                const double d = mdist(px[u], py[u], xa, ya);
                if (d < best.d) best = Best{d, w0 + u * SEARCH_THREADS, q[u]};
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const Best o{__shfl_xor(best.d, off), __shfl_xor(best.w, off),
                         __shfl_xor(best.q, off)};
            take_better(best, o);
        }
        if (lane == 0) wbest[wave] = best;
        __syncthreads();
        if (t == 0) {
            Best b = wbest[0];
            for (int k = 1; k < SEARCH_THREADS / 64; ++k) take_better(b, wbest[k]);
            if (b.w != INT32_MAX) {
                mdis[a] = b.d;
                mnum[a] = -b.q;
                flag[a] = 1;
            }
        }
        __syncthreads();  // wbest is rewritten for the next anchor
    }
}

struct MarkerFields {
    void* f[EREB_M_NFIELDS];
    int es[EREB_M_NFIELDS];  // element size in bytes: 4 or 8
};

// splitmix64 step: golden-ratio increment, then the finaliser
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// a counter-based uniform in [0, 1)
__device__ __forceinline__ double u01(uint64_t z) {
    return (double)(splitmix64(z) >> 11) * 0x1p-53;
}

// pass 3: append the new markers in column-major anchor order (E:5638-5683)
// field f of new marker k (0-based among all markers) from anchor a: the
// position of the anchor (jittered when randomized), every other field
// copied from the anchor's marker
__device__ __forceinline__ void append_field(const MGridK& G, const MarkerFields& F,
                                             const int32_t* __restrict__ mnum, int64_t a,
                                             int64_t k, int f, int randomized, uint64_t seed) {
    if (f == EREB_M_XM || f == EREB_M_YM) {
        const bool isx = f == EREB_M_XM;
        const int i = (int)(a % G.Nym) + 1, j = (int)(a / G.Nym) + 1;
        double v = isx ? G.xxm[j - 1] : G.yym[i - 1];
        if (randomized)  // E:5662-5663 with a counter-based generator
            v = v + (u01(seed + (2 * (uint64_t)a + (isx ? 0 : 1))) - 0.5) * (isx ? G.dxm : G.dym);
        ((double*)F.f[f])[k] = v;
        return;
    }
    const int64_t src = (int64_t)(-mnum[a]) - 1;
    if (F.es[f] == 8)
        ((double*)F.f[f])[k] = ((const double*)F.f[f])[src];
    else
        ((int32_t*)F.f[f])[k] = ((const int32_t*)F.f[f])[src];
}

__device__ __forceinline__ void append_one(const MGridK& G, const MarkerFields& F,
                                           const int32_t* __restrict__ mnum, int64_t a,
                                           int64_t k, int randomized, uint64_t seed) {
#pragma unroll
    for (int f = 0; f < EREB_M_NFIELDS; ++f) append_field(G, F, mnum, a, k, f, randomized, seed);
}

// positions from an exclusive scan of the flags of all anchors
__global__ __launch_bounds__(256) void k_anchor_append(MGridK G, MarkerFields F,
                                                       const int32_t* __restrict__ mnum,
                                                       const int32_t* __restrict__ flag,
                                                       const int32_t* __restrict__ pos,
                                                       int64_t M, int randomized, uint64_t seed) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= (int64_t)G.Nxm * G.Nym || !flag[a]) return;
    append_one(G, F, mnum, a, M + pos[a], randomized, seed);
}

// The new markers come only from the empty anchors.  When there are at most
// NEW_MAX of them (every timestep but the first of a sparse start), one
// workgroup sorts the empty-anchor list (filled in atomic order) into anchor
// order and compacts the flagged ones into list[0..total); res[0] = total,
// or -1 when the list is longer (the flags of all anchors are then scanned).
constexpr int NEW_MAX = 4096;
constexpr int NEW_THREADS = 1024;

__global__ __launch_bounds__(NEW_THREADS) void k_new_order(const int32_t* __restrict__ empty,
                                                           const int32_t* __restrict__ nempty,
                                                           const int32_t* __restrict__ flag,
                                                           int32_t* __restrict__ list,
                                                           int32_t* __restrict__ res) {
    constexpr int PER = NEW_MAX / NEW_THREADS;
    __shared__ uint32_t s[NEW_MAX];
    __shared__ int32_t part[NEW_THREADS];
    const int t = threadIdx.x;
    const int ne = *nempty;
    if (ne > NEW_MAX) {  // block-uniform: no thread reaches a barrier below
        if (t == 0) res[0] = -1;
        return;
    }
    int n = 1;
    while (n < ne) n <<= 1;
    for (int i = t; i < n; i += NEW_THREADS) s[i] = i < ne ? (uint32_t)empty[i] : 0xFFFFFFFFu;
    __syncthreads();
    bitonic(s, n);
    // thread t owns sorted entries [t*PER, t*PER + PER): flags, then an
    // exclusive scan of the per-thread counts (Hillis-Steele over the block)
    int f[PER], cnt = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int i = t * PER + u;
        f[u] = i < ne ? (flag[s[i]] != 0) : 0;
        cnt += f[u];
    }
    part[t] = cnt;
    __syncthreads();
    for (int off = 1; off < NEW_THREADS; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int k = part[t] - cnt;  // exclusive
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (f[u]) list[k++] = (int32_t)s[t * PER + u];
    if (t == NEW_THREADS - 1) res[0] = part[t];
}

// positions from the compacted list, one thread per (new marker, field): the
// copies' loads are independent; the count is read on the device
// (res[0] = -1 appends nothing), so the launch needs no host round trip
__global__ __launch_bounds__(256) void k_anchor_append_list(MGridK G, MarkerFields F,
                                                            const int32_t* __restrict__ mnum,
                                                            const int32_t* __restrict__ list,
                                                            const int32_t* __restrict__ res,
                                                            int64_t M, int randomized,
                                                            uint64_t seed) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k = idx / EREB_M_NFIELDS;
    const int f = (int)(idx % EREB_M_NFIELDS);
    if (k >= res[0]) return;
    append_field(G, F, mnum, list[k], M + k, f, randomized, seed);
}

int launch_replenish(ereb_ctx* c, const ereb_marker_grid& mg, int randomized, uint64_t seed,
                     int64_t* added) {
    const int64_t M = c->marknum;
    const int64_t A = mg.Nxm * mg.Nym;
    // the jitter stream of anchor a is u01(h(seed) + 2a + {0, 1}): the seed
    // is mixed once, so nearby seeds (a timestep counter, say) give unrelated
    // streams rather than the same draws moved between anchors
    seed = splitmix64(seed);
    if (mg.Nxm < 2 || mg.Nym < 2 || !mg.xxm || !mg.yym || A > INT32_MAX)
        return fail(c, EREB_ERR_ARG, "marker lattice needs Nxm, Nym >= 2 and both axes");
    int rc;
    // scratch: distance bits, picks, mdis, mnum, flags, positions, the
    // empty-anchor list and its count, axes
    const size_t need = A * (sizeof(uint64_t) + sizeof(double) + 5 * sizeof(int32_t)) +
                        2 * sizeof(int32_t) + sizeof(double) * (mg.Nxm + mg.Nym) +
                        sizeof(int32_t) * (NEW_MAX + 1) + 256 * 12;
    if (need > c->rep_bytes) {
        hipFree(c->rep_pool);
        c->rep_pool = nullptr;
        c->rep_bytes = 0;
        EREB_HIP(c, hipMalloc(&c->rep_pool, need));
        c->rep_bytes = need;
    }
    char* p = (char*)c->rep_pool;
    auto carve = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    uint64_t* dbits = (uint64_t*)carve(sizeof(uint64_t) * A);
    double* mdis = (double*)carve(sizeof(double) * A);
    int32_t* pick = (int32_t*)carve(sizeof(int32_t) * A);
    int32_t* mnum = (int32_t*)carve(sizeof(int32_t) * A);
    int32_t* flag = (int32_t*)carve(sizeof(int32_t) * A);
    int32_t* pos = (int32_t*)carve(sizeof(int32_t) * (A + 1));
    int32_t* empty = (int32_t*)carve(sizeof(int32_t) * A);
    int32_t* nempty = (int32_t*)carve(sizeof(int32_t));
    double* axes = (double*)carve(sizeof(double) * (mg.Nxm + mg.Nym));
    int32_t* list = (int32_t*)carve(sizeof(int32_t) * NEW_MAX);
    int32_t* res = (int32_t*)carve(sizeof(int32_t));
    if ((size_t)(p - (char*)c->rep_pool) > c->rep_bytes)
        return fail(c, EREB_ERR_NOMEM, "replenish scratch carve overflow");
    EREB_HIP(c, hipMemcpyAsync(axes, mg.xxm, sizeof(double) * mg.Nxm, hipMemcpyHostToDevice,
                               c->stream));
    EREB_HIP(c, hipMemcpyAsync(axes + mg.Nxm, mg.yym, sizeof(double) * mg.Nym,
                               hipMemcpyHostToDevice, c->stream));
    MGridK G;
    G.Nxm = (int)mg.Nxm;
    G.Nym = (int)mg.Nym;
    G.dii = (int)std::min<int64_t>(std::max<int64_t>(mg.dii, 0), mg.Nym);
    G.djj = (int)std::min<int64_t>(std::max<int64_t>(mg.djj, 0), mg.Nxm);
    G.dxm = mg.dxm;
    G.dym = mg.dym;
    G.idxm = 1.0 / mg.dxm;
    G.idym = 1.0 / mg.dym;
    G.mdis_init = mg.mdis_init;
    G.xxm = axes;
    G.yym = axes + mg.Nxm;
    hipEvent_t ev;
    prof_begin(c, EREB_K_REPLENISH, &ev);
    const dim3 ga((unsigned)((A + 255) / 256)), gm((unsigned)((M + 255) / 256));
    uint64_t init_bits;
    std::memcpy(&init_bits, &mg.mdis_init, sizeof init_bits);
    hipLaunchKernelGGL(k_anchor_init, ga, dim3(256), 0, c->stream, dbits, pick, init_bits, A);
    if (M > 0) {
        // the last completed sort, stale or not, while it orders exactly
        // these markers (the sort buffers are reallocated, and have_sort
        // cleared, when the capacity grows)
        const uint32_t* order = c->have_sort && c->sort_m == M ? c->d_perm : nullptr;
        hipLaunchKernelGGL(k_anchor_min, gm, dim3(256), 0, c->stream, G, mptr(c, EREB_M_XM),
                           mptr(c, EREB_M_YM), M, order, dbits);
        // the index pass in the same order: it re-reads the markers and the
        // anchors of pass 1a while they are in cache (storage order measured
        // 198 against 191 us for the whole replenishment)
        hipLaunchKernelGGL(k_anchor_pick, gm, dim3(256), 0, c->stream, G, mptr(c, EREB_M_XM),
                           mptr(c, EREB_M_YM), M, order, (const uint64_t*)dbits, pick);
    }
    if ((rc = fill_i32(c, nempty, 0, 1))) return rc;
    hipLaunchKernelGGL(k_anchor_fill, ga, dim3(256), 0, c->stream, G, (const uint64_t*)dbits,
                       (const int32_t*)pick, mdis, mnum, flag, empty, nempty);
    hipLaunchKernelGGL(k_anchor_search, dim3(2048), dim3(SEARCH_THREADS), 0, c->stream, G,
                       mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), (const int32_t*)pick,
                       (const int32_t*)empty, (const int32_t*)nempty, mdis, mnum, flag);
    hipLaunchKernelGGL(k_new_order, dim3(1), dim3(NEW_THREADS), 0, c->stream,
                       (const int32_t*)empty, (const int32_t*)nempty, (const int32_t*)flag, list,
                       res);
    auto fields = [&]() {
        MarkerFields F;
        for (int f = 0; f < EREB_M_NFIELDS; ++f) {
            F.f[f] = c->d_m[f];
            F.es[f] = (int)marker_es(c->precision, f);
        }
        return F;
    };
    // with room for NEW_MAX more markers the listed append goes ahead of the
    // host's read of the count (a longer list appends nothing there)
    const bool roomy = M + NEW_MAX <= c->cap;
    if (roomy)
        hipLaunchKernelGGL(k_anchor_append_list,
                           dim3((unsigned)((NEW_MAX * EREB_M_NFIELDS + 255) / 256)), dim3(256),
                           0, c->stream,
                           G, fields(), (const int32_t*)mnum, (const int32_t*)list,
                           (const int32_t*)res, M, randomized, seed);
    EREB_HIP(c, hipGetLastError());
    int32_t total = 0;
    EREB_HIP(c, hipMemcpyAsync(&total, res, sizeof total, hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    const bool listed = total >= 0;
    if (!listed) {  // more empty anchors than one workgroup orders: scan all flags
        if ((rc = scan_exclusive(c, flag, pos, A))) return rc;
        EREB_HIP(c, hipMemcpyAsync(&total, pos + A, sizeof total, hipMemcpyDeviceToHost,
                                   c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (M + total > (int64_t)INT32_MAX) return fail(c, EREB_ERR_ARG, "marknum exceeds 2^31-1");
    if (M + total > c->cap && (rc = reserve_markers(c, std::max<int64_t>(M + total,
                                                                         c->cap + c->cap / 4))))
        return rc;
    if (total > 0 && !(roomy && listed)) {
        if (listed)
            hipLaunchKernelGGL(k_anchor_append_list,
                               dim3((unsigned)(((int64_t)total * EREB_M_NFIELDS + 255) / 256)),
                               dim3(256), 0, c->stream, G, fields(), (const int32_t*)mnum,
                               (const int32_t*)list, (const int32_t*)res, M, randomized, seed);
        else
            hipLaunchKernelGGL(k_anchor_append, ga, dim3(256), 0, c->stream, G, fields(),
                               (const int32_t*)mnum, (const int32_t*)flag, (const int32_t*)pos,
                               M, randomized, seed);
        EREB_HIP(c, hipGetLastError());
    }
    prof_end(c, EREB_K_REPLENISH, ev);
    if (total > 0) {
        c->marknum = M + total;
        ++c->pos_gen;
    }
    c->rep_mdis = mdis;
    c->rep_mnum = mnum;
    c->rep_n = A;
    *added = total;
    return EREB_OK;
}

}  // namespace ereb
