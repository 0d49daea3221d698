// ereb_scatter.hip -- step 2 of the transfer pipeline: the bilinear
// marker->node scatter onto the basic, Vx, Vy and P grids
// (marker_to_{basic,vx,vy,p}_nodes!, E:1605-1863, loop E:6419-6517).  Step 1,
// the marker properties (compute_marker_properties!, E:700-750), is evaluated
// in storage order by the sort's key pass (k_fine_count in ereb_sort.hip),
// where the property vectors are written with coalesced stores; the scatter
// reads them through the sort order.
//
// Geometry of the reduction:
//   - A marker's four stencils follow from two cells: basic (ib, jb) and
//     P (ip, jp); Vx uses (ip, jb), Vy uses (ib, jp), because the staggered
//     axes (C:40-64) and clamps (C:66-101) coincide that way.  The pair
//     (ip-ib, jp-jb) is in {0,1}^2, so a marker's "fine cell" (a quarter of
//     its basic cell) fixes all 16 nodes it touches, and every marker of a
//     tile of TILE x TILE basic cells lands in a (TILE+2)^2 = 64-node window
//     of each grid.
//   - Markers arrive sorted by fine key (ereb_sort.hip): a tile's markers are
//     contiguous and grouped by fine cell in ascending marker index.
//
// One workgroup per tile (eight wavefronts in the Float64 form, four in the
// mixed form; ScatterShape below).  The tile is walked one basic
// column at a time: the column's 24 fine cells (fine columns 2k, 2k+1) hold a
// contiguous run of sorted markers, split into chunks of at most CH markers
// when it is longer.  Per chunk, between two workgroup barriers:
//   1. tasks  : wavefronts g and g + 4 own the first and second half of the
//               sums of grid g (basic, Vx, Vy, P; the second half includes
//               the weight sum); with four wavefronts, wavefront g owns them
//               all.  A task is (fine cell, corner column cj): the sums over
//               the fine cell's markers, in order, of property * weight for
//               the corners (0, cj) and (1, cj), and of the weights;
//   2. combine: lane (r, c) of a wavefront owns window node (r, c) of its
//               grid and adds the partials of the at most 8 fine cells of the
//               column that touch it, in a fixed (column, fine column, fine
//               row) order; then
//      stage  : the next chunk's markers (two threads per marker, one per
//               half of its inputs; one thread in the four-wavefront form)
//               stage their 18 property values and the fractions dxmj/dx,
//               dymi/dy of the basic and P stencils in LDS.  Their inputs
//               were loaded into registers before step 1, so the gather
//               latency overlaps steps 1 and 2, and the staging overlaps
//               other wavefronts' combine.
// No atomics: every sum has exactly one owner and a fixed order, so results
// are identical run to run.  Tile windows overlap by two node rows/columns;
// k_combine adds the at most 2 x 2 window partials of each node in a fixed
// tile order.  Sums accumulate with fused multiply-add, which changes the
// rounding of the sequential sums only.
#include "ereb_internal.h"

namespace ereb {

constexpr int CH = 256;      // markers per LDS chunk
constexpr int CHP = CH + 1;  // padded stride: fields of one marker in different banks
constexpr int NCF = 2 * FINE;  // fine cells in one basic column of a tile
constexpr int NLF = 18;        // property values staged per marker
// Partials of one column: value q of grid G for (half column h, corner row
// ci, corner column cj, fine row fi) at
//   part[poff + q*PQ + h*PH + ci*PCI + cj*PCJ + fi].
// In the combine, the 16 lanes that read one value hold 8 fine rows 2 apart
// and the two corner columns: fine rows 2 doubles apart and an odd corner
// column stride (13 doubles) put the 16 reads in 32 distinct LDS banks.
constexpr int PCJ = FINE + 1, PCI = 2 * PCJ, PH = 2 * PCI, PQ = 2 * PH;
constexpr int NPART = NSUM * PQ;

// staged property indices
enum {
    L_ETATOTAL = 0, L_ETAVP, L_INVGGG, L_SXY, L_COHES, L_TENS, L_FRICT,  // basic
    L_RHOTOTAL, L_RHOF, L_KTOTAL, L_PHIM, L_ETAFKPHI,                    // Vx, Vy
    L_SXX, L_RHOCP, L_ALPHAS, L_ALPHAF, L_HR, L_TKRHOCP                  // P
};

// R is the type of the property vectors, the LDS staging and the per-fine-cell
// sums: double in the Float64 form, float in the mixed form.  The marker
// state (coordinates, temperature, porosity, stresses) is double in both.
// Per-type values of the seven material vectors (E:659-665): 1/G (inv),
// cohesion, tension, friction, fluid density, solid and fluid expansivity
struct TypedTab {
    double v[7][3];
};

__host__ static TypedTab typed_tab(const ereb_materials& M) {
    TypedTab T;
    for (int k = 0; k < 3; ++k) {
        T.v[0][k] = 1.0 / M.gggsolidm[k];  // inv(gggsolidm[tm]), E:659
        T.v[1][k] = M.cohessolidm[k];
        T.v[2][k] = M.tenssolidm[k];
        T.v[3][k] = M.frictsolidm[k];
        T.v[4][k] = M.rhofluidm[k];
        T.v[5][k] = M.alphasolidm[k];
        T.v[6][k] = M.alphafluidm[k];
    }
    return T;
}

// row q of the table at type index t (0..2), by selects (no dynamic
// indexing of the kernel argument)
__device__ __forceinline__ double typed_at(const TypedTab& T, int q, int t) {
    return t == 0 ? T.v[q][0] : (t == 1 ? T.v[q][1] : T.v[q][2]);
}

template <class R>
struct ScatterArgs {
    GeomK g;
    int nty;
    const uint32_t* __restrict__ perm;
    const int32_t* __restrict__ fstart;
    const double *xm, *ym, *phim, *sxx, *sxy;
    const R *etavp, *invggg, *frict, *cohes, *tens, *rhof, *alphas, *alphaf;
    const R *rhototal, *rhocptotal, *etatotal, *hrtotal, *ktotal, *tkrhocp, *etafkphi;
    double* part;
    // typed: 1/G, cohesion, tension, friction, fluid density and the two
    // expansivities are the material table values of the marker's type, as
    // the reference sets them (E:659-665, checked on the device by
    // typed_refresh): read through tm from `tab` instead of seven vectors
    int typed;
    const int32_t* tm;
    TypedTab tab;
};

__device__ __forceinline__ double fma_r(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fma_r(float a, float b, float c) { return fmaf(a, b, c); }

// per-node accumulators of the combine: plain double sums in the Float64 form,
// compensated float sums (value + error term) in the mixed form
template <class R> struct NodeAcc;
template <> struct NodeAcc<double> {
    double s[10];
    __device__ void zero() {
#pragma unroll
        for (int q = 0; q < 10; ++q) s[q] = 0.0;
    }
    __device__ void add(int q, double v) { s[q] += v; }
    __device__ double get(int q) const { return s[q]; }
};
template <> struct NodeAcc<float> {
    float s[10], c[10];
    __device__ void zero() {
#pragma unroll
        for (int q = 0; q < 10; ++q) s[q] = c[q] = 0.0f;
    }
    __device__ void add(int q, float v) { two_sum_add(s[q], c[q], v); }
    __device__ double get(int q) const { return (double)s[q] + (double)c[q]; }
};

// per-grid SUM slot offset, partial offset and staged-property list (SUM
// order of reset_interpolated_properties!, E:968-999).
template <int G> struct GridSpec;
template <> struct GridSpec<0> {  // basic: ETA0 ETA GGG SXY COH TEN FRI (+WT)
    static constexpr int slot0 = 0, poff = 0, nf = 7;
    static constexpr bool prow = false, pcol = false;
    __device__ static constexpr int f(int q) {
        return q == 0 ? L_ETATOTAL : q == 1 ? L_ETAVP : q == 2 ? L_INVGGG : q == 3 ? L_SXY
             : q == 4 ? L_COHES : q == 5 ? L_TENS : L_FRICT;
    }
};
template <> struct GridSpec<1> {  // Vx: RHOX RHOFX KX PHIX RX (+WTX)
    static constexpr int slot0 = 8, poff = PQ * 8, nf = 5;
    static constexpr bool prow = true, pcol = false;
    __device__ static constexpr int f(int q) { return L_RHOTOTAL + q; }
};
template <> struct GridSpec<2> {  // Vy: RHOY RHOFY KY PHIY RY (+WTY)
    static constexpr int slot0 = 14, poff = PQ * 14, nf = 5;
    static constexpr bool prow = false, pcol = true;
    __device__ static constexpr int f(int q) { return L_RHOTOTAL + q; }
};
template <> struct GridSpec<3> {  // P: RHO RHOCP ALPHA ALPHAF HR GGGP SXX TK PHI (+WTP)
    static constexpr int slot0 = 20, poff = PQ * 20, nf = 9;
    static constexpr bool prow = true, pcol = true;
    __device__ static constexpr int f(int q) {
        return q == 0 ? L_RHOTOTAL : q == 1 ? L_RHOCP : q == 2 ? L_ALPHAS : q == 3 ? L_ALPHAF
             : q == 4 ? L_HR : q == 5 ? L_INVGGG : q == 6 ? L_SXX : q == 7 ? L_TKRHOCP
             : L_PHIM;
    }
};

template <class R>
struct ScatterLDS {
    R sf[NLF][CHP];
    R sfr[4][CHP];  // dxmj/dx basic, dymi/dy basic, dxmj/dx P, dymi/dy P
    R part[NPART];
    double ax[4][WIN];   // tile window of the x, y, xp, yp axes
    double typed[7][3];  // the typed material table (TypedTab)
    int32_t fs[NFINE + 1];
};

// Values q = Q0 .. Q0+QN-1 of grid G's NQ = nf + 1 sums (q == nf is the weight
// sum) for the task (fine cell fl of the column, corner column cj): lanes
// 0..47 of one wavefront.  The marker values of iteration s + 1 are read from
// LDS while iteration s accumulates.
template <class R, int G, int Q0, int QN>
__device__ __forceinline__ void tasks(ScatterLDS<R>& L, int lane, int f0, int c0, int cend) {
    using S = GridSpec<G>;
    constexpr int NQ = S::nf + 1;
    static_assert(Q0 + QN <= NQ, "value range");
    constexpr bool HASW = Q0 + QN == NQ;   // this part holds the weight sum
    constexpr int NV = HASW ? QN - 1 : QN;  // staged property values read per marker
    if (lane >= 2 * NCF) return;
    const int fl = lane >> 1, cj = lane & 1;
    const int k0 = max(L.fs[f0 + fl], c0) - c0, k1 = min(L.fs[f0 + fl + 1], cend) - c0;
    const R* fa = S::pcol ? L.sfr[2] : L.sfr[0];
    const R* fb = S::prow ? L.sfr[3] : L.sfr[1];
    R s0[QN], s1[QN];  // corners (0, cj) and (1, cj)
#pragma unroll
    for (int q = 0; q < QN; ++q) s0[q] = s1[q] = R(0);
    if (k0 < k1) {
        R v[NV], a = fa[k0], b = fb[k0];
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = L.sf[S::f(Q0 + q)][k0];
        // two markers per trip in the mixed form (measured 4 µs faster there;
        // in the Float64 form two gained nothing and four were 60% slower)
        constexpr int UNROLL = sizeof(R) == 4 ? 2 : 1;
#pragma unroll UNROLL
        for (int s = k0; s < k1; ++s) {
            const int sn = min(s + 1, k1 - 1);
            R vn[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) vn[q] = L.sf[S::f(Q0 + q)][sn];
            const R an = fa[sn], bn = fb[sn];
            // weights order (i,j),(i+1,j),(i,j+1),(i+1,j+1) = (1-b)(1-a), b(1-a), (1-b)a, ba
            const R wa = cj ? a : R(1) - a;
            const R w0 = (R(1) - b) * wa, w1 = b * wa;
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                s0[q] = fma_r(v[q], w0, s0[q]);
                s1[q] = fma_r(v[q], w1, s1[q]);
            }
            if (HASW) {
                s0[QN - 1] += w0;
                s1[QN - 1] += w1;
            }
#pragma unroll
            for (int q = 0; q < NV; ++q) v[q] = vn[q];
            a = an;
            b = bn;
        }
    }
    const int h = fl >= FINE ? 1 : 0, fi = fl - h * FINE;
    R* p = &L.part[S::poff + Q0 * PQ + h * PH + cj * PCJ + fi];
#pragma unroll
    for (int q = 0; q < QN; ++q) {
        p[q * PQ] = s0[q];
        p[q * PQ + PCI] = s1[q];
    }
}

// node (r, c) of grid G adds values Q0 .. Q0+QN-1 of the partials of the fine
// cells of basic column k that touch it: fine columns 2k + h, fine rows
// fi_lo .. fi_lo + 3
template <class R, int G, int Q0, int QN>
__device__ __forceinline__ void combine(const ScatterLDS<R>& L, NodeAcc<R>& acc, int r, int c,
                                        int k) {
    using S = GridSpec<G>;
    const int fi_lo = S::prow ? 2 * r - 3 : 2 * r - 2;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int fj = 2 * k + h;
        const int cj = c - (S::pcol ? (fj + 1) >> 1 : fj >> 1);
        if (cj < 0 || cj > 1) continue;
        R v[4][QN];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int fi = min(max(fi_lo + d, 0), FINE - 1);  // clamped; skipped below
            const int ci = 1 - (d >> 1);  // = r - (grid row of fine row fi)
            const R* p = &L.part[S::poff + Q0 * PQ + h * PH + ci * PCI + cj * PCJ + fi];
#pragma unroll
            for (int q = 0; q < QN; ++q) v[d][q] = p[q * PQ];
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int fi = fi_lo + d;
            if (fi < 0 || fi >= FINE) continue;
#pragma unroll
            for (int q = 0; q < QN; ++q) acc.add(q, v[d][q]);
        }
    }
}

template <class R, int G, int Q0, int QN>
__device__ __forceinline__ void write_part(double* part, int tile, const NodeAcc<R>& acc,
                                           int lane) {
    using S = GridSpec<G>;
    double* p = part + ((size_t)tile * NSUM + S::slot0 + Q0) * 64 + lane;
#pragma unroll
    for (int q = 0; q < QN; ++q) p[(size_t)q * 64] = acc.get(q);
}

// Inputs of one marker, loaded one chunk ahead of their use so that the
// global-memory latency of the gather through perm overlaps the previous
// chunk's task and combine phases.  The two halves of the workgroup load
// half of each marker's inputs each (part 0: xm, ym, phim, sxxm, sxym,
// etavp, 1/G, cohesion, tension, friction; part 1: rhofluid, alphasolid,
// alphafluid and the 7 marker properties, which the sort's key pass
// evaluated in storage order).  Values are held as double (float inputs of
// the mixed form convert exactly).
constexpr int NRAW = 10;

template <class R>
__device__ __forceinline__ void load_raw(const ScatterArgs<R>& A, const ScatterLDS<R>& L,
                                         int part, uint32_t m, double (&v)[NRAW]) {
    if (part == 0) {
        v[0] = A.xm[m];
        v[1] = A.ym[m];
        v[2] = A.phim[m];
        v[3] = A.sxx[m];
        v[4] = A.sxy[m];
        v[5] = A.etavp[m];
        if (A.typed) {
            const int t = min(max(A.tm[m], 1), 3) - 1;
            v[6] = L.typed[0][t];
            v[7] = L.typed[1][t];
            v[8] = L.typed[2][t];
            v[9] = L.typed[3][t];
        } else {
            v[6] = A.invggg[m];
            v[7] = A.cohes[m];
            v[8] = A.tens[m];
            v[9] = A.frict[m];
        }
    } else {
        if (A.typed) {
            const int t = min(max(A.tm[m], 1), 3) - 1;
            v[0] = L.typed[4][t];
            v[1] = L.typed[5][t];
            v[2] = L.typed[6][t];
        } else {
            v[0] = A.rhof[m];
            v[1] = A.alphas[m];
            v[2] = A.alphaf[m];
        }
        v[3] = A.rhototal[m];
        v[4] = A.rhocptotal[m];
        v[5] = A.etatotal[m];
        v[6] = A.hrtotal[m];
        v[7] = A.ktotal[m];
        v[8] = A.tkrhocp[m];
        v[9] = A.etafkphi[m];
    }
}

// one half of a marker's staged values into LDS slot t: part 0 also the
// stencil fractions.  (j0, i0): 0-based first basic column and row of the
// tile.  Every marker of the tile has jb - 1, ib - 1 in [0, TILE) and
// jp - 1, ip - 1 in [0, TILE] relative to them (its fine key was derived
// from these cells).
template <class R>
__device__ __forceinline__ void stage_raw(const ScatterArgs<R>& A, ScatterLDS<R>& L, int part,
                                          int t, const double (&v)[NRAW], int j0, int i0) {
    if (part == 0) {
        const double x = v[0], y = v[1];
        const Cells cl = marker_cells(A.g, x, y);
        // fractions in double from the double coordinates, then rounded to R
        L.sfr[0][t] = (R)((x - L.ax[0][cl.jb - 1 - j0]) / A.g.dx);
        L.sfr[1][t] = (R)((y - L.ax[1][cl.ib - 1 - i0]) / A.g.dy);
        L.sfr[2][t] = (R)((x - L.ax[2][cl.jp - 1 - j0]) / A.g.dx);
        L.sfr[3][t] = (R)((y - L.ax[3][cl.ip - 1 - i0]) / A.g.dy);
        L.sf[L_PHIM][t] = (R)v[2];
        L.sf[L_SXX][t] = (R)v[3];
        L.sf[L_SXY][t] = (R)v[4];
        L.sf[L_ETAVP][t] = (R)v[5];
        L.sf[L_INVGGG][t] = (R)v[6];
        L.sf[L_COHES][t] = (R)v[7];
        L.sf[L_TENS][t] = (R)v[8];
        L.sf[L_FRICT][t] = (R)v[9];
    } else {
        L.sf[L_RHOF][t] = (R)v[0];
        L.sf[L_ALPHAS][t] = (R)v[1];
        L.sf[L_ALPHAF][t] = (R)v[2];
        L.sf[L_RHOTOTAL][t] = (R)v[3];
        L.sf[L_RHOCP][t] = (R)v[4];
        L.sf[L_ETATOTAL][t] = (R)v[5];
        L.sf[L_HR][t] = (R)v[6];
        L.sf[L_KTOTAL][t] = (R)v[7];
        L.sf[L_TKRHOCP][t] = (R)v[8];
        L.sf[L_ETAFKPHI][t] = (R)v[9];
    }
}

// Chunks of a tile: basic column k, markers [c0, cend) with cend - c0 <= CH.
// k == TILE marks the end.  All fields are wave-uniform.
struct Chunk {
    int k, c0, cend;
};

template <class R>
__device__ __forceinline__ Chunk chunk_from(const ScatterLDS<R>& L, int k, int c0) {
    for (; k < TILE; ++k) {
        const int ce = __builtin_amdgcn_readfirstlane(L.fs[(k + 1) * NCF]);
        if (c0 < ce) return Chunk{k, c0, min(c0 + CH, ce)};
        if (k + 1 < TILE) c0 = max(c0, __builtin_amdgcn_readfirstlane(L.fs[(k + 1) * NCF]));
    }
    return Chunk{TILE, 0, 0};
}

template <class R>
__device__ __forceinline__ Chunk chunk_next(const ScatterLDS<R>& L, const Chunk& c) {
    return c.k < TILE ? chunk_from(L, c.k, c.cend) : c;
}

template <class R>
__device__ __forceinline__ uint32_t chunk_perm(const ScatterArgs<R>& A, const Chunk& c, int t) {
    return (c.k < TILE && t < c.cend - c.c0) ? A.perm[c.c0 + t] : 0u;
}

// Wavefronts per workgroup: 8 in the Float64 form (two per grid, each thread
// loading half of a marker's inputs; 116 VGPRs, 4 waves per SIMD), 4 in the
// mixed form (one per grid, each thread loading a whole marker; its smaller
// LDS footprint already fits 4 workgroups per CU, and halving the work per
// thread there measured 36% slower).
template <class R> struct ScatterShape;
template <> struct ScatterShape<double> { static constexpr int NW = 8; };
template <> struct ScatterShape<float> { static constexpr int NW = 4; };

template <class R>
__global__ __launch_bounds__(ScatterShape<R>::NW * 64, 2) void k_scatter(ScatterArgs<R> A) {
    constexpr int NW = ScatterShape<R>::NW, NTHR = NW * 64;
    constexpr int NSET = NW == 8 ? 1 : 2;  // input halves per thread
    __shared__ ScatterLDS<R> L;
    // XCD-contiguous: vertically neighbouring tiles (consecutive numbers)
    // read adjacent runs of the same marker columns
    const int tile = (int)xcd_block();
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int part = NW == 8 ? wave >> 2 : 0, ts = t % CH;  // first input half, marker slot
    const int r = lane & 7, c = lane >> 3;
    const int j0 = (tile / A.nty) * TILE, i0 = (tile % A.nty) * TILE;
    for (int i = t; i <= NFINE; i += NTHR) L.fs[i] = A.fstart[(size_t)tile * NFINE + i];
    if (t < 21) L.typed[t / 3][t % 3] = A.tab.v[t / 3][t % 3];
    if (t < 4 * WIN) {
        const int a = t / WIN, k = t % WIN;
        const double* axis = a == 0 ? A.g.x : a == 1 ? A.g.y : a == 2 ? A.g.xp : A.g.yp;
        const int n = a == 0 ? A.g.Nx : a == 1 ? A.g.Ny : a == 2 ? A.g.Nx1 : A.g.Ny1;
        const int k0 = (a & 1) ? i0 : j0;
        L.ax[a][k] = k0 + k < n ? axis[k0 + k] : 0.0;
    }
    __syncthreads();
    NodeAcc<R> acc;
    acc.zero();

    Chunk cur = chunk_from(L, 0, __builtin_amdgcn_readfirstlane(L.fs[0]));
    Chunk nx = chunk_next(L, cur);
    double raw[NSET][NRAW];
    // the first chunk is staged before the loop; each iteration then stages
    // the next chunk while its combine runs (the combine reads the partials,
    // the staging writes the staged values), two barriers per chunk
    if (cur.k < TILE && ts < cur.cend - cur.c0) {
        const uint32_t m = chunk_perm(A, cur, ts);
#pragma unroll
        for (int q = 0; q < NSET; ++q) load_raw(A, L, part + q, m, raw[q]);
#pragma unroll
        for (int q = 0; q < NSET; ++q) stage_raw(A, L, part + q, ts, raw[q], j0, i0);
    }
    uint32_t pnx = chunk_perm(A, nx, ts);
    __syncthreads();
    while (cur.k < TILE) {
        // ---- 1. loads of the next chunk's markers, perm of the one after ----
        const Chunk nx2 = chunk_next(L, nx);
        if (nx.k < TILE && ts < nx.cend - nx.c0) {
#pragma unroll
            for (int q = 0; q < NSET; ++q) load_raw(A, L, part + q, pnx, raw[q]);
        }
        pnx = chunk_perm(A, nx2, ts);
        // ---- 2. (fine cell, corner column) partials: wavefronts g and g + 4
        // own the first and second half of grid g's sums ---------------------
        const int f0 = cur.k * NCF;
        if constexpr (NW == 4) {
            switch (wave) {
                case 0: tasks<R, 0, 0, 8>(L, lane, f0, cur.c0, cur.cend); break;
                case 1: tasks<R, 1, 0, 6>(L, lane, f0, cur.c0, cur.cend); break;
                case 2: tasks<R, 2, 0, 6>(L, lane, f0, cur.c0, cur.cend); break;
                default: tasks<R, 3, 0, 10>(L, lane, f0, cur.c0, cur.cend); break;
            }
        } else switch (wave) {
            case 0: tasks<R, 0, 0, 4>(L, lane, f0, cur.c0, cur.cend); break;
            case 1: tasks<R, 1, 0, 3>(L, lane, f0, cur.c0, cur.cend); break;
            case 2: tasks<R, 2, 0, 3>(L, lane, f0, cur.c0, cur.cend); break;
            case 3: tasks<R, 3, 0, 5>(L, lane, f0, cur.c0, cur.cend); break;
            case 4: tasks<R, 0, 4, 4>(L, lane, f0, cur.c0, cur.cend); break;
            case 5: tasks<R, 1, 3, 3>(L, lane, f0, cur.c0, cur.cend); break;
            case 6: tasks<R, 2, 3, 3>(L, lane, f0, cur.c0, cur.cend); break;
            default: tasks<R, 3, 5, 5>(L, lane, f0, cur.c0, cur.cend); break;
        }
        __syncthreads();
        // ---- 3. node-owned combine -----------------------------------------
        if constexpr (NW == 4) {
            switch (wave) {
                case 0: combine<R, 0, 0, 8>(L, acc, r, c, cur.k); break;
                case 1: combine<R, 1, 0, 6>(L, acc, r, c, cur.k); break;
                case 2: combine<R, 2, 0, 6>(L, acc, r, c, cur.k); break;
                default: combine<R, 3, 0, 10>(L, acc, r, c, cur.k); break;
            }
        } else switch (wave) {
            case 0: combine<R, 0, 0, 4>(L, acc, r, c, cur.k); break;
            case 1: combine<R, 1, 0, 3>(L, acc, r, c, cur.k); break;
            case 2: combine<R, 2, 0, 3>(L, acc, r, c, cur.k); break;
            case 3: combine<R, 3, 0, 5>(L, acc, r, c, cur.k); break;
            case 4: combine<R, 0, 4, 4>(L, acc, r, c, cur.k); break;
            case 5: combine<R, 1, 3, 3>(L, acc, r, c, cur.k); break;
            case 6: combine<R, 2, 3, 3>(L, acc, r, c, cur.k); break;
            default: combine<R, 3, 5, 5>(L, acc, r, c, cur.k); break;
        }
        // ---- 4. stage the next chunk (the staged values of this one are no
        // longer read) ---------------------------------------------------------
        if (nx.k < TILE && ts < nx.cend - nx.c0) {
#pragma unroll
            for (int q = 0; q < NSET; ++q) stage_raw(A, L, part + q, ts, raw[q], j0, i0);
        }
        __syncthreads();
        cur = nx;
        nx = nx2;
    }
    if constexpr (NW == 4) {
        switch (wave) {
            case 0: write_part<R, 0, 0, 8>(A.part, tile, acc, lane); break;
            case 1: write_part<R, 1, 0, 6>(A.part, tile, acc, lane); break;
            case 2: write_part<R, 2, 0, 6>(A.part, tile, acc, lane); break;
            default: write_part<R, 3, 0, 10>(A.part, tile, acc, lane); break;
        }
        return;
    }
    switch (wave) {
        case 0: write_part<R, 0, 0, 4>(A.part, tile, acc, lane); break;
        case 1: write_part<R, 1, 0, 3>(A.part, tile, acc, lane); break;
        case 2: write_part<R, 2, 0, 3>(A.part, tile, acc, lane); break;
        case 3: write_part<R, 3, 0, 5>(A.part, tile, acc, lane); break;
        case 4: write_part<R, 0, 4, 4>(A.part, tile, acc, lane); break;
        case 5: write_part<R, 1, 3, 3>(A.part, tile, acc, lane); break;
        case 6: write_part<R, 2, 3, 3>(A.part, tile, acc, lane); break;
        default: write_part<R, 3, 5, 5>(A.part, tile, acc, lane); break;
    }
}

struct SumPtrs {
    double* p[NSUM];
};

// SUM[s][node] = 0 + sum of the window partials covering the node, tiles in
// ascending (tx, ty) order
__global__ __launch_bounds__(256) void k_combine(const double* __restrict__ part, int nty, int ntx,
                                                 int Ny, int Nx, SumPtrs out) {
    const int s = blockIdx.y;
    const bool basic = s < 8;
    const int rows = basic ? Ny : Ny + 1, cols = basic ? Nx : Nx + 1;
    // node counts stay far below 2^31: 32-bit index arithmetic
    const int n = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (n >= rows * cols) return;
    const int i0 = n % rows, j0 = n / rows;
    out.p[s][n] = win_sum(part, nty, win_tiles(i0, j0, nty, ntx), s, i0, j0);
}

static const int kSumGrid[NSUM] = {
    EREB_G_ETA0SUM, EREB_G_ETASUM, EREB_G_GGGSUM, EREB_G_SXYSUM, EREB_G_COHSUM,
    EREB_G_TENSUM, EREB_G_FRISUM, EREB_G_WTSUM,
    EREB_G_RHOXSUM, EREB_G_RHOFXSUM, EREB_G_KXSUM, EREB_G_PHIXSUM, EREB_G_RXSUM, EREB_G_WTXSUM,
    EREB_G_RHOYSUM, EREB_G_RHOFYSUM, EREB_G_KYSUM, EREB_G_PHIYSUM, EREB_G_RYSUM, EREB_G_WTYSUM,
    EREB_G_RHOSUM, EREB_G_RHOCPSUM, EREB_G_ALPHASUM, EREB_G_ALPHAFSUM, EREB_G_HRSUM,
    EREB_G_GGGPSUM, EREB_G_SXXSUM, EREB_G_TKSUM, EREB_G_PHISUM, EREB_G_WTPSUM};

int sum_grid_field(int s) { return kSumGrid[s]; }

// flag[0] := 1 if any marker's seven material values differ (bitwise) from
// its type's table values (in the vectors' precision)
__device__ __forceinline__ bool same_bits(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ __forceinline__ bool same_bits(float a, double b) {
    return __float_as_uint(a) == __float_as_uint((float)b);
}

template <class R>
__global__ __launch_bounds__(256) void k_typed_check(ScatterArgs<R> A, int64_t M,
                                                     int32_t* __restrict__ flag) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const int t = min(max(A.tm[m], 1), 3) - 1;
    const R* v[7] = {A.invggg, A.cohes, A.tens, A.frict, A.rhof, A.alphas, A.alphaf};
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 7; ++q) ok = ok && same_bits(v[q][m], typed_at(A.tab, q, t));
    if (!ok) flag[0] = 1;
}

template <class R>
static int typed_check_t(ereb_ctx* c) {
    auto P = [&](int f) { return (const R*)c->d_m[f]; };
    ScatterArgs<R> A{};
    A.invggg = P(EREB_M_INV_GGGTOTALM);
    A.cohes = P(EREB_M_COHESTOTALM);
    A.tens = P(EREB_M_TENSTOTALM);
    A.frict = P(EREB_M_FRICTTOTALM);
    A.rhof = P(EREB_M_RHOFLUIDCUR);
    A.alphas = P(EREB_M_ALPHASOLIDCUR);
    A.alphaf = P(EREB_M_ALPHAFLUIDCUR);
    A.tm = tm_ptr(c);
    A.tab = typed_tab(c->mat);
    int rc = fill_i32(c, c->d_typed, 0, 1);
    if (rc) return rc;
    const int64_t M = c->marknum;
    if (M > 0)
        hipLaunchKernelGGL(k_typed_check<R>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                           c->stream, A, M, c->d_typed);
    EREB_HIP(c, hipGetLastError());
    int32_t flag = 1;
    EREB_HIP(c, hipMemcpyAsync(&flag, c->d_typed, sizeof flag, hipMemcpyDeviceToHost,
                               c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    c->typed = flag ? 0 : 1;
    return EREB_OK;
}

int typed_refresh(ereb_ctx* c) {
    if (c->typed >= 0) return EREB_OK;
    if (c->capturing) return fail(c, EREB_ERR_STATE, "material check inside a captured step");
    // the vectors are float exactly when the context is in the mixed form
    if (marker_is_f32(c->precision, EREB_M_INV_GGGTOTALM)) return typed_check_t<float>(c);
    return typed_check_t<double>(c);
}

template <class R>
static int launch_scatter_t(ereb_ctx* c) {
    auto P = [&](int f) { return (const R*)c->d_m[f]; };
    ScatterArgs<R> A;
    A.g = geomk(c);
    A.nty = c->nty;
    A.perm = c->d_perm;
    A.fstart = c->d_fstart;
    A.xm = mptr(c, EREB_M_XM);
    A.ym = mptr(c, EREB_M_YM);
    A.phim = mptr(c, EREB_M_PHIM);
    A.sxx = mptr(c, EREB_M_SXXM);
    A.sxy = mptr(c, EREB_M_SXYM);
    A.etavp = P(EREB_M_ETAVPM);
    A.invggg = P(EREB_M_INV_GGGTOTALM);
    A.frict = P(EREB_M_FRICTTOTALM);
    A.cohes = P(EREB_M_COHESTOTALM);
    A.tens = P(EREB_M_TENSTOTALM);
    A.rhof = P(EREB_M_RHOFLUIDCUR);
    A.alphas = P(EREB_M_ALPHASOLIDCUR);
    A.alphaf = P(EREB_M_ALPHAFLUIDCUR);
    A.rhototal = P(EREB_M_RHOTOTALM);
    A.rhocptotal = P(EREB_M_RHOCPTOTALM);
    A.etatotal = P(EREB_M_ETATOTALM);
    A.hrtotal = P(EREB_M_HRTOTALM);
    A.ktotal = P(EREB_M_KTOTALM);
    A.tkrhocp = P(EREB_M_TKM_RHOCPTOTALM);
    A.etafkphi = P(EREB_M_ETAFLUIDCUR_INV_KPHIM);
    A.part = c->d_part;
    if (c->typed < 0 && !c->capturing) {
        const int rc = typed_refresh(c);
        if (rc) return rc;
    }
    A.typed = c->typed == 1 ? 1 : 0;
    A.tm = tm_ptr(c);
    A.tab = typed_tab(c->mat);
    hipEvent_t ev;
    prof_begin(c, EREB_K_SCATTER, &ev);
    // the scatter's completion as ev_scatter when asked (rec_scatter)
    const dim3 grid((unsigned)c->ntiles), block(ScatterShape<R>::NW * 64);
    if (c->rec_scatter)
        hipExtLaunchKernelGGL(k_scatter<R>, grid, block, 0, c->stream, nullptr, c->ev_scatter, 0,
                              A);
    else
        hipLaunchKernelGGL(k_scatter<R>, grid, block, 0, c->stream, A);
    c->rec_scatter = false;
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_SCATTER, ev);
    return EREB_OK;
}

int launch_scatter(ereb_ctx* c, const double* hrs, const double* hrf, int mode) {
    // compute_marker_properties! (E:700-750) in the sort's pass over the markers
    const PropsReq req{hrs, hrf, mode};
    int rc = launch_fine_sort(c, &req);
    if (rc) return rc;
    // the property vectors are float exactly when the context is in the mixed form
    if (marker_is_f32(c->precision, EREB_M_ETATOTALM)) return launch_scatter_t<float>(c);
    return launch_scatter_t<double>(c);
}

int launch_combine(ereb_ctx* c) {
    SumPtrs o;
    for (int s = 0; s < NSUM; ++s) o.p[s] = gptr(c, kSumGrid[s]);
    const int64_t nmax = (int64_t)(c->Ny + 1) * (c->Nx + 1);
    hipEvent_t ev;
    prof_begin(c, EREB_K_COMBINE, &ev);
    hipLaunchKernelGGL(k_combine, dim3((unsigned)((nmax + 255) / 256), NSUM), dim3(256), 0,
                       c->stream, c->d_part, c->nty, c->ntx, c->Ny, c->Nx, o);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_COMBINE, ev);
    return EREB_OK;
}

}  // namespace ereb
