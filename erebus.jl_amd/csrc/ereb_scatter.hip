// ereb_scatter.hip -- step 1 + 2 of the transfer pipeline: marker properties
// (compute_marker_properties!, E:700-750) fused into the bilinear
// marker->node scatter onto the basic, Vx, Vy and P grids
// (marker_to_{basic,vx,vy,p}_nodes!, E:1605-1863, loop E:6419-6517).
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
// One workgroup (4 wavefronts) per tile, streaming the tile in chunks of at
// most CH markers and NFMAX fine cells:
//   1. load   : one thread per marker stages the 18 property values and the
//               fractions dxmj/dx, dymi/dy of the basic and P stencils in
//               LDS (the 7 marker properties are computed here and written
//               back to HBM);
//   2. tasks  : wavefront g owns grid g; a task is (fine cell, corner): the
//               sum over the fine cell's markers, in order, of
//               property * weight for that corner (and of the weight);
//   3. combine: lane (r, c) of wavefront g owns window node (r, c) of grid g
//               and adds the partials of the at most 16 fine cells around it
//               in a fixed (chunk, fine column, fine row) order.
// No atomics: every sum has exactly one owner and a fixed order, so results
// are identical run to run.  Tile windows overlap by two node rows/columns;
// k_combine adds the at most 2 x 2 window partials of each node in a fixed
// tile order.
#include "ereb_internal.h"

namespace ereb {

constexpr int CH = 256;     // markers per LDS chunk
constexpr int CHP = CH + 1;  // padded stride: fields of one marker in different banks
constexpr int NFMAX = 32;    // fine cells per chunk
constexpr int NLF = 18;      // property values staged per marker
constexpr int NPART = 120;   // partials per fine cell: 4 corners x (8 + 6 + 6 + 10)

// staged property indices
enum {
    L_ETATOTAL = 0, L_ETAVP, L_INVGGG, L_SXY, L_COHES, L_TENS, L_FRICT,  // basic
    L_RHOTOTAL, L_RHOF, L_KTOTAL, L_PHIM, L_ETAFKPHI,                    // Vx, Vy
    L_SXX, L_RHOCP, L_ALPHAS, L_ALPHAF, L_HR, L_TKRHOCP                  // P
};

struct ScatterArgs {
    GeomK g;
    ereb_materials M;
    double hrs[3], hrf[3];
    int mode, write_props, nty;
    const uint32_t* perm;
    const int32_t* fstart;
    const double *xm, *ym;
    const int32_t* tm;
    const double *tkm, *phim, *XW, *etavp, *invggg, *frict, *cohes, *tens, *rhof, *alphas,
        *alphaf, *sxx, *sxy;
    double *rhototal, *rhocptotal, *etatotal, *hrtotal, *ktotal, *tkrhocp, *etafkphi;
    double* part;
};

// per-grid SUM slot offset, partial offset and staged-property list (SUM
// order of reset_interpolated_properties!, E:968-999)
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
    static constexpr int slot0 = 8, poff = 32, nf = 5;
    static constexpr bool prow = true, pcol = false;
    __device__ static constexpr int f(int q) { return L_RHOTOTAL + q; }
};
template <> struct GridSpec<2> {  // Vy: RHOY RHOFY KY PHIY RY (+WTY)
    static constexpr int slot0 = 14, poff = 56, nf = 5;
    static constexpr bool prow = false, pcol = true;
    __device__ static constexpr int f(int q) { return L_RHOTOTAL + q; }
};
template <> struct GridSpec<3> {  // P: RHO RHOCP ALPHA ALPHAF HR GGGP SXX TK PHI (+WTP)
    static constexpr int slot0 = 20, poff = 80, nf = 9;
    static constexpr bool prow = true, pcol = true;
    __device__ static constexpr int f(int q) {
        return q == 0 ? L_RHOTOTAL : q == 1 ? L_RHOCP : q == 2 ? L_ALPHAS : q == 3 ? L_ALPHAF
             : q == 4 ? L_HR : q == 5 ? L_INVGGG : q == 6 ? L_SXX : q == 7 ? L_TKRHOCP
             : L_PHIM;
    }
};

struct ScatterLDS {
    double sf[NLF][CHP];
    double sfr[4][CHP];  // dxmj/dx basic, dymi/dy basic, dxmj/dx P, dymi/dy P
    double part[NFMAX][NPART];
    int32_t fs[NFINE + 1];
};

// tasks of grid G: (fine cell fl of the chunk, corner) for this wavefront
template <int G>
__device__ __forceinline__ void tasks(ScatterLDS& L, int lane, int f_lo, int nf, int c0,
                                      int cend) {
    using S = GridSpec<G>;
    const double* fa = S::pcol ? L.sfr[2] : L.sfr[0];
    const double* fb = S::prow ? L.sfr[3] : L.sfr[1];
    for (int task = lane; task < nf * 4; task += 64) {
        const int fl = task >> 2, corner = task & 3;
        const int ci = corner & 1, cj = corner >> 1;  // weights order (i,j),(i+1,j),(i,j+1),(i+1,j+1)
        const int f = f_lo + fl;
        const int k0 = max(L.fs[f], c0) - c0, k1 = min(L.fs[f + 1], cend) - c0;
        double acc[S::nf + 1];
#pragma unroll
        for (int q = 0; q <= S::nf; ++q) acc[q] = 0.0;
        for (int s = k0; s < k1; ++s) {
            const double a = fa[s], b = fb[s];
            const double w = (ci ? b : 1.0 - b) * (cj ? a : 1.0 - a);
#pragma unroll
            for (int q = 0; q < S::nf; ++q) acc[q] += L.sf[S::f(q)][s] * w;
            acc[S::nf] += w;
        }
        double* p = &L.part[fl][S::poff + corner * (S::nf + 1)];
#pragma unroll
        for (int q = 0; q <= S::nf; ++q) p[q] = acc[q];
    }
}

// node (r, c) of grid G adds the partials of its fine cells in this chunk
template <int G>
__device__ __forceinline__ void combine(const ScatterLDS& L, double* acc, int r, int c, int f_lo,
                                        int f_hi) {
    using S = GridSpec<G>;
    const int fi_lo = S::prow ? 2 * r - 3 : 2 * r - 2;
    const int fj_lo = S::pcol ? 2 * c - 3 : 2 * c - 2;
    for (int fj = max(0, fj_lo); fj <= min(FINE - 1, fj_lo + 3); ++fj) {
        const int cj = c - (S::pcol ? (fj + 1) >> 1 : fj >> 1);
        for (int fi = max(0, fi_lo); fi <= min(FINE - 1, fi_lo + 3); ++fi) {
            const int f = fj * FINE + fi;
            if (f < f_lo || f > f_hi) continue;
            const int ci = r - (S::prow ? (fi + 1) >> 1 : fi >> 1);
            const double* p = &L.part[f - f_lo][S::poff + (ci + 2 * cj) * (S::nf + 1)];
#pragma unroll
            for (int q = 0; q <= S::nf; ++q) acc[q] += p[q];
        }
    }
}

template <int G>
__device__ __forceinline__ void write_part(double* part, int tile, const double* acc, int lane) {
    using S = GridSpec<G>;
    double* p = part + ((size_t)tile * NSUM + S::slot0) * 64 + lane;
#pragma unroll
    for (int q = 0; q <= S::nf; ++q) p[(size_t)q * 64] = acc[q];
}

__global__ __launch_bounds__(256, 2) void k_scatter_fp64(ScatterArgs A) {
    __shared__ ScatterLDS L;
    const int tile = blockIdx.x;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int r = lane & 7, c = lane >> 3;
    for (int i = t; i <= NFINE; i += 256) L.fs[i] = A.fstart[(size_t)tile * NFINE + i];
    __syncthreads();
    const int S0 = L.fs[0], E0 = L.fs[NFINE];
    double acc[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] = 0.0;

    int f_lo = 0;
    for (int c0 = S0; c0 < E0;) {
        // chunk [c0, cend): at most CH markers and NFMAX fine cells
        while (L.fs[f_lo + 1] <= c0) ++f_lo;
        int cend = min(c0 + CH, E0);
        if (f_lo + NFMAX < NFINE) cend = min(cend, L.fs[f_lo + NFMAX]);
        if (cend <= c0) cend = min(c0 + CH, E0);  // a single fine cell longer than CH
        int f_hi = f_lo;
        while (f_hi + 1 < NFINE && L.fs[f_hi + 1] < cend) ++f_hi;
        const int nf = f_hi - f_lo + 1;
        // ---- 1. load ------------------------------------------------------
        if (t < cend - c0) {
            const uint32_t m = A.perm[c0 + t];
            const double x = A.xm[m], y = A.ym[m];
            const Cells cl = marker_cells(A.g, x, y);
            L.sfr[0][t] = (x - A.g.x[cl.jb - 1]) / A.g.dx;
            L.sfr[1][t] = (y - A.g.y[cl.ib - 1]) / A.g.dy;
            L.sfr[2][t] = (x - A.g.xp[cl.jp - 1]) / A.g.dx;
            L.sfr[3][t] = (y - A.g.yp[cl.ip - 1]) / A.g.dy;
            const int tt = min(max(A.tm[m], 1), 3);  // set validates 1..3
            const double T = A.tkm[m], phi = A.phim[m];
            Props st;
            st.rhototal = A.rhototal[m];
            st.rhocptotal = A.rhocptotal[m];
            st.etatotal = A.etatotal[m];
            st.hrtotal = A.hrtotal[m];
            st.ktotal = A.ktotal[m];
            st.tkrhocp = 0.0;
            st.etafkphi = 0.0;
            const Props p = marker_props(A.M, tt, T, phi, A.XW[m], A.hrs, A.hrf, A.mode, st);
            if (A.write_props) {
                if (tt < 3) {
                    A.rhototal[m] = p.rhototal;
                    A.rhocptotal[m] = p.rhocptotal;
                    A.etatotal[m] = p.etatotal;
                    A.hrtotal[m] = p.hrtotal;
                    A.ktotal[m] = p.ktotal;
                }
                A.tkrhocp[m] = p.tkrhocp;
                A.etafkphi[m] = p.etafkphi;
            }
            L.sf[L_ETATOTAL][t] = p.etatotal;
            L.sf[L_ETAVP][t] = A.etavp[m];
            L.sf[L_INVGGG][t] = A.invggg[m];
            L.sf[L_SXY][t] = A.sxy[m];
            L.sf[L_COHES][t] = A.cohes[m];
            L.sf[L_TENS][t] = A.tens[m];
            L.sf[L_FRICT][t] = A.frict[m];
            L.sf[L_RHOTOTAL][t] = p.rhototal;
            L.sf[L_RHOF][t] = A.rhof[m];
            L.sf[L_KTOTAL][t] = p.ktotal;
            L.sf[L_PHIM][t] = phi;
            L.sf[L_ETAFKPHI][t] = p.etafkphi;
            L.sf[L_SXX][t] = A.sxx[m];
            L.sf[L_RHOCP][t] = p.rhocptotal;
            L.sf[L_ALPHAS][t] = A.alphas[m];
            L.sf[L_ALPHAF][t] = A.alphaf[m];
            L.sf[L_HR][t] = p.hrtotal;
            L.sf[L_TKRHOCP][t] = p.tkrhocp;
        }
        __syncthreads();
        // ---- 2. per (fine cell, corner) partial sums, wavefront g = grid g --
        switch (wave) {
            case 0: tasks<0>(L, lane, f_lo, nf, c0, cend); break;
            case 1: tasks<1>(L, lane, f_lo, nf, c0, cend); break;
            case 2: tasks<2>(L, lane, f_lo, nf, c0, cend); break;
            default: tasks<3>(L, lane, f_lo, nf, c0, cend); break;
        }
        __syncthreads();
        // ---- 3. node-owned combine -----------------------------------------
        switch (wave) {
            case 0: combine<0>(L, acc, r, c, f_lo, f_hi); break;
            case 1: combine<1>(L, acc, r, c, f_lo, f_hi); break;
            case 2: combine<2>(L, acc, r, c, f_lo, f_hi); break;
            default: combine<3>(L, acc, r, c, f_lo, f_hi); break;
        }
        __syncthreads();
        c0 = cend;
    }
    switch (wave) {
        case 0: write_part<0>(A.part, tile, acc, lane); break;
        case 1: write_part<1>(A.part, tile, acc, lane); break;
        case 2: write_part<2>(A.part, tile, acc, lane); break;
        default: write_part<3>(A.part, tile, acc, lane); break;
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
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)rows * cols) return;
    const int i0 = (int)(n % rows), j0 = (int)(n / rows);
    const int tylo = max(0, (i0 - (WIN - 1) + TILE - 1) / TILE), tyhi = min(nty - 1, i0 / TILE);
    const int txlo = max(0, (j0 - (WIN - 1) + TILE - 1) / TILE), txhi = min(ntx - 1, j0 / TILE);
    double acc = 0.0;
    for (int tx = txlo; tx <= txhi; ++tx)
        for (int ty = tylo; ty <= tyhi; ++ty) {
            const int lane = (j0 - tx * TILE) * 8 + (i0 - ty * TILE);
            acc += part[((size_t)(tx * nty + ty) * NSUM + s) * 64 + lane];
        }
    out.p[s][n] = acc;
}

static const int kSumGrid[NSUM] = {
    EREB_G_ETA0SUM, EREB_G_ETASUM, EREB_G_GGGSUM, EREB_G_SXYSUM, EREB_G_COHSUM,
    EREB_G_TENSUM, EREB_G_FRISUM, EREB_G_WTSUM,
    EREB_G_RHOXSUM, EREB_G_RHOFXSUM, EREB_G_KXSUM, EREB_G_PHIXSUM, EREB_G_RXSUM, EREB_G_WTXSUM,
    EREB_G_RHOYSUM, EREB_G_RHOFYSUM, EREB_G_KYSUM, EREB_G_PHIYSUM, EREB_G_RYSUM, EREB_G_WTYSUM,
    EREB_G_RHOSUM, EREB_G_RHOCPSUM, EREB_G_ALPHASUM, EREB_G_ALPHAFSUM, EREB_G_HRSUM,
    EREB_G_GGGPSUM, EREB_G_SXXSUM, EREB_G_TKSUM, EREB_G_PHISUM, EREB_G_WTPSUM};

int launch_scatter(ereb_ctx* c, const double* hrs, const double* hrf, int mode, bool write_props) {
    if (c->precision != EREB_FP64)
        return fail(c, EREB_ERR_ARG, "mixed-precision scatter is not available in this build");
    ScatterArgs A;
    A.g = geomk(c);
    A.M = c->mat;
    for (int k = 0; k < 3; ++k) {
        A.hrs[k] = hrs[k];
        A.hrf[k] = hrf[k];
    }
    A.mode = mode;
    A.write_props = write_props ? 1 : 0;
    A.nty = c->nty;
    A.perm = c->d_perm;
    A.fstart = c->d_fstart;
    A.xm = mptr(c, EREB_M_XM);
    A.ym = mptr(c, EREB_M_YM);
    A.tm = tm_ptr(c);
    A.tkm = mptr(c, EREB_M_TKM);
    A.phim = mptr(c, EREB_M_PHIM);
    A.XW = mptr(c, EREB_M_XWSOLIDM0);
    A.etavp = mptr(c, EREB_M_ETAVPM);
    A.invggg = mptr(c, EREB_M_INV_GGGTOTALM);
    A.frict = mptr(c, EREB_M_FRICTTOTALM);
    A.cohes = mptr(c, EREB_M_COHESTOTALM);
    A.tens = mptr(c, EREB_M_TENSTOTALM);
    A.rhof = mptr(c, EREB_M_RHOFLUIDCUR);
    A.alphas = mptr(c, EREB_M_ALPHASOLIDCUR);
    A.alphaf = mptr(c, EREB_M_ALPHAFLUIDCUR);
    A.sxx = mptr(c, EREB_M_SXXM);
    A.sxy = mptr(c, EREB_M_SXYM);
    A.rhototal = mptr(c, EREB_M_RHOTOTALM);
    A.rhocptotal = mptr(c, EREB_M_RHOCPTOTALM);
    A.etatotal = mptr(c, EREB_M_ETATOTALM);
    A.hrtotal = mptr(c, EREB_M_HRTOTALM);
    A.ktotal = mptr(c, EREB_M_KTOTALM);
    A.tkrhocp = mptr(c, EREB_M_TKM_RHOCPTOTALM);
    A.etafkphi = mptr(c, EREB_M_ETAFLUIDCUR_INV_KPHIM);
    A.part = c->d_part;
    hipEvent_t ev;
    prof_begin(c, EREB_K_SCATTER, &ev);
    hipLaunchKernelGGL(k_scatter_fp64, dim3((unsigned)c->ntiles), dim3(256), 0, c->stream, A);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_SCATTER, ev);
    return EREB_OK;
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
