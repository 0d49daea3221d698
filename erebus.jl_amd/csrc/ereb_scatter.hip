// ereb_scatter.hip -- step 1 + 2 of the transfer pipeline: marker properties
// (compute_marker_properties!, E:700-750) fused into the bilinear
// marker->node scatter onto the basic, Vx, Vy and P grids
// (marker_to_{basic,vx,vy,p}_nodes!, E:1605-1863, loop E:6419-6517).
//
// Layout and algorithm (one workgroup of 4 wavefronts per tile of
// TILE x TILE basic cells, markers taken in tile-sorted order):
//   - A marker's four stencils follow from two cells: basic (ib, jb) and
//     P (ip, jp); Vx uses (ip, jb), Vy uses (ib, jp) because the staggered
//     axes coincide (C:40-64) and so do the clamps (C:66-101).  The pair
//     (ip-ib, jp-jb) is in {0,1}^2, so every marker of a tile lands in a
//     (TILE+2)^2 = 64-node window of each grid, and its "fine cell"
//     (2*(ib-i0) + ip-ib, 2*(jb-j0) + jp-jb) fixes all 16 target nodes.
//   - Chunks of CH markers are staged in LDS (18 property values and the
//     four fractions dxmj/dx, dymi/dy of the basic and P stencils), then
//     grouped by fine cell with a stable in-LDS counting sort.
//   - Wavefront g owns grid g; lane (r, c) owns window node (r, c) and sums,
//     in a fixed order (chunk, fine column, fine row, marker), the products
//     property*weight of the 16 fine cells around it.  No atomics: each
//     window sum has one owner, so results are identical run to run.
//   - Tile windows overlap by two node rows/columns; k_combine adds the at
//     most 2 x 2 window partials of every node in a fixed tile order.
#include "ereb_internal.h"

namespace ereb {

constexpr int CH = 256;     // markers per LDS chunk
constexpr int NLF = 18;     // property values staged per marker
constexpr int NFINE = FINE * FINE;

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
    const int32_t* tstart;
    const double *xm, *ym;
    const int32_t* tm;
    const double *tkm, *phim, *XW, *etavp, *invggg, *frict, *cohes, *tens, *rhof, *alphas,
        *alphaf, *sxx, *sxy;
    double *rhototal, *rhocptotal, *etatotal, *hrtotal, *ktotal, *tkrhocp, *etafkphi;
    double* part;
};

// per-grid SUM slot offset and staged-property list (SUM order of
// reset_interpolated_properties!, E:968-999)
template <int G> struct GridSpec;
template <> struct GridSpec<0> {  // basic: ETA0 ETA GGG SXY COH TEN FRI (+WT)
    static constexpr int slot0 = 0, nf = 7;
    static constexpr bool prow = false, pcol = false;
    __device__ static constexpr int f(int q) {
        return q == 0 ? L_ETATOTAL : q == 1 ? L_ETAVP : q == 2 ? L_INVGGG : q == 3 ? L_SXY
             : q == 4 ? L_COHES : q == 5 ? L_TENS : L_FRICT;
    }
};
template <> struct GridSpec<1> {  // Vx: RHOX RHOFX KX PHIX RX (+WTX)
    static constexpr int slot0 = 8, nf = 5;
    static constexpr bool prow = true, pcol = false;
    __device__ static constexpr int f(int q) { return L_RHOTOTAL + q; }
};
template <> struct GridSpec<2> {  // Vy: RHOY RHOFY KY PHIY RY (+WTY)
    static constexpr int slot0 = 14, nf = 5;
    static constexpr bool prow = false, pcol = true;
    __device__ static constexpr int f(int q) { return L_RHOTOTAL + q; }
};
template <> struct GridSpec<3> {  // P: RHO RHOCP ALPHA ALPHAF HR GGGP SXX TK PHI (+WTP)
    static constexpr int slot0 = 20, nf = 9;
    static constexpr bool prow = true, pcol = true;
    __device__ static constexpr int f(int q) {
        return q == 0 ? L_RHOTOTAL : q == 1 ? L_RHOCP : q == 2 ? L_ALPHAS : q == 3 ? L_ALPHAF
             : q == 4 ? L_HR : q == 5 ? L_INVGGG : q == 6 ? L_SXX : q == 7 ? L_TKRHOCP
             : L_PHIM;
    }
};

struct ScatterLDS {
    double sf[NLF][CH];
    double sfr[4][CH];  // dxmj/dx basic, dymi/dy basic, dxmj/dx P, dymi/dy P
    int32_t fstart[NFINE + 1];
    int32_t wbase[4][NFINE];
    int16_t fine[CH];
    int16_t order[CH];
};

template <int G>
__device__ __forceinline__ void accumulate(const ScatterLDS& L, double* acc, int r, int c) {
    using S = GridSpec<G>;
    const int fi_lo = S::prow ? 2 * r - 3 : 2 * r - 2;
    const int fj_lo = S::pcol ? 2 * c - 3 : 2 * c - 2;
    const double* fa = S::pcol ? L.sfr[2] : L.sfr[0];
    const double* fb = S::prow ? L.sfr[3] : L.sfr[1];
    for (int fj = max(0, fj_lo); fj <= min(FINE - 1, fj_lo + 3); ++fj) {
        const int cj = c - (S::pcol ? (fj + 1) >> 1 : fj >> 1);
        for (int fi = max(0, fi_lo); fi <= min(FINE - 1, fi_lo + 3); ++fi) {
            const int ci = r - (S::prow ? (fi + 1) >> 1 : fi >> 1);
            const int fc = fj * FINE + fi;
            const int k1 = L.fstart[fc + 1];
            for (int k = L.fstart[fc]; k < k1; ++k) {
                const int s = L.order[k];
                const double a = fa[s], b = fb[s];
                const double wy = ci ? b : 1.0 - b;
                const double wx = cj ? a : 1.0 - a;
                const double w = wy * wx;
#pragma unroll
                for (int q = 0; q < S::nf; ++q) acc[q] += L.sf[S::f(q)][s] * w;
                acc[S::nf] += w;
            }
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

__global__ __launch_bounds__(256) void k_scatter_fp64(ScatterArgs A) {
    __shared__ ScatterLDS L;
    const int tile = blockIdx.x;
    const int ty = tile % A.nty, tx = tile / A.nty;
    const int i0 = ty * TILE, j0 = tx * TILE;  // first basic cell (0-based)
    const int S0 = A.tstart[tile], E0 = A.tstart[tile + 1];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int r = lane & 7, c = lane >> 3;
    double acc[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] = 0.0;

    for (int c0 = S0; c0 < E0; c0 += CH) {
        const int n = min(CH, E0 - c0);
        int f = -1;
        if (t < n) {
            const uint32_t m = A.perm[c0 + t];
            const double x = A.xm[m], y = A.ym[m];
            const Cells cl = marker_cells(A.g, x, y);
            const int fi = 2 * (cl.ib - 1 - i0) + (cl.ip - cl.ib);
            const int fj = 2 * (cl.jb - 1 - j0) + (cl.jp - cl.jb);
            if (fi >= 0 && fi < FINE && fj >= 0 && fj < FINE) f = fj * FINE + fi;
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
        // ---- stable grouping of the chunk by fine cell -----------------
        for (int i = t; i < 4 * NFINE; i += 256) (&L.wbase[0][0])[i] = 0;
        __syncthreads();
        int rank = 0;
        {
            uint64_t todo = __ballot(f >= 0);
            while (todo) {
                const int leader = __ffsll((unsigned long long)todo) - 1;
                const int lf = __shfl(f, leader);
                const uint64_t mm = __ballot(f == lf);
                if (f == lf) rank = (int)__popcll(mm & lanemask_lt());
                if (lane == leader) L.wbase[wave][lf] = (int)__popcll(mm);
                todo &= ~mm;
            }
        }
        __syncthreads();
        // fine-cell totals -> starts (wave 0: 3 fine cells per lane), and
        // per-wave bases inside each fine cell
        if (wave == 0) {
            int tot[3], s = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int fc = lane * 3 + k;
                tot[k] = fc < NFINE ? L.wbase[0][fc] + L.wbase[1][fc] + L.wbase[2][fc] +
                                          L.wbase[3][fc]
                                    : 0;
                s += tot[k];
            }
            int incl = s;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            int run = incl - s;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int fc = lane * 3 + k;
                if (fc < NFINE) {
                    L.fstart[fc] = run;
                    int b = run;
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        const int cw = L.wbase[w][fc];
                        L.wbase[w][fc] = b;
                        b += cw;
                    }
                }
                run += tot[k];
            }
            if (lane == 63) L.fstart[NFINE] = incl;
        }
        __syncthreads();
        if (f >= 0) L.order[L.wbase[wave][f] + rank] = (int16_t)t;
        __syncthreads();
        // ---- node-owned accumulation --------------------------------------
        switch (wave) {
            case 0: accumulate<0>(L, acc, r, c); break;
            case 1: accumulate<1>(L, acc, r, c); break;
            case 2: accumulate<2>(L, acc, r, c); break;
            default: accumulate<3>(L, acc, r, c); break;
        }
        __syncthreads();
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
    A.tstart = c->d_tstart;
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
