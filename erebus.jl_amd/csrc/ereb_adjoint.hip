// ereb_adjoint.hip -- reverse-mode (adjoint) kernels of the transfer path
// with respect to the marker temperature tkm and porosity phim.  The
// reference has no adjoint; these are the transposes of the forward kernels:
//   scatter^T  (marker_to_*_nodes!, E:1605-1863, with the property
//              derivatives of compute_marker_properties!, E:700-750): a
//              gather -- each marker reads the SUM cotangents at its 16
//              nodes, no atomics;
//   nodes^T    (compute_*_node_properties!, E:1970-2189, and the insulating
//              boundary copy on tk1, E:2432-2449).
// The gather^T and RK4^T kernels sit in ereb_markers.hip next to their
// forward kernels.  Marker positions and the grid weights do not depend on
// tkm or phim (the velocity is prescribed), so they carry no cotangent.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ereb_internal.h"

namespace ereb {

struct ScatterAdjArgs {
    GeomK g;
    const ereb_materials* __restrict__ M;
    double hrs[3], hrf[3];
    int mode;
    int64_t n;
    const double *xm, *ym, *tkm, *phim, *XW;
    const int32_t* tm;
    const double *st_rho, *st_rhocp, *st_eta, *st_hr, *st_k;  // stored properties (air)
    const double* gs;  // cotangents of the 30 SUM grids, concatenated (basic Ny*Nx, then Ny1*Nx1)
    int64_t goff[NSUM];
    double *gT, *gphi;  // accumulated (+=)
};

__device__ __forceinline__ double gdot4_at(const double* gr, int rows, int i, int j,
                                           const double* w) {
    return gdot4(gr, rows, i, j, w);
}

// scatter^T: one thread per marker (storage order).  FULL = false reads only
// the RHOCPSUM and TKSUM cotangents (all others zero), as for a loss that
// depends on the sums through tk1 alone.
template <bool FULL>
__global__ __launch_bounds__(256) void k_scatter_adj(ScatterAdjArgs A) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= A.n) return;
    const GeomK& g = A.g;
    const ereb_materials& M = *A.M;
    const double x = A.xm[m], y = A.ym[m], T = A.tkm[m], phi = A.phim[m];
    const int tt = min(max(A.tm[m], 1), 3);
    const Cells cl = marker_cells(g, x, y);
    const double ab = (x - g.x[cl.jb - 1]) / g.dx, bb = (y - g.y[cl.ib - 1]) / g.dy;
    const double ap = (x - g.xp[cl.jp - 1]) / g.dx, bp = (y - g.yp[cl.ip - 1]) / g.dy;
    double wvx[4], wvy[4], wp[4];
    weights4(ab, bp, wvx);  // Vx cell (ip, jb)
    weights4(ap, bb, wvy);  // Vy cell (ib, jp)
    weights4(ap, bp, wp);   // P cell (ip, jp)
    Props st{A.st_rho[m], A.st_rhocp[m], A.st_eta[m], A.st_hr[m], A.st_k[m], 0.0, 0.0};
    const Props p = marker_props(M, tt, T, phi, A.XW[m], A.hrs, A.hrf, A.mode, st);
    const PropsD d = marker_props_d(M, tt, T, phi, A.XW[m], A.hrs, A.hrf, A.mode, p);
    auto G = [&](int slot) { return A.gs + A.goff[slot]; };
    const int R = g.Ny1;
    double gT = 0.0, gp = 0.0;
    if (!FULL) {
        const double grc = gdot4_at(G(21), R, cl.ip, cl.jp, wp);
        const double gtk = gdot4_at(G(27), R, cl.ip, cl.jp, wp);
        A.gT[m] += grc * d.rhocp_t + gtk * d.tk_t;
        A.gphi[m] += grc * d.rhocp_p + gtk * d.tk_p;
        return;
    }
    // Vx (slots 8..13: RHO RHOF K PHI R WT) and Vy (14..19)
    for (int v = 0; v < 2; ++v) {
        const int s0 = v == 0 ? 8 : 14;
        const double* w = v == 0 ? wvx : wvy;
        const int i = v == 0 ? cl.ip : cl.ib, j = v == 0 ? cl.jb : cl.jp;
        const double grho = gdot4_at(G(s0 + 0), R, i, j, w);
        const double gk = gdot4_at(G(s0 + 2), R, i, j, w);
        const double gph = gdot4_at(G(s0 + 3), R, i, j, w);
        const double gr = gdot4_at(G(s0 + 4), R, i, j, w);
        gT += gk * d.k_t;
        gp += grho * d.rho_p + gk * d.k_p + gph + gr * d.ef_p;
    }
    // P (slots 20..29: RHO RHOCP ALPHA ALPHAF HR GGGP SXX TK PHI WT)
    {
        const double grho = gdot4_at(G(20), R, cl.ip, cl.jp, wp);
        const double grc = gdot4_at(G(21), R, cl.ip, cl.jp, wp);
        const double ghr = gdot4_at(G(24), R, cl.ip, cl.jp, wp);
        const double gtk = gdot4_at(G(27), R, cl.ip, cl.jp, wp);
        const double gph = gdot4_at(G(28), R, cl.ip, cl.jp, wp);
        gT += grc * d.rhocp_t + gtk * d.tk_t;
        gp += grho * d.rho_p + grc * d.rhocp_p + ghr * d.hr_p + gtk * d.tk_p + gph;
    }
    A.gT[m] += gT;
    A.gphi[m] += gp;
}

// nodes^T on the P grid: cotangents of RHO RHOCP ALPHA ALPHAF HR GGGP SXX0 tk1
// PHI BETAPHI -> cotangents of the 10 P SUM grids (WTPSUM: the weights carry
// no cotangent).  Zero-weight nodes keep their old values in the forward
// pass, so their SUM cotangents are zero.
struct NodesAdjArgs {
    int n;  // Ny1 * Nx1
    const double *RHOCPSUM, *TKSUM, *GGGPSUM, *PHISUM, *WTPSUM;
    const double* gnode[10];  // RHO RHOCP ALPHA ALPHAF HR GGGP SXX0 tk1 PHI BETAPHI
    double* gsum[10];         // RHOSUM ... PHISUM, WTPSUM (set to 0)
};

__device__ __forceinline__ void nodes_adj_p(const NodesAdjArgs& A, int k) {
    const double wt = A.WTPSUM[k];
    double out[10] = {};
    if (wt > 0.0) {
        const double iw = 1.0 / wt;
        const double* const* gn = A.gnode;
        out[0] = gn[0][k] * iw;  // RHO
        out[2] = gn[2][k] * iw;  // ALPHA
        out[3] = gn[3][k] * iw;  // ALPHAF
        out[4] = gn[4][k] * iw;  // HR
        out[6] = gn[6][k] * iw;  // SXX0
        // tk1 = TKSUM * inv(RHOCPSUM); RHOCP = RHOCPSUM * iw
        const double irc = 1.0 / A.RHOCPSUM[k];
        out[7] = gn[7][k] * irc;
        out[1] = gn[1][k] * iw - gn[7][k] * A.TKSUM[k] * irc * irc;
        // GGGP = inv(GGGPSUM) * wt; PHI = PHISUM * iw; BETAPHI = inv(GGGP) * PHI
        const double gggp = (1.0 / A.GGGPSUM[k]) * wt, ph = A.PHISUM[k] * iw;
        const double g_gggp = gn[5][k] - gn[9][k] * ph / (gggp * gggp);
        const double g_phi = gn[8][k] + gn[9][k] / gggp;
        out[5] = -g_gggp * wt / (A.GGGPSUM[k] * A.GGGPSUM[k]);
        out[8] = g_phi * iw;
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) A.gsum[q][k] = out[q];
}

// nodes^T on a Vx or Vy grid: RHO RHOF K PHI R = SUM * iw
struct NodesAdjV {
    const double* wts;   // the grid's weight sum
    const double* g[5];  // cotangents of RHO RHOF K PHI R
    double* s[6];        // of their SUM grids and of the weight sum (set to 0)
};

__device__ __forceinline__ void nodes_adj_v(const double* __restrict__ wts,
                                            const double* const (&g)[5], double* const (&s)[6],
                                            int k) {
    const double wt = wts[k];
    const double iw = wt > 0.0 ? 1.0 / wt : 0.0;
#pragma unroll
    for (int q = 0; q < 5; ++q) s[q][k] = g[q][k] * iw;
    s[5][k] = 0.0;
}

// grid row 0: the P grid, 1: Vx, 2: Vy
__global__ __launch_bounds__(256) void k_nodes_adj(NodesAdjArgs A, NodesAdjV X, NodesAdjV Y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.n) return;
    if (blockIdx.y == 0) nodes_adj_p(A, k);
    else if (blockIdx.y == 1) nodes_adj_v(X.wts, X.g, X.s, k);
    else nodes_adj_v(Y.wts, Y.g, Y.s, k);
}

// insulating boundary copy on a (Nyy, Nxx) array, transposed: the forward
// pass copies rows 2 and Nyy-1 into rows 1 and Nyy (columns 2..Nxx-1), then
// columns 2 and Nxx-1 into columns 1 and Nxx (all rows); the adjoint undoes
// the column step first, then the row step.
__global__ void k_bc_adj_cols(double* t, int Nyy, int Nxx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nyy) return;
    double* c1 = t + i;
    c1[(size_t)1 * Nyy] += c1[0];
    c1[0] = 0.0;
    c1[(size_t)(Nxx - 2) * Nyy] += c1[(size_t)(Nxx - 1) * Nyy];
    c1[(size_t)(Nxx - 1) * Nyy] = 0.0;
}

__global__ void k_bc_adj_rows(double* t, int Nyy, int Nxx) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;  // columns 2..Nxx-1 (0-based 1..)
    if (j >= Nxx - 1) return;
    double* col = t + (size_t)j * Nyy;
    col[1] += col[0];
    col[0] = 0.0;
    col[Nyy - 2] += col[Nyy - 1];
    col[Nyy - 1] = 0.0;
}

int launch_scatter_adj(ereb_ctx* c, const double* hrs, const double* hrf, int mode,
                       const double* gsum_base, const int64_t* gsum_off, double* gT,
                       double* gphi, bool only_tk) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    if (c->precision != EREB_FP64)
        return fail(c, EREB_ERR_ARG, "the adjoint kernels run in the Float64 form");
    ScatterAdjArgs A;
    A.g = geomk(c);
    A.M = c->d_mat;
    for (int k = 0; k < 3; ++k) {
        A.hrs[k] = hrs[k];
        A.hrf[k] = hrf[k];
    }
    A.mode = mode;
    A.n = M;
    A.xm = mptr(c, EREB_M_XM);
    A.ym = mptr(c, EREB_M_YM);
    A.tkm = mptr(c, EREB_M_TKM);
    A.phim = mptr(c, EREB_M_PHIM);
    A.XW = mptr(c, EREB_M_XWSOLIDM0);
    A.tm = tm_ptr(c);
    A.st_rho = mptr(c, EREB_M_RHOTOTALM);
    A.st_rhocp = mptr(c, EREB_M_RHOCPTOTALM);
    A.st_eta = mptr(c, EREB_M_ETATOTALM);
    A.st_hr = mptr(c, EREB_M_HRTOTALM);
    A.st_k = mptr(c, EREB_M_KTOTALM);
    A.gs = gsum_base;
    for (int s = 0; s < NSUM; ++s) A.goff[s] = gsum_off[s];
    A.gT = gT;
    A.gphi = gphi;
    hipLaunchKernelGGL(only_tk ? k_scatter_adj<false> : k_scatter_adj<true>,
                       dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream, A);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

int launch_nodes_adj(ereb_ctx* c, double* gbase, const int64_t* off) {
    // off: offsets (in doubles from gbase) of the grid fields, by ereb_grid_field
    auto G = [&](int f) { return gbase + off[f]; };
    const int n = c->Nx1 * c->Ny1;
    const unsigned nb = (unsigned)((n + 255) / 256);
    // tk1 boundary copy first (it was applied last in the forward pass)
    if (c->Ny1 > 2 && c->Nx1 > 2) {
        hipLaunchKernelGGL(k_bc_adj_cols, dim3((unsigned)((c->Ny1 + 255) / 256)), dim3(256), 0,
                           c->stream, G(EREB_G_TK1), c->Ny1, c->Nx1);
        hipLaunchKernelGGL(k_bc_adj_rows, dim3((unsigned)((c->Nx1 + 255) / 256)), dim3(256), 0,
                           c->stream, G(EREB_G_TK1), c->Ny1, c->Nx1);
    }
    NodesAdjArgs A;
    A.n = n;
    A.RHOCPSUM = gptr(c, EREB_G_RHOCPSUM);
    A.TKSUM = gptr(c, EREB_G_TKSUM);
    A.GGGPSUM = gptr(c, EREB_G_GGGPSUM);
    A.PHISUM = gptr(c, EREB_G_PHISUM);
    A.WTPSUM = gptr(c, EREB_G_WTPSUM);
    const int pin[10] = {EREB_G_RHO, EREB_G_RHOCP, EREB_G_ALPHA, EREB_G_ALPHAF, EREB_G_HR,
                         EREB_G_GGGP, EREB_G_SXX0, EREB_G_TK1, EREB_G_PHI, EREB_G_BETAPHI};
    const int pout[10] = {EREB_G_RHOSUM, EREB_G_RHOCPSUM, EREB_G_ALPHASUM, EREB_G_ALPHAFSUM,
                          EREB_G_HRSUM, EREB_G_GGGPSUM, EREB_G_SXXSUM, EREB_G_TKSUM,
                          EREB_G_PHISUM, EREB_G_WTPSUM};
    for (int q = 0; q < 10; ++q) {
        A.gnode[q] = G(pin[q]);
        A.gsum[q] = G(pout[q]);
    }
    const NodesAdjV X{gptr(c, EREB_G_WTXSUM),
                      {G(EREB_G_RHOX), G(EREB_G_RHOFX), G(EREB_G_KX), G(EREB_G_PHIX), G(EREB_G_RX)},
                      {G(EREB_G_RHOXSUM), G(EREB_G_RHOFXSUM), G(EREB_G_KXSUM), G(EREB_G_PHIXSUM),
                       G(EREB_G_RXSUM), G(EREB_G_WTXSUM)}};
    const NodesAdjV Y{gptr(c, EREB_G_WTYSUM),
                      {G(EREB_G_RHOY), G(EREB_G_RHOFY), G(EREB_G_KY), G(EREB_G_PHIY), G(EREB_G_RY)},
                      {G(EREB_G_RHOYSUM), G(EREB_G_RHOFYSUM), G(EREB_G_KYSUM), G(EREB_G_PHIYSUM),
                       G(EREB_G_RYSUM), G(EREB_G_WTYSUM)}};
    // the P, Vx and Vy grids are all Ny1 x Nx1
    hipLaunchKernelGGL(k_nodes_adj, dim3(nb, 3), dim3(256), 0, c->stream, A, X, Y);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

}  // namespace ereb

// ---- adjoint driver: gradient of sum(tk1^2) after n steps ------------------

namespace ereb {

__global__ void k_twice(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = 2.0 * a[i];
}

static const int kStateFields[5] = {EREB_M_XM, EREB_M_YM, EREB_M_TKM, EREB_M_SXXM,
                                    EREB_M_SXYM};

// up to five vectors of n doubles copied in one launch (grid row y = vector y)
struct CopyList {
    const double* src[5];
    double* dst[5];
};

__global__ __launch_bounds__(256) void k_copy_list(CopyList L, int64_t n) {
    // the row's pair by selects (no dynamic indexing of the argument)
    const int y = blockIdx.y;
    const double* __restrict__ s = y == 0 ? L.src[0] : y == 1 ? L.src[1] : y == 2 ? L.src[2]
                                 : y == 3 ? L.src[3] : L.src[4];
    double* __restrict__ d = y == 0 ? L.dst[0] : y == 1 ? L.dst[1] : y == 2 ? L.dst[2]
                           : y == 3 ? L.dst[3] : L.dst[4];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// copy the first nf of (xm, ym, tkm, sxxm, sxym) of the resident markers to
// (save) or from (restore) buf, M doubles each
static int state_copy(ereb_ctx* c, double* buf, int nf, bool save) {
    const int64_t M = c->marknum;
    if (M <= 0 || nf <= 0) return EREB_OK;
    CopyList L{};
    for (int q = 0; q < nf; ++q) {
        double* dev = mptr(c, kStateFields[q]);
        double* b = buf + (size_t)q * M;
        L.src[q] = save ? dev : b;
        L.dst[q] = save ? b : dev;
    }
    const unsigned nb = (unsigned)std::min<int64_t>((M + 255) / 256, 2048);
    hipLaunchKernelGGL(k_copy_list, dim3(nb, (unsigned)nf), dim3(256), 0, c->stream, L, M);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

// up to three ranges set to zero in one launch (grid row y = range y)
struct ZeroList {
    double* p[3];
    int64_t n[3];
};

__global__ __launch_bounds__(256) void k_zero_list(ZeroList Z) {
    const int y = blockIdx.y;
    double* __restrict__ p = y == 0 ? Z.p[0] : y == 1 ? Z.p[1] : Z.p[2];
    const int64_t n = y == 0 ? Z.n[0] : y == 1 ? Z.n[1] : Z.n[2];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

static int zero_list(ereb_ctx* c, const ZeroList& Z, int nr) {
    int64_t mx = 0;
    for (int r = 0; r < nr; ++r) mx = std::max(mx, Z.n[r]);
    if (mx <= 0) return EREB_OK;
    const unsigned nb = (unsigned)std::min<int64_t>((mx + 255) / 256, 2048);
    hipLaunchKernelGGL(k_zero_list, dim3(nb, (unsigned)nr), dim3(256), 0, c->stream, Z);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

static size_t grid_doubles(const ereb_ctx* c) {
    size_t n = 0;
    for (int f = 0; f < EREB_G_NFIELDS; ++f) n = std::max(n, c->goff[f] + (size_t)c->gnumel[f]);
    return n;
}

// one timestep of the differentiated map: scatter, node properties with the
// insulating tk1 boundary, tk2 := tk1 (no heat solve), the temperature
// gather (only at timestep 1: later it adds DT = tk2 - tk1 = 0), RK4;
// xb/yb record the fluid back-trace end points when given
static int adj_forward_step(ereb_ctx* c, int64_t k, double dt, int mode, const double* hrs,
                            const double* hrf, double* xb, double* yb) {
    int rc;
    if ((rc = launch_scatter(c, hrs, hrf, mode))) return rc;
    if ((rc = launch_combine_nodes(c))) return rc;
    EREB_HIP(c, hipMemcpyAsync(gptr(c, EREB_G_TK2), gptr(c, EREB_G_TK1),
                               sizeof(double) * c->gnumel[EREB_G_TK1], hipMemcpyDeviceToDevice,
                               c->stream));
    if (k == 0 && (rc = launch_gather_temperature(c, 1))) return rc;
    return launch_rk4(c, dt, mode, xb, yb);
}

struct AdjBuffers {
    double *save = nullptr, *cps = nullptr, *st = nullptr, *snap = nullptr, *xb = nullptr;
    double *gT = nullptr, *gT2 = nullptr, *gP = nullptr, *gg = nullptr;
    // per recorded step: the scatter's sort order of the step's start
    // positions, the order the transposed kernels walk the markers in
    uint32_t* perm = nullptr;
    // fixed-point sums of the tk2 cotangent of RK4^T (A) and of the gather^T
    // (B): bounds [0..1] (bits of the largest marker cotangent), flag [2],
    // then 2 words per P node for A and for B
    unsigned long long* fix = nullptr;
};

// fixed-point grids of the tk2 cotangent (ereb_fixsum.h) over the pool piece
static FixGrid fix_grid(const AdjBuffers& B, int which, int64_t n) {
    return FixGrid{B.fix + 8 + (size_t)which * 2 * n, B.fix + which, 1.0};
}

// out = value(A) + value(B) per node (B optional), the scales from the bounds
__global__ void k_fix_to_grid(const unsigned long long* __restrict__ wa,
                              const unsigned long long* __restrict__ ba,
                              const unsigned long long* __restrict__ wb,
                              const unsigned long long* __restrict__ bb, int64_t n,
                              double* __restrict__ out, double* __restrict__ out2) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double v = fix_value(wa, k, fix_up_of_bound(ba));
    if (wb) v += fix_value(wb, k, fix_up_of_bound(bb));
    out[k] = v;
    out2[k] = v;
}

// bits[0] = max(bits[0], max |a[i]|) over i < n, one atomic per wavefront:
// the bound of a fixed-point pass (non-negative doubles order as their bits)
// (one atomic per workgroup: many atomics on one address serialise)
__global__ __launch_bounds__(256) void k_absmax_bits(const double* __restrict__ a, int64_t n,
                                                     unsigned long long* __restrict__ bits) {
    __shared__ double wmax[4];
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = fabs(a[i]);
        mx = v > mx || v != v ? v : mx;  // a NaN wins, and is flagged by the pass
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(mx, o);
        mx = y > mx || y != y ? y : mx;
    }
    if (__lane_id() == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            mx = wmax[w] > mx || wmax[w] != wmax[w] ? wmax[w] : mx;
        if (mx != 0.0) atomicMax(bits, (unsigned long long)__double_as_longlong(mx));
    }
}

static int absmax_bits(ereb_ctx* c, const double* a, int64_t n, unsigned long long* bits) {
    const unsigned nb = (unsigned)std::min<int64_t>((n + 255) / 256, 256);
    hipLaunchKernelGGL(k_absmax_bits, dim3(std::max(nb, 1u)), dim3(256), 0, c->stream, a, n,
                       bits);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

// the buffers of one call, carved from the context's grow-only pool; nsnap:
// grid snapshots kept (every step's, or one segment's)
static int adj_buffers(ereb_ctx* c, AdjBuffers& B, bool grad, int64_t nseg, int64_t every,
                       int64_t nsnap, size_t gsz) {
    const size_t M = (size_t)c->marknum;
    const size_t np = (size_t)c->Nx1 * c->Ny1;
    const size_t sizes[11] = {5 * M,
                              grad ? 3 * M * (size_t)nseg : 0,
                              grad ? 3 * M * (size_t)every : 0,
                              grad ? gsz * (size_t)nsnap : 0,
                              grad ? 2 * M * (size_t)every : 0,
                              grad ? M : 0, grad ? M : 0, grad ? M : 0, grad ? gsz : 0,
                              grad ? 8 + 4 * np : 0,
                              grad ? (M * (size_t)every + 1) / 2 : 0};
    size_t tot = 0;
    for (size_t z : sizes) tot += (z + 31) / 32 * 32;  // 256-byte aligned pieces
    if (tot > c->adj_pool_n) {
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        hipFree(c->adj_pool);
        c->adj_pool = nullptr;
        c->adj_pool_n = 0;
        EREB_HIP(c, hipMalloc(&c->adj_pool, sizeof(double) * tot));
        c->adj_pool_n = tot;
    }
    double* p = c->adj_pool;
    double *fix = nullptr, *perm = nullptr;
    double** dst[11] = {&B.save, &B.cps, &B.st, &B.snap, &B.xb, &B.gT, &B.gT2, &B.gP, &B.gg,
                        &fix, &perm};
    for (int k = 0; k < 11; ++k) {
        *dst[k] = sizes[k] ? p : nullptr;
        p += (sizes[k] + 31) / 32 * 32;
    }
    B.fix = (unsigned long long*)fix;
    B.perm = (uint32_t*)perm;
    return EREB_OK;
}

static int adjoint_body(ereb_ctx* c, AdjBuffers& B, int64_t nsteps, int64_t every, int64_t nseg,
                        bool keep_grids, double dt, int md, const double* hrsolidm,
                        const double* hrfluidm, double* loss, double* tk1_out, double* grad_tkm,
                        double* grad_phim);

}  // namespace ereb

using namespace ereb;

// Forward with checkpoints of (xm, ym, tkm) every `every` steps; the reverse
// sweep recomputes each segment from its checkpoint, recording per step the
// start state, the sort order, the grid fields and the fluid back-trace end
// points (the last segment is recorded during the forward pass itself), then
// runs the adjoint kernels backwards through the segment.  The grid fields
// are grid-sized: while every step's fit the context's budget, the forward
// pass keeps them all, and the recomputation runs only the marker side of
// each step (sort of the start positions, temperature gather at timestep 1,
// RK4) instead of the scatter and the node pass again.
extern "C" int ereb_adjoint_set_grid_budget(ereb_ctx* c, int64_t bytes) {
    if (!c || bytes < 0) return EREB_ERR_ARG;
    c->adj_grid_budget = bytes;
    return EREB_OK;
}

extern "C" int ereb_adjoint_tk1_loss(ereb_ctx* c, int64_t nsteps, int64_t every, double dt,
                                     int64_t mode, const double hrsolidm[3],
                                     const double hrfluidm[3], double* loss, double* tk1_out,
                                     double* grad_tkm, double* grad_phim) {
    if (!c || !hrsolidm || !hrfluidm || !loss || nsteps < 0 || every < 1) return EREB_ERR_ARG;
    if (mode != 1 && mode != 9) return fail(c, EREB_ERR_MODE, "unknown mode");
    const bool grad = grad_tkm || grad_phim;
    if (grad && c->precision != EREB_FP64)
        return fail(c, EREB_ERR_ARG, "the adjoint kernels run in the Float64 form");
    const int64_t M = c->marknum;
    if (M <= 0) return fail(c, EREB_ERR_STATE, "no markers");
    const int md = (int)mode;
    const int64_t nseg = (nsteps + every - 1) / every;
    const int64_t seg = std::min(every, std::max<int64_t>(nsteps, 1));
    const size_t gsz = grid_doubles(c);
    const bool keep_grids =
        grad && nsteps > seg &&
        (double)gsz * sizeof(double) * (double)nsteps <= (double)c->adj_grid_budget;
    AdjBuffers B;
    int rc;
    if ((rc = adj_buffers(c, B, grad, nseg, seg, keep_grids ? nsteps : seg, gsz))) return rc;
    // the markers are restored at the end, on every exit path once saved
    if ((rc = state_copy(c, B.save, 5, true))) return rc;
    rc = adjoint_body(c, B, nsteps, every, nseg, keep_grids, dt, md, hrsolidm, hrfluidm, loss,
                      tk1_out, grad_tkm, grad_phim);
    const std::string err = c->err;
    const int rc_restore = state_copy(c, B.save, 5, false);
    ++c->pos_gen;  // the last sort was of moved positions
    if (rc) {
        c->err = err;  // the first failure is the one reported
        hipStreamSynchronize(c->stream);
        return rc;
    }
    if (rc_restore) return rc_restore;
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return check_flags(c);
}

namespace ereb {

static int adjoint_body(ereb_ctx* c, AdjBuffers& B, int64_t nsteps, int64_t every, int64_t nseg,
                        bool keep_grids, double dt, int md, const double* hrsolidm,
                        const double* hrfluidm, double* loss, double* tk1_out, double* grad_tkm,
                        double* grad_phim) {
    const bool grad = grad_tkm || grad_phim;
    const int64_t M = c->marknum;
    const size_t MB = sizeof(double) * (size_t)M;
    const size_t gsz = grid_doubles(c);
    const size_t GB = sizeof(double) * gsz;
    int rc;
    // the records of step k in the segment starting at k0: start state, fluid
    // back-trace end points, sort order (one segment's), grid fields (every
    // step's with keep_grids, else one segment's)
    auto st_of = [&](int64_t k, int64_t k0) { return B.st + 3 * (size_t)M * (k - k0); };
    auto xb_of = [&](int64_t k, int64_t k0) { return B.xb + 2 * (size_t)M * (k - k0); };
    auto perm_of = [&](int64_t k, int64_t k0) { return B.perm + (size_t)M * (k - k0); };
    auto snap_of = [&](int64_t k, int64_t k0) {
        return B.snap + gsz * (size_t)(keep_grids ? k : k - k0);
    };
    // forward step k of the segment starting at k0; record: the step's records
    // (keep_grids: its grid fields in any case)
    auto run_step = [&](int64_t k, int64_t k0, bool record) -> int {
        int r;
        double* xb = nullptr;
        double* yb = nullptr;
        if (record) {
            if ((r = state_copy(c, st_of(k, k0), 3, true))) return r;
            xb = xb_of(k, k0);
            yb = xb + M;
        }
        if ((r = adj_forward_step(c, k, dt, md, hrsolidm, hrfluidm, xb, yb))) return r;
        if (record || keep_grids)
            EREB_HIP(c, hipMemcpyAsync(snap_of(k, k0), c->d_grids, GB, hipMemcpyDeviceToDevice,
                                       c->stream));
        if (record)
            // the step's scatter sorted its start positions (RK4 moved them
            // afterwards without a new sort)
            EREB_HIP(c, hipMemcpyAsync(perm_of(k, k0), c->d_perm, sizeof(uint32_t) * (size_t)M,
                                       hipMemcpyDeviceToDevice, c->stream));
        return EREB_OK;
    };
    // step k recomputed from its start state with its grid fields kept: only
    // the marker side runs -- the sort of the start positions (the order the
    // transposed kernels walk), the temperature gather at timestep 1 and RK4
    // on the step's grids, recording the back-trace end points -- with the
    // results of run_step bit for bit (the scatter and the node pass only
    // produce the grid fields, and the sort depends on the positions alone)
    auto replay_step = [&](int64_t k, int64_t k0) -> int {
        int r;
        if ((r = state_copy(c, st_of(k, k0), 3, true))) return r;
        EREB_HIP(c, hipMemcpyAsync(c->d_grids, snap_of(k, k0), GB, hipMemcpyDeviceToDevice,
                                   c->stream));
        if ((r = launch_fine_sort(c, nullptr))) return r;
        EREB_HIP(c, hipMemcpyAsync(perm_of(k, k0), c->d_perm, sizeof(uint32_t) * (size_t)M,
                                   hipMemcpyDeviceToDevice, c->stream));
        if (k == 0 && (r = launch_gather_temperature(c, 1))) return r;
        double* xb = xb_of(k, k0);
        return launch_rk4(c, dt, md, xb, xb + M);
    };
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t sg = k / every;
        if (grad && k % every == 0 &&
            (rc = state_copy(c, B.cps + 3 * (size_t)M * sg, 3, true)))
            return rc;
        if ((rc = run_step(k, sg * every, grad && sg == nseg - 1))) return rc;
    }
    // loss on the final state
    if ((rc = launch_scatter(c, hrsolidm, hrfluidm, md)) || (rc = launch_combine_nodes(c)))
        return rc;
    const int64_t n = c->gnumel[EREB_G_TK1];
    std::vector<double> tk1((size_t)n);
    EREB_HIP(c, hipMemcpyAsync(tk1.data(), gptr(c, EREB_G_TK1), sizeof(double) * n,
                               hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    if ((rc = check_flags(c))) return rc;
    double L = 0.0;
    for (int64_t i = 0; i < n; ++i) L += tk1[i] * tk1[i];
    *loss = L;
    if (tk1_out) std::copy(tk1.begin(), tk1.end(), tk1_out);
    if (grad) {
        auto G = [&](int f) { return B.gg + c->goff[f]; };
        int64_t goff[EREB_G_NFIELDS];
        for (int f = 0; f < EREB_G_NFIELDS; ++f) goff[f] = (int64_t)c->goff[f];
        int64_t soff[NSUM];
        for (int s = 0; s < NSUM; ++s) soff[s] = goff[sum_grid_field(s)];
        const unsigned nbg = (unsigned)((n + 255) / 256);
        if ((rc = zero_list(c, ZeroList{{B.gT, B.gP, B.gg}, {M, M, (int64_t)gsz}}, 3)))
            return rc;
        // d L / d tk1 = 2 tk1 on the final state, then nodes^T and scatter^T there
        hipLaunchKernelGGL(k_twice, dim3(nbg), dim3(256), 0, c->stream,
                           (const double*)gptr(c, EREB_G_TK1), G(EREB_G_TK1), n);
        if ((rc = launch_nodes_adj(c, B.gg, goff))) return rc;
        if ((rc = launch_scatter_adj(c, hrsolidm, hrfluidm, md, B.gg, soff, B.gT, B.gP, true)))
            return rc;
        // set when a tk2 cotangent contribution falls outside the range of
        // its fixed-point scale (or is not finite); checked after the sweep
        int* flag = (int*)(B.fix + 2);
        if ((rc = fill_i32(c, flag, 0, 2))) return rc;
        for (int64_t sg = nseg - 1; sg >= 0; --sg) {
            const int64_t k0 = sg * every, k1 = std::min(nsteps, k0 + every);
            if (sg != nseg - 1) {  // the last segment was recorded by the forward pass
                if ((rc = state_copy(c, B.cps + 3 * (size_t)M * sg, 3, false))) return rc;
                for (int64_t k = k0; k < k1; ++k)
                    if ((rc = keep_grids ? replay_step(k, k0) : run_step(k, k0, true))) return rc;
            }
            for (int64_t k = k1 - 1; k >= k0; --k) {
                double* stk = st_of(k, k0);
                const double* xb = xb_of(k, k0);
                // the grids and the start state of step k
                EREB_HIP(c, hipMemcpyAsync(c->d_grids, snap_of(k, k0), GB,
                                           hipMemcpyDeviceToDevice, c->stream));
                if ((rc = state_copy(c, stk, 3, false))) return rc;
                if (k == 0 && (rc = launch_gather_temperature(c, 1))) return rc;
                // RK4^T (tkm as RK4 saw it), then gather^T at timestep 1; their
                // tk2 cotangents are summed in fixed point, each with a scale
                // from a bound on its contributions (the largest marker
                // cotangent, ereb_fixsum.h), so the gradient is the same on
                // every run and the sweep needs no host round trip.  Cleared
                // first: the grid cotangents, the fixed-point sums, the bounds
                if ((rc = zero_list(c, ZeroList{{B.gg, (double*)B.fix + 8, (double*)B.fix},
                                                {(int64_t)gsz, 4 * n, 2}},
                                    3)))
                    return rc;
                FixGrid FA = fix_grid(B, 0, n), FB = fix_grid(B, 1, n);
                FA.bound = FA.maxbits;
                FB.bound = FB.maxbits;
                // the transposed kernels walk the markers in the step's sort
                // order: a workgroup's stencils then cover a compact patch of
                // nodes, and its window needs few device-wide atomics
                const uint32_t* order = perm_of(k, k0);
                if ((rc = absmax_bits(c, B.gT, M, FA.maxbits)) ||
                    (rc = launch_rk4_adj(c, 1, dt, md, xb, xb + M, B.gT, B.gT2, B.gP, FA, flag,
                                         order)))
                    return rc;
                if (k == 0 && ((rc = absmax_bits(c, B.gT2, M, FB.maxbits)) ||
                               (rc = launch_gather_T_adj(c, 1, 1, B.gT2, FB, flag, order))))
                    return rc;
                // the tk2 cotangent, also the tk1 one (tk2 := tk1), then node
                // properties^T
                hipLaunchKernelGGL(k_fix_to_grid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                                   c->stream, (const unsigned long long*)FA.w, FA.bound,
                                   k == 0 ? (const unsigned long long*)FB.w : nullptr, FB.bound,
                                   n, G(EREB_G_TK2), G(EREB_G_TK1));
                EREB_HIP(c, hipGetLastError());
                if ((rc = launch_nodes_adj(c, B.gg, goff))) return rc;
                // scatter^T with the temperature the scatter saw
                if (k == 0)
                    EREB_HIP(c, hipMemcpyAsync(mptr(c, EREB_M_TKM), stk + 2 * (size_t)M, MB,
                                               hipMemcpyDeviceToDevice, c->stream));
                if ((rc = launch_scatter_adj(c, hrsolidm, hrfluidm, md, B.gg, soff, B.gT2, B.gP,
                                             true)))
                    return rc;
                std::swap(B.gT, B.gT2);
            }
        }
        int fl = 0;
        EREB_HIP(c, hipMemcpyAsync(&fl, flag, sizeof fl, hipMemcpyDeviceToHost, c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        if (fl)
            return fail(c, EREB_ERR_DOMAIN,
                        "adjoint: a tk2 cotangent contribution is not finite or outside the "
                        "fixed-point range of its scale");
        if (grad_tkm)
            EREB_HIP(c, hipMemcpyAsync(grad_tkm, B.gT, MB, hipMemcpyDeviceToHost, c->stream));
        if (grad_phim)
            EREB_HIP(c, hipMemcpyAsync(grad_phim, B.gP, MB, hipMemcpyDeviceToHost, c->stream));
    }
    return EREB_OK;
}

}  // namespace ereb
