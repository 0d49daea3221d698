// ereb_nodes.hip -- nodal normalisation of the interpolated sums:
// compute_{basic,vx,vy,p}_node_properties! (E:1970-2189) and
// apply_insulating_boundary_conditions! (E:2432-2449).  Float64 in both
// precision forms.  Nodes with zero weight keep their previous values, as
// in the reference (the writes sit under `if WT > 0.0`).
#include "ereb_internal.h"

namespace ereb {

struct NodeArgs {
    int nb, ns;  // basic and staggered node counts
    int do_basic, do_vx, do_vy, do_p;
    int bc, R, C;  // insulating conditions on tk1 (P grid of R rows, C columns) fused
    // sums (SUM order of reset_interpolated_properties!) and outputs:
    // basic  ETA0 ETA GGG SXY COH TEN FRI | WT   -> ETA0 ETA GGG SXY0 COH TEN FRI
    // Vx, Vy RHO RHOF K PHI R | WT               -> RHO RHOF K PHI R
    // P      RHO RHOCP ALPHA ALPHAF HR GGGP SXX TK PHI | WT
    //                                           -> RHO RHOCP ALPHA ALPHAF HR GGGP SXX0 tk1 PHI
    //                                              BETAPHI
    double *bsum[8], *vxs[6], *vys[6], *psum[10];
    double *bout[7], *vxo[5], *vyo[5], *pout[10];
    uint8_t* YNY;
};

// compute_basic_node_properties! (E:1970-2007) at node k; sum(q) gives the
// node's sum q
template <class S>
__device__ __forceinline__ void basic_props(const NodeArgs& A, int k, const S& sum) {
    const double wt = sum(7);
    if (wt > 0.0) {
        const double iw = 1.0 / wt;
        const double eta0 = sum(0) * iw, eta = sum(1) * iw;
        A.bout[0][k] = eta0;
        A.bout[1][k] = eta;
        if (eta < eta0) A.YNY[k] = 1;
        A.bout[2][k] = (1.0 / sum(2)) * wt;
        A.bout[3][k] = sum(3) * iw;
        A.bout[4][k] = sum(4) * iw;
        A.bout[5][k] = sum(5) * iw;
        A.bout[6][k] = sum(6) * iw;
    }
}

// compute_vx_node_properties! / compute_vy_node_properties! (E:2033-2060,
// E:2086-2113)
template <class S>
__device__ __forceinline__ void v_props(double* const* out, int k, const S& sum) {
    const double wt = sum(5);
    if (wt > 0.0) {
        const double iw = 1.0 / wt;
#pragma unroll
        for (int q = 0; q < 5; ++q) out[q][k] = sum(q) * iw;
    }
}

// the P-grid temperature of node k when its weight is positive
template <class S>
__device__ __forceinline__ double p_tk1(const S& sum) {
    return sum(7) * (1.0 / sum(1));
}

// compute_p_node_properties! (E:2148-2189) at node k, then, with A.bc, the
// insulating conditions of tk1 (E:2432-2449) fused: a boundary node takes the
// value of its nearest inner node (rows first, then columns, so a corner
// takes the inner diagonal): that node's temperature from its own sums
// (innersum) when its weight is positive, else its unchanged tk1
template <class S, class T>
__device__ __forceinline__ void p_props(const NodeArgs& A, int k, const S& sum,
                                        const T& innersum) {
    const double wt = sum(9);
    if (wt > 0.0) {
        const double iw = 1.0 / wt;
        A.pout[0][k] = sum(0) * iw;
        A.pout[1][k] = sum(1) * iw;
        A.pout[2][k] = sum(2) * iw;
        A.pout[3][k] = sum(3) * iw;
        A.pout[4][k] = sum(4) * iw;
        const double gggp = (1.0 / sum(5)) * wt;
        A.pout[5][k] = gggp;
        A.pout[6][k] = sum(6) * iw;
        A.pout[7][k] = p_tk1(sum);
        const double phi = sum(8) * iw;
        A.pout[8][k] = phi;
        A.pout[9][k] = (1.0 / gggp) * phi;
    }
    if (A.bc) {
        const int i = k % A.R, j = k / A.R;
        if (i == 0 || i == A.R - 1 || j == 0 || j == A.C - 1) {
            const int si = min(max(i, 1), A.R - 2), sj = min(max(j, 1), A.C - 2);
            const int q = sj * A.R + si;
            A.pout[7][k] = innersum(q, si, sj, 9) > 0.0
                               ? innersum(q, si, sj, 7) * (1.0 / innersum(q, si, sj, 1))
                               : A.pout[7][q];
        }
    }
}

// blockIdx.y picks the grid (0 basic, 1 Vx, 2 Vy, 3 P): four times the
// wavefronts of one thread per node doing all four, for a pass that is all
// loads and stores
__global__ __launch_bounds__(256) void k_node_props(NodeArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int grid = blockIdx.y;
    if (grid == 0) {
        if (A.do_basic && k < A.nb) basic_props(A, k, [&](int q) { return A.bsum[q][k]; });
        return;
    }
    if (k >= A.ns) return;
    if (grid == 1 && A.do_vx) v_props(A.vxo, k, [&](int q) { return A.vxs[q][k]; });
    if (grid == 2 && A.do_vy) v_props(A.vyo, k, [&](int q) { return A.vys[q][k]; });
    if (grid == 3 && A.do_p)
        p_props(A, k, [&](int q) { return A.psum[q][k]; },
                [&](int node, int, int, int q) { return A.psum[q][node]; });
}

// The scatter's window combine and the node properties in one pass (the
// transfer step, where they run back to back): one thread per node and grid
// adds the window partials of each of the grid's sums (the same additions
// as k_combine), writes the SUM grids and evaluates the node properties
// from the sums in registers.  The fused boundary conditions recompute the
// inner node's three sums from the partials.
template <int G>
__device__ __forceinline__ void combine_node(const NodeArgs& A, const double* __restrict__ part,
                                             int nty, int ntx, int k) {
    constexpr int S0 = G == 0 ? 0 : G == 1 ? 8 : G == 2 ? 14 : 20;
    constexpr int NS = G == 0 ? 8 : G == 3 ? 10 : 6;
    const int rows = G == 0 ? A.R - 1 : A.R;
    const int i0 = k % rows, j0 = k / rows;
    const WinTiles w = win_tiles(i0, j0, nty, ntx);
    double sm[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) sm[q] = win_sum(part, nty, w, S0 + q, i0, j0);
    double* const* sums = G == 0 ? A.bsum : G == 1 ? A.vxs : G == 2 ? A.vys : A.psum;
#pragma unroll
    for (int q = 0; q < NS; ++q) sums[q][k] = sm[q];
    const auto sum = [&](int q) { return sm[q]; };
    if constexpr (G == 0) basic_props(A, k, sum);
    if constexpr (G == 1) v_props(A.vxo, k, sum);
    if constexpr (G == 2) v_props(A.vyo, k, sum);
    if constexpr (G == 3)
        p_props(A, k, sum, [&](int, int si, int sj, int q) {
            return win_sum(part, nty, win_tiles(si, sj, nty, ntx), S0 + q, si, sj);
        });
}

__global__ __launch_bounds__(256) void k_combine_nodes(NodeArgs A, const double* __restrict__ part,
                                                       int nty, int ntx) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    switch (blockIdx.y) {
        case 0: if (k < A.nb) combine_node<0>(A, part, nty, ntx, k); break;
        case 1: if (k < A.ns) combine_node<1>(A, part, nty, ntx, k); break;
        case 2: if (k < A.ns) combine_node<2>(A, part, nty, ntx, k); break;
        default: if (k < A.ns) combine_node<3>(A, part, nty, ntx, k); break;
    }
}

// apply_insulating_boundary_conditions! (E:2432-2449): rows 1 and Nyy copy
// their inner neighbours for columns 2..Nxx-1, then columns 1 and Nxx copy
// columns 2 and Nxx-1 over all rows (so corners take the updated rows).
// One workgroup: the rows, a barrier (the workgroup's global stores are
// visible to it after the barrier), then the columns.
__global__ __launch_bounds__(1024) void k_bc(double* t, int Nyy, int Nxx) {
    for (int j = 1 + threadIdx.x; j <= Nxx - 2; j += blockDim.x) {  // columns 2..Nxx-1
        double* col = t + (size_t)j * Nyy;
        col[0] = col[1];
        col[Nyy - 1] = col[Nyy - 2];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < Nyy; i += blockDim.x) {
        t[i] = t[(size_t)Nyy + i];
        t[(size_t)(Nxx - 1) * Nyy + i] = t[(size_t)(Nxx - 2) * Nyy + i];
    }
}

int launch_insulating_bc(ereb_ctx* c, double* t, int64_t Nyy, int64_t Nxx) {
    if (!(Nyy > 2 && Nxx > 2)) return EREB_OK;
    hipLaunchKernelGGL(k_bc, dim3(1), dim3(1024), 0, c->stream, t, (int)Nyy, (int)Nxx);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

static NodeArgs node_args(ereb_ctx* c, bool basic, bool vx, bool vy, bool p, bool bc) {
    NodeArgs A;
    A.nb = c->Nx * c->Ny;
    A.ns = c->Nx1 * c->Ny1;
    A.do_basic = basic;
    A.do_vx = vx;
    A.do_vy = vy;
    A.do_p = p;
    // the boundary conditions ride on the P-grid pass when it runs
    A.R = (int)c->Ny1;
    A.C = (int)c->Nx1;
    A.bc = bc && p && A.R > 2 && A.C > 2;
    static const int bs[8] = {EREB_G_ETA0SUM, EREB_G_ETASUM, EREB_G_GGGSUM, EREB_G_SXYSUM,
                              EREB_G_COHSUM, EREB_G_TENSUM, EREB_G_FRISUM, EREB_G_WTSUM};
    static const int bo[7] = {EREB_G_ETA0, EREB_G_ETA, EREB_G_GGG, EREB_G_SXY0,
                              EREB_G_COH, EREB_G_TEN, EREB_G_FRI};
    static const int vxs[6] = {EREB_G_RHOXSUM, EREB_G_RHOFXSUM, EREB_G_KXSUM, EREB_G_PHIXSUM,
                               EREB_G_RXSUM, EREB_G_WTXSUM};
    static const int vys[6] = {EREB_G_RHOYSUM, EREB_G_RHOFYSUM, EREB_G_KYSUM, EREB_G_PHIYSUM,
                               EREB_G_RYSUM, EREB_G_WTYSUM};
    static const int vxo[5] = {EREB_G_RHOX, EREB_G_RHOFX, EREB_G_KX, EREB_G_PHIX, EREB_G_RX};
    static const int vyo[5] = {EREB_G_RHOY, EREB_G_RHOFY, EREB_G_KY, EREB_G_PHIY, EREB_G_RY};
    static const int ps[10] = {EREB_G_RHOSUM, EREB_G_RHOCPSUM, EREB_G_ALPHASUM,
                               EREB_G_ALPHAFSUM, EREB_G_HRSUM, EREB_G_GGGPSUM, EREB_G_SXXSUM,
                               EREB_G_TKSUM, EREB_G_PHISUM, EREB_G_WTPSUM};
    static const int po[10] = {EREB_G_RHO, EREB_G_RHOCP, EREB_G_ALPHA, EREB_G_ALPHAF,
                               EREB_G_HR, EREB_G_GGGP, EREB_G_SXX0, EREB_G_TK1,
                               EREB_G_PHI, EREB_G_BETAPHI};
    for (int q = 0; q < 8; ++q) A.bsum[q] = gptr(c, bs[q]);
    for (int q = 0; q < 7; ++q) A.bout[q] = gptr(c, bo[q]);
    for (int q = 0; q < 6; ++q) {
        A.vxs[q] = gptr(c, vxs[q]);
        A.vys[q] = gptr(c, vys[q]);
    }
    for (int q = 0; q < 5; ++q) {
        A.vxo[q] = gptr(c, vxo[q]);
        A.vyo[q] = gptr(c, vyo[q]);
    }
    for (int q = 0; q < 10; ++q) {
        A.psum[q] = gptr(c, ps[q]);
        A.pout[q] = gptr(c, po[q]);
    }
    A.YNY = c->d_yny;
    return A;
}

int launch_node_properties(ereb_ctx* c, bool basic, bool vx, bool vy, bool p, bool bc) {
    const NodeArgs A = node_args(c, basic, vx, vy, p, bc);
    const int n = A.ns > A.nb ? A.ns : A.nb;
    hipEvent_t ev;
    prof_begin(c, EREB_K_NODES, &ev);
    hipLaunchKernelGGL(k_node_props, dim3((unsigned)((n + 255) / 256), 4), dim3(256), 0,
                       c->stream, A);
    EREB_HIP(c, hipGetLastError());
    if (bc && !p) {
        int rc = launch_insulating_bc(c, gptr(c, EREB_G_TK1), c->Ny1, c->Nx1);
        if (rc) return rc;
    }
    prof_end(c, EREB_K_NODES, ev);
    return EREB_OK;
}

// the scatter's SUM grids from its window partials, the node properties of
// all four grids and the insulating tk1 boundary, in one launch
int launch_combine_nodes(ereb_ctx* c) {
    const NodeArgs A = node_args(c, true, true, true, true, true);
    const int n = A.ns > A.nb ? A.ns : A.nb;
    hipEvent_t ev;
    prof_begin(c, EREB_K_NODES, &ev);
    hipLaunchKernelGGL(k_combine_nodes, dim3((unsigned)((n + 255) / 256), 4), dim3(256), 0,
                       c->stream, A, (const double*)c->d_part, c->nty, c->ntx);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_NODES, ev);
    return EREB_OK;
}

}  // namespace ereb
