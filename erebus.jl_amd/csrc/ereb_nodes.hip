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
    // basic
    const double *ETA0SUM, *ETASUM, *GGGSUM, *SXYSUM, *COHSUM, *TENSUM, *FRISUM, *WTSUM;
    double *ETA0, *ETA, *GGG, *SXY0, *COH, *TEN, *FRI;
    uint8_t* YNY;
    // Vx / Vy (sums and outputs: RHO RHOF K PHI R)
    const double *vxs[6], *vys[6];
    double *vxo[5], *vyo[5];
    // P
    const double *RHOSUM, *RHOCPSUM, *ALPHASUM, *ALPHAFSUM, *HRSUM, *GGGPSUM, *SXXSUM, *TKSUM,
        *PHISUM, *WTPSUM;
    double *RHO, *RHOCP, *ALPHA, *ALPHAF, *HR, *GGGP, *SXX0, *tk1, *PHI, *BETAPHI;
};

// blockIdx.y picks the grid (0 basic, 1 Vx, 2 Vy, 3 P): four times the
// wavefronts of one thread per node doing all four, for a pass that is all
// loads and stores
__global__ __launch_bounds__(256) void k_node_props(NodeArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int grid = blockIdx.y;
    if (grid == 0) {
        if (!A.do_basic || k >= A.nb) return;
        const double wt = A.WTSUM[k];
        if (wt > 0.0) {
            const double iw = 1.0 / wt;
            const double eta0 = A.ETA0SUM[k] * iw, eta = A.ETASUM[k] * iw;
            A.ETA0[k] = eta0;
            A.ETA[k] = eta;
            if (eta < eta0) A.YNY[k] = 1;
            A.GGG[k] = (1.0 / A.GGGSUM[k]) * wt;
            A.SXY0[k] = A.SXYSUM[k] * iw;
            A.COH[k] = A.COHSUM[k] * iw;
            A.TEN[k] = A.TENSUM[k] * iw;
            A.FRI[k] = A.FRISUM[k] * iw;
        }
        return;
    }
    if (k >= A.ns) return;
    if (grid == 1 && A.do_vx) {
        const double wt = A.vxs[5][k];
        if (wt > 0.0) {
            const double iw = 1.0 / wt;
#pragma unroll
            for (int q = 0; q < 5; ++q) A.vxo[q][k] = A.vxs[q][k] * iw;
        }
    }
    if (grid == 2 && A.do_vy) {
        const double wt = A.vys[5][k];
        if (wt > 0.0) {
            const double iw = 1.0 / wt;
#pragma unroll
            for (int q = 0; q < 5; ++q) A.vyo[q][k] = A.vys[q][k] * iw;
        }
    }
    if (grid == 3 && A.do_p) {
        const double wt = A.WTPSUM[k];
        if (wt > 0.0) {
            const double iw = 1.0 / wt;
            A.RHO[k] = A.RHOSUM[k] * iw;
            A.RHOCP[k] = A.RHOCPSUM[k] * iw;
            A.ALPHA[k] = A.ALPHASUM[k] * iw;
            A.ALPHAF[k] = A.ALPHAFSUM[k] * iw;
            A.HR[k] = A.HRSUM[k] * iw;
            const double gggp = (1.0 / A.GGGPSUM[k]) * wt;
            A.GGGP[k] = gggp;
            A.SXX0[k] = A.SXXSUM[k] * iw;
            A.tk1[k] = A.TKSUM[k] * (1.0 / A.RHOCPSUM[k]);
            const double phi = A.PHISUM[k] * iw;
            A.PHI[k] = phi;
            A.BETAPHI[k] = (1.0 / gggp) * phi;
        }
        if (A.bc) {
            // apply_insulating_boundary_conditions! (E:2432-2449) fused: a
            // boundary node takes the value of its nearest inner node
            // (rows first, then columns, so a corner takes the inner
            // diagonal), which that node's thread writes from the same sums
            // when its weight is positive and leaves unchanged otherwise
            const int i = k % A.R, j = k / A.R;
            if (i == 0 || i == A.R - 1 || j == 0 || j == A.C - 1) {
                const int si = min(max(i, 1), A.R - 2), sj = min(max(j, 1), A.C - 2);
                const int q = sj * A.R + si;
                A.tk1[k] = A.WTPSUM[q] > 0.0 ? A.TKSUM[q] * (1.0 / A.RHOCPSUM[q]) : A.tk1[q];
            }
        }
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

int launch_node_properties(ereb_ctx* c, bool basic, bool vx, bool vy, bool p, bool bc) {
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
    A.ETA0SUM = gptr(c, EREB_G_ETA0SUM); A.ETASUM = gptr(c, EREB_G_ETASUM);
    A.GGGSUM = gptr(c, EREB_G_GGGSUM); A.SXYSUM = gptr(c, EREB_G_SXYSUM);
    A.COHSUM = gptr(c, EREB_G_COHSUM); A.TENSUM = gptr(c, EREB_G_TENSUM);
    A.FRISUM = gptr(c, EREB_G_FRISUM); A.WTSUM = gptr(c, EREB_G_WTSUM);
    A.ETA0 = gptr(c, EREB_G_ETA0); A.ETA = gptr(c, EREB_G_ETA); A.GGG = gptr(c, EREB_G_GGG);
    A.SXY0 = gptr(c, EREB_G_SXY0); A.COH = gptr(c, EREB_G_COH); A.TEN = gptr(c, EREB_G_TEN);
    A.FRI = gptr(c, EREB_G_FRI); A.YNY = c->d_yny;
    const int vxs[6] = {EREB_G_RHOXSUM, EREB_G_RHOFXSUM, EREB_G_KXSUM, EREB_G_PHIXSUM,
                        EREB_G_RXSUM, EREB_G_WTXSUM};
    const int vys[6] = {EREB_G_RHOYSUM, EREB_G_RHOFYSUM, EREB_G_KYSUM, EREB_G_PHIYSUM,
                        EREB_G_RYSUM, EREB_G_WTYSUM};
    const int vxo[5] = {EREB_G_RHOX, EREB_G_RHOFX, EREB_G_KX, EREB_G_PHIX, EREB_G_RX};
    const int vyo[5] = {EREB_G_RHOY, EREB_G_RHOFY, EREB_G_KY, EREB_G_PHIY, EREB_G_RY};
    for (int q = 0; q < 6; ++q) {
        A.vxs[q] = gptr(c, vxs[q]);
        A.vys[q] = gptr(c, vys[q]);
    }
    for (int q = 0; q < 5; ++q) {
        A.vxo[q] = gptr(c, vxo[q]);
        A.vyo[q] = gptr(c, vyo[q]);
    }
    A.RHOSUM = gptr(c, EREB_G_RHOSUM); A.RHOCPSUM = gptr(c, EREB_G_RHOCPSUM);
    A.ALPHASUM = gptr(c, EREB_G_ALPHASUM); A.ALPHAFSUM = gptr(c, EREB_G_ALPHAFSUM);
    A.HRSUM = gptr(c, EREB_G_HRSUM); A.GGGPSUM = gptr(c, EREB_G_GGGPSUM);
    A.SXXSUM = gptr(c, EREB_G_SXXSUM); A.TKSUM = gptr(c, EREB_G_TKSUM);
    A.PHISUM = gptr(c, EREB_G_PHISUM); A.WTPSUM = gptr(c, EREB_G_WTPSUM);
    A.RHO = gptr(c, EREB_G_RHO); A.RHOCP = gptr(c, EREB_G_RHOCP); A.ALPHA = gptr(c, EREB_G_ALPHA);
    A.ALPHAF = gptr(c, EREB_G_ALPHAF); A.HR = gptr(c, EREB_G_HR); A.GGGP = gptr(c, EREB_G_GGGP);
    A.SXX0 = gptr(c, EREB_G_SXX0); A.tk1 = gptr(c, EREB_G_TK1); A.PHI = gptr(c, EREB_G_PHI);
    A.BETAPHI = gptr(c, EREB_G_BETAPHI);
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

}  // namespace ereb
