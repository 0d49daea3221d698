// ereb_assembly.hip -- step 5: assemble_hydromechanical_lse! (E:3089-3572)
// and assemble_thermal_lse! (E:4526-4611) straight into CSC.
//
// ExtendableSparse's updateindex!(L, +, v, row, col) creates no entry when
// v == 0, so the pattern depends on the values.  Every value is computed by
// the same expression as the reference (no fma contraction), which makes
// the zero tests, the pattern and the stored values bit-exact.
//
// Every row of P cell (i', j') references columns of cells (i'+di, j'+dj)
// with |di|, |dj| <= 1 only.  Seen from a column of cell (i, j), its
// candidate rows are therefore the 9 x NU rows of the neighbouring cells,
// and ordering them by (dj, di, unknown) orders them by row index.  So:
//   1. mask : one thread per P cell evaluates its rows; each nonzero entry
//             sets bit ((dj+1)*3 + di+1)*NU + unknown in its column's 64-bit
//             mask (atomicOr: the result does not depend on order); R is
//             written here
//   2. scan : column counts = popcount(mask), column pointers
//   3. fill : the same evaluation places each entry at colptr + popcount of
//             the lower bits of its column's mask: rows come out ascending
//             in every column, with no sort and independent of thread order.
#include "ereb_internal.h"

namespace ereb {

struct AsmArgs {
    int Nx, Ny, Nx1, Ny1;
    double dx, dy, dt;
    double Kcont, psurface;
    double bcleft, bcright, bctop, bcbottom, bcfleft, bcfright, bcftop, bcfbottom;
    double vxleft, vxright, vytop, vybottom;
    const double *ETA, *ETAP, *GGG, *GGGP, *SXY0, *SXX0, *RHOX, *RHOY, *RHOFX, *RHOFY, *RX, *RY,
        *ETAPHI, *BETAPHI, *PHI, *gx, *gy, *pr0, *pf0, *DMP;
    // thermal
    const double *tk1, *RHOCP, *KX, *KY, *HR, *HA, *HS, *DHP;
    double* R;
    int32_t *colcnt, *colptr, *rowval;
    uint64_t* mask;
    double* nzval;
};

// bit of row `row` among the candidate rows of column `col` (1-based),
// NU unknowns per cell, Ny1 cells per grid column
template <int NU>
__device__ __forceinline__ int cand_bit(int64_t row, int64_t col, int Ny1) {
    const int64_t cr = (row - 1) / NU, cc = (col - 1) / NU;
    const int di = (int)(cr % Ny1) - (int)(cc % Ny1);
    const int dj = (int)(cr / Ny1) - (int)(cc / Ny1);
    return ((dj + 1) * 3 + (di + 1)) * NU + (int)((row - 1) % NU);
}

template <int NU>
struct MaskSink {
    uint64_t* mask;
    int Ny1;
    __device__ void operator()(double v, int64_t row, int64_t col) const {
        if (v != 0.0)
            atomicOr((unsigned long long*)&mask[col - 1],
                     1ull << cand_bit<NU>(row, col, Ny1));
    }
};

template <int NU>
struct FillSink {
    const int32_t* colptr;
    const uint64_t* mask;
    int32_t* rowval;
    double* nzval;
    int Ny1;
    __device__ void operator()(double v, int64_t row, int64_t col) const {
        if (v != 0.0) {
            const int b = cand_bit<NU>(row, col, Ny1);
            const int32_t p = colptr[col - 1] + (int32_t)__popcll(mask[col - 1] & ((1ull << b) - 1));
            rowval[p] = (int32_t)(row - 1);
            nzval[p] = v;
        }
    }
};

// One P cell (i, j) of assemble_hydromechanical_lse!; rows kvx..kpf.
template <class Sink>
__device__ __forceinline__ void stokes_cell(const AsmArgs& A, int i, int j, const Sink& add,
                                            bool writeR) {
    const int Nx = A.Nx, Ny = A.Ny, Nx1 = A.Nx1, Ny1 = A.Ny1;
    const double dx = A.dx, dy = A.dy, dt = A.dt, Kcont = A.Kcont;
    auto B = [&](const double* a, int ii, int jj) { return a[(size_t)(jj - 1) * Ny + (ii - 1)]; };
    auto P = [&](const double* a, int ii, int jj) { return a[(size_t)(jj - 1) * Ny1 + (ii - 1)]; };
    const int64_t kvx = ((int64_t)(j - 1) * Ny1 + i - 1) * 6 + 1;
    const int64_t kvy = kvx + 1, kpm = kvx + 2, kqx = kvx + 3, kqy = kvx + 4, kpf = kvx + 5;
    double rvx = 0.0, rvy = 0.0, rpm = 0.0, rqx = 0.0, rqy = 0.0, rpf = 0.0;
    // Vx equation E:3128-3262
    if (i == 1 || i == Ny1 || j == 1 || j == Nx || j == Nx1) {
        add(1.0, kvx, kvx);
        if (j == 1) rvx = A.vxleft;
        if (j == Nx) rvx = A.vxright;
        if (i == 1 && 1 < j && j < Nx) add(A.bctop, kvx, kvx + 6);
        if (i == Ny1 && 1 < j && j < Nx) add(A.bcbottom, kvx, kvx - 6);
    } else {
        const double ETA1 = B(A.ETA, i - 1, j) * B(A.GGG, i - 1, j) * dt /
                            (B(A.GGG, i - 1, j) * dt + B(A.ETA, i - 1, j));
        const double ETA2 = B(A.ETA, i, j) * B(A.GGG, i, j) * dt /
                            (B(A.GGG, i, j) * dt + B(A.ETA, i, j));
        const double ETAP1 = P(A.ETAP, i, j) * P(A.GGGP, i, j) * dt /
                             (P(A.GGGP, i, j) * dt + P(A.ETAP, i, j));
        const double ETAP2 = P(A.ETAP, i, j + 1) * P(A.GGGP, i, j + 1) * dt /
                             (P(A.GGGP, i, j + 1) * dt + P(A.ETAP, i, j + 1));
        const double SXY1 = B(A.SXY0, i - 1, j) * B(A.ETA, i - 1, j) /
                            (B(A.GGG, i - 1, j) * dt + B(A.ETA, i - 1, j));
        const double SXY2 = B(A.SXY0, i, j) * B(A.ETA, i, j) / (B(A.GGG, i, j) * dt + B(A.ETA, i, j));
        const double SXX1 = P(A.SXX0, i, j) * P(A.ETAP, i, j) /
                            (P(A.GGGP, i, j) * dt + P(A.ETAP, i, j));
        const double SXX2 = P(A.SXX0, i, j + 1) * P(A.ETAP, i, j + 1) /
                            (P(A.GGGP, i, j + 1) * dt + P(A.ETAP, i, j + 1));
        const double dRdx = 0.5 * (P(A.RHOX, i, j + 1) - P(A.RHOX, i, j - 1)) * (1.0 / dx);
        const double dRdy = 0.5 * (P(A.RHOX, i + 1, j) - P(A.RHOX, i - 1, j)) * (1.0 / dy);
        const double gxij = P(A.gx, i, j);
        const double idx = 1.0 / dx, idy = 1.0 / dy;
        add(ETAP1 / (dx * dx), kvx, kvx - 6 * Ny1);
        add(ETA1 / (dy * dy), kvx, kvx - 6);
        add(-(ETAP1 + ETAP2) * (1.0 / (dx * dx)) - (ETA1 + ETA2) * (1.0 / (dy * dy)) -
                dRdx * gxij * dt,
            kvx, kvx);
        add(ETA2 / (dy * dy), kvx, kvx + 6);
        add(ETAP2 / (dx * dx), kvx, kvx + 6 * Ny1);
        add(ETAP1 * idx * idy - ETA2 * idx * idy - dRdy * gxij * dt * 0.25, kvx, kvy);
        add(-ETAP2 * idx * idy + ETA2 * idx * idy - dRdy * gxij * dt * 0.25, kvx, kvy + 6 * Ny1);
        add(-ETAP1 * idx * idy + ETA1 * idx * idy - dRdy * gxij * dt * 0.25, kvx, kvy - 6);
        add(ETAP2 * idx * idy - ETA1 * idx * idy - dRdy * gxij * dt * 0.25, kvx,
            kvy + 6 * Ny1 - 6);
        add(Kcont * idx, kvx, kpm);
        add(-Kcont * idx, kvx, kpm + 6 * Ny1);
        rvx = -P(A.RHOX, i, j) * gxij - (SXY2 - SXY1) * idy - (SXX2 - SXX1) * idx;
    }
    // Vy equation E:3264-3396
    if (i == 1 || i == Ny || i == Ny1 || j == 1 || j == Nx1) {
        add(1.0, kvy, kvy);
        if (i == 1) rvy = A.vytop;
        if (i == Ny) rvy = A.vybottom;
        if (j == 1 && 1 < i && i < Ny) add(A.bcleft, kvy, kvy + 6 * Ny1);
        if (j == Nx1 && 1 < i && i < Ny) add(A.bcright, kvy, kvy - 6 * Ny1);
    } else {
        const double ETA1 = B(A.ETA, i, j - 1) * B(A.GGG, i, j - 1) * dt /
                            (B(A.GGG, i, j - 1) * dt + B(A.ETA, i, j - 1));
        const double ETA2 = B(A.ETA, i, j) * B(A.GGG, i, j) * dt /
                            (B(A.GGG, i, j) * dt + B(A.ETA, i, j));
        const double ETAP1 = P(A.ETAP, i, j) * P(A.GGGP, i, j) * dt /
                             (P(A.GGGP, i, j) * dt + P(A.ETAP, i, j));
        const double ETAP2 = P(A.ETAP, i + 1, j) * P(A.GGGP, i + 1, j) * dt /
                             (P(A.GGGP, i + 1, j) * dt + P(A.ETAP, i + 1, j));
        const double SXY1 = B(A.SXY0, i, j - 1) * B(A.ETA, i, j - 1) /
                            (B(A.GGG, i, j - 1) * dt + B(A.ETA, i, j - 1));
        const double SXY2 = B(A.SXY0, i, j) * B(A.ETA, i, j) / (B(A.GGG, i, j) * dt + B(A.ETA, i, j));
        const double SYY1 = -P(A.SXX0, i, j) * P(A.ETAP, i, j) /
                            (P(A.GGGP, i, j) * dt + P(A.ETAP, i, j));
        const double SYY2 = -P(A.SXX0, i + 1, j) * P(A.ETAP, i + 1, j) /
                            (P(A.GGGP, i + 1, j) * dt + P(A.ETAP, i + 1, j));
        const double dRdx = 0.5 * (P(A.RHOY, i, j + 1) - P(A.RHOY, i, j - 1)) / dx;
        const double dRdy = 0.5 * (P(A.RHOY, i + 1, j) - P(A.RHOY, i - 1, j)) / dy;
        const double gyij = P(A.gy, i, j);
        const double idx = 1.0 / dx, idy = 1.0 / dy;
        add(ETA1 / (dx * dx), kvy, kvy - 6 * Ny1);
        add(ETAP1 / (dy * dy), kvy, kvy - 6);
        add(-(ETAP1 + ETAP2) * (1.0 / (dy * dy)) - (ETA1 + ETA2) * (1.0 / (dx * dx)) -
                dRdy * gyij * dt,
            kvy, kvy);
        add(ETAP2 * (1.0 / (dy * dy)), kvy, kvy + 6);
        add(ETA2 * (1.0 / (dx * dx)), kvy, kvy + 6 * Ny1);
        add(ETAP1 * idx * idy - ETA2 * idx * idy - dRdx * gyij * dt * 0.25, kvy, kvx);
        add(-ETAP2 * idx * idy + ETA2 * idx * idy - dRdx * gyij * dt * 0.25, kvy, kvx + 6);
        add(-ETAP1 * idx * idy + ETA1 * idx * idy - dRdx * gyij * dt * 0.25, kvy, kvx - 6 * Ny1);
        add(ETAP2 * idx * idy - ETA1 * idx * idy - dRdx * gyij * dt * 0.25, kvy,
            kvx + 6 - 6 * Ny1);
        add(Kcont * idy, kvy, kpm);
        add(-Kcont * idy, kvy, kpm + 6);
        rvy = -P(A.RHOY, i, j) * gyij - (SXY2 - SXY1) * idx - (SYY2 - SYY1) * idy;
    }
    // P equation E:3398-3452
    const bool anchor = (i == 2 && 2 <= j && j <= Nx) || (j == 2 && 2 < i && i < Ny) ||
                        (i == Ny && 2 <= j && j <= Nx) || (j == Nx && 2 < i && i < Ny);
    const bool pedge = i == 1 || i == Ny1 || j == 1 || j == Nx1;
    if (pedge) {
        add(1.0, kpm, kpm);
    } else if (anchor) {
        add(Kcont, kpm, kpm);
        rpm = A.psurface;
    } else {
        const double phi = P(A.PHI, i, j), ep = P(A.ETAPHI, i, j), bp = P(A.BETAPHI, i, j);
        add(-1.0 / dx, kpm, kvx - 6 * Ny1);
        add(1.0 / dx, kpm, kvx);
        add(-1.0 / dy, kpm, kvy - 6);
        add(1.0 / dy, kpm, kvy);
        add(Kcont / (1.0 - phi) * ((1.0 / ep) + bp / dt), kpm, kpm);
        add(-Kcont / (1.0 - phi) * ((1.0 / ep) + bp / dt), kpm, kpf);
        rpm = (P(A.pr0, i, j) - P(A.pf0, i, j)) / (1.0 - phi) * bp / dt + P(A.DMP, i, j);
    }
    // qx Darcy equation E:3454-3480
    if (i == 1 || i == Ny1 || j == 1 || j == Nx || j == Nx1) {
        add(1.0, kqx, kqx);
        if (i == 1 && 1 < j && j < Nx) add(A.bcftop, kqx, kqx + 6);
        if (i == Ny1 && 1 < j && j < Nx) add(A.bcfbottom, kqx, kqx - 6);
    } else {
        add(P(A.RX, i, j), kqx, kqx);
        add(-Kcont * (1.0 / dx), kqx, kpf);
        add(Kcont * (1.0 / dx), kqx, kpf + 6 * Ny1);
        rqx = P(A.RHOFX, i, j) * P(A.gx, i, j);
    }
    // qy Darcy equation E:3482-3514
    if (i == 1 || i == Ny || i == Ny1 || j == 1 || j == Nx1) {
        add(1.0, kqy, kqy);
        if (j == 1 && 1 < i && i < Ny) add(A.bcfleft, kqy, kqy + 6 * Ny1);
        if (j == Nx1 && 1 < i && i < Ny) add(A.bcfright, kqy, kqy - 6 * Ny1);
    } else {
        add(P(A.RY, i, j), kqy, kqy);
        add(-Kcont * (1.0 / dy), kqy, kpf);
        add(Kcont * (1.0 / dy), kqy, kpf + 6);
        rqy = P(A.RHOFY, i, j) * P(A.gy, i, j);
    }
    // P fluid equation E:3516-3565
    if (pedge) {
        add(1.0, kpf, kpf);
    } else if (anchor) {
        add(Kcont, kpf, kpf);
        rpf = A.psurface;
    } else {
        const double phi = P(A.PHI, i, j), ep = P(A.ETAPHI, i, j), bp = P(A.BETAPHI, i, j);
        add(-(1.0 / dx), kpf, kqx - 6 * Ny1);
        add((1.0 / dx), kpf, kqx);
        add(-(1.0 / dy), kpf, kqy - 6);
        add((1.0 / dy), kpf, kqy);
        add(-Kcont / (1.0 - phi) * (1.0 / ep + bp / dt), kpf, kpm);
        add(Kcont / (1.0 - phi) * (1.0 / ep + bp / dt), kpf, kpf);
        rpf = -(P(A.pr0, i, j) - P(A.pf0, i, j)) / (1.0 - phi) * bp / dt;
    }
    if (writeR) {
        double* R = A.R + (kvx - 1);
        R[0] = rvx;
        R[1] = rvy;
        R[2] = rpm;
        R[3] = rqx;
        R[4] = rqy;
        R[5] = rpf;
    }
}

// One P cell of assemble_thermal_lse! (E:4534-4605)
template <class Sink>
__device__ __forceinline__ void thermal_cell(const AsmArgs& A, int i, int j, const Sink& add,
                                             bool writeR) {
    const int Nx1 = A.Nx1, Ny1 = A.Ny1;
    const double dx = A.dx, dy = A.dy, dt = A.dt;
    auto P = [&](const double* a, int ii, int jj) { return a[(size_t)(jj - 1) * Ny1 + (ii - 1)]; };
    const int64_t gk = (int64_t)(j - 1) * Ny1 + i;
    double rt = 0.0;
    if (i == 1 || i == Ny1 || j == 1 || j == Nx1) {
        add(1.0, gk, gk);
        if (j == 1) add(-1.0, gk, gk + Ny1);
        if (j == Nx1) add(-1.0, gk, gk - Ny1);
        if (i == 1 && 1 < j && j < Nx1) add(-1.0, gk, gk + 1);
        if (i == Ny1 && 1 < j && j < Nx1) add(-1.0, gk, gk - 1);
    } else {
        const double Kx1 = P(A.KX, i, j - 1), Kx2 = P(A.KX, i, j);
        const double Ky1 = P(A.KY, i - 1, j), Ky2 = P(A.KY, i, j);
        const double rc = P(A.RHOCP, i, j);
        add(-Kx1 * (1.0 / (dx * dx)), gk, gk - Ny1);
        add(-Ky1 * (1.0 / (dy * dy)), gk, gk - 1);
        add(rc / dt + (Kx1 + Kx2) * (1.0 / (dx * dx)) + (Ky1 + Ky2) * (1.0 / (dy * dy)), gk, gk);
        add(-Ky2 * (1.0 / (dy * dy)), gk, gk + 1);
        add(-Kx2 * (1.0 / (dx * dx)), gk, gk + Ny1);
        rt = rc / dt * P(A.tk1, i, j) + P(A.HR, i, j) + P(A.HA, i, j) + P(A.HS, i, j) +
             P(A.DHP, i, j);
    }
    if (writeR) A.R[gk - 1] = rt;
}

template <int SYS>
__global__ __launch_bounds__(256) void k_asm_mask(AsmArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.Nx1 * A.Ny1) return;
    const int i = k % A.Ny1 + 1, j = k / A.Ny1 + 1;
    if (SYS == 0)
        stokes_cell(A, i, j, MaskSink<6>{A.mask, A.Ny1}, true);
    else
        thermal_cell(A, i, j, MaskSink<1>{A.mask, A.Ny1}, true);
}

__global__ __launch_bounds__(256) void k_mask_count(const uint64_t* __restrict__ mask, int64_t n,
                                                    int32_t* __restrict__ cnt) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) cnt[c] = (int32_t)__popcll(mask[c]);
}

__global__ __launch_bounds__(256) void k_fill_u64(uint64_t* __restrict__ p, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) p[c] = 0;
}

template <int SYS>
__global__ __launch_bounds__(256) void k_asm_fill(AsmArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.Nx1 * A.Ny1) return;
    const int i = k % A.Ny1 + 1, j = k / A.Ny1 + 1;
    if (SYS == 0)
        stokes_cell(A, i, j, FillSink<6>{A.colptr, A.mask, A.rowval, A.nzval, A.Ny1}, false);
    else
        thermal_cell(A, i, j, FillSink<1>{A.colptr, A.mask, A.rowval, A.nzval, A.Ny1}, false);
}

static AsmArgs asm_args(ereb_ctx* c, double dt, Sys& s) {
    AsmArgs A;
    A.Nx = c->Nx;
    A.Ny = c->Ny;
    A.Nx1 = c->Nx1;
    A.Ny1 = c->Ny1;
    A.dx = c->dx;
    A.dy = c->dy;
    A.dt = dt;
    const ereb_materials& M = c->mat;
    A.Kcont = M.Kcont;
    A.psurface = M.psurface;
    A.bcleft = M.bcleft; A.bcright = M.bcright; A.bctop = M.bctop; A.bcbottom = M.bcbottom;
    A.bcfleft = M.bcfleft; A.bcfright = M.bcfright; A.bcftop = M.bcftop;
    A.bcfbottom = M.bcfbottom;
    A.vxleft = M.vxleft; A.vxright = M.vxright; A.vytop = M.vytop; A.vybottom = M.vybottom;
    A.ETA = gptr(c, EREB_G_ETA); A.ETAP = gptr(c, EREB_G_ETAP); A.GGG = gptr(c, EREB_G_GGG);
    A.GGGP = gptr(c, EREB_G_GGGP); A.SXY0 = gptr(c, EREB_G_SXY0); A.SXX0 = gptr(c, EREB_G_SXX0);
    A.RHOX = gptr(c, EREB_G_RHOX); A.RHOY = gptr(c, EREB_G_RHOY); A.RHOFX = gptr(c, EREB_G_RHOFX);
    A.RHOFY = gptr(c, EREB_G_RHOFY); A.RX = gptr(c, EREB_G_RX); A.RY = gptr(c, EREB_G_RY);
    A.ETAPHI = gptr(c, EREB_G_ETAPHI); A.BETAPHI = gptr(c, EREB_G_BETAPHI);
    A.PHI = gptr(c, EREB_G_PHI); A.gx = gptr(c, EREB_G_GX); A.gy = gptr(c, EREB_G_GY);
    A.pr0 = gptr(c, EREB_G_PR0); A.pf0 = gptr(c, EREB_G_PF0); A.DMP = gptr(c, EREB_G_DMP);
    A.tk1 = gptr(c, EREB_G_TK1); A.RHOCP = gptr(c, EREB_G_RHOCP); A.KX = gptr(c, EREB_G_KX);
    A.KY = gptr(c, EREB_G_KY); A.HR = gptr(c, EREB_G_HR); A.HA = gptr(c, EREB_G_HA);
    A.HS = gptr(c, EREB_G_HS); A.DHP = gptr(c, EREB_G_DHP);
    A.R = s.rhs;
    A.colcnt = s.colcnt;
    A.colptr = s.colptr;
    A.mask = s.mask;
    A.rowval = s.rowval;
    A.nzval = s.nzval;
    return A;
}

template <int SYS>
static int assemble(ereb_ctx* c, double dt) {
    const int64_t ncell = (int64_t)c->Nx1 * c->Ny1;
    // at most 40 entries per P cell (Stokes-Darcy: 11+11+6+3+3+6), 5 (thermal)
    const int64_t n = SYS == 0 ? 6 * ncell : ncell;
    const int64_t cap = SYS == 0 ? 40 * ncell : 5 * ncell;
    int rc = ensure_sys(c, SYS, n, cap);
    if (rc) return rc;
    Sys& s = c->sys[SYS];
    AsmArgs A = asm_args(c, dt, s);
    hipEvent_t ev;
    prof_begin(c, SYS == 0 ? EREB_K_ASM_SD : EREB_K_ASM_T, &ev);
    const unsigned ncol = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u64, dim3(ncol), dim3(256), 0, c->stream, s.mask, n);
    const unsigned nb = (unsigned)((ncell + 255) / 256);
    hipLaunchKernelGGL(k_asm_mask<SYS>, dim3(nb), dim3(256), 0, c->stream, A);
    hipLaunchKernelGGL(k_mask_count, dim3(ncol), dim3(256), 0, c->stream, s.mask, n, s.colcnt);
    if ((rc = scan_exclusive(c, s.colcnt, s.colptr, n))) return rc;
    hipLaunchKernelGGL(k_asm_fill<SYS>, dim3(nb), dim3(256), 0, c->stream, A);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, SYS == 0 ? EREB_K_ASM_SD : EREB_K_ASM_T, ev);
    s.valid = true;
    return EREB_OK;
}

int launch_assemble_stokes(ereb_ctx* c, double dt) { return assemble<0>(c, dt); }
int launch_assemble_thermal(ereb_ctx* c, double dt) { return assemble<1>(c, dt); }

}  // namespace ereb
