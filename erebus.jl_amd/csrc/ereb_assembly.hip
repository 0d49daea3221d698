// ereb_assembly.hip -- step 5: assemble_hydromechanical_lse! (E:3089-3572)
// and assemble_thermal_lse! (E:4526-4611) straight into CSC.
//
// ExtendableSparse's updateindex!(L, +, v, row, col) creates no entry when
// v == 0 and adds v to an existing entry, so the pattern depends on the
// values.  Every value is computed by the same expression as the reference
// (no fma contraction), which makes the zero tests, the pattern and the
// stored values bit-exact.
//
// Every equation references a fixed stencil of columns relative to its own
// row: at most 11 (row, column) offsets per Stokes-Darcy equation, 5 for the
// thermal one.  Each offset is a "slot" of that equation; the slot tables
// are discovered once per process by running the same equation code on the
// host with a recording sink, so they cannot drift from the kernels.
//   1. rows  : one thread per (P cell, equation) evaluates its row, writes
//              each nonzero entry into its slot (slot-major: consecutive
//              cells are consecutive in memory) and a presence bit per slot;
//              R is written here;
//   2. count : one thread per column walks the rows that can reference it
//              (the reverse slot table) in ascending row order and counts
//              the present entries, summed per block of columns;
//   3. fill  : the same walk gives each column's start (the block sums
//              before it plus a scan within the block: colptr), writes
//              rowval of each column contiguously and, for every stored
//              entry, the index of its slot value;
//   4. values: nzval, one gather per entry through those indices (and the
//              pattern flag of step 1 cleared for the next assembly).
// No atomics and no sort: rows come out ascending in every column.
// Step 1 also compares the presence bits with the previous assembly's (a
// device flag): when the pattern is unchanged
// -- the common case from one plastic iteration to the next --
// steps 2 and 3 return at once and step 4 alone refills nzval, colptr and
// rowval being those of the previous assembly.  The condition is evaluated
// on the device, so a captured step graph takes either path.
#include <algorithm>
#include <cstring>
#include <vector>

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
};

// Stencil slot tables of one system.  A slot key encodes the column of an
// entry relative to the row's cell: ((dcj + 1) * 3 + (dci + 1)) * NU + u, with
// (dci, dcj) in {-1, 0, 1}^2 the column cell minus the row cell and u the
// column's unknown.
constexpr int SLOT_E = 6;    // equations (unknowns) per cell, at most
constexpr int SLOT_S = 16;   // slots per equation, at most (presence bits in uint16)
constexpr int SLOT_K = 54;   // keys 0 .. 9*NU-1
constexpr int SLOT_R = 32;   // reverse entries per column unknown, at most
struct SlotTables {
    int ne, nstot;                       // equations per cell, slots per cell
    int slot0[SLOT_E];                   // first slot of equation e within its cell
    int ns[SLOT_E];                      // slots of equation e
    int8_t lut[SLOT_E][SLOT_K];          // key -> slot of equation e, -1 if none
    int nrev[SLOT_E];                    // reverse entries of column unknown u
    int8_t rev_e[SLOT_E][SLOT_R], rev_s[SLOT_E][SLOT_R];
    int8_t rev_di[SLOT_E][SLOT_R], rev_dj[SLOT_E][SLOT_R];  // column cell - row cell
};

// key of the column offset `off` = col - row of a row of equation e
__host__ __device__ __forceinline__ int slot_key(int64_t off, int e, int nu, int Ny1) {
    const int64_t t = off + e;  // column index relative to the first unknown of the row cell
    const int64_t dcell = t >= 0 ? t / nu : -((-t + nu - 1) / nu);
    const int u = (int)(t - dcell * nu);
    const int64_t dcj = dcell >= 0 ? (dcell + 1) / Ny1 : -((-dcell + 1) / Ny1);
    const int64_t dci = dcell - dcj * Ny1;
    if (dci < -1 || dci > 1 || dcj < -1 || dcj > 1) return -1;
    return (int)(((dcj + 1) * 3 + (dci + 1)) * nu + u);
}

// Row sink: the entry (row, col) goes to its slot; the offset col - row is the
// same for all lanes at every call site (the stencils are fixed), so it is
// read from the first active lane and the table lookups are scalar; a lane
// whose offset differs, or an offset outside the table, raises flag 1.
struct SlotSink {
    const SlotTables* __restrict__ T;
    double* slotv;     // value of slot 0 of the row; slot s at slotv[s * stride]
    int64_t stride;    // P cells: slots are stored slot-major
    uint32_t* pres;    // presence bits of the row (register)
    int32_t* flags;
    int e, nu, Ny1;
    __device__ void operator()(double v, int64_t row, int64_t col) const {
        const int64_t off = col - row;
        const int offu = __builtin_amdgcn_readfirstlane((int)off);
        const int key = slot_key(offu, e, nu, Ny1);
        const int sl = key < 0 ? -1 : (int)T->lut[e][key];
        if (off != offu || sl < 0) {
            flags[1] = 1;
            return;
        }
        if (v != 0.0) {
            const uint32_t bit = 1u << sl;
            if (*pres & bit)
                slotv[sl * stride] += v;
            else
                slotv[sl * stride] = v;
            *pres |= bit;
        }
    }
};

// One equation EQ (0 Vx, 1 Vy, 2 P, 3 qx, 4 qy, 5 P fluid) of P cell (i, j)
// of assemble_hydromechanical_lse!: row kvx + EQ.
template <int EQ, class Sink>
__host__ __device__ __forceinline__ void stokes_cell(const AsmArgs& A, int i, int j,
                                                     const Sink& add, bool writeR) {
    const int Nx = A.Nx, Ny = A.Ny, Nx1 = A.Nx1, Ny1 = A.Ny1;
    const double dx = A.dx, dy = A.dy, dt = A.dt, Kcont = A.Kcont;
    auto B = [&](const double* a, int ii, int jj) { return a[(size_t)(jj - 1) * Ny + (ii - 1)]; };
    auto P = [&](const double* a, int ii, int jj) { return a[(size_t)(jj - 1) * Ny1 + (ii - 1)]; };
    const int64_t kvx = ((int64_t)(j - 1) * Ny1 + i - 1) * 6 + 1;
    const int64_t kvy = kvx + 1, kpm = kvx + 2, kqx = kvx + 3, kqy = kvx + 4, kpf = kvx + 5;
    double rvx = 0.0, rvy = 0.0, rpm = 0.0, rqx = 0.0, rqy = 0.0, rpf = 0.0;
    if (EQ == 0) {
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
    }
    if (EQ == 1) {
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
    }
    // ghost and pressure-anchor cells of both continuity equations
    const bool anchor = (i == 2 && 2 <= j && j <= Nx) || (j == 2 && 2 < i && i < Ny) ||
                        (i == Ny && 2 <= j && j <= Nx) || (j == Nx && 2 < i && i < Ny);
    const bool pedge = i == 1 || i == Ny1 || j == 1 || j == Nx1;
    if (EQ == 2) {
        // P equation E:3398-3452
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
    }
    if (EQ == 3) {
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
    }
    if (EQ == 4) {
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
    }
    if (EQ == 5) {
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
    }
    if (writeR) {
        const double r[6] = {rvx, rvy, rpm, rqx, rqy, rpf};
        A.R[kvx - 1 + EQ] = r[EQ];
    }
}

// One P cell of assemble_thermal_lse! (E:4534-4605)
template <class Sink>
__host__ __device__ __forceinline__ void thermal_cell(const AsmArgs& A, int i, int j,
                                                      const Sink& add, bool writeR) {
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

template <class Sink>
__host__ __device__ __forceinline__ void stokes_eq(const AsmArgs& A, int e, int i, int j,
                                                   const Sink& add, bool writeR) {
    switch (e) {
        case 0: stokes_cell<0>(A, i, j, add, writeR); break;
        case 1: stokes_cell<1>(A, i, j, add, writeR); break;
        case 2: stokes_cell<2>(A, i, j, add, writeR); break;
        case 3: stokes_cell<3>(A, i, j, add, writeR); break;
        case 4: stokes_cell<4>(A, i, j, add, writeR); break;
        default: stokes_cell<5>(A, i, j, add, writeR); break;
    }
}

// 1. rows: thread per P cell, blockIdx.y = equation (wave-uniform)
template <int SYS>
// prev, chg (optional): the presence bits of the previous assembly; where
// they differ, prev is updated and chg[0] set to 1 (every writer stores the
// same value), so chg[0] tells whether the pattern changed
__global__ __launch_bounds__(256) void k_asm_rows(AsmArgs A0, const StepScalars* __restrict__ S,
                                                  const SlotTables* __restrict__ T,
                                                  uint16_t* __restrict__ pres_out,
                                                  double* __restrict__ slotv,
                                                  int32_t* __restrict__ flags,
                                                  uint16_t* __restrict__ prev,
                                                  int32_t* __restrict__ chg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A0.Nx1 * A0.Ny1) return;
    AsmArgs A = A0;
    A.dt = S->dt;  // the step's dt from the device block
    const int e = blockIdx.y;
    const int i = k % A.Ny1 + 1, j = k / A.Ny1 + 1;
    const int64_t ncell = (int64_t)A.Nx1 * A.Ny1;
    uint32_t pres = 0;
    SlotSink sink{T, slotv + (size_t)T->slot0[e] * ncell + k, ncell, &pres, flags, e,
                  SYS == 0 ? 6 : 1, A.Ny1};
    if (SYS == 0)
        stokes_eq(A, e, i, j, sink, true);
    else
        thermal_cell(A, i, j, sink, true);
    const size_t q = (size_t)e * ncell + k;
    pres_out[q] = (uint16_t)pres;
    if (prev && prev[q] != (uint16_t)pres) {
        prev[q] = (uint16_t)pres;
        chg[0] = 1;
    }
}

// 2. column counts per block of columns: one thread per column walks the
// rows that can hold an entry in it (the reverse slot table of its unknown),
// and the block writes the number of entries of its columns, bsum[block];
// the fill then computes each column's start from these and its own count,
// so the column pointers need no separate scan.  The slot tables are staged
// in LDS.  run (optional): the kernel does its work only when run[0] != 0.

// exclusive scan over the workgroup (blockDim.x <= 1024); sh holds 17 ints
__device__ __forceinline__ int32_t asm_block_scan(int32_t v, int32_t* sh, int32_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    int32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        const int32_t w = lane < nw ? sh[lane] : 0;
        int32_t t = w;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < nw) sh[lane] = t - w;
        if (lane == nw - 1) sh[16] = t;
    }
    __syncthreads();
    const int32_t r = sh[wid] + x - v;
    *total = sh[16];
    __syncthreads();  // sh is reused by the caller
    return r;
}

__global__ __launch_bounds__(256) void k_asm_count(int Nx1, int Ny1,
                                                   const SlotTables* __restrict__ Tg,
                                                   const uint16_t* __restrict__ pres,
                                                   int32_t* __restrict__ bsum,
                                                   const int32_t* __restrict__ run) {
    if (run && run[0] == 0) return;
    __shared__ SlotTables T;
    __shared__ int32_t sh[17];
    static_assert(sizeof(SlotTables) % 4 == 0, "table size");
    for (int q = threadIdx.x; q < (int)(sizeof(SlotTables) / 4); q += blockDim.x)
        ((int32_t*)&T)[q] = ((const int32_t*)Tg)[q];
    __syncthreads();
    const int ne = T.ne;
    const int64_t ncell = (int64_t)Nx1 * Ny1;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t p = 0;
    if (col < ncell * ne) {
        const int u = (int)(col % ne);
        const int64_t k = col / ne;
        const int ic = (int)(k % Ny1), jc = (int)(k / Ny1);
        const int nr = T.nrev[u];
        for (int r = 0; r < nr; ++r) {
            const int ir = ic - T.rev_di[u][r], jr = jc - T.rev_dj[u][r];
            if (ir < 0 || ir >= Ny1 || jr < 0 || jr >= Nx1) continue;
            const int e = T.rev_e[u][r], sl = T.rev_s[u][r];
            const int64_t cell = (int64_t)jr * Ny1 + ir;
            p += (pres[(size_t)e * ncell + cell] >> sl) & 1;
        }
    }
    int32_t tot;
    asm_block_scan(p, sh, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// 3. colptr/rowval/nzval: the same column walk; a column's start is the sum
// of the block sums of the blocks before its own plus the exclusive scan of
// the counts within the block (colptr written here, the last block adding
// colptr[n]).  The entries of the block's columns are staged in LDS at their
// CSC positions and then written out by consecutive threads (a thread's own
// entries are ~7 apart from its neighbours', so direct stores would fill a
// fraction of each cache line).  Rows come out ascending in every column
// with no sort.
// blockDim.x * (most entries of a column) <= ASM_STAGE, checked on the host.
constexpr int ASM_STAGE = 4096;

// MAXR: most candidate rows of a column (16, or SLOT_R); every candidate's
// presence bit is loaded before any is used, so a column costs one memory
// round trip instead of one per candidate.  VALUES: nzval written here too
// (each candidate's value loaded with its bit, a value whose bit is clear
// dropped); otherwise only rowval and the slot index map, and k_asm_values
// writes nzval by one gather per stored entry (a quarter of the candidate
// loads, and half the LDS: twice the workgroups per CU).
// IDX, BASE: the row index type and base of the output (int32 0-based for
// the context's device CSC, Int64 1-based for SparseMatrixCSC{Float64,Int64})
// run (optional): only when run[0] != 0; map (optional): the slot value's
// index of every stored entry, for the values-only refill
template <int MAXR, class IDX = int32_t, int BASE = 0, bool VALUES = false>
__global__ __launch_bounds__(256) void k_asm_fill(int Nx1, int Ny1,
                                                  const SlotTables* __restrict__ Tg,
                                                  const uint16_t* __restrict__ pres,
                                                  const double* __restrict__ slotv,
                                                  const int32_t* __restrict__ bsum,
                                                  int32_t* __restrict__ colptr,
                                                  IDX* __restrict__ rowval,
                                                  double* __restrict__ nzval,
                                                  const int32_t* __restrict__ run,
                                                  int32_t* __restrict__ map) {
    if (run && run[0] == 0) return;
    __shared__ SlotTables T;
    __shared__ int32_t srow[ASM_STAGE];
    __shared__ int32_t smap[ASM_STAGE];
    __shared__ double sval[VALUES ? ASM_STAGE : 1];
    __shared__ int32_t sh[17];
    for (int q = threadIdx.x; q < (int)(sizeof(SlotTables) / 4); q += blockDim.x)
        ((int32_t*)&T)[q] = ((const int32_t*)Tg)[q];
    // entries of the blocks before this one
    int32_t before = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += blockDim.x) before += bsum[b];
    int32_t base;
    asm_block_scan(before, sh, &base);  // synchronises: T is staged
    const int ne = T.ne;
    const int64_t ncell = (int64_t)Nx1 * Ny1, n = ncell * ne;
    const int64_t col0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t col = col0 + threadIdx.x;
    bool on[MAXR];
    int32_t rw[MAXR], mp[MAXR];
    double vv[MAXR];
    int32_t cnt = 0;
    if (col < n) {
        const int u = (int)(col % ne);
        const int64_t k = col / ne;
        const int ic = (int)(k % Ny1), jc = (int)(k / Ny1);
        const int nr = T.nrev[u];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            on[r] = false;
            rw[r] = mp[r] = 0;
            vv[r] = 0.0;
            const int ir = ic - T.rev_di[u][r], jr = jc - T.rev_dj[u][r];
            if (r < nr && ir >= 0 && ir < Ny1 && jr >= 0 && jr < Nx1) {
                const int e = T.rev_e[u][r], sl = T.rev_s[u][r];
                const int64_t cell = (int64_t)jr * Ny1 + ir;
                on[r] = (pres[(size_t)e * ncell + cell] >> sl) & 1;
                mp[r] = (int32_t)(((int64_t)T.slot0[e] + sl) * ncell + cell);
                if (VALUES) vv[r] = slotv[mp[r]];
                rw[r] = (int32_t)(cell * ne + e);
            }
            cnt += on[r] ? 1 : 0;
        }
    }
    int32_t total;
    int32_t p = asm_block_scan(cnt, sh, &total);
    if (col < n) {
        colptr[col] = base + p;
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            if (on[r]) {
                srow[p] = rw[r];
                smap[p] = mp[r];
                if (VALUES) sval[p] = vv[r];
                ++p;
            }
        }
    }
    if (col == n - 1) colptr[n] = base + total;
    __syncthreads();
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
        rowval[base + q] = (IDX)srow[q] + BASE;
        if (VALUES) nzval[base + q] = sval[q];
        if (map) map[base + q] = smap[q];
    }
}

// nzval through the stored slot indices (after every fill, and alone while
// the pattern is unchanged): one coalesced write, one gather per entry, in a
// grid-stride loop over the entries (nnz is known on the device only) with
// ASM_VALUES_PER entries in flight per thread.  chg (optional): the
// pattern-change flag, read by the kernels before this one and cleared here
// for the next assembly's rows kernel.
constexpr int ASM_VALUES_PER = 4;
constexpr int ASM_VALUES_BLOCKS = 8192;
__global__ __launch_bounds__(256) void k_asm_values(const int32_t* __restrict__ colptr_end,
                                                    const int32_t* __restrict__ map,
                                                    const double* __restrict__ slotv,
                                                    double* __restrict__ nzval,
                                                    int32_t* __restrict__ chg) {
    if (chg && blockIdx.x == 0 && threadIdx.x == 0) chg[0] = 0;
    const int64_t nnz = colptr_end[0];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q0 < nnz;
         q0 += ASM_VALUES_PER * stride) {
        int32_t mi[ASM_VALUES_PER];
        double v[ASM_VALUES_PER];
#pragma unroll
        for (int u = 0; u < ASM_VALUES_PER; ++u) {
            const int64_t q = q0 + u * stride;
            mi[u] = q < nnz ? map[q] : 0;
        }
#pragma unroll
        for (int u = 0; u < ASM_VALUES_PER; ++u) v[u] = slotv[mi[u]];
#pragma unroll
        for (int u = 0; u < ASM_VALUES_PER; ++u) {
            const int64_t q = q0 + u * stride;
            if (q < nnz) nzval[q] = v[u];
        }
    }
}

// 0-based int32 column pointers to 1-based Int64
__global__ void k_colptr_1based(const int32_t* __restrict__ a, int64_t* __restrict__ b,
                                int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int64_t)a[i] + 1;
}

// ---- slot tables, discovered on the host from the equation code -----------

struct RecordSink {
    std::vector<int>* keys;  // keys of equation e in first-use order
    int e, nu, Ny1;
    bool* bad;
    void operator()(double, int64_t row, int64_t col) const {
        const int key = slot_key(col - row, e, nu, Ny1);
        if (key < 0) {
            *bad = true;
            return;
        }
        for (int k : *keys)
            if (k == key) return;
        keys->push_back(key);
    }
};

template <int SYS>
static bool discover_slots(SlotTables& T) {
    // a small grid with every boundary, anchor and interior case; the
    // recording sink keeps every add, zero or not
    const int Nx = 9, Ny = 9, Nx1 = 10, Ny1 = 10;
    std::vector<double> one((size_t)Nx1 * Ny1, 1.0), rhs((size_t)Nx1 * Ny1 * 6, 0.0);
    AsmArgs A{};
    A.Nx = Nx; A.Ny = Ny; A.Nx1 = Nx1; A.Ny1 = Ny1;
    A.dx = A.dy = A.dt = A.Kcont = 1.0;
    const double* o = one.data();
    A.ETA = A.ETAP = A.GGG = A.GGGP = A.SXY0 = A.SXX0 = A.RHOX = A.RHOY = o;
    A.RHOFX = A.RHOFY = A.RX = A.RY = A.ETAPHI = A.BETAPHI = A.PHI = o;
    A.gx = A.gy = A.pr0 = A.pf0 = A.DMP = o;
    A.tk1 = A.RHOCP = A.KX = A.KY = A.HR = A.HA = A.HS = A.DHP = o;
    std::vector<double> phi((size_t)Nx1 * Ny1, 0.5);
    A.PHI = phi.data();
    A.R = rhs.data();
    const int ne = SYS == 0 ? 6 : 1, nu = ne;
    std::vector<int> keys[SLOT_E];
    bool bad = false;
    for (int e = 0; e < ne; ++e) {
        RecordSink rec{&keys[e], e, nu, Ny1, &bad};
        for (int j = 1; j <= Nx1; ++j)
            for (int i = 1; i <= Ny1; ++i) {
                if (SYS == 0)
                    stokes_eq(A, e, i, j, rec, false);
                else
                    thermal_cell(A, i, j, rec, false);
            }
    }
    if (bad) return false;
    T = SlotTables{};
    T.ne = ne;
    int tot = 0;
    for (int e = 0; e < ne; ++e) {
        if ((int)keys[e].size() > SLOT_S) return false;
        T.slot0[e] = tot;
        T.ns[e] = (int)keys[e].size();
        tot += T.ns[e];
        for (int k = 0; k < SLOT_K; ++k) T.lut[e][k] = -1;
        for (int sl = 0; sl < T.ns[e]; ++sl) T.lut[e][keys[e][sl]] = (int8_t)sl;
    }
    T.nstot = tot;
    // reverse table: rows referencing a column of unknown u, ascending row
    // order = (row cell column, row cell row, equation) ascending
    struct Rev { int dj, di, e, sl; };
    for (int u = 0; u < nu; ++u) {
        std::vector<Rev> r;
        for (int e = 0; e < ne; ++e)
            for (int sl = 0; sl < T.ns[e]; ++sl) {
                const int key = keys[e][sl];
                if (key % nu != u) continue;
                const int dci = (key / nu) % 3 - 1, dcj = key / (3 * nu) - 1;
                r.push_back(Rev{dcj, dci, e, sl});
            }
        std::sort(r.begin(), r.end(), [](const Rev& a, const Rev& b) {
            if (a.dj != b.dj) return a.dj > b.dj;  // row cell = column cell - offset
            if (a.di != b.di) return a.di > b.di;
            return a.e < b.e;
        });
        if ((int)r.size() > SLOT_R) return false;
        T.nrev[u] = (int)r.size();
        for (size_t k = 0; k < r.size(); ++k) {
            T.rev_e[u][k] = (int8_t)r[k].e;
            T.rev_s[u][k] = (int8_t)r[k].sl;
            T.rev_di[u][k] = (int8_t)r[k].di;
            T.rev_dj[u][k] = (int8_t)r[k].dj;
        }
    }
    return true;
}

int host_slot_check(int64_t* slots_sd, int64_t* slots_t) {
    SlotTables A, B;
    if (!discover_slots<0>(A) || !discover_slots<1>(B)) return EREB_ERR_STATE;
    // six unknowns per Stokes-Darcy cell; the thermal equation's 5-point stencil
    if (A.ne != 6 || B.ne != 1 || B.nstot != 5) return EREB_ERR_STATE;
    if (slots_sd) *slots_sd = A.nstot;
    if (slots_t) *slots_t = B.nstot;
    return EREB_OK;
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
    return A;
}

template <int SYS>
static int assemble(ereb_ctx* c, double dt) {
    Sys& s = c->sys[SYS];
    if (!s.tables) {
        SlotTables T;
        if (!discover_slots<SYS>(T))
            return fail(c, EREB_ERR_STATE, "assembly: stencil slot discovery failed");
        EREB_HIP(c, hipMalloc(&s.tables, sizeof(SlotTables)));
        EREB_HIP(c, hipMemcpy(s.tables, &T, sizeof T, hipMemcpyHostToDevice));
        s.ne = T.ne;
        s.nstot = T.nstot;
        s.maxrev = 0;
        for (int u = 0; u < T.ne; ++u) s.maxrev = std::max(s.maxrev, T.nrev[u]);
    }
    const int64_t ncell = (int64_t)c->Nx1 * c->Ny1;
    const int64_t n = ncell * s.ne;
    const int64_t cap = ncell * s.nstot;  // every slot present
    int rc = ensure_sys(c, SYS, n, cap, ncell);
    if (rc) return rc;
    if ((rc = set_dt(c, dt))) return rc;
    AsmArgs A = asm_args(c, dt, s);
    const SlotTables* T = (const SlotTables*)s.tables;
    hipEvent_t ev;
    prof_begin(c, SYS == 0 ? EREB_K_ASM_SD : EREB_K_ASM_T, &ev);
    // pattern unchanged since the previous assembly?  The rows kernel compares
    // the presence bits as it writes them and raises the flag, which the
    // values kernel clears at the end of every assembly; with no previous
    // assembly (a fresh system) it is raised here.
    if (!s.have_prev && (rc = fill_i32(c, s.chg, 1, 1))) return rc;
    const dim3 grid((unsigned)((ncell + 255) / 256), (unsigned)s.ne);
    hipLaunchKernelGGL(k_asm_rows<SYS>, grid, dim3(256), 0, c->stream, A, c->d_scal, T, s.pres,
                       s.slotv, c->d_flags, s.pres_prev, s.chg);
    s.have_prev = true;
    const int32_t* chg = s.chg;
    // a block's entries must fit the LDS stage: 256 columns of up to 16
    // entries, else 128 of up to 32 (SLOT_R); the count's blocks are the
    // fill's
    const bool narrow = s.maxrev <= ASM_STAGE / 256;
    static_assert(ASM_STAGE / 256 == 16, "16 candidates per column in the 256-column blocks");
    static_assert(ASM_STAGE / 128 >= SLOT_R, "stage holds 128 columns of SLOT_R entries");
    const int fb = narrow ? 256 : 128;
    const dim3 fgrid((unsigned)((n + fb - 1) / fb));
    hipLaunchKernelGGL(k_asm_count, fgrid, dim3(fb), 0, c->stream, c->Nx1, c->Ny1, T, s.pres,
                       s.colcnt, chg);
    if (narrow)
        hipLaunchKernelGGL(k_asm_fill<16>, fgrid, dim3(fb), 0, c->stream, c->Nx1, c->Ny1, T,
                           s.pres, s.slotv, (const int32_t*)s.colcnt, s.colptr, s.rowval,
                           s.nzval, chg, s.map);
    else
        hipLaunchKernelGGL(k_asm_fill<SLOT_R>, fgrid, dim3(fb), 0, c->stream, c->Nx1, c->Ny1, T,
                           s.pres, s.slotv, (const int32_t*)s.colcnt, s.colptr, s.rowval,
                           s.nzval, chg, s.map);
    hipLaunchKernelGGL(k_asm_values,
                       dim3((unsigned)std::min<int64_t>(
                           (cap + 256 * ASM_VALUES_PER - 1) / (256 * ASM_VALUES_PER),
                           ASM_VALUES_BLOCKS)),
                       dim3(256), 0, c->stream, (const int32_t*)s.colptr + n,
                       (const int32_t*)s.map, (const double*)s.slotv, s.nzval, s.chg);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, SYS == 0 ? EREB_K_ASM_SD : EREB_K_ASM_T, ev);
    s.valid = true;
    return EREB_OK;
}

int launch_assemble_stokes(ereb_ctx* c, double dt) { return assemble<0>(c, dt); }
int launch_assemble_thermal(ereb_ctx* c, double dt) { return assemble<1>(c, dt); }

// ---- thermal assembly into caller buffers, scratch supplied ------------------
// Scratch layout (each piece 256-byte aligned): the step scalars (dt), the
// flags, the slot tables, the presence bits, the slot values, the block sums
// of the column counts and the 0-based column pointers.
namespace {
struct ThermalScratch {
    StepScalars* scal;
    int32_t* flags;
    SlotTables* tables;
    uint16_t* pres;
    double* slotv;
    int32_t* colcnt;  // block sums of the column counts
    int32_t* colptr;
    size_t bytes;
};
constexpr int THERMAL_SLOTS = 5;  // slots per cell of the 5-point equation

ThermalScratch thermal_carve(char* base, int64_t ncell) {
    ThermalScratch t{};
    size_t off = 0;
    auto take = [&](size_t b) {
        char* q = base ? base + off : nullptr;
        off += (b + 255) / 256 * 256;
        return q;
    };
    t.scal = (StepScalars*)take(sizeof(StepScalars));
    t.flags = (int32_t*)take(4 * sizeof(int32_t));
    t.tables = (SlotTables*)take(sizeof(SlotTables));
    t.pres = (uint16_t*)take(sizeof(uint16_t) * ncell);
    t.slotv = (double*)take(sizeof(double) * THERMAL_SLOTS * ncell);
    t.colcnt = (int32_t*)take(sizeof(int32_t) * ncell);
    t.colptr = (int32_t*)take(sizeof(int32_t) * (ncell + 1));
    t.bytes = off;
    return t;
}
}  // namespace

size_t thermal_scratch_bytes(int64_t ncell) { return thermal_carve(nullptr, ncell).bytes; }

ThermalScratchHead thermal_scratch_head(char* scratch) {
    const ThermalScratch t = thermal_carve(scratch, 0);  // these pieces do not depend on ncell
    return ThermalScratchHead{t.scal, t.flags, (void*)t.tables};
}

int thermal_tables_bytes() { return (int)sizeof(SlotTables); }

int thermal_tables_host(void* out) {
    SlotTables T;
    if (!discover_slots<1>(T) || T.nstot != THERMAL_SLOTS || T.ne != 1) return EREB_ERR_STATE;
    for (int u = 0; u < T.ne; ++u)
        if (T.nrev[u] > ASM_STAGE / 256) return EREB_ERR_STATE;
    std::memcpy(out, &T, sizeof T);
    return EREB_OK;
}

// assemble_thermal_lse! (E:4526-4611) with every buffer supplied: the slot
// tables already in the scratch, dt in its StepScalars
int launch_thermal_noalloc(ereb_ctx* c, char* scratch, const double* const in[8], double* RT,
                           int64_t* colptr, int64_t* rowval, double* nzval) {
    const int64_t ncell = (int64_t)c->Nx1 * c->Ny1;
    const ThermalScratch t = thermal_carve(scratch, ncell);
    AsmArgs A{};
    A.Nx = c->Nx;
    A.Ny = c->Ny;
    A.Nx1 = c->Nx1;
    A.Ny1 = c->Ny1;
    A.dx = c->dx;
    A.dy = c->dy;
    A.tk1 = in[0];
    A.RHOCP = in[1];
    A.KX = in[2];
    A.KY = in[3];
    A.HR = in[4];
    A.HA = in[5];
    A.HS = in[6];
    A.DHP = in[7];
    A.R = RT;
    const dim3 grid((unsigned)((ncell + 255) / 256), 1);
    hipLaunchKernelGGL(k_asm_rows<1>, grid, dim3(256), 0, c->stream, A,
                       (const StepScalars*)t.scal, (const SlotTables*)t.tables, t.pres, t.slotv,
                       t.flags, (uint16_t*)nullptr, (int32_t*)nullptr);
    hipLaunchKernelGGL(k_asm_count, grid, dim3(256), 0, c->stream, c->Nx1, c->Ny1,
                       (const SlotTables*)t.tables, (const uint16_t*)t.pres, t.colcnt,
                       (const int32_t*)nullptr);
    hipLaunchKernelGGL((k_asm_fill<16, int64_t, 1, true>), grid, dim3(256), 0, c->stream, c->Nx1,
                       c->Ny1, (const SlotTables*)t.tables, (const uint16_t*)t.pres,
                       (const double*)t.slotv, (const int32_t*)t.colcnt, t.colptr, rowval,
                       nzval, (const int32_t*)nullptr, (int32_t*)nullptr);
    hipLaunchKernelGGL(k_colptr_1based, dim3((unsigned)((ncell + 1 + 255) / 256)), dim3(256), 0,
                       c->stream, (const int32_t*)t.colptr, colptr, ncell + 1);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

}  // namespace ereb
