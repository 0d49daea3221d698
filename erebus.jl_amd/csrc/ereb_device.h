// ereb_device.h -- device-side arithmetic shared by the gfx950 kernels.
//
// Every expression follows the reference operator by operator so that, with
// -ffp-contract=off, integer cell indices and Float64 values agree bit for bit
// with the CPU restatement (E: = src/Erebus.jl).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ereb.h"

namespace ereb {

// Geometry as the kernels see it.  The Vx axes equal (x, yp) and the Vy axes
// equal (xp, y) element for element (checked when the context is created),
// so four axis arrays describe all four grids.
struct GeomK {
    int Nx, Ny, Nx1, Ny1;
    double dx, dy, idx, idy;  // idx = inv(dx) = 1.0/dx as in fix (E:1457)
    const double* x;          // basic x axis (Nx), = xvx
    const double* y;          // basic y axis (Ny), = yvy
    const double* xp;         // P x axis (Nx1), = xvy
    const double* yp;         // P y axis (Ny1), = yvx
};

// XCD-contiguous workgroup number: the hardware deals the workgroups of a
// launch to the 8 XCDs round robin (workgroup b runs on XCD b mod 8), each
// XCD with its own L2; renumbered XCD-major, each XCD takes one contiguous
// run of the grid, so neighbouring workgroups (neighbouring tiles, runs of
// markers) share one L2 and the lines straddling their data are fetched
// once.  A bijection of [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x, q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

// Node k (1-based) of an axis with spacing h.  EXACT (the context found
// every node of every axis equal to the first node plus k*h rounded once,
// as for integer or half-integer metres): computed in one fused operation,
// which spares a dependent load before the stencil arithmetic; otherwise
// loaded.  Bit-identical either way.
template <bool EXACT>
__device__ __forceinline__ double axis_node(const double* __restrict__ ax, double h, int k) {
    if constexpr (EXACT) return fma((double)(k - 1), h, ax[0]);
    else return ax[k - 1];
}

// unsafe_trunc(Int, t) + 1 clamped to [lo, hi] (E:1457-1468).  NaN and
// |t| >= 2^63 give typemin(Int64) on x86-64, hence lo.
// Branch-free: NaN and t >= 2^63 go to -1e9 (hence lo), then the value is
// clamped to [-1e9, 1e9] so the conversion stays in range; for every input
// the clamped result equals the case analysis above.
__device__ __forceinline__ int jl_cell(double t, int lo, int hi) {
    const double t2 = t < 0x1p63 ? t : -1.0e9;
    const double tc = fmin(fmax(t2, -1.0e9), 1.0e9);
    const int c = (int)tc + 1;
    return min(max(c, lo), hi);
}

// Cell indices of a marker: basic (ib, jb) with the basic clamps and P
// (ip, jp) with the P clamps.  Vx cell = (ip, jb), Vy cell = (ib, jp).
struct Cells {
    int ib, jb, ip, jp;
};

__device__ __forceinline__ Cells marker_cells(const GeomK& g, double xm, double ym) {
    Cells c;
    c.jb = jl_cell((xm - g.x[0]) * g.idx, 1, g.Nx - 1);
    c.ib = jl_cell((ym - g.y[0]) * g.idy, 1, g.Ny - 1);
    c.jp = jl_cell((xm - g.xp[0]) * g.idx, 1, g.Nx);
    c.ip = jl_cell((ym - g.yp[0]) * g.idy, 1, g.Ny);
    return c;
}

// fix_weights (E:1388-1397) for a cell already fixed: returns the four
// weights [(1-b)(1-a), b(1-a), (1-b)a, ba] with a = dxmj/dx, b = dymi/dy.
__device__ __forceinline__ void weights4(double a, double b, double* w) {
    w[0] = (1.0 - b) * (1.0 - a);
    w[1] = b * (1.0 - a);
    w[2] = (1.0 - b) * a;
    w[3] = b * a;
}

// Address of grid element (i, j), 1-based, column-major.  Grid fields hold
// at most 2^24 elements (checked when the context is created), so the
// element index is computed in 32-bit arithmetic.
__device__ __forceinline__ const double* gptr_at(const double* __restrict__ base, int i, int j,
                                                 int rows) {
    return base + (uint32_t)((j - 1) * rows + (i - 1));
}

// Grid element at a byte offset below 2^31 (grid fields hold fewer than
// 2^24 elements): the 32-bit offset lets a load take the grid's base address
// from scalar registers and the offset from one vector register, with no
// 64-bit address arithmetic per element.
__device__ __forceinline__ double ldg_at(const double* __restrict__ base, uint32_t byte_off) {
    return *(const double*)((const char*)base + byte_off);
}

// byte offset of grid element (i, j), 1-based, column-major
__device__ __forceinline__ uint32_t goff(int i, int j, int rows) {
    return (uint32_t)((j - 1) * rows + (i - 1)) << 3;
}

// dot4(grid_vector(i, j, grid), w) (E:1137-1159), 1-based (i, j)
__device__ __forceinline__ double gdot4(const double* __restrict__ gr, int rows,
                                        int i, int j, const double* w) {
    const uint32_t o = goff(i, j, rows), oR = o + ((uint32_t)rows << 3);
    return ldg_at(gr, o) * w[0] + ldg_at(gr, o + 8) * w[1] + ldg_at(gr, oR) * w[2] +
           ldg_at(gr, oR + 8) * w[3];
}

// ---- marker properties: compute_marker_properties! (E:700-750) ----------

// per-type table entry for t in 1..3 without a dynamic index (a dynamic
// index into a by-value kernel argument forces a copy to scratch memory)
__device__ __forceinline__ double sel3(const double (&a)[3], int t) {
    return t == 1 ? a[0] : (t == 2 ? a[1] : a[2]);
}

__device__ __forceinline__ double jl_max(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}

// compute_rhocpfluidm (E:2272-2293); mode already validated (1 or 9)
__device__ __forceinline__ double rhocpfluidm(const ereb_materials& M, double T, int mode) {
    if (mode == 9) return M.rhocpfluidm[0];
    double tf = M.tmfluidphase;
    if (T < tf - 5.0) return M.rhoH2Ofi * (7.67 * T);
    if (T < tf) return M.rhoH2Ofi * (7.67 * T + 0.1 * M.Lf);
    if (T < tf + 5.0) return M.rhoH2Of * (4200.0 + 0.1 * M.Lf);
    if (T < 410.0) return M.rhoH2Of * 4200.0;
    return M.rhoH2Of * (-4.67e4 + 333.0 * T - 0.731 * (T * T) + 5.4e-4 * (T * T * T));
}

// compute_ksolidm (E:2311-2322)
__device__ __forceinline__ double ksolidm(const ereb_materials& M, double T, int mode) {
    return mode == 9 ? M.ksolidm[0] : 0.73 + 1293.0 / (T + 77.0);
}

// compute_kfluidm (E:2341-2358)
__device__ __forceinline__ double kfluidm(const ereb_materials& M, double T, int mode) {
    if (mode == 9) return M.kfluidm[0];
    if (T < M.tmfluidphase) return 0.465 + 488.0 / T;
    if (T < 410.0) return -0.581 + 6.34e-3 * T - 7.93e-6 * (T * T);
    return -0.142 + 4.12e-3 * T - 5.01e-6 * (T * T);
}

// total (E:1116)
__device__ __forceinline__ double total(double s, double f, double phi) {
    return s * (1.0 - phi) + f * phi;
}

// ktotal (E:1226-1234)
__device__ __forceinline__ double ktotal(double ks, double kf, double phi) {
    double t = ks * (3.0 * phi - 2.0) + kf * (1.0 - 3.0 * phi);
    return sqrt(ks * kf / 2.0 + (t * t) * (1.0 / 16.0)) - 0.25 * t;
}

// ηᶠcur_inv_kᵠ (E:1273-1275): ^3.0 and ^2.0 reduce to x*x*x and x*x in
// Julia's pow_body for integral exponents
__device__ __forceinline__ double etaf_inv_kphi(double phim0, double kphir, double phi,
                                                double etaf) {
    double q = phim0 * (1.0 / phi);
    double r = (1.0 - phi) * (1.0 / (1.0 - phim0));
    return etaf * (1.0 / kphir) * (q * q * q) * (r * r);
}

struct Props {
    double rhototal, rhocptotal, etatotal, hrtotal, ktotal, tkrhocp, etafkphi;
};

// One marker of compute_marker_properties!.  For air (t == 3) the stored
// rhototal, rhocptotal, etatotal, hrtotal, ktotal pass through (E:738-741).
__device__ __forceinline__ Props marker_props(const ereb_materials& M, int t, double T,
                                              double phi, double XW, const double (&hrs)[3],
                                              const double (&hrf)[3], int mode, Props stored) {
    Props p = stored;
    double etafluidcur;
    if (t < 3) {
        double XD = 1.0 - XW;
        double rhosolidm0 = (M.MD + M.MH2O * XW) / (M.VDs * XD + M.VWs * XW);
        double rhofluidm0 = T > M.tmfluidphase ? M.rhoH2Of : M.rhoH2Ofi;
        p.rhototal = total(rhosolidm0, rhofluidm0, phi);
        p.rhocptotal = total(sel3(M.rhocpsolidm, t), rhocpfluidm(M, T, mode), phi);
        double es = T > M.tmsolidphase ? sel3(M.etasolidmm, t) : sel3(M.etasolidm, t);
        etafluidcur = T > M.tmfluidphase ? sel3(M.etafluidmm, t) : sel3(M.etafluidm, t);
        p.etatotal = jl_max(jl_max(M.etamin, es), etafluidcur);
        p.hrtotal = total(sel3(hrs, t), sel3(hrf, t), phi);
        p.ktotal = ktotal(ksolidm(M, T, mode), kfluidm(M, T, mode), phi);
    } else {
        etafluidcur = M.etafluidm[2];
    }
    p.tkrhocp = T * p.rhocptotal;
    p.etafkphi = etaf_inv_kphi(M.phim0, sel3(M.kphim0, t), phi, etafluidcur);
    return p;
}

// ---- derivatives of the marker properties (adjoint kernels) -------------
// d/dT of the piecewise fluid heat capacity, conductivities (mode 1; mode 9
// uses constants).  The property switches at the phase temperatures are
// steps: their derivative is zero away from the switch.

__device__ __forceinline__ double d_rhocpfluidm(const ereb_materials& M, double T, int mode) {
    if (mode == 9) return 0.0;
    const double tf = M.tmfluidphase;
    if (T < tf) return M.rhoH2Ofi * 7.67;
    if (T < 410.0) return 0.0;
    return M.rhoH2Of * (333.0 - 2.0 * 0.731 * T + 3.0 * 5.4e-4 * (T * T));
}

__device__ __forceinline__ double d_ksolidm(double T, int mode) {
    if (mode == 9) return 0.0;
    const double u = T + 77.0;
    return -1293.0 / (u * u);
}

__device__ __forceinline__ double d_kfluidm(const ereb_materials& M, double T, int mode) {
    if (mode == 9) return 0.0;
    if (T < M.tmfluidphase) return -488.0 / (T * T);
    if (T < 410.0) return 6.34e-3 - 2.0 * 7.93e-6 * T;
    return 4.12e-3 - 2.0 * 5.01e-6 * T;
}

// partial derivatives of the properties that depend on T or phi
struct PropsD {
    double rho_p;                  // d rhototal / d phi (d/dT is zero: step in T)
    double rho_xw;                 // d rhototal / d XWsolidm0
    double rhocp_t, rhocp_p;       // rhocptotal
    double hr_p;                   // hrtotal
    double k_t, k_p;               // ktotal
    double tk_t, tk_p;             // tkm_rhocptotalm = T * rhocptotal
    double ef_p;                   // etafluidcur_inv_kphim
};

__device__ __forceinline__ PropsD marker_props_d(const ereb_materials& M, int t, double T,
                                                 double phi, double XW, const double (&hrs)[3],
                                                 const double (&hrf)[3], int mode,
                                                 const Props& p) {
    PropsD d{};
    if (t < 3) {
        const double XD = 1.0 - XW;
        const double rhosolidm0 = (M.MD + M.MH2O * XW) / (M.VDs * XD + M.VWs * XW);
        const double rhofluidm0 = T > M.tmfluidphase ? M.rhoH2Of : M.rhoH2Ofi;
        d.rho_p = rhofluidm0 - rhosolidm0;
        // rhosolidm0 = (MD + MH2O XW) / (VDs (1 - XW) + VWs XW)
        const double den = M.VDs * XD + M.VWs * XW;
        d.rho_xw = (1.0 - phi) * (M.MH2O * den - (M.MD + M.MH2O * XW) * (M.VWs - M.VDs)) /
                   (den * den);
        const double rcs = sel3(M.rhocpsolidm, t), rcf = rhocpfluidm(M, T, mode);
        d.rhocp_t = phi * d_rhocpfluidm(M, T, mode);
        d.rhocp_p = rcf - rcs;
        d.hr_p = sel3(hrf, t) - sel3(hrs, t);
        // ktotal = sqrt(A) - t/4, t = ks(3phi-2) + kf(1-3phi), A = ks kf/2 + t^2/16
        const double ks = ksolidm(M, T, mode), kf = kfluidm(M, T, mode);
        const double tt = ks * (3.0 * phi - 2.0) + kf * (1.0 - 3.0 * phi);
        const double sa = sqrt(ks * kf / 2.0 + (tt * tt) * (1.0 / 16.0));
        const double dks = (kf * 0.5 + tt * (3.0 * phi - 2.0) * 0.125) / (2.0 * sa) -
                           0.25 * (3.0 * phi - 2.0);
        const double dkf = (ks * 0.5 + tt * (1.0 - 3.0 * phi) * 0.125) / (2.0 * sa) -
                           0.25 * (1.0 - 3.0 * phi);
        const double dtp = 3.0 * (ks - kf);
        d.k_p = tt * dtp * 0.125 / (2.0 * sa) - 0.25 * dtp;
        d.k_t = dks * d_ksolidm(T, mode) + dkf * d_kfluidm(M, T, mode);
    }
    // tkm_rhocptotalm = T * rhocptotal (air keeps its stored rhocptotal)
    d.tk_t = p.rhocptotal + T * d.rhocp_t;
    d.tk_p = T * d.rhocp_p;
    // etafluidcur_inv_kphim = etaf / kphir * (phim0/phi)^3 * ((1-phi)/(1-phim0))^2
    d.ef_p = p.etafkphi * (-3.0 / phi - 2.0 / (1.0 - phi));
    return d;
}

// ---- FP32 property evaluation of the mixed-precision form ----------------
// The same formulas as the Float64 functions above, evaluated in single
// precision (inputs rounded to float once, Float64 table entries converted on
// use).  Only the mixed form calls these.

__device__ __forceinline__ float self3(const double (&a)[3], int t) {
    return (float)(t == 1 ? a[0] : (t == 2 ? a[1] : a[2]));
}

// FP32 forms of the temperature-dependent laws.  The branch conditions are
// evaluated on the Float64 temperature against the Float64 breakpoints, as in
// the Float64 form: rounding T to float first would move markers within half
// a float spacing (1.5e-5 K at 273 K) of a breakpoint to the other branch.
__device__ __forceinline__ float rhocpfluidm_f(const ereb_materials& M, double Td, int mode) {
    if (mode == 9) return (float)M.rhocpfluidm[0];
    const double tf = M.tmfluidphase;
    const float T = (float)Td, Lf = (float)M.Lf;
    const float rw = (float)M.rhoH2Of, ri = (float)M.rhoH2Ofi;
    if (Td < tf - 5.0) return ri * (7.67f * T);
    if (Td < tf) return ri * (7.67f * T + 0.1f * Lf);
    if (Td < tf + 5.0) return rw * (4200.0f + 0.1f * Lf);
    if (Td < 410.0) return rw * 4200.0f;
    return rw * (-4.67e4f + 333.0f * T - 0.731f * (T * T) + 5.4e-4f * (T * T * T));
}

__device__ __forceinline__ float ksolidm_f(const ereb_materials& M, double Td, int mode) {
    return mode == 9 ? (float)M.ksolidm[0] : 0.73f + 1293.0f / ((float)Td + 77.0f);
}

__device__ __forceinline__ float kfluidm_f(const ereb_materials& M, double Td, int mode) {
    if (mode == 9) return (float)M.kfluidm[0];
    const float T = (float)Td;
    if (Td < M.tmfluidphase) return 0.465f + 488.0f / T;
    if (Td < 410.0) return -0.581f + 6.34e-3f * T - 7.93e-6f * (T * T);
    return -0.142f + 4.12e-3f * T - 5.01e-6f * (T * T);
}

__device__ __forceinline__ float total_f(float s, float f, float phi) {
    return s * (1.0f - phi) + f * phi;
}

__device__ __forceinline__ float ktotal_f(float ks, float kf, float phi) {
    const float t = ks * (3.0f * phi - 2.0f) + kf * (1.0f - 3.0f * phi);
    return sqrtf(ks * kf * 0.5f + (t * t) * (1.0f / 16.0f)) - 0.25f * t;
}

__device__ __forceinline__ float etaf_inv_kphi_f(float phim0, float kphir, float phi, float etaf) {
    const float q = phim0 / phi;
    const float r = (1.0f - phi) / (1.0f - phim0);
    return etaf / kphir * (q * q * q) * (r * r);
}

struct PropsF {
    float rhototal, rhocptotal, etatotal, hrtotal, ktotal, tkrhocp, etafkphi;
};

__device__ __forceinline__ PropsF marker_props_f(const ereb_materials& M, int t, double Td,
                                                 float phi, float XW, const double (&hrs)[3],
                                                 const double (&hrf)[3], int mode,
                                                 PropsF stored) {
    const float T = (float)Td;
    PropsF p = stored;
    float etafluidcur;
    if (t < 3) {
        const float XD = 1.0f - XW;
        const float rhosolidm0 = ((float)M.MD + (float)M.MH2O * XW) /
                                 ((float)M.VDs * XD + (float)M.VWs * XW);
        const float rhofluidm0 = Td > M.tmfluidphase ? (float)M.rhoH2Of : (float)M.rhoH2Ofi;
        p.rhototal = total_f(rhosolidm0, rhofluidm0, phi);
        p.rhocptotal = total_f(self3(M.rhocpsolidm, t), rhocpfluidm_f(M, Td, mode), phi);
        const float es = Td > M.tmsolidphase ? self3(M.etasolidmm, t) : self3(M.etasolidm, t);
        etafluidcur = Td > M.tmfluidphase ? self3(M.etafluidmm, t) : self3(M.etafluidm, t);
        p.etatotal = fmaxf(fmaxf((float)M.etamin, es), etafluidcur);
        p.hrtotal = total_f(self3(hrs, t), self3(hrf, t), phi);
        p.ktotal = ktotal_f(ksolidm_f(M, Td, mode), kfluidm_f(M, Td, mode), phi);
    } else {
        etafluidcur = (float)M.etafluidm[2];
    }
    p.tkrhocp = T * p.rhocptotal;
    p.etafkphi = etaf_inv_kphi_f((float)M.phim0, self3(M.kphim0, t), phi, etafluidcur);
    return p;
}

// Per-step scalars read by the kernels from device memory rather than taken
// as launch arguments, so that a captured step graph stays valid while they
// change from step to step: dt (E:6619, E:6790, E:6933), the radiogenic
// heating rates of calculate_radioactive_heating(timesum) (E:6413-6414,
// E:1336-1357) and whether the temperature gather is the first timestep's.
struct StepScalars {
    double dt;
    double hrs[3], hrf[3];
    int32_t first;
    int32_t pad;
};

// compute_marker_properties! (E:700-750) for marker m in storage order:
// reads tm, tkm, phim, XWsolidm0 (rocks) or the stored rhototal, rhocptotal,
// etatotal, hrtotal, ktotal (air, which keeps them) and writes the property
// vectors.  R is the type of the property vectors: double in the Float64
// form, float in the mixed form.
template <class R>
struct PropArgs {
    const ereb_materials* __restrict__ M;
    const StepScalars* __restrict__ S;  // hrs, hrf
    // Sw set: the step's scalars come with the launch (Sv) instead of from S,
    // and the pass writes them to Sw for the kernels after it
    StepScalars Sv;
    StepScalars* Sw;
    int mode;
    const int32_t* __restrict__ tm;
    const double *__restrict__ tkm, *__restrict__ phim;
    const R* __restrict__ XW;
    R *rhototal, *rhocptotal, *etatotal, *hrtotal, *ktotal, *tkrhocp, *etafkphi;
    __device__ void apply(int64_t m) const;
    __device__ void publish() const {
        if (Sw) *Sw = Sv;
    }
    // the heating rates: values selected, not pointers (a pointer to either
    // the launch argument or the device block would make every load a
    // per-lane generic one)
    __device__ __forceinline__ void heating(double (&hrs)[3], double (&hrf)[3]) const {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            hrs[q] = Sw ? Sv.hrs[q] : S->hrs[q];
            hrf[q] = Sw ? Sv.hrf[q] : S->hrf[q];
        }
    }
};
template <>
__device__ __forceinline__ void PropArgs<double>::apply(int64_t m) const {
    const int t = min(max(tm[m], 1), 3);  // set validates 1..3
    Props st{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double xw = 0.0;
    if (t == 3)
        st = Props{rhototal[m], rhocptotal[m], etatotal[m], hrtotal[m], ktotal[m], 0.0, 0.0};
    else
        xw = XW[m];
    double hrs[3], hrf[3];
    heating(hrs, hrf);
    const Props p = marker_props(*M, t, tkm[m], phim[m], xw, hrs, hrf, mode, st);
    if (t < 3) {
        rhototal[m] = p.rhototal;
        rhocptotal[m] = p.rhocptotal;
        etatotal[m] = p.etatotal;
        hrtotal[m] = p.hrtotal;
        ktotal[m] = p.ktotal;
    }
    tkrhocp[m] = p.tkrhocp;
    etafkphi[m] = p.etafkphi;
}
template <>
__device__ __forceinline__ void PropArgs<float>::apply(int64_t m) const {
    const int t = min(max(tm[m], 1), 3);
    PropsF st{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float xw = 0.0f;
    if (t == 3)
        st = PropsF{rhototal[m], rhocptotal[m], etatotal[m], hrtotal[m], ktotal[m], 0.0f, 0.0f};
    else
        xw = XW[m];
    double hrs[3], hrf[3];
    heating(hrs, hrf);
    const PropsF p = marker_props_f(*M, t, tkm[m], (float)phim[m], xw, hrs, hrf, mode, st);
    if (t < 3) {
        rhototal[m] = p.rhototal;
        rhocptotal[m] = p.rhocptotal;
        etatotal[m] = p.etatotal;
        hrtotal[m] = p.hrtotal;
        ktotal[m] = p.ktotal;
    }
    tkrhocp[m] = p.tkrhocp;
    etafkphi[m] = p.etafkphi;
}
// no property evaluation
struct NoProps {
    __device__ void apply(int64_t) const {}
    __device__ void publish() const {}
};

// Compensated FP32 accumulation (Knuth's TwoSum): s + c carries the sum of
// the added values with an error of O(u^2) relative to their magnitudes.
__device__ __forceinline__ void two_sum_add(float& s, float& c, float x) {
    const float t = s + x;
    const float bp = t - s;
    const float e = (s - (t - bp)) + (x - bp);
    s = t;
    c += e;
}

// ---- wavefront helpers (64 lanes) ---------------------------------------

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// bitonic sort (ascending) of s[0..n), n a power of two, by the block
__device__ __forceinline__ void bitonic(uint32_t* s, int n) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t a = s[i], b = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace ereb
