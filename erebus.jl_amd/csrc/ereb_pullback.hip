// ereb_pullback.hip -- reverse-mode pullbacks of the transfer functions, the
// kernels behind ChainRulesCore rrules of:
//   marker_to_*_nodes! with compute_marker_properties! (E:700-750,
//       E:1605-1863, loop E:6419-6517): its transpose is a gather, each marker
//       reads the SUM cotangents at its 16 stencil nodes (no reduction);
//   update_marker_temperature! (E:4782-4801) and update_marker_stress!
//       (E:4354-4388): marker cotangents go to the grid nodes of the marker's
//       stencil (a scatter) and, through the interpolated field's gradient,
//       to the marker position;
//   move_markers_rk4! (E:4990-5257): per marker, the reverse sweep through
//       the stress rotation, the fluid back-trace, the solid RK4 stages and
//       the temperature correction, recomputing the stage positions; the
//       cotangents of the velocity, rotation-rate and tk2 grids are scattered
//       from every stage position.
// Grid cotangents are summed in fixed point (ereb_fixsum.h), so a pullback
// gives the same bits on every run.  Float64 form only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ereb_fixsum.h"
#include "ereb_internal.h"

namespace ereb {

// ---- stencils with gradients --------------------------------------------------

// cell (i, j) of a position on a grid with axes (xa, ya) and clamps, and the
// fractions a = (x - xa[j]) / dx, b = (y - ya[i]) / dy (fix_weights, E:1388)
struct Stn {
    int i, j;
    double a, b;
};

__device__ __forceinline__ Stn stn(const GeomK& g, const double* xa, const double* ya, double x,
                                   double y, int jlo, int jhi, int ilo, int ihi) {
    Stn s;
    s.j = jl_cell((x - xa[0]) * g.idx, jlo, jhi);
    s.i = jl_cell((y - ya[0]) * g.idy, ilo, ihi);
    s.a = (x - xa[s.j - 1]) / g.dx;
    s.b = (y - ya[s.i - 1]) / g.dy;
    return s;
}

__device__ __forceinline__ int64_t node(int i, int j, int rows) {
    return (int64_t)(j - 1) * rows + (i - 1);
}

// bilinear value of grid gr at the stencil and its gradient (d/dx, d/dy)
__device__ __forceinline__ double interp_grad(const GeomK& g, const double* gr, int rows,
                                              const Stn& s, double& gx, double& gy) {
    const double* p = gr + node(s.i, s.j, rows);
    const double g00 = p[0], g10 = p[1], g01 = p[rows], g11 = p[rows + 1];
    gx = ((g01 - g00) * (1.0 - s.b) + (g11 - g10) * s.b) / g.dx;
    gy = ((g10 - g00) * (1.0 - s.a) + (g11 - g01) * s.a) / g.dy;
    return g00 * ((1.0 - s.b) * (1.0 - s.a)) + g10 * (s.b * (1.0 - s.a)) +
           g01 * ((1.0 - s.b) * s.a) + g11 * (s.b * s.a);
}

// grid cotangent of a bilinear read: v times the four weights
template <int PASS>
__device__ __forceinline__ void spread4(const FixGrid& F, int rows, const Stn& s, double v,
                                        int* flag) {
    if (!F.w) return;
    const int64_t k = node(s.i, s.j, rows);
    fix_add<PASS>(F, k, v * ((1.0 - s.b) * (1.0 - s.a)), flag);
    fix_add<PASS>(F, k + 1, v * (s.b * (1.0 - s.a)), flag);
    fix_add<PASS>(F, k + rows, v * ((1.0 - s.b) * s.a), flag);
    fix_add<PASS>(F, k + rows + 1, v * (s.b * s.a), flag);
}

// ---- RK4 stages with partials (E:5057-5127) -------------------------------------

// Vx stage: value and gradient at (x, y); with vbar != 0 the node cotangents
// vbar * dv/d(node) go to F (grid v) and Fc (the centre node of the left-half
// correction, read from vc: vx also in the fluid loop, E:5173)
template <int PASS>
__device__ double vx_stage(const GeomK& g, const double* v, const double* vc, double x, double y,
                           double& dvx, double& dvy, double vbar, const FixGrid& F,
                           const FixGrid& Fc, int* flag) {
    const int R = g.Ny1;
    const Stn s = stn(g, g.x, g.yp, x, y, 1, g.Nx - 1, 1, g.Ny);
    const int i = s.i, j = s.j;
    const double a = s.a, b = s.b;
    const double* p = v + node(i, j, R);
    const double A0 = p[0], A1 = p[1], B0 = p[R], B1 = p[R + 1];
    double v13 = A0 * (1.0 - a) + B0 * a, v24 = A1 * (1.0 - a) + B1 * a;
    double d13 = B0 - A0, d24 = B1 - A1;  // d/da
    const bool right = a >= 0.5;
    const bool apply = right ? j < g.Nx - 1 : j > 1;
    const double t = a - 0.5, h = 0.5 * (t * t);
    // columns of the correction's three points (q0, centre, q2)
    const int jq0 = right ? j : j - 1, jc = right ? j + 1 : j, jq2 = right ? j + 2 : j + 1;
    if (apply) {
        const double q0 = v[node(i, jq0, R)], q1 = v[node(i + 1, jq0, R)];
        const double c13 = right ? v[node(i, jc, R)] : vc[node(i, jc, R)];
        const double c24 = v[node(i + 1, jc, R)];
        const double q20 = v[node(i, jq2, R)], q21 = v[node(i + 1, jq2, R)];
        const double D13 = q0 - 2.0 * c13 + q20, D24 = q1 - 2.0 * c24 + q21;
        v13 += h * D13;
        v24 += h * D24;
        d13 += t * D13;
        d24 += t * D24;
    }
    dvx = ((1.0 - b) * d13 + b * d24) / g.dx;
    dvy = (v24 - v13) / g.dy;
    if (vbar != 0.0 && (F.w || Fc.w)) {  // fix_add drops the grids not wanted
        fix_add<PASS>(F, node(i, j, R), vbar * ((1.0 - b) * (1.0 - a)), flag);
        fix_add<PASS>(F, node(i, j + 1, R), vbar * ((1.0 - b) * a), flag);
        fix_add<PASS>(F, node(i + 1, j, R), vbar * (b * (1.0 - a)), flag);
        fix_add<PASS>(F, node(i + 1, j + 1, R), vbar * (b * a), flag);
        if (apply) {
            const double r0 = vbar * (1.0 - b) * h, r1 = vbar * b * h;
            fix_add<PASS>(F, node(i, jq0, R), r0, flag);
            fix_add<PASS>(F, node(i, jq2, R), r0, flag);
            fix_add<PASS>(right ? F : Fc, node(i, jc, R), -2.0 * r0, flag);
            fix_add<PASS>(F, node(i + 1, jq0, R), r1, flag);
            fix_add<PASS>(F, node(i + 1, jq2, R), r1, flag);
            fix_add<PASS>(F, node(i + 1, jc, R), -2.0 * r1, flag);
        }
    }
    return v13 * (1.0 - b) + v24 * b;
}

// Vy stage (E:5093-5127): the transpose of the Vx stencil, no centre quirk
template <int PASS>
__device__ double vy_stage(const GeomK& g, const double* v, double x, double y, double& dvx,
                           double& dvy, double vbar, const FixGrid& F, int* flag) {
    const int R = g.Ny1;
    const Stn s = stn(g, g.xp, g.y, x, y, 1, g.Nx, 1, g.Ny - 1);
    const int i = s.i, j = s.j;
    const double a = s.a, b = s.b;
    const double* p = v + node(i, j, R);
    const double A0 = p[0], A1 = p[1], B0 = p[R], B1 = p[R + 1];
    double v12 = A0 * (1.0 - b) + A1 * b, v34 = B0 * (1.0 - b) + B1 * b;
    double d12 = A1 - A0, d34 = B1 - B0;  // d/db
    const bool down = b >= 0.5;
    const bool apply = down ? i < g.Ny - 1 : i > 1;
    const double t = b - 0.5, h = 0.5 * (t * t);
    const int iq0 = down ? i : i - 1, ic = down ? i + 1 : i, iq2 = down ? i + 2 : i + 1;
    if (apply) {
        const double D12 = v[node(iq0, j, R)] - 2.0 * v[node(ic, j, R)] + v[node(iq2, j, R)];
        const double D34 =
            v[node(iq0, j + 1, R)] - 2.0 * v[node(ic, j + 1, R)] + v[node(iq2, j + 1, R)];
        v12 += h * D12;
        v34 += h * D34;
        d12 += t * D12;
        d34 += t * D34;
    }
    dvy = ((1.0 - a) * d12 + a * d34) / g.dy;
    dvx = (v34 - v12) / g.dx;
    if (vbar != 0.0 && F.w) {
        fix_add<PASS>(F, node(i, j, R), vbar * ((1.0 - a) * (1.0 - b)), flag);
        fix_add<PASS>(F, node(i + 1, j, R), vbar * ((1.0 - a) * b), flag);
        fix_add<PASS>(F, node(i, j + 1, R), vbar * (a * (1.0 - b)), flag);
        fix_add<PASS>(F, node(i + 1, j + 1, R), vbar * (a * b), flag);
        if (apply) {
            const double r0 = vbar * (1.0 - a) * h, r1 = vbar * a * h;
            fix_add<PASS>(F, node(iq0, j, R), r0, flag);
            fix_add<PASS>(F, node(iq2, j, R), r0, flag);
            fix_add<PASS>(F, node(ic, j, R), -2.0 * r0, flag);
            fix_add<PASS>(F, node(iq0, j + 1, R), r1, flag);
            fix_add<PASS>(F, node(iq2, j + 1, R), r1, flag);
            fix_add<PASS>(F, node(ic, j + 1, R), -2.0 * r1, flag);
        }
    }
    return v12 * (1.0 - a) + v34 * a;
}

// ---- RK4 pullback ------------------------------------------------------------------

struct RK4PB {
    GeomK g;
    const ereb_materials* __restrict__ M;
    double dt;
    int mode;
    int64_t n;
    const double *xm, *ym, *tkm, *phim, *sxym, *sxxm;
    const int32_t* tm;
    const double *vx, *vy, *vxf, *vyf, *wyx, *tk2;
    // cotangents of the outputs (nullptr = 0) and of the inputs (nullptr = skip)
    const double *bx, *by, *bT, *bsxy, *bsxx;
    double *ox, *oy, *oT, *ophi, *osxy, *osxx;
    FixGrid Fvx, Fvy, Fvxf, Fvyf, Fwyx, Ftk2;
    int* flag;
};

template <int PASS>
__global__ __launch_bounds__(256) void k_rk4_pullback(RK4PB A) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= A.n) return;
    const GeomK& g = A.g;
    const ereb_materials& M = *A.M;
    const double dt = A.dt;
    const double* bk = M.brk4;
    const double* ck = M.crk4;
    auto bar = [&](const double* b) { return b ? b[m] : 0.0; };
    const double x0 = A.xm[m], y0 = A.ym[m];
    const FixGrid none{nullptr, nullptr, 1.0};
    double gx, gy;
    // forward: the solid stage positions and the fluid back-trace positions
    double X[4], Y[4], Xf[4], Yf[4], u[4], w[4];
    X[0] = x0;
    Y[0] = y0;
    for (int k = 0; k < 4; ++k) {
        u[k] = vx_stage<PASS>(g, A.vx, A.vx, X[k], Y[k], gx, gy, 0.0, none, none, A.flag);
        w[k] = vy_stage<PASS>(g, A.vy, X[k], Y[k], gx, gy, 0.0, none, A.flag);
        if (k < 3) {
            X[k + 1] = x0 + dt * ck[k] * u[k];
            Y[k + 1] = y0 + dt * ck[k] * w[k];
        }
    }
    const double xn = x0 + dt * (bk[0] * u[0] + bk[1] * u[1] + bk[2] * u[2] + bk[3] * u[3]);
    const double yn = y0 + dt * (bk[0] * w[0] + bk[1] * w[1] + bk[2] * w[2] + bk[3] * w[3]);
    Xf[0] = xn;
    Yf[0] = yn;
    for (int k = 0; k < 4; ++k) {
        u[k] = vx_stage<PASS>(g, A.vxf, A.vx, Xf[k], Yf[k], gx, gy, 0.0, none, none, A.flag);
        w[k] = vy_stage<PASS>(g, A.vyf, Xf[k], Yf[k], gx, gy, 0.0, none, A.flag);
        if (k < 3) {
            Xf[k + 1] = xn - dt * ck[k] * u[k];
            Yf[k + 1] = yn - dt * ck[k] * w[k];
        }
    }
    const double xb = xn - dt * (bk[0] * u[0] + bk[1] * u[1] + bk[2] * u[2] + bk[3] * u[3]);
    const double yb = yn - dt * (bk[0] * w[0] + bk[1] * w[1] + bk[2] * w[2] + bk[3] * w[3]);

    // temperature correction (E:5243-5252)
    const Stn sp0 = stn(g, g.xp, g.yp, x0, y0, 1, g.Nx, 1, g.Ny);
    const Stn spb = stn(g, g.xp, g.yp, xb, yb, 1, g.Nx, 1, g.Ny);
    double g0x, g0y, gbx, gby;
    const double tks0 = interp_grad(g, A.tk2, g.Ny1, sp0, g0x, g0y);
    const double tkf0 = interp_grad(g, A.tk2, g.Ny1, spb, gbx, gby);
    const double d = tkf0 - tks0;
    const int tt = min(max(A.tm[m], 1), 3);
    const double T = A.tkm[m], ph = A.phim[m];
    const double rcs = sel3(M.rhocpsolidm, tt);
    const double rcf = rhocpfluidm(M, T, A.mode), drcf = d_rhocpfluidm(M, T, A.mode);
    const double num = (1.0 - ph) * T * rcs + ph * (T + d) * rcf;
    const double den = (1.0 - ph) * rcs + ph * rcf;
    const double F_d = ph * rcf / den;
    const double F_T = ((1.0 - ph) * rcs + ph * rcf + ph * (T + d) * drcf) / den -
                       num * (ph * drcf) / (den * den);
    const double F_p = (-T * rcs + (T + d) * rcf) / den - num * (rcf - rcs) / (den * den);
    const double Tb = bar(A.bT);
    const double dbar = Tb * F_d;
    spread4<PASS>(A.Ftk2, g.Ny1, spb, dbar, A.flag);
    spread4<PASS>(A.Ftk2, g.Ny1, sp0, -dbar, A.flag);
    double xbb = dbar * gbx, ybb = dbar * gby;  // cotangent of the back-trace end point
    double x0b = -dbar * g0x, y0b = -dbar * g0y;

    // stress rotation (E:5042-5048): sxx' = sxx cos 2th - sxy sin 2th,
    // sxy' = sxx sin 2th + sxy cos 2th, th = dt * <wyx>
    const Stn sb0 = stn(g, g.x, g.y, x0, y0, 1, g.Nx - 1, 1, g.Ny - 1);
    double wx, wy;
    const double om = interp_grad(g, A.wyx, g.Ny, sb0, wx, wy);
    const double th = dt * om;
    double s1, c1;
    sincos(th, &s1, &c1);
    const double s2 = 2.0 * (s1 * c1), c2 = (c1 - s1) * (c1 + s1);
    const double sxx = A.sxxm[m], sxy = A.sxym[m];
    const double bsxx = bar(A.bsxx), bsxy = bar(A.bsxy);
    const double thb = bsxx * (-2.0 * sxx * s2 - 2.0 * sxy * c2) +
                       bsxy * (2.0 * sxx * c2 - 2.0 * sxy * s2);
    const double omb = dt * thb;
    spread4<PASS>(A.Fwyx, g.Ny, sb0, omb, A.flag);
    x0b += omb * wx;
    y0b += omb * wy;

    // fluid back-trace reverse: xb = xn - dt sum b_k u_k(Xf_k)
    double xnb = bar(A.bx) + xbb, ynb = bar(A.by) + ybb;
    double ub[4], wb[4];
    for (int k = 0; k < 4; ++k) {
        ub[k] = -dt * bk[k] * xbb;
        wb[k] = -dt * bk[k] * ybb;
    }
    for (int k = 3; k >= 0; --k) {
        double ux, uy, vxg, vyg;
        vx_stage<PASS>(g, A.vxf, A.vx, Xf[k], Yf[k], ux, uy, ub[k], A.Fvxf, A.Fvx, A.flag);
        vy_stage<PASS>(g, A.vyf, Xf[k], Yf[k], vxg, vyg, wb[k], A.Fvyf, A.flag);
        const double Xkb = ux * ub[k] + vxg * wb[k], Ykb = uy * ub[k] + vyg * wb[k];
        // Xf_k = xn - dt c_{k-1} u_{k-1} (k >= 1), Xf_0 = xn
        xnb += Xkb;
        ynb += Ykb;
        if (k > 0) {
            ub[k - 1] += -dt * ck[k - 1] * Xkb;
            wb[k - 1] += -dt * ck[k - 1] * Ykb;
        }
    }
    // solid reverse: xn = x0 + dt sum b_k u_k(X_k)
    x0b += xnb;
    y0b += ynb;
    for (int k = 0; k < 4; ++k) {
        ub[k] = dt * bk[k] * xnb;
        wb[k] = dt * bk[k] * ynb;
    }
    for (int k = 3; k >= 0; --k) {
        double ux, uy, vxg, vyg;
        vx_stage<PASS>(g, A.vx, A.vx, X[k], Y[k], ux, uy, ub[k], A.Fvx, A.Fvx, A.flag);
        vy_stage<PASS>(g, A.vy, X[k], Y[k], vxg, vyg, wb[k], A.Fvy, A.flag);
        const double Xkb = ux * ub[k] + vxg * wb[k], Ykb = uy * ub[k] + vyg * wb[k];
        x0b += Xkb;
        y0b += Ykb;
        if (k > 0) {
            ub[k - 1] += dt * ck[k - 1] * Xkb;
            wb[k - 1] += dt * ck[k - 1] * Ykb;
        }
    }
    if (PASS == 1) {
        if (A.ox) A.ox[m] = x0b;
        if (A.oy) A.oy[m] = y0b;
        if (A.oT) A.oT[m] = Tb * F_T;
        if (A.ophi) A.ophi[m] = Tb * F_p;
        if (A.osxx) A.osxx[m] = bsxx * c2 + bsxy * s2;
        if (A.osxy) A.osxy[m] = -bsxx * s2 + bsxy * c2;
    }
}

// ---- gather pullbacks --------------------------------------------------------------

// update_marker_temperature! (E:4782-4801): tkm' = <G> (timestep 1) or
// tkm + <G>, G = tk2 or DT on the P grid
template <int PASS>
__global__ __launch_bounds__(256) void k_gather_T_pullback(GeomK g, int64_t n,
                                                           const double* __restrict__ xm,
                                                           const double* __restrict__ ym,
                                                           const double* __restrict__ grid,
                                                           int first, const double* __restrict__ tb,
                                                           FixGrid F, int* flag,
                                                           double* __restrict__ otk,
                                                           double* __restrict__ ox,
                                                           double* __restrict__ oy) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double x = xm[m], y = ym[m], b = tb[m];
    const Stn s = stn(g, g.xp, g.yp, x, y, 1, g.Nx, 1, g.Ny);
    spread4<PASS>(F, g.Ny1, s, b, flag);
    if (PASS == 1) {
        double gx, gy;
        interp_grad(g, grid, g.Ny1, s, gx, gy);
        if (ox) ox[m] = b * gx;
        if (oy) oy[m] = b * gy;
        if (otk) otk[m] = first ? 0.0 : b;
    }
}

// update_marker_stress! (E:4354-4388): sxx' = sxx + <DSXX> on the P grid with
// the clamps 2..Nx-1, 2..Ny-1 (E:4364-4367), sxy' = sxy + <DSXY> (basic)
template <int PASS>
__global__ __launch_bounds__(256) void k_gather_S_pullback(GeomK g, int64_t n,
                                                           const double* __restrict__ xm,
                                                           const double* __restrict__ ym,
                                                           const double* __restrict__ DSXX,
                                                           const double* __restrict__ DSXY,
                                                           const double* __restrict__ bxx,
                                                           const double* __restrict__ bxy,
                                                           FixGrid Fxx, FixGrid Fxy, int* flag,
                                                           double* __restrict__ ox,
                                                           double* __restrict__ oy) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double x = xm[m], y = ym[m];
    const double a = bxx ? bxx[m] : 0.0, b = bxy ? bxy[m] : 0.0;
    const Stn sp = stn(g, g.xp, g.yp, x, y, 2, g.Nx - 1, 2, g.Ny - 1);
    const Stn sb = stn(g, g.x, g.y, x, y, 1, g.Nx - 1, 1, g.Ny - 1);
    spread4<PASS>(Fxx, g.Ny1, sp, a, flag);
    spread4<PASS>(Fxy, g.Ny, sb, b, flag);
    if (PASS == 1) {
        double px, py, qx, qy;
        interp_grad(g, DSXX, g.Ny1, sp, px, py);
        interp_grad(g, DSXY, g.Ny, sb, qx, qy);
        if (ox) ox[m] = a * px + b * qx;
        if (oy) oy[m] = a * py + b * qy;
    }
}

// ---- scatter pullback ----------------------------------------------------------------

struct ScatterPB {
    GeomK g;
    const ereb_materials* __restrict__ M;
    const StepScalars* __restrict__ S;  // hrs, hrf
    int mode;
    int64_t n;
    const double* f[EREB_M_NFIELDS];  // marker fields (Float64 form)
    const int32_t* tm;
    const double* sb[NSUM];  // SUM cotangents (nullptr = 0)
    const double* pb[7];     // cotangents of the 7 property outputs (nullptr = 0)
    double* out[EREB_M_NFIELDS];
};

// cotangent-weighted sums of slot s at a stencil: value (sum of bar * w) and
// the derivatives with respect to a and b
struct Gath {
    double v, da, db;
};
__device__ __forceinline__ Gath gath(const double* bar, int rows, const Stn& s) {
    if (!bar) return Gath{0.0, 0.0, 0.0};
    const double* p = bar + node(s.i, s.j, rows);
    const double g00 = p[0], g10 = p[1], g01 = p[rows], g11 = p[rows + 1];
    Gath r;
    r.v = g00 * ((1.0 - s.b) * (1.0 - s.a)) + g10 * (s.b * (1.0 - s.a)) +
          g01 * ((1.0 - s.b) * s.a) + g11 * (s.b * s.a);
    r.da = (g01 - g00) * (1.0 - s.b) + (g11 - g10) * s.b;
    r.db = (g10 - g00) * (1.0 - s.a) + (g11 - g01) * s.a;
    return r;
}

__global__ __launch_bounds__(256) void k_scatter_pullback(ScatterPB A) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= A.n) return;
    const GeomK& g = A.g;
    const ereb_materials& M = *A.M;
    auto F = [&](int f) { return A.f[f] ? A.f[f][m] : 0.0; };
    const double x = F(EREB_M_XM), y = F(EREB_M_YM), T = F(EREB_M_TKM), phi = F(EREB_M_PHIM);
    const double XW = F(EREB_M_XWSOLIDM0);
    const int tt = min(max(A.tm[m], 1), 3);
    const double hrs[3] = {A.S->hrs[0], A.S->hrs[1], A.S->hrs[2]};
    const double hrf[3] = {A.S->hrf[0], A.S->hrf[1], A.S->hrf[2]};
    const Props st{F(EREB_M_RHOTOTALM), F(EREB_M_RHOCPTOTALM), F(EREB_M_ETATOTALM),
                   F(EREB_M_HRTOTALM), F(EREB_M_KTOTALM), 0.0, 0.0};
    const Props p = marker_props(M, tt, T, phi, XW, hrs, hrf, A.mode, st);
    const PropsD d = marker_props_d(M, tt, T, phi, XW, hrs, hrf, A.mode, p);
    // stencils of the four grids (E:1605-1863)
    const Stn sB = stn(g, g.x, g.y, x, y, 1, g.Nx - 1, 1, g.Ny - 1);
    const Stn sVx = stn(g, g.x, g.yp, x, y, 1, g.Nx - 1, 1, g.Ny);
    const Stn sVy = stn(g, g.xp, g.y, x, y, 1, g.Nx, 1, g.Ny - 1);
    const Stn sP = stn(g, g.xp, g.yp, x, y, 1, g.Nx, 1, g.Ny);
    // staged value of each SUM slot (reset_interpolated_properties! order)
    const double val[NSUM] = {
        p.etatotal, F(EREB_M_ETAVPM), F(EREB_M_INV_GGGTOTALM), F(EREB_M_SXYM),
        F(EREB_M_COHESTOTALM), F(EREB_M_TENSTOTALM), F(EREB_M_FRICTTOTALM), 1.0,
        p.rhototal, F(EREB_M_RHOFLUIDCUR), p.ktotal, phi, p.etafkphi, 1.0,
        p.rhototal, F(EREB_M_RHOFLUIDCUR), p.ktotal, phi, p.etafkphi, 1.0,
        p.rhototal, p.rhocptotal, F(EREB_M_ALPHASOLIDCUR), F(EREB_M_ALPHAFLUIDCUR), p.hrtotal,
        F(EREB_M_INV_GGGTOTALM), F(EREB_M_SXXM), p.tkrhocp, phi, 1.0};
    double vb[NSUM];  // cotangent of each staged value
    double xb = 0.0, yb = 0.0;
#pragma unroll
    for (int s = 0; s < NSUM; ++s) {
        const Stn& st_ = s < 8 ? sB : s < 14 ? sVx : s < 20 ? sVy : sP;
        const int rows = s < 8 ? g.Ny : g.Ny1;
        const Gath r = gath(A.sb[s], rows, st_);
        vb[s] = r.v;
        xb += val[s] * r.da;
        yb += val[s] * r.db;
    }
    xb /= g.dx;
    yb /= g.dy;
    auto PB = [&](int k) { return A.pb[k] ? A.pb[k][m] : 0.0; };
    const double b_rho = vb[8] + vb[14] + vb[20] + PB(0);
    const double b_rhocp = vb[21] + PB(1);
    const double b_eta = vb[0] + PB(2);
    const double b_hr = vb[24] + PB(3);
    const double b_k = vb[10] + vb[16] + PB(4);
    const double b_tk = vb[27] + PB(5);
    const double b_ef = vb[12] + vb[18] + PB(6);
    // chain through compute_marker_properties! (rocks) or the stored values
    // that air keeps (E:738-741)
    const double Tb = b_rhocp * d.rhocp_t + b_k * d.k_t + b_tk * d.tk_t;
    const double phb = b_rho * d.rho_p + b_rhocp * d.rhocp_p + b_hr * d.hr_p + b_k * d.k_p +
                       b_tk * d.tk_p + b_ef * d.ef_p + vb[11] + vb[17] + vb[28];
    double o[EREB_M_NFIELDS] = {};
    o[EREB_M_XM] = xb;
    o[EREB_M_YM] = yb;
    o[EREB_M_TKM] = Tb;
    o[EREB_M_PHIM] = phb;
    o[EREB_M_XWSOLIDM0] = tt < 3 ? b_rho * d.rho_xw : 0.0;
    o[EREB_M_ETAVPM] = vb[1];
    o[EREB_M_INV_GGGTOTALM] = vb[2] + vb[25];
    o[EREB_M_SXYM] = vb[3];
    o[EREB_M_COHESTOTALM] = vb[4];
    o[EREB_M_TENSTOTALM] = vb[5];
    o[EREB_M_FRICTTOTALM] = vb[6];
    o[EREB_M_RHOFLUIDCUR] = vb[9] + vb[15];
    o[EREB_M_ALPHASOLIDCUR] = vb[22];
    o[EREB_M_ALPHAFLUIDCUR] = vb[23];
    o[EREB_M_SXXM] = vb[26];
    if (tt == 3) {  // the stored properties pass through (tkm_rhocp = T * rhocp)
        o[EREB_M_RHOTOTALM] = b_rho;
        o[EREB_M_RHOCPTOTALM] = b_rhocp + b_tk * T;
        o[EREB_M_ETATOTALM] = b_eta;
        o[EREB_M_HRTOTALM] = b_hr;
        o[EREB_M_KTOTALM] = b_k;
    }
#pragma unroll
    for (int f = 0; f < EREB_M_NFIELDS; ++f)
        if (A.out[f]) A.out[f][m] = o[f];
}

// value of a fixed-point grid into doubles
__global__ void k_fix_finalize(const unsigned long long* __restrict__ w, double up, int64_t n,
                               double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = fix_value(w, k, up);
}

// ---- host side ------------------------------------------------------------------------

// pieces of the context's grow-only pullback pool
struct PBPool {
    char* p;
    size_t off, cap;
    void* take(size_t bytes) {
        void* q = p + off;
        off += (bytes + 255) / 256 * 256;
        return off <= cap ? q : nullptr;
    }
};

static int pb_pool(ereb_ctx* c, size_t bytes, PBPool& P) {
    if (bytes > c->pb_bytes) {
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        hipFree(c->pb_pool);
        c->pb_pool = nullptr;
        c->pb_bytes = 0;
        EREB_HIP(c, hipMalloc(&c->pb_pool, bytes));
        c->pb_bytes = bytes;
    }
    P = PBPool{(char*)c->pb_pool, 0, c->pb_bytes};
    return EREB_OK;
}

static size_t rup(size_t b) { return (b + 255) / 256 * 256; }

// A set of fixed-point grids: words, max bits and the error flag in one
// zeroed block; after the first pass the scales are read back
struct FixSet {
    int ng = 0;
    int64_t numel[8] = {};
    FixGrid F[8];
    double* val[8] = {};  // the finished values
    unsigned long long* block = nullptr;
    size_t words = 0;  // 64-bit words in the block
    int* flag = nullptr;
};

static size_t fixset_bytes(const int64_t* numel, int ng) {
    size_t w = 8 + 32;  // max bits (8) and the flag (padded)
    for (int k = 0; k < ng; ++k) w += 2 * (size_t)numel[k];
    size_t v = 0;
    for (int k = 0; k < ng; ++k) v += rup(8 * (size_t)numel[k]);
    return rup(w * 8) + v;
}

// wanted[k]: grid k is accumulated (else its FixGrid has w == nullptr)
static int fixset_init(ereb_ctx* c, PBPool& P, FixSet& S, const int64_t* numel, const bool* wanted,
                       int ng) {
    S.ng = ng;
    size_t words = 8 + 32;
    for (int k = 0; k < ng; ++k) words += 2 * (size_t)numel[k];
    S.block = (unsigned long long*)P.take(rup(words * 8));
    if (!S.block) return fail(c, EREB_ERR_NOMEM, "pullback pool carve overflow");
    S.words = words;
    S.flag = (int*)(S.block + 8);
    unsigned long long* w = S.block + 8 + 32;
    for (int k = 0; k < ng; ++k) {
        S.numel[k] = numel[k];
        S.F[k] = FixGrid{wanted[k] ? w : nullptr, S.block + k, 1.0};
        w += 2 * (size_t)numel[k];
        S.val[k] = (double*)P.take(8 * (size_t)numel[k]);
        if (!S.val[k]) return fail(c, EREB_ERR_NOMEM, "pullback pool carve overflow");
    }
    // the block is cleared with a fill kernel (graph-safe, no memset node)
    return fill_f64(c, (double*)S.block, 0.0, (int64_t)S.words);
}

// after the first pass: the scale of each grid from its largest |contribution|
static int fixset_scale(ereb_ctx* c, FixSet& S) {
    unsigned long long mx[8];
    EREB_HIP(c, hipMemcpyAsync(mx, S.block, sizeof(unsigned long long) * 8,
                               hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < S.ng; ++k) {
        double v;
        std::memcpy(&v, &mx[k], sizeof v);
        if (!std::isfinite(v))
            return fail(c, EREB_ERR_DOMAIN, "pullback: a grid cotangent contribution is not finite");
        S.F[k].up = fix_scale_of(v);
    }
    return EREB_OK;
}

static int fixset_finish(ereb_ctx* c, FixSet& S, double* const* host_out) {
    int flag = 0;
    EREB_HIP(c, hipMemcpyAsync(&flag, S.flag, sizeof flag, hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    if (flag)
        return fail(c, EREB_ERR_DOMAIN, "pullback: a grid cotangent contribution is outside the "
                                        "fixed-point range of its pass");
    for (int k = 0; k < S.ng; ++k) {
        if (!host_out[k] || !S.F[k].w) continue;
        double* out = S.val[k];
        const int64_t n = S.numel[k];
        hipLaunchKernelGGL(k_fix_finalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           c->stream, (const unsigned long long*)S.F[k].w, S.F[k].up, n, out);
        EREB_HIP(c, hipGetLastError());
        EREB_HIP(c, hipMemcpyAsync(host_out[k], out, sizeof(double) * n, hipMemcpyDeviceToHost,
                                   c->stream));
    }
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}

static int fp64_only(ereb_ctx* c) {
    if (c->precision != EREB_FP64)
        return fail(c, EREB_ERR_ARG, "the pullbacks run in the Float64 form");
    return EREB_OK;
}

}  // namespace ereb

using namespace ereb;

#define PB_SETM(f, p)                                          \
    do {                                                       \
        int rc_ = ereb_markers_set(c, f, p, marknum);          \
        if (rc_) return rc_;                                   \
    } while (0)
#define PB_SETG(f, p)                                          \
    do {                                                       \
        int rc_ = ereb_grid_set(c, f, p);                      \
        if (rc_) return rc_;                                   \
    } while (0)

// host array -> pool (nullptr stays nullptr)
static double* pb_up(ereb_ctx* c, PBPool& P, const double* h, int64_t n, int* rc) {
    if (!h || *rc) return nullptr;
    double* d = (double*)P.take(sizeof(double) * n);
    if (!d) {
        *rc = fail(c, EREB_ERR_NOMEM, "pullback pool carve overflow");
        return nullptr;
    }
    const hipError_t e =
        hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) *rc = hipfail(c, e, "hipMemcpyAsync");
    return d;
}

static double* pb_out(ereb_ctx* c, PBPool& P, const double* h, int64_t n, int* rc) {
    if (!h || *rc) return nullptr;
    double* d = (double*)P.take(sizeof(double) * n);
    if (!d) *rc = fail(c, EREB_ERR_NOMEM, "pullback pool carve overflow");
    return d;
}

static int pb_down(ereb_ctx* c, double* h, const double* d, int64_t n) {
    if (!h || !d) return EREB_OK;
    EREB_HIP(c, hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    return EREB_OK;
}

extern "C" int ereb_move_markers_rk4_pullback(ereb_ctx* c, const double* xm, const double* ym,
    const int64_t* tm, const double* tkm, const double* phim, const double* sxym,
    const double* sxxm, const double* vx, const double* vy, const double* vxf,
    const double* vyf, const double* wyx, const double* tk2, int64_t marknum, double dt,
    int64_t mode, const double* xm_bar, const double* ym_bar, const double* tkm_bar,
    const double* sxym_bar, const double* sxxm_bar, double* xm_bar_in, double* ym_bar_in,
    double* tkm_bar_in, double* phim_bar, double* sxym_bar_in, double* sxxm_bar_in,
    double* vx_bar, double* vy_bar, double* vxf_bar, double* vyf_bar, double* wyx_bar,
    double* tk2_bar) {
    if (!c || marknum < 0) return EREB_ERR_ARG;
    if (mode != 1 && mode != 9) return fail(c, EREB_ERR_MODE, "unknown mode");
    int rc = fp64_only(c);
    if (rc) return rc;
    PB_SETM(EREB_M_XM, xm);
    PB_SETM(EREB_M_YM, ym);
    PB_SETM(EREB_M_TM, tm);
    PB_SETM(EREB_M_TKM, tkm);
    PB_SETM(EREB_M_PHIM, phim);
    PB_SETM(EREB_M_SXYM, sxym);
    PB_SETM(EREB_M_SXXM, sxxm);
    PB_SETG(EREB_G_VX, vx);
    PB_SETG(EREB_G_VY, vy);
    PB_SETG(EREB_G_VXF, vxf);
    PB_SETG(EREB_G_VYF, vyf);
    PB_SETG(EREB_G_WYX, wyx);
    PB_SETG(EREB_G_TK2, tk2);
    const int64_t M = marknum, NS = (int64_t)c->Ny1 * c->Nx1, NB = (int64_t)c->Ny * c->Nx;
    const int64_t numel[6] = {NS, NS, NS, NS, NB, NS};
    double* const gout[6] = {vx_bar, vy_bar, vxf_bar, vyf_bar, wyx_bar, tk2_bar};
    bool wanted[6];
    for (int k = 0; k < 6; ++k) wanted[k] = gout[k] != nullptr;
    PBPool P;
    if ((rc = pb_pool(c, 5 * rup(8 * M) + 6 * rup(8 * M) + fixset_bytes(numel, 6) + 4096, P)))
        return rc;
    RK4PB A{};
    A.g = geomk(c);
    A.M = c->d_mat;
    A.dt = dt;
    A.mode = (int)mode;
    A.n = M;
    A.xm = mptr(c, EREB_M_XM);
    A.ym = mptr(c, EREB_M_YM);
    A.tkm = mptr(c, EREB_M_TKM);
    A.phim = mptr(c, EREB_M_PHIM);
    A.sxym = mptr(c, EREB_M_SXYM);
    A.sxxm = mptr(c, EREB_M_SXXM);
    A.tm = tm_ptr(c);
    A.vx = gptr(c, EREB_G_VX);
    A.vy = gptr(c, EREB_G_VY);
    A.vxf = gptr(c, EREB_G_VXF);
    A.vyf = gptr(c, EREB_G_VYF);
    A.wyx = gptr(c, EREB_G_WYX);
    A.tk2 = gptr(c, EREB_G_TK2);
    A.bx = pb_up(c, P, xm_bar, M, &rc);
    A.by = pb_up(c, P, ym_bar, M, &rc);
    A.bT = pb_up(c, P, tkm_bar, M, &rc);
    A.bsxy = pb_up(c, P, sxym_bar, M, &rc);
    A.bsxx = pb_up(c, P, sxxm_bar, M, &rc);
    A.ox = pb_out(c, P, xm_bar_in, M, &rc);
    A.oy = pb_out(c, P, ym_bar_in, M, &rc);
    A.oT = pb_out(c, P, tkm_bar_in, M, &rc);
    A.ophi = pb_out(c, P, phim_bar, M, &rc);
    A.osxy = pb_out(c, P, sxym_bar_in, M, &rc);
    A.osxx = pb_out(c, P, sxxm_bar_in, M, &rc);
    if (rc) return rc;
    FixSet S;
    if ((rc = fixset_init(c, P, S, numel, wanted, 6))) return rc;
    A.flag = S.flag;
    const unsigned nb = (unsigned)((M + 255) / 256);
    auto set = [&]() {
        A.Fvx = S.F[0];
        A.Fvy = S.F[1];
        A.Fvxf = S.F[2];
        A.Fvyf = S.F[3];
        A.Fwyx = S.F[4];
        A.Ftk2 = S.F[5];
    };
    const bool any_grid = std::any_of(wanted, wanted + 6, [](bool b) { return b; });
    if (M > 0 && any_grid) {
        set();
        hipLaunchKernelGGL(k_rk4_pullback<0>, dim3(nb), dim3(256), 0, c->stream, A);
        EREB_HIP(c, hipGetLastError());
        if ((rc = fixset_scale(c, S))) return rc;
    }
    if (M > 0) {
        set();
        hipLaunchKernelGGL(k_rk4_pullback<1>, dim3(nb), dim3(256), 0, c->stream, A);
        EREB_HIP(c, hipGetLastError());
    }
    if ((rc = pb_down(c, xm_bar_in, A.ox, M)) || (rc = pb_down(c, ym_bar_in, A.oy, M)) ||
        (rc = pb_down(c, tkm_bar_in, A.oT, M)) || (rc = pb_down(c, phim_bar, A.ophi, M)) ||
        (rc = pb_down(c, sxym_bar_in, A.osxy, M)) || (rc = pb_down(c, sxxm_bar_in, A.osxx, M)))
        return rc;
    return fixset_finish(c, S, gout);
}

extern "C" int ereb_update_marker_temperature_pullback(ereb_ctx* c, const double* xm,
    const double* ym, const double* DT, const double* tk2, int64_t timestep, int64_t marknum,
    const double* tkm_bar, double* tkm_bar_in, double* DT_bar, double* tk2_bar,
    double* xm_bar, double* ym_bar) {
    if (!c || marknum < 0 || !tkm_bar) return EREB_ERR_ARG;
    int rc = fp64_only(c);
    if (rc) return rc;
    const bool first = timestep == 1;
    PB_SETM(EREB_M_XM, xm);
    PB_SETM(EREB_M_YM, ym);
    const int gfield = first ? EREB_G_TK2 : EREB_G_DT;
    PB_SETG(gfield, first ? tk2 : DT);
    const int64_t M = marknum, NS = (int64_t)c->Ny1 * c->Nx1;
    double* const gout[1] = {first ? tk2_bar : DT_bar};
    const bool wanted[1] = {gout[0] != nullptr};
    PBPool P;
    if ((rc = pb_pool(c, 4 * rup(8 * M) + fixset_bytes(&NS, 1) + 4096, P))) return rc;
    const double* tb = pb_up(c, P, tkm_bar, M, &rc);
    double* otk = pb_out(c, P, tkm_bar_in, M, &rc);
    double* ox = pb_out(c, P, xm_bar, M, &rc);
    double* oy = pb_out(c, P, ym_bar, M, &rc);
    if (rc) return rc;
    FixSet S;
    if ((rc = fixset_init(c, P, S, &NS, wanted, 1))) return rc;
    const unsigned nb = (unsigned)((M + 255) / 256);
    const GeomK g = geomk(c);
    if (M > 0 && wanted[0]) {
        hipLaunchKernelGGL(k_gather_T_pullback<0>, dim3(nb), dim3(256), 0, c->stream, g, M,
                           (const double*)mptr(c, EREB_M_XM), (const double*)mptr(c, EREB_M_YM),
                           (const double*)gptr(c, gfield), first ? 1 : 0, tb, S.F[0], S.flag,
                           otk, ox, oy);
        EREB_HIP(c, hipGetLastError());
        if ((rc = fixset_scale(c, S))) return rc;
    }
    if (M > 0) {
        hipLaunchKernelGGL(k_gather_T_pullback<1>, dim3(nb), dim3(256), 0, c->stream, g, M,
                           (const double*)mptr(c, EREB_M_XM), (const double*)mptr(c, EREB_M_YM),
                           (const double*)gptr(c, gfield), first ? 1 : 0, tb, S.F[0], S.flag,
                           otk, ox, oy);
        EREB_HIP(c, hipGetLastError());
    }
    if ((rc = pb_down(c, tkm_bar_in, otk, M)) || (rc = pb_down(c, xm_bar, ox, M)) ||
        (rc = pb_down(c, ym_bar, oy, M)))
        return rc;
    // the grid that timestep does not read has a zero cotangent
    double* other = first ? DT_bar : tk2_bar;
    if (other) std::fill(other, other + NS, 0.0);
    return fixset_finish(c, S, gout);
}

extern "C" int ereb_update_marker_stress_pullback(ereb_ctx* c, const double* xm,
    const double* ym, const double* DSXX, const double* DSXY, int64_t marknum,
    const double* sxxm_bar, const double* sxym_bar, double* DSXX_bar, double* DSXY_bar,
    double* xm_bar, double* ym_bar) {
    if (!c || marknum < 0) return EREB_ERR_ARG;
    if (c->Nx < 4 || c->Ny < 4)
        return fail(c, EREB_ERR_ARG, "update_marker_stress! needs Nx, Ny >= 4 (P clamps 2..N-1)");
    int rc = fp64_only(c);
    if (rc) return rc;
    PB_SETM(EREB_M_XM, xm);
    PB_SETM(EREB_M_YM, ym);
    PB_SETG(EREB_G_DSXX, DSXX);
    PB_SETG(EREB_G_DSXY, DSXY);
    const int64_t M = marknum;
    const int64_t numel[2] = {(int64_t)c->Ny1 * c->Nx1, (int64_t)c->Ny * c->Nx};
    double* const gout[2] = {DSXX_bar, DSXY_bar};
    const bool wanted[2] = {DSXX_bar != nullptr, DSXY_bar != nullptr};
    PBPool P;
    if ((rc = pb_pool(c, 4 * rup(8 * M) + fixset_bytes(numel, 2) + 4096, P))) return rc;
    const double* bxx = pb_up(c, P, sxxm_bar, M, &rc);
    const double* bxy = pb_up(c, P, sxym_bar, M, &rc);
    double* ox = pb_out(c, P, xm_bar, M, &rc);
    double* oy = pb_out(c, P, ym_bar, M, &rc);
    if (rc) return rc;
    FixSet S;
    if ((rc = fixset_init(c, P, S, numel, wanted, 2))) return rc;
    const unsigned nb = (unsigned)((M + 255) / 256);
    const GeomK g = geomk(c);
    for (int pass = (wanted[0] || wanted[1]) ? 0 : 1; pass < 2 && M > 0; ++pass) {
        if (pass == 0)
            hipLaunchKernelGGL(k_gather_S_pullback<0>, dim3(nb), dim3(256), 0, c->stream, g, M,
                               (const double*)mptr(c, EREB_M_XM),
                               (const double*)mptr(c, EREB_M_YM),
                               (const double*)gptr(c, EREB_G_DSXX),
                               (const double*)gptr(c, EREB_G_DSXY), bxx, bxy, S.F[0], S.F[1],
                               S.flag, ox, oy);
        else
            hipLaunchKernelGGL(k_gather_S_pullback<1>, dim3(nb), dim3(256), 0, c->stream, g, M,
                               (const double*)mptr(c, EREB_M_XM),
                               (const double*)mptr(c, EREB_M_YM),
                               (const double*)gptr(c, EREB_G_DSXX),
                               (const double*)gptr(c, EREB_G_DSXY), bxx, bxy, S.F[0], S.F[1],
                               S.flag, ox, oy);
        EREB_HIP(c, hipGetLastError());
        if (pass == 0 && (rc = fixset_scale(c, S))) return rc;
    }
    if ((rc = pb_down(c, xm_bar, ox, M)) || (rc = pb_down(c, ym_bar, oy, M))) return rc;
    return fixset_finish(c, S, gout);
}

extern "C" int ereb_marker_to_nodes_pullback(ereb_ctx* c, const void* const* markers,
    int64_t marknum, const double hrsolidm[3], const double hrfluidm[3], int64_t mode,
    const double* const* sum_bar, const double* const* prop_bar, double* const* marker_bar) {
    if (!c || !markers || !hrsolidm || !hrfluidm || !marker_bar || marknum < 0)
        return EREB_ERR_ARG;
    if (mode != 1 && mode != 9) return fail(c, EREB_ERR_MODE, "unknown mode");
    int rc = fp64_only(c);
    if (rc) return rc;
    static const int need[] = {EREB_M_XM, EREB_M_YM, EREB_M_TM, EREB_M_TKM, EREB_M_PHIM};
    for (int f : need)
        if (!markers[f]) return fail(c, EREB_ERR_ARG, "marker_to_nodes pullback: xm, ym, tm, "
                                                      "tkm and phim are required");
    for (int f = 0; f < EREB_M_NFIELDS; ++f)
        if (markers[f]) PB_SETM(f, markers[f]);
    if ((rc = set_heating(c, hrsolidm, hrfluidm))) return rc;
    const int64_t M = marknum, NS = (int64_t)c->Ny1 * c->Nx1, NB = (int64_t)c->Ny * c->Nx;
    PBPool P;
    if ((rc = pb_pool(c, 30 * rup(8 * NS) + 7 * rup(8 * M) + EREB_M_NFIELDS * rup(8 * M) + 4096,
                      P)))
        return rc;
    ScatterPB A{};
    A.g = geomk(c);
    A.M = c->d_mat;
    A.S = c->d_scal;
    A.mode = (int)mode;
    A.n = M;
    for (int f = 0; f < EREB_M_NFIELDS; ++f)
        A.f[f] = markers[f] && f != EREB_M_TM ? mptr(c, f) : nullptr;
    A.tm = tm_ptr(c);
    for (int s = 0; s < NSUM; ++s)
        A.sb[s] = sum_bar ? pb_up(c, P, sum_bar[s], s < 8 ? NB : NS, &rc) : nullptr;
    for (int k = 0; k < 7; ++k) A.pb[k] = prop_bar ? pb_up(c, P, prop_bar[k], M, &rc) : nullptr;
    for (int f = 0; f < EREB_M_NFIELDS; ++f) A.out[f] = pb_out(c, P, marker_bar[f], M, &rc);
    if (rc) return rc;
    if (M > 0) {
        hipLaunchKernelGGL(k_scatter_pullback, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                           c->stream, A);
        EREB_HIP(c, hipGetLastError());
    }
    for (int f = 0; f < EREB_M_NFIELDS; ++f)
        if ((rc = pb_down(c, marker_bar[f], A.out[f], M))) return rc;
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}
