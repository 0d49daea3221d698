// ereb_markers.hip -- node->marker transfers: update_marker_temperature!
// (E:4782-4801), update_marker_stress! (E:4354-4388) and RK4 advection
// move_markers_rk4! (E:4990-5257).  One thread per marker in storage order
// (coalesced SoA reads); the grids (at most a few MB each) are read through
// the L2/MALL.
//
// RK4 is bound by FP64 instruction issue: the reference divides by dx or dy
// 87 times per marker.  Here those quotients are products with the rounded
// reciprocals 1/dx, 1/dy (within an ulp of the quotient) and products and
// sums may fuse into fma, so positions and temperatures agree with the
// reference to ~1e-15 relative rather than bit for bit; cell indices
// computed from given positions stay exact.
#include "ereb_internal.h"

namespace ereb {

// R = double: the reference's Float64 interpolation.  R = float (mixed form):
// weights and the bilinear interpolation in FP32, the increment added to the
// Float64 marker value.
template <class R>
__device__ __forceinline__ R dot4_r(const double* __restrict__ gr, int rows, int i, int j,
                                    const R* w) {
    const double* p = gr + (size_t)(j - 1) * rows + (i - 1);
    return (R)p[0] * w[0] + (R)p[1] * w[1] + (R)p[rows] * w[2] + (R)p[rows + 1] * w[3];
}

template <class R>
__device__ __forceinline__ void weights4_r(double a, double b, R* w) {
    const R ar = (R)a, br = (R)b;
    w[0] = (R(1) - br) * (R(1) - ar);
    w[1] = br * (R(1) - ar);
    w[2] = (R(1) - br) * ar;
    w[3] = br * ar;
}

// update_marker_temperature! (E:4782-4801) for one marker: the new tkm
template <class R, bool EXACT>
__device__ __forceinline__ double gather_T_one(const GeomK& g, double x, double y, double T,
                                               const double* __restrict__ grid, bool first) {
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    R w[4];
    weights4_r<R>((x - axis_node<EXACT>(g.xp, g.dx, jp)) / g.dx,
                  (y - axis_node<EXACT>(g.yp, g.dy, ip)) / g.dy, w);
    const double v = (double)dot4_r<R>(grid, g.Ny1, ip, jp, w);
    return first ? v : T + v;
}

// update_marker_stress! (E:4354-4388) for one marker: sxxm, sxym updated in
// place; P grid with the clamps 2..Nx-1, 2..Ny-1 of E:4364-4367
template <class R, bool EXACT>
__device__ __forceinline__ void gather_S_one(const GeomK& g, double x, double y, double& sxx,
                                             double& sxy, const double* __restrict__ DSXX,
                                             const double* __restrict__ DSXY) {
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 2, g.Nx - 1);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 2, g.Ny - 1);
    const int jb = jl_cell((x - g.x[0]) * g.idx, 1, g.Nx - 1);
    const int ib = jl_cell((y - g.y[0]) * g.idy, 1, g.Ny - 1);
    R wp[4], wb[4];
    weights4_r<R>((x - axis_node<EXACT>(g.xp, g.dx, jp)) / g.dx,
                  (y - axis_node<EXACT>(g.yp, g.dy, ip)) / g.dy, wp);
    weights4_r<R>((x - axis_node<EXACT>(g.x, g.dx, jb)) / g.dx,
                  (y - axis_node<EXACT>(g.y, g.dy, ib)) / g.dy, wb);
    sxx += (double)dot4_r<R>(DSXX, g.Ny1, ip, jp, wp);
    sxy += (double)dot4_r<R>(DSXY, g.Ny, ib, jb, wb);
}

template <class R>
__global__ __launch_bounds__(256) void k_gather_T(GeomK g, const double* __restrict__ xm,
                                                  const double* __restrict__ ym,
                                                  double* __restrict__ tkm,
                                                  const double* __restrict__ grid, int64_t M,
                                                  int first) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    tkm[m] = gather_T_one<R, false>(g, xm[m], ym[m], first ? 0.0 : tkm[m], grid, first != 0);
}

template <class R>
__global__ __launch_bounds__(256) void k_gather_S(GeomK g, const double* __restrict__ xm,
                                                  const double* __restrict__ ym,
                                                  double* __restrict__ sxxm,
                                                  double* __restrict__ sxym,
                                                  const double* __restrict__ DSXX,
                                                  const double* __restrict__ DSXY, int64_t M) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double sxx = sxxm[m], sxy = sxym[m];
    gather_S_one<R, false>(g, xm[m], ym[m], sxx, sxy, DSXX, DSXY);
    sxxm[m] = sxx;
    sxym[m] = sxy;
}

// RK4 region: a*b + c may be fused (the quotients above already make the
// results agree to rounding rather than bit for bit)
#pragma clang fp contract(fast)

// RK4 stage velocity on the Vx grid (E:5057-5091).  `vc` supplies the
// centre value of the left-half correction: the reference reads vx there
// also in the fluid loop (E:5173).  Every load of the stencil (rows i, i+1
// of columns j-1 .. j+2) is issued before any is used and the second-order
// correction is selected, not branched to, so a stage costs one memory round
// trip.
template <bool EXACT>
__device__ __forceinline__ double rk_vx(const GeomK& g, const double* __restrict__ v,
                                        const double* __restrict__ vc, double x, double y) {
    const int R = g.Ny1;
    const double x0 = g.x[0], y0 = g.yp[0];
    const int j = jl_cell((x - x0) * g.idx, 1, g.Nx - 1);
    const int i = jl_cell((y - y0) * g.idy, 1, g.Ny);
    const double dxmj = x - axis_node<EXACT>(g.x, g.dx, j);
    const double dymi = y - axis_node<EXACT>(g.yp, g.dy, i);
    const double idx = g.idx, idy = g.idy;
    const double* p = gptr_at(v, i, j, R);  // v[i, j]
    // column j-1 (read at j only where j = 1: it is unused there); column
    // j+2 <= Nx+1 always exists
    const double* pl = j > 1 ? p - R : p;
    const double l0 = pl[0], l1 = pl[1];
    const double a0 = p[0], a1 = p[1], b0 = p[R], b1 = p[R + 1];
    const double c0 = p[2 * R], c1 = p[2 * R + 1];
    const double vcij = *gptr_at(vc, i, j, R);
    const double fx = dxmj * idx, fy = dymi * idy;
    double v13 = a0 * (1.0 - fx) + (b0 * dxmj) * idx;
    double v24 = a1 * (1.0 - fx) + (b1 * dxmj) * idx;
    // second-order correction on columns j..j+2 (fx >= 0.5) or j-1..j+1
    const bool right = fx >= 0.5;
    const bool apply = right ? j < g.Nx - 1 : j > 1;
    const double t = fx - 0.5, h = 0.5 * (t * t);
    const double q0 = right ? a0 : l0, q1 = right ? a1 : l1;    // first column
    const double q20 = right ? c0 : b0, q21 = right ? c1 : b1;  // third column
    const double c13 = right ? b0 : vcij, c24 = right ? b1 : a1;
    const double n13 = v13 + h * (q0 - 2.0 * c13 + q20);
    const double n24 = v24 + h * (q1 - 2.0 * c24 + q21);
    v13 = apply ? n13 : v13;
    v24 = apply ? n24 : v24;
    return v13 * (1.0 - fy) + (v24 * dymi) * idy;
}

// RK4 stage velocity on the Vy grid (E:5093-5127), rows i-1 .. i+2 of
// columns j, j+1 loaded together as in rk_vx.  Row i-1 at i = 1 is unused:
// the address is then the previous column's last row or, for the first
// column, one of the guard elements the context keeps before its grids.
template <bool EXACT>
__device__ __forceinline__ double rk_vy(const GeomK& g, const double* __restrict__ v, double x,
                                        double y) {
    const int R = g.Ny1;
    const double x0 = g.xp[0], y0 = g.y[0];
    const int j = jl_cell((x - x0) * g.idx, 1, g.Nx);
    const int i = jl_cell((y - y0) * g.idy, 1, g.Ny - 1);
    const double dxmj = x - axis_node<EXACT>(g.xp, g.dx, j);
    const double dymi = y - axis_node<EXACT>(g.y, g.dy, i);
    const double idx = g.idx, idy = g.idy;
    const double* p = gptr_at(v, i, j, R);
    const double am = p[-1], a0 = p[0], a1 = p[1], a2 = p[2];
    const double bm = p[R - 1], b0 = p[R], b1 = p[R + 1], b2 = p[R + 2];
    const double fx = dxmj * idx, fy = dymi * idy;
    double v12 = a0 * (1.0 - fy) + (a1 * dymi) * idy;
    double v34 = b0 * (1.0 - fy) + (b1 * dymi) * idy;
    // second-order correction on rows i..i+2 (fy >= 0.5) or i-1..i+1
    const bool down = fy >= 0.5;
    const bool apply = down ? i < g.Ny - 1 : i > 1;
    const double t = fy - 0.5, h = 0.5 * (t * t);
    const double q0 = down ? a0 : am, q1 = down ? a1 : a0, q2 = down ? a2 : a1;
    const double r0 = down ? b0 : bm, r1 = down ? b1 : b0, r2 = down ? b2 : b1;
    const double n12 = v12 + h * (q0 - 2.0 * q1 + q2);
    const double n34 = v34 + h * (r0 - 2.0 * r1 + r2);
    v12 = apply ? n12 : v12;
    v34 = apply ? n34 : v34;
    return v12 * (1.0 - fx) + (v34 * dxmj) * idx;
}

struct RK4Args {
    GeomK g;
    ereb_materials M;
    const StepScalars* __restrict__ S;  // dt and, for the fused gathers, first
    int mode;
    int64_t n;
    double *xm, *ym, *tkm, *sxxm, *sxym;
    const int32_t* tm;
    const double* phim;
    const double *vx, *vy, *vxf, *vyf, *wyx, *tk2;
    double *xb_out, *yb_out;  // optional: fluid back-trace end points (adjoint recording)
    // fused gathers (k_rk4<1 or 2>): DSXX, DSXY and the DT grid (the
    // temperature gather reads tk2 at timestep 1, else DT)
    const double *DSXX, *DSXY, *DTg;
};

// Marker paths of one RK4 step: the solid step (E:5055-5136) from (x0, y0)
// to (xn, yn) and the fluid back-trace (E:5138-5226) from (xn, yn) to
// (xb, yb).
template <bool EXACT>
__device__ __forceinline__ void rk4_paths(const RK4Args& A, double dt, double x0, double y0,
                                          double& xn, double& yn, double& xb, double& yb) {
    const GeomK& g = A.g;
    const double* b = A.M.brk4;
    const double* cr = A.M.crk4;
    double xmm = x0, ymm = y0, xr = x0, yr = y0;
    double vxr[4], vyr[4];
#pragma unroll
    for (int rk = 0; rk < 4; ++rk) {
        vxr[rk] = 0.0 + rk_vx<EXACT>(g, A.vx, A.vx, xr, yr);
        vyr[rk] = 0.0 + rk_vy<EXACT>(g, A.vy, xr, yr);
        if (rk < 3) {
            xr = xmm + dt * cr[rk] * vxr[rk];
            yr = ymm + dt * cr[rk] * vyr[rk];
        }
    }
    xn = xmm + dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
    yn = ymm + dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
    xmm = xr = xn;
    ymm = yr = yn;
#pragma unroll
    for (int rk = 0; rk < 4; ++rk) {
        vxr[rk] = 0.0 + rk_vx<EXACT>(g, A.vxf, A.vx, xr, yr);
        vyr[rk] = 0.0 + rk_vy<EXACT>(g, A.vyf, xr, yr);
        if (rk < 3) {
            xr = xmm - dt * cr[rk] * vxr[rk];
            yr = ymm - dt * cr[rk] * vyr[rk];
        }
    }
    xb = xmm - dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
    yb = ymm - dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
}

// bilinear P-grid stencil of a position: cell (ip, jp), weights w
template <bool EXACT>
__device__ __forceinline__ void p_stencil(const GeomK& g, double x, double y, int& ip, int& jp,
                                          double* w) {
    const double x0 = g.xp[0], y0 = g.yp[0];
    jp = jl_cell((x - x0) * g.idx, 1, g.Nx);
    ip = jl_cell((y - y0) * g.idy, 1, g.Ny);
    weights4((x - axis_node<EXACT>(g.xp, g.dx, jp)) * g.idx,
             (y - axis_node<EXACT>(g.yp, g.dy, ip)) * g.idy, w);
}

// GATHER = 0: RK4 alone.  GATHER = 1 (Float64) or 2 (FP32 interpolation, the
// mixed form): the stress gather, then the temperature gather, then RK4 on
// the gathered values, as the three separate kernels would compute them
// (the gathers' arithmetic is defined outside the fused-FP region), with
// each marker's state read and written once.
template <int GATHER, bool EXACT>
__global__ __launch_bounds__(256) void k_rk4(RK4Args A) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= A.n) return;
    const GeomK& g = A.g;
    const double dt = A.S->dt;
    const bool first = GATHER && A.S->first != 0;
    const double* tgrid = first ? A.tk2 : A.DTg;
    const double xmm = A.xm[m], ymm = A.ym[m];
    double s = A.sxxm[m], t = A.sxym[m];
    double T = first ? 0.0 : A.tkm[m];
    if (GATHER == 1) {
        gather_S_one<double, EXACT>(g, xmm, ymm, s, t, A.DSXX, A.DSXY);
        T = gather_T_one<double, EXACT>(g, xmm, ymm, T, tgrid, first);
    } else if (GATHER == 2) {
        gather_S_one<float, EXACT>(g, xmm, ymm, s, t, A.DSXX, A.DSXY);
        T = gather_T_one<float, EXACT>(g, xmm, ymm, T, tgrid, first);
    }
    double xr = xmm, yr = ymm;
    double w[4];
    // solid temperature at the start position (P grid) and rotation rate
    // (basic grid), E:5013-5040
    int jp, ip;
    p_stencil<EXACT>(g, xr, yr, ip, jp, w);
    const double tks0 = gdot4(A.tk2, g.Ny1, ip, jp, w);
    const double xb0 = g.x[0], yb0 = g.y[0];
    const int jb = jl_cell((xr - xb0) * g.idx, 1, g.Nx - 1);
    const int ib = jl_cell((yr - yb0) * g.idy, 1, g.Ny - 1);
    weights4((xr - axis_node<EXACT>(g.x, g.dx, jb)) * g.idx,
             (yr - axis_node<EXACT>(g.y, g.dy, ib)) * g.idy, w);
    const double om = gdot4(A.wyx, g.Ny, ib, jb, w);
    // analytic stress rotation E:5042-5048
    const double th = dt * om;
    // sin(2θ) and cos(2θ) from the double-angle identities (within a few ulp
    // of the library values; one argument reduction and polynomial instead
    // of two)
    double ct, st;
    sincos(th, &st, &ct);
    const double s2 = 2.0 * (st * ct), c2 = (ct - st) * (ct + st);
    A.sxxm[m] = s * (ct * ct) - s * (st * st) - t * s2;
    A.sxym[m] = s * s2 + t * c2;
    // solid RK4 step and fluid back-trace
    double xn, yn;
    rk4_paths<EXACT>(A, dt, xmm, ymm, xn, yn, xr, yr);
    A.xm[m] = xn;
    A.ym[m] = yn;
    if (A.xb_out) {
        A.xb_out[m] = xr;
        A.yb_out[m] = yr;
    }
    p_stencil<EXACT>(g, xr, yr, ip, jp, w);
    const double tkf0 = gdot4(A.tk2, g.Ny1, ip, jp, w);
    // fluid-solid temperature correction E:5243-5252
    const double d = tkf0 - tks0;
    const int tt = min(max(A.tm[m], 1), 3);  // set validates 1..3
    const double ph = A.phim[m];
    const double rcs = sel3(A.M.rhocpsolidm, tt);
    const double rcf = rhocpfluidm(A.M, T, A.mode);
    A.tkm[m] = ((1.0 - ph) * T * rcs + ph * (T + d) * rcf) / ((1.0 - ph) * rcs + ph * rcf);
}

// ---- adjoints (with respect to tkm, phim and the tk2 grid) ---------------

__device__ __forceinline__ void atomic_add4(double* gr, int rows, int i, int j, const double* w,
                                            double v) {
    double* p = gr + (size_t)(j - 1) * rows + (i - 1);
    atomicAdd(p, w[0] * v);
    atomicAdd(p + 1, w[1] * v);
    atomicAdd(p + rows, w[2] * v);
    atomicAdd(p + rows + 1, w[3] * v);
}

constexpr int ADJ_WIN = 2048;  // LDS window of the workgroup pre-aggregation (doubles)

struct AddWin {
    double v[ADJ_WIN];
    int box[4];  // imin, imax, jmin, jmax (1-based cells)
};

// Workgroup-wide gr[stencil of (i, j)] += w * v (bilinear, `rows` rows): the
// contributions are first summed in an LDS window over the workgroup's cells
// (LDS atomics), then added to gr with one atomic per touched node; when the
// cells span more than the window, every lane adds straight to gr.  Every
// thread of the workgroup must call it (it synchronises); `on` marks the
// lanes that contribute.
__device__ void block_add4(double* gr, int rows, bool on, int i, int j, const double* w,
                           double v, AddWin& W) {
    on = on && v != 0.0;
    int imin = on ? i : INT32_MAX, imax = on ? i : INT32_MIN;
    int jmin = on ? j : INT32_MAX, jmax = on ? j : INT32_MIN;
    for (int o = 32; o > 0; o >>= 1) {
        imin = min(imin, __shfl_xor(imin, o));
        imax = max(imax, __shfl_xor(imax, o));
        jmin = min(jmin, __shfl_xor(jmin, o));
        jmax = max(jmax, __shfl_xor(jmax, o));
    }
    __syncthreads();  // every thread is done with the previous call's box and window
    if (threadIdx.x == 0) {
        W.box[0] = INT32_MAX;
        W.box[1] = INT32_MIN;
        W.box[2] = INT32_MAX;
        W.box[3] = INT32_MIN;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&W.box[0], imin);
        atomicMax(&W.box[1], imax);
        atomicMin(&W.box[2], jmin);
        atomicMax(&W.box[3], jmax);
    }
    __syncthreads();
    const int i0 = W.box[0], i1 = W.box[1], j0 = W.box[2], j1 = W.box[3];
    if (i0 > i1) return;  // uniform: no contribution in the workgroup
    const int h = i1 - i0 + 2, wdt = j1 - j0 + 2;  // nodes i0..i1+1, j0..j1+1
    if ((int64_t)h * wdt > ADJ_WIN) {
        if (on) atomic_add4(gr, rows, i, j, w, v);
        return;
    }
    for (int q = threadIdx.x; q < h * wdt; q += blockDim.x) W.v[q] = 0.0;
    __syncthreads();
    if (on) {
        double* p = W.v + (j - j0) * h + (i - i0);
        atomicAdd(p, w[0] * v);
        atomicAdd(p + 1, w[1] * v);
        atomicAdd(p + h, w[2] * v);
        atomicAdd(p + h + 1, w[3] * v);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < h * wdt; q += blockDim.x) {
        const double a = W.v[q];
        if (a != 0.0) atomicAdd(&gr[(size_t)(j0 - 1 + q / h) * rows + (i0 - 1 + q % h)], a);
    }
    __syncthreads();  // the window is reused by the next call
}

// RK4^T of the temperature correction E:5243-5252: given the cotangent gout of
// tkm after the step, the cotangent of tkm before it (gin), the increment of
// the phim cotangent and the tk2 cotangent (both stencils, by atomics).
// Positions and stresses do not depend on tkm, phim or tk2; the fluid
// back-trace end points (xb, yb) were recorded by the forward step.
__global__ __launch_bounds__(256) void k_rk4_adj(RK4Args A, const double* __restrict__ xbs,
                                                 const double* __restrict__ ybs,
                                                 const double* __restrict__ gout,
                                                 double* __restrict__ gin,
                                                 double* __restrict__ gphi,
                                                 double* __restrict__ gtk2) {
    __shared__ AddWin W;
    const int64_t m0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = m0 < A.n;
    const int64_t m = on ? m0 : A.n - 1;  // idle lanes read a valid marker, add nothing
    const GeomK& g = A.g;
    const double x0 = A.xm[m], y0 = A.ym[m];
    int ips, jps, ipb, jpb;
    double ws[4], wb[4];
    p_stencil<false>(g, x0, y0, ips, jps, ws);
    const double tks0 = gdot4(A.tk2, g.Ny1, ips, jps, ws);
    p_stencil<false>(g, xbs[m], ybs[m], ipb, jpb, wb);
    const double tkf0 = gdot4(A.tk2, g.Ny1, ipb, jpb, wb);
    const double d = tkf0 - tks0;
    const int tt = min(max(A.tm[m], 1), 3);
    const double T = A.tkm[m], ph = A.phim[m];
    const double rcs = sel3(A.M.rhocpsolidm, tt);
    const double rcf = rhocpfluidm(A.M, T, A.mode), drcf = d_rhocpfluidm(A.M, T, A.mode);
    const double num = (1.0 - ph) * T * rcs + ph * (T + d) * rcf;
    const double den = (1.0 - ph) * rcs + ph * rcf;
    const double F_d = ph * rcf / den;
    const double F_T = ((1.0 - ph) * rcs + ph * rcf + ph * (T + d) * drcf) / den -
                       num * (ph * drcf) / (den * den);
    const double F_p = (-T * rcs + (T + d) * rcf) / den - num * (rcf - rcs) / (den * den);
    const double go = gout[m];
    if (on) {
        gin[m] = go * F_T;
        gphi[m] += go * F_p;
    }
    const double gd = go * F_d;
    block_add4(gtk2, g.Ny1, on, ipb, jpb, wb, gd, W);
    block_add4(gtk2, g.Ny1, on, ips, jps, ws, -gd, W);
}

#pragma clang fp contract(off)

// gather^T of update_marker_temperature! (E:4782-4801): the marker cotangent
// goes to the interpolated grid (tk2 at timestep 1, DT later); at timestep 1
// the marker value was overwritten, so its cotangent becomes zero.
__global__ __launch_bounds__(256) void k_gather_T_adj(GeomK g, const double* __restrict__ xm,
                                                      const double* __restrict__ ym,
                                                      double* __restrict__ gT,
                                                      double* __restrict__ ggrid, int64_t M,
                                                      int first) {
    __shared__ AddWin W;
    const int64_t m0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = m0 < M;
    const int64_t m = on ? m0 : M - 1;
    const double x = xm[m], y = ym[m];
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    double w[4];
    weights4((x - g.xp[jp - 1]) / g.dx, (y - g.yp[ip - 1]) / g.dy, w);
    const double v = gT[m];
    block_add4(ggrid, g.Ny1, on, ip, jp, w, v, W);
    if (on && first) gT[m] = 0.0;
}

int launch_gather_T_adj(ereb_ctx* c, int64_t timestep, double* gT, double* ggrid) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    hipLaunchKernelGGL(k_gather_T_adj, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                       c->stream, geomk(c), mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), gT, ggrid, M,
                       timestep == 1 ? 1 : 0);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

static RK4Args rk4_args(ereb_ctx* c, int mode);

int launch_rk4_adj(ereb_ctx* c, double dt, int mode, const double* xb, const double* yb,
                   const double* gout, double* gin, double* gphi, double* gtk2) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    int rc;
    if ((rc = set_dt(c, dt))) return rc;
    RK4Args A = rk4_args(c, mode);
    hipLaunchKernelGGL(k_rk4_adj, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                       A, xb, yb, gout, gin, gphi, gtk2);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

int launch_gather_temperature(ereb_ctx* c, int64_t timestep) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    hipEvent_t ev;
    prof_begin(c, EREB_K_GATHER_T, &ev);
    const double* grid = gptr(c, timestep == 1 ? EREB_G_TK2 : EREB_G_DT);
    if (c->precision == EREB_MIXED)
        hipLaunchKernelGGL(k_gather_T<float>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                           c->stream, geomk(c), mptr(c, EREB_M_XM), mptr(c, EREB_M_YM),
                           mptr(c, EREB_M_TKM), grid, M, timestep == 1 ? 1 : 0);
    else
        hipLaunchKernelGGL(k_gather_T<double>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                           c->stream, geomk(c), mptr(c, EREB_M_XM), mptr(c, EREB_M_YM),
                           mptr(c, EREB_M_TKM), grid, M, timestep == 1 ? 1 : 0);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_GATHER_T, ev);
    return EREB_OK;
}

int launch_gather_stress(ereb_ctx* c) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    if (c->Nx < 4 || c->Ny < 4)
        return fail(c, EREB_ERR_ARG, "update_marker_stress! needs Nx, Ny >= 4 (P clamps 2..N-1)");
    hipEvent_t ev;
    prof_begin(c, EREB_K_GATHER_S, &ev);
    auto kern = c->precision == EREB_MIXED ? k_gather_S<float> : k_gather_S<double>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                       geomk(c), mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), mptr(c, EREB_M_SXXM),
                       mptr(c, EREB_M_SXYM), gptr(c, EREB_G_DSXX), gptr(c, EREB_G_DSXY), M);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_GATHER_S, ev);
    return EREB_OK;
}

static RK4Args rk4_args(ereb_ctx* c, int mode) {
    const int64_t M = c->marknum;
    RK4Args A;
    A.g = geomk(c);
    A.M = c->mat;
    A.S = c->d_scal;
    A.mode = mode;
    A.n = M;
    A.xm = mptr(c, EREB_M_XM);
    A.ym = mptr(c, EREB_M_YM);
    A.tkm = mptr(c, EREB_M_TKM);
    A.sxxm = mptr(c, EREB_M_SXXM);
    A.sxym = mptr(c, EREB_M_SXYM);
    A.tm = tm_ptr(c);
    A.phim = mptr(c, EREB_M_PHIM);
    A.vx = gptr(c, EREB_G_VX);
    A.vy = gptr(c, EREB_G_VY);
    A.vxf = gptr(c, EREB_G_VXF);
    A.vyf = gptr(c, EREB_G_VYF);
    A.wyx = gptr(c, EREB_G_WYX);
    A.tk2 = gptr(c, EREB_G_TK2);
    A.xb_out = A.yb_out = nullptr;
    A.DSXX = A.DSXY = A.DTg = nullptr;
    return A;
}

int launch_rk4(ereb_ctx* c, double dt, int mode, double* xb_out, double* yb_out) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    int rc;
    if ((rc = set_dt(c, dt))) return rc;
    RK4Args A = rk4_args(c, mode);
    A.xb_out = xb_out;
    A.yb_out = yb_out;
    ++c->pos_gen;
    hipEvent_t ev;
    prof_begin(c, EREB_K_RK4, &ev);
    const dim3 grid((unsigned)((M + 255) / 256));
    if (c->axes_exact)
        hipLaunchKernelGGL((k_rk4<0, true>), grid, dim3(256), 0, c->stream, A);
    else
        hipLaunchKernelGGL((k_rk4<0, false>), grid, dim3(256), 0, c->stream, A);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_RK4, ev);
    return EREB_OK;
}

int launch_gathers_rk4(ereb_ctx* c, double dt, int mode, int64_t timestep) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    if (c->Nx < 4 || c->Ny < 4)
        return fail(c, EREB_ERR_ARG, "update_marker_stress! needs Nx, Ny >= 4 (P clamps 2..N-1)");
    int rc;
    if ((rc = set_dt(c, dt)) || (rc = set_first(c, timestep == 1))) return rc;
    RK4Args A = rk4_args(c, mode);
    A.DSXX = gptr(c, EREB_G_DSXX);
    A.DSXY = gptr(c, EREB_G_DSXY);
    A.DTg = gptr(c, EREB_G_DT);
    ++c->pos_gen;
    hipEvent_t ev;
    prof_begin(c, EREB_K_RK4, &ev);
    const dim3 grid((unsigned)((M + 255) / 256));
    // node coordinates computed (exact axes) or loaded, chosen per kernel
    if (c->precision == EREB_MIXED) {
        if (c->axes_exact)
            hipLaunchKernelGGL((k_rk4<2, true>), grid, dim3(256), 0, c->stream, A);
        else
            hipLaunchKernelGGL((k_rk4<2, false>), grid, dim3(256), 0, c->stream, A);
    } else {
        if (c->axes_exact)
            hipLaunchKernelGGL((k_rk4<1, true>), grid, dim3(256), 0, c->stream, A);
        else
            hipLaunchKernelGGL((k_rk4<1, false>), grid, dim3(256), 0, c->stream, A);
    }
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_RK4, ev);
    return EREB_OK;
}

}  // namespace ereb
