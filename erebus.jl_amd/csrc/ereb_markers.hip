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
#include "ereb_fixsum.h"
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

// Second differences of the stage velocities for the RK4 corrections
// (E:5068-5088, E:5104-5124), one thread per node of the (Ny1, Nx1) grids,
// in the reference's association (q0 - 2 c) + q2:
//   DX[i, j]  = vx[i, j-1] - 2 vx[i, j] + vx[i, j+1]     for 2 <= j <= Nx-1
//   DXf       = the same of vxf
//   DXq[i, j] = vxf[i, j-1] - 2 vx[i, j] + vxf[i, j+1]   (E:5173)
//   DY[i, j]  = vy[i-1, j] - 2 vy[i, j] + vy[i+1, j]     for 2 <= i <= Ny-1
//   DYf       = the same of vyf
// and zero elsewhere: a correction centred there is the one the reference
// skips (j == 1 left, j == Nx-1 right; i == 1 up, i == Ny-1 down).
__global__ __launch_bounds__(256) void k_rk4_prep(int Nx, int Ny, const double* __restrict__ vx,
                                                  const double* __restrict__ vy,
                                                  const double* __restrict__ vxf,
                                                  const double* __restrict__ vyf,
                                                  double* __restrict__ DX, double* __restrict__ DXf,
                                                  double* __restrict__ DXq, double* __restrict__ DY,
                                                  double* __restrict__ DYf) {
    const int R = Ny + 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= R * (Nx + 1)) return;
    const int i = k % R + 1, j = k / R + 1;
    const bool cx = j >= 2 && j <= Nx - 1, cy = i >= 2 && i <= Ny - 1;
    DX[k] = cx ? (vx[k - R] - 2.0 * vx[k]) + vx[k + R] : 0.0;
    DXf[k] = cx ? (vxf[k - R] - 2.0 * vxf[k]) + vxf[k + R] : 0.0;
    DXq[k] = cx ? (vxf[k - R] - 2.0 * vx[k]) + vxf[k + R] : 0.0;
    DY[k] = cy ? (vy[k - 1] - 2.0 * vy[k]) + vy[k + 1] : 0.0;
    DYf[k] = cy ? (vyf[k - 1] - 2.0 * vyf[k]) + vyf[k + 1] : 0.0;
}

// RK4 region: a*b + c may be fused (the quotients above already make the
// results agree to rounding rather than bit for bit)
#pragma clang fp contract(fast)

// Cell of an RK4 stage position: the clamp to [lo - 1, hi - 1] before the
// conversion equals fix's truncate-then-clamp (E:1457-1468) for every t
// below 2^63 in magnitude, NaN included (it maps to lo in both); stage
// positions are those of the marker moved by dt * velocity.
__device__ __forceinline__ int stage_cell(double t, int lo, int hi) {
    return (int)fmin(fmax(t, (double)(lo - 1)), (double)(hi - 1)) + 1;
}

// Fraction of a position within its cell along an axis: t = (x - x0) / h
// less the cell's index when the axis nodes are x0 + k h (EXACT), else from
// the loaded node.
template <bool EXACT>
__device__ __forceinline__ double stage_frac(const double* __restrict__ ax, double x, double t,
                                             double ih, int k) {
    if constexpr (EXACT) return t - (double)(k - 1);
    else return (x - ax[k - 1]) * ih;
}

// RK4 stage velocity on the Vx grid (E:5057-5091):
//   v13 = v[i,j] (1 - fx) + v[i,j+1] fx + h (v[i,j-1] - 2 v[i,j] + v[i,j+1])
// (the correction on columns j..j+2 instead when fx >= 0.5; only where
// those columns exist), the same for row i+1, then v13 (1 - fy) + v24 fy.
// The second differences are read from grids prepared once per step
// (k_rk4_prep): D at the correction's centre column, zero where the
// reference skips the correction, and Dq for the first row of the left-half
// correction, whose centre the reference reads from vx also in the fluid
// loop (E:5173); Dq == D in the solid loop.  Six loads, no selects of
// values.
template <bool EXACT>
__device__ __forceinline__ double rk_vx(const GeomK& g, const double* __restrict__ v,
                                        const double* __restrict__ D,
                                        const double* __restrict__ Dq, double x, double y) {
    const int R = g.Ny1;
    const double tx = (x - g.x[0]) * g.idx, ty = (y - g.yp[0]) * g.idy;
    const int j = stage_cell(tx, 1, g.Nx - 1);
    const int i = stage_cell(ty, 1, g.Ny);
    const double fx = stage_frac<EXACT>(g.x, x, tx, g.idx, j);
    const double fy = stage_frac<EXACT>(g.yp, y, ty, g.idy, i);
    const uint32_t k = (uint32_t)((j - 1) * R + (i - 1));
    const bool right = fx >= 0.5;
    const uint32_t kc = right ? k + R : k;
    const double a0 = v[k], a1 = v[k + 1], b0 = v[k + R], b1 = v[k + R + 1];
    const double d0 = (right ? D : Dq)[kc], d1 = D[kc + 1];
    const double t = fx - 0.5, h = 0.5 * (t * t);
    const double v13 = a0 + fx * (b0 - a0) + h * d0;
    const double v24 = a1 + fx * (b1 - a1) + h * d1;
    return v13 + fy * (v24 - v13);
}

// RK4 stage velocity on the Vy grid (E:5093-5127), the transpose: second
// differences along y at the correction's centre row (D, zero where the
// reference skips it).
template <bool EXACT>
__device__ __forceinline__ double rk_vy(const GeomK& g, const double* __restrict__ v,
                                        const double* __restrict__ D, double x, double y) {
    const int R = g.Ny1;
    const double tx = (x - g.xp[0]) * g.idx, ty = (y - g.y[0]) * g.idy;
    const int j = stage_cell(tx, 1, g.Nx);
    const int i = stage_cell(ty, 1, g.Ny - 1);
    const double fx = stage_frac<EXACT>(g.xp, x, tx, g.idx, j);
    const double fy = stage_frac<EXACT>(g.y, y, ty, g.idy, i);
    const uint32_t k = (uint32_t)((j - 1) * R + (i - 1));
    const bool down = fy >= 0.5;
    const uint32_t kc = down ? k + 1 : k;
    const double a0 = v[k], a1 = v[k + 1], b0 = v[k + R], b1 = v[k + R + 1];
    const double d0 = D[kc], d1 = D[kc + R];
    const double t = fy - 0.5, h = 0.5 * (t * t);
    const double v12 = a0 + fy * (a1 - a0) + h * d0;
    const double v34 = b0 + fy * (b1 - b0) + h * d1;
    return v12 + fx * (v34 - v12);
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
    // second differences of the stage velocities (k_rk4_prep): vx, vxf, the
    // fluid loop's left-half first row (centre from vx), vy, vyf
    const double *DX, *DXf, *DXq, *DY, *DYf;
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
        vxr[rk] = 0.0 + rk_vx<EXACT>(g, A.vx, A.DX, A.DX, xr, yr);
        vyr[rk] = 0.0 + rk_vy<EXACT>(g, A.vy, A.DY, xr, yr);
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
        vxr[rk] = 0.0 + rk_vx<EXACT>(g, A.vxf, A.DXf, A.DXq, xr, yr);
        vyr[rk] = 0.0 + rk_vy<EXACT>(g, A.vyf, A.DYf, xr, yr);
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

constexpr int ADJ_WIN = 2048;  // LDS words of the workgroup pre-aggregation (2 per node)

struct FixWin {
    unsigned long long v[ADJ_WIN];
    int box[4];  // imin, imax, jmin, jmax (1-based cells)
};

// Workgroup-wide F[stencil of (i, j)] += w * v (bilinear, `rows` rows) in
// fixed point (ereb_fixsum.h), so the sums do not depend on the order of
// the atomics.  PASS 0 records the largest |w v| (one atomic per wavefront).
// PASS 1 sums the contributions' words in an LDS window over the
// workgroup's cells (integer LDS atomics), then adds each touched node's
// words to F with one atomic pair; when the cells span more than the
// window, every lane adds straight to F.  Every thread of the workgroup
// must call it (PASS 1 synchronises); `on` marks the lanes that contribute.
template <int PASS>
__device__ void block_fix_add4(const FixGrid& F, int rows, bool on, int i, int j,
                               const double* w, double v, FixWin& W, int* flag) {
    on = on && v != 0.0;
    if (PASS == 0) {
        double mx = 0.0;
        if (on)
            for (int c = 0; c < 4; ++c) mx = fmax(mx, fabs(w[c] * v));
        for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
        if (__lane_id() == 0 && mx > 0.0)
            atomicMax(F.maxbits, (unsigned long long)__double_as_longlong(mx));
        return;
    }
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
    if (2 * (int64_t)h * wdt > ADJ_WIN) {
        if (on) {
            const int64_t k = (int64_t)(j - 1) * rows + (i - 1);
            fix_add<1>(F, k, w[0] * v, flag);
            fix_add<1>(F, k + 1, w[1] * v, flag);
            fix_add<1>(F, k + rows, w[2] * v, flag);
            fix_add<1>(F, k + rows + 1, w[3] * v, flag);
        }
        return;
    }
    for (int q = threadIdx.x; q < 2 * h * wdt; q += blockDim.x) W.v[q] = 0ull;
    __syncthreads();
    if (on) {
        const int q0 = (j - j0) * h + (i - i0);
        const int qs[4] = {q0, q0 + 1, q0 + h, q0 + h + 1};
        for (int c = 0; c < 4; ++c) {
            unsigned long long H, L;
            if (!fix_split(w[c] * v, F.up, H, L)) {
                atomicOr(flag, 1);
                continue;
            }
            atomicAdd(&W.v[2 * qs[c]], H);
            atomicAdd(&W.v[2 * qs[c] + 1], L);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < h * wdt; q += blockDim.x) {
        const unsigned long long H = W.v[2 * q], L = W.v[2 * q + 1];
        if (H | L) {
            const int64_t k = (int64_t)(j0 - 1 + q / h) * rows + (i0 - 1 + q % h);
            atomicAdd(&F.w[2 * k], H);
            atomicAdd(&F.w[2 * k + 1], L);
        }
    }
    __syncthreads();  // the window is reused by the next call
}

// RK4^T of the temperature correction E:5243-5252: given the cotangent gout of
// tkm after the step, the cotangent of tkm before it (gin), the increment of
// the phim cotangent (PASS 1) and the tk2 cotangent (both stencils, in fixed
// point).  Positions and stresses do not depend on tkm, phim or tk2; the
// fluid back-trace end points (xb, yb) were recorded by the forward step.
template <int PASS>
__global__ __launch_bounds__(256) void k_rk4_adj(RK4Args A, const double* __restrict__ xbs,
                                                 const double* __restrict__ ybs,
                                                 const double* __restrict__ gout,
                                                 double* __restrict__ gin,
                                                 double* __restrict__ gphi, FixGrid gtk2,
                                                 int* flag) {
    __shared__ FixWin W;
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
    if (on && PASS == 1) {
        gin[m] = go * F_T;
        gphi[m] += go * F_p;
    }
    const double gd = go * F_d;
    block_fix_add4<PASS>(gtk2, g.Ny1, on, ipb, jpb, wb, gd, W, flag);
    block_fix_add4<PASS>(gtk2, g.Ny1, on, ips, jps, ws, -gd, W, flag);
}

#pragma clang fp contract(off)

// gather^T of update_marker_temperature! (E:4782-4801): the marker cotangent
// goes to the interpolated grid (tk2 at timestep 1, DT later); at timestep 1
// the marker value was overwritten, so its cotangent becomes zero.
template <int PASS>
__global__ __launch_bounds__(256) void k_gather_T_adj(GeomK g, const double* __restrict__ xm,
                                                      const double* __restrict__ ym,
                                                      double* __restrict__ gT, FixGrid ggrid,
                                                      int* flag, int64_t M, int first) {
    __shared__ FixWin W;
    const int64_t m0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = m0 < M;
    const int64_t m = on ? m0 : M - 1;
    const double x = xm[m], y = ym[m];
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    double w[4];
    weights4((x - g.xp[jp - 1]) / g.dx, (y - g.yp[ip - 1]) / g.dy, w);
    const double v = gT[m];
    block_fix_add4<PASS>(ggrid, g.Ny1, on, ip, jp, w, v, W, flag);
    if (PASS == 1 && on && first) gT[m] = 0.0;
}

int launch_gather_T_adj(ereb_ctx* c, int pass, int64_t timestep, double* gT, const FixGrid& F,
                        int* flag) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    auto kern = pass == 0 ? k_gather_T_adj<0> : k_gather_T_adj<1>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                       geomk(c), (const double*)mptr(c, EREB_M_XM),
                       (const double*)mptr(c, EREB_M_YM), gT, F, flag, M,
                       timestep == 1 ? 1 : 0);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

static RK4Args rk4_args(ereb_ctx* c, int mode);

int launch_rk4_adj(ereb_ctx* c, int pass, double dt, int mode, const double* xb,
                   const double* yb, const double* gout, double* gin, double* gphi,
                   const FixGrid& gtk2, int* flag) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    int rc;
    if ((rc = set_dt(c, dt))) return rc;
    RK4Args A = rk4_args(c, mode);
    auto kern = pass == 0 ? k_rk4_adj<0> : k_rk4_adj<1>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream, A, xb,
                       yb, gout, gin, gphi, gtk2, flag);
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
    const size_t ns = (size_t)c->Nx1 * c->Ny1;
    A.DX = c->d_rkd;
    A.DXf = c->d_rkd + ns;
    A.DXq = c->d_rkd + 2 * ns;
    A.DY = c->d_rkd + 3 * ns;
    A.DYf = c->d_rkd + 4 * ns;
    A.xb_out = A.yb_out = nullptr;
    A.DSXX = A.DSXY = A.DTg = nullptr;
    return A;
}

// the second-difference grids of this step's velocities (k_rk4_prep)
static void launch_rk4_prep(ereb_ctx* c, const RK4Args& A) {
    const int n = c->Nx1 * c->Ny1;
    hipLaunchKernelGGL(k_rk4_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->Nx, c->Ny, A.vx, A.vy, A.vxf, A.vyf, (double*)A.DX, (double*)A.DXf,
                       (double*)A.DXq, (double*)A.DY, (double*)A.DYf);
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
    launch_rk4_prep(c, A);
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
    launch_rk4_prep(c, A);
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
