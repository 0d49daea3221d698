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
    const uint32_t o = goff(i, j, rows), oR = o + ((uint32_t)rows << 3);
    return (R)ldg_at(gr, o) * w[0] + (R)ldg_at(gr, o + 8) * w[1] + (R)ldg_at(gr, oR) * w[2] +
           (R)ldg_at(gr, oR + 8) * w[3];
}

template <class R>
__device__ __forceinline__ void weights4_r(double a, double b, R* w) {
    const R ar = (R)a, br = (R)b;
    w[0] = (R(1) - br) * (R(1) - ar);
    w[1] = br * (R(1) - ar);
    w[2] = (R(1) - br) * ar;
    w[3] = br * ar;
}

// Bilinear stencil of a marker: byte offset of its cell's first node and
// the four weights.  The RK4 kernel reuses the gathers' stencils for the
// start-point interpolations of tk2 (P cell, clamps 1..Nx, 1..Ny) and wyx
// (basic cell, clamps 1..Nx-1, 1..Ny-1), which use the same cells.
template <class R>
struct Stencil4 {
    uint32_t off;
    R w[4];
};

// update_marker_temperature! (E:4782-4801) for one marker: the new tkm; P
// stencil (clamps 1..Nx, 1..Ny) in *st
template <class R, bool EXACT>
__device__ __forceinline__ double gather_T_one(const GeomK& g, double x, double y, double T,
                                               const double* __restrict__ grid, bool first,
                                               Stencil4<R>* st = nullptr) {
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    R w[4];
    weights4_r<R>((x - axis_node<EXACT>(g.xp, g.dx, jp)) / g.dx,
                  (y - axis_node<EXACT>(g.yp, g.dy, ip)) / g.dy, w);
    const double v = (double)dot4_r<R>(grid, g.Ny1, ip, jp, w);
    if (st) {
        st->off = goff(ip, jp, g.Ny1);
        for (int c = 0; c < 4; ++c) st->w[c] = w[c];
    }
    return first ? v : T + v;
}

// update_marker_stress! (E:4354-4388) for one marker: sxxm, sxym updated in
// place; P grid with the clamps 2..Nx-1, 2..Ny-1 of E:4364-4367; basic
// stencil (clamps 1..Nx-1, 1..Ny-1) in *st
template <class R, bool EXACT>
__device__ __forceinline__ void gather_S_one(const GeomK& g, double x, double y, double& sxx,
                                             double& sxy, const double* __restrict__ DSXX,
                                             const double* __restrict__ DSXY,
                                             Stencil4<R>* st = nullptr) {
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
    if (st) {
        st->off = goff(ib, jb, g.Ny);
        for (int c = 0; c < 4; ++c) st->w[c] = wb[c];
    }
}

// The RK4 kernel's start-point stencils, computed as the Float64 gathers
// compute theirs (so every path of the kernel gets the same values): P cell
// with the clamps 1..Nx, 1..Ny (the temperature gather's) and basic cell
// with the clamps 1..Nx-1, 1..Ny-1 (the stress gather's DSXY stencil).
template <bool EXACT>
__device__ __forceinline__ void start_stencil_p(const GeomK& g, double x, double y,
                                                Stencil4<double>& st) {
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    weights4_r<double>((x - axis_node<EXACT>(g.xp, g.dx, jp)) / g.dx,
                       (y - axis_node<EXACT>(g.yp, g.dy, ip)) / g.dy, st.w);
    st.off = goff(ip, jp, g.Ny1);
}

template <bool EXACT>
__device__ __forceinline__ void start_stencil_b(const GeomK& g, double x, double y,
                                                Stencil4<double>& st) {
    const int jb = jl_cell((x - g.x[0]) * g.idx, 1, g.Nx - 1);
    const int ib = jl_cell((y - g.y[0]) * g.idy, 1, g.Ny - 1);
    weights4_r<double>((x - axis_node<EXACT>(g.x, g.dx, jb)) / g.dx,
                       (y - axis_node<EXACT>(g.y, g.dy, ib)) / g.dy, st.w);
    st.off = goff(ib, jb, g.Ny);
}

// dot4 of a grid with a stencil (E:1137-1159), products and sums rounded one
// by one, so every kernel that calls it gets the same bits
__device__ __forceinline__ double stencil_dot(const double* __restrict__ gr, int rows,
                                              const Stencil4<double>& st) {
    const uint32_t oR = st.off + ((uint32_t)rows << 3);
    return ldg_at(gr, st.off) * st.w[0] + ldg_at(gr, st.off + 8) * st.w[1] +
           ldg_at(gr, oR) * st.w[2] + ldg_at(gr, oR + 8) * st.w[3];
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

// RK4 stage velocities as quadratics, prepared once per step (k_rk4_coef).
// On the Vx grid (E:5057-5091) the stage velocity in cell (i, j) is
//   v = v_i(fx) + fy (v_{i+1}(fx) - v_i(fx)),
//   v_r(fx) = a_r + fx (b_r - a_r) + 0.5 (fx - 0.5)^2 d_r,
// a_r = v[r, j], b_r = v[r, j+1], and d_r the second difference
// v[r, c-1] - 2 v[r, c] + v[r, c+1] (the reference's association) at the
// correction's centre column c = j for fx < 0.5 ("left" half), c = j + 1
// otherwise, zero where the reference skips the correction (c == 1 left,
// c == Nx right, i.e. outside 2 <= c <= Nx-1).  In the fluid loop the
// centre of the first row's left-half correction is read from vx (E:5173).
// Expanding 0.5 (fx - 0.5)^2 = 0.125 - 0.5 fx + 0.5 fx^2, each half of each
// cell is v = P(fx) + fy Q(fx) with
//   P = (a_i + d_i/8) + fx ((b_i - a_i) - d_i/2) + fx^2 (d_i/2),
//   Q = the same of row i+1 minus P,
// six coefficients [P0 P1 P2 Q0 Q1 Q2]; the Vy grid (E:5093-5127) is the
// transpose (quadratic in fy, linear in fx).  Table layout: per cell of the
// (Ny1, Nx1) node numbering, 12 doubles (left half, then right half), so a
// stage reads its half with three 16-byte loads; four tables (Vx and Vy of
// the solid and of the fluid loop).
constexpr int RKC = 12;  // doubles per cell and table

struct RKCoefArgs {
    int Nx, Ny;
    const double *vx, *vy, *vxf, *vyf;
    double *cxs, *cys, *cxf, *cyf;  // tables: Vx solid, Vy solid, Vx fluid, Vy fluid
};

// second difference of g along a line at index k with stride s, if the
// centre's position c lies in [2, n-1] (else the correction is skipped)
__device__ __forceinline__ double d2(const double* __restrict__ q, const double* __restrict__ ctr,
                                     int k, int s, int c, int n) {
    return (c >= 2 && c <= n - 1) ? (q[k - s] - 2.0 * ctr[k]) + q[k + s] : 0.0;
}

// coefficients of one half: rows (a0, b0, d0) and (a1, b1, d1) along the
// quadratic direction
__device__ __forceinline__ void rk_half(double* __restrict__ out, double a0, double b0,
                                        double d0, double a1, double b1, double d1) {
    const double p0 = a0 + 0.125 * d0, p1 = (b0 - a0) - 0.5 * d0, p2 = 0.5 * d0;
    const double r0 = a1 + 0.125 * d1, r1 = (b1 - a1) - 0.5 * d1, r2 = 0.5 * d1;
    out[0] = p0;
    out[1] = p1;
    out[2] = p2;
    out[3] = r0 - p0;
    out[4] = r1 - p1;
    out[5] = r2 - p2;
}

// One thread per cell of one table (grid y = table t), cells in table order
// (k = (j - 1) Ny1 + (i - 1)): the thread loads the cell's velocity
// neighbourhood once and evaluates both halves; the workgroup's 256 cells
// (24 KB of coefficients) are assembled in LDS and written with consecutive
// 16-byte stores.
constexpr int RKCOEF_CELLS = 256;

__global__ __launch_bounds__(RKCOEF_CELLS) void k_rk4_coef(RKCoefArgs A) {
    __shared__ double sh[RKCOEF_CELLS * RKC];
    const int R = A.Ny + 1;
    const int N = R * (A.Nx + 1);
    const int b0 = blockIdx.x * RKCOEF_CELLS;
    const int k = b0 + (int)threadIdx.x;
    const int t = blockIdx.y;  // 0 Vx solid, 1 Vy solid, 2 Vx fluid, 3 Vy fluid
    const bool xg = (t & 1) == 0, fluid = t >= 2;
    double c[RKC];
#pragma unroll
    for (int q = 0; q < RKC; ++q) c[q] = 0.0;
    if (k < N) {
        const int i = k % R + 1, j = k / R + 1;
        const int k1 = k + 1, kR = k + R, kR1 = k + R + 1;
        if (xg && i <= A.Ny && j <= A.Nx - 1) {
            // Vx cell (i, j): rows i, i+1 of columns j, j+1; centre column j
            // (left half) or j+1 (right half); the fluid loop's first-row
            // left-half centre is read from vx (E:5173)
            const double* v = fluid ? A.vxf : A.vx;
            const double a0 = v[k], b0v = v[kR], a1 = v[k1], b1 = v[kR1];
            rk_half(c, a0, b0v, d2(v, fluid ? A.vx : v, k, R, j, A.Nx), a1, b1,
                    d2(v, v, k1, R, j, A.Nx));
            rk_half(c + 6, a0, b0v, d2(v, v, kR, R, j + 1, A.Nx), a1, b1,
                    d2(v, v, kR1, R, j + 1, A.Nx));
        } else if (!xg && i <= A.Ny - 1 && j <= A.Nx) {
            // Vy cell (i, j): rows i, i+1 of columns j, j+1; centre row i
            // (upper half) or i+1 (lower half)
            const double* v = fluid ? A.vyf : A.vy;
            const double a0 = v[k], b0v = v[k1], a1 = v[kR], b1 = v[kR1];
            rk_half(c, a0, b0v, d2(v, v, k, 1, i, A.Ny), a1, b1, d2(v, v, kR, 1, i, A.Ny));
            rk_half(c + 6, a0, b0v, d2(v, v, k1, 1, i + 1, A.Ny), a1, b1,
                    d2(v, v, kR1, 1, i + 1, A.Ny));
        }
    }
#pragma unroll
    for (int q = 0; q < RKC; q += 2)
        *(double2*)(sh + threadIdx.x * RKC + q) = make_double2(c[q], c[q + 1]);
    __syncthreads();
    double* out = t == 0 ? A.cxs : (t == 1 ? A.cys : (t == 2 ? A.cxf : A.cyf));
    const int npiece = min(RKCOEF_CELLS, N - b0) * (RKC / 2);
    for (int p = threadIdx.x; p < npiece; p += RKCOEF_CELLS)
        *(double2*)(out + (size_t)b0 * RKC + 2 * p) = *(const double2*)(sh + 2 * p);
}

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

// P(u) + w Q(u) of one half's coefficients [P0 P1 P2 Q0 Q1 Q2] in Horner form.
// The RK4 code fuses multiply-adds only where it says so (__builtin_fma; the
// library is built without contraction), so its rounding does not depend on
// the kernel it is inlined into.
__device__ __forceinline__ double rk_eval(double2 c01, double2 c23, double2 c45, double u,
                                          double w) {
    const double P = __builtin_fma(u, __builtin_fma(u, c23.x, c01.y), c01.x);
    const double Q = __builtin_fma(u, __builtin_fma(u, c45.y, c45.x), c23.y);
    return __builtin_fma(w, Q, P);
}

// The last half-cell of one table a thread looked up: consecutive RK4
// stages of a marker mostly fall in the same half-cell, which then costs no
// load (key = 2 node + half; ~0u = none yet).
struct RkCache {
    uint32_t key;
    double2 c01, c23, c45;
};

// The stage velocity of the half `second` of the cell whose coefficients start
// at node index `node` of table T
__device__ __forceinline__ double rk_poly(const double* __restrict__ T, uint32_t node,
                                          bool second, double u, double w, RkCache& C) {
    const uint32_t key = (node << 1) | (second ? 1u : 0u);
    if (key != C.key) {
        const uint32_t off = node * (uint32_t)(RKC * 8) + (second ? 48u : 0u);  // < 2^31 bytes
        const char* p = (const char*)T + off;
        C.key = key;
        C.c01 = *(const double2*)p;
        C.c23 = *(const double2*)(p + 16);
        C.c45 = *(const double2*)(p + 32);
    }
    return rk_eval(C.c01, C.c23, C.c45, u, w);
}

// RK4 stage velocity on the Vx grid (E:5057-5091) from the cell's quadratic
// coefficients (k_rk4_coef): cell (i, j) with the clamps 1..Ny, 1..Nx-1,
// fractions fx, fy, the half chosen by fx >= 0.5, then
// P(fx) + fy Q(fx).  The Vy grid (E:5093-5127) is the transpose.
template <bool EXACT>
__device__ __forceinline__ double rk_vx(const GeomK& g, const double* __restrict__ T, double x,
                                        double y, RkCache& C) {
    const double tx = (x - g.x[0]) * g.idx, ty = (y - g.yp[0]) * g.idy;
    const int j = stage_cell(tx, 1, g.Nx - 1);
    const int i = stage_cell(ty, 1, g.Ny);
    const double fx = stage_frac<EXACT>(g.x, x, tx, g.idx, j);
    const double fy = stage_frac<EXACT>(g.yp, y, ty, g.idy, i);
    return rk_poly(T, (uint32_t)((j - 1) * g.Ny1 + (i - 1)), fx >= 0.5, fx, fy, C);
}

template <bool EXACT>
__device__ __forceinline__ double rk_vy(const GeomK& g, const double* __restrict__ T, double x,
                                        double y, RkCache& C) {
    const double tx = (x - g.xp[0]) * g.idx, ty = (y - g.y[0]) * g.idy;
    const int j = stage_cell(tx, 1, g.Nx);
    const int i = stage_cell(ty, 1, g.Ny - 1);
    const double fx = stage_frac<EXACT>(g.xp, x, tx, g.idx, j);
    const double fy = stage_frac<EXACT>(g.y, y, ty, g.idy, i);
    return rk_poly(T, (uint32_t)((j - 1) * g.Ny1 + (i - 1)), fy >= 0.5, fy, fx, C);
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
    // quadratic coefficients of the stage velocities (k_rk4_coef): Vx and Vy
    // of the solid step, Vx and Vy of the fluid back-trace
    const double *CXs, *CYs, *CXf, *CYf;
    double *xb_out, *yb_out;  // optional: fluid back-trace end points (adjoint recording)
    // fused gathers (k_rk4<1 or 2>): DSXX, DSXY and the DT grid (the
    // temperature gather reads tk2 at timestep 1, else DT)
    const double *DSXX, *DSXY, *DTg;
};

// Marker paths of one RK4 step: the solid step (E:5055-5136) from (x0, y0)
// to (xn, yn) and the fluid back-trace (E:5138-5226) from (xn, yn) to
// (xb, yb).  The weighted stage sum b0 v0 + b1 v1 + b2 v2 + b3 v3 (dot4,
// E:5135-5136) is accumulated as the stages come, left to right with fmas,
// so no stage velocity stays live.
template <bool EXACT>
__device__ __forceinline__ void rk4_paths(const RK4Args& A, double dt, double x0, double y0,
                                          double& xn, double& yn, double& xb, double& yb) {
    const GeomK& g = A.g;
    const double* b = A.M.brk4;
    const double* cr = A.M.crk4;
    double xr = x0, yr = y0, sx = 0.0, sy = 0.0;
    RkCache cx{~0u, {}, {}, {}}, cy{~0u, {}, {}, {}};
#pragma unroll
    for (int rk = 0; rk < 4; ++rk) {
        const double vx = rk_vx<EXACT>(g, A.CXs, xr, yr, cx);
        const double vy = rk_vy<EXACT>(g, A.CYs, xr, yr, cy);
        sx = rk == 0 ? b[0] * vx : __builtin_fma(b[rk], vx, sx);
        sy = rk == 0 ? b[0] * vy : __builtin_fma(b[rk], vy, sy);
        if (rk < 3) {
            xr = __builtin_fma(dt * cr[rk], vx, x0);
            yr = __builtin_fma(dt * cr[rk], vy, y0);
        }
    }
    xn = __builtin_fma(dt, sx, x0);
    yn = __builtin_fma(dt, sy, y0);
    xr = xn;
    yr = yn;
    cx.key = cy.key = ~0u;  // the fluid tables
#pragma unroll
    for (int rk = 0; rk < 4; ++rk) {
        const double vx = rk_vx<EXACT>(g, A.CXf, xr, yr, cx);
        const double vy = rk_vy<EXACT>(g, A.CYf, xr, yr, cy);
        sx = rk == 0 ? b[0] * vx : __builtin_fma(b[rk], vx, sx);
        sy = rk == 0 ? b[0] * vy : __builtin_fma(b[rk], vy, sy);
        if (rk < 3) {
            xr = __builtin_fma(-dt * cr[rk], vx, xn);
            yr = __builtin_fma(-dt * cr[rk], vy, yn);
        }
    }
    xb = __builtin_fma(-dt, sx, xn);
    yb = __builtin_fma(-dt, sy, yn);
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

// GATHER = 0: RK4 alone.  GATHER = 1 (Float64) or 2 (the mixed form: the
// stress gather's interpolation in FP32; the temperature gather is Float64 in
// both forms): the stress gather, then the temperature gather, then RK4 on
// the gathered values, as the three separate kernels would compute them
// (the same helpers, no contraction), with each marker's state read and
// written once.
template <int GATHER, bool EXACT>
__device__ __forceinline__ void rk4_marker(const RK4Args& A, int64_t m) {
    const GeomK& g = A.g;
    const double dt = A.S->dt;
    const bool first = GATHER && A.S->first != 0;
    const double* tgrid = first ? A.tk2 : A.DTg;
    const double xmm = A.xm[m], ymm = A.ym[m];
    double s = A.sxxm[m], t = A.sxym[m];
    double T = first ? 0.0 : A.tkm[m];
    double xr, yr;
    double w[4];
    int jp, ip;
    // solid temperature at the start position (P grid) and rotation rate
    // (basic grid), E:5013-5040, on the stencils of the Float64 gathers (the
    // temperature gather's P stencil, the stress gather's basic one),
    // computed the same way when there is no gather, so that the fused and
    // the separate kernels give the same values
    Stencil4<double> sp, sb;
    if (GATHER)
        T = gather_T_one<double, EXACT>(g, xmm, ymm, T, tgrid, first, &sp);
    else
        start_stencil_p<EXACT>(g, xmm, ymm, sp);
    if (GATHER == 1) {
        gather_S_one<double, EXACT>(g, xmm, ymm, s, t, A.DSXX, A.DSXY, &sb);
    } else {
        if (GATHER == 2) gather_S_one<float, EXACT>(g, xmm, ymm, s, t, A.DSXX, A.DSXY);
        start_stencil_b<EXACT>(g, xmm, ymm, sb);
    }
    const double tks0 = stencil_dot(A.tk2, g.Ny1, sp);
    const double om = stencil_dot(A.wyx, g.Ny, sb);
    // analytic stress rotation E:5042-5048
    const double th = dt * om;
    // sin(2θ) and cos(2θ) from the double-angle identities (within a few ulp
    // of the library values)
    double ct, st;
    if (fabs(th) <= 0.78539816339744828) {
        // Taylor series to the 17th (sine) and 16th (cosine) power, Horner in
        // th^2: the first omitted terms are below 2^-60 of the result for
        // |th| <= pi/4; beyond that the library sincos
        const double t2 = th * th;
        double ps = 1.0 / 355687428096000.0, pc = 1.0 / 6402373705728000.0;
        ps = fma(ps, -t2, 1.0 / 1307674368000.0);
        pc = fma(pc, -t2, 1.0 / 20922789888000.0);
        ps = fma(ps, -t2, 1.0 / 6227020800.0);
        pc = fma(pc, -t2, 1.0 / 87178291200.0);
        ps = fma(ps, -t2, 1.0 / 39916800.0);
        pc = fma(pc, -t2, 1.0 / 479001600.0);
        ps = fma(ps, -t2, 1.0 / 362880.0);
        pc = fma(pc, -t2, 1.0 / 3628800.0);
        ps = fma(ps, -t2, 1.0 / 5040.0);
        pc = fma(pc, -t2, 1.0 / 40320.0);
        ps = fma(ps, -t2, 1.0 / 120.0);
        pc = fma(pc, -t2, 1.0 / 720.0);
        ps = fma(ps, -t2, 1.0 / 6.0);
        pc = fma(pc, -t2, 1.0 / 24.0);
        pc = fma(pc, -t2, 0.5);
        st = fma(-ps * t2, th, th);
        ct = fma(-pc, t2, 1.0);
    } else {
        sincos(th, &st, &ct);
    }
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

// One thread per marker in storage order (coalesced marker reads and writes),
// the stage tables read from global memory.
template <int GATHER, bool EXACT>
__global__ __launch_bounds__(256) void k_rk4(RK4Args A) {
    // XCD-contiguous block order: each XCD moves one band of the domain, whose
    // coefficient-table lines then stay in that XCD's L2 instead of being
    // fetched by all eight
    const int64_t m = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (m >= A.n) return;
    rk4_marker<GATHER, EXACT>(A, m);
}

// ---- adjoints (with respect to tkm, phim and the tk2 grid) ---------------

constexpr int ADJ_WIN = 2048;  // LDS words of the workgroup pre-aggregation (2 per node)

struct FixWin {
    unsigned long long v[ADJ_WIN];
    int box[4];  // imin, imax, jmin, jmax (1-based cells)
};

// Workgroup-wide F[stencil s of (i_s, j_s)] += w_s * v_s for NS stencils
// per lane (bilinear, `rows` rows) in fixed point (ereb_fixsum.h), so the
// sums do not depend on the order of the atomics.  PASS 0 records the
// largest |w v| (one atomic per wavefront).  PASS 1 sums the contributions'
// words in an LDS window over the workgroup's cells (integer LDS atomics),
// then adds each touched node's words to F with one atomic pair (device-wide
// atomics are carried out at the memory side on this chip, so their number
// sets the cost: one window for all of a lane's stencils); when the cells
// span more than the window, every lane adds straight to F.  Every thread of
// the workgroup must call it (PASS 1 synchronises); `on` marks the lanes
// that contribute.
template <int PASS, int NS>
__device__ void block_fix_add(const FixGrid& F, int rows, bool on, const int (&is)[NS],
                              const int (&js)[NS], const double (&ws)[NS][4],
                              const double (&vs)[NS], FixWin& W, int* flag) {
    if (PASS == 0) {
        double mx = 0.0;
        if (on)
            for (int q = 0; q < NS; ++q)
                for (int c = 0; c < 4; ++c) mx = fmax(mx, fabs(ws[q][c] * vs[q]));
        for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
        if (__lane_id() == 0 && mx > 0.0)
            atomicMax(F.maxbits, (unsigned long long)__double_as_longlong(mx));
        return;
    }
    bool any = false;
    int imin = INT32_MAX, imax = INT32_MIN, jmin = INT32_MAX, jmax = INT32_MIN;
    for (int q = 0; q < NS; ++q) {
        if (on && vs[q] != 0.0) {
            any = true;
            imin = min(imin, is[q]);
            imax = max(imax, is[q]);
            jmin = min(jmin, js[q]);
            jmax = max(jmax, js[q]);
        }
    }
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
        if (any) {
            for (int q = 0; q < NS; ++q) {
                if (vs[q] == 0.0) continue;
                const int64_t k = (int64_t)(js[q] - 1) * rows + (is[q] - 1);
                fix_add<1>(F, k, ws[q][0] * vs[q], flag);
                fix_add<1>(F, k + 1, ws[q][1] * vs[q], flag);
                fix_add<1>(F, k + rows, ws[q][2] * vs[q], flag);
                fix_add<1>(F, k + rows + 1, ws[q][3] * vs[q], flag);
            }
        }
        return;
    }
    for (int q = threadIdx.x; q < 2 * h * wdt; q += blockDim.x) W.v[q] = 0ull;
    __syncthreads();
    if (any) {
        for (int q = 0; q < NS; ++q) {
            if (vs[q] == 0.0) continue;
            const int q0 = (js[q] - j0) * h + (is[q] - i0);
            const int qs[4] = {q0, q0 + 1, q0 + h, q0 + h + 1};
            for (int c = 0; c < 4; ++c) {
                unsigned long long H, L;
                if (!fix_split(ws[q][c] * vs[q], F.up, H, L)) {
                    atomicOr(flag, 1);
                    continue;
                }
                atomicAdd(&W.v[2 * qs[c]], H);
                atomicAdd(&W.v[2 * qs[c] + 1], L);
            }
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
                                                 int* flag, const uint32_t* __restrict__ order) {
    __shared__ FixWin W;
    gtk2.up = fix_up(gtk2);
    const int64_t m0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = m0 < A.n;
    // idle lanes read a valid marker and add nothing
    const int64_t m = on ? (order ? (int64_t)order[m0] : m0) : A.n - 1;
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
    // both stencils (back-trace end: +gd, start: -gd) in one window
    const int is2[2] = {ipb, ips}, js2[2] = {jpb, jps};
    const double ws2[2][4] = {{wb[0], wb[1], wb[2], wb[3]}, {ws[0], ws[1], ws[2], ws[3]}};
    const double vs2[2] = {gd, -gd};
    block_fix_add<PASS, 2>(gtk2, g.Ny1, on, is2, js2, ws2, vs2, W, flag);
}

// gather^T of update_marker_temperature! (E:4782-4801): the marker cotangent
// goes to the interpolated grid (tk2 at timestep 1, DT later); at timestep 1
// the marker value was overwritten, so its cotangent becomes zero.
template <int PASS>
__global__ __launch_bounds__(256) void k_gather_T_adj(GeomK g, const double* __restrict__ xm,
                                                      const double* __restrict__ ym,
                                                      double* __restrict__ gT, FixGrid ggrid,
                                                      int* flag, int64_t M, int first,
                                                      const uint32_t* __restrict__ order) {
    __shared__ FixWin W;
    ggrid.up = fix_up(ggrid);
    const int64_t m0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = m0 < M;
    const int64_t m = on ? (order ? (int64_t)order[m0] : m0) : M - 1;
    const double x = xm[m], y = ym[m];
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    double w[4];
    weights4((x - g.xp[jp - 1]) / g.dx, (y - g.yp[ip - 1]) / g.dy, w);
    const double v = gT[m];
    const int is1[1] = {ip}, js1[1] = {jp};
    const double ws1[1][4] = {{w[0], w[1], w[2], w[3]}};
    const double vs1[1] = {v};
    block_fix_add<PASS, 1>(ggrid, g.Ny1, on, is1, js1, ws1, vs1, W, flag);
    if (PASS == 1 && on && first) gT[m] = 0.0;
}

int launch_gather_T_adj(ereb_ctx* c, int pass, int64_t timestep, double* gT, const FixGrid& F,
                        int* flag, const uint32_t* order) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    auto kern = pass == 0 ? k_gather_T_adj<0> : k_gather_T_adj<1>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                       geomk(c), (const double*)mptr(c, EREB_M_XM),
                       (const double*)mptr(c, EREB_M_YM), gT, F, flag, M,
                       timestep == 1 ? 1 : 0, order);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

static RK4Args rk4_args(ereb_ctx* c, int mode);

int launch_rk4_adj(ereb_ctx* c, int pass, double dt, int mode, const double* xb,
                   const double* yb, const double* gout, double* gin, double* gphi,
                   const FixGrid& gtk2, int* flag, const uint32_t* order) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    int rc;
    if ((rc = set_dt(c, dt))) return rc;
    RK4Args A = rk4_args(c, mode);
    auto kern = pass == 0 ? k_rk4_adj<0> : k_rk4_adj<1>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream, A, xb,
                       yb, gout, gin, gphi, gtk2, flag, order);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

int launch_gather_temperature(ereb_ctx* c, int64_t timestep) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    hipEvent_t ev;
    prof_begin(c, EREB_K_GATHER_T, &ev);
    const double* grid = gptr(c, timestep == 1 ? EREB_G_TK2 : EREB_G_DT);
    // Float64 in both forms: the property laws switch branch at fixed
    // temperatures (E:2272-2358), and an FP32 interpolation error (~3e-5 K at
    // 300 K) would move markers near a breakpoint to the other branch, a
    // jump of up to 16x in the fluid heat capacity
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
    const size_t ns = (size_t)c->Nx1 * c->Ny1 * RKC;
    A.CXs = c->d_rkd;
    A.CYs = c->d_rkd + ns;
    A.CXf = c->d_rkd + 2 * ns;
    A.CYf = c->d_rkd + 3 * ns;
    A.xb_out = A.yb_out = nullptr;
    A.DSXX = A.DSXY = A.DTg = nullptr;
    return A;
}

// the stage-velocity coefficient tables of this step's velocities (k_rk4_coef)
static void launch_rk4_prep(ereb_ctx* c, const RK4Args& A) {
    RKCoefArgs P{c->Nx, c->Ny, A.vx, A.vy, A.vxf, A.vyf, (double*)A.CXs, (double*)A.CYs,
                 (double*)A.CXf, (double*)A.CYf};
    const int64_t N = (int64_t)c->Nx1 * c->Ny1;
    hipLaunchKernelGGL(k_rk4_coef, dim3((unsigned)((N + RKCOEF_CELLS - 1) / RKCOEF_CELLS), 4),
                       dim3(RKCOEF_CELLS), 0, c->stream, P);
}

template <int GATHER>
static void launch_rk4_kernel(ereb_ctx* c, const RK4Args& A) {
    const dim3 grid((unsigned)((c->marknum + 255) / 256));
    if (c->axes_exact)
        hipLaunchKernelGGL((k_rk4<GATHER, true>), grid, dim3(256), 0, c->stream, A);
    else
        hipLaunchKernelGGL((k_rk4<GATHER, false>), grid, dim3(256), 0, c->stream, A);
}

int launch_rk4(ereb_ctx* c, double dt, int mode, double* xb_out, double* yb_out) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    int rc;
    if ((rc = set_dt(c, dt))) return rc;
    RK4Args A = rk4_args(c, mode);
    A.xb_out = xb_out;
    A.yb_out = yb_out;
    hipEvent_t ev;
    prof_begin(c, EREB_K_RK4, &ev);
    launch_rk4_prep(c, A);
    launch_rk4_kernel<0>(c, A);
    ++c->pos_gen;
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_RK4, ev);
    return EREB_OK;
}

int launch_gathers_rk4(ereb_ctx* c, double dt, int mode, int64_t timestep, int parts) {
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
    hipEvent_t ev;
    prof_begin(c, EREB_K_RK4, &ev);
    if (parts & RK4_TABLES) launch_rk4_prep(c, A);
    // node coordinates computed (exact axes) or loaded, chosen per kernel
    if (parts & RK4_MARKERS) {
        if (c->precision == EREB_MIXED) launch_rk4_kernel<2>(c, A);
        else launch_rk4_kernel<1>(c, A);
        ++c->pos_gen;
    }
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_RK4, ev);
    return EREB_OK;
}

}  // namespace ereb
