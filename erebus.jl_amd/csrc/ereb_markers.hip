// ereb_markers.hip -- node->marker transfers: update_marker_temperature!
// (E:4782-4801), update_marker_stress! (E:4354-4388) and RK4 advection
// move_markers_rk4! (E:4990-5257).  One thread per marker in storage order
// (coalesced SoA reads); the grids (at most a few MB each) are read through
// the L2/MALL.
#include "ereb_internal.h"

namespace ereb {

__global__ __launch_bounds__(256) void k_gather_T(GeomK g, const double* __restrict__ xm,
                                                  const double* __restrict__ ym,
                                                  double* __restrict__ tkm,
                                                  const double* __restrict__ grid, int64_t M,
                                                  int first) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const double x = xm[m], y = ym[m];
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    double w[4];
    weights4((x - g.xp[jp - 1]) / g.dx, (y - g.yp[ip - 1]) / g.dy, w);
    const double v = gdot4(grid, g.Ny1, ip, jp, w);
    if (first)
        tkm[m] = v;
    else
        tkm[m] += v;
}

__global__ __launch_bounds__(256) void k_gather_S(GeomK g, const double* __restrict__ xm,
                                                  const double* __restrict__ ym,
                                                  double* __restrict__ sxxm,
                                                  double* __restrict__ sxym,
                                                  const double* __restrict__ DSXX,
                                                  const double* __restrict__ DSXY, int64_t M) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const double x = xm[m], y = ym[m];
    // P grid with the clamps 2..Nx-1, 2..Ny-1 of E:4364-4367
    const int jp = jl_cell((x - g.xp[0]) * g.idx, 2, g.Nx - 1);
    const int ip = jl_cell((y - g.yp[0]) * g.idy, 2, g.Ny - 1);
    const int jb = jl_cell((x - g.x[0]) * g.idx, 1, g.Nx - 1);
    const int ib = jl_cell((y - g.y[0]) * g.idy, 1, g.Ny - 1);
    double wp[4], wb[4];
    weights4((x - g.xp[jp - 1]) / g.dx, (y - g.yp[ip - 1]) / g.dy, wp);
    weights4((x - g.x[jb - 1]) / g.dx, (y - g.y[ib - 1]) / g.dy, wb);
    sxxm[m] += gdot4(DSXX, g.Ny1, ip, jp, wp);
    sxym[m] += gdot4(DSXY, g.Ny, ib, jb, wb);
}

// RK4 stage velocity on the Vx grid (E:5057-5091).  `vc` supplies the
// centre value of the left-half correction: the reference reads vx there
// also in the fluid loop (E:5173).
__device__ __forceinline__ double rk_vx(const GeomK& g, const double* __restrict__ v,
                                        const double* __restrict__ vc, double x, double y) {
    const int R = g.Ny1;
    const int j = jl_cell((x - g.x[0]) * g.idx, 1, g.Nx - 1);
    const int i = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    const double dxmj = x - g.x[j - 1], dymi = y - g.yp[i - 1];
    const double dx = g.dx, dy = g.dy;
    const double* p = v + (size_t)(j - 1) * R + (i - 1);  // v[i, j]
    double v13 = p[0] * (1.0 - dxmj / dx) + p[R] * dxmj / dx;
    double v24 = p[1] * (1.0 - dxmj / dx) + p[R + 1] * dxmj / dx;
    if (dxmj / dx >= 0.5) {
        if (j < g.Nx - 1) {
            const double t = dxmj / dx - 0.5;
            v13 += 0.5 * (t * t) * (p[0] - 2.0 * p[R] + p[2 * R]);
            v24 += 0.5 * (t * t) * (p[1] - 2.0 * p[R + 1] + p[2 * R + 1]);
        }
    } else {
        if (j > 1) {
            const double t = dxmj / dx - 0.5;
            const double cc = vc[(size_t)(j - 1) * R + (i - 1)];
            v13 += 0.5 * (t * t) * (p[-R] - 2.0 * cc + p[R]);
            v24 += 0.5 * (t * t) * (p[1 - R] - 2.0 * p[1] + p[R + 1]);
        }
    }
    return v13 * (1.0 - dymi / dy) + v24 * dymi / dy;
}

// RK4 stage velocity on the Vy grid (E:5093-5127)
__device__ __forceinline__ double rk_vy(const GeomK& g, const double* __restrict__ v, double x,
                                        double y) {
    const int R = g.Ny1;
    const int j = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int i = jl_cell((y - g.y[0]) * g.idy, 1, g.Ny - 1);
    const double dxmj = x - g.xp[j - 1], dymi = y - g.y[i - 1];
    const double dx = g.dx, dy = g.dy;
    const double* p = v + (size_t)(j - 1) * R + (i - 1);
    double v12 = p[0] * (1.0 - dymi / dy) + p[1] * dymi / dy;
    double v34 = p[R] * (1.0 - dymi / dy) + p[R + 1] * dymi / dy;
    if (dymi / dy >= 0.5) {
        if (i < g.Ny - 1) {
            const double t = dymi / dy - 0.5;
            v12 += 0.5 * (t * t) * (p[0] - 2.0 * p[1] + p[2]);
            v34 += 0.5 * (t * t) * (p[R] - 2.0 * p[R + 1] + p[R + 2]);
        }
    } else {
        if (i > 1) {
            const double t = dymi / dy - 0.5;
            v12 += 0.5 * (t * t) * (p[-1] - 2.0 * p[0] + p[1]);
            v34 += 0.5 * (t * t) * (p[R - 1] - 2.0 * p[R] + p[R + 1]);
        }
    }
    return v12 * (1.0 - dxmj / dx) + v34 * dxmj / dx;
}

struct RK4Args {
    GeomK g;
    ereb_materials M;
    double dt;
    int mode;
    int64_t n;
    double *xm, *ym, *tkm, *sxxm, *sxym;
    const int32_t* tm;
    const double* phim;
    const double *vx, *vy, *vxf, *vyf, *wyx, *tk2;
};

__global__ __launch_bounds__(256) void k_rk4(RK4Args A) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= A.n) return;
    const GeomK& g = A.g;
    const double dt = A.dt;
    const double* b = A.M.brk4;
    const double* cr = A.M.crk4;
    double xmm = A.xm[m], ymm = A.ym[m];
    double xr = xmm, yr = ymm;
    double w[4];
    // solid temperature at the start position (P grid) and rotation rate
    // (basic grid), E:5013-5040
    int jp = jl_cell((xr - g.xp[0]) * g.idx, 1, g.Nx);
    int ip = jl_cell((yr - g.yp[0]) * g.idy, 1, g.Ny);
    weights4((xr - g.xp[jp - 1]) / g.dx, (yr - g.yp[ip - 1]) / g.dy, w);
    const double tks0 = gdot4(A.tk2, g.Ny1, ip, jp, w);
    const int jb = jl_cell((xr - g.x[0]) * g.idx, 1, g.Nx - 1);
    const int ib = jl_cell((yr - g.y[0]) * g.idy, 1, g.Ny - 1);
    weights4((xr - g.x[jb - 1]) / g.dx, (yr - g.y[ib - 1]) / g.dy, w);
    const double om = gdot4(A.wyx, g.Ny, ib, jb, w);
    // analytic stress rotation E:5042-5048
    const double th = dt * om;
    const double s = A.sxxm[m], t = A.sxym[m];
    const double ct = cos(th), st = sin(th), s2 = sin(2.0 * th), c2 = cos(2.0 * th);
    A.sxxm[m] = s * (ct * ct) - s * (st * st) - t * s2;
    A.sxym[m] = s * s2 + t * c2;
    // RK4 on the solid velocity E:5055-5136
    double vxr[4], vyr[4];
#pragma unroll
    for (int rk = 0; rk < 4; ++rk) {
        vxr[rk] = 0.0 + rk_vx(g, A.vx, A.vx, xr, yr);
        vyr[rk] = 0.0 + rk_vy(g, A.vy, xr, yr);
        if (rk < 3) {
            xr = xmm + dt * cr[rk] * vxr[rk];
            yr = ymm + dt * cr[rk] * vyr[rk];
        }
    }
    const double xn = xmm + dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
    const double yn = ymm + dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
    A.xm[m] = xn;
    A.ym[m] = yn;
    // RK4 back-trace on the fluid velocity E:5138-5226
    xmm = xr = xn;
    ymm = yr = yn;
#pragma unroll
    for (int rk = 0; rk < 4; ++rk) {
        vxr[rk] = 0.0 + rk_vx(g, A.vxf, A.vx, xr, yr);
        vyr[rk] = 0.0 + rk_vy(g, A.vyf, xr, yr);
        if (rk < 3) {
            xr = xmm - dt * cr[rk] * vxr[rk];
            yr = ymm - dt * cr[rk] * vyr[rk];
        }
    }
    xr = xmm - dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
    yr = ymm - dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
    jp = jl_cell((xr - g.xp[0]) * g.idx, 1, g.Nx);
    ip = jl_cell((yr - g.yp[0]) * g.idy, 1, g.Ny);
    weights4((xr - g.xp[jp - 1]) / g.dx, (yr - g.yp[ip - 1]) / g.dy, w);
    const double tkf0 = gdot4(A.tk2, g.Ny1, ip, jp, w);
    // fluid-solid temperature correction E:5243-5252
    const double d = tkf0 - tks0;
    const int tt = min(max(A.tm[m], 1), 3);  // set validates 1..3
    const double T = A.tkm[m], ph = A.phim[m];
    const double rcs = sel3(A.M.rhocpsolidm, tt);
    const double rcf = rhocpfluidm(A.M, T, A.mode);
    A.tkm[m] = ((1.0 - ph) * T * rcs + ph * (T + d) * rcf) / ((1.0 - ph) * rcs + ph * rcf);
}

int launch_gather_temperature(ereb_ctx* c, int64_t timestep) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    hipEvent_t ev;
    prof_begin(c, EREB_K_GATHER_T, &ev);
    const double* grid = gptr(c, timestep == 1 ? EREB_G_TK2 : EREB_G_DT);
    hipLaunchKernelGGL(k_gather_T, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                       geomk(c), mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), mptr(c, EREB_M_TKM),
                       grid, M, timestep == 1 ? 1 : 0);
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
    hipLaunchKernelGGL(k_gather_S, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream,
                       geomk(c), mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), mptr(c, EREB_M_SXXM),
                       mptr(c, EREB_M_SXYM), gptr(c, EREB_G_DSXX), gptr(c, EREB_G_DSXY), M);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_GATHER_S, ev);
    return EREB_OK;
}

int launch_rk4(ereb_ctx* c, double dt, int mode) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    RK4Args A;
    A.g = geomk(c);
    A.M = c->mat;
    A.dt = dt;
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
    hipEvent_t ev;
    prof_begin(c, EREB_K_RK4, &ev);
    hipLaunchKernelGGL(k_rk4, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream, A);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_RK4, ev);
    return EREB_OK;
}

}  // namespace ereb
