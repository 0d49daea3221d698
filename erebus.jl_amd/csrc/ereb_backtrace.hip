// ereb_backtrace.hip -- backtrace_pressures_rk4! (E:5284-5496): every
// interior P node is traced back through the staggered velocities with the
// RK4 stages and second-order-corrected interpolation of the marker RK4
// (E:5057-5127), and takes the total and solid pressure (solid velocities)
// and the fluid pressure (fluid velocities) at its departure point.
//
// One thread per P node (about 80k at 282x282 nodes; the grids stay in L2).
// Operation sequence as the reference with divisions (no reciprocals, no
// fused products): the results equal the CPU restatement bit for bit.
#include "ereb_internal.h"

namespace ereb {

// stage velocity on the Vx grid (axes x, yp; clamps j 1..Nx-1, i 1..Ny),
// E:5299-5332, combined as (1-dymi/dy)*v13 + (dymi/dy)*v24
__device__ __forceinline__ double bt_vx(const GeomK& g, const double* __restrict__ v, double x,
                                        double y) {
    const int R = g.Ny1;
    const int j = jl_cell((x - g.x[0]) * g.idx, 1, g.Nx - 1);
    const int i = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    const double dxmj = x - g.x[j - 1], dymi = y - g.yp[i - 1];
    const double* p = v + (size_t)(j - 1) * R + (i - 1);
    const double a = dxmj / g.dx;
    double v13 = p[0] * (1.0 - a) + p[R] * dxmj / g.dx;
    double v24 = p[1] * (1.0 - a) + p[R + 1] * dxmj / g.dx;
    if (a >= 0.5) {
        if (j < g.Nx - 1) {
            const double t = a - 0.5;
            v13 += 0.5 * (t * t) * (p[0] - 2.0 * p[R] + p[2 * R]);
            v24 += 0.5 * (t * t) * (p[1] - 2.0 * p[R + 1] + p[2 * R + 1]);
        }
    } else if (j > 1) {
        const double t = a - 0.5;
        v13 += 0.5 * (t * t) * (p[-R] - 2.0 * p[0] + p[R]);
        v24 += 0.5 * (t * t) * (p[1 - R] - 2.0 * p[1] + p[R + 1]);
    }
    const double b = dymi / g.dy;
    return (1.0 - b) * v13 + b * v24;
}

// stage velocity on the Vy grid (axes xp, y; clamps j 1..Nx, i 1..Ny-1),
// E:5333-5364, combined as (1-dxmj/dx)*v12 + (dxmj/dx)*v34
__device__ __forceinline__ double bt_vy(const GeomK& g, const double* __restrict__ v, double x,
                                        double y) {
    const int R = g.Ny1;
    const int j = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int i = jl_cell((y - g.y[0]) * g.idy, 1, g.Ny - 1);
    const double dxmj = x - g.xp[j - 1], dymi = y - g.y[i - 1];
    const double* p = v + (size_t)(j - 1) * R + (i - 1);
    const double b = dymi / g.dy;
    double v12 = p[0] * (1.0 - b) + p[1] * dymi / g.dy;
    double v34 = p[R] * (1.0 - b) + p[R + 1] * dymi / g.dy;
    if (b >= 0.5) {
        if (i < g.Ny - 1) {
            const double t = b - 0.5;
            v12 += 0.5 * (t * t) * (p[0] - 2.0 * p[1] + p[2]);
            v34 += 0.5 * (t * t) * (p[R] - 2.0 * p[R + 1] + p[R + 2]);
        }
    } else if (i > 1) {
        const double t = b - 0.5;
        v12 += 0.5 * (t * t) * (p[-1] - 2.0 * p[0] + p[1]);
        v34 += 0.5 * (t * t) * (p[R - 1] - 2.0 * p[R] + p[R + 1]);
    }
    const double a = dxmj / g.dx;
    return (1.0 - a) * v12 + a * v34;
}

// RK4 departure point of the node at (xA, yA) (E:5296-5376); the last step
// is xA - dt*1//6*(...) = xA - (dt*Float64(1//6))*(...)
__device__ __forceinline__ void bt_node(const GeomK& g, const double* __restrict__ vx,
                                        const double* __restrict__ vy, double xA, double yA,
                                        double dt, double& xo, double& yo) {
    double xc = xA, yc = yA, vxm[4], vym[4];
#pragma unroll
    for (int rk = 0; rk < 4; ++rk) {
        vxm[rk] = bt_vx(g, vx, xc, yc);
        vym[rk] = bt_vy(g, vy, xc, yc);
        if (rk < 2) {
            xc = xA - 0.5 * dt * vxm[rk];
            yc = yA - 0.5 * dt * vym[rk];
        } else if (rk == 2) {
            xc = xA - dt * vxm[rk];
            yc = yA - dt * vym[rk];
        }
    }
    const double sixth = 1.0 / 6.0;
    xo = xA - dt * sixth * (vxm[0] + 2.0 * vxm[1] + 2.0 * vxm[2] + vxm[3]);
    yo = yA - dt * sixth * (vym[0] + 2.0 * vym[1] + 2.0 * vym[2] + vym[3]);
}

// fix_weights on the P grid (clamps 1..Nx, 1..Ny), then dot4 of each field
__device__ __forceinline__ void p_weights(const GeomK& g, double x, double y, int& i, int& j,
                                          double* w) {
    j = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    i = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    weights4((x - g.xp[j - 1]) / g.dx, (y - g.yp[i - 1]) / g.dy, w);
}

__global__ __launch_bounds__(256) void k_backtrace(
    GeomK g, const double* __restrict__ pr, double* __restrict__ pr0,
    const double* __restrict__ ps, double* __restrict__ ps0, const double* __restrict__ pf,
    double* __restrict__ pf0, const double* __restrict__ vx, const double* __restrict__ vy,
    const double* __restrict__ vxf, const double* __restrict__ vyf, double dt) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.Ny1 * g.Nx1) return;
    const int ii = n % g.Ny1 + 1, jj = n / g.Ny1 + 1;
    if (ii < 2 || ii > g.Ny || jj < 2 || jj > g.Nx) {  // pr0 .= pr etc. (E:5292-5293, E:5393)
        pr0[n] = pr[n];
        ps0[n] = ps[n];
        pf0[n] = pf[n];
        return;
    }
    const double xA = g.xp[jj - 1], yA = g.yp[ii - 1];
    double x, y, w[4];
    int i, j;
    bt_node(g, vx, vy, xA, yA, dt, x, y);
    p_weights(g, x, y, i, j, w);
    pr0[n] = gdot4(pr, g.Ny1, i, j, w);
    ps0[n] = gdot4(ps, g.Ny1, i, j, w);
    bt_node(g, vxf, vyf, xA, yA, dt, x, y);
    p_weights(g, x, y, i, j, w);
    pf0[n] = gdot4(pf, g.Ny1, i, j, w);
}

int launch_backtrace_pressures(ereb_ctx* c, double dt) {
    const int n = c->Ny1 * c->Nx1;
    hipEvent_t ev;
    prof_begin(c, EREB_K_BACKTRACE, &ev);
    hipLaunchKernelGGL(k_backtrace, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       geomk(c), gptr(c, EREB_G_PR), gptr(c, EREB_G_PR0), gptr(c, EREB_G_PS),
                       gptr(c, EREB_G_PS0), gptr(c, EREB_G_PF), gptr(c, EREB_G_PF0),
                       gptr(c, EREB_G_VX), gptr(c, EREB_G_VY), gptr(c, EREB_G_VXF),
                       gptr(c, EREB_G_VYF), dt);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_BACKTRACE, ev);
    return EREB_OK;
}

}  // namespace ereb
