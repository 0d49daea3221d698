/*
 * ereb_dual.cpp -- forward-mode (dual-number) restatement of the oracle's
 * differentiable transfer functions.  TEST INFRASTRUCTURE ONLY: the exact
 * derivatives that the GPU pullbacks and the checkpointed adjoint are checked
 * against, in place of ForwardDiff of the reference (not installable here).
 *
 * Every function is a template over the scalar type S and follows
 * ereb_oracle.c operation by operation (which follows E: = src/Erebus.jl):
 * with S = double it reproduces the oracle bit for bit (tested), with
 * S = Dual<N> it also carries N exact directional derivatives.  Branches
 * and cell indices use the value part only, so the derivatives are those of
 * the piece of the piecewise function the point lies in, as ForwardDiff's.
 * Built with -ffp-contract=off like the oracle.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {
#include "../include/ereb.h"
}

namespace {

// ---- dual numbers -----------------------------------------------------------
template <int N>
struct Dual {
    double v;
    double d[N];
    Dual() : v(0.0) { std::memset(d, 0, sizeof d); }
    Dual(double x) : v(x) { std::memset(d, 0, sizeof d); }  // NOLINT: implicit on purpose
};

template <int N>
inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    const double ib = 1.0 / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N> inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r(a.v * b);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> inline Dual<N> operator*(double a, const Dual<N>& b) {
    Dual<N> r(a * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int N> inline Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> inline Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { return a = a + b; }
template <int N>
inline Dual<N> sqrt(const Dual<N>& a) {
    Dual<N> r(std::sqrt(a.v));
    const double h = 0.5 / r.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}
template <int N>
inline Dual<N> sin(const Dual<N>& a) {
    Dual<N> r(std::sin(a.v));
    const double c = std::cos(a.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * c;
    return r;
}
template <int N>
inline Dual<N> cos(const Dual<N>& a) {
    Dual<N> r(std::cos(a.v));
    const double s = -std::sin(a.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r;
}

inline double val(double x) { return x; }
template <int N> inline double val(const Dual<N>& x) { return x.v; }

inline double jl_sin(double x) { return std::sin(x); }
inline double jl_cos(double x) { return std::cos(x); }
template <int N> inline Dual<N> jl_sin(const Dual<N>& x) { return sin(x); }
template <int N> inline Dual<N> jl_cos(const Dual<N>& x) { return cos(x); }
inline double jl_sqrt(double x) { return std::sqrt(x); }
template <int N> inline Dual<N> jl_sqrt(const Dual<N>& x) { return sqrt(x); }

#define G_(a, i, j, rows) (a)[((int64_t)(j) - 1) * (rows) + ((int64_t)(i) - 1)]

// unsafe_trunc as x86-64 cvttsd2si (ereb_oracle.c)
inline int64_t jl_unsafe_trunc(double t) {
    if (!(t > -9223372036854775808.0 && t < 9223372036854775808.0))
        return INT64_MIN;
    return (int64_t)t;
}

// ---- marker properties (E:700-750 with E:1116-1357, E:2272-2358) ------------
template <class S> S total(S s, S f, S phi) { return s * (1.0 - phi) + f * phi; }

template <class S> S ktotal(S ks, S kf, S phi) {
    S t = ks * (3.0 * phi - 2.0) + kf * (1.0 - 3.0 * phi);
    return jl_sqrt(ks * kf / 2.0 + (t * t) * (1.0 / 16.0)) - 0.25 * t;
}

template <class S> S etaf_inv_kphi(double phim0, double kphir, S phi, double etaf) {
    S q = phim0 * (1.0 / phi);
    S r = (1.0 - phi) * (1.0 / (1.0 - phim0));
    return etaf * (1.0 / kphir) * (q * q * q) * (r * r);
}

template <class S> S rhocpfluidm(const ereb_materials* M, S T, int64_t mode) {
    if (mode == 9) return S(M->rhocpfluidm[0]);
    const double tf = M->tmfluidphase, t = val(T);
    if (t < tf - 5.0) return M->rhoH2Ofi * (7.67 * T);
    if (t < tf) return M->rhoH2Ofi * (7.67 * T + 0.1 * M->Lf);
    if (t < tf + 5.0) return S(M->rhoH2Of * (4200.0 + 0.1 * M->Lf));
    if (t < 410.0) return S(M->rhoH2Of * 4200.0);
    return M->rhoH2Of * (-4.67e4 + 333.0 * T - 0.731 * (T * T) + 5.4e-4 * (T * T * T));
}

template <class S> S ksolidm(const ereb_materials* M, S T, int64_t mode) {
    if (mode == 9) return S(M->ksolidm[0]);
    return 0.73 + 1293.0 / (T + 77.0);
}

template <class S> S kfluidm(const ereb_materials* M, S T, int64_t mode) {
    if (mode == 9) return S(M->kfluidm[0]);
    const double t = val(T);
    if (t < M->tmfluidphase) return 0.465 + 488.0 / T;
    if (t < 410.0) return -0.581 + 6.34e-3 * T - 7.93e-6 * (T * T);
    return -0.142 + 4.12e-3 * T - 5.01e-6 * (T * T);
}

inline double jl_max(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : (b > a ? b : (std::signbit(a) ? b : a));
}

// the seven property outputs of one marker, in this order
enum { P_RHO = 0, P_RHOCP, P_ETA, P_HR, P_K, P_TKRHOCP, P_ETAFK, NPROP };

// compute_marker_properties! for one marker; st[] are the stored values the
// air markers keep (rho, rhocp, eta, hr, k)
template <class S>
void marker_props(const ereb_materials* M, int64_t t, S T, S phi, S XW, const double* hrs,
                  const double* hrf, int64_t mode, const S* st, S* p) {
    double etafluidcur;
    if (t < 3) {
        S XD = 1.0 - XW;
        S rhosolidm0 = (M->MD + M->MH2O * XW) / (M->VDs * XD + M->VWs * XW);
        double rhofluidm0 = val(T) > M->tmfluidphase ? M->rhoH2Of : M->rhoH2Ofi;
        p[P_RHO] = total<S>(rhosolidm0, S(rhofluidm0), phi);
        p[P_RHOCP] = total<S>(S(M->rhocpsolidm[t - 1]), rhocpfluidm<S>(M, T, mode), phi);
        double es = val(T) > M->tmsolidphase ? M->etasolidmm[t - 1] : M->etasolidm[t - 1];
        etafluidcur = val(T) > M->tmfluidphase ? M->etafluidmm[t - 1] : M->etafluidm[t - 1];
        p[P_ETA] = S(jl_max(jl_max(M->etamin, es), etafluidcur));
        p[P_HR] = total<S>(S(hrs[t - 1]), S(hrf[t - 1]), phi);
        p[P_K] = ktotal<S>(ksolidm<S>(M, T, mode), kfluidm<S>(M, T, mode), phi);
    } else {
        for (int k = 0; k < 5; ++k) p[k] = st[k];
        etafluidcur = M->etafluidm[t - 1];
    }
    p[P_TKRHOCP] = T * p[P_RHOCP];
    p[P_ETAFK] = etaf_inv_kphi<S>(M->phim0, M->kphim0[t - 1], phi, etafluidcur);
}

// ---- interpolation primitives (E:1388-1470, E:1137-1159) --------------------
// cell (fix) from the value parts, distances and weights in S
template <class S>
void fix_weights(S x, S y, const double* xa, const double* ya, double dx, double dy,
                 int64_t jmin, int64_t jmax, int64_t imin, int64_t imax, int64_t* ri,
                 int64_t* rj, S* w) {
    int64_t j = jl_unsafe_trunc((val(x) - xa[0]) * (1.0 / dx)) + 1;
    int64_t i = jl_unsafe_trunc((val(y) - ya[0]) * (1.0 / dy)) + 1;
    if (j < jmin) j = jmin; else if (j > jmax) j = jmax;
    if (i < imin) i = imin; else if (i > imax) i = imax;
    *ri = i;
    *rj = j;
    S dxmj = x - xa[j - 1], dymi = y - ya[i - 1];
    w[0] = (1.0 - dymi / dy) * (1.0 - dxmj / dx);
    w[1] = (dymi / dy) * (1.0 - dxmj / dx);
    w[2] = (1.0 - dymi / dy) * (dxmj / dx);
    w[3] = (dymi / dy) * (dxmj / dx);
}

template <class S>
void fix_distances(S x, S y, const double* xa, const double* ya, double dx, double dy,
                   int64_t jmin, int64_t jmax, int64_t imin, int64_t imax, int64_t* ri,
                   int64_t* rj, S* dxmj, S* dymi) {
    int64_t j = jl_unsafe_trunc((val(x) - xa[0]) * (1.0 / dx)) + 1;
    int64_t i = jl_unsafe_trunc((val(y) - ya[0]) * (1.0 / dy)) + 1;
    if (j < jmin) j = jmin; else if (j > jmax) j = jmax;
    if (i < imin) i = imin; else if (i > imax) i = imax;
    *ri = i;
    *rj = j;
    *dxmj = x - xa[j - 1];
    *dymi = y - ya[i - 1];
}

// A grid whose values may carry tangents: value array and, per tangent
// direction k, an optional tangent array (NULL = zero)
template <class S> struct GridOf;
template <> struct GridOf<double> {
    const double* v;
    double at(int64_t k) const { return v[k]; }
};
template <int N> struct GridOf<Dual<N>> {
    const double* v;
    const double* const* t;  // N tangent arrays or nullptr
    Dual<N> at(int64_t k) const {
        Dual<N> r(v[k]);
        if (t)
            for (int q = 0; q < N; ++q) r.d[q] = t[q] ? t[q][k] : 0.0;
        return r;
    }
};

template <class S>
S gdot4(const GridOf<S>& g, int64_t rows, int64_t i, int64_t j, const S* w) {
    return g.at((j - 1) * rows + i - 1) * w[0] + g.at((j - 1) * rows + i) * w[1] +
           g.at(j * rows + i - 1) * w[2] + g.at(j * rows + i) * w[3];
}

// ---- RK4 stages (E:5057-5127) ------------------------------------------------
template <class S>
S rk_vx(const ereb_geometry* g, const GridOf<S>& v, const GridOf<S>& vc, S x, S y) {
    const int64_t Nx = g->Nx, Ny = g->Ny, R = Ny + 1;
    int64_t i, j;
    S dxmj, dymi;
    fix_distances<S>(x, y, g->xvx, g->yvx, g->dx, g->dy, 1, Nx - 1, 1, Ny, &i, &j, &dxmj,
                     &dymi);
    const double dx = g->dx, dy = g->dy;
    auto V = [&](const GridOf<S>& a, int64_t ii, int64_t jj) { return a.at((jj - 1) * R + ii - 1); };
    S v13 = V(v, i, j) * (1.0 - dxmj / dx) + V(v, i, j + 1) * dxmj / dx;
    S v24 = V(v, i + 1, j) * (1.0 - dxmj / dx) + V(v, i + 1, j + 1) * dxmj / dx;
    if (val(dxmj / dx) >= 0.5) {
        if (j < Nx - 1) {
            S t = dxmj / dx - 0.5;
            v13 += 0.5 * (t * t) * (V(v, i, j) - 2.0 * V(v, i, j + 1) + V(v, i, j + 2));
            v24 += 0.5 * (t * t) * (V(v, i + 1, j) - 2.0 * V(v, i + 1, j + 1) + V(v, i + 1, j + 2));
        }
    } else {
        if (j > 1) {
            S t = dxmj / dx - 0.5;
            v13 += 0.5 * (t * t) * (V(v, i, j - 1) - 2.0 * V(vc, i, j) + V(v, i, j + 1));
            v24 += 0.5 * (t * t) * (V(v, i + 1, j - 1) - 2.0 * V(v, i + 1, j) + V(v, i + 1, j + 1));
        }
    }
    return v13 * (1.0 - dymi / dy) + v24 * dymi / dy;
}

template <class S>
S rk_vy(const ereb_geometry* g, const GridOf<S>& v, S x, S y) {
    const int64_t Nx = g->Nx, Ny = g->Ny, R = Ny + 1;
    int64_t i, j;
    S dxmj, dymi;
    fix_distances<S>(x, y, g->xvy, g->yvy, g->dx, g->dy, 1, Nx, 1, Ny - 1, &i, &j, &dxmj,
                     &dymi);
    const double dx = g->dx, dy = g->dy;
    auto V = [&](int64_t ii, int64_t jj) { return v.at((jj - 1) * R + ii - 1); };
    S v12 = V(i, j) * (1.0 - dymi / dy) + V(i + 1, j) * dymi / dy;
    S v34 = V(i, j + 1) * (1.0 - dymi / dy) + V(i + 1, j + 1) * dymi / dy;
    if (val(dymi / dy) >= 0.5) {
        if (i < Ny - 1) {
            S t = dymi / dy - 0.5;
            v12 += 0.5 * (t * t) * (V(i, j) - 2.0 * V(i + 1, j) + V(i + 2, j));
            v34 += 0.5 * (t * t) * (V(i, j + 1) - 2.0 * V(i + 1, j + 1) + V(i + 2, j + 1));
        }
    } else {
        if (i > 1) {
            S t = dymi / dy - 0.5;
            v12 += 0.5 * (t * t) * (V(i - 1, j) - 2.0 * V(i, j) + V(i + 1, j));
            v34 += 0.5 * (t * t) * (V(i - 1, j + 1) - 2.0 * V(i, j + 1) + V(i + 1, j + 1));
        }
    }
    return v12 * (1.0 - dxmj / dx) + v34 * dxmj / dx;
}

// the grids RK4 reads
template <class S>
struct RK4Grids {
    GridOf<S> vx, vy, vxf, vyf, wyx, tk2;
};

// move_markers_rk4! (E:4990-5257) for one marker: in/out x, y, T, sxy, sxx
template <class S>
void rk4_one(const ereb_geometry* g, const ereb_materials* M, const RK4Grids<S>& G, double dt,
             int64_t mode, int64_t tt, S& xm, S& ym, S& T, S phi, S& sxym, S& sxxm) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    const double* b = M->brk4;
    const double* c = M->crk4;
    S xmm = xm, ymm = ym;
    S xr = xmm, yr = ymm;
    int64_t i, j;
    S w[4];
    fix_weights<S>(xr, yr, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
    S tks0 = gdot4<S>(G.tk2, Ny1, i, j, w);
    fix_weights<S>(xr, yr, g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1, Ny - 1, &i, &j, w);
    S om = gdot4<S>(G.wyx, Ny, i, j, w);
    S th = dt * om;
    S s = sxxm, t = sxym;
    S ct = jl_cos(th), st = jl_sin(th), s2 = jl_sin(2.0 * th), c2 = jl_cos(2.0 * th);
    sxxm = s * (ct * ct) - s * (st * st) - t * s2;
    sxym = s * s2 + t * c2;
    S vxr[4], vyr[4];
    for (int rk = 0; rk < 4; ++rk) {
        vxr[rk] = 0.0 + rk_vx<S>(g, G.vx, G.vx, xr, yr);
        vyr[rk] = 0.0 + rk_vy<S>(g, G.vy, xr, yr);
        if (rk < 3) {
            xr = xmm + dt * c[rk] * vxr[rk];
            yr = ymm + dt * c[rk] * vyr[rk];
        }
    }
    xm = xm + dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
    ym = ym + dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
    xmm = xr = xm;
    ymm = yr = ym;
    for (int rk = 0; rk < 4; ++rk) {
        vxr[rk] = 0.0 + rk_vx<S>(g, G.vxf, G.vx, xr, yr);
        vyr[rk] = 0.0 + rk_vy<S>(g, G.vyf, xr, yr);
        if (rk < 3) {
            xr = xmm - dt * c[rk] * vxr[rk];
            yr = ymm - dt * c[rk] * vyr[rk];
        }
    }
    xr = xmm - dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
    yr = ymm - dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
    fix_weights<S>(xr, yr, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
    S tkf0 = gdot4<S>(G.tk2, Ny1, i, j, w);
    S d = tkf0 - tks0;
    const double rcs = M->rhocpsolidm[tt - 1];
    S rcf = rhocpfluidm<S>(M, T, mode);
    T = ((1.0 - phi) * T * rcs + phi * (T + d) * rcf) / ((1.0 - phi) * rcs + phi * rcf);
}

// ---- the marker's contribution to the 30 SUM grids (E:1605-1863) ------------
// Sum over its 16 stencil nodes of bar[node] * (value * weight), i.e. the
// scalar whose derivative is the marker's row of the scatter's transpose.
struct SumBars {
    const double* s[30];  // NULL = zero
};

template <class S>
S scatter_dot(const ereb_geometry* g, const SumBars& B, S x, S y, const S* v /* 30 values */) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    S acc(0.0);
    int64_t i, j;
    S w[4];
    auto add = [&](int slot, int64_t rows) {
        const double* b = B.s[slot];
        if (!b) return;
        const int64_t n0 = (j - 1) * rows + i - 1;
        acc += b[n0] * (v[slot] * w[0]) + b[n0 + 1] * (v[slot] * w[1]) +
               b[n0 + rows] * (v[slot] * w[2]) + b[n0 + rows + 1] * (v[slot] * w[3]);
    };
    fix_weights<S>(x, y, g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1, Ny - 1, &i, &j, w);
    for (int k = 0; k < 8; ++k) add(k, Ny);
    fix_weights<S>(x, y, g->xvx, g->yvx, g->dx, g->dy, 1, Nx - 1, 1, Ny, &i, &j, w);
    for (int k = 8; k < 14; ++k) add(k, Ny1);
    fix_weights<S>(x, y, g->xvy, g->yvy, g->dx, g->dy, 1, Nx, 1, Ny - 1, &i, &j, w);
    for (int k = 14; k < 20; ++k) add(k, Ny1);
    fix_weights<S>(x, y, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
    for (int k = 20; k < 30; ++k) add(k, Ny1);
    return acc;
}

// the 30 staged values of one marker in SUM order, from its fields and
// properties (E:1605-1863)
template <class S>
void sum_values(const S* f /* by ereb_marker_field */, const S* p /* 7 props */, S* v) {
    v[0] = p[P_ETA];
    v[1] = f[EREB_M_ETAVPM];
    v[2] = f[EREB_M_INV_GGGTOTALM];
    v[3] = f[EREB_M_SXYM];
    v[4] = f[EREB_M_COHESTOTALM];
    v[5] = f[EREB_M_TENSTOTALM];
    v[6] = f[EREB_M_FRICTTOTALM];
    v[7] = S(1.0);
    for (int q = 0; q < 2; ++q) {
        const int o = q ? 14 : 8;
        v[o + 0] = p[P_RHO];
        v[o + 1] = f[EREB_M_RHOFLUIDCUR];
        v[o + 2] = p[P_K];
        v[o + 3] = f[EREB_M_PHIM];
        v[o + 4] = p[P_ETAFK];
        v[o + 5] = S(1.0);
    }
    v[20] = p[P_RHO];
    v[21] = p[P_RHOCP];
    v[22] = f[EREB_M_ALPHASOLIDCUR];
    v[23] = f[EREB_M_ALPHAFLUIDCUR];
    v[24] = p[P_HR];
    v[25] = f[EREB_M_INV_GGGTOTALM];
    v[26] = f[EREB_M_SXXM];
    v[27] = p[P_TKRHOCP];
    v[28] = f[EREB_M_PHIM];
    v[29] = S(1.0);
}

using D1 = Dual<1>;

}  // namespace

extern "C" {

/* Row m of the transpose of ereb_marker_to_nodes (properties + scatter),
 * differentiated with respect to marker field `field` (ereb_marker_field):
 * out[m] = d/d field_m of [ sum over nodes of sum_bar * contribution_m
 *                          + sum over the 7 outputs of prop_bar * property_m ].
 * markers[f] are the marker vectors by ereb_marker_field (tm as int64); the
 * stored properties of air markers are the property vectors' values. */
int dual_marker_to_nodes_vjp(const ereb_geometry* g, const ereb_materials* M,
                             const void* const* markers, int64_t marknum,
                             const double* hrs, const double* hrf, int64_t mode,
                             const double* const* sum_bar, const double* const* prop_bar,
                             int field, double* out) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    SumBars B;
    for (int k = 0; k < 30; ++k) B.s[k] = sum_bar ? sum_bar[k] : nullptr;
    const int64_t* tm = (const int64_t*)markers[EREB_M_TM];
    for (int64_t m = 0; m < marknum; ++m) {
        D1 f[EREB_M_NFIELDS];
        for (int q = 0; q < EREB_M_NFIELDS; ++q) {
            if (q == EREB_M_TM || !markers[q]) continue;
            f[q] = D1(((const double*)markers[q])[m]);
        }
        f[field].d[0] = 1.0;
        D1 st[5] = {f[EREB_M_RHOTOTALM], f[EREB_M_RHOCPTOTALM], f[EREB_M_ETATOTALM],
                    f[EREB_M_HRTOTALM], f[EREB_M_KTOTALM]};
        D1 p[NPROP], v[30];
        marker_props<D1>(M, tm[m], f[EREB_M_TKM], f[EREB_M_PHIM], f[EREB_M_XWSOLIDM0], hrs, hrf,
                         mode, st, p);
        sum_values<D1>(f, p, v);
        D1 s = scatter_dot<D1>(g, B, f[EREB_M_XM], f[EREB_M_YM], v);
        if (prop_bar)
            for (int k = 0; k < NPROP; ++k)
                if (prop_bar[k]) s += prop_bar[k][m] * p[k];
        out[m] = s.d[0];
    }
    return EREB_OK;
}

/* The same with S = double: the marker's SUM contributions dotted with
 * sum_bar (checks the template against the oracle's scatter). */
int dual_marker_to_nodes_dot(const ereb_geometry* g, const ereb_materials* M,
                             const void* const* markers, int64_t marknum,
                             const double* hrs, const double* hrf, int64_t mode,
                             const double* const* sum_bar, double* out, double* props_out) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    SumBars B;
    for (int k = 0; k < 30; ++k) B.s[k] = sum_bar ? sum_bar[k] : nullptr;
    const int64_t* tm = (const int64_t*)markers[EREB_M_TM];
    for (int64_t m = 0; m < marknum; ++m) {
        double f[EREB_M_NFIELDS] = {};
        for (int q = 0; q < EREB_M_NFIELDS; ++q)
            if (q != EREB_M_TM && markers[q]) f[q] = ((const double*)markers[q])[m];
        double st[5] = {f[EREB_M_RHOTOTALM], f[EREB_M_RHOCPTOTALM], f[EREB_M_ETATOTALM],
                        f[EREB_M_HRTOTALM], f[EREB_M_KTOTALM]};
        double p[NPROP], v[30];
        marker_props<double>(M, tm[m], f[EREB_M_TKM], f[EREB_M_PHIM], f[EREB_M_XWSOLIDM0], hrs,
                             hrf, mode, st, p);
        sum_values<double>(f, p, v);
        out[m] = scatter_dot<double>(g, B, f[EREB_M_XM], f[EREB_M_YM], v);
        if (props_out)
            for (int k = 0; k < NPROP; ++k) props_out[(int64_t)k * marknum + m] = p[k];
    }
    return EREB_OK;
}

/* update_marker_temperature! (E:4782-4801) differentiated: per marker,
 * d/d(seed) of tkm_bar[m] * tkm_new[m], the seed being xm (which = 0), ym (1)
 * or tkm (2), or, with which = 3, the grid (DT or tk2 at timestep 1) along the
 * tangent grid `gdot`; out[m] per marker (summed by the caller for which = 3). */
int dual_gather_T_vjp(const ereb_geometry* g, const double* xm, const double* ym,
                      const double* tkm, const double* DT, const double* tk2, int64_t timestep,
                      int64_t marknum, const double* tkm_bar, int which, const double* gdot,
                      double* out) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    const double* gv = timestep == 1 ? tk2 : DT;
    const double* tg[1] = {which == 3 ? gdot : nullptr};
    GridOf<D1> grid{gv, tg};
    for (int64_t m = 0; m < marknum; ++m) {
        D1 x(xm[m]), y(ym[m]), T(tkm[m]);
        if (which == 0) x.d[0] = 1.0;
        if (which == 1) y.d[0] = 1.0;
        if (which == 2) T.d[0] = 1.0;
        int64_t i, j;
        D1 w[4];
        fix_weights<D1>(x, y, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
        D1 r = timestep == 1 ? gdot4<D1>(grid, Ny1, i, j, w) : T + gdot4<D1>(grid, Ny1, i, j, w);
        out[m] = tkm_bar[m] * r.d[0];
    }
    return EREB_OK;
}

/* update_marker_stress! (E:4354-4388) differentiated: per marker, d/d(seed)
 * of sxx_bar*sxx_new + sxy_bar*sxy_new; seed xm (0), ym (1), sxxm (2), sxym
 * (3), DSXX along gdot (4), DSXY along gdot (5). */
int dual_gather_S_vjp(const ereb_geometry* g, const double* xm, const double* ym,
                      const double* sxxm, const double* sxym, const double* DSXX,
                      const double* DSXY, int64_t marknum, const double* sxx_bar,
                      const double* sxy_bar, int which, const double* gdot, double* out) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    const double* t4[1] = {which == 4 ? gdot : nullptr};
    const double* t5[1] = {which == 5 ? gdot : nullptr};
    GridOf<D1> gxx{DSXX, t4}, gxy{DSXY, t5};
    for (int64_t m = 0; m < marknum; ++m) {
        D1 x(xm[m]), y(ym[m]), sxx(sxxm[m]), sxy(sxym[m]);
        if (which == 0) x.d[0] = 1.0;
        if (which == 1) y.d[0] = 1.0;
        if (which == 2) sxx.d[0] = 1.0;
        if (which == 3) sxy.d[0] = 1.0;
        int64_t ip, jp, ib, jb;
        D1 wp[4], wb[4];
        fix_weights<D1>(x, y, g->xp, g->yp, g->dx, g->dy, 2, Nx - 1, 2, Ny - 1, &ip, &jp, wp);
        fix_weights<D1>(x, y, g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1, Ny - 1, &ib, &jb, wb);
        D1 a = sxx + gdot4<D1>(gxx, Ny1, ip, jp, wp);
        D1 b = sxy + gdot4<D1>(gxy, Ny, ib, jb, wb);
        out[m] = sxx_bar[m] * a.d[0] + sxy_bar[m] * b.d[0];
    }
    return EREB_OK;
}

/* move_markers_rk4! (E:4990-5257) differentiated: per marker, d/d(seed) of
 * xm_bar*x' + ym_bar*y' + tkm_bar*T' + sxym_bar*sxy' + sxxm_bar*sxx'; seed
 * (which) xm 0, ym 1, tkm 2, phim 3, sxym 4, sxxm 5, or a grid along its
 * tangent gdot: vx 6, vy 7, vxf 8, vyf 9, wyx 10, tk2 11.  A NULL output bar
 * is zero. */
int dual_rk4_vjp(const ereb_geometry* g, const ereb_materials* M, const double* xm,
                 const double* ym, const int64_t* tm, const double* tkm, const double* phim,
                 const double* sxym, const double* sxxm, const double* vx, const double* vy,
                 const double* vxf, const double* vyf, const double* wyx, const double* tk2,
                 int64_t marknum, double dt, int64_t mode, const double* xm_bar,
                 const double* ym_bar, const double* tkm_bar, const double* sxym_bar,
                 const double* sxxm_bar, int which, const double* gdot, double* out) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    const double* tg[6][1];
    for (int k = 0; k < 6; ++k) tg[k][0] = which == 6 + k ? gdot : nullptr;
    RK4Grids<D1> G{{vx, tg[0]}, {vy, tg[1]}, {vxf, tg[2]}, {vyf, tg[3]}, {wyx, tg[4]},
                   {tk2, tg[5]}};
    auto bar = [&](const double* b, int64_t m) { return b ? b[m] : 0.0; };
    for (int64_t m = 0; m < marknum; ++m) {
        D1 x(xm[m]), y(ym[m]), T(tkm[m]), ph(phim[m]), sxy(sxym[m]), sxx(sxxm[m]);
        D1* seed[6] = {&x, &y, &T, &ph, &sxy, &sxx};
        if (which < 6) seed[which]->d[0] = 1.0;
        rk4_one<D1>(g, M, G, dt, mode, tm[m], x, y, T, ph, sxy, sxx);
        out[m] = bar(xm_bar, m) * x.d[0] + bar(ym_bar, m) * y.d[0] + bar(tkm_bar, m) * T.d[0] +
                 bar(sxym_bar, m) * sxy.d[0] + bar(sxxm_bar, m) * sxx.d[0];
    }
    return EREB_OK;
}

/* move_markers_rk4! with S = double (checks the template against the oracle) */
int dual_rk4_double(const ereb_geometry* g, const ereb_materials* M, double* xm, double* ym,
                    const int64_t* tm, double* tkm, const double* phim, double* sxym,
                    double* sxxm, const double* vx, const double* vy, const double* vxf,
                    const double* vyf, const double* wyx, const double* tk2, int64_t marknum,
                    double dt, int64_t mode) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    RK4Grids<double> G{{vx}, {vy}, {vxf}, {vyf}, {wyx}, {tk2}};
    for (int64_t m = 0; m < marknum; ++m)
        rk4_one<double>(g, M, G, dt, mode, tm[m], xm[m], ym[m], tkm[m], phim[m], sxym[m],
                        sxxm[m]);
    return EREB_OK;
}

}  // extern "C"

// ---- the multi-step loss map with NT tangents -------------------------------
namespace {

constexpr int NT = 16;  // tangent directions per pass
using DT_ = Dual<NT>;

// SUM grids -> P node tk1 (compute_p_node_properties! E:2148-2189, only the
// fields the loss map needs: TKSUM, RHOCPSUM, WTPSUM), then the insulating
// boundary (E:2432-2449).  Nodes with WTPSUM == 0 keep their previous value.
void tk1_of(const ereb_geometry* g, const std::vector<DT_>& TKSUM,
            const std::vector<DT_>& RHOCPSUM, const std::vector<double>& WTPSUM,
            std::vector<DT_>& tk1) {
    const int64_t Nx1 = g->Nx + 1, Ny1 = g->Ny + 1, n = Nx1 * Ny1;
    for (int64_t k = 0; k < n; ++k)
        if (WTPSUM[k] > 0.0) tk1[k] = TKSUM[k] * (1.0 / RHOCPSUM[k]);
    if (!(Ny1 > 2 && Nx1 > 2)) return;
    auto T = [&](int64_t i, int64_t j) -> DT_& { return tk1[(j - 1) * Ny1 + i - 1]; };
    for (int64_t j = 2; j <= Nx1 - 1; ++j) T(1, j) = T(2, j);
    for (int64_t j = 2; j <= Nx1 - 1; ++j) T(Ny1, j) = T(Ny1 - 1, j);
    for (int64_t i = 1; i <= Ny1; ++i) T(i, 1) = T(i, 2);
    for (int64_t i = 1; i <= Ny1; ++i) T(i, Nx1) = T(i, Nx1 - 1);
}

}  // namespace

extern "C" {

/* Exact gradient entries of L = sum over P nodes of tk1^2 after nsteps steps
 * of the transfer map that ereb_adjoint_tk1_loss differentiates (marker
 * properties + scatter, P node properties with the insulating tk1, tk2 :=
 * tk1, the temperature gather at the first step, RK4 through vx, vy (= vxf,
 * vyf) and wyx = 0, tk1 after a final scatter), with respect to the initial
 * tkm (kind 0) or phim (kind 1) of the markers seed_m[k], k < nseed, in passes
 * of 16 tangent directions.  markers[f] by ereb_marker_field (positions,
 * stresses and properties are copied, not modified).  grad[k] receives
 * dL/d(kind_k of marker seed_m[k]); *loss the loss. */
int dual_tk1_loss_grad(const ereb_geometry* g, const ereb_materials* M,
                       const void* const* markers, int64_t marknum, const double* vx,
                       const double* vy, int64_t nsteps, double dt, int64_t mode,
                       const double* hrs, const double* hrf, int64_t nseed,
                       const int64_t* seed_m, const int32_t* seed_kind, double* grad,
                       double* loss) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    const int64_t Nx = g->Nx, Ny = g->Ny, Nx1 = Nx + 1, Ny1 = Ny + 1, n = Nx1 * Ny1;
    const int64_t* tm = (const int64_t*)markers[EREB_M_TM];
    auto F = [&](int f) { return (const double*)markers[f]; };
    for (int64_t k0 = 0; k0 < nseed || (k0 == 0 && nseed == 0); k0 += NT) {
        const int nk = (int)std::min<int64_t>(NT, nseed - k0);
        std::vector<double> x(F(EREB_M_XM), F(EREB_M_XM) + marknum);
        std::vector<double> y(F(EREB_M_YM), F(EREB_M_YM) + marknum);
        std::vector<DT_> T((size_t)marknum), ph((size_t)marknum);
        std::vector<DT_> st[5];
        static const int stf[5] = {EREB_M_RHOTOTALM, EREB_M_RHOCPTOTALM, EREB_M_ETATOTALM,
                                   EREB_M_HRTOTALM, EREB_M_KTOTALM};
        for (int q = 0; q < 5; ++q) {
            st[q].resize((size_t)marknum);
            for (int64_t m = 0; m < marknum; ++m) st[q][m] = DT_(F(stf[q])[m]);
        }
        for (int64_t m = 0; m < marknum; ++m) {
            T[m] = DT_(F(EREB_M_TKM)[m]);
            ph[m] = DT_(F(EREB_M_PHIM)[m]);
        }
        for (int k = 0; k < nk; ++k) {
            const int64_t m = seed_m[k0 + k];
            (seed_kind[k0 + k] == 0 ? T[m] : ph[m]).d[k] = 1.0;
        }
        std::vector<DT_> tk1((size_t)n), TKSUM((size_t)n), RHOCPSUM((size_t)n);
        std::vector<double> WTPSUM((size_t)n);
        // marker properties (air keeps, and so passes on, its stored values)
        // and the P scatter of TK, RHOCP and the weights (E:1814-1863)
        auto scatter = [&]() {
            std::fill(TKSUM.begin(), TKSUM.end(), DT_(0.0));
            std::fill(RHOCPSUM.begin(), RHOCPSUM.end(), DT_(0.0));
            std::fill(WTPSUM.begin(), WTPSUM.end(), 0.0);
            for (int64_t m = 0; m < marknum; ++m) {
                DT_ s5[5] = {st[0][m], st[1][m], st[2][m], st[3][m], st[4][m]};
                DT_ p[NPROP];
                marker_props<DT_>(M, tm[m], T[m], ph[m], DT_(F(EREB_M_XWSOLIDM0)[m]), hrs,
                                  hrf, mode, s5, p);
                for (int q = 0; q < 5; ++q) st[q][m] = p[q];
                int64_t i, j;
                double w[4];
                fix_weights<double>(x[m], y[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i,
                                    &j, w);
                const int64_t n0 = (j - 1) * Ny1 + i - 1;
                const int64_t nn[4] = {n0, n0 + 1, n0 + Ny1, n0 + Ny1 + 1};
                for (int c = 0; c < 4; ++c) {
                    RHOCPSUM[nn[c]] += p[P_RHOCP] * w[c];
                    TKSUM[nn[c]] += p[P_TKRHOCP] * w[c];
                    WTPSUM[nn[c]] += 1.0 * w[c];
                }
            }
            tk1_of(g, TKSUM, RHOCPSUM, WTPSUM, tk1);
        };
        for (int64_t step = 0; step < nsteps; ++step) {
            scatter();
            std::vector<DT_> tk2 = tk1;  // no heat solve inside the map
            for (int64_t m = 0; m < marknum; ++m) {
                int64_t i, j;
                double w[4];
                if (step == 0) {  // update_marker_temperature!(..., timestep = 1)
                    fix_weights<double>(x[m], y[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny,
                                        &i, &j, w);
                    const int64_t n0 = (j - 1) * Ny1 + i - 1;
                    T[m] = tk2[n0] * w[0] + tk2[n0 + 1] * w[1] + tk2[n0 + Ny1] * w[2] +
                           tk2[n0 + Ny1 + 1] * w[3];
                }
                // RK4 (E:4990-5257): the path in Float64 (positions do not
                // depend on tkm or phim; with wyx = 0 the stresses are
                // unchanged and do not enter the loss), the temperature
                // correction with tangents.  Start-point tk2 (P stencil):
                double xs = x[m], ys = y[m];
                fix_weights<double>(xs, ys, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
                int64_t n0 = (j - 1) * Ny1 + i - 1;
                DT_ tks0 = tk2[n0] * w[0] + tk2[n0 + 1] * w[1] + tk2[n0 + Ny1] * w[2] +
                           tk2[n0 + Ny1 + 1] * w[3];
                RK4Grids<double> Gd{{vx}, {vy}, {vx}, {vy}, {nullptr}, {nullptr}};
                {
                    const double* b = M->brk4;
                    const double* c = M->crk4;
                    double xmm = xs, ymm = ys, xr = xs, yr = ys, vxr[4], vyr[4];
                    for (int rk = 0; rk < 4; ++rk) {
                        vxr[rk] = 0.0 + rk_vx<double>(g, Gd.vx, Gd.vx, xr, yr);
                        vyr[rk] = 0.0 + rk_vy<double>(g, Gd.vy, xr, yr);
                        if (rk < 3) {
                            xr = xmm + dt * c[rk] * vxr[rk];
                            yr = ymm + dt * c[rk] * vyr[rk];
                        }
                    }
                    xs = xs + dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
                    ys = ys + dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
                    xmm = xr = xs;
                    ymm = yr = ys;
                    for (int rk = 0; rk < 4; ++rk) {
                        vxr[rk] = 0.0 + rk_vx<double>(g, Gd.vxf, Gd.vx, xr, yr);
                        vyr[rk] = 0.0 + rk_vy<double>(g, Gd.vyf, xr, yr);
                        if (rk < 3) {
                            xr = xmm - dt * c[rk] * vxr[rk];
                            yr = ymm - dt * c[rk] * vyr[rk];
                        }
                    }
                    xr = xmm - dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
                    yr = ymm - dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
                    fix_weights<double>(xr, yr, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i,
                                        &j, w);
                }
                n0 = (j - 1) * Ny1 + i - 1;
                DT_ tkf0 = tk2[n0] * w[0] + tk2[n0 + 1] * w[1] + tk2[n0 + Ny1] * w[2] +
                           tk2[n0 + Ny1 + 1] * w[3];
                DT_ d = tkf0 - tks0;
                const double rcs = M->rhocpsolidm[tm[m] - 1];
                DT_ rcf = rhocpfluidm<DT_>(M, T[m], mode);
                T[m] = ((1.0 - ph[m]) * T[m] * rcs + ph[m] * (T[m] + d) * rcf) /
                       ((1.0 - ph[m]) * rcs + ph[m] * rcf);
                x[m] = xs;
                y[m] = ys;
            }
        }
        scatter();
        DT_ L(0.0);
        for (int64_t k = 0; k < n; ++k) L += tk1[k] * tk1[k];
        if (loss) *loss = L.v;
        for (int k = 0; k < nk; ++k) grad[k0 + k] = L.d[k];
        if (nseed == 0) break;
    }
    return EREB_OK;
}

}  // extern "C"
