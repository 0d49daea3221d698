/*
 * ereb_oracle.c -- CPU Float64 restatement of the Erebus.jl marker-in-cell
 * transfer path.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP
 * library is compared against, and the CPU baseline of the benchmark.  No
 * product code path links, loads or calls it.
 *
 * Every function follows the reference source operator by operator
 * (E: = src/Erebus.jl, C: = src/constants.jl, T: = test/runtests.jl):
 *   - Julia evaluates a*b*c left to right and never contracts a*b+c into an
 *     fma; this file is compiled with -ffp-contract=off and SSE2 doubles.
 *   - `x^2` and `x^3` (literal powers) are x*x and x*x*x; for Float64
 *     exponents Julia 1.10's ^(Float64,Float64) dispatches integral
 *     exponents to pow_body, which returns x*x*x for 3 and the rounded
 *     square for 2, so `^3.0` and `^2.0` are restated the same way.
 *   - inv(d) is 1.0/d; the reference multiplies by inv(d) in fix (E:1457)
 *     but divides in fix_weights (E:1392) and the RK4 stencils (E:5070).
 *   - unsafe_trunc(Int, t) follows x86-64 cvttsd2si: NaN and |t| >= 2^63
 *     give typemin(Int64).
 *
 * Parity pins for each function are listed in the header of
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ereb.h"

#define G(a, i, j, rows) (a)[((int64_t)(j) - 1) * (rows) + ((int64_t)(i) - 1)]

/* ---------------------------------------------------------------- scalars */

static inline int64_t jl_unsafe_trunc(double t) {
    if (!(t >= -9223372036854775808.0 && t < 9223372036854775808.0))
        return INT64_MIN;
    return (int64_t)t;
}

/* Julia max(a, b) for Float64: NaN propagates */
static inline double jl_max(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}

/* total E:1116-1118 */
double orc_total(double solid, double fluid, double phi) {
    return solid * (1.0 - phi) + fluid * phi;
}

/* ktotal E:1226-1234 */
double orc_ktotal(double ks, double kf, double phi) {
    double t = ks * (3.0 * phi - 2.0) + kf * (1.0 - 3.0 * phi);
    return sqrt(ks * kf / 2.0 + (t * t) * (1.0 / 16.0)) - 0.25 * t;
}

/* kphi E:1251-1254 (not on the path; restated for the T:991 pin) */
double orc_kphi(double phim0, double kphim0m, double phimm) {
    double q = phimm * (1.0 / phim0);
    double r = (1.0 - phimm) * (1.0 / (1.0 - phim0));
    double rinv = 1.0 / r;  /* ^-2.0 -> pow_body n=-2: inv(x)*inv(x) */
    return kphim0m * (q * q * q) * (rinv * rinv);
}

/* ηᶠcur_inv_kᵠ E:1273-1275 */
double orc_etaf_inv_kphi(double phim0, double kphir, double phi, double etaf) {
    double q = phim0 * (1.0 / phi);
    double r = (1.0 - phi) * (1.0 / (1.0 - phim0));
    return etaf * (1.0 / kphir) * (q * q * q) * (r * r);
}

/* compute_rhocpfluidm E:2272-2293; *err = EREB_ERR_MODE on unknown mode */
double orc_compute_rhocpfluidm(const ereb_materials* M, double T, int64_t mode,
                               int* err) {
    if (mode == 1) {
        double tf = M->tmfluidphase;
        if (T < tf - 5.0)
            return M->rhoH2Ofi * (7.67 * T);
        else if (T < tf)
            return M->rhoH2Ofi * (7.67 * T + 0.1 * M->Lf);
        else if (T < tf + 5.0)
            return M->rhoH2Of * (4200.0 + 0.1 * M->Lf);
        else if (T < 410.0)
            return M->rhoH2Of * 4200.0;
        else
            return M->rhoH2Of *
                   (-4.67e4 + 333.0 * T - 0.731 * (T * T) + 5.4e-4 * (T * T * T));
    } else if (mode == 9) {
        return M->rhocpfluidm[0];
    }
    if (err) *err = EREB_ERR_MODE;
    return NAN;
}

/* compute_ksolidm E:2311-2322 */
double orc_compute_ksolidm(const ereb_materials* M, double T, int64_t mode,
                           int* err) {
    if (mode == 1) return 0.73 + 1293.0 / (T + 77.0);
    if (mode == 9) return M->ksolidm[0];
    if (err) *err = EREB_ERR_MODE;
    return NAN;
}

/* compute_kfluidm E:2341-2358 */
double orc_compute_kfluidm(const ereb_materials* M, double T, int64_t mode,
                           int* err) {
    if (mode == 1) {
        if (T < M->tmfluidphase)
            return 0.465 + 488.0 / T;
        else if (T < 410.0)
            return -0.581 + 6.34e-3 * T - 7.93e-6 * (T * T);
        else
            return -0.142 + 4.12e-3 * T - 5.01e-6 * (T * T);
    }
    if (mode == 9) return M->kfluidm[0];
    if (err) *err = EREB_ERR_MODE;
    return NAN;
}

/* etatotal_rocks E:1312-1318 */
double orc_etatotal_rocks(const ereb_materials* M, double T, int64_t tm) {
    double es = T > M->tmsolidphase ? M->etasolidmm[tm - 1] : M->etasolidm[tm - 1];
    double ef = T > M->tmfluidphase ? M->etafluidmm[tm - 1] : M->etafluidm[tm - 1];
    return jl_max(jl_max(M->etamin, es), ef);
}

/* Q_radiogenic E:1294-1296 */
double orc_q_radiogenic(double f, double ratio, double E, double tau,
                        double time) {
    return f * ratio * E * exp(-time * (1.0 / tau)) * (1.0 / tau);
}

/* ------------------------------------------------------- interpolation */

/* fix E:1456-1470 */
void orc_fix(double x, double y, const double* xa, const double* ya, double dx,
             double dy, int64_t jmin, int64_t jmax, int64_t imin, int64_t imax,
             int64_t* ri, int64_t* rj) {
    int64_t j = jl_unsafe_trunc((x - xa[0]) * (1.0 / dx)) + 1;
    int64_t i = jl_unsafe_trunc((y - ya[0]) * (1.0 / dy)) + 1;
    if (j < jmin) j = jmin; else if (j > jmax) j = jmax;
    if (i < imin) i = imin; else if (i > imax) i = imax;
    *ri = i;
    *rj = j;
}

/* fix_distances E:1424-1431 */
void orc_fix_distances(double x, double y, const double* xa, const double* ya,
                       double dx, double dy, int64_t jmin, int64_t jmax,
                       int64_t imin, int64_t imax, int64_t* ri, int64_t* rj,
                       double* dxmj, double* dymi) {
    orc_fix(x, y, xa, ya, dx, dy, jmin, jmax, imin, imax, ri, rj);
    *dxmj = x - xa[*rj - 1];
    *dymi = y - ya[*ri - 1];
}

/* fix_weights E:1388-1397 */
void orc_fix_weights(double x, double y, const double* xa, const double* ya,
                     double dx, double dy, int64_t jmin, int64_t jmax,
                     int64_t imin, int64_t imax, int64_t* ri, int64_t* rj,
                     double* w) {
    double dxmj, dymi;
    orc_fix_distances(x, y, xa, ya, dx, dy, jmin, jmax, imin, imax, ri, rj,
                      &dxmj, &dymi);
    w[0] = (1.0 - dymi / dy) * (1.0 - dxmj / dx);
    w[1] = (dymi / dy) * (1.0 - dxmj / dx);
    w[2] = (1.0 - dymi / dy) * (dxmj / dx);
    w[3] = (dymi / dy) * (dxmj / dx);
}

/* dot4(grid_vector(i, j, grid), weights) E:1137-1159 */
static inline double gdot4(const double* g, int64_t rows, int64_t i, int64_t j,
                           const double* w) {
    return G(g, i, j, rows) * w[0] + G(g, i + 1, j, rows) * w[1] +
           G(g, i, j + 1, rows) * w[2] + G(g, i + 1, j + 1, rows) * w[3];
}

/* interpolate_add_to_grid! E:1517-1525 */
static inline void add4(double* g, int64_t rows, int64_t i, int64_t j,
                        const double* w, double p) {
    G(g, i, j, rows) += p * w[0];
    G(g, i + 1, j, rows) += p * w[1];
    G(g, i, j + 1, rows) += p * w[2];
    G(g, i + 1, j + 1, rows) += p * w[3];
}

/* the same with |p * w|: the scale of a sum taken in another order */
static inline void add4abs(double* g, int64_t rows, int64_t i, int64_t j,
                           const double* w, double p) {
    G(g, i, j, rows) += fabs(p * w[0]);
    G(g, i + 1, j, rows) += fabs(p * w[1]);
    G(g, i, j + 1, rows) += fabs(p * w[2]);
    G(g, i + 1, j + 1, rows) += fabs(p * w[3]);
}

/* ------------------------------------------------------- step 1 + 2 */

/* compute_marker_properties! E:700-750 for one marker (0-based m) */
static int marker_properties_one(const ereb_materials* M, int64_t m,
    const int64_t* tm, const double* tkm, double* rhototalm,
    double* rhocptotalm, double* etatotalm, double* hrtotalm, double* ktotalm,
    double* tkm_rhocptotalm, double* etafluidcur_inv_kphim,
    const double* hrsolidm, const double* hrfluidm, const double* phim,
    const double* XW, int64_t mode) {
    int err = 0;
    int64_t t = tm[m];
    double T = tkm[m];
    double etafluidcur;
    if (t < 3) {
        double XD = 1.0 - XW[m];
        double rhosolidm0 = (M->MD + M->MH2O * XW[m]) / (M->VDs * XD + M->VWs * XW[m]);
        double rhofluidm0 = T > M->tmfluidphase ? M->rhoH2Of : M->rhoH2Ofi;
        rhototalm[m] = orc_total(rhosolidm0, rhofluidm0, phim[m]);
        rhocptotalm[m] = orc_total(M->rhocpsolidm[t - 1],
                                   orc_compute_rhocpfluidm(M, T, mode, &err), phim[m]);
        double etasolidcur = T > M->tmsolidphase ? M->etasolidmm[t - 1] : M->etasolidm[t - 1];
        etafluidcur = T > M->tmfluidphase ? M->etafluidmm[t - 1] : M->etafluidm[t - 1];
        etatotalm[m] = jl_max(jl_max(M->etamin, etasolidcur), etafluidcur);
        hrtotalm[m] = orc_total(hrsolidm[t - 1], hrfluidm[t - 1], phim[m]);
        ktotalm[m] = orc_ktotal(orc_compute_ksolidm(M, T, mode, &err),
                                orc_compute_kfluidm(M, T, mode, &err), phim[m]);
    } else {
        etafluidcur = M->etafluidm[t - 1];
    }
    tkm_rhocptotalm[m] = T * rhocptotalm[m];
    etafluidcur_inv_kphim[m] =
        orc_etaf_inv_kphi(M->phim0, M->kphim0[t - 1], phim[m], etafluidcur);
    return err;
}

int orc_compute_marker_properties(const ereb_materials* M, int64_t marknum,
    const int64_t* tm, const double* tkm, double* rhototalm,
    double* rhocptotalm, double* etatotalm, double* hrtotalm, double* ktotalm,
    double* tkm_rhocptotalm, double* etafluidcur_inv_kphim,
    const double* hrsolidm, const double* hrfluidm, const double* phim,
    const double* XW, int64_t mode) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    for (int64_t m = 0; m < marknum; ++m)
        marker_properties_one(M, m, tm, tkm, rhototalm, rhocptotalm, etatotalm,
                              hrtotalm, ktotalm, tkm_rhocptotalm,
                              etafluidcur_inv_kphim, hrsolidm, hrfluidm, phim,
                              XW, mode);
    return EREB_OK;
}

/* The marker loop E:6376-6517: reset_interpolated_properties! (E:968) then
 * for m = 1:marknum compute_marker_properties! and
 * marker_to_{basic,vx,vy,p}_nodes! (E:1605-1863).  sums[] in the order of
 * reset_interpolated_properties!: ETA0SUM ETASUM GGGSUM SXYSUM COHSUM
 * TENSUM FRISUM WTSUM | RHOXSUM RHOFXSUM KXSUM PHIXSUM RXSUM WTXSUM |
 * RHOYSUM RHOFYSUM KYSUM PHIYSUM RYSUM WTYSUM | RHOSUM RHOCPSUM ALPHASUM
 * ALPHAFSUM HRSUM GGGPSUM SXXSUM TKSUM PHISUM WTPSUM */
static int marker_to_nodes_impl(const ereb_geometry* g, const ereb_materials* M,
    const double* xm, const double* ym, const int64_t* tm, const double* tkm,
    const double* phim, const double* XW, const double* etavpm,
    const double* inv_gggtotalm, const double* fricttotalm,
    const double* cohestotalm, const double* tenstotalm,
    const double* rhofluidcur, const double* alphasolidcur,
    const double* alphafluidcur, const double* sxxm, const double* sxym,
    double* rhototalm, double* rhocptotalm, double* etatotalm,
    double* hrtotalm, double* ktotalm, double* tkm_rhocptotalm,
    double* etafluidcur_inv_kphim, const double* hrsolidm,
    const double* hrfluidm, int64_t mode, double* const* sums,
    int64_t marknum, int absval) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    void (*add)(double*, int64_t, int64_t, int64_t, const double*, double) =
        absval ? add4abs : add4;
    const int64_t Nx = g->Nx, Ny = g->Ny, Nx1 = Nx + 1, Ny1 = Ny + 1;
    for (int k = 0; k < 8; ++k) memset(sums[k], 0, sizeof(double) * Nx * Ny);
    for (int k = 8; k < 30; ++k) memset(sums[k], 0, sizeof(double) * Nx1 * Ny1);
    for (int64_t m = 0; m < marknum; ++m) {
        marker_properties_one(M, m, tm, tkm, rhototalm, rhocptotalm, etatotalm,
                              hrtotalm, ktotalm, tkm_rhocptotalm,
                              etafluidcur_inv_kphim, hrsolidm, hrfluidm, phim,
                              XW, mode);
        int64_t i, j;
        double w[4];
        /* marker_to_basic_nodes! E:1605-1648 */
        orc_fix_weights(xm[m], ym[m], g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1,
                        Ny - 1, &i, &j, w);
        add(sums[0], Ny, i, j, w, etatotalm[m]);
        add(sums[1], Ny, i, j, w, etavpm[m]);
        add(sums[2], Ny, i, j, w, inv_gggtotalm[m]);
        add(sums[3], Ny, i, j, w, sxym[m]);
        add(sums[4], Ny, i, j, w, cohestotalm[m]);
        add(sums[5], Ny, i, j, w, tenstotalm[m]);
        add(sums[6], Ny, i, j, w, fricttotalm[m]);
        add(sums[7], Ny, i, j, w, 1.0);
        /* marker_to_vx_nodes! E:1676-1713 */
        orc_fix_weights(xm[m], ym[m], g->xvx, g->yvx, g->dx, g->dy, 1, Nx - 1, 1,
                        Ny, &i, &j, w);
        add(sums[8], Ny1, i, j, w, rhototalm[m]);
        add(sums[9], Ny1, i, j, w, rhofluidcur[m]);
        add(sums[10], Ny1, i, j, w, ktotalm[m]);
        add(sums[11], Ny1, i, j, w, phim[m]);
        add(sums[12], Ny1, i, j, w, etafluidcur_inv_kphim[m]);
        add(sums[13], Ny1, i, j, w, 1.0);
        /* marker_to_vy_nodes! E:1741-1778 */
        orc_fix_weights(xm[m], ym[m], g->xvy, g->yvy, g->dx, g->dy, 1, Nx, 1,
                        Ny - 1, &i, &j, w);
        add(sums[14], Ny1, i, j, w, rhototalm[m]);
        add(sums[15], Ny1, i, j, w, rhofluidcur[m]);
        add(sums[16], Ny1, i, j, w, ktotalm[m]);
        add(sums[17], Ny1, i, j, w, phim[m]);
        add(sums[18], Ny1, i, j, w, etafluidcur_inv_kphim[m]);
        add(sums[19], Ny1, i, j, w, 1.0);
        /* marker_to_p_nodes! E:1814-1863 */
        orc_fix_weights(xm[m], ym[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny,
                        &i, &j, w);
        add(sums[25], Ny1, i, j, w, inv_gggtotalm[m]); /* GGGPSUM */
        add(sums[26], Ny1, i, j, w, sxxm[m]);          /* SXXSUM */
        add(sums[20], Ny1, i, j, w, rhototalm[m]);     /* RHOSUM */
        add(sums[21], Ny1, i, j, w, rhocptotalm[m]);   /* RHOCPSUM */
        add(sums[22], Ny1, i, j, w, alphasolidcur[m]); /* ALPHASUM */
        add(sums[23], Ny1, i, j, w, alphafluidcur[m]); /* ALPHAFSUM */
        add(sums[24], Ny1, i, j, w, hrtotalm[m]);      /* HRSUM */
        add(sums[28], Ny1, i, j, w, phim[m]);          /* PHISUM */
        add(sums[27], Ny1, i, j, w, tkm_rhocptotalm[m]); /* TKSUM */
        add(sums[29], Ny1, i, j, w, 1.0);              /* WTPSUM */
    }
    return EREB_OK;
}

int orc_marker_to_nodes(const ereb_geometry* g, const ereb_materials* M,
    const double* xm, const double* ym, const int64_t* tm, const double* tkm,
    const double* phim, const double* XW, const double* etavpm,
    const double* inv_gggtotalm, const double* fricttotalm,
    const double* cohestotalm, const double* tenstotalm,
    const double* rhofluidcur, const double* alphasolidcur,
    const double* alphafluidcur, const double* sxxm, const double* sxym,
    double* rhototalm, double* rhocptotalm, double* etatotalm,
    double* hrtotalm, double* ktotalm, double* tkm_rhocptotalm,
    double* etafluidcur_inv_kphim, const double* hrsolidm,
    const double* hrfluidm, int64_t mode, double* const* sums,
    int64_t marknum) {
    return marker_to_nodes_impl(g, M, xm, ym, tm, tkm, phim, XW, etavpm, inv_gggtotalm,
                                fricttotalm, cohestotalm, tenstotalm, rhofluidcur,
                                alphasolidcur, alphafluidcur, sxxm, sxym, rhototalm,
                                rhocptotalm, etatotalm, hrtotalm, ktotalm, tkm_rhocptotalm,
                                etafluidcur_inv_kphim, hrsolidm, hrfluidm, mode, sums,
                                marknum, 0);
}

/* Per node sum of |value * weight| over the same loop: the scale against
 * which the GPU's reordered sums are compared (tests only) */
int orc_marker_to_nodes_abs(const ereb_geometry* g, const ereb_materials* M,
    const double* xm, const double* ym, const int64_t* tm, const double* tkm,
    const double* phim, const double* XW, const double* etavpm,
    const double* inv_gggtotalm, const double* fricttotalm,
    const double* cohestotalm, const double* tenstotalm,
    const double* rhofluidcur, const double* alphasolidcur,
    const double* alphafluidcur, const double* sxxm, const double* sxym,
    double* rhototalm, double* rhocptotalm, double* etatotalm,
    double* hrtotalm, double* ktotalm, double* tkm_rhocptotalm,
    double* etafluidcur_inv_kphim, const double* hrsolidm,
    const double* hrfluidm, int64_t mode, double* const* sums,
    int64_t marknum) {
    return marker_to_nodes_impl(g, M, xm, ym, tm, tkm, phim, XW, etavpm, inv_gggtotalm,
                                fricttotalm, cohestotalm, tenstotalm, rhofluidcur,
                                alphasolidcur, alphafluidcur, sxxm, sxym, rhototalm,
                                rhocptotalm, etatotalm, hrtotalm, ktotalm, tkm_rhocptotalm,
                                etafluidcur_inv_kphim, hrsolidm, hrfluidm, mode, sums,
                                marknum, 1);
}

/* --------------------------------------------------- node properties */

/* compute_basic_node_properties! E:1970-2007 */
void orc_compute_basic_node_properties(const ereb_geometry* g,
    const double* ETA0SUM, const double* ETASUM, const double* GGGSUM,
    const double* SXYSUM, const double* COHSUM, const double* TENSUM,
    const double* FRISUM, const double* WTSUM, double* ETA0, double* ETA,
    double* GGG, double* SXY0, double* COH, double* TEN, double* FRI,
    uint8_t* YNY) {
    int64_t n = g->Nx * g->Ny;
    for (int64_t k = 0; k < n; ++k) {
        double wt = WTSUM[k];
        if (wt > 0.0) {
            double iw = 1.0 / wt;
            ETA0[k] = ETA0SUM[k] * iw;
            ETA[k] = ETASUM[k] * iw;
            if (ETA[k] < ETA0[k]) YNY[k] = 1;
            GGG[k] = (1.0 / GGGSUM[k]) * wt;
            SXY0[k] = SXYSUM[k] * iw;
            COH[k] = COHSUM[k] * iw;
            TEN[k] = TENSUM[k] * iw;
            FRI[k] = FRISUM[k] * iw;
        }
    }
}

/* compute_vx_node_properties! E:2033-2060 (also Vy, E:2086-2113) */
void orc_compute_v_node_properties(const ereb_geometry* g, const double* RHOSUM,
    const double* RHOFSUM, const double* KSUM, const double* PHISUM,
    const double* RSUM, const double* WTSUM, double* RHO, double* RHOF,
    double* K, double* PHI, double* R) {
    int64_t n = (g->Nx + 1) * (g->Ny + 1);
    for (int64_t k = 0; k < n; ++k) {
        double wt = WTSUM[k];
        if (wt > 0.0) {
            double iw = 1.0 / wt;
            RHO[k] = RHOSUM[k] * iw;
            RHOF[k] = RHOFSUM[k] * iw;
            K[k] = KSUM[k] * iw;
            PHI[k] = PHISUM[k] * iw;
            R[k] = RSUM[k] * iw;
        }
    }
}

/* compute_p_node_properties! E:2148-2189 */
void orc_compute_p_node_properties(const ereb_geometry* g, const double* RHOSUM,
    const double* RHOCPSUM, const double* ALPHASUM, const double* ALPHAFSUM,
    const double* HRSUM, const double* GGGPSUM, const double* SXXSUM,
    const double* TKSUM, const double* PHISUM, const double* WTPSUM,
    double* RHO, double* RHOCP, double* ALPHA, double* ALPHAF, double* HR,
    double* GGGP, double* SXX0, double* tk1, double* PHI, double* BETAPHI) {
    int64_t n = (g->Nx + 1) * (g->Ny + 1);
    for (int64_t k = 0; k < n; ++k) {
        double wt = WTPSUM[k];
        if (wt > 0.0) {
            double iw = 1.0 / wt;
            RHO[k] = RHOSUM[k] * iw;
            RHOCP[k] = RHOCPSUM[k] * iw;
            ALPHA[k] = ALPHASUM[k] * iw;
            ALPHAF[k] = ALPHAFSUM[k] * iw;
            HR[k] = HRSUM[k] * iw;
            GGGP[k] = (1.0 / GGGPSUM[k]) * wt;
            SXX0[k] = SXXSUM[k] * iw;
            tk1[k] = TKSUM[k] * (1.0 / RHOCPSUM[k]);
            PHI[k] = PHISUM[k] * iw;
            BETAPHI[k] = (1.0 / GGGP[k]) * PHI[k];
        }
    }
}

/* apply_insulating_boundary_conditions! E:2432-2449 */
void orc_apply_insulating_bc(double* t, int64_t Nyy, int64_t Nxx) {
    if (!(Nyy > 2 && Nxx > 2)) return;
    for (int64_t j = 2; j <= Nxx - 1; ++j) G(t, 1, j, Nyy) = G(t, 2, j, Nyy);
    for (int64_t j = 2; j <= Nxx - 1; ++j) G(t, Nyy, j, Nyy) = G(t, Nyy - 1, j, Nyy);
    for (int64_t i = 1; i <= Nyy; ++i) G(t, i, 1, Nyy) = G(t, i, 2, Nyy);
    for (int64_t i = 1; i <= Nyy; ++i) G(t, i, Nxx, Nyy) = G(t, i, Nxx - 1, Nyy);
}

/* ------------------------------------------------------------ step 3 */

/* update_marker_temperature! E:4782-4801 */
void orc_update_marker_temperature(const ereb_geometry* g, const double* xm,
    const double* ym, double* tkm, const double* DT, const double* tk2,
    int64_t timestep, int64_t marknum) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    /* the reference threads this loop (@threads, E:4786/E:4793); markers
     * are independent, so the result does not depend on the thread count */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t m = 0; m < marknum; ++m) {
        int64_t i, j;
        double w[4];
        orc_fix_weights(xm[m], ym[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny,
                        &i, &j, w);
        if (timestep == 1)
            tkm[m] = gdot4(tk2, Ny1, i, j, w);
        else
            tkm[m] += gdot4(DT, Ny1, i, j, w);
    }
}

/* update_marker_stress! E:4354-4388 (P clamps 2..Nx-1, 2..Ny-1) */
void orc_update_marker_stress(const ereb_geometry* g, const double* xm,
    const double* ym, double* sxxm, double* sxym, const double* DSXX,
    const double* DSXY, int64_t marknum) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    /* @threads in the reference (E:4356) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t m = 0; m < marknum; ++m) {
        int64_t ip, jp, ib, jb;
        double wp[4], wb[4];
        orc_fix_weights(xm[m], ym[m], g->xp, g->yp, g->dx, g->dy, 2, Nx - 1, 2,
                        Ny - 1, &ip, &jp, wp);
        orc_fix_weights(xm[m], ym[m], g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1,
                        Ny - 1, &ib, &jb, wb);
        sxxm[m] += gdot4(DSXX, Ny1, ip, jp, wp);
        sxym[m] += gdot4(DSXY, Ny, ib, jb, wb);
    }
}

/* ------------------------------------------------------------ step 4 */

/* one RK4 stage velocity on the Vx grid (E:5057-5091); `vc` is the grid
 * whose values feed the second-order correction at the (i, j) centre of the
 * left-half branch, which the reference takes from vx also in the fluid
 * loop (E:5173) */
static double rk_vx(const ereb_geometry* g, const double* v, const double* vc,
                    double x, double y) {
    const int64_t Nx = g->Nx, Ny = g->Ny, R = Ny + 1;
    int64_t i, j;
    double dxmj, dymi;
    orc_fix_distances(x, y, g->xvx, g->yvx, g->dx, g->dy, 1, Nx - 1, 1, Ny, &i,
                      &j, &dxmj, &dymi);
    double dx = g->dx, dy = g->dy;
    double v13 = G(v, i, j, R) * (1.0 - dxmj / dx) + G(v, i, j + 1, R) * dxmj / dx;
    double v24 = G(v, i + 1, j, R) * (1.0 - dxmj / dx) + G(v, i + 1, j + 1, R) * dxmj / dx;
    if (dxmj / dx >= 0.5) {
        if (j < Nx - 1) {
            double t = dxmj / dx - 0.5;
            v13 += 0.5 * (t * t) * (G(v, i, j, R) - 2.0 * G(v, i, j + 1, R) + G(v, i, j + 2, R));
            v24 += 0.5 * (t * t) * (G(v, i + 1, j, R) - 2.0 * G(v, i + 1, j + 1, R) + G(v, i + 1, j + 2, R));
        }
    } else {
        if (j > 1) {
            double t = dxmj / dx - 0.5;
            v13 += 0.5 * (t * t) * (G(v, i, j - 1, R) - 2.0 * G(vc, i, j, R) + G(v, i, j + 1, R));
            v24 += 0.5 * (t * t) * (G(v, i + 1, j - 1, R) - 2.0 * G(v, i + 1, j, R) + G(v, i + 1, j + 1, R));
        }
    }
    return v13 * (1.0 - dymi / dy) + v24 * dymi / dy;
}

/* one RK4 stage velocity on the Vy grid (E:5093-5127) */
static double rk_vy(const ereb_geometry* g, const double* v, double x,
                    double y) {
    const int64_t Nx = g->Nx, Ny = g->Ny, R = Ny + 1;
    int64_t i, j;
    double dxmj, dymi;
    orc_fix_distances(x, y, g->xvy, g->yvy, g->dx, g->dy, 1, Nx, 1, Ny - 1, &i,
                      &j, &dxmj, &dymi);
    double dx = g->dx, dy = g->dy;
    double v12 = G(v, i, j, R) * (1.0 - dymi / dy) + G(v, i + 1, j, R) * dymi / dy;
    double v34 = G(v, i, j + 1, R) * (1.0 - dymi / dy) + G(v, i + 1, j + 1, R) * dymi / dy;
    if (dymi / dy >= 0.5) {
        if (i < Ny - 1) {
            double t = dymi / dy - 0.5;
            v12 += 0.5 * (t * t) * (G(v, i, j, R) - 2.0 * G(v, i + 1, j, R) + G(v, i + 2, j, R));
            v34 += 0.5 * (t * t) * (G(v, i, j + 1, R) - 2.0 * G(v, i + 1, j + 1, R) + G(v, i + 2, j + 1, R));
        }
    } else {
        if (i > 1) {
            double t = dymi / dy - 0.5;
            v12 += 0.5 * (t * t) * (G(v, i - 1, j, R) - 2.0 * G(v, i, j, R) + G(v, i + 1, j, R));
            v34 += 0.5 * (t * t) * (G(v, i - 1, j + 1, R) - 2.0 * G(v, i, j + 1, R) + G(v, i + 1, j + 1, R));
        }
    }
    return v12 * (1.0 - dxmj / dx) + v34 * dxmj / dx;
}

/* move_markers_rk4! E:4990-5257 */
int orc_move_markers_rk4(const ereb_geometry* g, const ereb_materials* M,
    double* xm, double* ym, const int64_t* tm, double* tkm, const double* phim,
    double* sxym, double* sxxm, const double* vx, const double* vy,
    const double* vxf, const double* vyf, const double* wyx, const double* tk2,
    int64_t marknum, double dt, int64_t mode) {
    if (mode != 1 && mode != 9) return EREB_ERR_MODE;
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    const double* b = M->brk4;
    const double* c = M->crk4;
    for (int64_t m = 0; m < marknum; ++m) {
        double xmm = xm[m], ymm = ym[m];
        double xr = xmm, yr = ymm;
        int64_t i, j;
        double w[4];
        orc_fix_weights(xr, yr, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
        double tks0 = gdot4(tk2, Ny1, i, j, w);
        orc_fix_weights(xr, yr, g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1, Ny - 1, &i, &j, w);
        double om = gdot4(wyx, Ny, i, j, w);
        double th = dt * om;
        double s = sxxm[m], t = sxym[m];
        double ct = cos(th), st = sin(th), s2 = sin(2.0 * th), c2 = cos(2.0 * th);
        double sxx1 = s * (ct * ct) - s * (st * st) - t * s2;
        double sxy1 = s * s2 + t * c2;
        sxxm[m] = sxx1;
        sxym[m] = sxy1;
        double vxr[4], vyr[4];
        for (int rk = 0; rk < 4; ++rk) {
            vxr[rk] = 0.0 + rk_vx(g, vx, vx, xr, yr);
            vyr[rk] = 0.0 + rk_vy(g, vy, xr, yr);
            if (rk < 3) {
                xr = xmm + dt * c[rk] * vxr[rk];
                yr = ymm + dt * c[rk] * vyr[rk];
            }
        }
        xm[m] += dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
        ym[m] += dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
        xmm = xr = xm[m];
        ymm = yr = ym[m];
        for (int rk = 0; rk < 4; ++rk) {
            vxr[rk] = 0.0 + rk_vx(g, vxf, vx, xr, yr);
            vyr[rk] = 0.0 + rk_vy(g, vyf, xr, yr);
            if (rk < 3) {
                xr = xmm - dt * c[rk] * vxr[rk];
                yr = ymm - dt * c[rk] * vyr[rk];
            }
        }
        xr = xmm - dt * (b[0] * vxr[0] + b[1] * vxr[1] + b[2] * vxr[2] + b[3] * vxr[3]);
        yr = ymm - dt * (b[0] * vyr[0] + b[1] * vyr[1] + b[2] * vyr[2] + b[3] * vyr[3]);
        orc_fix_weights(xr, yr, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
        double tkf0 = gdot4(tk2, Ny1, i, j, w);
        double d = tkf0 - tks0;
        int64_t tt = tm[m];
        double T = tkm[m], ph = phim[m];
        double rcs = M->rhocpsolidm[tt - 1];
        double rcf = orc_compute_rhocpfluidm(M, T, mode, NULL);
        tkm[m] = ((1.0 - ph) * T * rcs + ph * (T + d) * rcf) /
                 ((1.0 - ph) * rcs + ph * rcf);
    }
    return EREB_OK;
}

/* ------------------------------------------------------------ step 5 */

/* Row-wise builder mirroring ExtendableSparse: updateindex!(L, +, v, r, c)
 * creates no entry when v == 0 (its documented zero skip), entries are
 * unique per (row, col) in these assemblies, flush! yields CSC with rows
 * ascending in each column. */
typedef struct {
    int64_t n, cnt, cap;
    int64_t* row;
    int64_t* col;
    double* val;
    int overflow;
} coo_t;

static void coo_add(coo_t* L, double v, int64_t r, int64_t c) {
    if (v == 0.0) return; /* -0.0 == 0.0 as in iszero */
    if (L->cnt >= L->cap) { L->overflow = 1; return; }
    L->row[L->cnt] = r;
    L->col[L->cnt] = c;
    L->val[L->cnt] = v;
    L->cnt++;
}

/* COO (row-major generation order) -> 1-based CSC, rows ascending per col */
static int64_t coo_to_csc(const coo_t* L, int64_t* colptr, int64_t* rowval,
                          double* nzval) {
    int64_t n = L->n;
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < L->cnt; ++k) cnt[L->col[k]]++;  /* cols 1-based */
    colptr[0] = 1;
    for (int64_t c = 1; c <= n; ++c) colptr[c] = colptr[c - 1] + cnt[c];
    int64_t* fill = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    /* generation order is by row ascending except within a row; collect then
       sort each column by row (columns hold at most a few dozen entries) */
    for (int64_t k = 0; k < L->cnt; ++k) {
        int64_t c = L->col[k];
        int64_t p = colptr[c - 1] - 1 + fill[c]++;
        rowval[p] = L->row[k];
        nzval[p] = L->val[k];
    }
    for (int64_t c = 1; c <= n; ++c) {
        int64_t a = colptr[c - 1] - 1, e = colptr[c] - 1;
        for (int64_t p = a + 1; p < e; ++p) {
            int64_t r = rowval[p];
            double v = nzval[p];
            int64_t q = p - 1;
            while (q >= a && rowval[q] > r) {
                rowval[q + 1] = rowval[q];
                nzval[q + 1] = nzval[q];
                --q;
            }
            rowval[q + 1] = r;
            nzval[q + 1] = v;
        }
    }
    free(cnt);
    free(fill);
    return L->cnt;
}

/* assemble_hydromechanical_lse! E:3089-3572.  Output CSC arrays need room
 * for `cap` entries; returns nnz, or -1 when cap is too small. */
int64_t orc_assemble_hydromechanical(const ereb_geometry* g,
    const ereb_materials* M, const double* ETA, const double* ETAP,
    const double* GGG, const double* GGGP, const double* SXY0,
    const double* SXX0, const double* RHOX, const double* RHOY,
    const double* RHOFX, const double* RHOFY, const double* RX,
    const double* RY, const double* ETAPHI, const double* BETAPHI,
    const double* PHI, const double* gx, const double* gy, const double* pr0,
    const double* pf0, const double* DMP, double dt, double* R,
    int64_t* colptr, int64_t* rowval, double* nzval, int64_t cap) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Nx1 = Nx + 1, Ny1 = Ny + 1;
    const double dx = g->dx, dy = g->dy, Kcont = M->Kcont;
    const int64_t n = Nx1 * Ny1 * 6;
    coo_t L = {n, 0, cap, (int64_t*)malloc(sizeof(int64_t) * cap),
               (int64_t*)malloc(sizeof(int64_t) * cap),
               (double*)malloc(sizeof(double) * cap), 0};
    for (int64_t k = 0; k < n; ++k) R[k] = 0.0;
#define B_(a, i, j) G(a, i, j, Ny)
#define P_(a, i, j) G(a, i, j, Ny1)
    for (int64_t j = 1; j <= Nx1; ++j) {
        for (int64_t i = 1; i <= Ny1; ++i) {
            int64_t kvx = ((j - 1) * Ny1 + i - 1) * 6 + 1;
            int64_t kvy = kvx + 1, kpm = kvx + 2, kqx = kvx + 3, kqy = kvx + 4,
                    kpf = kvx + 5;
            /* Vx equation E:3128-3262 */
            if (i == 1 || i == Ny1 || j == 1 || j == Nx || j == Nx1) {
                coo_add(&L, 1.0, kvx, kvx);
                if (j == 1) R[kvx - 1] = M->vxleft;
                if (j == Nx) R[kvx - 1] = M->vxright;
                if (i == 1 && 1 < j && j < Nx) coo_add(&L, M->bctop, kvx, kvx + 6);
                if (i == Ny1 && 1 < j && j < Nx) coo_add(&L, M->bcbottom, kvx, kvx - 6);
            } else {
                double ETA1 = B_(ETA, i - 1, j) * B_(GGG, i - 1, j) * dt /
                              (B_(GGG, i - 1, j) * dt + B_(ETA, i - 1, j));
                double ETA2 = B_(ETA, i, j) * B_(GGG, i, j) * dt /
                              (B_(GGG, i, j) * dt + B_(ETA, i, j));
                double ETAP1 = P_(ETAP, i, j) * P_(GGGP, i, j) * dt /
                               (P_(GGGP, i, j) * dt + P_(ETAP, i, j));
                double ETAP2 = P_(ETAP, i, j + 1) * P_(GGGP, i, j + 1) * dt /
                               (P_(GGGP, i, j + 1) * dt + P_(ETAP, i, j + 1));
                double SXY1 = B_(SXY0, i - 1, j) * B_(ETA, i - 1, j) /
                              (B_(GGG, i - 1, j) * dt + B_(ETA, i - 1, j));
                double SXY2 = B_(SXY0, i, j) * B_(ETA, i, j) /
                              (B_(GGG, i, j) * dt + B_(ETA, i, j));
                double SXX1 = P_(SXX0, i, j) * P_(ETAP, i, j) /
                              (P_(GGGP, i, j) * dt + P_(ETAP, i, j));
                double SXX2 = P_(SXX0, i, j + 1) * P_(ETAP, i, j + 1) /
                              (P_(GGGP, i, j + 1) * dt + P_(ETAP, i, j + 1));
                double dRdx = 0.5 * (P_(RHOX, i, j + 1) - P_(RHOX, i, j - 1)) * (1.0 / dx);
                double dRdy = 0.5 * (P_(RHOX, i + 1, j) - P_(RHOX, i - 1, j)) * (1.0 / dy);
                double gxij = P_(gx, i, j);
                double idx = 1.0 / dx, idy = 1.0 / dy;
                coo_add(&L, ETAP1 / (dx * dx), kvx, kvx - 6 * Ny1);
                coo_add(&L, ETA1 / (dy * dy), kvx, kvx - 6);
                coo_add(&L, -(ETAP1 + ETAP2) * (1.0 / (dx * dx)) -
                                (ETA1 + ETA2) * (1.0 / (dy * dy)) - dRdx * gxij * dt,
                        kvx, kvx);
                coo_add(&L, ETA2 / (dy * dy), kvx, kvx + 6);
                coo_add(&L, ETAP2 / (dx * dx), kvx, kvx + 6 * Ny1);
                coo_add(&L, ETAP1 * idx * idy - ETA2 * idx * idy - dRdy * gxij * dt * 0.25,
                        kvx, kvy);
                coo_add(&L, -ETAP2 * idx * idy + ETA2 * idx * idy - dRdy * gxij * dt * 0.25,
                        kvx, kvy + 6 * Ny1);
                coo_add(&L, -ETAP1 * idx * idy + ETA1 * idx * idy - dRdy * gxij * dt * 0.25,
                        kvx, kvy - 6);
                coo_add(&L, ETAP2 * idx * idy - ETA1 * idx * idy - dRdy * gxij * dt * 0.25,
                        kvx, kvy + 6 * Ny1 - 6);
                coo_add(&L, Kcont * idx, kvx, kpm);
                coo_add(&L, -Kcont * idx, kvx, kpm + 6 * Ny1);
                R[kvx - 1] = -P_(RHOX, i, j) * gxij - (SXY2 - SXY1) * idy - (SXX2 - SXX1) * idx;
            }
            /* Vy equation E:3264-3396 */
            if (i == 1 || i == Ny || i == Ny1 || j == 1 || j == Nx1) {
                coo_add(&L, 1.0, kvy, kvy);
                if (i == 1) R[kvy - 1] = M->vytop;
                if (i == Ny) R[kvy - 1] = M->vybottom;
                if (j == 1 && 1 < i && i < Ny) coo_add(&L, M->bcleft, kvy, kvy + 6 * Ny1);
                if (j == Nx1 && 1 < i && i < Ny) coo_add(&L, M->bcright, kvy, kvy - 6 * Ny1);
            } else {
                double ETA1 = B_(ETA, i, j - 1) * B_(GGG, i, j - 1) * dt /
                              (B_(GGG, i, j - 1) * dt + B_(ETA, i, j - 1));
                double ETA2 = B_(ETA, i, j) * B_(GGG, i, j) * dt /
                              (B_(GGG, i, j) * dt + B_(ETA, i, j));
                double ETAP1 = P_(ETAP, i, j) * P_(GGGP, i, j) * dt /
                               (P_(GGGP, i, j) * dt + P_(ETAP, i, j));
                double ETAP2 = P_(ETAP, i + 1, j) * P_(GGGP, i + 1, j) * dt /
                               (P_(GGGP, i + 1, j) * dt + P_(ETAP, i + 1, j));
                double SXY1 = B_(SXY0, i, j - 1) * B_(ETA, i, j - 1) /
                              (B_(GGG, i, j - 1) * dt + B_(ETA, i, j - 1));
                double SXY2 = B_(SXY0, i, j) * B_(ETA, i, j) /
                              (B_(GGG, i, j) * dt + B_(ETA, i, j));
                double SYY1 = -P_(SXX0, i, j) * P_(ETAP, i, j) /
                              (P_(GGGP, i, j) * dt + P_(ETAP, i, j));
                double SYY2 = -P_(SXX0, i + 1, j) * P_(ETAP, i + 1, j) /
                              (P_(GGGP, i + 1, j) * dt + P_(ETAP, i + 1, j));
                double dRdx = 0.5 * (P_(RHOY, i, j + 1) - P_(RHOY, i, j - 1)) / dx;
                double dRdy = 0.5 * (P_(RHOY, i + 1, j) - P_(RHOY, i - 1, j)) / dy;
                double gyij = P_(gy, i, j);
                double idx = 1.0 / dx, idy = 1.0 / dy;
                coo_add(&L, ETA1 / (dx * dx), kvy, kvy - 6 * Ny1);
                coo_add(&L, ETAP1 / (dy * dy), kvy, kvy - 6);
                coo_add(&L, -(ETAP1 + ETAP2) * (1.0 / (dy * dy)) -
                                (ETA1 + ETA2) * (1.0 / (dx * dx)) - dRdy * gyij * dt,
                        kvy, kvy);
                coo_add(&L, ETAP2 * (1.0 / (dy * dy)), kvy, kvy + 6);
                coo_add(&L, ETA2 * (1.0 / (dx * dx)), kvy, kvy + 6 * Ny1);
                coo_add(&L, ETAP1 * idx * idy - ETA2 * idx * idy - dRdx * gyij * dt * 0.25,
                        kvy, kvx);
                coo_add(&L, -ETAP2 * idx * idy + ETA2 * idx * idy - dRdx * gyij * dt * 0.25,
                        kvy, kvx + 6);
                coo_add(&L, -ETAP1 * idx * idy + ETA1 * idx * idy - dRdx * gyij * dt * 0.25,
                        kvy, kvx - 6 * Ny1);
                coo_add(&L, ETAP2 * idx * idy - ETA1 * idx * idy - dRdx * gyij * dt * 0.25,
                        kvy, kvx + 6 - 6 * Ny1);
                coo_add(&L, Kcont * idy, kvy, kpm);
                coo_add(&L, -Kcont * idy, kvy, kpm + 6);
                R[kvy - 1] = -P_(RHOY, i, j) * gyij - (SXY2 - SXY1) * idx - (SYY2 - SYY1) * idy;
            }
            /* P equation E:3398-3452 */
            int anchor = (i == 2 && 2 <= j && j <= Nx) || (j == 2 && 2 < i && i < Ny) ||
                         (i == Ny && 2 <= j && j <= Nx) || (j == Nx && 2 < i && i < Ny);
            if (i == 1 || i == Ny1 || j == 1 || j == Nx1) {
                coo_add(&L, 1.0, kpm, kpm);
            } else if (anchor) {
                coo_add(&L, Kcont, kpm, kpm);
                R[kpm - 1] = M->psurface;
            } else {
                double phi = P_(PHI, i, j), ep = P_(ETAPHI, i, j), bp = P_(BETAPHI, i, j);
                coo_add(&L, -1.0 / dx, kpm, kvx - 6 * Ny1);
                coo_add(&L, 1.0 / dx, kpm, kvx);
                coo_add(&L, -1.0 / dy, kpm, kvy - 6);
                coo_add(&L, 1.0 / dy, kpm, kvy);
                coo_add(&L, Kcont / (1.0 - phi) * ((1.0 / ep) + bp / dt), kpm, kpm);
                coo_add(&L, -Kcont / (1.0 - phi) * ((1.0 / ep) + bp / dt), kpm, kpf);
                R[kpm - 1] = (P_(pr0, i, j) - P_(pf0, i, j)) / (1.0 - phi) * bp / dt +
                             P_(DMP, i, j);
            }
            /* qx Darcy equation E:3454-3480 */
            if (i == 1 || i == Ny1 || j == 1 || j == Nx || j == Nx1) {
                coo_add(&L, 1.0, kqx, kqx);
                if (i == 1 && 1 < j && j < Nx) coo_add(&L, M->bcftop, kqx, kqx + 6);
                if (i == Ny1 && 1 < j && j < Nx) coo_add(&L, M->bcfbottom, kqx, kqx - 6);
            } else {
                coo_add(&L, P_(RX, i, j), kqx, kqx);
                coo_add(&L, -Kcont * (1.0 / dx), kqx, kpf);
                coo_add(&L, Kcont * (1.0 / dx), kqx, kpf + 6 * Ny1);
                R[kqx - 1] = P_(RHOFX, i, j) * P_(gx, i, j);
            }
            /* qy Darcy equation E:3482-3514 */
            if (i == 1 || i == Ny || i == Ny1 || j == 1 || j == Nx1) {
                coo_add(&L, 1.0, kqy, kqy);
                if (j == 1 && 1 < i && i < Ny) coo_add(&L, M->bcfleft, kqy, kqy + 6 * Ny1);
                if (j == Nx1 && 1 < i && i < Ny) coo_add(&L, M->bcfright, kqy, kqy - 6 * Ny1);
            } else {
                coo_add(&L, P_(RY, i, j), kqy, kqy);
                coo_add(&L, -Kcont * (1.0 / dy), kqy, kpf);
                coo_add(&L, Kcont * (1.0 / dy), kqy, kpf + 6);
                R[kqy - 1] = P_(RHOFY, i, j) * P_(gy, i, j);
            }
            /* P fluid equation E:3516-3565 */
            if (i == 1 || i == Ny1 || j == 1 || j == Nx1) {
                coo_add(&L, 1.0, kpf, kpf);
            } else if (anchor) {
                coo_add(&L, Kcont, kpf, kpf);
                R[kpf - 1] = M->psurface;
            } else {
                double phi = P_(PHI, i, j), ep = P_(ETAPHI, i, j), bp = P_(BETAPHI, i, j);
                coo_add(&L, -(1.0 / dx), kpf, kqx - 6 * Ny1);
                coo_add(&L, (1.0 / dx), kpf, kqx);
                coo_add(&L, -(1.0 / dy), kpf, kqy - 6);
                coo_add(&L, (1.0 / dy), kpf, kqy);
                coo_add(&L, -Kcont / (1.0 - phi) * (1.0 / ep + bp / dt), kpf, kpm);
                coo_add(&L, Kcont / (1.0 - phi) * (1.0 / ep + bp / dt), kpf, kpf);
                R[kpf - 1] = -(P_(pr0, i, j) - P_(pf0, i, j)) / (1.0 - phi) * bp / dt;
            }
        }
    }
#undef B_
#undef P_
    int64_t nnz = -1;
    if (!L.overflow) nnz = coo_to_csc(&L, colptr, rowval, nzval);
    free(L.row);
    free(L.col);
    free(L.val);
    return nnz;
}

/* assemble_thermal_lse! E:4526-4611 */
int64_t orc_assemble_thermal(const ereb_geometry* g, const double* tk1,
    const double* RHOCP, const double* KX, const double* KY, const double* HR,
    const double* HA, const double* HS, const double* DHP, double* RT,
    double dt, int64_t* colptr, int64_t* rowval, double* nzval, int64_t cap) {
    const int64_t Nx1 = g->Nx + 1, Ny1 = g->Ny + 1;
    const double dx = g->dx, dy = g->dy;
    const int64_t n = Nx1 * Ny1;
    coo_t L = {n, 0, cap, (int64_t*)malloc(sizeof(int64_t) * cap),
               (int64_t*)malloc(sizeof(int64_t) * cap),
               (double*)malloc(sizeof(double) * cap), 0};
    for (int64_t k = 0; k < n; ++k) RT[k] = 0.0;
    for (int64_t j = 1; j <= Nx1; ++j) {
        for (int64_t i = 1; i <= Ny1; ++i) {
            int64_t gk = (j - 1) * Ny1 + i;
            if (i == 1 || i == Ny1 || j == 1 || j == Nx1) {
                coo_add(&L, 1.0, gk, gk);
                if (j == 1) coo_add(&L, -1.0, gk, gk + Ny1);
                if (j == Nx1) coo_add(&L, -1.0, gk, gk - Ny1);
                if (i == 1 && 1 < j && j < Nx1) coo_add(&L, -1.0, gk, gk + 1);
                if (i == Ny1 && 1 < j && j < Nx1) coo_add(&L, -1.0, gk, gk - 1);
            } else {
                double Kx1 = G(KX, i, j - 1, Ny1), Kx2 = G(KX, i, j, Ny1);
                double Ky1 = G(KY, i - 1, j, Ny1), Ky2 = G(KY, i, j, Ny1);
                double rc = G(RHOCP, i, j, Ny1);
                coo_add(&L, -Kx1 * (1.0 / (dx * dx)), gk, gk - Ny1);
                coo_add(&L, -Ky1 * (1.0 / (dy * dy)), gk, gk - 1);
                coo_add(&L, rc / dt + (Kx1 + Kx2) * (1.0 / (dx * dx)) +
                                (Ky1 + Ky2) * (1.0 / (dy * dy)),
                        gk, gk);
                coo_add(&L, -Ky2 * (1.0 / (dy * dy)), gk, gk + 1);
                coo_add(&L, -Kx2 * (1.0 / (dx * dx)), gk, gk + Ny1);
                RT[gk - 1] = rc / dt * G(tk1, i, j, Ny1) + G(HR, i, j, Ny1) +
                             G(HA, i, j, Ny1) + G(HS, i, j, Ny1) + G(DHP, i, j, Ny1);
            }
        }
    }
    int64_t nnz = -1;
    if (!L.overflow) nnz = coo_to_csc(&L, colptr, rowval, nzval);
    free(L.row);
    free(L.col);
    free(L.val);
    return nnz;
}

/* ------------------------------------------------- sort specification */

/* ------------------------ viscosity, compaction and (de)hydration (next) */

/* Julia min(a, b) for Float64: NaN propagates */
static inline double jl_min(double a, double b) {
    if (a != a || b != b) return a + b;
    return a < b ? a : (b < a ? b : (signbit(a) ? a : b));
}

/* update_marker_viscosity! E:774-804 for m = 1..marknum (loop E:6950-6953) */
void orc_update_marker_viscosity(const ereb_geometry* g, const ereb_materials* M,
    const double* xm, const double* ym, const int64_t* tm, const double* tkm,
    double* etatotalm, double* etavpm, const uint8_t* YNY,
    const double* YNY_inv_ETA, int64_t marknum) {
    const int64_t Nx = g->Nx, Ny = g->Ny;
    /* @threads in the reference loop (E:6950) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t m = 0; m < marknum; ++m) {
        int64_t i, j;
        double w[4];
        orc_fix_weights(xm[m], ym[m], g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1,
                        Ny - 1, &i, &j, w);
        if (tm[m] < 3) etatotalm[m] = orc_etatotal_rocks(M, tkm[m], tm[m]);
        const int any = G(YNY, i, j, Ny) || G(YNY, i + 1, j, Ny) ||
                        G(YNY, i, j + 1, Ny) || G(YNY, i + 1, j + 1, Ny);
        if (any) {
            double v = gdot4(YNY_inv_ETA, Ny, i, j, w);
            v = 1.0 / v;
            etavpm[m] = v > etatotalm[m] ? etatotalm[m] : v;
        } else {
            etavpm[m] = etatotalm[m];
        }
    }
}

/* update_marker_porosity! E:4822-4856 */
void orc_update_marker_porosity(const ereb_geometry* g, const ereb_hydration* H,
    const double* xm, const double* ym, const int64_t* tm, double* phim,
    const double* APHI, double dt, int64_t marknum) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    for (int64_t m = 0; m < marknum; ++m) {
        if (!(tm[m] < 3)) continue;
        int64_t i, j;
        double w[4];
        orc_fix_weights(xm[m], ym[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny,
                        &i, &j, w);
        const double aphim = gdot4(APHI, Ny1, i, j, w);
        const double v = phim[m] / ((1.0 - phim[m]) * exp(aphim * dt) + phim[m]);
        phim[m] = jl_max(H->phimin, jl_min(H->phimax, v));
    }
}

/* update_p_nodes_melt_composition! E:1929-1939: molarfraction_marker_to_p_nodes!
 * (E:1883-1908) for every marker, then compute_molarfraction! E:2239-2252 */
void orc_update_p_nodes_melt_composition(const ereb_geometry* g, const double* xm,
    const double* ym, const double* XWsolidm0, double* XWS, double* XWSSUM,
    double* WTPSUM, int64_t marknum) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Nx1 = Nx + 1, Ny1 = Ny + 1;
    memset(XWSSUM, 0, sizeof(double) * Nx1 * Ny1);
    memset(WTPSUM, 0, sizeof(double) * Nx1 * Ny1);
    for (int64_t m = 0; m < marknum; ++m) {
        int64_t i, j;
        double w[4];
        orc_fix_weights(xm[m], ym[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny,
                        &i, &j, w);
        add4(XWSSUM, Ny1, i, j, w, XWsolidm0[m]);
        add4(WTPSUM, Ny1, i, j, w, 1.0);
    }
    for (int64_t k = 0; k < Nx1 * Ny1; ++k)
        XWS[k] = WTPSUM[k] > 0.0 ? XWSSUM[k] * (1.0 / WTPSUM[k]) : 0.0;
}

/* compute_Δtreaction E:2382-2397; *err = EREB_ERR_MODE on unknown mode.
 * Mode 3 divides by Sxo_B, as the reference does. */
double orc_dtreaction(const ereb_hydration* H, double T, double phi, int64_t mode,
                      int* err) {
    if (mode == 1) {
        const double d = T - H->c_I;
        return -H->log_completion_rate / (H->A_I * phi) * exp(H->b_I * (d * d));
    } else if (mode == 2) {
        return -H->log_completion_rate / (H->Sxo_B * phi) *
               pow(2.0, (H->To_B - T) / H->Tscl_B);
    } else if (mode == 3) {
        return -H->log_completion_rate / (H->Sxo_B * phi) *
               exp(H->Ea_T / H->RG * (1.0 / T - 1.0 / H->To_T));
    } else if (mode == 9) {
        return H->dtreaction;
    }
    if (err) *err = EREB_ERR_MODE;
    return NAN;
}

/* compute_gibbs_free_energy E:2470-2482 */
double orc_gibbs_free_energy(const ereb_hydration* H, double T, double pf, double XDs,
                             double XWs, double dt, double dtr) {
    if (dt < dtr)
        return (H->dHWD - T * H->dSWD + pf * H->dVWD + H->RG * T * log(XDs / XWs)) *
               (1.0 - dt / dtr);
    return 0.0;
}

/* compute_reaction_constant E:2521-2524 */
double orc_reaction_constant(const ereb_hydration* H, double T, double pf, double dG) {
    return exp(-(H->dHWD - T * H->dSWD + H->dVWD * pf - dG) / (H->RG * T));
}

/* compute_relative_enthalpy E:2502-2504 */
double orc_relative_enthalpy(const ereb_materials* M, const ereb_hydration* H,
                             double Xs, double XWs) {
    return -Xs * XWs * H->dHWD / (M->MD + M->MH2O);
}

/* perform_thermochemical_reaction! E:2558-2694 (without the @info lines):
 * the reaction of every rock marker, the ΔM, ΔH and weight sums onto P
 * nodes for the fluid-bearing ones, then compute_thermodynamic_xfer!
 * E:2208-2222.  Returns EREB_ERR_MODE for an unknown rate mode. */
int orc_perform_thermochemical_reaction(const ereb_geometry* g,
    const ereb_materials* M, const ereb_hydration* H, double* DMP, double* DHP,
    double* DMPSUM, double* DHPSUM, double* WTPSUM, const double* pf,
    const double* tk2, const int64_t* tm, const double* xm, const double* ym,
    double* XWsm0, double* XWsm, double* phim, double* phinewm, double* pfm0,
    int64_t marknum, double dt, int64_t timestep, int64_t titer) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Nx1 = Nx + 1, Ny1 = Ny + 1;
    int err = 0;
    memset(DMPSUM, 0, sizeof(double) * Nx1 * Ny1);
    memset(DHPSUM, 0, sizeof(double) * Nx1 * Ny1);
    memset(WTPSUM, 0, sizeof(double) * Nx1 * Ny1);
    for (int64_t m = 0; m < marknum; ++m) {
        if (!(tm[m] < 3)) continue;
        int64_t i, j;
        double w[4];
        orc_fix_weights(xm[m], ym[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny,
                        &i, &j, w);
        const double tknm = gdot4(tk2, Ny1, i, j, w);
        double pfnm = jl_max(0.0, gdot4(pf, Ny1, i, j, w));
        if (titer > 2) pfnm = pfnm * (1.0 - H->pfcoeff) + pfm0[m] * H->pfcoeff;
        pfm0[m] = pfnm;
        const double XDs0 = 1.0 - XWsm0[m];
        const double VH2O = tknm > M->tmfluidphase ? H->VH2Of : H->VH2Ofi;
        const double Xf0 = phim[m] * (XWsm0[m] * M->VWs + XDs0 * M->VDs) /
                           ((1.0 - phim[m]) * VH2O +
                            phim[m] * (XWsm0[m] * M->VWs + XDs0 * M->VDs));
        const double Xs0 = 1.0 - Xf0;
        const double XH2Ot = (XWsm0[m] * Xs0 + Xf0) / (1.0 + XWsm0[m] * Xs0);
        const double XDt = 1.0 - XH2Ot;
        const double rhos0 = (M->MD + M->MH2O * XWsm0[m]) / (M->VDs * XDs0 + M->VWs * XWsm0[m]);
        const double rhof0 = tknm > M->tmfluidphase ? M->rhoH2Of : M->rhoH2Ofi;
        const double dtr = orc_dtreaction(H, tknm, phim[m], H->rate_mode, &err);
        if (err) return EREB_ERR_MODE;
        const double Ht0 = orc_relative_enthalpy(M, H, Xs0, XWsm0[m]);
        const double dG0 = orc_gibbs_free_energy(H, tknm, pfnm, XDs0, XWsm0[m], dt, dtr);
        const double KWD = orc_reaction_constant(H, tknm, pfnm, dG0);
        const double XWs1 = 1.0 / (KWD + 1.0);
        const double XDs1 = 1.0 - XWs1;
        const double Xs1 = XDt / (1.0 - XDt * XWs1);
        const double Xf1 = 1.0 - Xs1;
        if (0.0 < Xf1 && Xf1 < 1.0) {
            const double phi1 = Xf1 * VH2O / (Xf1 * VH2O + Xs1 * (XWs1 * M->VWs + XDs1 * M->VDs));
            const double rhos1 = (M->MD + M->MH2O * XWs1) / (M->VDs * XDs1 + M->VWs * XWs1);
            const double rhof1 = tknm > M->tmfluidphase ? M->rhoH2Of : M->rhoH2Ofi;
            const double Ht1 = orc_relative_enthalpy(M, H, Xs1, XWs1);
            const double dHt = Ht1 - Ht0;
            const double RV = (rhos1 * (1.0 - phi1) + rhof1 * phi1) /
                              (rhos0 * (1.0 - phim[m]) + rhof0 * phim[m]);
            const double Gmass = (rhos0 * RV * (1.0 - phim[m]) - rhos1 * (1.0 - phi1)) / dt;
            const double dMm = (1.0 - RV) / dt;
            const double dHm = Gmass * dHt;
            XWsm[m] = XWs1;
            phinewm[m] = phi1;
            if (timestep == 1) {
                XWsm0[m] = XWsm[m];
                phim[m] = phinewm[m];
            }
            add4(DMPSUM, Ny1, i, j, w, dMm);
            add4(DHPSUM, Ny1, i, j, w, dHm);
            add4(WTPSUM, Ny1, i, j, w, 1.0);
        }
    }
    for (int64_t k = 0; k < Nx1 * Ny1; ++k) {
        if (WTPSUM[k] > 0.0) {
            DMP[k] = DMPSUM[k] * (1.0 / WTPSUM[k]);
            DHP[k] = DHPSUM[k] * (1.0 / WTPSUM[k]);
        } else {
            DMP[k] = DHP[k] = 0.0;
        }
    }
    return EREB_OK;
}

/* ------------------------------------------------ pressure back-trace (next) */

/* One stage velocity of backtrace_pressures_rk4! on the Vx grid (E:5299-5332):
 * the stencil of E:5057-5091 without the E:5173 centre quirk, combined as
 * (1-dymi/dy)*v13 + (dymi/dy)*v24 */
static double bt_vx(const ereb_geometry* g, const double* v, double x, double y) {
    const int64_t Nx = g->Nx, Ny = g->Ny, R = Ny + 1;
    int64_t i, j;
    double dxmj, dymi;
    orc_fix_distances(x, y, g->xvx, g->yvx, g->dx, g->dy, 1, Nx - 1, 1, Ny, &i,
                      &j, &dxmj, &dymi);
    const double dx = g->dx, dy = g->dy;
    double v13 = G(v, i, j, R) * (1.0 - dxmj / dx) + G(v, i, j + 1, R) * dxmj / dx;
    double v24 = G(v, i + 1, j, R) * (1.0 - dxmj / dx) + G(v, i + 1, j + 1, R) * dxmj / dx;
    if (dxmj / dx >= 0.5) {
        if (j < Nx - 1) {
            const double t = dxmj / dx - 0.5;
            v13 += 0.5 * (t * t) * (G(v, i, j, R) - 2.0 * G(v, i, j + 1, R) + G(v, i, j + 2, R));
            v24 += 0.5 * (t * t) * (G(v, i + 1, j, R) - 2.0 * G(v, i + 1, j + 1, R) + G(v, i + 1, j + 2, R));
        }
    } else if (j > 1) {
        const double t = dxmj / dx - 0.5;
        v13 += 0.5 * (t * t) * (G(v, i, j - 1, R) - 2.0 * G(v, i, j, R) + G(v, i, j + 1, R));
        v24 += 0.5 * (t * t) * (G(v, i + 1, j - 1, R) - 2.0 * G(v, i + 1, j, R) + G(v, i + 1, j + 1, R));
    }
    return (1.0 - dymi / dy) * v13 + (dymi / dy) * v24;
}

/* ... and on the Vy grid (E:5333-5364); `1/2*(...)` of the left branch
 * (E:5359-5360) is 0.5*(...) */
static double bt_vy(const ereb_geometry* g, const double* v, double x, double y) {
    const int64_t Nx = g->Nx, Ny = g->Ny, R = Ny + 1;
    int64_t i, j;
    double dxmj, dymi;
    orc_fix_distances(x, y, g->xvy, g->yvy, g->dx, g->dy, 1, Nx, 1, Ny - 1, &i,
                      &j, &dxmj, &dymi);
    const double dx = g->dx, dy = g->dy;
    double v12 = G(v, i, j, R) * (1.0 - dymi / dy) + G(v, i + 1, j, R) * dymi / dy;
    double v34 = G(v, i, j + 1, R) * (1.0 - dymi / dy) + G(v, i + 1, j + 1, R) * dymi / dy;
    if (dymi / dy >= 0.5) {
        if (i < Ny - 1) {
            const double t = dymi / dy - 0.5;
            v12 += 0.5 * (t * t) * (G(v, i, j, R) - 2.0 * G(v, i + 1, j, R) + G(v, i + 2, j, R));
            v34 += 0.5 * (t * t) * (G(v, i, j + 1, R) - 2.0 * G(v, i + 1, j + 1, R) + G(v, i + 2, j + 1, R));
        }
    } else if (i > 1) {
        const double t = dymi / dy - 0.5;
        v12 += 0.5 * (t * t) * (G(v, i - 1, j, R) - 2.0 * G(v, i, j, R) + G(v, i + 1, j, R));
        v34 += 0.5 * (t * t) * (G(v, i - 1, j + 1, R) - 2.0 * G(v, i, j + 1, R) + G(v, i + 1, j + 1, R));
    }
    return (1.0 - dxmj / dx) * v12 + (dxmj / dx) * v34;
}

/* RK4 back-trace of P node (ii, jj) through (vx, vy) (E:5296-5376): the
 * departure point (*xo, *yo).  The final step is xA - dt*1//6*(...), which
 * Julia parses as xA - (dt*(1//6))*(...) (// binds tighter than *). */
static void bt_node(const ereb_geometry* g, const double* vx, const double* vy,
                    int64_t ii, int64_t jj, double dt, double* xo, double* yo) {
    const double xA = g->xp[jj - 1], yA = g->yp[ii - 1];
    double xc = xA, yc = yA, vxm[4], vym[4];
    for (int rk = 0; rk < 4; ++rk) {
        vxm[rk] = bt_vx(g, vx, xc, yc);
        vym[rk] = bt_vy(g, vy, xc, yc);
        if (rk == 0 || rk == 1) {
            xc = xA - 0.5 * dt * vxm[rk];
            yc = yA - 0.5 * dt * vym[rk];
        } else if (rk == 2) {
            xc = xA - dt * vxm[rk];
            yc = yA - dt * vym[rk];
        }
    }
    const double sixth = 1.0 / 6.0; /* Float64(1//6) */
    *xo = xA - dt * sixth * (vxm[0] + 2.0 * vxm[1] + 2.0 * vxm[2] + vxm[3]);
    *yo = yA - dt * sixth * (vym[0] + 2.0 * vym[1] + 2.0 * vym[2] + vym[3]);
}

/* backtrace_pressures_rk4! E:5284-5496: pr0 .= pr, ps0 .= ps, pf0 .= pf, then
 * the interior P nodes (ii = 2..Ny, jj = 2..Nx) take the pressures at their
 * departure points: solid velocities for pr and ps, fluid ones for pf */
void orc_backtrace_pressures_rk4(const ereb_geometry* g, const double* pr,
    double* pr0, const double* ps, double* ps0, const double* pf, double* pf0,
    const double* vx, const double* vy, const double* vxf, const double* vyf,
    double dt) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Nx1 = Nx + 1, Ny1 = Ny + 1;
    memcpy(pr0, pr, sizeof(double) * Nx1 * Ny1);
    memcpy(ps0, ps, sizeof(double) * Nx1 * Ny1);
    for (int64_t jj = 2; jj <= Nx; ++jj)
        for (int64_t ii = 2; ii <= Ny; ++ii) {
            double x, y, w[4];
            int64_t i, j;
            bt_node(g, vx, vy, ii, jj, dt, &x, &y);
            orc_fix_weights(x, y, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
            G(pr0, ii, jj, Ny1) = gdot4(pr, Ny1, i, j, w);
            G(ps0, ii, jj, Ny1) = gdot4(ps, Ny1, i, j, w);
        }
    memcpy(pf0, pf, sizeof(double) * Nx1 * Ny1);
    for (int64_t jj = 2; jj <= Nx; ++jj)
        for (int64_t ii = 2; ii <= Ny; ++ii) {
            double x, y, w[4];
            int64_t i, j;
            bt_node(g, vxf, vyf, ii, jj, dt, &x, &y);
            orc_fix_weights(x, y, g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &i, &j, w);
            G(pf0, ii, jj, Ny1) = gdot4(pf, Ny1, i, j, w);
        }
}

/* -------------------------------------------- marker replenishment (next) */

/* distance E:1097-1099 */
static inline double orc_distance(double x1, double y1, double x2, double y2) {
    const double a = x1 - x2, b = y1 - y2;
    return sqrt(a * a + b * b);
}

/* replenish_markers! E:5586-5698 with randomized = false (the reference's
 * test mode): fields[f] is the marker vector of ereb_marker_field f
 * (int64_t* for EREB_M_TM, double* otherwise) with room for `cap` markers;
 * mdis (Nym*Nxm) and mnum (Nym*Nxm) are overwritten.  Returns the new
 * marker count, or -1 when it would exceed cap (vectors then partly
 * appended). */
int64_t orc_replenish_markers(const ereb_marker_grid* mg, void* const* fields,
                              int64_t marknum, int64_t cap, double* mdis,
                              int64_t* mnum) {
    const int64_t Nxm = mg->Nxm, Nym = mg->Nym;
    double* xm = (double*)fields[EREB_M_XM];
    double* ym = (double*)fields[EREB_M_YM];
    for (int64_t k = 0; k < Nxm * Nym; ++k) {
        mdis[k] = mg->mdis_init;
        mnum[k] = 0;
    }
    /* update_marker_population_geometry! E:5518-5544 for every marker */
    for (int64_t m = 0; m < marknum; ++m) {
        int64_t i, j;
        orc_fix(xm[m], ym[m], mg->xxm, mg->yym, mg->dxm, mg->dym, 1, Nxm - 1, 1,
                Nym - 1, &i, &j);
        const int64_t ci[4] = {i, i + 1, i, i + 1}, cj[4] = {j, j, j + 1, j + 1};
        for (int q = 0; q < 4; ++q) {
            const double d = orc_distance(xm[m], ym[m], mg->xxm[cj[q] - 1], mg->yym[ci[q] - 1]);
            if (d < G(mdis, ci[q], cj[q], Nym)) {
                G(mdis, ci[q], cj[q], Nym) = d;
                G(mnum, ci[q], cj[q], Nym) = m + 1;
            }
        }
    }
    int64_t n = marknum;
    for (int64_t j = 1; j <= Nxm; ++j)
        for (int64_t i = 1; i <= Nym; ++i) {
            if (G(mnum, i, j, Nym) != 0) continue;
            const int64_t jlo = j - mg->djj > 1 ? j - mg->djj : 1;
            const int64_t jhi = j + mg->djj < Nxm ? j + mg->djj : Nxm;
            const int64_t ilo = i - mg->dii > 1 ? i - mg->dii : 1;
            const int64_t ihi = i + mg->dii < Nym ? i + mg->dii : Nym;
            for (int64_t jj = jlo; jj <= jhi; ++jj)
                for (int64_t ii = ilo; ii <= ihi; ++ii)
                    if (G(mnum, ii, jj, Nym) > 0) {
                        const int64_t m = G(mnum, ii, jj, Nym);
                        const double d = orc_distance(xm[m - 1], ym[m - 1], mg->xxm[j - 1],
                                                      mg->yym[i - 1]);
                        if (d < G(mdis, i, j, Nym)) {
                            G(mdis, i, j, Nym) = d;
                            G(mnum, i, j, Nym) = -m;
                        }
                    }
            if (G(mnum, i, j, Nym) < 0) {
                if (n >= cap) return -1;
                const int64_t src = -G(mnum, i, j, Nym) - 1;
                for (int f = 0; f < EREB_M_NFIELDS; ++f) {
                    if (f == EREB_M_TM)
                        ((int64_t*)fields[f])[n] = ((int64_t*)fields[f])[src];
                    else
                        ((double*)fields[f])[n] = ((double*)fields[f])[src];
                }
                xm[n] = mg->xxm[j - 1];
                ym[n] = mg->yym[i - 1];
                ++n;
            }
        }
    return n;
}

/* Fine key of the library's sort: the basic cell (ib, jb) from fix with the
 * basic clamps (E:1456, C:68-74) refined by the P-cell offsets
 * (ip-ib, jp-jb) in {0,1}^2 (P clamps, C:94-101), numbered tile-major:
 * tiles of `tile` x `tile` basic cells in column-major order, then the
 * 2*tile x 2*tile fine cells of a tile in column-major order.  Offsets
 * outside {0,1} (coordinates beyond ~1e15 m) are clamped into it. */
void orc_fine_keys(const ereb_geometry* g, int64_t tile, const double* xm,
                   const double* ym, int64_t marknum, int32_t* key) {
    const int64_t Nx = g->Nx, Ny = g->Ny;
    const int64_t nty = (Ny - 1 + tile - 1) / tile;
    const int64_t fine = 2 * tile;
    for (int64_t m = 0; m < marknum; ++m) {
        int64_t ib, jb, ip, jp;
        orc_fix(xm[m], ym[m], g->x, g->y, g->dx, g->dy, 1, Nx - 1, 1, Ny - 1, &ib, &jb);
        orc_fix(xm[m], ym[m], g->xp, g->yp, g->dx, g->dy, 1, Nx, 1, Ny, &ip, &jp);
        int64_t di = ip - ib, dj = jp - jb;
        di = di < 0 ? 0 : (di > 1 ? 1 : di);
        dj = dj < 0 ? 0 : (dj > 1 ? 1 : dj);
        const int64_t ty = (ib - 1) / tile, tx = (jb - 1) / tile;
        const int64_t fi = 2 * (ib - 1 - ty * tile) + di;
        const int64_t fj = 2 * (jb - 1 - tx * tile) + dj;
        key[m] = (int32_t)((tx * nty + ty) * fine * fine + fj * fine + fi);
    }
}

/* Stable counting sort of 0..marknum-1 by key: perm[k] = marker index,
 * start[q] = first position of key q (start[nkeys] = marknum). */
void orc_key_sort(const int32_t* key, int64_t marknum, int64_t nkeys,
                  int32_t* perm, int64_t* start) {
    int64_t* cnt = (int64_t*)calloc((size_t)nkeys + 1, sizeof(int64_t));
    for (int64_t m = 0; m < marknum; ++m) cnt[key[m] + 1]++;
    for (int64_t t = 0; t < nkeys; ++t) cnt[t + 1] += cnt[t];
    for (int64_t t = 0; t <= nkeys; ++t) start[t] = cnt[t];
    for (int64_t m = 0; m < marknum; ++m) perm[cnt[key[m]]++] = (int32_t)m;
    free(cnt);
}

/* ----------------------------------------------- derived grid helpers */

/* compute_rotation_rate! E:4946-4956 */
void orc_compute_rotation_rate(const ereb_geometry* g, const double* vx,
                               const double* vy, double* wyx) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Ny1 = Ny + 1;
    for (int64_t j = 1; j <= Nx; ++j)
        for (int64_t i = 1; i <= Ny; ++i)
            G(wyx, i, j, Ny) = 0.5 * ((G(vy, i, j + 1, Ny1) - G(vy, i, j, Ny1)) * (1.0 / g->dx) -
                                      (G(vx, i + 1, j, Ny1) - G(vx, i, j, Ny1)) * (1.0 / g->dy));
}

int orc_version(void) { return 1; }
