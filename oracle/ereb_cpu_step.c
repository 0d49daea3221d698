/*
 * ereb_cpu_step.c -- CPU baseline of the benchmark: whole timesteps of the
 * transfer path run by the C restatement (ereb_oracle.c) with no Python in
 * the timed loop.  TEST INFRASTRUCTURE ONLY (linked into
 * _build/libereb_oracle_bench.so, loaded by bench.py's CPU-baseline leg).
 *
 * Step order, as simulation_loop (E:6376-7053) without the host solves:
 *   compute_marker_properties! + marker_to_*_nodes!  (loop E:6419-6517)
 *   compute_*_node_properties!, insulating tk1        (E:6522-6604)
 *   assemble_hydromechanical_lse!                     (E:6680)
 *   assemble_thermal_lse!                             (E:6922)
 *   update_marker_stress!                             (E:6980, @threads)
 *   update_marker_temperature!                        (E:7004, @threads)
 *   move_markers_rk4!                                 (E:7036)
 * Compiled with -O3 -fopenmp: only the two gathers carry OpenMP pragmas,
 * where the reference has @threads (E:4356, E:4786, E:4793); the scatter,
 * node properties, assemblies and RK4 run on one thread, as in the reference.
 */
#define _POSIX_C_SOURCE 199309L  /* clock_gettime */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ereb.h"

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
    int64_t marknum);
void orc_compute_basic_node_properties(const ereb_geometry* g,
    const double* ETA0SUM, const double* ETASUM, const double* GGGSUM,
    const double* SXYSUM, const double* COHSUM, const double* TENSUM,
    const double* FRISUM, const double* WTSUM, double* ETA0, double* ETA,
    double* GGG, double* SXY0, double* COH, double* TEN, double* FRI,
    uint8_t* YNY);
void orc_compute_v_node_properties(const ereb_geometry* g, const double* RHOSUM,
    const double* RHOFSUM, const double* KSUM, const double* PHISUM,
    const double* RSUM, const double* WTSUM, double* RHO, double* RHOF,
    double* K, double* PHI, double* R);
void orc_compute_p_node_properties(const ereb_geometry* g, const double* RHOSUM,
    const double* RHOCPSUM, const double* ALPHASUM, const double* ALPHAFSUM,
    const double* HRSUM, const double* GGGPSUM, const double* SXXSUM,
    const double* TKSUM, const double* PHISUM, const double* WTPSUM,
    double* RHO, double* RHOCP, double* ALPHA, double* ALPHAF, double* HR,
    double* GGGP, double* SXX0, double* tk1, double* PHI, double* BETAPHI);
void orc_apply_insulating_bc(double* t, int64_t Nyy, int64_t Nxx);
int64_t orc_assemble_hydromechanical(const ereb_geometry* g,
    const ereb_materials* M, const double* ETA, const double* ETAP,
    const double* GGG, const double* GGGP, const double* SXY0,
    const double* SXX0, const double* RHOX, const double* RHOY,
    const double* RHOFX, const double* RHOFY, const double* RX,
    const double* RY, const double* ETAPHI, const double* BETAPHI,
    const double* PHI, const double* gx, const double* gy, const double* pr0,
    const double* pf0, const double* DMP, double dt, double* R,
    int64_t* colptr, int64_t* rowval, double* nzval, int64_t cap);
int64_t orc_assemble_thermal(const ereb_geometry* g, const double* tk1,
    const double* RHOCP, const double* KX, const double* KY, const double* HR,
    const double* HA, const double* HS, const double* DHP, double* RT,
    double dt, int64_t* colptr, int64_t* rowval, double* nzval, int64_t cap);
void orc_update_marker_temperature(const ereb_geometry* g, const double* xm,
    const double* ym, double* tkm, const double* DT, const double* tk2,
    int64_t timestep, int64_t marknum);
void orc_update_marker_stress(const ereb_geometry* g, const double* xm,
    const double* ym, double* sxxm, double* sxym, const double* DSXX,
    const double* DSXY, int64_t marknum);
int orc_move_markers_rk4(const ereb_geometry* g, const ereb_materials* M,
    double* xm, double* ym, const int64_t* tm, double* tkm, const double* phim,
    double* sxym, double* sxxm, const double* vx, const double* vy,
    const double* vxf, const double* vyf, const double* wyx, const double* tk2,
    int64_t marknum, double dt, int64_t mode);

/* threads OpenMP will use in the threaded loops */
int orc_bench_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* Marker vectors, in this order (marknum each, updated in place):
 *   0 xm, 1 ym, 2 tkm, 3 phim, 4 sxxm, 5 sxym, 6 XWsolidm0, 7 etavpm,
 *   8 inv_gggtotalm, 9 fricttotalm, 10 cohestotalm, 11 tenstotalm,
 *   12 rhofluidcur, 13 alphasolidcur, 14 alphafluidcur, 15 rhototalm,
 *   16 rhocptotalm, 17 etatotalm, 18 hrtotalm, 19 ktotalm,
 *   20 tkm_rhocptotalm, 21 etafluidcur_inv_kphim
 * Fixed grids, in this order ((Ny1, Nx1) unless noted):
 *   0 vx, 1 vy, 2 wyx (Ny, Nx), 3 tk2, 4 DT, 5 DSXX, 6 DSXY (Ny, Nx),
 *   7 ETAP, 8 ETAPHI, 9 gx, 10 gy, 11 zeros (pr0, pf0, DMP, HA, HS, DHP)
 * Runs timesteps until `max_steps` are done or `budget_s` seconds have
 * passed; returns the number done (or a negative status) and the elapsed
 * seconds in *seconds. */
int64_t orc_bench_transfer(const ereb_geometry* g, const ereb_materials* M,
    double* const* mk, const int64_t* tm, int64_t marknum,
    const double* const* gr, const double* hrs, const double* hrf,
    int64_t mode, double dt, int64_t max_steps, double budget_s,
    double* seconds) {
    const int64_t Nx = g->Nx, Ny = g->Ny, Nb = Nx * Ny;
    const int64_t Ns = (Nx + 1) * (Ny + 1);
    const int64_t nsd = 6 * Ns, capsd = 40 * Ns, capt = 5 * Ns;
    /* node work arrays: 30 sums, 7 basic + 5 Vx + 5 Vy + 10 P outputs */
    double* pool = (double*)calloc((size_t)(8 * Nb + 22 * Ns + 7 * Nb + 20 * Ns) +
                                   (size_t)(nsd + capsd + Ns + capt), sizeof(double));
    int64_t* ipool = (int64_t*)malloc(sizeof(int64_t) *
                                      (size_t)(nsd + 1 + capsd + Ns + 1 + capt));
    uint8_t* YNY = (uint8_t*)calloc((size_t)Nb, 1);
    if (!pool || !ipool || !YNY) {
        free(pool);
        free(ipool);
        free(YNY);
        return -EREB_ERR_NOMEM;
    }
    double* p = pool;
    double* sums[30];
    for (int k = 0; k < 30; ++k) {
        sums[k] = p;
        p += k < 8 ? Nb : Ns;
    }
    double* ob[7];
    for (int k = 0; k < 7; ++k, p += Nb) ob[k] = p;
    double* ovx[5];
    double* ovy[5];
    for (int k = 0; k < 5; ++k, p += Ns) ovx[k] = p;
    for (int k = 0; k < 5; ++k, p += Ns) ovy[k] = p;
    double* op[10];
    for (int k = 0; k < 10; ++k, p += Ns) op[k] = p;
    double* R = p;
    p += nsd;
    double* nzsd = p;
    p += capsd;
    double* RT = p;
    p += Ns;
    double* nzt = p;
    int64_t* cpsd = ipool;
    int64_t* rvsd = cpsd + nsd + 1;
    int64_t* cpt = rvsd + capsd;
    int64_t* rvt = cpt + Ns + 1;
    const double* z = gr[11];
    const double t0 = now_s();
    int64_t done = 0;
    int rc = EREB_OK;
    while (done < max_steps) {
        rc = orc_marker_to_nodes(g, M, mk[0], mk[1], tm, mk[2], mk[3], mk[6], mk[7],
                                 mk[8], mk[9], mk[10], mk[11], mk[12], mk[13], mk[14],
                                 mk[4], mk[5], mk[15], mk[16], mk[17], mk[18], mk[19],
                                 mk[20], mk[21], hrs, hrf, mode, sums, marknum);
        if (rc) break;
        orc_compute_basic_node_properties(g, sums[0], sums[1], sums[2], sums[3], sums[4],
                                          sums[5], sums[6], sums[7], ob[0], ob[1], ob[2],
                                          ob[3], ob[4], ob[5], ob[6], YNY);
        orc_compute_v_node_properties(g, sums[8], sums[9], sums[10], sums[11], sums[12],
                                      sums[13], ovx[0], ovx[1], ovx[2], ovx[3], ovx[4]);
        orc_compute_v_node_properties(g, sums[14], sums[15], sums[16], sums[17], sums[18],
                                      sums[19], ovy[0], ovy[1], ovy[2], ovy[3], ovy[4]);
        orc_compute_p_node_properties(g, sums[20], sums[21], sums[22], sums[23], sums[24],
                                      sums[25], sums[26], sums[27], sums[28], sums[29],
                                      op[0], op[1], op[2], op[3], op[4], op[5], op[6],
                                      op[7], op[8], op[9]);
        orc_apply_insulating_bc(op[7], Ny + 1, Nx + 1);
        /* ETA GGG SXY0 basic; ETAP GGGP SXX0 RHOX RHOY RHOFX RHOFY RX RY
         * ETAPHI BETAPHI PHI gx gy pr0 pf0 DMP */
        if (orc_assemble_hydromechanical(g, M, ob[1], gr[7], ob[2], op[5], ob[3], op[6],
                                         ovx[0], ovy[0], ovx[1], ovy[1], ovx[4], ovy[4],
                                         gr[8], op[9], op[8], gr[9], gr[10], z, z, z, dt, R,
                                         cpsd, rvsd, nzsd, capsd) < 0 ||
            orc_assemble_thermal(g, op[7], op[1], ovx[2], ovy[2], op[4], z, z, z, RT, dt,
                                 cpt, rvt, nzt, capt) < 0) {
            rc = EREB_ERR_NOMEM;
            break;
        }
        orc_update_marker_stress(g, mk[0], mk[1], mk[4], mk[5], gr[5], gr[6], marknum);
        orc_update_marker_temperature(g, mk[0], mk[1], mk[2], gr[4], gr[3], 2, marknum);
        rc = orc_move_markers_rk4(g, M, mk[0], mk[1], tm, mk[2], mk[3], mk[5], mk[4],
                                  gr[0], gr[1], gr[0], gr[1], gr[2], gr[3], marknum, dt,
                                  mode);
        if (rc) break;
        ++done;
        if (now_s() - t0 > budget_s) break;
    }
    *seconds = now_s() - t0;
    free(pool);
    free(ipool);
    free(YNY);
    return rc ? -(int64_t)rc : done;
}
