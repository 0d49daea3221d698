/*
 * ereb.h -- C ABI of libereb_hip.so, the MI355X (gfx950) marker-in-cell
 * transfer library for Erebus.jl.
 *
 * Citations are file:line into the Erebus.jl sources
 * (E: = src/Erebus.jl, C: = src/constants.jl).
 *
 * Conventions (all entry points):
 *   - every function returns an int status, 0 = EREB_OK; on error
 *     ereb_last_error(ctx) returns a message.  No C++ exception crosses the ABI.
 *   - grids are column-major Float64 arrays exactly as Julia stores them:
 *     basic grids are (Ny, Nx), Vx/Vy/P grids are (Ny1, Nx1), element
 *     (i, j) (1-based) at offset (j-1)*rows + (i-1).
 *   - marker arrays are Julia Vector{Float64}; the marker type `tm` is
 *     Vector{Int64} (values 1 mantle, 2 crust, 3 sticky air).
 *   - host pointers are borrowed for the duration of one call only; the
 *     library owns every device allocation.
 *   - calls on one context are serialised by the caller (one HIP stream per
 *     context); different contexts are independent.
 *
 * Two call forms are offered for every step:
 *   1. drop-in entry points with the reference function's argument list
 *      (host arrays in, host arrays out), e.g. ereb_update_marker_temperature;
 *   2. resident entry points (ereb_step_*) that operate on the marker and
 *      grid state kept in HBM by the context, used by a GPU-resident time
 *      loop and by the benchmark.  Host <-> device traffic happens only in
 *      ereb_markers_set/get and ereb_grid_set/get.
 */
#ifndef EREB_H
#define EREB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define EREB_OK              0
#define EREB_ERR_ARG         1  /* bad argument (null pointer, size mismatch) */
#define EREB_ERR_MODE        2  /* unknown property mode: the reference throws
                                   "unknown mode" (E:2289, E:2318, E:2354) */
#define EREB_ERR_HIP         3  /* HIP runtime error */
#define EREB_ERR_DOMAIN      4  /* marker coordinate outside the range the
                                   tiled scatter supports (|x|,|y| > 1e15 m
                                   but finite) */
#define EREB_ERR_NOMEM       5
#define EREB_ERR_STATE       6  /* resident step called before its inputs */

/* ---- precision forms --------------------------------------------------- */
#define EREB_FP64   0  /* every operation in Float64, as the reference */
#define EREB_MIXED  1  /* properties, weights and interpolation in FP32 with
                          compensated accumulation; coordinates, RK4 updates,
                          node normalisation and matrix coefficients FP64 */

/* ---- geometry: the compile-time constants of C:14-101 ------------------- */
/* Axes are passed explicitly so that node coordinates equal the reference's
 * SVector constants (C:40-64) bit for bit.  Lengths: x Nx, y Ny, xvx Nx1,
 * yvx Ny1, xvy Nx1, yvy Ny1, xp Nx1, yp Ny1 (Nx1 = Nx+1, Ny1 = Ny+1).
 * The clamp ranges (C:66-101) are fixed functions of Nx, Ny:
 *   basic j 1..Nx-1, i 1..Ny-1;  Vx j 1..Nx-1, i 1..Ny;
 *   Vy j 1..Nx, i 1..Ny-1;       P  j 1..Nx,   i 1..Ny. */
typedef struct {
    int64_t Nx, Ny;
    double dx, dy;
    const double *x, *y;
    const double *xvx, *yvx;
    const double *xvy, *yvy;
    const double *xp, *yp;
} ereb_geometry;

/* ---- material tables and scalars: C:143-186, C:218-258, C:283-314 ------- */
typedef struct {
    /* per marker type 1..3 (index 0..2), C:146-184 */
    double rhosolidm[3], rhofluidm[3], etasolidm[3], etasolidmm[3];
    double etafluidm[3], etafluidmm[3], rhocpsolidm[3], rhocpfluidm[3];
    double alphasolidm[3], alphafluidm[3], ksolidm[3], kfluidm[3];
    double gggsolidm[3], frictsolidm[3], cohessolidm[3], tenssolidm[3];
    double kphim0[3], tkm0[3];
    /* scalars */
    double phim0;          /* C:225 */
    double tmsolidphase;   /* C:218 */
    double tmfluidphase;   /* C:220 */
    double Lf;             /* C:222 heat of fusion of ice */
    double MH2O, MD;       /* C:234, C:236 */
    double rhoH2Of, rhoH2Ofi; /* C:242, C:244 */
    double VDs, VWs;       /* C:246, C:248 (computed there as MD/rhoDs etc.) */
    double etamin;         /* C:343 */
    double Kcont;          /* C:143 */
    double psurface;       /* C:12 */
    /* boundary conditions, C:283-309 (free slip -1, no slip 1) */
    double bcleft, bcright, bctop, bcbottom;
    double bcfleft, bcfright, bcftop, bcfbottom;
    double vxleft, vxright, vytop, vybottom;
    /* RK4 Butcher coefficients C:312-314 (brk4 as Float64 of 1//6, 1//3) */
    double brk4[4], crk4[3];
} ereb_materials;

/* ---- context ------------------------------------------------------------ */
typedef struct ereb_ctx ereb_ctx;

int ereb_ctx_create(const ereb_geometry* geom, const ereb_materials* mat,
                    int precision, int device, ereb_ctx** out);
int ereb_ctx_destroy(ereb_ctx* ctx);
const char* ereb_last_error(const ereb_ctx* ctx);
int ereb_ctx_set_precision(ereb_ctx* ctx, int precision);
/* make room for `marknum` resident markers (grows, never shrinks) */
int ereb_ctx_reserve(ereb_ctx* ctx, int64_t marknum);
int ereb_sync(ereb_ctx* ctx);
int ereb_version(void);
/* Host-side set-up check, no device needed: discovers the stencil slot
 * tables of the Stokes-Darcy and thermal systems by running the equation
 * code (E:3128-3566, E:4544-4605) over a small grid on the host, as the
 * first assembly does, and returns the slots per P cell (44 and 5).
 * EREB_ERR_STATE when an equation references a column outside the 3x3-cell
 * neighbourhood or a table overflows. */
int ereb_host_selfcheck(int64_t* slots_stokes_darcy, int64_t* slots_thermal);

/* ---- resident marker state --------------------------------------------- */
/* Names follow the reference's marker vectors (E:55-71, E:593-668). */
enum ereb_marker_field {
    EREB_M_XM = 0, EREB_M_YM, EREB_M_TM /* int64 on the host */, EREB_M_TKM,
    EREB_M_PHIM, EREB_M_SXXM, EREB_M_SXYM, EREB_M_XWSOLIDM0, EREB_M_ETAVPM,
    EREB_M_INV_GGGTOTALM, EREB_M_FRICTTOTALM, EREB_M_COHESTOTALM,
    EREB_M_TENSTOTALM, EREB_M_RHOFLUIDCUR, EREB_M_ALPHASOLIDCUR,
    EREB_M_ALPHAFLUIDCUR, EREB_M_RHOTOTALM, EREB_M_RHOCPTOTALM,
    EREB_M_ETATOTALM, EREB_M_HRTOTALM, EREB_M_KTOTALM,
    EREB_M_TKM_RHOCPTOTALM, EREB_M_ETAFLUIDCUR_INV_KPHIM,
    /* hydration state (E:2558-2694, E:7011-7014): current wet-solid molar
     * fraction XWsolidm, next porosity generation phinewm and the previous
     * iteration's marker fluid pressure pfm0 */
    EREB_M_XWSOLIDM, EREB_M_PHINEWM, EREB_M_PFM0,
    EREB_M_NFIELDS
};
/* marknum becomes the context's active marker count on set */
int ereb_markers_set(ereb_ctx* ctx, int field, const void* host, int64_t marknum);
int ereb_markers_get(ereb_ctx* ctx, int field, void* host, int64_t marknum);
int64_t ereb_marknum(const ereb_ctx* ctx);

/* ---- resident grid state ------------------------------------------------ */
/* Names follow the reference (E:851-923 SUM grids, E:147-343 node grids). */
enum ereb_grid_field {
    /* basic nodes (Ny, Nx) */
    EREB_G_ETA0SUM = 0, EREB_G_ETASUM, EREB_G_GGGSUM, EREB_G_SXYSUM,
    EREB_G_COHSUM, EREB_G_TENSUM, EREB_G_FRISUM, EREB_G_WTSUM,
    EREB_G_ETA0, EREB_G_ETA, EREB_G_GGG, EREB_G_SXY0, EREB_G_COH, EREB_G_TEN,
    EREB_G_FRI, EREB_G_WYX, EREB_G_DSXY,
    /* Vx nodes (Ny1, Nx1) */
    EREB_G_RHOXSUM, EREB_G_RHOFXSUM, EREB_G_KXSUM, EREB_G_PHIXSUM,
    EREB_G_RXSUM, EREB_G_WTXSUM,
    EREB_G_RHOX, EREB_G_RHOFX, EREB_G_KX, EREB_G_PHIX, EREB_G_RX,
    EREB_G_VX, EREB_G_VXF, EREB_G_GX,
    /* Vy nodes (Ny1, Nx1) */
    EREB_G_RHOYSUM, EREB_G_RHOFYSUM, EREB_G_KYSUM, EREB_G_PHIYSUM,
    EREB_G_RYSUM, EREB_G_WTYSUM,
    EREB_G_RHOY, EREB_G_RHOFY, EREB_G_KY, EREB_G_PHIY, EREB_G_RY,
    EREB_G_VY, EREB_G_VYF, EREB_G_GY,
    /* P nodes (Ny1, Nx1) */
    EREB_G_RHOSUM, EREB_G_RHOCPSUM, EREB_G_ALPHASUM, EREB_G_ALPHAFSUM,
    EREB_G_HRSUM, EREB_G_GGGPSUM, EREB_G_SXXSUM, EREB_G_TKSUM,
    EREB_G_PHISUM, EREB_G_WTPSUM,
    EREB_G_RHO, EREB_G_RHOCP, EREB_G_ALPHA, EREB_G_ALPHAF, EREB_G_HR,
    EREB_G_GGGP, EREB_G_SXX0, EREB_G_TK1, EREB_G_PHI, EREB_G_BETAPHI,
    EREB_G_ETAP, EREB_G_ETAPHI, EREB_G_PR0, EREB_G_PF0, EREB_G_DMP,
    EREB_G_DT, EREB_G_TK2, EREB_G_DSXX, EREB_G_HA, EREB_G_HS, EREB_G_DHP,
    /* inverse viscoplastic viscosity at yielding basic nodes (Ny, Nx), read
     * by update_marker_viscosity! (E:774) */
    EREB_G_YNY_INV_ETA,
    /* P nodes: Dln[(1-phi)/phi]/Dt (E:4822), melt composition (E:1929),
     * fluid pressure and the thermochemical transfer sums (E:2558) */
    EREB_G_APHI, EREB_G_XWS, EREB_G_XWSSUM, EREB_G_PF, EREB_G_DMPSUM,
    EREB_G_DHPSUM,
    /* P nodes: total and solid pressure and the solid pressure of the
     * previous generation (E:5284) */
    EREB_G_PR, EREB_G_PS, EREB_G_PS0,
    EREB_G_NFIELDS
};
int ereb_grid_set(ereb_ctx* ctx, int field, const double* host);
int ereb_grid_get(ereb_ctx* ctx, int field, double* host);
/* YNY (Matrix{Bool}, (Ny, Nx)), one byte per node, 0/1 */
int ereb_yny_set(ereb_ctx* ctx, const uint8_t* host);
int ereb_yny_get(ereb_ctx* ctx, uint8_t* host);
/* number of elements of a grid field (Ny*Nx or Ny1*Nx1) */
int64_t ereb_grid_numel(const ereb_ctx* ctx, int field);

/* ---- save_state field set (E:5718-5957) --------------------------------- */
/* save_state writes one JLD2 record per output step (calls E:6193, E:7106).
 * Of its arrays, the library holds 76: 52 node grids, YNY and 23 marker
 * vectors.  Entry k of the set, in the order of the record's keyword list
 * (E:5864-5957), is described by ereb_save_field: the record's key, its kind
 * and its ereb_grid_field / ereb_marker_field value.  The record's scalars,
 * axes and the 17 arrays produced by the host solves and plastic iterations
 * (EXY, SXY, qxD, qyD, EXX, SXX, vxp, vyp, vxpf, vypf, FI, ETA5, ETA00, YNY5,
 * YNY00, EII, SII) stay with the caller, as does the JLD2 write. */
#define EREB_SAVE_GRID    0 /* double, ereb_grid_numel(field) entries */
#define EREB_SAVE_YNY     1 /* uint8, Ny*Nx entries */
#define EREB_SAVE_MARKER  2 /* marknum entries: double, tm as int64 */
#define EREB_SAVE_NFIELDS 76
int ereb_save_field(int k, const char** name, int* kind, int* field);
/* One call: dst[k] (EREB_SAVE_NFIELDS host pointers; NULL skips entry k)
 * receives entry k of the set, as ereb_grid_get / ereb_yny_get /
 * ereb_markers_get would write it; marker vectors get `marknum` entries. */
int ereb_save_state_fields(ereb_ctx* ctx, void* const* dst, int64_t marknum);

/* ---- resident steps ----------------------------------------------------- */
/* Step 1 + 2: the marker loop E:6419-6517 (compute_marker_properties! E:700
 * then marker_to_{basic,vx,vy,p}_nodes! E:1605-1863) after
 * reset_interpolated_properties! (E:968): writes the 7 marker property
 * vectors and the 30 *SUM grids.  hrsolidm/hrfluidm are the per-type
 * radiogenic heat productions from calculate_radioactive_heating (E:1336). */
int ereb_step_scatter(ereb_ctx* ctx, const double hrsolidm[3],
                      const double hrfluidm[3], int64_t mode);
/* compute_{basic,vx,vy,p}_node_properties! (E:1970, E:2033, E:2086, E:2148)
 * then apply_insulating_boundary_conditions!(tk1) (E:2432, call E:6604) */
int ereb_step_node_properties(ereb_ctx* ctx);
/* update_marker_temperature! (E:4782): tk2 grid at timestep 1, DT otherwise */
int ereb_step_gather_temperature(ereb_ctx* ctx, int64_t timestep);
/* update_marker_stress! (E:4354) with DSXX (P) and DSXY (basic) grids */
int ereb_step_gather_stress(ereb_ctx* ctx);
/* move_markers_rk4! (E:4990) with vx, vy, vxf, vyf, wyx, tk2 grids */
int ereb_step_advect_rk4(ereb_ctx* ctx, double dt, int64_t mode);
/* assemble_hydromechanical_lse! (E:3089); the CSC matrix and R stay on the
 * device until ereb_csc_get; *nnz receives the number of stored entries */
int ereb_step_assemble_stokes_darcy(ereb_ctx* ctx, double dt, int64_t* nnz);
/* assemble_thermal_lse! (E:4526) */
int ereb_step_assemble_thermal(ereb_ctx* ctx, double dt, int64_t* nnz);

/* The whole transfer step on the resident state, in the reference's order:
 * ereb_step_scatter, ereb_step_node_properties,
 * ereb_step_assemble_stokes_darcy, ereb_step_assemble_thermal,
 * ereb_step_gather_stress, ereb_step_gather_temperature(timestep),
 * ereb_step_advect_rk4 (E:6419-7053 without the host solves); the two
 * gathers and RK4 run as one kernel (kernel class EREB_K_RK4), on a second
 * stream of the context once the scatter is done, beside the node pass and
 * the assemblies, which read none of the marker state it writes; the call
 * returns with the context's stream ordered after both.  With graph replay
 * on (ereb_ctx_set_graphs), after the first call the step's launches are
 * captured once into four graphs (the two streams' parts) and replayed; they
 * are captured again when the mode, the marker count, the precision form or
 * a device buffer changes (not when dt, the heating rates or the timestep
 * change).  Results are those of the separate calls, bit for bit.  Every
 * launch in sequence on the context's stream while profiling is enabled. */
int ereb_step_transfer(ereb_ctx* ctx, const double hrsolidm[3],
                       const double hrfluidm[3], int64_t mode, double dt,
                       int64_t timestep);
/* graph replay of ereb_step_transfer on or off (default) */
int ereb_ctx_set_graphs(ereb_ctx* ctx, int on);
/* number of captures of the step graph and of replays since the context was
 * created.  dt, hrsolidm/hrfluidm and timestep == 1 are read by the kernels
 * from a device block written before each replay, so they change from step to
 * step without a new capture, as in the reference loop where dt follows the
 * Courant and temperature-change limits (E:6619, E:6790, E:6933) and the
 * heating rates follow calculate_radioactive_heating(timesum) (E:6413). */
int ereb_step_graph_stats(const ereb_ctx* ctx, int64_t* captures, int64_t* replays);

#define EREB_SYS_STOKES_DARCY 0
#define EREB_SYS_THERMAL      1
/* Copy the last assembled system to the host as a 1-based
 * SparseMatrixCSC{Float64,Int64}: colptr (n+1), rowval (nnz), nzval (nnz),
 * rhs (n).  Any pointer may be NULL to skip that part. */
int ereb_csc_get(ereb_ctx* ctx, int which, int64_t* colptr, int64_t* rowval,
                 double* nzval, double* rhs);
int ereb_csc_size(const ereb_ctx* ctx, int which, int64_t* n, int64_t* nnz);

/* Solver hand-off.  The reference gives L to LinearSolve's UMFPACK cache at
 * every plastic iteration (`hydromech_cache.A = L`, E:6719-6749) and solves
 * LT with `\` (E:6925).  UMFPACK keeps its symbolic analysis while the
 * sparsity pattern is unchanged, and the caller then needs only nzval and
 * the right-hand side: ereb_csc_get(ctx, which, NULL, NULL, nzval, rhs).
 * *changed = 0 when the pattern of the last assembled system equals the one
 * last fetched by ereb_csc_get with colptr or rowval non-NULL, 1 when it
 * differs or none was fetched since the system's buffers were (re)sized. */
int ereb_csc_pattern_changed(ereb_ctx* ctx, int which, int* changed);
/* Page-locked host memory on the context's device for hand-off arrays
 * (device-to-host copies into it run at the link's full rate); release it
 * with ereb_host_free.  bytes == 0 gives NULL. */
int ereb_host_alloc(ereb_ctx* ctx, int64_t bytes, void** out);
int ereb_host_free(ereb_ctx* ctx, void* p);

/* Sort order of the last ereb_step_scatter: perm[k] (0-based marker index)
 * of the k-th marker in sorted order, the fine key of every marker and the
 * number of keys.  The sort is a stable counting sort of marker indices by
 * fine key: the basic cell (fix, E:1456) refined by the P-cell offsets,
 * numbered tile-major (tiles of 6x6 basic cells, column-major; 12x12 fine
 * cells per tile, column-major). */
int ereb_sort_get(ereb_ctx* ctx, int32_t* perm, int32_t* fine_key,
                  int64_t* nkeys);

/* ---- drop-in entry points: the reference's argument lists -------------- */
/* Batched form of the loop E:6419-6517 (compute_marker_properties! +
 * marker_to_*_nodes! for m = 1..marknum).  Writes the property vectors
 * (rhototalm, rhocptotalm, etatotalm, hrtotalm, ktotalm, tkm_rhocptotalm,
 * etafluidcur_inv_kphim) and the 30 SUM grids in the order of
 * reset_interpolated_properties! (E:968-999). */
int ereb_marker_to_nodes(ereb_ctx* ctx,
    const double* xm, const double* ym, const int64_t* tm, const double* tkm,
    const double* phim, const double* XWsolidm0, const double* etavpm,
    const double* inv_gggtotalm, const double* fricttotalm,
    const double* cohestotalm, const double* tenstotalm,
    const double* rhofluidcur, const double* alphasolidcur,
    const double* alphafluidcur, const double* sxxm, const double* sxym,
    double* rhototalm, double* rhocptotalm, double* etatotalm,
    double* hrtotalm, double* ktotalm, double* tkm_rhocptotalm,
    double* etafluidcur_inv_kphim,
    const double hrsolidm[3], const double hrfluidm[3], int64_t mode,
    double* const sums[30], int64_t marknum);

/* compute_basic_node_properties! E:1970 (in: 8 SUM; in/out: 7 grids + YNY) */
int ereb_compute_basic_node_properties(ereb_ctx* ctx,
    const double* ETA0SUM, const double* ETASUM, const double* GGGSUM,
    const double* SXYSUM, const double* COHSUM, const double* TENSUM,
    const double* FRISUM, const double* WTSUM,
    double* ETA0, double* ETA, double* GGG, double* SXY0, double* COH,
    double* TEN, double* FRI, uint8_t* YNY);
/* compute_vx_node_properties! E:2033 */
int ereb_compute_vx_node_properties(ereb_ctx* ctx,
    const double* RHOXSUM, const double* RHOFXSUM, const double* KXSUM,
    const double* PHIXSUM, const double* RXSUM, const double* WTXSUM,
    double* RHOX, double* RHOFX, double* KX, double* PHIX, double* RX);
/* compute_vy_node_properties! E:2086 */
int ereb_compute_vy_node_properties(ereb_ctx* ctx,
    const double* RHOYSUM, const double* RHOFYSUM, const double* KYSUM,
    const double* PHIYSUM, const double* RYSUM, const double* WTYSUM,
    double* RHOY, double* RHOFY, double* KY, double* PHIY, double* RY);
/* compute_p_node_properties! E:2148 */
int ereb_compute_p_node_properties(ereb_ctx* ctx,
    const double* RHOSUM, const double* RHOCPSUM, const double* ALPHASUM,
    const double* ALPHAFSUM, const double* HRSUM, const double* GGGPSUM,
    const double* SXXSUM, const double* TKSUM, const double* PHISUM,
    const double* WTPSUM,
    double* RHO, double* RHOCP, double* ALPHA, double* ALPHAF, double* HR,
    double* GGGP, double* SXX0, double* tk1, double* PHI, double* BETAPHI);
/* apply_insulating_boundary_conditions! E:2432 on a (Nyy, Nxx) array */
int ereb_apply_insulating_boundary_conditions(ereb_ctx* ctx, double* t,
    int64_t Nyy, int64_t Nxx);

/* update_marker_temperature! E:4782 */
int ereb_update_marker_temperature(ereb_ctx* ctx, const double* xm,
    const double* ym, double* tkm, const double* DT, const double* tk2,
    int64_t timestep, int64_t marknum);
/* update_marker_stress! E:4354 */
int ereb_update_marker_stress(ereb_ctx* ctx, const double* xm,
    const double* ym, double* sxxm, double* sxym, const double* DSXX,
    const double* DSXY, int64_t marknum);
/* move_markers_rk4! E:4990 (argument order as the reference: sxym before
 * sxxm) */
int ereb_move_markers_rk4(ereb_ctx* ctx, double* xm, double* ym,
    const int64_t* tm, double* tkm, const double* phim, double* sxym,
    double* sxxm, const double* vx, const double* vy, const double* vxf,
    const double* vyf, const double* wyx, const double* tk2,
    int64_t marknum, double dt, int64_t mode);
/* assemble_hydromechanical_lse! E:3089: writes R (6*Ny1*Nx1) and keeps the
 * CSC matrix L on the device; *nnz receives its size, fetch it with
 * ereb_csc_get(ctx, EREB_SYS_STOKES_DARCY, ...). */
int ereb_assemble_hydromechanical_lse(ereb_ctx* ctx,
    const double* ETA, const double* ETAP, const double* GGG,
    const double* GGGP, const double* SXY0, const double* SXX0,
    const double* RHOX, const double* RHOY, const double* RHOFX,
    const double* RHOFY, const double* RX, const double* RY,
    const double* ETAPHI, const double* BETAPHI, const double* PHI,
    const double* gx, const double* gy, const double* pr0, const double* pf0,
    const double* DMP, double dt, double* R, int64_t* nnz);
/* assemble_thermal_lse! E:4526: writes RT (Ny1*Nx1), keeps LT on device */
int ereb_assemble_thermal_lse(ereb_ctx* ctx, const double* tk1,
    const double* RHOCP, const double* KX, const double* KY, const double* HR,
    const double* HA, const double* HS, const double* DHP, double* RT,
    double dt, int64_t* nnz);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight loops, harnesses that forbid
 * hipMalloc): ereb_assemble_thermal_lse with every buffer on the device and
 * supplied by the caller, allocating nothing.  The 8 input grids and RT are
 * device arrays of Ny1*Nx1 Float64 (column-major); the matrix is written as
 * a 1-based Int64 CSC into colptr (Ny1*Nx1 + 1), rowval and nzval (capacity
 * entries, at least 5*Ny1*Nx1); *nnz (optional, synchronises the stream)
 * receives the entry count.  The kernels' scratch is a static device array
 * per GPU, held until the process ends: grids of more than 524,288 P cells
 * (Ny1*Nx1) are refused with EREB_ERR_ARG.  Calls that overlap on one GPU
 * (other contexts or host threads) share that array and race: serialise
 * them.  Results equal ereb_assemble_thermal_lse's bit for bit. */
int ereb_assemble_thermal_lse_noalloc(ereb_ctx* ctx, const double* tk1,
    const double* RHOCP, const double* KX, const double* KY, const double* HR,
    const double* HA, const double* HS, const double* DHP, double* RT,
    double dt, int64_t* colptr, int64_t* rowval, double* nzval,
    int64_t capacity, int64_t* nnz);

/* ---- adjoint (no counterpart in the reference) -------------------------- */
/* Gradient of L = sum over the P nodes of tk1^2 after `nsteps` timesteps of
 * the transfer map on the resident markers, with respect to the initial
 * marker temperature tkm and porosity phim.  One timestep is: scatter
 * (compute_marker_properties! + marker_to_*_nodes!, E:6419-6517), node
 * properties with the insulating tk1 boundary (E:6522-6604), tk2 := tk1 (no
 * heat solve), the temperature gather at timestep 1 (E:7004; later steps add
 * DT = tk2 - tk1 = 0) and RK4 advection through the resident vx/vy/vxf/vyf
 * grids (E:7036).  tk1 is taken after one more scatter + node pass on the
 * final state.  The reverse pass runs the hand-written adjoint kernels
 * (scatter^T as a gather, node properties^T, gather^T and RK4^T), recomputing
 * the forward steps from checkpoints of (xm, ym, tkm) taken every
 * `checkpoint_every` steps.  While the grid fields of every step fit the
 * context's grid budget (ereb_adjoint_set_grid_budget), the forward pass keeps
 * them, and a recomputed step runs only its sort, temperature gather and RK4;
 * otherwise it runs the scatter and the node pass again.  The gradient is the
 * same bit for bit either way.  Positions do not depend on tkm or phim.
 * loss: required; tk1_out (Ny1*Nx1, column-major), grad_tkm and grad_phim
 * (marknum) may be NULL; with both gradients NULL only the forward map runs.
 * The resident markers are left as they were.  Float64 form only for the
 * gradient. */
int ereb_adjoint_tk1_loss(ereb_ctx* ctx, int64_t nsteps, int64_t checkpoint_every,
                          double dt, int64_t mode, const double hrsolidm[3],
                          const double hrfluidm[3], double* loss, double* tk1_out,
                          double* grad_tkm, double* grad_phim);
/* the most device memory, in bytes, that ereb_adjoint_tk1_loss keeps for the
 * grid fields of every step (default 4 GiB; nsteps times the size of the
 * context's grid fields is compared with it); 0 recomputes every step in
 * full.  Marker-sized records stay one checkpoint segment's either way. */
int ereb_adjoint_set_grid_budget(ereb_ctx* ctx, int64_t bytes);

/* ---- pullbacks: the reverse-mode kernels behind ChainRulesCore rrules --- */
/* Each takes the forward call's inputs (host arrays, as the drop-in entry
 * points) and the cotangents of its outputs, recomputes what it needs (no
 * tape is kept between calls) and writes the cotangents of its inputs.
 * Output cotangent arguments that are NULL are zero; input cotangent
 * arguments that are NULL are not computed.  Cotangents of grids (a marker's
 * cotangent spread to the nodes of its stencils) are summed in fixed point,
 * so every pullback gives the same bits on every run.  Float64 form only.
 * The derivatives are those of the piece of each piecewise function (cell
 * index, property branch) the inputs lie in, as forward-mode automatic
 * differentiation of the reference gives them. */

/* Pullback of ereb_marker_to_nodes (compute_marker_properties! E:700-750 and
 * marker_to_*_nodes! E:1605-1863 over the loop E:6419-6517).  markers[f] are
 * the marker vectors by ereb_marker_field (tm as int64; xm, ym, tm, tkm,
 * phim required, the others as the forward call used them, NULL = 0);
 * sum_bar[30] the cotangents of the 30 SUM grids in the order of
 * reset_interpolated_properties! (E:968-999); prop_bar[7] (may be NULL, or
 * hold NULLs) those of the property outputs rhototalm, rhocptotalm,
 * etatotalm, hrtotalm, ktotalm, tkm_rhocptotalm, etafluidcur_inv_kphim.
 * marker_bar[f] receives the cotangent of input field f: xm, ym, tkm, phim,
 * XWsolidm0, sxxm, sxym, etavpm, inv_gggtotalm, fricttotalm, cohestotalm,
 * tenstotalm, rhofluidcur, alphasolidcur, alphafluidcur, and, for air
 * markers (which keep them, E:738-741), the stored rhototalm, rhocptotalm,
 * etatotalm, hrtotalm, ktotalm (zero for rocks, whose values are replaced).
 * The scatter's transpose is a gather: no reduction across markers. */
int ereb_marker_to_nodes_pullback(ereb_ctx* ctx, const void* const* markers,
    int64_t marknum, const double hrsolidm[3], const double hrfluidm[3],
    int64_t mode, const double* const* sum_bar, const double* const* prop_bar,
    double* const* marker_bar);

/* Pullback of update_marker_temperature! E:4782-4801: tkm_bar is the
 * cotangent of the updated tkm; tkm_bar_in receives that of the tkm passed
 * in (zero at timestep 1), DT_bar or, at timestep 1, tk2_bar that of the
 * grid read ((Ny1, Nx1); the other one is zeroed), xm_bar and ym_bar those
 * of the positions. */
int ereb_update_marker_temperature_pullback(ereb_ctx* ctx, const double* xm,
    const double* ym, const double* DT, const double* tk2, int64_t timestep,
    int64_t marknum, const double* tkm_bar, double* tkm_bar_in,
    double* DT_bar, double* tk2_bar, double* xm_bar, double* ym_bar);

/* Pullback of update_marker_stress! E:4354-4388 (the stress cotangents pass
 * through unchanged): DSXX_bar ((Ny1, Nx1), P clamps 2..Nx-1, 2..Ny-1 as
 * E:4364-4367), DSXY_bar ((Ny, Nx)), xm_bar, ym_bar. */
int ereb_update_marker_stress_pullback(ereb_ctx* ctx, const double* xm,
    const double* ym, const double* DSXX, const double* DSXY, int64_t marknum,
    const double* sxxm_bar, const double* sxym_bar, double* DSXX_bar,
    double* DSXY_bar, double* xm_bar, double* ym_bar);

/* Pullback of move_markers_rk4! E:4990-5257 (argument order as the forward:
 * sxym before sxxm; the marker arrays are the values before the step).
 * Output cotangents xm_bar .. sxxm_bar; input cotangents of the markers
 * (xm, ym, tkm, phim, sxym, sxxm) and of the grids vx, vy, vxf, vyf
 * ((Ny1, Nx1)), wyx ((Ny, Nx)) and tk2 ((Ny1, Nx1)).  The reverse sweep
 * recomputes the stage positions of the solid step and of the fluid
 * back-trace; the centre node of the left-half correction in the fluid
 * stages is read from vx (E:5173), so its cotangent goes to vx_bar. */
int ereb_move_markers_rk4_pullback(ereb_ctx* ctx, const double* xm,
    const double* ym, const int64_t* tm, const double* tkm, const double* phim,
    const double* sxym, const double* sxxm, const double* vx, const double* vy,
    const double* vxf, const double* vyf, const double* wyx, const double* tk2,
    int64_t marknum, double dt, int64_t mode, const double* xm_bar,
    const double* ym_bar, const double* tkm_bar, const double* sxym_bar,
    const double* sxxm_bar, double* xm_bar_in, double* ym_bar_in,
    double* tkm_bar_in, double* phim_bar, double* sxym_bar_in,
    double* sxxm_bar_in, double* vx_bar, double* vy_bar, double* vxf_bar,
    double* vyf_bar, double* wyx_bar, double* tk2_bar);

/* ---- hydration, compaction and viscoplastic viscosity ------------------- */
/* Constants of the compaction and silicate (de)hydration model, C:227-279
 * (plain data so that test and production constant sets can be passed). */
typedef struct {
    double phimin, phimax;        /* C:227, C:229 */
    double RG;                    /* C:232 molar gas constant */
    double VH2Of, VH2Ofi;         /* C:250, C:252: MH2O/rhoH2Of, MH2O/rhoH2Ofi */
    double dHWD, dSWD, dVWD;      /* C:254-258 (dVWD = VDs + VH2Of - VWs) */
    double pfcoeff;               /* C:260 */
    double dtreaction;            /* C:266 */
    double log_completion_rate;   /* C:268, log(0.01) */
    double A_I, b_I, c_I;         /* C:273 */
    double Sxo_B, Tscl_B, To_B;   /* C:276 */
    double Sxo_T, To_T, Ea_T;     /* C:279 */
    int64_t rate_mode;            /* reaction_rate_coeff_mode C:271: 1, 2, 3, 9 */
} ereb_hydration;

/* Resident forms (marker and grid state in HBM; see the drop-in forms
 * below for the argument meaning). */
/* update_marker_viscosity! (E:774-804) for every marker: the @threads loop
 * E:6950-6953; reads the resident YNY and the YNY_INV_ETA grid */
int ereb_step_marker_viscosity(ereb_ctx* ctx);
/* update_marker_porosity! (E:4822-4856) with the APHI grid */
int ereb_step_marker_porosity(ereb_ctx* ctx, const ereb_hydration* hyd, double dt);
/* update_p_nodes_melt_composition! (E:1929-1939): XWSSUM, WTPSUM, XWS */
int ereb_step_melt_composition(ereb_ctx* ctx);
/* perform_thermochemical_reaction! (E:2558-2694) with the pf and tk2 grids:
 * DMPSUM, DHPSUM, WTPSUM, DMP, DHP and the marker hydration state.  An
 * unknown hyd->rate_mode is rejected before any marker is touched (the
 * reference throws from compute_Δtreaction, E:2382-2397, at the first rock
 * marker). */
int ereb_step_thermochemical_reaction(ereb_ctx* ctx, const ereb_hydration* hyd,
                                      double dt, int64_t timestep, int64_t titer);

/* Drop-in forms with the reference's argument lists. */
/* Batched update_marker_viscosity!(m, xm, ym, tm, tkm, etatotalm, etavpm,
 * YNY, YNY_inv_ETA) (E:774) over m = 1..marknum (E:6950-6953).  YNY is
 * Matrix{Bool} (Ny, Nx) as bytes. */
int ereb_update_marker_viscosity(ereb_ctx* ctx, const double* xm,
    const double* ym, const int64_t* tm, const double* tkm, double* etatotalm,
    double* etavpm, const uint8_t* YNY, const double* YNY_inv_ETA,
    int64_t marknum);
/* update_marker_porosity!(xm, ym, tm, phim, APHI, dt, marknum) E:4822 */
int ereb_update_marker_porosity(ereb_ctx* ctx, const ereb_hydration* hyd,
    const double* xm, const double* ym, const int64_t* tm, double* phim,
    const double* APHI, double dt, int64_t marknum);
/* update_p_nodes_melt_composition!(xm, ym, XWsolidm0, XWS, XWSSUM, WTPSUM,
 * marknum) E:1929 */
int ereb_update_p_nodes_melt_composition(ereb_ctx* ctx, const double* xm,
    const double* ym, const double* XWsolidm0, double* XWS, double* XWSSUM,
    double* WTPSUM, int64_t marknum);
/* perform_thermochemical_reaction!(DMP, DHP, DMPSUM, DHPSUM, WTPSUM, pf, tk2,
 * tm, xm, ym, XWsm0, XWsm, phim, phinewm, pfm0, marknum, dt, timestep,
 * titer) E:2558 (the reference's diagnostic @info lines are not emitted) */
int ereb_perform_thermochemical_reaction(ereb_ctx* ctx,
    const ereb_hydration* hyd, double* DMP, double* DHP, double* DMPSUM,
    double* DHPSUM, double* WTPSUM, const double* pf, const double* tk2,
    const int64_t* tm, const double* xm, const double* ym, double* XWsm0,
    double* XWsm, double* phim, double* phinewm, double* pfm0,
    int64_t marknum, double dt, int64_t timestep, int64_t titer);

/* ---- pressure back-trace ------------------------------------------------ */
/* backtrace_pressures_rk4! (E:5284-5496) on the resident grids: PR, PS, PF,
 * vx, vy, vxf, vyf in; PR0, PS0, PF0 out (pr0 .= pr etc., then every
 * interior P node takes the pressures at its RK4 departure point) */
int ereb_step_backtrace_pressures(ereb_ctx* ctx, double dt);
/* drop-in form: backtrace_pressures_rk4!(pr, pr0, ps, ps0, pf, pf0, vx, vy,
 * vxf, vyf, dt); all (Ny1, Nx1) */
int ereb_backtrace_pressures_rk4(ereb_ctx* ctx, const double* pr, double* pr0,
    const double* ps, double* ps0, const double* pf, double* pf0,
    const double* vx, const double* vy, const double* vxf, const double* vyf,
    double dt);

/* ---- marker replenishment ------------------------------------------------ */
/* The marker launch lattice (C:104-135): Nxm x Nym anchor points.  The
 * anchors are the Julia ranges xxm = dxm/2:dxm:xsize-dxm/2 and yym (C:116,
 * C:119), passed element by element; the search half-widths are dii =
 * 5*Nymc and djj = 5*Nxmc (E:5639-5640); mdis_init = 1e30 (C:122). */
typedef struct {
    int64_t Nxm, Nym;
    double dxm, dym;
    const double* xxm;   /* Nxm */
    const double* yym;   /* Nym */
    int64_t dii, djj;
    double mdis_init;
} ereb_marker_grid;

/* replenish_markers! (E:5586-5698) on the resident markers: every anchor
 * takes its nearest marker among those whose marker cell touches it (ties:
 * the lowest marker index, as the sequential strict-< loop); an anchor left
 * without one takes the nearest of the markers found by the anchors within
 * (dii, djj) and, if there is one, gets a new marker appended (anchors in
 * column-major order) with that marker's properties, at the anchor
 * (randomized = 0, the reference's test mode) or jittered by U(-1/2, 1/2)
 * cells (randomized = 1).  The jitter is drawn from a counter-based
 * generator keyed by (seed, anchor), not from Julia's MersenneTwister
 * stream.  mdis (Nym*Nxm) and mnum (Int64, Nym*Nxm; negative for the
 * anchors that received a marker) may be NULL.  *marknum_out receives the
 * new marker count; the resident arrays grow as needed. */
int ereb_step_replenish_markers(ereb_ctx* ctx, const ereb_marker_grid* mg,
    int randomized, uint64_t seed, double* mdis, int64_t* mnum,
    int64_t* marknum_out);
/* Drop-in form with the reference's argument list: the 26 marker vectors
 * (capacity `capacity` each; the first marknum are read), mdis and mnum.
 * Writes the new marker count to *marknum_out; when it exceeds capacity
 * the call returns EREB_ERR_ARG and leaves the host arrays unchanged (the
 * result stays resident: resize and read it with ereb_markers_get). */
int ereb_replenish_markers(ereb_ctx* ctx, double* xm, double* ym, int64_t* tm,
    double* tkm, double* phim, double* sxxm, double* sxym, double* etavpm,
    double* phinewm, double* pfm0, double* XWsolidm, double* XWsolidm0,
    double* rhototalm, double* rhocptotalm, double* etatotalm,
    double* hrtotalm, double* ktotalm, double* inv_gggtotalm,
    double* fricttotalm, double* cohestotalm, double* tenstotalm,
    double* rhofluidcur, double* alphasolidcur, double* alphafluidcur,
    double* tkm_rhocptotalm, double* etafluidcur_inv_kphim, double* mdis,
    int64_t* mnum, int64_t marknum, int64_t capacity,
    const ereb_marker_grid* mg, int randomized, uint64_t seed,
    int64_t* marknum_out);

/* ---- profiling: HIP events around the launches of one kernel class ------ */
enum ereb_kernel_class {
    EREB_K_SORT = 0, EREB_K_SCATTER, EREB_K_COMBINE, EREB_K_NODES,
    EREB_K_GATHER_T, EREB_K_GATHER_S, EREB_K_RK4, EREB_K_ASM_SD,
    EREB_K_ASM_T, EREB_K_VISCOSITY, EREB_K_POROSITY, EREB_K_MELT,
    EREB_K_REACTION, EREB_K_BACKTRACE, EREB_K_REPLENISH, EREB_K_NCLASSES
};
int ereb_profile_enable(ereb_ctx* ctx, int on);
/* total milliseconds and launch count accumulated since the last reset */
int ereb_profile_get(ereb_ctx* ctx, int kclass, double* total_ms,
                     int64_t* count);
int ereb_profile_reset(ereb_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* EREB_H */
