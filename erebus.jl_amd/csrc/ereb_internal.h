// ereb_internal.h -- host-side context of libereb_hip.so and the launch
// interface between the C-ABI layer and the kernel files.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ereb.h"
#include "ereb_device.h"
#include "ereb_fixsum.h"

namespace ereb {

// Scatter tiling: a tile is TILE x TILE basic cells; its node window is
// (TILE+2) x (TILE+2) = 64 nodes per grid, one wavefront per grid.
constexpr int TILE = 6;
constexpr int WIN = TILE + 2;
constexpr int NSUM = 30;  // SUM grids, order of reset_interpolated_properties!

// Scatter window partials: tile t's partial of sum s at node (r, c) of its
// (TILE+2)^2 window, r and c relative to the tile's first node, is
// part[(t * NSUM + s) * 64 + 8 c + r].  A node's sum adds the partials of
// the at most 2 x 2 tiles whose windows hold it, tiles in ascending (tx, ty)
// order; (i0, j0): 0-based node row and column.
struct WinTiles {
    int tylo, tyhi, txlo, txhi;
};
__device__ __forceinline__ WinTiles win_tiles(int i0, int j0, int nty, int ntx) {
    return WinTiles{max(0, (i0 - (WIN - 1) + TILE - 1) / TILE), min(nty - 1, i0 / TILE),
                    max(0, (j0 - (WIN - 1) + TILE - 1) / TILE), min(ntx - 1, j0 / TILE)};
}
__device__ __forceinline__ double win_sum(const double* __restrict__ part, int nty,
                                          const WinTiles& w, int s, int i0, int j0) {
    double acc = 0.0;
    for (int tx = w.txlo; tx <= w.txhi; ++tx)
        for (int ty = w.tylo; ty <= w.tyhi; ++ty) {
            const int lane = (j0 - tx * TILE) * 8 + (i0 - ty * TILE);
            acc += part[((size_t)(tx * nty + ty) * NSUM + s) * 64 + lane];
        }
    return acc;
}
constexpr int FINE = 2 * TILE;  // fine cells per tile side
constexpr int NFINE = FINE * FINE;  // fine cells (sort keys) per tile

struct Sys {
    int64_t n = 0, cap = 0, ncell = 0;
    int ne = 0, nstot = 0;       // equations and stencil slots per P cell
    int maxrev = 0;              // most entries of one column (slot tables)
    void* tables = nullptr;      // device copy of the stencil slot tables
    int32_t* colcnt = nullptr;   // n
    int32_t* colptr = nullptr;   // n + 1, 0-based
    uint16_t* pres = nullptr;    // n: presence bits of each row's slots
    double* slotv = nullptr;     // ncell * nstot: entry values by slot
    int32_t* rowval = nullptr;   // cap, 0-based
    double* nzval = nullptr;     // cap
    double* rhs = nullptr;       // n
    bool valid = false;
    // solver hand-off: the pattern (pres) as of the last host fetch of
    // colptr/rowval, a comparison flag, and 1-based Int64 index staging
    uint16_t* pres_fetched = nullptr;  // n
    int32_t* diff = nullptr;           // 1
    int64_t* stage = nullptr;          // cap + 1
    bool have_fetched = false;
    // values-only assembly: the pattern of the previous assembly, a device
    // flag (pattern changed), and for every stored entry its slot value's
    // index, so that an unchanged pattern refills nzval with one gather
    uint16_t* pres_prev = nullptr;     // n
    int32_t* chg = nullptr;            // 1
    int32_t* map = nullptr;            // cap
    bool have_prev = false;
};

struct ProfPair {
    int kclass;
    hipEvent_t a, b;
};

// What a captured transfer step depends on: its scalar arguments and every
// device buffer its kernels were launched with.  A replay is valid while
// all of them are unchanged.
struct StepKey {
    int64_t marknum, cap, mode;
    int precision, typed;
    const void* bufs[EREB_M_NFIELDS + 19];
};

// segments of the transfer step, each captured as a graph of its own
// (ereb_step_transfer)
constexpr int NSTEP_SEG = 4;

}  // namespace ereb

struct ereb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of the transfer step (default priority, like `stream`):
    // the gathers and RK4 run there beside the node pass and the assemblies;
    // scatter-done, join and key-pass-done events; while rec_keys /
    // rec_scatter is set, the next key pass / scatter launch records ev_keys /
    // ev_scatter as its own completion (hipExtLaunchKernelGGL's stop event,
    // no marker packet between two kernels of the stream) and clears the flag
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_scatter = nullptr, ev_join = nullptr, ev_keys = nullptr;
    bool rec_keys = false, rec_scatter = false;
    int precision = EREB_FP64;
    int Nx = 0, Ny = 0, Nx1 = 0, Ny1 = 0;
    double dx = 0, dy = 0;
    double* d_axes = nullptr;  // x (Nx) | y (Ny) | xp (Nx1) | yp (Ny1)
    ereb_materials mat{};
    ereb_materials* d_mat = nullptr;  // device copy, read by kernels through a pointer
    // per-step scalars (dt, heating rates, first-timestep flag): the kernels
    // read the device block; `scal` is what the host wants there, `scal_dev`
    // what was last written (a write is a one-thread kernel launch)
    ereb::StepScalars* d_scal = nullptr;
    ereb::StepScalars scal{}, scal_dev{};
    bool scal_written = false;
    // set by an eager transfer step with markers: the scalars go to the
    // device block with the sort's key pass (PropArgs::publish) instead of a
    // launch of their own; scal_sync leaves them until then
    bool scal_ride = false;
    bool capturing = false;  // a step is being captured into a HIP graph

    int64_t marknum = 0, cap = 0;
    void* d_m[EREB_M_NFIELDS] = {};

    double* d_grids = nullptr;
    double* d_rkd = nullptr;  // 4 x 12 x Ny1*Nx1: RK4 stage-velocity coefficient tables
    size_t goff[EREB_G_NFIELDS] = {};
    int64_t gnumel[EREB_G_NFIELDS] = {};
    uint8_t* d_yny = nullptr;

    // fine-key sort
    int ntx = 0, nty = 0, ntiles = 0;
    int64_t nkeys = 0;             // ntiles * NFINE
    uint32_t* d_key = nullptr;     // cap
    uint32_t* d_perm = nullptr;    // cap
    uint32_t* d_tmp = nullptr;     // 2*cap scratch: the sort's ranks, then oversize key segments
    int64_t tmp_cap = 0;
    int32_t* d_fcount = nullptr;   // nkeys
    int32_t* d_fstart = nullptr;   // nkeys + 1
    int32_t* d_big = nullptr;      // nkeys + 1: [0] = count, then key ids
    int32_t* d_flags = nullptr;    // [0] domain error
    double* d_part = nullptr;      // ntiles * NSUM * 64
    int32_t* d_scan = nullptr;     // scan block sums
    // thermochemical reaction: (dM, dH)[react_cap] as double2, contributes[react_cap]
    void* d_react = nullptr;
    int64_t react_cap = 0;
    int64_t scan_cap = 0;
    bool have_sort = false;
    // the seven material vectors of the scatter that the reference derives
    // from the marker type (E:659-665) hold exactly those values for every
    // resident marker: 1 yes, 0 no, -1 not checked since they last changed
    int typed = -1;
    int32_t* d_typed = nullptr;  // flag of the check kernel
    bool axes_exact = false;  // node k of every axis == fma(k, spacing, node 0)
    bool fcount_dirty = false;  // a sort stopped between counting and scanning
    // the sort is current while sort_gen == pos_gen; pos_gen advances
    // whenever positions or the marker count may have changed
    uint64_t pos_gen = 1, sort_gen = 0;
    int64_t sort_m = 0;  // marker count that the last completed sort ordered

    ereb::Sys sys[2];

    double* adj_pool = nullptr;  // grow-only device pool of the adjoint driver
    size_t adj_pool_n = 0;       // doubles
    // most bytes of per-step grid fields the adjoint driver keeps
    // (ereb_adjoint_set_grid_budget)
    int64_t adj_grid_budget = int64_t(4) << 30;
    void* pb_pool = nullptr;     // grow-only device pool of the pullbacks and the
                                 // save_state copy (widened marker vectors)
    size_t pb_bytes = 0;

    // marker replenishment: grow-only scratch, and the last call's lattice
    // results (views into it)
    void* rep_pool = nullptr;
    size_t rep_bytes = 0;
    double* rep_mdis = nullptr;
    int32_t* rep_mnum = nullptr;
    int64_t rep_n = 0;

    // graph replay of the resident transfer step (ereb_step_transfer), off
    // by default: its launches on the two streams take less time at every
    // size measured
    bool use_graphs = false;
    bool step_warm = false;  // one eager step has run (allocations made)
    // one graph per segment of the step, replayed on the two streams
    hipGraphExec_t step_exec[ereb::NSTEP_SEG] = {};
    bool step_graphs = false;  // step_exec holds the graphs of step_key
    ereb::StepKey step_key{};
    int64_t step_captures = 0, step_replays = 0;

    bool prof = false;
    std::vector<ereb::ProfPair> prof_pending;
    std::vector<hipEvent_t> ev_pool;
    double prof_ms[EREB_K_NCLASSES] = {};
    int64_t prof_n[EREB_K_NCLASSES] = {};

    std::string err;
};

namespace ereb {

// ---- helpers implemented in ereb_ctx.hip --------------------------------
int fail(ereb_ctx* c, int code, const std::string& msg);
int hipfail(ereb_ctx* c, hipError_t e, const char* what);
#define EREB_HIP(c, call)                                      \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return ereb::hipfail(c, e_, #call); \
    } while (0)

GeomK geomk(const ereb_ctx* c);
bool grid_is_basic(int field);
double* gptr(ereb_ctx* c, int field);
double* mptr(ereb_ctx* c, int field);
float* fptr(ereb_ctx* c, int field);
bool marker_is_f32(int precision, int field);
size_t marker_es(int precision, int field);
int32_t* tm_ptr(ereb_ctx* c);
int ensure_sys(ereb_ctx* c, int which, int64_t n, int64_t cap, int64_t ncell);
int ensure_scan(ereb_ctx* c, int64_t n);
int check_flags(ereb_ctx* c);
// write c->scal to the device block when it differs from the last write
// (stream-ordered; never called while a step is being captured: the values
// a capture needs are written before it)
int scal_sync(ereb_ctx* c);
int set_heating(ereb_ctx* c, const double* hrs, const double* hrf);
int set_dt(ereb_ctx* c, double dt);
int set_first(ereb_ctx* c, bool first);
// grow the resident marker arrays to hold n markers (contents kept)
int reserve_markers(ereb_ctx* c, int64_t n);
void prof_begin(ereb_ctx* c, int kclass, hipEvent_t* a);
void prof_end(ereb_ctx* c, int kclass, hipEvent_t a);
int prof_collect(ereb_ctx* c);

// ---- kernel launchers ---------------------------------------------------
// ereb_scan.hip: exclusive scan of n int32 into out[0..n], out[n] = total;
// optionally a copy of out[0..n) in out2 and in[] zeroed once read
// optionally only when run[0] != 0 (a device-side condition)
int scan_exclusive(ereb_ctx* c, const int32_t* in, int32_t* out, int64_t n,
                   int32_t* out2 = nullptr, bool zero_in = false,
                   const int32_t* run = nullptr);
// the same with a caller-provided block-sum buffer of scan_blocks(n) ints
int64_t scan_blocks(int64_t n);
int scan_exclusive_buf(ereb_ctx* c, const int32_t* in, int32_t* out, int64_t n,
                       int32_t* blocksums, int32_t* out2 = nullptr, bool zero_in = false,
                       const int32_t* run = nullptr);
// fill kernel (graph-safe replacement for hipMemsetAsync)
int fill_i32(ereb_ctx* c, int32_t* p, int32_t v, int64_t n);
int fill_f64(ereb_ctx* c, double* p, double v, int64_t n);

// ereb_sort.hip: stable counting sort of marker indices by fine key; with
// props, the pass that computes the keys also evaluates the marker
// properties (compute_marker_properties!) with these heating rates and mode
struct PropsReq {
    const double* hrs;
    const double* hrf;
    int mode;
};
int launch_fine_sort(ereb_ctx* c, const PropsReq* props = nullptr);
// runs the sort unless the last one still matches the resident positions
int ensure_sort(ereb_ctx* c);

// ereb_scatter.hip: the sort with the marker properties (launch_fine_sort),
// then the scatter of the properties
int launch_scatter(ereb_ctx* c, const double* hrs, const double* hrf, int mode);
int launch_combine(ereb_ctx* c);
// refresh ctx->typed when unknown (a check kernel and a host read; not under
// graph capture)
int typed_refresh(ereb_ctx* c);
// launch_combine + launch_node_properties(all grids, boundary) in one kernel
int launch_combine_nodes(ereb_ctx* c);

// ereb_nodes.hip
int launch_node_properties(ereb_ctx* c, bool basic, bool vx, bool vy, bool p, bool bc);
int launch_insulating_bc(ereb_ctx* c, double* t, int64_t Nyy, int64_t Nxx);

// ereb_markers.hip
int launch_gather_temperature(ereb_ctx* c, int64_t timestep);
int launch_gather_stress(ereb_ctx* c);
// xb_out/yb_out (optional): fluid back-trace end points, for the adjoint
int launch_rk4(ereb_ctx* c, double dt, int mode, double* xb_out = nullptr,
               double* yb_out = nullptr);
// stress gather, temperature gather (timestep) and RK4 in one kernel.  parts:
// RK4_TABLES launches the stage-velocity coefficient tables (they depend on
// the velocity grids only), RK4_MARKERS the marker kernel, which reads them
constexpr int RK4_TABLES = 1, RK4_MARKERS = 2;
int launch_gathers_rk4(ereb_ctx* c, double dt, int mode, int64_t timestep,
                       int parts = RK4_TABLES | RK4_MARKERS);

// ereb_assembly.hip
int launch_assemble_stokes(ereb_ctx* c, double dt);
int launch_assemble_thermal(ereb_ctx* c, double dt);
// Thermal assembly into caller buffers with no allocation: grids in (device
// pointers, (Ny1, Nx1)), RT and a 1-based Int64 CSC out; `scratch` holds at
// least thermal_scratch_bytes(ncell) bytes, 256-byte aligned; dt is read
// from the StepScalars at its start (written by the caller)
size_t thermal_scratch_bytes(int64_t ncell);
struct ThermalScratchHead {  // the pieces at the start of that scratch
    StepScalars* scal;
    int32_t* flags;  // 4 ints; [1] = a slot-table violation
    void* tables;
};
ThermalScratchHead thermal_scratch_head(char* scratch);
int thermal_tables_bytes();
int thermal_tables_host(void* out);  // the slot tables, discovered on the host
int launch_thermal_noalloc(ereb_ctx* c, char* scratch, const double* const in[8], double* RT,
                           int64_t* colptr, int64_t* rowval, double* nzval);

// stencil slot discovery of both systems on the host (no device work):
// slots per cell of the Stokes-Darcy and thermal systems
int host_slot_check(int64_t* slots_sd, int64_t* slots_t);

// grid field of SUM slot s (order of reset_interpolated_properties!)
int sum_grid_field(int s);

// ereb_hydration.hip
int launch_marker_viscosity(ereb_ctx* c);
int launch_marker_porosity(ereb_ctx* c, const ereb_hydration& H, double dt);
int launch_melt_composition(ereb_ctx* c);
int launch_thermochemical(ereb_ctx* c, const ereb_hydration& H, double dt, int64_t timestep,
                          int64_t titer);

// ereb_noalloc.hip: scratch address and thermal slot tables of the no-alloc
// entry point on the context's device (once per device; ereb_ctx_create)
int noalloc_init(ereb_ctx* c);

// ereb_backtrace.hip
int launch_backtrace_pressures(ereb_ctx* c, double dt);

// ereb_replenish.hip: *added = number of markers appended
int launch_replenish(ereb_ctx* c, const ereb_marker_grid& mg, int randomized, uint64_t seed,
                     int64_t* added);

// adjoints (ereb_adjoint.hip, ereb_markers.hip); cotangent grids live in a
// buffer with the layout of the grid fields (offsets goff)
int launch_scatter_adj(ereb_ctx* c, const double* hrs, const double* hrf, int mode,
                       const double* gsum_base, const int64_t* gsum_off, double* gT,
                       double* gphi, bool only_tk);
int launch_nodes_adj(ereb_ctx* c, double* gbase, const int64_t* off);
// the grid cotangent in fixed point (ereb_fixsum.h): pass 0 records the
// largest contribution, pass 1 adds (and writes the marker cotangents)
// order (optional): the markers are walked as order[0], order[1], ... (a
// sort by cell, so a workgroup's stencils cover few nodes)
int launch_gather_T_adj(ereb_ctx* c, int pass, int64_t timestep, double* gT,
                        const FixGrid& ggrid, int* flag, const uint32_t* order = nullptr);
int launch_rk4_adj(ereb_ctx* c, int pass, double dt, int mode, const double* xb,
                   const double* yb, const double* gout, double* gin, double* gphi,
                   const FixGrid& gtk2, int* flag, const uint32_t* order = nullptr);

}  // namespace ereb
