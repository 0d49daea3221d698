// ereb_ctx.hip -- context, resident state and the C-ABI entry points of
// libereb_hip.so (declared in include/ereb.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "ereb_internal.h"

using namespace ereb;

namespace ereb {

int fail(ereb_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hipfail(ereb_ctx* c, hipError_t e, const char* what) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
    return fail(c, EREB_ERR_HIP, buf);
}

GeomK geomk(const ereb_ctx* c) {
    GeomK g;
    g.Nx = c->Nx;
    g.Ny = c->Ny;
    g.Nx1 = c->Nx1;
    g.Ny1 = c->Ny1;
    g.dx = c->dx;
    g.dy = c->dy;
    g.idx = 1.0 / c->dx;
    g.idy = 1.0 / c->dy;
    g.x = c->d_axes;
    g.y = c->d_axes + c->Nx;
    g.xp = c->d_axes + c->Nx + c->Ny;
    g.yp = c->d_axes + c->Nx + c->Ny + c->Nx1;
    return g;
}

bool grid_is_basic(int f) {
    return (f >= EREB_G_ETA0SUM && f <= EREB_G_DSXY) || f == EREB_G_YNY_INV_ETA;
}

double* gptr(ereb_ctx* c, int f) { return c->d_grids + c->goff[f]; }
// In the mixed-precision form the 16 material and derived property vectors
// are stored as float; the state (xm, ym, tkm, phim, sxxm, sxym) stays
// double in both forms and tm is int32.
bool marker_is_f32(int precision, int f) {
    if (precision != EREB_MIXED) return false;
    return f == EREB_M_XWSOLIDM0 || (f >= EREB_M_ETAVPM && f <= EREB_M_ETAFLUIDCUR_INV_KPHIM);
}
size_t marker_es(int precision, int f) {
    return f == EREB_M_TM ? sizeof(int32_t)
                          : (marker_is_f32(precision, f) ? sizeof(float) : sizeof(double));
}
double* mptr(ereb_ctx* c, int f) { return (double*)c->d_m[f]; }
float* fptr(ereb_ctx* c, int f) { return (float*)c->d_m[f]; }
int32_t* tm_ptr(ereb_ctx* c) { return (int32_t*)c->d_m[EREB_M_TM]; }

__global__ void k_cvt_f64_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_cvt_f32_f64(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

int ensure_sys(ereb_ctx* c, int which, int64_t n, int64_t cap, int64_t ncell) {
    Sys& s = c->sys[which];
    if (s.n == n && s.cap >= cap && s.ncell == ncell) return EREB_OK;
    hipFree(s.colcnt);
    hipFree(s.colptr);
    hipFree(s.pres);
    hipFree(s.slotv);
    hipFree(s.rowval);
    hipFree(s.nzval);
    hipFree(s.rhs);
    hipFree(s.pres_fetched);
    hipFree(s.diff);
    hipFree(s.stage);
    hipFree(s.pres_prev);
    hipFree(s.chg);
    hipFree(s.map);
    s.colcnt = s.colptr = s.rowval = s.diff = s.chg = s.map = nullptr;
    s.pres = s.pres_fetched = s.pres_prev = nullptr;
    s.slotv = s.nzval = s.rhs = nullptr;
    s.stage = nullptr;
    s.valid = false;
    s.have_fetched = false;
    s.have_prev = false;
    EREB_HIP(c, hipMalloc(&s.colcnt, sizeof(int32_t) * n));
    EREB_HIP(c, hipMalloc(&s.colptr, sizeof(int32_t) * (n + 1)));
    EREB_HIP(c, hipMalloc(&s.pres, sizeof(uint16_t) * n));
    EREB_HIP(c, hipMalloc(&s.slotv, sizeof(double) * ncell * s.nstot));
    EREB_HIP(c, hipMalloc(&s.rowval, sizeof(int32_t) * cap));
    EREB_HIP(c, hipMalloc(&s.nzval, sizeof(double) * cap));
    EREB_HIP(c, hipMalloc(&s.rhs, sizeof(double) * n));
    EREB_HIP(c, hipMalloc(&s.pres_fetched, sizeof(uint16_t) * n));
    EREB_HIP(c, hipMalloc(&s.diff, sizeof(int32_t)));
    EREB_HIP(c, hipMalloc(&s.stage, sizeof(int64_t) * (std::max(n, cap) + 1)));
    EREB_HIP(c, hipMalloc(&s.pres_prev, sizeof(uint16_t) * n));
    EREB_HIP(c, hipMalloc(&s.chg, sizeof(int32_t)));
    EREB_HIP(c, hipMalloc(&s.map, sizeof(int32_t) * cap));
    s.n = n;
    s.cap = cap;
    s.ncell = ncell;
    return EREB_OK;
}

int ensure_scan(ereb_ctx* c, int64_t n) {
    int64_t nb = (n + 4095) / 4096 + 1;
    if (nb <= c->scan_cap) return EREB_OK;
    hipFree(c->d_scan);
    c->d_scan = nullptr;
    EREB_HIP(c, hipMalloc(&c->d_scan, sizeof(int32_t) * (nb + 1)));
    c->scan_cap = nb;
    return EREB_OK;
}

__global__ void k_set_scalars(StepScalars* __restrict__ d, StepScalars v) {
    if (threadIdx.x == 0) *d = v;
}

int scal_sync(ereb_ctx* c) {
    if (c->scal_ride) return EREB_OK;  // written by the key pass of the step
    if (c->scal_written && std::memcmp(&c->scal, &c->scal_dev, sizeof c->scal) == 0)
        return EREB_OK;
    if (c->capturing)
        return fail(c, EREB_ERR_STATE, "step scalars changed inside a captured step");
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c->stream, c->d_scal, c->scal);
    EREB_HIP(c, hipGetLastError());
    c->scal_dev = c->scal;
    c->scal_written = true;
    return EREB_OK;
}

int set_heating(ereb_ctx* c, const double* hrs, const double* hrf) {
    for (int k = 0; k < 3; ++k) {
        c->scal.hrs[k] = hrs[k];
        c->scal.hrf[k] = hrf[k];
    }
    return scal_sync(c);
}

int set_dt(ereb_ctx* c, double dt) {
    c->scal.dt = dt;
    return scal_sync(c);
}

int set_first(ereb_ctx* c, bool first) {
    c->scal.first = first ? 1 : 0;
    return scal_sync(c);
}

int check_flags(ereb_ctx* c) {
    int32_t f[4];
    EREB_HIP(c, hipMemcpyAsync(f, c->d_flags, sizeof f, hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    if (f[0] || f[1]) {
        int32_t z[4] = {0, 0, 0, 0};
        EREB_HIP(c, hipMemcpy(c->d_flags, z, sizeof z, hipMemcpyHostToDevice));
        if (f[1])
            return fail(c, EREB_ERR_STATE,
                        "assembly: an equation referenced a column outside its stencil "
                        "slot table");
        return fail(c, EREB_ERR_DOMAIN,
                    "marker coordinate outside the supported range (finite |x| or |y| "
                    "beyond about 1e15 m puts its P cell outside the basic-cell tile)");
    }
    return EREB_OK;
}

void prof_begin(ereb_ctx* c, int kclass, hipEvent_t* a) {
    *a = nullptr;
    if (!c->prof) return;
    if (c->ev_pool.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        c->ev_pool.push_back(e);
    }
    *a = c->ev_pool.back();
    c->ev_pool.pop_back();
    hipEventRecord(*a, c->stream);
}

void prof_end(ereb_ctx* c, int kclass, hipEvent_t a) {
    if (!c->prof || !a) return;
    hipEvent_t b;
    if (c->ev_pool.empty()) {
        if (hipEventCreate(&b) != hipSuccess) return;
    } else {
        b = c->ev_pool.back();
        c->ev_pool.pop_back();
    }
    hipEventRecord(b, c->stream);
    c->prof_pending.push_back({kclass, a, b});
}

int prof_collect(ereb_ctx* c) {
    for (auto& p : c->prof_pending) {
        EREB_HIP(c, hipEventSynchronize(p.b));
        float ms = 0.f;
        EREB_HIP(c, hipEventElapsedTime(&ms, p.a, p.b));
        c->prof_ms[p.kclass] += ms;
        c->prof_n[p.kclass] += 1;
        c->ev_pool.push_back(p.a);
        c->ev_pool.push_back(p.b);
    }
    c->prof_pending.clear();
    return EREB_OK;
}

}  // namespace ereb

static int valid_mode(ereb_ctx* c, int64_t mode) {
    if (mode == 1 || mode == 9) return EREB_OK;
    char buf[64];
    snprintf(buf, sizeof buf, "unknown mode %lld", (long long)mode);
    return fail(c, EREB_ERR_MODE, buf);
}

static int reserve(ereb_ctx* c, int64_t n) {
    if (n <= c->cap) return EREB_OK;
    if (n > (int64_t)INT32_MAX) return fail(c, EREB_ERR_ARG, "marknum exceeds 2^31-1");
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    for (int f = 0; f < EREB_M_NFIELDS; ++f) {
        const size_t es = marker_es(c->precision, f);
        void* p = nullptr;
        EREB_HIP(c, hipMalloc(&p, es * n));
        if (c->d_m[f]) {
            EREB_HIP(c, hipMemcpy(p, c->d_m[f], es * c->cap, hipMemcpyDeviceToDevice));
            EREB_HIP(c, hipMemset((char*)p + es * c->cap, 0, es * (n - c->cap)));
            hipFree(c->d_m[f]);
        } else {
            EREB_HIP(c, hipMemset(p, 0, es * n));
        }
        c->d_m[f] = p;
    }
    // the sort buffers: new ones first, so that a failed allocation leaves
    // the old (still large enough for the old capacity) in place
    uint32_t *key = nullptr, *perm = nullptr, *tmp = nullptr;
    hipError_t e = hipMalloc(&key, sizeof(uint32_t) * n);
    if (e == hipSuccess) e = hipMalloc(&perm, sizeof(uint32_t) * n);
    // oversize tile segments are sorted in [2*start, 2*start + pow2(len)),
    // disjoint between tiles and inside [0, 2n)
    if (e == hipSuccess) e = hipMalloc(&tmp, sizeof(uint32_t) * 2 * n);
    if (e != hipSuccess) {
        hipFree(key);
        hipFree(perm);
        EREB_HIP(c, e);
    }
    hipFree(c->d_key);
    hipFree(c->d_perm);
    hipFree(c->d_tmp);
    c->d_key = key;
    c->d_perm = perm;
    c->d_tmp = tmp;
    c->tmp_cap = 2 * n;
    c->cap = n;
    c->have_sort = false;
    ++c->pos_gen;
    return EREB_OK;
}

namespace ereb {
int reserve_markers(ereb_ctx* c, int64_t n) { return reserve(c, n); }
}  // namespace ereb

extern "C" {

int ereb_version(void) { return 1; }

int ereb_host_selfcheck(int64_t* slots_stokes_darcy, int64_t* slots_thermal) {
    return host_slot_check(slots_stokes_darcy, slots_thermal);
}

const char* ereb_last_error(const ereb_ctx* c) {
    return c ? c->err.c_str() : "null context";
}

int ereb_ctx_create(const ereb_geometry* g, const ereb_materials* mat, int precision,
                    int device, ereb_ctx** out) {
    if (!out) return EREB_ERR_ARG;
    *out = nullptr;
    if (!g || !mat) return EREB_ERR_ARG;
    if (precision != EREB_FP64 && precision != EREB_MIXED) return EREB_ERR_ARG;
    if (g->Nx < 3 || g->Ny < 3 || !(g->dx > 0) || !(g->dy > 0)) return EREB_ERR_ARG;
    if (!g->x || !g->y || !g->xvx || !g->yvx || !g->xvy || !g->yvy || !g->xp || !g->yp)
        return EREB_ERR_ARG;
    const int64_t Nx = g->Nx, Ny = g->Ny, Nx1 = Nx + 1, Ny1 = Ny + 1;
    if (Nx > 1000000 || Ny > 1000000) return EREB_ERR_ARG;
    // grids stay within 2^24 nodes (4096 x 4096): kernels index grid fields,
    // and the RK4 coefficient tables of 96 bytes per node, with 32-bit byte
    // offsets
    if (Nx1 * Ny1 > ((int64_t)1 << 24)) return EREB_ERR_ARG;
    // the staggered axes coincide element for element (C:40-64)
    for (int64_t k = 0; k < Nx; ++k)
        if (g->xvx[k] != g->x[k]) return EREB_ERR_ARG;
    for (int64_t k = 0; k < Ny1; ++k)
        if (g->yvx[k] != g->yp[k]) return EREB_ERR_ARG;
    for (int64_t k = 0; k < Nx1; ++k)
        if (g->xvy[k] != g->xp[k]) return EREB_ERR_ARG;
    for (int64_t k = 0; k < Ny; ++k)
        if (g->yvy[k] != g->y[k]) return EREB_ERR_ARG;

    ereb_ctx* c = new (std::nothrow) ereb_ctx();
    if (!c) return EREB_ERR_NOMEM;
    c->device = device;
    c->precision = precision;
    c->Nx = (int)Nx;
    c->Ny = (int)Ny;
    c->Nx1 = (int)Nx1;
    c->Ny1 = (int)Ny1;
    c->dx = g->dx;
    c->dy = g->dy;
    c->mat = *mat;
    int rc = EREB_OK;
    auto bail = [&](int code) {
        *out = c;  // caller reads the message, then destroys
        return code;
    };
#define CK(call)                                             \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return bail(hipfail(c, e_, #call)); \
    } while (0)
    CK(hipSetDevice(device));
    CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        // the second stream with every CU enabled: a stream with a CU mask gets
        // a hardware queue of its own, while plain streams share the runtime's
        // few queues by usage, so that with other streams open in the process
        // both streams of a context could land on one queue and run in turn
        hipDeviceProp_t prop;
        CK(hipGetDeviceProperties(&prop, device));
        std::vector<uint32_t> all((size_t)(prop.multiProcessorCount + 31) / 32, ~0u);
        CK(hipExtStreamCreateWithCUMask(&c->stream2, (uint32_t)all.size(), all.data()));
    }
    // recorded as kernels' stop events (hipExtLaunchKernelGGL): default flags
    CK(hipEventCreate(&c->ev_scatter));
    CK(hipEventCreate(&c->ev_keys));
    CK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    // the no-alloc entry point's scratch on this device (once per device); a
    // failure here is reported by that entry point, which tries again
    if (noalloc_init(c)) c->err.clear();
    // axes
    std::vector<double> ax;
    ax.insert(ax.end(), g->x, g->x + Nx);
    ax.insert(ax.end(), g->y, g->y + Ny);
    ax.insert(ax.end(), g->xp, g->xp + Nx1);
    ax.insert(ax.end(), g->yp, g->yp + Ny1);
    // are all nodes node 0 + k*spacing, rounded once?
    auto exact_axis = [](const double* a, int64_t n, double h) {
        for (int64_t k = 0; k < n; ++k)
            if (std::fma((double)k, h, a[0]) != a[k]) return false;
        return true;
    };
    c->axes_exact = exact_axis(g->x, Nx, g->dx) && exact_axis(g->y, Ny, g->dy) &&
                    exact_axis(g->xp, Nx1, g->dx) && exact_axis(g->yp, Ny1, g->dy);
    CK(hipMalloc(&c->d_mat, sizeof(ereb_materials)));
    CK(hipMemcpy(c->d_mat, &c->mat, sizeof(ereb_materials), hipMemcpyHostToDevice));
    CK(hipMalloc(&c->d_scal, sizeof(StepScalars)));
    CK(hipMalloc(&c->d_rkd, sizeof(double) * 4 * 12 * (size_t)(Nx1 * Ny1)));
    CK(hipMalloc(&c->d_axes, sizeof(double) * ax.size()));
    CK(hipMemcpy(c->d_axes, ax.data(), sizeof(double) * ax.size(), hipMemcpyHostToDevice));
    // grids, 256-byte aligned, after 32 guard elements (the RK4 stencils read
    // one element before a grid at its first node, and discard it)
    size_t off = 32;
    for (int f = 0; f < EREB_G_NFIELDS; ++f) {
        int64_t n = grid_is_basic(f) ? Nx * Ny : Nx1 * Ny1;
        c->gnumel[f] = n;
        c->goff[f] = off;
        off += ((size_t)n + 31) / 32 * 32;
    }
    CK(hipMalloc(&c->d_grids, sizeof(double) * off));
    CK(hipMemset(c->d_grids, 0, sizeof(double) * off));
    CK(hipMalloc(&c->d_yny, (size_t)Nx * Ny));
    CK(hipMemset(c->d_yny, 0, (size_t)Nx * Ny));
    // tiles
    c->nty = (int)((Ny - 1 + TILE - 1) / TILE);
    c->ntx = (int)((Nx - 1 + TILE - 1) / TILE);
    c->ntiles = c->ntx * c->nty;
    c->nkeys = (int64_t)c->ntiles * NFINE;
    CK(hipMalloc(&c->d_fcount, sizeof(int32_t) * c->nkeys));
    // the sort leaves the key counts at zero for the next sort
    CK(hipMemset(c->d_fcount, 0, sizeof(int32_t) * c->nkeys));
    CK(hipMalloc(&c->d_fstart, sizeof(int32_t) * (c->nkeys + 1)));
    CK(hipMalloc(&c->d_big, sizeof(int32_t) * (c->nkeys + 1)));
    CK(hipMalloc(&c->d_flags, sizeof(int32_t) * 4));
    CK(hipMemset(c->d_flags, 0, sizeof(int32_t) * 4));
    CK(hipMalloc(&c->d_part, sizeof(double) * (size_t)c->ntiles * NSUM * 64));
    CK(hipMalloc(&c->d_typed, sizeof(int32_t)));
#undef CK
    rc = ensure_scan(c, std::max<int64_t>(6 * Nx1 * Ny1, c->nkeys));
    if (rc) return bail(rc);
    *out = c;
    return EREB_OK;
}

int ereb_ctx_destroy(ereb_ctx* c) {
    if (!c) return EREB_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& p : c->prof_pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (auto e : c->ev_pool) hipEventDestroy(e);
    for (int f = 0; f < EREB_M_NFIELDS; ++f) hipFree(c->d_m[f]);
    hipFree(c->d_axes);
    hipFree(c->d_mat);
    hipFree(c->d_scal);
    hipFree(c->d_rkd);
    hipFree(c->pb_pool);
    for (hipGraphExec_t e : c->step_exec)
        if (e) hipGraphExecDestroy(e);
    hipFree(c->adj_pool);
    hipFree(c->rep_pool);
    hipFree(c->d_grids);
    hipFree(c->d_yny);
    hipFree(c->d_key);
    hipFree(c->d_perm);
    hipFree(c->d_tmp);
    hipFree(c->d_fcount);
    hipFree(c->d_fstart);
    hipFree(c->d_big);
    hipFree(c->d_flags);
    hipFree(c->d_part);
    hipFree(c->d_typed);
    hipFree(c->d_scan);
    hipFree(c->d_react);
    for (auto& s : c->sys) {
        hipFree(s.colcnt);
        hipFree(s.colptr);
        hipFree(s.pres);
        hipFree(s.slotv);
        hipFree(s.rowval);
        hipFree(s.nzval);
        hipFree(s.rhs);
        hipFree(s.tables);
        hipFree(s.pres_fetched);
        hipFree(s.diff);
        hipFree(s.stage);
        hipFree(s.pres_prev);
        hipFree(s.chg);
        hipFree(s.map);
    }
    for (hipEvent_t e : {c->ev_scatter, c->ev_join, c->ev_keys})
        if (e) hipEventDestroy(e);
    if (c->stream2) hipStreamDestroy(c->stream2);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return EREB_OK;
}

int ereb_ctx_set_precision(ereb_ctx* c, int precision) {
    if (!c) return EREB_ERR_ARG;
    if (precision != EREB_FP64 && precision != EREB_MIXED)
        return fail(c, EREB_ERR_ARG, "precision must be EREB_FP64 or EREB_MIXED");
    if (precision == c->precision) return EREB_OK;
    // convert the resident property vectors between float and double
    for (int f = 0; f < EREB_M_NFIELDS && c->cap > 0; ++f) {
        if (marker_is_f32(precision, f) == marker_is_f32(c->precision, f)) continue;
        void* p = nullptr;
        EREB_HIP(c, hipMalloc(&p, marker_es(precision, f) * c->cap));
        const unsigned nb = (unsigned)((c->cap + 255) / 256);
        if (marker_is_f32(precision, f))
            hipLaunchKernelGGL(k_cvt_f64_f32, dim3(nb), dim3(256), 0, c->stream,
                               (const double*)c->d_m[f], (float*)p, c->cap);
        else
            hipLaunchKernelGGL(k_cvt_f32_f64, dim3(nb), dim3(256), 0, c->stream,
                               (const float*)c->d_m[f], (double*)p, c->cap);
        EREB_HIP(c, hipGetLastError());
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        hipFree(c->d_m[f]);
        c->d_m[f] = p;
    }
    c->precision = precision;
    c->typed = -1;
    return EREB_OK;
}

int ereb_ctx_reserve(ereb_ctx* c, int64_t n) {
    if (!c || n < 0) return EREB_ERR_ARG;
    return reserve(c, n);
}

int ereb_sync(ereb_ctx* c) {
    if (!c) return EREB_ERR_ARG;
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return check_flags(c);
}

int64_t ereb_marknum(const ereb_ctx* c) { return c ? c->marknum : -1; }

int ereb_markers_set(ereb_ctx* c, int field, const void* host, int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    if (field < 0 || field >= EREB_M_NFIELDS || marknum < 0 || (!host && marknum > 0))
        return fail(c, EREB_ERR_ARG, "ereb_markers_set: bad field or size");
    int rc = reserve(c, std::max<int64_t>(marknum, 1));
    if (rc) return rc;
    if (field == EREB_M_TM) {
        const int64_t* t = (const int64_t*)host;
        std::vector<int32_t> t32((size_t)marknum);
        for (int64_t m = 0; m < marknum; ++m) {
            if (t[m] < 1 || t[m] > 3)
                return fail(c, EREB_ERR_ARG, "marker type tm must be 1, 2 or 3");
            t32[m] = (int32_t)t[m];
        }
        EREB_HIP(c, hipMemcpyAsync(c->d_m[field], t32.data(), sizeof(int32_t) * marknum,
                                   hipMemcpyHostToDevice, c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
    } else if (marker_is_f32(c->precision, field) && marknum > 0) {
        const double* h = (const double*)host;
        std::vector<float> f32((size_t)marknum);
        for (int64_t m = 0; m < marknum; ++m) f32[m] = (float)h[m];
        EREB_HIP(c, hipMemcpyAsync(c->d_m[field], f32.data(), sizeof(float) * marknum,
                                   hipMemcpyHostToDevice, c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
    } else if (marknum > 0) {
        EREB_HIP(c, hipMemcpyAsync(c->d_m[field], host, sizeof(double) * marknum,
                                   hipMemcpyHostToDevice, c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (field == EREB_M_XM || field == EREB_M_YM || marknum != c->marknum) ++c->pos_gen;
    if (field == EREB_M_TM || field == EREB_M_INV_GGGTOTALM || field == EREB_M_FRICTTOTALM ||
        field == EREB_M_COHESTOTALM || field == EREB_M_TENSTOTALM ||
        field == EREB_M_RHOFLUIDCUR || field == EREB_M_ALPHASOLIDCUR ||
        field == EREB_M_ALPHAFLUIDCUR || marknum != c->marknum)
        c->typed = -1;
    c->marknum = marknum;
    return EREB_OK;
}

int ereb_markers_get(ereb_ctx* c, int field, void* host, int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    if (field < 0 || field >= EREB_M_NFIELDS || marknum < 0 || marknum > c->cap ||
        (!host && marknum > 0))
        return fail(c, EREB_ERR_ARG, "ereb_markers_get: bad field or size");
    int rc = check_flags(c);
    if (rc) return rc;
    if (marknum == 0) return EREB_OK;
    if (field == EREB_M_TM) {
        std::vector<int32_t> t32((size_t)marknum);
        EREB_HIP(c, hipMemcpyAsync(t32.data(), c->d_m[field], sizeof(int32_t) * marknum,
                                   hipMemcpyDeviceToHost, c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        int64_t* t = (int64_t*)host;
        for (int64_t m = 0; m < marknum; ++m) t[m] = t32[m];
    } else if (marker_is_f32(c->precision, field)) {
        std::vector<float> f32((size_t)marknum);
        EREB_HIP(c, hipMemcpyAsync(f32.data(), c->d_m[field], sizeof(float) * marknum,
                                   hipMemcpyDeviceToHost, c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        double* h = (double*)host;
        for (int64_t m = 0; m < marknum; ++m) h[m] = (double)f32[m];
    } else {
        EREB_HIP(c, hipMemcpyAsync(host, c->d_m[field], sizeof(double) * marknum,
                                   hipMemcpyDeviceToHost, c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
    }
    return EREB_OK;
}

int64_t ereb_grid_numel(const ereb_ctx* c, int field) {
    if (!c || field < 0 || field >= EREB_G_NFIELDS) return -1;
    return c->gnumel[field];
}

int ereb_grid_set(ereb_ctx* c, int field, const double* host) {
    if (!c) return EREB_ERR_ARG;
    if (field < 0 || field >= EREB_G_NFIELDS || !host)
        return fail(c, EREB_ERR_ARG, "ereb_grid_set: bad field");
    EREB_HIP(c, hipMemcpyAsync(gptr(c, field), host, sizeof(double) * c->gnumel[field],
                               hipMemcpyHostToDevice, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}

int ereb_grid_get(ereb_ctx* c, int field, double* host) {
    if (!c) return EREB_ERR_ARG;
    if (field < 0 || field >= EREB_G_NFIELDS || !host)
        return fail(c, EREB_ERR_ARG, "ereb_grid_get: bad field");
    int rc = check_flags(c);
    if (rc) return rc;
    EREB_HIP(c, hipMemcpyAsync(host, gptr(c, field), sizeof(double) * c->gnumel[field],
                               hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}

int ereb_yny_set(ereb_ctx* c, const uint8_t* host) {
    if (!c || !host) return EREB_ERR_ARG;
    EREB_HIP(c, hipMemcpyAsync(c->d_yny, host, (size_t)c->Nx * c->Ny, hipMemcpyHostToDevice,
                               c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}

int ereb_yny_get(ereb_ctx* c, uint8_t* host) {
    if (!c || !host) return EREB_ERR_ARG;
    EREB_HIP(c, hipMemcpyAsync(host, c->d_yny, (size_t)c->Nx * c->Ny, hipMemcpyDeviceToHost,
                               c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}

// ---- resident steps -------------------------------------------------------

int ereb_step_scatter(ereb_ctx* c, const double hrsolidm[3], const double hrfluidm[3],
                      int64_t mode) {
    if (!c || !hrsolidm || !hrfluidm) return EREB_ERR_ARG;
    int rc = valid_mode(c, mode);
    if (rc) return rc;
    if ((rc = reserve(c, std::max<int64_t>(c->marknum, 1)))) return rc;
    if ((rc = launch_scatter(c, hrsolidm, hrfluidm, (int)mode))) return rc;
    return launch_combine(c);
}

// The transfer step in the reference's order: marker properties + scatter
// (E:6419-6517), node properties + insulating tk1 boundary (E:6522-6604),
// Stokes-Darcy and thermal assembly (E:6680, E:6922), stress and
// temperature gathers (E:6980, E:7004), RK4 (E:7036).
// Launches of c->stream go to `s` while it lives (the launch helpers all use
// c->stream)
struct OnStream {
    ereb_ctx* c;
    hipStream_t s0;
    OnStream(ereb_ctx* c_, hipStream_t s) : c(c_), s0(c_->stream) { c->stream = s; }
    ~OnStream() { c->stream = s0; }
};

// The step in four segments on two streams.  The stress and temperature
// gathers and RK4 (one pass over the markers) read the velocity, rotation,
// stress-change and temperature grids given to the step, which no kernel of
// the step writes, and write marker state that the sort and the scatter read;
// the node pass and the two assemblies read the scatter's window sums and
// node grids only.  So the marker pass (SEG_MARKERS) runs on the second
// stream as soon as the scatter (SEG_SCATTER) is done, beside the node pass
// and the assemblies (SEG_NODES), and its coefficient tables (SEG_TABLES,
// from the velocities alone) before it, from the end of the sort's key pass
// (beside the key pass itself, which is bound by HBM traffic, they stretched
// it by more than their own duration); the main stream waits for the second
// at the end.  Results are those of the launches in sequence: no kernel of
// one stream touches what the other's write.
enum { SEG_TABLES = 0, SEG_SCATTER, SEG_MARKERS, SEG_NODES };

static int seg_launches(ereb_ctx* c, int seg, const double* hrs, const double* hrf,
                        int64_t mode, double dt, int64_t timestep) {
    int rc;
    switch (seg) {
        case SEG_TABLES: return launch_gathers_rk4(c, dt, (int)mode, timestep, RK4_TABLES);
        case SEG_SCATTER: return launch_scatter(c, hrs, hrf, (int)mode);
        case SEG_MARKERS: return launch_gathers_rk4(c, dt, (int)mode, timestep, RK4_MARKERS);
        default:
            if ((rc = launch_combine_nodes(c))) return rc;
            if ((rc = launch_assemble_stokes(c, dt))) return rc;
            return launch_assemble_thermal(c, dt);
    }
}

// one segment on stream s: its launches, or the replay of its captured graph
static int seg_run(ereb_ctx* c, int seg, hipStream_t s, bool replay, const double* hrs,
                   const double* hrf, int64_t mode, double dt, int64_t timestep) {
    if (replay) {
        // a segment that launches nothing (no markers) has no graph
        if (c->step_exec[seg]) EREB_HIP(c, hipGraphLaunch(c->step_exec[seg], s));
        return EREB_OK;
    }
    OnStream on(c, s);
    return seg_launches(c, seg, hrs, hrf, mode, dt, timestep);
}

static int transfer_run(ereb_ctx* c, bool replay, const double* hrs, const double* hrf,
                        int64_t mode, double dt, int64_t timestep) {
    int rc;
    // with per-launch timing on: every launch in sequence on the main stream,
    // the coefficient tables and the marker pass timed together as one class
    if (c->prof) {
        for (int seg : {SEG_SCATTER, SEG_NODES})
            if ((rc = seg_launches(c, seg, hrs, hrf, mode, dt, timestep))) return rc;
        return launch_gathers_rk4(c, dt, (int)mode, timestep);
    }
    hipStream_t s1 = c->stream, s2 = c->stream2;
    // eager: the coefficient tables start once the sort's key pass is done,
    // beside the rest of the sort (short launches that leave most of the
    // chip idle); a replay runs them after the scatter (a segment graph
    // holds no event)
    // (the key pass and the scatter record their events themselves, as
    // their completion; a replay records the scatter's after its graph)
    const bool evs = !replay && c->marknum > 0;
    c->rec_keys = c->rec_scatter = evs;
    rc = seg_run(c, SEG_SCATTER, s1, replay, hrs, hrf, mode, dt, timestep);
    const bool keys = evs && !c->rec_keys, scattered = evs && !c->rec_scatter;
    c->rec_keys = c->rec_scatter = false;
    if (!rc && keys) {
        EREB_HIP(c, hipStreamWaitEvent(s2, c->ev_keys, 0));
        rc = seg_run(c, SEG_TABLES, s2, replay, hrs, hrf, mode, dt, timestep);
    }
    if (!rc) {
        if (!scattered) EREB_HIP(c, hipEventRecord(c->ev_scatter, s1));
        EREB_HIP(c, hipStreamWaitEvent(s2, c->ev_scatter, 0));
        if (!keys) rc = seg_run(c, SEG_TABLES, s2, replay, hrs, hrf, mode, dt, timestep);
        if (!rc) rc = seg_run(c, SEG_MARKERS, s2, replay, hrs, hrf, mode, dt, timestep);
    }
    if (!rc) rc = seg_run(c, SEG_NODES, s1, replay, hrs, hrf, mode, dt, timestep);
    // the join also after a failure: later work on the main stream follows
    // everything the second stream was given
    EREB_HIP(c, hipEventRecord(c->ev_join, s2));
    EREB_HIP(c, hipStreamWaitEvent(s1, c->ev_join, 0));
    return rc;
}

static void step_graphs_free(ereb_ctx* c) {
    for (hipGraphExec_t& e : c->step_exec) {
        if (e) hipGraphExecDestroy(e);
        e = nullptr;
    }
    c->step_graphs = false;
}

// each segment captured on the main stream as a graph of its own (a graph of
// one stream's launches replays them back to back)
static int step_capture(ereb_ctx* c, const double* hrs, const double* hrf, int64_t mode,
                        double dt, int64_t timestep) {
    step_graphs_free(c);
    const uint64_t pos0 = c->pos_gen;
    for (int seg = 0; seg < NSTEP_SEG; ++seg) {
        hipGraph_t g = nullptr;
        EREB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        const int rc = seg_launches(c, seg, hrs, hrf, mode, dt, timestep);
        c->capturing = false;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            step_graphs_free(c);
            c->pos_gen = pos0;
            return rc ? rc : hipfail(c, e, "hipStreamEndCapture");
        }
        size_t nodes = 0;
        hipError_t ei = hipGraphGetNodes(g, nullptr, &nodes);
        if (ei == hipSuccess && nodes > 0)
            ei = hipGraphInstantiate(&c->step_exec[seg], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->step_exec[seg] = nullptr;
            step_graphs_free(c);
            c->pos_gen = pos0;
            return hipfail(c, ei, "hipGraphInstantiate");
        }
    }
    c->pos_gen = pos0;  // the captured launches have not run yet
    c->step_graphs = true;
    return EREB_OK;
}

// dt, the heating rates and the first-timestep flag are not part of the key:
// the kernels read them from the context's device block (StepScalars),
// written before each replay
static StepKey step_key(const ereb_ctx* c, int64_t mode) {
    StepKey k;
    std::memset(&k, 0, sizeof k);  // padding included: keys are compared bytewise
    k.marknum = c->marknum;
    k.cap = c->cap;
    k.mode = mode;
    k.precision = c->precision;
    k.typed = c->typed;
    int n = 0;
    for (int f = 0; f < EREB_M_NFIELDS; ++f) k.bufs[n++] = c->d_m[f];
    const void* more[] = {c->d_key, c->d_perm, c->d_tmp, c->d_scan, c->d_grids, c->d_part,
                          c->d_scal};
    for (const void* p : more) k.bufs[n++] = p;
    for (const Sys& s : c->sys) {
        const void* sb[] = {s.colcnt, s.colptr, s.pres, s.slotv, s.rowval, s.nzval};
        for (const void* p : sb) k.bufs[n++] = p;
    }
    return k;
}

int ereb_step_transfer(ereb_ctx* c, const double hrsolidm[3], const double hrfluidm[3],
                       int64_t mode, double dt, int64_t timestep) {
    if (!c || !hrsolidm || !hrfluidm) return EREB_ERR_ARG;
    int rc = valid_mode(c, mode);
    if (rc) return rc;
    if ((rc = reserve(c, std::max<int64_t>(c->marknum, 1)))) return rc;
    // the step's scalars: written to the device block before a replay
    // (outside any capture), or taken by the eager launches (below)
    for (int q = 0; q < 3; ++q) {
        c->scal.hrs[q] = hrsolidm[q];
        c->scal.hrf[q] = hrfluidm[q];
    }
    c->scal.dt = dt;
    c->scal.first = timestep == 1 ? 1 : 0;
    // eager: when asked, while profiling (events around every launch), and
    // for the first step (it makes the allocations a capture must not)
    const bool eager = !c->use_graphs || c->prof || !c->step_warm;
    // eager with markers: the scalars reach the device with the sort's key
    // pass, the step's first kernel that reads them, instead of by a launch
    // of their own; graph replays read them from the block written before
    c->scal_ride = eager && c->marknum > 0;
    if ((rc = scal_sync(c))) return rc;
    // which scatter loads the step takes is part of its key (typed_refresh)
    if ((rc = typed_refresh(c))) return rc;
    if (eager) {
        rc = transfer_run(c, false, hrsolidm, hrfluidm, mode, dt, timestep);
        if (c->scal_ride) {  // not taken (a failure before the key pass)
            c->scal_ride = false;
            if (!rc) rc = fail(c, EREB_ERR_STATE, "step scalars not written");
        }
        if (!rc) c->step_warm = true;
        return rc;
    }
    const StepKey k = step_key(c, mode);
    if (!c->step_graphs || std::memcmp(&k, &c->step_key, sizeof k) != 0) {
        if ((rc = step_capture(c, hrsolidm, hrfluidm, mode, dt, timestep))) return rc;
        c->step_key = k;
        ++c->step_captures;
    }
    if ((rc = transfer_run(c, true, hrsolidm, hrfluidm, mode, dt, timestep))) return rc;
    ++c->step_replays;
    // host state of the replayed step, as the eager launches leave it: the
    // sort is of the positions before RK4, which then moved them
    c->have_sort = true;
    c->sort_gen = c->pos_gen;
    c->sort_m = c->marknum;
    ++c->pos_gen;
    c->sys[0].valid = c->sys[1].valid = true;
    return EREB_OK;
}

int ereb_ctx_set_graphs(ereb_ctx* c, int on) {
    if (!c) return EREB_ERR_ARG;
    c->use_graphs = on != 0;
    return EREB_OK;
}

int ereb_step_graph_stats(const ereb_ctx* c, int64_t* captures, int64_t* replays) {
    if (!c) return EREB_ERR_ARG;
    if (captures) *captures = c->step_captures;
    if (replays) *replays = c->step_replays;
    return EREB_OK;
}

int ereb_step_node_properties(ereb_ctx* c) {
    if (!c) return EREB_ERR_ARG;
    return launch_node_properties(c, true, true, true, true, true);
}

int ereb_step_gather_temperature(ereb_ctx* c, int64_t timestep) {
    if (!c) return EREB_ERR_ARG;
    return launch_gather_temperature(c, timestep);
}

int ereb_step_gather_stress(ereb_ctx* c) {
    if (!c) return EREB_ERR_ARG;
    return launch_gather_stress(c);
}

int ereb_step_advect_rk4(ereb_ctx* c, double dt, int64_t mode) {
    if (!c) return EREB_ERR_ARG;
    int rc = valid_mode(c, mode);
    if (rc) return rc;
    return launch_rk4(c, dt, (int)mode);
}

static int fetch_nnz(ereb_ctx* c, int which, int64_t* nnz) {
    if (!nnz) return EREB_OK;
    int32_t v = 0;
    Sys& s = c->sys[which];
    EREB_HIP(c, hipMemcpyAsync(&v, s.colptr + s.n, sizeof v, hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    *nnz = v;
    return EREB_OK;
}

int ereb_step_assemble_stokes_darcy(ereb_ctx* c, double dt, int64_t* nnz) {
    if (!c) return EREB_ERR_ARG;
    int rc = launch_assemble_stokes(c, dt);
    if (rc) return rc;
    return fetch_nnz(c, EREB_SYS_STOKES_DARCY, nnz);
}

int ereb_step_assemble_thermal(ereb_ctx* c, double dt, int64_t* nnz) {
    if (!c) return EREB_ERR_ARG;
    int rc = launch_assemble_thermal(c, dt);
    if (rc) return rc;
    return fetch_nnz(c, EREB_SYS_THERMAL, nnz);
}

// 0-based int32 CSC indices to the 1-based Int64 of SparseMatrixCSC{Float64,Int64}
__global__ void k_index_1based(const int32_t* __restrict__ a, int64_t* __restrict__ b,
                               int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int64_t)a[i] + 1;
}

// flag[0] = 1 when the two presence tables differ anywhere (every writer
// stores the same value)
__global__ void k_pres_diff(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                            int64_t n, int32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] != b[i]) flag[0] = 1;
}

static unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

static int sys_nnz(ereb_ctx* c, const Sys& s, int64_t* nnz) {
    int32_t v = 0;
    EREB_HIP(c, hipMemcpyAsync(&v, s.colptr + s.n, sizeof v, hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    *nnz = v;
    return EREB_OK;
}

int ereb_csc_size(const ereb_ctx* cc, int which, int64_t* n, int64_t* nnz) {
    if (!cc || which < 0 || which > 1) return EREB_ERR_ARG;
    ereb_ctx* c = const_cast<ereb_ctx*>(cc);
    const Sys& s = c->sys[which];
    if (!s.valid) return EREB_ERR_STATE;
    if (n) *n = s.n;
    if (nnz) return sys_nnz(c, s, nnz);
    return EREB_OK;
}

int ereb_csc_get(ereb_ctx* c, int which, int64_t* colptr, int64_t* rowval, double* nzval,
                 double* rhs) {
    if (!c || which < 0 || which > 1) return EREB_ERR_ARG;
    Sys& s = c->sys[which];
    if (!s.valid) return fail(c, EREB_ERR_STATE, "no system assembled yet");
    int64_t nnz = 0;
    int rc = sys_nnz(c, s, &nnz);
    if (rc) return rc;
    // stream order keeps the staging buffer's two uses apart
    if (colptr) {
        hipLaunchKernelGGL(k_index_1based, dim3(blocks256(s.n + 1)), dim3(256), 0, c->stream,
                           (const int32_t*)s.colptr, s.stage, s.n + 1);
        EREB_HIP(c, hipMemcpyAsync(colptr, s.stage, sizeof(int64_t) * (s.n + 1),
                                   hipMemcpyDeviceToHost, c->stream));
    }
    if (rowval && nnz > 0) {
        hipLaunchKernelGGL(k_index_1based, dim3(blocks256(nnz)), dim3(256), 0, c->stream,
                           (const int32_t*)s.rowval, s.stage, nnz);
        EREB_HIP(c, hipMemcpyAsync(rowval, s.stage, sizeof(int64_t) * nnz,
                                   hipMemcpyDeviceToHost, c->stream));
    }
    if (nzval && nnz > 0)
        EREB_HIP(c, hipMemcpyAsync(nzval, s.nzval, sizeof(double) * nnz, hipMemcpyDeviceToHost,
                                   c->stream));
    if (rhs)
        EREB_HIP(c, hipMemcpyAsync(rhs, s.rhs, sizeof(double) * s.n, hipMemcpyDeviceToHost,
                                   c->stream));
    if (colptr || rowval) {
        EREB_HIP(c, hipMemcpyAsync(s.pres_fetched, s.pres, sizeof(uint16_t) * s.n,
                                   hipMemcpyDeviceToDevice, c->stream));
        s.have_fetched = true;
    }
    EREB_HIP(c, hipGetLastError());
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}

int ereb_csc_pattern_changed(ereb_ctx* c, int which, int* changed) {
    if (!c || which < 0 || which > 1 || !changed) return EREB_ERR_ARG;
    Sys& s = c->sys[which];
    if (!s.valid) return fail(c, EREB_ERR_STATE, "no system assembled yet");
    if (!s.have_fetched) {
        *changed = 1;
        return EREB_OK;
    }
    int rc = fill_i32(c, s.diff, 0, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pres_diff, dim3(blocks256(s.n)), dim3(256), 0, c->stream,
                       (const uint16_t*)s.pres, (const uint16_t*)s.pres_fetched, s.n, s.diff);
    int32_t f = 0;
    EREB_HIP(c, hipMemcpyAsync(&f, s.diff, sizeof f, hipMemcpyDeviceToHost, c->stream));
    EREB_HIP(c, hipGetLastError());
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    *changed = f != 0;
    return EREB_OK;
}

int ereb_host_alloc(ereb_ctx* c, int64_t bytes, void** out) {
    if (!c || !out || bytes < 0) return EREB_ERR_ARG;
    *out = nullptr;
    if (bytes == 0) return EREB_OK;
    EREB_HIP(c, hipSetDevice(c->device));
    EREB_HIP(c, hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault));
    return EREB_OK;
}

int ereb_host_free(ereb_ctx* c, void* p) {
    if (!c) return EREB_ERR_ARG;
    if (p) EREB_HIP(c, hipHostFree(p));
    return EREB_OK;
}

int ereb_sort_get(ereb_ctx* c, int32_t* perm, int32_t* fine_key, int64_t* nkeys) {
    if (!c) return EREB_ERR_ARG;
    if (!c->have_sort) return fail(c, EREB_ERR_STATE, "no scatter has run");
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    if (perm && c->marknum)
        EREB_HIP(c, hipMemcpy(perm, c->d_perm, sizeof(int32_t) * c->marknum, hipMemcpyDeviceToHost));
    if (fine_key && c->marknum)
        EREB_HIP(c, hipMemcpy(fine_key, c->d_key, sizeof(int32_t) * c->marknum, hipMemcpyDeviceToHost));
    if (nkeys) *nkeys = c->nkeys;
    return EREB_OK;
}

// ---- profiling --------------------------------------------------------------

int ereb_profile_enable(ereb_ctx* c, int on) {
    if (!c) return EREB_ERR_ARG;
    c->prof = on != 0;
    return EREB_OK;
}

int ereb_profile_get(ereb_ctx* c, int k, double* total_ms, int64_t* count) {
    if (!c || k < 0 || k >= EREB_K_NCLASSES) return EREB_ERR_ARG;
    int rc = prof_collect(c);
    if (rc) return rc;
    if (total_ms) *total_ms = c->prof_ms[k];
    if (count) *count = c->prof_n[k];
    return EREB_OK;
}

int ereb_profile_reset(ereb_ctx* c) {
    if (!c) return EREB_ERR_ARG;
    int rc = prof_collect(c);
    if (rc) return rc;
    for (int k = 0; k < EREB_K_NCLASSES; ++k) {
        c->prof_ms[k] = 0;
        c->prof_n[k] = 0;
    }
    return EREB_OK;
}

// ---- drop-in entry points -----------------------------------------------------

#define SETM(f, p)                                              \
    do {                                                        \
        int rc_ = ereb_markers_set(c, f, p, marknum);           \
        if (rc_) return rc_;                                    \
    } while (0)
#define GETM(f, p)                                              \
    do {                                                        \
        int rc_ = ereb_markers_get(c, f, p, marknum);           \
        if (rc_) return rc_;                                    \
    } while (0)
#define SETG(f, p)                                              \
    do {                                                        \
        int rc_ = ereb_grid_set(c, f, p);                       \
        if (rc_) return rc_;                                    \
    } while (0)
#define GETG(f, p)                                              \
    do {                                                        \
        int rc_ = ereb_grid_get(c, f, p);                       \
        if (rc_) return rc_;                                    \
    } while (0)

static const int kSumField[NSUM] = {
    EREB_G_ETA0SUM, EREB_G_ETASUM, EREB_G_GGGSUM, EREB_G_SXYSUM, EREB_G_COHSUM,
    EREB_G_TENSUM, EREB_G_FRISUM, EREB_G_WTSUM,
    EREB_G_RHOXSUM, EREB_G_RHOFXSUM, EREB_G_KXSUM, EREB_G_PHIXSUM, EREB_G_RXSUM, EREB_G_WTXSUM,
    EREB_G_RHOYSUM, EREB_G_RHOFYSUM, EREB_G_KYSUM, EREB_G_PHIYSUM, EREB_G_RYSUM, EREB_G_WTYSUM,
    EREB_G_RHOSUM, EREB_G_RHOCPSUM, EREB_G_ALPHASUM, EREB_G_ALPHAFSUM, EREB_G_HRSUM,
    EREB_G_GGGPSUM, EREB_G_SXXSUM, EREB_G_TKSUM, EREB_G_PHISUM, EREB_G_WTPSUM};

int ereb_marker_to_nodes(ereb_ctx* c, const double* xm, const double* ym, const int64_t* tm,
    const double* tkm, const double* phim, const double* XWsolidm0, const double* etavpm,
    const double* inv_gggtotalm, const double* fricttotalm, const double* cohestotalm,
    const double* tenstotalm, const double* rhofluidcur, const double* alphasolidcur,
    const double* alphafluidcur, const double* sxxm, const double* sxym, double* rhototalm,
    double* rhocptotalm, double* etatotalm, double* hrtotalm, double* ktotalm,
    double* tkm_rhocptotalm, double* etafluidcur_inv_kphim, const double hrsolidm[3],
    const double hrfluidm[3], int64_t mode, double* const sums[30], int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    if (!sums || marknum < 0) return fail(c, EREB_ERR_ARG, "ereb_marker_to_nodes: bad arguments");
    int rc = valid_mode(c, mode);
    if (rc) return rc;
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_TM, tm);
    SETM(EREB_M_TKM, tkm);
    SETM(EREB_M_PHIM, phim);
    SETM(EREB_M_XWSOLIDM0, XWsolidm0);
    SETM(EREB_M_ETAVPM, etavpm);
    SETM(EREB_M_INV_GGGTOTALM, inv_gggtotalm);
    SETM(EREB_M_FRICTTOTALM, fricttotalm);
    SETM(EREB_M_COHESTOTALM, cohestotalm);
    SETM(EREB_M_TENSTOTALM, tenstotalm);
    SETM(EREB_M_RHOFLUIDCUR, rhofluidcur);
    SETM(EREB_M_ALPHASOLIDCUR, alphasolidcur);
    SETM(EREB_M_ALPHAFLUIDCUR, alphafluidcur);
    SETM(EREB_M_SXXM, sxxm);
    SETM(EREB_M_SXYM, sxym);
    // stored air properties are inputs (E:651-655), rock values are outputs
    SETM(EREB_M_RHOTOTALM, rhototalm);
    SETM(EREB_M_RHOCPTOTALM, rhocptotalm);
    SETM(EREB_M_ETATOTALM, etatotalm);
    SETM(EREB_M_HRTOTALM, hrtotalm);
    SETM(EREB_M_KTOTALM, ktotalm);
    rc = ereb_step_scatter(c, hrsolidm, hrfluidm, mode);
    if (rc) return rc;
    GETM(EREB_M_RHOTOTALM, rhototalm);
    GETM(EREB_M_RHOCPTOTALM, rhocptotalm);
    GETM(EREB_M_ETATOTALM, etatotalm);
    GETM(EREB_M_HRTOTALM, hrtotalm);
    GETM(EREB_M_KTOTALM, ktotalm);
    GETM(EREB_M_TKM_RHOCPTOTALM, tkm_rhocptotalm);
    GETM(EREB_M_ETAFLUIDCUR_INV_KPHIM, etafluidcur_inv_kphim);
    for (int s = 0; s < NSUM; ++s) {
        if (!sums[s]) continue;
        GETG(kSumField[s], sums[s]);
    }
    return EREB_OK;
}

int ereb_compute_basic_node_properties(ereb_ctx* c, const double* ETA0SUM, const double* ETASUM,
    const double* GGGSUM, const double* SXYSUM, const double* COHSUM, const double* TENSUM,
    const double* FRISUM, const double* WTSUM, double* ETA0, double* ETA, double* GGG,
    double* SXY0, double* COH, double* TEN, double* FRI, uint8_t* YNY) {
    if (!c || !YNY) return EREB_ERR_ARG;
    SETG(EREB_G_ETA0SUM, ETA0SUM); SETG(EREB_G_ETASUM, ETASUM); SETG(EREB_G_GGGSUM, GGGSUM);
    SETG(EREB_G_SXYSUM, SXYSUM); SETG(EREB_G_COHSUM, COHSUM); SETG(EREB_G_TENSUM, TENSUM);
    SETG(EREB_G_FRISUM, FRISUM); SETG(EREB_G_WTSUM, WTSUM);
    SETG(EREB_G_ETA0, ETA0); SETG(EREB_G_ETA, ETA); SETG(EREB_G_GGG, GGG);
    SETG(EREB_G_SXY0, SXY0); SETG(EREB_G_COH, COH); SETG(EREB_G_TEN, TEN); SETG(EREB_G_FRI, FRI);
    int rc = ereb_yny_set(c, YNY);
    if (rc) return rc;
    if ((rc = launch_node_properties(c, true, false, false, false, false))) return rc;
    GETG(EREB_G_ETA0, ETA0); GETG(EREB_G_ETA, ETA); GETG(EREB_G_GGG, GGG);
    GETG(EREB_G_SXY0, SXY0); GETG(EREB_G_COH, COH); GETG(EREB_G_TEN, TEN); GETG(EREB_G_FRI, FRI);
    return ereb_yny_get(c, YNY);
}

int ereb_compute_vx_node_properties(ereb_ctx* c, const double* RHOXSUM, const double* RHOFXSUM,
    const double* KXSUM, const double* PHIXSUM, const double* RXSUM, const double* WTXSUM,
    double* RHOX, double* RHOFX, double* KX, double* PHIX, double* RX) {
    if (!c) return EREB_ERR_ARG;
    SETG(EREB_G_RHOXSUM, RHOXSUM); SETG(EREB_G_RHOFXSUM, RHOFXSUM); SETG(EREB_G_KXSUM, KXSUM);
    SETG(EREB_G_PHIXSUM, PHIXSUM); SETG(EREB_G_RXSUM, RXSUM); SETG(EREB_G_WTXSUM, WTXSUM);
    SETG(EREB_G_RHOX, RHOX); SETG(EREB_G_RHOFX, RHOFX); SETG(EREB_G_KX, KX);
    SETG(EREB_G_PHIX, PHIX); SETG(EREB_G_RX, RX);
    int rc = launch_node_properties(c, false, true, false, false, false);
    if (rc) return rc;
    GETG(EREB_G_RHOX, RHOX); GETG(EREB_G_RHOFX, RHOFX); GETG(EREB_G_KX, KX);
    GETG(EREB_G_PHIX, PHIX); GETG(EREB_G_RX, RX);
    return EREB_OK;
}

int ereb_compute_vy_node_properties(ereb_ctx* c, const double* RHOYSUM, const double* RHOFYSUM,
    const double* KYSUM, const double* PHIYSUM, const double* RYSUM, const double* WTYSUM,
    double* RHOY, double* RHOFY, double* KY, double* PHIY, double* RY) {
    if (!c) return EREB_ERR_ARG;
    SETG(EREB_G_RHOYSUM, RHOYSUM); SETG(EREB_G_RHOFYSUM, RHOFYSUM); SETG(EREB_G_KYSUM, KYSUM);
    SETG(EREB_G_PHIYSUM, PHIYSUM); SETG(EREB_G_RYSUM, RYSUM); SETG(EREB_G_WTYSUM, WTYSUM);
    SETG(EREB_G_RHOY, RHOY); SETG(EREB_G_RHOFY, RHOFY); SETG(EREB_G_KY, KY);
    SETG(EREB_G_PHIY, PHIY); SETG(EREB_G_RY, RY);
    int rc = launch_node_properties(c, false, false, true, false, false);
    if (rc) return rc;
    GETG(EREB_G_RHOY, RHOY); GETG(EREB_G_RHOFY, RHOFY); GETG(EREB_G_KY, KY);
    GETG(EREB_G_PHIY, PHIY); GETG(EREB_G_RY, RY);
    return EREB_OK;
}

int ereb_compute_p_node_properties(ereb_ctx* c, const double* RHOSUM, const double* RHOCPSUM,
    const double* ALPHASUM, const double* ALPHAFSUM, const double* HRSUM, const double* GGGPSUM,
    const double* SXXSUM, const double* TKSUM, const double* PHISUM, const double* WTPSUM,
    double* RHO, double* RHOCP, double* ALPHA, double* ALPHAF, double* HR, double* GGGP,
    double* SXX0, double* tk1, double* PHI, double* BETAPHI) {
    if (!c) return EREB_ERR_ARG;
    SETG(EREB_G_RHOSUM, RHOSUM); SETG(EREB_G_RHOCPSUM, RHOCPSUM); SETG(EREB_G_ALPHASUM, ALPHASUM);
    SETG(EREB_G_ALPHAFSUM, ALPHAFSUM); SETG(EREB_G_HRSUM, HRSUM); SETG(EREB_G_GGGPSUM, GGGPSUM);
    SETG(EREB_G_SXXSUM, SXXSUM); SETG(EREB_G_TKSUM, TKSUM); SETG(EREB_G_PHISUM, PHISUM);
    SETG(EREB_G_WTPSUM, WTPSUM);
    SETG(EREB_G_RHO, RHO); SETG(EREB_G_RHOCP, RHOCP); SETG(EREB_G_ALPHA, ALPHA);
    SETG(EREB_G_ALPHAF, ALPHAF); SETG(EREB_G_HR, HR); SETG(EREB_G_GGGP, GGGP);
    SETG(EREB_G_SXX0, SXX0); SETG(EREB_G_TK1, tk1); SETG(EREB_G_PHI, PHI);
    SETG(EREB_G_BETAPHI, BETAPHI);
    int rc = launch_node_properties(c, false, false, false, true, false);
    if (rc) return rc;
    GETG(EREB_G_RHO, RHO); GETG(EREB_G_RHOCP, RHOCP); GETG(EREB_G_ALPHA, ALPHA);
    GETG(EREB_G_ALPHAF, ALPHAF); GETG(EREB_G_HR, HR); GETG(EREB_G_GGGP, GGGP);
    GETG(EREB_G_SXX0, SXX0); GETG(EREB_G_TK1, tk1); GETG(EREB_G_PHI, PHI);
    GETG(EREB_G_BETAPHI, BETAPHI);
    return EREB_OK;
}

int ereb_apply_insulating_boundary_conditions(ereb_ctx* c, double* t, int64_t Nyy, int64_t Nxx) {
    if (!c || !t || Nyy < 0 || Nxx < 0) return EREB_ERR_ARG;
    if (!(Nyy > 2 && Nxx > 2)) return EREB_OK;  // E:2435
    double* d = nullptr;
    size_t bytes = sizeof(double) * Nyy * Nxx;
    EREB_HIP(c, hipMalloc(&d, bytes));
    int rc = EREB_OK;
    if (hipMemcpy(d, t, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = EREB_ERR_HIP;
    if (!rc) rc = launch_insulating_bc(c, d, Nyy, Nxx);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = EREB_ERR_HIP;
    if (!rc && hipMemcpy(t, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = EREB_ERR_HIP;
    hipFree(d);
    if (rc == EREB_ERR_HIP) return fail(c, rc, "HIP error in insulating boundary conditions");
    return rc;
}

int ereb_update_marker_temperature(ereb_ctx* c, const double* xm, const double* ym, double* tkm,
    const double* DT, const double* tk2, int64_t timestep, int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_TKM, tkm);
    if (timestep == 1) { SETG(EREB_G_TK2, tk2); } else { SETG(EREB_G_DT, DT); }
    int rc = launch_gather_temperature(c, timestep);
    if (rc) return rc;
    GETM(EREB_M_TKM, tkm);
    return EREB_OK;
}

int ereb_update_marker_stress(ereb_ctx* c, const double* xm, const double* ym, double* sxxm,
    double* sxym, const double* DSXX, const double* DSXY, int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_SXXM, sxxm);
    SETM(EREB_M_SXYM, sxym);
    SETG(EREB_G_DSXX, DSXX);
    SETG(EREB_G_DSXY, DSXY);
    int rc = launch_gather_stress(c);
    if (rc) return rc;
    GETM(EREB_M_SXXM, sxxm);
    GETM(EREB_M_SXYM, sxym);
    return EREB_OK;
}

int ereb_move_markers_rk4(ereb_ctx* c, double* xm, double* ym, const int64_t* tm, double* tkm,
    const double* phim, double* sxym, double* sxxm, const double* vx, const double* vy,
    const double* vxf, const double* vyf, const double* wyx, const double* tk2, int64_t marknum,
    double dt, int64_t mode) {
    if (!c) return EREB_ERR_ARG;
    int rc = valid_mode(c, mode);
    if (rc) return rc;
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_TM, tm);
    SETM(EREB_M_TKM, tkm);
    SETM(EREB_M_PHIM, phim);
    SETM(EREB_M_SXYM, sxym);
    SETM(EREB_M_SXXM, sxxm);
    SETG(EREB_G_VX, vx);
    SETG(EREB_G_VY, vy);
    SETG(EREB_G_VXF, vxf);
    SETG(EREB_G_VYF, vyf);
    SETG(EREB_G_WYX, wyx);
    SETG(EREB_G_TK2, tk2);
    if ((rc = launch_rk4(c, dt, (int)mode))) return rc;
    GETM(EREB_M_XM, xm);
    GETM(EREB_M_YM, ym);
    GETM(EREB_M_TKM, tkm);
    GETM(EREB_M_SXYM, sxym);
    GETM(EREB_M_SXXM, sxxm);
    return EREB_OK;
}

int ereb_assemble_hydromechanical_lse(ereb_ctx* c, const double* ETA, const double* ETAP,
    const double* GGG, const double* GGGP, const double* SXY0, const double* SXX0,
    const double* RHOX, const double* RHOY, const double* RHOFX, const double* RHOFY,
    const double* RX, const double* RY, const double* ETAPHI, const double* BETAPHI,
    const double* PHI, const double* gx, const double* gy, const double* pr0, const double* pf0,
    const double* DMP, double dt, double* R, int64_t* nnz) {
    if (!c || !R) return EREB_ERR_ARG;
    SETG(EREB_G_ETA, ETA); SETG(EREB_G_ETAP, ETAP); SETG(EREB_G_GGG, GGG);
    SETG(EREB_G_GGGP, GGGP); SETG(EREB_G_SXY0, SXY0); SETG(EREB_G_SXX0, SXX0);
    SETG(EREB_G_RHOX, RHOX); SETG(EREB_G_RHOY, RHOY); SETG(EREB_G_RHOFX, RHOFX);
    SETG(EREB_G_RHOFY, RHOFY); SETG(EREB_G_RX, RX); SETG(EREB_G_RY, RY);
    SETG(EREB_G_ETAPHI, ETAPHI); SETG(EREB_G_BETAPHI, BETAPHI); SETG(EREB_G_PHI, PHI);
    SETG(EREB_G_GX, gx); SETG(EREB_G_GY, gy); SETG(EREB_G_PR0, pr0); SETG(EREB_G_PF0, pf0);
    SETG(EREB_G_DMP, DMP);
    int rc = ereb_step_assemble_stokes_darcy(c, dt, nnz);
    if (rc) return rc;
    return ereb_csc_get(c, EREB_SYS_STOKES_DARCY, nullptr, nullptr, nullptr, R);
}

// ---- viscosity, compaction and (de)hydration ----------------------------------

static int valid_hydration(ereb_ctx* c, const ereb_hydration* h) {
    if (!h) return fail(c, EREB_ERR_ARG, "ereb_hydration constants missing");
    const int64_t m = h->rate_mode;
    if (m == 1 || m == 2 || m == 3 || m == 9) return EREB_OK;
    char buf[64];
    snprintf(buf, sizeof buf, "unknown mode %lld", (long long)m);
    return fail(c, EREB_ERR_MODE, buf);
}

int ereb_step_marker_viscosity(ereb_ctx* c) {
    if (!c) return EREB_ERR_ARG;
    return launch_marker_viscosity(c);
}

int ereb_step_marker_porosity(ereb_ctx* c, const ereb_hydration* hyd, double dt) {
    if (!c) return EREB_ERR_ARG;
    if (!hyd) return fail(c, EREB_ERR_ARG, "ereb_hydration constants missing");
    return launch_marker_porosity(c, *hyd, dt);
}

int ereb_step_melt_composition(ereb_ctx* c) {
    if (!c) return EREB_ERR_ARG;
    return launch_melt_composition(c);
}

int ereb_step_thermochemical_reaction(ereb_ctx* c, const ereb_hydration* hyd, double dt,
                                      int64_t timestep, int64_t titer) {
    if (!c) return EREB_ERR_ARG;
    int rc = valid_hydration(c, hyd);
    if (rc) return rc;
    return launch_thermochemical(c, *hyd, dt, timestep, titer);
}

int ereb_update_marker_viscosity(ereb_ctx* c, const double* xm, const double* ym,
    const int64_t* tm, const double* tkm, double* etatotalm, double* etavpm, const uint8_t* YNY,
    const double* YNY_inv_ETA, int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    if (!YNY || !etavpm) return fail(c, EREB_ERR_ARG, "ereb_update_marker_viscosity: null array");
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_TM, tm);
    SETM(EREB_M_TKM, tkm);
    SETM(EREB_M_ETATOTALM, etatotalm);  // air markers keep theirs (E:789-792)
    int rc = ereb_yny_set(c, YNY);
    if (rc) return rc;
    SETG(EREB_G_YNY_INV_ETA, YNY_inv_ETA);
    if ((rc = launch_marker_viscosity(c))) return rc;
    GETM(EREB_M_ETATOTALM, etatotalm);
    GETM(EREB_M_ETAVPM, etavpm);
    return EREB_OK;
}

int ereb_update_marker_porosity(ereb_ctx* c, const ereb_hydration* hyd, const double* xm,
    const double* ym, const int64_t* tm, double* phim, const double* APHI, double dt,
    int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    if (!hyd) return fail(c, EREB_ERR_ARG, "ereb_hydration constants missing");
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_TM, tm);
    SETM(EREB_M_PHIM, phim);
    SETG(EREB_G_APHI, APHI);
    int rc = launch_marker_porosity(c, *hyd, dt);
    if (rc) return rc;
    GETM(EREB_M_PHIM, phim);
    return EREB_OK;
}

int ereb_update_p_nodes_melt_composition(ereb_ctx* c, const double* xm, const double* ym,
    const double* XWsolidm0, double* XWS, double* XWSSUM, double* WTPSUM, int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    if (!XWS || !XWSSUM || !WTPSUM)
        return fail(c, EREB_ERR_ARG, "ereb_update_p_nodes_melt_composition: null array");
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_XWSOLIDM0, XWsolidm0);
    int rc = launch_melt_composition(c);
    if (rc) return rc;
    GETG(EREB_G_XWS, XWS);
    GETG(EREB_G_XWSSUM, XWSSUM);
    GETG(EREB_G_WTPSUM, WTPSUM);
    return EREB_OK;
}

int ereb_perform_thermochemical_reaction(ereb_ctx* c, const ereb_hydration* hyd, double* DMP,
    double* DHP, double* DMPSUM, double* DHPSUM, double* WTPSUM, const double* pf,
    const double* tk2, const int64_t* tm, const double* xm, const double* ym, double* XWsm0,
    double* XWsm, double* phim, double* phinewm, double* pfm0, int64_t marknum, double dt,
    int64_t timestep, int64_t titer) {
    if (!c) return EREB_ERR_ARG;
    int rc = valid_hydration(c, hyd);
    if (rc) return rc;
    if (!DMP || !DHP || !DMPSUM || !DHPSUM || !WTPSUM)
        return fail(c, EREB_ERR_ARG, "ereb_perform_thermochemical_reaction: null array");
    SETM(EREB_M_XM, xm);
    SETM(EREB_M_YM, ym);
    SETM(EREB_M_TM, tm);
    SETM(EREB_M_XWSOLIDM0, XWsm0);
    SETM(EREB_M_XWSOLIDM, XWsm);
    SETM(EREB_M_PHIM, phim);
    SETM(EREB_M_PHINEWM, phinewm);
    SETM(EREB_M_PFM0, pfm0);
    SETG(EREB_G_PF, pf);
    SETG(EREB_G_TK2, tk2);
    if ((rc = launch_thermochemical(c, *hyd, dt, timestep, titer))) return rc;
    GETG(EREB_G_DMP, DMP);
    GETG(EREB_G_DHP, DHP);
    GETG(EREB_G_DMPSUM, DMPSUM);
    GETG(EREB_G_DHPSUM, DHPSUM);
    GETG(EREB_G_WTPSUM, WTPSUM);
    GETM(EREB_M_XWSOLIDM0, XWsm0);
    GETM(EREB_M_XWSOLIDM, XWsm);
    GETM(EREB_M_PHIM, phim);
    GETM(EREB_M_PHINEWM, phinewm);
    GETM(EREB_M_PFM0, pfm0);
    return EREB_OK;
}

int ereb_step_backtrace_pressures(ereb_ctx* c, double dt) {
    if (!c) return EREB_ERR_ARG;
    return launch_backtrace_pressures(c, dt);
}

int ereb_backtrace_pressures_rk4(ereb_ctx* c, const double* pr, double* pr0, const double* ps,
    double* ps0, const double* pf, double* pf0, const double* vx, const double* vy,
    const double* vxf, const double* vyf, double dt) {
    if (!c) return EREB_ERR_ARG;
    if (!pr0 || !ps0 || !pf0)
        return fail(c, EREB_ERR_ARG, "ereb_backtrace_pressures_rk4: null array");
    SETG(EREB_G_PR, pr); SETG(EREB_G_PS, ps); SETG(EREB_G_PF, pf);
    SETG(EREB_G_VX, vx); SETG(EREB_G_VY, vy); SETG(EREB_G_VXF, vxf); SETG(EREB_G_VYF, vyf);
    int rc = launch_backtrace_pressures(c, dt);
    if (rc) return rc;
    GETG(EREB_G_PR0, pr0);
    GETG(EREB_G_PS0, ps0);
    GETG(EREB_G_PF0, pf0);
    return EREB_OK;
}

int ereb_step_replenish_markers(ereb_ctx* c, const ereb_marker_grid* mg, int randomized,
                                uint64_t seed, double* mdis, int64_t* mnum,
                                int64_t* marknum_out) {
    if (!c) return EREB_ERR_ARG;
    if (!mg) return fail(c, EREB_ERR_ARG, "ereb_marker_grid missing");
    int64_t added = 0;
    int rc = launch_replenish(c, *mg, randomized, seed, &added);
    if (rc) return rc;
    const int64_t A = c->rep_n;
    if (mdis)
        EREB_HIP(c, hipMemcpyAsync(mdis, c->rep_mdis, sizeof(double) * A,
                                   hipMemcpyDeviceToHost, c->stream));
    std::vector<int32_t> mn;
    if (mnum) {
        mn.resize((size_t)A);
        EREB_HIP(c, hipMemcpyAsync(mn.data(), c->rep_mnum, sizeof(int32_t) * A,
                                   hipMemcpyDeviceToHost, c->stream));
    }
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t a = 0; mnum && a < A; ++a) mnum[a] = mn[(size_t)a];
    if (marknum_out) *marknum_out = c->marknum;
    return EREB_OK;
}

int ereb_replenish_markers(ereb_ctx* c, double* xm, double* ym, int64_t* tm, double* tkm,
    double* phim, double* sxxm, double* sxym, double* etavpm, double* phinewm, double* pfm0,
    double* XWsolidm, double* XWsolidm0, double* rhototalm, double* rhocptotalm,
    double* etatotalm, double* hrtotalm, double* ktotalm, double* inv_gggtotalm,
    double* fricttotalm, double* cohestotalm, double* tenstotalm, double* rhofluidcur,
    double* alphasolidcur, double* alphafluidcur, double* tkm_rhocptotalm,
    double* etafluidcur_inv_kphim, double* mdis, int64_t* mnum, int64_t marknum,
    int64_t capacity, const ereb_marker_grid* mg, int randomized, uint64_t seed,
    int64_t* marknum_out) {
    if (!c) return EREB_ERR_ARG;
    const struct {
        int f;
        void* p;
    } v[] = {{EREB_M_XM, xm}, {EREB_M_YM, ym}, {EREB_M_TM, tm}, {EREB_M_TKM, tkm},
             {EREB_M_PHIM, phim}, {EREB_M_SXXM, sxxm}, {EREB_M_SXYM, sxym},
             {EREB_M_ETAVPM, etavpm}, {EREB_M_PHINEWM, phinewm}, {EREB_M_PFM0, pfm0},
             {EREB_M_XWSOLIDM, XWsolidm}, {EREB_M_XWSOLIDM0, XWsolidm0},
             {EREB_M_RHOTOTALM, rhototalm}, {EREB_M_RHOCPTOTALM, rhocptotalm},
             {EREB_M_ETATOTALM, etatotalm}, {EREB_M_HRTOTALM, hrtotalm},
             {EREB_M_KTOTALM, ktotalm}, {EREB_M_INV_GGGTOTALM, inv_gggtotalm},
             {EREB_M_FRICTTOTALM, fricttotalm}, {EREB_M_COHESTOTALM, cohestotalm},
             {EREB_M_TENSTOTALM, tenstotalm}, {EREB_M_RHOFLUIDCUR, rhofluidcur},
             {EREB_M_ALPHASOLIDCUR, alphasolidcur}, {EREB_M_ALPHAFLUIDCUR, alphafluidcur},
             {EREB_M_TKM_RHOCPTOTALM, tkm_rhocptotalm},
             {EREB_M_ETAFLUIDCUR_INV_KPHIM, etafluidcur_inv_kphim}};
    static_assert(sizeof v / sizeof v[0] == EREB_M_NFIELDS, "every marker vector");
    for (const auto& e : v)
        if (!e.p) return fail(c, EREB_ERR_ARG, "ereb_replenish_markers: null marker vector");
    if (marknum < 0 || capacity < marknum)
        return fail(c, EREB_ERR_ARG, "ereb_replenish_markers: bad marknum or capacity");
    for (const auto& e : v) SETM(e.f, e.p);
    int64_t n = 0;
    int rc = ereb_step_replenish_markers(c, mg, randomized, seed, mdis, mnum, &n);
    if (rc) return rc;
    if (marknum_out) *marknum_out = n;
    if (n > capacity)
        return fail(c, EREB_ERR_ARG, "ereb_replenish_markers: the new marker count exceeds "
                                     "capacity; the result is resident (ereb_markers_get)");
    for (const auto& e : v)
        if ((rc = ereb_markers_get(c, e.f, e.p, n))) return rc;
    return EREB_OK;
}

int ereb_assemble_thermal_lse(ereb_ctx* c, const double* tk1, const double* RHOCP,
    const double* KX, const double* KY, const double* HR, const double* HA, const double* HS,
    const double* DHP, double* RT, double dt, int64_t* nnz) {
    if (!c || !RT) return EREB_ERR_ARG;
    SETG(EREB_G_TK1, tk1); SETG(EREB_G_RHOCP, RHOCP); SETG(EREB_G_KX, KX); SETG(EREB_G_KY, KY);
    SETG(EREB_G_HR, HR); SETG(EREB_G_HA, HA); SETG(EREB_G_HS, HS); SETG(EREB_G_DHP, DHP);
    int rc = ereb_step_assemble_thermal(c, dt, nnz);
    if (rc) return rc;
    return ereb_csc_get(c, EREB_SYS_THERMAL, nullptr, nullptr, nullptr, RT);
}

}  // extern "C"
