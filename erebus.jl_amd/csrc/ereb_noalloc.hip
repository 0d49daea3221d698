// ereb_noalloc.hip -- assemble_thermal_lse! (E:4526-4611) into caller-owned
// device buffers, allocating nothing: for callers that cannot allocate device
// memory (code under HIP graph capture, tight loops, harnesses that forbid
// hipMalloc).  The kernels are those of ereb_assemble_thermal_lse; their
// scratch (slot values, presence bits, column counts and pointers, scan block
// sums, slot tables, dt and the error flags) is the file-scope array
// `scratch` below, one copy per device, held until the process ends.
//
// Host code never passes `scratch` itself to a kernel or a copy: its name is
// a host-side stand-in.  Its device address on each device is looked up once
// with hipGetSymbolAddress while that device is current (noalloc_init, run
// by ereb_ctx_create) and kept in g_scratch[device].
//
// Every call on a device shares that device's array and leaves its contents
// behind; the parts a call reads before writing (the flags, dt) are written
// by the call itself.  Calls that overlap on one device (two contexts on two
// streams, two host threads) race on the array and must be serialised by
// the caller.
#include <mutex>
#include <vector>

#include "ereb_internal.h"

namespace ereb {

// largest P grid (Ny1 * Nx1 cells) the scratch holds: 724 x 724 nodes; the
// 562 x 562-node grid of the 20M-marker workload needs 315,844
constexpr int64_t NOALLOC_MAX_CELLS = (int64_t)1 << 19;
constexpr size_t NOALLOC_BYTES = (size_t)28 << 20;  // >= thermal_scratch_bytes(max), checked
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_mu;
static char* g_scratch[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

__global__ void k_noalloc_dt(StepScalars* __restrict__ s, double dt) {
    if (threadIdx.x == 0) s->dt = dt;
}

// Set-up, once per device, with c->device current: the scratch address and
// the thermal slot tables (discovered on the host) copied into it.
int noalloc_init(ereb_ctx* c) {
    const int dev = c->device;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(c, EREB_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[dev]) return EREB_OK;
    if (thermal_scratch_bytes(NOALLOC_MAX_CELLS) > NOALLOC_BYTES)
        return fail(c, EREB_ERR_STATE, "no-alloc scratch smaller than its layout");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hipfail(c, e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(c, EREB_ERR_STATE, "hipGetSymbolAddress(scratch) returned null");
    std::vector<char> tables((size_t)thermal_tables_bytes());
    int rc = thermal_tables_host(tables.data());
    if (rc) return fail(c, rc, "thermal stencil slot discovery failed");
    EREB_HIP(c, hipMemcpy(thermal_scratch_head((char*)p).tables, tables.data(), tables.size(),
                          hipMemcpyHostToDevice));
    g_scratch[dev] = (char*)p;
    return EREB_OK;
}

}  // namespace ereb

using namespace ereb;

extern "C" int ereb_assemble_thermal_lse_noalloc(ereb_ctx* c, const double* tk1,
    const double* RHOCP, const double* KX, const double* KY, const double* HR,
    const double* HA, const double* HS, const double* DHP, double* RT, double dt,
    int64_t* colptr, int64_t* rowval, double* nzval, int64_t capacity, int64_t* nnz) {
    if (!c) return EREB_ERR_ARG;
    const double* const in[8] = {tk1, RHOCP, KX, KY, HR, HA, HS, DHP};
    for (const double* p : in)
        if (!p) return fail(c, EREB_ERR_ARG, "null input grid");
    if (!RT || !colptr || !rowval || !nzval)
        return fail(c, EREB_ERR_ARG, "null output buffer");
    const int64_t ncell = (int64_t)c->Nx1 * c->Ny1;
    if (ncell > NOALLOC_MAX_CELLS)
        return fail(c, EREB_ERR_ARG,
                    "grid of " + std::to_string(ncell) + " P cells exceeds the no-alloc "
                    "scratch maximum of " + std::to_string(NOALLOC_MAX_CELLS) + " cells");
    if (capacity < 5 * ncell)
        return fail(c, EREB_ERR_ARG, "capacity must be at least 5 * Ny1 * Nx1 entries");
    EREB_HIP(c, hipSetDevice(c->device));
    char* s = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        s = g_scratch[c->device];
    }
    if (!s) {
        int rc = noalloc_init(c);
        if (rc) return rc;
        std::lock_guard<std::mutex> lock(g_mu);
        s = g_scratch[c->device];
    }
    // dt and cleared flags first: the scratch keeps the previous call's
    const ThermalScratchHead h = thermal_scratch_head(s);
    hipLaunchKernelGGL(k_noalloc_dt, dim3(1), dim3(64), 0, c->stream, h.scal, dt);
    int rc = fill_i32(c, h.flags, 0, 4);
    if (rc) return rc;
    if ((rc = launch_thermal_noalloc(c, s, in, RT, colptr, rowval, nzval))) return rc;
    if (nnz) {
        int64_t last = 0;
        int32_t flags[4];
        EREB_HIP(c, hipMemcpyAsync(&last, colptr + ncell, sizeof last, hipMemcpyDeviceToHost,
                                   c->stream));
        EREB_HIP(c, hipMemcpyAsync(flags, h.flags, sizeof flags, hipMemcpyDeviceToHost,
                                   c->stream));
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        if (flags[1])
            return fail(c, EREB_ERR_STATE,
                        "assembly: an equation referenced a column outside its stencil slot "
                        "table");
        *nnz = last - 1;
    }
    return EREB_OK;
}
