// ereb_save.hip -- the device side of save_state (E:5718-5957): the arrays of
// the JLD2 record that the context holds, copied into caller arrays in one
// call, in the record's order.  The JLD2 write itself stays with the caller.
#include "ereb_internal.h"

namespace ereb {
namespace {

struct SaveField {
    const char* name;  // the record's key (E:5864-5957)
    int kind;          // EREB_SAVE_GRID, EREB_SAVE_YNY or EREB_SAVE_MARKER
    int field;         // ereb_grid_field or ereb_marker_field
};

// The keyword list of jldsave (E:5864-5957) without the scalars, the axes and
// the 17 arrays the library does not hold (EXY, SXY, qxD, qyD, EXX, SXX, vxp,
// vyp, vxpf, vypf, FI, ETA5, ETA00, YNY5, YNY00, EII, SII: outputs of the host
// solves and the plastic iterations).
constexpr int G = EREB_SAVE_GRID, Y = EREB_SAVE_YNY, M = EREB_SAVE_MARKER;
const SaveField SAVE_FIELDS[EREB_SAVE_NFIELDS] = {
    {"ETA", G, EREB_G_ETA}, {"ETA0", G, EREB_G_ETA0}, {"GGG", G, EREB_G_GGG},
    {"SXY0", G, EREB_G_SXY0}, {"wyx", G, EREB_G_WYX}, {"COH", G, EREB_G_COH},
    {"TEN", G, EREB_G_TEN}, {"FRI", G, EREB_G_FRI}, {"YNY", Y, 0},
    {"RHOX", G, EREB_G_RHOX}, {"RHOFX", G, EREB_G_RHOFX}, {"KX", G, EREB_G_KX},
    {"PHIX", G, EREB_G_PHIX}, {"vx", G, EREB_G_VX}, {"vxf", G, EREB_G_VXF},
    {"RX", G, EREB_G_RX}, {"gx", G, EREB_G_GX}, {"RHOY", G, EREB_G_RHOY},
    {"RHOFY", G, EREB_G_RHOFY}, {"KY", G, EREB_G_KY}, {"PHIY", G, EREB_G_PHIY},
    {"vy", G, EREB_G_VY}, {"vyf", G, EREB_G_VYF}, {"RY", G, EREB_G_RY}, {"gy", G, EREB_G_GY},
    {"RHO", G, EREB_G_RHO}, {"RHOCP", G, EREB_G_RHOCP}, {"ALPHA", G, EREB_G_ALPHA},
    {"ALPHAF", G, EREB_G_ALPHAF}, {"HR", G, EREB_G_HR}, {"HA", G, EREB_G_HA},
    {"HS", G, EREB_G_HS}, {"ETAP", G, EREB_G_ETAP}, {"GGGP", G, EREB_G_GGGP},
    {"SXX0", G, EREB_G_SXX0}, {"tk1", G, EREB_G_TK1}, {"tk2", G, EREB_G_TK2},
    {"pr", G, EREB_G_PR}, {"pf", G, EREB_G_PF}, {"ps", G, EREB_G_PS}, {"pr0", G, EREB_G_PR0},
    {"pf0", G, EREB_G_PF0}, {"ps0", G, EREB_G_PS0}, {"ETAPHI", G, EREB_G_ETAPHI},
    {"BETAPHI", G, EREB_G_BETAPHI}, {"PHI", G, EREB_G_PHI}, {"APHI", G, EREB_G_APHI},
    {"YNY_inv_ETA", G, EREB_G_YNY_INV_ETA}, {"DSXY", G, EREB_G_DSXY},
    {"DSXX", G, EREB_G_DSXX}, {"DMP", G, EREB_G_DMP}, {"DHP", G, EREB_G_DHP},
    {"XWS", G, EREB_G_XWS},
    {"XWsolidm0", M, EREB_M_XWSOLIDM0}, {"xm", M, EREB_M_XM}, {"ym", M, EREB_M_YM},
    {"tm", M, EREB_M_TM}, {"tkm", M, EREB_M_TKM}, {"sxxm", M, EREB_M_SXXM},
    {"sxym", M, EREB_M_SXYM}, {"etavpm", M, EREB_M_ETAVPM}, {"phim", M, EREB_M_PHIM},
    {"rhototalm", M, EREB_M_RHOTOTALM}, {"rhocptotalm", M, EREB_M_RHOCPTOTALM},
    {"etatotalm", M, EREB_M_ETATOTALM}, {"hrtotalm", M, EREB_M_HRTOTALM},
    {"ktotalm", M, EREB_M_KTOTALM}, {"tkm_rhocptotalm", M, EREB_M_TKM_RHOCPTOTALM},
    {"etafluidcur_inv_kphim", M, EREB_M_ETAFLUIDCUR_INV_KPHIM},
    {"inv_gggtotalm", M, EREB_M_INV_GGGTOTALM}, {"fricttotalm", M, EREB_M_FRICTTOTALM},
    {"cohestotalm", M, EREB_M_COHESTOTALM}, {"tenstotalm", M, EREB_M_TENSTOTALM},
    {"rhofluidcur", M, EREB_M_RHOFLUIDCUR}, {"alphasolidcur", M, EREB_M_ALPHASOLIDCUR},
    {"alphafluidcur", M, EREB_M_ALPHAFLUIDCUR},
};

__global__ void k_widen_i32(const int32_t* __restrict__ a, int64_t* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_widen_f32(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

}  // namespace
}  // namespace ereb

using namespace ereb;

extern "C" int ereb_save_field(int k, const char** name, int* kind, int* field) {
    if (k < 0 || k >= EREB_SAVE_NFIELDS) return EREB_ERR_ARG;
    if (name) *name = SAVE_FIELDS[k].name;
    if (kind) *kind = SAVE_FIELDS[k].kind;
    if (field) *field = SAVE_FIELDS[k].field;
    return EREB_OK;
}

// Every copy is queued on the context's stream; marker vectors stored in
// another width (tm as int32, the mixed form's float properties) are widened
// on the device into a staging buffer first, which the stream order lets the
// next widening reuse.  One synchronisation at the end.
extern "C" int ereb_save_state_fields(ereb_ctx* c, void* const* dst, int64_t marknum) {
    if (!c) return EREB_ERR_ARG;
    if (!dst) return fail(c, EREB_ERR_ARG, "ereb_save_state_fields: null destination table");
    if (marknum < 0 || marknum > c->cap)
        return fail(c, EREB_ERR_ARG, "ereb_save_state_fields: marknum outside [0, capacity]");
    int rc = check_flags(c);
    if (rc) return rc;
    bool widen = false;
    for (int k = 0; k < EREB_SAVE_NFIELDS; ++k) {
        const SaveField& f = SAVE_FIELDS[k];
        if (dst[k] && f.kind == M &&
            (f.field == EREB_M_TM || marker_is_f32(c->precision, f.field)))
            widen = true;
    }
    const size_t stage = (size_t)8 * (size_t)std::max<int64_t>(marknum, 1);
    if (widen && marknum > 0 && stage > c->pb_bytes) {
        EREB_HIP(c, hipStreamSynchronize(c->stream));
        hipFree(c->pb_pool);
        c->pb_pool = nullptr;
        c->pb_bytes = 0;
        EREB_HIP(c, hipMalloc(&c->pb_pool, stage));
        c->pb_bytes = stage;
    }
    const dim3 grid((unsigned)((marknum + 255) / 256));
    for (int k = 0; k < EREB_SAVE_NFIELDS; ++k) {
        const SaveField& f = SAVE_FIELDS[k];
        void* out = dst[k];
        if (!out) continue;
        if (f.kind == G) {
            EREB_HIP(c, hipMemcpyAsync(out, gptr(c, f.field), sizeof(double) * c->gnumel[f.field],
                                       hipMemcpyDeviceToHost, c->stream));
        } else if (f.kind == Y) {
            EREB_HIP(c, hipMemcpyAsync(out, c->d_yny, (size_t)c->Nx * c->Ny,
                                       hipMemcpyDeviceToHost, c->stream));
        } else if (marknum > 0) {
            const void* src = c->d_m[f.field];
            if (f.field == EREB_M_TM) {
                hipLaunchKernelGGL(k_widen_i32, grid, dim3(256), 0, c->stream,
                                   (const int32_t*)src, (int64_t*)c->pb_pool, marknum);
                src = c->pb_pool;
            } else if (marker_is_f32(c->precision, f.field)) {
                hipLaunchKernelGGL(k_widen_f32, grid, dim3(256), 0, c->stream,
                                   (const float*)src, (double*)c->pb_pool, marknum);
                src = c->pb_pool;
            }
            EREB_HIP(c, hipGetLastError());
            EREB_HIP(c, hipMemcpyAsync(out, src, sizeof(double) * marknum, hipMemcpyDeviceToHost,
                                       c->stream));
        }
    }
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    return EREB_OK;
}
