// ereb_hydration.hip -- the remaining marker<->grid transfers of a timestep:
//   update_marker_viscosity!          E:774-804   (gather, basic grid)
//   update_marker_porosity!           E:4822-4856 (gather, P grid)
//   update_p_nodes_melt_composition!  E:1929-1939 (scatter onto P nodes)
//   perform_thermochemical_reaction!  E:2558-2694 (per-marker reaction and
//                                                  scatter onto P nodes)
//
// The gathers run one thread per marker in storage order, like the
// temperature and stress gathers.  The two P-grid scatters reuse the fine-key
// sort of the main scatter, re-run first when positions changed since it was
// made (a fine cell lies in exactly one P cell): one
// 192-thread workgroup per tile, one thread per fine cell summing its
// markers' weighted contributions to the four corners of that P cell in
// sorted order, then one thread per (window node, sum) adding the fine-cell
// partials that touch the node in a fixed order.  The tile windows are
// summed per node by k_pcombine, which also divides by the weight
// (compute_molarfraction! E:2239, compute_thermodynamic_xfer! E:2208).
// No atomics: results are reproducible run to run, and agree with the
// reference's marker-order sums to rounding.  The reaction's per-marker arithmetic
// (an exp, a log and a dozen divisions per rock marker) runs first, one
// thread per marker in storage order, and leaves each marker's ΔM, ΔH and
// whether it contributes; the scatter then sums those in the fine-cell order
// (one thread per fine cell would otherwise run its markers' arithmetic one
// after another at a fraction of the occupancy).
#include "ereb_internal.h"

namespace ereb {

// Julia min(a, b) for Float64: NaN propagates
__device__ __forceinline__ double jl_min(double a, double b) {
    if (a != a || b != b) return a + b;
    return a < b ? a : (b < a ? b : (signbit(a) ? a : b));
}

// etatotal_rocks E:1312-1318
__device__ __forceinline__ double etatotal_rocks(const ereb_materials& M, double T, int t) {
    const double es = T > M.tmsolidphase ? sel3(M.etasolidmm, t) : sel3(M.etasolidm, t);
    const double ef = T > M.tmfluidphase ? sel3(M.etafluidmm, t) : sel3(M.etafluidm, t);
    return jl_max(jl_max(M.etamin, es), ef);
}

// ---- update_marker_viscosity! (E:774-804) ---------------------------------
// R: storage type of etatotalm and etavpm (float in the mixed form; the
// arithmetic stays Float64 and reads back the stored etatotalm)
template <class R>
__global__ __launch_bounds__(256) void k_marker_viscosity(
    GeomK g, const ereb_materials* __restrict__ M, const double* __restrict__ xm,
    const double* __restrict__ ym, const int32_t* __restrict__ tm,
    const double* __restrict__ tkm, R* __restrict__ etatotalm, R* __restrict__ etavpm,
    const uint8_t* __restrict__ YNY, const double* __restrict__ YIE, int64_t n) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double x = xm[m], y = ym[m];
    const int j = jl_cell((x - g.x[0]) * g.idx, 1, g.Nx - 1);
    const int i = jl_cell((y - g.y[0]) * g.idy, 1, g.Ny - 1);
    double w[4];
    weights4((x - g.x[j - 1]) / g.dx, (y - g.y[i - 1]) / g.dy, w);
    const int t = tm[m];
    double eta;
    if (t < 3) {
        const R e = (R)etatotal_rocks(*M, tkm[m], t);
        etatotalm[m] = e;
        eta = (double)e;
    } else {
        eta = (double)etatotalm[m];
    }
    const uint8_t* yn = YNY + (size_t)(j - 1) * g.Ny + (i - 1);
    double v = eta;
    if (yn[0] | yn[1] | yn[g.Ny] | yn[g.Ny + 1]) {
        v = 1.0 / gdot4(YIE, g.Ny, i, j, w);
        v = v > eta ? eta : v;
    }
    etavpm[m] = (R)v;
}

// ---- update_marker_porosity! (E:4822-4856) --------------------------------
__global__ __launch_bounds__(256) void k_marker_porosity(
    GeomK g, double phimin, double phimax, const double* __restrict__ xm,
    const double* __restrict__ ym, const int32_t* __restrict__ tm, double* __restrict__ phim,
    const double* __restrict__ APHI, double dt, int64_t n) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n || !(tm[m] < 3)) return;
    const double x = xm[m], y = ym[m];
    const int j = jl_cell((x - g.xp[0]) * g.idx, 1, g.Nx);
    const int i = jl_cell((y - g.yp[0]) * g.idy, 1, g.Ny);
    double w[4];
    weights4((x - g.xp[j - 1]) / g.dx, (y - g.yp[i - 1]) / g.dy, w);
    const double aphim = gdot4(APHI, g.Ny1, i, j, w);
    const double p = phim[m];
    const double v = p / ((1.0 - p) * exp(aphim * dt) + p);
    phim[m] = jl_max(phimin, jl_min(phimax, v));
}

// ---- P-grid scatters ------------------------------------------------------
constexpr int PS_THREADS = 192;  // 144 fine cells; 64 nodes x up to 3 sums

struct PScatterBase {
    GeomK g;
    int nty;
    const uint32_t* __restrict__ perm;
    const int32_t* __restrict__ fstart;
    const double* __restrict__ xm;
    const double* __restrict__ ym;
    double* __restrict__ part;  // ntiles * S * 64
};

// molarfraction_marker_to_p_nodes! (E:1883-1908): XWsolidm0 and weight 1
template <class R>
struct MeltOp {
    static constexpr int S = 2;  // XWSSUM, WTPSUM
    static constexpr int GROUP = 4;  // markers of a fine cell loaded together
    static constexpr bool AHEAD = true;  // next group's loads before this group's sums
    const R* __restrict__ XW0;
    struct In {
        double xw;
    };
    __device__ __forceinline__ In load(uint32_t m) const { return In{(double)XW0[m]}; }
    __device__ __forceinline__ bool operator()(uint32_t, const In& in, const GeomK&,
                                               const double*, int, int, double* v) const {
        v[0] = in.xw;
        v[1] = 1.0;
        return true;
    }
};

// compute_Δtreaction E:2382-2397 (mode validated on the host; mode 3 divides
// by Sxo_B as the reference does)
__device__ __forceinline__ double dtreaction(const ereb_hydration& H, double T, double phi) {
    if (H.rate_mode == 1) {
        const double d = T - H.c_I;
        return -H.log_completion_rate / (H.A_I * phi) * exp(H.b_I * (d * d));
    }
    if (H.rate_mode == 2)
        return -H.log_completion_rate / (H.Sxo_B * phi) * pow(2.0, (H.To_B - T) / H.Tscl_B);
    if (H.rate_mode == 3)
        return -H.log_completion_rate / (H.Sxo_B * phi) *
               exp(H.Ea_T / H.RG * (1.0 / T - 1.0 / H.To_T));
    return H.dtreaction;
}

// the rock-marker body of perform_thermochemical_reaction! (E:2587-2686):
// writes pfm0 for every rock marker and XWsm, phinewm (and at timestep 1
// XWsm0, phim) for the fluid-bearing ones, which contribute ΔM, ΔH and 1
template <class R>
struct ReactionOp {
    const ereb_materials* __restrict__ M;
    ereb_hydration H;
    const int32_t* __restrict__ tm;
    const double* __restrict__ tk2;
    const double* __restrict__ pf;
    R* __restrict__ XWsm0;
    double* __restrict__ XWsm;
    double* __restrict__ phim;
    double* __restrict__ phinewm;
    double* __restrict__ pfm0;
    double dt;
    int timestep, titer;

    struct In {
        int t;
        double pf0, XW0, ph;
    };
    // every marker's inputs (rock or not): the loads are issued together
    __device__ __forceinline__ In load(uint32_t m) const {
        return In{tm[m], pfm0[m], (double)XWsm0[m], phim[m]};
    }
    __device__ __forceinline__ bool operator()(uint32_t m, const In& in, const GeomK& g,
                                               const double* w, int ip, int jp,
                                               double* v) const {
        if (!(in.t < 3)) return false;
        const ereb_materials& Mt = *M;
        const double tknm = gdot4(tk2, g.Ny1, ip, jp, w);
        double pfnm = jl_max(0.0, gdot4(pf, g.Ny1, ip, jp, w));
        if (titer > 2) pfnm = pfnm * (1.0 - H.pfcoeff) + in.pf0 * H.pfcoeff;
        pfm0[m] = pfnm;
        const double XW0 = in.XW0, ph = in.ph;
        const double XDs0 = 1.0 - XW0;
        const bool liquid = tknm > Mt.tmfluidphase;
        const double VH2O = liquid ? H.VH2Of : H.VH2Ofi;
        const double Xf0 = ph * (XW0 * Mt.VWs + XDs0 * Mt.VDs) /
                           ((1.0 - ph) * VH2O + ph * (XW0 * Mt.VWs + XDs0 * Mt.VDs));
        const double Xs0 = 1.0 - Xf0;
        const double XH2Ot = (XW0 * Xs0 + Xf0) / (1.0 + XW0 * Xs0);
        const double XDt = 1.0 - XH2Ot;
        const double rhos0 = (Mt.MD + Mt.MH2O * XW0) / (Mt.VDs * XDs0 + Mt.VWs * XW0);
        const double rhof = liquid ? Mt.rhoH2Of : Mt.rhoH2Ofi;
        const double dtr = dtreaction(H, tknm, ph);
        const double MDW = Mt.MD + Mt.MH2O;
        const double Ht0 = -Xs0 * XW0 * H.dHWD / MDW;  // compute_relative_enthalpy E:2502
        // compute_gibbs_free_energy E:2470-2482
        const double dG0 = dt < dtr ? (H.dHWD - tknm * H.dSWD + pfnm * H.dVWD +
                                       H.RG * tknm * log(XDs0 / XW0)) * (1.0 - dt / dtr)
                                    : 0.0;
        // compute_reaction_constant E:2521-2524
        const double KWD = exp(-(H.dHWD - tknm * H.dSWD + H.dVWD * pfnm - dG0) / (H.RG * tknm));
        const double XWs1 = 1.0 / (KWD + 1.0);
        const double XDs1 = 1.0 - XWs1;
        const double Xs1 = XDt / (1.0 - XDt * XWs1);
        const double Xf1 = 1.0 - Xs1;
        if (!(0.0 < Xf1 && Xf1 < 1.0)) return false;
        const double phi1 = Xf1 * VH2O / (Xf1 * VH2O + Xs1 * (XWs1 * Mt.VWs + XDs1 * Mt.VDs));
        const double rhos1 = (Mt.MD + Mt.MH2O * XWs1) / (Mt.VDs * XDs1 + Mt.VWs * XWs1);
        const double Ht1 = -Xs1 * XWs1 * H.dHWD / MDW;
        const double dHt = Ht1 - Ht0;
        const double RV = (rhos1 * (1.0 - phi1) + rhof * phi1) / (rhos0 * (1.0 - ph) + rhof * ph);
        const double Gmass = (rhos0 * RV * (1.0 - ph) - rhos1 * (1.0 - phi1)) / dt;
        XWsm[m] = XWs1;
        phinewm[m] = phi1;
        if (timestep == 1) {
            XWsm0[m] = (R)XWs1;
            phim[m] = phi1;
        }
        v[0] = (1.0 - RV) / dt;  // ΔMm (16.112e)
        v[1] = Gmass * dHt;      // ΔHm (16.113)
        v[2] = 1.0;
        return true;
    }
};

constexpr int PS_PERM = 4096;  // tile sort-order entries staged in LDS (16 KB)

// the reaction's per-marker pass (E:2587-2686) in storage order: the marker's
// P cell and weights are those that the fine cell of the scatter gives it
// (the fine key is derived from the same cells), so the values are the ones
// the scatter would compute
template <class R>
__global__ __launch_bounds__(256) void k_reaction_markers(ReactionOp<R> op, GeomK g,
                                                          const double* __restrict__ xm,
                                                          const double* __restrict__ ym,
                                                          int64_t M, double2* __restrict__ dMH,
                                                          uint8_t* __restrict__ ok) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const typename ReactionOp<R>::In in = op.load((uint32_t)m);
    const double x = xm[m], y = ym[m];
    const Cells cl = marker_cells(g, x, y);
    double w[4], v[3];
    weights4((x - g.xp[cl.jp - 1]) / g.dx, (y - g.yp[cl.ip - 1]) / g.dy, w);
    const bool r = op((uint32_t)m, in, g, w, cl.ip, cl.jp, v);
    ok[m] = r ? 1 : 0;
    if (r) dMH[m] = make_double2(v[0], v[1]);
}

// the reaction's scatter: ΔM, ΔH (one 16-byte load) and weight 1 of the
// contributing markers
struct ReactionSumOp {
    static constexpr int S = 3;  // DMPSUM, DHPSUM, WTPSUM
    static constexpr int GROUP = 4;
    static constexpr bool AHEAD = true;
    const double2* __restrict__ dMH;
    const uint8_t* __restrict__ ok;
    struct In {
        double dm, dh;
        bool ok;
    };
    // ΔM, ΔH of a marker that does not contribute hold stale values (read, not used)
    __device__ __forceinline__ In load(uint32_t m) const {
        const double2 p = dMH[m];
        return In{p.x, p.y, ok[m] != 0};
    }
    __device__ __forceinline__ bool operator()(uint32_t, const In& in, const GeomK&,
                                               const double*, int, int, double* v) const {
        if (!in.ok) return false;
        v[0] = in.dm;
        v[1] = in.dh;
        v[2] = 1.0;
        return true;
    }
};

template <class Op>
__global__ __launch_bounds__(PS_THREADS) void k_pscatter(PScatterBase B, Op op) {
    constexpr int S = Op::S;
    __shared__ double fl[NFINE * 4 * S];  // fine-cell partials [f][corner][s]
    __shared__ uint32_t sperm[PS_PERM];   // the tile's part of the sort order
    const int tile = (int)xcd_block(), t = threadIdx.x;  // XCD-contiguous, as k_scatter
    const int j0 = (tile / B.nty) * TILE, i0 = (tile % B.nty) * TILE;
    // the tile's markers are one contiguous run of the sort order: staged in
    // LDS with coalesced loads when it fits (read from memory otherwise)
    const int ks = B.fstart[(size_t)tile * NFINE], ke = B.fstart[(size_t)(tile + 1) * NFINE];
    const bool staged = ke - ks <= PS_PERM;
    if (staged)
        for (int k = ks + t; k < ke; k += PS_THREADS) sperm[k - ks] = B.perm[k];
    __syncthreads();
    if (t < NFINE) {
        // fine cell t = fj * FINE + fi lies in P cell (ip, jp) (1-based): the
        // basic cell refined by its P offset (the fine key's definition)
        const int fi = t % FINE, fj = t / FINE;
        const int ip = i0 + (fi >> 1) + (fi & 1) + 1, jp = j0 + (fj >> 1) + (fj & 1) + 1;
        double acc[4][S];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int s = 0; s < S; ++s) acc[c][s] = 0.0;
        const int k0 = B.fstart[(size_t)tile * NFINE + t];
        const int k1 = B.fstart[(size_t)tile * NFINE + t + 1];
        if (k0 < k1) {
            // markers exist only in fine cells whose P cell is on the grid
            const double xa = B.g.xp[jp - 1], ya = B.g.yp[ip - 1];
            // groups of Op::GROUP markers: a group's loads are issued together
            // (with Op::AHEAD, before the previous group is summed); the
            // markers are summed one after the other in sorted order
            constexpr int G = Op::GROUP;
            struct Group {
                uint32_t m[G];
                double x[G], y[G];
                typename Op::In in[G];
            };
            auto fetch = [&](int k, Group& q) {
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    const int kk = min(k + u, k1 - 1);
                    q.m[u] = staged ? sperm[kk - ks] : B.perm[kk];
                }
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    q.x[u] = B.xm[q.m[u]];
                    q.y[u] = B.ym[q.m[u]];
                    q.in[u] = op.load(q.m[u]);
                }
            };
            Group cur, nxt;
            if (Op::AHEAD) fetch(k0, cur);
            for (int k = k0; k < k1; k += G) {
                if (!Op::AHEAD) fetch(k, cur);
                else if (k + G < k1) fetch(k + G, nxt);
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    if (k + u >= k1) break;
                    double w[4];
                    weights4((cur.x[u] - xa) / B.g.dx, (cur.y[u] - ya) / B.g.dy, w);
                    double v[S];
                    if (op(cur.m[u], cur.in[u], B.g, w, ip, jp, v)) {
#pragma unroll
                        for (int c = 0; c < 4; ++c)
#pragma unroll
                            for (int s = 0; s < S; ++s) acc[c][s] += v[s] * w[c];
                    }
                }
                if (Op::AHEAD) cur = nxt;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int s = 0; s < S; ++s) fl[(t * 4 + c) * S + s] = acc[c][s];
    }
    __syncthreads();
    if (t < 64 * S) {
        // window node (li, lj) = corner (ci, cj) of P cell (li - ci, lj - cj);
        // P cell p holds the fine rows/columns 2p - 1 and 2p
        const int s = t >> 6, nd = t & 63, li = nd & 7, lj = nd >> 3;
        double sum = 0.0;
        for (int cj = 0; cj < 2; ++cj) {
            const int pj = lj - cj;
            if (pj < 0 || pj > TILE) continue;
            for (int ci = 0; ci < 2; ++ci) {
                const int pi = li - ci;
                if (pi < 0 || pi > TILE) continue;
                const int c = ci + 2 * cj;
                for (int fj = max(2 * pj - 1, 0); fj <= min(2 * pj, FINE - 1); ++fj)
                    for (int fi = max(2 * pi - 1, 0); fi <= min(2 * pi, FINE - 1); ++fi)
                        sum += fl[((fj * FINE + fi) * 4 + c) * S + s];
            }
        }
        B.part[((size_t)tile * S + s) * 64 + nd] = sum;
    }
}

// Sum of the tile windows covering each P node (tiles in ascending order),
// then the normalisation: melt (S = 2) XWS = XWSSUM * inv(WTPSUM);
// reaction (S = 3) DMP, DHP = DMPSUM, DHPSUM * inv(WTPSUM); 0 where the
// weight is not positive.
template <int S>
__global__ __launch_bounds__(256) void k_pcombine(const double* __restrict__ part, int nty,
                                                  int ntx, int Ny1, int Nx1, double* sum0,
                                                  double* sum1, double* sum2, double* out0,
                                                  double* out1) {
    // node counts stay far below 2^31: 32-bit index arithmetic
    const int n = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (n >= Ny1 * Nx1) return;
    const int i0 = n % Ny1, j0 = n / Ny1;
    const int tylo = max(0, (i0 - (WIN - 1) + TILE - 1) / TILE), tyhi = min(nty - 1, i0 / TILE);
    const int txlo = max(0, (j0 - (WIN - 1) + TILE - 1) / TILE), txhi = min(ntx - 1, j0 / TILE);
    double a[S];
#pragma unroll
    for (int s = 0; s < S; ++s) a[s] = 0.0;
    for (int tx = txlo; tx <= txhi; ++tx)
        for (int ty = tylo; ty <= tyhi; ++ty) {
            const int lane = (j0 - tx * TILE) * 8 + (i0 - ty * TILE);
#pragma unroll
            for (int s = 0; s < S; ++s)
                a[s] += part[((size_t)(tx * nty + ty) * S + s) * 64 + lane];
        }
    const double wt = a[S - 1];
    sum0[n] = a[0];
    sum1[n] = a[1];
    if (S == 3) sum2[n] = a[2];
    if (wt > 0.0) {
        const double iw = 1.0 / wt;
        out0[n] = a[0] * iw;
        if (S == 3) out1[n] = a[1] * iw;
    } else {
        out0[n] = 0.0;
        if (S == 3) out1[n] = 0.0;
    }
}

static PScatterBase pscatter_base(ereb_ctx* c) {
    PScatterBase B;
    B.g = geomk(c);
    B.nty = c->nty;
    B.perm = c->d_perm;
    B.fstart = c->d_fstart;
    B.xm = mptr(c, EREB_M_XM);
    B.ym = mptr(c, EREB_M_YM);
    B.part = c->d_part;  // consumed by k_pcombine on the same stream
    return B;
}

static_assert(3 <= NSUM, "the P-scatter partials fit the main scatter's buffer");

template <class Op>
static int run_pscatter(ereb_ctx* c, const Op& op, double* s0, double* s1, double* s2,
                        double* o0, double* o1) {
    const PScatterBase B = pscatter_base(c);
    if (c->ntiles > 0)
        hipLaunchKernelGGL(k_pscatter<Op>, dim3((unsigned)c->ntiles), dim3(PS_THREADS), 0,
                           c->stream, B, op);
    const int64_t nn = (int64_t)c->Ny1 * c->Nx1;
    hipLaunchKernelGGL(k_pcombine<Op::S>, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0,
                       c->stream, (const double*)c->d_part, c->nty, c->ntx, c->Ny1, c->Nx1, s0,
                       s1, s2, o0, o1);
    EREB_HIP(c, hipGetLastError());
    return EREB_OK;
}

// ---- launchers --------------------------------------------------------------

int launch_marker_viscosity(ereb_ctx* c) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    const dim3 grid((unsigned)((M + 255) / 256));
    hipEvent_t ev;
    prof_begin(c, EREB_K_VISCOSITY, &ev);
    const GeomK g = geomk(c);
    if (marker_is_f32(c->precision, EREB_M_ETATOTALM))
        hipLaunchKernelGGL(k_marker_viscosity<float>, grid, dim3(256), 0, c->stream, g, c->d_mat,
                           mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), tm_ptr(c),
                           mptr(c, EREB_M_TKM), fptr(c, EREB_M_ETATOTALM),
                           fptr(c, EREB_M_ETAVPM), c->d_yny, gptr(c, EREB_G_YNY_INV_ETA), M);
    else
        hipLaunchKernelGGL(k_marker_viscosity<double>, grid, dim3(256), 0, c->stream, g,
                           c->d_mat, mptr(c, EREB_M_XM), mptr(c, EREB_M_YM), tm_ptr(c),
                           mptr(c, EREB_M_TKM), mptr(c, EREB_M_ETATOTALM),
                           mptr(c, EREB_M_ETAVPM), c->d_yny, gptr(c, EREB_G_YNY_INV_ETA), M);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_VISCOSITY, ev);
    return EREB_OK;
}

int launch_marker_porosity(ereb_ctx* c, const ereb_hydration& H, double dt) {
    const int64_t M = c->marknum;
    if (M <= 0) return EREB_OK;
    hipEvent_t ev;
    prof_begin(c, EREB_K_POROSITY, &ev);
    hipLaunchKernelGGL(k_marker_porosity, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                       c->stream, geomk(c), H.phimin, H.phimax, mptr(c, EREB_M_XM),
                       mptr(c, EREB_M_YM), tm_ptr(c), mptr(c, EREB_M_PHIM),
                       gptr(c, EREB_G_APHI), dt, M);
    EREB_HIP(c, hipGetLastError());
    prof_end(c, EREB_K_POROSITY, ev);
    return EREB_OK;
}

int launch_melt_composition(ereb_ctx* c) {
    hipEvent_t ev;
    int rc = ensure_sort(c);
    if (rc) return rc;
    prof_begin(c, EREB_K_MELT, &ev);
    double *s0 = gptr(c, EREB_G_XWSSUM), *s1 = gptr(c, EREB_G_WTPSUM),
           *o0 = gptr(c, EREB_G_XWS);
    if (marker_is_f32(c->precision, EREB_M_XWSOLIDM0))
        rc = run_pscatter(c, MeltOp<float>{fptr(c, EREB_M_XWSOLIDM0)}, s0, s1, nullptr, o0,
                          nullptr);
    else
        rc = run_pscatter(c, MeltOp<double>{mptr(c, EREB_M_XWSOLIDM0)}, s0, s1, nullptr, o0,
                          nullptr);
    if (rc) return rc;
    prof_end(c, EREB_K_MELT, ev);
    return EREB_OK;
}

template <class R>
static ReactionOp<R> reaction_op(ereb_ctx* c, const ereb_hydration& H, R* XWsm0, double dt,
                                 int64_t timestep, int64_t titer) {
    ReactionOp<R> op;
    op.M = c->d_mat;
    op.H = H;
    op.tm = tm_ptr(c);
    op.tk2 = gptr(c, EREB_G_TK2);
    op.pf = gptr(c, EREB_G_PF);
    op.XWsm0 = XWsm0;
    op.XWsm = mptr(c, EREB_M_XWSOLIDM);
    op.phim = mptr(c, EREB_M_PHIM);
    op.phinewm = mptr(c, EREB_M_PHINEWM);
    op.pfm0 = mptr(c, EREB_M_PFM0);
    op.dt = dt;
    // only the comparisons timestep == 1 and titer > 2 matter
    op.timestep = timestep == 1 ? 1 : 0;
    op.titer = titer > 2 ? 3 : 0;
    return op;
}

// (ΔM, ΔH) pairs (double2) and the contribute flag (byte) of every marker,
// for the context's marker capacity (reallocated when it grows)
static int ensure_react(ereb_ctx* c) {
    if (c->react_cap >= c->cap) return EREB_OK;
    EREB_HIP(c, hipStreamSynchronize(c->stream));
    hipFree(c->d_react);
    c->d_react = nullptr;
    c->react_cap = 0;
    EREB_HIP(c, hipMalloc(&c->d_react, (size_t)c->cap * (2 * sizeof(double) + 1)));
    c->react_cap = c->cap;
    return EREB_OK;
}

template <class R>
static void launch_reaction_markers(ereb_ctx* c, const ReactionOp<R>& op, double2* dMH,
                                    uint8_t* ok) {
    const int64_t M = c->marknum;
    if (M <= 0) return;
    hipLaunchKernelGGL(k_reaction_markers<R>, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                       c->stream, op, geomk(c), (const double*)mptr(c, EREB_M_XM),
                       (const double*)mptr(c, EREB_M_YM), M, dMH, ok);
}

int launch_thermochemical(ereb_ctx* c, const ereb_hydration& H, double dt, int64_t timestep,
                          int64_t titer) {
    hipEvent_t ev;
    int rc = ensure_sort(c);
    if (rc) return rc;
    if ((rc = ensure_react(c))) return rc;
    prof_begin(c, EREB_K_REACTION, &ev);
    double *s0 = gptr(c, EREB_G_DMPSUM), *s1 = gptr(c, EREB_G_DHPSUM),
           *s2 = gptr(c, EREB_G_WTPSUM), *o0 = gptr(c, EREB_G_DMP), *o1 = gptr(c, EREB_G_DHP);
    double2* dMH = (double2*)c->d_react;
    uint8_t* ok = (uint8_t*)(dMH + c->react_cap);
    if (marker_is_f32(c->precision, EREB_M_XWSOLIDM0))
        launch_reaction_markers(c, reaction_op(c, H, fptr(c, EREB_M_XWSOLIDM0), dt, timestep,
                                               titer), dMH, ok);
    else
        launch_reaction_markers(c, reaction_op(c, H, mptr(c, EREB_M_XWSOLIDM0), dt, timestep,
                                               titer), dMH, ok);
    ReactionSumOp sum;
    sum.dMH = dMH;
    sum.ok = ok;
    rc = run_pscatter(c, sum, s0, s1, s2, o0, o1);
    if (rc) return rc;
    prof_end(c, EREB_K_REACTION, ev);
    return EREB_OK;
}

}  // namespace ereb
