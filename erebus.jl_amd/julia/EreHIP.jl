# EreHIP.jl -- Julia binding of libereb_hip.so for Erebus.jl: ccall wrappers
# with the reference functions' argument lists, and ChainRulesCore rrules of
# the differentiable transfer functions backed by the library's pullback
# kernels.  Include it next to Erebus.jl's sources (it reads the reference's
# constants: Nxm, Nym, dxm, dym, xxm, yym, Nxmc, Nymc, mdis_init,
# random_markers and the hydration constants of src/constants.jl).
# Citations: E: = src/Erebus.jl, C: = src/constants.jl.

module EreHIP

using ChainRulesCore
using SparseArrays

const LIB = get(ENV, "EREB_HIP_LIB", "libereb_hip.so")

struct Geometry                       # ereb_geometry
    Nx::Int64; Ny::Int64; dx::Float64; dy::Float64
    x::Ptr{Float64}; y::Ptr{Float64}; xvx::Ptr{Float64}; yvx::Ptr{Float64}
    xvy::Ptr{Float64}; yvy::Ptr{Float64}; xp::Ptr{Float64}; yp::Ptr{Float64}
end
# Materials mirrors ereb_materials field for field (tables as NTuple{3,Float64},
# then the scalars, brk4::NTuple{4,Float64}, crk4::NTuple{3,Float64}).

const EREB_FP64 = Cint(0)

mutable struct Ctx
    h::Ptr{Cvoid}
    pinned::Vector{Ptr{Cvoid}}    # page-locked hand-off arrays (section 2)
end

function check(ctx, rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:ereb_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx.h))
    rc == 2 && throw("unknown mode")          # EREB_ERR_MODE: the reference's throw (E:2289)
    error("ereb: $msg (code $rc)")
end

function Ctx(x, y, xvx, yvx, xvy, yvy, xp, yp, dx, dy, mat::Ref)
    GC.@preserve x y xvx yvx xvy yvy xp yp begin
        g = Ref(Geometry(length(x), length(y), dx, dy, pointer(x), pointer(y),
                         pointer(xvx), pointer(yvx), pointer(xvy), pointer(yvy),
                         pointer(xp), pointer(yp)))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:ereb_ctx_create, LIB), Cint,
                   (Ref{Geometry}, Ptr{Cvoid}, Cint, Cint, Ref{Ptr{Cvoid}}),
                   g, mat, EREB_FP64, 0, h)
        ctx = Ctx(h[], Ptr{Cvoid}[])
        check(ctx, rc)
        finalizer(ctx) do c
            for p in c.pinned
                ccall((:ereb_host_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), c.h, p)
            end
            ccall((:ereb_ctx_destroy, LIB), Cint, (Ptr{Cvoid},), c.h)
        end
        return ctx
    end
end

# update_marker_stress!(xm, ym, sxxm, sxym, DSXX, DSXY, marknum)   E:4354
function update_marker_stress!(ctx::Ctx, xm, ym, sxxm, sxym, DSXX, DSXY, marknum)
    GC.@preserve xm ym sxxm sxym DSXX DSXY check(ctx, ccall(
        (:ereb_update_marker_stress, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Int64),
        ctx.h, xm, ym, sxxm, sxym, DSXX, DSXY, marknum))
end

# update_marker_temperature!(xm, ym, tkm, DT, tk2, timestep, marknum)   E:4782
function update_marker_temperature!(ctx::Ctx, xm, ym, tkm, DT, tk2, timestep, marknum)
    GC.@preserve xm ym tkm DT tk2 check(ctx, ccall(
        (:ereb_update_marker_temperature, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Int64, Int64),
        ctx.h, xm, ym, tkm, DT, tk2, timestep, marknum))
end

# move_markers_rk4!(xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf, vyf, wyx,
#                   tk2, marknum, dt, mode)                            E:4990
function move_markers_rk4!(ctx::Ctx, xm, ym, tm::Vector{Int64}, tkm, phim, sxym,
                           sxxm, vx, vy, vxf, vyf, wyx, tk2, marknum, dt, mode)
    GC.@preserve xm ym tm tkm phim sxym sxxm vx vy vxf vyf wyx tk2 check(ctx, ccall(
        (:ereb_move_markers_rk4, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64,
         Int64),
        ctx.h, xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf, vyf, wyx, tk2,
        marknum, dt, mode))
end

# assemble_thermal_lse!(tk1, RHOCP, KX, KY, HR, HA, HS, DHP, RT, dt)    E:4526
# returns the matrix as SparseMatrixCSC (the reference returns the
# ExtendableSparseMatrix wrapping the same CSC)
function assemble_thermal_lse!(ctx::Ctx, tk1, RHOCP, KX, KY, HR, HA, HS, DHP, RT, dt)
    nnz = Ref{Int64}(0)
    GC.@preserve tk1 RHOCP KX KY HR HA HS DHP RT check(ctx, ccall(
        (:ereb_assemble_thermal_lse, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Float64, Ref{Int64}),
        ctx.h, tk1, RHOCP, KX, KY, HR, HA, HS, DHP, RT, dt, nnz))
    return csc_get(ctx, 1, length(RT), nnz[])
end

function csc_get(ctx::Ctx, which, n, nnz)
    colptr = Vector{Int64}(undef, n + 1)
    rowval = Vector{Int64}(undef, nnz)
    nzval = Vector{Float64}(undef, nnz)
    check(ctx, ccall((:ereb_csc_get, LIB), Cint,
                     (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, which, colptr, rowval, nzval, C_NULL))
    return SparseMatrixCSC(n, n, colptr, rowval, nzval)   # 1-based, rows sorted
end

# save_state (E:5718-5957): the 76 arrays of the JLD2 record that the device
# holds, fetched in one call, as a Dict keyed by the record's names; the
# caller merges its host-side entries (scalars, axes, the solver outputs)
# and writes the record with jldsave(fid; pairs...).
function save_state_fields(ctx::Ctx, Nx, Ny, marknum)
    n = 76
    name = Ref{Cstring}(); kind = Ref{Cint}(0); field = Ref{Cint}(0)
    out = Dict{Symbol,Any}()
    ptrs = Vector{Ptr{Cvoid}}(undef, n)
    for k in 0:n-1
        check(ctx, ccall((:ereb_save_field, LIB), Cint, (Cint, Ref{Cstring}, Ref{Cint}, Ref{Cint}),
                         k, name, kind, field))
        a = if kind[] == 0         # EREB_SAVE_GRID
            m = ccall((:ereb_grid_numel, LIB), Int64, (Ptr{Cvoid}, Cint), ctx.h, field[])
            m == Int64(Ny) * Nx ? Matrix{Float64}(undef, Ny, Nx) :
                                  Matrix{Float64}(undef, Ny + 1, Nx + 1)
        elseif kind[] == 1         # EREB_SAVE_YNY: one byte per node, 0/1
            Matrix{UInt8}(undef, Ny, Nx)
        else                       # EREB_SAVE_MARKER
            field[] == 2 ? Vector{Int64}(undef, marknum) : Vector{Float64}(undef, marknum)
        end
        out[Symbol(unsafe_string(name[]))] = a
        ptrs[k+1] = pointer(a)
    end
    GC.@preserve out check(ctx, ccall((:ereb_save_state_fields, LIB), Cint,
                                      (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Int64), ctx.h, ptrs, marknum))
    out[:YNY] = out[:YNY] .!= 0                   # Matrix{Bool}, as the reference stores it
    return out
end

# Solver hand-off (E:6719-6749): the reference sets hydromech_cache.A = L
# at every plastic iteration, and UMFPACK keeps its symbolic analysis while
# colptr/rowval are unchanged.  While the library reports an unchanged
# pattern only nzval and R are fetched, into the arrays of the previous L.
# The arrays are page-locked (ereb_host_alloc), so the copies run at the
# link's rate; the context's finalizer releases them (ereb_host_free) before
# ereb_ctx_destroy, so they must not outlive the context.
function pinned(ctx::Ctx, ::Type{T}, n) where {T}
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:ereb_host_alloc, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Ptr{Cvoid}}),
                     ctx.h, n * sizeof(T), p))
    push!(ctx.pinned, p[])                      # freed by the context's finalizer
    return unsafe_wrap(Array, Ptr{T}(p[]), n; own = false)
end

function csc_refresh!(ctx::Ctx, which, L::Union{Nothing,SparseMatrixCSC{Float64,Int64}}, R)
    changed = Ref{Cint}(1)
    check(ctx, ccall((:ereb_csc_pattern_changed, LIB), Cint, (Ptr{Cvoid}, Cint, Ref{Cint}),
                     ctx.h, which, changed))
    if L === nothing || changed[] != 0
        n, nnz = Ref{Int64}(0), Ref{Int64}(0)
        check(ctx, ccall((:ereb_csc_size, LIB), Cint,
                         (Ptr{Cvoid}, Cint, Ref{Int64}, Ref{Int64}), ctx.h, which, n, nnz))
        colptr = pinned(ctx, Int64, n[] + 1)
        rowval = pinned(ctx, Int64, nnz[])
        nzval = pinned(ctx, Float64, nnz[])
        GC.@preserve R check(ctx, ccall((:ereb_csc_get, LIB), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
            ctx.h, which, colptr, rowval, nzval, R))
        return SparseMatrixCSC(n[], n[], colptr, rowval, nzval)
    end
    GC.@preserve L R check(ctx, ccall((:ereb_csc_get, LIB), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
        ctx.h, which, C_NULL, C_NULL, L.nzval, R))
    return L
end
# at E:6719-6749, after the GPU assembly:
#     L = csc_refresh!(ctx, 0, L, R); hydromech_cache.A = L; hydromech_cache.b = R

# backtrace_pressures_rk4!(pr, pr0, ps, ps0, pf, pf0, vx, vy, vxf, vyf, dt)  E:5284
function backtrace_pressures_rk4!(ctx::Ctx, pr, pr0, ps, ps0, pf, pf0, vx, vy, vxf, vyf, dt)
    GC.@preserve pr pr0 ps ps0 pf pf0 vx vy vxf vyf check(ctx, ccall(
        (:ereb_backtrace_pressures_rk4, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Float64),
        ctx.h, pr, pr0, ps, ps0, pf, pf0, vx, vy, vxf, vyf, dt))
end

# The compaction and (de)hydration constants (C:227-279), once per run
struct Hydration                          # ereb_hydration
    phimin::Float64; phimax::Float64; RG::Float64; VH2Of::Float64; VH2Ofi::Float64
    dHWD::Float64; dSWD::Float64; dVWD::Float64; pfcoeff::Float64; dtreaction::Float64
    log_completion_rate::Float64; A_I::Float64; b_I::Float64; c_I::Float64
    Sxo_B::Float64; Tscl_B::Float64; To_B::Float64; Sxo_T::Float64; To_T::Float64
    Ea_T::Float64; rate_mode::Int64
end
const HYD = Ref(Hydration(phimin, phimax, RG, VH₂Oᶠ, VH₂Oᶠⁱ, ΔHWD, ΔSWD, ΔVWD, pfcoeff,
                          Δtreaction, log_completion_rate, A_I, b_I, c_I, Sxo_B, Tscl_B,
                          To_B, Sxo_T, To_T, Ea_T, reaction_rate_coeff_mode))

# update_marker_porosity!(xm, ym, tm, phim, APHI, dt, marknum)          E:4822
function update_marker_porosity!(ctx::Ctx, xm, ym, tm::Vector{Int64}, phim, APHI, dt, marknum)
    GC.@preserve xm ym tm phim APHI check(ctx, ccall(
        (:ereb_update_marker_porosity, LIB), Cint,
        (Ptr{Cvoid}, Ref{Hydration}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64},
         Ptr{Float64}, Ptr{Float64}, Float64, Int64),
        ctx.h, HYD, xm, ym, tm, phim, APHI, dt, marknum))
end

# the marker lattice of C:104-135, once per run
struct MarkerGrid                         # ereb_marker_grid
    Nxm::Int64; Nym::Int64; dxm::Float64; dym::Float64
    xxm::Ptr{Float64}; yym::Ptr{Float64}; dii::Int64; djj::Int64; mdis_init::Float64
end

# replenish_markers!(xm, ym, tm, …, etafluidcur_inv_kphim, mdis, mnum;
#                    randomized)                                     E:5586
# Julia grows the 26 vectors with push!; here they are grown to the upper
# bound (one new marker per anchor), filled, and shrunk to the new count.
function replenish_markers!(ctx::Ctx, vecs::Vararg{AbstractVector,26}; mdis, mnum,
                            randomized=random_markers, seed=rand(UInt64))
    n0 = length(vecs[1])
    cap = n0 + Nxm * Nym
    foreach(v -> resize!(v, cap), vecs)
    n = Ref{Int64}(0)
    xxm_, yym_ = collect(xxm), collect(yym)
    GC.@preserve vecs mdis mnum xxm_ yym_ begin
        mg = Ref(MarkerGrid(Nxm, Nym, dxm, dym, pointer(xxm_), pointer(yym_),
                            5 * Nymc, 5 * Nxmc, mdis_init))
        check(ctx, ccall((:ereb_replenish_markers, LIB), Cint,
            (Ptr{Cvoid}, ntuple(_ -> Ptr{Cvoid}, 26)..., Ptr{Float64}, Ptr{Int64},
             Int64, Int64, Ref{MarkerGrid}, Cint, UInt64, Ref{Int64}),
            ctx.h, map(pointer, vecs)..., mdis, mnum, n0, cap, mg, Cint(randomized),
            seed, n))
    end
    foreach(v -> resize!(v, n[]), vecs)
    return n[]
end

# ---- the 30 SUM grids and the marker loop E:6419-6517 ----------------------
# marker_to_nodes!(ctx, xm, ym, tm, tkm, phim, XWsolidm0, etavpm,
#     inv_gggtotalm, fricttotalm, cohestotalm, tenstotalm, rhofluidcur,
#     alphasolidcur, alphafluidcur, sxxm, sxym, rhototalm, rhocptotalm,
#     etatotalm, hrtotalm, ktotalm, tkm_rhocptotalm, etafluidcur_inv_kphim,
#     hrsolidm, hrfluidm, mode, sums::NTuple{30}, marknum)
function marker_to_nodes!(ctx::Ctx, inputs::NTuple{16,AbstractVector},
                          props::NTuple{7,Vector{Float64}}, hrsolidm, hrfluidm, mode,
                          sums::NTuple{30,Matrix{Float64}}, marknum)
    hs, hf = collect(Float64, hrsolidm), collect(Float64, hrfluidm)
    sp = collect(map(pointer, sums))
    GC.@preserve inputs props hs hf sums sp check(ctx, ccall(
        (:ereb_marker_to_nodes, LIB), Cint,
        (Ptr{Cvoid}, ntuple(_ -> Ptr{Cvoid}, 23)..., Ptr{Float64}, Ptr{Float64}, Int64,
         Ptr{Ptr{Float64}}, Int64),
        ctx.h, map(pointer, inputs)..., map(pointer, props)..., hs, hf, mode, sp, marknum))
end

# ---- ChainRulesCore rrules ---------------------------------------------------
# The reference functions mutate their arguments; the rrules are defined on
# non-mutating forms with the same arguments that return the updated arrays.
# Each pullback is one ccall of the library's pullback kernel, which
# recomputes what it needs from the forward inputs (no tape), sums grid
# cotangents in fixed point (the same bits on every run) and returns the
# cotangents of every differentiable input.  Cell indices, marker types and
# the property branches are piecewise constant (their tangents are zero);
# dt and mode are not differentiated.

_ptr_or_null(Δ) = Δ isa AbstractZero ? Ptr{Float64}(C_NULL) : pointer(Δ)
_dense(Δ) = Δ isa AbstractZero ? Δ : collect(Float64, unthunk(Δ))

# move_markers_rk4! E:4990-5257
function move_markers_rk4(ctx::Ctx, xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf, vyf,
                          wyx, tk2, dt, mode)
    x, y, T, sy, sx = copy(xm), copy(ym), copy(tkm), copy(sxym), copy(sxxm)
    move_markers_rk4!(ctx, x, y, tm, T, phim, sy, sx, vx, vy, vxf, vyf, wyx, tk2,
                      length(xm), dt, mode)
    return (x, y, T, sy, sx)
end

function ChainRulesCore.rrule(::typeof(move_markers_rk4), ctx::Ctx, xm, ym, tm, tkm, phim,
                              sxym, sxxm, vx, vy, vxf, vyf, wyx, tk2, dt, mode)
    out = move_markers_rk4(ctx, xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf, vyf, wyx,
                           tk2, dt, mode)
    function move_markers_rk4_pullback(Δ)
        Δx, Δy, ΔT, Δsy, Δsx = map(_dense, unthunk(Δ))
        n = length(xm)
        x̄, ȳ, T̄, φ̄, s̄y, s̄x = (zeros(n) for _ in 1:6)
        v̄x, v̄y, v̄xf, v̄yf, t̄k2 = (zeros(size(vx)) for _ in 1:5)
        w̄yx = zeros(size(wyx))
        GC.@preserve xm ym tm tkm phim sxym sxxm vx vy vxf vyf wyx tk2 Δx Δy ΔT Δsy Δsx begin
            check(ctx, ccall((:ereb_move_markers_rk4_pullback, LIB), Cint,
                (Ptr{Cvoid}, ntuple(_ -> Ptr{Cvoid}, 13)..., Int64, Float64, Int64,
                 ntuple(_ -> Ptr{Float64}, 17)...),
                ctx.h, xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf, vyf, wyx, tk2, n, dt,
                mode, _ptr_or_null(Δx), _ptr_or_null(Δy), _ptr_or_null(ΔT),
                _ptr_or_null(Δsy), _ptr_or_null(Δsx), x̄, ȳ, T̄, φ̄, s̄y, s̄x, v̄x, v̄y, v̄xf,
                v̄yf, w̄yx, t̄k2))
        end
        return (NoTangent(), NoTangent(), x̄, ȳ, NoTangent(), T̄, φ̄, s̄y, s̄x, v̄x, v̄y, v̄xf,
                v̄yf, w̄yx, t̄k2, @not_implemented("dt is not differentiated"), NoTangent())
    end
    return out, move_markers_rk4_pullback
end

# update_marker_temperature! E:4782-4801
function update_marker_temperature(ctx::Ctx, xm, ym, tkm, DT, tk2, timestep)
    T = copy(tkm)
    update_marker_temperature!(ctx, xm, ym, T, DT, tk2, timestep, length(xm))
    return T
end

function ChainRulesCore.rrule(::typeof(update_marker_temperature), ctx::Ctx, xm, ym, tkm, DT,
                              tk2, timestep)
    T = update_marker_temperature(ctx, xm, ym, tkm, DT, tk2, timestep)
    function update_marker_temperature_pullback(ΔT)
        ΔT = _dense(ΔT)
        ΔT isa AbstractZero && return (NoTangent(), NoTangent(), ZeroTangent(), ZeroTangent(),
                                       ZeroTangent(), ZeroTangent(), ZeroTangent(), NoTangent())
        n = length(xm)
        T̄, x̄, ȳ = zeros(n), zeros(n), zeros(n)
        D̄T, t̄k2 = zeros(size(tk2)), zeros(size(tk2))
        GC.@preserve xm ym DT tk2 ΔT check(ctx, ccall(
            (:ereb_update_marker_temperature_pullback, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Int64,
             ntuple(_ -> Ptr{Float64}, 6)...),
            ctx.h, xm, ym, DT, tk2, timestep, n, ΔT, T̄, D̄T, t̄k2, x̄, ȳ))
        return (NoTangent(), NoTangent(), x̄, ȳ, T̄, D̄T, t̄k2, NoTangent())
    end
    return T, update_marker_temperature_pullback
end

# update_marker_stress! E:4354-4388
function update_marker_stress(ctx::Ctx, xm, ym, sxxm, sxym, DSXX, DSXY)
    sx, sy = copy(sxxm), copy(sxym)
    update_marker_stress!(ctx, xm, ym, sx, sy, DSXX, DSXY, length(xm))
    return (sx, sy)
end

function ChainRulesCore.rrule(::typeof(update_marker_stress), ctx::Ctx, xm, ym, sxxm, sxym,
                              DSXX, DSXY)
    out = update_marker_stress(ctx, xm, ym, sxxm, sxym, DSXX, DSXY)
    function update_marker_stress_pullback(Δ)
        Δsx, Δsy = map(_dense, unthunk(Δ))
        n = length(xm)
        Δsx isa AbstractZero && (Δsx = zeros(n))
        Δsy isa AbstractZero && (Δsy = zeros(n))
        x̄, ȳ = zeros(n), zeros(n)
        D̄XX, D̄XY = zeros(size(DSXX)), zeros(size(DSXY))
        GC.@preserve xm ym DSXX DSXY Δsx Δsy check(ctx, ccall(
            (:ereb_update_marker_stress_pullback, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64,
             ntuple(_ -> Ptr{Float64}, 6)...),
            ctx.h, xm, ym, DSXX, DSXY, n, Δsx, Δsy, D̄XX, D̄XY, x̄, ȳ))
        # the stresses pass through unchanged
        return (NoTangent(), NoTangent(), x̄, ȳ, Δsx, Δsy, D̄XX, D̄XY)
    end
    return out, update_marker_stress_pullback
end

# the marker loop E:6419-6517: 16 marker inputs -> 7 properties, 30 SUM grids.
# inputs in the order xm, ym, tm, tkm, phim, XWsolidm0, etavpm, inv_gggtotalm,
# fricttotalm, cohestotalm, tenstotalm, rhofluidcur, alphasolidcur,
# alphafluidcur, sxxm, sxym; stored are the 5 property vectors air keeps
# (rhototalm, rhocptotalm, etatotalm, hrtotalm, ktotalm).
const M_FIELDS = (0, 1, 2, 3, 4, 7, 8, 9, 10, 11, 12, 13, 14, 15, 5, 6)  # ereb_marker_field
const P_FIELDS = (16, 17, 18, 19, 20, 21, 22)

function marker_to_nodes(ctx::Ctx, inputs::NTuple{16,AbstractVector}, stored::NTuple{5},
                         hrsolidm, hrfluidm, mode, shapes)
    n = length(inputs[1])
    props = (map(copy, stored)..., zeros(n), zeros(n))
    sums = ntuple(k -> zeros(shapes[k]), 30)
    marker_to_nodes!(ctx, inputs, props, hrsolidm, hrfluidm, mode, sums, n)
    return props, sums
end

function ChainRulesCore.rrule(::typeof(marker_to_nodes), ctx::Ctx, inputs, stored, hrsolidm,
                              hrfluidm, mode, shapes)
    props, sums = marker_to_nodes(ctx, inputs, stored, hrsolidm, hrfluidm, mode, shapes)
    function marker_to_nodes_pullback(Δ)
        Δprops, Δsums = unthunk(Δ)
        n = length(inputs[1])
        mk = fill(C_NULL, 26)
        for (k, f) in enumerate(M_FIELDS)
            mk[f + 1] = pointer(inputs[k])
        end
        for (k, f) in enumerate(P_FIELDS[1:5])
            mk[f + 1] = pointer(stored[k])
        end
        # dense copies of the cotangents, kept alive across the ccall
        sbv = [Δsums isa AbstractZero || Δsums[k] isa AbstractZero ? nothing :
               collect(Float64, Δsums[k]) for k in 1:30]
        pbv = [Δprops isa AbstractZero || Δprops[k] isa AbstractZero ? nothing :
               collect(Float64, Δprops[k]) for k in 1:7]
        sb = [a === nothing ? Ptr{Float64}(C_NULL) : pointer(a) for a in sbv]
        pb = [a === nothing ? Ptr{Float64}(C_NULL) : pointer(a) for a in pbv]
        bars = [zeros(n) for _ in 1:26]
        bp = map(pointer, bars)
        hs, hf = collect(Float64, hrsolidm), collect(Float64, hrfluidm)
        GC.@preserve inputs stored sbv pbv bars hs hf mk sb pb bp check(ctx, ccall(
            (:ereb_marker_to_nodes_pullback, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Int64, Ptr{Float64}, Ptr{Float64}, Int64,
             Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}),
            ctx.h, mk, n, hs, hf, mode, sb, pb, bp))
        inbar = ntuple(k -> M_FIELDS[k] == 2 ? NoTangent() : bars[M_FIELDS[k] + 1], 16)
        stbar = ntuple(k -> bars[P_FIELDS[k] + 1], 5)
        return (NoTangent(), NoTangent(), Tangent{typeof(inputs)}(inbar...),
                Tangent{typeof(stored)}(stbar...), NoTangent(), NoTangent(), NoTangent(),
                NoTangent())
    end
    return (props, sums), marker_to_nodes_pullback
end

end # module
