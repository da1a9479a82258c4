"""
    FermiHIP

Julia side of the MI355X DF-RCCSD(T) backend: registers `RCCSDhip <: RCCSDAlgorithm` and
`ijkhip <: RpTAlgorithm` with Fermi.jl and forwards the RCCSD / RCCSD(T) method family to
`libfermihip.so` through `ccall` (declarations in include/fermihip.h).

Selection: `@set cc_alg 2` picks `RCCSDhip` through `get_rccsd_alg(Val(2))` (the mechanism of
Fermi's RCCSD.jl:76-88); `@set pt_alg 4` picks `ijkhip` through the `get_rpt_alg` method
below (PerturbativeTriples.jl:3-11 keeps a fixed vector, so this module redefines it).
"""
module FermiHIP

using Fermi
using Fermi.Options
using Fermi.Integrals: IntegralHelper, AbstractDFERI
using Fermi.CoupledCluster: RCCSDAlgorithm, RpTAlgorithm, ijk, ijk2, abc
import Fermi.CoupledCluster: RCCSD, RCCSDpT, get_rccsd_alg, get_rpt_alg

const libfermihip = get(ENV, "FERMIHIP_LIB", joinpath(@__DIR__, "..", "fermihip", "libfermihip.so"))

struct RCCSDhip <: RCCSDAlgorithm end
struct ijkhip <: RpTAlgorithm end

get_rccsd_alg(::Val{2}) = RCCSDhip()
function get_rpt_alg()
    N = Options.get("pt_alg")
    N == 4 && return ijkhip()
    implemented = [ijk(), ijk2(), abc()]
    1 <= N <= 3 || throw(Fermi.Options.FermiException("implementation number $N not available for RCCSD(T)."))
    return implemented[N]
end

# --- C ABI mirrors (field order and types as in include/fermihip.h)
struct CCOpts
    cc_max_iter::Cint
    cc_e_conv::Cdouble
    cc_max_rms::Cdouble
    cc_diis::Cint
    cc_ndiis::Cint
    diis_start::Cint
    cc_diis_relax::Cint
    diis_prec::Cint
    cc_damp_ratio::Cdouble
    precision_override::Cint
    verbose::Cint
    precision_mode::Cint
    polish_iters::Cint
    mixed_e_switch::Cdouble
    mixed_rms_switch::Cdouble
    bf3_e_switch::Cdouble
    bf3_rms_switch::Cdouble
end

struct CCResult
    guess_energy::Cdouble
    correlation::Cdouble
    e_conv::Cdouble
    t_conv::Cdouble
    niter::Cint
    converged::Cint
    niter_mixed::Cint
    niter_bf3::Cint
end

mutable struct Context
    h::Ptr{Cvoid}
end

function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:fh_last_error, libfermihip), Cstring, ()))
    throw(Fermi.Options.FermiException("libfermihip error $rc: $msg"))
end

function Context(device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fh_create, libfermihip), Cint, (Ref{Ptr{Cvoid}}, Cint), h, device))
    ctx = Context(h[])
    finalizer(c -> ccall((:fh_destroy, libfermihip), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    return ctx
end

# one device context per MO integral helper (the DF tensors live in HBM until it is collected)
const CONTEXTS = WeakKeyDict{IntegralHelper,Context}()

function context_for(moints::IntegralHelper)
    haskey(CONTEXTS, moints) && return CONTEXTS[moints]
    Boo, Bov, Bvv = moints["BOO"], moints["BOV"], moints["BVV"]
    fo, fv = moints["Fii"], moints["Faa"]
    Q, o, v = size(Bov)
    ctx = Context(0)
    GC.@preserve Boo Bov Bvv fo fv begin
        check(ccall((:fh_set_df, libfermihip), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    ctx.h, o, v, Q, fo, fv, Boo, Bov, Bvv))
    end
    CONTEXTS[moints] = ctx
    return ctx
end

function cc_opts()
    prec = Options.get("diis_prec")
    prec in ("single", "double") || throw(Fermi.Options.FermiException("diis_prec $prec is not supported by RCCSDhip"))
    CCOpts(Options.get("cc_max_iter"), Options.get("cc_e_conv"), Options.get("cc_max_rms"), Options.get("cc_diis"),
           Options.get("cc_ndiis"), Options.get("diis_start"), Options.get("cc_diis_relax"), prec == "single" ? 1 : 0,
           Options.get("cc_damp_ratio"), Options.get("precision_override"), 0,
           0, -1, 1e-9, 1e-9, 1e-7, 1e-8)   # FP64 contractions; the mixed modes are selected by the Python/C callers
end

# RCCSD(moints, aoints, ::RCCSDa) for DF ERIs (RCCSDa.jl:37-44): only the three B blocks and F are needed
function RCCSD(moints::IntegralHelper{Float64,E1,O}, aoints::IntegralHelper{Float64,E2,Fermi.Orbitals.AtomicOrbitals},
               alg::RCCSDhip) where {E1<:AbstractDFERI,E2<:AbstractDFERI,O}
    for key in ("F", "BOO", "BOV", "BVV")
        Fermi.Integrals.compute!(moints, aoints, key)
    end
    RCCSD(moints, alg)
end

# MP2 guess on the device (RCCSDa.jl:46-90)
function RCCSD(moints::IntegralHelper{Float64}, alg::RCCSDhip)
    ctx = context_for(moints)
    Q, o, v = size(moints["BOV"])
    T1 = zeros(Float64, o, v)
    T2 = zeros(Float64, o, o, v, v)
    check(ccall((:fh_t2_guess, libfermihip), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}), ctx.h, 0, T1, T2))
    RCCSD(moints, T1, T2, alg)
end

# CC iterations from given amplitudes (RCCSDa.jl:92-245)
function RCCSD(moints::IntegralHelper{Float64}, newT1::AbstractArray{Float64,2}, newT2::AbstractArray{Float64,4},
               alg::RCCSDhip)
    ctx = context_for(moints)
    T1 = Array{Float64}(newT1)
    T2 = Array{Float64}(newT2)
    opts = Ref(cc_opts())
    res = Ref{CCResult}()
    GC.@preserve T1 T2 begin
        check(ccall((:fh_ccsd_solve, libfermihip), Cint, (Ptr{Cvoid}, Ref{CCOpts}, Ptr{Float64}, Ptr{Float64}, Ref{CCResult}),
                    ctx.h, opts, T1, T2, res))
    end
    r = res[]
    Eref = moints.orbitals.sd_energy
    out = Fermi.CoupledCluster.RCCSD{Float64}(r.guess_energy + Eref, r.correlation, r.correlation + Eref, T1, T2,
                                              r.e_conv, r.t_conv)
    return Options.get("return_ints") ? (out, moints) : out
end

# (T) (ijk.jl:20-150)
function RCCSDpT(ccsd::Fermi.CoupledCluster.RCCSD, moints::IntegralHelper{Float64}, alg::ijkhip)
    ctx = context_for(moints)
    Et = Ref{Float64}(0.0)
    T1 = Array{Float64}(ccsd.T1)
    T2 = Array{Float64}(ccsd.T2)
    GC.@preserve T1 T2 begin
        check(ccall((:fh_triples, libfermihip), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                    ctx.h, 0, T1, T2, Et))
    end
    return Fermi.CoupledCluster.RCCSDpT{Float64}(ccsd, Et[] + ccsd.energy, Et[])
end

function RCCSDpT(alg::ijkhip)
    val = Options.get("return_ints")
    Options.set("return_ints", true)
    ccsd, moints = RCCSD()
    Options.set("return_ints", val)
    return RCCSDpT(ccsd, moints, alg)
end

end # module
