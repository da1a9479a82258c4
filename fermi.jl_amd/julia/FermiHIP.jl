"""
    FermiHIP

Julia side of the MI355X DF-RCCSD(T) backend: registers `RCCSDhip <: RCCSDAlgorithm` and
`ijkhip <: RpTAlgorithm` with Fermi.jl and forwards the RCCSD / RCCSD(T) method family to
`libfermihip.so` through `ccall` (declarations in include/fermihip.h).

Selection: `@set cc_alg 2` picks `RCCSDhip` through `get_rccsd_alg(Val(2))` (the mechanism of
Fermi's RCCSD.jl:76-88); `@set pt_alg 4` picks `ijkhip` through the `get_rpt_alg` method
below (PerturbativeTriples.jl:3-11 keeps a fixed vector, so this module redefines it);
`@set mp2_alg 3` picks `RMP2hip` through `get_rmp2_alg(Val(3))` (RMP2.jl:60-77), whose DF energy
`RMP2_energy` (the method RMP2a.jl:80 calls) runs on the device; `@set rhf_alg 3` picks `RHFhip`
through `get_rhf_alg(Val(3))` (RHF.jl:97-118), the DF-RHF SCF of RHFa.jl:14-223 on the device.
"""
module FermiHIP

using Fermi
using Fermi.Options
using Fermi.Integrals: IntegralHelper, AbstractDFERI
using Fermi.CoupledCluster: RCCSDAlgorithm, RpTAlgorithm, ijk, ijk2, abc
import Fermi.CoupledCluster: RCCSD, RCCSDpT, get_rccsd_alg, get_rpt_alg
using Fermi.MollerPlesset: RMP2Algorithm
import Fermi.MollerPlesset: RMP2_energy, get_rmp2_alg
using Fermi.HartreeFock: RHFAlgorithm
import Fermi.HartreeFock: RHF, get_rhf_alg

const libfermihip = get(ENV, "FERMIHIP_LIB", joinpath(@__DIR__, "..", "fermihip", "libfermihip.so"))

struct RCCSDhip <: RCCSDAlgorithm end
struct ijkhip <: RpTAlgorithm end
struct RMP2hip <: RMP2Algorithm end
struct RHFhip <: RHFAlgorithm end

get_rccsd_alg(::Val{2}) = RCCSDhip()
get_rmp2_alg(::Val{3}) = RMP2hip()
get_rhf_alg(::Val{3}) = RHFhip()
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

# RCCSD(moints, aoints, ::RCCSDa) for DF ERIs (RCCSDa.jl:37-44): only the three B blocks and F are
# needed.  The AO -> MO transforms (DFERI.jl:15-69) run on the device unless the blocks are cached
# already; the MO blocks are copied back into the moints cache, where compute_BOO!/BOV!/BVV! put them.
function RCCSD(moints::IntegralHelper{Float64,E1,O}, aoints::IntegralHelper{Float64,E2,Fermi.Orbitals.AtomicOrbitals},
               alg::RCCSDhip) where {E1<:AbstractDFERI,E2<:AbstractDFERI,O}
    Fermi.Integrals.compute!(moints, aoints, "F")
    if all(k -> haskey(moints.cache, k), ("BOO", "BOV", "BVV"))
        context_for(moints)
    else
        transform_df!(moints, aoints)
    end
    RCCSD(moints, alg)
end

function transform_df!(moints::IntegralHelper{Float64}, aoints::IntegralHelper{Float64})
    b = Array{Float64}(aoints["ERI"])
    C = Array{Float64}(moints.orbitals.C)
    eps = Array{Float64}(moints.orbitals.eps)
    Q, nbf, _ = size(b)
    ndocc = moints.molecule.Nα
    core, inac = Options.get("drop_occ"), Options.get("drop_vir")
    ctx = Context(0)
    GC.@preserve b C eps begin
        check(ccall((:fh_set_df_ao, libfermihip), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    ctx.h, nbf, Q, ndocc, core, inac, eps, C, b))
    end
    o, v = ndocc - core, nbf - inac - ndocc
    Boo, Bov, Bvv = zeros(Float64, Q, o, o), zeros(Float64, Q, o, v), zeros(Float64, Q, v, v)
    check(ccall((:fh_get_df, libfermihip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx.h, Boo, Bov, Bvv))
    moints["BOO"], moints["BOV"], moints["BVV"] = Boo, Bov, Bvv
    CONTEXTS[moints] = ctx
    return ctx
end

# DF-MP2 energy (RMP2a.jl:91-144) on the device: an MP2-only context holds BOV and the orbital energies
function RMP2_energy(ints::IntegralHelper{Float64,<:AbstractDFERI,Fermi.Orbitals.RHFOrbitals}, alg::RMP2hip)
    Bov, fo, fv = ints["BOV"], ints["Fii"], ints["Faa"]
    Q, o, v = size(Bov)
    ctx = Context(0)
    E = Ref{Float64}(0.0)
    GC.@preserve Bov fo fv begin
        check(ccall((:fh_set_df, libfermihip), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    ctx.h, o, v, Q, fo, fv, C_NULL, Bov, C_NULL))
    end
    check(ccall((:fh_mp2_energy, libfermihip), Cint, (Ptr{Cvoid}, Ref{Float64}), ctx.h, E))
    return E[]
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

# --- DF-RHF SCF (RHFa.jl:14-223) through fh_rhf
struct SCFOpts
    scf_max_iter::Cint
    scf_e_conv::Cdouble
    scf_max_rms::Cdouble
    diis::Cint
    ndiis::Cint
    diis_start::Cint
    oda::Cint
    oda_cutoff::Cdouble
    oda_shutoff::Cint
    guess::Cint
end

struct SCFResult
    energy::Cdouble
    e_conv::Cdouble
    d_rms::Cdouble
    niter::Cint
    converged::Cint
end

function RHF(ints::IntegralHelper{Float64,<:AbstractDFERI,Fermi.Orbitals.AtomicOrbitals}, alg::RHFhip)
    molecule = ints.molecule
    ndocc = try
        Int((molecule.Nα + molecule.Nβ) / 2)
    catch InexactError
        throw(Fermi.Options.FermiException("Invalid number of electrons $(molecule.Nα + molecule.Nβ) for RHF method."))
    end
    guess = Options.get("scf_guess")
    guess in ("gwh", "core") || throw(Fermi.Options.FermiException("scf_guess $guess is not supported by RHFhip"))
    S = Array{Float64}(ints["S"])
    H = Array{Float64}(ints["T"] .+ ints["V"])
    b = Array{Float64}(ints["ERI"])
    Q, nbf, _ = size(b)
    Vnuc = Fermi.Molecules.nuclear_repulsion(molecule.atoms)
    opts = Ref(SCFOpts(Options.get("scf_max_iter"), Options.get("scf_e_conv"), Options.get("scf_max_rms"),
                       Options.get("diis") ? 1 : 0, Options.get("ndiis"), Options.get("diis_start"),
                       Options.get("oda") ? 1 : 0, Options.get("oda_cutoff"), Options.get("oda_shutoff"),
                       guess == "gwh" ? 0 : 1))
    eps = zeros(Float64, nbf)
    C = zeros(Float64, nbf, nbf)
    res = Ref{SCFResult}()
    ctx = Context(0)
    GC.@preserve S H b eps C begin
        check(ccall((:fh_rhf, libfermihip), Cint,
                    (Ptr{Cvoid}, Ref{SCFOpts}, Cint, Cint, Cint, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}, Ref{SCFResult}),
                    ctx.h, opts, nbf, Q, ndocc, Vnuc, S, H, b, eps, C, res))
    end
    r = res[]
    orbitals = Fermi.Orbitals.RHFOrbitals(molecule, ints.basis, eps, r.energy, C)
    return Fermi.HartreeFock.RHF(molecule, r.energy, ndocc, nbf - ndocc, orbitals, r.e_conv, r.d_rms)
end

end # module
