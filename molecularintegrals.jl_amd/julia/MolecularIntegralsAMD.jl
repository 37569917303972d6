"""
    MolecularIntegralsAMD

Julia front end of libmi_hgp.so, the MI355X HGP two-electron integral engine.
Drop-in GPU versions of MolecularIntegrals.jl's ERI drivers:

* `all_twoe_ints_chrr_gpu(bfs::MolecularIntegrals.Basis)` -- replaces
  `all_twoe_ints_chrr` (src/HGP.jl:22-34), same slot order (enumeration of
  `iiterator`, src/Basis.jl:203);
* `all_twoe_ints_gpu(bfs)` -- replaces `all_twoe_ints(bfs)` (src/ERI.jl:104-112),
  `iindex` order;
* `all_twoe_ints_rys_gpu(bfs)` -- replaces `all_twoe_ints_rys(bfs)`
  (src/Rys.jl:16-25): Rys quadrature, `iindex` order;
* `jk_gpu(bfs, D)` -- direct Coulomb/exchange (make2JmK's J and K, src/ERI.jl:114-128).

Values are the normalised contracted `coulomb(bfs[i],bfs[j],bfs[k],bfs[l])`.
The C ABI is include/mi_hgp.h; set `ENV["MI_HGP_LIB"]` to the library path.
"""
module MolecularIntegralsAMD

using MolecularIntegrals

export all_twoe_ints_chrr_gpu, all_twoe_ints_gpu, all_twoe_ints_rys_gpu, jk_gpu

const libmi = get(ENV, "MI_HGP_LIB", joinpath(@__DIR__, "..", "libmi_hgp.so"))

const MI_ORDER_IINDEX = Int32(0)
const MI_ORDER_CHRR = Int32(1)

# struct mi_basis (include/mi_hgp.h)
struct MiBasis
    nshell::Int32
    shell_xyz::Ptr{Float64}
    shell_l::Ptr{Int32}
    shell_prim_off::Ptr{Int32}
    prim_exp::Ptr{Float64}
    prim_coef::Ptr{Float64}
    prim_norm::Ptr{Float64}
    nbf::Int32
    bf_scale::Ptr{Float64}
end

# struct mi_opts
struct MiOpts
    order::Int32
    precision::Int32
    tau::Float64
    schwarz_cut::Float64
    stream::Ptr{Cvoid}
    method::Int32
end
const MI_METHOD_HGP = Int32(0)
const MI_METHOD_RYS = Int32(1)

mutable struct Ctx
    h::Ptr{Cvoid}
end

const _ctx = Ref{Union{Nothing,Ctx}}(nothing)

function context()
    if _ctx[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:mi_create, libmi), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, h)
        rc == 0 || error("mi_create failed with code $rc (no usable GPU?)")
        c = Ctx(h[])
        finalizer(c -> ccall((:mi_destroy, libmi), Cvoid, (Ptr{Cvoid},), c.h), c)
        _ctx[] = c
    end
    return _ctx[]
end

check(c::Ctx, rc) = rc == 0 || error(unsafe_string(ccall((:mi_last_error, libmi), Cstring, (Ptr{Cvoid},), c.h)))

"Flatten a Basis into the C-ABI arrays, reading the norms its objects hold."
function flatten(bfs::MolecularIntegrals.Basis)
    xyz, L, off, ex, co, pn = Float64[], Int32[], Int32[0], Float64[], Float64[], Float64[]
    first0 = Float64[]
    for sh in bfs.shells
        append!(xyz, sh.xyz)
        push!(L, sh.L)
        first = sh.cgbfs[1]
        for (p, (e, c)) in enumerate(zip(sh.expns, sh.coefs))
            push!(ex, e); push!(co, c); push!(pn, first.pgbfs[p].norm)
        end
        push!(off, length(ex))
        push!(first0, first.pgbfs[1].norm)
    end
    scale = [cg.norm * cg.pgbfs[1].norm / first0[bfs.ishell[i]] for (i, cg) in enumerate(bfs.cgbfs)]
    return xyz, L, off, ex, co, pn, scale
end

function _eri(bfs, order, method=MI_METHOD_HGP)
    c = context()
    xyz, L, off, ex, co, pn, scale = flatten(bfs)
    n = length(bfs.cgbfs)
    out = Vector{Float64}(undef, div(n * (n + 1) * (n * n + n + 2), 8))
    GC.@preserve xyz L off ex co pn scale out begin
        b = Ref(MiBasis(length(L), pointer(xyz), pointer(L), pointer(off), pointer(ex), pointer(co),
                        pointer(pn), n, pointer(scale)))
        o = Ref(MiOpts(order, 0, 0.0, 0.0, C_NULL, method))
        rc = ccall((:mi_eri_unique, libmi), Cint, (Ptr{Cvoid}, Ref{MiBasis}, Ref{MiOpts}, Ptr{Float64}, Int64),
                   c.h, b, o, out, length(out))
        check(c, rc)
    end
    return out
end

all_twoe_ints_chrr_gpu(bfs) = _eri(bfs, MI_ORDER_CHRR)
all_twoe_ints_gpu(bfs) = _eri(bfs, MI_ORDER_IINDEX)
# all_twoe_ints_rys (src/Rys.jl:16-25): Rys quadrature on the GPU, iindex order
all_twoe_ints_rys_gpu(bfs) = _eri(bfs, MI_ORDER_IINDEX, MI_METHOD_RYS)

function jk_gpu(bfs, D::Matrix{Float64}; schwarz_cut=0.0)
    c = context()
    xyz, L, off, ex, co, pn, scale = flatten(bfs)
    n = length(bfs.cgbfs)
    Dr = permutedims(D)  # row-major for the ABI
    J = Matrix{Float64}(undef, n, n)
    K = Matrix{Float64}(undef, n, n)
    GC.@preserve xyz L off ex co pn scale Dr J K begin
        b = Ref(MiBasis(length(L), pointer(xyz), pointer(L), pointer(off), pointer(ex), pointer(co),
                        pointer(pn), n, pointer(scale)))
        o = Ref(MiOpts(MI_ORDER_IINDEX, 0, 0.0, schwarz_cut, C_NULL, MI_METHOD_HGP))
        rc = ccall((:mi_jk, libmi), Cint,
                   (Ptr{Cvoid}, Ref{MiBasis}, Ref{MiOpts}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   c.h, b, o, Dr, J, K)
        check(c, rc)
    end
    return permutedims(J), permutedims(K)  # symmetric anyway
end

end # module
