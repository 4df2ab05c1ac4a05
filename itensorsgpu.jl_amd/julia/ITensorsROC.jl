# ITensorsROC — drop-in MI355X storage backend for ITensors.jl 0.2 dense
# tensors, the counterpart of ITensorsGPU's CuDense (src/tensor/cudense.jl).
#
# Storage is `Dense{ElT, RocVector{ElT}}` where RocVector is a thin owner of
# a libitg device buffer. Every method below ccalls libitg (include/itg.h);
# none of them computes on the host. Method list mirrors the generics the
# reference specialises (src/ITensorsGPU.jl:32-43): contract!!, _contract!,
# permutedims!!, permute!, +, -, *, /, norm, getindex, svd, eigen, qr,
# truncate!, plus roc/cpu conversions in place of cu/cpu
# (src/cuitensor.jl:1-33, src/mps/cumps.jl:1-9, src/mps/cumpo.jl:1-31).
module ITensorsROC

using ITensors, ITensors.NDTensors, LinearAlgebra
import ITensors: randn!
import ITensors.NDTensors: contract!!, permutedims!!, similartype, truncate!
import Base: +, -, *, /, permute!
import LinearAlgebra: norm, svd, eigen, qr

const libitg = get(ENV, "ITG_LIB", joinpath(@__DIR__, "..", "itg", "libitg.so"))

# ------------------------------------------------------------------ context
const CTX = Ref{Ptr{Cvoid}}(C_NULL)
function ctx()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:itg_init, libitg), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, r)
        st == 0 || error("itg_init failed: $st")
        CTX[] = r[]
    end
    return CTX[]
end

# status -> the exceptions the reference throws (cuitensor.jl:55,72; cumpo.jl:21)
function check(st::Cint)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:itg_last_error, libitg), Cstring, (Ptr{Cvoid},), ctx()))
    st == -1 && throw(DimensionMismatch(msg))
    (st == -2 || st == -6) && throw(ArgumentError(msg))
    error("libitg status $st: $msg")
end

dtype(::Type{Float64}) = Cint(0)
dtype(::Type{Float32}) = Cint(1)
dtype(::Type{ComplexF64}) = Cint(2)   # interleaved (re, im), Julia's layout

# ------------------------------------------------------------- device vector
mutable struct RocVector{T} <: AbstractVector{T}
    ptr::Ptr{T}
    len::Int
    function RocVector{T}(::UndefInitializer, n::Integer) where {T}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:itg_malloc, libitg), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}),
                    ctx(), max(n, 1) * sizeof(T), p))
        v = new{T}(Ptr{T}(p[]), n)
        finalizer(x -> ccall((:itg_free, libitg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx(), x.ptr), v)
        return v
    end
end
Base.size(v::RocVector) = (v.len,)
Base.pointer(v::RocVector) = v.ptr

# scalar read: one element copied to the host (synchronising), as the
# reference's `collect(data(D))[]` (src/tensor/cudense.jl:25-26)
function Base.getindex(v::RocVector{T}, i::Int) where {T}
    checkbounds(v, i)
    a = Ref{T}()
    check(ccall((:itg_memcpy_d2h, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                ctx(), a, v.ptr + (i - 1) * sizeof(T), sizeof(T)))
    return a[]
end

function RocVector{T}(v::RocVector{T}, n::Integer) where {T}   # first n elements, own storage
    w = RocVector{T}(undef, n)
    check(ccall((:itg_memcpy_d2d, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                ctx(), w.ptr, v.ptr, n * sizeof(T)))
    return w
end

function zeros_roc(::Type{T}, n::Integer) where {T}
    v = RocVector{T}(undef, n)
    check(ccall((:itg_fill, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, Cdouble, Ptr{T}),
                ctx(), dtype(T), n, 0.0, v.ptr))
    return v
end

function RocVector(a::Vector{T}) where {T}
    v = RocVector{T}(undef, length(a))
    check(ccall((:itg_memcpy_h2d, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                ctx(), v.ptr, a, sizeof(a)))
    return v
end
function Base.collect(v::RocVector{T}) where {T}
    a = Vector{T}(undef, v.len)
    check(ccall((:itg_memcpy_d2h, libitg), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                ctx(), a, v.ptr, sizeof(a)))
    return a
end

const RocDense{ElT} = Dense{ElT,RocVector{ElT}}
const RocDenseTensor{ElT,N} = Tensor{ElT,N,<:RocDense{ElT}}

Base.getindex(D::RocDense{<:Number}) = (length(data(D)) == 1 || throw(DimensionMismatch(
    "scalar read of a storage of length $(length(data(D)))")); data(D)[1])
Base.getindex(T::RocDenseTensor{<:Number,0}) = store(T)[]

# the generics of src/ITensorsGPU.jl:36-43 / cudense.jl:15-16 for the storage vector
similartype(::Type{<:RocVector}, eltype::Type) = RocVector{eltype}
NDTensors.similar(::Type{<:RocVector{T}}, dims) where {T} = RocVector{T}(undef, prod(dims))

# ---------------------------------------------------------------- conversion
roc(A::ITensor) = ITensor(Dense(RocVector(Vector(data(store(A))))), inds(A))
cpu(A::ITensor) = storage(A) isa RocDense ? ITensor(Dense(collect(data(storage(A)))), inds(A)) : A
roc(ψ::MPS) = (ϕ = copy(ψ); for j in eachindex(ϕ) ϕ[j] = roc(ψ[j]) end; ϕ)
roc(H::MPO) = (G = copy(H); for j in eachindex(G) G[j] = roc(H[j]) end; G)
cpu(ψ::Union{MPS,MPO}) = (ϕ = copy(ψ); for j in eachindex(ϕ) ϕ[j] = cpu(ψ[j]) end; ϕ)

function Base.similar(::Type{<:RocDenseTensor{ElT}}, inds) where {ElT}
    return tensor(Dense(RocVector{ElT}(undef, dim(inds))), inds)
end

# ------------------------------------------------------------------ random
# Philox Gaussian stream (itg_randn) in place of CURAND randn!
# (src/cuitensor.jl:35-49, src/mps/cumps.jl:44-53); successive calls draw
# successive parts of the stream of the current seed.
const SEED = Ref{UInt64}(0)
const OFFSET = Ref{UInt64}(0)
seed!(s::Integer) = (SEED[] = UInt64(s); OFFSET[] = 0; nothing)

# ComplexF64: unit-variance complex Gaussians, (g[2i] + i g[2i+1]) / sqrt(2)
function randn!(T::RocDenseTensor{ElT}) where {ElT<:Union{Float32,Float64,ComplexF64}}
    n = dim(T)
    check(ccall((:itg_randn, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, UInt64, UInt64, Ptr{ElT}),
                ctx(), dtype(ElT), n, SEED[], OFFSET[], pointer(data(store(T)))))
    OFFSET[] += n
    return T
end
randn!(T::ITensor) = (randn!(tensor(T)); T)

rocITensor(::Type{T}, inds::Indices) where {T<:Number} =
    ITensor(Dense(zeros_roc(T, dim(inds))), inds)
rocITensor(::Type{T}, inds::Index...) where {T<:Number} = rocITensor(T, IndexSet(inds...))
# real and ComplexF64 element types (randomCuITensor, src/cuitensor.jl:35-49)
randomRocITensor(::Type{S}, inds::Indices) where {S<:Union{Real,ComplexF64}} =
    randn!(rocITensor(S, inds))
randomRocITensor(::Type{S}, inds::Index...) where {S<:Union{Real,ComplexF64}} =
    randomRocITensor(S, IndexSet(inds...))
randomRocITensor(inds::Indices) = randomRocITensor(Float64, inds)
randomRocITensor(inds::Index...) = randomRocITensor(Float64, IndexSet(inds...))

function rocnormalize!(T::ITensor)
    st = tensor(T)
    nrm = norm(st)
    check(ccall((:itg_scal, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, Cdouble, Ptr{eltype(st)}),
                ctx(), dtype(eltype(st)), dim(st), 1 / nrm, pointer(data(store(st)))))
    return T
end

# cuMPS(T, sites; linkdims) builds dimension-1 links (src/mps/cumps.jl:12-38,
# quirk A10); linkdims is honoured here
function rocMPS(::Type{T}, sites::Vector{<:Index}; linkdims::Integer=1) where {T<:Number}
    N = length(sites)
    N == 1 && return MPS([rocITensor(T, sites[1])])
    l = [Index(linkdims, "Link,l=$ii") for ii in 1:N-1]
    v = Vector{ITensor}(undef, N)
    for ii in 1:N
        v[ii] = ii == 1 ? rocITensor(T, l[1], sites[1]) :
                ii == N ? rocITensor(T, l[N-1], sites[N]) :
                rocITensor(T, l[ii-1], sites[ii], l[ii])
    end
    return MPS(v, 0, N + 1)
end
rocMPS(sites::Vector{<:Index}; kwargs...) = rocMPS(Float64, sites; kwargs...)

function randomRocMPS(sites::Vector{<:Index}; linkdims::Integer=1)
    M = rocMPS(sites; linkdims=linkdims)
    for i in eachindex(sites)
        randn!(M[i])
        rocnormalize!(M[i])
    end
    M.llim = 1
    M.rlim = length(M)
    return M
end

# ------------------------------------------------------------- Diag storage
# Diag x Dense (src/tensor/cudiag.jl:118-162) through itg_diag_contract:
# only the diagonal of D is read; a uniform Diag passes its value instead
const RocDiag{ElT} = Diag{ElT,RocVector{ElT}}
function contract!!(R::RocDenseTensor{ElT,NR}, labelsR, D::Tensor{ElT,ND,<:RocDiag{ElT}},
                    labelsD, T::RocDenseTensor{ElT,NT}, labelsT) where {ElT,NR,ND,NT}
    dD = Int64[dims(D)...]; dT = Int64[dims(T)...]; dC = Int64[dims(R)...]
    lD = Int32[labelsD...]; lT = Int32[labelsT...]; lC = Int32[labelsR...]
    check(ccall((:itg_diag_contract, libitg), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}, Ptr{Int32}, Ptr{ElT}, Int64, Ptr{Cdouble},
                 Cint, Ptr{Int64}, Ptr{Int32}, Ptr{ElT}, Cint, Ptr{Int64}, Ptr{Int32}, Ptr{ElT},
                 Cdouble, Cdouble),
                ctx(), dtype(ElT), ND, dD, lD, pointer(data(store(D))), Int64(minimum(dD)),
                C_NULL, NT, dT, lT, pointer(data(store(T))), NR, dC, lC,
                pointer(data(store(R))), 1.0, 0.0))
    return R
end
contract!!(R::RocDenseTensor, labelsR, T::RocDenseTensor, labelsT,
           D::Tensor{<:Number,<:Any,<:RocDiag}, labelsD) = contract!!(R, labelsR, D, labelsD,
                                                                     T, labelsT)

# ----------------------------------------------------------------- contract
# Labels follow the einsum convention of itg_contract; NDTensors labels
# (negative = contracted) map directly.
function contract!!(R::RocDenseTensor{ElT,NR}, labelsR, T1::RocDenseTensor{ElT,N1}, labelsT1,
                    T2::RocDenseTensor{ElT,N2}, labelsT2, α::Number=one(ElT),
                    β::Number=zero(ElT)) where {ElT,NR,N1,N2}
    dA = Int64[dims(T1)...]; dB = Int64[dims(T2)...]; dC = Int64[dims(R)...]
    lA = Int32[labelsT1...]; lB = Int32[labelsT2...]; lC = Int32[labelsR...]
    check(ccall((:itg_contract, libitg), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Int64}, Ptr{Int32}, Ptr{ElT}, Cint, Ptr{Int64},
                 Ptr{Int32}, Ptr{ElT}, Cint, Ptr{Int64}, Ptr{Int32}, Ptr{ElT}, Cdouble, Cdouble),
                ctx(), dtype(ElT), 0, N1, dA, lA, pointer(data(store(T1))), N2, dB, lB,
                pointer(data(store(T2))), NR, dC, lC, pointer(data(store(R))), α, β))
    return R
end

function permutedims!!(B::RocDenseTensor{ElT,N}, A::RocDenseTensor{ElT,N}, perm::NTuple{N,Int},
                       f=(r, t) -> t) where {ElT,N}
    dA = Int64[dims(A)...]
    p = Int32[(perm .- 1)...]
    check(ccall((:itg_permute, libitg), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}, Ptr{Int32}, Cdouble, Ptr{ElT}, Ptr{ElT}),
                ctx(), dtype(ElT), N, dA, p, 1.0, pointer(data(store(A))), pointer(data(store(B)))))
    return B
end

function permute!(B::RocDenseTensor{ElT,N}, A::RocDenseTensor{ElT,N}) where {ElT,N}
    perm = ntuple(j -> findfirst(==(inds(B)[j]), inds(A)), N)
    return permutedims!!(B, A, perm)
end

function _add!(B::RocDenseTensor{ElT,N}, A::RocDenseTensor{ElT,N}, α) where {ElT,N}
    perm = Int32[(findfirst(==(inds(B)[j]), inds(A)) - 1 for j in 1:N)...]
    check(ccall((:itg_add_permuted, libitg), Cint,
                (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}, Ptr{Int32}, Cdouble, Ptr{ElT}, Cdouble, Ptr{ElT}),
                ctx(), dtype(ElT), N, Int64[dims(A)...], perm, α, pointer(data(store(A))), 1.0,
                pointer(data(store(B)))))
    return B
end
+(B::RocDenseTensor, A::RocDenseTensor) = _add!(copy(B), A, 1.0)
-(B::RocDenseTensor, A::RocDenseTensor) = _add!(copy(B), A, -1.0)

function *(D::RocDense{ElT}, x::Number) where {ElT}
    v = RocVector{ElT}(undef, length(data(D)))
    check(ccall((:itg_memcpy_d2d, libitg), Cint, (Ptr{Cvoid}, Ptr{ElT}, Ptr{ElT}, Csize_t),
                ctx(), v.ptr, pointer(data(D)), length(v) * sizeof(ElT)))
    if x isa Complex
        ElT === ComplexF64 || throw(ArgumentError("complex scale of real storage"))
        check(ccall((:itg_zscal, libitg), Cint, (Ptr{Cvoid}, Int64, Cdouble, Cdouble, Ptr{ElT}),
                    ctx(), length(v), real(x), imag(x), v.ptr))
    else
        check(ccall((:itg_scal, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, Cdouble, Ptr{ElT}),
                    ctx(), dtype(ElT), length(v), x, v.ptr))
    end
    return Dense(v)
end

# conj of complex storage (what dag does to the data); real storage is shared
function Base.conj(D::RocDense{ComplexF64})
    v = RocVector{ComplexF64}(undef, length(data(D)))
    check(ccall((:itg_memcpy_d2d, libitg), Cint,
                (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Csize_t),
                ctx(), v.ptr, pointer(data(D)), length(v) * sizeof(ComplexF64)))
    check(ccall((:itg_conj, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, Ptr{ComplexF64}),
                ctx(), dtype(ComplexF64), length(v), v.ptr))
    return Dense(v)
end
Base.conj(D::RocDense{<:Real}) = D
/(A::RocDenseTensor, x::Number) = A * inv(x)

function norm(T::RocDenseTensor{ElT}) where {ElT}
    r = Ref{Cdouble}(0)
    check(ccall((:itg_nrm2, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, Ptr{ElT}, Ref{Cdouble}),
                ctx(), dtype(ElT), dim(T), pointer(data(store(T))), r))
    return r[]
end

# -------------------------------------------------------------- truncate!
function truncate!(P::Vector{Float64}; maxdim=length(P), mindim=1, cutoff=0.0,
                   absoluteCutoff=false, doRelCutoff=true)
    k = Ref{Int64}(0); te = Ref{Cdouble}(0); dc = Ref{Cdouble}(0)
    check(ccall((:itg_truncate, libitg), Cint,
                (Ptr{Cvoid}, Ptr{Cdouble}, Int64, Int64, Int64, Cdouble, Cint, Cint, Cint,
                 Ref{Int64}, Ref{Cdouble}, Ref{Cdouble}),
                ctx(), P, length(P), maxdim, mindim, cutoff, absoluteCutoff, doRelCutoff, 0, k, te, dc))
    resize!(P, k[])
    return te[], dc[]
end

# -------------------------------------------------------- factorisations
function svd(T::RocDenseTensor{ElT,2}; maxdim=minimum(dims(T)), mindim=1, cutoff=0.0,
             absoluteCutoff=false, doRelCutoff=true, kwargs...) where {ElT}
    m, n = dims(T); r = min(m, n)
    A = RocVector{ElT}(undef, m * n)
    check(ccall((:itg_memcpy_d2d, libitg), Cint, (Ptr{Cvoid}, Ptr{ElT}, Ptr{ElT}, Csize_t),
                ctx(), A.ptr, pointer(data(store(T))), m * n * sizeof(ElT)))
    U = RocVector{ElT}(undef, m * r); V = RocVector{ElT}(undef, n * r)
    S = RocVector{Float64}(undef, r)
    k = Ref{Int64}(0); te = Ref{Cdouble}(0); dc = Ref{Cdouble}(0)
    check(ccall((:itg_svd_trunc, libitg), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{ElT}, Ptr{ElT}, Ptr{Float64}, Ptr{ElT}, Int64,
                 Int64, Cdouble, Cint, Cint, Ref{Int64}, Ref{Cdouble}, Ref{Cdouble}),
                ctx(), dtype(ElT), m, n, A.ptr, U.ptr, S.ptr, V.ptr, maxdim, mindim, cutoff,
                absoluteCutoff, doRelCutoff, k, te, dc))
    s = collect(S)[1:k[]]
    u = Index(k[], "Link,u"); v = Index(k[], "Link,v")
    # the leading k columns, in storage of exactly their length
    Ut = tensor(Dense(RocVector{ElT}(U, m * k[])), (ind(T, 1), u))
    Vt = tensor(Dense(RocVector{ElT}(V, n * k[])), (ind(T, 2), v))
    St = tensor(Dense(RocVector(vec(Matrix(Diagonal(s))))), (u, v))
    return Ut, St, Vt, Spectrum(s .^ 2, te[])
end

function eigen(T::Hermitian{ElT,<:RocDenseTensor{ElT,2}}; maxdim=minimum(dims(parent(T))),
               cutoff=0.0, absoluteCutoff=false, doRelCutoff=true, kwargs...) where {ElT}
    P = parent(T); n = dim(P, 1)
    A = RocVector{ElT}(undef, n * n)
    check(ccall((:itg_memcpy_d2d, libitg), Cint, (Ptr{Cvoid}, Ptr{ElT}, Ptr{ElT}, Csize_t),
                ctx(), A.ptr, pointer(data(store(P))), n * n * sizeof(ElT)))
    w = RocVector{Float64}(undef, n); U = RocVector{ElT}(undef, n * n)
    k = Ref{Int64}(0); te = Ref{Cdouble}(0); dc = Ref{Cdouble}(0)
    check(ccall((:itg_eigh_trunc, libitg), Cint,
                (Ptr{Cvoid}, Cint, Int64, Ptr{ElT}, Ptr{Float64}, Ptr{ElT}, Int64, Cdouble, Cint,
                 Cint, Ref{Int64}, Ref{Cdouble}, Ref{Cdouble}),
                ctx(), dtype(ElT), n, A.ptr, w.ptr, U.ptr, maxdim, cutoff, absoluteCutoff,
                doRelCutoff, k, te, dc))
    d = collect(w)[1:k[]]
    l = Index(k[], "Link,eigen"); r = Index(k[], "Link,eigen")
    Ut = tensor(Dense(RocVector{ElT}(U, n * k[])), (dag(ind(P, 2)), dag(r)))
    Dt = tensor(Diag(d), (l, dag(r)))
    return Dt, Ut, Spectrum(d, te[])
end

function qr(T::RocDenseTensor{ElT,2}; kwargs...) where {ElT}
    m, n = dims(T); r = min(m, n)
    Q = RocVector{ElT}(undef, m * r); R = RocVector{ElT}(undef, r * n)
    check(ccall((:itg_qr, libitg), Cint,
                (Ptr{Cvoid}, Cint, Int64, Int64, Ptr{ElT}, Ptr{ElT}, Ptr{ElT}),
                ctx(), dtype(ElT), m, n, pointer(data(store(T))), Q.ptr, R.ptr))
    q = sim(dim(ind(T, 1)) < dim(ind(T, 2)) ? ind(T, 1) : ind(T, 2))
    return tensor(Dense(Q), (ind(T, 1), q)), tensor(Dense(R), (q, ind(T, 2)))
end

# -------------------------------------------------- two-site ProjMPO product
# The library's index orders: L (a', wl, a), W1 (wl, wm, s1', s1),
# W2 (wm, wr, s2', s2), R (b', wr, b), psi (a, s1, s2, b). An absent
# environment (chain end) or MPO link counts as dimension 1, which leaves the
# memory layout of the others unchanged.
_dim(i) = isnothing(i) ? 1 : dim(i)
_ord(T, is...) = permute(T, filter(!isnothing, collect(is))...)
_ptr(T::ITensor) = pointer(data(store(tensor(T))))

function _heff2_operands(L, W1, W2, R, psi)
    s1 = commonind(W1, psi); s2 = commonind(W2, psi)
    a = isnothing(L) ? nothing : commonind(L, psi)
    b = isnothing(R) ? nothing : commonind(R, psi)
    wl = isnothing(L) ? nothing : commonind(L, W1)
    wm = commonind(W1, W2)
    wr = isnothing(R) ? nothing : commonind(W2, R)
    dims = Int64[_dim(a), _dim(wl), _dim(wm), _dim(wr), dim(s1), _dim(b)]
    ops = (isnothing(L) ? nothing : _ord(L, prime(a), wl, a), _ord(W1, wl, wm, prime(s1), s1),
           _ord(W2, wm, wr, prime(s2), s2), isnothing(R) ? nothing : _ord(R, prime(b), wr, b),
           _ord(psi, a, s1, s2, b))
    return dims, ops, (a, s1, s2, b)
end

function _one_env(::Type{ElT}) where {ElT}
    v = RocVector{ElT}(undef, 1)
    check(ccall((:itg_fill, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, Cdouble, Ptr{ElT}),
                ctx(), dtype(ElT), 1, 1.0, v.ptr))
    return v
end

"""
    heff2_product(L, W1, W2, R, psi; mode=0) -> ITensor

`noprime(((psi*L)*W1)*W2)*R`, the ProjMPO product dmrg runs in its
eigensolver, as one itg_heff2_apply (L or R `nothing` at a chain end).
mode: 0 FP64, 1 FP32, 2 split-BF16 contractions.
"""
function heff2_product(L, W1::ITensor, W2::ITensor, R, psi::ITensor; mode::Integer=0)
    ElT = eltype(tensor(psi))
    dims, (Lp, W1p, W2p, Rp, pp), (a, s1, s2, b) = _heff2_operands(L, W1, W2, R, psi)
    out = rocITensor(ElT, filter(!isnothing, [a, s1, s2, b])...)
    work = RocVector{Float32}(undef, ccall((:itg_heff2_work_elems, libitg), Int64,
                                           (Int64, Int64, Int64, Int64, Int64, Int64), dims...))
    one = _one_env(ElT)
    Lptr = isnothing(Lp) ? Ptr{ElT}(one.ptr) : _ptr(Lp)
    Rptr = isnothing(Rp) ? Ptr{ElT}(one.ptr) : _ptr(Rp)
    GC.@preserve Lp W1p W2p Rp pp out work one begin
        check(ccall((:itg_heff2_apply, libitg), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Int64, Int64, Int64, Int64, Ptr{ElT},
                     Ptr{ElT}, Ptr{ElT}, Ptr{ElT}, Ptr{ElT}, Ptr{ElT}, Ptr{Float32}),
                    ctx(), dtype(ElT), mode, dims..., Lptr, _ptr(W1p), _ptr(W2p), Rptr,
                    _ptr(pp), _ptr(out), work.ptr))
    end
    return out
end

"""
    roc_product(P::ProjMPO, v::ITensor; mode=0)

The two-site ProjMPO product through heff2_product, reading the
environments and MPO tensors the ProjMPO holds at its position.
"""
function roc_product(P::ProjMPO, v::ITensor; mode::Integer=0)
    nsite(P) == 2 || throw(ArgumentError("roc_product: two-site ProjMPO only"))
    L = ITensors.lproj(P); R = ITensors.rproj(P)
    return heff2_product(L, P.H[P.lpos + 1], P.H[P.lpos + 2], R, v; mode=mode)
end

"""
    roc_davidson(L, W1, W2, R, psi; maxiter=2, errgoal=1e-14, mode=0) -> (λ, ϕ)

ITensors' davidson for the two-site operator on one bond, run in the
library (itg_heff2_prepare + itg_davidson2): FP64 Krylov basis on the
device, split-BF16 applies replayed from HIP graphs when mode = 2.
"""
function roc_davidson(L, W1::ITensor, W2::ITensor, R, psi::ITensor; maxiter::Integer=2,
                      errgoal::Real=1e-14, mode::Integer=0)
    ElT = eltype(tensor(psi))
    ElT === Float64 || throw(ArgumentError("roc_davidson: Float64 vectors only"))
    dims, (Lp, W1p, W2p, Rp, pp), (a, s1, s2, b) = _heff2_operands(L, W1, W2, R, psi)
    phi = ITensor(Dense(RocVector{ElT}(data(store(tensor(pp))), dim(tensor(pp)))), inds(pp))
    one = _one_env(ElT)
    Lptr = isnothing(Lp) ? Ptr{ElT}(one.ptr) : _ptr(Lp)
    Rptr = isnothing(Rp) ? Ptr{ElT}(one.ptr) : _ptr(Rp)
    prep = RocVector{Float32}(undef, ccall((:itg_heff2_prep_elems, libitg), Int64,
                                           (Int64, Int64, Int64, Int64, Int64, Int64), dims...))
    work = RocVector{Float32}(undef, ccall((:itg_heff2_work_elems, libitg), Int64,
                                           (Int64, Int64, Int64, Int64, Int64, Int64), dims...))
    kwork = RocVector{Float64}(undef, ccall((:itg_davidson2_work_elems, libitg), Int64,
                                            (Int64, Int64, Int64, Cint), dims[1], dims[5],
                                            dims[6], maxiter))
    λ = Ref{Cdouble}(0); nmv = Ref{Cint}(0)
    GC.@preserve Lp W1p W2p Rp phi prep work kwork one begin
        check(ccall((:itg_heff2_prepare, libitg), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Int64, Int64, Int64, Int64, Ptr{ElT},
                     Ptr{ElT}, Ptr{ElT}, Ptr{ElT}, Ptr{Float32}),
                    ctx(), dtype(ElT), mode, dims..., Lptr, _ptr(W1p), _ptr(W2p), Rptr, prep.ptr))
        check(ccall((:itg_davidson2, libitg), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Int64, Int64, Int64, Int64, Int64, Int64, Ptr{ElT},
                     Ptr{ElT}, Ptr{Float32}, Ptr{ElT}, Cint, Cdouble, Ptr{ElT}, Ptr{Float32},
                     Ptr{Float64}, Ref{Cdouble}, Ref{Cint}),
                    ctx(), dtype(ElT), mode, dims..., Lptr, Rptr, prep.ptr, _ptr(phi), maxiter,
                    errgoal, C_NULL, work.ptr, kwork.ptr, λ, nmv))
    end
    return λ[], phi
end

export roc, cpu, RocVector, RocDense, seed!, randomRocITensor, rocITensor, rocMPS,
       randomRocMPS, heff2_product, roc_product, roc_davidson

end # module
