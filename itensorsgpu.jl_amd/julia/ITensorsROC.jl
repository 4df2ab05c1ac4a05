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
Base.getindex(v::RocVector, i::Int) = error("scalar indexing of device storage; use cpu()")
Base.pointer(v::RocVector) = v.ptr

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

# ---------------------------------------------------------------- conversion
roc(A::ITensor) = ITensor(Dense(RocVector(Vector(data(store(A))))), inds(A))
cpu(A::ITensor) = storage(A) isa RocDense ? ITensor(Dense(collect(data(storage(A)))), inds(A)) : A
roc(ψ::MPS) = (ϕ = copy(ψ); for j in eachindex(ϕ) ϕ[j] = roc(ψ[j]) end; ϕ)
roc(H::MPO) = (G = copy(H); for j in eachindex(G) G[j] = roc(H[j]) end; G)
cpu(ψ::Union{MPS,MPO}) = (ϕ = copy(ψ); for j in eachindex(ϕ) ϕ[j] = cpu(ψ[j]) end; ϕ)

function Base.similar(::Type{<:RocDenseTensor{ElT}}, inds) where {ElT}
    return tensor(Dense(RocVector{ElT}(undef, dim(inds))), inds)
end

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
    check(ccall((:itg_scal, libitg), Cint, (Ptr{Cvoid}, Cint, Int64, Cdouble, Ptr{ElT}),
                ctx(), dtype(ElT), length(v), x, v.ptr))
    return Dense(v)
end
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
    Ut = tensor(Dense(U), (ind(T, 1), u))   # leading k columns are used
    Vt = tensor(Dense(V), (ind(T, 2), v))
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
    Ut = tensor(Dense(U), (dag(ind(P, 2)), dag(r)))
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

export roc, cpu, RocVector, RocDense

end # module
