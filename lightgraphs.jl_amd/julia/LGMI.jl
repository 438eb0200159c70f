# LGMI.jl -- LightGraphs.jl drop-in for gdistances / bfs_parents /
# betweenness_centrality / stress_centrality / closeness_centrality /
# radiality_centrality / eccentricity and the measures built on it (serial
# and Parallel) on an MI355X through liblgmi.so (C ABI: include/lgmi.h).
#
# Loading this module adds methods that are more specific than the
# reference's (same names, same arguments, same results) for
# SimpleGraph{T}/SimpleDiGraph{T} with the default unit distances; every
# other case (weighted distmx, other graph types, more than 2^31-1 adjacency
# entries, no GPU) falls through to the reference method.
#
# This file is not executed in this repository's test suite (no Julia
# toolchain is installed); the Python package lgmi/ mirrors it call for call
# and is what the tests drive.
module LGMI

using LightGraphs
using LightGraphs: AbstractGraph, SimpleGraph, SimpleDiGraph, DefaultDistance, nv, ne,
    is_directed, vertices, degree, sample
import LightGraphs: gdistances, gdistances!, bfs_parents, bfs_tree, betweenness_centrality, stress_centrality,
    closeness_centrality, radiality_centrality, eccentricity, diameter, radius, center, periphery,
    connected_components, connected_components!, is_connected, has_path, period, neighborhood, neighborhood_dists
import LightGraphs.Parallel

const liblgmi = joinpath(@__DIR__, "..", "lgmi", "liblgmi.so")

const LGMI_OK = Cint(0)
const LGMI_E_ARG = Cint(-1)
const LGMI_E_BOUNDS = Cint(-2)
const LGMI_E_TOO_LARGE = Cint(-3)
const LGMI_E_NODEV = Cint(-4)

const SimpleG{T} = Union{SimpleGraph{T},SimpleDiGraph{T}}

struct Fallback <: Exception end

lasterror() = unsafe_string(ccall((:lgmi_last_error, liblgmi), Cstring, ()))

function check(rc::Cint)
    rc == LGMI_OK && return nothing
    (rc == LGMI_E_TOO_LARGE || rc == LGMI_E_NODEV) && throw(Fallback())
    rc == LGMI_E_BOUNDS && throw(BoundsError(lasterror()))
    rc == LGMI_E_ARG && throw(ArgumentError(lasterror()))
    error("liblgmi: ", lasterror())
end

# fadjlist / badjlist -> 0-based Int32 CSR (rows are already sorted,
# src/SimpleGraphs/simplegraph.jl:445-461)
function csr(adj::Vector{Vector{T}}) where T
    n = length(adj)
    rp = Vector{Int32}(undef, n + 1)
    rp[1] = 0
    @inbounds for v in 1:n
        rp[v+1] = rp[v] + Int32(length(adj[v]))
    end
    col = Vector{Int32}(undef, rp[end])
    @inbounds for v in 1:n, (k, u) in enumerate(adj[v])
        col[rp[v]+k] = Int32(u - 1)
    end
    return rp, col
end

mutable struct Handle
    ptr::Ptr{Cvoid}
end

function new_handle(g::SimpleG)
    length(g.fadjlist) > 0 && sum(length, g.fadjlist) >= typemax(Int32) && throw(Fallback())
    rp, col = csr(g.fadjlist)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    if is_directed(g)
        irp, icol = csr(g.badjlist)
        rc = ccall((:lgmi_graph_create, liblgmi), Cint,
                   (Int32, Int64, Ptr{Int32}, Ptr{Int32}, Int64, Ptr{Int32}, Ptr{Int32}, Cint, Ptr{Ptr{Cvoid}}),
                   nv(g), length(col), rp, col, length(icol), irp, icol, 1, out)
    else
        rc = ccall((:lgmi_graph_create, liblgmi), Cint,
                   (Int32, Int64, Ptr{Int32}, Ptr{Int32}, Int64, Ptr{Int32}, Ptr{Int32}, Cint, Ptr{Ptr{Cvoid}}),
                   nv(g), length(col), rp, col, 0, C_NULL, C_NULL, 0, out)
    end
    check(rc)
    h = Handle(out[])
    finalizer(h -> ccall((:lgmi_graph_destroy, liblgmi), Cvoid, (Ptr{Cvoid},), h.ptr), h)
    return h
end

# Opt-in handle cache.  By default every call converts fadjlist and uploads
# the graph (a LightGraphs graph is mutable, add_edge! / rem_edge!,
# src/SimpleGraphs/simplegraph.jl:445-461).  cache!(g) keeps g's device
# handle for later calls on g; the entry is dropped when nv(g) or ne(g)
# changes, and uncache!(g) drops it after a mutation that keeps both (for
# example rem_edge! followed by add_edge!).
const HANDLES = WeakKeyDict{Any,Tuple{Int,Int,Handle}}()
const HANDLES_LOCK = ReentrantLock()

function cache!(g::SimpleG)
    h = new_handle(g)
    lock(HANDLES_LOCK) do
        HANDLES[g] = (nv(g), ne(g), h)
    end
    return g
end

uncache!(g::SimpleG) = lock(() -> (delete!(HANDLES, g); g), HANDLES_LOCK)

function handle(g::SimpleG)
    c = lock(() -> get(HANDLES, g, nothing), HANDLES_LOCK)
    if c !== nothing
        c[1] == nv(g) && c[2] == ne(g) && return c[3]
        uncache!(g)
    end
    return new_handle(g)
end

# Launch thresholds chosen by the autotuner (lgmi/tuning.json, written by
# lgmi/tuning.py), applied when the module loads.  Layout: lgmi_tuning of
# include/lgmi.h.
struct Tuning
    wave_row::Int32
    lane_rounds::Int32
    range_front::Int32
    queue_front::Int32
    range_light::Int64
    queue_min_edges::Int64
    queue_max_edges::Int64
    bc_seg_len::Int32
    solo_front::Int32
    solo_edges::Int64
end

const TUNING_FILE = joinpath(@__DIR__, "..", "lgmi", "tuning.json")

# The file's "best" object is flat: "name": integer.
function load_tuning(path::AbstractString = TUNING_FILE)
    isfile(path) || return false
    txt = read(path, String)
    i = findfirst("\"best\"", txt)
    i === nothing && return false
    j = findnext('}', txt, last(i))
    vals = Dict(m.captures[1] => parse(Int64, m.captures[2])
                for m in eachmatch(r"\"(\w+)\"\s*:\s*(-?\d+)", txt[last(i):j]))
    cur = Ref{Tuning}()
    check(ccall((:lgmi_get_tuning, liblgmi), Cint, (Ptr{Tuning},), cur))
    t = cur[]
    f(name, dflt, T) = T(get(vals, name, dflt))
    new = Tuning(f("wave_row", t.wave_row, Int32), f("lane_rounds", t.lane_rounds, Int32),
                 f("range_front", t.range_front, Int32), f("queue_front", t.queue_front, Int32),
                 f("range_light", t.range_light, Int64), f("queue_min_edges", t.queue_min_edges, Int64),
                 f("queue_max_edges", t.queue_max_edges, Int64), f("bc_seg_len", t.bc_seg_len, Int32),
                 f("solo_front", t.solo_front, Int32), f("solo_edges", t.solo_edges, Int64))
    check(ccall((:lgmi_set_tuning, liblgmi), Cint, (Ref{Tuning},), new))
    return true
end

function __init__()
    try
        load_tuning()
    catch err
        @warn "LGMI: tuning.json not applied" exception = err
    end
end

# ---------------------------------------------------------------------
# gdistances! / gdistances  (src/traversals/bfs.jl:95-141)
# ---------------------------------------------------------------------
function device_dist(g::SimpleG, source)
    srcs = Int32[s - 1 for s in source]
    for s in source
        1 <= s <= nv(g) || throw(BoundsError(vertices(g), s))
    end
    h = handle(g)
    dist = Vector{Int32}(undef, nv(g))
    check(ccall((:lgmi_gdistances, liblgmi), Cint, (Ptr{Cvoid}, Ptr{Int32}, Int32, Ptr{Int32}),
                h.ptr, srcs, length(srcs), dist))
    return dist
end

function gdistances!(g::SimpleG{T}, source, vert_level; sort_alg = Base.Sort.QuickSort) where T
    dist = try
        device_dist(g, source)
    catch e
        e isa Fallback || rethrow()
        return invoke(gdistances!, Tuple{AbstractGraph{T},Any,Any}, g, source, vert_level; sort_alg = sort_alg)
    end
    @inbounds for v in eachindex(dist)
        dist[v] >= 0 && (vert_level[v] = T(dist[v]))    # reached entries only
    end
    return vert_level
end

gdistances(g::SimpleG{T}, source; sort_alg = Base.Sort.QuickSort) where T =
    gdistances!(g, source, fill(typemax(T), nv(g)); sort_alg = sort_alg)

# ---------------------------------------------------------------------
# bfs_parents / bfs_tree  (src/traversals/bfs.jl:35-76)
# ---------------------------------------------------------------------
function bfs_parents(g::SimpleG{T}, s::Integer; dir = :out) where T
    1 <= s <= nv(g) || throw(BoundsError(vertices(g), s))
    par = Vector{Int32}(undef, nv(g))
    try
        h = handle(g)
        check(ccall((:lgmi_bfs_parents, liblgmi), Cint, (Ptr{Cvoid}, Int32, Cint, Ptr{Int32}),
                    h.ptr, s - 1, dir == :out ? 0 : 1, par))
    catch e
        e isa Fallback || rethrow()
        return invoke(bfs_parents, Tuple{AbstractGraph,Integer}, g, s; dir = dir)
    end
    return T[p < 0 ? zero(T) : T(p + 1) for p in par]
end

# ---------------------------------------------------------------------
# betweenness_centrality  (src/centrality/betweenness.jl:46-78)
# ---------------------------------------------------------------------
function device_bc(g::SimpleG, srcs::AbstractVector, k::Integer, normalize::Bool, endpoints::Bool;
                   skip_degree0::Bool = true)
    for s in srcs
        1 <= s <= nv(g) || throw(BoundsError(vertices(g), s))
    end
    h = handle(g)
    bc = Vector{Float64}(undef, nv(g))
    s0 = Int32[s - 1 for s in srcs]
    check(ccall((:lgmi_betweenness, liblgmi), Cint,
                (Ptr{Cvoid}, Ptr{Int32}, Int32, Int64, Cint, Cint, Cint, Ptr{Float64}),
                h.ptr, s0, length(s0), k, skip_degree0, normalize, endpoints, bc))
    return bc
end

function betweenness_centrality(g::SimpleG, vs::AbstractVector = vertices(g), distmx::DefaultDistance = weights(g);
                                normalize = true, endpoints = false)
    try
        return device_bc(g, vs, length(vs), normalize, endpoints)
    catch e
        e isa Fallback || rethrow()
        return invoke(betweenness_centrality, Tuple{AbstractGraph,Any,AbstractMatrix}, g, vs, distmx;
                      normalize = normalize, endpoints = endpoints)
    end
end

betweenness_centrality(g::SimpleG, k::Integer, distmx::DefaultDistance = weights(g); normalize = true,
                       endpoints = false) =
    betweenness_centrality(g, sample(vertices(g), k), distmx; normalize = normalize, endpoints = endpoints)

# ---------------------------------------------------------------------
# stress_centrality  (src/centrality/stress.jl:33-51)
# ---------------------------------------------------------------------
function device_stress(g::SimpleG, srcs::AbstractVector)
    for s in srcs
        1 <= s <= nv(g) || throw(BoundsError(vertices(g), s))
    end
    h = handle(g)
    out = Vector{Int64}(undef, nv(g))
    s0 = Int32[s - 1 for s in srcs]
    check(ccall((:lgmi_stress, liblgmi), Cint, (Ptr{Cvoid}, Ptr{Int32}, Int32, Ptr{Int64}),
                h.ptr, s0, length(s0), out))
    return out
end

function stress_centrality(g::SimpleG, vs::AbstractVector = vertices(g))
    try
        return device_stress(g, vs)
    catch e
        e isa Fallback || rethrow()
        return invoke(stress_centrality, Tuple{AbstractGraph,Any}, g, vs)
    end
end
stress_centrality(g::SimpleG, k::Integer) = stress_centrality(g, sample(vertices(g), k))

# ---------------------------------------------------------------------
# closeness / radiality / eccentricity  (src/centrality/closeness.jl:32-57,
# src/centrality/radiality.jl:34-47, src/distance.jl:74-174).  With unit
# lengths each needs, per source, only (sum of finite distances, vertices
# reached, largest finite distance); an unreached vertex is typemax(Int) in
# the reference's distance vector and enters its Int sums the same way here.
# ---------------------------------------------------------------------
function device_dstats(g::SimpleG, srcs::AbstractVector)
    for s in srcs
        1 <= s <= nv(g) || throw(BoundsError(vertices(g), s))
    end
    h = handle(g)
    st = Matrix{Int64}(undef, 3, length(srcs))
    s0 = Int32[s - 1 for s in srcs]
    isempty(s0) || check(ccall((:lgmi_distance_stats, liblgmi), Cint, (Ptr{Cvoid}, Ptr{Int32}, Int32, Ptr{Int64}),
                               h.ptr, s0, length(s0), st))
    return st
end
ecc_of(st, n) = [st[2, i] == n ? st[3, i] : typemax(Int) for i in axes(st, 2)]
wrapped_sum(st, n) = [st[1, i] + (n - st[2, i]) * typemax(Int) for i in axes(st, 2)]   # Int wraps

function eccentricity(g::SimpleG, v::Integer, distmx::DefaultDistance = weights(g))
    e = ecc_of(device_dstats(g, [v]), nv(g))[1]
    e == typemax(Int) && @warn("Infinite path length detected for vertex $v")
    return e
end
eccentricity(g::SimpleG, vs::AbstractVector = vertices(g), distmx::DefaultDistance = weights(g)) =
    ecc_of(device_dstats(g, vs), nv(g))
diameter(g::SimpleG, distmx::DefaultDistance = weights(g)) = maximum(eccentricity(g, vertices(g), distmx))
radius(g::SimpleG, distmx::DefaultDistance = weights(g)) = minimum(eccentricity(g, vertices(g), distmx))
periphery(g::SimpleG, distmx::DefaultDistance = weights(g)) = periphery(eccentricity(g, vertices(g), distmx))
center(g::SimpleG, distmx::DefaultDistance = weights(g)) = center(eccentricity(g, vertices(g), distmx))

function closeness_centrality(g::SimpleG, distmx::DefaultDistance = weights(g); normalize = true)
    n_v = nv(g)
    closeness = zeros(n_v)
    act = [u for u in vertices(g) if degree(g, u) != 0]
    st = device_dstats(g, act)
    for (i, u) in enumerate(act)
        σ = st[1, i]
        l = st[2, i] - 1
        if σ > 0
            closeness[u] = l / σ
            if normalize
                n = l / (n_v - 1)
                closeness[u] *= n
            end
        end
    end
    return closeness
end

function radiality_centrality(g::SimpleG)::Vector{Float64}
    n_v = nv(g)
    st = device_dstats(g, vertices(g))
    dmtr = max(0.0, maximum(ecc_of(st, n_v)))
    meandists = wrapped_sum(st, n_v) ./ (n_v - 1)
    return ((dmtr + 1) .- meandists) ./ dmtr
end

# ---------------------------------------------------------------------
# device-side ingest: SimpleGraph(edge_list) / SimpleDiGraph(edge_list)
# (src/SimpleGraphs/simplegraph.jl:208-244, simpledigraph.jl:212-253).
# The CSR is sorted and de-duplicated on the GPU; the handle is kept for the
# calls above and the fadjlist is filled from the CSR it returns.
# ---------------------------------------------------------------------
function device_graph_ids(::Type{T}, s0::Vector{Int32}, d0::Vector{Int32}, directed::Bool, nv_min::Integer) where T
    ptr = Ref{Ptr{Cvoid}}(C_NULL)
    ne_ = Ref{Int64}(0)
    check(ccall((:lgmi_graph_from_edges, liblgmi), Cint,
                (Ptr{Int32}, Ptr{Int32}, Int64, Cint, Int32, Ptr{Ptr{Cvoid}}, Ptr{Int64}),
                s0, d0, length(s0), directed, nv_min, ptr, ne_))
    n = Ref{Int32}(0); nnz = Ref{Int64}(0)
    check(ccall((:lgmi_graph_size, liblgmi), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int64}), ptr[], n, nnz))
    rp = Vector{Int32}(undef, n[] + 1); col = Vector{Int32}(undef, nnz[])
    irp = Vector{Int32}(undef, n[] + 1); icol = Vector{Int32}(undef, nnz[])
    check(ccall((:lgmi_graph_csr, liblgmi), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}),
                ptr[], rp, col, irp, icol))
    ccall((:lgmi_graph_destroy, liblgmi), Cvoid, (Ptr{Cvoid},), ptr[])
    rows(p, c) = [T[c[k] + 1 for k in (p[v] + 1):p[v + 1]] for v in 1:n[]]
    return directed ? SimpleDiGraph{T}(ne_[], rows(rp, col), rows(irp, icol)) : SimpleGraph{T}(ne_[], rows(rp, col))
end

function device_graph(edge_list::Vector{<:LightGraphs.AbstractEdge{T}}, directed::Bool) where T
    s0 = Int32[src(e) - 1 for e in edge_list]
    d0 = Int32[dst(e) - 1 for e in edge_list]
    return device_graph_ids(T, s0, d0, directed, 0)
end
device_simplegraph(edge_list) = device_graph(edge_list, false)
device_simpledigraph(edge_list) = device_graph(edge_list, true)

# ---------------------------------------------------------------------
# LG files (src/persistence/lg.jl:35-112): headers parsed by the
# reference's _parse_header, edge lines by lgmi_lg_parse_edges (host C++),
# each graph built on the device from its in-range pairs (add_edge! drops
# the others, ignores duplicates and keeps self-loops).
# ---------------------------------------------------------------------
function lg_edges_device(text::Vector{UInt8}, pos::Int, h)
    T = h.dtype
    cap = max(h.ne, 1)
    s = Vector{Int64}(undef, cap); d = Vector{Int64}(undef, cap)
    np_ = Ref{Int64}(0); used = Ref{Int64}(0); bad = Ref{Int64}(0)
    avail = length(text) - pos + 1
    rc = avail <= 0 ? Cint(0) : GC.@preserve text ccall((:lgmi_lg_parse_edges, liblgmi), Cint,
        (Ptr{UInt8}, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}),
        pointer(text, pos), avail, h.ne, s, d, np_, used, bad)
    rc == -3 && throw(OverflowError("LG edge line $(bad[]) of graph $(h.name)"))
    rc == 0 || throw(ArgumentError("LG edge line $(bad[]) of graph $(h.name): expected src,dst"))
    s0 = Int32[]; d0 = Int32[]
    for i in 1:np_[]
        (typemin(T) <= s[i] <= typemax(T) && typemin(T) <= d[i] <= typemax(T)) ||
            throw(OverflowError("LG edge line of graph $(h.name) does not fit $T"))
        if 1 <= s[i] <= h.nv && 1 <= d[i] <= h.nv
            push!(s0, Int32(s[i] - 1)); push!(d0, Int32(d[i] - 1))
        end
    end
    return device_graph_ids(T, s0, d0, h.is_directed, h.nv), pos + used[]
end

function lg_next_line(text::Vector{UInt8}, pos::Int)
    e = something(findnext(==(UInt8('\n')), text, pos), length(text) + 1)
    return strip(String(text[pos:e-1])), e + 1
end

"""
    loadgraphs_device(file) / loadgraph_device(file, gname="graph")

loadgraphs / loadgraph (src/persistence/common.jl:3-39, lg.jl:87-112) for
LG files, with each graph's adjacency sorted and de-duplicated on the GPU.
"""
function loadgraphs_device(fn::AbstractString)
    text = open(fn, "r") do io read(LightGraphs.auto_decompress(io)) end
    graphs = Dict{String,AbstractGraph}()
    pos = 1
    while pos <= length(text)
        line, pos = lg_next_line(text, pos)
        (startswith(line, "#") || line == "") && continue
        h = LightGraphs._parse_header(line)
        g, pos = lg_edges_device(text, pos, h)
        graphs[h.name] = g
    end
    return graphs
end

function loadgraph_device(fn::AbstractString, gname::AbstractString="graph")
    text = open(fn, "r") do io read(LightGraphs.auto_decompress(io)) end
    pos = 1
    while pos <= length(text)
        line, pos = lg_next_line(text, pos)
        (startswith(line, "#") || line == "") && continue
        h = LightGraphs._parse_header(line)
        h.name == gname && return first(lg_edges_device(text, pos, h))
        for _ in 1:h.ne                        # _lg_skip_one_graph
            pos > length(text) && break
            _, pos = lg_next_line(text, pos)
        end
    end
    throw(ArgumentError("graph $gname not found"))
end

# ---------------------------------------------------------------------
# connectivity  (src/connectivity.jl:13-134, 495-516; has_path
# src/traversals/bfs.jl:150-172)
# ---------------------------------------------------------------------
function connected_components!(label::AbstractVector, g::SimpleG{T}) where T
    lab = Vector{Int32}(undef, nv(g))
    nv(g) > 0 && check(ccall((:lgmi_connected_components, liblgmi), Cint, (Ptr{Cvoid}, Ptr{Int32}),
                             handle(g).ptr, lab))
    for v in eachindex(lab)
        label[v] = lab[v] + 1          # the smallest vertex id of v's component
    end
    return label
end

function connected_components(g::SimpleG{T}) where T
    label = zeros(T, nv(g))
    connected_components!(label, g)
    c, d = LightGraphs.components(label)
    return c
end

is_connected(g::SimpleG) = (is_directed(g) ? 2 : 1) * ne(g) + 1 >= nv(g) && length(connected_components(g)) == 1

function has_path(g::SimpleG{T}, u::Integer, v::Integer; exclude_vertices::AbstractVector = Vector{T}()) where T
    for x in (u, v)
        1 <= x <= nv(g) || throw(BoundsError(vertices(g), x))
    end
    ex = Int32[x - 1 for x in exclude_vertices]
    res = Ref{Cint}(0)
    check(ccall((:lgmi_has_path, liblgmi), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Int32}, Int32, Ptr{Cint}),
                handle(g).ptr, u - 1, v - 1, ex, length(ex), res))
    return res[] != 0
end

function neighborhood_dists(g::SimpleG{T}, v::Integer, d, distmx::DefaultDistance = weights(g); dir = :out) where T
    1 <= v <= nv(g) || throw(BoundsError(vertices(g), v))
    d < 0 && return Tuple{T,Int}[]
    dd = isinf(d) ? nv(g) : min(floor(Int64, d), nv(g))
    verts = Vector{Int32}(undef, nv(g))
    dists = Vector{Int32}(undef, nv(g))
    cnt = Ref{Int32}(0)
    check(ccall((:lgmi_neighborhood, liblgmi), Cint,
                (Ptr{Cvoid}, Int32, Int64, Cint, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}),
                handle(g).ptr, v - 1, dd, dir == :out ? 0 : 1, verts, dists, cnt))
    return [(T(verts[i] + 1), Int(dists[i])) for i in 1:cnt[]]     # FIFO queue order, as _neighborhood
end
neighborhood(g::SimpleG, v::Integer, d, distmx::DefaultDistance = weights(g); dir = :out) =
    first.(neighborhood_dists(g, v, d, distmx; dir = dir))

function period(g::SimpleDiGraph)
    p = Ref{Int64}(0)
    check(ccall((:lgmi_period, liblgmi), Cint, (Ptr{Cvoid}, Ptr{Int64}), handle(g).ptr, p))   # ArgumentError if not strongly connected
    return p[]
end

# ---------------------------------------------------------------------
# Parallel.*  (src/Parallel/traversals/{gdistances,bfs}.jl,
#              src/Parallel/centrality/{betweenness,stress,closeness,
#              radiality}.jl, src/Parallel/distance.jl)
# ---------------------------------------------------------------------
function Parallel.gdistances!(g::SimpleG{T}, sources::Vector{<:Integer}, vert_level::Vector{T};
                              queue_segment_size::Integer = 20) where T
    fill!(vert_level, typemax(T))                      # gdistances.jl:46
    return gdistances!(g, sources, vert_level)
end
Parallel.gdistances(g::SimpleG{T}, sources::Vector{<:Integer}; queue_segment_size::Integer = 20) where T =
    Parallel.gdistances!(g, sources, Vector{T}(undef, nv(g)))
Parallel.gdistances(g::SimpleG{T}, source::Integer; queue_segment_size::Integer = 20) where T =
    Parallel.gdistances!(g, [source], Vector{T}(undef, nv(g)))

function Parallel.bfs_tree!(next::Parallel.ThreadQueue, g::SimpleG{T}, source::T,
                            parents::Array{Threads.Atomic{T}}) where T
    par = bfs_parents(g, source)
    for v in eachindex(par)
        parents[v][] = par[v]
    end
    return parents
end
Parallel.bfs_tree(g::SimpleG{T}, source::T) where T = LightGraphs.tree(bfs_parents(g, source))

function Parallel.betweenness_centrality(g::SimpleG, vs::AbstractVector = vertices(g),
                                         distmx::DefaultDistance = weights(g);
                                         normalize = true, endpoints = false, parallel = :distributed)
    k = length(vs)
    # :threads takes findall's positions as sources
    # (Parallel/centrality/betweenness.jl:53-56)
    parallel == :distributed && return device_bc(g, vs, k, normalize, endpoints)
    return device_bc(g, findall(x -> degree(g, x) > 0, vs), k, normalize, endpoints; skip_degree0 = false)
end
Parallel.betweenness_centrality(g::SimpleG, k::Integer, distmx::DefaultDistance = weights(g);
                                normalize = true, endpoints = false, parallel = :distributed) =
    Parallel.betweenness_centrality(g, sample(vertices(g), k), distmx; normalize = normalize,
                                    endpoints = endpoints, parallel = parallel)

# both modes sum the per-source Int64 vectors of the sources with degree > 0
# (Parallel/centrality/stress.jl:9-51), i.e. the serial result
Parallel.stress_centrality(g::SimpleG, vs::AbstractVector = vertices(g); parallel = :distributed) =
    stress_centrality(g, vs)
Parallel.stress_centrality(g::SimpleG, k::Integer; parallel = :distributed) =
    stress_centrality(g, sample(vertices(g), k))

Parallel.closeness_centrality(g::SimpleG, distmx::DefaultDistance = weights(g); normalize = true,
                              parallel = :distributed) = closeness_centrality(g, distmx; normalize = normalize)
Parallel.radiality_centrality(g::SimpleG; parallel = :distributed) = radiality_centrality(g)
Parallel.eccentricity(g::SimpleG, vs::AbstractVector = vertices(g), distmx::DefaultDistance = weights(g)) =
    eccentricity(g, vs, distmx)
Parallel.diameter(g::SimpleG, distmx::DefaultDistance = weights(g)) = diameter(g, distmx)
Parallel.radius(g::SimpleG, distmx::DefaultDistance = weights(g)) = radius(g, distmx)
Parallel.periphery(g::SimpleG, distmx::DefaultDistance = weights(g)) = periphery(g, distmx)
Parallel.center(g::SimpleG, distmx::DefaultDistance = weights(g)) = center(g, distmx)

end # module
