# madx.jl -- MAD-X lattice files for TrackPad's element types: read_madx.
#
# The reference has no lattice reader: its migration plan lists a MAD-X parser as future work
# (MIGRATION_PLAN.md:87-97) and lattices are built from element constructors
# (test/verify_dba.jl:20-60).  This reads the MAD-X subset that describes such lattices, with
# the rules and conversions of the Python reader (trackpad_hip/madx.py), so that a file gives
# the same elements in either language:
#
# * statements end with `;`; comments are `//` or `!` to the end of the line and `/* ... */`;
#   names and keywords are case-insensitive;
# * variables `name = expr;` (evaluated at once) and `name := expr;` (deferred); `const` / `real`
#   / `int` prefixes; expressions with + - * / ^ (or **), parentheses, `elem->attr`, the
#   constants pi, twopi, degrad, raddeg, e, clight, emass, pmass and the functions sqrt, exp,
#   log, log10, sin, cos, tan, asin, acos, atan, sinh, cosh, tanh, abs, floor, ceil, round;
# * element classes DRIFT, QUADRUPOLE, SEXTUPOLE, OCTUPOLE, SBEND, RBEND, MULTIPOLE, RFCAVITY,
#   MARKER (MONITOR, HMONITOR, VMONITOR, INSTRUMENT, PLACEHOLDER as markers or drifts), HKICKER,
#   VKICKER, KICKER, SOLENOID; `label: parent, attrs;` derives an element; `elem->attr = expr;`
#   and `elem, attr = expr;` change one;
# * `label: LINE = (a, b, 2*c, -d, (e, f));`, `label: SEQUENCE, l = L, refer = entry | centre |
#   exit; elem, at = s [, from = other]; ... ENDSEQUENCE;` (gaps become drifts), BEAM, USE,
#   CALL; other commands (option, title, twiss, ...) are ignored.
#
# Conversions: thick multipoles -> Quadrupole / Sextupole / Octupole with num_int_steps = nst
# (default 10); SBEND -> the bend model (SBend, ExactSBend or ExactSBendLinearEdge: the `bend`
# keyword or a `! trackpad: bend = ...` comment) with fint1 = fint, fint2 = fintx (default
# fint), gap = 2 hgap, k1 and k2/2 in polynom_b (ExactSBend: gk = fint hgap); RBEND -> the same
# with e1, e2 + angle/2 and the chord length l turned into the arc l (angle/2) / sin(angle/2);
# MULTIPOLE knl, ksl -> ThinMultipole with polynom_b[n] = knl[n]/n!, polynom_a[n] = ksl[n]/n!
# (orders above 3 through WithMultipoles); RFCAVITY volt [MV], freq [MHz] (or harmon: f =
# harmon beta c / C of the used line), lag [2 pi] -> RFCavity(l, volt, freq, 0; h = harmon,
# philag = 2 pi lag, energy = the beam's kinetic energy); kickers -> Corrector; SOLENOID ->
# Solenoid; BEAM (total energy in GeV) -> Beam(E - m) (TrackPad's Beam takes the kinetic
# energy in eV, src/lattice.jl:50-54).  Skew thick multipoles and other classes are refused
# with MadxError, so a lattice is never changed silently.

struct MadxError <: Exception
    msg::String
end
Base.showerror(io::IO, e::MadxError) = print(io, "MadxError: ", e.msg)

const MADX_CONST = Dict("pi" => pi * 1.0, "twopi" => 2.0 * pi, "degrad" => 180.0 / pi, "raddeg" => pi / 180.0,
                        "e" => Float64(ℯ), "clight" => C_LIGHT, "emass" => TrackPad.M_ELECTRON * 1e-9,
                        "pmass" => TrackPad.M_PROTON * 1e-9)
const MADX_FUNCS = Dict{String, Function}("sqrt" => sqrt, "exp" => exp, "log" => log, "log10" => log10,
                                          "sin" => sin, "cos" => cos, "tan" => tan, "asin" => asin,
                                          "acos" => acos, "atan" => atan, "sinh" => sinh, "cosh" => cosh,
                                          "tanh" => tanh, "abs" => abs, "floor" => floor, "ceil" => ceil,
                                          "round" => x -> round(x, RoundNearest))
const MADX_CLASSES = Set(["drift", "quadrupole", "sextupole", "octupole", "sbend", "rbend", "multipole",
                          "rfcavity", "marker", "monitor", "hmonitor", "vmonitor", "instrument", "placeholder",
                          "hkicker", "vkicker", "kicker", "solenoid"])
const MADX_IGNORED = Set(["option", "title", "set", "show", "value", "twiss", "select", "print", "system",
                          "assign", "exit", "stop", "return", "survey", "emit", "ptc_create_universe", "ptc_end",
                          "esave", "save", "seqedit", "flatten", "endedit", "help", "echo", "removefile",
                          "resbeam", "savebeta", "create", "write", "fill", "setvars", "setvars_lin",
                          "readtable", "match", "endmatch", "vary", "constraint", "lmdif", "simplex"])
const MADX_RESERVED = union(MADX_CLASSES, MADX_IGNORED, Set(keys(MADX_CONST)), Set(keys(MADX_FUNCS)),
                            Set(["line", "sequence", "endsequence", "beam", "use", "call", "const", "real", "int"]))

# ----------------------------------------------------------------------------------------
# Expressions: tokens and a recursive-descent evaluator (no eval of file text)
# ----------------------------------------------------------------------------------------
const MADX_TOKEN = r"\G\s*(?:(\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)|([A-Za-z_][\w.]*(?:->[A-Za-z_]\w*)?)|(\*\*|[-+*/^(),]))"

function madx_tokens(src::AbstractString)
    s = String(strip(src))
    out = Tuple{Symbol, Any}[]
    pos = 1
    while pos <= ncodeunits(s)
        m = match(MADX_TOKEN, s, pos)
        (m === nothing || m.offset != pos || isempty(m.match)) &&
            throw(MadxError("cannot read expression $(repr(s)) at $(repr(s[pos:end]))"))
        num, name, op = m.captures
        if num !== nothing
            push!(out, (:n, parse(Float64, num)))
        elseif name !== nothing
            push!(out, (:v, lowercase(name)))
        else
            push!(out, (:o, String(op)))
        end
        pos += ncodeunits(m.match)
    end
    return out
end

mutable struct MadxExpr
    t::Vector{Tuple{Symbol, Any}}
    i::Int
    lookup::Function
    src::String
end

madx_peek(x::MadxExpr) = x.i <= length(x.t) ? x.t[x.i] : (:none, nothing)
function madx_next!(x::MadxExpr)
    tok = madx_peek(x)
    x.i += 1
    return tok
end
isop(tok, ops...) = tok[1] == :o && tok[2] in ops

function madx_value(src::AbstractString, lookup::Function)
    x = MadxExpr(madx_tokens(src), 1, lookup, String(src))
    v = madx_sum(x)
    x.i <= length(x.t) && throw(MadxError("unexpected $(repr(madx_peek(x)[2])) in $(repr(x.src))"))
    return v
end

function madx_sum(x::MadxExpr)
    v = madx_prod(x)
    while isop(madx_peek(x), "+", "-")
        op = madx_next!(x)[2]
        w = madx_prod(x)
        v = op == "+" ? v + w : v - w
    end
    return v
end

function madx_prod(x::MadxExpr)
    v = madx_unary(x)
    while isop(madx_peek(x), "*", "/")
        op = madx_next!(x)[2]
        w = madx_unary(x)
        v = op == "*" ? v * w : v / w
    end
    return v
end

function madx_unary(x::MadxExpr)
    if isop(madx_peek(x), "-", "+")
        op = madx_next!(x)[2]
        v = madx_unary(x)
        return op == "-" ? -v : v
    end
    return madx_power(x)
end

function madx_power(x::MadxExpr)
    v = madx_atom(x)
    if isop(madx_peek(x), "^", "**")
        madx_next!(x)
        v = v^madx_unary(x)          # right-associative
    end
    return v
end

function madx_atom(x::MadxExpr)
    kind, val = madx_next!(x)
    kind == :n && return Float64(val)
    if kind == :o && val == "("
        v = madx_sum(x)
        madx_next!(x) == (:o, ")") || throw(MadxError("missing ')' in $(repr(x.src))"))
        return v
    end
    if kind == :v
        if madx_peek(x) == (:o, "(")
            haskey(MADX_FUNCS, val) || throw(MadxError("unknown function $(repr(val)) in $(repr(x.src))"))
            madx_next!(x)
            arg = madx_sum(x)
            madx_next!(x) == (:o, ")") || throw(MadxError("missing ')' in $(repr(x.src))"))
            return Float64(MADX_FUNCS[val](arg))
        end
        return Float64(x.lookup(val))
    end
    throw(MadxError("cannot read expression $(repr(x.src))"))
end

# ----------------------------------------------------------------------------------------
# Statements
# ----------------------------------------------------------------------------------------
function madx_strip_comments(text::AbstractString)
    text = replace(text, r"/\*.*?\*/"s => " ")
    out = String[]
    for line in split(text, '\n')
        q = nothing
        cut = lastindex(line) + 1
        i = firstindex(line)
        while i <= lastindex(line)
            ch = line[i]
            if q !== nothing
                ch == q && (q = nothing)
            elseif ch == '"' || ch == '\''
                q = ch
            elseif ch == '!' || (ch == '/' && i < lastindex(line) && line[nextind(line, i)] == '/')
                cut = i
                break
            end
            i = nextind(line, i)
        end
        push!(out, line[firstindex(line):prevind(line, cut)])
    end
    return join(out, "\n")
end

# split at sep outside (), {} and quotes
function madx_split_top(s::AbstractString, sep::Char = ',')
    parts = String[]
    depth = 0
    q = nothing
    cur = IOBuffer()
    for ch in s
        if q !== nothing
            print(cur, ch)
            ch == q && (q = nothing)
            continue
        end
        if ch == '"' || ch == '\''
            q = ch
        elseif ch == '(' || ch == '{'
            depth += 1
        elseif ch == ')' || ch == '}'
            depth -= 1
        end
        if ch == sep && depth == 0
            push!(parts, String(take!(cur)))
        else
            print(cur, ch)
        end
    end
    push!(parts, String(take!(cur)))
    return [String(strip(p)) for p in parts]
end

# `name = value` / `name := value` items -> Dict(name => (value text, deferred)); a bare name is
# a flag set to 1 (to 0 with a leading minus)
function madx_attrs(items)
    out = Dict{String, Tuple{String, Bool}}()
    for it in items
        isempty(it) && continue
        m = match(r"^([A-Za-z_][\w.]*)\s*(:=|=)\s*(.*)$"s, it)
        if m !== nothing
            out[lowercase(m.captures[1])] = (String(strip(m.captures[3])), m.captures[2] == ":=")
        elseif occursin(r"^-?[A-Za-z_][\w.]*$", it)
            neg = startswith(it, "-")
            out[lowercase(lstrip(it, '-'))] = (neg ? "0" : "1", false)
        else
            throw(MadxError("cannot read attribute $(repr(it))"))
        end
    end
    return out
end

mutable struct MadxModel
    vars::Dict{String, Tuple{String, Bool, Union{Nothing, Float64}}}
    elems::Dict{String, Tuple{String, Dict{String, Tuple{String, Bool}}}}   # label => (parent, attrs)
    lines::Dict{String, String}
    seqs::Dict{String, Tuple{Union{Nothing, Tuple{String, Bool}}, String, Vector{Tuple{String, Dict{String, Tuple{String, Bool}}}}}}
    seq_order::Vector{String}
    line_order::Vector{String}
    beam::Dict{String, Tuple{String, Bool}}
    use::Union{Nothing, String}
    directives::Dict{String, String}
    stack::Vector{String}
end
MadxModel() = MadxModel(Dict(), Dict(), Dict(), Dict(), String[], String[], Dict(), nothing, Dict(), String[])

function madx_lookup(M::MadxModel, name::AbstractString)
    if occursin("->", name)
        el, at = split(name, "->")
        return madx_attr_value(M, String(el), String(at))
    end
    if haskey(M.vars, name)
        text, deferred, val = M.vars[name]
        (val !== nothing && !deferred) && return val
        return madx_eval_guarded(M, text, String(name))
    end
    haskey(MADX_CONST, name) && return MADX_CONST[name]
    throw(MadxError("undefined name $(repr(name))"))
end

function madx_eval_guarded(M::MadxModel, text::AbstractString, what::String)
    what in M.stack && throw(MadxError("circular definition of $(repr(what))"))
    push!(M.stack, what)
    try
        return madx_value(text, n -> madx_lookup(M, n))
    finally
        pop!(M.stack)
    end
end

madx_eval(M::MadxModel, text::AbstractString) = madx_value(text, n -> madx_lookup(M, n))

# (base class, merged attributes) of an element, following its parents
function madx_resolved(M::MadxModel, label::AbstractString)
    chain = Dict{String, Tuple{String, Bool}}[]
    lab = String(label)
    seen = Set{String}()
    while !(lab in MADX_CLASSES)
        haskey(M.elems, lab) || throw(MadxError("undefined element or class $(repr(lab))"))
        lab in seen && throw(MadxError("circular element definition $(repr(lab))"))
        push!(seen, lab)
        parent, attrs = M.elems[lab]
        push!(chain, attrs)
        lab = parent
    end
    merged = Dict{String, Tuple{String, Bool}}()
    for a in reverse(chain)
        merge!(merged, a)
    end
    return lab, merged
end

function madx_attr_value(M::MadxModel, label::AbstractString, attr::AbstractString, default = nothing)
    _, at = madx_resolved(M, label)
    if !haskey(at, attr)
        if default === nothing
            attr == "l" && return 0.0
            throw(MadxError("$label has no attribute $(repr(attr))"))
        end
        return Float64(default)
    end
    return madx_eval_guarded(M, at[attr][1], "$label->$attr")
end

function madx_attr_list(M::MadxModel, label::AbstractString, attr::AbstractString)
    _, at = madx_resolved(M, label)
    haskey(at, attr) || return Float64[]
    text = strip(at[attr][1])
    (startswith(text, "{") && endswith(text, "}")) || return [madx_eval(M, text)]
    body = strip(text[2:end-1])
    return isempty(body) ? Float64[] : [madx_eval(M, x) for x in madx_split_top(body)]
end

function madx_statement!(M::MadxModel, st::AbstractString, base_dir::AbstractString, seq)
    s = String(strip(st))
    isempty(s) && return seq
    low = lowercase(s)
    if seq !== nothing
        low == "endsequence" && return nothing
        madx_sequence_item!(M, s, seq)
        return seq
    end
    m = match(r"^(?:(const|real|int)\s+)?([A-Za-z_][\w.]*)\s*(:=|=)\s*(.+)$"s, s)
    if m !== nothing && !occursin(r"^[A-Za-z_][\w.]*\s*:\s*line\b"i, s)
        name, op, text = lowercase(m.captures[2]), m.captures[3], String(strip(m.captures[4]))
        deferred = op == ":="
        M.vars[name] = (text, deferred, deferred ? nothing : madx_eval(M, text))
        return nothing
    end
    m = match(r"^([A-Za-z_][\w.]*)->([A-Za-z_]\w*)\s*(:=|=)\s*(.+)$"s, s)
    if m !== nothing
        el, at = lowercase(m.captures[1]), lowercase(m.captures[2])
        haskey(M.elems, el) || throw(MadxError("undefined element $(repr(el))"))
        M.elems[el][2][at] = (String(strip(m.captures[4])), m.captures[3] == ":=")
        return nothing
    end
    m = match(r"^([A-Za-z_][\w.]*)\s*:\s*(.+)$"s, s)
    m !== nothing && return madx_labelled!(M, lowercase(m.captures[1]), String(strip(m.captures[2])))
    parts = madx_split_top(s)
    cmd, rest = lowercase(parts[1]), parts[2:end]
    if cmd == "beam"
        merge!(M.beam, madx_attrs(rest))
    elseif cmd == "use"
        at = madx_attrs(rest)
        key = haskey(at, "sequence") ? "sequence" : haskey(at, "period") ? "period" : nothing
        key === nothing && throw(MadxError("USE needs sequence= or period="))
        M.use = lowercase(at[key][1])
    elseif cmd == "call"
        at = madx_attrs(rest)
        haskey(at, "file") || throw(MadxError("CALL needs file="))
        path = joinpath(base_dir, strip(at["file"][1], ['"', '\'']))
        madx_read_text!(M, read(path, String), dirname(path))
    elseif haskey(M.elems, cmd) && !isempty(rest)
        merge!(M.elems[cmd][2], madx_attrs(rest))        # "elem, attr = value": change the element
    elseif cmd in MADX_IGNORED || cmd == "endsequence"
        nothing
    else
        throw(MadxError("cannot read statement $(repr(first(s, 60)))"))
    end
    return nothing
end

function madx_labelled!(M::MadxModel, label::String, body::String)
    m = match(r"^line\s*=\s*\((.*)\)$"si, body)
    if m !== nothing
        M.lines[label] = m.captures[1]
        label in M.line_order || push!(M.line_order, label)
        return nothing
    end
    label in MADX_RESERVED && throw(MadxError("$(repr(label)) is a keyword, not a label"))
    parts = madx_split_top(body)
    cls, rest = lowercase(parts[1]), parts[2:end]
    if cls == "sequence"
        at = madx_attrs(rest)
        refer = lowercase(get(at, "refer", ("centre", false))[1])
        refer == "center" && (refer = "centre")
        refer in ("entry", "centre", "exit") || throw(MadxError("sequence $label: refer = $(repr(refer))"))
        M.seqs[label] = (get(at, "l", nothing), refer, Tuple{String, Dict{String, Tuple{String, Bool}}}[])
        label in M.seq_order || push!(M.seq_order, label)
        return label
    end
    M.elems[label] = (cls, madx_attrs(rest))
    madx_resolved(M, label)      # the parent must exist
    return nothing
end

function madx_sequence_item!(M::MadxModel, s::String, seq::String)
    m = match(r"^([A-Za-z_][\w.]*)\s*:\s*(.+)$"s, s)
    if m !== nothing                 # an element defined in place
        label = lowercase(m.captures[1])
        parts = madx_split_top(m.captures[2])
        at = madx_attrs(parts[2:end])
        place = Dict{String, Tuple{String, Bool}}()
        for k in ("at", "from")
            haskey(at, k) && (place[k] = pop!(at, k))
        end
        M.elems[label] = (lowercase(parts[1]), at)
        madx_resolved(M, label)
    else
        parts = madx_split_top(s)
        label = lowercase(parts[1])
        place = madx_attrs(parts[2:end])
        madx_resolved(M, label)
    end
    haskey(place, "at") || throw(MadxError("sequence $seq: $label has no at="))
    push!(M.seqs[seq][3], (label, place))
end

function madx_read_text!(M::MadxModel, text::AbstractString, base_dir::AbstractString)
    for m in eachmatch(r"trackpad:\s*(\w+)\s*=\s*(\w+)", text)
        M.directives[lowercase(m.captures[1])] = m.captures[2]
    end
    seq = nothing
    for st in madx_split_top(madx_strip_comments(text), ';')
        seq = madx_statement!(M, st, base_dir, seq)
    end
    seq === nothing || throw(MadxError("sequence $seq has no ENDSEQUENCE"))
end

# ----------------------------------------------------------------------------------------
# Flattening
# ----------------------------------------------------------------------------------------
function madx_expand_line(M::MadxModel, label::String, depth::Int = 0)
    depth > 100 && throw(MadxError("lines nested too deeply (recursive line?)"))
    out = String[]
    for item0 in madx_split_top(M.lines[label])
        isempty(item0) && continue
        item = item0
        rep, inv = 1, false
        m = match(r"^(\d+)\s*\*\s*(.+)$"s, item)
        if m !== nothing
            rep, item = parse(Int, m.captures[1]), String(strip(m.captures[2]))
        end
        if startswith(item, "-")
            inv, item = true, String(strip(item[2:end]))
        end
        if startswith(item, "(") && endswith(item, ")")
            anon = "__anon$(depth)_$(length(M.lines))"
            M.lines[anon] = item[2:end-1]
            sub = madx_expand_line(M, anon, depth + 1)
        elseif haskey(M.lines, lowercase(item))
            sub = madx_expand_line(M, lowercase(item), depth + 1)
        else
            madx_resolved(M, lowercase(item))
            sub = [lowercase(item)]
        end
        inv && (sub = reverse(sub))
        for _ in 1:rep
            append!(out, sub)
        end
    end
    return out
end

# the arc length of an element (RBEND: the chord l turned into the arc)
function madx_length(M::MadxModel, lab::AbstractString)
    L = madx_attr_value(M, lab, "l", 0.0)
    cls, _ = madx_resolved(M, lab)
    if cls == "rbend"
        ang = madx_attr_value(M, lab, "angle", 0.0)
        ang != 0.0 && (L = L * (ang / 2) / sin(ang / 2))
    end
    return L
end

# labels of a sequence with drift lengths (Float64) for the gaps between them
function madx_flatten_sequence(M::MadxModel, label::String)
    lsq, refer, items = M.seqs[label]
    pos = Dict{String, Float64}()
    placed = Tuple{Float64, Float64, String}[]
    for (lab, place) in items
        L = madx_length(M, lab)
        at = madx_eval(M, place["at"][1])
        if haskey(place, "from")
            ref = lowercase(place["from"][1])
            haskey(pos, ref) || throw(MadxError("sequence $label: from=$ref names no earlier element"))
            at += pos[ref]
        end
        centre = refer == "entry" ? at + L / 2 : refer == "centre" ? at : at - L / 2
        pos[lab] = centre
        push!(placed, (centre - L / 2, L, lab))
    end
    out = Any[]
    s = 0.0
    for (start, L, lab) in placed
        gap = start - s
        gap < -1e-9 && throw(MadxError("sequence $label: $lab overlaps the element before it by $(-gap) m"))
        gap > 1e-12 && push!(out, gap)
        push!(out, lab)
        s = start + L
    end
    total = lsq === nothing ? s : madx_eval(M, lsq[1])
    total - s > 1e-12 && push!(out, total - s)
    return out
end

# ----------------------------------------------------------------------------------------
# Conversion to TrackPad's element types
# ----------------------------------------------------------------------------------------
function madx_beam(M::MadxModel, energy)
    b = M.beam
    part = lowercase(strip(get(b, "particle", ("electron", false))[1], ['"', '\'']))
    mass = if haskey(b, "mass")
        madx_eval(M, b["mass"][1]) * 1e9
    elseif part in ("electron", "positron")
        TrackPad.M_ELECTRON
    elseif part in ("proton", "antiproton")
        TrackPad.M_PROTON
    else
        throw(MadxError("BEAM particle $(repr(part)): give mass="))
    end
    energy === nothing || return Beam(Float64(energy); mass = mass)
    etot = if haskey(b, "energy")
        madx_eval(M, b["energy"][1]) * 1e9
    elseif haskey(b, "pc")
        pc = madx_eval(M, b["pc"][1]) * 1e9
        sqrt(pc * pc + mass * mass)
    elseif haskey(b, "gamma")
        madx_eval(M, b["gamma"][1]) * mass
    else
        throw(MadxError("BEAM needs energy=, pc= or gamma= (or pass energy= to read_madx)"))
    end
    etot <= mass && throw(MadxError("BEAM total energy $etot eV is not above the rest mass"))
    return Beam(etot - mass; mass = mass)
end

function madx_element(M::MadxModel, lab::String, bend::String, beam::Beam, circumference::Float64)
    cls, at = madx_resolved(M, lab)
    name = uppercase(lab)
    g(a, d = 0.0) = madx_attr_value(M, lab, a, d)
    for skew in ("k1s", "k2s", "k3s", "tilt")
        haskey(at, skew) && g(skew) != 0.0 &&
            throw(MadxError("$name: $skew is not supported (the reference's magnets have no such field)"))
    end
    L = g("l")
    nst = round(Int, g("nst", 10.0))
    cls == "drift" && return Drift(L; name = name)
    if cls in ("marker", "monitor", "hmonitor", "vmonitor", "instrument", "placeholder")
        return L != 0.0 ? Drift(L; name = name) : Marker(; name = name)
    end
    cls == "quadrupole" && return Quadrupole(L, g("k1"); num_int_steps = nst, name = name)
    cls == "sextupole" && return Sextupole(L, g("k2"); num_int_steps = nst, name = name)
    cls == "octupole" && return Octupole(L, g("k3"); num_int_steps = nst, name = name)
    if cls in ("sbend", "rbend")
        ang, e1, e2 = g("angle"), g("e1"), g("e2")
        if cls == "rbend"
            ang != 0.0 && (L = L * (ang / 2) / sin(ang / 2))
            e1, e2 = e1 + ang / 2, e2 + ang / 2
        end
        fint = g("fint")
        fintx = g("fintx", fint)
        hgap = g("hgap")
        pb = zeros(4)
        pb[2], pb[3] = g("k1"), g("k2") / 2
        mo = pb[3] != 0.0 ? 2 : pb[2] != 0.0 ? 1 : 0
        if bend == "ExactSBend"
            return ExactSBend(L, ang, e1, e2; gk = fint * hgap, num_int_steps = nst, polynom_b = pb, max_order = mo,
                              name = name)
        elseif bend == "ExactSBendLinearEdge"
            return ExactSBendLinearEdge(L, ang, e1, e2; fint1 = fint, fint2 = fintx, gap = 2 * hgap,
                                        num_int_steps = nst, polynom_b = pb, max_order = mo, name = name)
        end
        return SBend(L, ang, e1, e2; fint1 = fint, fint2 = fintx, gap = 2 * hgap, num_int_steps = nst,
                     polynom_b = pb, max_order = mo, name = name)
    end
    if cls == "multipole"
        knl, ksl = madx_attr_list(M, lab, "knl"), madx_attr_list(M, lab, "ksl")
        n = max(length(knl), length(ksl), 1)
        n > MAXORD + 1 && throw(MadxError("$name: multipoles above order $MAXORD are not supported"))
        pb, pa = zeros(max(n, 4)), zeros(max(n, 4))
        for (i, v) in enumerate(knl)
            pb[i] = v / factorial(i - 1)
        end
        for (i, v) in enumerate(ksl)
            pa[i] = v / factorial(i - 1)
        end
        n <= 4 && return ThinMultipole(L, pa[1:4], pb[1:4]; max_order = n - 1, name = name)
        # orders above 3: the reference's SVector{4} holds the first four, the rest ride along
        return WithMultipoles(ThinMultipole(L, pa[1:4], pb[1:4]; max_order = 3, name = name), pa, pb, n - 1)
    end
    if cls == "rfcavity"
        harmon = g("harmon", 0.0)
        freq = if haskey(at, "freq")
            g("freq") * 1e6
        elseif harmon > 0.0
            harmon * beam.beta * C_LIGHT / circumference
        else
            throw(MadxError("$name: RFCAVITY needs freq= or harmon="))
        end
        return RFCavity(L, g("volt") * 1e6, freq, 0.0; h = harmon > 0.0 ? harmon : 1.0, philag = 2.0 * pi * g("lag"),
                        energy = beam.energy, name = name)
    end
    if cls in ("hkicker", "vkicker", "kicker")
        hk = cls == "hkicker" ? g("kick") : g("hkick")
        vk = cls == "vkicker" ? g("kick") : g("vkick")
        return Corrector(L, hk, vk; name = name)
    end
    cls == "solenoid" && return Solenoid(L, g("ks"); name = name)
    throw(MadxError("$name: element class $(uppercase(cls)) is not supported"))
end

"""
    read_madx(source; use = nothing, bend = nothing, energy = nothing) -> (Lattice, Beam)

Read a MAD-X file (a path, or the text itself if it holds a `;` or a newline) and return the
used sequence or line as a TrackPad `Lattice` with its `Beam`.  `use` overrides the file's
USE; `bend` is "SBend" (default), "ExactSBend" or "ExactSBendLinearEdge" (a `! trackpad: bend =
...` comment in the file sets it too); `energy` (kinetic, eV) overrides BEAM.  The result goes
to `compile_table` like any constructed lattice.
"""
function read_madx(source::AbstractString; use = nothing, bend = nothing, energy = nothing)
    if occursin(';', source) || occursin('\n', source)
        text, base = String(source), pwd()
    else
        text, base = read(source, String), dirname(abspath(source))
    end
    M = MadxModel()
    madx_read_text!(M, text, base)
    name = lowercase(something(use, M.use, ""))
    if isempty(name)
        cands = vcat(M.seq_order, M.line_order)
        length(cands) == 1 || throw(MadxError("no USE statement and not exactly one sequence or line: pass use="))
        name = cands[1]
    end
    bendmodel = something(bend, get(M.directives, "bend", "SBend"))
    bendmodel in ("SBend", "ExactSBend", "ExactSBendLinearEdge") ||
        throw(MadxError("bend model $(repr(bendmodel)) not in SBend, ExactSBend, ExactSBendLinearEdge"))
    beam = madx_beam(M, energy)
    flat = if haskey(M.seqs, name)
        madx_flatten_sequence(M, name)
    elseif haskey(M.lines, name)
        madx_expand_line(M, name)
    else
        throw(MadxError("no sequence or line named $(repr(name))"))
    end
    # the circumference, for cavities given by harmonic number
    circ = 0.0
    for it in flat
        circ += it isa Float64 ? it : madx_length(M, it)
    end
    els = AbstractElement[]
    k = 0
    for it in flat
        if it isa Float64
            k += 1
            push!(els, Drift(it; name = "DRIFT_$k"))
        else
            push!(els, madx_element(M, it, bendmodel, beam, circ))
        end
    end
    return Lattice(els; name = Symbol(uppercase(name))), beam
end
