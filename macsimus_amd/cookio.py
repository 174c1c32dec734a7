"""cook's input files for the engine: the force-field file `.ble` and the
parameter decks `.def` / `.get`.

`read_ble` restates readblend() (cook/simdef.c:484-1471) for the
Lennard-Jones form (sim/lj) with and without Drude polarizability: the
header data, the site-type table (`Lennard-Jones`, POLAR columns), the
`nbfixes` table and the species (`species`, `sites` with exclusions,
`bonds`).  Lines starting with `!` are comments and lines starting with
`***` are warnings written by blend (cook/simdef.c:137-168); a table is
found by its keyword as the first token of a line (find(),
cook/simdef.c:179-202); tokens are whitespace separated (tok(),
cook/simdef.c:205-218).  What the engine does not carry is rejected with
a clear error: nbfixes of the non-LJ forms, 1-4 pairs (`*k` exclusions), flexible angles,
torsions and vibrating bonds.

`read_get` reads the getdata syntax of `.def` and `.get` files
(name=value pairs up to `;`, values being numbers, quoted strings or
arithmetic expressions of names assigned before, as cook's calculator
evaluates them).

`system_from_cook(ble, def)` builds the engine's system from the two, as
cook's set-up does (initss, initrspace, Ewald parameters from el.epsr and
el.epsk, LJcutoff).

`write_rdf` / `read_rdf` write and read cook's binary RDF histograms
(`.rdf`, saverdf() of sim/simmeas.c:813-825) so that cook's own tools
(util/rdfg.c) take the engine's histograms.
"""
import ast
import operator
import re
import struct

import numpy as np


class CookInputError(ValueError):
    pass


# --------------------------------------------------------------- getdata
_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow, ast.USub: operator.neg,
        ast.UAdd: operator.pos}


def _ident(name):
    """a getdata name (el.epsr, L[0]) as a Python identifier"""
    return re.sub(r"\[(\d+)\]", r"__\1", name).replace(".", "_dot_")


def _eval(expr, env):
    """value of an arithmetic expression of numbers and of names in env
    (+ - * / ** and parentheses; nothing else is evaluated)"""
    names = {_ident(k): v for k, v in env.items()}

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return n.value
        if isinstance(n, ast.BinOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.operand))
        if isinstance(n, ast.Name):
            if n.id not in names:
                raise CookInputError("undefined name in %r" % expr)
            return names[n.id]
        raise CookInputError("unsupported expression %r" % expr)

    try:
        tree = ast.parse(_ident(expr), mode="eval")
    except SyntaxError as e:
        raise CookInputError("bad expression %r" % expr) from e
    return ev(tree)


def parse_getdata(text, env=None):
    """name=value pairs of a getdata block, up to `;` (or the end); `!`
    starts a comment to the end of the line.  Returns (dict, rest)."""
    env = dict(env or {})
    out = {}
    i, n = 0, len(text)
    buf = []
    while i < n:
        ch = text[i]
        if ch == "!":
            j = text.find("\n", i)
            i = n if j < 0 else j
            continue
        if ch == ";":
            i += 1
            break
        if ch == '"':
            j = text.find('"', i + 1)
            if j < 0:
                raise CookInputError("unterminated string")
            buf.append(text[i:j + 1])
            i = j + 1
            continue
        buf.append(ch)
        i += 1
    body = "".join(buf)
    # split into assignments: name=value, separated by whitespace; values
    # may contain spaces only inside quotes
    for m in re.finditer(r'([A-Za-z_][\w.\[\]]*)\s*=\s*("[^"]*"|[^\s]+)', body):
        name, val = m.group(1), m.group(2)
        if val.startswith('"'):
            v = val[1:-1]
        else:
            try:
                v = float(val) if re.search(r"[.eE]", val) else int(val)
            except ValueError:
                v = _eval(val, {**env, **out})
        out[name] = v
        env[name] = v
    return out, text[i:]


def read_get(path):
    """all getdata blocks of a .def or .get file, merged in order"""
    text = open(path).read()
    out = {}
    while text.strip():
        blk, text = parse_getdata(text, out)
        if not blk:
            break
        out.update(blk)
    return out


# ---------------------------------------------------------------- .ble
class _Lines:
    """mygetline(), find() and tok() of cook/simdef.c:137-218"""

    def __init__(self, text):
        self.lines = text.splitlines()
        self.i = 0
        self.toks = []
        self.found = ""

    def getline(self):
        while self.i < len(self.lines):
            line = self.lines[self.i].rstrip()
            self.i += 1
            if line.startswith("***") or line.startswith("!"):
                continue
            self.toks = line.split()
            return line
        self.toks = []
        return None

    def find(self, key):
        while True:
            line = self.getline()
            if line is None:
                raise CookInputError("table %s not found" % key)
            if self.toks and self.toks[0] == key:
                self.found = self.toks[1] if len(self.toks) > 1 else ""
                if self.getline() is None:
                    raise CookInputError("table %s is empty" % key)
                return

    def tok(self, required=True):
        if not self.toks:
            if required:
                raise CookInputError("missing data in line %d" % self.i)
            return None
        return self.toks.pop(0)


def read_ble(path, form=None):
    """the force field of a .ble file in the engine's description
    (types, type_mass, type_EvdW_kcal, type_RvdW, species, polar; for the
    exp-6 form also form="exp6" and type_parm, the stiffness a).

    The site-site form is cook's compile-time choice; the file has one
    parameter column per nparms.  nparms=0 is Lennard-Jones; nparms=1 is
    read as exp-6 when the file's table is the `exp-6` one (files with a
    polar key) or when form="exp6" is given; form="buck" and
    form="busing" read the Buckingham and Busing tables (POW = 0) of the
    same layout."""
    from .system import EUNIT_KCAL
    text = open(path).read()
    # header: getdata up to ';' after skipping comments
    head, _ = parse_getdata(text)
    nsites = int(head.get("nsites", 0))
    nspec = int(head.get("nspec", 0))
    nnbfixes = int(head.get("nnbfixes", 0))
    distance14 = int(head.get("distance14", 4))
    polar = int(head.get("polar", -1))
    nparms = int(head.get("nparms", 0))
    # energies of the file times eunit are K (cook/simdef.c:43,531,626;
    # default kcal/mol); the engine's description keeps kcal/mol
    escale = float(head.get("eunit", EUNIT_KCAL)) / EUNIT_KCAL
    comb_rule = int(head.get("comb_rule", 0))
    if nparms == 1:
        if form is None and polar >= 0:
            if re.search(r"^exp-6(\s|$)", text, re.M):
                form = "exp6"
            elif re.search(r"^Buckingham(\s|$)", text, re.M):
                form = "buck"
            elif re.search(r"^Busing(\s|$)", text, re.M):
                form = "busing"
        if form not in ("exp6", "buck", "busing"):
            raise CookInputError("nparms=1: name the site-site form (form='exp6', 'buck' or "
                                 "'busing')")
        if form == "exp6" and not comb_rule & 1:
            raise CookInputError("exp-6: comb_rule&1 = 0 (Kirkwood) is not implemented "
                                 "(sim/exp6/sitesite.c:52)")
    elif nparms == 0:
        if form not in (None, "lj"):
            raise CookInputError("form %r needs nparms=1 in the file" % form)
        form = "lj"
    else:
        raise CookInputError("nparms=%d: only the Lennard-Jones (0) and exp-6 (1) forms are "
                             "supported" % nparms)
    if nnbfixes and form != "lj":
        raise CookInputError("nbfixes are supported for the Lennard-Jones form only")
    n14 = 2 if distance14 else 1
    L = _Lines(text)
    L.find("site-types" if polar < 0 else
           {"exp6": "exp-6", "buck": "Buckingham", "busing": "Busing"}.get(form,
                                                                           "Lennard-Jones"))
    types, mass, evdw, rvdw, alphas, pol, parm, shell = [], [], [], [], [], {}, [], {}
    for i in range(nsites):
        L.tok()                                   # site number
        name = L.tok()
        m = float(L.tok())
        groups = []
        for j in range(n14):
            t = L.tok(required=(j == 0))
            if t is None or t == "=":
                break
            g = (float(t), float(L.tok()) * escale, float(L.tok()))
            # Buckingham: C_ii is an energy times A^6 (Epow 1, sim/buck/sitesite.c:16-24);
            # Busing: C_i is the square root of one (Epow 1/2, sim/busing/sitesite.c:5-16)
            ps = {"buck": escale, "busing": escale ** 0.5}.get(form, 1.0)
            groups.append(g + tuple(float(L.tok()) * ps for _ in range(nparms)))
        alpha, E, R = groups[0][:3]
        parm.append(groups[0][3:])
        if E > 0 and form not in ("buck", "busing"):
            E = -E                                # initcombrule(): EvdW is a minimum
        alphas.append(alpha)
        if polar == 1:
            apol = float(L.tok())
            qpol = float(L.tok())
            if qpol:
                pol[name] = (apol, qpol)
            # Esat, kappa and (cook built with POLAR&1) rep, each optional
            # (cook/simdef.c:646-694); kappa and rep are the shell-core data
            rest = []
            for _ in range(3):
                t = L.tok(required=False)
                if t is None:
                    break
                rest.append(t)
            kap = float(rest[1]) if len(rest) > 1 else 0.0
            rp = int(float(rest[2])) if len(rest) > 2 else 0
            if kap or rp:
                shell[name] = (kap, rp)
        types.append(name)
        mass.append(m)
        evdw.append(E)
        rvdw.append(R)
        L.getline()
    # nbfixes (cook/simdef.c:772-798): type, type, then EvdW RvdW of the
    # pair itself (Epow 1 and 0: scaled by eunit, not combined, not
    # sign-flipped) and optionally a 1-4 group, which the engine does not use
    nbfix = []
    if nnbfixes:
        L.find("NBFIX" if polar < 0 else "nbfixes")
        for _ in range(nnbfixes):
            names = []
            for _k in range(2):
                # by name only, as atomn() (cook/simdef.c:77-120)
                t = L.tok()
                if t not in types:
                    raise CookInputError("nbfix: unknown site type %s" % t)
                names.append(t)
            nbfix.append([names[0], names[1], float(L.tok()) * escale, float(L.tok())])
            L.getline()
    species = []
    for sp in range(nspec):
        L.find("species")
        sname = L.found
        # the species' getdata block (lines up to ';')
        blk_text = "\n".join(L.lines[L.i - 1:])
        blk, _ = parse_getdata(blk_text)
        ns = int(blk.get("ns", 0))
        nc = int(blk.get("nc", 0))
        for k in ("nangles", "ndihedrals", "nimpropers", "naromatics"):
            if int(blk.get(k, 0)):
                raise CookInputError("species %s: %s=%d: flexible terms are not supported"
                                     % (sname, k, int(blk[k])))
        if int(blk.get("config", 0)):
            raise CookInputError("species %s: config tables are not supported" % sname)
        L.find("sites")
        site_type, site_q, site_xyz, site_excl = [], [], [], []
        for i in range(ns):
            if int(L.tok()) != i:
                raise CookInputError("species %s: site number %d out of order" % (sname, i))
            t = L.tok()
            if t not in types:
                raise CookInputError("species %s: unknown site type %s" % (sname, t))
            site_type.append(types.index(t))
            site_q.append(float(L.tok()))
            site_xyz.append([float(L.tok()) for _ in range(3)])
            nexc = int(float(L.tok()))
            exc = []
            for _ in range(nexc):
                e = L.tok()
                if e.startswith("*"):
                    raise CookInputError("species %s: 1-4 pairs are not supported" % sname)
                exc.append(int(e))
            site_excl.append(exc)
            L.getline()
        L.find("bonds")
        bonds = []
        for _ in range(nc):
            i = int(L.tok()); L.tok(); j = int(L.tok()); L.tok()
            K = float(L.tok()); length = float(L.tok())
            if K != 0:
                raise CookInputError("species %s: bond %d-%d has K=%g: vibrating bonds are not "
                                     "supported (K=0 makes a constraint)" % (sname, i, j, K))
            bonds.append([i, j, length])
            L.getline()
        species.append(dict(name=sname, site_type=site_type, site_charge_e=site_q,
                            site_xyz=site_xyz, site_excl=site_excl, bonds=bonds))
    ff = dict(types=types, type_mass=mass, type_EvdW_kcal=evdw, type_RvdW=rvdw,
              type_alpha=alphas, species=species, comb_rule=comb_rule)
    if nbfix:
        ff["nbfix"] = nbfix
    if form in ("exp6", "buck", "busing"):
        ff["form"] = form
        ff["type_parm"] = [p[0] for p in parm]
    if pol:
        ff["polar"] = pol
    if shell:
        ff["shell"] = shell
    return ff


def system_from_cook(ble_path, def_path, provider=None, form=None, exp6_cook_units=True,
                     shell_core=False, **kw):
    """(BuiltSystem, parameters) from a .ble force field and a .def deck:
    N[i], L[0..2], cutoff, LJcutoff, el.epsr, el.epsk, el.epsinf.

    shell_core=True applies the kappa and rep columns of a polar file as
    cook built with POLAR&1 does (the shell-core repulsion); by default
    they are read and left unused, as in cook's POLAR=0 build.

    exp-6: cook's exp-6 initcombrule() multiplies EvdW by kcal/Eunit
    (sim/exp6/sitesite.c:52-54) after readblend() has converted it with
    eunit (cook/simdef.c:626), so cook's exp-6 runs with EvdW scaled twice;
    exp6_cook_units=True reproduces that, False converts once as the other
    forms do."""
    from .system import EUNIT_KCAL, build_system
    ff = read_ble(ble_path, form)
    if ff.get("form") == "exp6" and exp6_cook_units:
        kw.setdefault("exp6_evdw_scale", EUNIT_KCAL)
    d = read_get(def_path)
    counts = [int(d.get("N[%d]" % i, 0)) for i in range(len(ff["species"]))]
    L = [float(d["L[%d]" % k]) for k in range(3)]
    cutoff = float(d["cutoff"])
    lj = d.get("LJcutoff", cutoff)
    args = dict(epsr=float(d.get("el.epsr", 0.05)), epsk=float(d.get("el.epsk", 0.05)),
                lj_cutoff=float(lj))
    if "el.epsinf" in d:
        args["epsinf"] = float(d["el.epsinf"])
    if shell_core and ff.get("shell"):
        args["shell"] = ff["shell"]
    args.update(kw)
    return build_system(ff, counts, L, cutoff, provider=provider, **args), d


# ---------------------------------------------------------------- .rdf
def _var_record(data):
    """one record of cook's variable-length files (gen/varfile.c:61-74):
    int4 size, the bytes, int4 checksum 0x4A4B + the sum of the bytes"""
    return struct.pack("<i", len(data)) + data + struct.pack("<I", (0x4A4B + sum(data)) & 0xffffffff)


def _var_records(data):
    recs, o = [], 0
    while True:
        if o + 4 > len(data):
            raise CookInputError("varfile: no end record")
        (n,) = struct.unpack_from("<i", data, o)
        o += 4
        if n == 0:
            return recs
        rec = data[o:o + n]
        if len(rec) != n or o + n + 4 > len(data):
            raise CookInputError("varfile: truncated record")
        (chk,) = struct.unpack_from("<I", data, o + n)
        if (0x4A4B + sum(rec)) & 0xffffffff != chk:
            raise CookInputError("varfile: checksum")
        recs.append(rec)
        o += n + 4


def _rdf_index(ti, tj):
    hi, lo = max(ti, tj), min(ti, tj)
    return hi * (hi + 1) // 2 + lo


def rdf_npair(ff, counts, merged=False):
    """npair of each RDF record as cutcorr() counts it (sim/siminit.c:
    665-720): for every species pair and every site pair of the two
    species, N(N-1)/2 (same species) or N*M pairs of the sites' type pair;
    then N per intramolecular pair that is not excluded"""
    nt = len(ff["types"])
    out = np.zeros(1 if merged else nt * (nt + 1) // 2)
    sp = ff["species"]
    for n in range(len(sp)):
        for m in range(n + 1):
            npr = counts[n] * (counts[n] - 1) / 2 if m == n else counts[n] * counts[m]
            for ti in sp[n]["site_type"]:
                for tj in sp[m]["site_type"]:
                    out[0 if merged else _rdf_index(ti, tj)] += npr
        ns = len(sp[n]["site_type"])
        ex = set()
        for i, lst in enumerate(sp[n]["site_excl"]):
            for j in lst:
                ex.add((max(i, j), min(i, j)))
        for i in range(ns):
            for j in range(i):
                if (i, j) not in ex:
                    t = (sp[n]["site_type"][i], sp[n]["site_type"][j])
                    out[0 if merged else _rdf_index(*t)] += counts[n]
    return out


def write_rdf(path, rdf, ff, counts):
    """cook's .rdf file of the histograms rdf = Engine.rdf() (hist
    [records, bins], nmeas, V, grid as given to set_rdf) for the force
    field ff with counts molecules per species: the grid as given, the
    number of site types, then per record the rdf_t of sim/rdf.h without
    its size field (names, type indices, site counts, npair, grid, V,
    nmeas, nhist, 32-bit counts) and the end record.  Counts above
    2^32-1 do not fit cook's record and raise ValueError."""
    grid = rdf["grid"]
    g = int(grid)
    hist = np.asarray(rdf["hist"])
    nt = len(ff["types"])
    merged = g < 0
    nrec = 1 if merged else nt * (nt + 1) // 2
    if hist.ndim != 2 or hist.shape[0] != nrec:
        raise ValueError("hist must have %d records" % nrec)
    if hist.size and int(hist.max()) > 0xffffffff:
        raise ValueError("a histogram count exceeds cook's unsigned 32-bit field")
    sitecount = [0] * nt
    for sp, N in zip(ff["species"], counts):
        for t in sp["site_type"]:
            sitecount[t] += N
    npair = rdf_npair(ff, counts, merged)
    name8 = lambda s: s.encode()[:8].ljust(8, b"\0")
    out = [_var_record(struct.pack("<d", float(grid))), _var_record(struct.pack("<i", nt))]
    pairs = [(0, 0)] if merged else [(i, j) for i in range(nt) for j in range(i + 1)]
    for k, (i, j) in enumerate(pairs):
        if merged:
            names, ns = (name8("*"), name8("*")), (sum(sitecount),) * 2
        else:
            names, ns = (name8(ff["types"][i]), name8(ff["types"][j])), (sitecount[i], sitecount[j])
        rec = (names[0] + names[1] + struct.pack("<2i2i", i, j, *ns) + b"\0" * 4 +
               struct.pack("<3d2i", npair[k], float(abs(g)), rdf["V"], rdf["nmeas"],
                           hist.shape[1]) +
               hist[k].astype("<u4").tobytes())
        out.append(_var_record(rec))
    with open(path, "wb") as f:
        f.write(b"".join(out) + struct.pack("<i", 0))


def read_rdf(path):
    """(grid, number of site types, records) of a cook .rdf file; each
    record a dict of names, indx, ns, npair, grid, V, nmeas, hist"""
    recs = _var_records(open(path, "rb").read())
    if len(recs) < 2:
        raise CookInputError("%s: no header" % path)
    (grid,) = struct.unpack("<d", recs[0])
    (nsites,) = struct.unpack("<i", recs[1])
    out = []
    for rec in recs[2:]:
        nhist = struct.unpack_from("<i", rec, 64)[0]
        if len(rec) != 68 + 4 * nhist:
            raise CookInputError("%s: record length %d for %d bins" % (path, len(rec), nhist))
        npair, grd, V = struct.unpack_from("<3d", rec, 36)
        out.append(dict(names=[rec[0:8].split(b"\0")[0].decode(), rec[8:16].split(b"\0")[0].decode()],
                        indx=list(struct.unpack_from("<2i", rec, 16)),
                        ns=list(struct.unpack_from("<2i", rec, 24)), npair=npair, grid=grd, V=V,
                        nmeas=struct.unpack_from("<i", rec, 60)[0],
                        hist=np.frombuffer(rec, "<u4", nhist, 68).astype(np.uint64)))
    return grid, nsites, out


# ---------------------------------------------------------------- .plb
class PlaybackWriter:
    """cook's playback file (openplayback() and writeplayback(),
    sim/simils.c:1550-1680): a header of two floats (ns, -3; -6 for a
    centred file), then per frame the box L and the ns site vectors, all as
    little-endian float32.  write(L, r) takes the positions r(t) (not
    wrapped, as cook keeps them); for cook's -n2 pass r - r1/2 (r(t-h/2)),
    for a .vlb file r1/h (velocities v(t-h/2)).  With centered=True the
    sites are shifted by -L/2 (cook's -n4).  append=True continues a file
    whose header matches ns."""

    def __init__(self, path, ns, centered=False, append=False):
        self.ns = int(ns)
        self.centered = bool(centered)
        if append:
            with open(path, "rb") as f:
                head = np.frombuffer(f.read(8), "<f4")
            if len(head) != 2 or int(head[0]) != self.ns:
                raise CookInputError("%s: not a playback file of %d sites" % (path, self.ns))
            self.f = open(path, "ab")
        else:
            self.f = open(path, "wb")
            self.f.write(np.array([self.ns, -6 if centered else -3], "<f4").tobytes())

    def write(self, L, r):
        r = np.asarray(r, dtype=np.float64)
        if r.shape != (self.ns, 3):
            raise ValueError("r must be [%d, 3]" % self.ns)
        L = np.asarray(L, dtype=np.float64).reshape(3)
        if self.centered:
            r = r - L / 2
        self.f.write(L.astype("<f4").tobytes() + r.astype("<f4").tobytes())
        self.f.flush()

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_plb(path):
    """(ns, header flag, L [frames, 3], r [frames, ns, 3]) of a playback
    file (float32 as stored)"""
    data = open(path, "rb").read()
    if len(data) < 8:
        raise CookInputError("%s: too short" % path)
    ns, flag = np.frombuffer(data[:8], "<f4")
    ns = int(ns)
    fb = 12 * (ns + 1)
    if (len(data) - 8) % fb:
        raise CookInputError("%s: %d bytes after the header is not whole frames of %d sites"
                             % (path, len(data) - 8, ns))
    a = np.frombuffer(data[8:], "<f4").reshape(-1, ns + 1, 3)
    return ns, float(flag), a[:, 0].copy(), a[:, 1:].copy()


# ---------------------------------------------------------------- .cfg
def read_cfg(path):
    """cook's configuration file (savecfg(), sim/simils.c:1445-1517): the
    varfile records cfgkey, optionlist[32], nspec, spec[nspec] (N, ns),
    three key records (1: N, t, h, En.tot; 2: ns, L; 4: thermostat,
    sigvdW, tau.T, tau.P; 17 for Drude systems: tau.dip), then cfg[0],
    cfg[1], .. as ToInt blocks (sim/simglob.h:306-315: size, dep, logs,
    lambda[3], offdiag[3], rp[nrow][3]), closed by KeyClose(): an int4 0
    and the int4 time key.  Returns a dict with those fields; cfgs is a
    list of dict(dep, logs, lambda, offdiag, rp)."""
    data = open(path, "rb").read()
    recs, o = [], 0
    while True:
        if o + 4 > len(data):
            raise CookInputError("%s: no end record" % path)
        (n,) = struct.unpack_from("<i", data, o)
        o += 4
        if n == 0:
            break
        rec = data[o:o + n]
        if len(rec) != n or o + n + 4 > len(data):
            raise CookInputError("%s: truncated record" % path)
        (chk,) = struct.unpack_from("<I", data, o + n)
        if (0x4A4B + sum(rec)) & 0xffffffff != chk:
            raise CookInputError("%s: checksum" % path)
        recs.append(rec)
        o += n + 4
    if o + 4 != len(data):
        raise CookInputError("%s: %d bytes after the end record" % (path, len(data) - o))
    (timekey,) = struct.unpack_from("<i", data, o)
    out = dict(cfgkey=struct.unpack("<i", recs[0])[0],
               options=list(struct.unpack("<32i", recs[1])),
               nspec=struct.unpack("<i", recs[2])[0], timekey=timekey, keys={})
    out["spec"] = [struct.unpack_from("<2i", recs[3], 8 * i) for i in range(out["nspec"])]
    k = 4
    while k < len(recs) and len(recs[k]) == 32:
        key, iv = struct.unpack_from("<2i", recs[k])
        out["keys"][key] = (iv, list(struct.unpack_from("<3d", recs[k], 8)))
        k += 1
    cfgs = []
    for rec in recs[k:]:
        size, dep = struct.unpack_from("<2i", rec)
        if size != len(rec) or (size - 64) % 24:
            raise CookInputError("%s: bad configuration record" % path)
        v = np.frombuffer(rec, "<f8", offset=8)
        cfgs.append(dict(dep=dep, logs=float(v[0]), **{"lambda": v[1:4].copy()},
                         offdiag=v[4:7].copy(), rp=v[7:].reshape(-1, 3).copy()))
    out["cfgs"] = cfgs
    if 1 in out["keys"]:
        N, (t, h, Etot) = out["keys"][1]
        out.update(N=N, t=t, h=h, Etot=Etot)
    if 2 in out["keys"]:
        out.update(ns=out["keys"][2][0], L=out["keys"][2][1])
    if 4 in out["keys"]:
        th, (sig, tauT, tauP) = out["keys"][4]
        out.update(thermostat=th, sigvdW=sig, tauT=tauT, tauP=tauP)
    return out


def write_cfg(path, cfg):
    """the configuration file of a dict as read_cfg returns it (records in
    the same order; a cfg level's ToInt is rebuilt from dep, logs, lambda,
    offdiag and rp), e.g. a cook file with the engine's downloaded state
    put into cfgs[0]["rp"] (r) and cfgs[1]["rp"] (r1 = r(t)-r(t-h))"""
    out = [_var_record(struct.pack("<i", cfg["cfgkey"])),
           _var_record(struct.pack("<32i", *cfg["options"])),
           _var_record(struct.pack("<i", len(cfg["spec"]))),
           _var_record(b"".join(struct.pack("<2i", *s) for s in cfg["spec"]))]
    for key, (iv, vec) in cfg["keys"].items():
        out.append(_var_record(struct.pack("<2i3d", key, iv, *vec)))
    for c in cfg["cfgs"]:
        rp = np.ascontiguousarray(c["rp"], dtype="<f8").reshape(-1, 3)
        body = np.concatenate([[c["logs"]], c["lambda"], c["offdiag"], rp.ravel()]).astype("<f8")
        out.append(_var_record(struct.pack("<2i", 8 + body.nbytes, c["dep"]) + body.tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out) + struct.pack("<2i", 0, cfg["timekey"]))
