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
a clear error: nbfixes, 1-4 pairs (`*k` exclusions), flexible angles,
torsions and vibrating bonds.

`read_get` reads the getdata syntax of `.def` and `.get` files
(name=value pairs up to `;`, values being numbers, quoted strings or
arithmetic expressions of names assigned before, as cook's calculator
evaluates them).

`system_from_cook(ble, def)` builds the engine's system from the two, as
cook's set-up does (initss, initrspace, Ewald parameters from el.epsr and
el.epsk, LJcutoff).
"""
import ast
import operator
import re


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


def read_ble(path):
    """the force field of a .ble file in the engine's description
    (types, type_mass, type_EvdW_kcal, type_RvdW, species, polar)"""
    text = open(path).read()
    # header: getdata up to ';' after skipping comments
    head, _ = parse_getdata(text)
    nsites = int(head.get("nsites", 0))
    nspec = int(head.get("nspec", 0))
    nnbfixes = int(head.get("nnbfixes", 0))
    distance14 = int(head.get("distance14", 4))
    polar = int(head.get("polar", -1))
    nparms = int(head.get("nparms", 0))
    eunit = float(head.get("eunit", 1.0))
    if nparms != 0:
        raise CookInputError("nparms=%d: only the Lennard-Jones form (SS_PARMS=0) is supported" % nparms)
    if nnbfixes:
        raise CookInputError("nbfixes are not supported by the engine")
    n14 = 2 if distance14 else 1
    L = _Lines(text)
    L.find("site-types" if polar < 0 else "Lennard-Jones")
    types, mass, evdw, rvdw, alphas, pol = [], [], [], [], [], {}
    for i in range(nsites):
        L.tok()                                   # site number
        name = L.tok()
        m = float(L.tok())
        groups = []
        for j in range(n14):
            t = L.tok(required=(j == 0))
            if t is None or t == "=":
                break
            groups.append((float(t), float(L.tok()) * eunit, float(L.tok())))
        alpha, E, R = groups[0]
        if E > 0:
            E = -E                                # initcombrule(): EvdW is a minimum
        alphas.append(alpha)
        if polar == 1:
            apol = float(L.tok())
            qpol = float(L.tok())
            if qpol:
                pol[name] = (apol, qpol)
        types.append(name)
        mass.append(m)
        evdw.append(E)
        rvdw.append(R)
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
              type_alpha=alphas, species=species, comb_rule=int(head.get("comb_rule", 0)))
    if pol:
        ff["polar"] = pol
    return ff


def system_from_cook(ble_path, def_path, provider=None, **kw):
    """(BuiltSystem, parameters) from a .ble force field and a .def deck:
    N[i], L[0..2], cutoff, LJcutoff, el.epsr, el.epsk, el.epsinf"""
    from .system import build_system
    ff = read_ble(ble_path)
    d = read_get(def_path)
    counts = [int(d.get("N[%d]" % i, 0)) for i in range(len(ff["species"]))]
    L = [float(d["L[%d]" % k]) for k in range(3)]
    cutoff = float(d["cutoff"])
    lj = d.get("LJcutoff", cutoff)
    args = dict(epsr=float(d.get("el.epsr", 0.05)), epsk=float(d.get("el.epsk", 0.05)),
                lj_cutoff=float(lj))
    if "el.epsinf" in d:
        args["epsinf"] = float(d["el.epsinf"])
    args.update(kw)
    return build_system(ff, counts, L, cutoff, provider=provider, **args), d
