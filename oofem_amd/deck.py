"""Reader for the reference's text input decks, restricted to the records the
assembly-and-solve path uses, and the deck's errorcheck block.

Record syntax follows OOFEMTXTDataReader / OOFEMTXTInputRecord
(oofemtxtdatareader.C:42-178): one record per line, keyword then number then
``key value`` pairs; arrays are ``count v1 v2 ...``; ranges ``{(a b) c}``;
``#`` starts a comment.  Section order: output file, description, engineering
model, export modules, ``domain``, ``OutputManager``, the component counts,
then dof managers, elements, cross sections, materials, boundary conditions,
initial conditions, load-time functions and sets.

Supported records: node; LSpace / LTRSpace / QSpace; SimpleCS; IsoLE,
MisesMat and TutorialMaterial (the same J2 return as MisesMat; with density
``d``); BoundaryCondition; ConstantSurfaceLoad;
DeadWeight; ConstantFunction and
PiecewiseLinFunction; Set with nodes / elements / elementranges /
elementBoundaries / noderanges.

The errorcheck block (errorcheckingexportmodule.C:78-735) is a list of
``#NODE``, ``#REACTION``, ``#ELEMENT`` and ``#LOADLEVEL`` lines, each an absolute window
``|value - expected| <= tolerance`` (errorcheckingexportmodule.C:66).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

from .meshgen import LSPACE, LTRSPACE, QSPACE, MAT_ISOLE, MAT_MISES, MATP, Deck, body_load_vector

_ETYPES = {"lspace": LSPACE, "ltrspace": LTRSPACE, "qspace": QSPACE}


class DeckError(ValueError):
    pass


def _tokens(line):
    # braces of range lists become separate tokens; "f(t)" stays one token
    line = line.replace("{", " { ").replace("}", " } ")
    line = re.sub(r"\(\s*(-?\d+)\s+(-?\d+)\s*\)", r" ( \1 \2 ) ", line)
    return line.split()


class Record:
    """One input record: keyword, number and the raw token list."""

    def __init__(self, line):
        self.line = line
        self.tok = _tokens(line)
        self.keyword = self.tok[0].lower()
        self.number = int(self.tok[1]) if len(self.tok) > 1 and re.fullmatch(r"-?\d+", self.tok[1]) else None

    def _find(self, key):
        key = key.lower()
        for i, t in enumerate(self.tok):
            if t.lower() == key:
                return i
        return -1

    def has(self, key):
        return self._find(key) >= 0

    def scalar(self, key, default=None, cast=float):
        i = self._find(key)
        if i < 0:
            if default is None:
                raise DeckError(f"missing field '{key}' in record: {self.line}")
            return default
        return cast(self.tok[i + 1])

    def array(self, key, cast=float, default=None):
        i = self._find(key)
        if i < 0:
            if default is None:
                raise DeckError(f"missing array '{key}' in record: {self.line}")
            return default
        n = int(self.tok[i + 1])
        return [cast(v) for v in self.tok[i + 2: i + 2 + n]]

    def ranges(self, key):
        """``{(a b) c ...}`` -> list of ints (inclusive ranges expanded)."""
        i = self._find(key)
        if i < 0:
            return None
        out, j = [], i + 1
        if self.tok[j] != "{":
            raise DeckError(f"bad range list in: {self.line}")
        j += 1
        while self.tok[j] != "}":
            if self.tok[j] == "(":
                a, b = int(self.tok[j + 1]), int(self.tok[j + 2])
                out.extend(range(a, b + 1))
                j += 4
            else:
                out.append(int(self.tok[j]))
                j += 1
        return out


@dataclass
class LoadTimeFunction:
    kind: str
    value: float = 1.0
    t: list = field(default_factory=list)
    f: list = field(default_factory=list)

    def __call__(self, time):
        if self.kind == "constantfunction":
            return self.value
        # PiecewiseLinFunction: linear interpolation, constant outside
        return float(np.interp(time, self.t, self.f))


@dataclass
class Check:
    kind: str          # "node", "reaction", "element", "loadlevel"
    tstep: int
    number: int
    dof: int = 0
    gp: int = 0
    keyword: int = 0
    component: int = 0
    value: float = 0.0
    line: str = ""


@dataclass
class DeckProblem:
    """Geometry + per-step loading of a deck, and its errorcheck block."""
    deck: Deck
    model: str
    nsteps: int
    engng: Record
    bcs: list
    loads: list
    ltf: dict
    node_numbers: np.ndarray
    elem_numbers: np.ndarray
    checks: list
    tolerance: float
    bodyloads: list = field(default_factory=list)
    density: np.ndarray = None     # per material index (record ``d``)
    nodalloads: list = field(default_factory=list)

    def time_of_step(self, k):
        """StaticStructural / LinearStatic: step k (1-based) at t = k * deltaT."""
        dt = self.engng.scalar("deltat", 1.0) if self.engng.has("deltat") else 1.0
        return k * dt

    def step_loading(self, k, body_load=None):
        """(prescribed values per prescribed equation, (f_free, f_presc)) of
        step k. ``body_load(deck, elems, dens, b, factor)`` assembles each
        DeadWeight (default: meshgen.body_load_vector)."""
        return self.loading_at(self.time_of_step(k), body_load=body_load)

    def loading_at(self, t, body_load=None):
        """(prescribed values, (f_free, f_presc)) at time t."""
        d = self.deck
        val = np.zeros_like(d.bc_value)
        for bc in self.bcs:
            fac = self.ltf[bc["ltf"]](t)
            for dof, v in zip(bc["dofs"], bc["values"]):
                val[bc["nodes"], dof - 1] = v * fac
        up = np.zeros(d.npeq)
        m = d.eqn < 0
        up[-d.eqn[m] - 1] = val[m]
        return up, self.loads_at(t, body_load=body_load)

    def loads_at(self, t, body_load=None):
        """(f_free, f_presc): every load of the deck with its load-time
        function evaluated at time t."""
        d = self.deck
        f_free = np.zeros(d.neq)
        f_presc = np.zeros(d.npeq)
        for ld in self.loads:
            fac = self.ltf[ld["ltf"]](t)
            d.faces = ld["faces"]
            d.traction = np.tile(ld["traction"], (len(ld["faces"]), 1))
            ff, fp = d.external_load(fac)
            f_free += ff
            f_presc += fp
        for nl in self.nodalloads:
            fac = self.ltf[nl["ltf"]](t)
            for dof, c in zip(nl["dofs"], nl["comps"]):
                q = d.eqn[nl["nodes"], dof - 1]
                np.add.at(f_free, q[q > 0] - 1, c * fac)
                np.add.at(f_presc, -q[q < 0] - 1, c * fac)
        bl_fn = body_load or body_load_vector
        for bl in self.bodyloads:
            fac = self.ltf[bl["ltf"]](t)
            el = bl["elements"]
            ff, fp = bl_fn(d, el, self.density[d.emat[el]], bl["b"], fac)
            f_free += ff
            f_presc += fp
        return f_free, f_presc


def _records(path):
    lines = []
    checks_txt = []
    in_check = False
    with open(path) as fh:
        for raw in fh:
            s = raw.strip()
            if s.startswith("#%BEGIN_CHECK%"):
                in_check = True
                checks_txt.append(s)
                continue
            if s.startswith("#%END_CHECK%"):
                in_check = False
                continue
            if in_check:
                checks_txt.append(s)
                continue
            if not s or s.startswith("#"):
                continue
            lines.append(s)
    return lines, checks_txt


def _parse_checks(txt):
    tol = 0.0
    out = []
    for s in txt:
        if s.startswith("#%BEGIN_CHECK%"):
            m = re.search(r"tolerance\s+(\S+)", s)
            tol = float(m.group(1)) if m else 1e-6
            continue
        if s.startswith("##"):
            continue
        t = s.lstrip("#").split()
        if not t:
            continue
        kind = t[0].lower()
        kv = {t[i].lower(): t[i + 1] for i in range(1, len(t) - 1)}
        if kind == "node":
            out.append(Check("node", int(kv["tstep"]), int(kv["number"]), dof=int(kv["dof"]),
                             value=float(kv["value"]), line=s))
        elif kind == "reaction":
            out.append(Check("reaction", int(kv["tstep"]), int(kv["number"]), dof=int(kv["dof"]),
                             value=float(kv["value"]), line=s))
        elif kind == "loadlevel":
            # errorcheckingexportmodule.C:660-703: the engineering model's load level
            out.append(Check("loadlevel", int(kv["tstep"]), 0, value=float(kv["value"]), line=s))
        elif kind == "element":
            out.append(Check("element", int(kv["tstep"]), int(kv["number"]), gp=int(kv["gp"]),
                             keyword=int(kv["keyword"]), component=int(kv["component"]),
                             value=float(kv["value"]), line=s))
    return tol, out


def read_deck(path) -> DeckProblem:
    lines, checks_txt = _records(path)
    if len(lines) < 6:
        raise DeckError("deck too short")
    pos = 2                                  # output file, description
    engng = Record(lines[pos]); pos += 1
    # metastep records follow the engineering-model record (EngngModel::
    # instanciateMetaSteps reads nmsteps of them); they carry the solver
    # keywords of their steps
    nmsteps = engng.scalar("nmsteps", 0, int)
    metasteps = [Record("metastep 0 " + lines[pos + i]) for i in range(nmsteps)]
    pos += nmsteps
    nmodules = engng.scalar("nmodules", 0, int)
    pos += nmodules                          # export module records
    if not lines[pos].lower().startswith("domain"):
        raise DeckError(f"expected 'domain', got: {lines[pos]}")
    pos += 1
    if lines[pos].lower().startswith("outputmanager"):
        pos += 1
    counts = Record("counts 0 " + lines[pos]); pos += 1
    nd = counts.scalar("ndofman", cast=int)
    ne = counts.scalar("nelem", cast=int)
    ncs = counts.scalar("ncrosssect", cast=int)
    nmat = counts.scalar("nmat", cast=int)
    nbc = counts.scalar("nbc", cast=int)
    nic = counts.scalar("nic", 0, int)
    nltf = counts.scalar("nltf", cast=int)
    nset = counts.scalar("nset", 0, int)

    def take(n):
        nonlocal pos
        out = [Record(lines[pos + i]) for i in range(n)]
        pos += n
        return out

    nodes, elems, css, mats, bcs, ics, ltfs, sets = (take(nd), take(ne), take(ncs), take(nmat), take(nbc),
                                                      take(nic), take(nltf), take(nset))
    node_numbers = np.array([r.number for r in nodes])
    nid = {n: i for i, n in enumerate(node_numbers)}
    coords = np.array([r.array("coords") for r in nodes], dtype=np.float64)
    etypes = {r.keyword for r in elems}
    if len(etypes) != 1 or next(iter(etypes)) not in _ETYPES:
        raise DeckError(f"unsupported element types {etypes}")
    etype = _ETYPES[next(iter(etypes))]
    for r in elems:
        # NLStructuralElement nlGeometry 1 (Green-Lagrange strain,
        # nlstructuralelement.C:151-241, 332-448) is outside this small-strain path
        if r.scalar("nlgeo", 0, int) != 0:
            raise DeckError(f"element {r.number}: nlgeo {r.scalar('nlgeo', 0, int)} (geometric nonlinearity) "
                            "is not supported")
    elem_numbers = np.array([r.number for r in elems])
    eid = {n: i for i, n in enumerate(elem_numbers)}
    conn = np.array([[nid[v] for v in r.array("nodes", int)] for r in elems], dtype=np.int32)
    setd = {}
    for s in sets:
        d = {}
        if s.has("nodes"):
            d["nodes"] = [nid[v] for v in s.array("nodes", int)]
        if s.has("noderanges"):
            d["nodes"] = [nid[v] for v in s.ranges("noderanges")]
        if s.has("elements"):
            d["elements"] = [eid[v] for v in s.array("elements", int)]
        if s.has("elementranges"):
            d["elements"] = [eid[v] for v in s.ranges("elementranges")]
        if s.has("elementboundaries"):
            v = s.array("elementboundaries", int)
            d["faces"] = np.array([[eid[v[2 * i]], v[2 * i + 1]] for i in range(len(v) // 2)], dtype=np.int32)
        setd[s.number] = d
    # materials
    mrows = []
    mnum = {}
    dens = []
    for i, r in enumerate(mats):
        row = np.zeros(MATP)
        dens.append(r.scalar("d", 0.0))
        if r.keyword == "isole":
            row[0] = MAT_ISOLE
        elif r.keyword == "misesmat":
            row[0] = MAT_MISES
            if r.scalar("htype", 0, int) != 0:
                raise DeckError("MisesMat: only htype 0 (linear hardening) is supported")
            if r.scalar("omega_crit", 0.0) != 0.0:
                # misesmat.C:121-133, 451-478, 535-546: stress (1 - omega) sigma_eff
                # and an unsymmetric tangent term; not part of this build
                raise DeckError("MisesMat: damage (omega_crit != 0) is not supported")
            row[3] = r.scalar("sig0")
            row[4] = r.scalar("h", 0.0)
        elif r.keyword == "tutorialmaterial":
            # TutorialMaterial (tutorialmaterial.C:80-163) is the same J2
            # radial return with linear isotropic hardening and the same
            # consistent tangent as MisesMat: yield sqrt(3 J2) - (sigy + H k),
            # multiplier phi / (3G + H); its records map onto the MisesMat
            # kernels (sigy -> sig0, H -> h)
            row[0] = MAT_MISES
            row[3] = r.scalar("sigy")
            row[4] = r.scalar("h", 0.0)
        else:
            raise DeckError(f"unsupported material {r.keyword}")
        row[1] = r.scalar("e")
        row[2] = r.scalar("n")
        mrows.append(row)
        mnum[r.number] = i
    emat = np.full(ne, -1, dtype=np.int32)
    csmat = {cs.number: mnum[cs.scalar("material", cast=int)] for cs in css if cs.has("material")}
    for i, r in enumerate(elems):
        if r.has("mat"):
            emat[i] = mnum[r.scalar("mat", cast=int)]
        elif r.has("crosssect") and r.scalar("crosssect", cast=int) in csmat:
            # the element's cross section names the material (Element::giveMaterial
            # through SimpleCrossSection, simplecrosssection.C)
            emat[i] = csmat[r.scalar("crosssect", cast=int)]
    for cs in css:
        if cs.keyword != "simplecs":
            raise DeckError(f"unsupported cross section {cs.keyword}")
        if cs.has("material") and cs.has("set"):
            for e in setd[cs.scalar("set", cast=int)].get("elements", []):
                emat[e] = mnum[cs.scalar("material", cast=int)]
    if (emat < 0).any():
        raise DeckError("elements without material")
    # load-time functions
    ltf = {}
    for r in ltfs:
        if r.keyword == "constantfunction":
            ltf[r.number] = LoadTimeFunction("constantfunction", value=r.scalar("f(t)"))
        elif r.keyword == "piecewiselinfunction":
            ltf[r.number] = LoadTimeFunction("piecewise", t=r.array("t"), f=r.array("f(t)"))
        else:
            raise DeckError(f"unsupported load-time function {r.keyword}")
    # boundary conditions and loads
    bc_mask = np.zeros((nd, 3), dtype=bool)
    blist, llist, bodies, nodal = [], [], [], []
    for r in bcs:
        f = r.scalar("loadtimefunction", cast=int)
        if r.keyword == "boundarycondition":
            dofs = r.array("dofs", int)
            vals = r.array("values")
            nodes_s = setd[r.scalar("set", cast=int)].get("nodes", [])
            for dof in dofs:
                bc_mask[nodes_s, dof - 1] = True
            blist.append(dict(nodes=np.array(nodes_s, dtype=np.int64), dofs=dofs, values=vals, ltf=f,
                              number=r.number))
        elif r.keyword == "constantsurfaceload":
            dofs = r.array("dofs", int)
            comps = r.array("components")
            tr = np.zeros(3)
            for dof, c in zip(dofs, comps):
                tr[dof - 1] = c
            faces = setd[r.scalar("set", cast=int)]["faces"]
            llist.append(dict(faces=faces, traction=tr, ltf=f))
        elif r.keyword == "nodalload":
            # NodalLoad (nodalload.C): Components on the listed dofs of every
            # node of the set, times the load-time function; on a prescribed
            # dof it enters the reactions as an external load
            dofs = r.array("dofs", int)
            comps = r.array("components")
            if len(comps) != len(dofs):
                raise DeckError("NodalLoad: one component per listed dof")
            nodes_s = np.array(setd[r.scalar("set", cast=int)].get("nodes", []), dtype=np.int64)
            nodal.append(dict(nodes=nodes_s, dofs=dofs, comps=comps, ltf=f))
        elif r.keyword == "deadweight":
            # deadweight.C: Components are the force per unit mass in the
            # global c.s.; the element set carries it
            comps = np.asarray(r.array("components"), dtype=np.float64)
            if comps.size != 3:
                raise DeckError("DeadWeight needs three components in 3d")
            el = np.array(setd[r.scalar("set", cast=int)].get("elements", []), dtype=np.int64)
            bodies.append(dict(elements=el, b=comps, ltf=f))
        else:
            raise DeckError(f"unsupported boundary condition / load {r.keyword}")
    mat = np.array(mrows)
    deck = Deck(str(path), etype, coords, conn, emat, mat, bc_mask, np.zeros((nd, 3)))
    tol, checks = _parse_checks(checks_txt)
    prob = DeckProblem(deck, engng.keyword, engng.scalar("nsteps", 1, int), engng, blist, llist, ltf,
                       node_numbers, elem_numbers, checks, tol, bodies, np.array(dens), nodal)
    prob.output_file = lines[0].strip()
    prob.metasteps = metasteps
    if metasteps:
        # one metastep is what the drivers here run: its step count and
        # increment replace the engineering model's
        if len(metasteps) > 1:
            raise DeckError("decks with more than one metastep are not supported")
        prob.nsteps = metasteps[0].scalar("nsteps", prob.nsteps, int)
        if metasteps[0].has("deltat"):
            prob.engng = Record(engng.line + f" deltat {metasteps[0].scalar('deltat')}")
    prob.description = lines[1].strip()
    return prob


def evaluate_checks(prob: DeckProblem, results: dict):
    """results[k] = dict(u_node (nnode x 3 total displacements),
    reactions (node, dof) -> value, gp_sig / gp_eps (nelem x ngp x 6)).
    Returns a list of (check, computed, ok)."""
    nid = {n: i for i, n in enumerate(prob.node_numbers)}
    eid = {n: i for i, n in enumerate(prob.elem_numbers)}
    out = []
    for c in prob.checks:
        r = results.get(c.tstep)
        if r is None:
            continue
        if c.kind == "node":
            v = r["u_node"][nid[c.number], c.dof - 1]
        elif c.kind == "reaction":
            v = r["reactions"].get((nid[c.number], c.dof), 0.0)
        elif c.kind == "loadlevel":
            if "load_level" not in r:
                raise DeckError("a #LOADLEVEL check needs a load-level (arc-length) analysis")
            v = r["load_level"]
        else:
            arr = r["gp_sig"] if c.keyword == 1 else r["gp_eps"] if c.keyword == 4 else None
            if arr is None:
                raise DeckError(f"unsupported ELEMENT keyword {c.keyword}")
            v = arr[eid[c.number], c.gp - 1, c.component - 1]
        out.append((c, float(v), abs(v - c.value) <= prob.tolerance))
    return out
