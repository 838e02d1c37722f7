"""The plug-in boundary compiles against the reference's own declarations.

For every reference declaration the plug-ins and the batched evaluator rely
on (cited by file:line), the mirror header under oofem_amd/host/ holds the
same declaration, compared with whitespace, comments and export macros
removed.  Runs where the reference tree is present; the host library's
build (g++ -fsyntax-only of the plug-in sources against the mirrors) is
checked everywhere.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "oofem_amd", "host")
REF = "/root/reference/src"

# (reference file, first line, last line, mirror header)
DECLS = [
    # EngngModel hooks and the members the plug-ins call
    ("core/engngm.h", 356, 356, "engngm.h"),            # Domain *giveDomain(int n);
    ("core/engngm.h", 527, 527, "engngm.h"),            # giveNumberOfDomainEquations
    ("core/engngm.h", 852, 852, "engngm.h"),            # updateInternalRHS
    ("core/engngm.h", 957, 958, "engngm.h"),            # assemble(SparseMtrx&, TimeStep*, MatrixAssembler&, ...)
    ("sm/EngineeringModels/structengngmodel.h", 167, 167, "structengngmodel.h"),
    ("core/assemblercallback.h", 189, 189, "engngm.h"),  # MatResponseMode rmode;
    ("core/assemblercallback.h", 192, 192, "engngm.h"),  # TangentAssembler(MatResponseMode m = TangentStiffness)
    ("core/element.h", 228, 228, "engngm.h"),           # giveLocationArray
    ("core/domain.h", 293, 294, "engngm.h"),            # giveElement / giveElements
    ("core/domain.h", 463, 463, "engngm.h"),            # giveNumberOfElements
    ("core/timestep.h", 211, 211, "engngm.h"),
    ("core/timestep.h", 213, 213, "engngm.h"),
    # InputRecord and the reading macros
    ("core/inputrecord.h", 59, 59, "inputrecord.h"),    # typedef const char *InputFieldType;
    ("core/inputrecord.h", 67, 67, "inputrecord.h"),    # IR_GIVE_FIELD
    ("core/inputrecord.h", 75, 75, "inputrecord.h"),    # IR_GIVE_OPTIONAL_FIELD
    ("core/inputrecord.h", 112, 112, "inputrecord.h"),  # giveRecordAsString
    ("core/inputrecord.h", 114, 114, "inputrecord.h"),  # giveLocation
    ("core/inputrecord.h", 127, 137, "inputrecord.h"),  # giveField(int / double / bool / string / FloatArray / IntArray)
    ("core/inputrecord.h", 192, 193, "inputrecord.h"),  # template giveOptionalField
    ("core/inputrecord.h", 231, 231, "inputrecord.h"),  # hasField
    ("core/dynamicinputrecord.h", 121, 126, "dynamicinputrecord.h"),   # setField
    ("core/iml/imlsolver.h", 48, 52, "imlsolver.h"),    # _IFT_IMLSolver_* keywords
    # SparseMtrx and SparseLinearSystemNM
    ("core/sparsemtrx.h", 94, 94, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 137, 137, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 143, 143, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 149, 149, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 160, 160, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 166, 166, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 198, 198, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 220, 220, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 231, 231, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 234, 236, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 239, 239, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 254, 254, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 265, 269, "sparsemtrx.h"),
    ("core/sparsemtrx.h", 279, 283, "sparsemtrx.h"),
    ("core/sparselinsystemnm.h", 66, 66, "sparsemtrx.h"),
    ("core/sparselinsystemnm.h", 72, 72, "sparsemtrx.h"),
    ("core/sparselinsystemnm.h", 81, 81, "sparsemtrx.h"),
    ("core/sparselinsystemnm.h", 91, 91, "sparsemtrx.h"),
]


def _norm(text):
    """Declarations only: comments, export macros and whitespace removed;
    bodies and the pure / override markers kept out of the comparison."""
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = text.replace("OOFEM_EXPORT", "")
    return re.sub(r"\s+", "", text)


def _signatures(text):
    """Split normalized text into declarations: each ends at ';' or at the
    '{' of an inline body; '=0', 'override' and 'NULL'/'nullptr' defaults
    are made uniform."""
    t = _norm(text)
    t = t.replace("=NULL", "=nullptr").replace("override", "").replace("=0;", ";")
    out, depth, cur = [], 0, ""
    for ch in t:
        if depth == 0 and ch in ";{":
            out.append(cur)
            cur = ""
            if ch == "{":
                depth = 1
            continue
        if depth > 0:
            depth += (ch == "{") - (ch == "}")
            continue
        if ch == "}":
            continue
        cur += ch
    if cur:
        out.append(cur)
    return [s for s in out if s]


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference tree is not present on this machine")
@pytest.mark.parametrize("ref,lo,hi,mirror", DECLS, ids=[f"{d[0]}:{d[1]}" for d in DECLS])
def test_mirror_declares_what_the_reference_declares(ref, lo, hi, mirror):
    with open(os.path.join(REF, ref)) as fh:
        lines = fh.read().split("\n")[lo - 1:hi]
    want = _signatures("\n".join(lines))
    assert want, (ref, lo, hi)
    with open(os.path.join(HOST, mirror)) as fh:
        have = _norm(fh.read()).replace("=NULL", "=nullptr").replace("override", "").replace("=0;", ";")
    for sig in want:
        # a '#define' line keeps its macro body; compare it whole
        assert sig in have, f"{ref}:{lo}-{hi}: '{sig}' not declared in oofem_amd/host/{mirror}"


def test_plugin_sources_compile_against_the_mirrors():
    """hipcsr.C, hippcg.C and the driver with its EngngModel overrides
    compile against the mirror headers (syntax and types only)."""
    srcs = [os.path.join(HOST, f) for f in ("hipcsr.C", "hippcg.C", "structmodel.C", "host_capi.C")]
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror=overloaded-virtual",
                        "-I", os.path.join(ROOT, "include"), *srcs], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
