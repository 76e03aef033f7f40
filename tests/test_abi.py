"""The C-ABI library: built, loadable without a GPU, exporting every entry point the
public header declares, and not linked against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import fvg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fvg.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "fvg_assemble" in names and "fvg_cg" in names and "fvg_bicgstab" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    out = subprocess.check_output(["nm", "-D", "--defined-only", fvg.LIB_PATH]).decode()
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    L = fvg.lib()
    for n in declared():
        assert getattr(L, n) is not None


def test_library_does_not_link_the_oracle():
    out = subprocess.check_output(["readelf", "-d", fvg.LIB_PATH]).decode()
    assert "oracle" not in out
    syms = subprocess.check_output(["nm", "-D", fvg.LIB_PATH]).decode()
    assert "orc_" not in syms


def test_context_fails_loudly_without_gpu_or_works_with_one():
    try:
        c = fvg.Context(0)
    except fvg.FvgError as e:
        assert e.status == 2   # FVG_E_HIP: no device on this machine
    else:
        c.close()


def test_status_strings():
    L = fvg.lib()
    assert L.fvg_status_string(4).decode() == "zero pivot in ILU"
    assert L.fvg_status_string(0).decode() == "ok"


def test_term_program_compiles_like_the_reference_algebra():
    """-laplacian(k) - source(f) -> postfix [LAP, NEG, SRC, SUB] with deduplicated fields."""
    k = np.ones(5)
    f = np.zeros(3)
    eq = -fvg.laplacian(fvg.Orthogonal(), k) - fvg.source(f)
    ops, ff, cf = eq.compile()
    assert [o[0] for o in ops] == [fvg.OP_LAPLACIAN_ORTHOGONAL, fvg.OP_NEG, fvg.OP_SOURCE, fvg.OP_SUB]
    assert len(ff) == 1 and len(cf) == 1
    flux = np.arange(5.0)
    eq = fvg.convection(fvg.Upwind(flux), flux) - fvg.laplacian(fvg.Orthogonal(), k)
    ops, ff, cf = eq.compile()
    assert ops[0][:3] == (fvg.OP_CONVECTION_UPWIND, 0, 0) and len(ff) == 2


def test_backward_and_user_term_compile():
    """time(Backward) compiles to its opcode plus a parameter word carrying dt_2, with the
    three densities in consecutive cell slots; a user term puts its face arrays (li, lj, r)
    and cell arrays (l, r) in consecutive slots, -1 for a part left at the zero default."""
    p1, p2, r = np.zeros(4), np.ones(4), np.full(4, 2.0)
    eq = fvg.time(fvg.Backward(p1, p2, 0.1, 0.2, r, r, r)) - fvg.source(p1)
    ops, ff, cf = eq.compile()
    assert ops == [(fvg.OP_TIME_BACKWARD, 0, 1, 2, 0.1), (fvg.OP_PARAM, 0, 0, 0, 0.2),
                   (fvg.OP_SOURCE, 0, 0, 0, 0.0), (fvg.OP_SUB, 0, 0, 0, 0.0)]
    assert len(cf) == 5 and all(cf[k] is r for k in (2, 3, 4))
    with pytest.raises(fvg.FvgError):
        fvg.Backward(p1, p2, 0.1, 0.2, r)
    with pytest.raises(fvg.FvgError):
        fvg.time(object())
    fa = (np.zeros(6), np.ones(6), np.zeros(6))
    eq = fvg.user_term(face_terms=fa) + fvg.user_term(face_terms=fa, cell_terms=(p1, p2)) + fvg.user_term(
        cell_terms=(p2, p1))
    ops, ff, cf = eq.compile()
    assert ops == [(fvg.OP_USER_TERM, 0, -1, 0, 0.0), (fvg.OP_USER_TERM, 3, 0, 0, 0.0), (fvg.OP_ADD, 0, 0, 0, 0.0),
                   (fvg.OP_USER_TERM, -1, 2, 0, 0.0), (fvg.OP_ADD, 0, 0, 0, 0.0)]
    assert len(ff) == 6 and len(cf) == 4
    with pytest.raises(fvg.FvgError):
        fvg.user_term(face_terms=fa[:2])


def test_sweep_launch_rule():
    """The co-resident sweeps launch ordinarily only while their context is the only one
    alive on the device in the process, none of its batches is in flight and none of its
    sweep waits has timed out; otherwise cooperatively.  Knob values 1 and 2 force the
    mode; arguments out of range give 0."""
    R = fvg.sweep_launch_rule
    assert R(0, 1, 0, False) == 1
    assert R(0, 2, 0, False) == 2          # another context of the process
    assert R(0, 1, 1, False) == 2          # a batch of this context in flight
    assert R(0, 1, 0, True) == 2           # a wait of this context timed out before
    assert R(0, 3, 2, True) == 2
    assert R(1, 4, 3, True) == 1 and R(2, 1, 0, False) == 2
    assert R(3, 1, 0, False) == 0 and R(-1, 1, 0, False) == 0
    assert R(0, -1, 0, False) == 0 and R(0, 1, -1, False) == 0
