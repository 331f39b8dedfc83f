"""CPU checks of the native library: it loads, exports every symbol the header declares, and
its expression compiler agrees with the oracle's parser (no GPU needed)."""
import math
import os
import re

import numpy as np
import pytest

import oracle
from pygranite_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "granite_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s+(gr_\w+)\s*\(", src, re.M)))


def test_library_exports_every_header_symbol():
    L = capi.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), f"{s} declared in granite_hip.h but not exported"
    assert sorted(capi.EXPORTS) == syms


def test_abi_version_and_device_count_without_gpu():
    L = capi.lib()
    import re
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "granite_hip.h")).read()
    assert L.gr_abi_version() == int(re.search(r"#define GR_ABI_VERSION (\d+)", hdr).group(1))
    n = capi.device_count()
    assert n >= 0


def test_ctx_create_fails_cleanly_without_gpu():
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.GraniteError):
        capi.Context(0)


EXPRS = ["sqrt(u*u+v*v+w*w)", "w*1000/(abs(u)+abs(v)+1)", "1-2-3", "x*y+z/2-cv",
         "atan2(y, x) + sin(z)*cos(x) - tan(0.25)", "exp(-1.5)*ln(2.5)+asin(0.5)-acos(0.25)",
         "atan(x) / (1 + u*u)", "((x))", "2*-3+x", "temp*rho - c_2", "(u+v)*(u-v)/(w+10)",
         "u*u+v*v", "x*y+z*w+u*v", "1+u*u+v*v", "u*u+v*v+w*w+x*x", "x*(u*u+v*v)",
         "u*v+w*temp", "sqrt(u*u+v*v)-sqrt(w*w+x*x+y*y)"]


@pytest.mark.parametrize("src", EXPRS)
def test_expression_compiler_matches_oracle(src):
    rng = np.random.default_rng(abs(hash(src)) % 2 ** 32)
    consts = ["temp", "rho", "c_2"]
    env = {k: 13 + i for i, k in enumerate(consts)}
    prog = capi.compile_expr(src, dim=3, curvature=3, constants=consts, allow_abs=True)
    assert prog.n > 0 and prog.depth <= 8
    for _ in range(20):
        ref = rng.uniform(0.1, 2.0, 19)
        got = capi.eval_expr_host(prog, ref)
        f32, f64 = oracle.eval_expr(src, ref, env, allow_abs=True)
        assert got == pytest.approx(f64, rel=2e-5, abs=1e-6)


@pytest.mark.parametrize("src,kw", [
    ("velocity", {}), ("f0+1", {}), ("c1*2", {}), ("abs(x)", {}), ("1.", {}), (".5", {}),
    ("x y", {}), ("", {}), ("sqrt(x", {}), ("nope", {}), ("z+1", {"dim": 2}),
    ("w", {"dim": 2}), ("cv", {"curvature": 0}), ("(((((((((x)))))))))+" * 1 + "1", {}),
])
def test_expression_compiler_rejects(src, kw):
    if src.startswith("((((("):
        # deep parentheses alone do not deepen the stack: this one is valid
        capi.compile_expr(src, **kw)
        return
    with pytest.raises(capi.GraniteError):
        capi.compile_expr(src, **kw)


def test_expression_stack_limit():
    deep = "1+(" * 8 + "1" + ")" * 8
    with pytest.raises(capi.GraniteError):
        capi.compile_expr(deep)
    ok = "1+(" * 6 + "1" + ")" * 6
    capi.compile_expr(ok)


def test_expression_postfix_layout():
    """sqrt(u*u+v*v+w*w): operands before operators, 12 instructions like the reference's 12
    flattened nodes."""
    p = capi.compile_expr("sqrt(u*u+v*v+w*w)")
    ops = [p.code[i] & 0xff for i in range(p.n)]
    args = [p.code[i] >> 8 for i in range(p.n)]
    assert p.n == 12
    assert ops == [1, 1, 5, 1, 1, 5, 3, 1, 1, 5, 3, 7]
    assert [a for o, a in zip(ops, args) if o == 1] == [16, 16, 17, 17, 18, 18]


def _random_expr(rng, depth):
    """A random expression of the AdditionalCompute language (registers, constants, + - * /,
    the functions) at most `depth` levels deep."""
    if depth == 0 or rng.random() < 0.25:
        if rng.random() < 0.7:
            return str(rng.choice(["x", "y", "z", "cv", "u", "v", "w"]))
        return f"{rng.uniform(0.1, 3.0):.3f}"
    r = rng.random()
    if r < 0.55:
        op = str(rng.choice(["+", "-", "*", "/"]))
        return f"({_random_expr(rng, depth - 1)}{op}{_random_expr(rng, depth - 1)})"
    if r < 0.65:
        return f"atan2({_random_expr(rng, depth - 1)}, {_random_expr(rng, depth - 1)})"
    fn = str(rng.choice(["sqrt", "exp", "ln", "sin", "cos", "tan", "atan", "abs"]))
    return f"{fn}({_random_expr(rng, depth - 1)})"


def test_device_form_equals_postfix_bit_for_bit():
    """The device form of a program (leaf operands folded into their operations, the two top
    stack values in registers, reversed forms keeping operand order) computes exactly what the
    postfix bytecode computes, on the bench expression, the listed expressions and 400 random
    ones; and agrees with the oracle's own parser in FP32 bit for bit."""
    rng = np.random.default_rng(7)
    srcs = list(EXPRS) + [_random_expr(rng, 4) for _ in range(400)]
    consts = ["temp", "rho", "c_2"]
    env = {k: 13 + i for i, k in enumerate(consts)}
    checked = 0
    for src in srcs:
        try:
            prog = capi.compile_expr(src, dim=3, curvature=3, constants=consts, allow_abs=True)
        except capi.GraniteError:
            continue  # too long / too deep for the limits
        for _ in range(6):
            ref = rng.uniform(-2.0, 2.0, 19).astype(np.float32)
            a = np.float32(capi.eval_expr_host(prog, ref))
            b = np.float32(capi.eval_expr_device_form(prog, ref))
            f32, _ = oracle.eval_expr(src, ref, env, allow_abs=True)
            assert (a == b) or (np.isnan(a) and np.isnan(b)), src
            assert (a == np.float32(f32)) or (np.isnan(a) and np.isnan(f32)), src
        checked += 1
    assert checked > 300


def test_job_bytes_hand_counted():
    """gr_job_bytes without a GPU, against byte counts derived by hand from the job layout:
    SoA positions, status and abort step, curvature / expression state, the saved-column ring
    (capacity x particles), two AoS scratch arrays, and a reorder's index map plus one shadow
    copy per state array and of the map."""
    d = capi.JobDesc()
    d.dim, d.integrator, d.windfield_mode, d.dt, d.dtd = 3, 1, 1, 0.01, 1.0
    n = 1000
    # positions 3x4, status 4, abort step 4 -> 20 B; scratch 2x12 -> 24 B
    assert capi.job_bytes(d, n, 1, reorder=False) == 44 * n
    # 5 state arrays: index map + 6 shadow copies
    assert capi.job_bytes(d, n, 1, reorder=True) == 44 * n + (5 + 2) * 4 * n
    d.curvature, d.save_interval, d.nprog = 3, 1, 1
    d.prog[0] = capi.compile_expr("sqrt(u*u+v*v+w*w)", dim=3, curvature=3)
    # state: p, q (2x12), status, abort, total, curv, expr (5x4) = 44; ring of 10 columns:
    # positions 12, curvature 4, expression 4 -> 200; scratch 24
    assert capi.job_bytes(d, n, 10, reorder=False) == (44 + 200 + 24) * n
    d.dim = 2
    assert capi.job_bytes(d, n, 10, reorder=False) == ((8 + 8 + 20) + 10 * (8 + 4 + 4) + 16) * n
