"""The C-ABI library loads (no GPU needed) and exports every function that
include/umhip.h declares; host-side tables match the oracle."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "umhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(um_[a-z0-9_]+)\s*\(", src)))


def _lib():
    from umagnus_amd import build, umhip
    if not os.path.exists(umhip.LIB_PATH):
        build.build()
    return umhip.load()


def test_header_and_binding_agree():
    from umagnus_amd import umhip
    assert _declared() == sorted(umhip.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_calls_without_gpu_fail_loudly():
    """No CPU fallback: with no device the library reports an error."""
    import torch
    if torch.cuda.is_available():
        return
    lib = _lib()
    lib.um_last_error.restype = ctypes.c_char_p
    rc = lib.um_init(0)
    assert rc != 0 and lib.um_last_error()


def test_exchange_lut_matches_oracle(oracle):
    from umagnus_amd.params import exchange_lut
    vals = {0: 1.3e-11, 1: 2e-11, 2: -1e-11, 7: 5e-12}
    got = exchange_lut(vals, dtype=np.float64)
    want = oracle.exchange_lut(vals)
    assert np.array_equal(got, want)


def test_field_args_layout_matches_header(tmp_path):
    """The ctypes mirror of um_field_args has the C compiler's field offsets
    and size (a C program including the header prints them)."""
    import subprocess
    from umagnus_amd import umhip
    fields = [f[0] for f in umhip.um_field_args._fields_]
    cnames = ["lambda" if f == "lambda_" else f for f in fields]
    src = ["#include <stddef.h>", "#include <stdio.h>", '#include "umhip.h"', "int main(void) {",
           '  printf("%zu\\n", sizeof(um_field_args));']
    src += [f'  printf("%zu\\n", offsetof(um_field_args, {c}));' for c in cnames]
    src += ["  return 0;", "}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(umhip.um_field_args)
    for name, off in zip(fields, out[1:]):
        assert getattr(umhip.um_field_args, name).offset == off, name


def _split_top(s):
    """Split an argument list at top-level commas."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def _call_args(text, start):
    depth, i = 1, start
    while depth:
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
        i += 1
    return text[start:i - 1]


def test_integration_go_stubs_match_header():
    """Every C.um_* call in the cgo stubs of INTEGRATION.md names a declared
    function with the declared number of arguments, and every C.UM_* constant
    exists (the stubs cannot be compiled here: no Go toolchain)."""
    hdr = open(os.path.join(ROOT, "include", "umhip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    nparams = {}
    for mt in re.finditer(r"\b(um_[a-z0-9_]+)\s*\(", hdr):
        if hdr[:mt.start()].rstrip().endswith("typedef"):
            continue
        args = _call_args(hdr, mt.end())
        nparams[mt.group(1)] = 0 if args.strip() in ("", "void") else len(_split_top(args))
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    go = "\n".join(re.findall(r"```go\n(.*?)```", doc, flags=re.S))
    types = set(re.findall(r"\(\s*\*\s*(um_[a-z0-9_]+)\s*\)", hdr)) | set(re.findall(r"}\s*(um_[a-z0-9_]+)\s*;", hdr))
    calls = [mt for mt in re.finditer(r"\bC\.(um_[a-z0-9_]+)\(", go) if mt.group(1) not in types]
    assert len(calls) > 20
    for mt in calls:
        name = mt.group(1)
        assert name in nparams, name
        n = len(_split_top(_call_args(go, mt.end())))
        assert n == nparams[name], (name, n, nparams[name])
    for const in set(re.findall(r"\bC\.(UM_[A-Z0-9_]+)", go)):
        assert re.search(r"\b" + const + r"\b", hdr), const
