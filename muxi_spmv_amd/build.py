"""Builds libmuxi_spmv_amd.so (gfx950) in-tree with hipcc.

Usage: python -m muxi_spmv_amd.build [--jobs N]
Every source under csrc/ is compiled as HIP for gfx950 and linked into one
shared library next to this file. Packed-FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``). When the package holds
trained selector weights (selector_weights.json), they are written to a
generated header that selector.cpp compiles in, so MUXI_KERNEL_AUTO plans use
the learned selector without any set-up call.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libmuxi_spmv_amd.so"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("MUXI_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    f"-I{BUILD}",
]

WEIGHTS = PKG / "selector_weights.json"
_KERNEL_IDS = {"scalar": 1, "vector": 2, "vwave": 3, "adaptive": 4, "merge": 5, "csr5": 6, "ell": 7, "xslab": 8, "lslab": 9}


def write_selector_header() -> Path | None:
    """Writes _build/selector_weights.inc from selector_weights.json (or removes a stale one)."""
    inc = BUILD / "selector_weights.inc"
    if not WEIGHTS.exists():
        inc.unlink(missing_ok=True)
        return None
    st = json.loads(WEIGHTS.read_text())
    if st["n_in"] not in (10, 11) or len(st["candidates"]) != st["n_out"]:
        raise RuntimeError(f"{WEIGHTS}: expected 10 or 11 inputs and one output per candidate")
    ks, ps = [], []
    for name in st["candidates"]:
        k, p = name.split(":")
        ks.append(_KERNEL_IDS[k])
        ps.append(int(p))

    def arr(ctype, name, vals, fmt):
        body = ",".join(fmt(v) for v in vals)
        return f"static const {ctype} {name}[] = {{{body}}};\n"

    f32 = lambda v: f"{float(v)!r}f"  # noqa: E731  (repr round-trips the float32 value)
    txt = ("// Generated by muxi_spmv_amd/build.py from selector_weights.json; do not edit.\n"
           f"static const int kSelIn = {st['n_in']}, kSelHidden = {st['hidden']}, kSelOut = {st['n_out']};\n"
           + arr("float", "kSelParams", st["params"], f32) + arr("float", "kSelMu", st["mu"], f32)
           + arr("float", "kSelSd", st["sd"], f32) + arr("int", "kSelKernel", ks, str)
           + arr("int", "kSelParam0", ps, str))
    if not inc.exists() or inc.read_text() != txt:
        inc.write_text(txt)
    return inc


SCHED = PKG / "sched_tables.json"


def write_sched_header() -> Path | None:
    """Writes _build/sched_tables.inc from sched_tables.json (or removes a stale one)."""
    inc = BUILD / "sched_tables.inc"
    if not SCHED.exists():
        inc.unlink(missing_ok=True)
        return None
    t = json.loads(SCHED.read_text())
    b, m = t["batch"], t["spmm"]
    nb_len, nb_rows = len(b["row_len_edges"]) + 1, len(b["rows_edges"]) + 1
    ns_len = len(m["path_edges"]) + 1
    if len(b["table"]) != nb_len * nb_rows or len(m["table"]) != len(m["k"]) * ns_len:
        raise RuntimeError(f"{SCHED}: table sizes do not match the bins")

    def arr(ctype, name, vals):
        return f"static const {ctype} {name}[] = {{{','.join(repr(v) for v in vals)}}};\n"

    txt = ("// Generated by muxi_spmv_amd/build.py from sched_tables.json; do not edit.\n"
           f"static const int kBatchRowLenBins = {nb_len}, kBatchRowsBins = {nb_rows};\n"
           + arr("double", "kBatchRowLenEdges", [float(v) for v in b["row_len_edges"]])
           + arr("double", "kBatchRowsEdges", [float(v) for v in b["rows_edges"]])
           + arr("int", "kBatchIpt", [int(e[0]) for e in b["table"]])
           + arr("int", "kBatchWgPerCu", [int(e[1]) for e in b["table"]])
           + f"static const int kSpmmKCount = {len(m['k'])}, kSpmmPathBins = {ns_len};\n"
           + arr("int", "kSpmmK", [int(v) for v in m["k"]])
           + arr("double", "kSpmmPathEdges", [float(v) for v in m["path_edges"]])
           + arr("int", "kSpmmIpt", [int(v) for v in m["table"]]))
    if not inc.exists() or inc.read_text() != txt:
        inc.write_text(txt)
    return inc


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, "-x", "hip", "-c", str(src), "-o", str(obj)] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    write_selector_header()
    write_sched_header()
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-o", str(tmp)] + [str(o) for o in objs] + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    return LIB


CPP_TESTS = [ROOT / "tests" / "cpp" / "test_spmv_known_answer.cpp", ROOT / "tests" / "cpp" / "test_formats_io.cpp",
             ROOT / "tests" / "cpp" / "test_csr5_fields.cpp", ROOT / "tests" / "cpp" / "test_benchmark_layer.cpp"]


# C++ programs beside the library: rocSPARSE timed outside a PyTorch process (bench.py runs it)
CPP_TOOLS = [PKG / "rocsparse_timing.cpp"]


def build_cpp_tests(verbose: bool = True) -> list[Path]:
    """Reference-style C++ callers compiled against include/muxi_spmv/muxi_spmv.hpp."""
    out = []
    for src in CPP_TESTS + CPP_TOOLS:
        exe = src.with_suffix("")
        rpath = "$ORIGIN" if src.parent == PKG else "$ORIGIN/../../muxi_spmv_amd"
        cmd = [HIPCC, "-std=c++17", "-O2", f"--offload-arch={ARCH}", f"-I{ROOT / 'include'}", str(src),
               f"-L{PKG}", "-lmuxi_spmv_amd", f"-Wl,-rpath,{rpath}", "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {src.name}\n{r.stdout}\n{r.stderr}")
        out.append(exe)
    if verbose:
        print(f"built {len(out)} C++ caller(s)", flush=True)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    build(a.jobs)
    build_cpp_tests()
    return 0


if __name__ == "__main__":
    sys.exit(main())
