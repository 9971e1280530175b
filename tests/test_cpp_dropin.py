"""The C++ drop-in as a reference caller uses it, and the reference CLI.

* tests/cpp/loader_check (no GPU): the CSV loader of include/txtReader.hpp
  against a restatement of the reference's rule (txtReader.hpp:35-40, 77-82:
  a point equal to the previous kept one on x, y, or x, y, z, is dropped), and
  the Curve helpers the reference declares (recenter, operator<<).
* tests/cpp/dropin_check (GPU): a C++ program that builds Curve objects and
  calls compute_distance_parallel (wrapper.hpp:4) and distance_cuda
  (frechet.hpp:14), linked against librfd.so, on the 2,000 x 2,500 pair,
  against the oracle under the reference's stop rule (absolute FLT_EPSILON,
  compute_distance_wrapper.cpp:6) and decision path (two-phase when its
  arrays fit, .cu:87-96; fused under forceOversize).
* robotfd_amd/fd (GPU): the CLI of main.cpp:13-44 on CSV files holding
  consecutive duplicates and a longer first curve.
"""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from robotfd_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
FLT_EPS = float(np.finfo(np.float32).eps)


def need(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: run __graft_entry__.build()")
    return path


def reference_load(lines, dims):
    """txtLoader2/3 (txtReader.hpp:6-88): comma-separated doubles per line;
    a line without any double is skipped; a point equal to the last kept
    point on its first `dims` coordinates is dropped."""
    out = []
    for line in lines:
        vals = []
        for tok in line.split(","):
            m = re.match(r"\s*[-+]?(\d+\.?\d*([eE][-+]?\d+)?|\.\d+([eE][-+]?\d+)?)", tok)
            if m:
                vals.append(float(m.group(0)))
        if not vals:
            continue
        if out and all(out[-1][k] == vals[k] for k in range(dims)):
            continue
        out.append(vals)
    return out


def hexfloats(line):
    return [float.fromhex(t) for t in line.split()]


@pytest.mark.parametrize("dims", [2, 3])
def test_loader_and_curve_helpers(tmp_path, dims):
    exe = need(os.path.join(CPP, "loader_check"))
    rng = np.random.default_rng(7 + dims)
    pts = np.round(rng.standard_normal((30, 4)) * 100, 3)
    lines = []
    for k, p in enumerate(pts):
        row = ",".join(repr(float(v)) for v in p[:dims + (k % 2)])  # some lines carry an extra column
        lines.append(row)
        if k % 5 == 0:
            lines.append(row)  # exact consecutive duplicate: dropped
        if k % 7 == 0:
            # equal on the first dims coordinates, different beyond: dropped too
            lines.append(",".join(repr(float(v)) for v in p[:dims]) + ",12345.5")
        if k % 6 == 3:
            lines.append("")  # a line without values: skipped
    lines.append(lines[0])  # equal to an earlier, not the last, point: kept
    path = tmp_path / "curve.csv"
    path.write_text("\n".join(lines) + "\n")
    out = subprocess.run([exe, str(path), str(dims)], capture_output=True, text=True, check=True).stdout
    got_lines = out.splitlines()
    n = int(got_lines[0].split()[1])
    got = [hexfloats(l) for l in got_lines[1:1 + n]]
    want = reference_load(lines, dims)
    assert len(got) == len(want) < len(lines)
    for g, w in zip(got, want):
        assert g == w
    curve_txt = got_lines[2 + n:2 + 2 * n]
    assert [list(map(float, l.split(","))) for l in curve_txt] == want
    assert got_lines[-1] == "FIRST 0x0p+0 0x0p+0"


def write_bin(path, P, Q):
    with open(path, "wb") as f:
        f.write(struct.pack("<qqq", len(P), len(Q), P.shape[1]))
        f.write(np.ascontiguousarray(P, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(Q, dtype="<f8").tobytes())


@pytest.mark.gpu
def test_cpp_dropin_caller_matches_oracle(tmp_path):
    exe = need(os.path.join(CPP, "dropin_check"))
    P, Q = W.small_2d(seed=0)
    write_bin(tmp_path / "c1.bin", P, Q)
    out = subprocess.run([exe, str(tmp_path / "c1.bin"), "1e-6"], capture_output=True, text=True,
                         timeout=120, check=True).stdout
    res = [l for l in out.splitlines() if l.startswith("RESULT")]
    a, b, c = hexfloats(res[0][len("RESULT"):])
    # Without forceOversize the 2,000 x 2,500 pair takes the two-phase path
    # (its interval arrays, 288 MB, fit in device memory); with it, the fused.
    two = O.distance(P, Q, stop_mode=0, eps=FLT_EPS, form=1)
    fused = O.distance(P, Q, stop_mode=0, eps=FLT_EPS, form=0)
    fine = O.distance(P, Q, stop_mode=0, eps=1e-6, form=0)
    assert a == two["distance"]
    assert b == fused["distance"]
    assert c == fine["distance"]
    # The progress lines of frechet.cpp:21,31,34,65-66, once per call.
    assert out.count("2000, 2500\n") == 3
    assert out.count("Naive LB: ") == 3 and out.count("Iteration: ") == 3
    iters = [int(x) for x in re.findall(r"Iteration: (\d+)", out)]
    assert iters == [two["iterations"], fused["iterations"], fine["iterations"]]
    results = re.findall(r"Result: ([-0-9.]+)", out)
    assert results == [f"{two['distance']:.7f}", f"{fused['distance']:.7f}", f"{fine['distance']:.7f}"]


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [2, 3])
def test_cli_dedupes_and_puts_the_shorter_curve_first(tmp_path, dims):
    exe = need(os.path.join(ROOT, "robotfd_amd", "fd"))
    rng = np.random.default_rng(40 + dims)
    long_c = np.cumsum(rng.standard_normal((1000, dims)), axis=0)
    short_c = long_c[::2][:480] + rng.standard_normal((480, dims)) * 0.3
    # Consecutive duplicates in both files (the loader drops them).
    dup_l = np.repeat(long_c, np.where(np.arange(1000) % 9 == 0, 2, 1), axis=0)
    dup_s = np.repeat(short_c, np.where(np.arange(480) % 11 == 0, 3, 1), axis=0)
    f1, f2 = tmp_path / "long.txt", tmp_path / "short.txt"
    f1.write_text("\n".join(",".join(repr(float(v)) for v in p) for p in dup_l) + "\n")
    f2.write_text("\n".join(",".join(repr(float(v)) for v in p) for p in dup_s) + "\n")
    for force, form in (("0", 1), ("1", 0)):
        out = subprocess.run([exe, str(f1), str(f2), str(dims), force], capture_output=True, text=True,
                             timeout=120, check=True).stdout
        lines = out.splitlines()
        assert lines[0] == force                      # main.cpp:29
        assert lines[1] == "480, 1000"                # shorter first (main.cpp:34), deduplicated
        ref = O.distance(short_c, long_c, stop_mode=0, eps=FLT_EPS, form=form)
        assert f"Result: {ref['distance']:.7f}" in lines
        assert f"Iteration: {ref['iterations']}" in lines
