"""The reference's own annealing programs on libodp_hip.

sample/SA/{general,general_s,grid,grid_s}.c of the reference, compiled unchanged
against the reference's include/odp.h (oracle/Makefile, `make sa`), run once
linked with the reference library (CPU) and once linked with libodp_hip (every
ODP_Set_aspl on the GPU).  Start graph, moves, restores and acceptance draws
all come from the same libc random() stream, so the two runs must print the
same trajectory -- every progress line, the best diameter, sum and ASPL -- and
differ only in the timing line.
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "_ref")

CASES = [
    ("general", "-N 128 -D 4 -s 7 -n 400 -w 50 -c 0.22"),
    ("general", "-N 256 -D 6 -s 3 -n 300 -w 30 -c 0.2"),
    ("general", "-N 96 -D 3 -s 2 -n 300 -w 40 -c 0.2 -Y"),            # hill climbing
    ("general", "-N 128 -D 4 -s 11 -n 200"),                          # temperature found by its own short run
    ("general_s", "-N 96 -D 4 -S 4 -s 5 -n 300 -w 50 -c 0.2"),
    ("general_s", "-N 120 -D 3 -S 2 -s 9 -n 300 -w 50 -c 0.2 -A"),   # ASPL priority
    ("grid", "-W 10 -H 10 -D 4 -L 3 -s 5 -n 300 -w 50 -c 0.2"),
    ("grid_s", "-W 12 -H 12 -D 4 -L 3 -S 4 -s 5 -n 300 -w 50 -c 0.2"),
    ("grid_s", "-W 10 -H 8 -D 3 -L 4 -S 2 -s 6 -n 300 -w 50 -c 0.2"),
]


def run(prog, args, tmp):
    out = subprocess.run([os.path.join(REF, prog), *args.split()], capture_output=True, text=True, timeout=120,
                         cwd=tmp)
    assert out.returncode == 0, (prog, out.stderr[-2000:])
    return [ln for ln in out.stdout.splitlines() if not ln.startswith("Time")]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0] + ":" + c[1].replace(" ", ""))
def test_reference_annealing_program(case, tmp_path):
    prog, args = case
    if not (os.path.exists(os.path.join(REF, f"sa_{prog}_ref")) and os.path.exists(os.path.join(REF, f"sa_{prog}_hip"))):
        pytest.skip("reference annealing programs not built (oracle/Makefile `sa` needs the reference sources)")
    env_profile = os.environ.pop("ODP_PROFILE", None)
    try:
        ref = run(f"sa_{prog}_ref", args, tmp_path)
        hip = run(f"sa_{prog}_hip", args, tmp_path)
    finally:
        if env_profile is not None:
            os.environ["ODP_PROFILE"] = env_profile
    assert any(ln.startswith("ASPL ") for ln in ref), ref[-5:]
    diff = [(i, a, b) for i, (a, b) in enumerate(zip(ref, hip)) if a != b]
    assert len(ref) == len(hip) and not diff, (len(ref), len(hip), diff[:3])


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# the reference's sample programs (sample/<prog>.c): the edge file they read, or
# the file they write
SAMPLES = [
    ("general", os.path.join(GOLDEN, "general", "n16d4.edges"), None),
    ("general", os.path.join(GOLDEN, "general", "n256d10.edges"), None),
    ("grid", os.path.join(GOLDEN, "grid", "w10h10d4r4g4.edges"), None),
    ("regular_general", None, None),
    ("non_regular_general", None, None),
    ("random_general", None, "general.edges"),
    ("random_grid", None, "grid.edges"),
]


@pytest.mark.parametrize("case", SAMPLES, ids=lambda c: c[0] + ("-" + os.path.basename(c[1]) if c[1] else ""))
def test_reference_sample_program(case, tmp_path):
    """sample/<prog>.c linked with the reference library and with libodp_hip:
    the same printed results and the same written edge file."""
    prog, infile, outfile = case
    if not (os.path.exists(os.path.join(REF, f"smp_{prog}_ref")) and os.path.exists(os.path.join(REF, f"smp_{prog}_hip"))):
        pytest.skip("reference sample programs not built (oracle/Makefile `sa` needs the reference sources)")
    outs, files = [], []
    for lib in ("ref", "hip"):
        d = tmp_path / lib
        d.mkdir()
        args = [infile] if infile else []
        r = subprocess.run([os.path.join(REF, f"smp_{prog}_{lib}"), *args], capture_output=True, text=True,
                           timeout=120, cwd=d)
        assert r.returncode == 0, (prog, lib, r.stderr[-2000:])
        outs.append(r.stdout)
        if outfile:
            files.append((d / outfile).read_text())
    assert "ASPL" in outs[0] and outs[0] == outs[1]
    if outfile:
        assert files[0] and files[0] == files[1]


def test_libc_random_stream_untouched():
    """Evaluations leave the caller's libc random() stream where it was: draws
    after seeding are the same with or without Init / Set_aspl / Finalize (and
    the other device entry points) in between."""
    import ctypes
    import odp_amd as odp
    libc = ctypes.CDLL(None)
    libc.srand.argtypes = [ctypes.c_uint]
    libc.rand.restype = ctypes.c_int
    libc.srand(7)
    want = [libc.rand() for _ in range(5)]
    libc.srand(7)
    n, d = 300, 5
    adj = odp.generate_regular(n, d, 3)
    odp.init_aspl_general(n, d)
    odp.set_aspl(adj)
    odp.upload_adjacency(adj)
    odp.set_aspl_resident()
    odp.set_plan(plan=odp.PLAN_LEVEL, kw=2, upw=4, seg_major=0, xcd=1)
    odp.set_aspl(adj)
    odp.finalize_aspl()
    odp.batch_eval(odp.pack([adj, adj]))
    got = [libc.rand() for _ in range(5)]
    assert got == want


def _profile_block(text):
    lines = text.splitlines()
    i = lines.index("------ Profile for SET_ASPL ------")
    j = lines.index("--------- End of Profile ---------")
    return lines[i:j + 1]


def test_profile_report_layout(tmp_path):
    """ODP_PROFILE=1: the report printed at ODP_Finalize_aspl has the reference's
    layout (utils.c:1268-1294) line for line.  Labels, the call count, the
    symmetry factor, the process and thread counts are equal to the reference
    program's; the algorithm line names this library's flavour; the memory line
    is the reference's GPU-branch estimate (utils.c:1206-1236 with is_cpu =
    false, as aspl_cuda.cu:137 computes it) to three decimals."""
    prog = "general"
    if not (os.path.exists(os.path.join(REF, f"smp_{prog}_ref")) and os.path.exists(os.path.join(REF, f"smp_{prog}_hip"))):
        pytest.skip("reference sample programs not built (oracle/Makefile `sa` needs the reference sources)")
    infile = os.path.join(GOLDEN, "general", "n256d10.edges")
    env = dict(os.environ, ODP_PROFILE="1")
    env.pop("ODP_ASPL", None)
    blocks = []
    for lib in ("ref", "hip"):
        r = subprocess.run([os.path.join(REF, f"smp_{prog}_{lib}"), infile], capture_output=True, text=True,
                           timeout=120, cwd=tmp_path, env=env)
        assert r.returncode == 0, r.stderr[-2000:]
        blocks.append(_profile_block(r.stdout))
    ref, hip = blocks
    assert len(ref) == len(hip) == 12
    label = lambda ln: ln.split("=")[0]
    assert [label(a) for a in ref] == [label(b) for b in hip]
    value = lambda blk, name: next(ln.split("=", 1)[1].strip() for ln in blk if ln.startswith(name))
    for name in ("Number of Times", "Symmetries", "Num of Procs", "Num of Threads"):
        assert value(ref, name) == value(hip, name), name
    assert value(ref, "Algorithm") == "MATRIX (SERIAL)" and value(hip, "Algorithm") == "MATRIX (HIP)"
    for name in ("Total Time", "Average Time"):
        assert re.fullmatch(r"[0-9]+\.[0-9]{6} sec\.", value(hip, name)), name
    kind_exe = os.path.join(REF, "odp_ref_kind")
    if os.path.exists(kind_exe):
        r = subprocess.run([kind_exe], input="256 10 0 1 0\n", capture_output=True, text=True, timeout=60)
        mem = float.fromhex(r.stdout.split()[1])
        assert value(hip, "Memory Usage") == f"{mem:.3f} MB"
