"""Input/output formats of the drop-in boundary and the CLI's error paths
(none of these reach the GPU)."""
import os
import subprocess

import numpy as np
import pytest

import sanepic_amd
from sanepic_amd import synth

EXE = sanepic_amd.EXE_PATH


def test_segment_list_matches_generate_segfile():
    # helpers/generate_segfile.c:22-27: start = i*L, end = (i+1)*L - 1, int64 pairs
    seg = synth.segment_list(12, 256)
    assert seg.dtype == np.int64 and seg.size == 24
    assert np.array_equal(seg[0::2], np.arange(12) * 256)
    assert np.array_equal(seg[1::2], np.arange(1, 13) * 256 - 1)


def _tree(tmp_path, **kw):
    cfg = synth.Config(nside=4, ndet=2, nseg=2, ns=256, **kw)
    ds = synth.make_dataset(cfg, workers=1)
    args = synth.write_tree(str(tmp_path), cfg, ds)
    return cfg, ds, args


def test_tree_formats(tmp_path):
    cfg, ds, args = _tree(tmp_path)
    n = cfg.nseg * cfg.ns
    for key in ("ra", "dec", "psi"):
        a = np.fromfile(args["pointing"] + key + ".bi", dtype="<f8")
        assert np.array_equal(a, ds[key]) and a.size == n
    for k, name in enumerate(cfg.det_names):
        t = np.fromfile(os.path.join(args["tod_dir"], f"data_{name}output.bi"), dtype="<f8")
        assert np.array_equal(t, ds["tod"][k])
        sp = np.fromfile(args["ispf_prefix"] + name, dtype="<f8")
        assert sp.size == cfg.ns and np.all(sp > 0)
    lines = open(args["bolofile"]).read().split("\n")
    f = lines[1].split()
    assert len(f) == 7 and f[6] == "det1" and float(f[4]) == 22.5
    ra, dec = ds["ra"], ds["dec"]
    assert ra.min() > 0 and ra.max() <= 2 * np.pi and abs(dec).max() <= np.pi / 2


def test_run_script_flags():
    argv = synth.sanepic_argv(dict(tod_dir="T/", suffix="output", segfile="S", ndet=1, dets=["d0"],
                                   ispf_prefix="K", ns=256, nseg=12, nside=256, bolofile="X",
                                   pointing="Z", out="O", ext="output_N256_.bin"))
    flags = [a for a in argv if a.startswith("-")]
    # the 18 flags of scripts/run_sanepic.sh:55-73
    assert sorted(set(flags)) == sorted(["-F", "-B", "-f", "-d", "-C", "-e", "-h", "-w", "-k", "-u",
                                         "-O", "-p", "-l", "-n", "-N", "-X", "-Z", "-i"])


def _run(argv, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([EXE] + argv, capture_output=True, text=True, env=e, timeout=60)


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in executable not built")
def test_cli_unknown_option_exits_1():
    r = _run(["-Q", "1"])
    assert r.returncode == 1 and "not valid" in r.stderr


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in executable not built")
def test_cli_nonzero_mpi_rank_exits_0(tmp_path):
    cfg, ds, args = _tree(tmp_path)
    r = _run(synth.sanepic_argv(args)[1:], env={"OMPI_COMM_WORLD_RANK": "2"})
    assert r.returncode == 0 and r.stdout == ""


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in executable not built")
def test_cli_errors_before_device(tmp_path):
    cfg, ds, args = _tree(tmp_path)
    argv = synth.sanepic_argv(args)[1:]
    # unknown detector (3184-3187)
    bad = list(argv)
    bad[bad.index("-C") + 1] = "nosuchdet"
    r = _run(bad)
    assert r.returncode == 1 and "not found in offset list" in r.stderr
    # missing bolometer file (3169-3171)
    bad = list(argv)
    bad[bad.index("-X") + 1] = str(tmp_path / "missing.txt")
    r = _run(bad)
    assert r.returncode == 1 and "Can't find offset file" in r.stderr
    # segment length differs from -l (668-672)
    bad = list(argv)
    bad[bad.index("-l") + 1] = "128"
    r = _run(bad)
    assert r.returncode == 1 and "Problem with segment size" in r.stdout


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in executable not built")
def test_cli_unsupported_flags_fail(tmp_path):
    """-r with -b forces the leakage map to zero at high latitude inside the
    loop (forceLeakzero, 3262-3275), which this build does not do: it refuses
    the combination rather than ignore it, before touching the device."""
    cfg, ds, args = _tree(tmp_path)
    argv = synth.sanepic_argv(args)[1:]
    r = _run(argv + ["-r", "1", "-b", "1", "-K", str(tmp_path / "k.bi")])
    assert r.returncode == 1 and "not supported" in r.stderr
