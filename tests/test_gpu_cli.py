"""End-to-end drop-in test: SANEPIC's input tree -> bin/sanepic (run with the
flags of run_sanepic.sh) -> SANEPIC's output files, checked against the CPU
oracle run on the same inputs."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import sanepic_amd
from sanepic_amd import synth
from tests.helpers import oracle_problem, rel

pytestmark = pytest.mark.gpu


def _case(tmp_path, nside=8, ndet=2, nseg=4, ns=2048):
    tmin = nseg * ns / 19.0 / 60.0
    cfg = synth.Config(nside=nside, ndet=ndet, nseg=nseg, ns=ns, fknee=0.2,
                       spin_min=tmin / 40, prec_min=tmin / 5)
    ds = synth.make_dataset(cfg, workers=1)
    args = synth.write_tree(str(tmp_path), cfg, ds)
    return cfg, ds, args


def _run(argv, env=None):
    e = dict(os.environ)
    e.update(env or {})
    print("running", " ".join(argv), flush=True)
    r = subprocess.run(argv, capture_output=True, text=True, env=e, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def test_drop_in_outputs_match_oracle(tmp_path):
    cfg, ds, args = _case(tmp_path)
    r = _run(synth.sanepic_argv(args, exe=sanepic_amd.EXE_PATH))
    crit = [float(m) for m in re.findall(r"^iter = \d+, crit = (\S+),", r.stdout, re.M)]
    assert crit, r.stdout[-2000:]
    P = oracle_problem(cfg, ds)
    ref = P.solve(stop=1e-15, maxiter=2000)
    npix = P.npix
    # the reference's loop: same stop rule, same first crit values
    assert np.allclose(crit[:5], ref["crit"][:5], rtol=1e-8)
    assert crit[-1] <= 1e-15 or len(crit) == 2000
    _, _, _, ind = P.pixels()
    nsky = 12 * cfg.nside ** 2
    O, E = args["out"], args["ext"]

    def rd(name, n=nsky):
        a = np.fromfile(name, dtype="<f8")
        assert a.size == n, name
        return a

    hits = ref["hits"]
    full = lambda v: oracle.expand_map(cfg.nside, ind, v)
    assert np.array_equal(rd(f"{O}_HITS{E}"), full(hits[:npix].astype(float)))
    assert np.array_equal(rd(f"{O}_POL_HITS{E}"), full(hits[npix:].astype(float)))
    mpt = ref["Mptot"]
    assert rel(rd(f"{O}_NVAR{E}"), full(mpt[:npix])) < 1e-12
    assert rel(rd(f"{O}_Q_NVAR{E}"), full(mpt[npix:2 * npix])) < 1e-12
    assert rel(rd(f"{O}_U_NVAR{E}"), full(mpt[2 * npix:])) < 1e-12
    sc = ref["sumCross"]
    assert rel(rd(f"{O}_COS2P_{E}"), full(sc[:npix])) < 1e-12
    assert rel(rd(f"{O}_SIN2P_{E}"), full(sc[npix:2 * npix])) < 1e-12
    A = rd(f"{O}_A_{E}_1", 3 * npix)
    for k, c in enumerate("IQU"):
        assert np.array_equal(rd(f"{O}_{c}_{E}_1"), full(A[k * npix:(k + 1) * npix]))
    # the written map solves the oracle's system to the reference's tolerance
    b = ref["b"]
    ax = P.apply_A(A)
    m = hits[npix:] < 4
    ax[npix:2 * npix][m] = 0
    ax[2 * npix:][m] = 0
    assert np.linalg.norm(b - ax) / np.linalg.norm(b) < 1e-6


def test_warm_start_from_A_file(tmp_path):
    cfg, ds, args = _case(tmp_path)
    argv = synth.sanepic_argv(args, exe=sanepic_amd.EXE_PATH)
    pat = r"^iter = \d+, crit = (\S+),"
    cold = [float(m) for m in re.findall(pat, _run(argv).stdout, re.M)]
    afile = f"{args['out']}_A_{args['ext']}_1"
    warm = [float(m) for m in re.findall(pat, _run(argv + ["-I", afile]).stdout, re.M)]
    # the _A_ file is the restart state (1209-1227): the warm run resumes
    # where the cold one stopped, not from the start
    assert warm[0] < 1e-6 * cold[0]
    assert warm[0] < 10 * cold[-1]


def test_run_script_and_mpirun_compat(tmp_path):
    cfg, ds, args = _case(tmp_path, ndet=1, nseg=12, ns=256)
    env = dict(base_dir=str(tmp_path), NSIDE=str(cfg.nside), det_num="1", det_name="det0",
               nn=str(cfg.nseg), seglength=str(cfg.ns), fknee=str(cfg.fknee),
               fsamp2=str(cfg.fsamp / 2), indexN=str(cfg.alpha))
    script = os.path.join(os.path.dirname(sanepic_amd.EXE_PATH), "..", "run_sanepic.sh")
    r = _run(["bash", script], env=env)
    assert "Map-making Ends" in r.stdout
    assert os.path.exists(os.path.join(str(tmp_path), "out_dir", f"map__I_output_N{cfg.nside}_.bin_1"))
    assert os.path.exists(os.path.join(str(tmp_path), "out_dir", f"map__HITSoutput_N{cfg.nside}_.bin"))
