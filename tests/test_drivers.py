"""The C++ drivers (the reference's program-level interface: points file
in, parameters as flags, factor / pivot / eigenpair / bw0.out files out)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from conftest import clustered_points
from kernel_evd_amd import delay_embed, ks_generate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "kernel_evd_amd", "bin")


def run(*args, check=True):
    r = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError("%s failed (%d):\n%s" % (args[0], r.returncode, r.stdout))
    return r


@pytest.mark.parametrize("exe", ["kevd-gen", "kevd-arpc", "kevd-evd", "kevd-bandwidth"])
def test_help(exe):
    assert run(exe, "--help").returncode == 0


def test_gen_matches_library(tmp_path):
    out = tmp_path / "X.npy"
    run("kevd-gen", "--n", 300, "--discard", 200, "--seed", 4, "--out", out)
    X = np.load(out)
    assert X.shape == (300, 64)
    assert np.array_equal(X, ks_generate(300, n_discard=200, seed=4))


def test_delay_embedding_index_map(tmp_path):
    out = tmp_path / "Xd.npy"
    run("kevd-gen", "--n", 40, "--discard", 50, "--M", 32, "--delay", 8, "--out", out)
    Xd = np.load(out)
    U = ks_generate(47, M=32, n_discard=50)
    assert Xd.shape == (40 * 32, 8)
    # point g*N + t is the window U[t:t+8, g] (ks_arpclm.cu:263-268)
    for g, t in [(0, 0), (5, 17), (31, 39)]:
        assert np.array_equal(Xd[g * 40 + t], U[t:t + 8, g])
    assert np.array_equal(Xd, delay_embed(U, 8))


def test_missing_file_exit_code(tmp_path):
    r = run("kevd-arpc", "--points", tmp_path / "nope.npy", "--rank", 4, check=False)
    assert r.returncode == 1


@pytest.mark.gpu
def test_pipeline_end_to_end(tmp_path):
    X = tmp_path / "X.npy"
    run("kevd-gen", "--n", 3000, "--discard", 2000, "--seed", 0, "--out", X)
    F, S, Li = tmp_path / "F.npy", tmp_path / "S.npy", tmp_path / "Linv.npy"
    pts = np.load(X)
    eps = orc.median_sqdist(pts)  # exact
    r = run("kevd-arpc", "--points", X, "--eps", repr(eps), "--rank", 200, "--block", 32,
            "--precision", "fp64", "--factor-out", F, "--pivots-out", S, "--linv-out", Li)
    print(r.stdout)
    assert "# round accepted cumrank relerr" in r.stdout
    Fa, Sa = np.load(F), np.load(S)
    assert Fa.shape == (3000, 200) and Sa.shape == (200,)
    lam_f = tmp_path / "lam.npy"
    U_f = tmp_path / "U.npy"
    r = run("kevd-evd", "--points", X, "--factor", F, "--kout", 16, "--precision", "fp64",
            "--lam-out", lam_f, "--U-out", U_f)
    lam = np.load(lam_f)
    ref_lam, _, _, _ = orc.lowrank_evd(Fa, 16)
    assert np.max(np.abs(lam[-16:] - ref_lam[-16:])) < 1e-9
    assert np.load(U_f).shape == (3000, 16)
    # rebuild from the pivot list (the reference's spiv-driven kEVD)
    lam2_f = tmp_path / "lam2.npy"
    run("kevd-evd", "--points", X, "--pivots", S, "--eps", repr(eps), "--kout", 16,
        "--precision", "fp64", "--lam-out", lam2_f)
    assert np.max(np.abs(np.load(lam2_f)[-16:] - lam[-16:])) < 1e-8
    # and with no factor stored at all (K(R,S) panels regenerated per pass)
    lam3_f = tmp_path / "lam3.npy"
    run("kevd-evd", "--points", X, "--pivots", S, "--eps", repr(eps), "--kout", 16,
        "--precision", "fp64", "--regenerate", "--lam-out", lam3_f)
    assert np.max(np.abs(np.load(lam3_f)[-16:] - lam[-16:])) < 1e-8
    bw = tmp_path / "bw0.out"
    run("kevd-bandwidth", "--points", X, "--rows", 500, "--grid", "ref36", "--out", bw)
    lines = open(bw).read().strip().split("\n")
    assert len(lines) == 34 + 1 + 2 and lines[34] == ""


@pytest.mark.gpu
def test_evd_dump_directory(tmp_path):
    """kevd-evd --dump DIR: the reference's kevd_dump datasets as flat .npy
    files, against the oracle's reference-basis restatement."""
    X = tmp_path / "X.npy"
    run("kevd-gen", "--n", 2000, "--discard", 1000, "--seed", 4, "--out", X)
    pts = np.load(X)
    eps = orc.median_sqdist(pts)
    S = tmp_path / "S.npy"
    run("kevd-arpc", "--points", X, "--eps", repr(eps), "--rank", 120, "--block", 16, "--precision", "fp64",
        "--pivots-out", S, "--factor-out", tmp_path / "F.npy")
    d = tmp_path / "dump"
    run("kevd-evd", "--points", X, "--pivots", S, "--eps", repr(eps), "--kout", 8, "--precision", "fp64",
        "--dump", d)
    ref = orc.kevd_record(pts, eps, np.load(S), 8, extended=True)
    r = 120
    for name in ("c1", "dtil", "cq", "qtil", "lam"):
        v = np.load(d / (name + ".npy"))
        assert v.ndim == 1
        np.testing.assert_allclose(v, ref[name], rtol=1e-9, atol=1e-10)
    for name in ("Atil", "A", "R", "G", "Btil", "W"):
        v = np.load(d / (name + ".npy")).reshape(r, r, order="F")
        assert np.linalg.norm(v - ref[name]) < 1e-8 * np.linalg.norm(ref[name]), name
    assert np.load(d / "V.npy").size == r * r and np.load(d / "Z.npy").size == r * 8
    assert np.load(d / "U.npy").dtype == np.float32 and np.load(d / "U.npy").size == 2000 * 8
    # a factor file carries no pivots: no kernel basis to record in
    bad = run("kevd-evd", "--points", X, "--factor", tmp_path / "F.npy", "--kout", 8, "--precision", "fp64",
              "--dump", tmp_path / "dump2", check=False)
    assert bad.returncode == 1, bad.stdout


@pytest.mark.gpu
def test_bw0_out_matches_reference_columns(tmp_path):
    """--all-pairs --grid ref36: column 2 of bw0.out is the reference's
    log(2 S_lower / NM + 1) (ks_bandwidth.cu:615-618) = log(S_full / NM),
    column 3 the centred slope, then eopt and mdim."""
    X = tmp_path / "X.npy"
    pts = clustered_points(700, 6, seed=3)
    np.save(X, pts)
    bw, S_f = tmp_path / "bw0.out", tmp_path / "S.npy"
    run("kevd-bandwidth", "--points", X, "--all-pairs", "--grid", "ref36", "--precision", "fp64",
        "--out", bw, "--sums-out", S_f)
    S = np.load(S_f)
    n, dim = pts.shape
    bws = np.array([m * 10.0 ** e for e in range(-3, 1) for m in range(1, 10)])
    # the reference's sums: strictly lower triangle of exp(-|dx|^2 / (dim bw)) / NM
    from scipy.spatial.distance import pdist
    D = pdist(pts, "sqeuclidean")
    S_lower = np.array([np.exp(-D / (dim * b)).sum() for b in bws]) / n
    ref_col2 = np.log(2.0 * S_lower + 1.0)
    np.testing.assert_allclose(S, 2.0 * S_lower * n + n, rtol=1e-11)
    lines = open(bw).read().strip().split("\n")
    vals = np.array([[float(v) for v in l.split()] for l in lines[:34]])
    np.testing.assert_allclose(vals[:, 0], bws[1:-1], rtol=1e-5)
    np.testing.assert_allclose(vals[:, 1], ref_col2[1:-1], rtol=2e-5, atol=1e-9)
    slope = (ref_col2[2:] - ref_col2[:-2]) / (np.log(bws[2:]) - np.log(bws[:-2]))
    np.testing.assert_allclose(vals[:, 2], slope, rtol=2e-5, atol=1e-9)
    idmax = int(np.argmax(slope))
    assert abs(float(lines[35]) - bws[idmax + 1]) <= 1e-5 * bws[idmax + 1]
    assert abs(float(lines[36]) - 2.0 * slope[idmax]) <= 2e-5 * abs(slope[idmax]) + 1e-9
