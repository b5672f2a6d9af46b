"""The drivers on the reference's HDF5 files (when built with libhdf5):
the snapshot file /u,/t (ks_datagen.py:91-103) in, /spiv and /linv
(ks_arpclm.cu:636-644, 1094-1143) and /lam,/U (ks_kevd.cu:488-501,
1930-1938, 2300-2306) out.  Files are read back with the HDF5 tools'
h5dump (no Python HDF5 module here).

Bars: the .h5 and .npy routes give bit-identical points, pivots and L^-1
(and the LM form, --regenerate, the same pivots and L^-1 to 1e-9);
kevd-evd on (/u, /spiv, /linv) -- the reference kEVD's inputs -- matches the
same EVD run on the factor file to 1e-10 in the eigenvalues and reports a
rebuilt L^-1 within 1e-10 of /linv."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from kernel_evd_amd import ks_generate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "kernel_evd_amd", "bin")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def run(*args, check=True):
    r = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError("%s failed (%d):\n%s" % (args[0], r.returncode, r.stdout))
    return r


def h5read(path, dset):
    """Values of a 1-D dataset and its HDF5 type name, through h5dump."""
    out = subprocess.run([H5DUMP, "-d", dset, "-y", "-w", "0", "-m", "%.17g", str(path)],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120).stdout
    dtype = re.search(r"DATATYPE\s+(\S+)", out).group(1)
    body = out[out.index("DATA {") + 6:]
    body = body[:body.index("}")]
    vals = [v for v in re.split(r"[,\s]+", body.strip()) if v]
    return np.array([float(v) for v in vals]), dtype


@pytest.fixture(scope="module")
def h5_ok(tmp_path_factory):
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    p = tmp_path_factory.mktemp("probe") / "p.h5"
    r = run("kevd-gen", "--n", 4, "--M", 32, "--discard", 10, "--out", p, check=False)
    if r.returncode != 0 and "without HDF5" in r.stdout:
        pytest.skip("drivers built without HDF5")
    assert r.returncode == 0, r.stdout
    return True


def test_gen_h5_snapshot_layout(h5_ok, tmp_path):
    run("kevd-gen", "--n", 50, "--M", 32, "--discard", 100, "--seed", 3, "--out", tmp_path / "u.h5")
    run("kevd-gen", "--n", 50, "--M", 32, "--discard", 100, "--seed", 3, "--f32", "--out", tmp_path / "u32.h5")
    U = ks_generate(50, M=32, n_discard=100, seed=3)
    u, ty = h5read(tmp_path / "u.h5", "/u")
    assert ty == "H5T_IEEE_F64LE" and u.size == 50 * 32
    # time index fastest: u[g * Ntrue + t] = U[t, g]
    assert np.array_equal(u.reshape(32, 50).T, U)
    u32, ty32 = h5read(tmp_path / "u32.h5", "/u")
    assert ty32 == "H5T_IEEE_F32LE"
    assert np.array_equal(u32.reshape(32, 50).T, U.astype(np.float32).astype(np.float64))
    t, _ = h5read(tmp_path / "u.h5", "/t")
    assert np.allclose(t, (100 + 4 * np.arange(50)) * 0.25)


def test_gen_h5_rejects_delay(h5_ok, tmp_path):
    r = run("kevd-gen", "--n", 50, "--M", 32, "--delay", 8, "--out", tmp_path / "u.h5", check=False)
    assert r.returncode == 1


@pytest.mark.gpu
def test_arpc_and_evd_on_reference_files(h5_ok, tmp_path):
    nq, n, M = 8, 700, 32
    # the same 707 snapshots: as the reference's snapshot file, and delay
    # embedded by kevd-gen into N*M = 700*32 points of dimension 8
    run("kevd-gen", "--n", n + nq - 1, "--M", M, "--discard", 500, "--out", tmp_path / "u.h5")
    run("kevd-gen", "--n", n, "--M", M, "--discard", 500, "--delay", nq, "--out", tmp_path / "X.npy")
    # the seed also keys the sampled median (22,400 points), so the EVD
    # runs below take the same one to get the same eps
    common = ["--eps-median", "--rank", 150, "--block", 32, "--seed", 5, "--precision", "fp64"]
    print("step kevd-arpc h5", flush=True)
    r1 = run("kevd-arpc", "--points", tmp_path / "u.h5", "--delay", nq, *common,
             "--pivots-out", tmp_path / "spiv.h5", "--linv-out", tmp_path / "linv.h5")
    print("step kevd-arpc npy", flush=True)
    r2 = run("kevd-arpc", "--points", tmp_path / "X.npy", *common, "--pivots-out", tmp_path / "S.npy",
             "--linv-out", tmp_path / "L.npy", "--factor-out", tmp_path / "F.npy")
    assert "Points: 22400 x 8" in r1.stdout and "Points: 22400 x 8" in r2.stdout
    spiv, ty = h5read(tmp_path / "spiv.h5", "/spiv")
    assert ty == "H5T_STD_U64LE"
    S = np.load(tmp_path / "S.npy")
    assert np.array_equal(spiv.astype(np.int64), S)
    linv, _ = h5read(tmp_path / "linv.h5", "/linv")
    assert np.array_equal(linv, np.load(tmp_path / "L.npy"))
    print("step kevd-arpc --regenerate", flush=True)
    run("kevd-arpc", "--points", tmp_path / "u.h5", "--delay", nq, *common, "--regenerate",
        "--pivots-out", tmp_path / "spiv_lm.h5", "--linv-out", tmp_path / "linv_lm.h5")
    assert np.array_equal(h5read(tmp_path / "spiv_lm.h5", "/spiv")[0], spiv)
    linv_lm, _ = h5read(tmp_path / "linv_lm.h5", "/linv")
    assert np.linalg.norm(linv_lm - linv) <= 1e-9 * np.linalg.norm(linv)
    r = run("kevd-arpc", "--points", tmp_path / "u.h5", "--delay", nq, *common, "--regenerate",
            "--factor-out", tmp_path / "F_lm.npy", check=False)
    assert r.returncode == 1
    print("step kevd-evd on /u /spiv /linv", flush=True)
    r3 = run("kevd-evd", "--points", tmp_path / "u.h5", "--delay", nq, "--pivots", tmp_path / "spiv.h5",
             "--linv", tmp_path / "linv.h5", "--eps-median", "--seed", 5, "--kout", 10, "--precision", "fp64",
             "--out", tmp_path / "evd.h5")
    m = re.search(r"relative Frobenius difference (\S+)", r3.stdout)
    assert m and float(m.group(1)) < 1e-10, r3.stdout
    print("step kevd-evd on the factor", flush=True)
    run("kevd-evd", "--points", tmp_path / "X.npy", "--factor", tmp_path / "F.npy", "--kout", 10,
        "--precision", "fp64", "--all-eigenvalues", "--lam-out", tmp_path / "lam.npy", "--U-out", tmp_path / "U.npy")
    lam, ty = h5read(tmp_path / "evd.h5", "/lam")
    assert ty == "H5T_IEEE_F64LE" and lam.size == 150
    ref = np.load(tmp_path / "lam.npy")
    assert np.all(np.diff(lam) >= -1e-12)  # ascending, all r values
    np.testing.assert_allclose(lam[-10:], ref[-10:], rtol=1e-10, atol=1e-13)
    U, ty = h5read(tmp_path / "evd.h5", "/U")
    assert ty == "H5T_IEEE_F32LE" and U.size == 22400 * 10
    U = U.reshape(22400, 10)
    Ur = np.load(tmp_path / "U.npy").astype(np.float64)
    # same columns up to sign (row-major N x k_out, columns ascending)
    for j in range(10):
        c = np.dot(U[:, j], Ur[:, j]) / (np.linalg.norm(U[:, j]) * np.linalg.norm(Ur[:, j]))
        assert abs(abs(c) - 1.0) < 1e-5, (j, c)


@pytest.mark.gpu
def test_evd_dump_file(h5_ok, tmp_path):
    """--dump: the reference's kevd_dump datasets, flat 1-D, FP64 but U
    FP32, and the same values as the .npy directory form."""
    X = tmp_path / "X.npy"
    run("kevd-gen", "--n", 1500, "--discard", 1000, "--seed", 2, "--out", X)
    run("kevd-arpc", "--points", X, "--eps-median", "--rank", 80, "--block", 16, "--precision", "fp64",
        "--pivots-out", tmp_path / "spiv.h5")
    args = ["--points", X, "--pivots", tmp_path / "spiv.h5", "--eps-median", "--kout", 6, "--precision", "fp64"]
    print("step kevd-evd --dump h5", flush=True)
    run("kevd-evd", *args, "--dump", tmp_path / "dump.h5")
    print("step kevd-evd --dump npy", flush=True)
    run("kevd-evd", *args, "--dump", tmp_path / "dump")
    sizes = {"c1": 80, "p1": 80, "dtil": 1500, "cq": 80, "pq": 80, "Atil": 6400, "A": 6400, "R": 6400,
             "G": 6400, "qtil": 1500, "Btil": 6400, "W": 6400, "V": 6400, "lam": 80, "Z": 480, "U": 9000}
    for name, size in sizes.items():
        v, ty = h5read(tmp_path / "dump.h5", "/" + name)
        assert ty == ("H5T_IEEE_F32LE" if name == "U" else "H5T_IEEE_F64LE"), name
        assert v.size == size, name
        np.testing.assert_array_equal(v, np.load(tmp_path / "dump" / (name + ".npy")).astype(np.float64))
