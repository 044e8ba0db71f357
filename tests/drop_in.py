"""Builds and runs tests/cpp/link_test.cpp (a translation unit written like
the reference's LevelSet.cpp) against libftreinit.so and the oracle's answer
for its inputs."""
import os
import subprocess

import numpy as np

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_link_test(tmp_path, copies=False, n=12, env=None, check=True):
    exe = tmp_path / "link_test"
    lib_dir = os.path.join(ROOT, "fluidtissue_amd")
    subprocess.run(["g++", "-std=c++14", "-O2", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "link_test.cpp"), "-L" + lib_dir,
                    "-lftreinit", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    args = [str(exe), str(n)] + (["copy"] if copies else [])
    run = subprocess.run(args, capture_output=True, text=True, env=env, timeout=120)
    if not check:
        return run
    assert run.returncode == 0, run.stderr
    tok = run.stdout.split()
    N, P = int(tok[0]), int(tok[1])
    ifp = np.array([float.fromhex(x) for x in tok[2:2 + 3 * P]], np.float32).reshape(P, 3)
    cols = 5 if copies else 2
    rest = np.array([float.fromhex(x) for x in tok[2 + 3 * P:]]).reshape(N, cols)
    phi = rest[:, 0]
    h = np.float32(1.0 / (n - 1))
    g = np.arange(n, dtype=np.float32) * h
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    ref, _ = oracle.reset_phi_points(xyz, ifp, phi)
    return dict(phi_cols=[rest[:, k] for k in range(1, cols)], stderr=run.stderr), ref


def run_level2(tmp_path, n, mode):
    """Builds and runs tests/cpp/level2_test.cpp (the Level-2 integration of
    LevelSet::resetPhi_ through the C ABI) and returns its arrays."""
    exe = tmp_path / "level2_test"
    lib_dir = os.path.join(ROOT, "fluidtissue_amd")
    subprocess.run(["g++", "-std=c++14", "-O2", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "level2_test.cpp"), "-L" + lib_dir,
                    "-lftreinit", "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    out = tmp_path / f"l2_{mode}"
    out.mkdir()
    run = subprocess.run([str(exe), str(n), str(out), mode], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    N, T, E, P, F, X = (int(x) for x in run.stdout.split())

    def load(name, dtype, cols=None):
        a = np.fromfile(out / f"{name}.bin", dtype=dtype)
        return a.reshape(-1, cols) if cols else a
    r = dict(N=N, T=T, E=E, P=P, F=F, X=X, dir=str(out),
             xyz=load("xyz", np.float32, 3), tets=load("tets", np.int32, 4), edges=load("edges", np.int32, 2),
             t2e=load("t2e", np.int32, 6), ifp=load("ifp", np.float32, 3), ifp2e=load("ifp2e", np.int32),
             tris=load("tris", np.int32, 4), elems=load("elems", np.int32), area=load("area", np.float32),
             p2t_off=load("p2t_off", np.int32), p2t=load("p2t", np.int32))
    for k in ("phi0", "phi1", "ifn1", "phi2", "ifn2"):
        r[k] = load(k, np.float64)
    return r
