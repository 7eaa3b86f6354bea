"""Write tests/golden/dat_fixtures.npz: the reference's own `.dat` writers
(EulerTests / MHDTests / HPRTests / EulerRGFMTests ::outputSolution[2D]) run by
the compiled reference harness (`oracle/_ref/ref_harness dat ...`) on seeded
states, with the inputs and every file's bytes stored as plain arrays (no
pickles); and tests/golden/driver_fixtures.npz: the reference's test drivers
run end to end (`ref_harness driver ...`: their initial conditions, solves and
writers), every file's bytes.  Run in the build container only:

    python tests/golden/gen_dat.py
"""

from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# is_thermal rho0 p0 t_ref t0 cv gamma p_inf cs tau1 tau2 alpha (the harness's HPR argument order)
HPR_PARAMS = np.array([1, 1.0, 1.0 / 1.4, 0.05, 1.0, 2.5, 1.4, 0.3, 1.0, 8.4e-3, 3.1e-3, 2.0])


def euler_states(rng, n):
    w = np.zeros((n, 6))
    w[:, 0] = 10.0 ** rng.uniform(-3, 3, n)
    w[:, 1:4] = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-8, 4, (n, 3))
    w[:, 4] = 10.0 ** rng.uniform(-6, 6, n)
    w[:, 5] = rng.uniform(0, 1, n)
    w[0, 1] = -0.0
    w[1, 1] = 0.0
    w[2, 2] = np.nan
    w[3, 2] = np.copysign(np.nan, -1.0)
    w[4, 4] = 123456.5
    w[5, 4] = 1e-300
    w[6, 0] = 9.9999995
    return w


def mhd_states(rng, n):
    w = np.zeros((n, 9))
    w[:, 0] = 10.0 ** rng.uniform(-1, 1, n)
    w[:, 1:4] = rng.uniform(-2, 2, (n, 3))
    w[:, 4] = 10.0 ** rng.uniform(-2, 2, n)
    w[:, 5:8] = rng.uniform(-2, 2, (n, 3))
    w[0, 6] = -0.0
    return w


def hpr_states(rng, n):
    w = np.zeros((n, 17))
    w[:, 0] = rng.uniform(0.5, 2.0, n)
    w[:, 1] = rng.uniform(0.1, 2.0, n)
    w[:, 2:11] = np.eye(3).reshape(9) + 0.05 * rng.standard_normal((n, 9))
    w[:, 11:14] = rng.uniform(-1, 1, (n, 3))
    w[:, 14:17] = 0.1 * rng.standard_normal((n, 3))
    return w


def rgfm_system_1d(rng, n, phi0_zero):
    m1, m2 = euler_states(rng, n), euler_states(rng, n)
    x = (np.arange(n) + 0.5) / n
    phi = np.abs(x - 0.5) - 0.2
    if phi0_zero:
        phi[0] = 0.0
    return m1, m2, phi


def rgfm_system_2d(rng, ny, nx):
    m1 = euler_states(rng, ny * nx).reshape(ny, nx, 6)
    m2 = euler_states(rng, ny * nx).reshape(ny, nx, 6)
    y = (np.arange(ny) + 0.5)[:, None] / ny
    x = (np.arange(nx) + 0.5)[None, :] / nx
    phi = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2) - 0.3
    return m1, m2, phi


def run(args, tmp):
    r = subprocess.run([HARNESS, "dat"] + [str(a) for a in args], capture_output=True, text=True, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(f"harness failed: {r.stderr}")


# (driver, cellCount, sub-cycles or reinitialisation frequency): the reference's
# EulerTests, EulerRGFMTests (HLLC), MHDTests and HPRTests drivers
DRIVERS = [("euler_toro1", 100, 0), ("euler_2d_toro1", 32, 0), ("rgfm_toro1_hllc", 100, 0),
           ("rgfm_fedkiw_hllc", 100, 0), ("rgfm_2d_toro1_hllc", 32, 0), ("rgfm_2d_fedkiw_hllc", 32, 0),
           ("mhd_dumbser1", 100, 0), ("mhd_dumbser2", 100, 1), ("mhd_dumbser3", 100, 0), ("mhd_2d_dumbser1", 32, 0),
           ("mhd_2d_dumbser2", 32, 1), ("hpr_stokes", 50, 1), ("hpr_heat", 100, 1)]


def drivers():
    out = {}
    env = dict(os.environ, OMP_NUM_THREADS="1")
    for name, n, k in DRIVERS:
        with tempfile.TemporaryDirectory() as tmp:
            r = subprocess.run([HARNESS, "driver", name, str(n), str(k)], capture_output=True, text=True, cwd=tmp,
                               env=env)
            if r.returncode != 0:
                raise RuntimeError(f"driver {name} failed: {r.stderr}")
            files = sorted(f for f in os.listdir(tmp) if f.endswith(".dat"))
            out[f"{name}__args"] = np.array([n, k])
            out[f"{name}__files"] = np.array(files)
            for f in files:
                with open(os.path.join(tmp, f), "rb") as fh:
                    out[f"{name}__file__{f}"] = np.frombuffer(fh.read(), dtype=np.uint8)
        print(name, files)
    np.savez_compressed(os.path.join(HERE, "driver_fixtures.npz"), **out)


def main():
    drivers()
    rng = np.random.default_rng(20261202)
    cases = {
        "euler1d": ("euler", 1, euler_states(rng, 23)[None]),
        "euler2d": ("euler", 2, euler_states(rng, 35).reshape(5, 7, 6)),
        "mhd1d": ("mhd", 1, mhd_states(rng, 19)[None]),
        "mhd2d": ("mhd", 2, mhd_states(rng, 24).reshape(4, 6, 9)),
        "hpr1d": ("hpr", 1, hpr_states(rng, 17)[None]),
        "hpr2d": ("hpr", 2, hpr_states(rng, 15).reshape(3, 5, 17)),
        "rgfm1d": ("rgfm", 1, rgfm_system_1d(rng, 29, False)),
        "rgfm1d_phi0": ("rgfm", 1, rgfm_system_1d(rng, 29, True)),
        "rgfm2d": ("rgfm", 2, rgfm_system_2d(rng, 6, 9)),
        "rgfm2d_odd": ("rgfm", 2, rgfm_system_2d(rng, 5, 8)),
    }
    out = {"hpr_params": HPR_PARAMS}
    for name, (kind, dim, data) in cases.items():
        with tempfile.TemporaryDirectory() as tmp:
            if kind == "rgfm":
                m1, m2, phi = data
                if dim == 1:
                    m1, m2, phi = m1[None], m2[None], phi[None]
                ny, nx = phi.shape
                for arr, f in ((m1, "m1.f64"), (m2, "m2.f64"), (phi, "phi.f64")):
                    np.ascontiguousarray(arr, dtype=np.float64).tofile(os.path.join(tmp, f))
                run([kind, dim, ny, nx, "m1.f64", "m2.f64", "phi.f64"], tmp)
                out[f"{name}__m1"], out[f"{name}__m2"], out[f"{name}__phi"] = data
            else:
                ny, nx = data.shape[:2]
                np.ascontiguousarray(data, dtype=np.float64).tofile(os.path.join(tmp, "in.f64"))
                extra = [repr(float(v)) if i else int(v) for i, v in enumerate(HPR_PARAMS)] if kind == "hpr" else []
                run([kind, dim, ny, nx, "in.f64"] + extra, tmp)
                out[f"{name}__in"] = data[0] if dim == 1 else data
            files = sorted(f for f in os.listdir(tmp) if f.endswith(".dat"))
            out[f"{name}__files"] = np.array(files)
            for f in files:
                with open(os.path.join(tmp, f), "rb") as fh:
                    out[f"{name}__file__{f}"] = np.frombuffer(fh.read(), dtype=np.uint8)
        print(name, files)
    np.savez_compressed(os.path.join(HERE, "dat_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
