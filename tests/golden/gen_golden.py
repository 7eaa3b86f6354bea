"""Write the golden fixtures under tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_harness (built by `make -C oracle ref` from the reference
sources under /root/reference) single-threaded -- the reference's OpenMP
wave-speed maximum is racy, so only OMP_NUM_THREADS=1 runs are deterministic --
and stores inputs and outputs as compressed .npz fixtures (plain float64 arrays,
no pickles).  Run in the build container only:

    python tests/golden/gen_golden.py [name ...]
"""

from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vanguard_amd import problems  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
IDEAL = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)
# Stiffened gas with an energy of formation, so every EOS term is exercised.
STIFF = (3.0, 0.1, 25.0, 0.5, 0.25, 250.0)


def run(args, tmp):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([HARNESS] + [str(a) for a in args], env=env, capture_output=True, text=True, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(f"harness failed: {r.stderr}")
    return r.stderr


def euler_run(name, w, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, params, check_api):
    with tempfile.TemporaryDirectory() as tmp:
        ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
        np.ascontiguousarray(w, dtype=np.float64).tofile(os.path.join(tmp, "in.f64"))
        t = "inf" if np.isinf(tfinal) else repr(float(tfinal))
        log = run(["euler_run", order, dim, ny, nx, repr(dx), repr(cfl), t, max_steps, repr(bias), limiter,
                   *[repr(p) for p in params], "in.f64", "state.f64", "dt.f64", int(check_api)], tmp)
        state = np.fromfile(os.path.join(tmp, "state.f64")).reshape(w.shape)
        dt = np.fromfile(os.path.join(tmp, "dt.f64"))
    api = None
    for line in log.splitlines():
        if line.startswith("API_MATCH"):
            api = int(line.split()[1])
    if check_api and max_steps == 0 and api != 1:
        raise RuntimeError(f"{name}: mirror loop differs from the reference's own solve")
    meta = dict(order=order, dim=dim, dx=dx, cfl=cfl, tfinal=tfinal, max_steps=max_steps, bias=bias,
                limiter=limiter, params=list(params), api_checked=bool(check_api and max_steps == 0))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=w, prim_out=state, dt=dt,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {len(dt)} steps")


def random_states(rng, n, stiff):
    """Five primitive states per record, with neighbours close enough that many
    conserved differences fall below the reference's 1e-5 slope clamp and that
    every limiter branch is reached."""
    base = np.empty((n, 6))
    base[:, 0] = rng.uniform(0.2, 2.0, n)
    base[:, 1:4] = rng.uniform(-1.0, 1.0, (n, 3))
    base[:, 4] = rng.uniform(0.2, 2.0, n)
    base[:, 5] = rng.uniform(0.3, 1.0, n) if stiff else 1.0
    out = np.empty((n, 5, 6))
    for k in range(5):
        scale = rng.choice([1e-7, 1e-6, 1e-3, 0.05, 0.3], size=(n, 1))
        pert = rng.uniform(-1.0, 1.0, (n, 6)) * scale
        s = base + pert * np.abs(base).clip(0.1)
        s[:, 0] = np.abs(s[:, 0]) + 0.05
        s[:, 4] = np.abs(s[:, 4]) + 0.05
        if not stiff:
            s[:, 5] = 1.0
        # some records with a repeated neighbour (zero differences)
        same = rng.random(n) < 0.1
        if k > 0:
            s[same] = out[same, k - 1]
        out[:, k] = s
    return out.reshape(n, 30)


def euler_units(name, n, seed, bias, limiter, params, stiff):
    rng = np.random.default_rng(seed)
    states = random_states(rng, n, stiff)
    dx, ts = 1e-2, 2e-3
    with tempfile.TemporaryDirectory() as tmp:
        states.tofile(os.path.join(tmp, "in.f64"))
        run(["euler_units", n, repr(bias), limiter, repr(dx), repr(ts), *[repr(p) for p in params],
             "in.f64", "out.f64"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.f64")).reshape(n, 92)
    meta = dict(bias=bias, limiter=limiter, dx=dx, ts=ts, params=list(params))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), states=states, out=out,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {n} records")


def limiter_units(name, bias):
    r = np.concatenate([np.array([-1.0, -0.0, 0.0, 1e-300, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 1e300,
                                  np.nextafter(0.5, 0), np.nextafter(1.0, 0), np.nextafter(1.0, 2)]),
                        np.random.default_rng(7).uniform(-2, 5, 500)])
    with tempfile.TemporaryDirectory() as tmp:
        r.tofile(os.path.join(tmp, "in.f64"))
        run(["limiter_units", r.size, repr(bias), "in.f64", "out.f64"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.f64")).reshape(r.size, 5)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), r=r, out=out, bias=np.array(bias))
    print(f"{name}: {r.size} values")


def stiff_1d(n):
    w = np.empty((n, 6))
    x = (np.arange(n) + 0.5) / n
    left = x < 0.4
    w[left] = (1.0, 0.2, 0.1, -0.05, 1.0, 1.0)
    w[~left] = (0.5, -0.1, 0.0, 0.02, 0.2, 0.3)
    return w


def rect_2d(ny, nx):
    """Off-centre, anisotropic 2D problem with flow in x and y (pins the i = y,
    j = x orientation and the y-direction flux)."""
    w = np.empty((ny, nx, 6))
    i = np.arange(ny)[:, None]
    j = np.arange(nx)[None, :]
    inside = (i - 0.3 * ny) ** 2 + (0.5 * (j - 0.6 * nx)) ** 2 <= (0.2 * ny) ** 2
    w[..., 0] = np.where(inside, 1.0, 0.2)
    w[..., 1] = np.where(inside, 0.3, -0.1)
    w[..., 2] = np.where(inside, -0.2, 0.15)
    w[..., 3] = 0.05
    w[..., 4] = np.where(inside, 1.5, 0.3)
    w[..., 5] = 1.0
    return w


def mhd_call(w, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, gamma, p_inf, check_api, subcycles=0):
    with tempfile.TemporaryDirectory() as tmp:
        ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
        np.ascontiguousarray(w, dtype=np.float64).tofile(os.path.join(tmp, "in.f64"))
        t = "inf" if np.isinf(tfinal) else repr(float(tfinal))
        log = run(["mhd_run", order, dim, ny, nx, repr(dx), repr(cfl), t, max_steps, repr(bias), limiter, repr(gamma),
                   repr(p_inf), "in.f64", "state.f64", "dt.f64", int(check_api), subcycles], tmp)
        state = np.fromfile(os.path.join(tmp, "state.f64")).reshape(w.shape)
        dt = np.fromfile(os.path.join(tmp, "dt.f64"))
    if check_api and max_steps == 0 and "API_MATCH 1" not in log:
        raise RuntimeError("MHD mirror loop differs from the reference's own solve")
    return state, dt


def mhd_run(name, w, **kw):
    state, dt = mhd_call(w, **kw)
    meta = dict(kw)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=w, prim_out=state, dt=dt,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {len(dt)} steps")


def mhd_batch_fixture(name, n_inst, seed, n_min, n_max, tfinal):
    prim, offsets, gamma, dx = problems.mhd_batch(n_inst, seed=seed, n_min=n_min, n_max=n_max)
    out = np.empty_like(prim)
    steps = np.empty(n_inst, dtype=np.int64)
    tend = np.empty(n_inst)
    for q in range(n_inst):
        a, b = offsets[q], offsets[q + 1]
        s, d = mhd_call(prim[a:b], order=2, dim=1, dx=float(dx[q]), cfl=0.8, tfinal=tfinal, max_steps=0, bias=0.0,
                        limiter=1, gamma=float(gamma[q]), p_inf=0.0, check_api=False)
        out[a:b] = s
        steps[q] = len(d)
        t = 0.0
        for x in d:
            t += x
        tend[q] = t
    meta = dict(n_inst=n_inst, seed=seed, n_min=n_min, n_max=n_max, tfinal=tfinal, cfl=0.8, limiter=1, bias=0.0)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=prim, prim_out=out, offsets=offsets, gamma=gamma,
                        dx=dx, steps=steps, t=tend, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {n_inst} instances, {int(steps.sum())} steps")


def mhd_units(name, n, seed, bias, limiter, gamma, ch):
    rng = np.random.default_rng(seed)
    base = np.empty((n, 9))
    base[:, 0] = rng.uniform(0.2, 2.0, n)
    base[:, 1:4] = rng.uniform(-1.0, 1.0, (n, 3))
    base[:, 4] = rng.uniform(0.2, 2.0, n)
    base[:, 5:8] = rng.uniform(-1.5, 1.5, (n, 3))
    base[:, 8] = rng.uniform(-0.1, 0.1, n)
    st = np.empty((n, 5, 9))
    for k in range(5):
        scale = rng.choice([1e-7, 1e-6, 1e-3, 0.05, 0.3], size=(n, 1))
        s = base + rng.uniform(-1.0, 1.0, (n, 9)) * scale * np.abs(base).clip(0.1)
        s[:, 0] = np.abs(s[:, 0]) + 0.05
        s[:, 4] = np.abs(s[:, 4]) + 0.05
        st[:, k] = s
    states = st.reshape(n, 45)
    dx, ts = 1e-2, 2e-3
    with tempfile.TemporaryDirectory() as tmp:
        states.tofile(os.path.join(tmp, "in.f64"))
        run(["mhd_units", n, repr(bias), limiter, repr(dx), repr(ts), repr(gamma), "0.0", repr(ch), "in.f64",
             "out.f64"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.f64")).reshape(n, 85)
    meta = dict(bias=bias, limiter=limiter, dx=dx, ts=ts, gamma=gamma, p_inf=0.0, ch=ch)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), states=states, out=out,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {n} records")


def mhd_psi_1d(n):
    """Brio-Wu with a Gaussian bump of the GLM field psi and a ripple in Bx, so
    the damping source and the psi/Bx coupling of the flux both act."""
    w = problems.brio_wu(n)
    x = (np.arange(n) + 0.5) / n
    w[:, 8] = 0.2 * np.exp(-((x - 0.3) / 0.05) ** 2)
    w[:, 5] = 0.75 + 0.1 * np.sin(2 * np.pi * x)
    return w


def mhd_2d_ic(n):
    """Orszag-Tang-like smooth vortex with a density step (2D MHD fixture)."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    w = np.zeros((n, n, 9))
    w[..., 0] = np.where(X < 0.5, 1.0, 0.5)
    w[..., 1] = -np.sin(2 * np.pi * Y)
    w[..., 2] = np.sin(2 * np.pi * X)
    w[..., 3] = 0.1
    w[..., 4] = 0.6
    w[..., 5] = -np.sin(2 * np.pi * Y) * 0.5
    w[..., 6] = np.sin(4 * np.pi * X) * 0.5
    w[..., 7] = 0.2
    return w


def hpr_params(thermal, *, rho0=1.0, p0=1.0 / 1.4, t_ref=0.0, t0=1.0, cv=2.5, gamma=1.4, p_inf=0.0, cs=1.0, mu=1e-2,
               alpha=2.0, kappa=1e-2):
    """HPR parameters as the reference's configure* helpers set the relaxation
    times (hprmaterialparameters.cpp:101-119)."""
    tau1 = 6.0 * (mu / (p0 * (cs * cs)))
    tau2 = kappa * (rho0 / (t0 * (alpha * alpha))) if thermal else 0.0
    return dict(is_thermal=int(thermal), rho0=rho0, p0=p0, t_ref=t_ref, t0=t0, cv=cv, gamma=gamma, p_inf=p_inf, cs=cs,
                tau1=tau1, tau2=tau2, alpha=alpha)


HPR_ARG_ORDER = ("is_thermal", "rho0", "p0", "t_ref", "t0", "cv", "gamma", "p_inf", "cs", "tau1", "tau2", "alpha")


def hpr_call(w, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, subcycles, params, check_api):
    with tempfile.TemporaryDirectory() as tmp:
        ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
        np.ascontiguousarray(w, dtype=np.float64).tofile(os.path.join(tmp, "in.f64"))
        t = "inf" if np.isinf(tfinal) else repr(float(tfinal))
        log = run(["hpr_run", order, dim, ny, nx, repr(dx), repr(cfl), t, max_steps, repr(bias), limiter, subcycles,
                   *[repr(params[k]) if k != "is_thermal" else params[k] for k in HPR_ARG_ORDER], "in.f64",
                   "state.f64", "dt.f64", int(check_api)], tmp)
        state = np.fromfile(os.path.join(tmp, "state.f64")).reshape(w.shape)
        dt = np.fromfile(os.path.join(tmp, "dt.f64"))
    if check_api and max_steps == 0 and order == 2 and dim == 1 and "API_MATCH 1" not in log:
        raise RuntimeError("HPR mirror loop differs from the reference's own solve")
    return state, dt


def hpr_run(name, w, **kw):
    state, dt = hpr_call(w, **kw)
    meta = dict(kw)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=w, prim_out=state, dt=dt,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {len(dt)} steps")


def hpr_state(rho, p, u, v, w=0.0, A=None, J=(0.0, 0.0, 0.0)):
    s = np.zeros(17)
    s[0], s[1] = rho, p
    s[2:11] = (np.eye(3) if A is None else A).ravel()
    s[11:14] = (u, v, w)
    s[14:17] = J
    return s


def hpr_stokes(n):
    """HPRTests::solveStokesFirstProblem (hprtests.cpp:7-31): v = -0.1 / +0.1 shear layer."""
    return np.stack([hpr_state(1.0, 1.0 / 1.4, 0.0, -0.1 if i <= 0.5 * n else 0.1) for i in range(n)])


def hpr_heat(n):
    """HPRTests::solveHeatConductionProblem (hprtests.cpp:33-57): density jump at equal pressure."""
    rows = []
    for i in range(n):
        rho = 2.0 if i <= 0.5 * n else 0.5
        rows.append(hpr_state(rho, 1.0, 0.0, 0.0, A=(rho ** (1.0 / 3.0)) * np.eye(3)))
    return np.stack(rows)


def hpr_2d(n, seed=0):
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    w = np.zeros((n, n, 17))
    w[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * X)
    w[..., 1] = 1.0 / 1.4 + 0.05 * np.cos(2 * np.pi * Y)
    A = np.eye(3)[None, None] + 0.02 * rng.standard_normal((n, n, 3, 3))
    w[..., 2:11] = A.reshape(n, n, 9)
    w[..., 11] = 0.1 * np.sin(2 * np.pi * Y)
    w[..., 12] = -0.1 * np.sin(2 * np.pi * X)
    w[..., 13] = 0.01
    return w


def hpr_units(name, n, seed, thermal):
    rng = np.random.default_rng(seed)
    params = hpr_params(thermal)
    base = np.zeros((n, 17))
    base[:, 0] = rng.uniform(0.8, 1.2, n)
    base[:, 1] = rng.uniform(0.5, 1.0, n)
    base[:, 2:11] = (np.eye(3).ravel()[None] + 0.05 * rng.standard_normal((n, 9)))
    base[:, 11:14] = rng.uniform(-0.2, 0.2, (n, 3))
    if thermal:
        base[:, 14:17] = rng.uniform(-0.1, 0.1, (n, 3))
    st = np.empty((n, 5, 17))
    for k in range(5):
        scale = rng.choice([1e-7, 1e-3, 0.02], size=(n, 1))
        s = base + rng.standard_normal((n, 17)) * scale * np.abs(base).clip(0.05)
        if not thermal:
            s[:, 14:17] = 0.0
        st[:, k] = s
    states = st.reshape(n, 85)
    dx, ts = 1e-2, 2e-3
    with tempfile.TemporaryDirectory() as tmp:
        states.tofile(os.path.join(tmp, "in.f64"))
        run(["hpr_units", n, "0.0", 0, repr(dx), repr(ts),
             *[repr(params[k]) if k != "is_thermal" else params[k] for k in HPR_ARG_ORDER], "in.f64", "out.f64"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.f64")).reshape(n, 156)
    meta = dict(bias=0.0, limiter=0, dx=dx, ts=ts, params=params)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), states=states, out=out,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {n} records")


# ---- HPR with the full material parameters (all three EOS, plastic relaxation) ----

def hpr_full_vec(eos, *, thermal=0, plastic=0, rho0=1.0, p_ref=0.0, t_ref=0.0, t0=1.0, cv=2.5, gamma=1.4, p_inf=0.0,
                 c0=0.0, gamma0=0.0, s_h=0.0, cs=1.0, tau1=8.4e-3, tau2=0.0, sigma_y=0.0, n_pow=1.0, alpha=2.0,
                 alpha_gr=1.0):
    """The 26 values of the reference's full HPRMaterialParameters constructor as
    the harness reads them (EOS code, is_thermal, is_plastic, then its 23 doubles)."""
    return [float(eos), float(thermal), float(plastic), 8.31445985, rho0, p_ref, t_ref, t0, cv, gamma, p_inf, c0, gamma0,
            s_h, 0.0, cs, tau1, tau2, 0.0, 0.75, sigma_y, n_pow, alpha, alpha_gr, 3.0, 2.0]


def _oracle_params(pv):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    return oracle_lib.hpr_params_from_full(pv)


def hpr_run_full(name, w, pv, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, subcycles):
    with tempfile.TemporaryDirectory() as tmp:
        ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
        np.ascontiguousarray(w, dtype=np.float64).tofile(os.path.join(tmp, "in.f64"))
        np.asarray(pv, dtype=np.float64).tofile(os.path.join(tmp, "params.f64"))
        t = "inf" if np.isinf(tfinal) else repr(float(tfinal))
        run(["hpr_run_p", order, dim, ny, nx, repr(dx), repr(cfl), t, max_steps, repr(bias), limiter, subcycles,
             "params.f64", "in.f64", "state.f64", "dt.f64", 0], tmp)
        state = np.fromfile(os.path.join(tmp, "state.f64")).reshape(w.shape)
        dt = np.fromfile(os.path.join(tmp, "dt.f64"))
    if not np.isfinite(state).all():
        raise RuntimeError(f"{name}: non-finite reference state")
    meta = dict(order=order, dim=dim, dx=dx, cfl=cfl, tfinal=tfinal, max_steps=max_steps, bias=bias, limiter=limiter,
                subcycles=subcycles, params=_oracle_params(pv), params_full=list(pv))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=w, prim_out=state, dt=dt,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {len(dt)} steps")


def hpr_eos_units_fixture(name, pv, n, seed, rho_lo, rho_hi):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 11))
    st[:, 0] = rng.uniform(rho_lo, rho_hi, n)
    st[:, 1] = rng.uniform(0.05, 2.0, n)
    st[:, 2:11] = np.eye(3).ravel()[None] + 0.1 * rng.standard_normal((n, 9))
    with tempfile.TemporaryDirectory() as tmp:
        st.tofile(os.path.join(tmp, "in.f64"))
        np.asarray(pv, dtype=np.float64).tofile(os.path.join(tmp, "params.f64"))
        run(["hpr_eos_units", n, "params.f64", "in.f64", "out.f64"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.f64")).reshape(n, 8)
    meta = dict(params=_oracle_params(pv), params_full=list(pv))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), states=st, out=out,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {n} records")


def hpr_driver_ic(name, which, n):
    """The initial condition and parameters of HPRTests::solveBartonTest1 /
    solveZhangElasticPlasticTest, built by the reference's own constructors."""
    with tempfile.TemporaryDirectory() as tmp:
        log = run(["hpr_driver", which, n, 0, -1, "params.f64", "in.f64", "state.f64", "dt.f64"], tmp)
        pv = np.fromfile(os.path.join(tmp, "params.f64"))
        w = np.fromfile(os.path.join(tmp, "in.f64")).reshape(n, 17)
    dx = tfinal = None
    for line in log.splitlines():
        if line.startswith("DX"):
            f = line.split()
            dx, tfinal = float(f[1]), float(f[3])
    meta = dict(params=_oracle_params(pv), params_full=[float(x) for x in pv], dx=dx, tfinal=tfinal, cfl=0.8,
                limiter=0, bias=0.0)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=w,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: initial condition of {n} cells")


def hpr_tube(n, left, right, shear=0.0):
    """A 1D tube: (rho, p, u) left / right of the middle, distortion A = I plus an
    off-diagonal shear (which loads the plastic relaxation)."""
    rows = []
    for i in range(n):
        rho, p, u = left if i <= 0.5 * n else right
        A = np.eye(3)
        A[0, 1] = shear
        rows.append(hpr_state(rho, p, u, 0.0, A=A))
    return np.stack(rows)


SMG = dict(rho0=1.0, p_ref=0.1, t_ref=0.0, cv=1.0, c0=1.5, gamma0=2.0, s_h=1.4, cs=0.5)
GR = dict(rho0=1.0, t_ref=300.0, t0=300.0, cv=1e-3, gamma=2.0, c0=3.0, cs=1.0, alpha_gr=1.0)


def rgfm_call(m1, m2, phi, *, dim, dx, cfl, tfinal, max_steps, bias, limiter, p1, p2, check_api):
    with tempfile.TemporaryDirectory() as tmp:
        ny, nx = (1, phi.shape[0]) if dim == 1 else phi.shape[:2]
        for nm, a in (("in1.f64", m1), ("in2.f64", m2), ("phi.f64", phi)):
            np.ascontiguousarray(a, dtype=np.float64).tofile(os.path.join(tmp, nm))
        t = "inf" if np.isinf(tfinal) else repr(float(tfinal))
        log = run(["rgfm_run", dim, ny, nx, repr(dx), repr(cfl), t, max_steps, repr(bias), limiter,
                   *[repr(v) for v in p1], *[repr(v) for v in p2], "in1.f64", "in2.f64", "phi.f64", "o1.f64",
                   "o2.f64", "ophi.f64", "dt.f64", int(check_api)], tmp)
        o1 = np.fromfile(os.path.join(tmp, "o1.f64")).reshape(m1.shape)
        o2 = np.fromfile(os.path.join(tmp, "o2.f64")).reshape(m2.shape)
        ophi = np.fromfile(os.path.join(tmp, "ophi.f64")).reshape(phi.shape)
        dt = np.fromfile(os.path.join(tmp, "dt.f64"))
    if check_api and max_steps == 0 and "API_MATCH 1" not in log:
        raise RuntimeError("RGFM mirror loop differs from the reference's own solveHLLC")
    return o1, o2, ophi, dt


def rgfm_run(name, m1, m2, phi, **kw):
    o1, o2, ophi, dt = rgfm_call(m1, m2, phi, **kw)
    meta = dict(kw)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), m1_in=m1, m2_in=m2, phi_in=phi, m1_out=o1, m2_out=o2,
                        phi_out=ophi, dt=dt, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {len(dt)} steps")


def rgfm_toro_1d(n):
    """EulerRGFMTests::solveToroTest1HLLC (eulerrgfmtests.cpp:27-45)."""
    m1 = np.tile([1.0, 0.0, 0.0, 0.0, 1.0, 1.0], (n, 1))
    m2 = np.tile([0.125, 0.0, 0.0, 0.0, 0.1, 1.0], (n, 1))
    phi = (1.0 / n) * np.arange(n) - 0.5
    return m1, m2, phi


def rgfm_fedkiw_1d(n):
    """EulerRGFMTests::solveFedkiwTestHLLC (eulerrgfmtests.cpp:76-103)."""
    x = (1.0 / n) * np.arange(n)
    m1 = np.tile([1.0, 0.0, 0.0, 0.0, 1e5, 1.0], (n, 1))
    m1[x <= 0.05] = [1.3333, 0.3535 * np.sqrt(1e5), 0.0, 0.0, 1.5e5, 1.0]
    m2 = np.tile([0.1379, 0.0, 0.0, 0.0, 1e5, 1.0], (n, 1))
    return m1, m2, x - 0.5


def rgfm_disc_2d(n, fedkiw):
    """EulerRGFMTests::solve2DToroTest1HLLC / solve2DFedkiwTestHLLC
    (eulerrgfmtests.cpp:128-148, 183-213)."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((i - 0.5 * n) ** 2 + (j - 0.5 * n) ** 2)
    m1 = np.zeros((n, n, 6))
    m2 = np.zeros((n, n, 6))
    m1[..., 5] = m2[..., 5] = 1.0
    if fedkiw:
        m1[..., 0], m1[..., 4] = 1.0, 1e5
        inner = r <= 0.02 * n
        m1[inner, 0], m1[inner, 1], m1[inner, 4] = 1.3333, 0.3535 * np.sqrt(1e5), 1.5e5
        m2[..., 0], m2[..., 4] = 0.1379, 1e5
    else:
        m1[..., 0], m1[..., 4] = 1.0, 1.0
        m2[..., 0], m2[..., 4] = 0.125, 0.1
    return m1, m2, r / n - 0.2


def hllc_units(name, n, seed):
    rng = np.random.default_rng(seed)
    sets = [((1.4, 0.0, 25.0, 0.0, 0.25, 250.0), (1.648, 0.0, 25.0, 0.0, 0.25, 250.0)),
            ((1.4, 0.0, 25.0, 0.0, 0.25, 250.0), (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)),
            ((4.4, 0.5, 25.0, 0.0, 0.25, 250.0), (1.4, 0.0, 25.0, 0.0, 0.25, 250.0))]
    out = {}
    for k, (p1, p2) in enumerate(sets):
        st = np.empty((n, 12))
        for s in (0, 6):
            st[:, s + 0] = 10.0 ** rng.uniform(-1, 1, n)
            st[:, s + 1:s + 4] = rng.uniform(-1.5, 1.5, (n, 3))
            st[:, s + 4] = 10.0 ** rng.uniform(-1, 1.5, n)
            st[:, s + 5] = 1.0
        # a third of the pairs with nearly equal pressures (the PVRS region)
        near = rng.random(n) < 0.33
        st[near, 10] = st[near, 4] * rng.uniform(0.8, 1.25, near.sum())
        with tempfile.TemporaryDirectory() as tmp:
            st.tofile(os.path.join(tmp, "in.f64"))
            run(["hllc_units", n, *[repr(v) for v in p1], *[repr(v) for v in p2], "in.f64", "out.f64"], tmp)
            out[f"out{k}"] = np.fromfile(os.path.join(tmp, "out.f64")).reshape(n, 24)
        out[f"states{k}"] = st
        out[f"p1_{k}"] = np.array(p1)
        out[f"p2_{k}"] = np.array(p2)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: {n} x {len(sets)} records")


def rgfm_bubble_small():
    m1, m2, phi, dx, dy = problems.shock_bubble_2d(96, 32, length=(0.325, 0.325 * 32 / 96), center=(0.175, 0.325 / 6),
                                                  radius=0.03)
    return m1, m2, phi, dx


# ---- two-material diffuse-interface Euler (kind 0 multiphysics, 1 reduced) ----
MP_NP = (8, 7)
MP_NF = (9, 8)
AIR = (1.4, 0.0, 25.0, 0.0, 0.25, 250.0)       # EulerMaterialParameters(1.4)
HELIUM = (1.67, 0.0, 25.0, 0.0, 0.25, 250.0)   # EulerMaterialParameters(1.67)
STIFF_B = (4.4, 0.6, 25.0, 0.0, 0.25, 250.0)


def mp_run(name, kind, w, *, order, dim, dx, cfl, tfinal, max_steps, bias, limiter, reinit, p1, p2, check_api):
    with tempfile.TemporaryDirectory() as tmp:
        ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
        np.ascontiguousarray(w, dtype=np.float64).tofile(os.path.join(tmp, "in.f64"))
        t = "inf" if np.isinf(tfinal) else repr(float(tfinal))
        log = run(["mp_run", kind, order, dim, ny, nx, repr(dx), repr(cfl), t, max_steps, repr(bias), limiter, reinit,
                   *[repr(p) for p in p1], *[repr(p) for p in p2], "in.f64", "state.f64", "dt.f64", int(check_api)],
                  tmp)
        state = np.fromfile(os.path.join(tmp, "state.f64")).reshape(w.shape)
        dt = np.fromfile(os.path.join(tmp, "dt.f64"))
    api = None
    for line in log.splitlines():
        if line.startswith("API_MATCH"):
            api = int(line.split()[1])
    if check_api and max_steps == 0 and api != 1:
        raise RuntimeError(f"{name}: mirror loop differs from the reference's own solve")
    meta = dict(kind=kind, order=order, dim=dim, dx=dx, cfl=cfl, tfinal=tfinal, max_steps=max_steps, bias=bias,
                limiter=limiter, reinit=reinit, p1=list(p1), p2=list(p2),
                api_checked=bool(check_api and max_steps == 0))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=w, prim_out=state, dt=dt,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {len(dt)} steps")


def mp_random_states(rng, kind, n):
    """Records of four primitive states: volume fractions across the clamps,
    non-zero w (the multiphysics conversion adds w + w), moderate jumps."""
    npf = MP_NP[kind]
    st = np.empty((n, 4, npf))
    base = np.empty((n, npf))
    base[:, 0] = rng.choice([0.0004, 0.001, 0.3, 0.5, 0.7, 0.999, 0.9996], n)
    base[:, 1:4] = rng.uniform(-0.8, 0.8, (n, 3))
    base[:, 4:] = rng.uniform(0.4, 2.0, (n, npf - 4))
    for k in range(4):
        st[:, k] = base * (1.0 + rng.uniform(-0.3, 0.3, (n, npf)))
        st[:, k, 0] = np.clip(base[:, 0] + rng.uniform(-0.2, 0.2, n), 0.0002, 0.9998)
    return st.reshape(n, 4 * npf)


def mp_units_fixture(name, kind, n, seed, bias, limiter, p1, p2):
    rng = np.random.default_rng(seed)
    states = mp_random_states(rng, kind, n)
    dx, ts = 0.01, 0.002
    npf, nf = MP_NP[kind], MP_NF[kind]
    with tempfile.TemporaryDirectory() as tmp:
        states.tofile(os.path.join(tmp, "in.f64"))
        run(["mp_units", kind, n, repr(bias), limiter, repr(dx), repr(ts), *[repr(p) for p in p1],
             *[repr(p) for p in p2], "in.f64", "out.f64"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.f64")).reshape(n, 7 * nf + 3 * npf + 2)
    meta = dict(kind=kind, bias=bias, limiter=limiter, dx=dx, ts=ts, p1=list(p1), p2=list(p2))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), states=states, out=out,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {n} records")


def mp_reinit_field(kind, ny, nx, seed):
    """A volume-fraction field with crossings of both kinds along rows and
    columns, none within two cells of an edge (where the reference indexes
    outside its vectors), on varied densities and pressures."""
    rng = np.random.default_rng(seed)
    npf = MP_NP[kind]
    w = np.empty((ny, nx, npf))
    w[..., 1:4] = rng.uniform(-0.5, 0.5, (ny, nx, 3))
    w[..., 4:] = rng.uniform(0.3, 2.0, (ny, nx, npf - 4))
    i = np.arange(ny)[:, None]
    j = np.arange(nx)[None, :]
    blob = ((i - ny / 2) ** 2 / (0.3 * ny) ** 2 + (j - nx / 2) ** 2 / (0.3 * nx) ** 2) <= 1.0
    stripe = (j >= 3) & (j <= 4) & (i >= 3) & (i < ny - 3)   # two crossings three cells apart
    inside = blob | stripe
    w[..., 0] = np.where(inside, rng.uniform(0.6, 0.999, (ny, nx)), rng.uniform(0.001, 0.4, (ny, nx)))
    w[3:, :, 0] = np.where(np.abs(w[3:, :, 0] - 0.5) < 1e-3, 0.6, w[3:, :, 0])
    w[:, :2, 0] = 0.9          # no crossing within two cells of the row ends
    w[:, -2:, 0] = 0.9
    w[:2, :, 0] = 0.9
    w[-2:, :, 0] = 0.9
    return w


def mp_reinit_fixture(name, kind, w, dim, p1, p2):
    ny, nx = (1, w.shape[0]) if dim == 1 else w.shape[:2]
    with tempfile.TemporaryDirectory() as tmp:
        np.ascontiguousarray(w, dtype=np.float64).tofile(os.path.join(tmp, "in.f64"))
        run(["mp_reinit", kind, dim, ny, nx, *[repr(p) for p in p1], *[repr(p) for p in p2], "in.f64", "out.f64"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.f64")).reshape(w.shape)
    meta = dict(kind=kind, dim=dim, p1=list(p1), p2=list(p2))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), prim_in=w, prim_out=out,
                        meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: reinitialised")


def mp_job(name, kind, ic, n, *, order=2, dim=1, tfinal, max_steps=0, reinit=0, p1=AIR, p2=AIR, limiter=0,
           bias=0.0, cfl=0.8):
    return lambda: mp_run(name, kind, ic(n, kind), order=order, dim=dim, dx=1.0 / n, cfl=cfl, tfinal=tfinal,
                          max_steps=max_steps, bias=bias, limiter=limiter, reinit=reinit, p1=p1, p2=p2,
                          check_api=True)


MP_JOBS = {
    "mp_toro100": mp_job("mp_toro100", 0, problems.two_material_toro_1d, 100, tfinal=0.25),
    "mp_toro100_r5": mp_job("mp_toro100_r5", 0, problems.two_material_toro_1d, 100, tfinal=0.25, reinit=5),
    "mp_fedkiw100_r10": mp_job("mp_fedkiw100_r10", 0, problems.two_material_fedkiw_1d, 100, tfinal=0.0012,
                               reinit=10, p2=HELIUM, limiter=1),
    "mp_force_fedkiw100": mp_job("mp_force_fedkiw100", 0, problems.two_material_fedkiw_1d, 100, order=1,
                                 tfinal=0.0012, p2=HELIUM),
    "rd_toro100_r5": mp_job("rd_toro100_r5", 1, problems.two_material_toro_1d, 100, tfinal=0.25, reinit=5),
    "rd_fedkiw100_r10": mp_job("rd_fedkiw100_r10", 1, problems.two_material_fedkiw_1d, 100, tfinal=0.0012,
                               reinit=10, p2=HELIUM, limiter=2, bias=0.3, max_steps=40),
    "rd_force_toro100": mp_job("rd_force_toro100", 1, problems.two_material_toro_1d, 100, order=1, tfinal=0.25),
    "mp_2d_toro32_r4": mp_job("mp_2d_toro32_r4", 0, problems.two_material_toro_2d, 32, dim=2, tfinal=0.125,
                              reinit=4),
    "mp_2d_fedkiw32_r3": mp_job("mp_2d_fedkiw32_r3", 0, problems.two_material_fedkiw_2d, 32, dim=2, tfinal=0.0006,
                                reinit=3, p2=HELIUM),
    "rd_2d_toro32_r4": mp_job("rd_2d_toro32_r4", 1, problems.two_material_toro_2d, 32, dim=2, tfinal=0.125,
                              reinit=4, limiter=1),
    "rd_2d_fedkiw32_r3": mp_job("rd_2d_fedkiw32_r3", 1, problems.two_material_fedkiw_2d, 32, dim=2, tfinal=0.0006,
                                reinit=3, p2=HELIUM),
    "mp_force_2d_toro24": mp_job("mp_force_2d_toro24", 0, problems.two_material_toro_2d, 24, order=1, dim=2,
                                 tfinal=0.125),
    "mp_units": lambda: mp_units_fixture("mp_units", 0, 400, 41, 0.0, 0, AIR, STIFF_B),
    "mp_units_vl_bias": lambda: mp_units_fixture("mp_units_vl_bias", 0, 300, 42, 0.4, 1, STIFF_B, HELIUM),
    "rd_units": lambda: mp_units_fixture("rd_units", 1, 400, 43, 0.0, 0, AIR, STIFF_B),
    "rd_units_mb_bias": lambda: mp_units_fixture("rd_units_mb_bias", 1, 300, 44, -0.3, 2, HELIUM, STIFF_B),
    "mp_reinit_2d": lambda: mp_reinit_fixture("mp_reinit_2d", 0, mp_reinit_field(0, 20, 27, 5), 2, AIR, HELIUM),
    "rd_reinit_2d": lambda: mp_reinit_fixture("rd_reinit_2d", 1, mp_reinit_field(1, 23, 18, 6), 2, AIR, HELIUM),
    "mp_reinit_1d": lambda: mp_reinit_fixture("mp_reinit_1d", 0, mp_reinit_field(0, 7, 40, 7)[3], 1, AIR, HELIUM),
    "rd_reinit_1d": lambda: mp_reinit_fixture("rd_reinit_1d", 1, mp_reinit_field(1, 7, 40, 8)[3], 1, AIR, HELIUM),
}


JOBS = {
    "rgfm_toro1d_hllc": lambda: rgfm_run("rgfm_toro1d_hllc", *rgfm_toro_1d(200), dim=1, dx=1 / 200, cfl=0.8,
                                         tfinal=0.25, max_steps=0, bias=0.0, limiter=0, p1=IDEAL, p2=IDEAL,
                                         check_api=True),
    "rgfm_fedkiw1d_hllc": lambda: rgfm_run("rgfm_fedkiw1d_hllc", *rgfm_fedkiw_1d(200), dim=1, dx=1 / 200, cfl=0.8,
                                           tfinal=0.0012, max_steps=0, bias=0.0, limiter=0, p1=IDEAL,
                                           p2=(1.67, 0.0, 25.0, 0.0, 0.25, 250.0), check_api=True),
    "rgfm_2d_toro32": lambda: rgfm_run("rgfm_2d_toro32", *rgfm_disc_2d(32, False), dim=2, dx=1 / 32, cfl=0.8,
                                       tfinal=np.inf, max_steps=12, bias=0.0, limiter=0, p1=IDEAL, p2=IDEAL,
                                       check_api=False),
    "rgfm_2d_fedkiw32": lambda: rgfm_run("rgfm_2d_fedkiw32", *rgfm_disc_2d(32, True), dim=2, dx=1 / 32, cfl=0.8,
                                         tfinal=np.inf, max_steps=12, bias=0.0, limiter=1, p1=IDEAL,
                                         p2=(1.67, 0.0, 25.0, 0.0, 0.25, 250.0), check_api=False),
    "rgfm_2d_fedkiw24_api": lambda: rgfm_run("rgfm_2d_fedkiw24_api", *rgfm_disc_2d(24, True), dim=2, dx=1 / 24,
                                             cfl=0.8, tfinal=2e-5, max_steps=0, bias=0.0, limiter=0, p1=IDEAL,
                                             p2=(1.67, 0.0, 25.0, 0.0, 0.25, 250.0), check_api=True),
    "rgfm_2d_bubble96": lambda: rgfm_run("rgfm_2d_bubble96", *rgfm_bubble_small()[:3], dim=2, dx=rgfm_bubble_small()[3],
                                         cfl=0.8, tfinal=np.inf, max_steps=15, bias=0.0, limiter=0,
                                         p1=problems.HELIUM, p2=problems.AIR, check_api=False),
    "hllc_units": lambda: hllc_units("hllc_units", 300, 41),
    "hpr_stokes100": lambda: hpr_run("hpr_stokes100", hpr_stokes(100), order=2, dim=1, dx=0.01, cfl=0.8, tfinal=0.05,
                                     max_steps=0, bias=0.0, limiter=0, subcycles=0, params=hpr_params(False),
                                     check_api=True),
    "hpr_stokes100_sub1": lambda: hpr_run("hpr_stokes100_sub1", hpr_stokes(100), order=2, dim=1, dx=0.01, cfl=0.8,
                                          tfinal=0.05, max_steps=0, bias=0.0, limiter=0, subcycles=1,
                                          params=hpr_params(False), check_api=True),
    "hpr_heat100": lambda: hpr_run("hpr_heat100", hpr_heat(100), order=2, dim=1, dx=0.01, cfl=0.8, tfinal=0.02,
                                   max_steps=0, bias=0.0, limiter=1, subcycles=0, params=hpr_params(True),
                                   check_api=True),
    "hpr_heat100_sub1": lambda: hpr_run("hpr_heat100_sub1", hpr_heat(100), order=2, dim=1, dx=0.01, cfl=0.8,
                                        tfinal=0.02, max_steps=0, bias=0.0, limiter=1, subcycles=1,
                                        params=hpr_params(True), check_api=True),
    "hpr_force_stokes64": lambda: hpr_run("hpr_force_stokes64", hpr_stokes(64), order=1, dim=1, dx=1 / 64, cfl=0.8,
                                          tfinal=0.05, max_steps=0, bias=0.0, limiter=0, subcycles=0,
                                          params=hpr_params(False), check_api=False),
    "hpr_2d16": lambda: hpr_run("hpr_2d16", hpr_2d(16), order=2, dim=2, dx=1 / 16, cfl=0.8, tfinal=np.inf,
                                max_steps=8, bias=0.0, limiter=0, subcycles=0, params=hpr_params(False),
                                check_api=False),
    "hpr_2d16_thermal_sub1": lambda: hpr_run("hpr_2d16_thermal_sub1", hpr_2d(16, 1), order=2, dim=2, dx=1 / 16,
                                             cfl=0.8, tfinal=np.inf, max_steps=5, bias=0.0, limiter=1, subcycles=1,
                                             params=hpr_params(True), check_api=False),
    "hpr_units": lambda: hpr_units("hpr_units", 300, 31, False),
    "hpr_eos_units_sg_plastic": lambda: hpr_eos_units_fixture(
        "hpr_eos_units_sg_plastic", hpr_full_vec(0, plastic=1, p_inf=0.3, sigma_y=0.05, n_pow=1.7), 200, 41, 0.5, 2.0),
    "hpr_eos_units_smg": lambda: hpr_eos_units_fixture(
        "hpr_eos_units_smg", hpr_full_vec(1, plastic=1, sigma_y=0.2, n_pow=2.0, **SMG), 300, 42, 0.7, 1.4),
    "hpr_eos_units_gr": lambda: hpr_eos_units_fixture(
        "hpr_eos_units_gr", hpr_full_vec(2, plastic=1, sigma_y=0.3, n_pow=1.338, **GR), 300, 43, 0.8, 1.25),
    "hpr_smg_tube100": lambda: hpr_run_full(
        "hpr_smg_tube100", hpr_tube(100, (1.15, 1.0, 0.0), (0.92, 0.1, 0.0)), hpr_full_vec(1, **SMG), order=2, dim=1,
        dx=0.01, cfl=0.8, tfinal=0.1, max_steps=0, bias=0.0, limiter=0, subcycles=0),
    "hpr_smg_plastic_sub1": lambda: hpr_run_full(
        "hpr_smg_plastic_sub1", hpr_tube(100, (1.1, 1.0, 0.2), (0.95, 0.3, -0.2), shear=0.05),
        hpr_full_vec(1, plastic=1, sigma_y=0.2, n_pow=2.0, tau1=0.05, **SMG), order=2, dim=1, dx=0.01, cfl=0.8,
        tfinal=0.05, max_steps=0, bias=0.0, limiter=1, subcycles=1),
    "hpr_sg_plastic_sub1": lambda: hpr_run_full(
        "hpr_sg_plastic_sub1", hpr_tube(100, (1.0, 1.0 / 1.4, 0.1), (1.0, 1.0 / 1.4, -0.1), shear=0.08),
        hpr_full_vec(0, plastic=1, sigma_y=0.05, n_pow=1.5, tau1=0.02), order=2, dim=1, dx=0.01, cfl=0.8,
        tfinal=0.05, max_steps=0, bias=0.0, limiter=0, subcycles=1),
    "hpr_2d16_plastic_sub1": lambda: hpr_run_full(
        "hpr_2d16_plastic_sub1", hpr_2d(16, 2), hpr_full_vec(0, plastic=1, sigma_y=0.5, n_pow=1.5, tau1=0.05),
        order=2, dim=2, dx=1 / 16, cfl=0.8, tfinal=np.inf, max_steps=6, bias=0.0, limiter=0, subcycles=1),
    "hpr_driver_barton100": lambda: hpr_driver_ic("hpr_driver_barton100", "barton", 100),
    "hpr_driver_zhang100": lambda: hpr_driver_ic("hpr_driver_zhang100", "zhang", 100),
    "hpr_units_thermal": lambda: hpr_units("hpr_units_thermal", 300, 32, True),
    "mhd_briowu256": lambda: mhd_run("mhd_briowu256", problems.brio_wu(256), order=2, dim=1, dx=1.0 / 256, cfl=0.8,
                                     tfinal=0.1, max_steps=0, bias=0.0, limiter=1, gamma=2.0, p_inf=0.0,
                                     check_api=True),
    "mhd_force_briowu128": lambda: mhd_run("mhd_force_briowu128", problems.brio_wu(128), order=1, dim=1,
                                           dx=1.0 / 128, cfl=0.8, tfinal=0.1, max_steps=0, bias=0.0, limiter=0,
                                           gamma=2.0, p_inf=0.0, check_api=True),
    # first order never sets c_h: with a varying Bx the GLM terms stay switched off
    "mhd_force_2d_ot16": lambda: mhd_run("mhd_force_2d_ot16", mhd_2d_ic(16), order=1, dim=2, dx=1.0 / 16, cfl=0.6,
                                         tfinal=0.05, max_steps=0, bias=0.0, limiter=0, gamma=5.0 / 3.0, p_inf=0.0,
                                         check_api=True),
    # RK4 GLM sub-cycles (MHDForcingSolver) around every SLIC step
    "mhd_psi128_sub2": lambda: mhd_run("mhd_psi128_sub2", mhd_psi_1d(128), order=2, dim=1, dx=1.0 / 128,
                                       cfl=0.8, tfinal=0.05, max_steps=0, bias=0.0, limiter=1, gamma=2.0,
                                       p_inf=0.0, check_api=True, subcycles=2),
    "mhd_2d_ot16_sub1": lambda: mhd_run("mhd_2d_ot16_sub1", mhd_2d_ic(16), order=2, dim=2, dx=1.0 / 16, cfl=0.6,
                                        tfinal=0.03, max_steps=0, bias=0.0, limiter=0, gamma=5.0 / 3.0, p_inf=0.0,
                                        check_api=True, subcycles=1),
    "mhd_2d_ot32": lambda: mhd_run("mhd_2d_ot32", mhd_2d_ic(32), order=2, dim=2, dx=1.0 / 32, cfl=0.6,
                                   tfinal=np.inf, max_steps=30, bias=0.0, limiter=0, gamma=5.0 / 3.0, p_inf=0.0,
                                   check_api=False),
    "mhd_batch12": lambda: mhd_batch_fixture("mhd_batch12", 12, 3, 64, 256, 0.02),
    "mhd_units": lambda: mhd_units("mhd_units", 500, 21, 0.0, 1, 2.0, 1.7),

    # the Sod benchmark exactly (1000 cells, minbee, CFL 0.9, T 0.25); also checks the
    # reference's own SecondOrderSolver::solve against the mirror loop.
    "euler_c1_sod1000": lambda: euler_run("euler_c1_sod1000", problems.sod_1d(1000), order=2, dim=1, dx=1e-3,
                                          cfl=0.9, tfinal=0.25, max_steps=0, bias=0.0, limiter=2,
                                          params=IDEAL, check_api=True),
    # EulerTests::solveToroTest1 settings (eulertests.cpp:26) at 100 cells.
    "euler_toro100": lambda: euler_run("euler_toro100", problems.sod_1d(100), order=2, dim=1, dx=1e-2, cfl=0.8,
                                       tfinal=0.25, max_steps=0, bias=0.0, limiter=0, params=IDEAL,
                                       check_api=True),
    "euler_1d_stiff_bias": lambda: euler_run("euler_1d_stiff_bias", stiff_1d(200), order=2, dim=1, dx=5e-3,
                                             cfl=0.7, tfinal=0.05, max_steps=0, bias=0.4, limiter=1,
                                             params=STIFF, check_api=True),
    "euler_force1d": lambda: euler_run("euler_force1d", problems.sod_1d(200), order=1, dim=1, dx=5e-3, cfl=0.9,
                                       tfinal=0.2, max_steps=0, bias=0.0, limiter=0, params=IDEAL,
                                       check_api=True),
    "euler_2d_disc64_l0": lambda: euler_run("euler_2d_disc64_l0", problems.disc_2d(64, gaussians=0, length=1.0),
                                            order=2, dim=2, dx=1.0 / 64, cfl=0.8, tfinal=np.inf, max_steps=200,
                                            bias=0.0, limiter=0, params=IDEAL, check_api=False),
    "euler_2d_disc64_l1": lambda: euler_run("euler_2d_disc64_l1", problems.disc_2d(64, gaussians=0, length=1.0),
                                            order=2, dim=2, dx=1.0 / 64, cfl=0.8, tfinal=np.inf, max_steps=120,
                                            bias=0.0, limiter=1, params=IDEAL, check_api=False),
    "euler_2d_disc64_l2": lambda: euler_run("euler_2d_disc64_l2", problems.disc_2d(64, gaussians=0, length=1.0),
                                            order=2, dim=2, dx=1.0 / 64, cfl=0.8, tfinal=np.inf, max_steps=120,
                                            bias=0.0, limiter=2, params=IDEAL, check_api=False),
    # the 2D Euler headline's initial condition (disc + 64 Gaussians, [0,2]^2) at 64^2.
    "euler_2d_c2ic64": lambda: euler_run("euler_2d_c2ic64", problems.disc_2d(64), order=2, dim=2, dx=2.0 / 64,
                                         cfl=0.8, tfinal=np.inf, max_steps=200, bias=0.0, limiter=0,
                                         params=IDEAL, check_api=False),
    "euler_2d_rect": lambda: euler_run("euler_2d_rect", rect_2d(40, 72), order=2, dim=2, dx=1.0 / 72, cfl=0.6,
                                       tfinal=np.inf, max_steps=60, bias=0.2, limiter=0, params=STIFF,
                                       check_api=False),
    "euler_force2d": lambda: euler_run("euler_force2d", problems.disc_2d(48, gaussians=0, length=1.0), order=1,
                                       dim=2, dx=1.0 / 48, cfl=0.8, tfinal=np.inf, max_steps=100, bias=0.0,
                                       limiter=0, params=IDEAL, check_api=False),
    # solve2D to a final time through the reference's own entry point.
    "euler_2d_api32": lambda: euler_run("euler_2d_api32", problems.disc_2d(32, gaussians=0, length=1.0), order=2,
                                        dim=2, dx=1.0 / 32, cfl=0.8, tfinal=0.05, max_steps=0, bias=0.0,
                                        limiter=0, params=IDEAL, check_api=True),
    "euler_units_ideal": lambda: euler_units("euler_units_ideal", 1000, 11, 0.0, 0, IDEAL, False),
    "euler_units_vl_bias": lambda: euler_units("euler_units_vl_bias", 1000, 12, 0.3, 1, STIFF, True),
    "euler_units_minbee": lambda: euler_units("euler_units_minbee", 1000, 13, -0.2, 2, STIFF, True),
    "limiter_units_b0": lambda: limiter_units("limiter_units_b0", 0.0),
    "limiter_units_b03": lambda: limiter_units("limiter_units_b03", 0.3),
}


JOBS.update(MP_JOBS)


# ---- two-material MHD RGFM ------------------------------------------------------
def mhd_rgfm_call(m1, m2, phi, *, dim, dx, cfl, tfinal, max_steps, bias, limiter, subcycles, p1, p2, check_api):
    with tempfile.TemporaryDirectory() as tmp:
        ny, nx = (1, phi.shape[0]) if dim == 1 else phi.shape[:2]
        for nm, a in (("in1.f64", m1), ("in2.f64", m2), ("phi.f64", phi)):
            np.ascontiguousarray(a, dtype=np.float64).tofile(os.path.join(tmp, nm))
        t = "inf" if np.isinf(tfinal) else repr(float(tfinal))
        log = run(["mhd_rgfm_run", dim, ny, nx, repr(dx), repr(cfl), t, max_steps, repr(bias), limiter, subcycles,
                   *[repr(v) for v in p1], *[repr(v) for v in p2], "in1.f64", "in2.f64", "phi.f64", "o1.f64",
                   "o2.f64", "ophi.f64", "dt.f64", int(check_api)], tmp)
        o1 = np.fromfile(os.path.join(tmp, "o1.f64")).reshape(m1.shape)
        o2 = np.fromfile(os.path.join(tmp, "o2.f64")).reshape(m2.shape)
        ophi = np.fromfile(os.path.join(tmp, "ophi.f64")).reshape(phi.shape)
        dt = np.fromfile(os.path.join(tmp, "dt.f64"))
    if check_api and max_steps == 0 and "API_MATCH 1" not in log:
        raise RuntimeError("MHD RGFM mirror loop differs from the reference's own solve")
    return o1, o2, ophi, dt


def mhd_rgfm_run(name, m1, m2, phi, **kw):
    o1, o2, ophi, dt = mhd_rgfm_call(m1, m2, phi, **kw)
    meta = dict(kw)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), m1_in=m1, m2_in=m2, phi_in=phi, m1_out=o1, m2_out=o2,
                        phi_out=ophi, dt=dt, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    print(f"{name}: {len(dt)} steps, finite {bool(np.isfinite(o1).all() and np.isfinite(o2).all())}")


def mhd_hllc_units_fixture(name, n, seed):
    """Random state pairs across the HLLC branches (supersonic both ways, star
    left and right), for three material pairs with GLM speeds."""
    rng = np.random.default_rng(seed)
    sets = [((2.0, 0.0, 0.0, 0.0), (2.0, 0.0, 0.0, 0.0)), ((1.4, 0.0, 1.3, 0.5), (1.67, 0.0, 0.7, 0.4)),
            ((3.0, 0.4, 2.0, 0.6), (1.4, 0.0, 2.5, 0.7))]
    out = {}
    for k, (p1, p2) in enumerate(sets):
        st = np.empty((n, 18))
        for s in (0, 9):
            st[:, s + 0] = 10.0 ** rng.uniform(-1, 0.5, n)
            st[:, s + 1:s + 4] = rng.uniform(-1.5, 1.5, (n, 3))
            st[:, s + 4] = 10.0 ** rng.uniform(-1, 1, n)
            st[:, s + 5:s + 8] = rng.uniform(-1.2, 1.2, (n, 3))
            st[:, s + 8] = rng.uniform(-0.3, 0.3, n)
        fast = rng.random(n) < 0.2      # a fifth with a strong common drift: supersonic cases
        st[fast, 1] += 6.0 * np.sign(rng.uniform(-1, 1, fast.sum()))
        st[fast, 10] = st[fast, 1]
        st[fast, 2] += 6.0 * np.sign(rng.uniform(-1, 1, fast.sum()))
        st[fast, 11] = st[fast, 2]
        with tempfile.TemporaryDirectory() as tmp:
            st.tofile(os.path.join(tmp, "in.f64"))
            run(["mhd_hllc_units", n, *[repr(v) for v in p1], *[repr(v) for v in p2], "in.f64", "out.f64"], tmp)
            out[f"out{k}"] = np.fromfile(os.path.join(tmp, "out.f64")).reshape(n, 24)
        out[f"states{k}"] = st
        out[f"p1_{k}"] = np.array(p1)
        out[f"p2_{k}"] = np.array(p2)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: {n} x {len(sets)} records")


MHD_RGFM_JOBS = {
    "mhd_rgfm_dumbser200": lambda: mhd_rgfm_run(
        "mhd_rgfm_dumbser200", *problems.mhd_rgfm_dumbser_1d(200), dim=1, dx=1 / 200, cfl=0.8, tfinal=0.1,
        max_steps=0, bias=0.0, limiter=0, subcycles=0, p1=(2.0, 0.0, 0.0, 0.0), p2=(2.0, 0.0, 0.0, 0.0),
        check_api=True),
    "mhd_rgfm_dumbser_multi200": lambda: mhd_rgfm_run(
        "mhd_rgfm_dumbser_multi200", *problems.mhd_rgfm_dumbser_1d(200), dim=1, dx=1 / 200, cfl=0.8, tfinal=0.2,
        max_steps=0, bias=0.0, limiter=0, subcycles=0, p1=(1.4, 0.0, 0.0, 0.0), p2=(1.67, 0.0, 0.0, 0.0),
        check_api=True),
    "mhd_rgfm_vl_glm100": lambda: mhd_rgfm_run(
        "mhd_rgfm_vl_glm100", *problems.mhd_rgfm_dumbser_1d(100), dim=1, dx=1 / 100, cfl=0.7, tfinal=0.1,
        max_steps=0, bias=0.2, limiter=1, subcycles=0, p1=(1.4, 0.0, 1.5, 0.5), p2=(1.67, 0.2, 1.1, 0.45),
        check_api=True),
    "mhd_rgfm_2d_dumbser32": lambda: mhd_rgfm_run(
        "mhd_rgfm_2d_dumbser32", *problems.mhd_rgfm_dumbser_2d(32), dim=2, dx=1 / 32, cfl=0.8, tfinal=0.05,
        max_steps=0, bias=0.0, limiter=0, subcycles=0, p1=(2.0, 0.0, 0.0, 0.0), p2=(2.0, 0.0, 0.0, 0.0),
        check_api=True),
    "mhd_rgfm_2d_multi32_sub1": lambda: mhd_rgfm_run(
        "mhd_rgfm_2d_multi32_sub1", *problems.mhd_rgfm_dumbser_2d(32), dim=2, dx=1 / 32, cfl=0.8, tfinal=0.05,
        max_steps=0, bias=0.0, limiter=0, subcycles=1, p1=(1.4, 0.0, 0.0, 0.0), p2=(1.67, 0.0, 0.0, 0.0),
        check_api=True),
    "mhd_hllc_units": lambda: mhd_hllc_units_fixture("mhd_hllc_units", 300, 51),
}
JOBS.update(MHD_RGFM_JOBS)



def main(names):
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    for n in names or JOBS:
        JOBS[n]()


if __name__ == "__main__":
    main(sys.argv[1:])
