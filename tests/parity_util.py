"""Run one workload through the GPU level stepper and through the CPU oracle
on the same primitive input and compare (test infrastructure)."""
import numpy as np

import oracle_bind as ob

VALID = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))


def compare(gpu, ref, tol=1e-12, scale=None):
    """Relative error of each component against that component's scale
    (max |ref| of the component unless `scale` gives one per component; inf
    when either side holds a non-finite value).
    Returns (max_err, n_not_bitwise, n_total)."""
    assert gpu.shape == ref.shape
    nb = int(np.count_nonzero(gpu != ref))
    if not (np.isfinite(gpu).all() and np.isfinite(ref).all()):
        return float("inf"), nb, gpu.size
    if scale is None:
        scale = np.max(np.abs(ref.reshape(ref.shape[0], -1)), axis=1)
    scale = np.where(scale > 0, scale, 1.0)
    d = np.abs(gpu - ref).reshape(ref.shape[0], -1)
    rel = np.max(d / scale[:, None]) if d.size else 0.0
    return float(rel), nb, gpu.size


def oracle_level(w, prm_over=None):
    prm = dict(w.params)
    prm.update(prm_over or {})
    H = ob.HostLevel(w.nspec, w.boxes, w.domlo, w.domhi, w.bclo, w.bchi, w.dx,
                     periodic_self=w.periodic_self)
    H.set_prims(w)
    return H, ob.params(**prm)


def run_oracle(w, nsteps, prm_over=None, dt0=None, nthreads=0, times=None):
    H, prm = oracle_level(w, prm_over)
    if dt0 is None:
        dt0 = H.init_dt(prm)
    dts = [dt0]
    dt = dt0
    for _ in range(nsteps):
        dt = H.step(prm, 1, dt, nthreads)
        dts.append(dt)
        if times is not None:
            times.append(H.time)
    return [H.fab(b).copy() for b in range(len(w.boxes))], dts


def run_gpu(w, nsteps, prm_over=None, dt0=None, times=None):
    from particleclem_amd.hydro import DeviceLevel, HydroContext
    ctx = HydroContext(w.nspec)
    lev = DeviceLevel.from_workload(ctx, w, **(prm_over or {}))
    if dt0 is None:
        lev.init_dt()
    else:
        lev.set_dt(dt0)
    dts = [lev.get_dt()]
    for _ in range(nsteps):
        lev.step(1)
        dts.append(lev.get_dt())
        if times is not None:
            times.append(lev.get_time())
    rc, bad, small = lev.status()
    assert rc == 0, (rc, bad, small)
    out = [lev.read_box(b) for b in range(len(w.boxes))]
    lev.close()
    ctx.close()
    return out, dts


def per_cell(gpu, ref, rtol=1e-12, floor=None):
    """Largest per-value error |gpu - ref| / (rtol |ref| + floor_n), where
    floor_n (per component, default 1e-30 of the component's largest |ref|)
    keeps values that round to zero from dividing by zero: <= 1 passes."""
    if floor is None:
        s = np.max(np.abs(ref.reshape(ref.shape[0], -1)), axis=1)
        floor = 1e-30 * np.where(s > 0, s, 1.0)
    floor = np.broadcast_to(np.asarray(floor, dtype=np.float64).reshape(-1, *([1] * (ref.ndim - 1))),
                            (ref.shape[0],) + (1,) * (ref.ndim - 1))
    d = np.abs(gpu - ref) / (rtol * np.abs(ref) + floor)
    return float(d.max()) if d.size else 0.0


def assert_parity(w, nsteps, prm_over=None, tol=1e-12, label="", bitwise=True):
    """FP64 level steps on the GPU against the oracle: every valid value and
    every dt (and level time) of the chain equal bit for bit (`bitwise`), or
    else each value within tol relative to itself (floor 1e-30 of its
    component's scale)."""
    gt, rt = [], []
    g, gdt = run_gpu(w, nsteps, prm_over, times=gt)
    r, rdt = run_oracle(w, nsteps, prm_over, times=rt)
    worst = 0.0
    cellw = 0.0
    nbits = 0
    ntot = 0
    for b in range(len(w.boxes)):
        e, nb, nt = compare(g[b][VALID], r[b][VALID])
        worst = max(worst, e)
        cellw = max(cellw, per_cell(g[b][VALID], r[b][VALID], tol))
        nbits += nb
        ntot += nt
    dterr = max(abs(a - b) / b if b else abs(a) for a, b in zip(gdt, rdt))
    ndt = sum(a != b for a, b in zip(gdt + gt, rdt + rt))
    print(f"[{label}] steps={nsteps} max rel err {worst:.3e} (per value "
          f"{cellw * tol:.3e}), {nbits}/{ntot} values not bitwise equal, "
          f"dt rel err {dterr:.3e}, {ndt} dt/time values differ")
    assert np.all(np.isfinite(np.concatenate([x[VALID].ravel() for x in g])))
    if bitwise:
        assert nbits == 0, f"{nbits} of {ntot} values differ (max {worst:.3e})"
        assert ndt == 0, (gdt, rdt, gt, rt)
    assert cellw <= 1.0, cellw * tol
    assert dterr <= tol, (gdt, rdt)
    return worst, nbits


def rate_scale(U, dt):
    """Per-component scale for comparing a rate dU/dt: the error it makes in
    the updated state U + dt dU/dt relative to max |U| of the component.  A
    flux divergence cancels most of its face fluxes, so its own relative
    rounding is ~|F|/|dF| times that of the fluxes; the updated state is what
    the step keeps."""
    s = np.max(np.abs(U.reshape(U.shape[0], -1)), axis=1)
    return np.where(s > 0, s, 1.0) / dt
