"""Synthetic inputs for the hydro step (primitive fields + box tables).

Every workload is generated on the host from a counter-based RNG (numpy's
Philox) and a fixed seed, so the GPU path and the CPU oracle start from the
same bits.  Fields are primitive: the conservative state is built by the
library's pc_prim_to_cons_v1 (device) with the EOS of the run, the analogue
of the reference's pc_initdata (Exec/RegTests/Sod/prob.H:25-117,
Exec/RegTests/Sod/prob.cpp:182-243).

Primitive component layout (PRIM_*), per cell:
  mode 0 ("rho_p"): rho, p, u, v, w, Y[0..NS-1]    (T from e = RYP2E, EY2T)
  mode 1 ("p_T"):   p,   T, u, v, w, Y[0..NS-1]    (rho = p Wbar / (RU T))
"""
import numpy as np

DEFAULT_SEED = 20261203

# species order H2 O2 H2O H O OH HO2 H2O2 N2 (LiDryer)
SPECIES = ["H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "N2"]
MW = np.array([2.01594, 31.9988, 18.01534, 1.00797, 15.9994, 17.00737,
               33.00677, 34.01474, 28.0134])

BC = dict(interior=0, inflow=1, outflow=2, symmetry=3, slipwall=4,
          noslipwall=5, userbc=6)


def mole_to_mass(X):
    X = np.asarray(X, dtype=np.float64)
    Y = X * MW
    return Y / Y.sum()


def unburnt_h2_air(phi):
    """Mass fractions of a phi H2/air mixture (air = O2 + 3.76 N2)."""
    X = np.zeros(9)
    X[0] = 2.0 * phi
    X[1] = 1.0
    X[8] = 3.76
    return mole_to_mass(X / X.sum())


def burnt_h2_air(phi):
    """Complete-combustion products of a phi H2/air mixture (excess O2 when
    lean, excess H2 when rich) with small radical pools so that every
    species is present ("equilibrium-like")."""
    X = np.zeros(9)
    X[2] = 2.0 * min(phi, 1.0)       # H2O
    X[1] = max(1.0 - phi, 0.0)       # excess O2
    X[0] = 2.0 * max(phi - 1.0, 0.0)  # excess H2
    X[8] = 3.76                      # N2
    X = X / X.sum()
    X += np.array([1e-4, 0.0, 0.0, 1e-5, 1e-5, 1e-3, 1e-6, 1e-7, 0.0])
    return mole_to_mass(X / X.sum())


class Workload:
    """A single-level problem: box table, domain, BCs, dx and primitive
    fields per box (valid cells only; ghosts are filled by the level)."""

    def __init__(self, name, nspec, boxes, domlo, domhi, bclo, bchi, dx,
                 prim_mode, prims, periodic_self=False, params=None,
                 nsteps=1):
        self.name = name
        self.nspec = nspec
        self.boxes = boxes
        self.domlo = domlo
        self.domhi = domhi
        self.bclo = bclo
        self.bchi = bchi
        self.dx = dx
        self.prim_mode = prim_mode
        self.prims = prims            # list of (5+NS, nz, ny, nx) arrays
        self.periodic_self = periodic_self
        self.params = params or {}
        self.nsteps = nsteps

    @property
    def nvar(self):
        return 7 + self.nspec

    def valid_cells(self):
        return int(sum(np.prod([hi[d] - lo[d] + 1 for d in range(3)])
                       for lo, hi in self.boxes))


def _box_centres(lo, hi, dx, problo=(0.0, 0.0, 0.0)):
    xs = [problo[d] + (np.arange(lo[d], hi[d] + 1) + 0.5) * dx[d]
          for d in range(3)]
    return xs


def sod(n=64, seed=DEFAULT_SEED, nsteps=10, cfl=0.3, max_grid=None,
        ny=None, nz=None, idir=0, xbc="outflow", pert=1e-3, mirror=False):
    """3D Sod shock tube, gamma-law gas, 1 species.  Left (x <= 0.5):
    rho=1, p=1; right: rho=0.125, p=0.1; u=v=w=0; rho carries a seeded
    uniform +-1e-3 relative perturbation.  PPM, CFL 0.3, periodic in the
    transverse directions, `xbc` on the two ends of the tube."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    ncell = [n, ny, nz]
    ncell = ncell[-idir:] + ncell[:-idir] if idir else ncell
    dx = [1.0 / n] * 3
    domlo = [0, 0, 0]
    domhi = [ncell[0] - 1, ncell[1] - 1, ncell[2] - 1]
    mg = max_grid or max(ncell)
    boxes = []
    for k0 in range(0, ncell[2], mg):
        for j0 in range(0, ncell[1], mg):
            for i0 in range(0, ncell[0], mg):
                boxes.append(([i0, j0, k0],
                              [min(i0 + mg, ncell[0]) - 1,
                               min(j0 + mg, ncell[1]) - 1,
                               min(k0 + mg, ncell[2]) - 1]))
    rng = np.random.Generator(np.random.Philox(seed))
    pert_amp = pert
    pert = rng.uniform(-1.0, 1.0, size=(ncell[2], ncell[1], ncell[0]))
    prims = []
    for lo, hi in boxes:
        xs = _box_centres(lo, hi, dx)
        X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")[2 - idir]
        left = (X >= 0.5) if mirror else (X <= 0.5)
        sl = (slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1),
              slice(lo[0], hi[0] + 1))
        rho = np.where(left, 1.0, 0.125) * (1.0 + pert_amp * pert[sl])
        p = np.where(left, 1.0, 0.1)
        P = np.zeros((6,) + rho.shape)
        P[0] = rho
        P[1] = p
        P[5] = 1.0
        prims.append(np.ascontiguousarray(P))
    bclo = [BC["interior"]] * 3
    bchi = [BC["interior"]] * 3
    bclo[idir] = BC[xbc]
    bchi[idir] = BC[xbc]
    return Workload("sod", 1, boxes, domlo, domhi, bclo, bchi, dx, 0, prims,
                    params=dict(ppm_type=1, cfl=cfl), nsteps=nsteps)


def premixed(n=128, box=32, seed=DEFAULT_SEED, nsteps=50, phi=0.8,
             cfl=0.3, T0=300.0, T1=1800.0, width_cells=8.0, w0=100.0,
             sigma=5.0):
    """Premixed H2/air slab, 9 species (LiDryer), ideal-gas mixture EOS, no
    reactions.  T = T0 + (T1-T0) * s(z), s = (1 + tanh((z - 0.5)/(w dz)))/2;
    Y = (1-s) Y_unburnt(phi) + s Y_burnt(phi); p = 1 atm; velocity
    (0, 0, w0) cm/s plus seeded Gaussian noise (sigma cm/s, all three
    components).  Domain [0,1]^3 cm as (n/box)^3 boxes of box^3; periodic in
    x and y, outflow in z.  PPM."""
    dx = [1.0 / n] * 3
    nb = n // box
    boxes = []
    for kb in range(nb):
        for jb in range(nb):
            for ib in range(nb):
                lo = [ib * box, jb * box, kb * box]
                boxes.append((lo, [lo[0] + box - 1, lo[1] + box - 1,
                                   lo[2] + box - 1]))
    rng = np.random.Generator(np.random.Philox(seed))
    noise = rng.normal(0.0, sigma, size=(3, n, n, n))
    Yu = unburnt_h2_air(phi)
    Yb = burnt_h2_air(phi)
    prims = []
    for lo, hi in boxes:
        xs = _box_centres(lo, hi, dx)
        Z = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")[0]
        s = 0.5 * (1.0 + np.tanh((Z - 0.5) / (width_cells * dx[2])))
        sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1),
              slice(lo[0], hi[0] + 1))
        P = np.zeros((14,) + Z.shape)
        P[0] = 1013250.0
        P[1] = T0 + (T1 - T0) * s
        P[2:5] = noise[sl]
        P[4] += w0
        Y = (1.0 - s)[None] * Yu[:, None, None, None] + \
            s[None] * Yb[:, None, None, None]
        P[5:] = Y
        prims.append(np.ascontiguousarray(P))
    bclo = [BC["interior"], BC["interior"], BC["outflow"]]
    bchi = [BC["interior"], BC["interior"], BC["outflow"]]
    return Workload("premixed_h2air", 9, boxes, [0, 0, 0], [n - 1] * 3, bclo,
                    bchi, dx, 1, prims, params=dict(ppm_type=1, cfl=cfl),
                    nsteps=nsteps)


def ragged(nboxes=300, seed=DEFAULT_SEED, nsteps=100, phi=0.8,
           sides=(8, 16, 24, 32, 48, 64),
           probs=(0.1, 0.2, 0.2, 0.25, 0.15, 0.1), ppm_type=1, cfl=0.3):
    """One level as a ragged table of boxes of uneven sizes, each carrying
    the premixed generator over its own extent and acting as its own
    periodic domain (ghosts are its own interior values, wrapped)."""
    rng = np.random.Generator(np.random.Philox(seed))
    sz = rng.choice(np.asarray(sides), size=nboxes, p=np.asarray(probs))
    boxes = []
    prims = []
    Yu = unburnt_h2_air(phi)
    Yb = burnt_h2_air(phi)
    for b in range(nboxes):
        m = int(sz[b])
        lo = [0, 0, 0]
        hi = [m - 1] * 3
        dx = 1.0 / m
        z = (np.arange(m) + 0.5) * dx
        s = 0.5 * (1.0 + np.tanh((z - 0.5) / (8.0 * dx)))
        s = np.broadcast_to(s[:, None, None], (m, m, m))
        P = np.zeros((14, m, m, m))
        P[0] = 1013250.0
        P[1] = 300.0 + 1500.0 * s
        P[2:5] = rng.normal(0.0, 5.0, size=(3, m, m, m))
        P[4] += 100.0
        P[5:] = (1.0 - s)[None] * Yu[:, None, None, None] + \
            s[None] * Yb[:, None, None, None]
        boxes.append((lo, hi))
        prims.append(np.ascontiguousarray(P))
    bc = [BC["interior"]] * 3
    return Workload("ragged", 9, boxes, [0, 0, 0], [0, 0, 0], bc, bc,
                    [1.0 / 32] * 3, 1, prims, periodic_self=True,
                    params=dict(ppm_type=ppm_type, cfl=cfl), nsteps=nsteps)


def density_wave(n=32, ny=4, nz=4, amp=0.2, u=1.0, p=1.0, cfl=0.3,
                 nsteps=10, max_grid=None):
    """Smooth periodic advection, gamma-law gas: rho = 1 + amp sin(2 pi x),
    uniform velocity u along x and pressure p, periodic in every direction.
    The exact solution is the initial profile translated by u t, so the
    error after a fixed time measures the order of the scheme."""
    dx = [1.0 / n] * 3
    domhi = [n - 1, ny - 1, nz - 1]
    mg = max_grid or max(n, ny, nz)
    boxes = []
    for k0 in range(0, nz, mg):
        for j0 in range(0, ny, mg):
            for i0 in range(0, n, mg):
                boxes.append(([i0, j0, k0], [min(i0 + mg, n) - 1,
                                             min(j0 + mg, ny) - 1,
                                             min(k0 + mg, nz) - 1]))
    prims = []
    for lo, hi in boxes:
        xs = _box_centres(lo, hi, dx)
        X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")[2]
        P = np.zeros((6,) + X.shape)
        P[0] = 1.0 + amp * np.sin(2.0 * np.pi * X)
        P[1] = p
        P[2] = u
        P[5] = 1.0
        prims.append(np.ascontiguousarray(P))
    bc = [BC["interior"]] * 3
    return Workload("density_wave", 1, boxes, [0, 0, 0], domhi, bc, list(bc),
                    dx, 0, prims, params=dict(cfl=cfl), nsteps=nsteps)
