"""Case builders: turn a workload description into the reference's deck form
(scalars + property table + per-cell property ids), the input the solver's
C ABI takes.

Unit mapping. The reference solver takes SI inputs (R, gamma, mu, Pr,
T_ref, Tc_Tref, dx, CFL; cpuLBM.cpp:113-265) and its WENO5 epsilon (1e-6)
acts on |e|*f in those units, so results depend on the unit system. The
lattice-unit workloads (Couette, shear wave, the ensembles) are mapped to
one fixed input set: R = 1, rho_ref = 1, T_ref = the case temperature,
p_ref = rho_ref*R*T_ref, mu = tau*p_ref, Tc_Tref = 1 (so c = sqrt(R*T)),
dt = CFL*dx/(2c). The compressible workloads (flat plate, Riemann) use the
SI decks of the reference's own case generators (makeBoundaryLayer.py,
makeLaxLiu3.py; Items/11.04.06-BoundaryLayer-Jacobs/input.mat).

Maps are [x][y] with y = 0 the bottom row (already y-flipped, as
cpuLBM::loadInput leaves bndArray, cpuLBM.cpp:254-266).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

FLUID, PERMANENT, SOLID = 0, 1, 2


@dataclass
class Case:
    name: str
    params: dict
    bnd: np.ndarray            # int32 [nx, ny] property ids
    props: dict                # columns of the property table
    nsteps: int
    meta: dict = field(default_factory=dict)

    @property
    def nx(self) -> int:
        return int(self.bnd.shape[0])

    @property
    def ny(self) -> int:
        return int(self.bnd.shape[1])

    @property
    def cells(self) -> int:
        return self.nx * self.ny

    def with_steps(self, nsteps: int) -> "Case":
        p = dict(self.params)
        p["steps"] = int(nsteps)
        return Case(self.name, p, self.bnd, self.props, int(nsteps), dict(self.meta))


def _params(**kw) -> dict:
    base = dict(
        rk_method=1, flux_method=1, cfl=0.3, dtau=0.0, tt_tref=0.0, tt_time=0.0,
        steps=0, tol=0.0, periodic_x=0, periodic_y=0, mirror_north=0,
        mirror_south=0, mirror_east=0, mirror_west=0, gamma=1.4, R=1.0,
        mu=0.0, Pr=0.71, dx=0.0, dy=0.0, Lx=0.0, Ly=0.0, rho_ref=0.0,
        p_ref=0.0, T_ref=0.0, Tc_Tref=1.0,
    )
    for k, v in kw.items():
        if k not in base:
            raise KeyError(k)
        base[k] = v
    return base


def _props(rows) -> dict:
    """rows: iterable of (density, pressure, temp, ux, uy, cell_type)."""
    rows = list(rows)
    return {
        "density": np.array([r[0] for r in rows], dtype=np.float64),
        "pressure": np.array([r[1] for r in rows], dtype=np.float64),
        "temp": np.array([r[2] for r in rows], dtype=np.float64),
        "velocity_x": np.array([r[3] for r in rows], dtype=np.float64),
        "velocity_y": np.array([r[4] for r in rows], dtype=np.float64),
        "cell_type": np.array([r[5] for r in rows], dtype=np.int32),
    }


def couette(n: int = 64, u_wall: float = 0.1, T: float = 1.0, rho: float = 1.0,
            tau: float = 0.05, Pr: float = 0.71, gamma: float = 1.4, R: float = 1.0,
            cfl: float = 0.3, steps: int = 2000, flux_method: int = 1,
            rk_method: int = 1, length: float = 1.0, nx: int | None = None) -> Case:
    """Thermal Couette flow: periodic in x, fixed isothermal bottom wall
    (row 0), isothermal top wall (row n-1) moving at u_wall, fluid between
    at rest and in equilibrium. The wall-to-wall gap (between the wall
    faces) is `length`, so dx = dy = length/(n-2)."""
    nx = n if nx is None else nx
    ny = n
    p = rho * R * T
    bnd = np.zeros((nx, ny), dtype=np.int32)
    bnd[:, 0] = 1
    bnd[:, ny - 1] = 2
    props = _props([
        (rho, p, T, 0.0, 0.0, FLUID),
        (rho, p, T, 0.0, 0.0, SOLID),
        (rho, p, T, u_wall, 0.0, SOLID),
    ])
    d = length / (ny - 2)
    params = _params(rk_method=rk_method, flux_method=flux_method, cfl=cfl,
                     steps=steps, periodic_x=1, gamma=gamma, R=R, mu=tau * p,
                     Pr=Pr, dx=d, dy=d, rho_ref=rho, p_ref=p, T_ref=T, Tc_Tref=1.0)
    meta = dict(kind="couette", u_wall=u_wall, T=T, rho=rho, tau=tau, Pr=Pr,
                gamma=gamma, R=R, length=length)
    return Case(f"couette{nx}x{ny}", params, bnd, props, steps, meta)


def shear_wave(n: int = 64, u0: float = 0.1, T: float = 1.0, rho: float = 1.0,
               tau: float = 0.05, Pr: float = 0.71, gamma: float = 1.4, R: float = 1.0,
               cfl: float = 0.3, steps: int = 2000, flux_method: int = 1,
               rk_method: int = 1, length: float = 1.0) -> Case:
    """Doubly periodic decaying shear wave u_x = u0*sin(2*pi*(y+1/2)/n):
    the builder's periodic shear-layer instance (one property row per
    grid row). dx = dy = length/n."""
    nx = ny = n
    p = rho * R * T
    ys = (np.arange(ny) + 0.5) / ny
    ux = u0 * np.sin(2.0 * math.pi * ys)
    rows = [(rho, p, T, float(ux[j]), 0.0, FLUID) for j in range(ny)]
    bnd = np.tile(np.arange(ny, dtype=np.int32), (nx, 1))
    d = length / n
    params = _params(rk_method=rk_method, flux_method=flux_method, cfl=cfl,
                     steps=steps, periodic_x=1, periodic_y=1, gamma=gamma, R=R,
                     mu=tau * p, Pr=Pr, dx=d, dy=d, rho_ref=rho, p_ref=p,
                     T_ref=T, Tc_Tref=1.0)
    meta = dict(kind="shear", u_wall=u0, T=T, rho=rho, tau=tau, Pr=Pr, gamma=gamma,
                R=R, length=length)
    return Case(f"shear{n}", params, bnd, _props(rows), steps, meta)


def obstacle(n: int = 48, block=(18, 20, 8, 6), u0: float = 0.1, T: float = 1.0, rho: float = 1.0,
             tau: float = 0.05, Pr: float = 0.71, gamma: float = 1.4, R: float = 1.0,
             cfl: float = 0.3, steps: int = 200, flux_method: int = 1, rk_method: int = 1,
             length: float = 1.0, mirror_south: bool = False) -> Case:
    """Uniform flow u0 along x in an n x n box, periodic in x and in y (with
    `mirror_south`: mirrored at the south edge, clamped at the north edge),
    around a solid block (isothermal, at rest) of w x h cells with its
    lower-left cell at (x0, y0), block = (x0, y0, w, h).
    The block's corner cells are wall cells with fluid neighbours on both
    axes; h = 1 or w = 1 gives a one-cell-thick plate, and y0 = 0 with
    `mirror_south` a solid on a mirror edge. dx = dy = length/n."""
    x0, y0, w, h = block
    nx = ny = n
    p = rho * R * T
    bnd = np.zeros((nx, ny), dtype=np.int32)
    bnd[x0:x0 + w, y0:y0 + h] = 1
    props = _props([
        (rho, p, T, u0, 0.0, FLUID),
        (rho, p, T, 0.0, 0.0, SOLID),
    ])
    d = length / n
    params = _params(rk_method=rk_method, flux_method=flux_method, cfl=cfl,
                     steps=steps, periodic_x=1, periodic_y=0 if mirror_south else 1,
                     mirror_south=1 if mirror_south else 0, gamma=gamma, R=R, mu=tau * p,
                     Pr=Pr, dx=d, dy=d, rho_ref=rho, p_ref=p, T_ref=T, Tc_Tref=1.0)
    meta = dict(kind="obstacle", u_wall=u0, T=T, rho=rho, tau=tau, Pr=Pr, gamma=gamma, R=R,
                length=length, block=list(block))
    return Case(f"obstacle{n}_{x0}_{y0}_{w}x{h}", params, bnd, props, steps, meta)


def flat_plate(nx: int = 512, ny: int = 256, mach: float = 2.0, re_x: float = 1e5,
               plate_start: int = 64, steps: int = 20000, flux_method: int = 0,
               rk_method: int = 1, cfl: float = 0.3, wall: str = "recovery") -> Case:
    """Compressible flat-plate boundary layer after the reference's Jacobs
    deck (Items/11.04.06-BoundaryLayer-Jacobs/input.mat; makeBoundaryLayer.py):
    SI air (R = 287, gamma = 1.4, Pr = 0.72), free stream T = 222 K,
    rho = 0.0404 kg/m^3, u = mach*sqrt(gamma*R*T), dx = dy = 1 mm,
    Tc_Tref = 2.5, T_ref = 293 K. Column x = 0 and row y = ny-1 are
    permanent free-stream cells (cellType 1); row y = 0 is a no-slip wall
    (cellType 2) from x = plate_start on and a symmetry plane before it
    (mirrorSouth); the east edge uses zeroth-order extrapolation.
    mu is set so that Re_x = re_x at the plate end.

    The reference's only wall closure emits the wall cell's frozen
    equilibrium, i.e. an isothermal wall (cpuLBM.cpp:1502-1590). wall =
    "recovery" holds it at the adiabatic recovery temperature
    T*(1 + sqrt(Pr)*(gamma-1)/2*M^2); wall = "freestream" at T.
    """
    R, gamma, Pr = 287.0, 1.4, 0.72
    T0, rho0 = 222.0, 0.0404
    p0 = rho0 * R * T0
    u0 = mach * math.sqrt(gamma * R * T0)
    dx = 1e-3
    plate_len = (nx - plate_start) * dx
    mu = rho0 * u0 * plate_len / re_x
    if wall == "recovery":
        Tw = T0 * (1.0 + math.sqrt(Pr) * (gamma - 1.0) / 2.0 * mach * mach)
    elif wall == "freestream":
        Tw = T0
    else:
        raise ValueError(wall)
    props = _props([
        (rho0, p0, T0, u0, 0.0, PERMANENT),  # 0: inflow column and top row
        (rho0, p0, T0, u0, 0.0, FLUID),      # 1: fluid
        (p0 / (R * Tw), p0, Tw, 0.0, 0.0, SOLID),  # 2: plate
    ])
    bnd = np.ones((nx, ny), dtype=np.int32)
    bnd[0, :] = 0
    bnd[:, ny - 1] = 0
    bnd[plate_start:, 0] = 2
    params = _params(rk_method=rk_method, flux_method=flux_method, cfl=cfl,
                     steps=steps, mirror_south=1, gamma=gamma, R=R, mu=mu, Pr=Pr,
                     dx=dx, dy=dx, rho_ref=rho0, p_ref=0.0, T_ref=293.0, Tc_Tref=2.5)
    meta = dict(kind="flat_plate", mach=mach, re_x=re_x, u_inf=u0, T_inf=T0,
                rho_inf=rho0, T_wall=Tw, plate_start=plate_start)
    return Case(f"plate{nx}x{ny}", params, bnd, props, steps, meta)


# Lax-Liu quadrant states, (rho, u, v, p) in units of rho_ref, u_ref, p_ref:
# column 0 of the tables in makeLaxLiu3.py:115-118, quadrants numbered as
# in makeLaxLiu3.py:79-89 (0 top-right, 1 top-left, 2 bottom-left,
# 3 bottom-right).
RIEMANN_STATES = (
    (1.5, 0.0, 0.0, 1.5),
    (0.5323, 1.206, 0.0, 0.3),
    (0.138, 1.206, 1.206, 0.029),
    (0.5323, 0.0, 1.206, 0.3),
)


def riemann(n: int = 4096, steps: int = 2000, flux_method: int = 1,
            rk_method: int = 1, cfl: float = 0.2) -> Case:
    """Four-quadrant Riemann problem on an n x n grid with the SI set-up of
    the reference's Lax-Liu generator (makeLaxLiu3.py:15-35, 120-123):
    R = 287, gamma = 1.4, mu = 1.86e-5, Pr = 0.71, Tc_Tref = 2.5, unit
    square, rho_ref = 1.165, p_ref = 101310, u_ref = sqrt(R*T_ref), all
    edges zeroth-order extrapolation."""
    R, gamma, mu, Pr = 287.0, 1.4, 1.86e-5, 0.71
    rho_ref, p_ref = 1.165, 101310.0
    T_ref = p_ref / (rho_ref * R)
    u_ref = math.sqrt(R * T_ref)
    rows = []
    for (r, u, v, pr) in RIEMANN_STATES:
        rho = r * rho_ref
        p = pr * p_ref
        rows.append((rho, p, p / (rho * R), u * u_ref, v * u_ref, FLUID))
    mid = n // 2
    bnd = np.zeros((n, n), dtype=np.int32)
    bnd[mid:, mid:] = 0
    bnd[:mid, mid:] = 1
    bnd[:mid, :mid] = 2
    bnd[mid:, :mid] = 3
    d = 1.0 / n
    params = _params(rk_method=rk_method, flux_method=flux_method, cfl=cfl,
                     steps=steps, gamma=gamma, R=R, mu=mu, Pr=Pr, dx=d, dy=d,
                     rho_ref=rho_ref, p_ref=p_ref, T_ref=T_ref, Tc_Tref=2.5)
    meta = dict(kind="riemann", n=n)
    return Case(f"riemann{n}", params, bnd, _props(rows), steps, meta)


def uq_sample(rng: np.random.Generator) -> dict:
    """One draw of the uncertain physical parameters: wall speed
    U(0.05, 0.2), Mach U(0.1, 0.5), Pr U(0.6, 1.0), ln tau ~ N(ln 0.05, 0.2)."""
    return dict(
        u_wall=float(rng.uniform(0.05, 0.2)),
        mach=float(rng.uniform(0.1, 0.5)),
        Pr=float(rng.uniform(0.6, 1.0)),
        tau=float(math.exp(rng.normal(math.log(0.05), 0.2))),
    )


def uq_instance(kind: str, n: int, s: dict, steps: int, gamma: float = 1.4,
                R: float = 1.0, cfl: float = 0.3, tc_tref: float = 2.5) -> Case:
    """Map a parameter draw to a deck: T from the wall speed and Mach number,
    T = u_wall^2/(gamma*R*Ma^2), and the lattice temperature Tc = tc_tref*T.

    tc_tref = 2.5 is the ratio of the reference's own deck
    (Items/11.04.06-BoundaryLayer-Jacobs/input.mat). With Tc = T (ratio 1)
    the scheme itself is unstable for the faster walls: 25 of the 256
    ensemble Couette instances (Mach > 0.31) blow up within 5000 steps, in
    the CPU oracle and on the GPU at the same step; at 1.5 or 2.5 none do."""
    T = s["u_wall"] ** 2 / (gamma * R * s["mach"] ** 2)
    if kind == "couette":
        c = couette(n=n, u_wall=s["u_wall"], T=T, tau=s["tau"], Pr=s["Pr"],
                    gamma=gamma, R=R, cfl=cfl, steps=steps)
    elif kind == "shear":
        c = shear_wave(n=n, u0=s["u_wall"], T=T, tau=s["tau"], Pr=s["Pr"],
                       gamma=gamma, R=R, cfl=cfl, steps=steps)
    else:
        raise ValueError(kind)
    c.params["Tc_Tref"] = float(tc_tref)
    c.meta.update(s)
    c.meta["mach"] = s["mach"]
    c.meta["tc_tref"] = float(tc_tref)
    return c


def uq_ensemble(n_instances: int = 256, seed: int = 2026, sides=(64, 128, 256, 512),
                steps: int = 5000, kinds=("couette", "shear"), tc_tref: float = 2.5) -> list[Case]:
    """The UQ ensemble: square sides drawn uniformly from `sides`, the kind
    (Couette or periodic shear wave) drawn uniformly, parameters from
    uq_sample. Seeded."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_instances):
        n = int(rng.choice(sides))
        kind = str(rng.choice(kinds))
        s = uq_sample(rng)
        c = uq_instance(kind, n, s, steps, tc_tref=tc_tref)
        c.meta.update(dict(index=k, seed=seed))
        out.append(c)
    return out


def uq_member(index: int, n_instances: int = 256, seed: int = 2026, steps: int = 5000,
              tc_tref: float = 2.5) -> Case:
    """Member `index` of uq_ensemble(n_instances, seed, steps, tc_tref=...)."""
    return uq_ensemble(n_instances, seed=seed, steps=steps, tc_tref=tc_tref)[index]


@dataclass
class Request:
    tag: int
    arrival_s: float
    case: Case


def request_stream(rate: float = 50.0, duration: float = 60.0, seed: int = 7,
                   side_min: int = 64, side_max: int = 1024, steps_min: int = 500,
                   steps_max: int = 5000, kinds=("couette", "shear"),
                   tc_tref: float = 2.5) -> list[Request]:
    """Poisson arrivals at `rate` per second over `duration` seconds; each
    request one solve with side log-uniform in [side_min, side_max], steps
    uniform in [steps_min, steps_max], parameters as in the UQ ensemble."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    tag = 0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(side_min), math.log(side_max)))))
        steps = int(rng.integers(steps_min, steps_max + 1))
        kind = str(rng.choice(kinds))
        s = uq_sample(rng)
        c = uq_instance(kind, n, s, steps, tc_tref=tc_tref)
        out.append(Request(tag, t, c))
        tag += 1
    return out


def with_spacing(case: Case, dx, dy) -> Case:
    """The case on a non-uniform grid (the OpenCL front end's per-column dx
    and per-row dy, PY/codeCL.py:611-618, 676-679): applied by Solver and the
    oracle at construction. Decks with solid cells are refused there."""
    dx = [float(v) for v in dx]
    dy = [float(v) for v in dy]
    if len(dx) != case.nx or len(dy) != case.ny:
        raise ValueError("spacing arrays must have nx and ny entries")
    meta = dict(case.meta)
    meta["spacing"] = {"dx": dx, "dy": dy}
    return Case(case.name + "_spaced", dict(case.params), case.bnd, case.props, case.nsteps, meta)


def stretched(case: Case, amp: float = 0.3) -> Case:
    """A smoothly stretched grid for `case`: dx[x] = dx (1 + amp sin(2 pi
    x / nx)), dy[y] = dy (1 + amp cos(2 pi y / ny)) (periodic in both
    directions, so a doubly periodic deck stays periodic)."""
    x = np.arange(case.nx)
    y = np.arange(case.ny)
    dx = case.params["dx"] * (1 + amp * np.sin(2 * np.pi * x / case.nx))
    dy = case.params["dy"] * (1 + amp * np.cos(2 * np.pi * y / case.ny))
    return with_spacing(case, dx, dy)


def couette_analytic(case: Case, y_cells: np.ndarray, c_p: float | None = None):
    """Steady thermal Couette profile with first-order Knudsen slip and
    temperature jump, after the reference's own check
    (Items/11.04.18 - Couette flow check for macrossan/couetteConstT.m:1-30),
    written for a fixed bottom wall and a top wall moving at u_wall.

    The wall faces sit half a cell from the wall-cell centres (rows 0 and
    n-1), so the gap is L = (n-2)*dy and eta = (y - 1/2)/(n-2) in [0, 1].
    Returns (u, T) at the cell centres y_cells (row indices).
    """
    m = case.meta
    R, gamma, Pr = m["R"], m["gamma"], m["Pr"]
    T0, rho, tau, U = m["T"], m["rho"], m["tau"], m["u_wall"]
    n = case.ny
    L = m["length"]
    mu = tau * rho * R * T0
    cp = gamma * R / (gamma - 1.0) if c_p is None else c_p
    k = mu * cp / Pr
    lam = 2.0 * mu / (rho * math.sqrt(8.0 * R * T0 / math.pi))
    Kn = lam / L
    # couetteConstT.m works on y in [-1/2, 1/2] with walls at -U (right) and
    # +U (left); here the bottom wall is at rest and the top moves at U.
    A = U / (1.0 + 2.0 * Kn)           # velocity slope in eta
    eta = (np.asarray(y_cells, dtype=np.float64) - 0.5) / (n - 2)
    u = U / 2.0 + A * (eta - 0.5)
    hK = (2.0 * gamma / (gamma + 1.0)) / Pr
    # temperature: T = T0 + C*s^2 + E with s = eta - 1/2, C = -mu/(2k)*A^2
    C = -(mu / (2.0 * k)) * A * A
    E = (mu / (8.0 * k)) * A * A * (1.0 + 4.0 * hK * Kn)
    s = eta - 0.5
    T = T0 + C * s * s + E
    return u, T, dict(Kn=Kn, A=A, C=C, E=E)


def config_cases() -> dict:
    """The single-instance workloads by name (full sizes). The Riemann deck
    runs its 2000 steps at CFL 0.1: the same end time as the reference
    generator's own 1000 steps at CFL 0.2 (makeLaxLiu3.py:22, 26); at CFL 0.2
    the scheme itself trips its equilibrium check near step 1050 on every
    grid of 128^2 and finer, the oracle included."""
    return {
        "couette64": couette(64),
        "plate512x256": flat_plate(),
        "riemann4096": riemann(4096, steps=2000, cfl=0.1),
    }
