"""Fault problems: friction fields, loading and initial state.

The initial state restates the reference's steady-state formulas
(src/QDsim.py:623 and :642, :652-654):
    Tt    = Tno a asinh(V/(2 V0) exp((f0 + b ln(V0/Vpl))/a))
    state = a ln(2 V0/V sinh(Tt/(Tno a)))
    Tt1o, Tt2o = Tt cos(rake), Tt sin(rake);  V1, V2 = V cos, V sin (+1e-16)
Tractions are in MPa as in the reference; rates in m/s; lengths in m.

Workloads (synthetic, seeded):
  planar_c1       40 x 20 km vertical strike-slip fault, ~2,000 triangles,
                  a = 0.004 in a 20 x 10 km VW patch tapering over 2 km to 0.04
  undulating_c3   100 x 25 km undulating fault, BP5-like heterogeneous a, b,
                  Dc from seeded Gaussian random fields (5 km correlation)
Both add a seeded nucleation patch (V = 100 Vpl in a disc inside the VW
region): from an exact steady state the first instability would otherwise
grow out of rounding noise.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from .mesh import FaultMesh, get_eleVec, planar_fault, undulating_fault


@dataclass
class Problem:
    mesh: FaultMesh
    mu: float
    lambda_: float
    density: float
    V0: float
    a: np.ndarray
    b: np.ndarray
    dc: np.ndarray
    f0: np.ndarray
    Tno: np.ndarray            # MPa
    slipv: np.ndarray          # m/s
    rake: np.ndarray           # rad
    Vpl: np.ndarray            # m/s, per cell (the reference's slipvC)
    shear_loading: np.ndarray = None
    normal_loading: np.ndarray = None
    fix_Tn: bool = False
    name: str = ""
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        n = self.mesh.n
        if self.shear_loading is None:
            self.shear_loading = np.zeros(n)
        if self.normal_loading is None:
            self.normal_loading = np.zeros(n)

    @property
    def n(self) -> int:
        return self.mesh.n

    @property
    def Cs(self) -> float:
        return float(np.sqrt(self.mu / self.density))

    def initial_state(self):
        """(Tno, Tt1o, Tt2o, state, slipv1, slipv2, slipv) in MPa / m/s."""
        a, b, f0, V0 = self.a, self.b, self.f0, self.V0
        V, Tno = self.slipv, self.Tno
        Tt = self.meta.get("Tt")   # shear traction given by a stress tensor, else steady state
        if Tt is None:
            Tt = Tno * a * np.arcsinh(V / (2.0 * V0) * np.exp((f0 + b * np.log(V0 / self.Vpl)) / a))
        state = np.log(np.sinh(Tt / Tno / a) * 2.0 * V0 / V) * a
        Tt1o = Tt * np.cos(self.rake)
        Tt2o = Tt * np.sin(self.rake)
        slipv1 = V * np.cos(self.rake) + 1e-16
        slipv2 = V * np.sin(self.rake) + 1e-16
        return Tno.copy(), Tt1o, Tt2o, state, slipv1, slipv2, V.copy()


def _taper(d, width):
    return np.clip(d / width, 0.0, 1.0)


def planar_c1(nx=45, nz=22, seed=0, nucleation=True, mu=32e9, nu=0.25, Vs=3464.0,
              sigma=25.0, Vpl=1e-9) -> Problem:
    mesh = planar_fault(length=40e3, depth=20e3, nx=nx, nz=nz, seed=seed)
    lam = 2 * mu * nu / (1 - 2 * nu)
    _, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    n = mesh.n
    # velocity-weakening patch 20 x 10 km centred at (0, -10 km), 2 km taper
    d = np.maximum(np.abs(xg[:, 0]) - 10e3, np.abs(xg[:, 2] + 10e3) - 5e3)
    a = 0.004 + (0.04 - 0.004) * _taper(d, 2e3)
    b = np.full(n, 0.03)
    dc = np.full(n, 0.14)
    f0 = np.full(n, 0.6)
    V = np.full(n, Vpl)
    if nucleation:
        rng = np.random.default_rng(seed + 101)
        cx, cz = rng.uniform(-5e3, 5e3), rng.uniform(-12e3, -8e3)
        V = np.where((xg[:, 0] - cx) ** 2 + (xg[:, 2] - cz) ** 2 < (2.5e3) ** 2, 100 * Vpl, V)
    return Problem(mesh, mu, lam, mu / Vs ** 2, 1e-6, a, b, dc, f0, np.full(n, sigma), V,
                   np.zeros(n), np.full(n, Vpl), name=f"planar_c1_N{n}")


def gaussian_field(xg, corr, seed, grid=256):
    """Zero-mean, unit-variance Gaussian random field with a Gaussian
    correlation of length ``corr``, FFT-filtered white noise on a regular
    (x, z) grid, sampled at the element centroids (nearest grid node)."""
    rng = np.random.default_rng(seed)
    x0, x1 = xg[:, 0].min(), xg[:, 0].max()
    z0, z1 = xg[:, 2].min(), xg[:, 2].max()
    L = max(x1 - x0, z1 - z0) * 1.25 + 4 * corr
    dx = L / grid
    w = rng.standard_normal((grid, grid))
    k = np.fft.fftfreq(grid, d=dx) * 2 * np.pi
    kx, kz = np.meshgrid(k, k, indexing="ij")
    filt = np.exp(-0.25 * (kx ** 2 + kz ** 2) * corr ** 2)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    f = (f - f.mean()) / f.std()
    ix = np.clip(((xg[:, 0] - x0) / dx).astype(int), 0, grid - 1)
    iz = np.clip(((xg[:, 2] - z0) / dx).astype(int), 0, grid - 1)
    return f[ix, iz]


def undulating_c3(nx=448, nz=112, seed=0, nucleation=True, mu=32e9, nu=0.25, Vs=3464.0,
                  sigma=25.0, Vpl=1e-9, friction_seed=None, mesh=None) -> Problem:
    """BP5-like heterogeneous friction on the undulating 100 x 25 km fault.
    a, b, Dc: BP5 values (a_VW 0.004, a_VS 0.04, b 0.03, Dc 0.14 m) with a
    VW region x in [-30, 30] km, depth 3..18 km (2 km tapers), perturbed by
    seeded Gaussian fields (5 km correlation): a * (1 + 0.1 g1),
    b * (1 + 0.05 g2), Dc * exp(0.2 g3). The mesh follows ``seed``; the
    friction draw and nucleation patch follow ``friction_seed`` (default
    ``seed``), so that draws can share one mesh (pass it as ``mesh``)."""
    if mesh is None:
        mesh = undulating_fault(length=100e3, depth=25e3, nx=nx, nz=nz, amplitude=300.0,
                                wavelength=5e3, seed=seed)
    if friction_seed is not None:
        seed = friction_seed
    lam = 2 * mu * nu / (1 - 2 * nu)
    _, xg = get_eleVec(mesh.nodelst, mesh.elelst)
    n = mesh.n
    depth = -xg[:, 2]
    d = np.maximum.reduce([np.abs(xg[:, 0]) - 30e3, 3e3 - depth, depth - 18e3])
    a = 0.004 + (0.04 - 0.004) * _taper(d, 2e3)
    g1 = gaussian_field(xg, 5e3, seed + 11)
    g2 = gaussian_field(xg, 5e3, seed + 12)
    g3 = gaussian_field(xg, 5e3, seed + 13)
    a = a * (1 + 0.1 * g1)
    b = 0.03 * (1 + 0.05 * g2)
    dc = 0.14 * np.exp(0.2 * g3)
    f0 = np.full(n, 0.6)
    V = np.full(n, Vpl)
    if nucleation:
        rng = np.random.default_rng(seed + 101)
        cx, cz = rng.uniform(-20e3, 20e3), rng.uniform(-14e3, -8e3)
        V = np.where((xg[:, 0] - cx) ** 2 + (xg[:, 2] - cz) ** 2 < (3e3) ** 2, 100 * Vpl, V)
    return Problem(mesh, mu, lam, mu / Vs ** 2, 1e-6, a, b, dc, f0, np.full(n, sigma), V,
                   np.zeros(n), np.full(n, Vpl), name=f"undulating_c3_N{n}")


_ARRAYS = ("a", "b", "dc", "f0", "Tno", "slipv", "rake", "Vpl", "shear_loading",
           "normal_loading")


def save_problem(path, p: Problem, halfspace: bool = False) -> None:
    """A model as one compressed .npz (mesh, per-element fields, scalars):
    what ``load_problem`` reads back, for runs where the original model files
    are not at hand."""
    extra = {} if p.meta.get("Tt") is None else {"Tt": np.asarray(p.meta["Tt"])}
    np.savez_compressed(
        path, nodelst=p.mesh.nodelst, elelst=p.mesh.elelst.astype(np.int32),
        scalars=np.array([p.mu, p.lambda_, p.density, p.V0, float(halfspace), float(p.fix_Tn)]),
        name=np.array(p.name), **{k: getattr(p, k) for k in _ARRAYS}, **extra)


def load_problem(path):
    """(Problem, halfspace) from a ``save_problem`` archive."""
    with np.load(path, allow_pickle=False) as z:
        mesh = FaultMesh(np.array(z["nodelst"]), np.array(z["elelst"]).astype(np.int64),
                         name=str(z["name"]))
        mu, lam, rho, V0, hs, fix = (float(v) for v in z["scalars"])
        arr = {k: np.array(z[k]) for k in _ARRAYS}
        Tt = np.array(z["Tt"]) if "Tt" in z.files else None
    p = Problem(mesh, mu, lam, rho, V0, arr["a"], arr["b"], arr["dc"], arr["f0"], arr["Tno"],
                arr["slipv"], arr["rake"], arr["Vpl"], arr["shear_loading"],
                arr["normal_loading"], fix_Tn=bool(fix), name=mesh.name, meta={"Tt": Tt})
    return p, bool(hs)


def with_friction(p: Problem, **kw) -> Problem:
    return replace(p, **kw)


def ensemble_c4(count=64, seed=0, n_min=2000, n_max=20000, mu=32e9, nu=0.25, Vs=3464.0,
                sigma=25.0, Vpl=1e-9) -> list:
    """Independent simulations on meshes of uneven size: N log-uniform in
    [n_min, n_max], a planar or an undulating 40 x 20 km fault by coin flip
    (2:1 rectangles split in two), and a friction draw each: a uniform in
    [0.004, 0.02] inside a 20 x 10 km velocity-weakening patch (2 km taper to
    0.04 outside), Dc uniform in [0.05, 0.3] m, b = 0.03, with a seeded
    nucleation disc at 100 Vpl."""
    rng = np.random.default_rng(seed)
    lam = 2 * mu * nu / (1 - 2 * nu)
    out = []
    for k in range(count):
        n_target = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        nz = max(2, int(round(np.sqrt(n_target / 4.0))))
        nx = 2 * nz
        if rng.random() < 0.5:
            mesh = planar_fault(length=40e3, depth=20e3, nx=nx, nz=nz, seed=seed + k)
        else:
            mesh = undulating_fault(length=40e3, depth=20e3, nx=nx, nz=nz, amplitude=300.0,
                                    wavelength=5e3, seed=seed + k)
        _, xg = get_eleVec(mesh.nodelst, mesh.elelst)
        n = mesh.n
        d = np.maximum(np.abs(xg[:, 0]) - 10e3, np.abs(xg[:, 2] + 10e3) - 5e3)
        a_vw = rng.uniform(0.004, 0.02)
        a = a_vw + (0.04 - a_vw) * _taper(d, 2e3)
        dc = np.full(n, rng.uniform(0.05, 0.3))
        cx, cz = rng.uniform(-5e3, 5e3), rng.uniform(-12e3, -8e3)
        V = np.where((xg[:, 0] - cx) ** 2 + (xg[:, 2] - cz) ** 2 < (2.5e3) ** 2, 100 * Vpl, Vpl)
        out.append(Problem(mesh, mu, lam, mu / Vs ** 2, 1e-6, a, np.full(n, 0.03), dc,
                           np.full(n, 0.6), np.full(n, sigma), V, np.zeros(n),
                           np.full(n, Vpl), name=f"c4_{k}_N{n}"))
    return out
