"""Synthetic spherical-shell workloads of the Hall-MHD step (physics, mesh sizes, seeds).

shell_60k      FP64 ideal two-ion MHD, Hall off, one explicit stage; 24x48 net, 9 layers (59,616 tets)
shell_2m_hall  Hall + uniform resistivity, GMRES(30) to 1e-8, all FP64; 128x256 net, 10 layers (1.95M)
shell_8m_mixed Hall + resistivity, mixed-precision GMRES (Jacobi; the agglomeration-multigrid V-cycle with
               precision 2); 256x512 net, 10 layers (7.83M)
shell_4m_mars  ten-ion Mars-like fluid, crustal dipoles, exponential-atmosphere ion production, Hall on;
               shell r in [1, 3], 192x384 net, 9 layers (3.98M)

Hall runs hold the EMF of the conducting wall at zero in the implicit solve as well (Dirichlet wall edges,
the MarsSolver convention): with free wall edges (the Mercury convention, whose wall is a solid-planet
interface) the semi-implicit Hall step grows without bound at the wall next to the poles of this mesh.
Common physics: rho_H = 1, rho_Na = 1e-3, v = (-1, 0, 0), p_H = 0.1 with Na+ at the same temperature
(p_Na = 1e-4); centred dipole of equatorial surface strength 0.2 (m = -0.2 z); IMF (0, 0, -0.05); wall at
r = 1, far field at r = 4 (inflow on the day side); gamma = 5/3, inver_MA2 = 1, M_Na/M_H = 23; CFL 0.4,
second order with relative limiter 0.25, cuct 0.5.  Hall runs: hall_coef 0.5 (taper off), inver_Rem 1e-3
everywhere.  Initial state: 1% seeded Gaussian noise, induced flux d1 of a seeded edge potential.
Seeds: 20261202 + workload index.

shell_4m_mars: ion masses 1, 4, 16, 32, 44, 2, 14, 28, 30, 12 (M_ref 1); upstream v = (-1, 0, 0), density 1 for
the first ion and 1e-4 for the others, p = 0.1 rho; no planetary dipole: the IMF plus 100 crustal point dipoles
at r = 0.98 (seeded uniform positions and orientations on the sphere, moments lognormal(mu=-3, sigma=1) x
0.02^3, i.e. surface fields of order exp(-3) above each source); ion production q_s exp(-(r - 1) / 0.02) with
seeded rates q_s = 10^U(-3, -1) for the nine heavy ions (none for the first), created at temperature 1e-3;
Hall as above, 5 steps.
"""
import numpy as np

from .mesh import Mesh, default_config

WORKLOADS = {
    "shell_60k": dict(index=1, net=(24, 48, 9), hall=False),
    "shell_2m_hall": dict(index=2, net=(128, 256, 10), hall=True),
    "shell_8m_mixed": dict(index=3, net=(256, 512, 10), hall=True),
    "shell_4m_mars": dict(index=5, net=(192, 384, 9), hall=True, mars=True),
}
MARS_ION_MASS = (1.0, 4.0, 16.0, 32.0, 44.0, 2.0, 14.0, 28.0, 30.0, 12.0)

DIPOLE_M = (0.0, 0.0, -0.2)
B_IMF = (0.0, 0.0, -0.05)
U_INF = (-1.0, 0.0, 0.0)
SEED0 = 20261202


def physics_config(hall, **kw):
    base = dict(gamma=5.0 / 3.0, inver_MA2=1.0, M_ref=1.0, M_H=1.0, M_Na=23.0, ne_floor=0.01,
                cfl=0.4, second_order=1, recon_rel_limit=0.25, lorentz_force=1, cuct=0.5,
                fs=((1.0, -1.0, 0.0, 0.0, 0.1), (1e-3, -1.0, 0.0, 0.0, 1e-4)), u_inf=U_INF, B_imf=B_IMF,
                ind_rtol=1e-8, ind_max_it=2000, ind_restart=30)
    if hall:
        base.update(hall=1, hall_coef=0.5, hall_r_min=-100.0, hall_r_max=-99.0,
                    resistive=1, inver_Rem=1e-3, eta_r_inner=-100.0, eta_r_outer=100.0, eta_width=1.0,
                    wall_dirichlet=1)
    base.update(kw)
    return default_config(**base)


def build(name=None, net=None, hall=None, seed=None, noise=0.01, order=None, **cfg_kw):
    """Mesh, config, initial state and background of a workload (or of a custom net size).

    order="hilbert": the same workload on the mesh renumbered in the Hilbert order of the cell centres (cells in key
    order, faces and edges by first appearance); the state and background are drawn on the generator's numbering and
    permuted, so every entity carries the same values.  w["maps"] = (cell_map, face_map, edge_map), new id -> old."""
    if order not in (None, "hilbert"):
        raise ValueError("order must be None or 'hilbert'")
    if order == "hilbert":
        return renumber(build(name, net, hall, seed, noise, None, **cfg_kw))
    spec = dict(WORKLOADS[name]) if name else dict(index=0, net=net, hall=bool(hall))
    if net is not None:
        spec["net"] = net
    if hall is not None:
        spec["hall"] = hall
    seed = SEED0 + spec["index"] if seed is None else seed
    if spec.get("mars"):
        return _build_mars(spec, seed, noise, **cfg_kw)
    n_lat, n_lon, n_r = spec["net"]
    mesh = Mesh.shell(n_lat, n_lon, n_r, seed=seed).boundary({1: "wall", 2: "farfield"}, u_inf=U_INF)
    cfg = physics_config(spec["hall"], **cfg_kw)
    PhiAdd, Badd = mesh.background(B_IMF, DIPOLE_M, (0.0, 0.0, 0.0), 1e-3)
    UH, UNa, Phi = mesh.seeded_state(cfg, noise, seed, Badd)
    return dict(mesh=mesh, cfg=cfg, state=(UH, UNa, Phi), background=(PhiAdd, Badd), seed=seed, spec=spec)


def mars_config(hall=True, n_dipoles=100, seed=SEED0 + 5, **kw):
    """Ten-ion configuration of shell_4m_mars (seeded ion production rates)."""
    rng = np.random.default_rng(seed)
    rates = [0.0] + list(10.0 ** rng.uniform(-3.0, -1.0, size=9))
    fs = [(1.0 if s == 0 else 1e-4, U_INF[0], U_INF[1], U_INF[2], 0.1 * (1.0 if s == 0 else 1e-4)) for s in range(10)]
    base = dict(n_species=10, ion_mass=MARS_ION_MASS, ion_fs=fs, fs=fs[:2], M_H=MARS_ION_MASS[0], M_Na=MARS_ION_MASS[1],
                src_rate=rates, src_scale_height=0.02, src_r0=1.0, src_T=1e-3)
    base.update(kw)
    return physics_config(hall, **base)


# MARS::ParticleMass (MarsUns/physics/MarsSpecies.h:34-36) [kg]: Ar, CO2, CO, C, H, N2, N, NO, O, O2; the ions and the
# reactive neutrals of MarsSolver share the table
MARS_PARTICLE_MASS = (6.645e-26, 7.309e-26, 4.651e-26, 1.993e-26, 1.673e-27, 4.651e-26, 2.326e-26, 4.983e-26, 2.658e-26, 5.316e-26)


def mars_params(n_neutral=10, neutral_stride=12, **kw):
    """MarsParams of the point-implicit sources with the reference's defaults (MarsSolver.cpp:40-62: NA, R_uni, q_e,
    U_ref, n_ref, B_ref, L_ref, Molecular_mass 0.001, floors 1e-10); keyword overrides."""
    from ._lib import MarsParams

    NA, R = 6.02214076e23, 8.314462618
    p = MarsParams()
    p.kB, p.q = R / NA, 1.602176634e-19
    p.U, p.n, p.B, p.L = 3.27e4, 4.0e6, 3.0e-9, 3.396e6
    p.rho_ref = 0.001 / NA * p.n
    p.p_ref = p.rho_ref * p.U * p.U
    p.time_ref = p.L / p.U
    p.rho_floor = p.p_floor = 1e-10
    for s in range(10):
        p.ion_mass[s] = p.neutral_mass[s] = MARS_PARTICLE_MASS[s]
    p.n_neutral, p.neutral_stride = n_neutral, neutral_stride
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def neutral_atmosphere(mesh, seed=SEED0 + 5, stride=12, scale_height=0.02, r0=1.0):
    """Seeded exponential neutral atmosphere [cm^-3] per cell (stride values per cell): surface densities
    10^U(6, 11) per neutral, falling off as exp(-(|x| - r0) / scale_height)."""
    rng = np.random.default_rng(seed + 2000)
    n0 = 10.0 ** rng.uniform(6.0, 11.0, size=stride)
    r = np.linalg.norm(mesh.array("cell_center").reshape(-1, 3), axis=1)
    return np.ascontiguousarray((np.exp(-(r - r0) / scale_height)[:, None] * n0[None, :]).ravel())


def crustal_dipoles(n=100, seed=SEED0 + 5, r=0.98, depth=0.02):
    """Seeded point dipoles at radius r: positions and orientations uniform on the sphere, moments
    lognormal(-3, 1) * depth^3."""
    rng = np.random.default_rng(seed + 1000)
    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1)[:, None]
    pos = r * unit(n)
    mom = unit(n) * (rng.lognormal(-3.0, 1.0, size=n) * depth ** 3)[:, None]
    return pos, mom


def _build_mars(spec, seed, noise, **cfg_kw):
    n_lat, n_lon, n_r = spec["net"]
    mesh = Mesh.shell(n_lat, n_lon, n_r, r_in=1.0, r_out=3.0, seed=seed).boundary({1: "wall", 2: "farfield"}, u_inf=U_INF)
    cfg = mars_config(spec["hall"], seed=seed, **cfg_kw)
    pos, mom = crustal_dipoles(seed=seed)
    PhiAdd, Badd = mesh.background_dipoles(B_IMF, pos, mom)
    fs = [tuple(cfg.ion_fs[s][m] for m in range(5)) for s in range(cfg.n_species)]
    states, Phi = mesh.seeded_species_state(cfg, fs, noise, seed, Badd)
    return dict(mesh=mesh, cfg=cfg, state=(states[0], states[1], Phi), species_state=states, background=(PhiAdd, Badd),
                dipoles=(pos, mom), seed=seed, spec=spec)


def renumber(w):
    """Workload w on its mesh renumbered in Hilbert order (Mesh.hilbert_order, Mesh.renumbered): same boundary rule,
    state and background permuted entity by entity."""
    m = w["mesh"]
    m2, cm, fm, em = m.renumbered(m.hilbert_order())
    m2.boundary({1: "wall", 2: "farfield"}, u_inf=U_INF)
    cells5 = lambda U: np.ascontiguousarray(U.reshape(-1, 5)[cm].ravel())
    UH, UNa, Phi = w["state"]
    PhiAdd, Badd = w["background"]
    out = dict(w, mesh=m2, state=(cells5(UH), cells5(UNa), np.ascontiguousarray(Phi[fm])),
               background=(np.ascontiguousarray(PhiAdd[fm]), np.ascontiguousarray(Badd.reshape(-1, 3)[cm].ravel())),
               maps=(cm, fm, em), order="hilbert")
    if "species_state" in w:
        out["species_state"] = [cells5(U) for U in w["species_state"]]
    return out


def load(dev, w):
    """Background and initial state of workload `w` onto Device `dev` (every species of a ten-ion fluid)."""
    dev.set_background(*w["background"])
    dev.set_state(*w["state"])
    for s, U in enumerate(w.get("species_state", [])[2:], start=2):
        dev.set_species_state(s, U)
