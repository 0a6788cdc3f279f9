"""Synthetic spherical-shell workloads of the Hall-MHD step (physics, mesh sizes, seeds).

shell_60k      FP64 ideal two-ion MHD, Hall off, one explicit stage; 24x48 net, 9 layers (59,616 tets)
shell_2m_hall  Hall + uniform resistivity, GMRES(30) to 1e-8, all FP64; 128x256 net, 10 layers (1.95M)
shell_8m_mixed Hall + resistivity, mixed-precision GMRES + multigrid; 256x512 net, 10 layers (7.83M)

Hall runs hold the EMF of the conducting wall at zero in the implicit solve as well (Dirichlet wall edges,
the MarsSolver convention): with free wall edges (the Mercury convention, whose wall is a solid-planet
interface) the semi-implicit Hall step grows without bound at the wall next to the poles of this mesh.
Common physics: rho_H = 1, rho_Na = 1e-3, v = (-1, 0, 0), p_H = 0.1 with Na+ at the same temperature
(p_Na = 1e-4); centred dipole of equatorial surface strength 0.2 (m = -0.2 z); IMF (0, 0, -0.05); wall at
r = 1, far field at r = 4 (inflow on the day side); gamma = 5/3, inver_MA2 = 1, M_Na/M_H = 23; CFL 0.4,
second order with relative limiter 0.25, cuct 0.5.  Hall runs: hall_coef 0.5 (taper off), inver_Rem 1e-3
everywhere.  Initial state: 1% seeded Gaussian noise, induced flux d1 of a seeded edge potential.
Seeds: 20261202 + workload index.
"""
from .mesh import Mesh, default_config

WORKLOADS = {
    "shell_60k": dict(index=1, net=(24, 48, 9), hall=False),
    "shell_2m_hall": dict(index=2, net=(128, 256, 10), hall=True),
    "shell_8m_mixed": dict(index=3, net=(256, 512, 10), hall=True),
}

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


def build(name=None, net=None, hall=None, seed=None, noise=0.01, **cfg_kw):
    """Mesh, config, initial state and background of a workload (or of a custom net size)."""
    spec = dict(WORKLOADS[name]) if name else dict(index=0, net=net, hall=bool(hall))
    if net is not None:
        spec["net"] = net
    if hall is not None:
        spec["hall"] = hall
    seed = SEED0 + spec["index"] if seed is None else seed
    n_lat, n_lon, n_r = spec["net"]
    mesh = Mesh.shell(n_lat, n_lon, n_r, seed=seed).boundary({1: "wall", 2: "farfield"}, u_inf=U_INF)
    cfg = physics_config(spec["hall"], **cfg_kw)
    PhiAdd, Badd = mesh.background(B_IMF, DIPOLE_M, (0.0, 0.0, 0.0), 1e-3)
    UH, UNa, Phi = mesh.seeded_state(cfg, noise, seed, Badd)
    return dict(mesh=mesh, cfg=cfg, state=(UH, UNa, Phi), background=(PhiAdd, Badd), seed=seed, spec=spec)
