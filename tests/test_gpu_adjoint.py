"""Reverse mode (adjoint) of the explicit step: mhdg_step_vjp.

Objective as in the tangent tests: J = sum |B|^2 V over the shell 1.2 <= r <= 1.5 after 5 explicit steps.
The adjoint is pinned by the tangent (forward mode, the Dual instantiation of the step's own kernels) through
the dot-product identity <grad J, v> = dJ[v] for random seeds of every input, and by central finite
differences of the double step for the 3 upstream IMF components and the upstream density.  The objective is also
wrapped in a torch.autograd.Function whose forward runs the double step and whose backward runs the adjoint
(mhdg_step_vjp); and the dot-product identity is checked on the 1.95M-cell shell of the differentiation workload.
"""
import numpy as np
import pytest

from test_gpu_explicit_tangent import R0, R1, STEPS, _inputs, energy, p0, tangent

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    from mpcns_uns_amd import workloads

    return workloads.build(net=(12, 48, 6), hall=False)


def _device(w, p=None, cfg_kw=None):
    from mpcns_uns_amd import Device

    if p is None:
        cfg, bg, st = w["cfg"], w["background"], w["state"]
    else:
        cfg, bg, st = _inputs(w, p)
    if cfg_kw:
        cfg = type(cfg).from_buffer_copy(cfg)
        for k, v in cfg_kw.items():
            setattr(cfg, k, v)
    dev = Device(w["mesh"], cfg)
    dev.set_background(*bg)
    dev.set_state(*st)
    return dev


def _param_gradient(w, p, g):
    """dJ/dp for p = (B_imf x, y, z, upstream rho_H) from the adjoint's input gradients."""
    m = w["mesh"]
    cfg, _, st = _inputs(w, p)
    out = np.zeros(4)
    for k in range(3):
        dPhiAdd = np.ascontiguousarray(m.array("face_area").reshape(-1, 3)[:, k])
        e = np.zeros(3)
        e[k] = 1.0
        out[k] = g["PhiAdd"] @ dPhiAdd + g["Badd"] @ m.cell_vector(dPhiAdd) + g["E_inf"] @ (-np.cross(np.array(cfg.u_inf[:]), e))
    U = st[0].reshape(-1, 5)
    ke = 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0]
    dU = np.column_stack([U[:, 0], U[:, 1], U[:, 2], U[:, 3], ke]) / p[3]
    out[3] = g["U_H"] @ dU.ravel() + g["fs"][0, 0]
    return out


@pytest.mark.parametrize("second_order", [1, 0])
def test_dot_product_with_the_tangent(problem, second_order):
    _check_dot_product(problem, second_order, R0, R1)


@pytest.mark.parametrize("second_order", [1, 0])
def test_dot_product_on_a_hexahedron_pyramid_box(second_order):
    """The same identity on a box of hexahedra and pyramids (cells of five and six faces, faces of three and four
    edges; every boundary face outflow), J over 1.5 <= r <= 3 of the box, which lies in 1.5 <= r <= 3.9."""
    from meshes import mixed_workload

    _check_dot_product(mixed_workload(False, nx=8, ny=6, nz=6), second_order, 1.5, 3.0)


def _check_dot_product(w, second_order, r0, r1):
    m = w["mesh"]
    dev = _device(w, cfg_kw=dict(second_order=second_order))
    before = dev.get_state()
    J, g = dev.step_vjp(STEPS, r0, r1)
    for a, b in zip(before, dev.get_state()):
        assert np.array_equal(a, b)  # the resident state is left unchanged
    rng = np.random.default_rng(57)
    seeds = dict(U_H=rng.normal(size=5 * m.n_cells) * 1e-3, U_Na=rng.normal(size=5 * m.n_cells) * 1e-6,
                 Phi=rng.normal(size=m.n_faces) * 1e-3, PhiAdd=rng.normal(size=m.n_faces) * 1e-3,
                 Badd=rng.normal(size=3 * m.n_cells) * 1e-3, fs=rng.normal(size=10) * 1e-3, E_inf=rng.normal(size=3) * 1e-3)
    Jt, dJ = dev.explicit_tangent(STEPS, r0, r1, dU_H=seeds["U_H"], dU_Na=seeds["U_Na"], dPhi=seeds["Phi"], dPhiAdd=seeds["PhiAdd"],
                                  dBadd=seeds["Badd"], dfs=seeds["fs"], dE_inf=seeds["E_inf"])
    assert J == Jt
    parts = {k: float(np.ravel(g[k]) @ np.ravel(seeds[k])) for k in seeds}
    total = sum(parts.values())
    scale = sum(abs(v) for v in parts.values())
    assert abs(total - dJ) <= 1e-9 * scale, (total, dJ, parts)
    # each input separately (a wrong transposed kernel shows up in its own term)
    for k in ("U_H", "Phi", "PhiAdd", "Badd", "E_inf"):
        kw = {"d" + k: seeds[k]}
        _, dk = dev.explicit_tangent(STEPS, r0, r1, **kw)
        assert abs(parts[k] - dk) <= 1e-9 * (abs(dk) + 1e-12 * scale), (k, parts[k], dk)


def test_parameter_gradient_against_differences_and_tangent(problem):
    w = problem
    p = p0(w)
    J, g = _device(w, p).step_vjp(STEPS, R0, R1)
    assert J == energy(w, p)
    grad = _param_gradient(w, p, g)
    h = 1e-6
    for k in range(4):
        e = np.zeros(4)
        e[k] = h
        fd = (energy(w, p + e) - energy(w, p - e)) / (2 * h)
        _, dJ = tangent(w, p, k)
        assert abs(grad[k] - dJ) <= 1e-9 * abs(dJ), (k, grad[k], dJ)
        assert abs(grad[k] - fd) <= 1e-5 * abs(grad[k]) + 1e-9 * J, (k, grad[k], fd)


def test_vjp_rejects_ten_ion_contexts():
    from mpcns_uns_amd import Device, workloads

    w = workloads.build("shell_4m_mars", net=(6, 12, 2), hall=False)
    dev = Device(w["mesh"], w["cfg"])
    workloads.load(dev, w)
    with pytest.raises(RuntimeError):
        dev.step_vjp(1, R0, R1)


def _shell_energy_function(w):
    """J(p) as a torch.autograd.Function: forward = the double explicit steps on the device, backward = one reverse
    sweep (mhdg_step_vjp) mapped to the 4 parameters by the chain rule through the inputs they set."""
    torch = pytest.importorskip("torch")

    class ShellEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, p):
            ctx.save_for_backward(p)
            return torch.tensor(energy(w, p.detach().cpu().double().numpy()), dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            (p,) = ctx.saved_tensors
            pn = p.detach().cpu().double().numpy()
            _, grads = _device(w, pn).step_vjp(STEPS, R0, R1)
            return g * torch.from_numpy(_param_gradient(w, pn, grads))

    return ShellEnergy


def test_torch_autograd_function_backward_is_the_adjoint(problem):
    torch = pytest.importorskip("torch")

    w = problem
    p = torch.tensor(p0(w), dtype=torch.float64, requires_grad=True)
    J = _shell_energy_function(w).apply(p)
    (2.0 * J).backward()
    grad = p.grad.numpy() / 2.0
    h = 1e-6
    pn = p0(w)
    for k in range(4):
        e = np.zeros(4)
        e[k] = h
        fd = (energy(w, pn + e) - energy(w, pn - e)) / (2 * h)
        _, dJ = tangent(w, pn, k)
        assert abs(grad[k] - dJ) <= 1e-9 * abs(dJ), (k, grad[k], dJ)
        assert abs(grad[k] - fd) <= 1e-5 * abs(grad[k]) + 1e-9 * float(J), (k, grad[k], fd)


def test_dot_product_at_differentiation_size():
    """Config-size check: <grad J, v> = dJ[v] for a random seed of the state and the background on the 1.95M-cell
    shell (5 explicit steps), one reverse sweep against one tangent sweep."""
    from mpcns_uns_amd import Device, workloads

    w = workloads.build("shell_2m_hall", hall=False)
    m = w["mesh"]
    print(f"\n[adjoint] shell_2m: {m.n_cells} cells", flush=True)
    dev = Device(m, w["cfg"])
    workloads.load(dev, w)
    J, g = dev.step_vjp(STEPS, R0, R1)
    rng = np.random.default_rng(91)
    vU, vPhi, vPA = rng.normal(size=5 * m.n_cells) * 1e-3, rng.normal(size=m.n_faces) * 1e-3, rng.normal(size=m.n_faces) * 1e-3
    Jt, dJ = dev.explicit_tangent(STEPS, R0, R1, dU_H=vU, dPhi=vPhi, dPhiAdd=vPA)
    parts = [float(g["U_H"] @ vU), float(g["Phi"] @ vPhi), float(g["PhiAdd"] @ vPA)]
    print(f"[adjoint] shell_2m: J {J:.6e}, <grad J, v> {sum(parts):.12e}, dJ[v] {dJ:.12e}", flush=True)
    assert J == Jt and J > 0
    assert abs(sum(parts) - dJ) <= 1e-9 * sum(abs(x) for x in parts)
