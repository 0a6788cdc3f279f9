"""One mesh resident on the GPU (include/mhdg.h).

Methods mirror the reference's mhd_common calls (host arrays in and out) and the fused Mercury step on the
resident state.  Every call goes through the HIP kernels of libmhdg.so; a failure raises.
"""
import ctypes as C

import numpy as np

from ._lib import lib, ptr

DERIVED = {"W_H": 0, "W_Na": 1, "B_ind": 2, "B_tot": 3, "J": 4, "u_plus": 5, "E": 6}


class Device:
    def __init__(self, mesh, cfg):
        self.mesh = mesh
        self.cfg = cfg
        self.nc, self.nf, self.ne = mesh.n_cells, mesh.n_faces, mesh.n_edges
        h = C.c_void_p()
        if lib().mhdg_create(C.byref(mesh.desc()), C.byref(cfg), C.byref(h)):
            raise RuntimeError("mhdg_create: " + lib().mhdg_last_error(None).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mhdg_destroy(self._h)
            self._h = None

    __del__ = close

    def _ok(self, rc):
        if rc:
            raise RuntimeError(lib().mhdg_last_error(self._h).decode())

    # Inputs crossing the C-ABI: contiguous float64 (or int8) arrays of the exact size the entry point reads, so a
    # float32, list or strided input is converted (and kept alive by ptr() for the call) and a wrong size raises
    # here instead of letting the library read past the caller's buffer.
    _SIZES = {"c": lambda d: d.nc, "c3": lambda d: 3 * d.nc, "c5": lambda d: 5 * d.nc, "c15": lambda d: 15 * d.nc,
              "f": lambda d: d.nf, "f5": lambda d: 5 * d.nf, "e": lambda d: d.ne, "fs": lambda d: 5}

    def _in(self, a, kind, dtype=np.float64):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype).ravel()
        n = self._SIZES[kind](self)
        if a.size != n:
            raise ValueError(f"array of {a.size} values where {n} are expected")
        return a

    @staticmethod
    def _sized(a, n):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != n:
            raise ValueError(f"array of {a.size} values where {n} are expected")
        return a

    @staticmethod
    def info():
        buf = C.create_string_buffer(256)
        lib().mhdg_device_info(buf, 256)
        return buf.value.decode()

    # ------------------------------------------------------------------ mhd_common calls
    def cell_vector(self, flux):
        out = np.zeros(3 * self.nc)
        self._ok(lib().mhdg_cell_vector(self._h, ptr(self._in(flux, 'f')), ptr(out)))
        return out

    def cell_current(self, B):
        out = np.zeros(3 * self.nc)
        self._ok(lib().mhdg_cell_current(self._h, ptr(self._in(B, 'c3')), ptr(out)))
        return out

    def limited_gradients(self, W, species, fs, limit=True, ghost=None):
        out = np.zeros(15 * self.nc)
        g = self._in(ghost, 'f5')
        self._ok(lib().mhdg_limited_gradients(self._h, ptr(self._in(W, 'c5')), species, ptr(self._in(fs, 'fs')), int(limit), ptr(g), ptr(out)))
        return out

    def flux_rhs(self, species, fs, W, grad, ghost=None):
        out = np.zeros(5 * self.nc)
        g = self._in(ghost, 'f5')
        self._ok(lib().mhdg_flux_rhs(self._h, species, ptr(self._in(fs, 'fs')), ptr(self._in(W, 'c5')), ptr(self._in(grad, 'c15')), ptr(g), ptr(out)))
        return out

    def ideal_emf(self, u, B_tot, B_ind, phi_tot, phi_ind, cuct=0.5, E=None):
        E = np.zeros(self.ne) if E is None else self._in(E, 'e').copy()
        self._ok(lib().mhdg_ideal_emf(self._h, ptr(self._in(u, 'c3')), ptr(self._in(B_tot, 'c3')), ptr(self._in(B_ind, 'c3')),
                                      ptr(self._in(phi_tot, 'f')), ptr(self._in(phi_ind, 'f')), cuct, ptr(E)))
        return E

    def faraday(self, E, Phi, dt):
        Phi = self._in(Phi, 'f').copy()
        self._ok(lib().mhdg_faraday(self._h, ptr(self._in(E, 'e')), ptr(Phi), dt))
        return Phi

    def faraday_edge_dt(self, E, dt_edge, Phi):
        Phi = self._in(Phi, 'f').copy()
        self._ok(lib().mhdg_faraday_edge_dt(self._h, ptr(self._in(E, 'e')), ptr(self._in(dt_edge, 'e')), ptr(Phi)))
        return Phi

    def div_b(self, B, extra=None):
        a, r, w = C.c_double(), C.c_double(), C.c_int64()
        self._ok(lib().mhdg_div_b(self._h, ptr(self._in(B, 'f')), ptr(self._in(extra, 'f')), C.byref(a), C.byref(r), C.byref(w)))
        return a.value, r.value, w.value

    def local_dt(self, W_H, W_Na, B_tot):
        dt = C.c_double()
        self._ok(lib().mhdg_local_dt(self._h, ptr(self._in(W_H, 'c5')), ptr(self._in(W_Na, 'c5')), ptr(self._in(B_tot, 'c3')), C.byref(dt)))
        return dt.value

    def add_sources(self, species, W_H, W_Na, B_tot, J, dt, rhs):
        rhs = self._in(rhs, 'c5').copy()
        self._ok(lib().mhdg_add_sources(self._h, species, ptr(self._in(W_H, 'c5')), ptr(self._in(W_Na, 'c5')), ptr(self._in(B_tot, 'c3')),
                                        ptr(self._in(J, 'c3')), dt, ptr(rhs)))
        return rhs

    def ind_set_frozen(self, B0, ne_eff):
        self._ok(lib().mhdg_ind_set_frozen(self._h, ptr(self._in(B0, 'c3')), ptr(self._in(ne_eff, 'c'))))

    def ind_apply_F(self, phi):
        out = np.zeros(self.ne)
        self._ok(lib().mhdg_ind_apply_F(self._h, ptr(self._in(phi, 'f')), ptr(out)))
        return out

    def ind_flags(self):
        """(active, has_ideal, has_hall) of ImplicitInduction (ImplicitInduction.h:31-33)."""
        a, i, h = C.c_int32(), C.c_int32(), C.c_int32()
        self._ok(lib().mhdg_ind_flags(self._h, C.byref(a), C.byref(i), C.byref(h)))
        return bool(a.value), bool(i.value), bool(h.value)

    def ind_set_ideal_state(self, u, mask):
        self._ok(lib().mhdg_ind_set_ideal_state(self._h, ptr(self._in(u, 'c3')), ptr(self._in(np.asarray(mask) != 0, 'e', np.int8))))

    def ind_ideal_linear(self, phi):
        out = np.zeros(self.ne)
        self._ok(lib().mhdg_ind_ideal_linear(self._h, ptr(self._in(phi, 'f')), ptr(out)))
        return out

    def ind_solve(self, Phi_n, E_expl, dirichlet, dt, precision=0, dt_edge=None):
        E = np.zeros(self.ne)
        it, rr = C.c_int32(), C.c_double()
        d = self._in(np.asarray(dirichlet) != 0, 'e', np.int8)
        self._ok(lib().mhdg_ind_solve(self._h, ptr(self._in(Phi_n, 'f')), ptr(self._in(E_expl, 'e')), ptr(d), dt,
                                      ptr(self._in(dt_edge, 'e')), precision, ptr(E), C.byref(it), C.byref(rr)))
        return E, it.value, rr.value

    def mg_apply(self, v, dirichlet, dt):
        """One V-cycle z = M v of the multigrid preconditioner (precision 2) with the frozen state of ind_set_frozen,
        for the Dirichlet set and time step given (the cycle the flexible solve applies to each Krylov vector)."""
        z = np.zeros(self.ne)
        d = self._in(np.asarray(dirichlet) != 0, 'e', np.int8)
        self._ok(lib().mhdg_mg_apply(self._h, ptr(d), float(dt), ptr(self._in(v, 'e')), ptr(z)))
        return z

    def mg_info(self, dirichlet):
        """Unknowns per level of the multigrid hierarchy for this Dirichlet set (built on first use)."""
        lv, n = C.c_int32(), np.zeros(16, np.int32)
        d = self._in(np.asarray(dirichlet) != 0, 'e', np.int8)
        self._ok(lib().mhdg_mg_info(self._h, ptr(d), 16, C.byref(lv), ptr(n)))
        return n[:lv.value].tolist()

    def ind_apply_G(self, phi):
        """Induction operator with the Hall field B_add + R phi of phi itself (quadratic in phi)."""
        out = np.zeros(self.ne)
        self._ok(lib().mhdg_ind_apply_G(self._h, ptr(self._in(phi, 'f')), ptr(out)))
        return out

    def ind_jvp(self, phi_lin, dphi):
        """Tangent-linear product G'(phi_lin) dphi."""
        out = np.zeros(self.ne)
        self._ok(lib().mhdg_ind_jvp(self._h, ptr(self._in(phi_lin, 'f')), ptr(self._in(dphi, 'f')), ptr(out)))
        return out

    def ind_newton(self, Phi_n, E_expl, dirichlet, dt, E0=None, rtol=1e-8, max_newton=20, eta=1e-4):
        """Newton-Krylov solve of the nonlinear Hall induction step: (E, newton steps, GMRES iterations, rel. residual)."""
        E = np.zeros(self.ne)
        k, g, rr = C.c_int32(), C.c_int32(), C.c_double()
        d = self._in(np.asarray(dirichlet) != 0, 'e', np.int8)
        self._ok(lib().mhdg_ind_newton(self._h, ptr(self._in(Phi_n, 'f')), ptr(self._in(E_expl, 'e')), ptr(d), dt, ptr(self._in(E0, 'e')),
                                       rtol, max_newton, eta, ptr(E), C.byref(k), C.byref(g), C.byref(rr)))
        return E, k.value, g.value, rr.value

    def shell_energy(self, r0, r1):
        """sum over fluid cells with r0 <= |x_c| <= r1 of |B_ind + B_add|^2 V of the resident state."""
        J = C.c_double()
        self._ok(lib().mhdg_shell_energy(self._h, r0, r1, C.byref(J)))
        return J.value

    def explicit_tangent(self, n_steps, r0, r1, dU_H=None, dU_Na=None, dPhi=None, dPhiAdd=None, dBadd=None, dfs=None, dE_inf=None,
                         dt_max=1e9, states=False):
        """Forward-mode derivative of n explicit steps from the resident state along the given seeds:
        (J, dJ) of the shell energy of the final state, and the final state tangents when `states`."""
        J, dJ = C.c_double(), C.c_double()
        seeds = [self._in(a, k) for a, k in ((dU_H, 'c5'), (dU_Na, 'c5'), (dPhi, 'f'), (dPhiAdd, 'f'), (dBadd, 'c3'))]
        seeds += [None if dfs is None else self._sized(dfs, 10), None if dE_inf is None else self._sized(dE_inf, 3)]
        outs = [np.zeros(5 * self.nc), np.zeros(5 * self.nc), np.zeros(self.nf)] if states else [None, None, None]
        self._ok(lib().mhdg_explicit_tangent(self._h, n_steps, dt_max, *[ptr(a) for a in seeds], r0, r1, C.byref(J), C.byref(dJ),
                                             *[ptr(a) for a in outs]))
        return (J.value, dJ.value, *outs) if states else (J.value, dJ.value)

    def step_vjp(self, n_steps, r0, r1, dt_max=1e9):
        """Reverse mode of explicit_tangent: J and its gradient with respect to the initial state and the
        background / free-stream parameters, as a dict (U_H, U_Na, Phi, PhiAdd, Badd, fs (2x5), E_inf)."""
        J = C.c_double()
        g = dict(U_H=np.zeros(5 * self.nc), U_Na=np.zeros(5 * self.nc), Phi=np.zeros(self.nf), PhiAdd=np.zeros(self.nf),
                 Badd=np.zeros(3 * self.nc), fs=np.zeros(10), E_inf=np.zeros(3))
        self._ok(lib().mhdg_step_vjp(self._h, n_steps, dt_max, r0, r1, C.byref(J), *[ptr(g[k]) for k in
                                     ("U_H", "U_Na", "Phi", "PhiAdd", "Badd", "fs", "E_inf")]))
        g["fs"] = g["fs"].reshape(2, 5)
        return J.value, g

    # ------------------------------------------------------------------ resident state
    def set_state(self, U_H, U_Na, Phi):
        self._ok(lib().mhdg_set_state(self._h, ptr(self._in(U_H, 'c5')), ptr(self._in(U_Na, 'c5')), ptr(self._in(Phi, 'f'))))

    def set_state_from_fields(self, path):
        """Resident state from a restart file in the layout of mercury::Solver::write_global_fields
        (MercuryUns/driver/Solver.cpp:721-745): U_H, U_Na (cell, 5) and B (face flux)."""
        from .mesh import read_fields

        _, f = read_fields(path)
        for name, dim, nc, n in (("U_H", 3, 5, self.nc), ("U_Na", 3, 5, self.nc), ("B", 2, 1, self.nf)):
            if name not in f or f[name][0] != dim or f[name][1] != nc or f[name][2].size != nc * n:
                raise RuntimeError(f"{path}: no {name} record of {n} entities x {nc} on this mesh")
        self.set_state(f["U_H"][2], f["U_Na"][2], f["B"][2])

    def get_state(self):
        UH, UNa, Phi = np.zeros(5 * self.nc), np.zeros(5 * self.nc), np.zeros(self.nf)
        self._ok(lib().mhdg_get_state(self._h, ptr(UH), ptr(UNa), ptr(Phi)))
        return UH, UNa, Phi

    def set_species_state(self, species, U):
        """Conservative state (5 per cell) of ion species `species` < n_species (ten-ion fluids)."""
        self._ok(lib().mhdg_set_species_state(self._h, int(species), ptr(self._in(U, 'c5'))))

    def get_species_state(self, species):
        U = np.zeros(5 * self.nc)
        self._ok(lib().mhdg_get_species_state(self._h, int(species), ptr(U)))
        return U

    def get_species_prim(self, species):
        """Primitive state of `species` written by the last step's prepare."""
        W = np.zeros(5 * self.nc)
        self._ok(lib().mhdg_get_species_prim(self._h, int(species), ptr(W)))
        return W

    def set_background(self, PhiAdd, BaddCell):
        self._ok(lib().mhdg_set_background(self._h, ptr(self._in(PhiAdd, 'f')), ptr(self._in(BaddCell, 'c3'))))

    def step(self, dt_max=1e9, precision=0):
        dt, it, rr = C.c_double(), C.c_int32(), C.c_double()
        self._ok(lib().mhdg_step(self._h, dt_max, precision, C.byref(dt), C.byref(it), C.byref(rr)))
        return dt.value, it.value, rr.value

    def step_explicit(self, dt_max=1e9):
        dt = C.c_double()
        self._ok(lib().mhdg_step_explicit(self._h, dt_max, C.byref(dt)))
        return dt.value

    def step_hall(self, dt_max=1e9, precision=0):
        dt, it, rr = C.c_double(), C.c_int32(), C.c_double()
        self._ok(lib().mhdg_step_hall(self._h, dt_max, precision, C.byref(dt), C.byref(it), C.byref(rr)))
        return dt.value, it.value, rr.value

    def state_save(self):
        """Device-side copy of the resident state (every species, Phi)."""
        self._ok(lib().mhdg_state_save(self._h))

    def state_restore(self):
        """Back to the saved state, without the implicit solve's warm start."""
        self._ok(lib().mhdg_state_restore(self._h))

    def prepare(self):
        """Solver::prepare on the resident state: derived() then gives the current state's fields."""
        self._ok(lib().mhdg_prepare(self._h))

    def mars_implicit_sources(self, prm, neutral, dt):
        """MarsSolver::implicit_sources_ on the resident state with B_tot of the last prepare (prm: MarsParams;
        neutral: densities [cm^-3], prm.neutral_stride per cell)."""
        nd = None if prm.n_neutral == 0 else self._sized(np.ascontiguousarray(neutral, dtype=np.float64), self.nc * prm.neutral_stride)
        self._ok(lib().mhdg_mars_implicit_sources(self._h, C.byref(prm), ptr(nd), dt))

    def derived(self, name):
        n = {0: 5, 1: 5, 6: None}.get(DERIVED[name], 3)
        out = np.zeros(self.ne if n is None else n * self.nc)
        self._ok(lib().mhdg_get_derived(self._h, DERIVED[name], ptr(out)))
        return out

    def sync(self):
        self._ok(lib().mhdg_sync(self._h))

    def profile(self, enable=True):
        """Per-kernel HIP-event timing on the context stream (True on, False off, None clear)."""
        self._ok(lib().mhdg_profile(self._h, -1 if enable is None else int(bool(enable))))

    def profile_only(self, name=None):
        """Time only the kernels called `name` (None: every kernel)."""
        self._ok(lib().mhdg_profile_only(self._h, None if name is None else name.encode()))

    def profile_report(self):
        """{kernel name: (launches, total milliseconds)} accumulated while profiling was on."""
        buf = C.create_string_buffer(1 << 16)
        self._ok(lib().mhdg_profile_report(self._h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out
