"""ORACLE — test infrastructure only (imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg).  ctypes loaders of

  * oracle/_build/liboracle.so  — the CPU restatement of the reference path (oracle/oracle.cpp)
  * oracle/_ref/libmpcns_ref.so — the reference's own sources compiled in place (oracle/ref_harness.cpp)

with the same method names as mpcns_uns_amd.Device so that parity tests read side by side.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libmpcns_ref.so")
# the reference's drivers linked against the device library through the C++ shim (mpcns_uns_amd/shim)
REF_GPU_LIB = os.path.join(os.path.dirname(REF_LIB), "libmpcns_ref_gpu.so")

_P, _D = C.c_void_p, C.c_double


def _ptr(a):
    """ctypes pointer to a contiguous array's data; the pointer object holds a reference to the array, so a
    converted temporary (_ptr(_f(x))) stays alive until the C call has returned."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def have_ref():
    return os.path.exists(REF_LIB)


def have_ref_on_device():
    return os.path.exists(REF_GPU_LIB)


class Oracle:
    """CPU restatement (one rank) of the reference step; same call surface as Device."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_LIB):
                raise RuntimeError("oracle not built: make -C oracle")
            L = C.CDLL(ORACLE_LIB)
            L.orc_create.argtypes = [_P, _P, C.POINTER(_P)]
            L.orc_destroy.argtypes = [_P]
            L.orc_last_error.restype = C.c_char_p
            L.orc_last_error.argtypes = [_P]
            for name, args in {
                "orc_cell_vector": [_P, _P, _P], "orc_cell_current": [_P, _P, _P],
                "orc_limited_gradients": [_P, _P, C.c_int, _P, C.c_int, _P, _P],
                "orc_flux_rhs": [_P, C.c_int, _P, _P, _P, _P, _P],
                "orc_ideal_emf": [_P, _P, _P, _P, _P, _P, _D, _P],
                "orc_faraday": [_P, _P, _P, _D], "orc_faraday_edge_dt": [_P, _P, _P, _P],
                "orc_div_b": [_P, _P, _P, _P, _P, _P], "orc_local_dt": [_P, _P, _P, _P, _P],
                "orc_add_sources": [_P, C.c_int, _P, _P, _P, _P, _D, _P],
                "orc_ind_set_frozen": [_P, _P, _P], "orc_ind_apply_F": [_P, _P, _P],
                "orc_ind_solve": [_P, _P, _P, _P, _D, _P, _P, _P, _P],
                "orc_ind_set_ideal_state": [_P, _P, _P], "orc_ind_ideal_linear": [_P, _P, _P],
                "orc_set_state": [_P, _P, _P, _P], "orc_get_state": [_P, _P, _P, _P],
                "orc_set_background": [_P, _P, _P], "orc_step": [_P, _D, _P, _P, _P],
                "orc_get_derived": [_P, C.c_int, _P], "orc_prepare": [_P],
                "orc_set_species_state": [_P, C.c_int, _P], "orc_get_species_state": [_P, C.c_int, _P],
                "orc_get_species_prim": [_P, C.c_int, _P], "orc_mars_implicit_sources": [_P, _P, _P, _D],
            }.items():
                getattr(L, name).argtypes = args
            cls._lib = L
        return cls._lib

    def __init__(self, mesh, cfg):
        self.mesh, self.cfg = mesh, cfg
        self.nc, self.nf, self.ne = mesh.n_cells, mesh.n_faces, mesh.n_edges
        h = _P()
        if self.lib().orc_create(C.addressof(mesh.desc()), C.addressof(cfg), C.byref(h)):
            raise RuntimeError("orc_create failed")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self.lib().orc_destroy(self._h)
            self._h = None

    def mars_implicit_sources(self, prm, neutral, dt):
        """MarsSolver::implicit_sources_ (MarsSolver.cpp:422-527) on the current state with B_tot of the last prepare."""
        nd = None if prm.n_neutral == 0 else _f(neutral)
        if nd is not None and nd.size != self.nc * prm.neutral_stride:
            raise ValueError("neutral densities: neutral_stride values per cell")
        self.lib().orc_mars_implicit_sources(self._h, C.addressof(prm), _ptr(nd), dt)

    def cell_vector(self, flux):
        out = np.zeros(3 * self.nc)
        self.lib().orc_cell_vector(self._h, _ptr(_f(flux)), _ptr(out))
        return out

    def cell_current(self, B):
        out = np.zeros(3 * self.nc)
        self.lib().orc_cell_current(self._h, _ptr(_f(B)), _ptr(out))
        return out

    def limited_gradients(self, W, species, fs, limit=True, ghost=None):
        out = np.zeros(15 * self.nc)
        self.lib().orc_limited_gradients(self._h, _ptr(_f(W)), species, _ptr(_f(fs)), int(limit),
                                         _ptr(None if ghost is None else _f(ghost)), _ptr(out))
        return out

    def flux_rhs(self, species, fs, W, grad, ghost=None):
        out = np.zeros(5 * self.nc)
        self.lib().orc_flux_rhs(self._h, species, _ptr(_f(fs)), _ptr(_f(W)), _ptr(_f(grad)),
                                _ptr(None if ghost is None else _f(ghost)), _ptr(out))
        return out

    def ideal_emf(self, u, B_tot, B_ind, phi_tot, phi_ind, cuct=0.5, E=None):
        E = np.zeros(self.ne) if E is None else _f(E).copy()
        self.lib().orc_ideal_emf(self._h, _ptr(_f(u)), _ptr(_f(B_tot)), _ptr(_f(B_ind)), _ptr(_f(phi_tot)),
                                 _ptr(_f(phi_ind)), cuct, _ptr(E))
        return E

    def faraday(self, E, Phi, dt):
        Phi = _f(Phi).copy()
        self.lib().orc_faraday(self._h, _ptr(_f(E)), _ptr(Phi), dt)
        return Phi

    def faraday_edge_dt(self, E, dt_edge, Phi):
        Phi = _f(Phi).copy()
        self.lib().orc_faraday_edge_dt(self._h, _ptr(_f(E)), _ptr(_f(dt_edge)), _ptr(Phi))
        return Phi

    def div_b(self, B, extra=None):
        a, r, w = C.c_double(), C.c_double(), C.c_int64()
        self.lib().orc_div_b(self._h, _ptr(_f(B)), _ptr(None if extra is None else _f(extra)), C.byref(a), C.byref(r), C.byref(w))
        return a.value, r.value, w.value

    def local_dt(self, W_H, W_Na, B_tot):
        dt = C.c_double()
        self.lib().orc_local_dt(self._h, _ptr(_f(W_H)), _ptr(_f(W_Na)), _ptr(_f(B_tot)), C.byref(dt))
        return dt.value

    def add_sources(self, species, W_H, W_Na, B_tot, J, dt, rhs):
        rhs = _f(rhs).copy()
        self.lib().orc_add_sources(self._h, species, _ptr(_f(W_H)), _ptr(_f(W_Na)), _ptr(_f(B_tot)), _ptr(_f(J)), dt, _ptr(rhs))
        return rhs

    def ind_set_frozen(self, B0, ne_eff):
        self.lib().orc_ind_set_frozen(self._h, _ptr(_f(B0)), _ptr(_f(ne_eff)))

    def ind_apply_F(self, phi):
        out = np.zeros(self.ne)
        self.lib().orc_ind_apply_F(self._h, _ptr(_f(phi)), _ptr(out))
        return out

    def ind_solve(self, Phi_n, E_expl, dirichlet, dt, dt_edge=None):
        E = np.zeros(self.ne)
        it, rr = C.c_int32(), C.c_double()
        d = np.ascontiguousarray(dirichlet, dtype=np.int8)
        self.lib().orc_ind_solve(self._h, _ptr(_f(Phi_n)), _ptr(_f(E_expl)), _ptr(d), dt,
                                 _ptr(None if dt_edge is None else _f(dt_edge)), _ptr(E), C.byref(it), C.byref(rr))
        return E, it.value, rr.value

    def ind_set_ideal_state(self, u, mask):
        self.lib().orc_ind_set_ideal_state(self._h, _ptr(_f(u)), _ptr(np.ascontiguousarray(mask, dtype=np.int8)))

    def ind_ideal_linear(self, phi):
        out = np.zeros(self.ne)
        self.lib().orc_ind_ideal_linear(self._h, _ptr(_f(phi)), _ptr(out))
        return out

    def set_state(self, U_H, U_Na, Phi):
        self.lib().orc_set_state(self._h, _ptr(_f(U_H)), _ptr(_f(U_Na)), _ptr(_f(Phi)))

    def get_state(self):
        UH, UNa, Phi = np.zeros(5 * self.nc), np.zeros(5 * self.nc), np.zeros(self.nf)
        self.lib().orc_get_state(self._h, _ptr(UH), _ptr(UNa), _ptr(Phi))
        return UH, UNa, Phi

    def set_species_state(self, species, U):
        if self.lib().orc_set_species_state(self._h, int(species), _ptr(_f(U))):
            raise RuntimeError("species out of range")

    def get_species_state(self, species):
        U = np.zeros(5 * self.nc)
        if self.lib().orc_get_species_state(self._h, int(species), _ptr(U)):
            raise RuntimeError("species out of range")
        return U

    def get_species_prim(self, species):
        W = np.zeros(5 * self.nc)
        if self.lib().orc_get_species_prim(self._h, int(species), _ptr(W)):
            raise RuntimeError("species out of range")
        return W

    def set_background(self, PhiAdd, Badd):
        self.lib().orc_set_background(self._h, _ptr(_f(PhiAdd)), _ptr(_f(Badd)))

    def step(self, dt_max=1e9, precision=0):
        dt, it, rr = C.c_double(), C.c_int32(), C.c_double()
        if self.lib().orc_step(self._h, dt_max, C.byref(dt), C.byref(it), C.byref(rr)):
            raise RuntimeError(self.lib().orc_last_error(self._h).decode())
        return dt.value, it.value, rr.value

    def prepare(self):
        self.lib().orc_prepare(self._h)

    def derived(self, name):
        idx = {"W_H": 0, "W_Na": 1, "B_ind": 2, "B_tot": 3, "J": 4, "u_plus": 5, "E": 6}[name]
        n = {0: 5 * self.nc, 1: 5 * self.nc, 6: self.ne}.get(idx, 3 * self.nc)
        out = np.zeros(n)
        self.lib().orc_get_derived(self._h, idx, _ptr(out))
        return out


class Reference:
    """The compiled reference (oracle/_ref) on a mesh given as cells of outward loops."""

    _lib = None
    _path = REF_LIB

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(cls._path):
                raise RuntimeError(f"{cls._path} not built (make -C oracle ref)")
            L = C.CDLL(cls._path)
            L.ref_create.argtypes = [C.c_int32, _P, C.c_int64, _P, _P, _P, _P, C.c_int64, _P, _P, _P, C.c_int,
                                     _P, C.c_int, _P, _P, _D, C.POINTER(_P)]
            L.ref_destroy.argtypes = [_P]
            L.ref_local_dt.restype = _D
            L.ref_last_induction_residual.restype = _D
            for name, args in {
                "ref_counts": [_P, _P], "ref_incidence": [_P, C.c_int, _P, _P, _P], "ref_face_tags": [_P, _P],
                "ref_validate": [_P, _P], "ref_geometry": [_P, C.c_char_p, _P], "ref_boundary": [_P, _P, _P],
                "ref_background": [_P, _P, _P], "ref_cell_vector": [_P, _P, _P], "ref_cell_current": [_P, _P, _P],
                "ref_limited_gradients": [_P, _P, C.c_int, _P, C.c_int, _P],
                "ref_flux_rhs": [_P, C.c_int, _P, _P, _P, _P],
                "ref_limited_gradients_ghost": [_P, _P, C.c_int, _P, C.c_int, _P, _P, _P],
                "ref_flux_rhs_ghost": [_P, C.c_int, _P, _P, _P, _P, _P, _P], "ref_faraday_edge_dt": [_P, _P, _P, _P],
                "ref_limited_gradients_wall_ghost": [_P, _P, C.c_int, _P, C.c_int, _P, _P], "ref_faraday_masked": [_P, _P, _P, _D, _P],
                "ref_flux_rhs_wall_ghost": [_P, C.c_int, _P, _P, _P, _P, _P],
                "ref_ideal_emf": [_P, _P, _P, _P, _P, _P, _D, _P], "ref_faraday": [_P, _P, _P, _D],
                "ref_div_b": [_P, _P, _P, _P, _P], "ref_local_dt": [_P, _P, _P, _P],
                "ref_add_sources": [_P, C.c_int, _P, _P, _P, _P, _P, _D, _P],
                "ref_ind_set_frozen": [_P, _P, _P], "ref_ind_apply_F": [_P, _P, _P],
                "ref_ind_solve": [_P, _P, _P, _P, _D, _P, _P, _P], "ref_solver_init": [_P],
                "ref_ind_set_ideal_state": [_P, _P, _P], "ref_ind_ideal_linear": [_P, _P, _P],
                "ref_set_state": [_P, _P, _P, _P], "ref_get_state": [_P, _P, _P, _P], "ref_step": [_P, _D, _P],
                "ref_get_derived": [_P, C.c_int, _P], "ref_last_induction_residual": [_P],
                "ref_write_mesh": [_P, C.c_char_p], "ref_set_output_dir": [_P, C.c_char_p], "ref_write_vtk": [_P], "ref_read_mesh": [C.c_char_p, _P, _P], "ref_topology_hash": [_P],
                "ref_write_fields": [C.c_char_p, C.c_uint64, C.c_int, _P, _P, _P, _P, _P],
                "ref_read_field": [C.c_char_p, C.c_char_p, _P, _P, _P, _P, _P],
                "ref_dipole_edge_integrals": [C.c_int, _P, _P, _P, _D, C.c_int64, _P, _P, _P],
                "ref_hilbert_order": [C.c_int32, _P, _P],
            }.items():
                getattr(L, name).argtypes = args
            L.ref_ind_solve.restype = C.c_int
            L.ref_topology_hash.restype = C.c_uint64
            cls._lib = L
        return cls._lib

    def __init__(self, xyz, cell_loops, cfg, tags=(), material=None, tagset=0, dipole_m=None, dipole_c=(0, 0, 0), soft=1e-3):
        xyz = _f(xyz).reshape(-1, 3)
        cell_ptr, loop_ptr, nodes = [0], [0], []
        for loops in cell_loops:
            for lp in loops:
                nodes.extend(int(v) for v in lp)
                loop_ptr.append(len(nodes))
            cell_ptr.append(len(loop_ptr) - 1)
        tag_ptr, tag_nodes, tag_id = [0], [], []
        for nds, t in tags:
            tag_nodes.extend(int(v) for v in nds)
            tag_ptr.append(len(tag_nodes))
            tag_id.append(int(t))
        a = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.int32))
        self._keep = [a(cell_ptr), a(loop_ptr), a(nodes), a(tag_ptr), a(tag_nodes), a(tag_id)]
        mat = None if material is None else a(material)
        self._keep.append(mat)
        cp, lp, nd, tp, tn, ti = self._keep[:6]
        dm = _f(dipole_m) if dipole_m is not None else None
        dc = _f(dipole_c)
        self._keep += [dm, dc, xyz]
        h = _P()
        if self.lib().ref_create(len(xyz), _ptr(xyz), len(cell_loops), _ptr(cp), _ptr(lp), _ptr(nd), _ptr(mat), len(ti), _ptr(tp),
                                 _ptr(tn), _ptr(ti), tagset, C.addressof(cfg), int(dm is not None), _ptr(dm), _ptr(dc), soft,
                                 C.byref(h)):
            raise RuntimeError("ref_create failed")
        self._h = h
        cnt = np.zeros(8, np.int64)
        self.lib().ref_counts(self._h, _ptr(cnt))
        self.counts = cnt
        self.nn, self.ne, self.nf, self.nc = (int(x) for x in cnt[:4])

    def incidence(self, which):
        k = {"e2n": 0, "f2e": 1, "c2f": 2, "e2f": 3, "f2c": 4}[which]
        rows = {0: self.ne, 1: self.nf, 2: self.nc, 3: self.ne, 4: self.nf}[k]
        nnz = {0: self.counts[4], 1: self.counts[5], 2: self.counts[6], 3: self.counts[5], 4: self.counts[6]}[k]
        ptr_, idx, sgn = np.zeros(rows + 1, np.int64), np.zeros(nnz, np.int32), np.zeros(nnz, np.int8)
        self.lib().ref_incidence(self._h, k, _ptr(ptr_), _ptr(idx), _ptr(sgn))
        return ptr_, idx, sgn

    def face_tags(self):
        t = np.zeros(self.nf, np.int32)
        self.lib().ref_face_tags(self._h, _ptr(t))
        return t

    def validate(self):
        o = np.zeros(5, np.int64)
        self.lib().ref_validate(self._h, _ptr(o))
        return dict(euler=int(o[0]), boundary_faces=int(o[1]), d1d0=int(o[2]), d2d1=int(o[3]), ok=bool(o[4]))

    def geometry(self, name):
        sizes = {"edge_vec": 3 * self.ne, "edge_mid": 3 * self.ne, "edge_len": self.ne, "face_area": 3 * self.nf,
                 "face_center": 3 * self.nf, "face_area_mag": self.nf, "cell_center": 3 * self.nc, "cell_volume": self.nc,
                 "edge_dual_area": 3 * self.ne, "ec_ptr": self.ne + 1, "ec_cell": int(self.counts[7]),
                 "ec_weight": int(self.counts[7]), "ef_weight": int(self.counts[5]), "star1": self.ne, "star2": self.nf,
                 "star3": self.nc}
        dt = np.int64 if name == "ec_ptr" else np.int32 if name == "ec_cell" else np.float64
        out = np.zeros(sizes[name], dt)
        if self.lib().ref_geometry(self._h, name.encode(), _ptr(out)):
            raise KeyError(name)
        return out

    def boundary(self):
        fk, em = np.zeros(self.nf, np.int8), np.zeros(self.ne, np.int8)
        self.lib().ref_boundary(self._h, _ptr(fk), _ptr(em))
        return fk, em

    def background(self):
        PA, BA = np.zeros(self.nf), np.zeros(3 * self.nc)
        self.lib().ref_background(self._h, _ptr(PA), _ptr(BA))
        return PA, BA

    def cell_vector(self, flux):
        out = np.zeros(3 * self.nc)
        self.lib().ref_cell_vector(self._h, _ptr(_f(flux)), _ptr(out))
        return out

    def cell_current(self, B):
        out = np.zeros(3 * self.nc)
        self.lib().ref_cell_current(self._h, _ptr(_f(B)), _ptr(out))
        return out

    def limited_gradients(self, W, species, fs, limit=True, ghost=None):
        """Stencil::limited_gradients; `ghost` (5 per face) is served through a GhostFn that returns the table entry.
        With a table, also returns the largest |rho| + |p| of the `inside` states the GhostFn was handed."""
        out = np.zeros(15 * self.nc)
        if ghost is None:
            self.lib().ref_limited_gradients(self._h, _ptr(_f(W)), species, _ptr(_f(fs)), int(limit), _ptr(out))
            return out
        seen = C.c_double(0.0)
        self.lib().ref_limited_gradients_ghost(self._h, _ptr(_f(W)), species, _ptr(_f(fs)), int(limit), _ptr(_f(ghost)), _ptr(out),
                                               C.byref(seen))
        return out, seen.value

    def flux_rhs(self, species, fs, W, grad, ghost=None):
        """fluid_flux_rhs; `ghost` as in limited_gradients (then also returns the GhostFn's largest inside state)."""
        out = np.zeros(5 * self.nc)
        if ghost is None:
            self.lib().ref_flux_rhs(self._h, species, _ptr(_f(fs)), _ptr(_f(W)), _ptr(_f(grad)), _ptr(out))
            return out
        seen = C.c_double(0.0)
        self.lib().ref_flux_rhs_ghost(self._h, species, _ptr(_f(fs)), _ptr(_f(W)), _ptr(_f(grad)), _ptr(_f(ghost)), _ptr(out),
                                      C.byref(seen))
        return out, seen.value

    def faraday_masked(self, E, Phi, dt, mask):
        """faraday_update with a caller-given face mask (nonzero = active): inactive faces keep their values."""
        out = _f(Phi).copy()
        self.lib().ref_faraday_masked(self._h, _ptr(_f(E)), _ptr(out), dt, _ptr(np.ascontiguousarray(mask, dtype=np.int8)))
        return out

    def limited_gradients_wall_ghost(self, W, species, fs, wall, limit=True):
        """Stencil::limited_gradients with a GhostFn of MarsSolver's form (MarsSolver.cpp:124-128): `wall` (5 per face)
        on wall faces, ghost_state of the inside state the reference hands it on the other boundary faces."""
        out = np.zeros(15 * self.nc)
        self.lib().ref_limited_gradients_wall_ghost(self._h, _ptr(_f(W)), species, _ptr(_f(fs)), int(limit), _ptr(_f(wall)), _ptr(out))
        return out

    def flux_rhs_wall_ghost(self, species, fs, W, grad, wall):
        """fluid_flux_rhs with the GhostFn of limited_gradients_wall_ghost (its outflow ghosts are the reconstructed
        face states)."""
        out = np.zeros(5 * self.nc)
        self.lib().ref_flux_rhs_wall_ghost(self._h, species, _ptr(_f(fs)), _ptr(_f(W)), _ptr(_f(grad)), _ptr(_f(wall)), _ptr(out))
        return out

    def ideal_emf(self, u, B_tot, B_ind, phi_tot, phi_ind, cuct=0.5, E=None):
        E = np.zeros(self.ne) if E is None else _f(E).copy()
        self.lib().ref_ideal_emf(self._h, _ptr(_f(u)), _ptr(_f(B_tot)), _ptr(_f(B_ind)), _ptr(_f(phi_tot)), _ptr(_f(phi_ind)),
                                 cuct, _ptr(E))
        return E

    def faraday(self, E, Phi, dt):
        Phi = _f(Phi).copy()
        self.lib().ref_faraday(self._h, _ptr(_f(E)), _ptr(Phi), dt)
        return Phi

    def faraday_edge_dt(self, E, dt_edge, Phi):
        Phi = _f(Phi).copy()
        self.lib().ref_faraday_edge_dt(self._h, _ptr(_f(E)), _ptr(_f(dt_edge)), _ptr(Phi))
        return Phi

    def div_b(self, B):
        a, r, w = C.c_double(), C.c_double(), C.c_int64()
        self.lib().ref_div_b(self._h, _ptr(_f(B)), C.byref(a), C.byref(r), C.byref(w))
        return a.value, r.value, w.value

    def local_dt(self, W_H, W_Na, B_tot):
        return self.lib().ref_local_dt(self._h, _ptr(_f(W_H)), _ptr(_f(W_Na)), _ptr(_f(B_tot)))

    def add_sources(self, species, W_H, W_Na, u_plus, B_tot, J, dt, rhs):
        rhs = _f(rhs).copy()
        self.lib().ref_add_sources(self._h, species, _ptr(_f(W_H)), _ptr(_f(W_Na)), _ptr(_f(u_plus)), _ptr(_f(B_tot)), _ptr(_f(J)),
                                   dt, _ptr(rhs))
        return rhs

    def ind_set_frozen(self, B0, ne_eff):
        self.lib().ref_ind_set_frozen(self._h, _ptr(_f(B0)), _ptr(_f(ne_eff)))

    def ind_apply_F(self, phi):
        out = np.zeros(self.ne)
        self.lib().ref_ind_apply_F(self._h, _ptr(_f(phi)), _ptr(out))
        return out

    def ind_solve(self, Phi_n, E_expl, dirichlet, dt, dt_edge=None):
        E = np.zeros(self.ne)
        rr = C.c_double()
        d = np.ascontiguousarray(dirichlet, dtype=np.int8)
        it = self.lib().ref_ind_solve(self._h, _ptr(_f(Phi_n)), _ptr(_f(E_expl)), _ptr(d), dt,
                                      _ptr(None if dt_edge is None else _f(dt_edge)), _ptr(E), C.byref(rr))
        return E, it, rr.value

    def ind_set_ideal_state(self, u, mask):
        self.lib().ref_ind_set_ideal_state(self._h, _ptr(_f(u)), _ptr(np.ascontiguousarray(mask, dtype=np.int8)))

    def ind_ideal_linear(self, phi):
        out = np.zeros(self.ne)
        self.lib().ref_ind_ideal_linear(self._h, _ptr(_f(phi)), _ptr(out))
        return out

    def write_vtk(self, out_dir):
        """mercury::Solver::write_vtk (MercuryUns/driver/Solver.cpp:677-718) into out_dir (set before solver_init)."""
        if self.lib().ref_write_vtk(self._h):
            raise RuntimeError("ref_write_vtk failed")

    def set_output_dir(self, out_dir):
        self.lib().ref_set_output_dir(self._h, str(out_dir).encode())

    def solver_init(self):
        if self.lib().ref_solver_init(self._h):
            raise RuntimeError("ref_solver_init failed")

    def set_state(self, U_H, U_Na, Phi):
        self.lib().ref_set_state(self._h, _ptr(_f(U_H)), _ptr(_f(U_Na)), _ptr(_f(Phi)))

    def get_state(self):
        UH, UNa, Phi = np.zeros(5 * self.nc), np.zeros(5 * self.nc), np.zeros(self.nf)
        self.lib().ref_get_state(self._h, _ptr(UH), _ptr(UNa), _ptr(Phi))
        return UH, UNa, Phi

    def step(self, dt_max=1e9):
        dt = C.c_double()
        if self.lib().ref_step(self._h, dt_max, C.byref(dt)):
            raise RuntimeError("reference step failed")
        return dt.value

    def derived(self, name):
        idx = {"W_H": 0, "W_Na": 1, "B_ind": 2, "B_tot": 3, "J": 4}[name]
        out = np.zeros((5 if idx < 2 else 3) * self.nc)
        self.lib().ref_get_derived(self._h, idx, _ptr(out))
        return out


def tet_cell_loops(tets):
    """Outward loops of positively oriented tetrahedra, in the order the mesh builder uses."""
    out = []
    for a, b, c, d in np.asarray(tets).reshape(-1, 4):
        out.append([(a, c, b), (a, b, d), (a, d, c), (b, c, d)])
    return out


class ReferenceOnDevice(Reference):
    """The reference's drivers (mercury::Solver, harness) with the mhd_common calls of the hot path served by the
    device library through the C++ shim (mpcns_uns_amd/shim/mhd_gpu_shim.cpp): the drop-in configuration."""

    _lib = None
    _path = REF_GPU_LIB


def reference_for_mesh(mesh, cfg, dipole_m=None, dipole_c=(0, 0, 0), soft=1e-3, on_device=False):
    """Reference objects on the same tetrahedral mesh (same node order, cell order and loops)."""
    xyz = mesh.array("xyz")
    tets = mesh.array("tets")
    ftag = mesh.array("face_tag")
    f2e_ptr, f2e_idx, f2e_sgn = mesh.array("f2e.ptr"), mesh.array("f2e.idx"), mesh.array("f2e.sgn")
    e2n_idx = mesh.array("e2n.idx").reshape(-1, 2)
    tags = []
    for f in np.nonzero(ftag)[0]:
        es = f2e_idx[f2e_ptr[f]:f2e_ptr[f + 1]]
        nodes = sorted(set(int(x) for x in e2n_idx[es].ravel()))
        tags.append((nodes, int(ftag[f])))
    cls = ReferenceOnDevice if on_device else Reference
    return cls(xyz, tet_cell_loops(tets), cfg, tags=tags, dipole_m=dipole_m, dipole_c=dipole_c, soft=soft)


def reference_hilbert_order(centres):
    """The reference partitioner's Hilbert order of points (uns::partition_cells, one part per point)."""
    c = _f(centres).reshape(-1, 3)
    order = np.zeros(len(c), np.int32)
    if Reference.lib().ref_hilbert_order(len(c), _ptr(c), _ptr(order)):
        raise RuntimeError("ref_hilbert_order failed")
    return order
