"""Python host API of the gfx950 implementation of DTC's per-control-volume micro-FE path.

Mirrors the reference's per-domain operations (f-src/mod_dtc_main_subroutines.f90:52
exec_single_domain and what it calls) over the C ABI of libdtc_amd.so:

  ddc_params / domain_bounds   calc_general_ddc_params, generate_geometry index math
  extract_block                the voxel block read of generate_geometry
  mesh                         gen_quadmesh_from_phi (bit-exact numbering)
  boundaries                   generate_boundaries + determine_prescribed_displ
  element_stiffness            Hexe08 / Hexe20
  solve_domain / Batch         assembly + 24 KSPSolves + reactions + effective stiffness
  rotation_search              CR_1 / CR_2 search of calc_effective_material_parameters

Voxel arrays are numpy arrays of shape (nz, ny, nx) (x fastest in memory, the
reference's Phi(x, y, z) layout).  Matrices returned here are numpy arrays in
mathematical orientation (ff[dof, lc], C[i, j]).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import (DTC_ERR_EMPTY_MESH, DTC_ERR_FEW_CONSTRAINTS, DTC_ERR_NO_CONVERGENCE, DTC_FP32_REFINE, DTC_FP64,
                   DTC_KSTAT_BRES, DTC_KSTAT_CHEB, DTC_KSTAT_MV, DTC_KSTAT_RESID, DTC_PC_JACOBI, DTC_PC_MG, DTC_OK,
                   KSTAT_NAMES, DTCError, DomainDesc, DomainResult, SolveOpts, check, load)

KIND = {np.dtype(np.int8): 1, np.dtype(np.int16): 2, np.dtype(np.int32): 4}

__all__ = ["Domain", "Batch", "mesh", "mesh_count", "boundaries", "element_stiffness", "solve_domain", "solve_batch",
           "rotation_search", "ddc_params", "domain_bounds", "extract_block", "DTCError", "DTC_FP64",
           "DTC_FP32_REFINE", "load"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i64(*v):
    return np.ascontiguousarray(np.array(v, np.int64).reshape(-1))


def _f64(*v):
    return np.ascontiguousarray(np.array(v, np.float64).reshape(-1))


def _phi(phi):
    phi = np.ascontiguousarray(phi)
    if phi.dtype not in KIND:
        raise TypeError(f"voxels must be int8/int16/int32 (signed ik1/ik2/ik4), got {phi.dtype}")
    if phi.ndim != 3:
        raise ValueError("voxels must be a 3-D array (nz, ny, nx)")
    return phi


def ddc_params(size_domain, delta, vdim):
    x_D, nn_D, xdp = np.zeros(3, np.int64), np.zeros(3, np.int64), np.zeros(3)
    check("dtc_ddc_params", load().dtc_ddc_params(_p(_f64(*size_domain)), _p(_f64(*delta)), _p(_i64(*vdim)),
                                                  _p(x_D), _p(nn_D), _p(xdp)))
    return x_D, nn_D, xdp


def domain_bounds(domain, x_D, nn_D):
    nn_i, xa, xe = np.zeros(3, np.int64), np.zeros(3, np.int64), np.zeros(3, np.int64)
    check("dtc_domain_bounds", load().dtc_domain_bounds(int(domain), _p(_i64(*x_D)), _p(_i64(*nn_D)), _p(nn_i),
                                                        _p(xa), _p(xe)))
    return nn_i, xa, xe


def extract_block(vol, xa_n, xe_n):
    vol = _phi(vol)
    vdim = (vol.shape[2], vol.shape[1], vol.shape[0])
    n = [int(xe_n[d] - xa_n[d] + 1) for d in range(3)]
    out = np.zeros((n[2], n[1], n[0]), vol.dtype)
    check("dtc_extract_block", load().dtc_extract_block(_p(vol), KIND[vol.dtype], _p(_i64(*vdim)), _p(_i64(*xa_n)),
                                                        _p(_i64(*xe_n)), _p(out)))
    return out


def mesh_count(phi, llimit, nen=8):
    phi = _phi(phi)
    ne, nn = C.c_int64(), C.c_int64()
    rc = load().dtc_mesh_count(_p(phi), KIND[phi.dtype], _p(_i64(phi.shape[2], phi.shape[1], phi.shape[0])),
                               int(llimit), nen, C.byref(ne), C.byref(nn))
    check("dtc_mesh_count", rc, (DTC_OK, DTC_ERR_EMPTY_MESH))
    return ne.value, nn.value


def mesh(phi, llimit, nen=8, delta=(1.0, 1.0, 1.0), xa_n=(1, 1, 1)):
    """Returns (nodes (N_n, 3) float64, elems (N_e, nen) int64, 1-based ids)."""
    phi = _phi(phi)
    ne, nn = mesh_count(phi, llimit, nen)
    if ne == 0:
        return np.zeros((0, 3)), np.zeros((0, nen), np.int64)
    nodes = np.zeros((nn, 3))
    elems = np.zeros((ne, nen), np.int64)
    one, two = C.c_int64(), C.c_int64()
    check("dtc_mesh_build", load().dtc_mesh_build(
        _p(phi), KIND[phi.dtype], _p(_i64(phi.shape[2], phi.shape[1], phi.shape[0])), int(llimit), nen,
        _p(_f64(*delta)), _p(_i64(*xa_n)), ne, nn, _p(nodes), _p(elems), C.byref(one), C.byref(two)))
    return nodes, elems


def boundaries(nodes, min_c, max_c, dim_c, macro_order=1):
    """Returns (ids (n_b,), vals (R, 3*n_b)) like bnode_vals(lc, 3*(b-1)+d), R = 24
    (HEX08 macro element) or 60 (HEX20)."""
    nodes = np.ascontiguousarray(nodes, np.float64)
    nb = C.c_int64()
    lib = load()
    R = _lcs(macro_order)
    check("dtc_boundaries", lib.dtc_boundaries(_p(nodes), nodes.shape[0], _p(_f64(*min_c)), _p(_f64(*max_c)),
                                               _p(_f64(*dim_c)), macro_order, C.byref(nb), None, None))
    ids = np.zeros(nb.value, np.int64)
    vals = np.zeros(R * 3 * nb.value)
    check("dtc_boundaries", lib.dtc_boundaries(_p(nodes), nodes.shape[0], _p(_f64(*min_c)), _p(_f64(*max_c)),
                                               _p(_f64(*dim_c)), macro_order, C.byref(nb), _p(ids), _p(vals)))
    return ids, vals.reshape(3 * nb.value, R).T.copy()


def element_stiffness(nen, E, nu, a):
    nd = 3 * nen
    K = np.zeros(nd * nd)
    check("dtc_element_stiffness", load().dtc_element_stiffness(nen, float(E), float(nu), float(a), _p(K)))
    return K.reshape(nd, nd).T.copy()


@dataclass
class Domain:
    """One control volume: the voxel block and the parameters the reference reads
    from its .meta file (f-src_main/struct_process.f90:245-260)."""
    phi: np.ndarray                    # (nz, ny, nx) int8/int16/int32
    llimit: int = 1                    # BINARIZE_LO
    delta: Sequence[float] = (0.01, 0.01, 0.01)   # SPACING
    xa_n: Sequence[int] = (2, 2, 2)    # first voxel of the domain (1-based)
    size_domain: Optional[float] = None  # SIZE_DOMAIN; default n * delta
    E: float = 5600.0                  # YOUNG_MODULUS
    nu: float = 0.3                    # POISSON_RATIO
    rve_strain: float = 1e-6           # RVE_STRAIN
    modulus: Optional[np.ndarray] = None   # per-voxel modulus (nz, ny, nx) float64
    micro_nen: int = 8
    macro_order: int = 1
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n(self):
        return (self.phi.shape[2], self.phi.shape[1], self.phi.shape[0])

    def box(self):
        n = self.n
        mn = [(self.xa_n[d] - 1) * self.delta[d] for d in range(3)]
        mx = [(self.xa_n[d] + n[d] - 1) * self.delta[d] for d in range(3)]
        dim = [n[d] * self.delta[d] for d in range(3)]
        return mn, mx, dim

    def a(self):
        sd = self.size_domain if self.size_domain is not None else self.n[0] * self.delta[0]
        return sd * self.delta[0]

    def desc(self) -> DomainDesc:
        phi = _phi(self.phi)
        self._keep = [phi]
        d = DomainDesc()
        d.phi = phi.ctypes.data
        d.phi_kind = KIND[phi.dtype]
        d.micro_nen = self.micro_nen
        d.macro_order = self.macro_order
        d.n = (C.c_int64 * 3)(*self.n)
        d.llimit = int(self.llimit)
        d.xa_n = (C.c_int64 * 3)(*[int(v) for v in self.xa_n])
        d.delta = (C.c_double * 3)(*[float(v) for v in self.delta])
        d.size_domain = float(self.size_domain if self.size_domain is not None else self.n[0] * self.delta[0])
        d.E, d.nu, d.rve_strain = float(self.E), float(self.nu), float(self.rve_strain)
        if self.modulus is not None:
            m = np.ascontiguousarray(self.modulus, np.float64)
            if m.shape != phi.shape:
                raise ValueError("modulus must have the voxel array's shape")
            self._keep.append(m)
            d.modulus = m.ctypes.data
        else:
            d.modulus = None
        return d


def opts(precision=DTC_FP32_REFINE, rel_tol=1e-8, inner_tol=0.0, max_iter=200000, rotation=True, rot_step_deg=1.0,
         num_zero=1e-9, precond=DTC_PC_JACOBI, mg_degree=0) -> SolveOpts:
    o = SolveOpts()
    o.precision, o.rotation, o.rel_tol, o.inner_tol = int(precision), int(bool(rotation)), float(rel_tol), float(inner_tol)
    o.max_iter, o.rot_step_deg, o.num_zero = int(max_iter), float(rot_step_deg), float(num_zero)
    o.precond, o.mg_degree = int(precond), int(mg_degree)
    return o


def _lcs(macro_order):
    return 24 if macro_order == 1 else 60


def _result(r: DomainResult, macro_order=1) -> dict:
    a = lambda x, *s: np.array(x[:int(np.prod(s))], dtype=np.float64).reshape(*s)
    R = _lcs(macro_order)
    out = dict(status=r.status, precision_used=r.precision_used, n_elems=r.n_elems, n_nodes=r.n_nodes,
               n_bnodes=r.n_bnodes, iterations=r.iterations,
               outer_iterations=r.outer_iterations, relres=r.relres_max,
               ff=a(r.ff, R, R).T.copy(), K_macro=a(r.K_macro, R, R).T.copy(), C=a(r.C, 6, 6).T.copy(),
               EE=a(r.EE, 6, 6).T.copy(), density=r.density)
    out["rot"] = [dict(alpha=r.rot_alpha[c], n=np.array(r.rot_n[3 * c:3 * c + 3]),
                       aa=np.array(r.rot_aa[9 * c:9 * c + 9]).reshape(3, 3).T.copy(),
                       EE_opt=np.array(r.rot_EE[36 * c:36 * c + 36]).reshape(6, 6).T.copy(),
                       jj=r.rot_jj[c], crit=r.rot_crit[c]) for c in range(2)]
    return out


class Batch:
    """Domains resident in HBM; run() = mesh -> BCs -> solve -> fit -> rotation on the GPU."""

    def __init__(self, domains: Sequence[Domain], **kw):
        self.domains = list(domains)
        self._descs = (DomainDesc * len(self.domains))(*[d.desc() for d in self.domains])
        self._opts = opts(**kw)
        self._h = C.c_void_p()
        check("dtc_batch_create", load().dtc_batch_create(C.byref(self._descs), len(self.domains),
                                                          C.byref(self._opts), C.byref(self._h)))

    def run(self):
        res = (DomainResult * len(self.domains))()
        check("dtc_batch_run", load().dtc_batch_run(self._h, C.byref(res)))
        return [_result(r, d.macro_order) for r, d in zip(res, self.domains)]

    def grad(self, i, dL_dC):
        d = self.domains[i]
        G = np.ascontiguousarray(np.asarray(dL_dC, np.float64).T)      # column-major
        out = np.zeros(d.phi.size)
        check("dtc_batch_grad", load().dtc_batch_grad(self._h, int(i), _p(G), _p(out)))
        return out.reshape(d.phi.shape)

    def stats(self):
        avg, n, by, fl, run = C.c_double(), C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        check("dtc_batch_stats", load().dtc_batch_stats(self._h, C.byref(avg), C.byref(n), C.byref(by), C.byref(fl),
                                                        C.byref(run)))
        return dict(mv_avg_ms=avg.value, mv_launches=n.value, mv_bytes=by.value, mv_flops=fl.value, run_ms=run.value)

    def timing(self, enable=True):
        """Kernel timing (HIP events on the solve stream) of the following runs on/off."""
        check("dtc_batch_timing", load().dtc_batch_timing(self._h, int(bool(enable))))

    def kernel_stats(self, kernel):
        """Average launch duration (ms), launches, average algorithmic bytes and average
        dense element-product flops of one timed kernel (DTC_KSTAT_MV / _CHEB / _BRES /
        _RESID) over the last run."""
        avg, n, by, fl = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
        check("dtc_batch_kernel_stats", load().dtc_batch_kernel_stats(self._h, int(kernel), C.byref(avg), C.byref(n),
                                                                      C.byref(by)))
        check("dtc_batch_kernel_flops", load().dtc_batch_kernel_flops(self._h, int(kernel), C.byref(fl)))
        return dict(avg_ms=avg.value, launches=n.value, bytes=by.value, flops=fl.value)

    def close(self):
        if self._h:
            load().dtc_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def solve_domain(domain: Domain, want_u=False, want_r=False, **kw):
    d = domain.desc()
    o = opts(**kw)
    res = DomainResult()
    nn = None
    if want_u or want_r:
        _, nn = mesh_count(domain.phi, domain.llimit, domain.micro_nen)
    R = _lcs(domain.macro_order)
    u = np.zeros(3 * nn * R) if want_u else None
    r = np.zeros(3 * nn * R) if want_r else None
    check("dtc_solve_domain", load().dtc_solve_domain(C.byref(d), C.byref(o), C.byref(res),
                                                      _p(u) if u is not None else None,
                                                      _p(r) if r is not None else None))
    out = _result(res, domain.macro_order)
    if want_u:
        out["u"] = u.reshape(R, 3 * nn)         # u[lc, 3*(node-1)+d]
    if want_r:
        out["r"] = r.reshape(R, 3 * nn)
    return out


def solve_batch(domains: Sequence[Domain], **kw):
    descs = (DomainDesc * len(domains))(*[d.desc() for d in domains])
    o = opts(**kw)
    res = (DomainResult * len(domains))()
    check("dtc_solve_batch", load().dtc_solve_batch(C.byref(descs), len(domains), C.byref(o), C.byref(res)))
    return [_result(r, d.macro_order) for r, d in zip(res, domains)]


def rotation_search(EE, criterion, step_deg=1.0, num_zero=1e-9, noalloc=False):
    """CR_1 (criterion 1) / CR_2 (criterion 2) search.  noalloc: the entry point
    that allocates nothing (fixed per-GPU work space, bounded sweep size)."""
    EEc = np.ascontiguousarray(np.asarray(EE, np.float64).T)
    alpha, jj, crit = C.c_double(), C.c_int64(), C.c_double()
    n, aa, EEo = np.zeros(3), np.zeros(9), np.zeros(36)
    fn = "dtc_rotation_search_noalloc" if noalloc else "dtc_rotation_search"
    check(fn, getattr(load(), fn)(_p(EEc), int(criterion), float(step_deg), float(num_zero),
                                  C.byref(alpha), _p(n), _p(aa), _p(EEo), C.byref(jj), C.byref(crit)))
    return dict(alpha=alpha.value, n=n, aa=aa.reshape(3, 3).T.copy(), EE_opt=EEo.reshape(6, 6).T.copy(),
                jj=jj.value, crit=crit.value)
