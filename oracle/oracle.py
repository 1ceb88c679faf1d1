"""ctypes wrapper of the CPU FP64 oracle (oracle/dtc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product library
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libdtc_oracle.so")
# DTC_ORACLE_FAST=1 (the CPU baseline's timing runs): the -O3 build of the same source
if os.environ.get("DTC_ORACLE_FAST") == "1":
    _LIB_PATH = os.path.join(_HERE, "_build", "libdtc_oracle_fast.so")
_lib = None

KIND = {np.dtype(np.int8): 1, np.dtype(np.int16): 2, np.dtype(np.int32): 4}


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.dtco_density.restype = C.c_double
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i64(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.int64).reshape(-1))


def _f64(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float64).reshape(-1))


class OracleError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} returned {rc}")
        self.rc = rc


def _chk(fn, rc):
    if rc != 0:
        raise OracleError(fn, rc)


def ddc_params(size_domain, delta, vdim):
    x_D = np.zeros(3, np.int64); nn_D = np.zeros(3, np.int64); xdp = np.zeros(3)
    _chk("ddc_params", lib().dtco_ddc_params(_p(_f64(*size_domain)), _p(_f64(*delta)), _p(_i64(*vdim)),
                                             _p(x_D), _p(nn_D), _p(xdp)))
    return x_D, nn_D, xdp


def domain_bounds(domain, x_D, nn_D):
    nn_i = np.zeros(3, np.int64); xa = np.zeros(3, np.int64); xe = np.zeros(3, np.int64)
    _chk("domain_bounds", lib().dtco_domain_bounds(C.c_int64(domain), _p(_i64(*x_D)), _p(_i64(*nn_D)),
                                                   _p(nn_i), _p(xa), _p(xe)))
    return nn_i, xa, xe


def extract_block(vol, xa_n, xe_n):
    """vol: numpy (nz, ny, nx) signed int array; returns the sub-block (nz', ny', nx')."""
    vol = np.ascontiguousarray(vol)
    vdim = (vol.shape[2], vol.shape[1], vol.shape[0])
    n = [int(xe_n[d] - xa_n[d] + 1) for d in range(3)]
    out = np.zeros((n[2], n[1], n[0]), vol.dtype)
    _chk("extract_block", lib().dtco_extract_block(_p(vol), KIND[vol.dtype], _p(_i64(*vdim)),
                                                   _p(_i64(*xa_n)), _p(_i64(*xe_n)), _p(out)))
    return out


def mesh(phi, llimit, nen=8, delta=(1.0, 1.0, 1.0), xa_n=(1, 1, 1)):
    """phi: numpy (nz, ny, nx) int8/int16/int32 (x fastest).
    Returns (nodes (N_n,3) float64, elems (N_e,nen) int64 1-based)."""
    phi = np.ascontiguousarray(phi)
    n = _i64(phi.shape[2], phi.shape[1], phi.shape[0])
    nn = C.c_int64(); ne = C.c_int64(); pn = C.c_void_p(); pe = C.c_void_p()
    rc = lib().dtco_mesh(_p(phi), KIND[phi.dtype], _p(n), C.c_int64(int(llimit)), C.c_int(nen),
                         _p(_f64(*delta)), _p(_i64(*xa_n)), C.byref(nn), C.byref(ne), C.byref(pn), C.byref(pe))
    if rc == 1:
        return np.zeros((0, 3)), np.zeros((0, nen), np.int64)
    _chk("mesh", rc)
    nodes = np.ctypeslib.as_array(C.cast(pn, C.POINTER(C.c_double)), shape=(nn.value, 3)).copy()
    elems = np.ctypeslib.as_array(C.cast(pe, C.POINTER(C.c_int64)), shape=(ne.value, nen)).copy()
    lib().dtco_free(pn); lib().dtco_free(pe)
    return nodes, elems


def boundaries(nodes, min_c, max_c, dim_c, macro_order=1):
    nodes = np.ascontiguousarray(nodes, np.float64)
    nb = C.c_int64(); pi = C.c_void_p(); pv = C.c_void_p()
    _chk("boundaries", lib().dtco_boundaries(_p(nodes), C.c_int64(nodes.shape[0]), _p(_f64(*min_c)),
                                             _p(_f64(*max_c)), _p(_f64(*dim_c)), C.c_int(macro_order),
                                             C.byref(nb), C.byref(pi), C.byref(pv)))
    R = 24 if macro_order == 1 else 60
    ids = np.ctypeslib.as_array(C.cast(pi, C.POINTER(C.c_int64)), shape=(nb.value,)).copy()
    vals = np.ctypeslib.as_array(C.cast(pv, C.POINTER(C.c_double)), shape=(3 * nb.value, R)).copy()
    lib().dtco_free(pi); lib().dtco_free(pv)
    return ids, vals.T.copy()          # vals[lc, 3*b+d]


def macro_shape(x, min_c, max_c, macro_order=1):
    N = np.zeros(8 if macro_order == 1 else 20)
    _chk("macro_shape", lib().dtco_macro_shape(C.c_int(macro_order), _p(_f64(*x)), _p(_f64(*min_c)),
                                               _p(_f64(*max_c)), _p(N)))
    return N


def uu(macro_order=1):
    nen = 8 if macro_order == 1 else 20
    out = np.zeros(nen * 3 * 3 * nen)
    _chk("uu", lib().dtco_uu(C.c_int(macro_order), _p(out)))
    return out.reshape(3 * nen, 3, nen).transpose(2, 1, 0).copy()   # uu[l, d, lc]


def vv_lin(rve_strain):
    out = np.zeros(24 * 24)
    _chk("vv", lib().dtco_vv_lin(C.c_double(rve_strain), _p(out)))
    return out.reshape(24, 24).T.copy()                      # vv[dof, lc]


def vv_quad_consistent(rve_strain):
    out = np.zeros(60 * 60)
    _chk("vv", lib().dtco_vv_quad_consistent(C.c_double(rve_strain), _p(out)))
    return out.reshape(60, 60).T.copy()                      # vv[dof, lc]


def ke(nen, E, nu, a):
    nd = 3 * nen
    K = np.zeros(nd * nd)
    _chk("ke", lib().dtco_ke(C.c_int(nen), C.c_double(E), C.c_double(nu), C.c_double(a), _p(K)))
    return K.reshape(nd, nd).T.copy()


def solve_domain(nodes, elems, E, nu, a, min_c, max_c, dim_c, tol=1e-10, max_it=100000,
                 norm_type=0, emod=None, lc_first=0, lc_count=None, want_u=False, want_r=False,
                 macro_order=1):
    nodes = np.ascontiguousarray(nodes, np.float64)
    elems = np.ascontiguousarray(elems, np.int64)
    nn, nen = nodes.shape[0], elems.shape[1]
    m = 3 * nn
    R = 24 if macro_order == 1 else 60
    if lc_count is None:
        lc_count = R - lc_first
    u = np.zeros((R, m)) if want_u else None
    r = np.zeros((R, m)) if want_r else None
    ff = np.zeros((R, R))
    iters = np.zeros(R, np.int64); relres = np.zeros(R)
    em = None if emod is None else np.ascontiguousarray(emod, np.float64)
    _chk("solve_domain", lib().dtco_solve_domain(
        _p(nodes), C.c_int64(nn), _p(elems), C.c_int64(elems.shape[0]), C.c_int(nen),
        _p(em) if em is not None else None, C.c_double(E), C.c_double(nu), C.c_double(a),
        _p(_f64(*min_c)), _p(_f64(*max_c)), _p(_f64(*dim_c)), C.c_int(macro_order), C.c_double(tol),
        C.c_int64(max_it), C.c_int(norm_type), C.c_int(lc_first), C.c_int(lc_count),
        _p(u) if u is not None else None, _p(r) if r is not None else None, _p(ff), _p(iters), _p(relres)))
    # ff is column-major (R,R): element [a + R*lc] -> array row lc, col a
    return dict(ff=ff.T.copy(), iters=iters, relres=relres, u=u, r=r)


def effective(ff, x_D_phy, rve_strain=1e-6, macro_order=1):
    ffc = np.ascontiguousarray(np.asarray(ff, np.float64).T)      # column-major
    R = ffc.shape[0]
    KM = np.zeros(R * R); Cm = np.zeros(36); EE = np.zeros(36)
    _chk("effective", lib().dtco_effective(_p(ffc), C.c_int(macro_order), C.c_double(rve_strain),
                                           _p(_f64(*x_D_phy)), _p(KM), _p(Cm), _p(EE)))
    return KM.reshape(R, R).T.copy(), Cm.reshape(6, 6).T.copy(), EE.reshape(6, 6).T.copy()


def density(n_elems, x_D_phy, delta):
    return lib().dtco_density(C.c_int64(n_elems), _p(_f64(*x_D_phy)), _p(_f64(*delta)))


def rotation_search(EE, criterion, step_deg=1.0, num_zero=1e-9):
    EEc = np.ascontiguousarray(np.asarray(EE, np.float64).T)
    alpha = C.c_double(); n = np.zeros(3); aa = np.zeros(9); EEo = np.zeros(36)
    jj = C.c_int64(); crit = C.c_double()
    _chk("rotation_search", lib().dtco_rotation_search(
        _p(EEc), C.c_int(criterion), C.c_double(step_deg), C.c_double(num_zero), C.byref(alpha), _p(n),
        _p(aa), _p(EEo), C.byref(jj), C.byref(crit)))
    return dict(alpha=alpha.value, n=n, aa=aa.reshape(3, 3).T.copy(), EE_opt=EEo.reshape(6, 6).T.copy(),
                jj=jj.value, crit=crit.value)


def rotate66(EE, n, alpha):
    EEc = np.ascontiguousarray(np.asarray(EE, np.float64).T)
    T = np.zeros(36)
    _chk("rotate66", lib().dtco_rotate66(_p(EEc), _p(_f64(*n)), C.c_double(alpha), _p(T)))
    return T.reshape(6, 6).T.copy()


def grad_modulus(nodes, elems, nu, a, min_c, max_c, x_D_phy, u, G, rve_strain=1e-6, macro_order=1):
    nodes = np.ascontiguousarray(nodes, np.float64)
    elems = np.ascontiguousarray(elems, np.int64)
    uc = np.ascontiguousarray(u, np.float64)          # (24 or 60, 3N)
    Gc = np.ascontiguousarray(np.asarray(G, np.float64).T)
    out = np.zeros(elems.shape[0])
    _chk("grad_modulus", lib().dtco_grad_modulus(
        _p(nodes), C.c_int64(nodes.shape[0]), _p(elems), C.c_int64(elems.shape[0]), C.c_int(elems.shape[1]),
        C.c_int(macro_order), C.c_double(nu),
        C.c_double(a), _p(_f64(*min_c)), _p(_f64(*max_c)), _p(_f64(*x_D_phy)), C.c_double(rve_strain),
        _p(uc), _p(Gc), _p(out)))
    return out
