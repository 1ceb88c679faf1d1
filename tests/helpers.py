"""Shared helpers of the test suite (case builders; the oracle is the checker)."""
import numpy as np


def domain_box(n, delta, xa):
    mn = [(xa[d] - 1) * delta for d in range(3)]
    mx = [(xa[d] + n[d] - 1) * delta for d in range(3)]
    dim = [n[d] * delta for d in range(3)]
    return mn, mx, dim


def random_mask(shape, p, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.int8)


def iso_voigt(E, nu):
    C = np.zeros((6, 6))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


# ---------------------------------------------------------------- puredat
def read_puredat_head(path):
    """The tree of a puredat .head file (f-src/mod_puredat.f90:4646-4700): a list
    of branches in file order, each {desc, no_branches, no_leaves, streams
    (dict of size_* or None), leaves: [{desc, no, type, lb, ub}]}."""
    branches, cur, leaf = [], None, None
    for line in open(path).read().splitlines():
        if line.startswith("<==branch==>"):
            cur = dict(desc=None, no_branches=None, no_leaves=None, streams=None, leaves=[])
            branches.append(cur)
            leaf = None
        elif line.startswith("<--leaf-->"):
            leaf = dict(desc=None, no=None, type=None, lb=None, ub=None)
            cur["leaves"].append(leaf)
        else:
            key, _, val = line.partition(">")
            key, val = key.lstrip("<").strip(), val.strip()
            tgt = leaf if leaf is not None else cur
            if key == "description":
                tgt["desc"] = val.strip("'")
            elif key == "no_of_branches":
                cur["no_branches"] = int(val)
            elif key == "no_of_leaves":
                cur["no_leaves"] = int(val)
            elif key == "streams_allocated":
                cur["streams"] = {} if val == "T" else None
            elif key.startswith("size_"):
                cur["streams"][key] = int(val)
            elif key == "no_of_data":
                leaf["no"] = int(val)
            elif key == "type_of_data":
                leaf["type"] = int(val)
            elif key == "lower_bound":
                leaf["lb"] = int(val)
            elif key == "upper_bound":
                leaf["ub"] = int(val)
    return branches


_PD_STREAM = {4: (".int8.st", np.int64), 5: (".real8.st", np.float64), 6: (".char.st", np.uint8)}


def read_puredat_leaf(stem, leaf):
    """The data of one leaf from the typed stream files <stem>.int8.st / .real8.st /
    .char.st (elements lb..ub, 1-based)."""
    suf, dt = _PD_STREAM[leaf["type"]]
    a = np.fromfile(stem + suf, dt, count=leaf["ub"], offset=0)
    v = a[leaf["lb"] - 1:leaf["ub"]]
    return v.tobytes().decode() if leaf["type"] == 6 else v
