"""The puredat .head layout rules that the Fortran driver writes and the test
helpers read, checked on the reference's own output header
(tests/golden/solid_cube_Hex08.head, a copy of the data file
regression/solid_cube/solid_cube_Hex08/solid_cube_Hex08.head): branch and leaf
records as f-src/mod_puredat.f90:4646-4700 writes them, and leaf bounds numbered
per stream type in the order set_bounds_in_branch visits them (child branches
first, then the branch's own leaves, f-src/mod_puredat.f90:2007-2035), the
stream sizes being the totals."""
import os

from helpers import read_puredat_head

HEAD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "solid_cube_Hex08.head")


def test_reference_head_parses():
    tree = read_puredat_head(HEAD)
    assert [b["desc"] for b in tree] == ["solid_cube_Hex08", "Input parameters", "Global domain decomposition",
                                         "Averaged Material Properties"]
    assert tree[0]["no_branches"] == 3 and tree[0]["no_leaves"] == 1 and tree[0]["streams"] is not None
    assert [b["no_leaves"] for b in tree] == [len(b["leaves"]) for b in tree] == [1, 15, 8, 18]
    assert all(b["streams"] is None for b in tree[1:])


def test_reference_bounds_follow_the_visit_order():
    tree = read_puredat_head(HEAD)
    nxt = {}
    # depth-1 tree: the children's leaves in order, then the root's own leaves
    for leaf in [l for b in tree[1:] for l in b["leaves"]] + tree[0]["leaves"]:
        assert leaf["lb"] == nxt.get(leaf["type"], 1), leaf
        assert leaf["ub"] == leaf["lb"] + leaf["no"] - 1, leaf
        nxt[leaf["type"]] = leaf["ub"] + 1
    st = tree[0]["streams"]
    names = {1: "size_int1_stream", 2: "size_int2_stream", 3: "size_int4_stream", 4: "size_int8_stream",
             5: "size_real_stream", 6: "size_char_stream", 7: "size_log_stream"}
    for t, name in names.items():
        assert st[name] == nxt.get(t, 1) - 1, name
