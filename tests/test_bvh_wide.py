"""4-wide quantized nodes of the culling-stack traversal (no GPU).

Scenes whose traversal set spills the L2 traverse the mesh tree as 4-wide
48-B nodes (mcrt_amd/csrc/bvh_wide.h, decoded by dev_common.h
node_step_wide): each node holds up to 4 children on one quantized grid, the
interior children of a node take consecutive indices and the leaf records
are reordered so that the leaves of a node follow one another.
tests/bvh_wide_check.cpp (built with hipcc as a host program) builds refined
trees of this library's builder over voxel quads, random and degenerate
triangles, converts them to binary pairs with shuffled indices, and checks
the record order (a permutation; every leaf keeps its records), the decoded
boxes (each contains its binary box), the references, the stack bound, that
no unused slot passes the slab test (axis-parallel rays included), and
that an exact emulation of the kernel's wide node step finds the closest
hits of the binary traversal wherever that agrees with brute force.  A leaf
child's record offset moved by one must be caught.
"""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("bvh_wide") / "bvh_wide_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-o", exe,
                    os.path.join(HERE, "bvh_wide_check.cpp"), os.path.join(ROOT, "mcrt_amd", "csrc", "bvh_build.cpp")],
                   check=True, capture_output=True)
    return exe


def run(exe, seed, plant=False):
    out = subprocess.run([exe, str(seed)] + (["plant"] if plant else []), check=True, capture_output=True,
                         text=True).stdout
    m = re.search(r"cases (\d+) fail (\d+) diff (\d+) diffq (\d+) order (\d+) box (\d+) ref (\d+) stack (\d+) "
                  r"nodes (\d+) full (\d+) empty (\d+)", out)
    assert m, out
    return dict(zip(["cases", "fail", "diff", "diffq", "order", "box", "ref", "stack", "nodes", "full", "empty"],
                    (int(v) for v in m.groups())))


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_wide_tree_is_valid_and_hits_match_the_binary_tree(checker, seed):
    r = run(checker, seed)
    assert r["cases"] == 12000
    assert r["fail"] == 0 and r["order"] == 0 and r["box"] == 0 and r["ref"] == 0 and r["stack"] == 0
    assert r["empty"] == 0  # an unused slot (a point box) never passes the slab test
    # rays aimed exactly at vertices and edges: the slab test rounds at box
    # corners, so a looser box may find a hit the exact boxes miss, but rarely
    assert r["diff"] * 50 <= r["cases"]
    # most nodes hold 4 children (those with 2 are mostly at the fringe of a
    # balanced subtree, where a path crosses the same number of levels)
    assert r["full"] * 5 >= r["nodes"] * 2


def test_a_moved_leaf_offset_is_caught(checker):
    r = run(checker, 1, plant=True)
    assert r["fail"] > 500
