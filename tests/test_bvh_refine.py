"""Leaf refinement of the mesh BVH at upload (no GPU).

mcrt_upload_scene splits every leaf of 3 to 8 triangles of the given tree
into leaves of at most 2 (mcrt_amd/csrc/bvh_refine.h): an exact SAH sweep over
the leaf's triangles, boxes widened by a scene-scale margin and clipped to the
given leaf box, triangles permuted within their leaf.  The render and trace
parity suites compare the GPU on the refined tree with the oracle on the
given one.  tests/bvh_refine_check.cpp (host compiler only) refines trees of
this library's builder over voxel quads, random and degenerate triangles and
checks the pre-order shape, leaf sizes and membership, box containment, that
no split leaves two children with equal boxes (which would split a voxel
face's two triangles for one more node step and no fewer tests), and
that closest hits through the refined tree (with the traversal's rule: slab
test on the query interval, entry at most best t * (1 + 2^-10), lowest
(t, original id)) equal those through the given tree, for rays aimed at
centroids, vertices and edges.  A leaf box moved off its triangles, and
refinement without the margin (new box planes through vertices, where the
slab test rounds), must both be caught.
"""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not shutil.which(CXX):
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("bvh_refine") / "bvh_refine_check")
    subprocess.run([CXX, "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "bvh_refine_check.cpp"),
                    os.path.join(ROOT, "mcrt_amd", "csrc", "bvh_build.cpp")], check=True, capture_output=True)
    return exe


def run(exe, seed, mode=None):
    out = subprocess.run([exe, str(seed)] + ([mode] if mode else []), check=True, capture_output=True,
                         text=True).stdout
    m = re.search(r"cases (\d+) fail (\d+) shape (\d+) contain (\d+) hits (\d+) dup (\d+) big (\d+) "
                  r"leaves (\d+) voxbig (\d+)", out)
    assert m, out
    return tuple(int(v) for v in m.groups())


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_refined_tree_is_valid_and_hits_match_the_given_tree(checker, seed):
    cases, fail, shape, contain, hits, dup, big, leaves, voxbig = run(checker, seed)
    assert cases == 12000 and hits > cases // 2
    assert fail == 0 and shape == 0 and contain == 0 and dup == 0
    assert voxbig == 0 and big * 100 <= leaves


def test_a_leaf_box_off_its_triangles_is_caught(checker):
    cases, fail, *_ = run(checker, 4, "plant")
    assert fail > 100


def test_refinement_without_the_margin_is_caught(checker):
    cases, fail, *_ = run(checker, 1, "nomargin")
    assert fail > 0
