"""BVH builders on the GPU (APH_BUILDER_LBVH, APH_BUILDER_SAH): structural
invariants of the exported trees and their SAH quality.

Pinned by construction-independent properties:
  * every triangle sits in exactly one leaf and the leaf-ordered triangle array
    is a permutation of the mesh's triangles with their vertices;
  * every child box contains its subtree (triangle vertices and child boxes);
  * the reported depth bounds every root-to-leaf path;
  * SAH quality: the binned-SAH tree's cost is at most 1.10 x the cost of a
    full-sweep SAH tree built on the CPU by the oracle, and below the LBVH's.
Traversal results for both builders are checked against the oracle in
test_gpu_trace.py.
"""
import numpy as np
import pytest

import pyoracle
from aurora_amd import capi, scenes

pytestmark = pytest.mark.gpu

COST_NODE, COST_TRI, SAH_MAX_LEAF = 1.0, 1.2, 4
LEAF_BIT = 0x80000000


def _ref(f):
    return np.float32(f).view(np.uint32).item()


def walk(nodes):
    """Yields (node index, depth) of internal nodes and (slot box lo, hi, ref, depth) of leaves."""
    inner, leaves = [], []
    stack = [(0, 1)]
    while stack:
        ni, d = stack.pop()
        nd = nodes[ni]
        inner.append((ni, d))
        for slot in range(2):
            lo = nd[0:3] if slot == 0 else nd[8:11]
            hi = nd[4:7] if slot == 0 else nd[12:15]
            ref = _ref(nd[3] if slot == 0 else nd[7])
            if lo[0] > hi[0]:  # inverted box: empty slot
                continue
            if ref & LEAF_BIT:
                leaves.append((lo, hi, ref, d))
            else:
                stack.append((ref, d + 1))
    return inner, leaves


def subtree_boxes(nodes, tris):
    """Checks containment bottom-up; returns the exact bounds of each child slot's content."""
    def content(ref):
        if ref & LEAF_BIT:
            first, cnt = ref & 0x07FFFFFF, ((ref >> 27) & 0xF) + 1
            v = tris[first:first + cnt].reshape(-1, 3, 4)[:, :, :3].reshape(-1, 3)
            return v.min(0), v.max(0)
        nd = nodes[ref]
        los, his = [], []
        for slot in range(2):
            lo = nd[0:3] if slot == 0 else nd[8:11]
            hi = nd[4:7] if slot == 0 else nd[12:15]
            r = _ref(nd[3] if slot == 0 else nd[7])
            if lo[0] > hi[0]:
                continue
            clo, chi = content(r)
            assert (lo <= clo).all() and (hi >= chi).all(), f"node {ref} slot {slot} does not contain its subtree"
            los.append(lo)
            his.append(hi)
        return np.min(los, 0), np.max(his, 0)
    return content(0)


def area(lo, hi):
    d = np.maximum(np.asarray(hi, np.float64) - np.asarray(lo, np.float64), 0.0)
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def sah_cost(nodes):
    inner, leaves = walk(nodes)
    root = nodes[0]
    rlo = np.minimum(root[0:3], root[8:11])
    rhi = np.maximum(root[4:7], root[12:15])
    ra = area(rlo, rhi)
    cost = 0.0
    for ni, _ in inner:
        nd = nodes[ni]
        cost += COST_NODE * area(np.minimum(nd[0:3], nd[8:11]), np.maximum(nd[4:7], nd[12:15])) / ra
    for lo, hi, ref, _ in leaves:
        cost += COST_TRI * (((ref >> 27) & 0xF) + 1) * area(lo, hi) / ra
    return cost


def meshes(teapot_path):
    out = []
    tp = scenes.load_teapot(teapot_path)
    out.append(("teapot", tp[0].reshape(-1, 3), tp[2].reshape(-1, 3)))
    p, f = scenes.icosphere(4)
    out.append(("icosphere4", p, f))
    p, _, f = scenes.displaced_icosphere(4, np.random.default_rng(3))
    out.append(("displaced4", p, f))
    rng = np.random.default_rng(5)  # triangle soup: long thin and duplicated triangles
    soup = rng.uniform(-5, 5, (3000, 3)).astype(np.float32)
    idx = rng.integers(0, 3000, (4000, 3)).astype(np.uint32)
    idx[3000:] = idx[:1000]
    out.append(("soup", soup, idx))
    return out


@pytest.mark.parametrize("builder", [capi.APH_BUILDER_LBVH, capi.APH_BUILDER_SAH])
def test_bvh_invariants(renderer, teapot_path, builder):
    for name, pos, idx in meshes(teapot_path):
        m = renderer.mesh(pos, idx)
        renderer.build_blas(m, builder)
        nodes, tris, depth = renderer.blas_export(m)
        n = len(idx)
        assert len(nodes) <= max(n - 1, 1)
        inner, leaves = walk(nodes)
        assert len(inner) == len(nodes), f"{name}: unreachable or shared nodes"
        covered = np.zeros(n, dtype=np.int32)
        for lo, hi, ref, d in leaves:
            first, cnt = ref & 0x07FFFFFF, ((ref >> 27) & 0xF) + 1
            assert cnt <= (1 if builder == capi.APH_BUILDER_LBVH else SAH_MAX_LEAF)
            covered[first:first + cnt] += 1
            assert d <= depth, f"{name}: leaf at depth {d} > reported {depth}"
        assert (covered == 1).all(), f"{name}: leaf ranges do not partition the triangles"
        prim = tris[:, 3].view(np.uint32)
        assert np.array_equal(np.sort(prim), np.arange(n, dtype=np.uint32))
        verts = tris.reshape(n, 3, 4)[:, :, :3]
        np.testing.assert_array_equal(verts, pos.reshape(-1, 3)[idx[prim].astype(np.int64)])
        subtree_boxes(nodes, tris)


def test_sah_quality_against_full_sweep(renderer, teapot_path):
    lib = pyoracle.load()
    for name, pos, idx in meshes(teapot_path):
        v = pos.reshape(-1, 3)[idx.astype(np.int64)]
        lo = np.ascontiguousarray(v.min(1), dtype=np.float32)
        hi = np.ascontiguousarray(v.max(1), dtype=np.float32)
        ref = lib.ora_sah_sweep_cost(pyoracle._p(lo), pyoracle._p(hi), len(idx), SAH_MAX_LEAF, COST_NODE, COST_TRI)
        costs = {}
        for builder in (capi.APH_BUILDER_LBVH, capi.APH_BUILDER_SAH):
            m = renderer.mesh(pos, idx)
            renderer.build_blas(m, builder)
            costs[builder] = sah_cost(renderer.blas_export(m)[0])
        assert costs[capi.APH_BUILDER_SAH] <= 1.10 * ref, (name, costs, ref)
        assert costs[capi.APH_BUILDER_SAH] < costs[capi.APH_BUILDER_LBVH], (name, costs)
