"""CPU restatement of ngsxfem's node and dof marking (numpy).

TEST INFRASTRUCTURE ONLY: the checker for xg_nodes.hip.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use this module; the
product package never does.

Every function returns plain boolean / uint8 numpy arrays (one entry per
node, element or dof) and follows the reference loop by loop:

* node_domains       xfem/cutinfo.cpp:135-224 (cut_neighboring_node,
                     dom_of_node inside CutInformation::Update)
* elements_with_neighbor_facets   xfem/cutinfo.cpp:540-564
* elements_with_shared_vertex     xfem/cutinfo.cpp:566-596
* dofs_of_items      xfem/cutinfo.cpp:598-622 (GetDofsOfElements) and
                     :624-647 (GetDofsOfFacets)

Pinned in tests/test_nodes_cpu.py against brute-force set loops over tiny
meshes and closed-form counts on structured meshes.
"""
import numpy as np

NEG, POS, IF = 0, 1, 2


def dofs_of_items(item2dof, mark, ndof):
    """bit d set for every dof d of every marked item; negative dofs skipped"""
    out = np.zeros(ndof, bool)
    rows = item2dof[np.asarray(mark, bool)]
    d = rows[rows >= 0]
    out[d] = True
    return out


def elements_with_neighbor_facets(f2el, facet_mark, ne):
    """elements on either side of a marked facet"""
    return dofs_of_items(f2el, facet_mark, ne)


def elements_with_shared_vertex(el2v, elem_mark, nv):
    """elements sharing a vertex with a marked element"""
    vmark = dofs_of_items(el2v, elem_mark, nv)
    return vmark[el2v].any(axis=1)


def node_domains(el2node, dt, cut_ratio, nn, cutnb_prev=None):
    """(cut-neighbour flags, dom codes) of one node type.  el2node None: the
    element level (cells), where dom follows the cut ratio.  cutnb_prev: the
    flags of earlier updates (they are never cleared between updates)."""
    dt = np.asarray(dt)
    cut = np.zeros(nn, bool) if cutnb_prev is None else np.asarray(cutnb_prev, bool).copy()
    if el2node is None:
        cut |= dt == IF
        dom = np.where(np.asarray(cut_ratio) > 0.5, NEG, POS).astype(np.uint8)
        return cut, dom
    cut |= dofs_of_items(el2node, dt == IF, nn)
    neg = dofs_of_items(el2node, dt == NEG, nn)
    return cut, np.where(neg, NEG, POS).astype(np.uint8)


def bits_to_bool(words, n):
    """BitArray words (uint32, bit i of word i/32) -> bool[n]"""
    w = np.asarray(words).view(np.uint32)
    return ((w[np.arange(n) >> 5] >> (np.arange(n) & 31).astype(np.uint32)) & 1).astype(bool)


def bool_to_bits(b):
    """bool[n] -> BitArray words (uint32)"""
    b = np.asarray(b, bool)
    n = len(b)
    w = np.zeros((n + 31) // 32, np.uint32)
    idx = np.nonzero(b)[0]
    np.bitwise_or.at(w, idx >> 5, (np.uint32(1) << (idx & 31).astype(np.uint32)))
    return w
