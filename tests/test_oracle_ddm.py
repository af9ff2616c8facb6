"""Pins of the oracle's Direct Delta Mush functions (CPU only): the known
answers of DirectDeltaMush.lean:115-165, Features/KnownKnowns/DirectDeltaMush.lean
and DirectDeltaMushBind.lean:169-262, plus closed forms (rotation, linearity,
the dense smoothing operator, argsort on distinct values) and planted
mistakes that those checks must reject."""
import math

import numpy as np
import pytest

import oracle as O

I4 = np.eye(4).reshape(16)


def _tr(tx, ty, tz):
    M = np.eye(4)
    M[:3, 3] = (tx, ty, tz)
    return M.reshape(16)


def _lbs1(T, infl, p):
    """lbsMatvec of one vertex: infl = [(handle, weight), ...]."""
    start = np.array([0, len(infl)], np.int32)
    idx = np.array([i for i, _ in infl], np.int32)
    w = np.array([x for _, x in infl], np.float64)
    return O.lbs_matvec(np.asarray(p, float).reshape(1, 3), np.asarray(T).reshape(-1, 16),
                        start, idx, w)[0]


def _close(a, b, eps):                              # DirectDeltaMush.vclose
    return all(abs(x - y) < eps for x, y in zip(a, b))


# ---- runtime matvec (DirectDeltaMush.lean:115-165) --------------------------
def test_lbs_kats():
    assert _close(_lbs1([I4], [(0, 1.0)], (1.5, -2.0, 0.75)), (1.5, -2.0, 0.75), 1e-12)
    assert _close(_lbs1([_tr(2.0, -3.0, 1.0)], [(0, 1.0)], (1, 1, 1)), (3.0, -2.0, 2.0), 1e-12)
    assert _close(_lbs1([I4, I4], [(0, 0.5), (1, 0.5)], (4.0, -1.0, 2.0)), (4.0, -1.0, 2.0),
                  1e-12)
    assert _close(_lbs1([_tr(1, 0, 0), _tr(-1, 0, 0)], [(0, 0.7), (1, 0.3)], (0, 0, 0)),
                  (0.4, 0.0, 0.0), 1e-12)
    assert _close(_lbs1([I4], [], (1, 1, 1)), (0.0, 0.0, 0.0), 1e-12)
    assert _close(_lbs1([I4] * 4, [(i, 0.25) for i in range(4)], (7.0, 8.0, 9.0)),
                  (7.0, 8.0, 9.0), 1e-12)


def test_apply_transform_rotation_and_translation():
    c, s = math.cos(0.3), math.sin(0.3)
    M = np.eye(4)
    M[:2, :2] = [[c, -s], [s, c]]
    M[:3, 3] = (0.5, -1.0, 2.0)
    p = np.array([1.25, -0.5, 3.0])
    assert np.allclose(O.apply_transform4x4(M.reshape(16), *p), M[:3, :3] @ p + M[:3, 3],
                       atol=1e-14)


def test_lbs_linear_in_weights_and_out_of_range_handle():
    rng = np.random.default_rng(1)
    nv, nt = 50, 7
    rest = rng.normal(size=(nv, 3))
    T = rng.normal(size=(nt, 16))
    start = np.arange(0, 4 * nv + 1, 4, dtype=np.int32)
    idx = rng.integers(0, nt, size=4 * nv).astype(np.int32)
    w1, w2 = rng.normal(size=4 * nv), rng.normal(size=4 * nv)
    a = O.lbs_matvec(rest, T, start, idx, w1) + O.lbs_matvec(rest, T, start, idx, w2)
    assert np.allclose(O.lbs_matvec(rest, T, start, idx, w1 + w2), a, atol=1e-12)
    # per vertex: sum_q w_q (M_q [p; 1])[:3]
    ref = np.zeros((nv, 3))
    for v in range(nv):
        for q in range(start[v], start[v + 1]):
            M = T[idx[q]].reshape(4, 4)
            ref[v] += w1[q] * (M[:3, :3] @ rest[v] + M[:3, 3])
    assert np.allclose(O.lbs_matvec(rest, T, start, idx, w1), ref, atol=1e-12)
    # an index past the transforms reads the zero matrix (Array default)
    idx2 = idx.copy()
    idx2[0] = nt + 3
    out = O.lbs_matvec(rest, T, start, idx2, w1)
    M = T[idx[0]].reshape(4, 4)
    assert np.allclose(out[0], ref[0] - w1[0] * (M[:3, :3] @ rest[0] + M[:3, 3]), atol=1e-12)


def test_lbs_planted_transpose_fails_kat():
    """Reading T column-major (a plausible mistake) breaks the translation KAT."""
    T = _tr(2.0, -3.0, 1.0).reshape(4, 4).T.reshape(16)
    assert not _close(_lbs1([T], [(0, 1.0)], (1, 1, 1)), (3.0, -2.0, 2.0), 1e-12)


# ---- smoothWeights (DirectDeltaMushBind.lean:169-197) ------------------------
PATH4 = (np.array([0, 1, 3, 5, 6], np.int32), np.array([1, 0, 2, 1, 3, 2], np.int32))
IDW = np.eye(4)
HALF = np.array([[0.5, 0.5, 0.0, 0.0]] * 4)


def _pou(W, eps):
    return bool(np.all(np.abs(W.sum(axis=1) - 1.0) < eps))


def test_smooth_kats():
    assert _pou(IDW, 1e-12)
    W0 = O.smooth_weights(IDW, *PATH4, 0, 0.5)
    assert abs(W0[0, 0] - 1.0) < 1e-12 and abs(W0[1, 1] - 1.0) < 1e-12
    assert _pou(HALF, 1e-12)
    W5 = O.smooth_weights(HALF, *PATH4, 5, 0.5)
    assert _pou(W5, 1e-12)
    assert abs(W5[0, 0] - 0.5) < 1e-12 and abs(W5[2, 1] - 0.5) < 1e-12


def _random_graph(n, rng, isolated=()):
    nb = [set() for _ in range(n)]
    for _ in range(3 * n):
        a, b = rng.integers(0, n, size=2)
        if a != b and a not in isolated and b not in isolated:
            nb[a].add(b)
            nb[b].add(a)
    ptr = np.zeros(n + 1, np.int32)
    adj = []
    for v in range(n):
        lst = list(nb[v])
        rng.shuffle(lst)                              # adjacency order is the caller's
        adj += lst
        ptr[v + 1] = len(adj)
    return ptr, np.array(adj, np.int32)


@pytest.mark.parametrize("omega", [0.5, 0.3, 0.9])
def test_smooth_equals_dense_operator(omega):
    """W_nu = ((1 - omega) I + omega D^-1 A)^nu W, rows without neighbours
    scaled by (1 - omega)."""
    rng = np.random.default_rng(7)
    n, nh, nu = 40, 6, 4
    ptr, adj = _random_graph(n, rng, isolated=(5,))
    W = rng.random((n, nh))
    M = (1.0 - omega) * np.eye(n)
    for v in range(n):
        c = ptr[v + 1] - ptr[v]
        for u in adj[ptr[v]:ptr[v + 1]]:
            M[v, u] += omega / c
    ref = np.linalg.matrix_power(M, nu) @ W
    assert np.allclose(O.smooth_weights(W, ptr, adj, nu, omega), ref, rtol=1e-13, atol=1e-14)
    if omega != 0.5:
        # planted: omega and 1 - omega swapped
        bad = O.smooth_weights(W, ptr, adj, nu, 1.0 - omega)
        assert not np.allclose(bad, ref, rtol=1e-13, atol=1e-14)


def test_smooth_out_of_range_neighbour_reads_zero():
    W = np.array([[1.0, 2.0], [3.0, 4.0]])
    ptr = np.array([0, 2, 3], np.int32)
    adj = np.array([1, 9, 0], np.int32)             # vertex 0: neighbours 1 and (invalid) 9
    out = O.smooth_weights(W, ptr, adj, 1, 0.5)
    assert np.array_equal(out[0], 0.5 * W[0] + 0.5 * ((0.0 + W[1] + 0.0) / 2))
    assert np.array_equal(out[1], 0.5 * W[1] + 0.5 * W[0])


# ---- sparsifyTopK (DirectDeltaMushBind.lean:199-262) -------------------------
def test_topk_kats():
    idx, w = O.sparsify_topk(np.array([[0.4, 0.3, 0.2, 0.1]]), 2)
    assert idx.shape == (1, 2)
    assert abs(w[0].sum() - 1.0) < 1e-12
    assert abs(w[0, 0] - 4.0 / 7.0) < 1e-12 and abs(w[0, 1] - 3.0 / 7.0) < 1e-12
    assert list(idx[0]) == [0, 1]
    idx, w = O.sparsify_topk(IDW, 4)
    assert abs(w[0].sum() - 1.0) < 1e-12


def _dense(idx, w, cols):                           # sparseToWeightMatrix (:230-236)
    out = np.zeros((idx.shape[0], cols))
    for v in range(idx.shape[0]):
        for j, x in zip(idx[v], w[v]):
            if j < cols:
                out[v, j] = x
    return out


def test_smooth_then_topk_partition_of_unity():     # :238-262
    W1 = O.smooth_weights(IDW, *PATH4, 2, 0.5)
    assert _pou(_dense(*O.sparsify_topk(W1, 2), 4), 1e-9)
    assert _pou(_dense(*O.sparsify_topk(W1, 4), 4), 1e-9)
    W1 = O.smooth_weights(HALF, *PATH4, 1, 0.3)
    assert _pou(_dense(*O.sparsify_topk(W1, 1), 2), 1e-9)


def test_topk_swap_order_decides_ties():
    """[1, 1, 3], K = 2: step 0 swaps 3 into place, moving index 0 to
    position 2; step 1 then meets index 1 first, so the kept pairs are
    (2, 1) and not the lowest-index tie (2, 0)."""
    idx, w = O.sparsify_topk(np.array([[1.0, 1.0, 3.0]]), 2)
    assert list(idx[0]) == [2, 1]
    assert np.allclose(w[0], [0.75, 0.25])
    # [1, 2, 2, 1], K = 3 -> positions pick 1, 2, then index 0 (moved to 2)
    idx, _ = O.sparsify_topk(np.array([[1.0, 2.0, 2.0, 1.0]]), 3)
    assert list(idx[0]) == [1, 2, 0]
    # |w| decides, the sign is kept; a zero kept sum leaves the weights as is
    idx, w = O.sparsify_topk(np.array([[0.1, -0.5, 0.5, 0.2]]), 2)
    assert list(idx[0]) == [1, 2] and list(w[0]) == [-0.5, 0.5]


def test_topk_matches_argsort_on_distinct_values():
    rng = np.random.default_rng(3)
    W = rng.normal(size=(30, 17))
    for K in (1, 4, 8, 17, 25):
        idx, w = O.sparsify_topk(W, K)
        take = min(K, 17)
        ref = np.argsort(-np.abs(W), axis=1, kind="stable")[:, :take]
        assert np.array_equal(idx, ref)
        kept = np.take_along_axis(W, ref, axis=1)
        assert np.allclose(w, kept / kept.sum(axis=1, keepdims=True), rtol=1e-14)
