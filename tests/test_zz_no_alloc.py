"""kh_sys_solve_static: the per-system seam's Solve (solver.Matrix.Solve,
solver/matrix.go:23-36) with the system's pattern, values, solution and LU
workspace in the library's static device array instead of device memory it
allocates, for callers that cannot allocate device memory. Its results must
equal kh_sys_solve's (the allocating entry point, the same factorization)
bit for bit, on consecutive calls whose systems change size, whose second
system is smaller, and when another system used the array in between.
"""
import numpy as np
import pytest

from kronspice_amd import api

pytestmark = pytest.mark.gpu


def solve_static(s, b):
    return s.solve_static(b)


def grid_pair(w, h, nv, seed):
    """Two handles stamped with the same MNA grid (conductances, a small
    conductance to ground at every node, nv voltage sources)."""
    rng = np.random.default_rng(seed)
    nn = w * h
    n = nn + nv
    a, c = api.Sparse(n), api.Sparse(n)
    st = []
    for y in range(h):
        for x in range(w):
            i = y * w + x
            for j, ok in ((i + 1, x + 1 < w), (i + w, y + 1 < h)):
                if ok:
                    g = 1 / (10 + 990 * rng.random())
                    st += [(i, i, g), (j, j, g), (i, j, -g), (j, i, -g)]
            st.append((i, i, 1e-9))
    b = np.zeros(n)
    for k in range(nv):
        node, br = (k * 7919) % nn, nn + k
        st += [(node, br, 1.0), (br, node, 1.0)]
        b[br] = 1 + rng.random()
    for i, j, v in st:
        a.add(i, j, v)
        c.add(i, j, v)
    return a, c, b


def test_static_seam_equals_allocating_seam():
    big_a, big_c, big_b = grid_pair(40, 40, 4, 1)
    small_a, small_c, small_b = grid_pair(12, 9, 2, 2)
    # two calls in a row, the second with a smaller system
    x1 = solve_static(big_a, big_b)
    assert np.array_equal(x1, big_c.solve(big_b))
    x2 = solve_static(small_a, small_b)
    assert np.array_equal(x2, small_c.solve(small_b))
    # the same system again (its pattern already in the array), then after
    # another system used it, then with new values on the same pattern
    assert np.array_equal(solve_static(small_a, small_b), x2)
    assert np.array_equal(solve_static(big_a, big_b), x1)
    big_a.clear(); big_c.clear()
    _, _, b3 = grid_pair(40, 40, 4, 3)
    a3, c3, _ = grid_pair(40, 40, 4, 3)
    assert np.array_equal(solve_static(a3, b3), c3.solve(b3))


def test_static_and_allocating_solves_on_one_handle():
    """Either entry point after the other on the same handle: the static call
    analyses the pattern on the host only, so the allocating call that follows
    still sizes its own device buffers for it (and the reverse)."""
    a, c, b = grid_pair(20, 15, 3, 4)
    xs = solve_static(a, b)
    assert np.array_equal(a.solve(b), xs)
    assert np.array_equal(c.solve(b), xs)
    assert np.array_equal(solve_static(c, b), xs)
    # a new entry grows the pattern: both entry points see the new system
    for s in (a, c):
        s.add(0, 5, 1e-3)
    xs2 = solve_static(a, b)
    assert np.array_equal(a.solve(b), xs2)
    assert np.array_equal(c.solve(b), xs2)
    assert np.array_equal(solve_static(c, b), xs2)
    assert not np.array_equal(xs2, xs)


def test_static_seam_error_classes():
    s = api.Sparse(3)
    for j in range(3):
        s.add(0, j, 1); s.add(1, j, 2)
    s.add(2, 2, 1)
    with pytest.raises(api.KhError) as e:
        solve_static(s, np.zeros(3))
    assert e.value.status == api.SINGULAR


def test_static_seam_rejects_oversize():
    n = 200_000  # the workspace of a 200,000-unknown system exceeds the static array
    s = api.Sparse(n)
    i = np.arange(n, dtype=np.int32)
    v = np.ones(n)
    api._check(api.lib().kh_sys_add_batch(s.h, api._p(i), api._p(i), api._p(v), n))
    with pytest.raises(api.KhError) as e:
        solve_static(s, np.ones(n))
    assert e.value.status == api.BADSIZE
