"""Box-mesh fixture: sizes, orientation, BC codes and CSR graph."""
import numpy as np
import pytest

from phasta_amd import config, mesh


@pytest.mark.parametrize("N", [2, 4, 16])
def test_counts_and_nnz(N):
    c = mesh.box_case(N)
    assert c.nshg == (N + 1) ** 3
    assert c.nel == 6 * N ** 3
    assert c.nelb == 4 * N * N
    assert c.nnz_tot == mesh.expected_nnz(N)


def test_config_sizes_match_the_workloads():
    # sizes quoted for the 16^3 / 64^3 / 128^3 / 256^3 boxes
    assert mesh.expected_nnz(16) == 66961
    assert mesh.expected_nnz(64) == 4018753
    assert mesh.expected_nnz(128) == 31802497
    assert mesh.expected_nnz(256) == 253036801


@pytest.mark.parametrize("grade", [1.0, 1.1])
def test_positive_volumes_and_unit_box(grade):
    c = mesh.box_case(8, jitter=0.1, grade=grade)
    x, ien = c.x, c.ien - 1
    a = x[ien[:, 0]] - x[ien[:, 3]]
    b = x[ien[:, 1]] - x[ien[:, 3]]
    d = x[ien[:, 2]] - x[ien[:, 3]]
    det = np.einsum("ij,ij->i", a, np.cross(b, d))
    assert det.min() > 0
    assert abs(det.sum() / 6 - 1.0) < 1e-13


def test_wall_faces_outward_and_on_walls():
    c = mesh.box_case(4, jitter=0.1)
    x, ienb = c.x, c.ienb - 1
    y = x[:, 1]
    v1 = x[ienb[:, 1]] - x[ienb[:, 0]]
    v2 = x[ienb[:, 2]] - x[ienb[:, 0]]
    nrm = np.cross(v1, v2)
    face_y = y[ienb[:, :3]]
    assert np.all((face_y == 0).all(1) | (face_y == 1).all(1))
    outward = np.where(face_y[:, 0] == 0, -1.0, 1.0)
    assert np.all(nrm[:, 1] * outward > 0)
    assert np.allclose(np.abs(nrm).sum() / 2, 2.0)  # two walls of unit area


def test_bc_codes_and_periodic_masters():
    N = 4
    c = mesh.box_case(N)
    n = np.arange(c.nshg)
    i, j, k = n % (N + 1), (n // (N + 1)) % (N + 1), n // (N + 1) ** 2
    wall = (j == 0) | (j == N)
    slave = (i == N) | (k == N)
    assert np.array_equal((c.iBC & 56) == 56, wall)
    assert np.array_equal((c.iBC & 1024) != 0, slave)
    m = c.iper - 1
    assert np.all(m[m] == m)                       # final masters
    assert np.array_equal(m == n, ~slave)
    assert np.allclose(c.x[m, 1], c.x[:, 1])       # same y


def test_csr_graph_is_sorted_adjacency():
    c = mesh.box_case(3)
    colm, rowp = c.colm - 1, c.rowp - 1
    adj = [set() for _ in range(c.nshg)]
    for e in c.ien - 1:
        for a in e:
            adj[a].update(e)
    for r in range(c.nshg):
        cols = rowp[colm[r]:colm[r + 1]]
        assert list(cols) == sorted(adj[r])


def test_state_ranges():
    c = mesh.box_case(4, mach=0.3)
    u, v, w, p, T = c.yold.T
    smax = 8 / np.sqrt(8)            # |s| <= sum |a_m|, |a_m| <= 1/sqrt(8)
    s = (p / 1e5 - 1) / 0.01
    assert np.all(np.abs(s) <= smax)
    assert np.allclose(T / 300 - 1, 0.01 * s)
    assert np.allclose(v, w)


@pytest.mark.parametrize("grade", [1.0, 1.1])
def test_time_step_uses_the_smallest_grid_spacing(grade):
    """dt = CFL h_min / (c0 (1 + M)), h_min the first wall cell of the graded
    grid (measured on an unjittered mesh), 1/N on a uniform one."""
    c = mesh.box_case(16, jitter=0.0, grade=grade, mach=0.3, cfl=3.0)
    ys = np.unique(np.round(c.x[:, 1], 14))
    hmin = np.min(np.diff(ys))
    assert abs(hmin - config.min_spacing(16, grade)) <= 1e-12
    assert abs(ys[1] - ys[0] - hmin) <= 1e-12          # at the wall
    dt = 1.0 / c.params["Dtgl"]
    assert abs(dt - 3.0 * hmin / (config.C0 * 1.3)) <= 1e-12 * dt
