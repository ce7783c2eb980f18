"""Pins the oracle's red-black SOR and residual norm, which have no
counterpart in the reference (js/solver.js:1397-1404 sweeps lexicographically
for a fixed count), by the mathematics of the system they solve:

  - the Poisson matrix is rebuilt here from the cell types alone, following
    the rule of js/solver.js:1330-1365 (diag = fluid + air neighbours, -1 per
    fluid neighbour, solid = Neumann, air = Dirichlet 0);
  - a sparse FP64 direct solve of that matrix is the converged pressure;
  - the oracle's residual is checked against ||rhs - A q|| / ||rhs|| computed
    with scipy on the same q;
  - red-black SOR and the reference's lexicographic SOR must converge to the
    direct solution, and one red-black sweep must leave every red cell
    satisfying its own row exactly before the black half runs.
A wrong colour parity, a dropped neighbour, a wrong sign of the relaxation or
an air neighbour treated as Neumann fails one of these checks."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import golden
from scenes import oracle_from_golden


def poisson_matrix(cell_type, nx, ny, nz):
    """Rows over fluid cells in (k, j, i) order; columns over the same list."""
    t = cell_type.reshape(nz, ny, nx)
    fl = np.flatnonzero(cell_type == 1)
    pos = -np.ones(cell_type.size, np.int64)
    pos[fl] = np.arange(fl.size)
    k, j, i = np.unravel_index(fl, (nz, ny, nx))
    rows, cols, vals = [], [], []
    diag = np.zeros(fl.size)
    for dk, dj, di in [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]:
        kk, jj, ii = k + dk, j + dj, i + di
        ok = (kk >= 0) & (kk < nz) & (jj >= 0) & (jj < ny) & (ii >= 0) & (ii < nx)
        nt = np.full(fl.size, 2)
        nt[ok] = t[kk[ok], jj[ok], ii[ok]]
        diag += (nt != 2)
        f = nt == 1
        nb = ((kk * ny + jj) * nx + ii)[f]
        rows.append(np.flatnonzero(f))
        cols.append(pos[nb])
        vals.append(-np.ones(f.sum()))
    rows.append(np.arange(fl.size))
    cols.append(np.arange(fl.size))
    vals.append(diag)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(fl.size, fl.size))
    return A, fl


def assembled(scene, bits=64, omega=None):
    g = golden.load(scene)
    st = oracle_from_golden(g, bits=bits)
    if omega is not None:
        st.set_f("omega", omega)  # the relaxation weights omega/diag are built at assembly
    st.arrays["cellType"][:] = g["mark_cellType"]
    for k in "uvw":
        st.arrays[k][:] = g["grav_" + k]
    st.arrays["q"][:] = g["in_q"]
    nF = st.call("oc_pressure_assemble")
    assert nF == g.meta["nFluid"]
    return g, st


@pytest.mark.parametrize("scene", ["micro_16", "cfg1_32"])
def test_rhs_and_diag_follow_the_matrix(scene):
    g, st = assembled(scene)
    w = g.meta["world"]
    A, fl = poisson_matrix(st["cellType"], w["nx"], w["ny"], w["nz"])
    assert np.array_equal(st["diag"][fl], A.diagonal())
    # rhs = -dx^2 div u (the 1/dx of the divergence times dx^2): recompute it
    nx, ny, nz, dx = w["nx"], w["ny"], w["nz"], w["dx"]
    u = np.asarray(g["grav_u"], np.float64).reshape(nz, ny, nx + 1)
    v = np.asarray(g["grav_v"], np.float64).reshape(nz, ny + 1, nx)
    ww = np.asarray(g["grav_w"], np.float64).reshape(nz + 1, ny, nx)
    div = (u[:, :, 1:] - u[:, :, :-1] + v[:, 1:, :] - v[:, :-1, :] + ww[1:] - ww[:-1]) / dx
    rhs = (-dx * dx * div).ravel()[fl]
    assert np.abs(st["rhs"][fl] - rhs).max() <= 1e-12 * max(1.0, np.abs(rhs).max())


@pytest.mark.parametrize("scene", ["micro_16", "cfg1_32"])
def test_residual_norm_matches_scipy(scene):
    g, st = assembled(scene)
    w = g.meta["world"]
    A, fl = poisson_matrix(st["cellType"], w["nx"], w["ny"], w["nz"])
    rng = np.random.default_rng(7)
    st.arrays["q"][fl] = rng.standard_normal(fl.size)
    b = st["rhs"][fl].astype(np.float64)
    want = np.linalg.norm(b - A @ st["q"][fl]) / np.linalg.norm(b)
    got = st.call("oc_residual", None)
    assert abs(got - want) <= 1e-12 * want


def test_red_black_half_sweep_solves_its_own_rows():
    g, st = assembled("micro_16", omega=1.0)  # Gauss-Seidel: each updated row holds exactly
    w = g.meta["world"]
    nx, ny, nz = w["nx"], w["ny"], w["nz"]
    A, fl = poisson_matrix(st["cellType"], nx, ny, nz)
    st.call("oc_sor_redblack", 1)
    q = st["q"][fl].astype(np.float64)
    r = st["rhs"][fl] - A @ q
    k, j, i = np.unravel_index(fl, (nz, ny, nx))
    black = ((i + j + k) & 1) == 1
    # black cells were updated last, with all their (red) neighbours final;
    # the weight 1/diag is stored in float32 (js/solver.js:1383), so a row
    # holds to float rounding of that weight
    assert np.abs(r[black]).max() <= 1e-6 * np.abs(st["rhs"][fl]).max()
    assert np.abs(r[~black]).max() > 1e-6  # the red rows went stale


@pytest.mark.parametrize("lexicographic", [False, True])
def test_sor_converges_to_direct_solve(lexicographic):
    g, st = assembled("micro_16", omega=1.7)
    w = g.meta["world"]
    A, fl = poisson_matrix(st["cellType"], w["nx"], w["ny"], w["nz"])
    b = st["rhs"][fl].astype(np.float64)
    exact = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(b - A @ exact) <= 1e-10 * np.linalg.norm(b)
    st.call("oc_sor_lex" if lexicographic else "oc_sor_redblack", 3000)
    q = st["q"][fl]
    # the fixed point solves diag' q - sum = rhs with diag' = omega / fl32(omega / diag),
    # a relative change of the diagonal below 2^-24
    assert np.abs(q - exact).max() <= 1e-6 * np.abs(exact).max()
    assert st.call("oc_residual", None) <= 1e-6
