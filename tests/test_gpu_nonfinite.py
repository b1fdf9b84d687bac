"""A state that stops being finite is reported (GPE_ENONFINITE -> GPENonFiniteError).

The block end divides by the norm (lib/@GPEtask/solve_split.m:67-73: phi*sqrt(NNN/ncur))
and the chemical potential enters the next block's phase; a NaN there would otherwise
run silently to the end.  Each test uploads a state with one NaN point: every kernel
runs to completion (NaN only propagates through the arithmetic), and the solver
returns the error at the first block-end read."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(grid):
    r2 = grid.mesh.x ** 2 + grid.mesh.y ** 2 + (grid.mesh.z ** 2 if grid.ndims == 3 else 0)
    phi = np.exp(-r2 / 4).astype(np.complex128)
    phi.flat[7] = np.nan
    return phi


@pytest.mark.parametrize("dims,gamma", [(2, 0.0), (2, 0.01), (3, 0.01)])
def test_solve_split_reports_nan_state(dims, gamma):
    import gpe_hip
    grid = gpe_hip.grid2d(6, 32, 5, 24) if dims == 2 else gpe_hip.grid3d(5, 16, 5, 16, 4, 12)
    t = gpe_hip.GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g, t.gamma, t.Ntotal = 0.5, gamma, 3.0
    t.init_state = _state(grid)
    with pytest.raises(gpe_hip.GPENonFiniteError, match="ENONFINITE"):
        t.solve_split(0.002, 20, 3)


def test_finite_run_is_not_flagged():
    import gpe_hip
    grid = gpe_hip.grid2d(6, 32, 5, 24)
    t = gpe_hip.GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g, t.gamma, t.Ntotal = 0.5, 0.01, 3.0
    phi = _state(grid)
    phi.flat[7] = 0.0
    t.init_state = phi
    out = t.solve_split(0.002, 20, 3)
    assert np.isfinite(out).all()


def test_spinor_reports_nan_state():
    import gpe_hip
    grid = gpe_hip.grid2d(6, 32, 5, 24)
    r2 = grid.mesh.x ** 2 + grid.mesh.y ** 2
    t = gpe_hip.GPEtaskSpinor(grid, [0.5 * r2, 0.5 * r2])
    t.g = np.array([[0.5, 0.45], [0.4, 0.6]])
    t.coupling = 0.3 * np.exp(0.5j)
    t.Ntotal = 4.0
    t.init_state = [np.exp(-r2 / 3) + 0j, _state(grid)]
    with pytest.raises(gpe_hip.GPENonFiniteError):
        t.solve_split(0.002, 20, 2)


def test_itp_reports_nan_state():
    import gpe_hip
    grid = gpe_hip.grid2d(6, 32, 5, 24)
    t = gpe_hip.GPEtask(grid, lambda X, Y, Z: 0.5 * (X ** 2 + Y ** 2))
    t.g, t.Ntotal = 0.5, 3.0
    with pytest.raises(gpe_hip.GPENonFiniteError):
        t.groundstate_itp(0.01, 1e-9, _state(grid))


def test_batch_reports_nan_trajectory():
    import gpe_hip
    from gpe_hip.batch import GPEEnsemble
    tasks = []
    for i, n in enumerate((32, 64)):
        G = gpe_hip.grid2d(6, n, 6, n)
        r2 = G.mesh.x ** 2 + G.mesh.y ** 2
        t = gpe_hip.GPEtask(G, 0.5 * r2)
        t.g, t.Ntotal = 0.3, 2.0
        psi = np.exp(-r2 / 4).astype(np.complex128)
        if i == 1:
            psi.flat[3] = np.nan
        t.init_state = psi
        tasks.append(t)
    ens = GPEEnsemble(tasks, tune=False)
    with pytest.raises(gpe_hip.GPENonFiniteError, match="trajectory 1"):
        ens.solve_split(0.002, 20, 2)
