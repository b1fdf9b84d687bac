"""GPU runs of the benchmark configurations themselves, checked against the oracle,
finite differences and single-trajectory runs.

  c1  2D 256^2 (examples/example1_vortex.m:11-35): groundstate_itp(0.02, 1e-6) from the
      seeded random guess, vortex imprint at x0 = 5, then 20 x 50 split steps.
  c3  adjoint on 512^2 with 8 bumps x 8 intervals: the full T = 2000 gradient against
      central differences, and the oracle adjoint on the same problem with 30-step
      intervals (T = 240).
  c4  the 256-trajectory ragged ensemble, 100 steps in one graph-replayed outer
      iteration: one trajectory of every grid shape bit-identical to its single-task
      run, three trajectories against the oracle.
  c5  the (Omega0, k) gradient of the population imbalance at 192^3 after 1000 steps
      against a fourth-order central difference.
Tolerances are written in each test.
"""
import gc

import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_c1_itp_then_vortex_dynamics():
    import gpe_hip
    from gpe_hip import workloads
    cfg = workloads.c1_vortex_2d()
    res = {}
    for name, mod in (("gpu", gpe_hip), ("oracle", oracle)):
        G = mod.grid2d(*cfg["grid"][1])
        t = mod.GPEtask(G, cfg["V"], seed=cfg["seed"])
        t.Ntotal, t.g = cfg["props"]["Ntotal"], cfg["props"]["g"]
        phi, MU = t.groundstate_itp(*cfg["itp"], nargout=2)
        res[name + "_itp"] = (phi, MU, t.itp_iterations)
        t.init_state = workloads.imprint_vortex(G, phi, t.g)
        out = t.solve_split(cfg["split"][0], cfg["split"][1], 20)
        res[name] = (out, t.history["mu"].copy(), t.history["n"].copy())
    (pg, MUg, ig), (po, MUo, io) = res["gpu_itp"], res["oracle_itp"]
    e_itp = rel_l2(pg, po)
    print(f"c1 ITP: {ig} iterations (oracle {io}); state rel-L2 {e_itp:.3e}", flush=True)
    assert ig == io
    assert e_itp < 1e-11
    assert np.allclose(MUg, MUo, rtol=1e-11, atol=0)
    (sg, mug, ng), (so, muo, no) = res["gpu"], res["oracle"]
    e = rel_l2(sg, so)
    print(f"c1 vortex, 1000 steps: rel-L2 {e:.3e}", flush=True)
    assert e < 1e-11
    assert np.allclose(mug, muo, rtol=1e-11, atol=0)
    assert np.allclose(ng, no, rtol=1e-11, atol=0)


def _c3(nsteps, spi):
    import gpe_hip
    from gpe_hip import workloads
    from gpe_hip.adjoint import SplitStepControl
    w = workloads.c3_control(nsteps=nsteps, spi=spi)
    t = gpe_hip.GPEtask(w["grid"], w["V0"])
    t.g, t.Ntotal = w["g"], w["Ntotal"]
    t.init_state = w["psi0"]
    return w, SplitStepControl(t, w["V0"], w["basis"], w["ddt"], w["nsteps"], w["spi"])


def test_c3_full_horizon_gradient_vs_central_fd():
    w, ctl = _c3(2000, 250)
    psi_t = ctl.forward(w["c_target"])
    L, g, rerr = ctl.loss_and_grad(w["c"], psi_t)
    G = w["grid"]
    h = 1e-6
    fd, ad = [], []
    for (k, m) in [(0, 0), (3, 2), (5, 5), (7, 7), (1, 4)]:
        cp_, cm_ = w["c"].copy(), w["c"].copy()
        cp_[k, m] += h
        cm_[k, m] -= h
        Lp = 1 - abs(G.inner(psi_t, ctl.forward(cp_))) ** 2
        Lm = 1 - abs(G.inner(psi_t, ctl.forward(cm_))) ** 2
        fd.append((Lp - Lm) / (2 * h))
        ad.append(g[k, m])
    err = rel_l2(ad, fd)
    print(f"c3 T=2000: loss {L:.6f}, gradient vs central FD (h=1e-6) rel {err:.2e}, "
          f"backward reconstruction {rerr:.1e}", flush=True)
    assert err < 1e-6
    assert rerr < 1e-11


def test_c3_gradient_vs_oracle_adjoint_short_horizon():
    from oracle.adjoint import ControlProblem
    w, ctl = _c3(240, 30)
    O = oracle.grid2d(*w["grid_args"])
    cp = ControlProblem(O, w["V0"], w["basis"], g=w["g"], ddt=w["ddt"], nsteps=w["nsteps"],
                        steps_per_interval=w["spi"], psi0=w["psi0"], Ntotal=w["Ntotal"])
    psi_t = cp.forward(w["c_target"])
    Lg, gg, _ = ctl.loss_and_grad(w["c"], psi_t)
    Lo, go, _ = cp.loss_and_grad(w["c"], psi_t)
    err = rel_l2(gg, go)
    print(f"c3 T=240 (8 x 30 steps): loss {Lg:.9f} vs oracle {Lo:.9f}; gradient rel {err:.2e}", flush=True)
    assert abs(Lg - Lo) < 1e-11
    assert err < 1e-9


def test_c4_ensemble_trajectories():
    import gpe_hip
    from gpe_hip import workloads
    from gpe_hip.batch import GPEEnsemble
    ens = workloads.c4_ensemble()
    tasks = []
    for G, V, psi, pr in ens:
        t = gpe_hip.GPEtask(G, V)
        t.Ntotal, t.g = pr["Ntotal"], pr["g"]
        t.init_state = psi
        tasks.append(t)
    e = GPEEnsemble(tasks)
    outs = e.solve_split(0.004, 100, 1)
    print(f"c4: {len(tasks)} trajectories, {sum(o.size for o in outs)} points, 100 steps", flush=True)
    first = {}
    for i, (G, _, _, _) in enumerate(ens):
        first.setdefault(G.nx, i)
    assert sorted(first) == [128, 192, 256, 384]
    for n, i in sorted(first.items()):
        G, V, psi, pr = ens[i]
        s = gpe_hip.GPEtask(G, V)
        s.Ntotal, s.g = pr["Ntotal"], pr["g"]
        s.init_state = psi
        ref = s.solve_split(0.004, 100, 1)
        same = np.array_equal(outs[i], ref) and np.array_equal(tasks[i].history["mu"], s.history["mu"])
        print(f"  trajectory {i} ({n}^2): bit-identical to its single run: {same}", flush=True)
        assert same
        del s
        gc.collect()
    for i in (first[128], first[256], first[384]):
        G, V, psi, pr = ens[i]
        O = oracle.grid2d(G.xmax, G.nx, G.ymax, G.ny)
        to = oracle.GPEtask(O, V)
        to.Ntotal, to.g = pr["Ntotal"], pr["g"]
        to.init_state = psi
        po = to.solve_split(0.004, 100, 1)
        err = rel_l2(outs[i], po)
        print(f"  trajectory {i} ({G.nx}^2) vs oracle: rel-L2 {err:.3e}", flush=True)
        assert err < 1e-11
        assert np.allclose(tasks[i].history["mu"], to.history["mu"], rtol=1e-11, atol=0)


def test_c5_rabi_gradient_full_size_vs_fourth_order_fd():
    """192^3, 1000 steps.  The loss (N1-N2)/N is ~1e-5 with a rounding noise of ~1e-15,
    so the fourth-order stencil at H = 1e-3 (truncation ~H^4, rounding ~1e-12 / H) is
    the reference; the adjoint agrees to ~1e-7 relative."""
    import gpe_hip
    from gpe_hip import workloads
    from gpe_hip.adjoint import RabiControl
    w = workloads.c5_spinor()
    t = gpe_hip.GPEtaskSpinor(w["grid"], [w["V"], w["V"]])
    t.g, t.Ntotal = w["g"], w["Ntotal"]
    t.init_state = w["psi"]
    ctl = RabiControl(t, 0.002, 1000)
    om0, k = w["omega0"], w["k"]
    L, g = ctl.loss_and_grad(om0, k)
    H = 1e-3
    fd4 = []
    for i in range(2):
        vals = []
        for m in (2, 1, -1, -2):
            p = [om0, k]
            p[i] += m * H
            vals.append(ctl.loss(*p))
        fd4.append((-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * H))
    rel = [abs(g[i] - fd4[i]) / abs(fd4[i]) for i in range(2)]
    print(f"c5 gradient: L {L:.6e}, adjoint {g}, 4th-order FD {fd4}, rel {rel}", flush=True)
    assert L == ctl.loss(om0, k)
    assert max(rel) < 1e-6
