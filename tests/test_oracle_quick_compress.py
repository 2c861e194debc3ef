"""The oracle's QuickCompress (OracleQuickCompress in tests/_oracle.py,
hoomd-mc/src/quick_compress.rs:19-444) and its Uniform draw: closed forms of
the draw, and the state machine's contract on a placement run under the
overlap penalty (benchmarks/src/place.rs:31-118): it ends Complete exactly at
the target volume, in a state of zero energy (no pair closer than delta),
after passing through Waiting and Compressing."""
import numpy as np

import _oracle as O
import _pkg

wl = _pkg.workloads()
P = (1e4, 0.2, 100.0, 1.0)


def test_uniform_draw_closed_forms():
    assert O.uniform_f64(0, 0.0, 0.05) == 0.0
    assert O.uniform_f64(1 << 63, 0.0, 0.05) == 0.025
    top = O.uniform_f64((1 << 64) - 1, 0.0, 0.05)
    assert top == (1.0 - 2.0**-52) * 0.05 and top < 0.05
    # a range whose top would round up: the scale is lowered by an ulp
    assert O.uniform_f64((1 << 64) - 1, 3.0, 5.0) < 5.0


def test_placement_run_completes_at_the_target():
    n, rho_t = 400, 0.6
    L = [(n / (0.7 * rho_t)) ** (1 / 3)] * 3
    vt = n / rho_t
    s = O.OracleSystem(L, wl.random_dilute(n, L, 3, 1.0, 3), model="overlap_penalty", params=P, seed=5)
    qc = O.OracleQuickCompress(vt)
    seen = set()
    for _ in range(3000):
        step, _sub = s.counter()
        if step % 20 == 0:
            s, L = qc.apply(s, L, 5, "overlap_penalty", P)
            seen.add(qc.state)
            if qc.state == "complete":
                break
        s.parallel_apply(1.0, 0.1)
        s.increment_step()
    assert qc.state == "complete" and {"waiting", "compressing"} <= seen
    assert np.prod(L) == vt or abs(np.prod(L) / vt - 1.0) < 1e-14
    assert s.total_energy() == 0.0
    body, _ = s.positions()
    d = body[:, None, :] - body[None, :, :]
    d -= np.asarray(L) * np.round(d / np.asarray(L))
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    assert r2.min() >= 1.0
    # Complete is kept while the boundary is unchanged
    step, _ = s.counter()
    s2, L2 = qc.apply(s, L, 5, "overlap_penalty", P)
    assert qc.state == "complete" and s2 is s and L2 == L
