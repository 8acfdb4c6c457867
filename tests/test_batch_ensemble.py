"""The batched ensemble's draw (host logic, no GPU): splitmix64 against its published
first output for seed 0, and the composition of the 32-instance ensemble."""
from sweeping_fd_weno_amd import batch as B


def test_splitmix64_known_answer():
    # splitmix64 (Vigna): the first output from state 0 is 0xe220a8397b1dcdaf
    assert B.splitmix64(0)[1] == 0xE220A8397B1DCDAF
    s, a = B.splitmix64(0)
    _, b = B.splitmix64(s)
    assert b == 0x6E789E6AA1B965F4


def test_ensemble_composition():
    ens = B.ensemble("fp32_bulk")
    assert len(ens) == 48
    sedov = [e for e in ens if e[0].startswith("sedov")]
    vort = [e for e in ens if e[0].startswith("vortex")]
    step = [e for e in ens if e[0].startswith("step")]
    assert len(sedov) == len(vort) == len(step) == 16
    for fam in (sedov, vort):
        sizes = [(c.nx, c.ny) for _, c, _, _ in fam]
        assert sizes[:8] == B.SIZES and sizes[8:] == B.SIZES
    assert [c.step_total for _, c, _, _ in step] == B.STEP_TOTALS * 2
    for label, c, u, tf in sedov:
        e0 = float(label.split("_E")[1])
        assert 0.1 <= e0 <= 1.0 and tf == 1.0
        assert abs(u[3:-3, 3:-3, 3].max() * c.dx * c.dy - e0) < 1e-3
    for label, c, u, tf in vort:
        mag = float(label.split("_mag")[1])
        assert 5.0 <= mag <= 10.0828 and tf == 0.01
    for label, c, u, tf in step:
        mach = float(label.split("_M")[1])
        assert 3.0 <= mach <= 6.0 and tf == 2.3
        assert (c.nx, c.ny) == (13 * c.step_total, 11 * c.step_total) and c.bc == 4
    # the families alternate and the draw is deterministic
    assert [e[0][:4] for e in ens[:6]] == ["sedo", "vort", "step", "sedo", "vort", "step"]
    assert [e[0] for e in B.ensemble("fp64")] == [e[0] for e in ens]
