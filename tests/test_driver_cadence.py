"""The host driver's output schedule (harm_harm_amd.driver.Cadence), CPU
only: diag()'s rule (get_dodumps diag.c:549-700, post_dump :502-548)."""
from harm_harm_amd import driver


def test_cadence_rule():
    """post_dump's next time: the first multiple of DT after t, never a repeat."""
    c = driver.Cadence(1.0, 0.0)
    assert c.due(0.0, "init")
    c.done(0.0)
    assert c.tnext == 1.0 and not c.due(0.0, "final")
    assert not c.due(0.9, "step") and c.due(1.05, "step")
    c.done(3.4)  # jumped past two periods
    assert c.tnext == 4.0 and c.count == 2


def test_restart_cadence_rule():
    """Every DTr steps the next of rdump-0 / rdump-1, and always at the final
    call (diag.c:95-136, :574-577, post_dump :511-523); a fresh run writes
    rdump-0 at its first call (nrestart = nstep, :111)."""
    c = driver.RestartCadence(3, 0)
    written = []
    for n in range(0, 8):
        if c.due(n, "init" if n == 0 else "step"):
            written.append((n, c.which))
            c.done(n)
    assert written == [(0, 0), (3, 1), (6, 0)]
    assert c.due(7, "final") and c.which == 1
    # resumed at step 5 after rdump-0: the next is at 6 (not 5), into rdump-1
    c = driver.RestartCadence(3, 5, resume=True, which=1)
    assert not c.due(5, "init") and c.due(6, "step") and c.which == 1
    assert not driver.RestartCadence(None, 0).due(10, "final")


def test_resumed_cadence():
    """A resumed run writes its next main dump at the first multiple of DT
    after t0, not at t0 (diag.c:137-139, :586), and numbers it on; the energy
    file also writes at the initial call (:683)."""
    c = driver.Cadence(1.0, 2.5, resume=True, count=3)
    assert c.tnext == 3.0 and not c.due(2.5, "init") and c.due(3.0, "step")
    c.done(3.0)
    assert c.count == 4 and c.tnext == 4.0
    e = driver.Cadence(1.0, 2.5, resume=True, at_init=True)
    assert e.due(2.5, "init")
    f = driver.Cadence(1.0, 0.0)
    assert f.due(0.0, "init")
