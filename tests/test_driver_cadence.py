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
    call (diag.c:95-136, :574-577, post_dump :511-523)."""
    c = driver.RestartCadence(3, 0)
    written = []
    for n in range(0, 8):
        if c.due(n, "step"):
            written.append((n, c.which))
            c.done(n)
    assert written == [(3, 0), (6, 1)]
    assert c.due(7, "final") and c.which == 0
    # started from a restart at step 5: the next is at 6, not 5
    c = driver.RestartCadence(3, 5)
    assert not c.due(5, "init") and c.due(6, "step")
    assert not driver.RestartCadence(None, 0).due(10, "final")
