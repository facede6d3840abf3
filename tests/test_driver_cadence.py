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
