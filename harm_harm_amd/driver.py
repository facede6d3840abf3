"""HARM's time loop and output cadence around the device step.

Mirrors, on the host, what main.c does around step_ch_full():

* the loop runs whole steps until the end-of-run rule of post_stepch
  (step_ch.c:200-250) stops it: after each step the next dt (already the
  device's min(ndt, SAFE dt)) is clipped to land on tf and one more step is
  taken; a step that would land exactly on tf ends the run without it; a dt
  below SMALL ends it;
* outputs follow diag()'s cadence (get_dodumps diag.c:549-700, post_dump
  :502-548): an output type with period DT is written when t reaches its next
  time (and at the initial and final calls), never twice at the same t, and
  its next time then moves to the first multiple of DT after t.

Setting the clipped dt goes through harm_set_time, which also restarts the
Newton sums (harm_get_newton_stats' totals) at the last steps.

* restart files follow the step-count rule (diag.c:95-136, :574-577,
  post_dump :511-523): every DTr steps (and at the final call) the next of
  dumps/rdump-0 and dumps/rdump-1 is written, and every main dump also
  writes dumps/rdump--NNNN with its number (diag.c:319-328).

Files, as HARM names them: dumps/gdump once at the start, dumps/dumpNNNN
(main dump), the restart files above and ener.out (one energy-file line
per ener output).
"""
import math
import os

from .grid import ener_line

SMALL = 1.0e-300  # SMALL for doubles (global.depnmemonics.h)


class Cadence:
    """One output type's schedule: tdumpgen, DTdumpgen, dumpcntgen, tlastgen."""

    def __init__(self, DT, t0):
        self.DT = DT
        self.tnext = t0
        self.count = 0
        self.tlast = None

    def due(self, t, call):
        """call: "init", "step" or "final" (INIT_OUT, DUMP_OUT, FINAL_OUT)"""
        if self.DT is None or t == self.tlast:
            return False
        return t >= self.tnext or call in ("init", "final")

    def done(self, t):
        # post_dump: dumpc = 1 + max(0, (long long)((t - tdump)/DT));
        # tdump = (ROUND2LONGLONGINT(tdump/DT) + dumpc) DT
        dumpc = 1 + max(0, int((t - self.tnext) / self.DT))
        self.tnext = (math.floor(self.tnext / self.DT + 0.5) + dumpc) * self.DT
        self.count += 1
        self.tlast = t


class RestartCadence:
    """Restart dumps by step count: nrestart, nlastrestart, whichrestart."""

    def __init__(self, DTr, n0):
        self.DTr = DTr
        self.which = 0
        self.count = 0
        if DTr is not None:
            self.nlast = DTr * (n0 // DTr) - 1
            self.nnext = DTr * (n0 // DTr + 1)

    def due(self, n, call):
        if self.DTr is None:
            return False
        return (self.nlast != self.nnext and n >= self.nnext) or call == "final"

    def done(self, n):
        self.which = 1 - self.which
        c = 1 + max(0, int((n - self.nnext) / self.DTr))
        self.nnext = (math.floor(n / self.DTr + 0.5) + c) * self.DTr
        self.nlast = n
        self.count += 1


def run(grid, tf, DTdump=None, DTener=None, outdir=".", gdump=True, max_steps=None, which="p", DTr=None,
        restart_binary=True):
    """Advance `grid` (harm_harm_amd.grid.Grid, clock already set) to tf with
    HARM's end-of-run rule, writing dumps every DTdump, energy-file lines
    every DTener and restart files every DTr steps into outdir (restart
    files MIXEDOUTPUT when restart_binary, else TEXTOUTPUT).  Returns a
    summary dict."""
    os.makedirs(os.path.join(outdir, "dumps"), exist_ok=True)
    t0, _ = grid.time()
    dump_c, ener_c = Cadence(DTdump, t0), Cadence(DTener, t0)
    rdump_c = RestartCadence(DTr, grid.nstep())
    if DTener is not None:
        grid.set_ener_tracking(True)
    ener_path = os.path.join(outdir, "ener.out")
    if DTener is not None:
        open(ener_path, "w").close()
    if gdump:
        grid.write_gdump(os.path.join(outdir, "dumps", "gdump"))

    def diag(call):
        t, _ = grid.time()
        if dump_c.due(t, call):
            grid.write_dump(os.path.join(outdir, "dumps", "dump%04d" % dump_c.count), which)
            if DTr is not None:
                grid.write_restart(os.path.join(outdir, "dumps", "rdump--%04d" % dump_c.count), restart_binary, tf)
            dump_c.done(t)
        n = grid.nstep()
        if rdump_c.due(n, call):
            grid.write_restart(os.path.join(outdir, "dumps", "rdump-%01d" % rdump_c.which), restart_binary, tf)
            rdump_c.done(n)
        if ener_c.due(t, call):
            with open(ener_path, "a") as f:
                f.write(ener_line(grid.ener(which)))
            ener_c.done(t)

    diag("init")
    steps, onemorestep = 0, False
    while True:
        grid.step(1)
        steps += 1
        t, dt = grid.time()
        dt0, last = dt, False
        if onemorestep:
            last, dt = True, SMALL
        elif t + dt > tf:
            dt, onemorestep = tf - t, True
        elif t + dt == tf:
            last = True
        if dt < SMALL:
            last, dt = True, SMALL
        if dt != dt0:
            grid.set_time(t, dt, grid.nstep())
        diag("step")  # diag(DUMP_OUT) follows every step, the last one included
        if last or (max_steps is not None and steps >= max_steps):
            break
    diag("final")
    t, _ = grid.time()
    return {"t": t, "steps": steps, "dumps": dump_c.count, "ener_lines": ener_c.count, "rdumps": rdump_c.count}
