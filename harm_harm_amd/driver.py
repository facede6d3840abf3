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
    """One output type's schedule: tdumpgen, DTdumpgen, dumpcntgen, tlastgen.

    A fresh run writes at t0 (RESTARTMODE 0, diag.c:111-113); a resumed run
    writes next at the first multiple of DT after t0 (diag.c:137-139) and
    numbers its files from `count` (dumpcntgen as the restart file kept it).
    at_init: the type is also written by the initial call whatever its next
    time (the energy file, diag.c:683; main dumps only when due, :586)."""

    def __init__(self, DT, t0, resume=False, count=0, at_init=False):
        self.DT = DT
        self.count = count
        self.at_init = at_init
        self.tlast = None
        if DT is not None and resume:
            self.tnext = DT * (int(t0 / DT) + 1)
            self.tlast = DT * int(t0 / DT) - SMALL
        else:
            self.tnext = t0

    def due(self, t, call):
        """call: "init", "step" or "final" (INIT_OUT, DUMP_OUT, FINAL_OUT)"""
        if self.DT is None or t == self.tlast:
            return False
        return t >= self.tnext or call == "final" or (call == "init" and self.at_init)

    def done(self, t):
        # post_dump: dumpc = 1 + max(0, (long long)((t - tdump)/DT));
        # tdump = (ROUND2LONGLONGINT(tdump/DT) + dumpc) DT
        dumpc = 1 + max(0, int((t - self.tnext) / self.DT))
        self.tnext = (math.floor(self.tnext / self.DT + 0.5) + dumpc) * self.DT
        self.count += 1
        self.tlast = t


class RestartCadence:
    """Restart dumps by step count: nrestart, nlastrestart, whichrestart.

    nlastrestart starts at DTr (n0 / DTr) - 1 (diag.c:98); a fresh run has
    nrestart = n0, so it writes rdump-0 at once (:111), a resumed one the next
    multiple of DTr (:134) and continues the 0/1 alternation from `which`."""

    def __init__(self, DTr, n0, resume=False, which=0):
        self.DTr = DTr
        self.which = which
        self.count = 0
        if DTr is not None:
            self.nlast = DTr * (n0 // DTr) - 1
            self.nnext = DTr * (n0 // DTr + 1) if resume else n0

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


def _resume_counters(outdir):
    """what a resumed run continues from in outdir/dumps: the number of the
    next main dump (one past the highest dumpNNNN) and the next restart slot
    (the other one than the rdump-0/1 with the later step count)"""
    d = os.path.join(outdir, "dumps")
    nums = [int(n[4:]) for n in os.listdir(d) if n.startswith("dump") and n[4:].isdigit()] if os.path.isdir(d) else []
    steps = {}
    for w in (0, 1):
        f = os.path.join(d, "rdump-%d" % w)
        if os.path.exists(f):
            with open(f, "rb") as fh:
                steps[w] = int(fh.readline().split()[5])
    which = 0 if not steps else 1 - max(steps, key=lambda w: steps[w])
    return (max(nums) + 1 if nums else 0), which


def run(grid, tf, DTdump=None, DTener=None, outdir=".", gdump=True, max_steps=None, which="p", DTr=None,
        restart_binary=True, resume=False):
    """Advance `grid` (harm_harm_amd.grid.Grid, clock already set) to tf with
    HARM's end-of-run rule, writing dumps every DTdump, energy-file lines
    every DTener and restart files every DTr steps into outdir (restart
    files MIXEDOUTPUT when restart_binary, else TEXTOUTPUT).

    resume: the grid continues a run of the same outdir (RESTARTMODE 1,
    after harm_read_restart with tracking switched on): ener.out is appended
    to and the energy sums the restart file restored go on accumulating, the
    next output times are the first multiples of DT after t (no initial
    outputs), main dumps are numbered on from the highest existing one, and
    restart files continue the rdump-0/1 alternation.  Returns a summary
    dict."""
    os.makedirs(os.path.join(outdir, "dumps"), exist_ok=True)
    t0, _ = grid.time()
    first_dump, first_which = _resume_counters(outdir) if resume else (0, 0)
    dump_c = Cadence(DTdump, t0, resume, first_dump)
    ener_c = Cadence(DTener, t0, resume, at_init=True)
    rdump_c = RestartCadence(DTr, grid.nstep(), resume, first_which)
    if DTener is not None:
        grid.set_ener_tracking(True)  # keeps the restored sums when tracking is on already
    ener_path = os.path.join(outdir, "ener.out")
    if DTener is not None and not resume:
        open(ener_path, "w").close()
    if gdump and not resume:
        grid.write_gdump(os.path.join(outdir, "dumps", "gdump"))

    def diag(call):
        t, _ = grid.time()
        if dump_c.due(t, call):
            if call == "step":
                # post_stepch computes the dump's four-current after a step
                # that ends at a main dump (step_ch.c:156-170)
                grid.currents()
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

    grid.currents(at_start=True)  # post_init's compute_currents_t0 (initbase.c:1373)
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
    return {"t": t, "steps": steps, "dumps": dump_c.count - first_dump, "ener_lines": ener_c.count,
            "rdumps": rdump_c.count}
