"""PIC_loop: the step orchestration of chimeraCL (chimeraCL/pic_loop.py:30-142).

The order of operations is the reference's.  Adjacent reference calls are
fused, with identical arithmetic:
  * push_coords + the cell keys of the following sort_parts -> one pass;
  * with one solver, the J deposit + the second push_coords + its keys ->
    one pass;
  * fb_transform(dir=0) of rho, J + fields_smooth -> filter applied in the
    transform's last pass;
  * the dN0 <- dN1 copy + field_grad -> one assembly kernel;
  * push_fields + damp_fields -> one row kernel (PSATD, then the damping
    round trip through (kr, x) space).
restore_B_fb computes rot(G)/w^2 in one pass.  With ``graph=True`` the step
is captured once into a HIP graph (after one eager step) and replayed; the
step then makes no host synchronisation and no allocation.

Timing (timit=True) uses HIP events around each phase, keyed by the
reference's loop_steps names.
"""
from __future__ import annotations

import os

import numpy as np
import torch


def counter_collection_active():
    """True inside `rocprofv3 --pmc ...` (the profiler exports
    ROCPROF_COUNTER_COLLECTION to the program it runs)."""
    v = os.environ.get("ROCPROF_COUNTER_COLLECTION", "")
    return v.strip().lower() not in ("", "0", "false", "no")

loop_steps = ["frame", "push-x", "sort", "depose", "transform", "smooth", "data_copy",
              "grad", "push-eb", "damp-eb", "restore_B", "gather + push-p"]


class PIC_loop:
    def __init__(self, solvers=[], species=[], frames=[], diags=[], timit=False, graph=False,
                 align_every=0):
        """align_every > 0 (an addition, off by default): every align_every
        steps, at the start of the step, the mobile species are permuted into
        their current sorted order (Particles.reorder_parts), so that their
        memory order keeps following the cells."""
        self.align_every = align_every
        self.solvers = solvers
        self.mainsolver = self.solvers[0]
        self.species = species
        self.frames = frames
        self.diags = diags
        self.timit = timit
        self.it = 0
        self.graph = graph
        self._graphs = {}    # (consume, produce) -> captured step
        self._pre = False    # the last step left the next first half push
        self.Timer = {s: 0.0 for s in loop_steps}
        self._ev = []

    # ---------------------------------------------------------------- timing
    def timer_start(self):
        if self.timit:
            self._t0 = torch.cuda.Event(enable_timing=True)
            self._t0.record()

    def timer_record(self, name):
        if self.timit:
            t1 = torch.cuda.Event(enable_timing=True)
            t1.record()
            self._ev.append((name, self._t0, t1))

    def _flush_timers(self):
        if self._ev:
            torch.cuda.synchronize()
            for name, a, b in self._ev:
                self.Timer[name] += a.elapsed_time(b) * 1e-3
            self._ev = []

    # ---------------------------------------------------------------- step
    def _fused(self):
        """One solver: its J deposit also applies the second half push and
        the sort keys (the same positions and keys as push_coords).  Its
        writes go through sort_indx, so it pays while the particles' memory
        order follows their cells: always in FP64, in FP32 with align_every
        (measured on MI355X, 16.6M electrons)."""
        return len(self.solvers) == 1 and (self.align_every > 0 or
                                           not self.mainsolver.comm.mixed)

    def _body(self, consume=False, produce=False):
        """One step.  consume: the previous step's gather left this step's
        first half push (positions xh, yh, zh and their keys); produce: this
        step's gather leaves the next one's (Grid.gather_and_push(prepush))."""
        self.timer_start()
        for frame in self.frames:
            if np.mod(self.it, frame.Args["Steps"]) == 0:
                frame.shift_grids(grids=self.solvers)
                frame.inject_plasma(species=self.species, grid=self.mainsolver)
        self.timer_record("frame")
        for parts in self.species:
            self.timer_start()
            if consume:
                parts.use_prepushed(self.mainsolver)
            else:
                parts.push_coords(mode="half", grid=self.mainsolver)
            self.timer_record("push-x")
            self.timer_start()
            parts.sort_parts(grid=self.mainsolver)
            self.timer_record("sort")
        fused = self._fused()
        self.timer_start()
        for solver in self.solvers:
            solver.depose_currents(species=self.species, push_half=fused, half_in=consume)
        self.timer_record("depose")
        for parts in self.species:
            if not fused:
                self.timer_start()
                parts.push_coords(mode="half", grid=self.mainsolver)
                self.timer_record("push-x")
            self.timer_start()
            parts.sort_parts(grid=self.mainsolver)
            self.timer_record("sort")
        for solver in self.solvers:
            self.timer_start()
            solver.depose_charge(species=self.species)
            self.timer_record("depose")
            self.timer_start()
            solver.fb_transform_smooth(scals=["rho"], vects=["J"])
            self.timer_record("transform")
            self.timer_start()
            solver.field_grad("rho", "dN1", copy_to="dN0")
            self.timer_record("grad")
            self.timer_start()
            solver.damp_edges(("E", "G"), advance_with=("J", "dN0", "dN1"))
            self.timer_record("push-eb")
            self.timer_start()
            solver.restore_B_fb()
            self.timer_record("restore_B")
            self.timer_start()
            solver.fb_transform(vects=["E", "B"], dir=1)
            self.timer_record("transform")
            self.timer_start()
            solver.gather_and_push(species=self.species, prepush=produce)
            self.timer_record("gather + push-p")

    def step(self):
        """One PIC step (pic_loop.py:57-142)."""
        return self._step(self._pre, False)

    def run(self, nsteps):
        """nsteps consecutive steps.  Inside the call nothing else touches the
        particles between two steps, so (one solver, the fused deposit) each
        step's gather also applies the next step's first half push and its
        sort keys (chm_gather_push_pre_*), except before a step with a frame
        event or a diagnostics record, and before the call returns: the state
        after run(n) is the state after n calls of step(), bit for bit."""
        for i in range(nsteps):
            produce = (i < nsteps - 1 and self._fused() and self._plain_step(self.it + 1))
            self._step(self._pre, produce)
        return self.it

    def _plain_step(self, it):
        """No frame event and no diagnostics record at step it."""
        for frame in self.frames:
            if np.mod(it, frame.Args["Steps"]) == 0:
                return False
        for diag in self.diags:
            if np.mod(it, diag.Args.get("Interval", 1)) == 0:
                return False
        return True

    def _step(self, consume, produce):
        self.pre_step(consume)
        key = (consume, produce)
        if (self.graph and key not in self._graphs and self.it >= 1 and not self.frames
                and not self.timit):
            self._capture(key)
        g = self._graphs.get(key) if self.graph else None
        if g is not None:
            g.replay()
            for parts in self.species:
                if parts.Args["Np"] > 0 and "Immobile" not in parts.Args:
                    parts.Args.set_lazy("Np_stay", (lambda t: lambda: int(t.item()))(parts.DataDev["Np_stay"]))
        else:
            self._body(consume, produce)
            self._flush_timers()
        self._pre = produce
        self.it += 1
        return self.it

    def pre_step(self, consume=False):
        """What a step does before its kernels: diagnostics records, then
        (align_every) the reorder of the mobile species (with the pre-pushed
        positions and keys when the step consumes them)."""
        for diag in self.diags:
            diag.make_record(self.it)
        if self.align_every:
            for parts in self.species:
                if "Immobile" not in parts.Args:
                    # the gather reads pid from the first step on, so that a
                    # captured step does not change when the first reorder
                    # creates it
                    parts.ensure_pid()
                    if self.it > 0 and self.it % self.align_every == 0:
                        # consuming the pre-push, the step's J deposit
                        # writes every x, y, z before anything reads them
                        parts.reorder_parts(
                            ("xh", "yh", "zh", "indx_in_cell") if consume else (),
                            ("x", "y", "z") if consume else ())

    def _capture(self, key):
        # A rocprofv3 counter-collection run of the graph-replayed step ended
        # in a segmentation fault in one of the profiler's threads as soon as
        # the first captured steps were replayed (the same steps run eagerly
        # profile normally), so graph mode refuses to run under counter
        # collection: collect counters with graph=False, which launches the
        # same kernels in the same order.
        if counter_collection_active():
            raise RuntimeError(
                "PIC_loop(graph=True) under rocprofv3 counter collection "
                "(ROCPROF_COUNTER_COLLECTION is set): HIP graph replay is not supported by the "
                "profiler's counter collection here; run with graph=False (eager steps)")
        # one eager step has run, so every buffer exists; ions stay sorted
        # (they are never pushed), electrons are sorted twice per step
        for parts in self.species:
            if key[1] and "Immobile" not in parts.Args and parts.Args["Np"] > 0:
                parts.ensure_prepush_buffers(self.mainsolver)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self._body(*key)
        torch.cuda.current_stream().wait_stream(s)
        self._graphs[key] = g
