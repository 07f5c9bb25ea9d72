"""OOMMF .odt data tables of a running simulation.

Counterpart of the reference's DataTableLog / LogStepHandler
(src/magnum/micromagnetics/stephandler/data_table_log.py:20-78,
log_stephandler.py:22-130): columns of the simulation time, the step size, the
average magnetisation and the wall time by default, energy columns on request,
one row per handled state, in the same header and row layout (values with
their column's format, tab-separated).

A "state" here is anything with the attributes the column functions read; for
the fused stepper, `StepperState(stepper)` gives t, h, step, M.average() and
the energies E_stray, E_exch, E_aniso, E_ext, E_tot (computed when read).
"""
import os
import time

import numpy as np


class _Avg:
    def __init__(self, stepper):
        self._st = stepper

    def average(self):
        return tuple(float(v) for v in self._st.average_m())


class StepperState:
    """The attributes of the reference's State that the columns read, from a
    micromagnum_amd.solver.Stepper."""

    def __init__(self, stepper, step=None):
        s = stepper.stats()
        self.t = float(s["t"])
        self.h = float(s["h"])
        self.step = step
        self.M = _Avg(stepper)
        self._stepper = stepper
        self._energy = None

    def __getattr__(self, name):
        if name.startswith("E_"):
            if self._energy is None:
                self._energy = self._stepper.energy()
            if name in self._energy:
                return float(self._energy[name])
        raise AttributeError(name)


class LogStepHandler:
    """Columns of (id, description, unit, format) filled by one function per
    group of columns (log_stephandler.py:22-130)."""

    class MultiColumn:
        class Column:
            def __init__(self, id, desc="(no descr.)", unit="unknown", fmt="%r"):
                self.id, self.desc, self.unit, self.fmt = id, desc, unit, fmt

        def __init__(self, fn):
            self.columns = []
            self.func = fn

        def addColumn(self, *spec):
            self.columns.append(LogStepHandler.MultiColumn.Column(*spec))

    def __init__(self, out_stream):
        self.f = out_stream
        self.first_call = True
        self.columns = []
        self._walltime0 = time.time()

    def addColumn(self, *args):
        if not len(args) >= 2:
            raise TypeError("Not enough arguments.")
        func, col_spec = args[-1], args[:-1]
        if not callable(func):
            raise TypeError("Last argument must be a callable function")
        mc = LogStepHandler.MultiColumn(func)
        for spec in col_spec:
            mc.addColumn(*(spec if isinstance(spec, tuple) else (spec,)))
        self.columns.append(mc)

    def addTimeColumn(self):
        self.addColumn(("t", "time", "s", "%r"), lambda state: state.t)

    def addWallTimeColumn(self):
        self.addColumn(("t_wall", "wall clock time", "s", "%.3f"), lambda state: time.time() - self._walltime0)

    def addStepSizeColumn(self):
        self.addColumn(("h", "step size", "s", "%r"), lambda state: state.h)

    def addAverageMagColumn(self):
        self.addColumn(("Mx", "average magn. x", "A/m", "%r"), ("My", "average magn. y", "A/m", "%r"),
                       ("Mz", "average magn. z", "A/m", "%r"), lambda state: state.M.average())

    def addEnergyColumn(self, energy_variable):
        self.addColumn((energy_variable, "%s-energy" % energy_variable, "J", "%r"),
                       lambda state: getattr(state, energy_variable))

    def handle(self, state):
        if self.first_call:
            self.writeHeader()
            self.first_call = False
        self.f.write(self.generateRow(state) + "\n")
        self.f.flush()

    def generateRow(self, state):
        raise NotImplementedError

    def writeHeader(self):
        raise NotImplementedError

    def writeAppendix(self):
        self.f.write("# done\n")

    def done(self):
        self.writeAppendix()


def _py(v):
    # repr of a Python number (a numpy scalar's repr names its type)
    return v.item() if isinstance(v, np.generic) else v


class DataTableLog(LogStepHandler):
    """OOMMF .odt table (data_table_log.py:20-78): time, step size, average
    magnetisation and wall time columns by default."""

    def __init__(self, path, title="(no title)"):
        path = os.path.normpath(path)
        d = os.path.dirname(path)
        if d and not os.path.exists(d):
            os.makedirs(d)
        self.file = open(path, "w+")
        super().__init__(self.file)
        self.addTimeColumn()
        self.addStepSizeColumn()
        self.addAverageMagColumn()
        self.addWallTimeColumn()
        self._title = title

    def done(self):
        super().done()
        self.file.close()

    def writeHeader(self):
        def pad(txt, width):
            return txt + " " * (width - len(txt))

        num_columns = 0
        col_spec, unit_spec = "", ""
        for mc in self.columns:
            for col in mc.columns:
                col_spec += pad("{" + col.desc + "}", 22) + "    "
                unit_spec += pad("{" + col.unit + "}", 22) + "    "
                num_columns += 1
        self.file.write("# ODT 1.0\n")
        self.file.write("# Table Start\n")
        self.file.write("# Title: %s\n" % self._title)
        self.file.write("#\n")
        self.file.write("## Generated by magnum.DataTableStepHandler.\n")
        self.file.write("## Number of columns: %s\n" % num_columns)
        self.file.write("#\n")
        self.file.write("# Columns:   " + col_spec + "\n")
        self.file.write("# Units:     " + unit_spec + "\n")

    def writeAppendix(self):
        self.file.write("# Table End\n")

    def generateRow(self, state):
        row = ""
        for mc in self.columns:
            values = mc.func(state)
            if type(values) != tuple:
                values = (values,)
            for n, col in enumerate(mc.columns):
                row += (col.fmt % _py(values[n])) + "\t"
        return row


def read_odt(path):
    """Column descriptions and the rows of an .odt table as a float array."""
    cols, rows = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("# Columns:"):
                body = line[len("# Columns:"):]
                cols = [c.strip() for c in body.split("}") if c.strip()]
                cols = [c.lstrip("{") for c in cols]
            elif line.startswith("#") or not line.strip():
                continue
            else:
                rows.append([float(v) for v in line.split()])
    return cols, np.array(rows)


def run_logged(stepper, steps, log, every=1):
    """Run `steps` accepted steps, handing the state to `log` before the
    first step and after every `every` steps."""
    log.handle(StepperState(stepper, 0))
    done = 0
    while done < steps:
        n = min(every, steps - done)
        stepper.run(n)
        done += n
        log.handle(StepperState(stepper, done))
