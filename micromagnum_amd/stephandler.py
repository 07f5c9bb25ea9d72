"""Step handlers: objects the solver hands selected states to
(src/magnum/solver/step_handler.py:19-36, src/magnum/micromagnetics/stephandler/)."""
import sys


class StepHandler(object):
    """Base of all step handlers: handle(state) per selected state, done()
    when the solver is finished with it."""

    def handle(self, state):
        raise NotImplementedError("StepHandler.handle(): Implement me!")

    def done(self):
        pass


class ScreenLog(StepHandler):
    """One line per handled state: step, t, h, <M>."""

    def __init__(self, out=None):
        self.out = out or sys.stdout

    def handle(self, state):
        a = state.M.average()
        self.out.write("step=%d t=%.6e h=%.6e <M>=(%.6e, %.6e, %.6e)\n" % (state.step, state.t, state.h, *a))
