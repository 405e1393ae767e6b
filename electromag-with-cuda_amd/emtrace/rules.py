"""Step rules of the tracer and their parameters.

* ``PLAIN`` -- x += E/(|E|*resolution): the reference's plain step
  (ElectroMag/src/CPU_Implement.cpp:89-90, include/Vector.h:158-164).
* ``REF``   -- the reference's live default, CalcField_CPU_T_Curvature<float>
  (CPU_Implement.cpp:322-331): k = |(E-P) x P| / |E|^3 with P the line's seed
  point, x += E/(|E|*(k+1)*resolution), E including k_e = 8.987551787e9.
* ``ANGLE`` -- curvature-adaptive rule with no reference counterpart: the chord
  c = |u_s - u_{s-1}| between successive unit field directions bounds the step,
  h_s = clamp(turn*h_{s-1}/c, hmin, hmax), h_1 = h0.  With ``stop`` (the
  default for this rule) a line ends at the first step whose direction
  reverses (u_s . u_{s-1} < 0: it went through a sink); its remaining rows
  repeat the end point and its step count is the last step that moved it.

Rows always run 0..len-1 (row 0 = seeds); without stop every line takes
exactly len-1 steps, as in the reference (CPU_Implement.cpp:49,75).
"""
from __future__ import annotations

from dataclasses import dataclass

PLAIN, REF, ANGLE = 0, 1, 2
NAMES = {PLAIN: "plain", REF: "ref", ANGLE: "angle"}
K_E = 8.987551787e9  # include/Electrostatics.h:27


@dataclass(frozen=True)
class StepRule:
    rule: int = ANGLE
    resolution: float = 10.0   # PLAIN/REF: step = 1/resolution
    h0: float = 0.1            # ANGLE: first step
    hmin: float = 0.005
    hmax: float = 0.5
    turn: float = 0.05         # ANGLE: target chord between successive directions
    stop: bool = False         # ANGLE: end lines whose direction reverses

    @property
    def field_scale(self) -> float:
        """Factor applied to the Coulomb sum: k_e where the rule sees |E|."""
        return K_E if self.rule in (PLAIN, REF) else 1.0


def angle(h0=0.1, hmin=0.005, hmax=0.5, turn=0.05, stop=True) -> StepRule:
    return StepRule(ANGLE, 10.0, h0, hmin, hmax, turn, bool(stop))


def plain(resolution=10.0) -> StepRule:
    return StepRule(PLAIN, resolution)


def ref(resolution=10.0) -> StepRule:
    return StepRule(REF, resolution)
