"""Headless counterpart of the reference's interactive loop: the render loop
of its examples (main/examples/hy.rs:185-205, eu.rs:139-159), the viewer
crate's isotropic camera controller (viewer/src/controllers/isotropic.rs,
controllers/mod.rs) and its input handler (viewer/src/handler.rs:63-181).

There is no window here: a frame's input is an `Input` record (keys held,
relative mouse motion, left button, wheel, and the discrete events quit /
escape / tab), which an interactive front end would fill from its event
pump, and a finished frame goes to a `present(rgba8)` callback instead of
`Window::draw` (viewer/src/window.rs:80-95). The loop's calls into the
pipeline are the reference's: a camera change cleans the canvas (seeds
kept), every frame re-uploads the scene, then renders passes for the frame
time and reads the RGBA8 image.

Camera maps: hyperbolic ones are Moebius transforms built from the
generators of geometry/hyperbolic.cc:35-62 (objects.hy_*); Euclidean ones
are Homogenous3 (shift after rotation) built like eu_shift / eu_rotate
(geometry/euclidean.cc:35-40) with quaternions as transform/rotation.cc:60-76.
The reference's host maps come from its `ccgeom` crate, which is not in
the reference tree; these follow the kernel's own generator formulas.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, FrozenSet, Iterable, Optional, Tuple

import numpy as np

from .objects import (Homogenous3, Moebius, Rotation3, hy_xrotate, hy_xshift, hy_yrotate, hy_yshift, hy_zrotate,
                      hy_zshift)

# viewer/src/controllers/isotropic.rs:51-54
SPEED = 1.0
SENS = 1.0
DT = 0.04


@dataclass(frozen=True)
class Motion:
    """A control input (controllers/mod.rs:7-11): an absolute amount
    (`position`, e.g. mouse motion) or a rate applied over one DT step
    (`velocity`, e.g. a held key)."""
    kind: str
    value: float

    @staticmethod
    def position(x: float) -> "Motion":
        return Motion("position", float(x))

    @staticmethod
    def velocity(x: float) -> "Motion":
        return Motion("velocity", float(x))

    def amount(self) -> float:
        """isotropic.rs:56-61"""
        return self.value if self.kind == "position" else self.value * DT


# ------------------------------------------------------------------ geometry maps

def _qmul(a, b):
    """Hamilton product of quaternions stored (w, x, y, z)."""
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return (w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
            w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
            w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
            w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2)


def rotate_vector(q, v) -> Tuple[float, float, float]:
    """q (0, v) q* (rot3_apply_pos, transform/rotation.cc:64-66)."""
    c = (q[0], -q[1], -q[2], -q[3])
    r = _qmul(_qmul(q, (0.0, v[0], v[1], v[2])), c)
    return (r[1], r[2], r[3])


class Euclidean3:
    """Euclidean camera maps as Homogenous3 (shift after rotation)."""

    @staticmethod
    def identity() -> Homogenous3:
        return Homogenous3((0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0))

    @staticmethod
    def _shift(v) -> Homogenous3:
        return Homogenous3(tuple(float(x) for x in v), (1.0, 0.0, 0.0, 0.0))

    @staticmethod
    def _rotate(axis, phi) -> Homogenous3:
        s = np.sin(phi / 2)
        return Homogenous3((0.0, 0.0, 0.0), (np.cos(phi / 2), s * axis[0], s * axis[1], s * axis[2]))

    shift_x = staticmethod(lambda l: Euclidean3._shift((l, 0.0, 0.0)))
    shift_y = staticmethod(lambda l: Euclidean3._shift((0.0, l, 0.0)))
    shift_z = staticmethod(lambda l: Euclidean3._shift((0.0, 0.0, l)))
    rotate_x = staticmethod(lambda phi: Euclidean3._rotate((1.0, 0.0, 0.0), phi))
    rotate_y = staticmethod(lambda phi: Euclidean3._rotate((0.0, 1.0, 0.0), phi))
    rotate_z = staticmethod(lambda phi: Euclidean3._rotate((0.0, 0.0, 1.0), phi))

    @staticmethod
    def chain(a: Homogenous3, b: Homogenous3) -> Homogenous3:
        """a after b: x -> a(b(x)) = a.shift + Ra (b.shift + Rb x)."""
        moved = rotate_vector(a.rot.q, b.shift.v)
        shift = tuple(a.shift.v[i] + moved[i] for i in range(3))
        return Homogenous3(shift, Rotation3(_qmul(a.rot.q, b.rot.q)))

    @staticmethod
    def apply(m: Homogenous3, p) -> Tuple[float, float, float]:
        r = rotate_vector(m.rot.q, p)
        return tuple(m.shift.v[i] + r[i] for i in range(3))


class Hyperbolic3:
    """Hyperbolic camera maps as Moebius transforms (geometry/hyperbolic.cc:35-62)."""

    @staticmethod
    def identity() -> Moebius:
        return Moebius.identity()

    shift_x = staticmethod(hy_xshift)
    shift_y = staticmethod(hy_yshift)
    shift_z = staticmethod(hy_zshift)
    rotate_x = staticmethod(hy_xrotate)
    rotate_y = staticmethod(hy_yrotate)
    rotate_z = staticmethod(hy_zrotate)

    @staticmethod
    def chain(a: Moebius, b: Moebius) -> Moebius:
        return a.chain(b)


# ------------------------------------------------------------------ controller

class IsotropicController:
    """The viewer's isotropic camera controller (isotropic.rs): every motion
    chains a map onto the camera's map in the camera's own frame; zoom
    scales by 1.2 per wheel step; `updated` stays true until the next
    `step`. `geometry` is Euclidean3 or Hyperbolic3."""

    def __init__(self, geometry, map=None, zoom: float = 1.0):
        self.geometry = geometry
        self._map = map if map is not None else geometry.identity()
        self._zoom = float(zoom)
        self._updated = True

    def update_map(self, m) -> None:
        self._map = self.geometry.chain(self._map, m)
        self._updated = True

    def map(self):
        return self._map

    def zoom(self) -> float:
        return self._zoom

    def updated(self) -> bool:
        return self._updated

    # Controller (controllers/mod.rs:13-26)
    def move_forward(self, m: Motion) -> None:
        self.update_map(self.geometry.shift_z(SPEED * -m.amount()))

    def move_right(self, m: Motion) -> None:
        self.update_map(self.geometry.shift_x(SPEED * m.amount()))

    def move_up(self, m: Motion) -> None:
        self.update_map(self.geometry.shift_y(SPEED * m.amount()))

    def rotate_yaw(self, m: Motion) -> None:
        self.update_map(self.geometry.rotate_y(SENS * -m.amount()))

    def rotate_pitch(self, m: Motion) -> None:
        self.update_map(self.geometry.rotate_x(SENS * -m.amount()))

    def rotate_roll(self, m: Motion) -> None:
        self.update_map(self.geometry.rotate_z(SENS * -m.amount()))

    def zoom_by(self, m: Motion) -> None:
        """`zoom` of the reference's trait (the accessor keeps the name here)."""
        self._zoom *= 1.2 ** m.amount()
        self._updated = True

    def zoom_alt(self, m: Motion) -> None:
        """Not implemented by the reference either (isotropic.rs:87-89)."""

    def step(self, dt: float) -> None:
        self._updated = False


# ------------------------------------------------------------------ input

@dataclass
class Input:
    """One frame of input state: `keys` held (names of the reference's
    scancodes: "W", "A", "S", "D", "Up", "Down", "Left", "Right", "Space",
    "C", "E", "Q", "LShift", "RShift"), relative mouse motion, the left
    button, the wheel's steps since the last frame, and the discrete events
    of the frame ("quit", "escape", "tab")."""
    keys: FrozenSet[str] = frozenset()
    mouse_rel: Tuple[int, int] = (0, 0)
    left: bool = False
    wheel: int = 0
    events: Tuple[str, ...] = field(default_factory=tuple)


class Handler:
    """The viewer's input handler (handler.rs:63-181) over `Input` records:
    events first (quit or escape end the loop, tab toggles mouse capture and
    drops the next relative motion), then held keys, then mouse."""

    def __init__(self, size: Tuple[int, int]):
        self.size = (int(size[0]), int(size[1]))
        self.capture = False
        self.drop_move = False

    def set_capture(self, capture: bool) -> None:
        self.capture = capture
        if capture:
            self.drop_move = True

    def poll(self, controller, inp: Input) -> str:
        """Returns "exit" or "continue" (PollStatus)."""
        toggle = False
        for ev in inp.events:
            if ev in ("quit", "escape"):
                return "exit"
            if ev == "tab":
                toggle = True
        self.handle_keys(controller, inp.keys)
        if toggle:
            self.set_capture(not self.capture)
        if not self.drop_move:
            self.handle_mouse(controller, inp)
        else:
            self.drop_move = False
        return "continue"

    def handle_mouse(self, controller, inp: Input) -> None:
        shift = "LShift" in inp.keys or "RShift" in inp.keys
        dx, dy = inp.mouse_rel
        if self.capture or inp.left:
            if dx != 0:
                controller.rotate_yaw(Motion.position(2.0 * dx / self.size[0]))
            if dy != 0:
                controller.rotate_pitch(Motion.position(2.0 * dy / self.size[1]))
        if inp.wheel != 0:
            if not shift:
                controller.zoom_by(Motion.position(inp.wheel))
            else:
                controller.zoom_alt(Motion.position(inp.wheel))

    @staticmethod
    def handle_keys(controller, keys: FrozenSet[str]) -> None:
        if "W" in keys or "Up" in keys:
            controller.move_forward(Motion.velocity(1.0))
        if "A" in keys or "Left" in keys:
            controller.move_right(Motion.velocity(-1.0))
        if "S" in keys or "Down" in keys:
            controller.move_forward(Motion.velocity(-1.0))
        if "D" in keys or "Right" in keys:
            controller.move_right(Motion.velocity(1.0))
        if "Space" in keys:
            controller.move_up(Motion.velocity(1.0))
        if "C" in keys:
            controller.move_up(Motion.velocity(-1.0))
        if "E" in keys:
            controller.rotate_roll(Motion.velocity(1.0))
        if "Q" in keys:
            controller.rotate_roll(Motion.velocity(-1.0))


# ------------------------------------------------------------------ loop

def run(pipeline, scene, controller: IsotropicController, handler: Handler, inputs: Iterable[Input],
        present: Callable[[np.ndarray], None], delay: float = 0.04,
        passes_per_frame: Optional[int] = None) -> int:
    """The examples' render loop (hy.rs:185-205): per frame, poll the input
    (stop on exit or when `inputs` runs out); if the camera changed, put the
    controller's map and 1 / zoom into the scene's view (a `Mapped` over a
    `PointView`) and clean the canvas (the seeds stay); upload the scene;
    render single passes for `delay` seconds (or exactly `passes_per_frame`
    of them, for a reproducible frame); hand the RGBA8 image to `present`.
    `pipeline` is a `processing.Pipeline` (or any object with its
    `clean`, `set_scene`, `render` and `get_image`). Returns the number of
    frames presented."""
    frames = 0
    for inp in inputs:
        if handler.poll(controller, inp) == "exit":
            break
        updated = controller.updated()
        controller.step(delay)
        if updated:
            scene.view.map = controller.map()
            scene.view.inner.fov = 1.0 / controller.zoom()
            pipeline.clean()
        pipeline.set_scene(scene)
        if passes_per_frame is not None:
            for _ in range(int(passes_per_frame)):
                pipeline.render()
        else:
            start = time.monotonic()
            while time.monotonic() - start < delay:
                pipeline.render()
        present(pipeline.get_image())
        frames += 1
    return frames
