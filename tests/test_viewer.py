"""The headless interactive loop (hypertrace_amd/viewer.py): the isotropic
controller's maps against the geometry they must realise, the input
handler's key and mouse bindings (viewer/src/handler.rs:63-181), and the
loop's calls into the pipeline (main/examples/hy.rs:185-205), on the CPU."""
import numpy as np
import pytest

from hypertrace_amd import scenes
from hypertrace_amd.objects import hy_zshift
from hypertrace_amd.viewer import (DT, Euclidean3, Handler, Hyperbolic3, Input, IsotropicController, Motion,
                                   rotate_vector, run)


def test_motion_amounts():
    assert Motion.position(0.5).amount() == 0.5
    assert Motion.velocity(1.0).amount() == pytest.approx(DT)


def test_euclidean_rotations_and_chain_act_on_points():
    th = 0.7
    # a right-handed rotation about x takes y to (0, cos, sin)
    np.testing.assert_allclose(Euclidean3.apply(Euclidean3.rotate_x(th), (0, 1, 0)), (0, np.cos(th), np.sin(th)),
                               atol=1e-15)
    np.testing.assert_allclose(Euclidean3.apply(Euclidean3.rotate_y(th), (0, 0, 1)), (np.sin(th), 0, np.cos(th)),
                               atol=1e-15)
    np.testing.assert_allclose(Euclidean3.apply(Euclidean3.rotate_z(th), (1, 0, 0)), (np.cos(th), np.sin(th), 0),
                               atol=1e-15)
    # chain(a, b) applies b first
    rng = np.random.default_rng(3)
    a = Euclidean3.chain(Euclidean3.shift_x(0.3), Euclidean3.rotate_y(1.1))
    b = Euclidean3.chain(Euclidean3.rotate_z(-0.4), Euclidean3.shift_z(2.0))
    for _ in range(5):
        p = rng.normal(size=3)
        np.testing.assert_allclose(Euclidean3.apply(Euclidean3.chain(a, b), p),
                                   Euclidean3.apply(a, Euclidean3.apply(b, p)), atol=1e-12)
    # rotations preserve length
    v = rotate_vector(Euclidean3.rotate_x(0.3).rot.q, (1.0, 2.0, 3.0))
    assert np.linalg.norm(v) == pytest.approx(np.sqrt(14.0))


def test_controller_chains_in_the_camera_frame():
    m0 = Hyperbolic3.rotate_x(0.4)
    c = IsotropicController(Hyperbolic3, m0, 1.0)
    assert c.updated()
    c.step(0.04)
    assert not c.updated()
    c.move_forward(Motion.velocity(1.0))  # forward = -z in the camera frame, SPEED * DT
    assert c.updated()
    np.testing.assert_allclose(c.map().m, m0.m @ hy_zshift(-DT).m, rtol=1e-15)
    c.zoom_by(Motion.position(2))
    assert c.zoom() == pytest.approx(1.44)
    c.zoom_alt(Motion.position(1))  # not implemented by the reference either
    assert c.zoom() == pytest.approx(1.44)
    e = IsotropicController(Euclidean3)
    e.move_right(Motion.position(0.5))
    e.move_up(Motion.position(-0.25))
    np.testing.assert_allclose(e.map().shift.v, (0.5, -0.25, 0.0))


class Recorder:
    """Stands in for the controller and records the handler's calls."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        return lambda m: self.calls.append((name, m.kind, m.value))


def test_handler_bindings():
    h = Handler((200, 100))
    r = Recorder()
    assert h.poll(r, Input(keys=frozenset({"W", "A", "Space", "E"}))) == "continue"
    assert r.calls == [("move_forward", "velocity", 1.0), ("move_right", "velocity", -1.0),
                       ("move_up", "velocity", 1.0), ("rotate_roll", "velocity", 1.0)]
    r.calls.clear()
    h.poll(r, Input(keys=frozenset({"Down", "Right", "C", "Q"})))
    assert r.calls == [("move_forward", "velocity", -1.0), ("move_right", "velocity", 1.0),
                       ("move_up", "velocity", -1.0), ("rotate_roll", "velocity", -1.0)]
    # mouse motion rotates only while dragging (left button) or captured
    r.calls.clear()
    h.poll(r, Input(mouse_rel=(10, 5)))
    assert r.calls == []
    h.poll(r, Input(mouse_rel=(10, 5), left=True))
    assert r.calls == [("rotate_yaw", "position", 0.1), ("rotate_pitch", "position", 0.1)]
    # the wheel zooms; with shift held it is the unimplemented alternative zoom
    r.calls.clear()
    h.poll(r, Input(wheel=-2))
    h.poll(r, Input(wheel=1, keys=frozenset({"LShift"})))
    assert r.calls == [("zoom_by", "position", -2.0), ("zoom_alt", "position", 1.0)]
    # tab captures the mouse and drops that frame's motion, then motion rotates
    r.calls.clear()
    h.poll(r, Input(mouse_rel=(4, 0), events=("tab",)))
    assert h.capture and r.calls == []
    h.poll(r, Input(mouse_rel=(4, 0)))
    assert r.calls == [("rotate_yaw", "position", 0.04)]
    assert h.poll(r, Input(events=("escape",))) == "exit"
    assert h.poll(r, Input(events=("quit",))) == "exit"


class FakePipeline:
    def __init__(self):
        self.log = []

    def clean(self):
        self.log.append("clean")

    def set_scene(self, scene):
        self.log.append(("scene", scene.view.inner.fov))

    def render(self):
        self.log.append("render")

    def get_image(self):
        self.log.append("image")
        return np.zeros((1, 1, 4), np.uint8)


def test_loop_cleans_on_camera_changes_only():
    wl = scenes.reference_hyperbolic_example(32, 24)
    ctrl = IsotropicController(Hyperbolic3, wl.scene.view.map, 1.0)
    pipe = FakePipeline()
    frames = []
    inputs = [Input(), Input(), Input(wheel=1), Input(), Input(events=("quit",)), Input()]
    n = run(pipe, wl.scene, ctrl, Handler((32, 24)), inputs, frames.append, passes_per_frame=2)
    assert n == 4 and len(frames) == 4
    frame = ["render", "render", "image"]
    fov = 1.0 / 1.2
    assert pipe.log == (["clean", ("scene", 1.0)] + frame + [("scene", 1.0)] + frame
                        + ["clean", ("scene", pytest.approx(fov))] + frame + [("scene", pytest.approx(fov))] + frame)
    assert wl.scene.view.map is ctrl.map()
