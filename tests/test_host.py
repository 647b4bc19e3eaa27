"""Host logic of the fused run: boundary-plan compilation, inlet packing,
chunking, and the observable buffer against a literal restatement."""
import math

import numpy as np
import pytest

from lbm_amd import _lib
from lbm_amd.domain import pack_inlets
from lbm_amd.run import PlanError, compile_plan


class _Rec:
    def __init__(self):
        self._recording = []

    def __getattr__(self, name):
        kinds = {"zou_he_left_wall_velocity": "left", "zou_he_right_wall_velocity": "right",
                 "zou_he_right_wall_pressure": "right_pressure",
                 "zou_he_top_wall_velocity": "top", "zou_he_bottom_wall_velocity": "bottom",
                 "zou_he_bottom_left_corner": "bottom_left", "zou_he_top_left_corner": "top_left",
                 "zou_he_top_right_corner": "top_right",
                 "zou_he_bottom_right_corner": "bottom_right"}
        if name in kinds:
            return lambda: self._recording.append((kinds[name], None))
        if name == "bounce_back_obstacle":
            return lambda o: self._recording.append(("obstacle", o))
        raise AttributeError(name)


@pytest.mark.parametrize("name,right,nobs", [("cavity", _lib.BC_VELOCITY, 0),
                                             ("poiseuille", _lib.BC_PRESSURE, 0),
                                             ("turek", _lib.BC_PRESSURE, 1),
                                             ("array", _lib.BC_PRESSURE, 8),
                                             ("step", _lib.BC_PRESSURE, 2)])
def test_reference_apps_compile_to_fused_plan(name, right, nobs):
    from lbm_amd.app.app import app_factory
    app = app_factory.create(name)
    rec = _Rec()
    app.set_bc(rec)
    bc, obstacles, first = compile_plan(rec._recording)
    assert bc == [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, right]
    assert len(obstacles) == nobs and first


def test_unsupported_plans_are_refused():
    with pytest.raises(PlanError):
        compile_plan([("left", None), ("right", None)])            # no top/bottom/corners
    full = [("bottom", None), ("left", None), ("top", None), ("right", None),
            ("bottom_left", None), ("top_left", None), ("top_right", None),
            ("bottom_right", None)]
    with pytest.raises(PlanError):
        compile_plan(full[:4] + [("bottom_left", None)] + [("top", None)] + full[5:])
    with pytest.raises(PlanError):
        compile_plan([("obstacle", 1)] + full[:2] + [("obstacle", 2)] + full[2:])
    bc, obs, first = compile_plan(full[:4] + [("obstacle", 3)] + full[4:])
    assert obs == [3] and not first


def test_pack_inlets_deduplicates_by_identity():
    a, b = np.arange(4.0), np.ones(3)
    pool, offs = pack_inlets([(a, b, a, b, a), (a, b, a, b, b)])
    assert offs.shape == (2, 5)
    assert pool.size == 1 + 4 + 3
    assert np.array_equal(pool[offs[0, 0]:offs[0, 0] + 4], a)
    assert np.array_equal(pool[offs[1, 4]:offs[1, 4] + 3], b)


def _literal_buff_series(values, dt, ct, nb):
    """buff.py:35-77 as written there (np.append, O(it) per call)."""
    buff = np.zeros([2]); a1 = np.zeros([2]); a2 = np.zeros([2]); a3 = np.zeros([2])
    it, cnt, cv, out = 0, 0, False, []
    for v in values:
        buff = np.append(buff, v)
        it += 1
        it_s, it_e = math.floor(it * 4 / 5), it
        obs = np.sum(buff[it_s:it_e]) / (float(it_e - it_s + 1))
        a1 = np.append(a1, obs)
        obs = np.sum(a1[it_s:it_e]) / (float(it_e - it_s + 1))
        a2 = np.append(a2, obs)
        obs = np.sum(a2[it_s:it_e]) / (float(it_e - it_s + 1))
        a3 = np.append(a3, obs)
        obs = np.sum(a3[it_s:it_e]) / (float(it_e - it_s + 1))
        growth = 0.0
        if it > 10:
            growth = (a3[it_e] - a3[it_s]) / ((it_e - it_s + 1) * dt)
            cnt = cnt + 1 if abs(growth) < ct else 0
            if cnt > nb:
                cv = True
        out.append((obs, growth, cv))
    return out


def test_buff_matches_reference_semantics():
    from lbm_amd.buff import buff
    rng = np.random.default_rng(0)
    vals = 3.0 + 0.1 * np.sin(np.arange(3000) / 40.0) + 1e-3 * rng.standard_normal(3000)
    b = buff("drag", 0.01, 1.0e-1, 50, "/tmp")
    ref = _literal_buff_series(vals, 0.01, 1.0e-1, 50)
    for v, (obs, growth, cv) in zip(vals, ref):
        b.add(v)
        o, g = b.mv_avg()
        assert o == obs and g == growth and b.obs_cv == cv


def test_chunking_follows_output_frequency():
    from lbm_amd.run import FusedRun

    class A:
        output_freq = 500
        stop = 'it'
        it_max = 1234

    fr = FusedRun.__new__(FusedRun)
    fr.app, fr.chunk = A(), 1000
    ends, it = [], 0
    while it <= A.it_max:
        it = fr._chunk_end(it)
        ends.append(it)
    assert ends == [500, 1000, 1235]


def test_pack_inlets_checks_side_shapes():
    nx, ny = 6, 4
    row = [np.zeros((2, nx)), np.zeros((2, ny)), np.zeros((2, nx)), np.zeros((2, ny)),
           np.ones(ny)]
    pool, offs = pack_inlets([tuple(row)], nx, ny)
    assert offs.shape == (1, 5)
    for sd, bad in enumerate([np.zeros((2, nx + 1)), np.zeros((2, ny, 1)), np.zeros(nx),
                              np.zeros((3, ny)), np.ones(ny + 1)]):
        r = list(row)
        r[sd] = bad
        with pytest.raises(ValueError):
            pack_inlets([tuple(r)], nx, ny)
    with pytest.raises(ValueError):
        pack_inlets([tuple(row[:4])], nx, ny)


def test_lattice_walls_check_side_shapes():
    """The per-call Zou-He methods refuse wall arrays of the wrong shape
    before any kernel launch (the reference raises on a mis-shaped array)."""
    from lbm_amd.lattice import lattice
    from tests.cases import small_cavity
    lat = lattice(small_cavity())
    lat.u_left = np.zeros((2, lat.ny - 1))
    with pytest.raises(ValueError, match="left"):
        lat.zou_he_left_wall_velocity()
    lat.u_top = np.zeros(lat.nx)
    with pytest.raises(ValueError, match="top"):
        lat.zou_he_top_wall_velocity()
    lat.u_right = np.zeros((2, lat.ny))
    lat.rho_right = np.zeros(lat.ny + 2)
    with pytest.raises(ValueError, match="rho_right"):
        lat.zou_he_right_wall_pressure()
