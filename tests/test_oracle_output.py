"""Pins of the oracle's output accumulation (src/mod_output.f90:94-160):
time means of constant fields are the fields, the A-grid u (v) averages
D-grid u (v) along j (i), and write_output resets the accumulators."""
import numpy as np

import oracle as O

H = O.HALO


def _state(o, dp, pt, ud, vd, ts, nf):
    for fid, a in ((O.DP, dp), (O.PT, pt), (O.UD, ud), (O.VD, vd), (O.NET_FLUX, nf)):
        o.field(fid)[...] = a
    o.field(O.TS)[...] = ts


def test_constant_fields_mean_to_themselves():
    nx, ny, L = 6, 5, 3
    o = O.Oracle(nx, ny, L, 1e5, 1e5)
    shp = o.field(O.DP).shape
    _state(o, np.full(shp, 3000.0), np.full(shp, 42.5), np.full(shp, -7.25), np.full(shp, 3.5),
           np.full(o.field(O.TS).shape, 288.0), np.full(o.field(O.NET_FLUX).shape, -12.0))
    o.init_output()
    for dt in (60.0, 30.0, 17.0):
        o.accumulate_output(dt)
    m = o.output_mean()
    for f, v in ((0, 3000.0), (1, 42.5), (2, -7.25), (3, 3.5)):
        np.testing.assert_allclose(m[f, :L], v, rtol=1e-15)
        assert np.all(m[f, L] == 0.0)  # no layer L+1
    np.testing.assert_allclose(m[4, 0], 288.0, rtol=1e-15)
    assert np.all(m[4, 1:] == 0.0)
    np.testing.assert_allclose(m[5], -12.0, rtol=1e-15)


def test_a_grid_winds_and_reset():
    nx, ny, L = 7, 6, 2
    o = O.Oracle(nx, ny, L, 1e5, 1e5)
    shp = o.field(O.UD).shape
    jj = np.arange(shp[1], dtype=float)[None, :, None]
    ii = np.arange(shp[2], dtype=float)[None, None, :]
    ud = np.broadcast_to(2.0 + 0.5 * jj, shp).copy()  # varies along j only
    vd = np.broadcast_to(-1.0 + 0.25 * ii, shp).copy()  # varies along i only
    zero = np.zeros(shp)
    _state(o, zero, zero, ud, vd, np.zeros(o.field(O.TS).shape), np.zeros(o.field(O.NET_FLUX).shape))
    o.init_output()
    o.accumulate_output(10.0)
    m = o.output_mean()
    j = np.arange(ny)[:, None] + H
    i = np.arange(nx)[None, :] + H
    np.testing.assert_allclose(m[2, 0], np.broadcast_to(2.0 + 0.5 * (j + 0.5), (ny, nx)), rtol=1e-15)
    np.testing.assert_allclose(m[3, 0], np.broadcast_to(-1.0 + 0.25 * (i + 0.5), (ny, nx)), rtol=1e-15)
    # after write_output the sums restart
    o.field(O.UD)[...] = 1.0
    o.accumulate_output(5.0)
    np.testing.assert_allclose(o.output_mean()[2, :L], 1.0, rtol=1e-15)
