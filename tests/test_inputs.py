"""goSPL's YAML + npz input (gospl_amd.inputs) — CPU only; the GPU test that
runs a model built from a file is in test_gpu_inputs.py."""

from __future__ import annotations

import os
import textwrap
import warnings

import numpy as np
import pytest

from gospl_amd import inputs, mesh as gmesh


def _write_case(tmp_path, extra="", n=12):
    v, c = gmesh.planar_jittered(n, n, 500.0, 100.0, seed=1)
    z = np.random.default_rng(0).uniform(0, 10, len(v))
    np.savez(tmp_path / "mesh.npz", v=v, c=c, z=z)
    np.savez(tmp_path / "forcing.npz", r=np.full(len(v), 2.0), u=np.full(len(v), 1e-3), s=np.full(len(v), 0.05))
    y = textwrap.dedent("""\
        name: test
        domain:
          npdata: ['mesh', 'v', 'c', 'z']
          flowdir: 3
          bc: 'ffff'
          nodep: True
        time:
          start: 0.
          end: 10000.
          dt: 1000.
          tout: 5000.
        spl:
          K: 1.e-5
          d: 0.5
        diffusion:
          hillslopeKa: 0.1
          hillslopeKm: 0.2
        sea:
          position: -10.
        climate:
          - start: 0.
            uniform: 1.
          - start: 3000.
            map: ['forcing', 'r']
        tectonics:
          - start: 0.
            end: 5000.
            upsub: ['forcing', 'u']
        sedfactor:
          - start: 0.
            map: ['forcing', 's']
        """) + extra
    (tmp_path / "input.yml").write_text(y)
    return tmp_path / "input.yml", v, c, z


def test_read_input_keys_and_defaults(tmp_path):
    fn, v, c, z = _write_case(tmp_path)
    inp = inputs.read_input(str(fn))
    assert np.array_equal(inp["coords"], v) and np.array_equal(inp["cells"], c) and np.array_equal(inp["elev"], z)
    assert inp["flowdir"] == 3 and inp["K"] == 1e-5 and inp["m"] == 0.5 and inp["d"] == 0.5
    assert (inp["cda"], inp["cdm"], inp["ofill"], inp["sealevel"]) == (0.1, 0.2, -6000.0, -10.0)
    assert (inp["tstart"], inp["tend"], inp["dt"]) == (0.0, 10000.0, 1000.0)
    assert [t for t, _ in inp["climate"]] == [0.0, 3000.0] and inp["climate"][0][1] == 1.0
    assert np.all(inp["climate"][1][1] == 2.0)
    assert inp["tectonics"][0][:2] == (0.0, 5000.0) and np.all(inp["tectonics"][0][2] == 1e-3)
    assert np.all(inp["sedfactor"][0][1] == 0.05)


def test_sediment_routing_keys(tmp_path):
    """nodep False routes sediment, with goSPL's default seadepo True
    (marine deposition) unless it is set False; the lake-deposit and marine
    diffusion keys are read with goSPL's defaults."""
    fn, *_ = _write_case(tmp_path)
    text = fn.read_text().replace("  nodep: True\n", "")
    fn.write_text(text)
    inp = inputs.read_input(str(fn))
    assert inp["nodep"] is False and inp["seadepo"] is True
    fn.write_text(text.replace("  bc: 'ffff'\n", "  bc: 'ffff'\n  seadepo: False\n"))
    inp = inputs.read_input(str(fn))
    assert inp["nodep"] is False and inp["seadepo"] is False
    assert inp["sed"] == dict(nonlinKm=10.0, nlPitVolume=1.0e9, nlPitDepth=100.0, nlPitK=None, pitInletBias=0.10,
                              marineSolver="ts", picardSub=10, picardIts=2, clinSlp=1.0e-6, offshore=100.0e5,
                              Gmar=0.0)
    fn.write_text(text.replace("  hillslopeKm: 0.2\n", "  hillslopeKm: 0.2\n  clinSlp: 1.e-5\n  Gmar: 0.5\n"
                                                          "  offshore: 5.e5\n"))
    inp = inputs.read_input(str(fn))
    assert (inp["sed"]["clinSlp"], inp["sed"]["Gmar"], inp["sed"]["offshore"]) == (1.0e-5, 0.5, 5.0e5)
    fn.write_text(text.replace("  bc: 'ffff'\n", "  bc: 'ffff'\n  seadepo: False\n").replace(
        "  hillslopeKm: 0.2\n", "  hillslopeKm: 0.2\n  nlPitK: 3.\n  marineSolver: picard\n  picardSub: 4\n"))
    inp = inputs.read_input(str(fn))
    assert inp["sed"]["nlPitK"] == 3.0 and inp["sed"]["marineSolver"] == "picard" and inp["sed"]["picardSub"] == 4


@pytest.mark.parametrize("old,new,exc", [
    ("  d: 0.5\n", "  d: 0.5\n  n: -1.\n", ValueError),
    ("  d: 0.5\n", "  d: 0.5\n  G: -1.\n", ValueError),
    ("  position: -10.\n", "  position: -10.\n  curve: 'sea.csv'\n", NotImplementedError),
    ("  K: 1.e-5\n", "  m: 0.4\n", KeyError),
    ("  dt: 1000.\n", "", KeyError),
    ("['mesh', 'v', 'c', 'z']", "['nomesh', 'v', 'c', 'z']", FileNotFoundError),
])
def test_unsupported_or_missing_settings_raise(tmp_path, old, new, exc):
    fn, *_ = _write_case(tmp_path)
    text = fn.read_text()
    assert old in text
    fn.write_text(text.replace(old, new, 1))
    with pytest.raises(exc):
        inputs.read_input(str(fn))


def test_slope_exponent_is_read(tmp_path):
    fn, *_ = _write_case(tmp_path)
    fn.write_text(fn.read_text().replace("  d: 0.5\n", "  d: 0.5\n  n: 1.5\n", 1))
    assert inputs.read_input(str(fn))["n"] == 1.5


def test_deposition_coefficient_is_read(tmp_path):
    fn, *_ = _write_case(tmp_path)
    assert inputs.read_input(str(fn))["G"] == 0.0
    fn.write_text(fn.read_text().replace("  d: 0.5\n", "  d: 0.5\n  G: 1.5\n", 1))
    assert inputs.read_input(str(fn))["G"] == 1.5


def test_border_outlets_follow_bc(tmp_path):
    v, c = gmesh.planar_jittered(10, 10, 100.0, 0.0, seed=0)
    bv = gmesh.boundary_vertices(c)
    out = inputs.border_outlets(v, c, "ffff", bv)
    assert np.array_equal(np.nonzero(out)[0], np.sort(bv))
    out = inputs.border_outlets(v, c, "fwff", bv)  # east wall
    east = np.isclose(v[:, 0], v[:, 0].max())
    assert not out[east].any() and out[bv[~east[bv]]].all()
    with pytest.raises(NotImplementedError):
        inputs.border_outlets(v, c, "offf", bv)
