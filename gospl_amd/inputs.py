"""goSPL's input file (YAML + npz mesh) for this path.

``read_input(filename)`` parses the subset of goSPL's YAML keys that the
per-step surface-process path uses, with goSPL's defaults
(gospl/tools/inputparser.py):

* ``domain`` (`_readDomain`, :175-341): ``npdata: [file, coords, cells,
  elev]`` -- an npz (loaded without pickles) holding the vertex coordinates,
  the triangles and the initial elevation; ``flowdir`` (default 8); ``bc``
  (default ``oooo``; [North, East, South, West] edges of a planar mesh);
  ``nodep`` (default False: continental sediment routing runs),
  ``seadepo`` (default True);
* ``time`` (`_readTime`, :343-452): ``start``, ``end``, ``dt``, ``tout``,
  ``rstep`` (restart from that output);
* ``output`` (`_readOut`, :2095-2118): ``dir`` ("output"), ``makedir``
  (False), ``compression`` ("gzip");
* ``spl`` (`_readSPL`, :454-496): ``K``, ``m`` (0.5), ``n`` (1), ``d`` (0),
  ``G`` (0);
* ``diffusion`` (`_readHillslope`, :498-599): ``hillslopeKa``,
  ``hillslopeKm``, ``oFill`` (-6000), and the lake-deposit keys
  ``nonlinKm`` (10), ``nlPitVolume`` (1e9), ``nlPitDepth`` (100), ``nlPitK``,
  ``pitInletBias`` (0.1), ``marineSolver``, ``picardSub`` (10),
  ``picardIts`` (2), and the marine keys ``clinSlp`` (1e-6), ``offshore``
  (1e7 m), ``Gmar`` (0);
* ``sea`` (`_readSealevel`, :710-787): ``position`` (0);
* ``climate`` (`_readRain`, :1352-1520): events ``{start, uniform | map:
  [file, key]}``;
* ``tectonics`` (`_readTectonics`, :904-942): events ``{start, end, upsub:
  [file, key]}`` (vertical rate, m/yr);
* ``sedfactor`` (`_readErofactor`, :999-1094): events ``{start, uniform |
  map: [file, key]}``.

Everything else in goSPL's input describes subsystems outside this path and
is ignored.  Settings that would change what this path computes and that it
does not implement raise ``NotImplementedError`` (a
sea-level curve, open or cyclic planar edges); n != 1 selects the non-linear
stream-power solver, as goSPL's runProcesses does; ``nodep: False`` runs
continental sediment routing, and with ``seadepo: True`` the marine
deposition of what reaches the sea.  Relative file names are resolved
against the YAML file's folder, then the working directory, as goSPL opens
them from the working directory.
"""

from __future__ import annotations

import os

import numpy as np
import yaml


def _resolve(path, base):
    for cand in (os.path.join(base, path), path):
        if os.path.exists(cand):
            return cand
    raise FileNotFoundError("goSPL input refers to %s, which is not found (looked in %s and the working "
                            "directory)" % (path, base))


def _npz_field(spec, base, what):
    """A [file, key] pair of goSPL's input: the field ``key`` of file.npz."""
    if not (isinstance(spec, (list, tuple)) and len(spec) == 2):
        raise ValueError("%s must be [file, key] (got %r)" % (what, spec))
    fn = _resolve(str(spec[0]) + ".npz", base)
    with np.load(fn, allow_pickle=False) as d:
        if spec[1] not in d.files:
            raise KeyError("field %s is missing from %s (it holds %s)" % (spec[1], fn, d.files))
        return np.asarray(d[spec[1]], dtype=np.float64)


def _events(lst, what):
    if lst is None:
        return []
    if not isinstance(lst, list):
        raise ValueError("%s must be a list of events" % what)
    for k, ev in enumerate(lst):
        if "start" not in ev:
            raise ValueError("%s event %d has no parameter start" % (what, k))
    return sorted(lst, key=lambda e: e["start"])


def read_input(filename):
    """Parse goSPL's YAML file.  Returns a dict with the mesh arrays
    (``coords``, ``cells``, ``elev``; input order), the flat-model border
    flags and the ``Model`` parameters."""
    base = os.path.dirname(os.path.abspath(filename))
    with open(filename) as fh:
        inp = yaml.safe_load(fh)
    if not isinstance(inp, dict):
        raise ValueError("%s is not a goSPL input file" % filename)
    if "domain" not in inp:
        raise KeyError("Key domain is required in the input file!")
    dom = inp["domain"]
    if "npdata" not in dom:
        raise KeyError("Compressed numpy dataset definition is not defined!")
    npd = dom["npdata"]
    mfile = _resolve(str(npd[0]) + ".npz", base)
    with np.load(mfile, allow_pickle=False) as d:
        coords = np.asarray(d[npd[1]], dtype=np.float64)
        cells = np.asarray(d[npd[2]], dtype=np.int32)
        elev = np.asarray(d[npd[3]], dtype=np.float64)
    bc = str(dom.get("bc", "oooo")).replace("0", "o").replace("1", "f")
    if len(bc) != 4 or any(c not in "ofwc" for c in bc):
        raise ValueError("domain.bc must be 4 characters from {o, f, w, c}; got %r" % dom.get("bc"))
    nodep = bool(dom.get("nodep", False))
    seadepo = bool(dom.get("seadepo", True))
    tdict = inp.get("time")
    if tdict is None:
        raise KeyError("Key time is required in the input file!")
    for k in ("start", "end", "dt"):
        if k not in tdict:
            raise KeyError("Key %s is required in the time declaration" % k)
    spl = inp.get("spl")
    if spl is None:
        K, m, n, d, G = 1.0e-12, 0.5, 1.0, 0.0, 0.0
    else:
        if "K" not in spl:
            raise KeyError("Key K is required in the spl declaration")
        K, m, n, d, G = (float(spl["K"]), float(spl.get("m", 0.5)), float(spl.get("n", 1.0)),
                         float(spl.get("d", 0.0)), float(spl.get("G", 0.0)))
    if not n > 0.0:
        raise ValueError("spl.n = %g: the slope exponent must be positive" % n)
    if G < 0.0:
        raise ValueError("spl.G = %g: the deposition coefficient must be >= 0" % G)
    hill = inp.get("diffusion") or {}
    cda = float(hill.get("hillslopeKa", 0.0))
    cdm = float(hill.get("hillslopeKm", 0.0))
    ofill = float(hill.get("oFill", -6000.0))
    # _extraHillslope (inputparser.py:546-597): the lake-deposit diffusion
    sedkw = dict(nonlinKm=float(hill.get("nonlinKm", 10.0)), nlPitVolume=float(hill.get("nlPitVolume", 1.0e9)),
                 nlPitDepth=float(hill.get("nlPitDepth", 100.0)),
                 nlPitK=float(hill["nlPitK"]) if hill.get("nlPitK") is not None else None,
                 pitInletBias=float(hill.get("pitInletBias", 0.10 if "diffusion" in inp else 0.50)),
                 marineSolver=str(hill.get("marineSolver", "ts")), picardSub=int(hill.get("picardSub", 10)),
                 picardIts=int(hill.get("picardIts", 2)), clinSlp=float(hill.get("clinSlp", 1.0e-6)),
                 offshore=float(hill.get("offshore", 100.0e5)), Gmar=float(hill.get("Gmar", 0.0)))
    sea = inp.get("sea") or {}
    if sea.get("curve") is not None:
        raise NotImplementedError("sea.curve: a time-varying sea level is not on this path (sea.position only)")
    sealevel = float(sea.get("position", 0.0))
    climate = []
    for ev in _events(inp.get("climate"), "climate"):
        if ev.get("map") is not None:
            climate.append((float(ev["start"]), _npz_field(ev["map"], base, "climate map")))
        elif ev.get("uniform") is not None:
            climate.append((float(ev["start"]), float(ev["uniform"])))
        else:
            raise NotImplementedError("climate event at %s: only uniform or map rainfall is on this path"
                                      % ev["start"])
    tecto = []
    for ev in _events(inp.get("tectonics"), "tectonics"):
        if "end" not in ev:
            raise ValueError("Tectonic event at %s has no parameter end" % ev["start"])
        if ev.get("hdisp") is not None or ev.get("zfit") is not None:
            raise NotImplementedError("tectonics hdisp/zfit (horizontal advection) are not on this path")
        up = _npz_field(ev["upsub"], base, "tectonics upsub") if ev.get("upsub") is not None else None
        tecto.append((float(ev["start"]), float(ev["end"]), up))
    sedfac = []
    for ev in _events(inp.get("sedfactor"), "sedfactor"):
        if ev.get("map") is not None:
            sedfac.append((float(ev["start"]), _npz_field(ev["map"], base, "sedfactor map")))
        elif ev.get("uniform") is not None:
            sedfac.append((float(ev["start"]), float(ev["uniform"])))
    out = inp.get("output") or {}
    return dict(name=inp.get("name"), coords=coords, cells=cells, elev=elev, bc=bc, nodep=nodep, seadepo=seadepo,
                sed=sedkw,
                outputDir=str(out.get("dir", "output")), makedir=bool(out.get("makedir", False)),
                compression=out.get("compression", "gzip"), rstep=int(tdict.get("rstep", 0)),
                flowdir=int(dom.get("flowdir", 8)), tstart=float(tdict["start"]), tend=float(tdict["end"]),
                dt=float(tdict["dt"]), tout=float(tdict.get("tout", float(tdict["end"]) - float(tdict["start"]))),
                K=K, m=m, d=d, n=n, G=G, cda=cda, cdm=cdm, ofill=ofill, sealevel=sealevel, climate=climate,
                tectonics=tecto, sedfactor=sedfac)


def border_outlets(coords, cells, bc, bverts):
    """The draining borders of a planar mesh from goSPL's bc flags
    (unstructuredmesh.py:470-533): fixed ('f') edges drain, walls ('w') do
    not; open ('o') edges need goSPL's per-step base-level reset
    (_drainOpenEdges, :796-870) and cyclic ('c') edges a periodic mesh, which
    this path does not provide."""
    bad = sorted({c for c in bc if c in "oc"})
    if bad and len(bverts):
        raise NotImplementedError("domain.bc = %r: open ('o') and cyclic ('c') edges are not on this path; use "
                                  "fixed ('f') or wall ('w') edges" % bc)
    out = np.zeros(len(coords), bool)
    out[bverts] = True
    tol = 1.0e-9
    x, y = coords[:, 0], coords[:, 1]
    edges = {"N": np.isclose(y, y.max(), atol=tol), "E": np.isclose(x, x.max(), atol=tol),
             "S": np.isclose(y, y.min(), atol=tol), "W": np.isclose(x, x.min(), atol=tol)}
    for flag, key in zip(bc, "NESW"):
        if flag == "w":
            out &= ~edges[key]
    return out
