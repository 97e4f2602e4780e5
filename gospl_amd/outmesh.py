"""goSPL's model outputs and restart (gospl/tools/outmesh.py).

``WriteMesh`` is mixed into ``Model`` and keeps goSPL's method names:

* ``_createOutputDir`` (outmesh.py:106-142): ``<cwd>/<dir>`` with ``h5/``
  and ``xmf/``; an existing folder is removed, or with ``makedir`` the run
  goes to the first free ``<dir>_<n>``; the input file is copied in;
* ``visModel`` (:89-104): an output at the start time, then whenever
  ``tNow >= saveTime``, ``saveTime`` advancing by ``tout``;
* ``_outputMesh`` (:256-507): ``h5/topology.p0.h5`` (float32 ``coords``,
  1-based int32 ``cells``) at the first output, and per output
  ``h5/gospl.<step>.p0.h5`` with the per-vertex float32 ``(N, 1)`` datasets
  ``elev``, ``erodep``, ``EDrate``, ``waterFill``, ``fillFA``, ``FA``
  (discharges floored at 1e-8 and set to 1 on the sea), ``sedLoad``,
  ``uplift`` and ``rain`` when those forcings exist;
* ``_save_DMPlex_XMF`` / ``_save_XDMF`` (:627-909): the per-output XMF and
  the ``gospl.xdmf`` time series ParaView opens;
* ``readData`` (:509-625): a restart (``time.rstep`` > 0) reads ``elev``,
  ``erodep``, ``sedLoad``, ``EDrate``, ``FA`` and ``fillFA`` of output
  ``rstep`` back (as float32 values, as goSPL does).

Arrays are written in input-mesh order (goSPL's one-rank local order).  The
files are HDF5 through the HDF5 C library (``gospl_amd.h5``); one rank, so
every file carries the suffix ``.p0``.
"""

from __future__ import annotations

import os
import shutil

import numpy as np
import torch

from . import h5

DISCHARGE_FLOOR = 1.0e-8   # gospl/tools/constants.py:51


class WriteMesh:
    """Output and restart methods of goSPL's ``WriteMesh`` for ``Model``."""

    def _init_output(self, outputDir=None, makedir=False, tout=None, rstep=0, compression="gzip", finput=None):
        """WriteMesh.__init__ (outmesh.py:31-87) and goSPL's time keys
        (inputparser.py:343-452).  outputDir None: no outputs (a model built
        from keywords); Model(filename) passes goSPL's default "output"."""
        self.file = "gospl"
        # goSPL's output counter (its attribute `step`; Model.step() here is
        # the body of the time loop)
        self.outputStep = 0
        self.rStep = int(rstep or 0)
        self.outputDir = outputDir
        self.makedir = bool(makedir) and self.rStep == 0
        self.outCompress = compression
        self.finput = finput
        self.elems = self.nodes = None
        tout = float(tout) if tout is not None else self.tEnd - self.tStart
        if self.tStart + tout > self.tEnd:
            tout = self.tEnd - self.tStart
        if tout < self.dt:
            tout = self.dt
        self.tout = tout
        self.saveTime = self.tNow
        self._seaOut = None
        if self.rStep > 0:
            self.outputStep = self.rStep + 1
            self.tNow = self.tStart + self.rStep * self.tout
            self.saveTime = self.tNow + self.tout
        if self.outputDir is not None:
            self._createOutputDir()

    # ------------------------------------------------------------ folders
    def _createOutputDir(self):
        d = self.outputDir
        if not os.path.isabs(d):
            d = os.path.join(os.getcwd(), d)
        if self.rStep == 0:
            if self.makedir and os.path.exists(d):
                n = 0
                while os.path.exists("%s_%d" % (d, n)):
                    n += 1
                d = "%s_%d" % (d, n)
            elif os.path.exists(d):
                shutil.rmtree(d, ignore_errors=True)
            os.makedirs(os.path.join(d, "h5"))
            os.makedirs(os.path.join(d, "xmf"))
        elif not os.path.isdir(os.path.join(d, "h5")):
            raise ValueError("Restart file is missing... (no output folder %s)" % d)
        if self.finput is not None:
            shutil.copy(self.finput, d)
        self.outputDir = d

    # ------------------------------------------------------------ writing
    def _h5opts(self):
        c = self.outCompress
        if c in (None, False, "none", "None", ""):
            return dict(compression=None)
        if isinstance(c, bool):
            return dict(compression="gzip")
        if isinstance(c, int):
            return dict(compression="gzip", level=c)
        return dict(compression=c)

    def _field(self, t):
        """Per-vertex tensor (device order) -> float32 (N, 1) in input order."""
        a = t.detach().to("cpu", torch.float64).numpy() if isinstance(t, torch.Tensor) else np.asarray(t, np.float64)
        a = np.broadcast_to(a, (self.lpoints,))
        return self.dm.mesh.to_input_order(a).astype(np.float32).reshape(-1, 1)

    def _sea_mask(self):
        """goSPL's seaID when the output is written: set by the last
        getHillslope (h <= sea level before diffusion, hillslope.py:620) or,
        before the first step, by the flow directions (lFill <= sea level,
        flowplex.py:493)."""
        if self._seaOut is not None:
            return self._seaOut
        return self.lFill <= self.sealevel

    def _discharge(self, t, sea):
        d = t.clone()
        d[d <= DISCHARGE_FLOOR] = DISCHARGE_FLOOR
        d[sea] = 1.0
        return d

    def visModel(self):
        if self.outputDir is None:
            return
        if self.saveTime == self.tStart or self.tNow >= self.saveTime:
            self._outputMesh()
            self.saveTime += self.tout

    def _outputMesh(self):
        opts = self._h5opts()
        mesh = self.dm.mesh
        cells = mesh.extras.get("cells")
        if cells is None:
            raise RuntimeError("outputs need the mesh triangles (build the mesh with gospl_amd.mesh.build_mesh)")
        if self.outputStep == 0:
            with h5.File(os.path.join(self.outputDir, "h5", "topology.p0.h5"), "w") as f:
                f.write("coords", mesh.to_input_order(mesh.coords).astype(np.float32), **opts)
                f.write("cells", (np.asarray(cells) + 1).astype(np.int32), **opts)
        self.elems = [len(cells)]
        self.nodes = [self.lpoints]
        sea = self._sea_mask()
        name = os.path.join(self.outputDir, "h5", "%s.%d.p0.h5" % (self.file, self.outputStep))
        with h5.File(name, "w") as f:
            f.write("elev", self._field(self.hLocal), **opts)
            f.write("erodep", self._field(self.cumED), **opts)
            f.write("EDrate", self._field(self.EbLocal), **opts)
            f.write("waterFill", self._field(self.waterFilled), **opts)
            f.write("fillFA", self._field(self._discharge(self.fillFAL, sea)), **opts)
            f.write("FA", self._field(self._discharge(self.FAL, sea)), **opts)
            vs = self.vSedLocal.clone()
            vs[vs <= DISCHARGE_FLOOR] = DISCHARGE_FLOOR
            f.write("sedLoad", self._field(vs), **opts)
            if self.upsub is not None:
                f.write("uplift", self._field(self.upsub), **opts)
            if self.rainVal is not None:
                f.write("rain", self._field(self.rainVal), **opts)
        self._save_DMPlex_XMF()
        self._save_XDMF()
        if self.verbose:
            print("+++ Output Simulation Time: %0.02f years" % self.tNow, flush=True)
        self.outputStep += 1

    def _save_DMPlex_XMF(self):
        n = self.nodes[0]
        pfile = "h5/%s.%d.p0.h5" % (self.file, self.outputStep)
        tfile = "h5/topology.p0.h5"

        def attr(label, ds):
            return ('         <Attribute Type="Scalar" Center="Node" Name="%s">\n'
                    '          <DataItem Format="HDF" NumberType="Float" Precision="4" '
                    'Dimensions="%d 1">%s:/%s</DataItem>\n'
                    '         </Attribute>\n' % (label, n, pfile, ds))

        out = ['<?xml version="1.0" encoding="UTF-8"?>\n', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd">\n',
               '<Xdmf Version="2.0" xmlns:xi="http://www.w3.org/2001/XInclude">\n', ' <Domain>\n',
               '    <Grid GridType="Collection" CollectionType="Spatial">\n',
               '      <Time Type="Single" Value="%0.02f"/>\n' % self.saveTime,
               '      <Grid Name="Block.0">\n',
               '         <Topology Type="Triangle" NumberOfElements="%d" BaseOffset="1">\n' % self.elems[0],
               '          <DataItem Format="HDF" DataType="Int" Dimensions="%d 3">%s:/cells</DataItem>\n'
               % (self.elems[0], tfile),
               '         </Topology>\n', '         <Geometry Type="XYZ">\n',
               '          <DataItem Format="HDF" NumberType="Float" Precision="4" Dimensions="%d 3">%s:/coords'
               '</DataItem>\n' % (n, tfile),
               '         </Geometry>\n']
        out += [attr("Z", "elev"), attr("ED", "erodep"), attr("waterFill", "waterFill"), attr("EDrate", "EDrate"),
                attr("FA", "FA"), attr("fillFA", "fillFA"), attr("SL", "sedLoad")]
        if self.upsub is not None:
            out.append(attr("vTec", "uplift"))
        if self.rainVal is not None:
            out.append(attr("Rain", "rain"))
        out += ['         <Attribute Type="Scalar" Center="Node" Name="sea">\n',
                '          <DataItem ItemType="Function" Function="$0 * 0.00000000000000001 + %f" '
                'Dimensions="%d 1">\n' % (self.sealevel, n),
                '           <DataItem Format="HDF" NumberType="Float" Precision="4" Dimensions="%d 1">%s:/erodep'
                '</DataItem>\n' % (n, pfile),
                '          </DataItem>\n', '         </Attribute>\n', '      </Grid>\n', '    </Grid>\n',
                ' </Domain>\n', '</Xdmf>\n']
        with open(os.path.join(self.outputDir, "xmf", "%s%d.xmf" % (self.file, self.outputStep)), "w") as f:
            f.write("".join(out))

    def _save_XDMF(self):
        out = ['<?xml version="1.0" encoding="UTF-8"?>\n', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd">\n',
               '<Xdmf Version="2.0" xmlns:xi="http://www.w3.org/2001/XInclude">\n', ' <Domain>\n',
               '    <Grid GridType="Collection" CollectionType="Temporal">\n']
        for s in range(self.outputStep + 1):
            out.append('      <xi:include href="xmf/%s%d.xmf" xpointer="xpointer(//Xdmf/Domain/Grid)"/>\n'
                       % (self.file, s))
        out += ['    </Grid>\n', ' </Domain>\n', '</Xdmf>\n']
        with open(os.path.join(self.outputDir, self.file + ".xdmf"), "w") as f:
            f.write("".join(out))

    # ------------------------------------------------------------ restart
    def readData(self):
        name = os.path.join(self.outputDir, "h5", "%s.%d.p0.h5" % (self.file, self.outputStep - 1))
        if not os.path.exists(name):
            raise ValueError("Restart file is missing... (%s)" % name)
        mesh = self.dm.mesh
        dev = self.device

        def get(f, ds):
            a = f.read(ds, np.float32)[:, 0]
            if len(a) != self.lpoints:
                raise ValueError("%s:/%s holds %d values for a %d-vertex mesh" % (name, ds, len(a), self.lpoints))
            return torch.as_tensor(mesh.to_device_order(a.astype(np.float64))).to(dev)

        with h5.File(name, "r") as f:
            self.hLocal = get(f, "elev").contiguous()
            self.cumED = get(f, "erodep")
            self.vSedLocal = get(f, "sedLoad")
            self.EbLocal = get(f, "EDrate")
            self.FAL.copy_(get(f, "FA"))
            self.fillFAL.copy_(get(f, "fillFA"))
        cells = mesh.extras.get("cells")
        self.elems = [len(cells)] if cells is not None else None
        self.nodes = [self.lpoints]
