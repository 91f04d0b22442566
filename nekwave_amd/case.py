"""Case lifecycle over the solver (reference Case, src/case.hpp:48-153, case.cpp:94-728).

preprocess(): box mesh or mesh file, dt rule, initial condition hook evaluated at the nodes.
simulate():   DgSolver.initialize + uploadState, then the step loop with the
              reference's termination and last-dt clamp (case.cpp:562-583).
              Unlike the reference (which downloads the whole state every step,
              case.cpp:588) probes are recorded and diagnostics reduced on the
              device, the steps between two outputs run as one multi-step
              launch, and the state is downloaded once at the end.
postprocess(): user hook, then DgSolver.finalize.

Config keeps the subset of the reference's keys that shape the step
(config.hpp:21-254): order, cfl, dt, c0/flux_type, nelx/nely/nelz, bounds,
periodic_*, mesh file, num_steps/max_steps, final_time, output_frequency, probes,
relative_probes, num_probes, output_dir, export_fields, export_format,
save_frequency, export_continuous_vtk; plus the extensions precision, pipeline,
jitter, seed.

Field snapshots (case.cpp:438-464, 603-625): with export_fields on and the
format "hdf5" or "vtk", <output_dir>/fields.bin, fields_step_<n>.vtu and
fields.pvd as Hdf5Writer writes them without the HDF5 library, at t = 0, every
save_frequency steps and at the last step; the state is downloaded for those
steps only.

Probes (probe.cpp:33-216, case.cpp:381-424, 473-474, 594-596) are recorded on the
device every step; the small history comes back at output frequency and is
written to <output_dir>/probe_history.csv in the reference's format, beside
probes.json.
"""
import json
import math
import os
import sys

import numpy as np

from . import DgSolver, FieldWriter, Mesh


def _six_sig(v):
    """The reference round-trips JSON numbers through a default-precision string
    (json.hpp:52-61), keeping 6 significant digits."""
    return float("%g" % float(v))


def _round_half_away(v):
    """std::round: halves away from zero."""
    return math.floor(v + 0.5) if v >= 0 else -math.floor(-v + 0.5)


def probes_json_text(order, num_modes, points, domain_lx, elements_x, mode_k=()):
    """probes.json as ProbeManager::init writes it (probe.cpp:135-173). The probe
    is placed by interpolation, so nodeIdx is -1, actual = target and offset 0.
    The stream turns std::scientific with precision 8 at the separation line,
    which exists only for two or more probes: with one probe the numbers print
    in the default (%g) format."""
    pts = [tuple(float(v) for v in p) for p in points]
    sci = len(pts) >= 2
    num = (lambda v: "%.8e" % v) if sci else (lambda v: "%g" % v)
    out = ["{", '  "order": %d,' % order, '  "num_modes": %d,' % num_modes, '  "num_probes": %d,' % len(pts)]
    if sci:
        dx, dy, dz = pts[1][0] - pts[0][0], pts[1][1] - pts[0][1], pts[1][2] - pts[0][2]
        dist = math.sqrt(dx * dx + dy * dy + dz * dz)
        out += ['  "separation_distance": %.8e,' % dist, '  "delta_x": %.8e,' % dx, '  "delta_y": %.8e,' % dy,
                '  "delta_z": %.8e,' % dz, '  "domain_Lx": %.8e,' % domain_lx, '  "elements_x": %d,' % elements_x]
        if len(mode_k):
            out.append('  "k_modes": [' + ", ".join("%.8e" % k for k in mode_k) + "],")
    out.append('  "probes": [')
    for q, p in enumerate(pts):
        xyz = ", ".join(num(v) for v in p)
        out += ["    {", '      "id": %d,' % (q + 1), '      "nodeIdx": -1,', '      "target": [%s],' % xyz,
                '      "actual": [%s],' % xyz, '      "offset": %s' % num(0.0), "    }" + ("," if q + 1 < len(pts) else "")]
    out += ["  ]", "}"]
    return "\n".join(out) + "\n"


class Config:
    _ALIASES = {"n": "order", "polynomial_order": "order", "num_steps": "num_steps", "nsteps": "num_steps",
                "max_steps": "num_steps", "steps": "num_steps", "t_final": "final_time", "final_time": "final_time",
                "output_frequency": "output_frequency", "outputfreq": "output_frequency",
                "flux": "flux_type", "mesh": "mesh_file", "file": "mesh_file", "path": "mesh_file",
                "meshfile": "mesh_file", "probe": "probes", "probe_rel": "relative_probes",
                "outputdir": "output_dir", "sigma": "pulse_sigma", "save_freq": "save_frequency",
                "export_freq": "save_frequency", "save_format": "export_format", "save_fields": "export_fields",
                "continuous_vtk": "export_continuous_vtk", "export_continuous": "export_continuous_vtk"}

    def __init__(self, **kw):
        self.order = 3
        self.cfl = -1.0
        self.dt = -1.0
        self.num_steps = 100
        # whether num_steps / max_steps / steps was given (Config::hasExplicitMaxSteps,
        # config.hpp:29, 182-185): without it a positive final_time alone ends the run
        self.explicit_steps = False
        self.final_time = -1.0
        self.output_frequency = 10
        self.c0 = 1.0
        self.nelx = self.nely = self.nelz = 0
        self.xmin, self.xmax, self.ymin, self.ymax, self.zmin, self.zmax = -1.0, 1.0, -1.0, 1.0, -1.0, 1.0
        self.periodic_x = self.periodic_y = self.periodic_z = False
        self.mesh_file = ""
        self.precision = "fp64"
        self.pipeline = "fused"
        self.jitter = 0.0
        self.seed = 0
        self.probes = []
        self.relative_probes = []
        self.num_probes = 0
        self.num_modes = 1
        # built-in initial conditions (case.cpp:247-372; the wave type is set on the Case)
        self.pulse_sigma = 20.0
        self.carrier_k = None
        self.angle_deg = 0.0
        self.k_start = 0.0
        self.k_end = 0.0
        self.packet_sigma = 0.30
        self.packet_x0 = None
        self.output_dir = "output"
        # field snapshots (config.hpp:32-48, 193-226): written through the
        # VTU/PVD/.bin files of Hdf5Writer when export_fields is on and the format
        # is "hdf5" or "vtk"; save_frequency <= 0 means output_frequency
        self.export_fields = False
        self.export_format = "hdf5"
        self.save_frequency = -1
        self.export_continuous_vtk = False
        # the registry entry "wave_type" (Config::set, config.hpp:127-129); Case uses it
        # when setWaveType was not called (case.cpp:163)
        self.wave_type = ""
        for k, v in kw.items():
            self.set(k, v)

    def set(self, key, val):
        k = self._ALIASES.get(key.lower(), key.lower())
        if k == "flux_type":
            self.c0 = 0.0 if str(val).lower() in ("central", "0", "0.0") else 1.0
            return
        if k in ("cfl", "dt") and isinstance(val, str) and val.lower() == "auto":
            setattr(self, k, -1.0)
            return
        if k in ("probes", "relative_probes"):  # config.hpp:291-322: [x, y, z] items or "x,y,z" strings
            items = [val] if isinstance(val, str) else list(val)
            if items and not isinstance(items[0], (list, tuple, str)):
                items = [items]
            for it in items:
                v3 = [float(t) for t in it.replace(",", " ").split()] if isinstance(it, str) else [float(t) for t in it]
                if len(v3) >= 3:
                    getattr(self, k).append(tuple(v3[:3]))
            return
        if not hasattr(self, k) or k == "explicit_steps":
            raise KeyError("unknown config key %r" % key)
        if k == "num_steps":
            self.explicit_steps = True
        cur = getattr(self, k)
        if cur is None:
            val = float(val)
        elif isinstance(cur, bool):
            val = str(val).lower() in ("1", "true", "yes", "on") if not isinstance(val, bool) else val
        elif isinstance(cur, int) and not isinstance(cur, bool):
            val = int(val)
        elif isinstance(cur, float):
            val = float(val)
        setattr(self, k, val)

    @classmethod
    def from_json(cls, path_or_dict):
        """Flatten JSON sections like Config::loadFromJson (config.hpp:261-340)."""
        d = path_or_dict if isinstance(path_or_dict, dict) else json.load(open(path_or_dict))
        c = cls()
        flat = {}
        for k, v in d.items():
            if isinstance(v, dict):
                flat.update(v)
            elif not k.startswith("_"):
                flat[k] = v
        for k, v in flat.items():
            if k.lower() == "wave_type":  # config.hpp:146-149: a file cannot choose the wave type
                print("[CONFIG] Notice: 'wave_type' in configuration file is ignored (wave type must be "
                      "specified directly in code).")
                continue
            if isinstance(v, float):
                v = _six_sig(v)
            try:
                c.set(k, v)
            except KeyError:
                pass
        return c

    def max_steps(self):
        """Step limit as Case::preprocess settles it (case.cpp:429-436 after
        config.hpp:446-450): an explicit or non-default step count wins; a positive
        final_time with no step count given runs to final_time (-1: no limit)."""
        if self.num_steps != 100:
            return self.num_steps
        if self.explicit_steps or self.final_time <= 0.0:
            return 100
        return -1


def step_schedule(dt, max_steps, final_time):
    """The reference's time loop (case.cpp:562-583): yields (step, dt, time after
    the step, whether it is the last step).  It stops at max_steps (when > 0) or
    once time reaches final_time (when > 0), the last dt clamped to land on it."""
    step, t = 0, 0.0
    while True:
        if max_steps > 0 and step >= max_steps:
            return
        if final_time > 0.0 and t >= final_time - 1e-13:
            return
        if max_steps <= 0 and final_time <= 0.0:
            raise ValueError("no termination: neither a step limit nor a final time")
        d = dt
        if final_time > 0.0 and t + d > final_time:
            d = final_time - t
        step += 1
        t += d
        final = (max_steps > 0 and step >= max_steps) or (final_time > 0.0 and t >= final_time - 1e-13)
        yield step, d, t, final


class Case:
    PROBE_CAPACITY = 4096  # device probe records kept between fetches

    def __init__(self, config=None):
        self.config = config or Config()
        self.ic = None
        self.post = None
        self.mesh = None
        self.solver = None
        self.state = None
        self.dt = 0.0
        self.cfl = 0.0
        self.current_time = 0.0
        self.current_step = 0
        self.history = []
        self.custom_probes = []
        self.custom_relative_probes = []
        self.wave_type = ""
        self.num_modes = 1  # as used (the bloch wave clamps it), written to probes.json
        self.mode_k = []
        self.probe_points = np.zeros((0, 3))
        self.probe_steps = []
        self.probe_times = []
        self.probe_values = np.zeros((0, 0, 6))
        self.save_fields = ["E", "H"]
        self.field_writer = None

    def setFieldSaveOptions(self, fields):
        """Point fields of each snapshot (Case::setFieldSaveOptions with
        Hdf5Writer::FieldSaveOptions): any of "E", "H", "magnitude_E",
        "magnitude_H", "energy_density"."""
        self.save_fields = list(fields)

    def setWaveType(self, wave_type):
        """Built-in initial condition when no hook is set (case.cpp:59-62, 164-176,
        247-372): "gaussian", "bloch"/"periodic", "multimode", "packet"; "3dboxper"
        maps a [-1, 1]^3 mesh onto [0, 2 pi]^3 (Mesh::rescale, mesh.cpp:639-677)."""
        self.wave_type = str(wave_type)

    def waveType(self):
        """The wave type in force: setWaveType's, else the config's (case.cpp:163)."""
        return self.wave_type or self.config.wave_type

    def addProbe(self, x, y, z):
        self.custom_probes.append((float(x), float(y), float(z)))

    def addRelativeProbe(self, rx, ry, rz):
        self.custom_relative_probes.append((float(rx), float(ry), float(rz)))

    def _probe_list(self):
        """Probe coordinates in the order of case.cpp:383-422."""
        c = self.config
        pts = list(c.probes) + list(self.custom_probes)
        lx, ly, lz = c.xmax - c.xmin, c.ymax - c.ymin, c.zmax - c.zmin
        xc, yc, zc = 0.5 * (c.xmin + c.xmax), 0.5 * (c.ymin + c.ymax), 0.5 * (c.zmin + c.zmax)
        for r in list(c.relative_probes) + list(self.custom_relative_probes):
            pts.append((xc + r[0] * lx, yc + r[1] * ly, zc + r[2] * lz))
        if not pts and c.num_probes > 0:
            if c.nelx == 1 and c.nely == 1 and c.nelz == 1:
                z = self.mesh.getGllZ()
                for q in range(min(c.num_probes, self.mesh.getN())):
                    pts.append((c.xmin + 0.5 * (1.0 + z[q]) * lx, yc, zc))
            else:
                x0, x1 = c.xmin + 0.05 * lx, c.xmax - 0.05 * lx
                dx = (x1 - x0) / (c.num_probes - 1) if c.num_probes > 1 else 0.0
                for q in range(c.num_probes):
                    pts.append((x0 + q * dx, yc, zc))
        return np.array(pts, dtype=np.float64).reshape(-1, 3)

    def _fetch_probes(self, pending, out):
        if not pending:
            return
        h = self.solver.probeHistory()
        if h.shape[0] != len(pending):
            raise RuntimeError("probe history holds %d records, expected %d" % (h.shape[0], len(pending)))
        self._probe_chunks.append(h)
        for (st, t), rec in zip(pending, h):
            self.probe_steps.append(st)
            self.probe_times.append(t)
            if out is not None:
                out.write("%d,%.8e" % (st, t))  # std::scientific, setprecision(8) (probe.cpp:133)
                for v in rec.reshape(-1):
                    out.write(",%.8e" % v)
                out.write("\n")
        pending.clear()

    def setInitialConditionHook(self, fn):
        """fn(x, y, z) -> (Ex, Ey, Ez, Hx, Hy, Hz), vectorised over node arrays."""
        self.ic = fn

    def setPostprocessingHook(self, fn):
        self.post = fn

    def preprocess(self):
        c = self.config
        self.mesh = Mesh(c.order)
        if c.nelx > 0 and c.nely > 0 and c.nelz > 0:
            self.mesh.createBoxMesh(c.nelx, c.nely, c.nelz, c.xmin, c.xmax, c.ymin, c.ymax, c.zmin, c.zmax,
                                    c.periodic_x, c.periodic_y, c.periodic_z, jitter=c.jitter, seed=c.seed)
        elif c.mesh_file and c.mesh_file != "box":  # case.cpp:144-156
            if not self.mesh.loadFromRea(c.mesh_file):
                print("Warning: Could not load mesh. Falling back to default 3x3x3 mesh.", file=sys.stderr)
                self.mesh.createBoxMesh(3, 3, 3, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
        else:  # the reference's default 3x3x3 box (case.cpp:158-161)
            self.mesh.createBoxMesh(3, 3, 3, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
        if self.waveType() == "3dboxper":  # case.cpp:165-176
            corners = self.mesh.getElementCorners()
            if abs(corners[..., 0].min() + 1.0) < 1e-3 and abs(corners[..., 0].max() - 1.0) < 1e-3:
                lo, hi = corners.reshape(-1, 3).min(0), corners.reshape(-1, 3).max(0)
                span = np.where(hi - lo < 1e-14, 1.0, hi - lo)
                scaled = 0.0 + (corners - lo) / span * (2.0 * math.pi - 0.0)
                kind, ne, nf = self.mesh.getFaces()
                self.mesh = Mesh.fromCorners(c.order, scaled, kind, ne, nf)
        dxmin = self.mesh.computeMinNodeDistance()
        if c.dt > 0.0:  # case.cpp:201-210
            self.dt = c.dt
            self.cfl = self.dt / dxmin
        elif c.cfl > 0.0:
            self.cfl = c.cfl
            self.dt = self.cfl * dxmin
        else:
            self.cfl = Mesh.computeAutomaticCFL(0.4)
            self.dt = self.mesh.computeAutomaticDt(1.0, 0.4)
        x, y, z = self.mesh.coords()
        self.num_modes, self.mode_k = c.num_modes, []
        if self.ic is not None:
            f = self.ic(x, y, z)
            self.state = np.concatenate([np.broadcast_to(np.asarray(v, dtype=np.float64), x.shape) for v in f])
        else:
            self.state = np.zeros(6 * x.size)
            self._builtin_ic(x, y, z)
        self.current_time = 0.0
        self.current_step = 0

    def _builtin_ic(self, x, y, z):
        """The reference's built-in fields (case.cpp:247-372), evaluated at the nodes."""
        c, n = self.config, x.size
        st = self.state
        lx, ly = c.xmax - c.xmin, c.ymax - c.ymin
        xt, yt = x - c.xmin, y - c.ymin
        w = self.waveType()
        if w in ("bloch", "periodic"):
            ang = c.angle_deg
            rad = ang * math.pi / 180.0
            if abs(ang - 0.0) < 1e-4:
                kx, ky = 2.0 * math.pi / lx, 0.0
            elif abs(ang - 45.0) < 1e-4:
                kx, ky = 2.0 * math.pi / lx, 2.0 * math.pi / ly
            elif abs(ang - 90.0) < 1e-4:
                kx, ky = 0.0, 2.0 * math.pi / ly
            elif abs(ang - 26.6) < 1.0 or abs(ang - 30.0) < 5.0:
                kx, ky = 4.0 * math.pi / lx, 2.0 * math.pi / ly
            elif abs(ang - 18.4) < 1.0 or abs(ang - 15.0) < 5.0:
                kx, ky = 6.0 * math.pi / lx, 2.0 * math.pi / ly
            else:
                mx = max(1, int(_round_half_away(math.cos(rad) * 4.0)))
                my = int(_round_half_away(math.sin(rad) * 4.0))
                kx, ky = 2.0 * math.pi * mx / lx, 2.0 * math.pi * my / ly
            kmag = math.sqrt(kx * kx + ky * ky)
            ca, sa = kx / kmag, ky / kmag
            max_modes = max(1, (c.nelx * (self.mesh.getN() - 1)) // 2)
            nm = min(c.num_modes, max_modes)
            if c.k_start > 0.0 and c.k_end > 0.0:
                ks = [c.k_start + m * (c.k_end - c.k_start) / (nm - 1) if nm > 1 else c.k_start for m in range(nm)]
            else:
                ks = [m * kmag for m in range(1, nm + 1)]
            if nm > 1:
                e = np.zeros(n)
                for k in ks:
                    e += np.cos(k * (ca * xt + sa * yt))
                e /= math.sqrt(nm)
            else:
                carrier = c.carrier_k if c.carrier_k is not None else 2.0 * math.pi / lx
                e = np.cos(carrier * xt if (c.carrier_k is not None and abs(ang) < 1e-4) else kx * xt + ky * yt)
            st[2 * n:3 * n], st[3 * n:4 * n], st[4 * n:5 * n] = e, sa * e, -ca * e
            self.num_modes, self.mode_k = nm, ks
        elif w == "multimode":
            e = np.zeros(n)
            for m in range(1, 9):
                e += (1.0 / m) * np.sin(m * math.pi / lx * xt)
            st[2 * n:3 * n] = e * np.sin(math.pi * yt / ly)
        elif w == "packet":
            carrier = c.carrier_k if c.carrier_k is not None else 4.0 * math.pi
            x0 = c.packet_x0 if c.packet_x0 is not None else c.xmin + 0.25 * lx
            env = np.exp(-((x - x0) ** 2) / (2.0 * c.packet_sigma * c.packet_sigma))
            st[2 * n:3 * n] = env * np.sin(carrier * (x - c.xmin)) * np.sin(math.pi * yt / ly)
        elif w == "gaussian":
            st[2 * n:3 * n] = np.exp(-c.pulse_sigma * (x * x + y * y + z * z))

    def simulate(self):
        c = self.config
        self.solver = DgSolver(precision=c.precision, pipeline=c.pipeline)
        self.solver.initialize(self.mesh, c.c0)
        self.solver.uploadState(self.state)
        max_steps, final_time = c.max_steps(), c.final_time
        out_freq = max(1, c.output_frequency)
        self.probe_points = self._probe_list()
        nprobe = self.probe_points.shape[0]
        pending, out, self._probe_chunks = [], None, []
        self.probe_steps, self.probe_times = [], []
        if nprobe:
            cap = min(out_freq + 1, self.PROBE_CAPACITY)
            self.solver.setProbes(self.mesh, self.probe_points, capacity=cap)
            save = c.output_dir and c.output_dir.lower() != "none"
            if save:  # probe.cpp:43-48, 117-133
                os.makedirs(c.output_dir, exist_ok=True)
                out = open(os.path.join(c.output_dir, "probe_history.csv"), "w")
                out.write("step,time" + "".join(",Ex_%d,Ey_%d,Ez_%d,Hx_%d,Hy_%d,Hz_%d" % ((q + 1,) * 6)
                                                 for q in range(nprobe)) + "\n")
                with open(os.path.join(c.output_dir, "probes.json"), "w") as fh:
                    fh.write(probes_json_text(self.mesh.getN(), self.num_modes, self.probe_points, c.xmax - c.xmin,
                                              c.nelx, self.mode_k))
            self.solver.sampleProbes()  # t = 0 (case.cpp:473-474)
            pending.append((0, 0.0))
        # field snapshots: the t = 0 fields, then every save_frequency steps and at
        # the last step (case.cpp:438-464, 603-625); the state comes off the
        # device only for those steps
        save_dir = c.output_dir and c.output_dir.lower() != "none"
        writer = None
        if save_dir and c.export_fields and c.export_format.lower() in ("hdf5", "vtk"):
            os.makedirs(c.output_dir, exist_ok=True)
            writer = FieldWriter()
            writer.setFieldSaveOptions(self.save_fields)
            writer.initialize(os.path.join(c.output_dir, "fields.h5"), self.mesh, True, c.export_continuous_vtk)
            writer.writeStep(0, 0.0, self.state)
        self.field_writer = writer
        save_freq = max(1, c.save_frequency if c.save_frequency > 0 else out_freq)
        snap = np.zeros(6 * self.mesh.getTotalPoints()) if writer else None
        # The reference's step schedule (case.cpp:562-583); the steps between two
        # outputs with equal dt go to the device as one multi-step launch with
        # one synchronisation instead of a synchronised step() each.
        chunk = [0.0, 0]  # dt, steps not launched yet

        def flush():
            if chunk[1]:
                self.solver.run(chunk[0], chunk[1])
                chunk[1] = 0

        for step, dt, t, final in step_schedule(self.dt, max_steps, final_time):
            if chunk[1] and dt != chunk[0]:
                flush()
            chunk[0] = dt
            chunk[1] += 1
            self.current_time = t
            self.current_step = step
            if nprobe:
                pending.append((step, self.current_time))
                if len(pending) == cap:  # the device history is full: fetch it before the next step
                    flush()
                    self._fetch_probes(pending, out)
            if writer is not None and (step % save_freq == 0 or final):
                flush()
                self.solver.downloadState(snap)
                writer.writeStep(step, self.current_time, snap)
            if step % out_freq == 0 or final:
                flush()
                d = self.solver.diagnostics()
                self.history.append((step, self.current_time, d["max_e"], d["max_h"], d["energy"]))
                self._fetch_probes(pending, out)
        flush()
        self._fetch_probes(pending, out)
        if out is not None:
            out.close()
        self.probe_values = (np.concatenate(self._probe_chunks) if self._probe_chunks
                             else np.zeros((0, nprobe, 6)))
        self.state = self.solver.downloadState()

    def postprocess(self):
        if self.field_writer is not None:  # case.cpp:689-693
            self.field_writer.close()
            self.field_writer = None
        if self.post is not None:
            self.post(self)
        if self.solver is not None:
            self.solver.finalize()

    def run(self):
        self.preprocess()
        self.simulate()
        self.postprocess()

    def computeTotalEnergy(self):
        """Host energy of the downloaded state (case.cpp:733-766)."""
        w3 = self.mesh.getW3()
        m = self.mesh.info()
        # jacobian from the product's geometry: J = dV / w3 per node via the
        # solver's device reduction is the cheap path; here use diagnostics
        if self.solver is not None and self.solver._h is not None:
            return self.solver.diagnostics()["energy"]
        raise RuntimeError("computeTotalEnergy needs a live solver (%d elements, %d weights)" % (m.nelt, w3.size))
