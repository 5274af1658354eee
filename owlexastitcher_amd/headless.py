"""Headless benchmark harness and command line of the reference viewer.

    python -m owlexastitcher_amd.headless model.exum -grids model.grids \\
        -scalars model.scalars -xf cloud.xf --camera px py pz ix iy iz ux uy uz \\
        -fovy 70 --size 1024 1024 -o out -fps fps.log

Follows viewer.cpp:119-158 (defaults), 755-885 (flags), 919-975 (renderer
set-up), 1476-1513 (headless set-up) and headless.cpp:97-186 (frame loop, log
and fps files), 250-298 (.xf file, xfRange).  The element mesh is this
library's EXAELEM1 file (the reference reads a umesh file); ``-cells`` with
``-scalars`` builds the dual mesh in memory instead of reading mesh files;
``-bricks`` with ``-scalars`` renders with the ExaBricks sampler
(OWLRenderer.cpp:145-148) when no mesh or grids file is given.

Where the reference's headless binary parses a flag without applying it
(-spp, --heat-map, -dt, --range), this harness applies it, as the reference's
interactive viewer does through its widgets.  Two flags are additions:
``--frames N`` (frames rendered, default 510) and ``--screenshot-frame N``
(default 10).  Flags that select samplers or modes that are not built (-bm,
-qc, -kdtree, -majorants, -mesh) are rejected.
"""
import datetime
import shutil
import struct
import subprocess
import sys
import time
import zlib

import numpy as np

from .host import AmrScene, DualMesh, ExaBrickModel, ExaStitchModel, OWLRenderer, camera_look_at, xf_load

CLIP_PLANES_MAX = 1  # common.h:34
UNSUPPORTED = {"-bm": "BigMesh", "-qc": "QuickClusters", "-kdtree": "k-d tree majorants",
               "-majorants": "own majorants", "-mesh": "triangle meshes"}


class Cmdline:
    """viewer.cpp:119-158 defaults."""

    def __init__(self):
        self.inFileName = ""
        self.scalarFileName = ""
        self.gridsFileName = ""
        self.amrCellFileName = ""
        self.exaBrickFileName = ""
        self.xfFileName = ""
        self.outFileName = "Witcher3.png"
        self.fpsFileName = "fps.log"
        self.valueRange = (1e30, -1e30)
        self.subImage = (1e30, 1e30, -1e30, -1e30)
        self.camera = {"vp": (0.0, 0.0, 0.0), "vi": (0.0, 0.0, 0.0), "vu": (0.0, 0.0, 0.0), "fovy": 70.0}
        self.clipPlanes = [{"enabled": False, "N": (0.0, 0.0, 1.0), "d": 1000.0} for _ in range(CLIP_PLANES_MAX)]
        self.remap_from = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
        self.remap_to = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
        self.rendererType = 0
        self.shadeMode = 0
        self.light = {"pos": (0.0, 0.0, 0.0), "intensity": 0.0, "on": False}
        self.numMCs = (128, 128, 128)
        self.windowSize = (1024, 1024)
        self.dt = None
        self.spp = 1
        self.heatMapEnabled = False
        self.heatMapScale = 1.0
        self.frames = 510
        self.screenshotFrame = 10


class UsageError(Exception):
    pass


def parse(argv):
    """viewer.cpp:755-885 flag parsing."""
    c = Cmdline()
    clip = 0
    i = 0

    def take(n, conv=float):
        nonlocal i
        if i + n >= len(argv):
            raise UsageError(f"missing value(s) after '{argv[i]}'")
        vals = [conv(v) for v in argv[i + 1:i + 1 + n]]
        i += n
        return vals

    while i < len(argv):
        a = argv[i]
        if not a.startswith("-"):
            c.inFileName = a
        elif a in ("-scalar", "-scalars"):
            c.scalarFileName = take(1, str)[0]
        elif a in ("-grids", "-gridlets"):
            c.gridsFileName = take(1, str)[0]
        elif a == "-cells":
            c.amrCellFileName = take(1, str)[0]
        elif a == "-bricks":
            c.exaBrickFileName = take(1, str)[0]
        elif a in UNSUPPORTED:
            raise UsageError(f"'{a}' selects the {UNSUPPORTED[a]} mode, which is not built")
        elif a == "-xf":
            c.xfFileName = take(1, str)[0]
        elif a == "--range":
            c.valueRange = tuple(take(2))
        elif a == "--subimg":
            c.subImage = tuple(take(4))
        elif a in ("-fovy", "--fov"):
            c.camera["fovy"] = take(1)[0]
        elif a == "--camera":
            v = take(9)
            c.camera["vp"], c.camera["vi"], c.camera["vu"] = tuple(v[0:3]), tuple(v[3:6]), tuple(v[6:9])
        elif a in ("-win", "--win", "--size"):
            c.windowSize = tuple(take(2, int))
        elif a == "-o":
            c.outFileName = take(1, str)[0]
        elif a == "-fps":
            c.fpsFileName = take(1, str)[0]
        elif a in ("-spp", "--spp"):
            c.spp = take(1, int)[0]
        elif a == "--heat-map":
            c.heatMapEnabled = True
            c.heatMapScale = take(1)[0]
        elif a == "-dt":
            c.dt = take(1)[0]
        elif a == "--remap-from":
            c.remap_from = tuple(take(6))
        elif a == "--remap-to":
            c.remap_to = tuple(take(6))
        elif a == "--clip-plane":
            if clip >= CLIP_PLANES_MAX:
                raise UsageError(f"at most {CLIP_PLANES_MAX} clip plane(s)")
            v = take(4)
            c.clipPlanes[clip] = {"enabled": True, "N": tuple(v[0:3]), "d": v[3]}
            clip += 1
        elif a == "-rt":
            c.rendererType = take(1, int)[0]
        elif a == "-sm":
            c.shadeMode = take(1, int)[0]
        elif a == "--num-mcs":
            c.numMCs = tuple(take(3, int))
        elif a == "--light":
            v = take(4)
            c.light = {"pos": tuple(v[0:3]), "intensity": v[3], "on": True}
        elif a == "--frames":
            c.frames = take(1, int)[0]
        elif a == "--screenshot-frame":
            c.screenshotFrame = take(1, int)[0]
        else:
            raise UsageError(f"unknown cmdline arg '{a}'")
        i += 1
    if not c.amrCellFileName and not c.inFileName and not c.gridsFileName and not c.exaBrickFileName:
        raise UsageError("no model: give an element-mesh file and/or -grids, -bricks, or -cells")
    if c.frames < 1:
        raise UsageError("--frames must be at least 1")
    return c


def default_color_map(n=128):
    """Diverging blue-grey-red colour map with a linear alpha ramp.  The
    reference starts from map 0 of its TF editor's library (absent from the
    reference); this is the library's own default."""
    x = np.linspace(0.0, 1.0, n)
    stops = np.array([[59, 76, 192], [221, 221, 221], [180, 4, 38]], np.float64) / 255.0
    rgb = np.stack([np.interp(x, [0.0, 0.5, 1.0], stops[:, k]) for k in range(3)], -1)
    return np.concatenate([rgb, x[:, None]], -1).astype(np.float32)


def write_png(path, rgba):
    """8-bit RGBA PNG (no filtering, zlib level 6)."""
    h, w, _ = rgba.shape
    raw = b"".join(b"\0" + rgba[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def xf_range(abs_domain, rel_domain):
    """headless.cpp:285-298 xfRange: the ordered absolute domain lerped at the
    ordered relative domain / 100, in float32 as the reference computes it."""
    f = np.float32
    a = sorted(f(v) for v in abs_domain)
    r = sorted(f(0.01) * f(v) for v in rel_domain)
    lerp = lambda t: (f(1.0) - t) * a[0] + t * a[1]  # noqa: E731
    return float(lerp(r[0])), float(lerp(r[1]))


def gpu_info():
    """The reference logs nvidia-smi; the ROCm counterpart is rocm-smi."""
    tool = shutil.which("rocm-smi")
    if not tool:
        return ""
    try:
        return subprocess.run([tool, "--showproductname"], capture_output=True, text=True, timeout=30).stdout
    except (OSError, subprocess.SubprocessError):
        return ""


class Headless:
    """headless.cpp: a fixed-size frame buffer, a camera, the .xf state and
    the benchmark loop."""

    def __init__(self, renderer, size):
        self.renderer = renderer
        self.outFileName = ""
        self.outFileNameFPS = ""
        self.xf = {"opacityScale": 100.0, "absDomain": np.array([0.0, 1.0], np.float32),
                   "relDomain": np.array([0.0, 99.0], np.float32), "colorMap": None}
        self.camera = None
        self.fb = None
        self.resize(size)

    def setOutFileName(self, name):
        self.outFileName = name

    def setOutFileNameFPS(self, name):
        self.outFileNameFPS = name

    def resize(self, size):
        import torch
        self.fbSize = (int(size[0]), int(size[1]))
        self.fb = torch.zeros((self.fbSize[1], self.fbSize[0]), dtype=torch.int32, device="cuda")
        self.renderer.resize(*self.fbSize)
        self.renderer.resetAccum()
        if self.camera is not None:
            self.cameraChanged()

    def setCameraOrientation(self, origin, interest, up, fovy):
        self.camera = (tuple(origin), tuple(interest), tuple(up), float(fovy))
        self.cameraChanged()

    def cameraChanged(self):
        """viewer.cpp:550-560: SimpleCamera -> renderer camera, accumulation reset."""
        frm, at, up, fovy = self.camera
        self.renderer.setCamera(*camera_look_at(frm, at, up, fovy, *self.fbSize))
        self.renderer.resetAccum()

    def loadTransferFunction(self, path):
        self.xf = xf_load(path)

    def xfRange(self):
        return xf_range(self.xf["absDomain"], self.xf["relDomain"])

    def screenShot(self, path):
        """Rows flipped so that the file starts with the top row, alpha 255."""
        px = self.fb.cpu().numpy().view(np.uint32)
        px = (px | np.uint32(0xFF << 24))[::-1]
        write_png(path, np.ascontiguousarray(px).view(np.uint8).reshape(px.shape[0], px.shape[1], 4))

    def run(self, frames=510, screenshot_frame=10):
        """headless.cpp:97-186: render `frames` frames, time each, save a PNG at
        screenshot_frame, average the frames after it."""
        import torch
        log = []

        def out(s):
            log.append(s)
            sys.stdout.write(s)
            sys.stdout.flush()

        out("Benchmark summary:\n==================\n\n")
        out("\n")
        out(gpu_info())
        out("\nBechmark:\n")
        out("FRAME_ID;SEC.\n")
        shot = (self.outFileName + ".png") if self.outFileName else "offline.png"
        total, count = 0.0, 0
        for frameID in range(1, frames + 1):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            self.renderer.render(self.fb)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            out(f"{frameID};{t2 - t1:g}\n")
            if frameID == screenshot_frame:
                self.screenShot(shot)
                print(f"#owl.viewer: frame buffer written to {shot}")
            if frameID > screenshot_frame:
                total += t2 - t1
                count += 1
        avg = total / count if count else float("nan")
        fps = 1.0 / avg if count else float("nan")
        out(f"{avg:g} avg. secs over {count} time\n")
        out(f"{fps:g} FPS\n")
        with open((self.outFileName + ".log") if self.outFileName else "benchmark.log", "w") as f:
            f.write("".join(log))
        fps_name = self.outFileNameFPS if self.outFileName else "fps.log"
        with open(fps_name, "a") as f:
            f.write(f"{datetime.datetime.now():%Y-%m-%d %H:%M:%S},{self.outFileName},{fps:g}\n")
        return {"avg_secs": avg, "fps": fps, "frames": frames, "timed_frames": count}


def build_renderer(c):
    """viewer.cpp:919-975 without the GUI: model, renderer, colour map, clip
    planes, integrator, shade mode, light."""
    if c.exaBrickFileName and not (c.inFileName or c.gridsFileName):
        # OWLRenderer.cpp:141-148: mesh or grids first, then bricks + scalars
        if not c.scalarFileName:
            raise UsageError("-bricks needs -scalars")
        model = ExaBrickModel.load(c.exaBrickFileName, c.scalarFileName)
    elif c.amrCellFileName:
        if not c.scalarFileName:
            raise UsageError("-cells needs -scalars")
        model = ExaStitchModel.from_dual_mesh(DualMesh(AmrScene.load(c.amrCellFileName, c.scalarFileName)))
    else:
        model = ExaStitchModel.load(c.inFileName or None, c.gridsFileName or None, c.scalarFileName or None)
    r = OWLRenderer(model, c.numMCs, c.remap_from, c.remap_to)
    r.setColorMap(default_color_map())
    for i, cp in enumerate(c.clipPlanes):
        r.setClipPlane(i, cp["enabled"], cp["N"], cp["d"])
    r.setType(c.rendererType)
    r.setShadeMode(c.shadeMode)
    b = r.modelBounds
    light = dict(c.light)
    if not light["on"]:
        light["pos"], light["intensity"] = tuple(b[3:]), 60.0
    r.setLightSource(0, light["pos"], light["intensity"], light["on"])
    lo_x, lo_y, hi_x, hi_y = c.subImage
    if lo_x <= hi_x and lo_y <= hi_y:
        r.setSubImage((lo_x, lo_y, hi_x, hi_y), True)
    r.setSpp(c.spp)
    if c.heatMapEnabled:
        r.setHeatMap(True, c.heatMapScale)
    if c.dt is not None:
        r.setDt(c.dt)
    return model, r


def run(argv):
    """The headless viewer: returns the run summary (see Headless.run)."""
    c = parse(argv)
    model, r = build_renderer(c)
    b = r.modelBounds
    viewer = Headless(r, c.windowSize)
    viewer.setOutFileName(c.outFileName)
    viewer.setOutFileNameFPS(c.fpsFileName)
    if any(v != 0.0 for v in c.camera["vu"]):
        viewer.setCameraOrientation(c.camera["vp"], c.camera["vi"], c.camera["vu"], c.camera["fovy"])
    else:
        center = 0.5 * (b[:3] + b[3:])
        span = b[3:] - b[:3]
        viewer.setCameraOrientation(center + np.array([-0.3, 0.7, 1.0]) * span, center, (0.0, 1.0, 0.0), 70.0)
    abs_domain = r.valueRange
    if c.valueRange[0] <= c.valueRange[1]:
        abs_domain = np.array(c.valueRange, np.float32)
    if c.xfFileName:
        viewer.loadTransferFunction(c.xfFileName)
        r.setColorMap(viewer.xf["colorMap"])
        viewer.xf["absDomain"] = abs_domain
        lo, hi = viewer.xfRange()
        r.setRange(min(lo, hi), max(lo, hi))
        r.setRelDomain(*viewer.xf["relDomain"])
        r.setOpacityScale(viewer.xf["opacityScale"])
        r.resetAccum()
    else:
        r.setRange(float(abs_domain[0]), float(abs_domain[1]))
    summary = viewer.run(c.frames, c.screenshotFrame)
    summary["renderer"] = r
    summary["model"] = model
    summary["viewer"] = viewer
    return summary


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    try:
        run(argv)
    except UsageError as e:
        sys.stderr.write(f"\nFatal error: {e}\n\nUsage: python -m owlexastitcher_amd.headless <elements.exum> "
                         "[-grids f.grids] [-scalars f.scalars] [-cells f.cells] [-xf f.xf] [--camera 9 floats] "
                         "[-fovy deg] [--size W H] [-o out] [-fps file] [-spp n] [-rt type] [-sm mode] "
                         "[--num-mcs x y z] [--light x y z I] [--clip-plane nx ny nz d] [--frames n]\n")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
