"""Rendered runs for a human look: the counterpart of the reference's visual
suite (tests/visual_suite.py). The same 18 cases (:97-142: every built-in
initial condition, the integrator/order/fallback/blending/NAD knobs, the
magnetic loop) are run through this package's driver on the GPU and drawn as
PNG panels, with a markdown index of titles, notes and command lines.

No plotting library is needed: fields become images through a colour ramp
and nearest-neighbour scaling, profiles are drawn as polylines on a raster,
and the PNG file is written with zlib. Panels carry no text; the index names
what each panel shows (left to right, top to bottom).

    python -m spd_k_amd.gallery --out docs/gallery [--only 'sod*'] [--fast] [--skip-run]
"""
import argparse
import fnmatch
import os
import shutil
import struct
import subprocess
import sys
import zlib

import numpy as np

from . import outputs

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(HERE, "_build", "spdk")
INPUTS = os.path.join(ROOT, "tests", "inputs")


# ---- PNG ---------------------------------------------------------------------
def png_bytes(img):
    """8-bit RGB image (h, w, 3) as a PNG file (one IDAT, filter 0 rows)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w, c = img.shape
    if c != 3 or h < 1 or w < 1:
        raise ValueError("png_bytes: expected an (h, w, 3) image")
    raw = np.zeros((h, 1 + 3 * w), np.uint8)
    raw[:, 1:] = img.reshape(h, 3 * w)

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw.tobytes(), 9)) + chunk(b"IEND", b""))


def read_png(data):
    """Inverse of png_bytes (8-bit RGB, filter 0 rows only): checks the
    signature and every chunk's CRC, returns the (h, w, 3) image."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        if zlib.crc32(tag + body) & 0xFFFFFFFF != crc:
            raise ValueError(f"bad CRC in chunk {tag!r}")
        if tag == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[:4]
    if depth != 8 or ctype != 2:
        raise ValueError("only 8-bit RGB is read")
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    if raw[:, 0].any():
        raise ValueError("only filter-0 rows are read")
    return raw[:, 1:].reshape(h, w, 3).copy()


# ---- drawing -------------------------------------------------------------------
# colour ramps as anchor colours at equal spacing, linearly interpolated
RAMPS = {
    "viridis": [(68, 1, 84), (59, 82, 139), (33, 145, 140), (94, 201, 98), (253, 231, 37)],
    "magma": [(0, 0, 4), (81, 18, 124), (183, 55, 121), (252, 137, 97), (252, 253, 191)],
}
WHITE, BLACK, GREY = (255, 255, 255), (0, 0, 0), (200, 200, 200)
LINE_COLOURS = [(31, 119, 180), (214, 39, 40), (44, 160, 44)]


def colourise(f, vmin, vmax, ramp="viridis"):
    """Values to RGB through a ramp; values outside [vmin, vmax] clip, and a
    non-finite value is drawn magenta."""
    f = np.asarray(f, float)
    a = np.asarray(RAMPS[ramp], float)
    span = vmax - vmin
    t = np.clip((f - vmin) / span, 0.0, 1.0) if span > 0 else np.full(f.shape, 0.5)
    t = np.where(np.isfinite(t), t, 0.0)
    x = t * (len(a) - 1)
    i = np.minimum(np.floor(x).astype(int), len(a) - 2)
    r = (x - i)[..., None]
    rgb = a[i] * (1.0 - r) + a[i + 1] * r
    rgb[~np.isfinite(f)] = (255, 0, 255)
    return np.round(rgb).astype(np.uint8)


def frame(img, colour=BLACK, width=1):
    out = np.empty((img.shape[0] + 2 * width, img.shape[1] + 2 * width, 3), np.uint8)
    out[:] = colour
    out[width:-width, width:-width] = img
    return out


def heat(field, vmin=None, vmax=None, ramp="viridis", size=240):
    """A 2-D field (ny, nx), y up, scaled by an integer factor to about
    `size` pixels along its longer side, framed, with a colour bar on the right."""
    field = np.asarray(field, float)
    fin = field[np.isfinite(field)]
    vmin = float(fin.min()) if vmin is None and fin.size else (0.0 if vmin is None else vmin)
    vmax = float(fin.max()) if vmax is None and fin.size else (1.0 if vmax is None else vmax)
    s = max(1, int(round(size / max(field.shape))))
    img = colourise(field[::-1], vmin, vmax, ramp)
    img = np.repeat(np.repeat(img, s, axis=0), s, axis=1)
    h = img.shape[0]
    bar = colourise(np.linspace(vmax, vmin, h)[:, None] * np.ones((1, 10)), vmin, vmax, ramp)
    return hstack([frame(img), frame(bar)], gap=4)


def chart(x, series, w=320, h=220, pad=6):
    """Polylines on a white raster. series: list of (y, colour, dashed)."""
    img = np.full((h, w, 3), 255, np.uint8)
    x = np.asarray(x, float)
    ys = [np.asarray(y, float) for y, _, _ in series]
    fin = np.concatenate([y[np.isfinite(y)] for y in ys]) if ys else np.zeros(1)
    lo, hi = (float(fin.min()), float(fin.max())) if fin.size else (0.0, 1.0)
    if hi <= lo:
        lo, hi = lo - 0.5, hi + 0.5
    lo, hi = lo - 0.05 * (hi - lo), hi + 0.05 * (hi - lo)
    x0, x1 = float(x.min()), float(x.max()) if x.size > 1 and x.max() > x.min() else float(x.min()) + 1.0
    px = lambda v: pad + (v - x0) / (x1 - x0) * (w - 1 - 2 * pad)
    py = lambda v: h - 1 - pad - (v - lo) / (hi - lo) * (h - 1 - 2 * pad)
    if lo < 0 < hi:  # zero line
        img[int(round(py(0.0))), pad:w - pad] = GREY
    for y, colour, dashed in series:
        X, Y = px(x), py(y)
        run = 0
        for k in range(len(X) - 1):
            if not (np.isfinite(Y[k]) and np.isfinite(Y[k + 1])):
                continue
            n = int(max(abs(X[k + 1] - X[k]), abs(Y[k + 1] - Y[k]))) + 1
            for t in np.linspace(0.0, 1.0, n + 1):
                run += 1
                if dashed and (run // 6) % 2:
                    continue
                c, r = int(round(X[k] + t * (X[k + 1] - X[k]))), int(round(Y[k] + t * (Y[k + 1] - Y[k])))
                img[max(r - 1, 0):r + 1, max(c, 0):c + 1] = colour
    return frame(img)


def hstack(imgs, gap=8):
    h = max(i.shape[0] for i in imgs)
    parts = []
    for k, i in enumerate(imgs):
        p = np.full((h, i.shape[1], 3), 255, np.uint8)
        p[:i.shape[0]] = i
        parts.append(p)
        if k < len(imgs) - 1:
            parts.append(np.full((h, gap, 3), 255, np.uint8))
    return np.concatenate(parts, axis=1)


def vstack(imgs, gap=8):
    w = max(i.shape[1] for i in imgs)
    parts = []
    for k, i in enumerate(imgs):
        p = np.full((i.shape[0], w, 3), 255, np.uint8)
        p[:, :i.shape[1]] = i
        parts.append(p)
        if k < len(imgs) - 1:
            parts.append(np.full((gap, w, 3), 255, np.uint8))
    return np.concatenate(parts, axis=0)


def margin(img, m=10):
    out = np.full((img.shape[0] + 2 * m, img.shape[1] + 2 * m, 3), 255, np.uint8)
    out[m:-m, m:-m] = img
    return out


# ---- output access -------------------------------------------------------------
def ndim(g):
    return len(g["faces"])


def centres(g, d):
    f = g["faces"][d]
    return 0.5 * (f[1:] + f[:-1])


def field(outdir, g, i, var, prefix="W_cv", nvar=6):
    """One variable as a 2-D (y, x) array (mid-z plane in 3-D) or a 1-D profile."""
    f = outputs.load_sd(outdir, g, i, var, prefix, nvar)
    if ndim(g) == 3:
        return f[f.shape[0] // 2]
    if ndim(g) == 2:
        return f[0]
    return f[0, 0]


def troubles(outdir, g, i):
    """The aggregate trouble slot (theta in [0, 1]) on the FV cells, as field()
    shapes it, or None when the run wrote none."""
    n = g["p"] + 1
    act = [d in g["faces"] for d in range(3)]
    cells = [n * g["N"][d] + 2 * outputs.nGH if act[d] else 1 for d in range(3)]
    name = os.path.join(outdir, f"troubles_N{cells[0]}_{i}_0.dat")
    if not os.path.exists(name):
        return None
    A = np.fromfile(name)
    slots = A.size // (cells[0] * cells[1] * cells[2])
    A = A.reshape(slots, cells[2], cells[1], cells[0])[-1]
    sl = lambda d: slice(outputs.nGH, -outputs.nGH) if act[d] else slice(None)
    t = A[sl(2), sl(1), sl(0)]
    if ndim(g) == 3:
        return t[t.shape[0] // 2]
    if ndim(g) == 2:
        return t[0]
    return t[0, 0]


# ---- cases (tests/visual_suite.py:97-142) --------------------------------------
KNOB_NOTE = ("density at t = 0 (top) and t = 0.2 (bottom), with the trouble map when the run has the fallback "
             "(the sine wave flags no cell)")


def _c(cid, section, title, deck, overrides, system="hydro", note=""):
    return {"id": cid, "section": section, "title": title, "input": deck, "overrides": overrides,
            "system": system, "note": note or (KNOB_NOTE if section == "knobs" else "")}


_SINE16 = ["mesh/nx1=16", "mesh/nx2=16", "mesh/nx3=1", "time/tlim=0.2", "output/dt=0.1"]
CASES = [
    _c("sine_wave_2d", "ics", "Sine wave, 2-D, one period", "sine_wave.athinput",
       ["mesh/nx3=1", "mesh/nx1=32", "mesh/nx2=32", "time/tlim=1.0", "output/dt=1.0"],
       note="density at t = 0 and t = 1 on one scale, and the y-centre cut of both (t = 0 dashed)"),
    _c("square_2d", "ics", "Square (top hat), 2-D", "square.athinput", [],
       note="density and trouble map at t = 0 (top) and the end (bottom)"),
    _c("sod_1d", "ics", "Sod shock tube", "sod.athinput", [],
       note="density, pressure, velocity: t = 0 dashed, the end solid"),
    _c("shu_osher_1d", "ics", "Shu-Osher shock-entropy interaction", "shu_osher.athinput", [],
       note="density, pressure, velocity: t = 0 dashed, the end solid"),
    _c("kelvin_helmholtz_2d", "ics", "Kelvin-Helmholtz, 2-D", "kelvin_helmholtz.athinput", [],
       note="density and trouble map at t = 0 (top) and the end (bottom)"),
    _c("implosion_2d", "ics", "Implosion, 2-D, reflective walls", "implosion.athinput",
       ["mesh/nx1=100", "mesh/nx2=100", "output/dt=2.5"],
       note="density and trouble map at t = 0 (top) and t = 2.5 (bottom)"),
    _c("user_2d", "ics", "User initial condition with viscosity", "user.athinput",
       ["mesh/nx1=32", "mesh/nx2=32", "mesh/nx3=1", "time/tlim=1.0", "output/dt=1.0", "problem/sigma=0.08",
        "problem/amp=0.5", "problem/v1=1.0", "hydro/nu=1e-3"],
       note="v_y at t = 0 and t = 1 on one scale, and the y-centre cut (t = 0 dashed): the pulse spreads"),
    _c("sedov_3d", "ics", "Sedov blast, 3-D", "sedov.athinput",
       ["mesh/nx1=32", "mesh/nx2=32", "mesh/nx3=32", "time/tlim=0.5", "output/dt=0.5"],
       note="mid-z density and trouble map at t = 0 (top) and t = 0.5 (bottom)"),
    _c("spherical_blast_2d", "ics", "Blast in a 1 x 1.5 periodic box, 2-D", "spherical_blast.athinput",
       ["mesh/nx1=100", "mesh/nx2=150", "mesh/x2len=1.5", "time/tlim=1.5", "output/dt=1.5"],
       note="density and trouble map at t = 0 (top) and t = 1.5 (bottom)"),
    _c("rti_2d", "ics", "Rayleigh-Taylor under gravity, 2-D", "rti.athinput", [],
       note="density and trouble map at t = 0 (top) and the end (bottom)"),
    _c("knob_ader", "knobs", "ADER integrator", "sine_wave.athinput",
       _SINE16 + ["time/integrator=ader", "job/fallback=true"]),
    _c("knob_rk3", "knobs", "SSP-RK3 integrator", "sine_wave.athinput",
       _SINE16 + ["time/integrator=rk3", "job/fallback=true"]),
    _c("knob_p3", "knobs", "Degree p = 3", "sine_wave.athinput", ["mesh/p=3"] + _SINE16 + ["job/fallback=true"]),
    _c("knob_p7", "knobs", "Degree p = 7", "sine_wave.athinput",
       ["mesh/p=7", "mesh/nx1=8", "mesh/nx2=8", "mesh/nx3=1", "time/tlim=0.2", "output/dt=0.1",
        "job/fallback=true"]),
    _c("knob_no_fallback", "knobs", "SD only (fallback off)", "sine_wave.athinput", _SINE16 + ["job/fallback=false"]),
    _c("knob_no_blending", "knobs", "Binary theta (blending off)", "sine_wave.athinput",
       _SINE16 + ["fallback/blending=false"]),
    _c("knob_nad_delta", "knobs", "NAD band from the local range", "sine_wave.athinput",
       _SINE16 + ["fallback/NAD=delta"]),
    _c("induction_loop_3d", "induction", "Magnetic loop, 3-D induction", "induction_loop.athinput", [],
       system="induction", note="|B_xy| in the mid-z plane at t = 0 and the end"),
]
FIELD_CUT = {"sine_wave_2d": 0, "user_2d": 2}  # case -> primitive variable drawn with a cut
FAST = {"mesh/nx1": "8", "mesh/nx2": "8", "mesh/nx3": "8", "time/tlim": "0.05", "output/dt": "0.05"}


def fast_case(case):
    """A shrunken case for a quick look (the knob section is left out)."""
    if case["section"] == "knobs":
        return None
    ov = list(case["overrides"])
    for k, v in FAST.items():
        if not any(o.split("=", 1)[0] == k for o in ov):
            ov.append(f"{k}={v}")
    if case["id"] == "implosion_2d":
        ov += ["time/tlim=0.15", "output/dt=0.075"]
    return dict(case, overrides=ov)


def command(case, driver=DRIVER):
    return [driver, "-i", os.path.join(INPUTS, case["input"])] + case["overrides"]


def run_case(case, outdir, driver=DRIVER, timeout=900):
    if os.path.isdir(outdir):
        shutil.rmtree(outdir)
    os.makedirs(outdir)
    r = subprocess.run(command(case, driver), env=dict(os.environ, SPD_OUTPUT_DIR=outdir), capture_output=True,
                       text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{case['id']}: driver exit {r.returncode}\n{r.stdout[-1500:]}{r.stderr[-1500:]}")
    return r.stdout


# ---- panels --------------------------------------------------------------------
def render(case, outdir):
    """The case's panel as an RGB image."""
    induction = case["system"] == "induction"
    idx = outputs.output_indices(outdir, "B2_cv" if induction else "W_cv")
    if not idx:
        raise RuntimeError(f"no outputs in {outdir}")
    g = outputs.grid(outdir)
    first, last = idx[0], idx[-1]
    if induction:
        mags = []
        for i in (first, last):
            bx, by = (field(outdir, g, i, v, "B2_cv", 4) for v in (0, 1))
            mags.append(np.sqrt(bx * bx + by * by))
        return margin(hstack([heat(m) for m in mags]))
    if case["id"] in FIELD_CUT:
        v = FIELD_CUT[case["id"]]
        f0, f1 = field(outdir, g, first, v), field(outdir, g, last, v)
        lo, hi = float(min(f0.min(), f1.min())), float(max(f0.max(), f1.max()))
        j = f0.shape[0] // 2
        cut = chart(centres(g, 0), [(f0[j], LINE_COLOURS[0], True), (f1[j], LINE_COLOURS[1], False)],
                    w=360, h=300)
        return margin(hstack([vstack([heat(f0, lo, hi), heat(f1, lo, hi)]), cut]))
    if ndim(g) == 1:
        x = centres(g, 0)
        plots = [chart(x, [(field(outdir, g, first, v), LINE_COLOURS[0], True),
                           (field(outdir, g, last, v), LINE_COLOURS[1], False)]) for v in (0, 4, 1)]
        return margin(hstack(plots))
    rows = []
    for i in (first, last):
        row = [heat(field(outdir, g, i, 0))]
        t = troubles(outdir, g, i)
        if t is not None:
            row.append(heat(t, 0.0, 1.0, "magma"))
        rows.append(hstack(row, gap=16))
    return margin(vstack(rows, gap=12))


def write_index(path, cases, rel):
    lines = ["# Gallery", "",
             "Runs of this package's driver drawn without a plotting library: colour ramp low (dark) to high "
             "(bright), each map with its colour bar on the right; profiles with t = 0 dashed and the final "
             "output solid.", ""]
    section = None
    for c in cases:
        if c["section"] != section:
            section = c["section"]
            lines += [f"## {section}", ""]
        cmd = " ".join(["spd_k_amd/_build/spdk", "-i", "tests/inputs/" + c["input"]] + c["overrides"])
        lines += [f"### {c['title']}", ""]
        if c["note"]:
            lines += [c["note"][0].upper() + c["note"][1:] + ".", ""]
        lines += [f"`{cmd}`", "", f"![{c['id']}]({rel}/{c['id']}.png)", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(ROOT, "docs", "gallery"))
    ap.add_argument("--runs", default=None, help="run directories (default: OUT/runs)")
    ap.add_argument("--only", default=None, help="fnmatch pattern on case ids")
    ap.add_argument("--fast", action="store_true", help="shrunken cases, knob section left out")
    ap.add_argument("--skip-run", action="store_true", help="render existing run directories")
    ap.add_argument("--driver", default=DRIVER)
    a = ap.parse_args(argv)
    runs = a.runs or os.path.join(a.out, "runs")
    os.makedirs(a.out, exist_ok=True)
    cases = [fast_case(c) if a.fast else c for c in CASES]
    cases = [c for c in cases if c and (a.only is None or fnmatch.fnmatch(c["id"], a.only))]
    done = []
    for c in cases:
        d = os.path.join(runs, c["id"])
        if not a.skip_run:
            print(f"{c['id']}: running", flush=True)
            run_case(c, d, a.driver)
        with open(os.path.join(a.out, c["id"] + ".png"), "wb") as f:
            f.write(png_bytes(render(c, d)))
        print(f"{c['id']}: drawn", flush=True)
        done.append(c)
    write_index(os.path.join(a.out, "README.md"), done, ".")
    return 0


if __name__ == "__main__":
    sys.exit(main())
