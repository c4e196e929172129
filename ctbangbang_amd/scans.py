"""Standard synthetic helical scans (prm token sets + phantom parameters).

Shapes follow the reference's planning arithmetic (src/setup.cu:291-297,
483-679).  "Pitch p" is expressed as the table feed per rotation the prm
expects (PitchValue = p * Nrows * CollSlicewidth, reference src/setup.cu:295).

  c0   no FFS, 736 ch x 16 rows, feed 19.2 mm, 256^2 over 250 mm, 16 slices of
       1.0 mm, 5 rotations (the reference's allowed-range check needs >= 5,
       its own range rule src/setup.cu:514-545), Ram-Lak, no adaptive filter.
  c1   phi+z FFS, 32 -> 64 rows of 0.6 mm, pitch 0.8 (30.72 mm), 10
       rotations, 512^2 over 250 mm, 40 final slices of 3 mm (224 raw slices
       in 7 blocks), adaptive filtering on, HU output.
  c4   z FFS, 64 rows of 0.5 mm, pitch 0.6 (19.2 mm), 40 rotations, 1024^2
       over 400 mm, 150 final slices of 2 mm (608 raw slices in 19 blocks).
  mini_* small scans on a 128-channel scanner for CPU-sized parity tests.
"""
import os

from . import PKG_DIR, PhantomParams, ramp_filter, write_prm

SCANNER_DIR = os.path.join(PKG_DIR, "resources", "scanners")

MINI_SCANNER = {
    "RSrcToIso": 595.0, "RSrcToDet": 1085.6, "AnodeAngle": 0.1222, "FanAngleInc": 0.0068,
    "ThetaCone": 0.030249, "CentralChannel": 63.75, "NProjTurn": 192, "NChannels": 128,
    "ReverseRowInterleave": 0, "ReverseChanInterleave": 0,
}


def base_tokens():
    return {
        "RawDataDir": ".", "RawDataFile": "synthetic.raw", "OutputDir": ".", "OutputFile": "synthetic.img",
        "Scanner": "definitionas.scanner", "FileType": 0, "FileSubType": 0, "RawOffset": 0,
        "Xorigin": 0.0, "Yorigin": 0.0, "TubeStartAngle": 0.0, "AcqFOV": 500.0, "ReconKernel": 3,
        "TableDir": "out",
    }


def c0_tokens():
    t = base_tokens()
    t.update({"Nrows": 16, "CollSlicewidth": 1.2, "Zffs": 0, "Phiffs": 0, "PitchValue": 19.2,
              "Readings": 5 * 1152, "Nx": 256, "Ny": 256, "ReconFOV": 250.0, "SliceThickness": 1.0,
              "StartPos": 40.0, "NSlices": 16, "EndPos": 55.0, "AdaptiveFiltration": 0.0})
    return t


def c1_tokens():
    t = base_tokens()
    t.update({"Nrows": 64, "CollSlicewidth": 0.6, "Zffs": 1, "Phiffs": 1, "PitchValue": 30.72,
              "Readings": 10 * 4608, "Nx": 512, "Ny": 512, "ReconFOV": 250.0, "SliceThickness": 3.0,
              "StartPos": 90.0, "NSlices": 40, "EndPos": 207.0, "AdaptiveFiltration": 1.0, "MuWater": 0.02})
    return t


def c4_tokens():
    t = base_tokens()
    t.update({"Nrows": 64, "CollSlicewidth": 0.5, "Zffs": 1, "Phiffs": 0, "PitchValue": 19.2,
              "Readings": 40 * 2304, "Nx": 1024, "Ny": 1024, "ReconFOV": 400.0, "SliceThickness": 2.0,
              "StartPos": 230.0, "NSlices": 150, "EndPos": 528.0, "AdaptiveFiltration": 1.0, "MuWater": 0.02})
    return t


def c3_tokens(i, n_rot_max=20):
    """Scan i of the batching workload (seed 3000 + i): phi+z FFS on the
    64-row scanner, pitch U(0.5, 1.5), reconstruction FOV U(200, 500) mm,
    matrix 512^2 or 768^2, U(50, 400) slices of 1 mm, with just enough
    rotations (at most n_rot_max; fewer slices when that is not enough)."""
    import random
    r = random.Random(3000 + i)
    pitch = r.uniform(0.5, 1.5)
    fov = round(r.uniform(200.0, 500.0), 1)
    nx = r.choice([512, 768])
    n_slices = r.randint(50, 400)
    feed = round(pitch * 64 * 0.6, 4)
    # slices span (n_slices - 1) mm; the pulls need about one rotation of
    # margin on either side of the slice range (with 0.5 extra as headroom)
    rot = int(min(n_rot_max, max(2, -(-((n_slices - 1) + 3.0 * feed) // feed))))
    n_slices = int(min(n_slices, max(1, (rot - 3.0) * feed + 1)))
    start = round(1.5 * feed, 3)
    t = base_tokens()
    t.update({"Nrows": 64, "CollSlicewidth": 0.6, "Zffs": 1, "Phiffs": 1, "PitchValue": feed,
              "Readings": rot * 4608, "Nx": nx, "Ny": nx, "ReconFOV": fov, "SliceThickness": 1.0,
              "StartPos": start, "NSlices": n_slices, "EndPos": start + (n_slices - 1) * 1.0,
              "AdaptiveFiltration": 1.0, "MuWater": 0.02})
    return t


def phantom_for(tokens, config_index=0, noise=True):
    """Phantom parameters: seeds 1000+config (phantom) and 2000+config (noise)."""
    pp = PhantomParams()
    pp.phantom_seed = 1000 + config_index
    pp.noise_seed = 2000 + config_index
    pp.n_random = 8
    pp.i0 = 1e5 if noise else 0.0
    pp.scale_mm = 100.0
    start, end = float(tokens["StartPos"]), float(tokens["EndPos"])
    pp.z_center = 0.5 * (start + end)
    pp.mu_water = float(tokens.get("MuWater", 0.02))
    return pp


def write_scanner(path, params):
    with open(path, "w") as f:
        f.write("% synthetic test scanner (lengths in mm, angles in radians)\n")
        for k, v in params.items():
            f.write(f"{k}:\t{v}\n")
    return path


def mini_install(dirpath, params=None):
    """Install directory with resources/ for the 128-channel test scanner:
    the scanner file and its closed-form filter tables (sharp = Ram-Lak,
    medium, smooth)."""
    import math
    params = params or MINI_SCANNER
    sdir = os.path.join(dirpath, "resources", "scanners")
    fdir = os.path.join(dirpath, "resources", "filters")
    os.makedirs(sdir, exist_ok=True)
    os.makedirs(fdir, exist_ok=True)
    write_scanner(os.path.join(sdir, "mini.scanner"), params)
    n = 2 * int(params["NChannels"])
    ds = float(params["RSrcToIso"]) * math.sin(float(params["FanAngleInc"]) / 2.0)
    for name, (c, a) in {"sharp": (1, 1), "ramp": (1, 1), "medium": (0.65, 0.5), "smooth": (0.4, 0.5)}.items():
        ramp_filter(n, ds, c, a).tofile(os.path.join(fdir, f"f_{params['NChannels']}_{name}.txt"))
    return dirpath


def mini_tokens(ffs="none", nrows=8, nx=32, table_dir="out", rotations=None, n_slices=4):
    """Small scan on the 128-channel scanner: ffs in none/phi/z/both;
    n_slices final slices of 2.4 mm centred on the scan."""
    t = base_tokens()
    phi = 1 if ffs in ("phi", "both") else 0
    z = 1 if ffs in ("z", "both") else 0
    n_ffs = (2 if phi else 1) * (2 if z else 1)
    coll = 1.2
    feed = 0.8 * nrows * coll
    rot = rotations or 12
    t.update({"Scanner": "mini.scanner", "Nrows": nrows, "CollSlicewidth": coll, "Zffs": z, "Phiffs": phi,
              "PitchValue": feed, "Readings": rot * 192 * n_ffs, "Nx": nx, "Ny": nx, "ReconFOV": 250.0,
              "SliceThickness": 2.4, "TableDir": table_dir, "AdaptiveFiltration": 1.0, "MuWater": 0.02})
    length = rot * feed
    t["StartPos"] = round(0.5 * length - 8.0 - 1.2 * (n_slices - 4), 3)
    t["NSlices"] = n_slices
    t["EndPos"] = t["StartPos"] + (n_slices - 1) * 2.4
    return t


def write_config(path, tokens):
    return write_prm(path, tokens)
