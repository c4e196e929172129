"""Analytic ground truth for reconstruction tests: line integrals of one
ellipsoid through the helical fan-beam geometry without flying focal spot,
evaluated in numpy (FP64), independently of the library's own scan
generator.

Geometry (reading n, n_ffs = 1):
  tube angle   a = TubeStartAngle + 360 n / n_proj_ffs degrees
  source       S = r_f (cos a, sin a), at table position z_n (the planner's
               table_positions, reference src/setup.cu:405-415)
  channel ch   fan angle b = (ch - central_channel) * fan_angle_increment,
               in-plane direction cos b * e1 + sin b * e2 with
               e1 = -(cos a, sin a) (towards the isocentre) and
               e2 = (sin a, -cos a)
  row j        normalised row coordinate q = 2 j / (n_rows - 1) - 1; the ray
               climbs dz = -q tan(theta_cone / 2) per mm in the plane (the row
               model of the weighted backprojection,
               include/backproject.cuh:82-85).
Voxel (xi, yi) of the output stack lies at the world point
  x = dx (xi - (nx - 1) / 2) * (-1 unless table_direction == -1) + x_origin
  y = dy (yi - (ny - 1) / 2) + y_origin
(include/backproject.cuh:59-64), final slice k at
  z = StartPos + direction * k * SliceThickness
(src/finalize_image_stack.cu:83-85).
"""
import numpy as np


class Ellipsoid:
    def __init__(self, centre, semi_axes, yaw_deg, mu):
        self.c = np.asarray(centre, dtype=np.float64)
        self.ax = np.asarray(semi_axes, dtype=np.float64)
        t = np.deg2rad(yaw_deg)
        # rows: body axes in world coordinates (rotation about z)
        self.R = np.array([[np.cos(t), np.sin(t), 0.0], [-np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]])
        self.mu = float(mu)

    def inside(self, x, y, z):
        """Normalised radius (< 1 inside) of world points."""
        p = np.stack([np.asarray(x, np.float64) - self.c[0], np.asarray(y, np.float64) - self.c[1],
                      np.asarray(z, np.float64) - self.c[2]], axis=-1)
        u = p @ self.R.T / self.ax
        return np.sqrt((u ** 2).sum(axis=-1))

    def line_integral(self, p0, d):
        """mu * chord length of the lines p0 + t d (d unit vectors), arrays [..., 3]."""
        u = ((p0 - self.c) @ self.R.T) / self.ax
        v = (d @ self.R.T) / self.ax
        A = (v * v).sum(-1)
        B = (u * v).sum(-1)
        C = (u * u).sum(-1) - 1.0
        disc = B * B - A * C
        return np.where(disc > 0, self.mu * 2.0 * np.sqrt(np.maximum(disc, 0.0)) / A, 0.0)


def scan(mr, ellipsoids, first, count):
    """Raw line integrals [count][n_rows_raw][n_channels] float32 of readings
    first..first+count-1 (readings outside the scan are zero)."""
    cg, rp = mr.cg, mr.rp
    assert mr.ri.n_ffs == 1, "no flying focal spot"
    nch, nrows = int(cg.n_channels), int(cg.n_rows_raw)
    n_read = int(rp.n_readings)
    out = np.zeros((count, nrows, nch), dtype=np.float32)
    ch = np.arange(nch, dtype=np.float64)
    b = (ch - float(cg.central_channel)) * float(cg.fan_angle_increment)
    q = 2.0 * np.arange(nrows, dtype=np.float64) / (nrows - 1) - 1.0
    slope = -q * np.tan(float(cg.theta_cone) / 2.0)
    nrm = np.sqrt(1.0 + slope ** 2)
    for i in range(count):
        n = first + i
        if n < 0 or n >= n_read:
            continue
        a = np.deg2rad(float(rp.tube_start_angle) + 360.0 * n / float(cg.n_proj_ffs))
        e1 = -np.array([np.cos(a), np.sin(a)])
        e2 = np.array([np.sin(a), -np.cos(a)])
        u = np.cos(b)[:, None] * e1[None, :] + np.sin(b)[:, None] * e2[None, :]     # [nch, 2]
        d = np.empty((nrows, nch, 3))
        d[:, :, 0] = u[None, :, 0] / nrm[:, None]
        d[:, :, 1] = u[None, :, 1] / nrm[:, None]
        d[:, :, 2] = slope[:, None] / nrm[:, None]
        p0 = np.array([float(cg.r_f) * np.cos(a), float(cg.r_f) * np.sin(a), float(mr.table_positions[n])])
        p0 = np.broadcast_to(p0, d.shape)
        acc = np.zeros((nrows, nch))
        for e in ellipsoids:
            acc += e.line_integral(p0, d)
        out[i] = acc
    return out


def voxel_world(mr):
    """World x (per xi) and y (per yi) of the output voxels."""
    rp, cg = mr.rp, mr.cg
    nx, ny = int(rp.nx), int(rp.ny)
    dx, dy = float(rp.recon_fov) / nx, float(rp.recon_fov) / ny
    sx = 1.0 if int(cg.table_direction) == -1 else -1.0
    x = sx * dx * (np.arange(nx) - (nx - 1) / 2.0) + float(rp.x_origin)
    y = dy * (np.arange(ny) - (ny - 1) / 2.0) + float(rp.y_origin)
    return x, y


def final_slice_z(mr, n_final):
    rp = mr.rp
    d = float(np.float32(rp.end_pos) - np.float32(rp.start_pos))
    direction = 1.0 if d >= 0 else -1.0
    return float(rp.start_pos) + direction * np.arange(n_final) * float(rp.slice_thickness)


def check_stack(mr, stack_hu, e, label=""):
    """Ground-truth figures of a reconstructed HU stack [n_final][nx][ny] of
    the single ellipsoid e (mu = mu_water) in air: interior and exterior HU,
    the in-plane centroid of the object (world mm) and its z centroid."""
    nf = stack_hu.shape[0]
    x, y = voxel_world(mr)
    z = final_slice_z(mr, nf)
    X, Y = np.meshgrid(x, y, indexing="ij")
    r_in = np.stack([e.inside(X, Y, np.full_like(X, zk)) for zk in z])
    interior = stack_hu[r_in < 0.6]
    # exterior: away from the object, inside the reconstructed circle
    fov_r = 0.45 * float(mr.rp.recon_fov)
    ring = (np.hypot(X - float(mr.rp.x_origin), Y - float(mr.rp.y_origin)) < fov_r)[None]
    exterior = stack_hu[(r_in > 1.5) & ring]
    mu = np.clip(stack_hu + 1000.0, 0.0, None)          # attenuation relative to air, HU units
    w = mu * (mu > 500.0)
    cx = float((w * X[None]).sum() / w.sum())
    cy = float((w * Y[None]).sum() / w.sum())
    # z profile: object area per final slice
    area = (stack_hu > -500.0).reshape(nf, -1).sum(axis=1).astype(np.float64)
    cz = float((area * z).sum() / area.sum())
    res = {"interior_mean": float(interior.mean()), "interior_rms": float(np.sqrt((interior ** 2).mean())),
           "exterior_mean": float(exterior.mean()), "cx": cx, "cy": cy, "cz": cz, "n_interior": int(interior.size)}
    print(f"ground truth {label}: " + ", ".join(f"{k} {v:.3f}" if isinstance(v, float) else f"{k} {v}"
                                               for k, v in res.items()), flush=True)
    return res


# --------------------------------------------------------------------------
# the standard case of the ground-truth tests

MU_WATER = 0.02
BARS = {"interior_mean": 2.0, "interior_rms": 3.0, "exterior": 3.0, "xy_mm": 0.5, "z_mm": 0.5}


def make_case(cb, scans, install, tmpdir, table_dir):
    """Mini scanner (128 channels, 16 rows, no flying focal spot), 64^2 over
    250 mm, 26 final slices of 2.4 mm; one water ellipsoid (semi-axes 45, 30,
    22 mm, yawed 25 deg) centred off axis at (55, -35) mm, 3 mm above the
    middle of the slice range; no noise.  Returns (mr, ellipsoid, raw scan)."""
    import os
    tok = scans.mini_tokens("none", nrows=16, nx=64, table_dir=table_dir, rotations=14, n_slices=26)
    tok["AdaptiveFiltration"] = 0.0
    tok["MuWater"] = MU_WATER
    prm = os.path.join(str(tmpdir), f"truth_{table_dir}.prm")
    cb.write_prm(prm, tok)
    mr = cb.metadata_from_prm(prm, install_dir=install)
    z_mid = 0.5 * (tok["StartPos"] + tok["EndPos"])
    e = Ellipsoid((55.0, -35.0, z_mid + 3.0), (45.0, 30.0, 22.0), 25.0, MU_WATER)
    raw = scan(mr, [e], 0, int(tok["Readings"]))
    return mr, e, raw


def failures(res, e):
    """The bars a result misses (empty list: all met)."""
    bad = []
    if abs(res["interior_mean"]) > BARS["interior_mean"]:
        bad.append("interior_mean")
    if res["interior_rms"] > BARS["interior_rms"]:
        bad.append("interior_rms")
    if abs(res["exterior_mean"] + 1000.0) > BARS["exterior"]:
        bad.append("exterior")
    if abs(res["cx"] - e.c[0]) > BARS["xy_mm"] or abs(res["cy"] - e.c[1]) > BARS["xy_mm"]:
        bad.append("centroid_xy")
    if abs(res["cz"] - e.c[2]) > BARS["z_mm"]:
        bad.append("centroid_z")
    return bad
