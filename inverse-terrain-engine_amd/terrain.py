"""Host-side terrain preparation: the inputs of the batch evaluator.

Mirrors the reference's prepared-terrain pipeline so a caller holding a DEM gets the
exact f32 face jumps the reference's kernels consume:
  FineTerrain::new          crates/itr-hydro/src/terrain.rs:204-234
  TerrainBase::build        crates/itr-hydro/src/terrain.rs:262-321 (Signed and Crest)
  PreparedTerrain::apply_edit  crates/itr-hydro/src/terrain.rs:460-505
  Boundaries::from_segments crates/itr-hydro/src/terrain.rs:151-187
  DesignSpace::materialize  crates/itr-core/src/design.rs:318-385
All elevation arithmetic is f64 (numpy float64 is IEEE); face jumps are rounded to f32
once.  Sums that the reference takes sequentially are taken sequentially here too.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

G = 9.81  # terrain.rs:259

GHOST_WALL, GHOST_TRANSMISSIVE, GHOST_STAGE, GHOST_INFLOW = 0, 1, 2, 3
SIDES = ("north", "south", "west", "east")


@dataclass
class GeoTransform:
    """raster.rs:10-32: row 0 is north; cell centre x = ox + (c+0.5)dx, y = oy - (r+0.5)dy."""
    origin_x: float
    origin_y: float
    dx: float
    dy: float

    def cell_center(self, col, row):
        return self.origin_x + (col + 0.5) * self.dx, self.origin_y - (row + 0.5) * self.dy

    def to_grid(self, x, y):
        return (x - self.origin_x) / self.dx, (self.origin_y - y) / self.dy

    def cell_area(self):
        return self.dx * self.dy


@dataclass
class FineTerrain:
    nx: int
    ny: int
    geo: GeoTransform
    z: np.ndarray          # (ny, nx) re-based, walls at z_wall
    wall: np.ndarray       # (ny, nx) bool
    z_ref: float
    z_wall: float
    n: np.ndarray          # (ny, nx) Manning n
    dem: np.ndarray = None  # (ny, nx) absolute elevations as given

    @staticmethod
    def new(dem, walls=None, manning=0.035, geo: GeoTransform | None = None, dx=1.0) -> "FineTerrain":
        dem = np.asarray(dem, dtype=np.float64)
        ny, nx = dem.shape
        if geo is None:
            geo = GeoTransform(0.0, ny * dx, dx, dx)
        walls = np.zeros((ny, nx), bool) if walls is None else np.asarray(walls, bool)
        ok = ~walls & np.isfinite(dem)
        if ok.any():
            zmin, zmax = float(dem[ok].min()), float(dem[ok].max())
        else:
            zmin = zmax = 0.0
        z_ref = math.floor(zmin)
        z_wall = (zmax - z_ref) + 100.0
        wall = walls | ~np.isfinite(dem)
        z = np.where(wall, z_wall, dem - z_ref)
        n = np.broadcast_to(np.asarray(manning, np.float64), (ny, nx)).copy()
        return FineTerrain(nx, ny, geo, z, wall, z_ref, z_wall, n, dem)


@dataclass
class TerrainEdit:
    """Sparse fine-resolution edit (design.rs:144-166): bbox + dz (h, w) + volumes."""
    c0: int = 0
    r0: int = 0
    dz: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    cut_m3: float = 0.0
    fill_m3: float = 0.0

    @property
    def w(self):
        return self.dz.shape[1]

    @property
    def h(self):
        return self.dz.shape[0]

    def is_empty(self):
        return self.w == 0 or self.h == 0 or not np.any(self.dz != 0.0)

    def as_dict(self):
        return {"c0": self.c0, "r0": self.r0, "dz": self.dz, "cut_m3": self.cut_m3, "fill_m3": self.fill_m3}

    def dense(self, nx, ny):
        out = np.zeros((ny, nx))
        if self.w and self.h:
            out[self.r0:self.r0 + self.h, self.c0:self.c0 + self.w] = self.dz
        return out


def _coarse(fine: FineTerrain, lv: int, zfine: np.ndarray, c0=0, r0=0, c1=None, r1=None):
    """coarse_cell (terrain.rs:357-373) over the coarse cells [r0, r1) x [c0, c1) (default:
    the whole level), sequential block sums."""
    c1 = -(-fine.nx // lv) if c1 is None else c1
    r1 = -(-fine.ny // lv) if r1 is None else r1
    nx, ny = c1 - c0, r1 - r0
    s = np.zeros((ny, nx))
    sn = np.zeros((ny, nx))
    sn_all = np.zeros((ny, nx))
    cnt = np.zeros((ny, nx), np.int64)
    tot = np.zeros((ny, nx), np.int64)
    for dr in range(lv):
        for dc in range(lv):
            rr = np.arange(r0, r1) * lv + dr
            cc = np.arange(c0, c1) * lv + dc
            vr = rr < fine.ny
            vc = cc < fine.nx
            valid = vr[:, None] & vc[None, :]
            rri = np.minimum(rr, fine.ny - 1)
            cci = np.minimum(cc, fine.nx - 1)
            zz = zfine[np.ix_(rri, cci)]
            ww = fine.wall[np.ix_(rri, cci)]
            nn = fine.n[np.ix_(rri, cci)]
            tot += valid
            sn_all = np.where(valid, sn_all + nn, sn_all)
            take = valid & ~ww
            s = np.where(take, s + zz, s)
            sn = np.where(take, sn + nn, sn)
            cnt += take
    with np.errstate(invalid="ignore", divide="ignore"):
        z = np.where(cnt == 0, fine.z_wall, s / np.maximum(cnt, 1))
        n = np.where(cnt == 0, sn_all / np.maximum(tot, 1), sn / np.maximum(cnt, 1))
    return z, cnt == 0, n


@dataclass
class TerrainBase:
    """One fidelity level: grid, face jumps (f64, rounded per use), walls, wet, gn2, boundaries."""
    fine: FineTerrain
    level: int
    nx: int
    ny: int
    dx: float
    dy: float
    crest: bool
    z: np.ndarray         # (ny, nx) coarse z
    wall: np.ndarray
    nval: np.ndarray
    fx0: np.ndarray       # (ny, nx+1) face jumps, f32
    fx1: np.ndarray
    fy0: np.ndarray       # (ny+1, nx)
    fy1: np.ndarray
    wet: np.ndarray       # (ny, nx) f32
    gn2: np.ndarray       # (ny, nx) f32
    wet_count: int
    ghosts: dict          # side -> int32 codes
    g: float = G
    segments: tuple = ()  # the BoundarySegments the ghosts were built from

    @property
    def z_ref(self):
        return self.fine.z_ref

    @staticmethod
    def build(fine: FineTerrain, level: int = 1, segments=()) -> "TerrainBase":
        lv = max(int(level), 1)
        crest = lv != 1
        z, wall, nval = _coarse(fine, lv, fine.z)
        ny, nx = z.shape
        fx0, fx1, fy0, fy1 = _faces(fine, lv, crest, z, fine.z, nx, ny)
        wet = np.where(wall, 0.0, 1.0).astype(np.float32)
        gn2 = (G * nval * nval).astype(np.float32)
        ghosts = boundaries_from_segments(segments, nx, ny, fine.geo.dx * lv, fine.geo.dy * lv, fine.z_ref)
        return TerrainBase(fine, lv, nx, ny, fine.geo.dx * lv, fine.geo.dy * lv, crest, z, wall, nval, fx0, fx1,
                           fy0, fy1, wet, gn2, int((~wall).sum()), ghosts, segments=tuple(segments))

    def ghost_codes(self) -> np.ndarray:
        return np.concatenate([self.ghosts[s]["code"] for s in SIDES]).astype(np.int32)

    def ghost_params(self) -> np.ndarray:
        """Per boundary cell, in ghost-code order: stage level minus z_ref (Stage) or
        1 / segment length (Inflow)."""
        return np.concatenate([self.ghosts[s]["param"] for s in SIDES]).astype(np.float64)

    def ghost_hydro(self) -> np.ndarray:
        return np.concatenate([self.ghosts[s]["hydro"] for s in SIDES]).astype(np.int32)

    @property
    def has_stage(self) -> bool:
        return bool((self.ghost_codes() == GHOST_STAGE).any())


def _crest_line_x(fine: FineTerrain, lv, zfine, r0, r1, ia, ib):
    """The crest lines of face_x (terrain.rs:383-397) for the x-faces i in [ia, ib) of
    coarse rows [r0, r1), as an array: the max over the non-wall fine cells either side of
    each face (fine columns i*lv-1 and min(i*lv, nx-1); walls and rows past the grid read
    -inf), z_wall where all of them are walls.  Maxima of finite values do not depend on
    their order, so this equals the reference's loop."""
    rr = np.arange(r0, r1)[:, None] * lv + np.arange(lv)[None, :]          # [nr][lv] fine rows
    out = rr >= fine.ny
    rri = np.minimum(rr, fine.ny - 1)
    ii = np.arange(ia, ib)
    cols = np.stack([ii * lv - 1, np.minimum(ii * lv, fine.nx - 1)], axis=-1)  # [nf][2] fine columns
    R, Cc = rri[:, :, None, None], cols[None, None, :, :]
    wall = fine.wall[R, Cc] | out[:, :, None, None]
    crest = np.where(wall, -np.inf, zfine[R, Cc]).max(axis=(1, 3))
    return np.where(wall.all(axis=(1, 3)), fine.z_wall, crest)


def _crest_line_y(fine: FineTerrain, lv, zfine, fa, fb, c0, c1):
    """The crest lines of face_y (terrain.rs:410-424) for the y-faces fa .. fb-1 (face f
    joins coarse rows f-1 and f; fine rows f*lv-1 and min(f*lv, ny-1)) of columns [c0, c1),
    the same way as _crest_line_x."""
    cc = np.arange(c0, c1)[:, None] * lv + np.arange(lv)[None, :]          # [nc][lv] fine columns
    out = cc >= fine.nx
    cci = np.minimum(cc, fine.nx - 1)
    ff = np.arange(fa, fb)
    rows = np.stack([ff * lv - 1, np.minimum(ff * lv, fine.ny - 1)], axis=-1)  # [nf][2] fine rows
    R, Cc = rows[:, :, None, None], cci[None, None, :, :]                     # [nf][2][nc][lv]
    wall = fine.wall[R, Cc] | out[None, None, :, :]
    crest = np.where(wall, -np.inf, zfine[R, Cc]).max(axis=(1, 3))
    return np.where(wall.all(axis=(1, 3)), fine.z_wall, crest)


def _crest_x_block(fine: FineTerrain, lv, zfine, z, r0, r1, ia, ib):
    """face_x crest (terrain.rs:378-402) for the x-faces i in [ia, ib) of coarse rows
    [r0, r1), as arrays: the crest line (_crest_line_x), then max with both coarse levels;
    the rises to the crest from the left and right levels, rounded to f32."""
    crest = _crest_line_x(fine, lv, zfine, r0, r1, ia, ib)
    zl, zr = z[r0:r1, ia - 1:ib - 1], z[r0:r1, ia:ib]
    zf = np.maximum(np.maximum(crest, zl), zr)
    return (zf - zl).astype(np.float32), (zf - zr).astype(np.float32)


def _crest_y_block(fine: FineTerrain, lv, zfine, z, fa, fb, c0, c1):
    """face_y crest (terrain.rs:406-429) for the y-faces fa .. fb-1 of columns [c0, c1), as
    arrays, the same way as _crest_x_block."""
    crest = _crest_line_y(fine, lv, zfine, fa, fb, c0, c1)
    zt, zb = z[fa - 1:fb - 1, c0:c1], z[fa:fb, c0:c1]
    zf = np.maximum(np.maximum(crest, zt), zb)
    return (zf - zt).astype(np.float32), (zf - zb).astype(np.float32)


def _faces(fine, lv, crest, z, zfine, nx, ny):
    fx0 = np.zeros((ny, nx + 1), np.float32)
    fx1 = np.zeros((ny, nx + 1), np.float32)
    fy0 = np.zeros((ny + 1, nx), np.float32)
    fy1 = np.zeros((ny + 1, nx), np.float32)
    if not crest:
        fx0[:, 1:nx] = (z[:, 1:] - z[:, :-1]).astype(np.float32)
        fy0[1:ny, :] = (z[1:, :] - z[:-1, :]).astype(np.float32)
        return fx0, fx1, fy0, fy1
    if nx > 1:
        fx0[:, 1:nx], fx1[:, 1:nx] = _crest_x_block(fine, lv, zfine, z, 0, ny, 1, nx)
    if ny > 1:
        fy0[1:ny, :], fy1[1:ny, :] = _crest_y_block(fine, lv, zfine, z, 1, ny, 0, nx)
    return fx0, fx1, fy0, fy1


def boundaries_from_segments(segments, nx, ny, dx, dy, z_ref):
    """Boundaries::from_segments (terrain.rs:151-187).  segment: dict(kind, side, range_m,
    stage_m, hydrograph)."""
    out = {s: {"code": np.zeros(nx if s in ("north", "south") else ny, np.int32),
               "param": np.zeros(nx if s in ("north", "south") else ny),
               "hydro": np.zeros(nx if s in ("north", "south") else ny, np.int32)} for s in SIDES}
    hydro = []
    for seg in segments:
        side = seg["side"]
        n, d = (nx, dx) if side in ("north", "south") else (ny, dy)
        rng = seg.get("range_m")
        if rng is None:
            a, e = 0, n
        else:
            a = min(int(max(math.floor(rng[0] / d), 0.0)), n)
            e = min(int(max(math.ceil(rng[1] / d), 0.0)), n)
        len_m = max(e - a, 0) * d
        kind = seg["kind"]
        if kind == "stage":
            code, param = GHOST_STAGE, seg["stage_m"] - z_ref
        elif kind == "transmissive":
            code, param = GHOST_TRANSMISSIVE, 0.0
        else:
            hydro.append(seg["hydrograph"])
            code, param = GHOST_INFLOW, (1.0 / len_m if len_m > 0.0 else 0.0)
            out[side]["hydro"][a:e] = len(hydro) - 1
        out[side]["code"][a:e] = code
        out[side]["param"][a:e] = param
    out["hydrographs"] = hydro
    return out


class PreparedTerrain:
    """Per-candidate terrain: base faces + the faces an edit changes (terrain.rs:431-515)."""

    def __init__(self, base: TerrainBase):
        self.base = base
        self.fx0, self.fx1, self.fy0, self.fy1 = base.fx0, base.fx1, base.fy0, base.fy1
        self.z_level = base.z
        self.edit = None

    def apply_edit(self, edit: TerrainEdit | None):
        b = self.base
        self.fx0, self.fx1, self.fy0, self.fy1 = b.fx0, b.fx1, b.fy0, b.fy1
        self.z_level = b.z
        self.edit = edit
        if edit is None or edit.is_empty():
            return self
        lv, nx, ny, fine = b.level, b.nx, b.ny, b.fine
        c0, r0 = edit.c0 // lv, edit.r0 // lv
        c1 = min(-(-(edit.c0 + edit.w) // lv), nx)
        r1 = min(-(-(edit.r0 + edit.h) // lv), ny)
        z = b.z.copy()
        if lv == 1:
            # coarse_cell at level 1: 0 + (z + dz) for open cells, z_wall for walls
            dzw = edit.dz[:r1 - r0, :c1 - c0]
            win = 0.0 + (fine.z[r0:r1, c0:c1] + dzw)
            z[r0:r1, c0:c1] = np.where(fine.wall[r0:r1, c0:c1], fine.z_wall, win)
            zfine = None
        else:
            zfine = fine.z + edit.dense(fine.nx, fine.ny)
            zc, _, _ = _coarse(fine, lv, zfine, c0, r0, c1, r1)
            z[r0:r1, c0:c1] = zc
        fx0, fx1, fy0, fy1 = b.fx0.copy(), b.fx1.copy(), b.fy0.copy(), b.fy1.copy()
        if not b.crest:
            ia, ib = max(c0 - 1, 0) + 1, min(c1, nx - 1) + 1
            if ia < ib and r0 < r1:
                fx0[r0:r1, ia:ib] = (z[r0:r1, ia:ib] - z[r0:r1, ia - 1:ib - 1]).astype(np.float32)
            ra, rb = max(r0 - 1, 0) + 1, min(r1, ny - 1) + 1
            if ra < rb and c0 < c1:
                fy0[ra:rb, c0:c1] = (z[ra:rb, c0:c1] - z[ra - 1:rb - 1, c0:c1]).astype(np.float32)
            self.fx0, self.fx1, self.fy0, self.fy1 = fx0, fx1, fy0, fy1
            self.z_level = z
            return self
        # the faces of the edited coarse window (terrain.rs:480-500): x-faces
        # max(c0-1, 0)+1 ..= min(c1, nx-1) of rows r0..r1, y-faces max(r0-1, 0)+1 ..= min(r1, ny-1)
        ia, ib = max(c0 - 1, 0) + 1, min(c1, nx - 1) + 1
        if ia < ib and r0 < r1:
            fx0[r0:r1, ia:ib], fx1[r0:r1, ia:ib] = _crest_x_block(fine, lv, zfine, z, r0, r1, ia, ib)
        fa, fb = max(r0 - 1, 0) + 1, min(r1, ny - 1) + 1
        if fa < fb and c0 < c1:
            fy0[fa:fb, c0:c1], fy1[fa:fb, c0:c1] = _crest_y_block(fine, lv, zfine, z, fa, fb, c0, c1)
        self.fx0, self.fx1, self.fy0, self.fy1 = fx0, fx1, fy0, fy1
        self.z_level = z
        return self


def fine_gradient(base: TerrainBase, edit: TerrainEdit | None, dJ_dz_cell, dJ_dz_face_x=None, dJ_dz_face_y=None):
    """d J / d(fine z) on the fine raster from the gradients the adjoint reports on a level.

    Level 1 (signed faces): the level is the fine raster, so this is dJ_dz_cell.  A
    face-crest level (terrain.rs:357-429): a level z is the mean of its block's open fine
    cells, so dJ_dz_cell (crest lines held) spreads evenly over them; a face's crest line is
    the maximum of the open fine cells either side of the face, so its gradient goes to the
    cell that holds that maximum (the first in the reference's loop order, rows then the two
    columns, on a tie).  ``edit``: the candidate's edit on the fine raster (the maxima are
    taken over the edited terrain)."""
    fine, lv = base.fine, base.level
    if not base.crest:
        return np.asarray(dJ_dz_cell, np.float64).copy()
    nx, ny = base.nx, base.ny
    zfine = fine.z + (edit.dense(fine.nx, fine.ny) if edit is not None and not edit.is_empty() else 0.0)
    cnt = np.zeros((ny, nx), np.int64)
    for dr in range(lv):
        for dc in range(lv):
            rr = np.arange(ny) * lv + dr
            cc = np.arange(nx) * lv + dc
            valid = (rr < fine.ny)[:, None] & (cc < fine.nx)[None, :]
            ww = fine.wall[np.ix_(np.minimum(rr, fine.ny - 1), np.minimum(cc, fine.nx - 1))]
            cnt += valid & ~ww
    g = np.zeros((fine.ny, fine.nx))
    fr, fc = np.arange(fine.ny) // lv, np.arange(fine.nx) // lv
    share = np.where(cnt > 0, np.asarray(dJ_dz_cell, np.float64) / np.maximum(cnt, 1), 0.0)
    g += np.where(fine.wall, 0.0, share[np.ix_(fr, fc)])
    gx = np.zeros((ny, nx + 1)) if dJ_dz_face_x is None else np.asarray(dJ_dz_face_x, np.float64)
    gy = np.zeros((ny + 1, nx)) if dJ_dz_face_y is None else np.asarray(dJ_dz_face_y, np.float64)
    for r, i in zip(*np.nonzero(gx)):
        if not 1 <= i <= nx - 1:
            continue
        best, at = -np.inf, None
        for rr in range(r * lv, min((r + 1) * lv, fine.ny)):
            for cc in (i * lv - 1, min(i * lv, fine.nx - 1)):
                if not fine.wall[rr, cc] and zfine[rr, cc] > best:
                    best, at = zfine[rr, cc], (rr, cc)
        if at is not None:
            g[at] += gx[r, i]
    for f, c in zip(*np.nonzero(gy)):
        if not 1 <= f <= ny - 1:
            continue
        best, at = -np.inf, None
        for cc in range(c * lv, min((c + 1) * lv, fine.nx)):
            for rr in (f * lv - 1, min(f * lv, fine.ny - 1)):
                if not fine.wall[rr, cc] and zfine[rr, cc] > best:
                    best, at = zfine[rr, cc], (rr, cc)
        if at is not None:
            g[at] += gy[f, c]
    return g


# ---------------------------------------------------------------- design.rs:95-385
PRIM_KINDS = ("berm", "swale", "cut", "mound")


@dataclass
class Primitive:
    kind: str
    cx: float
    cy: float
    length: float
    width: float
    angle_deg: float
    height: float

    def segment(self):
        """design.rs:97-105."""
        if self.kind == "mound":
            return (self.cx, self.cy), (self.cx, self.cy)
        a = self.angle_deg * (math.pi / 180.0)
        s, c = math.sin(a), math.cos(a)
        hl = 0.5 * self.length
        return (self.cx - c * hl, self.cy - s * hl), (self.cx + c * hl, self.cy + s * hl)

    def profile(self, x, y):
        """design.rs:108-134, vectorised over x, y arrays."""
        (x0, y0), (x1, y1) = self.segment()
        vx, vy = x1 - x0, y1 - y0
        l2 = vx * vx + vy * vy
        if l2 > 0.0:
            t = np.clip(((x - x0) * vx + (y - y0) * vy) / l2, 0.0, 1.0)
        else:
            t = np.zeros_like(x)
        px, py = x0 + t * vx - x, y0 + t * vy - y
        d = np.sqrt(px * px + py * py)
        s = d / (0.5 * self.width)
        if self.kind == "cut":
            u = (s - 0.5) / 0.5
            w = 1.0 - u * u
            v = np.where(s <= 0.5, 1.0, w * w)
        else:
            w = 1.0 - s * s
            v = w * w
        return np.where(s >= 1.0, 0.0, v)

    def bbox(self):
        (x0, y0), (x1, y1) = self.segment()
        r = 0.5 * self.width
        return min(x0, x1) - r, min(y0, y1) - r, max(x0, x1) + r, max(y0, y1) + r


def prim_window(p, geo: GeoTransform, nx, ny):
    """Cells (c0, r0, c1, r1), half-open, that can hold the primitive's support
    (design.rs:322-331): its world bbox snapped outward to whole cells, inside the grid."""
    x0, y0, x1, y1 = p.bbox()
    fc0, fr1 = geo.to_grid(x0, y0)
    fc1, fr0 = geo.to_grid(x1, y1)
    a = min(int(max(math.floor(fc0), 0.0)), nx)
    b = min(int(max(math.floor(fr0), 0.0)), ny)
    c = min(int(max(math.ceil(fc1), 0.0)), nx)
    d = min(int(max(math.ceil(fr1), 0.0)), ny)
    return a, b, c, d


def design_rows(prims, geo: GeoTransform, nx, ny):
    """One candidate's primitives as the device materializes them: rows of (profile kind,
    segment x0 y0 x1 y1, width, height, cell window c0 r0 c1 r1, 0), and the candidate's
    union window (c0, r0, w, h) of materialize's first loop (0 0 0 0: no edit)."""
    rows = []
    c0 = r0 = 1 << 62
    c1 = r1 = 0
    for p in prims:
        (x0, y0), (x1, y1) = p.segment()
        a, b, c, d = prim_window(p, geo, nx, ny)
        rows.append([0.0 if p.kind == "cut" else 1.0, x0, y0, x1, y1, p.width, p.height, a, b, c, d, 0.0])
        if p.height != 0.0 and a < c and b < d:
            c0, r0, c1, r1 = min(c0, a), min(r0, b), max(c1, c), max(r1, d)
    box = [c0, r0, c1 - c0, r1 - r0] if (c0 < c1 and r0 < r1) else [0, 0, 0, 0]
    return rows, box


def materialize(prims, geo: GeoTransform, nx, ny, editable=None, max_abs_dz=1e9) -> TerrainEdit:
    """DesignSpace::materialize (design.rs:318-385): 4x4 sub-sample quadrature per cell,
    summed over primitives in order, masked, clamped; cut/fill volumes."""
    c0 = r0 = 1 << 62
    c1 = r1 = 0
    for p in prims:
        if p.height == 0.0:
            continue
        a, b, c, d = prim_window(p, geo, nx, ny)
        if a < c and b < d:
            c0, r0, c1, r1 = min(c0, a), min(r0, b), max(c1, c), max(r1, d)
    if c0 >= c1 or r0 >= r1:
        return TerrainEdit()
    w, h = c1 - c0, r1 - r0
    S = 4
    inv = 1.0 / (S * S)
    acc = np.zeros((h, w))
    for p in prims:
        if p.height == 0.0:
            continue
        # The profile is exactly 0 outside the primitive's own bbox, and adding
        # height * 0 * inv (= +-0) leaves acc unchanged, so only that window is summed.
        wa, wb, wc, wd = prim_window(p, geo, nx, ny)
        pa, pb, pc, pd = max(wa, c0), max(wb, r0), min(wc, c1), min(wd, r1)
        if pa >= pc or pb >= pd:
            continue
        cols = np.arange(pa, pc, dtype=np.float64)[None, :]
        rows = np.arange(pb, pd, dtype=np.float64)[:, None]
        hh, ww = pd - pb, pc - pa
        s = np.zeros((hh, ww))
        for sy in range(S):
            for sx in range(S):
                x = geo.origin_x + (cols + (sx + 0.5) / S) * geo.dx
                y = geo.origin_y - (rows + (sy + 0.5) / S) * geo.dy
                s = s + p.profile(np.broadcast_to(x, (hh, ww)), np.broadcast_to(y, (hh, ww)))
        win = acc[pb - r0:pd - r0, pa - c0:pc - c0]
        acc[pb - r0:pd - r0, pa - c0:pc - c0] = win + p.height * s * inv
    dz = np.clip(acc, -max_abs_dz, max_abs_dz)
    if editable is not None:
        dz = np.where(np.asarray(editable, bool)[r0:r1, c0:c1], dz, 0.0)
    a = geo.cell_area()
    fill = 0.0
    cut = 0.0
    for v in dz.ravel():
        if v > 0.0:
            fill += v * a
        else:
            cut -= v * a
    return TerrainEdit(c0, r0, dz, cut, fill)

def materialize_jacobian(prims, geo: GeoTransform, nx, ny, editable=None, max_abs_dz=1e9):
    """d dz / d height of every primitive, dense (ny, nx) per primitive: the primitive's
    4x4-quadrature mean profile (design.rs:356-371 is linear in each height), zero where
    the sum is clamped to +-max_abs_dz or the cell is not editable."""
    total = materialize(prims, geo, nx, ny, None, 1e300).dense(nx, ny)
    free = np.abs(total) < max_abs_dz
    if editable is not None:
        free &= np.asarray(editable, bool)
    S = 4
    out = []
    for p in prims:
        J = np.zeros((ny, nx))
        if p.height != 0.0:
            x0, y0, x1, y1 = p.bbox()
            fc0, fr1 = geo.to_grid(x0, y0)
            fc1, fr0 = geo.to_grid(x1, y1)
            pa, pb = min(int(max(math.floor(fc0), 0.0)), nx), min(int(max(math.floor(fr0), 0.0)), ny)
            pc, pd = min(int(max(math.ceil(fc1), 0.0)), nx), min(int(max(math.ceil(fr1), 0.0)), ny)
            if pa < pc and pb < pd:
                cols = np.arange(pa, pc, dtype=np.float64)[None, :]
                rows = np.arange(pb, pd, dtype=np.float64)[:, None]
                hh, ww = pd - pb, pc - pa
                s = np.zeros((hh, ww))
                for sy in range(S):
                    for sx in range(S):
                        x = geo.origin_x + (cols + (sx + 0.5) / S) * geo.dx
                        y = geo.origin_y - (rows + (sy + 0.5) / S) * geo.dy
                        s = s + p.profile(np.broadcast_to(x, (hh, ww)), np.broadcast_to(y, (hh, ww)))
                J[pb:pd, pa:pc] = s / (S * S)
        out.append(np.where(free, J, 0.0))
    return out
