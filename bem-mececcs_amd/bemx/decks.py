"""Deck I/O and generated flat-patch problems.

Reads and writes the reference's input decks:
  * E deck (static): Inputece.for:34-105 -- title; ``N NBE NP NPG GE RNU``;
    NP lines ``x y z``; N lines of 4 node ids written as reals (truncated
    with INT); the first NBE elements are boundary elements, the rest are
    enclosing elements.
  * D deck (dynamic): Inputecd.for:26-134 -- title; ``L NFR DAM RHO``; NFR
    frequencies; NBE lines ``KODE_x BC_x KODE_y BC_y KODE_z BC_z``; L lines of
    interior point coordinates.
Every READ statement of the reference is a Fortran list-directed read: it
starts on a new record, may span records, and drops what is left of its last
record.  :class:`ListDirectedReader` follows that rule (blank lines and CR LF
line ends are accepted, ``D`` exponents and ``r*value`` repeats too).

The generator builds the flat free-surface patch problems: a uniform nx x ny
grid of 4-node quads on z = 0 with outward normal +z (domain z < 0), closed by
five enclosing elements (four sloped sides and a bottom) so that the
rigid-body diagonal of the static H matrix is well defined.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

__all__ = [
    "Mesh", "Dynamic", "ListDirectedReader", "read_e_deck", "read_d_deck",
    "write_e_deck", "write_d_deck", "flat_patch", "flat_patch_problem",
    "config_problem", "CONFIGS",
]


@dataclasses.dataclass
class Mesh:
    """Static mesh data (the outputs of INPUTECE)."""
    title: str
    n: int          # elements, boundary + enclosing
    nbe: int        # boundary elements
    npg: int        # Gauss points per direction
    ge: float       # shear modulus G (RMU = GE, Inputece.for:49)
    rnu: float      # Poisson ratio
    cx: np.ndarray  # (NP,) float64
    cy: np.ndarray
    cz: np.ndarray
    cone: np.ndarray  # (N, 4) int32, 1-based node ids

    @property
    def np_(self) -> int:
        return int(self.cx.shape[0])

    def cone_f(self) -> np.ndarray:
        """Connectivity in the reference's column-major CONE(N,4) layout."""
        return np.asfortranarray(self.cone.astype(np.int32))

    def centroids(self):
        """Collocation points (Inputece.for:100-102)."""
        c = self.cone - 1
        cxm = (self.cx[c[:, 0]] + self.cx[c[:, 1]] + self.cx[c[:, 2]] + self.cx[c[:, 3]]) * 0.25
        cym = (self.cy[c[:, 0]] + self.cy[c[:, 1]] + self.cy[c[:, 2]] + self.cy[c[:, 3]]) * 0.25
        czm = (self.cz[c[:, 0]] + self.cz[c[:, 1]] + self.cz[c[:, 2]] + self.cz[c[:, 3]]) * 0.25
        return cxm, cym, czm


@dataclasses.dataclass
class Dynamic:
    """Dynamic problem data (the outputs of INPUTECD)."""
    title: str
    l: int
    dam: float
    rho: float
    freqs: np.ndarray   # (NFR,) float64
    kode: np.ndarray    # (3 NBE,) int32; 1 = traction prescribed, 0 = displacement
    bc: np.ndarray      # (3 NBE,) float64 prescribed values
    cxi: np.ndarray     # (L,) interior points
    cyi: np.ndarray
    czi: np.ndarray
    load_cases: Optional[np.ndarray] = None  # (nrhs, 3 NBE) extra BC columns

    @property
    def nfr(self) -> int:
        return int(self.freqs.shape[0])


class ListDirectedReader:
    """Fortran list-directed READ semantics over a text deck."""

    def __init__(self, text: str):
        text = text.replace("\r\n", "\n").replace("\r", "\n")
        self.lines: List[str] = text.split("\n")
        self.pos = 0

    def read_line(self) -> str:
        """A formatted ``READ(u,'(A)')`` of one record."""
        if self.pos >= len(self.lines):
            raise EOFError("deck ended early")
        line = self.lines[self.pos]
        self.pos += 1
        return line

    @staticmethod
    def _tokens(line: str) -> List[str]:
        out: List[str] = []
        for tok in line.replace(",", " ").replace("\t", " ").split():
            if "*" in tok:
                cnt, val = tok.split("*", 1)
                out.extend([val] * int(cnt))
            else:
                out.append(tok)
        return out

    def read(self, count: int) -> List[str]:
        """One list-directed READ of ``count`` items."""
        vals: List[str] = []
        while len(vals) < count:
            if self.pos >= len(self.lines):
                raise EOFError("deck ended while reading %d values" % count)
            vals.extend(self._tokens(self.lines[self.pos]))
            self.pos += 1
        return vals[:count]


def _f(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def _i(tok: str) -> int:
    return int(_f(tok)) if any(c in tok for c in ".eEdD") else int(tok)


def read_e_deck(path_or_text: str, is_text: bool = False) -> Mesh:
    """INPUTECE (Inputece.for:16-108)."""
    text = path_or_text if is_text else open(path_or_text, "rb").read().decode("latin-1")
    rd = ListDirectedReader(text)
    title = rd.read_line()[:75]
    h = rd.read(6)
    n, nbe, npts, npg = (_i(t) for t in h[:4])
    ge, rnu = _f(h[4]), _f(h[5])
    xyz = np.array([_f(t) for t in rd.read(3 * npts)], dtype=np.float64).reshape(npts, 3)
    cone = np.empty((n, 4), dtype=np.int32)
    for e in range(n):
        # CONE(I,1:4) = INT(cone_in(1:4)): truncation toward zero
        cone[e] = [int(math.trunc(_f(t))) for t in rd.read(4)]
    if cone.min() < 1 or cone.max() > npts:
        raise ValueError("connectivity refers to a node outside 1..NP")
    return Mesh(title, n, nbe, npg, ge, rnu, xyz[:, 0].copy(), xyz[:, 1].copy(),
                xyz[:, 2].copy(), cone)


def read_d_deck(path_or_text: str, nbe: int, is_text: bool = False) -> Dynamic:
    """INPUTECD (Inputecd.for:7-139)."""
    text = path_or_text if is_text else open(path_or_text, "rb").read().decode("latin-1")
    rd = ListDirectedReader(text)
    title = rd.read_line()[:75]
    h = rd.read(4)
    l, nfr = _i(h[0]), _i(h[1])
    dam, rho = _f(h[2]), _f(h[3])
    freqs = np.array([_f(t) for t in rd.read(nfr)], dtype=np.float64)
    kode = np.empty(3 * nbe, dtype=np.int32)
    bc = np.empty(3 * nbe, dtype=np.float64)
    for e in range(nbe):
        t = rd.read(6)
        for j in range(3):
            kode[3 * e + j] = _i(t[2 * j])
            bc[3 * e + j] = _f(t[2 * j + 1])
    if l > 0:
        p = np.array([_f(t) for t in rd.read(3 * l)], dtype=np.float64).reshape(l, 3)
    else:
        p = np.zeros((0, 3))
    return Dynamic(title, l, dam, rho, freqs, kode, bc, p[:, 0].copy(), p[:, 1].copy(),
                   p[:, 2].copy())


def write_e_deck(mesh: Mesh, path: str) -> None:
    with open(path, "w") as f:
        f.write(mesh.title[:75] + "\n\n")
        f.write("%d %d %d %d %r %r\n\n" % (mesh.n, mesh.nbe, mesh.np_, mesh.npg, mesh.ge, mesh.rnu))
        for x, y, z in zip(mesh.cx, mesh.cy, mesh.cz):
            f.write("%.17g %.17g %.17g\n" % (x, y, z))
        f.write("\n")
        for row in mesh.cone:
            f.write("%d. %d. %d. %d.\n" % tuple(int(v) for v in row))


def write_d_deck(dyn: Dynamic, path: str) -> None:
    with open(path, "w") as f:
        f.write(dyn.title[:75] + "\n\n")
        f.write("%d %d %r %r\n\n" % (dyn.l, dyn.nfr, dyn.dam, dyn.rho))
        f.write(" ".join("%.17g" % w for w in dyn.freqs) + "\n\n")
        for e in range(dyn.kode.shape[0] // 3):
            f.write("  ".join("%d %.17g" % (dyn.kode[3 * e + j], dyn.bc[3 * e + j])
                              for j in range(3)) + "\n")
        f.write("\n")
        for x, y, z in zip(dyn.cxi, dyn.cyi, dyn.czi):
            f.write("%.17g %.17g %.17g\n" % (x, y, z))


def flat_patch(nx: int, ny: int, half: float = 5.0, npg: int = 8, ge: float = 1.0,
               rnu: float = 0.25, bottom_scale: float = 1.1, depth_scale: float = 1.2,
               title: Optional[str] = None) -> Mesh:
    """Flat free-surface patch [-half, half]^2 at z = 0 plus 5 enclosing elements.

    Boundary element e = i + nx*j (i along x) has nodes
    (i,j), (i+1,j), (i+1,j+1), (i,j+1) so NORMVEC gives n = +z.  The enclosure
    is a frustum down to z = -depth_scale*half whose bottom square has
    half-width bottom_scale*half; all enclosing normals point outward.
    """
    xs = np.linspace(-half, half, nx + 1)
    ys = np.linspace(-half, half, ny + 1)
    gx, gy = np.meshgrid(xs, ys, indexing="xy")       # (ny+1, nx+1)
    cx = gx.ravel().astype(np.float64)
    cy = gy.ravel().astype(np.float64)
    cz = np.zeros_like(cx)

    def nid(i, j):
        return 1 + i + (nx + 1) * j

    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    cone = np.stack([nid(ii, jj), nid(ii + 1, jj), nid(ii + 1, jj + 1), nid(ii, jj + 1)],
                    axis=1).astype(np.int32)
    a = bottom_scale * half
    d = depth_scale * half
    base = cx.shape[0]
    bx = np.array([-a, a, a, -a])
    by = np.array([-a, -a, a, a])
    cx = np.concatenate([cx, bx])
    cy = np.concatenate([cy, by])
    cz = np.concatenate([cz, np.full(4, -d)])
    b00, b10, b11, b01 = base + 1, base + 2, base + 3, base + 4
    c00, c10, c11, c01 = nid(0, 0), nid(nx, 0), nid(nx, ny), nid(0, ny)
    ee = np.array([
        [c10, c00, b00, b10],   # south side, outward -y
        [c11, c10, b10, b11],   # east side, outward +x
        [c01, c11, b11, b01],   # north side, outward +y
        [c00, c01, b01, b00],   # west side, outward -x
        [b00, b01, b11, b10],   # bottom, outward -z
    ], dtype=np.int32)
    cone = np.concatenate([cone, ee], axis=0)
    nbe = nx * ny
    if title is None:
        title = "FLAT PATCH %dx%d [-%g,%g]^2 (%d BE + 5 EE)" % (nx, ny, half, half, nbe)
    return Mesh(title, cone.shape[0], nbe, npg, float(ge), float(rnu), cx, cy, cz, cone)


def central_elements(nx: int, ny: int, w: int = 2) -> List[int]:
    """The w x w block of elements at the centre of an nx x ny grid."""
    i0 = (nx - 1) // 2 - (w - 2) // 2
    j0 = (ny - 1) // 2 - (w - 2) // 2
    return [i + nx * j for j in range(j0, j0 + w) for i in range(i0, i0 + w)]


def random_load_cases(nx: int, ny: int, nbe: int, count: int, rng) -> np.ndarray:
    """``count`` random central traction patches, (count, 3 NBE): a w x w
    block of elements (w in 1..3) near the centre of the nx x ny patch
    loaded with unit traction along one axis."""
    cases = np.zeros((count, 3 * nbe), dtype=np.float64)
    for c in range(count):
        w = int(rng.integers(1, 4))
        i0 = int(rng.integers(max(0, nx // 2 - 4), min(nx - w, nx // 2 + 3) + 1))
        j0 = int(rng.integers(max(0, ny // 2 - 4), min(ny - w, ny // 2 + 3) + 1))
        comp = int(rng.integers(0, 3))
        for j in range(j0, j0 + w):
            for i in range(i0, i0 + w):
                cases[c, 3 * (i + nx * j) + comp] = 1.0
    return cases


def flat_patch_problem(nx: int, ny: int, half: float = 5.0, npg: int = 8,
                       ge: float = 1.0, rnu: float = 0.25, dam: float = 0.01,
                       rho: float = 1.0, freqs: Sequence[float] = (1.0,),
                       n_interior: int = 64, seed: int = 0,
                       loaded: Optional[Sequence[int]] = None,
                       load_cases: int = 0) -> tuple:
    """Mesh + dynamic data: unit vertical traction on the central 2x2 elements,
    zero traction elsewhere (all KODE = 1), interior points drawn uniformly
    from [-0.8h,0.8h]^2 x [-0.6h,-0.1h] with ``numpy.random.default_rng(seed)``.
    ``load_cases`` > 0 adds that many random central traction patches as extra
    right-hand sides (same matrix)."""
    mesh = flat_patch(nx, ny, half, npg, ge, rnu)
    nbe = mesh.nbe
    kode = np.ones(3 * nbe, dtype=np.int32)
    bc = np.zeros(3 * nbe, dtype=np.float64)
    for e in (loaded if loaded is not None else central_elements(nx, ny)):
        bc[3 * e + 2] = 1.0
    rng = np.random.default_rng(seed)
    h = half
    if n_interior > 0:
        pts = rng.uniform([-0.8 * h, -0.8 * h, -0.6 * h], [0.8 * h, 0.8 * h, -0.1 * h],
                          (n_interior, 3))
    else:
        pts = np.zeros((0, 3))
    cases = random_load_cases(nx, ny, nbe, load_cases, rng) if load_cases > 0 else None
    dyn = Dynamic("DYNAMIC DATA " + mesh.title, pts.shape[0], float(dam), float(rho),
                  np.asarray(freqs, dtype=np.float64), kode, bc, pts[:, 0].copy(),
                  pts[:, 1].copy(), pts[:, 2].copy(), cases)
    return mesh, dyn


# The benchmark workloads: (nx, ny, half, npg, frequencies, extra).
CONFIGS = {
    "c1": dict(nx=16, ny=15, half=5.0, npg=8, freqs=(1.0,)),
    "c2": dict(nx=48, ny=45, half=5.0, npg=8, freqs=(2.0,)),
    "c3": dict(nx=96, ny=90, half=10.0, npg=6, freqs=(0.5,), random_material=True),
    "c4": dict(nx=48, ny=45, half=5.0, npg=8, freqs="sweep256"),
}


def config_problem(name: str, seed: int = 0):
    """Mesh + dynamic data of a named workload (c1..c4)."""
    cfg = dict(CONFIGS[name])
    freqs = cfg.pop("freqs")
    if freqs == "sweep256":
        freqs = tuple(np.logspace(-1.0, 1.0, 256))
    rnu, ge = 0.25, 1.0
    if cfg.pop("random_material", False):
        rng = np.random.default_rng(1000 + seed)
        rnu = float(rng.uniform(0.2, 0.35))
        ge = float(rng.uniform(0.5, 2.0))
    return flat_patch_problem(cfg["nx"], cfg["ny"], cfg["half"], cfg["npg"], ge=ge,
                              rnu=rnu, freqs=freqs, seed=seed)


# Serving stream: mesh size classes of the request mix.
SERVE_CLASSES = ((16, 15), (24, 22), (32, 30), (48, 45), (64, 60))


def serving_stream(n_requests: int = 2000, rate: float = 20.0, seed: int = 0,
                   classes=SERVE_CLASSES):
    """The request mix of the serving workload: Poisson arrivals at ``rate``
    requests/s, a mesh class drawn uniformly from ``classes`` (flat patches,
    half-width 5, 8 x 8 Gauss), omega ~ U[0.1, 10], damping ~ U[0.005, 0.05]
    and 1..4 load cases (random central traction patches).  Returns
    (meshes, kodes, requests) with requests a list of dicts (t_arrival, cls,
    omega, dam, zdfi (3 NBE, nrhs))."""
    rng = np.random.default_rng(seed)
    meshes = [flat_patch(nx, ny) for (nx, ny) in classes]
    kodes = [np.ones(3 * m.nbe, dtype=np.int32) for m in meshes]
    t = 0.0
    reqs = []
    for _ in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        c = int(rng.integers(0, len(classes)))
        nx, ny = classes[c]
        omega = float(rng.uniform(0.1, 10.0))
        dam = float(rng.uniform(0.005, 0.05))
        nrhs = int(rng.integers(1, 5))
        cases = random_load_cases(nx, ny, meshes[c].nbe, nrhs, rng)
        reqs.append(dict(t_arrival=t, cls=c, omega=omega, dam=dam,
                         zdfi=np.asfortranarray(cases.T.astype(np.complex128))))
    return meshes, kodes, reqs
