"""Seeded synthetic REPEAT workloads (cube + parameter files) for tests and benchmarks.

Every workload is a periodic cell of point charges whose ESP on the cube grid
is the exact Ewald potential (splitting eta = 0.3 Bohr^-1, real cutoff 20 Bohr,
reciprocal sphere |k| <= 2.5 Bohr^-1, k = 0 dropped) plus seeded N(0, 1e-5) a.u.
noise.  The real-space part is summed in C (``csrc/synth_esp.c``), the
reciprocal part by placing the weighted structure factor on the FFT grid, which
is exact because k . g = 2 pi (h1 i/n1 + h2 j/n2 + h3 k/n3).

Workloads (random streams ``numpy.random.default_rng([cfg, req])``):

- ``water``      one water molecule, cubic 12 Bohr, 48^3, fp64
- ``crystal200`` 200 atoms, triclinic 24/22/26 Bohr, 85/95/100 deg, 96x96x100, fp64,
                 20 symmetry groups of 2-4 same-element atoms
- ``mof1000``    1000 atoms, cubic 40 Bohr, 160^3, fp32, RESP ``H 0.1 0.0``
- ``stream``     request ``r`` of the Poisson serving stream (20-800 atoms, 0.3 Bohr spacing, fp32);
                 its ESP uses the splitting eta = 0.6 Bohr^-1, real cutoff 8 Bohr, |k| <= 5 Bohr^-1, which
                 sums the same potential with the same truncation (erfc(eta Rc) and exp(-k^2/4 eta^2) at
                 the cutoffs are as small or smaller) at 1/16 of the real-space work
- ``large5000``  5000 atoms, cubic 70 Bohr, 232^3, fp32
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SYNTH_SO = os.path.join(_HERE, "libsynth_esp.so")
_synth = None

ETA = 0.3
RCUT = 20.0
KCUT = 2.5
NOISE = 1e-5

Z_OF = {"H": 1, "C": 6, "N": 7, "O": 8, "Si": 14, "Zn": 30}


def build_helper(force=False):
    src = os.path.join(_HERE, "csrc", "synth_esp.c")
    if force or not os.path.exists(_SYNTH_SO) or os.path.getmtime(_SYNTH_SO) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-fopenmp", "-shared", "-o", _SYNTH_SO, src, "-lm"])
    return _SYNTH_SO


def _lib():
    global _synth
    if _synth is None:
        if not os.path.exists(_SYNTH_SO):
            build_helper()
        L = ctypes.CDLL(_SYNTH_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        L.synth_esp_real.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int64, dp, dp, dp, dp,
                                     ctypes.c_double, ctypes.c_double, dp, ctypes.c_int]
        L.synth_esp_real.restype = None
        _synth = L
    return _synth


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class System:
    """One charge-fit input, in the layout the cube reader produces."""
    name: str
    n_grid: np.ndarray            # (3,) int
    axis: np.ndarray              # (3,3) grid step vectors (rows)
    atomic_numbers: np.ndarray    # (N,) int
    atom_pos: np.ndarray          # (N,3) Bohr
    true_charges: np.ndarray      # (N,)
    esp_cube: np.ndarray          # (nx*ny*nz,) cube file order: x outer, z inner; float32 or float64
    symm_groups: list = field(default_factory=list)   # list of lists of 1-based atom ids
    resp_text: str | None = None
    q_tot: float = 0.0

    @property
    def box(self):
        return self.axis * self.n_grid[:, None]

    @property
    def n_atoms(self):
        return len(self.atomic_numbers)

    def esp_flat(self):
        """ESP in flat order ``i + j*nx + k*nx*ny``."""
        nx, ny, nz = (int(v) for v in self.n_grid)
        return self.esp_cube.reshape(nx, ny, nz).transpose(2, 1, 0).ravel()

    def symm_text(self):
        return "".join(",".join(str(i) for i in g) + "\n" for g in self.symm_groups)

    def write_cube(self, path):
        """Gaussian cube text: origin 0, atom lines ``Z 0.0 x y z``, values %.10E, 6 per line."""
        from .cube import cube_from_system, write_gauss_cube
        write_gauss_cube(path, cube_from_system(self), comment=(f"{self.name} synthetic ESP",
                                                               "exact Ewald potential plus noise"))

# --------------------------------------------------------------------------- geometry helpers

def triclinic_box(a, b, c, alpha_deg, beta_deg, gamma_deg):
    """Cell rows: a along x, b in the xy plane."""
    al, be, ga = (math.radians(v) for v in (alpha_deg, beta_deg, gamma_deg))
    ax = np.array([a, 0.0, 0.0])
    bx = np.array([b * math.cos(ga), b * math.sin(ga), 0.0])
    cx = c * math.cos(be)
    cy = c * (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    cz = math.sqrt(max(c * c - cx * cx - cy * cy, 0.0))
    return np.array([ax, bx, [cx, cy, cz]])


def place_atoms(rng, box, n, dmin, max_tries=2000000):
    """Random sequential addition with a minimum-image separation check."""
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=np.float64)
    frac = np.zeros((n, 3))
    count = 0
    tries = 0
    d2min = dmin * dmin
    while count < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms")
        f = rng.random(3)
        if count:
            df = frac[:count] - f
            df -= np.round(df)
            dd = (df[:, None, :] + shifts[None]) @ box       # every neighbour of the wrapped image
            if np.min(np.einsum("ijk,ijk->ij", dd, dd)) < d2min:
                continue
        frac[count] = f
        count += 1
    return frac @ box


def element_list(rng, n, mix):
    """Counts by rounded percentages, remainder to the largest share, shuffled."""
    names = list(mix)
    counts = [int(round(n * mix[e])) for e in names]
    diff = n - sum(counts)
    counts[int(np.argmax([mix[e] for e in names]))] += diff
    out = []
    for e, c in zip(names, counts):
        out += [e] * c
    out = np.array(out)
    rng.shuffle(out)
    return out


# --------------------------------------------------------------------------- exact Ewald ESP

def ewald_esp(n_grid, axis, atom_pos, charges, eta=ETA, rcut=RCUT, kcut=KCUT, nthreads=None):
    """Exact Ewald potential of the point charges on the grid, flat order ``i + j*nx + k*nx*ny``."""
    n_grid = np.asarray(n_grid, dtype=np.int64)
    nx, ny, nz = (int(v) for v in n_grid)
    axis = np.ascontiguousarray(axis, dtype=np.float64)
    box = axis * n_grid[:, None]
    volume = abs(float(np.dot(box[0], np.cross(box[1], box[2]))))
    recip = 2 * np.pi * np.linalg.inv(box).T          # rows b_i with b_i . box_j = 2 pi delta_ij
    frac = np.asarray(atom_pos) @ np.linalg.inv(box)
    q = np.asarray(charges, dtype=np.float64)

    # real space over images whose cutoff sphere reaches the cell
    dfrac = rcut * np.linalg.norm(recip, axis=1) / (2 * np.pi)
    nimg = np.ceil(dfrac).astype(int) + 1
    img_xyz, img_frac, img_q = [], [], []
    for s1 in range(-nimg[0], nimg[0] + 1):
        for s2 in range(-nimg[1], nimg[1] + 1):
            for s3 in range(-nimg[2], nimg[2] + 1):
                fs = frac + np.array([s1, s2, s3])
                ok = np.all((fs + dfrac >= 0.0) & (fs - dfrac <= 1.0), axis=1)
                if np.any(ok):
                    img_frac.append(fs[ok])
                    img_xyz.append(fs[ok] @ box)
                    img_q.append(q[ok])
    img_xyz = np.ascontiguousarray(np.concatenate(img_xyz))
    img_frac = np.ascontiguousarray(np.concatenate(img_frac))
    img_q = np.ascontiguousarray(np.concatenate(img_q))
    out = np.zeros(nx * ny * nz, dtype=np.float64)
    dfr = np.ascontiguousarray(dfrac, dtype=np.float64)
    _lib().synth_esp_real(nx, ny, nz, _dp(axis), len(img_q), _dp(img_xyz), _dp(img_frac), _dp(img_q), _dp(dfr),
                          float(eta), float(rcut), _dp(out), int(nthreads or len(os.sched_getaffinity(0))))

    # reciprocal space: structure factor on the FFT grid
    H = np.floor(kcut * np.linalg.norm(box, axis=1) / (2 * np.pi)).astype(int)
    if np.any(2 * H + 1 > n_grid):
        raise ValueError("grid too coarse for the reciprocal cutoff")
    h1 = np.arange(-H[0], H[0] + 1)
    h2 = np.arange(-H[1], H[1] + 1)
    h3 = np.arange(-H[2], H[2] + 1)
    E1 = np.exp(-2j * np.pi * frac[:, 0:1] * h1[None, :])
    E2 = np.exp(-2j * np.pi * frac[:, 1:2] * h2[None, :])
    E3 = np.exp(-2j * np.pi * frac[:, 2:3] * h3[None, :])
    S = np.empty((len(h1), len(h2), len(h3)), dtype=np.complex128)
    for a, _ in enumerate(h1):
        S[a] = ((q * E1[:, a])[:, None] * E2).T @ E3
    hh = np.stack(np.meshgrid(h1, h2, h3, indexing="ij"), axis=-1).astype(np.float64)
    kvec = hh @ recip
    k2 = np.einsum("...i,...i->...", kvec, kvec)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = (4 * np.pi / volume) * np.exp(-k2 / (4 * eta * eta)) / k2
    w[(k2 == 0) | (k2 > kcut * kcut)] = 0.0
    X = np.zeros((nx, ny, nz), dtype=np.complex128)
    X[np.ix_(h1 % nx, h2 % ny, h3 % nz)] = w * S
    rec = np.fft.ifftn(X).real * (nx * ny * nz)     # indices [i, j, k]
    out += rec.transpose(2, 1, 0).ravel()
    out -= np.pi / (eta * eta * volume) * q.sum()
    return out


STREAM_SPLIT = {"eta": 0.6, "rcut": 8.0, "kcut": 5.0}


def _finish(name, rng, n_grid, axis, Z, pos, q, dtype, split=None, nthreads=None, **kw):
    flat = ewald_esp(n_grid, axis, pos, q, nthreads=nthreads, **(split or {}))
    flat = flat + rng.normal(0.0, NOISE, size=flat.shape)
    nx, ny, nz = (int(v) for v in n_grid)
    cube = flat.reshape(nz, ny, nx).transpose(2, 1, 0).ravel().astype(dtype)
    return System(name=name, n_grid=np.asarray(n_grid, dtype=np.int64), axis=np.asarray(axis, dtype=np.float64),
                  atomic_numbers=np.asarray(Z, dtype=np.int64), atom_pos=np.asarray(pos, dtype=np.float64),
                  true_charges=np.asarray(q, dtype=np.float64), esp_cube=np.ascontiguousarray(cube), **kw)


def _symmetry_groups(rng, elements, n_groups, min_size=2, max_size=4):
    """Disjoint groups of same-element atoms, 1-based ids, first id is the base."""
    used = set()
    groups = []
    by_el = {}
    for i, e in enumerate(elements):
        by_el.setdefault(e, []).append(i)
    tries = 0
    while len(groups) < n_groups and tries < 100 * n_groups:
        tries += 1
        e = elements[rng.integers(len(elements))]
        free = [i for i in by_el[e] if i not in used]
        size = int(rng.integers(min_size, max_size + 1))
        if len(free) < size:
            continue
        pick = [free[i] for i in rng.choice(len(free), size=size, replace=False)]
        used.update(pick)
        groups.append([p + 1 for p in pick])
    return groups


def water():
    rng = np.random.default_rng([1, 0])
    L, n = 12.0, 48
    o = np.array([6.0, 6.0, 6.0])
    pos = np.array([o, o + [1.43, 1.11, 0.0], o + [-1.43, 1.11, 0.0]])
    axis = np.eye(3) * (L / n)
    return _finish("water", rng, [n, n, n], axis, [8, 1, 1], pos, [-0.82, 0.41, 0.41], np.float64)


def crystal200():
    rng = np.random.default_rng([2, 0])
    box = triclinic_box(24.0, 22.0, 26.0, 85.0, 95.0, 100.0)
    n_grid = np.array([96, 96, 100])
    el = element_list(rng, 200, {"C": 0.4, "H": 0.4, "O": 0.1, "N": 0.1})
    pos = place_atoms(rng, box, 200, 2.0)
    q = rng.uniform(-0.8, 0.8, 200)
    q -= q.mean()
    groups = _symmetry_groups(rng, list(el), 20)
    return _finish("crystal200", rng, n_grid, box / n_grid[:, None], [Z_OF[e] for e in el], pos, q, np.float64,
                   symm_groups=groups)


def mof1000():
    rng = np.random.default_rng([3, 0])
    box = np.eye(3) * 40.0
    n_grid = np.array([160, 160, 160])
    el = element_list(rng, 1000, {"Zn": 0.05, "O": 0.30, "C": 0.40, "H": 0.20, "N": 0.05})
    pos = place_atoms(rng, box, 1000, 1.9)
    q = np.where(el == "Zn", rng.uniform(1.0, 1.6, 1000), rng.uniform(-0.9, 0.5, 1000))
    q -= q.mean()
    return _finish("mof1000", rng, n_grid, box / n_grid[:, None], [Z_OF[e] for e in el], pos, q, np.float32,
                   resp_text="H 0.1 0.0\n")


def large5000():
    rng = np.random.default_rng([5, 0])
    box = np.eye(3) * 70.0
    n_grid = np.array([232, 232, 232])
    el = element_list(rng, 5000, {"C": 0.45, "H": 0.35, "O": 0.12, "N": 0.05, "Si": 0.03})
    pos = place_atoms(rng, box, 5000, 1.9)
    q = rng.uniform(-1.0, 1.0, 5000)
    q -= q.mean()
    return _finish("large5000", rng, n_grid, box / n_grid[:, None], [Z_OF[e] for e in el], pos, q, np.float32)


def stream_schedule(n_requests=512, rate=20.0):
    """Arrival times (s) and atom counts of the serving stream."""
    rng = np.random.default_rng([4, 0])
    gaps = rng.exponential(1.0 / rate, n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    n_atoms = np.round(np.exp(rng.uniform(np.log(20), np.log(800), n_requests))).astype(int)
    return arrivals, n_atoms


def stream_request(r, n_atoms=None, nthreads=None):
    """Request ``r`` of the serving stream: cubic cell at 0.01 atoms/Bohr^3, 0.3 Bohr spacing, fp32."""
    if n_atoms is None:
        n_atoms = int(stream_schedule()[1][r])
    rng = np.random.default_rng([4, r + 1])
    L = (n_atoms / 0.01) ** (1.0 / 3.0)
    n = int(round(L / 0.3))
    box = np.eye(3) * L
    el = element_list(rng, n_atoms, {"C": 0.4, "H": 0.4, "O": 0.1, "N": 0.1})
    pos = place_atoms(rng, box, n_atoms, 2.0)
    q_tot = float(rng.integers(-1, 2))
    q = rng.uniform(-0.8, 0.8, n_atoms)
    q += (q_tot - q.sum()) / n_atoms
    groups = _symmetry_groups(rng, list(el), min(20, n_atoms // 10)) if rng.random() < 0.5 else []
    return _finish(f"stream{r}", rng, [n, n, n], box / n, [Z_OF[e] for e in el], pos, q, np.float32,
                   split=STREAM_SPLIT, nthreads=nthreads, symm_groups=groups, q_tot=q_tot)


def stream_requests(count=512, workers=4):
    """The first ``count`` requests of the serving stream, generated by ``workers`` threads."""
    from concurrent.futures import ThreadPoolExecutor
    per = max(1, len(os.sched_getaffinity(0)) // workers)
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda r: stream_request(r, nthreads=per), range(count)))


def tiny(seed=0, n_atoms=3, n=10, L=9.0, triclinic=False):
    """Small random cell for brute-force checks."""
    rng = np.random.default_rng([9, seed])
    box = triclinic_box(L, L * 1.05, L * 0.95, 80.0, 97.0, 104.0) if triclinic else np.eye(3) * L
    n_grid = np.array([n, n + 1, n + 2])
    el = element_list(rng, n_atoms, {"C": 0.5, "H": 0.3, "O": 0.2})
    pos = place_atoms(rng, box, n_atoms, 1.5)
    q = rng.uniform(-0.5, 0.5, n_atoms)
    q -= q.mean()
    return _finish(f"tiny{seed}", rng, n_grid, box / n_grid[:, None], [Z_OF[e] for e in el], pos, q, np.float64)


def cluster(n_atoms=400, sub=18.0, L=36.0, n=64, seed=1):
    """A compact cluster of ``n_atoms`` (C/H/O, min separation 1.6 Bohr) in the middle of an ``L`` Bohr
    cubic box: the grid points lie outside the cluster, so the buried atoms' basis columns are nearly
    dependent (the classic ill-conditioned ESP fit; cond(A) ~ 1e11-1e12 at 200-400 atoms)."""
    rng = np.random.default_rng([77, seed])
    pos = place_atoms(rng, np.eye(3) * sub, n_atoms, 1.6) + (L - sub) / 2
    el = element_list(rng, n_atoms, {"C": 0.5, "H": 0.3, "O": 0.2})
    q = rng.uniform(-0.5, 0.5, n_atoms)
    q -= q.mean()
    return _finish(f"cluster{n_atoms}", rng, np.array([n, n, n]), np.eye(3) * (L / n), [Z_OF[e] for e in el], pos, q,
                   np.float64)


WORKLOADS = {"water": water, "crystal200": crystal200, "mof1000": mof1000, "large5000": large5000}
