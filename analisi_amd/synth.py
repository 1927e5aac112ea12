"""Seeded synthetic inputs for the g(r,t) / MSD / Green-Kubo workloads.

Counter-based so that any machine with this numpy produces identical arrays:
every random number is SplitMix64(seed, stream, index) (Steele, Lea & Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014), turned into a
uniform double on (0, 1] from its top 53 bits and into normals with the
Box-Muller transform.  Trajectories are random walks: frame 0 uniform in the
cell, every later frame adds an i.i.d. Gaussian step per axis, so positions are
unwrapped (the min-image loops of the pair path do real work).
"""
import hashlib

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, stream, idx):
    """SplitMix64 output for counter ``idx`` (uint64 array) of ``stream``."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(0x632BE59BD9B4E019) + np.uint64(stream) * np.uint64(0x8CB92BA72F3D8DD7)
        z = base + (idx.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n, offset=0):
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    return ((splitmix64(seed, stream, idx) >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)


def normal(seed, stream, n, offset=0):
    """n standard normals (Box-Muller on consecutive uniform pairs)."""
    m = (n + 1) // 2
    u = uniform(seed, stream, 2 * m, 2 * offset)
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    th = 2.0 * np.pi * u[1::2]
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(th)
    z[1::2] = r * np.sin(th)
    return z[:n]


def lammps_box(L, tilt=None):
    """LammpsOrtho [xlo,xhi,ylo,yhi,zlo,zhi] or LammpsTriclinic (+ xy,xz,yz)."""
    L = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))
    b = [0.0, L[0], 0.0, L[1], 0.0, L[2]]
    if tilt is not None:
        b += list(np.asarray(tilt, dtype=np.float64))
    return np.array(b)


def types_by_fraction(natoms, fractions):
    """Contiguous type blocks: counts = floor(f*N), remainder to the last."""
    counts = [int(f * natoms) for f in fractions]
    counts[-1] = natoms - sum(counts[:-1])
    return np.repeat(np.arange(len(fractions), dtype=np.int32), counts)


def random_walk(seed, natoms, nframes, box, sigma, dtype=np.float64, progress=None):
    """Positions [nframes, natoms, 3]: frame 0 uniform in the (possibly tilted)
    cell, each later frame adds N(0, sigma^2) per axis.  ``progress(f)`` is
    called after every 1,000th frame (long trajectories take minutes)."""
    box = np.asarray(box, dtype=np.float64)
    lo = box[[0, 2, 4]]
    L = box[[1, 3, 5]] - lo
    xy, xz, yz = (box[6], box[7], box[8]) if box.size == 9 else (0.0, 0.0, 0.0)
    s = uniform(seed, 1, natoms * 3).reshape(natoms, 3)
    x0 = np.empty((natoms, 3))
    x0[:, 0] = lo[0] + s[:, 0] * L[0] + s[:, 1] * xy + s[:, 2] * xz
    x0[:, 1] = lo[1] + s[:, 1] * L[1] + s[:, 2] * yz
    x0[:, 2] = lo[2] + s[:, 2] * L[2]
    pos = np.empty((nframes, natoms, 3), dtype=np.float64)
    pos[0] = x0
    # the steps of a chunk of frames are generated on a few threads (numpy
    # releases the GIL; the values depend only on their counters), then added
    # frame after frame: pos[f] = pos[f-1] + sigma * step[f-1], in order
    from concurrent.futures import ThreadPoolExecutor
    chunk = max(1, (1 << 22) // max(natoms * 3, 1))
    per = natoms * 3

    def steps(f0, f1):
        # frame f's step is normal(seed, 2, per, offset=(f - 1) * per): its
        # uniforms start at counter 2 (f - 1) per
        m = (per + 1) // 2
        idx = (2 * per * np.arange(f0 - 1, f1 - 1, dtype=np.uint64))[:, None] + np.arange(2 * m, dtype=np.uint64)
        u = ((splitmix64(seed, 2, idx) >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
        r = np.sqrt(-2.0 * np.log(u[:, 0::2]))
        th = 2.0 * np.pi * u[:, 1::2]
        z = np.empty((f1 - f0, 2 * m))
        z[:, 0::2] = r * np.cos(th)
        z[:, 1::2] = r * np.sin(th)
        return sigma * z[:, :per].reshape(f1 - f0, natoms, 3)

    bounds = [(f0, min(f0 + chunk, nframes)) for f0 in range(1, nframes, chunk)]
    done = 0
    with ThreadPoolExecutor(max_workers=8) as ex:
        for k in range(0, len(bounds), 8):
            group = bounds[k:k + 8]
            for (f0, f1), st in zip(group, ex.map(lambda b: steps(*b), group)):
                for i in range(f1 - f0):
                    pos[f0 + i] = pos[f0 + i - 1] + st[i]
                if progress is not None and f1 // 1000 > done // 1000:
                    progress(f1 - 1)
                done = f1
    return pos.astype(dtype, copy=False)


def ar2_currents(seed, nsteps, ncur=2, a1=1.6, a2=-0.7, coupling=0.3):
    """ncur vector currents [nsteps, 3*ncur] from seeded AR(2) processes
    x_t = a1 x_{t-1} + a2 x_{t-2} + e_t; current k>0 is
    coupling*current0 + sqrt(1-coupling^2)*independent AR(2)."""
    from scipy.signal import lfilter

    def ar2(stream):
        e = normal(seed, stream, nsteps * 3).reshape(nsteps, 3)
        return lfilter([1.0], [1.0, -a1, -a2], e, axis=0)

    base = ar2(10)
    cols = [base]
    for k in range(1, ncur):
        cols.append(coupling * base + np.sqrt(1.0 - coupling ** 2) * ar2(10 + k))
    return np.ascontiguousarray(np.concatenate(cols, axis=1))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]
