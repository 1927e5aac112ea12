"""LAMMPS binary dump files, restated in numpy for the tests.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the product's
reader and writer (analisi_amd/csrc/host/lammps_traj.h).  The product package
``analisi_amd`` never imports it.

Format (lammps_struct.h:31-62, :65-97, :128-187, :104-122 of rikigigi/analisi
v0.5.2), little-endian, no padding between fields.  Each timestep is
  header  old:   int64 timestep (>= 0), int64 natoms, int32 triclinic,
                 int32 boundary[6], float64 box[6] (xlo xhi ylo yhi zlo zhi),
                 [float64 xy xz yz if triclinic], int32 size_one, int32 nchunk
          2020:  int64 -len(magic), magic, int32 endian, int32 revision,
                 int64 timestep, int64 natoms, int32 triclinic, int32 boundary[6],
                 float64 box[6], [float64 xy xz yz], int32 size_one,
                 [revision > 1: int32 unit_len, unit, char time_flag,
                  [float64 time], int32 columns_len, columns],
                 int32 nchunk
  chunks  nchunk x (int32 ndouble, ndouble float64 = ndouble/8 atoms of
          id, type, x, y, z, vx, vy, vz)
and size_one must be 8 (lammps_struct.h:212-214).

Atom order (trajectory.cpp:136-207): atom k is the k-th distinct id met in
the first timestep's chunks; later timesteps place atoms by id.  The box goes
to the internal form xlo, ylo, zlo, lx/2, ly/2, lz/2 [, xy, xz, yz]
(basetrajectory.h:94-105); the per-type centre of mass is the running mean of
the raw positions (and velocities) in file order (trajectory.cpp:604-617).
"""
import struct

import numpy as np

NDOUBLE = 8


def _header(b, off):
    """Parses one timestep header at off; returns (dict, offset of the first chunk)."""
    ts, = struct.unpack_from("<q", b, off)
    h = {}
    o = off
    if ts < 0:  # lammps_struct.h:147-187
        mlen = -ts
        o += 8
        h["magic"] = bytes(b[o:o + mlen])
        o += mlen
        h["endian"], h["revision"] = struct.unpack_from("<ii", b, o)
        o += 8
    h["timestep"], h["natoms"] = struct.unpack_from("<qq", b, o)
    o += 16
    h["triclinic"], = struct.unpack_from("<i", b, o)
    o += 4
    h["boundary"] = struct.unpack_from("<6i", b, o)
    o += 24
    h["box"] = np.array(struct.unpack_from("<6d", b, o))
    o += 48
    h["tilt"] = None
    if h["triclinic"]:
        h["tilt"] = np.array(struct.unpack_from("<3d", b, o))
        o += 24
    h["size_one"], = struct.unpack_from("<i", b, o)
    o += 4
    if ts < 0 and h["revision"] > 1:
        ulen, = struct.unpack_from("<i", b, o)
        o += 4 + max(ulen, 0)
        tflag = b[o]
        o += 1
        if tflag:
            o += 8
        clen, = struct.unpack_from("<i", b, o)
        o += 4 + clen
    h["nchunk"], = struct.unpack_from("<i", b, o)
    o += 4
    if h["size_one"] != NDOUBLE:
        raise ValueError("size_one %d" % h["size_one"])
    return h, o


def read_frames(path):
    """All timesteps: list of (header, atoms[natoms][8] in file order)."""
    b = open(path, "rb").read()
    off, out = 0, []
    while off < len(b):
        h, o = _header(b, off)
        rows = []
        for _ in range(h["nchunk"]):
            nd, = struct.unpack_from("<i", b, o)
            o += 4
            rows.append(np.frombuffer(b, dtype="<f8", count=nd, offset=o).reshape(-1, NDOUBLE))
            o += nd * 8
        out.append((h, np.concatenate(rows)))
        off = o
    return out


def lammps_to_internal(box6):
    c = np.array(box6, dtype=np.float64)
    return np.array([c[0], c[2], c[4], (c[1] - c[0]) / 2, (c[3] - c[2]) / 2, (c[5] - c[4]) / 2])


def window(path, start, n, wrap=False):
    """The reader's window [start, start + n): positions and velocities
    [n][natoms][3] in atom order, boxes [n][6|9] internal, raw lammps types
    per atom, the per-type centre of mass [n][ntypes][3] of positions and
    velocities, atom ids.  wrap is not restated here (the tests keep it off)."""
    if wrap:
        raise NotImplementedError
    fr = read_frames(path)
    h0, a0 = fr[0]
    ids = [int(round(x)) for x in a0[:, 0]]
    idmap = {}
    for i in ids:
        if i not in idmap:
            idmap[i] = len(idmap)
    natoms = len(idmap)
    raw_type = np.empty(natoms, dtype=np.int64)
    for row in a0:
        raw_type[idmap[int(round(row[0]))]] = int(round(row[1]))
    types = sorted(set(raw_type.tolist()))
    tid = np.array([types.index(t) for t in raw_type])
    nt = len(types)
    tri = bool(h0["triclinic"])
    pos = np.zeros((n, natoms, 3))
    vel = np.zeros((n, natoms, 3))
    box = np.zeros((n, 9 if tri else 6))
    pcm = np.zeros((n, nt, 3))
    vcm = np.zeros((n, nt, 3))
    for k in range(n):
        h, a = fr[start + k]
        box[k, :6] = lammps_to_internal(h["box"])
        if tri:
            box[k, 6:] = h["tilt"]
        cont = np.zeros(nt, dtype=np.int64)
        for row in a:
            i = idmap[int(round(row[0]))]
            pos[k, i] = row[2:5]
            vel[k, i] = row[5:8]
            t = tid[i]
            cont[t] += 1
            for c in range(3):
                pcm[k, t, c] += (pos[k, i, c] - pcm[k, t, c]) / cont[t]
                vcm[k, t, c] += (vel[k, i, c] - vcm[k, t, c]) / cont[t]
    lammps_id = np.empty(natoms, dtype=np.int64)
    for i, k in idmap.items():
        lammps_id[k] = i
    return dict(pos=pos, vel=vel, box=box, raw_type=raw_type, type_id=tid, pcm=pcm, vcm=vcm,
                lammps_id=lammps_id, timesteps=[f[0]["timestep"] for f in fr], nframes=len(fr))


def write(path, frames, fmt2020=False, chunks=1, order=None, magic=b"DUMPATOM", revision=2, timestep0=0):
    """Writes frames = list of (box6 lammps, tilt or None, ids[N], types[N],
    pos[N][3], vel[N][3]); order(k) gives the file order of frame k's atoms
    (a permutation), chunks splits each frame into that many chunks."""
    out = bytearray()
    for k, (box6, tilt, ids, types, pos, vel) in enumerate(frames):
        N = len(ids)
        if fmt2020:
            out += struct.pack("<q", -len(magic)) + magic + struct.pack("<ii", 1, revision)
        out += struct.pack("<qqi", timestep0 + k, N, 1 if tilt is not None else 0)
        out += struct.pack("<6i", 0, 0, 0, 0, 0, 0) + struct.pack("<6d", *box6)
        if tilt is not None:
            out += struct.pack("<3d", *tilt)
        out += struct.pack("<i", NDOUBLE)
        if fmt2020 and revision > 1:
            unit = b"metal"
            out += struct.pack("<i", len(unit)) + unit + bytes([1]) + struct.pack("<d", 0.5 * k)
            cols = b"id type x y z vx vy vz"
            out += struct.pack("<i", len(cols)) + cols
        out += struct.pack("<i", chunks)
        perm = np.arange(N) if order is None else np.asarray(order(k))
        data = np.empty((N, NDOUBLE))
        data[:, 0] = np.asarray(ids)[perm]
        data[:, 1] = np.asarray(types)[perm]
        data[:, 2:5] = np.asarray(pos)[perm]
        data[:, 5:8] = np.asarray(vel)[perm]
        cuts = np.linspace(0, N, chunks + 1).astype(int)
        for c in range(chunks):
            part = data[cuts[c]:cuts[c + 1]]
            out += struct.pack("<i", part.size) + part.astype("<f8").tobytes()
    open(path, "wb").write(bytes(out))
