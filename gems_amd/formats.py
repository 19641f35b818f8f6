"""GEMS's on-disk formats, read and written as the reference does, so a run on
the device can start from, and hand back to, the reference's own files.

* Checkpoints (`checkpoint write|read`, write_chp / read_chp,
  src/Checkpoint.F90:154-235): Fortran unformatted sequential records (4-byte
  length markers) in this order: (current, step, nsteps, nlocal) int32; the
  box tensor tbox(dm, dm) FP64; the PRNG name ('lcg') and its state (idum, iy,
  ix) int32 (lcg_write_chpseed, src/Random.F90:178-184); the piston flags
  logical(dm) and, per piston axis, three FP64 records
  (write_chppiston, src/Integration.F90:1487-1500); (nframe, time, dm_steps,
  pnframe, ptime) FP64 + frame int32; then six records of dm FP64 per atom:
  pos, vel, acel, acel2, acel3, acel4.
* xyz frames (`outfile ... pos|vel|fce 1`, write_pos / write_vel / write_fce,
  src/Output.F90:900-1000): the atom count and (nframe, time) list-directed,
  then per atom the symbol in 5 characters and three `2x,e25.12` fields
  (forces in kcal/mol/A).  `>< read` accepts the same layout.
* Energy columns (`outfile ... cols energy 1`, write_energy,
  src/Output.F90:824-831): [epot, ekin, total] in eV, ` e20.7` per field by
  default.

Checkpoint contents map to a context with `load_checkpoint` / `checkpoint_of`
(positions, velocities, box, the lcg stream's state); atom types and masses
come from the script, as in the reference.
"""
import struct

import numpy as np

DM = 3


# ---------------------------------------------------------------------------
# Fortran unformatted sequential records
# ---------------------------------------------------------------------------
def read_records(path):
    """The records of a Fortran unformatted sequential file (4-byte markers)."""
    b = open(path, "rb").read()
    out, i = [], 0
    while i < len(b):
        n = struct.unpack("<i", b[i:i + 4])[0]
        if n < 0 or i + 8 + n > len(b) or struct.unpack("<i", b[i + 4 + n:i + 8 + n])[0] != n:
            raise ValueError("%s: not a Fortran unformatted sequential file (record at byte %d)"
                             % (path, i))
        out.append(b[i + 4:i + 4 + n])
        i += 8 + n
    return out


def write_records(path, records):
    with open(path, "wb") as f:
        for r in records:
            f.write(struct.pack("<i", len(r)) + r + struct.pack("<i", len(r)))


# ---------------------------------------------------------------------------
# checkpoints
# ---------------------------------------------------------------------------
def read_chp(path):
    """write_chp's records (src/Checkpoint.F90:154-187) as a dict."""
    rec = read_records(path)
    k = 0
    current, step, nsteps, nlocal = struct.unpack("<4i", rec[k]); k += 1
    tbox = np.frombuffer(rec[k], "<f8").reshape(DM, DM, order="F").copy(); k += 1
    prng = rec[k].decode("ascii"); k += 1
    if prng != "lcg":
        raise NotImplementedError("checkpoint PRNG %r: only the lcg stream is supported" % prng)
    lcg = struct.unpack("<3i", rec[k]); k += 1
    piston = struct.unpack("<%di" % DM, rec[k]); k += 1
    pst = []
    for d in range(DM):
        if piston[d]:
            pst.append(tuple(struct.unpack("<d", rec[k + u])[0] for u in range(3)))
            k += 3
    nframe, time, dm_steps, pnframe, ptime = struct.unpack("<5d", rec[k][:40])
    frame = struct.unpack("<i", rec[k][40:44])[0]
    k += 1
    per = np.array([np.frombuffer(r, "<f8") for r in rec[k:k + 6 * nlocal]])
    if per.shape != (6 * nlocal, DM):
        raise ValueError("%s: %d atoms announced, %d atom records found"
                         % (path, nlocal, len(rec) - k))
    per = per.reshape(nlocal, 6, DM)
    names = ("pos", "vel", "acel", "acel2", "acel3", "acel4")
    out = dict(current=current, step=step, nsteps=nsteps, nlocal=nlocal, tbox=tbox,
               prng=prng, lcg=lcg, piston=tuple(bool(p) for p in piston), piston_state=pst,
               nframe=nframe, time=time, dm_steps=dm_steps, pnframe=pnframe, ptime=ptime,
               frame=frame)
    for u, nm in enumerate(names):
        out[nm] = per[:, u, :].copy()
    return out


def write_chp(path, ck):
    """The inverse of read_chp: the same records in the same order."""
    n = int(ck["nlocal"])
    recs = [struct.pack("<4i", ck["current"], ck["step"], ck["nsteps"], n),
            np.asarray(ck["tbox"], "<f8").reshape(DM, DM).tobytes(order="F"),
            ck.get("prng", "lcg").encode("ascii"),
            struct.pack("<3i", *ck["lcg"]),
            struct.pack("<%di" % DM, *[1 if p else 0 for p in ck.get("piston", (False,) * DM)])]
    for st in ck.get("piston_state", []):
        recs += [struct.pack("<d", v) for v in st]
    recs.append(struct.pack("<5di", ck["nframe"], ck["time"], ck["dm_steps"], ck["pnframe"],
                            ck["ptime"], ck["frame"]))
    arrs = [np.asarray(ck[nm], "<f8").reshape(n, DM)
            for nm in ("pos", "vel", "acel", "acel2", "acel3", "acel4")]
    for i in range(n):
        for a in arrs:
            recs.append(a[i].tobytes())
    write_records(path, recs)


# ---------------------------------------------------------------------------
# xyz frames and energy columns
# ---------------------------------------------------------------------------
def fortran_e(v, w, d):
    """Fortran `Ew.d` editing: 0.<d digits>E+ee, right-justified in w (the
    exponent loses its E beyond two digits)."""
    v = float(v)
    if v == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, ex = ("%.*e" % (d - 1, abs(v))).split("e")
        ex = int(ex) + 1
        digits = m.replace(".", "")
        sign = "+" if ex >= 0 else "-"
        s = "0." + digits + ("E%s%02d" % (sign, abs(ex)) if abs(ex) <= 99
                             else "%s%03d" % (sign, abs(ex)))
        if v < 0:
            s = "-" + s
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def read_xyz(path):
    """Frames of a GEMS xyz output: a list of (symbols, values[n, 3], nframe,
    time)."""
    lines = open(path).read().splitlines()
    frames, k = [], 0
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        n = int(lines[k].split()[0])
        hdr = lines[k + 1].split()
        nframe = float(hdr[0]) if hdr else 0.0
        time = float(hdr[1]) if len(hdr) > 1 else 0.0
        syms, vals = [], []
        for ln in lines[k + 2:k + 2 + n]:
            t = ln.split()
            syms.append(t[0])
            vals.append([float(x.replace("D", "E")) for x in t[1:1 + DM]])
        frames.append((syms, np.array(vals), nframe, time))
        k += 2 + n
    return frames


def write_xyz(path, syms, values, nframe=0.0, time=0.0, append=False):
    """One frame in write_pos's layout (a5, 3(2x,e25.12) per atom)."""
    with open(path, "a" if append else "w") as f:
        f.write(" %d\n" % len(syms))
        f.write(" %r %r\n" % (float(nframe), float(time)))
        for s, v in zip(syms, np.asarray(values, float)):
            f.write("%-5s" % s[:5] + "".join("  " + fortran_e(x, 25, 12) for x in v[:DM]) + "\n")


def read_energy(path):
    """Energy columns (eV) as an array [rows, 3]."""
    return np.loadtxt(path, ndmin=2)


def write_energy(path, rows, w=20, d=7):
    """write_energy's rows: ` e20.7` per field (the default `out float`)."""
    with open(path, "w") as f:
        for r in np.asarray(rows, float).reshape(-1, 3):
            f.write("".join(" " + fortran_e(x, w, d) for x in r) + "\n")


# ---------------------------------------------------------------------------
# a context <-> a checkpoint
# ---------------------------------------------------------------------------
def load_checkpoint(ctx, ck, types=None, pbc=(1, 1, 1)):
    """Positions, velocities, box and the lcg stream of a checkpoint into a
    context (orthorhombic boxes; types as the script declares them)."""
    tbox = np.asarray(ck["tbox"])
    if np.any(np.abs(tbox - np.diag(np.diag(tbox))) > 0):
        raise NotImplementedError("only orthorhombic checkpoint boxes are supported")
    ctx.set_box(list(np.diag(tbox)), pbc)
    ctx.set_atoms(ck["pos"], types)
    ctx.set_velocities(ck["vel"])
    ctx.lcg_state(ck["lcg"])


def checkpoint_of(ctx, L, masses_per_atom, step=1, nsteps=1, nframe=0.0, time=0.0):
    """A checkpoint of a context's state in the reference's layout (acel =
    F/m as velocity Verlet leaves it; the higher derivatives zero)."""
    x, v, F, _ = ctx.state()
    n = len(x)
    acel = F / np.asarray(masses_per_atom, float)[:, None]
    z = np.zeros((n, DM))
    return dict(current=1, step=step, nsteps=nsteps, nlocal=n, tbox=np.diag(np.asarray(L, float)),
                prng="lcg", lcg=ctx.lcg_state(), piston=(False,) * DM, piston_state=[],
                nframe=float(nframe), time=float(time), dm_steps=0.0, pnframe=0.0, ptime=0.0,
                frame=0, pos=x, vel=v, acel=acel, acel2=z, acel3=z, acel4=z)
