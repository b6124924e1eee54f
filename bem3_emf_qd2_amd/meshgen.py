"""Input generation for the benchmark and test scenes: cube-sphere meshes in the Gmsh 2.2
ASCII paired-element format the reference reads (d3qd2_src/d3qd2_setup.c:662-959), and
the three text input files (ipw, periodicity, medium; d3qd2_src/m_ipw.c:9-106,
d3qd2_src/d3qd2_setup.c:557-650).

Cube-sphere: each cube face is split into an n x n grid of equal angles (gnomonic
projection) and pushed radially onto the sphere, like mesh_sample/sphere_m4.geo does with
transfinite circle arcs; 6 n^2 bilinear quads, 6 n^2 + 2 nodes.

Mesh convention (checked against mesh_sample/sphere_m4.msh): each surface element is
written twice, first for the inner domain with its normal pointing out of that domain, then
for the outer domain with the node order reversed (0, 3, 2, 1); the open region is
physical 99.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

OPEN = 99


def cube_sphere(n: int, radius: float, center=(0.0, 0.0, 0.0)):
    """Return (nodes[N,3], quads[M,4]) with 0-based node ids and outward normals."""
    t = np.tan(np.linspace(-math.pi / 4, math.pi / 4, n + 1))
    faces = []
    # (axis, sign): face normal direction; local u, v axes chosen so (u x v) = outward
    for axis in range(3):
        for sign in (1.0, -1.0):
            u_ax, v_ax = [(1, 2), (2, 0), (0, 1)][axis]
            if sign < 0:
                u_ax, v_ax = v_ax, u_ax
            faces.append((axis, sign, u_ax, v_ax))
    key_to_id: dict[tuple, int] = {}
    nodes: list[np.ndarray] = []
    quads = []

    def node_id(p):
        k = tuple(np.round(p, 12))
        if k not in key_to_id:
            key_to_id[k] = len(nodes)
            nodes.append(p)
        return key_to_id[k]

    for axis, sign, ua, va in faces:
        ids = np.zeros((n + 1, n + 1), int)
        for i in range(n + 1):
            for j in range(n + 1):
                p = np.zeros(3)
                p[axis] = sign
                p[ua] = t[i]
                p[va] = t[j]
                p = p / np.linalg.norm(p)
                ids[i, j] = node_id(p)
        for i in range(n):
            for j in range(n):
                quads.append([ids[i, j], ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]])
    X = np.array(nodes) * radius + np.asarray(center, float)
    Q = np.array(quads, int)
    # orient every quad outward (robust against the per-face axis bookkeeping)
    c = X[Q].mean(1) - np.asarray(center, float)
    nrm = np.cross(X[Q[:, 2]] - X[Q[:, 0]], X[Q[:, 3]] - X[Q[:, 1]])
    flip = np.einsum("ij,ij->i", nrm, c) < 0
    Q[flip] = Q[flip][:, ::-1]
    return X, Q


def _ring(radius: float, z: float, m: int, phase: float):
    ang = phase + 2.0 * math.pi * np.arange(m) / m
    return np.c_[radius * np.cos(ang), radius * np.sin(ang), np.full(m, z)], ang


def _strip(ia, aa, ib, ab):
    """Triangulate the band between two closed rings of vertex ids ia (angles aa) and ib
    (angles ab), merging by angle; a ring of one vertex is a pole."""
    tris = []
    if len(ia) == 1 or len(ib) == 1:
        pole, ring = (ia[0], ib) if len(ia) == 1 else (ib[0], ia)
        return [[pole, ring[k], ring[(k + 1) % len(ring)]] for k in range(len(ring))]
    na, nb = len(ia), len(ib)
    i = j = 0
    ua = np.unwrap(np.r_[aa, aa[0] + 2 * math.pi])
    ub = np.unwrap(np.r_[ab, ab[0] + 2 * math.pi])
    while i < na or j < nb:
        if j >= nb or (i < na and ua[i + 1] <= ub[j + 1]):
            tris.append([ia[i % na], ia[(i + 1) % na], ib[j % nb]])
            i += 1
        else:
            tris.append([ia[i % na], ib[(j + 1) % nb], ib[j % nb]])
            j += 1
    return tris


def cone(radius: float, height: float, h: float, z0: float = 0.0, jitter: float = 0.0, seed: int = 0):
    """Right circular cone (base disk at z0, apex at z0 + height) as linear triangles of edge
    about h: rings of nearly uniform spacing on the lateral surface and the base, bands
    between rings triangulated by angle.  jitter moves every vertex off the apex, the base
    centre and the rim by jitter * h in a seeded random direction.  Returns (nodes[N, 3],
    tris[M, 3]) with 0-based ids and outward normals."""
    slant = math.hypot(radius, height)
    n_l = max(2, round(slant / h))
    n_b = max(2, round(radius / h))
    m_rim = max(6, round(2 * math.pi * radius / h))
    X = [np.array([[0.0, 0.0, z0 + height]])]
    rings = [(np.array([0]), np.array([0.0]))]
    nid = 1
    for i in range(1, n_l + 1):  # lateral, apex -> rim
        s = i / n_l
        m = m_rim if i == n_l else max(3, round(2 * math.pi * radius * s / h))
        P, ang = _ring(radius * s, z0 + height * (1 - s), m, 0.5 * (i % 2) * 2 * math.pi / m)
        X.append(P)
        rings.append((np.arange(nid, nid + m), ang))
        nid += m
    rim = rings[-1]
    tris = []
    for a, b in zip(rings[:-1], rings[1:]):
        tris += _strip(a[0], a[1], b[0], b[1])
    base = [rim]
    for j in range(n_b - 1, 0, -1):  # base, rim -> centre
        s = j / n_b
        m = max(3, round(2 * math.pi * radius * s / h))
        P, ang = _ring(radius * s, z0, m, 0.5 * (j % 2) * 2 * math.pi / m)
        X.append(P)
        base.append((np.arange(nid, nid + m), ang))
        nid += m
    X.append(np.array([[0.0, 0.0, z0]]))
    base.append((np.array([nid]), np.array([0.0])))
    nid += 1
    for a, b in zip(base[:-1], base[1:]):
        tris += _strip(a[0], a[1], b[0], b[1])
    X = np.vstack(X)
    T = np.array(tris, int)
    if jitter:
        fixed = np.zeros(len(X), bool)
        fixed[0] = fixed[nid - 1] = True
        fixed[rim[0]] = True
        rng = np.random.default_rng(seed)
        v = rng.normal(size=(len(X), 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        X = X + np.where(fixed[:, None], 0.0, jitter * h * v)
    # outward orientation: the cone is convex, so a face normal points away from its
    # interior point (0, 0, z0 + height / 4)
    c = X[T].mean(1) - np.array([0.0, 0.0, z0 + height / 4])
    nrm = np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]])
    flip = np.einsum("ij,ij->i", nrm, c) < 0
    T[flip] = T[flip][:, ::-1]
    return X, T


def split_quads(Q: np.ndarray, every: int):
    """Mixed mesh: every `every`-th quad (a, b, c, d) becomes the triangles (a, b, c) and
    (a, c, d), which keep its orientation.  Returns a list of 4- and 3-node rows."""
    out = []
    for i, q in enumerate(Q):
        if every and i % every == 0:
            out.append([q[0], q[1], q[2]])
            out.append([q[0], q[2], q[3]])
        else:
            out.append(list(q))
    return out


def write_msh(path, nodes: np.ndarray, surfaces) -> None:
    """surfaces: list of (elements[M,k] 0-based, inner_domain, outer_domain, entity_id);
    element normals point out of inner_domain.  Domain 0 is written as physical 99."""
    path = Path(path)
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(nodes))]
    for i, p in enumerate(nodes, 1):
        lines.append(f"{i} {p[0]:.16g} {p[1]:.16g} {p[2]:.16g}")
    lines += ["$EndNodes", "$Elements"]
    body = []
    eid = 1
    for elems, din, dout, ent in surfaces:
        pin, pout = (din or OPEN), (dout or OPEN)
        for e in elems:
            e1 = [int(v) + 1 for v in e]
            typ = 3 if len(e1) == 4 else 2
            rev = [e1[0], e1[3], e1[2], e1[1]] if typ == 3 else [e1[0], e1[2], e1[1]]
            body.append(f"{eid} {typ} 2 {pin} {ent} " + " ".join(map(str, e1)))
            body.append(f"{eid + 1} {typ} 2 {pout} {ent} " + " ".join(map(str, rev)))
            eid += 2
    lines.append(str(len(body)))
    lines += body
    lines.append("$EndElements")
    path.write_text("\n".join(lines) + "\n")


def write_ipw(path, waves) -> None:
    """waves: list of dicts with lambda0, ni, power, e0x, e0y (complex), theta, phi."""
    out = ["## plane wave data", "# number of wave", str(len(waves)),
           "# lambda ni power real(e0x) imag(e0x) real(e0y) imag(e0y) fx fy fz theta phi"]
    for w in waves:
        ex, ey = complex(w.get("e0x", 1.0)), complex(w.get("e0y", 0.0))
        out.append(f"{w['lambda0']:.17g} {w['ni']:.17g} {w.get('power', 0.5 * w['ni']):.17g} "
                   f"{ex.real:.17g} {ex.imag:.17g} {ey.real:.17g} {ey.imag:.17g} 0 0 0 "
                   f"{w['theta']:.17g} {w.get('phi', 0.0):.17g}")
    Path(path).write_text("\n".join(out) + "\n")


def write_periodicity(path, d1, d2, ps=1) -> None:
    Path(path).write_text("## periodicity datafile\n# ps d1x d1y d2x d2y\n"
                          f"{ps} {d1[0]:.17g} {d1[1]:.17g} {d2[0]:.17g} {d2[1]:.17g}\n")


def write_medium(path, indices) -> None:
    lines = ["## medium datafile", "# number of medium", str(len(indices)), "# real(n) imag(n)"]
    lines += [f"{complex(n).real:.17g} {complex(n).imag:.17g}" for n in indices]
    Path(path).write_text("\n".join(lines) + "\n")


def sphere_scene(dirpath, n: int, radius: float, n_sphere, lam: float, n0: float = 1.0,
                 theta: float = 1e-6, phi: float = 0.0, pol: str = "TM", d1=(1.0, 0.0), d2=(0.0, 1.0),
                 core=None, tri_every: int = 0, ps: int = 1):
    """Write a complete input set for a (optionally two-layer) sphere on a lattice, or alone
    in free space with ps = 0.

    core = (radius, index) adds a concentric core (domain 2 inside the shell domain 1).
    tri_every = k > 0 splits every k-th quad of the outer surface into two linear triangles.
    pol: "TM" -> e0x = 1 (p-polarised in the beam frame), "TE" -> e0y = 1
    (d3qd2_src/ipw.c:139-144).  Amplitude 1 means power = n0 / 2 (d3qd2_src/ipw.c:131).
    """
    d = Path(dirpath)
    d.mkdir(parents=True, exist_ok=True)
    X, Q = cube_sphere(n, radius)
    surfaces = [(split_quads(Q, tri_every) if tri_every else Q, 1, 0, 1)]
    media = [n_sphere]
    if core is not None:
        Xc, Qc = cube_sphere(n, core[0])
        Qc = Qc + len(X)
        X = np.vstack([X, Xc])
        surfaces.append((Qc, 2, 1, 2))
        media.append(core[1])
    write_msh(d / "mesh.msh", X, surfaces)
    ex, ey = (1.0, 0.0) if pol.upper() == "TM" else (0.0, 1.0)
    write_ipw(d / "ipw.txt", [dict(lambda0=lam, ni=n0, power=0.5 * n0, e0x=ex, e0y=ey, theta=theta, phi=phi)])
    write_periodicity(d / "periodicity_data.txt", d1, d2, ps=ps)
    write_medium(d / "medium_data.txt", media)
    return d


def sphere_cone_scene(dirpath, n: int = 16, core=(0.3, 1.6 + 0.01j), shell=(0.40, 1.45),
                      cone_=(0.3, 0.6, 2.0), cone_h: float = 0.03, cone_z0: float = 0.5, jitter: float = 0.01,
                      seed: int = 2, lam: float = 0.55, n0: float = 1.0, theta: float = math.radians(30),
                      phi: float = math.radians(15), pol: str = "TM", d1=(0.9, 0.0), d2=(0.45, 0.779)):
    """Two-layer cube-sphere (6 n^2 quads per surface) centred at the origin and a linear-
    triangle cone (base radius, height, index) standing on the plane z = cone_z0 above it, in
    one hexagonal lattice cell.  Domains: 1 shell, 2 core, 3 cone.  The default shell radius
    0.40 um keeps the sphere off its own periodic images: with |a2| = 0.8996 um a 0.45 um
    shell would overlap them."""
    d = Path(dirpath)
    d.mkdir(parents=True, exist_ok=True)
    Xs, Qs = cube_sphere(n, shell[0])
    Xc, Qc = cube_sphere(n, core[0])
    Xk, Tk = cone(cone_[0], cone_[1], cone_h, z0=cone_z0, jitter=jitter, seed=seed)
    X = np.vstack([Xs, Xc, Xk])
    surfaces = [(Qs, 1, 0, 1), (Qc + len(Xs), 2, 1, 2), (Tk + len(Xs) + len(Xc), 3, 0, 3)]
    write_msh(d / "mesh.msh", X, surfaces)
    ex, ey = (1.0, 0.0) if pol.upper() == "TM" else (0.0, 1.0)
    write_ipw(d / "ipw.txt", [dict(lambda0=lam, ni=n0, power=0.5 * n0, e0x=ex, e0y=ey, theta=theta, phi=phi)])
    write_periodicity(d / "periodicity_data.txt", d1, d2)
    write_medium(d / "medium_data.txt", [shell[1], core[1], cone_[2]])
    return d


# ---- transforms of an existing Gmsh 2.2 ASCII mesh (the reference's mesh_sample/tools) ----

def _msh_map_nodes(src, dst, fn) -> None:
    """Copy a Gmsh 2.2 ASCII file, replacing every node's coordinates by fn(x, y, z); node
    lines are written as the reference's tools write them ('%d' then ' %16.15f' x 3), all
    other lines unchanged."""
    lines = Path(src).read_text().splitlines(keepends=True)
    out = []
    i = 0
    while i < len(lines):
        out.append(lines[i])
        if lines[i].strip() == "$Nodes":
            n = int(lines[i + 1])
            out.append(lines[i + 1])
            for k in range(n):
                f = lines[i + 2 + k].split()
                v = fn(float(f[1]), float(f[2]), float(f[3]))
                out.append(f[0] + "".join(f" {c:16.15f}" for c in v) + "\n")
            i += 2 + n
            continue
        i += 1
    Path(dst).write_text("".join(out))


def msh_rotate(src, dst, axis, angle: float) -> None:
    """Rotate every node by `angle` (rad) about the unit vector along `axis` (Rodrigues'
    formula), as mesh_sample/tools/rotation/main.c:41-93 does."""
    n = [float(a) for a in axis]
    a = math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
    if a == 0.0:
        raise ValueError("msh_rotate: the rotation axis is the zero vector")
    n = [n[0] / a, n[1] / a, n[2] / a]
    st, ct = math.sin(angle), math.cos(angle)
    R = [ct + n[0] * n[0] * (1.0 - ct), n[0] * n[1] * (1.0 - ct) - n[2] * st, n[2] * n[0] * (1.0 - ct) + n[1] * st,
         n[0] * n[1] * (1.0 - ct) + n[2] * st, ct + n[1] * n[1] * (1.0 - ct), n[1] * n[2] * (1.0 - ct) - n[0] * st,
         n[2] * n[0] * (1.0 - ct) - n[1] * st, n[1] * n[2] * (1.0 - ct) + n[0] * st, ct + n[2] * n[2] * (1.0 - ct)]
    _msh_map_nodes(src, dst, lambda x, y, z: [R[3 * j] * x + R[3 * j + 1] * y + R[3 * j + 2] * z for j in range(3)])


def msh_stretch(src, dst, vector) -> None:
    """Scale every node by |vector| along the direction of `vector` (the perpendicular part is
    kept), as mesh_sample/tools/stretch/main.c:27-84 does."""
    v = [float(c) for c in vector]
    s = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    if s == 0.0:
        raise ValueError("msh_stretch: the stretch vector is the zero vector")
    u = [v[0] / s, v[1] / s, v[2] / s]

    def f(x, y, z):
        p = (x, y, z)
        d = p[0] * u[0] + p[1] * u[1] + p[2] * u[2]
        return [s * (d * u[j]) + (p[j] - d * u[j]) for j in range(3)]

    _msh_map_nodes(src, dst, f)


if __name__ == "__main__":
    import sys
    a = sys.argv[1:]
    if len(a) == 7 and a[0] == "rotate":
        # python -m bem3_emf_qd2_amd.meshgen rotate in.msh nx ny nz angle_rad out.msh
        msh_rotate(a[1], a[6], (float(a[2]), float(a[3]), float(a[4])), float(a[5]))
        print(f"Done writing {a[6]}")
    elif len(a) == 6 and a[0] == "stretch":
        # python -m bem3_emf_qd2_amd.meshgen stretch in.msh sx sy sz out.msh
        msh_stretch(a[1], a[5], (float(a[2]), float(a[3]), float(a[4])))
        print(f"Done writing {a[5]}")
    else:
        print("usage: python -m bem3_emf_qd2_amd.meshgen rotate in.msh nx ny nz angle_rad out.msh\n"
              "       python -m bem3_emf_qd2_amd.meshgen stretch in.msh sx sy sz out.msh")
        sys.exit(1)
