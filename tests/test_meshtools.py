"""Mesh transforms of the reference's mesh_sample/tools (rotation: Rodrigues' formula about a
unit axis; stretch: scaling by |v| along v) on a Gmsh 2.2 file written by meshgen: node
coordinates against scipy's rotation and the closed form of the stretch, the inverse
transform, the file layout (node lines '%d' + ' %16.15f' x 3, every other line unchanged),
and that the library reads the result as a model."""
from pathlib import Path

import numpy as np
import pytest

from bem3_emf_qd2_amd import meshgen


def _nodes(path):
    lines = Path(path).read_text().splitlines()
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    rows = [l.split() for l in lines[i + 2:i + 2 + n]]
    return lines, i, np.array([[float(v) for v in r[1:]] for r in rows]), [r[0] for r in rows]


def _mesh(tmp_path):
    V, Q = meshgen.cube_sphere(3, 0.4)
    p = tmp_path / "s.msh"
    meshgen.write_msh(p, V, [(Q, 1, 0, 1)])
    return p


@pytest.mark.parametrize("axis,angle", [((1, 0, 0), np.pi), ((0.3, -0.5, 0.8), 0.7), ((0, 0, 2), -1.2)])
def test_msh_rotate(tmp_path, axis, angle):
    from scipy.spatial.transform import Rotation
    src = _mesh(tmp_path)
    dst, back = tmp_path / "r.msh", tmp_path / "b.msh"
    meshgen.msh_rotate(src, dst, axis, angle)
    l0, i0, X0, ids0 = _nodes(src)
    l1, i1, X1, ids1 = _nodes(dst)
    u = np.array(axis, float) / np.linalg.norm(axis)
    assert np.abs(X1 - Rotation.from_rotvec(angle * u).apply(X0)).max() < 1e-14
    assert ids1 == ids0 and i1 == i0 and l1[:i0 + 2] == l0[:i0 + 2]
    n = len(X0)
    assert l1[i0 + 2 + n:] == l0[i0 + 2 + n:]
    assert all(len(v.split(".")[1]) == 15 for v in l1[i0 + 2].split()[1:])
    meshgen.msh_rotate(dst, back, axis, -angle)
    assert np.abs(_nodes(back)[2] - X0).max() < 1e-14


def test_msh_stretch(tmp_path):
    src = _mesh(tmp_path)
    dst = tmp_path / "s2.msh"
    v = np.array([0.0, 1.0, 1.0])
    meshgen.msh_stretch(src, dst, v)
    X0, X1 = _nodes(src)[2], _nodes(dst)[2]
    s, u = np.linalg.norm(v), v / np.linalg.norm(v)
    want = X0 + (s - 1.0) * (X0 @ u)[:, None] * u[None, :]
    assert np.abs(X1 - want).max() < 1e-14
    meshgen.msh_stretch(src, tmp_path / "x2.msh", (2.0, 0, 0))
    X2 = _nodes(tmp_path / "x2.msh")[2]
    assert np.allclose(X2[:, 0], 2 * X0[:, 0], atol=1e-15) and np.allclose(X2[:, 1:], X0[:, 1:], atol=1e-15)
    with pytest.raises(ValueError):
        meshgen.msh_stretch(src, dst, (0, 0, 0))


def test_rotated_mesh_is_a_model(tmp_path):
    """A sphere scene whose mesh was rotated reads as a model with the rotated nodes."""
    from bem3_emf_qd2_amd.solver import Model
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from datfile import read_dat
    meshgen.sphere_scene(tmp_path, 3, 0.4, 1.5, 0.6)
    msh = next(tmp_path.glob("*.msh"))
    X0 = _nodes(msh)[2]
    meshgen.msh_rotate(msh, msh, (0, 1, 0), 0.5)
    m = Model.from_dir(tmp_path)
    m.write(tmp_path / "setup.dat")
    m.close()
    rn = read_dat(tmp_path / "setup.dat")["rn"][1:]
    from scipy.spatial.transform import Rotation
    assert np.abs(rn - Rotation.from_rotvec([0, 0.5, 0]).apply(X0)).max() < 1e-14
