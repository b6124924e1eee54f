"""dat_read -> dat_write reproduces the reference's own result files byte for byte, outside
the two blocks named below.

The file layout is that of d3qd2_src/d3qd2_setup.c:247-512 (structs stored raw, then the mesh
arrays, the incident field and every domain's U, dU, E, dE, H, dH).  The only bytes that may
differ are
* the 8 bytes of MIPW's `bd.ipw`, a heap pointer stored raw with the struct (its value is
  meaningless in a file and is replaced by the reader), and
* each Ipw's derived block IpwD (E0, ki, direction cosines, polarisation vectors): the reader
  recomputes it from the stored wave inputs, as the reference's dat_read does
  (d3qd2_src/d3qd2_setup.c:363 -> setup_mipw), so it carries this machine's sin/cos and may
  differ from the file by an ulp or in the sign of a zero.  It must agree to 2 ulp.
Host code only: no GPU call.
"""
import struct
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
FILES = sorted(p for p in GOLDEN.glob("*/*.dat"))

# offsets in the file: three 128-byte names, rv[3], th, tv[3], MN
_MN_AT = 3 * 128 + 7 * 8
_MIPW_PTR = 96          # offsetof(MIPW, bd.ipw): fn_ipw[64], n_ipw (+4 padding), n_0, lambda_0, omega
_MIPW_SIZE, _NIPW_AT = 104, 64
_IPW_SIZE, _IPWD_AT = 240, 96    # Ipw: wave inputs (96 B), then IpwD (144 B)


def _layout(raw):
    """(offset of MIPW, n_ipw)"""
    mn = struct.unpack_from("<i", raw, _MN_AT)[0]
    at = _MN_AT + 4 + 2 * 16 * (mn + 1)      # after n[MN+1] and kn[MN+1]
    return at, struct.unpack_from("<i", raw, at + _NIPW_AT)[0]


def test_reference_files_present():
    assert len(FILES) >= 4, FILES


@pytest.mark.parametrize("path", FILES, ids=[p.parent.name + "/" + p.name for p in FILES])
def test_dat_round_trip_is_byte_exact(path, tmp_path):
    from bem3_emf_qd2_amd.solver import Model
    raw = path.read_bytes()
    m = Model.from_dat(path)
    m.write(tmp_path / "rt.dat")
    m.close()
    out = (tmp_path / "rt.dat").read_bytes()
    assert len(out) == len(raw)
    at, nipw = _layout(raw)
    assert nipw >= 1
    ptr = at + _MIPW_PTR
    assert out[:ptr] == raw[:ptr]
    assert out[ptr + 8:at + _MIPW_SIZE] == raw[ptr + 8:at + _MIPW_SIZE]
    for i in range(nipw):
        o = at + _MIPW_SIZE + i * _IPW_SIZE
        assert out[o:o + _IPWD_AT] == raw[o:o + _IPWD_AT]          # stored inputs: exact
        x = np.frombuffer(out[o + _IPWD_AT:o + _IPW_SIZE], "<f8")
        y = np.frombuffer(raw[o + _IPWD_AT:o + _IPW_SIZE], "<f8")
        assert np.all(np.abs(x - y) <= 2 * np.spacing(np.maximum(np.abs(x), np.abs(y)))), (x - y)
    end = at + _MIPW_SIZE + nipw * _IPW_SIZE
    assert out[end:] == raw[end:]                                   # mesh, fields, solutions: exact
