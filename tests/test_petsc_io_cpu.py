"""PETSc binary Vec files (the format of the reference's cp%04d.dat, Restart0x.dat and RestartSol0x.dat,
written with PetscViewerBinary VecView: LinearElasticity.cc:1128-1137, 1261-1276, 2474-2486;
TopOpt.cc:614-735).  Host-only entry points of the C-ABI, so these run without a GPU.

The expected bytes are packed here with `struct` from the documented layout (per vector: big-endian
int32 class id 1211214, big-endian int32 length, big-endian IEEE doubles).  The reference tree holds
no PETSc-written file to compare against.
"""
import struct

import numpy as np
import pytest

from topopt_in_petsc_transient_amd import capi

VEC_CLASSID = 1211214


def packed(*vecs):
    out = b""
    for v in vecs:
        out += struct.pack(">ii", VEC_CLASSID, len(v)) + struct.pack(">%dd" % len(v), *v)
    return out


def test_writer_bytes_match_the_petsc_layout(tmp_path):
    a = [1.0, -2.5, 3e-300, float("inf")]
    b = [0.1, 2.0 ** 60]
    p = tmp_path / "cp0000.dat"
    capi.petsc_write_vecs(p, [a, b])
    assert p.read_bytes() == packed(a, b)


def test_reader_reads_a_hand_packed_file(tmp_path):
    rng = np.random.default_rng(3)
    u, v, acc = (rng.standard_normal(n) for n in (12, 12, 12))
    p = tmp_path / "state.dat"
    p.write_bytes(packed(u, v, acc))
    ru, rv, ra = capi.petsc_read_vecs(p, [12, 12, 12])
    assert np.array_equal(ru, u) and np.array_equal(rv, v) and np.array_equal(ra, acc)
    # VecLoad reads vectors in order: the first ones of a longer file
    (first,) = capi.petsc_read_vecs(p, [12])
    assert np.array_equal(first, u)


def test_round_trip_keeps_every_bit(tmp_path):
    rng = np.random.default_rng(4)
    vecs = [rng.standard_normal(1000) * 10.0 ** rng.integers(-300, 300, 1000), np.zeros(0), rng.random(7)]
    p = tmp_path / "Restart00.dat"
    capi.petsc_write_vecs(p, vecs)
    back = capi.petsc_read_vecs(p, [v.size for v in vecs])
    for v, w in zip(vecs, back):
        assert v.tobytes() == w.tobytes()


@pytest.mark.parametrize("case", ["missing", "classid", "length", "short_header", "short_values"])
def test_reader_rejects_bad_files(tmp_path, case):
    p = tmp_path / "bad.dat"
    good = packed([1.0, 2.0, 3.0])
    if case == "classid":  # e.g. a Mat file, or a Vec written with 64-bit PetscInt
        p.write_bytes(struct.pack(">ii", 1211216, 3) + good[8:])
    elif case == "length":
        p.write_bytes(good)
    elif case == "short_header":
        p.write_bytes(good[:6])
    elif case == "short_values":
        p.write_bytes(good[:-4])
    lens = [4] if case == "length" else [3]
    with pytest.raises(capi.TTError) as e:
        capi.petsc_read_vecs(p, lens)
    assert "tt_petsc_read_vecs" in str(e.value)


def test_writer_rejects_a_bad_path(tmp_path):
    with pytest.raises(capi.TTError):
        capi.petsc_write_vecs(tmp_path / "no" / "such" / "dir.dat", [[1.0]])
