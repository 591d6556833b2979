"""Visibility egress (host code, runs without a GPU): the CSV writer against a
restatement of MDT_OneStation_client.cu:1239-1266 written here (byte for
byte), both input layouts, the complete cross-polarisation enumeration, the
timestamped name (client.cu:1209) and the binary dump's round trip."""
import os
import time

import numpy as np
import pytest

import mdtfx


def _vis(rng, Ne, Np, N, Ns=16):
    NN = Ne * Np
    x = (rng.standard_normal((N, Ns, NN)) + 1j * rng.standard_normal((N, Ns, NN))) * 3.7
    V = np.einsum("csa,csb->cab", x, x.conj()) / Ns  # V[c][a][b] = mean x_a conj(x_b)
    full = np.ascontiguousarray(np.transpose(V, (0, 2, 1))).astype(np.complex64)  # column-major: [c][b][a]
    a, b = np.tril_indices(NN)
    tri = V[:, a, b].astype(np.complex64)  # packed lower triangle: V[a][b], a >= b
    return full, tri, V


def _reference_csv(full, Ne, Np, N, complete=False):
    """client.cu:1239-1266: j, p, k >= j, q >= p (q from 0 for k > j when complete);
    channels N/2..N-1 then 0..N/2-1; value XOutput[c*NN*NN + (j*Np+p) + (k*Np+q)*NN]"""
    NN = Ne * Np
    flat = full.reshape(-1)
    out = []
    for j in range(Ne):
        for p in range(Np):
            for k in range(j, Ne):
                for q in range(0 if (complete and k > j) else p, Np):
                    z = 0
                    for i in list(range(N // 2, N)) + list(range(N // 2)):
                        v = flat[i * NN * NN + (k * Np + q) * NN + (j * Np + p)]
                        out.append("%d,%d,%d,%d,%d,%.6f,%.6f\n" % (j, p, k, q, z, float(v.real), float(v.imag)))
                        z += 1
    return "".join(out)


@pytest.mark.parametrize("Ne,Np,N", [(3, 2, 8), (5, 1, 16), (2, 2, 32)])
def test_csv_matches_reference_writer(tmp_path, Ne, Np, N):
    full, tri, _ = _vis(np.random.default_rng(Ne * 10 + N), Ne, Np, N)
    want = _reference_csv(full, Ne, Np, N)
    mdtfx.write_csv(tmp_path / "f.csv", full, mdtfx.VIS_FULL_COLMAJOR, Ne, Np, N)
    assert (tmp_path / "f.csv").read_text() == want
    # the native triangle gives the same file (V[a][b] = conj of the stored V[b][a])
    mdtfx.write_csv(tmp_path / "t.csv", tri, mdtfx.VIS_TRIANGLE, Ne, Np, N)
    assert (tmp_path / "t.csv").read_text() == want


def test_csv_complete_pol_pairs(tmp_path):
    Ne, Np, N = 3, 2, 8
    full, tri, V = _vis(np.random.default_rng(5), Ne, Np, N)
    mdtfx.write_csv(tmp_path / "c.csv", tri, mdtfx.VIS_TRIANGLE, Ne, Np, N, all_pol_pairs=True)
    text = (tmp_path / "c.csv").read_text()
    assert text == _reference_csv(full, Ne, Np, N, complete=True)
    rows = [l.split(",") for l in text.splitlines()]
    # every (element pair, pol pair) with (j,p) <= (k,q) in the element order, N rows each
    pairs = {(int(r[0]), int(r[1]), int(r[2]), int(r[3])) for r in rows}
    assert (0, 1, 1, 0) in pairs and len(rows) == len(pairs) * N
    ref_pairs = _reference_csv(full, Ne, Np, N).count("\n") // N
    assert len(pairs) == ref_pairs + Ne * (Ne - 1) // 2  # the (j, 1, k, 0) rows the reference skips
    # values: row (0,1,1,0,z) is V[a=1][b=2] at source channel (z + N/2) % N
    r = [x for x in rows if x[:4] == ["0", "1", "1", "0"]][3]
    v = V[(3 + N // 2) % N, 1, 2]
    assert r[5] == "%.6f" % float(np.float32(v.real)) and r[6].strip() == "%.6f" % float(np.float32(v.imag))


def test_csv_filename():
    old = os.environ.get("TZ")
    os.environ["TZ"] = "UTC"
    time.tzset()
    try:
        assert mdtfx.csv_filename(0) == "MDT_19700101_000000.csv"
        assert mdtfx.csv_filename(1733270399) == "MDT_20241203_235959.csv"
    finally:
        if old is None:
            del os.environ["TZ"]
        else:
            os.environ["TZ"] = old
        time.tzset()


def test_binary_round_trip(tmp_path):
    full, tri, _ = _vis(np.random.default_rng(9), 4, 2, 16)
    mdtfx.write_binary(tmp_path / "v.bin", tri, mdtfx.VIS_TRIANGLE, 4, 2, 16, 512, unix_time=1234567890)
    hdr, data = mdtfx.read_binary(tmp_path / "v.bin")
    assert hdr == dict(layout=1, dtype=0, Nelements=4, Npols=2, Nchannels=16, Nspectra=512, unix_time=1234567890)
    assert np.array_equal(data.reshape(-1), tri.reshape(-1))
    acc = tri.astype(np.complex128) * 3
    mdtfx.write_binary(tmp_path / "a.bin", acc, mdtfx.VIS_TRIANGLE, 4, 2, 16, 512)
    hdr, data = mdtfx.read_binary(tmp_path / "a.bin")
    assert hdr["dtype"] == 1 and np.array_equal(data.reshape(-1), acc.reshape(-1))
    assert os.path.getsize(tmp_path / "v.bin") == 64 + tri.nbytes


def test_egress_errors(tmp_path):
    full, tri, _ = _vis(np.random.default_rng(1), 2, 1, 8)
    with pytest.raises(mdtfx.MdtError):
        mdtfx.write_csv(tmp_path / "x.csv", tri, 7, 2, 1, 8)
    with pytest.raises(mdtfx.MdtError):
        mdtfx.write_csv(tmp_path / "x.csv", np.zeros(8 * 36, np.complex64), mdtfx.VIS_FULL_COLMAJOR, 2, 3, 8)
    with pytest.raises(mdtfx.MdtError):  # odd channel count: no fftshift halves
        mdtfx.write_csv(tmp_path / "x.csv", full, mdtfx.VIS_FULL_COLMAJOR, 2, 1, 7)
    with pytest.raises(mdtfx.MdtError):
        mdtfx.write_csv(tmp_path / "no_such_dir" / "x.csv", full, mdtfx.VIS_FULL_COLMAJOR, 2, 1, 8)
    with pytest.raises(ValueError):
        mdtfx.write_csv(tmp_path / "x.csv", tri[:3], mdtfx.VIS_TRIANGLE, 2, 1, 8)
