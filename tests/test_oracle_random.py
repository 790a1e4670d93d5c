"""Oracle Philox stream vs the reference's Random123 known-answer vectors and
the statistical checks of test/test_random.jl."""
import numpy as np
import pytest

import fcc_hip
import oracle


def test_philox_known_answers():
    # test/test_random.jl:5-10 (Random123 KATs)
    assert oracle.philox4x32((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert oracle.philox4x32((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2) == (
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert oracle.philox4x32((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
                             (0xA4093822, 0x299F31D0)) == (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)


def test_uniform_open_bounds():
    # test/test_random.jl:55 and the closed form of random.jl:42-44
    lo, hi = oracle.uniform_open64(0, 0), oracle.uniform_open64(0xFFFFFFFF, 0xFFFFFFFF)
    assert 0 < lo < hi < 1
    assert lo == 0.5 * 2.0 ** -52 and hi == (2.0 ** 52 - 0.5) * 2.0 ** -52
    lo32, hi32 = oracle.uniform_open32(0), oracle.uniform_open32(0xFFFFFFFF)
    assert 0 < lo32 < hi32 < 1
    assert np.float32(lo32) == np.float32(0.5 * 2.0 ** -23)


def test_host_philox_matches_oracle():
    # the host helper used for initial orientations is the same Philox stream
    u = fcc_hip.philox_uniforms(1234, 5, stream=7)
    for i in range(5):
        w = oracle.philox4x32((i + 1, 7, 0, 0), (1234, 0))
        assert u[i] == oracle.uniform_open64(w[0], w[1])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_normal_noise_statistics(oracle_engine, prec):
    # test/test_random.jl:13-34
    xi = fcc_hip.philox_noise(100_000, 42, 7, prec, engine=oracle_engine)
    assert xi.dtype == (np.float32 if prec == "fp32" else np.float64)
    for r in range(3):
        assert abs(xi[r].mean()) < 0.015
        assert abs(xi[r].var() - 1) < 0.02
        assert abs((xi[r].astype(np.float64) ** 4).mean() - 3) < 0.1
    assert abs(np.corrcoef(xi[0], xi[1])[0, 1]) < 0.015
    assert abs(np.corrcoef(xi[0], xi[2])[0, 1]) < 0.015
    xi8 = fcc_hip.philox_noise(100_000, 42, 8, prec, engine=oracle_engine)
    assert abs(np.corrcoef(xi.ravel(), xi8.ravel())[0, 1]) < 0.01
    np.testing.assert_array_equal(fcc_hip.philox_noise(10, 42, 7, prec, engine=oracle_engine), xi[:, :10])
    assert not np.array_equal(fcc_hip.philox_noise(10, 43, 7, prec, engine=oracle_engine), xi[:, :10])
    assert np.all(np.isfinite(xi))


def test_draw_index_uses_64_bits(oracle_engine):
    # test/test_random.jl:46-49
    a = fcc_hip.philox_noise(1, 1, 1, "fp64", engine=oracle_engine)
    b = fcc_hip.philox_noise(1, 1, 1 + (1 << 32), "fp64", engine=oracle_engine)
    assert not np.array_equal(a, b)


def test_box_muller_closed_form(oracle_engine):
    # box_muller (random.jl:47-51) from the oracle's words, recomputed in numpy
    xi = fcc_hip.philox_noise(4, 9, 12, "fp64", engine=oracle_engine)
    for i in range(4):
        w = oracle.philox4x32((i + 1, 12, 0, 0), (9, 0))
        u1, u2 = oracle.uniform_open64(w[0], w[1]), oracle.uniform_open64(w[2], w[3])
        r = np.sqrt(-2 * np.log(u1))
        np.testing.assert_allclose(xi[:2, i], [r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)],
                                   rtol=4e-16, atol=1e-300)
