"""GPU parity of the corner turns either side of the beamformer, through the
C ABI, against the oracle: the TUSE beam corner turn
(meerkat/tuse/src/transpose_to_dada.cpp:19-48) and the F-engine heap reorder
(meerkat/tools/src/feng_to_dada.cu:54-80).  Byte moves: bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(capi, a):
    return capi.DeviceBuffer.from_array(np.ascontiguousarray(a))


@pytest.mark.parametrize("C,S,F,B,G,b0,nb", [
    (128, 64, 32, 4, 10, 0, 4),     # the reference's default geometry (transpose_to_dada.cpp:13-16), 4 beams
    (64, 128, 16, 6, 2, 0, 6),      # 16-byte units, full 16 x 16 element tiles
    (64, 40, 13, 3, 2, 1, 2),       # partial tiles, a beam range
    (12, 33, 17, 2, 3, 1, 1),       # 4-byte units, element 12 bytes
    (3, 9, 70, 4, 2, 2, 2),         # 1-byte units, odd element size, > 64 bands
    (4096, 4, 3, 2, 2, 0, 2),       # large elements: one element per tile
    (16, 1, 1, 1, 1, 0, 1),         # degenerate 1 x 1 matrix
])
def test_tuse_transpose_bit_exact(gpu, oracle, C, S, F, B, G, b0, nb):
    capi = gpu
    rng = np.random.default_rng(C * 7 + S)
    blk = rng.integers(-128, 128, G * B * F * S * C).astype(np.int8)
    d_in = _dev(capi, blk)
    d_out = capi.DeviceBuffer(nb * G * S * F * C)
    capi.tuse_transpose(d_in, d_out, C, S, F, B, G, b0, nb)
    np.testing.assert_array_equal(d_out.download(), oracle.tuse_transpose(blk, C, S, F, B, G, b0, nb))


def test_tuse_transpose_full_size_against_numpy(gpu):
    """The bench workload: 512 beams x 16 bands x 512 samples (4 groups of
    128) x 64 channels, 256 MiB; checked against numpy's transpose of the
    same block."""
    capi = gpu
    C, S, F, B, G = 64, 128, 16, 512, 4
    n = G * B * F * S * C
    blk = capi.synth_gaussian_int8(99, 30.0, n)
    d_in = _dev(capi, blk)
    d_out = capi.DeviceBuffer(n)
    capi.tuse_transpose(d_in, d_out, C, S, F, B, G)
    got = d_out.download()
    ref = blk.reshape(G, B, F, S, C).transpose(1, 0, 3, 2, 4).reshape(-1)
    assert np.array_equal(got, ref)


def test_tuse_transpose_rejects_bad_sizes(gpu):
    capi = gpu
    C, S, F, B, G = 16, 8, 4, 3, 2
    d_in = capi.DeviceBuffer(G * B * F * S * C)
    with pytest.raises(RuntimeError, match="geometry needs"):
        capi.tuse_transpose(capi.DeviceBuffer(100), capi.DeviceBuffer(B * G * S * F * C), C, S, F, B, G)
    with pytest.raises(RuntimeError, match="output must be"):
        capi.tuse_transpose(d_in, capi.DeviceBuffer(7), C, S, F, B, G)
    with pytest.raises(RuntimeError, match="beam range"):
        capi.tuse_transpose(d_in, capi.DeviceBuffer(2 * G * S * F * C), C, S, F, B, G, 2, 2)


@pytest.mark.parametrize("nchans,nts", [(32, 3), (4096, 2), (96, 5), (40, 1), (1, 4)])
def test_feng_to_dada_bit_exact(gpu, oracle, nchans, nts):
    capi = gpu
    rng = np.random.default_rng(nchans + nts)
    w = rng.integers(-2**31, 2**31, nts * nchans * 256, dtype=np.int64).astype(np.int32)
    d_in = _dev(capi, w)
    d_out = capi.DeviceBuffer(w.nbytes)
    capi.feng_to_dada(d_in, d_out, nchans)
    np.testing.assert_array_equal(d_out.download(np.int32), oracle.feng_to_dada(w, nchans))


def test_feng_to_dada_rejects_bad_sizes(gpu):
    capi = gpu
    d = capi.DeviceBuffer(32 * 256 * 4 * 2)
    with pytest.raises(RuntimeError, match="bytes per timestamp"):
        capi.call("fbf_feng_to_dada", d.ptr, capi.DeviceBuffer(d.nbytes).ptr, d.nbytes - 4, 32, None)
    with pytest.raises(RuntimeError, match="distinct"):
        capi.feng_to_dada(d, d, 32)
