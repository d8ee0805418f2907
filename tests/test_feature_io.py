"""Feature files through the host C ABI (gmmhost_features_*), against
writers and readers written here from the reference's layouts:
  SVES  "SV-ES-PARAM", int32 D, int32 N, float32 CF[N], float32 [N*D]
        (Param::Save / Load, param/OL_Param.cpp:177-220, :246-354;
        IOscripts/writePRM.m)
  HTK   big-endian uint32 nSamples, uint32 sampPeriod, int16 sampSize = 4 D,
        int16 parmKind, big-endian float32 rows (Param::LoadHTK, :477-576;
        IOscripts/writeHTK.m)
  RAW   int32 N, int32 D, float32 [N*D]; N = 0 means "count from the file
        size" (Param::LoadRaw, :358-453)
Down-sampling keeps frames 0, k, 2k, ... (ceil(N/k) of them, :284-318).
Frame intervals (Param::Load with frameIdxs, :581-1037, HTK variant
:897-1000): [x0, x1) in sample-period units covers frames
[x0 // p, x1 // p), clipped to the file; intervals ascending."""
import struct

import numpy as np
import pytest

from gmm_amd import host

PERIOD = 100000  # 10 ms in HTK's 100 ns units


def _frames(N=23, D=5, seed=0):
    x = np.random.default_rng(seed).normal(0, 3, (N, D)).astype(np.float32)
    if N >= 4 and D >= 3:
        # values whose bits must survive: signed zero, infinity, a denormal
        x[1, 0] = -0.0
        x[2, 1] = np.inf
        x[3, 2] = np.float32(1e-42)
    return x


def write_sves(path, x, cf=None):
    N, D = x.shape
    cf = np.ones(N, np.float32) if cf is None else np.asarray(cf, np.float32)
    with open(path, "wb") as f:
        f.write(b"SV-ES-PARAM")
        f.write(struct.pack("<ii", D, N))
        f.write(cf.astype("<f4").tobytes())
        f.write(x.astype("<f4").tobytes())


def write_htk(path, x, period=PERIOD, kind=6 + 64):
    N, D = x.shape
    with open(path, "wb") as f:
        f.write(struct.pack(">IIhh", N, period, 4 * D, kind))
        f.write(x.astype(">f4").tobytes())


def write_raw(path, x, n_header=None):
    N, D = x.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", N if n_header is None else n_header, D))
        f.write(x.astype("<f4").tobytes())


WRITERS = {host.SVES: write_sves, host.HTK: write_htk, host.RAW: write_raw}


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("kind", [host.SVES, host.HTK, host.RAW])
def test_load_matches_layout(tmp_path, kind):
    x = _frames()
    p = tmp_path / "f"
    if kind == host.SVES:
        # confidence factors are read past, whatever they hold
        write_sves(p, x, cf=np.linspace(0, 1, x.shape[0]))
    else:
        WRITERS[kind](p, x)
    y, period = host.load_features(p, kind)
    assert same_bits(x, y)
    assert period == (PERIOD if kind == host.HTK else 0)


@pytest.mark.parametrize("kind", [host.SVES, host.HTK, host.RAW])
def test_save_matches_layout(tmp_path, kind):
    x = _frames(17, 4, 1)
    p = tmp_path / "f"
    host.save_features(p, kind, x, sample_period=PERIOD)
    b = p.read_bytes()
    N, D = x.shape
    if kind == host.SVES:
        assert b[:11] == b"SV-ES-PARAM"
        assert struct.unpack("<ii", b[11:19]) == (D, N)
        assert np.array_equal(np.frombuffer(b[19:19 + 4 * N], "<f4"), np.ones(N))
        body = np.frombuffer(b[19 + 4 * N:], "<f4")
    elif kind == host.HTK:
        assert struct.unpack(">IIhh", b[:12]) == (N, PERIOD, 4 * D, 9)
        body = np.frombuffer(b[12:], ">f4")
    else:
        assert struct.unpack("<ii", b[:8]) == (N, D)
        body = np.frombuffer(b[8:], "<f4")
    assert same_bits(body.astype(np.float32).reshape(N, D), x)
    y, _ = host.load_features(p, kind)
    assert same_bits(x, y)


@pytest.mark.parametrize("kind", [host.SVES, host.HTK, host.RAW])
@pytest.mark.parametrize("k", [2, 3, 7, 23, 40])
def test_downsampling(tmp_path, kind, k):
    x = _frames()
    p = tmp_path / "f"
    WRITERS[kind](p, x)
    y, _ = host.load_features(p, kind, dwnsmp=k)
    assert y.shape[0] == -(-x.shape[0] // k)
    assert same_bits(x[::k], y)


def test_raw_count_from_file_size(tmp_path):
    x = _frames(11, 3)
    write_raw(tmp_path / "f", x, n_header=0)
    y, _ = host.load_features(tmp_path / "f", host.RAW)
    assert same_bits(x, y)


@pytest.mark.parametrize("kind,body", [
    (host.SVES, b"SV-ES-PARAM" + struct.pack("<ii", 5, 0)),
    (host.SVES, b"SV-ES-PARAM" + struct.pack("<ii", 0, 3) + b"\0" * 12),
    (host.HTK, struct.pack(">IIhh", 0, PERIOD, 20, 9)),
    (host.HTK, struct.pack(">IIhh", 4, PERIOD, 2, 9)),
    (host.RAW, struct.pack("<ii", 4, 0)),
])
def test_empty_files(tmp_path, kind, body):
    """No frames or no dimension: an empty set (N = D = 0), no error, as
    Param leaves it (:328-331, :427-430, :552-555)."""
    (tmp_path / "f").write_bytes(body)
    y, _ = host.load_features(tmp_path / "f", kind)
    assert y.shape == (0, 0)


def test_errors(tmp_path):
    with pytest.raises(host.HostError) as e:
        host.load_features(tmp_path / "missing", host.SVES)
    assert e.value.code == host.FEAT_MISSING
    x = _frames(6, 4)
    write_raw(tmp_path / "r", x)
    with pytest.raises(host.HostError) as e:  # not SVES
        host.load_features(tmp_path / "r", host.SVES)
    assert e.value.code == 1 and "SV-ES-PARAM" in str(e.value)
    for kind in (host.SVES, host.HTK, host.RAW):
        WRITERS[kind](tmp_path / "t", x)
        (tmp_path / "t").write_bytes((tmp_path / "t").read_bytes()[:-5])
        with pytest.raises(host.HostError) as e:
            host.load_features(tmp_path / "t", kind)
        assert e.value.code == 1
    with pytest.raises(host.HostError):
        host.load_features(tmp_path / "r", 7)


def _expect(x, ranges, p):
    """Brute force: the frames of every interval, clipped to the file."""
    rows = []
    for a, b in ranges:
        rows.extend(range(a // p, min(b // p, x.shape[0])))
    return x[rows]


@pytest.mark.parametrize("kind", [host.SVES, host.HTK, host.RAW])
def test_ranges(tmp_path, kind, capfd):
    x = _frames(40, 6, 3)
    p = tmp_path / "f"
    WRITERS[kind](p, x)
    arg = 0 if kind == host.HTK else PERIOD  # HTK uses its header's period
    ranges = [[0, 250000], [250000, 800000], [1234567, 1399999], [2000000, 2100000],
              [3950000, 4000000]]
    y, _ = host.load_feature_ranges(p, kind, ranges, arg)
    assert same_bits(y, _expect(x, ranges, PERIOD))
    # an interval running past the end is clipped, with a warning
    y, _ = host.load_feature_ranges(p, kind, [[100000, 300000], [3800000, 4600000]], arg)
    assert same_bits(y, _expect(x, [[100000, 300000], [3800000, 4600000]], PERIOD))
    assert "too short: 6 frames missing" in capfd.readouterr().err
    # SVES / RAW carry no period: the caller's, in other units
    if kind != host.HTK:
        y, _ = host.load_feature_ranges(p, kind, [[3, 9], [12, 13]], 1)
        assert same_bits(y, x[[3, 4, 5, 6, 7, 8, 12]])


@pytest.mark.parametrize("kind", [host.SVES, host.HTK, host.RAW])
@pytest.mark.parametrize("ranges", [
    [[300000, 300000]],                      # begin == end
    [[500000, 200000]],                      # begin > end
    [[4000000, 4100000]],                    # begins at the end of the file
    [[500000, 900000], [100000, 200000]],    # not ascending
    [[500000, 900000], [800000, 1000000]],   # overlapping
])
def test_bad_ranges(tmp_path, kind, ranges):
    x = _frames(40, 6, 3)
    WRITERS[kind](tmp_path / "f", x)
    with pytest.raises(host.HostError) as e:
        host.load_feature_ranges(tmp_path / "f", kind, ranges, 0 if kind == host.HTK else PERIOD)
    assert e.value.code == 1


@pytest.mark.parametrize("kind", [host.SVES, host.RAW])
def test_ranges_need_a_period(tmp_path, kind):
    WRITERS[kind](tmp_path / "f", _frames(10, 2))
    with pytest.raises(host.HostError, match="sample period"):
        host.load_feature_ranges(tmp_path / "f", kind, [[0, 5]], 0)
