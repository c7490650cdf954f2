"""CPU tests of the host-side readers of libmpchol: the settings.json options
file and the big-endian idx container of the MNIST data (no GPU needed).

tests/golden/reference_settings.json is the reference's own options file
(src/settings.json, a data fixture).  The MNIST label files are read from the
reference tree when it is present; the images are not in it
(.MISSING_LARGE_BLOBS), so the image path is checked on files written here.
"""
import json
import os
import struct

import numpy as np
import pytest

from cholesky_gpu_amd import _lib as L
from cholesky_gpu_amd import solver as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_SETTINGS = os.path.join(HERE, "golden", "reference_settings.json")
REF_MNIST = "/root/reference/src/MNIST_data"


# ---------------------------------------------------------------- settings.json
def test_reference_settings_file():
    s = S.settings_read(GOLDEN_SETTINGS)
    assert s.n == 8192
    assert s.condition_number == 100.0
    assert s.dist_type == L.DIST_ARITHMETIC
    assert s.block_size == 1024
    assert s.eps_prime == pytest.approx(0.000005, rel=1e-15)
    assert s.floor == pytest.approx(0.01, rel=1e-15)
    assert s.diag_pert == 0 and s.left == 0
    mo, fo = S.settings_options(s)
    assert mo.dist == L.DIST_ARITHMETIC and mo.cond == 100.0
    # block 1024 = four 256-column panels aggregated into one K = 1024 update
    assert fo.nb == 256 and fo.agg_panels == 4
    assert fo.eps_prime == pytest.approx(0.000005, rel=1e-15)


def test_settings_matches_python_json():
    """Every value agrees with Python's own JSON reader of the same file."""
    ref = json.load(open(GOLDEN_SETTINGS))
    s = S.settings_read(GOLDEN_SETTINGS)
    m, f = ref["matrix_settings"], ref["factorization_settings"]
    assert s.n == int(m["n"])
    assert s.condition_number == float(m["condition_number"])
    assert s.block_size == int(f["block_size"])
    assert s.eps_prime == float(f["eps_prime"])
    assert s.floor == float(f["floor"])
    assert s.diag_pert == int(f["diag_pert"]) and s.left == int(f["left"])


def test_settings_spellings_numbers_and_defaults():
    # the key spellings of the reference's reader (json_utils.hpp), numbers
    # instead of strings, escapes and nesting the parser must get through
    text = ('{"matrix settings": {"n": 1024, "condition number": 1e6, "dist type": "Geometric",'
            ' "note": "tab\\tquote\\" \\u00e9\\ud83d\\ude00", "list": [1, [2, {"x": null}], true]},'
            ' "factorization settings": {"block size": "128", "eps prime": 0.5, "left": true}}')
    s = S.settings_parse(text)
    assert (s.n, s.condition_number, s.dist_type) == (1024, 1e6, L.DIST_GEOMETRIC)
    assert (s.block_size, s.eps_prime, s.left) == (128, 0.5, 1)
    _, fo = S.settings_options(s)
    assert fo.nb == 128 and fo.agg_panels == 1
    d = S.settings_parse("{}")        # everything absent: the defaults
    dflt = L.Settings()
    L.load().mpc_settings_default(__import__("ctypes").byref(dflt))
    assert (d.n, d.block_size, d.dist_type) == (dflt.n, dflt.block_size, dflt.dist_type)
    for name, code in [("clustered", L.DIST_CLUSTERED), ("RandomLogUniform", L.DIST_RANDOM_LOG_UNIFORM),
                       ("logspaced", L.DIST_LOGSPACED), ("2", L.DIST_ARITHMETIC)]:
        assert S.settings_parse('{"matrix_settings": {"dist_type": "%s"}}' % name).dist_type == code


@pytest.mark.parametrize("bad", [
    '{"matrix_settings": {"n": "12x"}}',                        # not a number
    '{"matrix_settings": {"n": 0}}',                            # out of range
    '{"matrix_settings": {"dist_type": "gaussian"}}',           # unknown distribution
    '{"matrix_settings": {"condition_number": "0.5"}}',          # cond < 1
    '{"factorization_settings": {"eps_prime": "-1"}}',
    '{"matrix_settings": {"n": 1}',                             # unterminated object
    '{"a": "\\x"}',                                             # bad escape
    '{"a": 1} trailing',
    '[1, 2]',                                                   # not an object
    '{"a": "\\ud800"}',                                         # unpaired surrogate
    "",
])
def test_settings_rejects(bad):
    with pytest.raises(L.MpcError) as e:
        S.settings_parse(bad)
    assert e.value.code == L.ERR_ARG


def test_settings_block_sizes():
    for b, nb, agg in [(64, 64, 1), (256, 256, 1), (512, 256, 2), (768, 256, 3), (1024, 256, 4)]:
        _, fo = S.settings_options(S.settings_parse('{"factorization_settings": {"block_size": %d}}' % b))
        assert (fo.nb, fo.agg_panels) == (nb, agg)
    for b in (32, 300, 1280, 2048):
        s = S.settings_parse('{"factorization_settings": {"block_size": %d}}' % b)
        with pytest.raises(L.MpcError) as e:
            S.settings_options(s)
        assert e.value.code == L.ERR_UNSUPPORTED


def test_settings_missing_file(tmp_path):
    with pytest.raises(L.MpcError):
        S.settings_read(tmp_path / "absent.json")


# ---------------------------------------------------------------- idx
def write_idx(path, arr, type_code=0x08):
    arr = np.asarray(arr)
    with open(path, "wb") as f:
        f.write(bytes([0, 0, type_code, arr.ndim]))
        for d in arr.shape:
            f.write(struct.pack(">I", d))
        f.write(arr.astype(">u1" if type_code == 0x08 else ">i4").tobytes())


def test_idx_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    imgs = rng.integers(0, 256, size=(7, 28, 28), dtype=np.uint8)
    p = tmp_path / "imgs-idx3-ubyte"
    write_idx(p, imgs)
    t, dims, off = S.idx_header(p)
    assert (t, dims, off) == (0x08, (7, 28, 28), 16)
    assert np.array_equal(S.idx_read_u8(p), imgs)
    assert np.array_equal(S.idx_read_u8(p, 3), imgs[:3])
    labels = rng.integers(0, 10, size=11, dtype=np.uint8)
    q = tmp_path / "labels-idx1-ubyte"
    write_idx(q, labels)
    assert S.idx_header(q)[:2] == (0x08, (11,))
    assert np.array_equal(S.idx_read_u8(q), labels)


def test_idx_rejects(tmp_path):
    p = tmp_path / "short"
    write_idx(p, np.zeros((4, 3), np.uint8))
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-1])                   # size no longer matches the header
    with pytest.raises(L.MpcError):
        S.idx_header(p)
    q = tmp_path / "magic"
    open(q, "wb").write(b"\x01\x00\x08\x01" + struct.pack(">I", 0))
    with pytest.raises(L.MpcError):
        S.idx_header(q)
    r = tmp_path / "ints"
    write_idx(r, np.arange(6).reshape(2, 3), type_code=0x0C)
    assert S.idx_header(r)[0] == 0x0C
    with pytest.raises(L.MpcError):                  # not unsigned bytes
        S.idx_read_u8(r)


# the published class counts of the MNIST label sets
MNIST_COUNTS = {
    "train-labels-idx1-ubyte": (60000, [5923, 6742, 5958, 6131, 5842, 5421, 5918, 6265, 5851, 5949]),
    "t10k-labels-idx1-ubyte": (10000, [980, 1135, 1032, 1010, 982, 892, 958, 1028, 974, 1009]),
}


@pytest.mark.parametrize("name", sorted(MNIST_COUNTS))
def test_reference_mnist_labels(name):
    path = os.path.join(REF_MNIST, name)
    if not os.path.exists(path):
        pytest.skip("reference MNIST label files not present")
    raw = open(path, "rb").read(4)
    assert struct.unpack(">I", raw)[0] == 0x0801
    t, dims, off = S.idx_header(path)
    count, hist = MNIST_COUNTS[name]
    assert (t, dims, off) == (0x08, (count,), 8)
    lab = S.idx_read_u8(path)
    assert np.bincount(lab, minlength=10).tolist() == hist
