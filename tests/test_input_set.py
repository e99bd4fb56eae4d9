"""Input sets on disk (bgw_inputs_write / bgw_inputs_open, include/bgw.h): the
library's writer and reader agree with each other and with an independent numpy
parser of the documented layout and checksum; damaged or incomplete files are
refused with BGW_EIO and bad descriptors with BGW_EINVAL.  Host only, no GPU."""
import struct

import numpy as np
import pytest

import bgw

MASK = (1 << 64) - 1
GOLD = 0x9E3779B97F4A7C15


def _mix64_np(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _checksum_np(payload: bytes) -> int:
    pad = (-len(payload)) % 8
    w = np.frombuffer(payload + b"\0" * pad, dtype="<u8")
    idx = np.arange(w.size, dtype=np.uint64) * np.uint64(GOLD)
    with np.errstate(over="ignore"):
        return int(_mix64_np(w ^ idx).sum(dtype=np.uint64))


def _checksum_py(payload: bytes) -> int:
    """Pure-Python restatement for small payloads (pins the numpy one)."""
    payload += b"\0" * ((-len(payload)) % 8)
    h = 0
    for i in range(len(payload) // 8):
        z = int.from_bytes(payload[8 * i:8 * i + 8], "little") ^ ((i * GOLD) & MASK)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        h = (h + (z ^ (z >> 31))) & MASK
    return h


_NP = {1: "<f8", 2: "<c16", 3: "<i4", 4: "<i8"}


def _parse(raw: bytes) -> tuple[int, dict]:
    """Independent reader of the layout in include/bgw.h."""
    assert raw[:8] == b"BGWINPUT"
    version, kind, count, zero, diroff = struct.unpack_from("<IIIIQ", raw, 8)
    assert (version, zero, diroff) == (1, 0, 32)
    out = {}
    for i in range(count):
        e = 32 + 96 * i
        name = raw[e:e + 32].rstrip(b"\0").decode()
        dtype, ndim = struct.unpack_from("<II", raw, e + 32)
        dims = struct.unpack_from("<4Q", raw, e + 40)
        off, nbytes, csum = struct.unpack_from("<QQQ", raw, e + 72)
        assert off % 4096 == 0
        payload = raw[off:off + nbytes]
        assert _checksum_np(payload) == csum, name
        a = np.frombuffer(payload, dtype=_NP[dtype])
        out[name] = a.reshape(dims[:ndim]) if ndim else a[0]
        assert all(d == 1 for d in dims[ndim:])
    return kind, out


def test_checksum_restatements_agree():
    rng = np.random.default_rng(3)
    for n in (0, 1, 7, 8, 9, 63, 200):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert _checksum_np(b) == _checksum_py(b)


def test_gpp_round_trip_and_independent_parser(tmp_path):
    d = bgw.gpp_inputs(7, 2, 40, 5, 123)
    p = tmp_path / "gpp.bgwin"
    bgw.save_inputs(p, d, "gpp")
    e = bgw.load_inputs(p)
    for k in ("aqsm", "aqsn", "I_eps", "wtilde", "vcoul", "inv", "indinv", "wx"):
        assert e[k].dtype == np.asarray(d[k]).dtype and np.array_equal(e[k], d[k]), k
    assert (e["nb"], e["nvband"], e["ncouls"], e["ngpown"], e["occ"]) == (7, 2, 40, 5, 1.0)
    kind, arrs = _parse(p.read_bytes())
    assert kind == bgw.BGW_INPUTS_GPP
    assert arrs["aqsmtemp"].shape == (7, 40) and arrs["I_eps_array"].shape == (5, 40)
    assert arrs["wx_array"].shape == (7, 3) and arrs["indinv"].shape == (41,)
    assert np.array_equal(arrs["wtilde_array"].reshape(-1), d["wtilde"])
    assert int(arrs["number_bands"]) == 7 and float(arrs["occ"]) == 1.0


def test_ff_round_trip_and_independent_parser(tmp_path):
    f = bgw.ff_inputs(6, 2, 30, 4, 5, 3, 77)
    p = tmp_path / "ff.bgwin"
    bgw.save_inputs(p, f, "ff")
    e = bgw.load_inputs(p)
    for k in ("aqsm", "aqsn", "I_epsR", "I_epsA", "vcoul", "inv", "indinv", "ekq", "dFreqGrid"):
        assert np.array_equal(e[k], np.asarray(f[k]).reshape(-1)), k
    assert (e["nFreq"], e["nfreqeval"], e["freqevalmin"], e["freqevalstep"]) == (5, 3, 0.0, 0.5)
    kind, arrs = _parse(p.read_bytes())
    assert kind == bgw.BGW_INPUTS_FF and arrs["I_epsA_array"].shape == (5, 4, 30)


def test_mapped_views_are_read_only_and_zero_copy(tmp_path):
    d = bgw.gpp_inputs(3, 1, 16, 2, 9)
    p = tmp_path / "g.bgwin"
    bgw.save_inputs(p, d, "gpp")
    with bgw.InputSet(p) as s:
        a = s.arrays["I_eps_array"]
        assert not a.flags.writeable and not a.flags.owndata
        assert a.ctypes.data % 4096 == 0
        with pytest.raises(ValueError):
            a[0, 0] = 0


def test_views_outlive_close(tmp_path):
    import gc

    d = bgw.gpp_inputs(3, 1, 16, 2, 9)
    p = tmp_path / "g.bgwin"
    bgw.save_inputs(p, d, "gpp")
    s = bgw.InputSet(p)
    v = s.arrays["wtilde_array"]
    s.close()
    del s
    gc.collect()
    assert np.array_equal(v.reshape(-1), d["wtilde"])  # still mapped through the view
    maps = lambda: open("/proc/self/maps").read().count(str(p))  # noqa: E731
    assert maps() == 1
    del v
    gc.collect()
    assert maps() == 0  # unmapped with the last view


def _gpp_file(tmp_path, name="g.bgwin"):
    p = tmp_path / name
    bgw.save_inputs(p, bgw.gpp_inputs(4, 1, 24, 3, 5), "gpp")
    return p


def _expect_eio(p, verify=True, match=None):
    with pytest.raises(bgw.BgwError) as ei:
        bgw.InputSet(p, verify)
    assert ei.value.code == bgw.BGW_EIO
    if match:
        assert match in str(ei.value), str(ei.value)


def test_checksum_mismatch_names_the_entry(tmp_path):
    p = _gpp_file(tmp_path)
    raw = bytearray(p.read_bytes())
    kind, arrs = _parse(bytes(raw))
    # locate I_eps_array's payload through the directory and flip one bit in it
    for i in range(len(arrs)):
        e = 32 + 96 * i
        if raw[e:e + 32].rstrip(b"\0") == b"I_eps_array":
            off = struct.unpack_from("<Q", raw, e + 72)[0]
    raw[off + 5] ^= 0x10
    p.write_bytes(bytes(raw))
    _expect_eio(p, True, "checksum mismatch in 'I_eps_array'")
    with bgw.InputSet(p, verify=False) as s:  # structure is still valid
        assert s.kind == bgw.BGW_INPUTS_GPP


def test_malformed_files_are_refused(tmp_path):
    good = _gpp_file(tmp_path).read_bytes()
    cases = {
        "magic": b"BGWINPUX" + good[8:],
        "version": good[:8] + struct.pack("<I", 2) + good[12:],
        "kind": good[:12] + struct.pack("<I", 9) + good[16:],
        "reserved": good[:20] + struct.pack("<I", 1) + good[24:],
        "unused_dim": good[:32 + 40 + 24] + struct.pack("<Q", 2) + good[32 + 40 + 32:],
        "truncated_directory": good[:40],
        "truncated_payload": good[:len(good) - 100],
        "short": good[:16],
    }
    for name, raw in cases.items():
        q = tmp_path / f"{name}.bgwin"
        q.write_bytes(raw)
        _expect_eio(q)
    _expect_eio(tmp_path / "does_not_exist.bgwin")


def test_mutated_files_are_refused_or_read_in_bounds(tmp_path):
    """Seeded random damage to the header and directory (byte overwrites, bit flips,
    large and small 64-bit values, truncation): every file is either refused with
    BGW_EIO or opens with each entry inside the file, whose bytes then read back as
    the file holds them."""
    good = _gpp_file(tmp_path).read_bytes()
    rng = np.random.default_rng(20261204)
    head = 32 + 96 * 13  # header + directory of the GPP set
    q = tmp_path / "mut.bgwin"
    opened = 0
    for _ in range(400):
        raw = bytearray(good)
        for _ in range(int(rng.integers(1, 5))):
            pos = int(rng.integers(0, head))
            kind = int(rng.integers(0, 4))
            if kind == 0:
                raw[pos] = int(rng.integers(0, 256))
            elif kind == 1:
                raw[pos] ^= 1 << int(rng.integers(0, 8))
            elif kind == 2 and pos + 8 <= len(raw):
                v = (2**64 - 1) >> int(rng.integers(0, 64)) if rng.integers(0, 2) else int(rng.integers(0, 100000))
                raw[pos:pos + 8] = struct.pack("<Q", v)
            else:
                raw[pos] = 0
        if rng.integers(0, 10) == 0:
            raw = raw[:int(rng.integers(0, len(raw)))]
        q.write_bytes(bytes(raw))
        try:
            s = bgw.InputSet(q, verify=bool(rng.integers(0, 2)))
        except bgw.BgwError as e:
            assert e.code == bgw.BGW_EIO
            continue
        opened += 1
        assert all(k.isascii() and k.isprintable() and " " not in k for k in s.arrays)
        for v in s.arrays.values():
            a = np.atleast_1d(v)
            b = np.frombuffer(a.tobytes(), dtype=np.uint8)
            assert b.size == a.nbytes and b.size <= len(raw)
            assert bytes(b) in bytes(raw)
        s.close()
    assert opened < 400  # the damage was seen at least somewhere


def test_kind_checks_required_entries_and_shapes(tmp_path):
    d = bgw.gpp_inputs(4, 1, 24, 3, 5)
    base = {"number_bands": np.int64(4), "nvband": np.int64(1), "ncouls": np.int64(24), "ngpown": np.int64(3),
            "occ": np.float64(1.0), "inv_igp_index": d["inv"], "indinv": d["indinv"],
            "wx_array": d["wx"].reshape(4, 3), "wtilde_array": d["wtilde"].reshape(3, 24),
            "aqsmtemp": d["aqsm"].reshape(4, 24), "aqsntemp": d["aqsn"].reshape(4, 24),
            "I_eps_array": d["I_eps"].reshape(3, 24), "vcoul": d["vcoul"]}
    p = tmp_path / "ok.bgwin"
    bgw.write_input_set(p, bgw.BGW_INPUTS_GPP, base)
    bgw.InputSet(p).close()
    # the reference form of wx ([3]) is accepted too
    bgw.write_input_set(p, bgw.BGW_INPUTS_GPP, {**base, "wx_array": d["wx"][:3].copy()})
    bgw.InputSet(p).close()
    bad = {
        "missing vcoul": ({k: v for k, v in base.items() if k != "vcoul"}, "lacks 'vcoul'"),
        "aqsm shape": ({**base, "aqsmtemp": d["aqsm"].reshape(24, 4)}, "'aqsmtemp' has the wrong shape"),
        "indinv size": ({**base, "indinv": d["indinv"][:24].copy()}, "'indinv' has the wrong shape"),
        "inv dtype": ({**base, "inv_igp_index": d["inv"].astype(np.int64)}, "'inv_igp_index' has the wrong dtype"),
        "wx shape": ({**base, "wx_array": d["wx"].reshape(12)}, "'wx_array' must be"),
        "scalar": ({**base, "ngpown": np.int64(30)}, "scalars out of range"),
    }
    for name, (arrays, msg) in bad.items():
        q = tmp_path / "bad.bgwin"
        bgw.write_input_set(q, bgw.BGW_INPUTS_GPP, arrays)
        _expect_eio(q, True, msg)
        # the same entries load as an unchecked set
        bgw.write_input_set(q, bgw.BGW_INPUTS_OTHER, arrays)
        bgw.InputSet(q).close()


def test_writer_rejects_bad_descriptors(tmp_path):
    p = tmp_path / "w.bgwin"
    for arrays in ({"x" * 32: np.zeros(2)},             # name too long
                   {"a b": np.zeros(2)},                # names are printable ASCII without spaces
                   {"a\u00e9": np.zeros(2)},
                   {"a": np.zeros(2, np.float32)},      # unsupported dtype
                   {"a": np.zeros((1, 1, 1, 1, 2))}):   # rank above 4
        with pytest.raises((bgw.BgwError, TypeError, ValueError)):
            bgw.write_input_set(p, bgw.BGW_INPUTS_OTHER, arrays)
    L = bgw.lib()
    assert L.bgw_inputs_write(str(p).encode(), 7, None, 0) == bgw.BGW_EINVAL
    d = (bgw._ArrayDesc * 2)()
    for i in range(2):
        d[i].name, d[i].dtype, d[i].ndim, d[i].data = b"same", bgw.BGW_DT_F64, 1, 0
        d[i].dims[0] = 0
    assert L.bgw_inputs_write(str(p).encode(), bgw.BGW_INPUTS_OTHER, d, 2) == bgw.BGW_EINVAL
    assert b"duplicate" in L.bgw_io_last_error()


def test_oracle_sees_identical_inputs_after_a_round_trip(tmp_path):
    import oracle as O

    d = bgw.gpp_inputs(5, 2, 32, 4, 31)
    p = tmp_path / "g.bgwin"
    bgw.save_inputs(p, d, "gpp")
    a, b = O.gpp_forward(d), O.gpp_forward(bgw.load_inputs(p))
    assert np.array_equal(a["achtemp"], b["achtemp"])
