""".nsys snapshots: the engine's reader and writer against the reference.

Pins (CPU):
  * reference-written fixtures (tests/golden/nsys_*.nsys, written by the
    compiled reference's to_nsys, see make_nsys_fixtures.py) parsed by the
    oracle's independent layout restatement give back the bodies the
    fixtures were made from (t0) and the oracle's long-double KDK state after
    3 steps (t3, itself bit-exact to the reference), time 30 s, dt 10 s and
    the G ratios of "M3:2,D5:7,T11:13";
  * nb_nsys_read returns exactly those values;
  * nb_nsys_write reproduces the reference's bytes (x87 padding bytes masked)
    for the same state, in both value types;
  * the compiled reference reads the engine's files and writes them back
    unchanged (when oracle/_ref is built);
  * load_from_nsys's error paths and messages.
Pins (GPU): nb_save_nsys / nb_load_nsys round trips through a context, and
the time elapsed bookkeeping of evolve (iterations*dt) and advance.
"""
import os
import struct

import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

GOLD = os.path.join(os.path.dirname(__file__), "golden")
G_SI = 6.6743e-11
LD = np.longdouble


def fixture(name):
    return os.path.join(GOLD, name)


def bodies():
    m, pos, vel = capi.ic_uniform_ball(41, 5, 1e20, 1e22, 1e9, 1e3)
    r = 1e3 * np.arange(1, 6, dtype=np.float64)
    e = 0.5 + 0.1 * np.arange(5, dtype=np.float64)
    return m, r, e, pos, vel


def ld_mask(raw, chunk, ld_offsets):
    """Zero the 6 padding bytes of every 16-byte x87 field (the reference
    copies whatever its long double objects hold there)."""
    b = bytearray(raw)
    for off in (4 + 10, 20 + 10):
        b[off:off + 6] = bytes(6)
    n = (len(raw) - 96) // chunk
    for i in range(n):
        for o in ld_offsets:
            p = 96 + i * chunk + o + 10
            b[p:p + 6] = bytes(6)
    return bytes(b)


def test_reference_fixtures_match_inputs_and_oracle():
    m, r, e, pos, vel = bodies()
    for name in ("nsys_ld_n5_t0.nsys", "nsys_d_n5_t0.nsys"):
        d = O.parse_nsys(fixture(name))
        assert d["signature"] == b"NS" and d["d_types"] == 15 and d["time_size"] == 16
        assert d["file_size"] == d["expected_size"]
        assert list(d["g_vals"]) == [2, 3, 5, 7, 13, 11]
        assert d["time_elapsed"] == 0 and d["dt"] == 10
        assert np.array_equal(d["m"].astype(np.float64), m) and np.array_equal(d["r"].astype(np.float64), r)
        assert np.array_equal(d["pos"].astype(np.float64), pos) and np.array_equal(d["vel"].astype(np.float64), vel)
        assert np.array_equal(d["rest_c"].astype(np.float64), e)
    d = O.parse_nsys(fixture("nsys_ld_n5_t3.nsys"))
    assert d["time_elapsed"] == 30 and d["dt"] == 10
    # units "M3:2,D5:7,T11:13": G = 66743 * 2*5^3*13^2 / (3*7^3*11^2) / 1e15
    G = LD(66743) * LD(2 * 5 ** 3 * 13 ** 2) / LD(3 * 7 ** 3 * 11 ** 2) / LD(10 ** 15)
    ref = O.kdk(m, pos, vel, G, 10.0, 3)
    assert np.array_equal(d["pos"], ref["pos"]) and np.array_equal(d["vel"], ref["vel"])


@pytest.mark.parametrize("name", ["nsys_ld_n5_t0.nsys", "nsys_ld_n5_t3.nsys", "nsys_d_n5_t0.nsys"])
def test_engine_reader_matches_layout_restatement(name):
    d = O.parse_nsys(fixture(name))
    got = capi.nsys_read(fixture(name))
    info = got["info"]
    assert (info.d_types, info.time_size, info.m_size, info.r_size, info.t_size, info.rest_size) == \
        (d["d_types"], d["time_size"], d["m_size"], d["r_size"], d["t_size"], d["rest_size"])
    assert list(info.g_vals) == list(d["g_vals"]) and info.num_bodies == d["num_bodies"]
    assert info.time_elapsed == float(d["time_elapsed"]) and info.dt == float(d["dt"])
    for k in ("m", "r", "pos", "vel", "rest_c"):
        assert np.array_equal(got[k], d[k].astype(np.float64)), k


@pytest.mark.parametrize("value_size,name", [(16, "nsys_ld_n5_t0.nsys"), (8, "nsys_d_n5_t0.nsys")])
def test_engine_writer_reproduces_reference_bytes(tmp_path, value_size, name):
    m, r, e, pos, vel = bodies()
    out = str(tmp_path / "w.nsys")
    size = capi.nsys_write(out, m, r, pos, vel, e, 0.0, 10.0, (2, 3, 5, 7, 13, 11), value_size)
    ours, theirs = open(out, "rb").read(), open(fixture(name), "rb").read()
    assert size == len(ours) == len(theirs)
    chunk = O.parse_nsys(out)["chunk"]
    lds = [0] + ([16 + 16 * k for k in range(8)] if value_size == 16 else [])
    assert ld_mask(ours, chunk, lds) == ld_mask(theirs, chunk, lds)


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference (oracle/_ref) not built")
@pytest.mark.parametrize("value_size,vt", [(16, "ld"), (8, "d")])
def test_reference_reads_engine_files(tmp_path, value_size, vt):
    rng = np.random.default_rng(7)
    n = 37
    m, r, e = rng.uniform(1, 1e22, n), rng.uniform(0, 1e5, n), rng.uniform(0, 1, n)
    pos, vel = rng.normal(0, 1e9, (n, 3)), rng.normal(0, 1e3, (n, 3))
    ours, back = str(tmp_path / "ours.nsys"), str(tmp_path / "back.nsys")
    capi.nsys_write(ours, m, r, pos, vel, e, 1234.5, 0.25, (1, 2, 3, 4, 5, 6), value_size)
    O.ref_nsys_copy(ours, back, vt)
    a, b = O.parse_nsys(ours), O.parse_nsys(back)
    for k in ("time_elapsed", "dt", "num_bodies"):
        assert a[k] == b[k]
    assert list(a["g_vals"]) == list(b["g_vals"])
    for k in ("m", "r", "pos", "vel", "rest_c"):
        assert np.array_equal(a[k], b[k]), k


def write_header(path, body_bytes=b"", **kw):
    h = dict(sig=b"NS", d_types=15, time_size=16, time=0.0, dt=1.0, sizes=(8, 8, 8), rest=16,
             g=(1, 1, 1, 1, 1, 1), n=0)
    h.update(kw)
    b = bytearray(96)
    b[0:2] = h["sig"]
    b[2], b[3] = h["d_types"], h["time_size"]
    b[4:4 + 16] = np.array([h["time"]], LD).tobytes()
    b[20:20 + 16] = np.array([h["dt"]], LD).tobytes()
    b[36], b[37], b[38] = h["sizes"]
    b[39] = h["rest"]
    b[40:88] = struct.pack("<6Q", *h["g"])
    b[88:96] = struct.pack("<Q", h["n"])
    with open(path, "wb") as f:
        f.write(bytes(b) + body_bytes)


def expect(code, text, path, check=True):
    with pytest.raises(capi.NBodyError) as ei:
        capi.nsys_read(path, check)
    assert ei.value.code == code and text in str(ei.value), str(ei.value)


def test_load_errors_mirror_reference(tmp_path):
    p = str(tmp_path / "x.nsys")
    expect(capi.NB_ERR_NSYS_LOAD, "Error obtaining .nsys file information", str(tmp_path / "missing.nsys"))
    expect(capi.NB_ERR_NSYS_LOAD, "not a regular file", str(tmp_path))
    open(p, "wb").write(b"NS" + bytes(10))
    expect(capi.NB_ERR_NSYS_LOAD, "Insufficient .nsys file size", p)
    write_header(p, sig=b"XY")
    expect(capi.NB_ERR_NSYS_FORMAT, "invalid file signature", p)
    write_header(p, sig=b"NY")  # the reference only rejects when both bytes differ
    assert capi.nsys_read(p)["info"].num_bodies == 0
    write_header(p, d_types=15 | 64)
    expect(capi.NB_ERR_NSYS_FORMAT, "msb occupied in d_types field", p)
    write_header(p, rest=32)
    expect(capi.NB_ERR_NSYS_FORMAT, "larger than widest floating point data type", p)
    write_header(p, rest=2)
    expect(capi.NB_ERR_NSYS_FORMAT, "does not match any floating point data type", p)
    write_header(p, n=3)
    expect(capi.NB_ERR_NSYS_FORMAT, "file size does not match expected file size", p)
    write_header(p, time_size=2)
    expect(capi.NB_ERR_NSYS_FORMAT, "reported size of data type for time values", p)
    write_header(p, dt=float("nan"))
    expect(capi.NB_ERR_NSYS_FORMAT, "time-step is NaN or INF", p)
    write_header(p, dt=0.0)
    expect(capi.NB_ERR_INVALID_ARGUMENT, "Time-step cannot be zero or negative", p)
    # a NaN in body 2's z velocity: index and field name in the message, or
    # accepted with the check off
    m, r, e, pos, vel = bodies()
    capi.nsys_write(p, m, r, pos, vel, e, 0.0, 1.0, (1,) * 6, 8)
    raw = bytearray(open(p, "rb").read())
    struct.pack_into("<d", raw, 96 + 2 * 80 + 16 + 8 * 7, float("nan"))
    open(p, "wb").write(bytes(raw))
    expect(capi.NB_ERR_NSYS_LOAD, 'chunk with index 2. Invalid field: "zvel"', p)
    assert np.isnan(capi.nsys_read(p, check_nan_inf=False)["vel"][2, 2])


def test_integral_and_mixed_types_read(tmp_path):
    """Types the header can describe besides the writer's: float M, unsigned
    16-bit R, double T (natural alignment: mass at 16, radius at 20, T at
    24, chunk 72 bytes), with an integral time of 4 bytes."""
    p = str(tmp_path / "mix.nsys")
    n = 2
    body = b""
    for i in range(n):
        c = bytearray(72)
        c[0:8] = struct.pack("<d", 0.25 * (i + 1))
        struct.pack_into("<f", c, 16, 1.5 * (i + 1))
        struct.pack_into("<H", c, 20, 7 + i)
        struct.pack_into("<6d", c, 24, *[float(i * 10 + k) for k in range(6)])
        body += bytes(c)
    b = bytearray(96)
    b[0:2] = b"NS"
    b[2] = 0b1010  # integral time, floating M and T, unsigned R
    b[3] = 4
    struct.pack_into("<I", b, 4, 600)
    struct.pack_into("<I", b, 20, 3)
    b[36], b[37], b[38], b[39] = 4, 2, 8, 8
    struct.pack_into("<6Q", b, 40, 1, 1, 1, 1, 1, 1)
    struct.pack_into("<Q", b, 88, n)
    open(p, "wb").write(bytes(b) + body)
    d = capi.nsys_read(p)
    assert d["info"].time_elapsed == 600 and d["info"].dt == 3
    assert list(d["m"]) == [1.5, 3.0] and list(d["r"]) == [7, 8] and list(d["rest_c"]) == [0.25, 0.5]
    assert np.array_equal(d["pos"][1], [10, 11, 12]) and np.array_equal(d["vel"][1], [13, 14, 15])


@pytest.mark.gpu
def test_context_save_load_round_trip(tmp_path):
    print("case nsys context round trip", flush=True)
    m, pos, vel = capi.ic_uniform_ball(42, 300, 1e20, 1e22, 1e9, 1e3)
    s = capi.System(0)
    s.set_units("M3:2,D5:7,T11:13")
    s.set_timestep(10.0)
    s.set_softening(1e6)
    s.set_iterations(20)
    s.set_bodies(m, pos, vel)
    s.evolve(capi.LEAPFROG_KDK)
    assert s.time_elapsed() == 200.0
    s.advance(5)
    assert s.time_elapsed() == 250.0
    for vs in (8, 16):
        p = str(tmp_path / f"s{vs}.nsys")
        size = s.save_nsys(p, vs)
        assert size == os.path.getsize(p) == 96 + 300 * (80 if vs == 8 else 144)
        t = capi.System(0)
        t.load_nsys(p)
        a, b = s.state(), t.state()
        for k in ("m", "r", "pos", "vel"):
            assert np.array_equal(a[k], b[k]), k
        assert t.time_elapsed() == 250.0 and t.G == s.G
        # the loaded context continues as the original does
        t.set_softening(1e6)
        t.set_iterations(7)
        s2 = capi.System(0)
        s2.load_nsys(p)
        s2.set_softening(1e6)
        s2.set_iterations(7)
        t.evolve(capi.LEAPFROG_KDK)
        s2.evolve(capi.LEAPFROG_KDK)
        assert np.array_equal(t.state()["pos"], s2.state()["pos"])
    s.load_nsys(fixture("nsys_ld_n5_t3.nsys"))
    assert s.time_elapsed() == 30.0
