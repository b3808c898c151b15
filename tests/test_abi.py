"""CPU tests of libbh.so: it loads, exports every entry point include/bh.h
declares, and its host-only parts (generator, CSV I/O, time parsing, step
counting, argument checks) behave.  No compute call needs a GPU here."""
import math
import pathlib
import re

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HOUR = 1.0 / 24.0


def declared_functions():
    text = (ROOT / "include" / "bh.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(bh):
    lib = bh.load()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_binding_covers_header(bh):
    assert sorted(n for n, _, _ in bh.SIGNATURES) == declared_functions()


def test_generator_is_deterministic(bh):
    a = bh.generate(bh.GEN_BELT, 0, 2000)
    b = bh.generate(bh.GEN_BELT, 0, 2000)
    c = bh.generate(bh.GEN_BELT, 1, 2000)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])


def _elements(pos, vel, mu):
    r = np.linalg.norm(pos, axis=1)
    v2 = (vel * vel).sum(1)
    a = 1.0 / (2.0 / r - v2 / mu)
    h = np.cross(pos, vel)
    e_vec = np.cross(vel, h) / mu - pos / r[:, None]
    inc = np.arccos(np.clip(h[:, 2] / np.linalg.norm(h, axis=1), -1, 1))
    return a, np.linalg.norm(e_vec, axis=1), inc


def test_generator_populations(bh):
    mu = bh.G * bh.CENTRAL_MASS
    pos, vel, m = bh.generate(bh.GEN_BELT, 3, 5009)
    assert np.all(pos[0] == 0) and np.all(vel[0] == 0) and m[0] == bh.CENTRAL_MASS
    ap = np.linalg.norm(pos[1:9], axis=1)
    assert np.allclose(ap, [0.39, 0.72, 1.0, 1.52, 5.2, 9.54, 19.2, 30.1], rtol=1e-12)
    assert np.allclose(np.linalg.norm(vel[1:9], axis=1), np.sqrt(mu / ap), rtol=1e-12)
    assert np.all(pos[1:9, 2] == 0)
    a, e, inc = _elements(pos[9:], vel[9:], mu)
    assert a.min() > 2.1 - 1e-6 and a.max() < 3.3 + 1e-6
    assert e.max() < 0.3 + 1e-6
    assert np.all(m[9:] >= 1e12) and np.all(m[9:] <= 1e20)
    # Rayleigh(5 deg): mean sigma * sqrt(pi/2)
    assert abs(np.degrees(inc).mean() - 5.0 * math.sqrt(math.pi / 2)) < 0.3

    pos, vel, m = bh.generate(bh.GEN_KUIPER, 3, 10009)
    a, e, inc = _elements(pos[9:], vel[9:], mu)
    assert np.sum(a > 29.9) == 3000
    pos, vel, m = bh.generate(bh.GEN_TROJAN, 3, 10009)
    jup = np.arctan2(pos[5, 1], pos[5, 0])
    ang = np.degrees(np.angle(np.exp(1j * (np.arctan2(pos[9:, 1], pos[9:, 0]) - jup))))
    r = np.linalg.norm(pos[9:, :2], axis=1)
    near = np.abs(r - 5.2) < 1.0
    assert np.sum(near & (np.abs(ang - 60) < 10)) > 1400 and np.sum(near & (np.abs(ang + 60) < 10)) > 1400


def test_ensemble_sizes(bh):
    s = bh.ensemble_sizes(0, 64, 10_000, 200_000)
    assert s.min() >= 10_000 and s.max() <= 200_000 and len(np.unique(s)) > 50
    assert np.array_equal(s, bh.ensemble_sizes(0, 64, 10_000, 200_000))


def test_csv_round_trip_and_limits(bh, tmp_path):
    pos, vel, m = bh.generate(bh.GEN_BELT, 9, 300)
    p = tmp_path / "s.csv"
    bh.write_csv(p, pos, vel, m)
    ids, m2, p2, v2 = bh.read_csv(p)
    assert np.array_equal(ids, np.arange(300)) and np.array_equal(m2, m)
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel)
    ids, m3, _, _ = bh.read_csv(p, body_limit=17)
    assert len(m3) == 17
    # malformed rows are skipped and ids follow the parsed rows (src/io.cpp:51-77)
    lines = p.read_text().splitlines()
    lines.insert(3, "x,bad,AST,notanumber,1,2,3,4,5,6")
    lines.insert(6, "y,short,AST,1.0,2.0")
    p.write_text("\n".join(lines) + "\n")
    ids, m4, p4, _ = bh.read_csv(p)
    assert len(m4) == 300 and np.array_equal(p4, pos) and ids[-1] == 299
    with pytest.raises(bh.BHError) as ei:
        bh.read_csv(tmp_path / "missing.csv")
    assert ei.value.code == bh.ERR_IO


def test_reference_csv_format(bh, tmp_path):
    # header + id,name,class,mass,pos,vel rows as in the reference data files
    p = tmp_path / "ref.csv"
    p.write_text("id,name,class,mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z\n"
                 "0,Sun,STA,1.988469999999999977e+30,0,0,0,0,0,0\n"
                 "1,Mercury,PLA,3.3010110000000003e+23,-0.14072807971083467,-0.44390095802703367,"
                 "-0.023345559199712027,0.021168726145230923,-0.0070979267300425707,-0.0025228137246656006\r\n")
    ids, m, pos, vel = bh.read_csv(p)
    assert list(ids) == [0, 1] and m[0] == 1.98847e30 and pos[1, 0] == -0.14072807971083467
    assert vel[1, 2] == -0.0025228137246656006


def test_parse_time_and_step_count(bh):
    assert bh.parse_time("1h") == 1.0 / 24.0
    assert bh.parse_time("2.5d") == 2.5
    assert bh.parse_time("1y") == 365.25
    with pytest.raises(bh.BHError):
        bh.parse_time("5x")
    assert bh.count_steps(HOUR, 100 * HOUR) == 101   # the t < t_end trap
    assert bh.count_steps(bh.parse_time("1h"), bh.parse_time("200h")) == 200
    assert bh.count_steps(HOUR, 168 * HOUR) == 168


def test_create_rejects_let_policy(bh):
    with pytest.raises(bh.BHError) as ei:
        bh.Engine(fc_policy=bh.FC_LET)
    assert ei.value.code == bh.ERR_UNSUPPORTED


def test_create_without_device_fails_cleanly(bh):
    import ctypes as C
    lib = bh.load()
    n = C.c_int(0)
    # only meaningful where no GPU is visible
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is visible")
    with pytest.raises(bh.BHError) as ei:
        bh.Engine()
    assert ei.value.code == bh.ERR_HIP
    assert n.value == 0
