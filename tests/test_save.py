"""GTOSOLVE2 save files (crates/solver/src/save.rs) written and read by the
native library: layout, header, lock validation, legacy and pre-validation
saves, atomic replacement. Mirrors the reference's convergence.rs:524-700,
save_locks.rs, save_atomic.rs and save_compat.rs on the CPU oracle's arenas.
"""
import json
import os
import struct

import numpy as np
import pytest

MAGIC = b"GTOSOLVE2\n"


def _tiny():
    from gtopen_amd import SpotConfig, StreetSizing, TreeConfig

    s = StreetSizing.of
    tree = TreeConfig(starting_pot=100.0, effective_stack=100.0, oop=[s("", ""), s("", ""), s("100", "")],
                      ip=[s("", ""), s("", ""), s("", "")])
    return SpotConfig("QcJc9c3d2d", "AcKc,8h7h", "QdQh", tree)


def _solved(cfg, iters, storage="f32"):
    import oracle
    from gtopen_amd import Spot

    spot = Spot(cfg)
    o = oracle.Oracle(spot, storage=storage)
    o.iterate(iters)
    return spot, o


def _save(tmp_path, spot, o, name="spot.gto", locks=None, labels=None):
    from gtopen_amd.save import write_save

    path = str(tmp_path / name)
    r, s = o.arenas()
    write_save(path, spot, o.iteration, r, s, locks=locks, labels=labels)
    return path


def _split(path):
    b = open(path, "rb").read()
    nl = 10 + b[10:].index(b"\n")
    return b[:10], b[10:nl], b[nl + 1:]


def _doctor(path, fn):
    magic, hdr, rest = _split(path)
    h = json.loads(hdr)
    fn(h)
    with open(path, "wb") as f:
        f.write(magic + json.dumps(h).encode() + b"\n" + rest)


def _err(fn):
    from gtopen_amd import GtoError

    with pytest.raises(GtoError) as e:
        fn()
    return str(e.value)


def test_layout_and_roundtrip(built, tmp_path):
    import oracle
    from gtopen_amd.save import SaveFile

    spot, o = _solved(_tiny(), 200)
    root_sigma = np.tile(np.array([[0.7], [0.3]], np.float32), (1, spot.num_hands[0])).reshape(-1)
    path = _save(tmp_path, spot, o, locks={0: root_sigma}, labels={0: "root 30% bet"})
    assert not os.path.exists(path + ".tmp")
    magic, hdr, rest = _split(path)
    assert magic == MAGIC
    h = json.loads(hdr)
    assert h["iteration"] == 200
    t = h["config"]["tree"]
    # serde layout of TreeConfig (tree.rs:70-95): every field present
    assert set(t) == {"starting_pot", "effective_stack", "rake_pct", "rake_cap", "oop", "ip", "allin_threshold",
                      "add_allin", "max_raises"}
    assert t["oop"][2]["bet"] == [{"PotPct": 100.0}] and t["max_raises"] == 10
    assert h["locks"] == [[0, [0.7, 0.3]]] or h["locks"][0][0] == 0
    assert np.array_equal(np.array(h["locks"][0][1], np.float32), root_sigma)
    assert h["lock_labels"] == [[0, "root 30% bet"]]
    # four (u64 LE length, f32 LE data) arenas: R0, R1, S0, S1
    r, s = o.arenas()
    pos = 0
    for a in (r[0], r[1], s[0], s[1]):
        (n,) = struct.unpack_from("<Q", rest, pos)
        pos += 8
        assert n == a.size
        assert np.array_equal(np.frombuffer(rest, dtype="<f4", count=n, offset=pos), a)
        pos += 4 * n
    assert pos == len(rest)
    # read back through the library: a rebuilt spot, identical arenas
    sf = SaveFile(path)
    assert sf.iteration == 200 and sf.labels == {0: "root 30% bet"}
    assert np.array_equal(sf.locks[0], root_sigma)
    spot2 = sf.spot()
    assert spot2.num_nodes == spot.num_nodes and spot2.data_size == spot.data_size
    sf.validate_locks(spot2)
    r2, s2 = sf.read_arenas(spot2)
    for a, b in zip(r + s, r2 + s2):
        assert np.array_equal(a, b)
    o2 = oracle.Oracle(spot2)
    o2.set_arenas(r2, s2)
    o2.iteration = sf.iteration
    assert abs(o2.exploitability() - o.exploitability()) < 1e-9


def test_compressed_roundtrip_is_exact(built, tmp_path):
    """convergence.rs:524-556: compressed -> file (f32) -> compressed keeps
    the exploitability exactly; -> f32 within 1e-4."""
    import oracle
    from gtopen_amd import workloads
    from gtopen_amd.save import SaveFile

    spot, o = _solved(workloads.river_spot(), 50, storage="compressed")
    e0 = o.exploitability()
    path = _save(tmp_path, spot, o)
    sf = SaveFile(path)
    r, s = sf.read_arenas(spot)
    oc = oracle.Oracle(spot, storage="compressed")
    oc.set_arenas(r, s)
    assert abs(oc.exploitability() - e0) < 1e-9
    of = oracle.Oracle(spot)
    of.set_arenas(r, s)
    assert abs(of.exploitability() - e0) < 1e-4


def test_legacy_header_without_locks_loads(built, tmp_path):
    from gtopen_amd.save import SaveFile

    spot, o = _solved(_tiny(), 50)
    path = _save(tmp_path, spot, o, locks={0: np.full(2 * spot.num_hands[0], 0.5, np.float32)}, labels={0: "x"})

    def strip(h):
        del h["locks"]
        del h["lock_labels"]

    _doctor(path, strip)
    sf = SaveFile(path)
    assert sf.locks == {} and sf.labels == {} and sf.iteration == 50
    r, _ = sf.read_arenas(sf.spot())
    assert np.array_equal(r[0], o.arenas()[0][0])


@pytest.mark.parametrize("case", ["oob", "nonaction", "short", "negative"])
def test_malformed_locks_are_rejected(built, tmp_path, case):
    """save_locks.rs:95-174: refused at load with a clear error."""
    from gtopen_amd.save import SaveFile

    spot, o = _solved(_tiny(), 50)
    path = _save(tmp_path, spot, o)
    if case == "oob":
        locks, want = [[999_999, [0.25, 0.25, 0.25, 0.25]]], ("lock", "out of range")
    elif case == "nonaction":
        idx = int(np.nonzero(spot.kind != 0)[0][0])
        locks, want = [[idx, [0.5, 0.5]]], ("not an action node",)
    elif case == "short":
        locks, want = [[0, [0.5]]], ("expected", "lock")
    else:
        sig = [0.5] * (int(spot.num_children[0]) * spot.num_hands[int(spot.player[0])])
        sig[0] = -1.0
        locks, want = [[0, sig]], ("invalid",)

    def put(h):
        h["locks"] = locks

    _doctor(path, put)
    sf = SaveFile(path)
    msg = _err(lambda: sf.validate_locks(sf.spot()))
    assert "bad lock section" in msg and all(w in msg for w in want), msg


def test_bad_magic_and_arena_mismatch(built, tmp_path):
    from gtopen_amd.save import SaveFile

    spot, o = _solved(_tiny(), 20)
    path = _save(tmp_path, spot, o)
    bad = str(tmp_path / "bad.gto")
    with open(bad, "wb") as f:
        f.write(b"GTOSOLVE1\n{}\n")
    assert "not a solver save file" in _err(lambda: SaveFile(bad))

    def bigger(h):
        h["config"]["tree"]["oop"][2]["bet"] = [{"PotPct": 50.0}, {"PotPct": 100.0}]

    _doctor(path, bigger)
    sf = SaveFile(path)
    msg = _err(lambda: sf.read_arenas(sf.spot()))
    assert "arena size mismatch" in msg and "tree config changed" in msg, msg
    with open(bad, "wb") as f:
        f.write(MAGIC + b"{not json}\n")
    assert "bad header" in _err(lambda: SaveFile(bad))


def test_failed_save_leaves_previous_save_intact(built, tmp_path):
    """save_atomic.rs: a blocked staging path fails the save and the old
    file stays byte-identical."""
    from gtopen_amd import GtoError

    spot, o = _solved(_tiny(), 20)
    path = _save(tmp_path, spot, o)
    before = open(path, "rb").read()
    os.mkdir(path + ".tmp")
    o.iterate(10)
    with pytest.raises(GtoError):
        _save(tmp_path, spot, o)
    assert open(path, "rb").read() == before
    os.rmdir(path + ".tmp")


def test_pre_validation_save_loads_leniently(built, tmp_path):
    """save_compat.rs: a stored config holding a raise multiple in a bet list
    is rejected by the strict builder and rebuilt leniently by load."""
    import oracle
    from gtopen_amd import GtoError, Spot, SpotConfig, StreetSizing, TreeConfig
    from gtopen_amd.save import SaveFile

    s = StreetSizing.of
    tree = TreeConfig(starting_pot=60.0, effective_stack=200.0, oop=[s("50", "60"), s("50", ""), s("50", "")],
                      ip=[s("50", "60"), s("50", ""), s("50", "")])
    spot, o = _solved(SpotConfig("Th9h2c", "QQ,JJ,TT,99", "AA,KK", tree), 20)
    e0 = o.exploitability()
    path = _save(tmp_path, spot, o)
    _doctor(path, lambda h: h["config"]["tree"]["oop"][0]["bet"].insert(0, {"PrevMult": 2.5}))
    sf = SaveFile(path)
    with pytest.raises(GtoError) as e:
        Spot(sf.config_json)
    assert "2.5x" in str(e.value)
    spot2 = sf.spot()
    r, st = sf.read_arenas(spot2)
    o2 = oracle.Oracle(spot2)
    o2.set_arenas(r, st)
    assert abs(o2.exploitability() - e0) < 1e-9
