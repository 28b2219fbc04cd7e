"""The autotuner's host logic (no GPU): which launch choices it times for a
shape of domains, the cache format and its lookup."""
from richmol_amd import autotune
from richmol_amd.plan import ell_configs


class _Plan:
    def __init__(self, n):
        self.n = n


def _labels(c):
    return [k + "".join(f" {a}={b}" for a, b in sorted(t.items())) for k, t in c]


def test_candidates_follow_the_domain_shape():
    p = _Plan(1681)
    # 512 pairable domains, fixed Krylov dimension: pairing on and off; one class only, no fork choice
    c = _labels(autotune.candidates(p, 512, 12, (512, 0, 0, 0, 21)))
    assert c[0] == "auto" and "auto pair=0" in c and "auto fork=0" not in c
    assert not any("coop" in x or "wave_max" in x for x in c)
    # adaptive: no pairing choice
    assert "auto pair=0" not in _labels(autotune.candidates(p, 512, 0, (512, 0, 0, 0, 21)))
    # pairable and one-wave domains together: the classes forked or in order
    assert "auto fork=0" in _labels(autotune.candidates(p, 40, 12, (30, 10, 0, 0, 60)))
    # domains of 65-1024 rows: one wave each up to 1024 rows, or the long forms' choices
    c = _labels(autotune.candidates(p, 4, 0, (0, 0, 4, 0, 441)))
    assert "auto wave_max_rows=1024" in c and "auto coop_max=0" in c and "auto coop_min_entries=0" in c
    # one long domain of 3101 rows (49 row groups): one workgroup, or the grid capped below 49
    c = _labels(autotune.candidates(_Plan(12341), 1, 0, (0, 0, 0, 1, 3101)))
    assert "auto coop_max=0" in c and "auto coop_grid=16 coop_min_entries=0" in c
    assert "auto coop_grid=64 coop_min_entries=0" not in c
    # more long domains than the cooperative limit: all of them cooperative, in turn
    assert "auto coop_max=12 coop_min_entries=0" in _labels(autotune.candidates(_Plan(12341), 12, 0, (0, 0, 0, 12, 3101)))
    # full-vector configurations only where they fit N
    n_fit = sum(1 for ntr, r, cc, opl in ell_configs() if 1681 <= ntr * r)
    assert sum(1 for x in _labels(autotune.candidates(p, 512, 12, (512, 0, 0, 0, 21))) if x.startswith("ell:")) == n_fit


def test_cache_lookup_formats(monkeypatch):
    mem = {}
    monkeypatch.setattr(autotune, "_MEM", mem)
    monkeypatch.setattr(autotune, "key", lambda plan, b, m, shape=None: f"k|{b}|{m}" + ("" if shape is None else f"|{shape}"))
    p = _Plan(10)
    assert autotune.lookup(p, 4, 12, (4, 0, 0, 0, 5)) is None
    mem["k|4|12"] = "ell:3"                                   # the older form: a kernel name
    assert autotune.lookup(p, 4, 12, (4, 0, 0, 0, 5)) == ("ell:3", {})
    mem["k|4|12|(4, 0, 0, 0, 5)"] = {"kernel": "auto", "tuning": {"pair": 0}}
    assert autotune.lookup(p, 4, 12, (4, 0, 0, 0, 5)) == ("auto", {"pair": 0})
    assert autotune.lookup(p, 4, 12, (3, 0, 0, 0, 5)) == ("ell:3", {})   # other shape: the shape-free entry
