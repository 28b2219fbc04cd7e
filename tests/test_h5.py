"""CarTens HDF5 store/read (richmol/field.py:1304-1848) and state filters
(field.py:187-401, 1859-1883), pinned by the reference's own data:

* tests/golden/water.h5 is the file richmol wrote (h5py) in its documentation
  notebook; reading it must give exactly the matrices that notebook printed
  before storing them (tests/golden/water_h5_notebook.json: the entries of
  the full dipole-x, polarizability-xz and field-free matrices, the energy
  table, the full-form assignment of the first 20 states) and, with an m
  filter, the matrix sizes it printed for every m;
* store -> read of tensors built here returns every Cartesian component
  unchanged, and the file holds what the reference's reader expects (type
  tag, JSON attributes, complex compound data, split sizes);
* the reference's error behaviour (missing name, unknown group, other type,
  existing K/M datasets without replace_k / replace_m);
* the attribute JSON follows json_ext's key rules.
"""
import json
import os

import numpy as np
import pytest

from richmol_amd import field, h5, rot
from richmol_amd.field import CarTens

HERE = os.path.dirname(os.path.abspath(__file__))
WATER = os.path.join(HERE, "golden", "water.h5")
NB = json.load(open(os.path.join(HERE, "golden", "water_h5_notebook.json")))


@pytest.fixture(scope="module")
def water():
    return {n: CarTens(WATER, name=n) for n in ("dip", "pol", "h0")}


def _check_entries(mat, entries):
    for i, j, re_, im_ in entries:
        z = mat[i, j]
        assert z.real == re_ and z.imag == im_, (i, j, z, re_, im_)


def test_reads_the_matrices_the_notebook_printed(water):
    mx = water["dip"].tomat(form="full", cart="x")
    assert list(mx.shape) == NB["shape"]
    _check_entries(mx, NB["dip_x"])
    _check_entries(water["pol"].tomat(form="full", cart="xz"), NB["pol_xz"])
    _check_entries(water["h0"].tomat(form="full", cart="0"), NB["h0"])


def test_energy_table_and_assignment(water):
    h0 = water["h0"]
    for J, sym, i, e in NB["energies"]:
        if e == 0.0:   # the ground state's zero energy leaves an empty block, which is not stored
            assert (J, J) not in h0.kmat
            continue
        assert abs(h0.kmat[(J, J)][(sym, sym)][0][i, i].real - e) < 5e-7
    # full-form order: J, symmetry, then m slowest and k fastest inside a block
    states = [(J, s, m, k) for J in h0.Jlist1 for s in h0.symlist1[J]
              for m in h0.quanta_m1[J][s] for k in h0.quanta_k1[J][s]]
    for i, J, s, m, kq, ke in NB["assign"]:
        assert states[i][:3] == (J, s, m)
        assert states[i][3] == (kq, ke)


def test_m_filter_dimensions(water):
    for m_, n in NB["m_filter_dims"]:
        flt = lambda **kw: kw["m"] == m_ if "m" in kw else True   # noqa: E731
        h = field.filter(water["h0"], bra=flt, ket=flt)
        d = field.filter(water["dip"], bra=flt, ket=flt)
        assert h.tomat(form="full", cart="0").shape == (n, n)
        assert d.tomat(form="full", cart="z").shape == (n, n)


def test_filter_is_the_submatrix(water):
    """Filtered tensor = rows/columns of the kept states of the full one, for
    an m filter on the bra side and a k (energy) filter on the ket side."""
    dip = water["dip"]
    bra = lambda **kw: kw["m"] >= 0 if "m" in kw else True                     # noqa: E731
    ket = lambda **kw: kw["enr"] < 300.0 if "enr" in kw else kw.get("J", 0) != 2  # noqa: E731
    sub = field.filter(dip, bra=bra, ket=ket)
    rows = [i for i, (J, s, m, k) in enumerate(_states(dip, 1)) if m >= 0]
    cols = [i for i, (J, s, m, k) in enumerate(_states(dip, 2)) if k[1] < 300.0 and J != 2]
    for c in dip.cart:
        full = dip.tomat(form="full", cart=c).toarray()
        got = sub.tomat(form="full", cart=c).toarray()
        assert np.array_equal(got, full[np.ix_(rows, cols)])
    # the source is unchanged (filter works on a copy)
    assert dip.tomat(form="full", cart="x").shape == tuple(NB["shape"])


def _states(T, side):
    Jl, sl = getattr(T, f"Jlist{side}"), getattr(T, f"symlist{side}")
    qm, qk = getattr(T, f"quanta_m{side}"), getattr(T, f"quanta_k{side}")
    return [(J, s, m, k) for J in Jl for s in sl[J] for m in qm[J][s] for k in qk[J][s]]


def test_read_with_filters_equals_filter_after_read(water):
    flt = lambda **kw: kw["m"] == 1 if "m" in kw else True   # noqa: E731
    a = CarTens(WATER, name="pol", bra=flt, ket=flt)
    b = field.filter(water["pol"], bra=flt, ket=flt)
    for c in a.cart:
        assert np.array_equal(a.tomat(form="full", cart=c).toarray(), b.tomat(form="full", cart=c).toarray())


def test_read_thresh_drops_small_elements(water):
    t = 0.1
    a = CarTens(WATER, name="dip", thresh=t)
    for Jp, blk in a.kmat.items():
        for sp, d in blk.items():
            for w, k in d.items():
                assert np.all(np.abs(k.data) >= t)
    for Jp, blk in a.mmat.items():
        for sp, d in blk.items():
            for w, mm in d.items():
                for c, m in mm.items():
                    assert np.all(np.abs(m.data) >= t)


def _asym():
    bas = rot.asymtop_basis(9.28, 27.88, 14.51, 3, watson="watson_a", distortion={"DeltaJ": 1e-3})
    return bas, rot.tensor([0.1, -0.3, 0.7], bas), rot.tensor(np.diag([9.1, 9.9, 9.4]) + 0.2, bas), \
        rot.hamiltonian(bas), rot.named("cos2theta", bas)


@pytest.mark.parametrize("which", [1, 2, 3, 4])
def test_store_read_round_trip(tmp_path, which):
    T = _asym()[which]
    fn = tmp_path / "t.h5"
    T.store(fn, name="tens", comment="round  trip\n test")
    R = CarTens(fn, name="tens")
    for c in T.cart:
        assert np.array_equal(R.tomat(form="full", cart=c).toarray(), T.tomat(form="full", cart=c).toarray())
    for a in ("Jlist1", "Jlist2", "symlist1", "symlist2", "dim1", "dim2", "dim_k1", "dim_m1", "quanta_m1",
              "rank", "cart", "os"):
        assert getattr(R, a) == getattr(T, a), a
    assert R.quanta_k1 == {J: {s: [tuple(q) for q in v] for s, v in d.items()} for J, d in T.quanta_k1.items()}
    assert R.__doc__.startswith("Cartesian tensor operator, store date: ")
    assert R.__doc__.endswith(", comment: round trip test")


def test_file_layout_is_the_references(tmp_path):
    _, mu, *_ = _asym()
    fn = tmp_path / "t.h5"
    mu.store(fn, name="mu")
    with h5.File(fn) as f:
        g = f["mu"]
        assert g.attrs["__class_name__"] == "richmol.field.CarTens"
        keys = set(g.attrs.keys())
        assert {"Jlist1", "cart", "os", "rank", "dim1__json", "symlist1__json", "quanta_k1__json"} <= keys
        assert not any(k.startswith("_") and not k.startswith("__") for k in keys)
        assert json.loads(g.attrs["dim1__json"]) == {str(J): d for J, d in mu.dim1.items()}
        assert g.attrs["rank"] == 1 and np.ndim(g.attrs["rank"]) == 0
        gs = g["J:1.0,2.0"]
        s = gs[gs.keys()[0]]
        assert s["kmat_data"].read().dtype == np.complex128
        assert s["kmat_indices"].read().dtype == np.int32
        assert np.asarray(s.attrs["kmat_nnz"]).sum() == len(s["kmat_data"].read())
        ic = np.asarray(s.attrs["mmat_irrep_cart"])
        assert ic.dtype.kind == "S" and ic.shape[1] == 2


def test_errors(tmp_path, water):
    with pytest.raises(ValueError, match="please specify the name"):
        CarTens(WATER)
    with pytest.raises(KeyError, match="has no dataset with the name"):
        CarTens(WATER, name="nope")
    _, mu, _, h, _ = _asym()
    fn = tmp_path / "t.h5"
    mu.store(fn, name="mu")
    with pytest.raises(RuntimeError, match="replace_k"):
        mu.store(fn, name="mu")
    with pytest.raises(RuntimeError, match="replace_m"):
        mu.store(fn, name="mu", store_k=False)
    mu.store(fn, name="mu", replace_k=True, replace_m=True)
    (mu * 2.0).store(fn, name="mu", replace=True)
    R = CarTens(fn, name="mu")
    assert np.allclose(R.tomat(form="full", cart="z").toarray(), 2 * mu.tomat(form="full", cart="z").toarray())
    with h5.File(fn, "a") as f:
        f["mu"].attrs["__class_name__"] = "richmol.rot.Molecule"
    with pytest.raises(TypeError, match="has different type"):
        CarTens(fn, name="mu")
    # a file that is not HDF5 gives an empty tensor, as the reference's base class
    txt = tmp_path / "x.txt"
    txt.write_text("not hdf5")
    assert not hasattr(CarTens(txt, name="x"), "kmat")


def test_store_thresh_and_default_name(tmp_path):
    _, mu, *_ = _asym()
    fn = tmp_path / "t.h5"
    big = 0.05
    mu.store(fn, name="m", thresh=big)
    R = CarTens(fn, name="m")
    for c in mu.cart:
        full = mu.tomat(form="full", cart=c).toarray()
        assert np.abs(R.tomat(form="full", cart=c).toarray()).max() <= np.abs(full).max()
    # K elements under thresh are not written; the tensor itself is unchanged
    n_in = sum(k.nnz for blk in mu.kmat.values() for d in blk.values() for k in d.values())
    n_out = sum(k.nnz for blk in R.kmat.values() for d in blk.values() for k in d.values())
    n_big = sum(int(np.sum(np.abs(k.data) >= big)) for blk in mu.kmat.values() for d in blk.values()
                for k in d.values())
    assert n_out == n_big < n_in
    dipole_of_water = _asym()[1]
    dipole_of_water.store(fn)      # the group is named after the variable
    with h5.File(fn) as f:
        assert "dipole_of_water" in f


def test_json_key_rules():
    val = {0: {"A": [1, 2]}, 2: {"B1": {0.5: (1 + 2j), 1.04: np.arange(3)}}}
    s = h5.json_dumps(val)
    assert json.loads(s) == {"0": {"A": [1, 2]},
                             "2": {"B1": {"0.5": {"__complex__": True, "imag": 2.0, "real": 1.0},
                                          "1.0": {"__ndarray__": True, "val": [0, 1, 2]}}}}
    back = h5.json_loads(s)
    assert back == {0: {"A": [1, 2]}, 2: {"B1": {0.5: 1 + 2j, 1.0: [0, 1, 2]}}}
    assert all(isinstance(k, int) for k in back) and all(isinstance(k, float) for k in back[2]["B1"])
    # keys of one dictionary must sort together, as json_ext's sort_keys requires
    with pytest.raises(TypeError):
        h5.json_dumps({0: 1, "A": 2})


def test_h5_values_round_trip(tmp_path):
    fn = tmp_path / "v.h5"
    vals = {"s": "unicode ✓", "i": 7, "f": 2.5, "c": np.array([1 + 2j, 3 - 4j]), "l": ["a", "bb", "ccc"],
            "n": np.arange(6, dtype=np.int32).reshape(2, 3), "b": np.array([[b"1", b"x"], [b"2", b"yz"]])}
    with h5.File(fn, "w") as f:
        g = f.create_group("g")
        for k, v in vals.items():
            g.attrs[k] = v
        g.create_dataset("d", np.linspace(0, 1, 5))
        g.create_dataset("z", np.array([1j, 2.0]))
    with h5.File(fn) as f:
        g = f["g"]
        assert g.attrs["s"] == "unicode ✓" and g.attrs["i"] == 7 and g.attrs["f"] == 2.5
        assert np.array_equal(g.attrs["c"], vals["c"])
        assert list(g.attrs["l"]) == vals["l"]
        assert np.array_equal(g.attrs["n"], vals["n"]) and g.attrs["n"].dtype == np.int32
        assert np.array_equal(g.attrs["b"], vals["b"])
        assert np.array_equal(g["d"].read(), np.linspace(0, 1, 5))
        assert np.array_equal(g["z"].read(), np.array([1j, 2.0]))
        assert sorted(g.keys()) == ["d", "z"]
    assert h5.is_hdf5(fn) and not h5.is_hdf5(__file__)
