"""Result files of a saved run (wakeSolver.py:411-416, 705-718, 771-781,
827-835, 1362-1507): names, two-line headers, exact float64/complex128 round
trip through read_txt and load_results, and the header keys read_txt builds."""

import numpy as np

import wakis_amd as wa


def _saved(tmp_path):
    w = wa.WakeSolver(q=1e-9, sigmaz=5e-3, save=True, results_folder=str(tmp_path), verbose=0)
    rng = np.random.default_rng(3)
    s = np.linspace(-0.02, 0.1, 57)
    f = np.linspace(0.0, 3e9, 21)
    vals = dict(s=s, WP=rng.standard_normal(57), WPx=rng.standard_normal(57),
                WPy=rng.standard_normal(57), lambdas=rng.random(57), f=f,
                Z=rng.standard_normal(21) + 1j * rng.standard_normal(21),
                Zx=rng.standard_normal(21) + 1j * rng.standard_normal(21),
                Zy=rng.standard_normal(21) + 1j * rng.standard_normal(21),
                lambdaf=rng.standard_normal(21) + 1j * rng.standard_normal(21))
    w._save_txt("WP.txt", s, vals["WP"], "WP [V/pC]")
    w._save_txt("WPx.txt", s, vals["WPx"], "WP [V/pC]")
    w._save_txt("WPy.txt", s, vals["WPy"], "WP [V/pC]")
    w._save_txt("lambda.txt", s, vals["lambdas"], "Charge distribution [C/m]",
                x_label="s [Hz]")
    w._save_txt("Z.txt", f, vals["Z"], "Z [Ohm]", x_label="f [Hz]")
    w._save_txt("spectrum.txt", f, vals["lambdaf"], "Charge distribution spectrum [C/s]",
                x_label="f [Hz]")
    w._save_txt("Zx.txt", f, vals["Zx"], "Zx [Ohm]", x_label="f [Hz]")
    w._save_txt("Zy.txt", f, vals["Zy"], "Zy [Ohm]", x_label="f [Hz]")
    return w, vals


def test_result_files_round_trip(tmp_path):
    _, vals = _saved(tmp_path)
    r = wa.WakeSolver(save=False, verbose=0)
    r.load_results(str(tmp_path))
    for k, v in vals.items():
        assert np.array_equal(getattr(r, k), v), k
    assert r.wakelength == vals["s"][-1]
    assert r.Ez_file == str(tmp_path) + "/Ez.h5"


def test_headers_and_keys(tmp_path):
    w, vals = _saved(tmp_path)
    with open(tmp_path / "Z.txt") as fh:
        first, second = fh.readline(), fh.readline()
    assert first == "#    f [Hz]" + " " * 20 + "Z [Ohm]\n"
    assert second == "# " + "-" * 48 + "\n"
    d = w.read_txt(str(tmp_path / "WP.txt"))
    assert list(d) == ["s[m]", "WP[V/pC]"]
    d = w.read_txt(str(tmp_path / "spectrum.txt"))
    assert list(d) == ["f[Hz]", "Chargedistributionspectrum[C/s]"]
    assert d["Chargedistributionspectrum[C/s]"].dtype == np.complex128


def test_save_txt_and_integer_keys(tmp_path):
    w = wa.WakeSolver(save=False, verbose=0)
    x, y = np.arange(5.0), np.arange(5.0) ** 2
    w.save_txt(str(tmp_path / "data.txt"), x, y, x_name="t [s]", y_name="A [V]")
    d = w.read_txt(str(tmp_path / "data.txt"))
    assert list(d) == ["t[s]", "A[V]"] and np.array_equal(d["A[V]"], y)
    np.savetxt(tmp_path / "plain.txt", np.c_[x, y, y])
    d = w.read_txt(str(tmp_path / "plain.txt"), skiprows=0)
    assert list(d) == [0, 1, 2] and np.array_equal(d[2], y)
    w.save_txt(str(tmp_path / "none"), None, y)
    assert not (tmp_path / "none.txt").exists()
