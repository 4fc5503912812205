"""The library's G4EMLOW readers and the table generator's use of them.

The readers follow io/SeltzerBergerReader.cc and io/LivermorePEReader.cc.
They are checked against an independent numpy parse of the same fixture
files (br29 and the pe-*-19 files of the reference's tests), for their
errors (missing files, unset G4LEDATA, Z out of range), and through the
interactor tests, whose reference golden samples (SeltzerBerger.test.cc,
LivermorePE.test.cc) run on tables these readers produce.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import celeritas_amd as amd
from celeritas_amd import workloads as wl
from tests import g4emlow


def test_sb_reader_matches_independent_parse():
    got = amd.read_seltzer_berger(29, g4emlow.DATA)
    ref = g4emlow.read_sb_table(29)
    assert len(got["x"]) == 57 and len(got["y"]) == 32
    assert len(got["values"]) == 57 * 32
    for k in ("x", "y", "values"):
        assert np.array_equal(np.array(got[k]), np.array(ref[k])), k
    # x is log(E/MeV) from 1 keV to 10 GeV; y the reduced photon energy
    assert got["x"][0] == pytest.approx(np.log(1e-3), abs=1e-4)
    assert got["x"][-1] == pytest.approx(np.log(1e4), abs=1e-4)
    assert got["y"][0] == 1e-12 and got["y"][-1] == 1.0


def test_pe_reader_matches_independent_parse():
    got = amd.read_livermore_pe(19, g4emlow.DATA)
    ref = g4emlow.read_livermore_pe(19)
    assert set(got) == set(ref)
    assert got["thresh_lo"] == ref["thresh_lo"] and got["thresh_hi"] == ref["thresh_hi"]
    for k in ("xs_hi", "xs_lo"):
        assert got[k]["x"] == ref[k]["x"] and got[k]["y"] == ref[k]["y"]
    assert len(got["shells"]) == len(ref["shells"]) == 8
    for gs, rs in zip(got["shells"], ref["shells"]):
        assert gs == rs
    # K shell of potassium: 3.5833 keV, the first energy of the subshell table
    assert got["shells"][0]["binding_energy"] == pytest.approx(3.5833e-3)
    assert got["shells"][0]["x"][0] == got["shells"][0]["binding_energy"]


def test_reader_errors(tmp_path, monkeypatch):
    with pytest.raises(amd.CeleritasError, match="failed to open"):
        amd.read_seltzer_berger(30, g4emlow.DATA)
    with pytest.raises(amd.CeleritasError, match="failed to open"):
        amd.read_livermore_pe(20, g4emlow.DATA)
    with pytest.raises(amd.CeleritasError, match="1 <= Z <= 100"):
        amd.read_seltzer_berger(0, g4emlow.DATA)
    monkeypatch.delenv("G4LEDATA", raising=False)
    with pytest.raises(amd.CeleritasError, match="G4LEDATA"):
        amd.read_seltzer_berger(29)
    bad = tmp_path / "br29"
    bad.write_text("4 32 57\n0.1 0.2\n")
    with pytest.raises(amd.CeleritasError, match="truncated"):
        amd.read_seltzer_berger(29, str(tmp_path))


def _g4ledata(tmp_path):
    """A G4LEDATA tree holding the fixture files in the Geant4 layout."""
    sb = tmp_path / "brem_SB"
    pe = tmp_path / "livermore" / "phot_epics2014"
    sb.mkdir(parents=True)
    pe.mkdir(parents=True)
    os.symlink(os.path.join(g4emlow.DATA, "br29"), sb / "br29")
    for f in os.listdir(g4emlow.DATA):
        if f.startswith("pe-"):
            os.symlink(os.path.join(g4emlow.DATA, f), pe / f)
    return tmp_path


def _spec_with(z_list, g4emlow_opt):
    spec = wl.testem3_physics_spec()
    spec = dict(spec)
    spec["elements"] = [dict(spec["elements"][0], z=z) for z in z_list]
    spec["materials"] = [{"label": "m", "density": 1.0,
                          "elements": [{"element": i, "fraction": 1.0}
                                       for i in range(len(z_list))]}]
    spec["cutoffs"] = {k: v[:1] for k, v in spec["cutoffs"].items()}
    if g4emlow_opt is not None:
        spec["g4emlow"] = g4emlow_opt
    return spec


def test_generator_uses_g4emlow_tables(tmp_path, monkeypatch):
    root = _g4ledata(tmp_path)
    monkeypatch.setenv("G4LEDATA", str(root))
    synth = amd.generate_physics(_spec_with([29], None))
    real = amd.generate_physics(_spec_with([29], {"seltzer_berger": True}))
    assert real["elements"][0]["sb_table"] == amd.read_seltzer_berger(29, g4emlow.DATA)
    assert real["elements"][0]["sb_table"] != synth["elements"][0]["sb_table"]
    # The SB-model brems cross sections below 1 GeV follow the file's DCS;
    # above it the relativistic model's are unchanged
    def brems_xs(phys):
        proc = next(p for p in phys["processes"]
                    if p["particle"] == "electron" and p["class"] == "e_brems")
        return proc["xs"][0]
    xs_s, xs_r = brems_xs(synth), brems_xs(real)
    assert xs_s["emin"] == xs_r["emin"] and xs_s["emax"] == xs_r["emax"]
    vs, vr = np.array(xs_s["values"]), np.array(xs_r["values"])
    e = np.exp(np.linspace(np.log(xs_s["emin"]), np.log(xs_s["emax"]), len(vs)))
    assert np.all(vr[(e < 900) & (vs > 0)] > 0)
    assert not np.allclose(vs[e < 900], vr[e < 900], rtol=1e-6)
    assert np.array_equal(vs[e > 1100], vr[e > 1100])

    spec = _spec_with([19], {"livermore_pe": str(root / "livermore" / "phot_epics2014")})
    for el in spec["elements"]:
        el.pop("shells", None)
        el.pop("pe_reference", None)
    pe = amd.generate_physics(spec)
    assert pe["elements"][0]["livermore_pe"] == amd.read_livermore_pe(19, g4emlow.DATA)


def test_generator_g4emlow_missing_element(tmp_path):
    root = _g4ledata(tmp_path)
    with pytest.raises(amd.CeleritasError, match="failed to open"):
        amd.generate_physics(_spec_with([82], {"seltzer_berger": str(root / "brem_SB")}))
