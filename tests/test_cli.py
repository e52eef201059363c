"""The ms-stem-fem command line: flags, parameter-file keys, required
parameters, defaults and derived settings (tem_io.cpp:4063-4500,
:5277-5700, :5858-6060; tem.cpp:1320-1480), checked through --plan_only,
which stops before any device work."""
import os
import subprocess

import numpy as np
import pytest

import msf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ms-stem-fem_amd", "ms-stem-fem")

MODEL = "4\n20 20 10\nSi 1 1 1\nSi 5 5 5\nCu 9 2 3\nZr 3 8 9\n"


@pytest.fixture
def model(tmp_path):
    p = tmp_path / "m.xyz"
    p.write_text(MODEL)
    return str(p)


def run(args, cwd):
    r = subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=60)
    plan = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(" ")
        plan.setdefault(k, []).append(v)
    return r.returncode, plan, r.stdout + r.stderr


def test_uncorrected_fem_plan(tmp_path, model):
    rc, plan, out = run(["-m", "128", "128", "200000", "-a", model, "--output_prefix", "o",
                         "--adfstemuncorrfem", "--cs3", "1.2e7", "--scherzer_defocus",
                         "--scherzer_alphamax", "--rasterspacing", "1.25", "--minslice", "2",
                         "--dr", "0.1414213562", "--plan_only"], tmp_path)
    assert rc == 0, out
    lam = msf.wavelength(200e3)[0]
    mode = plan["mode"][0].split()
    assert mode[0] == "adfstem_uncorr"
    assert float(mode[6]) == msf.scherzer_defocus(lam, 1.2e7)
    assert float(mode[8]) == msf.scherzer_alphamax(lam, 1.2e7)
    assert plan["raster"][0].split()[:2] == ["16", "16"]          # floor(20 / 1.25)
    assert [float(x) for x in plan["detector"][0].split()] == [0.040, 0.200]  # >= 200 kV
    assert plan["fem"] == ["1"] and plan["images"] == ["0"]
    assert float(plan["minslice"][0]) == 2.0 and float(plan["dr"][0]) == 0.1414213562


def test_parameter_file_keys_and_precedence(tmp_path, model):
    pf = tmp_path / "p.txt"
    pf.write_text(f"SAMPLES_X 256\nsamples_y 192\nmicroscope_voltage 100000\nmodel_file {model}\n"
                  "output_prefix pf\n\nadfstemcorr\ncs5 5e7\nscherzer_cs3\nscherzer_defocus\n"
                  "alphamax 0.02\ndetectorangles 0.3 0.05\nraster_spacing 2.5\nimages\n")
    rc, plan, out = run(["--parameter_file", str(pf), "--alphamax", "0.015", "--plan_only"], tmp_path)
    assert rc == 0, out
    lam = msf.wavelength(100e3)[0]
    mode = plan["mode"][0].split()
    assert mode[0] == "adfstem_corr"
    assert float(mode[2]) == msf.scherzer_cs3(lam, 5e7)
    assert float(mode[6]) == msf.scherzer_defocus(lam, 5e7, 1)
    assert float(mode[8]) == 0.015                     # the later command-line value wins
    assert [float(x) for x in plan["detector"][0].split()] == [0.05, 0.3]  # swapped
    assert plan["grid"][0].split()[:2] == ["256", "192"]
    assert plan["raster"][0].split()[:2] == ["8", "8"]
    assert plan["images"] == ["1"] and plan["fem"] == ["0"]
    # a command-line value given before the parameter file is overridden by it
    rc, plan, _ = run(["--alphamax", "0.015", "--parameter_file", str(pf), "--plan_only"], tmp_path)
    assert float(plan["mode"][0].split()[8]) == 0.02


def test_defaults(tmp_path, model):
    base = ["-m", "64", "64", "100000", "-a", model, "--output_prefix", "o", "--cs3", "1e6",
            "--defocus", "10", "--alphamax", "0.01", "--plan_only"]
    rc, plan, out = run(base + ["--adfstemuncorr"], tmp_path)
    assert rc == 0, out
    assert [float(x) for x in plan["detector"][0].split()] == [0.045, 0.200]  # below 200 kV
    assert float(plan["raster"][0].split()[2]) == 1.5
    assert float(plan["minslice"][0]) == 1.0
    assert float(plan["dr"][0]) == 0.70710678118654746
    rc, plan, _ = run(base + ["--adfstemuncorrfem"], tmp_path)
    assert float(plan["raster"][0].split()[2]) == 10.0   # FEM default spacing
    assert plan["raster"][0].split()[:2] == ["2", "2"]


def test_dupe_replicates_for_fem_only(tmp_path, model):
    base = ["-m", "64", "64", "300000", "-a", model, "--output_prefix", "o", "--cs3", "0",
            "--defocus", "0", "--alphamax", "0.01", "--dupe", "2", "1", "3", "--rasterspacing",
            "5", "--plan_only"]
    rc, plan, out = run(base + ["--adfstemuncorrfem"], tmp_path)
    assert rc == 0, out
    assert plan["atoms"] == ["24"]
    g = plan["grid"][0].split()
    assert [float(x) for x in g[2:]] == [40.0, 20.0, 30.0]
    assert plan["raster"][0].split()[:2] == ["4", "4"]     # raster of the undupled cell
    rc, plan, _ = run(base + ["--adfstemuncorr"], tmp_path)
    assert plan["atoms"] == ["4"]


@pytest.mark.parametrize("args, why", [
    (["-a", "MODEL", "--output_prefix", "o", "--adfstemuncorr", "--cs3", "0", "--defocus", "0",
      "--alphamax", "0.01"], "no -m"),
    (["-m", "64", "64", "3e5", "-a", "MODEL", "--adfstemuncorr", "--cs3", "0", "--defocus", "0",
      "--alphamax", "0.01"], "no output prefix"),
    (["-m", "64", "64", "3e5", "-a", "MODEL", "--output_prefix", "o", "--adfstemuncorr",
      "--defocus", "0", "--alphamax", "0.01"], "uncorrected without cs3"),
    (["-m", "64", "64", "3e5", "-a", "MODEL", "--output_prefix", "o", "--adfstemcorr",
      "--cs3", "0", "--defocus", "0", "--alphamax", "0.01"], "corrected without cs5"),
    (["-m", "64", "64", "3e5", "-a", "MODEL", "--output_prefix", "o", "--adfstemuncorr",
      "--cs3", "0", "--alphamax", "0.01"], "no defocus"),
    (["-m", "64", "64", "3e5", "-a", "MODEL", "--output_prefix", "o", "--adfstemuncorr",
      "--cs3", "0", "--defocus", "0", "--alphamax", "0.01", "--bogus"], "unknown flag"),
    (["-m", "64", "64", "3e5", "-a", "MODEL", "--output_prefix", "o", "--adfstemuncorrfem",
      "--cs3", "0", "--defocus", "0", "--alphamax", "0.01", "--paptif"],
     "mode outside the build"),
    (["-m", "64", "64", "3e5", "-a", "missing.xyz", "--output_prefix", "o", "--adfstemuncorr",
      "--cs3", "0", "--defocus", "0", "--alphamax", "0.01"], "unreadable model"),
    (["-m", "64", "64", "3e5", "-a", "MODEL", "--output_prefix", "o", "--adfstemuncorr",
      "--cs3", "0", "--defocus", "0", "--alphamax", "0.01", "--detectorangles", "2.0", "3.0"],
     "inner angle beyond the band limit"),
])
def test_failures(tmp_path, model, args, why):
    args = [model if a == "MODEL" else a for a in args] + ["--plan_only"]
    rc, _, out = run(args, tmp_path)
    assert rc == 1, (why, out)


def test_parameter_file_rejects_unknown_key_and_drops_unterminated_last_line(tmp_path, model):
    pf = tmp_path / "p.txt"
    pf.write_text(f"samples_x 64\nsamples_y 64\nmicroscope_voltage 3e5\nmodel_file {model}\n"
                  "output_prefix o\nadfstemuncorr\ncs3 0\ndefocus 0\nalphamax 0.01\nfrobnicate 1\n")
    rc, _, out = run(["--parameter_file", str(pf), "--plan_only"], tmp_path)
    assert rc == 1 and "unexpected argument in parameter file" in out
    # the reference's read loop stops at a line without a newline: "alphamax"
    # is never read, so the run lacks a required parameter
    pf.write_text(f"samples_x 64\nsamples_y 64\nmicroscope_voltage 3e5\nmodel_file {model}\n"
                  "output_prefix o\nadfstemuncorr\ncs3 0\ndefocus 0\nalphamax 0.01")
    rc, _, out = run(["--parameter_file", str(pf), "--plan_only"], tmp_path)
    assert rc == 1 and "lacking required parameters" in out


def test_fem_mode_selection(tmp_path, model):
    base = ["-m", "64", "64", "300000", "-a", model, "--output_prefix", "o", "--cs3", "0",
            "--defocus", "0", "--alphamax", "0.01", "--plan_only"]
    # V_omega is the default only when no FEM mode is named (tem_io.cpp:5705-5712)
    rc, plan, out = run(base + ["--adfstemuncorrfem"], tmp_path)
    assert rc == 0, out
    assert plan["fem_modes"] == ["V_omega"]
    rc, plan, _ = run(base + ["--adfstemuncorrfem", "--V_re", "--V_gt17"], tmp_path)
    assert plan["fem_modes"][0].split() == ["V_re", "V_gt17"]
    # the mode flags do nothing without an FEM run mode
    rc, plan, _ = run(base + ["--adfstemuncorr", "--Vbar_r"], tmp_path)
    assert plan["fem_modes"][0].split() == []


def test_correlograph_flags(tmp_path, model):
    base = ["-m", "64", "64", "300000", "-a", model, "--output_prefix", "o", "--cs3", "0",
            "--defocus", "0", "--alphamax", "0.01", "--plan_only", "--adfstemuncorrfem"]
    # naming only the correlograph drops the V_omega default (tem_io.cpp:5705-5712)
    rc, plan, out = run(base + ["--correlograph"], tmp_path)
    assert rc == 0, out
    assert plan["fem_modes"][0].split() == ["correlograph"]
    # the variance flag implies the correlograph (tem_io.cpp:5646-5650)
    rc, plan, _ = run(base + ["--correlograph_variance", "--V_re"], tmp_path)
    assert plan["fem_modes"][0].split() == ["V_re", "correlograph", "correlograph_variance"]
    pf = tmp_path / "p.txt"
    pf.write_text("correlograph_variance\n")
    rc, plan, _ = run(["--parameter_file", str(pf)] + base, tmp_path)
    assert plan["fem_modes"][0].split() == ["correlograph", "correlograph_variance"]
    for flag in ("--correlograph_everyimage", "--correlograph_everytxt"):
        rc, plan, out = run(base + [flag], tmp_path)
        assert rc == 0 and plan["fem_modes"][0].split() == ["correlograph"], flag


def test_mtf_flags(tmp_path, model):
    (tmp_path / "mtf.txt").write_text("0, 0.5, 1\n1, 0.5, 0.25\n")
    base = ["-m", "64", "64", "300000", "-a", model, "--output_prefix", "o", "--cs3", "0",
            "--defocus", "0", "--alphamax", "0.01", "--plan_only", "--adfstemuncorrfem"]
    rc, plan, out = run(base + ["--mtf_file", "mtf.txt", "--mtf_sampling_rate", "8"], tmp_path)
    assert rc == 0, out
    assert plan["mtf"] == ["3 8"]
    # a file without a sampling rate (or with a rate <= 0) lacks a parameter
    rc, _, out = run(base + ["--mtf_file", "mtf.txt"], tmp_path)
    assert rc == 1 and "lacking required parameters" in out
    rc, _, out = run(base + ["--mtf_file", "mtf.txt", "--mtf_sampling_rate", "-1"], tmp_path)
    assert rc == 1 and "sampling rate <= 0" in out
    rc, _, out = run(base + ["--mtf_file", "nothere.txt", "--mtf_sampling_rate", "8"], tmp_path)
    assert rc == 1 and "could not read MTF file" in out
    # parameter-file keys; a negative rate is accepted there (tem_io.cpp:4249-4258)
    pf = tmp_path / "p.txt"
    pf.write_text("mtf_file mtf.txt\nmtf_sampling_rate -2.5\n")
    rc, plan, out = run(["--parameter_file", str(pf)] + base, tmp_path)
    assert rc == 0, out
    assert plan["mtf"] == ["3 -2.5"]
    rc, plan, _ = run(base, tmp_path)
    assert plan["mtf"] == ["0 0"]
