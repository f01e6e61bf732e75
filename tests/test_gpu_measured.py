"""Measured (real-scanner) input on the GPU: air-scan normalisation and the
helical window against the oracle (which is pinned to the reference's own
loadProjectionData in test_oracle.py), and the MBFPA driver reading the
detector and air-scan files its input txt names."""
from __future__ import annotations

import os
import shutil

import numpy as np
import pytest

from conftest import HELICAL_CASES, HELICAL_N, MEASURED_CASES, helical_txt, measured_inputs, rel_max
from oracle.oracle import Geom
from test_drivers import _built, _mbfpa_dir, run

pytestmark = pytest.mark.gpu


def _ops():
    from virtualct_amd import ops
    return ops


def _write(tmp_path, air, det):
    air.tofile(tmp_path / "air.raw")
    det.tofile(tmp_path / "det.raw")
    return str(tmp_path / "det.raw"), str(tmp_path / "air.raw")


@pytest.mark.parametrize("case", MEASURED_CASES, ids=[c[0] for c in MEASURED_CASES])
def test_load_measured_matches_oracle(gpu, tmp_path, oracle, case):
    ops = _ops()
    name, inpr, H_txt, rate, nskip, rot, W, seed = case
    H = H_txt // nskip
    air, det, raw = measured_inputs(inpr, H, rate, rot, seed)
    dp, ap = _write(tmp_path, air, det)
    skip = 4 if inpr == "IN" and rate >= 200 else 0
    m = ops.measured_shape(inpr == "IN", rot, H, W, nskip, skip)
    print(f"case {name}: {inpr} rot {rot} H {H} W {W} sino_fact {nskip}", flush=True)
    vol, mask, mx = ops.load_measured(m, dp, ap)
    ref_vol, ref_mask, ref_mx = oracle.load_measured(int(inpr == "IN"), rot, H, W, nskip, air, raw)
    np.testing.assert_array_equal(vol, ref_vol)
    np.testing.assert_array_equal(mask, ref_mask)
    if inpr == "IN":
        assert mx == ref_mx


def test_load_measured_errors(gpu, tmp_path):
    ops = _ops()
    air, det, _ = measured_inputs("PR", 800, 80.0, 1, 21)
    dp, ap = _write(tmp_path, air, det)
    with pytest.raises(RuntimeError, match="skipped channels"):
        ops.load_measured(ops.measured_shape(False, 1, 800, 583), dp, ap)
    with pytest.raises(RuntimeError, match="shorter"):
        ops.load_measured(ops.measured_shape(False, 2, 800, 500), dp, ap)  # one rotation on file
    with pytest.raises(RuntimeError, match="not found"):
        ops.load_measured(ops.measured_shape(False, 1, 800, 500), dp, str(tmp_path / "missing.raw"))
    # IN values above 1 (the reference's "Maximum projection ... STOP")
    rows = np.full((800, 609), 3.0, np.float32)
    rows.tofile(tmp_path / "hot.raw")
    with pytest.raises(RuntimeError, match="Maximum projection"):
        ops.load_measured(ops.measured_shape(True, 1, 800, 500), str(tmp_path / "hot.raw"), ap)


@pytest.mark.parametrize("rot", [1, 2, 3])
def test_helical_window_matches_oracle(gpu, oracle, rot):
    import torch
    ops = _ops()
    H, W = 800, 500
    vol = np.random.default_rng(30 + rot).uniform(0, 1, (rot * H, W)).astype(np.float32)
    win = ops.helical_window(torch.from_numpy(vol).to(gpu), rot, H).cpu().numpy()
    np.testing.assert_array_equal(win, oracle.helical_window(vol, rot, H, W))


def _measured_txt(d, inpr, width, height, det="det.raw", air="air.raw"):
    txt = (d / "in.txt").read_text()
    txt = txt.replace("// Energy Label (IN (for treatment beam) or PR (for CT))\nPR",
                      f"// Energy Label (IN (for treatment beam) or PR (for CT))\n{inpr}")
    txt = txt.replace("// sinogram width \n609", f"// sinogram width \n{width}")
    txt = txt.replace("// sinogram height \n800", f"// sinogram height \n{height}")
    txt = txt.replace("// sinogram image data for input \n*", f"// sinogram image data for input \n{det}")
    txt = txt.replace("// Air scan image data for input // parameters are fixed\n*",
                      f"// Air scan image data for input // parameters are fixed\n{air}")
    txt = txt.replace("//Couch speed [cm/rotation] // 0.4=fine, 0.8=normal, 1.2=coarse\n0.0",
                      "//Couch speed [cm/rotation] // 0.4=fine, 0.8=normal, 1.2=coarse\n0.8")
    txt = txt.replace("// reconstruction start z [cm]\n0.0", "// reconstruction start z [cm]\n1.3")
    (d / "in.txt").write_text(txt)


def test_mbfpa_driver_npf_from_measured_files(gpu, tmp_path, golden, oracle):
    """PR detector and air-scan files named in the txt (3 rotations of 800
    views, 500 channels): projections.raw is the normalised volume and
    npf_float.raw the npf of its helical window for z = 1.3 cm."""
    _built()
    d = _mbfpa_dir(tmp_path, golden, noise="0.0")
    rot, H, W = 3, 800, 500
    _measured_txt(d, "PR", W, rot * H)
    air, det, raw = measured_inputs("PR", H, 80.0, rot, 40)
    air.tofile(d / "air.raw")
    det.tofile(d / "det.raw")
    r = run("mvct_mbfpa", ["in.txt"], d)
    assert r.returncode == 0, r.stderr
    assert "Maximum projection" in r.stderr and "Z scan step" in r.stderr
    vol, _, _ = oracle.load_measured(0, rot, H, W, 1, air, raw)
    np.testing.assert_array_equal(np.fromfile(d / "projections.raw", np.float32).reshape(rot * H, W), vol)
    win = oracle.helical_window(vol, rot, H, W)
    # the window holds 500 of the 609-channel detector's channels
    ref = oracle.lengthproj_ref(Geom(), win, 1.3, 0.8)
    npf = np.fromfile(d / "npf_float.raw", np.float32).reshape(512, 512)
    assert rel_max(npf, ref) < 1e-5
    assert np.fromfile(d / "reprojection_float.raw", np.float32).size == H * W


@pytest.mark.parametrize("case", HELICAL_CASES, ids=[c[0] for c in HELICAL_CASES])
def test_mbfpa_driver_helical_npf_against_reference_program(gpu, tmp_path, golden, oracle, case):
    """The driver on the measured files with couch motion that the reference's
    whole MBFPA program was run on (256^2, 500 channels): npf_float.raw within
    1e-5 of the oracle chain, which equals the reference program's output bit
    for bit (test_oracle.py)."""
    _built()
    name, inpr, rot, H, W, seed, z, couch = case
    d = tmp_path / name
    os.makedirs(d / "Spectrum")
    shutil.copyfile(golden.path("spectrum_120kV.txt"), d / "Spectrum" / "PhaseSpace_photon_120kV_0cm.txt")
    c0 = (512 - HELICAL_N) // 2
    golden.icru110()[:, c0:c0 + HELICAL_N, c0:c0 + HELICAL_N].copy().tofile(d / "Weight_input_ICRU110.raw")
    air, det, raw = measured_inputs(inpr, H, 80.0, rot, seed)
    air.tofile(d / "air.raw")
    det.tofile(d / "det.raw")
    (d / "in.txt").write_text(helical_txt(open(golden.path("input_mbfpa_120kV.txt")).read(), inpr, W, rot * H, z,
                                          couch, HELICAL_N))
    print(f"case {name}", flush=True)
    r = run("mvct_mbfpa", ["in.txt"], d)
    assert r.returncode == 0, r.stderr
    vol, _, _ = oracle.load_measured(int(inpr == "IN"), rot, H, W, 1, air, raw)
    ref = oracle.lengthproj_ref(Geom(N=HELICAL_N, pix=0.1), oracle.helical_window(vol, rot, H, W), z, couch)
    npf = np.fromfile(d / "npf_float.raw", np.float32).reshape(HELICAL_N, HELICAL_N)
    assert rel_max(npf, ref) < 1e-5


def test_mbfpa_driver_measured_stop(gpu, tmp_path, golden):
    """IN data above twice the air scan: the reference prints its STOP line
    and ends with status 0 before writing any output."""
    _built()
    d = _mbfpa_dir(tmp_path, golden, noise="0.0")
    _measured_txt(d, "IN", 500, 800)
    air, _, _ = measured_inputs("PR", 800, 80.0, 1, 41)
    air.tofile(d / "air.raw")
    np.full((800, 609), 3.0, np.float32).tofile(d / "det.raw")
    r = run("mvct_mbfpa", ["in.txt"], d)
    assert r.returncode == 0 and "STOP" in r.stderr, r.stderr
    assert not (d / "npf_float.raw").exists() and not (d / "reprojection_float.raw").exists()
