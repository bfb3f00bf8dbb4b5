"""js/sensor.mjs, the sensor / ISP parameters and the UVC loop of gpu/sensor.js, on the CPU.

Inputs: tests/golden/physics_sensor.json, the leaves of the reference's configs/twin/physics.json that sensorParams
reads (written by tests/golden/make_physics_sensor.py). Pins: every value and status comes from the file; the
derived uniform numbers (optics constant pi T / 4N^2, the ESF remainder after the pixel aperture, 2^bits - 1, the
glare sigma at the 1/16 level) against their closed forms; the ISP mode strings map to the kernel codes; auto
exposure drives a simulated linear camera to the target luma; grey-world AWB follows its closed form
log g_n = (1 - (1 - d)^n) log target.
"""
import json
import math
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE = os.path.join(ROOT, "veridien-twin_amd", "js", "sensor.mjs")
FIXTURE = os.path.join(ROOT, "tests", "golden", "physics_sensor.json")
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None, reason="node is not available")


def _run(body):
    code = ("import('%s').then(m => { const phys = JSON.parse(require('fs').readFileSync('%s', 'utf8')); %s })"
            ".catch(e => { console.error(e); process.exit(1); })" % (MODULE, FIXTURE, body))
    r = subprocess.run([NODE, "--input-type=commonjs", "-e", code], capture_output=True, text=True, timeout=120,
                       cwd=ROOT)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def _val(d, path):
    for k in path.split("."):
        d = d[k]
    return d["value"]


def test_sensor_params_from_the_physics_file():
    phys = json.load(open(FIXTURE))
    d = _run("const p = m.sensorParams(phys); console.log(JSON.stringify({p, d: m.describe(p)}));")
    p = d["p"]
    assert p["fNumber"] == _val(phys, "sensor.lens.f_number")
    assert p["readNoise"] == _val(phys, "sensor.read_noise")
    assert p["blackLevel"] == _val(phys, "sensor.black_level")
    assert p["cfa"] == _val(phys, "sensor.bayer_order")
    assert p["demosaic"] == _val(phys, "isp.demosaic")
    assert p["sCurveC"] == _val(phys, "isp.s_curve_c")
    assert p["glareFraction"] == _val(phys, "sensor.lens.veiling_glare.fraction")
    assert p["analogGainMax"] == _val(phys, "sensor.analog_gain_range")[1]
    assert p["exposureUnits0"] == _val(phys, "sensor.uvc.exposure_default")
    M = _val(phys, "sensor.response_matrix.M")
    src = phys["sensor"]["response_matrix"]["M"]["source"]
    if "placeholder" in src.lower() or max(abs(v) for row in M for v in row) < 1e3:
        assert p["provenance"]["response_matrix"]["status"] == "derived"
    else:
        assert p["responseMatrix"] == M
    assert p["provenance"]["read_noise"]["status"] == phys["sensor"]["read_noise"]["status"]
    assert sum(d["d"]["counts"].values()) == len(p["provenance"])


def test_placeholder_response_matrix_is_derived():
    d = _run("const p = m.sensorParams({sensor: {response_matrix: {M: {value: [[1,0,0],[0,1,0],[0,0,1]], "
             "source: 'placeholder'}}}}); console.log(JSON.stringify(p.responseMatrix));")
    h, c = 6.62607015e-34, 299792458.0
    want = [3e-6 * 3e-6 * q * nm * 1e-9 / (h * c) for q, nm in zip((0.45, 0.55, 0.5), (600, 540, 460))]
    for k in range(3):
        assert d[k][k] == pytest.approx(want[k], rel=1e-12)
        assert sum(abs(v) for v in d[k]) == pytest.approx(want[k], rel=1e-12)


def test_native_parameters_and_mode_codes():
    d = _run("const p = m.sensorParams(phys); console.log(JSON.stringify({s: m.nativeSensor(p, {width: 640, "
             "height: 480, frame: 3, seed: 9, exposure: 0.0157, gain: 2}), q: m.nativeIsp(p), "
             "modes: ['bilinear', 'edge_aware', 'hamilton_adams'].map(x => m.nativeIsp({...p, demosaic: x}).demosaic), "
             "tones: [['srgb', null], ['srgb', 1.3], ['linear', null], ['gamma', null]].map(([t, c]) => "
             "m.nativeIsp({...p, tone: t, sCurveC: c}).tone), "
             "nr: ['none', 'edge_preserving', 'bilateral', 'on'].map(x => m.nativeIsp({...p, noiseReduction: x}).nr)}));")
    phys = json.load(open(FIXTURE))
    s, q = d["s"], d["q"]
    assert s["opticsK"] == pytest.approx(math.pi * 0.9 / (4 * 2.0 * 2.0), rel=1e-15)
    esf = _val(phys, "sensor.lens.esf_sigma_px")
    assert s["esfSigma"] == pytest.approx(math.sqrt(max(0.0, esf * esf - 1.0 / 12.0)), rel=1e-12)
    assert s["maxDn"] == 2 ** _val(phys, "sensor.adc_bits") - 1
    assert s["black"] == _val(phys, "sensor.black_level") and s["cfa"] == 0 and s["frame"] == 3 and s["seed"] == 9
    assert s["glareSigma"] == pytest.approx(0.25 * 800.0 / 16.0) and s["glareLevel"] == 16
    assert len(s["m"]) == 9 and s["exposure"] == 0.0157 and s["gain"] == 2
    assert q["demosaic"] == 1 and q["nr"] == 1 and q["tone"] == 3 and q["chroma420"] == 1
    assert q["sCurveC"] == _val(phys, "isp.s_curve_c") and q["blackOffset"] == _val(phys, "isp.black_offset_linear")
    # the reference tests the string with /edge|bilateral|on/, which 'none' matches too (gpu/sensor.js:171)
    assert d["modes"] == [0, 1, 2] and d["tones"] == [0, 3, 2, 1] and d["nr"] == [1, 1, 1, 1]


def test_auto_exposure_reaches_the_target():
    """A linear camera behind an sRGB-like curve: luma = min(1, (k t g)^(1/2.2)); the UVC loop settles at the
    target, exposure in whole units first and gain only past the frame time."""
    d = _run("const p = m.sensorParams(phys); const out = [];"
             "for (const k of [40, 0.5]) { const ae = new m.AutoExposure(p, {exposureUnits: p.exposureUnits0}); let L = 0;"
             "  for (let i = 0; i < 60; i++) { L = Math.min(1, Math.pow(k * ae.exposure * ae.gain, 1 / 2.2)); ae.update(L); }"
             "  out.push({L, units: ae.units, gain: ae.gain, gu: ae.gainUnits}); }"
             "console.log(JSON.stringify({out, g0: m.analogGain(p, 0), g1: m.analogGain(p, 100), max: p.analogGainMax}));")
    bright, dark = d["out"]
    assert bright["L"] == pytest.approx(0.45, rel=0.03) and bright["gain"] == 1.0
    assert dark["units"] == 500                      # 1 / (20 fps x 100 us): the frame time caps the exposure
    assert dark["gain"] > 1.0 and dark["L"] == pytest.approx(0.45, rel=0.05)
    assert d["g0"] == 1.0 and d["g1"] == d["max"]


def test_awb_closed_form():
    d = _run("let g = null; const mean = [0.2, 0.3, 0.15]; const out = [];"
             "for (let i = 0; i < 5; i++) { g = m.updateAWB(g, mean, 0.3); out.push(g); }"
             "console.log(JSON.stringify({out, same: m.updateAWB([1.2, 1, 0.9], [0.3, 0.3, 0.3], 0.3), "
             "dark: m.updateAWB([1.1, 1, 1], [0, 0.3, 0.3], 0.3)}));")
    target = [0.3 / 0.2, 1.0, 0.3 / 0.15]
    for n, g in enumerate(d["out"], start=1):
        for c in range(3):
            assert math.log(g[c]) == pytest.approx((1 - 0.7 ** n) * math.log(target[c]), rel=1e-12, abs=1e-15)
    for c, v in enumerate([1.2, 1.0, 0.9]):
        assert d["same"][c] == pytest.approx(math.exp(0.7 * math.log(v)), rel=1e-12)
    assert d["dark"] == [1.1, 1, 1]
