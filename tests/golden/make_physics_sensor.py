"""Writes tests/golden/physics_sensor.json: the leaves of the reference's configs/twin/physics.json that
gpu/sensor.js:17-84 (sensorParams) reads, each kept as the file holds it ({value, unit, status, source}).
Run here, where the reference is mounted: python tests/golden/make_physics_sensor.py"""
import json
import os

SRC = "/root/reference/configs/twin/physics.json"
PATHS = [
    "sensor.pixel_pitch", "sensor.qe_peak", "sensor.response_matrix.M", "sensor.lens.f_number",
    "sensor.lens.transmission", "sensor.lens.relative_illumination.pupil_obliquity_exponent",
    "sensor.lens.relative_illumination.vignetting_theta_deg", "sensor.lens.relative_illumination.vignetting",
    "sensor.lens.esf_sigma_px", "sensor.lens.veiling_glare.fraction", "sensor.lens.veiling_glare.sigma_frac_diag",
    "sensor.read_noise", "sensor.full_well", "sensor.dark_current", "sensor.conversion_gain_base", "sensor.adc_bits",
    "sensor.black_level", "sensor.bayer_order", "sensor.analog_gain_range", "sensor.uvc.exposure_unit",
    "sensor.uvc.exposure_range", "sensor.uvc.gain_range", "sensor.uvc.frame_rate", "sensor.uvc.exposure_default",
    "sensor.uvc.white_balance_auto_default", "sensor.auto_exposure.target_mean", "sensor.auto_exposure.damping",
    "sensor.auto_exposure.max_step", "isp.white_balance_gains", "isp.ccm", "isp.tone_curve", "isp.s_curve_c",
    "isp.black_offset_linear", "isp.demosaic", "isp.noise_reduction", "isp.noise_reduction_sigma_px", "isp.gamma",
    "isp.sharpening.amount", "isp.sharpening.radius_px", "isp.chroma_subsampling",
]


def main():
    with open(SRC) as f:
        phys = json.load(f)
    out = {}
    for path in PATHS:
        node, keys = phys, path.split(".")
        for k in keys:
            node = node.get(k) if isinstance(node, dict) else None
            if node is None:
                break
        if node is None:
            continue
        if isinstance(node, dict):
            node = {k: node[k] for k in ("value", "unit", "status", "source") if k in node}
        dst = out
        for k in keys[:-1]:
            dst = dst.setdefault(k, {})
        dst[keys[-1]] = node
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "physics_sensor.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
