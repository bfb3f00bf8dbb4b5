"""Writes tests/golden/kb4_scene_cameras.json: the three tracking cameras' KB4 lens (K, D, size) and the numbers the
reference's scene export records for them (assets/scene.json cameras[]: theta_corner_rad, theta_fold_rad,
checks.aim_target_px, checks.aim_target_off_axis_deg, checks.aim_target_off_principal_px). Data only; run where the
reference is mounted:

    python tests/golden/make_kb4_fixture.py [/root/reference]
"""
import json
import os
import sys


def main(ref="/root/reference"):
    scene = json.load(open(os.path.join(ref, "assets", "scene.json")))
    out = []
    for c in scene["cameras"]:
        out.append(dict(name=c["name"], K=c["lens"]["K"], D=c["lens"]["D"], size=c["lens"]["size"],
                        theta_corner_rad=c["theta_corner_rad"], theta_fold_rad=c["theta_fold_rad"],
                        aim_target_px=c["checks"]["aim_target_px"],
                        aim_target_off_axis_deg=c["checks"]["aim_target_off_axis_deg"],
                        aim_target_off_principal_px=c["checks"]["aim_target_off_principal_px"]))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kb4_scene_cameras.json")
    with open(dst, "w") as f:
        json.dump({"source": "assets/scene.json cameras[] (lens, theta_corner_rad, theta_fold_rad, checks)",
                   "cameras": out}, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(*sys.argv[1:])
