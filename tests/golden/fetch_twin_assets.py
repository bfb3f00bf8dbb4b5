"""Copies the real twin scene the tests and `bench.py --config twin` render into <repo>/twin_assets/ (kept out of
git: 13.5 MB of mesh data): the reference's assets/scene.json, the 44 mesh files it lists, and
configs/twin/physics.json (materials, emitter, light profile, scope), and the baked node table of the lite renderer
(assets/lite.json, assets/lite_nodes.bin). Run in a container where the reference is
mounted; the folder then travels with the working tree. Also the boards' print textures (assets/textures/*.png).

    python tests/golden/fetch_twin_assets.py [/root/reference]
"""
import json
import os
import shutil
import sys


def main(ref="/root/reference"):
    repo = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    dst = os.path.join(repo, "twin_assets")
    os.makedirs(os.path.join(dst, "meshes"), exist_ok=True)
    src = os.path.join(ref, "assets")
    with open(os.path.join(src, "scene.json")) as f:
        scene = json.load(f)
    shutil.copyfile(os.path.join(src, "scene.json"), os.path.join(dst, "scene.json"))
    for m in scene["meshes"]:
        shutil.copyfile(os.path.join(src, m["file"]), os.path.join(dst, m["file"]))
    shutil.copyfile(os.path.join(ref, "configs", "twin", "physics.json"), os.path.join(dst, "physics.json"))
    # the boards' print textures (the ChArUco mask the board's mesh samples)
    for b in scene.get("boards") or []:
        if b.get("texture"):
            os.makedirs(os.path.dirname(os.path.join(dst, b["texture"])), exist_ok=True)
            shutil.copyfile(os.path.join(src, b["texture"]), os.path.join(dst, b["texture"]))
    # the lite renderer's baked node table (12,976 nodes x 256 gather hits): real sparse input for the Jacobi step
    os.makedirs(os.path.join(dst, "lite"), exist_ok=True)
    for name in ("lite.json", "lite_nodes.bin"):
        shutil.copyfile(os.path.join(src, name), os.path.join(dst, "lite", name))
    print("twin assets: %d meshes, the baked node table -> %s" % (len(scene["meshes"]), dst))


if __name__ == "__main__":
    main(*sys.argv[1:])
