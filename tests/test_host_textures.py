"""Material textures named by a scene's material json (Scene::LoadMaterial,
scene.cpp:155-203): string values of albedo / opacity / reflectiveness /
emissivity are texture files relative to the scene's directory, loaded once
each.  TGA (uncompressed and RLE, bottom-up and top-down), binary PPM and PFM
decode to the expected texels with row 0 at the bottom; the json round trip
keeps the file names; a "reflectiveness" constant replaces the exponent."""
import json
import os
import struct

import numpy as np
import pytest


def tga(path, rgb, rle=False, top_down=False, alpha=None):
    h, w, _ = rgb.shape
    px = rgb[..., ::-1]  # BGR
    if alpha is not None:
        px = np.concatenate([px, alpha[..., None]], -1)
    rows = px if top_down else px[::-1]  # stored bottom row first unless top-down
    bpp = px.shape[2] * 8
    head = struct.pack("<BBBHHBHHHHBB", 0, 0, 10 if rle else 2, 0, 0, 0, 0, 0, w, h, bpp, 0x20 if top_down else 0)
    flat = rows.reshape(-1, px.shape[2])
    if not rle:
        body = flat.tobytes()
    else:  # one run packet for equal neighbours, raw packets otherwise
        body, i = b"", 0
        while i < len(flat):
            j = i
            while j + 1 < len(flat) and j - i < 127 and (flat[j + 1] == flat[i]).all():
                j += 1
            if j > i:
                body += bytes([0x80 | (j - i)]) + flat[i].tobytes()
                i = j + 1
            else:
                body += bytes([0]) + flat[i].tobytes()
                i += 1
    with open(path, "wb") as f:
        f.write(head + body)


@pytest.fixture
def scene_dir(rt, tmp_path):
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    mats = [rt.default_material("A"), rt.default_material("B")]
    hs = rt.HostScene.build(pos, np.array([[0, 2, 1], [0, 3, 2]]), material_ids=np.array([0, 1]), materials=mats)
    hs.save(str(tmp_path / "s.rawscene"), str(tmp_path / "s.json"))
    os.makedirs(tmp_path / "tex")
    return tmp_path


def load(rt, d, materials):
    with open(d / "s.json") as f:
        j = json.load(f)
    for name, vals in materials.items():
        j["materials"][name].update(vals)
    with open(d / "s.json", "w") as f:
        json.dump(j, f)
    return rt.HostScene.load(str(d / "s.rawscene"), str(d / "s.json"))


def test_tga_ppm_pfm_decode_and_bindings(rt, scene_dir):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    a[:, 2:5] = a[0, 0]  # runs for the RLE encoder
    tga(scene_dir / "tex" / "a.tga", a)
    tga(scene_dir / "tex" / "b.tga", a, rle=True, top_down=True, alpha=np.full((5, 7), 77, np.uint8))
    c = rng.integers(0, 256, (3, 4, 3), dtype=np.uint8)
    with open(scene_dir / "tex" / "c.ppm", "wb") as f:
        f.write(b"P6\n# comment\n4 3\n255\n" + c.tobytes())  # rows top-down
    e = rng.uniform(0, 4, (2, 3, 3)).astype(np.float32)
    with open(scene_dir / "tex" / "e.pfm", "wb") as f:
        f.write(b"PF\n3 2\n-1.0\n" + e.astype("<f4").tobytes())  # rows bottom-up already
    hs = load(rt, scene_dir, {"A": {"albedo": "tex\\a.tga", "emissivity": "tex/e.pfm"},
                              "B": {"albedo": "tex/b.tga", "opacity": "tex/c.ppm", "reflectiveness": "tex/a.tga"}})
    texs = hs.textures()
    mt = hs.material_textures()
    assert len(texs) == 4  # a.tga is loaded once for both materials
    ia, ie = mt[0, 0], mt[0, 3]
    ib, ic = mt[1, 0], mt[1, 1]
    assert mt[1, 2] == ia and mt[0, 1] == -1 and mt[0, 2] == -1 and mt[1, 3] == -1
    # 8-bit images: RGBA8, row 0 = the bottom row of the picture (the last row of a top-down file)
    want_a = np.concatenate([a, np.full((5, 7, 1), 255, np.uint8)], -1)[::-1]
    assert texs[ia].dtype == np.uint8 and np.array_equal(texs[ia], want_a)
    want_b = np.concatenate([a, np.full((5, 7, 1), 77, np.uint8)], -1)[::-1]
    assert np.array_equal(texs[ib], want_b)
    assert np.array_equal(texs[ic][..., :3], c[::-1]) and (texs[ic][..., 3] == 255).all()
    assert texs[ie].dtype == np.float32 and np.array_equal(texs[ie][..., :3], e) and (texs[ie][..., 3] == 1).all()


def test_json_round_trip_keeps_texture_files(rt, scene_dir):
    a = np.full((2, 2, 3), 100, np.uint8)
    tga(scene_dir / "tex" / "a.tga", a)
    hs = load(rt, scene_dir, {"B": {"albedo": "tex/a.tga", "reflectiveness": [1, 1, 1, 42.0]}})
    hs.save(str(scene_dir / "t.rawscene"), str(scene_dir / "t.json"))
    with open(scene_dir / "t.json") as f:
        j = json.load(f)
    assert j["materials"]["B"]["albedo"] == "tex/a.tga"
    assert j["materials"]["B"]["reflectiveness"] == [1, 1, 1, 42]
    hs2 = rt.HostScene.load(str(scene_dir / "t.rawscene"), str(scene_dir / "t.json"))
    assert np.array_equal(hs2.material_textures(), hs.material_textures())
    m = hs2.arrays()["materials"]
    assert m["reflectiveness"][1].tolist() == [1, 1, 1, 42] and m["reflectiveness"][0][3] == 1.0  # roughness 1


def test_missing_or_bad_texture_is_an_error(rt, scene_dir):
    with pytest.raises(rt.GpugiError, match="cannot open texture"):
        load(rt, scene_dir, {"A": {"albedo": "tex/none.tga"}})
    with open(scene_dir / "tex" / "bad.tga", "wb") as f:
        f.write(b"\x00\x01\x01" + bytes(15))
    with pytest.raises(rt.GpugiError, match="colour-mapped"):
        load(rt, scene_dir, {"A": {"albedo": "tex/bad.tga"}})
