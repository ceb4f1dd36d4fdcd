"""CPU checks of the C-ABI boundary: the library loads, exports every entry point
include/aph.h declares, uses the reference's 400-byte MaterialConstants layout
(fixture from Tests/Assets/TextFiles/DefaultMaterialUniformBuffer.slang),
reports errors as status codes, and its host-side helpers (inverse, alias map,
material defaults) behave without a GPU."""
import ctypes as C
import json
import os
import re

import numpy as np

from aurora_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "aph.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|double|uint64_t|const char\*)\s+(aph_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)
    assert lib.aph_version() == 1


def test_material_layout_matches_reference_fixture():
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "material_layout.json")))
    ours = {name: getattr(capi.Material, name).offset for name, _ in capi.Material._fields_}
    prop_to_field = lambda p: p.lower()  # noqa: E731
    checked = 0
    for m in fx["members"]:
        prop = m["property"]
        if prop.startswith("has_") or "image" in prop:
            continue  # texture block, kept as raw words
        field = prop_to_field(prop)
        assert field in ours, prop
        assert ours[field] == m["offset"], (prop, ours[field], m["offset"])
        checked += 1
    assert checked == 32  # base .. thin_walled
    # texture flags at the fixture offsets map into texture_block
    tb = capi.Material.texture_block.offset
    for m in fx["members"]:
        if m["property"].startswith("has_"):
            assert (m["offset"] - tb) % 4 == 0 and m["offset"] >= tb
    assert C.sizeof(capi.Material) == 400


def test_material_defaults_are_standard_surface_defaults():
    m = capi.default_material()
    assert m.base == np.float32(0.8) and list(m.base_color) == [1, 1, 1]
    assert m.specular == 1.0 and m.specular_roughness == np.float32(0.2) and m.specular_ior == np.float32(1.5)
    assert m.sheen_roughness == np.float32(0.3) and m.coat_roughness == np.float32(0.1)
    assert m.coat_ior == np.float32(1.5) and list(m.opacity) == [1, 1, 1] and m.thin_walled == 0


def test_errors_are_status_codes_without_a_gpu_context():
    lib = capi.load()
    ctx = C.c_void_p()
    rc = lib.aph_create(9999, C.byref(ctx))
    assert rc == capi.APH_E_DEVICE and not ctx.value
    assert lib.aph_last_error(None) == b"null context"
    assert lib.aph_blas_build(None, 1, 0) == capi.APH_E_INVALID
    assert lib.aph_inverse3x4(None, None) == capi.APH_E_INVALID
    singular = np.zeros(12, dtype=np.float32)
    out = np.ones(12, dtype=np.float32)
    assert lib.aph_inverse3x4(capi.ptr(singular), capi.ptr(out)) == capi.APH_E_INVALID
    assert (out == 0).all()
    assert lib.aph_alias_build(None, 4, 4, None, None) == capi.APH_E_INVALID


def test_session_fails_loudly_without_gpu(monkeypatch):
    import pytest
    from aurora_amd.session import AphError, Renderer

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(AphError):
        Renderer(0)
