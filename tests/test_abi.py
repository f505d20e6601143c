"""The C-ABI library loads without a GPU and exports every declared entry point."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


@pytest.mark.parametrize("header", ["rsh.h", "rsh_engine.h"])
def test_every_declared_symbol_is_exported(rsh, header):
    L = rsh.lib()
    names = declared_functions(header)
    assert len(names) >= (17 if header == "rsh.h" else 39)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_wrapper_names_are_all_present():
    """The 18 stage wrappers and 15 binders of RayTracer.cu:2042-3301."""
    names = set(declared_functions("rsh_engine.h"))
    ref = {"TriangleUpdateWrapper", "BoundingSphereUpdateWrapper", "MemoryPreparationWrapper",
           "ShadowRayCreationWrapper", "ReflectionRayCreationWrapper", "RefractionRayCreationWrapper", "rsh_sort_pairs", "RayTrimmingWrapper", "RayCompressionWrapper",
           "RaySortingWrapper", "RayDecompressionWrapper", "HierarchyCreationWrapper",
           "HierarchyTraversalWarmUpWrapper", "HierarchyTraversalWrapper", "ShadowRayPreparationWrapper",
           "ShadowRayIntersectionWrapper", "ShadowRayColoringWrapper", "ReflectionRayPreparationWrapper",
           "ReflectionRayIntersectionWrapper", "ReflectionRayColoringWrapper"}
    binders = {"bindDiffuseTextureArray", "bindSpecularTextureArray", "bindFragmentPositionArray",
               "bindFragmentNormalArray", "bindTrianglePositions", "bindTriangleNormals",
               "bindTriangleTextureCoordinates", "bindTriangleObjectIDs", "bindTriangleMaterialIDs",
               "bindMaterialDiffuseProperties", "bindMaterialSpecularProperties", "bindBoundingSpheres",
               "bindLightPositions", "bindLightColors", "bindLightIntensities"}
    assert ref <= names and binders <= names


def test_no_gpu_device_count_is_safe(rsh):
    # Counting devices needs no GPU; on a CPU-only host it is 0.
    assert rsh.lib().rsh_device_count() >= 0


def test_missing_library_fails_loudly(monkeypatch):
    import rsh_amd
    monkeypatch.setattr(rsh_amd, "_lib", None)
    monkeypatch.setattr(rsh_amd, "LIB_PATH", "/nonexistent/librsh.so")
    with pytest.raises(rsh_amd.LibraryMissing):
        rsh_amd.lib()
