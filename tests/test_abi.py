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


@pytest.mark.parametrize("header", ["rsh.h", "rsh_engine.h", "rsh_ctx.h"])
def test_every_declared_symbol_is_exported(rsh, header):
    L = rsh.lib()
    names = declared_functions(header)
    assert len(names) >= {"rsh.h": 17, "rsh_engine.h": 39, "rsh_ctx.h": 26}[header]
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


CTX_STAGES = {
    "TriangleUpdateWrapper": "rsh_triangle_update", "BoundingSphereUpdateWrapper": "rsh_bounding_sphere_update",
    "MemoryPreparationWrapper": "rsh_memory_preparation", "ShadowRayCreationWrapper": "rsh_shadow_ray_creation",
    "ReflectionRayCreationWrapper": "rsh_reflection_ray_creation",
    "RefractionRayCreationWrapper": "rsh_refraction_ray_creation", "RayTrimmingWrapper": "rsh_ray_trimming",
    "RayCompressionWrapper": "rsh_ray_compression", "RaySortingWrapper": "rsh_ray_sorting",
    "rsh_sort_pairs": "rsh_ctx_sort_pairs", "RayDecompressionWrapper": "rsh_ray_decompression",
    "HierarchyCreationWrapper": "rsh_hierarchy_creation",
    "HierarchyTraversalWarmUpWrapper": "rsh_hierarchy_traversal_warmup",
    "HierarchyTraversalWrapper": "rsh_hierarchy_traversal", "ShadowRayPreparationWrapper": "rsh_shadow_ray_preparation",
    "ShadowRayIntersectionWrapper": "rsh_shadow_ray_intersection", "ShadowRayColoringWrapper": "rsh_shadow_ray_coloring",
    "ReflectionRayPreparationWrapper": "rsh_reflection_ray_preparation",
    "ReflectionRayIntersectionWrapper": "rsh_reflection_ray_intersection",
    "ReflectionRayColoringWrapper": "rsh_reflection_ray_coloring",
}


def declared_params(header, name):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"\b" + name + r"\s*\(([^;{]*)\)\s*;", text)
    return [p.strip() for p in m.group(1).split(",")]


def test_every_stage_has_a_context_form():
    """Each drop-in stage has a context form taking (ctx, stream) and then the
    wrapper's own parameters in order (plus device-count words)."""
    ctx = set(declared_functions("rsh_ctx.h"))
    for wrapper, native in CTX_STAGES.items():
        assert native in ctx, native
        p = declared_params("rsh_ctx.h", native)
        assert p[0].startswith("rsh_ctx*") and p[1].startswith("void*"), (native, p[:2])
        w = [x.split()[-1].lstrip("*") for x in declared_params("rsh_engine.h", wrapper)]
        n = [x.split()[-1].lstrip("*") for x in p[2:]]
        if wrapper == "MemoryPreparationWrapper":
            w = w[:3]            # the caller's buffers are not the library's to touch
        extra = [x for x in n if x.endswith("Device")]
        assert [x for x in n if not x.endswith("Device")] == [x for x in w if x not in extra], (native, n, w)


def test_no_gpu_device_count_is_safe(rsh):
    # Counting devices needs no GPU; on a CPU-only host it is 0.
    assert rsh.lib().rsh_device_count() >= 0


def test_missing_library_fails_loudly(monkeypatch):
    import rsh_amd
    monkeypatch.setattr(rsh_amd, "_lib", None)
    monkeypatch.setattr(rsh_amd, "LIB_PATH", "/nonexistent/librsh.so")
    with pytest.raises(rsh_amd.LibraryMissing):
        rsh_amd.lib()
