"""The plane extents the GPU plane tests cover are exactly the ones the
library instantiates: the registry lists in plane_kernels.hip (one-workgroup
kernels) and plane_kernels_wide.hip (wide split kernels), parsed here,
against tests/test_gpu_planes.py's REGULAR_PLANES and WIDE_PLANES."""
import importlib.util
import os
import re

CSRC = os.path.join(os.path.dirname(__file__), "..", "fourierconvolutioncudalib_amd", "csrc")


def _gpu_plane_lists():
    path = os.path.join(os.path.dirname(__file__), "test_gpu_planes.py")
    spec = importlib.util.spec_from_file_location("gpu_planes_lists", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.REGULAR_PLANES, mod.WIDE_PLANES


def test_plane_lists_match_registry():
    reg = open(os.path.join(CSRC, "plane_kernels.hip")).read().split("const PlaneEntry kPlanes[]")[1]
    reg = reg.split("};")[0]
    rows = [int(n) for n in re.findall(r"FC_PLANE_ROW\((\d+)\)", reg)]
    regular = [(nx, ny) for nx in rows for ny in (64, 80, 96, 112, 120, 128)]
    regular += [(int(a), int(b)) for a, b in re.findall(r"FC_PLANE\((\d+), (\d+)\)", reg)]
    wide_src = open(os.path.join(CSRC, "plane_kernels_wide.hip")).read()
    wide_src = wide_src.split("static const PlaneEntry table[]")[1].split("};")[0]
    wide = [(int(a), int(b)) for a, b in re.findall(r"FC_PLANE_WIDE(?:_INV)?\((\d+), (\d+)\)", wide_src)]
    t_regular, t_wide = _gpu_plane_lists()
    assert len(regular) == len(set(regular)) == 69
    assert len(wide) == len(set(wide)) == 20
    assert sorted(regular) == sorted(t_regular)
    assert sorted(wide) == sorted(t_wide)
    assert not set(regular) & set(wide)
