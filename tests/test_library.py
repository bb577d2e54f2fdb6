"""CPU checks of the product library: it is built for gfx950, loads, and exports every C entry point that
include/*.h declares (no compute call without a GPU)."""

import glob
import os
import re

import ryoanji_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(ryo_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"ryo_sfc_keys_f32", "ryo_sort_pairs_u64_i32", "ryo_build_leaves_u64", "ryo_link_octree_u64",
            "ryo_upsweep_f32", "ryo_traverse_f32", "ryo_ewald_f32", "ryo_direct_sum_f64"} <= names


def test_library_exports_every_declared_symbol():
    lib = R.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert declared_symbols() <= set(R.SIGNATURES)


def test_library_is_built_for_gfx950():
    # the embedded HIP fat binary names its offload target
    assert b"amdgcn-amd-amdhsa--gfx950" in open(R.LIB_PATH, "rb").read()


def test_sort_temp_bytes_is_host_only():
    assert R.lib().ryo_sort_temp_bytes(1 << 20) > 0


def test_params_mirror_the_header():
    """ryoanji_amd.Params lists ryo_params' fields in the header's order, and ryo_default_params (host only) fills
    the reference demo's settings"""
    src = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "ryoanji_amd.h")).read(), flags=re.S)
    body = re.search(r"typedef struct\s*\{([^}]*)\}\s*ryo_params;", src).group(1)
    fields = [re.match(r"\s*[\w\s]+?\b(\w+)\s*(?:\[\d+\])?\s*$", d).group(1) for d in body.split(";") if d.strip()]
    assert fields == [f[0] for f in R.Params._fields_]
    p = R.default_params()
    assert (p.bucket_size, p.theta, p.G, p.group_size, p.num_shells, p.ewald) == (64, 0.5, 1.0, 64, 0, 0)
    assert list(p.ewald_settings) == [1, 2.6, 2.8, 2.0, 3.0e-3]
    assert (p.init_node, p.soft_mac, p.group_split_tol, p.multipole) == (0, 0, 0.0, 0)


def test_server_order_arguments_are_checked_on_the_host():
    lib = R.lib()
    assert lib.ryo_server_set_order(None, 1, 0.0) != 0
