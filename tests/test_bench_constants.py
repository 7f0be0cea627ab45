"""The HBM traffic figures that bench.py reports are constants taken from rocprofv3 counter runs
whose per-kernel tables are committed under profiles/.  These checks keep the two in agreement:
a constant changed without its table (or the reverse) fails here."""
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _row(path, kernel):
    """(FETCH_SIZE KiB, WRITE_SIZE KiB) of the table row that starts with `kernel`."""
    with open(os.path.join(ROOT, "profiles", path)) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 3 and cells[0] == kernel:
                return int(cells[1]), int(cells[2])
    raise AssertionError("no row %r in %s" % (kernel, path))


def test_headline_stencil_traffic_matches_its_profile():
    f, w = _row("k_stencil_elasticity_hex8_64_sphere_array_hbm_traffic.md",
                "k_stencil<3,3,64,false,float> (mixed path)")
    assert bench.STENCIL_TRAFFIC_BYTES == (2 * f + w) * 1024


def test_hex27_traffic_matches_its_profile():
    f, w = _row("k_generic_hex27_elasticity_p2_gyroid_32_hbm_traffic.md",
                "k_generic_hex27<3,double,false> (FP64)")
    assert bench.HEX27_32_TRAFFIC_BYTES == (2 * f + w) * 1024
    # the writes are the K_e pool of the 32^3 set: 188,116,992 FP64 entries, to within 0.1 %
    assert abs(w * 1024 - 188116992 * 8) < 1e-3 * 188116992 * 8


def test_c2_stencil_traffic_per_variant_matches_its_profile():
    rows = {0: 64, 1: 128, 2: 256, 3: 64, 4: 128, 5: 256}
    for v, (f, w) in bench.C2_STENCIL_TRAFFIC_KIB.items():
        kernel = "k_stencil<3,1,%d,%s,double> (variant %d)" % (rows[v], "true" if v >= 3 else "false", v)
        assert _row("k_stencil_diffusion_hex8_128_sphere_hbm_traffic.md", kernel) == (f, w), v


def test_profiles_name_their_commands():
    """Every summary under profiles/ gives the rocprofv3 command that produced it."""
    for name in sorted(os.listdir(os.path.join(ROOT, "profiles"))):
        with open(os.path.join(ROOT, "profiles", name)) as f:
            text = f.read()
        assert re.search(r"rocprofv3 |mgpu_hbm_calibrate", text), name


def test_thb48_scatter_traffic_matches_its_profile():
    f, w = _row("k_generic_hex27_scatter_elasticity_p2_thb_gyroid_48_hbm_traffic.md",
                "k_generic_hex27<3,double,true,false>, 42 launches (FP64 subdivision)")
    assert bench.THB48_SCATTER_TRAFFIC_BYTES == (2 * f + w) * 1024
    # the zero fill writes the 949,141,152 FP64 values once (0.11 % over, measured), to within 0.5 %
    _, wz = _row("k_generic_hex27_scatter_elasticity_p2_thb_gyroid_48_hbm_traffic.md", "k_generic_zero")
    assert abs(wz * 1024 - 949141152 * 8) < 5e-3 * 949141152 * 8
