"""A DEM file through the GPU LEM and back to files: the reference's
GeoTIFF fixture (DEFLATE, floating-point predictor) read by the library,
three LEM steps on the device, the per-step statistics CSV and the surface
written and read back (host formats checked on the CPU in test_formats.py)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "rasters")


def test_gpu_lem_from_geotiff_to_files(tmp_path):
    print("[gpu formats] lakeland crop 128x128, 3 steps", flush=True)
    import lsdtt3_amd
    from lsdtt3_amd import formats as F
    assert lsdtt3_amd.gpu_available()
    z, mask, meta = F.read_raster(os.path.join(GOLD, "flat_resolution_lakeland_crop_dem.tif"))
    lem = lsdtt3_amd.LandscapeModel(z.astype(np.float64), meta["cell_size"], mask=mask, timestep_years=100.0,
                                    erodibility=1e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-4,
                                    hillslope_diffusivity=0.01, boundary_conditions="oooo")
    csv = tmp_path / "run_model_elevation_stats.csv"
    rows = []
    table = F.ElevationStatsCSV.for_model(lem)
    for _ in range(3):
        st = lem.step()
        rows.append(st)
        table.add(lem, st)
    table.write(csv)
    lines = csv.read_text().splitlines()
    assert len(lines) == 4 and lines[0].startswith("iteration,time_yr,delta_max_z")
    for k, st in enumerate(rows):
        f = lines[k + 1].split(",")
        assert f[0] == str(k + 1) and f[1] == "%f" % float(np.float32(100.0 * (k + 1)))
        assert f[4] == "%f" % float(np.float32(st["max_local_elevation_change"]))
    out = tmp_path / "run_model_iteration_3.tif"
    zz = lem.elevation.astype(np.float32)
    F.write_raster(out, zz, meta, mask=mask)
    back, bmask, bmeta = F.read_raster(out)
    assert np.array_equal(bmask, mask)
    assert np.array_equal(back[mask == 1], zz[mask == 1])
    assert (bmeta["x_min"], bmeta["y_min"], bmeta["cell_size"], bmeta["epsg"]) == \
        (meta["x_min"], meta["y_min"], meta["cell_size"], meta["epsg"])
