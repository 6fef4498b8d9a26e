"""GPU BSDFs (k_bsdf_eval / k_bsdf_sample through lrt_bsdf_*) against the
oracle on 1M seeded (material, wi, wo, front_face) tuples: eval (f*cos) and
pdf agree within 1e-5 relative (absolute floor 1e-7 for values that are zero
up to rounding), for every reference BSDF class and both measures."""
from pathlib import Path

import numpy as np
import pytest

import linrt
from common import NAMES, bsdf_tuples, material_zoo
from oracle_lib import oracle
from oracle_lib import sky_eval as oracle_sky

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
REL, ABS = 1e-5, 1e-7


@pytest.fixture(scope="module")
def zoo():
    tables = np.load(GOLD / "misc_reference.npz")["slices"]
    return material_zoo(tables), tables


def _close(a, b):
    return np.abs(a - b) <= REL * np.maximum(np.abs(a), np.abs(b)) + ABS


@pytest.mark.parametrize("measure", [linrt.MEASURE_SOLID_ANGLE, linrt.MEASURE_DISCRETE])
@pytest.mark.parametrize("mode", [linrt.MODE_RADIANCE, linrt.MODE_IMPORTANCE])
def test_eval_pdf_1m(zoo, measure, mode):
    mats, tables = zoo
    n = 1 << 20
    mat, wi, wo, front, _, _ = bsdf_tuples(n, seed=7 + measure + 10 * mode, n_mats=len(mats))
    f_g, p_g = linrt.bsdf_eval_pdf(mats, tables, mat, wi, wo, front, measure=measure, mode=mode)
    f_o, p_o = oracle.bsdf_eval_pdf(mats, tables, mat, wi, wo, front, measure=measure, mode=mode)
    ok = np.all(_close(f_g, f_o), axis=1) & _close(p_g, p_o)
    exact = np.all(f_g == f_o, axis=1) & (p_g == p_o)
    for m in range(len(mats)):
        sel = mat == m
        bad = np.nonzero(sel & ~ok)[0]
        print(f"{NAMES[m]:>22}: {sel.sum()} tuples, nonzero {(np.any(f_o[sel] != 0, axis=1)).sum()}, "
              f"bit-exact {exact[sel].mean():.4f}, out of tolerance {len(bad)}")
        for i in bad[:5]:
            print(f"    wi={wi[i]} wo={wo[i]} front={front[i]} gpu={f_g[i]} {p_g[i]!r} oracle={f_o[i]} {p_o[i]!r}")
    assert ok.all()


def test_reference_vectors_on_gpu(zoo):
    """The committed reference outputs, evaluated on the GPU."""
    mats, tables = zoo
    d = np.load(GOLD / "bsdf_reference.npz")
    for meas, tag in ((linrt.MEASURE_SOLID_ANGLE, "sa"), (linrt.MEASURE_DISCRETE, "disc")):
        f, p = linrt.bsdf_eval_pdf(mats, tables, d["mat"], d["wi"], d["wo"], d["front"], measure=meas)
        assert np.all(_close(f, d[f"f_{tag}"])) and np.all(_close(p, d[f"pdf_{tag}"]))


def test_sample_1m(zoo):
    mats, tables = zoo
    n = 1 << 20
    mat, wi, _, front, u, lobe = bsdf_tuples(n, seed=99, n_mats=len(mats))
    g = linrt.bsdf_sample(mats, tables, mat, wi, front, u, lobe)
    o = oracle.bsdf_sample(mats, tables, mat, wi, front, u, lobe)
    same_lobe = (g["sampled_type"] == o["sampled_type"]) & (g["sampled_component"] == o["sampled_component"])
    agree = same_lobe & np.all(np.abs(g["wo"] - o["wo"]) <= 1e-5, axis=1)
    w_ok = np.all(np.abs(g["weight"] - o["weight"]) <= 1e-4 * np.abs(o["weight"]) + 1e-6, axis=1)
    p_ok = np.abs(g["pdf"] - o["pdf"]) <= 1e-4 * np.abs(o["pdf"]) + 1e-6
    for m in range(len(mats)):
        sel = mat == m
        print(f"{NAMES[m]:>22}: same lobe {same_lobe[sel].mean():.5f}, wo within 1e-5 {agree[sel].mean():.5f}, "
              f"weight {w_ok[sel & agree].mean():.5f}, pdf {p_ok[sel & agree].mean():.5f}")
    assert same_lobe.mean() > 0.9999
    assert agree.mean() > 0.999
    assert (w_ok & p_ok)[agree].mean() > 0.999
    assert np.array_equal(g["eta"][agree], o["eta"][agree])


def test_flags(zoo):
    mats, _ = zoo
    for i in range(len(mats)):
        assert linrt.bsdf_flags(mats, i) == oracle.bsdf_flags(mats, i)


def test_sky_gpu_vs_reference_vectors():
    d = np.load(GOLD / "misc_reference.npz")
    sky = linrt.load_sky()
    rgb = linrt.sky_eval(sky, d["sky_dirs"])
    assert np.all(np.abs(rgb - d["sky_rgb"]) <= 1e-5 * np.abs(d["sky_rgb"]) + 1e-9)


def test_sky_gpu_grazing_nan():
    """Grazing escapes (0 < y < 1.6e-8) are NaN on the device exactly where
    the oracle and the reference are NaN; just above they agree."""
    from test_cpu_golden import GRAZING
    sky = linrt.load_sky()
    rgb = linrt.sky_eval(sky, GRAZING)
    want = oracle_sky(sky, GRAZING)
    assert np.array_equal(np.isnan(rgb), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.all(np.abs(rgb[ok] - want[ok]) <= 1e-5 * np.abs(want[ok]))


def test_invalid_material_rejected(zoo):
    mats, tables = zoo
    bad = list(mats) + [linrt.material(linrt.BSDF_COATING, nested=7)]  # coating over a coating
    with pytest.raises(linrt.LinrtError) as e:
        linrt.bsdf_eval_pdf(bad, tables, np.array([0], np.int32), np.zeros((1, 3), np.float32),
                            np.zeros((1, 3), np.float32), np.ones(1, np.uint8))
    assert e.value.code == linrt.ERR_UNSUPPORTED
