"""The oracle against the committed vectors of the reference (tests/golden/,
written by make_golden.py from the reference's own sources): BSDF eval/pdf/
sample, ray hits, Hosek sky radiance, fdr, rough-transmittance slices, the
reference sampler sequence and per-sample path radiance.  These run anywhere
(no reference tree, no GPU)."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

import linrt
import oracle_lib as OL
from common import material_zoo

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def tables():
    return np.load(GOLD / "misc_reference.npz")["slices"]


def test_slices_and_product_data(tables):
    """The committed displaced-sphere tables are the reference's slices."""
    t, _ = linrt.load_tables()
    assert np.array_equal(t, tables)


def test_fdr_matches_reference():
    d = np.load(GOLD / "misc_reference.npz")
    mine = np.array([linrt.fresnel_diffuse_reflectance(e) for e in d["fdr_eta"]], np.float32)
    assert np.array_equal(mine, d["fdr"])


def test_fdr_against_double_quadrature():
    """fdr is the cosine-weighted hemispherical average of Fresnel: compare
    the FP32 adaptive Gauss-Lobatto with a double-precision Gauss-Legendre
    integral (independent of the reference code)."""
    def fresnel(ci, eta):
        s = np.where(ci > 0, 1 / eta, eta)
        ct2 = 1 - (1 - ci * ci) * s * s
        out = np.ones_like(ci)
        ok = ct2 > 0
        ai, at = np.abs(ci[ok]), np.sqrt(ct2[ok])
        rs = (ai - eta * at) / (ai + eta * at)
        rp = (eta * ai - at) / (eta * ai + at)
        out[ok] = 0.5 * (rs * rs + rp * rp)
        return out
    x, w = np.polynomial.legendre.leggauss(4000)
    for eta in (1 / 1.49, 1 / 1.5, 1.5, 1 / 1.7):
        xi = 0.5 * (x + 1)
        ref = 0.5 * np.sum(w * fresnel(np.sqrt(xi), eta))
        assert abs(linrt.fresnel_diffuse_reflectance(eta) - ref) < 2e-5 * max(ref, 1e-3)


def test_sky_matches_reference():
    d = np.load(GOLD / "misc_reference.npz")
    sky = linrt.load_sky()
    mine = OL.sky_eval(sky, d["sky_dirs"])
    assert np.array_equal(mine, d["sky_rgb"])


# escaped directions the path tracer produced on the 1080p displaced-sphere
# frame whose samples came out NaN (grazing, 0 < y < 1.6e-8), and two just above
GRAZING = np.array([[0.948603, 1.39698e-08, -0.31647], [-0.189868, 1.49012e-08, 0.98181],
                    [-0.569782, 2.79397e-09, 0.821796], [0.999903, 7.45058e-09, -0.0139272],
                    [0.948603, 2.0e-08, -0.31647], [0.6, 1.0e-6, 0.8]], np.float32)


def test_sky_grazing_nan_is_the_reference():
    """For 0 < y < 1.6e-8 the float acos of y rounds to float(pi/2), which is
    above pi/2, so cos(theta) < 0 and the zenith term sqrt(cos(theta)) of the
    Hosek model is NaN (env_emitter.h:79, ArHosekSkyModel.cc:300)."""
    assert np.cos(np.float64(np.float32(np.pi / 2))) < 0.0
    assert np.float32(np.arccos(np.float64(np.float32(1.39698e-08)))) == np.float32(np.pi / 2)
    assert np.float32(np.arccos(np.float64(np.float32(2.0e-08)))) < np.float32(np.pi / 2)
    sky = linrt.load_sky()
    mine = OL.sky_eval(sky, GRAZING)
    assert np.isnan(mine[:4]).all() and np.isfinite(mine[4:]).all()
    if OL.REF is not None:
        ref = np.zeros_like(GRAZING)
        sun = np.array(sky.sun_dir[:], np.float32)
        assert OL.REF.ref_sky_eval(OL._p(sun), C.c_float(sky.turbidity), C.c_float(sky.ground_albedo),
                                   C.c_float(sky.scale), OL._p(GRAZING), C.c_uint32(len(GRAZING)), OL._p(ref)) == 0
        assert np.array_equal(np.isnan(ref), np.isnan(mine))
        assert np.array_equal(ref[4:], mine[4:])


def test_bsdf_oracle_is_the_reference(tables):
    d = np.load(GOLD / "bsdf_reference.npz")
    mats = material_zoo(tables)
    o = OL.oracle
    for meas, tag in ((linrt.MEASURE_SOLID_ANGLE, "sa"), (linrt.MEASURE_DISCRETE, "disc")):
        f, p = o.bsdf_eval_pdf(mats, tables, d["mat"], d["wi"], d["wo"], d["front"], measure=meas)
        assert np.array_equal(f, d[f"f_{tag}"]), tag
        assert np.array_equal(p, d[f"pdf_{tag}"]), tag
    f, p = o.bsdf_eval_pdf(mats, tables, d["mat"], d["wi"], d["wo"], d["front"], mode=linrt.MODE_IMPORTANCE)
    assert np.array_equal(f, d["f_imp"]) and np.array_equal(p, d["pdf_imp"])
    s = o.bsdf_sample(mats, tables, d["mat"], d["wi"], d["front"], d["u"], d["lobe"])
    for k in ("weight", "wo", "pdf", "eta", "sampled_type", "sampled_component"):
        assert np.array_equal(s[k], d["s_" + k]), k
    flags = [o.bsdf_flags(mats, i) for i in range(len(mats))]
    assert np.array_equal(np.array(flags), d["flags"])
    # the tuples exercise every lobe: nonzero values for every material
    nz = [(d["f_sa"][d["mat"] == m].any() or d["f_disc"][d["mat"] == m].any()) for m in range(len(mats))]
    assert all(nz)


def test_hits_oracle_is_the_reference():
    d = np.load(GOLD / "hits_box_icosphere_l4.npz")
    g = linrt.box_icosphere(seed=1, subdivisions=4)
    op, t = OL.oracle.scene(g.desc_ptr).trace(d["rays"])
    assert np.array_equal(op, d["obj_prim"]) and np.array_equal(t, d["t"])


def test_render_oracle_is_the_reference():
    d = np.load(GOLD / "render_box_icosphere_16x16x4.npz")
    g = linrt.box_icosphere(seed=3, subdivisions=3)
    r = OL.oracle.scene(g.desc_ptr).render(16, 16, 4, 50, 5, mode=1, per_sample=True, fsum=True)
    assert np.array_equal(r["per_sample"], d["per_sample"])
    assert np.array_equal(r["fsum"], d["fsum"])


def pooled_z(m_a, v_a, n_a, m_b, v_b, n_b):
    """Per pixel-channel z and the pooled z of the image sums: sum(m_a - m_b) /
    sqrt(sum(se^2)).  The per-pixel mean(z) of right-skewed radiance samples is
    biased (a low sample mean comes with a low sample variance), the pooled z
    is not: it is the test of a bias of the whole image."""
    se2 = v_a / n_a + v_b / n_b
    with np.errstate(divide="ignore", invalid="ignore"):
        z = np.where(se2 > 0, (m_a - m_b) / np.sqrt(se2), np.where(np.abs(m_a - m_b) > 1e-6 * np.abs(m_b) + 1e-9,
                                                                    np.inf, 0.0))
    return z, (m_a - m_b).sum() / np.sqrt(se2.sum())


def test_oracle_image_vs_reference_render():
    """Independent streams: the oracle's Philox render (128 spp, depth 50, RR
    from 5) of the middle 128 rows of the box-room scene against the
    reference's own 256-spp render (ref_image_box_icosphere_256x256.npz, its
    Sobol sampler).  Per pixel-channel z = (m_o - m_r) / sqrt(s2_o / n_o +
    s2_r / n_r): |z| > 3 on no more than a binomial bound of 0.27%; the pooled
    z of the image sums within 4 (a 0.3% bias of the image gives ~15); mean(z)
    within 0.15 (its skew bias) and var(z) below 1.1."""
    d = np.load(GOLD / "ref_image_box_icosphere_256x256.npz")
    mr, vr, nr = d["mean"].astype(np.float64), d["var"].astype(np.float64), int(d["spp"])
    H, W, _ = mr.shape
    n = 128
    rows = (64, 192)
    g = linrt.box_icosphere(seed=1, subdivisions=5)
    o = OL.oracle.scene(g.desc_ptr).render(W, H, n, int(d["max_depth"]), int(d["rr_depth"]), seed=99, mode=0,
                                           rows=rows)
    sl = slice(*rows)
    mo = o["sum"][sl] / n
    vo = np.maximum(o["sumsq"][sl] / n - mo * mo, 0.0) * n / (n - 1)
    z, zp = pooled_z(mo, vo, n, mr[sl], vr[sl], nr)
    z = z.ravel()
    frac3 = np.mean(np.abs(z) > 3)
    bound = 0.0027 + 5 * np.sqrt(0.0027 * 0.9973 / z.size)
    print(f"oracle vs reference render: |z|>3 {frac3:.5f} (bound {bound:.5f}), pooled z {zp:+.3f}, "
          f"mean z {z.mean():+.4f}, var z {z.var():.4f}")
    assert frac3 <= bound
    assert abs(zp) < 4
    assert abs(z.mean()) < 0.15
    assert z.var() < 1.1


def _sky_json_pairs():
    import json
    data = Path(linrt.__file__).resolve().parent.parent / "data"
    out = [json.loads((data / "sky_t3_a0.3_e30.json").read_text())]
    out += json.loads((data / "skies_train.json").read_text())["skies"]
    return out


def test_sky_cook_equals_reference_cooked_skies():
    """lrt_sky_cook (the library's own cook from the vendored RGB datasets)
    reproduces, bit for bit, every sky state the reference's
    arhosek_rgb_skymodelstate_alloc_init cooked into the committed data."""
    skies = _sky_json_pairs()
    for d in skies:
        L = linrt.cook_sky(d["sun_dir"], d["turbidity"], d["ground_albedo"], d["scale"])
        cfg = np.array([[L.config[c][k] for k in range(9)] for c in range(3)])
        rad = np.array(L.channel_radiance[:])
        assert np.array_equal(cfg, np.array(d["config"])), d
        assert np.array_equal(rad, np.array(d["channel_radiance"])), d
    print(f"{len(skies)} skies bit-identical")


def test_sky_cook_layout_against_control_points():
    """At elevation pi/2 the quintic's parameter is 1, so only the last control
    point counts; at integer turbidity and albedo 0 or 1 no blending is left:
    the state is that control point of the dataset itself.  Between corners it
    is the bilinear blend (checked in float64 with numpy, to 1e-15)."""
    coef, rad = linrt.hosek_dataset()
    c4 = coef.reshape(3, 2, 10, 6, 9)
    r4 = rad.reshape(3, 2, 10, 6)
    for t in (1, 4, 10):
        for a in (0, 1):
            cfg, r3 = linrt.sky_cook(float(t), float(a), math.pi / 2)
            assert np.array_equal(cfg, c4[:, a, t - 1, 5, :])
            assert np.array_equal(r3, r4[:, a, t - 1, 5])
    t, a = 3.25, 0.4
    cfg, r3 = linrt.sky_cook(t, a, math.pi / 2)
    w = np.array([[(1 - a) * 0.75, a * 0.75], [(1 - a) * 0.25, a * 0.25]])  # [turbidity 3/4][albedo]
    want = sum(w[i, j] * c4[:, j, 2 + i, 5, :] for i in range(2) for j in range(2))
    assert np.allclose(cfg, want, rtol=1e-15, atol=1e-15)
    # elevation 0: the first control point
    cfg0, _ = linrt.sky_cook(2.0, 0.0, 0.0)
    assert np.array_equal(cfg0, c4[:, 0, 1, 0, :])
    # out of range and below the horizon
    with pytest.raises(linrt.LinrtError):
        linrt.sky_cook(0.5, 0.3, 0.5)
    with pytest.raises(linrt.LinrtError):
        linrt.sky_cook(3.0, 1.5, 0.5)
    assert np.isnan(linrt.sky_cook(3.0, 0.3, -0.1)[0]).all()
