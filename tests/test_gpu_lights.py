"""Area lights on meshes and spheres on the GPU (area_emitter.h over
triangle_mesh::sample_position / sphere::sample_position, the pdf's ray cast
of the emitter's own shape, scene::sample_emitter_ray), against the oracle,
whose path and particles on this scene equal the reference's own bit for bit
(test_mesh_and_sphere_lights_oracle_is_the_reference).

The scene (tests/common.py EmitterRoom) has a non-convex emissive mesh, so the
shadow test toward it must be the reference's closest-hit form: blocked iff the
nearest surface in the window is not the emitter's own.

Bars: identical Philox streams >= 99.9% of pixel-channels within 1e-4;
reference-sampler samples >= 99% within 1e-4; occlusion answers equal on all
but 1e-5 of the rays; photons >= 99% of records within 1e-3, <= 1% of
particles forked, flux density within 2%; eye pass >= 99% within 1e-4.
"""
import numpy as np
import pytest

import linrt
import oracle_lib as OL
from common import EmitterRoom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room():
    er = EmitterRoom()
    return er, linrt.Scene(er.desc_ptr), OL.oracle.scene(er.desc_ptr)


def test_image_vs_oracle(room):
    er, sc, orc = room
    W, H, spp = 96, 96, 16
    s, q, st = sc.render(W, H, spp, 8, 3, seed=5)
    o = orc.render(W, H, spp, 8, 3, seed=5, mode=0)
    close = np.isclose(s, o["sum"], rtol=1e-4, atol=1e-6).mean()
    print(f"emitter room {W}x{H} {spp}spp: equal to 1e-4 {close:.5f}, ext {st.extension_rays} vs {o['ext']}, "
          f"shadow {st.shadow_rays} vs {o['shadow']}, mean {s.mean() / spp:.5f} vs {o['sum'].mean() / spp:.5f}")
    assert close >= 0.999
    assert s.mean() > 0


def test_reference_sampler_vs_oracle(room):
    er, sc, orc = room
    ps = orc.render(24, 24, 2, 50, 5, mode=1, per_sample=True)["per_sample"]
    got = np.zeros_like(ps)
    for k in range(2):
        sm, _, _ = sc.render(24, 24, 1, 50, 5, spp_offset=k, reference_sampler=True)
        got[:, :, k, :] = sm.astype(np.float32)
    within = np.isclose(got, ps, rtol=1e-4, atol=1e-6).mean()
    print(f"reference sampler: within 1e-4 {within:.5f}, bit-exact {(got == ps).mean():.5f}")
    assert within >= 0.99


def test_shadow_rays_toward_the_mesh_light(room):
    """Shadow queries from points in the room to points sampled on the
    non-convex emissive mesh, ignoring that emitter: the GPU's answers equal
    the oracle's closest-hit form, and some rays meet the emitter's own surface
    before the sampled point (where an any-hit that skips the emitter would
    report whatever lies behind it)."""
    er, sc, orc = room
    rng = np.random.default_rng(3)
    n = 200_000
    o = rng.uniform(-0.9, 0.9, (n, 3))
    tri = rng.integers(0, len(er.mesh_idx), n)
    b = rng.dirichlet((1.0, 1.0, 1.0), n)
    p = er.mesh_pos.astype(np.float64)
    q = (b[:, :1] * p[er.mesh_idx[tri, 0]] + b[:, 1:2] * p[er.mesh_idx[tri, 1]] + b[:, 2:] * p[er.mesh_idx[tri, 2]])
    d = q - o
    dist = np.linalg.norm(d, axis=1)
    d /= dist[:, None]
    rays = np.empty((8, n), np.float32)
    rays[0:3], rays[3:6], rays[6], rays[7] = o.T, d.T, 0.001, dist * (1 - 1e-4)
    ignore = np.zeros(n, np.int32)  # emitter 0: the mesh
    a = sc.trace_occluded(rays, ignore)
    c = orc.occluded(rays, ignore)
    op, _ = orc.trace(rays)
    self_first = np.mean(op[:, 0] == 1)
    print(f"occluded {c.mean():.4f}; nearest surface is the emitter itself on {self_first:.4f} of the rays; "
          f"GPU differs on {(a != c).sum()}")
    assert self_first > 0.01
    assert (a != c).sum() <= n * 1e-5


def test_photons_from_mesh_and_sphere_lights(room):
    er, sc, orc = room
    req = 40000
    st = sc.photon_build(req, 8, 5, seed=7, radius=0.05, k=64)
    ph, dp, shot = sc.photon_export()
    oph, odp, oshot = orc.gather_photons(req, 8, 5, seed=7)
    assert st.full == 1 and len(ph) == req and len(oph) == req
    kd = dp[:, 1].astype(np.int64) << 8 | dp[:, 0]
    ko = odp[:, 1].astype(np.int64) << 8 | odp[:, 0]
    _, ia, ib = np.intersect1d(kd, ko, return_indices=True)
    rel = np.abs(ph[ia] - oph[ib]) / (np.abs(oph[ib]) + 1e-3)
    near = np.all(rel < 1e-3, axis=1).mean()
    forked = np.unique(np.concatenate([odp[ib[np.any(rel > 1e-2, axis=1)], 1], np.setxor1d(kd, ko) >> 8]))
    parts = np.unique(odp[:, 1])
    print(f"{len(ia)} common records, {np.mean(np.all(rel == 0, axis=1)):.4f} bit-exact, {near:.4f} within 1e-3, "
          f"forked {len(forked)} of {len(parts)}, shot {shot} vs {oshot}")
    assert near >= 0.99 and len(forked) <= 0.01 * len(parts)
    from test_gpu_photon import _density
    pick = np.random.default_rng(2).integers(0, len(oph), 4000)
    qp = oph[pick, :3] + np.random.default_rng(3).normal(scale=0.05, size=(4000, 3)).astype(np.float32)
    ratio = _density(ph, qp).sum() * oshot / (_density(oph, qp).sum() * shot)
    print(f"flux density ratio {ratio:.5f}")
    assert abs(ratio - 1) < 0.02


def test_eye_pass_vs_oracle(room):
    er, sc, orc = room
    R = sc.info.bounding_radius
    sc.photon_build(30000, 127, 5, seed=5, radius=0.05 * R, k=120, kind=linrt.PHOTON_SURFACE)
    gp, gd, gs = sc.photon_export(linrt.PHOTON_SURFACE)
    W, H, spp = 24, 24, 2
    s, q, st = sc.render_photon(W, H, spp, 128, seed=3)
    maps = [(gp, gd, 1.0 / gs, 0.05 * R), None]
    ps = orc.render_photon(W, H, spp, maps, mode=0, seed=3, tree=1)
    m = ps.astype(np.float64).mean(2)
    close = np.isclose(s / spp, m, rtol=1e-4, atol=1e-6).mean()
    print(f"eye pass: {close:.5f} of pixel-channels within 1e-4, mean {s.mean() / spp:.5f} vs {m.mean():.5f}")
    assert close >= 0.99
