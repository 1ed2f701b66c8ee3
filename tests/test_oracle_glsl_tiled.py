"""The oracle's restatement of the renderer's texture path (convert-2d-index.glsl, interpolate-2d.glsl,
make-2d-index-from-4d.glsl, interpolate-4d.glsl, ray-scatter-outer.glsl over the tiled .scatter layout of
image.clj:299-312) against the multilinear interpolation of the untiled tables (interpolate.clj:75-110):
with exact bilinear texture weights the two are the same function, so every random probe agrees to
FP64 rounding.  A reversed dimension, a missing half-texel offset or a wrong tile stride breaks it."""
import numpy as np


def test_interpolate_4d_over_tiles_is_multilinear(oracle):
    rng = np.random.default_rng(11)
    for shape in ((3, 5, 4, 6), (2, 7, 3, 2), (32, 9, 5, 8)):
        table = rng.uniform(-1, 1, shape + (3,))
        tex = oracle.convert_4d_to_2d(table).reshape(-1, 3)
        for _ in range(40):
            glsl_idx = rng.uniform(0, 1, 4)  # (heading, sun elevation, elevation, height), normalised
            glsl_idx[rng.integers(4)] = rng.choice([0.0, 1.0])  # edges too
            clj = [c * (n - 1) for c, n in zip(reversed(glsl_idx), shape)]
            want = oracle.interpolate(table.reshape(-1, 3), list(shape), clj)
            got = oracle.glsl_interpolate_4d(tex, *shape, glsl_idx)
            assert np.allclose(got, want, rtol=0, atol=1e-12), (shape, glsl_idx)


def test_interpolate_2d_is_multilinear(oracle):
    rng = np.random.default_rng(12)
    for shape in ((4, 7), (64, 255), (16, 63)):
        table = rng.uniform(-1, 1, shape + (3,))
        for _ in range(40):
            glsl_idx = rng.uniform(0, 1, 2)  # (elevation or sun elevation, height)
            clj = [glsl_idx[1] * (shape[0] - 1), glsl_idx[0] * (shape[1] - 1)]
            want = oracle.interpolate(table.reshape(-1, 3), list(shape), clj)
            got = oracle.glsl_interpolate_2d(table.reshape(-1, 3), shape[0], shape[1], glsl_idx)
            assert np.allclose(got, want, rtol=0, atol=1e-12)


def test_ray_scatter_outer_over_tiles_matches_untiled_lookup(oracle):
    """ray-scatter-outer.glsl on the tiled S and M textures equals S + M phase(0.76, mu) looked up in the
    untiled tables at ray-scatter-forward of the same point, direction and light."""
    rng = np.random.default_rng(13)
    earth = oracle.planet(6378000.0, 35000.0)
    shape = (6, 11, 7, 4)
    S = rng.uniform(0, 1, shape + (3,))
    M = rng.uniform(0, 1, shape + (3,))
    St, Mt = oracle.convert_4d_to_2d(S).reshape(-1, 3), oracle.convert_4d_to_2d(M).reshape(-1, 3)
    for _ in range(50):
        h = rng.uniform(0, 35000.0)
        point = np.array([6378000.0 + h, 0, 0])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        l = rng.normal(size=3)
        l /= np.linalg.norm(l)
        if not oracle.is_above_horizon(earth, point, d):
            continue
        got = oracle.glsl_tiled_ray_scatter_outer(earth, shape, St, Mt, 0.76, l, point, d)
        idx = oracle.space_forward(earth, 4, shape, point, d, l, True)
        want = (oracle.interpolate(S.reshape(-1, 3), list(shape), idx) +
                oracle.interpolate(M.reshape(-1, 3), list(shape), idx) * oracle.phase(0.76, float(d @ l)))
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
