"""Small end-to-end invocation of the engine, used by the smoke entry point."""


def run():
    import numpy as np
    from egsnrc_amd import _lib
    from tests import oracle_lib
    xb = np.linspace(-2, 2, 9)
    g, keep = _lib.make_geometry(xb, xb, np.linspace(0, 4, 9))
    x = np.array([[0.1, 0.2, -1.0], [-3.0, 0.3, 1.1]])
    u = np.array([[0.1, 0.2, 0.97], [1.0, 0.0, 0.0]])
    u /= np.linalg.norm(u, axis=1)[:, None]
    a = _lib.trace_rays(g, x, u, 64, 32)
    b = oracle_lib.trace_rays(g, x, u, 64, 32)
    for p, q in zip(a, b):
        assert np.array_equal(p, q), "voxel trace mismatch"
    print("smoke ok")
