"""The closure's box filter on the CPU: coarse values of each staggered variable are the
means of the fine values that make them up (faces sampled on the coarse faces)."""
import numpy as np


def test_box_filter_is_consistent_with_the_staggering():
    from simpletcsolver_amd import closure
    rng = np.random.default_rng(0)
    nz, ny, nx, r = 8, 12, 16, 4
    f = {"ux": rng.standard_normal((nz, ny, nx + 1)), "uy": rng.standard_normal((nz, ny, nx + 2)),
         "uz": rng.standard_normal((nz, ny, nx + 2)), "t": rng.standard_normal((nz, ny, nx + 2))}
    c = closure.box_filter(f, r)
    assert c["ux"].shape == (nz // r, ny // r, nx // r + 1)
    assert c["uy"].shape == (nz // r, ny // r, nx // r + 2)
    # coarse ux face I is the mean of fine face r I over its r x r fine faces
    assert np.isclose(c["ux"][1, 2, 3], f["ux"][4:8, 8:12, 12].mean())
    # coarse uy face J: fine face r J, averaged over the cells in x and z
    assert np.isclose(c["uy"][1, 2, 3], f["uy"][4:8, 8, 9:13].mean())
    assert np.isclose(c["uz"][1, 2, 3], f["uz"][4, 8:12, 9:13].mean())
    assert np.isclose(c["t"][1, 2, 3], f["t"][4:8, 8:12, 9:13].mean())
    assert np.isclose(c["t"][1, 2, 0], f["t"][4:8, 8:12, 0].mean())
    assert np.isclose(c["t"][1, 2, -1], f["t"][4:8, 8:12, -1].mean())
