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


def test_box_filter_weights_cells_by_volume_on_a_clustered_grid():
    """on a non-uniform radial grid the coarse centre value is the volume-weighted mean of its
    r fine cells (width times radius: the exact integral of r dr over the cell), so the fine
    cells' exact means of a field filter to the exact mean over the coarse cell; face
    variables are unweighted"""
    from simpletcsolver_amd import closure
    nz, ny, nx, r = 4, 8, 16, 4
    xf = 1.0 + 0.4 * (1.0 - np.cos(np.pi * np.arange(nx + 1) / nx)) / 2.0
    w = closure.radial_weights(xf, True)
    assert np.allclose(w, (xf[1:] - xf[:-1]) * 0.5 * (xf[1:] + xf[:-1]))
    def cell_mean(a, b):                     # exact mean of r^2 / 2 over [a, b], weight r dr
        return ((b ** 4 - a ** 4) / 8.0) / ((b ** 2 - a ** 2) / 2.0)
    fine = cell_mean(xf[:-1], xf[1:])
    t = np.zeros((nz, ny, nx + 2))
    t[:, :, 1:-1] = fine
    c = closure.box_filter({"t": t}, r, w)
    xfc = xf[::r]
    assert np.allclose(c["t"][0, 0, 1:-1], cell_mean(xfc[:-1], xfc[1:]), rtol=1e-13)
    # face variables are sampled on the coarse faces and averaged over y and z only
    rng = np.random.default_rng(1)
    ux = rng.standard_normal((nz, ny, nx + 1))
    cu = closure.box_filter({"ux": ux}, r, w)
    assert np.isclose(cu["ux"][0, 1, 2], ux[0:4, 4:8, 8].mean())


def test_model_power_is_the_j_weighted_work_of_the_corrections():
    """closure.model_power with a stand-in network of constant outputs: the ux correction at a
    face is the mean of its two neighbouring centres (faces 2..nx), so with unit outputs and
    ux = 1 on the faces the power is the sum of JDXF over the inner faces, plus the centre
    terms with uy = 2, uz = 3"""
    from simpletcsolver_amd import closure
    nz, ny, nx = 4, 6, 8
    lengths = (0.4, 1.0, 2.0)
    xf = 1.0 + 0.4 * np.linspace(0.0, 1.0, nx + 1) ** 1.3
    xc = np.concatenate([[xf[0]], 0.5 * (xf[:-1] + xf[1:]), [xf[-1]]])

    class Fields:
        def fields(self):
            return {"ux": np.ones((nz, ny, nx + 1)), "uy": 2.0 * np.ones((nz, ny, nx + 2)),
                    "uz": 3.0 * np.ones((nz, ny, nx + 2)), "t": np.zeros((nz, ny, nx + 2))}

    class Net:
        def predict(self, x):
            assert x.shape == (1, nz, ny, nx, 5)
            return np.ones((1, nz, ny, nx, 4), dtype=np.float32)

    class Norm:
        out_scale = np.array([0.5, 0.25, 0.125, 1.0], dtype=np.float32)

        def x(self, x):
            return x

    p = closure.model_power(Fields(), Net(), Norm(), xf, xc, lengths)
    hz = lengths[2] / nz
    jf = (xc[1:] - xc[:-1]) * (xf * lengths[1] / ny) * hz
    jc = (xf[1:] - xf[:-1]) * (xc[1:-1] * lengths[1] / ny) * hz
    want = ny * nz * (0.5 * jf[1:nx].sum() + (2.0 * 0.25 + 3.0 * 0.125) * jc.sum())
    assert np.isclose(p, want, rtol=1e-12)
