"""The uneven-batch job generator (random orthorhombic crystals on 360-720
point grids): parameter ranges, determinism, periodicity of the specimen."""
import numpy as np

from jmultislicelib_amd import configs


def test_job_ranges_and_determinism():
    a = configs.uneven_jobs(12, seed=2)
    b = configs.uneven_jobs(12, seed=2)
    for ja, jb in zip(a, b):
        sa, sb = ja.setup, jb.setup
        assert sa.nx == sa.ny and sa.nx in configs.JOB_EDGES
        assert 50 <= len(sa.objslc) <= 400 and 16 <= len(ja.dx) <= 256
        assert 1 <= len(ja.species) <= 4 and ja.species.min() >= 6 and ja.species.max() <= 79
        assert abs(sa.a0[0] * 10 - 0.06 * sa.nx) < 1e-9
        assert np.array_equal(sa.objslc, np.arange(len(sa.objslc)) % 2)
        assert all(np.array_equal(g, h) for g, h in zip(sa.gratings, sb.gratings))
        assert np.array_equal(ja.dx, jb.dx) and np.array_equal(ja.dy, jb.dy)


def test_job_grating_is_cell_periodic():
    """The super-cell holds whole cells: shifting a grating by one cell (when
    the cell is a whole number of pixels) or comparing it with its own
    Fourier-space lattice (otherwise) must show the lattice period."""
    rng = np.random.default_rng(7)
    job = configs.random_job(rng, edge=360, nslices=50, npos=16)
    sp = job.setup
    ax, by, _ = job.cell_ang
    L = sp.a0[0] * 10
    ncx, ncy = round(L / ax), round(L / by)
    for g in sp.gratings:
        p = np.abs(np.fft.fft2(np.angle(g[0])))
        p[0, 0] = 0
        # only reciprocal-lattice frequencies (multiples of the cell count) carry power
        ky, kx = np.meshgrid(np.arange(sp.ny), np.arange(sp.nx), indexing="ij")
        on = (kx % ncx == 0) & (ky % ncy == 0)
        assert p[~on].max() < 1e-3 * max(p[on].max(), 1e-30) or p.max() == 0
