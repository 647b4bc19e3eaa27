"""lbm_amd -- the D2Q9 lattice-Boltzmann time step of jviquerat/lbm as
hand-written HIP kernels for MI355X (gfx950), behind the reference's
lattice/app/run interface.

    from lbm_amd.app.app import app_factory
    from lbm_amd.lattice import lattice
    from lbm_amd.run import run
    app = app_factory.create("cavity"); ltc = lattice(app); run(ltc, app)
"""
__all__ = ["lattice", "run"]
