"""App registry (lbm/src/app/app.py:10-17) plus the benchmark configurations."""
from lbm_amd.factory import factory
from lbm_amd.app.cavity import cavity
from lbm_amd.app.turek import turek
from lbm_amd.app.poiseuille import poiseuille
from lbm_amd.app.array import array
from lbm_amd.app.step import step
from lbm_amd.app.configs import cavity200, turek200, obstacle_array

app_factory = factory()
app_factory.register("cavity", cavity)
app_factory.register("turek", turek)
app_factory.register("poiseuille", poiseuille)
app_factory.register("array", array)
app_factory.register("step", step)
app_factory.register("cavity200", cavity200)
app_factory.register("turek200", turek200)
app_factory.register("obstacle_array", obstacle_array)
