"""Command line: ``python3 start.py <app_name>`` (the reference's start.py:13-29)."""
import sys

from lbm_amd.app.app import app_factory
from lbm_amd.lattice import lattice
from lbm_amd.run import run

if __name__ == '__main__':
    if len(sys.argv) != 2:
        print('Command line error, please use as follows:')
        print('python3 start.py app_name')
        sys.exit(1)
    app = app_factory.create(sys.argv[1])
    ltc = lattice(app)
    run(ltc, app)
