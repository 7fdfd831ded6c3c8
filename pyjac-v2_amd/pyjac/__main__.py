"""``python -m pyjac --lang hip -i mech.inp ...`` (reference pyjac/__main__.py)."""

import sys

from pyjac import utils


def main(argv=None):
    utils.create(argv)
    return 0


if __name__ == '__main__':
    sys.exit(main())
