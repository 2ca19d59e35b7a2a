"""``python -m fastcorticalwiring_amd``: the command line of :mod:`fastcorticalwiring_amd.cli`."""

from .cli import main

main()
