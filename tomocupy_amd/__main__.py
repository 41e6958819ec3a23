"""``python -m tomocupy_amd recon ...``: see :mod:`tomocupy_amd.cli`."""
import sys

from .cli import main

sys.exit(main())
