"""`python -m ym_amd <command>`: see ym_amd.cli."""
import sys

from .cli import main

sys.exit(main())
