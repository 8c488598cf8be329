"""``python -m cfd_amd.benchmarking`` -- the cfd-bench command line."""
import sys

from .cli import main

sys.exit(main())
