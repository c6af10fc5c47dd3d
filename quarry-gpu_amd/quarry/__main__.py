"""`python -m quarry selftest | run | query`: the command line of the GPU path."""
import sys

from .cli import main

sys.exit(main())
