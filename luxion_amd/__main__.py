import sys

from .cli import main

sys.exit(main(["luxion"] + sys.argv[1:]))
